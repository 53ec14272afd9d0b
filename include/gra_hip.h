/*
 * gra_hip.h -- C ABI of the MI355X (gfx950) event-generation inner loop.
 *
 * The library evaluates the per-point event weight W = |M|^2 x Jacobian of the
 * GRANIITTI <C> continuum class (pp -> p X p, X = 2 central particles) on the
 * GPU, runs VEGAS iterations over it, and trains / samples a rational-quadratic
 * spline normalizing flow used as an alternative importance sampler.
 *
 * Conventions
 *   - every entry point returns an int status: GRA_OK (0) or a negative
 *     GRA_E_* code; the message of the last failure is in gra_hip_last_error();
 *   - no C++ exception crosses this boundary;
 *   - host arrays are owned by the caller and are only touched during the call;
 *   - a handle owns all its device memory and one HIP stream, and is used by one
 *     host thread at a time (the reference keeps one process copy per thread for
 *     the same reason, src/MGraniitti.cc:278-309);
 *   - calls are synchronous.
 *
 * Each declaration names the reference interface it replaces
 * (mieskolainen/GRANIITTI v1.051, paths relative to the reference root).
 */
#ifndef GRA_HIP_H
#define GRA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gra_handle gra_handle; /* opaque: device state of one process */

/* ------------------------------------------------------------------------- */
/* Status codes                                                              */
/* ------------------------------------------------------------------------- */
#define GRA_OK 0
#define GRA_E_ARG (-1)     /* invalid argument (reference: std::invalid_argument) */
#define GRA_E_HIP (-2)     /* HIP runtime failure */
#define GRA_E_NOMEM (-3)   /* device allocation failed */
#define GRA_E_STATE (-4)   /* call not valid in the handle's state */
#define GRA_E_NUMERIC (-5) /* more than GRA_MAX_NONFINITE NaN (or Inf) weights
                              (include/Graniitti/MProcess.h:317-340), or a VEGAS integral that is
                              NaN/Inf or below 1e-60 (src/MGraniitti.cc:1258-1273) */
#define GRA_E_IO (-6)      /* file missing, unreadable or malformed */

/* Phase-space classes */
#define GRA_PS_C 0 /* <C>: MContinuum, 2 -> 4 in (pt_i, phi_i, kt, phi_kt, y_3, y_4) */
#define GRA_PS_F 1 /* <F>: MFactorized B51, system (M^2, Y) x isotropic 2-body decay */

/* ------------------------------------------------------------------------- */
/* Limits                                                                    */
/* ------------------------------------------------------------------------- */
#define GRA_MAX_DIM 16     /* hypercube dimension limit (2->4: 8, rho rho -> 4 pi: 14) */
#define GRA_VEGAS_BINS 128 /* VEGASPARAM::BINS default, include/Graniitti/MVEGAS.h:19 */

/* Processes (sub-process registry, include/Graniitti/MSubProc.h) */
#define GRA_PROC_REGGE_CON 15  /* PROC_15 "PP[CON]" 2-body: MRegge::ME4, include/Graniitti/MSubProc.h:361-380 */
#define GRA_PROC_TENSOR_CON 12 /* PROC_12 "PP[CONTENSOR]": MTensorPomeron::ME4, MSubProc.h:288-300;
                                  2xV (vector-meson pair) or 2xS (pseudoscalar pair) by central_spin2 */
#define GRA_PROC_RES_CON 17    /* PROC_17 "PP[RES+CON]" 2-body: MRegge::ME4 + resonances, MSubProc.h:398-432 */
#define GRA_PROC_RES 9         /* PROC_9 "PP[RES]": resonances only, MSubProc.h:233-262 */

/* Resonances of PP[RES+CON] */
#define GRA_MAX_RES 8
#define GRA_MAX_VETO 8 /* veto domains (the reference's card reader takes up to 100) */
/* Breit-Wigner forms, PARAM_RES "BW" (form::CBW, src/MForm.cc:743-828) */
#define GRA_BW_FIXED 1   /* fixed-width relativistic */
#define GRA_BW_RUNNING 2 /* running-width relativistic */
#define GRA_BW_BARRIER 3 /* spin-barrier (Blatt-Weisskopf type) relativistic */
#define GRA_BW_SPIN 4    /* spin-dependent relativistic */
/* Rest frame of the spin density, PARAM_SPIN "FRAME" (spin::GetRhoRotation, src/MSpin.cc:544-610) */
#define GRA_FRAME_CM 0 /* no rotation */
#define GRA_FRAME_HX 1 /* helicity */
#define GRA_FRAME_CS 2 /* Collins-Soper */
#define GRA_FRAME_GJ 3 /* Gottfried-Jackson */

/* Screening-loop quadratures */
#define GRA_LOOP_SIMPSON38 0
#define GRA_LOOP_GAUSS 1

/* Amplitude arithmetic */
#define GRA_AMP_FP64 0 /* everything FP64, operation order of the reference */
#define GRA_AMP_FP32 1 /* FP64 kinematics/invariants/Jacobian, FP32 complex amplitude + compensated sums */

/* Per-point flag bits (AuxIntData, include/Graniitti/MProcess.h:37-52) */
#define GRA_FLAG_KIN_OK 0x01u /* kinematics_ok */
#define GRA_FLAG_FID_OK 0x02u /* fidcuts_ok   */
#define GRA_FLAG_VETO_OK 0x04u /* vetocuts_ok */
#define GRA_FLAG_AMP_OK 0x08u /* amplitude_ok (finite weight, CheckInfNan) */
#define GRA_FLAG_INF 0x10u    /* with GRA_FLAG_AMP_OK clear: the weight was +-Inf (else NaN) */

/* CheckInfNan's limit: a run fails once more than this many NaN, or more than
 * this many Inf, weights have been seen (include/Graniitti/MProcess.h:317-340) */
#define GRA_MAX_NONFINITE 50

/* ------------------------------------------------------------------------- */
/* Configuration (plain old data, passed by value through the ABI)           */
/* ------------------------------------------------------------------------- */
/* One resonance (form::ReadResonance, src/MForm.cc:213-330; modeldata/TUNE0/RES_*.json)
 * and its decay to the two central particles (MProcess::ProcessHelicityDecay,
 * src/MProcess.cc:534-740; modeldata/TUNE0/BRANCHING.json). Spins J = 0, 1, 2
 * with parity (-1)^J: the states that decay to two spinless central particles. */
typedef struct gra_res {
  double mass, width;            /* "M", "W" (GeV) */
  double g_A, g_phi;             /* production coupling g_A exp(i g_phi) */
  double g_FF;                   /* resonance form-factor parameter (GeV) */
  double BR;                     /* branching fraction to the central pair */
  double alpha_re, alpha_im;     /* decay helicity coupling alpha_{l = J, s = 0} */
  double rho_re[25], rho_im[25]; /* spin density, (2J+1) x (2J+1) row-major, projections -J..J */
  int32_t J, P;                  /* spin, parity */
  int32_t BW;                    /* GRA_BW_* */
  int32_t pdg;
} gra_res;

typedef struct gra_cfg {
  int32_t abi_size; /* = sizeof(gra_cfg), checked by every entry point */
  int32_t process;  /* GRA_PROC_* */
  int32_t amp_precision; /* GRA_AMP_* */
  int32_t dim;      /* hypercube dimension (output of gra_hip_cfg_default) */

  /* Beams (MProcess::SetBeamEnergies, src/MProcess.cc:313-334). Unequal energies
   * boost the <C> kinematics to the lab (src/MContinuum.cc:386-395); they are
   * accepted for the unscreened <C> class only. */
  double beam_E1, beam_E2; /* GeV */
  double m_beam;           /* proton mass, modeldata/mass_width_2018.mcd:169 */

  /* Central particles: first-level decaytree legs */
  double m_central;     /* pole mass of each central particle */
  double w_central;     /* full width (0 => fixed mass, GetOffShellMass src/MProcess.cc:755) */
  double offshell;      /* BW window in widths, MProcess::OFFSHELL (src/MProcess.h:363) */
  double m_daughter;    /* decay daughter mass (rho -> pi pi); 0 if no decay */
  double S_factor;      /* identical-particle factor (CalculateSymmetryFactor) */

  /* Generation cuts (GENCUT, include/Graniitti/MKinematics.h:1554-1576) */
  double fwd_pt_min, fwd_pt_max; /* forward leg pt; defaults 0, 2.5 (src/MProcess.cc:1369-1386) */
  double kt_min, kt_max;         /* intermediate kt; default 0, 10/Kf (src/MContinuum.cc:71-82) */
  double rap_min, rap_max;       /* central final state rapidity */

  /* Fiducial cuts (FIDCUT, include/Graniitti/MKinematics.h:1579-1611) */
  int32_t fid_active;
  int32_t pad0;
  double fid_eta_min, fid_eta_max, fid_rap_min, fid_rap_max;
  double fid_pt_min, fid_pt_max, fid_et_min, fid_et_max;
  double sys_M_min, sys_M_max, sys_Y_min, sys_Y_max, sys_Pt_min, sys_Pt_max;
  double fwd_t_min, fwd_t_max, fwd_M_min, fwd_M_max;

  /* Soft model (PARAM_SOFT, modeldata/TUNE0/GENERAL.json) */
  double delta_P, alpha_P, gN_P, fc1, fc2, fc3;
  double mpi_const; /* PDG::mpi used in the pion loop of S3PomAlpha, include/Graniitti/MPDG.h:56 */

  /* Regge continuum (PARAM_REGGE + PARAM_PI) */
  double regge_a0[3], regge_ap[3], regge_sgn[3], regge_c[3];
  int32_t regge_on[3];
  int32_t offshellFF; /* 0 = exp, 1 = orear, 2 = pow (src/MRegge.cc:1155-1174) */
  double regge_s0, b_EXP, a_OREAR, b_OREAR, b_POW;
  double m_prop; /* off-shell propagator nominal mass, decaytree[0].p.mass (src/MRegge.cc:396) */

  /* Tensor pomeron (PARAM_TENSORPOM) */
  double tp_delta_P, tp_ap_P, tp_gPNN, tp_gPvv[2];
  double tp_mp_ff;  /* proton mass used by the Kelly/F1 form factor */
  double tp_gPpipi, tp_gPKK; /* Pomeron-pion-pion and Pomeron-kaon-kaon couplings (2xS) */
  /* 2J of each central particle (lts.decaytree[i].p.spinX2): selects the tensor-pomeron
   * spin mode, 0 -> "2xS" (pseudoscalar pair), 2 -> "2xV" (vector-meson pair)
   * (MTensorPomeron::ME4, src/MTensorPomeron.cc:643-654) */
  int32_t central_spin2;
  int32_t pad_tp;

  /* Eikonal screening (MEikonalNumerics, include/Graniitti/MEikonal.h:26-57) */
  int32_t screening;
  int32_t n_loop_phi; /* intervals, multiple of 3 */
  int32_t n_loop_kt;  /* intervals, multiple of 3 */
  int32_t n_bt;       /* NumberBT: density table intervals (modeldata/NUMERICS.json) */
  int32_t n_kt2;      /* NumberKT2: screening table intervals */
  int32_t n_fb;       /* FBIntegralN: Fourier-Bessel quadrature intervals (even) */
  double min_loop_kt, max_loop_kt;
  /* Loop quadrature: GRA_LOOP_SIMPSON38, the reference's (n_loop_phi + 1) x
   * (n_loop_kt + 1) Simpson 3/8 grid (S3ScreenedAmp2, src/MProcess.cc:187-261), or
   * GRA_LOOP_GAUSS: n_loop_kt + 1 Gauss-Legendre nodes in kt on [MinLoopKT,
   * MaxLoopKT] times n_loop_phi + 1 equally spaced azimuths (the periodic trapezoid
   * rule), any node counts. eik_log_nodes = 0 takes A_eik(kt^2) from the MSA table
   * by linear interpolation (IArray1D::Interpolate1D, src/MEikonal.cc:489-512);
   * eik_log_nodes = M > 1 evaluates A_eik's Fourier-Bessel integral at M
   * log-spaced kt on [MinLoopKT, MaxLoopKT] and interpolates linearly in log kt. */
  int32_t loop_rule;
  int32_t eik_log_nodes;
  /* Phase-space class, the <C> / <F> suffix of the process string (MProcess
   * classes MContinuum / MFactorized). With GRA_PS_F, rap_min/rap_max bound the
   * central system's rapidity and gen_M_min/gen_M_max its mass (GENCUTS "<F>",
   * e.g. tests/run_compare_PS/card.json); the reference has no default for the
   * mass range (GENCUT::M_max = 0, include/Graniitti/MKinematics.h:1565). */
  int32_t phase_space; /* GRA_PS_C or GRA_PS_F */
  int32_t pad_ps;
  double gen_M_min, gen_M_max;
  /* Resonances of PP[RES+CON], summed with the continuum amplitude in this
   * order (PROC_17::Amp2, include/Graniitti/MSubProc.h:398-432); the spin
   * switches and frame of PARAM_SPIN and the flux exponent PARAM_REGGE::omega
   * (modeldata/TUNE0/GENERAL.json) */
  int32_t n_res;
  int32_t spin_gen;   /* SPINGEN: 2 -> 1 production spin correlations */
  int32_t spin_dec;   /* SPINDEC: 1 -> 2 decay spin correlations */
  int32_t spin_frame; /* GRA_FRAME_* */
  double regge_omega;
  gra_res res[GRA_MAX_RES];
  /* Veto domains (VETOCUT, include/Graniitti/MKinematics.h:1613-1626; read from
   * the card's "VETOCUTS" block, src/MGraniitti.cc:910-962). With veto_active, a
   * point whose final-state central particles include one with pt in
   * [veto_pt_min, veto_pt_max] and eta in [veto_eta_min, veto_eta_max] of any of
   * the n_veto domains is vetoed: W = 0 and GRA_FLAG_VETO_OK clear
   * (MProcess::VetoCuts / FindVetoCuts, src/MProcess.cc:1069-1082, 1157-1175).
   * Every domain needs max > min in both ranges (aux::AssertCut). The elastic
   * forward protons are never vetoed: the reference tests its forward decay
   * branches, which stay empty (a zero four-vector, whose pseudorapidity is NaN)
   * unless the proton is excited. */
  int32_t veto_active;
  int32_t n_veto;
  double veto_eta_min[GRA_MAX_VETO], veto_eta_max[GRA_MAX_VETO];
  double veto_pt_min[GRA_MAX_VETO], veto_pt_max[GRA_MAX_VETO];
} gra_cfg;

/* ------------------------------------------------------------------------- */
/* VEGAS                                                                     */
/* ------------------------------------------------------------------------- */
/* Grid: bin upper edges, dimension-major: xmat[j * GRA_VEGAS_BINS + i] is the
 * reference's VD.xmat[i][j] (include/Graniitti/MVEGAS.h:212-214). */
typedef struct gra_grid {
  int32_t dim;
  int32_t bins;
  double xmat[GRA_MAX_DIM * GRA_VEGAS_BINS];
} gra_grid;

/* One VEGAS iteration's accumulators (the locked section of
 * MGraniitti::VEGASMultiThread, src/MGraniitti.cc:1389-1419). Sums are kept as
 * (value, compensation) pairs; fmat/f2mat are dimension-major like xmat. */
typedef struct gra_iter_stats {
  double fsum, fsum_c, f2sum, f2sum_c;
  double maxW, maxf;
  uint64_t n_calls, n_kin_ok, n_fid_ok, n_valid;
  uint64_t n_nan, n_inf; /* weights that were NaN / +-Inf (set to 0, CheckInfNan) */
  double fmat[GRA_MAX_DIM * GRA_VEGAS_BINS];
  double f2mat[GRA_MAX_DIM * GRA_VEGAS_BINS];
} gra_iter_stats;

/* Fixed, clock-independent schedule for MGraniitti::SampleVegas
 * (src/MGraniitti.cc:1074-1135). */
typedef struct gra_vegas_cfg {
  uint64_t ncall;        /* calls per iteration */
  uint64_t seed;
  int32_t burnin_iter;   /* init = 0 iterations (grid adaptation, integral discarded) */
  int32_t iter;          /* init = 1 iterations (integral accumulated) */
  int32_t max_iter;      /* cap when extending iterations for precision (>= iter) */
  int32_t bins;          /* <= GRA_VEGAS_BINS, even */
  double lambda;         /* VEGASPARAM::LAMBDA */
  double precision;      /* stop when sigma_err/sigma <= precision (0 = fixed schedule) */
  double chi2max;        /* VEGASPARAM::CHI2MAX */
} gra_vegas_cfg;

typedef struct gra_vegas_result {
  double sigma, sigma_err, chi2;   /* barn */
  double maxW, maxf;
  double seconds;                  /* wall time of the init=1 phase (incl. host grid updates) */
  double seconds_total;            /* burn-in + main */
  double seconds_to_precision;     /* wall time until sigma_err/sigma <= precision, or -1 */
  uint64_t calls_total;
  int32_t iterations;              /* init = 1 iterations done */
  int32_t restarts;                /* chi2 > 50 restarts (burn-in with 10 x calls, or with twice
                                      the burn-in iterations), src/MGraniitti.cc:1093-1126, 1276-1282 */
  double kernel_ms;                /* sum of integrand-kernel durations (HIP events) */
} gra_vegas_result;

/* ------------------------------------------------------------------------- */
/* Normalizing flow (rational-quadratic spline coupling layers)              */
/* ------------------------------------------------------------------------- */
typedef struct gra_flow_cfg {
  int32_t dim;       /* = integrand dimension */
  int32_t layers;    /* coupling layers */
  int32_t bins;      /* spline bins K */
  int32_t hidden;    /* MLP width (multiple of 32) */
  int32_t depth;     /* hidden layers per MLP (1..4); a negative value selects the FP32
                        VALU MLP path (gradient checks) instead of BF16 MFMA */
  int32_t steps;     /* Adam steps */
  uint64_t batch;    /* points per training step */
  double lr;         /* Adam learning rate */
  double beta1, beta2, eps;
  uint64_t seed;
  double ess_fallback; /* ESS fraction below which sampling falls back to VEGAS */
  int32_t spline_loops; /* nonzero: spline kernels loop over a runtime bin count for every K
                           (0: unrolled kernels for K = 8, 16, 32; used to check them) */
  int32_t out_groups;   /* transformed dimensions per output group of the fused sampling layer
                           (0: the largest group that fits two blocks per CU in LDS) */
  int32_t wgrad_split;  /* nonzero: BF16 weight gradients in split 64 x 64 tiles for every width
                           (0: 128 x 128 tiles for widths that are multiples of 128, and for
                           the input layer; used to check them) */
  int32_t pad1;
} gra_flow_cfg;

typedef struct gra_flow_stats {
  double sigma, sigma_err;  /* plain mean of w = f/q and its standard error (barn) */
  double cv2;               /* Var(w)/<w>^2 */
  double ess;               /* (sum w)^2 / (N sum w^2) */
  double maxw;
  double loss;              /* last training-loss value */
  double seconds;           /* wall time of the call */
  uint64_t n;               /* points */
  int32_t fallback;         /* 1 if ESS < ess_fallback */
  int32_t pad0;
} gra_flow_stats;

/* One unweighted event: the final-state 4-momenta (E, px, py, pz) in GeV as
 * MProcess::EventRecord lists them (src/MProcess.cc:1287-1314): p[0], p[1] the
 * forward protons, p[2], p[3] the central particles, p[4..7] their two-body
 * decay daughters (p[4], p[5] from p[2]; p[6], p[7] from p[3]) when the central
 * particles decay. n_p = 4 or 8. f is the point's VEGAS-weighted integrand
 * (barn), index its Philox point index within the iteration. */
typedef struct {
  double p[8][4];
  double f;
  uint64_t index;
  int32_t n_p;
  int32_t pad0;
} gra_event;

/* Counters of one generation pass (MGraniitti::Stats, include/Graniitti/MGraniitti.h:46-93) */
typedef struct {
  uint64_t trials;    /* points sampled */
  uint64_t valid;     /* AuxIntData::Valid() points */
  uint64_t accepted;  /* hit-or-miss hits among valid points */
  uint64_t overflow;  /* points with f > maxf (never accepted) */
  uint64_t stored;    /* events written (min(accepted, max_events)) */
} gra_gen_stats;

/* ------------------------------------------------------------------------- */
/* Entry points                                                              */
/* ------------------------------------------------------------------------- */

/* Fill a configuration with one of the preset workloads:
 *   1: PP[CON]<C> -> pi+ pi-, 13 TeV, |eta|<0.9, pT>0.15, no screening, FP64
 *   2: preset 1 + eikonal screening (16 phi x 31 kt nodes, MaxLoopKT 3), FP32 amplitude
 *   3: PP[CONTENSOR]<C> -> rho0 rho0 -> 4 pi, |eta|<2.5, pT>0.1, screening, FP32
 *   4: PP[CON]<F> -> pi+ pi-, preset 1's cuts, system |Y| < 0.9 and M in [0, 4] GeV
 *      (GENCUTS "<F>" of tests/run_compare_PS/card.json), no screening, FP64
 *   5: preset 4 + eikonal screening as preset 2, FP32 amplitude
 *   6: PP[RES+CON]<C> -> pi+ pi- (tests/processes/ALICECUTS.json): preset 1's continuum and
 *      cuts plus f0(500), rho(770), f0(980), phi(1020), f2(1270), f0(1500), f2'(1525),
 *      f0(1710) with spin correlations in the Collins-Soper frame, FP64 amplitude
 *   7: preset 6 + eikonal screening as preset 2 (FP64 amplitude)
 *   8: PP[RES]<F> -> pi+ pi- with f2(1270) alone, |J_z| = 2 (rho = diag(1/2, 0, 0, 0, 1/2)),
 *      Collins-Soper frame, preset 4's generation cuts, no fiducial cuts
 *      (tests/run_JW_frames/run.sh), no screening, FP64
 *   9: PP[CONTENSOR]<C> -> pi+ pi- (tensor pomeron, 2xS), preset 1's cuts, no screening, FP64
 *  10: preset 9 + eikonal screening as preset 2, FP64 amplitude
 *  11: preset 2 in the stated loop discretisation: 32 Gauss-Legendre kt nodes on
 *      [1e-4, 3] GeV x 16 azimuths, A_eik at 64 log-spaced kt nodes (the azimuthally
 *      symmetric content of a 64 x 64 (kt, phi) table), FP32 amplitude
 * Replaces the steering-card reader MGraniitti::ReadInput (src/MGraniitti.cc:318-329)
 * for the keys these workloads use, with TUNE0 values (modeldata/TUNE0/GENERAL.json). */
int gra_hip_cfg_default(gra_cfg* cfg, int32_t preset);

/* Create a handle: uploads parameters, builds the eikonal tables on the GPU if
 * screening is on (MEikonal::S3Constructor, src/MEikonal.cc:46-132), and sets
 * up the process (MContinuum ctor + post_Constructor, src/MContinuum.cc:43-89). */
int gra_hip_create(const gra_cfg* cfg, gra_handle** out);

/* gra_hip_create with an eikonal table cache directory, as MEikonal::S3Constructor
 * keeps its tables (src/MEikonal.cc:62-93, 98-129): each of MBT and MSA is read
 * from eikonal_dir/<gra_hip_eikonal_cache_name> when that file holds n+1 rows
 * "x,re,im"; otherwise it is computed on the GPU, written there with 15
 * significant digits (IArray1D::WriteArray, src/MEikonal.cc:404-436) and read
 * back (IArray1D::ReadArray, 439-485). A field that is not a number fails with
 * GRA_E_ARG, as the reference's parse throws. Files the reference wrote for the
 * same parameters are read as they are. */
int gra_hip_create_cached(const gra_cfg* cfg, const char* eikonal_dir, gra_handle** out);

/* File name of table 0 (MBT, proton density) or 1 (MSA, screening amplitude)
 * for cfg: "MBT_2212_2212_<sqrt s, %.0f>_<djb2 hash>" (MSA_... likewise), the
 * hash (src/MAux.cc:280-287) taken over the same string of std::to_string
 * values of s, beams, PARAM_SOFT (src/MForm.cc:94-100) and MEikonalNumerics
 * (include/Graniitti/MEikonal.h:59-67) as S3Constructor forms (src/MEikonal.cc:70-79,
 * 107-116). */
int gra_hip_eikonal_cache_name(const gra_cfg* cfg, int32_t table, char* name, size_t len);

/* Host-only table file helpers (no GPU): n_rows rows of (x, re, im) in the
 * format above. Read returns GRA_E_IO for a missing file or one without
 * 3 n_rows values in its first n_rows lines (the reference's "Corrupted file"),
 * GRA_E_ARG for a field that is not a number. */
int gra_eikonal_table_write(const char* path, const double* rows, int32_t n_rows);
int gra_eikonal_table_read(const char* path, double* rows, int32_t n_rows);
/* djb2 string hash (gra::aux::djb2hash, src/MAux.cc:280-287), 64-bit */
uint64_t gra_djb2(const char* s);

/* Release every device resource of the handle. */
void gra_hip_destroy(gra_handle* h);

/* Message of the last failed call on h (or of the last failed create when h is NULL). */
const char* gra_hip_last_error(const gra_handle* h);

/* Batched MContinuum::EventWeight(randvec, aux) (src/MContinuum.cc:164-193;
 * include/Graniitti/MProcess.h:83-84): u is n x dim row-major in [0,1); W[i] is
 * the weight in barn, amp2[i] (nullable) the screened |M|^2, flags[i] (nullable)
 * the AuxIntData bits. Host pointers. */
int gra_hip_eval_points(gra_handle* h, const double* u, size_t n, double* W, double* amp2,
                        uint8_t* flags);

/* Same as gra_hip_eval_points with device pointers, on the handle's stream. */
int gra_hip_eval_points_device(gra_handle* h, const double* d_u, size_t n, double* d_W,
                               double* d_amp2, uint8_t* d_flags);

/* Size of the per-device scratch array of gra_hip_eval_points_noalloc (64 MiB):
 * every GPU the library runs on holds it from the first handle creation on that
 * device until the process ends. */
#define GRA_NOALLOC_SCRATCH_BYTES (64 << 20)

/* gra_hip_eval_points_device without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Intermediate buffers (the
 * candidate counter and records of the screening stage, per-point bin bytes) live
 * in one static per-device array of GRA_NOALLOC_SCRATCH_BYTES, so n is limited to
 * what fits there for the process (about 2.8e5 points for the screened FP32
 * continuum); larger n fails with GRA_E_ARG. Inputs and outputs are the caller's
 * device buffers, as for gra_hip_eval_points_device; the call runs on the handle's
 * stream and returns when it is done. Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (several handles or threads) race. */
int gra_hip_eval_points_noalloc(gra_handle* h, const double* d_u, size_t n, double* d_W,
                                double* d_amp2, uint8_t* d_flags);

/* Initialize a uniform grid (VEGASData::Init(0) + InitGridDependent,
 * include/Graniitti/MVEGAS.h:38-84). */
int gra_hip_grid_init(gra_grid* grid, int32_t dim, int32_t bins);

/* Grid update from one iteration's histograms (VEGASData::OptimizeGrid,
 * include/Graniitti/MVEGAS.h:107-142). f2mat is consumed (smoothed in place). */
int gra_hip_grid_optimize(gra_grid* grid, gra_iter_stats* st, double lambda);

/* One VEGAS iteration on the GPU: n_calls points drawn with Philox4x32-10 keyed
 * by (seed, iter, point), mapped through the grid, weighted and accumulated
 * (MGraniitti::VEGASMultiThread, src/MGraniitti.cc:1341-1443). */
int gra_hip_vegas_iterate(gra_handle* h, const gra_grid* grid, uint64_t n_calls, uint64_t seed,
                          uint32_t iter, gra_iter_stats* out);

/* Full VEGAS integration (MGraniitti::SampleVegas + VEGAS, src/MGraniitti.cc:1074-1338):
 * the schedule and guards of gra_vegas_run_host over gra_hip_vegas_iterate, with a
 * fixed clock-independent call count. grid is in/out. */
int gra_hip_vegas_run(gra_handle* h, const gra_vegas_cfg* vc, gra_grid* grid,
                      gra_vegas_result* res);

/* One VEGAS pass over an integrand: n_calls points of iteration iter drawn through
 * grid, with their accumulators written to out (the role of
 * MGraniitti::VEGASMultiThread, src/MGraniitti.cc:1341-1443). Returns GRA_OK or a
 * negative status, which ends the run with that status. */
typedef int (*gra_vegas_pass_fn)(void* ctx, const gra_grid* grid, uint64_t n_calls,
                                 uint64_t seed, uint32_t iter, gra_iter_stats* out);

/* The schedule of gra_hip_vegas_run over a caller's pass, host only (no GPU):
 * MGraniitti::SampleVegas + VEGAS in integration mode (src/MGraniitti.cc:1074-1338)
 * with the reference's guards after every iteration but the first of a pass
 * sequence: an integral that is NaN/Inf or below 1e-60 fails with GRA_E_NUMERIC;
 * chi2 > 50 restarts the burn-in from a uniform grid, with 10 x the calls when it
 * happens during burn-in, or with twice the burn-in iterations when it happens
 * while integrating (at most 6 restarts; the reference retries without limit).
 * grid is in/out; dim is the integrand's dimension. */
int gra_vegas_run_host(const gra_vegas_cfg* vc, gra_grid* grid, int32_t dim,
                       gra_vegas_pass_fn pass, void* ctx, gra_vegas_result* res);

/* Points drawn from a grid (no weights): u -> (x, vegas weight, bin) for tests
 * of the map (src/MGraniitti.cc:1357-1375). Host pointers; x is n x dim, vw n,
 * bins n x dim (int32, 1-based as in the reference). */
int gra_hip_vegas_map(gra_handle* h, const gra_grid* grid, const double* u, size_t n, double* x,
                      double* vw, int32_t* bins);

/* Unweighted event generation (generation mode of MGraniitti::VEGASMultiThread
 * with MGraniitti::SaveEvent, src/MGraniitti.cc:1424-1436, 1651-1726): n_calls
 * points of iteration iter are drawn through the grid and weighted; a valid
 * point is accepted when U < f / maxf (U from Philox, its own counter stream)
 * and never when f > maxf (counted as overflow). The first max_events accepted
 * points, in increasing point index, are written to events (host memory; pinned
 * memory from gra_hip_host_alloc is written by the event-record kernel itself,
 * pageable memory through a device buffer and a pinned staging buffer of the
 * handle). Entries past st->stored are left untouched; after an error return the
 * first max_events entries are unspecified. */
int gra_hip_generate(gra_handle* h, const gra_grid* grid, uint64_t n_calls, uint64_t seed,
                     uint32_t iter, double maxf, gra_event* events, uint64_t max_events,
                     gra_gen_stats* st);

/* Pinned (page-locked) host memory for event buffers: gra_hip_generate's kernel
 * writes its events there directly, and gra_hip_flow_generate's event copies run
 * at DMA rate. No counterpart
 * in the reference (its events never leave host memory). gra_hip_host_free
 * releases a buffer of gra_hip_host_alloc (NULL is ignored). */
int gra_hip_host_alloc(size_t bytes, void** p);
void gra_hip_host_free(void* p);

/* Hit-or-miss unweighting of n weights against wmax (MGraniitti::SaveEvent,
 * src/MGraniitti.cc:1651-1663): accept[i] = U_i < w_i / wmax with U from
 * Philox keyed by (seed, i). Returns accepted count and overflow count. */
int gra_hip_unweight(gra_handle* h, const double* w, size_t n, double wmax, uint64_t seed,
                     uint8_t* accept, uint64_t* n_accepted, uint64_t* n_overflow);

/* HepMC3 ASCII event output (host only, no GPU). Replaces the HepMC3 3.2
 * WriterAscii the reference opens in MGraniitti::InitFileOutput
 * (src/MGraniitti.cc:190-207) and fills in SaveEvent (1665-1712), with the event
 * record of MProcess::CommonRecord (src/MProcess.cc:1588-1706) and
 * WriteDecayKinematics (1287-1314): beams, forward protons, the two generic
 * propagators (pdg 99), the central system (pdg 90), the central particles and
 * their decay daughters. */
typedef struct gra_hepmc3 gra_hepmc3;

/* PDG ids of an event record */
typedef struct {
  int32_t beam[2];        /* beam 1 (+z), beam 2 (-z); the forward protons keep these ids */
  int32_t central[2];     /* central particles in decaytree order (gra_event p[2], p[3]) */
  int32_t daughter[2][2]; /* daughters of central[0] (p[4], p[5]) and central[1] (p[6], p[7]) */
} gra_record_pdg;

/* Open path for writing and write the header: format version and the run's
 * tools (generator name and version, steering card; card may be NULL). */
int gra_hepmc3_open(const char* path, const char* generator, const char* version,
                    const char* card, gra_hepmc3** out);
/* Append n events. sigma, sigma_err in barn are written in picobarn as each
 * event's GenCrossSection; the event weight is 1 (unweighted) or, when weighted
 * is nonzero, the event's f. Beam 4-momenta come from cfg (beam_E1, beam_E2,
 * m_beam). Events are numbered consecutively from 0 across calls. */
int gra_hepmc3_write(gra_hepmc3* w, const gra_cfg* cfg, const gra_record_pdg* pdg,
                     const gra_event* events, uint64_t n, double sigma, double sigma_err,
                     int32_t weighted);
/* Write the footer, close the file and free the writer. */
int gra_hepmc3_close(gra_hepmc3* w);

/* Copy the screening-node eikonal amplitudes A_eik(kt_j^2) (complex, interleaved
 * re,im; n_loop_kt+1 values) and the density / amplitude tables (MBT, MSA of
 * src/MEikonal.cc:46-132; x, re, im triplets, n+1 rows) to host. */
int gra_hip_eikonal_tables(gra_handle* h, double* node_amp, double* mbt, double* msa);

/* Create / train the flow sampler in the NEURO integrator slot
 * (MGraniitti::CallIntegrator, src/MGraniitti.cc:1056-1071; SampleNeuro 1527-1648). */
int gra_hip_flow_init(gra_handle* h, const gra_flow_cfg* fc);
/* VEGAS grid the flow composes with (x = grid(u), weight f(x) vw(u) / q(u));
 * uniform after flow_init. */
int gra_hip_flow_set_grid(gra_handle* h, const gra_grid* grid);
int gra_hip_flow_train(gra_handle* h, int32_t steps, gra_flow_stats* last);
/* Sample n points from the flow, evaluate the integrand and report sigma, CV^2,
 * ESS (the per-point loop of SampleNeuro, src/MGraniitti.cc:1586-1622). Host
 * arrays, nullable: w (n) receives each point's weight w_i = f(x_i) vw_i / q_i in
 * barn, x (n x dim, row-major) the point x_i = grid(u_i) the integrand was
 * evaluated at. st->fallback is set when the ESS is below the flow's ess_fallback. */
int gra_hip_flow_sample(gra_handle* h, uint64_t n, uint64_t seed, double* w, double* x,
                        gra_flow_stats* st);

/* The NEURO integrator run (MGraniitti::SampleNeuro integration, src/MGraniitti.cc:
 * 1527-1648) with plain VEGAS as the fallback and as the cross-check of sigma. */
typedef struct gra_flow_run_cfg {
  int32_t train_steps; /* Adam steps before sampling (0: sample the flow as it is), at most */
  int32_t check_every; /* training window in steps for the plateau test (0: no test) */
  uint64_t n_max;      /* sampling budget in points */
  uint64_t seed;       /* sampling stream */
  double precision;    /* stop when sigma_err / sigma <= precision (0: use the budget) */
  double ess_min;      /* VEGAS takes over when the running ESS falls below this */
  gra_vegas_cfg vegas; /* the fallback VEGAS run */
  double plateau;      /* stop training when a window's mean CV^2 of the training batches
                          falls by less than this fraction of the previous window's (0: no
                          test) */
} gra_flow_run_cfg;

typedef struct gra_flow_result {
  double sigma, sigma_err;     /* barn: the flow's plain mean, or VEGAS's after a fallback */
  double cv2, ess, maxw;       /* of the flow's sample (maxw: the largest w) */
  double train_seconds;        /* training part of the call */
  double sample_seconds;       /* flow sampling part of the call */
  double seconds_to_precision; /* from the call's start, training included; -1 if not reached */
  uint64_t n;                  /* flow points sampled */
  int32_t fallback;            /* 1: ESS < ess_min, sigma and sigma_err come from VEGAS */
  int32_t reached;             /* 1: sigma_err / sigma <= precision */
  gra_vegas_result vegas;      /* the fallback run (zero without one) */
  uint64_t train_steps_run;    /* Adam steps taken (fewer than train_steps after a plateau) */
  double train_cv2;            /* mean CV^2 of the last training window (or the last batch) */
} gra_flow_result;

/* Train (rc->train_steps, or fewer when the plateau test stops it), then sample
 * in chunks until the precision or the budget is reached; if the running effective sample size (sum w)^2 / (N sum w^2)
 * drops below rc->ess_min, run gra_hip_vegas_run(rc->vegas) on grid (in/out) and
 * report its result with fallback = 1. */
int gra_hip_flow_integrate(gra_handle* h, const gra_flow_run_cfg* rc, gra_grid* grid,
                           gra_flow_result* res);

/* Unweighted events from the flow (the generation mode of SampleNeuro with
 * MGraniitti::SaveEvent, src/MGraniitti.cc:1624-1632, 1651-1726): n_calls flow
 * points of generation pass `pass` (pass 0 draws the points gra_hip_flow_sample
 * draws with the same seed), each valid point accepted when U < w / maxw (U from
 * Philox, its own stream) and never when w > maxw (overflow). The first
 * max_events accepted points, in increasing point index, are written to events
 * (host memory) with f = w and index = the point index. */
int gra_hip_flow_generate(gra_handle* h, uint64_t n_calls, uint64_t seed, uint32_t pass,
                          double maxw, gra_event* events, uint64_t max_events,
                          gra_gen_stats* st);

/* Parameter access for tests: host copy in/out of all flow parameters (FP32,
 * packed per layer: W1,b1,...,Wout,bout) and of their gradients. */
int gra_hip_flow_param_count(gra_handle* h, uint64_t* count);
int gra_hip_flow_get_params(gra_handle* h, float* params);
int gra_hip_flow_set_params(gra_handle* h, const float* params);
/* Sampling direction on caller points: z (n x dim, n a multiple of 64) ->
 * u (n x dim) in the grid's unit cube, logq (n) = log q(u). Host pointers. */
int gra_hip_flow_forward(gra_handle* h, const float* z, size_t n, float* x, float* logq);
/* Gradient of L = -sum_i c_i log q(u_i) with respect to every parameter, at
 * caller points u (n a multiple of 64) and coefficients c; logq (nullable)
 * receives log q(u). Host pointers. */
int gra_hip_flow_grad(gra_handle* h, const float* x, const float* c, size_t n, float* grad,
                      float* logq);

/* Op count of the FP32 screening node of the Regge continuum (screened PP[CON]
 * in the FP32-amplitude mode). Measurement only: no counterpart in the
 * reference. An op-counting instance of the node code re-runs the loop of the
 * first min(max_points, screened points) records of the handle's last integrand
 * evaluation through its own buffers (any entry point but
 * gra_hip_eval_points_noalloc; max_points <= 65536) and tallies every
 * operation it executes; a
 * rejected node counts what ran before its rejection. out[12]: points, nodes,
 * accepted nodes, add/sub, mul, fma, transcendental (log2, exp2, rcp, sqrt, sin,
 * cos, and the FP64 exp of each point's epilogue), compare; then add, mul, fma,
 * transcendental of one accepted node. FLOP = add + mul + 2 fma + transcendental.
 * The counting launch uses a static device tally: calls that overlap on one GPU
 * race. */
int gra_hip_screen_opcount(gra_handle* h, uint64_t max_points, double* out);

/* The FP32 FMA rate (TFLOP/s, FMA = 2) the current device sustains on
 * independent FMA chains without packed FP32 instructions, which this library
 * does not use: the attainable ceiling of its FP32 vector kernels, measured with
 * HIP events on a private stream (about 10 ms). Measurement only: no
 * counterpart in the reference. */
int gra_hip_fp32_fma_rate(double* tflops);

/* Accumulated duration of the integrand kernels (HIP events on the handle's
 * stream) since creation or the last reset, and the number of launches. */
int gra_hip_kernel_time(gra_handle* h, int32_t which, double* total_ms, uint64_t* launches,
                        int32_t reset);

#ifdef __cplusplus
}
#endif

#endif /* GRA_HIP_H */
