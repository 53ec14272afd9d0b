/*
 * gra_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 CPU restatement of the GRANIITTI (v1.051) event-generation inner
 * loop, used as the parity checker of the HIP library and as the CPU baseline of
 * the benchmark. It is never linked into, loaded by, or called from the product
 * path. Every function cites the reference file:line it follows (paths relative
 * to the reference root). Floating-point operations follow the reference's order
 * and are compiled without FMA contraction (the reference is built for baseline
 * x86-64, Makefile:196-204), so that the FP64 restatement reproduces the
 * reference's rounding wherever libm agrees.
 *
 * Pinning status of each part (what fixes it beyond this text; the checks are
 * in tests/test_oracle_core.py and tests/test_oracle_physics.py):
 *   philox4x32_10          Random123 known-answer vectors
 *   vegas grid/rebin/opt   golden vectors written by the reference's own MVEGAS.h
 *                          compiled from source (oracle/_ref ->
 *                          tests/golden/ref_vectors.json) and a live comparison
 *   vegas map              uniform grid is the identity; exact integration of
 *                          polynomials through an adapted map
 *   simpson 3/8, CS rule,  golden vectors from the reference's MMath.h/extmath.hpp
 *   CheckEMC, BESSJ0       (oracle/_ref) + exactness on cubics + scipy j0
 *   SolvePz                residuals of the conservation equations + mpmath root
 *   kinematics/scalars     4-momentum conservation, on-shell masses
 *   phase-space weight     equals the finite-difference Jacobian of the 2->4 map
 *   Regge ME4              ReggeEtaLinear vs ReggeEta (the reference's testbench3
 *                          5% check), |P| = |eta| (s/s0)^alpha, t<->u symmetry;
 *                          TUNE0 amplitude values themselves: parity unpinned
 *   pomeron trajectory     alpha(0) = 1 + Delta, small-t pion-loop slope (closed form)
 *   eikonal Fourier-Bessel Hankel pair of a Gaussian (closed form); table
 *                          convergence; sigma_tot/sigma_el ballpark (sanity only)
 *   screening loop         node amplitudes equal the table's interpolation; small
 *                          table limit; survival factor in (0.05, 0.5)
 *   resonances (PP[RES+CON], tests/test_oracle_resonance.py)
 *     Wigner d, traces     closed-form d^1, d^2 values and orthogonality; the
 *                          spin-density traces equal sum_j rho_j |d^J_{j lam}(beta)|^2
 *                          (beta: angle between the two axes); unpolarised rho
 *                          gives no angular dependence
 *     Collins-Soper frame  f2 and rho spin factors at generated events equal the
 *                          closed forms in the CS angle computed from the 4-vectors
 *     decay coupling       g^2 p* / (8 pi M^2) = BR Gamma
 *   form factors (tests/test_oracle_pins.py)
 *     Kelly F1             F1(0) = 1; within 40% of the dipole Dirac form factor
 *                          (the reference's testbench3 check)
 *     CBW                  all forms equal i / (M0 Gamma) at the pole; Lorentzian
 *                          fixed form; running/barrier widths; J = 1, 2 numerators
 *     resonance FF         Gaussian in s_hat of width S0^2
 *     CohFlux              its Q^2 -> 0 limits for pt >> xi m_p and pt << xi m_p
 *                          against CODATA alpha and mu_p
 *     PhotoProp            phase of the signature factor; slope B0 + 2 alpha' log s
 *                          with the reference's stated B0 (11 rho, 4.8 J/psi)
 *   CohFlux away from Q^2 = 0 (its Q^2 dependence through Kelly's G_E, G_M):
 *   parity unpinned beyond the pinned F1
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/gra_hip.h"

#define ORA_EXPORT __attribute__((visibility("default")))

static const double PI = 3.141592653589793238462643383279502884197169399375105820974944;

/* ------------------------------------------------------------------------- */
/* Complex numbers with the arithmetic of std::complex<double> (libstdc++)   */
/* ------------------------------------------------------------------------- */
typedef struct {
  double re, im;
} cx;

static inline cx cx_make(double re, double im) {
  cx z = {re, im};
  return z;
}
static inline cx cx_add(cx a, cx b) { return cx_make(a.re + b.re, a.im + b.im); }
static inline cx cx_sub(cx a, cx b) { return cx_make(a.re - b.re, a.im - b.im); }
/* complex * complex, finite case of __muldc3 */
static inline cx cx_mul(cx a, cx b) {
  return cx_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
/* complex * real and real * complex are component-wise in libstdc++ */
static inline cx cx_scale(cx a, double d) { return cx_make(a.re * d, a.im * d); }
static inline cx cx_div_real(cx a, double d) { return cx_make(a.re / d, a.im / d); }
/* std::exp(complex) -> glibc cexp, finite arguments */
static inline cx cx_exp(cx a) {
  const double e = exp(a.re);
  return cx_make(e * cos(a.im), e * sin(a.im));
}
/* gra::math::abs2 = pow2(std::abs(z)), include/Graniitti/MMath.h:253 */
static inline double cx_abs2(cx a) {
  const double h = hypot(a.re, a.im);
  return h * h;
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based RNG      */
/* ------------------------------------------------------------------------- */
ORA_EXPORT void ora_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                  uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words */
static inline double u53(uint32_t hi, uint32_t lo) {
  const uint64_t x = (((uint64_t)hi << 32) | lo) >> 11;
  return (double)x * (1.0 / 9007199254740992.0);
}

/* Uniforms of point `idx` in iteration `iter`: counter = (block, idx_lo, idx_hi, iter),
 * key = seed; each Philox block yields two doubles. */
ORA_EXPORT void ora_point_uniforms(uint64_t seed, uint32_t iter, uint64_t idx, int dim,
                                   double* u) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int b = 0; 2 * b < dim; ++b) {
    const uint32_t ctr[4] = {(uint32_t)b, (uint32_t)idx, (uint32_t)(idx >> 32), iter};
    uint32_t o[4];
    ora_philox4x32_10(ctr, key, o);
    u[2 * b] = u53(o[0], o[1]);
    if (2 * b + 1 < dim) u[2 * b + 1] = u53(o[2], o[3]);
  }
}

/* ------------------------------------------------------------------------- */
/* VEGAS grid (include/Graniitti/MVEGAS.h)                                   */
/* ------------------------------------------------------------------------- */
#define XM(g, i, j) ((g)->xmat[(j)*GRA_VEGAS_BINS + (i)])

/* VEGASData::Rebin, include/Graniitti/MVEGAS.h:87-104 */
static void vg_rebin(gra_grid* g, double ac, int j, const double* rvec) {
  const int BINS = g->bins;
  unsigned int k = 0;
  double dr = 0.0, zn = 0.0, zo = 0.0;
  double xcache[GRA_VEGAS_BINS];
  for (int i = 0; i < BINS - 1; ++i) {
    while (ac > dr) { dr += rvec[(++k) - 1]; }
    if (k > 1) { zo = XM(g, k - 2, j); }
    zn = XM(g, k - 1, j);
    dr -= ac;
    xcache[i] = zn - (zn - zo) * dr / rvec[k - 1];
  }
  for (int i = 0; i < BINS - 1; ++i) { XM(g, i, j) = xcache[i]; }
  XM(g, BINS - 1, j) = 1.0;
}

/* VEGASData::ClearAll + InitGridDependent with BINS_prev = 1,
 * include/Graniitti/MVEGAS.h:72-84, 163-187 */
ORA_EXPORT int ora_grid_init(gra_grid* g, int dim, int bins) {
  if (dim < 1 || dim > GRA_MAX_DIM || bins < 2 || bins > GRA_VEGAS_BINS || (bins % 2)) return -1;
  memset(g, 0, sizeof(*g));
  g->dim = dim;
  g->bins = bins;
  for (int j = 0; j < dim; ++j) XM(g, 0, j) = 1.0;
  double rvec[GRA_VEGAS_BINS];
  for (int i = 0; i < GRA_VEGAS_BINS; ++i) rvec[i] = 1.0;
  for (int j = 0; j < dim; ++j) vg_rebin(g, 1.0 / (double)bins, j, rvec);
  return 0;
}

/* VEGASData::OptimizeGrid, include/Graniitti/MVEGAS.h:107-142 (f2mat consumed) */
ORA_EXPORT void ora_grid_optimize(gra_grid* g, double* f2mat, double lambda) {
  const int BINS = g->bins;
  const double EPS = 1.0e-30;
  double dcache[GRA_MAX_DIM];
  double rvec[GRA_VEGAS_BINS];
#define F2(i, j) f2mat[(j)*GRA_VEGAS_BINS + (i)]
  for (int j = 0; j < g->dim; ++j) {
    double zo = F2(0, j), zn = F2(1, j), ac;
    F2(0, j) = (zo + zn) / 2.0;
    dcache[j] = F2(0, j);
    for (int i = 2; i < BINS; ++i) {
      ac = zo + zn;
      zo = zn;
      zn = F2(i, j);
      F2(i - 1, j) = (ac + zn) / 3.0;
      dcache[j] += F2(i - 1, j);
    }
    F2(BINS - 1, j) = (zo + zn) / 2.0;
    dcache[j] += F2(BINS - 1, j);
  }
  for (int j = 0; j < g->dim; ++j) {
    double ac = 0.0;
    for (int i = 0; i < BINS; ++i) {
      F2(i, j) = (F2(i, j) < EPS) ? EPS : F2(i, j);
      rvec[i] = pow((1.0 - F2(i, j) / dcache[j]) / (log(dcache[j]) - log(F2(i, j)) + EPS), lambda);
      if (isnan(rvec[i]) || isinf(rvec[i])) { rvec[i] = EPS; }
      ac += rvec[i];
    }
    vg_rebin(g, ac / BINS, j, rvec);
  }
#undef F2
}

/* Map of one point through the grid, MGraniitti::VEGASMultiThread,
 * src/MGraniitti.cc:1355-1375. bins are 1-based like indvec. */
ORA_EXPORT void ora_vegas_map(const gra_grid* g, const double* u, double* x, double* vw_out,
                              int32_t* bin) {
  const unsigned int BINS = (unsigned int)g->bins;
  double vw = 1.0;
  for (int j = 0; j < g->dim; ++j) {
    const double zn = u[j] * BINS + 1.0;
    unsigned int k = (unsigned int)zn;
    if (k > BINS) k = BINS;
    if (k < 1) k = 1;
    double zo, ac;
    if (k > 1) {
      zo = XM(g, k - 1, j) - XM(g, k - 2, j);
      ac = XM(g, k - 2, j) + (zn - k) * zo;
    } else {
      zo = XM(g, k - 1, j);
      ac = (zn - k) * zo;
    }
    x[j] = 0.0 + ac * 1.0; /* region[j] + ac * dxvec[j] on [0,1]^D */
    vw *= zo * BINS;
    bin[j] = (int32_t)k;
  }
  *vw_out = vw;
}

/* ------------------------------------------------------------------------- */
/* Lorentz 4-vectors, include/Graniitti/M4Vec.h (storage order E, x, y, z)   */
/* ------------------------------------------------------------------------- */
typedef struct {
  double e, x, y, z;
} v4;

static inline double msqrt(double x) { return sqrt(x > 0.0 ? x : 0.0); }
static inline v4 v4_make(double x, double y, double z, double e) {
  v4 v = {e, x, y, z};
  return v;
}
static inline v4 v4_add(v4 a, v4 b) { return v4_make(a.x + b.x, a.y + b.y, a.z + b.z, a.e + b.e); }
static inline v4 v4_sub(v4 a, v4 b) { return v4_make(a.x - b.x, a.y - b.y, a.z - b.z, a.e - b.e); }
static inline v4 v4_scale(v4 a, double d) { return v4_make(a.x * d, a.y * d, a.z * d, a.e * d); }
static inline double v4_p3mod2(v4 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double v4_m2(v4 a) { return a.e * a.e - v4_p3mod2(a); }
static inline double v4_m(v4 a) {
  const double m2 = v4_m2(a);
  return (m2 > 0.0) ? msqrt(m2) : -msqrt(-m2);
}
static inline double v4_pt2(v4 a) { return a.x * a.x + a.y * a.y; }
static inline double v4_pt(v4 a) { return msqrt(v4_pt2(a)); }
static inline double v4_p3mod(v4 a) { return msqrt(v4_p3mod2(a)); }
static inline double v4_mt(v4 a) { return msqrt(v4_m2(a) + v4_pt2(a)); }
static inline double v4_eta(v4 a) {
  return 0.5 * log((v4_p3mod(a) + a.z) / (v4_p3mod(a) - a.z));
}
static inline double v4_rap(v4 a) { return 0.5 * log((a.e + a.z) / (a.e - a.z)); }
static inline double v4_dot(v4 a, v4 b) { return a.e * b.e - (a.x * b.x + a.y * b.y + a.z * b.z); }

/* gra::math::CheckEMC, include/Graniitti/MMath.h:191-198 */
static inline int check_emc(v4 d) {
  const double eps = 1e-6;
  if (fabs(d.x) > eps || fabs(d.y) > eps || fabs(d.z) > eps || d.e > eps) return 0;
  return 1;
}
ORA_EXPORT int ora_check_emc(const double d[4]) {
  v4 v = {d[0], d[1], d[2], d[3]};
  return check_emc(v);
}

/* kinematics::SolvePz3_A (m3 = m4), include/Graniitti/MKinematics.h:33-64 */
static double solvepz_a(double m3, double pt3, double pt4, double pz5, double E5, double s) {
  const double a2 = E5 * E5, a3 = pt3 * pt3, a4 = pt4 * pt4, a5 = pz5 * pz5, a6 = m3 * m3;
  const double rs = sqrt(s);
  const double c8 = s * a2 * 6.0;
  const double s32 = pow(s, 3.0 / 2.0);
  const double a10 = a2 * a2, a11 = a3 * a3, a12 = a4 * a4, a13 = a5 * a5, a14 = s * s;
  const double c15 = a5 * a6 * 4.0, c16 = a3 * a5 * 2.0, c17 = a4 * a5 * 2.0;
  const double c18 = E5 * a6 * rs * 8.0, c19 = E5 * a3 * rs * 4.0, c20 = E5 * a4 * rs * 4.0;
  const double c21 = E5 * a5 * rs * 4.0;
  const double disc = c8 + a10 + a11 + a12 + a13 + a14 + c15 + c16 + c17 + c18 + c19 + c20 + c21 -
                      E5 * s32 * 4.0 - s * a3 * 2.0 - s * a4 * 2.0 - s * a5 * 2.0 - s * a6 * 4.0 -
                      a2 * a3 * 2.0 - a2 * a4 * 2.0 - a2 * a5 * 2.0 - a3 * a4 * 2.0 -
                      a2 * a6 * 4.0 - E5 * a2 * rs * 4.0;
  const double sq = sqrt(disc);
  return pz5 * (-1.0 / 2.0) -
         (E5 * sq * (1.0 / 2.0) - rs * sq * (1.0 / 2.0) + pz5 * (a3 * (1.0 / 2.0) - a4 * (1.0 / 2.0))) /
             (s + a2 - a5 - E5 * rs * 2.0);
}

/* kinematics::SolvePz3_B (general masses), include/Graniitti/MKinematics.h:117-159 */
static double solvepz_b(double m3, double m4, double pt3, double pt4, double pz5, double E5,
                        double s) {
  const double a2 = E5 * E5, a3 = m3 * m3, a4 = m4 * m4, a5 = pt3 * pt3, a6 = pt4 * pt4,
               a7 = pz5 * pz5;
  const double rs = sqrt(s);
  const double c9 = s * a2 * 6.0;
  const double s32 = pow(s, 3.0 / 2.0);
  const double a11 = a2 * a2, a12 = a3 * a3, a13 = a4 * a4, a14 = a5 * a5, a15 = a6 * a6,
               a16 = a7 * a7, a17 = s * s;
  const double c18 = a3 * a5 * 2.0, c19 = a4 * a6 * 2.0, c20 = a3 * a7 * 2.0, c21 = a4 * a7 * 2.0,
               c22 = a5 * a7 * 2.0, c23 = a6 * a7 * 2.0;
  const double c24 = E5 * a3 * rs * 4.0, c25 = E5 * a4 * rs * 4.0, c26 = E5 * a5 * rs * 4.0,
               c27 = E5 * a6 * rs * 4.0, c28 = E5 * a7 * rs * 4.0;
  const double disc = c9 + a11 + a12 + a13 + a14 + a15 + a16 + a17 + c18 + c19 + c20 + c21 + c22 +
                      c23 + c24 + c25 + c26 + c27 + c28 - E5 * s32 * 4.0 - s * a3 * 2.0 -
                      s * a4 * 2.0 - s * a5 * 2.0 - s * a6 * 2.0 - s * a7 * 2.0 - a2 * a3 * 2.0 -
                      a2 * a4 * 2.0 - a2 * a5 * 2.0 - a3 * a4 * 2.0 - a2 * a6 * 2.0 -
                      a2 * a7 * 2.0 - a3 * a6 * 2.0 - a4 * a5 * 2.0 - a5 * a6 * 2.0 -
                      E5 * a2 * rs * 4.0;
  const double sq = sqrt(disc);
  return pz5 * (-1.0 / 2.0) -
         (E5 * sq * (1.0 / 2.0) - rs * sq * (1.0 / 2.0) +
          pz5 * (a3 * (1.0 / 2.0) - a4 * (1.0 / 2.0) + a5 * (1.0 / 2.0) - a6 * (1.0 / 2.0))) /
             (s + a2 - a7 - E5 * rs * 2.0);
}

/* kinematics::SolvePz, include/Graniitti/MKinematics.h:220-229 */
ORA_EXPORT double ora_solvepz(double m3, double m4, double pt3, double pt4, double pz5, double E5,
                              double s) {
  if (fabs(m3 - m4) < 1e-9) return solvepz_a(m3, pt3, pt4, pz5, E5, s);
  return solvepz_b(m3, m4, pt3, pt4, pz5, E5, s);
}

/* ------------------------------------------------------------------------- */
/* Run constants derived from the configuration                              */
/* ------------------------------------------------------------------------- */
typedef struct {
  gra_cfg c;
  v4 pbeam1, pbeam2, beamsum;
  double s, sqrt_s;
  double GeV2barn;
  /* screening node tables */
  int nphi, nkt;
  double step_kt, step_phi;
  /* eikonal tables */
  int n_bt, n_kt2, n_fb;
  double mbt_min, mbt_max, mbt_step, msa_min, msa_max, msa_step;
  double *mbt_x, *msa_x; /* n+1 */
  cx *mbt_y, *msa_y;
  int tables_ready;
  /* <F>: system mass range, src/MFactorized.cc:259-285 */
  double f_Mmin, f_Mmax;
  /* Gauss loop rule: radial nodes and weights on [MinLoopKT, MaxLoopKT]; the
   * screening amplitude at log-spaced kt nodes (eik_log_nodes > 0) */
  double *g_kt, *g_w;
  double* l_lkt; /* log kt of the log-spaced nodes */
  cx* l_amp;
} ora_run;

/* Gauss-Legendre nodes x and weights w on [-1, 1]: roots of P_n by Newton's
 * iteration from cos(pi (i + 3/4) / (n + 1/2)), w = 2 / ((1 - x^2) P_n'(x)^2) */
ORA_EXPORT void ora_gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = cos(PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, pm = 0.0;
      for (int k = 1; k <= n; ++k) {
        const double pk = ((2.0 * k - 1.0) * z * p0 - (k - 1.0) * pm) / k;
        pm = p0;
        p0 = pk;
      }
      dp = n * (z * p0 - pm) / (z * z - 1.0);
      const double dz = p0 / dp;
      z -= dz;
      if (fabs(dz) < 1e-15) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

/* MParticle masses and units, include/Graniitti/MPDG.h:54-78 */
static double gev2barn(void) {
  const double c = 2.99792458E8, hbar = 6.582119514E-25;
  const double GeV2m = hbar * c;
  const double GeV2m2 = GeV2m * GeV2m;
  return GeV2m2 / 1E-28;
}
static double gev2fm(void) {
  const double c = 2.99792458E8, hbar = 6.582119514E-25;
  return hbar * c * 1E15;
}

/* ------------------------------------------------------------------------- */
/* Form factors and Regge factors (src/MForm.cc, src/MRegge.cc)              */
/* ------------------------------------------------------------------------- */
/* form::ReggeEta, src/MForm.cc:382-385 */
static cx regge_eta(double alpha_t, double sigma) {
  const double denom = sin(PI * alpha_t);
  /* -zi * PI * alpha_t = (-0, -PI*alpha_t) */
  const cx e = cx_exp(cx_make(-0.0 * PI * alpha_t, -PI * alpha_t));
  const cx num = cx_make(1.0 + sigma * e.re, sigma * e.im);
  return cx_div_real(cx_make(-num.re, -num.im), denom);
}
/* form::ReggeEtaLinear, src/MForm.cc:394-396 */
static cx regge_eta_linear(double t, double a0, double ap, double sigma) {
  /* -zi * PI / 2.0 * ap * t */
  const double th = ((-PI / 2.0) * ap) * t;
  const cx ph = cx_exp(cx_make(((-0.0 * PI / 2.0) * ap) * t, th));
  return cx_mul(regge_eta(a0, sigma), ph);
}
ORA_EXPORT void ora_regge_eta(double alpha_t, double sigma, double* out) {
  const cx z = regge_eta(alpha_t, sigma);
  out[0] = z.re;
  out[1] = z.im;
}
ORA_EXPORT void ora_regge_eta_linear(double t, double a0, double ap, double sigma, double* out) {
  const cx z = regge_eta_linear(t, a0, ap, sigma);
  out[0] = z.re;
  out[1] = z.im;
}

/* form::S3F, src/MForm.cc:433-435 */
static double s3f(const gra_cfg* c, double t) {
  return (1.0 / (1 - t / c->fc1)) * (1.0 / (1 - t / c->fc2));
}
/* form::S3HPL, src/MForm.cc:416-426 */
static double s3hpl(const gra_cfg* c, double tau, double t) {
  const double m = 1.0;
  const double sqrtau = msqrt(1 + tau);
  const double F_pi = 1.0 / (1.0 - t / c->fc3);
  const double mpi = c->mpi_const;
  return (4.0 / tau) * (F_pi * F_pi) *
         (2.0 * tau - pow(1.0 + tau, 3.0 / 2.0) * log((sqrtau + 1.0) / (sqrtau - 1.0)) +
          log((m * m) / (mpi * mpi)));
}
/* form::S3PomAlpha, src/MForm.cc:402-413 */
ORA_EXPORT double ora_pom_alpha(const gra_cfg* c, double t) {
  const double mpi = c->mpi_const;
  const double BETApi = (2.0 / 3.0) * c->gN_P;
  const double PIC = ((BETApi * BETApi) * (mpi * mpi)) / (32.0 * (PI * (PI * PI)));
  const double h = PIC * s3hpl(c, 4.0 * (mpi * mpi) / fabs(t), t);
  return 1.0 + c->delta_P + c->alpha_P * t - h;
}

/* MRegge::PropOnly, src/MRegge.cc:888-906 */
static cx regge_prop(const gra_cfg* c, double s, double t) {
  cx M = cx_make(0, 0);
  for (int k = 0; k < 3; ++k) {
    if (c->regge_on[k]) {
      const double alpha = c->regge_a0[k] + c->regge_ap[k] * t;
      const cx eta = regge_eta_linear(t, c->regge_a0[k], c->regge_ap[k], c->regge_sgn[k]);
      M = cx_add(M, cx_scale(eta, pow(s / c->regge_s0, alpha)));
    }
  }
  return M;
}
ORA_EXPORT void ora_regge_prop(const gra_cfg* c, double s, double t, double* out) {
  const cx z = regge_prop(c, s, t);
  out[0] = z.re;
  out[1] = z.im;
}
/* PARAM_REGGE::Meson_FF / Meson_prop, src/MRegge.cc:1155-1207 */
static double meson_ff(const gra_cfg* c, double t_hat, double M2) {
  const double tprime = t_hat - M2;
  switch (c->offshellFF) {
    case 0: return exp(-c->b_EXP * fabs(tprime));
    case 1:
      return exp(-c->b_OREAR * msqrt(fabs(tprime) + c->a_OREAR * c->a_OREAR) +
                 c->a_OREAR * c->b_OREAR);
    default: return 1.0 / (1.0 - tprime / c->b_POW);
  }
}
static double meson_prop(double t_hat, double M2) { return 1.0 / (t_hat - M2); }

/* ------------------------------------------------------------------------- */
/* Kinematic state (LORENTZSCALAR subset, include/Graniitti/MKinematics.h)   */
/* ------------------------------------------------------------------------- */
typedef struct {
  v4 pf[5];      /* pfinal: 0 system, 1,2 forward, 3,4 central */
  v4 pf_orig[5]; /* pfinal_orig */
  v4 dt[2];      /* decaytree[i].p4 (tree level) */
  v4 legs[2][2]; /* decay daughters of decaytree[i] (rho -> pi pi) */
  v4 pkt;        /* pkt_[0] */
  v4 q1, q2;
  double ss[5][5];
  double s1, s2, t1, t2, t_hat, u_hat, m2, Y, Pt;
  double m_off[2];
  double dw; /* <F>: lts.DW, the two-body phase-space weight */
} kin_t;

/* MProcess::GetLorentzScalars(Nf=4), src/MProcess.cc:1499-1584 (fields used by
 * the amplitudes and the cuts) */
static int lorentz_scalars(const ora_run* R, kin_t* K) {
  for (int i = 0; i <= 4; ++i) {
    for (int j = 0; j <= 4; ++j) {
      if (i < j) K->ss[i][j] = v4_m2(v4_add(K->pf[i], K->pf[j]));
      if (K->ss[i][j] < 0) return 0;
    }
  }
  for (int i = 0; i <= 4; ++i)
    for (int j = i; j <= 4; ++j)
      if (i != j) K->ss[j][i] = K->ss[i][j];
  K->s1 = K->ss[1][0];
  K->s2 = K->ss[2][0];
  if (K->s1 < 0 || K->s2 < 0) return 0;
  K->q1 = v4_sub(R->pbeam1, K->pf[1]);
  K->q2 = v4_sub(R->pbeam2, K->pf[2]);
  K->t1 = v4_m2(K->q1);
  K->t2 = v4_m2(K->q2);
  if (K->t1 > 0 || K->t2 > 0) return 0;
  K->t_hat = v4_m2(v4_sub(K->q1, K->dt[0]));
  K->u_hat = v4_m2(v4_sub(K->q1, K->dt[1]));
  K->m2 = v4_m2(K->pf[0]);
  K->Y = v4_rap(K->pf[0]);
  K->Pt = v4_pt(K->pf[0]);
  return 1;
}

/* MContinuum::BLinearSystem for Kf = 2, src/MContinuum.cc:438-550:
 * A = [[1/2, 0], [0, 1/2]], b0 = q0 - w, b1 = -q0 - w, w = p1f + p2f */
static void blinear2(v4 p[2], v4 q, v4 p1f, v4 p2f) {
  const v4 w = v4_add(p1f, p2f);
  const v4 b0 = v4_sub(q, w);
  const v4 b1 = v4_sub(v4_scale(q, -1.0), w);
  const double A[2][2] = {{1.0 / 2.0, 0.0}, {0.0, 1.0 / 2.0}};
  for (int i = 0; i < 2; ++i) {
    v4 acc = v4_make(0, 0, 0, 0);
    acc = v4_add(acc, v4_scale(b0, A[i][0]));
    acc = v4_add(acc, v4_scale(b1, A[i][1]));
    p[i] = acc;
  }
}

/* kinematics::LorentzBoost(mother, M0, p, sign=+1) for the rho -> pi pi decay,
 * include/Graniitti/MKinematics.h:384-401 */
static v4 lorentz_boost(v4 boost, double M0, v4 p) {
  const double bx = boost.x / boost.e, by = boost.y / boost.e, bz = boost.z / boost.e;
  const double gamma = boost.e / M0;
  const double kappa1 = bx * p.x + by * p.y + bz * p.z;
  const double kappa2 = gamma * (gamma * kappa1 / (1.0 + gamma) + p.e);
  return v4_make(p.x + kappa2 * bx, p.y + kappa2 * by, p.z + kappa2 * bz, gamma * (p.e + kappa1));
}

/* MContinuum::BNRandomKin + BNBuildKin for Kf = 2 central particles,
 * src/MContinuum.cc:276-435. Hidden variables of the rho decays (config 3) are
 * taken from x[8..13] (see the tensor section). Returns kinematics_ok. */
static int build_kin_c(const ora_run* R, const double* x, kin_t* K) {
  const gra_cfg* c = &R->c;
  const double ZERO_EPS = 1e-12;
  const double lmin = log(c->fwd_pt_min + ZERO_EPS), lmax = log(c->fwd_pt_max);
  const double u1 = lmin + (lmax - lmin) * x[0];
  const double u2 = lmin + (lmax - lmin) * x[1];
  const double pt1 = exp(u1), pt2 = exp(u2);
  const double phi1 = 2.0 * PI * x[2], phi2 = 2.0 * PI * x[3];

  /* GetOffShellMass, src/MProcess.cc:753-812, with the BW draw lifted to x[8], x[9]
   * (RelativisticBWRandom, src/MRandom.cc:82-104) */
  for (int i = 0; i < 2; ++i) {
    if (c->w_central <= 0.0) {
      K->m_off[i] = c->m_central;
    } else {
      const double m0 = c->m_central, G = c->w_central, L = c->offshell;
      const double dmin = 2.0 * c->m_daughter + 1e-5; /* daughters at pole + safe margin */
      const double m2 = m0 * m0;
      const double m2max = (m0 + L * G) * (m0 + L * G);
      double lo = dmin * dmin;
      if (lo < 0.0) lo = 0.0;
      const double m2min = fmax(lo, (m0 - L * G) * (m0 - L * G));
      const double a1 = atan2(m2min - m2, m0 * G), a2 = atan2(m2max - m2, m0 * G);
      const double m2val = m2 + m0 * G * tan(a1 + x[8 + i] * (a2 - a1));
      K->m_off[i] = fmax(dmin, msqrt(m2val));
    }
  }
  const double kt = c->kt_min + (c->kt_max - c->kt_min) * x[4];
  const double phi = 2.0 * PI * x[5];
  const double y[2] = {c->rap_min + (c->rap_max - c->rap_min) * x[6],
                       c->rap_min + (c->rap_max - c->rap_min) * x[7]};
  const double m1 = c->m_beam, m2 = c->m_beam;

  v4 p1 = v4_make(pt1 * cos(phi1), pt1 * sin(phi1), 0, 0);
  v4 p2 = v4_make(pt2 * cos(phi2), pt2 * sin(phi2), 0, 0);
  K->pkt = v4_make(kt * cos(phi), kt * sin(phi), 0, 0);
  v4 p[2];
  blinear2(p, K->pkt, p1, p2);
  v4 sumP = v4_make(0, 0, 0, 0);
  for (int i = 0; i < 2; ++i) {
    const double m = K->m_off[i];
    const double mt = msqrt(m * m + v4_pt2(p[i]));
    p[i].z = mt * sinh(y[i]);
    p[i].e = mt * cosh(y[i]);
    sumP = v4_add(sumP, p[i]);
  }
  if (sumP.e > R->sqrt_s) return 0;
  const double p1z = ora_solvepz(m1, m2, pt1, pt2, sumP.z, sumP.e, R->s);
  const double p2z = -(sumP.z + p1z);
  if (p1z < 0 || p2z > 0) return 0;
  p1.z = p1z;
  p1.e = msqrt(m1 * m1 + pt1 * pt1 + p1z * p1z);
  p2.z = p2z;
  p2.e = msqrt(m2 * m2 + pt2 * pt2 + p2z * p2z);
  /* asymmetric beams: the centre-of-mass momenta boosted to the lab
   * (src/MContinuum.cc:386-395, kinematics::LorentzBoost with sign +1) */
  if (fabs(R->beamsum.z) > 1e-6) {
    p1 = lorentz_boost(R->beamsum, R->sqrt_s, p1);
    p2 = lorentz_boost(R->beamsum, R->sqrt_s, p2);
    sumP = lorentz_boost(R->beamsum, R->sqrt_s, sumP);
    for (int i = 0; i < 2; ++i) p[i] = lorentz_boost(R->beamsum, R->sqrt_s, p[i]);
  }
  K->pf[1] = p1;
  K->pf[2] = p2;
  K->pf[0] = sumP;
  double sumM = 0;
  for (int i = 0; i < 2; ++i) {
    K->dt[i] = p[i];
    K->pf[3 + i] = p[i];
    sumM += v4_m(p[i]);
  }
  if (v4_m(sumP) < sumM) return 0;
  if (!check_emc(v4_sub(R->beamsum, v4_add(v4_add(K->pf[1], K->pf[2]), K->pf[0])))) return 0;

  /* ConstructDecayKinematics, src/MProcess.cc:1231-1283: rho -> pi+ pi-,
   * daughters at pole mass, isotropic in the rest frame
   * (TwoBodyPhaseSpace, include/Graniitti/MKinematics.h:405-459), angles lifted to
   * x[10..13] = (cos theta_1, phi_1, cos theta_2, phi_2) */
  if (c->m_daughter > 0.0) {
    for (int i = 0; i < 2; ++i) {
      const double M0 = v4_m(K->dt[i]);
      const double md = c->m_daughter;
      if (M0 < 2.0 * md) return 0;
      const double pn = msqrt((M0 * M0 - (md + md) * (md + md)) * (M0 * M0 - (md - md) * (md - md))) /
                        (2.0 * M0);
      const double ct = 2.0 * x[10 + 2 * i] - 1.0;
      const double ph = 2.0 * PI * x[11 + 2 * i];
      const double st = msqrt(1.0 - ct * ct);
      const double px = pn * st * cos(ph), py = pn * st * sin(ph), pz = pn * ct;
      const double ed = msqrt(md * md + pn * pn);
      K->legs[i][0] = lorentz_boost(K->dt[i], M0, v4_make(px, py, pz, ed));
      K->legs[i][1] = lorentz_boost(K->dt[i], M0, v4_make(-px, -py, -pz, ed));
    }
  }
  return lorentz_scalars(R, K);
}

/* MFactorized::B51RandomKin + B51BuildKin, src/MFactorized.cc:243-393, system
 * decaying to two stable particles (pions at their pole mass). x[0..4] as there,
 * x[5] the system's M^2 uniform on [M_MIN^2, M_MAX^2] (268); TwoBodyPhaseSpace
 * (include/Graniitti/MKinematics.h:443-459) with its FlatIsotropic draws lifted:
 * cos theta = U(-1, 1) from x[6], phi = U(0, 2 pi) from x[7] (405-409;
 * MRandom::U(a, b) = a + (b - a) u, src/MRandom.cc:61-64); DecayMomentum 294-297;
 * dPhi2 267-271. */
static int build_kin_f(const ora_run* R, const double* x, kin_t* K) {
  const gra_cfg* c = &R->c;
  const double ZERO_EPS = 1e-12;
  const double lmin = log(c->fwd_pt_min + ZERO_EPS), lmax = log(c->fwd_pt_max);
  const double u1 = lmin + (lmax - lmin) * x[0];
  const double u2 = lmin + (lmax - lmin) * x[1];
  const double pt1 = exp(u1), pt2 = exp(u2);
  const double phi1 = 2.0 * PI * x[2], phi2 = 2.0 * PI * x[3];
  const double yX = c->rap_min + (c->rap_max - c->rap_min) * x[4];
  const double Mn = R->f_Mmin, Mx = R->f_Mmax;
  const double m2X = Mn * Mn + (Mx * Mx - Mn * Mn) * x[5];
  const double m1 = c->m_beam, m2 = c->m_beam;
  v4 p1 = v4_make(pt1 * cos(phi1), pt1 * sin(phi1), 0, 0);
  v4 p2 = v4_make(pt2 * cos(phi2), pt2 * sin(phi2), 0, 0);
  v4 pX = v4_make(-(p1.x + p2.x), -(p1.y + p2.y), 0, 0);
  const double mtX = msqrt(m2X + v4_pt2(pX));
  pX.z = mtX * sinh(yX);
  pX.e = mtX * cosh(yX);
  if (pX.e > R->sqrt_s - (m1 + m2)) return 0;
  const double p1z = ora_solvepz(m1, m2, pt1, pt2, pX.z, pX.e, R->s);
  const double p2z = -(pX.z + p1z);
  if (p1z < 0 || p2z > 0) return 0;
  p1.z = p1z;
  p1.e = msqrt(m1 * m1 + pt1 * pt1 + p1z * p1z);
  p2.z = p2z;
  p2.e = msqrt(m2 * m2 + pt2 * pt2 + p2z * p2z);
  K->pf[1] = p1;
  K->pf[2] = p2;
  K->pf[0] = pX;
  if (!check_emc(v4_sub(R->beamsum, v4_add(v4_add(K->pf[1], K->pf[2]), K->pf[0])))) return 0;
  const double M0 = msqrt(m2X), ma = c->m_central, mb = c->m_central;
  const double pn = msqrt((M0 - ma - mb) * (M0 + ma + mb) * (M0 + ma - mb) * (M0 - ma + mb)) /
                    (2.0 * M0);
  const double costh = -1.0 + (1.0 - (-1.0)) * x[6];
  const double sinth = msqrt(1.0 - costh * costh);
  const double phi = 0.0 + (2.0 * PI - 0.0) * x[7];
  const double k[3] = {pn * sinth * cos(phi), pn * sinth * sin(phi), pn * costh};
  v4 d0 = v4_make(k[0], k[1], k[2], msqrt(ma * ma + pn * pn));
  v4 d1 = v4_make(-k[0], -k[1], -k[2], msqrt(mb * mb + pn * pn));
  K->dt[0] = lorentz_boost(pX, M0, d0);
  K->dt[1] = lorentz_boost(pX, M0, d1);
  const double V = 2.0 * 2.0 * PI;
  K->dw = V * pn / (16.0 * (PI * PI) * M0);
  if (K->dw < 0) return 0;
  K->m_off[0] = ma;
  K->m_off[1] = mb;
  K->pf[3] = K->dt[0];
  K->pf[4] = K->dt[1];
  return lorentz_scalars(R, K);
}

/* MFactorized::LoopKinematics, src/MFactorized.cc:77-139: the system keeps
 * pfinal_orig[0]'s mass and rapidity with pT = -(p1T + p2T); the decay products
 * stay those of the tree point */
static int loop_kinematics_f(const ora_run* R, kin_t* K, const double p1p[2], const double p2p[2]) {
  K->pf[1].x = p1p[0];
  K->pf[1].y = p1p[1];
  K->pf[2].x = p2p[0];
  K->pf[2].y = p2p[1];
  const double pt1 = v4_pt(K->pf[1]), pt2 = v4_pt(K->pf[2]);
  const double m1 = v4_m(K->pf_orig[1]), m2 = v4_m(K->pf_orig[2]);
  const double mX = v4_m(K->pf_orig[0]), yX = v4_rap(K->pf_orig[0]);
  K->pf[0].x = -(K->pf[1].x + K->pf[2].x);
  K->pf[0].y = -(K->pf[1].y + K->pf[2].y);
  const double mtX = msqrt(mX * mX + v4_pt2(K->pf[0]));
  K->pf[0].z = mtX * sinh(yX);
  K->pf[0].e = mtX * cosh(yX);
  if (K->pf[0].e > R->sqrt_s - (m1 + m2)) return 0;
  const double p1z = ora_solvepz(m1, m2, pt1, pt2, K->pf[0].z, K->pf[0].e, R->s);
  const double p2z = -(K->pf[0].z + p1z);
  if (isnan(p1z)) return 0;
  if (p1z < 0 || p2z > 0) return 0;
  K->pf[1].z = p1z;
  K->pf[1].e = msqrt(m1 * m1 + pt1 * pt1 + p1z * p1z);
  K->pf[2].z = p2z;
  K->pf[2].e = msqrt(m2 * m2 + pt2 * pt2 + p2z * p2z);
  if (!check_emc(v4_sub(R->beamsum, v4_add(v4_add(K->pf[1], K->pf[2]), K->pf[0])))) return 0;
  return lorentz_scalars(R, K);
}

/* The phase-space class's kinematics: <C> (MContinuum) or <F> (MFactorized) */
static int build_kin(const ora_run* R, const double* x, kin_t* K) {
  return R->c.phase_space == GRA_PS_F ? build_kin_f(R, x, K) : build_kin_c(R, x, K);
}

/* lts.DW of the <F> point x (-1 if its kinematics fail) */
ORA_EXPORT double ora_two_body_weight(ora_run* R, const double* x) {
  kin_t K;
  memset(&K, 0, sizeof(K));
  return build_kin_f(R, x, &K) ? K.dw : -1.0;
}

/* Fiducial check of one final-state leg, MProcess::FindDecayCuts, src/MProcess.cc:1138-1154 */
static int leg_cuts(const gra_cfg* c, v4 p) {
  const double pt = v4_pt(p), et = v4_mt(p), eta = v4_eta(p), rap = v4_rap(p);
  return (pt >= c->fid_pt_min && pt <= c->fid_pt_max && et >= c->fid_et_min && et <= c->fid_et_max &&
          eta >= c->fid_eta_min && eta <= c->fid_eta_max && rap >= c->fid_rap_min &&
          rap <= c->fid_rap_max);
}

/* MProcess::VetoCuts and FindVetoCuts, src/MProcess.cc:1069-1082, 1157-1175: the
 * recursion reaches the final states of the central decay tree (the forward
 * decay branches of elastic protons are empty: a zero four-vector, whose Eta()
 * is NaN, triggers no domain). 1 = no final state in any domain. */
static int veto_cuts(const ora_run* R, const kin_t* K) {
  const gra_cfg* c = &R->c;
  if (!c->veto_active) return 1;
  int ok = 1;
  for (int i = 0; i < 2; ++i) {
    const int nl = c->m_daughter > 0.0 ? 2 : 1;
    for (int l = 0; l < nl; ++l) {
      const v4 p = c->m_daughter > 0.0 ? K->legs[i][l] : K->dt[i];
      for (int d = 0; d < c->n_veto; ++d) {
        if (v4_pt(p) >= c->veto_pt_min[d] && v4_pt(p) <= c->veto_pt_max[d] &&
            v4_eta(p) >= c->veto_eta_min[d] && v4_eta(p) <= c->veto_eta_max[d]) {
          ok = 0; /* VETO, does not pass */
        }
      }
    }
  }
  return ok;
}

/* MProcess::CommonCuts, src/MProcess.cc:1085-1135 (USERCUTS = 0, elastic legs) */
static int fiducial_cuts(const ora_run* R, const kin_t* K) {
  const gra_cfg* c = &R->c;
  if (!c->fid_active) return 1;
  if (!(fabs(K->t1) >= c->fwd_t_min && fabs(K->t1) <= c->fwd_t_max && fabs(K->t2) >= c->fwd_t_min &&
        fabs(K->t2) <= c->fwd_t_max))
    return 0;
  const double M = msqrt(K->m2);
  if (!(M >= c->sys_M_min && M <= c->sys_M_max && K->Y >= c->sys_Y_min && K->Y <= c->sys_Y_max &&
        K->Pt >= c->sys_Pt_min && K->Pt <= c->sys_Pt_max))
    return 0;
  int ok = 1;
  for (int i = 0; i < 2; ++i) {
    if (c->m_daughter > 0.0) {
      if (!leg_cuts(c, K->legs[i][0])) ok = 0;
      if (!leg_cuts(c, K->legs[i][1])) ok = 0;
    } else {
      if (!leg_cuts(c, K->dt[i])) ok = 0;
    }
  }
  return ok;
}

/* MContinuum::LoopKinematics for Kf = 2, src/MContinuum.cc:94-161 */
static int loop_kinematics(const ora_run* R, kin_t* K, const double p1p[2], const double p2p[2]) {
  if (R->c.phase_space == GRA_PS_F) return loop_kinematics_f(R, K, p1p, p2p);
  K->pf[1].x = p1p[0];
  K->pf[1].y = p1p[1];
  K->pf[2].x = p2p[0];
  K->pf[2].y = p2p[1];
  v4 p[2];
  blinear2(p, K->pkt, K->pf[1], K->pf[2]);
  v4 sumP = v4_make(0, 0, 0, 0);
  for (int i = 0; i < 2; ++i) {
    const double y = v4_rap(K->pf_orig[i + 3]);
    const double m = v4_m(K->pf_orig[i + 3]);
    K->pf[i + 3].x = p[i].x;
    K->pf[i + 3].y = p[i].y;
    const double mt = msqrt(m * m + v4_pt2(K->pf[i + 3]));
    K->pf[i + 3].z = mt * sinh(y);
    K->pf[i + 3].e = mt * cosh(y);
    sumP = v4_add(sumP, K->pf[i + 3]);
  }
  const double pt1 = v4_pt(K->pf[1]), pt2 = v4_pt(K->pf[2]);
  const double m1 = v4_m(K->pf_orig[1]), m2 = v4_m(K->pf_orig[2]);
  const double p1z = ora_solvepz(m1, m2, pt1, pt2, sumP.z, sumP.e, R->s);
  const double p2z = -(sumP.z + p1z);
  if (p1z < 0 || p2z > 0) return 0;
  K->pf[1].z = p1z;
  K->pf[1].e = msqrt(m1 * m1 + pt1 * pt1 + p1z * p1z);
  K->pf[2].z = p2z;
  K->pf[2].e = msqrt(m2 * m2 + pt2 * pt2 + p2z * p2z);
  K->pf[0] = sumP;
  if (!check_emc(v4_sub(R->beamsum, v4_add(v4_add(K->pf[1], K->pf[2]), K->pf[0])))) return 0;
  return lorentz_scalars(R, K);
}

/* ------------------------------------------------------------------------- */
/* Amplitudes                                                                */
/* ------------------------------------------------------------------------- */
/* MRegge::ME4(lts, sign = +1) with PomPomProtonVertex (elastic legs),
 * src/MRegge.cc:394-438, 637-644 */
static cx regge_me4(const ora_run* R, const kin_t* K) {
  const gra_cfg* c = &R->c;
  const double M2_ = c->m_prop * c->m_prop;
  const double FF_A = c->gN_P * s3f(c, K->t1);
  const double FF_B = c->gN_P * s3f(c, K->t2);
  const double gpp_P = c->regge_c[0] / c->gN_P;
  cx A_t = regge_prop(c, K->ss[1][3], K->t1);
  A_t = cx_scale(A_t, FF_A);
  A_t = cx_scale(A_t, meson_ff(c, K->t_hat, M2_));
  A_t = cx_scale(A_t, gpp_P);
  A_t = cx_scale(A_t, meson_prop(K->t_hat, M2_));
  A_t = cx_scale(A_t, meson_ff(c, K->t_hat, M2_));
  A_t = cx_scale(A_t, gpp_P);
  A_t = cx_mul(A_t, regge_prop(c, K->ss[2][4], K->t2));
  A_t = cx_scale(A_t, FF_B);
  cx A_u = cx_scale(regge_prop(c, K->ss[1][4], K->t1), 1.0);
  A_u = cx_scale(A_u, FF_A);
  A_u = cx_scale(A_u, meson_ff(c, K->u_hat, M2_));
  A_u = cx_scale(A_u, gpp_P);
  A_u = cx_scale(A_u, meson_prop(K->u_hat, M2_));
  A_u = cx_scale(A_u, meson_ff(c, K->u_hat, M2_));
  A_u = cx_scale(A_u, gpp_P);
  A_u = cx_mul(A_u, regge_prop(c, K->ss[2][3], K->t2));
  A_u = cx_scale(A_u, FF_B);
  return cx_add(A_t, A_u);
}

/* ------------------------------------------------------------------------- */
/* Resonances of PP[RES+CON] (PROC_17::Amp2, include/Graniitti/MSubProc.h:398-432) */
/* ------------------------------------------------------------------------- */
static double o_fact(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= (double)i;
  return f;
}
/* spin::WignerSmalld, src/MSpin.cc:927-960 (integer J) */
static double o_wigner_d(double th, double m, double mp, double J) {
  const double factor = msqrt(o_fact((int)(J + m)) * o_fact((int)(J - m)) * o_fact((int)(J + mp)) *
                              o_fact((int)(J - mp)));
  double s = 0.0;
  for (int k = 0; k <= 2 * J; ++k) {
    double value = 0.0;
    if (J - mp - k >= 0 && J + m - k >= 0 && k + mp - m >= 0 && k >= 0) {
      const double val1 = pow(-1, mp - m + k) /
                          (o_fact((int)(J - mp - k)) * o_fact((int)(J + m - k)) *
                           o_fact((int)(k + mp - m)) * o_fact(k));
      const double val2 = pow(cos(th / 2.0), 2 * J + m - mp - 2 * k) * pow(sin(th / 2.0), mp - m + 2 * k);
      value = val1 * val2;
    }
    s += value;
  }
  return factor * s;
}
ORA_EXPORT double ora_wigner_d(double th, int m, int mp, int J) { return o_wigner_d(th, m, mp, J); }
/* spin::WignerD, src/MSpin.cc:968-970 */
static cx o_wigner_D(double th, double ph, double m, double mp, double J) {
  const cx e = cx_exp(cx_make(0.0, mp * ph));
  return cx_scale(e, o_wigner_d(th, m, mp, J));
}
static void o_theta_phi(v4 p, double* th, double* ph) {
  *ph = (p.x == 0.0 && p.y == 0.0) ? 0.0 : atan2(p.y, p.x);
  *th = (p.x == 0.0 && p.y == 0.0 && p.z == 0.0) ? 0.0 : atan2(v4_pt(p), p.z);
}
/* kinematics::LorentzBoost(b, M0, p, -1), include/Graniitti/MKinematics.h:384-401 */
static v4 o_boost_rest(v4 b, double M0, v4 p) {
  const double beta[3] = {-1 * b.x / b.e, -1 * b.y / b.e, -1 * b.z / b.e};
  const double gamma = b.e / M0;
  const double k1 = beta[0] * p.x + beta[1] * p.y + beta[2] * p.z;
  const double k2 = gamma * (gamma * k1 / (1.0 + gamma) + p.e);
  return v4_make(p.x + k2 * beta[0], p.y + k2 * beta[1], p.z + k2 * beta[2], gamma * (p.e + k1));
}
/* spin::GetRhoRotation, src/MSpin.cc:544-610 */
static void o_rho_rotation(const ora_run* R, const kin_t* K, double* thR, double* phR) {
  const double M = v4_m(K->pf[0]);
  *thR = 0.0;
  *phR = 0.0;
  if (R->c.spin_frame == GRA_FRAME_HX) {
    o_theta_phi(K->pf[0], thR, phR);
  } else if (R->c.spin_frame == GRA_FRAME_CS) {
    const v4 b1 = o_boost_rest(K->pf[0], M, R->pbeam1), b2 = o_boost_rest(K->pf[0], M, R->pbeam2);
    const double n1 = sqrt(v4_p3mod2(b1)), n2 = sqrt(v4_p3mod2(b2));
    double z[3] = {b1.x / n1 - b2.x / n2, b1.y / n1 - b2.y / n2, b1.z / n1 - b2.z / n2};
    const double nz = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    o_theta_phi(v4_make(z[0] / nz, z[1] / nz, z[2] / nz, 0.0), thR, phR);
  } else if (R->c.spin_frame == GRA_FRAME_GJ) {
    o_theta_phi(o_boost_rest(K->pf[0], M, K->q1), thR, phR);
  }
}
/* Re Tr[f rho_ROT f^dagger], rho_ROT = D^dagger rho D, in the reference's order
 * (spin::fMatrix, DMatrix, src/MSpin.cc:631-705; ProdAmp / DecayAmp 380-398, 460-484).
 * Rows r of f: final-state helicity lam[r] with amplitude T[r]. */
static double o_spin_trace(int J, int nrow, const int* lam, const cx* T, double th, double ph,
                           double thR, double phR, const double* rho_re, const double* rho_im) {
  const int n = 2 * J + 1;
  cx f[3][5], D[5][5], X[5][5], Rr[5][5], Y[3][5];
  for (int i = 0; i < nrow; ++i)
    for (int j = 0; j < n; ++j) f[i][j] = cx_mul(o_wigner_D(th, ph, lam[i], j - J, J), T[i]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) D[i][j] = o_wigner_D(thR, phR, i - J, j - J, J);
  for (int i = 0; i < n; ++i) /* D^dagger * rho */
    for (int j = 0; j < n; ++j) {
      cx a = cx_make(0, 0);
      for (int k = 0; k < n; ++k)
        a = cx_add(a, cx_mul(cx_make(D[k][i].re, -D[k][i].im), cx_make(rho_re[k * n + j], rho_im[k * n + j])));
      X[i][j] = a;
    }
  for (int i = 0; i < n; ++i) /* (D^dagger rho) * D */
    for (int j = 0; j < n; ++j) {
      cx a = cx_make(0, 0);
      for (int k = 0; k < n; ++k) a = cx_add(a, cx_mul(X[i][k], D[k][j]));
      Rr[i][j] = a;
    }
  for (int i = 0; i < nrow; ++i) /* f * rho_ROT */
    for (int j = 0; j < n; ++j) {
      cx a = cx_make(0, 0);
      for (int k = 0; k < n; ++k) a = cx_add(a, cx_mul(f[i][k], Rr[k][j]));
      Y[i][j] = a;
    }
  double tr = 0.0; /* Tr[(f rho_ROT) f^dagger] */
  for (int i = 0; i < nrow; ++i) {
    cx a = cx_make(0, 0);
    for (int k = 0; k < n; ++k) a = cx_add(a, cx_mul(Y[i][k], cx_make(f[i][k].re, -f[i][k].im)));
    tr += a.re;
  }
  return tr;
}
ORA_EXPORT double ora_spin_trace(int J, int nrow, const int* lam, const double* T, double th,
                                 double ph, double thR, double phR, const double* rho_re,
                                 const double* rho_im) {
  cx Tc[3];
  for (int i = 0; i < nrow; ++i) Tc[i] = cx_make(T[i], 0.0);
  return o_spin_trace(J, nrow, lam, Tc, th, ph, thR, phR, rho_re, rho_im);
}
/* decay helicity amplitude T_00 (spin::InitTMatrix, src/MSpin.cc:54-262, two spinless
 * daughters: only (l, s) = (J, 0), NORM = CG = 1, alpha renormalised to |alpha| = 1) */
static cx o_res_T(const gra_res* q) {
  const double a2 = q->alpha_re * q->alpha_re + q->alpha_im * q->alpha_im;
  const double an = fabs(a2 - 1.0) > 1e-6 ? sqrt(a2) : 1.0;
  return cx_make(q->alpha_re / an, q->alpha_im / an);
}
/* hc.g_decay = msqrt(BR W / PS), PS = PDW2body(M0^2, m^2, m^2, 1, 1)
 * (src/MProcess.cc:668-713, include/Graniitti/MKinematics.h:331-334) */
static double o_res_gdecay(const ora_run* R, const gra_res* q) {
  const double M02 = q->mass * q->mass, m2 = R->c.m_central * R->c.m_central;
  const double lam = M02 * M02 + m2 * m2 + m2 * m2 - 2.0 * (M02 * m2 + M02 * m2 + m2 * m2);
  const double PS = msqrt(lam) * 1.0 / (16.0 * PI * 1.0 * pow(msqrt(M02), 3));
  return msqrt(q->BR * q->width / PS);
}
ORA_EXPORT double ora_res_gdecay(const ora_run* R, int r) { return o_res_gdecay(R, &R->c.res[r]); }
/* spin::ProdAmp, src/MSpin.cc:280-403 (J^P = 0+, 1-, 2+) */
static cx o_prod_amp(const ora_run* R, const gra_res* q, const kin_t* K, double thR, double phR) {
  cx A0 = cx_make(q->g_A * cos(q->g_phi), q->g_A * sin(q->g_phi));
  if (!R->c.spin_gen || q->J == 0) return A0;
  int lam[3] = {0, 0, 0};
  cx T[3] = {cx_make(1, 0), cx_make(0, 0), cx_make(0, 0)};
  int nrow = 1;
  if (q->J == 1) { /* s1 = 0, s2 = 1: lambda = 0 - lambda2, lambda2 = -1, 0, 1 */
    nrow = 3;
    lam[0] = 1;
    lam[1] = 0;
    lam[2] = -1;
    T[0] = cx_make(-0.697, 0);
    T[1] = cx_make(-0.169, 0);
    T[2] = cx_make(0.697, 0);
  }
  double th, ph;
  o_theta_phi(o_boost_rest(K->pf[0], v4_m(K->pf[0]), K->q1), &th, &ph);
  const double tr = o_spin_trace(q->J, nrow, lam, T, th, ph, thR, phR, q->rho_re, q->rho_im);
  return cx_scale(A0, msqrt(2.0 * q->J + 1.0) * msqrt(tr));
}
/* spin::DecayAmp, src/MSpin.cc:406-486 (two spinless daughters, no recursion) */
static cx o_decay_amp(const ora_run* R, const gra_res* q, const kin_t* K, double thR, double phR) {
  const double A0 = o_res_gdecay(R, q);
  if (!R->c.spin_dec) return cx_make(A0, 0);
  const cx T = o_res_T(q);
  if (q->J == 0) return cx_scale(T, A0);
  double th, ph;
  o_theta_phi(o_boost_rest(K->pf[0], v4_m(K->pf[0]), K->dt[0]), &th, &ph);
  const int lam[1] = {0};
  const double tr = o_spin_trace(q->J, 1, lam, &T, th, ph, thR, phR, q->rho_re, q->rho_im);
  return cx_make(A0 * (msqrt(2.0 * q->J + 1.0) * msqrt(tr)), 0);
}
ORA_EXPORT int ora_res_spin(ora_run* R, const double* u, int r, double* out);
/* form::CBW, src/MForm.cc:743-828 */
static cx o_cbw(const gra_res* q, double m2, double mA, double mB) {
  const double M0 = q->mass, G = q->width;
  cx den = cx_make(m2 - M0 * M0, M0 * G);
  double num = -1.0;
  if (q->BW == GRA_BW_RUNNING) den = cx_make(m2 - M0 * M0, sqrt(m2) * G);
  if (q->BW == GRA_BW_BARRIER) {
    const int J = q->J;
    const double u = msqrt((m2 - (mA + mB) * (mA + mB)) * (m2 - (mA - mB) * (mA - mB))) / (2 * msqrt(m2));
    const double d = msqrt((M0 * M0 - (mA + mB) * (mA + mB)) * (M0 * M0 - (mA - mB) * (mA - mB))) / (2 * M0);
    den = cx_make(m2 - M0 * M0, G * M0 * M0 / msqrt(m2) * pow(u / d, 2 * J + 1));
  }
  if (q->BW == GRA_BW_SPIN) {
    if (q->J == 1) num = -(1.0 - m2 / (M0 * M0));
    if (q->J == 2)
      num = -(7.0 / 6.0 - (4.0 / 3.0) * (m2 / (M0 * M0)) + (2.0 / 3.0) * (m2 * m2) / pow(M0, 4));
  }
  const double d2 = den.re * den.re + den.im * den.im;
  return cx_make(num * den.re / d2, -num * den.im / d2);
}
/* form::CohFlux with the Kelly G_E, G_M (src/MForm.cc:607-687), alpha_QED(0) */
static double o_coh_flux(double xi, double t, double pt, double mp) {
  const double pt2 = pt * pt, xi2 = xi * xi, mp2 = mp * mp, Q2 = fabs(t);
  const double tau = Q2 / ((2 * mp) * (2 * mp));
  const double GE = (1.0 - 0.24 * tau) / (1.0 + 10.98 * tau + 12.82 * tau * tau + 21.97 * tau * tau * tau);
  const double GM = 2.792847337 * (1.0 + 0.12 * tau) / (1.0 + 10.97 * tau + 18.86 * tau * tau + 6.55 * tau * tau * tau);
  const double PART1 = (4.0 * mp2 * GE * GE + Q2 * GM * GM) / (4.0 * mp2 + Q2);
  const double PART2 = GM * GM;
  const double DELTA = pt2 / (pt2 + xi2 * mp2);
  double f = 0.0072973525693 / PI * ((1.0 - xi) * DELTA * DELTA * PART1 + (xi2 / 4.0) * DELTA * PART2);
  f /= xi;
  f /= pt2;
  f *= 16.0 * PI * PI;
  return f;
}
/* MRegge::PhotoProp, elastic leg (src/MRegge.cc:849-880) */
static cx o_photo_prop(const gra_cfg* c, double s, double t, double m2) {
  const double B0 = pow(fmax(0.7 * 0.7, m2) / 2.0, -PI / 2.0) + 3 * PI / 2.0;
  const double alpha = c->regge_a0[0] + c->regge_ap[0] * t;
  const cx eta = regge_eta_linear(t, c->regge_a0[0], c->regge_ap[0], c->regge_sgn[0]);
  const double FF = c->gN_P * exp(B0 * t / 2.0);
  return cx_scale(cx_scale(eta, pow(s / 1.0, alpha)), FF);
}
/* PARAM_REGGE::ResonanceFormFactor, src/MRegge.cc:1233-1235 */
static double o_res_ff(double s_hat, double M2, double S0) {
  return exp(-((s_hat - M2) * (s_hat - M2)) / pow(S0, 4));
}
/* sum over resonances of MRegge::ME3 (src/MRegge.cc:677-706) and, for 1-,
 * MRegge::PhotoME3 (src/MRegge.cc:781-843) */
static cx res_sum(const ora_run* R, const kin_t* K) {
  const gra_cfg* c = &R->c;
  double thR, phR;
  o_rho_rotation(R, K, &thR, &phR);
  const double FF_A = c->gN_P * s3f(c, K->t1), FF_B = c->gN_P * s3f(c, K->t2);
  cx A = cx_make(0, 0);
  for (int r = 0; r < c->n_res; ++r) {
    const gra_res* q = &c->res[r];
    const cx bw = o_cbw(q, K->m2, v4_m(K->dt[0]), v4_m(K->dt[1]));
    const double M2 = q->mass * q->mass;
    const double rff = o_res_ff(K->m2, M2, q->g_FF);
    const cx spin = cx_mul(o_prod_amp(R, q, K, thR, phR), o_decay_amp(R, q, K, thR, phR));
    if (q->J == 1 && q->P == -1) {
      const cx common = cx_scale(bw, rff);
      const double x1 = (1 - K->pf[1].z / R->pbeam1.z), x2 = (1 - K->pf[2].z / R->pbeam2.z);
      double gf1 = o_coh_flux(x1, K->t1, v4_pt(K->q1), c->tp_mp_ff);
      double gf2 = o_coh_flux(x2, K->t2, v4_pt(K->q2), c->tp_mp_ff);
      gf1 = msqrt(gf1 / x1);
      gf2 = msqrt(gf2 / x2);
      const cx A1 = cx_mul(cx_scale(common, gf2), o_photo_prop(c, K->s1, K->t1, M2));
      const cx A2 = cx_mul(cx_scale(common, gf1), o_photo_prop(c, K->s2, K->t2, M2));
      A = cx_add(A, cx_mul(cx_add(A1, A2), spin));
    } else {
      cx Ap = cx_scale(regge_prop(c, K->s1, K->t1), 2.0);
      Ap = cx_scale(Ap, FF_A);
      Ap = cx_mul(Ap, bw);
      Ap = cx_scale(Ap, rff);
      Ap = cx_mul(Ap, regge_prop(c, K->s2, K->t2));
      Ap = cx_scale(Ap, FF_B);
      const double V = pow(1.0 / K->m2, c->regge_omega);
      A = cx_add(A, cx_scale(cx_mul(Ap, spin), V));
    }
  }
  return A;
}

/* Tensor-pomeron ME4 (defined further below). Writes nh helicity amplitudes. */
static int tensor_me4(const ora_run* R, const kin_t* K, cx* hamp);

/* Helicity amplitudes of the bare process; returns their count. */
static int bare_amplitudes(const ora_run* R, const kin_t* K, cx* hamp) {
  if (R->c.process == GRA_PROC_TENSOR_CON) return tensor_me4(R, K, hamp);
  if (R->c.process == GRA_PROC_RES) { /* PROC_9: resonances only */
    hamp[0] = res_sum(R, K);
    return 1;
  }
  hamp[0] = regge_me4(R, K);
  if (R->c.process == GRA_PROC_RES_CON) hamp[0] = cx_add(hamp[0], res_sum(R, K));
  return 1;
}

/* |M|^2 of the bare process as MSubProc returns it (PROC_15: abs2(A); PROC_12:
 * sum |A_h|^2 / 4) */
static double bare_amp2(int nh, const cx* hamp) {
  if (nh == 1) return cx_abs2(hamp[0]);
  double a = 0.0;
  for (int h = 0; h < nh; ++h) a += cx_abs2(hamp[h]);
  return a / 4;
}

/* ------------------------------------------------------------------------- */
/* Eikonal tables (src/MEikonal.cc)                                          */
/* ------------------------------------------------------------------------- */
/* BESSJ0, libs/extmath.hpp:12-49 (Abramowitz-Stegun / Clenshaw rational fits) */
ORA_EXPORT double ora_bessj0(double X) {
  const double P1 = 1.0, P2 = -0.1098628627E-2, P3 = 0.2734510407E-4, P4 = -0.2073370639E-5,
               P5 = 0.2093887211E-6, Q1 = -0.1562499995E-1, Q2 = 0.1430488765E-3,
               Q3 = -0.6911147651E-5, Q4 = 0.7621095161E-6, Q5 = -0.9349451520E-7,
               R1 = 57568490574.0, R2 = -13362590354.0, R3 = 651619640.7, R4 = -11214424.18,
               R5 = 77392.33017, R6 = -184.9052456, S1 = 57568490411.0, S2 = 1029532985.0,
               S3 = 9494680.718, S4 = 59272.64853, S5 = 267.8532712, S6 = 1.0;
  if (X == 0.0) return 1.0;
  const double AX = fabs(X);
  if (AX < 8.0) {
    const double Y = X * X;
    const double FR = R1 + Y * (R2 + Y * (R3 + Y * (R4 + Y * (R5 + Y * R6))));
    const double FS = S1 + Y * (S2 + Y * (S3 + Y * (S4 + Y * (S5 + Y * S6))));
    return FR / FS;
  }
  const double Z = 8.0 / AX;
  const double Y = Z * Z;
  const double XX = AX - 0.785398164;
  const double FP = P1 + Y * (P2 + Y * (P3 + Y * (P4 + Y * P5)));
  const double FQ = Q1 + Y * (Q2 + Y * (Q3 + Y * (Q4 + Y * Q5)));
  return sqrt(0.636619772 / AX) * (FP * cos(XX) - Z * FQ * sin(XX));
}

/* gra::math::CSIntegral, include/Graniitti/MMath.h:378-394 (complex) */
static cx cs_integral(const cx* f, int N, double h) {
  cx Ie = cx_make(0, 0), Io = cx_make(0, 0);
  for (int j = 1; j <= N / 2 - 1; ++j) Ie = cx_add(Ie, f[2 * j]);
  for (int j = 1; j <= N / 2; ++j) Io = cx_add(Io, f[2 * j - 1]);
  const cx sum = cx_add(cx_add(cx_add(f[0], cx_scale(Ie, 2.0)), cx_scale(Io, 4.0)), f[N]);
  return cx_div_real(cx_scale(sum, h), 3.0);
}
ORA_EXPORT void ora_cs_integral(const double* f_reim, int N, double h, double* out) {
  cx* f = (cx*)malloc(sizeof(cx) * (N + 1));
  for (int i = 0; i <= N; ++i) f[i] = cx_make(f_reim[2 * i], f_reim[2 * i + 1]);
  const cx r = cs_integral(f, N, h);
  free(f);
  out[0] = r.re;
  out[1] = r.im;
}

/* Hankel transform (2 pi) ∫ g(k) J0(b k) k dk on the Fourier-Bessel grid of
 * S3Density (k in [kmin, kmax], N intervals) for a real test function given on
 * the grid; used to pin the quadrature against closed forms. */
ORA_EXPORT double ora_fb_transform(const double* g, int N, double kmin, double kmax, double b) {
  const double h = (kmax - kmin) / N;
  cx* f = (cx*)malloc(sizeof(cx) * (N + 1));
  for (int i = 0; i <= N; ++i) {
    const double k = kmin + i * h;
    f[i] = cx_make(g[i] * ora_bessj0(b * k) * k, 0.0);
  }
  const cx r = cs_integral(f, N, h);
  free(f);
  return r.re;
}

/* MEikonal::SingleAmpElastic(s, t, +1), src/MEikonal.cc:139-154 */
static cx single_amp_elastic(const ora_run* R, double t) {
  const gra_cfg* c = &R->c;
  const double F = s3f(c, t);
  const double aP = ora_pom_alpha(c, t);
  const double s0 = 1.0;
  const cx eta = regge_eta(aP, 1);
  const double pre = ((c->gN_P * c->gN_P) * F) * F;
  return cx_scale(cx_scale(eta, pre), pow(R->s / s0, aP));
}

/* IArray1D::Interpolate1D (linear), src/MEikonal.cc:489-512 */
static cx interp1d(const double* X, const cx* Y, int N, double MIN, double MAX, double STEP,
                   double a) {
  if (a < MIN) a = MIN;
  (void)MAX;
  int i = (int)floor((a - MIN) / STEP);
  if (i < 0) i = 0;
  if (i >= N) i = N - 1;
  const cx dy = cx_sub(Y[i + 1], Y[i]);
  const double dx = X[i + 1] - X[i];
  return cx_add(Y[i], cx_scale(cx_div_real(dy, dx), a - X[i]));
}

/* WriteArray/ReadArray round trip: std::setprecision(15) then std::stod,
 * src/MEikonal.cc:404-485 */
static double round15(double v) {
  char buf[64];
  snprintf(buf, sizeof(buf), "%.15g", v);
  return strtod(buf, NULL);
}

typedef struct {
  ora_run* R;
  int which; /* 0 = MBT density, 1 = MSA screening */
  int i0, i1;
  const cx* pre; /* hoisted per-quadrature-node factors */
} eik_job;

static void* eik_worker(void* arg) {
  eik_job* J = (eik_job*)arg;
  ora_run* R = J->R;
  const int NF = R->n_fb;
  cx* f = (cx*)malloc(sizeof(cx) * (NF + 1));
  for (int i = J->i0; i < J->i1; ++i) {
    if (J->which == 0) {
      /* MEikonal::S3Density(bt), src/MEikonal.cc:177-216 */
      const double bt = R->mbt_min + i * R->mbt_step;
      const double kmin = 1E-9, kmax = 30.0;
      const double kstep = (kmax - kmin) / NF;
      for (int q = 0; q <= NF; ++q) {
        const double kt = kmin + q * kstep;
        f[q] = cx_scale(cx_scale(J->pre[q], ora_bessj0(bt * kt)), kt);
      }
      const double FACTOR = 1 / (2.0 * PI * R->s);
      R->mbt_y[i] = cx_scale(cs_integral(f, NF, kstep), FACTOR);
    } else {
      /* MEikonal::S3Screening(kt2), src/MEikonal.cc:222-245 */
      const double kt2 = R->msa_min + i * R->msa_step;
      const double STEP = (R->mbt_max - R->mbt_min) / NF;
      const double kt = msqrt(kt2);
      for (int q = 0; q <= NF; ++q) {
        const double bt = R->mbt_min + q * STEP;
        f[q] = cx_scale(cx_scale(J->pre[q], ora_bessj0(bt * kt)), bt);
      }
      const double FACTOR = 2.0 * PI;
      R->msa_y[i] = cx_scale(cs_integral(f, NF, STEP), (2.0 * R->s) * FACTOR);
    }
  }
  free(f);
  return NULL;
}

static void run_eik_jobs(ora_run* R, int which, int n, const cx* pre, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  pthread_t th[64];
  eik_job jobs[64];
  if (nthreads > 64) nthreads = 64;
  for (int t = 0; t < nthreads; ++t) {
    jobs[t].R = R;
    jobs[t].which = which;
    jobs[t].i0 = (int)((long long)n * t / nthreads);
    jobs[t].i1 = (int)((long long)n * (t + 1) / nthreads);
    jobs[t].pre = pre;
    pthread_create(&th[t], NULL, eik_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* MEikonal::S3Constructor tables, src/MEikonal.cc:46-132 */
static void build_tables(ora_run* R, int nthreads) {
  const int NF = R->n_fb;
  cx* pre = (cx*)malloc(sizeof(cx) * (NF + 1));
  /* MBT: f = SingleAmpElastic(s, -kt^2) * J0(bt kt) * kt; the amplitude does not
   * depend on bt and is evaluated once per quadrature node. */
  {
    const double kmin = 1E-9, kmax = 30.0, kstep = (kmax - kmin) / NF;
    for (int q = 0; q <= NF; ++q) {
      const double kt = kmin + q * kstep;
      pre[q] = single_amp_elastic(R, -(kt * kt));
    }
  }
  run_eik_jobs(R, 0, R->n_bt + 1, pre, nthreads);
  for (int i = 0; i <= R->n_bt; ++i) {
    R->mbt_x[i] = round15(R->mbt_min + i * R->mbt_step);
    R->mbt_y[i] = cx_make(round15(R->mbt_y[i].re), round15(R->mbt_y[i].im));
  }
  /* MSA: f = i (1 - exp(i Omega(bt) / 2)) * J0(bt kt) * bt */
  {
    const double STEP = (R->mbt_max - R->mbt_min) / NF;
    for (int q = 0; q <= NF; ++q) {
      const double bt = R->mbt_min + q * STEP;
      const cx Om = interp1d(R->mbt_x, R->mbt_y, R->n_bt, R->mbt_min, R->mbt_max, R->mbt_step, bt);
      /* zi * Omega / 2.0 = (-Om.im, Om.re) / 2 */
      const cx e = cx_exp(cx_div_real(cx_make(0.0 * Om.re - 1.0 * Om.im, 0.0 * Om.im + 1.0 * Om.re), 2.0));
      const cx one_m = cx_make(1.0 - e.re, -e.im);
      pre[q] = cx_make(0.0 * one_m.re - 1.0 * one_m.im, 0.0 * one_m.im + 1.0 * one_m.re);
    }
  }
  run_eik_jobs(R, 1, R->n_kt2 + 1, pre, nthreads);
  for (int i = 0; i <= R->n_kt2; ++i) {
    R->msa_x[i] = round15(R->msa_min + i * R->msa_step);
    R->msa_y[i] = cx_make(round15(R->msa_y[i].re), round15(R->msa_y[i].im));
  }
  /* A_eik at log-spaced kt: the S3Screening integral at those kt^2 */
  const int NL = R->c.eik_log_nodes;
  if (NL > 0) {
    const double STEP = (R->mbt_max - R->mbt_min) / NF;
    const double l0 = log(R->c.min_loop_kt), l1 = log(R->c.max_loop_kt);
    cx* f = (cx*)malloc(sizeof(cx) * (NF + 1));
    for (int m = 0; m < NL; ++m) {
      R->l_lkt[m] = l0 + (l1 - l0) * m / (NL - 1);
      const double e = exp(R->l_lkt[m]);
      const double kt = msqrt(e * e); /* as S3Screening takes sqrt(kt^2) */
      for (int q = 0; q <= NF; ++q) {
        const double bt = R->mbt_min + q * STEP;
        f[q] = cx_scale(cx_scale(pre[q], ora_bessj0(bt * kt)), bt);
      }
      R->l_amp[m] = cx_scale(cs_integral(f, NF, STEP), (2.0 * R->s) * (2.0 * PI));
    }
    free(f);
  }
  free(pre);
  R->tables_ready = 1;
}

/* A_eik(kt) interpolated linearly in log kt between the log-spaced nodes */
static cx log_node_amp(const ora_run* R, double kt) {
  const int NL = R->c.eik_log_nodes;
  const double lk = log(kt), step = (R->l_lkt[NL - 1] - R->l_lkt[0]) / (NL - 1);
  int m = (int)floor((lk - R->l_lkt[0]) / step);
  if (m < 0) m = 0;
  if (m > NL - 2) m = NL - 2;
  const double f = (lk - R->l_lkt[m]) / (R->l_lkt[m + 1] - R->l_lkt[m]);
  return cx_add(R->l_amp[m], cx_scale(cx_sub(R->l_amp[m + 1], R->l_amp[m]), f));
}

/* ------------------------------------------------------------------------- */
/* Screened amplitude and event weight                                       */
/* ------------------------------------------------------------------------- */
#define MAX_HAMP 64

/* Simpson 3/8 weight vector, include/Graniitti/MMath.h:502-516 */
static double simpson38_w(int i, int N) {
  if (i == 0 || i == N) return 1.0;
  return (i % 3 == 0) ? 2.0 : 3.0;
}
ORA_EXPORT double ora_simpson38_weight(int i, int N) { return simpson38_w(i, N); }

/* MProcess::S3ScreenedAmp2, src/MProcess.cc:157-288 (non-gg branch) */
static double screened_amp2(ora_run* R, kin_t* K) {
  cx hamp0[MAX_HAMP], h[MAX_HAMP];
  const int nh = bare_amplitudes(R, K, hamp0);
  if (!R->c.screening) return bare_amp2(nh, hamp0);

  const double p1T[2] = {K->pf[1].x, K->pf[1].y};
  const double p2T[2] = {K->pf[2].x, K->pf[2].y};
  for (int i = 0; i < 5; ++i) K->pf_orig[i] = K->pf[i];
  const int NP = R->nphi, NK = R->nkt;
  const double StepKT = R->step_kt, StepPhi = R->step_phi;
  const int gauss = R->c.loop_rule == GRA_LOOP_GAUSS;
  cx I[MAX_HAMP];
  for (int q = 0; q < nh; ++q) I[q] = cx_make(0, 0);
  /* The reference stores f[h][i][j] and sums W[i][j] * f[i][j] in (i, j) order;
   * accumulating in the same order gives the same sums. The Gauss rule: radial
   * Gauss-Legendre nodes, equally spaced azimuths with weight 2 pi / (NP + 1). */
  for (int i = 0; i <= NP; ++i) {
    const double phi = gauss ? 2.0 * PI * i / (NP + 1) : 0.0 + i * StepPhi;
    for (int j = 0; j <= NK; ++j) {
      const double kt = gauss ? R->g_kt[j] : R->c.min_loop_kt + j * StepKT;
      const double kt2 = kt * kt;
      const cx A_eik =
          R->c.eik_log_nodes > 0
              ? log_node_amp(R, kt)
              : interp1d(R->msa_x, R->msa_y, R->n_kt2, R->msa_min, R->msa_max, R->msa_step, kt2);
      const double ktx = kt * cos(phi), kty = kt * sin(phi);
      const double p1p[2] = {p1T[0] - ktx, p1T[1] - kty};
      const double p2p[2] = {p2T[0] + ktx, p2T[1] + kty};
      const double w = gauss ? (2.0 * PI / (NP + 1)) * R->g_w[j]
                             : simpson38_w(i, NP) * simpson38_w(j, NK);
      if (!loop_kinematics(R, K, p1p, p2p)) {
        continue; /* f[i][j] = 0 still enters the sum as W * 0 */
      }
      bare_amplitudes(R, K, h);
      for (int q = 0; q < nh; ++q) {
        const cx f = cx_scale(cx_mul(A_eik, h[q]), kt);
        I[q] = cx_add(I[q], cx_scale(f, w));
      }
    }
  }
  const double nrm = 1.0 / (8.0 * (PI * PI) * R->s);
  const double hfac = gauss ? 1.0 : 9 * (StepPhi * StepKT) / 64;
  loop_kinematics(R, K, p1T, p2T); /* back to tree level */
  double amp2 = 0.0;
  for (int q = 0; q < nh; ++q) {
    const cx Iq = cx_scale(I[q], hfac);
    const cx loop = cx_make(0.0 * Iq.re - nrm * Iq.im, 0.0 * Iq.im + nrm * Iq.re);
    amp2 += cx_abs2(cx_add(hamp0[q], loop));
  }
  if (nh != 1) amp2 /= 4;
  return amp2;
}

/* MContinuum::EventWeight, src/MContinuum.cc:164-193, with BNPhaseSpaceWeight
 * (639-656), BNIntegralVolume (619-636), ForwardVolume (src/MProcess.cc:1390-1415),
 * MollerFlux = 2 s (src/MProcess.h:214) and CheckInfNan (src/MProcess.h:317-340). */
static double event_weight(ora_run* R, const double* x, double* amp2_out, uint8_t* flags) {
  const gra_cfg* c = &R->c;
  kin_t K;
  memset(&K, 0, sizeof(K));
  uint8_t fl = 0;
  double W = 0.0, amp2 = 0.0;
  const int kin_ok = build_kin(R, x, &K);
  if (kin_ok) fl |= GRA_FLAG_KIN_OK;
  const int fid_ok = kin_ok ? fiducial_cuts(R, &K) : 0;
  if (fid_ok) fl |= GRA_FLAG_FID_OK;
  /* tested on valid kinematics only, like the fiducial flag */
  const int veto_ok = kin_ok ? veto_cuts(R, &K) : 1;
  if (veto_ok) fl |= GRA_FLAG_VETO_OK;
  if (kin_ok && fid_ok && veto_ok) {
    amp2 = screened_amp2(R, &K);
    if (c->phase_space == GRA_PS_F) {
      /* MFactorized::EventWeight, src/MFactorized.cc:142-182: C_space = DW / (2 pi)
       * times CascadePS (1 for stable products); B51PhaseSpaceWeight 441-452 with
       * pow5(2 pi) = x pow2(pow2(x)) (include/Graniitti/MMath.h:162-176);
       * B51IntegralVolume 434-438 with ForwardVolume (src/MProcess.cc:1390-1400) */
      const double C_space = (K.dw / (2 * PI)) * 1.0;
      const double J = 1.0 / fabs(K.pf[1].z / K.pf[1].e - K.pf[2].z / K.pf[2].e);
      const double tp = 2.0 * PI;
      const double psw = (1.0 / 2.0) * (1.0 / (tp * ((tp * tp) * (tp * tp)))) *
                         (v4_pt(K.pf[1]) / (2.0 * K.pf[1].e)) *
                         (v4_pt(K.pf[2]) / (2.0 * K.pf[2].e)) * J;
      const double ZERO_EPS = 1e-12;
      const double PHI_vol = (2.0 * PI) * (2.0 * PI);
      const double J_pt = v4_pt(K.pf[1]) * v4_pt(K.pf[2]);
      const double lspan = log(c->fwd_pt_max) - log(c->fwd_pt_min + ZERO_EPS);
      const double fwd = PHI_vol * (J_pt * (lspan * lspan));
      const double Mn = R->f_Mmin, Mx = R->f_Mmax;
      const double vol = (Mx * Mx - Mn * Mn) * (c->rap_max - c->rap_min) * fwd;
      W = C_space * (1.0 / c->S_factor) * psw * vol * amp2 * R->GeV2barn / (2.0 * R->s);
    } else {
    const double C_space = 1.0;
    const double J = 1.0 / fabs(K.pf[1].z / K.pf[1].e - K.pf[2].z / K.pf[2].e);
    const double PROD = 1.0 * v4_pt(K.pkt);
    const double Nf = 4;
    const double psw = (1.0 / pow(2, 2 * (Nf - 2))) * (1.0 / pow(2.0 * PI, 3 * Nf - 4)) *
                       (v4_pt(K.pf[1]) / (2.0 * K.pf[1].e)) * (v4_pt(K.pf[2]) / (2.0 * K.pf[2].e)) *
                       J * PROD;
    const double ZERO_EPS = 1e-12;
    const double PHI_vol = (2.0 * PI) * (2.0 * PI);
    const double J_pt = v4_pt(K.pf[1]) * v4_pt(K.pf[2]);
    const double lspan = log(c->fwd_pt_max) - log(c->fwd_pt_min + ZERO_EPS);
    const double PT_vol = J_pt * (lspan * lspan);
    double KT_vol = pow(c->kt_max - c->kt_min, 1);
    KT_vol *= pow(2.0 * PI, 1);
    KT_vol *= pow(c->rap_max - c->rap_min, 2);
    const double vol = KT_vol * (PHI_vol * PT_vol);
    W = C_space * (1.0 / c->S_factor) * psw * vol * amp2 * R->GeV2barn / (2.0 * R->s);
    }
  }
  if (isnan(W) || isinf(W)) {
    if (isinf(W)) fl |= GRA_FLAG_INF;
    W = 0;
  } else {
    fl |= GRA_FLAG_AMP_OK;
  }
  if (amp2_out) *amp2_out = amp2;
  if (flags) *flags = fl;
  return W;
}

/* ------------------------------------------------------------------------- */
/* Run object                                                                */
/* ------------------------------------------------------------------------- */
ORA_EXPORT ora_run* ora_create(const gra_cfg* cfg, int nthreads) {
  if (cfg->abi_size != (int32_t)sizeof(gra_cfg)) return NULL;
  ora_run* R = (ora_run*)calloc(1, sizeof(ora_run));
  R->c = *cfg;
  const double E1 = cfg->beam_E1, E2 = cfg->beam_E2, m = cfg->m_beam;
  R->pbeam1 = v4_make(0, 0, msqrt(E1 * E1 - m * m), E1);
  R->pbeam2 = v4_make(0, 0, -msqrt(E2 * E2 - m * m), E2);
  R->beamsum = v4_add(R->pbeam1, R->pbeam2);
  R->s = v4_m2(R->beamsum);
  R->sqrt_s = msqrt(R->s);
  R->GeV2barn = gev2barn();
  R->nphi = cfg->n_loop_phi;
  R->nkt = cfg->n_loop_kt;
  R->step_kt = (cfg->max_loop_kt - cfg->min_loop_kt) / cfg->n_loop_kt;
  R->step_phi = (2.0 * PI - 0.0) / cfg->n_loop_phi;
  if (cfg->phase_space == GRA_PS_F) {
    /* M_MIN = sum of decay-tree masses + 1e-4 GeV, M_MAX = sqrt(s) - beam masses,
     * clipped to the GENCUTS (src/MFactorized.cc:259-272) */
    R->f_Mmin = fmax((0.0 + cfg->m_central) + cfg->m_central + 1e-4, cfg->gen_M_min);
    R->f_Mmax = fmin(R->sqrt_s - (v4_m(R->pbeam1) + v4_m(R->pbeam2)), cfg->gen_M_max);
  }
  if (cfg->screening) {
    R->n_bt = cfg->n_bt;
    R->n_kt2 = cfg->n_kt2;
    R->n_fb = cfg->n_fb;
    R->mbt_min = 1E-6;
    R->mbt_max = 10.0 / gev2fm();
    R->mbt_step = (R->mbt_max - R->mbt_min) / R->n_bt;
    R->msa_min = 1E-6;
    R->msa_max = 25.0;
    R->msa_step = (R->msa_max - R->msa_min) / R->n_kt2;
    R->mbt_x = (double*)malloc(sizeof(double) * (R->n_bt + 1));
    R->msa_x = (double*)malloc(sizeof(double) * (R->n_kt2 + 1));
    R->mbt_y = (cx*)malloc(sizeof(cx) * (R->n_bt + 1));
    R->msa_y = (cx*)malloc(sizeof(cx) * (R->n_kt2 + 1));
    if (cfg->eik_log_nodes > 0) {
      R->l_lkt = (double*)malloc(sizeof(double) * cfg->eik_log_nodes);
      R->l_amp = (cx*)malloc(sizeof(cx) * cfg->eik_log_nodes);
    }
    if (cfg->loop_rule == GRA_LOOP_GAUSS) {
      const int n = cfg->n_loop_kt + 1;
      const double half = 0.5 * (cfg->max_loop_kt - cfg->min_loop_kt);
      R->g_kt = (double*)malloc(sizeof(double) * n);
      R->g_w = (double*)malloc(sizeof(double) * n);
      ora_gauss_legendre(n, R->g_kt, R->g_w);
      for (int j = 0; j < n; ++j) {
        R->g_kt[j] = cfg->min_loop_kt + half * (1.0 + R->g_kt[j]);
        R->g_w[j] *= half;
      }
    }
    build_tables(R, nthreads);
  }
  return R;
}

ORA_EXPORT void ora_destroy(ora_run* R) {
  if (!R) return;
  free(R->mbt_x);
  free(R->msa_x);
  free(R->mbt_y);
  free(R->msa_y);
  free(R->g_kt);
  free(R->g_w);
  free(R->l_lkt);
  free(R->l_amp);
  free(R);
}

/* A_eik at kt from the log-spaced nodes (test access; eik_log_nodes > 0) */
ORA_EXPORT void ora_log_node_amp(const ora_run* R, double kt, double* out) {
  const cx a = log_node_amp(R, kt);
  out[0] = a.re;
  out[1] = a.im;
}

/* Table access: rows of (x, re, im) */
ORA_EXPORT int ora_tables(const ora_run* R, double* mbt, double* msa) {
  if (!R->tables_ready) return -1;
  for (int i = 0; i <= R->n_bt; ++i) {
    mbt[3 * i] = R->mbt_x[i];
    mbt[3 * i + 1] = R->mbt_y[i].re;
    mbt[3 * i + 2] = R->mbt_y[i].im;
  }
  for (int i = 0; i <= R->n_kt2; ++i) {
    msa[3 * i] = R->msa_x[i];
    msa[3 * i + 1] = R->msa_y[i].re;
    msa[3 * i + 2] = R->msa_y[i].im;
  }
  return 0;
}

/* Interpolated screening amplitude A_eik(kt2) */
ORA_EXPORT void ora_msa(const ora_run* R, double kt2, double* out) {
  const cx z = interp1d(R->msa_x, R->msa_y, R->n_kt2, R->msa_min, R->msa_max, R->msa_step, kt2);
  out[0] = z.re;
  out[1] = z.im;
}
ORA_EXPORT void ora_mbt(const ora_run* R, double bt, double* out) {
  const cx z = interp1d(R->mbt_x, R->mbt_y, R->n_bt, R->mbt_min, R->mbt_max, R->mbt_step, bt);
  out[0] = z.re;
  out[1] = z.im;
}

/* Batched EventWeight on caller points (n x dim row-major) */
ORA_EXPORT void ora_eval_points(ora_run* R, const double* u, int64_t n, double* W, double* amp2,
                                uint8_t* flags) {
  const int D = R->c.dim;
  for (int64_t i = 0; i < n; ++i) {
    W[i] = event_weight(R, u + i * D, amp2 ? amp2 + i : NULL, flags ? flags + i : NULL);
  }
}

/* Kinematic state of one point for tests: pfinal[0..4], decaytree legs and the
 * scalars (t1, t2, t_hat, u_hat, s13, s24, s14, s23). Returns kinematics_ok. */
ORA_EXPORT int ora_kinematics(ora_run* R, const double* x, double* pf /* 5x4 E,x,y,z */,
                              double* legs /* 4x4 */, double* sc /* 8 */) {
  kin_t K;
  memset(&K, 0, sizeof(K));
  const int ok = build_kin(R, x, &K);
  for (int i = 0; i < 5; ++i) {
    pf[4 * i] = K.pf[i].e;
    pf[4 * i + 1] = K.pf[i].x;
    pf[4 * i + 2] = K.pf[i].y;
    pf[4 * i + 3] = K.pf[i].z;
  }
  for (int i = 0; i < 2; ++i)
    for (int k = 0; k < 2; ++k) {
      const v4 v = K.legs[i][k];
      legs[8 * i + 4 * k] = v.e;
      legs[8 * i + 4 * k + 1] = v.x;
      legs[8 * i + 4 * k + 2] = v.y;
      legs[8 * i + 4 * k + 3] = v.z;
    }
  sc[0] = K.t1;
  sc[1] = K.t2;
  sc[2] = K.t_hat;
  sc[3] = K.u_hat;
  sc[4] = K.ss[1][3];
  sc[5] = K.ss[2][4];
  sc[6] = K.ss[1][4];
  sc[7] = K.ss[2][3];
  return ok;
}

/* Bare (unscreened) helicity amplitudes at one point: returns their count, or 0
 * if the kinematics fail. */
ORA_EXPORT int ora_bare_amplitudes(ora_run* R, const double* x, double* hamp_reim) {
  kin_t K;
  memset(&K, 0, sizeof(K));
  if (!build_kin(R, x, &K)) return 0;
  cx h[MAX_HAMP];
  const int nh = bare_amplitudes(R, &K, h);
  for (int q = 0; q < nh; ++q) {
    hamp_reim[2 * q] = h[q].re;
    hamp_reim[2 * q + 1] = h[q].im;
  }
  return nh;
}

/* ------------------------------------------------------------------------- */
/* VEGAS iterations: multithreaded CPU run (reference threading model)        */
/* ------------------------------------------------------------------------- */
typedef struct {
  ora_run* R;
  const gra_grid* g;
  uint64_t seed;
  uint32_t iter;
  uint64_t i0, i1;
  int use_mutex;
  pthread_mutex_t* mtx;
  gra_iter_stats* shared; /* mutex mode */
  gra_iter_stats local;   /* per-thread mode */
} vg_job;

static inline void kahan_add(double* s, double* c, double v) {
  const double y = v - *c;
  const double t = *s + y;
  *c = (t - *s) - y;
  *s = t;
}

static void accumulate_point(gra_iter_stats* st, const gra_grid* g, double W, double vw,
                             const int32_t* bin, uint8_t fl) {
  const double f = W * vw;
  const double f2 = f * f;
  kahan_add(&st->fsum, &st->fsum_c, f);
  kahan_add(&st->f2sum, &st->f2sum_c, f2);
  for (int j = 0; j < g->dim; ++j) {
    st->fmat[j * GRA_VEGAS_BINS + bin[j] - 1] += f;
    st->f2mat[j * GRA_VEGAS_BINS + bin[j] - 1] += f2;
  }
  if (W > st->maxW) st->maxW = W;
  if (f > st->maxf) st->maxf = f;
  st->n_calls += 1;
  if (fl & GRA_FLAG_KIN_OK) st->n_kin_ok += 1;
  if (fl & GRA_FLAG_FID_OK) st->n_fid_ok += 1;
  if ((fl & GRA_FLAG_KIN_OK) && (fl & GRA_FLAG_FID_OK) && (fl & GRA_FLAG_VETO_OK)) st->n_valid += 1;
  if (!(fl & GRA_FLAG_AMP_OK)) {
    if (fl & GRA_FLAG_INF)
      st->n_inf += 1;
    else
      st->n_nan += 1;
  }
}

static void* vg_worker(void* arg) {
  vg_job* J = (vg_job*)arg;
  const int D = J->g->dim;
  double u[GRA_MAX_DIM], x[GRA_MAX_DIM];
  int32_t bin[GRA_MAX_DIM];
  for (uint64_t i = J->i0; i < J->i1; ++i) {
    ora_point_uniforms(J->seed, J->iter, i, D, u);
    double vw;
    ora_vegas_map(J->g, u, x, &vw, bin);
    uint8_t fl;
    const double W = event_weight(J->R, x, NULL, &fl);
    if (J->use_mutex) {
      pthread_mutex_lock(J->mtx);
      accumulate_point(J->shared, J->g, W, vw, bin, fl);
      pthread_mutex_unlock(J->mtx);
    } else {
      accumulate_point(&J->local, J->g, W, vw, bin, fl);
    }
  }
  return NULL;
}

/* One VEGAS iteration over points [first, first + n) of the Philox stream
 * (first > 0 lets the benchmark time a bounded sample of a larger iteration). */
ORA_EXPORT void ora_vegas_iterate(ora_run* R, const gra_grid* g, uint64_t first, uint64_t n,
                                  uint64_t seed, uint32_t iter, int nthreads, int use_mutex,
                                  gra_iter_stats* out) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  memset(out, 0, sizeof(*out));
  vg_job* jobs = (vg_job*)calloc(nthreads, sizeof(vg_job));
  pthread_t* th = (pthread_t*)calloc(nthreads, sizeof(pthread_t));
  pthread_mutex_t mtx = PTHREAD_MUTEX_INITIALIZER;
  for (int t = 0; t < nthreads; ++t) {
    jobs[t].R = R;
    jobs[t].g = g;
    jobs[t].seed = seed;
    jobs[t].iter = iter;
    jobs[t].i0 = first + n * t / nthreads;
    jobs[t].i1 = first + n * (t + 1) / nthreads;
    jobs[t].use_mutex = use_mutex;
    jobs[t].mtx = &mtx;
    jobs[t].shared = out;
  }
  /* Each thread needs its own process state: the run object is read-only during
   * evaluation (all mutable state lives on the stack). */
  for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, vg_worker, &jobs[t]);
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  if (!use_mutex) {
    for (int t = 0; t < nthreads; ++t) {
      const gra_iter_stats* L = &jobs[t].local;
      kahan_add(&out->fsum, &out->fsum_c, L->fsum - L->fsum_c);
      kahan_add(&out->f2sum, &out->f2sum_c, L->f2sum - L->f2sum_c);
      for (int k = 0; k < GRA_MAX_DIM * GRA_VEGAS_BINS; ++k) {
        out->fmat[k] += L->fmat[k];
        out->f2mat[k] += L->f2mat[k];
      }
      if (L->maxW > out->maxW) out->maxW = L->maxW;
      if (L->maxf > out->maxf) out->maxf = L->maxf;
      out->n_calls += L->n_calls;
      out->n_kin_ok += L->n_kin_ok;
      out->n_fid_ok += L->n_fid_ok;
      out->n_valid += L->n_valid;
      out->n_nan += L->n_nan;
      out->n_inf += L->n_inf;
    }
  }
  free(jobs);
  free(th);
}

/* Iteration combination of MGraniitti::VEGAS, src/MGraniitti.cc:1222-1251.
 * state = {sumdata, sumchi2, sweight}; out = {I_this, I2_this, sigma, sigma_err, chi2}. */
ORA_EXPORT void ora_vegas_combine(double fsum, double f2sum, double calls, int iter, double* state,
                                  double* out) {
  const double EPS = 1.0e-30;
  fsum *= 1.0 / calls;
  f2sum *= (1.0 / calls) * (1.0 / calls);
  const double I_this = fsum;
  f2sum = msqrt(f2sum * calls);
  f2sum = (f2sum - fsum) * (f2sum + fsum);
  if (f2sum <= 0.0) f2sum = EPS;
  const double I2_this = f2sum * 1.0 / calls;
  const double alpha = 1.0 / I2_this;
  state[0] += alpha * I_this;
  state[1] += alpha * (I_this * I_this);
  state[2] += alpha;
  const double sigma = state[0] / state[2];
  const double sigma_err = msqrt(1.0 / state[2]);
  double chi2 = (state[1] - (state[0] * state[0]) / state[2]) / (iter + 1E-5);
  chi2 = (chi2 < 0) ? 0.0 : chi2;
  out[0] = I_this;
  out[1] = I2_this;
  out[2] = sigma;
  out[3] = sigma_err;
  out[4] = chi2;
}

/* ------------------------------------------------------------------------- */
/* Tensor pomeron, MTensorPomeron::ME4 for a vector-meson pair ("2xV"),       */
/* src/MTensorPomeron.cc:572-898 (2xV block 808-888). Lorentz tensors are     */
/* stored as plain arrays with the component order (t, x, y, z) and          */
/* contracted with plain index sums, as the reference's FTensor expressions   */
/* do; the metric g = diag(1, -1, -1, -1) appears explicitly where the        */
/* reference writes g[u][v], and `%` (covariant component) is v_mu = g_mumu v^mu. */
/* ------------------------------------------------------------------------- */
/* The tensor network is accumulated in long double: evaluated in double, as the
 * reference does, its sums of 256 large terms lose up to ~1e-8 relative to
 * cancellation at some points, which would dominate any FP64 comparison. */
typedef struct {
  long double re, im;
} lcx;
static inline lcx lc_make(long double re, long double im) {
  lcx z;
  z.re = re;
  z.im = im;
  return z;
}
static inline lcx lc_add(lcx a, lcx b) { return lc_make(a.re + b.re, a.im + b.im); }
static inline lcx lc_mul(lcx a, lcx b) {
  return lc_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline lcx lc_scale(lcx a, long double d) { return lc_make(a.re * d, a.im * d); }
typedef lcx t2_t[4][4];
typedef lcx t4_t[4][4][4][4];

static const long double GM[4][4] = {{1, 0, 0, 0}, {0, -1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, -1}};

static inline long double v4_c(v4 a, int mu) { return mu == 0 ? a.e : (mu == 1 ? a.x : (mu == 2 ? a.y : a.z)); }
/* M4Vec::operator% (lowered component), include/Graniitti/M4Vec.h:216-223 */
static inline long double v4_lo(v4 a, int mu) { return GM[mu][mu] * v4_c(a, mu); }

/* std::pow(std::complex<double>, double) of libstdc++: real pow for a positive
 * real base, else polar(exp(y log|z|), y arg z) */
static lcx lc_pow_real(lcx z, long double y) {
  if (z.im == 0.0L && z.re > 0.0L) return lc_make(powl(z.re, y), 0.0L);
  const long double r = expl(y * logl(hypotl(z.re, z.im)));
  const long double th = y * atan2l(z.im, z.re);
  return lc_make(r * cosl(th), r * sinl(th));
}
static inline long double ld_dot(v4 a, v4 b) {
  return (long double)a.e * b.e -
         ((long double)a.x * b.x + (long double)a.y * b.y + (long double)a.z * b.z);
}
static inline long double ld_m2(v4 a) { return ld_dot(a, a); }

/* form::F1 with the Kelly G_E, G_M parametrisation (PARAM_STRUCTURE EM = KELLY),
 * src/MForm.cc:542-548, 611-645; mu_p = 2.792847337 (src/MForm.cc:521) */
static long double kelly_f1(long double Q2, long double mp) {
  Q2 = fabsl(Q2);
  const long double tau = Q2 / ((2 * mp) * (2 * mp));
  const long double mu = 2.792847337L;
  long double num = 0.0L, den = 1.0L;
  num += 1.0;
  num += -0.24L * tau;
  den += 10.98L * tau;
  den += 12.82L * (tau * tau);
  den += 21.97L * (tau * tau * tau);
  const long double GE = num / den;
  num = 0.0;
  den = 1.0;
  num += 1.0;
  num += 0.12L * tau;
  den += 10.97L * tau;
  den += 18.86L * (tau * tau);
  den += 6.55L * (tau * tau * tau);
  const long double GM_ = mu * num / den;
  return 1.0 / (tau + 1) * GE + tau / (tau + 1) * GM_;
}

/* iD_P (src/MTensorPomeron.cc:2288-2298): Pomeron propagator */
static void tp_iD_P(const gra_cfg* c, double s, double t, t4_t T) {
  const long double alpha = 1.0L + c->tp_delta_P + (long double)c->tp_ap_P * t;
  const lcx base = lc_make(-0.0L, -1.0L * s * c->tp_ap_P); /* -i s ap */
  const lcx F = lc_scale(lc_pow_real(base, alpha - 1.0L), 1.0L / (4.0L * s));
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v)
      for (int k = 0; k < 4; ++k)
        for (int l = 0; l < 4; ++l)
          T[u][v][k][l] =
              lc_scale(F, GM[u][k] * GM[v][l] + GM[u][l] * GM[v][k] - 0.5L * GM[u][v] * GM[k][l]);
}

/* iG_PppHE (src/MTensorPomeron.cc:1388-1399): high-energy proton-Pomeron-proton vertex */
static void tp_iG_PppHE(const gra_cfg* c, v4 prime, v4 p, t2_t T) {
  const v4 ps = v4_add(prime, p);
  const long double F = -3.0L * c->tp_gPNN * kelly_f1(ld_m2(v4_sub(prime, p)), c->tp_mp_ff);
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v) T[u][v] = lc_make(0.0L, F * (v4_lo(ps, u) * v4_lo(ps, v)));
}

/* Gamma0 and Gamma2 (src/MTensorPomeron.cc:2500-2529) */
static long double tp_gamma0(v4 k1, v4 k2, int u, int v, int k, int l) {
  const long double k1k2 = ld_dot(k1, k2);
  return (k1k2 * GM[u][v] - v4_lo(k2, u) * v4_lo(k1, v)) *
         (v4_lo(k1, k) * v4_lo(k2, l) + v4_lo(k2, k) * v4_lo(k1, l) - 0.5L * k1k2 * GM[k][l]);
}
static long double tp_gamma2(v4 k1, v4 k2, int u, int v, int k, int l) {
  const long double k1k2 = ld_dot(k1, k2);
  return k1k2 * (GM[u][k] * GM[v][l] + GM[u][l] * GM[v][k]) +
         GM[u][v] * (v4_lo(k1, k) * v4_lo(k2, l) + v4_lo(k2, k) * v4_lo(k1, l)) -
         v4_lo(k1, v) * v4_lo(k2, l) * GM[u][k] - v4_lo(k1, v) * v4_lo(k2, k) * GM[u][l] -
         v4_lo(k2, u) * v4_lo(k1, l) * GM[v][k] - v4_lo(k2, u) * v4_lo(k1, k) * GM[v][l] -
         (k1k2 * GM[u][v] - v4_lo(k2, u) * v4_lo(k1, v)) * GM[k][l];
}

/* iG_Pvv (src/MTensorPomeron.cc:2224-2237) with FM (2486), LAMBDA^2 = 0.5 */
static void tp_iG_Pvv(v4 prime, v4 p, long double g1, long double g2, t4_t T) {
  const v4 mp = v4_scale(p, -1.0);
  const long double FM = 1.0L / (1.0L - ld_m2(v4_sub(prime, p)) / 0.5L);
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v)
      for (int k = 0; k < 4; ++k)
        for (int l = 0; l < 4; ++l)
          T[u][v][k][l] = lc_make(0.0L, FM * (2.0L * g1 * tp_gamma0(prime, mp, u, v, k, l) -
                                               g2 * tp_gamma2(prime, mp, u, v, k, l)));
}

/* iD_V (src/MTensorPomeron.cc:2353-2369): vector propagator with naive Reggeization */
static void tp_iD_V(v4 p, long double M0, long double s34, t2_t T) {
  const long double PIL = 3.141592653589793238462643383279502884L;
  const long double m2 = ld_m2(p);
  const long double alpha = 0.5L + 0.9L * m2;
  const long double s_th = 4 * (M0 * M0);
  const long double phi = PIL / 2.0L * expl((s_th - s34) / s_th) - PIL / 2.0L;
  const lcx base = lc_scale(lc_make(cosl(phi), sinl(phi)), s34 / s_th);
  const lcx R = lc_pow_real(base, alpha - 1.0L);
  const long double den = m2 - M0 * M0;
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v) T[u][v] = lc_scale(R, -GM[u][v] / den);
}

/* MDirac::EpsMassiveSpin1 / EpsSpin1 (src/MDirac.cc:393-456), conjugated, index up,
 * helicities -1, 0, +1 (MassiveSpin1States, src/MDirac.cc:661-679) */
static void tp_eps_conj(v4 k, lcx e[3][4]) {
  const long double pt = sqrtl((long double)k.x * k.x + (long double)k.y * k.y);
  const long double theta = (k.x == 0.0 && k.y == 0.0 && k.z == 0.0) ? 0.0L : atan2l(pt, k.z);
  const long double phi = (k.x == 0.0 && k.y == 0.0) ? 0.0L : atan2l(k.y, k.x);
  const long double N = 1.0L / sqrtl(2.0L);
  for (int h = 0; h < 3; ++h) {
    const int m = h - 1;
    if (m != 0) {
      const long double neg = -(long double)m, pos = (long double)m;
      e[h][0] = lc_make(0.0L, 0.0L);
      e[h][1] = lc_make((neg * cosl(theta) * cosl(phi)) * N, -(sinl(phi) * N));
      e[h][2] = lc_make((neg * cosl(theta) * sinl(phi)) * N, (cosl(phi) * N));
      e[h][3] = lc_make(pos * sinl(theta) * N, 0.0L);
    } else {
      const long double E = k.e, M = sqrtl(ld_m2(k));
      const long double p3 = sqrtl((long double)k.x * k.x + (long double)k.y * k.y +
                                   (long double)k.z * k.z);
      e[h][0] = lc_make(p3 / M, 0.0L);
      e[h][1] = lc_make(E * sinl(theta) * cosl(phi) / M, 0.0L);
      e[h][2] = lc_make(E * sinl(theta) * sinl(phi) / M, 0.0L);
      e[h][3] = lc_make(E * cosl(theta) / M, 0.0L);
    }
  }
}

/* sum_{a,b} G(a,b) D(a,b,k,l) -> X(k,l) (first pair contracted) */
static void tp_vd_first(t2_t G, t4_t Dp, t2_t X) {
  for (int k = 0; k < 4; ++k)
    for (int l = 0; l < 4; ++l) {
      lcx s = lc_make(0, 0);
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) s = lc_add(s, lc_mul(G[a][b], Dp[a][b][k][l]));
      X[k][l] = s;
    }
}
/* sum_{a,b} G(a,b) D(k,l,a,b) -> X(k,l) (second pair contracted) */
static void tp_vd_second(t2_t G, t4_t Dp, t2_t X) {
  for (int k = 0; k < 4; ++k)
    for (int l = 0; l < 4; ++l) {
      lcx s = lc_make(0, 0);
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) s = lc_add(s, lc_mul(G[a][b], Dp[k][l][a][b]));
      X[k][l] = s;
    }
}
/* sum_{a,b} X(a,b) V(r,q,a,b) -> Y(r,q) */
static void tp_xv(t2_t X, t4_t V, t2_t Y) {
  for (int r = 0; r < 4; ++r)
    for (int q = 0; q < 4; ++q) {
      lcx s = lc_make(0, 0);
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) s = lc_add(s, lc_mul(X[a][b], V[r][q][a][b]));
      Y[r][q] = s;
    }
}

/* Helicity amplitudes of pp -> p V V p (ME4, 2xV), written to hamp in the
 * reference's order: 4 proton-helicity replicas (ha = h1, hb = h2, identical in
 * the high-energy vertex) x 9 (h3, h4) pairs. Returns 36. */
/* M(r3, r4) = -i (M_t + M_u)(r3, r4) of the 2xV block (indices down) */
static void tensor_mtensor(const ora_run* R, v4 pa, v4 pb, const kin_t* K, t2_t Mout) {
  const gra_cfg* c = &R->c;
  const v4 p1 = K->pf[1], p2 = K->pf[2];
  const v4 p3 = K->dt[0], p4 = K->dt[1];
  const long double M_ = c->m_prop;
  const long double g1 = c->tp_gPvv[0], g2 = c->tp_gPvv[1];
  const v4 pt = v4_sub(v4_sub(pa, p1), p3);
  const v4 pu = v4_add(v4_sub(p4, pa), p1);
  t4_t W[8]; /* D13, D24, D14, D23, GtA, GtB, GuA, GuB */
  tp_iD_P(c, K->ss[1][3], K->t1, W[0]);
  tp_iD_P(c, K->ss[2][4], K->t2, W[1]);
  tp_iD_P(c, K->ss[1][4], K->t1, W[2]);
  tp_iD_P(c, K->ss[2][3], K->t2, W[3]);
  t2_t G1a, G2b;
  tp_iG_PppHE(c, p1, pa, G1a);
  tp_iG_PppHE(c, p2, pb, G2b);
  tp_iG_Pvv(pt, v4_scale(p3, -1.0), g1, g2, W[4]);
  tp_iG_Pvv(p4, pt, g1, g2, W[5]);
  tp_iG_Pvv(p4, pu, g1, g2, W[6]);
  tp_iG_Pvv(pu, v4_scale(p3, -1.0), g1, g2, W[7]);
  const long double s34 = ld_m2(K->pf[0]);
  t2_t DVt, DVu;
  tp_iD_V(pt, M_, s34, DVt);
  tp_iD_V(pu, M_, s34, DVu);
  /* PARAM_REGGE::Meson_FF, pow form (offshellFF = 2) or as configured */
  const long double M2_ = M_ * M_;
  long double fft, ffu;
  if (c->offshellFF == 2) {
    fft = 1.0L / (1.0L - (ld_m2(pt) - M2_) / c->b_POW);
    ffu = 1.0L / (1.0L - (ld_m2(pu) - M2_) / c->b_POW);
  } else {
    fft = meson_ff(c, v4_m2(pt), (double)M2_);
    ffu = meson_ff(c, v4_m2(pu), (double)M2_);
  }
  const long double fft2 = fft * fft, ffu2 = ffu * ffu;
  t2_t X, A, B, Mt, Mu, T;
  /* t-channel: A(r1, r3) = G1a(m1,n1) D13(m1,n1,a1,b1) GtA(r1,r3,a1,b1) */
  tp_vd_first(G1a, W[0], X);
  tp_xv(X, W[4], A);
  /* B(r4, r2) = G2b(m2,n2) D24(a2,b2,m2,n2) GtB(r4,r2,a2,b2) */
  tp_vd_second(G2b, W[1], X);
  tp_xv(X, W[5], B);
  /* Mt(r3, r4) = A(r1, r3) DVt(r1, r2) B(r4, r2) ff^2 */
  for (int r3 = 0; r3 < 4; ++r3)
    for (int r2 = 0; r2 < 4; ++r2) {
      lcx s = lc_make(0, 0);
      for (int r1 = 0; r1 < 4; ++r1) s = lc_add(s, lc_mul(A[r1][r3], DVt[r1][r2]));
      T[r3][r2] = s;
    }
  for (int r3 = 0; r3 < 4; ++r3)
    for (int r4 = 0; r4 < 4; ++r4) {
      lcx s = lc_make(0, 0);
      for (int r2 = 0; r2 < 4; ++r2) s = lc_add(s, lc_mul(T[r3][r2], B[r4][r2]));
      Mt[r3][r4] = lc_scale(s, fft2);
    }
  /* u-channel: A(r4, r1) = G1a D14 GuA(r4,r1,..); B(r2, r3) = G2b D23 GuB(r2,r3,..) */
  tp_vd_first(G1a, W[2], X);
  tp_xv(X, W[6], A);
  tp_vd_second(G2b, W[3], X);
  tp_xv(X, W[7], B);
  /* Mu(r3, r4) = A(r4, r1) DVu(r1, r2) B(r2, r3) ff^2 */
  for (int r4 = 0; r4 < 4; ++r4)
    for (int r2 = 0; r2 < 4; ++r2) {
      lcx s = lc_make(0, 0);
      for (int r1 = 0; r1 < 4; ++r1) s = lc_add(s, lc_mul(A[r4][r1], DVu[r1][r2]));
      T[r4][r2] = s;
    }
  for (int r4 = 0; r4 < 4; ++r4)
    for (int r3 = 0; r3 < 4; ++r3) {
      lcx s = lc_make(0, 0);
      for (int r2 = 0; r2 < 4; ++r2) s = lc_add(s, lc_mul(T[r4][r2], B[r2][r3]));
      Mu[r3][r4] = lc_scale(s, ffu2);
    }
  for (int r3 = 0; r3 < 4; ++r3)
    for (int r4 = 0; r4 < 4; ++r4) {
      const lcx m = lc_add(Mt[r3][r4], Mu[r3][r4]);
      Mout[r3][r4] = lc_make(m.im, -m.re); /* (-i) m */
    }
}

static int tensor_me4_2xv(const ora_run* R, v4 pa, v4 pb, const kin_t* K, cx* hamp) {
  t2_t M;
  tensor_mtensor(R, pa, pb, K, M);
  /* helicity amplitudes eps3*(r3) eps4*(r4) M(r3, r4) (MassiveSpin1PolSum) */
  lcx e3[3][4], e4[3][4];
  tp_eps_conj(K->dt[0], e3);
  tp_eps_conj(K->dt[1], e4);
  lcx h9[9];
  for (int h3 = 0; h3 < 3; ++h3)
    for (int h4 = 0; h4 < 3; ++h4) {
      lcx s = lc_make(0, 0);
      for (int r3 = 0; r3 < 4; ++r3)
        for (int r4 = 0; r4 < 4; ++r4)
          s = lc_add(s, lc_mul(lc_mul(e3[h3][r3], e4[h4][r4]), M[r3][r4]));
      h9[3 * h3 + h4] = s;
    }
  for (int rep = 0; rep < 4; ++rep)
    for (int q = 0; q < 9; ++q) hamp[9 * rep + q] = cx_make((double)h9[q].re, (double)h9[q].im);
  return 36;
}

/* iG_Ppsps (src/MTensorPomeron.cc:2203-2215): Pomeron-pseudoscalar-pseudoscalar
 * vertex -i 2 g FM((k' - k)^2) [(k' + k)_u (k' + k)_v - g_uv (k' + k)^2 / 4] */
static void tp_iG_Ppsps(v4 prime, v4 p, long double g1, t2_t T) {
  const v4 ps = v4_add(prime, p);
  const long double M2 = ld_m2(ps);
  const long double FM = 1.0L / (1.0L - ld_m2(v4_sub(prime, p)) / 0.5L);
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v)
      T[u][v] = lc_make(0.0L, -2.0L * g1 * FM *
                                  (v4_lo(ps, u) * v4_lo(ps, v) - 0.25L * GM[u][v] * M2));
}
/* sum_{a,b} X(a,b) Y(a,b) */
static lcx tp_full(t2_t X, t2_t Y) {
  lcx s = lc_make(0, 0);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) s = lc_add(s, lc_mul(X[a][b], Y[a][b]));
  return s;
}

/* MTensorPomeron::ME4 for a pseudoscalar pair ("2xS", src/MTensorPomeron.cc:726-763):
 * the literal index sums of G1a D13 GtA . GtB D24 G2b (t-channel) and of the
 * u-channel, with iD_MES0 = i / (p^2 - M^2) (2375-2377) and Meson_FF^2; four
 * identical proton-helicity replicas (ha = h1, hb = h2). Returns 4. */
static int tensor_me4_2xs(const ora_run* R, v4 pa, v4 pb, const kin_t* K, cx* hamp) {
  const gra_cfg* c = &R->c;
  const v4 p1 = K->pf[1], p2 = K->pf[2];
  const v4 p3 = K->dt[0], p4 = K->dt[1];
  const long double M_ = c->m_prop;
  const long double gP = (c->m_central < 0.2) ? c->tp_gPpipi : c->tp_gPKK;
  const v4 pt = v4_sub(v4_sub(pa, p1), p3);
  const v4 pu = v4_add(v4_sub(p4, pa), p1);
  const v4 mp3 = v4_scale(p3, -1.0);
  t4_t D;
  t2_t G1a, G2b, Ga, Gb, X;
  tp_iG_PppHE(c, p1, pa, G1a);
  tp_iG_PppHE(c, p2, pb, G2b);
  const long double M2_ = M_ * M_;
  long double ff[2];
  const v4 ex[2] = {pt, pu};
  for (int k = 0; k < 2; ++k)
    ff[k] = (c->offshellFF == 2) ? 1.0L / (1.0L - (ld_m2(ex[k]) - M2_) / c->b_POW)
                                 : meson_ff(c, v4_m2(ex[k]), (double)M2_);
  lcx M[2];
  for (int ch = 0; ch < 2; ++ch) {
    /* t: iG_Ppsps(pt, -p3), iG_Ppsps(p4, pt); u: iG_Ppsps(p4, pu), iG_Ppsps(pu, -p3) */
    if (ch == 0) {
      tp_iG_Ppsps(pt, mp3, gP, Ga);
      tp_iG_Ppsps(p4, pt, gP, Gb);
    } else {
      tp_iG_Ppsps(p4, pu, gP, Ga);
      tp_iG_Ppsps(pu, mp3, gP, Gb);
    }
    tp_iD_P(c, ch == 0 ? K->ss[1][3] : K->ss[1][4], K->t1, D);
    tp_vd_first(G1a, D, X);
    const lcx A = tp_full(X, Ga);
    tp_iD_P(c, ch == 0 ? K->ss[2][4] : K->ss[2][3], K->t2, D);
    tp_vd_second(G2b, D, X);
    const lcx B = tp_full(Gb, X);
    const long double den = ld_m2(ex[ch]) - M2_;
    const lcx iD = lc_make(0.0L, 1.0L / den); /* iD_MES0 */
    M[ch] = lc_scale(lc_mul(lc_mul(A, B), iD), ff[ch] * ff[ch]);
  }
  const lcx m = lc_add(M[0], M[1]);
  for (int rep = 0; rep < 4; ++rep) hamp[rep] = cx_make((double)m.im, (double)-m.re); /* -i m */
  return 4;
}

/* Spin mode by the central particles' spins (src/MTensorPomeron.cc:643-654) */
static int tensor_me4(const ora_run* R, const kin_t* K, cx* hamp) {
  if (R->c.central_spin2 == 0) return tensor_me4_2xs(R, R->pbeam1, R->pbeam2, K, hamp);
  return tensor_me4_2xv(R, R->pbeam1, R->pbeam2, K, hamp);
}

static void tensor_kin(const double* p, v4 q[6], kin_t* K) {
  for (int i = 0; i < 6; ++i) q[i] = v4_make(p[4 * i + 1], p[4 * i + 2], p[4 * i + 3], p[4 * i]);
  memset(K, 0, sizeof(*K));
  K->pf[1] = q[2];
  K->pf[2] = q[3];
  K->pf[3] = q[4];
  K->pf[4] = q[5];
  K->pf[0] = v4_add(q[4], q[5]);
  K->dt[0] = q[4];
  K->dt[1] = q[5];
  for (int i = 0; i <= 4; ++i)
    for (int j = 0; j <= 4; ++j) K->ss[i][j] = v4_m2(v4_add(K->pf[i], K->pf[j]));
  K->t1 = v4_m2(v4_sub(q[0], q[2]));
  K->t2 = v4_m2(v4_sub(q[1], q[3]));
}

/* The amplitude tensor M(r3, r4) (indices down, (t, x, y, z) order, 4 x 4
 * complex as re, im pairs) at caller momenta. Test entry point. */
ORA_EXPORT void ora_tensor_mtensor(ora_run* R, const double* p, double* out) {
  v4 q[6];
  kin_t K;
  tensor_kin(p, q, &K);
  t2_t M;
  tensor_mtensor(R, q[0], q[1], &K, M);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      out[2 * (4 * a + b)] = (double)M[a][b].re;
      out[2 * (4 * a + b) + 1] = (double)M[a][b].im;
    }
}

/* Helicity amplitudes (9, one replica) at caller momenta pa, pb, p1, p2, p3, p4
 * (6 x (E, x, y, z)); invariants are formed from them. Test entry point. */
ORA_EXPORT int ora_tensor_amplitudes(ora_run* R, const double* p, double* hamp_reim) {
  v4 q[6];
  kin_t K;
  tensor_kin(p, q, &K);
  cx h[36];
  tensor_me4_2xv(R, q[0], q[1], &K, h);
  for (int k = 0; k < 9; ++k) {
    hamp_reim[2 * k] = h[k].re;
    hamp_reim[2 * k + 1] = h[k].im;
  }
  return 9;
}

/* form::F1 with the Kelly fits at Q2 (test access) */
ORA_EXPORT double ora_kelly_f1(double Q2, double mp) { return (double)kelly_f1(Q2, mp); }

/* PARAM_REGGE::ResonanceFormFactor, form::CohFlux and the elastic leg of
 * MRegge::PhotoProp (out = (re, im)) for tests */
ORA_EXPORT double ora_res_ff(double s_hat, double M2, double S0) { return o_res_ff(s_hat, M2, S0); }
ORA_EXPORT double ora_coh_flux(double xi, double t, double pt, double mp) {
  return o_coh_flux(xi, t, pt, mp);
}
ORA_EXPORT void ora_photo_prop(const gra_cfg* c, double s, double t, double m2, double* out) {
  const cx z = o_photo_prop(c, s, t, m2);
  out[0] = z.re;
  out[1] = z.im;
}

/* form::CBW of resonance q at m2 with decay masses mA, mB: out = (re, im) */
ORA_EXPORT void ora_cbw(const gra_res* q, double m2, double mA, double mB, double* out) {
  const cx z = o_cbw(q, m2, mA, mB);
  out[0] = z.re;
  out[1] = z.im;
}

/* The 2xS amplitude -i (M_t + M_u) (one of the four identical replicas) at caller
 * momenta pa, pb, p1, p2, p3, p4 (6 x (E, x, y, z)). Test entry point. */
ORA_EXPORT void ora_tensor_2xs_amplitude(ora_run* R, const double* p, double* out) {
  v4 q[6];
  kin_t K;
  tensor_kin(p, q, &K);
  cx h[4];
  tensor_me4_2xs(R, q[0], q[1], &K, h);
  out[0] = h[0].re;
  out[1] = h[0].im;
}

/* Phase-space part of the weight, BNPhaseSpaceWeight x BNIntegralVolume
 * (src/MContinuum.cc:619-656), at tree-level kinematics, without cuts; 0 if the
 * kinematics fail. */
ORA_EXPORT double ora_phase_space_weight(ora_run* R, const double* x) {
  const gra_cfg* c = &R->c;
  kin_t K;
  memset(&K, 0, sizeof(K));
  if (!build_kin(R, x, &K)) return 0.0;
  const double J = 1.0 / fabs(K.pf[1].z / K.pf[1].e - K.pf[2].z / K.pf[2].e);
  const double psw = (1.0 / pow(2, 4)) * (1.0 / pow(2.0 * PI, 8)) *
                     (v4_pt(K.pf[1]) / (2.0 * K.pf[1].e)) * (v4_pt(K.pf[2]) / (2.0 * K.pf[2].e)) *
                     J * (1.0 * v4_pt(K.pkt));
  const double lspan = log(c->fwd_pt_max) - log(c->fwd_pt_min + 1e-12);
  const double vol = pow(c->kt_max - c->kt_min, 1) * pow(2.0 * PI, 1) *
                     pow(c->rap_max - c->rap_min, 2) *
                     ((2.0 * PI) * (2.0 * PI) * (v4_pt(K.pf[1]) * v4_pt(K.pf[2]) * (lspan * lspan)));
  return psw * vol;
}

/* Spin factors of resonance r at hypercube point x (test access):
 * out[0] = |ProdAmp| / g_A, out[1] = |DecayAmp| / g_decay. Returns 1 if the
 * kinematics are valid. */
ORA_EXPORT int ora_res_spin(ora_run* R, const double* x, int r, double* out) {
  kin_t K;
  memset(&K, 0, sizeof(K));
  if (!build_kin(R, x, &K)) return 0;
  double thR, phR;
  o_rho_rotation(R, &K, &thR, &phR);
  const gra_res* q = &R->c.res[r];
  const cx pa = o_prod_amp(R, q, &K, thR, phR), da = o_decay_amp(R, q, &K, thR, phR);
  out[0] = sqrt(cx_abs2(pa)) / q->g_A;
  out[1] = sqrt(cx_abs2(da)) / o_res_gdecay(R, q);
  return 1;
}
