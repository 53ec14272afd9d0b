// Host-side configuration and VEGAS grid arithmetic of the C ABI.
//
// gra_hip_cfg_default fills the preset workloads with the TUNE0 model values
// (modeldata/TUNE0/GENERAL.json), the PDG table (modeldata/mass_width_2018.mcd)
// and the eikonal numerics (modeldata/NUMERICS.json, include/Graniitti/MEikonal.h:26-57).
// The grid routines restate VEGASData (include/Graniitti/MVEGAS.h); they run on
// the host once per iteration over BINS x D = 128 x 8 numbers.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>

#include "gra_internal.h"

namespace gra_hip {

namespace {

void set_common(gra_cfg* c) {
  std::memset(c, 0, sizeof(*c));
  c->abi_size = (int32_t)sizeof(gra_cfg);

  // Beams: PROCESSPARAM BEAM ["p+","p+"], ENERGY [6500, 6500] (tests/processes/ALICECUTS.json)
  c->beam_E1 = 6500.0;
  c->beam_E2 = 6500.0;
  c->m_beam = 0.938272081;  // p, mass_width_2018.mcd:169

  // MProcess::OFFSHELL default (include/Graniitti/MProcess.h:363)
  c->offshell = 5.0;

  // GENCUT defaults and technical boundaries (src/MProcess.cc:1369-1386,
  // src/MContinuum.cc:71-82 with Kf = 2)
  c->fwd_pt_min = 0.0;
  c->fwd_pt_max = 2.5;
  c->kt_min = 0.0;
  c->kt_max = 10.0 / 2.0;

  // VETOCUT: off; unused domains hold VETODOMAIN's defaults
  // (include/Graniitti/MKinematics.h:1613-1626)
  c->veto_active = 0;
  c->n_veto = 0;
  for (int d = 0; d < GRA_MAX_VETO; ++d) {
    c->veto_eta_min[d] = -30.0;
    c->veto_eta_max[d] = 30.0;
    c->veto_pt_min[d] = 0.0;
    c->veto_pt_max[d] = 1000000.0;
  }

  // FIDCUT (ALICECUTS.json values where the presets state no other)
  c->fid_active = 1;
  c->fid_rap_min = -10.0;
  c->fid_rap_max = 10.0;
  c->fid_et_min = 0.0;
  c->fid_et_max = 100000.0;
  c->fid_pt_max = 100000.0;
  c->sys_M_min = 0.0;
  c->sys_M_max = 100000.0;
  c->sys_Y_min = -10.0;
  c->sys_Y_max = 10.0;
  c->sys_Pt_min = 0.0;
  c->sys_Pt_max = 100000.0;
  c->fwd_t_min = 0.0;
  c->fwd_t_max = 100000.0;
  c->fwd_M_min = 0.0;
  c->fwd_M_max = 100000.0;

  // PARAM_SOFT
  c->delta_P = 0.098;
  c->alpha_P = 0.065;
  c->gN_P = 8.20;
  c->fc1 = 0.34;
  c->fc2 = 5.0;
  c->fc3 = 10.0;
  c->mpi_const = 0.13957018;  // PDG::mpi, include/Graniitti/MPDG.h:56

  // PARAM_REGGE
  const double a0[3] = {1.088, 0.55, 0.55};
  const double ap[3] = {0.20, 0.93, 0.93};
  const double sg[3] = {1.0, -1.0, 1.0};
  for (int k = 0; k < 3; ++k) {
    c->regge_a0[k] = a0[k];
    c->regge_ap[k] = ap[k];
    c->regge_sgn[k] = sg[k];
  }
  c->regge_s0 = 1.0;
  c->offshellFF = 2;
  c->b_EXP = 0.944;
  c->a_OREAR = 0.638;
  c->b_OREAR = 1.886;
  c->b_POW = 0.500;

  // PARAM_TENSORPOM
  c->tp_delta_P = 0.0808;
  c->tp_ap_P = 0.25;
  c->tp_gPNN = 1.87;
  c->tp_gPvv[0] = 0.70;
  c->tp_gPvv[1] = 6.20;
  c->tp_mp_ff = 0.938272081;
  c->tp_gPpipi = 1.76;
  c->tp_gPKK = 1.54;

  // Eikonal numerics: NumberBT = NumberKT2 = 10000 (NUMERICS.json),
  // FBIntegralN = 10000, MinLoopKT = 1e-4 (MEikonal.h:37-40)
  c->n_bt = 10000;
  c->n_kt2 = 10000;
  c->n_fb = 10000;
  c->min_loop_kt = 1e-4;
  c->max_loop_kt = 3.0;
  c->n_loop_phi = 15;
  c->n_loop_kt = 30;
}

void set_pipi(gra_cfg* c) {
  c->process = GRA_PROC_REGGE_CON;
  c->dim = 8;  // 3 (Kf + 2) - 4, src/MContinuum.cc:85
  c->m_central = 0.13957061;  // pi+-, mass_width_2018.mcd:57
  c->w_central = 0.0;  // width 2.5284e-17 GeV moves the mass by < 1 ulp: pole mass
  c->m_daughter = 0.0;
  c->S_factor = 1.0;  // pi+ pi-
  c->m_prop = 0.13957061;
  c->central_spin2 = 0;
  // PARAM_PI couplings and active trajectories
  c->regge_c[0] = 35.00184;
  c->regge_c[1] = 81.63672;
  c->regge_c[2] = 10.86264;
  c->regge_on[0] = 1;
  c->regge_on[1] = 0;
  c->regge_on[2] = 0;
  c->rap_min = -0.9;
  c->rap_max = 0.9;
  c->fid_eta_min = -0.9;
  c->fid_eta_max = 0.9;
  c->fid_pt_min = 0.15;
}

// PP[CON]<F>: system rapidity and mass generation ranges, GENCUTS "<F>" of
// tests/run_compare_PS/card.json; 8 dimensions: the B51 variables (pt_1, pt_2,
// phi_1, phi_2, Y, M^2) and the isotropic decay angles (cos theta, phi) that
// TwoBodyPhaseSpace draws (include/Graniitti/MKinematics.h:405-459), lifted
void set_factorized(gra_cfg* c) {
  c->phase_space = GRA_PS_F;
  c->dim = 8;  // LIPSDIM = 5 + 1 (src/MFactorized.cc:57), + 2 decay angles
  c->rap_min = -0.9;
  c->rap_max = 0.9;
  c->gen_M_min = 0.0;
  c->gen_M_max = 4.0;
}

void set_rhorho(gra_cfg* c) {
  c->process = GRA_PROC_TENSOR_CON;
  c->dim = 14;  // 8 phase-space + 2 BW masses + 2 x (cos theta, phi) decay angles
  c->m_central = 0.775;
  c->w_central = 0.149;
  c->m_daughter = 0.13957061;
  c->S_factor = 2.0;  // rho0 rho0
  c->m_prop = 0.775;
  c->central_spin2 = 2;
  // PARAM_RHO couplings (ansatz), used only if the Regge amplitude is selected
  c->regge_c[0] = 25.0;
  c->regge_on[0] = 1;
  c->rap_min = -2.5;
  c->rap_max = 2.5;
  c->fid_eta_min = -2.5;
  c->fid_eta_max = 2.5;
  c->fid_pt_min = 0.1;
}

// One resonance: RES_<name>.json (modeldata/TUNE0) and its pi+ pi- entry of
// BRANCHING.json (BR, alpha_ls = [[J, 0, 1, 0]]); diagonal spin density
void add_res(gra_cfg* c, int pdg, double M, double W, int J, double g_A, double g_phi, double BR,
             double rho_pm) {
  gra_res& r = c->res[c->n_res++];
  r.pdg = pdg;
  r.mass = M;
  r.width = W;
  r.J = J;
  r.P = (J & 1) ? -1 : 1;
  r.BW = GRA_BW_RUNNING;
  r.g_A = g_A;
  r.g_phi = g_phi;
  r.g_FF = 1.0;
  r.BR = BR;
  r.alpha_re = 1.0;
  r.alpha_im = 0.0;
  const int n = 2 * J + 1;
  if (J == 0) r.rho_re[0] = 1.0;
  else {
    r.rho_re[0] = rho_pm;                  // m = -J
    r.rho_re[(n - 1) * n + (n - 1)] = rho_pm;  // m = +J
  }
}

// PP[RES+CON]<C> -> pi+ pi- of tests/processes/ALICECUTS.json: the continuum of
// preset 1 and the eight resonances of its "RES" list. PROC_17 sums
// lts.RESONANCES, a std::map keyed by name (include/Graniitti/MKinematics.h:1467),
// so in name order. PARAM_SPIN: SPINGEN, SPINDEC on, FRAME "CS"; PARAM_REGGE omega 0.5.
void set_res_con(gra_cfg* c) {
  c->process = GRA_PROC_RES_CON;
  c->spin_gen = 1;
  c->spin_dec = 1;
  c->spin_frame = GRA_FRAME_CS;
  c->regge_omega = 0.5;
  c->n_res = 0;
  add_res(c, 9030221, 1.506, 0.112, 0, 0.03, -3.14159, 0.2299, 0.0);     // f0_1500
  add_res(c, 10331, 1.704, 0.123, 0, 0.03, -3.14159, 0.0552, 0.0);       // f0_1710
  add_res(c, 9000221, 0.5, 0.4, 0, 0.05, -3.14159, 1.0, 0.0);            // f0_500
  add_res(c, 9010221, 0.98, 0.08, 0, 0.065, -3.14159, 0.5, 0.0);         // f0_980
  add_res(c, 225, 1.2755, 0.1867, 2, 0.7, -1.57, 0.562, 0.5);            // f2_1270
  add_res(c, 335, 1.525, 0.073, 2, 0.19, -1.57, 0.00546, 0.5);           // f2_1525
  add_res(c, 333, 1.019461, 0.004249, 1, 0.165, -2.36, 7.3e-05, 0.5);    // phi_1020
  add_res(c, 113, 0.77526, 0.1491, 1, 0.19, -2.36, 1.0, 0.5);            // rho_770
}

}  // namespace

}  // namespace gra_hip

using namespace gra_hip;

extern "C" int gra_hip_cfg_default(gra_cfg* cfg, int32_t preset) {
  if (!cfg) return set_global_error(GRA_E_ARG, "gra_hip_cfg_default: cfg is NULL");
  set_common(cfg);
  switch (preset) {
    case 1:
      set_pipi(cfg);
      cfg->screening = 0;
      cfg->amp_precision = GRA_AMP_FP64;
      break;
    case 2:
      set_pipi(cfg);
      cfg->screening = 1;
      cfg->amp_precision = GRA_AMP_FP32;
      break;
    case 3:
      set_rhorho(cfg);
      cfg->screening = 1;
      cfg->amp_precision = GRA_AMP_FP32;
      break;
    case 4:
    case 5:
      set_pipi(cfg);
      set_factorized(cfg);
      cfg->screening = preset == 5;
      cfg->amp_precision = preset == 5 ? GRA_AMP_FP32 : GRA_AMP_FP64;
      break;
    case 6:
    case 7:
      set_pipi(cfg);
      set_res_con(cfg);
      cfg->screening = preset == 7;
      cfg->amp_precision = GRA_AMP_FP64;
      break;
    case 8:
      // tests/run_JW_frames/run.sh: "PP[RES]<F> -> pi+ pi- @RES{f2_1270:1}
      // @R[f2_1270]{JZ0:0, JZ1:0, JZ2:1} @FRAME:CS"
      set_pipi(cfg);
      set_factorized(cfg);
      set_res_con(cfg);
      cfg->process = GRA_PROC_RES;
      cfg->n_res = 0;
      add_res(cfg, 225, 1.2755, 0.1867, 2, 0.7, -1.57, 0.562, 0.5);
      cfg->fid_active = 0;
      cfg->screening = 0;
      cfg->amp_precision = GRA_AMP_FP64;
      break;
    case 9:
    case 10:
      // "PP[CONTENSOR]<C> -> pi+ pi-": the tensor-pomeron continuum of a pion pair
      // (2xS, src/MTensorPomeron.cc:726-763) with preset 1's beams and cuts
      set_pipi(cfg);
      cfg->process = GRA_PROC_TENSOR_CON;
      cfg->screening = preset == 10;
      cfg->amp_precision = GRA_AMP_FP64;
      break;
    case 11:
      // the screened pi+ pi- workload in the loop discretisation stated for it: 32
      // radial x 16 azimuthal loop nodes, a 64-node log-spaced kt table of A_eik
      set_pipi(cfg);
      cfg->screening = 1;
      cfg->amp_precision = GRA_AMP_FP32;
      cfg->loop_rule = GRA_LOOP_GAUSS;
      cfg->n_loop_phi = 15;
      cfg->n_loop_kt = 31;
      cfg->eik_log_nodes = 64;
      break;
    default:
      return set_global_error(GRA_E_ARG, "gra_hip_cfg_default: unknown preset %d", preset);
  }
  return GRA_OK;
}

namespace gra_hip {

#define XM(g, i, j) ((g)->xmat[(j)*GRA_VEGAS_BINS + (i)])

// VEGASData::Rebin, include/Graniitti/MVEGAS.h:87-104
static void grid_rebin(gra_grid* g, double ac, int j, const double* rvec) {
  const int B = g->bins;
  unsigned int k = 0;
  double dr = 0.0, zn = 0.0, zo = 0.0;
  double xc[GRA_VEGAS_BINS];
  for (int i = 0; i < B - 1; ++i) {
    while (ac > dr) dr += rvec[(++k) - 1];
    if (k > 1) zo = XM(g, k - 2, j);
    zn = XM(g, k - 1, j);
    dr -= ac;
    xc[i] = zn - (zn - zo) * dr / rvec[k - 1];
  }
  for (int i = 0; i < B - 1; ++i) XM(g, i, j) = xc[i];
  XM(g, B - 1, j) = 1.0;
}

int grid_init(gra_grid* g, int dim, int bins) {
  if (!g) return set_global_error(GRA_E_ARG, "grid_init: grid is NULL");
  if (dim < 1 || dim > GRA_MAX_DIM)
    return set_global_error(GRA_E_ARG, "grid_init: dim %d outside [1, %d]", dim, GRA_MAX_DIM);
  if (bins < 2 || bins > GRA_VEGAS_BINS || (bins % 2) != 0)
    return set_global_error(GRA_E_ARG, "grid_init: bins %d must be even and <= %d", bins,
                            GRA_VEGAS_BINS);
  std::memset(g, 0, sizeof(*g));
  g->dim = dim;
  g->bins = bins;
  for (int j = 0; j < dim; ++j) XM(g, 0, j) = 1.0;
  double rvec[GRA_VEGAS_BINS];
  for (int i = 0; i < GRA_VEGAS_BINS; ++i) rvec[i] = 1.0;
  for (int j = 0; j < dim; ++j) grid_rebin(g, 1.0 / (double)bins, j, rvec);
  return GRA_OK;
}

// VEGASData::OptimizeGrid, include/Graniitti/MVEGAS.h:107-142
int grid_optimize(gra_grid* g, double* f2mat, double lambda) {
  const int B = g->bins;
  const double EPS = 1.0e-30;
  double dcache[GRA_MAX_DIM];
  double rvec[GRA_VEGAS_BINS];
#define F2(i, j) f2mat[(j)*GRA_VEGAS_BINS + (i)]
  for (int j = 0; j < g->dim; ++j) {
    double zo = F2(0, j), zn = F2(1, j), ac;
    F2(0, j) = (zo + zn) / 2.0;
    dcache[j] = F2(0, j);
    for (int i = 2; i < B; ++i) {
      ac = zo + zn;
      zo = zn;
      zn = F2(i, j);
      F2(i - 1, j) = (ac + zn) / 3.0;
      dcache[j] += F2(i - 1, j);
    }
    F2(B - 1, j) = (zo + zn) / 2.0;
    dcache[j] += F2(B - 1, j);
  }
  for (int j = 0; j < g->dim; ++j) {
    double ac = 0.0;
    for (int i = 0; i < B; ++i) {
      F2(i, j) = (F2(i, j) < EPS) ? EPS : F2(i, j);
      rvec[i] = std::pow((1.0 - F2(i, j) / dcache[j]) /
                             (std::log(dcache[j]) - std::log(F2(i, j)) + EPS),
                         lambda);
      if (std::isnan(rvec[i]) || std::isinf(rvec[i])) rvec[i] = EPS;
      ac += rvec[i];
    }
    grid_rebin(g, ac / B, j, rvec);
  }
#undef F2
  return GRA_OK;
}

}  // namespace gra_hip

extern "C" int gra_hip_grid_init(gra_grid* grid, int32_t dim, int32_t bins) {
  return grid_init(grid, dim, bins);
}

extern "C" int gra_hip_grid_optimize(gra_grid* grid, gra_iter_stats* st, double lambda) {
  if (!grid || !st) return set_global_error(GRA_E_ARG, "gra_hip_grid_optimize: NULL argument");
  return grid_optimize(grid, st->f2mat, lambda);
}

// ---------------------------------------------------------------------------
// VEGAS schedule (MGraniitti::SampleVegas + MGraniitti::VEGAS in integration
// mode, src/MGraniitti.cc:1074-1135, 1146-1338) over a caller-supplied pass.
// ---------------------------------------------------------------------------
namespace gra_hip {

// Iteration combination, src/MGraniitti.cc:1222-1251
void vegas_combine(double fsum, double f2sum, double calls, int iter, double* state,
                   double* sigma, double* sigma_err, double* chi2) {
  const double EPS = 1.0e-30;
  fsum *= 1.0 / calls;
  f2sum *= (1.0 / calls) * (1.0 / calls);
  const double I_this = fsum;
  f2sum = std::sqrt(std::max(0.0, f2sum * calls));
  f2sum = (f2sum - fsum) * (f2sum + fsum);
  if (f2sum <= 0.0) f2sum = EPS;
  const double I2_this = f2sum * 1.0 / calls;
  const double alpha = 1.0 / I2_this;
  state[0] += alpha * I_this;
  state[1] += alpha * (I_this * I_this);
  state[2] += alpha;
  *sigma = state[0] / state[2];
  *sigma_err = std::sqrt(std::max(0.0, 1.0 / state[2]));
  double c2 = (state[1] - (state[0] * state[0]) / state[2]) / (iter + 1E-5);
  *chi2 = (c2 < 0) ? 0.0 : c2;
}

// The run's guards after every iteration but the first of a VEGAS call
// (src/MGraniitti.cc:1258-1283): 0 = go on, 1 = restart (chi2 > 50),
// GRA_E_NUMERIC with a message for an integral that is NaN/Inf or < 1e-60.
static int vegas_guard(int iter, double sigma, double chi2) {
  if (iter == 0) return 0;
  if (std::isnan(sigma) || std::isinf(sigma))
    return set_global_error(GRA_E_NUMERIC,
                            "VEGAS: integral inf/nan - cuts or parameters need fixing");
  if (sigma < 1e-60)
    return set_global_error(GRA_E_NUMERIC,
                            "VEGAS: integral %g < 1e-60 - check generation and fiducial cuts",
                            sigma);
  return chi2 > kVegasChi2Restart ? 1 : 0;
}

int vegas_schedule(const gra_vegas_cfg* vc, gra_grid* grid, int dim, VegasPass pass, void* ctx,
                   gra_vegas_result* res) {
  if (!vc || !grid || !res) return set_global_error(GRA_E_ARG, "vegas_run: NULL argument");
  if (vc->ncall == 0 || vc->ncall > 0xFFFFFFFFull || vc->iter < 1 || vc->burnin_iter < 0)
    return set_global_error(GRA_E_ARG,
                            "vegas_run: 1 <= ncall < 2^32, iter >= 1, burnin_iter >= 0 required");
  std::memset(res, 0, sizeof(*res));
  res->seconds_to_precision = -1.0;
  std::unique_ptr<gra_iter_stats> st(new (std::nothrow) gra_iter_stats);
  if (!st) return set_global_error(GRA_E_NOMEM, "vegas_run: host stats");
  const auto t0 = std::chrono::steady_clock::now();
  auto since = [&t0]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  const int bins = vc->bins > 0 ? vc->bins : GRA_VEGAS_BINS;
  uint64_t ncall = vc->ncall;
  int burnin = vc->burnin_iter;
  uint32_t counter = 0;
  double state[3];
  double sigma = 0, err = 0, chi2 = 0;
  double t_main = 0.0;
  int done = 0;
  // CheckInfNan over the run (include/Graniitti/MProcess.h:317-340): the passes
  // report the NaN and Inf weights they set to zero
  uint64_t n_nan = 0, n_inf = 0;
  auto run_pass = [&](uint64_t calls) -> int {
    const int r = pass(ctx, grid, calls, vc->seed, counter++, st.get());
    if (r) return r;
    n_nan += st->n_nan;
    n_inf += st->n_inf;
    if (n_nan > GRA_MAX_NONFINITE)
      return set_global_error(GRA_E_NUMERIC,
                              "CheckInfNan: too many NaN weights (%llu) - check model parameters "
                              "and cuts", (unsigned long long)n_nan);
    if (n_inf > GRA_MAX_NONFINITE)
      return set_global_error(GRA_E_NUMERIC,
                              "CheckInfNan: too many Inf weights (%llu) - check model parameters "
                              "and cuts", (unsigned long long)n_inf);
    return GRA_OK;
  };
  for (;;) {
    // init = 0: burn-in from a fresh uniform grid (VEGASData::Init(0)); repeated
    // with 10 x the calls while an iteration's chi2 exceeds 50
    for (bool again = true; again;) {
      again = false;
      int r = grid_init(grid, dim, bins);
      if (r) return r;
      state[0] = state[1] = state[2] = 0;
      for (int it = 0; it < burnin; ++it) {
        r = run_pass(ncall);
        if (r) return r;
        vegas_combine(st->fsum, st->f2sum, (double)ncall, it, state, &sigma, &err, &chi2);
        res->calls_total += ncall;
        r = vegas_guard(it, sigma, chi2);
        if (r < 0) return r;
        if (r == 1) {
          if (res->restarts >= kVegasMaxRestarts || ncall * 10 > 0xFFFFFFFFull)
            return set_global_error(GRA_E_NUMERIC,
                                    "VEGAS: chi2 = %.1f > 50 after %d restarts (%llu calls)", chi2,
                                    res->restarts, (unsigned long long)ncall);
          ncall *= 10;
          ++res->restarts;
          again = true;
          break;
        }
        grid_optimize(grid, st->f2mat, vc->lambda);
      }
    }
    // init = 1: keep the grid, clear the integral (VEGASData::Init(1))
    t_main = since();
    state[0] = state[1] = state[2] = 0;
    res->maxW = res->maxf = 0.0;
    res->seconds_to_precision = -1.0;
    int itermin = vc->iter;
    const int itermax = std::max(vc->max_iter, vc->iter);
    done = 0;
    bool restart = false;
    for (int it = 0; it < itermin; ++it) {
      int r = run_pass(ncall);
      if (r) return r;
      vegas_combine(st->fsum, st->f2sum, (double)ncall, it, state, &sigma, &err, &chi2);
      res->calls_total += ncall;
      res->maxW = std::max(res->maxW, st->maxW);
      res->maxf = std::max(res->maxf, st->maxf);
      ++done;
      r = vegas_guard(it, sigma, chi2);
      if (r < 0) return r;
      if (r == 1) {
        // SampleVegas: twice the burn-in iterations, then integrate again
        if (res->restarts >= kVegasMaxRestarts)
          return set_global_error(GRA_E_NUMERIC, "VEGAS: chi2 = %.1f > 50 after %d restarts",
                                  chi2, res->restarts);
        burnin = std::max(1, 2 * burnin);
        ++res->restarts;
        restart = true;
        break;
      }
      if (vc->precision > 0 && res->seconds_to_precision < 0 && sigma > 0 &&
          err / sigma <= vc->precision)
        res->seconds_to_precision = since();
      if (it == itermin - 1 && itermin < itermax &&
          ((vc->chi2max > 0 && chi2 > vc->chi2max) ||
           (vc->precision > 0 && err / sigma > vc->precision)))
        ++itermin;
      grid_optimize(grid, st->f2mat, vc->lambda);
    }
    if (!restart) break;
  }
  res->sigma = sigma;
  res->sigma_err = err;
  res->chi2 = chi2;
  res->iterations = done;
  res->seconds_total = since();
  res->seconds = res->seconds_total - t_main;
  if (!std::isfinite(sigma) || sigma < 1e-60)
    return set_global_error(GRA_E_NUMERIC, "VEGAS: integral %g is not finite or < 1e-60", sigma);
  return GRA_OK;
}

}  // namespace gra_hip

extern "C" int gra_vegas_run_host(const gra_vegas_cfg* vc, gra_grid* grid, int32_t dim,
                                  gra_vegas_pass_fn pass, void* ctx, gra_vegas_result* res) {
  if (!pass) return set_global_error(GRA_E_ARG, "gra_vegas_run_host: pass is NULL");
  return vegas_schedule(vc, grid, dim, pass, ctx, res);
}
