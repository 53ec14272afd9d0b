// C ABI of the integrand and VEGAS integrator (declarations in include/gra_hip.h).
//
// Host responsibilities: run constants (every transcendental constant is taken
// from the host libm so it equals the reference's value), the eikonal-table
// pipeline around the GPU Fourier-Bessel kernel, buffer management, the VEGAS
// iteration schedule and grid update (MGraniitti::VEGAS, src/MGraniitti.cc:1146-1338).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>

#include "gra_internal.h"
#include "gra_kernels.h"

namespace gra_hip {

static thread_local std::string g_err;

static std::string vformat(const char* fmt, va_list ap) {
  char buf[1024];
  vsnprintf(buf, sizeof(buf), fmt, ap);
  return std::string(buf);
}

int set_global_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  g_err = vformat(fmt, ap);
  va_end(ap);
  return code;
}
const char* global_error() { return g_err.c_str(); }

int fail(gra_handle* h, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  const std::string m = vformat(fmt, ap);
  va_end(ap);
  g_err = m;
  if (h) h->err = m;
  return code;
}

int DevBuf::ensure(size_t need) {
  if (need <= bytes && p) return 0;
  release();
  if (need == 0) need = 16;
  if (hipMalloc(&p, need) != hipSuccess) {
    p = nullptr;
    bytes = 0;
    return -1;
  }
  bytes = need;
  return 0;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// ---------------------------------------------------------------------------
// Host restatement of the per-run scalar functions (same operations as the
// reference so the constants are bit-identical to its values)
// ---------------------------------------------------------------------------
namespace {

inline double msqrt(double x) { return std::sqrt(std::max(0.0, x)); }

struct hcx {
  double re, im;
};
inline hcx hexp(hcx a) {
  const double e = std::exp(a.re);
  return {e * std::cos(a.im), e * std::sin(a.im)};
}
inline hcx hmul(hcx a, hcx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// form::ReggeEta, src/MForm.cc:382-385
hcx regge_eta(double alpha_t, double sigma) {
  const double denom = std::sin(GRA_PI * alpha_t);
  const hcx e = hexp({-0.0 * GRA_PI * alpha_t, -GRA_PI * alpha_t});
  const hcx num = {1.0 + sigma * e.re, sigma * e.im};
  return {-num.re / denom, -num.im / denom};
}

double s3f(const gra_cfg& c, double t) { return (1.0 / (1 - t / c.fc1)) * (1.0 / (1 - t / c.fc2)); }

// form::S3HPL and form::S3PomAlpha, src/MForm.cc:402-426
double s3hpl(const gra_cfg& c, double tau, double t) {
  const double m = 1.0;
  const double sqrtau = msqrt(1 + tau);
  const double F_pi = 1.0 / (1.0 - t / c.fc3);
  const double mpi = c.mpi_const;
  return (4.0 / tau) * (F_pi * F_pi) *
         (2.0 * tau - std::pow(1.0 + tau, 3.0 / 2.0) * std::log((sqrtau + 1.0) / (sqrtau - 1.0)) +
          std::log((m * m) / (mpi * mpi)));
}
double pom_alpha(const gra_cfg& c, double t) {
  const double mpi = c.mpi_const;
  const double BETApi = (2.0 / 3.0) * c.gN_P;
  const double PIC = ((BETApi * BETApi) * (mpi * mpi)) / (32.0 * (GRA_PI * (GRA_PI * GRA_PI)));
  const double h = PIC * s3hpl(c, 4.0 * (mpi * mpi) / std::fabs(t), t);
  return 1.0 + c.delta_P + c.alpha_P * t - h;
}

// MEikonal::SingleAmpElastic(s, t, +1), src/MEikonal.cc:139-154
hcx single_amp_elastic(const gra_cfg& c, double s, double t) {
  const double F = s3f(c, t);
  const double aP = pom_alpha(c, t);
  const hcx eta = regge_eta(aP, 1);
  const double pre = ((c.gN_P * c.gN_P) * F) * F;
  const double pw = std::pow(s / 1.0, aP);
  return {(eta.re * pre) * pw, (eta.im * pre) * pw};
}

// IArray1D::Interpolate1D on rows (x, re, im), src/MEikonal.cc:489-512
hcx interp_rows(const std::vector<double>& T, int N, double MIN, double STEP, double a) {
  if (a < MIN) a = MIN;
  int i = (int)std::floor((a - MIN) / STEP);
  if (i < 0) i = 0;
  if (i >= N) i = N - 1;
  const double x0 = T[3 * i], x1 = T[3 * (i + 1)];
  const hcx y0 = {T[3 * i + 1], T[3 * i + 2]}, y1 = {T[3 * (i + 1) + 1], T[3 * (i + 1) + 2]};
  const double dx = x1 - x0;
  const hcx q = {(y1.re - y0.re) / dx, (y1.im - y0.im) / dx};
  return {y0.re + q.re * (a - x0), y0.im + q.im * (a - x0)};
}

// IArray1D::WriteArray / ReadArray round trip (std::setprecision(15), std::stod)
double round15(double v) {
  char buf[64];
  snprintf(buf, sizeof(buf), "%.15g", v);
  return std::strtod(buf, nullptr);
}

double gev2barn() {
  const double c = 2.99792458E8, hbar = 6.582119514E-25;
  const double GeV2m = hbar * c;
  const double GeV2m2 = GeV2m * GeV2m;
  return GeV2m2 / 1E-28;
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton iteration on P_n from the
// Chebyshev-like first guesses, as in Numerical Recipes' gauleg)
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(GRA_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int k = 0; k < n; ++k) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * k + 1.0) * z * p2 - k * p3) / (k + 1.0);
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-15) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

double simpson38_w(int i, int N) {
  if (i == 0 || i == N) return 1.0;
  return (i % 3 == 0) ? 2.0 : 3.0;
}

// Spin density of a resonance with J > 0 (spin::Positivity, src/MSpin.cc:988-1110, called
// from form::ReadResonance, src/MForm.cc:361): |Tr rho| within 1e-2 of 1 as the
// reference requires; Hermitian and without negative eigenvalue. The reference's
// eigenvalue test compares against -1e6 (src/MSpin.cc:1012), so it accepts any
// matrix; here a negative eigenvalue below -1e-9 is rejected. The eigenvalues
// come from the real symmetric form [[Re, -Im], [Im, Re]] (each eigenvalue of
// rho twice) by cyclic Jacobi rotations.
int check_rho(const gra_res& q, int r) {
  if (q.J == 0) return GRA_OK;
  const int n = 2 * q.J + 1;
  double tr_re = 0.0, tr_im = 0.0;
  for (int a = 0; a < n; ++a) {
    tr_re += q.rho_re[a * n + a];
    tr_im += q.rho_im[a * n + a];
  }
  const double tr = std::hypot(tr_re, tr_im);
  if (std::fabs(tr - 1.0) > 1e-2)
    return set_global_error(GRA_E_ARG,
                            "resonance %d: spin density not normalised (|Tr rho| = %g, should be 1)",
                            r, tr);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b)
      if (std::fabs(q.rho_re[a * n + b] - q.rho_re[b * n + a]) > 1e-9 ||
          std::fabs(q.rho_im[a * n + b] + q.rho_im[b * n + a]) > 1e-9)
        return set_global_error(GRA_E_ARG, "resonance %d: spin density is not Hermitian", r);
  const int m = 2 * n;
  double A[10][10];
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      const double re = q.rho_re[a * n + b], im = q.rho_im[a * n + b];
      A[a][b] = re;
      A[n + a][n + b] = re;
      A[a][n + b] = -im;
      A[n + a][b] = im;
    }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < m; ++p)
      for (int k = p + 1; k < m; ++k) off += A[p][k] * A[p][k];
    if (off < 1e-30) break;
    for (int p = 0; p < m; ++p)
      for (int k = p + 1; k < m; ++k) {
        if (A[p][k] == 0.0) continue;
        const double th = (A[k][k] - A[p][p]) / (2.0 * A[p][k]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int i = 0; i < m; ++i) {
          const double ap = A[i][p], ak = A[i][k];
          A[i][p] = cs * ap - sn * ak;
          A[i][k] = sn * ap + cs * ak;
        }
        for (int i = 0; i < m; ++i) {
          const double ap = A[p][i], ak = A[k][i];
          A[p][i] = cs * ap - sn * ak;
          A[k][i] = sn * ap + cs * ak;
        }
      }
  }
  for (int p = 0; p < m; ++p)
    if (A[p][p] < -1e-9)
      return set_global_error(GRA_E_ARG,
                              "resonance %d: spin density has a negative eigenvalue (%g)", r,
                              A[p][p]);
  return GRA_OK;
}

int validate(const gra_cfg* c) {
  if (!c) return set_global_error(GRA_E_ARG, "configuration is NULL");
  if (c->abi_size != (int32_t)sizeof(gra_cfg))
    return set_global_error(GRA_E_ARG, "gra_cfg.abi_size = %d, library expects %d", c->abi_size,
                            (int)sizeof(gra_cfg));
  if (c->process != GRA_PROC_REGGE_CON && c->process != GRA_PROC_TENSOR_CON &&
      c->process != GRA_PROC_RES_CON && c->process != GRA_PROC_RES)
    return set_global_error(GRA_E_ARG, "unknown process %d", c->process);
  if (c->process == GRA_PROC_TENSOR_CON) {
    // MTensorPomeron::ME4 deduces its spin mode from the central particles' spins
    // (src/MTensorPomeron.cc:643-654)
    if (c->central_spin2 == 2) {
      if (!(c->w_central > 0.0) || !(c->m_daughter > 0.0))
        return set_global_error(GRA_E_ARG,
                                "PP[CONTENSOR] 2xV: built for vector mesons with a width and a "
                                "two-body decay (w_central > 0, m_daughter > 0)");
    } else if (c->central_spin2 == 0) {
      if (c->w_central > 0.0 || c->m_daughter > 0.0)
        return set_global_error(GRA_E_ARG,
                                "PP[CONTENSOR] 2xS: built for a stable pseudoscalar pair "
                                "(w_central = 0, m_daughter = 0)");
      if (c->amp_precision != GRA_AMP_FP64)
        return set_global_error(GRA_E_ARG, "PP[CONTENSOR] 2xS: the amplitude runs in FP64 only");
    } else if (c->central_spin2 == 1) {
      return set_global_error(GRA_E_ARG,
                              "PP[CONTENSOR] 2xF (fermion pair) is not built");
    } else {
      return set_global_error(GRA_E_ARG,
                              "MTensorPomeron::ME4: invalid daughter spin 2J = %d (J = 0, 1/2, 1 "
                              "pairs supported)", c->central_spin2);
    }
  }
  if (c->process == GRA_PROC_RES_CON || c->process == GRA_PROC_RES) {
    if (c->w_central > 0.0 || c->m_daughter > 0.0)
      return set_global_error(GRA_E_ARG,
                              "PP[RES(+CON)]: built for stable central particles");
    if (c->n_res < (c->process == GRA_PROC_RES ? 1 : 0) || c->n_res > GRA_MAX_RES)
      return set_global_error(GRA_E_ARG, "n_res %d outside %d..%d", c->n_res,
                              c->process == GRA_PROC_RES ? 1 : 0, GRA_MAX_RES);
    if (c->spin_frame < GRA_FRAME_CM || c->spin_frame > GRA_FRAME_GJ)
      return set_global_error(GRA_E_ARG, "spin_frame %d unknown", c->spin_frame);
    for (int r = 0; r < c->n_res; ++r) {
      const gra_res& q = c->res[r];
      if (q.J < 0 || q.J > 2 || q.P != ((q.J & 1) ? -1 : 1))
        return set_global_error(GRA_E_ARG,
                                "resonance %d: J^P = %d^%d; built for 0+, 1-, 2+ (decays to two "
                                "spinless particles)", r, q.J, q.P);
      if (q.BW < GRA_BW_FIXED || q.BW > GRA_BW_SPIN)
        return set_global_error(GRA_E_ARG, "resonance %d: BW %d unknown", r, q.BW);
      if (!(q.mass > 2.0 * c->m_central) || !(q.width > 0.0) || !(q.BR > 0.0) || !(q.g_FF > 0.0))
        return set_global_error(GRA_E_ARG,
                                "resonance %d: needs M > 2 m_central, W > 0, BR > 0, g_FF > 0", r);
      if (q.alpha_re * q.alpha_re + q.alpha_im * q.alpha_im <= 0.0)
        return set_global_error(GRA_E_ARG, "resonance %d: alpha_{J0} = 0 (decay impossible)", r);
      const int rv = check_rho(q, r);
      if (rv != GRA_OK) return rv;
    }
  }
  const int dmin = (c->m_daughter > 0.0) ? 14 : (c->w_central > 0.0 ? 10 : 8);
  if (c->dim < dmin || c->dim > GRA_MAX_DIM)
    return set_global_error(GRA_E_ARG, "dim %d: this process needs %d..%d", c->dim, dmin,
                            GRA_MAX_DIM);
  if (c->phase_space != GRA_PS_C && c->phase_space != GRA_PS_F)
    return set_global_error(GRA_E_ARG, "phase_space %d unknown", c->phase_space);
  if (c->phase_space == GRA_PS_F) {
    if ((c->process != GRA_PROC_REGGE_CON && c->process != GRA_PROC_RES_CON &&
         c->process != GRA_PROC_RES &&
         !(c->process == GRA_PROC_TENSOR_CON && c->central_spin2 == 0)) ||
        c->w_central > 0.0 || c->m_daughter > 0.0)
      return set_global_error(GRA_E_ARG,
                              "<F> phase space: built for two stable central particles");
    const double Mmax = std::min(c->beam_E1 + c->beam_E2 - 2.0 * c->m_beam, c->gen_M_max);
    if (!(std::max(2.0 * c->m_central + 1e-4, c->gen_M_min) < Mmax))
      return set_global_error(GRA_E_ARG,
                              "<F> phase space: empty system mass range [%g, %g] (set gen_M_max; "
                              "MFactorized::B51RandomKin throws here)", c->gen_M_min, c->gen_M_max);
  }
  if (c->amp_precision != GRA_AMP_FP64 && c->amp_precision != GRA_AMP_FP32)
    return set_global_error(GRA_E_ARG, "amp_precision %d unknown", c->amp_precision);
  // the boost to the lab applies when |p_z(beam 1) + p_z(beam 2)| > 1e-6 GeV
  // (src/MContinuum.cc:386-387)
  const double pz_sum = msqrt(c->beam_E1 * c->beam_E1 - c->m_beam * c->m_beam) -
                        msqrt(c->beam_E2 * c->beam_E2 - c->m_beam * c->m_beam);
  if (std::fabs(pz_sum) > 1e-6 && (c->screening || c->phase_space != GRA_PS_C))
    return set_global_error(GRA_E_ARG,
                            "asymmetric beams (E1 != E2): built for the unscreened <C> class (the "
                            "screening nodes and the <F> kinematics assume the centre-of-mass "
                            "frame)");
  if (!(c->fwd_pt_max > c->fwd_pt_min) || !(c->kt_max > c->kt_min) || !(c->rap_max > c->rap_min))
    return set_global_error(GRA_E_ARG, "empty generation-cut interval");
  if (c->veto_active) {
    if (c->n_veto < 0 || c->n_veto > GRA_MAX_VETO)
      return set_global_error(GRA_E_ARG, "n_veto %d outside 0..%d", c->n_veto, GRA_MAX_VETO);
    // aux::AssertCut(.., dothrow = true) on each domain's Eta and Pt, src/MGraniitti.cc:950-956
    for (int d = 0; d < c->n_veto; ++d)
      if (!(c->veto_eta_max[d] > c->veto_eta_min[d]) || !(c->veto_pt_max[d] > c->veto_pt_min[d]))
        return set_global_error(GRA_E_ARG,
                                "VETOCUTS::%d: Eta [%g, %g], Pt [%g, %g] (maximum value not above "
                                "minimum value)", d, c->veto_eta_min[d], c->veto_eta_max[d],
                                c->veto_pt_min[d], c->veto_pt_max[d]);
  }
  if (c->S_factor <= 0.0) return set_global_error(GRA_E_ARG, "S_factor must be > 0");
  if (c->screening) {
    if (c->loop_rule == GRA_LOOP_SIMPSON38) {
      if (c->n_loop_phi < 3 || c->n_loop_phi % 3 || c->n_loop_kt < 3 || c->n_loop_kt % 3)
        return set_global_error(GRA_E_ARG,
                                "loop intervals must be multiples of 3 (Simpson 3/8): phi %d, kt %d",
                                c->n_loop_phi, c->n_loop_kt);
    } else if (c->loop_rule == GRA_LOOP_GAUSS) {
      if (c->n_loop_phi < 0 || c->n_loop_kt < 0)
        return set_global_error(GRA_E_ARG, "loop node counts must be >= 1");
    } else {
      return set_global_error(GRA_E_ARG, "loop_rule %d unknown", c->loop_rule);
    }
    if (c->eik_log_nodes != 0 &&
        (c->eik_log_nodes < 2 || c->eik_log_nodes > 4096 || !(c->min_loop_kt > 0.0)))
      return set_global_error(GRA_E_ARG,
                              "eik_log_nodes %d: 0, or 2..4096 log-spaced nodes with MinLoopKT > 0",
                              c->eik_log_nodes);
    if ((c->n_loop_phi + 1) * (c->n_loop_kt + 1) > 1024)
      return set_global_error(GRA_E_ARG, "too many loop nodes (max 1024, LDS node table)");
    if (c->n_fb < 2 || c->n_fb % 2)
      return set_global_error(GRA_E_ARG, "n_fb %d must be even (composite Simpson)", c->n_fb);
    if (c->n_bt < 2 || c->n_kt2 < 2) return set_global_error(GRA_E_ARG, "tables need >= 2 intervals");
    if (!(c->max_loop_kt > c->min_loop_kt))
      return set_global_error(GRA_E_ARG, "MaxLoopKT must exceed MinLoopKT");
    if (c->max_loop_kt * c->max_loop_kt > 25.0 * (1 + 1e-5))
      return set_global_error(GRA_E_ARG, "Interpolate1D: loop kt^2 outside the table [1e-6, 25]");
  }
  int on = 0;
  for (int k = 0; k < 3; ++k) on += c->regge_on[k] ? 1 : 0;
  if (c->process == GRA_PROC_REGGE_CON && c->amp_precision == GRA_AMP_FP32 && on != 1)
    return set_global_error(GRA_E_ARG, "FP32 amplitude needs exactly one active Regge trajectory");
  return GRA_OK;
}

// Run constants of the resonances: the production coupling g_A e^{i g_phi}
// (src/MForm.cc:276-279); the 2-body decay coupling sqrt(BR W / PS) with
// PS = PDW2body(M0^2, m^2, m^2, 1, 1) (src/MProcess.cc:668-713,
// include/Graniitti/MKinematics.h:331-334); T_{00} = alpha_{J0}, renormalised to
// |alpha| = 1 when it is not (spin::InitTMatrix, src/MSpin.cc:54-262: for two
// spinless daughters only (l, s) = (J, 0) passes parity, with NORM = CG = 1);
// PhotoProp's running slope B0 at m^2 = M0^2 (src/MRegge.cc:849-880).
void res_table(const gra_cfg& c, std::vector<ResDev>& tab) {
  tab.assign((size_t)c.n_res, ResDev{});
  const double m2 = c.m_central * c.m_central;
  for (int r = 0; r < c.n_res; ++r) {
    const gra_res& q = c.res[r];
    ResDev& d = tab[r];
    d.M0 = q.mass;
    d.W = q.width;
    d.M02 = q.mass * q.mass;
    d.gFF4 = (q.g_FF * q.g_FF) * (q.g_FF * q.g_FF);
    d.g_re = q.g_A * std::cos(q.g_phi);
    d.g_im = q.g_A * std::sin(q.g_phi);
    const double lam = d.M02 * d.M02 + m2 * m2 + m2 * m2 - 2.0 * d.M02 * m2 - 2.0 * d.M02 * m2 -
                       2.0 * m2 * m2;
    const double M0 = std::sqrt(d.M02 > 0.0 ? d.M02 : 0.0);
    const double PS = std::sqrt(lam > 0.0 ? lam : 0.0) * 1.0 / (16.0 * GRA_PI * 1.0 * (M0 * M0 * M0));
    const double g2 = q.BR * q.width / PS;
    d.gdec = std::sqrt(g2 > 0.0 ? g2 : 0.0);
    const double a2 = q.alpha_re * q.alpha_re + q.alpha_im * q.alpha_im;
    const double an = std::fabs(a2 - 1.0) > 1e-6 ? std::sqrt(a2) : 1.0;
    d.T_re = q.alpha_re / an;
    d.T_im = q.alpha_im / an;
    const double mf = std::max(0.7 * 0.7, d.M02);
    d.B0 = std::pow(mf / 2.0, -GRA_PI / 2.0) + 3.0 * GRA_PI / 2.0;
    for (int k = 0; k < 25; ++k) {
      d.rho_re[k] = q.rho_re[k];
      d.rho_im[k] = q.rho_im[k];
    }
    d.J = q.J;
    d.P = q.P;
    d.BW = q.BW;
    d.photo = (q.J == 1 && q.P == -1) ? 1 : 0;
  }
}

void fill_params(const gra_cfg& c, RunParams& P) {
  std::memset(&P, 0, sizeof(P));
  P.process = c.process;
  P.amp_mode = c.amp_precision;
  P.dim = c.dim;
  P.screening = c.screening;
  const double E1 = c.beam_E1, E2 = c.beam_E2, m = c.m_beam;
  P.pbeam1 = {E1, 0.0, 0.0, msqrt(E1 * E1 - m * m)};
  P.pbeam2 = {E2, 0.0, 0.0, -msqrt(E2 * E2 - m * m)};
  P.beamsum = {P.pbeam1.e + P.pbeam2.e, P.pbeam1.x + P.pbeam2.x, P.pbeam1.y + P.pbeam2.y,
               P.pbeam1.z + P.pbeam2.z};
  P.s = P.beamsum.e * P.beamsum.e -
        (P.beamsum.x * P.beamsum.x + P.beamsum.y * P.beamsum.y + P.beamsum.z * P.beamsum.z);
  P.sqrt_s = msqrt(P.s);
  P.rs = std::sqrt(P.s);
  P.asym = std::fabs(P.beamsum.z) > 1e-6 ? 1 : 0;
  P.s32 = std::pow(P.s, 3.0 / 2.0);
  P.c_beam1 = (m * m) / (P.pbeam1.e + P.pbeam1.z);
  P.c_beam2 = (m * m) / (P.pbeam2.e - P.pbeam2.z);
  const double ZERO_EPS = 1e-12;
  P.lmin = std::log(c.fwd_pt_min + ZERO_EPS);
  P.lmax = std::log(c.fwd_pt_max);
  P.kt_min = c.kt_min;
  P.kt_max = c.kt_max;
  P.rap_min = c.rap_min;
  P.rap_max = c.rap_max;
  P.m_beam = c.m_beam;
  P.m_central = c.m_central;
  P.w_central = c.w_central;
  P.m_daughter = c.m_daughter;
  if (c.w_central > 0.0) {
    // MProcess::GetOffShellMass + MRandom::RelativisticBWRandom constants
    const double m0 = c.m_central, G = c.w_central, L = c.offshell;
    const double dm = (c.m_daughter > 0.0) ? (0.0 + c.m_daughter) + c.m_daughter + 1e-5 : 0.0;
    const double m2 = m0 * m0;
    const double m2max = (m0 + L * G) * (m0 + L * G);
    const double m2min = std::max(std::max(0.0, dm * dm), (m0 - L * G) * (m0 - L * G));
    P.bw_a1 = std::atan2(m2min - m2, m0 * G);
    P.bw_a2 = std::atan2(m2max - m2, m0 * G);
    P.bw_dmin = dm;
  }
  P.fid_active = c.fid_active;
  P.veto_active = c.veto_active && c.n_veto > 0;
  P.n_veto = c.veto_active ? c.n_veto : 0;
  for (int d = 0; d < GRA_MAX_VETO; ++d) {
    P.veto_eta_min[d] = c.veto_eta_min[d];
    P.veto_eta_max[d] = c.veto_eta_max[d];
    P.veto_pt_min[d] = c.veto_pt_min[d];
    P.veto_pt_max[d] = c.veto_pt_max[d];
  }
  P.fid_eta_min = c.fid_eta_min;
  P.fid_eta_max = c.fid_eta_max;
  P.fid_rap_min = c.fid_rap_min;
  P.fid_rap_max = c.fid_rap_max;
  P.fid_pt_min = c.fid_pt_min;
  P.fid_pt_max = c.fid_pt_max;
  P.fid_et_min = c.fid_et_min;
  P.fid_et_max = c.fid_et_max;
  P.sys_M_min = c.sys_M_min;
  P.sys_M_max = c.sys_M_max;
  P.sys_Y_min = c.sys_Y_min;
  P.sys_Y_max = c.sys_Y_max;
  P.sys_Pt_min = c.sys_Pt_min;
  P.sys_Pt_max = c.sys_Pt_max;
  P.fwd_t_min = c.fwd_t_min;
  P.fwd_t_max = c.fwd_t_max;
  // BNPhaseSpaceWeight / BNIntegralVolume / ForwardVolume constants (Nf = 4, Kf = 2)
  P.psw_c1 = 1.0 / std::pow(2, 2 * (4 - 2));
  P.psw_c2 = 1.0 / std::pow(2.0 * GRA_PI, 3 * 4 - 4);
  const double lspan = std::log(c.fwd_pt_max) - std::log(c.fwd_pt_min + ZERO_EPS);
  P.lspan2 = lspan * lspan;
  double KT_vol = std::pow(c.kt_max - c.kt_min, 1);
  KT_vol *= std::pow(2.0 * GRA_PI, 1);
  KT_vol *= std::pow(c.rap_max - c.rap_min, 2);
  P.KT_vol = KT_vol;
  P.PHI_vol = (2.0 * GRA_PI) * (2.0 * GRA_PI);
  P.invS = 1.0 / c.S_factor;
  P.GeV2barn = gev2barn();
  P.flux = 2.0 * P.s;
  // <F>: system mass range from the decay-tree masses (pions at their pole mass)
  // and the GENCUTS, src/MFactorized.cc:259-285
  P.phase_space = c.phase_space;
  if (c.phase_space == GRA_PS_F) {
    const double MARGIN = 1e-4;
    double Mmin = (0.0 + c.m_central) + c.m_central + MARGIN;
    const double mb1 = P.pbeam1.e * P.pbeam1.e -
                       (P.pbeam1.x * P.pbeam1.x + P.pbeam1.y * P.pbeam1.y + P.pbeam1.z * P.pbeam1.z);
    const double mb2 = P.pbeam2.e * P.pbeam2.e -
                       (P.pbeam2.x * P.pbeam2.x + P.pbeam2.y * P.pbeam2.y + P.pbeam2.z * P.pbeam2.z);
    double Mmax = P.sqrt_s - ((mb1 > 0.0 ? msqrt(mb1) : -msqrt(-mb1)) +
                              (mb2 > 0.0 ? msqrt(mb2) : -msqrt(-mb2)));
    Mmin = std::max(Mmin, c.gen_M_min);
    Mmax = std::min(Mmax, c.gen_M_max);
    P.f_Mmin = Mmin;
    P.f_Mmax = Mmax;
    P.f_vol = (Mmax * Mmax - Mmin * Mmin) * (c.rap_max - c.rap_min);
    const double tp = 2.0 * GRA_PI;
    P.f_psw = (1.0 / 2.0) * (1.0 / (tp * ((tp * tp) * (tp * tp))));
  }
  // Regge
  int on = 0, kk = -1;
  for (int k = 0; k < 3; ++k) {
    P.regge_on[k] = c.regge_on[k];
    P.regge_a0[k] = c.regge_a0[k];
    P.regge_ap[k] = c.regge_ap[k];
    const hcx e = regge_eta(c.regge_a0[k], c.regge_sgn[k]);
    P.eta0_re[k] = e.re;
    P.eta0_im[k] = e.im;
    if (c.regge_on[k]) {
      ++on;
      kk = k;
    }
  }
  P.regge_k = (on == 1) ? kk : -1;
  P.regge_s0 = c.regge_s0;
  P.gN_P = c.gN_P;
  P.fc1 = c.fc1;
  P.fc2 = c.fc2;
  P.gpp_P = c.regge_c[0] / c.gN_P;
  P.m_prop2 = c.m_prop * c.m_prop;
  P.offshellFF = c.offshellFF;
  P.b_EXP = c.b_EXP;
  P.a_OREAR = c.a_OREAR;
  P.b_OREAR = c.b_OREAR;
  P.b_POW = c.b_POW;
  if (kk >= 0) {
    P.ph_c = (-GRA_PI / 2.0) * c.regge_ap[kk];
    const double e2 = P.eta0_re[kk] * P.eta0_re[kk] + P.eta0_im[kk] * P.eta0_im[kk];
    const double g2 = (c.gN_P * P.gpp_P) * (c.gN_P * P.gpp_P);
    P.amp_c2 = (e2 * e2) * (g2 * g2);
  }
  // screening loop
  P.nphi = c.n_loop_phi;
  P.nkt = c.n_loop_kt;
  P.step_kt = (c.max_loop_kt - c.min_loop_kt) / c.n_loop_kt;
  P.step_phi = (2.0 * GRA_PI - 0.0) / c.n_loop_phi;
  P.min_loop_kt = c.min_loop_kt;
  // Simpson 3/8 in both variables (src/MProcess.cc:254-261); the Gauss rule's
  // node weights carry their step sizes themselves
  P.hfac = (c.loop_rule == GRA_LOOP_GAUSS) ? 1.0 : 9 * (P.step_phi * P.step_kt) / 64;
  P.nrm = 1.0 / (8.0 * (GRA_PI * GRA_PI) * P.s);
  P.kappa = P.nrm * P.hfac;
  // tensor pomeron
  P.tp_delta_P = c.tp_delta_P;
  P.tp_ap_P = c.tp_ap_P;
  P.tp_gPNN = c.tp_gPNN;
  P.tp_gPvv0 = c.tp_gPvv[0];
  P.tp_gPvv1 = c.tp_gPvv[1];
  P.tp_mp_ff = c.tp_mp_ff;
  P.tp_mode = (c.process == GRA_PROC_TENSOR_CON && c.central_spin2 == 0) ? kTp2xS : kTp2xV;
  // 2xS: pion or kaon coupling by the central mass (src/MTensorPomeron.cc:730)
  P.tp_gP = (c.m_central < 0.2) ? c.tp_gPpipi : c.tp_gPKK;
  P.res = nullptr;
  P.nres = (c.process == GRA_PROC_RES_CON || c.process == GRA_PROC_RES) ? c.n_res : 0;
  P.spin_gen = c.spin_gen;
  P.spin_dec = c.spin_dec;
  P.spin_frame = c.spin_frame;
  P.omega = c.regge_omega;
  P.m_rho = c.m_central;
}

// One eikonal table through the cache directory dir (none: nullptr), as
// MEikonal::S3Constructor handles MBT and MSA (src/MEikonal.cc:62-93, 98-129):
// read the table's file; if it is missing or corrupt, compute the table, write
// it and read it back. compute() fills T already rounded as that write-read
// cycle rounds it (15 significant digits).
template <typename F>
int cached_table(gra_handle* h, const char* dir, int table, int n_rows, std::vector<double>& T,
                 F compute) {
  std::string path;
  if (dir) {
    path = std::string(dir) + "/" + eikonal_cache_name(h->cfg, h->P.s, table);
    const int r = eikonal_table_read(path.c_str(), n_rows, T);
    if (r == GRA_OK) return GRA_OK;
    if (r != GRA_E_IO) return fail(h, r, "%s", global_error());
  }
  const int r = compute();
  if (r != GRA_OK || !dir) return r;
  int w = eikonal_table_write(path.c_str(), T);
  if (w == GRA_OK) w = eikonal_table_read(path.c_str(), n_rows, T);
  return w == GRA_OK ? GRA_OK : fail(h, w, "%s", global_error());
}

// MEikonal::S3Constructor tables on the GPU, src/MEikonal.cc:46-132
int build_eikonal(gra_handle* h, const char* cache_dir) {
  const gra_cfg& c = h->cfg;
  const RunParams& P = h->P;
  const int NF = c.n_fb;
  std::vector<double> pre(2 * (size_t)(NF + 1));
  DevBuf d_pre, d_out;
  const int nrow_max = std::max(std::max(c.n_bt, c.n_kt2) + 1, c.eik_log_nodes);
  if (d_pre.ensure(sizeof(double) * pre.size()) || d_out.ensure(sizeof(double) * 2 * nrow_max))
    return fail(h, GRA_E_NOMEM, "eikonal table buffers");
  std::vector<double> out(2 * (size_t)nrow_max);
  // Density Omega(bt): pre_q = SingleAmpElastic(s, -kt_q^2)
  const double kmin = 1E-9, kmax = 30.0, kstep = (kmax - kmin) / NF;
  for (int q = 0; q <= NF; ++q) {
    const double kt = kmin + q * kstep;
    const hcx a = single_amp_elastic(c, P.s, -(kt * kt));
    pre[2 * q] = a.re;
    pre[2 * q + 1] = a.im;
  }
  const double bmin = 1E-6, bmax = 10.0 / gev2fm(), bstep = (bmax - bmin) / c.n_bt;
  int r = cached_table(h, cache_dir, 0, c.n_bt + 1, h->mbt, [&]() -> int {
    GRA_HIP_CHECK(h, hipMemcpyAsync(d_pre.p, pre.data(), sizeof(double) * pre.size(),
                                    hipMemcpyHostToDevice, h->stream));
    GRA_HIP_CHECK(h, launch_fb_table(0, d_pre.as<double>(), NF, kmin, kstep, bmin, bstep,
                                     c.n_bt + 1, 1 / (2.0 * GRA_PI * P.s), d_out.as<double>(),
                                     h->stream, nullptr));
    GRA_HIP_CHECK(h, hipMemcpyAsync(out.data(), d_out.p, sizeof(double) * 2 * (c.n_bt + 1),
                                    hipMemcpyDeviceToHost, h->stream));
    GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
    h->mbt.resize(3 * (size_t)(c.n_bt + 1));
    for (int i = 0; i <= c.n_bt; ++i) {
      h->mbt[3 * i] = round15(bmin + i * bstep);
      h->mbt[3 * i + 1] = round15(out[2 * i]);
      h->mbt[3 * i + 2] = round15(out[2 * i + 1]);
    }
    return GRA_OK;
  });
  if (r != GRA_OK) return r;
  // Screening amplitude A_eik(kt^2): pre_q = i (1 - exp(i Omega(bt_q) / 2))
  const double STEP = (bmax - bmin) / NF;
  for (int q = 0; q <= NF; ++q) {
    const double bt = bmin + q * STEP;
    const hcx Om = interp_rows(h->mbt, c.n_bt, bmin, bstep, bt);
    const hcx z = {(0.0 * Om.re - 1.0 * Om.im) / 2.0, (0.0 * Om.im + 1.0 * Om.re) / 2.0};
    const hcx e = hexp(z);
    const hcx om = {1.0 - e.re, -e.im};
    pre[2 * q] = 0.0 * om.re - 1.0 * om.im;
    pre[2 * q + 1] = 0.0 * om.im + 1.0 * om.re;
  }
  const double amin = 1E-6, amax = 25.0, astep = (amax - amin) / c.n_kt2;
  r = cached_table(h, cache_dir, 1, c.n_kt2 + 1, h->msa, [&]() -> int {
    GRA_HIP_CHECK(h, hipMemcpyAsync(d_pre.p, pre.data(), sizeof(double) * pre.size(),
                                    hipMemcpyHostToDevice, h->stream));
    GRA_HIP_CHECK(h, launch_fb_table(1, d_pre.as<double>(), NF, bmin, STEP, amin, astep,
                                     c.n_kt2 + 1, (2.0 * P.s) * (2.0 * GRA_PI),
                                     d_out.as<double>(), h->stream, nullptr));
    GRA_HIP_CHECK(h, hipMemcpyAsync(out.data(), d_out.p, sizeof(double) * 2 * (c.n_kt2 + 1),
                                    hipMemcpyDeviceToHost, h->stream));
    GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
    h->msa.resize(3 * (size_t)(c.n_kt2 + 1));
    for (int i = 0; i <= c.n_kt2; ++i) {
      h->msa[3 * i] = round15(amin + i * astep);
      h->msa[3 * i + 1] = round15(out[2 * i]);
      h->msa[3 * i + 2] = round15(out[2 * i + 1]);
    }
    return GRA_OK;
  });
  if (r != GRA_OK) return r;
  // A_eik at log-spaced kt nodes: the same Fourier-Bessel integral as the MSA
  // table's rows, evaluated at those kt^2 exactly
  const int NL = c.eik_log_nodes;
  std::vector<double> lkt, lre, lim;
  if (NL > 0) {
    const double l0 = std::log(c.min_loop_kt), l1 = std::log(c.max_loop_kt);
    std::vector<double> xs((size_t)NL);
    lkt.resize(NL);
    for (int m = 0; m < NL; ++m) {
      lkt[m] = l0 + (l1 - l0) * m / (NL - 1);
      const double kt = std::exp(lkt[m]);
      xs[m] = kt * kt;
    }
    DevBuf d_xs;
    if (d_xs.ensure(sizeof(double) * NL)) return fail(h, GRA_E_NOMEM, "log-node buffer");
    GRA_HIP_CHECK(h, hipMemcpyAsync(d_pre.p, pre.data(), sizeof(double) * pre.size(),
                                    hipMemcpyHostToDevice, h->stream));
    GRA_HIP_CHECK(h, hipMemcpyAsync(d_xs.p, xs.data(), sizeof(double) * NL, hipMemcpyHostToDevice,
                                    h->stream));
    GRA_HIP_CHECK(h, launch_fb_table(1, d_pre.as<double>(), NF, bmin, STEP, 0.0, 0.0, NL,
                                     (2.0 * P.s) * (2.0 * GRA_PI), d_out.as<double>(), h->stream,
                                     d_xs.as<double>()));
    std::vector<double> o(2 * (size_t)NL);
    GRA_HIP_CHECK(h, hipMemcpyAsync(o.data(), d_out.p, sizeof(double) * 2 * NL,
                                    hipMemcpyDeviceToHost, h->stream));
    GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
    lre.resize(NL);
    lim.resize(NL);
    for (int m = 0; m < NL; ++m) {
      lre[m] = o[2 * m];
      lim[m] = o[2 * m + 1];
    }
  }
  auto aeik_at = [&](double kt) -> hcx {
    if (NL == 0) return interp_rows(h->msa, c.n_kt2, amin, astep, kt * kt);
    const double lk = std::log(kt);
    const double step = (lkt[NL - 1] - lkt[0]) / (NL - 1);
    int m = (int)std::floor((lk - lkt[0]) / step);
    m = std::max(0, std::min(NL - 2, m));
    const double f = (lk - lkt[m]) / (lkt[m + 1] - lkt[m]);
    return {lre[m] + f * (lre[m + 1] - lre[m]), lim[m] + f * (lim[m + 1] - lim[m])};
  };
  d_pre.release();
  d_out.release();
  // Loop nodes (phi_i, kt_j), i-major as the reference's (i, j) loops
  const int NP = c.n_loop_phi, NK = c.n_loop_kt;
  std::vector<double> gx, gw;
  if (c.loop_rule == GRA_LOOP_GAUSS) gauss_legendre(NK + 1, gx, gw);
  h->nodes.resize((size_t)(NP + 1) * (NK + 1));
  std::vector<NodeRow> rows((size_t)NP + 1);
  for (int i = 0; i <= NP; ++i) {
    const bool gauss = c.loop_rule == GRA_LOOP_GAUSS;
    const double phi = gauss ? 2.0 * GRA_PI * i / (NP + 1) : 0.0 + i * P.step_phi;
    rows[i].c = (float)std::cos(phi);
    rows[i].s = (float)std::sin(phi);
    for (int j = 0; j <= NK; ++j) {
      const double half = 0.5 * (c.max_loop_kt - c.min_loop_kt);
      const double kt = gauss ? c.min_loop_kt + half * (1.0 + gx[j]) : c.min_loop_kt + j * P.step_kt;
      const hcx A = aeik_at(kt);
      NodeConst& nd = h->nodes[(size_t)i * (NK + 1) + j];
      nd.kx = kt * std::cos(phi);
      nd.ky = kt * std::sin(phi);
      nd.kt = kt;
      nd.w = gauss ? (2.0 * GRA_PI / (NP + 1)) * (half * gw[j])
                   : simpson38_w(i, NP) * simpson38_w(j, NK);
      nd.aeik_re = A.re;
      nd.aeik_im = A.im;
      nd.cf_re = (float)(nd.w * kt * A.re);
      nd.cf_im = (float)(nd.w * kt * A.im);
      nd.kxf = (float)nd.kx;
      nd.kyf = (float)nd.ky;
      nd.k2f = (float)(kt * kt);
      nd.pad = 0.f;
    }
  }
  // device table: the NodeConst entries, then the compact FP32 nodes
  const size_t NN = h->nodes.size();
  std::vector<NodeF32> n32(NN);
  for (size_t k = 0; k < NN; ++k) {
    const NodeConst& nd = h->nodes[k];
    const double cre = nd.w * nd.kt * nd.aeik_re, cim = nd.w * nd.kt * nd.aeik_im;
    NodeF32& f = n32[k];
    f.kt = (float)nd.kt;
    f.k2 = nd.k2f;
    f.arg = (float)(std::atan2(cim, cre) / (2.0 * GRA_PI));
    f.mod = (float)std::hypot(cre, cim);
  }
  const size_t b0 = sizeof(NodeConst) * NN, b1 = sizeof(NodeF32) * NN,
               b2 = sizeof(NodeRow) * rows.size();
  if (h->d_nodes.ensure(b0 + b1 + b2)) return fail(h, GRA_E_NOMEM, "node table");
  GRA_HIP_CHECK(h, hipMemcpy(h->d_nodes.p, h->nodes.data(), b0, hipMemcpyHostToDevice));
  GRA_HIP_CHECK(h, hipMemcpy((char*)h->d_nodes.p + b0, n32.data(), b1, hipMemcpyHostToDevice));
  GRA_HIP_CHECK(h, hipMemcpy((char*)h->d_nodes.p + b0 + b1, rows.data(), b2, hipMemcpyHostToDevice));
  return GRA_OK;
}

}  // namespace

// The FP32 unit serves the FP32-amplitude mode, except for the resonance
// processes, whose FP32 screening node (FP64 kinematics, FP32 amplitude) is built
// in the FP64 unit beside their FP64 tree amplitude
static bool f32_kernels(const RunParams& P) {
  return P.amp_mode == GRA_AMP_FP32 && P.process != GRA_PROC_RES_CON && P.process != GRA_PROC_RES;
}

int ensure_point_buffers(gra_handle* h, uint64_t n) {
  const int D = h->P.dim;
  const size_t recb = f32_kernels(h->P) ? k32::loop_record_bytes(h->P)
                                        : k64::loop_record_bytes(h->P);
  if (h->d_W.ensure(sizeof(double) * n) || h->d_amp2.ensure(sizeof(double) * n) ||
      h->d_flags.ensure(n) || h->d_vw.ensure(sizeof(double) * n) ||
      h->d_bins.ensure((size_t)D * n) || h->d_ncand.ensure(16) ||
      (h->P.screening && h->d_rec.ensure(recb * n)))
    return fail(h, GRA_E_NOMEM, "device buffers for %llu points", (unsigned long long)n);
  return GRA_OK;
}

int ensure_gen_buffers(gra_handle* h, uint64_t n, uint64_t cap) {
  if (h->d_gen_idx.ensure(sizeof(uint64_t) * std::max<uint64_t>(cap, 1)) ||
      h->d_gen_mask.ensure(sizeof(uint64_t) * accept_mask_words(n)) ||
      h->d_gen_tile.ensure(sizeof(uint32_t) * accept_tile_words(n)) ||
      h->d_gen_cnt.ensure(6 * sizeof(unsigned long long)) ||
      h->d_gen_ev.ensure(sizeof(gra_event) * std::max<uint64_t>(cap, 1)))
    return fail(h, GRA_E_NOMEM, "generation buffers for %llu points", (unsigned long long)n);
  if (!h->h_gen_cnt &&
      hipHostMalloc((void**)&h->h_gen_cnt, 6 * sizeof(unsigned long long), hipHostMallocDefault) !=
          hipSuccess) {
    h->h_gen_cnt = nullptr;
    return fail(h, GRA_E_NOMEM, "pinned generation counters");
  }
  return GRA_OK;
}

int events_to_host(gra_handle* h, gra_event* dst, const void* d_src, uint64_t n) {
  const size_t bytes = sizeof(gra_event) * n;
  if (bytes == 0) return GRA_OK;
  hipPointerAttribute_t a;
  const bool pinned = hipPointerGetAttributes(&a, dst) == hipSuccess && a.type == hipMemoryTypeHost;
  (void)hipGetLastError();  // an unregistered pointer leaves an error behind
  if (pinned) {
    GRA_HIP_CHECK(h, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, h->stream));
    GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return GRA_OK;
  }
  if (bytes > h->h_ev_stage_bytes) {
    // grown by half again as much, so that passes whose event counts drift
    // upwards do not reallocate the pinned buffer each time
    if (h->h_ev_stage) (void)hipHostFree(h->h_ev_stage);
    h->h_ev_stage = nullptr;
    h->h_ev_stage_bytes = 0;
    const size_t want = bytes + bytes / 2;
    if (hipHostMalloc(&h->h_ev_stage, want, hipHostMallocDefault) != hipSuccess) {
      h->h_ev_stage = nullptr;
      return fail(h, GRA_E_NOMEM, "pinned event staging (%zu bytes)", want);
    }
    h->h_ev_stage_bytes = want;
  }
  GRA_HIP_CHECK(h, hipMemcpyAsync(h->h_ev_stage, d_src, bytes, hipMemcpyDeviceToHost, h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  std::memcpy(dst, h->h_ev_stage, bytes);
  return GRA_OK;
}

static void time_begin(gra_handle* h, int slot) {
  (void)hipEventRecord(h->ev[2 * slot], h->stream);
}
static void time_end(gra_handle* h, int slot) {
  (void)hipEventRecord(h->ev[2 * slot + 1], h->stream);
}
// After a stream synchronization: add the elapsed time of the recorded slots
static void time_collect(gra_handle* h, bool tree, bool screen, bool accum) {
  const bool on[3] = {tree, screen, accum};
  for (int s = 0; s < 3; ++s) {
    if (!on[s]) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev[2 * s], h->ev[2 * s + 1]) == hipSuccess) {
      h->time_ms[s] += ms;
      h->launches[s] += 1;
    }
  }
}

// The integrand kernels on the handle's stream: the tree stage appends the
// points that need screening to rec through the counter ncand (zero on entry),
// the screening stage completes them. timed: record the kernel durations.
static int launch_integrand(gra_handle* h, const double* d_grid, unsigned int bins, uint64_t seed,
                            uint32_t iter, uint64_t first, const double* d_x, uint64_t n,
                            double* d_W, double* d_amp2, uint8_t* d_flags, double* d_vw,
                            uint8_t* d_bins, void* rec, uint32_t* ncand, bool timed) {
  const RunParams& P = h->P;
  const bool f32 = f32_kernels(P);
  if (timed) time_begin(h, TS_TREE);
  GRA_HIP_CHECK(h, f32 ? k32::launch_tree(P, d_grid, bins, seed, iter, first, d_x, n, d_W, d_amp2,
                                          d_flags, d_vw, d_bins, rec, ncand, h->stream)
                       : k64::launch_tree(P, d_grid, bins, seed, iter, first, d_x, n, d_W, d_amp2,
                                          d_flags, d_vw, d_bins, rec, ncand, h->stream));
  if (timed) time_end(h, TS_TREE);
  if (P.screening) {
    if (timed) time_begin(h, TS_SCREEN);
    GRA_HIP_CHECK(h, f32 ? k32::launch_screen(P, h->d_nodes.p, rec, ncand, n, d_W, d_amp2, d_flags,
                                              h->stream)
                         : k64::launch_screen(P, h->d_nodes.p, rec, ncand, n, d_W, d_amp2, d_flags,
                                              h->stream));
    if (timed) time_end(h, TS_SCREEN);
  }
  return GRA_OK;
}

int run_integrand(gra_handle* h, const double* d_grid, unsigned int bins, uint64_t seed,
                  uint32_t iter, uint64_t first, const double* d_x, uint64_t n, double* d_W,
                  double* d_amp2, uint8_t* d_flags, double* d_vw, uint8_t* d_bins) {
  uint32_t* ncand = h->d_ncand.as<uint32_t>();
  GRA_HIP_CHECK(h, hipMemsetAsync(ncand, 0, sizeof(uint32_t), h->stream));
  return launch_integrand(h, d_grid, bins, seed, iter, first, d_x, n, d_W, d_amp2, d_flags, d_vw,
                          d_bins, h->d_rec.p, ncand, true);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free evaluation: one array per device, in the code
// object, so that gra_hip_eval_points_noalloc calls no allocator. Its device
// address comes from hipGetSymbolAddress once per device, at the first handle
// creation on that device; host code uses that address, never the symbol.
// Layout: [0, 256) the candidate counter, then the candidate records, the VEGAS
// weights (unused here: the points come from the caller) and the bin bytes, each
// 256-byte aligned.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[GRA_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mtx;
void* g_scratch_ptr[kMaxDevices] = {};

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of scratch n points of this process need (counter, records, weights, bins)
size_t noalloc_bytes(const gra_handle* h, uint64_t n) {
  const size_t recb = f32_kernels(h->P) ? k32::loop_record_bytes(h->P)
                                        : k64::loop_record_bytes(h->P);
  return 256 + (h->P.screening ? align256(recb * n) : 0) + align256(sizeof(double) * n) +
         align256((size_t)h->P.dim * n);
}
// Largest n with noalloc_bytes(h, n) <= the scratch size
uint64_t noalloc_max_points(const gra_handle* h) {
  const size_t recb = f32_kernels(h->P) ? k32::loop_record_bytes(h->P)
                                        : k64::loop_record_bytes(h->P);
  const size_t per = (h->P.screening ? recb : 0) + sizeof(double) + (size_t)h->P.dim;
  uint64_t n = (GRA_NOALLOC_SCRATCH_BYTES - 256 - 3 * 256) / per;
  while (n > 0 && noalloc_bytes(h, n) > (size_t)GRA_NOALLOC_SCRATCH_BYTES) --n;
  return n;
}
}  // namespace

int scratch_setup(gra_handle* h) {
  if (h->device < 0 || h->device >= kMaxDevices)
    return set_global_error(GRA_E_HIP, "device %d beyond the scratch table", h->device);
  std::lock_guard<std::mutex> lk(g_scratch_mtx);
  if (!g_scratch_ptr[h->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_global_error(GRA_E_HIP, "hipGetSymbolAddress(scratch): %s",
                              hipGetErrorString(e));
    g_scratch_ptr[h->device] = p;
  }
  h->scratch = g_scratch_ptr[h->device];
  return GRA_OK;
}

}  // namespace gra_hip

using namespace gra_hip;

// ---------------------------------------------------------------------------
// Entry points
// ---------------------------------------------------------------------------
extern "C" const char* gra_hip_last_error(const gra_handle* h) {
  return h ? h->err.c_str() : global_error();
}

namespace {
int create_handle(const gra_cfg* cfg, const char* eikonal_dir, gra_handle** out) {
  if (!out) return set_global_error(GRA_E_ARG, "gra_hip_create: out is NULL");
  *out = nullptr;
  const int v = validate(cfg);
  if (v != GRA_OK) return v;
  gra_handle* h = new (std::nothrow) gra_handle();
  if (!h) return set_global_error(GRA_E_NOMEM, "gra_hip_create: host allocation");
  h->cfg = *cfg;
  fill_params(h->cfg, h->P);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    delete h;
    return set_global_error(GRA_E_HIP, "gra_hip_create: no HIP device available");
  }
  h->device = dev;
  if (scratch_setup(h) != GRA_OK) {
    const std::string m = global_error();
    delete h;
    return set_global_error(GRA_E_HIP, "gra_hip_create: %s", m.c_str());
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return set_global_error(GRA_E_HIP, "gra_hip_create: stream creation failed");
  }
  for (int k = 0; k < 2 * TS_N; ++k) {
    if (hipEventCreate(&h->ev[k]) != hipSuccess) {
      gra_hip_destroy(h);
      return set_global_error(GRA_E_HIP, "gra_hip_create: event creation failed");
    }
  }
  if (h->P.nres > 0) {
    std::vector<ResDev> tab;
    res_table(h->cfg, tab);
    const size_t b = sizeof(ResDev) * tab.size();
    if (h->d_res.ensure(b) ||
        hipMemcpy(h->d_res.p, tab.data(), b, hipMemcpyHostToDevice) != hipSuccess) {
      gra_hip_destroy(h);
      return set_global_error(GRA_E_HIP, "gra_hip_create: resonance table upload failed");
    }
    h->P.res = (const ResDev*)h->d_res.p;
  }
  if (h->cfg.screening) {
    const int r = build_eikonal(h, eikonal_dir);
    if (r != GRA_OK) {
      const std::string m = h->err;
      gra_hip_destroy(h);
      return set_global_error(r, "gra_hip_create: %s", m.c_str());
    }
  }
  *out = h;
  return GRA_OK;
}
}  // namespace

extern "C" int gra_hip_create(const gra_cfg* cfg, gra_handle** out) {
  return create_handle(cfg, nullptr, out);
}

extern "C" int gra_hip_create_cached(const gra_cfg* cfg, const char* eikonal_dir,
                                     gra_handle** out) {
  if (!eikonal_dir || !*eikonal_dir)
    return set_global_error(GRA_E_ARG, "gra_hip_create_cached: eikonal_dir is empty");
  return create_handle(cfg, eikonal_dir, out);
}

extern "C" int gra_hip_eikonal_cache_name(const gra_cfg* cfg, int32_t table, char* name,
                                          size_t len) {
  const int v = validate(cfg);
  if (v != GRA_OK) return v;
  if ((table != 0 && table != 1) || !name || len == 0)
    return set_global_error(GRA_E_ARG, "gra_hip_eikonal_cache_name: table must be 0 or 1");
  RunParams P;
  fill_params(*cfg, P);
  const std::string n = eikonal_cache_name(*cfg, P.s, table);
  if (n.size() + 1 > len)
    return set_global_error(GRA_E_ARG, "gra_hip_eikonal_cache_name: buffer of %zu < %zu", len,
                            n.size() + 1);
  std::memcpy(name, n.c_str(), n.size() + 1);
  return GRA_OK;
}

extern "C" void gra_hip_destroy(gra_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->flow) flow_release(h->flow);
  DevBuf* bufs[] = {&h->d_nodes,     &h->d_res,       &h->d_x,         &h->d_W,
                    &h->d_amp2,      &h->d_flags,     &h->d_vw,        &h->d_bins,
                    &h->d_rec,       &h->d_ncand,     &h->d_grid,      &h->d_part_hist,
                    &h->d_part_scal, &h->d_stats,     &h->d_aux,       &h->d_gen_idx,
                    &h->d_gen_mask,  &h->d_gen_tile,  &h->d_gen_cnt,   &h->d_gen_ev};
  for (DevBuf* b : bufs) b->release();
  if (h->h_gen_cnt) (void)hipHostFree(h->h_gen_cnt);
  if (h->h_ev_stage) (void)hipHostFree(h->h_ev_stage);
  for (int k = 0; k < 2 * TS_N; ++k)
    if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

extern "C" int gra_hip_eval_points_device(gra_handle* h, const double* d_u, size_t n, double* d_W,
                                          double* d_amp2, uint8_t* d_flags) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  if (n == 0) return GRA_OK;
  if (!d_u || !d_W || !d_flags) return fail(h, GRA_E_ARG, "eval_points_device: NULL buffer");
  if (n > 0xFFFFFFFFull) return fail(h, GRA_E_ARG, "eval_points_device: n too large");
  int r = ensure_point_buffers(h, n);
  if (r) return r;
  r = run_integrand(h, nullptr, 0, 0, 0, 0, d_u, n, d_W, d_amp2, d_flags, h->d_vw.as<double>(),
                    h->d_bins.as<uint8_t>());
  if (r) return r;
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  time_collect(h, true, h->P.screening != 0, false);
  return GRA_OK;
}

extern "C" int gra_hip_eval_points_noalloc(gra_handle* h, const double* d_u, size_t n,
                                           double* d_W, double* d_amp2, uint8_t* d_flags) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  if (n == 0) return GRA_OK;
  if (!d_u || !d_W || !d_flags) return fail(h, GRA_E_ARG, "eval_points_noalloc: NULL buffer");
  if (!h->scratch) return fail(h, GRA_E_STATE, "eval_points_noalloc: no scratch on this device");
  const uint64_t nmax = noalloc_max_points(h);
  if (n > nmax)
    return fail(h, GRA_E_ARG,
                "eval_points_noalloc: %zu points exceed the %llu this process fits in the "
                "%d-byte scratch", n, (unsigned long long)nmax, (int)GRA_NOALLOC_SCRATCH_BYTES);
  const size_t recb = f32_kernels(h->P) ? k32::loop_record_bytes(h->P)
                                        : k64::loop_record_bytes(h->P);
  unsigned char* base = (unsigned char*)h->scratch;
  uint32_t* ncand = (uint32_t*)base;
  size_t off = 256;
  void* rec = base + off;
  off += h->P.screening ? align256(recb * n) : 0;
  double* vw = (double*)(base + off);
  off += align256(sizeof(double) * n);
  uint8_t* bins = base + off;
  // the counter is cleared by a kernel (no memset node, so the call can be captured)
  GRA_HIP_CHECK(h, launch_fill_u32(ncand, 0u, 1, h->stream));
  const int r = launch_integrand(h, nullptr, 0, 0, 0, 0, d_u, n, d_W, d_amp2, d_flags, vw, bins,
                                 rec, ncand, false);
  if (r) return r;
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return GRA_OK;
}

extern "C" int gra_hip_eval_points(gra_handle* h, const double* u, size_t n, double* W,
                                   double* amp2, uint8_t* flags) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  if (n == 0) return GRA_OK;
  if (!u || !W) return fail(h, GRA_E_ARG, "eval_points: u and W must not be NULL");
  if (n > 0xFFFFFFFFull) return fail(h, GRA_E_ARG, "eval_points: n too large");
  const int D = h->P.dim;
  int r = ensure_point_buffers(h, n);
  if (r) return r;
  if (h->d_x.ensure(sizeof(double) * D * n)) return fail(h, GRA_E_NOMEM, "input buffer");
  GRA_HIP_CHECK(h, hipMemcpyAsync(h->d_x.p, u, sizeof(double) * D * n, hipMemcpyHostToDevice,
                                  h->stream));
  r = run_integrand(h, nullptr, 0, 0, 0, 0, h->d_x.as<double>(), n, h->d_W.as<double>(),
                    h->d_amp2.as<double>(), h->d_flags.as<uint8_t>(), h->d_vw.as<double>(),
                    h->d_bins.as<uint8_t>());
  if (r) return r;
  GRA_HIP_CHECK(h, hipMemcpyAsync(W, h->d_W.p, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  if (amp2)
    GRA_HIP_CHECK(h, hipMemcpyAsync(amp2, h->d_amp2.p, sizeof(double) * n, hipMemcpyDeviceToHost,
                                    h->stream));
  std::vector<uint8_t> own;
  if (!flags) {
    own.resize(n);
    flags = own.data();
  }
  GRA_HIP_CHECK(h, hipMemcpyAsync(flags, h->d_flags.p, n, hipMemcpyDeviceToHost, h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  time_collect(h, true, h->P.screening != 0, false);
  uint64_t n_nan = 0, n_inf = 0;
  for (size_t i = 0; i < n; ++i)
    if (!(flags[i] & GRA_FLAG_AMP_OK)) ++((flags[i] & GRA_FLAG_INF) ? n_inf : n_nan);
  return count_nonfinite(h, n_nan, n_inf);
}

// MProcess::CheckInfNan (include/Graniitti/MProcess.h:317-340): NaN and Inf
// weights are set to 0 and counted over the life of the process object (here the
// handle); the run fails once either count exceeds 50.
int gra_hip::count_nonfinite(gra_handle* h, uint64_t n_nan, uint64_t n_inf) {
  h->n_nan_seen += n_nan;
  h->n_inf_seen += n_inf;
  if (h->n_nan_seen > GRA_MAX_NONFINITE)
    return fail(h, GRA_E_NUMERIC,
                "CheckInfNan: too many NaN weights (%llu) - check model parameters and cuts",
                (unsigned long long)h->n_nan_seen);
  if (h->n_inf_seen > GRA_MAX_NONFINITE)
    return fail(h, GRA_E_NUMERIC,
                "CheckInfNan: too many Inf weights (%llu) - check model parameters and cuts",
                (unsigned long long)h->n_inf_seen);
  return GRA_OK;
}

static int check_grid(gra_handle* h, const gra_grid* g) {
  if (!g) return fail(h, GRA_E_ARG, "grid is NULL");
  if (g->dim != h->P.dim)
    return fail(h, GRA_E_ARG, "grid dim %d != process dim %d", g->dim, h->P.dim);
  if (g->bins < 2 || g->bins > GRA_VEGAS_BINS || (g->bins % 2))
    return fail(h, GRA_E_ARG, "grid bins %d invalid", g->bins);
  return GRA_OK;
}

// Copies the grid's edges (dim x GRA_VEGAS_BINS) to h->d_grid, unless the buffer
// already holds exactly these values from an earlier call.
static int upload_grid(gra_handle* h, const gra_grid* grid) {
  if (h->d_grid.ensure(sizeof(double) * GRA_MAX_DIM * GRA_VEGAS_BINS))
    return fail(h, GRA_E_NOMEM, "grid buffer");
  const size_t n = (size_t)grid->dim * GRA_VEGAS_BINS;
  if (h->grid_dev_at == h->d_grid.p && h->grid_host.size() == n &&
      std::memcmp(h->grid_host.data(), grid->xmat, n * sizeof(double)) == 0)
    return GRA_OK;
  h->grid_dev_at = nullptr;
  GRA_HIP_CHECK(h, hipMemcpyAsync(h->d_grid.p, grid->xmat, n * sizeof(double),
                                  hipMemcpyHostToDevice, h->stream));
  h->grid_host.assign(grid->xmat, grid->xmat + n);
  h->grid_dev_at = h->d_grid.p;
  return GRA_OK;
}

extern "C" int gra_hip_vegas_iterate(gra_handle* h, const gra_grid* grid, uint64_t n_calls,
                                     uint64_t seed, uint32_t iter, gra_iter_stats* out) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  int r = check_grid(h, grid);
  if (r) return r;
  if (!out) return fail(h, GRA_E_ARG, "vegas_iterate: out is NULL");
  if (n_calls == 0 || n_calls > 0xFFFFFFFFull)
    return fail(h, GRA_E_ARG, "vegas_iterate: n_calls %llu outside [1, 2^32)",
                (unsigned long long)n_calls);
  const int D = h->P.dim;
  r = ensure_point_buffers(h, n_calls);
  if (r) return r;
  // one wave per 512 points, at most 8192 blocks; +32 rows of stage-1 reduction space
  const int nblk = (int)std::min<uint64_t>(8192, std::max<uint64_t>(1, (n_calls + 511) / 512));
  const size_t H = 2 * (size_t)D * GRA_VEGAS_BINS;
  if (h->d_part_hist.ensure(sizeof(double) * (H * (nblk + 32) + 32 * kIterScalars)) ||
      h->d_part_scal.ensure(sizeof(double) * kIterScalars * nblk) ||
      h->d_stats.ensure(sizeof(double) * (kIterScalars + H)))
    return fail(h, GRA_E_NOMEM, "VEGAS buffers");
  r = upload_grid(h, grid);
  if (r) return r;
  r = run_integrand(h, h->d_grid.as<double>(), (unsigned int)grid->bins, seed, iter, 0, nullptr,
                    n_calls, h->d_W.as<double>(), nullptr, h->d_flags.as<uint8_t>(),
                    h->d_vw.as<double>(), h->d_bins.as<uint8_t>());
  if (r) return r;
  time_begin(h, TS_ACCUM);
  GRA_HIP_CHECK(h, launch_accumulate(h->d_W.as<double>(), h->d_vw.as<double>(),
                                     h->d_bins.as<uint8_t>(), h->d_flags.as<uint8_t>(), n_calls, D,
                                     nblk, h->d_part_hist.as<double>(),
                                     h->d_part_scal.as<double>(), h->d_stats.as<double>(),
                                     h->stream));
  time_end(h, TS_ACCUM);
  h->h_stats.resize(kIterScalars + H);
  GRA_HIP_CHECK(h, hipMemcpyAsync(h->h_stats.data(), h->d_stats.p, sizeof(double) * (kIterScalars + H),
                                  hipMemcpyDeviceToHost, h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  time_collect(h, true, h->P.screening != 0, true);
  const double* s = h->h_stats.data();
  std::memset(out, 0, sizeof(*out));
  out->fsum = s[0] - s[1];
  out->fsum_c = s[1];
  out->f2sum = s[2] - s[3];
  out->f2sum_c = s[3];
  out->maxW = s[4];
  out->maxf = s[5];
  out->n_calls = n_calls;
  out->n_kin_ok = (uint64_t)s[6];
  out->n_fid_ok = (uint64_t)s[7];
  out->n_valid = (uint64_t)s[8];
  out->n_nan = (uint64_t)s[9];
  out->n_inf = (uint64_t)s[10];
  for (int j = 0; j < D; ++j)
    for (int b = 0; b < GRA_VEGAS_BINS; ++b) {
      out->fmat[j * GRA_VEGAS_BINS + b] = s[kIterScalars + j * GRA_VEGAS_BINS + b];
      out->f2mat[j * GRA_VEGAS_BINS + b] = s[kIterScalars + (D + j) * GRA_VEGAS_BINS + b];
    }
  return count_nonfinite(h, out->n_nan, out->n_inf);
}

// One integration pass on the GPU for the host schedule
static int gpu_pass(void* ctx, const gra_grid* grid, uint64_t n_calls, uint64_t seed, uint32_t iter,
                    gra_iter_stats* out) {
  gra_handle* h = (gra_handle*)ctx;
  const int r = gra_hip_vegas_iterate(h, grid, n_calls, seed, iter, out);
  if (r) set_global_error(r, "%s", h->err.c_str());
  return r;
}

extern "C" int gra_hip_vegas_run(gra_handle* h, const gra_vegas_cfg* vc, gra_grid* grid,
                                 gra_vegas_result* res) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  const double kt0 = h->time_ms[TS_TREE] + h->time_ms[TS_SCREEN] + h->time_ms[TS_ACCUM];
  const int r = vegas_schedule(vc, grid, h->P.dim, gpu_pass, h, res);
  if (res) res->kernel_ms = h->time_ms[TS_TREE] + h->time_ms[TS_SCREEN] + h->time_ms[TS_ACCUM] - kt0;
  if (r) return fail(h, r, "%s", global_error());
  return GRA_OK;
}

extern "C" int gra_hip_vegas_map(gra_handle* h, const gra_grid* grid, const double* u, size_t n,
                                 double* x, double* vw, int32_t* bins) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  int r = check_grid(h, grid);
  if (r) return r;
  if (n == 0) return GRA_OK;
  if (!u || !x || !vw || !bins) return fail(h, GRA_E_ARG, "vegas_map: NULL buffer");
  const int D = grid->dim;
  DevBuf du, dx, dv, db;
  if (du.ensure(sizeof(double) * D * n) || dx.ensure(sizeof(double) * D * n) ||
      dv.ensure(sizeof(double) * n) || db.ensure(sizeof(int32_t) * D * n))
    return fail(h, GRA_E_NOMEM, "vegas_map buffers");
  r = upload_grid(h, grid);
  if (r) return r;
  GRA_HIP_CHECK(h, hipMemcpyAsync(du.p, u, sizeof(double) * D * n, hipMemcpyHostToDevice, h->stream));
  GRA_HIP_CHECK(h, launch_map(h->d_grid.as<double>(), D, (unsigned int)grid->bins, du.as<double>(),
                              n, dx.as<double>(), dv.as<double>(), db.as<int32_t>(), h->stream));
  GRA_HIP_CHECK(h, hipMemcpyAsync(x, dx.p, sizeof(double) * D * n, hipMemcpyDeviceToHost, h->stream));
  GRA_HIP_CHECK(h, hipMemcpyAsync(vw, dv.p, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  GRA_HIP_CHECK(h, hipMemcpyAsync(bins, db.p, sizeof(int32_t) * D * n, hipMemcpyDeviceToHost,
                                  h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  du.release();
  dx.release();
  dv.release();
  db.release();
  return GRA_OK;
}

extern "C" int gra_hip_generate(gra_handle* h, const gra_grid* grid, uint64_t n_calls,
                                uint64_t seed, uint32_t iter, double maxf, gra_event* events,
                                uint64_t max_events, gra_gen_stats* st) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  int r = check_grid(h, grid);
  if (r) return r;
  if (!st) return fail(h, GRA_E_ARG, "generate: st is NULL");
  if (max_events > 0 && !events) return fail(h, GRA_E_ARG, "generate: events is NULL");
  if (!(maxf > 0.0)) return fail(h, GRA_E_ARG, "generate: maxf must be > 0");
  if (n_calls == 0 || n_calls > 0xFFFFFFFFull)
    return fail(h, GRA_E_ARG, "generate: n_calls %llu outside [1, 2^32)",
                (unsigned long long)n_calls);
  r = ensure_point_buffers(h, n_calls);
  if (r) return r;
  // the accepted indices come out in ascending order (ordered compaction), so
  // the events written are the accepted points of smallest index
  const uint64_t cap = std::min<uint64_t>(max_events, n_calls);
  r = ensure_gen_buffers(h, n_calls, cap);
  if (r) return r;
  r = upload_grid(h, grid);
  if (r) return r;
  r = run_integrand(h, h->d_grid.as<double>(), (unsigned int)grid->bins, seed, iter, 0, nullptr,
                    n_calls, h->d_W.as<double>(), nullptr, h->d_flags.as<uint8_t>(),
                    h->d_vw.as<double>(), h->d_bins.as<uint8_t>());
  if (r) return r;
  unsigned long long* d_cnt = h->d_gen_cnt.as<unsigned long long>();
  GRA_HIP_CHECK(h, launch_accept(h->d_W.as<double>(), h->d_vw.as<double>(),
                                 h->d_flags.as<uint8_t>(), n_calls, maxf, seed, iter, 0,
                                 h->d_gen_mask.as<uint64_t>(), h->d_gen_tile.as<uint32_t>(),
                                 h->d_gen_idx.as<uint64_t>(), cap, d_cnt, h->stream));
  // The event records follow on the same stream: the kernel takes the accepted
  // count (d_cnt[2]) from device memory and writes the first min(count, cap)
  // records. A pinned events buffer that the device sees at the same address is
  // written directly, so the pass ends with one synchronisation; any other
  // buffer is filled from a device copy afterwards.
  bool direct = false;
  if (cap > 0) {
    hipPointerAttribute_t a;
    direct = hipPointerGetAttributes(&a, events) == hipSuccess && a.type == hipMemoryTypeHost &&
             a.devicePointer == (void*)events;
    (void)hipGetLastError();  // an unregistered pointer leaves an error behind
    GRA_HIP_CHECK(h, k64::launch_event_record(h->P, h->d_grid.as<double>(),
                                              (unsigned int)grid->bins, seed, iter, 0, nullptr,
                                              h->d_gen_idx.as<uint64_t>(), h->d_W.as<double>(),
                                              h->d_vw.as<double>(), cap, d_cnt + 2,
                                              direct ? (void*)events : h->d_gen_ev.p, h->stream));
  }
  unsigned long long* c = h->h_gen_cnt;
  GRA_HIP_CHECK(h, hipMemcpyAsync(c, d_cnt, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  time_collect(h, true, h->P.screening != 0, false);
  r = count_nonfinite(h, c[4], c[5]);
  if (r) return r;
  const uint64_t stored = std::min<uint64_t>(c[2], cap);
  std::memset(st, 0, sizeof(*st));
  st->trials = c[0];
  st->valid = c[1];
  st->accepted = c[2];
  st->overflow = c[3];
  st->stored = stored;
  if (stored == 0 || direct) return GRA_OK;
  return events_to_host(h, events, h->d_gen_ev.p, stored);
}

extern "C" int gra_hip_host_alloc(size_t bytes, void** p) {
  if (!p) return set_global_error(GRA_E_ARG, "host_alloc: p is NULL");
  *p = nullptr;
  if (bytes == 0) return set_global_error(GRA_E_ARG, "host_alloc: 0 bytes");
  if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) {
    *p = nullptr;
    return set_global_error(GRA_E_NOMEM, "host_alloc: %zu bytes of pinned memory", bytes);
  }
  return GRA_OK;
}

extern "C" void gra_hip_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

extern "C" int gra_hip_unweight(gra_handle* h, const double* w, size_t n, double wmax,
                                uint64_t seed, uint8_t* accept, uint64_t* n_accepted,
                                uint64_t* n_overflow) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  if (!w || !accept) return fail(h, GRA_E_ARG, "unweight: NULL buffer");
  if (!(wmax > 0.0)) return fail(h, GRA_E_ARG, "unweight: wmax must be > 0");
  if (n == 0) {
    if (n_accepted) *n_accepted = 0;
    if (n_overflow) *n_overflow = 0;
    return GRA_OK;
  }
  DevBuf dw, da, dc;
  if (dw.ensure(sizeof(double) * n) || da.ensure(n) || dc.ensure(2 * sizeof(unsigned long long)))
    return fail(h, GRA_E_NOMEM, "unweight buffers");
  GRA_HIP_CHECK(h, hipMemcpyAsync(dw.p, w, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  GRA_HIP_CHECK(h, hipMemsetAsync(dc.p, 0, 2 * sizeof(unsigned long long), h->stream));
  GRA_HIP_CHECK(h, launch_unweight(dw.as<double>(), n, wmax, seed, da.as<uint8_t>(),
                                   dc.as<unsigned long long>(), h->stream));
  unsigned long long c[2] = {0, 0};
  GRA_HIP_CHECK(h, hipMemcpyAsync(accept, da.p, n, hipMemcpyDeviceToHost, h->stream));
  GRA_HIP_CHECK(h, hipMemcpyAsync(c, dc.p, sizeof(c), hipMemcpyDeviceToHost, h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  if (n_accepted) *n_accepted = c[0];
  if (n_overflow) *n_overflow = c[1];
  return GRA_OK;
}

extern "C" int gra_hip_eikonal_tables(gra_handle* h, double* node_amp, double* mbt, double* msa) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  if (!h->cfg.screening) return fail(h, GRA_E_STATE, "eikonal tables: screening is off");
  if (node_amp) {
    const int NK = h->cfg.n_loop_kt;
    for (int j = 0; j <= NK; ++j) {
      node_amp[2 * j] = h->nodes[j].aeik_re;
      node_amp[2 * j + 1] = h->nodes[j].aeik_im;
    }
  }
  if (mbt) std::memcpy(mbt, h->mbt.data(), sizeof(double) * h->mbt.size());
  if (msa) std::memcpy(msa, h->msa.data(), sizeof(double) * h->msa.size());
  return GRA_OK;
}

extern "C" int gra_hip_screen_opcount(gra_handle* h, uint64_t max_points, double* out) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  if (!out) return fail(h, GRA_E_ARG, "screen_opcount: out is NULL");
  if (max_points == 0 || max_points > 65536)
    return fail(h, GRA_E_ARG, "screen_opcount: max_points %llu outside [1, 65536]",
                (unsigned long long)max_points);
  if (!h->P.screening || h->P.process != GRA_PROC_REGGE_CON || !f32_kernels(h->P))
    return fail(h, GRA_E_STATE,
                "screen_opcount: counts the FP32 Regge screening node (screened PP[CON], "
                "FP32 amplitude)");
  if (!h->d_rec.p || !h->d_ncand.p)
    return fail(h, GRA_E_STATE, "screen_opcount: no screened evaluation has run on this handle");
  DevBuf d_out;
  if (d_out.ensure(12 * sizeof(unsigned long long)))
    return fail(h, GRA_E_NOMEM, "screen_opcount buffer");
  GRA_HIP_CHECK(h, launch_fill_u32(d_out.as<uint32_t>(), 0u, 24, h->stream));
  GRA_HIP_CHECK(h, k32::launch_screen_opcount(h->P, h->d_nodes.p, h->d_rec.p,
                                              h->d_ncand.as<uint32_t>(), (uint32_t)max_points,
                                              d_out.as<unsigned long long>(), h->stream));
  unsigned long long c[12];
  GRA_HIP_CHECK(h, hipMemcpyAsync(c, d_out.p, sizeof(c), hipMemcpyDeviceToHost, h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  for (int k = 0; k < 12; ++k) out[k] = (double)c[k];
  return GRA_OK;
}

extern "C" int gra_hip_kernel_time(gra_handle* h, int32_t which, double* total_ms,
                                   uint64_t* launches, int32_t reset) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  if (which < 0 || which >= TS_N) return fail(h, GRA_E_ARG, "kernel_time: slot %d", which);
  if (total_ms) *total_ms = h->time_ms[which];
  if (launches) *launches = h->launches[which];
  if (reset) {
    h->time_ms[which] = 0;
    h->launches[which] = 0;
  }
  return GRA_OK;
}
