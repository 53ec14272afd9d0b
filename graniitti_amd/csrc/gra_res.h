// Resonance amplitudes of PP[RES+CON] in FP64: A = MRegge::ME4(+1) + sum_r
// {MRegge::ME3 | MRegge::PhotoME3}(r) (PROC_17::Amp2, include/Graniitti/MSubProc.h:398-432),
// evaluated at the tree point and at every screening-loop node. Resonances are
// summed in table order.
#pragma once

#include "gra_device.h"

__device__ __forceinline__ double fact_small(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= (double)i;
  return f;
}
__device__ __forceinline__ double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; ++i) r *= x;
  return r;
}
// Wigner small d^J_{m m'}(theta) as the finite sum over k (spin::WignerSmalld,
// src/MSpin.cc:927-960), integer J <= 2
// with c = cos(theta / 2), s = sin(theta / 2)
__device__ __forceinline__ double wigner_d(double c, double s, int m, int mp, int J) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k <= 2 * J; ++k) {
    if (J - mp - k >= 0 && J + m - k >= 0 && k + mp - m >= 0) {
      const double sg = ((mp - m + k) & 1) ? -1.0 : 1.0;
      const double v1 = sg / (fact_small(J - mp - k) * fact_small(J + m - k) *
                              fact_small(k + mp - m) * fact_small(k));
      acc += v1 * (ipow(c, 2 * J + m - mp - 2 * k) * ipow(s, mp - m + 2 * k));
    }
  }
  return sqrt(fact_small(J + m) * fact_small(J - m) * fact_small(J + mp) * fact_small(J - mp)) *
         acc;
}
// M4Vec::Theta / Phi (include/Graniitti/M4Vec.h:150-155)
__device__ __forceinline__ void theta_phi(dv4 p, double& th, double& ph) {
  ph = (p.x == 0.0 && p.y == 0.0) ? 0.0 : atan2(p.y, p.x);
  th = (p.x == 0.0 && p.y == 0.0 && p.z == 0.0) ? 0.0 : atan2(ptof(p), p.z);
}
// kinematics::LorentzBoost(b, M0, p, -1): p into the rest frame of b
// (include/Graniitti/MKinematics.h:384-401)
__device__ __forceinline__ dv4 boost_rest(dv4 b, double M0, dv4 p) {
  const double bx = -1.0 * b.x / b.e, by = -1.0 * b.y / b.e, bz = -1.0 * b.z / b.e;
  const double g = b.e / M0;
  const double k1 = (bx * p.x + by * p.y) + bz * p.z;
  const double k2 = g * (g * k1 / (1.0 + g) + p.e);
  return v4m(p.x + k2 * bx, p.y + k2 * by, p.z + k2 * bz, g * (p.e + k1));
}
// spin::GetRhoRotation (src/MSpin.cc:544-610): the angles of the quantisation
// axis of the spin density in the system rest frame
__device__ __forceinline__ void rho_rotation(const RunParams& P, const Kin& K, double M, dv4 q1b,
                                             double& thR, double& phR) {
  thR = 0.0;
  phR = 0.0;
  if (P.spin_frame == GRA_FRAME_HX) {
    theta_phi(K.pf[0], thR, phR);
  } else if (P.spin_frame == GRA_FRAME_CS) {
    const dv4 b1 = boost_rest(K.pf[0], M, P.pbeam1), b2 = boost_rest(K.pf[0], M, P.pbeam2);
    const double n1 = sqrt(p3mod2(b1)), n2 = sqrt(p3mod2(b2));
    const double zx = b1.x / n1 - b2.x / n2, zy = b1.y / n1 - b2.y / n2, zz = b1.z / n1 - b2.z / n2;
    const double nz = sqrt(zx * zx + zy * zy + zz * zz);
    theta_phi(v4m(zx / nz, zy / nz, zz / nz, 0.0), thR, phR);
  } else if (P.spin_frame == GRA_FRAME_GJ) {
    theta_phi(q1b, thR, phR);
  }
}
// Wigner d-matrices of the rotation, dR[a][j] = d^J_{M_a M_j}(thR) for J = 1, 2 (rows and
// columns -J..J), shared by every resonance and by production and decay at one point
struct RotD {
  double d1[3][3], d2[5][5];
  double phR;
};
__device__ __forceinline__ void rot_d(double thR, double phR, RotD& D) {
  const double c = cos(thR / 2.0), s = sin(thR / 2.0);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int j = 0; j < 3; ++j) D.d1[a][j] = wigner_d(c, s, a - 1, j - 1, 1);
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int j = 0; j < 5; ++j) D.d2[a][j] = wigner_d(c, s, a - 2, j - 2, 2);
  D.phR = phR;
}
// Re Tr[f rho_rot f^dagger] with rho_rot = D^dagger rho D, D_aj = D^J(thR, phR, M_a, M_j),
// and rows f_rj = T_r D^J(th, ph, lam_r, M_j) (spin::fMatrix, DMatrix, WignerD,
// src/MSpin.cc:631-705, 968-970), contracted as sum_ab rho_ab g_a conj(g_b) with
// g_a = sum_j f_rj conj(D_aj)
template <int J, int NROW>
__device__ __forceinline__ double spin_trace(const ResDev& R, const int* lam, const double* T,
                                             double th, double ph, const RotD& D) {
  constexpr int n = 2 * J + 1;
  const double c = cos(th / 2.0), s = sin(th / 2.0);
  // e^{i M (ph - phR)} for M = -J..J: the rotation's phase folded into the rows
  const double dp = ph - D.phR;
  const cxd e1 = cd(cos(dp), sin(dp));
  cxd ph_[n];
  ph_[J] = cd(1.0, 0.0);
#pragma unroll
  for (int k = 1; k <= J; ++k) {
    ph_[J + k] = cmul(ph_[J + k - 1], e1);
    ph_[J - k] = cd(ph_[J + k].re, -ph_[J + k].im);
  }
  double tr = 0.0;
#pragma unroll
  for (int r = 0; r < NROW; ++r) {
    cxd f[n];
#pragma unroll
    for (int j = 0; j < n; ++j) f[j] = cscale(ph_[j], T[r] * wigner_d(c, s, lam[r], j - J, J));
    cxd g[n];
#pragma unroll
    for (int a = 0; a < n; ++a) {
      cxd acc = cd(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < n; ++j) acc = cadd(acc, cscale(f[j], J == 1 ? D.d1[a][j] : D.d2[a][j]));
      g[a] = acc;
    }
#pragma unroll
    for (int a = 0; a < n; ++a)
#pragma unroll
      for (int b = 0; b < n; ++b) {
        const cxd rab = cd(R.rho_re[a * n + b], R.rho_im[a * n + b]);
        tr += cmul(rab, cmul(g[a], cd(g[b].re, -g[b].im))).re;
      }
  }
  return tr;
}
// spin::ProdAmp (src/MSpin.cc:280-403) for J^P = 0+, 1-, 2+: "Pomeron effective
// spins" s1 = s2 = 0 with T0 = 1 (tensors), s1 = 0, s2 = 1 with the fixed T0 row (vectors)
__device__ __forceinline__ cxd prod_amp(const RunParams& P, const ResDev& R, double thq,
                                        double phq, const RotD& D) {
  const cxd A0 = cd(R.g_re, R.g_im);
  if (!P.spin_gen || R.J == 0) return A0;
  double tr;
  if (R.J == 1) {
    // s1 = 0, s2 = 1: lambda = 0 - lambda2 for lambda2 = -1, 0, 1, with the fixed T0 row
    const int lam[3] = {1, 0, -1};
    const double T[3] = {-0.697, -0.169, 0.697};
    tr = spin_trace<1, 3>(R, lam, T, thq, phq, D);
  } else {
    const int lam[1] = {0};
    const double T[1] = {1.0};
    tr = spin_trace<2, 1>(R, lam, T, thq, phq, D);
  }
  return cscale(A0, sqrt(2.0 * R.J + 1.0) * msq(tr));
}
// spin::DecayAmp (src/MSpin.cc:406-486) for two spinless daughters: f = fX, 1 x (2J+1).
// |T_00| = 1 leaves the trace unchanged, so only J = 0 carries T_00's phase.
__device__ __forceinline__ cxd decay_amp(const RunParams& P, const ResDev& R, double thd,
                                         double phd, const RotD& D) {
  const double A0 = R.gdec;
  if (!P.spin_dec) return cd(A0, 0.0);
  if (R.J == 0) return cd(R.T_re * A0, R.T_im * A0);
  const int lam[1] = {0};
  const double T[1] = {sqrt(R.T_re * R.T_re + R.T_im * R.T_im)};
  const double tr = R.J == 1 ? spin_trace<1, 1>(R, lam, T, thd, phd, D)
                             : spin_trace<2, 1>(R, lam, T, thd, phd, D);
  return cd(A0 * (sqrt(2.0 * R.J + 1.0) * msq(tr)), 0.0);
}
// form::CBW (src/MForm.cc:743-828)
__device__ __forceinline__ cxd cbw(const ResDev& R, double m2, double mA, double mB) {
  const double a = m2 - R.M0 * R.M0;
  double b, num = -1.0;
  if (R.BW == GRA_BW_FIXED) {
    b = R.M0 * R.W;
  } else if (R.BW == GRA_BW_RUNNING) {
    b = sqrt(m2) * R.W;
  } else if (R.BW == GRA_BW_BARRIER) {
    const double u = msq((m2 - (mA + mB) * (mA + mB)) * (m2 - (mA - mB) * (mA - mB))) / (2 * msq(m2));
    const double d = msq((R.M0 * R.M0 - (mA + mB) * (mA + mB)) * (R.M0 * R.M0 - (mA - mB) * (mA - mB))) /
                     (2 * R.M0);
    b = R.W * R.M0 * R.M0 / msq(m2) * pow(u / d, 2 * R.J + 1);
  } else {
    b = R.M0 * R.W;
    const double r = m2 / (R.M0 * R.M0);
    if (R.J == 1) num = -(1.0 - r);
    if (R.J == 2) num = -(7.0 / 6.0 - (4.0 / 3.0) * r + (2.0 / 3.0) * (m2 * m2) / (R.M02 * R.M02));
  }
  const double den = a * a + b * b;
  return cd(num * a / den, -num * b / den);
}
// form::G_E_KELLY, G_M_KELLY (src/MForm.cc:607-645) and form::CohFlux (src/MForm.cc:664-687)
__device__ __forceinline__ double coh_flux(double xi, double t, double pt, double mp) {
  const double pt2 = pt * pt, xi2 = xi * xi, mp2 = mp * mp, Q2 = fabs(t);
  const double tau = Q2 / ((2 * mp) * (2 * mp));
  const double GE = (1.0 + -0.24 * tau) / (1.0 + 10.98 * tau + 12.82 * (tau * tau) + 21.97 * (tau * tau * tau));
  const double GM = 2.792847337 * ((1.0 + 0.12 * tau) /
                                   (1.0 + 10.97 * tau + 18.86 * (tau * tau) + 6.55 * (tau * tau * tau)));
  const double PART1 = (4.0 * mp2 * (GE * GE) + Q2 * (GM * GM)) / (4.0 * mp2 + Q2);
  const double PART2 = GM * GM;
  const double DELTA = pt2 / (pt2 + xi2 * mp2);
  double f = 0.0072973525693 / GRA_PI *
             ((1.0 - xi) * (DELTA * DELTA) * PART1 + (xi2 / 4.0) * DELTA * PART2);
  f /= xi;
  f /= pt2;
  f *= 16.0 * (GRA_PI * GRA_PI);
  return f;
}
// MRegge::PhotoProp, elastic forward leg (src/MRegge.cc:849-880): the leading trajectory
__device__ __forceinline__ cxd photo_prop(const RunParams& P, double s, double t, double B0) {
  const double alpha = P.regge_a0[0] + P.regge_ap[0] * t;
  const double th = ((-GRA_PI / 2.0) * P.regge_ap[0]) * t;
  const cxd eta = cmul(cd(P.eta0_re[0], P.eta0_im[0]), cd(cos(th), sin(th)));
  const double FF = P.gN_P * exp(B0 * t / 2.0);
  return cscale(cscale(eta, pow(s / 1.0, alpha)), FF);
}

__device__ __forceinline__ cxd res_sum(const RunParams& P, const Kin& K) {
  const dv4 X = K.pf[0];
  const double m2 = m2of(X), M = mof(X);
  const double s1 = m2of(v4add(X, K.pf[1])), s2 = m2of(v4add(X, K.pf[2]));
  const double FF_A = P.gN_P * s3f(P, K.t1), FF_B = P.gN_P * s3f(P, K.t2);
  const cxd pr1 = regge_prop(P, s1, K.t1), pr2 = regge_prop(P, s2, K.t2);
  const double V = pow(1.0 / m2, P.omega);
  const dv4 q1b = boost_rest(X, M, K.q1);
  double thq, phq, thd, phd, thR, phR;
  theta_phi(q1b, thq, phq);
  theta_phi(boost_rest(X, M, K.dt[0]), thd, phd);
  rho_rotation(P, K, M, q1b, thR, phR);
  RotD D;
  rot_d(thR, phR, D);
  const double mA = mof(K.dt[0]), mB = mof(K.dt[1]);
  cxd A = cd(0.0, 0.0);
  for (int r = 0; r < P.nres; ++r) {
    const ResDev& R = P.res[r];
    const cxd bw = cbw(R, m2, mA, mB);
    const double rff = exp(-((m2 - R.M02) * (m2 - R.M02)) / R.gFF4);
    const cxd spin = cmul(prod_amp(P, R, thq, phq, D), decay_amp(P, R, thd, phd, D));
    if (R.photo) {
      const dv4 q2 = v4sub(P.pbeam2, K.pf[2]);
      const double x1 = 1 - K.pf[1].z / P.pbeam1.z, x2 = 1 - K.pf[2].z / P.pbeam2.z;
      const double gf1 = msq(coh_flux(x1, K.t1, ptof(K.q1), P.tp_mp_ff) / x1);
      const double gf2 = msq(coh_flux(x2, K.t2, ptof(q2), P.tp_mp_ff) / x2);
      const cxd common = cscale(bw, rff);
      const cxd A1 = cmul(cscale(common, gf2), photo_prop(P, s1, K.t1, R.B0));
      const cxd A2 = cmul(cscale(common, gf1), photo_prop(P, s2, K.t2, R.B0));
      // identical beams (pp): A_1 + A_2 (src/MRegge.cc:813)
      A = cadd(A, cmul(cadd(A1, A2), spin));
    } else {
      cxd Ap = cscale(pr1, 2.0);
      Ap = cscale(Ap, FF_A);
      Ap = cmul(Ap, bw);
      Ap = cscale(Ap, rff);
      Ap = cmul(Ap, pr2);
      Ap = cscale(Ap, FF_B);
      A = cadd(A, cscale(cmul(Ap, spin), V));
    }
  }
  return A;
}

// ---------------------------------------------------------------------------
// The same amplitude in FP32 on FP64 kinematics (the screening node of the
// FP32-amplitude mode): every quantity formed from the 6.5 TeV beam momenta or
// sensitive to cancellation (m^2, s_1 = (X + p1)^2, s_2, t1, t2, the rest-frame
// angles, the photon momentum fractions x = 1 - p_z / p_beam,z) is computed in
// FP64 by res_kin (the rest-frame vectors in FP64, their angles in FP32), and
// the amplitude from these scalars in FP32, with the
// transcendentals on the hardware units and FMA contraction on (the unit that
// includes this file compiles its FP64 code without contraction).
// ---------------------------------------------------------------------------
struct ResKin {
  double m2, M, s1, s2, t1, t2, mA, mB, x1, x2, pt1, pt2;
  float thq, phq, thd, phd, thR, phR;
};
// M4Vec::Theta / Phi of an FP64 vector in FP32 (the vectors are O(GeV) momenta
// in the system rest frame, or unit vectors)
__device__ __forceinline__ void theta_phi32(dv4 p, float& th, float& ph) {
  const float x = (float)p.x, y = (float)p.y, z = (float)p.z;
  ph = (x == 0.f && y == 0.f) ? 0.f : atan2f(y, x);
  th = (x == 0.f && y == 0.f && z == 0.f) ? 0.f : atan2f(sqrtf(x * x + y * y), z);
}
// spin::GetRhoRotation with the boosts in FP64 and the angles in FP32
__device__ __forceinline__ void rho_rotation32(const RunParams& P, const Kin& K, double M, dv4 q1b,
                                               float& thR, float& phR) {
  thR = 0.f;
  phR = 0.f;
  if (P.spin_frame == GRA_FRAME_HX) {
    theta_phi32(K.pf[0], thR, phR);
  } else if (P.spin_frame == GRA_FRAME_CS) {
    const dv4 b1 = boost_rest(K.pf[0], M, P.pbeam1), b2 = boost_rest(K.pf[0], M, P.pbeam2);
    const double n1 = sqrt(p3mod2(b1)), n2 = sqrt(p3mod2(b2));
    theta_phi32(v4m(b1.x / n1 - b2.x / n2, b1.y / n1 - b2.y / n2, b1.z / n1 - b2.z / n2, 0.0),
                thR, phR);
  } else if (P.spin_frame == GRA_FRAME_GJ) {
    theta_phi32(q1b, thR, phR);
  }
}
__device__ __forceinline__ void res_kin(const RunParams& P, const Kin& K, ResKin& k) {
  const dv4 X = K.pf[0];
  k.m2 = m2of(X);
  k.M = mof(X);
  k.s1 = m2of(v4add(X, K.pf[1]));
  k.s2 = m2of(v4add(X, K.pf[2]));
  k.t1 = K.t1;
  k.t2 = K.t2;
  const dv4 q1b = boost_rest(X, k.M, K.q1);
  theta_phi32(q1b, k.thq, k.phq);
  theta_phi32(boost_rest(X, k.M, K.dt[0]), k.thd, k.phd);
  rho_rotation32(P, K, k.M, q1b, k.thR, k.phR);
  k.mA = mof(K.dt[0]);
  k.mB = mof(K.dt[1]);
  k.x1 = 1 - K.pf[1].z / P.pbeam1.z;
  k.x2 = 1 - K.pf[2].z / P.pbeam2.z;
  k.pt1 = ptof(K.q1);
  k.pt2 = ptof(v4sub(P.pbeam2, K.pf[2]));
}

__device__ __forceinline__ cxf rf_add(cxf a, cxf b) { return cf(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cxf rf_scale(cxf a, float s) { return cf(a.re * s, a.im * s); }
__device__ __forceinline__ float rf_pow(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
// x^n for a small non-negative integer n
__device__ __forceinline__ float rf_ipow(float x, int n) {
  float r = 1.f;
  for (int i = 0; i < n; ++i) r *= x;
  return r;
}

__device__ __forceinline__ float wigner_d32(float c, float s, int m, int mp, int J) {
#pragma clang fp contract(fast)
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k <= 2 * J; ++k) {
    if (J - mp - k >= 0 && J + m - k >= 0 && k + mp - m >= 0) {
      const float sg = ((mp - m + k) & 1) ? -1.f : 1.f;
      const float v1 = sg / (float)(fact_small(J - mp - k) * fact_small(J + m - k) *
                                    fact_small(k + mp - m) * fact_small(k));
      acc += v1 * (rf_ipow(c, 2 * J + m - mp - 2 * k) * rf_ipow(s, mp - m + 2 * k));
    }
  }
  return sqrtf((float)(fact_small(J + m) * fact_small(J - m) * fact_small(J + mp) *
                       fact_small(J - mp))) *
         acc;
}
struct RotD32 {
  float d1[3][3], d2[5][5];
  float phR;
};
__device__ __forceinline__ void rot_d32(float thR, float phR, RotD32& D) {
#pragma clang fp contract(fast)
  float s, c;
  __sincosf(0.5f * thR, &s, &c);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int j = 0; j < 3; ++j) D.d1[a][j] = wigner_d32(c, s, a - 1, j - 1, 1);
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int j = 0; j < 5; ++j) D.d2[a][j] = wigner_d32(c, s, a - 2, j - 2, 2);
  D.phR = phR;
}
template <int J, int NROW>
__device__ __forceinline__ float spin_trace32(const ResDev& R, const int* lam, const float* T,
                                              float th, float ph, const RotD32& D) {
#pragma clang fp contract(fast)
  constexpr int n = 2 * J + 1;
  float s, c, sd, cd_;
  __sincosf(0.5f * th, &s, &c);
  __sincosf(ph - D.phR, &sd, &cd_);
  cxf ph_[n];
  ph_[J] = cf(1.f, 0.f);
#pragma unroll
  for (int k = 1; k <= J; ++k) {
    ph_[J + k] = cfmul(ph_[J + k - 1], cf(cd_, sd));
    ph_[J - k] = cf(ph_[J + k].re, -ph_[J + k].im);
  }
  float tr = 0.f;
#pragma unroll
  for (int r = 0; r < NROW; ++r) {
    cxf f[n];
#pragma unroll
    for (int j = 0; j < n; ++j) f[j] = rf_scale(ph_[j], T[r] * wigner_d32(c, s, lam[r], j - J, J));
    cxf g[n];
#pragma unroll
    for (int a = 0; a < n; ++a) {
      cxf acc = cf(0.f, 0.f);
#pragma unroll
      for (int j = 0; j < n; ++j) acc = rf_add(acc, rf_scale(f[j], J == 1 ? D.d1[a][j] : D.d2[a][j]));
      g[a] = acc;
    }
#pragma unroll
    for (int a = 0; a < n; ++a)
#pragma unroll
      for (int b = 0; b < n; ++b) {
        const cxf rab = cf((float)R.rho_re[a * n + b], (float)R.rho_im[a * n + b]);
        tr += cfmul(rab, cfmul(g[a], cf(g[b].re, -g[b].im))).re;
      }
  }
  return tr;
}
__device__ __forceinline__ cxf prod_amp32(const RunParams& P, const ResDev& R, float thq,
                                          float phq, const RotD32& D) {
#pragma clang fp contract(fast)
  const cxf A0 = cf((float)R.g_re, (float)R.g_im);
  if (!P.spin_gen || R.J == 0) return A0;
  float tr;
  if (R.J == 1) {
    const int lam[3] = {1, 0, -1};
    const float T[3] = {-0.697f, -0.169f, 0.697f};
    tr = spin_trace32<1, 3>(R, lam, T, thq, phq, D);
  } else {
    const int lam[1] = {0};
    const float T[1] = {1.f};
    tr = spin_trace32<2, 1>(R, lam, T, thq, phq, D);
  }
  return rf_scale(A0, sqrtf(2.f * R.J + 1.f) * sqrtf(fmaxf(tr, 0.f)));
}
__device__ __forceinline__ cxf decay_amp32(const RunParams& P, const ResDev& R, float thd,
                                           float phd, const RotD32& D) {
#pragma clang fp contract(fast)
  const float A0 = (float)R.gdec;
  if (!P.spin_dec) return cf(A0, 0.f);
  if (R.J == 0) return cf((float)R.T_re * A0, (float)R.T_im * A0);
  const int lam[1] = {0};
  const float T[1] = {(float)sqrt(R.T_re * R.T_re + R.T_im * R.T_im)};
  const float tr = R.J == 1 ? spin_trace32<1, 1>(R, lam, T, thd, phd, D)
                            : spin_trace32<2, 1>(R, lam, T, thd, phd, D);
  return cf(A0 * (sqrtf(2.f * R.J + 1.f) * sqrtf(fmaxf(tr, 0.f))), 0.f);
}
__device__ __forceinline__ cxf cbw32(const ResDev& R, float m2, float mA, float mB) {
#pragma clang fp contract(fast)
  const float M0 = (float)R.M0, W = (float)R.W;
  const float a = m2 - (float)R.M02;
  float b, num = -1.f;
  if (R.BW == GRA_BW_FIXED) {
    b = M0 * W;
  } else if (R.BW == GRA_BW_RUNNING) {
    b = sqrtf(m2) * W;
  } else if (R.BW == GRA_BW_BARRIER) {
    const float sp = (mA + mB) * (mA + mB), sm = (mA - mB) * (mA - mB), M02 = (float)R.M02;
    const float u = sqrtf(fmaxf((m2 - sp) * (m2 - sm), 0.f)) / (2.f * sqrtf(m2));
    const float d = sqrtf(fmaxf((M02 - sp) * (M02 - sm), 0.f)) / (2.f * M0);
    b = W * M02 / sqrtf(m2) * rf_ipow(u / d, 2 * R.J + 1);
  } else {
    b = M0 * W;
    const float r = m2 / (float)R.M02;
    if (R.J == 1) num = -(1.f - r);
    if (R.J == 2) num = -(7.f / 6.f - (4.f / 3.f) * r + (2.f / 3.f) * (r * r));
  }
  const float iden = __builtin_amdgcn_rcpf(a * a + b * b);
  return cf(num * a * iden, -num * b * iden);
}
__device__ __forceinline__ float coh_flux32(float xi, float t, float pt, float mp) {
#pragma clang fp contract(fast)
  const float pt2 = pt * pt, xi2 = xi * xi, mp2 = mp * mp, Q2 = fabsf(t);
  const float tau = Q2 / (4.f * mp2);
  const float GE = (1.f - 0.24f * tau) / (1.f + tau * (10.98f + tau * (12.82f + 21.97f * tau)));
  const float GM = 2.792847337f * ((1.f + 0.12f * tau) /
                                   (1.f + tau * (10.97f + tau * (18.86f + 6.55f * tau))));
  const float PART1 = (4.f * mp2 * (GE * GE) + Q2 * (GM * GM)) / (4.f * mp2 + Q2);
  const float DELTA = pt2 / (pt2 + xi2 * mp2);
  const float f = (float)(0.0072973525693 / GRA_PI * 16.0 * GRA_PI * GRA_PI) *
                  ((1.f - xi) * (DELTA * DELTA) * PART1 + (xi2 / 4.f) * DELTA * (GM * GM));
  return f / (xi * pt2);
}
__device__ __forceinline__ cxf photo_prop32(const RunParams& P, float s, float t, float B0) {
#pragma clang fp contract(fast)
  const float alpha = (float)P.regge_a0[0] + (float)P.regge_ap[0] * t;
  float sn, cs;
  __sincosf((float)((-GRA_PI / 2.0) * P.regge_ap[0]) * t, &sn, &cs);
  const cxf eta = cfmul(cf((float)P.eta0_re[0], (float)P.eta0_im[0]), cf(cs, sn));
  const float FF = (float)P.gN_P * __expf(0.5f * B0 * t);
  return rf_scale(eta, rf_pow(s, alpha) * FF);
}
__device__ __forceinline__ float s3f32(const RunParams& P, float t) {
#pragma clang fp contract(fast)
  return __builtin_amdgcn_rcpf((1.f - t / (float)P.fc1) * (1.f - t / (float)P.fc2));
}
__device__ __forceinline__ cxf regge_prop32(const RunParams& P, float s, float t) {
#pragma clang fp contract(fast)
  cxf M = cf(0.f, 0.f);
  const float ls = __builtin_amdgcn_logf(s * (float)(1.0 / P.regge_s0));
  for (int k = 0; k < 3; ++k) {
    if (P.regge_on[k]) {
      const float alpha = (float)P.regge_a0[k] + (float)P.regge_ap[k] * t;
      float sn, cs;
      __sincosf((float)((-GRA_PI / 2.0) * P.regge_ap[k]) * t, &sn, &cs);
      const cxf eta = cfmul(cf((float)P.eta0_re[k], (float)P.eta0_im[k]), cf(cs, sn));
      M = rf_add(M, rf_scale(eta, __builtin_amdgcn_exp2f(alpha * ls)));
    }
  }
  return M;
}
// MRegge::ME4 in FP32 from the FP64 invariants of K
__device__ __forceinline__ cxf regge_me4_32(const RunParams& P, const Kin& K) {
#pragma clang fp contract(fast)
  const float t1 = (float)K.t1, t2 = (float)K.t2;
  const float FF = (float)P.gN_P, g = (float)P.gpp_P;
  const float th_ = (float)(K.t_hat - P.m_prop2), tu_ = (float)(K.u_hat - P.m_prop2);
  float fft, ffu;
  if (P.offshellFF == 2) {
    fft = __builtin_amdgcn_rcpf(1.f - th_ / (float)P.b_POW);
    ffu = __builtin_amdgcn_rcpf(1.f - tu_ / (float)P.b_POW);
  } else if (P.offshellFF == 0) {
    fft = __expf(-(float)P.b_EXP * fabsf(th_));
    ffu = __expf(-(float)P.b_EXP * fabsf(tu_));
  } else {
    const float ao = (float)P.a_OREAR, bo = (float)P.b_OREAR;
    fft = __expf(-bo * sqrtf(fabsf(th_) + ao * ao) + ao * bo);
    ffu = __expf(-bo * sqrtf(fabsf(tu_) + ao * ao) + ao * bo);
  }
  const float cA = FF * s3f32(P, t1), cB = FF * s3f32(P, t2);
  const cxf At = rf_scale(cfmul(regge_prop32(P, (float)K.ss13, t1), regge_prop32(P, (float)K.ss24, t2)),
                          cA * cB * (g * g) * (fft * fft) / th_);
  const cxf Au = rf_scale(cfmul(regge_prop32(P, (float)K.ss14, t1), regge_prop32(P, (float)K.ss23, t2)),
                          cA * cB * (g * g) * (ffu * ffu) / tu_);
  return rf_add(At, Au);
}
// The resonance sum of res_sum, FP64 kinematics and FP32 amplitude
__device__ __forceinline__ cxf res_sum32(const RunParams& P, const Kin& K) {
#pragma clang fp contract(fast)
  ResKin k;
  res_kin(P, K, k);
  const float m2 = (float)k.m2, s1 = (float)k.s1, s2 = (float)k.s2;
  const float t1 = (float)k.t1, t2 = (float)k.t2;
  const float FF_A = (float)P.gN_P * s3f32(P, t1), FF_B = (float)P.gN_P * s3f32(P, t2);
  const cxf pr12 = cfmul(regge_prop32(P, s1, t1), regge_prop32(P, s2, t2));
  const float V = rf_pow(1.f / m2, (float)P.omega);
  RotD32 D;
  rot_d32(k.thR, k.phR, D);
  const float thq = k.thq, phq = k.phq, thd = k.thd, phd = k.phd;
  const float mA = (float)k.mA, mB = (float)k.mB;
  const float cont = 2.f * FF_A * FF_B * V;
  float gf1 = 0.f, gf2 = 0.f;
  bool flux = false;
  cxf A = cf(0.f, 0.f);
  for (int r = 0; r < P.nres; ++r) {
    const ResDev& R = P.res[r];
    const cxf bw = cbw32(R, m2, mA, mB);
    const float dm = m2 - (float)R.M02;
    const float rff = __expf(-(dm * dm) / (float)R.gFF4);
    const cxf spin = cfmul(prod_amp32(P, R, thq, phq, D), decay_amp32(P, R, thd, phd, D));
    if (R.photo) {
      if (!flux) {
        const float x1 = (float)k.x1, x2 = (float)k.x2, mp = (float)P.tp_mp_ff;
        gf1 = sqrtf(fmaxf(coh_flux32(x1, t1, (float)k.pt1, mp) / x1, 0.f));
        gf2 = sqrtf(fmaxf(coh_flux32(x2, t2, (float)k.pt2, mp) / x2, 0.f));
        flux = true;
      }
      const float B0 = (float)R.B0;
      const cxf common = rf_scale(bw, rff);
      const cxf A12 = rf_add(rf_scale(cfmul(common, photo_prop32(P, s1, t1, B0)), gf2),
                             rf_scale(cfmul(common, photo_prop32(P, s2, t2, B0)), gf1));
      A = rf_add(A, cfmul(A12, spin));
    } else {
      A = rf_add(A, rf_scale(cfmul(cfmul(pr12, bw), spin), cont * rff));
    }
  }
  return A;
}
