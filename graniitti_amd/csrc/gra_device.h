// Device-side physics of the continuum 2 -> 4 event weight (gfx950).
//
// Included by each kernel translation unit inside an unnamed namespace, so every
// unit keeps its own copy compiled with its own floating-point contraction mode:
// the FP64 reference-order unit is built with -ffp-contract=off, the mixed
// FP64/FP32 unit with contraction on.
//
// 4-vectors are stored as (E, x, y, z) like gra::M4Vec (include/Graniitti/M4Vec.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gra_params.h"

namespace {

// ---------------------------------------------------------------------------
// Small types
// ---------------------------------------------------------------------------
typedef GV4 dv4;
__device__ __forceinline__ dv4 v4m(double x, double y, double z, double e) {
  dv4 v;
  v.e = e;
  v.x = x;
  v.y = y;
  v.z = z;
  return v;
}
__device__ __forceinline__ dv4 v4add(dv4 a, dv4 b) {
  return v4m(a.x + b.x, a.y + b.y, a.z + b.z, a.e + b.e);
}
__device__ __forceinline__ dv4 v4sub(dv4 a, dv4 b) {
  return v4m(a.x - b.x, a.y - b.y, a.z - b.z, a.e - b.e);
}
__device__ __forceinline__ dv4 v4scale(dv4 a, double d) {
  return v4m(a.x * d, a.y * d, a.z * d, a.e * d);
}
__device__ __forceinline__ double msq(double x) { return sqrt(x > 0.0 ? x : 0.0); }
__device__ __forceinline__ double p3mod2(dv4 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ double m2of(dv4 a) { return a.e * a.e - p3mod2(a); }
__device__ __forceinline__ double mof(dv4 a) {
  const double m2 = m2of(a);
  return (m2 > 0.0) ? msq(m2) : -msq(-m2);
}
__device__ __forceinline__ double pt2of(dv4 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double ptof(dv4 a) { return msq(pt2of(a)); }
__device__ __forceinline__ double rapof(dv4 a) { return 0.5 * log((a.e + a.z) / (a.e - a.z)); }
__device__ __forceinline__ double etaof(dv4 a) {
  const double p = msq(p3mod2(a));
  return 0.5 * log((p + a.z) / (p - a.z));
}
__device__ __forceinline__ double mtof(dv4 a) { return msq(m2of(a) + pt2of(a)); }

struct cxd {
  double re, im;
};
__device__ __forceinline__ cxd cd(double r, double i) {
  cxd z;
  z.re = r;
  z.im = i;
  return z;
}
__device__ __forceinline__ cxd cadd(cxd a, cxd b) { return cd(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cxd cmul(cxd a, cxd b) {
  return cd(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cxd cscale(cxd a, double s) { return cd(a.re * s, a.im * s); }

struct cxf {
  float re, im;
};
__device__ __forceinline__ cxf cf(float r, float i) {
  cxf z;
  z.re = r;
  z.im = i;
  return z;
}
__device__ __forceinline__ cxf cfmul(cxf a, cxf b) {
  return cf(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and 53-bit uniforms
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                         uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
}
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
  const uint64_t x = ((((uint64_t)hi) << 32) | lo) >> 11;
  return (double)x * (1.0 / 9007199254740992.0);
}
// Uniforms of point idx in iteration iter (counter = (block, idx_lo, idx_hi, iter), key = seed)
__device__ __forceinline__ void point_uniforms(uint64_t seed, uint32_t iter, uint64_t idx, int dim,
                                               double* u) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int b = 0; 2 * b < dim; ++b) {
    uint32_t c0 = (uint32_t)b, c1 = (uint32_t)idx, c2 = (uint32_t)(idx >> 32), c3 = iter;
    philox10(c0, c1, c2, c3, k0, k1);
    u[2 * b] = u53(c0, c1);
    if (2 * b + 1 < dim) u[2 * b + 1] = u53(c2, c3);
  }
}

// ---------------------------------------------------------------------------
// VEGAS map, MGraniitti::VEGASMultiThread (src/MGraniitti.cc:1355-1375)
// xm: grid edges in LDS, dimension-major with row stride GRA_VEGAS_BINS
// ---------------------------------------------------------------------------
__device__ __forceinline__ double vegas_map(const double* xm, int dim, unsigned int bins,
                                            const double* u, double* x, uint8_t* bin) {
  double vw = 1.0;
  for (int j = 0; j < dim; ++j) {
    const double zn = u[j] * bins + 1.0;
    unsigned int k = (unsigned int)zn;
    k = k > bins ? bins : k;
    k = k < 1u ? 1u : k;
    const double* xj = xm + j * GRA_VEGAS_BINS;
    double zo, ac;
    if (k > 1) {
      zo = xj[k - 1] - xj[k - 2];
      ac = xj[k - 2] + (zn - k) * zo;
    } else {
      zo = xj[0];
      ac = (zn - k) * zo;
    }
    x[j] = 0.0 + ac * 1.0;
    vw *= zo * bins;
    bin[j] = (uint8_t)(k - 1);
  }
  return vw;
}

// ---------------------------------------------------------------------------
// Kinematics (src/MContinuum.cc, src/MProcess.cc, include/Graniitti/MKinematics.h)
// ---------------------------------------------------------------------------
// kinematics::SolvePz3_A / _B / SolvePz, include/Graniitti/MKinematics.h:33-229.
// rs = sqrt(s) and s32 = pow(s, 1.5) are run constants.
__device__ __forceinline__ double solvepz_a(double m3, double pt3, double pt4, double pz5,
                                            double E5, double s, double rs, double s32) {
  const double a2 = E5 * E5, a3 = pt3 * pt3, a4 = pt4 * pt4, a5 = pz5 * pz5, a6 = m3 * m3;
  const double c8 = s * a2 * 6.0;
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
__device__ __forceinline__ double solvepz_b(double m3, double m4, double pt3, double pt4, double pz5,
                                            double E5, double s, double rs, double s32) {
  const double a2 = E5 * E5, a3 = m3 * m3, a4 = m4 * m4, a5 = pt3 * pt3, a6 = pt4 * pt4,
               a7 = pz5 * pz5;
  const double c9 = s * a2 * 6.0;
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
__device__ __forceinline__ double solvepz(double m3, double m4, double pt3, double pt4, double pz5,
                                          double E5, const RunParams& P) {
  if (fabs(m3 - m4) < 1e-9) return solvepz_a(m3, pt3, pt4, pz5, E5, P.s, P.rs, P.s32);
  return solvepz_b(m3, m4, pt3, pt4, pz5, E5, P.s, P.rs, P.s32);
}

// gra::math::CheckEMC, include/Graniitti/MMath.h:191-198 (NaN passes, as there)
__device__ __forceinline__ bool check_emc(dv4 d) {
  const double eps = 1e-6;
  return !(fabs(d.x) > eps || fabs(d.y) > eps || fabs(d.z) > eps || d.e > eps);
}

// Kinematic state of one point (LORENTZSCALAR subset)
struct Kin {
  dv4 pf[5];   // 0 system, 1,2 forward protons, 3,4 central
  dv4 dt[2];   // decaytree[i].p4 at tree level
  dv4 legs[2][2];
  dv4 pkt;
  dv4 q1;
  double ss13, ss24, ss14, ss23;
  double t1, t2, t_hat, u_hat, m2, Y, Pt;
  double dw;  // <F>: two-body phase-space weight of the central system (lts.DW)
};

// MProcess::GetLorentzScalars(Nf = 4), src/MProcess.cc:1499-1584. Only the
// scalars the amplitudes and cuts read are kept; every s_ij is still formed and
// checked as in the reference (lower triangle entries are >= 0 copies there).
__device__ __forceinline__ bool lorentz_scalars(const RunParams& P, Kin& K, bool cut_vars) {
#pragma unroll
  for (int i = 0; i < 5; ++i) {
#pragma unroll
    for (int j = i + 1; j < 5; ++j) {
      const double v = m2of(v4add(K.pf[i], K.pf[j]));
      if (v < 0) return false;
      if (i == 1 && j == 3) K.ss13 = v;
      if (i == 2 && j == 4) K.ss24 = v;
      if (i == 1 && j == 4) K.ss14 = v;
      if (i == 2 && j == 3) K.ss23 = v;
    }
  }
  K.q1 = v4sub(P.pbeam1, K.pf[1]);
  const dv4 q2 = v4sub(P.pbeam2, K.pf[2]);
  K.t1 = m2of(K.q1);
  K.t2 = m2of(q2);
  if (K.t1 > 0 || K.t2 > 0) return false;
  K.t_hat = m2of(v4sub(K.q1, K.dt[0]));
  K.u_hat = m2of(v4sub(K.q1, K.dt[1]));
  if (cut_vars) {
    K.m2 = m2of(K.pf[0]);
    K.Y = rapof(K.pf[0]);
    K.Pt = ptof(K.pf[0]);
  }
  return true;
}

// kinematics::LorentzBoost(mother, M0, p, +1), include/Graniitti/MKinematics.h:384-401
__device__ __forceinline__ dv4 lorentz_boost(dv4 b, double M0, dv4 p) {
  const double bx = b.x / b.e, by = b.y / b.e, bz = b.z / b.e;
  const double gamma = b.e / M0;
  const double kappa1 = bx * p.x + by * p.y + bz * p.z;
  const double kappa2 = gamma * (gamma * kappa1 / (1.0 + gamma) + p.e);
  return v4m(p.x + kappa2 * bx, p.y + kappa2 * by, p.z + kappa2 * bz, gamma * (p.e + kappa1));
}

// Phase-space class of a kernel: GRA_PS_C or GRA_PS_F fixed at compile time
// (each instantiation then holds only its own kinematics), or kPsAny, read
// from the run's parameters
constexpr int kPsAny = -1;
template <int PS>
__device__ __forceinline__ bool ps_f(const RunParams& P) {
  return PS == kPsAny ? P.phase_space == GRA_PS_F : PS == GRA_PS_F;
}

// MFactorized::B51RandomKin + B51BuildKin, src/MFactorized.cc:243-393, for a
// central system decaying to two stable particles: x[0..4] as there, x[5] the
// system's M^2 on [M_MIN^2, M_MAX^2], and the isotropic decay angles that
// TwoBodyPhaseSpace draws (include/Graniitti/MKinematics.h:405-459) lifted to
// x[6] (cos theta = -1 + 2 x[6]) and x[7] (phi = 2 pi x[7]).
__device__ __forceinline__ bool build_kin_f(const RunParams& P, const double* x, Kin& K) {
  const double u1 = P.lmin + (P.lmax - P.lmin) * x[0];
  const double u2 = P.lmin + (P.lmax - P.lmin) * x[1];
  const double pt1 = exp(u1), pt2 = exp(u2);
  const double phi1 = 2.0 * GRA_PI * x[2], phi2 = 2.0 * GRA_PI * x[3];
  const double yX = P.rap_min + (P.rap_max - P.rap_min) * x[4];
  const double m2X = P.f_Mmin * P.f_Mmin + (P.f_Mmax * P.f_Mmax - P.f_Mmin * P.f_Mmin) * x[5];
  const double m1 = P.m_beam, m2 = P.m_beam;
  dv4 p1 = v4m(pt1 * cos(phi1), pt1 * sin(phi1), 0, 0);
  dv4 p2 = v4m(pt2 * cos(phi2), pt2 * sin(phi2), 0, 0);
  dv4 pX = v4m(-(p1.x + p2.x), -(p1.y + p2.y), 0, 0);
  const double mtX = msq(m2X + pt2of(pX));
  pX.z = mtX * sinh(yX);
  pX.e = mtX * cosh(yX);
  if (pX.e > P.sqrt_s - (m1 + m2)) return false;
  const double p1z = solvepz(m1, m2, pt1, pt2, pX.z, pX.e, P);
  const double p2z = -(pX.z + p1z);
  if (p1z < 0 || p2z > 0) return false;
  p1.z = p1z;
  p1.e = msq(m1 * m1 + pt1 * pt1 + p1z * p1z);
  p2.z = p2z;
  p2.e = msq(m2 * m2 + pt2 * pt2 + p2z * p2z);
  K.pf[1] = p1;
  K.pf[2] = p2;
  K.pf[0] = pX;
  if (!check_emc(v4sub(P.beamsum, v4add(v4add(K.pf[1], K.pf[2]), K.pf[0])))) return false;
  // TwoBodyPhaseSpace(pX, sqrt(m2X), {m, m}): DecayMomentum, Isotropic, boost, dPhi2
  const double M0 = msq(m2X), ma = P.m_central, mb = P.m_central;
  const double pn = msq((M0 - ma - mb) * (M0 + ma + mb) * (M0 + ma - mb) * (M0 - ma + mb)) /
                    (2.0 * M0);
  const double ct = -1.0 + 2.0 * x[6];
  const double st = msq(1.0 - ct * ct);
  const double ph = 2.0 * GRA_PI * x[7];
  const double kx = pn * st * cos(ph), ky = pn * st * sin(ph), kz = pn * ct;
  K.dt[0] = lorentz_boost(pX, M0, v4m(kx, ky, kz, msq(ma * ma + pn * pn)));
  K.dt[1] = lorentz_boost(pX, M0, v4m(-kx, -ky, -kz, msq(mb * mb + pn * pn)));
  const double V = 2.0 * 2.0 * GRA_PI;
  K.dw = V * pn / (16.0 * (GRA_PI * GRA_PI) * M0);
  if (K.dw < 0) return false;
  K.pf[3] = K.dt[0];
  K.pf[4] = K.dt[1];
  return lorentz_scalars(P, K, true);
}

// MContinuum::BNRandomKin + BNBuildKin (Kf = 2), src/MContinuum.cc:276-435
template <int PS = kPsAny>
__device__ __forceinline__ bool build_kin(const RunParams& P, const double* x, Kin& K) {
  if (ps_f<PS>(P)) return build_kin_f(P, x, K);
  const double u1 = P.lmin + (P.lmax - P.lmin) * x[0];
  const double u2 = P.lmin + (P.lmax - P.lmin) * x[1];
  const double pt1 = exp(u1), pt2 = exp(u2);
  const double phi1 = 2.0 * GRA_PI * x[2], phi2 = 2.0 * GRA_PI * x[3];
  double moff[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (P.w_central <= 0.0) {
      moff[i] = P.m_central;
    } else {
      // RelativisticBWRandom (src/MRandom.cc:82-104) with the draw lifted to x[8 + i]
      const double m0 = P.m_central, G = P.w_central;
      const double m2 = m0 * m0;
      const double m2val = m2 + m0 * G * tan(P.bw_a1 + x[8 + i] * (P.bw_a2 - P.bw_a1));
      moff[i] = fmax(P.bw_dmin, msq(m2val));
    }
  }
  const double kt = P.kt_min + (P.kt_max - P.kt_min) * x[4];
  const double phi = 2.0 * GRA_PI * x[5];
  const double y0 = P.rap_min + (P.rap_max - P.rap_min) * x[6];
  const double y1 = P.rap_min + (P.rap_max - P.rap_min) * x[7];
  const double m1 = P.m_beam, m2 = P.m_beam;
  dv4 p1 = v4m(pt1 * cos(phi1), pt1 * sin(phi1), 0, 0);
  dv4 p2 = v4m(pt2 * cos(phi2), pt2 * sin(phi2), 0, 0);
  K.pkt = v4m(kt * cos(phi), kt * sin(phi), 0, 0);
  // BLinearSystem, Kf = 2: A = diag(1/2, 1/2)
  const dv4 w = v4add(p1, p2);
  const dv4 b0 = v4sub(K.pkt, w);
  const dv4 b1 = v4sub(v4scale(K.pkt, -1.0), w);
  dv4 p[2];
  p[0] = v4add(v4add(v4m(0, 0, 0, 0), v4scale(b0, 0.5)), v4scale(b1, 0.0));
  p[1] = v4add(v4add(v4m(0, 0, 0, 0), v4scale(b0, 0.0)), v4scale(b1, 0.5));
  const double yy[2] = {y0, y1};
  dv4 sumP = v4m(0, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double mt = msq(moff[i] * moff[i] + pt2of(p[i]));
    p[i].z = mt * sinh(yy[i]);
    p[i].e = mt * cosh(yy[i]);
    sumP = v4add(sumP, p[i]);
  }
  if (sumP.e > P.sqrt_s) return false;
  const double p1z = solvepz(m1, m2, pt1, pt2, sumP.z, sumP.e, P);
  const double p2z = -(sumP.z + p1z);
  if (p1z < 0 || p2z > 0) return false;
  p1.z = p1z;
  p1.e = msq(m1 * m1 + pt1 * pt1 + p1z * p1z);
  p2.z = p2z;
  p2.e = msq(m2 * m2 + pt2 * pt2 + p2z * p2z);
  if (P.asym) {
    // asymmetric beams: the centre-of-mass momenta boosted to the lab
    // (src/MContinuum.cc:386-395, kinematics::LorentzBoost with sign +1)
    p1 = lorentz_boost(P.beamsum, P.sqrt_s, p1);
    p2 = lorentz_boost(P.beamsum, P.sqrt_s, p2);
    sumP = lorentz_boost(P.beamsum, P.sqrt_s, sumP);
#pragma unroll
    for (int i = 0; i < 2; ++i) p[i] = lorentz_boost(P.beamsum, P.sqrt_s, p[i]);
  }
  K.pf[1] = p1;
  K.pf[2] = p2;
  K.pf[0] = sumP;
  double sumM = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    K.dt[i] = p[i];
    K.pf[3 + i] = p[i];
    sumM += mof(p[i]);
  }
  if (mof(sumP) < sumM) return false;
  if (!check_emc(v4sub(P.beamsum, v4add(v4add(K.pf[1], K.pf[2]), K.pf[0])))) return false;
  if (P.m_daughter > 0.0) {
    // ConstructDecayKinematics + TwoBodyPhaseSpace, src/MProcess.cc:1231-1283,
    // include/Graniitti/MKinematics.h:405-459; decay angles lifted to x[10..13]
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double M0 = mof(K.dt[i]);
      const double md = P.m_daughter;
      if (M0 < 2.0 * md) return false;
      const double pn = msq((M0 * M0 - (md + md) * (md + md)) * (M0 * M0 - (md - md) * (md - md))) /
                        (2.0 * M0);
      const double ct = 2.0 * x[10 + 2 * i] - 1.0;
      const double ph = 2.0 * GRA_PI * x[11 + 2 * i];
      const double st = msq(1.0 - ct * ct);
      const double px = pn * st * cos(ph), py = pn * st * sin(ph), pz = pn * ct;
      const double ed = msq(md * md + pn * pn);
      K.legs[i][0] = lorentz_boost(K.dt[i], M0, v4m(px, py, pz, ed));
      K.legs[i][1] = lorentz_boost(K.dt[i], M0, v4m(-px, -py, -pz, ed));
    }
  }
  return lorentz_scalars(P, K, true);
}

// MProcess::FindDecayCuts leaf test, src/MProcess.cc:1138-1154
__device__ __forceinline__ bool leg_cuts(const RunParams& P, dv4 p) {
  const double pt = ptof(p), et = mtof(p), eta = etaof(p), rap = rapof(p);
  return pt >= P.fid_pt_min && pt <= P.fid_pt_max && et >= P.fid_et_min && et <= P.fid_et_max &&
         eta >= P.fid_eta_min && eta <= P.fid_eta_max && rap >= P.fid_rap_min &&
         rap <= P.fid_rap_max;
}

// MProcess::VetoCuts with MProcess::FindVetoCuts, src/MProcess.cc:1069-1082,
// 1157-1175: false when a final-state central particle lies in a veto domain
// (pt and eta both inside). The forward protons are not tested: the reference
// tests its forward decay branches, which are empty for elastic protons.
__device__ __forceinline__ bool veto_hit(const RunParams& P, dv4 p) {
  const double pt = ptof(p), eta = etaof(p);
  bool hit = false;
  for (int d = 0; d < P.n_veto; ++d)
    if (pt >= P.veto_pt_min[d] && pt <= P.veto_pt_max[d] && eta >= P.veto_eta_min[d] &&
        eta <= P.veto_eta_max[d])
      hit = true;
  return hit;
}
__device__ __forceinline__ bool veto_cuts(const RunParams& P, const Kin& K) {
  if (!P.veto_active) return true;
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (P.m_daughter > 0.0) {
      if (veto_hit(P, K.legs[i][0]) || veto_hit(P, K.legs[i][1])) ok = false;
    } else {
      if (veto_hit(P, K.dt[i])) ok = false;
    }
  }
  return ok;
}

// MProcess::CommonCuts, src/MProcess.cc:1085-1135
__device__ __forceinline__ bool fiducial_cuts(const RunParams& P, const Kin& K) {
  if (!P.fid_active) return true;
  if (!(fabs(K.t1) >= P.fwd_t_min && fabs(K.t1) <= P.fwd_t_max && fabs(K.t2) >= P.fwd_t_min &&
        fabs(K.t2) <= P.fwd_t_max))
    return false;
  const double M = msq(K.m2);
  if (!(M >= P.sys_M_min && M <= P.sys_M_max && K.Y >= P.sys_Y_min && K.Y <= P.sys_Y_max &&
        K.Pt >= P.sys_Pt_min && K.Pt <= P.sys_Pt_max))
    return false;
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (P.m_daughter > 0.0) {
      if (!leg_cuts(P, K.legs[i][0])) ok = false;
      if (!leg_cuts(P, K.legs[i][1])) ok = false;
    } else {
      if (!leg_cuts(P, K.dt[i])) ok = false;
    }
  }
  return ok;
}

// Per-point constants of the screening loop (pfinal_orig-derived)
struct LoopPt {
  double p1x, p1y, p2x, p2y;  // tree proton pT vectors
  double pktx, pkty;
  double sh[2], ch[2], mc2[2];  // sinh y, cosh y, m^2 of the central particles (<F>: [0]
                                // holds the central system's)
  double m1, m2;                // proton masses from pfinal_orig[].M()
  dv4 dt[2];
  double dw;  // <F>: two-body phase-space weight (unchanged by the loop)
};

template <int PS = kPsAny>
__device__ __forceinline__ void loop_point(const RunParams& P, const Kin& K, LoopPt& L) {
  L.p1x = K.pf[1].x;
  L.p1y = K.pf[1].y;
  L.p2x = K.pf[2].x;
  L.p2y = K.pf[2].y;
  L.pktx = K.pkt.x;
  L.pkty = K.pkt.y;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double y = rapof(K.pf[3 + i]);
    const double m = mof(K.pf[3 + i]);
    L.sh[i] = sinh(y);
    L.ch[i] = cosh(y);
    L.mc2[i] = m * m;
    L.dt[i] = K.dt[i];
  }
  L.m1 = mof(K.pf[1]);
  L.m2 = mof(K.pf[2]);
  L.dw = K.dw;
  if (ps_f<PS>(P)) {
    // MFactorized::LoopKinematics keeps the system's mass and rapidity
    const double mX = mof(K.pf[0]), yX = rapof(K.pf[0]);
    L.mc2[0] = mX * mX;
    L.sh[0] = sinh(yX);
    L.ch[0] = cosh(yX);
  }
}

// MFactorized::LoopKinematics, src/MFactorized.cc:77-139: the system keeps its
// mass and rapidity, takes pT = -(p1T + p2T), and the decay products (pfinal[3],
// pfinal[4] and the decay tree) stay those of the tree point.
__device__ __forceinline__ bool loop_kinematics_f(const RunParams& P, const LoopPt& L, double kx,
                                                  double ky, Kin& K, bool cut_vars) {
  const double p1x = L.p1x - kx, p1y = L.p1y - ky;
  const double p2x = L.p2x + kx, p2y = L.p2y + ky;
  const double pt1 = msq(p1x * p1x + p1y * p1y), pt2 = msq(p2x * p2x + p2y * p2y);
  dv4 pX = v4m(-(p1x + p2x), -(p1y + p2y), 0, 0);
  const double mtX = msq(L.mc2[0] + pt2of(pX));
  pX.z = mtX * L.sh[0];
  pX.e = mtX * L.ch[0];
  if (pX.e > P.sqrt_s - (L.m1 + L.m2)) return false;
  const double p1z = solvepz(L.m1, L.m2, pt1, pt2, pX.z, pX.e, P);
  const double p2z = -(pX.z + p1z);
  if (isnan(p1z)) return false;
  if (p1z < 0 || p2z > 0) return false;
  K.pf[1] = v4m(p1x, p1y, p1z, msq(L.m1 * L.m1 + pt1 * pt1 + p1z * p1z));
  K.pf[2] = v4m(p2x, p2y, p2z, msq(L.m2 * L.m2 + pt2 * pt2 + p2z * p2z));
  K.pf[0] = pX;
  if (!check_emc(v4sub(P.beamsum, v4add(v4add(K.pf[1], K.pf[2]), K.pf[0])))) return false;
  K.dt[0] = L.dt[0];
  K.dt[1] = L.dt[1];
  K.pf[3] = L.dt[0];
  K.pf[4] = L.dt[1];
  K.dw = L.dw;
  return lorentz_scalars(P, K, cut_vars);
}

// MContinuum::LoopKinematics (Kf = 2), src/MContinuum.cc:94-161, for the node
// with loop momentum (kx, ky). pf[1], pf[2] carry E and z of the previous node
// into BLinearSystem only through their x, y components, which are set here.
template <int PS = kPsAny>
__device__ __forceinline__ bool loop_kinematics(const RunParams& P, const LoopPt& L, double kx,
                                                double ky, Kin& K, bool cut_vars) {
  if (ps_f<PS>(P)) return loop_kinematics_f(P, L, kx, ky, K, cut_vars);
  const double p1x = L.p1x - kx, p1y = L.p1y - ky;
  const double p2x = L.p2x + kx, p2y = L.p2y + ky;
  // BLinearSystem: b0 = q - w, b1 = -q - w, p_i = sum_j A_ij b_j
  const double wx = p1x + p2x, wy = p1y + p2y;
  const double b0x = L.pktx - wx, b0y = L.pkty - wy;
  const double b1x = L.pktx * -1.0 - wx, b1y = L.pkty * -1.0 - wy;
  const double cx[2] = {(0.0 + b0x * 0.5) + b1x * 0.0, (0.0 + b0x * 0.0) + b1x * 0.5};
  const double cy[2] = {(0.0 + b0y * 0.5) + b1y * 0.0, (0.0 + b0y * 0.0) + b1y * 0.5};
  dv4 sumP = v4m(0, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double mt = msq(L.mc2[i] + (cx[i] * cx[i] + cy[i] * cy[i]));
    K.pf[3 + i] = v4m(cx[i], cy[i], mt * L.sh[i], mt * L.ch[i]);
    sumP = v4add(sumP, K.pf[3 + i]);
  }
  const double pt1 = msq(p1x * p1x + p1y * p1y), pt2 = msq(p2x * p2x + p2y * p2y);
  const double p1z = solvepz(L.m1, L.m2, pt1, pt2, sumP.z, sumP.e, P);
  const double p2z = -(sumP.z + p1z);
  if (p1z < 0 || p2z > 0) return false;
  K.pf[1] = v4m(p1x, p1y, p1z, msq(L.m1 * L.m1 + pt1 * pt1 + p1z * p1z));
  K.pf[2] = v4m(p2x, p2y, p2z, msq(L.m2 * L.m2 + pt2 * pt2 + p2z * p2z));
  K.pf[0] = sumP;
  if (!check_emc(v4sub(P.beamsum, v4add(v4add(K.pf[1], K.pf[2]), K.pf[0])))) return false;
  K.dt[0] = L.dt[0];
  K.dt[1] = L.dt[1];
  return lorentz_scalars(P, K, cut_vars);
}

// ---------------------------------------------------------------------------
// Regge continuum amplitude, MRegge::ME4 (src/MRegge.cc:394-438)
// ---------------------------------------------------------------------------
// form::S3F, src/MForm.cc:433-435
__device__ __forceinline__ double s3f(const RunParams& P, double t) {
  return (1.0 / (1 - t / P.fc1)) * (1.0 / (1 - t / P.fc2));
}
// PARAM_REGGE::Meson_FF, src/MRegge.cc:1155-1174
__device__ __forceinline__ double meson_ff(const RunParams& P, double t_hat) {
  const double tprime = t_hat - P.m_prop2;
  if (P.offshellFF == 0) return exp(-P.b_EXP * fabs(tprime));
  if (P.offshellFF == 1)
    return exp(-P.b_OREAR * msq(fabs(tprime) + P.a_OREAR * P.a_OREAR) + P.a_OREAR * P.b_OREAR);
  return 1.0 / (1.0 - tprime / P.b_POW);
}
// MRegge::PropOnly with form::ReggeEtaLinear, src/MRegge.cc:888-906, src/MForm.cc:394-396.
// ReggeEta(alpha(0), sigma) is a run constant (P.eta0).
__device__ __forceinline__ cxd regge_prop(const RunParams& P, double s, double t) {
  cxd M = cd(0, 0);
  for (int k = 0; k < 3; ++k) {
    if (P.regge_on[k]) {
      const double alpha = P.regge_a0[k] + P.regge_ap[k] * t;
      const double th = ((-GRA_PI / 2.0) * P.regge_ap[k]) * t;
      const cxd ph = cd(cos(th), sin(th));
      const cxd eta = cmul(cd(P.eta0_re[k], P.eta0_im[k]), ph);
      const double pw = pow(s / P.regge_s0, alpha);
      M = cadd(M, cd(eta.re * pw, eta.im * pw));
    }
  }
  return M;
}
// Full-precision amplitude in the reference's operation order
__device__ __forceinline__ cxd regge_me4_fp64(const RunParams& P, const Kin& K) {
  const double FF_A = P.gN_P * s3f(P, K.t1);
  const double FF_B = P.gN_P * s3f(P, K.t2);
  const double g = P.gpp_P;
  const double fft = meson_ff(P, K.t_hat), ffu = meson_ff(P, K.u_hat);
  const double prt = 1.0 / (K.t_hat - P.m_prop2), pru = 1.0 / (K.u_hat - P.m_prop2);
  cxd At = regge_prop(P, K.ss13, K.t1);
  At = cscale(At, FF_A);
  At = cscale(At, fft);
  At = cscale(At, g);
  At = cscale(At, prt);
  At = cscale(At, fft);
  At = cscale(At, g);
  At = cmul(At, regge_prop(P, K.ss24, K.t2));
  At = cscale(At, FF_B);
  cxd Au = cscale(regge_prop(P, K.ss14, K.t1), 1.0);
  Au = cscale(Au, FF_A);
  Au = cscale(Au, ffu);
  Au = cscale(Au, g);
  Au = cscale(Au, pru);
  Au = cscale(Au, ffu);
  Au = cscale(Au, g);
  Au = cmul(Au, regge_prop(P, K.ss23, K.t2));
  Au = cscale(Au, FF_B);
  return cadd(At, Au);
}

// Reduced FP32 amplitude with a single active trajectory (TUNE0 for pi, K):
//   A = eta0^2 e^{i(th1+th2)} [R_t e^{a1 L13 + a2 L24 - E0} + R_u e^{a1 L14 + a2 L23 - E0}] e^{E0}
// where E0 is a per-point exponent offset that keeps the FP32 values near 1;
// the common factor e^{E0} eta0^2 is applied by the caller in FP64.
__device__ __forceinline__ cxf regge_me4_fp32(const RunParams& P, const Kin& K, float E0) {
  const float t1 = (float)K.t1, t2 = (float)K.t2;
  const float th = (float)P.ph_c * (t1 + t2);
  const float a1 = (float)P.regge_a0[P.regge_k] + (float)P.regge_ap[P.regge_k] * t1;
  const float a2 = (float)P.regge_a0[P.regge_k] + (float)P.regge_ap[P.regge_k] * t2;
  const float inv_s0 = (float)(1.0 / P.regge_s0);
  const float L13 = __logf((float)K.ss13 * inv_s0), L24 = __logf((float)K.ss24 * inv_s0);
  const float L14 = __logf((float)K.ss14 * inv_s0), L23 = __logf((float)K.ss23 * inv_s0);
  const float fc1 = (float)P.fc1, fc2 = (float)P.fc2;
  const float FA = 1.0f / ((1.0f - t1 / fc1) * (1.0f - t1 / fc2));
  const float FB = 1.0f / ((1.0f - t2 / fc1) * (1.0f - t2 / fc2));
  const float m2p = (float)P.m_prop2;
  const float th_ = (float)K.t_hat - m2p, tu_ = (float)K.u_hat - m2p;
  float fft, ffu;
  if (P.offshellFF == 2) {
    const float bp = (float)P.b_POW;
    fft = 1.0f / (1.0f - th_ / bp);
    ffu = 1.0f / (1.0f - tu_ / bp);
  } else if (P.offshellFF == 0) {
    fft = __expf(-(float)P.b_EXP * fabsf(th_));
    ffu = __expf(-(float)P.b_EXP * fabsf(tu_));
  } else {
    const float ao = (float)P.a_OREAR, bo = (float)P.b_OREAR;
    fft = __expf(-bo * sqrtf(fabsf(th_) + ao * ao) + ao * bo);
    ffu = __expf(-bo * sqrtf(fabsf(tu_) + ao * ao) + ao * bo);
  }
  const float Rt = fft * fft / th_, Ru = ffu * ffu / tu_;
  const float et = __expf(a1 * L13 + a2 * L24 - E0);
  const float eu = __expf(a1 * L14 + a2 * L23 - E0);
  const float mag = FA * FB * (Rt * et + Ru * eu);
  float sn, cs;
  __sincosf(th, &sn, &cs);
  return cf(mag * cs, mag * sn);
}

// Exponent offset of a point: a1 L13 + a2 L24 at tree level (FP32)
__device__ __forceinline__ float regge_e0(const RunParams& P, const Kin& K) {
  const float t1 = (float)K.t1, t2 = (float)K.t2;
  const float a1 = (float)P.regge_a0[P.regge_k] + (float)P.regge_ap[P.regge_k] * t1;
  const float a2 = (float)P.regge_a0[P.regge_k] + (float)P.regge_ap[P.regge_k] * t2;
  const float inv_s0 = (float)(1.0 / P.regge_s0);
  return a1 * __logf((float)K.ss13 * inv_s0) + a2 * __logf((float)K.ss24 * inv_s0);
}

// ---------------------------------------------------------------------------
// Tensor-pomeron continuum, MTensorPomeron::ME4 for a vector-meson pair (2xV),
// src/MTensorPomeron.cc:572-898 (2xV 808-888).
//
// The reference contracts rank-4 complex tensors (Pomeron propagators iD_P,
// vertices iG_PppHE and iG_Pvv = i FM (2 g1 Gamma0 - g2 Gamma2)) component by
// component. Contracting the proton vertex c P_mu P_nu with the propagator
// F (g g + g g - g g / 2) leaves V^{ab} = 2 c F (P^a P^b - P^2 g^{ab} / 4), and
// Gamma0, Gamma2 are traceless in their Pomeron index pair, so every
// (proton vertex) x (propagator) x (meson vertex) block is the 4 x 4 matrix
//   X_{mu nu} = i FM 2 c F [Cg g_{mu nu} + C21 k2_mu k1_nu + CPP P_mu P_nu
//                           + CP1 P_mu k1_nu + C2P k2_mu P_nu]
// with (k1, k2) the meson-vertex momenta, and the helicity amplitudes only need
// these matrices contracted with one polarisation vector each: a handful of
// Minkowski products instead of ~2e4 complex multiply-adds per point.
// ---------------------------------------------------------------------------
struct cv4 {
  cxd t, x, y, z;
};
__device__ __forceinline__ double mdot(dv4 a, dv4 b) {
  return a.e * b.e - a.x * b.x - a.y * b.y - a.z * b.z;
}
__device__ __forceinline__ cxd mdotc(dv4 a, const cv4& e) {
  return cd(a.e * e.t.re - a.x * e.x.re - a.y * e.y.re - a.z * e.z.re,
            a.e * e.t.im - a.x * e.x.im - a.y * e.y.im - a.z * e.z.im);
}
__device__ __forceinline__ cxd mdotcc(const cv4& a, const cv4& b) {
  cxd s = cmul(a.t, b.t);
  const cxd sx = cmul(a.x, b.x), sy = cmul(a.y, b.y), sz = cmul(a.z, b.z);
  return cd(s.re - sx.re - sy.re - sz.re, s.im - sx.im - sy.im - sz.im);
}

// form::F1 with the Kelly G_E, G_M fits (src/MForm.cc:542-548, 611-645)
__device__ __forceinline__ double kelly_f1(double Q2, double mp) {
  Q2 = fabs(Q2);
  const double tau = Q2 / ((2 * mp) * (2 * mp));
  const double GE = (1.0 - 0.24 * tau) / (1.0 + tau * (10.98 + tau * (12.82 + tau * 21.97)));
  const double GM = 2.792847337 * (1.0 + 0.12 * tau) /
                    (1.0 + tau * (10.97 + tau * (18.86 + tau * 6.55)));
  return (GE + tau * GM) / (tau + 1.0);
}

struct TBlk {
  cxd pre;
  double Cg, C21, CPP, CP1, C2P;
  dv4 k1, k2, P;
};
// iG_PppHE(prime, pa) . iD_P(s, t) . iG_Pvv(k1, -k2) reduced to the dyadic form;
// c = -3 i gPNN F1(t), F = (-i s a')^{alpha(t) - 1} / (4 s), FM = 1 / (1 - (k1 + k2)^2 / 0.5)
__device__ __forceinline__ TBlk tensor_block(const RunParams& P, dv4 Pv, double F1, double s,
                                             double t, dv4 k1, dv4 k2) {
  TBlk b;
  const double g1 = P.tp_gPvv0, g2 = P.tp_gPvv1;
  const double k1k2 = mdot(k1, k2), k1P = mdot(k1, Pv), k2P = mdot(k2, Pv), P2 = mdot(Pv, Pv);
  const double S0 = 2.0 * k1P * k2P - 0.5 * k1k2 * P2;
  b.Cg = 2.0 * g1 * S0 * k1k2 - 2.0 * g2 * k1P * k2P + g2 * P2 * k1k2;
  b.C21 = -2.0 * g1 * S0 - g2 * P2;
  b.CPP = -2.0 * g2 * k1k2;
  b.CP1 = 2.0 * g2 * k2P;
  b.C2P = 2.0 * g2 * k1P;
  const dv4 q = v4add(k1, k2);
  const double FM = 1.0 / (1.0 - mdot(q, q) / 0.5);
  // 2 c F with c = -3 i gPNN F1: (-i s a')^{a-1} = exp((a - 1)(log(s a') - i pi / 2))
  const double am1 = P.tp_delta_P + P.tp_ap_P * t;
  const double mag = exp(am1 * log(s * P.tp_ap_P)) / (4.0 * s);
  const double ph = -0.5 * GRA_PI * am1;
  double sn, cs;
  sincos(ph, &sn, &cs);
  // i FM 2 (-3 i gPNN F1) F = 6 gPNN F1 FM F
  const double r = 6.0 * P.tp_gPNN * F1 * FM * mag;
  b.pre = cd(r * cs, r * sn);
  b.k1 = k1;
  b.k2 = k2;
  b.P = Pv;
  return b;
}
__device__ __forceinline__ cv4 cv4_comb(const TBlk& b, const cv4& e, cxd ak1, cxd ak2, cxd aP) {
  cv4 r;
  r.t = cadd(cadd(cscale(e.t, b.Cg), cscale(ak1, b.k1.e)), cadd(cscale(ak2, b.k2.e), cscale(aP, b.P.e)));
  r.x = cadd(cadd(cscale(e.x, b.Cg), cscale(ak1, b.k1.x)), cadd(cscale(ak2, b.k2.x), cscale(aP, b.P.x)));
  r.y = cadd(cadd(cscale(e.y, b.Cg), cscale(ak1, b.k1.y)), cadd(cscale(ak2, b.k2.y), cscale(aP, b.P.y)));
  r.z = cadd(cadd(cscale(e.z, b.Cg), cscale(ak1, b.k1.z)), cadd(cscale(ak2, b.k2.z), cscale(aP, b.P.z)));
  return r;
}
// sum_nu X_{mu nu} e^nu (without the block prefactor), as a contravariant vector
__device__ __forceinline__ cv4 tensor_right(const TBlk& b, const cv4& e) {
  const cxd k1e = mdotc(b.k1, e), Pe = mdotc(b.P, e);
  const cxd ak2 = cadd(cscale(k1e, b.C21), cscale(Pe, b.C2P));
  const cxd aP = cadd(cscale(Pe, b.CPP), cscale(k1e, b.CP1));
  return cv4_comb(b, e, cd(0, 0), ak2, aP);
}
// sum_mu e^mu X_{mu nu}
__device__ __forceinline__ cv4 tensor_left(const TBlk& b, const cv4& e) {
  const cxd k2e = mdotc(b.k2, e), Pe = mdotc(b.P, e);
  const cxd ak1 = cadd(cscale(k2e, b.C21), cscale(Pe, b.CP1));
  const cxd aP = cadd(cscale(Pe, b.CPP), cscale(k2e, b.C2P));
  return cv4_comb(b, e, ak1, cd(0, 0), aP);
}
// Conjugated massive spin-1 polarisation vectors for helicities -1, 0, +1
// (MDirac::MassiveSpin1States "conj", index up; src/MDirac.cc:393-456, 661-679),
// with the angles' cosines and sines formed from the momentum components
__device__ __forceinline__ void spin1_conj(dv4 k, cv4 e[3]) {
  const double pt = sqrt(k.x * k.x + k.y * k.y), pm = sqrt(pt * pt + k.z * k.z);
  const double ct = pm > 0.0 ? k.z / pm : 1.0, st = pm > 0.0 ? pt / pm : 0.0;
  const double cp = pt > 0.0 ? k.x / pt : 1.0, sp = pt > 0.0 ? k.y / pt : 0.0;
  const double N = 0.70710678118654752440;
#pragma unroll
  for (int h = 0; h < 3; h += 2) {
    const double m = (double)(h - 1);
    e[h].t = cd(0.0, 0.0);
    e[h].x = cd(-m * ct * cp * N, -sp * N);
    e[h].y = cd(-m * ct * sp * N, cp * N);
    e[h].z = cd(m * st * N, 0.0);
  }
  const double M = sqrt(mdot(k, k));
  e[1].t = cd(pm / M, 0.0);
  e[1].x = cd(k.e * st * cp / M, 0.0);
  e[1].y = cd(k.e * st * sp / M, 0.0);
  e[1].z = cd(k.e * ct / M, 0.0);
}
// Vector-meson propagator factor of iD_V (src/MTensorPomeron.cc:2353-2369) times
// the two off-shell form factors: d = -R ff^2 / (m2 - M^2), R = (e^{i phi} s34 / s_th)^{alpha - 1}
__device__ __forceinline__ cxd tensor_vprop(const RunParams& P, double m2, double ls, double phi) {
  const double M2 = P.m_prop2;
  const double am1 = 0.5 + 0.9 * m2 - 1.0;
  const double mag = exp(am1 * ls);
  double sn, cs;
  sincos(am1 * phi, &sn, &cs);
  const double ff = meson_ff(P, m2);
  const double r = -mag * ff * ff / (m2 - M2);
  return cd(r * cs, r * sn);
}

// The 9 helicity amplitudes h[3 h3 + h4] (h3, h4 = -1, 0, +1) at kinematics K.
// The meson momenta are decaytree[].p4 (tree level even inside the screening
// loop, as in the reference), the protons and s_ij those of K.
__device__ __forceinline__ void tensor_me4_fp64(const RunParams& P, const Kin& K, cxd h[9]) {
  const dv4 pa = P.pbeam1, pb = P.pbeam2;
  const dv4 p1 = K.pf[1], p2 = K.pf[2], p3 = K.dt[0], p4 = K.dt[1];
  const dv4 P1 = v4add(p1, pa), P2 = v4add(p2, pb);
  const dv4 pt = v4sub(v4sub(pa, p1), p3);
  const dv4 pu = v4add(v4sub(p4, pa), p1);
  const double F1a = kelly_f1(K.t1, P.tp_mp_ff), F1b = kelly_f1(K.t2, P.tp_mp_ff);
  const TBlk tA = tensor_block(P, P1, F1a, K.ss13, K.t1, pt, p3);
  const TBlk tB = tensor_block(P, P2, F1b, K.ss24, K.t2, p4, v4scale(pt, -1.0));
  const TBlk uA = tensor_block(P, P1, F1a, K.ss14, K.t1, p4, v4scale(pu, -1.0));
  const TBlk uB = tensor_block(P, P2, F1b, K.ss23, K.t2, pu, p3);
  const double s34 = mdot(K.pf[0], K.pf[0]);
  const double sth = 4.0 * P.m_prop2;
  const double phi = 0.5 * GRA_PI * exp((sth - s34) / sth) - 0.5 * GRA_PI;
  const double ls = log(s34 / sth);
  const cxd dt = cmul(tensor_vprop(P, mdot(pt, pt), ls, phi), cmul(tA.pre, tB.pre));
  const cxd du = cmul(tensor_vprop(P, mdot(pu, pu), ls, phi), cmul(uA.pre, uB.pre));
  cv4 e3[3], e4[3];
  spin1_conj(p3, e3);
  spin1_conj(p4, e4);
  cv4 a3[3], b3[3], a4[3], b4[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    a3[i] = tensor_right(tA, e3[i]);  // t-channel upper block . eps3
    b3[i] = tensor_right(uB, e3[i]);  // u-channel lower block . eps3
    a4[i] = tensor_left(tB, e4[i]);   // eps4 . t-channel lower block
    b4[i] = tensor_left(uA, e4[i]);   // eps4 . u-channel upper block
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const cxd m = cadd(cmul(dt, mdotcc(a3[i], a4[j])), cmul(du, mdotcc(b4[j], b3[i])));
      h[3 * i + j] = cd(m.im, -m.re);  // -i (M_t + M_u)
    }
}

// ---------------------------------------------------------------------------
// Tensor-pomeron continuum for a pseudoscalar pair (2xS), src/MTensorPomeron.cc:726-763.
//
// Each (proton vertex) x (Pomeron propagator) x (meson vertex) chain of the
// reference's tensor network is a scalar: with the proton vertex
// -3 i gPNN F1(t) P_mu P_nu, the propagator F (g g + g g - g g / 2) and the meson
// vertex iG_Ppsps = -2 i gP FM (K_a K_b - g_ab K^2 / 4) (K = k' + k), the index
// sums give
//   A = -6 gPNN gP F1 FM F [2 (P.K)^2 - P^2 K^2 / 2],
// since the propagator's trace part meets the traceless meson vertex. The t- and
// u-channel amplitudes are M_t = A_t B_t i / (t^ - M^2) ff(t^)^2 and likewise
// with u^; the four proton-helicity replicas are identical (high-energy vertex),
// so one amplitude -i (M_t + M_u) carries |M|^2.
// ---------------------------------------------------------------------------
// 2 c F d [2 (P.K)^2 - P^2 K^2 / 2] without the couplings: F1 FM F times the
// bracket, with F = (-i s a')^{alpha(t) - 1} / (4 s) and FM = 1 / (1 - q^2 / 0.5)
__device__ __forceinline__ cxd tensor_chain_2xs(const RunParams& P, dv4 Pv, dv4 K, double q2,
                                                double F1, double s, double t) {
  const double PK = mdot(Pv, K);
  const double br = 2.0 * (PK * PK) - 0.5 * mdot(Pv, Pv) * mdot(K, K);
  const double FM = 1.0 / (1.0 - q2 / 0.5);
  const double am1 = P.tp_delta_P + P.tp_ap_P * t;
  const double mag = exp(am1 * log(s * P.tp_ap_P)) / (4.0 * s);
  double sn, cs;
  sincos(-0.5 * GRA_PI * am1, &sn, &cs);
  const double r = F1 * FM * mag * br;
  return cd(r * cs, r * sn);
}
__device__ __forceinline__ cxd tensor_me4_2xs_fp64(const RunParams& P, const Kin& K) {
  const dv4 pa = P.pbeam1, pb = P.pbeam2;
  const dv4 p1 = K.pf[1], p2 = K.pf[2], p3 = K.dt[0], p4 = K.dt[1];
  const dv4 P1 = v4add(p1, pa), P2 = v4add(p2, pb);
  const dv4 pt = v4sub(v4sub(pa, p1), p3);
  const dv4 pu = v4add(v4sub(p4, pa), p1);
  const dv4 mp3 = v4scale(p3, -1.0);
  const double F1a = kelly_f1(K.t1, P.tp_mp_ff), F1b = kelly_f1(K.t2, P.tp_mp_ff);
  // meson vertices iG_Ppsps(k', k): K = k' + k, FM of (k' - k)^2
  const cxd At = tensor_chain_2xs(P, P1, v4add(pt, mp3), mdot(v4sub(pt, mp3), v4sub(pt, mp3)),
                                  F1a, K.ss13, K.t1);
  const cxd Bt = tensor_chain_2xs(P, P2, v4add(p4, pt), mdot(v4sub(p4, pt), v4sub(p4, pt)), F1b,
                                  K.ss24, K.t2);
  const cxd Au = tensor_chain_2xs(P, P1, v4add(p4, pu), mdot(v4sub(p4, pu), v4sub(p4, pu)), F1a,
                                  K.ss14, K.t1);
  const cxd Bu = tensor_chain_2xs(P, P2, v4add(pu, mp3), mdot(v4sub(pu, mp3), v4sub(pu, mp3)),
                                  F1b, K.ss23, K.t2);
  const double pt2 = mdot(pt, pt), pu2 = mdot(pu, pu);
  const double fft = meson_ff(P, pt2), ffu = meson_ff(P, pu2);
  // couplings (-6 gPNN gP)^2, the propagators i / (m^2 - M^2) and the overall -i
  const double g = 6.0 * P.tp_gPNN * P.tp_gP;
  const double ct = (g * g) * (fft * fft) / (pt2 - P.m_prop2);
  const double cu = (g * g) * (ffu * ffu) / (pu2 - P.m_prop2);
  return cadd(cscale(cmul(At, Bt), ct), cscale(cmul(Au, Bu), cu));
}

// |z|^2 as gra::math::abs2 (hypot then square)
__device__ __forceinline__ double abs2d(cxd z) {
  const double h = hypot(z.re, z.im);
  return h * h;
}

// Event weight from the screened |M|^2 (MContinuum::EventWeight,
// src/MContinuum.cc:164-193; BNPhaseSpaceWeight 639-656; BNIntegralVolume 619-636;
// ForwardVolume src/MProcess.cc:1390-1415; MollerFlux = 2s).
template <int PS = kPsAny>
__device__ __forceinline__ double weight_from_amp2(const RunParams& P, const Kin& K, double amp2) {
  if (ps_f<PS>(P)) {
    // MFactorized::EventWeight, src/MFactorized.cc:142-182: C_space = DW / (2 pi)
    // times CascadePS (1: stable decay products); B51PhaseSpaceWeight 441-452;
    // B51IntegralVolume 434-438 with ForwardVolume
    const double C_space = (K.dw / (2 * GRA_PI)) * 1.0;
    const double J = 1.0 / fabs(K.pf[1].z / K.pf[1].e - K.pf[2].z / K.pf[2].e);
    const double pt1 = ptof(K.pf[1]), pt2 = ptof(K.pf[2]);
    const double psw = P.f_psw * (pt1 / (2.0 * K.pf[1].e)) * (pt2 / (2.0 * K.pf[2].e)) * J;
    const double PT_vol = (pt1 * pt2) * P.lspan2;
    const double vol = P.f_vol * (P.PHI_vol * PT_vol);
    return C_space * P.invS * psw * vol * amp2 * P.GeV2barn / P.flux;
  }
  const double J = 1.0 / fabs(K.pf[1].z / K.pf[1].e - K.pf[2].z / K.pf[2].e);
  const double PROD = 1.0 * ptof(K.pkt);
  const double pt1 = ptof(K.pf[1]), pt2 = ptof(K.pf[2]);
  const double psw = P.psw_c1 * P.psw_c2 * (pt1 / (2.0 * K.pf[1].e)) * (pt2 / (2.0 * K.pf[2].e)) *
                     J * PROD;
  const double PT_vol = (pt1 * pt2) * P.lspan2;
  const double vol = P.KT_vol * (P.PHI_vol * PT_vol);
  return 1.0 * P.invS * psw * vol * amp2 * P.GeV2barn / P.flux;
}

}  // namespace
