// Run constants shared by host and device code. Filled once per handle on the
// host (so every transcendental constant comes from the host libm, exactly as
// the reference computes it) and passed to kernels by value.
#pragma once

#include <stdint.h>

#include "gra_hip.h"

#define GRA_PI 3.141592653589793238462643383279502884197169399375105820974944

struct GV4 {
  double e, x, y, z;
};

// One resonance of PP[RES+CON] with its run constants (MRegge::ME3 / PhotoME3,
// src/MRegge.cc:677-706, 781-843; spin::ProdAmp / DecayAmp, src/MSpin.cc:280-486)
struct ResDev {
  double M0, W, M02, gFF4;  // pole mass, width, M0^2, g_FF^4
  double g_re, g_im;        // production coupling g_A e^{i g_phi}
  double gdec;              // decay coupling sqrt(BR W / PS) (src/MProcess.cc:713)
  double T_re, T_im;        // decay helicity amplitude T_{00} of InitTMatrix
  double B0;                // PhotoProp slope at m^2 = M0^2
  double rho_re[25], rho_im[25];
  int32_t J, P, BW, photo;  // photo: J^P = 1^- through MRegge::PhotoME3
};

// Tensor-pomeron spin modes of MTensorPomeron::ME4 (src/MTensorPomeron.cc:643-654)
constexpr int32_t kTp2xV = 0;  // vector-meson pair
constexpr int32_t kTp2xS = 1;  // pseudoscalar pair

struct RunParams {
  int32_t process, amp_mode, dim, screening;
  // beams and invariants
  GV4 pbeam1, pbeam2, beamsum;
  double s, sqrt_s, rs, s32;
  double c_beam1, c_beam2;  // E - |p_z| of each beam, m^2 / (E + |p_z|)
  // sampling map (BNRandomKin)
  double lmin, lmax, kt_min, kt_max, rap_min, rap_max;
  double m_beam, m_central, w_central, m_daughter;
  double bw_a1, bw_a2, bw_dmin;  // Breit-Wigner inverse-CDF constants
  // fiducial cuts
  int32_t fid_active, pad0;
  double fid_eta_min, fid_eta_max, fid_rap_min, fid_rap_max;
  double fid_pt_min, fid_pt_max, fid_et_min, fid_et_max;
  double sys_M_min, sys_M_max, sys_Y_min, sys_Y_max, sys_Pt_min, sys_Pt_max;
  double fwd_t_min, fwd_t_max;
  // veto domains (elastic forward protons are never tested)
  int32_t veto_active, n_veto;
  double veto_eta_min[GRA_MAX_VETO], veto_eta_max[GRA_MAX_VETO];
  double veto_pt_min[GRA_MAX_VETO], veto_pt_max[GRA_MAX_VETO];
  // weight constants
  double psw_c1, psw_c2, lspan2, KT_vol, PHI_vol, invS, GeV2barn, flux;
  // <F> phase space (MFactorized B51): M_MIN, M_MAX of the system (B51RandomKin,
  // src/MFactorized.cc:256-285), B51IntegralVolume's (M_MAX^2 - M_MIN^2)(Y_max - Y_min)
  // and B51PhaseSpaceWeight's (1/2)(2 pi)^-5
  int32_t phase_space, pad_ps;
  // unequal beam energies: the <C> kinematics are boosted from the centre-of-mass
  // frame to the lab (|beamsum.z| > 1e-6 GeV, as the reference tests)
  int32_t asym, pad_asym;
  double f_Mmin, f_Mmax, f_vol, f_psw;
  // Regge amplitude
  int32_t regge_on[3];
  int32_t regge_k;  // index of the single active trajectory, -1 if several
  double regge_a0[3], regge_ap[3], regge_s0;
  double eta0_re[3], eta0_im[3];  // ReggeEta(alpha(0), sigma)
  double ph_c;                    // -pi/2 alpha'_k of the single trajectory
  double amp_c2;  // |eta0^2 gN_P^2 gpp_P^2|^2: squared common factor of the FP32 amplitude
  double gN_P, fc1, fc2, gpp_P, m_prop2;
  int32_t offshellFF, pad1;
  double b_EXP, a_OREAR, b_OREAR, b_POW;
  // screening loop
  int32_t nphi, nkt;  // intervals
  double step_kt, step_phi, min_loop_kt, hfac, nrm;
  double kappa;  // nrm * hfac (FP32 loop sum normalisation)
  // tensor pomeron: spin mode (kTp2xV, kTp2xS) and couplings
  int32_t tp_mode, pad_tp;
  double tp_delta_P, tp_ap_P, tp_gPNN, tp_gPvv0, tp_gPvv1, tp_mp_ff, m_rho;
  double tp_gP;  // Pomeron-pseudoscalar-pseudoscalar coupling of the 2xS mode
  // resonances (PP[RES+CON]): device table of nres entries
  const ResDev* res;
  int32_t nres, spin_gen, spin_dec, spin_frame;
  double omega;  // PARAM_REGGE::omega
};

// Screening node constants: one entry per (phi_i, kt_j) node, i-major.
struct NodeConst {
  double kx, ky;   // kt cos(phi_i), kt sin(phi_i)
  double kt;       // kt_j
  double w;        // Simpson 3/8 weight U_i V_j
  double aeik_re, aeik_im;  // A_eik(kt_j^2) from the interpolated table
  float cf_re, cf_im;       // FP32: w * kt * A_eik (scaled by the run's table scale)
  float kxf, kyf, k2f, pad; // FP32 kx, ky, kt^2
};

// Compact FP32 node of the Regge screening loop (FP32 amplitude kernels). On the
// device the node table holds the NodeConst entries, then one NodeF32 per node in
// the same order, then one NodeRow per phi row (the node's direction).
struct NodeF32 {
  float kt, k2;  // kt, kt^2
  float arg;     // arg(w kt A_eik) / (2 pi), in revolutions
  float mod;     // |w kt A_eik|
};
struct NodeRow {
  float c, s;  // cos(phi_i), sin(phi_i)
};
