// Normalizing-flow importance sampler: rational-quadratic spline (RQS) coupling
// layers (Durkan, Bekasov, Murray, Papamakarios 2019) over the unit hypercube of
// the VEGAS grid, trained by minimising the variance of w = f / q with the
// score-function gradient  grad Var_q(w) = -E_q[w^2 grad log q]
// (Mueller et al. 2019, neural importance sampling). It occupies the NEURO
// integrator slot of the reference (MGraniitti::SampleNeuro, src/MGraniitti.cc:
// 1527-1648; include/Graniitti/MNeuroJacobian.h).
//
// Direction convention: F = F_L o ... o F_1 maps a point u of the VEGAS unit
// cube to z; log q(u) = sum_l log dF_l. Sampling draws z ~ U(0,1)^D and inverts
// the layers (analytic RQS inverse). Training re-uses the sampling pass's stored
// layer inputs and MLP activations and back-propagates -sum_b c_b log q(u_b),
// c_b = w_b^2 / sum w^2, through the forward direction.
//
// MLP: input (FP32, VALU) -> depth hidden ReLU layers -> output, the hidden and
// output products as BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation;
// a pure FP32 path (same kernels, VALU FMA) serves gradient checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "gra_device.h"
#include "gra_internal.h"
#include "gra_kernels.h"

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

namespace {

constexpr float kEpsW = 1e-3f;  // minimum bin width / height
constexpr float kEpsD = 1e-3f;  // minimum knot derivative
constexpr int kMaxK = 32;       // spline bins
constexpr int kMaxDB = 8;       // transformed dimensions per layer
// Parameter columns per transformed dimension: the 3K - 1 spline parameters
// (K width logits, K height logits, K - 1 interior derivative logits) padded to
// a multiple of 8, so that every dimension's block starts 32-byte aligned in
// FP32 and 16-byte aligned in BF16. The padding columns are frozen at zero.
__host__ __device__ constexpr int prm_stride(int K) { return (3 * K - 1 + 7) / 8 * 8; }

__device__ __forceinline__ float softplusf(float x) {
  return x > 20.f ? x : log1pf(__expf(x));
}
__device__ __forceinline__ float sigmoidf(float x) { return 1.f / (1.f + __expf(-x)); }

// ---------------------------------------------------------------------------
// RQS kernels. One spline (one point, one transformed dimension) is handled by
// a group of G lanes (G = 8, 16 or 32, K <= G): lane m owns bin m. The softmax
// of the widths and heights is a group max / sum, the knot positions are an
// exclusive prefix sum over the group, the bin of a point is the population
// count of a ballot, and the raw parameters and their gradients are read and
// written as contiguous runs (coalesced), not one 3K-1 row per thread.
// ---------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ float grp_max(float v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, G));
  return v;
}
template <int G>
__device__ __forceinline__ float grp_sum(float v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}
template <int G>
__device__ __forceinline__ float grp_excl_scan(float v, int m) {
  float s = v;
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    const float t = __shfl_up(s, d, G);
    if (m >= d) s += t;
  }
  return s - v;
}
// number of lanes of this lane's group whose predicate holds
template <int G>
__device__ __forceinline__ int grp_count(bool pred) {
  const unsigned long long bl = __ballot(pred);
  const int base = (threadIdx.x & 63) & ~(G - 1);
  const unsigned long long msk = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
  return __popcll((bl >> base) & msk);
}

// Knots of one spline: lane m (< K) holds bin m's width, height, lower knot
// (X, Y) and normalised softmax weights; p[0..K) widths, p[K..2K) heights,
// p[2K..3K-1) interior derivatives (raw).
struct BinLane {
  float w, h, X, Y, sw, sh;
};
template <int G>
__device__ __forceinline__ void grp_knots(const float* __restrict__ p, int K, int m, BinLane& q) {
  const bool act = m < K;
  const float aw = act ? p[m] : -3.0e38f;
  const float ah = act ? p[K + m] : -3.0e38f;
  const float mw = grp_max<G>(aw), mh = grp_max<G>(ah);
  const float ew = act ? __expf(aw - mw) : 0.f;
  const float eh = act ? __expf(ah - mh) : 0.f;
  const float iw = 1.f / grp_sum<G>(ew), ih = 1.f / grp_sum<G>(eh);
  const float cw = 1.f - K * kEpsW;
  q.sw = ew * iw;
  q.sh = eh * ih;
  q.w = act ? kEpsW + cw * q.sw : 0.f;
  q.h = act ? kEpsW + cw * q.sh : 0.f;
  q.X = grp_excl_scan<G>(q.w, m);
  q.Y = grp_excl_scan<G>(q.h, m);
}

// The selected bin's parameters, identical in every lane of the group
struct Bin {
  int k;
  float X, w, Y, h, dk, dk1, s;
};
template <int G>
__device__ __forceinline__ void grp_bin(const float* __restrict__ p, int K, const BinLane& q, int k,
                                        Bin& e) {
  e.k = k;
  e.X = __shfl(q.X, k, G);
  e.w = __shfl(q.w, k, G);
  e.Y = __shfl(q.Y, k, G);
  e.h = __shfl(q.h, k, G);
  e.dk = (k == 0) ? 1.f : kEpsD + softplusf(p[2 * K + k - 1]);
  e.dk1 = (k == K - 1) ? 1.f : kEpsD + softplusf(p[2 * K + k]);
  e.s = e.h / e.w;
}

// Forward value and log-derivative at xi
__device__ __forceinline__ void rqs_value(const Bin& e, float xi, float& z, float& logJ) {
  const float t = xi * (1.f - xi);
  const float Dn = e.s + (e.dk1 + e.dk - 2.f * e.s) * t;
  const float N = e.h * (e.s * xi * xi + e.dk * t);
  z = e.Y + N / Dn;
  const float Pn = e.dk1 * xi * xi + 2.f * e.s * t + e.dk * (1.f - xi) * (1.f - xi);
  logJ = 2.f * __logf(e.s) + __logf(Pn) - 2.f * __logf(Dn);
}

struct LayerIdx {
  int dA, dB;
  int A[16], B[16];
};

// Sampling direction: u_{l-1}[B] = RQS^{-1}(u_l[B]), u_{l-1}[A] = u_l[A];
// ld[b] += log dF_l at the pre-image.
template <int G>
__global__ __launch_bounds__(256) void k_rqs_inverse(const float* __restrict__ prm, int ldp,
                                                     const float* __restrict__ u_in,
                                                     float* __restrict__ u_out, float* __restrict__ ld,
                                                     int n, int D, LayerIdx L, int K) {
  const int m = threadIdx.x & (G - 1);
  const int b = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) / G);
  if (b >= n) return;
  const float* uin = u_in + (size_t)b * D;
  float* uo = u_out + (size_t)b * D;
  if (m < L.dA) uo[L.A[m]] = uin[L.A[m]];
  float acc = 0.f;
  const int nps = prm_stride(K);
  for (int j = 0; j < L.dB; ++j) {
    const float* p = prm + (size_t)b * ldp + j * nps;
    BinLane q;
    grp_knots<G>(p, K, m, q);
    const float zv = fminf(fmaxf(uin[L.B[j]], 0.f), 1.f);
    const int k = max(grp_count<G>(m < K && q.Y <= zv) - 1, 0);
    Bin e;
    grp_bin<G>(p, K, q, k, e);
    const float zeta = zv - e.Y;
    const float Dl = e.dk1 + e.dk - 2.f * e.s;
    const float qa = e.h * (e.s - e.dk) + zeta * Dl;
    const float qb = e.h * e.dk - zeta * Dl;
    const float qc = -e.s * zeta;
    const float disc = fmaxf(qb * qb - 4.f * qa * qc, 0.f);
    float xi = (2.f * qc) / (-qb - sqrtf(disc));
    xi = fminf(fmaxf(xi, 0.f), 1.f);
    float zz, lj;
    rqs_value(e, xi, zz, lj);
    if (m == 0) uo[L.B[j]] = fminf(fmaxf(e.X + xi * e.w, 0.f), 0.99999994f);
    acc += lj;
  }
  if (m == 0) ld[b] += acc;
}

// Density direction: u_l[B] = RQS(u_{l-1}[B]) (reconstructs a sampling pass's
// layer inputs from the flow's output points)
template <int G>
__global__ __launch_bounds__(256) void k_rqs_forward(const float* __restrict__ prm, int ldp,
                                                     const float* __restrict__ u_in,
                                                     float* __restrict__ u_out, int n, int D,
                                                     LayerIdx L, int K) {
  const int m = threadIdx.x & (G - 1);
  const int b = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) / G);
  if (b >= n) return;
  const float* uin = u_in + (size_t)b * D;
  float* uo = u_out + (size_t)b * D;
  if (m < L.dA) uo[L.A[m]] = uin[L.A[m]];
  const int nps = prm_stride(K);
  for (int j = 0; j < L.dB; ++j) {
    const float* p = prm + (size_t)b * ldp + j * nps;
    BinLane q;
    grp_knots<G>(p, K, m, q);
    const float x = fminf(fmaxf(uin[L.B[j]], 0.f), 1.f);
    const int k = max(grp_count<G>(m < K && q.X <= x) - 1, 0);
    Bin e;
    grp_bin<G>(p, K, q, k, e);
    const float xi = fminf(fmaxf((x - e.X) / e.w, 0.f), 1.f);
    float z, lj;
    rqs_value(e, xi, z, lj);
    if (m == 0) uo[L.B[j]] = z;
  }
}

// Training direction, backward of layer l with one thread per (point,
// transformed dimension). Tile form: the block's spline-parameter rows are
// staged in LDS (coalesced), each thread overwrites its own 3K - 1 parameters
// with their gradients, and the rows are written back as FP32 and BF16
// (coalesced). Direct form (compile-time K): each thread reads its parameter
// block (32-byte aligned) with 16-byte loads and writes its gradients with
// 16-byte stores of 8 BF16 (and 4 FP32) values, padding included. The softmax
// backward needs sum_m sw_m G_m, which for the width (height) gradient
// G_m = G_X (m < k), G_w (m = k), 0 (m > k) is
//   G_X (X_k - k eps) / (1 - K eps) + G_w sw_k,
// so no per-bin arrays are kept.
template <int KT, bool DIRECT>
__global__ __launch_bounds__(256) void k_rqs_backward_t(
    const float* __restrict__ prm, int ldp, const float* __restrict__ u_prev,
    const float* __restrict__ g_in, const float* __restrict__ coef, float* __restrict__ g_out,
    float* __restrict__ gp, __bf16* __restrict__ gp16, int n, int D, LayerIdx L, int Kin, int R,
    float* __restrict__ logq_chk) {
  // KT > 0: compile-time K, logits and their exponentials held in registers.
  // DIRECT: each thread reads its parameters from prm and writes its gradients
  // to gp16 (and gp) itself, without the LDS tile, whose size otherwise limits
  // the blocks per CU when the parameter rows are long
  static_assert(KT > 0 || !DIRECT, "the direct form needs a compile-time K");
  const int K = KT > 0 ? KT : Kin;
  extern __shared__ float tile[];  // [R][ldp + 1]
  __shared__ float sLj[256];
  const int ldt = ldp + 1;
  const int row0 = blockIdx.x * R;
  // the block's rows are contiguous in prm: walk them with an incremental (row, column)
  const int nrow = min(R, n - row0);
  if constexpr (!DIRECT) {
    int r = threadIdx.x / ldp, c = threadIdx.x % ldp;
    const int dr = blockDim.x / ldp, dc = blockDim.x % ldp;
    for (int e = threadIdx.x; e < nrow * ldp; e += blockDim.x) {
      tile[r * ldt + c] = prm[(size_t)row0 * ldp + e];
      r += dr;
      c += dc;
      if (c >= ldp) {
        c -= ldp;
        ++r;
      }
    }
  }
  __syncthreads();
  const int t = threadIdx.x;
  const int r = t / L.dB, j = t % L.dB;
  const int b = row0 + r;
  const int nps = prm_stride(K);
  const float cw = 1.f - K * kEpsW;
  float lj = 0.f;
  if (r < R && b < n) {
    const float* up = u_prev + (size_t)b * D;
    const float* gi = g_in + (size_t)b * D;
    float* go = g_out + (size_t)b * D;
    if (j == 0)
      for (int i = 0; i < L.dA; ++i) go[L.A[i]] = gi[L.A[i]];
    const float gld = -coef[b];
    const size_t o = (size_t)b * ldp + j * nps;  // this spline's parameter block
    const float* p = DIRECT ? prm + o : tile + r * ldt + j * nps;
    float* pw = DIRECT ? nullptr : tile + r * ldt + j * nps;
    if constexpr (DIRECT) {
      // zero the row's padding columns past the last dimension's block
      if (j == 0) {
        bf16x8 z16;
#pragma unroll
        for (int e = 0; e < 8; ++e) z16[e] = (__bf16)0.f;
        for (int c = L.dB * nps; c < ldp; c += 8) {
          *(bf16x8*)(gp16 + (size_t)b * ldp + c) = z16;
          if (gp) {
            *(float4*)(gp + (size_t)b * ldp + c) = make_float4(0.f, 0.f, 0.f, 0.f);
            *(float4*)(gp + (size_t)b * ldp + c + 4) = make_float4(0.f, 0.f, 0.f, 0.f);
          }
        }
      }
    }
    const float x = fminf(fmaxf(up[L.B[j]], 0.f), 1.f);
    float ew[KT > 0 ? KT : 1], eh[KT > 0 ? KT : 1];
    float mw, mh, iw, ih;
    float X = 0.f, Y = 0.f, swk = 0.f, shk = 0.f;
    int k = 0;
    if constexpr (KT > 0) {
      if constexpr (DIRECT) {
#pragma unroll
        for (int m = 0; m < KT; m += 4) {
          const float4 a = *(const float4*)(p + m), c = *(const float4*)(p + KT + m);
          ew[m] = a.x, ew[m + 1] = a.y, ew[m + 2] = a.z, ew[m + 3] = a.w;
          eh[m] = c.x, eh[m + 1] = c.y, eh[m + 2] = c.z, eh[m + 3] = c.w;
        }
      } else {
#pragma unroll
        for (int m = 0; m < KT; ++m) {
          ew[m] = p[m];
          eh[m] = p[KT + m];
        }
      }
      mw = ew[0];
      mh = eh[0];
#pragma unroll
      for (int m = 1; m < KT; ++m) {
        mw = fmaxf(mw, ew[m]);
        mh = fmaxf(mh, eh[m]);
      }
      float sw = 0.f, sh = 0.f;
#pragma unroll
      for (int m = 0; m < KT; ++m) {
        ew[m] = __expf(ew[m] - mw);
        eh[m] = __expf(eh[m] - mh);
        sw += ew[m];
        sh += eh[m];
      }
      iw = 1.f / sw;
      ih = 1.f / sh;
      float Xa = 0.f, Ya = 0.f;
      bool found = false;
#pragma unroll
      for (int m = 0; m < KT; ++m) {
        const float swm = ew[m] * iw, shm = eh[m] * ih;
        const float w = kEpsW + cw * swm;
        const bool hit = !found && (m == KT - 1 || Xa + w > x);
        k = hit ? m : k;
        X = hit ? Xa : X;
        Y = hit ? Ya : Y;
        swk = hit ? swm : swk;
        shk = hit ? shm : shk;
        found = found || hit;
        Xa += w;
        Ya += kEpsW + cw * shm;
      }
    } else {
      mw = p[0];
      mh = p[K];
      for (int m = 1; m < K; ++m) {
        mw = fmaxf(mw, p[m]);
        mh = fmaxf(mh, p[K + m]);
      }
      float sw = 0.f, sh = 0.f;
      for (int m = 0; m < K; ++m) {
        sw += __expf(p[m] - mw);
        sh += __expf(p[K + m] - mh);
      }
      iw = 1.f / sw;
      ih = 1.f / sh;
      for (; k < K; ++k) {
        swk = __expf(p[k] - mw) * iw;
        shk = __expf(p[K + k] - mh) * ih;
        const float w = kEpsW + cw * swk;
        if (k == K - 1 || X + w > x) break;
        X += w;
        Y += kEpsW + cw * shk;
      }
    }
    Bin e;
    e.k = k;
    e.X = X;
    e.Y = Y;
    e.w = kEpsW + cw * swk;
    e.h = kEpsW + cw * shk;
    const float sgk = (k > 0) ? sigmoidf(p[2 * K + k - 1]) : 0.f;
    const float sgk1 = (k < K - 1) ? sigmoidf(p[2 * K + k]) : 0.f;
    e.dk = (k == 0) ? 1.f : kEpsD + softplusf(p[2 * K + k - 1]);
    e.dk1 = (k == K - 1) ? 1.f : kEpsD + softplusf(p[2 * K + k]);
    e.s = e.h / e.w;
    const float xi = fminf(fmaxf((x - e.X) / e.w, 0.f), 1.f);
    const float tt = xi * (1.f - xi), om = 1.f - xi;
    const float Dl = e.dk1 + e.dk - 2.f * e.s;
    const float Dn = e.s + Dl * tt;
    const float N = e.h * (e.s * xi * xi + e.dk * tt);
    const float Pn = e.dk1 * xi * xi + 2.f * e.s * tt + e.dk * om * om;
    lj = 2.f * __logf(e.s) + __logf(Pn) - 2.f * __logf(Dn);
    const float gu = gi[L.B[j]];
    const float N_xi = e.h * (2.f * e.s * xi + e.dk * (1.f - 2.f * xi));
    const float N_s = e.h * xi * xi, N_h = e.s * xi * xi + e.dk * tt, N_dk = e.h * tt;
    const float D_xi = Dl * (1.f - 2.f * xi), D_s = 1.f - 2.f * tt, D_dk = tt, D_dk1 = tt;
    const float P_xi = 2.f * e.dk1 * xi + 2.f * e.s * (1.f - 2.f * xi) - 2.f * e.dk * om;
    const float P_s = 2.f * tt, P_dk = om * om, P_dk1 = xi * xi;
    const float iD = 1.f / Dn, iD2 = iD * iD, iP = 1.f / Pn;
    const float G_xi = gu * (N_xi * Dn - N * D_xi) * iD2 + gld * (P_xi * iP - 2.f * D_xi * iD);
    const float G_s = gu * (N_s * Dn - N * D_s) * iD2 + gld * (2.f / e.s + P_s * iP - 2.f * D_s * iD);
    const float G_h0 = gu * N_h * iD;
    const float G_dk = gu * (N_dk * Dn - N * D_dk) * iD2 + gld * (P_dk * iP - 2.f * D_dk * iD);
    const float G_dk1 = gu * (-N * D_dk1) * iD2 + gld * (P_dk1 * iP - 2.f * D_dk1 * iD);
    const float G_Y = gu;
    const float iwk = 1.f / e.w;
    const float G_X = -G_xi * iwk;
    const float G_w = -G_xi * xi * iwk - G_s * e.s * iwk;
    const float G_h = G_h0 + G_s * iwk;
    go[L.B[j]] = G_xi * iwk;
    const float icw = 1.f / cw;
    const float sdw = G_X * (X - k * kEpsW) * icw + G_w * swk;
    const float sdh = G_Y * (Y - k * kEpsW) * icw + G_h * shk;
    // gradient of width / height logit m and of interior derivative logit m
    auto gwid = [&](int m, float swm) {
      return cw * swm * (((m < k) ? G_X : (m == k ? G_w : 0.f)) - sdw);
    };
    auto ghei = [&](int m, float shm) {
      return cw * shm * (((m < k) ? G_Y : (m == k ? G_h : 0.f)) - sdh);
    };
    auto gder = [&](int m) {
      return (m == k - 1 ? G_dk * sgk : 0.f) + (m == k ? G_dk1 * sgk1 : 0.f);
    };
    if constexpr (DIRECT) {
      // 8 gradients per 16-byte BF16 store: widths, heights, then the K - 1
      // derivative logits and the block's padding (K % 8 == 0: the block is 3K)
#pragma unroll
      for (int m0 = 0; m0 < KT; m0 += 8) {
        float vw[8], vh[8], vd[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int m = m0 + e;
          vw[e] = gwid(m, ew[m] * iw);
          vh[e] = ghei(m, eh[m] * ih);
          vd[e] = m < KT - 1 ? gder(m) : 0.f;
        }
        bf16x8 a16, b16, c16;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          a16[e] = (__bf16)vw[e];
          b16[e] = (__bf16)vh[e];
          c16[e] = (__bf16)vd[e];
        }
        *(bf16x8*)(gp16 + o + m0) = a16;
        *(bf16x8*)(gp16 + o + KT + m0) = b16;
        *(bf16x8*)(gp16 + o + 2 * KT + m0) = c16;
        if (gp) {
#pragma unroll
          for (int e = 0; e < 8; e += 4) {
            *(float4*)(gp + o + m0 + e) = make_float4(vw[e], vw[e + 1], vw[e + 2], vw[e + 3]);
            *(float4*)(gp + o + KT + m0 + e) =
                make_float4(vh[e], vh[e + 1], vh[e + 2], vh[e + 3]);
            *(float4*)(gp + o + 2 * KT + m0 + e) =
                make_float4(vd[e], vd[e + 1], vd[e + 2], vd[e + 3]);
          }
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < K; ++m) {
        float swm, shm;
        if constexpr (KT > 0) {
          swm = ew[m] * iw;
          shm = eh[m] * ih;
        } else {
          swm = __expf(p[m] - mw) * iw;
          shm = __expf(p[K + m] - mh) * ih;
        }
        pw[m] = gwid(m, swm);
        pw[K + m] = ghei(m, shm);
      }
      for (int m = 0; m < K - 1; ++m) pw[2 * K + m] = gder(m);
      // the block's padding columns keep the parameters' zeros
    }
  }
  sLj[t] = lj;
  __syncthreads();
  if constexpr (!DIRECT) {
    int rr = threadIdx.x / ldp, c = threadIdx.x % ldp;
    const int dr = blockDim.x / ldp, dc = blockDim.x % ldp;
    for (int e2 = threadIdx.x; e2 < nrow * ldp; e2 += blockDim.x) {
      const float v = tile[rr * ldt + c];
      if (gp) gp[(size_t)row0 * ldp + e2] = v;
      gp16[(size_t)row0 * ldp + e2] = (__bf16)v;
      rr += dr;
      c += dc;
      if (c >= ldp) {
        c -= ldp;
        ++rr;
      }
    }
  }
  if (logq_chk && t < R && row0 + t < n) {
    float acc = 0.f;
    for (int jj = 0; jj < L.dB; ++jj) acc += sLj[t * L.dB + jj];
    logq_chk[row0 + t] += acc;
  }
}

// ---------------------------------------------------------------------------
// MLP kernels
// ---------------------------------------------------------------------------
// Input layer (FP32 VALU): h[b][j] = relu(sum_i u[b][A_i] W1[i][j] + b1[j])
__global__ void k_mlp_in(const float* __restrict__ u, int D, LayerIdx L, const float* __restrict__ W1,
                         const float* __restrict__ b1, int H, int n, __bf16* __restrict__ h16,
                         float* __restrict__ h32) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = (int)(idx / H), j = (int)(idx % H);
  if (b >= n) return;
  float acc = b1[j];
  for (int i = 0; i < L.dA; ++i) acc += u[(size_t)b * D + L.A[i]] * W1[i * H + j];
  acc = fmaxf(acc, 0.f);
  if (h16) h16[(size_t)b * H + j] = (__bf16)acc;
  if (h32) h32[(size_t)b * H + j] = acc;
}

// Input layer backward, dx part: du[b][A_i] += sum_j dh[b][j] W1[i][j], with dh
// in FP32 (FP32 MLP path) or BF16 (the BF16 path's masked dX product). A block
// takes 256 points; 32-column slices of dh are staged in LDS with coalesced
// loads (rows padded to 33 words; BF16 rows as 16-byte pieces of 8 columns),
// then each thread accumulates the dot products of its own point.
template <class T>
__global__ __launch_bounds__(256) void k_mlp_in_dx(const T* __restrict__ dh,
                                                   const float* __restrict__ W1, int H, int n, int D,
                                                   LayerIdx L, float* __restrict__ g) {
  __shared__ float tile[256][33];
  __shared__ float wsl[kMaxDB][32];
  const int t = threadIdx.x;
  const int b0 = blockIdx.x * 256;
  float acc[kMaxDB];
#pragma unroll
  for (int i = 0; i < kMaxDB; ++i) acc[i] = 0.f;
  for (int c0 = 0; c0 < H; c0 += 32) {
    if constexpr (sizeof(T) == 2) {
      for (int e = t; e < 256 * 4; e += 256) {
        const int r = e >> 2, c = (e & 3) * 8;
        bf16x8 v;
        if (b0 + r < n) v = *(const bf16x8*)(dh + (size_t)(b0 + r) * H + c0 + c);
        else
          for (int q = 0; q < 8; ++q) v[q] = (__bf16)0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) tile[r][c + q] = (float)v[q];
      }
    } else {
      for (int e = t; e < 256 * 32; e += 256) {
        const int r = e >> 5, c = e & 31;
        tile[r][c] = (b0 + r < n) ? (float)dh[(size_t)(b0 + r) * H + c0 + c] : 0.f;
      }
    }
    if (t < kMaxDB * 32) {
      const int i = t >> 5, c = t & 31;
      wsl[i][c] = (i < L.dA) ? W1[i * H + c0 + c] : 0.f;
    }
    __syncthreads();
#pragma unroll 4
    for (int c = 0; c < 32; ++c) {
      const float d = tile[t][c];
#pragma unroll
      for (int i = 0; i < kMaxDB; ++i) acc[i] += d * wsl[i][c];
    }
    __syncthreads();
  }
  const int b = b0 + t;
  if (b < n) {
#pragma unroll
    for (int i = 0; i < kMaxDB; ++i)
      if (i < L.dA) g[(size_t)b * D + L.A[i]] += acc[i];
  }
}

// Sum of the per-chunk weight-gradient partials: rows < kw go to gW[row][col],
// row kw (the all-ones input row) to the bias gradient gb[col].
// Fixed-order sum of the per-chunk partials: 32 outputs per block of 256 threads;
// group g of 8 sums chunks g, g + 8, g + 16, ... in order, and the 8 group sums
// are then added in order (deterministic, independent of the schedule)
constexpr int kReduceGroups = 8;
__global__ __launch_bounds__(256) void k_part_reduce(const float* __restrict__ part, int nchunk,
                                                     int kin_t, int N, int kw,
                                                     float* __restrict__ gW,
                                                     float* __restrict__ gb) {
  __shared__ float s[kReduceGroups][33];
  const int o = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int idx = blockIdx.x * 32 + o;
  const bool on = idx < (kw + 1) * N;
  const int row = on ? idx / N : 0, col = on ? idx % N : 0;
  float acc = 0.f;
  // unrolled so that eight loads are in flight before their (in-order) adds
  if (on) {
#pragma unroll 8
    for (int k = g; k < nchunk; k += kReduceGroups) acc += part[((size_t)k * kin_t + row) * N + col];
  }
  s[g][o] = acc;
  __syncthreads();
  if (g != 0 || !on) return;
  float t = s[0][o];
#pragma unroll
  for (int j = 1; j < kReduceGroups; ++j) t += s[j][o];
  if (row < kw) gW[(size_t)row * N + col] = t;
  else gb[col] = t;
}

// Dense product C[M x N] = A[M x K] * Bt[N x K]^T with row-major A and Bt,
// 64 x 64 tiles, 4 waves each computing 16 rows x 64 columns.
// Epilogue: + bias[col], ReLU, multiply by (mask[row][col] > 0), stores bf16 and/or fp32.
struct Epi {
  const float* bias;
  int relu;
  const __bf16* mask;  // optional, ldm
  int ldm;
  __bf16* out16;
  float* out32;
  int ldc;
  int ncols_valid;  // columns >= this are written as zero
};

template <bool BF16>
__global__ __launch_bounds__(256) void k_gemm_nt(const void* __restrict__ Av, int lda,
                                                 const void* __restrict__ Bv, int ldb, int M, int N,
                                                 int K, Epi ep) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = blockIdx.x * 64 + wave * 16, n0 = blockIdx.y * 64;
  f32x4 acc[4];
  for (int t = 0; t < 4; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  if (BF16) {
    const __bf16* A = (const __bf16*)Av;
    const __bf16* Bt = (const __bf16*)Bv;
    const int ar = m0 + (lane & 15), kq = 8 * (lane >> 4);
    for (int k0 = 0; k0 < K; k0 += 32) {
      const bf16x8 a = *(const bf16x8*)(A + (size_t)ar * lda + k0 + kq);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x8 bb = *(const bf16x8*)(Bt + (size_t)(n0 + 16 * t + (lane & 15)) * ldb + k0 + kq);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[t], 0, 0, 0);
      }
    }
  } else {
    // FP32 VALU path with the same output ownership as the MFMA layout
    const float* A = (const float*)Av;
    const float* Bt = (const float*)Bv;
    for (int t = 0; t < 4; ++t) {
      const int col = n0 + 16 * t + (lane & 15);
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + 4 * (lane >> 4) + r;
        float s = 0.f;
        for (int k = 0; k < K; ++k) s += A[(size_t)row * lda + k] * Bt[(size_t)col * ldb + k];
        acc[t][r] = s;
      }
    }
  }
  if (BF16 && (ep.ldc & 7) == 0 && (ep.ldm & 7) == 0) {
    // the wave's 16 x 64 tile through LDS, so that each lane then reads its
    // mask and writes its outputs as 16-byte vectors: lane -> row lane / 4,
    // columns 16 (lane % 4) .. + 15 (full 128-byte row segments)
    __shared__ float sT[4][16][64 + 4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) sT[wave][4 * (lane >> 4) + r][16 * t + (lane & 15)] = acc[t][r];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int rr = lane >> 2, c0 = 16 * (lane & 3);
    const size_t row = (size_t)(m0 + rr);
    const int col0 = n0 + c0;
    float v[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      v[c] = sT[wave][rr][c0 + c];
      if (ep.bias) v[c] += ep.bias[col0 + c];
      if (ep.relu) v[c] = fmaxf(v[c], 0.f);
    }
    if (ep.mask) {
      bf16x8 mk[2];
      mk[0] = *(const bf16x8*)(ep.mask + row * ep.ldm + col0);
      mk[1] = *(const bf16x8*)(ep.mask + row * ep.ldm + col0 + 8);
#pragma unroll
      for (int c = 0; c < 16; ++c)
        if (!((float)mk[c >> 3][c & 7] > 0.f)) v[c] = 0.f;
    }
#pragma unroll
    for (int c = 0; c < 16; ++c)
      if (col0 + c >= ep.ncols_valid) v[c] = 0.f;
    if (ep.out16) {
      bf16x8 o[2];
#pragma unroll
      for (int c = 0; c < 16; ++c) o[c >> 3][c & 7] = (__bf16)v[c];
      *(bf16x8*)(ep.out16 + row * ep.ldc + col0) = o[0];
      *(bf16x8*)(ep.out16 + row * ep.ldc + col0 + 8) = o[1];
    }
    if (ep.out32) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        *(f32x4*)(ep.out32 + row * ep.ldc + col0 + 4 * q) =
            (f32x4){v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
    }
    return;
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int col = n0 + 16 * t + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = m0 + 4 * (lane >> 4) + r;
      float v = acc[t][r];
      if (ep.bias) v += ep.bias[col];
      if (ep.relu) v = fmaxf(v, 0.f);
      if (ep.mask && !((float)ep.mask[(size_t)row * ep.ldm + col] > 0.f)) v = 0.f;
      if (col >= ep.ncols_valid) v = 0.f;
      if (ep.out16) ep.out16[(size_t)row * ep.ldc + col] = (__bf16)v;
      if (ep.out32) ep.out32[(size_t)row * ep.ldc + col] = v;
    }
  }
}

// BF16 C[M x N] = A[M x K] Bt[N x K]^T with every output column in one block of
// 128 rows: each wave takes 32 rows (two m-tiles) x all 16 NT columns, so A is
// read once, and each k-step's slice of Bt (N x 32) is staged once per block in
// LDS (double-buffered: the next slice is loaded into registers while the
// current one feeds the MFMAs). The epilogue of k_gemm_nt, 16 rows x 64
// columns at a time through LDS (over the B buffers once the products are done).
template <int NT>
__global__ __launch_bounds__(256) void k_gemm_nt_wide(const __bf16* __restrict__ A, int lda,
                                                      const __bf16* __restrict__ Bt, int ldb,
                                                      int K, Epi ep) {
  constexpr int N = 16 * NT, kLd = 40;        // B slice row stride (32 + 8) in BF16
  constexpr int kChunks = N * 4 / 256;        // 16-byte pieces of a slice per thread
  // two B slices, or the four waves' 16 x 68 FP32 epilogue tiles, whichever is larger
  constexpr int kBytes = 2 * N * kLd * 2 > 4 * 16 * 68 * 4 ? 2 * N * kLd * 2 : 4 * 16 * 68 * 4;
  __shared__ __attribute__((aligned(16))) unsigned char smem_nt[kBytes];
  __bf16 (*sB)[N * kLd] = (__bf16 (*)[N * kLd])smem_nt;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = blockIdx.x * 128 + wave * 32;
  const int kq = 8 * (lane >> 4);
  f32x4 acc[2][NT];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[i][t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  bf16x8 pre[kChunks];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      const int q = threadIdx.x + 256 * c, n = q >> 2, seg = q & 3;
      pre[c] = *(const bf16x8*)(Bt + (size_t)n * ldb + k0 + 8 * seg);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      const int q = threadIdx.x + 256 * c, n = q >> 2, seg = q & 3;
      *(bf16x8*)(&sB[buf][n * kLd + 8 * seg]) = pre[c];
    }
  };
  fetch(0);
  stash(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bool more = k0 + 32 < K;
    if (more) fetch(k0 + 32);
    bf16x8 a[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
      a[i] = *(const bf16x8*)(A + (size_t)(m0 + 16 * i + (lane & 15)) * lda + k0 + kq);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x8 bb = *(const bf16x8*)(&sB[cur][(16 * t + (lane & 15)) * kLd + kq]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
        acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bb, acc[i][t], 0, 0, 0);
    }
    if (more) stash(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  // epilogue: the wave's 16 x 64 tiles through its own slice of the B buffers
  float* sT = (float*)smem_nt + wave * (16 * 68);
  const int rr = lane >> 2, c0 = 16 * (lane & 3);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const size_t row = (size_t)(m0 + 16 * i + rr);
#pragma unroll
    for (int ch = 0; ch < NT / 4; ++ch) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          sT[(4 * (lane >> 4) + r) * 68 + 16 * t + (lane & 15)] = acc[i][4 * ch + t][r];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int col0 = 64 * ch + c0;
      float v[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        v[c] = sT[rr * 68 + c0 + c];
        if (ep.bias) v[c] += ep.bias[col0 + c];
        if (ep.relu) v[c] = fmaxf(v[c], 0.f);
      }
      if (ep.mask) {
        bf16x8 mk[2];
        mk[0] = *(const bf16x8*)(ep.mask + row * ep.ldm + col0);
        mk[1] = *(const bf16x8*)(ep.mask + row * ep.ldm + col0 + 8);
#pragma unroll
        for (int c = 0; c < 16; ++c)
          if (!((float)mk[c >> 3][c & 7] > 0.f)) v[c] = 0.f;
      }
#pragma unroll
      for (int c = 0; c < 16; ++c)
        if (col0 + c >= ep.ncols_valid) v[c] = 0.f;
      if (ep.out16) {
        bf16x8 o[2];
#pragma unroll
        for (int c = 0; c < 16; ++c) o[c >> 3][c & 7] = (__bf16)v[c];
        *(bf16x8*)(ep.out16 + row * ep.ldc + col0) = o[0];
        *(bf16x8*)(ep.out16 + row * ep.ldc + col0 + 8) = o[1];
      }
      if (ep.out32) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          *(f32x4*)(ep.out32 + row * ep.ldc + col0 + 4 * q) =
              (f32x4){v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

// Weight-gradient partials over batch chunks: part[chunk][m][n] =
// sum_{b in chunk} X[b][m] dY[b][n], 64 x 64 tiles, four 16-row waves. X is
// either a dense activation matrix [rows x Kin] followed by one extra m-tile
// whose first row is all ones (so row Kin of the result is the bias gradient,
// the column sum of dY), or, when ug is given, the input layer's conditioning
// coordinates gathered from the points: X[b][i] = ug[b][A_i] for i < dA,
// X[b][dA] = 1, zero beyond. Tiles of 32 batch rows are staged transposed in
// LDS so that each MFMA operand is one 16-byte read.
template <bool BF16>
__global__ __launch_bounds__(256) void k_gemm_tn_part(const void* __restrict__ Xv, int ldx,
                                                      const void* __restrict__ Yv, int ldy, int Kin,
                                                      int N, int nrows, int chunk,
                                                      float* __restrict__ part,
                                                      const float* __restrict__ ug, int D,
                                                      LayerIdx L) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64, ch = blockIdx.z;
  const int kin_t = gridDim.x * 64;
  const int b0 = ch * chunk, b1 = min(nrows, b0 + chunk);
  const bool gather = ug != nullptr;
  const bool ones = !gather && m0 >= Kin;
  f32x4 acc[4];
  for (int t = 0; t < 4; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  if (BF16) {
    __shared__ __attribute__((aligned(16))) __bf16 sx[64][40];
    __shared__ __attribute__((aligned(16))) __bf16 sy[64][40];
    const __bf16* X = (const __bf16*)Xv;
    const __bf16* Y = (const __bf16*)Yv;
    const int r = threadIdx.x >> 3, c8 = (threadIdx.x & 7) * 8;
    for (int bb = b0; bb < b1; bb += 32) {
      const int brow = bb + r;
      const bool valid = brow < b1;
      bf16x8 yv, xv;
      if (valid) yv = *(const bf16x8*)(Y + (size_t)brow * ldy + n0 + c8);
      else
        for (int e = 0; e < 8; ++e) yv[e] = (__bf16)0.f;
      if (gather) {
        for (int e = 0; e < 8; ++e) {
          const int col = c8 + e;
          float v = 0.f;
          if (valid) v = col < L.dA ? ug[(size_t)brow * D + L.A[col]] : (col == L.dA ? 1.f : 0.f);
          xv[e] = (__bf16)v;
        }
      } else if (ones) {
        for (int e = 0; e < 8; ++e) xv[e] = (__bf16)((valid && c8 + e == 0) ? 1.f : 0.f);
      } else if (valid) {
        xv = *(const bf16x8*)(X + (size_t)brow * ldx + m0 + c8);
      } else {
        for (int e = 0; e < 8; ++e) xv[e] = (__bf16)0.f;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        sx[c8 + e][r] = xv[e];
        sy[c8 + e][r] = yv[e];
      }
      __syncthreads();
      const int kq = 8 * (lane >> 4);
      const bf16x8 a = *(const bf16x8*)&sx[wave * 16 + (lane & 15)][kq];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x8 bv = *(const bf16x8*)&sy[16 * t + (lane & 15)][kq];
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[t], 0, 0, 0);
      }
      __syncthreads();
    }
  } else {
    const float* X = (const float*)Xv;
    const float* Y = (const float*)Yv;
    for (int t = 0; t < 4; ++t) {
      const int col = n0 + 16 * t + (lane & 15);
      for (int rr = 0; rr < 4; ++rr) {
        const int row = m0 + wave * 16 + 4 * (lane >> 4) + rr;
        float s = 0.f;
        for (int b = b0; b < b1; ++b) {
          float xb;
          if (gather) xb = row < L.dA ? ug[(size_t)b * D + L.A[row]] : (row == L.dA ? 1.f : 0.f);
          else xb = row < Kin ? X[(size_t)b * ldx + row] : (row == Kin ? 1.f : 0.f);
          s += xb * Y[(size_t)b * ldy + col];
        }
        acc[t][rr] = s;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int col = n0 + 16 * t + (lane & 15);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int row = m0 + wave * 16 + 4 * (lane >> 4) + rr;
      part[((size_t)ch * kin_t + row) * N + col] = acc[t][rr];
    }
  }
}

// Weight-gradient partials of a BF16 product, one block per (row chunk, group of
// 64 output columns): part[ch][m][n] = sum_{b in chunk} X'[b][m] dY[b][n] for all
// m < 16 MT, where X' is X [rows x Kin] followed by a column of ones (row Kin of
// the result is the bias gradient, the column sum of dY) and zeros up to 16 MT.
// Each step of 32 rows stages X'^T [16 MT][32] and dY^T [64][32] in LDS: a thread
// gathers 8 consecutive rows of one column with coalesced 2-byte loads (a wave
// reads 64 consecutive columns of a row per load) and writes them as one 16-byte
// MFMA operand. Wave w then multiplies m-tiles w, w + 4, ... against the four
// n-tiles, whose fragments it reads once per step.
template <int MW>
__global__ __launch_bounds__(256) void k_wgrad(const __bf16* __restrict__ X, int ldx,
                                               const __bf16* __restrict__ Y, int ldy, int Kin,
                                               int N, int nrows, int chunk,
                                               float* __restrict__ part) {
  constexpr int kLd = 40;  // BF16 row stride of the staged tiles (32 rows + 8)
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int MT = Kin / 16 + 1, KT = 16 * MT;  // KT columns of X' (Kin, the ones, zeros)
  __bf16* sx = (__bf16*)smem;                 // [KT][kLd]
  __bf16* sy = sx + KT * kLd;                 // [64][kLd]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ch = blockIdx.x, n0 = blockIdx.y * 64;
  const int b0 = ch * chunk, b1 = min(nrows, b0 + chunk);
  f32x4 acc[MW][4];
#pragma unroll
  for (int i = 0; i < MW; ++i)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[i][t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int kq = 8 * (lane >> 4);
  for (int bb = b0; bb < b1; bb += 32) {
    // X'^T: 4 KT groups of (column, 8 rows)
    for (int q = threadIdx.x; q < 4 * KT; q += 256) {
      const int c = q % KT, g = q / KT;
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int b = bb + 8 * g + j;
        float one = (c == Kin && b < b1) ? 1.f : 0.f;
        v[j] = (c < Kin && b < b1) ? X[(size_t)b * ldx + c] : (__bf16)one;
      }
      *(bf16x8*)(sx + c * kLd + 8 * g) = v;
    }
    {
      const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int b = bb + 8 * g + j;
        v[j] = b < b1 ? Y[(size_t)b * ldy + n0 + c] : (__bf16)0.f;
      }
      *(bf16x8*)(sy + c * kLd + 8 * g) = v;
    }
    __syncthreads();
    bf16x8 bv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) bv[t] = *(const bf16x8*)(sy + (16 * t + (lane & 15)) * kLd + kq);
#pragma unroll
    for (int i = 0; i < MW; ++i) {
      const int mt = wave + 4 * i;
      if (mt < MT) {
        const bf16x8 a = *(const bf16x8*)(sx + (16 * mt + (lane & 15)) * kLd + kq);
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv[t], acc[i][t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < MW; ++i) {
    const int mt = wave + 4 * i;
    if (mt >= MT) continue;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int col = n0 + 16 * t + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * mt + 4 * (lane >> 4) + r;
        part[((size_t)ch * KT + row) * N + col] = acc[i][t][r];
      }
    }
  }
}

// Weight-gradient partials of a BF16 product in output tiles of TM x TN, one block
// per (row chunk, tile): part[ch][m][n] = sum_{b in chunk} X'[b][m] dY[b][n].
// Plain form: X' = X [rows x Kin]; the blocks of m-tile 0 also multiply their dY
// operands by an all-ones operand, which gives row m = Kin, the bias gradient
// sum_b dY[b][n]. Gather form (the input layer, TM = 16): X'[b][m] = u[b][A[m]]
// for m < dA, 1 for m = dA (the bias row), 0 above.
// Each step of 64 rows stages the X' slab [64 x TM] and the dY slab [64 x TN] in
// LDS column-major (a column's 64 rows contiguous), so that every MFMA operand
// is one 16-byte LDS read: a thread loads an 8 x 8 block as 8 row pieces of 16
// bytes, transposes it in registers and stores 8 column pieces; the next step's
// blocks are loaded while the current step multiplies. Wave (wm, wn) computes
// MI x NI MFMA tiles of 16 x 16. The hardware hands consecutive block ids to the
// 8 XCDs in turn; the blocks are renumbered so that the tiles of one row chunk
// run together on one XCD and share its L2.
template <int MI, int NI, int WM, int WN, bool GATHER>
__global__ __launch_bounds__(256, 2) void k_wgrad_tile(
    const __bf16* __restrict__ X, int ldx, const float* __restrict__ ug, int D, LayerIdx L,
    const __bf16* __restrict__ Y, int ldy, int Kin, int N, int nrows, int chunk, int tiles_m,
    int ntile, int nblk, int kin_t, float* __restrict__ part) {
  constexpr int TM = 16 * MI * WM, TN = 16 * NI * WN, kLd = 72;  // kLd: 64 rows + 8
  constexpr int NB = NI / WM;  // n-tiles of the bias row per wave
  static_assert(WM * WN == 4 && TN <= 256 && (GATHER ? TM == 16 : TM + TN <= 256 && NB * WM == NI),
                "tile shape");
  __shared__ __attribute__((aligned(16))) __bf16 sx[TM * kLd];
  __shared__ __attribute__((aligned(16))) __bf16 sy[TN * kLd];
  // the grid is a multiple of 8 blocks; block b runs on XCD b % 8
  const int per = (int)(gridDim.x / 8);
  const int lb = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (lb >= nblk) return;  // block-uniform
  const int tile = lb % ntile, ch = lb / ntile;
  const int m0 = (tile % tiles_m) * TM, n0 = (tile / tiles_m) * TN;
  const int b0 = ch * chunk, b1 = min(nrows, b0 + chunk);
  // the wave index in a scalar register: wave-uniform branches, no per-lane selects
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int wm = wave % WM, wn = wave / WM;
  // staging roles. Plain form: 8 x 8 block q of the X slab (q < TM) or of the dY
  // slab. Gather form: every thread gathers 4 values of the X' slab (row q & 63,
  // columns 4 (q >> 6) ..), and threads q < TN load block q of the dY slab.
  const int q = threadIdx.x;
  const bool stx = !GATHER && q < TM;
  const bool sty = GATHER ? q < TN : !stx && q < TM + TN;
  const int qq = (GATHER || stx) ? q : q - TM;
  const int rb = qq & 7, cb = qq >> 3;  // rows 8 rb .. 8 rb + 7, columns 8 cb .. 8 cb + 7
  bf16x8 v[8];
  float gx[4];
  auto load = [&](int bb) {
    if constexpr (GATHER) {
      const int b = bb + (q & 63);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = 4 * (q >> 6) + e;
        float x = 0.f;
        if (b < b1) x = m < L.dA ? ug[(size_t)b * D + L.A[m]] : (m == L.dA ? 1.f : 0.f);
        gx[e] = x;
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int b = bb + 8 * rb + j;
      bf16x8 z;
#pragma unroll
      for (int e = 0; e < 8; ++e) z[e] = (__bf16)0.f;
      if (stx) {
        const int c = m0 + 8 * cb;
        if (b < b1 && c < Kin) z = *(const bf16x8*)(X + (size_t)b * ldx + c);
      } else if (sty) {
        const int c = n0 + 8 * cb;
        if (b < b1 && c < N) z = *(const bf16x8*)(Y + (size_t)b * ldy + c);
      }
      v[j] = z;
    }
  };
  auto store = [&]() {
    if constexpr (GATHER) {
#pragma unroll
      for (int e = 0; e < 4; ++e) sx[(4 * (q >> 6) + e) * kLd + (q & 63)] = (__bf16)gx[e];
    }
    if (!stx && !sty) return;
    __bf16* s = stx ? sx : sy;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      bf16x8 o;
#pragma unroll
      for (int j = 0; j < 8; ++j) o[j] = v[j][e];
      *(bf16x8*)(s + (8 * cb + e) * kLd + 8 * rb) = o;
    }
  };
  f32x4 acc[MI][NI];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int t = 0; t < NI; ++t) acc[i][t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  f32x4 accb[NB > 0 ? NB : 1];
#pragma unroll
  for (int u = 0; u < (NB > 0 ? NB : 1); ++u) accb[u] = (f32x4){0.f, 0.f, 0.f, 0.f};
  bf16x8 ones;
#pragma unroll
  for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.f;
  const int mrows = GATHER ? L.dA + 1 : Kin;
  const bool active = m0 + 16 * MI * wm < mrows && n0 + 16 * NI * wn < N;
  const bool bias = !GATHER && m0 == 0;
  const int kq = 8 * (lane >> 4), r16 = lane & 15;
  load(b0);
  for (int bb = b0; bb < b1; bb += 64) {
    store();
    __syncthreads();
    if (bb + 64 < b1) load(bb + 64);
    if (active) {
#pragma unroll
      for (int ks = 0; ks < 64; ks += 32) {
        bf16x8 a[MI], bv[NI];
#pragma unroll
        for (int i = 0; i < MI; ++i)
          a[i] = *(const bf16x8*)(sx + (16 * (MI * wm + i) + r16) * kLd + ks + kq);
#pragma unroll
        for (int t = 0; t < NI; ++t)
          bv[t] = *(const bf16x8*)(sy + (16 * (NI * wn + t) + r16) * kLd + ks + kq);
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
          for (int t = 0; t < NI; ++t)
            acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bv[t], acc[i][t], 0, 0, 0);
        if constexpr (!GATHER) {
          // n-tiles NB wm .. NB wm + NB - 1 of the wave's dY operands (constant
          // register indices in each branch)
          if (bias) {
#pragma unroll
            for (int w = 0; w < WM; ++w)
              if (wm == w) {
#pragma unroll
                for (int u = 0; u < NB; ++u)
                  accb[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ones, bv[NB * w + u], accb[u],
                                                                    0, 0, 0);
              }
          }
        }
      }
    }
    __syncthreads();
  }
  if (!active) return;
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int t = 0; t < NI; ++t) {
      const int col = n0 + 16 * (NI * wn + t) + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + 16 * (MI * wm + i) + 4 * (lane >> 4) + r;
        if (row < mrows) part[((size_t)ch * kin_t + row) * N + col] = acc[i][t][r];
      }
    }
  if constexpr (!GATHER) {
    // row 0 of the all-ones product holds the column sums
    if (bias && (lane >> 4) == 0) {
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int col = n0 + 16 * (NI * wn + NB * wm + u) + r16;
        part[((size_t)ch * kin_t + Kin) * N + col] = accb[u][0];
      }
    }
  }
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; refreshes the bf16 copies
__global__ void k_adam(float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, const unsigned char* __restrict__ frozen) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || frozen[i]) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// W[Kin x N] (row-major, fp32) -> Wt16[N x Kin] and W16[Kin x N] (bf16)
__global__ void k_weights_bf16(const float* __restrict__ W, int Kin, int N, __bf16* __restrict__ Wt16,
                               __bf16* __restrict__ W16, float* __restrict__ Wt32) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= Kin * N) return;
  const int k = idx / N, c = idx % N;
  const float w = W[idx];
  Wt16[(size_t)c * Kin + k] = (__bf16)w;
  W16[idx] = (__bf16)w;
  Wt32[(size_t)c * Kin + k] = w;
}

// z ~ U(0,1)^D from Philox (counter = (block, idx, 0, step) under key = seed)
__global__ void k_flow_z(uint64_t seed, uint32_t step, uint64_t first, int n, int D,
                         float* __restrict__ z) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  double u[GRA_MAX_DIM];
  point_uniforms(seed, 0x80000000u | step, first + b, D, u);
  for (int j = 0; j < D; ++j) z[(size_t)b * D + j] = fminf((float)u[j], 0.99999994f);
}

// u (flow output, FP32) -> x through the VEGAS grid (FP64), vegas weight
__global__ void k_u_to_x(const float* __restrict__ u, const double* __restrict__ grid, int D,
                         unsigned int bins, int n, double* __restrict__ x, double* __restrict__ vw) {
  __shared__ double xm[GRA_MAX_DIM * GRA_VEGAS_BINS];
  for (int k = threadIdx.x; k < D * GRA_VEGAS_BINS; k += blockDim.x) xm[k] = grid[k];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  double uu[GRA_MAX_DIM], xx[GRA_MAX_DIM];
  uint8_t bb[GRA_MAX_DIM];
  for (int j = 0; j < D; ++j) uu[j] = (double)u[(size_t)b * D + j];
  vw[b] = vegas_map(xm, D, bins, uu, xx, bb);
  for (int j = 0; j < D; ++j) x[(size_t)b * D + j] = xx[j];
}

// w_b = f_b vw_b / q_b in FP64 and per-block partial sums (sum w, sum w^2, max w)
__global__ void k_weights(const double* __restrict__ W, const double* __restrict__ vw,
                          const float* __restrict__ logq, int n, double* __restrict__ w,
                          double* __restrict__ part) {
  __shared__ double s1[256], s2[256], smx[256];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  double wv = 0.0;
  if (b < n) {
    wv = W[b] * vw[b] * exp(-(double)logq[b]);
    if (!(wv >= 0.0) || isinf(wv)) wv = 0.0;
    w[b] = wv;
  }
  s1[threadIdx.x] = wv;
  s2[threadIdx.x] = wv * wv;
  smx[threadIdx.x] = wv;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
      smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s1[0];
    part[3 * blockIdx.x + 1] = s2[0];
    part[3 * blockIdx.x + 2] = smx[0];
  }
}

__global__ void k_coef(const double* __restrict__ w, int n, double inv_s2, float* __restrict__ c) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < n) c[b] = (float)(w[b] * w[b] * inv_s2);
}

// ---------------------------------------------------------------------------
// Fused coupling layer, sampling direction (BF16 MFMA MLP): one block of 256
// threads takes R points through the whole layer. The conditioner MLP runs
// out of LDS (input layer on the VALU in FP32, hidden and output products on
// v_mfma_f32_16x16x32_bf16 with the transposed BF16 weights read through L2),
// the spline parameters stay in LDS, and each thread then inverts one
// (point, transformed dimension) spline with the knots accumulated in order.
// Only u_{l-1}, log q and, when the layer's activations are needed for the
// backward pass, the activations and spline parameters are written to HBM.
// LDS holds two activation tiles; the FP32 parameter tile of the output
// product overlays the one that product does not read. R = 64 rows per block
// (the output product's B fragments, read through L2, each feed R / 16 MFMA row
// tiles; at H = 256 it is two thirds of the layer's time). At H = 256 one
// block of 512 threads per CU: 216 VGPRs without spills; 32 rows at two blocks
// per CU (128 VGPRs, 108-byte spill) sampled 10% slower, 128 rows spill 448 bytes.
// ---------------------------------------------------------------------------
#ifndef GRA_FUSED_ROWS_256
#define GRA_FUSED_ROWS_256 64
#endif
#ifndef GRA_FUSED_THREADS_256
#define GRA_FUSED_THREADS_256 512
#endif
#ifndef GRA_FUSED_THREADS
#define GRA_FUSED_THREADS 256
#endif
// Below H = 256 a block takes 64 rows: each B fragment read through L2 then
// feeds four MFMA row tiles instead of two (+10% sampling rate at H = 128
// against 32 rows; 128 rows is slower, at one block per CU)
#ifndef GRA_FUSED_ROWS
#define GRA_FUSED_ROWS 64
#endif
__host__ __device__ constexpr int fused_rows(int H) {
  return H == 256 ? GRA_FUSED_ROWS_256 : GRA_FUSED_ROWS;
}
// Threads per block: more waves share the block's LDS tiles and split its
// n-tiles, so each SIMD has several waves to switch between while the B
// fragments are in flight (one wave per SIMD leaves it waiting most of the time)
__host__ __device__ constexpr int fused_threads(int H) {
  return H == 256 ? GRA_FUSED_THREADS_256 : GRA_FUSED_THREADS;
}
// Blocks per CU the register allocation must allow: one at H = 256; three
// below (168 VGPRs, a 40-byte spill outside the products), with the LDS tile
// kept to a third of the CU (fused_lds_target)
#ifndef GRA_FUSED_BLOCKS
#define GRA_FUSED_BLOCKS 3
#endif
#ifndef GRA_FUSED_BLOCKS_256
#define GRA_FUSED_BLOCKS_256 1
#endif
__host__ __device__ constexpr int fused_blocks(int H) {
  return H == 256 ? GRA_FUSED_BLOCKS_256 : GRA_FUSED_BLOCKS;
}
__host__ __device__ constexpr int fused_waves(int H) {
  return fused_blocks(H) * fused_threads(H) / 256;
}

struct FusedLayer {
  const float* W1;            // [dA x H]
  const float* b1;            // [H]
  const __bf16* wt[4];        // transposed weights of products 1..depth: [N_i x H]
  const float* bias[4];       // biases of products 1..depth
  __bf16* act_out[4];         // optional stores of activations 0..depth-1 ([n x H])
  float* prm_out;             // optional store of the spline parameters ([n x ldp])
};

// One product of the MLP on the block's R rows: out = A[R x H] Wt^T,
// with bias (+ ReLU into the next BF16 activation buffer, or into the FP32
// parameter tile for the output product)
// TT n-tiles nt0, nt0 + NW, ... of one product on the block's R rows: every
// load is unconditional, so the compiler keeps the B fragments of a k-step in
// flight together instead of waiting for each before its MFMA
template <int H, int R, int NW, int TT>
__device__ __forceinline__ void product_tiles(const __bf16* sA, int lda,
                                              const __bf16* __restrict__ wt, int nt0,
                                              f32x4 (&acc)[R / 16][4]) {
  constexpr int MT = R / 16;
  const int lane = threadIdx.x & 63;
  const int kq = 8 * (lane >> 4);
#pragma unroll
  for (int k0 = 0; k0 < H; k0 += 32) {
    bf16x8 a[MT], b[TT];
#pragma unroll
    for (int t = 0; t < TT; ++t)
      b[t] = *(const bf16x8*)(wt + (size_t)((nt0 + NW * t) * 16 + (lane & 15)) * H + k0 + kq);
#pragma unroll
    for (int m = 0; m < MT; ++m)
      a[m] = *(const bf16x8*)(sA + (size_t)(16 * m + (lane & 15)) * lda + k0 + kq);
#pragma unroll
    for (int t = 0; t < TT; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[t], acc[m][t], 0, 0, 0);
  }
}

template <int H, int R, int NW>
__device__ __forceinline__ void fused_product(const __bf16* sA, int lda, const __bf16* __restrict__ wt,
                                              const float* __restrict__ bias, int tb, int N,
                                              int nvalid, bool out_layer, __bf16* sOut, int ldo,
                                              float* sP, int ldpp, __bf16* gact, float* gprm,
                                              int row0, int n, int ldp) {
  // output n-tiles tb .. tb + N/16 - 1; tile tb is column 0 of sP / sOut
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ntiles = N / 16;
  wt += (size_t)tb * 16 * H;
  for (int nt0 = wave; nt0 < ntiles; nt0 += 4 * NW) {
    // each of the NW waves owns n-tiles nt0 + NW t (t < 4) of this chunk; nv of
    // them exist (wave-uniform)
    constexpr int MT = R / 16;
    const int nv = min(4, (ntiles - nt0 + NW - 1) / NW);
    f32x4 acc[MT][4];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[m][t] = (f32x4){0.f, 0.f, 0.f, 0.f};
    if (nv == 4)
      product_tiles<H, R, NW, 4>(sA, lda, wt, nt0, acc);
    else if (nv == 3)
      product_tiles<H, R, NW, 3>(sA, lda, wt, nt0, acc);
    else if (nv == 2)
      product_tiles<H, R, NW, 2>(sA, lda, wt, nt0, acc);
    else
      product_tiles<H, R, NW, 1>(sA, lda, wt, nt0, acc);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t >= nv) break;
      const int nt = nt0 + NW * t;
      const int cl = nt * 16 + (lane & 15), col = tb * 16 + cl;
      const float bv = bias[col];
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * m + 4 * (lane >> 4) + r;
          float v = acc[m][t][r] + bv;
          if (out_layer) {
            if (col >= nvalid) v = 0.f;
            sP[row * ldpp + cl] = v;
            if (gprm && row0 + row < n) gprm[(size_t)(row0 + row) * ldp + col] = v;
          } else {
            v = fmaxf(v, 0.f);
            const __bf16 hv = (__bf16)v;
            sOut[row * ldo + cl] = hv;
            if (gact && row0 + row < n) gact[(size_t)(row0 + row) * H + col] = hv;
          }
        }
    }
  }
}

// Bin of the inverse spline (sampling direction) for one (point, dimension):
// the softmax of the width and height logits p[0..K), p[K..2K), and the first
// bin k whose cumulative height passes z (the last bin otherwise). KT > 0 is a
// compile-time K: the logits are read once into registers and the loops are
// unrolled; KT = 0 loops over a runtime K. Both orders of operations are the
// same, so both give the same bin and values.
template <int KT>
__device__ __forceinline__ void inverse_bin(const float* p, int K, float zv, Bin& e) {
  float X = 0.f, Y = 0.f, w = 0.f, h = 0.f;
  int k = 0;
  if constexpr (KT > 0) {
    // heights in registers: their softmax and the bin search; the widths are
    // then walked from LDS (two exponentials each) to keep the registers low
    float eh[KT];
#pragma unroll
    for (int m = 0; m < KT; ++m) eh[m] = p[KT + m];
    float mw = p[0], mh = eh[0];
#pragma unroll
    for (int m = 1; m < KT; ++m) {
      mw = fmaxf(mw, p[m]);
      mh = fmaxf(mh, eh[m]);
    }
    float sw = 0.f, sh = 0.f;
#pragma unroll
    for (int m = 0; m < KT; ++m) {
      sw += __expf(p[m] - mw);
      eh[m] = __expf(eh[m] - mh);
      sh += eh[m];
    }
    const float cw = 1.f - KT * kEpsW;
    const float iw = cw / sw, ih = cw / sh;
    float Ya = 0.f;
    bool found = false;
#pragma unroll
    for (int m = 0; m < KT; ++m) {
      const float hm = kEpsW + eh[m] * ih;
      const bool hit = !found && (m == KT - 1 || Ya + hm > zv);
      k = hit ? m : k;
      Y = hit ? Ya : Y;
      h = hit ? hm : h;
      found = found || hit;
      Ya += hm;
    }
#pragma unroll
    for (int m = 0; m < KT; ++m) {
      const float wm = kEpsW + __expf(p[m] - mw) * iw;
      w = (m == k) ? wm : w;
      X = (m < k) ? X + wm : X;
    }
  } else {
    float mw = p[0], mh = p[K];
    for (int m = 1; m < K; ++m) {
      mw = fmaxf(mw, p[m]);
      mh = fmaxf(mh, p[K + m]);
    }
    float sw = 0.f, sh = 0.f;
    for (int m = 0; m < K; ++m) {
      sw += __expf(p[m] - mw);
      sh += __expf(p[K + m] - mh);
    }
    const float cw = 1.f - K * kEpsW;
    const float iw = cw / sw, ih = cw / sh;
    for (; k < K; ++k) {
      w = kEpsW + __expf(p[k] - mw) * iw;
      h = kEpsW + __expf(p[K + k] - mh) * ih;
      if (k == K - 1 || Y + h > zv) break;
      X += w;
      Y += h;
    }
  }
  e.k = k;
  e.X = X;
  e.w = w;
  e.Y = Y;
  e.h = h;
}

template <int H, int KT>
__global__ __launch_bounds__(fused_threads(H), fused_waves(H)) void k_flow_layer_inverse(
    const float* __restrict__ u_in, float* __restrict__ u_out, float* __restrict__ ld, int n,
    int D, LayerIdx L, int Kin, int depth, int ldp, int gd, int ldg, FusedLayer W) {
  const int K = KT > 0 ? KT : Kin;
  constexpr int R = fused_rows(H);
  constexpr int NT = fused_threads(H);
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  constexpr int lda = H + 8;  // BF16 row stride (16-byte multiple, staggers banks)
  const int ldpp = ldg;  // parameter tile row stride (one group of gd dimensions)
  // X: activation tile; Y: activation tile, overlaid by the parameter tile
  __bf16* sX = (__bf16*)smem;
  __bf16* sY = sX + R * lda;
  float* sP = (float*)sY;
  const int ybytes = max(R * lda * (int)sizeof(__bf16), R * ldpp * (int)sizeof(float));
  float* sLj = (float*)((unsigned char*)sY + ybytes);
  float* sU = sLj + R * 8;  // the block's input rows [R x D]
  const int row0 = blockIdx.x * R;
  for (int e = threadIdx.x; e < R * D; e += blockDim.x) {
    const int row = row0 + e / D;
    sU[e] = row < n ? u_in[(size_t)row0 * D + e] : 0.f;
  }
  __syncthreads();
  // activation i sits in X when depth - 1 - i is even, so the output product
  // reads X and writes the parameter tile over Y
  __bf16* cur = (depth & 1) ? sX : sY;
  __bf16* nxt = (depth & 1) ? sY : sX;
  // input layer (FP32 VALU): act0 = relu(u[A] W1 + b1); each thread owns one
  // column, with its weights in registers, for R / (NT / H) rows
  {
    constexpr int nrg = NT / H;
    const int j = threadIdx.x % H, rg = threadIdx.x / H;
    // indices past dA read column 0 with weight 0, so no load is conditional
    float w[kMaxDB];
    int ai[kMaxDB];
#pragma unroll
    for (int i = 0; i < kMaxDB; ++i) {
      const bool on = i < L.dA;
      w[i] = on ? W.W1[(on ? i : 0) * H + j] : 0.f;
      ai[i] = on ? L.A[i] : 0;
    }
    const float bj = W.b1[j];
    for (int r = rg; r < R; r += nrg) {
      float acc = bj;
#pragma unroll
      for (int i = 0; i < kMaxDB; ++i) acc += sU[r * D + ai[i]] * w[i];
      const __bf16 hv = (__bf16)fmaxf(acc, 0.f);
      cur[r * lda + j] = hv;
      if (W.act_out[0] && row0 + r < n) W.act_out[0][(size_t)(row0 + r) * H + j] = hv;
    }
  }
  __syncthreads();
  const int nps = prm_stride(K);
  for (int i = 1; i < depth; ++i) {
    fused_product<H, R, NT / 64>(cur, lda, W.wt[i - 1], W.bias[i - 1], 0, H, H, false, nxt, lda,
                                 sP, ldpp, W.act_out[i], nullptr, row0, n, ldp);
    __syncthreads();
    __bf16* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  // output product and spline inverses by groups of gd transformed dimensions:
  // a group's columns [j0 nps, j1 nps) span the n-tiles [tb, te); thread -> (row,
  // dimension of the group)
  const int t = threadIdx.x;
  for (int j0 = 0; j0 < L.dB; j0 += gd) {
    const int j1 = min(L.dB, j0 + gd), g = j1 - j0;
    const int tb = (j0 * nps) / 16, te = (j1 * nps + 15) / 16;
    fused_product<H, R, NT / 64>(cur, lda, W.wt[depth - 1], W.bias[depth - 1], tb, 16 * (te - tb),
                                 L.dB * nps, true, nxt, lda, sP, ldpp, nullptr, W.prm_out, row0, n,
                                 ldp);
    __syncthreads();
    // (row, dimension) pairs of the group; R g may exceed the block's threads
    for (int q = t; q < R * g; q += NT) {
      const int r = q / g, j = j0 + q % g;
      const int row = row0 + r;
      if (row >= n) continue;
      if (j == 0)
        for (int i = 0; i < L.dA; ++i) u_out[(size_t)row * D + L.A[i]] = sU[r * D + L.A[i]];
      const float* p = sP + r * ldpp + j * nps - 16 * tb;
      const float zv = fminf(fmaxf(sU[r * D + L.B[j]], 0.f), 1.f);
      Bin e;
      inverse_bin<KT>(p, K, zv, e);
      const int k = e.k;
      e.dk = (k == 0) ? 1.f : kEpsD + softplusf(p[2 * K + k - 1]);
      e.dk1 = (k == K - 1) ? 1.f : kEpsD + softplusf(p[2 * K + k]);
      e.s = e.h / e.w;
      const float zeta = zv - e.Y;
      const float Dl = e.dk1 + e.dk - 2.f * e.s;
      const float qa = e.h * (e.s - e.dk) + zeta * Dl;
      const float qb = e.h * e.dk - zeta * Dl;
      const float qc = -e.s * zeta;
      const float disc = fmaxf(qb * qb - 4.f * qa * qc, 0.f);
      float xi = (2.f * qc) / (-qb - sqrtf(disc));
      xi = fminf(fmaxf(xi, 0.f), 1.f);
      float zz, lj;
      rqs_value(e, xi, zz, lj);
      u_out[(size_t)row * D + L.B[j]] = fminf(fmaxf(e.X + xi * e.w, 0.f), 0.99999994f);
      sLj[r * 8 + j] = lj;
    }
    __syncthreads();
  }
  for (int r = t; r < R; r += NT) {
    if (row0 + r >= n) break;
    float acc = 0.f;
    for (int jj = 0; jj < L.dB; ++jj) acc += sLj[r * 8 + jj];
    ld[row0 + r] += acc;
  }
}

}  // namespace

// one spline per group of G >= K lanes (G = 8, 16, 32)
#define LAUNCH_SPLINE(KERNEL, K, n, ...)                                                        \
  do {                                                                                          \
    const int g_ = (K) <= 8 ? 8 : ((K) <= 16 ? 16 : 32);                                         \
    const dim3 grd_((unsigned)(((size_t)(n) * g_ + 255) / 256));                                \
    if (g_ == 8) hipLaunchKernelGGL(KERNEL<8>, grd_, dim3(256), 0, h->stream, __VA_ARGS__);      \
    else if (g_ == 16) hipLaunchKernelGGL(KERNEL<16>, grd_, dim3(256), 0, h->stream, __VA_ARGS__); \
    else hipLaunchKernelGGL(KERNEL<32>, grd_, dim3(256), 0, h->stream, __VA_ARGS__);             \
  } while (0)

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace gra_hip {
struct FlowState;
static FlowState* flow_of(gra_handle* h);

struct FlowLayerHost {
  LayerIdx idx;
  int gd = 1, ldg = 1;  // fused sampling layer: dimensions per output group, tile stride
  size_t lds = 0;       // and its LDS bytes
  size_t oW1, ob1, oWo, obo;
  size_t oWh[4], obh[4];
};

struct FlowState {
  gra_flow_cfg fc;
  int D = 0, L = 0, K = 0, H = 0, depth = 0, np = 0, nps = 0, ldp = 0;
  int mlp_fp32 = 0;
  std::vector<FlowLayerHost> layers;
  size_t nparam = 0;
  DevBuf theta, grad, adam_m, adam_v, frozen;
  DevBuf wt16, w16, wt32;  // per layer, per GEMM weight (hidden and output)
  std::vector<size_t> wofs;  // element offsets of each GEMM weight copy
  // batch buffers
  int cap = 0;
  DevBuf z, us, ld, acts16, acts32, prm, gprm, gprm16, dact16, dact32, gu0, gu1, coef, xx, vw, W,
      wts, part, tmp, dh32, grid;
  int step = 0;
  std::vector<double> hpart;
};

void flow_release(FlowState* f) {
  if (!f) return;
  DevBuf* b[] = {&f->theta, &f->grad,   &f->adam_m, &f->adam_v, &f->frozen, &f->wt16, &f->w16,
                 &f->wt32,  &f->z,      &f->us,     &f->ld,     &f->acts16, &f->acts32, &f->prm,
                 &f->gprm,  &f->gprm16, &f->dact16, &f->dact32, &f->gu0,    &f->gu1,   &f->coef,
                 &f->xx,    &f->vw,     &f->W,      &f->wts,    &f->part,   &f->tmp,   &f->dh32,
                 &f->grid};
  for (DevBuf* x : b) x->release();
  delete f;
}

static FlowState* flow_of(gra_handle* h) { return h ? h->flow : nullptr; }

static LayerIdx make_mask(int l, int D) {
  LayerIdx m;
  std::memset(&m, 0, sizeof(m));
  const int pat = l % 4;
  for (int i = 0; i < D; ++i) {
    bool inA;
    if (pat == 0) inA = (i % 2) == 0;
    else if (pat == 1) inA = (i % 2) == 1;
    else if (pat == 2) inA = i < D / 2;
    else inA = i >= D / 2;
    if (inA) m.A[m.dA++] = i;
    else m.B[m.dB++] = i;
  }
  return m;
}

static uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Rows per block of k_wgrad: about 1024 blocks for n rows and N output columns,
// between 512 and 4096 rows, a multiple of 32
static int wgrad_chunk(int n, int N) {
  const long c = (long)n * (N / 64) / 1024;
  return (int)std::max(512L, std::min(4096L, (c / 32) * 32));
}
// m-tiles per wave of k_wgrad (0: no instance for this width)
static int wgrad_mw(int Kin) { return Kin == 64 ? 2 : Kin == 128 ? 3 : 0; }
// Rows per block of k_wgrad_tile: about 1024 blocks over ntile output tiles, a
// multiple of 64 between 512 and 8192
static int wgt_chunk(int n, int ntile) {
  const long c = (long)n * ntile / 1024;
  return (int)std::max(512L, std::min(8192L, (c / 64) * 64));
}
// Output tiles of k_wgrad_tile: 128 x 128 (plain), 16 x 128 (gather)
static int wgt_tiles(int Kin, int N, bool gather) {
  return (gather ? 1 : (Kin + 127) / 128) * ((N + 127) / 128);
}
// Floats of k_wgrad_tile's partials for n rows
static size_t wgt_part_floats(int n, int Kin, int N, bool gather) {
  const int chunk = wgt_chunk(n, wgt_tiles(Kin, N, gather));
  return (size_t)((n + chunk - 1) / chunk) * (gather ? 16 : Kin + 16) * N;
}

static int flow_alloc_batch(gra_handle* h, FlowState* F, int n) {
  if (n <= F->cap) return GRA_OK;
  const int D = F->D, L = F->L, H = F->H, dep = F->depth, ldp = F->ldp;
  const size_t N = (size_t)n;
  bool bad = false;
  bad |= F->z.ensure(sizeof(float) * N * D) != 0;
  bad |= F->us.ensure(sizeof(float) * N * D * (L + 1)) != 0;
  bad |= F->ld.ensure(sizeof(float) * N) != 0;
  bad |= F->acts16.ensure(sizeof(__bf16) * N * H * dep * L) != 0;
  bad |= F->acts32.ensure(F->mlp_fp32 ? sizeof(float) * N * H * dep * L : 16) != 0;
  bad |= F->prm.ensure(sizeof(float) * N * ldp * L) != 0;
  bad |= F->gprm.ensure(sizeof(float) * N * ldp) != 0;
  bad |= F->gprm16.ensure(sizeof(__bf16) * N * ldp) != 0;
  bad |= F->dact16.ensure(sizeof(__bf16) * N * H * 2) != 0;
  bad |= F->dact32.ensure(F->mlp_fp32 ? sizeof(float) * N * H * 2 : 16) != 0;
  bad |= F->gu0.ensure(sizeof(float) * N * D) != 0;
  bad |= F->gu1.ensure(sizeof(float) * N * D) != 0;
  bad |= F->coef.ensure(sizeof(float) * N) != 0;
  bad |= F->xx.ensure(sizeof(double) * N * D) != 0;
  bad |= F->vw.ensure(sizeof(double) * N) != 0;
  bad |= F->W.ensure(sizeof(double) * N) != 0;
  bad |= F->wts.ensure(sizeof(double) * N) != 0;
  const size_t nch = (N + 2047) / 2048;
  const size_t big = (size_t)(std::max(H, ldp) + 64) * std::max(H, ldp);
  // k_wgrad's partials: ceil(N / chunk) x (H + 16) x columns, for both product widths
  size_t wg = 0;
  for (int cols : {H, ldp})
    wg = std::max(wg, (size_t)((N + wgrad_chunk(N, cols) - 1) / wgrad_chunk(N, cols)) *
                          (size_t)(H + 16) * cols);
  // k_wgrad_tile's partials: hidden and output products, and the input layer
  if (H % 128 == 0) {
    wg = std::max(wg, wgt_part_floats(n, H, H, false));
    wg = std::max(wg, wgt_part_floats(n, H, ldp, false));
  }
  wg = std::max(wg, wgt_part_floats(n, 16, H, true));
  bad |= F->part.ensure(sizeof(float) * std::max(std::max(nch * big, wg),
                                                 (size_t)3 * ((N + 255) / 256) * 2)) != 0;
  bad |= F->tmp.ensure(sizeof(double) * 3 * ((N + 255) / 256) + 64) != 0;
  if (bad) return fail(h, GRA_E_NOMEM, "flow batch buffers for %d points", n);
  F->cap = n;
  return GRA_OK;
}

static float* prm_of(FlowState* F, int l) { return F->prm.as<float>() + (size_t)l * F->cap * F->ldp; }
static float* u_of(FlowState* F, int l) { return F->us.as<float>() + (size_t)l * F->cap * F->D; }
static __bf16* act16_of(FlowState* F, int l, int i) {
  return F->acts16.as<__bf16>() + ((size_t)l * F->depth + i) * F->cap * F->H;
}
static float* act32_of(FlowState* F, int l, int i) {
  return F->acts32.as<float>() + ((size_t)l * F->depth + i) * F->cap * F->H;
}

// bf16 / fp32 copies of the GEMM weights (hidden layers and output) of every layer
static int refresh_weights(gra_handle* h, FlowState* F) {
  int slot = 0;
  for (int l = 0; l < F->L; ++l) {
    const FlowLayerHost& Ly = F->layers[l];
    for (int i = 1; i <= F->depth; ++i) {
      const bool out = (i == F->depth);
      const size_t off = out ? Ly.oWo : Ly.oWh[i];
      const int N = out ? F->ldp : F->H;
      const int n = F->H * N;
      hipLaunchKernelGGL(k_weights_bf16, dim3((n + 255) / 256), dim3(256), 0, h->stream,
                         F->theta.as<float>() + off, F->H, N, F->wt16.as<__bf16>() + F->wofs[slot],
                         F->w16.as<__bf16>() + F->wofs[slot], F->wt32.as<float>() + F->wofs[slot]);
      ++slot;
    }
  }
  GRA_HIP_CHECK(h, hipGetLastError());
  return GRA_OK;
}

static void gemm_nt(gra_handle* h, FlowState* F, const void* A, int lda, const void* Bt, int ldb,
                    int M, int N, int K, const Epi& ep) {
  if (!F->mlp_fp32 && (ep.ldc & 7) == 0 && (ep.ldm & 7) == 0 && M % 128 == 0 && K % 32 == 0 &&
      (N == 64 || N == 128 || N == 256)) {
    const __bf16* Ab = (const __bf16*)A;
    const __bf16* Bb = (const __bf16*)Bt;
    if (N == 64)
      hipLaunchKernelGGL(k_gemm_nt_wide<4>, dim3(M / 128), dim3(256), 0, h->stream, Ab, lda, Bb,
                         ldb, K, ep);
    else if (N == 128)
      hipLaunchKernelGGL(k_gemm_nt_wide<8>, dim3(M / 128), dim3(256), 0, h->stream, Ab, lda, Bb,
                         ldb, K, ep);
    else
      hipLaunchKernelGGL(k_gemm_nt_wide<16>, dim3(M / 128), dim3(256), 0, h->stream, Ab, lda, Bb,
                         ldb, K, ep);
    return;
  }
  const dim3 grid(M / 64, N / 64);
  if (F->mlp_fp32)
    hipLaunchKernelGGL(k_gemm_nt<false>, grid, dim3(256), 0, h->stream, A, lda, Bt, ldb, M, N, K, ep);
  else
    hipLaunchKernelGGL(k_gemm_nt<true>, grid, dim3(256), 0, h->stream, A, lda, Bt, ldb, M, N, K, ep);
}

// dW[Kin x N] = X^T dY and the bias gradient db[N] = sum_b dY[b], accumulated
// deterministically (per-chunk partials, then a fixed-order sum)
// k_wgrad_tile over n rows, then the fixed-order sum of its partials into gW (kw
// rows) and gb; gather: the input layer (X' from u through L)
static void wgrad_tile(gra_handle* h, FlowState* F, const __bf16* X, int ldx, const float* ug,
                       const LayerIdx& L, const __bf16* Y, int ldy, int Kin, int N, int n,
                       bool gather, float* gW, float* gb) {
  const int tiles_m = gather ? 1 : (Kin + 127) / 128;
  const int ntile = wgt_tiles(Kin, N, gather);
  // chunks sized for the allocated batch, so the partials fit for every n <= cap
  const int chunk = wgt_chunk(F->cap, ntile);
  const int nch = (n + chunk - 1) / chunk;
  const int nblk = nch * ntile;
  const int kin_t = gather ? 16 : Kin + 16;
  const unsigned grid = (unsigned)((nblk + 7) / 8 * 8);
  float* part = F->part.as<float>();
  if (gather)
    hipLaunchKernelGGL((k_wgrad_tile<1, 2, 1, 4, true>), dim3(grid), dim3(256), 0, h->stream,
                       nullptr, 0, ug, F->D, L, Y, ldy, Kin, N, n, chunk, tiles_m, ntile, nblk,
                       kin_t, part);
  else
    hipLaunchKernelGGL((k_wgrad_tile<4, 4, 2, 2, false>), dim3(grid), dim3(256), 0, h->stream, X,
                       ldx, nullptr, 0, L, Y, ldy, Kin, N, n, chunk, tiles_m, ntile, nblk, kin_t,
                       part);
  const int kw = gather ? L.dA : Kin;
  const int tot = (kw + 1) * N;
  hipLaunchKernelGGL(k_part_reduce, dim3((tot + 31) / 32), dim3(256), 0, h->stream, part, nch,
                     kin_t, N, kw, gW, gb);
}

static void weight_grad(gra_handle* h, FlowState* F, const void* X, int ldx, const void* Y, int ldy,
                        int Kin, int N, int n, float* gW, float* gb) {
  if (!F->mlp_fp32 && !F->fc.wgrad_split && Kin % 128 == 0) {
    LayerIdx none;
    std::memset(&none, 0, sizeof(none));
    wgrad_tile(h, F, (const __bf16*)X, ldx, nullptr, none, (const __bf16*)Y, ldy, Kin, N, n, false,
               gW, gb);
    return;
  }
  if (!F->mlp_fp32 && wgrad_mw(Kin) > 0) {
    const int chunk = wgrad_chunk(F->cap, N);  // sized for the allocated batch
    const int nch = (n + chunk - 1) / chunk;
    const int KT = Kin + 16;
    const size_t lds = sizeof(__bf16) * (size_t)(KT + 64) * 40;
    const dim3 grid(nch, N / 64);
    const __bf16* Xb = (const __bf16*)X;
    const __bf16* Yb = (const __bf16*)Y;
    float* part = F->part.as<float>();
    switch (wgrad_mw(Kin)) {
      case 2:
        hipLaunchKernelGGL(k_wgrad<2>, grid, dim3(256), lds, h->stream, Xb, ldx, Yb, ldy, Kin, N, n,
                           chunk, part);
        break;
      default:
        hipLaunchKernelGGL(k_wgrad<3>, grid, dim3(256), lds, h->stream, Xb, ldx, Yb, ldy, Kin, N, n,
                           chunk, part);
    }
    const int tot = (Kin + 1) * N;
    hipLaunchKernelGGL(k_part_reduce, dim3((tot + 31) / 32), dim3(256), 0, h->stream, part, nch,
                       KT, N, Kin, gW, gb);
    return;
  }
  const int chunk = 2048;
  const int nch = (n + chunk - 1) / chunk;
  const dim3 grid(Kin / 64 + 1, N / 64, nch);
  LayerIdx none;
  std::memset(&none, 0, sizeof(none));
  if (F->mlp_fp32)
    hipLaunchKernelGGL(k_gemm_tn_part<false>, grid, dim3(256), 0, h->stream, X, ldx, Y, ldy, Kin, N,
                       n, chunk, F->part.as<float>(), nullptr, 0, none);
  else
    hipLaunchKernelGGL(k_gemm_tn_part<true>, grid, dim3(256), 0, h->stream, X, ldx, Y, ldy, Kin, N,
                       n, chunk, F->part.as<float>(), nullptr, 0, none);
  const int tot = (Kin + 1) * N;
  hipLaunchKernelGGL(k_part_reduce, dim3((tot + 31) / 32), dim3(256), 0, h->stream,
                     F->part.as<float>(), nch, Kin + 64, N, Kin, gW, gb);
}

// Input-layer weight and bias gradients: dW1[i][j] = sum_b u[b][A_i] dh[b][j], db1
static void input_weight_grad(gra_handle* h, FlowState* F, const float* u, const LayerIdx& idx,
                              const void* Y, int ldy, int N, int n, float* gW1, float* gb1) {
  if (!F->mlp_fp32 && !F->fc.wgrad_split && N % 128 == 0) {
    wgrad_tile(h, F, nullptr, 0, u, idx, (const __bf16*)Y, ldy, 16, N, n, true, gW1, gb1);
    return;
  }
  const int chunk = 2048;
  const int nch = (n + chunk - 1) / chunk;
  const dim3 grid(1, N / 64, nch);
  if (F->mlp_fp32)
    hipLaunchKernelGGL(k_gemm_tn_part<false>, grid, dim3(256), 0, h->stream, nullptr, 0, Y, ldy, 64,
                       N, n, chunk, F->part.as<float>(), u, F->D, idx);
  else
    hipLaunchKernelGGL(k_gemm_tn_part<true>, grid, dim3(256), 0, h->stream, nullptr, 0, Y, ldy, 64, N,
                       n, chunk, F->part.as<float>(), u, F->D, idx);
  const int tot = (idx.dA + 1) * N;
  hipLaunchKernelGGL(k_part_reduce, dim3((tot + 31) / 32), dim3(256), 0, h->stream,
                     F->part.as<float>(), nch, 64, N, idx.dA, gW1, gb1);
}

// MLP of layer l on u[A] -> raw spline parameters (prm_of(l)); activations stored
static void mlp_forward(gra_handle* h, FlowState* F, int l, const float* u, int n) {
  const FlowLayerHost& Ly = F->layers[l];
  const int H = F->H;
  const float* th = F->theta.as<float>();
  hipLaunchKernelGGL(k_mlp_in, dim3((unsigned)(((size_t)n * H + 255) / 256)), dim3(256), 0, h->stream, u, F->D, Ly.idx,
                     th + Ly.oW1, th + Ly.ob1, H, n, act16_of(F, l, 0),
                     F->mlp_fp32 ? act32_of(F, l, 0) : nullptr);
  int slot = l * F->depth;
  for (int i = 1; i <= F->depth; ++i, ++slot) {
    const bool out = i == F->depth;
    Epi ep;
    std::memset(&ep, 0, sizeof(ep));
    ep.bias = th + (out ? Ly.obo : Ly.obh[i]);
    ep.relu = out ? 0 : 1;
    ep.ldc = out ? F->ldp : H;
    ep.ncols_valid = out ? Ly.idx.dB * F->nps : H;
    if (out) ep.out32 = prm_of(F, l);
    else {
      ep.out16 = act16_of(F, l, i);
      if (F->mlp_fp32) ep.out32 = act32_of(F, l, i);
    }
    const int N = out ? F->ldp : H;
    if (F->mlp_fp32)
      gemm_nt(h, F, act32_of(F, l, i - 1), H, F->wt32.as<float>() + F->wofs[slot], H, n, N, H, ep);
    else
      gemm_nt(h, F, act16_of(F, l, i - 1), H, F->wt16.as<__bf16>() + F->wofs[slot], H, n, N, H, ep);
  }
}

// Points per block of the spline backward: one thread per transformed
// dimension, the parameter tile within 159 KB of LDS
static int bwd_rows(const FlowState* F, int dB) {
  const int by_lds = (int)((159 * 1024) / (4 * (size_t)(F->ldp + 1)));
  return std::max(1, std::min(256 / dB, by_lds));
}
// The spline backward reads and writes each parameter block directly with
// 16-byte accesses (no LDS tile) for the unrolled bin counts; the runtime-K
// loops use the tile
static bool bwd_direct(const FlowState* F, int dB) {
  (void)dB;
  return F->fc.spline_loops == 0 && (F->K == 8 || F->K == 16 || F->K == 32);
}

// Kernel instances of the fused layer by (hidden, bins) and of the spline
// backward by bins: 8, 16 and 32 bins run the unrolled register forms, other
// bin counts the runtime loops
template <int H>
static const void* layer_inverse_kernel_h(int K) {
  switch (K) {
    case 8: return (const void*)k_flow_layer_inverse<H, 8>;
    case 16: return (const void*)k_flow_layer_inverse<H, 16>;
    case 32: return (const void*)k_flow_layer_inverse<H, 32>;
    default: return (const void*)k_flow_layer_inverse<H, 0>;
  }
}
static const void* layer_inverse_kernel(int H, int K) {
  return H == 64 ? layer_inverse_kernel_h<64>(K)
                 : (H == 128 ? layer_inverse_kernel_h<128>(K) : layer_inverse_kernel_h<256>(K));
}
static const void* spline_backward_kernel(int K, bool direct) {
  switch (K) {
    case 8: return direct ? (const void*)k_rqs_backward_t<8, true> : (const void*)k_rqs_backward_t<8, false>;
    case 16: return direct ? (const void*)k_rqs_backward_t<16, true> : (const void*)k_rqs_backward_t<16, false>;
    case 32: return direct ? (const void*)k_rqs_backward_t<32, true> : (const void*)k_rqs_backward_t<32, false>;
    default: return (const void*)k_rqs_backward_t<0, false>;
  }
}

// Bin count that selects the kernel instance: 0 (the runtime-K loops) when the
// configuration asks for them
static int spline_k(const FlowState* F) { return F->fc.spline_loops ? 0 : F->K; }

// LDS bytes of the fused layer with a parameter tile of row stride ldg
static size_t fused_lds_bytes(const FlowState* F, int ldg) {
  const size_t R = (size_t)fused_rows(F->H);
  const size_t act = R * (F->H + 8) * sizeof(__bf16);
  return act + std::max(act, R * (size_t)ldg * sizeof(float)) + R * 8 * 4 + R * F->D * 4;
}

// Output groups of the fused layer for a layer with dB transformed dimensions:
// the group of gd dimensions [j0, j1) computes the n-tiles covering columns
// [j0 nps, j1 nps); the tile stride is that of the widest group. Without a
// requested size the largest group that keeps fused_blocks(H) blocks per CU is taken.
// LDS per block that lets fused_blocks(H) blocks share a CU's 160 KB
static size_t fused_lds_target(int H) {
  return (size_t)(160 * 1024 / fused_blocks(H)) - 512;
}
static void fused_groups(const FlowState* F, int dB, int want, int& gd, int& ldg, size_t& lds) {
  const int nps = prm_stride(F->K);
  auto stride = [&](int g) {
    int w = 0;
    for (int j0 = 0; j0 < dB; j0 += g) {
      const int j1 = std::min(dB, j0 + g);
      w = std::max(w, (j1 * nps + 15) / 16 - (j0 * nps) / 16);
    }
    return 16 * w + 1;
  };
  gd = want > 0 ? std::min(want, dB) : dB;
  if (want <= 0)
    while (gd > 1 && fused_lds_bytes(F, stride(gd)) > fused_lds_target(F->H)) --gd;
  ldg = stride(gd);
  lds = fused_lds_bytes(F, ldg);
}

// Sampling pass: z -> u_0 (stored per layer), ld = log q(u_0). With store set,
// the MLP activations and spline parameters of every layer are kept for the
// backward pass.
static int flow_sample_pass(gra_handle* h, FlowState* F, int n, uint64_t seed, uint32_t step,
                            uint64_t first, const float* zin, bool store) {
  const int D = F->D, L = F->L;
  if (zin)
    GRA_HIP_CHECK(h, hipMemcpyAsync(u_of(F, L), zin, sizeof(float) * n * D, hipMemcpyHostToDevice,
                                    h->stream));
  else
    hipLaunchKernelGGL(k_flow_z, dim3((n + 255) / 256), dim3(256), 0, h->stream, seed, step, first,
                       n, D, u_of(F, L));
  GRA_HIP_CHECK(h, hipMemsetAsync(F->ld.p, 0, sizeof(float) * n, h->stream));
  if (!F->mlp_fp32) {
    for (int l = L - 1; l >= 0; --l) {
      const FlowLayerHost& Ly = F->layers[l];
      const size_t lds = Ly.lds;
      FusedLayer W;
      std::memset(&W, 0, sizeof(W));
      const float* th = F->theta.as<float>();
      W.W1 = th + Ly.oW1;
      W.b1 = th + Ly.ob1;
      for (int i = 1; i <= F->depth; ++i) {
        W.wt[i - 1] = F->wt16.as<__bf16>() + F->wofs[l * F->depth + i - 1];
        W.bias[i - 1] = th + (i == F->depth ? Ly.obo : Ly.obh[i]);
      }
      if (store) {
        for (int i = 0; i < F->depth; ++i) W.act_out[i] = act16_of(F, l, i);
        W.prm_out = prm_of(F, l);
      }
      const int R = fused_rows(F->H);
      const dim3 grd((n + R - 1) / R);
      const float* uin = u_of(F, l + 1);
      float* uout = u_of(F, l);
      float* ldp_ = F->ld.as<float>();
      LayerIdx idx = Ly.idx;
      int nn = n, DD = D, KK = F->K, dep = F->depth, ldp = F->ldp, gd = Ly.gd, ldg = Ly.ldg;
      void* args[] = {(void*)&uin, (void*)&uout, (void*)&ldp_, (void*)&nn,  (void*)&DD, (void*)&idx,
                      (void*)&KK,  (void*)&dep,  (void*)&ldp,  (void*)&gd, (void*)&ldg, (void*)&W};
      GRA_HIP_CHECK(h, hipLaunchKernel(layer_inverse_kernel(F->H, spline_k(F)), grd,
                                       dim3(fused_threads(F->H)), args, lds, h->stream));
    }
    GRA_HIP_CHECK(h, hipGetLastError());
    return GRA_OK;
  }
  for (int l = L - 1; l >= 0; --l) {
    // layer l maps u_l (stored at slot l) to u_{l+1}; inverting reads slot l+1
    mlp_forward(h, F, l, u_of(F, l + 1), n);
    LAUNCH_SPLINE(k_rqs_inverse, F->K, n, prm_of(F, l), F->ldp, u_of(F, l + 1), u_of(F, l),
                  F->ld.as<float>(), n, D, F->layers[l].idx, F->K);
  }
  GRA_HIP_CHECK(h, hipGetLastError());
  return GRA_OK;
}

// Backward of L = -sum_b coef_b log q(u_b) into F->grad (overwritten);
// requires the stored sampling-pass state of the same points. If logq_chk is
// given it receives the forward-direction log q recomputed from the splines.
static int flow_backward(gra_handle* h, FlowState* F, int n, float* logq_chk) {
  const int D = F->D, L = F->L, H = F->H, ldp = F->ldp;
  float* g = F->gu0.as<float>();
  float* g2 = F->gu1.as<float>();
  float* grad = F->grad.as<float>();
  GRA_HIP_CHECK(h, hipMemsetAsync(g, 0, sizeof(float) * n * D, h->stream));
  GRA_HIP_CHECK(h, hipMemsetAsync(grad, 0, sizeof(float) * F->nparam, h->stream));
  if (logq_chk) GRA_HIP_CHECK(h, hipMemsetAsync(logq_chk, 0, sizeof(float) * n, h->stream));
  const float* th = F->theta.as<float>();
  for (int l = L - 1; l >= 0; --l) {
    const FlowLayerHost& Ly = F->layers[l];
    // g = dL/du_{l+1}; spline backward gives dL/dprm and dL/du_l (B dims; A passes through)
    {
      const bool direct = bwd_direct(F, Ly.idx.dB);
      const int R = direct ? 256 / Ly.idx.dB : bwd_rows(F, Ly.idx.dB);
      const size_t lds = direct ? 0 : sizeof(float) * (size_t)R * (ldp + 1);
      const float* a_prm = prm_of(F, l);
      const float* a_up = u_of(F, l);
      const float* a_gin = g;
      const float* a_coef = F->coef.as<float>();
      // the FP32 copy of the parameter gradients feeds only the FP32 MLP path
      float* a_gp = F->mlp_fp32 ? F->gprm.as<float>() : nullptr;
      __bf16* a_gp16 = F->gprm16.as<__bf16>();
      LayerIdx idx = Ly.idx;
      int a_ldp = ldp, nn = n, DD = D, KK = F->K, RR = R;
      void* args[] = {(void*)&a_prm, (void*)&a_ldp, (void*)&a_up,   (void*)&a_gin,
                      (void*)&a_coef, (void*)&g2,   (void*)&a_gp,   (void*)&a_gp16,
                      (void*)&nn,     (void*)&DD,   (void*)&idx,    (void*)&KK,
                      (void*)&RR,     (void*)&logq_chk};
      GRA_HIP_CHECK(h, hipLaunchKernel(spline_backward_kernel(spline_k(F), direct),
                                       dim3((n + R - 1) / R), dim3(256), args, lds, h->stream));
    }
    // MLP backward from the output layer down
    const float* dy32 = F->gprm.as<float>();
    const __bf16* dy16 = F->gprm16.as<__bf16>();
    int ldy = ldp, ncol = ldp;
    int slot = l * F->depth + F->depth - 1;
    int cur = 0;
    for (int i = F->depth; i >= 1; --i, --slot) {
      const bool out = i == F->depth;
      // weight and bias gradients of product i: X = act_{i-1} [n x H], dY [n x ncol]
      float* gW = grad + (out ? Ly.oWo : Ly.oWh[i]);
      float* gb = grad + (out ? Ly.obo : Ly.obh[i]);
      if (F->mlp_fp32)
        weight_grad(h, F, act32_of(F, l, i - 1), H, dy32, ldy, H, ncol, n, gW, gb);
      else
        weight_grad(h, F, act16_of(F, l, i - 1), H, dy16, ldy, H, ncol, n, gW, gb);
      // dX = dY W^T, masked by ReLU'(act_{i-1})
      Epi ep;
      std::memset(&ep, 0, sizeof(ep));
      ep.mask = act16_of(F, l, i - 1);
      ep.ldm = H;
      ep.out16 = F->dact16.as<__bf16>() + (size_t)cur * F->cap * H;
      // the FP32 copy feeds the FP32 path only (the BF16 path's input-layer dx
      // kernel reads the BF16 product)
      ep.out32 = F->mlp_fp32 ? F->dact32.as<float>() + (size_t)cur * F->cap * H : nullptr;
      ep.ldc = H;
      ep.ncols_valid = H;
      if (F->mlp_fp32)
        gemm_nt(h, F, dy32, ldy, F->theta.as<float>() + (out ? Ly.oWo : Ly.oWh[i]), ncol, n, H,
                ncol, ep);
      else
        gemm_nt(h, F, dy16, ldy, F->w16.as<__bf16>() + F->wofs[slot], ncol, n, H, ncol, ep);
      dy32 = ep.out32;
      dy16 = ep.out16;
      ldy = H;
      ncol = H;
      cur ^= 1;
    }
    // input layer: dW1[i][j] = sum_b u[b][A_i] dh[b][j], db1, and dL/du[A] += dh W1^T
    if (F->mlp_fp32)
      input_weight_grad(h, F, u_of(F, l), Ly.idx, dy32, H, H, n, grad + Ly.oW1, grad + Ly.ob1);
    else
      input_weight_grad(h, F, u_of(F, l), Ly.idx, dy16, H, H, n, grad + Ly.oW1, grad + Ly.ob1);
    if (F->mlp_fp32)
      hipLaunchKernelGGL(k_mlp_in_dx<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                         h->stream, dy32, th + Ly.oW1, H, n, D, Ly.idx, g2);
    else
      hipLaunchKernelGGL(k_mlp_in_dx<__bf16>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                         h->stream, dy16, th + Ly.oW1, H, n, D, Ly.idx, g2);
    std::swap(g, g2);
  }
  GRA_HIP_CHECK(h, hipGetLastError());
  return GRA_OK;
}

// Integrand weights of the current sample: x = grid(u_0), w = f vw / q (FP64)
static int flow_weights(gra_handle* h, FlowState* F, int n, double stats[3]) {
  const int D = F->D;
  hipLaunchKernelGGL(k_u_to_x, dim3((n + 255) / 256), dim3(256), 0, h->stream, u_of(F, 0),
                     F->grid.as<double>(), D, (unsigned int)GRA_VEGAS_BINS, n,
                     F->xx.as<double>(), F->vw.as<double>());
  int r = ensure_point_buffers(h, n);
  if (r) return r;
  r = run_integrand(h, nullptr, 0, 0, 0, 0, F->xx.as<double>(), n, F->W.as<double>(), nullptr,
                    h->d_flags.as<uint8_t>(), h->d_vw.as<double>(), h->d_bins.as<uint8_t>());
  if (r) return r;
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_weights, dim3(nb), dim3(256), 0, h->stream, F->W.as<double>(),
                     F->vw.as<double>(), F->ld.as<float>(), n, F->wts.as<double>(),
                     F->tmp.as<double>());
  F->hpart.resize(3 * (size_t)nb);
  GRA_HIP_CHECK(h, hipMemcpyAsync(F->hpart.data(), F->tmp.p, sizeof(double) * 3 * nb,
                                  hipMemcpyDeviceToHost, h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  double s1 = 0, s2 = 0, mx = 0;
  for (int k = 0; k < nb; ++k) {
    s1 += F->hpart[3 * k];
    s2 += F->hpart[3 * k + 1];
    mx = std::max(mx, F->hpart[3 * k + 2]);
  }
  stats[0] = s1;
  stats[1] = s2;
  stats[2] = mx;
  return GRA_OK;
}

static int flow_adam(gra_handle* h, FlowState* F) {
  F->step += 1;
  const gra_flow_cfg& c = F->fc;
  const float c1 = (float)(1.0 / (1.0 - std::pow(c.beta1, F->step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow(c.beta2, F->step)));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((F->nparam + 255) / 256)), dim3(256), 0, h->stream,
                     F->theta.as<float>(), F->grad.as<float>(), F->adam_m.as<float>(),
                     F->adam_v.as<float>(), F->nparam, (float)c.lr, (float)c.beta1,
                     (float)c.beta2, (float)c.eps, c1, c2, F->frozen.as<unsigned char>());
  return refresh_weights(h, F);
}

static void fill_stats(gra_flow_stats* st, const double s[3], uint64_t n, double seconds) {
  const double N = (double)n;
  const double mean = s[0] / N;
  const double var = std::max(s[1] / N - mean * mean, 0.0);
  st->sigma = mean;
  st->sigma_err = std::sqrt(var / N);
  st->cv2 = mean > 0 ? var / (mean * mean) : 0.0;
  st->ess = s[1] > 0 ? (s[0] * s[0]) / (N * s[1]) : 0.0;
  st->maxw = s[2];
  st->n = n;
  st->seconds = seconds;
}

// `steps` Adam steps on batches of n flow points (buffers allocated): st receives
// the last batch's sums (sum w, sum w^2, max w) and *cv2sum the sum over the steps
// of each batch's CV^2 of w
static int flow_train_loop(gra_handle* h, FlowState* F, int n, int steps, double st[3],
                           double* cv2sum) {
  for (int it = 0; it < steps; ++it) {
    int r = flow_sample_pass(h, F, n, F->fc.seed, (uint32_t)F->step, 0, nullptr, true);
    if (r) return r;
    r = flow_weights(h, F, n, st);
    if (r) return r;
    if (!(st[1] > 0)) return fail(h, GRA_E_NUMERIC, "flow_train: all weights are zero");
    const double mean = st[0] / n;
    *cv2sum += std::max(st[1] / n - mean * mean, 0.0) / (mean * mean);
    hipLaunchKernelGGL(k_coef, dim3((n + 255) / 256), dim3(256), 0, h->stream, F->wts.as<double>(),
                       n, 1.0 / st[1], F->coef.as<float>());
    r = flow_backward(h, F, n, nullptr);
    if (r) return r;
    r = flow_adam(h, F);
    if (r) return r;
  }
  return GRA_OK;
}

}  // namespace gra_hip

using namespace gra_hip;

extern "C" int gra_hip_flow_init(gra_handle* h, const gra_flow_cfg* fc_in) {
  if (!h) return set_global_error(GRA_E_ARG, "handle is NULL");
  if (!fc_in) return fail(h, GRA_E_ARG, "flow_init: cfg is NULL");
  gra_flow_cfg fc = *fc_in;
  const int D = h->P.dim;
  const bool mlp32 = fc.depth < 0;  // negative depth selects the FP32 MLP path
  if (mlp32) fc.depth = -fc.depth;
  if (fc.dim != D) return fail(h, GRA_E_ARG, "flow dim %d != process dim %d", fc.dim, D);
  if (fc.layers < 1 || fc.layers > 32) return fail(h, GRA_E_ARG, "flow layers %d", fc.layers);
  if (fc.bins < 2 || fc.bins > kMaxK) return fail(h, GRA_E_ARG, "flow bins %d (2..%d)", fc.bins, kMaxK);
  if (fc.hidden < 64 || fc.hidden % 64 || fc.hidden > 1024)
    return fail(h, GRA_E_ARG, "flow hidden %d: multiple of 64 up to 1024", fc.hidden);
  if (fc.depth < 1 || fc.depth > 4) return fail(h, GRA_E_ARG, "flow depth %d (1..4)", fc.depth);
  if (D < 2 || D / 2 > kMaxDB + 0 || (D - D / 2) > kMaxDB)
    return fail(h, GRA_E_ARG, "flow needs 2 <= dim <= %d", 2 * kMaxDB);
  if (h->flow) {
    flow_release(h->flow);
    h->flow = nullptr;
  }
  FlowState* F = new (std::nothrow) FlowState();
  if (!F) return fail(h, GRA_E_NOMEM, "flow state");
  F->fc = fc;
  F->mlp_fp32 = mlp32 ? 1 : 0;
  F->D = D;
  F->L = fc.layers;
  F->K = fc.bins;
  F->H = fc.hidden;
  F->depth = fc.depth;
  F->np = 3 * fc.bins - 1;
  F->nps = prm_stride(fc.bins);
  int maxdB = 0;
  for (int l = 0; l < F->L; ++l) {
    FlowLayerHost Ly;
    std::memset(&Ly, 0, sizeof(Ly));
    Ly.idx = make_mask(l, D);
    maxdB = std::max(maxdB, Ly.idx.dB);
    F->layers.push_back(Ly);
  }
  F->ldp = ((maxdB * F->nps + 63) / 64) * 64;
  // parameter layout
  size_t off = 0;
  for (int l = 0; l < F->L; ++l) {
    FlowLayerHost& Ly = F->layers[l];
    Ly.oW1 = off;
    off += (size_t)Ly.idx.dA * F->H;
    Ly.ob1 = off;
    off += F->H;
    for (int i = 1; i < F->depth; ++i) {
      Ly.oWh[i] = off;
      off += (size_t)F->H * F->H;
      Ly.obh[i] = off;
      off += F->H;
    }
    Ly.oWo = off;
    off += (size_t)F->H * F->ldp;
    Ly.obo = off;
    off += F->ldp;
  }
  F->nparam = off;
  // GEMM weight copies
  size_t wo = 0;
  for (int l = 0; l < F->L; ++l)
    for (int i = 1; i <= F->depth; ++i) {
      F->wofs.push_back(wo);
      wo += (size_t)F->H * ((i == F->depth) ? F->ldp : F->H);
    }
  if (F->theta.ensure(4 * off) || F->grad.ensure(4 * off) || F->adam_m.ensure(4 * off) ||
      F->adam_v.ensure(4 * off) || F->frozen.ensure(off) || F->wt16.ensure(2 * wo) ||
      F->w16.ensure(2 * wo) || F->wt32.ensure(4 * wo)) {
    flow_release(F);
    return fail(h, GRA_E_NOMEM, "flow parameters");
  }
  // initialisation: He-uniform hidden weights, zero output weights, output biases
  // giving the identity spline (uniform bins, unit derivatives)
  std::vector<float> th(off, 0.f);
  std::vector<unsigned char> fr(off, 0);
  uint64_t s = fc.seed ^ 0x5DEECE66Dull;
  auto unif = [&s]() { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); };
  const float dinv = std::log(std::exp(1.0 - kEpsD) - 1.0);
  for (int l = 0; l < F->L; ++l) {
    FlowLayerHost& Ly = F->layers[l];
    const double a1 = std::sqrt(6.0 / std::max(1, Ly.idx.dA));
    for (int k = 0; k < Ly.idx.dA * F->H; ++k) th[Ly.oW1 + k] = (float)((2 * unif() - 1) * a1);
    const double ah = std::sqrt(6.0 / F->H);
    for (int i = 1; i < F->depth; ++i)
      for (int k = 0; k < F->H * F->H; ++k) th[Ly.oWh[i] + k] = (float)((2 * unif() - 1) * ah);
    for (int j = 0; j < Ly.idx.dB; ++j)
      for (int m = 0; m < F->K - 1; ++m) th[Ly.obo + j * F->nps + 2 * F->K + m] = dinv;
    // padding columns of the output layer (each dimension's block past its 3K - 1
    // parameters, and the row past the last block) stay zero and frozen
    for (int c = 0; c < F->ldp; ++c) {
      const int j = c / F->nps;
      if (j < Ly.idx.dB && c - j * F->nps < F->np) continue;
      for (int hrow = 0; hrow < F->H; ++hrow) fr[Ly.oWo + (size_t)hrow * F->ldp + c] = 1;
      fr[Ly.obo + c] = 1;
    }
  }
  hipError_t e1 = hipMemcpy(F->theta.p, th.data(), 4 * off, hipMemcpyHostToDevice);
  hipError_t e2 = hipMemcpy(F->frozen.p, fr.data(), off, hipMemcpyHostToDevice);
  hipError_t e3 = hipMemset(F->adam_m.p, 0, 4 * off);
  hipError_t e4 = hipMemset(F->adam_v.p, 0, 4 * off);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
    flow_release(F);
    return fail(h, GRA_E_HIP, "flow parameter upload");
  }
  if (!F->mlp_fp32) {
    size_t lds = 0;
    for (FlowLayerHost& Ly : F->layers) {
      fused_groups(F, Ly.idx.dB, fc.out_groups, Ly.gd, Ly.ldg, Ly.lds);
      lds = std::max(lds, Ly.lds);
    }
    if (lds > 160 * 1024) {
      flow_release(F);
      return fail(h, GRA_E_ARG, "flow layer tile needs %zu B of LDS (hidden %d, %d spline columns)",
                  lds, fc.hidden, F->ldp);
    }
    if (fc.hidden != 64 && fc.hidden != 128 && fc.hidden != 256) {
      flow_release(F);
      return fail(h, GRA_E_ARG, "flow hidden %d: the BF16 path supports 64, 128, 256", fc.hidden);
    }
    const void* kern = layer_inverse_kernel(fc.hidden, spline_k(F));
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess) {
      flow_release(F);
      return fail(h, GRA_E_HIP, "flow: cannot reserve %zu B of LDS per block", lds);
    }
  }
  {
    size_t bl = 0;
    for (int l = 0; l < F->L; ++l)
      bl = std::max(bl, sizeof(float) * (size_t)bwd_rows(F, F->layers[l].idx.dB) * (F->ldp + 1));
    if (hipFuncSetAttribute(spline_backward_kernel(spline_k(F), false),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bl) != hipSuccess) {
      flow_release(F);
      return fail(h, GRA_E_HIP, "flow: cannot reserve %zu B of LDS for the spline backward", bl);
    }
  }
  h->flow = F;
  int r = refresh_weights(h, F);
  if (r) return r;
  // the VEGAS grid the flow composes with: uniform unless a grid was set
  if (F->grid.ensure(sizeof(double) * GRA_MAX_DIM * GRA_VEGAS_BINS))
    return fail(h, GRA_E_NOMEM, "grid buffer");
  gra_grid g;
  grid_init(&g, D, GRA_VEGAS_BINS);
  GRA_HIP_CHECK(h, hipMemcpy(F->grid.p, g.xmat, sizeof(double) * D * GRA_VEGAS_BINS,
                             hipMemcpyHostToDevice));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return GRA_OK;
}

extern "C" int gra_hip_flow_set_grid(gra_handle* h, const gra_grid* g) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow_set_grid: flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (!g || g->dim != h->P.dim || g->bins != GRA_VEGAS_BINS)
    return fail(h, GRA_E_ARG, "flow_set_grid: grid must have the process dim and %d bins",
                GRA_VEGAS_BINS);
  GRA_HIP_CHECK(h, hipMemcpy(F->grid.p, g->xmat, sizeof(double) * g->dim * GRA_VEGAS_BINS,
                             hipMemcpyHostToDevice));
  return GRA_OK;
}

extern "C" int gra_hip_flow_train(gra_handle* h, int32_t steps, gra_flow_stats* last) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow_train: flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (steps < 0) return fail(h, GRA_E_ARG, "flow_train: steps < 0");
  const uint64_t B64 = F->fc.batch;
  if (B64 < 256 || B64 % 256 || B64 > (1u << 24))
    return fail(h, GRA_E_ARG, "flow batch %llu: multiple of 256 up to 2^24",
                (unsigned long long)B64);
  const int n = (int)B64;
  int r = flow_alloc_batch(h, F, n);
  if (r) return r;
  const auto t0 = std::chrono::steady_clock::now();
  double st[3] = {0, 0, 0}, cv2sum = 0.0;
  r = flow_train_loop(h, F, n, steps, st, &cv2sum);
  if (r) return r;
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  if (last) {
    std::memset(last, 0, sizeof(*last));
    fill_stats(last, st, (uint64_t)n,
               std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    last->loss = st[1] > 0 ? st[1] / n : 0.0;
  }
  return GRA_OK;
}

namespace gra_hip {
// Sampling chunk size: a multiple of 256 points (the MLP products tile rows by
// 64), the training batch by default, at most 2^20
static int sample_chunk(const FlowState* F) {
  uint64_t c0 = F->fc.batch > 0 ? F->fc.batch : (1u << 18);
  return (int)std::min<uint64_t>(std::max<uint64_t>(256, (c0 / 256) * 256), 1u << 20);
}

// One sampling chunk of pass `pass`: n points (padded to a multiple of 256 in
// the flow, the padding never weighted) with their flow-stream index starting at
// `first`, the integrand and the weights w = f vw / q in F->wts, x in F->xx
static int flow_chunk(gra_handle* h, FlowState* F, int n, uint64_t seed, uint32_t pass,
                      uint32_t call, uint64_t first, double st[3]) {
  const int npad = ((n + 255) / 256) * 256;
  const uint32_t stream = 0x40000000u | ((pass & 0x3FFFu) << 16) | (call & 0xFFFFu);
  int r = flow_sample_pass(h, F, npad, seed, stream, first, nullptr, false);
  if (r) return r;
  return flow_weights(h, F, n, st);
}
}  // namespace gra_hip

extern "C" int gra_hip_flow_sample(gra_handle* h, uint64_t n_total, uint64_t seed, double* w,
                                   double* x, gra_flow_stats* out) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow_sample: flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (!out) return fail(h, GRA_E_ARG, "flow_sample: out is NULL");
  if (n_total == 0) return fail(h, GRA_E_ARG, "flow_sample: n = 0");
  const int chunk = sample_chunk(F);
  if ((n_total + chunk - 1) / chunk > 0xFFFF)
    return fail(h, GRA_E_ARG, "flow_sample: n = %llu needs more than 65535 chunks of %d",
                (unsigned long long)n_total, chunk);
  int r = flow_alloc_batch(h, F, chunk);
  if (r) return r;
  const auto t0 = std::chrono::steady_clock::now();
  double acc[3] = {0, 0, 0};
  uint64_t done = 0;
  uint32_t call = 0;
  const int D = F->D;
  while (done < n_total) {
    const int n = (int)std::min<uint64_t>(chunk, n_total - done);
    double st[3];
    r = flow_chunk(h, F, n, seed, 0, call, done, st);
    if (r) return r;
    if (w)
      GRA_HIP_CHECK(h, hipMemcpyAsync(w + done, F->wts.p, sizeof(double) * n,
                                      hipMemcpyDeviceToHost, h->stream));
    if (x)
      GRA_HIP_CHECK(h, hipMemcpyAsync(x + done * D, F->xx.p, sizeof(double) * n * D,
                                      hipMemcpyDeviceToHost, h->stream));
    if (w || x) GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
    acc[0] += st[0];
    acc[1] += st[1];
    acc[2] = std::max(acc[2], st[2]);
    done += n;
    ++call;
  }
  std::memset(out, 0, sizeof(*out));
  fill_stats(out, acc, n_total,
             std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  out->fallback = (out->ess < F->fc.ess_fallback) ? 1 : 0;
  return GRA_OK;
}

extern "C" int gra_hip_flow_integrate(gra_handle* h, const gra_flow_run_cfg* rc, gra_grid* grid,
                                      gra_flow_result* res) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow_integrate: flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (!rc || !res) return fail(h, GRA_E_ARG, "flow_integrate: NULL argument");
  if (rc->n_max == 0 || rc->train_steps < 0)
    return fail(h, GRA_E_ARG, "flow_integrate: n_max >= 1 and train_steps >= 0 required");
  std::memset(res, 0, sizeof(*res));
  res->seconds_to_precision = -1.0;
  const auto t0 = std::chrono::steady_clock::now();
  auto since = [&t0]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  int r = GRA_OK;
  if (rc->train_steps > 0 && rc->check_every > 0 && rc->plateau > 0.0) {
    // training in windows of check_every steps, stopped when the mean CV^2 of a
    // window's training batches improves on the previous window's by less than
    // the fraction rc->plateau
    const uint64_t B64 = F->fc.batch;
    if (B64 < 256 || B64 % 256 || B64 > (1u << 24))
      return fail(h, GRA_E_ARG, "flow batch %llu: multiple of 256 up to 2^24",
                  (unsigned long long)B64);
    r = flow_alloc_batch(h, F, (int)B64);
    if (r) return r;
    double prev = -1.0, st[3];
    int done_steps = 0;
    while (done_steps < rc->train_steps) {
      const int w = std::min(rc->check_every, rc->train_steps - done_steps);
      double cv2sum = 0.0;
      r = flow_train_loop(h, F, (int)B64, w, st, &cv2sum);
      if (r) return r;
      done_steps += w;
      const double cur = cv2sum / w;
      res->train_cv2 = cur;
      if (prev > 0.0 && prev - cur < rc->plateau * prev) break;
      prev = cur;
    }
    res->train_steps_run = (uint64_t)done_steps;
  } else if (rc->train_steps > 0) {
    gra_flow_stats st;
    r = gra_hip_flow_train(h, rc->train_steps, &st);
    if (r) return r;
    res->train_steps_run = (uint64_t)rc->train_steps;
    res->train_cv2 = st.cv2;
  }
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  res->train_seconds = since();
  const int chunk = sample_chunk(F);
  if ((rc->n_max + chunk - 1) / chunk > 0xFFFF)
    return fail(h, GRA_E_ARG, "flow_integrate: n_max needs more than 65535 chunks");
  r = flow_alloc_batch(h, F, chunk);
  if (r) return r;
  // SampleNeuro's integration (src/MGraniitti.cc:1586-1622): plain mean of
  // w = f vw / q over the flow's points until the relative error reaches the
  // requested precision or the budget ends
  double acc[3] = {0, 0, 0};
  uint64_t done = 0;
  uint32_t call = 0;
  gra_flow_stats fs;
  std::memset(&fs, 0, sizeof(fs));
  bool low_ess = false;
  while (done < rc->n_max) {
    const int n = (int)std::min<uint64_t>(chunk, rc->n_max - done);
    double st[3];
    r = flow_chunk(h, F, n, rc->seed, 0, call++, done, st);
    if (r) return r;
    acc[0] += st[0];
    acc[1] += st[1];
    acc[2] = std::max(acc[2], st[2]);
    done += n;
    fill_stats(&fs, acc, done, 0.0);
    if (fs.ess < rc->ess_min) {
      low_ess = true;
      break;
    }
    if (rc->precision > 0 && fs.sigma > 0 && fs.sigma_err / fs.sigma <= rc->precision) {
      res->seconds_to_precision = since();
      res->reached = 1;
      break;
    }
  }
  res->sample_seconds = since() - res->train_seconds;
  res->sigma = fs.sigma;
  res->sigma_err = fs.sigma_err;
  res->cv2 = fs.cv2;
  res->ess = fs.ess;
  res->maxw = fs.maxw;
  res->n = done;
  if (!low_ess) return GRA_OK;
  // effective sample size below the threshold: plain VEGAS takes over and its
  // result is the one reported
  if (!grid) return fail(h, GRA_E_ARG, "flow_integrate: ESS %.3g < %.3g and no grid for VEGAS",
                         fs.ess, rc->ess_min);
  res->fallback = 1;
  r = gra_hip_vegas_run(h, &rc->vegas, grid, &res->vegas);
  if (r) return r;
  res->sigma = res->vegas.sigma;
  res->sigma_err = res->vegas.sigma_err;
  res->reached = (rc->precision > 0 && res->vegas.seconds_to_precision >= 0) ? 1 : 0;
  res->seconds_to_precision =
      res->reached ? res->train_seconds + res->sample_seconds + res->vegas.seconds_to_precision
                   : -1.0;
  return GRA_OK;
}

extern "C" int gra_hip_flow_generate(gra_handle* h, uint64_t n_calls, uint64_t seed,
                                     uint32_t pass, double maxw, gra_event* events,
                                     uint64_t max_events, gra_gen_stats* st) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow_generate: flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (!st) return fail(h, GRA_E_ARG, "flow_generate: st is NULL");
  if (max_events > 0 && !events) return fail(h, GRA_E_ARG, "flow_generate: events is NULL");
  if (!(maxw > 0.0)) return fail(h, GRA_E_ARG, "flow_generate: maxw must be > 0");
  if (n_calls == 0) return fail(h, GRA_E_ARG, "flow_generate: n_calls = 0");
  if (pass > 0x3FFFu) return fail(h, GRA_E_ARG, "flow_generate: pass %u > 16383", pass);
  const int chunk = sample_chunk(F);
  if ((n_calls + chunk - 1) / chunk > 0xFFFF)
    return fail(h, GRA_E_ARG, "flow_generate: n_calls needs more than 65535 chunks");
  int r = flow_alloc_batch(h, F, chunk);
  if (r) return r;
  r = ensure_gen_buffers(h, (uint64_t)chunk, (uint64_t)chunk);
  if (r) return r;
  unsigned long long* d_cnt = h->d_gen_cnt.as<unsigned long long>();
  unsigned long long* c = h->h_gen_cnt;
  std::memset(st, 0, sizeof(*st));
  uint64_t done = 0;
  uint32_t call = 0;
  while (done < n_calls) {
    const int n = (int)std::min<uint64_t>(chunk, n_calls - done);
    double s3[3];
    r = flow_chunk(h, F, n, seed, pass, call++, done, s3);
    if (r) return r;
    // SaveEvent's hit-or-miss on w = f vw / q: U < w / maxw, never when w > maxw;
    // the accepted indices come out in ascending order
    const uint64_t room = max_events > st->stored ? max_events - st->stored : 0;
      GRA_HIP_CHECK(h, launch_accept(F->wts.as<double>(), nullptr, h->d_flags.as<uint8_t>(),
                                   (uint64_t)n, maxw, seed, 0x10000000u | pass, done,
                                   h->d_gen_mask.as<uint64_t>(), h->d_gen_tile.as<uint32_t>(),
                                   h->d_gen_idx.as<uint64_t>(), std::min<uint64_t>(room, n),
                                   d_cnt, h->stream));
    GRA_HIP_CHECK(h, hipMemcpyAsync(c, d_cnt, 6 * sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost, h->stream));
    GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
    r = count_nonfinite(h, c[4], c[5]);
    if (r) return r;
    st->trials += c[0];
    st->valid += c[1];
    st->accepted += c[2];
    st->overflow += c[3];
    const uint64_t take = std::min<uint64_t>(c[2], room);
    if (take > 0) {
      GRA_HIP_CHECK(h, k64::launch_event_record(h->P, nullptr, 0, 0, 0, 0, F->xx.as<double>(),
                                                h->d_gen_idx.as<uint64_t>(), F->wts.as<double>(),
                                                nullptr, take, nullptr, h->d_gen_ev.p,
                                                h->stream));
      r = events_to_host(h, events + st->stored, h->d_gen_ev.p, take);
      if (r) return r;
      for (uint64_t e = 0; e < take; ++e) events[st->stored + e].index += done;
      st->stored += take;
    }
    done += n;
  }
  return GRA_OK;
}

extern "C" int gra_hip_flow_param_count(gra_handle* h, uint64_t* count) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (count) *count = F->nparam;
  return GRA_OK;
}

extern "C" int gra_hip_flow_get_params(gra_handle* h, float* params) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (!params) return fail(h, GRA_E_ARG, "params is NULL");
  GRA_HIP_CHECK(h, hipMemcpy(params, F->theta.p, 4 * F->nparam, hipMemcpyDeviceToHost));
  return GRA_OK;
}

extern "C" int gra_hip_flow_set_params(gra_handle* h, const float* params) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (!params) return fail(h, GRA_E_ARG, "params is NULL");
  GRA_HIP_CHECK(h, hipMemcpy(F->theta.p, params, 4 * F->nparam, hipMemcpyHostToDevice));
  return refresh_weights(h, F);
}

extern "C" int gra_hip_flow_forward(gra_handle* h, const float* z, size_t n, float* x,
                                    float* logq) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (!z || !x || !logq) return fail(h, GRA_E_ARG, "flow_forward: NULL buffer");
  if (n == 0 || n % 64 || n > (1u << 24)) return fail(h, GRA_E_ARG, "flow_forward: n %% 64 != 0");
  int r = flow_alloc_batch(h, F, (int)n);
  if (r) return r;
  r = flow_sample_pass(h, F, (int)n, 0, 0, 0, z, false);
  if (r) return r;
  GRA_HIP_CHECK(h, hipMemcpyAsync(x, u_of(F, 0), sizeof(float) * n * F->D, hipMemcpyDeviceToHost,
                                  h->stream));
  GRA_HIP_CHECK(h, hipMemcpyAsync(logq, F->ld.p, sizeof(float) * n, hipMemcpyDeviceToHost,
                                  h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return GRA_OK;
}

// Gradient at caller points u (flow output space) and coefficients c. The
// sampling pass is reconstructed by running the forward direction layer by
// layer (MLP on the conditioning half, forward spline on the other half).

extern "C" int gra_hip_flow_grad(gra_handle* h, const float* u, const float* c, size_t n,
                                 float* grad, float* logq) {
  FlowState* F = flow_of(h);
  if (!F) return h ? fail(h, GRA_E_STATE, "flow not initialised")
                   : set_global_error(GRA_E_ARG, "handle is NULL");
  if (!u || !c || !grad) return fail(h, GRA_E_ARG, "flow_grad: NULL buffer");
  if (n == 0 || n % 64 || n > (1u << 24)) return fail(h, GRA_E_ARG, "flow_grad: n %% 64 != 0");
  const int N = (int)n, D = F->D;
  int r = flow_alloc_batch(h, F, N);
  if (r) return r;
  GRA_HIP_CHECK(h, hipMemcpyAsync(u_of(F, 0), u, sizeof(float) * n * D, hipMemcpyHostToDevice,
                                  h->stream));
  GRA_HIP_CHECK(h, hipMemcpyAsync(F->coef.p, c, sizeof(float) * n, hipMemcpyHostToDevice,
                                  h->stream));
  for (int l = 0; l < F->L; ++l) {
    mlp_forward(h, F, l, u_of(F, l), N);
    LAUNCH_SPLINE(k_rqs_forward, F->K, N, prm_of(F, l), F->ldp, u_of(F, l), u_of(F, l + 1), N, D,
                  F->layers[l].idx, F->K);
  }
  r = flow_backward(h, F, N, F->ld.as<float>());
  if (r) return r;
  GRA_HIP_CHECK(h, hipMemcpyAsync(grad, F->grad.p, 4 * F->nparam, hipMemcpyDeviceToHost, h->stream));
  if (logq)
    GRA_HIP_CHECK(h, hipMemcpyAsync(logq, F->ld.p, sizeof(float) * n, hipMemcpyDeviceToHost,
                                    h->stream));
  GRA_HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return GRA_OK;
}
