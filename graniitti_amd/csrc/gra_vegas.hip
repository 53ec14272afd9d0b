// VEGAS accumulation, grid-map and unweighting kernels (mode independent).
#include <hip/hip_runtime.h>

#include "gra_device.h"
#include "gra_kernels.h"

namespace {

// Deterministic accumulation of one iteration: each block (one wave) owns a
// contiguous range of points; lane l owns histogram bins l and l + 64; lanes
// read each point's (bin, f) with wave broadcasts in point order.
constexpr int kAccWave = 64;
__global__ __launch_bounds__(kAccWave) void k_accumulate(const double* __restrict__ W,
                                                         const double* __restrict__ vw,
                                                         const uint8_t* __restrict__ bins,
                                                         const uint8_t* __restrict__ flags,
                                                         uint64_t n, uint64_t chunk, int D,
                                                         double* __restrict__ part_hist,
                                                         double* __restrict__ part_scal) {
  extern __shared__ double hist[];  // [2][D][128]
  const int lane = threadIdx.x;
  for (int k = lane; k < 2 * D * GRA_VEGAS_BINS; k += kAccWave) hist[k] = 0.0;
  __syncthreads();
  const uint64_t p0 = (uint64_t)blockIdx.x * chunk;
  uint64_t p1 = p0 + chunk;
  if (p1 > n) p1 = n;
  double fs = 0, fc = 0, f2s = 0, f2c = 0, mW = 0, mf = 0;
  uint64_t c_kin = 0, c_fid = 0, c_val = 0, c_nan = 0;
  for (uint64_t base = p0; base < p1; base += kAccWave) {
    const uint64_t i = base + lane;
    double f = 0.0;
    if (i < p1) {
      const double Wi = W[i];
      f = Wi * vw[i];
      const double f2 = f * f;
      double y = f - fc, t = fs + y;
      fc = (t - fs) - y;
      fs = t;
      y = f2 - f2c;
      t = f2s + y;
      f2c = (t - f2s) - y;
      f2s = t;
      mW = Wi > mW ? Wi : mW;
      mf = f > mf ? f : mf;
      const uint8_t fl = flags[i];
      c_kin += (fl & GRA_FLAG_KIN_OK) ? 1 : 0;
      c_fid += (fl & GRA_FLAG_FID_OK) ? 1 : 0;
      c_val += ((fl & GRA_FLAG_KIN_OK) && (fl & GRA_FLAG_FID_OK)) ? 1 : 0;
      c_nan += (fl & GRA_FLAG_AMP_OK) ? 0 : 1;
    }
    // Non-zero contributions of this tile, in point order
    uint64_t m = __ballot(f != 0.0);
    while (m) {
      const int src = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const double fv = __shfl(f, src);
      const uint64_t ip = base + src;
      for (int j = 0; j < D; ++j) {
        const int b = bins[ip * D + j];
        if ((b & 63) == lane) {
          hist[j * GRA_VEGAS_BINS + b] += fv;
          hist[(D + j) * GRA_VEGAS_BINS + b] += fv * fv;
        }
      }
    }
  }
  __syncthreads();
  for (int k = lane; k < 2 * D * GRA_VEGAS_BINS; k += kAccWave)
    part_hist[(uint64_t)blockIdx.x * 2 * D * GRA_VEGAS_BINS + k] = hist[k];
  // Wave reduction of the scalars in a fixed lane order
  for (int off = 32; off > 0; off >>= 1) {
    const double ofs = __shfl_down(fs, off), ofc = __shfl_down(fc, off);
    const double of2s = __shfl_down(f2s, off), of2c = __shfl_down(f2c, off);
    // two-sum of the partial sums, carrying the compensations
    double s = fs + ofs, bp = s - fs, e = (fs - (s - bp)) + (ofs - bp);
    fs = s;
    fc = fc + ofc - e;
    s = f2s + of2s;
    bp = s - f2s;
    e = (f2s - (s - bp)) + (of2s - bp);
    f2s = s;
    f2c = f2c + of2c - e;
    const double omW = __shfl_down(mW, off), omf = __shfl_down(mf, off);
    mW = omW > mW ? omW : mW;
    mf = omf > mf ? omf : mf;
    c_kin += __shfl_down(c_kin, off);
    c_fid += __shfl_down(c_fid, off);
    c_val += __shfl_down(c_val, off);
    c_nan += __shfl_down(c_nan, off);
  }
  if (lane == 0) {
    double* o = part_scal + (uint64_t)blockIdx.x * 10;
    o[0] = fs;
    o[1] = fc;
    o[2] = f2s;
    o[3] = f2c;
    o[4] = mW;
    o[5] = mf;
    o[6] = (double)c_kin;
    o[7] = (double)c_fid;
    o[8] = (double)c_val;
    o[9] = (double)c_nan;
  }
}

// Fixed-order reduction of the block partials into gra_iter_stats layout:
// out[0..9] scalars, out[10 ..] hist [2][D][128].
__global__ void k_reduce(const double* __restrict__ part_hist, const double* __restrict__ part_scal,
                         int nblk, int D, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int H = 2 * D * GRA_VEGAS_BINS;
  if (t < H) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part_hist[(uint64_t)b * H + t];
    out[10 + t] = s;
  } else if (t == H) {
    double fs = 0, fc = 0, f2s = 0, f2c = 0, mW = 0, mf = 0, ck = 0, cf_ = 0, cv = 0, cn = 0;
    for (int b = 0; b < nblk; ++b) {
      const double* p = part_scal + (uint64_t)b * 10;
      double y = (p[0] - p[1]) - fc, tt = fs + y;
      fc = (tt - fs) - y;
      fs = tt;
      y = (p[2] - p[3]) - f2c;
      tt = f2s + y;
      f2c = (tt - f2s) - y;
      f2s = tt;
      mW = p[4] > mW ? p[4] : mW;
      mf = p[5] > mf ? p[5] : mf;
      ck += p[6];
      cf_ += p[7];
      cv += p[8];
      cn += p[9];
    }
    out[0] = fs;
    out[1] = fc;
    out[2] = f2s;
    out[3] = f2c;
    out[4] = mW;
    out[5] = mf;
    out[6] = ck;
    out[7] = cf_;
    out[8] = cv;
    out[9] = cn;
  }
}

// Map only (tests of the VEGAS map)
__global__ void k_map(const double* __restrict__ grid, int D, unsigned int bins,
                      const double* __restrict__ u, uint64_t n, double* __restrict__ x,
                      double* __restrict__ vw, int32_t* __restrict__ b) {
  __shared__ double xm[GRA_MAX_DIM * GRA_VEGAS_BINS];
  for (int k = threadIdx.x; k < D * GRA_VEGAS_BINS; k += blockDim.x) xm[k] = grid[k];
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double uu[GRA_MAX_DIM], xx[GRA_MAX_DIM];
  uint8_t bb[GRA_MAX_DIM];
  for (int j = 0; j < D; ++j) uu[j] = u[i * D + j];
  vw[i] = vegas_map(xm, D, bins, uu, xx, bb);
  for (int j = 0; j < D; ++j) {
    x[i * D + j] = xx[j];
    b[i * D + j] = (int32_t)bb[j] + 1;
  }
}

// Hit-or-miss unweighting, MGraniitti::SaveEvent (src/MGraniitti.cc:1651-1663):
// accept = U < w / wmax, overflow when w > wmax (then rejected).
__global__ void k_unweight(const double* __restrict__ w, uint64_t n, double wmax, uint64_t seed,
                           uint8_t* __restrict__ acc, unsigned long long* __restrict__ counts) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool hit = false, over = false;
  if (i < n) {
    double u;
    point_uniforms(seed, 0xFFFFFFFFu, i, 1, &u);
    hit = u < (w[i] / wmax);
    if (w[i] > wmax) {
      over = true;
      hit = false;
    }
    acc[i] = hit ? 1 : 0;
  }
  const uint64_t mh = __ballot(hit), mo = __ballot(over);
  if ((threadIdx.x & 63) == 0) {
    if (mh) atomicAdd(&counts[0], (unsigned long long)__popcll(mh));
    if (mo) atomicAdd(&counts[1], (unsigned long long)__popcll(mo));
  }
}

}  // namespace

namespace gra_hip {

hipError_t launch_accumulate(const double* W, const double* vw, const uint8_t* bins,
                             const uint8_t* flags, uint64_t n, int D, int nblk, double* part_hist,
                             double* part_scal, double* out, hipStream_t st) {
  const uint64_t chunk = (n + nblk - 1) / nblk;
  const size_t lds = sizeof(double) * 2 * D * GRA_VEGAS_BINS;
  hipLaunchKernelGGL(k_accumulate, dim3(nblk), dim3(kAccWave), lds, st, W, vw, bins, flags, n,
                     chunk, D, part_hist, part_scal);
  const int H = 2 * D * GRA_VEGAS_BINS + 1;
  hipLaunchKernelGGL(k_reduce, dim3((H + 255) / 256), dim3(256), 0, st, part_hist, part_scal, nblk,
                     D, out);
  return hipGetLastError();
}

hipError_t launch_map(const double* grid, int D, unsigned int bins, const double* u, uint64_t n,
                      double* x, double* vw, int32_t* b, hipStream_t st) {
  const unsigned int nb = (unsigned int)((n + 255) / 256);
  hipLaunchKernelGGL(k_map, dim3(nb), dim3(256), 0, st, grid, D, bins, u, n, x, vw, b);
  return hipGetLastError();
}

hipError_t launch_unweight(const double* w, uint64_t n, double wmax, uint64_t seed, uint8_t* acc,
                           unsigned long long* counts, hipStream_t st) {
  const unsigned int nb = (unsigned int)((n + 255) / 256);
  hipLaunchKernelGGL(k_unweight, dim3(nb), dim3(256), 0, st, w, n, wmax, seed, acc, counts);
  return hipGetLastError();
}

}  // namespace gra_hip
