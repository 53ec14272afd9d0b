// VEGAS accumulation, grid-map and unweighting kernels (mode independent).
#include <hip/hip_runtime.h>

#include "gra_device.h"
#include "gra_kernels.h"

namespace {

using gra_hip::kIterScalars;

// Deterministic accumulation of one iteration. Each block owns a contiguous
// range of points and has kAccWaves waves that walk the same points; wave w
// owns the dimensions w, w + kAccWaves, ... and keeps their histograms of f and
// f^2 in its own LDS slice. For each batch of 64 points and each owned
// dimension the wave sorts the keys (bin, lane) with a bitonic network, sums f
// and f^2 over each run of equal bins with a segmented scan, and the last lane
// of each run adds the run's sums to its bin. Keys are unique, so the order of
// every addition is fixed and the histograms are the same on every run, without
// atomics. Wave 0 also sums the scalars.
constexpr int kAccWave = 64;
constexpr int kAccWaves = 4;

// Ascending bitonic sort of one int key per lane across the wave
__device__ __forceinline__ int wave_sort(int key, int lane) {
#pragma unroll
  for (int k = 2; k <= kAccWave; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int other = __shfl_xor(key, j);
      const bool up = (lane & k) == 0, lower = (lane & j) == 0;
      key = (lower == up) ? min(key, other) : max(key, other);
    }
  }
  return key;
}

template <int DT>
__global__ __launch_bounds__(kAccWave* kAccWaves) void k_accumulate(
    const double* __restrict__ W, const double* __restrict__ vw, const uint8_t* __restrict__ bins,
    const uint8_t* __restrict__ flags, uint64_t n, uint64_t chunk, int D,
    double* __restrict__ part_hist, double* __restrict__ part_scal) {
  constexpr int DPW = (DT + kAccWaves - 1) / kAccWaves;  // dimensions per wave
  constexpr int NB = GRA_VEGAS_BINS;
  __shared__ double hist[kAccWaves][DPW][2][NB];
  const int lane = threadIdx.x & (kAccWave - 1);
  const int wave = threadIdx.x / kAccWave;  // wave-uniform
  double(*hw)[2][NB] = hist[wave];
#pragma unroll
  for (int q = 0; q < DPW; ++q)
    for (int b = lane; b < NB; b += kAccWave) hw[q][0][b] = hw[q][1][b] = 0.0;
  const uint64_t p0 = (uint64_t)blockIdx.x * chunk;
  uint64_t p1 = p0 + chunk;
  if (p1 > n) p1 = n;
  double fs = 0, fc = 0, f2s = 0, f2c = 0, mW = 0, mf = 0;
  uint64_t c_kin = 0, c_fid = 0, c_val = 0, c_nan = 0, c_inf = 0;
  for (uint64_t base = p0; base < p1; base += kAccWave) {
    const uint64_t i = base + lane;
    double f = 0.0;
    int bb[DPW];
#pragma unroll
    for (int q = 0; q < DPW; ++q) bb[q] = 0;
    if (i < p1) {
      const double Wi = W[i];
      f = Wi * vw[i];
      if (wave == 0) {
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
        c_val += ((fl & GRA_FLAG_KIN_OK) && (fl & GRA_FLAG_FID_OK) && (fl & GRA_FLAG_VETO_OK)) ? 1
                                                                                           : 0;
        const bool bad = !(fl & GRA_FLAG_AMP_OK);
        c_nan += (bad && !(fl & GRA_FLAG_INF)) ? 1 : 0;
        c_inf += (bad && (fl & GRA_FLAG_INF)) ? 1 : 0;
      }
#pragma unroll
      for (int q = 0; q < DPW; ++q) {
        const int j = wave + q * kAccWaves;
        if (j < D) bb[q] = bins[i * D + j];
      }
    }
    if (__ballot(f != 0.0) == 0) continue;  // wave-uniform
#pragma unroll
    for (int q = 0; q < DPW; ++q) {
      if (wave + q * kAccWaves >= D) break;  // wave-uniform
      // zero contributions sort behind every bin (key >= 128 * 64) and are dropped
      const int key = wave_sort(f != 0.0 ? (bb[q] << 6) | lane : (NB << 6) | lane, lane);
      const int bin = key >> 6;
      double s = __shfl(f, key & 63);
      double s2 = s * s;
      // segmented inclusive scan over runs of equal bins (sorted, so a run is contiguous)
#pragma unroll
      for (int d = 1; d < kAccWave; d <<= 1) {
        const int ob = __shfl_up(bin, d);
        const double os = __shfl_up(s, d), os2 = __shfl_up(s2, d);
        if (lane >= d && ob == bin) {
          s += os;
          s2 += os2;
        }
      }
      const int nb = __shfl_down(bin, 1);
      if (bin < NB && (lane == kAccWave - 1 || nb != bin)) {
        hw[q][0][bin] += s;
        hw[q][1][bin] += s2;
      }
    }
  }
  const uint64_t H = 2 * (uint64_t)D * NB;
  double* ph = part_hist + (uint64_t)blockIdx.x * H;
#pragma unroll
  for (int q = 0; q < DPW; ++q) {
    const int j = wave + q * kAccWaves;
    if (j < D)
      for (int b = lane; b < NB; b += kAccWave) {
        ph[j * NB + b] = hw[q][0][b];
        ph[(D + j) * NB + b] = hw[q][1][b];
      }
  }
  if (wave != 0) return;
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
    c_inf += __shfl_down(c_inf, off);
  }
  if (lane == 0) {
    double* o = part_scal + (uint64_t)blockIdx.x * kIterScalars;
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
    o[10] = (double)c_inf;
  }
}

// Fixed-order two-level reduction of the block partials. Stage 1: group g sums
// blocks [g nblk / G, (g + 1) nblk / G) in order; stage 2 sums the groups in
// order into gra_iter_stats layout: out[0 .. kIterScalars - 1] scalars, then the
// histograms [2][D][128].
constexpr int kRedGroups = 32;

// Kahan-add of one block's (sum, compensation) pair and the other scalars
__device__ __forceinline__ void scal_add(double* a, const double* p) {
  double y = (p[0] - p[1]) - a[1], tt = a[0] + y;
  a[1] = (tt - a[0]) - y;
  a[0] = tt;
  y = (p[2] - p[3]) - a[3];
  tt = a[2] + y;
  a[3] = (tt - a[2]) - y;
  a[2] = tt;
  a[4] = p[4] > a[4] ? p[4] : a[4];
  a[5] = p[5] > a[5] ? p[5] : a[5];
  for (int k = 6; k < kIterScalars; ++k) a[k] += p[k];
}

__global__ void k_reduce1(const double* __restrict__ part_hist, const double* __restrict__ part_scal,
                          int nblk, int H, double* __restrict__ tmp, double* __restrict__ tmp_scal) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  const int b0 = (int)((long long)nblk * g / kRedGroups);
  const int b1 = (int)((long long)nblk * (g + 1) / kRedGroups);
  if (t < H) {
    double s = 0.0;
    for (int b = b0; b < b1; ++b) s += part_hist[(uint64_t)b * H + t];
    tmp[(uint64_t)g * H + t] = s;
  } else if (t == H) {
    double a[kIterScalars] = {};
    for (int b = b0; b < b1; ++b) scal_add(a, part_scal + (uint64_t)b * kIterScalars);
    for (int k = 0; k < kIterScalars; ++k) tmp_scal[g * kIterScalars + k] = a[k];
  }
}

__global__ void k_reduce2(const double* __restrict__ tmp, const double* __restrict__ tmp_scal, int H,
                          double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < H) {
    double s = 0.0;
    for (int g = 0; g < kRedGroups; ++g) s += tmp[(uint64_t)g * H + t];
    out[kIterScalars + t] = s;
  } else if (t == H) {
    double a[kIterScalars] = {};
    for (int g = 0; g < kRedGroups; ++g) scal_add(a, tmp_scal + g * kIterScalars);
    for (int k = 0; k < kIterScalars; ++k) out[k] = a[k];
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

// Hit-or-miss of the generation pass: f = W vw of valid points against maxf.
// Each wave stores five ballots of its 64 points (word 5 g + m of masks for
// group g): hits, valid points, overflows (f > maxf), NaN weights and Inf
// weights. k_tile_count, k_tile_scan and k_tile_write then count them and
// compact the hit indices in index order, with no atomics, so the result does
// not depend on the schedule.
constexpr int kAccMasks = 5;
__global__ void k_accept(const double* __restrict__ W, const double* __restrict__ vw,
                         const uint8_t* __restrict__ flags, uint64_t n, double maxf, uint64_t seed,
                         uint32_t iter, uint64_t first, uint64_t* __restrict__ masks) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = false, hit = false, over = false, is_nan = false, is_inf = false;
  if (i < n) {
    const uint8_t fl = flags[i];
    is_nan = !(fl & GRA_FLAG_AMP_OK) && !(fl & GRA_FLAG_INF);
    is_inf = !(fl & GRA_FLAG_AMP_OK) && (fl & GRA_FLAG_INF);
    valid = (fl & (GRA_FLAG_KIN_OK | GRA_FLAG_FID_OK | GRA_FLAG_VETO_OK)) ==
            (GRA_FLAG_KIN_OK | GRA_FLAG_FID_OK | GRA_FLAG_VETO_OK);
    const double f = vw ? W[i] * vw[i] : W[i];
    double u;
    point_uniforms(seed, 0xC0000000u | (iter & 0x3FFFFFFFu), first + i, 1, &u);
    hit = u < (f / maxf);
    if (f > maxf) {
      over = true;
      hit = false;
    }
    hit = hit && valid;
  }
  const uint64_t m[kAccMasks] = {__ballot(hit), __ballot(valid), __ballot(over), __ballot(is_nan),
                                 __ballot(is_inf)};
  const int lane = threadIdx.x & 63;
  // lane 0 holds the wave's smallest index: the wave has points iff lane 0 does
  if (lane < kAccMasks && (i - lane) < n) {
    uint64_t v = m[0];
#pragma unroll
    for (int k = 1; k < kAccMasks; ++k) v = (lane == k) ? m[k] : v;
    masks[kAccMasks * ((i - lane) >> 6) + lane] = v;
  }
}

// Ordered compaction of the hit masks: a tile is 256 wpt groups of 64 points
// (one block of 256 threads, wpt consecutive groups per thread). wpt is 1 up to
// 2^28 points, so that a pass of 2^20 points spreads over 64 blocks, and 16
// above (accept_wpt)
constexpr int kAccMaxTiles = 1024 * 16;  // k_tile_scan: 1024 threads x 16 tiles

// Exclusive prefix sum of v over the 256 threads of a block; total in *tot.
// sh holds 4 entries; a second call in the same kernel must follow a barrier.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  uint32_t base = 0, all = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    if (w < wave) base += sh[w];
    all += sh[w];
  }
  *tot = all;
  return base + x - v;
}

// Per tile: the five mask counts (tile_cnt[5 t + m])
__global__ __launch_bounds__(256) void k_tile_count(const uint64_t* __restrict__ masks,
                                                    uint64_t ngroups, int wpt,
                                                    uint32_t* __restrict__ tile_cnt) {
  __shared__ uint32_t sh[kAccMasks][4];
  const uint64_t g0 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * wpt;
  uint32_t c[kAccMasks] = {};
  for (int j = 0; j < wpt; ++j) {
    if (g0 + j >= ngroups) break;
#pragma unroll
    for (int k = 0; k < kAccMasks; ++k) c[k] += (uint32_t)__popcll(masks[kAccMasks * (g0 + j) + k]);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kAccMasks; ++k) {
    uint32_t x = c[k];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d);
    if (lane == 0) sh[k][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x < kAccMasks)
    tile_cnt[kAccMasks * blockIdx.x + threadIdx.x] =
        sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
}

// One block: exclusive offsets of the tiles' hits, and the pass counters
// (trials, valid, accepted, overflow, NaN, Inf)
__global__ __launch_bounds__(1024) void k_tile_scan(const uint32_t* __restrict__ tile_cnt,
                                                    int ntiles, uint64_t n,
                                                    uint32_t* __restrict__ tile_off,
                                                    unsigned long long* __restrict__ counts) {
  __shared__ uint32_t sh[16];
  __shared__ unsigned long long tot[kAccMasks];
  if (threadIdx.x < kAccMasks) tot[threadIdx.x] = 0;
  const int t0 = threadIdx.x * 16;
  uint32_t c = 0;
  unsigned long long o[kAccMasks - 1] = {};
  for (int j = 0; j < 16; ++j)
    if (t0 + j < ntiles) {
      c += tile_cnt[kAccMasks * (t0 + j)];
#pragma unroll
      for (int k = 1; k < kAccMasks; ++k) o[k - 1] += tile_cnt[kAccMasks * (t0 + j) + k];
    }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
#pragma unroll
  for (int k = 1; k < kAccMasks; ++k)
    if (o[k - 1]) atomicAdd(&tot[k], o[k - 1]);
  uint32_t base = 0, all = 0;
  for (int w = 0; w < 16; ++w) {
    if (w < wave) base += sh[w];
    all += sh[w];
  }
  uint32_t run = base + x - c;
  for (int j = 0; j < 16; ++j)
    if (t0 + j < ntiles) {
      tile_off[t0 + j] = run;
      run += tile_cnt[kAccMasks * (t0 + j)];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    counts[0] = n;
    counts[1] = tot[1];
    counts[2] = all;
    counts[3] = tot[2];
    counts[4] = tot[3];
    counts[5] = tot[4];
  }
}

__global__ __launch_bounds__(256) void k_tile_write(const uint64_t* __restrict__ masks,
                                                    uint64_t ngroups, int wpt,
                                                    const uint32_t* __restrict__ tile_off,
                                                    uint64_t* __restrict__ idx, uint64_t max_out) {
  __shared__ uint32_t sh[4];
  const uint64_t g0 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * wpt;
  uint32_t c = 0;
  for (int j = 0; j < wpt; ++j) {
    if (g0 + j >= ngroups) break;
    c += (uint32_t)__popcll(masks[kAccMasks * (g0 + j)]);
  }
  uint32_t tot;
  uint64_t pos = (uint64_t)tile_off[blockIdx.x] + block_exclusive_scan(c, sh, &tot);
  for (int j = 0; j < wpt; ++j) {
    if (g0 + j >= ngroups || pos >= max_out) break;
    uint64_t m = masks[kAccMasks * (g0 + j)];
    while (m) {
      const int b = __ffsll((unsigned long long)m) - 1;
      if (pos < max_out) idx[pos] = ((g0 + j) << 6) + (uint64_t)b;
      ++pos;
      m &= m - 1;
    }
  }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

namespace gra_hip {

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const unsigned int nb = (unsigned int)((n + 255) / 256);
  hipLaunchKernelGGL(k_fill_u32, dim3(nb), dim3(256), 0, st, p, v, (uint64_t)n);
  return hipGetLastError();
}

size_t accept_mask_words(uint64_t n) { return kAccMasks * (size_t)((n + 63) / 64); }
static int accept_wpt(uint64_t n) { return n <= (1ull << 28) ? 1 : 16; }
static size_t accept_tiles(uint64_t n) {
  const uint64_t tile_words = 256 * (uint64_t)accept_wpt(n);
  return (size_t)(((n + 63) / 64 + tile_words - 1) / tile_words);
}
size_t accept_tile_words(uint64_t n) { return (kAccMasks + 1) * accept_tiles(n); }

hipError_t launch_accept(const double* W, const double* vw, const uint8_t* flags, uint64_t n,
                         double maxf, uint64_t seed, uint32_t iter, uint64_t first, uint64_t* masks,
                         uint32_t* tiles, uint64_t* idx, uint64_t max_out,
                         unsigned long long* counts, hipStream_t st) {
  if (n == 0 || n > 0xFFFFFFFFull) return hipErrorInvalidValue;
  const uint64_t ng = (n + 63) / 64;
  const size_t nt = accept_tiles(n);
  if (nt > (size_t)kAccMaxTiles) return hipErrorInvalidValue;
  const unsigned int nb = (unsigned int)((n + 255) / 256);
  hipLaunchKernelGGL(k_accept, dim3(nb), dim3(256), 0, st, W, vw, flags, n, maxf, seed, iter, first,
                     masks);
  const int wpt = accept_wpt(n);
  hipLaunchKernelGGL(k_tile_count, dim3((unsigned int)nt), dim3(256), 0, st, masks, ng, wpt, tiles);
  uint32_t* off = tiles + kAccMasks * nt;
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, st, tiles, (int)nt, n, off, counts);
  hipLaunchKernelGGL(k_tile_write, dim3((unsigned int)nt), dim3(256), 0, st, masks, ng, wpt, off,
                     idx, max_out);
  return hipGetLastError();
}

hipError_t launch_accumulate(const double* W, const double* vw, const uint8_t* bins,
                             const uint8_t* flags, uint64_t n, int D, int nblk, double* part_hist,
                             double* part_scal, double* out, hipStream_t st) {
  const uint64_t chunk = (n + nblk - 1) / nblk;
  if (D <= 8)
    hipLaunchKernelGGL(k_accumulate<8>, dim3(nblk), dim3(kAccWave * kAccWaves), 0, st, W, vw,
                       bins, flags, n, chunk, D, part_hist, part_scal);
  else if (D <= 14)
    hipLaunchKernelGGL(k_accumulate<14>, dim3(nblk), dim3(kAccWave * kAccWaves), 0, st, W, vw,
                       bins, flags, n, chunk, D, part_hist, part_scal);
  else
    hipLaunchKernelGGL(k_accumulate<16>, dim3(nblk), dim3(kAccWave * kAccWaves), 0, st, W, vw,
                       bins, flags, n, chunk, D, part_hist, part_scal);
  const int H = 2 * D * GRA_VEGAS_BINS;
  // stage-1 partial sums use the tail of part_hist's allocation: [G][H] then [G][kIterScalars]
  double* tmp = part_hist + (uint64_t)nblk * H;
  double* tmp_scal = tmp + (uint64_t)kRedGroups * H;
  hipLaunchKernelGGL(k_reduce1, dim3((H + 1 + 255) / 256, kRedGroups), dim3(256), 0, st,
                     part_hist, part_scal, nblk, H, tmp, tmp_scal);
  hipLaunchKernelGGL(k_reduce2, dim3((H + 1 + 255) / 256), dim3(256), 0, st, tmp, tmp_scal, H, out);
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
