// Launchers of the device kernels (one set per floating-point mode).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "gra_params.h"

namespace gra_hip {

#define GRA_DECLARE_INTEGRAND_LAUNCHERS(NS)                                                        \
  namespace NS {                                                                                   \
  size_t loop_record_bytes(const RunParams& P);                                                    \
  hipError_t launch_tree(const RunParams& P, const double* grid, unsigned int bins, uint64_t seed, \
                         uint32_t iter, uint64_t first, const double* xin, uint64_t n, double* W,  \
                         double* amp2, uint8_t* flags, double* vw, uint8_t* binsout, void* rec,    \
                         uint32_t* n_cand, hipStream_t st);                                        \
  hipError_t launch_screen(const RunParams& P, const void* nodes, const void* rec,                 \
                           const uint32_t* n_cand, uint64_t n_max, double* W, double* amp2,        \
                           uint8_t* flags, hipStream_t st);                                        \
  hipError_t launch_event_record(const RunParams& P, const double* grid, unsigned int bins,         \
                                 uint64_t seed, uint32_t iter, uint64_t first, const double* xin,   \
                                 const uint64_t* idx, const double* W, const double* vw,            \
                                 uint64_t n_ev, const unsigned long long* n_dev, void* out,         \
                                 hipStream_t st);                                                   \
  }

GRA_DECLARE_INTEGRAND_LAUNCHERS(k64)
GRA_DECLARE_INTEGRAND_LAUNCHERS(k32)

namespace k32 {
// Op-counting run of the FP32 Regge screening node over the first
// min(*n_cand, max_pts) records (max_pts <= 65536); out[12] (zeroed by the
// caller): points, nodes, accepted nodes, add, mul, fma, transcendental, compare
// counts, then add, mul, fma, transcendental of one accepted node
hipError_t launch_screen_opcount(const RunParams& P, const void* nodes, const void* rec,
                                 const uint32_t* n_cand, uint32_t max_pts, unsigned long long* out,
                                 hipStream_t st);
}  // namespace k32

// Scalars of one accumulated iteration, ahead of the [2][D][128] histograms:
// fsum, its compensation, f2sum, its compensation, max W, max f, and the counts of
// kinematics-valid, fiducial-valid, valid, NaN-weight and Inf-weight points
constexpr int kIterScalars = 11;

hipError_t launch_accumulate(const double* W, const double* vw, const uint8_t* bins,
                             const uint8_t* flags, uint64_t n, int D, int nblk, double* part_hist,
                             double* part_scal, double* out, hipStream_t st);
// p[0 .. n) = v (a kernel, not a memset node)
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);
hipError_t launch_map(const double* grid, int D, unsigned int bins, const double* u, uint64_t n,
                      double* x, double* vw, int32_t* b, hipStream_t st);
// Hit-or-miss of a generation pass: f = W vw (W alone if vw is NULL) of valid
// points against maxf, U from Philox (seed, 0xC0000000 | iter, first + i); the
// first max_out accepted indices (pass-local, ascending) go to idx; counts:
// trials, valid, accepted, overflow, NaN weights, Inf weights (all six
// written, no atomics). masks holds accept_mask_words(n) words, tiles
// accept_tile_words(n) entries; 0 < n < 2^32.
size_t accept_mask_words(uint64_t n);
size_t accept_tile_words(uint64_t n);
hipError_t launch_accept(const double* W, const double* vw, const uint8_t* flags, uint64_t n,
                         double maxf, uint64_t seed, uint32_t iter, uint64_t first, uint64_t* masks,
                         uint32_t* tiles, uint64_t* idx, uint64_t max_out,
                         unsigned long long* counts, hipStream_t st);
hipError_t launch_unweight(const double* w, uint64_t n, double wmax, uint64_t seed, uint8_t* acc,
                           unsigned long long* counts, hipStream_t st);

// Eikonal Fourier-Bessel tables: rows i = 0..nrows-1 at x_i = xmin + i xstep (or
// x_i = xs[i] when xs is given), out_i = factor * CSIntegral_q[pre_q J0(m(x_i) q_v) q_v],
// q_v = qmin + q qstep, m(x) = x (density, which = 0) or sqrt(max(x, 0)) (screening,
// which = 1).
hipError_t launch_fb_table(int which, const double* pre, int nf, double qmin, double qstep,
                           double xmin, double xstep, int nrows, double factor, double* out,
                           hipStream_t st, const double* xs);

}  // namespace gra_hip
