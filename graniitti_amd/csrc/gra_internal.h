// Host-side internals of the C ABI: the handle, error reporting, buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "gra_hip.h"
#include "gra_params.h"

namespace gra_hip {

// Global (handle-less) error text, for gra_hip_create failures and host-only calls
int set_global_error(int code, const char* fmt, ...);
const char* global_error();

int grid_init(gra_grid* g, int dim, int bins);
int grid_optimize(gra_grid* g, double* f2mat, double lambda);

// VEGAS schedule of MGraniitti::SampleVegas + VEGAS (integration mode) over a
// pass that fills one iteration's accumulators; errors go to the global message
typedef gra_vegas_pass_fn VegasPass;
constexpr double kVegasChi2Restart = 50.0;  // src/MGraniitti.cc:1276
constexpr int kVegasMaxRestarts = 6;        // the reference retries without limit
int vegas_schedule(const gra_vegas_cfg* vc, gra_grid* grid, int dim, VegasPass pass, void* ctx,
                   gra_vegas_result* res);

// Device buffer that grows on demand
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);
  void release();
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct FlowState;  // normalizing-flow sampler (gra_flow.hip)
void flow_release(FlowState* f);

// 1 GeV^-1 in fermi: hbar c (include/Graniitti/MPDG.h:65-74)
inline double gev2fm() {
  const double c = 2.99792458E8, hbar = 6.582119514E-25;
  return hbar * c * 1E15;
}

// Eikonal table cache in the reference's file format (gra_eikcache.cpp):
// file name of table 0 (MBT) or 1 (MSA) for a configuration of Mandelstam s;
// a table of n_rows rows "x,re,im" (GRA_E_IO: missing, short or malformed;
// GRA_E_ARG: a field that is not a number); a table written with 15 significant
// digits per value
std::string eikonal_cache_name(const gra_cfg& c, double s, int table);
int eikonal_table_read(const char* path, int n_rows, std::vector<double>& rows);
int eikonal_table_write(const char* path, const std::vector<double>& rows);

// Kernel timing slots
enum TimeSlot { TS_TREE = 0, TS_SCREEN = 1, TS_ACCUM = 2, TS_FLOW = 3, TS_N = 4 };

}  // namespace gra_hip

struct gra_handle {
  gra_cfg cfg;
  RunParams P;
  hipStream_t stream = nullptr;
  int device = 0;
  std::string err;

  // eikonal tables (x, re, im rows, host copies after the 15-digit round trip)
  std::vector<double> mbt, msa;
  std::vector<NodeConst> nodes;
  gra_hip::DevBuf d_nodes;
  gra_hip::DevBuf d_res;  // ResDev table of PP[RES+CON]

  // per-point buffers
  gra_hip::DevBuf d_x, d_W, d_amp2, d_flags, d_vw, d_bins, d_rec, d_ncand, d_grid;
  gra_hip::DevBuf d_part_hist, d_part_scal, d_stats, d_aux;
  // generation passes: accepted indices, hit masks, tile counts/offsets, event
  // records (grow-only), and the pass counters in pinned host memory
  gra_hip::DevBuf d_gen_idx, d_gen_mask, d_gen_tile, d_gen_cnt, d_gen_ev;
  unsigned long long* h_gen_cnt = nullptr;
  // pinned staging of event records bound for pageable caller memory (grow-only)
  void* h_ev_stage = nullptr;
  size_t h_ev_stage_bytes = 0;
  std::vector<double> h_stats;
  // the grid edges last copied to d_grid: consecutive generation passes draw from
  // one grid, which is then not copied again
  std::vector<double> grid_host;
  const void* grid_dev_at = nullptr;

  // timing
  hipEvent_t ev[2 * gra_hip::TS_N] = {};
  double time_ms[gra_hip::TS_N] = {};
  uint64_t launches[gra_hip::TS_N] = {};

  // this device's allocation-free scratch (device address of the scratch array)
  void* scratch = nullptr;

  // CheckInfNan counters over the handle's life (include/Graniitti/MProcess.h:317-340)
  uint64_t n_nan_seen = 0, n_inf_seen = 0;

  gra_hip::FlowState* flow = nullptr;
};

namespace gra_hip {
int fail(gra_handle* h, int code, const char* fmt, ...);
// Integrand evaluation on device-resident inputs (x given) or a grid draw.
int run_integrand(gra_handle* h, const double* d_grid, unsigned int bins, uint64_t seed,
                  uint32_t iter, uint64_t first, const double* d_x, uint64_t n, double* d_W,
                  double* d_amp2, uint8_t* d_flags, double* d_vw, uint8_t* d_bins);
int ensure_point_buffers(gra_handle* h, uint64_t n);
// Buffers of a generation pass of n points storing up to cap events
int ensure_gen_buffers(gra_handle* h, uint64_t n, uint64_t cap);
// Copies n event records from device memory to the caller's host buffer and
// synchronizes the stream: directly into pinned memory, else through the
// handle's pinned staging buffer
int events_to_host(gra_handle* h, gra_event* dst, const void* d_src, uint64_t n);
// Adds a pass's NaN and Inf weight counts to the handle's; GRA_E_NUMERIC once
// either total exceeds GRA_MAX_NONFINITE
int count_nonfinite(gra_handle* h, uint64_t n_nan, uint64_t n_inf);
}  // namespace gra_hip

#define GRA_HIP_CHECK(h, call)                                                        \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return gra_hip::fail((h), GRA_E_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                       \
  } while (0)
