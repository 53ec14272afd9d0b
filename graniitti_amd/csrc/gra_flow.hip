// Normalizing-flow sampler entry points (rational-quadratic spline coupling
// layers). This unit is filled in by the flow implementation.
#include <hip/hip_runtime.h>

#include "gra_internal.h"

namespace gra_hip {
struct FlowState {};
void flow_release(FlowState* f) { delete f; }
}  // namespace gra_hip

using namespace gra_hip;

#define FLOW_UNAVAILABLE(h)                                               \
  do {                                                                    \
    if (!(h)) return set_global_error(GRA_E_ARG, "handle is NULL");       \
    return fail((h), GRA_E_STATE, "flow sampler is not available");       \
  } while (0)

extern "C" int gra_hip_flow_init(gra_handle* h, const gra_flow_cfg*) { FLOW_UNAVAILABLE(h); }
extern "C" int gra_hip_flow_train(gra_handle* h, int32_t, gra_flow_stats*) { FLOW_UNAVAILABLE(h); }
extern "C" int gra_hip_flow_sample(gra_handle* h, uint64_t, uint64_t, gra_flow_stats*) {
  FLOW_UNAVAILABLE(h);
}
extern "C" int gra_hip_flow_param_count(gra_handle* h, uint64_t*) { FLOW_UNAVAILABLE(h); }
extern "C" int gra_hip_flow_get_params(gra_handle* h, float*) { FLOW_UNAVAILABLE(h); }
extern "C" int gra_hip_flow_set_params(gra_handle* h, const float*) { FLOW_UNAVAILABLE(h); }
extern "C" int gra_hip_flow_forward(gra_handle* h, const float*, size_t, float*, float*) {
  FLOW_UNAVAILABLE(h);
}
extern "C" int gra_hip_flow_grad(gra_handle* h, const float*, const float*, size_t, float*, float*) {
  FLOW_UNAVAILABLE(h);
}
