// Integrand kernels with FP64 kinematics and an FP32 amplitude in the screening
// loop (contraction on).
#define GRA_KERNEL_FP32 1
#define GRA_KNS k32
#include "gra_kernels.inc"
