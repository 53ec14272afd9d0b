// Integrand kernels in FP64 with the reference's operation order (built with
// -ffp-contract=off).
#define GRA_KERNEL_FP32 0
#define GRA_KNS k64
#include "gra_kernels.inc"
