// Eikonal Fourier-Bessel tables (MEikonal::S3Density / S3Screening,
// src/MEikonal.cc:177-245), built with -ffp-contract=off so each row's
// composite-Simpson sum runs in the reference's order.
#include <hip/hip_runtime.h>

#include "gra_device.h"
#include "gra_kernels.h"

namespace {

// J0 by the rational / asymptotic fits of Abramowitz-Stegun 9.4 as tabulated by
// Clenshaw (NPL Math. Tables 5, 1962); the same approximation the reference
// evaluates (BESSJ0, libs/extmath.hpp:12-49).
__device__ __forceinline__ double bessj0(double X) {
  const double P1 = 1.0, P2 = -0.1098628627E-2, P3 = 0.2734510407E-4, P4 = -0.2073370639E-5,
               P5 = 0.2093887211E-6, Q1 = -0.1562499995E-1, Q2 = 0.1430488765E-3,
               Q3 = -0.6911147651E-5, Q4 = 0.7621095161E-6, Q5 = -0.9349451520E-7,
               R1 = 57568490574.0, R2 = -13362590354.0, R3 = 651619640.7, R4 = -11214424.18,
               R5 = 77392.33017, R6 = -184.9052456, S1 = 57568490411.0, S2 = 1029532985.0,
               S3 = 9494680.718, S4 = 59272.64853, S5 = 267.8532712, S6 = 1.0;
  if (X == 0.0) return 1.0;
  const double AX = fabs(X);
  if (AX < 8.0) {
    const double Y = X * X;
    const double FR = R1 + Y * (R2 + Y * (R3 + Y * (R4 + Y * (R5 + Y * R6))));
    const double FS = S1 + Y * (S2 + Y * (S3 + Y * (S4 + Y * (S5 + Y * S6))));
    return FR / FS;
  }
  const double Z = 8.0 / AX;
  const double Y = Z * Z;
  const double XX = AX - 0.785398164;
  const double FP = P1 + Y * (P2 + Y * (P3 + Y * (P4 + Y * P5)));
  const double FQ = Q1 + Y * (Q2 + Y * (Q3 + Y * (Q4 + Y * Q5)));
  return sqrt(0.636619772 / AX) * (FP * cos(XX) - Z * FQ * sin(XX));
}

// One thread per table row; the quadrature nodes' prefactors are read as
// wave-uniform loads.
__global__ __launch_bounds__(256) void k_fb_table(int which, const double* __restrict__ pre, int nf,
                                                  double qmin, double qstep, double xmin,
                                                  double xstep, int nrows, double factor,
                                                  const double* __restrict__ xs,
                                                  double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const double xv = xs ? xs[i] : xmin + i * xstep;
  const double m = (which == 0) ? xv : msq(xv);
  double f0r = 0, f0i = 0, fNr = 0, fNi = 0, er = 0, ei = 0, odr = 0, odi = 0;
  for (int q = 0; q <= nf; ++q) {
    const double qv = qmin + q * qstep;
    const double j0 = bessj0(m * qv);
    const double fr = (pre[2 * q] * j0) * qv, fi = (pre[2 * q + 1] * j0) * qv;
    if (q == 0) {
      f0r = fr;
      f0i = fi;
    } else if (q == nf) {
      fNr = fr;
      fNi = fi;
    } else if (q & 1) {
      odr += fr;
      odi += fi;
    } else {
      er += fr;
      ei += fi;
    }
  }
  // CSIntegral: h (f0 + 2 I_even + 4 I_odd + fN) / 3, include/Graniitti/MMath.h:378-394
  const double sr = ((f0r + 2.0 * er) + 4.0 * odr) + fNr;
  const double si = ((f0i + 2.0 * ei) + 4.0 * odi) + fNi;
  out[2 * i] = ((sr * qstep) / 3.0) * factor;
  out[2 * i + 1] = ((si * qstep) / 3.0) * factor;
}

}  // namespace

namespace gra_hip {

hipError_t launch_fb_table(int which, const double* pre, int nf, double qmin, double qstep,
                           double xmin, double xstep, int nrows, double factor, double* out,
                           hipStream_t st, const double* xs) {
  hipLaunchKernelGGL(k_fb_table, dim3((nrows + 255) / 256), dim3(256), 0, st, which, pre, nf, qmin,
                     qstep, xmin, xstep, nrows, factor, xs, out);
  return hipGetLastError();
}

}  // namespace gra_hip
