// ref_harness.cc -- TEST INFRASTRUCTURE ONLY.
//
// C entry points around the parts of the reference (GRANIITTI v1.051) that
// compile from their own headers without LHAPDF/HepMC3:
//   include/Graniitti/MVEGAS.h   (VEGASData grid, Rebin, OptimizeGrid)
//   include/Graniitti/MMath.h    (Simpson 3/8 weights, CSIntegral, CheckEMC)
//   include/Graniitti/M4Vec.h    (4-vector algebra)
//   libs/extmath.hpp             (BESSJ0)
// Built by oracle/Makefile from the sources where they lie under the reference
// root into oracle/_ref/ (never committed); used only to write and check the
// golden vectors of tests/golden/ and to cross-check the C restatement.
#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "Graniitti/M4Vec.h"
#include "Graniitti/MMath.h"
#include "Graniitti/MVEGAS.h"
#include "extmath.hpp"

extern "C" {

// Uniform grid after VEGASData::InitRegion + Init(0): xmat dimension-major [dim][bins]
int ref_grid_init(int dim, int bins, double* xmat_out) {
  gra::VEGASPARAM p;
  p.BINS = bins;
  gra::VEGASData v;
  v.InitRegion(dim);
  v.Init(0, p);
  for (int j = 0; j < dim; ++j)
    for (int i = 0; i < bins; ++i) xmat_out[j * bins + i] = v.xmat[i][j];
  return 0;
}

// One OptimizeGrid step from a given grid and f2mat (both [dim][bins]); writes
// the new grid and the smoothed f2mat back.
int ref_grid_optimize(int dim, int bins, double lambda, double* xmat, double* f2mat) {
  gra::VEGASPARAM p;
  p.BINS = bins;
  p.LAMBDA = lambda;
  gra::VEGASData v;
  v.InitRegion(dim);
  v.Init(0, p);
  for (int j = 0; j < dim; ++j)
    for (int i = 0; i < bins; ++i) {
      v.xmat[i][j] = xmat[j * bins + i];
      v.f2mat[i][j] = f2mat[j * bins + i];
    }
  v.OptimizeGrid(p);
  for (int j = 0; j < dim; ++j)
    for (int i = 0; i < bins; ++i) {
      xmat[j * bins + i] = v.xmat[i][j];
      f2mat[j * bins + i] = v.f2mat[i][j];
    }
  return 0;
}

// Simpson 3/8 2D weight matrix, (M+1) x (N+1) row-major
int ref_simpson38_weight2d(int M, int N, double* W) {
  try {
    const auto w = gra::math::Simpson38Weight2D(M, N);
    for (int i = 0; i <= M; ++i)
      for (int j = 0; j <= N; ++j) W[i * (N + 1) + j] = w[i][j];
  } catch (...) {
    return -1;
  }
  return 0;
}

// Simpson 3/8 2D integral of a complex matrix f (interleaved re, im)
int ref_simpson38_integral2d(int M, int N, const double* f_reim, double hM, double hN,
                             double* out) {
  try {
    gra::MMatrix<std::complex<double>> f(M + 1, N + 1, 0.0);
    for (int i = 0; i <= M; ++i)
      for (int j = 0; j <= N; ++j)
        f[i][j] = std::complex<double>(f_reim[2 * (i * (N + 1) + j)],
                                       f_reim[2 * (i * (N + 1) + j) + 1]);
    const auto W = gra::math::Simpson38Weight2D(M, N);
    const auto I = gra::math::Simpson38Integral2D(f, W, hM, hN);
    out[0] = I.real();
    out[1] = I.imag();
  } catch (...) {
    return -1;
  }
  return 0;
}

// Composite Simpson integral of complex samples (N intervals, N even)
int ref_csintegral(int N, const double* f_reim, double h, double* out) {
  try {
    std::vector<std::complex<double>> f(N + 1);
    for (int i = 0; i <= N; ++i) f[i] = std::complex<double>(f_reim[2 * i], f_reim[2 * i + 1]);
    const auto I = gra::math::CSIntegral(f, h);
    out[0] = I.real();
    out[1] = I.imag();
  } catch (...) {
    return -1;
  }
  return 0;
}

int ref_checkemc(double e, double x, double y, double z) {
  return gra::math::CheckEMC(gra::M4Vec(x, y, z, e)) ? 1 : 0;
}

double ref_bessj0(double x) { return BESSJ0(x); }

// M4Vec observables of (E, x, y, z): M2, M, Pt, Mt, Eta, Rap
void ref_m4vec(double e, double x, double y, double z, double* out) {
  const gra::M4Vec v(x, y, z, e);
  out[0] = v.M2();
  out[1] = v.M();
  out[2] = v.Pt();
  out[3] = v.Mt();
  out[4] = v.Eta();
  out[5] = v.Rap();
}

}  // extern "C"
