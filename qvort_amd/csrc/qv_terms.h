// qv_terms.h -- the FP64 multipole term shared by the velocity walk and the
// probe walk.
#pragma once
#include "qv_internal.h"

namespace qv {

// 1/sqrt(x) to about an ulp: the hardware estimate (relative error below
// 2^-22) refined by two Newton steps y <- y + y (1/2 - (x/2) y^2), each of
// which squares the error
__device__ inline double rsqrt64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = __builtin_fma(y, __builtin_fma(-(h * y), y, 0.5), y);
  y = __builtin_fma(y, __builtin_fma(-(h * y), y, 0.5), y);
  return y;
}

// cell term, tree.f90:400-411 (without Gamma/(4 pi)), regrouped as
// [circ x r - eps:C] / d^3 + 3 r x (C^T r) / d^5; r = x_i - (cent + shift),
// d2 = |r|^2; zero where not `on`
__device__ inline void cell64(bool on, const double r[3], double d2,
                              const double *__restrict__ m, double t[3]) {
  // the regrouped multipole has no reference rounding to keep: fused
  // multiply-adds allowed here
#pragma clang fp contract(fast)
  // lanes that are off use d2 = 1 (finite terms) and weight 0
  const double id = rsqrt64(on ? d2 : 1.);
  const double id2 = id * id;
  const double id3 = on ? id2 * id : 0.;
  const double id5 = id3 * id2;
  const double c0 = m[0], c1 = m[1], c2 = m[2];
  const double *C = m + 3;  // C[3*i + k] = Cij(i,k)
  // eps(:,k,l) C(k,l): antisymmetric part
  const double A0 = C[5] - C[7], A1 = C[6] - C[2], A2 = C[1] - C[3];
  // v_l = dot(r, C(:,l))
  const double v0 = (r[0] * C[0] + r[1] * C[3]) + r[2] * C[6];
  const double v1 = (r[0] * C[1] + r[1] * C[4]) + r[2] * C[7];
  const double v2 = (r[0] * C[2] + r[1] * C[5]) + r[2] * C[8];
  t[0] = ((c1 * r[2] - c2 * r[1]) - A0) * id3 + (3. * (r[1] * v2 - r[2] * v1)) * id5;
  t[1] = ((c2 * r[0] - c0 * r[2]) - A1) * id3 + (3. * (r[2] * v0 - r[0] * v2)) * id5;
  t[2] = ((c0 * r[1] - c1 * r[0]) - A2) * id3 + (3. * (r[0] * v1 - r[1] * v0)) * id5;
}

}  // namespace qv
