// qv_lia.h -- local (LIA) term of calc_velocity as a template over the scalar
// type: double for the velocity stage, Dual (value + one directional
// derivative, forward mode) for the adjoint's Jacobian of the local term.
//
// The double instantiation keeps the reference's operation order
// (derivatives.f90:11-86, timestep.f90:221-225 / 270-274); callers compile it
// without contraction.
#pragma once
#include <hip/hip_runtime.h>

namespace qv {
namespace ad {

// forward-mode dual number: f(x + eps dx) = f + eps f' dx.  Its own
// namespace, so that sqrt/log below are found by argument-dependent lookup
// and do not hide ::sqrt/::log for doubles.
struct Dual {
  double v, d;
  __device__ Dual() : v(0.), d(0.) {}
  __device__ Dual(double a) : v(a), d(0.) {}
  __device__ Dual(double a, double b) : v(a), d(b) {}
};
__device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ inline Dual sqrt(Dual a) {
  const double r = ::sqrt(a.v);
  return Dual(r, a.d / (2. * r));
}
__device__ inline Dual log(Dual a) { return Dual(::log(a.v), a.d / a.v); }

}  // namespace ad
using ad::Dual;

// The templates below are host and device code: the double instantiation
// also serves the host line operations (pinsert's curvature, precon's
// tangents), which must form s' and s'' exactly as the velocity does.
template <class T>
__host__ __device__ inline T lia_sq(T a) { return a * a; }

// dist_gen, derivatives.f90:79-86
template <class T>
__host__ __device__ inline T lia_dist(const T *a, const T *b) {
  return sqrt((lia_sq(a[0] - b[0]) + lia_sq(a[1] - b[1])) + lia_sq(a[2] - b[2]));
}

// get_deriv_1 (s1) and get_deriv_2 (s2), derivatives.f90:11-75
template <class T>
__host__ __device__ void lia_derivs(int order, const T x[3], const T gi[3], const T gb[3],
                                    const T gii[3], const T gbb[3], T s1[3], T s2[3]) {
  const T di = lia_dist(x, gi), db = lia_dist(x, gb);
  if (order != 4) {
    const T den = (T(2.) * db) * di;
    for (int k = 0; k < 3; ++k) s1[k] = ((db * gi[k] + (di - db) * x[k]) - di * gb[k]) / den;
    for (int k = 0; k < 3; ++k)
      s2[k] = T(2.) * ((gi[k] / (di * (di + db)) - x[k] / (di * db)) + gb[k] / (db * (di + db)));
  } else {
    const T dii = lia_dist(gii, gi), dbb = lia_dist(gbb, gb);
    {
      T A = db * (di * di) + (db * di) * dii;
      A = A / (((dbb * (dbb + db)) * ((dbb + db) + di)) * (((dbb + db) + di) + dii));
      T B = ((-(dbb * (di * di)) - db * (di * di)) - (dbb * di) * dii) - (db * di) * dii;
      B = B / (((dbb * db) * (db + di)) * ((db + di) + dii));
      T D = (((dbb * db) * di + (db * db) * di) + (dbb * db) * dii) + (db * db) * dii;
      D = D / (((di * dii) * (db + di)) * ((dbb + db) + di));
      T E = -(di * (db * db)) - (dbb * db) * di;
      E = E / (((dii * (di + dii)) * ((db + di) + dii)) * (((dbb + db) + di) + dii));
      const T C = -(((A + B) + D) + E);
      for (int k = 0; k < 3; ++k)
        s1[k] = (((A * gbb[k] + B * gb[k]) + C * x[k]) + D * gi[k]) + E * gii[k];
    }
    {
      T A = T(2.) * (((-((T(2.) * db) * di) + di * di) - db * dii) + di * dii);
      A = A / (((dbb * (dbb + db)) * ((dbb + db) + di)) * (((dbb + db) + di) + dii));
      T B = T(2.) * ((((((T(2.) * dbb) * di + (T(2.) * db) * di) - di * di) + dbb * dii) +
                      db * dii) - di * dii);
      B = B / (((dbb * db) * (db + di)) * ((db + di) + dii));
      T D = T(2.) * (((((-(dbb * db) - db * db) + dbb * di) + (T(2.) * db) * di) + dbb * dii) +
                     (T(2.) * db) * dii);
      D = D / (((di * dii) * (db + di)) * ((dbb + db) + di));
      T E = T(2.) * (((dbb * db + db * db) - dbb * di) - (T(2.) * db) * di);
      E = E / (((dii * (di + dii)) * ((db + di) + dii)) * (((dbb + db) + di) + dii));
      const T C = -(((A + B) + D) + E);
      for (int k = 0; k < 3; ++k)
        s2[k] = (((A * gbb[k] + B * gb[k]) + C * x[k]) + D * gi[k]) + E * gii[k];
    }
  }
}

// u_LI = Gamma/(4 pi) ln((1.21/a) sqrt(l_i l_b)) s' x s''
template <class T>
__device__ void lia_term(double qc, double corea, int order, const T x[3], const T gi[3],
                         const T gb[3], const T gii[3], const T gbb[3], T u[3]) {
  T s1[3], s2[3];
  lia_derivs<T>(order, x, gi, gb, gii, gbb, s1, s2);
  const T li = lia_dist(x, gi), lb = lia_dist(x, gb);
  const T beta = T(qc / (4. * 3.14159265358979324)) * log(T(1.21 / corea) * sqrt(li * lb));
  u[0] = beta * (s1[1] * s2[2] - s1[2] * s2[1]);
  u[1] = beta * (s1[2] * s2[0] - s1[0] * s2[2]);
  u[2] = beta * (s1[0] * s2[1] - s1[1] * s2[0]);
}

}  // namespace qv
