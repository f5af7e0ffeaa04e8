// qv_adjoint.hip -- derivatives of the direct-sum velocity u = U(x)
// (velocity='BS': LIA term + biot_savart + biot_savart_shift,
// timestep.f90:218-259, 462-507) with respect to the point positions.
//
//   VJP  xbar = (dU/dx)^T w      (adjoint, for gradients of objectives)
//   JVP  jv   = (dU/dx) v        (tangent, for implicit solvers)
//
// There is no counterpart in the reference.  Each source pair term is the
// straight-segment field K(R1, R2) = Gamma/(4 pi) F c with R1 = x_j + s - x_i,
// R2 = g_j + s - x_i, a = |R1|, b = |R2|, c = R1 x R2,
// F = (a+b) / (a b (a b + R1.R2))  (the same function as the pair formula of
// biot_savart, in the form of tree_walk's leaf, tree.f90:380-392).  With
// L = w.K:
//   dL/dR1 = G [ (w.c) dF/dR1 + F (R2 x w) ],  dL/dR2 = G [ (w.c) dF/dR2 + F (w x R1) ]
//   dF/dR1 = F [ -b/(a^2 (a+b)) R1 - b/(a q) R1 - R2/q ],  q = a b + R1.R2
//   dF/dR2 = F [ -a/(b^2 (a+b)) R2 - a/(b q) R2 - R1/q ]
// x_i receives -(dL/dR1 + dL/dR2), x_j receives dL/dR1, the point infront(j)
// (whose position is g_j) receives dL/dR2.  Target and source sums run as two
// passes with one thread per point each, so no atomics are needed and the
// result is deterministic.  The local term's Jacobian comes from forward-mode
// duals over the same code as the velocity stage (qv_lia.h).
#include "qv_internal.h"
#include "qv_lia.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 256;

struct AdjArgs {
  int n, order, mixed;
  double qc, corea, L;
  bool px, py, pz;
  const double *X, *Y, *Z;
  const double *GX, *GY, *GZ;  // ghosti (unwrapped x of infront)
  const double *BX, *BY, *BZ;  // ghostb
  const double *IX, *IY, *IZ;  // ghostii
  const double *JX, *JY, *JZ;  // ghostbb
  const int32_t *inf, *beh;
  const double *wx, *wy, *wz;  // cotangent (VJP) or tangent (JVP), slot order
  double *out;                 // scratch, 24 doubles per slot (see k_adj_combine)
  double *rx, *ry, *rz;        // result
};

// scratch layout per slot, 24 doubles: [0..2] target part, [3..5] source
// part at x_j, [6..8] source part at g_j, [9..23] local term w.r.t. x, gi,
// gb, gii, gbb
__device__ inline double *slot_out(const AdjArgs &A, int i) { return A.out + 24 * (int64_t)i; }

// dL/dR1, dL/dR2 of one segment term (without Gamma/(4 pi)); false when the
// term is skipped (R1 x R2 == 0, the D == 0 case of timestep.f90:476)
__device__ inline bool pair_grad64(const double R1[3], const double R2[3], const double w[3],
                                   double g1[3], double g2[3]) {
  const double c[3] = {R1[1] * R2[2] - R1[2] * R2[1], R1[2] * R2[0] - R1[0] * R2[2],
                       R1[0] * R2[1] - R1[1] * R2[0]};
  const double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  if (cc == 0.) return false;
  const double a2 = R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2];
  const double b2 = R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2];
  const double a = sqrt(a2), b = sqrt(b2);
  const double p = R1[0] * R2[0] + R1[1] * R2[1] + R1[2] * R2[2];
  const double ab = a * b;
  const double q = p >= 0. ? ab + p : cc / (ab - p);  // a b + R1.R2 without cancellation
  const double iq = 1. / q, iab = 1. / (a + b);
  const double F = (a + b) / (ab * q);
  const double wc = w[0] * c[0] + w[1] * c[1] + w[2] * c[2];
  const double k1 = -b * iab / a2 - (b / a) * iq;
  const double m2 = -a * iab / b2 - (a / b) * iq;
  const double s = wc * F;
  // R2 x w and w x R1
  const double r2w[3] = {R2[1] * w[2] - R2[2] * w[1], R2[2] * w[0] - R2[0] * w[2],
                         R2[0] * w[1] - R2[1] * w[0]};
  const double wr1[3] = {w[1] * R1[2] - w[2] * R1[1], w[2] * R1[0] - w[0] * R1[2],
                         w[0] * R1[1] - w[1] * R1[0]};
  for (int k = 0; k < 3; ++k) {
    g1[k] = s * (k1 * R1[k] - iq * R2[k]) + F * r2w[k];
    g2[k] = s * (m2 * R2[k] - iq * R1[k]) + F * wr1[k];
  }
  return true;
}

// the same in FP32 from FP64-formed R1, R2 (the mixed-precision adjoint)
__device__ inline bool pair_grad32(const double R1d[3], const double R2d[3], const double wd[3],
                                   double g1[3], double g2[3]) {
  const float R1[3] = {(float)R1d[0], (float)R1d[1], (float)R1d[2]};
  const float e[3] = {(float)(R2d[0] - R1d[0]), (float)(R2d[1] - R1d[1]),
                      (float)(R2d[2] - R1d[2])};
  const float R2[3] = {R1[0] + e[0], R1[1] + e[1], R1[2] + e[2]};
  const float w[3] = {(float)wd[0], (float)wd[1], (float)wd[2]};
  // c = R1 x e (= R1 x R2 without cancellation)
  const float c[3] = {R1[1] * e[2] - R1[2] * e[1], R1[2] * e[0] - R1[0] * e[2],
                      R1[0] * e[1] - R1[1] * e[0]};
  const float cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  if (cc == 0.f) return false;
  const float a2 = R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2];
  const float b2 = R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2];
  const float ia = __builtin_amdgcn_rsqf(a2), ib = __builtin_amdgcn_rsqf(b2);
  const float a = a2 * ia, b = b2 * ib;
  const float p = R1[0] * R2[0] + R1[1] * R2[1] + R1[2] * R2[2];
  const float ab = a * b;
  const float iq = p >= 0.f ? __builtin_amdgcn_rcpf(ab + p) : (ab - p) * __builtin_amdgcn_rcpf(cc);
  const float iab = __builtin_amdgcn_rcpf(a + b);
  const float F = (a + b) * (ia * ib) * iq;
  const float wc = w[0] * c[0] + w[1] * c[1] + w[2] * c[2];
  const float k1 = -b * iab * ia * ia - b * ia * iq;
  const float m2 = -a * iab * ib * ib - a * ib * iq;
  const float s = wc * F;
  const float r2w[3] = {R2[1] * w[2] - R2[2] * w[1], R2[2] * w[0] - R2[0] * w[2],
                        R2[0] * w[1] - R2[1] * w[0]};
  const float wr1[3] = {w[1] * R1[2] - w[2] * R1[1], w[2] * R1[0] - w[0] * R1[2],
                        w[0] * R1[1] - w[1] * R1[0]};
  for (int k = 0; k < 3; ++k) {
    g1[k] = (double)(s * (k1 * R1[k] - iq * R2[k]) + F * r2w[k]);
    g2[k] = (double)(s * (m2 * R2[k] - iq * R1[k]) + F * wr1[k]);
  }
  return true;
}

struct Tile {
  double x[kTile], y[kTile], z[kTile];
  double a[kTile], b[kTile], c[kTile];  // ghosti (sources) or w (targets)
  int32_t k0[kTile], k1[kTile];          // active flag; behind (targets)
};

// pass 1: one thread per target i, sum over sources and shells of
// -(dL/dR1 + dL/dR2); the sources of chunk blockIdx.y only (the source
// range is cut into chunks when the targets alone do not fill the GPU):
// the sums go to part (3 per target per chunk), or straight to the slot's
// scratch when there is one chunk
__global__ void __launch_bounds__(kBlock) k_adj_target(AdjArgs A, Shells S, int len,
                                                       double *__restrict__ part) {
  __shared__ Tile t;
  const int lo = blockIdx.y * len, hi = min(A.n, lo + len);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < A.n && A.inf[i] != 0;
  double xi[3] = {0., 0., 0.}, w[3] = {0., 0., 0.};
  int bi = -1;
  if (act) {
    xi[0] = A.X[i]; xi[1] = A.Y[i]; xi[2] = A.Z[i];
    w[0] = A.wx[i]; w[1] = A.wy[i]; w[2] = A.wz[i];
    bi = A.beh[i] - 1;
  }
  double acc[3] = {0., 0., 0.};
  for (int base = lo; base < hi; base += kTile) {
    __syncthreads();
    {
      const int j = base + threadIdx.x;
      const bool in = j < hi;
      t.x[threadIdx.x] = in ? A.X[j] : 0.;
      t.y[threadIdx.x] = in ? A.Y[j] : 0.;
      t.z[threadIdx.x] = in ? A.Z[j] : 0.;
      t.a[threadIdx.x] = in ? A.GX[j] : 0.;
      t.b[threadIdx.x] = in ? A.GY[j] : 0.;
      t.c[threadIdx.x] = in ? A.GZ[j] : 0.;
      t.k0[threadIdx.x] = in ? (A.inf[j] != 0) : 0;
    }
    __syncthreads();
    if (!act) continue;
    const int m = min(kTile, hi - base);
    for (int sh = 0; sh < S.n; ++sh) {
      const double s[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
      for (int jj = 0; jj < m; ++jj) {
        const int j = base + jj;
        if (!t.k0[jj] || j == bi || (sh == 0 && j == i)) continue;
        const double R1[3] = {(t.x[jj] + s[0]) - xi[0], (t.y[jj] + s[1]) - xi[1],
                              (t.z[jj] + s[2]) - xi[2]};
        const double R2[3] = {(t.a[jj] + s[0]) - xi[0], (t.b[jj] + s[1]) - xi[1],
                              (t.c[jj] + s[2]) - xi[2]};
        double g1[3], g2[3];
        const bool ok = A.mixed ? pair_grad32(R1, R2, w, g1, g2) : pair_grad64(R1, R2, w, g1, g2);
        if (ok)
          for (int k = 0; k < 3; ++k) acc[k] -= g1[k] + g2[k];
      }
    }
  }
  if (i >= A.n) return;
  double *o = part ? part + 3 * ((int64_t)blockIdx.y * A.n + i) : slot_out(A, i);
  for (int k = 0; k < 3; ++k) o[k] = acc[k];
}

// pass 2: one thread per source j, sum over targets and shells of dL/dR1
// (at x_j) and dL/dR2 (at g_j); the targets of chunk blockIdx.y only: the
// sums go to part (6 per source per chunk) or straight to the scratch
__global__ void __launch_bounds__(kBlock) k_adj_source(AdjArgs A, Shells S, int len,
                                                       double *__restrict__ part) {
  __shared__ Tile t;
  const int lo = blockIdx.y * len, hi = min(A.n, lo + len);
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = j < A.n && A.inf[j] != 0;
  double xj[3] = {0., 0., 0.}, gj[3] = {0., 0., 0.};
  if (act) {
    xj[0] = A.X[j]; xj[1] = A.Y[j]; xj[2] = A.Z[j];
    gj[0] = A.GX[j]; gj[1] = A.GY[j]; gj[2] = A.GZ[j];
  }
  double a1[3] = {0., 0., 0.}, a2[3] = {0., 0., 0.};
  for (int base = lo; base < hi; base += kTile) {
    __syncthreads();
    {
      const int i = base + threadIdx.x;
      const bool in = i < hi;
      t.x[threadIdx.x] = in ? A.X[i] : 0.;
      t.y[threadIdx.x] = in ? A.Y[i] : 0.;
      t.z[threadIdx.x] = in ? A.Z[i] : 0.;
      t.a[threadIdx.x] = in ? A.wx[i] : 0.;
      t.b[threadIdx.x] = in ? A.wy[i] : 0.;
      t.c[threadIdx.x] = in ? A.wz[i] : 0.;
      t.k0[threadIdx.x] = in ? (A.inf[i] != 0) : 0;
      t.k1[threadIdx.x] = in ? A.beh[i] - 1 : -1;
    }
    __syncthreads();
    if (!act) continue;
    const int m = min(kTile, hi - base);
    for (int sh = 0; sh < S.n; ++sh) {
      const double s[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
      for (int ii = 0; ii < m; ++ii) {
        const int i = base + ii;
        if (!t.k0[ii] || j == t.k1[ii] || (sh == 0 && j == i)) continue;
        const double xi[3] = {t.x[ii], t.y[ii], t.z[ii]};
        const double w[3] = {t.a[ii], t.b[ii], t.c[ii]};
        const double R1[3] = {(xj[0] + s[0]) - xi[0], (xj[1] + s[1]) - xi[1],
                              (xj[2] + s[2]) - xi[2]};
        const double R2[3] = {(gj[0] + s[0]) - xi[0], (gj[1] + s[1]) - xi[1],
                              (gj[2] + s[2]) - xi[2]};
        double g1[3], g2[3];
        const bool ok = A.mixed ? pair_grad32(R1, R2, w, g1, g2) : pair_grad64(R1, R2, w, g1, g2);
        if (ok)
          for (int k = 0; k < 3; ++k) {
            a1[k] += g1[k];
            a2[k] += g2[k];
          }
      }
    }
  }
  if (j >= A.n) return;
  if (part) {
    double *o = part + 6 * ((int64_t)blockIdx.y * A.n + j);
    for (int k = 0; k < 3; ++k) {
      o[k] = a1[k];
      o[3 + k] = a2[k];
    }
  } else {
    double *o = slot_out(A, j);
    for (int k = 0; k < 3; ++k) {
      o[3 + k] = a1[k];
      o[6 + k] = a2[k];
    }
  }
}

// the chunks' sums of both passes in chunk order into the slot's scratch
__global__ void __launch_bounds__(kBlock) k_adj_chunks(AdjArgs A, int nchunk,
                                                       const double *__restrict__ pt,
                                                       const double *__restrict__ ps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double s[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  for (int c = 0; c < nchunk; ++c) {
    const double *a = pt + 3 * ((int64_t)c * A.n + i);
    const double *b = ps + 6 * ((int64_t)c * A.n + i);
    for (int k = 0; k < 3; ++k) s[k] = s[k] + a[k];
    for (int k = 0; k < 6; ++k) s[3 + k] = s[3 + k] + b[k];
  }
  double *o = slot_out(A, i);
  for (int k = 0; k < 9; ++k) o[k] = s[k];
}

// local term: w . d u_LI / d(coordinate) for each of the inputs x, gi, gb
// (and gii, gbb at 4th order), one forward dual pass per coordinate
__global__ void __launch_bounds__(kBlock) k_adj_lia(AdjArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double *o = slot_out(A, i);
  for (int k = 9; k < 24; ++k) o[k] = 0.;
  if (A.inf[i] == 0) return;
  const double in[5][3] = {{A.X[i], A.Y[i], A.Z[i]},
                           {A.GX[i], A.GY[i], A.GZ[i]},
                           {A.BX[i], A.BY[i], A.BZ[i]},
                           {A.IX[i], A.IY[i], A.IZ[i]},
                           {A.JX[i], A.JY[i], A.JZ[i]}};
  const double w[3] = {A.wx[i], A.wy[i], A.wz[i]};
  const int nin = A.order == 4 ? 5 : 3;
  for (int v = 0; v < nin; ++v) {
    for (int k = 0; k < 3; ++k) {
      Dual d[5][3];
      for (int a = 0; a < 5; ++a)
        for (int c = 0; c < 3; ++c) d[a][c] = Dual(in[a][c], (a == v && c == k) ? 1. : 0.);
      Dual u[3];
      lia_term<Dual>(A.qc, A.corea, A.order, d[0], d[1], d[2], d[3], d[4], u);
      o[9 + 3 * v + k] = (w[0] * u[0].d + w[1] * u[1].d) + w[2] * u[2].d;
    }
  }
}

// xbar_k = target part + source parts whose endpoint is x_k (x_j at j = k,
// g_j at j = behind(k)) + local-term parts (x at k, gi at behind(k), gb at
// infront(k), gii at behind(behind(k)), gbb at infront(infront(k)));
// times Gamma/(4 pi) for the pair parts
__global__ void __launch_bounds__(kBlock) k_adj_combine(AdjArgs A) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n) return;
  if (A.inf[k] == 0) {
    A.rx[k] = A.ry[k] = A.rz[k] = 0.;
    return;
  }
  const int f = A.inf[k] - 1, b = A.beh[k] - 1;
  const int ff = A.inf[f] - 1, bb = A.beh[b] - 1;
  const double *ok_ = slot_out(A, k), *ob = slot_out(A, b), *of = slot_out(A, f);
  const double *obb = slot_out(A, bb), *off = slot_out(A, ff);
  const double g4 = A.qc / (4. * kPi);
  double r[3];
  for (int c = 0; c < 3; ++c) {
    double pair = ok_[c] + ok_[3 + c] + ob[6 + c];
    double loc = ok_[9 + c] + ob[12 + c] + of[15 + c];
    if (A.order == 4) loc += obb[18 + c] + off[21 + c];
    r[c] = g4 * pair + loc;
  }
  A.rx[k] = r[0]; A.ry[k] = r[1]; A.rz[k] = r[2];
}

// ---------------------------------------------------------------------------
// JVP: du_i = dU_i/dx . v, by forward duals: the local term with all inputs
// seeded by their tangents, the pair terms with R1' = v_j - v_i,
// R2' = v_infront(j) - v_i

template <class T>
__device__ inline void seg_term(const T R1[3], const T R2[3], T u[3]) {
  const T a = sqrt((R1[0] * R1[0] + R1[1] * R1[1]) + R1[2] * R1[2]);
  const T b = sqrt((R2[0] * R2[0] + R2[1] * R2[1]) + R2[2] * R2[2]);
  const T c = (a * b) * ((a * b) + ((R1[0] * R2[0] + R1[1] * R2[1]) + R1[2] * R2[2]));
  const T f = (a + b) / c;
  u[0] = f * (R1[1] * R2[2] - R1[2] * R2[1]);
  u[1] = f * (R1[2] * R2[0] - R1[0] * R2[2]);
  u[2] = f * (R1[0] * R2[1] - R1[1] * R2[0]);
}

__global__ void __launch_bounds__(kBlock) k_jvp(AdjArgs A, Shells S) {
  __shared__ Tile t;
  __shared__ double tv[3][kTile], tg[3][kTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < A.n && A.inf[i] != 0;
  double xi[3] = {0., 0., 0.}, vi[3] = {0., 0., 0.};
  int bi = -1;
  if (act) {
    xi[0] = A.X[i]; xi[1] = A.Y[i]; xi[2] = A.Z[i];
    vi[0] = A.wx[i]; vi[1] = A.wy[i]; vi[2] = A.wz[i];
    bi = A.beh[i] - 1;
  }
  double acc[3] = {0., 0., 0.};
  for (int sh = 0; sh < S.n; ++sh) {
    const double s[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
    for (int base = 0; base < A.n; base += kTile) {
      __syncthreads();
      {
        const int j = base + threadIdx.x;
        const bool in = j < A.n && A.inf[j] != 0;
        t.x[threadIdx.x] = in ? A.X[j] : 0.;
        t.y[threadIdx.x] = in ? A.Y[j] : 0.;
        t.z[threadIdx.x] = in ? A.Z[j] : 0.;
        t.a[threadIdx.x] = in ? A.GX[j] : 0.;
        t.b[threadIdx.x] = in ? A.GY[j] : 0.;
        t.c[threadIdx.x] = in ? A.GZ[j] : 0.;
        t.k0[threadIdx.x] = in;
        const int fj = in ? A.inf[j] - 1 : 0;
        tv[0][threadIdx.x] = in ? A.wx[j] : 0.;
        tv[1][threadIdx.x] = in ? A.wy[j] : 0.;
        tv[2][threadIdx.x] = in ? A.wz[j] : 0.;
        tg[0][threadIdx.x] = in ? A.wx[fj] : 0.;
        tg[1][threadIdx.x] = in ? A.wy[fj] : 0.;
        tg[2][threadIdx.x] = in ? A.wz[fj] : 0.;
      }
      __syncthreads();
      if (!act) continue;
      const int m = min(kTile, A.n - base);
      for (int jj = 0; jj < m; ++jj) {
        const int j = base + jj;
        if (!t.k0[jj] || j == bi || (sh == 0 && j == i)) continue;
        const double r1[3] = {(t.x[jj] + s[0]) - xi[0], (t.y[jj] + s[1]) - xi[1],
                              (t.z[jj] + s[2]) - xi[2]};
        const double r2[3] = {(t.a[jj] + s[0]) - xi[0], (t.b[jj] + s[1]) - xi[1],
                              (t.c[jj] + s[2]) - xi[2]};
        const double cx = r1[1] * r2[2] - r1[2] * r2[1], cy = r1[2] * r2[0] - r1[0] * r2[2],
                     cz = r1[0] * r2[1] - r1[1] * r2[0];
        if (cx * cx + cy * cy + cz * cz == 0.) continue;
        Dual R1[3], R2[3], u[3];
        for (int k = 0; k < 3; ++k) {
          R1[k] = Dual(r1[k], tv[k][jj] - vi[k]);
          R2[k] = Dual(r2[k], tg[k][jj] - vi[k]);
        }
        seg_term<Dual>(R1, R2, u);
        for (int k = 0; k < 3; ++k) acc[k] += u[k].d;
      }
    }
  }
  if (i >= A.n) return;
  if (!act) {
    A.rx[i] = A.ry[i] = A.rz[i] = 0.;
    return;
  }
  // local term, all inputs seeded with their tangents
  const int f = A.inf[i] - 1, b = bi, ff = A.inf[f] - 1, bb = A.beh[b] - 1;
  const double in[5][3] = {{A.X[i], A.Y[i], A.Z[i]},
                           {A.GX[i], A.GY[i], A.GZ[i]},
                           {A.BX[i], A.BY[i], A.BZ[i]},
                           {A.IX[i], A.IY[i], A.IZ[i]},
                           {A.JX[i], A.JY[i], A.JZ[i]}};
  const int src[5] = {i, f, b, ff, bb};
  Dual d[5][3];
  for (int a = 0; a < 5; ++a) {
    const int q = src[a];
    const double tq[3] = {A.wx[q], A.wy[q], A.wz[q]};
    for (int c = 0; c < 3; ++c) d[a][c] = Dual(in[a][c], tq[c]);
  }
  Dual ul[3];
  lia_term<Dual>(A.qc, A.corea, A.order, d[0], d[1], d[2], d[3], d[4], ul);
  const double g4 = A.qc / (4. * kPi);
  A.rx[i] = g4 * acc[0] + ul[0].d;
  A.ry[i] = g4 * acc[1] + ul[1].d;
  A.rz[i] = g4 * acc[2] + ul[2].d;
}

inline int grid(int n) { return (n + kBlock - 1) / kBlock; }

AdjArgs make_args(Ctx &c, Vec3 w, Vec3 r, int mixed, double *scratch) {
  AdjArgs A;
  A.n = c.pcount;
  A.order = c.p.deriv_order;
  A.mixed = mixed;
  A.qc = c.p.quant_circ;
  A.corea = c.p.corea;
  A.L = c.p.box_size;
  A.px = c.p.boundary == QV_BC_PERIODIC;
  A.py = A.px || c.p.boundary == QV_BC_OPENX;
  A.pz = c.p.boundary != QV_BC_OPEN;
  A.X = c.x.x; A.Y = c.x.y; A.Z = c.x.z;
  A.GX = c.gi.x; A.GY = c.gi.y; A.GZ = c.gi.z;
  A.BX = c.gb.x; A.BY = c.gb.y; A.BZ = c.gb.z;
  A.IX = c.gii.x; A.IY = c.gii.y; A.IZ = c.gii.z;
  A.JX = c.gbb.x; A.JY = c.gbb.y; A.JZ = c.gbb.z;
  A.inf = c.infront; A.beh = c.behind;
  A.wx = w.x; A.wy = w.y; A.wz = w.z;
  A.out = scratch;
  A.rx = r.x; A.ry = r.y; A.rz = r.z;
  return A;
}

}  // namespace

int vjp_chunks(int n, int &len) {
  // about 4096 wavefronts, 16 per CU, in whole tiles
  const int waves = grid(n) * (kBlock / 64);
  const int want = (4096 + waves - 1) / waves;
  len = (n + want - 1) / want;
  len = ((len + kTile - 1) / kTile) * kTile;
  return (n + len - 1) / len;
}

hipError_t launch_vjp(Ctx &c, Vec3 w, Vec3 r, int mixed, double *scratch, double *part,
                      int64_t part_doubles) {
  AdjArgs A = make_args(c, w, r, mixed, scratch);
  Shells S = make_shells(c.p);
  const int g = grid(c.pcount);
  int len = 0;
  int nchunk = vjp_chunks(c.pcount, len);
  if (!part || 9 * (int64_t)nchunk * c.pcount > part_doubles) {
    nchunk = 1;  // no room for partial sums: one chunk, straight to the scratch
    len = c.pcount;
    part = nullptr;
  }
  double *pt = part, *ps = part ? part + 3 * (int64_t)nchunk * c.pcount : nullptr;
  hipLaunchKernelGGL(k_adj_target, dim3(g, nchunk), dim3(kBlock), 0, c.stream, A, S, len, pt);
  hipLaunchKernelGGL(k_adj_source, dim3(g, nchunk), dim3(kBlock), 0, c.stream, A, S, len, ps);
  if (part)
    hipLaunchKernelGGL(k_adj_chunks, dim3(g), dim3(kBlock), 0, c.stream, A, nchunk, pt, ps);
  hipLaunchKernelGGL(k_adj_lia, dim3(g), dim3(kBlock), 0, c.stream, A);
  hipLaunchKernelGGL(k_adj_combine, dim3(g), dim3(kBlock), 0, c.stream, A);
  return hipGetLastError();
}

hipError_t launch_jvp(Ctx &c, Vec3 v, Vec3 r) {
  AdjArgs A = make_args(c, v, r, 0, nullptr);
  Shells S = make_shells(c.p);
  hipLaunchKernelGGL(k_jvp, dim3(grid(c.pcount)), dim3(kBlock), 0, c.stream, A, S);
  return hipGetLastError();
}

}  // namespace qv
