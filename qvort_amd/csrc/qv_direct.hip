// qv_direct.hip -- tiled direct Biot-Savart sum (velocity='BS').
//
// One thread per target slot; sources are staged through LDS in tiles of
// kTile slots.  The FP64 kernel visits sources in slot order, shell by
// shell, exactly as biot_savart / biot_savart_shift do (timestep.f90:462-507,
// called from :227-259), and keeps their operation order, so its sums are
// bit-identical to the reference.  The FP32C kernel evaluates each pair with
// the equivalent segment form of tree_walk's leaf (tree.f90:380-392) in
// FP32 with fused multiply-adds, relative vectors formed in FP64, and sums
// per target in FP64.
#include "qv_internal.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 256;

struct DirectArgs {
  int n;
  double qc;
  const double *X, *Y, *Z;      // positions
  const double *GX, *GY, *GZ;   // ghosti
  const int32_t *inf, *beh;
  const double *ulx, *uly, *ulz;
  double *ux, *uy, *uz;          // velocity (u)
  double *bx, *by, *bz;          // nonlocal part (u_s_BS)
};

// one source of biot_savart (shift == nullptr) or biot_savart_shift
template <bool kShift>
__device__ inline void bs_pair(const double xi[3], const double xj[3], const double gj[3],
                               const double s[3], double qc, double u[3]) {
  double r[3], e[3], a;
  if (kShift) {
    const double xs[3] = {xj[0] + s[0], xj[1] + s[1], xj[2] + s[2]};
    a = ((xi[0] - xs[0]) * (xi[0] - xs[0]) + (xi[1] - xs[1]) * (xi[1] - xs[1])) +
        (xi[2] - xs[2]) * (xi[2] - xs[2]);
    for (int k = 0; k < 3; ++k) r[k] = xs[k] - xi[k];
  } else {
    a = ((xj[0] - xi[0]) * (xj[0] - xi[0]) + (xj[1] - xi[1]) * (xj[1] - xi[1])) +
        (xj[2] - xi[2]) * (xj[2] - xi[2]);
    for (int k = 0; k < 3; ++k) r[k] = xj[k] - xi[k];
  }
  for (int k = 0; k < 3; ++k) e[k] = gj[k] - xj[k];
  const double b = 2. * ((r[0] * e[0] + r[1] * e[1]) + r[2] * e[2]);
  const double c = ((gj[0] - xj[0]) * (gj[0] - xj[0]) + (gj[1] - xj[1]) * (gj[1] - xj[1])) +
                   (gj[2] - xj[2]) * (gj[2] - xj[2]);
  const double D = (4. * a) * c - b * b;
  if (D == 0.) return;
  const double den = (2. * kPi) * D;
  const double ub0 = ((r[1] * e[2] - r[2] * e[1]) * qc) / den;
  const double ub1 = ((r[2] * e[0] - r[0] * e[2]) * qc) / den;
  const double ub2 = ((r[0] * e[1] - r[1] * e[0]) * qc) / den;
  const double fac = (2. * c + b) / sqrt((a + b) + c) - b / sqrt(a);
  u[0] = u[0] + ub0 * fac;
  u[1] = u[1] + ub1 * fac;
  u[2] = u[2] + ub2 * fac;
}

struct SrcTile {
  double x[kTile], y[kTile], z[kTile];
  double gx[kTile], gy[kTile], gz[kTile];
  int32_t ok[kTile];
};

// the FP32C kernel's tile: positions in FP64 (the relative vector is formed
// in FP64), the segment g - x already rounded to FP32 (it does not depend
// on the target)
struct SrcTile32 {
  double x[kTile], y[kTile], z[kTile];
  float ex[kTile], ey[kTile], ez[kTile];
  int32_t ok[kTile];
};

__device__ inline void load_tile(SrcTile &t, const DirectArgs &A, int base) {
  const int j = base + threadIdx.x;
  if (threadIdx.x < kTile) {
    const bool in = j < A.n;
    t.x[threadIdx.x] = in ? A.X[j] : 0.;
    t.y[threadIdx.x] = in ? A.Y[j] : 0.;
    t.z[threadIdx.x] = in ? A.Z[j] : 0.;
    t.gx[threadIdx.x] = in ? A.GX[j] : 0.;
    t.gy[threadIdx.x] = in ? A.GY[j] : 0.;
    t.gz[threadIdx.x] = in ? A.GZ[j] : 0.;
    t.ok[threadIdx.x] = in ? (A.inf[j] != 0) : 0;
  }
}

__global__ void __launch_bounds__(kBlock) k_direct_fp64(DirectArgs A, Shells S) {
  __shared__ SrcTile t;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < A.n && A.inf[i] != 0;
  double xi[3] = {0., 0., 0.};
  int bi = -1;
  if (act) {
    xi[0] = A.X[i]; xi[1] = A.Y[i]; xi[2] = A.Z[i];
    bi = A.beh[i] - 1;
  }
  double uc[3] = {0., 0., 0.}, us[3] = {0., 0., 0.};
  for (int sh = 0; sh < S.n; ++sh) {
    const double s[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
    for (int base = 0; base < A.n; base += kTile) {
      __syncthreads();
      load_tile(t, A, base);
      __syncthreads();
      if (!act) continue;
      const int m = min(kTile, A.n - base);
      for (int jj = 0; jj < m; ++jj) {
        const int j = base + jj;
        if (!t.ok[jj] || j == bi) continue;
        const double xj[3] = {t.x[jj], t.y[jj], t.z[jj]};
        const double gj[3] = {t.gx[jj], t.gy[jj], t.gz[jj]};
        if (sh == 0) {
          if (j == i) continue;
          bs_pair<false>(xi, xj, gj, s, A.qc, uc);
        } else {
          bs_pair<true>(xi, xj, gj, s, A.qc, us);
        }
      }
    }
  }
  if (i >= A.n) return;
  if (!act) {
    A.ux[i] = A.uy[i] = A.uz[i] = 0.;
    A.bx[i] = A.by[i] = A.bz[i] = 0.;
    return;
  }
  // u = u_LI + u_bs(central); u = u + u_bs(images)  (timestep.f90:225-240)
  double u0 = A.ulx[i] + uc[0], u1 = A.uly[i] + uc[1], u2 = A.ulz[i] + uc[2];
  double b0 = uc[0], b1 = uc[1], b2 = uc[2];
  if (S.n > 1) {
    u0 = u0 + us[0]; u1 = u1 + us[1]; u2 = u2 + us[2];
    b0 = b0 + us[0]; b1 = b1 + us[1]; b2 = b2 + us[2];
  }
  A.ux[i] = u0; A.uy[i] = u1; A.uz[i] = u2;
  A.bx[i] = b0; A.by[i] = b1; A.bz[i] = b2;
}

__device__ inline void load_tile32(SrcTile32 &t, const DirectArgs &A, int base) {
  const int j = base + threadIdx.x;
  if (threadIdx.x < kTile) {
    const bool in = j < A.n;
    const double x = in ? A.X[j] : 0., y = in ? A.Y[j] : 0., z = in ? A.Z[j] : 0.;
    t.x[threadIdx.x] = x;
    t.y[threadIdx.x] = y;
    t.z[threadIdx.x] = z;
    t.ex[threadIdx.x] = in ? (float)(A.GX[j] - x) : 0.f;
    t.ey[threadIdx.x] = in ? (float)(A.GY[j] - y) : 0.f;
    t.ez[threadIdx.x] = in ? (float)(A.GZ[j] - z) : 0.f;
    t.ok[threadIdx.x] = in ? (A.inf[j] != 0) : 0;
  }
}

__device__ inline float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// seg32 with fused multiply-adds (the FP32 path has no reference rounding to
// keep), as a weight and a direction: the term is W (c0, c1, c2) with W = 0
// where `on` is false or |R1 x e| == 0 (the case the FP64 path skips)
__device__ inline void seg32w(bool on, float r0, float r1, float r2, float e0, float e1,
                              float e2, float &W, float &c0, float &c1, float &c2) {
  c0 = fmaf_(r1, e2, -r2 * e1);
  c1 = fmaf_(r2, e0, -r0 * e2);
  c2 = fmaf_(r0, e1, -r1 * e0);
  const float cc = fmaf_(c2, c2, fmaf_(c1, c1, c0 * c0));
  const float s0 = r0 + e0, s1 = r1 + e1, s2 = r2 + e2;
  const float a2 = fmaf_(r2, r2, fmaf_(r1, r1, r0 * r0));
  const float b2 = fmaf_(s2, s2, fmaf_(s1, s1, s0 * s0));
  const float ia = __builtin_amdgcn_rsqf(a2), ib = __builtin_amdgcn_rsqf(b2);
  const float a = a2 * ia, b = b2 * ib;
  const float ab = a * b;
  const float dt = fmaf_(r2, s2, fmaf_(r1, s1, r0 * s0));
  const float g = (a + b) * (ia * ib);
  // one reciprocal, of the selected denominator
  const bool in = dt >= 0.f;
  const float f = (in ? g : g * (ab - dt)) * __builtin_amdgcn_rcpf(in ? ab + dt : cc);
  W = (on && cc != 0.f) ? f : 0.f;
}

__global__ void __launch_bounds__(kBlock) k_direct_fp32c(DirectArgs A, Shells S) {
  __shared__ SrcTile32 t;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < A.n && A.inf[i] != 0;
  double xi[3] = {0., 0., 0.};
  int bi = -1;
  if (act) {
    xi[0] = A.X[i]; xi[1] = A.Y[i]; xi[2] = A.Z[i];
    bi = A.beh[i] - 1;
  }
  // per-target sums in FP64 (on CDNA4 an FP64 fused multiply-add issues at
  // the FP32 rate: the cheapest compensated sum of FP32 terms)
  double uc[3] = {0., 0., 0.}, us[3] = {0., 0., 0.};
  for (int sh = 0; sh < S.n; ++sh) {
    const double xs0 = xi[0] - S.s[sh][0], xs1 = xi[1] - S.s[sh][1], xs2 = xi[2] - S.s[sh][2];
    double v0 = 0., v1 = 0., v2 = 0.;
    for (int base = 0; base < A.n; base += kTile) {
      __syncthreads();
      load_tile32(t, A, base);
      __syncthreads();
      if (!act) continue;
      const int m = min(kTile, A.n - base);
      for (int jj = 0; jj < m; ++jj) {
        const int j = base + jj;
        const bool on = t.ok[jj] && j != bi && !(sh == 0 && j == i);
        const float r0 = (float)(t.x[jj] - xs0), r1 = (float)(t.y[jj] - xs1),
                    r2 = (float)(t.z[jj] - xs2);
        float W, c0, c1, c2;
        seg32w(on, r0, r1, r2, t.ex[jj], t.ey[jj], t.ez[jj], W, c0, c1, c2);
        const double w = (double)W;
        v0 = fma(w, (double)c0, v0);
        v1 = fma(w, (double)c1, v1);
        v2 = fma(w, (double)c2, v2);
      }
    }
    if (sh == 0) {
      uc[0] = v0; uc[1] = v1; uc[2] = v2;
    } else {
      us[0] = us[0] + v0; us[1] = us[1] + v1; us[2] = us[2] + v2;
    }
  }
  if (i >= A.n) return;
  if (!act) {
    A.ux[i] = A.uy[i] = A.uz[i] = 0.;
    A.bx[i] = A.by[i] = A.bz[i] = 0.;
    return;
  }
  const double g4 = A.qc / (4. * kPi);
  for (int k = 0; k < 3; ++k) { uc[k] *= g4; us[k] *= g4; }
  double u0 = A.ulx[i] + uc[0], u1 = A.uly[i] + uc[1], u2 = A.ulz[i] + uc[2];
  double b0 = uc[0], b1 = uc[1], b2 = uc[2];
  if (S.n > 1) {
    u0 = u0 + us[0]; u1 = u1 + us[1]; u2 = u2 + us[2];
    b0 = b0 + us[0]; b1 = b1 + us[1]; b2 = b2 + us[2];
  }
  A.ux[i] = u0; A.uy[i] = u1; A.uz[i] = u2;
  A.bx[i] = b0; A.by[i] = b1; A.bz[i] = b2;
}

}  // namespace

hipError_t launch_direct(Ctx &c) {
  DirectArgs A;
  A.n = c.pcount;
  A.qc = c.p.quant_circ;
  A.X = c.x.x; A.Y = c.x.y; A.Z = c.x.z;
  A.GX = c.gi.x; A.GY = c.gi.y; A.GZ = c.gi.z;
  A.inf = c.infront; A.beh = c.behind;
  A.ulx = c.uli.x; A.uly = c.uli.y; A.ulz = c.uli.z;
  A.ux = c.u.x; A.uy = c.u.y; A.uz = c.u.z;
  A.bx = c.ubs.x; A.by = c.ubs.y; A.bz = c.ubs.z;
  Shells S = make_shells(c.p);
  const int g = (c.pcount + kBlock - 1) / kBlock;
  if (c.p.precision == QV_FP32C)
    hipLaunchKernelGGL(k_direct_fp32c, dim3(g), dim3(kBlock), 0, c.stream, A, S);
  else
    hipLaunchKernelGGL(k_direct_fp64, dim3(g), dim3(kBlock), 0, c.stream, A, S);
  return hipGetLastError();
}

}  // namespace qv
