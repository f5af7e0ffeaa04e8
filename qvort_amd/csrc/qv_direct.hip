// qv_direct.hip -- tiled direct Biot-Savart sum (velocity='BS').
//
// Sources are staged through LDS in tiles of kTile slots.  The FP64 kernel
// evaluates biot_savart / biot_savart_shift's pair term (timestep.f90:462-507,
// called from :227-259) in the reference's operation order except for one
// division of the denominator in place of three, each thread summing two
// targets; the source range is cut into chunks over a second grid dimension
// when the targets alone do not fill the GPU (16 wavefronts per CU), and
// k_direct_combine adds the chunks' partial sums in chunk order
// (deterministic; the order of the sum differs from the reference's
// sequential one, within 1e-12 of it per point).  The FP32C kernel evaluates each pair with the equivalent
// segment form of tree_walk's leaf (tree.f90:380-392) in FP32 with fused
// multiply-adds, relative vectors formed in FP64, and sums per target in
// FP64.
#include "qv_internal.h"
#include "qv_terms.h"

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

// the FP32C kernel's tile: positions in FP64 (the relative vector is formed
// in FP64), the segment g - x already rounded to FP32 (it does not depend
// on the target)
struct SrcTile32 {
  double x[kTile], y[kTile], z[kTile];
  float ex[kTile], ey[kTile], ez[kTile];
  int32_t ok[kTile];
};

// the FP64 kernel's tile: source positions and their segments e = g - x
// with |e|^2 (target-independent), FP64
struct SrcTileE {
  double x[kTile], y[kTile], z[kTile];
  double ex[kTile], ey[kTile], ez[kTile], cc[kTile];
  int32_t ok[kTile];
};

constexpr int kTargets = 2;  // targets per thread in the FP64 kernel

// biot_savart's pair term (timestep.f90:470-480): for r = x_j + s - x_i,
// e = g_j - x_j, a = |r|^2, b = 2 r.e, c = |e|^2, D = 4ac - b^2:
// u += (r x e) qc / (2 pi D) * ((2c + b)/sqrt(a + b + c) - b/sqrt(a)).
// a, b, D and the factor in brackets are formed exactly as the reference
// forms them: the bracket cancels for distant segments and D for nearly
// collinear ones, so their rounding must be the reference's for the sum to
// stay within 1e-12 of it.  The reference's three divisions of the cross
// product by (2 pi) D become one division and three products (well
// conditioned, an ulp per term).  Skipped (weight 0) where D == 0
// (timestep.f90:476) or not `on`.
__device__ inline void pair64(bool on, const double r[3], const double e[3], double c, double qc,
                              double u[3]) {
  const double a = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
  const double b = 2. * ((r[0] * e[0] + r[1] * e[1]) + r[2] * e[2]);
  const double D = (4. * a) * c - b * b;
  const bool use = on && D != 0.;
  // lanes that skip use a = c = D = 1 (finite) and weight 0
  const double Ds = use ? D : 1., as = use ? a : 1., abc = use ? (a + b) + c : 1.;
  const double g = use ? qc / ((2. * kPi) * Ds) : 0.;
  const double fac = (2. * c + b) / sqrt(abc) - b / sqrt(as);
  u[0] = u[0] + ((r[1] * e[2] - r[2] * e[1]) * g) * fac;
  u[1] = u[1] + ((r[2] * e[0] - r[0] * e[2]) * g) * fac;
  u[2] = u[2] + ((r[0] * e[1] - r[1] * e[0]) * g) * fac;
}

__device__ inline void load_tile_e(SrcTileE &t, const DirectArgs &A, int base, int end) {
  const int j = base + threadIdx.x;
  if (threadIdx.x < kTile) {
    const bool in = j < end;
    const double x = in ? A.X[j] : 0., y = in ? A.Y[j] : 0., z = in ? A.Z[j] : 0.;
    const double ex = in ? A.GX[j] - x : 0., ey = in ? A.GY[j] - y : 0.,
                 ez = in ? A.GZ[j] - z : 0.;
    t.x[threadIdx.x] = x; t.y[threadIdx.x] = y; t.z[threadIdx.x] = z;
    t.ex[threadIdx.x] = ex; t.ey[threadIdx.x] = ey; t.ez[threadIdx.x] = ez;
    t.cc[threadIdx.x] = (ex * ex + ey * ey) + ez * ez;
    t.ok[threadIdx.x] = in ? (A.inf[j] != 0) : 0;
  }
}

// partial sums of a chunk of sources [chunk * len, ...) for kTargets
// targets per thread: the central shell's (6 per target: central, images)
__global__ void __launch_bounds__(kBlock) k_direct_fp64(DirectArgs A, Shells S, int len,
                                                        double *__restrict__ part) {
  __shared__ SrcTileE t;
  const int lo = blockIdx.y * len, hi = min(A.n, lo + len);
  int ti[kTargets], bi[kTargets];
  bool act[kTargets];
  double xi[kTargets][3];
  for (int q = 0; q < kTargets; ++q) {
    ti[q] = (blockIdx.x * kTargets + q) * kBlock + threadIdx.x;
    act[q] = ti[q] < A.n && A.inf[ti[q]] != 0;
    bi[q] = act[q] ? A.beh[ti[q]] - 1 : -1;
    xi[q][0] = act[q] ? A.X[ti[q]] : 0.;
    xi[q][1] = act[q] ? A.Y[ti[q]] : 0.;
    xi[q][2] = act[q] ? A.Z[ti[q]] : 0.;
  }
  double uc[kTargets][3] = {}, us[kTargets][3] = {};
  for (int base = lo; base < hi; base += kTile) {
    __syncthreads();
    load_tile_e(t, A, base, hi);
    __syncthreads();
    const int m = min(kTile, hi - base);
    for (int sh = 0; sh < S.n; ++sh) {
      for (int jj = 0; jj < m; ++jj) {
        const int j = base + jj;
        const double e[3] = {t.ex[jj], t.ey[jj], t.ez[jj]};
        for (int q = 0; q < kTargets; ++q) {
          // central shell: not the target itself; every shell: not its behind
          const bool on = act[q] && t.ok[jj] && j != bi[q] && (sh > 0 || j != ti[q]);
          const double r[3] = {(t.x[jj] + S.s[sh][0]) - xi[q][0],
                               (t.y[jj] + S.s[sh][1]) - xi[q][1],
                               (t.z[jj] + S.s[sh][2]) - xi[q][2]};
          pair64(on, r, e, t.cc[jj], A.qc, sh == 0 ? uc[q] : us[q]);
        }
      }
    }
  }
  for (int q = 0; q < kTargets; ++q) {
    if (ti[q] >= A.n) continue;
    double *o = part + 6 * ((int64_t)blockIdx.y * A.n + ti[q]);
    o[0] = uc[q][0]; o[1] = uc[q][1]; o[2] = uc[q][2];
    o[3] = us[q][0]; o[4] = us[q][1]; o[5] = us[q][2];
  }
}

// the chunks' sums in chunk order and calc_velocity's composition
// u = u_LI + u_bs(central) + u_bs(images)  (timestep.f90:225-240)
__global__ void __launch_bounds__(kBlock) k_direct_combine(DirectArgs A, int nchunk,
                                                           int nshells,
                                                           const double *__restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  if (A.inf[i] == 0) {
    A.ux[i] = A.uy[i] = A.uz[i] = 0.;
    A.bx[i] = A.by[i] = A.bz[i] = 0.;
    return;
  }
  double uc[3] = {0., 0., 0.}, us[3] = {0., 0., 0.};
  for (int c = 0; c < nchunk; ++c) {
    const double *o = part + 6 * ((int64_t)c * A.n + i);
    for (int k = 0; k < 3; ++k) {
      uc[k] = uc[k] + o[k];
      us[k] = us[k] + o[3 + k];
    }
  }
  double u0 = A.ulx[i] + uc[0], u1 = A.uly[i] + uc[1], u2 = A.ulz[i] + uc[2];
  double b0 = uc[0], b1 = uc[1], b2 = uc[2];
  if (nshells > 1) {
    u0 = u0 + us[0]; u1 = u1 + us[1]; u2 = u2 + us[2];
    b0 = b0 + us[0]; b1 = b1 + us[1]; b2 = b2 + us[2];
  }
  A.ux[i] = u0; A.uy[i] = u1; A.uz[i] = u2;
  A.bx[i] = b0; A.by[i] = b1; A.bz[i] = b2;
}

__device__ inline void load_tile32(SrcTile32 &t, const DirectArgs &A, int base, int end) {
  const int j = base + threadIdx.x;
  if (threadIdx.x < kTile) {
    const bool in = j < end;
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

__global__ void __launch_bounds__(kBlock) k_direct_fp32c(DirectArgs A, Shells S, int len,
                                                         double *__restrict__ part) {
  __shared__ SrcTile32 t;
  const int lo = blockIdx.y * len, hi = min(A.n, lo + len);
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
  for (int base = lo; base < hi; base += kTile) {
    __syncthreads();
    load_tile32(t, A, base, hi);
    __syncthreads();
    if (!act) continue;
    const int m = min(kTile, hi - base);
    for (int sh = 0; sh < S.n; ++sh) {
      const double xs0 = xi[0] - S.s[sh][0], xs1 = xi[1] - S.s[sh][1], xs2 = xi[2] - S.s[sh][2];
      double *v = sh == 0 ? uc : us;
      for (int jj = 0; jj < m; ++jj) {
        const int j = base + jj;
        const bool on = t.ok[jj] && j != bi && !(sh == 0 && j == i);
        const float r0 = (float)(t.x[jj] - xs0), r1 = (float)(t.y[jj] - xs1),
                    r2 = (float)(t.z[jj] - xs2);
        float W, c0, c1, c2;
        seg32w(on, r0, r1, r2, t.ex[jj], t.ey[jj], t.ez[jj], W, c0, c1, c2);
        const double w = (double)W;
        v[0] = fma(w, (double)c0, v[0]);
        v[1] = fma(w, (double)c1, v[1]);
        v[2] = fma(w, (double)c2, v[2]);
      }
    }
  }
  if (i >= A.n) return;
  // Gamma/(4 pi) of the segment form, applied to each chunk's sums
  const double g4 = A.qc / (4. * kPi);
  double *o = part + 6 * ((int64_t)blockIdx.y * A.n + i);
  for (int k = 0; k < 3; ++k) {
    o[k] = uc[k] * g4;
    o[3 + k] = us[k] * g4;
  }
}

// the direct sum: enough chunks of the source range (whole tiles each)
// for about 4096 wavefronts, 16 per CU
int chunk_plan(int pcount, int targets_per_thread, int &len) {
  const int gx = (pcount + kBlock * targets_per_thread - 1) / (kBlock * targets_per_thread);
  const int waves = gx * (kBlock / kWave);
  const int want = (4096 + waves - 1) / waves;
  len = (pcount + want - 1) / want;
  len = ((len + kTile - 1) / kTile) * kTile;
  return (pcount + len - 1) / len;
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
  const bool fp32 = c.p.precision == QV_FP32C;
  int len = 0;
  const int nchunk = chunk_plan(c.pcount, fp32 ? 1 : kTargets, len);
  if (6 * (int64_t)nchunk * c.pcount > c.direct_part_doubles) return hipErrorInvalidValue;
  if (fp32) {
    hipLaunchKernelGGL(k_direct_fp32c, dim3(g, nchunk), dim3(kBlock), 0, c.stream, A, S, len,
                       c.direct_part);
  } else {
    const int gx = (c.pcount + kBlock * kTargets - 1) / (kBlock * kTargets);
    hipLaunchKernelGGL(k_direct_fp64, dim3(gx, nchunk), dim3(kBlock), 0, c.stream, A, S, len,
                       c.direct_part);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_direct_combine, dim3(g), dim3(kBlock), 0, c.stream, A, nchunk, S.n,
                     c.direct_part);
  return hipGetLastError();
}

int direct_chunks(int pcount, bool fp32) {
  int len = 0;
  return chunk_plan(pcount, fp32 ? 1 : kTargets, len);
}

}  // namespace qv
