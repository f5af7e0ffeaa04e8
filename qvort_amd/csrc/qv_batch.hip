// qv_batch.hip -- ensembles of independent filament systems advanced
// together (direct sum, velocity='BS'), one persistent launch per
// Adams-Bashforth step.
//
// No counterpart in the reference: Qvort runs one system per process
// (multiple initial conditions run one after another).  Here a step is one
// launch of k_batch_step over a queue of work units (system, tile of
// targets), units of the largest systems first; every wavefront (a
// workgroup of 64 lanes) takes units from a global counter until the queue
// is empty.  A unit computes, for each
// target of its tile, the LIA term and the direct sum over the sources of
// its own system (timestep.f90:218-259 with biot_savart / biot_savart_shift,
// :462-507): in FP64 in the reference's operation order, or with FP32 pair
// terms (FP64 relative vectors, FP64 sums) and FP64 local terms.  The
// wavefront that completes the last unit of a system applies the system's
// AB ladder (timestep.f90:48-71) and wrap (periodic.f90:188-276): every
// other unit of that system has finished reading its positions by then.
// A step can be captured into a HIP graph (a one-thread kernel resets the
// queue head; no memset nodes), and the tile shape is chosen by timing the
// candidates on the ensemble itself.
#include <algorithm>
#include <new>
#include <vector>

#include "qv_internal.h"
#include "qv_lia.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

struct BatchArgs {
  int nunits;
  double qc, corea, L, dt;
  int order;
  bool px, py, pz;
  const int32_t *unit_sys, *unit_first;  // system, first target slot of each unit
  const int32_t *sys_off, *sys_n, *sys_units;
  double *X, *Y, *Z;
  double *ux, *uy, *uz, *u1x, *u1y, *u1z, *u2x, *u2y, *u2z, *u3x, *u3y, *u3z;
  const int32_t *inf, *beh;  // 1-based slots of the whole ensemble (0: empty slot)
  int32_t *next_unit;        // queue head, reset before each launch
  int32_t *sys_done;         // finished units per system, reset by the finisher
  int32_t do_ab;             // 0: velocity only
};

// get_ghost_p's unwrap (periodic.f90:36-55) for one coordinate
__device__ inline double unwrap1(double xi, double g, double L, bool on) {
  if (!on) return g;
  if ((xi - g) > (L / 2.)) return g + L;
  if ((xi - g) < (-L / 2.)) return g - L;
  return g;
}

__device__ inline void ghost_of(const BatchArgs &A, int j, const double x[3], double g[3]) {
  const bool per[3] = {A.px, A.py, A.pz};
  const double *P[3] = {A.X, A.Y, A.Z};
  for (int k = 0; k < 3; ++k) g[k] = unwrap1(x[k], P[k][j], A.L, per[k]);
}

// one source term of biot_savart (kShift false) / biot_savart_shift, in the
// reference's operation order
template <bool kShift>
__device__ inline void pair64(const double xi[3], const double xj[3], const double gj[3],
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

// FP32 straight-segment term with R1 = x_j + s - x_i and e = g_j - x_j
// (each rounded once from FP64), in the form of tree_walk's leaf
// (tree.f90:380-392): R1 x R2 = R1 x e, and for targets that project inside
// the segment ab + R1.R2 = |R1 x e|^2 / (ab - R1.R2), which keeps the FP32
// arithmetic free of cancellation.  The term (without Gamma/4pi) goes into
// an FP64 sum; |R1 x e| == 0 is skipped as the FP64 path skips D == 0.
__device__ inline void pair32(float r0, float r1, float r2, float e0, float e1, float e2,
                              AccF64 &acc) {
  const float c0 = __builtin_fmaf(r1, e2, -(r2 * e1));
  const float c1 = __builtin_fmaf(r2, e0, -(r0 * e2));
  const float c2 = __builtin_fmaf(r0, e1, -(r1 * e0));
  const float cc = __builtin_fmaf(c2, c2, __builtin_fmaf(c1, c1, c0 * c0));
  if (cc == 0.f) return;
  const float s0 = r0 + e0, s1 = r1 + e1, s2 = r2 + e2;
  const float a2 = __builtin_fmaf(r2, r2, __builtin_fmaf(r1, r1, r0 * r0));
  const float b2 = __builtin_fmaf(s2, s2, __builtin_fmaf(s1, s1, s0 * s0));
  const float ia = __builtin_amdgcn_rsqf(a2), ib = __builtin_amdgcn_rsqf(b2);
  const float a = a2 * ia, b = b2 * ib;
  const float ab = a * b;
  const float dt = __builtin_fmaf(r2, s2, __builtin_fmaf(r1, s1, r0 * s0));
  const float g = (a + b) * (ia * ib);
  // one reciprocal of the selected denominator
  const bool in = dt >= 0.f;
  const float f = (in ? g : g * (ab - dt)) * __builtin_amdgcn_rcpf(in ? ab + dt : cc);
  const double fd = (double)f;
  acc.s0 = fma(fd, (double)c0, acc.s0);
  acc.s1 = fma(fd, (double)c1, acc.s1);
  acc.s2 = fma(fd, (double)c2, acc.s2);
}

__device__ inline double maxabs3(const double a[3]) {
  double m = fabs(a[0]);
  if (fabs(a[1]) > m) m = fabs(a[1]);
  if (fabs(a[2]) > m) m = fabs(a[2]);
  return m;
}

// pmotion's ladder keyed on exact zeros of the history, then the wrap
__device__ inline void ab_point(const BatchArgs &A, int i) {
  if (A.inf[i] == 0) return;
  const double dt = A.dt;
  double x[3] = {A.X[i], A.Y[i], A.Z[i]};
  const double u[3] = {A.ux[i], A.uy[i], A.uz[i]};
  const double a1[3] = {A.u1x[i], A.u1y[i], A.u1z[i]};
  const double a2[3] = {A.u2x[i], A.u2y[i], A.u2z[i]};
  const double a3[3] = {A.u3x[i], A.u3y[i], A.u3z[i]};
  if (maxabs3(a1) == 0.) {
    for (int k = 0; k < 3; ++k) x[k] = x[k] + dt * u[k];
  } else if (maxabs3(a2) == 0.) {
    for (int k = 0; k < 3; ++k) x[k] = (x[k] + ((3. / 2.) * dt) * u[k]) - ((1. / 2.) * dt) * a1[k];
  } else if (maxabs3(a3) == 0.) {
    for (int k = 0; k < 3; ++k)
      x[k] = ((x[k] + ((23. / 12.) * dt) * u[k]) - ((4. / 3.) * dt) * a1[k]) +
             ((5. / 12.) * dt) * a2[k];
  } else {
    for (int k = 0; k < 3; ++k)
      x[k] = x[k] + (dt * (((55. * u[k] - 59. * a1[k]) + 37. * a2[k]) - 9. * a3[k])) / 24.;
  }
  const bool per[3] = {A.px, A.py, A.pz};
  for (int k = 0; k < 3; ++k) {
    if (!per[k]) continue;
    if (x[k] > (A.L / 2.)) x[k] = x[k] - A.L;
    else if (x[k] < (-A.L / 2.)) x[k] = x[k] + A.L;
  }
  A.X[i] = x[0]; A.Y[i] = x[1]; A.Z[i] = x[2];
  A.u3x[i] = a2[0]; A.u3y[i] = a2[1]; A.u3z[i] = a2[2];
  A.u2x[i] = a1[0]; A.u2y[i] = a1[1]; A.u2z[i] = a1[2];
  A.u1x[i] = u[0]; A.u1y[i] = u[1]; A.u1z[i] = u[2];
}

__global__ void k_batch_reset(int32_t *next_unit) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *next_unit = 0;
}

// source chunk in LDS.  FP64: positions and ghosts of the sources (flag 0 for
// an empty slot).  FP32C: positions and the segment e = g - x rounded to
// FP32 (zero for an empty slot, so that its term vanishes through
// |R1 x e| == 0).
template <int kSrc, bool kF64>
struct SrcChunk {
  double x[kSrc], y[kSrc], z[kSrc];
  double gx[kF64 ? kSrc : 1], gy[kF64 ? kSrc : 1], gz[kF64 ? kSrc : 1];
  int32_t ok[kF64 ? kSrc : 1];
  float ex[kF64 ? 1 : kSrc], ey[kF64 ? 1 : kSrc], ez[kF64 ? 1 : kSrc];
};

// One wavefront (one workgroup of 64 lanes) per work unit of 64 * kPer
// targets, lane l holding targets first + l + 64 k; kSrc sources per LDS
// chunk.  Every branch that decides which unit a wavefront works on, or
// whether it leaves the loop, is on readfirstlane values: the queue's
// atomicAdd is executed by all lanes (lane 0 adds 1, the others 0) and the
// loop exits on a scalar compare, so the loop has no divergent exit and no
// lane can be separated from the atomic.
// The register budget is that of 4 waves per SIMD for every tile shape;
// the 512-source chunks use enough LDS (18-25 KB per one-wave workgroup) to
// allow only 2-3, and the compiler reports that it missed the target.
// Budgeting those shapes for their real occupancy made the fastest shape
// (2 targets per lane, 512 sources) 4% slower, so they keep the 4-wave
// budget.
template <int kPer, int kSrc, bool kF64>
__global__ void __launch_bounds__(64, 4) k_batch_step(BatchArgs A, Shells S) {
  __shared__ SrcChunk<kSrc, kF64> t;
  const int lane = threadIdx.x;
  while (true) {
    const int unit = __builtin_amdgcn_readfirstlane(atomicAdd(A.next_unit, lane == 0 ? 1 : 0));
    if (unit >= A.nunits) return;  // the queue only grows: every wavefront gets here
    const int sys = A.unit_sys[unit];
    const int off = A.sys_off[sys], end = off + A.sys_n[sys];
    int ti[kPer], bi[kPer];
    bool act[kPer];
    // cur: the sum of the shell being summed (a fixed accumulator per pair,
    // no select between the central and the image sums); the central
    // shell's is moved to uc / ac when that shell ends
    double x[kPer][3], ul[kPer][3], uc[kPer][3], cur[kPer][3];
    AccF64 ac[kPer], acur[kPer];
    for (int k = 0; k < kPer; ++k) {
      const int i = A.unit_first[unit] + lane + 64 * k;
      ti[k] = i;
      bi[k] = -1;
      act[k] = i < end && A.inf[i] != 0;
      for (int c = 0; c < 3; ++c) x[k][c] = ul[k][c] = uc[k][c] = cur[k][c] = 0.;
      if (act[k]) {
        const int a = A.inf[i] - 1, b = A.beh[i] - 1;
        const int aa = A.inf[a] - 1, bb = A.beh[b] - 1;
        x[k][0] = A.X[i]; x[k][1] = A.Y[i]; x[k][2] = A.Z[i];
        double gi[3], gb[3], gii[3], gbb[3];
        ghost_of(A, a, x[k], gi);
        ghost_of(A, b, x[k], gb);
        ghost_of(A, aa, x[k], gii);
        ghost_of(A, bb, x[k], gbb);
        lia_term<double>(A.qc, A.corea, A.order, x[k], gi, gb, gii, gbb, ul[k]);
        bi[k] = b;
      }
    }
    for (int sh = 0; sh < S.n; ++sh) {
      const double s[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
      double xs[kPer][3];
      for (int k = 0; k < kPer; ++k)
        for (int c = 0; c < 3; ++c) xs[k][c] = x[k][c] - s[c];
      for (int base = off; base < end; base += kSrc) {
        __syncthreads();  // the previous chunk has been read
        for (int q = lane; q < kSrc; q += 64) {
          const int j = base + q;
          const bool in = j < end && A.inf[j] != 0;
          double xj[3] = {0., 0., 0.}, gj[3] = {0., 0., 0.};
          if (in) {
            xj[0] = A.X[j]; xj[1] = A.Y[j]; xj[2] = A.Z[j];
            ghost_of(A, A.inf[j] - 1, xj, gj);
          }
          t.x[q] = xj[0]; t.y[q] = xj[1]; t.z[q] = xj[2];
          if constexpr (kF64) {
            t.gx[q] = gj[0]; t.gy[q] = gj[1]; t.gz[q] = gj[2];
            t.ok[q] = in;
          } else {
            t.ex[q] = in ? (float)(gj[0] - xj[0]) : 0.f;
            t.ey[q] = in ? (float)(gj[1] - xj[1]) : 0.f;
            t.ez[q] = in ? (float)(gj[2] - xj[2]) : 0.f;
          }
        }
        __syncthreads();  // the chunk is in LDS
        const int m = min(kSrc, end - base);
        for (int jj = 0; jj < m; ++jj) {
          const int j = base + jj;
          if constexpr (kF64) {
            if (!t.ok[jj]) continue;
            const double xj[3] = {t.x[jj], t.y[jj], t.z[jj]};
            const double gj[3] = {t.gx[jj], t.gy[jj], t.gz[jj]};
            for (int k = 0; k < kPer; ++k) {
              if (!act[k] || j == bi[k] || (sh == 0 && j == ti[k])) continue;
              if (sh == 0) pair64<false>(x[k], xj, gj, s, A.qc, cur[k]);
              else pair64<true>(x[k], xj, gj, s, A.qc, cur[k]);
            }
          } else {
            const double sx = t.x[jj], sy = t.y[jj], sz = t.z[jj];
            const float ex = t.ex[jj], ey = t.ey[jj], ez = t.ez[jj];
            for (int k = 0; k < kPer; ++k) {
              if (!act[k] || j == bi[k] || (sh == 0 && j == ti[k])) continue;
              pair32((float)(sx - xs[k][0]), (float)(sy - xs[k][1]), (float)(sz - xs[k][2]),
                     ex, ey, ez, acur[k]);
            }
          }
        }
      }
      if (sh == 0) {
        for (int k = 0; k < kPer; ++k) {
          for (int c = 0; c < 3; ++c) {
            uc[k][c] = cur[k][c];
            cur[k][c] = 0.;
          }
          ac[k] = acur[k];
          acur[k] = AccF64();
        }
      }
    }
    for (int k = 0; k < kPer; ++k) {
      const int i = ti[k];
      double us[3] = {cur[k][0], cur[k][1], cur[k][2]};  // the image shells' sum
      if (act[k]) {
        if constexpr (!kF64) {
          // the two segments next to the target's own (infront(i) ->
          // infront(infront(i)) and behind(behind(i)) -> behind(i)) in FP64:
          // their FP32 terms (exact FP64 products) are taken back out of
          // the central sum and the FP64 terms added
          double adj[3] = {0., 0., 0.};
          const int fi = A.inf[i] - 1, bb = A.beh[bi[k]] - 1;
          const int js[2] = {fi, bb};
          const double z3[3] = {0., 0., 0.};
          for (int q = 0; q < (bb == fi ? 1 : 2); ++q) {
            const int j = js[q];
            if (j == i || j == bi[k]) continue;
            const double xj[3] = {A.X[j], A.Y[j], A.Z[j]};
            double gj[3];
            ghost_of(A, A.inf[j] - 1, xj, gj);
            AccF64 t;
            pair32((float)(xj[0] - x[k][0]), (float)(xj[1] - x[k][1]), (float)(xj[2] - x[k][2]),
                   (float)(gj[0] - xj[0]), (float)(gj[1] - xj[1]), (float)(gj[2] - xj[2]), t);
            ac[k].s0 -= t.s0; ac[k].s1 -= t.s1; ac[k].s2 -= t.s2;
            pair64<false>(x[k], xj, gj, z3, A.qc, adj);
          }
          const double g4 = A.qc / (4. * kPi);
          ac[k].get(uc[k][0], uc[k][1], uc[k][2]);
          acur[k].get(us[0], us[1], us[2]);
          for (int c = 0; c < 3; ++c) { uc[k][c] = uc[k][c] * g4 + adj[c]; us[c] *= g4; }
        }
        // u = u_LI + u_bs(central); u = u + u_bs(images)  (timestep.f90:225-240)
        double u0 = ul[k][0] + uc[k][0], u1 = ul[k][1] + uc[k][1], u2 = ul[k][2] + uc[k][2];
        if (S.n > 1) { u0 = u0 + us[0]; u1 = u1 + us[1]; u2 = u2 + us[2]; }
        A.ux[i] = u0; A.uy[i] = u1; A.uz[i] = u2;
      } else if (i < end) {
        A.ux[i] = A.uy[i] = A.uz[i] = 0.;
      }
    }
    // the wavefront that finishes the last unit of a system moves its points
    __threadfence();
    const int done =
        __builtin_amdgcn_readfirstlane(atomicAdd(A.sys_done + sys, lane == 0 ? 1 : 0));
    if (done + 1 == A.sys_units[sys]) {
      __threadfence();
      if (A.do_ab)
        for (int p = off + lane; p < end; p += 64) ab_point(A, p);
      atomicExch(A.sys_done + sys, 0);  // every lane stores the same 0
    }
  }
}

// tile shapes: targets per lane, sources per LDS chunk
struct Variant {
  int per, src;
};
constexpr Variant kVariants[] = {{1, 128}, {1, 256}, {1, 512}, {2, 256}, {2, 512}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

template <bool kF64>
hipError_t launch_variant(int v, int grid, hipStream_t st, const BatchArgs &A, const Shells &S) {
  const dim3 g(grid), b(64);
  switch (v) {
    case 0: hipLaunchKernelGGL((k_batch_step<1, 128, kF64>), g, b, 0, st, A, S); break;
    case 1: hipLaunchKernelGGL((k_batch_step<1, 256, kF64>), g, b, 0, st, A, S); break;
    case 2: hipLaunchKernelGGL((k_batch_step<1, 512, kF64>), g, b, 0, st, A, S); break;
    case 3: hipLaunchKernelGGL((k_batch_step<2, 256, kF64>), g, b, 0, st, A, S); break;
    default: hipLaunchKernelGGL((k_batch_step<2, 512, kF64>), g, b, 0, st, A, S); break;
  }
  return hipGetLastError();
}

}  // namespace

struct Batch {
  qv_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  int32_t nsys = 0, ntot = 0;
  std::vector<int32_t> h_off, h_n;
  int32_t *sys_off = nullptr, *sys_n = nullptr, *sys_units = nullptr, *sys_done = nullptr;
  int32_t *unit_sys = nullptr, *unit_first = nullptr, *next_unit = nullptr;
  int32_t nunits = 0;
  int variant = -1;
  int grid = 0;
  bool uploaded = false;
  Vec3 x, u, u1, u2, u3;
  int32_t *inf = nullptr, *beh = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0, graph_variant = -1;
  float last_ms = 0.f;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

}  // namespace qv

using namespace qv;

namespace {

#define BTRY(expr)                                 \
  do {                                             \
    if ((expr) != hipSuccess) return QV_ERR_HIP;   \
  } while (0)

template <class T>
hipError_t balloc(T *&p, size_t n) {
  return hipMalloc(&p, (n ? n : 1) * sizeof(T));
}

hipError_t balloc3(Vec3 &v, size_t n) {
  hipError_t e;
  if ((e = balloc(v.x, n)) != hipSuccess) return e;
  if ((e = balloc(v.y, n)) != hipSuccess) return e;
  return balloc(v.z, n);
}

BatchArgs make_args(const Batch &b, int do_ab) {
  BatchArgs A;
  A.nunits = b.nunits;
  A.qc = b.p.quant_circ;
  A.corea = b.p.corea;
  A.L = b.p.box_size;
  A.dt = b.p.dt;
  A.order = b.p.deriv_order;
  A.px = b.p.boundary == QV_BC_PERIODIC;
  A.py = A.px || b.p.boundary == QV_BC_OPENX;
  A.pz = b.p.boundary != QV_BC_OPEN;
  A.unit_sys = b.unit_sys;
  A.unit_first = b.unit_first;
  A.sys_off = b.sys_off;
  A.sys_n = b.sys_n;
  A.sys_units = b.sys_units;
  A.X = b.x.x; A.Y = b.x.y; A.Z = b.x.z;
  A.ux = b.u.x; A.uy = b.u.y; A.uz = b.u.z;
  A.u1x = b.u1.x; A.u1y = b.u1.y; A.u1z = b.u1.z;
  A.u2x = b.u2.x; A.u2y = b.u2.y; A.u2z = b.u2.z;
  A.u3x = b.u3.x; A.u3y = b.u3.y; A.u3z = b.u3.z;
  A.inf = b.inf;
  A.beh = b.beh;
  A.next_unit = b.next_unit;
  A.sys_done = b.sys_done;
  A.do_ab = do_ab;
  return A;
}

// the work-unit queue for a tile size: units of the largest systems first
// (longest-processing-time order over the persistent workgroups)
int build_units(Batch &b, int tile) {
  std::vector<int32_t> order(b.nsys);
  for (int s = 0; s < b.nsys; ++s) order[s] = s;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return b.h_n[a] > b.h_n[c]; });
  std::vector<int32_t> us, uf, units(b.nsys);
  for (int s : order) {
    const int k = (b.h_n[s] + tile - 1) / tile;
    units[s] = k;
    for (int q = 0; q < k; ++q) {
      us.push_back(s);
      uf.push_back(b.h_off[s] + q * tile);
    }
  }
  BTRY(hipStreamSynchronize(b.stream));
  (void)hipFree(b.unit_sys);
  (void)hipFree(b.unit_first);
  b.unit_sys = b.unit_first = nullptr;
  b.nunits = (int32_t)us.size();
  BTRY(balloc(b.unit_sys, us.size()));
  BTRY(balloc(b.unit_first, uf.size()));
  BTRY(hipMemcpy(b.unit_sys, us.data(), us.size() * 4, hipMemcpyHostToDevice));
  BTRY(hipMemcpy(b.unit_first, uf.data(), uf.size() * 4, hipMemcpyHostToDevice));
  BTRY(hipMemcpy(b.sys_units, units.data(), units.size() * 4, hipMemcpyHostToDevice));
  return QV_OK;
}

int select_variant(Batch &b, int v) {
  if (v == b.variant) return QV_OK;
  const int rc = build_units(b, 64 * kVariants[v].per);
  if (rc == QV_OK) b.variant = v;
  return rc;
}

int launch_step(Batch &b, int do_ab) {
  const BatchArgs A = make_args(b, do_ab);
  const Shells S = make_shells(b.p);
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, b.stream, b.next_unit);
  BTRY(hipGetLastError());
  if (b.p.precision == QV_FP64)
    BTRY(launch_variant<true>(b.variant, b.grid, b.stream, A, S));
  else
    BTRY(launch_variant<false>(b.variant, b.grid, b.stream, A, S));
  return QV_OK;
}

int get3(const Vec3 &v, int n, double *out) {
  std::vector<double> c0(n), c1(n), c2(n);
  BTRY(hipMemcpy(c0.data(), v.x, n * 8, hipMemcpyDeviceToHost));
  BTRY(hipMemcpy(c1.data(), v.y, n * 8, hipMemcpyDeviceToHost));
  BTRY(hipMemcpy(c2.data(), v.z, n * 8, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) {
    out[3 * i] = c0[i];
    out[3 * i + 1] = c1[i];
    out[3 * i + 2] = c2[i];
  }
  return QV_OK;
}

int put3(Vec3 &v, int n, const double *in) {
  std::vector<double> c0(n), c1(n), c2(n);
  for (int i = 0; i < n; ++i) {
    c0[i] = in ? in[3 * i] : 0.;
    c1[i] = in ? in[3 * i + 1] : 0.;
    c2[i] = in ? in[3 * i + 2] : 0.;
  }
  BTRY(hipMemcpy(v.x, c0.data(), n * 8, hipMemcpyHostToDevice));
  BTRY(hipMemcpy(v.y, c1.data(), n * 8, hipMemcpyHostToDevice));
  BTRY(hipMemcpy(v.z, c2.data(), n * 8, hipMemcpyHostToDevice));
  return QV_OK;
}

}  // namespace

extern "C" {

int qv_batch_create(const qv_params *params, int32_t nsys, const int32_t *counts,
                    qv_batch **out) {
  if (!out) return QV_ERR_ARG;
  *out = nullptr;
  if (!params || nsys < 1 || !counts) return QV_ERR_ARG;
  const qv_params &p = *params;
  if (p.velocity != QV_VEL_BS || !(p.box_size > 0.) || !(p.corea > 0.) ||
      (p.deriv_order != 2 && p.deriv_order != 4) ||
      (p.precision != QV_FP64 && p.precision != QV_FP32C) || p.boundary < QV_BC_OPEN ||
      p.boundary > QV_BC_OPENXY || p.flags != 0)
    return QV_ERR_ARG;
  // the persistent launch wraps periodic axes but cannot remove a loop that
  // leaves through an open face (boundary_loop_remove): open and periodic
  // ensembles only
  if (p.boundary == QV_BC_OPENX || p.boundary == QV_BC_OPENXY) return QV_ERR_ARG;
  int64_t tot = 0;
  for (int s = 0; s < nsys; ++s) {
    if (counts[s] < 3) return QV_ERR_ARG;
    tot += counts[s];
  }
  if (tot > (int64_t)1 << 30) return QV_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return QV_ERR_NO_DEVICE;
  Batch *b = new (std::nothrow) Batch();
  if (!b) return QV_ERR_HIP;
  *out = reinterpret_cast<qv_batch *>(b);  // qv_batch_destroy frees a partial create
  b->p = p;
  b->nsys = nsys;
  b->ntot = (int32_t)tot;
  b->h_off.resize(nsys);
  b->h_n.assign(counts, counts + nsys);
  for (int s = 0, o = 0; s < nsys; o += counts[s], ++s) b->h_off[s] = o;
  BTRY(hipGetDevice(&b->device));
  BTRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  BTRY(hipEventCreate(&b->e0));
  BTRY(hipEventCreate(&b->e1));
  for (Vec3 *v : {&b->x, &b->u, &b->u1, &b->u2, &b->u3}) BTRY(balloc3(*v, b->ntot));
  BTRY(balloc(b->inf, b->ntot));
  BTRY(balloc(b->beh, b->ntot));
  BTRY(balloc(b->sys_off, nsys));
  BTRY(balloc(b->sys_n, nsys));
  BTRY(balloc(b->sys_units, nsys));
  BTRY(balloc(b->sys_done, nsys));
  BTRY(balloc(b->next_unit, 1));
  BTRY(hipMemcpy(b->sys_off, b->h_off.data(), nsys * 4, hipMemcpyHostToDevice));
  BTRY(hipMemcpy(b->sys_n, b->h_n.data(), nsys * 4, hipMemcpyHostToDevice));
  BTRY(hipMemsetAsync(b->sys_done, 0, nsys * 4, b->stream));
  BTRY(hipStreamSynchronize(b->stream));
  int cus = 0;
  BTRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  b->grid = 8 * std::max(cus, 1);
  return select_variant(*b, 1);
}

int qv_batch_destroy(qv_batch *batch) {
  if (!batch) return QV_OK;
  Batch *b = reinterpret_cast<Batch *>(batch);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (Vec3 *v : {&b->x, &b->u, &b->u1, &b->u2, &b->u3}) {
    (void)hipFree(v->x);
    (void)hipFree(v->y);
    (void)hipFree(v->z);
  }
  for (void *q : {(void *)b->inf, (void *)b->beh, (void *)b->sys_off, (void *)b->sys_n,
                  (void *)b->sys_units, (void *)b->sys_done, (void *)b->next_unit,
                  (void *)b->unit_sys, (void *)b->unit_first})
    (void)hipFree(q);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return QV_OK;
}

int qv_batch_upload(qv_batch *batch, const double *x, const int32_t *infront,
                    const int32_t *behind) {
  if (!batch || !x || !infront || !behind) return QV_ERR_ARG;
  Batch *b = reinterpret_cast<Batch *>(batch);
  std::vector<int32_t> gi(b->ntot), gb(b->ntot);
  for (int s = 0; s < b->nsys; ++s) {
    const int off = b->h_off[s], n = b->h_n[s];
    for (int q = 0; q < n; ++q) {
      const int32_t a = infront[off + q], c = behind[off + q];
      if (a == 0) {
        if (c != 0) return QV_ERR_ARG;
        gi[off + q] = gb[off + q] = 0;
        continue;
      }
      if (a < 1 || a > n || c < 1 || c > n) return QV_ERR_ARG;
      if (infront[off + a - 1] == 0 || infront[off + c - 1] == 0) return QV_ERR_ARG;
      gi[off + q] = a + off;
      gb[off + q] = c + off;
    }
  }
  BTRY(hipStreamSynchronize(b->stream));
  BTRY(hipMemcpy(b->inf, gi.data(), b->ntot * 4, hipMemcpyHostToDevice));
  BTRY(hipMemcpy(b->beh, gb.data(), b->ntot * 4, hipMemcpyHostToDevice));
  int rc = put3(b->x, b->ntot, x);
  for (Vec3 *v : {&b->u, &b->u1, &b->u2, &b->u3})
    if (rc == QV_OK) rc = put3(*v, b->ntot, nullptr);
  b->uploaded = rc == QV_OK;
  return rc;
}

int qv_batch_velocity(qv_batch *batch, double *u) {
  if (!batch) return QV_ERR_ARG;
  Batch *b = reinterpret_cast<Batch *>(batch);
  if (!b->uploaded) return QV_ERR_STATE;
  const int rc = launch_step(*b, 0);
  if (rc) return rc;
  BTRY(hipStreamSynchronize(b->stream));
  return u ? get3(b->u, b->ntot, u) : QV_OK;
}

int qv_batch_step(qv_batch *batch, int32_t nsteps, int32_t use_graph) {
  if (!batch || nsteps < 0) return QV_ERR_ARG;
  Batch *b = reinterpret_cast<Batch *>(batch);
  if (!b->uploaded) return QV_ERR_STATE;
  BTRY(hipEventRecord(b->e0, b->stream));
  int k = 0;
  if (use_graph && nsteps > 0) {
    const int per = std::min<int32_t>(nsteps, 10);
    if (!b->exec || b->graph_steps != per || b->graph_variant != b->variant) {
      if (b->exec) (void)hipGraphExecDestroy(b->exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
      BTRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = QV_OK;
      for (int q = 0; q < per && rc == QV_OK; ++q) rc = launch_step(*b, 1);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc != QV_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc != QV_OK ? rc : QV_ERR_HIP;
      }
      b->graph = g;
      BTRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_steps = per;
      b->graph_variant = b->variant;
    }
    for (; k + per <= nsteps; k += per) BTRY(hipGraphLaunch(b->exec, b->stream));
  }
  for (; k < nsteps; ++k) {
    const int rc = launch_step(*b, 1);
    if (rc) return rc;
  }
  BTRY(hipEventRecord(b->e1, b->stream));
  BTRY(hipStreamSynchronize(b->stream));
  BTRY(hipEventElapsedTime(&b->last_ms, b->e0, b->e1));
  return QV_OK;
}

int qv_batch_download(qv_batch *batch, double *x, double *u) {
  if (!batch) return QV_ERR_ARG;
  Batch *b = reinterpret_cast<Batch *>(batch);
  if (!b->uploaded) return QV_ERR_STATE;
  BTRY(hipStreamSynchronize(b->stream));
  int rc = x ? get3(b->x, b->ntot, x) : QV_OK;
  if (rc == QV_OK && u) rc = get3(b->u, b->ntot, u);
  return rc;
}

int qv_batch_tune(qv_batch *batch, int32_t trial_steps, int32_t *chosen, float *ms) {
  if (!batch || trial_steps < 1) return QV_ERR_ARG;
  Batch *b = reinterpret_cast<Batch *>(batch);
  if (!b->uploaded) return QV_ERR_STATE;
  float best = 0.f;
  int best_v = -1;
  for (int v = 0; v < kNumVariants; ++v) {
    int rc = select_variant(*b, v);
    if (rc == QV_OK) rc = launch_step(*b, 0);  // warm-up
    if (rc) return rc;
    BTRY(hipEventRecord(b->e0, b->stream));
    for (int k = 0; k < trial_steps; ++k)
      if ((rc = launch_step(*b, 0)) != QV_OK) return rc;
    BTRY(hipEventRecord(b->e1, b->stream));
    BTRY(hipEventSynchronize(b->e1));
    float t = 0.f;
    BTRY(hipEventElapsedTime(&t, b->e0, b->e1));
    t /= trial_steps;
    if (ms) ms[v] = t;
    if (best_v < 0 || t < best) {
      best = t;
      best_v = v;
    }
  }
  if (chosen) *chosen = best_v;
  return select_variant(*b, best_v);
}

int qv_batch_set_variant(qv_batch *batch, int32_t variant) {
  if (!batch || variant < 0 || variant >= kNumVariants) return QV_ERR_ARG;
  return select_variant(*reinterpret_cast<Batch *>(batch), variant);
}

int qv_batch_info(qv_batch *batch, int32_t *nvariants, int32_t *variant, int32_t *tile,
                  int32_t *nunits, float *last_ms) {
  if (!batch) return QV_ERR_ARG;
  const Batch *b = reinterpret_cast<const Batch *>(batch);
  if (nvariants) *nvariants = kNumVariants;
  if (variant) *variant = b->variant;
  if (tile) *tile = 64 * kVariants[b->variant].per;
  if (nunits) *nunits = b->nunits;
  if (last_ms) *last_ms = b->last_ms;
  return QV_OK;
}

}  // extern "C"
