// qv_walk.hip -- Barnes-Hut tree walk (tree_walk, tree.f90:336-421) for every
// active point and every periodic image shell (timestep.f90:276-307), fused
// with the final composition of calc_velocity.
//
// Wavefront-cooperative DFS.  A wavefront owns 64 targets that are adjacent
// in the tree's sorted order, so their walks overlap.  The wavefront keeps
// one DFS stack in LDS whose entries are (node, end of the sibling list,
// lane mask); the node index is uniform across the wave, so every node
// record is read with scalar loads and broadcast to the lanes, and each lane
// evaluates the node only if its bit is set.  A lane that accepts a cell or
// meets a leaf leaves the subtree; the lanes that open the cell form the
// mask of the children entry.  Each lane therefore visits exactly the nodes
// of its own recursive walk, in the same DFS order, and accumulates its
// terms in the reference's order.
//
// Decisions: a node is accepted when the target's squared distance
// d2 = (dx^2 + dy^2) + dz^2 (the reference's operand of sqrt) is >= the
// node's threshold T, the smallest d2 for which the reference's test
// w/(d-g) < theta holds (precomputed exactly at build time); nodes with
// T < 0 take the full test.  The decisions are bit-identical to the
// reference's.
//
// QV_FP64: leaf terms in the reference's operation order (bit-identical),
// multipoles as [circ x r - eps:C] / d^3 + 3 r x (C^T r) / d^5 (equal to the
// reference's double loop up to rounding).
// QV_FP32C: relative vectors in FP64, terms in FP32 with compensated sums.
#include "qv_internal.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

struct WalkArgs {
  int nact;
  int64_t nnodes;
  double qc, theta;
  int extra;
  const int32_t *slots;
  const double *xs, *ys, *zs;       // target positions, sorted order
  const int32_t *inf, *beh;
  const double4 *hot;
  const int2 *meta;
  const double *mp;
  const float *mp32;
  const double2 *geo;
  const double *ulx, *uly, *ulz;    // slot order
  double *ux, *uy, *uz, *bx, *by, *bz;
  unsigned long long *evals;
  int32_t *flags;
};

__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t uni64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// tree_walk's acceptance test in the reference's operation order
__device__ inline bool accept_full(double d2, double w, double g, double theta, int extra) {
  const double d = sqrt(d2);
  double t;
  if (extra) {
    if (d - g > 0.) t = w / (d - g);
    else t = w / d;
  } else {
    t = w / d;
  }
  return t < theta;
}

// leaf term, tree.f90:380-392, reference operation order
__device__ inline void leaf64(const double xi[3], double cx, double cy, double cz, double gx,
                              double gy, double gz, const double s[3], double qc, double u[3]) {
  const double R1[3] = {(cx + s[0]) - xi[0], (cy + s[1]) - xi[1], (cz + s[2]) - xi[2]};
  const double R2[3] = {(gx + s[0]) - xi[0], (gy + s[1]) - xi[1], (gz + s[2]) - xi[2]};
  const double a = sqrt((R1[0] * R1[0] + R1[1] * R1[1]) + R1[2] * R1[2]);
  const double b = sqrt((R2[0] * R2[0] + R2[1] * R2[1]) + R2[2] * R2[2]);
  const double c = (a * b) * ((a * b) + ((R1[0] * R2[0] + R1[1] * R2[1]) + R1[2] * R2[2]));
  const double f = (a + b) / c;
  const double w0 = f * (R1[1] * R2[2] - R1[2] * R2[1]);
  const double w1 = f * (R1[2] * R2[0] - R1[0] * R2[2]);
  const double w2 = f * (R1[0] * R2[1] - R1[1] * R2[0]);
  u[0] = u[0] + ((w0 * qc) / 4.) / kPi;
  u[1] = u[1] + ((w1 * qc) / 4.) / kPi;
  u[2] = u[2] + ((w2 * qc) / 4.) / kPi;
}

// cell term, tree.f90:400-411; r = x_i - (cent + shift), d = |r|
__device__ inline void cell64(const double r[3], double d2, const double *__restrict__ m,
                              double qc, double u[3]) {
  const double d = sqrt(d2);
  const double d3 = (d * d) * d;
  const double d5 = (((d * d) * d) * d) * d;
  const double id3 = 1. / d3, id5 = 1. / d5;
  const double c0 = m[0], c1 = m[1], c2 = m[2];
  const double *C = m + 3;  // C[3*i + k] = Cij(i,k)
  // eps(:,k,l) C(k,l): antisymmetric part
  const double A0 = C[5] - C[7], A1 = C[6] - C[2], A2 = C[1] - C[3];
  // v_l = dot(r, C(:,l))
  const double v0 = (r[0] * C[0] + r[1] * C[3]) + r[2] * C[6];
  const double v1 = (r[0] * C[1] + r[1] * C[4]) + r[2] * C[7];
  const double v2 = (r[0] * C[2] + r[1] * C[5]) + r[2] * C[8];
  const double t0 = ((c1 * r[2] - c2 * r[1]) - A0) * id3 + (3. * (r[1] * v2 - r[2] * v1)) * id5;
  const double t1 = ((c2 * r[0] - c0 * r[2]) - A1) * id3 + (3. * (r[2] * v0 - r[0] * v2)) * id5;
  const double t2 = ((c0 * r[1] - c1 * r[0]) - A2) * id3 + (3. * (r[0] * v1 - r[1] * v0)) * id5;
  u[0] = u[0] + ((t0 * qc) / 4.) / kPi;
  u[1] = u[1] + ((t1 * qc) / 4.) / kPi;
  u[2] = u[2] + ((t2 * qc) / 4.) / kPi;
}

// FP32 leaf: R1 = x_p + s - x_i formed in FP64, e = g_p - x_p in FP32
template <class Acc>
__device__ inline void leaf32(double r0d, double r1d, double r2d, const float *__restrict__ m,
                              Acc &acc) {
  const float r0 = (float)r0d, r1 = (float)r1d, r2 = (float)r2d;
  const float e0 = m[0], e1 = m[1], e2 = m[2];
  const float c0 = r1 * e2 - r2 * e1, c1 = r2 * e0 - r0 * e2, c2 = r0 * e1 - r1 * e0;
  const float s0 = r0 + e0, s1 = r1 + e1, s2 = r2 + e2;
  const float a2 = (r0 * r0 + r1 * r1) + r2 * r2;
  const float b2 = (s0 * s0 + s1 * s1) + s2 * s2;
  const float ia = __builtin_amdgcn_rsqf(a2), ib = __builtin_amdgcn_rsqf(b2);
  const float a = a2 * ia, b = b2 * ib;
  const float ab = a * b;
  const float dt = (r0 * s0 + r1 * s1) + r2 * s2;
  const float g = (a + b) * (ia * ib);
  float f;
  if (dt >= 0.f) {
    f = g * __builtin_amdgcn_rcpf(ab + dt);
  } else {
    // target projects inside the segment: ab + R1.R2 = |R1 x e|^2 / (ab - R1.R2)
    const float cc = (c0 * c0 + c1 * c1) + c2 * c2;
    f = g * (ab - dt) * __builtin_amdgcn_rcpf(cc);
  }
  acc.add(f * c0, f * c1, f * c2);
}

// FP32 cell: mp32 = circ(3), A(3) = eps:C, C(9)
template <class Acc>
__device__ inline void cell32(const double rd[3], double d2, const float *__restrict__ m,
                              Acc &acc) {
  const float r0 = (float)rd[0], r1 = (float)rd[1], r2 = (float)rd[2];
  const float id = __builtin_amdgcn_rsqf((float)d2);
  const float id2 = id * id;
  const float id3 = id2 * id;
  const float id5 = id3 * id2;
  const float *C = m + 6;
  const float v0 = (r0 * C[0] + r1 * C[3]) + r2 * C[6];
  const float v1 = (r0 * C[1] + r1 * C[4]) + r2 * C[7];
  const float v2 = (r0 * C[2] + r1 * C[5]) + r2 * C[8];
  const float t0 = ((m[1] * r2 - m[2] * r1) - m[3]) * id3 + (3.f * id5) * (r1 * v2 - r2 * v1);
  const float t1 = ((m[2] * r0 - m[0] * r2) - m[4]) * id3 + (3.f * id5) * (r2 * v0 - r0 * v2);
  const float t2 = ((m[0] * r1 - m[1] * r0) - m[5]) * id3 + (3.f * id5) * (r0 * v1 - r1 * v0);
  acc.add(t0, t1, t2);
}

struct Acc64 {
  double s[3] = {0., 0., 0.};
};

template <int kMode, class Acc>
__global__ void __launch_bounds__(kBlock) k_walk(WalkArgs A, Shells S) {
  __shared__ int st_node[kWavesPerBlock][kStackDepth];
  __shared__ int st_end[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t st_mask[kWavesPerBlock][kStackDepth];
  // XCD-aware order: consecutive blocks of the target list go to one XCD
  const int G = gridDim.x;
  const int per = (G + 7) / 8;
  const int b = blockIdx.x;
  const int lb = (b % 8) * per + (b / 8);
  const int wid = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int s = lb * kBlock + threadIdx.x;
  const bool valid = lb < G && s < A.nact;
  int t = 0, bi = -1, fi = -1;
  double xi[3] = {0., 0., 0.};
  if (valid) {
    t = A.slots[s];
    xi[0] = A.xs[s]; xi[1] = A.ys[s]; xi[2] = A.zs[s];
    bi = A.beh[t] - 1;
    fi = A.inf[t] - 1;
  }
  const uint64_t active = __ballot(valid);
  if (active == 0) return;
  double uc[3] = {0., 0., 0.}, us[3] = {0., 0., 0.};
  Acc ac, as;
  uint32_t ev = 0;
  bool singular = false;
  const double qc = A.qc;
  for (int sh = 0; sh < S.n; ++sh) {
    const double sv[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
    double *u = sh == 0 ? uc : us;
    Acc &acc = sh == 0 ? ac : as;
    int sp = 0;
    st_node[wid][0] = 0;
    st_end[wid][0] = (int)A.nnodes;
    st_mask[wid][0] = active;
    sp = 1;
    while (sp > 0) {
      --sp;
      const int n = uni(st_node[wid][sp]);
      const int e = uni(st_end[wid][sp]);
      const uint64_t M = uni64(st_mask[wid][sp]);
      const int2 meta = A.meta[n];
      const int skip = uni(meta.x), j = uni(meta.y);
      if (skip != e) {
        st_node[wid][sp] = skip;
        st_end[wid][sp] = e;
        st_mask[wid][sp] = M;
        ++sp;
      }
      const double4 h = A.hot[n];
      const bool mine = (M >> lane) & 1ull;
      const double cs0 = h.x + sv[0], cs1 = h.y + sv[1], cs2 = h.z + sv[2];
      const double r[3] = {xi[0] - cs0, xi[1] - cs1, xi[2] - cs2};
      const double d2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
      if (j >= 0) {
        // leaf: the single point's segment unless it is i, behind(i), infront(i)
        if (mine && j != t && j != bi && j != fi) {
          if (d2 < 1e-30 && sqrt(d2) < 2.220446049250313e-16) singular = true;
          if (kMode == QV_FP64) {
            const double *m = A.mp + 12 * (int64_t)n;
            leaf64(xi, h.x, h.y, h.z, m[0], m[1], m[2], sv, qc, u);
          } else {
            leaf32(-r[0], -r[1], -r[2], A.mp32 + 16 * (int64_t)n, acc);
          }
          ++ev;
        }
      } else {
        bool open = false;
        if (mine) {
          bool acc_ok;
          if (h.w >= 0.) {
            acc_ok = d2 >= h.w;
          } else {
            const double2 g = A.geo[n];
            acc_ok = accept_full(d2, g.x, g.y, A.theta, A.extra);
          }
          if (acc_ok) {
            if (kMode == QV_FP64) cell64(r, d2, A.mp + 12 * (int64_t)n, qc, u);
            else cell32(r, d2, A.mp32 + 16 * (int64_t)n, acc);
            ++ev;
          } else {
            open = true;
          }
        }
        const uint64_t O = __ballot(open);
        if (O != 0 && n + 1 != skip) {
          st_node[wid][sp] = n + 1;
          st_end[wid][sp] = skip;
          st_mask[wid][sp] = O;
          ++sp;
        }
      }
    }
  }
  // evaluation count (64-bit) and error flag
  const unsigned long long evw = (unsigned long long)ev;
  unsigned long long tot = evw;
  for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
  if (lane == 0 && tot) atomicAdd(A.evals, tot);
  if (singular) atomicOr(A.flags, kFlagSingular);
  if (!valid) return;
  if (kMode != QV_FP64) {
    const double g4 = qc / (4. * kPi);
    double a0, a1, a2;
    ac.get(a0, a1, a2);
    uc[0] = a0 * g4; uc[1] = a1 * g4; uc[2] = a2 * g4;
    as.get(a0, a1, a2);
    us[0] = a0 * g4; us[1] = a1 * g4; us[2] = a2 * g4;
  }
  // calc_velocity composition (timestep.f90:273-288)
  double u0 = A.ulx[t] + uc[0], u1 = A.uly[t] + uc[1], u2 = A.ulz[t] + uc[2];
  double b0 = uc[0], b1 = uc[1], b2 = uc[2];
  if (S.n > 1) {
    u0 = u0 + us[0]; u1 = u1 + us[1]; u2 = u2 + us[2];
    b0 = b0 + us[0]; b1 = b1 + us[1]; b2 = b2 + us[2];
  }
  A.ux[t] = u0; A.uy[t] = u1; A.uz[t] = u2;
  A.bx[t] = b0; A.by[t] = b1; A.bz[t] = b2;
}

}  // namespace

hipError_t launch_walk(Ctx &c) {
  WalkArgs A;
  A.nact = c.nact;
  A.nnodes = c.tree.nnodes;
  A.qc = c.p.quant_circ;
  A.theta = c.p.tree_theta;
  A.extra = c.p.tree_extra_correction;
  A.slots = c.slot_s;
  A.xs = c.xs.x; A.ys = c.xs.y; A.zs = c.xs.z;
  A.inf = c.infront; A.beh = c.behind;
  A.hot = c.tree.hot;
  A.meta = c.tree.meta;
  A.mp = c.tree.mp;
  A.mp32 = c.tree.mp32;
  A.geo = c.tree.geo;
  A.ulx = c.uli.x; A.uly = c.uli.y; A.ulz = c.uli.z;
  A.ux = c.u.x; A.uy = c.u.y; A.uz = c.u.z;
  A.bx = c.ubs.x; A.by = c.ubs.y; A.bz = c.ubs.z;
  A.evals = reinterpret_cast<unsigned long long *>(c.evals);
  A.flags = c.counters + 1;
  Shells S = make_shells(c.p);
  int g = (c.nact + kBlock - 1) / kBlock;
  g = ((g + 7) / 8) * 8;
  if (c.p.precision == QV_FP32C)
    hipLaunchKernelGGL((k_walk<QV_FP32C, AccKahan>), dim3(g), dim3(kBlock), 0, c.stream, A, S);
  else
    hipLaunchKernelGGL((k_walk<QV_FP64, AccF64>), dim3(g), dim3(kBlock), 0, c.stream, A, S);
  return hipGetLastError();
}

}  // namespace qv
