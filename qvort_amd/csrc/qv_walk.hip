// qv_walk.hip -- Barnes-Hut tree walk (tree_walk, tree.f90:336-421) for every
// active point and every periodic image shell (timestep.f90:276-307), fused
// with the final composition of calc_velocity.
//
// Wavefront-cooperative DFS.  A wavefront owns 64 targets that are adjacent
// on a Hilbert curve through the box, so their walks overlap (4% fewer wave
// steps than 64 targets adjacent in the tree's Morton order).  Each group of
// targets is three work items, launched as one grid: the central shell's walk
// (about two thirds of the steps) in the first third of the blocks, the first
// half of the image shells in the second, the other half in the last, so that
// shorter items fill the GPU as the launch drains; k_walk_join composes the
// sums (the two image halves added before the image sum joins the central
// one, as calc_velocity adds its image sum, timestep.f90:280-288).  The walk state is
// (node, end of its sibling list, lane mask), uniform across the wave: the
// current entry lives in scalar registers and an LDS stack keeps the sibling
// lists left behind when the walk descends.  Every node record (64 bytes)
// is read with one scalar load and broadcast to the lanes, and each lane
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
// T < 0 take the full test.  Given a node's centroid the decisions are
// bit-identical to the reference's (the centroid is summed in key order
// here, in slot order by the reference, and may differ in its last bits).
//
// Per node step the record arrives with one s_load_dwordx16, the lane's
// bit of the walk mask is taken with inverse_ballot (no VALU), T < 0 is
// read from the sign of T's high word (a scalar test), and the accepting
// lanes are M & ballot(d2 >= T).  In FP32C every internal
// node's multipole is evaluated with weight 0 on the lanes that do not
// accept: a uniform branch around it saves nothing measurable on the 7% of
// steps in which no lane accepts.  Each step's term (leaf or cell) is added
// to the target's sum at one place after the branches, so the sums stay in
// the same registers on every path (no copies at the join).  The image
// shells' shift lives in vector registers in FP32C, so that cent + shift
// reads the node's centroid as the one scalar operand of each add.  The
// kernel's output pointers are read from the kernel-argument segment after
// the walk: held in scalar registers through it they would be spilled.
//
// QV_FP64: leaf terms in the reference's operation order, multipoles as
// [circ x r - eps:C] / d^3 + 3 r x (C^T r) / d^5 (the reference's double
// loop regrouped, with fused multiply-adds) with 1/d from a refined
// reciprocal square root (a couple of ulp instead of the correctly rounded
// sqrt and two divisions), the factor Gamma/(4 pi) applied once per sum.
// QV_FP32C: relative vectors in FP64, terms in FP32, per-target sums in FP64.
#include "qv_internal.h"
#include "qv_terms.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

struct WalkArgs {
  int nact;
  int groups;  // blocks of one part: the central shell's walks, then the images'
  int64_t nnodes;
  double qc, theta;
  int extra;
  const int32_t *order;             // walk order -> sorted position
  const int32_t *slots;
  const double *xs, *ys, *zs;       // target positions, sorted order
  const int32_t *inf, *beh;
  const NodeRec *rec;
  const double *mp;
  const float *mp32;
  const double2 *geo;
  const double *ulx, *uly, *ulz;    // slot order
  double *ux, *uy, *uz, *bx, *by, *bz;
  double *p2;  // the second half of the image shells' raw sums, 3 per slot
  unsigned long long *evals;  // [0] leaf + cell interactions, [1] lane visits, [2] leaves,
                              // [3] wave steps (nodes fetched by a wavefront)
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

// leaf term, tree.f90:380-392, in the reference's operation order up to the
// factor Gamma/(4 pi), which is applied once to the target's sum.  Added to u
// only where `on` (the other lanes may hold inf/nan terms).
template <bool kShift>
__device__ inline void leaf64(bool on, const double xi[3], double cx, double cy, double cz,
                              double gx, double gy, double gz, const double s[3], double t[3]) {
  // x + shift with shift = 0 is x itself, so the central shell skips the add
  const double R1[3] = {(kShift ? cx + s[0] : cx) - xi[0], (kShift ? cy + s[1] : cy) - xi[1],
                        (kShift ? cz + s[2] : cz) - xi[2]};
  const double R2[3] = {(kShift ? gx + s[0] : gx) - xi[0], (kShift ? gy + s[1] : gy) - xi[1],
                        (kShift ? gz + s[2] : gz) - xi[2]};
  const double a = sqrt((R1[0] * R1[0] + R1[1] * R1[1]) + R1[2] * R1[2]);
  const double b = sqrt((R2[0] * R2[0] + R2[1] * R2[1]) + R2[2] * R2[2]);
  const double c = (a * b) * ((a * b) + ((R1[0] * R2[0] + R1[1] * R2[1]) + R1[2] * R2[2]));
  const double f = on ? (a + b) / c : 0.;  // finite cross product, so 0 * it = 0
  t[0] = f * (R1[1] * R2[2] - R1[2] * R2[1]);
  t[1] = f * (R1[2] * R2[0] - R1[0] * R2[2]);
  t[2] = f * (R1[0] * R2[1] - R1[1] * R2[0]);
}

// FP32 terms.  The parity code of this file is compiled without
// contraction; the FP32 path has no reference rounding to keep, so it uses
// explicit fused multiply-adds.
__device__ inline float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// FP32 leaf: R1 = x_p + s - x_i and e = g_p - x_p formed in FP64, rounded once.
// Straight-segment Biot-Savart (tree.f90:380-392) with R1 x R2 = R1 x e and,
// for targets projecting inside the segment, ab + R1.R2 = |R1 x e|^2 /
// (ab - R1.R2), which keeps FP32 free of cancellation.
__device__ inline void leaf32(bool on, double r0d, double r1d, double r2d, double e0d,
                              double e1d, double e2d, float T[3], float &W) {
  const float r0 = (float)r0d, r1 = (float)r1d, r2 = (float)r2d;
  const float e0 = (float)e0d, e1 = (float)e1d, e2 = (float)e2d;
  const float c0 = fmaf_(r1, e2, -r2 * e1), c1 = fmaf_(r2, e0, -r0 * e2),
              c2 = fmaf_(r0, e1, -r1 * e0);
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
  const float cc = fmaf_(c2, c2, fmaf_(c1, c1, c0 * c0));
  const float f = (in ? g : g * (ab - dt)) * __builtin_amdgcn_rcpf(in ? ab + dt : cc);
  // lanes that are off may hold inf/nan in f (e.g. their own point); R1 x e
  // is finite, so selecting f alone zeroes their contribution
  W = on ? f : 0.f;
  T[0] = c0; T[1] = c1; T[2] = c2;
}

// FP32 cell: m6 = circ(3), eps:C(3) from the node record, C(9) from mp32.
// t = [(circ x r - eps:C) + 3 r x (C^T r) / d^2] / d^3, the 1/d^3 applied in
// the FP64 accumulation.
__device__ inline void cell32(bool on, const double rd[3], double d2, const float *m6,
                              const float *__restrict__ C, float T[3], float &W) {
  const float r0 = (float)rd[0], r1 = (float)rd[1], r2 = (float)rd[2];
  // lanes that are off use d2 = 1 (finite terms) and weight 0
  const float id = __builtin_amdgcn_rsqf(on ? (float)d2 : 1.f);
  const float id2 = id * id;
  const float q = 3.f * id2;
  W = on ? id2 * id : 0.f;
  const float v0 = fmaf_(r2, C[6], fmaf_(r1, C[3], r0 * C[0]));
  const float v1 = fmaf_(r2, C[7], fmaf_(r1, C[4], r0 * C[1]));
  const float v2 = fmaf_(r2, C[8], fmaf_(r1, C[5], r0 * C[2]));
  const float X0 = fmaf_(m6[1], r2, fmaf_(-m6[2], r1, -m6[3]));
  const float X1 = fmaf_(m6[2], r0, fmaf_(-m6[0], r2, -m6[4]));
  const float X2 = fmaf_(m6[0], r1, fmaf_(-m6[1], r0, -m6[5]));
  const float Y0 = fmaf_(r1, v2, -r2 * v1);
  const float Y1 = fmaf_(r2, v0, -r0 * v2);
  const float Y2 = fmaf_(r0, v1, -r1 * v0);
  T[0] = fmaf_(Y0, q, X0);
  T[1] = fmaf_(Y1, q, X1);
  T[2] = fmaf_(Y2, q, X2);
}

struct alignas(16) StackEntry {
  int32_t node, end;
  uint64_t mask;
};

// per-wavefront walk state
struct WalkCounts {
  uint32_t nvis = 0, nleaf = 0, ncell = 0, nstep = 0;
  bool singular = false;
};

// one shell's walk of the whole tree for the wavefront's lanes (tree_walk
// called with one shift, timestep.f90:277 / :285)
template <int kMode, bool kShift>
__device__ inline void walk_shell(const WalkArgs &A, const double sv[3], const double xi[3],
                                  int t, int bi, int fi, uint64_t active,
                                  StackEntry *st, double u[3], AccF64 &acc, WalkCounts &wc) {
  // current entry (node, end of its sibling list, lanes) in SGPRs; the LDS
  // stack holds only sibling lists left behind when descending
  int n = 0, e = (int)A.nnodes;
  uint64_t M = active;
  int sp = 0;
  typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));
  const char *recb = reinterpret_cast<const char *>(A.rec);
  const char *m32b = reinterpret_cast<const char *>(A.mp32);
  // the shift in vector registers (a lane shuffle makes it a vector value):
  // cent + shift then reads one scalar operand, with no copy of the
  // centroid into vector registers
  // (FP32C only: the FP64 walk has no vector registers to spare)
  double s_[3] = {sv[0], sv[1], sv[2]};
  if (kShift && kMode != QV_FP64) {
    s_[0] = __shfl(sv[0], 0); s_[1] = __shfl(sv[1], 0); s_[2] = __shfl(sv[2], 0);
  }
  while (true) {
    // byte offset of record n (64 B per record, 16 floats of mp32 per node:
    // the same offset), 64-bit: trees beyond 2^26 nodes
    const uint64_t off = (uint64_t)(uint32_t)n << 6;
    const u32x16 raw = *reinterpret_cast<const u32x16 *>(recb + off);
    NodeRec rec;
    __builtin_memcpy(&rec, &raw, sizeof(NodeRec));
    ++wc.nstep;
    const int skip = uni(rec.skip), j = uni(rec.j);
    const bool mine = __builtin_amdgcn_inverse_ballot_w64(M);
    wc.nvis += (uint32_t)__popcll(M);
    // r = x_i - (cent + shift), d2 as the reference forms dist (tree.f90:351-353)
    const double cs0 = kShift ? rec.cx + s_[0] : rec.cx;
    const double cs1 = kShift ? rec.cy + s_[1] : rec.cy;
    const double cs2 = kShift ? rec.cz + s_[2] : rec.cz;
    const double r[3] = {xi[0] - cs0, xi[1] - cs1, xi[2] - cs2};
    const double d2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    uint64_t O = 0;
    // this step's term, added once after the branches: FP64 t, or the FP32
    // vector T with weight W
    double t64[3] = {0., 0., 0.};
    float T[3] = {0.f, 0.f, 0.f}, W = 0.f;
    if (j >= 0) {
      // leaf: the single point's segment unless it is i, behind(i), infront(i)
      const bool on = mine && j != t && j != bi && j != fi;
      const uint64_t L = __ballot(on);
      if (L) {
        // sqrt(d2) < epsilon(0.) (tree.f90:372) is exactly d2 < 2^-104
        if (__ballot(on && d2 < 0x1p-104)) wc.singular = true;
        if (kMode == QV_FP64) {
          leaf64<kShift>(on, xi, rec.cx, rec.cy, rec.cz, rec.g[0], rec.g[1], rec.g[2], s_, t64);
        } else {
          leaf32(on, -r[0], -r[1], -r[2], rec.g[0] - rec.cx, rec.g[1] - rec.cy,
                 rec.g[2] - rec.cz, T, W);
        }
        wc.nleaf += (uint32_t)__popcll(L);
      }
    } else {
      // T < 0 (full test) read from the sign of T's high word, a scalar test
      uint64_t Acc_;
      if (uni((int32_t)(raw[7])) >= 0) {
        Acc_ = M & __ballot(d2 >= rec.T);
      } else {
        const double2 g = A.geo[n];
        Acc_ = M & __ballot(accept_full(d2, g.x, g.y, A.theta, A.extra));
      }
      const bool ok = __builtin_amdgcn_inverse_ballot_w64(Acc_);
      if (kMode == QV_FP64) {
        if (Acc_) cell64(ok, r, d2, A.mp + 12 * (int64_t)n, t64);
      } else {
        cell32(ok, r, d2, rec.m6, reinterpret_cast<const float *>(m32b + off), T, W);
      }
      wc.ncell += (uint32_t)__popcll(Acc_);
      O = M & ~Acc_;
    }
    if (kMode == QV_FP64) {
      u[0] = u[0] + t64[0]; u[1] = u[1] + t64[1]; u[2] = u[2] + t64[2];
    } else {
      const double w = (double)W;
      acc.s0 = fma(w, (double)T[0], acc.s0);
      acc.s1 = fma(w, (double)T[1], acc.s1);
      acc.s2 = fma(w, (double)T[2], acc.s2);
    }
    if (O != 0 && n + 1 != skip) {
      // descend into the children [n+1, skip) with the opening lanes
      if (skip != e) {
        st[sp].node = skip;
        st[sp].end = e;
        st[sp].mask = M;
        ++sp;
      }
      e = skip;
      n = n + 1;
      M = O;
    } else if (skip != e) {
      n = skip;  // next sibling, same lanes
    } else {
      if (sp == 0) break;
      --sp;
      n = uni(st[sp].node);
      e = uni(st[sp].end);
      M = uni64(st[sp].mask);
    }
  }
}

// at least 8 wavefronts per SIMD: the register budget that leaves hides more
// of the node loads' latency than 7 waves with spare registers (the FP32C
// walk now takes 47 VGPRs and 78 SGPRs, no spills in its node loop, which
// leaves room for 10)
template <int kMode>
__global__ void __launch_bounds__(kBlock, 8) k_walk(WalkArgs A, Shells S) {
  __shared__ StackEntry stack[kWavesPerBlock][kStackDepth];
  // blocks [0, groups) walk the central shell of their targets, blocks
  // [groups, 2 groups) and [2 groups, 3 groups) the two halves of the image
  // shells of the same targets: shorter work items per wavefront's targets,
  // the long ones first, so that the last blocks leave less of the GPU idle
  const int G = A.groups;
  const int part = (int)blockIdx.x / G;
  const bool images = part > 0;
  const int b = (int)blockIdx.x - part * G;
  // XCD-aware order: consecutive blocks of the target list go to one XCD
  const int per = G / 8;
  const int lb = (b % 8) * per + (b / 8);
  const int wid = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int w = lb * kBlock + threadIdx.x;
  const bool valid = w < A.nact;
  int t = 0, bi = -1, fi = -1;
  double xi[3] = {0., 0., 0.};
  if (valid) {
    const int s = A.order[w];
    t = A.slots[s];
    xi[0] = A.xs[s]; xi[1] = A.ys[s]; xi[2] = A.zs[s];
    bi = A.beh[t] - 1;
    fi = A.inf[t] - 1;
  }
  const uint64_t active = __ballot(valid);
  if (active == 0) return;
  WalkCounts wc;
  StackEntry *st = stack[wid];
  // the central shell's sum, or the image shells' second sum
  // (timestep.f90:276-288)
  double r[3] = {0., 0., 0.};
  {
    double u[3] = {0., 0., 0.};
    AccF64 acc;
    if (!images) {
      const double z[3] = {0., 0., 0.};
      walk_shell<kMode, false>(A, z, xi, t, bi, fi, active, st, u, acc, wc);
    } else {
      const int half = 1 + (S.n - 1) / 2;
      const int sh0 = part == 1 ? 1 : half, sh1 = part == 1 ? half : S.n;
      for (int sh = sh0; sh < sh1; ++sh) {
        const double sv[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
        walk_shell<kMode, true>(A, sv, xi, t, bi, fi, active, st, u, acc, wc);
      }
    }
    if (kMode == QV_FP64) { r[0] = u[0]; r[1] = u[1]; r[2] = u[2]; }
    else acc.get(r[0], r[1], r[2]);
  }
  // the output pointers are read from the kernel-argument segment here, after
  // the walk, so that they do not occupy scalar registers during it
  typedef const WalkArgs __attribute__((address_space(4))) KernArgs;
  KernArgs &O = *(KernArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  // interaction, visit, leaf and step counts (64-bit) and the error flag
  if (lane == 0) {
    atomicAdd(O.evals + 0, (unsigned long long)wc.nleaf + wc.ncell);
    atomicAdd(O.evals + 1, (unsigned long long)wc.nvis);
    atomicAdd(O.evals + 2, (unsigned long long)wc.nleaf);
    atomicAdd(O.evals + 3, (unsigned long long)wc.nstep);
  }
  if (wc.singular) atomicOr(O.flags, kFlagSingular);
  if (!valid) return;
  // raw sums for k_walk_join: the central one in u_bs, the images' halves in
  // u and in the part buffer
  if (part == 0) {
    O.bx[t] = r[0]; O.by[t] = r[1]; O.bz[t] = r[2];
  } else if (part == 1) {
    O.ux[t] = r[0]; O.uy[t] = r[1]; O.uz[t] = r[2];
  } else {
    O.p2[3 * (int64_t)t] = r[0]; O.p2[3 * (int64_t)t + 1] = r[1]; O.p2[3 * (int64_t)t + 2] = r[2];
  }
}

// calc_velocity's composition (timestep.f90:273-288) from the two sums
__global__ void __launch_bounds__(kBlock) k_walk_join(WalkArgs A, int nshells) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.nact) return;
  const int t = A.slots[s];
  double uc[3] = {A.bx[t], A.by[t], A.bz[t]};
  double us[3] = {0., 0., 0.};
  if (nshells > 1) {
    us[0] = A.ux[t]; us[1] = A.uy[t]; us[2] = A.uz[t];
    if (nshells > 2)
      for (int k = 0; k < 3; ++k) us[k] = us[k] + A.p2[3 * (int64_t)t + k];
  }
  // u_bs * quant_circ / 4 / pi (tree.f90:390, 411), once per sum
  for (int k = 0; k < 3; ++k) {
    uc[k] = ((uc[k] * A.qc) / 4.) / kPi;
    us[k] = ((us[k] * A.qc) / 4.) / kPi;
  }
  double u0 = A.ulx[t] + uc[0], u1 = A.uly[t] + uc[1], u2 = A.ulz[t] + uc[2];
  double b0 = uc[0], b1 = uc[1], b2 = uc[2];
  if (nshells > 1) {
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
  A.order = c.walk_order;
  A.slots = c.slot_s;
  A.xs = c.xs.x; A.ys = c.xs.y; A.zs = c.xs.z;
  A.inf = c.infront; A.beh = c.behind;
  A.rec = c.tree.rec;
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
  A.groups = g;
  A.p2 = c.walk_part;
  const dim3 grid(S.n > 2 ? 3 * g : (S.n > 1 ? 2 * g : g));
  if (c.p.precision == QV_FP32C)
    hipLaunchKernelGGL((k_walk<QV_FP32C>), grid, dim3(kBlock), 0, c.stream, A, S);
  else
    hipLaunchKernelGGL((k_walk<QV_FP64>), grid, dim3(kBlock), 0, c.stream, A, S);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_walk_join, dim3((c.nact + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream,
                     A, S.n);
  return hipGetLastError();
}

}  // namespace qv
