// qv_tree.hip -- octree build on the device (construct_tree, tree.f90:44-186).
//
// The reference builds the tree top-down by recursion: a cell with more than
// one point is split into 8 children in the order fbl fbr ftl ftr bbl bbr btl
// btr, i.e. (dx,dy,dz) = (0,0,0) (1,0,0) (0,0,1) (1,0,1) (0,1,0) (1,1,0)
// (0,1,1) (1,1,1), child index 4*dy + 2*dz + dx (tree.f90:92-114).  A child
// keeps the points of its parent that pass `x >= pos .and. x < pos+width`
// per axis (tree.f90:146-150), where pos and width follow the FP64
// recurrence corner + width/2, width/2 from (-L/2, L).
//
// Here every point computes its own path through that recurrence, one 4-bit
// digit per level: the child index, or 8 when the point fails its child's
// bounds (rounding of the recurrence or a point outside the box) and belongs
// to no cell below.  Sorting the points by path puts every cell's members in
// one contiguous range, in the reference's DFS order, with the dropped
// points of a cell after all its children.  The non-empty cells then follow
// from the longest common prefixes of neighbouring keys: the point at sorted
// position s is the first member of the cells at levels A = lcp(s-1,s)+1 ...
// min(drop-1, max(A, lcp(s,s+1)+1)), so nodes are enumerated in DFS preorder by a
// prefix sum, and each node knows its member range and the index after its
// subtree (the skip pointer of a stackless DFS).
//
// Occupancy, cell bounds and the leaf/internal split are bit-identical to
// the reference.  Moments (tree.f90:154-179): cent = sum x / n, circ =
// sum ((circ + g) - x), C = sum x (g - x)^T - cent circ^T.  The reference
// sums each node's members in slot order (a child's member list is a stable
// filter of its parent's, and the root's is f(:) itself).  Every node of at
// most 8 members does the same here, so its moments are the reference's bit
// for bit: one thread sorts the node's member slots in registers (a sorting
// network) and folds the members in that order.  Larger nodes are summed in
// key order (up to 64 members, one thread each) or in chunks in parallel,
// and differ from the reference's sums in their last bits.  circ is a sum
// of segment vectors that cancels over every closed loop inside a node, so
// its rounding is a large share of the value, and the cell term of a node
// carries it to the velocity: at 1M points, sums in key order moved the
// worst target's velocity by 1.4e-12 relative; slot order for the nodes of
// up to 8 members brings the worst point to about 2e-13 (a CPU emulation of
// the rounding: 1.6e-13 with 8, 1.2e-13 with 64 on 50,000 targets).
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "qv_sort.h"

#include "qv_internal.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

constexpr int kBlock = 256;
constexpr int kSmallNode = 64;  // nodes up to this size: one thread each
constexpr int kTinyNode = 8;    // nodes up to this size: summed in slot order

__host__ __device__ inline int grid_of(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

__device__ inline int digit_at(uint64_t k0, uint64_t k1, int lev) {
  // lev 1..32
  return lev <= 16 ? (int)((k0 >> (64 - 4 * lev)) & 15u)
                   : (int)((k1 >> (64 - 4 * (lev - 16))) & 15u);
}

// first level whose digit is 8 (the point leaves the tree there), else 33
__device__ inline int drop_level(uint64_t k0, uint64_t k1) {
  for (int lev = 1; lev <= kMaxLevels; ++lev)
    if (digit_at(k0, k1, lev) == 8) return lev;
  return kMaxLevels + 1;
}

// number of equal leading levels of two keys (0..32)
__device__ inline int lcp(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1) {
  const uint64_t d0 = a0 ^ b0;
  if (d0) return __clzll(d0) >> 2;
  const uint64_t d1 = a1 ^ b1;
  if (d1) return 16 + (__clzll(d1) >> 2);
  return 32;
}

// ---------------------------------------------------------------------------
// path keys

__global__ void __launch_bounds__(kBlock)
k_keys(int n, double L, const double *__restrict__ X, const double *__restrict__ Y,
       const double *__restrict__ Z, const int32_t *__restrict__ inf,
       uint64_t *__restrict__ key0, uint64_t *__restrict__ key1, int32_t *__restrict__ slot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  slot[i] = i;
  if (inf[i] == 0) {
    key0[i] = ~0ull;
    key1[i] = ~0ull;
    return;
  }
  const double x[3] = {X[i], Y[i], Z[i]};
  double P[3] = {-L / 2., -L / 2., -L / 2.};  // construct_tree root corner
  double w = L;
  uint64_t k0 = 0, k1 = 0;
  bool dropped = false;
  for (int lev = 1; lev <= kMaxLevels; ++lev) {
    int d = 0;
    if (!dropped) {
      const double h = w / 2.;
      int bit[3];
      bool in = true;
      double lo[3];
      for (int a = 0; a < 3; ++a) {
        const double mid = P[a] + (w / 2);  // posx+(vtree%width/2)
        if (x[a] < mid) {
          bit[a] = 0;
          lo[a] = P[a];
          in = in && (x[a] >= P[a]) && (x[a] < (P[a] + h));
        } else {
          bit[a] = 1;
          lo[a] = mid;
          in = in && (x[a] >= mid) && (x[a] < (mid + h));
        }
      }
      if (in) {
        d = 4 * bit[1] + 2 * bit[2] + bit[0];
        P[0] = lo[0]; P[1] = lo[1]; P[2] = lo[2];
        w = h;
      } else {
        d = 8;
        dropped = true;
      }
    }
    if (lev <= 16) k0 |= (uint64_t)d << (64 - 4 * lev);
    else k1 |= (uint64_t)d << (64 - 4 * (lev - 16));
  }
  key0[i] = k0;
  key1[i] = k1;
}

// points sharing all 16 levels of key0 (and still in the tree) are ordered
// by key1 then slot: one thread per such run (rare and short)
__global__ void k_deep_runs(int nact, const uint64_t *__restrict__ k0s,
                            const uint64_t *__restrict__ key1, int32_t *__restrict__ slots) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nact - 1) return;
  const uint64_t k = k0s[s];
  if (k0s[s + 1] != k) return;
  if (s > 0 && k0s[s - 1] == k) return;
  for (int lev = 1; lev <= 16; ++lev)
    if (((k >> (64 - 4 * lev)) & 15u) == 8) return;
  int e = s + 1;
  while (e < nact && k0s[e] == k) ++e;
  for (int a = s + 1; a < e; ++a) {
    const int32_t v = slots[a];
    const uint64_t kv = key1[v];
    int b = a - 1;
    while (b >= s) {
      const int32_t u = slots[b];
      const uint64_t ku = key1[u];
      if (ku < kv || (ku == kv && u < v)) break;
      slots[b + 1] = u;
      --b;
    }
    slots[b + 1] = v;
  }
}

__global__ void __launch_bounds__(kBlock)
k_gather_sorted(int nact, const int32_t *__restrict__ slots, const uint64_t *__restrict__ key1,
                uint64_t *__restrict__ key1s, const double *__restrict__ X,
                const double *__restrict__ Y, const double *__restrict__ Z,
                const double *__restrict__ GX, const double *__restrict__ GY,
                const double *__restrict__ GZ, double *__restrict__ xs, double *__restrict__ ys,
                double *__restrict__ zs, double *__restrict__ gxs, double *__restrict__ gys,
                double *__restrict__ gzs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nact) return;
  const int i = slots[s];
  key1s[s] = key1[i];
  xs[s] = X[i]; ys[s] = Y[i]; zs[s] = Z[i];
  gxs[s] = GX[i]; gys[s] = GY[i]; gzs[s] = GZ[i];
}

// nodes whose first member is sorted position s
__global__ void __launch_bounds__(kBlock)
k_count_nodes(int nact, int pcount, const uint64_t *__restrict__ k0s,
              const uint64_t *__restrict__ k1s, int32_t *__restrict__ nn,
              int32_t *__restrict__ flags) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nact) return;
  const uint64_t a0 = k0s[s], a1 = k1s[s];
  const int lp = s > 0 ? lcp(k0s[s - 1], k1s[s - 1], a0, a1) : 0;
  const int ln = s + 1 < nact ? lcp(a0, a1, k0s[s + 1], k1s[s + 1]) : 0;
  const int f = drop_level(a0, a1);
  if (ln == kMaxLevels && f > kMaxLevels) atomicOr(flags, kFlagDepth);
  // the cell at level lp+1 always exists (its parent holds s-1 and s);
  // deeper ones need the next point as a second member
  const int A = lp + 1;
  int B = min(f - 1, max(A, ln + 1));
  B = min(B, kMaxLevels);
  int c = B - A + 1;
  if (pcount < 2 || c < 0) c = 0;  // a root holding <= 1 slot is a leaf
  nn[s] = c;
}

struct NodeArgs {
  int nact, pcount;
  int64_t nnodes;
  double L;
  const uint64_t *k0s, *k1s;
  const int32_t *slots, *nn, *off;
  const int32_t *inf, *beh;
  int8_t *level, *digit;
  int32_t *first, *count;
  double *corner;
  int2 *meta;
  double2 *geo;
  int32_t *linf;
};

// end of the level-`lev` cell whose first member is s: first t > s with
// lcp(s, t) < lev (lcp(s, t) is non-increasing in t).  Most cells are a few
// points long, so the search gallops from s (s+1, s+2, s+4, ...) before it
// bisects: a few loads next to s instead of a bisection of all positions
__device__ inline int cell_end(const NodeArgs &A, int s, int lev) {
  const uint64_t a0 = A.k0s[s], a1 = A.k1s[s];
  int lo = s + 1, hi = A.nact;  // answer in [lo, hi]
  for (int64_t step = 1;; step <<= 1) {
    const int64_t t = s + step;
    if (t >= A.nact) break;
    if (lcp(a0, a1, A.k0s[t], A.k1s[t]) < lev) {
      hi = (int)t;
      break;
    }
    lo = (int)t + 1;
  }
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (lcp(a0, a1, A.k0s[mid], A.k1s[mid]) >= lev) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(kBlock) k_make_nodes(NodeArgs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) {
    // root (construct_tree, tree.f90:49-61): all slots, corner -L/2, width L
    A.level[0] = 0;
    A.digit[0] = 0;
    A.first[0] = 0;
    A.count[0] = A.pcount;
    A.corner[0] = A.corner[1] = A.corner[2] = -A.L / 2.;
    A.geo[0].x = A.L;
    int j = -1;
    if (A.pcount == 1 && A.inf[0] != 0) j = A.beh[A.inf[0] - 1] - 1;
    A.meta[0] = make_int2((int)A.nnodes, j);
    A.linf[0] = (A.pcount == 1 && A.inf[0] != 0) ? A.inf[0] - 1 : -1;
  }
  if (s >= A.nact) return;
  const int c = A.nn[s];
  if (c == 0) return;
  const uint64_t a0 = A.k0s[s], a1 = A.k1s[s];
  const int lp = s > 0 ? lcp(A.k0s[s - 1], A.k1s[s - 1], a0, a1) : 0;
  const int Alev = lp + 1;
  // replay the corner recurrence down to level Alev-1
  double P[3] = {-A.L / 2., -A.L / 2., -A.L / 2.};
  double w = A.L;
  for (int lev = 1; lev < Alev; ++lev) {
    const int d = digit_at(a0, a1, lev);
    const int b[3] = {d & 1, (d >> 2) & 1, (d >> 1) & 1};
    for (int a = 0; a < 3; ++a)
      if (b[a]) P[a] = P[a] + (w / 2);
    w = w / 2.;
  }
  const int p = A.slots[s];
  for (int q = 0; q < c; ++q) {
    const int lev = Alev + q;
    const int d = digit_at(a0, a1, lev);
    const int b[3] = {d & 1, (d >> 2) & 1, (d >> 1) & 1};
    for (int a = 0; a < 3; ++a)
      if (b[a]) P[a] = P[a] + (w / 2);
    w = w / 2.;
    const int64_t node = A.off[s] + q;
    const int e = cell_end(A, s, lev);
    A.level[node] = (int8_t)lev;
    A.digit[node] = (int8_t)d;
    A.first[node] = s;
    A.count[node] = e - s;
    A.corner[3 * node + 0] = P[0];
    A.corner[3 * node + 1] = P[1];
    A.corner[3 * node + 2] = P[2];
    A.geo[node].x = w;
    int j = -1;
    if (e - s == 1) j = A.beh[A.inf[p] - 1] - 1;  // f(parray(1)%infront)%behind
    const int skip = e < A.nact ? A.off[e] : (int)A.nnodes;
    A.meta[node] = make_int2(skip, j);
    A.linf[node] = e - s == 1 ? A.inf[p] - 1 : -1;
  }
}

struct MomArgs {
  int64_t nnodes;
  int nact;
  double theta;
  int extra;
  const double *xs, *ys, *zs, *gxs, *gys, *gzs;
  const int32_t *first, *count;
  const double *corner;
  const int2 *meta;
  double2 *geo;
  NodeRec *rec;
  double *mp;
  float *mp32;
  int32_t *big_list;   // big node ids (count > kSmallNode), in discovery order
  int32_t *big_base;   // first chunk of each big node
  int32_t *big_count;  // [0] big nodes, [1] chunks
  int32_t *chunk_big;  // big-list index of each chunk
  double *partial;     // 15 sums per chunk
  const int32_t *slots;  // slot of each sorted position
};

constexpr int kChunk = 4096;  // members per workgroup in the big-node sums

// tree_walk's acceptance test (tree.f90:351-365, 394) for a squared target
// distance d2, evaluated in the reference's operation order
__device__ inline bool accept_exact(double d2, double w, double g, double theta, int extra) {
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

__device__ inline uint64_t dbits(double v) { return (uint64_t)__double_as_longlong(v); }
__device__ inline double bitsd(uint64_t b) { return __longlong_as_double((long long)b); }

// smallest d2 >= 0 with accept_exact(d2): acceptance is monotone in d2 when
// the cell-centre branch (d <= g) cannot accept, which holds for theta <=
// 2/sqrt(3) because g <= sqrt(3)/2 w for a centroid inside its cell.
// Returns -1 when the node needs the exact test at every visit, +inf when
// no distance is accepted.
__device__ double accept_threshold(double w, double g, double theta, int extra) {
  if (!(theta > 0.) || !(w > 0.)) return INFINITY;
  if (extra && !(theta <= 1.1547)) return -1.;
  if (extra && !(g <= 0.8660254 * w)) return -1.;
  const double t = (extra ? g : 0.) + w / theta;
  const double guess = t * t;
  if (!(guess < 1e300)) return -1.;
  // the threshold is within a few ulps of the real-number one: walk from
  // the guess one ulp at a time, then fall back to bisection on the bit
  // patterns (non-negative doubles order like their integer images)
  uint64_t lo, hi;
  if (accept_exact(guess, w, g, theta, extra)) {
    hi = dbits(guess);
    for (int k = 0; k < 16; ++k) {
      if (hi == 0 || !accept_exact(bitsd(hi - 1), w, g, theta, extra)) return bitsd(hi);
      --hi;
    }
    lo = 0;
  } else {
    lo = dbits(guess);
    for (int k = 0; k < 16; ++k) {
      if (accept_exact(bitsd(lo + 1), w, g, theta, extra)) return bitsd(lo + 1);
      ++lo;
    }
    hi = dbits(4. * guess + 1e-300);
    if (!accept_exact(bitsd(hi), w, g, theta, extra)) return -1.;
  }
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (accept_exact(bitsd(mid), w, g, theta, extra)) hi = mid;
    else lo = mid;
  }
  return bitsd(hi);
}

// the probe walk's plain test (tree_walk_general, tree.f90:437-444) as a
// threshold on d2; the root takes the exact test
__global__ void __launch_bounds__(kBlock) k_plain_thresholds(int64_t nnodes,
                                                             const double2 *__restrict__ geo,
                                                             const int32_t *__restrict__ count,
                                                             double theta,
                                                             double *__restrict__ tplain) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nnodes) return;
  double T = 0.;  // a one-point node is never tested
  if (node == 0) T = -1.;
  else if (count[node] > 1) T = accept_threshold(geo[node].x, 0., theta, 0);
  tplain[node] = T;
}

__device__ inline void finalize_node(const MomArgs &A, int64_t node, int n, double s[15],
                                     bool is_root) {
  const double cx = s[0] / n, cy = s[1] / n, cz = s[2] / n;
  double circ[3] = {s[3], s[4], s[5]};
  double C[9];
  const double cent[3] = {cx, cy, cz};
  for (int m = 0; m < 3; ++m)
    for (int k = 0; k < 3; ++k) C[3 * m + k] = s[6 + 3 * m + k] - cent[m] * circ[k];
  if (is_root) {  // construct_tree zeroes the root's moments (tree.f90:57-58)
    circ[0] = circ[1] = circ[2] = 0.;
    for (int k = 0; k < 9; ++k) C[k] = 0.;
  }
  const double w = A.geo[node].x;
  const double* P = A.corner + 3 * node;
  const double geo = sqrt(((cx - (P[0] + w / 2.)) * (cx - (P[0] + w / 2.)) +
                           (cy - (P[1] + w / 2.)) * (cy - (P[1] + w / 2.))) +
                          (cz - (P[2] + w / 2.)) * (cz - (P[2] + w / 2.)));
  A.geo[node].y = geo;
  const int2 meta = A.meta[node];
  double T;
  if (meta.y >= 0) T = 0.;  // leaf: acceptance is never tested
  else if (is_root) T = -1.;
  else T = accept_threshold(w, geo, A.theta, A.extra);
  NodeRec r;
  r.cx = cx; r.cy = cy; r.cz = cz; r.T = T;
  r.skip = meta.x; r.j = meta.y;
  r.m6[0] = (float)circ[0]; r.m6[1] = (float)circ[1]; r.m6[2] = (float)circ[2];
  r.m6[3] = (float)(C[5] - C[7]);   // eps(:,k,l) C(k,l), the antisymmetric part
  r.m6[4] = (float)(C[6] - C[2]);
  r.m6[5] = (float)(C[1] - C[3]);
  A.rec[node] = r;
  double *mp = A.mp + 12 * node;
  float *m32 = A.mp32 + 16 * node;
  for (int k = 0; k < 3; ++k) mp[k] = circ[k];
  for (int k = 0; k < 9; ++k) mp[3 + k] = C[k];
  for (int k = 0; k < 9; ++k) m32[k] = (float)C[k];
  for (int k = 9; k < 16; ++k) m32[k] = 0.f;
}

// one member's terms in new_tree's operation order (tree.f90:154-165):
// cent + x, (circ + g) - x, C(:,k) + x (g_k - x_k)
__device__ inline void add_member(const double x[3], const double g[3], double s[15]) {
  s[0] = s[0] + x[0]; s[1] = s[1] + x[1]; s[2] = s[2] + x[2];
  s[3] = (s[3] + g[0]) - x[0]; s[4] = (s[4] + g[1]) - x[1]; s[5] = (s[5] + g[2]) - x[2];
  for (int k = 0; k < 3; ++k) {
    const double e = g[k] - x[k];
    for (int m = 0; m < 3; ++m) s[6 + 3 * m + k] = s[6 + 3 * m + k] + x[m] * e;
  }
}

__device__ inline void add_point(const MomArgs &A, int t, double s[15]) {
  const double x[3] = {A.xs[t], A.ys[t], A.zs[t]};
  const double g[3] = {A.gxs[t], A.gys[t], A.gzs[t]};
  add_member(x, g, s);
}

// exclusive prefix of v over the wavefront's lanes, and one atomic add of
// the wavefront's total to *ctr (every lane must call it): the base of the
// lane's entries in the list that *ctr counts
__device__ inline int wave_reserve(int v, int32_t *ctr) {
  const int lane = threadIdx.x % kWave;
  int incl = v;
  for (int off = 1; off < kWave; off <<= 1) {
    const int u = __shfl_up(incl, off);
    if (lane >= off) incl += u;
  }
  const int total = __shfl(incl, kWave - 1);
  int base = 0;
  if (lane == kWave - 1 && total > 0) base = atomicAdd(ctr, total);
  base = __shfl(base, kWave - 1);
  return base + incl - v;
}

// compare-exchange of (slot, position) pairs by slot
__device__ inline void cswap(int &ka, int &va, int &kb, int &vb) {
  if (kb < ka) {
    const int k = ka, v = va;
    ka = kb; va = vb; kb = k; vb = v;
  }
}

// one thread per node.  Big nodes (more than kSmallNode members) are queued
// for k_moments_chunks / k_moments_big (one atomic per wavefront).  Nodes of
// at most kTinyNode members are summed in slot order, as new_tree meets
// them: their member slots are sorted in registers by a sorting network.
// Nodes of kTinyNode+1..kSmallNode members are summed in key order.  (One
// wavefront per node of 9..256 members, lane-strided sums and a butterfly of
// the 15 FP64 sums, measured slower: 0.35 ms at 1M points for those nodes
// alone, against 0.31 ms for this whole kernel.)
__global__ void __launch_bounds__(kBlock) k_moments_small(MomArgs A) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = node < A.nnodes;
  const bool is_root = node == 0;
  const int n = !in ? 0 : (is_root ? A.nact : A.count[node]);
  const bool big = in && n > kSmallNode;
  const int nch = big ? (n + kChunk - 1) / kChunk : 0;
  const int q = wave_reserve(big ? 1 : 0, A.big_count);
  const int base = wave_reserve(nch, A.big_count + 1);
  if (!in) return;
  if (big) {
    A.big_list[q] = (int)node;
    A.big_base[q] = base;
    for (int k = 0; k < nch; ++k) A.chunk_big[base + k] = q;
    return;
  }
  const int f = A.first[node];
  double s[15];
  if (n == 1 && !is_root) {
    // leaf: centroid x/1 == x, the record keeps x and ghosti exactly
    s[3] = A.gxs[f]; s[4] = A.gys[f]; s[5] = A.gzs[f];
    const double cent[3] = {A.xs[f], A.ys[f], A.zs[f]};
    const int2 meta = A.meta[node];
    const double w = A.geo[node].x;
    const double *P = A.corner + 3 * node;
    const double geo = sqrt(((cent[0] - (P[0] + w / 2.)) * (cent[0] - (P[0] + w / 2.)) +
                             (cent[1] - (P[1] + w / 2.)) * (cent[1] - (P[1] + w / 2.))) +
                            (cent[2] - (P[2] + w / 2.)) * (cent[2] - (P[2] + w / 2.)));
    A.geo[node].y = geo;
    NodeRec r;
    r.cx = cent[0]; r.cy = cent[1]; r.cz = cent[2]; r.T = 0.;
    r.skip = meta.x; r.j = meta.y;
    r.g[0] = s[3]; r.g[1] = s[4]; r.g[2] = s[5];
    A.rec[node] = r;
    double *mp = A.mp + 12 * node;
    mp[0] = s[3]; mp[1] = s[4]; mp[2] = s[5];
    for (int k = 3; k < 12; ++k) mp[k] = 0.;
    return;
  }
  if (n == 0) {
    NodeRec r;
    r.cx = r.cy = r.cz = 0.;
    r.T = INFINITY;
    r.skip = A.meta[node].x;
    r.j = A.meta[node].y;
    r.g[0] = r.g[1] = r.g[2] = 0.;
    A.rec[node] = r;
    return;
  }
  for (int k = 0; k < 15; ++k) s[k] = 0.;
  if (n <= kTinyNode) {
    // the members' slots sorted with their positions by Batcher's 19-
    // comparator network for 8 inputs (positions past n sort last)
    int k0 = A.slots[f], v0 = 0;
    int k1 = 1 < n ? A.slots[f + 1] : INT32_MAX, v1 = 1;
    int k2 = 2 < n ? A.slots[f + 2] : INT32_MAX, v2 = 2;
    int k3 = 3 < n ? A.slots[f + 3] : INT32_MAX, v3 = 3;
    int k4 = 4 < n ? A.slots[f + 4] : INT32_MAX, v4 = 4;
    int k5 = 5 < n ? A.slots[f + 5] : INT32_MAX, v5 = 5;
    int k6 = 6 < n ? A.slots[f + 6] : INT32_MAX, v6 = 6;
    int k7 = 7 < n ? A.slots[f + 7] : INT32_MAX, v7 = 7;
    cswap(k0, v0, k1, v1); cswap(k2, v2, k3, v3); cswap(k4, v4, k5, v5); cswap(k6, v6, k7, v7);
    cswap(k0, v0, k2, v2); cswap(k1, v1, k3, v3); cswap(k4, v4, k6, v6); cswap(k5, v5, k7, v7);
    cswap(k1, v1, k2, v2); cswap(k5, v5, k6, v6);
    cswap(k0, v0, k4, v4); cswap(k1, v1, k5, v5); cswap(k2, v2, k6, v6); cswap(k3, v3, k7, v7);
    cswap(k2, v2, k4, v4); cswap(k3, v3, k5, v5);
    cswap(k1, v1, k2, v2); cswap(k3, v3, k4, v4); cswap(k5, v5, k6, v6);
    const int order[8] = {v0, v1, v2, v3, v4, v5, v6, v7};
#pragma unroll
    for (int r = 0; r < kTinyNode; ++r)
      if (r < n) add_point(A, f + order[r], s);
  } else {
#pragma unroll 4
    for (int t = f; t < f + n; ++t) add_point(A, t, s);
  }
  finalize_node(A, node, n, s, is_root);
}

// one workgroup per chunk of a big node's member range; partial sums per
// chunk, combined in chunk order by k_moments_big (deterministic)
__global__ void __launch_bounds__(kBlock) k_moments_chunks(MomArgs A) {
  __shared__ double red[15][kBlock / kWave];
  const int nchunks = A.big_count[1];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int q = A.chunk_big[c];
    const int64_t node = A.big_list[q];
    const int n = node == 0 ? A.nact : A.count[node];
    const int f = A.first[node];
    const int lo = f + (c - A.big_base[q]) * kChunk;
    const int hi = min(f + n, lo + kChunk);
    double s[15];
    for (int k = 0; k < 15; ++k) s[k] = 0.;
    for (int t = lo + threadIdx.x; t < hi; t += blockDim.x) add_point(A, t, s);
    for (int k = 0; k < 15; ++k) {
      double v = s[k];
      for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red[k][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < 15) {
      double v = 0.;
      for (int w = 0; w < kBlock / kWave; ++w) v += red[threadIdx.x][w];
      A.partial[15 * (int64_t)c + threadIdx.x] = v;
    }
    __syncthreads();
  }
}

// one wavefront per big node: lane l sums chunk partials l, l+64, ... in
// chunk order, a butterfly combines the lanes (deterministic), lane 0 writes
// the node record
__global__ void __launch_bounds__(kBlock) k_moments_big(MomArgs A) {
  const int nbig = A.big_count[0];
  const int lane = threadIdx.x % kWave;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int nwaves = gridDim.x * blockDim.x / kWave;
  for (int q = wave; q < nbig; q += nwaves) {
    const int64_t node = A.big_list[q];
    const bool is_root = node == 0;
    const int n = is_root ? A.nact : A.count[node];
    const int nch = (n + kChunk - 1) / kChunk;
    const int base = A.big_base[q];
    double s[15];
    for (int k = 0; k < 15; ++k) s[k] = 0.;
    for (int c = base + lane; c < base + nch; c += kWave)
      for (int k = 0; k < 15; ++k) s[k] += A.partial[15 * (int64_t)c + k];
    for (int k = 0; k < 15; ++k)
      for (int off = kWave / 2; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off);
    if (lane == 0) finalize_node(A, node, n, s, is_root);
  }
}

// 21-bit Hilbert-curve index of each position on a 128^3 grid of the box
// (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 381
// (2004): axes to transposed index, then bit interleave); slots with
// infront = 0 (when inf is given) get 2^21 and sort last.  The index itself
// is the sort's value.  Only the grouping of targets into wavefronts uses
// this order (finer grids group the 1M-point workload no better; ties keep
// the input order): each target's own walk and results do not depend on it.
__global__ void __launch_bounds__(kBlock) k_hilbert_keys(int n, double L,
                                                         const double *__restrict__ xs,
                                                         const double *__restrict__ ys,
                                                         const double *__restrict__ zs,
                                                         int stride,
                                                         const int32_t *__restrict__ inf,
                                                         uint32_t *__restrict__ key,
                                                         int32_t *__restrict__ idx) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  idx[s] = s;
  if (inf && inf[s] == 0) {
    key[s] = 1u << 21;
    return;
  }
  constexpr int kBits = 7;
  const double scale = (double)(1 << kBits) / L;
  const int64_t o = (int64_t)s * stride;
  const double p[3] = {xs[o], ys[o], zs[o]};
  uint32_t X[3];
  for (int k = 0; k < 3; ++k) {
    const double q = (p[k] + L / 2.) * scale;
    X[k] = q <= 0. ? 0u : (q >= (double)((1 << kBits) - 1) ? (uint32_t)((1 << kBits) - 1)
                                                              : (uint32_t)q);
  }
  for (uint32_t Q = 1u << (kBits - 1); Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0;
  for (uint32_t Q = 1u << (kBits - 1); Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  X[0] ^= t; X[1] ^= t; X[2] ^= t;
  uint32_t h = 0;
  for (int b = kBits - 1; b >= 0; --b)
    for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
  key[s] = h;
}

hipError_t grow_tree(Tree &t, int64_t need) {
  if (need <= t.cap) return hipSuccess;
  int64_t cap = need + need / 4 + 1024;
  (void)hipFree(t.rec); (void)hipFree(t.meta); (void)hipFree(t.mp); (void)hipFree(t.mp32); (void)hipFree(t.geo);
  (void)hipFree(t.first); (void)hipFree(t.count); (void)hipFree(t.level); (void)hipFree(t.digit); (void)hipFree(t.corner);
  (void)hipFree(t.linf); (void)hipFree(t.tplain); (void)hipFree(t.tclose);
  t = Tree{};
  hipError_t e = hipSuccess;
  if ((e = hipMalloc(&t.rec, cap * sizeof(NodeRec)))) return e;
  if ((e = hipMalloc(&t.meta, cap * sizeof(int2)))) return e;
  if ((e = hipMalloc(&t.mp, cap * 12 * sizeof(double)))) return e;
  if ((e = hipMalloc(&t.mp32, cap * 16 * sizeof(float)))) return e;
  if ((e = hipMalloc(&t.geo, cap * sizeof(double2)))) return e;
  if ((e = hipMalloc(&t.first, cap * sizeof(int32_t)))) return e;
  if ((e = hipMalloc(&t.count, cap * sizeof(int32_t)))) return e;
  if ((e = hipMalloc(&t.level, cap * sizeof(int8_t)))) return e;
  if ((e = hipMalloc(&t.digit, cap * sizeof(int8_t)))) return e;
  if ((e = hipMalloc(&t.corner, cap * 3 * sizeof(double)))) return e;
  if ((e = hipMalloc(&t.linf, cap * sizeof(int32_t)))) return e;
  if ((e = hipMalloc(&t.tplain, cap * sizeof(double)))) return e;
  if ((e = hipMalloc(&t.tclose, cap * sizeof(double)))) return e;
  t.cap = cap;
  return hipSuccess;
}

}  // namespace

#define QV_TRY(expr)                       \
  do {                                     \
    hipError_t e_ = (expr);                \
    if (e_ != hipSuccess) return QV_ERR_HIP; \
  } while (0)

hipError_t launch_hilbert_keys(hipStream_t st, int n, double L, const double *x, const double *y,
                               const double *z, int stride, const int32_t *inf, uint32_t *key,
                               int32_t *idx) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_hilbert_keys, dim3(grid_of(n)), dim3(kBlock), 0, st, n, L, x, y, z, stride,
                     inf, key, idx);
  return hipGetLastError();
}

hipError_t launch_plain_thresholds(Ctx &c) {
  const int64_t nn = c.tree.nnodes;
  if (nn <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_plain_thresholds, dim3((unsigned)((nn + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, c.stream, nn, c.tree.geo, c.tree.count, c.p.tree_theta,
                     c.tree.tplain);
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) c.tree.plain_valid = true;
  return e;
}

hipError_t launch_close_thresholds(Ctx &c) {
  const int64_t nn = c.tree.nnodes;
  if (nn <= 0) return hipSuccess;
  // closest_tree takes a many-point node when width/dist < 0.5
  // (tree.f90:247-249): the plain test with theta = 0.5
  hipLaunchKernelGGL(k_plain_thresholds, dim3((unsigned)((nn + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, c.stream, nn, c.tree.geo, c.tree.count, 0.5,
                     c.tree.tclose);
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) c.tree.close_valid = true;
  return e;
}

int hilbert_order(Ctx &c, int n, Vec3 pos, const int32_t *inf, int32_t *order, hipStream_t st,
                  void *&tmp, size_t &tmp_bytes) {
  if (n <= 0) return QV_OK;
  // the key and slot buffers of the build are free outside tree_build's
  // steps 1-4: keys in key0 (2 cap 32-bit words), indices in slot
  uint32_t *hk = reinterpret_cast<uint32_t *>(c.key0);
  uint32_t *hk_s = hk + c.cap;
  QV_TRY(launch_hilbert_keys(st, n, c.p.box_size, pos.x, pos.y, pos.z, 1, inf, hk, c.slot));
  const unsigned end_bit = inf ? 22 : 21;
  size_t need = 0;
  QV_TRY(rocprim::radix_sort_pairs<SortConfig>(nullptr, need, hk, hk_s, c.slot, order, n, 0,
                                               end_bit, st));
  if (need > tmp_bytes) {
    QV_TRY(hipStreamSynchronize(st));
    (void)hipFree(tmp);
    tmp = nullptr;
    tmp_bytes = 0;
    QV_TRY(hipMalloc(&tmp, need));
    tmp_bytes = need;
  }
  size_t have = tmp_bytes;
  QV_TRY(rocprim::radix_sort_pairs<SortConfig>(tmp, have, hk, hk_s, c.slot, order, n, 0,
                                               end_bit, st));
  return QV_OK;
}

int join_aux(Ctx &c) {
  if (c.aux_live) QV_TRY(hipStreamWaitEvent(c.stream, c.aux_ev[1], 0));
  return QV_OK;
}

int tree_build(Ctx &c) {
  const int n = c.pcount;
  const int nact = c.nact;
  hipStream_t st = c.stream;
  // the previous build's sort on the second stream used key0 and slot
  if (int rc = join_aux(c)) return rc;
  // 1. keys
  hipLaunchKernelGGL(k_keys, dim3(grid_of(n)), dim3(kBlock), 0, st, n, c.p.box_size, c.x.x,
                     c.x.y, c.x.z, c.infront, c.key0, c.key1, c.slot);
  QV_TRY(hipGetLastError());
  // 2. sort slots by key0 (stable: ties keep slot order)
  size_t need = 0;
  QV_TRY(rocprim::radix_sort_pairs(nullptr, need, c.key0, c.key0_s, c.slot, c.slot_s, n, 0, 64,
                                   st));
  if (need > c.sort_tmp_bytes) {
    (void)hipFree(c.sort_tmp);
    c.sort_tmp = nullptr;
    c.sort_tmp_bytes = 0;
    QV_TRY(hipMalloc(&c.sort_tmp, need));
    c.sort_tmp_bytes = need;
  }
  size_t have = c.sort_tmp_bytes;
  QV_TRY(rocprim::radix_sort_pairs(c.sort_tmp, have, c.key0, c.key0_s, c.slot, c.slot_s, n, 0,
                                   64, st));
  // 3. order the rare runs that share all 16 levels of key0 by key1
  hipLaunchKernelGGL(k_deep_runs, dim3(grid_of(nact)), dim3(kBlock), 0, st, nact, c.key0_s,
                     c.key1, c.slot_s);
  QV_TRY(hipGetLastError());
  // 4. sorted copies of key1, positions and ghosts
  uint64_t *key1_s = c.key1 + n;  // second half of the key1 buffer
  hipLaunchKernelGGL(k_gather_sorted, dim3(grid_of(nact)), dim3(kBlock), 0, st, nact, c.slot_s,
                     c.key1, key1_s, c.x.x, c.x.y, c.x.z, c.gi.x, c.gi.y, c.gi.z, c.xs.x, c.xs.y,
                     c.xs.z, c.gs.x, c.gs.y, c.gs.z);
  QV_TRY(hipGetLastError());
  // 4b. the walk's order of the sorted targets (Hilbert curve), on the second
  // stream: it needs the sorted positions and key0/slot, which the steps
  // below leave alone, and overlaps them; the stream joins before the walk
  if (nact > 0) {
    QV_TRY(hipEventRecord(c.aux_ev[0], st));
    QV_TRY(hipStreamWaitEvent(c.aux_stream, c.aux_ev[0], 0));
    if (int rc = hilbert_order(c, nact, c.xs, nullptr, c.walk_order, c.aux_stream, c.aux_tmp,
                               c.aux_tmp_bytes))
      return rc;
    QV_TRY(hipEventRecord(c.aux_ev[1], c.aux_stream));
    c.aux_live = true;
  }
  // 5. nodes per sorted position, prefix sum -> DFS preorder offsets
  QV_TRY(hipMemsetAsync(c.nnode_of + nact, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_count_nodes, dim3(grid_of(nact)), dim3(kBlock), 0, st, nact, n, c.key0_s,
                     key1_s, c.nnode_of, c.counters + 1);
  QV_TRY(hipGetLastError());
  size_t scan_need = 0;
  QV_TRY(rocprim::exclusive_scan(nullptr, scan_need, c.nnode_of, c.node_off, (int32_t)1, nact + 1,
                                 rocprim::plus<int32_t>(), st));
  if (scan_need > c.sort_tmp_bytes) {
    (void)hipFree(c.sort_tmp);
    c.sort_tmp = nullptr;
    c.sort_tmp_bytes = 0;
    QV_TRY(hipMalloc(&c.sort_tmp, scan_need));
    c.sort_tmp_bytes = scan_need;
  }
  have = c.sort_tmp_bytes;
  // nnode_of[nact] is kept at 0 so that node_off[nact] is the node total
  QV_TRY(rocprim::exclusive_scan(c.sort_tmp, have, c.nnode_of, c.node_off, (int32_t)1, nact + 1,
                                 rocprim::plus<int32_t>(), st));
  QV_TRY(hipMemcpyAsync(c.h_counters, c.node_off + nact, sizeof(int32_t), hipMemcpyDeviceToHost,
                        st));
  QV_TRY(hipMemcpyAsync(c.h_counters + 1, c.counters + 1, sizeof(int32_t),
                        hipMemcpyDeviceToHost, st));
  QV_TRY(hipStreamSynchronize(st));
  if (c.h_counters[1] & kFlagDepth) return QV_ERR_TREE_DEPTH;
  const int64_t nnodes = c.h_counters[0];
  // node indices are 32-bit throughout (the count itself is an int32 sum)
  if (nnodes < 1 || nnodes >= (int64_t)INT32_MAX) return QV_ERR_ARG;
  QV_TRY(grow_tree(c.tree, nnodes));
  c.tree.nnodes = nnodes;
  c.tree.plain_valid = false;
  c.tree.close_valid = false;
  // 6. node records
  NodeArgs NA;
  NA.nact = nact;
  NA.pcount = n;
  NA.nnodes = nnodes;
  NA.L = c.p.box_size;
  NA.k0s = c.key0_s;
  NA.k1s = key1_s;
  NA.slots = c.slot_s;
  NA.nn = c.nnode_of;
  NA.off = c.node_off;
  NA.inf = c.infront;
  NA.beh = c.behind;
  NA.level = c.tree.level;
  NA.digit = c.tree.digit;
  NA.first = c.tree.first;
  NA.count = c.tree.count;
  NA.corner = c.tree.corner;
  NA.meta = c.tree.meta;
  NA.geo = c.tree.geo;
  NA.linf = c.tree.linf;
  hipLaunchKernelGGL(k_make_nodes, dim3(grid_of(nact > 0 ? nact : 1)), dim3(kBlock), 0, st, NA);
  QV_TRY(hipGetLastError());
  // 7. moments, geo distance, acceptance thresholds
  MomArgs MA;
  MA.nnodes = nnodes;
  MA.nact = nact;
  MA.theta = c.p.tree_theta;
  MA.extra = c.p.tree_extra_correction;
  MA.xs = c.xs.x; MA.ys = c.xs.y; MA.zs = c.xs.z;
  MA.gxs = c.gs.x; MA.gys = c.gs.y; MA.gzs = c.gs.z;
  MA.first = c.tree.first;
  MA.count = c.tree.count;
  MA.corner = c.tree.corner;
  MA.meta = c.tree.meta;
  MA.geo = c.tree.geo;
  MA.rec = c.tree.rec;
  MA.mp = c.tree.mp;
  MA.mp32 = c.tree.mp32;
  // nodes of one level are disjoint and there are at most 33 levels: at
  // most 33 * nact / (kSmallNode + 1) big nodes, and at most that many plus
  // 33 * nact / kChunk chunks
  const int64_t big_cap = 33 * (int64_t)nact / (kSmallNode + 1) + 1;
  const int64_t chunk_cap = big_cap + 33 * (int64_t)nact / kChunk + 1;
  const int64_t list_ints = big_cap * 2 + chunk_cap;
  if (list_ints > c.big_cap_ints || chunk_cap > c.partial_cap) {
    (void)hipFree(c.big_buf);
    (void)hipFree(c.partial);
    c.big_buf = nullptr;
    c.partial = nullptr;
    c.big_cap_ints = c.partial_cap = 0;
    QV_TRY(hipMalloc(&c.big_buf, list_ints * sizeof(int32_t)));
    QV_TRY(hipMalloc(&c.partial, chunk_cap * 15 * sizeof(double)));
    c.big_cap_ints = list_ints;
    c.partial_cap = chunk_cap;
  }
  MA.big_list = c.big_buf;
  MA.big_base = c.big_buf + big_cap;
  MA.chunk_big = c.big_buf + 2 * big_cap;
  MA.partial = c.partial;
  MA.big_count = c.counters + 2;
  MA.slots = c.slot_s;
  hipLaunchKernelGGL(k_moments_small, dim3(grid_of(nnodes)), dim3(kBlock), 0, st, MA);
  QV_TRY(hipGetLastError());
  // nodes of more than kSmallNode members: partials per chunk of up to kChunk
  // members, then one wavefront per node (grid-stride) combines them
  hipLaunchKernelGGL(k_moments_chunks, dim3(2048), dim3(kBlock), 0, st, MA);
  QV_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_moments_big, dim3(1024), dim3(kBlock), 0, st, MA);
  QV_TRY(hipGetLastError());
  // the walk reads the order: the second stream joins here
  if (int rc = join_aux(c)) return rc;
  return QV_OK;
}

}  // namespace qv
