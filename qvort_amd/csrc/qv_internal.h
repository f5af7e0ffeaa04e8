// qv_internal.h -- device data layout and kernel launchers shared by the
// translation units of libqvort_gpu.so (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdlib>
#include <thread>
#include <vector>

#include "../../include/qvort_gpu.h"

namespace qv {

constexpr double kPi = 3.14159265358979324;  // cdata.f90:176
constexpr int kWave = 64;
constexpr int kMaxLevels = 32;      // two 64-bit key words, 16 levels each
constexpr int kStackDepth = 40;     // per-wave DFS stack (levels + 1 <= 33)

// structure-of-arrays vectors of the filament state, one array per axis
struct Vec3 {
  double *x = nullptr, *y = nullptr, *z = nullptr;
};

// shells of the periodic images in the loop order of calc_velocity
// (timestep.f90:283-286): (0,0,0) first, then p outer .. r inner
struct Shells {
  int n = 1;
  double s[27][3] = {};
};

// one node as the walk reads it: a single 64-byte scalar load per visit
//   cx, cy, cz   centroid (leaf: the point's x, since x/1 == x)
//   T            acceptance threshold on the squared target distance
//                (< 0: evaluate the full test; leaves: unused)
//   skip         index of the node after this subtree (DFS preorder)
//   j            leaf: 0-based slot j = f(parray(1)%infront)%behind, whose
//                exclusion the walk tests; internal node: -1
//   leaf:        g[3] = ghosti of the point (FP64)
//   internal:    m6[0..2] = circ, m6[3..5] = eps:C (FP32 copies)
struct alignas(64) NodeRec {
  double cx, cy, cz, T;
  int32_t skip, j;
  union {
    double g[3];
    float m6[6];
  };
};
static_assert(sizeof(NodeRec) == 64, "node record is one 64-byte line");

// tree nodes in DFS preorder, empty cells omitted (they contribute nothing
// and are skipped by tree_walk, tree.f90:349).
//   rec    the walk's record (above)
//   meta   skip, j (as in rec; used while building)
//   mp     internal: circ(3), C(9) row-major C[i][k]; leaf: ghosti(3)
//   mp32   internal: C(9) in FP32 (FP32C path)
//   geo    width, geo_dist
struct Tree {
  int64_t nnodes = 0;
  int64_t cap = 0;
  NodeRec *rec = nullptr;
  int2 *meta = nullptr;
  double *mp = nullptr;     // 12 per node
  float *mp32 = nullptr;    // 16 per node
  double2 *geo = nullptr;   // width, geo_dist
  // smallest d2 accepted by the probe walk's plain test width/dist < theta
  // (tree_walk_general), per node; computed on the first probe call after a
  // build (plain_valid); -1: exact test at every visit
  double *tplain = nullptr;
  bool plain_valid = false;
  // the same for the neighbour search's test width/dist < 0.5 (closest_tree);
  // computed on the first search after a build (close_valid)
  double *tclose = nullptr;
  bool close_valid = false;
  // build-only arrays (per node)
  int32_t *first = nullptr;  // first sorted position
  int32_t *count = nullptr;  // member count (root: pcount slots)
  int8_t *level = nullptr;
  int8_t *digit = nullptr;   // child index within the parent (0..7)
  double *corner = nullptr;  // 3 per node
  // neighbour search: leaf -> 0-based slot infront(p) - 1 of its point p at
  // build time (the search resolves j = behind(that slot) with the links of
  // its own call); other nodes -1
  int32_t *linf = nullptr;
};

struct Ctx {
  qv_params p{};
  int device = 0;
  hipStream_t stream = nullptr;      // stream of every launch (own or the caller's)
  hipStream_t own_stream = nullptr;
  int32_t pcount = 0, cap = 0;
  int32_t nact = 0;
  bool have_velocity = false;
  // the tree of the last velocity stage (built whenever tree_theta > 0, as
  // construct_tree is, run.f90:43-45) stays usable by qv_closest after later
  // uploads; tree_pcount = slots at build time
  bool tree_valid = false;
  int32_t tree_pcount = 0;
  int32_t *close_i = nullptr;   // closest (1-based), per slot
  int32_t *close_order = nullptr;  // the search's slot order (Hilbert curve)
  double *close_d = nullptr;    // closestd, per slot
  int32_t close_cap = 0;
  // ghosts of the last ghostp (velocity stage or a probe call with
  // QV_PROBE_GHOSTP); an upload invalidates them
  bool ghosts_valid = false;
  double *probe_buf = nullptr;  // probes (3n), u (3n), gradient (9n), sort keys (3n)
  int32_t probe_cap = 0;
  double *probe_part = nullptr;  // Tree probes: partial sums per shell chunk
  size_t probe_part_cap = 0;     // doubles
  // state (SoA), slot order
  Vec3 x, gi, gb, gii, gbb, u, uli, ubs, u1, u2, u3;
  int32_t *infront = nullptr, *behind = nullptr;
  // staging for host <-> device copies of one (pcount,3) array
  double *stage = nullptr;
  // tree build
  uint64_t *key0 = nullptr, *key1 = nullptr, *key0_s = nullptr;
  int32_t *slot = nullptr, *slot_s = nullptr;  // sorted active slots
  // the walk's target order: sorted positions in Hilbert-curve order, so
  // that the 64 targets of a wavefront are spatially compact
  int32_t *walk_order = nullptr;
  // the walk's raw sums of the second half of the image shells, 3 per slot
  double *walk_part = nullptr;
  int32_t *nnode_of = nullptr, *node_off = nullptr;  // per sorted position
  Vec3 xs, gs;  // positions and ghosts in sorted order
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int32_t *counters = nullptr;   // [1] error flags, [2] big nodes, [3] chunks, [6] open faces
  int32_t *big_buf = nullptr;    // big-node list, chunk bases, chunk owners
  double *partial = nullptr;     // per-chunk partial sums of the big nodes
  double *direct_part = nullptr;  // FP64 direct sum: 6 sums per slot per chunk
  int64_t direct_part_doubles = 0;
  int64_t big_cap_ints = 0, partial_cap = 0;
  int64_t *evals = nullptr;
  Tree tree;
  // pinned host scalars
  int32_t *h_counters = nullptr;
  int64_t *h_evals = nullptr;
  qv_timings timings{};
  // adjoint: 24 doubles of scratch per slot, tangent/cotangent and result
  double *adj = nullptr;
  double *adj_part = nullptr;  // partial sums of the VJP's chunks
  int64_t adj_part_doubles = 0;
  Vec3 aw, ar;
  int32_t adj_cap = 0;
  bool links_consistent = false;  // infront(behind(i)) == i for every point
  // boundary_loop_remove_count / _length of the boundary steps so far
  int64_t boundary_count = 0;
  double boundary_length = 0.;
  hipEvent_t ev[6] = {};
  // a second stream (default priority) for the walk order's sort, which
  // overlaps the rest of the tree build: aux_ev[0] = build data ready,
  // aux_ev[1] = order sorted; aux_live once aux_ev[1] has been recorded
  hipStream_t aux_stream = nullptr;
  hipEvent_t aux_ev[2] = {};
  bool aux_live = false;
  void *aux_tmp = nullptr;
  size_t aux_tmp_bytes = 0;
};

// error flags accumulated in counters[1]
enum : int32_t { kFlagSingular = 1, kFlagDepth = 2 };

Shells make_shells(const qv_params &p);

// per-target accumulator of FP32 terms: every term folded into an FP64 sum
// (on CDNA4 an FP64 add costs the same issue slot as an FP32 add, so this is
// the cheapest compensated sum; Kahan or two-sum in FP32 costs three times
// the instructions per term).
struct AccF64 {
  double s0 = 0., s1 = 0., s2 = 0.;
  __device__ inline void add(float a, float b, float c) {
    s0 += (double)a; s1 += (double)b; s2 += (double)c;
  }
  __device__ inline void get(double &a, double &b, double &c) const { a = s0; b = s1; c = s2; }
};

// the per-slot arrays clear_particle zeroes besides x and the links
struct ClearArgs {
  static constexpr int kArrays = 27;  // u, u1, u2, u_LI, u_BS, 4 ghosts; 3 axes each
  double *a[kArrays];
};

// the boundary step of openx / openxy boxes on host arrays (x(3,pcount),
// 1-based links), in the reference's slot order: enforce_periodic_yz / _z
// (periodic.f90:217-276) with boundary_loop_remove (periodic.f90:306-327).
// Cleared slots get x = 0 and links 0 and are appended to `cleared`;
// count / length accumulate boundary_loop_remove_count / _length.
void boundary_step_host(const qv_params &p, int32_t pcount, double *x, int32_t *inf,
                        int32_t *beh, int64_t &count, double &length, int32_t *cleared,
                        int32_t &ncleared);

// launchers (each returns hipError_t of the launch)
hipError_t launch_aos_to_soa(Ctx &c, const double *d_aos, Vec3 dst);
hipError_t launch_soa_to_aos(Ctx &c, Vec3 src, double *d_aos);
hipError_t launch_ghost_lia(Ctx &c);
hipError_t launch_ab_update(Ctx &c);
hipError_t launch_rk3_update(Ctx &c, int stage);
hipError_t launch_enforce_periodic(Ctx &c);
// flag |= 1 when an active point lies beyond an open face (openx, openxy)
hipError_t launch_open_check(Ctx &c, int32_t *flag);
// the per-slot rules of a topology upload: new points (pinsert) and cleared
// slots, from the device's old links and the uploaded new ones (device
// buffers); mark: n_new ints of scratch
// keep_x: the positions were not uploaded; cleared slots get x = 0 and a
// slot that becomes a point sets *flag = 1 (it has no position)
hipError_t launch_topology_apply(Ctx &c, int n_new, int n_old, const int32_t *new_inf,
                                 const int32_t *new_beh, int32_t *mark, bool keep_x,
                                 int32_t *flag);
// clear_particle's zeroes in the m slots of d_list (device)
hipError_t launch_clear_slots(Ctx &c, int m, const int32_t *d_list);
hipError_t launch_direct(Ctx &c);
// source chunks of the direct sum at pcount slots (6 pcount doubles of
// partial sums each)
int direct_chunks(int pcount, bool fp32);
// 21-bit Hilbert-curve keys of n points (x[i*stride], y[..], z[..]) on a
// 128^3 grid of the box, idx[i] = i; with inf, points with inf[i] = 0 get
// the key 2^21
hipError_t launch_hilbert_keys(hipStream_t st, int n, double L, const double *x, const double *y,
                               const double *z, int stride, const int32_t *inf, uint32_t *key,
                               int32_t *idx);
// the probe walk's per-node thresholds (Tree::tplain) for the current tree
hipError_t launch_plain_thresholds(Ctx &c);
// the neighbour search's per-node thresholds (Tree::tclose)
hipError_t launch_close_thresholds(Ctx &c);
// slots 0..n-1 (positions pos) in Hilbert-curve order into order[0..n);
// with inf, slots with infront = 0 go last.  Uses the build's key buffers
// and the sort scratch (tmp, tmp_bytes; grown as needed) on stream st.
int hilbert_order(Ctx &c, int n, Vec3 pos, const int32_t *inf, int32_t *order, hipStream_t st,
                  void *&tmp, size_t &tmp_bytes);
// make c.stream wait for the walk order's sort on the second stream (it
// reads and writes the build's key buffers)
int join_aux(Ctx &c);
int tree_build(Ctx &c);          // returns qv_status
hipError_t launch_walk(Ctx &c);
// scratch: 24 doubles per slot, every one written before it is read;
// part (may be NULL): room for the partial sums of source / target chunks,
// 9 doubles per slot per chunk (vjp_chunks); without room, one chunk
hipError_t launch_vjp(Ctx &c, Vec3 w, Vec3 r, int mixed, double *scratch, double *part,
                      int64_t part_doubles);
int vjp_chunks(int n, int &len);
hipError_t launch_jvp(Ctx &c, Vec3 v, Vec3 r);
hipError_t launch_closest(Ctx &c, bool tree);
hipError_t launch_probe(Ctx &c, int nprobe, const double *d_y, double *d_u, double *d_g,
                        double *d_work,
                        bool tree, bool ref_scale, bool empty_core);

// QV_HOST_THREADS=0: the line operations evaluate every slot's test in the
// slot loop, without the pre-pass
inline bool host_prepass() {
  const char *e = std::getenv("QV_HOST_THREADS");
  return !(e && std::atoi(e) == 0);
}

// host: f(lo, hi) over [0, n) on up to 16 threads, or QV_HOST_THREADS of
// them when that is set (one thread for small n)
template <class F>
void par_for(int n, F f) {
  const unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min(16u, hw > 0 ? hw : 1u);
  if (const char *e = std::getenv("QV_HOST_THREADS")) nt = std::max(1, std::min(64, std::atoi(e)));
  if (n < (1 << 16) || nt <= 1) {
    f(0, n);
    return;
  }
  const int chunk = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) {
    const int lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo < hi) th.emplace_back(f, lo, hi);
  }
  f(0, std::min(n, chunk));
  for (auto &x : th) x.join();
}

}  // namespace qv
