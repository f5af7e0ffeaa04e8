// qv_closest.hip -- neighbour search of the reconnection step.
//
// Tree mode: pclose_tree / closest_tree (tree.f90:219-278).  Each slot
// walks the octree of the last velocity stage (built, like construct_tree's,
// from the positions before the step's motion) with the slot's current
// position, in DFS preorder: a taken node is left for its skip pointer, an
// opened node continues at the next record (its first stored child; empty
// children are not stored and would return).  A node is taken when it
// holds one point or width/dist < 0.5; when dist < closestd, a many-point
// node sets closest = 0 (closestd unchanged) and a one-point node offers
// j = behind(infront_at_build(point)) with the current behind, taken unless
// j is i, infront(i) or behind(i).  closestd starts at 100 for every slot,
// empty or not (tree.f90:227).
//
// The walk is wavefront-cooperative, as the velocity walk is: the 64 slots
// of a wavefront (adjacent on a Hilbert curve through the box) share one DFS whose current node is read with scalar loads; each lane
// takes part only in the nodes of its own walk (its bit of the mask), so it
// meets them in the same order and updates closest / closestd exactly as
// its own recursive walk would.
//
// Direct mode: pclose (line.f90:205-234), all pairs in slot order through
// LDS tiles; empty slots keep closest and closestd.
//
// dist = sqrt((dx^2 + dy^2) + dz^2) and width/dist in FP64 as the reference
// forms them, so closest and closestd are bit-identical to the reference's.
// width/dist < 0.5 is decided by a per-node threshold on d2, and a lane
// leaves out subtrees farther than its best distance (neither changes a
// result; see k_closest_tree).
#include "qv_internal.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

constexpr int kBlock = 256;

struct CloseArgs {
  int n;                 // slots of the current state
  int64_t nnodes;
  const int32_t *order;  // tree mode: the slots in the search's order
  const NodeRec *rec;
  const double2 *geo;    // width
  const int32_t *count;  // points per node (root: all slots)
  const double *tclose;  // smallest d2 with width/dist < 0.5 (-1: exact test)
  const double *corner;  // 3 per node: the cell's lower corner
  const int32_t *linf;
  const double *X, *Y, *Z;
  const int32_t *inf, *beh;
  int32_t *cl;
  double *cd;
};

struct alignas(16) CloseEntry {
  int32_t node, end;
  uint64_t mask;
};

__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t uni64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__global__ void __launch_bounds__(kBlock) k_closest_tree(CloseArgs A) {
  __shared__ CloseEntry stack[kBlock / kWave][kStackDepth];
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = w < A.n;
  const int i = valid ? A.order[w] : 0;
  if (valid) A.cd[i] = 100.;
  const bool act = valid && A.inf[i] != 0;
  int32_t fi = 0, bi = 0, best = 0;
  double x0 = 0., x1 = 0., x2 = 0.;
  if (act) {
    fi = A.inf[i];
    bi = A.beh[i];
    best = A.cl[i];
    x0 = A.X[i]; x1 = A.Y[i]; x2 = A.Z[i];
  }
  double bd = 100.;
  uint64_t M = __ballot(act);
  if (M == 0) return;  // the whole wavefront: M is the same in every lane
  CloseEntry *st = stack[threadIdx.x / kWave];
  int n = 0, e = (int)A.nnodes, sp = 0;
  while (true) {
    const NodeRec &r = A.rec[n];
    const double cx = r.cx, cy = r.cy, cz = r.cz;
    const int skip = uni(r.skip);
    const int32_t cnt = uni(A.count[n]);
    const double w = A.geo[n].x;
    const double T = A.tclose[n];
    const bool mine = __builtin_amdgcn_inverse_ballot_w64(M);
    const double d0 = x0 - cx, d1 = x1 - cy, d2 = x2 - cz;
    const double dsq = (d0 * d0 + d1 * d1) + d2 * d2;
    // width/sqrt(dsq) < 0.5 is monotone in dsq (a correctly rounded square
    // root and division): dsq >= T decides it where T >= 0
    const bool open_test = T >= 0. ? dsq >= T : w / sqrt(dsq) < .5;
    const bool taken = mine && (cnt == 1 || open_test);
    // sqrt(dsq) < bd implies dsq < bd^2, and bd * bd * (1 + 1e-7) exceeds bd^2
    // despite its two roundings: lanes failing this cannot be closer
    if (taken && dsq < (bd * bd) * 1.0000001) {
      const double dist = sqrt(dsq);
      if (dist < bd) {
        if (cnt > 1) {
          best = 0;
        } else {
          const int32_t q = A.linf[n];
          const int32_t j = q >= 0 ? A.beh[q] : 0;  // q < 0 only for a 1-slot root
          if (j != i + 1 && fi != j && bi != j) {
            best = j;
            bd = dist;
          }
        }
      }
    }
    // an opened node's subtree is left out for a lane whose best distance
    // is below the distance to the node's cell box: every centroid below
    // lies in the box (to rounding), so no test there can pass dist <
    // closestd, and closestd only falls.  The margin 1e-6 on the square
    // covers the rounding of the box distance and of the centroids; below
    // bd = 1e-8 nothing is left out.
    bool open_lane = mine && !taken;
    if (open_lane && bd > 1e-8) {
      const double *P = A.corner + 3 * (int64_t)n;
      const double q0 = fmax(fmax(P[0] - x0, x0 - (P[0] + w)), 0.);
      const double q1 = fmax(fmax(P[1] - x1, x1 - (P[1] + w)), 0.);
      const double q2 = fmax(fmax(P[2] - x2, x2 - (P[2] + w)), 0.);
      if ((q0 * q0 + q1 * q1) + q2 * q2 > (bd * bd) * 1.000001) open_lane = false;
    }
    const uint64_t O = __ballot(open_lane);
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
  if (act) {
    A.cl[i] = best;
    A.cd[i] = bd;
  }
}

constexpr int kTile = 256;

__global__ void __launch_bounds__(kBlock) k_closest_direct(CloseArgs A) {
  __shared__ double sx[kTile], sy[kTile], sz[kTile];
  __shared__ int32_t sok[kTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < A.n && A.inf[i] != 0;
  double x0 = 0., x1 = 0., x2 = 0.;
  int32_t fi = 0, bi = 0, best = 0;
  if (act) {
    x0 = A.X[i]; x1 = A.Y[i]; x2 = A.Z[i];
    fi = A.inf[i];
    bi = A.beh[i];
    best = A.cl[i];
  }
  double bd = 100.;
  for (int base = 0; base < A.n; base += kTile) {
    __syncthreads();
    const int j = base + threadIdx.x;
    const bool in = j < A.n;
    sx[threadIdx.x] = in ? A.X[j] : 0.;
    sy[threadIdx.x] = in ? A.Y[j] : 0.;
    sz[threadIdx.x] = in ? A.Z[j] : 0.;
    sok[threadIdx.x] = in ? A.inf[j] != 0 : 0;
    __syncthreads();
    const int m = act ? min(kTile, A.n - base) : 0;
    for (int jj = 0; jj < m; ++jj) {
      const int j1 = base + jj + 1;  // 1-based slot
      if (!sok[jj] || j1 == i + 1 || fi == j1 || bi == j1) continue;
      // distf (general.f90:27-34)
      const double d0 = x0 - sx[jj], d1 = x1 - sy[jj], d2 = x2 - sz[jj];
      const double dist = sqrt((d0 * d0 + d1 * d1) + d2 * d2);
      if (dist < bd) {
        best = j1;
        bd = dist;
      }
    }
  }
  if (act) {
    A.cl[i] = best;
    A.cd[i] = bd;
  }
}

}  // namespace

hipError_t launch_closest(Ctx &c, bool tree) {
  CloseArgs A;
  A.n = c.pcount;
  A.nnodes = tree ? c.tree.nnodes : 0;
  A.order = c.close_order;
  A.rec = c.tree.rec;
  A.geo = c.tree.geo;
  A.count = c.tree.count;
  A.tclose = c.tree.tclose;
  A.corner = c.tree.corner;
  A.linf = c.tree.linf;
  A.X = c.x.x; A.Y = c.x.y; A.Z = c.x.z;
  A.inf = c.infront;
  A.beh = c.behind;
  A.cl = c.close_i;
  A.cd = c.close_d;
  const dim3 g((c.pcount + kBlock - 1) / kBlock);
  if (tree && !c.tree.close_valid) {
    const hipError_t e = launch_close_thresholds(c);
    if (e != hipSuccess) return e;
  }
  if (tree)
    hipLaunchKernelGGL(k_closest_tree, g, dim3(kBlock), 0, c.stream, A);
  else
    hipLaunchKernelGGL(k_closest_direct, g, dim3(kBlock), 0, c.stream, A);
  return hipGetLastError();
}

}  // namespace qv
