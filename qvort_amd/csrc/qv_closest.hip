// qv_closest.hip -- neighbour search of the reconnection step.
//
// Tree mode: pclose_tree / closest_tree (tree.f90:219-278).  One thread per
// slot walks the octree of the last velocity stage (built, like
// construct_tree's, from the positions before the step's motion) with the
// slot's current position, in DFS preorder without a stack: a taken node
// continues at its skip pointer, an opened node at the next record (its
// first stored child; empty children are not stored and would return).  A
// node is taken when it holds one point or width/dist < 0.5; when dist <
// closestd, a many-point node sets closest = 0 (closestd unchanged) and a
// one-point node offers j = behind(infront_at_build(point)) with the current
// behind, taken unless j is i, infront(i) or behind(i).  closestd starts at
// 100 for every slot, empty or not (tree.f90:227).
//
// Direct mode: pclose (line.f90:205-234), all pairs in slot order through
// LDS tiles; empty slots keep closest and closestd.
//
// dist = sqrt((dx^2 + dy^2) + dz^2) and width/dist in FP64 as the reference
// forms them, so closest and closestd are bit-identical to the reference's.
#include "qv_internal.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

constexpr int kBlock = 256;

struct CloseArgs {
  int n;                 // slots of the current state
  int64_t nnodes;
  const NodeRec *rec;
  const double2 *geo;    // width
  const int32_t *count;  // points per node (root: all slots)
  const int32_t *linf;
  const double *X, *Y, *Z;
  const int32_t *inf, *beh;
  int32_t *cl;
  double *cd;
};

__global__ void __launch_bounds__(kBlock) k_closest_tree(CloseArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  A.cd[i] = 100.;
  const int32_t fi = A.inf[i];
  if (fi == 0) return;
  const int32_t bi = A.beh[i];
  const double x0 = A.X[i], x1 = A.Y[i], x2 = A.Z[i];
  int32_t best = A.cl[i];
  double bd = 100.;
  int64_t n = 0;
  while (n < A.nnodes) {
    const NodeRec &r = A.rec[n];
    const double d0 = x0 - r.cx, d1 = x1 - r.cy, d2 = x2 - r.cz;
    const double dist = sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    const int32_t cnt = A.count[n];
    const double theta = A.geo[n].x / dist;
    if (cnt == 1 || theta < .5) {
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
      n = r.skip;
    } else {
      n = n + 1;
    }
  }
  A.cl[i] = best;
  A.cd[i] = bd;
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
  A.rec = c.tree.rec;
  A.geo = c.tree.geo;
  A.count = c.tree.count;
  A.linf = c.tree.linf;
  A.X = c.x.x; A.Y = c.x.y; A.Z = c.x.z;
  A.inf = c.infront;
  A.beh = c.behind;
  A.cl = c.close_i;
  A.cd = c.close_d;
  const dim3 g((c.pcount + kBlock - 1) / kBlock);
  if (tree)
    hipLaunchKernelGGL(k_closest_tree, g, dim3(kBlock), 0, c.stream, A);
  else
    hipLaunchKernelGGL(k_closest_direct, g, dim3(kBlock), 0, c.stream, A);
  return hipGetLastError();
}

}  // namespace qv
