// qv_probe.hip -- superfluid velocity (and its gradient) at probe points.
//
// The sums of mesh_velocity (timestep.f90:374-452) and of the quasiparticle
// Hamiltonian (velocity_quasip, hamiltonian.f90:26-97): the central shell,
// then the image shells of the boundary in the order p outer .. r inner, one
// running sum per probe.
//   velocity BS:   biot_savart_general / _shift (timestep.f90:516-563) over
//                  the current positions with the context's ghosts, in the
//                  reference's operation order (bit-identical sums).  The
//                  optional gradient du_k/dy_m is the analytic derivative of
//                  the same terms with respect to the probe position y; it
//                  replaces the six extra evaluations and central
//                  differences of velocity_quasip (hamiltonian.f90:64-79).
//   velocity Tree: tree_walk_general (tree.f90:425-488) over the tree of the
//                  last velocity stage.  The probes are put in the order of
//                  their 21-bit Hilbert-curve keys in the box (a radix sort),
//                  and the 64 probes of a wavefront share one DFS per shell,
//                  node records by scalar loads, each lane taking part only
//                  in its own walk's nodes, in the same order as its own
//                  recursive walk.  Fewer than about half a million probes
//                  cannot fill the GPU that way: their shells are split over
//                  several wavefronts per group of 64 probes, whose partial
//                  sums are added in shell order after the walk.  Probes
//                  much sparser than the points walk one per lane.  The
//                  reference adds leaves as
//                  u_bs/4/pi (no Gamma) and cells without Gamma/(4 pi); by
//                  default both get Gamma/(4 pi) here, and
//                  QV_PROBE_REFERENCE_SCALE keeps the reference's factors.
#include "qv_internal.h"
#include "qv_terms.h"

#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>

#include "qv_sort.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 256;

struct ProbeArgs {
  int n, nprobe;
  double qc, theta;
  int grad, ref_scale;
  double core;  // > 0: probes closer than this to any slot's x get u = 0
  const double *X, *Y, *Z, *GX, *GY, *GZ;
  const int32_t *inf;
  const double *py;  // probes, (nprobe,3) row-major
  double *u;         // (nprobe,3)
  double *g;         // (nprobe,9): g[9q + 3k + m] = du_k/dy_m
  // tree
  int64_t nnodes;
  const NodeRec *rec;
  const double2 *geo;
  const int32_t *count;
  const double *tplain;  // plain-test thresholds on d2 (-1: exact test)
  const double *mp;  // internal: circ(3), C(9) row-major C[i][k]; leaf: ghosti(3)
  // shells split over nsplit wavefronts per group of 64 probes: partial sums
  // part[(q * nsplit + c) * 3 + k] of chunk c, added in chunk order after
  int nsplit;
  double *part;
};

// one term of biot_savart_general(_shift) at probe y, source x (+ shift s),
// ghost g; adds its derivative with respect to y when G != nullptr
template <bool kGrad>
__device__ inline void general_term(const double y[3], const double xj[3], const double gj[3],
                                    const double s[3], bool shifted, double qc, double u[3],
                                    double G[9]) {
  double xs[3], r[3], e[3];
  for (int k = 0; k < 3; ++k) xs[k] = shifted ? xj[k] + s[k] : xj[k];
  // dist_gen_sq(y, xs) (general.f90:142-148)
  const double a = ((y[0] - xs[0]) * (y[0] - xs[0]) + (y[1] - xs[1]) * (y[1] - xs[1])) +
                   (y[2] - xs[2]) * (y[2] - xs[2]);
  for (int k = 0; k < 3; ++k) r[k] = xs[k] - y[k];
  for (int k = 0; k < 3; ++k) e[k] = gj[k] - xj[k];
  const double b = 2. * ((r[0] * e[0] + r[1] * e[1]) + r[2] * e[2]);
  const double c = ((gj[0] - xj[0]) * (gj[0] - xj[0]) + (gj[1] - xj[1]) * (gj[1] - xj[1])) +
                   (gj[2] - xj[2]) * (gj[2] - xj[2]);
  const double D = (4. * a) * c - b * b;
  if (D == 0.) return;
  const double den = (2. * kPi) * D;
  const double v[3] = {r[1] * e[2] - r[2] * e[1], r[2] * e[0] - r[0] * e[2],
                       r[0] * e[1] - r[1] * e[0]};
  const double sa = sqrt(a), s2 = sqrt((a + b) + c);
  const double F = (2. * c + b) / s2 - b / sa;
  for (int k = 0; k < 3; ++k) u[k] = u[k] + ((v[k] * qc) / den) * F;
  if constexpr (kGrad) {
    // u = (qc/(2 pi)) v F / D with da/dy = -2r, db/dy = -2e, dc/dy = 0,
    // dv_k/dy_m = -eps_kmn e_n
    const double w = r[0] + e[0], w1 = r[1] + e[1], w2 = r[2] + e[2];
    const double ww[3] = {w, w1, w2};
    const double s23 = s2 * s2 * s2, sa3 = sa * sa * sa;
    const double h = qc / (2. * kPi);
    double dF[3], dD[3];
    for (int m = 0; m < 3; ++m) {
      dF[m] = -2. * e[m] / s2 + (2. * c + b) * ww[m] / s23 + 2. * e[m] / sa - b * r[m] / sa3;
      dD[m] = -8. * c * r[m] + 4. * b * e[m];
    }
    const double iD = 1. / D;
    for (int k = 0; k < 3; ++k) {
      for (int m = 0; m < 3; ++m) {
        double dv = 0.;  // -eps_kmn e_n
        if (k != m) {
          const int nn = 3 - k - m;
          const double sgn = ((k + 1) % 3 == m) ? -1. : 1.;  // eps_kmn = +1 for cyclic (k,m,n)
          dv = sgn * e[nn];
        }
        G[3 * k + m] += h * iD * (dv * F + v[k] * dF[m] - v[k] * F * dD[m] * iD);
      }
    }
  }
}

template <bool kGrad>
__global__ void __launch_bounds__(kBlock) k_probe_bs(ProbeArgs A, Shells S) {
  __shared__ double sx[kTile], sy[kTile], sz[kTile], gx[kTile], gy[kTile], gz[kTile];
  __shared__ int32_t sok[kTile];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = q < A.nprobe;
  double y[3] = {0., 0., 0.};
  if (act)
    for (int k = 0; k < 3; ++k) y[k] = A.py[3 * q + k];
  double u[3] = {0., 0., 0.};
  double G[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  double md = 1e4;  // squared min_dist = 100 (hamiltonian.f90:44)
  for (int sh = 0; sh < S.n; ++sh) {
    const double s[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
    for (int base = 0; base < A.n; base += kTile) {
      __syncthreads();
      const int j = base + threadIdx.x;
      const bool in = j < A.n && A.inf[j] != 0;
      // every slot's x (empty ones too: the empty-core test reads them);
      // the terms use only the active ones
      sx[threadIdx.x] = j < A.n ? A.X[j] : 0.;
      sy[threadIdx.x] = j < A.n ? A.Y[j] : 0.;
      sz[threadIdx.x] = j < A.n ? A.Z[j] : 0.;
      gx[threadIdx.x] = in ? A.GX[j] : 0.;
      gy[threadIdx.x] = in ? A.GY[j] : 0.;
      gz[threadIdx.x] = in ? A.GZ[j] : 0.;
      sok[threadIdx.x] = in;
      __syncthreads();
      const int m = act ? min(kTile, A.n - base) : 0;
      if (A.core > 0. && sh == 0)
        for (int jj = 0; jj < m; ++jj) {
          // dist_gen(pos, f(j)%x) over every slot (hamiltonian.f90:45-48)
          const double d0 = y[0] - sx[jj], d1 = y[1] - sy[jj], d2 = y[2] - sz[jj];
          md = fmin(md, (d0 * d0 + d1 * d1) + d2 * d2);
        }
      for (int jj = 0; jj < m; ++jj) {
        if (!sok[jj]) continue;
        const double xj[3] = {sx[jj], sy[jj], sz[jj]};
        const double gj[3] = {gx[jj], gy[jj], gz[jj]};
        general_term<kGrad>(y, xj, gj, s, sh != 0, A.qc, u, G);
      }
    }
  }
  if (!act) return;
  // min_dist < corea: the empty core, zero velocity and gradient
  // (hamiltonian.f90:53-58); min of sqrt = sqrt of min (sqrt is monotone)
  if (A.core > 0. && sqrt(md) < A.core) {
    for (int k = 0; k < 3; ++k) u[k] = 0.;
    for (int k = 0; k < 9; ++k) G[k] = 0.;
  }
  for (int k = 0; k < 3; ++k) A.u[3 * q + k] = u[k];
  if constexpr (kGrad)
    for (int k = 0; k < 9; ++k) A.g[9 * q + k] = G[k];
}

// probes fewer than points / kLaneWalkRatio take one walk per lane
constexpr int64_t kLaneWalkRatio = 16;

struct alignas(16) ProbeEntry {
  int32_t node, end;
  uint64_t mask;
};

__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t uni64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// wavefront gw walks shell chunk gw % nsplit for the 64 probes of group
// gw / nsplit; lane l of group p works on probe order[64 p + l]
__global__ void __launch_bounds__(kBlock) k_probe_tree(ProbeArgs A, Shells S,
                                                       const int32_t *__restrict__ order) {
  __shared__ ProbeEntry stack[kBlock / kWave][kStackDepth];
  const int gw = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const int chunk = gw % A.nsplit;
  const int qs = (gw / A.nsplit) * kWave + threadIdx.x % kWave;
  const bool act = qs < A.nprobe;
  const int q = act ? order[qs] : 0;
  double y[3] = {0., 0., 0.};
  if (act)
    for (int k = 0; k < 3; ++k) y[k] = A.py[3 * q + k];
  const uint64_t all = __ballot(act);
  if (all == 0) return;  // the whole wavefront
  ProbeEntry *st = stack[threadIdx.x / kWave];
  // leaf and cell terms in two sums: the factors (Gamma/(4 pi) for both, or
  // the reference's 1/(4 pi) for leaves and 1 for cells) are applied once
  double ul[3] = {0., 0., 0.}, uc[3] = {0., 0., 0.};
  const int sh0 = chunk * S.n / A.nsplit, sh1 = (chunk + 1) * S.n / A.nsplit;
  for (int sh = sh0; sh < sh1; ++sh) {
    const double s[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
    int n = 0, e = (int)A.nnodes, sp = 0;
    uint64_t M = all;
    while (true) {
      const NodeRec &r = A.rec[n];
      const double c3[3] = {r.cx, r.cy, r.cz};
      const int skip = uni(r.skip);
      const int32_t cnt = uni(A.count[n]);
      const bool mine = __builtin_amdgcn_inverse_ballot_w64(M);
      const double d0 = y[0] - (c3[0] + s[0]), d1 = y[1] - (c3[1] + s[1]),
                   d2 = y[2] - (c3[2] + s[2]);
      const double dsq = (d0 * d0 + d1 * d1) + d2 * d2;
      // dist < epsilon(0.): the node is left alone (avoids 1/0,
      // tree.f90:441); sqrt(dsq) < 2^-52 is exactly dsq < 2^-104
      const bool live = mine && !(dsq < 0x1p-104);
      bool taken;
      if (cnt == 1) {
        taken = mine;
        if (live) {
          // circ of a one-point cell is (0 + ghosti) - x (tree.f90:160-165)
          double R1[3], R2[3];
          for (int k = 0; k < 3; ++k) {
            const double circ = r.g[k] - c3[k];
            R1[k] = (c3[k] + s[k]) - y[k];
            R2[k] = ((c3[k] + circ) + s[k]) - y[k];
          }
          const double a = sqrt((R1[0] * R1[0] + R1[1] * R1[1]) + R1[2] * R1[2]);
          const double b = sqrt((R2[0] * R2[0] + R2[1] * R2[1]) + R2[2] * R2[2]);
          const double cc =
              (a * b) * ((a * b) + ((R1[0] * R2[0] + R1[1] * R2[1]) + R1[2] * R2[2]));
          const double f = (a + b) / cc;
          const double ub[3] = {f * (R1[1] * R2[2] - R1[2] * R2[1]),
                                f * (R1[2] * R2[0] - R1[0] * R2[2]),
                                f * (R1[0] * R2[1] - R1[1] * R2[0])};
          for (int k = 0; k < 3; ++k) ul[k] = ul[k] + ub[k];
        }
      } else {
        // width/dist < theta (tree.f90:443) as dsq >= the node's exact
        // threshold; the root (and any node without one) takes the test
        const double T = A.tplain[n];
        bool acc;
        if (uni((int32_t)(__double_as_longlong(T) >> 32)) >= 0) {
          acc = dsq >= T;
        } else {
          acc = A.geo[n].x / sqrt(dsq) < A.theta;
        }
        const bool cell = live && acc;
        taken = mine && (!live || cell);
        if (__ballot(cell)) {
          // the multipole regrouped (as in the velocity walk): within a few
          // ulp of the reference's k, l, m loops (tree.f90:466-478)
          const double v[3] = {d0, d1, d2};
          double ub[3];
          cell64(cell, v, dsq, A.mp + 12 * (int64_t)n, ub);
          for (int k = 0; k < 3; ++k) uc[k] = uc[k] + ub[k];
        }
      }
      const uint64_t O = M & ~__ballot(taken);
      if (O != 0 && n + 1 != skip) {
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
        n = skip;
      } else {
        if (sp == 0) break;
        --sp;
        n = uni(st[sp].node);
        e = uni(st[sp].end);
        M = uni64(st[sp].mask);
      }
    }
  }
  if (!act) return;
  double u[3];
  for (int k = 0; k < 3; ++k)
    u[k] = A.ref_scale ? (ul[k] / 4.) / kPi + uc[k] : ((ul[k] + uc[k]) * A.qc / 4.) / kPi;
  if (A.nsplit == 1) {
    for (int k = 0; k < 3; ++k) A.u[3 * q + k] = u[k];
  } else {
    double *pp = A.part + 3 * ((int64_t)q * A.nsplit + chunk);
    for (int k = 0; k < 3; ++k) pp[k] = u[k];
  }
}

// one walk per lane, for probes too scattered for a wavefront to share one:
// stackless, since an opened node continues at n + 1 (its first stored
// child) and any other node at its skip pointer (the next node of the
// preorder outside its subtree), so every lane meets its own nodes in the
// DFS order of tree_walk_general.  Shell chunks and output as k_probe_tree.
__global__ void __launch_bounds__(kBlock) k_probe_lane(ProbeArgs A, Shells S,
                                                       const int32_t *__restrict__ order) {
  const int gw = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const int chunk = gw % A.nsplit;
  const int qs = (gw / A.nsplit) * kWave + threadIdx.x % kWave;
  const bool act = qs < A.nprobe;
  const int q = act ? order[qs] : 0;
  double y[3] = {0., 0., 0.};
  if (act)
    for (int k = 0; k < 3; ++k) y[k] = A.py[3 * q + k];
  const int nn = (int)A.nnodes;
  double ul[3] = {0., 0., 0.}, uc[3] = {0., 0., 0.};
  const int sh0 = chunk * S.n / A.nsplit, sh1 = (chunk + 1) * S.n / A.nsplit;
  for (int sh = sh0; sh < sh1; ++sh) {
    const double s[3] = {S.s[sh][0], S.s[sh][1], S.s[sh][2]};
    int n = act ? 0 : nn;
    while (n < nn) {
      const NodeRec r = A.rec[n];
      const double c3[3] = {r.cx, r.cy, r.cz};
      const double d0 = y[0] - (c3[0] + s[0]), d1 = y[1] - (c3[1] + s[1]),
                   d2 = y[2] - (c3[2] + s[2]);
      const double dsq = (d0 * d0 + d1 * d1) + d2 * d2;
      // dist < epsilon(0.) leaves the node alone (tree.f90:441)
      const bool live = !(dsq < 0x1p-104);
      bool open = false;
      if (A.count[n] == 1) {
        if (live) {
          double R1[3], R2[3];
          for (int k = 0; k < 3; ++k) {
            const double circ = r.g[k] - c3[k];
            R1[k] = (c3[k] + s[k]) - y[k];
            R2[k] = ((c3[k] + circ) + s[k]) - y[k];
          }
          const double a = sqrt((R1[0] * R1[0] + R1[1] * R1[1]) + R1[2] * R1[2]);
          const double b = sqrt((R2[0] * R2[0] + R2[1] * R2[1]) + R2[2] * R2[2]);
          const double cc =
              (a * b) * ((a * b) + ((R1[0] * R2[0] + R1[1] * R2[1]) + R1[2] * R2[2]));
          const double f = (a + b) / cc;
          ul[0] = ul[0] + f * (R1[1] * R2[2] - R1[2] * R2[1]);
          ul[1] = ul[1] + f * (R1[2] * R2[0] - R1[0] * R2[2]);
          ul[2] = ul[2] + f * (R1[0] * R2[1] - R1[1] * R2[0]);
        }
      } else {
        const double T = A.tplain[n];
        const bool acc = T >= 0. ? dsq >= T : A.geo[n].x / sqrt(dsq) < A.theta;
        if (live && acc) {
          const double v[3] = {d0, d1, d2};
          double ub[3];
          cell64(true, v, dsq, A.mp + 12 * (int64_t)n, ub);
          for (int k = 0; k < 3; ++k) uc[k] = uc[k] + ub[k];
        }
        open = live && !acc;
      }
      n = open ? n + 1 : r.skip;
    }
  }
  if (!act) return;
  double u[3];
  for (int k = 0; k < 3; ++k)
    u[k] = A.ref_scale ? (ul[k] / 4.) / kPi + uc[k] : ((ul[k] + uc[k]) * A.qc / 4.) / kPi;
  if (A.nsplit == 1) {
    for (int k = 0; k < 3; ++k) A.u[3 * q + k] = u[k];
  } else {
    double *pp = A.part + 3 * ((int64_t)q * A.nsplit + chunk);
    for (int k = 0; k < 3; ++k) pp[k] = u[k];
  }
}

// u of each probe: its chunks' partial sums in chunk (= shell) order
__global__ void __launch_bounds__(kBlock) k_probe_join(ProbeArgs A) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.nprobe) return;
  const double *pp = A.part + 3 * (int64_t)q * A.nsplit;
  double u[3] = {pp[0], pp[1], pp[2]};
  for (int c = 1; c < A.nsplit; ++c)
    for (int k = 0; k < 3; ++k) u[k] = u[k] + pp[3 * c + k];
  for (int k = 0; k < 3; ++k) A.u[3 * q + k] = u[k];
}

}  // namespace

hipError_t launch_probe(Ctx &c, int nprobe, const double *d_y, double *d_u, double *d_g,
                        double *d_work, bool tree, bool ref_scale, bool empty_core) {
  ProbeArgs A;
  A.n = c.pcount;
  A.nprobe = nprobe;
  A.qc = c.p.quant_circ;
  A.theta = c.p.tree_theta;
  A.grad = d_g != nullptr;
  A.ref_scale = ref_scale;
  A.core = empty_core ? c.p.corea : 0.;
  A.X = c.x.x; A.Y = c.x.y; A.Z = c.x.z;
  A.GX = c.gi.x; A.GY = c.gi.y; A.GZ = c.gi.z;
  A.inf = c.infront;
  A.py = d_y;
  A.u = d_u;
  A.g = d_g;
  A.nnodes = tree ? c.tree.nnodes : 0;
  A.rec = c.tree.rec;
  A.geo = c.tree.geo;
  A.count = c.tree.count;
  A.tplain = c.tree.tplain;
  A.mp = c.tree.mp;
  A.nsplit = 1;
  A.part = nullptr;
  const Shells S = make_shells(c.p);
  const dim3 g((nprobe + kBlock - 1) / kBlock);
  if (tree) {
    if (!c.tree.plain_valid) {
      const hipError_t e = launch_plain_thresholds(c);
      if (e != hipSuccess) return e;
    }
    // probes in Hilbert-curve order: 2 x nprobe keys, then 2 x nprobe indices
    uint32_t *key = reinterpret_cast<uint32_t *>(d_work), *key_s = key + nprobe;
    int32_t *idx = reinterpret_cast<int32_t *>(key_s + nprobe), *idx_s = idx + nprobe;
    hipError_t err = launch_hilbert_keys(c.stream, nprobe, c.p.box_size, d_y, d_y + 1, d_y + 2,
                                         3, nullptr, key, idx);
    if (err != hipSuccess) return err;
    size_t need = 0;
    err = rocprim::radix_sort_pairs<SortConfig>(nullptr, need, key, key_s, idx, idx_s, nprobe, 0,
                                                21, c.stream);
    if (err != hipSuccess) return err;
    if (need > c.sort_tmp_bytes) {
      if ((err = hipStreamSynchronize(c.stream)) != hipSuccess) return err;
      (void)hipFree(c.sort_tmp);
      c.sort_tmp = nullptr;
      c.sort_tmp_bytes = 0;
      if ((err = hipMalloc(&c.sort_tmp, need)) != hipSuccess) return err;
      c.sort_tmp_bytes = need;
    }
    size_t have = c.sort_tmp_bytes;
    err = rocprim::radix_sort_pairs<SortConfig>(c.sort_tmp, have, key, key_s, idx, idx_s, nprobe,
                                                0, 21, c.stream);
    if (err != hipSuccess) return err;
    // a few probes cannot fill the GPU one wavefront per 64 of them: their
    // shells are then split over up to S.n wavefronts (about 8 per SIMD in
    // all), each with its own partial sums
    const int64_t groups = (nprobe + kWave - 1) / kWave;
    const int64_t want = 8 * 1024;
    A.nsplit = (int)std::min<int64_t>(S.n, std::max<int64_t>(1, (want + groups - 1) / groups));
    A.part = nullptr;
    if (A.nsplit > 1) {
      const size_t need_part = (size_t)3 * nprobe * A.nsplit;
      if (need_part > c.probe_part_cap) {
        if ((err = hipStreamSynchronize(c.stream)) != hipSuccess) return err;
        (void)hipFree(c.probe_part);
        c.probe_part = nullptr;
        c.probe_part_cap = 0;
        if ((err = hipMalloc(&c.probe_part, need_part * sizeof(double))) != hipSuccess) return err;
        c.probe_part_cap = need_part;
      }
      A.part = c.probe_part;
    }
    const int64_t waves = groups * A.nsplit;
    const dim3 gt((unsigned)((waves + kBlock / kWave - 1) / (kBlock / kWave)));
    // a wavefront shares one walk among 64 probes adjacent on the curve when
    // they are about as dense as the points; scattered probes walk alone
    if ((int64_t)nprobe * kLaneWalkRatio < (int64_t)c.nact)
      hipLaunchKernelGGL(k_probe_lane, gt, dim3(kBlock), 0, c.stream, A, S, idx_s);
    else
      hipLaunchKernelGGL(k_probe_tree, gt, dim3(kBlock), 0, c.stream, A, S, idx_s);
    if (A.nsplit > 1) {
      if ((err = hipGetLastError()) != hipSuccess) return err;
      hipLaunchKernelGGL(k_probe_join, g, dim3(kBlock), 0, c.stream, A);
    }
  }
  else if (d_g)
    hipLaunchKernelGGL(k_probe_bs<true>, g, dim3(kBlock), 0, c.stream, A, S);
  else
    hipLaunchKernelGGL(k_probe_bs<false>, g, dim3(kBlock), 0, c.stream, A, S);
  return hipGetLastError();
}

}  // namespace qv
