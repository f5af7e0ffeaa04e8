// qv_pointwise.hip -- per-slot kernels of the velocity stage: ghost gather,
// local (LIA) term, Adams-Bashforth ladder, periodic wrap, layout transposes.
//
// All FP64 arithmetic keeps the reference's operation order and is compiled
// without contraction, so these kernels reproduce the flang build of the
// reference bit for bit (log() aside, which may differ in the last ulp).
#include "qv_internal.h"
#include "qv_lia.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

constexpr int kBlock = 256;

__global__ void k_aos_to_soa(const double *__restrict__ aos, double *__restrict__ x,
                             double *__restrict__ y, double *__restrict__ z, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = aos[3 * i + 0];
  y[i] = aos[3 * i + 1];
  z[i] = aos[3 * i + 2];
}

__global__ void k_soa_to_aos(const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, double *__restrict__ aos, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  aos[3 * i + 0] = x[i];
  aos[3 * i + 1] = y[i];
  aos[3 * i + 2] = z[i];
}

// periodic unwrap of one ghost coordinate (periodic.f90:36-55 and repeats)
__device__ inline double unwrap(double xi, double g, double L, bool on) {
  if (!on) return g;
  if ((xi - g) > (L / 2.)) return g + L;
  if ((xi - g) < (-L / 2.)) return g - L;
  return g;
}

struct Lia {
  double qc, corea, L;
  int order;
  bool px, py, pz;  // periodic axes
};

// ghostp (periodic.f90:8-184) fused with the local term of calc_velocity
// (timestep.f90:221-225 / 270-274): u_LI = Gamma/(4 pi) ln((1.21/a)
// sqrt(l_i l_b)) s' x s''
__global__ void __launch_bounds__(kBlock)
k_ghost_lia(Lia P, int n, const double *__restrict__ X, const double *__restrict__ Y,
            const double *__restrict__ Z, const int32_t *__restrict__ inf,
            const int32_t *__restrict__ beh, double *__restrict__ gix,
            double *__restrict__ giy, double *__restrict__ giz, double *__restrict__ gbx,
            double *__restrict__ gby, double *__restrict__ gbz, double *__restrict__ giix,
            double *__restrict__ giiy, double *__restrict__ giiz, double *__restrict__ gbbx,
            double *__restrict__ gbby, double *__restrict__ gbbz, double *__restrict__ ulx,
            double *__restrict__ uly, double *__restrict__ ulz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int fi = inf[i];
  if (fi == 0) {
    gix[i] = giy[i] = giz[i] = 0.;
    gbx[i] = gby[i] = gbz[i] = 0.;
    giix[i] = giiy[i] = giiz[i] = 0.;
    gbbx[i] = gbby[i] = gbbz[i] = 0.;
    ulx[i] = uly[i] = ulz[i] = 0.;
    return;
  }
  const int a = fi - 1, b = beh[i] - 1;
  const int aa = inf[a] - 1, bb = beh[b] - 1;
  const double x[3] = {X[i], Y[i], Z[i]};
  double gi[3] = {X[a], Y[a], Z[a]};
  double gb[3] = {X[b], Y[b], Z[b]};
  double gii[3] = {X[aa], Y[aa], Z[aa]};
  double gbb[3] = {X[bb], Y[bb], Z[bb]};
  const bool per[3] = {P.px, P.py, P.pz};
  for (int k = 0; k < 3; ++k) {
    gi[k] = unwrap(x[k], gi[k], P.L, per[k]);
    gb[k] = unwrap(x[k], gb[k], P.L, per[k]);
    gii[k] = unwrap(x[k], gii[k], P.L, per[k]);
    gbb[k] = unwrap(x[k], gbb[k], P.L, per[k]);
  }
  gix[i] = gi[0]; giy[i] = gi[1]; giz[i] = gi[2];
  gbx[i] = gb[0]; gby[i] = gb[1]; gbz[i] = gb[2];
  giix[i] = gii[0]; giiy[i] = gii[1]; giiz[i] = gii[2];
  gbbx[i] = gbb[0]; gbby[i] = gbb[1]; gbbz[i] = gbb[2];
  double ul[3];
  lia_term<double>(P.qc, P.corea, P.order, x, gi, gb, gii, gbb, ul);
  ulx[i] = ul[0];
  uly[i] = ul[1];
  ulz[i] = ul[2];
}

__device__ inline double maxabs3(double a, double b, double c) {
  double m = fabs(a);
  if (fabs(b) > m) m = fabs(b);
  if (fabs(c) > m) m = fabs(c);
  return m;
}

// Adams-Bashforth ladder keyed on exact zeros of the history
// (timestep.f90:48-71): Euler, AB2, AB3, AB4, then u3<-u2<-u1<-u
__global__ void __launch_bounds__(kBlock)
k_ab_update(int n, double dt, const int32_t *__restrict__ inf, double *__restrict__ X,
            double *__restrict__ Y, double *__restrict__ Z, const double *__restrict__ ux,
            const double *__restrict__ uy, const double *__restrict__ uz,
            double *__restrict__ u1x, double *__restrict__ u1y, double *__restrict__ u1z,
            double *__restrict__ u2x, double *__restrict__ u2y, double *__restrict__ u2z,
            double *__restrict__ u3x, double *__restrict__ u3y, double *__restrict__ u3z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || inf[i] == 0) return;
  double x[3] = {X[i], Y[i], Z[i]};
  const double u[3] = {ux[i], uy[i], uz[i]};
  const double a1[3] = {u1x[i], u1y[i], u1z[i]};
  const double a2[3] = {u2x[i], u2y[i], u2z[i]};
  const double a3[3] = {u3x[i], u3y[i], u3z[i]};
  const double three_twos = 3. / 2., one_half = 1. / 2.;
  const double twenty_three_twelve = 23. / 12., four_thirds = 4. / 3., five_twelths = 5. / 12.;
  if (maxabs3(a1[0], a1[1], a1[2]) == 0.) {
    for (int k = 0; k < 3; ++k) x[k] = x[k] + dt * u[k];
  } else if (maxabs3(a2[0], a2[1], a2[2]) == 0.) {
    for (int k = 0; k < 3; ++k) x[k] = (x[k] + (three_twos * dt) * u[k]) - (one_half * dt) * a1[k];
  } else if (maxabs3(a3[0], a3[1], a3[2]) == 0.) {
    for (int k = 0; k < 3; ++k)
      x[k] = ((x[k] + (twenty_three_twelve * dt) * u[k]) - (four_thirds * dt) * a1[k]) +
             (five_twelths * dt) * a2[k];
  } else {
    for (int k = 0; k < 3; ++k)
      x[k] = x[k] + (dt * (((55. * u[k] - 59. * a1[k]) + 37. * a2[k]) - 9. * a3[k])) / 24.;
  }
  X[i] = x[0]; Y[i] = x[1]; Z[i] = x[2];
  u3x[i] = a2[0]; u3y[i] = a2[1]; u3z[i] = a2[2];
  u2x[i] = a1[0]; u2y[i] = a1[1]; u2z[i] = a1[2];
  u1x[i] = u[0]; u1y[i] = u[1]; u1z[i] = u[2];
}

// the low-storage update of one pmotion_RK3 stage (timestep.f90:95-176):
// stage 0: u1 = u dt, x += u1/3; stage 1: u1 = -(2/3) u1 + u dt, x += u1;
// stage 2: u1 = -u1 + u dt, x += u1/2
__global__ void __launch_bounds__(kBlock)
k_rk3_update(int n, double dt, int stage, const int32_t *__restrict__ inf,
             double *__restrict__ X, double *__restrict__ Y, double *__restrict__ Z,
             const double *__restrict__ ux, const double *__restrict__ uy,
             const double *__restrict__ uz, double *__restrict__ u1x, double *__restrict__ u1y,
             double *__restrict__ u1z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || inf[i] == 0) return;
  double *P[3] = {X, Y, Z}, *A[3] = {u1x, u1y, u1z};
  const double *U[3] = {ux, uy, uz};
  for (int k = 0; k < 3; ++k) {
    double a = A[k][i];
    const double v = U[k][i];
    double x = P[k][i];
    if (stage == 0) {
      a = v * dt;
      x = x + (a / 3.);
    } else if (stage == 1) {
      a = -((2. / 3.) * a) + v * dt;
      x = x + a;
    } else {
      a = -a + v * dt;
      x = x + 0.5 * a;
    }
    A[k][i] = a;
    P[k][i] = x;
  }
}

// enforce_periodic / _yz / _z wrap part (periodic.f90:188-276), for states
// with no point beyond an open face (k_open_check); loops that leave through
// one are removed on the host (boundary_step_host)
__global__ void k_wrap(int n, double L, int k0, const int32_t *__restrict__ inf,
                       double *__restrict__ X, double *__restrict__ Y, double *__restrict__ Z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || inf[i] == 0) return;
  double *A[3] = {X, Y, Z};
  for (int k = k0; k < 3; ++k) {
    const double v = A[k][i];
    if (v > (L / 2.)) A[k][i] = v - L;
    else if (v < (-L / 2.)) A[k][i] = v + L;
  }
}

// an active point beyond an open face of an openx (x) or openxy (x, y) box:
// the loops to remove (enforce_periodic_yz / _z, periodic.f90:217-276)
__global__ void k_open_check(int n, double L, int nopen, const int32_t *__restrict__ inf,
                             const double *__restrict__ X, const double *__restrict__ Y,
                             int32_t *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool out = false;
  if (i < n && inf[i] != 0) {
    const double a = X[i];
    out = a > (L / 2.) || a < (-L / 2.);
    if (nopen > 1) {
      const double b = Y[i];
      out = out || b > (L / 2.) || b < (-L / 2.);
    }
  }
  if (__ballot(out) && (threadIdx.x % kWave) == 0) atomicOr(flag, 1);
}

// zero what clear_particle zeroes (general.f90:11-22) in the listed slots
// whose positions and links came back from the host: velocity, AB history
// u1 and u2 (not u3), the LIA and nonlocal parts, the ghosts
__global__ void k_clear_slots(int m, const int32_t *__restrict__ list, ClearArgs A) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const int i = list[q];
  for (int k = 0; k < ClearArgs::kArrays; ++k) A.a[k][i] = 0.;
}

// the topology upload after host line operations.  A slot that held no
// point before (or lies past the old slot count) and holds one now is a
// point pinsert made: u1 = u2 = 0 (line.f90:93; an appended slot's u3 = 0
// too); a slot that held a point and is empty now was cleared by
// clear_particle (general.f90:11-22): velocity, u1, u2, the LIA and
// nonlocal parts and the ghosts are zeroed.  mark[s] = 1 for a new point.
// With keep_x (positions not uploaded), a cleared slot's x is zeroed here
// and a new point sets *flag.
__global__ void k_topo_mark(int n_new, int n_old, const int32_t *__restrict__ old_inf,
                            const int32_t *__restrict__ new_inf, int32_t *__restrict__ mark,
                            ClearArgs A, double *__restrict__ u3x, double *__restrict__ u3y,
                            double *__restrict__ u3z, int keep_x, double *__restrict__ xx,
                            double *__restrict__ xy, double *__restrict__ xz,
                            int32_t *__restrict__ flag) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_new) return;
  const bool was = s < n_old && old_inf[s] != 0;
  const bool now = new_inf[s] != 0;
  mark[s] = (!was && now) ? 1 : 0;
  if (keep_x && !was && now) atomicOr(flag, 1);
  if (was && !now) {
    for (int k = 0; k < ClearArgs::kArrays; ++k) A.a[k][s] = 0.;
    if (keep_x) { xx[s] = 0.; xy[s] = 0.; xz[s] = 0.; }
  } else if (!was && now) {
    // u1, u2: entries 3..8 of the list (u, u1, u2, ...)
    for (int k = 3; k < 9; ++k) A.a[k][s] = 0.;
    if (s >= n_old) { u3x[s] = 0.; u3y[s] = 0.; u3z[s] = 0.; }
  }
}

// pinsert's velocity of a new point (line.f90:86): the mean of the
// velocities of the point it was inserted after and of that point's infront
// at the time.  A point inserted after a new one continues a chain: after
// pinsert the links read old -> new -> new ... -> old, each new point made
// by its behind and inserted in front of the chain's end, so walking a
// chain from its old start gives every value in the order pinsert set them.
__global__ void k_topo_new_velocity(int n, const int32_t *__restrict__ mark,
                                    const int32_t *__restrict__ inf,
                                    const int32_t *__restrict__ beh, double *__restrict__ ux,
                                    double *__restrict__ uy, double *__restrict__ uz) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !mark[p]) return;
  const int b = beh[p] - 1;
  if (mark[b]) return;  // not the start of a chain
  int f = p;
  for (int it = 0; it < n && mark[f]; ++it) f = inf[f] - 1;
  const double uf[3] = {ux[f], uy[f], uz[f]};
  double prev[3] = {ux[b], uy[b], uz[b]};
  int q = p;
  for (int it = 0; it < n && mark[q]; ++it) {
    const double v[3] = {0.5 * (prev[0] + uf[0]), 0.5 * (prev[1] + uf[1]),
                         0.5 * (prev[2] + uf[2])};
    ux[q] = v[0]; uy[q] = v[1]; uz[q] = v[2];
    prev[0] = v[0]; prev[1] = v[1]; prev[2] = v[2];
    q = inf[q] - 1;
  }
}

// remove_mean_pos (timestep.f90:86-92), first pass: per block the sums of
// x and y over the active slots of its range (fixed order: a strided sum
// per thread, then a tree over the block)
constexpr int kMeanBlocks = 512;

__global__ void __launch_bounds__(kBlock) k_mean_partial(int n, const int32_t *__restrict__ inf,
                                                         const double *__restrict__ X,
                                                         const double *__restrict__ Y,
                                                         double *__restrict__ part) {
  __shared__ double sx[kBlock], sy[kBlock];
  const int per = (n + kMeanBlocks - 1) / kMeanBlocks;
  const int lo = blockIdx.x * per, hi = min(n, lo + per);
  double a = 0., b = 0.;
  for (int i = lo + (int)threadIdx.x; i < hi; i += kBlock)
    if (inf[i] > 0) {
      a = a + X[i];
      b = b + Y[i];
    }
  sx[threadIdx.x] = a;
  sy[threadIdx.x] = b;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sx[threadIdx.x] = sx[threadIdx.x] + sx[threadIdx.x + h];
      sy[threadIdx.x] = sy[threadIdx.x] + sy[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sx[0];
    part[2 * blockIdx.x + 1] = sy[0];
  }
}

// second pass: every block adds the partial sums in block order (the same
// sums everywhere), mean = sum / count, subtracted from every slot (empty
// ones too, as f(:)%x(1) - mean_x does)
__global__ void __launch_bounds__(kBlock) k_mean_apply(int n, int nact,
                                                       const double *__restrict__ part,
                                                       double *__restrict__ X,
                                                       double *__restrict__ Y) {
  __shared__ double m[2];
  if (threadIdx.x == 0) {
    double a = 0., b = 0.;
    for (int k = 0; k < kMeanBlocks; ++k) {
      a = a + part[2 * k];
      b = b + part[2 * k + 1];
    }
    m[0] = a / (double)nact;
    m[1] = b / (double)nact;
  }
  __syncthreads();
  const double mx = m[0], my = m[1];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    X[i] = X[i] - mx;
    Y[i] = Y[i] - my;
  }
}

inline int grid(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

Shells make_shells(const qv_params &p) {
  Shells s;
  const double L = p.box_size;
  s.n = 1;
  s.s[0][0] = s.s[0][1] = s.s[0][2] = 0.;
  auto add = [&](int a, int b, int c) {
    s.s[s.n][0] = a * L;
    s.s[s.n][1] = b * L;
    s.s[s.n][2] = c * L;
    ++s.n;
  };
  if (p.boundary == QV_BC_PERIODIC) {
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c)
          if (a || b || c) add(a, b, c);
  } else if (p.boundary == QV_BC_OPENX) {
    for (int b = -1; b <= 1; ++b)
      for (int c = -1; c <= 1; ++c)
        if (b || c) add(0, b, c);
  } else if (p.boundary == QV_BC_OPENXY) {
    for (int c = -1; c <= 1; ++c)
      if (c) add(0, 0, c);
  }
  return s;
}

hipError_t launch_aos_to_soa(Ctx &c, const double *d_aos, Vec3 dst) {
  if (c.pcount == 0) return hipSuccess;
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid(c.pcount)), dim3(kBlock), 0, c.stream, d_aos, dst.x,
                     dst.y, dst.z, c.pcount);
  return hipGetLastError();
}

hipError_t launch_soa_to_aos(Ctx &c, Vec3 src, double *d_aos) {
  if (c.pcount == 0) return hipSuccess;
  hipLaunchKernelGGL(k_soa_to_aos, dim3(grid(c.pcount)), dim3(kBlock), 0, c.stream, src.x, src.y,
                     src.z, d_aos, c.pcount);
  return hipGetLastError();
}

hipError_t launch_ghost_lia(Ctx &c) {
  Lia P;
  P.qc = c.p.quant_circ;
  P.corea = c.p.corea;
  P.L = c.p.box_size;
  P.order = c.p.deriv_order;
  P.px = c.p.boundary == QV_BC_PERIODIC;
  P.py = c.p.boundary == QV_BC_PERIODIC || c.p.boundary == QV_BC_OPENX;
  P.pz = c.p.boundary != QV_BC_OPEN;
  hipLaunchKernelGGL(k_ghost_lia, dim3(grid(c.pcount)), dim3(kBlock), 0, c.stream, P, c.pcount,
                     c.x.x, c.x.y, c.x.z, c.infront, c.behind, c.gi.x, c.gi.y, c.gi.z, c.gb.x,
                     c.gb.y, c.gb.z, c.gii.x, c.gii.y, c.gii.z, c.gbb.x, c.gbb.y, c.gbb.z,
                     c.uli.x, c.uli.y, c.uli.z);
  return hipGetLastError();
}

hipError_t launch_ab_update(Ctx &c) {
  hipLaunchKernelGGL(k_ab_update, dim3(grid(c.pcount)), dim3(kBlock), 0, c.stream, c.pcount,
                     c.p.dt, c.infront, c.x.x, c.x.y, c.x.z, c.u.x, c.u.y, c.u.z, c.u1.x, c.u1.y,
                     c.u1.z, c.u2.x, c.u2.y, c.u2.z, c.u3.x, c.u3.y, c.u3.z);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !(c.p.flags & QV_FLAG_REMOVE_MEAN_POS) || c.nact == 0) return e;
  // the walk's part buffer (3 cap >= 2 kMeanBlocks doubles) holds the
  // partial sums; the walk is done with it once a stage has finished
  hipLaunchKernelGGL(k_mean_partial, dim3(kMeanBlocks), dim3(kBlock), 0, c.stream, c.pcount,
                     c.infront, c.x.x, c.x.y, c.walk_part);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_mean_apply, dim3(grid(c.pcount) < 1024 ? grid(c.pcount) : 1024),
                     dim3(kBlock), 0, c.stream, c.pcount, c.nact, c.walk_part, c.x.x, c.x.y);
  return hipGetLastError();
}

hipError_t launch_open_check(Ctx &c, int32_t *flag) {
  const int nopen = c.p.boundary == QV_BC_OPENX ? 1 : (c.p.boundary == QV_BC_OPENXY ? 2 : 0);
  if (nopen == 0 || c.pcount == 0) return hipSuccess;
  hipLaunchKernelGGL(k_open_check, dim3(grid(c.pcount)), dim3(kBlock), 0, c.stream, c.pcount,
                     c.p.box_size, nopen, c.infront, c.x.x, c.x.y, flag);
  return hipGetLastError();
}

hipError_t launch_topology_apply(Ctx &c, int n_new, int n_old, const int32_t *new_inf,
                                 const int32_t *new_beh, int32_t *mark, bool keep_x,
                                 int32_t *flag) {
  if (n_new <= 0) return hipSuccess;
  ClearArgs A;
  Vec3 vs[] = {c.u, c.u1, c.u2, c.uli, c.ubs, c.gi, c.gb, c.gii, c.gbb};
  int k = 0;
  for (const Vec3 &v : vs) {
    A.a[k++] = v.x;
    A.a[k++] = v.y;
    A.a[k++] = v.z;
  }
  hipLaunchKernelGGL(k_topo_mark, dim3(grid(n_new)), dim3(kBlock), 0, c.stream, n_new, n_old,
                     c.infront, new_inf, mark, A, c.u3.x, c.u3.y, c.u3.z, keep_x ? 1 : 0,
                     c.x.x, c.x.y, c.x.z, flag);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_topo_new_velocity, dim3(grid(n_new)), dim3(kBlock), 0, c.stream, n_new,
                     mark, new_inf, new_beh, c.u.x, c.u.y, c.u.z);
  return hipGetLastError();
}

hipError_t launch_clear_slots(Ctx &c, int m, const int32_t *d_list) {
  if (m <= 0) return hipSuccess;
  ClearArgs A;
  Vec3 vs[] = {c.u, c.u1, c.u2, c.uli, c.ubs, c.gi, c.gb, c.gii, c.gbb};
  int k = 0;
  for (const Vec3 &v : vs) {
    A.a[k++] = v.x;
    A.a[k++] = v.y;
    A.a[k++] = v.z;
  }
  hipLaunchKernelGGL(k_clear_slots, dim3(grid(m)), dim3(kBlock), 0, c.stream, m, d_list, A);
  return hipGetLastError();
}

hipError_t launch_rk3_update(Ctx &c, int stage) {
  hipLaunchKernelGGL(k_rk3_update, dim3(grid(c.pcount)), dim3(kBlock), 0, c.stream, c.pcount,
                     c.p.dt, stage, c.infront, c.x.x, c.x.y, c.x.z, c.u.x, c.u.y, c.u.z, c.u1.x,
                     c.u1.y, c.u1.z);
  return hipGetLastError();
}

hipError_t launch_enforce_periodic(Ctx &c) {
  int k0 = 3;
  if (c.p.boundary == QV_BC_PERIODIC) k0 = 0;
  else if (c.p.boundary == QV_BC_OPENX) k0 = 1;
  else if (c.p.boundary == QV_BC_OPENXY) k0 = 2;
  if (k0 == 3) return hipSuccess;
  hipLaunchKernelGGL(k_wrap, dim3(grid(c.pcount)), dim3(kBlock), 0, c.stream, c.pcount,
                     c.p.box_size, k0, c.infront, c.x.x, c.x.y, c.x.z);
  return hipGetLastError();
}

}  // namespace qv
