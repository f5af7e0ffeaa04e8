// qv_api.hip -- C ABI of libqvort_gpu.so: context, device buffers, the
// velocity stage, the AB step and the tree dump.  Declarations and the
// reference routine each entry point replaces: include/qvort_gpu.h.
#include <string.h>

#include <atomic>
#include <mutex>
#include <new>
#include <vector>

#include "qv_internal.h"

namespace qv {
namespace {

// Scratch of qv_vjp_device_noalloc: per slot the cotangent and the result
// (3 + 3 doubles, SoA) and the adjoint kernels' 24 partial sums, then the
// partial sums of their source / target chunks (9 doubles per slot per
// chunk, as qv_vjp_device uses them), for up to kNoAllocSlots slots: 5 Mi
// doubles (42 MB) hold that for every slot count up to the limit.  Each
// device the library runs on holds its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, not a device address: its address on a device comes
// from hipGetSymbolAddress with that device current (scratch_setup, called
// by qv_create) and only that address is passed to kernels.
constexpr int32_t kNoAllocSlots = 65536;
constexpr size_t kScratchDoubles = (size_t)5 << 20;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double *scratch_dev[kMaxDevices] = {};

// the address of `scratch` on `device` (current), looked up once per device
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return QV_ERR_HIP;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_dev[device]) return QV_OK;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return QV_ERR_HIP;
  scratch_dev[device] = static_cast<double *>(ptr);
  return QV_OK;
}

double *scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(scratch_mu);
  return scratch_dev[device];
}

int check_params(const qv_params *p) {
  if (!p) return QV_ERR_ARG;
  if (p->velocity != QV_VEL_BS && p->velocity != QV_VEL_TREE) return QV_ERR_ARG;
  if (p->boundary < QV_BC_OPEN || p->boundary > QV_BC_OPENXY) return QV_ERR_ARG;
  if (!(p->box_size > 0.)) return QV_ERR_ARG;  // initial.f90:155-156
  if (p->velocity == QV_VEL_TREE && !(p->tree_theta > 0.)) return QV_ERR_ARG;  // initial.f90:356-359
  if (p->deriv_order != 2 && p->deriv_order != 4) return QV_ERR_ARG;
  if (p->precision != QV_FP64 && p->precision != QV_FP32C) return QV_ERR_ARG;
  if (!(p->quant_circ > 0.) || !(p->corea > 0.)) return QV_ERR_ARG;
  if (p->flags & ~QV_FLAG_REMOVE_MEAN_POS) return QV_ERR_ARG;
  return QV_OK;
}

template <class T>
hipError_t dalloc(T *&p, size_t n) {
  (void)hipFree(p);
  p = nullptr;
  return hipMalloc(&p, (n ? n : 1) * sizeof(T));
}

hipError_t valloc(Vec3 &v, size_t n) {
  hipError_t e;
  if ((e = dalloc(v.x, n))) return e;
  if ((e = dalloc(v.y, n))) return e;
  return dalloc(v.z, n);
}

void vfree(Vec3 &v) {
  (void)hipFree(v.x); (void)hipFree(v.y); (void)hipFree(v.z);
  v = Vec3{};
}

hipError_t vzero(Vec3 &v, size_t n, hipStream_t s) {
  hipError_t e;
  if ((e = hipMemsetAsync(v.x, 0, n * sizeof(double), s))) return e;
  if ((e = hipMemsetAsync(v.y, 0, n * sizeof(double), s))) return e;
  return hipMemsetAsync(v.z, 0, n * sizeof(double), s);
}

#define TRY(expr)                              \
  do {                                         \
    hipError_t e_ = (expr);                    \
    if (e_ != hipSuccess) return QV_ERR_HIP;   \
  } while (0)

// slot buffers grow by a quarter at a time: pinsert's appended slots make
// pcount creep up from one upload to the next.  keep: the state arrays
// (positions, links, velocities, AB history, ghosts) keep their first
// pcount entries; the rest are scratch of a stage.
int reserve(Ctx &c, int32_t n, bool keep = false) {
  if (n <= c.cap) return QV_OK;
  // the buffers below are freed: the walk order's sort may still use some
  if (c.aux_stream) TRY(hipStreamSynchronize(c.aux_stream));
  TRY(hipStreamSynchronize(c.stream));
  const size_t cap = (size_t)n + (size_t)n / 4 + 1024;
  if (cap > (size_t)INT32_MAX) return QV_ERR_ARG;
  const size_t old = keep ? (size_t)c.pcount : 0;
  Vec3 *vs[] = {&c.x, &c.gi, &c.gb, &c.gii, &c.gbb, &c.u, &c.uli, &c.ubs,
                &c.u1, &c.u2, &c.u3};
  for (Vec3 *v : vs) {
    Vec3 nv;
    TRY(valloc(nv, cap));
    if (old) {
      TRY(hipMemcpy(nv.x, v->x, old * sizeof(double), hipMemcpyDeviceToDevice));
      TRY(hipMemcpy(nv.y, v->y, old * sizeof(double), hipMemcpyDeviceToDevice));
      TRY(hipMemcpy(nv.z, v->z, old * sizeof(double), hipMemcpyDeviceToDevice));
    }
    vfree(*v);
    *v = nv;
  }
  int32_t *links[2] = {c.infront, c.behind};
  for (int q = 0; q < 2; ++q) {
    int32_t *nl = nullptr;
    TRY(dalloc(nl, cap));
    if (old) TRY(hipMemcpy(nl, links[q], old * sizeof(int32_t), hipMemcpyDeviceToDevice));
    (void)hipFree(links[q]);
    links[q] = nl;
  }
  c.infront = links[0];
  c.behind = links[1];
  TRY(valloc(c.xs, cap));
  TRY(valloc(c.gs, cap));
  TRY(dalloc(c.stage, 3 * cap));
  TRY(dalloc(c.key0, cap));
  TRY(dalloc(c.key1, 2 * cap));
  TRY(dalloc(c.key0_s, cap));
  TRY(dalloc(c.slot, cap));
  TRY(dalloc(c.slot_s, cap));
  TRY(dalloc(c.walk_order, cap));
  TRY(dalloc(c.walk_part, 3 * cap));
  TRY(dalloc(c.nnode_of, cap + 1));
  TRY(dalloc(c.node_off, cap + 1));
  c.cap = (int32_t)cap;
  return QV_OK;
}

// device (pcount,3) SoA -> host AoS through the staging buffer
int download_vec(Ctx &c, Vec3 v, double *host) {
  if (!host) return QV_OK;
  TRY(launch_soa_to_aos(c, v, c.stage));
  TRY(hipMemcpyAsync(host, c.stage, (size_t)3 * c.pcount * sizeof(double),
                     hipMemcpyDeviceToHost, c.stream));
  return QV_OK;
}

int upload_vec(Ctx &c, const double *host, Vec3 v) {
  if (!host) return QV_OK;
  TRY(hipMemcpyAsync(c.stage, host, (size_t)3 * c.pcount * sizeof(double),
                     hipMemcpyHostToDevice, c.stream));
  TRY(launch_aos_to_soa(c, c.stage, v));
  return QV_OK;
}

// the launches of one velocity stage (ghostp + LIA, tree build, walk or
// direct sum); the error flags (counters[1]) are not reset here, so that the
// stages of one qv_step call accumulate them
int velocity_stage_launch(Ctx &c) {
  if (c.pcount == 0) return QV_ERR_STATE;
  TRY(hipEventRecord(c.ev[0], c.stream));
  TRY(hipMemsetAsync(c.counters + 2, 0, 6 * sizeof(int32_t), c.stream));
  TRY(hipMemsetAsync(c.evals, 0, 4 * sizeof(int64_t), c.stream));
  TRY(launch_ghost_lia(c));
  c.ghosts_valid = true;
  TRY(hipEventRecord(c.ev[1], c.stream));
  int rc = QV_OK;
  // construct_tree runs whenever tree_theta > 0 (run.f90:43-45), also with
  // velocity='BS', where only the neighbour search uses it
  c.tree_valid = false;
  if (c.p.tree_theta > 0. && c.nact > 0) {
    rc = tree_build(c);
    if (rc != QV_OK) return rc;
  } else {
    c.tree.nnodes = 0;
  }
  TRY(hipEventRecord(c.ev[2], c.stream));
  if (c.p.velocity == QV_VEL_TREE) {
    if (c.nact > 0) TRY(launch_walk(c));
  } else {
    const int64_t need =
        6 * (int64_t)direct_chunks(c.pcount, c.p.precision == QV_FP32C) * c.pcount;
    if (need > c.direct_part_doubles) {
      TRY(hipStreamSynchronize(c.stream));
      TRY(dalloc(c.direct_part, (size_t)need));
      c.direct_part_doubles = need;
    }
    TRY(launch_direct(c));
  }
  TRY(hipEventRecord(c.ev[3], c.stream));
  c.tree_valid = c.p.tree_theta > 0.;
  c.tree_pcount = c.pcount;
  return QV_OK;
}

// after the last stage of a call: one copy of the counters and flags, the
// timings of that stage, the singular-leaf error of any stage since the
// flags were reset
int velocity_stage_finish(Ctx &c) {
  TRY(hipMemcpyAsync(c.h_counters, c.counters, 4 * sizeof(int32_t), hipMemcpyDeviceToHost,
                     c.stream));
  TRY(hipMemcpyAsync(c.h_evals, c.evals, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  TRY(hipStreamSynchronize(c.stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c.ev[0], c.ev[1]);
  c.timings.ghost_lia_ms = ms;
  (void)hipEventElapsedTime(&ms, c.ev[1], c.ev[2]);
  c.timings.build_ms = ms;
  (void)hipEventElapsedTime(&ms, c.ev[2], c.ev[3]);
  c.timings.walk_ms = ms;
  (void)hipEventElapsedTime(&ms, c.ev[0], c.ev[3]);
  c.timings.total_ms = ms;
  c.timings.nodes = c.p.tree_theta > 0. ? c.tree.nnodes : 0;
  if (c.p.velocity == QV_VEL_TREE) {
    // eval_counter starts at 2N (tree.f90:77-78) and counts every leaf and
    // accepted-cell interaction
    c.timings.evals = 2 * (int64_t)c.nact + c.h_evals[0];
    c.timings.visits = c.h_evals[1];
    c.timings.leaf_evals = c.h_evals[2];
    c.timings.cell_evals = c.h_evals[0] - c.h_evals[2];
    c.timings.wave_steps = c.h_evals[3];
  } else {
    const int64_t S = make_shells(c.p).n;
    c.timings.evals = S * (int64_t)c.nact * (int64_t)c.nact;
    c.timings.visits = 0;
    c.timings.leaf_evals = c.timings.evals;
    c.timings.cell_evals = 0;
    c.timings.wave_steps = 0;
  }
  if (c.h_counters[1] & kFlagSingular) return QV_ERR_SINGULAR;
  c.have_velocity = true;
  return QV_OK;
}

int velocity_stage(Ctx &c) {
  TRY(hipMemsetAsync(c.counters + 1, 0, sizeof(int32_t), c.stream));
  int rc = velocity_stage_launch(c);
  if (rc) return rc;
  return velocity_stage_finish(c);
}

struct DumpCtx {
  const std::vector<int8_t> &level, &digit;
  const std::vector<int32_t> &count;
  const std::vector<double> &corner, &mp;
  const std::vector<double2> &geo;
  const std::vector<NodeRec> &rec;
  const std::vector<int2> &meta;
  double *out;
  int64_t max_nodes, k;
};

void dump_emit(DumpCtx &D, int lev, int pc, int leafj, const double *pos, double w,
               const double *cent, const double *circ, const double *C /* row-major */) {
  if (D.out && D.k < D.max_nodes) {
    double *o = D.out + 22 * D.k;
    o[0] = lev; o[1] = pc; o[2] = leafj;
    o[3] = pos[0]; o[4] = pos[1]; o[5] = pos[2]; o[6] = w;
    for (int q = 0; q < 3; ++q) { o[7 + q] = cent[q]; o[10 + q] = circ[q]; }
    for (int kk = 0; kk < 3; ++kk)
      for (int m = 0; m < 3; ++m) o[13 + 3 * kk + m] = C[3 * m + kk];
  }
  ++D.k;
}

// children of a split node n are n+1, skip(n+1), ... before skip(n); the
// recursion depth is at most 33
void dump_rec(DumpCtx &D, int64_t n) {
  static const double zero[9] = {0};
  const double cent[3] = {D.rec[n].cx, D.rec[n].cy, D.rec[n].cz};
  const double *pos = &D.corner[3 * n];
  const double w = D.geo[n].x;
  if (D.meta[n].y >= 0) {
    // leaf: circ = (0 + ghosti) - x, C = x (g-x) - x (g-x) = 0 (tree.f90:160-178)
    const double *g = &D.mp[12 * n];
    const double circ[3] = {g[0] - cent[0], g[1] - cent[1], g[2] - cent[2]};
    dump_emit(D, D.level[n], D.count[n], D.meta[n].y + 1, pos, w, cent, circ, zero);
    return;
  }
  dump_emit(D, D.level[n], D.count[n], 0, pos, w, cent, &D.mp[12 * n], &D.mp[12 * n + 3]);
  if (D.count[n] <= 1) return;
  const int64_t end = D.meta[n].x;
  int64_t ch = n + 1;
  for (int d = 0; d < 8; ++d) {
    if (ch < end && D.digit[ch] == d && D.level[ch] == D.level[n] + 1) {
      dump_rec(D, ch);
      ch = D.meta[ch].x;
    } else {
      const int b[3] = {d & 1, (d >> 2) & 1, (d >> 1) & 1};
      double cp[3];
      for (int a = 0; a < 3; ++a) cp[a] = b[a] ? pos[a] + (w / 2) : pos[a];
      dump_emit(D, D.level[n] + 1, 0, 0, cp, w / 2., zero, zero, zero);
    }
  }
}

int adj_reserve(Ctx &c) {
  int len = 0;
  const int64_t need = 9 * (int64_t)vjp_chunks(c.pcount, len) * c.pcount;
  if (need > c.adj_part_doubles) {
    TRY(hipStreamSynchronize(c.stream));
    TRY(dalloc(c.adj_part, (size_t)need));
    c.adj_part_doubles = need;
  }
  if (c.adj_cap >= c.pcount) return QV_OK;
  TRY(dalloc(c.adj, (size_t)24 * c.pcount));
  TRY(valloc(c.aw, c.pcount));
  TRY(valloc(c.ar, c.pcount));
  c.adj_cap = c.pcount;
  return QV_OK;
}

// derivative of the direct-sum velocity at the device positions; `v` and
// the result are AoS (pcount,3) device buffers
int derivative(Ctx &c, const double *d_in, double *d_out, bool vjp) {
  if (c.pcount == 0) return QV_ERR_STATE;
  if (c.p.velocity != QV_VEL_BS || !c.links_consistent) return QV_ERR_ARG;
  int rc = adj_reserve(c);
  if (rc) return rc;
  TRY(launch_ghost_lia(c));
  TRY(launch_aos_to_soa(c, d_in, c.aw));
  if (vjp) TRY(launch_vjp(c, c.aw, c.ar, c.p.precision == QV_FP32C ? 1 : 0, c.adj, c.adj_part,
                          c.adj_part_doubles));
  else TRY(launch_jvp(c, c.aw, c.ar));
  TRY(launch_soa_to_aos(c, c.ar, d_out));
  return QV_OK;
}

int host_derivative(qv_ctx *ctx, const double *in, double *out, bool vjp) {
  if (!ctx || !in || !out) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  int rc = adj_reserve(*c);
  if (rc) return rc;
  // the staging buffer carries the input; the result comes back through it
  TRY(hipMemcpyAsync(c->stage, in, (size_t)3 * c->pcount * sizeof(double),
                     hipMemcpyHostToDevice, c->stream));
  rc = derivative(*c, c->stage, c->stage, vjp);
  if (rc) return rc;
  TRY(hipMemcpyAsync(out, c->stage, (size_t)3 * c->pcount * sizeof(double),
                     hipMemcpyDeviceToHost, c->stream));
  TRY(hipStreamSynchronize(c->stream));
  return QV_OK;
}

// the boundary step of run.f90:82-90 on the device state.  Periodic: the
// wrap.  openx / openxy: a device check for points beyond an open face
// first; without any, the wrap of the periodic axes; with some, the state's
// positions and links go to the host, boundary_step_host runs in slot order
// and they come back, and the cleared slots' other arrays are zeroed.
int boundary_step(Ctx &c) {
  if (c.p.boundary != QV_BC_OPENX && c.p.boundary != QV_BC_OPENXY) {
    TRY(launch_enforce_periodic(c));
    return QV_OK;
  }
  int32_t *flag = c.counters + 6;
  TRY(hipMemsetAsync(flag, 0, sizeof(int32_t), c.stream));
  TRY(launch_open_check(c, flag));
  TRY(hipMemcpyAsync(c.h_counters + 6, flag, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  TRY(hipStreamSynchronize(c.stream));
  if (c.h_counters[6] == 0) {
    TRY(launch_enforce_periodic(c));
    return QV_OK;
  }
  const size_t n = (size_t)c.pcount;
  std::vector<double> hx(3 * n);
  std::vector<int32_t> hinf(n), hbeh(n), cleared(n);
  int rc = download_vec(c, c.x, hx.data());
  if (rc) return rc;
  TRY(hipMemcpyAsync(hinf.data(), c.infront, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                     c.stream));
  TRY(hipMemcpyAsync(hbeh.data(), c.behind, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                     c.stream));
  TRY(hipStreamSynchronize(c.stream));
  int32_t ncleared = 0;
  boundary_step_host(c.p, c.pcount, hx.data(), hinf.data(), hbeh.data(), c.boundary_count,
                     c.boundary_length, cleared.data(), ncleared);
  rc = upload_vec(c, hx.data(), c.x);
  if (rc) return rc;
  TRY(hipMemcpyAsync(c.infront, hinf.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                     c.stream));
  TRY(hipMemcpyAsync(c.behind, hbeh.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                     c.stream));
  // the build's slot buffer (cap >= pcount ints) carries the list; no tree
  // build runs between here and the next velocity stage
  TRY(hipMemcpyAsync(c.slot, cleared.data(), (size_t)ncleared * sizeof(int32_t),
                     hipMemcpyHostToDevice, c.stream));
  TRY(launch_clear_slots(c, ncleared, c.slot));
  TRY(hipStreamSynchronize(c.stream));
  c.nact -= ncleared;
  return QV_OK;
}

int set_positions_device(Ctx &c, const double *d_x) {
  if (!d_x) return QV_OK;
  TRY(launch_aos_to_soa(c, d_x, c.x));
  return QV_OK;
}

}  // namespace
}  // namespace qv

using namespace qv;

extern "C" {

int qv_create(const qv_params *params, qv_ctx **out) {
  if (!out) return QV_ERR_ARG;
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return QV_ERR_NO_DEVICE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return QV_ERR_HIP;
  // the no-allocation entry point's scratch on this device
  rc = scratch_setup(dev);
  if (rc) return rc;
  Ctx *c = new (std::nothrow) Ctx();
  if (!c) return QV_ERR_HIP;
  c->p = *params;
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return QV_ERR_HIP;
  }
  c->stream = c->own_stream;
  for (auto &e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete c;
      return QV_ERR_HIP;
    }
  }
  if (hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->aux_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->aux_ev[1], hipEventDisableTiming) != hipSuccess) {
    qv_destroy(reinterpret_cast<qv_ctx *>(c));
    return QV_ERR_HIP;
  }
  if (hipMalloc(&c->counters, 8 * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->evals, 4 * sizeof(int64_t)) != hipSuccess ||
      hipHostMalloc(&c->h_counters, 8 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&c->h_evals, 4 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess) {
    qv_destroy(reinterpret_cast<qv_ctx *>(c));
    return QV_ERR_HIP;
  }
  *out = reinterpret_cast<qv_ctx *>(c);
  return QV_OK;
}

int qv_destroy(qv_ctx *ctx) {
  if (!ctx) return QV_OK;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->own_stream && c->own_stream != c->stream) (void)hipStreamSynchronize(c->own_stream);
  if (c->aux_stream) (void)hipStreamSynchronize(c->aux_stream);
  Vec3 *vs[] = {&c->x, &c->gi, &c->gb, &c->gii, &c->gbb, &c->u, &c->uli, &c->ubs,
                &c->u1, &c->u2, &c->u3, &c->xs, &c->gs, &c->aw, &c->ar};
  for (Vec3 *v : vs) vfree(*v);
  void *ps[] = {c->infront, c->behind, c->stage, c->key0, c->key1, c->key0_s, c->slot,
                c->slot_s, c->walk_order, c->walk_part, c->nnode_of, c->node_off, c->sort_tmp,
                c->counters, c->evals, c->direct_part, c->adj_part,
                c->big_buf, c->partial, c->adj,
                c->tree.rec, c->tree.meta, c->tree.mp, c->tree.mp32, c->tree.geo,
                c->tree.first, c->tree.count, c->tree.level, c->tree.digit, c->tree.corner,
                c->tree.linf, c->tree.tplain, c->tree.tclose, c->close_i, c->close_d, c->close_order, c->probe_buf, c->probe_part,
                c->aux_tmp};
  for (void *p : ps) (void)hipFree(p);
  if (c->h_counters) (void)hipHostFree(c->h_counters);
  if (c->h_evals) (void)hipHostFree(c->h_evals);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  if (c->aux_stream) (void)hipStreamDestroy(c->aux_stream);
  for (auto &e : c->aux_ev)
    if (e) (void)hipEventDestroy(e);
  delete c;
  return QV_OK;
}

int qv_set_params(qv_ctx *ctx, const qv_params *params) {
  if (!ctx) return QV_ERR_ARG;
  int rc = check_params(params);
  if (rc) return rc;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  c->p = *params;
  c->have_velocity = false;
  c->tree.plain_valid = false;  // the probe thresholds depend on tree_theta
  return QV_OK;
}

int qv_upload(qv_ctx *ctx, int32_t pcount, const double *x, const int32_t *infront,
              const int32_t *behind) {
  if (!ctx || !x || !infront || !behind || pcount < 2) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  // links must stay inside f(:) and point at occupied slots
  int32_t nact = 0;
  for (int32_t i = 0; i < pcount; ++i) {
    const int32_t a = infront[i], b = behind[i];
    if (a == 0) continue;
    if (a < 1 || a > pcount || b < 1 || b > pcount) return QV_ERR_ARG;
    if (infront[a - 1] == 0 || infront[b - 1] == 0) return QV_ERR_ARG;
    ++nact;
  }
  bool consistent = true;
  for (int32_t i = 0; i < pcount && consistent; ++i)
    if (infront[i] != 0 && (behind[infront[i] - 1] != i + 1 || infront[behind[i] - 1] != i + 1))
      consistent = false;
  if (hipSetDevice(c->device) != hipSuccess) return QV_ERR_HIP;
  int rc = reserve(*c, pcount);
  if (rc) return rc;
  c->pcount = pcount;
  c->nact = nact;
  c->have_velocity = false;
  c->ghosts_valid = false;
  c->links_consistent = consistent;
  TRY(hipMemcpyAsync(c->infront, infront, pcount * sizeof(int32_t), hipMemcpyHostToDevice,
                     c->stream));
  TRY(hipMemcpyAsync(c->behind, behind, pcount * sizeof(int32_t), hipMemcpyHostToDevice,
                     c->stream));
  rc = upload_vec(*c, x, c->x);
  if (rc) return rc;
  TRY(vzero(c->u1, pcount, c->stream));
  TRY(vzero(c->u2, pcount, c->stream));
  TRY(vzero(c->u3, pcount, c->stream));
  TRY(vzero(c->u, pcount, c->stream));
  TRY(hipStreamSynchronize(c->stream));
  return QV_OK;
}

int qv_upload_topology(qv_ctx *ctx, int32_t pcount, const double *x, const int32_t *infront,
                       const int32_t *behind) {
  if (!ctx || !infront || !behind) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  if (pcount < c->pcount) return QV_ERR_ARG;  // slots are never given back
  std::atomic<int32_t> nact_sum{0};
  std::atomic<bool> bad{false}, inconsistent{false};
  par_for(pcount, [&](int lo, int hi) {
    int32_t k = 0;
    bool ok = true, cons = true;
    for (int32_t i = lo; i < hi && ok; ++i) {
      const int32_t a = infront[i], b = behind[i];
      if (a == 0) continue;
      if (a < 1 || a > pcount || b < 1 || b > pcount) ok = false;
      else if (infront[a - 1] == 0 || infront[b - 1] == 0) ok = false;
      else if (behind[a - 1] != i + 1 || infront[b - 1] != i + 1) cons = false;
      ++k;
    }
    nact_sum += k;
    if (!ok) bad.store(true);
    if (!cons) inconsistent.store(true);
  });
  if (bad.load()) return QV_ERR_ARG;
  const int32_t nact = nact_sum.load();
  const bool consistent = !inconsistent.load();
  if (hipSetDevice(c->device) != hipSuccess) return QV_ERR_HIP;
  const int32_t n_old = c->pcount;
  int rc = reserve(*c, pcount, true);
  if (rc) return rc;
  c->pcount = pcount;
  // the new links go to the build's slot buffers (cap ints each), the
  // marks to its node-count buffer; no tree build runs until the next stage
  TRY(hipMemcpyAsync(c->slot, infront, pcount * sizeof(int32_t), hipMemcpyHostToDevice,
                     c->stream));
  TRY(hipMemcpyAsync(c->slot_s, behind, pcount * sizeof(int32_t), hipMemcpyHostToDevice,
                     c->stream));
  // counters[7]: a slot became a point while the positions were kept
  int32_t *flag = c->counters + 7;
  if (!x) TRY(hipMemsetAsync(flag, 0, sizeof(int32_t), c->stream));
  TRY(launch_topology_apply(*c, pcount, n_old, c->slot, c->slot_s, c->nnode_of, !x, flag));
  TRY(hipMemcpyAsync(c->infront, c->slot, pcount * sizeof(int32_t), hipMemcpyDeviceToDevice,
                     c->stream));
  TRY(hipMemcpyAsync(c->behind, c->slot_s, pcount * sizeof(int32_t), hipMemcpyDeviceToDevice,
                     c->stream));
  if (x) {
    rc = upload_vec(*c, x, c->x);
    if (rc) return rc;
  } else {
    TRY(hipMemcpyAsync(c->h_counters + 7, flag, sizeof(int32_t), hipMemcpyDeviceToHost,
                       c->stream));
  }
  TRY(hipStreamSynchronize(c->stream));
  if (!x && c->h_counters[7] != 0) return QV_ERR_ARG;
  c->nact = nact;
  c->links_consistent = consistent;
  c->have_velocity = false;
  c->ghosts_valid = false;
  return QV_OK;
}

int qv_upload_history(qv_ctx *ctx, const double *u1, const double *u2, const double *u3) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  int rc;
  if ((rc = upload_vec(*c, u1, c->u1))) return rc;
  if ((rc = upload_vec(*c, u2, c->u2))) return rc;
  if ((rc = upload_vec(*c, u3, c->u3))) return rc;
  TRY(hipStreamSynchronize(c->stream));
  return QV_OK;
}

int qv_velocity(qv_ctx *ctx, double *u, double *u_s_li, double *u_s_bs, int64_t *evals) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (hipSetDevice(c->device) != hipSuccess) return QV_ERR_HIP;
  int rc = velocity_stage(*c);
  if (rc) return rc;
  if ((rc = download_vec(*c, c->u, u))) return rc;
  if (u) TRY(hipStreamSynchronize(c->stream));
  if ((rc = download_vec(*c, c->uli, u_s_li))) return rc;
  if (u_s_li) TRY(hipStreamSynchronize(c->stream));
  if ((rc = download_vec(*c, c->ubs, u_s_bs))) return rc;
  TRY(hipStreamSynchronize(c->stream));
  if (evals) *evals = c->timings.evals;
  return QV_OK;
}

int qv_ab_step(qv_ctx *ctx) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c->have_velocity) return QV_ERR_STATE;
  TRY(launch_ab_update(*c));
  TRY(hipStreamSynchronize(c->stream));
  c->have_velocity = false;
  return QV_OK;
}

int qv_enforce_periodic(qv_ctx *ctx) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  const int rc = boundary_step(*c);
  if (rc) return rc;
  TRY(hipStreamSynchronize(c->stream));
  return QV_OK;
}

int qv_boundary_removed(qv_ctx *ctx, int64_t *count, double *length) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (count) *count = c->boundary_count;
  if (length) *length = c->boundary_length;
  return QV_OK;
}

int qv_step(qv_ctx *ctx, int32_t nsteps) {
  if (!ctx || nsteps < 0) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  // device-resident: the host waits once per stage only for the tree's node
  // count (the node arrays are sized by it) and, in openx / openxy boxes,
  // for the open-face check of the boundary step; the counters and the
  // error flags are read once, after the last step
  if (nsteps == 0) return QV_OK;
  TRY(hipMemsetAsync(c->counters + 1, 0, sizeof(int32_t), c->stream));
  for (int32_t k = 0; k < nsteps; ++k) {
    int rc = velocity_stage_launch(*c);
    if (rc) return rc;
    TRY(launch_ab_update(*c));
    if ((rc = boundary_step(*c))) return rc;
  }
  const int rc = velocity_stage_finish(*c);
  c->have_velocity = false;
  return rc;
}

int qv_download(qv_ctx *ctx, double *x, double *u, double *u1, double *u2, double *u3) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  double *hs[] = {x, u, u1, u2, u3};
  Vec3 vs[] = {c->x, c->u, c->u1, c->u2, c->u3};
  for (int k = 0; k < 5; ++k) {
    if (!hs[k]) continue;
    int rc = download_vec(*c, vs[k], hs[k]);
    if (rc) return rc;
    TRY(hipStreamSynchronize(c->stream));
  }
  return QV_OK;
}

int qv_download_links(qv_ctx *ctx, int32_t *infront, int32_t *behind) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  const size_t n = (size_t)c->pcount;
  if (infront)
    TRY(hipMemcpyAsync(infront, c->infront, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                       c->stream));
  if (behind)
    TRY(hipMemcpyAsync(behind, c->behind, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                       c->stream));
  TRY(hipStreamSynchronize(c->stream));
  return QV_OK;
}

int qv_host_alloc(size_t bytes, void **ptr) {
  if (!ptr) return QV_ERR_ARG;
  *ptr = nullptr;
  if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    *ptr = nullptr;
    return QV_ERR_HIP;
  }
  return QV_OK;
}

int qv_host_free(void *ptr) {
  if (!ptr) return QV_OK;
  return hipHostFree(ptr) == hipSuccess ? QV_OK : QV_ERR_HIP;
}

int qv_download_ghosts(qv_ctx *ctx, double *gi, double *gb, double *gii, double *gbb) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0 || !c->ghosts_valid) return QV_ERR_STATE;
  if (hipSetDevice(c->device) != hipSuccess) return QV_ERR_HIP;
  double *hs[] = {gi, gb, gii, gbb};
  Vec3 vs[] = {c->gi, c->gb, c->gii, c->gbb};
  for (int k = 0; k < 4; ++k) {
    if (!hs[k]) continue;
    int rc = download_vec(*c, vs[k], hs[k]);
    if (rc) return rc;
    TRY(hipStreamSynchronize(c->stream));
  }
  return QV_OK;
}

int qv_rk3_step(qv_ctx *ctx) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (hipSetDevice(c->device) != hipSuccess) return QV_ERR_HIP;
  for (int stage = 0; stage < 3; ++stage) {
    const int rc = velocity_stage(*c);
    if (rc) return rc;
    TRY(launch_rk3_update(*c, stage));
  }
  TRY(hipStreamSynchronize(c->stream));
  return QV_OK;
}

int qv_closest(qv_ctx *ctx, int32_t *closest, double *closestd) {
  if (!ctx || !closest || !closestd) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  const bool tree = c->p.tree_theta > 0.;
  // the tree must exist and its leaves' slots must still be slots
  if (tree && (!c->tree_valid || c->pcount < c->tree_pcount)) return QV_ERR_STATE;
  if (hipSetDevice(c->device) != hipSuccess) return QV_ERR_HIP;
  if (c->close_cap < c->pcount) {
    TRY(hipStreamSynchronize(c->stream));
    const int32_t cap = c->cap;  // the slot buffers' capacity (>= pcount)
    TRY(dalloc(c->close_i, cap));
    TRY(dalloc(c->close_d, cap));
    TRY(dalloc(c->close_order, cap));
    c->close_cap = cap;
  }
  const size_t n = c->pcount;
  TRY(hipMemcpyAsync(c->close_i, closest, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  // pclose_tree sets closestd = 100 for every slot first (tree.f90:224-225), and
  // so does the tree search; pclose keeps the empty slots' values
  if (!tree)
    TRY(hipMemcpyAsync(c->close_d, closestd, n * sizeof(double), hipMemcpyHostToDevice,
                       c->stream));
  if (tree) {
    // the search's slot order; the slot buffers hold cap >= pcount entries
    int rc = join_aux(*c);
    if (!rc) rc = hilbert_order(*c, c->pcount, c->x, c->infront, c->close_order, c->stream,
                                c->sort_tmp, c->sort_tmp_bytes);
    if (rc) return rc;
  }
  TRY(launch_closest(*c, tree));
  TRY(hipMemcpyAsync(closest, c->close_i, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  TRY(hipMemcpyAsync(closestd, c->close_d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TRY(hipStreamSynchronize(c->stream));
  return QV_OK;
}

int qv_probe_velocity(qv_ctx *ctx, int32_t nprobe, const double *y, double *u, double *grad,
                      int32_t flags) {
  if (!ctx || nprobe < 0 || (nprobe > 0 && (!y || !u))) return QV_ERR_ARG;
  if (flags & ~(QV_PROBE_GHOSTP | QV_PROBE_REFERENCE_SCALE | QV_PROBE_EMPTY_CORE))
    return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  const bool tree = c->p.velocity == QV_VEL_TREE;
  // velocity_quasip has no Tree form (hamiltonian.f90:49-51, 85-86): the
  // gradient and the empty core are BS only
  if (tree && (grad || (flags & QV_PROBE_EMPTY_CORE))) return QV_ERR_ARG;
  if (hipSetDevice(c->device) != hipSuccess) return QV_ERR_HIP;
  if (flags & QV_PROBE_GHOSTP) {
    TRY(launch_ghost_lia(*c));
    c->ghosts_valid = true;
  }
  if (tree ? !c->tree_valid : !c->ghosts_valid) return QV_ERR_STATE;
  if (nprobe == 0) return QV_OK;
  if (c->probe_cap < nprobe) {
    TRY(hipStreamSynchronize(c->stream));
    TRY(dalloc(c->probe_buf, (size_t)18 * nprobe));
    c->probe_cap = nprobe;
  }
  double *d_y = c->probe_buf, *d_u = d_y + 3 * (size_t)nprobe, *d_g = d_u + 3 * (size_t)nprobe;
  TRY(hipMemcpyAsync(d_y, y, 3 * (size_t)nprobe * sizeof(double), hipMemcpyHostToDevice,
                     c->stream));
  TRY(launch_probe(*c, nprobe, d_y, d_u, grad ? d_g : nullptr, d_g + 9 * (size_t)nprobe, tree,
                   (flags & QV_PROBE_REFERENCE_SCALE) != 0,
                   (flags & QV_PROBE_EMPTY_CORE) != 0));
  TRY(hipMemcpyAsync(u, d_u, 3 * (size_t)nprobe * sizeof(double), hipMemcpyDeviceToHost,
                     c->stream));
  if (grad)
    TRY(hipMemcpyAsync(grad, d_g, 9 * (size_t)nprobe * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream));
  TRY(hipStreamSynchronize(c->stream));
  return QV_OK;
}

int qv_tree_dump(qv_ctx *ctx, double *out, int64_t max_nodes, int64_t *nnodes) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c->tree_valid || c->tree.nnodes == 0) return QV_ERR_STATE;
  const int64_t nn = c->tree.nnodes;
  std::vector<int8_t> level(nn), digit(nn);
  std::vector<int32_t> count(nn);
  std::vector<double> corner(3 * nn), mp(12 * nn);
  std::vector<double2> geo(nn);
  std::vector<NodeRec> rec(nn);
  std::vector<int2> meta(nn);
  TRY(hipMemcpy(level.data(), c->tree.level, nn, hipMemcpyDeviceToHost));
  TRY(hipMemcpy(digit.data(), c->tree.digit, nn, hipMemcpyDeviceToHost));
  TRY(hipMemcpy(count.data(), c->tree.count, nn * 4, hipMemcpyDeviceToHost));
  TRY(hipMemcpy(corner.data(), c->tree.corner, nn * 24, hipMemcpyDeviceToHost));
  TRY(hipMemcpy(mp.data(), c->tree.mp, nn * 96, hipMemcpyDeviceToHost));
  TRY(hipMemcpy(geo.data(), c->tree.geo, nn * 16, hipMemcpyDeviceToHost));
  TRY(hipMemcpy(rec.data(), c->tree.rec, nn * sizeof(NodeRec), hipMemcpyDeviceToHost));
  TRY(hipMemcpy(meta.data(), c->tree.meta, nn * 8, hipMemcpyDeviceToHost));
  // expand to the reference's node list: every split cell has 8 children,
  // empty ones included, in DFS order (tree.f90:89-114)
  DumpCtx D{level, digit, count, corner, mp, geo, rec, meta, out, max_nodes, 0};
  dump_rec(D, 0);
  const int64_t k = D.k;
  if (nnodes) *nnodes = k;
  return QV_OK;
}

int qv_vjp(qv_ctx *ctx, const double *w, double *xbar) {
  return host_derivative(ctx, w, xbar, true);
}

int qv_jvp(qv_ctx *ctx, const double *v, double *jv) {
  return host_derivative(ctx, v, jv, false);
}

int qv_set_stream(qv_ctx *ctx, void *stream) {
  if (!ctx) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  // used as given: NULL is HIP's null stream, which orders with the
  // caller's other blocking streams
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (s != c->stream) TRY(hipStreamSynchronize(c->stream));
  c->stream = s;
  return QV_OK;
}

int qv_velocity_device(qv_ctx *ctx, const double *d_x, double *d_u) {
  if (!ctx || !d_u) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  int rc = set_positions_device(*c, d_x);
  if (rc) return rc;
  rc = velocity_stage(*c);
  if (rc) return rc;
  TRY(launch_soa_to_aos(*c, c->u, d_u));
  return QV_OK;
}

int qv_vjp_device(qv_ctx *ctx, const double *d_x, const double *d_w, double *d_xbar) {
  if (!ctx || !d_w || !d_xbar) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  int rc = set_positions_device(*c, d_x);
  if (rc) return rc;
  return derivative(*c, d_w, d_xbar, true);
}

int qv_vjp_device_noalloc(qv_ctx *ctx, const double *d_x, const double *d_w,
                          double *d_xbar) {
  if (!ctx || !d_w || !d_xbar) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  int len = 0;
  const int64_t part = 9 * (int64_t)vjp_chunks(c->pcount, len) * c->pcount;
  if (c->pcount > kNoAllocSlots || 30 * (int64_t)c->pcount + part > (int64_t)kScratchDoubles)
    return QV_ERR_SCRATCH;
  if (c->p.velocity != QV_VEL_BS || !c->links_consistent) return QV_ERR_ARG;
  double *base = scratch_of(c->device);
  if (!base) return QV_ERR_STATE;
  const size_t n = (size_t)c->pcount;
  const Vec3 w{base, base + n, base + 2 * n};
  const Vec3 r{base + 3 * n, base + 4 * n, base + 5 * n};
  double *adj = base + 6 * n;  // 24 n doubles, then the chunks' partial sums
  int rc = set_positions_device(*c, d_x);
  if (rc) return rc;
  TRY(launch_ghost_lia(*c));
  TRY(launch_aos_to_soa(*c, d_w, w));
  TRY(launch_vjp(*c, w, r, c->p.precision == QV_FP32C ? 1 : 0, adj, base + 30 * n,
                 (int64_t)kScratchDoubles - 30 * (int64_t)n));
  TRY(launch_soa_to_aos(*c, r, d_xbar));
  return QV_OK;
}

int qv_jvp_device(qv_ctx *ctx, const double *d_x, const double *d_v, double *d_jv) {
  if (!ctx || !d_v || !d_jv) return QV_ERR_ARG;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->pcount == 0) return QV_ERR_STATE;
  int rc = set_positions_device(*c, d_x);
  if (rc) return rc;
  return derivative(*c, d_v, d_jv, false);
}

int qv_get_timings(qv_ctx *ctx, qv_timings *t) {
  if (!ctx || !t) return QV_ERR_ARG;
  *t = reinterpret_cast<Ctx *>(ctx)->timings;
  return QV_OK;
}

const char *qv_status_string(int status) {
  switch (status) {
    case QV_OK: return "ok";
    case QV_ERR_ARG: return "invalid argument";
    case QV_ERR_HIP: return "HIP runtime error";
    case QV_ERR_SINGULAR:
      return "tree.mod:tree_walk: singularity in BS (tree) velocity field - this is normally "
             "caused by having recon_shots too large";
    case QV_ERR_TREE_DEPTH: return "points closer than box_size/2^32: octree deeper than 32 levels";
    case QV_ERR_STATE: return "call out of order";
    case QV_ERR_NO_DEVICE: return "no HIP device";
    case QV_ERR_LINE_EMPTY: return "line.mod:pinsert: vortex line length is 0, run over";
    case QV_ERR_CAPACITY: return "line operations need more slots than the arrays hold";
    case QV_ERR_SCRATCH:
      return "qv_vjp_device_noalloc: more than 65536 slots, the size of its static scratch";
    default: return "unknown status";
  }
}

}  // extern "C"
