/*
 * qvort_gpu.h -- C ABI of the MI355X (gfx950) filament-velocity library.
 *
 * Drop-in for Qvort's per-timestep velocity hot path.  The reference has no
 * FFI: the hot path is a set of Fortran module procedures acting on global
 * state (cdata.f90:20-41).  Each entry point below replaces one of them and
 * is what a Fortran `bind(C)` interface (qvort_amd/fortran/qv_gpu.f90) binds:
 *
 *   qv_upload            the f(:)%x / infront / behind arrays the hot path
 *                        reads (cdata.f90:20-41), 1-based links, 0 = empty
 *   qv_velocity          ghostp (periodic.f90:8-17) + construct_tree
 *                        (tree.f90:44-80) + the velocity loop of pmotion
 *                        (timestep.f90:41-47 -> calc_velocity :180-370,
 *                        tree_walk tree.f90:336-421, biot_savart
 *                        timestep.f90:462-507)
 *   qv_ab_step           the Adams-Bashforth ladder of pmotion
 *                        (timestep.f90:48-71)
 *   qv_enforce_periodic  enforce_periodic (periodic.f90:188-213)
 *   qv_tree_dump         the tree as construct_tree builds it, node by node
 *                        in tree_walk's DFS order (tree.f90:85-125, 416-419)
 *   qv_destroy           empty_tree + deallocate(vtree) (run.f90:132-138)
 *
 * Conventions:
 *   - arrays are host pointers in the Fortran layout x(3,pcount): x[3*i+k]
 *   - links are 1-based; infront[i]==0 marks an empty slot (holes allowed)
 *   - every call is synchronous on return, ctx is not thread-safe; call it
 *     from the master thread, outside OpenMP regions
 *   - every entry point returns a qv_status; nothing calls exit()
 *   - arithmetic: QV_FP64 reproduces the reference FP64 velocity (decisions
 *     of the tree walk bit-identical, values within 1e-12 relative);
 *     QV_FP32C evaluates leaf segments and cell multipoles in FP32 with
 *     compensated per-target sums and keeps the LIA term, the walk decisions
 *     and the AB update in FP64.
 */
#ifndef QVORT_GPU_H
#define QVORT_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qv_ctx qv_ctx;
typedef struct qv_batch qv_batch;

/* run.in `velocity` (cdata.f90:489-491) */
enum qv_velocity { QV_VEL_BS = 1, QV_VEL_TREE = 2 };
/* run.in `boundary` -> periodic_bc / _notx / _notxy (initial.f90:104-136) */
enum qv_boundary { QV_BC_OPEN = 0, QV_BC_PERIODIC = 1, QV_BC_OPENX = 2, QV_BC_OPENXY = 3 };
/* arithmetic of the nonlocal sum */
enum qv_precision { QV_FP64 = 0, QV_FP32C = 1 };
/* qv_params.flags: run.in switches of pmotion.  QV_FLAG_REMOVE_MEAN_POS:
 * remove_mean_pos T, the mean x and y of the active points subtracted from
 * every slot after each AB update (timestep.f90:86-92).  (pmotion's
 * dt_adapt, timestep.f90:62-85, has no run.in label in the reference,
 * cdata.f90:420-807, so it is always off there and has no flag here.) */
enum qv_param_flags { QV_FLAG_REMOVE_MEAN_POS = 1 };
/* flags of qv_probe_velocity */
enum qv_probe_flags { QV_PROBE_GHOSTP = 1, QV_PROBE_REFERENCE_SCALE = 2, QV_PROBE_EMPTY_CORE = 4 };

enum qv_status {
  QV_OK = 0,
  QV_ERR_ARG = 1,        /* invalid argument or parameter combination */
  QV_ERR_HIP = 2,        /* HIP runtime error (device, allocation, launch) */
  QV_ERR_SINGULAR = 3,   /* target coincides with a leaf: the fatal_error of
                            tree.f90:372-376 ("singularity in BS (tree)
                            velocity field") */
  QV_ERR_TREE_DEPTH = 4, /* two points closer than box_size/2^32: the octree
                            would need more than 32 levels */
  QV_ERR_STATE = 5,      /* call out of order (e.g. qv_ab_step before any
                            qv_velocity) */
  QV_ERR_NO_DEVICE = 6,  /* no HIP device visible */
  QV_ERR_LINE_EMPTY = 7, /* no active point: the fatal_error of line.f90:27
                            ("vortex line length is 0, run over") */
  QV_ERR_CAPACITY = 8,   /* pinsert needs more slots than the arrays hold */
  QV_ERR_SCRATCH = 9     /* qv_vjp_device_noalloc: more slots than its static
                            scratch holds (65536) */
};

/* the run.in parameters of the hot path (cdata.f90:196-416) */
typedef struct qv_params {
  double quant_circ;             /* Gamma, `quant_circ` */
  double corea;                  /* a, `corea` */
  double box_size;               /* L, `box_size` (> 0 for every boundary) */
  double delta;                  /* `delta` (not used by the velocity) */
  double tree_theta;             /* `tree_theta` (Tree: > 0) */
  double dt;                     /* `dt` */
  int32_t velocity;              /* enum qv_velocity */
  int32_t boundary;              /* enum qv_boundary */
  int32_t tree_extra_correction; /* `tree_extra_correction` (reference default 1) */
  int32_t deriv_order;           /* `deriv_order`: 2 ('second') or 4 ('fourth') */
  int32_t precision;             /* enum qv_precision */
  int32_t flags;                 /* enum qv_param_flags; other bits must be 0 */
} qv_params;

/* per-stage timings of the last qv_velocity (device events, milliseconds) */
typedef struct qv_timings {
  float ghost_lia_ms;  /* ghost gather + LIA */
  float build_ms;      /* keys + sort + nodes + moments (Tree only) */
  float walk_ms;       /* tree walk or direct sum, incl. final composition */
  float total_ms;      /* whole qv_velocity on the device */
  int64_t nodes;       /* non-empty tree nodes (Tree only) */
  int64_t evals;       /* interactions (reference eval_counter, 64-bit) */
  int64_t visits;      /* node visits of all walks (Tree only) */
  int64_t leaf_evals;  /* leaf segment interactions (BS: source pairs) */
  int64_t cell_evals;  /* accepted-cell (multipole) interactions */
  int64_t wave_steps;  /* node records fetched by the walk's wavefronts */
} qv_timings;

int qv_create(const qv_params *params, qv_ctx **out);
int qv_destroy(qv_ctx *ctx);
int qv_set_params(qv_ctx *ctx, const qv_params *params);

/* copy f(:)%x, infront, behind to the device; resets the AB history to 0 */
int qv_upload(qv_ctx *ctx, int32_t pcount, const double *x,
              const int32_t *infront, const int32_t *behind);
/* positions and links of f(:) after the host line operations of a step
 * (pinsert, precon, premove), keeping the device's velocities and AB
 * history: pcount may have grown (never shrunk); a slot that held no point
 * and holds one now is a point pinsert made (u1 = u2 = 0, u the mean of
 * the velocities of the point it follows and of that point's infront,
 * line.f90:86-93, in pinsert's order along chains of new points); a slot
 * that held a point and is empty now is cleared as clear_particle clears
 * it (general.f90:11-22).  Call it right after pinsert (the new points'
 * velocities are read off the links pinsert left) and after precon /
 * premove.  x may be NULL when no slot becomes a point (after precon and
 * premove, which only relink and clear): the device keeps its positions and
 * zeroes the cleared slots' x; a slot that becomes a point then returns
 * QV_ERR_ARG, with the links already applied.  The copies of a topology step
 * are then x down, the ghosts down, x and the links up after pinsert, and
 * the links up after premove, instead of the whole state both ways. */
int qv_upload_topology(qv_ctx *ctx, int32_t pcount, const double *x,
                       const int32_t *infront, const int32_t *behind);
/* AB history f(:)%u1..u3 (NULL = leave as is) */
int qv_upload_history(qv_ctx *ctx, const double *u1, const double *u2,
                      const double *u3);

/* velocity of every slot: u = f(:)%u, u_s_li = f(:)%u_s_LI, u_s_bs =
 * f(:)%u_s_BS (any output may be NULL; empty slots get 0).  evals (may be
 * NULL): eval_counter of the stage (Tree), 64-bit. */
int qv_velocity(qv_ctx *ctx, double *u, double *u_s_li, double *u_s_bs,
                int64_t *evals);
/* AB ladder + history shift with the velocity of the last qv_velocity
 * (then remove_mean_pos with QV_FLAG_REMOVE_MEAN_POS) */
int qv_ab_step(qv_ctx *ctx);
/* the boundary step of run.f90:82-90 on the device state: enforce_periodic
 * (periodic.f90:188-213) for periodic; for openx / openxy
 * enforce_periodic_yz / _z (periodic.f90:217-276), which wrap the periodic
 * axes and remove every loop with a point beyond an open face
 * (boundary_loop_remove, periodic.f90:306-327, slots cleared as
 * clear_particle clears them).  The device checks for such points; only
 * when one exists is the state brought to the host, where the removal runs
 * in the reference's slot order (its length counter depends on it), and
 * sent back.  open: nothing. */
int qv_enforce_periodic(qv_ctx *ctx);
/* boundary_loop_remove_count and _length (cdata.f90:316-317) summed over
 * the context's qv_enforce_periodic / qv_step calls */
int qv_boundary_removed(qv_ctx *ctx, int64_t *count, double *length);
/* nsteps of: velocity, AB step, boundary step -- on the device (openx /
 * openxy: one device-to-host flag read per step) */
int qv_step(qv_ctx *ctx, int32_t nsteps);
/* device state back to the host (NULL = skip) */
int qv_download(qv_ctx *ctx, double *x, double *u, double *u1, double *u2,
                double *u3);
/* the device links f(:)%infront / %behind (NULL = skip): they change on the
 * device only when qv_enforce_periodic / qv_step remove loops at an open
 * face (qv_boundary_removed counts them) */
int qv_download_links(qv_ctx *ctx, int32_t *infront, int32_t *behind);

/* the tree of the last Tree-mode qv_velocity, every node (empty children
 * included) in DFS order, 22 doubles per node: level, pcount, leaf index j
 * (1-based, 0 if none), posx, posy, posz, width, cent(3), circ(3), Cij(9)
 * column-major.  nnodes receives the node count; at most max_nodes records
 * are written (out may be NULL to query the count). */
int qv_tree_dump(qv_ctx *ctx, double *out, int64_t max_nodes, int64_t *nnodes);

/* pclose_tree / closest_tree (tree.f90:219-278) when tree_theta > 0, else
 * pclose (line.f90:205-234): for every slot, the closest point that is not
 * i, infront(i) or behind(i), with the context's current positions and
 * links.  The tree is the one of the last qv_velocity (construct_tree runs
 * whenever tree_theta > 0, run.f90:43-45, also with velocity='BS'), built
 * from the positions before that step's motion: the reference's order is
 * construct_tree, pmotion, pinsert, pclose_tree (run.f90:43-66), so a caller
 * may qv_upload the state after its host pinsert (more slots, new links)
 * before this call.  A one-point leaf offers j = behind(infront(p)) with
 * infront as at the build and behind as now; a many-point cell taken by
 * width/dist < 0.5 sets closest = 0.  closest (1-based) and closestd are
 * f(:)%closest / %closestd, in and out: tree mode sets closestd = 100 for
 * every slot before searching; a slot where nothing is taken keeps its
 * closest; pclose leaves empty slots untouched.  Bit-identical to the
 * reference.  QV_ERR_STATE: tree mode without a tree, or fewer slots than
 * at the build. */
int qv_closest(qv_ctx *ctx, int32_t *closest, double *closestd);

/* pmotion_RK3 (timestep.f90:95-176), the integrator of run.in
 * runge_kutta T (run.f90:50-52): three stages of ghostp + construct_tree +
 * the velocity (as qv_velocity) with the low-storage update u1 = u dt,
 * x += u1/3; u1 = -(2/3) u1 + u dt, x += u1; u1 = -u1 + u dt, x += u1/2.
 * u1 is the Runge-Kutta storage (f(:)%u1); no wrap (qv_enforce_periodic). */
int qv_rk3_step(qv_ctx *ctx);

/* superfluid velocity at nprobe points y (layout y(3,nprobe)), summed as
 * mesh_velocity does (timestep.f90:374-452) and velocity_quasip
 * (hamiltonian.f90:26-97): central shell, then the image shells of the
 * boundary.  velocity='BS': biot_savart_general / _shift
 * (timestep.f90:516-563) over the current positions with the ghosts of the
 * last ghostp -- the velocity stage's, or fresh ones with QV_PROBE_GHOSTP
 * (mesh_velocity runs after pmotion with the step's ghosts; the
 * quasiparticles after the later ghostp, run.f90:75) -- bit-identical to the
 * reference.  grad (may be NULL; BS only) receives du_k/dy_m at
 * grad[9*q + 3*k + m], the analytic derivative of the same sum: it replaces
 * the six extra evaluations and central differences of velocity_quasip
 * (hamiltonian.f90:64-79).  velocity='Tree': tree_walk_general
 * (tree.f90:425-488) over the tree of the last qv_velocity.  The reference
 * adds leaves as u/(4 pi) without Gamma and cells without Gamma/(4 pi)
 * (tree.f90:460, 479); by default both get Gamma/(4 pi), the velocity of
 * the BS sum, and QV_PROBE_REFERENCE_SCALE keeps the reference's factors.
 * QV_PROBE_EMPTY_CORE (BS only) applies velocity_quasip's empty core
 * (hamiltonian.f90:42-58): a probe closer than corea to the position of any
 * slot -- empty slots included, as the reference loops over all of f(:) --
 * gets u = 0 and a zero gradient.
 * QV_ERR_STATE: no ghosts (BS) or no tree (Tree) yet.  No counterpart for
 * the gradient in the reference. */
int qv_probe_velocity(qv_ctx *ctx, int32_t nprobe, const double *y, double *u, double *grad,
                      int32_t flags);

/* ---- derivatives of the direct-sum velocity (velocity = QV_VEL_BS) ----
 * No counterpart in the reference: the adjoint, for gradients of objectives
 * (line length, energy) through many steps, and the tangent map, for
 * implicit solvers.  U(x) is the velocity of qv_velocity (LIA term +
 * biot_savart + biot_savart_shift) at the context's positions; the loops
 * must be consistent (infront(behind(i)) == i), else QV_ERR_ARG.
 * precision QV_FP64: FP64 pair terms; QV_FP32C: FP32 pair terms with FP64
 * relative vectors and sums.  Arrays in the layout of x(3,pcount). */
/* xbar = (dU/dx)^T w */
int qv_vjp(qv_ctx *ctx, const double *w, double *xbar);
/* jv = (dU/dx) v */
int qv_jvp(qv_ctx *ctx, const double *v, double *jv);

/* ---- device-pointer variants for callers that keep the state in device
 * memory (e.g. a PyTorch harness).  No counterpart in the reference.  All
 * pointers are device buffers of (pcount,3) doubles; d_x (may be NULL to
 * keep the context's positions) replaces the context's positions. */
/* launch every later call on `hip_stream` (a hipStream_t, used as given:
 * NULL is HIP's null stream); a context starts on a stream of its own */
int qv_set_stream(qv_ctx *ctx, void *hip_stream);
int qv_velocity_device(qv_ctx *ctx, const double *d_x, double *d_u);
int qv_vjp_device(qv_ctx *ctx, const double *d_x, const double *d_w, double *d_xbar);
int qv_jvp_device(qv_ctx *ctx, const double *d_x, const double *d_v, double *d_jv);
/* qv_vjp_device without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the
 * scratch of the adjoint kernels is a static device array of the library
 * (5 Mi doubles, 42 MB per device, held until the process ends: at most
 * 65536 slots) instead of buffers qv_vjp_device allocates on its first
 * call.  Same arguments, same result bit for bit.  More than 65536 slots:
 * QV_ERR_SCRATCH.  Every call on one device shares that device's array:
 * calls that overlap on one GPU (two contexts on different streams, or two
 * threads) race. */
int qv_vjp_device_noalloc(qv_ctx *ctx, const double *d_x, const double *d_w,
                          double *d_xbar);

/* ---- ensembles of independent systems.  No counterpart in the reference,
 * which runs one system per process (run.f90).  Many filament systems of
 * uneven point counts, velocity='BS' with the same run.in parameters, are
 * held in one context and advanced together: one persistent launch per
 * step computes every system's LIA term + direct sum and applies each
 * system's AB ladder and wrap.  precision QV_FP64 gives, per system, the
 * velocity qv_velocity gives for that system alone (bit-identical);
 * QV_FP32C uses FP32 pair terms with FP64 relative vectors and sums.
 * Arrays are the systems' x(3,n_s) arrays one after another; links are
 * 1-based within each system. */
int qv_batch_create(const qv_params *params, int32_t nsys, const int32_t *counts,
                    qv_batch **out);
int qv_batch_destroy(qv_batch *batch);
/* positions and links; clears the AB history */
int qv_batch_upload(qv_batch *batch, const double *x, const int32_t *infront,
                    const int32_t *behind);
/* velocity of every system at the current positions (u may be NULL) */
int qv_batch_velocity(qv_batch *batch, double *u);
/* nsteps of velocity + AB ladder + wrap; use_graph != 0 replays a captured
 * HIP graph of up to 10 steps */
int qv_batch_step(qv_batch *batch, int32_t nsteps, int32_t use_graph);
/* positions and the last velocity (either may be NULL) */
int qv_batch_download(qv_batch *batch, double *x, double *u);
/* times each tile shape on trial_steps velocity launches and keeps the
 * fastest; ms (may be NULL) receives one time per shape */
int qv_batch_tune(qv_batch *batch, int32_t trial_steps, int32_t *chosen, float *ms);
int qv_batch_set_variant(qv_batch *batch, int32_t variant);
/* number of tile shapes, current shape and its targets per work unit,
 * work units per launch, device time of the last qv_batch_step call */
int qv_batch_info(qv_batch *batch, int32_t *nvariants, int32_t *variant, int32_t *tile,
                  int32_t *nunits, float *last_ms);

/* ---- line operations of a step (run.f90:60-72), host code on host arrays.
 * The reference's topology step: pinsert, then every recon_shots steps the
 * neighbour search (qv_closest on the device tree), precon and premove.
 * They are branchy list surgery on a few points per step and run on the
 * host as in the reference; they read the ghost zones the step's ghostp
 * left (qv_download_ghosts), which pinsert refreshes only for the points it
 * touches, exactly as the reference does.  Run.in options they implement:
 * recon_type 'original', delta_adapt F, phonon_emission F, recon_info F,
 * no mirror; the diagnostic logs (removed_curv.log, small_loops.log) are
 * not written.  The slot loops run in the reference's order; the per-slot
 * tests they read, and the validation of the links, are computed first on
 * up to 16 host threads (the environment variable QV_HOST_THREADS sets the
 * number; 0 evaluates every test inside the loop, without the pre-pass),
 * which changes no result. */
typedef struct qv_line_state {
  int32_t pcount;    /* slots in use (size(f)); pinsert may raise it */
  int32_t capacity;  /* slots allocated in every array below */
  double *x, *u, *u1, *u2, *u3;                /* (3, capacity) each */
  double *ghosti, *ghostb, *ghostii, *ghostbb; /* (3, capacity) each */
  int32_t *infront, *behind, *closest;         /* (capacity), 1-based */
  double *closestd;                            /* (capacity) */
} qv_line_state;

/* the reference's module counters (cdata.f90:146-154, 314-315); calls add
 * to the counts, pinsert overwrites total_length and avg_sep */
typedef struct qv_line_counts {
  int64_t inserted;                /* points added by pinsert */
  int64_t recon_count;             /* reconnections (recon_count) */
  int64_t remove_count;            /* points removed by premove (remove_count) */
  int64_t small_loop_remove_count; /* loops of 3-4 points removed by loop_killer */
  double total_length;             /* sum of segment lengths seen by pinsert */
  double avg_sep;                  /* total_length / active points (pinsert) */
  double small_loop_remove_length; /* length of the loops loop_killer removed */
} qv_line_counts;

/* ghost zones of the last ghostp (velocity stage or QV_PROBE_GHOSTP), in
 * the layout of x(3,pcount); NULL skips an array.  QV_ERR_STATE if none. */
int qv_download_ghosts(qv_ctx *ctx, double *gi, double *gb, double *gii, double *gbb);
/* pinsert (line.f90:18-108): a point between i and infront(i) where their
 * separation exceeds delta, curvature-placed, into the first empty slot or
 * appended.  Each slot adds at most one point, so capacity >= 2 pcount
 * always suffices; with less, QV_ERR_CAPACITY is returned when a point does
 * not fit and the points inserted before it stay (pcount counts them). */
int qv_line_pinsert(const qv_params *params, qv_line_state *s, qv_line_counts *c);
/* precon_original (reconnection.f90:145-256) with loop_killer
 * (reconnection.f90:421-461): reconnects i with closest(i) when closestd(i)
 * < delta/2, the tangents are not parallel and the line does not grow */
int qv_line_precon(const qv_params *params, qv_line_state *s, qv_line_counts *c);
/* premove (line.f90:114-161) with loop_killer: removes infront(i) when i and
 * infront(infront(i)) are closer than 0.998 delta */
int qv_line_premove(const qv_params *params, qv_line_state *s, qv_line_counts *c);

/* pinned (page-locked) host memory for the caller's f(:) arrays, so that
 * qv_upload / qv_download / qv_download_ghosts copy at the full PCIe rate
 * (pageable arrays are staged by the runtime).  No counterpart in the
 * reference, whose f(:) lives in ordinary Fortran allocatables.
 * qv_host_alloc: *ptr = `bytes` bytes, or NULL and QV_ERR_HIP (e.g. no
 * device), in which case ordinary memory serves as well, only slower.
 * qv_host_free: frees what qv_host_alloc returned; NULL is ignored. */
int qv_host_alloc(size_t bytes, void **ptr);
int qv_host_free(void *ptr);

int qv_get_timings(qv_ctx *ctx, qv_timings *t);
const char *qv_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif /* QVORT_GPU_H */
