/*
 * qv_oracle.h -- CPU restatement of Qvort's nonlocal filament velocity.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * the timed CPU baseline.  The product (libqvort_gpu.so) never links it.
 *
 * Conventions follow the reference's data model (cdata.f90:20-41):
 *   x[3*i+k]    position of slot i (0-based slot, k = axis), FP64
 *   infront[i], behind[i]  1-based slot links, 0 marks an empty slot
 */
#ifndef QV_ORACLE_H
#define QV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { QVO_VEL_BS = 1, QVO_VEL_TREE = 2 };
enum { QVO_BC_OPEN = 0, QVO_BC_PERIODIC = 1, QVO_BC_OPENX = 2, QVO_BC_OPENXY = 3 };

typedef struct qvo_params {
  double quant_circ;   /* Gamma        cdata.f90:202 */
  double corea;        /* a            cdata.f90:203 */
  double box_size;     /* L            cdata.f90:200 */
  double delta;        /* delta        cdata.f90:198 */
  double tree_theta;   /* theta        cdata.f90:331 */
  double dt;           /* dt           cdata.f90:199 */
  int32_t velocity;    /* QVO_VEL_*    cdata.f90:210 */
  int32_t boundary;    /* QVO_BC_*     initial.f90:104-136 */
  int32_t tree_extra_correction; /*    cdata.f90:333 */
  int32_t deriv_order; /* 2 or 4       cdata.f90:218 */
  int32_t remove_mean_pos; /* 0 or 1   cdata.f90:210, pmotion timestep.f90:86-92 */
} qvo_params;

/* node record written by qvo_tree_dump: 22 doubles per node, DFS order,
 * empty children included (the order of tree_walk's recursion) */
enum { QVO_NODE_DOUBLES = 22 };

int qvo_ghostp(const qvo_params *p, int32_t pcount, const double *x,
               const int32_t *infront, const int32_t *behind,
               double *gi, double *gb, double *gii, double *gbb);

/* full velocity of every active slot (calc_velocity, timestep.f90:180-309);
 * outputs are zero for empty slots.  evals: eval_counter after the stage
 * (Tree mode, 64-bit).  nthreads<=0: OpenMP default. */
int qvo_velocity(const qvo_params *p, int32_t pcount, const double *x,
                 const int32_t *infront, const int32_t *behind,
                 double *u, double *u_li, double *u_bs, int64_t *evals,
                 int32_t nthreads);

/* velocity of the listed 0-based targets only (bounded CPU baseline);
 * writes u[3*k] for target k; returns 0 or a negative error.
 * build_s / walk_s receive the wall times of the tree build and the walk. */
int qvo_velocity_targets(const qvo_params *p, int32_t pcount, const double *x,
                         const int32_t *infront, const int32_t *behind,
                         const int32_t *targets, int32_t ntargets, double *u,
                         int64_t *evals, int32_t nthreads,
                         double *build_s, double *walk_s);

/* nodes of the octree (construct_tree, tree.f90:44-186) in DFS order.
 * Returns the node count; writes at most max_nodes records into out. */
int64_t qvo_tree_dump(const qvo_params *p, int32_t pcount, const double *x,
                      const int32_t *infront, const int32_t *behind,
                      double *out, int64_t max_nodes);

/* Adams-Bashforth ladder of pmotion (timestep.f90:48-71), in place */
int qvo_ab_update(const qvo_params *p, int32_t pcount, double *x,
                  const int32_t *infront, const double *u,
                  double *u1, double *u2, double *u3);

/* wrap after a step: enforce_periodic / _yz / _z (periodic.f90:188-276) */
int qvo_enforce_periodic(const qvo_params *p, int32_t pcount, double *x,
                         const int32_t *infront);
/* the boundary step of run.f90:82-90: enforce_periodic (periodic), or
 * enforce_periodic_yz (openx) / enforce_periodic_z (openxy), which wrap the
 * periodic axes and remove every loop with a point beyond an open face
 * (boundary_loop_remove, periodic.f90:306-327), slot by slot in the
 * reference's order; removed slots are cleared as clear_particle clears them
 * (x, u1, u2, links; u3 kept).  count / length (may be NULL) accumulate
 * boundary_loop_remove_count / _length. */
int qvo_enforce_boundary(const qvo_params *p, int32_t pcount, double *x, int32_t *infront,
                         int32_t *behind, double *u1, double *u2, int64_t *count,
                         double *length);

/* one straight segment x_j -> g_j (+shift) seen from x_i; form 0 is the
 * biot_savart pair formula, form 1 the tree leaf formula */
int qvo_segment_velocity(const qvo_params *p, const double *xi, const double *xj,
                         const double *gj, const double *shift, int32_t form,
                         double *u);

/* nsteps of the reference main loop without line operations
 * (run.f90:41-54, 75-107): ghostp, tree, velocity, AB, ghostp, wrap.
 * x/u1/u2/u3 updated in place; u_first receives the first step's velocity. */
int qvo_run(const qvo_params *p, int32_t pcount, double *x,
            int32_t *infront, int32_t *behind,
            double *u1, double *u2, double *u3, int32_t nsteps,
            double *u_first, int32_t nthreads, int64_t *bcount, double *blength);

/* nsteps of the main loop with the Runge-Kutta integrator (runge_kutta T,
 * run.f90:50-52): pmotion_RK3 (timestep.f90:95-176) -- three stages of
 * ghostp, construct_tree and the velocity, with the low-storage update
 * u1 = u dt; x += u1/3; u1 = -(2/3) u1 + u dt; x += u1; u1 = -u1 + u dt;
 * x += u1/2 -- then the wrap.  u1 is the Runge-Kutta storage, in/out;
 * u_last receives the last stage's velocity of the last step. */
int qvo_run_rk3(const qvo_params *p, int32_t pcount, double *x,
                int32_t *infront, int32_t *behind, double *u1,
                int32_t nsteps, double *u_last, int32_t nthreads);

/* theta -> 0+ limit of the Tree velocity (every leaf with tree_walk's
 * exclusions, slot order) for the listed targets: the reference for the
 * multipole truncation error.  u[3*k] for target k. */
int qvo_leaf_sum_targets(const qvo_params *p, int32_t pcount, const double *x,
                         const int32_t *infront, const int32_t *behind,
                         const int32_t *targets, int32_t ntargets, double *u,
                         int64_t *evals, int32_t nthreads);

/* neighbour search of the reconnection step.
 * qvo_closest_tree: pclose_tree / closest_tree (tree.f90:219-278).  The tree
 * is built from the build state (x_b, infront_b, behind_b: the positions and
 * links of construct_tree, before pmotion and pinsert); the search uses the
 * current state (pcount >= pcount_b slots).  For every slot closestd = 100,
 * then an active slot i walks the tree: a node with one point, or with
 * width/dist < 0.5, is taken when dist < closestd; a many-point node sets
 * closest = 0 (closestd unchanged); a one-point node's candidate is
 * j = behind(infront_b(point)) with the current behind, taken unless j is i,
 * infront(i) or behind(i).  closest is in/out (kept where nothing is taken).
 * qvo_closest_direct: pclose (line.f90:205-234), all pairs of the current
 * state; empty slots keep closest and closestd.  Returns 0 or -1. */
int qvo_closest_tree(const qvo_params *p, int32_t pcount_b, const double *x_b,
                     const int32_t *infront_b, const int32_t *behind_b,
                     int32_t pcount, const double *x, const int32_t *infront,
                     const int32_t *behind, int32_t *closest, double *closestd,
                     int32_t nthreads);
int qvo_closest_direct(int32_t pcount, const double *x, const int32_t *infront,
                       const int32_t *behind, int32_t *closest, double *closestd,
                       int32_t nthreads);

/* superfluid velocity at probe points, as mesh_velocity computes it
 * (timestep.f90:374-452): central shell, then the image shells of the
 * boundary in the order (p outer .. r inner), one running sum per probe.
 * velocity BS: biot_savart_general / _shift (timestep.f90:516-563) over the
 * current positions x with the ghosts gi given (mesh_velocity runs after
 * pmotion, so they are the ghosts of the step's ghostp); no exclusions.
 * velocity Tree: tree_walk_general (tree.f90:425-488) over the tree of the
 * build state (x_b, links_b).  ref_scaling = 1 keeps the reference's factors
 * there (leaf u_bs/4/pi without Gamma, cell without Gamma/(4 pi)); 0 applies
 * Gamma/(4 pi) to both as tree_walk does.  Returns 0 or -1. */
int qvo_probe_velocity(const qvo_params *p, int32_t pcount_b, const double *x_b,
                       const int32_t *infront_b, const int32_t *behind_b,
                       int32_t pcount, const double *x, const int32_t *infront,
                       const double *gi, int32_t nprobe, const double *probe,
                       int32_t ref_scaling, double *u, int32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
