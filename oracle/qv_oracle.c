/*
 * qv_oracle.c -- CPU restatement of Qvort's nonlocal filament velocity.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  This file is the oracle: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * as the checker or the timed CPU baseline.  libqvort_gpu.so never links it.
 *
 * Every routine restates the reference Fortran (compiled there with
 * -fdefault-real-8, so all arithmetic is FP64) and keeps its operation order
 * so that results can be compared bit for bit with the flang build of the
 * reference (oracle/_ref).  Compile with -ffp-contract=off.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   ghostp / derivatives / LIA    golden vectors of the reference build, the
 *                                 ring closed form, exactness of the 3-point
 *                                 stencil on quadratics
 *   biot_savart(_shift)           golden vectors, exact straight-segment field,
 *                                 brute force on tiny N
 *   construct_tree / tree_walk    node-by-node dump of the reference tree
 *                                 (bounds, occupancy, moments), golden
 *                                 velocities, theta->0 reduces to the pair sum
 *   pmotion AB ladder             golden trajectories, polynomial exactness
 *   leaf sum (theta -> 0+)        the theta = 1e-9 walk (same leaf terms and
 *                                 count) and a numpy brute force of the
 *                                 tree's exclusion rule
 */
#include "qv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define QVO_PI 3.14159265358979324 /* cdata.f90:176 */

/* ------------------------------------------------------------------ */
/* helpers: general.f90:38-65,142-158, derivatives.f90:79-86           */

static inline double sq(double a) { return a * a; }

/* dist_gen(a,b) derivatives.f90:79-86 */
static inline double dist_gen(const double *a, const double *b) {
  return sqrt((sq(a[0] - b[0]) + sq(a[1] - b[1])) + sq(a[2] - b[2]));
}

/* dist_gen_sq(a,b) general.f90:58-65 */
static inline double dist_gen_sq(const double *a, const double *b) {
  return (sq(a[0] - b[0]) + sq(a[1] - b[1])) + sq(a[2] - b[2]);
}

/* cross_product(a,b) general.f90:142-149 */
static inline void cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static inline double dot3(const double *a, const double *b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* vector_norm(a) general.f90:153-158 */
static inline double vnorm(const double *a) { return sqrt(dot3(a, a)); }

/* x**3 and x**5: flang expands small integer powers as a left-to-right
 * product ((x*x)*x)..., which the bit-exact comparison with the flang build
 * of tree_walk (tree.f90:402-407) confirms */
static inline double pow3(double x) { return (x * x) * x; }
static inline double pow5(double x) { return (((x * x) * x) * x) * x; }

/* ------------------------------------------------------------------ */
/* ghost points: periodic.f90:23-184                                   */

static void unwrap_axis(double xi, double *g, double L) {
  if ((xi - *g) > (L / 2.)) {
    *g = *g + L;
  } else if ((xi - *g) < (-L / 2.)) {
    *g = *g - L;
  }
}

static void ghost_one(const qvo_params *p, const double *x,
                      const int32_t *infront, const int32_t *behind, int i,
                      double *gi, double *gb, double *gii, double *gbb) {
  const int a = infront[i] - 1, b = behind[i] - 1;
  const int aa = infront[a] - 1, bb = behind[b] - 1;
  for (int k = 0; k < 3; ++k) {
    gi[k] = x[3 * a + k];
    gb[k] = x[3 * b + k];
    gii[k] = x[3 * aa + k];
    gbb[k] = x[3 * bb + k];
  }
  int k0 = 3;
  if (p->boundary == QVO_BC_PERIODIC) k0 = 0;
  else if (p->boundary == QVO_BC_OPENX) k0 = 1;
  else if (p->boundary == QVO_BC_OPENXY) k0 = 2;
  const double L = p->box_size;
  for (int k = k0; k < 3; ++k) {
    const double xi = x[3 * i + k];
    unwrap_axis(xi, &gi[k], L);
    unwrap_axis(xi, &gb[k], L);
    unwrap_axis(xi, &gii[k], L);
    unwrap_axis(xi, &gbb[k], L);
  }
}

int qvo_ghostp(const qvo_params *p, int32_t pcount, const double *x,
               const int32_t *infront, const int32_t *behind, double *gi,
               double *gb, double *gii, double *gbb) {
  for (int i = 0; i < pcount; ++i) {
    if (infront[i] == 0) {
      for (int k = 0; k < 3; ++k)
        gi[3 * i + k] = gb[3 * i + k] = gii[3 * i + k] = gbb[3 * i + k] = 0.;
      continue;
    }
    ghost_one(p, x, infront, behind, i, gi + 3 * i, gb + 3 * i, gii + 3 * i,
              gbb + 3 * i);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* finite differences: derivatives.f90:11-75                           */

static void deriv_1(int order, const double *x, const double *gi,
                    const double *gb, const double *gii, const double *gbb,
                    double *s) {
  const double di = dist_gen(x, gi), db = dist_gen(x, gb);
  if (order != 4) {
    for (int k = 0; k < 3; ++k) s[k] = (db * gi[k] + (di - db) * x[k]) - di * gb[k];
    const double den = (2. * db) * di;
    for (int k = 0; k < 3; ++k) s[k] = s[k] / den;
    return;
  }
  const double dii = dist_gen(gii, gi), dbb = dist_gen(gbb, gb);
  double A = db * (di * di) + (db * di) * dii;
  A = A / (((dbb * (dbb + db)) * ((dbb + db) + di)) * (((dbb + db) + di) + dii));
  double B = ((-(dbb * (di * di)) - db * (di * di)) - (dbb * di) * dii) - (db * di) * dii;
  B = B / (((dbb * db) * (db + di)) * ((db + di) + dii));
  double D = (((dbb * db) * di + (db * db) * di) + (dbb * db) * dii) + (db * db) * dii;
  D = D / (((di * dii) * (db + di)) * ((dbb + db) + di));
  double E = -(di * (db * db)) - (dbb * db) * di;
  E = E / (((dii * (di + dii)) * ((db + di) + dii)) * (((dbb + db) + di) + dii));
  const double C = -(((A + B) + D) + E);
  for (int k = 0; k < 3; ++k)
    s[k] = (((A * gbb[k] + B * gb[k]) + C * x[k]) + D * gi[k]) + E * gii[k];
}

static void deriv_2(int order, const double *x, const double *gi,
                    const double *gb, const double *gii, const double *gbb,
                    double *s) {
  const double di = dist_gen(x, gi), db = dist_gen(x, gb);
  if (order != 4) {
    for (int k = 0; k < 3; ++k)
      s[k] = 2. * ((gi[k] / (di * (di + db)) - x[k] / (di * db)) +
                   gb[k] / (db * (di + db)));
    return;
  }
  const double dii = dist_gen(gii, gi), dbb = dist_gen(gbb, gb);
  double A = 2. * (((-((2. * db) * di) + di * di) - db * dii) + di * dii);
  A = A / (((dbb * (dbb + db)) * ((dbb + db) + di)) * (((dbb + db) + di) + dii));
  double B = 2. * ((((((2. * dbb) * di + (2. * db) * di) - di * di) + dbb * dii) +
                    db * dii) - di * dii);
  B = B / (((dbb * db) * (db + di)) * ((db + di) + dii));
  double D = 2. * (((((-(dbb * db) - db * db) + dbb * di) + (2. * db) * di) +
                    dbb * dii) + (2. * db) * dii);
  D = D / (((di * dii) * (db + di)) * ((dbb + db) + di));
  double E = 2. * (((dbb * db + db * db) - dbb * di) - (2. * db) * di);
  E = E / (((dii * (di + dii)) * ((db + di) + dii)) * (((dbb + db) + di) + dii));
  const double C = -(((A + B) + D) + E);
  for (int k = 0; k < 3; ++k)
    s[k] = (((A * gbb[k] + B * gb[k]) + C * x[k]) + D * gi[k]) + E * gii[k];
}

/* local (LIA) term of calc_velocity, timestep.f90:221-225 / 270-274 */
static void lia(const qvo_params *p, const double *x, const double *gi,
                const double *gb, const double *gii, const double *gbb,
                double *u) {
  double s1[3], s2[3], c[3];
  deriv_1(p->deriv_order, x, gi, gb, gii, gbb, s1);
  deriv_2(p->deriv_order, x, gi, gb, gii, gbb, s2);
  const double beta = (p->quant_circ / (4. * QVO_PI)) *
                      log((1.21 / p->corea) * sqrt(dist_gen(x, gi) * dist_gen(x, gb)));
  cross(s1, s2, c);
  for (int k = 0; k < 3; ++k) u[k] = beta * c[k];
}

/* ------------------------------------------------------------------ */
/* direct Biot-Savart: timestep.f90:462-507                            */

typedef struct {
  const qvo_params *p;
  int32_t pcount;
  const double *x;
  const int32_t *infront, *behind;
  double *gi, *gb, *gii, *gbb; /* ghosts of every slot */
} qvo_state;

static void bs_pair(const qvo_state *s, int i, int j, const double *shift,
                    double *u) {
  const double *xi = s->x + 3 * i, *xj = s->x + 3 * j, *gj = s->gi + 3 * j;
  double r[3], e[3], ub[3];
  double a, b, c;
  if (shift) {
    double xs[3] = {xj[0] + shift[0], xj[1] + shift[1], xj[2] + shift[2]};
    a = dist_gen_sq(xi, xs);                      /* timestep.f90:497 */
    for (int k = 0; k < 3; ++k) r[k] = xs[k] - xi[k];
  } else {
    a = (sq(xj[0] - xi[0]) + sq(xj[1] - xi[1])) + sq(xj[2] - xi[2]); /* distfsq(j,i) */
    for (int k = 0; k < 3; ++k) r[k] = xj[k] - xi[k];
  }
  for (int k = 0; k < 3; ++k) e[k] = gj[k] - xj[k];
  b = 2. * dot3(r, e);
  c = dist_gen_sq(gj, xj);
  const double D = (4. * a) * c - b * b;
  if (D == 0.) return;                            /* :476 / :501 */
  cross(r, e, ub);
  const double q = s->p->quant_circ;
  const double den = (2. * QVO_PI) * D;
  for (int k = 0; k < 3; ++k) ub[k] = (ub[k] * q) / den;
  const double fac = (2. * c + b) / sqrt((a + b) + c) - b / sqrt(a);
  for (int k = 0; k < 3; ++k) u[k] = u[k] + ub[k] * fac;
}

static void biot_savart(const qvo_state *s, int i, double *u) {
  const int bi = s->behind[i] - 1;
  for (int j = 0; j < s->pcount; ++j) {
    if (s->infront[j] == 0 || i == j || bi == j) continue;   /* :471 */
    bs_pair(s, i, j, NULL, u);
  }
}

static void biot_savart_shift(const qvo_state *s, int i, double *u,
                              const double *shift) {
  const int bi = s->behind[i] - 1;
  for (int j = 0; j < s->pcount; ++j) {
    if (s->infront[j] == 0 || bi == j) continue;             /* :496 */
    bs_pair(s, i, j, shift, u);
  }
}

/* ------------------------------------------------------------------ */
/* octree: tree.f90:22-186                                             */

typedef struct {
  double pos[3], width;
  int32_t pcount;
  int32_t level;
  int64_t child[8]; /* node index, -1 when not split */
  int32_t rec;      /* pcount==1: 0-based slot of parray(1) */
  double cent[3], circ[3];
  double C[3][3];   /* C[i][k] = Cij(i+1,k+1) */
} qvo_node;

typedef struct {
  qvo_node *n;
  int64_t count, cap;
} qvo_tree;

static int64_t tree_alloc(qvo_tree *t) {
  if (t->count == t->cap) {
    int64_t nc = t->cap ? 2 * t->cap : 1024;
    qvo_node *nn = (qvo_node *)realloc(t->n, (size_t)nc * sizeof(qvo_node));
    if (!nn) return -1;
    t->n = nn;
    t->cap = nc;
  }
  return t->count++;
}

/* child corner offsets in DFS order fbl fbr ftl ftr bbl bbr btl btr
 * (tree.f90:92-114): (dx,dy,dz) */
static const int child_off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1},
                                    {0, 1, 0}, {1, 1, 0}, {0, 1, 1}, {1, 1, 1}};

static int build_node(const qvo_state *s, qvo_tree *t, int64_t self,
                      const int32_t *list, int32_t n);

/* new_tree, tree.f90:129-186 */
static int new_tree(const qvo_state *s, qvo_tree *t, int64_t self,
                    const double *pos, double width, const int32_t *plist,
                    int32_t pn, int32_t level) {
  int32_t *list = (int32_t *)malloc((size_t)(pn > 0 ? pn : 1) * sizeof(int32_t));
  if (!list) return -1;
  qvo_node nd;
  memset(&nd, 0, sizeof(nd));
  for (int k = 0; k < 3; ++k) nd.pos[k] = pos[k];
  nd.width = width;
  nd.level = level;
  nd.rec = -1;
  for (int c = 0; c < 8; ++c) nd.child[c] = -1;
  int32_t counter = 0;
  for (int32_t q = 0; q < pn; ++q) {
    const int i = plist[q];
    if (s->infront[i] == 0) continue;
    const double *x = s->x + 3 * i, *g = s->gi + 3 * i;
    if ((x[0] >= pos[0]) && (x[0] < (pos[0] + width)) && (x[1] >= pos[1]) &&
        (x[1] < (pos[1] + width)) && (x[2] >= pos[2]) && (x[2] < (pos[2] + width))) {
      list[counter++] = i;
      for (int k = 0; k < 3; ++k) nd.cent[k] = nd.cent[k] + x[k];
      for (int k = 0; k < 3; ++k) nd.circ[k] = (nd.circ[k] + g[k]) - x[k];
      for (int k = 0; k < 3; ++k)
        for (int m = 0; m < 3; ++m) nd.C[m][k] = nd.C[m][k] + x[m] * (g[k] - x[k]);
    }
  }
  if (counter != 0) {
    for (int k = 0; k < 3; ++k) nd.cent[k] = nd.cent[k] / counter;
    for (int m = 0; m < 3; ++m)
      for (int k = 0; k < 3; ++k) nd.C[m][k] = nd.C[m][k] - nd.cent[m] * nd.circ[k];
  }
  nd.pcount = counter;
  if (counter == 1) nd.rec = list[0];
  t->n[self] = nd;
  int rc = build_node(s, t, self, list, counter);
  free(list);
  return rc;
}

/* build_tree, tree.f90:85-125 */
static int build_node(const qvo_state *s, qvo_tree *t, int64_t self,
                      const int32_t *list, int32_t n) {
  if (t->n[self].pcount <= 1) return 0;
  const double w = t->n[self].width;
  const double h = w / 2.;
  const int32_t level = t->n[self].level + 1;
  double base[3] = {t->n[self].pos[0], t->n[self].pos[1], t->n[self].pos[2]};
  for (int c = 0; c < 8; ++c) {
    double pos[3];
    for (int k = 0; k < 3; ++k) pos[k] = child_off[c][k] ? base[k] + (w / 2) : base[k];
    int64_t ci = tree_alloc(t);
    if (ci < 0) return -1;
    t->n[self].child[c] = ci;
    if (new_tree(s, t, ci, pos, h, list, n, level)) return -1;
  }
  return 0;
}

/* construct_tree, tree.f90:44-80 */
static int construct_tree(const qvo_state *s, qvo_tree *t) {
  t->n = NULL;
  t->count = t->cap = 0;
  int64_t r = tree_alloc(t);
  if (r < 0) return -1;
  qvo_node *nd = &t->n[r];
  memset(nd, 0, sizeof(*nd));
  const double L = s->p->box_size;
  for (int k = 0; k < 3; ++k) nd->pos[k] = -L / 2.;
  nd->width = L;
  nd->pcount = s->pcount;
  nd->rec = (s->pcount == 1) ? 0 : -1;
  for (int c = 0; c < 8; ++c) nd->child[c] = -1;
  double sum[3] = {0., 0., 0.};
  int32_t nact = 0;
  for (int i = 0; i < s->pcount; ++i) {
    if (s->infront[i] <= 0) continue;
    for (int k = 0; k < 3; ++k) sum[k] = sum[k] + s->x[3 * i + k];
    ++nact;
  }
  for (int k = 0; k < 3; ++k) nd->cent[k] = sum[k] / nact;
  int32_t *all = (int32_t *)malloc((size_t)(s->pcount > 0 ? s->pcount : 1) * sizeof(int32_t));
  if (!all) return -1;
  for (int i = 0; i < s->pcount; ++i) all[i] = i;
  int rc = build_node(s, t, r, all, s->pcount);
  free(all);
  return rc;
}

/* tree_walk, tree.f90:336-421.  Returns nonzero on the singularity error. */
/* the leaf term of tree_walk (tree.f90:377-391): segment x_p -> ghosti_p,
 * shifted, seen from x_i */
static void leaf_term(const qvo_state *s, int p, int i, const double *shift, double *u) {
  const double *xi = s->x + 3 * i;
  const double *xp = s->x + 3 * p, *gp = s->gi + 3 * p;
  double R1[3], R2[3], ub[3];
  for (int k = 0; k < 3; ++k) {
    R1[k] = (xp[k] + shift[k]) - xi[k];
    R2[k] = (gp[k] + shift[k]) - xi[k];
  }
  const double a = vnorm(R1), b = vnorm(R2);
  const double c = (a * b) * ((a * b) + dot3(R1, R2));
  cross(R1, R2, ub);
  const double f = (a + b) / c;
  const double q = s->p->quant_circ;
  for (int k = 0; k < 3; ++k) {
    ub[k] = f * ub[k];
    ub[k] = ((ub[k] * q) / 4.) / QVO_PI;
    u[k] = u[k] + ub[k];
  }
}

static int tree_walk(const qvo_state *s, const qvo_tree *t, int64_t ni, int i,
                     const double *shift, double *u, int64_t *evals) {
  const qvo_node *nd = &t->n[ni];
  const double *xi = s->x + 3 * i;
  if (s->infront[i] == 0) return 0;
  if (nd->pcount == 0) return 0;
  const double dist = sqrt((sq(xi[0] - (nd->cent[0] + shift[0])) +
                            sq(xi[1] - (nd->cent[1] + shift[1]))) +
                           sq(xi[2] - (nd->cent[2] + shift[2])));
  double theta;
  if (s->p->tree_extra_correction) {
    const double geo = sqrt((sq(nd->cent[0] - (nd->pos[0] + nd->width / 2.)) +
                             sq(nd->cent[1] - (nd->pos[1] + nd->width / 2.))) +
                            sq(nd->cent[2] - (nd->pos[2] + nd->width / 2.)));
    if (dist - geo > 0.) theta = nd->width / (dist - geo);
    else theta = nd->width / dist;
  } else {
    theta = nd->width / dist;
  }
  if (nd->pcount == 1) {
    const int p = nd->rec;
    const int j = s->behind[s->infront[p] - 1] - 1;   /* f(parray(1)%infront)%behind */
    if (j == i) return 0;
    if (j == s->behind[i] - 1) return 0;
    if (j == s->infront[i] - 1) return 0;
    if (dist < 2.220446049250313e-16) return 1;       /* :372-376 */
    leaf_term(s, p, i, shift, u);
    *evals += 1;
  } else if (theta < s->p->tree_theta) {
    double v[3], ub[3];
    for (int k = 0; k < 3; ++k) v[k] = xi[k] - (nd->cent[k] + shift[k]);
    cross(nd->circ, v, ub);
    const double d3 = pow3(dist), d5 = pow5(dist);
    for (int k = 0; k < 3; ++k) ub[k] = ub[k] / d3;
    /* eps(:,k,l) with eps(1,2,3)=1 (cdata.f90:185-189) */
    for (int k = 0; k < 3; ++k) {
      for (int l = 0; l < 3; ++l) {
        const double dcl = (v[0] * nd->C[0][l] + v[1] * nd->C[1][l]) + v[2] * nd->C[2][l];
        for (int m = 0; m < 3; ++m) {
          double e = 0.;
          if (m == k || k == l || m == l) e = 0.;
          else e = (((m + 1) % 3 == k) ? 1. : -1.);
          ub[m] = ub[m] - (e * nd->C[k][l]) / d3;
        }
        for (int m = 0; m < 3; ++m) {
          double e = 0.;
          if (m == k || k == l || m == l) e = 0.;
          else e = (((m + 1) % 3 == k) ? 1. : -1.);
          ub[m] = ub[m] + (((3. * e) * v[k]) * dcl) / d5;
        }
      }
    }
    const double q = s->p->quant_circ;
    for (int k = 0; k < 3; ++k) u[k] = u[k] + ((ub[k] * q) / 4.) / QVO_PI;
    *evals += 1;
  } else {
    for (int c = 0; c < 8; ++c) {
      int rc = tree_walk(s, t, nd->child[c], i, shift, u, evals);
      if (rc) return rc;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* calc_velocity for the BS and Tree modes, timestep.f90:180-309       */

/* image shifts in the loop order of calc_velocity (p outer .. r inner),
 * (0,0,0) first */
static int shell_shifts(const qvo_params *p, double sh[27][3]) {
  const double L = p->box_size;
  int n = 0;
  sh[n][0] = sh[n][1] = sh[n][2] = 0.;
  ++n;
  if (p->boundary == QVO_BC_PERIODIC) {
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c) {
          if (a == 0 && b == 0 && c == 0) continue;
          sh[n][0] = a * L; sh[n][1] = b * L; sh[n][2] = c * L; ++n;
        }
  } else if (p->boundary == QVO_BC_OPENX) {
    for (int b = -1; b <= 1; ++b)
      for (int c = -1; c <= 1; ++c) {
        if (b == 0 && c == 0) continue;
        sh[n][0] = 0.; sh[n][1] = b * L; sh[n][2] = c * L; ++n;
      }
  } else if (p->boundary == QVO_BC_OPENXY) {
    for (int c = -1; c <= 1; ++c) {
      if (c == 0) continue;
      sh[n][0] = 0.; sh[n][1] = 0.; sh[n][2] = c * L; ++n;
    }
  }
  return n;
}

static int calc_velocity(const qvo_state *s, const qvo_tree *t, int i,
                         double *u, double *u_li, double *u_bs, int64_t *evals) {
  const qvo_params *p = s->p;
  double ul[3], ub[3], sh[27][3];
  const int ns = shell_shifts(p, sh);
  lia(p, s->x + 3 * i, s->gi + 3 * i, s->gb + 3 * i, s->gii + 3 * i, s->gbb + 3 * i, ul);
  for (int k = 0; k < 3; ++k) u[k] = ul[k];
  if (u_li) for (int k = 0; k < 3; ++k) u_li[k] = ul[k];
  ub[0] = ub[1] = ub[2] = 0.;
  if (p->velocity == QVO_VEL_TREE) {
    if (tree_walk(s, t, 0, i, sh[0], ub, evals)) return -2;
  } else {
    biot_savart(s, i, ub);
  }
  for (int k = 0; k < 3; ++k) u[k] = u[k] + ub[k];
  double ubs[3] = {ub[0], ub[1], ub[2]};
  if (ns > 1) {
    ub[0] = ub[1] = ub[2] = 0.;
    for (int q = 1; q < ns; ++q) {
      if (p->velocity == QVO_VEL_TREE) {
        if (tree_walk(s, t, 0, i, sh[q], ub, evals)) return -2;
      } else {
        biot_savart_shift(s, i, ub, sh[q]);
      }
    }
    for (int k = 0; k < 3; ++k) u[k] = u[k] + ub[k];
    for (int k = 0; k < 3; ++k) ubs[k] = ubs[k] + ub[k];
  }
  if (u_bs) for (int k = 0; k < 3; ++k) u_bs[k] = ubs[k];
  return 0;
}

static int state_init(qvo_state *s, const qvo_params *p, int32_t pcount,
                      const double *x, const int32_t *infront,
                      const int32_t *behind) {
  s->p = p;
  s->pcount = pcount;
  s->x = x;
  s->infront = infront;
  s->behind = behind;
  size_t nb = (size_t)3 * (size_t)(pcount > 0 ? pcount : 1) * sizeof(double);
  s->gi = (double *)malloc(nb);
  s->gb = (double *)malloc(nb);
  s->gii = (double *)malloc(nb);
  s->gbb = (double *)malloc(nb);
  if (!s->gi || !s->gb || !s->gii || !s->gbb) return -1;
  return qvo_ghostp(p, pcount, x, infront, behind, s->gi, s->gb, s->gii, s->gbb);
}

static void state_free(qvo_state *s) {
  free(s->gi); free(s->gb); free(s->gii); free(s->gbb);
}

static int check_params(const qvo_params *p, int32_t pcount) {
  if (!p || pcount < 1) return -1;
  if (p->velocity != QVO_VEL_BS && p->velocity != QVO_VEL_TREE) return -1;
  if (p->velocity == QVO_VEL_TREE && !(p->tree_theta > 0.)) return -1;
  if (!(p->box_size > 0.)) return -1;
  return 0;
}

int qvo_velocity(const qvo_params *p, int32_t pcount, const double *x,
                 const int32_t *infront, const int32_t *behind, double *u,
                 double *u_li, double *u_bs, int64_t *evals, int32_t nthreads) {
  if (check_params(p, pcount)) return -1;
  qvo_state s;
  if (state_init(&s, p, pcount, x, infront, behind)) { state_free(&s); return -1; }
  qvo_tree t = {0};
  int64_t nact = 0;
  for (int i = 0; i < pcount; ++i) nact += (infront[i] != 0);
  if (p->velocity == QVO_VEL_TREE && construct_tree(&s, &t)) {
    state_free(&s); free(t.n); return -1;
  }
  int64_t ev = (p->velocity == QVO_VEL_TREE) ? 2 * nact : 0;  /* tree.f90:77-78 */
  int err = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : ev) reduction(| : err)
  for (int i = 0; i < pcount; ++i) {
    double ui[3] = {0., 0., 0.}, li[3] = {0., 0., 0.}, bi[3] = {0., 0., 0.};
    if (infront[i] != 0) {
      int64_t e = 0;
      if (calc_velocity(&s, &t, i, ui, li, bi, &e)) err |= 1;
      ev += e;
    }
    for (int k = 0; k < 3; ++k) {
      u[3 * i + k] = ui[k];
      if (u_li) u_li[3 * i + k] = li[k];
      if (u_bs) u_bs[3 * i + k] = bi[k];
    }
  }
  if (evals) *evals = ev;
  free(t.n);
  state_free(&s);
  return err ? -2 : 0;
}

int qvo_velocity_targets(const qvo_params *p, int32_t pcount, const double *x,
                         const int32_t *infront, const int32_t *behind,
                         const int32_t *targets, int32_t ntargets, double *u,
                         int64_t *evals, int32_t nthreads, double *build_s,
                         double *walk_s) {
  if (check_params(p, pcount)) return -1;
  qvo_state s;
  if (state_init(&s, p, pcount, x, infront, behind)) { state_free(&s); return -1; }
  qvo_tree t = {0};
  double t0 = omp_get_wtime();
  if (p->velocity == QVO_VEL_TREE && construct_tree(&s, &t)) {
    state_free(&s); free(t.n); return -1;
  }
  double t1 = omp_get_wtime();
  int64_t ev = 0;
  int err = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : ev) reduction(| : err)
  for (int k = 0; k < ntargets; ++k) {
    double ui[3] = {0., 0., 0.};
    const int i = targets[k];
    if (i >= 0 && i < pcount && infront[i] != 0) {
      int64_t e = 0;
      if (calc_velocity(&s, &t, i, ui, NULL, NULL, &e)) err |= 1;
      ev += e;
    }
    for (int c = 0; c < 3; ++c) u[3 * k + c] = ui[c];
  }
  double t2 = omp_get_wtime();
  if (build_s) *build_s = t1 - t0;
  if (walk_s) *walk_s = t2 - t1;
  if (evals) *evals = ev;
  free(t.n);
  state_free(&s);
  return err ? -2 : 0;
}

/* The theta -> 0+ limit of the Tree velocity for the listed targets: LIA +
 * the leaf term of every active slot inside the root cell (the tree's
 * membership test, tree.f90:146-150, on the root bounds) with tree_walk's
 * exclusions (j = i, behind(i), infront(i)), central shell then the image
 * shells composed as calc_velocity does (timestep.f90:268-288).  Sources
 * are summed in slot order, not in the tree's DFS order, so the result
 * equals a theta -> 0+ walk up to round-off: it is the reference for the
 * truncation error of the multipole acceptance, not a bit-exact restatement.
 * u[3*k] for target k; evals counts leaf terms. */
int qvo_leaf_sum_targets(const qvo_params *p, int32_t pcount, const double *x,
                         const int32_t *infront, const int32_t *behind,
                         const int32_t *targets, int32_t ntargets, double *u,
                         int64_t *evals, int32_t nthreads) {
  if (check_params(p, pcount)) return -1;
  qvo_state s;
  if (state_init(&s, p, pcount, x, infront, behind)) { state_free(&s); return -1; }
  const double lo = -p->box_size / 2.;
  char *inbox = (char *)malloc((size_t)pcount);
  if (!inbox) { state_free(&s); return -1; }
  for (int j = 0; j < pcount; ++j) {
    const double *xj = x + 3 * j;
    int in = infront[j] != 0;
    for (int k = 0; k < 3; ++k) in = in && xj[k] >= lo && xj[k] < lo + p->box_size;
    inbox[j] = (char)in;
  }
  double sh[27][3];
  const int ns = shell_shifts(p, sh);
  int64_t ev = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : ev)
  for (int k = 0; k < ntargets; ++k) {
    double ui[3] = {0., 0., 0.};
    const int i = targets[k];
    if (i >= 0 && i < pcount && infront[i] != 0) {
      const int bi = behind[i] - 1, fi = infront[i] - 1;
      lia(p, s.x + 3 * i, s.gi + 3 * i, s.gb + 3 * i, s.gii + 3 * i, s.gbb + 3 * i, ui);
      double ub[3] = {0., 0., 0.}, ui2[3] = {0., 0., 0.};
      for (int q = 0; q < ns; ++q) {
        double *acc = q == 0 ? ub : ui2;
        for (int j = 0; j < pcount; ++j) {
          if (!inbox[j]) continue;
          const int jj = behind[infront[j] - 1] - 1;
          if (jj == i || jj == bi || jj == fi) continue;
          leaf_term(&s, j, i, sh[q], acc);
          ++ev;
        }
      }
      for (int c = 0; c < 3; ++c) ui[c] = ui[c] + ub[c];
      if (ns > 1) for (int c = 0; c < 3; ++c) ui[c] = ui[c] + ui2[c];
    }
    for (int c = 0; c < 3; ++c) u[3 * k + c] = ui[c];
  }
  if (evals) *evals = ev;
  free(inbox);
  state_free(&s);
  return 0;
}

static void dump_rec(const qvo_state *s, const qvo_tree *t, int64_t ni,
                     double *out, int64_t max_nodes, int64_t *k) {
  const qvo_node *nd = &t->n[ni];
  if (*k < max_nodes && out) {
    double *o = out + (*k) * QVO_NODE_DOUBLES;
    int32_t leafj = 0;
    if (nd->pcount == 1 && nd->rec >= 0 && s->infront[nd->rec] > 0)
      leafj = s->behind[s->infront[nd->rec] - 1];
    o[0] = nd->level; o[1] = nd->pcount; o[2] = leafj;
    o[3] = nd->pos[0]; o[4] = nd->pos[1]; o[5] = nd->pos[2]; o[6] = nd->width;
    o[7] = nd->cent[0]; o[8] = nd->cent[1]; o[9] = nd->cent[2];
    o[10] = nd->circ[0]; o[11] = nd->circ[1]; o[12] = nd->circ[2];
    /* Cij column-major, as Fortran writes it */
    for (int kk = 0; kk < 3; ++kk)
      for (int m = 0; m < 3; ++m) o[13 + 3 * kk + m] = nd->C[m][kk];
  }
  ++(*k);
  if (nd->pcount > 1)
    for (int c = 0; c < 8; ++c) dump_rec(s, t, nd->child[c], out, max_nodes, k);
}

int64_t qvo_tree_dump(const qvo_params *p, int32_t pcount, const double *x,
                      const int32_t *infront, const int32_t *behind,
                      double *out, int64_t max_nodes) {
  if (!p || pcount < 1 || !(p->box_size > 0.)) return -1;
  qvo_state s;
  if (state_init(&s, p, pcount, x, infront, behind)) { state_free(&s); return -1; }
  qvo_tree t = {0};
  if (construct_tree(&s, &t)) { state_free(&s); free(t.n); return -1; }
  int64_t k = 0;
  dump_rec(&s, &t, 0, out, max_nodes, &k);
  free(t.n);
  state_free(&s);
  return k;
}

/* ------------------------------------------------------------------ */
/* pmotion AB ladder, timestep.f90:48-71                               */

static inline double maxabs3(const double *v) {
  double m = fabs(v[0]);
  if (fabs(v[1]) > m) m = fabs(v[1]);
  if (fabs(v[2]) > m) m = fabs(v[2]);
  return m;
}

int qvo_ab_update(const qvo_params *p, int32_t pcount, double *x,
                  const int32_t *infront, const double *u, double *u1,
                  double *u2, double *u3) {
  const double dt = p->dt;
  const double three_twos = 3. / 2., one_half = 1. / 2.;
  const double twenty_three_twelve = 23. / 12., four_thirds = 4. / 3.,
               five_twelths = 5. / 12.;
  for (int i = 0; i < pcount; ++i) {
    if (infront[i] == 0) continue;
    double *xi = x + 3 * i;
    const double *ui = u + 3 * i;
    double *a1 = u1 + 3 * i, *a2 = u2 + 3 * i, *a3 = u3 + 3 * i;
    if (maxabs3(a1) == 0.) {
      for (int k = 0; k < 3; ++k) xi[k] = xi[k] + dt * ui[k];
    } else if (maxabs3(a2) == 0.) {
      for (int k = 0; k < 3; ++k)
        xi[k] = (xi[k] + (three_twos * dt) * ui[k]) - (one_half * dt) * a1[k];
    } else if (maxabs3(a3) == 0.) {
      for (int k = 0; k < 3; ++k)
        xi[k] = ((xi[k] + (twenty_three_twelve * dt) * ui[k]) -
                 (four_thirds * dt) * a1[k]) + (five_twelths * dt) * a2[k];
    } else {
      for (int k = 0; k < 3; ++k)
        xi[k] = xi[k] + (dt * (((55. * ui[k] - 59. * a1[k]) + 37. * a2[k]) - 9. * a3[k])) / 24.;
    }
    for (int k = 0; k < 3; ++k) {
      a3[k] = a2[k];
      a2[k] = a1[k];
      a1[k] = ui[k];
    }
  }
  if (p->remove_mean_pos) {
    /* timestep.f90:86-92: the mean x and y of the active points (SUM with a
     * mask, a sequential sum in slot order), subtracted from every slot */
    double sx = 0., sy = 0.;
    int32_t n = 0;
    for (int i = 0; i < pcount; ++i) {
      if (infront[i] <= 0) continue;
      sx = sx + x[3 * i];
      sy = sy + x[3 * i + 1];
      ++n;
    }
    const double mx = sx / n, my = sy / n;
    for (int i = 0; i < pcount; ++i) {
      x[3 * i] = x[3 * i] - mx;
      x[3 * i + 1] = x[3 * i + 1] - my;
    }
  }
  return 0;
}

int qvo_enforce_periodic(const qvo_params *p, int32_t pcount, double *x,
                         const int32_t *infront) {
  /* wrap part of enforce_periodic / _yz / _z; the loop removal of the open
   * axes (boundary_loop_remove) is host topology and not restated here */
  int k0 = 3;
  if (p->boundary == QVO_BC_PERIODIC) k0 = 0;
  else if (p->boundary == QVO_BC_OPENX) k0 = 1;
  else if (p->boundary == QVO_BC_OPENXY) k0 = 2;
  const double L = p->box_size;
  for (int i = 0; i < pcount; ++i) {
    if (infront[i] == 0) continue;
    for (int k = k0; k < 3; ++k) {
      if (x[3 * i + k] > (L / 2.)) x[3 * i + k] = x[3 * i + k] - L;
      else if (x[3 * i + k] < (-L / 2.)) x[3 * i + k] = x[3 * i + k] + L;
    }
  }
  return 0;
}

/* distf (general.f90:27-34) */
static double distf(const double *x, int i, int j) {
  return sqrt((sq(x[3 * i] - x[3 * j]) + sq(x[3 * i + 1] - x[3 * j + 1])) +
              sq(x[3 * i + 2] - x[3 * j + 2]));
}

/* boundary_loop_remove (periodic.f90:306-327): walk the loop from
 * `particle`, adding each segment's length but the last one's, clearing
 * every slot (clear_particle, general.f90:11-22) */
static void boundary_loop_remove(int32_t pcount, double *x, int32_t *infront, int32_t *behind,
                                 double *u1, double *u2, int particle, int64_t *count,
                                 double *length) {
  int next = particle;
  *count += 1;
  for (int it = 0; it < pcount; ++it) {
    const int store_next = infront[next] - 1;
    if (store_next < 0) break; /* not a closed loop: the reference reads f(0) */
    if (store_next != particle) *length = *length + distf(x, next, store_next);
    for (int k = 0; k < 3; ++k) {
      x[3 * next + k] = 0.;
      u1[3 * next + k] = 0.;
      u2[3 * next + k] = 0.;
    }
    infront[next] = 0;
    behind[next] = 0;
    next = store_next;
    if (next == particle) break;
  }
}

int qvo_enforce_boundary(const qvo_params *p, int32_t pcount, double *x, int32_t *infront,
                         int32_t *behind, double *u1, double *u2, int64_t *count,
                         double *length) {
  int64_t cnt = 0;
  double len = 0.;
  if (count) cnt = *count;
  if (length) len = *length;
  const double L = p->box_size;
  for (int i = 0; i < pcount; ++i) {
    if (infront[i] == 0) continue;
    double *xi = x + 3 * i;
    if (p->boundary == QVO_BC_PERIODIC) {
      for (int k = 0; k < 3; ++k) {
        if (xi[k] > (L / 2.)) xi[k] = xi[k] - L;
        else if (xi[k] < (-L / 2.)) xi[k] = xi[k] + L;
      }
    } else if (p->boundary == QVO_BC_OPENX) {
      /* enforce_periodic_yz (periodic.f90:217-243): wrap y and z, then x */
      for (int k = 1; k < 3; ++k) {
        if (xi[k] > (L / 2.)) xi[k] = xi[k] - L;
        else if (xi[k] < (-L / 2.)) xi[k] = xi[k] + L;
      }
      if (xi[0] > (L / 2.) || xi[0] < (-L / 2.))
        boundary_loop_remove(pcount, x, infront, behind, u1, u2, i, &cnt, &len);
    } else if (p->boundary == QVO_BC_OPENXY) {
      /* enforce_periodic_z (periodic.f90:247-272): x, then y on the
       * (possibly cleared) point, then wrap z */
      for (int k = 0; k < 2; ++k)
        if (xi[k] > (L / 2.) || xi[k] < (-L / 2.))
          boundary_loop_remove(pcount, x, infront, behind, u1, u2, i, &cnt, &len);
      if (xi[2] > (L / 2.)) xi[2] = xi[2] - L;
      else if (xi[2] < (-L / 2.)) xi[2] = xi[2] + L;
    }
  }
  if (count) *count = cnt;
  if (length) *length = len;
  return 0;
}

/* one segment's velocity at a target: form 0 = biot_savart_shift
 * (timestep.f90:497-505, without the Gamma/(2 pi D) prefactor folded in
 * differently), form 1 = tree_walk's leaf (tree.f90:380-392); both include
 * Gamma.  For unit tests of the pair formulas. */
int qvo_segment_velocity(const qvo_params *p, const double *xi, const double *xj,
                         const double *gj, const double *shift, int32_t form,
                         double *u) {
  u[0] = u[1] = u[2] = 0.;
  if (form == 0) {
    double xs[6] = {xi[0], xi[1], xi[2], xj[0], xj[1], xj[2]};
    double gs[6] = {0., 0., 0., gj[0], gj[1], gj[2]};
    int32_t inf[2] = {2, 1}, beh[2] = {2, 1};
    qvo_state s = {p, 2, xs, inf, beh, gs, gs, gs, gs};
    s.behind = beh;
    bs_pair(&s, 0, 1, shift, u);
    return 0;
  }
  double R1[3], R2[3], ub[3];
  for (int k = 0; k < 3; ++k) {
    R1[k] = (xj[k] + shift[k]) - xi[k];
    R2[k] = (gj[k] + shift[k]) - xi[k];
  }
  const double a = vnorm(R1), b = vnorm(R2);
  const double c = (a * b) * ((a * b) + dot3(R1, R2));
  cross(R1, R2, ub);
  const double f = (a + b) / c;
  for (int k = 0; k < 3; ++k) u[k] = ((f * ub[k] * p->quant_circ) / 4.) / QVO_PI;
  return 0;
}

int qvo_run(const qvo_params *p, int32_t pcount, double *x,
            int32_t *infront, int32_t *behind, double *u1,
            double *u2, double *u3, int32_t nsteps, double *u_first,
            int32_t nthreads, int64_t *bcount, double *blength) {
  double *u = (double *)malloc((size_t)3 * (size_t)pcount * sizeof(double));
  if (!u) return -1;
  for (int step = 0; step < nsteps; ++step) {
    int rc = qvo_velocity(p, pcount, x, infront, behind, u, NULL, NULL, NULL, nthreads);
    if (rc) { free(u); return rc; }
    if (step == 0 && u_first) memcpy(u_first, u, (size_t)3 * pcount * sizeof(double));
    qvo_ab_update(p, pcount, x, infront, u, u1, u2, u3);
    qvo_enforce_boundary(p, pcount, x, infront, behind, u1, u2, bcount, blength);
  }
  free(u);
  return 0;
}

int qvo_run_rk3(const qvo_params *p, int32_t pcount, double *x,
                int32_t *infront, int32_t *behind, double *u1,
                int32_t nsteps, double *u_last, int32_t nthreads) {
  double *u = (double *)malloc((size_t)3 * (size_t)pcount * sizeof(double));
  if (!u) return -1;
  const double dt = p->dt;
  for (int step = 0; step < nsteps; ++step) {
    for (int stage = 0; stage < 3; ++stage) {
      int rc = qvo_velocity(p, pcount, x, infront, behind, u, NULL, NULL, NULL, nthreads);
      if (rc) { free(u); return rc; }
      for (int i = 0; i < pcount; ++i) {
        if (infront[i] == 0) continue;
        double *a = u1 + 3 * i, *xi = x + 3 * i;
        const double *v = u + 3 * i;
        for (int k = 0; k < 3; ++k) {
          if (stage == 0) {                              /* :110, :117 */
            a[k] = v[k] * dt;
            xi[k] = xi[k] + (a[k] / 3.);
          } else if (stage == 1) {                       /* :138, :145 */
            a[k] = -((2. / 3.) * a[k]) + v[k] * dt;
            xi[k] = xi[k] + a[k];
          } else {                                       /* :166, :173 */
            a[k] = -a[k] + v[k] * dt;
            xi[k] = xi[k] + 0.5 * a[k];
          }
        }
      }
    }
    /* the boundary step; u2 is not part of the Runge-Kutta state */
    double *u2 = (double *)calloc((size_t)3 * (size_t)pcount, sizeof(double));
    if (!u2) { free(u); return -1; }
    qvo_enforce_boundary(p, pcount, x, infront, behind, u1, u2, NULL, NULL);
    free(u2);
  }
  if (u_last) memcpy(u_last, u, (size_t)3 * pcount * sizeof(double));
  free(u);
  return 0;
}

/* ------------------------------------------------------------------ */
/* neighbour search: pclose_tree / closest_tree (tree.f90:219-278) and */
/* pclose (line.f90:205-234)                                           */

typedef struct {
  const qvo_tree *t;
  const int32_t *inf_b;            /* links of the build state */
  const double *x;                 /* current state */
  const int32_t *inf, *beh;
} qvo_close;

static void closest_tree(const qvo_close *c, int64_t ni, int i, int32_t *closest,
                         double *closestd) {
  const qvo_node *nd = &c->t->n[ni];
  if (c->inf[i] == 0) return;                    /* :247 */
  if (nd->pcount == 0) return;                   /* :248 */
  const double *xi = c->x + 3 * i;
  const double dist = sqrt((sq(xi[0] - nd->cent[0]) + sq(xi[1] - nd->cent[1])) +
                           sq(xi[2] - nd->cent[2]));   /* :250 */
  const double theta = nd->width / dist;
  if (nd->pcount == 1 || theta < .5) {
    if (dist < *closestd) {
      if (nd->pcount > 1) {
        *closest = 0;                            /* :258, closestd kept */
      } else {
        /* j = f(parray(1)%infront)%behind: build-time infront, current behind */
        const int32_t q = c->inf_b[nd->rec];
        const int32_t j = q > 0 ? c->beh[q - 1] : 0;  /* q = 0 only for a 1-slot root */
        if ((i + 1 != j) && (c->inf[i] != j) && (c->beh[i] != j)) {
          *closest = j;
          *closestd = dist;
        }
      }
    }
  } else {
    for (int k = 0; k < 8; ++k) closest_tree(c, nd->child[k], i, closest, closestd);
  }
}

int qvo_closest_tree(const qvo_params *p, int32_t pcount_b, const double *x_b,
                     const int32_t *infront_b, const int32_t *behind_b,
                     int32_t pcount, const double *x, const int32_t *infront,
                     const int32_t *behind, int32_t *closest, double *closestd,
                     int32_t nthreads) {
  if (!p || pcount_b < 1 || pcount < pcount_b || !(p->box_size > 0.)) return -1;
  qvo_state s;
  if (state_init(&s, p, pcount_b, x_b, infront_b, behind_b)) { state_free(&s); return -1; }
  qvo_tree t = {0};
  if (construct_tree(&s, &t)) { state_free(&s); free(t.n); return -1; }
  qvo_close c = {&t, infront_b, x, infront, behind};
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < pcount; ++i) {
    closestd[i] = 100.;                          /* :227, before the empty test */
    if (infront[i] == 0) continue;
    closest_tree(&c, 0, i, &closest[i], &closestd[i]);
  }
  free(t.n);
  state_free(&s);
  return 0;
}

int qvo_closest_direct(int32_t pcount, const double *x, const int32_t *infront,
                       const int32_t *behind, int32_t *closest, double *closestd,
                       int32_t nthreads) {
  if (pcount < 1) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < pcount; ++i) {
    if (infront[i] == 0) continue;               /* :211, closestd kept */
    double cd = 100.;
    int32_t cl = closest[i];
    const double *xi = x + 3 * i;
    for (int j = 0; j < pcount; ++j) {
      if (infront[j] == 0) continue;
      if ((i != j) && (infront[i] != j + 1) && (behind[i] != j + 1)) {
        const double *xj = x + 3 * j;
        /* distf (general.f90:27-34) */
        const double dist = sqrt((sq(xi[0] - xj[0]) + sq(xi[1] - xj[1])) + sq(xi[2] - xj[2]));
        if (dist < cd) {
          cl = j + 1;
          cd = dist;
        }
      }
    }
    closest[i] = cl;
    closestd[i] = cd;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* velocity at probe points: mesh_velocity (timestep.f90:374-452)      */

/* biot_savart_general(_shift) term of source i at probe y */
static void bs_general(const double *y, const double *xi, const double *gi, const double *shift,
                       double q, double *u) {
  double xs[3], r[3], e[3], ub[3];
  for (int k = 0; k < 3; ++k) xs[k] = shift ? xi[k] + shift[k] : xi[k];
  const double a = dist_gen_sq(y, xs);                       /* :527 / :553 */
  for (int k = 0; k < 3; ++k) r[k] = xs[k] - y[k];
  for (int k = 0; k < 3; ++k) e[k] = gi[k] - xi[k];
  const double b = 2. * dot3(r, e);
  const double c = dist_gen_sq(gi, xi);
  const double D = (4. * a) * c - b * b;
  if (D == 0.) return;                                       /* :531 */
  cross(r, e, ub);
  const double den = (2. * QVO_PI) * D;
  for (int k = 0; k < 3; ++k) ub[k] = (ub[k] * q) / den;
  const double fac = (2. * c + b) / sqrt((a + b) + c) - b / sqrt(a);
  for (int k = 0; k < 3; ++k) u[k] = u[k] + ub[k] * fac;
}

static void walk_general(const qvo_params *p, const qvo_tree *t, int64_t ni, const double *y,
                         const double *shift, int ref_scaling, double *u) {
  const qvo_node *nd = &t->n[ni];
  if (nd->pcount == 0) return;                               /* :436 */
  const double dist = sqrt((sq(y[0] - (nd->cent[0] + shift[0])) +
                            sq(y[1] - (nd->cent[1] + shift[1]))) +
                           sq(y[2] - (nd->cent[2] + shift[2])));
  if (dist < 2.220446049250313e-16) return;                  /* :441 */
  const double theta = nd->width / dist;
  const double q = p->quant_circ;
  if (nd->pcount == 1) {
    double R1[3], R2[3], ub[3];
    for (int k = 0; k < 3; ++k) {
      R1[k] = (nd->cent[k] + shift[k]) - y[k];
      R2[k] = ((nd->cent[k] + nd->circ[k]) + shift[k]) - y[k];
    }
    const double a = vnorm(R1), b = vnorm(R2);
    const double c = (a * b) * ((a * b) + dot3(R1, R2));
    cross(R1, R2, ub);
    const double f = (a + b) / c;
    for (int k = 0; k < 3; ++k) {
      ub[k] = f * ub[k];
      if (ref_scaling) ub[k] = (ub[k] / 4.) / QVO_PI;        /* :460 */
      else ub[k] = ((ub[k] * q) / 4.) / QVO_PI;
      u[k] = u[k] + ub[k];
    }
  } else if (theta < p->tree_theta) {
    double v[3], ub[3];
    for (int k = 0; k < 3; ++k) v[k] = y[k] - (nd->cent[k] + shift[k]);
    cross(nd->circ, v, ub);
    const double d3 = pow3(dist), d5 = pow5(dist);
    for (int k = 0; k < 3; ++k) ub[k] = ub[k] / d3;
    for (int k = 0; k < 3; ++k) {
      for (int l = 0; l < 3; ++l) {
        const double dcl = (v[0] * nd->C[0][l] + v[1] * nd->C[1][l]) + v[2] * nd->C[2][l];
        for (int m = 0; m < 3; ++m) {
          double e = 0.;
          if (!(m == k || k == l || m == l)) e = (((m + 1) % 3 == k) ? 1. : -1.);
          ub[m] = ub[m] - (e * nd->C[k][l]) / d3;
        }
        for (int m = 0; m < 3; ++m) {
          double e = 0.;
          if (!(m == k || k == l || m == l)) e = (((m + 1) % 3 == k) ? 1. : -1.);
          ub[m] = ub[m] + (((3. * e) * v[k]) * dcl) / d5;
        }
      }
    }
    for (int k = 0; k < 3; ++k)                              /* :479 */
      u[k] = ref_scaling ? u[k] + ub[k] : u[k] + ((ub[k] * q) / 4.) / QVO_PI;
  } else {
    for (int c = 0; c < 8; ++c) walk_general(p, t, nd->child[c], y, shift, ref_scaling, u);
  }
}

int qvo_probe_velocity(const qvo_params *p, int32_t pcount_b, const double *x_b,
                       const int32_t *infront_b, const int32_t *behind_b,
                       int32_t pcount, const double *x, const int32_t *infront,
                       const double *gi, int32_t nprobe, const double *probe,
                       int32_t ref_scaling, double *u, int32_t nthreads) {
  if (!p || nprobe < 0 || !(p->box_size > 0.)) return -1;
  double sh[27][3];
  const int ns = shell_shifts(p, sh);
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (p->velocity == QVO_VEL_BS) {
    if (pcount < 1) return -1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int q = 0; q < nprobe; ++q) {
      double v[3] = {0., 0., 0.};
      for (int s = 0; s < ns; ++s)
        for (int i = 0; i < pcount; ++i) {
          if (infront[i] == 0) continue;                     /* :526 */
          bs_general(probe + 3 * q, x + 3 * i, gi + 3 * i, s ? sh[s] : NULL, p->quant_circ, v);
        }
      for (int k = 0; k < 3; ++k) u[3 * q + k] = v[k];
    }
    return 0;
  }
  if (pcount_b < 1) return -1;
  qvo_state st;
  if (state_init(&st, p, pcount_b, x_b, infront_b, behind_b)) { state_free(&st); return -1; }
  qvo_tree t = {0};
  if (construct_tree(&st, &t)) { state_free(&st); free(t.n); return -1; }
#pragma omp parallel for schedule(dynamic, 16)
  for (int q = 0; q < nprobe; ++q) {
    double v[3] = {0., 0., 0.};
    for (int s = 0; s < ns; ++s) walk_general(p, &t, 0, probe + 3 * q, sh[s], ref_scaling, v);
    for (int k = 0; k < 3; ++k) u[3 * q + k] = v[k];
  }
  free(t.n);
  state_free(&st);
  return 0;
}
