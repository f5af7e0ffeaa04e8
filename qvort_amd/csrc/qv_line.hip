// qv_line.hip -- the line operations of a step on host arrays: pinsert
// (line.f90:18-108), precon_original (reconnection.f90:145-256), premove
// (line.f90:114-161) and loop_killer (reconnection.f90:421-461).
//
// Host code.  These loops visit the slots in order and each change of the
// links is seen by the slots visited after it, so the result depends on the
// order (which slot a new point takes, which reconnection happens first);
// they run sequentially, in the reference's order, and keep its FP64
// operation order (no contraction), so that the state after each operation
// is the reference's bit for bit.  Slot indices are 0-based here, links
// 1-based as in f(:).
//
// What a slot's test reads before the loop reaches it is computed first, in
// parallel (pinsert: each segment's length; premove: whether the point two
// ahead is within 0.499 (delta_i + delta_j)), and the sequential loop takes
// that value unless an earlier slot's operation changed one of its inputs
// (the slot's position, ghost or link, or those of the slots it reads),
// which it marks in a flag per slot; then it evaluates the test again, with
// the same expression.  The validation of the links is parallel too.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "qv_internal.h"
#include "qv_lia.h"
#include "qvort_gpu.h"

#pragma clang fp contract(off)

namespace qv {
namespace {

// per-thread scratch of the pre-pass, kept between calls
thread_local std::vector<double> tl_pre;
thread_local std::vector<uint8_t> tl_flag, tl_dirty;

struct Line {
  const qv_params &p;
  qv_line_state &s;
  qv_line_counts &c;
  bool per[3];
  uint8_t *dirty = nullptr;  // slots whose inputs changed in this call

  Line(const qv_params &p_, qv_line_state &s_, qv_line_counts &c_) : p(p_), s(s_), c(c_) {
    per[0] = p.boundary == QV_BC_PERIODIC;
    per[1] = p.boundary == QV_BC_PERIODIC || p.boundary == QV_BC_OPENX;
    per[2] = p.boundary != QV_BC_OPEN;
  }
  double *X(int i) { return s.x + 3 * (size_t)i; }
  double *U(int i) { return s.u + 3 * (size_t)i; }
  double *GI(int i) { return s.ghosti + 3 * (size_t)i; }
  double *GB(int i) { return s.ghostb + 3 * (size_t)i; }
  double *GII(int i) { return s.ghostii + 3 * (size_t)i; }
  double *GBB(int i) { return s.ghostbb + 3 * (size_t)i; }
  int inf(int i) const { return s.infront[i] - 1; }
  int beh(int i) const { return s.behind[i] - 1; }

  // get_ghost_p (periodic.f90:23-184): the neighbours' positions, each
  // moved by one box length where it lies more than half a box away from
  // x(i) along a periodic axis
  void ghost_p(int i) {
    if (dirty) dirty[i] = 1;
    const int a = inf(i), b = beh(i), aa = inf(a), bb = beh(b);
    const double *x = X(i);
    const int src[4] = {a, b, aa, bb};
    double *dst[4] = {GI(i), GB(i), GII(i), GBB(i)};
    const double L = p.box_size;
    for (int q = 0; q < 4; ++q) {
      const double *g = X(src[q]);
      for (int k = 0; k < 3; ++k) {
        double v = g[k];
        if (per[k]) {
          if ((x[k] - v) > (L / 2.)) v = v + L;
          else if ((x[k] - v) < (-L / 2.)) v = v - L;
        }
        dst[q][k] = v;
      }
    }
  }

  void derivs(int i, double s1[3], double s2[3]) {
    lia_derivs<double>(p.deriv_order, X(i), GI(i), GB(i), GII(i), GBB(i), s1, s2);
  }

  // distf (general.f90:27-34)
  double distf(int i, int j) {
    const double *a = X(i), *b = X(j);
    return std::sqrt(((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1])) +
                     (a[2] - b[2]) * (a[2] - b[2]));
  }

  // clear_particle (general.f90:11-22); u3 is left as it is
  void clear(int i) {
    if (dirty) dirty[i] = 1;
    for (double *v : {X(i), U(i), s.u1 + 3 * (size_t)i, s.u2 + 3 * (size_t)i, GI(i), GB(i),
                      GII(i), GBB(i)})
      v[0] = v[1] = v[2] = 0.;
    s.infront[i] = s.behind[i] = 0;
    s.closest[i] = 0;
    s.closestd[i] = 0.;
  }

  // loop_killer (reconnection.f90:421-461): a loop of fewer than 5 points
  // is removed.  On a slot that is already empty the reference follows the
  // link 0 out of the array; here that call does nothing.
  // A chain that reaches an empty slot (not a closed loop, which the
  // operations never leave behind) is not followed past it.
  void loop_killer(int particle) {
    if (s.infront[particle] == 0) return;
    // the count only decides whether the loop has fewer than 5 points (the
    // reference also logs loops of 5-40 points), so it stops at 5
    int counter = 1, next = particle;
    for (int it = 0; it < s.pcount && counter < 5; ++it) {
      if (s.infront[next] == 0) return;
      next = inf(next);
      if (next == particle) break;
      ++counter;
    }
    if (counter >= 5) return;
    next = particle;
    if (counter > 2) c.small_loop_remove_count += 1;
    for (int it = 0; it < s.pcount; ++it) {
      const int store_next = inf(next);
      if (store_next < 0) {
        clear(next);
        break;
      }
      if (counter > 2 && store_next != particle)
        c.small_loop_remove_length = c.small_loop_remove_length + distf(next, store_next);
      clear(next);
      next = store_next;
      if (next == particle) break;
    }
  }

  int pinsert() {
    int nact = 0;
    for (int i = 0; i < s.pcount; ++i) nact += s.infront[i] > 0;
    if (nact == 0) return QV_ERR_LINE_EMPTY;
    const int old = s.pcount;
    // each active slot's segment length (the test's input), in parallel
    const bool prepass = host_prepass();
    std::vector<double> &pre = tl_pre;
    pre.resize(old);
    if (prepass)
      par_for(old, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i)
          if (s.infront[i] != 0) pre[i] = lia_dist(X(i), GI(i));
      });
    // written by ghost_p (the ghosts, and the new point's position); every
    // slot marked without the pre-pass
    tl_dirty.assign((size_t)s.capacity, prepass ? 0 : 1);
    dirty = tl_dirty.data();
    double total = 0.;
    // the first empty slot (minloc of infront): empty slots are only filled
    // here, never made, so it only moves forward and one cursor finds it
    int hole = 0;
    for (int i = 0; i < old; ++i) {
      if (s.infront[i] == 0) continue;
      const double disti = dirty[i] ? lia_dist(X(i), GI(i)) : pre[i];
      total = total + disti;
      // the point's delta factor is 1 (delta_adapt F)
      if (!(disti > 1. * p.delta)) continue;
      // the first empty slot, else one more slot
      while (hole < s.pcount && s.infront[hole] != 0) ++hole;
      int nw = hole < s.pcount ? hole : -1;
      if (nw < 0) {
        if (s.pcount >= s.capacity) return QV_ERR_CAPACITY;
        nw = s.pcount++;
        s.infront[nw] = 0;
        s.u3[3 * (size_t)nw] = s.u3[3 * (size_t)nw + 1] = s.u3[3 * (size_t)nw + 2] = 0.;
        s.closest[nw] = 0;
        s.closestd[nw] = 0.;
      }
      double s1[3], fdd[3];
      derivs(i, s1, fdd);
      double curv = std::sqrt((fdd[0] * fdd[0] + fdd[1] * fdd[1]) + fdd[2] * fdd[2]);
      const double *xi = X(i), *gi = GI(i);
      double *xn = X(nw);
      bool linear = true;
      if (curv > 1e-5) {
        curv = 1. / curv;
        if (curv * curv - 0.25 * (disti * disti) > 0.) {
          const double f = (std::sqrt(curv * curv - 0.25 * (disti * disti)) - curv) * curv;
          for (int k = 0; k < 3; ++k) xn[k] = 0.5 * (xi[k] + gi[k]) + f * fdd[k];
          linear = false;
        }
      }
      if (linear)
        for (int k = 0; k < 3; ++k) xn[k] = 0.5 * (xi[k] + gi[k]);
      const int a = inf(i);
      for (int k = 0; k < 3; ++k) U(nw)[k] = 0.5 * (U(i)[k] + U(a)[k]);
      for (int k = 0; k < 3; ++k) s.u1[3 * (size_t)nw + k] = s.u2[3 * (size_t)nw + k] = 0.;
      s.behind[nw] = i + 1;
      s.infront[nw] = s.infront[i];
      ghost_p(nw);
      s.behind[a] = nw + 1;
      ghost_p(a);
      s.infront[i] = nw + 1;
      ghost_p(i);
      c.inserted += 1;
    }
    int empty = 0;
    for (int k = 0; k < s.pcount; ++k) empty += s.infront[k] == 0;
    c.total_length = total;
    c.avg_sep = total / (double)(old - empty);
    return QV_OK;
  }

  // norm_tanf (general.f90:94-107)
  void norm_tan(int i, double t[3]) {
    double s2[3];
    derivs(i, t, s2);
    const double len = std::sqrt((t[0] * t[0] + t[1] * t[1]) + t[2] * t[2]);
    for (int k = 0; k < 3; ++k) t[k] = t[k] / len;
  }

  int precon() {
    const double eps = 2.220446049250313e-16;  // epsilon(1.) in FP64
    for (int i = 0; i < s.pcount; ++i) {
      if (s.infront[i] == 0) continue;
      const int pari = inf(i), parb = beh(i);
      if (!(s.closestd[i] < p.delta / 2. && s.closestd[i] > eps)) continue;
      const int j1 = s.closest[i];  // 1-based, 0 = none
      if (j1 == pari + 1 || j1 == parb + 1 || j1 == 0) continue;
      const int j = j1 - 1;
      if (s.infront[j] == 0) continue;
      const int parjb = beh(j);
      double t1[3], t2[3];
      norm_tan(i, t1);
      norm_tan(j, t2);
      const double dot = (t1[0] * t2[0] + t1[1] * t2[1]) + t1[2] * t2[2];
      if (dot > 0.9) continue;  // parallel filaments do not reconnect
      const double lb = ((lia_dist(X(i), GI(i)) + lia_dist(X(i), GB(i))) +
                         lia_dist(X(j), GI(j))) + lia_dist(X(j), GB(j));
      const double la = ((lia_dist(X(i), GB(i)) + lia_dist(X(i), X(j))) +
                         lia_dist(X(j), GI(j))) + lia_dist(X(pari), X(parjb));
      if (!(la <= lb)) continue;
      c.recon_count += 1;
      s.infront[parjb] = pari + 1;
      s.behind[pari] = parjb + 1;
      s.infront[i] = j + 1;
      s.behind[j] = i + 1;
      loop_killer(pari);
      loop_killer(i);
    }
    return QV_OK;
  }

  // premove's test for slot i with the current links (both points' delta
  // factors are 1: delta_adapt F)
  bool premove_test(int i) {
    const int aa = inf(inf(i));
    const double distii = distf(i, aa);
    return distii < 0.499 * p.delta * (1. + 1.);
  }

  int premove() {
    const int n = s.pcount;
    const bool prepass = host_prepass();
    std::vector<uint8_t> &hit = tl_flag;
    hit.resize(n);
    if (prepass)
      par_for(n, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) hit[i] = s.infront[i] != 0 && premove_test(i);
      });
    // the test reads infront(i), infront(a), x(i) and x(aa): a change of
    // any of them marks its slot (the link set below, clear); every slot
    // marked without the pre-pass
    tl_dirty.assign((size_t)n, prepass ? 0 : 1);
    dirty = tl_dirty.data();
    for (int i = 0; i < n; ++i) {
      if (s.infront[i] == 0) continue;
      const int a = inf(i), aa = inf(a);
      const bool close = (dirty[i] || dirty[a] || dirty[aa]) ? premove_test(i) : hit[i] != 0;
      if (!close) continue;
      s.behind[aa] = i + 1;
      s.infront[i] = aa + 1;
      dirty[i] = 1;
      clear(a);
      loop_killer(i);
      c.remove_count += 1;
    }
    return QV_OK;
  }
};

}  // namespace

// The boundary step of an openx / openxy box on host arrays, slot by slot
// as enforce_periodic_yz / _z run (periodic.f90:217-276): a point's periodic
// axes are wrapped before its open axes are tested (openx: y, z, then x;
// openxy: x, then y, then z is wrapped), and a point beyond an open face
// removes its whole loop at once (boundary_loop_remove, periodic.f90:306-327).
// The removal's length counter adds the segments of the loop with the
// positions of the moment -- its points before the slot already wrapped,
// those after it not yet -- so the order is part of the result.
void boundary_step_host(const qv_params &p, int32_t pcount, double *x, int32_t *inf,
                        int32_t *beh, int64_t &count, double &length, int32_t *cleared,
                        int32_t &ncleared) {
  const double L = p.box_size;
  const bool openx = p.boundary == QV_BC_OPENX;
  auto wrap = [L](double &v) {
    if (v > (L / 2.)) v = v - L;
    else if (v < (-L / 2.)) v = v + L;
  };
  auto outside = [L](double v) { return v > (L / 2.) || v < (-L / 2.); };
  // the loop through `start`, cleared slot by slot (clear_particle's x and
  // links; the device zeroes the rest of the cleared slots)
  auto remove_loop = [&](int start) {
    count += 1;
    int cur = start;
    for (int32_t it = 0; it < pcount; ++it) {
      const int nxt = inf[cur] - 1;
      if (nxt < 0) break;  // not a closed loop; the reference would read f(0)
      if (nxt != start) {
        const double *a = x + 3 * (size_t)cur, *b = x + 3 * (size_t)nxt;
        length = length + std::sqrt(((a[0] - b[0]) * (a[0] - b[0]) +
                                     (a[1] - b[1]) * (a[1] - b[1])) +
                                    (a[2] - b[2]) * (a[2] - b[2]));
      }
      double *c = x + 3 * (size_t)cur;
      c[0] = c[1] = c[2] = 0.;
      inf[cur] = beh[cur] = 0;
      cleared[ncleared++] = cur;
      cur = nxt;
      if (cur == start) break;
    }
  };
  for (int i = 0; i < pcount; ++i) {
    if (inf[i] == 0) continue;
    double *xi = x + 3 * (size_t)i;
    if (openx) {
      wrap(xi[1]);
      wrap(xi[2]);
      if (outside(xi[0])) remove_loop(i);
    } else {
      // a cleared point has x = 0 and passes the later tests
      if (outside(xi[0])) remove_loop(i);
      if (outside(xi[1])) remove_loop(i);
      wrap(xi[2]);
    }
  }
}

namespace {

// the arrays the operations index must be present and the links must stay
// inside the slots in use
int check(const qv_params *p, qv_line_state *s, qv_line_counts *c) {
  if (!p || !s || !c) return QV_ERR_ARG;
  if (s->pcount < 0 || s->capacity < s->pcount) return QV_ERR_ARG;
  if (!s->x || !s->u || !s->u1 || !s->u2 || !s->u3 || !s->ghosti || !s->ghostb ||
      !s->ghostii || !s->ghostbb || !s->infront || !s->behind || !s->closest || !s->closestd)
    return QV_ERR_ARG;
  if (p->deriv_order != 2 && p->deriv_order != 4) return QV_ERR_ARG;
  if (p->boundary < QV_BC_OPEN || p->boundary > QV_BC_OPENXY || !(p->box_size > 0.))
    return QV_ERR_ARG;
  std::atomic<bool> bad{false};
  const int n = s->pcount;
  par_for(n, [&](int lo, int hi) {
    bool ok = true;
    for (int i = lo; i < hi && ok; ++i) {
      const int a = s->infront[i], b = s->behind[i];
      if (a < 0 || a > n || b < 0 || b > n) ok = false;
      else if ((a == 0) != (b == 0)) ok = false;
      else if (a != 0 && (s->infront[a - 1] == 0 || s->behind[b - 1] == 0)) ok = false;
      else if (s->closest[i] < 0 || s->closest[i] > n) ok = false;
    }
    if (!ok) bad.store(true, std::memory_order_relaxed);
  });
  return bad.load() ? QV_ERR_ARG : QV_OK;
}

}  // namespace
}  // namespace qv

extern "C" {

int qv_line_pinsert(const qv_params *params, qv_line_state *s, qv_line_counts *c) {
  const int rc = qv::check(params, s, c);
  if (rc) return rc;
  return qv::Line(*params, *s, *c).pinsert();
}

int qv_line_precon(const qv_params *params, qv_line_state *s, qv_line_counts *c) {
  const int rc = qv::check(params, s, c);
  if (rc) return rc;
  return qv::Line(*params, *s, *c).precon();
}

int qv_line_premove(const qv_params *params, qv_line_state *s, qv_line_counts *c) {
  const int rc = qv::check(params, s, c);
  if (rc) return rc;
  return qv::Line(*params, *s, *c).premove();
}

}  // extern "C"
