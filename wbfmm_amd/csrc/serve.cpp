// serve.cpp -- the serving engine declared in wbfmm_serve.h: a queue of
// independent FMM solves run by worker threads, one HIP stream each.
//
// A worker runs the reference's call sequence for the request's problem
// (tools/wbfmm-laplace-tree-calc.c.in:380-460 for Laplace,
// tools/wbfmm-tree-calc.c.in:306-417 for Helmholtz) through the drop-in
// API with its own stream selected (wbfmm_hip_set_stream), so the solves
// of different workers overlap on the GPU.  Operator tables are kept per
// worker: Laplace sets by order, and for Helmholtz one set per largest
// order whose coaxial tables are rebuilt when the wavenumber or box width
// changes (the rotation tables depend on the order only).

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "wbfmm_serve.h"

// the host layer's recovery point (wbfmm_api.c): runs fn(arg); a failure
// that the drop-in API would turn into an abort returns its status here
// instead (message in msg), with the library's locks released
extern "C" int wbfmm_guarded_call(int (*fn)(void *), void *arg, char *msg, size_t msglen);

namespace {

using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// output doubles per target for a field request (0: invalid)
int field_width(int problem, unsigned field) {
  const bool p = field & WBFMM_FIELD_POTENTIAL, g = field & WBFMM_FIELD_GRADIENT;
  if (field & ~(unsigned)(WBFMM_FIELD_POTENTIAL | WBFMM_FIELD_GRADIENT)) return 0;
  const int c = problem == WBFMM_PROBLEM_HELMHOLTZ ? 2 : 1;
  return c * ((p ? 1 : 0) + (g ? 3 : 0));
}

// Laplace order for a relative accuracy.  The error of the octree FMM with
// one box of separation falls by a factor between 0.3 and 0.55 per order,
// depending on the depth and the point distribution: 4.6e-7 (potential) at
// order 12 on 1e6 uniform points at depth 7, 6.9e-9 at order 17 on 6e4
// points at depth 4, 1.9e-8 at order 17 on 2192 points at depth 2.  The
// order is the larger of the estimates from either rate, plus one.
int laplace_order(double tol) {
  if (!(tol > 0.0) || tol >= 1.0) return 2;
  const int p_fast = (int)std::ceil(std::log(tol) / std::log(0.3)) + 1;
  const int p_slow = 12 + (int)std::ceil(std::log(tol / 4.6e-7) / std::log(0.55)) + 1;
  const int p = std::max(p_fast, p_slow);
  return std::min(std::max(p, 2), WBFMM_HIP_MAX_ORDER);
}

struct Plan {
  int precision = 64;
  int depth = 0;
  // kernel variants set on the worker's thread for the solve (-1: the
  // library's choice); from the autotuner's table
  int var[WBFMM_VARIANT_COUNT] = {-1, -1};
  int orders[WBFMM_TREE_MAX_DEPTH + 1] = {0};
  int lmax = 0;
  double x0[3] = {0, 0, 0};
  double D = 1.0;
  size_t footprint = 0;
};

struct Job {
  int64_t id = 0;
  wbfmm_serve_request_t req{};
  Plan plan;
  clk::time_point t_submit, t_start, t_done;
  wbfmm_serve_result_t res{};
  bool done = false;
  bool tune = false;  // an autotuner run: not counted in the statistics
  int fail_allocs = 0;  // device allocations forced to fail (tests)
};

// Helmholtz operator sets stay resident in an engine-wide cache, one per
// (largest order, k, domain width, depth, orders of the levels): the
// coaxial tables depend on k and the box widths, so requests of another
// wavenumber or domain need their own set.  A set is completed by the
// first solve that uses it (the passes build some tiles on first use), so
// a worker holds the set's `use` lock for its whole solve; solves end with
// a stream synchronisation, so the next holder, on any stream, finds the
// tables complete.
// the autotuner's choice for a request class
struct Tuned {
  int depth = 0;
  int var[WBFMM_VARIANT_COUNT] = {-1, -1};
};

struct HelKey {
  int lmax = 0, depth = 0;
  double k = 0.0, D = 0.0;
  std::array<int, WBFMM_TREE_MAX_DEPTH + 1> orders{};
  bool operator<(const HelKey &o) const {
    return std::tie(lmax, depth, k, D, orders) < std::tie(o.lmax, o.depth, o.k, o.D, o.orders);
  }
};

struct HelEntry {
  std::mutex use;
  wbfmm_shift_operators_t *op = nullptr;
  bool built = false;  // coaxial tables of every level present
  size_t bytes = 0;    // estimate (the tables term of the footprint)
  uint64_t last = 0;   // last use (engine tick), for eviction
};

struct Worker {
  std::thread th;
  void *stream = nullptr;
  std::map<int, wbfmm_shift_operators_t *> lap_ops;  // by order
};

}  // namespace

struct wbfmm_serve_s {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job *> queue;
  std::map<int64_t, std::unique_ptr<Job>> jobs;
  std::vector<std::unique_ptr<Worker>> workers;
  size_t budget = 0, inflight = 0, peak = 0;
  int fail_allocs = 0;  // for the next request: allocations forced to fail
  int64_t next_id = 0;
  bool stop = false;
  // statistics since the last reset
  std::vector<double> lat;
  uint64_t targets = 0;
  bool have_first = false;
  clk::time_point first_submit, last_done;
  // autotuned depth and kernel variants per request class:
  // (problem, precision, order, log2 n) -> Tuned
  std::map<std::tuple<int, int, int, int>, Tuned> tuned;
  // resident Helmholtz operator sets (at most hel_cap bytes of them idle)
  std::map<HelKey, std::shared_ptr<HelEntry>> hel_cache;
  size_t hel_bytes = 0, hel_cap = 0;
  uint64_t tick = 0;
};

namespace {

std::tuple<int, int, int, int> tune_key(const wbfmm_serve_request_t &r, int precision,
                                        int order) {
  const unsigned n = std::max(r.nsrc, 1u);
  return std::make_tuple(r.problem, precision, order, (int)std::floor(std::log2((double)n)));
}

// FP32 (Laplace at orders 2..15; Helmholtz in mixed precision, FP32 at
// the low-frequency levels) when asked for, or by default when the
// accuracy asked for allows it: 1e-5 or looser for Laplace (every level in
// FP32), 1e-6 or looser for Helmholtz (FP32 only where k a <= 2, the near
// field in FP32 with compensated sums); FP64 otherwise
int choose_precision(const wbfmm_serve_request_t &r, int lap_order) {
  const bool fp32_ok = r.problem == WBFMM_PROBLEM_HELMHOLTZ ||
                       (r.problem == WBFMM_PROBLEM_LAPLACE && lap_order >= 2 && lap_order <= 15);
  if (r.precision == 32) return fp32_ok ? 32 : 0;
  if (r.precision == 64) return 64;
  const double loosest = r.problem == WBFMM_PROBLEM_HELMHOLTZ ? 1e-6 : 1e-5;
  return fp32_ok && r.order == 0 && r.tol >= loosest ? 32 : 64;
}

// device bytes of a solve:
//  - dense coefficient arrays of every level (as the tree allocates them)
//    and their box ranges and flags;
//  - sorted points, strengths, codes and indices; the targets' staging and
//    the engine's field buffer; the check's copies;
//  - the largest transient of one pass: the M2L occupancy bits and class
//    list of a level, the partial expansions of a level whose offsets are
//    split (fewer than 1024 chunks of 16 targets: up to 8 copies of at most
//    16384 targets' expansions), and the radix sort's temporaries;
//  - the operator tables a worker builds for this order (rotation and
//    coaxial tiles; kept by the worker and reused by later requests)
// tables: rotation tiles of the order (49 angles, one 64-double tile per
// k-step, KT <= (N+1)^2 / 2), coaxial tiles of every level (15 lengths,
// complex for Helmholtz) and the Helmholtz host-side coaxial tables
size_t table_bytes(int problem, const Plan &p) {
  const size_t cplx = problem == WBFMM_PROBLEM_HELMHOLTZ ? 2 : 1;
  const size_t N = (size_t)p.lmax;
  const size_t kt = (N + 1) * (N + 1) / 2 + 8;
  size_t tables = 49 * kt * 64 * 8 * 2 + (size_t)p.depth * 15 * 2 * kt * 64 * 8 * cplx;
  if (cplx == 2) tables += (size_t)p.depth * 15 * 2 * (2 * N + 1) * (2 * N + 2) * (2 * N + 3) / 6 * 8;
  return tables;
}

size_t footprint(const wbfmm_serve_request_t &r, const Plan &p) {
  const size_t cplx = r.problem == WBFMM_PROBLEM_HELMHOLTZ ? 2 : 1;
  size_t b = 0, transient = 0;
  for (int l = 1; l <= p.depth; l++) {
    const size_t nb = (size_t)1 << (3 * l);
    const size_t C = (size_t)(p.orders[l] + 1) * (p.orders[l] + 1);
    // mixed-precision Helmholtz keeps the high-frequency levels in FP64
    const double ka = r.k * 0.5 * std::sqrt(3.0) * p.D / (double)(1u << l);
    const size_t es = p.precision == 32 && (cplx == 1 || ka <= 2.0) ? 4 : 8;
    b += nb * (2 * C * es * cplx + 16);
    const size_t split_targets = std::min<size_t>(nb, 16384);
    const size_t level_tr = nb / 8 + 4 * nb + 8 * split_targets * C * es * cplx;
    transient = std::max(transient, level_tr);
  }
  const size_t nt = r.xt ? r.ntgt : r.nsrc;
  b += (size_t)r.nsrc * (2 * 32 + 12 + 16 * cplx + (r.normals ? 32 : 0));
  b += nt * (32 + 12 + 8 * (size_t)std::max(r.fstr, 1));
  if (r.check) b += (size_t)r.nsrc * (32 + 16 * cplx);
  transient = std::max(transient, (size_t)std::max(r.nsrc, (unsigned)nt) * 24);
  return b + transient + table_bytes(r.problem, p) + ((size_t)16 << 20);
}

int validate(const wbfmm_serve_request_t &r) {
  if (r.problem != WBFMM_PROBLEM_LAPLACE && r.problem != WBFMM_PROBLEM_HELMHOLTZ)
    return WBFMM_HIP_ERR_ARG;
  if (r.precision != 0 && r.precision != 32 && r.precision != 64) return WBFMM_HIP_ERR_ARG;
  if (r.nsrc == 0 || !r.xs || r.pstr < 3 * sizeof(double)) return WBFMM_HIP_ERR_ARG;
  if (!r.q && !r.dip) return WBFMM_HIP_ERR_ARG;
  if (r.dip && (!r.normals || r.nstr < 3 * sizeof(double))) return WBFMM_HIP_ERR_ARG;
  if (r.xt && r.tstr < 3 * sizeof(double)) return WBFMM_HIP_ERR_ARG;
  const int fw = field_width(r.problem, r.field);
  if (fw == 0 || r.fstr < fw || !r.f) return WBFMM_HIP_ERR_ARG;
  if (r.problem == WBFMM_PROBLEM_HELMHOLTZ && !(r.k > 0.0)) return WBFMM_HIP_ERR_ARG;
  if (r.order < 0 || r.order > WBFMM_HIP_MAX_ORDER) return WBFMM_HIP_ERR_ORDER;
  if (r.depth < 0 || r.depth > 10) return WBFMM_HIP_ERR_ARG;
  if (r.order == 0 && !(r.tol > 0.0 && r.tol < 1.0)) return WBFMM_HIP_ERR_ARG;
  return 0;
}

// default depth: about 16 (Laplace) or 32 (Helmholtz) points per occupied
// leaf, counting the leaves a uniform cloud would fill
int default_depth(const wbfmm_serve_request_t &r) {
  const double per = r.problem == WBFMM_PROBLEM_HELMHOLTZ ? 32.0 : 16.0;
  const double n = (double)std::max(r.nsrc, r.xt ? r.ntgt : 0u);
  int d = (int)std::lround(std::log(std::max(n / per, 1.0)) / std::log(8.0));
  return std::min(std::max(d, 2), 8);
}

// domain, depth and orders; runs on the submitting thread (the bounds of
// device points are reduced on its current stream)
int make_plan(wbfmm_serve_t *s, const wbfmm_serve_request_t &r, Plan &p) {
  int rc = validate(r);
  if (rc) return rc;
  double lo[3], hi[3], lt[3], ht[3];
  rc = wbfmm_hip_bounds(r.xs, r.pstr, r.nsrc, lo, hi);
  if (rc) return rc;
  if (r.xt && r.ntgt) {
    rc = wbfmm_hip_bounds(r.xt, r.tstr, r.ntgt, lt, ht);
    if (rc) return rc;
    for (int c = 0; c < 3; c++) {
      lo[c] = std::min(lo[c], lt[c]);
      hi[c] = std::max(hi[c], ht[c]);
    }
  }
  double e = 0.0;
  for (int c = 0; c < 3; c++) e = std::max(e, hi[c] - lo[c]);
  if (!(e > 0.0) || !std::isfinite(e)) e = 1.0;
  // the reference drivers take the bounding cube widened by 0.01 on each
  // side (tools/wbfmm-tree-calc.c.in:261-267); here 1% of the width
  for (int c = 0; c < 3; c++) p.x0[c] = lo[c] - 0.01 * e;
  p.D = 1.02 * e;

  const int order_fixed = r.order;
  const int lap_order = order_fixed ? order_fixed : laplace_order(r.tol);
  p.precision = choose_precision(r, lap_order);
  if (p.precision == 0) return WBFMM_HIP_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tuned.find(
        tune_key(r, p.precision, r.problem == WBFMM_PROBLEM_LAPLACE ? lap_order : order_fixed));
    const bool have = it != s->tuned.end();
    p.depth = r.depth ? r.depth : (have && it->second.depth > 0 ? it->second.depth : default_depth(r));
    if (have)
      for (int v = 0; v < WBFMM_VARIANT_COUNT; v++) p.var[v] = it->second.var[v];
  }
  p.lmax = 0;
  for (int l = 0; l <= p.depth; l++) {
    int o;
    if (order_fixed) {
      o = order_fixed;
    } else if (r.problem == WBFMM_PROBLEM_LAPLACE) {
      o = lap_order;
    } else {
      // Gumerov & Duraiswami's truncation number of the level
      // (src/util.c:474-503) for the level's box width
      wbfmm_tree_t tt{};
      tt.D = p.D;
      o = wbfmm_truncation_number(&tt, r.k, (guint)std::max(l, 2), r.tol);
    }
    p.orders[l] = std::min(std::max(o, 2), WBFMM_HIP_MAX_ORDER);
    if (l >= 2) p.lmax = std::max(p.lmax, p.orders[l]);
  }
  p.orders[1] = p.orders[2];
  p.footprint = footprint(r, p);
  return 0;
}

struct OpsNew {
  guint L;
  wbfmm_shift_operators_t *op;
};

int ops_new_body(void *a) {
  auto *x = static_cast<OpsNew *>(a);
  x->op = wbfmm_shift_operators_new(x->L, FALSE, nullptr);
  return x->op ? 0 : WBFMM_HIP_ERR_NOMEM;
}

// the worker's Laplace operator set of this order (guarded: a failure
// returns a status)
int laplace_ops(Worker &w, int lmax, wbfmm_shift_operators_t **out, char *msg, size_t ml) {
  auto it = w.lap_ops.find(lmax);
  if (it != w.lap_ops.end()) {
    *out = it->second;
    return 0;
  }
  OpsNew x{(guint)lmax, nullptr};
  const int rc = wbfmm_guarded_call(ops_new_body, &x, msg, ml);
  if (rc) return rc;
  w.lap_ops[lmax] = x.op;
  *out = x.op;
  return 0;
}

struct CoaxJob {
  wbfmm_shift_operators_t *op;
  void *stream;
  guint lsr, lss;
  int level;
  double D, k;
  int rc;
  char msg[256];
};

int coax_body(void *a) {
  auto *c = static_cast<CoaxJob *>(a);
  wbfmm_shift_operators_coaxial_SR_init(c->op, c->D, (guint)c->level, c->lsr, c->k, nullptr);
  if (c->level >= 3)
    wbfmm_shift_operators_coaxial_SS_init(c->op, c->D, (guint)c->level, c->lss, c->k, nullptr);
  return 0;
}

// Helmholtz operator set of this plan from the engine's cache, locked for
// the caller's solve (`held`); built when missing: *tables_ms gets the
// milliseconds spent on tables.  Coaxial tables of the levels are built in
// parallel host threads that share the worker's stream.  Failures (in any
// of the threads) return a status.
int helmholtz_ops(wbfmm_serve_t *s, Worker &w, const Plan &p, double k, double *tables_ms,
                  std::shared_ptr<HelEntry> &ent, std::unique_lock<std::mutex> &held,
                  char *msg, size_t ml) {
  const auto t0 = clk::now();
  HelKey key;
  key.lmax = p.lmax;
  key.depth = p.depth;
  key.k = k;
  key.D = p.D;
  for (int l = 0; l <= p.depth; l++) key.orders[(size_t)l] = p.orders[l];
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto &e = s->hel_cache[key];
    if (!e) {
      e = std::make_shared<HelEntry>();
      e->bytes = table_bytes(WBFMM_PROBLEM_HELMHOLTZ, p);
      s->hel_bytes += e->bytes;
    }
    e->last = ++s->tick;
    ent = e;
  }
  held = std::unique_lock<std::mutex>(ent->use);
  HelEntry &h = *ent;
  if (!h.op) {
    OpsNew x{(guint)p.lmax, nullptr};
    const int rc = wbfmm_guarded_call(ops_new_body, &x, msg, ml);
    if (rc) return rc;
    h.op = x.op;
  }
  if (!h.built) {
    wbfmm_shift_operators_coaxial_reset(h.op);  // after a failed build
    std::vector<CoaxJob> jobs((size_t)std::max(p.depth - 1, 0));
    for (int l = 2; l <= p.depth; l++) {
      CoaxJob &c = jobs[(size_t)(l - 2)];
      c.op = h.op;
      c.stream = w.stream;
      c.lsr = (guint)p.orders[l];
      c.lss = (guint)std::max(p.orders[l], p.orders[l - 1]);
      c.level = l;
      c.D = p.D;
      c.k = k;
      c.rc = 0;
      c.msg[0] = 0;
    }
    std::vector<std::thread> th;
    for (auto &c : jobs)
      th.emplace_back([&c]() {
        wbfmm_hip_set_stream(c.stream, nullptr);
        c.rc = wbfmm_guarded_call(coax_body, &c, c.msg, sizeof(c.msg));
      });
    for (auto &t : th) t.join();
    for (auto &c : jobs)
      if (c.rc) {
        snprintf(msg, ml, "%s", c.msg);
        return c.rc;
      }
    h.built = true;
  }
  *tables_ms = ms_between(t0, clk::now());
  return 0;
}

// frees least recently used idle operator sets until the idle ones fit
// hel_cap (on the calling worker's stream; an entry in use is skipped)
void hel_evict(wbfmm_serve_t *s) {
  std::vector<std::shared_ptr<HelEntry>> drop;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    while (s->hel_bytes > s->hel_cap) {
      auto victim = s->hel_cache.end();
      for (auto it = s->hel_cache.begin(); it != s->hel_cache.end(); ++it) {
        if (it->second.use_count() > 1) continue;  // held by a worker
        if (victim == s->hel_cache.end() || it->second->last < victim->second->last) victim = it;
      }
      if (victim == s->hel_cache.end()) break;
      s->hel_bytes -= victim->second->bytes;
      drop.push_back(victim->second);
      s->hel_cache.erase(victim);
    }
  }
  for (auto &e : drop)
    if (e->op) wbfmm_shift_operators_free(e->op);
}

// relative l2 errors of the checked targets against the FP64 direct sum
int check_errors(const wbfmm_serve_request_t &r, const double *dfield,
                 unsigned ntgt, int fw, double *ep, double *eg) {
  *ep = *eg = -1.0;
  const unsigned m = std::min<unsigned>(r.check, ntgt);
  if (m == 0) return 0;
  const unsigned step = ntgt / m;
  const void *xt = r.xt ? (const void *)r.xt : (const void *)r.xs;
  const size_t tstr = r.xt ? r.tstr : r.pstr;
  const bool helm = r.problem == WBFMM_PROBLEM_HELMHOLTZ;
  const int cplx = helm ? 2 : 1;
  double *xyz = nullptr, *nrm = nullptr, *tx = nullptr, *fd = nullptr, *fs = nullptr;
  double *q = nullptr, *dip = nullptr;
  uint32_t *iota = nullptr;
  int rc = 0;
  std::vector<uint32_t> hi(m);
  for (unsigned i = 0; i < m; i++) hi[i] = i;
  std::vector<double> hd((size_t)m * fw), hs((size_t)m * fw);
  if ((rc = wbfmm_hip_malloc((void **)&xyz, sizeof(double) * 4 * r.nsrc))) goto out;
  if ((rc = wbfmm_hip_stage4(r.xs, r.pstr, r.nsrc, xyz))) goto out;
  if (r.dip) {
    if ((rc = wbfmm_hip_malloc((void **)&nrm, sizeof(double) * 4 * r.nsrc))) goto out;
    if ((rc = wbfmm_hip_stage4(r.normals, r.nstr, r.nsrc, nrm))) goto out;
  }
  // strengths on the device (host arrays are copied)
  if (r.q) {
    if (wbfmm_hip_is_device(r.q)) {
      q = const_cast<double *>(r.q);
    } else {
      if ((rc = wbfmm_hip_malloc((void **)&q, sizeof(double) * cplx * r.nsrc))) goto out;
      if ((rc = wbfmm_hip_memcpy(q, r.q, sizeof(double) * cplx * r.nsrc))) goto out;
    }
  }
  if (r.dip) {
    if (wbfmm_hip_is_device(r.dip)) {
      dip = const_cast<double *>(r.dip);
    } else {
      if ((rc = wbfmm_hip_malloc((void **)&dip, sizeof(double) * cplx * r.nsrc))) goto out;
      if ((rc = wbfmm_hip_memcpy(dip, r.dip, sizeof(double) * cplx * r.nsrc))) goto out;
    }
  }
  if ((rc = wbfmm_hip_malloc((void **)&tx, sizeof(double) * 4 * m))) goto out;
  if ((rc = wbfmm_hip_stage4(xt, tstr * step, m, tx))) goto out;
  if ((rc = wbfmm_hip_malloc((void **)&fd, sizeof(double) * fw * m))) goto out;
  if ((rc = wbfmm_hip_memset(fd, 0, sizeof(double) * fw * m))) goto out;
  if (helm)
    rc = wbfmm_hip_hel_direct(xyz, q, nrm, dip, r.nsrc, tx, m, r.k, (int)r.field, 1e-6, fd, fw);
  else
    rc = wbfmm_hip_lap_direct(xyz, q, nrm, dip, r.nsrc, tx, m, (int)r.field, 1e-6, fd, fw);
  if (rc) goto out;
  // the FMM values of the same targets
  if ((rc = wbfmm_hip_malloc((void **)&iota, sizeof(uint32_t) * m))) goto out;
  if ((rc = wbfmm_hip_memcpy(iota, hi.data(), sizeof(uint32_t) * m))) goto out;
  if ((rc = wbfmm_hip_malloc((void **)&fs, sizeof(double) * fw * m))) goto out;
  if ((rc = wbfmm_hip_gather_strided(dfield, fw * (int)step, fw, iota, m, fs))) goto out;
  if ((rc = wbfmm_hip_memcpy(hd.data(), fd, sizeof(double) * fw * m))) goto out;
  if ((rc = wbfmm_hip_memcpy(hs.data(), fs, sizeof(double) * fw * m))) goto out;
  {
    // potential: the first cplx values when requested; gradient: the rest
    const int np = (r.field & WBFMM_FIELD_POTENTIAL) ? cplx : 0;
    double pn = 0, pd = 0, gn = 0, gd = 0;
    for (unsigned i = 0; i < m; i++)
      for (int c = 0; c < fw; c++) {
        const double d = hd[(size_t)i * fw + c], e = hs[(size_t)i * fw + c] - d;
        if (c < np) {
          pn += e * e;
          pd += d * d;
        } else {
          gn += e * e;
          gd += d * d;
        }
      }
    if (np) *ep = pd > 0 ? std::sqrt(pn / pd) : std::sqrt(pn);
    if (fw > np) *eg = gd > 0 ? std::sqrt(gn / gd) : std::sqrt(gn);
  }
out:
  wbfmm_hip_free(xyz);
  wbfmm_hip_free(nrm);
  wbfmm_hip_free(tx);
  wbfmm_hip_free(fd);
  wbfmm_hip_free(fs);
  wbfmm_hip_free(iota);
  if (q && q != r.q) wbfmm_hip_free(q);
  if (dip && dip != r.dip) wbfmm_hip_free(dip);
  return rc;
}

// state of one solve, plain data only: the guarded body below may be left
// by the host layer's recovery jump, so nothing in it has a destructor
struct Solve {
  const wbfmm_serve_request_t *r;
  const Plan *p;
  wbfmm_shift_operators_t *op;
  wbfmm_tree_t *t;
  wbfmm_target_list_t *tl;
  double *df;
  int fw;
  unsigned ntgt;
};

// the reference's call sequence for the request, on the worker's stream
int solve_body(void *arg) {
  Solve &S = *static_cast<Solve *>(arg);
  const wbfmm_serve_request_t &r = *S.r;
  const Plan &p = *S.p;
  const bool helm = r.problem == WBFMM_PROBLEM_HELMHOLTZ;
  const int cplx = helm ? 2 : 1;
  const void *xt = r.xt ? (const void *)r.xt : (const void *)r.xs;
  const size_t tstr = r.xt ? r.tstr : r.pstr;
  double x0[3] = {p.x0[0], p.x0[1], p.x0[2]};
  S.t = wbfmm_tree_new(x0, p.D, r.nsrc);
  wbfmm_tree_t *t = S.t;
  // the engine's trees are private: no host box arrays (8^l entries per
  // level), which FP32 coefficients do not allow anyway
  wbfmm_tree_host_mirror(t, FALSE);
  if (p.precision == 32) wbfmm_tree_set_precision(t, 32);
  wbfmm_tree_add_points(t, (gpointer)r.xs, r.pstr, (gpointer)r.normals,
                        r.normals ? r.nstr : 0, r.nsrc, FALSE);
  for (int l = 0; l < p.depth; l++) wbfmm_tree_refine(t);
  t->problem = helm ? WBFMM_PROBLEM_HELMHOLTZ : WBFMM_PROBLEM_LAPLACE;
  t->nq = 1;
  for (int l = 1; l <= p.depth; l++) {
    if (helm)
      wbfmm_tree_coefficient_init(t, (guint)l, (guint)p.orders[l], (guint)p.orders[l]);
    else
      wbfmm_tree_laplace_coefficient_init(t, (guint)l, (guint)p.orders[l], (guint)p.orders[l]);
  }
  S.tl = wbfmm_target_list_new(t, S.ntgt);
  wbfmm_target_list_add_points(S.tl, (gpointer)xt, tstr, S.ntgt);

  // the field is evaluated into a zeroed buffer of the engine and then
  // added to the caller's array, so the check can read the FMM values
  int rc = wbfmm_hip_malloc((void **)&S.df, sizeof(double) * S.fw * std::max(S.ntgt, 1u));
  if (rc) return rc;
  rc = wbfmm_hip_memset(S.df, 0, sizeof(double) * S.fw * std::max(S.ntgt, 1u));
  if (rc) return rc;

  for (int l = 1; l <= p.depth; l++) wbfmm_tree_coefficients_zero(t, (guint)l);
  if (helm) {
    wbfmm_tree_leaf_expansions(t, r.k, const_cast<double *>(r.q), cplx,
                               const_cast<double *>(r.dip), cplx, TRUE, nullptr);
    for (int l = p.depth; l >= 3; l--) wbfmm_upward_pass(t, S.op, (guint)l, nullptr);
    for (int l = 2; l <= p.depth; l++) wbfmm_downward_pass_ref(t, S.op, (guint)l, nullptr, 0);
    wbfmm_target_list_coefficients_init(S.tl, r.field);
    wbfmm_target_list_local_coefficients(S.tl, r.k, nullptr);
  } else {
    wbfmm_tree_laplace_leaf_expansions(t, const_cast<double *>(r.q), 1,
                                       const_cast<double *>(r.dip), 1, TRUE, nullptr);
    for (int l = p.depth; l >= 3; l--) wbfmm_laplace_upward_pass(t, S.op, (guint)l, nullptr);
    for (int l = 2; l <= p.depth; l++)
      wbfmm_laplace_downward_pass_ref(t, S.op, (guint)l, nullptr, 0);
    wbfmm_target_list_coefficients_init(S.tl, r.field);
    wbfmm_laplace_target_list_local_coefficients(S.tl, 1, nullptr);
  }
  wbfmm_target_list_local_field(S.tl, const_cast<double *>(r.q), cplx,
                                const_cast<double *>(r.dip), cplx, S.df, S.fw);
  return wbfmm_hip_sync();
}

// the solve's device objects, released also after a failed solve
int release_body(void *arg) {
  Solve &S = *static_cast<Solve *>(arg);
  if (S.df) wbfmm_hip_free(S.df);
  if (S.tl) wbfmm_target_list_free(S.tl);
  if (S.t) wbfmm_tree_free(S.t);
  S.df = nullptr;
  S.tl = nullptr;
  S.t = nullptr;
  return wbfmm_hip_sync();
}

// one request on the worker's stream; a failure sets res.status and leaves
// the engine running
void run(wbfmm_serve_t *s, Worker &w, Job &j) {
  const wbfmm_serve_request_t &r = j.req;
  const Plan &p = j.plan;
  const bool helm = r.problem == WBFMM_PROBLEM_HELMHOLTZ;
  double tables_ms = 0.0;
  char msg[512] = "";
  Solve S{&r, &p, nullptr, nullptr, nullptr, nullptr, field_width(r.problem, r.field),
          r.xt ? r.ntgt : r.nsrc};
  // (Laplace tables are built on first use inside the passes, so their
  // time is part of the first solve of each order on a worker)
  std::shared_ptr<HelEntry> hent;
  std::unique_lock<std::mutex> hheld;
  int rc = helm ? helmholtz_ops(s, w, p, r.k, &tables_ms, hent, hheld, msg, sizeof(msg))
                : laplace_ops(w, p.lmax, &S.op, msg, sizeof(msg));
  if (helm && !rc) S.op = hent->op;
  for (int v = 0; v < WBFMM_VARIANT_COUNT; v++) wbfmm_hip_set_variant(v, p.var[v]);
  if (!rc) rc = wbfmm_guarded_call(solve_body, &S, msg, sizeof(msg));
  for (int v = 0; v < WBFMM_VARIANT_COUNT; v++) wbfmm_hip_set_variant(v, -1);
  // the operator set is complete after a solve that ended well; after a
  // failure it is built again by its next user
  if (helm && hent) {
    if (rc) hent->built = false;
    if (hheld.owns_lock()) hheld.unlock();
    hent.reset();
    hel_evict(s);
  }
  // add into the caller's output
  if (!rc) {
    if (wbfmm_hip_is_device(r.f)) {
      rc = wbfmm_hip_accumulate(r.f, r.fstr, S.df, S.fw, S.fw, S.ntgt);
      if (!rc) rc = wbfmm_hip_sync();
    } else {
      std::vector<double> h((size_t)S.fw * S.ntgt);
      rc = wbfmm_hip_memcpy(h.data(), S.df, sizeof(double) * h.size());
      if (!rc)
        for (unsigned i = 0; i < S.ntgt; i++)
          for (int c = 0; c < S.fw; c++) r.f[(size_t)i * r.fstr + c] += h[(size_t)i * S.fw + c];
    }
  }
  j.t_done = clk::now();
  j.res.tables_ms = tables_ms;
  if (!rc && r.check) {
    const auto c0 = clk::now();
    rc = check_errors(r, S.df, S.ntgt, S.fw, &j.res.err_potential, &j.res.err_gradient);
    j.res.check_ms = ms_between(c0, clk::now());
  } else {
    j.res.err_potential = j.res.err_gradient = -1.0;
  }
  // allocation failures forced for a test end with the solve
  wbfmm_hip_debug_fail_allocs(0);
  char msg2[256];
  const int rrc = wbfmm_guarded_call(release_body, &S, msg2, sizeof(msg2));
  if (rc && msg[0]) fprintf(stderr, "wbfmm serve: request failed (%d): %s\n", rc, msg);
  j.res.status = rc ? rc : rrc;
}

void worker_main(wbfmm_serve_t *s, Worker *w) {
  wbfmm_hip_set_stream(w->stream, nullptr);
  for (;;) {
    Job *j = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        if (s->stop && s->queue.empty()) return true;
        if (s->queue.empty()) return false;
        const size_t need = s->queue.front()->plan.footprint;
        return s->inflight == 0 || s->inflight + need <= s->budget;
      });
      if (s->queue.empty()) break;  // stopping
      j = s->queue.front();
      s->queue.pop_front();
      s->inflight += j->plan.footprint;
      s->peak = std::max(s->peak, s->inflight);
      j->t_start = clk::now();
    }
    // another waiting worker may fit beside this one
    s->cv_work.notify_all();
    wbfmm_hip_debug_fail_allocs(j->fail_allocs);
    run(s, *w, *j);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->inflight -= j->plan.footprint;
      j->res.queue_ms = ms_between(j->t_submit, j->t_start);
      j->res.solve_ms = ms_between(j->t_start, j->t_done);
      j->res.latency_ms = ms_between(j->t_submit, j->t_done);
      if (!j->tune && j->res.status == 0) {
        s->lat.push_back(j->res.latency_ms);
        s->targets += j->req.xt ? j->req.ntgt : j->req.nsrc;
        s->last_done = std::max(s->last_done, j->t_done);
      }
      j->done = true;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
  // release the worker's operator sets on its own stream
  for (auto &kv : w->lap_ops) wbfmm_shift_operators_free(kv.second);
  wbfmm_hip_sync();
  wbfmm_hip_set_stream(nullptr, nullptr);
}

// plan on the caller's thread (the bounds pass runs on the library stream)
// var (an autotuner run): the kernel variants to time, in place of the
// table's
int64_t enqueue(wbfmm_serve_t *s, const wbfmm_serve_request_t *r, bool tune,
                const int *var = nullptr) {
  auto j = std::make_unique<Job>();
  j->req = *r;
  j->tune = tune;
  j->t_submit = clk::now();
  int rc = make_plan(s, *r, j->plan);
  if (rc) return rc;
  if (var)
    for (int v = 0; v < WBFMM_VARIANT_COUNT; v++) j->plan.var[v] = var[v];
  j->res.depth = j->plan.depth;
  j->res.order = j->plan.lmax;
  j->res.precision = j->plan.precision;
  j->res.footprint = j->plan.footprint;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return WBFMM_HIP_ERR_ARG;
  // a request that cannot fit the budget even alone is refused
  if (j->plan.footprint > s->budget) return WBFMM_HIP_ERR_NOMEM;
  j->fail_allocs = s->fail_allocs;
  s->fail_allocs = 0;
  const int64_t id = s->next_id++;
  j->id = id;
  if (!tune && !s->have_first) {
    s->have_first = true;
    s->first_submit = j->t_submit;
    s->last_done = j->t_submit;
  }
  s->queue.push_back(j.get());
  s->jobs[id] = std::move(j);
  s->cv_work.notify_one();
  return id;
}

}  // namespace

extern "C" {

wbfmm_serve_t *wbfmm_serve_new(gint nstreams, gsize budget) {
  if (nstreams < 1 || nstreams > 64) return nullptr;
  if (wbfmm_hip_init()) return nullptr;
  auto *s = new wbfmm_serve_s();
  if (budget == 0) {
    // 90% of the device memory free now
    size_t fr = 0, tot = 0;
    if (wbfmm_hip_mem_info(&fr, &tot) == 0) budget = fr / 10 * 9;
    if (budget == 0) budget = (size_t)64 << 30;
  }
  s->budget = budget;
  // idle Helmholtz operator sets kept resident: up to 1/16 of the budget
  s->hel_cap = budget / 16;
  // process-wide tables of the drop-in API, set once before any worker
  wbfmm_shift_angle_table_init();
  wbfmm_laplace_coaxial_translate_init(WBFMM_HIP_MAX_ORDER + 2);
  for (int i = 0; i < nstreams; i++) {
    auto w = std::make_unique<Worker>();
    if (wbfmm_hip_stream_create(&w->stream)) {
      s->stop = true;
      for (auto &x : s->workers) wbfmm_hip_stream_destroy(x->stream);
      delete s;
      return nullptr;
    }
    s->workers.push_back(std::move(w));
  }
  for (auto &w : s->workers) w->th = std::thread(worker_main, s, w.get());
  return s;
}

gint64 wbfmm_serve_submit(wbfmm_serve_t *s, const wbfmm_serve_request_t *r) {
  if (!s || !r) return WBFMM_HIP_ERR_ARG;
  return enqueue(s, r, false);
}

gint wbfmm_serve_wait(wbfmm_serve_t *s, gint64 ticket, wbfmm_serve_result_t *res) {
  if (!s) return WBFMM_HIP_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end()) return WBFMM_HIP_ERR_ARG;
  Job *j = it->second.get();
  s->cv_done.wait(lk, [&] { return j->done; });
  if (res) *res = j->res;
  const int st = j->res.status;
  s->jobs.erase(it);
  return st;
}

gint wbfmm_serve_stats(wbfmm_serve_t *s, wbfmm_serve_stats_t *st) {
  if (!s || !st) return WBFMM_HIP_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st, 0, sizeof(*st));
  st->completed = s->lat.size();
  st->targets = s->targets;
  st->peak_bytes = s->peak;
  if (!s->lat.empty()) {
    std::vector<double> v = s->lat;
    std::sort(v.begin(), v.end());
    // nearest-rank percentiles
    auto pct = [&](double q) {
      size_t k = (size_t)std::ceil(q * v.size());
      return v[std::min(std::max(k, (size_t)1), v.size()) - 1];
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    double sum = 0;
    for (double x : v) sum += x;
    st->mean_ms = sum / v.size();
    st->max_ms = v.back();
    st->span_s = ms_between(s->first_submit, s->last_done) / 1e3;
  }
  return 0;
}

gint wbfmm_serve_reset_stats(wbfmm_serve_t *s) {
  if (!s) return WBFMM_HIP_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->lat.clear();
  s->targets = 0;
  s->have_first = false;
  s->peak = s->inflight;
  return 0;
}

gint wbfmm_serve_plan(wbfmm_serve_t *s, const wbfmm_serve_request_t *r, gint *depth,
                      gint *order, gsize *fp) {
  if (!s || !r) return WBFMM_HIP_ERR_ARG;
  Plan p;
  int rc = make_plan(s, *r, p);
  if (rc) return rc;
  if (depth) *depth = p.depth;
  if (order) *order = p.lmax;
  if (fp) *fp = p.footprint;
  return 0;
}

gint wbfmm_serve_tune(wbfmm_serve_t *s, const wbfmm_serve_request_t *r, const gint *depths,
                      gint ndepth, gint *best) {
  if (!s || !r || !depths || ndepth < 1) return WBFMM_HIP_ERR_ARG;
  int rc = validate(*r);
  if (rc) return rc;
  const int fw = field_width(r->problem, r->field);
  const unsigned ntgt = r->xt ? r->ntgt : r->nsrc;
  // as many copies as the engine has workers, each with its own scratch
  // output: a class is timed the way a batch runs it, its solves sharing
  // the GPU (alone on an idle GPU, fewer and larger launches would win)
  const int ncopy = (int)s->workers.size();
  std::vector<double *> scratch(ncopy, nullptr);
  for (int c = 0; c < ncopy && !rc; c++)
    rc = wbfmm_hip_malloc((void **)&scratch[c], sizeof(double) * fw * std::max(ntgt, 1u));
  auto release = [&]() {
    wbfmm_hip_sync();
    for (double *p : scratch)
      if (p) wbfmm_hip_free(p);
  };
  if (rc) {
    release();
    return rc;
  }
  // wall time per solve of q at depth d with the kernel variants var, the
  // copies submitted at once (two rounds: the first builds the workers'
  // tables)
  auto timed = [&](int d, const int *var, double &ms) {
    ms = INFINITY;
    for (int rep = 0; rep < 2; rep++) {
      std::vector<int64_t> ids;
      const auto t0 = clk::now();
      for (int c = 0; c < ncopy; c++) {
        wbfmm_serve_request_t q = *r;
        q.depth = d;
        q.f = scratch[c];
        q.fstr = fw;
        q.check = 0;
        const int64_t id = enqueue(s, &q, true, var);
        if (id < 0) return (int)id;
        ids.push_back(id);
      }
      int err = 0;
      for (int64_t id : ids) {
        wbfmm_serve_result_t res;
        const int e = wbfmm_serve_wait(s, id, &res);
        if (!err) err = e ? e : res.status;
      }
      if (err) return err;
      ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count() / ncopy;
    }
    return 0;
  };
  const int dflt[WBFMM_VARIANT_COUNT] = {-1, -1};
  Tuned t;
  double best_ms = INFINITY;
  // the depth with the library's kernels, then each kernel variant at that
  // depth, one kind at a time: the sparse-level M2L tiles (chunks, offset
  // tiles, class tiles; Helmholtz chunks or class tiles), the Helmholtz
  // field (thread per target, wavefront per target box)
  for (int i = 0; i < ndepth && !rc; i++) {
    double ms;
    rc = timed(depths[i], dflt, ms);
    if (!rc && ms < best_ms) {
      best_ms = ms;
      t.depth = depths[i];
    }
  }
  const bool helm = r->problem == WBFMM_PROBLEM_HELMHOLTZ;
  for (int kind = 0; kind < WBFMM_VARIANT_COUNT && !rc; kind++) {
    if (kind == WBFMM_VARIANT_FIELD && !helm) continue;
    const int cand_m2l[] = {0, 1, 2}, cand_hm2l[] = {0, 2}, cand_field[] = {0, 1};
    const int *cand = kind == WBFMM_VARIANT_M2L ? (helm ? cand_hm2l : cand_m2l) : cand_field;
    const int nc = kind == WBFMM_VARIANT_M2L ? (helm ? 2 : 3) : 2;
    int bv = -1;
    double bms = best_ms;
    for (int c = 0; c < nc && !rc; c++) {
      int var[WBFMM_VARIANT_COUNT];
      for (int v = 0; v < WBFMM_VARIANT_COUNT; v++) var[v] = t.var[v];
      var[kind] = cand[c];
      double ms;
      rc = timed(t.depth, var, ms);
      if (!rc && ms < bms) {
        bms = ms;
        bv = cand[c];
      }
    }
    if (bv >= 0) {
      t.var[kind] = bv;
      best_ms = bms;
    }
  }
  release();
  if (rc) return rc;
  {
    Plan p;
    rc = make_plan(s, *r, p);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    const int key_order = r->problem == WBFMM_PROBLEM_LAPLACE
                              ? (r->order ? r->order : laplace_order(r->tol))
                              : r->order;
    s->tuned[tune_key(*r, p.precision, key_order)] = t;
  }
  if (best) *best = t.depth;
  return 0;
}

gint wbfmm_serve_tuned(wbfmm_serve_t *s, const wbfmm_serve_request_t *r, gint *depth,
                       gint *var_m2l, gint *var_field) {
  if (!s || !r) return WBFMM_HIP_ERR_ARG;
  Plan p;
  const int rc = make_plan(s, *r, p);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(s->mu);
  const int key_order = r->problem == WBFMM_PROBLEM_LAPLACE
                            ? (r->order ? r->order : laplace_order(r->tol))
                            : r->order;
  auto it = s->tuned.find(tune_key(*r, p.precision, key_order));
  const Tuned t = it != s->tuned.end() ? it->second : Tuned{};
  if (depth) *depth = t.depth;
  if (var_m2l) *var_m2l = t.var[WBFMM_VARIANT_M2L];
  if (var_field) *var_field = t.var[WBFMM_VARIANT_FIELD];
  return 0;
}

gint wbfmm_serve_debug_fail_allocations(wbfmm_serve_t *s, gint count) {
  if (!s || count < 0) return WBFMM_HIP_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->fail_allocs = count;
  return 0;
}

gint wbfmm_serve_free(wbfmm_serve_t *s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &w : s->workers) w->th.join();
  // the resident Helmholtz operator sets (idle now) on the library stream
  for (auto &kv : s->hel_cache)
    if (kv.second->op) wbfmm_shift_operators_free(kv.second->op);
  s->hel_cache.clear();
  wbfmm_hip_sync();
  for (auto &w : s->workers) wbfmm_hip_stream_destroy(w->stream);
  delete s;
  return 0;
}

}  // extern "C"
