// Request loop of the solver core (no counterpart in the reference, which
// solves one case per process).
//
// A request is a mesh, its frequencies and the normal velocities of every
// problem at each frequency; the reply is the forces of every problem.  The
// unit of work is one (request, frequency).  Work is handed out while the
// device-memory budget allows: a worker continues the request its last item
// came from, else takes the oldest request no other worker is on, else the
// oldest (so distinct requests spread over the workers, each handle keeps
// its mesh across a request's frequencies, and a lone request is shared by
// every worker).  The budget bounds
// the device memory that the workers' handles hold: a busy worker is
// charged the larger of what its handle holds and what its frequency needs
// (FP64 D and S, the complex64 factor copy in mixed precision, the pivot
// blocks, right-hand sides and geometry: about 32 N^2 bytes per symmetry
// block in FP64, 40 N^2 in mixed), an idle worker what its handle still
// holds.  When only idle workers' buffers stand in the way, they are freed;
// with nothing running the head of the queue is always admitted, so an
// oversized request still runs.  Each worker thread owns one solver handle,
// and so one HIP stream and one set of resident matrices and factors; a
// handle is re-bound to the mesh of whichever request its worker takes
// next, reusing its device buffers.  So the frequencies of one request, and
// of requests that arrive while others run, overlap on the GPU as soon as a
// worker and budget are free.
//
// With a batch size above 1 (nemoh_server_set_batch; 1 by default) a worker
// may take several frequencies of a wave request at once, as many as the
// budget allows at the request's footprint each: it then binds as many
// handles of its own to the mesh and solves them as one batch
// (nemoh_batch_solve_waves_multi), every heavy step one launch over all of
// them.  That pays for many frequencies of one small mesh; when each request
// brings a new mesh, re-binding and re-tabling the handles costs more.  A
// batch gives each frequency the arithmetic of a lone handle, so results do
// not change either.
//
// A frequency's result does not depend on the schedule: every kernel is
// deterministic, handles share nothing but the read-only wave tables, and
// the module NQ that the reference carries from one frequency to the next
// (an infinite-depth frequency reuses the NQ of the last finite-depth one)
// is fixed per request before any of its frequencies runs, as a sequential
// session over the request's frequency list would see it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "core_internal.h"
#include "nemoh_core.h"

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  long long id = 0;
  // inputs (copied at submission)
  int np = 0, nfa = 0, nsym = 0, nfreq = 0, nrhs = 0, nint = 0, precision = NEMOH_MIXED;
  std::vector<double> x, y, z, periods, nvel, nds;
  std::vector<int> m1, m2, m3, m4;
  double depth = 0, rho = 1025, g = 9.81;
  // wave requests: problems generated on the device (nemoh_core_solve_waves)
  bool waves = false;
  int nbeta = 0;
  std::vector<double> omegas, betas;
  std::vector<int> body;  // body of each panel (empty: all body 1)
  std::vector<nemoh_dof> rad, force;
  double xeff = 0, yeff = 0;
  std::vector<double> fk;  // (nint, nbeta) per frequency
  double footprint = 0;  // device bytes a handle needs for one of its frequencies
  // set-up shared by the frequencies, made once by the first worker
  std::mutex pm;
  std::condition_variable pcv;
  int prep = 0;  // 0 not started, 1 running, 2 done, -1 failed
  std::vector<double> geo;  // XG YG ZG XN YN ZN AIRE TDIS, nfa each
  double zer = 0;
  std::vector<int> nq_before;
  // progress (guarded by the server mutex)
  int next = 0, finished = 0;
  bool done = false;
  double t_submit = 0, t_start = -1, t_done = 0;
  // results; each frequency writes its own slice
  std::vector<double> forces;
  std::mutex em;  // guards status/err
  int status = NEMOH_OK;
  std::string err;

  void set_error(int code, const std::string& msg) {
    std::lock_guard<std::mutex> lk(em);
    if (status == NEMOH_OK) {
      status = code;
      err = msg;
    }
  }
  bool failed() {
    std::lock_guard<std::mutex> lk(em);
    return status != NEMOH_OK;
  }
};

// One solver handle of a worker.
struct Slot {
  nemoh_core* h = nullptr;
  long long bound = -1;     // request the handle's mesh belongs to
  long long dofs_for = -1;  // request whose degrees of freedom the handle holds
};

struct Worker {
  Slot main;               // the handle of single frequencies (and the first of a batch)
  std::vector<Slot> more;  // the batch's further handles
  nemoh_batch* batch = nullptr;  // over the handles it was made for
  std::vector<nemoh_core*> batch_handles;
  int batch_prec = -1;
  // guarded by the server mutex:
  bool busy = false;
  double held = 0;     // device bytes the handles hold (updated after each item)
  double charged = 0;  // bytes counted against the budget (held, or more while busy)
  bool release = false;  // free the handles' matrices before the next item
  long long claim = -1;  // request of this worker's current or last item
  size_t pick = 0;       // queue position of the request admitted for its next item
  int take = 1;          // frequencies admitted for the next item
};

double held_bytes(const Worker& w) {
  double b = w.main.h ? (double)nemoh::core_held_bytes(w.main.h) : 0.0;
  for (const Slot& sl : w.more) b += sl.h ? (double)nemoh::core_held_bytes(sl.h) : 0.0;
  return b;
}

void release_all(Worker& w) {
  if (w.main.h) (void)nemoh::core_release_matrices(w.main.h);
  for (Slot& sl : w.more)
    if (sl.h) (void)nemoh::core_release_matrices(sl.h);
}

}  // namespace

struct nemoh_server {
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;  // requests with frequencies not yet handed out
  std::unordered_map<long long, std::shared_ptr<Request>> table;
  std::vector<Worker> workers;
  double budget = 0;
  int batch = 1;    // most frequencies of a wave request one worker solves at once
  double peak = 0;  // largest total charge admitted with other work running
  int running = 0;
  bool stop = false;
  long long next_id = 0;
  std::vector<double> xr, xz, apd1x, apd1z, apd2x, apd2z;
  std::string err;
};

namespace {

void geo_ptrs(const Request& r, const double* out[8]) {
  for (int k = 0; k < 8; ++k) out[k] = r.geo.data() + (size_t)k * r.nfa;
}

// Points a handle at r's mesh (creating the handle on first use).
int bind(nemoh_server* s, Slot& w, Request& r) {
  if (w.h && w.bound == r.id) return NEMOH_OK;
  const double* geo[8];
  geo_ptrs(r, geo);
  int e;
  if (!w.h) {
    e = nemoh_core_create(&w.h, r.np, r.nfa, r.nsym, r.x.data(), r.y.data(), r.z.data(), r.m1.data(), r.m2.data(),
                          r.m3.data(), r.m4.data(), geo[0], geo[1], geo[2], geo[3], geo[4], geo[5], geo[6], geo[7],
                          r.zer, r.g, r.rho, r.depth, s->xr.data(), s->xz.data(), s->apd1x.data(), s->apd1z.data(),
                          s->apd2x.data(), s->apd2z.data());
    if (e) {
      w.h = nullptr;
      r.set_error(e, std::string("could not create a solver handle: ") + nemoh_core_last_error(nullptr));
      return e;
    }
  } else {
    e = nemoh::core_bind_mesh(w.h, r.np, r.nfa, r.nsym, r.x.data(), r.y.data(), r.z.data(), r.m1.data(),
                              r.m2.data(), r.m3.data(), r.m4.data(), geo, r.zer, r.g, r.rho, r.depth);
    if (e) {
      w.bound = -1;
      r.set_error(e, nemoh_core_last_error(w.h));
      return e;
    }
  }
  w.bound = r.id;
  w.dofs_for = -1;  // binding clears the handle's body-condition set-up
  return NEMOH_OK;
}

// Panel geometry and the per-frequency NQ, once per request.
bool prepare(nemoh_server* s, Worker& w, Request& r) {
  std::unique_lock<std::mutex> lk(r.pm);
  if (r.prep == 0) {
    r.prep = 1;
    lk.unlock();
    const int n = r.nfa;
    r.geo.assign(8 * (size_t)n, 0.0);
    double* g = r.geo.data();
    int e = nemoh_core_prepare_mesh(r.np, n, r.nsym, r.x.data(), r.y.data(), r.z.data(), r.m1.data(), r.m2.data(),
                                    r.m3.data(), r.m4.data(), g, g + n, g + 2 * n, g + 3 * n, g + 4 * n, g + 5 * n,
                                    g + 6 * n, g + 7 * n, &r.zer);
    if (e) r.set_error(e, "degenerate panel in the request mesh");
    if (!e) e = bind(s, w.main, r);
    if (!e) {
      r.nq_before.assign(r.nfreq, 0);
      e = nemoh::core_nq_sequence(w.main.h, r.nfreq, r.periods.data(), r.nq_before.data());
      if (e) r.set_error(e, nemoh_core_last_error(w.main.h));
    }
    lk.lock();
    r.prep = e ? -1 : 2;
    r.pcv.notify_all();
  } else {
    r.pcv.wait(lk, [&] { return r.prep != 1; });
  }
  return r.prep == 2;
}

void run_item(nemoh_server* s, Worker& ww, Request& r, int f) {
  if (!prepare(s, ww, r) || r.failed()) return;
  Slot& w = ww.main;
  if (bind(s, w, r)) return;
  const size_t ntot = (size_t)r.nfa << r.nsym;
  const size_t per_f = 2 * ntot * r.nrhs, per_force = 2 * (size_t)r.nint * r.nrhs;
  int e = nemoh_core_set_precision(w.h, r.precision, 0.0, 0);
  nemoh::core_set_nq_state(w.h, r.nq_before[f]);
  if (r.waves) {
    if (!e && w.dofs_for != r.id) {
      e = nemoh_core_set_dofs(w.h, r.body.empty() ? nullptr : r.body.data(), (int)r.rad.size(), r.rad.data(),
                              (int)r.force.size(), r.force.data(), r.xeff, r.yeff);
      if (!e) w.dofs_for = r.id;
    }
    if (!e)
      e = nemoh_core_solve_waves(w.h, r.omegas[f], r.nbeta, r.betas.data(), r.forces.data() + f * per_force,
                                 r.nbeta > 0 ? r.fk.data() + f * 2 * (size_t)r.nint * r.nbeta : nullptr);
  } else {
    nemoh_freq fs;
    if (!e) e = nemoh_core_freq_setup(w.h, r.periods[f], &fs);
    if (!e)
      e = nemoh_core_solve_multi(w.h, r.periods[f], fs.amh, fs.nexp, fs.ambda, fs.ar, fs.nq, fs.qq, fs.cq, r.nrhs,
                                 r.nvel.data() + f * per_f, nullptr, r.nint, r.nds.data(),
                                 r.forces.data() + f * per_force);
  }
  if (e) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "frequency %d: ", f);
    r.set_error(e, buf + std::string(nemoh_core_last_error(w.h)));
  }
}

// Frequencies f0 .. f0 + k - 1 of wave request r as one batch of the
// worker's handles, one per frequency, each with the module NQ that
// frequency has in the request's sequence.
void run_batch(nemoh_server* s, Worker& w, Request& r, int f0, int k) {
  if (!prepare(s, w, r) || r.failed()) return;
  while ((int)w.more.size() < k - 1) w.more.emplace_back();
  std::vector<nemoh_core*> hs;
  int e = NEMOH_OK;
  std::string msg;
  for (int i = 0; i < k && !e; ++i) {
    Slot& sl = i == 0 ? w.main : w.more[i - 1];
    if ((e = bind(s, sl, r))) return;  // bind reported it
    e = nemoh_core_set_precision(sl.h, r.precision, 0.0, 0);
    nemoh::core_set_nq_state(sl.h, r.nq_before[f0 + i]);
    if (!e && sl.dofs_for != r.id) {
      e = nemoh_core_set_dofs(sl.h, r.body.empty() ? nullptr : r.body.data(), (int)r.rad.size(), r.rad.data(),
                              (int)r.force.size(), r.force.data(), r.xeff, r.yeff);
      if (!e) sl.dofs_for = r.id;
    }
    if (e) msg = nemoh_core_last_error(sl.h);
    hs.push_back(sl.h);
  }
  if (!e && (!w.batch || w.batch_handles != hs || w.batch_prec != r.precision)) {
    if (w.batch) nemoh_batch_destroy(w.batch);
    w.batch = nullptr;
    w.batch_handles.clear();
    e = nemoh_batch_create(&w.batch, k, hs.data(), r.precision, 16);
    if (e) {
      msg = nemoh_batch_last_error(nullptr);
      w.batch = nullptr;
    } else {
      w.batch_handles = hs;
      w.batch_prec = r.precision;
    }
  }
  if (!e) {
    const size_t per_force = 2 * (size_t)r.nint * r.nrhs, per_fk = 2 * (size_t)r.nint * r.nbeta;
    e = nemoh_batch_solve_waves_multi(w.batch, r.omegas.data() + f0, r.nbeta, r.nbeta > 0 ? r.betas.data() : nullptr,
                                      r.forces.data() + f0 * per_force, r.nbeta > 0 ? r.fk.data() + f0 * per_fk : nullptr);
    if (e) msg = nemoh_batch_last_error(w.batch);
  }
  if (e) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "frequencies %d-%d: ", f0, f0 + k - 1);
    r.set_error(e, buf + msg);
  }
}

// The queued request a worker takes its next frequency from (server mutex
// held): the one its last item came from while it has frequencies left, so
// the handle keeps its mesh, factors-to-be and graphs; else the first that no
// other worker's last item came from, so that distinct requests (a batch of
// bodies) spread over the workers; else the oldest, so that a lone request's
// frequencies are shared by every worker.
size_t pick_request(const nemoh_server* s, int self) {
  const long long mine = s->workers[self].claim;
  for (size_t k = 0; k < s->queue.size(); ++k)
    if (s->queue[k]->id == mine) return k;
  for (size_t k = 0; k < s->queue.size(); ++k) {
    bool taken = false;
    for (size_t v = 0; v < s->workers.size() && !taken; ++v)
      taken = (int)v != self && s->workers[v].claim == s->queue[k]->id;
    if (!taken) return k;
  }
  return 0;
}

// Under the server mutex: may worker `self` take its next frequency (from
// the request pick_request chooses, recorded in w.pick)?  Frees idle
// workers' matrices when that is what stands in the way.  On success the
// worker is charged and marked busy.
bool try_admit(nemoh_server* s, int self, bool freed = false) {
  Worker& w = s->workers[self];
  w.pick = pick_request(s, self);
  const Request& q = *s->queue[w.pick];
  double others = 0, idle_held = 0;
  for (size_t v = 0; v < s->workers.size(); ++v) {
    if ((int)v == self) continue;
    others += s->workers[v].charged;
    if (!s->workers[v].busy) idle_held += s->workers[v].charged;
  }
  int running = 0;
  for (const Worker& o : s->workers) running += o.busy;
  auto admit = [&](int k, double need, bool shrink) {
    w.take = k;
    w.release = shrink && w.held > need;
    w.charged = w.release ? need : std::max(w.held, need);
    w.busy = true;
    if (running > 0) s->peak = std::max(s->peak, others + w.charged);
    return true;
  };
  // the most frequencies of the request that fit, halving from the batch size
  const int want = q.waves ? std::max(1, std::min(s->batch, q.nfreq - q.next)) : 1;
  for (int k = want;; k = std::max(1, k / 2)) {
    const double need = k * q.footprint;
    if (others + std::max(w.held, need) <= s->budget) return admit(k, need, false);
    if (others + need <= s->budget) return admit(k, need, true);  // after freeing this worker's own buffers
    if (k == 1) break;
  }
  const double need = q.footprint;
  if (!freed && idle_held > 0 && others - idle_held + need <= s->budget) {
    // the idle workers' buffers are what stands in the way: free them (an
    // idle worker never touches its handles without holding this mutex)
    for (size_t v = 0; v < s->workers.size(); ++v) {
      Worker& o = s->workers[v];
      if ((int)v == self || o.busy || o.held == 0) continue;
      release_all(o);
      o.held = o.charged = held_bytes(o);
    }
    return try_admit(s, self, true);
  }
  if (running == 0) return admit(1, need, true);  // alone on the device: always admitted
  return false;
}

void worker_loop(nemoh_server* s, int self) {
  Worker& w = s->workers[self];
  for (;;) {
    std::shared_ptr<Request> r;
    int f = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // queued work is finished before a stop takes effect
      s->cv_work.wait(lk, [&] { return (!s->queue.empty() && try_admit(s, self)) || (s->stop && s->queue.empty()); });
      if (!w.busy) break;  // stopped with the queue empty
      r = s->queue[w.pick];
      f = r->next;
      r->next += w.take;
      w.claim = r->id;
      if (r->next == r->nfreq) s->queue.erase(s->queue.begin() + (std::ptrdiff_t)w.pick);
      if (r->t_start < 0) r->t_start = now_ms();
    }
    if (w.release) release_all(w);
    if (w.take == 1) run_item(s, w, *r, f);
    else run_batch(s, w, *r, f, w.take);
    bool last;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      w.busy = false;
      w.release = false;
      w.held = w.charged = held_bytes(w);
      r->finished += w.take;
      last = r->finished == r->nfreq;
      if (last) {
        r->t_done = now_ms();
        r->done = true;
      }
    }
    if (last) s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

}  // namespace

namespace {

bool mesh_ok(int np, int nfa, int nsym, const double* x, const double* y, const double* z, const int* m1,
             const int* m2, const int* m3, const int* m4, double depth, int precision) {
  return np > 0 && nfa > 0 && (nsym == 0 || nsym == 1) && x && y && z && m1 && m2 && m3 && m4 && depth >= 0 &&
         (precision == NEMOH_FP64 || precision == NEMOH_MIXED);
}

std::shared_ptr<Request> new_request(int np, int nfa, int nsym, const double* x, const double* y, const double* z,
                                     const int* m1, const int* m2, const int* m3, const int* m4, double depth,
                                     double rho, double g, int precision, int nfreq) {
  auto r = std::make_shared<Request>();
  r->np = np;
  r->nfa = nfa;
  r->nsym = nsym;
  r->x.assign(x, x + np);
  r->y.assign(y, y + np);
  r->z.assign(z, z + np);
  r->m1.assign(m1, m1 + nfa);
  r->m2.assign(m2, m2 + nfa);
  r->m3.assign(m3, m3 + nfa);
  r->m4.assign(m4, m4 + nfa);
  r->depth = depth;
  r->rho = rho;
  r->g = g;
  r->precision = precision;
  r->nfreq = nfreq;
  return r;
}

int enqueue(nemoh_server* s, const std::shared_ptr<Request>& r, long long* id) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return NEMOH_ERR_ARG;
    r->id = s->next_id++;
    r->t_submit = now_ms();
    s->table[r->id] = r;
    s->queue.push_back(r);
    *id = r->id;
  }
  s->cv_work.notify_all();
  return NEMOH_OK;
}

}  // namespace

extern "C" {

int nemoh_server_create(nemoh_server** out, int workers, double budget_bytes, const double* xr, const double* xz,
                        const double* apd1x, const double* apd1z, const double* apd2x, const double* apd2z) {
  if (!out) return NEMOH_ERR_ARG;
  *out = nullptr;
  if (workers < 1 || workers > 64 || !(budget_bytes > 0) || !xr || !xz || !apd1x || !apd1z || !apd2x || !apd2z)
    return NEMOH_ERR_ARG;
  if (nemoh::core_device_count() < 1) return NEMOH_ERR_HIP;
  nemoh_server* s = new nemoh_server();
  s->budget = budget_bytes;
  const size_t nr = 328, nz = 46, nt = nr * nz;
  s->xr.assign(xr, xr + nr);
  s->xz.assign(xz, xz + nz);
  s->apd1x.assign(apd1x, apd1x + nt);
  s->apd1z.assign(apd1z, apd1z + nt);
  s->apd2x.assign(apd2x, apd2x + nt);
  s->apd2z.assign(apd2z, apd2z + nt);
  s->workers.resize(workers);
  for (int k = 0; k < workers; ++k) s->threads.emplace_back(worker_loop, s, k);
  *out = s;
  return NEMOH_OK;
}

int nemoh_server_submit(nemoh_server* s, int np, int nfa, int nsym, const double* x, const double* y, const double* z,
                        const int* m1, const int* m2, const int* m3, const int* m4, double depth, double rho,
                        double g, int precision, int nfreq, const double* periods, int nrhs, const double* nvel,
                        int nint, const double* nds, long long* id) {
  if (!s || !id || !mesh_ok(np, nfa, nsym, x, y, z, m1, m2, m3, m4, depth, precision) || nfreq <= 0 || !periods ||
      nrhs <= 0 || nrhs > 16 || !nvel || nint <= 0 || !nds)
    return NEMOH_ERR_ARG;
  auto r = new_request(np, nfa, nsym, x, y, z, m1, m2, m3, m4, depth, rho, g, precision, nfreq);
  const size_t ntot = (size_t)nfa << nsym;
  r->periods.assign(periods, periods + nfreq);
  r->nrhs = nrhs;
  r->nint = nint;
  r->nvel.assign(nvel, nvel + 2 * ntot * nrhs * nfreq);
  r->nds.assign(nds, nds + (size_t)nint * ntot);
  r->forces.assign(2 * (size_t)nint * nrhs * nfreq, 0.0);
  r->footprint = (double)nemoh::core_footprint_bytes(np, nfa, nsym, precision, nrhs);
  return enqueue(s, r, id);
}

int nemoh_server_submit_waves(nemoh_server* s, int np, int nfa, int nsym, const double* x, const double* y,
                              const double* z, const int* m1, const int* m2, const int* m3, const int* m4,
                              double depth, double rho, double g, int precision, int nfreq, const double* omegas,
                              int nbeta, const double* betas, const int* body_of_panel, int nrad,
                              const nemoh_dof* rad, int nforce, const nemoh_dof* force, double xeff, double yeff,
                              long long* id) {
  if (!s || !id || !mesh_ok(np, nfa, nsym, x, y, z, m1, m2, m3, m4, depth, precision) || nfreq <= 0 || !omegas ||
      nbeta < 0 || nrad < 0 || nbeta + nrad < 1 || nbeta + nrad > 16 || (nbeta > 0 && !betas) ||
      (nrad > 0 && !rad) || nforce <= 0 || !force)
    return NEMOH_ERR_ARG;
  auto r = new_request(np, nfa, nsym, x, y, z, m1, m2, m3, m4, depth, rho, g, precision, nfreq);
  r->waves = true;
  r->omegas.assign(omegas, omegas + nfreq);
  const double pi = 4.0 * std::atan(1.0);
  for (int f = 0; f < nfreq; ++f) r->periods.push_back(2.0 * pi / omegas[f]);  // as nemoh_core_solve_waves
  r->nbeta = nbeta;
  if (nbeta > 0) r->betas.assign(betas, betas + nbeta);
  if (body_of_panel) r->body.assign(body_of_panel, body_of_panel + nfa);
  if (nrad > 0) r->rad.assign(rad, rad + nrad);
  r->force.assign(force, force + nforce);
  r->xeff = xeff;
  r->yeff = yeff;
  r->nrhs = nbeta + nrad;
  r->nint = nforce;
  r->forces.assign(2 * (size_t)nforce * r->nrhs * nfreq, 0.0);
  r->fk.assign(2 * (size_t)nforce * nbeta * nfreq, 0.0);
  r->footprint = (double)nemoh::core_footprint_bytes(np, nfa, nsym, precision, r->nrhs);
  return enqueue(s, r, id);
}

int nemoh_server_wait(nemoh_server* s, long long id, double* forces, double* fk, double* timing_ms) {
  if (!s) return NEMOH_ERR_ARG;
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->table.find(id);
    if (it == s->table.end()) {
      s->err = "unknown request id";
      return NEMOH_ERR_ARG;
    }
    r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->table.erase(id);
  }
  if (timing_ms) {
    timing_ms[0] = r->t_done - r->t_submit;   // latency
    timing_ms[1] = r->t_start - r->t_submit;  // wait for the first frequency to start
    timing_ms[2] = r->t_done - r->t_start;    // first start to last finish
  }
  std::lock_guard<std::mutex> lk(r->em);
  if (r->status) {
    std::lock_guard<std::mutex> lk2(s->mu);
    s->err = r->err;
    return r->status;
  }
  if (forces) std::memcpy(forces, r->forces.data(), r->forces.size() * sizeof(double));
  if (fk && !r->fk.empty()) std::memcpy(fk, r->fk.data(), r->fk.size() * sizeof(double));
  return NEMOH_OK;
}

int nemoh_server_set_batch(nemoh_server* s, int frequencies) {
  if (!s || frequencies < 1 || frequencies > 64) return NEMOH_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->batch = frequencies;
  return NEMOH_OK;
}

int nemoh_server_memory(nemoh_server* s, double* held, double* peak) {
  if (!s) return NEMOH_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  double h = 0;
  for (const Worker& w : s->workers) h += w.held;
  if (held) *held = h;
  if (peak) *peak = s->peak;
  return NEMOH_OK;
}

const char* nemoh_server_last_error(nemoh_server* s) {
  if (!s) return "null handle";
  std::lock_guard<std::mutex> lk(s->mu);
  return s->err.c_str();
}

void nemoh_server_destroy(nemoh_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->threads) t.join();
  for (Worker& w : s->workers) {
    if (w.batch) nemoh_batch_destroy(w.batch);
    if (w.main.h) nemoh_core_destroy(w.main.h);
    for (Slot& sl : w.more)
      if (sl.h) nemoh_core_destroy(sl.h);
  }
  delete s;
}

}  // extern "C"
