// ifdda_runner.hip -- many independent solves on one GPU at once.
//
// A pool of worker threads takes jobs (object, permittivity through alpha,
// wavelength, lattice, incidences) from a queue in submission order.  A free
// worker takes the job at the head and, with it, every queued job with the
// same solve options and a box within a factor 2 in cells that fits the pack
// (at most max_members right-hand sides) and the memory budget next to the packs already running (a job
// larger than the budget runs alone); the group is solved as one pack
// (ifdda_pack.hip): all its right-hand sides share every matvec launch and
// one lockstep Krylov solve, while each keeps its own iterations and
// convergence.  Requests that arrive while packs run queue up and start
// together when a worker frees; packs of different workers overlap on the
// device.  Each worker keeps its FFT tables, recent Green tensors and buffers
// between packs.  Every kernel is deterministic and a member's arithmetic
// does not depend on the rest of its pack, so a job's result does not depend
// on what else runs.
//
// This stands for the reference's callers that loop over objects or
// wavelengths and call cdmlib once per case (cdmlib.f:1-71, the GUI's
// cdmRun worker): here the cases run concurrently.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ifdda.h"
#include "ifdda_internal.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Job {
  ifdda_job spec{};
  std::vector<int32_t> tab;
  std::vector<double> alpha, einc;  // or eps, wave parameters
  double est = 0.0;
  int state = 0;  // 0 queued, 1 running, 2 done
  int rc = IFDDA_OK;
  std::string err;
  std::vector<double> FFloc, FF;
  std::vector<ifdda_solve_info> infos;
  double t_submit = 0.0, t_start = 0.0, t_end = 0.0;
};

// Device bytes a job needs in a pack: FFT scratch and vectors per
// right-hand side, and its Green tensor.
double job_bytes(const ifdda_job& j) {
  return (double)ifdda::pack_member_bytes(j) * j.nrhs + (double)ifdda::pack_green_bytes(j) +
         (double)(1 << 20);
}

bool same_opts(const ifdda_solve_opts& a, const ifdda_solve_opts& b) {
  return a.method == b.method && a.precision == b.precision && a.x0_zero == b.x0_zero &&
         a.maxit == b.maxit && a.tol == b.tol && a.check_every == b.check_every;
}

// Jobs that may share a pack: the same solve options and boxes within a
// factor 2 in cells.  A pack runs in lockstep until its slowest member has
// converged; a small job packed with a much larger one (more cells, most
// often more iterations) would wait for it.
// Jobs of at least this many box cells (an 80^3 box) run alone through a
// context of their own (ifdda_solve_multi_device: the compile-time z pass
// of their grid, the tuned tile widths, the fused GPBICG products) rather
// than in a pack, whose launches are generic; they never share a pack, so a
// large job's result is the same whatever else is queued.
constexpr double kLargeCells = 512000.0;
bool is_large(const ifdda_job& j) { return (double)j.nx * j.ny * j.nz >= kLargeCells; }

bool packable(const ifdda_job& a, const ifdda_job& head) {
  const double na = (double)a.nx * a.ny * a.nz, nh = (double)head.nx * head.ny * head.nz;
  return same_opts(a.opts, head.opts) && na <= 2.0 * nh && nh <= 2.0 * na && !is_large(a) &&
         !is_large(head);
}

// A worker's context for large jobs: kept for the last box, with the Green
// tensor of the last (d, k0, nquad) and grow-only device vectors.
struct CtxSlot {
  ifdda_ctx* c = nullptr;
  int nx = 0, ny = 0, nz = 0;
  bool green = false;
  double d = 0.0, k0 = 0.0;
  int nquad = -1;
  double2* buf = nullptr;  // E_inc, FF, FFloc
  size_t cap = 0;          // complex values
  void release() {
    if (c) ifdda_destroy(c);
    (void)hipFree(buf);
    *this = CtxSlot{};
  }
};

// One large job through its context: Green tensor, map, polarisability and
// incident waves on the device, one multi-RHS solve, results to the job.
int ctx_solve(CtxSlot& w, Job& j) {
  const ifdda_job& s = j.spec;
  if (!w.c || w.nx != s.nx || w.ny != s.ny || w.nz != s.nz) {
    w.release();
    int rc = ifdda_create(s.nx, s.ny, s.nz, &w.c);
    if (rc) return rc;
    w.nx = s.nx, w.ny = s.ny, w.nz = s.nz;
  }
  ifdda_ctx* c = w.c;
  int rc = IFDDA_OK;
  if (!w.green || w.d != s.d || w.k0 != s.k0 || w.nquad != s.nquad) {
    w.green = false;
    if ((rc = ifdda_green_build(c, s.d, s.k0, s.nquad, s.tolinit))) return rc;
    w.green = true, w.d = s.d, w.k0 = s.k0, w.nquad = s.nquad;
  }
  if ((rc = ifdda_set_dipoles(c, s.tabdip, s.nd))) return rc;
  rc = s.alpha ? ifdda_set_polarizability(c, s.alpha, s.nd, 1)
               : ifdda_polarizability_device(c, s.eps, s.d, s.k0, s.eps0 > 0.0 ? s.eps0 : 1.0, s.pola);
  if (rc) return rc;
  const size_t n = 3 * (size_t)s.nd * s.nrhs;
  if (3 * n > w.cap) {
    (void)hipFree(w.buf);
    w.buf = nullptr;
    w.cap = 0;
    if (hipMalloc(&w.buf, sizeof(double2) * 3 * n) != hipSuccess) {
      ifdda::set_error("runner: hipMalloc of a large job's vectors failed");
      return IFDDA_ERR_HIP;
    }
    w.cap = 3 * n;
  }
  double2 *E = w.buf, *FF = w.buf + n, *FL = w.buf + 2 * n;
  if (s.E_inc) {
    if (hipMemcpy(E, s.E_inc, sizeof(double2) * n, hipMemcpyHostToDevice) != hipSuccess) {
      ifdda::set_error("runner: copy of a large job's incident field failed");
      return IFDDA_ERR_HIP;
    }
  } else if ((rc = ifdda_incident_device(c, s.wave_kind, s.nrhs, s.waves, s.E0, s.k0, s.d, s.origin,
                                         E, nullptr))) {
    return rc;
  }
  volatile int stop = 0;
  if ((rc = ifdda_solve_multi_device(c, &s.opts, s.nrhs, E, FF, FL, j.infos.data(), &stop, nullptr)))
    return rc;
  if (hipMemcpy(j.FFloc.data(), FL, sizeof(double2) * n, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(j.FF.data(), FF, sizeof(double2) * n, hipMemcpyDeviceToHost) != hipSuccess) {
    ifdda::set_error("runner: copy of a large job's results failed");
    return IFDDA_ERR_HIP;
  }
  return IFDDA_OK;
}

}  // namespace

struct ifdda_runner {
  int nworkers = 1;
  int max_members = 64;
  double budget = 0.0;
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<long long> queue;
  std::map<long long, Job> jobs;
  double used = 0.0;
  int running = 0;
  bool stop = false;
  long long next = 1;
  std::vector<std::thread> threads;
};

namespace {

void worker(ifdda_runner* r) {
  (void)hipSetDevice(r->device);
  ifdda::PackWs* ws = ifdda::pack_ws_new();
  CtxSlot cs;
  // Green tensors kept between packs: a share of the budget per worker
  if (ws)
    ifdda::pack_ws_green_limit(
        ws, (size_t)std::min(16e9, std::max(1e9, r->budget / (4.0 * r->nworkers))));
  for (;;) {
    std::vector<Job*> group;
    {
      std::unique_lock<std::mutex> lk(r->mu);
      r->cv_work.wait(lk, [&] {
        if (r->stop) return true;
        if (r->queue.empty()) return false;
        const Job& head = r->jobs.at(r->queue.front());
        return r->running == 0 || r->used + head.est <= r->budget;
      });
      if (r->stop && r->queue.empty()) break;
      if (r->queue.empty()) continue;
      Job* head = &r->jobs.at(r->queue.front());
      r->queue.pop_front();
      group.push_back(head);
      r->used += head->est;
      int members = is_large(head->spec) ? r->max_members : head->spec.nrhs;
      for (auto it = r->queue.begin(); it != r->queue.end() && members < r->max_members;) {
        Job* j = &r->jobs.at(*it);
        if (packable(j->spec, head->spec) && members + j->spec.nrhs <= r->max_members &&
            r->used + j->est <= r->budget) {
          group.push_back(j);
          r->used += j->est;
          members += j->spec.nrhs;
          it = r->queue.erase(it);
        } else {
          ++it;
        }
      }
      const double t = now_s();
      for (Job* j : group) j->state = 1, j->t_start = t;
      r->running++;
    }
    std::vector<ifdda::PackIn> in;
    for (Job* j : group) {
      const size_t n = 2 * 3 * (size_t)j->spec.nd * j->spec.nrhs;
      j->FFloc.assign(n, 0.0);
      j->FF.assign(n, 0.0);
      j->infos.assign(j->spec.nrhs, ifdda_solve_info{});
      in.push_back(ifdda::PackIn{&j->spec, j->FFloc.data(), j->FF.data(), j->infos.data()});
    }
    int rc = IFDDA_ERR_HIP;
    std::string err = "no pack workspace";
    if (is_large(group[0]->spec)) {
      rc = ctx_solve(cs, *group[0]);
      err = rc == IFDDA_OK ? "" : ifdda_last_error();
    } else if (ws) {
      volatile int stop = 0;
      rc = ifdda::pack_solve(*ws, in, &group[0]->spec.opts, &stop);
      err = rc == IFDDA_OK ? "" : ifdda_last_error();
    }
    {
      std::lock_guard<std::mutex> lk(r->mu);
      const double t = now_s();
      for (Job* j : group) {
        j->rc = rc;
        j->err = err;
        j->state = 2;
        j->t_end = t;
        r->used -= j->est;
      }
      r->running--;
    }
    r->cv_done.notify_all();
    r->cv_work.notify_all();
  }
  cs.release();
  ifdda::pack_ws_delete(ws);
}

}  // namespace

extern "C" int ifdda_runner_create(int nworkers, double mem_budget, ifdda_runner** out) {
  if (!out || nworkers < 1 || nworkers > 64 || !(mem_budget > 0.0)) {
    ifdda::set_error("ifdda_runner_create: bad argument (1..64 workers, budget > 0)");
    return IFDDA_ERR_ARG;
  }
  ifdda_runner* r = new ifdda_runner();
  r->nworkers = nworkers;
  r->budget = mem_budget;
  if (hipGetDevice(&r->device) != hipSuccess) r->device = 0;
  for (int i = 0; i < nworkers; ++i) r->threads.emplace_back(worker, r);
  *out = r;
  return IFDDA_OK;
}

extern "C" int ifdda_runner_submit(ifdda_runner* r, const ifdda_job* j, long long* ticket) {
  if (!r || !j || !ticket || j->nx < 1 || j->ny < 1 || j->nz < 1 || j->nd < 1 || j->nrhs < 1 ||
      (!j->alpha && !j->eps) || (!j->E_inc && !j->waves) || !(j->d > 0.0) || !(j->k0 > 0.0)) {
    ifdda::set_error("ifdda_runner_submit: bad job");
    return IFDDA_ERR_ARG;
  }
  const size_t N = (size_t)j->nx * j->ny * j->nz;
  if (!j->tabdip && (size_t)j->nd != N) {
    ifdda::set_error("ifdda_runner_submit: bad job (a dense box needs nd = nx ny nz)");
    return IFDDA_ERR_ARG;
  }
  if (j->tabdip) {  // the map must name each dipole 1..nd once (a bad job fails alone)
    std::vector<char> seen((size_t)j->nd + 1, 0);
    size_t count = 0;
    for (size_t i = 0; i < N; ++i) {
      const int32_t t = j->tabdip[i];
      if (t < 0 || t > j->nd || (t > 0 && seen[t])) {
        ifdda::set_error("ifdda_runner_submit: bad job (Tabdip entry out of range or repeated)");
        return IFDDA_ERR_ARG;
      }
      if (t > 0) seen[t] = 1, ++count;
    }
    if (count != (size_t)j->nd) {
      ifdda::set_error("ifdda_runner_submit: bad job (Tabdip maps a different number of dipoles)");
      return IFDDA_ERR_ARG;
    }
  }
  Job jb;
  jb.spec = *j;
  if (j->tabdip) jb.tab.assign(j->tabdip, j->tabdip + N);
  if (j->alpha)
    jb.alpha.assign(j->alpha, j->alpha + 2 * (size_t)j->nd);
  else
    jb.alpha.assign(j->eps, j->eps + 2 * (size_t)j->nd);
  if (j->E_inc)
    jb.einc.assign(j->E_inc, j->E_inc + 2 * 3 * (size_t)j->nd * j->nrhs);
  else
    jb.einc.assign(j->waves, j->waves + 3 * (size_t)j->nrhs);
  jb.est = job_bytes(*j);
  jb.t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(r->mu);
    const long long t = r->next++;
    Job& q = r->jobs.emplace(t, std::move(jb)).first->second;
    q.spec.tabdip = j->tabdip ? q.tab.data() : nullptr;  // the copies are used
    (j->alpha ? q.spec.alpha : q.spec.eps) = q.alpha.data();
    (j->E_inc ? q.spec.E_inc : q.spec.waves) = q.einc.data();
    r->queue.push_back(t);
    *ticket = t;
  }
  r->cv_work.notify_all();
  return IFDDA_OK;
}

extern "C" int ifdda_runner_wait(ifdda_runner* r, long long ticket, double* FFloc, double* FF,
                                 ifdda_solve_info* infos, double* times) {
  if (!r) {
    ifdda::set_error("ifdda_runner_wait: NULL runner");
    return IFDDA_ERR_ARG;
  }
  Job jb;
  {
    std::unique_lock<std::mutex> lk(r->mu);
    auto it = r->jobs.find(ticket);
    if (it == r->jobs.end()) {
      ifdda::set_error("ifdda_runner_wait: unknown ticket");
      return IFDDA_ERR_ARG;
    }
    r->cv_done.wait(lk, [&] { return r->jobs.at(ticket).state == 2; });
    jb = std::move(r->jobs.at(ticket));
    r->jobs.erase(ticket);
  }
  if (times) {
    times[0] = jb.t_start - jb.t_submit;  // queued
    times[1] = jb.t_end - jb.t_start;     // running
    times[2] = jb.t_end - jb.t_submit;    // latency
  }
  if (jb.rc != IFDDA_OK) {
    ifdda::set_error("ifdda_runner job: " + jb.err);
    return jb.rc;
  }
  if (FFloc) std::memcpy(FFloc, jb.FFloc.data(), sizeof(double) * jb.FFloc.size());
  if (FF) std::memcpy(FF, jb.FF.data(), sizeof(double) * jb.FF.size());
  if (infos) std::memcpy(infos, jb.infos.data(), sizeof(ifdda_solve_info) * jb.infos.size());
  return IFDDA_OK;
}

extern "C" int ifdda_runner_pack_limit(ifdda_runner* r, int max_members) {
  if (!r || max_members < 1) {
    ifdda::set_error("ifdda_runner_pack_limit: bad argument");
    return IFDDA_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(r->mu);
  r->max_members = max_members;
  return IFDDA_OK;
}

extern "C" int ifdda_runner_destroy(ifdda_runner* r) {
  if (!r) return IFDDA_OK;
  {
    std::lock_guard<std::mutex> lk(r->mu);
    r->stop = true;
  }
  r->cv_work.notify_all();
  for (auto& t : r->threads) t.join();
  delete r;
  return IFDDA_OK;
}
