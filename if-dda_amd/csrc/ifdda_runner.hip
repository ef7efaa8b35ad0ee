// ifdda_runner.hip -- many independent solves on one GPU at once.
//
// A pool of worker threads, each with its own contexts and HIP streams,
// takes jobs (object, permittivity through alpha, wavelength, lattice,
// incidences) from a queue in submission order.  A job is admitted when its
// estimated device footprint fits the memory budget next to the jobs already
// running (a job larger than the budget runs alone), so requests that arrive
// while others are on the GPU start as soon as a worker and memory are free
// and their work overlaps on the device.  Each job is one multi-RHS solve
// (all its incidences share every matvec launch); a worker keeps the
// contexts of the last grids it served and rebuilds the Green tensor only
// when (d, k0, nquad) change.  Every kernel is deterministic, so a job's
// result does not depend on what else runs.
//
// This stands for the reference's callers that loop over objects or
// wavelengths and call cdmlib once per case (cdmlib.f:1-71, the GUI's
// cdmRun worker): here the cases run concurrently.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
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
  std::vector<double> alpha, einc;
  double est = 0.0;
  int state = 0;  // 0 queued, 1 running, 2 done
  int rc = IFDDA_OK;
  std::string err;
  std::vector<double> FFloc, FF;
  std::vector<ifdda_solve_info> infos;
  double t_submit = 0.0, t_start = 0.0, t_end = 0.0;
};

// Device bytes a job needs: Green tensor (FP64 + FP32 octant), FFT scratch
// for R right-hand sides (FP64), Krylov vectors (14 per precision and RHS),
// host-array staging.
double job_bytes(const ifdda_job& j) {
  const int L1 = ifdda::smooth235_at_least(2 * j.nx), L2 = ifdda::smooth235_at_least(2 * j.ny),
            L3 = ifdda::smooth235_at_least(2 * j.nz);
  const double F = (double)(L1 / 2 + 1) * (L2 / 2 + 1) * (L3 / 2 + 1);
  const double R = j.nrhs, n3 = 3.0 * j.nd;
  const double grid = (double)j.nz * j.ny * L1 * 3 + (double)j.nz * L2 * L1 * 3;
  const double kry = 14.0 * n3 * R * (16.0 + (j.opts.precision == 32 ? 8.0 : 0.0));
  return F * 6 * 24 + grid * 16 * R + kry + 2.0 * n3 * R * 16 + 4.0 * j.nx * j.ny * j.nz + (1 << 20);
}

}  // namespace

struct ifdda_runner {
  int nworkers = 1;
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

struct CtxEntry {
  std::array<int, 3> box;
  ifdda_ctx* ctx;
  double d, k0;
  int nquad;
  double tolinit;
};

// Run one job on this worker's contexts (at most kKeep kept, least recent
// first out).
constexpr size_t kKeep = 2;

int run_job(std::list<CtxEntry>& cache, Job& jb) {
  const ifdda_job& s = jb.spec;
  const std::array<int, 3> box{s.nx, s.ny, s.nz};
  auto it = std::find_if(cache.begin(), cache.end(), [&](const CtxEntry& e) { return e.box == box; });
  if (it == cache.end()) {
    ifdda_ctx* c = nullptr;
    int rc = ifdda_create(s.nx, s.ny, s.nz, &c);
    if (rc != IFDDA_OK) return rc;
    cache.push_front({box, c, -1.0, -1.0, -1, -1.0});
    while (cache.size() > kKeep) {
      ifdda_destroy(cache.back().ctx);
      cache.pop_back();
    }
    it = cache.begin();
  } else {
    cache.splice(cache.begin(), cache, it);
    it = cache.begin();
  }
  CtxEntry& e = *it;
  ifdda_ctx* c = e.ctx;
  int rc = IFDDA_OK;
  if (e.d != s.d || e.k0 != s.k0 || e.nquad != s.nquad || e.tolinit != s.tolinit) {
    rc = ifdda_green_build(c, s.d, s.k0, s.nquad, s.tolinit);
    if (rc != IFDDA_OK) return rc;
    e.d = s.d, e.k0 = s.k0, e.nquad = s.nquad, e.tolinit = s.tolinit;
  }
  rc = ifdda_set_dipoles(c, s.tabdip ? jb.tab.data() : nullptr, s.nd);
  if (rc == IFDDA_OK) rc = ifdda_set_polarizability(c, jb.alpha.data(), s.nd, 1);
  if (rc != IFDDA_OK) return rc;
  jb.FFloc.assign(2 * 3 * (size_t)s.nd * s.nrhs, 0.0);
  jb.FF.assign(jb.FFloc.size(), 0.0);
  jb.infos.assign(s.nrhs, ifdda_solve_info{});
  volatile int stop = 0;
  return ifdda_solve_multi(c, &s.opts, s.nrhs, jb.einc.data(), jb.FF.data(), jb.FFloc.data(),
                           jb.infos.data(), &stop);
}

void worker(ifdda_runner* r) {
  (void)hipSetDevice(r->device);
  std::list<CtxEntry> cache;
  for (;;) {
    long long t = 0;
    Job* jb = nullptr;
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
      t = r->queue.front();
      r->queue.pop_front();
      jb = &r->jobs.at(t);
      jb->state = 1;
      jb->t_start = now_s();
      r->used += jb->est;
      r->running++;
    }
    const int rc = run_job(cache, *jb);
    std::string err = rc == IFDDA_OK ? "" : ifdda_last_error();
    {
      std::lock_guard<std::mutex> lk(r->mu);
      jb->rc = rc;
      jb->err = err;
      jb->state = 2;
      jb->t_end = now_s();
      r->used -= jb->est;
      r->running--;
    }
    r->cv_done.notify_all();
    r->cv_work.notify_all();
  }
  for (auto& e : cache) ifdda_destroy(e.ctx);
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
      !j->alpha || !j->E_inc || !(j->d > 0.0) || !(j->k0 > 0.0)) {
    ifdda::set_error("ifdda_runner_submit: bad job");
    return IFDDA_ERR_ARG;
  }
  Job jb;
  jb.spec = *j;
  const size_t N = (size_t)j->nx * j->ny * j->nz;
  if (j->tabdip) jb.tab.assign(j->tabdip, j->tabdip + N);
  jb.alpha.assign(j->alpha, j->alpha + 2 * (size_t)j->nd);
  jb.einc.assign(j->E_inc, j->E_inc + 2 * 3 * (size_t)j->nd * j->nrhs);
  jb.spec.tabdip = nullptr;  // the copies are used
  if (j->tabdip) jb.spec.tabdip = jb.tab.data();
  jb.spec.alpha = nullptr;
  jb.spec.E_inc = nullptr;
  jb.est = job_bytes(*j);
  jb.t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(r->mu);
    const long long t = r->next++;
    auto ins = r->jobs.emplace(t, std::move(jb));
    if (j->tabdip) ins.first->second.spec.tabdip = ins.first->second.tab.data();
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
