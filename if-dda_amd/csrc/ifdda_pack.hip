// ifdda_pack.hip -- independent problems solved together.
//
// A pack holds P members; a member is one right-hand side (incidence) of one
// problem (job), and every job brings its own box, padded FFT lengths, Green
// tensor (wavelength, lattice spacing), dipole map and polarisability.  Each
// matvec of the pack is five packed launches over the tiles of all members
// (launch_matvec_pack_*), and the Krylov solve runs the members in lockstep
// with per-member state, length and convergence test (solve_core on a ragged
// vector layout), so a member's iterations and results are those of its own
// solve: the member's tiles, reductions and their order are the same.
// Members that have converged drop out of the launches.
//
// This stands for the reference's callers that solve many independent cases
// one cdmlib call at a time (cdmlib.f:1-71; the GUI's cdmRun worker): here
// their launches are shared.  A workspace keeps, between packs, the FFT
// tables per length, the Green tensors per (box, d, k0, nquad) (least
// recently used out past a byte limit), and grow-only device buffers, so that
// a stream of packs allocates and frees nothing in steady state; every copy
// is ordered on the workspace's own non-blocking stream.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/ifdda.h"
#include "ifdda_internal.h"

namespace ifdda {

namespace {

struct PackTable {
  double2* tw64 = nullptr;
  float2* tw32 = nullptr;
  double2* tws64 = nullptr;
  float2* tws32 = nullptr;
  int* pos = nullptr;
};

struct GreenKey {
  int nx, ny, nz, nquad;
  double d, k0;
  bool operator<(const GreenKey& o) const {
    return std::tie(nx, ny, nz, nquad, d, k0) < std::tie(o.nx, o.ny, o.nz, o.nquad, o.d, o.k0);
  }
};

struct GreenEntry {
  GridPlan g{};
  double2* G64 = nullptr;
  float2* G32 = nullptr;
  size_t bytes = 0;
  unsigned long long used = 0;  // pack counter of the last use
};

size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

struct PackWs {
  hipStream_t s = nullptr;
  size_t lds_limit = 65536;
  std::map<int, PackTable> tables;
  std::map<GreenKey, GreenEntry> greens;
  size_t green_bytes = 0;
  size_t green_limit = (size_t)16 << 30;
  unsigned long long tick = 0;
  char* arena = nullptr;  // per-pack data, grow-only
  size_t arena_cap = 0;
  char* blob = nullptr;  // packed-launch argument blocks, grow-only
  size_t blob_cap = 0;
  char* hstage = nullptr;  // pinned host image of a pack's inputs / outputs, grow-only
  size_t hstage_cap = 0;
  Solver* S = nullptr;
  std::vector<long long> lens, offs;  // host layout of the last pack (read by s)
  // Green tensors of a pack's new (box, d, k0) keys are built side by side on
  // these streams, each with its own full-grid scratch; s waits for them.
  static constexpr int kAux = 4;
  hipStream_t aux[kAux] = {};
  hipEvent_t ev_aux[kAux] = {}, ev_main = nullptr;
  char* afull[kAux] = {};
  size_t afull_cap[kAux] = {};
  GreenScratch gs[kAux];  // near-cell quadrature buffers per stream
};

namespace {

// IFDDA_DEBUG set in the environment: one line on stderr per buffer growth.
bool debug_on() {
  static const bool on = std::getenv("IFDDA_DEBUG") != nullptr;
  return on;
}

// IFDDA_DEBUG: the seconds each phase of a pack took (the stream is drained
// at each mark, so the marks cost time of their own).
struct PhaseClock {
  hipStream_t s;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  std::string line;
  void mark(const char* what) {
    if (!debug_on()) return;
    (void)hipStreamSynchronize(s);
    const auto n = std::chrono::steady_clock::now();
    char b[64];
    std::snprintf(b, sizeof(b), " %s %.1f ms", what,
                  std::chrono::duration<double, std::milli>(n - t).count());
    line += b;
    t = n;
  }
};

// Grow-only device buffer: at least doubles, so a worker meets O(log) growths
// (each frees the old buffer, which waits for the device).
bool grow(char** p, size_t* cap, size_t need, const char* what) {
  if (need <= *cap) return true;
  const size_t c = std::max(need + need / 4, 2 * *cap);
  if (debug_on()) std::fprintf(stderr, "ifdda pack: %s %zu -> %zu MB\n", what, *cap >> 20, c >> 20);
  (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (!hip_ok(hipMalloc(p, c), what)) return false;
  *cap = c;
  return true;
}

// Grow-only pinned host buffer (the old one is no longer read: callers
// drain the stream first).
bool grow_host(char** p, size_t* cap, size_t need) {
  if (need <= *cap) return true;
  const size_t c = std::max(need + need / 4, 2 * *cap);
  if (debug_on()) std::fprintf(stderr, "ifdda pack: pinned staging %zu -> %zu MB\n", *cap >> 20, c >> 20);
  (void)hipHostFree(*p);
  *p = nullptr;
  *cap = 0;
  if (!hip_ok(hipHostMalloc(reinterpret_cast<void**>(p), c, hipHostMallocDefault),
              "hipHostMalloc pack staging"))
    return false;
  *cap = c;
  return true;
}

// f(0) .. f(n - 1) on up to 8 host threads when the work moves at least
// 32 MB, else in order on this thread.
template <typename F> void par_for(int n, size_t bytes, F&& f) {
  const int nt = bytes >= ((size_t)32 << 20) ? std::min(8, n) : 1;
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (int i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto& t : th) t.join();
}

const PackTable* table_for(PackWs& w, const AxisPlan& ax) {
  auto it = w.tables.find(ax.L);
  if (it != w.tables.end()) return &it->second;
  const int L = ax.L;
  std::vector<double2> t64(L), s64(L);
  std::vector<float2> t32(L), s32(L);
  std::vector<int> pos(L);
  axis_tables(ax, t64.data(), t32.data(), s64.data(), s32.data(), pos.data());
  PackTable t;
  bool ok = hip_ok(hipMalloc(&t.tw64, sizeof(double2) * L), "hipMalloc tables") &&
            hip_ok(hipMalloc(&t.tw32, sizeof(float2) * L), "hipMalloc tables") &&
            hip_ok(hipMalloc(&t.tws64, sizeof(double2) * L), "hipMalloc tables") &&
            hip_ok(hipMalloc(&t.tws32, sizeof(float2) * L), "hipMalloc tables") &&
            hip_ok(hipMalloc(&t.pos, sizeof(int) * L), "hipMalloc tables") &&
            hip_ok(hipMemcpyAsync(t.tw64, t64.data(), sizeof(double2) * L, hipMemcpyHostToDevice, w.s), "copy tables") &&
            hip_ok(hipMemcpyAsync(t.tw32, t32.data(), sizeof(float2) * L, hipMemcpyHostToDevice, w.s), "copy tables") &&
            hip_ok(hipMemcpyAsync(t.tws64, s64.data(), sizeof(double2) * L, hipMemcpyHostToDevice, w.s), "copy tables") &&
            hip_ok(hipMemcpyAsync(t.tws32, s32.data(), sizeof(float2) * L, hipMemcpyHostToDevice, w.s), "copy tables") &&
            hip_ok(hipMemcpyAsync(t.pos, pos.data(), sizeof(int) * L, hipMemcpyHostToDevice, w.s), "copy tables") &&
            hip_ok(hipStreamSynchronize(w.s), "copy tables");  // the host vectors end here
  if (!ok) {
    (void)hipFree(t.tw64), (void)hipFree(t.tw32), (void)hipFree(t.tws64), (void)hipFree(t.tws32);
    (void)hipFree(t.pos);
    return nullptr;
  }
  return &w.tables.emplace(L, t).first->second;
}

// G32 lives in the same allocation as G64, after it.
void green_free(GreenEntry& e) {
  (void)hipFree(e.G64);
  e.G64 = nullptr;
  e.G32 = nullptr;
}

// The scalar polarisability of a job (nd complex) in FP64 and FP32.
struct JobDev {
  const GreenEntry* green = nullptr;
  const PackTable* tab3[3] = {};
  int32_t* tab = nullptr;  // null: dense box
  unsigned long long* rows = nullptr;  // line masks of tab (null: none)
  double2* al64 = nullptr;
  float2* al32 = nullptr;
  long long nd = 0;
};

struct PackOp : MvOp {
  mutable PackLayout lay64, lay32;  // block tables follow compaction
  const void* blob64 = nullptr;
  const void* blob32 = nullptr;
  // the same launches without alpha, for QMRCLA's A^T = I - alpha^T T
  // (inverserig.f:559-590; alpha is one scalar per dipole in a pack)
  const void* blobt64 = nullptr;
  const void* blobt32 = nullptr;
  std::vector<const double2*> al;  // per member
  std::vector<const float2*> al32;
  std::vector<long long> nd, off;
  bool mv64(const double2* x, const double2* b, double2* y, hipStream_t s, const int* skip, int,
            long long, int skip_stride) const override {
    return launch_matvec_pack_f64(lay64, blob64, x, b, y, s, skip, skip_stride);
  }
  bool mv32(const float2* x, const float2* b, float2* y, hipStream_t s, const int* skip, int,
            long long, int skip_stride) const override {
    return launch_matvec_pack_f32(lay32, blob32, x, b, y, s, skip, skip_stride);
  }
  bool alpha_apply(const double2* x, double2* FF, int R, long long, hipStream_t s) const override {
    for (int m = 0; m < R; ++m)
      if (!alpha_apply_one(x + off[m], FF + off[m], al[m], kAlphaScalar, nd[m], s)) return false;
    return true;
  }
  template <typename V>
  bool transposed(const V* x, const V* zero, V* y, hipStream_t s, const int* skip, int R,
                  int skip_stride) const {
    bool ok;
    if constexpr (std::is_same<V, double2>::value)
      ok = launch_matvec_pack_f64(lay64, blobt64, x, zero, y, s, skip, skip_stride);
    else
      ok = launch_matvec_pack_f32(lay32, blobt32, x, zero, y, s, skip, skip_stride);
    for (int m = 0; ok && m < R; ++m) {
      const V* a;
      if constexpr (std::is_same<V, double2>::value) a = al[m];
      else a = al32[m];
      ok = alpha_t_launch<V>(x + off[m], y + off[m], a, kAlphaScalar, nd[m], s,
                             skip ? skip + (long long)m * skip_stride : nullptr);
    }
    return ok;
  }
  bool mvt64(const double2* x, const double2* zero, double2* y, hipStream_t s, const int* skip,
             int R, long long, int skip_stride) const override {
    return transposed<double2>(x, zero, y, s, skip, R, skip_stride);
  }
  // Finished members dispatch no blocks: every blob's tables are rewritten
  // (the transposed blobs share the layouts of the plain ones).
  bool compact(const KState* st, const int* host_done, int, bool all, hipStream_t s) const override {
    return pack_compact(lay64, const_cast<void*>(blob64), st, host_done, all, s) &&
           pack_compact(lay32, const_cast<void*>(blob32), st, host_done, all, s) &&
           (!blobt64 || pack_compact(lay64, const_cast<void*>(blobt64), st, host_done, all, s)) &&
           (!blobt32 || pack_compact(lay32, const_cast<void*>(blobt32), st, host_done, all, s));
  }
  bool mvt32(const float2* x, const float2* zero, float2* y, hipStream_t s, const int* skip, int R,
             long long, int skip_stride) const override {
    return transposed<float2>(x, zero, y, s, skip, R, skip_stride);
  }
};

}  // namespace

PackWs* pack_ws_new() {
  PackWs* w = new PackWs();
  int dev = 0, lds = 65536;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    lds = 65536;
  w->lds_limit = (size_t)std::min(lds, 65536);
  bool ok = hip_ok(hipStreamCreateWithFlags(&w->s, hipStreamNonBlocking), "hipStreamCreate") &&
            hip_ok(hipEventCreateWithFlags(&w->ev_main, hipEventDisableTiming), "hipEventCreate");
  for (int k = 0; ok && k < PackWs::kAux; ++k)
    ok = hip_ok(hipStreamCreateWithFlags(&w->aux[k], hipStreamNonBlocking), "hipStreamCreate") &&
         hip_ok(hipEventCreateWithFlags(&w->ev_aux[k], hipEventDisableTiming), "hipEventCreate");
  if (!ok) {
    pack_ws_delete(w);
    return nullptr;
  }
  return w;
}

void pack_ws_green_limit(PackWs* w, size_t bytes) { w->green_limit = bytes; }

void pack_ws_delete(PackWs* w) {
  if (!w) return;
  if (w->s) (void)hipStreamSynchronize(w->s);
  for (auto& t : w->tables) {
    (void)hipFree(t.second.tw64), (void)hipFree(t.second.tw32);
    (void)hipFree(t.second.tws64), (void)hipFree(t.second.tws32), (void)hipFree(t.second.pos);
  }
  for (auto& g : w->greens) green_free(g.second);
  for (int k = 0; k < PackWs::kAux; ++k) {
    if (w->aux[k]) (void)hipStreamSynchronize(w->aux[k]);
    (void)hipFree(w->afull[k]);
    green_scratch_free(w->gs[k]);
    if (w->aux[k]) (void)hipStreamDestroy(w->aux[k]);
    if (w->ev_aux[k]) (void)hipEventDestroy(w->ev_aux[k]);
  }
  if (w->ev_main) (void)hipEventDestroy(w->ev_main);
  (void)hipFree(w->arena);
  (void)hipFree(w->blob);
  (void)hipHostFree(w->hstage);
  solver_delete(w->S);
  if (w->s) (void)hipStreamDestroy(w->s);
  delete w;
}

size_t pack_member_bytes(const ifdda_job& j) {
  const int L1 = smooth235_at_least(2 * j.nx), L2 = smooth235_at_least(2 * j.ny);
  const size_t grid = ((size_t)j.nz * j.ny * L1 + (size_t)j.nz * L2 * L1) * 3 * sizeof(double2);
  const size_t vec = (size_t)3 * j.nd * (14 * 24 + 16 + 3 * 16);  // Krylov, refinement, E/FF/FFloc
  return grid + vec;
}

size_t pack_green_bytes(const ifdda_job& j) {
  const size_t F = (size_t)(smooth235_at_least(2 * j.nx) / 2 + 1) *
                   (smooth235_at_least(2 * j.ny) / 2 + 1) * (smooth235_at_least(2 * j.nz) / 2 + 1);
  return F * 6 * (sizeof(double2) + sizeof(float2));
}

int pack_solve(PackWs& w, const std::vector<PackIn>& in, const ifdda_solve_opts* o,
               volatile int* nstop) {
  const auto t0 = std::chrono::steady_clock::now();
  PhaseClock clk{w.s};
  const int J = (int)in.size();
  if (J == 0) return IFDDA_OK;
  if (o->method < IFDDA_GPBICG1 || o->method > IFDDA_QMRCLA || (o->precision != 64 && o->precision != 32)) {
    set_error("pack solve: unknown method or precision");
    return IFDDA_ERR_ARG;
  }
  // ---- validate the jobs, plan their boxes
  std::vector<GridPlan> plans(J);
  std::vector<GreenKey> keys(J);
  std::vector<char> dense(J, 0);
  for (int j = 0; j < J; ++j) {
    const ifdda_job& s = *in[j].job;
    if (s.nx < 1 || s.ny < 1 || s.nz < 1 || s.nd < 1 || s.nrhs < 1 || !(s.d > 0.0) ||
        !(s.k0 > 0.0) || s.nquad < 0 || s.nquad > 10 || (!s.alpha && !s.eps) ||
        (!s.E_inc && !s.waves)) {
      set_error("pack solve: bad job " + std::to_string(j));
      return IFDDA_ERR_ARG;
    }
    if (!s.alpha && (s.pola < IFDDA_POLA_CM || s.pola > IFDDA_POLA_LR || s.eps0 < 0.0)) {
      set_error("pack solve: job " + std::to_string(j) + ": unknown polarisability method");
      return IFDDA_ERR_ARG;
    }
    const size_t N = (size_t)s.nx * s.ny * s.nz;
    if (!s.tabdip && (size_t)s.nd != N) {
      set_error("pack solve: a dense box needs nd = nx ny nz (job " + std::to_string(j) + ")");
      return IFDDA_ERR_ARG;
    }
    if (s.tabdip) {
      std::vector<char> seen((size_t)s.nd + 1, 0);
      bool identity = (size_t)s.nd == N;
      size_t count = 0;
      for (size_t i = 0; i < N; ++i) {
        const int32_t t = s.tabdip[i];
        if (t < 0 || t > s.nd || (t > 0 && seen[t])) {
          set_error("pack solve: Tabdip of job " + std::to_string(j) +
                    " out of range or mapping a dipole twice at cell " + std::to_string(i));
          return IFDDA_ERR_ARG;
        }
        if (t > 0) seen[t] = 1, ++count;
        identity = identity && t == (int32_t)(i + 1);
      }
      if (count != (size_t)s.nd) {
        set_error("pack solve: Tabdip of job " + std::to_string(j) + " maps " +
                  std::to_string(count) + " dipoles, nd = " + std::to_string(s.nd));
        return IFDDA_ERR_ARG;
      }
      dense[j] = identity;
    } else {
      dense[j] = 1;
    }
    std::string why;
    if (!make_grid_plan(s.nx, s.ny, s.nz, w.lds_limit, &plans[j], &why)) {
      set_error("pack solve: job " + std::to_string(j) + ": " + why);
      return IFDDA_ERR_ARG;
    }
    keys[j] = GreenKey{s.nx, s.ny, s.nz, s.nquad, s.d, s.k0};
  }
  // incident waves formed on the device: (k, e) per right-hand side
  std::vector<std::vector<double>> wk(J);
  std::vector<std::vector<double2>> we(J);
  for (int j = 0; j < J; ++j) {
    const ifdda_job& s = *in[j].job;
    if (s.E_inc) continue;
    wk[j].resize(3 * (size_t)s.nrhs);
    we[j].resize(3 * (size_t)s.nrhs);
    for (int r = 0; r < s.nrhs; ++r) {
      const std::string msg = wave_vectors(s.wave_kind, s.waves + 3 * r,
                                           make_double2(s.E0[0], s.E0[1]), s.k0, &wk[j][3 * r],
                                           &we[j][3 * r]);
      if (!msg.empty()) {
        set_error("pack solve: job " + std::to_string(j) + " wave " + std::to_string(r) + ": " + msg);
        return IFDDA_ERR_ARG;
      }
    }
  }
  ++w.tick;
  clk.mark("plan");
  // ---- Green tensors: cached per key; evict unused entries past the limit
  std::vector<JobDev> jd(J);
  {
    size_t need = 0;
    std::map<GreenKey, size_t> fresh;
    for (int j = 0; j < J; ++j)
      if (!w.greens.count(keys[j]) && !fresh.count(keys[j])) {
        const GridPlan& g = plans[j];
        fresh[keys[j]] = (size_t)g.F[0] * g.F[1] * g.F[2] * 6 * (sizeof(double2) + sizeof(float2));
        need += fresh[keys[j]];
      }
    for (int j = 0; j < J; ++j)
      if (w.greens.count(keys[j])) w.greens[keys[j]].used = w.tick;
    while (w.green_bytes + need > w.green_limit) {
      auto victim = w.greens.end();
      for (auto it = w.greens.begin(); it != w.greens.end(); ++it)
        if (it->second.used != w.tick && (victim == w.greens.end() || it->second.used < victim->second.used))
          victim = it;
      if (victim == w.greens.end()) break;
      if (!hip_ok(hipStreamSynchronize(w.s), "green evict")) return IFDDA_ERR_HIP;
      w.green_bytes -= victim->second.bytes;
      green_free(victim->second);
      w.greens.erase(victim);
    }
  }
  // FFT tables of every length the pack uses (uploaded and synchronised)
  for (int j = 0; j < J; ++j)
    for (int a = 0; a < 3; ++a)
      if (!(jd[j].tab3[a] = table_for(w, plans[j].ax[a]))) return IFDDA_ERR_HIP;
  // new keys: built round-robin on the auxiliary streams, each with its own
  // full-grid scratch, after everything already queued on s
  std::vector<int> build;  // first job of each new key
  {
    std::map<GreenKey, int> seen;
    for (int j = 0; j < J; ++j)
      if (!w.greens.count(keys[j]) && !seen.count(keys[j])) seen[keys[j]] = j, build.push_back(j);
  }
  if (!build.empty()) {
    size_t need[PackWs::kAux] = {};
    for (size_t i = 0; i < build.size(); ++i) {
      const GridPlan& g = plans[build[i]];
      need[i % PackWs::kAux] = std::max(need[i % PackWs::kAux],
                                        sizeof(double2) * green_build_comps(g) * (size_t)g.L[0] *
                                            g.L[1] * g.L[2]);
    }
    if (!hip_ok(hipEventRecord(w.ev_main, w.s), "green event")) return IFDDA_ERR_HIP;
    for (int k = 0; k < PackWs::kAux; ++k) {
      if (!need[k]) continue;
      // a growth frees the old scratch: the stream's earlier builds end first
      if (need[k] > w.afull_cap[k] && !hip_ok(hipStreamSynchronize(w.aux[k]), "green scratch"))
        return IFDDA_ERR_HIP;
      if (!grow(&w.afull[k], &w.afull_cap[k], need[k], "hipMalloc green scratch") ||
          !hip_ok(hipStreamWaitEvent(w.aux[k], w.ev_main, 0), "green wait"))
        return IFDDA_ERR_HIP;
    }
    for (size_t i = 0; i < build.size(); ++i) {
      const int j = build[i], k = (int)(i % PackWs::kAux);
      const GridPlan& g = plans[j];
      GreenEntry e;
      e.g = g;
      const size_t nf = (size_t)g.F[0] * g.F[1] * g.F[2] * 6;
      e.bytes = nf * (sizeof(double2) + sizeof(float2));
      if (!hip_ok(hipMalloc(&e.G64, sizeof(double2) * nf + sizeof(float2) * nf), "hipMalloc green")) {
        green_free(e);
        return IFDDA_ERR_HIP;
      }
      e.G32 = reinterpret_cast<float2*>(e.G64 + nf);
      const double2* tw[3] = {jd[j].tab3[0]->tw64, jd[j].tab3[1]->tw64, jd[j].tab3[2]->tw64};
      const int* pos[3] = {jd[j].tab3[0]->pos, jd[j].tab3[1]->pos, jd[j].tab3[2]->pos};
      if (!green_build_device(g, tw, pos, keys[j].d, keys[j].k0, keys[j].nquad,
                              reinterpret_cast<double2*>(w.afull[k]), e.G64, e.G32, &w.gs[k],
                              w.aux[k], green_build_comps(g))) {
        green_free(e);
        return IFDDA_ERR_HIP;
      }
      e.used = w.tick;
      w.green_bytes += e.bytes;
      w.greens.emplace(keys[j], e);
    }
    for (int k = 0; k < PackWs::kAux; ++k)
      if (need[k] && (!hip_ok(hipEventRecord(w.ev_aux[k], w.aux[k]), "green event") ||
                      !hip_ok(hipStreamWaitEvent(w.s, w.ev_aux[k], 0), "green wait")))
        return IFDDA_ERR_HIP;
  }
  for (int j = 0; j < J; ++j) {
    jd[j].green = &w.greens.find(keys[j])->second;
    jd[j].nd = in[j].job->nd;
  }
  clk.mark("green");
  // ---- members and the ragged vector layout (offsets 64-element aligned)
  std::vector<int> mjob, mrhs;
  for (int j = 0; j < J; ++j)
    for (int r = 0; r < in[j].job->nrhs; ++r) mjob.push_back(j), mrhs.push_back(r);
  const int P = (int)mjob.size();
  w.lens.assign(P, 0);
  w.offs.assign(P, 0);
  long long total = 0;
  for (int m = 0; m < P; ++m) {
    w.lens[m] = 3LL * jd[mjob[m]].nd;
    w.offs[m] = total;
    total += (w.lens[m] + 63) & ~63LL;
  }
  // ---- arena: per job map and polarisability, per member FFT scratch, E/FF/FFloc
  std::vector<size_t> o_tab(J), o_al64(J), o_al32(J), o_A(P), o_B(P), o_cell(J), o_wk(J), o_we(J);
  std::vector<size_t> o_rows(J, 0);
  size_t used = 0;
  auto take = [&](size_t bytes) {
    const size_t at = used;
    used = al256(used + bytes);
    return at;
  };
  for (int j = 0; j < J; ++j) {
    const ifdda_job& s = *in[j].job;
    o_tab[j] = dense[j] ? 0 : take(sizeof(int32_t) * (size_t)s.nx * s.ny * s.nz);
    // line masks of a sparse map (MatvecDev::rows), boxes of up to 64 rows
    if (!dense[j] && s.ny <= 64) o_rows[j] = take(sizeof(unsigned long long) * (size_t)s.nz);
    o_al64[j] = take(sizeof(double2) * (size_t)s.nd);
    o_al32[j] = take(sizeof(float2) * (size_t)s.nd);
    if (!s.E_inc) {
      o_cell[j] = dense[j] ? 0 : take(sizeof(int32_t) * (size_t)s.nd);
      o_wk[j] = take(sizeof(double) * 3 * (size_t)s.nrhs);
      o_we[j] = take(sizeof(double2) * 3 * (size_t)s.nrhs);
    }
  }
  const size_t o_off = take(sizeof(long long) * (size_t)P);
  for (int m = 0; m < P; ++m) {
    const GridPlan& g = plans[mjob[m]];
    o_A[m] = take(sizeof(double2) * (size_t)g.nz * g.ny * g.L[0] * 3);
    o_B[m] = take(sizeof(double2) * (size_t)g.nz * g.L[1] * g.L[0] * 3);
  }
  const size_t o_E = take(sizeof(double2) * (size_t)total);
  const size_t o_FF = take(sizeof(double2) * (size_t)total);
  const size_t o_FL = take(sizeof(double2) * (size_t)total);
  if (used > w.arena_cap && !hip_ok(hipStreamSynchronize(w.s), "arena")) return IFDDA_ERR_HIP;
  if (!grow(&w.arena, &w.arena_cap, used, "hipMalloc pack arena")) return IFDDA_ERR_HIP;
  char* A = w.arena;
  // Inputs: a host image of the arena's input part (maps, polarisabilities,
  // wave parameters, dipole cells, member offsets) and of the incident
  // fields, filled in pinned memory by several threads, then two copies.
  const size_t in_bytes = o_off + sizeof(long long) * (size_t)P;
  const size_t e_bytes = sizeof(double2) * (size_t)total;
  const size_t out_bytes = (o_FL - o_FF) + e_bytes;  // the FF and FFloc blocks
  if (std::max(in_bytes + e_bytes, out_bytes) > w.hstage_cap &&
      (!hip_ok(hipStreamSynchronize(w.s), "staging") ||
       !grow_host(&w.hstage, &w.hstage_cap, std::max(in_bytes + e_bytes, out_bytes))))
    return IFDDA_ERR_HIP;
  char* H = w.hstage;
  par_for(J, in_bytes, [&](int j) {
    const ifdda_job& s = *in[j].job;
    const size_t N = (size_t)s.nx * s.ny * s.nz;
    if (!dense[j]) std::memcpy(H + o_tab[j], s.tabdip, sizeof(int32_t) * N);
    if (o_rows[j]) {
      std::vector<unsigned long long> rm;
      if (!row_masks(s.tabdip, s.nx, s.ny, s.nz, rm)) rm.assign((size_t)s.nz, ~0ull);
      std::memcpy(H + o_rows[j], rm.data(), sizeof(unsigned long long) * (size_t)s.nz);
    }
    double2* h64 = reinterpret_cast<double2*>(H + o_al64[j]);
    float2* h32 = reinterpret_cast<float2*>(H + o_al32[j]);
    for (int i = 0; i < s.nd; ++i) {
      h64[i] = s.alpha ? make_double2(s.alpha[2 * i], s.alpha[2 * i + 1])
                       : poladiff_host(make_double2(s.eps[2 * i], s.eps[2 * i + 1]), s.d, s.k0,
                                       s.eps0 > 0.0 ? s.eps0 : 1.0, s.pola);
      h32[i] = make_float2((float)h64[i].x, (float)h64[i].y);
    }
    if (!s.E_inc) {  // plane waves: k and e per RHS, the cell of each dipole
      std::memcpy(H + o_wk[j], wk[j].data(), sizeof(double) * wk[j].size());
      std::memcpy(H + o_we[j], we[j].data(), sizeof(double2) * we[j].size());
      if (!dense[j]) {
        int32_t* cell = reinterpret_cast<int32_t*>(H + o_cell[j]);
        for (size_t i = 0; i < N; ++i)
          if (s.tabdip[i] > 0) cell[s.tabdip[i] - 1] = (int32_t)i;
      }
    }
  });
  std::memcpy(H + o_off, w.offs.data(), sizeof(long long) * P);
  par_for(P, e_bytes, [&](int m) {
    const ifdda_job& s = *in[mjob[m]].job;
    if (s.E_inc)
      std::memcpy(H + in_bytes + sizeof(double2) * w.offs[m],
                  s.E_inc + 2 * (size_t)w.lens[m] * mrhs[m], sizeof(double2) * w.lens[m]);
  });
  if (!hip_ok(hipMemcpyAsync(A, H, in_bytes, hipMemcpyHostToDevice, w.s), "copy inputs") ||
      !hip_ok(hipMemcpyAsync(A + o_E, H + in_bytes, e_bytes, hipMemcpyHostToDevice, w.s),
              "copy E_inc"))
    return IFDDA_ERR_HIP;
  for (int j = 0; j < J; ++j) {
    jd[j].tab = dense[j] ? nullptr : reinterpret_cast<int32_t*>(A + o_tab[j]);
    jd[j].rows = o_rows[j] ? reinterpret_cast<unsigned long long*>(A + o_rows[j]) : nullptr;
    jd[j].al64 = reinterpret_cast<double2*>(A + o_al64[j]);
    jd[j].al32 = reinterpret_cast<float2*>(A + o_al32[j]);
  }
  double2* E = reinterpret_cast<double2*>(A + o_E);
  double2* FF = reinterpret_cast<double2*>(A + o_FF);
  double2* FL = reinterpret_cast<double2*>(A + o_FL);
  // waves formed on the device at each job's dipoles
  const long long* doff = reinterpret_cast<const long long*>(A + o_off);
  for (int j = 0, m0 = 0; j < J; m0 += in[j].job->nrhs, ++j) {
    const ifdda_job& s = *in[j].job;
    if (s.E_inc) continue;
    const int32_t* dcell = dense[j] ? nullptr : reinterpret_cast<const int32_t*>(A + o_cell[j]);
    if (!launch_waves(E, doff + m0, s.nrhs, s.nd, dcell, s.nx, s.ny, s.origin, s.d,
                      reinterpret_cast<const double*>(A + o_wk[j]),
                      reinterpret_cast<const double2*>(A + o_we[j]), w.s))
      return IFDDA_ERR_HIP;
  }
  clk.mark("inputs");
  // ---- packed-launch argument blocks
  PackOp op;
  {
    std::vector<const GridPlan*> gp(P);
    std::vector<MatvecDev<double2>> m64(P);
    std::vector<MatvecDev<float2>> m32(P);
    op.al.resize(P);
    op.al32.resize(P);
    op.nd.resize(P);
    op.off = w.offs;
    for (int m = 0; m < P; ++m) {
      const JobDev& d = jd[mjob[m]];
      gp[m] = &plans[mjob[m]];
      MatvecDev<double2>& a = m64[m];
      MatvecDev<float2>& b = m32[m];
      a.A = reinterpret_cast<double2*>(A + o_A[m]);
      a.B = reinterpret_cast<double2*>(A + o_B[m]);
      b.A = reinterpret_cast<float2*>(A + o_A[m]);  // the FP64-sized scratch holds either
      b.B = reinterpret_cast<float2*>(A + o_B[m]);
      a.G = d.green->G64;
      b.G = d.green->G32;
      for (int ax = 0; ax < 3; ++ax) {
        a.tw[ax] = d.tab3[ax]->tw64, a.tws[ax] = d.tab3[ax]->tws64, a.pos[ax] = d.tab3[ax]->pos;
        b.tw[ax] = d.tab3[ax]->tw32, b.tws[ax] = d.tab3[ax]->tws32, b.pos[ax] = d.tab3[ax]->pos;
      }
      a.tab = b.tab = d.tab;
      a.rows = b.rows = d.rows;
      a.alpha = d.al64;
      b.alpha = d.al32;
      a.alpha_mode = b.alpha_mode = kAlphaScalar;
      op.al[m] = d.al64;
      op.al32[m] = d.al32;
      op.nd[m] = d.nd;
    }
    std::vector<unsigned char> b64, b32, t64, t32;
    const size_t n64 = pack_blob_f64(gp, m64, w.offs.data(), op.lay64, b64);
    const size_t n32 = pack_blob_f32(gp, m32, w.offs.data(), op.lay32, b32);
    size_t nt64 = 0, nt32 = 0;
    if (o->method == IFDDA_QMRCLA) {
      for (int m = 0; m < P; ++m) m64[m].alpha_mode = m32[m].alpha_mode = kAlphaNone;
      PackLayout l64, l32;
      nt64 = pack_blob_f64(gp, m64, w.offs.data(), l64, t64);
      nt32 = pack_blob_f32(gp, m32, w.offs.data(), l32, t32);
    }
    const size_t need = al256(n64) + al256(n32) + al256(nt64) + al256(nt32);
    if (need > w.blob_cap && !hip_ok(hipStreamSynchronize(w.s), "blob")) return IFDDA_ERR_HIP;
    if (!grow(&w.blob, &w.blob_cap, need, "hipMalloc pack blob") ||
        !hip_ok(hipMemcpyAsync(w.blob, b64.data(), n64, hipMemcpyHostToDevice, w.s), "copy blob") ||
        !hip_ok(hipMemcpyAsync(w.blob + al256(n64), b32.data(), n32, hipMemcpyHostToDevice, w.s), "copy blob") ||
        (nt64 && !hip_ok(hipMemcpyAsync(w.blob + al256(n64) + al256(n32), t64.data(), nt64,
                                        hipMemcpyHostToDevice, w.s), "copy blob")) ||
        (nt32 && !hip_ok(hipMemcpyAsync(w.blob + al256(n64) + al256(n32) + al256(nt64), t32.data(),
                                        nt32, hipMemcpyHostToDevice, w.s), "copy blob")) ||
        !hip_ok(hipStreamSynchronize(w.s), "pack setup"))  // host staging ends here
      return IFDDA_ERR_HIP;
    op.blob64 = w.blob;
    op.blob32 = w.blob + al256(n64);
    if (nt64) {
      op.blobt64 = w.blob + al256(n64) + al256(n32);
      op.blobt32 = w.blob + al256(n64) + al256(n32) + al256(nt64);
    }
  }
  // ---- Krylov workspace (reused when large enough) and the solve
  const bool need32 = o->precision == 32;
  if (!solver_refit(w.S, P, need32, w.lens.data(), w.offs.data(), w.s)) {
    if (!hip_ok(hipStreamSynchronize(w.s), "solver")) return IFDDA_ERR_HIP;
    // room for twice the elements and right-hand sides of the last workspace
    long long cap_total = 0;
    int cap_R = 0;
    solver_capacity(w.S, &cap_total, &cap_R);
    solver_delete(w.S);
    w.S = solver_new(P, need32, w.lens.data(), w.offs.data(), 2 * cap_total, 2 * cap_R);
    if (!w.S) return IFDDA_ERR_HIP;
    if (debug_on()) std::fprintf(stderr, "ifdda pack: solver for %d RHS\n", P);
  }
  std::vector<ifdda_solve_info> info(P);
  clk.mark("blob");
  const int rc = solve_core(op, w.S, o, P, E, FF, FL, info.data(), nstop, w.s, t0);
  if (rc != IFDDA_OK) return rc;
  clk.mark("solve");
  // ---- outputs: one copy of the FF and FFloc blocks into the pinned image,
  // then each member's slices to its job's arrays on several threads
  (void)FF;
  (void)FL;
  if (!hip_ok(hipMemcpyAsync(H, A + o_FF, out_bytes, hipMemcpyDeviceToHost, w.s), "copy outputs") ||
      !hip_ok(hipStreamSynchronize(w.s), "pack outputs"))
    return IFDDA_ERR_HIP;
  par_for(P, out_bytes, [&](int m) {
    const PackIn& p = in[mjob[m]];
    const size_t at = 2 * (size_t)w.lens[m] * mrhs[m];
    const size_t bytes = sizeof(double2) * w.lens[m];
    if (p.FF) std::memcpy(p.FF + at, H + sizeof(double2) * w.offs[m], bytes);
    if (p.FFloc) std::memcpy(p.FFloc + at, H + (o_FL - o_FF) + sizeof(double2) * w.offs[m], bytes);
  });
  for (int m = 0; m < P; ++m)
    if (in[mjob[m]].infos) in[mjob[m]].infos[mrhs[m]] = info[m];
  clk.mark("outputs");
  if (debug_on()) std::fprintf(stderr, "ifdda pack: %d jobs %d RHS:%s\n", J, P, clk.line.c_str());
  return IFDDA_OK;
}

}  // namespace ifdda

using namespace ifdda;

namespace {
constexpr int kBatchDevices = 64;
std::mutex g_batch_mu[kBatchDevices];
PackWs* g_batch_ws[kBatchDevices] = {};
}  // namespace

extern "C" int ifdda_solve_batch(int njobs, const ifdda_job* jobs, double* const* FFloc,
                                 double* const* FF, ifdda_solve_info* const* infos,
                                 volatile int* nstop) {
  if (njobs < 0 || (njobs > 0 && (!jobs || !infos))) {
    set_error("ifdda_solve_batch: NULL argument");
    return IFDDA_ERR_ARG;
  }
  std::vector<PackIn> in(njobs);
  for (int j = 0; j < njobs; ++j) {
    const ifdda_solve_opts& a = jobs[j].opts;
    const ifdda_solve_opts& b = jobs[0].opts;
    if (a.method != b.method || a.precision != b.precision || a.x0_zero != b.x0_zero ||
        a.maxit != b.maxit || a.tol != b.tol || a.check_every != b.check_every) {
      set_error("ifdda_solve_batch: all jobs need the same solve options");
      return IFDDA_ERR_ARG;
    }
    if (!infos[j]) {
      set_error("ifdda_solve_batch: NULL infos");
      return IFDDA_ERR_ARG;
    }
    in[j] = PackIn{&jobs[j], FFloc ? FFloc[j] : nullptr, FF ? FF[j] : nullptr, infos[j]};
  }
  if (njobs == 0) return IFDDA_OK;
  int dev = 0;
  if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return IFDDA_ERR_HIP;
  if (dev < 0 || dev >= kBatchDevices) {
    set_error("ifdda_solve_batch: device index out of range");
    return IFDDA_ERR_ARG;
  }
  // one workspace per device, kept between calls (FFT tables, recent Green
  // tensors, grow-only buffers), one batch at a time on it
  std::lock_guard<std::mutex> lk(g_batch_mu[dev]);
  if (!g_batch_ws[dev] && !(g_batch_ws[dev] = pack_ws_new())) return IFDDA_ERR_HIP;
  return pack_solve(*g_batch_ws[dev], in, &jobs[0].opts, nstop);
}

extern "C" int ifdda_solve_batch_release(void) {
  for (int dev = 0; dev < kBatchDevices; ++dev) {
    std::lock_guard<std::mutex> lk(g_batch_mu[dev]);
    if (!g_batch_ws[dev]) continue;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(dev);
    pack_ws_delete(g_batch_ws[dev]);
    g_batch_ws[dev] = nullptr;
    (void)hipSetDevice(cur);
  }
  return IFDDA_OK;
}
