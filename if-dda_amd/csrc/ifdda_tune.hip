// ifdda_tune.hip -- run-time tile-width autotuner for the five matvec passes.
//
// The reference plans its 3D FFTs once per run with FFTW_ESTIMATE
// (cdmlib.f:1153-1158): a fixed plan chosen without timing.  Here the
// transforms are the LDS-tiled passes P1..P5, whose one free parameter per
// pass and precision is the tile width B (box lines per workgroup).  B changes
// only how lines are grouped into workgroups, never the arithmetic of a line:
// with the generic kernels every width gives the same results bit for bit
// (a z-pass width without a compile-time kernel runs the generic one, which
// agrees with it to rounding).  B changes occupancy and the overlap of load,
// transform and store phases, which this tuner measures on the device at
// hand.
//
// The tuned widths are those of launches of one right-hand side (single
// solves and matvecs).  Batched launches and packs keep the widths of the LDS
// budgets: their right-hand sides already fill the device.
//
// ifdda_autotune times every pass of a context's grid at each candidate width
// (the widths that the LDS budgets 6..64 KB give, plus the current one) with
// HIP events around the pass launches, keeps the fastest width per pass, and
// records the choice in a process-wide table keyed by box size and precision.
// Contexts created afterwards for the same box (including those of the
// reference-shaped entries) start from the table; the table can be read and
// written through ifdda_tune_get / ifdda_tune_set, so a caller can keep it
// between runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ifdda_internal.h"

namespace ifdda {

namespace {
std::mutex g_tune_mu;
std::map<std::array<int, 4>, std::array<int, 5>> g_tuned;  // (nx, ny, nz, prec) -> B of P1..P5

std::array<int, 4> tune_key(int nx, int ny, int nz, int prec) { return {nx, ny, nz, prec}; }
}  // namespace

bool tuned_widths(int nx, int ny, int nz, int prec, int B[5]) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find(tune_key(nx, ny, nz, prec));
  if (it == g_tuned.end()) return false;
  for (int p = 0; p < 5; ++p) B[p] = it->second[p];
  return true;
}

static void store_widths(int nx, int ny, int nz, int prec, const int B[5]) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  std::array<int, 5> v;
  for (int p = 0; p < 5; ++p) v[p] = B[p];
  g_tuned[tune_key(nx, ny, nz, prec)] = v;
}

// The tile shape of pass p (0..4) of precision index pi (0 = FP64) in
// launches of one right-hand side, the ones the tuner times.
static TileShape* pass_tile(GridPlan& g, int pi, int p) { return &g.one[pi][p]; }
static const TileShape& pass_tile(const GridPlan& g, int pi, int p) {
  return *pass_tile(const_cast<GridPlan&>(g), pi, p);
}

// Candidate shapes of pass p: one per distinct width that an LDS budget of
// 6..64 KB (capped at the device limit) gives, and the current width.
static std::vector<TileShape> candidates(const GridPlan& g, int pi, int p, size_t lds_limit) {
  static const size_t budgets[] = {6144, 8192, 12288, 16384, 20480, 26624,
                                   33792, 40960, 53248, 65536};
  std::vector<TileShape> out;
  for (size_t b : budgets) {
    if (b > lds_limit) break;
    const TileShape t = pass_tile_shape(g, pi, p, 0, b);
    if (t.B == 0) continue;
    if (std::none_of(out.begin(), out.end(), [&](const TileShape& o) { return o.B == t.B; }))
      out.push_back(t);
  }
  const TileShape cur = pass_tile(g, pi, p);
  if (std::none_of(out.begin(), out.end(), [&](const TileShape& o) { return o.B == cur.B; }))
    out.push_back(cur);
  std::sort(out.begin(), out.end(), [](const TileShape& a, const TileShape& b) { return a.B < b.B; });
  return out;
}

template <typename V>
static int tune_one(ifdda_ctx* c, int pi, int reps, size_t lds_limit, int chosen[5]) {
  GridPlan& g = c->g;
  const long long n = 3 * c->nd;
  std::vector<TileShape> cand[5];
  size_t ncand = 0;
  for (int p = 0; p < 5; ++p) {
    cand[p] = candidates(g, pi, p, lds_limit);
    ncand = std::max(ncand, cand[p].size());
  }
  V *x = nullptr, *y = nullptr;
  PassTimer tm;
  tm.cap = 10 * reps;  // two matvecs per repetition when P5 is timed in its fused forms
  std::vector<hipEvent_t> ev(2 * (size_t)tm.cap, nullptr);
  std::vector<int> tag(tm.cap, 0);
  auto cleanup = [&]() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    (void)hipFree(x);
    (void)hipFree(y);
  };
  for (auto& e : ev)
    if (!hip_ok(hipEventCreate(&e), "hipEventCreate")) return cleanup(), IFDDA_ERR_HIP;
  tm.ev = ev.data();
  tm.tag = tag.data();
  if (!hip_ok(hipMalloc(&x, sizeof(V) * n), "hipMalloc tune x") ||
      !hip_ok(hipMalloc(&y, sizeof(V) * n), "hipMalloc tune y") ||
      // any finite, normal input: the passes' time does not depend on it
      !hip_ok(hipMemsetAsync(x, 0x3c, sizeof(V) * n, c->stream), "hipMemsetAsync tune x") ||
      !hip_ok(hipMemsetAsync(y, 0, sizeof(V) * n, c->stream), "hipMemsetAsync tune y"))
    return cleanup(), IFDDA_ERR_HIP;

  MatvecDev<V> m;
  if constexpr (sizeof(V) == 16) m = c->mv64(); else m = c->mv32();
  m.timer = nullptr;
  // Solves of this precision with GPBICG products fused into P5
  // (ifdda_fuse_krylov) run P5 as its fused forms, one gathering <r0*, Ap>,
  // the other the step-4 products (and the look-ahead's three): time those,
  // one after the other in every repetition, in place of the plain gather.
  const int pbit = sizeof(V) == 16 ? 0 : 1;
  const bool fz = c->specialized && ((c->fuse_mask >> pbit) & 1);
  const bool la = fz && ((c->fuse_mask >> (pbit == 0 ? 4 : 3)) & 1);
  const long long np5 = (long long)g.ny * g.nz;  // P5 blocks at width 1, an upper bound
  double2* part = nullptr;
  KFuse<V> f1, f2;
  if (fz) {
    if (!hip_ok(hipMalloc(&part, sizeof(double2) * 8 * np5), "hipMalloc tune partials"))
      return cleanup(), IFDDA_ERR_HIP;
    f1.p5 = 1;
    f1.r0s = x;
    f2.p5 = la ? 3 : 2;
    f2.y = x;
    f2.r0s = x;
    f1.part = f2.part = part;
    f1.pstride = f2.pstride = (int)np5;
  }
  const TileShape keep[5] = {*pass_tile(g, pi, 0), *pass_tile(g, pi, 1), *pass_tile(g, pi, 2),
                             *pass_tile(g, pi, 3), *pass_tile(g, pi, 4)};
  std::vector<double> best(5, 1e300);
  for (int p = 0; p < 5; ++p) chosen[p] = keep[p].B;
  auto run = [&](PassTimer* t) {
    if (fz) {
      m.timer = t;
      bool ok;
      if constexpr (sizeof(V) == 16)
        ok = launch_matvec_fused_f64(g, m, x, x, y, c->stream, nullptr, f1) &&
             launch_matvec_fused_f64(g, m, x, x, y, c->stream, nullptr, f2);
      else
        ok = launch_matvec_fused_f32(g, m, x, x, y, c->stream, nullptr, f1) &&
             launch_matvec_fused_f32(g, m, x, x, y, c->stream, nullptr, f2);
      m.timer = nullptr;
      return ok;
    }
    if constexpr (sizeof(V) == 16)
      return launch_matvec_f64(g, m, x, x, y, c->stream, nullptr, t);
    else
      return launch_matvec_f32(g, m, x, x, y, c->stream, nullptr, t);
  };
  int rc = IFDDA_OK;
  for (size_t k = 0; k < ncand && rc == IFDDA_OK; ++k) {
    for (int p = 0; p < 5; ++p) *pass_tile(g, pi, p) = cand[p][std::min(k, cand[p].size() - 1)];
    tm.used = 0;
    tm.seen = 0;
    if (!run(nullptr)) {  // warm-up: first launch of each kernel shape
      rc = IFDDA_ERR_HIP;
      break;
    }
    for (int r = 0; r < reps && rc == IFDDA_OK; ++r)
      if (!run(&tm)) rc = IFDDA_ERR_HIP;
    if (rc != IFDDA_OK || !hip_ok(hipStreamSynchronize(c->stream), "tune sync")) {
      rc = IFDDA_ERR_HIP;
      break;
    }
    std::vector<float> ms[5];
    for (int i = 0; i < tm.used; ++i) {
      float t = 0.f;
      if (!hip_ok(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]), "hipEventElapsedTime")) {
        rc = IFDDA_ERR_HIP;
        break;
      }
      ms[tm.tag[i] % 5].push_back(t);
    }
    for (int p = 0; p < 5 && rc == IFDDA_OK; ++p) {
      if (k >= cand[p].size() || ms[p].empty()) continue;  // this pass's list is exhausted
      std::sort(ms[p].begin(), ms[p].end());
      const double med = ms[p][ms[p].size() / 2];
      if (med < best[p]) {
        best[p] = med;
        chosen[p] = cand[p][k].B;
      }
    }
  }
  for (int p = 0; p < 5; ++p) {
    *pass_tile(g, pi, p) = keep[p];
    if (rc != IFDDA_OK) continue;
    for (const TileShape& t : cand[p])
      if (t.B == chosen[p]) *pass_tile(g, pi, p) = t;
  }
  if (part) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(part);
  cleanup();
  if (rc == IFDDA_OK) store_widths(g.nx, g.ny, g.nz, pi == 0 ? 64 : 32, chosen);
  return rc;
}

}  // namespace ifdda

using namespace ifdda;

extern "C" int ifdda_autotune(ifdda_ctx* c, int prec, int reps, int widths[10]) {
  if (!c) return set_error("ifdda_autotune: NULL context"), IFDDA_ERR_ARG;
  if (prec != 0 && prec != 32 && prec != 64)
    return set_error("ifdda_autotune: prec must be 0 (both), 32 or 64"), IFDDA_ERR_ARG;
  if (reps < 0 || reps > 100)
    return set_error("ifdda_autotune: reps must be 0 (default 5) .. 100"), IFDDA_ERR_ARG;
  if (!c->green_ready) return set_error("ifdda_autotune: no Green tensor"), IFDDA_ERR_STATE;
  if (!c->have_dipoles) return set_error("ifdda_autotune: no dipoles"), IFDDA_ERR_STATE;
  if (reps == 0) reps = 5;
  int B[2][5];
  for (int pi = 0; pi < 2; ++pi)
    for (int p = 0; p < 5; ++p) B[pi][p] = pass_tile(c->g, pi, p)->B;
  if (c->nd > 0) {
    if (!hip_ok(hipStreamSynchronize(c->stream), "ifdda_autotune: stream")) return IFDDA_ERR_HIP;
    int rc = IFDDA_OK;
    if (prec != 32) rc = tune_one<double2>(c, 0, reps, c->lds_limit, B[0]);
    if (rc == IFDDA_OK && prec != 64) rc = tune_one<float2>(c, 1, reps, c->lds_limit, B[1]);
    if (rc != IFDDA_OK) return rc;
  }
  if (widths)
    for (int pi = 0; pi < 2; ++pi)
      for (int p = 0; p < 5; ++p) widths[5 * pi + p] = B[pi][p];
  return IFDDA_OK;
}

extern "C" int ifdda_tile_widths(ifdda_ctx* c, int widths[10]) {
  if (!c || !widths) return set_error("ifdda_tile_widths: NULL argument"), IFDDA_ERR_ARG;
  for (int pi = 0; pi < 2; ++pi)
    for (int p = 0; p < 5; ++p) widths[5 * pi + p] = pass_tile(c->g, pi, p)->B;
  return IFDDA_OK;
}

extern "C" int ifdda_tune_get(int nx, int ny, int nz, int prec, int widths[5]) {
  if (!widths || (prec != 32 && prec != 64))
    return set_error("ifdda_tune_get: bad argument"), IFDDA_ERR_ARG;
  return tuned_widths(nx, ny, nz, prec, widths) ? 1 : 0;
}

extern "C" int ifdda_tune_set(int nx, int ny, int nz, int prec, const int widths[5]) {
  if (!widths || (prec != 32 && prec != 64) || nx < 1 || ny < 1 || nz < 1)
    return set_error("ifdda_tune_set: bad argument"), IFDDA_ERR_ARG;
  // every width must give a feasible tile of this box at the device's LDS limit
  GridPlan g{};
  std::string why;
  if (!make_grid_plan(nx, ny, nz, 65536, &g, &why))
    return set_error("ifdda_tune_set: " + why), IFDDA_ERR_ARG;
  const int pi = prec == 64 ? 0 : 1;
  for (int p = 0; p < 5; ++p)
    if (pass_tile_shape(g, pi, p, widths[p], 65536).B != widths[p])
      return set_error("ifdda_tune_set: width " + std::to_string(widths[p]) + " of pass P" +
                       std::to_string(p + 1) + " does not fit this box"),
             IFDDA_ERR_ARG;
  store_widths(nx, ny, nz, prec, widths);
  return IFDDA_OK;
}
