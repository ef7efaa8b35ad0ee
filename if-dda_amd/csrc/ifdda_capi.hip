// ifdda_capi.hip -- the C ABI of include/ifdda.h: context set-up, the device
// Green tensor, the dipole map and polarisability, the matvec entry points,
// and the reference-shaped entries that Fortran callers reach through
// if-dda_amd/fortran/ifdda_dropin.f90.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ifdda.h"
#include "ifdda_internal.h"

namespace ifdda {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

}  // namespace ifdda

using namespace ifdda;

#define CHECK_HIP(call, what)                 \
  do {                                        \
    if (!hip_ok((call), what)) return IFDDA_ERR_HIP; \
  } while (0)

extern "C" const char* ifdda_last_error(void) { return g_err.c_str(); }

static int arg_error(const std::string& m) {
  set_error(m);
  return IFDDA_ERR_ARG;
}

static int state_error(const std::string& m) {
  set_error(m);
  return IFDDA_ERR_STATE;
}

extern "C" void ifdda_destroy(ifdda_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  solver_free(c);
  for (int i = 0; i < 3; ++i) {
    (void)hipFree(c->tw64[i]);
    (void)hipFree(c->tw32[i]);
    (void)hipFree(c->tws64[i]);
    (void)hipFree(c->tws32[i]);
    (void)hipFree(c->pos[i]);
  }
  (void)hipFree(c->G64);
  (void)hipFree(c->G32);
  (void)hipFree(c->Gd64);
  (void)hipFree(c->tab);
  (void)hipFree(c->rows);
  (void)hipFree(c->al64);
  (void)hipFree(c->al32);
  (void)hipFree(c->scrA);
  (void)hipFree(c->scrB);
  (void)hipFree(c->sstage);
  (void)hipFree(c->hx);
  (void)hipFree(c->hb);
  (void)hipFree(c->ffbuf);
  (void)hipFree(c->dipcell);
  (void)hipFree(c->mbuf);
  (void)hipFree(c->gfull);
  (void)hipFree(c->xbuf);
  green_scratch_free(c->gscr);
  ifdda_timing_enable(c, 0);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" int ifdda_create(int nx, int ny, int nz, ifdda_ctx** out) {
  if (!out) return arg_error("ifdda_create: out is NULL");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1) return arg_error("ifdda_create: box sizes must be >= 1");
  if ((long long)nx * ny * nz > (1LL << 31) / 8) return arg_error("ifdda_create: box too large");
  ifdda_ctx* c = new ifdda_ctx();
  GridPlan& g = c->g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    g.L[a] = smooth235_at_least(2 * n[a]);
    g.F[a] = g.L[a] / 2 + 1;
    if (!make_axis_plan(g.L[a], &g.ax[a])) {
      delete c;
      return arg_error("ifdda_create: no radix plan for L=" + std::to_string(g.L[a]));
    }
  }
  if (!hip_ok(hipGetDevice(&c->device), "hipGetDevice")) {
    delete c;
    return IFDDA_ERR_HIP;
  }
  int lds = 65536;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
      hipSuccess)
    lds = 65536;
  c->lds_limit = (size_t)std::min(lds, 65536);
  if (!make_tiles(g, c->lds_limit)) {
    delete c;
    return arg_error("ifdda_create: padded FFT length above the supported maximum");
  }

  auto fail = [&](int code) {
    ifdda_destroy(c);
    return code;
  };
  if (!hip_ok(hipStreamCreate(&c->stream), "hipStreamCreate")) return fail(IFDDA_ERR_HIP);
  for (int a = 0; a < 3; ++a) {
    const int L = g.L[a];
    std::vector<double2> t64(L), s64(L);
    std::vector<float2> t32(L), s32(L);
    std::vector<int> pos(L);
    axis_tables(g.ax[a], t64.data(), t32.data(), s64.data(), s32.data(), pos.data());
    if (!hip_ok(hipMalloc(&c->pos[a], sizeof(int) * L), "hipMalloc positions") ||
        !hip_ok(hipMemcpy(c->pos[a], pos.data(), sizeof(int) * L, hipMemcpyHostToDevice),
                "copy positions") ||
        !hip_ok(hipMalloc(&c->tw64[a], sizeof(double2) * L), "hipMalloc tw64") ||
        !hip_ok(hipMalloc(&c->tw32[a], sizeof(float2) * L), "hipMalloc tw32") ||
        !hip_ok(hipMemcpy(c->tw64[a], t64.data(), sizeof(double2) * L, hipMemcpyHostToDevice),
                "copy tw64") ||
        !hip_ok(hipMemcpy(c->tw32[a], t32.data(), sizeof(float2) * L, hipMemcpyHostToDevice),
                "copy tw32") ||
        !hip_ok(hipMalloc(&c->tws64[a], sizeof(double2) * L), "hipMalloc tws64") ||
        !hip_ok(hipMalloc(&c->tws32[a], sizeof(float2) * L), "hipMalloc tws32") ||
        !hip_ok(hipMemcpy(c->tws64[a], s64.data(), sizeof(double2) * L, hipMemcpyHostToDevice),
                "copy tws64") ||
        !hip_ok(hipMemcpy(c->tws32[a], s32.data(), sizeof(float2) * L, hipMemcpyHostToDevice),
                "copy tws32"))
      return fail(IFDDA_ERR_HIP);
  }
  const size_t nf = (size_t)g.F[0] * g.F[1] * g.F[2] * 6;
  const size_t nA = (size_t)nz * ny * g.L[0] * 3;
  const size_t nB = (size_t)nz * g.L[1] * g.L[0] * 3;
  if (!hip_ok(hipMalloc(&c->G64, sizeof(double2) * nf), "hipMalloc G64") ||
      !hip_ok(hipMalloc(&c->G32, sizeof(float2) * nf), "hipMalloc G32") ||
      !hip_ok(hipMalloc(&c->scrA, sizeof(double2) * nA), "hipMalloc scratch A") ||
      !hip_ok(hipMalloc(&c->scrB, sizeof(double2) * nB), "hipMalloc scratch B"))
    return fail(IFDDA_ERR_HIP);
  *out = c;
  return IFDDA_OK;
}

int ifdda::ensure_scratch(ifdda_ctx* c, int R) {
  if (R <= c->scr_rhs) return IFDDA_OK;
  const GridPlan& g = c->g;
  const size_t nA = (size_t)g.nz * g.ny * g.L[0] * 3 * R;
  const size_t nB = (size_t)g.nz * g.L[1] * g.L[0] * 3 * R;
  CHECK_HIP(hipStreamSynchronize(c->stream), "scratch sync");
  (void)hipFree(c->scrA);
  (void)hipFree(c->scrB);
  c->scrA = c->scrB = nullptr;
  c->scr_rhs = 0;
  CHECK_HIP(hipMalloc(&c->scrA, sizeof(double2) * nA), "hipMalloc scratch A");
  CHECK_HIP(hipMalloc(&c->scrB, sizeof(double2) * nB), "hipMalloc scratch B");
  c->scr_rhs = R;
  return IFDDA_OK;
}

extern "C" int ifdda_fft_lengths(const ifdda_ctx* c, int L[3]) {
  if (!c || !L) return arg_error("ifdda_fft_lengths: NULL argument");
  for (int a = 0; a < 3; ++a) L[a] = c->g.L[a];
  return IFDDA_OK;
}

// The context's full-grid Green scratch, grown to M complex values (the
// stream drains before a smaller buffer is freed).
double2* ifdda::green_full(ifdda_ctx* c, size_t M) {
  const size_t bytes = sizeof(double2) * M;
  if (bytes > c->gfull_cap) {
    if (c->gfull && !hip_ok(hipStreamSynchronize(c->stream), "green scratch")) return nullptr;
    (void)hipFree(c->gfull);
    c->gfull = nullptr;
    c->gfull_cap = 0;
    if (!hip_ok(hipMalloc(&c->gfull, bytes), "hipMalloc green scratch")) return nullptr;
    c->gfull_cap = bytes;
  }
  return c->gfull;
}

extern "C" int ifdda_green_build(ifdda_ctx* c, double aretecube, double k0, int nquad,
                                 double tolinit) {
  (void)tolinit;  // RELREQ is unused by the reference quadrature (propaesplibreintsim.f:2-107)
  if (!c) return arg_error("ifdda_green_build: NULL context");
  if (!(aretecube > 0.0) || !(k0 > 0.0)) return arg_error("ifdda_green_build: need d > 0, k0 > 0");
  if (nquad < 0 || nquad > 10) return arg_error("ifdda_green_build: nquad must be in [0,10]");
  const size_t M = (size_t)c->g.L[0] * c->g.L[1] * c->g.L[2];
  const int nc = green_build_comps(c->g);
  double2* full = green_full(c, nc * M);
  if (!full) return IFDDA_ERR_HIP;
  const double2* tw[3] = {c->tw64[0], c->tw64[1], c->tw64[2]};
  const int* pos[3] = {c->pos[0], c->pos[1], c->pos[2]};
  c->green_ready = false;
  c->gd_ready = false;
  bool ok = green_build_device(c->g, tw, pos, aretecube, k0, nquad, full, c->G64, c->G32,
                               &c->gscr, c->stream, nc);
  ok = ok && hip_ok(hipStreamSynchronize(c->stream), "green build");
  if (!ok) return IFDDA_ERR_HIP;
  c->green_ready = true;
  for (auto& h : c->green_host) h = nullptr;
  return IFDDA_OK;
}

// Sample positions used to check the even/odd symmetry of a caller tensor
// and to fingerprint the caller's arrays.
static size_t sample_index(const GridPlan& g, int s) {
  const size_t M = (size_t)g.L[0] * g.L[1] * g.L[2];
  return (size_t)(((unsigned long long)(s + 1) * 2654435761ULL) % M);
}

extern "C" int ifdda_green_set(ifdda_ctx* c, const double* const G[6]) {
  if (!c || !G) return arg_error("ifdda_green_set: NULL argument");
  const GridPlan& g = c->g;
  if (g.L[0] != 2 * g.nx || g.L[1] != 2 * g.ny || g.L[2] != 2 * g.nz)
    return arg_error("ifdda_green_set: the reference layout needs L = 2n on every axis, but 2n "
                     "is not 2,3,5-smooth here; build the tensor with ifdda_green_build");
  const int L1 = g.L[0], L2 = g.L[1], L3 = g.L[2];
  const size_t M = (size_t)L1 * L2 * L3;
  // Parity per component under k -> L - k on each axis: xy odd in x,y; xz
  // odd in x,z; yz odd in y,z; the diagonal even (calculgreenfft.f:29-65).
  const int par[6][3] = {{1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}, {1, -1, -1}, {1, 1, 1}};
  for (int comp = 0; comp < 6; ++comp) {
    const double* a = G[comp];
    if (!a) return arg_error("ifdda_green_set: NULL component");
    double amax = 0.0;
    for (int s = 0; s < 64; ++s) {
      const size_t idx = sample_index(g, s);
      amax = std::max(amax, std::hypot(a[2 * idx], a[2 * idx + 1]));
    }
    for (int s = 0; s < 64; ++s) {
      const size_t idx = sample_index(g, s);
      const int i = idx % L1, j = (idx / L1) % L2, k = idx / ((size_t)L1 * L2);
      for (int ax = 0; ax < 3; ++ax) {  // mirror one axis at a time
        const int mi = ax == 0 ? (L1 - i) % L1 : i;
        const int mj = ax == 1 ? (L2 - j) % L2 : j;
        const int mk = ax == 2 ? (L3 - k) % L3 : k;
        const size_t m = ((size_t)mk * L2 + mj) * L1 + mi;
        const double sg = par[comp][ax];
        const double dr = a[2 * idx] - sg * a[2 * m], di = a[2 * idx + 1] - sg * a[2 * m + 1];
        if (std::hypot(dr, di) > 1e-9 * amax + 1e-300)
          return arg_error("ifdda_green_set: the tensor lacks the even/odd symmetry of a Green "
                           "tensor (component " + std::to_string(comp) + ")");
      }
    }
  }
  double2* full = green_full(c, M);
  if (!full) return IFDDA_ERR_HIP;
  bool ok = true;
  for (int comp = 0; ok && comp < 6; ++comp) {
    ok = hip_ok(hipMemcpyAsync(full, G[comp], sizeof(double2) * M, hipMemcpyHostToDevice,
                               c->stream),
                "copy green component") &&
         green_fold_component(g, full, comp, c->G64, c->G32, c->stream) &&
         hip_ok(hipStreamSynchronize(c->stream), "green fold");
  }
  c->gd_ready = false;
  if (!ok) return IFDDA_ERR_HIP;
  c->green_ready = true;
  for (auto& h : c->green_host) h = nullptr;
  return IFDDA_OK;
}

extern "C" int ifdda_green_get(ifdda_ctx* c, double* const G[6]) {
  if (!c || !G) return arg_error("ifdda_green_get: NULL argument");
  if (!c->green_ready) return state_error("ifdda_green_get: no Green tensor built");
  const GridPlan& g = c->g;
  const int L1 = g.L[0], L2 = g.L[1], L3 = g.L[2], F1 = g.F[0], F3 = g.F[2];
  const size_t nf = (size_t)g.F[0] * g.F[1] * g.F[2] * 6;
  std::vector<double2> h(nf);
  CHECK_HIP(hipMemcpy(h.data(), c->G64, sizeof(double2) * nf, hipMemcpyDeviceToHost),
            "copy green");
  auto fold = [](int k, int L, int& s) {
    if (2 * k > L) { s = -1; return L - k; }
    s = 1;
    return k;
  };
  for (int k = 0; k < L3; ++k)
    for (int j = 0; j < L2; ++j)
      for (int i = 0; i < L1; ++i) {
        int sx, sy, sz;
        const int fi = fold(i, L1, sx), fj = fold(j, L2, sy), fk = fold(k, L3, sz);
        const double2* v = &h[(((size_t)fj * F3 + fk) * F1 + fi) * 6];
        const double sgn[6] = {1.0, (double)(sx * sy), (double)(sx * sz), 1.0, (double)(sy * sz), 1.0};
        const size_t idx = ((size_t)k * L2 + j) * L1 + i;
        for (int comp = 0; comp < 6; ++comp) {
          G[comp][2 * idx] = sgn[comp] * v[comp].x;
          G[comp][2 * idx + 1] = sgn[comp] * v[comp].y;
        }
      }
  return IFDDA_OK;
}

extern "C" int ifdda_timing_enable(ifdda_ctx* c, int cap) {
  if (!c || cap < 0) return arg_error("ifdda_timing_enable: bad argument");
  PassTimer& t = c->timer;
  if (t.ev) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 2 * t.cap; ++i) (void)hipEventDestroy(t.ev[i]);
    delete[] t.ev;
    delete[] t.tag;
  }
  const int mask = t.mask, every = t.every;
  t = PassTimer();
  t.mask = mask;
  t.every = every;
  if (cap == 0) return IFDDA_OK;
  t.ev = new hipEvent_t[2 * (size_t)cap];
  t.tag = new int[cap];
  for (int i = 0; i < 2 * cap; ++i) CHECK_HIP(hipEventCreate(&t.ev[i]), "hipEventCreate");
  t.cap = cap;
  return IFDDA_OK;
}

extern "C" int ifdda_use_specialized(ifdda_ctx* c, int on) {
  if (!c) return arg_error("ifdda_use_specialized: NULL context");
  c->specialized = on != 0;
  return IFDDA_OK;
}

extern "C" int ifdda_zpass_registers(ifdda_ctx* c, int mask) {
  if (!c || mask < 0 || mask > 3) return arg_error("ifdda_zpass_registers: bad argument");
  c->zreg_mask = mask;
  return IFDDA_OK;
}

extern "C" int ifdda_fuse_krylov(ifdda_ctx* c, int mask) {
  if (!c || mask < 0 || mask > 31) return arg_error("ifdda_fuse_krylov: bad argument");
  c->fuse_mask = mask;
  return IFDDA_OK;
}

extern "C" int ifdda_specialized_z(ifdda_ctx* c, int prec, int* has) {
  if (!c || !has || (prec != 32 && prec != 64)) return arg_error("ifdda_specialized_z: bad argument");
  *has = has_specialized_z(c->g, prec) ? 1 : 0;
  return IFDDA_OK;
}

extern "C" int ifdda_timing_passes(ifdda_ctx* c, int mask, int every) {
  if (!c || mask < 0 || mask > 31 || every < 1) return arg_error("ifdda_timing_passes: bad argument");
  c->timer.mask = mask;
  c->timer.every = every;
  c->timer.seen = 0;
  return IFDDA_OK;
}

extern "C" int ifdda_timing_read(ifdda_ctx* c, double ms_sum[10], int count[10]) {
  if (!c || !ms_sum || !count) return arg_error("ifdda_timing_read: NULL argument");
  for (int i = 0; i < 10; ++i) {
    ms_sum[i] = 0.0;
    count[i] = 0;
  }
  PassTimer& t = c->timer;
  if (!t.cap) return IFDDA_OK;
  CHECK_HIP(hipStreamSynchronize(c->stream), "timing sync");
  for (int i = 0; i < t.used; ++i) {
    float ms = 0.f;
    CHECK_HIP(hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]), "hipEventElapsedTime");
    ms_sum[t.tag[i]] += ms;
    count[t.tag[i]] += 1;
  }
  t.used = 0;
  return IFDDA_OK;
}

static int ensure_staging(ifdda_ctx* c, size_t n) {
  if (c->hcap >= n) return IFDDA_OK;
  (void)hipFree(c->hx);
  (void)hipFree(c->hb);
  c->hx = c->hb = nullptr;
  c->hcap = 0;
  CHECK_HIP(hipMalloc(&c->hx, sizeof(double2) * n), "hipMalloc staging");
  CHECK_HIP(hipMalloc(&c->hb, sizeof(double2) * n), "hipMalloc staging");
  c->hcap = n;
  return IFDDA_OK;
}

extern "C" int ifdda_set_dipoles(ifdda_ctx* c, const int32_t* tabdip, int ndipole) {
  if (!c) return arg_error("ifdda_set_dipoles: NULL context");
  const GridPlan& g = c->g;
  const size_t N = (size_t)g.nx * g.ny * g.nz;
  (void)hipFree(c->rows);
  c->rows = nullptr;
  if (!tabdip) {
    (void)hipFree(c->tab);
    c->tab = nullptr;
    (void)hipFree(c->dipcell);
    c->dipcell = nullptr;
    c->nd = (long long)N;
    c->have_dipoles = true;
    c->tab_dense = true;
    const int box[6] = {0, g.nx - 1, 0, g.ny - 1, 0, g.nz - 1};
    std::memcpy(c->cellbox, box, sizeof(box));
    return IFDDA_OK;
  }
  if (ndipole < 0) return arg_error("ifdda_set_dipoles: ndipole < 0");
  std::vector<char> seen((size_t)ndipole + 1, 0);
  for (size_t i = 0; i < N; ++i) {
    const int32_t t = tabdip[i];
    if (t < 0 || t > ndipole)
      return arg_error("ifdda_set_dipoles: Tabdip entry out of range at cell " + std::to_string(i));
    if (t > 0) {
      if (seen[t]) return arg_error("ifdda_set_dipoles: dipole mapped twice");
      seen[t] = 1;
    }
  }
  // Every cell holding dipole cell+1 is the dense box of
  // produitfftmatvectopt.f:43-44: the passes then index dipoles by cell and
  // read no map.
  bool identity = (size_t)ndipole == N;
  for (size_t i = 0; identity && i < N; ++i) identity = tabdip[i] == (int32_t)(i + 1);
  // cell of each dipole (-1: not in the map) and the occupied-cell box
  std::vector<int32_t> cell((size_t)ndipole, -1);
  int box[6] = {g.nx, -1, g.ny, -1, g.nz, -1};
  for (size_t i = 0; i < N; ++i)
    if (tabdip[i] > 0) {
      cell[tabdip[i] - 1] = (int32_t)i;
      const int ii = (int)(i % g.nx), jj = (int)((i / g.nx) % g.ny), kk = (int)(i / ((size_t)g.nx * g.ny));
      box[0] = std::min(box[0], ii), box[1] = std::max(box[1], ii);
      box[2] = std::min(box[2], jj), box[3] = std::max(box[3], jj);
      box[4] = std::min(box[4], kk), box[5] = std::max(box[5], kk);
    }
  std::memcpy(c->cellbox, box, sizeof(box));
  (void)hipFree(c->dipcell);
  c->dipcell = nullptr;
  if (identity) {
    (void)hipFree(c->tab);
    c->tab = nullptr;
  } else {
    if (!c->tab) CHECK_HIP(hipMalloc(&c->tab, sizeof(int32_t) * N), "hipMalloc tabdip");
    CHECK_HIP(hipMemcpy(c->tab, tabdip, sizeof(int32_t) * N, hipMemcpyHostToDevice), "copy tabdip");
    std::vector<unsigned long long> rm;
    if (row_masks(tabdip, g.nx, g.ny, g.nz, rm)) {
      CHECK_HIP(hipMalloc(&c->rows, sizeof(unsigned long long) * rm.size()), "hipMalloc line masks");
      CHECK_HIP(hipMemcpy(c->rows, rm.data(), sizeof(unsigned long long) * rm.size(),
                          hipMemcpyHostToDevice),
                "copy line masks");
    }
    if (ndipole > 0) {
      CHECK_HIP(hipMalloc(&c->dipcell, sizeof(int32_t) * ndipole), "hipMalloc dipole cells");
      CHECK_HIP(hipMemcpy(c->dipcell, cell.data(), sizeof(int32_t) * ndipole, hipMemcpyHostToDevice),
                "copy dipole cells");
    }
  }
  c->nd = ndipole;
  c->have_dipoles = true;
  c->tab_dense = false;  // a caller's map (sync_tab compares its hash)
  return IFDDA_OK;
}

extern "C" int ifdda_set_polarizability(ifdda_ctx* c, const double* pola, int nmax, int layout) {
  if (!c || !pola) return arg_error("ifdda_set_polarizability: NULL argument");
  if (!c->have_dipoles) return state_error("ifdda_set_polarizability: set the dipoles first");
  const long long nd = c->nd;
  std::vector<double2> h;
  int mode;
  if (layout == 1) {
    mode = kAlphaScalar;
    h.resize(nd);
    for (long long i = 0; i < nd; ++i) h[i] = make_double2(pola[2 * i], pola[2 * i + 1]);
  } else if (layout == 0) {
    if (nmax < nd) return arg_error("ifdda_set_polarizability: nmax < number of dipoles");
    // polarisa(i,a,b) at i + nmax*(a + 3b) (column-major, 0-based)
    auto P = [&](long long i, int a, int b) {
      const size_t o = 2 * ((size_t)i + (size_t)nmax * (a + 3 * b));
      return make_double2(pola[o], pola[o + 1]);
    };
    bool iso = true;
    for (long long i = 0; i < nd && iso; ++i) {
      const double2 d0 = P(i, 0, 0);
      for (int a = 0; a < 3 && iso; ++a)
        for (int b = 0; b < 3; ++b) {
          const double2 v = P(i, a, b);
          if (a == b ? (v.x != d0.x || v.y != d0.y) : (v.x != 0.0 || v.y != 0.0)) {
            iso = false;
            break;
          }
        }
    }
    if (iso) {
      mode = kAlphaScalar;
      h.resize(nd);
      for (long long i = 0; i < nd; ++i) h[i] = P(i, 0, 0);
    } else {
      mode = kAlphaTensor;
      h.resize(nd * 9);
      for (long long i = 0; i < nd; ++i)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) h[i * 9 + a * 3 + b] = P(i, a, b);
    }
  } else {
    return arg_error("ifdda_set_polarizability: layout must be 0 or 1");
  }
  std::vector<float2> h32(h.size());
  for (size_t i = 0; i < h.size(); ++i) h32[i] = make_float2((float)h[i].x, (float)h[i].y);
  (void)hipFree(c->al64);
  (void)hipFree(c->al32);
  c->al64 = nullptr;
  c->al32 = nullptr;
  c->alpha_mode = kAlphaNone;
  if (!h.empty()) {
    CHECK_HIP(hipMalloc(&c->al64, sizeof(double2) * h.size()), "hipMalloc alpha");
    CHECK_HIP(hipMalloc(&c->al32, sizeof(float2) * h.size()), "hipMalloc alpha");
    CHECK_HIP(hipMemcpy(c->al64, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice),
              "copy alpha");
    CHECK_HIP(hipMemcpy(c->al32, h32.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice),
              "copy alpha");
  }
  c->alpha_mode = mode;
  return IFDDA_OK;
}

static int ready_for_matvec(const ifdda_ctx* c, const char* who) {
  if (!c) return arg_error(std::string(who) + ": NULL context");
  if (!c->green_ready) return state_error(std::string(who) + ": no Green tensor");
  if (!c->have_dipoles) return state_error(std::string(who) + ": no dipoles");
  return IFDDA_OK;
}

extern "C" int ifdda_matvec_device(ifdda_ctx* c, int prec, int apply_alpha, const void* x,
                                   const void* b, void* y, void* stream) {
  int rc = ready_for_matvec(c, "ifdda_matvec_device");
  if (rc) return rc;
  if (!x || !b || !y) return arg_error("ifdda_matvec_device: NULL vector");
  if (apply_alpha && c->alpha_mode == kAlphaNone)
    return state_error("ifdda_matvec_device: no polarisability set");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  if (c->nd == 0) return IFDDA_OK;
  if (prec == 64) {
    MatvecDev<double2> m = c->mv64();
    if (!apply_alpha) m.alpha_mode = kAlphaNone;
    return launch_matvec_f64(c->g, m, (const double2*)x, (const double2*)b, (double2*)y, s)
               ? IFDDA_OK
               : IFDDA_ERR_HIP;
  }
  if (prec == 32) {
    MatvecDev<float2> m = c->mv32();
    if (!apply_alpha) m.alpha_mode = kAlphaNone;
    return launch_matvec_f32(c->g, m, (const float2*)x, (const float2*)b, (float2*)y, s)
               ? IFDDA_OK
               : IFDDA_ERR_HIP;
  }
  return arg_error("ifdda_matvec_device: prec must be 64 or 32");
}

extern "C" int ifdda_matvec(ifdda_ctx* c, const double* FFX, double* FFB) {
  int rc = ready_for_matvec(c, "ifdda_matvec");
  if (rc) return rc;
  if (!FFX || !FFB) return arg_error("ifdda_matvec: NULL vector");
  const size_t n = 3 * (size_t)c->nd;
  if (n == 0) return IFDDA_OK;
  if ((rc = ensure_staging(c, n))) return rc;
  CHECK_HIP(hipMemcpyAsync(c->hx, FFX, sizeof(double2) * n, hipMemcpyHostToDevice, c->stream),
            "copy FFX");
  CHECK_HIP(hipMemcpyAsync(c->hb, FFB, sizeof(double2) * n, hipMemcpyHostToDevice, c->stream),
            "copy FFB");
  MatvecDev<double2> m = c->mv64();
  m.alpha_mode = kAlphaNone;
  if (!launch_matvec_f64(c->g, m, c->hx, c->hb, c->hb, c->stream)) return IFDDA_ERR_HIP;
  CHECK_HIP(hipMemcpyAsync(FFB, c->hb, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream),
            "copy FFB back");
  CHECK_HIP(hipStreamSynchronize(c->stream), "matvec");
  return IFDDA_OK;
}

// ---------------------------------------------------------------------------
// Reference-shaped entries.  One context per box size, created on first use
// and kept for the life of the process (as the reference keeps its FFTW
// plans, cdmlib.f:1155-1164).  The device Green tensor is tied to the host
// arrays it was last built for or taken from, by pointer and a 16-point
// fingerprint per component, so callers that rebuild the tensor between
// solves (cdmlib.f:3682-3730) are seen.
// ---------------------------------------------------------------------------

// Grow-only device staging of the host-array far-field entries (these run
// synchronously, so a growth frees a buffer no work still reads).
static void* x_stage(ifdda_ctx* c, size_t bytes) {
  if (bytes > c->xcap) {
    (void)hipFree(c->xbuf);
    c->xbuf = nullptr;
    c->xcap = 0;
    if (!hip_ok(hipMalloc(&c->xbuf, bytes), "hipMalloc far-field staging")) return nullptr;
    c->xcap = bytes;
  }
  return c->xbuf;
}

static std::mutex g_reg_mu;
static std::map<std::array<int, 3>, ifdda_ctx*> g_reg;

static ifdda_ctx* ctx_for(int nx, int ny, int nz) {
  std::lock_guard<std::mutex> lk(g_reg_mu);
  auto key = std::array<int, 3>{nx, ny, nz};
  auto it = g_reg.find(key);
  if (it != g_reg.end()) return it->second;
  ifdda_ctx* c = nullptr;
  if (ifdda_create(nx, ny, nz, &c) != IFDDA_OK) return nullptr;
  g_reg[key] = c;
  return c;
}

// A hash of every 61st complex value of a component (and of its last one):
// a Green tensor rebuilt for another wavelength, lattice step or box changes
// every value, so any rebuild between calls changes it.
static uint64_t green_hash(const double* G, size_t M) {
  uint64_t h = 1469598103934665603ULL;
  auto mix = [&](size_t i) {
    uint64_t a, b;
    std::memcpy(&a, G + 2 * i, sizeof(a));
    std::memcpy(&b, G + 2 * i + 1, sizeof(b));
    h = (h ^ a) * 1099511628211ULL;
    h = (h ^ b) * 1099511628211ULL;
  };
  for (size_t i = 0; i < M; i += 61) mix(i);
  if (M) mix(M - 1);
  return h;
}

static void take_fingerprint(ifdda_ctx* c, const double* const G[6], size_t M) {
  for (int comp = 0; comp < 6; ++comp) {
    c->green_host[comp] = G[comp];
    c->green_sig[comp] = green_hash(G[comp], M);
  }
}

static bool fingerprint_matches(const ifdda_ctx* c, const double* const G[6], size_t M) {
  if (!c->green_ready) return false;
  for (int comp = 0; comp < 6; ++comp)
    if (c->green_host[comp] != G[comp] || c->green_sig[comp] != green_hash(G[comp], M)) return false;
  return true;
}

// Make the device tensor match the caller's arrays.
static int sync_green(ifdda_ctx* c, const double* const G[6], int nx2, int ny2, int nz2) {
  const size_t M = (size_t)nx2 * ny2 * nz2;
  if (fingerprint_matches(c, G, M)) return IFDDA_OK;
  int rc = ifdda_green_set(c, G);
  if (rc) return rc;
  take_fingerprint(c, G, M);
  return IFDDA_OK;
}

static uint64_t hash_tab(const int32_t* t, size_t n) {
  uint64_t h = 1469598103934665603ULL;
  for (size_t i = 0; i < n; ++i) h = (h ^ (uint32_t)t[i]) * 1099511628211ULL;
  return h;
}

static int sync_tab(ifdda_ctx* c, const int32_t* Tabdip) {
  const size_t N = (size_t)c->g.nx * c->g.ny * c->g.nz;
  if (!Tabdip) {
    if (c->have_dipoles && c->tab_dense) return IFDDA_OK;
    return ifdda_set_dipoles(c, nullptr, 0);
  }
  const uint64_t h = hash_tab(Tabdip, N);
  if (c->have_dipoles && !c->tab_dense && h == c->tab_hash) return IFDDA_OK;
  int32_t nd = 0;
  for (size_t i = 0; i < N; ++i) nd = std::max(nd, Tabdip[i]);
  int rc = ifdda_set_dipoles(c, Tabdip, nd);
  if (rc == IFDDA_OK) c->tab_hash = h;
  return rc;
}

static void report(const char* who) {
  fprintf(stderr, "%s: %s\n", who, g_err.c_str());
}

static void matvec_entry(const char* who, const double* const G[6], const int32_t* Tabdip, int nx,
                         int ny, int nz, int nx2, int ny2, int nz2, const double* FFX,
                         double* FFB) {
  ifdda_ctx* c = ctx_for(nx, ny, nz);
  if (!c) return report(who);
  if (sync_green(c, G, nx2, ny2, nz2) || sync_tab(c, Tabdip) || ifdda_matvec(c, FFX, FFB))
    report(who);
}

extern "C" void ifdda_produitfftmatvect3(
    const double* Gxx, const double* Gxy, const double* Gxz, const double* Gyy, const double* Gyz,
    const double* Gzz, double* vectx, double* vecty, double* vectz, const int32_t* Tabdip,
    const int32_t* ntotalm, const int32_t* ntotal, const int32_t* nmax, const int32_t* ndipole,
    const int32_t* nxm, const int32_t* nym, const int32_t* nzm, const int32_t* nx,
    const int32_t* ny, const int32_t* nz, const int32_t* nx2, const int32_t* ny2,
    const int32_t* nxy2, const int32_t* nz2, const double* FFX, double* FFB, const int64_t* planb,
    const int64_t* planf) {
  (void)vectx; (void)vecty; (void)vectz; (void)ntotalm; (void)ntotal; (void)nmax;
  (void)ndipole; (void)nxm; (void)nym; (void)nzm; (void)nxy2; (void)planb; (void)planf;
  const double* G[6] = {Gxx, Gxy, Gxz, Gyy, Gyz, Gzz};
  matvec_entry("produitfftmatvect3", G, Tabdip, *nx, *ny, *nz, *nx2, *ny2, *nz2, FFX, FFB);
}

extern "C" void ifdda_produitfftmatvectopt(
    const double* Gxx, const double* Gxy, const double* Gxz, const double* Gyy, const double* Gyz,
    const double* Gzz, double* vectx, double* vecty, double* vectz, const int32_t* ntotalm,
    const int32_t* ntotal, const int32_t* nmax, const int32_t* nxm, const int32_t* nym,
    const int32_t* nzm, const int32_t* nx, const int32_t* ny, const int32_t* nz,
    const int32_t* nx2, const int32_t* ny2, const int32_t* nxy2, const int32_t* nz2,
    const double* FFX, double* FFB, const int64_t* planf, const int64_t* planb) {
  (void)vectx; (void)vecty; (void)vectz; (void)ntotalm; (void)ntotal; (void)nmax;
  (void)nxm; (void)nym; (void)nzm; (void)nxy2; (void)planb; (void)planf;
  const double* G[6] = {Gxx, Gxy, Gxz, Gyy, Gyz, Gzz};
  matvec_entry("produitfftmatvectopt", G, nullptr, *nx, *ny, *nz, *nx2, *ny2, *nz2, FFX, FFB);
}

// The optical-force FFTs (calculforcefft.f / calculforcefftopt.f, then
// derivativefield2.f, as cdmlib.f:3159-3300 calls them): the first keeps the
// moments, placed in box order, for the box; the second forms one derivative
// of their field at every box cell on the device.  The caller's vect and
// DFFTTENSOR arrays are scratch there and are left alone here.
static std::mutex g_mom_mu;
static std::map<ifdda_ctx*, std::vector<double>> g_mom;

static void force_moments(const char* who, int nx, int ny, int nz, const int32_t* Tabdip,
                          const double* FF) {
  ifdda_ctx* c = ctx_for(nx, ny, nz);
  if (!c) return report(who);
  const size_t N = (size_t)nx * ny * nz;
  std::vector<double> m(6 * N, 0.0);
  for (size_t i = 0; i < N; ++i) {
    const long long j = Tabdip ? (long long)Tabdip[i] - 1 : (long long)i;
    if (j < 0) continue;
    for (int q = 0; q < 6; ++q) m[6 * i + q] = FF[6 * j + q];
  }
  std::lock_guard<std::mutex> lk(g_mom_mu);
  g_mom[c] = std::move(m);
}

extern "C" void ifdda_calculforcefft(const int32_t* nx, const int32_t* ny, const int32_t* nz,
                                     const int32_t* nx2, const int32_t* ny2, const int32_t* nz2,
                                     const int32_t* nxy2, const int32_t* nxm, const int32_t* nym,
                                     const int32_t* nzm, const int32_t* Tabdip, double* vectx,
                                     double* vecty, double* vectz, const double* FF,
                                     const int64_t* planb) {
  (void)nx2; (void)ny2; (void)nz2; (void)nxy2; (void)nxm; (void)nym; (void)nzm;
  (void)vectx; (void)vecty; (void)vectz; (void)planb;
  force_moments("calculforcefft", *nx, *ny, *nz, Tabdip, FF);
}

extern "C" void ifdda_calculforcefftopt(const int32_t* nx, const int32_t* ny, const int32_t* nz,
                                        const int32_t* nx2, const int32_t* ny2, const int32_t* nz2,
                                        const int32_t* nxy2, const int32_t* nxm, const int32_t* nym,
                                        const int32_t* nzm, double* vectx, double* vecty,
                                        double* vectz, const double* FF, const int64_t* planb) {
  (void)nx2; (void)ny2; (void)nz2; (void)nxy2; (void)nxm; (void)nym; (void)nzm;
  (void)vectx; (void)vecty; (void)vectz; (void)planb;
  force_moments("calculforcefftopt", *nx, *ny, *nz, nullptr, FF);
}

extern "C" void ifdda_derivativefield2(
    const double* aretecube, const double* k0, const int32_t* nx, const int32_t* ny,
    const int32_t* nz, const int32_t* nx2, const int32_t* ny2, const int32_t* nz2,
    const int32_t* nxy2, const int32_t* ntotal, const int32_t* ntotalm, const int32_t* nmax,
    double* Dxx, double* Dxy, double* Dxz, double* Dyy, double* Dyz, double* Dzz, double* vectx,
    double* vecty, double* vectz, const int32_t* test, double* FF0, const int64_t* planb,
    const int64_t* planf) {
  (void)nx2; (void)ny2; (void)nz2; (void)nxy2; (void)ntotal; (void)ntotalm; (void)nmax;
  (void)Dxx; (void)Dxy; (void)Dxz; (void)Dyy; (void)Dyz; (void)Dzz;
  (void)vectx; (void)vecty; (void)vectz; (void)planb; (void)planf;
  ifdda_ctx* c = ctx_for(*nx, *ny, *nz);
  if (!c) return report("derivativefield2");
  std::vector<double> m;
  {
    std::lock_guard<std::mutex> lk(g_mom_mu);
    auto it = g_mom.find(c);
    if (it == g_mom.end()) {
      set_error("derivativefield2: call calculforcefft for this box first");
      return report("derivativefield2");
    }
    m = it->second;
  }
  if (derivative_field_box(c, m.data(), *aretecube, *k0, *test - 1, FF0))
    report("derivativefield2");
}

static void green_entry(const char* who, int nx, int ny, int nz, int nx2, int ny2, int nz2,
                        int nquad, double tolinit, double d, double k0, double* const G[6]) {
  ifdda_ctx* c = ctx_for(nx, ny, nz);
  if (!c) return report(who);
  if (ifdda_green_build(c, d, k0, nquad, tolinit)) return report(who);
  const size_t M = (size_t)nx2 * ny2 * nz2;
  const GridPlan& g = c->g;
  if (g.L[0] == nx2 && g.L[1] == ny2 && g.L[2] == nz2) {
    if (ifdda_green_get(c, G)) return report(who);
  } else {
    // The padded grid differs from the caller's 2n grid (2n not 2,3,5-smooth):
    // the device copy is the tensor; the caller's arrays are cleared.
    for (int comp = 0; comp < 6; ++comp) std::memset(G[comp], 0, sizeof(double) * 2 * M);
  }
  const double* Gc[6] = {G[0], G[1], G[2], G[3], G[4], G[5]};
  take_fingerprint(c, Gc, M);
}

extern "C" void ifdda_greencalculfft(const int32_t* nx, const int32_t* ny, const int32_t* nz,
                                     const int32_t* nx2, const int32_t* ny2, const int32_t* nz2,
                                     const int32_t* nxy2, const int32_t* nxm, const int32_t* nym,
                                     const int32_t* nzm, const double* aretecube, const double* k0,
                                     double* Gxx, double* Gxy, double* Gxz, double* Gyy,
                                     double* Gyz, double* Gzz, const int64_t* planb) {
  (void)nxy2; (void)nxm; (void)nym; (void)nzm; (void)planb;
  double* G[6] = {Gxx, Gxy, Gxz, Gyy, Gyz, Gzz};
  green_entry("greencalculfft", *nx, *ny, *nz, *nx2, *ny2, *nz2, 0, 0.0, *aretecube, *k0, G);
}

extern "C" void ifdda_greencalculquadfft(const int32_t* nx, const int32_t* ny, const int32_t* nz,
                                         const int32_t* nx2, const int32_t* ny2,
                                         const int32_t* nz2, const int32_t* nxy2,
                                         const int32_t* nxm, const int32_t* nym,
                                         const int32_t* nzm, const int32_t* nquad,
                                         const double* tolinit, const double* aretecube,
                                         const double* k0, double* Gxx, double* Gxy, double* Gxz,
                                         double* Gyy, double* Gyz, double* Gzz,
                                         const int64_t* planb) {
  (void)nxy2; (void)nxm; (void)nym; (void)nzm; (void)planb;
  double* G[6] = {Gxx, Gxy, Gxz, Gyy, Gyz, Gzz};
  green_entry("greencalculquadfft", *nx, *ny, *nz, *nx2, *ny2, *nz2, *nquad, *tolinit,
              *aretecube, *k0, G);
}

// Driver entries (inverserig / inverserigopt) share the solve of
// ifdda_krylov.hip through ifdda_solve().
static void fill_info(char* infostr, size_t len, const char* msg) {
  if (!infostr) return;
  const size_t n = std::min(len, std::strlen(msg));
  std::memcpy(infostr, msg, n);
  if (len > n) std::memset(infostr + n, ' ', len - n);
}

// inverserig.f:44-1076 dispatches on methodeit(1:k) .eq. 'NAME' in this
// order (Fortran compares blank-padded to k characters; the first match
// wins, so e.g. 'CG' must be followed by blanks).  Methods without a device
// implementation return 0 ('Iterative method not correct!').
static int method_from_name(const char* m, size_t len) {
  std::string s(m, len);
  auto eq = [&](size_t k, const char* name) {
    std::string a = s.substr(0, std::min(k, s.size())), b(name);
    a.resize(k, ' ');
    b.resize(k, ' ');
    return a == b;
  };
  if (eq(7, "GPBICG1")) return IFDDA_GPBICG1;
  if (eq(7, "GPBICG2")) return IFDDA_GPBICG2;
  if (eq(10, "GPBICGsafe")) return IFDDA_GPBICGSAFE;
  if (eq(10, "GPBICGplus")) return IFDDA_GPBICGPLUS;
  if (eq(8, "GPBICGAR")) return IFDDA_GPBICGAR;  // 'GPBICGAR2' too (inverserig.f:306 comes first)
  if (eq(12, "BICGSTARPLUS")) return IFDDA_BICGSTARPLUS;
  if (eq(6, "QMRCLA")) return IFDDA_QMRCLA;
  if (eq(5, "TFQMR")) return IFDDA_TFQMR;
  if (eq(5, "CG")) return IFDDA_CG;
  if (eq(8, "BICGSTAB")) return IFDDA_BICGSTAB;
  if (eq(12, "QMRBICGSTAB1")) return IFDDA_QMRBICGSTAB1;
  if (eq(12, "QMRBICGSTAB2")) return IFDDA_QMRBICGSTAB2;
  if (eq(7, "GPBICOR")) return IFDDA_GPBICOR;
  if (eq(7, "CORS")) return IFDDA_CORS;
  return 0;
}

static void driver_entry(const double* const G[6], const int32_t* Tabdip, int nx, int ny, int nz,
                         int nx2, int ny2, int nz2, int nmax, int nbsphere, double* FF,
                         const double* FF0, double* FFloc, const double* polarisa,
                         const char* methodeit, double tol, double* tol1, int32_t* nloop,
                         int32_t* ncompte, volatile int32_t* nstop, char* infostr,
                         size_t len_methodeit, size_t len_infostr) {
  const int method = method_from_name(methodeit, len_methodeit);
  if (!method) {
    *nstop = 1;
    fill_info(infostr, len_infostr, "Iterative method not correct!");
    return;
  }
  ifdda_ctx* c = ctx_for(nx, ny, nz);
  int rc = c ? IFDDA_OK : IFDDA_ERR_HIP;
  if (!rc) rc = sync_green(c, G, nx2, ny2, nz2);
  if (!rc) rc = sync_tab(c, Tabdip);
  if (!rc && c->nd != nbsphere) rc = arg_error("inverserig: nbsphere differs from the dipole map");
  if (!rc) rc = ifdda_set_polarizability(c, polarisa, nmax, 0);
  ifdda_solve_info info{};
  if (!rc) {
    ifdda_solve_opts o{};
    o.method = method;
    o.precision = 64;
    o.x0_zero = 0;
    o.maxit = 10000;
    o.tol = tol;
    rc = ifdda_solve(c, &o, FF0, FF, FFloc, &info, (volatile int*)nstop);
  }
  if (rc) {  // a device or argument failure, not a solver breakdown
    report("inverserig");
    *nstop = 1;
    fill_info(infostr, len_infostr, ("GPU: " + g_err).c_str());
    return;
  }
  *tol1 = info.tol1;
  *nloop = info.nloop;
  *ncompte += info.ncompte;
  if (info.nstop != 0) {
    *nstop = info.nstop;
    fill_info(infostr, len_infostr, info.infostr);
  }
}

extern "C" void ifdda_inverserig(
    const double* Gxx, const double* Gxy, const double* Gxz, const double* Gyy, const double* Gyz,
    const double* Gzz, double* vectx, double* vecty, double* vectz, const int32_t* Tabdip,
    const int32_t* ntotalm, const int32_t* ntotal, const int32_t* ldabi, const int32_t* nlar,
    const int32_t* nmax, const int32_t* ndipole, const int32_t* nxm, const int32_t* nym,
    const int32_t* nzm, const int32_t* nx, const int32_t* ny, const int32_t* nz,
    const int32_t* nx2, const int32_t* ny2, const int32_t* nxy2, const int32_t* nz2,
    const int32_t* nbsphere, const int32_t* nbsphere3, double* XI, double* XR, double* wrk,
    double* FF, const double* FF0, double* FFloc, const double* polarisa, const char* methodeit,
    const double* tol, double* tol1, int32_t* nloop, int32_t* ncompte, const int64_t* planf,
    const int64_t* planb, volatile int32_t* nstop, char* infostr, size_t len_methodeit,
    size_t len_infostr) {
  (void)vectx; (void)vecty; (void)vectz; (void)ntotalm; (void)ntotal; (void)ldabi; (void)nlar;
  (void)ndipole; (void)nxy2; (void)nbsphere3; (void)XI; (void)XR; (void)wrk; (void)planf;
  (void)planb;
  const double* G[6] = {Gxx, Gxy, Gxz, Gyy, Gyz, Gzz};
  const int nm = (*nxm) * (*nym) * (*nzm);
  (void)nmax;
  driver_entry(G, Tabdip, *nx, *ny, *nz, *nx2, *ny2, *nz2, nm, *nbsphere, FF, FF0, FFloc,
               polarisa, methodeit, *tol, tol1, nloop, ncompte, nstop, infostr, len_methodeit,
               len_infostr);
}

extern "C" void ifdda_inverserigopt(
    const double* Gxx, const double* Gxy, const double* Gxz, const double* Gyy, const double* Gyz,
    const double* Gzz, double* vectx, double* vecty, double* vectz, const int32_t* ntotalm,
    const int32_t* ntotal, const int32_t* ldabi, const int32_t* nlar, const int32_t* nmax,
    const int32_t* nxm, const int32_t* nym, const int32_t* nzm, const int32_t* nx,
    const int32_t* ny, const int32_t* nz, const int32_t* nx2, const int32_t* ny2,
    const int32_t* nxy2, const int32_t* nz2, const int32_t* nbsphere, const int32_t* nbsphere3,
    double* XI, double* XR, double* wrk, double* FF, const double* FF0, double* FFloc,
    const double* polarisa, const char* methodeit, const double* tol, double* tol1,
    int32_t* nloop, int32_t* ncompte, const int64_t* planf, const int64_t* planb,
    volatile int32_t* nstop, char* infostr, size_t len_methodeit, size_t len_infostr) {
  (void)vectx; (void)vecty; (void)vectz; (void)ntotalm; (void)ntotal; (void)ldabi; (void)nlar;
  (void)nxy2; (void)nbsphere3; (void)XI; (void)XR; (void)wrk; (void)planf; (void)planb;
  (void)nmax;
  const double* G[6] = {Gxx, Gxy, Gxz, Gyy, Gyz, Gzz};
  const int nm = (*nxm) * (*nym) * (*nzm);
  driver_entry(G, nullptr, *nx, *ny, *nz, *nx2, *ny2, *nz2, nm, *nbsphere, FF, FF0, FFloc,
               polarisa, methodeit, *tol, tol1, nloop, ncompte, nstop, infostr, len_methodeit,
               len_infostr);
}

// ------------------------------------------------- far field, reference-shaped

namespace ifdda {
std::string farfield_geometry(int nfft2d, double k0, double d, int* imax, double* dk);
int farfield_dev(ifdda_ctx* c, int R, const double2* FF, const int32_t* tab, long long nd,
                 int nfft2d, double k0, double d, double x0, double y0, double z0, double2* Epos,
                 double2* Eneg, int* imax_out, double* dk_out, hipStream_t s, int mode,
                 double numaper, int nside);
int gc_dev(ifdda_ctx* c, int R, const double2* Epos, const double2* Eneg, int imax, double dk,
           double k0, const double* I0, const double* theta, const double* phi, double* out,
           hipStream_t s);
}  // namespace ifdda

// diffractefft2d.f:1-3.  FF is box-ordered (the reference forces the dense
// box for this method, cdmlib.f:630); xs, ys, zs per cell; Ediffkzpos/neg are
// the caller's (nfft2d, nfft2d, 3) arrays (only the (2 imax + 1)^2 corner is
// filled, the rest zeroed, as the reference).  The Eloin scratch arrays and
// plan handles are accepted and not used.
extern "C" void ifdda_diffractefft2d(
    const int32_t* nx, const int32_t* ny, const int32_t* nz, const int32_t* nxm,
    const int32_t* nym, const int32_t* nzm, const int32_t* nfft2d, int32_t* tabfft2,
    const double* k0, const double* xs, const double* ys, const double* zs,
    const double* aretecube, double* Eloinx, double* Eloiny, double* Eloinz, const double* FF,
    int32_t* imax, double* deltakx, double* deltaky, double* Ediffkzpos, double* Ediffkzneg,
    const int64_t* plan2f, const int64_t* plan2b, int32_t* nstop, char* infostr,
    size_t len_infostr) {
  (void)nxm; (void)nym; (void)nzm; (void)Eloinx; (void)Eloiny; (void)Eloinz; (void)plan2f;
  (void)plan2b;
  const int N = *nfft2d;
  const int h = N / 2;
  for (int i = 1; i <= N; ++i) tabfft2[i - 1] = i - h - 1 >= 0 ? i - h : h + i;
  const size_t full = (size_t)N * N * 3;
  std::memset(Ediffkzpos, 0, sizeof(double) * 2 * full);
  std::memset(Ediffkzneg, 0, sizeof(double) * 2 * full);
  int im = 0;
  double dk = 0.0;
  const std::string msg = farfield_geometry(N, *k0, *aretecube, &im, &dk);
  *deltakx = *deltaky = dk;
  *imax = im;
  if (!msg.empty()) {
    *nstop = 1;
    fill_info(infostr, len_infostr, msg.c_str());
    return;
  }
  ifdda_ctx* c = ctx_for(*nx, *ny, *nz);
  const long long ncell = (long long)(*nx) * (*ny) * (*nz);
  const int n = 2 * im + 1;
  double2* dF = c ? (double2*)x_stage(c, sizeof(double2) * (3 * (size_t)ncell + 6 * (size_t)n * n))
                  : nullptr;
  double2* dE = dF ? dF + 3 * ncell : nullptr;
  bool ok = dF && hip_ok(hipMemcpy(dF, FF, sizeof(double2) * 3 * ncell, hipMemcpyHostToDevice),
                         "copy FF");
  if (ok)
    ok = farfield_dev(c, 1, dF, nullptr, ncell, N, *k0, *aretecube, xs[0], ys[0], zs[0], dE,
                      dE + 3 * (size_t)n * n, &im, &dk, c->stream, 0, 0.0, 1) == IFDDA_OK;
  std::vector<double2> hE(ok ? 6 * (size_t)n * n : 0);
  if (ok)
    ok = hip_ok(hipMemcpy(hE.data(), dE, sizeof(double2) * hE.size(), hipMemcpyDeviceToHost),
                "copy Ediff");
  if (!ok) {
    report("diffractefft2d");
    *nstop = 1;
    fill_info(infostr, len_infostr, "GPU far field failed");
    return;
  }
  // device [c][jj][ii] compact -> Fortran E(ii, jj, c) with leading dimension N
  for (int h2 = 0; h2 < 2; ++h2) {
    double* E = h2 == 0 ? Ediffkzpos : Ediffkzneg;
    const double2* src = hE.data() + (size_t)h2 * 3 * n * n;
    for (int cc = 0; cc < 3; ++cc)
      for (int jj = 0; jj < n; ++jj)
        for (int ii = 0; ii < n; ++ii) {
          const double2 v = src[((size_t)cc * n + jj) * n + ii];
          const size_t o = 2 * ((size_t)ii + (size_t)N * jj + (size_t)N * N * cc);
          E[o] = v.x;
          E[o + 1] = v.y;
        }
  }
}

// computegcfft2d.f:1-3 on the caller's (nfft2d, nfft2d, 3) far-field arrays.
extern "C" void ifdda_computegcfft2d(const int32_t* imax, const double* deltakx,
                                     const double* deltaky, const double* k0, const double* I0,
                                     const double* theta, const double* phi,
                                     const int32_t* nfft2d, const double* Ediffkzpos,
                                     const double* Ediffkzneg, double* gasym, double* Cscai) {
  (void)deltaky;  // the reference's grid is square (deltaky = deltakx)
  const int N = *nfft2d, n = 2 * (*imax) + 1;
  std::vector<double2> h(6 * (size_t)n * n);
  for (int h2 = 0; h2 < 2; ++h2) {
    const double* E = h2 == 0 ? Ediffkzpos : Ediffkzneg;
    for (int cc = 0; cc < 3; ++cc)
      for (int jj = 0; jj < n; ++jj)
        for (int ii = 0; ii < n; ++ii) {
          const size_t o = 2 * ((size_t)ii + (size_t)N * jj + (size_t)N * N * cc);
          h[(size_t)h2 * 3 * n * n + ((size_t)cc * n + jj) * n + ii] = make_double2(E[o], E[o + 1]);
        }
  }
  *gasym = 0.0;
  *Cscai = 0.0;
  ifdda_ctx* c = ctx_for(1, 1, 1);
  double2* dE = c ? (double2*)x_stage(c, sizeof(double2) * h.size()) : nullptr;
  double out[2] = {0.0, 0.0};
  bool ok = dE &&
            hip_ok(hipMemcpy(dE, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice),
                   "copy Ediff") &&
            gc_dev(c, 1, dE, dE + 3 * (size_t)n * n, *imax, *deltakx, *k0, I0, theta, phi, out,
                   c->stream) == IFDDA_OK;
  if (!ok) {
    report("computegcfft2d");
    return;
  }
  *Cscai = out[0];
  *gasym = out[1];
}
