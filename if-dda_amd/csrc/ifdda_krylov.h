// ifdda_krylov.h -- what the Krylov translation units share: the per-RHS
// device state, FP64 complex helpers, vector element access in either
// precision, the fixed-order reductions, the ragged RHS geometry, the solver
// workspace and the loop context the host drivers of the solvers run in.
#pragma once
#include <functional>
#include <vector>

#include "ifdda_internal.h"

namespace ifdda {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kMaxDots = 8;
static_assert(kMaxBlocks % kThreads == 0, "finish_sum splits the partials evenly");

struct KState {
  double norm;    // |b|
  double residu;  // recurrence residual estimate
  double tol;
  double2 alpha, beta, eta, dzeta, r0rn;    // GPBICG (eta/alpha shared by all)
  double2 ct_la, beta_la;  // GPBICG look-ahead: <r0*, r_new> and the next beta
  double2 rho, omega, tau, taut, theta, thetat, etat;  // QMRBICGSTAB / TFQMR
  double2 rho0, xxi;  // CG (zcg.f), BICGSTAB (pimzbicgstab.f)
  double rr;          // BICGSTAB: sum conj(r) r of the latest residual
  double2 gam, ksi, epsq, mu, kappa, lambda;  // QMRCLA (qmrpim.f GAMMA ... LAMBDA)
  double2 dots[kMaxDots];                   // last finished reduction
  int it;          // ITNO of the iteration in progress
  int done;        // solve ended (converged, breakdown or iteration cap)
  int status;      // 1 converged / cap, -1 breakdown
  int steperr;
  int maxit;
  int methode;     // QMRBICGSTAB omega formula
  int ntrace, tracecap;
  int lapend;      // GPBICG look-ahead: a residual check is due (g_f2xL)
};

// ---------------------------------------------------------------- complex FP64
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ bool zzero(double2 a) { return a.x == 0.0 && a.y == 0.0; }
__device__ __forceinline__ double zabs(double2 a) { return hypot(a.x, a.y); }
// Smith's division
__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.y + b.x * r;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
// principal square root
__device__ __forceinline__ double2 zsqrt(double2 z) {
  if (z.x == 0.0 && z.y == 0.0) return make_double2(0.0, 0.0);
  const double r = hypot(z.x, z.y);
  if (z.x >= 0.0) {
    const double t = sqrt(0.5 * (r + z.x));
    return make_double2(t, z.y / (2.0 * t));
  }
  const double t = sqrt(0.5 * (r - z.x));
  return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}
__device__ __forceinline__ double2 zone() { return make_double2(1.0, 0.0); }

// ------------------------------------------------------- vector element access
template <typename V> struct Vt;
template <> struct Vt<double2> {
  static __device__ __forceinline__ double2 ld(const double2* p, long long i) { return p[i]; }
  static __device__ __forceinline__ void st(double2* p, long long i, double2 v) { p[i] = v; }
  static __device__ __forceinline__ double2 rnd(double2 v) { return v; }
};
template <> struct Vt<float2> {
  static __device__ __forceinline__ double2 ld(const float2* p, long long i) {
    const float2 v = p[i];
    return make_double2((double)v.x, (double)v.y);
  }
  static __device__ __forceinline__ void st(float2* p, long long i, double2 v) {
    p[i] = make_float2((float)v.x, (float)v.y);
  }
  // the value a store followed by a load returns
  static __device__ __forceinline__ double2 rnd(double2 v) {
    return make_double2((double)(float)v.x, (double)(float)v.y);
  }
};
// Vector updates are evaluated in FP64 registers and stored in V.
#define LD(p, i) Vt<V>::ld(p, i)
#define ST(p, i, v) Vt<V>::st(p, i, v)

// -------------------------------------------------------------- reductions
template <int NR>
__device__ __forceinline__ void block_partials(double2 (&v)[NR], double2* part, int nblk, int bx) {
  for (int off = 32; off > 0; off >>= 1)
    for (int r = 0; r < NR; ++r) {
      v[r].x += __shfl_down(v[r].x, off, 64);
      v[r].y += __shfl_down(v[r].y, off, 64);
    }
  __shared__ double2 sh[NR][kThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int r = 0; r < NR; ++r) sh[r][wid] = v[r];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int r = 0; r < NR; ++r) {
      double2 s = make_double2(0.0, 0.0);
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s = zadd(s, sh[r][w]);
      part[r * nblk + bx] = s;
    }
}

// Sum the partials of NR reductions into st->dots.  Launched as one block of
// kThreads: every thread issues its (at most kMaxBlocks / kThreads) partial
// loads back to back, then a fixed-order reduction (thread, LDS, first
// wavefront shuffles) so the result does not depend on timing.
template <int NR>
__device__ __forceinline__ void finish_sum(KState* st, const double2* part, int nblk) {
  constexpr int kPer = kMaxBlocks / kThreads;
  __shared__ double2 red[NR][kThreads];
  const int tid = threadIdx.x;
  double2 v[NR][kPer];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int b = tid + k * kThreads;
      v[r][k] = b < nblk ? part[r * nblk + b] : make_double2(0.0, 0.0);
    }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double2 s = v[r][0];
#pragma unroll
    for (int k = 1; k < kPer; ++k) s = zadd(s, v[r][k]);
    red[r][tid] = s;
  }
  __syncthreads();
  // reduction r on wavefront r mod 4 (all four are full): the same sums in
  // the same order as on one wavefront, the NR of them side by side; thread
  // 0 stores them all (the thread that goes on to read them)
  __shared__ double2 res[NR];
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r % (kThreads / 64) != wv) continue;
    double2 s = red[r][lane];
#pragma unroll
    for (int k = 1; k < kThreads / 64; ++k) s = zadd(s, red[r][lane + 64 * k]);
    for (int off = 32; off > 0; off >>= 1) {
      s.x += __shfl_down(s.x, off, 64);
      s.y += __shfl_down(s.y, off, 64);
    }
    if (lane == 0) res[r] = s;
  }
  __syncthreads();
  if (tid == 0)
    for (int r = 0; r < NR; ++r) st->dots[r] = res[r];
  __syncthreads();
}

// As finish_sum for any number of partials (one per block of a matvec pass):
// each thread sums its strided share in order, then the same fixed-order
// reduction over threads.
template <int NR>
__device__ __forceinline__ void finish_sum_any(KState* st, const double2* part, int nblk) {
  __shared__ double2 red[NR][kThreads];
  const int tid = threadIdx.x;
  constexpr int kU = 8;  // loads issued before their sums
  double2 s[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) s[r] = make_double2(0.0, 0.0);
  for (int b0 = tid; b0 < nblk; b0 += kU * kThreads) {
    double2 v[NR][kU];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int k = 0; k < kU; ++k) {
        const int b = b0 + k * kThreads;
        v[r][k] = b < nblk ? part[r * nblk + b] : make_double2(0.0, 0.0);
      }
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int k = 0; k < kU; ++k) s[r] = zadd(s[r], v[r][k]);
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) red[r][tid] = s[r];
  __syncthreads();
  // reduction r on wavefront r mod 4 (all four are full): the same sums in
  // the same order as on one wavefront, the NR of them side by side; thread
  // 0 stores them all (the thread that goes on to read them)
  __shared__ double2 res[NR];
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r % (kThreads / 64) != wv) continue;
    double2 s = red[r][lane];
#pragma unroll
    for (int k = 1; k < kThreads / 64; ++k) s = zadd(s, red[r][lane + 64 * k]);
    for (int off = 32; off > 0; off >>= 1) {
      s.x += __shfl_down(s.x, off, 64);
      s.y += __shfl_down(s.y, off, 64);
    }
    if (lane == 0) res[r] = s;
  }
  __syncthreads();
  if (tid == 0)
    for (int r = 0; r < NR; ++r) st->dots[r] = res[r];
  __syncthreads();
}

// Vector geometry of a launch over right-hand sides: RHS r starts at element
// off(r) and holds len(r) elements -- r * n and n, or, for packed problems
// of different sizes, offs[r] and lens[r] (ragged, back to back).
//
// Blocks of a vector kernel over n elements (kThreads each, at most kMaxBlocks).
__host__ __device__ __forceinline__ int blocks_of(long long n) {
  const long long b = (n + kThreads - 1) / kThreads;
  return (int)(b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b));
}
//
// Ragged layouts launch a flat grid: kb (R + 1 ints) gives each RHS its
// blocks [kb[r], kb[r + 1]), as many as a solve of that RHS alone launches
// (or none once it has converged), so no block of the grid idles.  Equal
// lengths launch (nblk, R) with the RHS in blockIdx.y.
struct VN {
  long long n;
  const long long* lens;
  const long long* offs;
  const int* kb = nullptr;
  int R = 0;
  __device__ __forceinline__ long long len(int r) const { return lens ? lens[r] : n; }
  __device__ __forceinline__ long long off(int r) const { return offs ? offs[r] : r * n; }
  // the RHS of this block (block-uniform: scalar loads)
  __device__ __forceinline__ int rhs() const {
    if (!kb) return blockIdx.y;
    int lo = 0, hi = R;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (kb[mid] <= (int)blockIdx.x)
        lo = mid;
      else
        hi = mid;
    }
    return lo;
  }
  // the block's index among the blocks of RHS r
  __device__ __forceinline__ int bx(int r) const { return kb ? (int)blockIdx.x - kb[r] : blockIdx.x; }
};

// Blocks that serve right-hand side r: the whole grid row when all RHS have
// the same length, else the blocks a solve of RHS r alone would launch
// (flat grid, VN::kb).  Either way element i goes to the same block as in
// that solve, so the partial sums and their order are the same.  Partial
// slots past an RHS's block count hold zeros (written when the workspace is
// laid out), which the fixed-order finish adds exactly.
__device__ __forceinline__ int rhs_blocks(const VN& n, int r) {
  if (!n.lens) return gridDim.x;
  return n.kb ? n.kb[r + 1] - n.kb[r] : blocks_of(n.lens[r]);
}
#define RHS_BLOCKS(n, r, NR)                                                   \
  const int nbr_ = rhs_blocks(n, r);                                           \
  const int bx_ = (n).bx(r);                                                   \
  if (bx_ >= nbr_) {                                                           \
    if (threadIdx.x == 0)                                                      \
      for (int q_ = 0; q_ < NR; ++q_) part[q_ * nblk + bx_] = make_double2(0.0, 0.0); \
    return;                                                                    \
  }
#define RHS_BLOCKS_NOSUM(n, r)          \
  const int nbr_ = rhs_blocks(n, r);    \
  const int bx_ = (n).bx(r);            \
  if (bx_ >= nbr_) return;
// Grid-stride loop over the elements of the block's RHS (nbr_ blocks).
#define GRID_LOOP(i, n) \
  for (long long i = bx_ * (long long)blockDim.x + threadIdx.x; i < (n); i += (long long)nbr_ * blockDim.x)

__device__ __forceinline__ void push_trace(KState* st, double2* trace, double2 a, double2 b,
                                           double2 c, double2 d, double2 e, double res) {
  if (trace && st->ntrace < st->tracecap) {
    double2* t = trace + 6 * st->ntrace;
    t[0] = a; t[1] = b; t[2] = c; t[3] = d; t[4] = e; t[5] = make_double2(res, 0.0);
    st->ntrace++;
  }
}

// Workspace for up to R right-hand sides of n = 3*nd complex values each.
struct Solver {
  long long n = 0;  // 3 * dipoles per RHS
  int R = 0;        // right-hand sides the buffers hold
  int nblk = 1;     // blocks per RHS in vector kernels
  double2* v64[14] = {};
  float2* v32[14] = {};
  double2* part = nullptr;   // R * kMaxDots * nblk partial sums
  KState* st = nullptr;      // R states
  KState* st_host = nullptr; // pinned copy
  int* act = nullptr;        // R refinement flags: 1 = converged (skip)
  double2* trace = nullptr;  // per-iteration scalars of RHS 0 (single-RHS solves)
  int tracecap = 0;
  double2* rr64 = nullptr;   // refinement residual (FP64), R * n
  long long* lens = nullptr; // ragged layout (packed problems): per-RHS lengths
  long long* offs = nullptr; // and offsets, else null (RHS r at r * n)
  long long total = 0;       // elements per vector array (R * n, or the ragged sum)
  long long cap_total = 0;   // capacity of the buffers: elements, right-hand sides
  int cap_R = 0;
  int part_blocks = 0;       // blocks per RHS the partial-sum array holds
  double2* part5 = nullptr;  // partials of inner products formed in fused matvec ends
  long long part5_cap = 0;   // elements of part5
  int* kb = nullptr;         // ragged layouts: block offsets of the flat vector grid (cap_R + 1)
  int kb_full = 0;           // blocks of that grid with every RHS in
  std::vector<long long> lens_h;  // ragged layouts: the lengths on the host
};

// Outcome of one Krylov run for one right-hand side.
struct KRun {
  int status = 0, steperr = -1, nloop = 0, nmv = 0;
};

// Batched matvec over R right-hand sides (vectors n apart); skip flags
// skip[r * skip_stride] drop converged RHS on the device.
template <typename V> bool mv(const MvOp& op, const V* x, const V* b, V* y, hipStream_t s,
                                     const int* skip, int R, long long n, int skip_stride);
template <> inline bool mv<double2>(const MvOp& op, const double2* x, const double2* b, double2* y,
                             hipStream_t s, const int* skip, int R, long long n, int skip_stride) {
  return op.mv64(x, b, y, s, skip, R, n, skip_stride);
}
template <> inline bool mv<float2>(const MvOp& op, const float2* x, const float2* b, float2* y,
                            hipStream_t s, const int* skip, int R, long long n, int skip_stride) {
  return op.mv32(x, b, y, s, skip, R, n, skip_stride);
}

// The matvec with GPBICG work fused into its ends (MvOp::mvf64 / mvf32).
template <typename V> bool mvf(const MvOp& op, const V* x, const V* b, V* y, hipStream_t s,
                               const int* skip, int R, long long n, int skip_stride,
                               const KFuse<V>& f);
template <> inline bool mvf<double2>(const MvOp& op, const double2* x, const double2* b, double2* y,
                                     hipStream_t s, const int* skip, int R, long long n,
                                     int skip_stride, const KFuse<double2>& f) {
  return op.mvf64(x, b, y, s, skip, R, n, skip_stride, f);
}
template <> inline bool mvf<float2>(const MvOp& op, const float2* x, const float2* b, float2* y,
                                    hipStream_t s, const int* skip, int R, long long n,
                                    int skip_stride, const KFuse<float2>& f) {
  return op.mvf32(x, b, y, s, skip, R, n, skip_stride, f);
}

// The transposed operator over R right-hand sides (MvOp::mvt64 / mvt32).
template <typename V> bool mvt(const MvOp& op, const V* x, const V* zero, V* y, hipStream_t s,
                               const int* skip, int R, long long n, int skip_stride);
template <> inline bool mvt<double2>(const MvOp& op, const double2* x, const double2* zero,
                                     double2* y, hipStream_t s, const int* skip, int R, long long n,
                                     int skip_stride) {
  return op.mvt64(x, zero, y, s, skip, R, n, skip_stride);
}
template <> inline bool mvt<float2>(const MvOp& op, const float2* x, const float2* zero, float2* y,
                                    hipStream_t s, const int* skip, int R, long long n,
                                    int skip_stride) {
  return op.mvt32(x, zero, y, s, skip, R, n, skip_stride);
}

// What a solver's host loop needs: the operator, the workspace, launch
// geometry and the poll of the device state (step_poll(it): 1 stop, 0 go on,
// -1 HIP error), called after every queued iteration.
struct KLoop {
  const MvOp& op;
  Solver* S;
  VN vn;
  const dim3* grid;  // vector kernels (the flat grid of a ragged layout shrinks as RHS finish)
  int nb, R;
  long long n;
  hipStream_t s;
  double2* trace;
  const int* done;  // &S->st->done: matvec skip flags, dstride ints apart
  int dstride;
  int nmv0;         // matvecs before the loop (A x0)
  int x0zero;
  const char* active;
  std::function<int(int)> step_poll;
};

// GPBICGsafe, GPBICGplus, GPBICGAR, BICGSTARPLUS, GPBICOR, CORS, QMRCLA
// (ifdda_krylov_gp.hip); vv as for the other solvers (slot 0 = x, 1 = b,
// 13 = A x0).  False on a HIP error.
bool krylov_gp_method(int method);
template <typename V> bool krylov_gp(KLoop& L, V** vv, int method, KRun* out);

}  // namespace ifdda
