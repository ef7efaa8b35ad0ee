// ifdda_fft.hip -- LDS-staged mixed-radix (2,3,4,5,8) Stockham FFT passes for
// gfx950 and the five fused passes of the DDA matvec
//
//   P1 scatter(alpha.x) + FFT-x   (k_xpass, MODE 1)
//   P2 FFT-y                      (k_strided)
//   P3 FFT-z . G . IFFT-z         (k_zconv)
//   P4 IFFT-y                     (k_strided)
//   P5 IFFT-x + gather, b - v/M   (k_xpass, MODE 2)
//
// which together compute produitfftmatvect3.f:2-129 / produitfftmatvectopt.f
// on the zero-padded (L1,L2,L3) grid without touching the all-zero lines of
// the padding (only nz z-planes and ny y-rows are ever stored).
//
// One workgroup transforms a tile of NL = B*NC lines held in LDS as
// [point][line] (line fastest, so a wavefront reads contiguous LDS words);
// the FFT runs in place in that one buffer (decimation in time from a
// digit-reversed load; its transpose, decimation in frequency, for the
// inverse inside P3), one butterfly per thread at a time.  For the strided axes the B
// lines of a tile are B consecutive kx values (contiguous in HBM), so every
// global row access is a contiguous B*NC-element segment.
#include "ifdda_fftdev.h"
#include "ifdda_internal.h"
#include "ifdda_krylov.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace ifdda {

// In-place mixed-radix FFT of a tile held in LDS as [pos][line].
//
// Radices r_1..r_k (AxisPlan order), m_s = r_1...r_{s-1}.  Position
// pos = sum_s q_s m_s (q_s < r_s) corresponds to signal index
// n = sum_s q_s r_{s+1}...r_k; perm[n] gives pos.  Decimation in time (DIT)
// takes the signal in permuted positions and leaves the transform in natural
// order; its transpose, decimation in frequency (DIF), takes natural order
// and leaves the transform permuted.  Every butterfly of stage s reads and
// writes the same positions base + j + q*m_s (base = block*m_s*R, j < m_s), so
// a thread holds one butterfly (R values) at a time.  tw[e] = exp(+2 pi i e/L).
//
// Thread jt of a line takes butterflies t = jt, jt + JT, ...; (block, j) =
// divmod(t, m), the LDS element of point 0 and the twiddle exponent are
// advanced by constant steps, so the loop has no integer division and no
// per-point multiply.
//
// PRUNE 1 (first DIT stage, m = 1): points q >= R/2 of every butterfly are
// signal indices n >= L/2, zero padding when the line holds n_in <= L/2
// values; they are neither read nor need to be initialised.  PRUNE 2 (last
// DIF stage, m = 1): outputs q >= R/2 are transform indices >= L/2, which no
// caller keeps (n_out <= L/2); they are not computed or written.
template <int R, int SGN, bool DIT, int PRUNE, typename V>
__device__ __forceinline__ void stage_ip(V* buf, const V* __restrict__ tw, int L, int m, int NL,
                                         int line, int jt, int JT) {
  const unsigned um = (unsigned)m, uJT = (unsigned)JT, uNL = (unsigned)NL;
  const unsigned nb = (unsigned)L / R;
  const unsigned stride = um * uNL;  // LDS elements between the points of a butterfly
  const unsigned tws = (unsigned)L / (um * R);
  const unsigned dq = uJT / um, dr = uJT - dq * um;  // divmod(JT, m)
  unsigned j = (unsigned)jt % um;
  unsigned e0 = (((unsigned)jt - j) * R + j) * uNL + (unsigned)line;
  unsigned ex = j * tws;
  const unsigned de = (dq * um * R + dr) * uNL, dwrap = (um * (R - 1)) * uNL;
  const unsigned dx = dr * tws, xwrap = um * tws;
  for (unsigned t = (unsigned)jt; t < nb; t += uJT) {
    V* p = buf + e0;
    V a[R];
    constexpr int RIN = PRUNE == 1 ? R / 2 : R;
    constexpr int ROUT = PRUNE == 2 ? R / 2 : R;
#pragma unroll
    for (int q = 0; q < RIN; ++q) a[q] = p[q * stride];
    if constexpr (DIT && PRUNE == 0) if (um > 1) {
      unsigned xq = ex;
#pragma unroll
      for (int q = 1; q < R; ++q) {
        V w = tw[xq];
        xq += ex;
        if (SGN < 0) w.y = -w.y;
        a[q] = cmul(a[q], w);
      }
    }
    if constexpr (PRUNE == 1)
      DftHalfIn<R, SGN, V>::run(a);
    else if constexpr (PRUNE == 2)
      DftHalfOut<R, SGN, V>::run(a);
    else
      Dft<R, SGN, V>::run(a);
    if constexpr (!DIT && PRUNE == 0) if (um > 1) {
      unsigned xq = ex;
#pragma unroll
      for (int q = 1; q < R; ++q) {
        V w = tw[xq];
        xq += ex;
        if (SGN < 0) w.y = -w.y;
        a[q] = cmul(a[q], w);
      }
    }
#pragma unroll
    for (int q = 0; q < ROUT; ++q) p[q * stride] = a[q];
    j += dr;
    e0 += de;
    ex += dx;
    if (j >= um) {
      j -= um;
      e0 += dwrap;
      ex -= xwrap;
    }
  }
  lds_sync();
}

template <int SGN, bool DIT, typename V>
__device__ __forceinline__ void stage_any(int r, V* buf, const V* tw, int L, int m, int NL,
                                          int line, int jt, int JT) {
  switch (r) {
    case 2: stage_ip<2, SGN, DIT, 0>(buf, tw, L, m, NL, line, jt, JT); break;
    case 3: stage_ip<3, SGN, DIT, 0>(buf, tw, L, m, NL, line, jt, JT); break;
    case 4: stage_ip<4, SGN, DIT, 0>(buf, tw, L, m, NL, line, jt, JT); break;
    case 5: stage_ip<5, SGN, DIT, 0>(buf, tw, L, m, NL, line, jt, JT); break;
    default: stage_ip<8, SGN, DIT, 0>(buf, tw, L, m, NL, line, jt, JT); break;
  }
}

// Pruned first DIT / last DIF stage (m = 1, even radix).
template <int SGN, bool DIT, typename V>
__device__ __forceinline__ void stage_pruned(int r, V* buf, const V* tw, int L, int NL, int line,
                                             int jt, int JT) {
  constexpr int P = DIT ? 1 : 2;
  switch (r) {
    case 2: stage_ip<2, SGN, DIT, P>(buf, tw, L, 1, NL, line, jt, JT); break;
    case 4: stage_ip<4, SGN, DIT, P>(buf, tw, L, 1, NL, line, jt, JT); break;
    default: stage_ip<8, SGN, DIT, P>(buf, tw, L, 1, NL, line, jt, JT); break;
  }
}

// Whether the first radix of the plan allows the half-line pruning.
__device__ __forceinline__ bool prunable(const AxisPlan& ax) {
  return (ax.rad[0] & 1) == 0;
}

// Thread layout: line = tid % NL, jt = tid / NL, JT = blockDim / NL.
// half_in: the signal is zero at n >= L/2 (caller checked 2 n_in <= L).
template <int SGN, typename V>
__device__ __forceinline__ void tile_fft_dit(V* buf, const V* tw, const AxisPlan& ax, int NL, int line, int jt,
                             int JT, bool half_in) {
  int m = 1, s = 0;
  if (half_in && prunable(ax)) {
    stage_pruned<SGN, true>(ax.rad[0], buf, tw, ax.L, NL, line, jt, JT);
    m = ax.rad[0];
    s = 1;
  }
  for (; s < ax.nst; ++s) {
    stage_any<SGN, true>(ax.rad[s], buf, tw, ax.L, m, NL, line, jt, JT);
    m *= ax.rad[s];
  }
}

// half_out: only transform indices < L/2 are read afterwards.
template <int SGN, typename V>
__device__ __forceinline__ void tile_fft_dif(V* buf, const V* tw, const AxisPlan& ax, int NL, int line, int jt,
                             int JT, bool half_out) {
  int m = ax.L;
  const int last = half_out && prunable(ax) ? 1 : 0;
  for (int s = ax.nst - 1; s >= last; --s) {
    m /= ax.rad[s];
    stage_any<SGN, false>(ax.rad[s], buf, tw, ax.L, m, NL, line, jt, JT);
  }
  if (last) stage_pruned<SGN, false>(ax.rad[0], buf, tw, ax.L, NL, line, jt, JT);
}

// Rows the load must initialise: with the first-stage pruning, positions of
// signal indices >= L/2 are never read.
__device__ __forceinline__ int rows_to_fill(const AxisPlan& ax, bool half_in) {
  return half_in && prunable(ax) ? ax.L / 2 : ax.L;
}

// Forward transforms (SGN > 0) run DIT from a permuted load with the
// first-stage pruning; inverse transforms (SGN < 0) run DIF from a natural
// load with the last-stage pruning and store through pos.
//
// strided_tile computes logical tile lg (outer index lg / gx, kx tile
// lg % gx) with nthr threads.  In a packed launch (PK) the block has more
// threads than the tile uses: those at or above nthr take the loop start
// kIdle, beyond every loop bound, so they only meet the barriers.
template <typename V, int NC, int SGN, bool PK>
__device__ __forceinline__ void strided_tile(const SArgs<V>& a, const V* zsrc, V* zdst, int lg,
                                             int gx, int tid0, int nthr) {
  extern __shared__ double4 smem_d4[];
  V* tw = reinterpret_cast<V*>(smem_d4);
  const int L = a.ax.L;
  int* pos = reinterpret_cast<int*>(tw + L);
  V* buf = tile_base(tw, L);
  const int NL = a.B * NC;
  const int tid = PK && tid0 >= nthr ? kIdle : tid0;
  const int line = tid0 % NL, jt = tid / NL, JT = nthr / NL;
  const int o = lg / gx;
  const int kx0 = (lg - o * gx) * a.B;
  const int nvalid = min(a.B, a.nbx - kx0) * NC;
  // rows of plane o that hold dipoles (SArgs::rows; the block is uniform)
  const unsigned long long rm = a.rows ? a.rows[o] : ~0ull;
  if (SGN < 0 && rm == 0) return;

  load_tables(tw, pos, a.tw, a.pos, L, tid, nthr);
  lds_sync();
  const V* src = zsrc + (long long)o * a.so_src + (long long)kx0 * NC;
  V* dst = zdst + (long long)o * a.so_dst + (long long)kx0 * NC;
  if (SGN > 0) {
    const bool half = 2 * a.n_in <= L;
    load_strided<true>(buf, pos, src, a.sp_src, rows_to_fill(a.ax, half), a.n_in, NL, line,
                       nvalid, jt, JT, rm);
    tile_fft_dit<SGN>(buf, tw, a.ax, NL, line, jt, JT, half);
    if (line < nvalid) {
      V* d = dst + line + (long long)jt * a.sp_dst;
      const long long step = (long long)JT * a.sp_dst;
      for (int p = jt; p < a.n_out; p += JT, d += step) *d = buf[p * NL + line];
    }
  } else {
    load_strided<false>(buf, pos, src, a.sp_src, L, a.n_in, NL, line, nvalid, jt, JT);
    tile_fft_dif<SGN>(buf, tw, a.ax, NL, line, jt, JT, 2 * a.n_out <= L);
    if (line < nvalid) {
      V* d = dst + line + (long long)jt * a.sp_dst;
      const long long step = (long long)JT * a.sp_dst;
      for (int p = jt; p < a.n_out; p += JT, d += step)
        if (row_on(rm, p)) *d = buf[pos[p] * NL + line];
    }
  }
}

template <typename V, int NC, int SGN>
__global__ void __launch_bounds__(256) k_strided(SArgs<V> a) {
  // a stays unmodified (kernel-argument memory); per-RHS offsets go to locals
  if (a.skip && a.skip[blockIdx.z * a.skip_stride]) return;
  const int lg = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  strided_tile<V, NC, SGN, false>(a, a.src + blockIdx.z * a.zs_src, a.dst + blockIdx.z * a.zs_dst,
                                  lg, gridDim.x, threadIdx.x, blockDim.x);
}

// ---------------------------------------------------------------------------
// P3: FFT-z . G . IFFT-z on the y-transformed planes, in place.  One block
// takes a tile of B consecutive kx and the two mirror rows ky' and L2-ky'
// (one after the other); each thread loads the folded Green entries of its
// items (kz', kxl), kz' <= L3/2, and applies them to kz' and L3-kz' with the
// parity signs, so the octant array is read once per (kx tile, ky').  The
// forward transform is DIT (natural output for the product), the inverse DIF
// (natural input, permuted output read back through pos).
// ---------------------------------------------------------------------------
// zconv_tile: logical tile lg (ky' = lg / gx, kx tile lg % gx), nthr threads
// (PK: as strided_tile; DER: the tensor of a.dpar, see SArgs).
template <typename V, int NG, bool PK, bool DER = false>
__device__ __forceinline__ void zconv_tile(const SArgs<V>& a, const V* zsrc, V* zdst, int lg,
                                           int gx, int tid0, int nthr) {
  extern __shared__ double4 smem_d4[];
  V* tw = reinterpret_cast<V*>(smem_d4);
  const int L = a.ax.L;  // L3
  int* pos = reinterpret_cast<int*>(tw + L);
  V* buf = tile_base(tw, L);
  const int NL = a.B * 3;
  const int tid = PK && tid0 >= nthr ? kIdle : tid0;
  const int line = tid0 % NL, jt = tid / NL, JT = nthr / NL;
  const int kyf = lg / gx;  // 0 .. L2/2
  const int kx0 = (lg - kyf * gx) * a.B;
  const int nvalid = min(a.B, a.nbx - kx0) * 3;
  const int F3 = L / 2 + 1;

  load_tables(tw, pos, a.tw, a.pos, L, tid, nthr);
  lds_sync();
  const bool half_in = 2 * a.n_in <= L, half_out = 2 * a.n_out <= L;
  const int nrow = (kyf == 0 || 2 * kyf == a.L2) ? 1 : 2;
  for (int rr = 0; rr < nrow; ++rr) {
    const int ky = rr == 0 ? kyf : a.L2 - kyf;
    const bool sy = rr == 1;  // y odd-parity sign for the mirror row
    load_strided<true>(buf, pos, zsrc + (long long)ky * a.so_src + (long long)kx0 * 3, a.sp_src,
                       rows_to_fill(a.ax, half_in), a.n_in, NL, line, nvalid, jt, JT);
    tile_fft_dit<+1>(buf, tw, a.ax, NL, line, jt, JT, half_in);
    // Green entries of this thread's items (kz', kxl), all loads issued
    // first.  (Keeping the item indices live across the FFTs instead of
    // recomputing them here measured slower.)
    V gr[NG][6];
    int gx[NG];  // bit0 valid, bit1 x odd-parity sign
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      const int idx = tid + it * nthr;
      gx[it] = 0;
      if (idx < F3 * a.B) {
        const int kzf = idx / a.B, kxl = idx - kzf * a.B;
        const int kx = kx0 + kxl;
        if (kx < a.nbx) {
          int sx;
          const int kxf = fold_idx(kx, a.L1, sx);
          const V* g = a.G + ((long long)(kyf * a.F3 + kzf) * a.F1 + kxf) * 6;
#pragma unroll
          for (int c = 0; c < 6; ++c) gr[it][c] = g[c];
          gx[it] = 1 | ((sx < 0) << 1);
        }
      }
    }
    // Symmetric 3x3 product (produitfftmatvect3.f:76-90) at kz' and L3-kz';
    // the off-diagonal entries carry the parity signs of their two axes.
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      if (!(gx[it] & 1)) continue;
      const int idx = tid + it * nthr;
      const int kzf = idx / a.B, kxl = idx - kzf * a.B;
      const bool sxn = gx[it] & 2;
      const int nz2 = (kzf == 0 || 2 * kzf == L) ? 1 : 2;
      for (int zz = 0; zz < nz2; ++zz) {
        const int kz = zz == 0 ? kzf : L - kzf;
        const bool szn = zz == 1;
        V* v = buf + kz * NL + kxl * 3;
        sym_product(v, gr[it], sxn != sy, sxn != szn, sy != szn);
        if constexpr (DER) {
          if ((a.dpar == 1 && sxn) || (a.dpar == 2 && sy) || (a.dpar == 3 && szn))
            for (int c = 0; c < 3; ++c) v[c] = mk(-v[c].x, -v[c].y);
        }
      }
    }
    lds_sync();
    tile_fft_dif<-1>(buf, tw, a.ax, NL, line, jt, JT, half_out);
    if (line < nvalid) {
      V* d = zdst + (long long)ky * a.so_dst + (long long)kx0 * 3 + line + (long long)jt * a.sp_dst;
      const long long step = (long long)JT * a.sp_dst;
      for (int p = jt; p < a.n_out; p += JT, d += step) *d = buf[pos[p] * NL + line];
    }
    lds_sync();  // the next row reuses the LDS tile
  }
}

template <typename V, int NG>
__global__ void __launch_bounds__(256) k_zconv(SArgs<V> a) {
  if (a.skip && a.skip[blockIdx.z * a.skip_stride]) return;
  const int lg = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  zconv_tile<V, NG, false>(a, a.src + blockIdx.z * a.zs_src, a.dst + blockIdx.z * a.zs_dst, lg,
                           gridDim.x, threadIdx.x, blockDim.x);
}

// P3 for the derivative tensors of the optical force (derivativefield2.f)
template <typename V, int NG>
__global__ void __launch_bounds__(256) k_zconv_der(SArgs<V> a) {
  if (a.skip && a.skip[blockIdx.z * a.skip_stride]) return;
  const int lg = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  zconv_tile<V, NG, false, true>(a, a.src + blockIdx.z * a.zs_src, a.dst + blockIdx.z * a.zs_dst,
                                 lg, gridDim.x, threadIdx.x, blockDim.x);
}

// Rows p = jt, jt + JT, ... < FILL of one tile line at element stride sp into
// registers (zero for p >= n_in or an invalid line), and from registers to
// LDS at their permuted positions: load_strided in two halves, so that the
// loads of one row can be in flight while another is transformed.
template <int FILL, int JT, typename V>
__device__ __forceinline__ void fetch_rows(V (&v)[(FILL + JT - 1) / JT], const V* src,
                                           long long sp, int n_in, int line, int nvalid, int jt) {
#pragma unroll
  for (int u = 0; u < (FILL + JT - 1) / JT; ++u) {
    const int p = jt + u * JT;
    v[u] = czero<V>();
    if (p < n_in && line < nvalid) v[u] = src[(long long)p * sp + line];
  }
}
template <int FILL, int JT, int NL, typename V>
__device__ __forceinline__ void put_rows(V* buf, const int* pos,
                                         const V (&v)[(FILL + JT - 1) / JT], int line, int jt) {
#pragma unroll
  for (int u = 0; u < (FILL + JT - 1) / JT; ++u) {
    const int p = jt + u * JT;
    if (p < FILL) buf[pos[p] * NL + line] = v[u];
  }
}

// Green items per thread of k_zconv_ct<V, L, B>.
constexpr int zct_ng(int L, int B) {
  return ((L / 2 + 1) * B + (256 / (3 * B)) * 3 * B - 1) / ((256 / (3 * B)) * 3 * B);
}

// Minimum waves per SIMD of k_zconv_ct<V, L, B>: FP32 tiles of 8 lines or
// more at L <= 128 are held to 128 VGPRs, four waves instead of three (134
// unbounded, 8 values spilled), and the FP32 width-5 tile at L = 128 (the
// batched launches of 128^3 grids) to 96, five waves instead of four (98
// unbounded, no spill): the 256-RHS batch went from 1270 to 1290 RHS/s with
// the first and from 1371 to 1405 with width 5 and the second.  At L = 200
// the bound that would add a wave spills 14-46 values.
template <typename V, int L, int B> constexpr int zct_min_waves() {
  return sizeof(V) != 8 ? 1 : B >= 8 && L <= 128 ? 4 : (L == 128 && B == 5) ? 5 : 1;
}

// k_zconv for L3 = L, tile width B; launched with (256 / 3B) * 3B threads.
// Requires 2 n_in <= L and 2 n_out <= L (true for every box: L >= 2n).
template <typename V, int L, int B>
__global__ void __launch_bounds__(256, (zct_min_waves<V, L, B>())) k_zconv_ct(SArgs<V> a) {
  constexpr int NL = 3 * B, JT = 256 / NL, NTHR = NL * JT, F3 = L / 2 + 1;
  constexpr int NG = zct_ng(L, B);
  constexpr PlanC P = plan_c(L);
  static_assert(P.nst > 0 && NG >= 1 && NG <= kMaxG, "unsupported compile-time z plan");
  int lg, rhs;
  rhs_group_tile(lg, rhs);
  if (a.skip && a.skip[rhs * a.skip_stride]) return;
  const V* const zsrc = a.src + rhs * a.zs_src;
  V* const zdst = a.dst + rhs * a.zs_dst;
  extern __shared__ double4 smem_d4[];
  V* tws = reinterpret_cast<V*>(smem_d4);
  int* pos = reinterpret_cast<int*>(tws + L);
  V* buf = tile_base(tws, L);
  const int tid = threadIdx.x;
  const int line = tid % NL, jt = tid / NL;
  const int kyf = lg / gridDim.x;  // 0 .. L2/2
  // tiles of a row in the order 0, gx-1, 1, gx-2, ...: kx and its mirror
  // L1 - kx read the same folded Green column and now run side by side
  const int u = lg - kyf * gridDim.x;
  const int kx0 = ((u & 1) ? (int)gridDim.x - 1 - (u >> 1) : (u >> 1)) * B;
  const int nvalid = min(B, a.nbx - kx0) * 3;

  // FP64 at widths up to 5 (latency-bound, 3 blocks per CU by LDS): each
  // thread's rows of a tile line go through registers; the first row's loads
  // are issued before the tables', the mirror row's while the first row is
  // transformed (rows ky and L2 - ky are disjoint, so reading the second
  // before the first is stored in place is safe).  200^3, width 5: 158.7 ->
  // 155.2 us.  These rows keep the load / LDS / first-stage order and the
  // store from LDS: fed from registers to the first stage and stored from the
  // last, the FP64 pass measured 166.8 against 159.7 us.  Wider FP64 tiles
  // and FP32 load each row when they transform it (holding a row in
  // registers cost a wave per SIMD: 128^3 FP64 at width 8, the batch's: 9.9
  // -> 11.4 ms per 256-RHS launch; FP32 200^3: 83.6 -> 88.0 us), straight
  // into the first stage's registers (ct_load_first), and store from the
  // last stage's registers (ct_last_store): FP32 200^3 88.3 -> 85.4 us.  The
  // arithmetic is the same either way (the build contracts multiply-adds
  // only within an expression), so every width gives the same bits.
  // The folded Green entries of this thread's items (kz', kxl): the same for
  // both rows ky' and L2 - ky' (only the y parity sign differs).  FP32 loads
  // them once, before the first row's transform, and keeps them (fewer VGPRs
  // than loading per row, and the loads' latency hides behind the forward
  // FFT: 200^3 85.2 -> 84.1 us); FP64 loads them per row after its forward
  // FFT (64 more bytes per item held across it would cost a wave).
  V gr[NG][6];
  int gx[NG];  // bit0 valid, bit1 x odd-parity sign
  auto load_green = [&]() {
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      const int idx = tid + it * NTHR;
      gx[it] = 0;
      if (idx < F3 * B) {
        const int kzf = idx / B, kxl = idx - kzf * B;
        const int kx = kx0 + kxl;
        if (kx < a.nbx) {
          int sx;
          const int kxf = fold_idx(kx, a.L1, sx);
          const V* g = a.G + ((long long)(kyf * F3 + kzf) * a.F1 + kxf) * 6;
#pragma unroll
          for (int c = 0; c < 6; ++c) gr[it][c] = g[c];
          gx[it] = 1 | ((sx < 0) << 1);
        }
      }
    }
  };
  constexpr bool GONCE = sizeof(V) == 8;
  if constexpr (GONCE) load_green();
  constexpr bool PREF = sizeof(V) == 16 && B <= 5;
  constexpr int FILL = P.rad[0] % 2 == 0 ? L / 2 : L;
  constexpr int NR = PREF ? (FILL + JT - 1) / JT : 1;
  [[maybe_unused]] V nxt[NR];
  const int nrow = (kyf == 0 || 2 * kyf == a.L2) ? 1 : 2;
  if constexpr (PREF)
    fetch_rows<FILL, JT>(nxt, zsrc + (long long)kyf * a.so_src + (long long)kx0 * 3, a.sp_src,
                         a.n_in, line, nvalid, jt);
  load_tables(tws, pos, a.tws, a.pos, L, tid, NTHR);
  lds_sync();
  for (int rr = 0; rr < nrow; ++rr) {
    const int ky = rr == 0 ? kyf : a.L2 - kyf;
    const bool sy = rr == 1;  // y odd-parity sign for the mirror row
    if constexpr (PREF) {
      put_rows<FILL, JT, NL>(buf, pos, nxt, line, jt);
      lds_sync();
      if (rr + 1 < nrow)
        fetch_rows<FILL, JT>(nxt, zsrc + (long long)(a.L2 - kyf) * a.so_src + (long long)kx0 * 3,
                             a.sp_src, a.n_in, line, nvalid, jt);
      ct_dit<L, NL, JT, +1, 0>(buf, tws, line, jt);
    } else {
      ct_load_first<L, NL, JT, +1>(buf, pos, zsrc + (long long)ky * a.so_src + (long long)kx0 * 3,
                                   a.sp_src, a.n_in, line, nvalid, jt);
      ct_dit<L, NL, JT, +1, 1>(buf, tws, line, jt);
    }
    if constexpr (!GONCE) load_green();
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      if (!(gx[it] & 1)) continue;
      const int idx = tid + it * NTHR;
      const int kzf = idx / B, kxl = idx - kzf * B;
      const bool sxn = gx[it] & 2;
      sym_product(buf + kzf * NL + kxl * 3, gr[it], sxn != sy, sxn, sy);
      if (kzf != 0 && 2 * kzf != L)
        sym_product(buf + (L - kzf) * NL + kxl * 3, gr[it], sxn != sy, !sxn, !sy);
    }
    lds_sync();
    if constexpr (PREF) {
      ct_dif<L, NL, JT, -1, P.nst - 1>(buf, tws, line, jt);
      if (line < nvalid) {
        V* d = zdst + (long long)ky * a.so_dst + (long long)kx0 * 3 + line + (long long)jt * a.sp_dst;
        const long long step = (long long)JT * a.sp_dst;
        for (int p = jt; p < a.n_out; p += JT, d += step) *d = buf[pos[p] * NL + line];
      }
      lds_sync();  // the next row reuses the LDS tile
    } else {  // the last stage fused with the store (its barrier ends the row)
      ct_dif<L, NL, JT, -1, P.nst - 1, V, 1>(buf, tws, line, jt);
      ct_last_store<L, NL, JT, -1>(buf, pos, zdst + (long long)ky * a.so_dst + (long long)kx0 * 3,
                                   a.sp_dst, a.n_out, line, nvalid, jt);
    }
  }
}

// The compile-time z convolutions built into the library: (L, B) of the
// 40^3, 128^3, 192^3 and 200^3 grids at their default tile widths (L = 40:
// width 5 for one right-hand side, 27 for batches) and at the neighbouring
// widths the autotuner may pick (4 to 8 lines).
template <typename V> using ZconvKernel = void (*)(SArgs<V>);
template <typename V> ZconvKernel<V> zconv_ct_for(int L, int B) {
  if (L == 200 && B == 4) return k_zconv_ct<V, 200, 4>;
  if (L == 200 && B == 5) return k_zconv_ct<V, 200, 5>;
  if (L == 200 && B == 6) return k_zconv_ct<V, 200, 6>;
  if (L == 200 && B == 8) return k_zconv_ct<V, 200, 8>;
  if (L == 128 && B == 4) return k_zconv_ct<V, 128, 4>;
  if (L == 128 && B == 5) return k_zconv_ct<V, 128, 5>;
  if (L == 128 && B == 6) return k_zconv_ct<V, 128, 6>;
  if (L == 128 && B == 8) return k_zconv_ct<V, 128, 8>;
  if (L == 192 && B == 4) return k_zconv_ct<V, 192, 4>;
  if (L == 192 && B == 5) return k_zconv_ct<V, 192, 5>;
  if (L == 192 && B == 6) return k_zconv_ct<V, 192, 6>;
  if (L == 40 && B == 5) return k_zconv_ct<V, 40, 5>;
  if (L == 40 && B == 27) return k_zconv_ct<V, 40, 27>;
  return nullptr;
}

// ---------------------------------------------------------------------------
// x-axis pass: lines are contiguous in HBM (ld elements apart).  MODE 0 plain,
// MODE 1 scatter from the dipole list (alpha applied), MODE 2 gather to the
// dipole list: out = b - v/M.  Box line g covers cells i + nx*g.
// ---------------------------------------------------------------------------
template <typename V> struct XArgs {
  const V* src;
  V* dst;
  const V* tw;
  const V* tws;  // stage-major twiddles (k_xpass_ct)
  const int* pos;
  AxisPlan ax;
  int n_in, n_out;
  int nlines;  // total lines
  int B;       // lines per tile
  long long ld_src, ld_dst;
  // MODE 1 / 2
  const int32_t* tab;  // box -> 1-based dipole (0 = empty); null = dense box
  int nx;
  const V* xin;        // MODE 1: x (3 per dipole); MODE 2: b
  const V* alpha;      // MODE 1
  int alpha_mode;
  V* xout;             // MODE 2
  decltype(V::x) inv_m;
  // right-hand side blockIdx.z: element offsets per RHS (grids, dipole
  // vectors) and a per-RHS skip flag as in SArgs
  long long zs_src, zs_dst, zs_x;
  const int* skip;
  int skip_stride;
  int nthr;  // packed launches: threads of the tile shape
  // sparse box (null: every line): rows[k] bit j is set when box line
  // g = j + ny k holds a dipole (ny <= 64).  P1 writes only those lines of
  // A and P5 reads only those; a tile without any is skipped.
  const unsigned long long* rows;
  int ny;
};

// Bit l of the result: line g0 + l (l < nl <= 64) holds a dipole (all ones
// without a row mask).  g0 is uniform over the block, so this is scalar work.
__device__ __forceinline__ unsigned long long tile_lines(const unsigned long long* rows, int ny,
                                                         int g0, int nl) {
  if (!rows) return ~0ull;
  unsigned long long tm = 0;
  int k = g0 / ny, j = g0 - k * ny;
  unsigned long long w = rows[k];
  for (int l = 0; l < nl; ++l) {
    tm |= ((w >> j) & 1ull) << l;
    if (++j == ny && l + 1 < nl) {
      j = 0;
      w = rows[++k];
    }
  }
  return tm;
}

template <typename V>
__device__ __forceinline__ void alpha_apply(const V* __restrict__ al, int mode, long long dip, V& e0,
                                            V& e1, V& e2) {
  if (mode == kAlphaScalar) {
    const V s = al[dip];
    e0 = cmul(s, e0);
    e1 = cmul(s, e1);
    e2 = cmul(s, e2);
  } else if (mode == kAlphaTensor) {
    const V* m = al + dip * 9;
    const V f0 = cadd(cadd(cmul(m[0], e0), cmul(m[1], e1)), cmul(m[2], e2));
    const V f1 = cadd(cadd(cmul(m[3], e0), cmul(m[4], e1)), cmul(m[5], e2));
    const V f2 = cadd(cadd(cmul(m[6], e0), cmul(m[7], e1)), cmul(m[8], e2));
    e0 = f0;
    e1 = f1;
    e2 = f2;
  }
}

// Dipole index of cell i of box line g, or -1.
__device__ __forceinline__ long long dip_of(const int32_t* __restrict__ tab, int nx, int g, int i) {
  const long long cell = (long long)i + (long long)nx * g;
  return tab ? (long long)tab[cell] - 1 : cell;
}

// (line, point) = divmod(e, n) for e = e0 + k * step, advanced without
// division: step = dl * n + di.
struct DivStep {
  int l, i;
  int dl, di, n;
  __device__ __forceinline__ DivStep(int e0, int step, int n_) : n(n_) {
    l = e0 / n;
    i = e0 - l * n;
    dl = step / n;
    di = step - dl * n;
  }
  __device__ __forceinline__ void next() {
    l += dl;
    i += di;
    if (i >= n) {
      i -= n;
      ++l;
    }
  }
};

// A complex value of V in FP64 (the fused inner products are summed in FP64).
template <typename V> __device__ __forceinline__ double2 to_d(V v) {
  return make_double2((double)v.x, (double)v.y);
}
// An FP64 complex value rounded to V (as the Krylov kernels store it).
template <typename V> __device__ __forceinline__ V from_d(double2 v) {
  V o;
  o.x = (decltype(o.x))v.x;
  o.y = (decltype(o.y))v.y;
  return o;
}

// Forward (SGN > 0): DIT from a permuted load, first stage pruned when
// 2 n_in <= L.  Inverse (SGN < 0): DIF from a natural load, last stage pruned
// when 2 n_out <= L, output read through pos.
//
// xpass_tile: the tile of box lines g0 .. g0 + B - 1, nthr threads (PK: as
// strided_tile).
// MODE 1: P1 (scatter alpha x, FFT-x); MODE 3: P1 that first forms
// x = p = r + beta (p - u) and stores it over p (KFuse p1 = 1, gpbicg.f:161-169).
// MODE 2 / 5 / 6 / 7: P5 (IFFT-x, gather y = b - v/M); 5, 6 and 7 also form
// the inner products of y (KFuse p5 = 1, 2, 3).  MODE 0: plain x-FFT of full
// lines.  f and zo (the RHS's element offset in the Krylov vectors) are used
// by modes 3, 5, 6 and 7 only.
template <typename V, int NC, int SGN, int MODE, bool PK, int LC = 0, int BC = 0>
__device__ __forceinline__ void xpass_tile(const XArgs<V>& a, const V* zsrc, V* zdst,
                                           const V* zxin, V* zxout, int g0, int tid0, int nthr,
                                           const KFuse<V>* f = nullptr, long long zo = 0) {
  // LC > 0: compile-time line length LC and tile width BC (k_xpass_ct), FFT
  // stages of the z pass on the axis's stage-major twiddles; launched with
  // (256 / (BC NC)) * BC NC threads, 2 n_in <= LC (P1) / 2 n_out <= LC (P5)
  constexpr bool CT = LC > 0;
  constexpr bool kP5 = MODE == 2 || MODE == 5 || MODE == 6 || MODE == 7;
  // inner products formed
  constexpr int kND = MODE == 5 ? 1 : MODE == 6 ? 5 : MODE == 7 ? 8 : 0;
  constexpr bool DIT = SGN > 0;
  extern __shared__ double4 smem_d4[];
  V* tw = reinterpret_cast<V*>(smem_d4);
  const int L = CT ? LC : a.ax.L;
  const int BT = CT ? BC : a.B;
  int* pos = reinterpret_cast<int*>(tw + L);
  V* buf = tile_base(tw, L);
  const int NL = BT * NC;
  const int NLS = NC == 3 ? x_stride(NL) : NL;  // LDS row stride
  const int tid = PK && tid0 >= nthr ? kIdle : tid0;
  const int line = tid0 % NL, jt = tid / NL, JT = nthr / NL;
  const int nl = min(BT, a.nlines - g0);
  // lines of the tile that hold dipoles; a tile without any writes nothing
  // (P1) or gathers nothing (P5), and only the gathering pass with inner
  // products still runs, for its zero block partials
  const unsigned long long tm = tile_lines(a.rows, a.ny, g0, nl);
  if (tm == 0 && !(kP5 && kND > 0)) return;
  const bool half_in = DIT && 2 * a.n_in <= L;
  const bool half_out = !DIT && 2 * a.n_out <= L;
  const int fill = DIT ? (CT ? ct_fill(LC) : rows_to_fill(a.ax, half_in)) : L;

  load_tables(tw, pos, CT ? a.tws : a.tw, a.pos, L, tid, nthr);
  lds_sync();
  {
    constexpr int U = MODE == 1 || MODE == 3 ? kBatch / 2 : kBatch / NC;
    double2 beta = make_double2(0.0, 0.0);
    if constexpr (MODE == 3) beta = f->st[blockIdx.z].beta;
    DivStep ix(tid, nthr, fill);  // (tile line l, point i)
    for (int e0 = tid; e0 < BT * fill; e0 += U * nthr) {
      int ll[U], ii[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ll[u] = ix.l;
        ii[u] = ix.i;
        ix.next();
      }
      V x[U][NC];
      if constexpr (MODE == 1 || MODE == 3) {
        long long dip[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          dip[u] = -1;
          if (ll[u] < nl && ii[u] < a.n_in && row_on(tm, ll[u]))
            dip[u] = dip_of(a.tab, a.nx, g0 + ll[u], ii[u]);
        }
        if constexpr (MODE == 3) {
          // every load of the batch before the stores over p
          V rv[U][3], uv[U][3];
#pragma unroll
          for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int c = 0; c < NC; ++c) x[u][c] = czero<V>();
            if (dip[u] >= 0) {
              const long long e = zo + dip[u] * 3;
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                x[u][c] = zxin[dip[u] * 3 + c];
                rv[u][c] = f->r[e + c];
                uv[u][c] = f->u[e + c];
              }
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (dip[u] >= 0) {
              const long long e = zo + dip[u] * 3;
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                x[u][c] = from_d<V>(zadd(to_d(rv[u][c]), zmul(beta, zsub(to_d(x[u][c]), to_d(uv[u][c])))));
                f->p[e + c] = x[u][c];
              }
              alpha_apply(a.alpha, a.alpha_mode, dip[u], x[u][0], x[u][1], x[u][2]);
            }
        } else {
#pragma unroll
          for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int c = 0; c < NC; ++c) x[u][c] = czero<V>();
            if (dip[u] >= 0) {
              const V* xp = zxin + dip[u] * 3;
              x[u][0] = xp[0];
              x[u][1] = xp[1];
              x[u][2] = xp[2];
              alpha_apply(a.alpha, a.alpha_mode, dip[u], x[u][0], x[u][1], x[u][2]);
            }
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
          for (int c = 0; c < NC; ++c) x[u][c] = czero<V>();
          if (ll[u] < nl && ii[u] < a.n_in && row_on(tm, ll[u])) {
            const V* sp = zsrc + (long long)(g0 + ll[u]) * a.ld_src + (long long)ii[u] * NC;
#pragma unroll
            for (int c = 0; c < NC; ++c) x[u][c] = sp[c];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (ll[u] < BT) {
          V* d = buf + (DIT ? pos[ii[u]] : ii[u]) * NLS + ll[u] * NC;
#pragma unroll
          for (int c = 0; c < NC; ++c) d[c] = x[u][c];
        }
      }
    }
  }
  lds_sync();

  if (tm == 0) {
    // no dipole in the tile: the gathering pass's partials only (zeros)
  } else if constexpr (CT) {
    // the stages take the row stride as their NL (it only places rows)
    constexpr int NLC = BC * NC, JTC = 256 / NLC, NLSC = NC == 3 ? x_stride(NLC) : NLC;
    if constexpr (DIT)
      ct_dit<LC, NLSC, JTC, SGN, 0>(buf, tw, line, jt);
    else
      ct_dif<LC, NLSC, JTC, SGN, plan_c(LC).nst - 1>(buf, tw, line, jt);
  } else if (DIT) {
    tile_fft_dit<SGN>(buf, tw, a.ax, NLS, line, jt, JT, half_in);
  } else {
    tile_fft_dif<SGN>(buf, tw, a.ax, NLS, line, jt, JT, half_out);
  }

  if constexpr (kP5) {
    // gathers per batch: 4, or 2 for the eight products of MODE 7 (144 -> fewer
    // VGPRs: four gathers in flight took it to 3 waves per SIMD)
    constexpr int U = MODE == 7 ? kBatch / 4 : kBatch / 2;
    double2 acc[kND > 0 ? kND : 1];
#pragma unroll
    for (int q = 0; q < (kND > 0 ? kND : 1); ++q) acc[q] = make_double2(0.0, 0.0);
    DivStep ix(tid, nthr, a.n_out);
    for (int e0 = tid; e0 < BT * a.n_out; e0 += U * nthr) {
      int ll[U], ii[U];
      long long dip[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ll[u] = ix.l;
        ii[u] = ix.i;
        ix.next();
        dip[u] = -1;
        if (ll[u] < nl && row_on(tm, ll[u])) dip[u] = dip_of(a.tab, a.nx, g0 + ll[u], ii[u]);
      }
      V b[U][3];
      // r0* (MODE 5) or y (MODE 6, 7), and r0* (MODE 7), loaded before the
      // stores of y
      V q[kND > 0 ? U : 1][3];
      V q2[MODE == 7 ? U : 1][3];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (dip[u] >= 0) {
          const V* bp = zxin + dip[u] * 3;
          b[u][0] = bp[0];
          b[u][1] = bp[1];
          b[u][2] = bp[2];
          if constexpr (kND > 0) {
            const V* qp = (MODE == 5 ? f->r0s : f->y) + zo + dip[u] * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) q[u][c] = qp[c];
          }
          if constexpr (MODE == 7) {
            const V* qp = f->r0s + zo + dip[u] * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) q2[u][c] = qp[c];
          }
        }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (dip[u] >= 0) {
          const V* v = buf + (DIT ? ii[u] : pos[ii[u]]) * NLS + ll[u] * 3;
          V* y = zxout + dip[u] * 3;
          V yo[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            yo[c] = csub(b[u][c], cscal(v[c], a.inv_m));
            y[c] = yo[c];
          }
          if constexpr (MODE == 5) {  // <r0*, Ap> (gpbicg.f:174-181)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[0] = zadd(acc[0], zmul(to_d(q[u][c]), to_d(yo[c])));
          } else if constexpr (MODE == 6 || MODE == 7) {  // gpbicg.f:222-253
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const double2 At = to_d(yo[c]), yv = to_d(q[u][c]), t = to_d(b[u][c]);
              const double2 cAt = zconj(At), cy = zconj(yv);
              acc[0] = zadd(acc[0], zmul(cAt, At));
              acc[1] = zadd(acc[1], zmul(cy, yv));
              acc[2] = zadd(acc[2], zmul(cAt, yv));
              acc[3] = zadd(acc[3], zmul(cAt, t));
              acc[4] = zadd(acc[4], zmul(cy, t));
              if constexpr (MODE == 7) {  // r0* against t, y and At: <r0*, r_new>
                const double2 r0 = to_d(q2[u][c]);
                acc[5] = zadd(acc[5], zmul(r0, t));
                acc[6] = zadd(acc[6], zmul(r0, yv));
                acc[7] = zadd(acc[7], zmul(r0, At));
              }
            }
          }
        }
    }
    if constexpr (kND > 0) {
      // one partial per block and inner product, in a fixed order: every
      // thread's value in LDS, then lane l of a full wavefront sums the values
      // of threads l, l + 64, l + 128, l + 192 and the wavefront shuffles.
      // The tile is free once every thread has gathered, so the products
      // are staged there as many at a time as fit (two barriers per group),
      // and the full wavefronts take one product each.
      const int nt = (int)blockDim.x, nfull = nt >> 6;
      const size_t dyn = ((size_t)L * sizeof(V) + (size_t)L * sizeof(int) + 15 & ~(size_t)15) +
                         (size_t)L * NLS * sizeof(V);
      const int per = max(1, min(kND, (int)(dyn / (sizeof(double2) * (size_t)nt))));
      double2* red = reinterpret_cast<double2*>(smem_d4);
      double2* part = f->part + (long long)blockIdx.z * kND * f->pstride + blockIdx.x;
      const int lane = tid0 & 63, wv = tid0 >> 6;
      for (int q0 = 0; q0 < kND; q0 += per) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kND; ++q)
          if (q >= q0 && q < q0 + per) red[(q - q0) * nt + tid0] = acc[q];
        __syncthreads();
        for (int q = q0 + wv; q < min(kND, q0 + per) && wv < nfull; q += nfull) {
          const double2* rq = red + (q - q0) * nt;
          double2 sum = rq[lane];
          for (int k = 1; k < 4; ++k)
            if (lane + 64 * k < nt) sum = zadd(sum, rq[lane + 64 * k]);
          for (int off = 32; off > 0; off >>= 1) {
            sum.x += __shfl_down(sum.x, off, 64);
            sum.y += __shfl_down(sum.y, off, 64);
          }
          if (lane == 0) part[(long long)q * f->pstride] = sum;
        }
      }
    }
  } else {
    DivStep ix(tid, nthr, a.n_out);
    for (int e = tid; e < BT * a.n_out; e += nthr, ix.next()) {
      if (ix.l >= nl || !row_on(tm, ix.l)) continue;
      const V* v = buf + (DIT ? ix.i : pos[ix.i]) * NLS + ix.l * NC;
      V* d = zdst + (long long)(g0 + ix.l) * a.ld_dst + (long long)ix.i * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) d[c] = v[c];
    }
  }
}

template <typename V, int NC, int SGN, int MODE>
__global__ void __launch_bounds__(256) k_xpass(XArgs<V> a) {
  if (a.skip && a.skip[blockIdx.z * a.skip_stride]) return;
  // a stays unmodified (kernel-argument memory); per-RHS offsets go to locals
  xpass_tile<V, NC, SGN, MODE, false>(
      a, a.src ? a.src + blockIdx.z * a.zs_src : nullptr,
      a.dst ? a.dst + blockIdx.z * a.zs_dst : nullptr,
      a.xin ? a.xin + blockIdx.z * a.zs_x : nullptr,
      a.xout ? a.xout + blockIdx.z * a.zs_x : nullptr, xcd_remap(blockIdx.x, gridDim.x) * a.B,
      threadIdx.x, blockDim.x);
}

// P1 with the GPBICG p update (MODE 3) or P5 with GPBICG inner products
// (MODE 5, 6) fused in (KFuse; one context, R right-hand sides zs_x elements
// apart).
template <typename V, int NC, int SGN, int MODE>
__global__ void __launch_bounds__(256) k_xpass_f(XArgs<V> a, KFuse<V> f) {
  if (a.skip && a.skip[blockIdx.z * a.skip_stride]) return;
  xpass_tile<V, NC, SGN, MODE, false>(
      a, a.src ? a.src + blockIdx.z * a.zs_src : nullptr,
      a.dst ? a.dst + blockIdx.z * a.zs_dst : nullptr,
      a.xin ? a.xin + blockIdx.z * a.zs_x : nullptr,
      a.xout ? a.xout + blockIdx.z * a.zs_x : nullptr, xcd_remap(blockIdx.x, gridDim.x) * a.B,
      threadIdx.x, blockDim.x, &f, (long long)blockIdx.z * a.zs_x);
}

// x passes with a compile-time line length L and tile width B (P1: MODE 1;
// P5: MODE 2, and 5 / 6 with the GPBICG products): the z pass's FFT stages,
// twiddle offsets and trip counts as constants.  f is used by modes 5 and 6.
template <typename V, int NC, int SGN, int MODE, int L, int B>
__global__ void __launch_bounds__(256) k_xpass_ct(XArgs<V> a, KFuse<V> f) {
  if (a.skip && a.skip[blockIdx.z * a.skip_stride]) return;
  xpass_tile<V, NC, SGN, MODE, false, L, B>(
      a, a.src ? a.src + blockIdx.z * a.zs_src : nullptr,
      a.dst ? a.dst + blockIdx.z * a.zs_dst : nullptr,
      a.xin ? a.xin + blockIdx.z * a.zs_x : nullptr,
      a.xout ? a.xout + blockIdx.z * a.zs_x : nullptr, xcd_remap(blockIdx.x, gridDim.x) * B,
      threadIdx.x, blockDim.x, &f, (long long)blockIdx.z * a.zs_x);
}

template <typename V> using XpassKernel = void (*)(XArgs<V>, KFuse<V>);
template <typename V, int SGN, int MODE, int L>
static XpassKernel<V> xpass_ct_width(int B) {
  switch (B) {
    case 2: return k_xpass_ct<V, 3, SGN, MODE, L, 2>;
    case 3: return k_xpass_ct<V, 3, SGN, MODE, L, 3>;
    case 4: return k_xpass_ct<V, 3, SGN, MODE, L, 4>;
    case 5: return k_xpass_ct<V, 3, SGN, MODE, L, 5>;
    case 8: if constexpr (L == 128) return k_xpass_ct<V, 3, SGN, MODE, L, 8>; else return nullptr;
    default: return nullptr;
  }
}
// The compile-time x passes built into the library: the x lengths of the
// 128^3, 192^3 and 200^3 grids at widths 2 to 5 (and 8 at 128, the width of
// batched launches); nullptr elsewhere.
template <typename V, int SGN, int MODE> XpassKernel<V> xpass_ct_for(int L, int B) {
  if (L == 200) return xpass_ct_width<V, SGN, MODE, 200>(B);
  if (L == 128) return xpass_ct_width<V, SGN, MODE, 128>(B);
  if (L == 192) return xpass_ct_width<V, SGN, MODE, 192>(B);
  return nullptr;
}

// ---------------------------------------------------------------------------
// Packed launches: one grid over the tiles of P independent problems
// (members) of possibly different boxes, FFT lengths, Green tensors, dipole
// maps and polarisabilities.  Member m owns blocks [boff[m], boff[m+1]) and
// the argument block av[m] (built as for its own launch); its dipole vectors
// start at element zs_x of its P1 / P5 blocks (P1: alpha x input; P5: b
// input, y output), the members' vectors lying back to back.  The
// block size is 256 for all members; a member whose tile shape uses fewer
// threads leaves the rest idle.  A member's skip flag drops its blocks.
// ---------------------------------------------------------------------------
template <typename V> struct PkArgs {
  const void* av;
  const int* boff;
  int P;
  const V* x;
  V* y;
  const int* skip;
  int skip_stride;
};

// m with boff[m] <= b < boff[m+1] (block-uniform: scalar loads)
__device__ __forceinline__ int pk_member(const int* __restrict__ boff, int P, int b) {
  int lo = 0, hi = P;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (boff[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Argument block m, read through the constant address space: the blob does
// not change during a launch, so its fields are scalar loads (neither copied
// to the stack nor reloaded into vector registers after the tile's stores).
template <typename A> __device__ __forceinline__ const A& pk_args(const void* av, int m) {
  using CA = __attribute__((address_space(4))) const A;
  const CA* p = (const CA*)(av) + m;
  return *(const A*)p;
}

template <typename V, int NC, int SGN, int MODE>
__global__ void __launch_bounds__(256) k_xpass_pk(PkArgs<V> k) {
  const int m = pk_member(k.boff, k.P, blockIdx.x);
  if (k.skip && k.skip[m * k.skip_stride]) return;
  const XArgs<V>& a = pk_args<XArgs<V>>(k.av, m);
  const long long vo = a.zs_x;
  xpass_tile<V, NC, SGN, MODE, true>(a, a.src, a.dst, MODE ? k.x + vo : nullptr,
                                     MODE == 2 ? k.y + vo : nullptr,
                                     (blockIdx.x - k.boff[m]) * a.B, threadIdx.x, a.nthr);
}

template <typename V, int NC, int SGN>
__global__ void __launch_bounds__(256) k_strided_pk(PkArgs<V> k) {
  const int m = pk_member(k.boff, k.P, blockIdx.x);
  if (k.skip && k.skip[m * k.skip_stride]) return;
  const SArgs<V>& a = pk_args<SArgs<V>>(k.av, m);
  strided_tile<V, NC, SGN, true>(a, a.src, a.dst, blockIdx.x - k.boff[m], a.gx, threadIdx.x,
                                 a.nthr);
}

template <typename V, int NG>
__global__ void __launch_bounds__(256) k_zconv_pk(PkArgs<V> k) {
  const int m = pk_member(k.boff, k.P, blockIdx.x);
  if (k.skip && k.skip[m * k.skip_stride]) return;
  const SArgs<V>& a = pk_args<SArgs<V>>(k.av, m);
  zconv_tile<V, NG, true>(a, a.src, a.dst, blockIdx.x - k.boff[m], a.gx, threadIdx.x, a.nthr);
}

// ---------------------------------------------------------------------------
// Green tensor (FP64): fill one component on the circulant grid
// (calculgreenfft.f:22-71), near-cell Simpson integration
// (propaesplibreintsim.f:40-101), fold to one octant.
// ---------------------------------------------------------------------------

// T(R) for R = -(ox,oy,oz)*d (espace_libre.f:2-43 with the observation at the
// origin); comps xx,xy,xz,yy,yz,zz.
__device__ __forceinline__ void green_T(double rx, double ry, double rz, double k0, double2 T[6]) {
  const double r2 = rx * rx + ry * ry + rz * rz;
  if (r2 == 0.0) {
    for (int c = 0; c < 6; ++c) T[c] = make_double2(0.0, 0.0);
    return;
  }
  const double R = sqrt(r2);
  double sn, cs;
  sincos(k0 * R, &sn, &cs);
  // c1 = (1/R - i k0)/r2, c2 = k0^2/R (real)
  const double c1r = 1.0 / (R * r2), c1i = -k0 / r2;
  const double c2 = k0 * k0 / R;
  const double rv[3] = {rx, ry, rz};
  const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double rr = rv[ia[c]] * rv[ib[c]] / r2;
    const double id = (ia[c] == ib[c]) ? 1.0 : 0.0;
    const double f = 3.0 * rr - id, g = id - rr;
    const double ar = f * c1r + g * c2, ai = f * c1i;
    T[c] = make_double2(ar * cs - ai * sn, ar * sn + ai * cs);
  }
}

// d T / d r_dd at r (the observation minus the source: propesplibdermat.f
// restated in Cartesian form).  With T = A(R) I + B(R) r r^T,
//   A = e^{ikR} (k^2/R - 1/R^3 + ik/R^2), B = e^{ikR} (3/R^5 - 3ik/R^4 - k^2/R^3),
//   d T_ab = delta_ab A' r_dd / R + (delta_a,dd r_b + delta_b,dd r_a) B + r_a r_b B' r_dd / R,
//   A' = e^{ikR} (ik^3/R - 2k^2/R^2 - 3ik/R^3 + 3/R^4),
//   B' = e^{ikR} (-15/R^6 + 15ik/R^5 + 6k^2/R^4 - ik^3/R^3);  0 at r = 0.
__device__ __forceinline__ void green_dT(double rx, double ry, double rz, double k, int dd,
                                         double2 T[6]) {
  const double r2 = rx * rx + ry * ry + rz * rz;
  if (r2 == 0.0) {
    for (int c = 0; c < 6; ++c) T[c] = make_double2(0.0, 0.0);
    return;
  }
  const double R = sqrt(r2), iR = 1.0 / R, k2 = k * k;
  const double iR2 = iR * iR, iR3 = iR2 * iR, iR4 = iR2 * iR2, iR5 = iR4 * iR, iR6 = iR3 * iR3;
  double sn, cs;
  sincos(k * R, &sn, &cs);
  const double2 e = make_double2(cs, sn);
  auto cm = [](double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
  // T = A I + B r r^T; its derivative needs B, A' and B' (A itself drops out)
  const double2 B = cm(e, make_double2(3.0 * iR5 - k2 * iR3, -3.0 * k * iR4));
  const double2 Ap = cm(e, make_double2(-2.0 * k2 * iR2 + 3.0 * iR4, k2 * k * iR - 3.0 * k * iR3));
  const double2 Bp = cm(e, make_double2(-15.0 * iR6 + 6.0 * k2 * iR4, 15.0 * k * iR5 - k2 * k * iR3));
  const double rv[3] = {rx, ry, rz};
  const double rd = rv[dd] * iR;
  const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int a = ia[c], b = ib[c];
    const double fa = (a == b ? rd : 0.0);
    const double fb = (a == dd ? rv[b] : 0.0) + (b == dd ? rv[a] : 0.0);
    const double fc = rv[a] * rv[b] * rd;
    T[c] = make_double2(fa * Ap.x + fb * B.x + fc * Bp.x, fa * Ap.y + fb * B.y + fc * Bp.y);
  }
}

__device__ __forceinline__ int circ_off(int idx, int n, int L, bool& ok) {
  if (idx < n) { ok = true; return idx; }
  if (idx > L - n) { ok = true; return idx - L; }
  ok = false;
  return 0;
}

__global__ void k_green_fill(double2* __restrict__ out, int comp, int nx, int ny, int nz, int L1,
                             int L2, int L3, double d, double k0,
                             const double2* __restrict__ near, int nquad) {
  const long long M = (long long)L1 * L2 * L3;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < M;
       idx += (long long)gridDim.x * blockDim.x) {
    const int ii = idx % L1;
    const int jj = (idx / L1) % L2;
    const int kk = idx / ((long long)L1 * L2);
    bool okx, oky, okz;
    const int i = circ_off(ii, nx, L1, okx), j = circ_off(jj, ny, L2, oky),
              k = circ_off(kk, nz, L3, okz);
    double2 v = make_double2(0.0, 0.0);
    if (okx && oky && okz) {
      if (nquad != 0 && i * i + j * j + k * k <= nquad * nquad) {
        const int w = 2 * nquad + 1;
        v = near[(((k + nquad) * w + (j + nquad)) * w + (i + nquad)) * 6 + comp];
      } else {
        double2 T[6];
        green_T(-(double)i * d, -(double)j * d, -(double)k * d, k0, T);
        v = T[comp];
      }
    }
    out[idx] = v;
  }
}

// All six components of the circulant fill at once, component c at out +
// c M: one T evaluation per grid point where k_green_fill makes six, and one
// launch where it takes six.  The values are k_green_fill's.
__global__ void k_green_fill6(double2* __restrict__ out, int nx, int ny, int nz, int L1, int L2,
                              int L3, double d, double k0, const double2* __restrict__ near,
                              int nquad) {
  const long long M = (long long)L1 * L2 * L3;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < M;
       idx += (long long)gridDim.x * blockDim.x) {
    const int ii = idx % L1;
    const int jj = (idx / L1) % L2;
    const int kk = idx / ((long long)L1 * L2);
    bool okx, oky, okz;
    const int i = circ_off(ii, nx, L1, okx), j = circ_off(jj, ny, L2, oky),
              k = circ_off(kk, nz, L3, okz);
    double2 T[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) T[c] = make_double2(0.0, 0.0);
    if (okx && oky && okz) {
      if (nquad != 0 && i * i + j * j + k * k <= nquad * nquad) {
        const int w = 2 * nquad + 1;
        const double2* nv = near + (((k + nquad) * w + (j + nquad)) * w + (i + nquad)) * 6;
#pragma unroll
        for (int c = 0; c < 6; ++c) T[c] = nv[c];
      } else {
        green_T(-(double)i * d, -(double)j * d, -(double)k * d, k0, T);
      }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) out[c * M + idx] = T[c];
  }
}

// One component of d T / d r_dd on the circulant grid at r = +offset * d
// (derivativefield2.f:22-60: the observation minus the source, point values
// only, zero on the planes at index n).
__global__ void k_green_dfill(double2* __restrict__ out, int comp, int dd, int nx, int ny, int nz,
                              int L1, int L2, int L3, double d, double k0) {
  const long long M = (long long)L1 * L2 * L3;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < M;
       idx += (long long)gridDim.x * blockDim.x) {
    const int ii = idx % L1;
    const int jj = (idx / L1) % L2;
    const int kk = idx / ((long long)L1 * L2);
    bool okx, oky, okz;
    const int i = circ_off(ii, nx, L1, okx), j = circ_off(jj, ny, L2, oky),
              k = circ_off(kk, nz, L3, okz);
    double2 v = make_double2(0.0, 0.0);
    if (okx && oky && okz) {
      double2 T[6];
      green_dT((double)i * d, (double)j * d, (double)k * d, k0, dd, T);
      v = T[comp];
    }
    out[idx] = v;
  }
}

// Simpson partial sums: block (i-plane, cell).  Nodes x_a = -d/2 + h*a of the
// cube centred on the observation point (origin), integrand T(x - o*d).
__global__ void __launch_bounds__(256) k_green_near_partial(double2* __restrict__ part,
                                                            const int* __restrict__ offs,
                                                            double d, double k0) {
  constexpr int NS = 128;
  const int ip = blockIdx.x;  // 0..NS
  const int cell = blockIdx.y;
  const int ox = offs[cell * 3 + 0], oy = offs[cell * 3 + 1], oz = offs[cell * 3 + 2];
  const double h = d / NS;
  const double A = -0.5 * d;
  auto wgt = [](int a) { return (a == 0 || a == NS) ? 1.0 : ((a & 1) ? 4.0 : 2.0); };
  double2 acc[6];
  for (int c = 0; c < 6; ++c) acc[c] = make_double2(0.0, 0.0);
  const double wi = wgt(ip);
  const double xa = A + h * ip;
  for (int t = threadIdx.x; t < (NS + 1) * (NS + 1); t += blockDim.x) {
    const int jn = t / (NS + 1), kn = t - jn * (NS + 1);
    const double ya = A + h * jn, za = A + h * kn;
    const double ww = wi * wgt(jn) * wgt(kn);
    double2 T[6];
    green_T(xa - ox * d, ya - oy * d, za - oz * d, k0, T);
    for (int c = 0; c < 6; ++c) {
      acc[c].x += T[c].x * ww;
      acc[c].y += T[c].y * ww;
    }
  }
  __shared__ double red[256][12];
  for (int c = 0; c < 6; ++c) {
    red[threadIdx.x][2 * c] = acc[c].x;
    red[threadIdx.x][2 * c + 1] = acc[c].y;
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 12; ++q) red[threadIdx.x][q] += red[threadIdx.x + s][q];
    __syncthreads();
  }
  if (threadIdx.x < 6)
    part[((long long)cell * (NS + 1) + ip) * 6 + threadIdx.x] =
        make_double2(red[0][2 * threadIdx.x], red[0][2 * threadIdx.x + 1]);
}

// Finish the near cells: sum plane partials, scale, zero off-diagonals along
// axes with |offset| <= d/10, zero the self cell, point value beyond lambda.
__global__ void k_green_near_final(double2* __restrict__ near, const double2* __restrict__ part,
                                   const int* __restrict__ offs, int ncell, int nquad, double d,
                                   double k0, double lim) {
  constexpr int NS = 128;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= ncell) return;
  const int ox = offs[cell * 3 + 0], oy = offs[cell * 3 + 1], oz = offs[cell * 3 + 2];
  const int w = 2 * nquad + 1;
  double2* out = near + ((long long)((oz + nquad) * w + (oy + nquad)) * w + (ox + nquad)) * 6;
  const double rx = ox * d, ry = oy * d, rz = oz * d;
  const double R = sqrt(rx * rx + ry * ry + rz * rz);
  const double dm = 0.1 * d;
  const double lambda = 6.283185307179586476925286766559 / k0;
  if (R <= dm) {
    for (int c = 0; c < 6; ++c) out[c] = make_double2(0.0, 0.0);
    return;
  }
  if (R >= lim * lambda) {
    double2 T[6];
    green_T(-rx, -ry, -rz, k0, T);
    for (int c = 0; c < 6; ++c) out[c] = T[c];
    return;
  }
  double2 s[6];
  for (int c = 0; c < 6; ++c) s[c] = make_double2(0.0, 0.0);
  for (int ip = 0; ip <= NS; ++ip)
    for (int c = 0; c < 6; ++c) {
      const double2 p = part[((long long)cell * (NS + 1) + ip) * 6 + c];
      s[c].x += p.x;
      s[c].y += p.y;
    }
  const double h = d / NS;
  const double sc = (h / 3.0) * (h / 3.0) * (h / 3.0) / (d * d * d);
  const double fx = fabs(rx) <= dm ? 0.0 : 1.0, fy = fabs(ry) <= dm ? 0.0 : 1.0,
               fz = fabs(rz) <= dm ? 0.0 : 1.0;
  const double f[6] = {1.0, fx * fy, fx * fz, 1.0, fy * fz, 1.0};
  for (int c = 0; c < 6; ++c) out[c] = make_double2(s[c].x * sc * f[c], s[c].y * sc * f[c]);
}

// Fold one FFT'd component (full L grid) into the octant array (FP64 + FP32).
__global__ void k_green_fold(double2* __restrict__ g64, float2* __restrict__ g32,
                             const double2* __restrict__ full, int comp, int L1, int L2, int F1,
                             int F2, int F3) {
  const long long n = (long long)F1 * F2 * F3;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int kx = t % F1;
    const int kz = (t / F1) % F3;
    const int ky = t / ((long long)F1 * F3);
    const double2 v = full[((long long)kz * L2 + ky) * L1 + kx];
    g64[t * 6 + comp] = v;
    if (g32) g32[t * 6 + comp] = make_float2((float)v.x, (float)v.y);
  }
}

// Fold all six FFT'd components (component c at full + c M) into the octant
// arrays: the six values of an octant point are written side by side.
__global__ void k_green_fold6(double2* __restrict__ g64, float2* __restrict__ g32,
                              const double2* __restrict__ full, long long M, int L1, int L2,
                              int F1, int F2, int F3) {
  const long long n = (long long)F1 * F2 * F3;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int kx = t % F1;
    const int kz = (t / F1) % F3;
    const int ky = t / ((long long)F1 * F3);
    const long long src = ((long long)kz * L2 + ky) * L1 + kx;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double2 v = full[c * M + src];
      g64[t * 6 + c] = v;
      if (g32) g32[t * 6 + c] = make_float2((float)v.x, (float)v.y);
    }
  }
}

// ---------------------------------------------------------------------------
// Host side: plans, tile shapes and launchers.
// ---------------------------------------------------------------------------

int smooth235_at_least(int n) {
  for (int m = std::max(n, 1);; ++m) {
    int r = m;
    for (int p : {2, 3, 5})
      while (r % p == 0) r /= p;
    if (r == 1) return m;
  }
}

// Radices in execution order: 8s first, then a 4 or 2 for the rest of the
// power of two (8*2 is taken as 4*4), then 5s and 3s.
bool make_axis_plan(int L, AxisPlan* out) {
  AxisPlan a{};
  a.L = L;
  int r = L, n = 0;
  while (r % 8 == 0 && n < kMaxStages) { a.rad[n++] = 8; r /= 8; }
  if (r % 4 == 0 && n < kMaxStages) { a.rad[n++] = 4; r /= 4; }
  if (r % 2 == 0 && n < kMaxStages) {
    if (n > 0 && a.rad[n - 1] == 8 && n + 1 < kMaxStages) {
      a.rad[n - 1] = 4;  // 8*2 -> 4*4: no radix-2 stage (half the butterflies)
      a.rad[n++] = 4;
    } else {
      a.rad[n++] = 2;
    }
    r /= 2;
  }
  while (r % 5 == 0 && n < kMaxStages) { a.rad[n++] = 5; r /= 5; }
  while (r % 3 == 0 && n < kMaxStages) { a.rad[n++] = 3; r /= 3; }
  if (r != 1) return false;
  a.nst = n;
  *out = a;
  return true;
}

bool has_specialized_z(const GridPlan& g, int prec) {
  const int pi = prec == 32 ? 1 : 0;
  if (2 * g.nz > g.L[2]) return false;
  return prec == 32 ? zconv_ct_for<float2>(g.L[2], g.one[pi][2].B) != nullptr
                    : zconv_ct_for<double2>(g.L[2], g.one[pi][2].B) != nullptr;
}

// Largest batch width B (lines per tile = B*NC) such that the tile fits the
// LDS budget and, for the convolution pass, every thread holds at most kMaxG
// Green-tensor items ((L/2+1)*B <= kMaxG*threads).
static TileShape tile_for(const AxisPlan& ax, int NC, int esz, int nbatch, size_t lds_limit,
                          bool conv, bool xpass = false) {
  const int L = ax.L;
  TileShape t{0, 0, 0};  // B = 0: no feasible tile
  for (int B = 1; B <= 64 && B <= std::max(1, nbatch); ++B) {
    const int NL = B * NC;
    if (NL > 256) break;
    const int JT = 256 / NL;
    const int threads = NL * JT;
    const int NLS = xpass ? x_stride(NL) : NL;  // LDS row stride
    const size_t lds = esz == 16 ? tile_lds_bytes<double2>(L, NLS) : tile_lds_bytes<float2>(L, NLS);
    if (lds > lds_limit) break;
    if (conv && (L / 2 + 1) * B > kMaxG * threads) break;
    t = TileShape{B, threads, lds};
  }
  return t;
}

TileShape pass_tile_shape(const GridPlan& g, int pi, int p, int B, size_t lds_limit) {
  static const int axis[5] = {0, 1, 2, 1, 0};  // P1..P5: x, y, z, y, x
  const int esz = pi == 0 ? 16 : 8;
  const int nbatch = (p == 0 || p == 4) ? g.ny * g.nz : g.L[0];
  const bool conv = p == 2;
  const bool xp = p == 0 || p == 4;
  if (B == 0) return tile_for(g.ax[axis[p]], 3, esz, nbatch, lds_limit, conv, xp);
  if (B < 0 || B > std::max(1, nbatch)) return TileShape{0, 0, 0};
  // feasibility is monotone in B: the widest tile up to B is B when B fits
  const TileShape t = tile_for(g.ax[axis[p]], 3, esz, B, lds_limit, conv, xp);
  return t.B == B ? t : TileShape{0, 0, 0};
}

void axis_positions(const AxisPlan& ax, int* pos) {
  // pos[n] for n = sum_s q_s r_{s+1}...r_k, position sum_s q_s m_s
  const int k = ax.nst;
  for (int p = 0; p < ax.L; ++p) {
    int rem = p, n = 0;
    for (int s = 0; s < k; ++s) {
      const int q = rem % ax.rad[s];
      rem /= ax.rad[s];
      int P = 1;
      for (int t = s + 1; t < k; ++t) P *= ax.rad[t];
      n += q * P;
    }
    pos[n] = p;
  }
}

void stage_twiddles(const AxisPlan& ax, double* re, double* im) {
  int m = 1;
  for (int s = 0; s < ax.nst; ++s) {
    const int R = ax.rad[s], tws = ax.L / (m * R);
    for (int j = 0; j < m; ++j)
      for (int q = 1; q < R; ++q) {
        // the same exponent e = j q tws (< L) as the natural table tw[e]
        const double ang = 6.283185307179586476925286766559 * (double)(j * q * tws) / (double)ax.L;
        const int i = m - 1 + j * (R - 1) + q - 1;
        re[i] = std::cos(ang);
        im[i] = std::sin(ang);
      }
    m *= R;
  }
}

void axis_tables(const AxisPlan& ax, double2* t64, float2* t32, double2* s64, float2* s32,
                 int* pos) {
  const int L = ax.L;
  for (int e = 0; e < L; ++e) {
    // exp(+2 pi i e / L) from the host math library in FP64
    const double ang = 6.283185307179586476925286766559 * (double)e / (double)L;
    t64[e] = make_double2(std::cos(ang), std::sin(ang));
    t32[e] = make_float2((float)t64[e].x, (float)t64[e].y);
  }
  std::vector<double> sre(L, 0.0), sim(L, 0.0);
  stage_twiddles(ax, sre.data(), sim.data());
  for (int e = 0; e < L; ++e) {
    s64[e] = make_double2(sre[e], sim[e]);
    s32[e] = make_float2((float)sre[e], (float)sim[e]);
  }
  axis_positions(ax, pos);
}

bool make_grid_plan(int nx, int ny, int nz, size_t lds_limit, GridPlan* out, std::string* why) {
  GridPlan g{};
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    g.L[a] = smooth235_at_least(2 * n[a]);
    g.F[a] = g.L[a] / 2 + 1;
    if (!make_axis_plan(g.L[a], &g.ax[a])) {
      if (why) *why = "no radix plan for L=" + std::to_string(g.L[a]);
      return false;
    }
  }
  if (!make_tiles(g, lds_limit)) {
    if (why) *why = "padded FFT length above the supported maximum";
    return false;
  }
  *out = g;
  return true;
}

bool make_tiles(GridPlan& g, size_t lds_limit) {
  const int L1 = g.L[0];
  // LDS budgets per block from a tile-width sweep on MI355X (200^3 grid):
  // FP32 passes are fastest at ~26 KB (B = 5 for L = 200), FP64 y/x passes
  // at ~33 KB (B = 3) and the FP64 z convolution at ~52 KB (B = 5); smaller
  // tiles leave more workgroups per CU to overlap load, FFT and store phases.
  // On short lines the budgets allow wide tiles (B = 17..64 at L = 16..48),
  // which a sweep of 8^3..100^3 boxes measured 10-35% slower per matvec than
  // narrow ones: widths are further capped at 4 lines (x passes) and 5 (y
  // and z passes), which came within ~2% of the best width per pass on
  // average.  A z pass keeps its budget width where a compile-time kernel is
  // built for it.
  // Batched launches (several right-hand sides, packs) keep the budget
  // widths: the right-hand sides already give the grid its workgroups, and
  // narrower tiles there only add per-workgroup table loads (measured: the
  // capped widths made the 256-RHS batch and the packed uneven batch 8-15%
  // slower).
  constexpr int kCapX = 4, kCapYZ = 5;
  const size_t b32 = std::min<size_t>(lds_limit, 26624), b64 = std::min<size_t>(lds_limit, 33792),
               b32c = std::min<size_t>(lds_limit, 26624), b64c = std::min<size_t>(lds_limit, 53248);
  const int L3 = g.L[2];
  for (int p = 0; p < 2; ++p) {
    const int esz = p == 0 ? 16 : 8;
    const size_t bs = p == 0 ? b64 : b32, bc = p == 0 ? b64c : b32c;
    g.p1[p] = tile_for(g.ax[0], 3, esz, g.ny * g.nz, bs, false, true);
    g.p2[p] = tile_for(g.ax[1], 3, esz, L1, bs, false);
    // FP32 batched z pass at L3 = 128: width 5 at five waves per SIMD
    // (zct_min_waves) rather than the budget's 8 at four (VALU-bound either
    // way; the fifth wave hides more of its latency)
    g.p3[p] = tile_for(g.ax[2], 3, esz, (p == 1 && L3 == 128) ? std::min(L1, 5) : L1, bc, true);
    g.p4[p] = tile_for(g.ax[1], 3, esz, L1, bs, false);
    g.p5[p] = tile_for(g.ax[0], 3, esz, g.ny * g.nz, bs, false, true);
    // single right-hand side: capped widths
    TileShape* o = g.one[p];
    o[0] = tile_for(g.ax[0], 3, esz, std::min(g.ny * g.nz, kCapX), bs, false, true);
    o[1] = tile_for(g.ax[1], 3, esz, std::min(L1, kCapYZ), bs, false);
    o[2] = tile_for(g.ax[2], 3, esz, std::min(L1, kCapYZ), bc, true);
    if (!zconv_ct_for<float2>(L3, o[2].B) && zconv_ct_for<float2>(L3, g.p3[p].B)) o[2] = g.p3[p];
    o[3] = tile_for(g.ax[1], 3, esz, std::min(L1, kCapYZ), bs, false);
    o[4] = tile_for(g.ax[0], 3, esz, std::min(g.ny * g.nz, kCapX), bs, false, true);
    // widths measured by the autotuner for this box, where they fit
    int tb[5];
    if (tuned_widths(g.nx, g.ny, g.nz, p == 0 ? 64 : 32, tb)) {
      for (int q = 0; q < 5; ++q) {
        const TileShape u = pass_tile_shape(g, p, q, tb[q], lds_limit);
        if (u.B == tb[q]) o[q] = u;
      }
    }
  }
  g.gx = tile_for(g.ax[0], 1, 16, g.L[1] * g.L[2], lds_limit, false);
  g.gy = tile_for(g.ax[1], 1, 16, L1, lds_limit, false);
  g.gz = tile_for(g.ax[2], 1, 16, L1, lds_limit, false);
  for (const TileShape* t : {&g.p1[0], &g.p2[0], &g.p3[0], &g.p4[0], &g.p5[0], &g.p1[1], &g.p2[1],
                             &g.p3[1], &g.p4[1], &g.p5[1], &g.gx, &g.gy, &g.gz})
    if (t->B == 0) return false;
  for (int p = 0; p < 2; ++p)
    for (int q = 0; q < 5; ++q)
      if (g.one[p][q].B == 0) return false;
  return true;
}

template <typename K> static bool allow_lds(K kernel, size_t bytes) {
  static_assert(sizeof(K) > 0, "");
  if (bytes <= 65536) return true;
  return hip_ok(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes),
                "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

// Argument blocks of the five passes of one problem: everything but the
// vectors, the right-hand-side strides and the skip flags of a call.
template <typename V> struct PassSet {
  XArgs<V> p1, p5;
  SArgs<V> p2, p3, p4;
  const TileShape* t[5];
  int gx[5], gy[5];  // tiles of one right-hand side: x and y extents
};

template <typename V>
static void pass_set(const GridPlan& g, const MatvecDev<V>& m, int pi, bool single,
                     PassSet<V>& o) {
  using S = decltype(V::x);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int L1 = g.L[0], L2 = g.L[1], L3 = g.L[2];
  const long long l1c = (long long)L1 * 3;
  o.t[0] = single ? &g.one[pi][0] : &g.p1[pi];
  o.t[1] = single ? &g.one[pi][1] : &g.p2[pi];
  o.t[2] = single ? &g.one[pi][2] : &g.p3[pi];
  o.t[3] = single ? &g.one[pi][3] : &g.p4[pi];
  o.t[4] = single ? &g.one[pi][4] : &g.p5[pi];
  // the line masks of a sparse map, for x tiles of at most 64 lines (a mask
  // in all four passes or in none)
  const unsigned long long* rows =
      m.tab && m.rows && ny <= 64 && o.t[0]->B <= 64 && o.t[4]->B <= 64 ? m.rows : nullptr;
  {  // P1: scatter(alpha x) + FFT-x
    XArgs<V>& a = o.p1;
    a = XArgs<V>{};
    a.dst = m.A;
    a.tw = m.tw[0];
    a.tws = m.tws[0];
    a.pos = m.pos[0];
    a.ax = g.ax[0];
    a.n_in = nx;
    a.n_out = L1;
    a.nlines = ny * nz;
    a.B = o.t[0]->B;
    a.nthr = o.t[0]->threads;
    a.ld_dst = l1c;
    a.tab = m.tab;
    a.nx = nx;
    a.alpha = m.alpha;
    a.alpha_mode = m.alpha_mode;
    a.rows = rows;
    a.ny = ny;
    o.gx[0] = (a.nlines + a.B - 1) / a.B;
    o.gy[0] = 1;
  }
  {  // P2: FFT-y over the nz non-zero planes
    SArgs<V>& a = o.p2;
    a = SArgs<V>{};
    a.src = m.A;
    a.dst = m.B;
    a.tw = m.tw[1];
    a.pos = m.pos[1];
    a.ax = g.ax[1];
    a.n_in = ny;
    a.n_out = L2;
    a.nbx = L1;
    a.B = o.t[1]->B;
    a.nthr = o.t[1]->threads;
    a.so_src = (long long)ny * l1c;
    a.sp_src = l1c;
    a.so_dst = (long long)L2 * l1c;
    a.sp_dst = l1c;
    a.rows = rows;
    o.gx[1] = a.gx = (L1 + a.B - 1) / a.B;
    o.gy[1] = nz;
  }
  {  // P3: FFT-z . G . IFFT-z, in place, keeps kz < nz
    SArgs<V>& a = o.p3;
    a = SArgs<V>{};
    a.src = m.B;
    a.dst = m.B;
    a.tw = m.tw[2];
    a.pos = m.pos[2];
    a.ax = g.ax[2];
    a.n_in = nz;
    a.n_out = nz;
    a.nbx = L1;
    a.B = o.t[2]->B;
    a.nthr = o.t[2]->threads;
    a.so_src = a.so_dst = l1c;
    a.sp_src = a.sp_dst = (long long)L2 * l1c;
    a.G = m.G;
    a.F1 = g.F[0];
    a.F3 = g.F[2];
    a.L1 = L1;
    a.L2 = L2;
    a.tws = m.tws[2];
    o.gx[2] = a.gx = (L1 + a.B - 1) / a.B;
    o.gy[2] = L2 / 2 + 1;
  }
  {  // P4: IFFT-y, keeps j < ny
    SArgs<V>& a = o.p4;
    a = SArgs<V>{};
    a.src = m.B;
    a.dst = m.A;
    a.tw = m.tw[1];
    a.pos = m.pos[1];
    a.ax = g.ax[1];
    a.n_in = L2;
    a.n_out = ny;
    a.nbx = L1;
    a.B = o.t[3]->B;
    a.nthr = o.t[3]->threads;
    a.so_src = (long long)L2 * l1c;
    a.sp_src = l1c;
    a.so_dst = (long long)ny * l1c;
    a.sp_dst = l1c;
    a.rows = rows;
    o.gx[3] = a.gx = (L1 + a.B - 1) / a.B;
    o.gy[3] = nz;
  }
  {  // P5: IFFT-x + gather, y = b - v/M
    XArgs<V>& a = o.p5;
    a = XArgs<V>{};
    a.src = m.A;
    a.tw = m.tw[0];
    a.tws = m.tws[0];
    a.pos = m.pos[0];
    a.ax = g.ax[0];
    a.n_in = L1;
    a.n_out = nx;
    a.nlines = ny * nz;
    a.B = o.t[4]->B;
    a.nthr = o.t[4]->threads;
    a.ld_src = l1c;
    a.tab = m.tab;
    a.nx = nx;
    a.rows = rows;
    a.ny = ny;
    a.inv_m = (S)(1.0 / ((double)L1 * (double)L2 * (double)L3));
    o.gx[4] = (a.nlines + a.B - 1) / a.B;
    o.gy[4] = 1;
  }
}

bool row_masks(const int32_t* tab, int nx, int ny, int nz, std::vector<unsigned long long>& out) {
  if (!tab || ny > 64 || nx <= 0 || ny <= 0 || nz <= 0) return false;
  std::vector<unsigned long long> m((size_t)nz, 0ull);
  bool gap = false;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      const int32_t* t = tab + ((size_t)k * ny + j) * nx;
      bool any = false;
      for (int i = 0; i < nx && !any; ++i) any = t[i] > 0;
      if (any)
        m[k] |= 1ull << j;
      else
        gap = true;
    }
  if (!gap) return false;
  out.swap(m);
  return true;
}

template <typename V>
static bool matvec_impl(const GridPlan& g, const MatvecDev<V>& m, const V* x, const V* b, V* y,
                        hipStream_t s, int pi, const int* skip, PassTimer* tm,
                        const KFuse<V>* f = nullptr) {
  const int R = std::max(1, m.nrhs);
  const long long zA = (long long)g.nz * g.ny * g.L[0] * 3;  // A per RHS
  const long long zB = (long long)g.nz * g.L[1] * g.L[0] * 3;  // B per RHS
  PassSet<V> ps;
  pass_set(g, m, pi, R == 1, ps);
  auto grid = [&](int p) { return dim3(ps.gx[p], ps.gy[p], R); };
  // compile-time x passes (both need 2 nx <= L1, which every plan keeps), in
  // FP64: they took config 2's P1 / P5 from 49.5 / 42.5 to 44.6 / 38.7 us,
  // while in FP32 P1 stayed at 26.8 us and the fused P5 (more VGPRs than the
  // runtime kernel, fewer waves) went from 33.0 to 36.5 us
  const bool ct_x = pi == 0 && m.specialized && m.tws[0] && 2 * g.nx <= g.L[0];
  {  // P1: scatter(alpha x) + FFT-x
    XArgs<V> a = ps.p1;
    a.skip = skip;
    a.skip_stride = m.skip_stride;
    a.xin = x;
    a.zs_x = m.vstride;
    a.zs_dst = zA;
    if (f && f->p1) {
      if (f->p != x || !f->r || !f->u || !f->st) return false;
      auto k = k_xpass_f<V, 3, +1, 3>;
      if (!allow_lds(k, ps.t[0]->lds)) return false;
      if (tm) tm->begin(pi * 5 + 0, s);
      hipLaunchKernelGGL(k, grid(0), dim3(ps.t[0]->threads), ps.t[0]->lds, s, a, *f);
      if (tm) tm->end(s);
    } else if (XpassKernel<V> kc = ct_x ? xpass_ct_for<V, +1, 1>(g.L[0], ps.t[0]->B) : nullptr) {
      if (!allow_lds(kc, ps.t[0]->lds)) return false;
      if (tm) tm->begin(pi * 5 + 0, s);
      hipLaunchKernelGGL(kc, grid(0), dim3(ps.t[0]->threads), ps.t[0]->lds, s, a, KFuse<V>{});
      if (tm) tm->end(s);
    } else {
      auto k = k_xpass<V, 3, +1, 1>;
      if (!allow_lds(k, ps.t[0]->lds)) return false;
      if (tm) tm->begin(pi * 5 + 0, s);
      hipLaunchKernelGGL(k, grid(0), dim3(ps.t[0]->threads), ps.t[0]->lds, s, a);
      if (tm) tm->end(s);
    }
  }
  {  // P2: FFT-y over the nz non-zero planes
    SArgs<V> a = ps.p2;
    a.skip = skip;
    a.skip_stride = m.skip_stride;
    a.zs_src = zA;
    a.zs_dst = zB;
    const int B = ps.t[1]->B;
    if (m.specialized && m.tws[1] && 2 * g.ny <= g.L[1] &&
        has_strided_ct(g.L[1], B, pi == 0 ? 64 : 32, +1)) {
      a.tws = m.tws[1];
      if (tm) tm->begin(pi * 5 + 1, s);
      if (!launch_strided_ct(a, g.L[1], B, +1, grid(1), tile_lds_bytes<V>(g.L[1], 3 * B), s))
        return false;
      if (tm) tm->end(s);
    } else {
      auto k = k_strided<V, 3, +1>;
      if (!allow_lds(k, ps.t[1]->lds)) return false;
      if (tm) tm->begin(pi * 5 + 1, s);
      hipLaunchKernelGGL(k, grid(1), dim3(ps.t[1]->threads), ps.t[1]->lds, s, a);
      if (tm) tm->end(s);
    }
  }
  {  // P3: FFT-z . G . IFFT-z, in place, keeps kz < nz
    SArgs<V> a = ps.p3;
    a.skip = skip;
    a.skip_stride = m.skip_stride;
    a.zs_src = a.zs_dst = zB;
    const TileShape& t = *ps.t[2];
    const int items = g.F[2] * t.B;
    const int ng = (items + t.threads - 1) / t.threads;
    ZconvKernel<V> k = m.specialized && !m.dpar && 2 * g.nz <= g.L[2]
                           ? zconv_ct_for<V>(g.L[2], t.B) : nullptr;
    if (m.dpar) {
      a.dpar = m.dpar;
      k = ng <= 1 ? k_zconv_der<V, 1> : ng == 2 ? k_zconv_der<V, 2> : ng == 3 ? k_zconv_der<V, 3>
                                                                             : k_zconv_der<V, kMaxG>;
    }
    if (!k)
      k = ng <= 1 ? k_zconv<V, 1> : ng == 2 ? k_zconv<V, 2> : ng == 3 ? k_zconv<V, 3>
                                                                     : k_zconv<V, kMaxG>;
    // the register-staged z pass (ifdda_zreg.hip) where selected and built
    const bool zr = m.zreg && m.specialized && !m.dpar && 2 * g.nz <= g.L[2] &&
                    has_zreg(g.L[2], t.B, pi == 0 ? 64 : 32);
    if (zr) {
      if (tm) tm->begin(pi * 5 + 2, s);
      if (!launch_zreg(a, g.L[2], t.B, grid(2), s)) return false;
      if (tm) tm->end(s);
    } else {
      if (!allow_lds(k, t.lds)) return false;
      if (tm) tm->begin(pi * 5 + 2, s);
      hipLaunchKernelGGL(k, grid(2), dim3(t.threads), t.lds, s, a);
      if (tm) tm->end(s);
    }
  }
  {  // P4: IFFT-y, keeps j < ny
    SArgs<V> a = ps.p4;
    a.skip = skip;
    a.skip_stride = m.skip_stride;
    a.zs_src = zB;
    a.zs_dst = zA;
    const int B = ps.t[3]->B;
    if (m.specialized && m.tws[1] && 2 * g.ny <= g.L[1] &&
        has_strided_ct(g.L[1], B, pi == 0 ? 64 : 32, -1)) {
      a.tws = m.tws[1];
      if (tm) tm->begin(pi * 5 + 3, s);
      if (!launch_strided_ct(a, g.L[1], B, -1, grid(3), tile_lds_bytes<V>(g.L[1], 3 * B), s))
        return false;
      if (tm) tm->end(s);
    } else {
      auto k = k_strided<V, 3, -1>;
      if (!allow_lds(k, ps.t[3]->lds)) return false;
      if (tm) tm->begin(pi * 5 + 3, s);
      hipLaunchKernelGGL(k, grid(3), dim3(ps.t[3]->threads), ps.t[3]->lds, s, a);
      if (tm) tm->end(s);
    }
  }
  {  // P5: IFFT-x + gather, y = b - v/M
    XArgs<V> a = ps.p5;
    a.skip = skip;
    a.skip_stride = m.skip_stride;
    a.xin = b;
    a.xout = y;
    a.zs_x = m.vstride;
    a.zs_src = zA;
    const int mode5 = f && f->p5 ? 4 + f->p5 : 2;  // 5, 6, 7
    XpassKernel<V> kc = nullptr;
    if (ct_x)
      kc = mode5 == 2   ? xpass_ct_for<V, -1, 2>(g.L[0], ps.t[4]->B)
           : mode5 == 5 ? xpass_ct_for<V, -1, 5>(g.L[0], ps.t[4]->B)
           : mode5 == 6 ? xpass_ct_for<V, -1, 6>(g.L[0], ps.t[4]->B)
                        : xpass_ct_for<V, -1, 7>(g.L[0], ps.t[4]->B);
    if (mode5 != 2 && (ps.t[4]->threads < 64 || f->pstride < ps.gx[4]))
      return false;  // the block sum of the fused inner products needs a full first wavefront
    if (kc) {
      if (!allow_lds(kc, ps.t[4]->lds)) return false;
      if (tm) tm->begin(pi * 5 + 4, s);
      hipLaunchKernelGGL(kc, grid(4), dim3(ps.t[4]->threads), ps.t[4]->lds, s, a,
                         mode5 == 2 ? KFuse<V>{} : *f);
      if (tm) tm->end(s);
    } else if (mode5 != 2) {
      auto k = mode5 == 5   ? k_xpass_f<V, 3, -1, 5>
               : mode5 == 6 ? k_xpass_f<V, 3, -1, 6>
                            : k_xpass_f<V, 3, -1, 7>;
      if (!allow_lds(k, ps.t[4]->lds)) return false;
      if (tm) tm->begin(pi * 5 + 4, s);
      hipLaunchKernelGGL(k, grid(4), dim3(ps.t[4]->threads), ps.t[4]->lds, s, a, *f);
      if (tm) tm->end(s);
    } else {
      auto k = k_xpass<V, 3, -1, 2>;
      if (!allow_lds(k, ps.t[4]->lds)) return false;
      if (tm) tm->begin(pi * 5 + 4, s);
      hipLaunchKernelGGL(k, grid(4), dim3(ps.t[4]->threads), ps.t[4]->lds, s, a);
      if (tm) tm->end(s);
    }
  }
  return hip_ok(hipGetLastError(), "matvec launch");
}

// ----------------------------------------------------------- packed matvec
template <typename V>
static size_t pack_blob_impl(const std::vector<const GridPlan*>& g,
                             const std::vector<MatvecDev<V>>& m, const long long* offs, int pi,
                             PackLayout& lay, std::vector<unsigned char>& blob) {
  const int P = (int)g.size();
  const size_t asz[5] = {sizeof(XArgs<V>), sizeof(SArgs<V>), sizeof(SArgs<V>), sizeof(SArgs<V>),
                         sizeof(XArgs<V>)};
  auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
  size_t off = 0;
  for (int p = 0; p < 5; ++p) {
    lay.av[p] = off;
    off = al(off + asz[p] * P);
    lay.boff[p] = off;
    off = al(off + sizeof(int) * (P + 1));
  }
  lay.mb = off;
  off = al(off + sizeof(int) * 5 * P);
  blob.assign(off, 0);
  lay.mbh.assign(5 * (size_t)P, 0);
  for (int p = 0; p < 5; ++p) {
    lay.nblk[p] = 0;
    lay.lds[p] = 0;
  }
  lay.ng3 = 1;
  for (int i = 0; i < P; ++i) {
    PassSet<V> ps;
    pass_set(*g[i], m[i], pi, false, ps);
    ps.p1.zs_x = ps.p5.zs_x = offs[i];
    const void* src[5] = {&ps.p1, &ps.p2, &ps.p3, &ps.p4, &ps.p5};
    for (int p = 0; p < 5; ++p) {
      std::memcpy(blob.data() + lay.av[p] + asz[p] * i, src[p], asz[p]);
      int* bo = reinterpret_cast<int*>(blob.data() + lay.boff[p]);
      bo[i] = lay.nblk[p];
      lay.nblk[p] += ps.gx[p] * ps.gy[p];
      bo[i + 1] = lay.nblk[p];
      lay.mbh[(size_t)p * P + i] = ps.gx[p] * ps.gy[p];
      reinterpret_cast<int*>(blob.data() + lay.mb)[(size_t)p * P + i] = ps.gx[p] * ps.gy[p];
      lay.lds[p] = std::max(lay.lds[p], ps.t[p]->lds);
    }
    // Green items per thread of the z pass (as the single-problem launch)
    const int items = g[i]->F[2] * ps.t[2]->B;
    lay.ng3 = std::max(lay.ng3, (items + ps.t[2]->threads - 1) / ps.t[2]->threads);
  }
  lay.P = P;
  return off;
}

size_t pack_blob_f64(const std::vector<const GridPlan*>& g,
                     const std::vector<MatvecDev<double2>>& m, const long long* offs,
                     PackLayout& lay, std::vector<unsigned char>& blob) {
  return pack_blob_impl(g, m, offs, 0, lay, blob);
}
size_t pack_blob_f32(const std::vector<const GridPlan*>& g,
                     const std::vector<MatvecDev<float2>>& m, const long long* offs,
                     PackLayout& lay, std::vector<unsigned char>& blob) {
  return pack_blob_impl(g, m, offs, 1, lay, blob);
}

template <typename V>
static bool matvec_pack_impl(const PackLayout& lay, const void* dev, const V* x, const V* b, V* y,
                             hipStream_t s, const int* skip, int skip_stride) {
  const unsigned char* base = static_cast<const unsigned char*>(dev);
  auto args = [&](int p, const V* in, V* out) {
    PkArgs<V> k{};
    k.av = base + lay.av[p];
    k.boff = reinterpret_cast<const int*>(base + lay.boff[p]);
    k.P = lay.P;
    k.x = in;
    k.y = out;
    k.skip = skip;
    k.skip_stride = skip_stride;
    return k;
  };
  auto k1 = k_xpass_pk<V, 3, +1, 1>;
  auto k2 = k_strided_pk<V, 3, +1>;
  auto k3 = lay.ng3 <= 1 ? k_zconv_pk<V, 1> : lay.ng3 == 2 ? k_zconv_pk<V, 2>
                            : lay.ng3 == 3 ? k_zconv_pk<V, 3> : k_zconv_pk<V, kMaxG>;
  auto k4 = k_strided_pk<V, 3, -1>;
  auto k5 = k_xpass_pk<V, 3, -1, 2>;
  if (!allow_lds(k1, lay.lds[0]) || !allow_lds(k2, lay.lds[1]) || !allow_lds(k3, lay.lds[2]) ||
      !allow_lds(k4, lay.lds[3]) || !allow_lds(k5, lay.lds[4]))
    return false;
  const dim3 T(256);
  if (lay.nblk[0] == 0) return true;  // every member has finished (compacted away)
  hipLaunchKernelGGL(k1, dim3(lay.nblk[0]), T, lay.lds[0], s, args(0, x, nullptr));
  hipLaunchKernelGGL(k2, dim3(lay.nblk[1]), T, lay.lds[1], s, args(1, nullptr, nullptr));
  hipLaunchKernelGGL(k3, dim3(lay.nblk[2]), T, lay.lds[2], s, args(2, nullptr, nullptr));
  hipLaunchKernelGGL(k4, dim3(lay.nblk[3]), T, lay.lds[3], s, args(3, nullptr, nullptr));
  hipLaunchKernelGGL(k5, dim3(lay.nblk[4]), T, lay.lds[4], s, args(4, b, y));
  return hip_ok(hipGetLastError(), "packed matvec launch");
}

// Block offsets of the five packed passes from the blocks per member (mb,
// pass-major) and the members' solver state: members that have finished get
// none (unless all).  One thread per pass, members in order.
struct BoffOut {
  int* bo[5];
};
__global__ void k_pack_boff(const int* __restrict__ mb, int P, const KState* __restrict__ st,
                            int all, BoffOut o) {
  const int p = threadIdx.x;
  if (p >= 5) return;
  int* bo = o.bo[p];
  int acc = 0;
  bo[0] = 0;
  for (int m = 0; m < P; ++m) {
    if (all || !st[m].done) acc += mb[(size_t)p * P + m];
    bo[m + 1] = acc;
  }
}

bool pack_compact(PackLayout& lay, void* dev, const KState* st, const int* host_done, bool all,
                  hipStream_t s) {
  if (!dev || lay.P == 0) return true;
  unsigned char* base = static_cast<unsigned char*>(dev);
  BoffOut o;
  for (int p = 0; p < 5; ++p) {
    o.bo[p] = reinterpret_cast<int*>(base + lay.boff[p]);
    int n = 0;
    for (int m = 0; m < lay.P; ++m)
      if (all || !host_done[m]) n += lay.mbh[(size_t)p * lay.P + m];
    lay.nblk[p] = n;
  }
  hipLaunchKernelGGL(k_pack_boff, dim3(1), dim3(64), 0, s,
                     reinterpret_cast<const int*>(base + lay.mb), lay.P, st, all ? 1 : 0, o);
  return hip_ok(hipGetLastError(), "pack compaction");
}

bool launch_matvec_pack_f64(const PackLayout& lay, const void* dev, const double2* x,
                            const double2* b, double2* y, hipStream_t s, const int* skip,
                            int skip_stride) {
  return matvec_pack_impl(lay, dev, x, b, y, s, skip, skip_stride);
}
bool launch_matvec_pack_f32(const PackLayout& lay, const void* dev, const float2* x,
                            const float2* b, float2* y, hipStream_t s, const int* skip,
                            int skip_stride) {
  return matvec_pack_impl(lay, dev, x, b, y, s, skip, skip_stride);
}

bool launch_matvec_f64(const GridPlan& g, const MatvecDev<double2>& m, const double2* x,
                       const double2* b, double2* y, hipStream_t s, const int* skip,
                       PassTimer* timer) {
  return matvec_impl(g, m, x, b, y, s, 0, skip, timer ? timer : m.timer);
}
bool launch_matvec_f32(const GridPlan& g, const MatvecDev<float2>& m, const float2* x,
                       const float2* b, float2* y, hipStream_t s, const int* skip,
                       PassTimer* timer) {
  return matvec_impl(g, m, x, b, y, s, 1, skip, timer ? timer : m.timer);
}

int matvec_p5_blocks(const GridPlan& g, int prec, int R) {
  const int pi = prec == 64 ? 0 : 1;
  const TileShape& t = R <= 1 ? g.one[pi][4] : g.p5[pi];
  if (t.B < 1 || t.threads < 64) return 0;
  return (g.ny * g.nz + t.B - 1) / t.B;
}
bool launch_matvec_fused_f64(const GridPlan& g, const MatvecDev<double2>& m, const double2* x,
                             const double2* b, double2* y, hipStream_t s, const int* skip,
                             const KFuse<double2>& f) {
  return matvec_impl(g, m, x, b, y, s, 0, skip, m.timer, &f);
}
bool launch_matvec_fused_f32(const GridPlan& g, const MatvecDev<float2>& m, const float2* x,
                             const float2* b, float2* y, hipStream_t s, const int* skip,
                             const KFuse<float2>& f) {
  return matvec_impl(g, m, x, b, y, s, 1, skip, m.timer, &f);
}

// In-place forward 3D FFT (sign +1) of one FP64 component on the full grid.
// In-place forward 3D FFT of nc full grids M = L1 L2 L3 values apart (the
// grids are the launches' right-hand sides).
static bool fft3d_full_f64(const GridPlan& g, const double2* tw[3], const int* pos[3], double2* v,
                           hipStream_t s, int nc = 1) {
  const int L1 = g.L[0], L2 = g.L[1], L3 = g.L[2];
  const long long M = (long long)L1 * L2 * L3;
  {
    XArgs<double2> a{};
    a.src = v;
    a.dst = v;
    a.zs_src = a.zs_dst = M;
    a.tw = tw[0];
    a.pos = pos[0];
    a.ax = g.ax[0];
    a.n_in = a.n_out = L1;
    a.nlines = L2 * L3;
    a.B = g.gx.B;
    a.ld_src = a.ld_dst = L1;
    auto k = k_xpass<double2, 1, +1, 0>;
    if (!allow_lds(k, g.gx.lds)) return false;
    hipLaunchKernelGGL(k, dim3((a.nlines + a.B - 1) / a.B, 1, nc), dim3(g.gx.threads), g.gx.lds, s,
                       a);
  }
  {
    SArgs<double2> a{};
    a.src = v;
    a.dst = v;
    a.zs_src = a.zs_dst = M;
    a.tw = tw[1];
    a.pos = pos[1];
    a.ax = g.ax[1];
    a.n_in = a.n_out = L2;
    a.nbx = L1;
    a.B = g.gy.B;
    a.so_src = a.so_dst = (long long)L1 * L2;
    a.sp_src = a.sp_dst = L1;
    auto k = k_strided<double2, 1, +1>;
    if (!allow_lds(k, g.gy.lds)) return false;
    hipLaunchKernelGGL(k, dim3((L1 + a.B - 1) / a.B, L3, nc), dim3(g.gy.threads), g.gy.lds, s, a);
  }
  {
    SArgs<double2> a{};
    a.src = v;
    a.dst = v;
    a.zs_src = a.zs_dst = M;
    a.tw = tw[2];
    a.pos = pos[2];
    a.ax = g.ax[2];
    a.n_in = a.n_out = L3;
    a.nbx = L1;
    a.B = g.gz.B;
    a.so_src = a.so_dst = L1;
    a.sp_src = a.sp_dst = (long long)L1 * L2;
    auto k = k_strided<double2, 1, +1>;
    if (!allow_lds(k, g.gz.lds)) return false;
    hipLaunchKernelGGL(k, dim3((L1 + a.B - 1) / a.B, L2, nc), dim3(g.gz.threads), g.gz.lds, s, a);
  }
  return hip_ok(hipGetLastError(), "fft3d launch");
}

int green_build_comps(const GridPlan& g) {
  const size_t M = (size_t)g.L[0] * g.L[1] * g.L[2];
  return 6 * M * sizeof(double2) <= ((size_t)512 << 20) ? 6 : 1;
}

bool green_fold_component(const GridPlan& g, const double2* full, int comp, double2* g64,
                          float2* g32, hipStream_t s) {
  const long long nf = (long long)g.F[0] * g.F[1] * g.F[2];
  const int blocks = (int)std::min<long long>((nf + 255) / 256, 65535);
  hipLaunchKernelGGL(k_green_fold, dim3(blocks), dim3(256), 0, s, g64, g32, full, comp, g.L[0],
                     g.L[1], g.F[0], g.F[1], g.F[2]);
  return hip_ok(hipGetLastError(), "green fold");
}

bool green_dbuild_device(const GridPlan& g, const double2* tw64[3], const int* pos[3], double d,
                         double k0, int dd, double2* full, double2* g64, hipStream_t s) {
  const long long M = (long long)g.L[0] * g.L[1] * g.L[2];
  bool ok = true;
  for (int c = 0; ok && c < 6; ++c) {
    const int blocks = (int)std::min<long long>((M + 255) / 256, 65535);
    hipLaunchKernelGGL(k_green_dfill, dim3(blocks), dim3(256), 0, s, full, c, dd, g.nx, g.ny, g.nz,
                       g.L[0], g.L[1], g.L[2], d, k0);
    ok = hip_ok(hipGetLastError(), "green derivative fill") && fft3d_full_f64(g, tw64, pos, full, s) &&
         green_fold_component(g, full, c, g64, nullptr, s);
  }
  return ok;
}

__global__ void k_fill_zero(double2* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = make_double2(0.0, 0.0);
}

void green_scratch_free(GreenScratch& gs) {
  (void)hipFree(gs.near);
  (void)hipFree(gs.part);
  (void)hipFree(gs.offs);
  gs = GreenScratch{};
}

// Grow-only device buffer for a build on stream s (the stream's earlier work
// ends before the old buffer is freed).
template <typename T>
static bool scratch_fit(T** p, size_t* cap, size_t bytes, hipStream_t s, const char* what) {
  if (bytes <= *cap) return true;
  if (*p && !hip_ok(hipStreamSynchronize(s), what)) return false;
  (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (!hip_ok(hipMalloc(reinterpret_cast<void**>(p), bytes), what)) return false;
  *cap = bytes;
  return true;
}

bool green_build_device(const GridPlan& g, const double2* tw64[3], const int* pos[3], double d,
                        double k0, int nquad, double2* full, double2* g64, float2* g32,
                        GreenScratch* gs, hipStream_t s, int ncomp) {
  const long long M = (long long)g.L[0] * g.L[1] * g.L[2];
  double2* near = nullptr;
  bool ok = true;
  if (nquad > 0) {
    // Near cells: |offset|^2 <= nquad^2 inside the box reach (calculgreenquadfft.f:50).
    std::vector<int> h;
    for (int k = -nquad; k <= nquad; ++k)
      for (int j = -nquad; j <= nquad; ++j)
        for (int i = -nquad; i <= nquad; ++i)
          if (i * i + j * j + k * k <= nquad * nquad) {
            h.push_back(i);
            h.push_back(j);
            h.push_back(k);
          }
    const int ncell = (int)h.size() / 3;
    const int w = 2 * nquad + 1;
    const long long nnear = 6LL * w * w * w;
    ok = gs && scratch_fit(&gs->near, &gs->near_cap, sizeof(double2) * nnear, s, "green near") &&
         scratch_fit(&gs->part, &gs->part_cap, sizeof(double2) * 6 * 129 * (size_t)ncell, s,
                     "green near") &&
         scratch_fit(&gs->offs, &gs->offs_cap, sizeof(int) * h.size(), s, "green near") &&
         hip_ok(hipMemcpyAsync(gs->offs, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s),
                "copy offs");
    if (ok) {
      near = gs->near;
      hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((nnear + 255) / 256)), dim3(256), 0, s, near,
                         nnear);
      hipLaunchKernelGGL(k_green_near_partial, dim3(129, ncell), dim3(256), 0, s, gs->part,
                         gs->offs, d, k0);
      hipLaunchKernelGGL(k_green_near_final, dim3((ncell + 63) / 64), dim3(64), 0, s, near,
                         gs->part, gs->offs, ncell, nquad, d, k0, 1.0);
      ok = hip_ok(hipGetLastError(), "green near") &&
           hip_ok(hipStreamSynchronize(s), "green near");  // h (copied from) ends here
    }
  }
  const int blocks = (int)std::min<long long>((M + 255) / 256, 65535);
  if (ok && ncomp == 6) {  // five launches for the six components (full holds 6 M)
    const long long nf = (long long)g.F[0] * g.F[1] * g.F[2];
    hipLaunchKernelGGL(k_green_fill6, dim3(blocks), dim3(256), 0, s, full, g.nx, g.ny, g.nz,
                       g.L[0], g.L[1], g.L[2], d, k0, near, nquad);
    ok = hip_ok(hipGetLastError(), "green fill") && fft3d_full_f64(g, tw64, pos, full, s, 6);
    if (ok)
      hipLaunchKernelGGL(k_green_fold6, dim3((int)std::min<long long>((nf + 255) / 256, 65535)),
                         dim3(256), 0, s, g64, g32, full, M, g.L[0], g.L[1], g.F[0], g.F[1],
                         g.F[2]);
    return ok && hip_ok(hipGetLastError(), "green fold");
  }
  for (int c = 0; ok && c < 6; ++c) {
    hipLaunchKernelGGL(k_green_fill, dim3(blocks), dim3(256), 0, s, full, c, g.nx, g.ny, g.nz,
                       g.L[0], g.L[1], g.L[2], d, k0, near, nquad);
    ok = hip_ok(hipGetLastError(), "green fill") && fft3d_full_f64(g, tw64, pos, full, s) &&
         green_fold_component(g, full, c, g64, g32, s);
  }
  return ok;
}

}  // namespace ifdda
