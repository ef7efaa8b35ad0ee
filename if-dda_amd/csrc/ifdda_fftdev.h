// ifdda_fftdev.h -- device primitives shared by the FFT passes of the matvec
// (ifdda_fft.hip) and the register-staged z pass (ifdda_zreg.hip): radix-r
// DFTs on values held by one thread, the workgroup barrier, the strided-pass
// argument block, the Green-tensor folding, the symmetric 3x3 product, the
// XCD-aware block order, the LDS tile layout and strided loads, and the
// compile-time radix plans and stages of the LDS-staged transforms.
#pragma once

#include <stddef.h>

#include "ifdda_common.h"

namespace ifdda {

// Radix-r DFT constants (exact trigonometric values).
#define IFDDA_C51 0.309016994374947424102293417183    // cos(2 pi/5)
#define IFDDA_C52 (-0.809016994374947424102293417183) // cos(4 pi/5)
#define IFDDA_S51 0.951056516295153572116439333379    // sin(2 pi/5)
#define IFDDA_S52 0.587785252292473129168705954639    // sin(4 pi/5)
#define IFDDA_S3 0.866025403784438646763723170753     // sin(2 pi/3)
#define IFDDA_R2 0.707106781186547524400844362105     // 1/sqrt(2)

// In-place DFT of R values, y_q = sum_p a_p exp(SGN 2 pi i p q / R).
template <int R, int SGN, typename V> struct Dft;

template <int SGN, typename V> struct Dft<2, SGN, V> {
  static __device__ __forceinline__ void run(V* a) {
    V t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <int SGN, typename V> struct Dft<3, SGN, V> {
  static __device__ __forceinline__ void run(V* a) {
    using S = decltype(V::x);
    V t = cadd(a[1], a[2]), u = csub(a[1], a[2]);
    V m = csub(a[0], cscal(t, S(0.5)));
    V v = mul_si<SGN>(cscal(u, S(IFDDA_S3)));
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, v);
    a[2] = csub(m, v);
  }
};

template <int SGN, typename V> struct Dft<4, SGN, V> {
  static __device__ __forceinline__ void run(V* a) {
    V t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    V t2 = cadd(a[1], a[3]), t3 = mul_si<SGN>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <int SGN, typename V> struct Dft<5, SGN, V> {
  static __device__ __forceinline__ void run(V* a) {
    using S = decltype(V::x);
    V t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
    V u1 = csub(a[1], a[4]), u2 = csub(a[2], a[3]);
    // (a0 + c t1) + c' t2: two fused multiply-adds per component (a0 added
    // last took a multiply, a multiply-add and an add)
    V m1 = cadd(cadd(a[0], cscal(t1, S(IFDDA_C51))), cscal(t2, S(IFDDA_C52)));
    V m2 = cadd(cadd(a[0], cscal(t1, S(IFDDA_C52))), cscal(t2, S(IFDDA_C51)));
    V v1 = mul_si<SGN>(cadd(cscal(u1, S(IFDDA_S51)), cscal(u2, S(IFDDA_S52))));
    V v2 = mul_si<SGN>(csub(cscal(u1, S(IFDDA_S52)), cscal(u2, S(IFDDA_S51))));
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(m1, v1);
    a[4] = csub(m1, v1);
    a[2] = cadd(m2, v2);
    a[3] = csub(m2, v2);
  }
};

template <int SGN, typename V> struct Dft<8, SGN, V> {
  static __device__ __forceinline__ void run(V* a) {
    using S = decltype(V::x);
    V e[4] = {a[0], a[2], a[4], a[6]};
    V o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, SGN, V>::run(e);
    Dft<4, SGN, V>::run(o);
    const S h = S(IFDDA_R2);
    // o1 *= (1 + SGN i)/sqrt2, o2 *= SGN i, o3 *= (-1 + SGN i)/sqrt2
    V w1 = SGN > 0 ? mk((o[1].x - o[1].y) * h, (o[1].y + o[1].x) * h)
                   : mk((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    V w2 = mul_si<SGN>(o[2]);
    V w3 = SGN > 0 ? mk((-o[3].x - o[3].y) * h, (-o[3].y + o[3].x) * h)
                   : mk((-o[3].x + o[3].y) * h, (-o[3].y - o[3].x) * h);
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], w1);
    a[5] = csub(e[1], w1);
    a[2] = cadd(e[2], w2);
    a[6] = csub(e[2], w2);
    a[3] = cadd(e[3], w3);
    a[7] = csub(e[3], w3);
  }
};

// Wait for this thread's LDS operations, then a workgroup barrier.  Global
// loads already issued stay in flight across it (the compiler inserts the
// vmcnt wait at their first use).
__device__ __forceinline__ void lds_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Global loads issued back to back before their LDS stores.
constexpr int kBatch = 8;
// Green-tensor items per thread in the convolution pass (k_zconv<V, NG> is
// instantiated for NG = 1..kMaxG and launched with the smallest that fits).
constexpr int kMaxG = 4;
// Loop start of the threads of a packed launch that a tile does not use:
// above every per-thread loop bound of the FFT tiles (L < 2048, B <= 64).
constexpr int kIdle = 1 << 24;

// R-point DFT of a half-zero input: a[q] = 0 for q >= R/2 (only the first R/2
// are read).  Even outputs are the R/2-point DFT of a, odd outputs that of
// a[p] W^p, W = exp(SGN 2 pi i / R).
template <int R, int SGN, typename V> struct DftHalfIn;
template <int SGN, typename V> struct DftHalfIn<2, SGN, V> {
  static __device__ __forceinline__ void run(V* a) { a[1] = a[0]; }
};
template <int SGN, typename V> struct DftHalfIn<4, SGN, V> {
  static __device__ __forceinline__ void run(V* a) {
    const V w = mul_si<SGN>(a[1]);
    const V x = a[0];
    a[0] = cadd(x, a[1]);
    a[2] = csub(x, a[1]);
    a[1] = cadd(x, w);
    a[3] = csub(x, w);
  }
};
template <int SGN, typename V> struct DftHalfIn<8, SGN, V> {
  static __device__ __forceinline__ void run(V* a) {
    using S = decltype(V::x);
    const S h = S(IFDDA_R2);
    V e[4] = {a[0], a[1], a[2], a[3]};
    V o[4];
    o[0] = a[0];
    o[1] = SGN > 0 ? mk((a[1].x - a[1].y) * h, (a[1].y + a[1].x) * h)
                   : mk((a[1].x + a[1].y) * h, (a[1].y - a[1].x) * h);
    o[2] = mul_si<SGN>(a[2]);
    o[3] = SGN > 0 ? mk((-a[3].x - a[3].y) * h, (-a[3].y + a[3].x) * h)
                   : mk((-a[3].x + a[3].y) * h, (-a[3].y - a[3].x) * h);
    Dft<4, SGN, V>::run(e);
    Dft<4, SGN, V>::run(o);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[2 * q] = e[q];
      a[2 * q + 1] = o[q];
    }
  }
};

// R-point DFT of which only outputs q < R/2 are wanted (left in a[0..R/2)):
// y_q = sum_{p<R/2} (a_p + (-1)^q a_{p+R/2}) W^{pq}.
template <int R, int SGN, typename V> struct DftHalfOut;
template <int SGN, typename V> struct DftHalfOut<2, SGN, V> {
  static __device__ __forceinline__ void run(V* a) { a[0] = cadd(a[0], a[1]); }
};
template <int SGN, typename V> struct DftHalfOut<4, SGN, V> {
  static __device__ __forceinline__ void run(V* a) {
    const V b0 = cadd(a[0], a[2]), b1 = cadd(a[1], a[3]);
    const V c0 = csub(a[0], a[2]), c1 = mul_si<SGN>(csub(a[1], a[3]));
    a[0] = cadd(b0, b1);
    a[1] = cadd(c0, c1);
  }
};
template <int SGN, typename V> struct DftHalfOut<8, SGN, V> {
  static __device__ __forceinline__ void run(V* a) {
    using S = decltype(V::x);
    const S h = S(IFDDA_R2);
    V b[4], c[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      b[p] = cadd(a[p], a[p + 4]);
      c[p] = csub(a[p], a[p + 4]);
    }
    c[1] = SGN > 0 ? mk((c[1].x - c[1].y) * h, (c[1].y + c[1].x) * h)
                   : mk((c[1].x + c[1].y) * h, (c[1].y - c[1].x) * h);
    c[2] = mul_si<SGN>(c[2]);
    c[3] = SGN > 0 ? mk((-c[3].x - c[3].y) * h, (-c[3].y + c[3].x) * h)
                   : mk((-c[3].x + c[3].y) * h, (-c[3].y - c[3].x) * h);
    // outputs 0 and 1 of the 4-point DFTs of b (-> y0, y2) and c (-> y1, y3)
    const V bs = cadd(cadd(b[0], b[2]), cadd(b[1], b[3]));
    const V bd = cadd(csub(b[0], b[2]), mul_si<SGN>(csub(b[1], b[3])));
    const V cs = cadd(cadd(c[0], c[2]), cadd(c[1], c[3]));
    const V cd = cadd(csub(c[0], c[2]), mul_si<SGN>(csub(c[1], c[3])));
    a[0] = bs;
    a[1] = cs;
    a[2] = bd;
    a[3] = cd;
  }
};

// Strided-axis pass (y or z): tile = B consecutive kx x NC components, lines
// along the axis with element stride sp; blockIdx.y selects the outer index.
// ---------------------------------------------------------------------------
template <typename V> struct SArgs {
  const V* src;
  V* dst;
  const V* tw;
  const V* tws;     // stage-major twiddles (k_zconv_ct)
  const int* pos;   // perm[n] of the axis
  AxisPlan ax;
  int n_in, n_out;  // points loaded (rest zero) / stored
  int nbx;          // extent of the contiguous kx dimension
  int B;            // kx per tile
  long long so_src, sp_src, so_dst, sp_dst;
  // k_zconv only: folded G[((ky'*F3 + kz')*F1 + kx')*6 + c]
  const V* G;
  int F1, F3, L1, L2;
  // right-hand side blockIdx.z: element offsets per RHS, and a per-RHS skip
  // flag (when non-null and skip[z*skip_stride] != 0 the pass does nothing)
  long long zs_src, zs_dst;
  const int* skip;
  int skip_stride;
  // packed launches: tiles along kx (the grid x extent) and threads of the
  // tile shape
  int gx, nthr;
  // k_zconv_der only: the tensor is a derivative along axis dpar (1 x, 2 y,
  // 3 z), odd under that axis's mirror on top of the Green tensor's parity
  int dpar;
  // y passes of a sparse box (null: every row): rows[k] bit j is set when box
  // line (j, k) holds a dipole.  P2 reads only those rows of A (the x pass
  // writes no others) and P4 writes only those (the gathering x pass reads no
  // others); needs ny <= 64.
  const unsigned long long* rows;
};

// Row p of a tile under the row mask rm (all ones: every row; a mask only
// ever comes with p < 64).
__device__ __forceinline__ bool row_on(unsigned long long rm, int p) {
  return p >= 64 || ((rm >> p) & 1ull);
}

__device__ __forceinline__ int fold_idx(int k, int L, int& sgn) {
  if (2 * k > L) {
    sgn = -1;
    return L - k;
  }
  sgn = 1;
  return k;
}

// Block ids are dispatched round-robin over the 8 XCDs (id % 8), each with its
// own L2.  Remap the linear id so that the blocks of one XCD take a
// contiguous range of logical tiles: neighbouring tiles, which share cache
// lines at their edges, then meet in the same L2.  A bijection of [0, total).
__device__ __forceinline__ int xcd_remap(int lin, int total) {
  const int q = total >> 3, r = total & 7;
  const int xcd = lin & 7, j = lin >> 3;
  return xcd * q + min(xcd, r) + j;
}

// v[0..2] = G v with G symmetric: xx, xy, xz, yy, yz, zz = g[0..5]; nxy, nxz,
// nyz negate the off-diagonal entries.
template <typename V>
__device__ __forceinline__ void sym_product(V* v, const V (&g)[6], bool nxy, bool nxz, bool nyz) {
  const V gxy = nxy ? mk(-g[1].x, -g[1].y) : g[1];
  const V gxz = nxz ? mk(-g[2].x, -g[2].y) : g[2];
  const V gyz = nyz ? mk(-g[4].x, -g[4].y) : g[4];
  const V ex = v[0], ey = v[1], ez = v[2];
  v[0] = cadd(cadd(cmul(g[0], ex), cmul(gxy, ey)), cmul(gxz, ez));
  v[1] = cadd(cadd(cmul(gxy, ex), cmul(g[3], ey)), cmul(gyz, ez));
  v[2] = cadd(cadd(cmul(gxz, ex), cmul(gyz, ey)), cmul(g[5], ez));
}

// Tile lg (over gridDim.x * gridDim.y) and right-hand side rhs (over
// gridDim.z) of this block.  The right-hand sides are taken in groups of RG
// (8, 4, 2 or 1, dividing gridDim.z) and the blocks are ordered (group,
// tile, rhs in group), in XCD order: the RG blocks of one tile run
// back to back on one XCD and read that tile's Green columns from its L2,
// and neighbouring tiles of one right-hand side stay RG blocks apart.  With
// one right-hand side this is xcd_remap over the tiles.
__device__ __forceinline__ void rhs_group_tile(int& lg, int& rhs) {
  const int T = gridDim.x * gridDim.y, R = gridDim.z;
  const int g = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), T * R);
  const int RG = R % 8 == 0 ? 8 : R % 4 == 0 ? 4 : R % 2 == 0 ? 2 : 1;
  const int q = g / RG;
  lg = q % T;
  rhs = (q / T) * RG + (g - q * RG);
}

// LDS layout of every FFT pass: twiddles (L values), positions (L ints),
// then the tile (L * NL values).
template <typename V> __device__ __forceinline__ V* tile_base(V* tw, int L) {
  const size_t off = ((size_t)L * sizeof(V) + (size_t)L * sizeof(int) + 15) & ~(size_t)15;
  return reinterpret_cast<V*>(reinterpret_cast<char*>(tw) + off);
}
// Row stride (in values) of the matvec x passes' LDS tile of NL lines.  The
// scatter and gather phases put a wavefront's 64 lanes on 64 different rows
// (permuted positions, same line), so the rows' bank offsets are multiples of
// the stride: an even stride of FP32 values (8 B, two banks) shares its
// factor 2^k with the bank count and leaves 64 / 2^k distinct offsets, i.e.
// 2^k-way conflicts.  Wide tiles (16 lines or more: batched launches) are
// therefore given an odd stride, one value of padding when NL is even.  A
// 24-line FP32 tile unpadded spent 44% of its LDS cycles in conflict; the
// 21-line tile of the 256-RHS batch (128^3) padded to 22 still spent 39%.
// Narrow single-RHS tiles (B <= 5) keep NL: padded, they measured no faster.
__host__ __device__ constexpr int x_stride(int NL) { return NL >= 16 && NL % 2 == 0 ? NL + 1 : NL; }
template <typename V> size_t tile_lds_bytes(int L, int NL) {
  return (((size_t)L * sizeof(V) + (size_t)L * sizeof(int) + 15) & ~(size_t)15) +
         (size_t)L * NL * sizeof(V);
}

template <typename V>
__device__ __forceinline__ void load_tables(V* tw, int* pos, const V* gtw, const int* gpos, int L,
                                            int tid, int nthr) {
  for (int e = tid; e < L; e += nthr) {
    tw[e] = gtw[e];
    pos[e] = gpos[e];
  }
}

// ---------------------------------------------------------------------------
// Load rows p < n_in of a strided tile into LDS (zeros for n_in <= p < fill):
// at permuted positions pos[p] (PERM, input of a DIT transform) or at p.
template <bool PERM, typename V>
__device__ __forceinline__ void load_strided(V* buf, const int* pos, const V* src, long long sp,
                                             int fill, int n_in, int NL, int line, int nvalid,
                                             int jt, int JT, unsigned long long rm = ~0ull) {
  const long long ustep = (long long)JT * sp;
  for (int p0 = jt; p0 < fill; p0 += kBatch * JT) {
    const V* s0 = src + (long long)p0 * sp + line;
    V v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int p = p0 + u * JT;
      v[u] = czero<V>();
      if (p < n_in && line < nvalid && row_on(rm, p)) v[u] = s0[u * ustep];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int p = p0 + u * JT;
      if (p < fill) buf[(PERM ? pos[p] : p) * NL + line] = v[u];
    }
  }
  lds_sync();
}

// ---------------------------------------------------------------------------
// LDS-staged transforms with the FFT length L and tile width fixed at
// compile time (the z pass k_zconv_ct, the FP64 x passes k_xpass_ct, the y
// passes k_strided_ct), for the grid sizes of the benchmark workloads (other
// sizes take the runtime-plan kernels): the same algorithm and LDS layout,
// with every butterfly loop of known trip count, every LDS offset inside a
// butterfly an instruction immediate, and a butterfly's twiddles read from
// the stage-major table at one base address.
// ---------------------------------------------------------------------------
struct PlanC {
  int nst;
  int rad[kMaxStages];
};
// The radix plan of make_axis_plan, evaluated at compile time.
constexpr PlanC plan_c(int L) {
  PlanC a{0, {}};
  int r = L, n = 0;
  while (r % 8 == 0 && n < kMaxStages) a.rad[n++] = 8, r /= 8;
  if (r % 4 == 0 && n < kMaxStages) a.rad[n++] = 4, r /= 4;
  if (r % 2 == 0 && n < kMaxStages) {
    if (n > 0 && a.rad[n - 1] == 8 && n + 1 < kMaxStages) {
      a.rad[n - 1] = 4;
      a.rad[n++] = 4;
    } else {
      a.rad[n++] = 2;
    }
    r /= 2;
  }
  while (r % 5 == 0 && n < kMaxStages) a.rad[n++] = 5, r /= 5;
  while (r % 3 == 0 && n < kMaxStages) a.rad[n++] = 3, r /= 3;
  a.nst = r == 1 ? n : -1;
  return a;
}
// Rows a compile-time DIT pass loads: L/2 when its first (pruned) stage has
// an even radix, else L.
constexpr int ct_fill(int L) { return plan_c(L).rad[0] % 2 == 0 ? L / 2 : L; }
constexpr int plan_m(const PlanC& p, int s) {
  int m = 1;
  for (int i = 0; i < s; ++i) m *= p.rad[i];
  return m;
}

// One stage (radix R, span M) of the in-place FFT of stage_ip, for a tile of
// NL lines and JT threads per line.  PRUNE as in stage_ip.
template <int L, int R, int M, int NL, int JT, int SGN, bool DIT, int PRUNE, typename V>
__device__ __forceinline__ void ct_stage(V* buf, const V* __restrict__ tws, int line, int jt) {
  constexpr int NB = L / R;
  constexpr int NIT = (NB + JT - 1) / JT;
  constexpr int STRIDE = M * NL;
  constexpr int DQ = JT / M, DR = JT % M;
  constexpr int DE = (DQ * M * R + DR) * NL, DWRAP = M * (R - 1) * NL;
  constexpr int RIN = PRUNE == 1 ? R / 2 : R, ROUT = PRUNE == 2 ? R / 2 : R;
  int j = jt % M;
  int e0 = ((jt - j) * R + j) * NL + line;
  int tb = (M - 1) + j * (R - 1);
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    if (NB % JT == 0 || i + 1 < NIT || jt + i * JT < NB) {
      V* p = buf + e0;
      V a[R];
#pragma unroll
      for (int q = 0; q < RIN; ++q) a[q] = p[q * STRIDE];
      if constexpr (DIT && PRUNE == 0 && M > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          const V w = tws[tb + q - 1];
          a[q] = SGN > 0 ? cmul(a[q], w) : cmulc(w, a[q]);
        }
      }
      if constexpr (PRUNE == 1)
        DftHalfIn<R, SGN, V>::run(a);
      else if constexpr (PRUNE == 2)
        DftHalfOut<R, SGN, V>::run(a);
      else
        Dft<R, SGN, V>::run(a);
      if constexpr (!DIT && PRUNE == 0 && M > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          const V w = tws[tb + q - 1];
          a[q] = SGN > 0 ? cmul(a[q], w) : cmulc(w, a[q]);
        }
      }
#pragma unroll
      for (int q = 0; q < ROUT; ++q) p[q * STRIDE] = a[q];
    }
    j += DR;
    e0 += DE;
    tb += DR * (R - 1);
    if constexpr (M > 1) {
      if (j >= M) {
        j -= M;
        e0 += DWRAP;
        tb -= M * (R - 1);
      }
    }
  }
  lds_sync();
}

// DIT over stages S..nst-1 (first stage half-input pruned when its radix is even).
template <int L, int NL, int JT, int SGN, int S, typename V>
__device__ __forceinline__ void ct_dit(V* buf, const V* tws, int line, int jt) {
  constexpr PlanC P = plan_c(L);
  if constexpr (S < P.nst) {
    constexpr int R = P.rad[S], M = plan_m(P, S);
    ct_stage<L, R, M, NL, JT, SGN, true, (S == 0 && R % 2 == 0) ? 1 : 0>(buf, tws, line, jt);
    ct_dit<L, NL, JT, SGN, S + 1>(buf, tws, line, jt);
  }
}
// DIF over stages S down to SL (last stage half-output pruned when even).
template <int L, int NL, int JT, int SGN, int S, typename V, int SL = 0>
__device__ __forceinline__ void ct_dif(V* buf, const V* tws, int line, int jt) {
  constexpr PlanC P = plan_c(L);
  if constexpr (S >= SL) {
    constexpr int R = P.rad[S], M = plan_m(P, S);
    ct_stage<L, R, M, NL, JT, SGN, false, (S == 0 && R % 2 == 0) ? 2 : 0>(buf, tws, line, jt);
    ct_dif<L, NL, JT, SGN, S - 1, V, SL>(buf, tws, line, jt);
  }
}

// The global load of a strided tile fused with the first (pruned) DIT stage,
// for plans whose first radix R is even and lines of n_in <= L/2 values.
// That stage's butterfly at LDS positions R b .. R b + R - 1 (span 1) takes
// the signal samples n' + (L/R) q, q < R, where pos[n'] = R b (the first
// digit of a position is its own): n' = u, u < L/R, is loaded straight from
// HBM (rows q < R/2; the rest are zero padding), transformed in registers,
// and written to LDS once.  Saves the load's LDS round trip and barrier.
// ct_fetch_first issues the loads (a thread's FirstRows), ct_first_stage
// transforms and stores them, so a caller may fetch one row while it
// transforms another.
template <int L, int JT, typename V> struct FirstRows {
  static constexpr int R = plan_c(L).rad[0], NB = L / R, NIT = (NB + JT - 1) / JT;
  V a[NIT][R / 2];  // the non-zero inputs of the thread's butterflies
};
template <int L, int JT, typename V>
__device__ __forceinline__ void ct_fetch_first(FirstRows<L, JT, V>& f, const V* src, long long sp,
                                               int n_in, int line, int nvalid, int jt,
                                               unsigned long long rm = ~0ull) {
  using F = FirstRows<L, JT, V>;
  static_assert(F::R % 2 == 0, "fused first stage needs an even first radix");
#pragma unroll
  for (int i = 0; i < F::NIT; ++i) {
    const int u = jt + i * JT;
#pragma unroll
    for (int q = 0; q < F::R / 2; ++q) {
      const int n = u + F::NB * q;
      f.a[i][q] = czero<V>();
      if ((F::NB % JT == 0 || u < F::NB) && n < n_in && line < nvalid && row_on(rm, n))
        f.a[i][q] = src[(long long)n * sp + line];
    }
  }
}
template <int L, int NL, int JT, int SGN, typename V>
__device__ __forceinline__ void ct_first_stage(V* buf, const int* pos, FirstRows<L, JT, V>& f,
                                               int line, int jt) {
  using F = FirstRows<L, JT, V>;
#pragma unroll
  for (int i = 0; i < F::NIT; ++i) {
    const int u = jt + i * JT;
    if (F::NB % JT == 0 || u < F::NB) {
      V b[F::R];
#pragma unroll
      for (int q = 0; q < F::R / 2; ++q) b[q] = f.a[i][q];
      DftHalfIn<F::R, SGN, V>::run(b);
      V* p = buf + pos[u] * NL + line;
#pragma unroll
      for (int q = 0; q < F::R; ++q) p[q * NL] = b[q];
    }
  }
  lds_sync();
}
template <int L, int NL, int JT, int SGN, typename V>
__device__ __forceinline__ void ct_load_first(V* buf, const int* pos, const V* src, long long sp,
                                              int n_in, int line, int nvalid, int jt,
                                              unsigned long long rm = ~0ull) {
  FirstRows<L, JT, V> f;
  ct_fetch_first<L, JT>(f, src, sp, n_in, line, nvalid, jt, rm);
  ct_first_stage<L, NL, JT, SGN>(buf, pos, f, line, jt);
}

// The last (pruned) DIF stage fused with the global store of rows p < n_out
// (<= L/2): the butterfly at positions R b + q (span 1) leaves transform
// index n' + (L/R) q at R b + q, where pos[n'] = R b; its R/2 kept outputs go
// from registers to HBM.  Saves the store's LDS round trip.  Ends with a
// barrier, after which the tile may be reused.
template <int L, int NL, int JT, int SGN, typename V>
__device__ __forceinline__ void ct_last_store(const V* buf, const int* pos, V* dst, long long sp,
                                              int n_out, int line, int nvalid, int jt,
                                              unsigned long long rm = ~0ull) {
  constexpr PlanC P = plan_c(L);
  constexpr int R = P.rad[0], NB = L / R, NIT = (NB + JT - 1) / JT;
  static_assert(R % 2 == 0, "fused last stage needs an even first radix");
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    const int u = jt + i * JT;
    if (NB % JT == 0 || u < NB) {
      V a[R];
      const V* p = buf + pos[u] * NL + line;
#pragma unroll
      for (int q = 0; q < R; ++q) a[q] = p[q * NL];
      DftHalfOut<R, SGN, V>::run(a);
      if (line < nvalid) {
#pragma unroll
        for (int q = 0; q < R / 2; ++q) {
          const int n = u + NB * q;
          if (n < n_out && row_on(rm, n)) dst[(long long)n * sp + line] = a[q];
        }
      }
    }
  }
  lds_sync();
}

}  // namespace ifdda
