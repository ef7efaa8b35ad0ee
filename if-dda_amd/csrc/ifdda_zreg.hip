// ifdda_zreg.hip -- P3 (FFT-z . G . IFFT-z, produitfftmatvect3.f:62-107) with
// register-staged transforms, for z lengths L = 8 Q (Q = 25, 16, 24: the
// 200^3, 128^3 and 192^3 grids).
//
// The z line of a tile holds nz <= L/2 values, and only kz' < nz is kept
// after the inverse.  With n = Q n1 + n2 and k = k1 + 8 k2 (n1, k1 < 8,
// n2, k2 < Q) the length-L transforms split into 8- and Q-point ones:
//   forward  X[k1 + 8 k2] = sum_n2 W_Q^(n2 k2) W_L^(n2 k1) sum_(n1<4) x[Q n1 + n2] W_8^(n1 k1)
//   inverse  y[Q n1 + n2] = sum_k1 W_8^(-n1 k1) W_L^(-n2 k1) sum_k2 Y[k1 + 8 k2] W_Q^(-n2 k2)
// for n1 < 4 (W_N = exp(2 pi i / N), the sign of the other passes).  The
// 8-point steps see four non-zero inputs (forward) and keep four outputs
// (inverse); each Q-point step is one thread's whole line segment, done in
// registers as P x M with constant twiddles (25 = 5 x 5, 16 = 4 x 4,
// 24 = 8 x 3).  A tile is B consecutive kx of the three components
// (NL = 3B lines) of one ky row and then of its mirror row L2 - ky (the
// folded Green octant serves both), with 8 NL threads:
//   A  (n2, line): 4 rows from HBM, 8-point DFT, W_L twiddles -> LDS [k1][n2]
//   B  (k1, line): Q values from LDS, Q-point DFT -> LDS [k][line] (natural k)
//   G  the symmetric 3x3 product with the folded Green octant at kz and
//      L - kz, in LDS (the three components of a kx are three lines)
//   B' (k1, line): Q values, inverse Q-point DFT, W_L twiddles -> LDS [k1][n2]
//   A' (n2, line): 8 values, the four kept outputs of the 8-point DFT -> HBM
// Four LDS round trips per line and direction pair, against eight stage
// passes (and their barriers) in k_zconv_ct, whose butterflies are spread
// over 17 threads per line with uneven trip counts.
#include "ifdda_fftdev.h"
#include "ifdda_internal.h"

namespace ifdda {

// -------------------------------------------- compile-time roots of unity
struct CSd {
  double c, s;
};
// sin and cos of x, |x| <= pi/4, by their Taylor series (18 terms: below
// double rounding there)
constexpr double ct_sin(double x) {
  double t = x, s = x;
  for (int i = 1; i < 18; ++i) {
    t *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    s += t;
  }
  return s;
}
constexpr double ct_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int i = 1; i < 18; ++i) {
    t *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    s += t;
  }
  return s;
}
// exp(2 pi i a / n): a reduced to the nearest quarter turn k, the rest
// (pi/2) m / n with |m| <= n/2 from the series
constexpr CSd ct_root(int a, int n) {
  a %= n;
  if (a < 0) a += n;
  const int k = (4 * a + n / 2) / n;  // nearest quarter turn
  const int m = 4 * a - k * n;        // |m| <= n/2
  const double t = 1.57079632679489661923132169163975144 * (double)m / (double)n;
  const double c = ct_cos(t), s = ct_sin(t);
  switch (k & 3) {
    case 0: return {c, s};
    case 1: return {-s, c};
    case 2: return {-c, -s};
    default: return {s, -c};
  }
}
template <int N> struct RootTab {
  CSd v[N];
  constexpr RootTab() : v() {
    for (int e = 0; e < N; ++e) v[e] = ct_root(e, N);
  }
};
template <int N> inline constexpr RootTab<N> kRoots{};

// x * W_N^(SGN e) with a compile-time root
template <int N, int SGN, typename V> __device__ __forceinline__ V mul_root(V x, int e) {
  using S = decltype(V::x);
  const CSd w = kRoots<N>.v[e % N];
  const S wc = S(w.c), ws = S(SGN > 0 ? w.s : -w.s);
  return mk(x.x * wc - x.y * ws, x.x * ws + x.y * wc);
}

// The P and M of the Q-point register DFT (Q = P M)
template <int Q> struct RegPlan;
template <> struct RegPlan<25> { static constexpr int P = 5, M = 5; };
template <> struct RegPlan<16> { static constexpr int P = 4, M = 4; };
template <> struct RegPlan<24> { static constexpr int P = 8, M = 3; };

// In-place Q-point DFT of values held by one thread, natural order in and
// out: c[k] = sum_n c[n] W_Q^(SGN n k).  n = M p + m, k = kp + P km:
// P-point DFTs over p for each m, twiddles W_Q^(m kp), M-point DFTs over m.
template <int Q, int SGN, typename V> __device__ __forceinline__ void reg_dft(V (&c)[Q]) {
  constexpr int P = RegPlan<Q>::P, M = RegPlan<Q>::M;
  V d[M][P];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    V a[P];
#pragma unroll
    for (int p = 0; p < P; ++p) a[p] = c[M * p + m];
    Dft<P, SGN, V>::run(a);
#pragma unroll
    for (int kp = 0; kp < P; ++kp) d[m][kp] = (m == 0 || kp == 0) ? a[kp] : mul_root<Q, SGN>(a[kp], m * kp);
  }
#pragma unroll
  for (int kp = 0; kp < P; ++kp) {
    V b[M];
#pragma unroll
    for (int m = 0; m < M; ++m) b[m] = d[m][kp];
    Dft<M, SGN, V>::run(b);
#pragma unroll
    for (int km = 0; km < M; ++km) c[kp + P * km] = b[km];
  }
}

// LDS bytes of k_zconv_rg<V, L, B>: the twiddles exp(2 pi i e / L), then the
// tile (L x 3B values)
template <typename V> size_t zreg_lds(int L, int B) {
  return (((size_t)L * sizeof(V) + 15) & ~(size_t)15) + (size_t)L * 3 * B * sizeof(V);
}

// Launched with 24 B threads over the grid of k_zconv_ct (tiles of B kx by
// ky' <= L2/2, right-hand sides in z); requires 2 n_in <= L, 2 n_out <= L.
template <typename V, int L, int B>
__global__ void __launch_bounds__(24 * B, 3) k_zconv_rg(SArgs<V> a) {
  constexpr int Q = L / 8, NL = 3 * B, NT = 8 * NL, F3 = L / 2 + 1;
  constexpr int NA = (Q * NL + NT - 1) / NT;     // step A / A' items per thread
  constexpr int NG = (F3 * B + NT - 1) / NT;     // Green items per thread
  static_assert(L % 8 == 0 && NG <= 8, "unsupported register z pass");
  int lg, rhs;
  rhs_group_tile(lg, rhs);
  if (a.skip && a.skip[rhs * a.skip_stride]) return;
  const V* const zsrc = a.src + rhs * a.zs_src;
  V* const zdst = a.dst + rhs * a.zs_dst;
  extern __shared__ double4 smem_d4[];
  V* tw = reinterpret_cast<V*>(smem_d4);
  V* buf = reinterpret_cast<V*>(reinterpret_cast<char*>(smem_d4) +
                                (((size_t)L * sizeof(V) + 15) & ~(size_t)15));
  const int tid = threadIdx.x;
  const int kyf = lg / gridDim.x;  // 0 .. L2/2
  // tiles of a row in the order 0, gx-1, 1, gx-2, ...: kx and its mirror
  // L1 - kx read the same folded Green column and run side by side
  const int u = lg - kyf * gridDim.x;
  const int kx0 = ((u & 1) ? (int)gridDim.x - 1 - (u >> 1) : (u >> 1)) * B;
  const int nvalid = min(B, a.nbx - kx0) * 3;
  const long long sp_s = a.sp_src, sp_d = a.sp_dst;
  for (int e = tid; e < L; e += NT) tw[e] = a.tw[e];
  const int k1 = tid / NL, lb = tid - k1 * NL;  // step B: (k1, line)
  const int nrow = (kyf == 0 || 2 * kyf == a.L2) ? 1 : 2;
  lds_sync();
  for (int rr = 0; rr < nrow; ++rr) {
    const int ky = rr == 0 ? kyf : a.L2 - kyf;
    const bool sy = rr == 1;  // y odd-parity sign for the mirror row
    const V* src = zsrc + (long long)ky * a.so_src + (long long)kx0 * 3;
    // ---- A: rows Q n1 + n2 (n1 < 4), 8-point DFT of 4 inputs, W_L^(n2 k1)
    {
      V x[NA][8];
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const int it = tid + i * NT, n2 = it / NL, line = it - n2 * NL;
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) {
          const int n = Q * n1 + n2;
          x[i][n1] = czero<V>();
          if ((Q * NL % NT == 0 || it < Q * NL) && n < a.n_in && line < nvalid)
            x[i][n1] = src[(long long)n * sp_s + line];
        }
      }
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const int it = tid + i * NT, n2 = it / NL, line = it - n2 * NL;
        if (Q * NL % NT == 0 || it < Q * NL) {
          DftHalfIn<8, +1, V>::run(x[i]);
#pragma unroll
          for (int q = 1; q < 8; ++q) x[i][q] = cmul(x[i][q], tw[n2 * q]);
#pragma unroll
          for (int q = 0; q < 8; ++q) buf[(q * Q + n2) * NL + line] = x[i][q];
        }
      }
    }
    lds_sync();
    // ---- B: Q-point DFT of line lb's k1 column -> natural kz order
    V c[Q];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) c[n2] = buf[(k1 * Q + n2) * NL + lb];
    reg_dft<Q, +1>(c);
    lds_sync();  // every column read before the tile is overwritten
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) buf[(k1 + 8 * k2) * NL + lb] = c[k2];
    lds_sync();
    // ---- G: folded Green entries (kz', kx'), applied at kz' and L - kz'
    {
      V gr[NG][6];
      int gx[NG];  // bit0 valid, bit1 x odd-parity sign
#pragma unroll
      for (int it = 0; it < NG; ++it) {
        const int idx = tid + it * NT;
        gx[it] = 0;
        if (idx < F3 * B) {
          const int kzf = idx / B, kxl = idx - kzf * B;
          const int kx = kx0 + kxl;
          if (kx < a.nbx) {
            int sx;
            const int kxf = fold_idx(kx, a.L1, sx);
            const V* g = a.G + ((long long)(kyf * F3 + kzf) * a.F1 + kxf) * 6;
#pragma unroll
            for (int q = 0; q < 6; ++q) gr[it][q] = g[q];
            gx[it] = 1 | ((sx < 0) << 1);
          }
        }
      }
#pragma unroll
      for (int it = 0; it < NG; ++it) {
        if (!(gx[it] & 1)) continue;
        const int idx = tid + it * NT;
        const int kzf = idx / B, kxl = idx - kzf * B;
        const bool sxn = gx[it] & 2;
        sym_product(buf + kzf * NL + kxl * 3, gr[it], sxn != sy, sxn, sy);
        if (kzf != 0 && 2 * kzf != L)
          sym_product(buf + (L - kzf) * NL + kxl * 3, gr[it], sxn != sy, !sxn, !sy);
      }
    }
    lds_sync();
    // ---- B': inverse Q-point DFT of the k1 column, W_L^(-n2 k1)
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) c[k2] = buf[(k1 + 8 * k2) * NL + lb];
    reg_dft<Q, -1>(c);
    if (k1 != 0) {
#pragma unroll
      for (int n2 = 1; n2 < Q; ++n2) c[n2] = cmulc(tw[k1 * n2], c[n2]);
    }
    lds_sync();  // every k column read before the tile is overwritten
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) buf[(k1 * Q + n2) * NL + lb] = c[n2];
    lds_sync();
    // ---- A': outputs n1 < 4 of the 8-point inverse DFT -> rows Q n1 + n2
    {
      V* dst = zdst + (long long)ky * a.so_dst + (long long)kx0 * 3;
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const int it = tid + i * NT, n2 = it / NL, line = it - n2 * NL;
        if (Q * NL % NT == 0 || it < Q * NL) {
          V f[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) f[q] = buf[(q * Q + n2) * NL + line];
          DftHalfOut<8, -1, V>::run(f);
          if (line < nvalid) {
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) {
              const int n = Q * n1 + n2;
              if (n < a.n_out) dst[(long long)n * sp_d + line] = f[n1];
            }
          }
        }
      }
    }
    lds_sync();  // the next row reuses the tile
  }
}

template <typename V> using ZregKernel = void (*)(SArgs<V>);
template <typename V> static ZregKernel<V> zreg_for(int L, int B) {
  if (L == 200 && B == 4) return k_zconv_rg<V, 200, 4>;
  if (L == 200 && B == 5) return k_zconv_rg<V, 200, 5>;
  if (L == 200 && B == 8) return k_zconv_rg<V, 200, 8>;
  if (L == 128 && B == 5) return k_zconv_rg<V, 128, 5>;
  if (L == 128 && B == 8) return k_zconv_rg<V, 128, 8>;
  if (L == 192 && B == 5) return k_zconv_rg<V, 192, 5>;
  if (L == 192 && B == 8) return k_zconv_rg<V, 192, 8>;
  return nullptr;
}

// Built for (L, B) and within the 64 KB of LDS a workgroup is given without
// opting in (FP64 tiles of 8 lines at L = 200 are not: 80 KB).
bool has_zreg(int L, int B, int prec) {
  if (prec == 64) return zreg_for<double2>(L, B) != nullptr && zreg_lds<double2>(L, B) <= 65536;
  return zreg_for<float2>(L, B) != nullptr && zreg_lds<float2>(L, B) <= 65536;
}

template <typename V>
static bool launch_zreg_impl(const SArgs<V>& a, int L, int B, dim3 grid, hipStream_t s) {
  ZregKernel<V> k = zreg_for<V>(L, B);
  const size_t lds = zreg_lds<V>(L, B);
  if (!k || lds > 65536) return false;
  hipLaunchKernelGGL(k, grid, dim3(24 * B), lds, s, a);
  return true;
}
bool launch_zreg(const SArgs<double2>& a, int L, int B, dim3 grid, hipStream_t s) {
  return launch_zreg_impl(a, L, B, grid, s);
}
bool launch_zreg(const SArgs<float2>& a, int L, int B, dim3 grid, hipStream_t s) {
  return launch_zreg_impl(a, L, B, grid, s);
}

}  // namespace ifdda
