// ifdda_ystride.hip -- the y passes of the matvec (P2 FFT-y, P4 IFFT-y;
// produitfftmatvect3.f:64-73, :92-101) with the FFT length and the tile width
// fixed at compile time, for the 128^3, 192^3 and 200^3 grids.
//
// Same tile and algorithm as k_strided: B consecutive kx of the three
// components (NL = 3B lines, contiguous in HBM) of one z plane, the lines
// along y held in LDS as [point][line]; P2 loads the ny rows into digit-
// reversed positions and runs decimation in time with its first stage
// pruned (rows >= L/2 are zero padding; that stage runs on the loaded
// registers), P4 loads all L rows and runs decimation in frequency with its
// last stage pruned (rows >= ny are not kept; that stage's outputs go from
// registers to HBM); in a sparse box both touch only the rows of A whose box
// line holds a dipole (SArgs::rows).  Here the radix plan, trip counts and LDS offsets are constants and
// each butterfly's twiddles come from the stage-major table at one base
// address (the stages of the compile-time z pass).  On the 200^3 grid the
// runtime-plan y pass spends ~1.6x the VALU instructions per line of those
// stages, and in FP32 half of its time is VALU issue.
#include "ifdda_fftdev.h"
#include "ifdda_internal.h"

namespace ifdda {

// Launched with (256 / 3B) * 3B threads on the grid of k_strided (tiles of B
// kx by outer index in y, right-hand sides in z); P2 (SGN > 0) needs
// 2 n_in <= L, P4 (SGN < 0) 2 n_out <= L.
template <typename V, int L, int B, int SGN>
__global__ void __launch_bounds__(256) k_strided_ct(SArgs<V> a) {
  constexpr int NL = 3 * B, JT = 256 / NL, NTHR = NL * JT;
  constexpr PlanC P = plan_c(L);
  static_assert(P.nst > 0, "unsupported compile-time y plan");
  if (a.skip && a.skip[blockIdx.z * a.skip_stride]) return;
  const int lg = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const V* const zsrc = a.src + blockIdx.z * a.zs_src;
  V* const zdst = a.dst + blockIdx.z * a.zs_dst;
  extern __shared__ double4 smem_d4[];
  V* tws = reinterpret_cast<V*>(smem_d4);
  int* pos = reinterpret_cast<int*>(tws + L);
  V* buf = tile_base(tws, L);
  const int tid = threadIdx.x;
  const int line = tid % NL, jt = tid / NL;
  const int o = lg / gridDim.x;
  const int kx0 = (lg - o * (int)gridDim.x) * B;
  const int nvalid = min(B, a.nbx - kx0) * 3;
  // rows of plane o that hold dipoles (a plane without any has nothing to keep)
  const unsigned long long rm = a.rows ? a.rows[o] : ~0ull;
  if (SGN < 0 && rm == 0) return;
  load_tables(tws, pos, a.tws, a.pos, L, tid, NTHR);
  lds_sync();
  const V* src = zsrc + (long long)o * a.so_src + (long long)kx0 * 3;
  V* dst = zdst + (long long)o * a.so_dst + (long long)kx0 * 3;
  if constexpr (SGN > 0) {  // the load fused with the first stage
    ct_load_first<L, NL, JT, +1>(buf, pos, src, a.sp_src, a.n_in, line, nvalid, jt, rm);
    ct_dit<L, NL, JT, +1, 1>(buf, tws, line, jt);
    if (line < nvalid) {
      V* d = dst + line + (long long)jt * a.sp_dst;
      const long long step = (long long)JT * a.sp_dst;
      for (int p = jt; p < a.n_out; p += JT, d += step) *d = buf[p * NL + line];
    }
  } else {
    load_strided<false>(buf, pos, src, a.sp_src, L, a.n_in, NL, line, nvalid, jt, JT);
    // the last stage fused with the store
    ct_dif<L, NL, JT, -1, P.nst - 1, V, 1>(buf, tws, line, jt);
    ct_last_store<L, NL, JT, -1>(buf, pos, dst, a.sp_dst, a.n_out, line, nvalid, jt, rm);
  }
}

template <typename V> using StridedKernel = void (*)(SArgs<V>);
template <typename V, int L, int SGN> static StridedKernel<V> strided_ct_width(int B) {
  switch (B) {
    case 2: return k_strided_ct<V, L, 2, SGN>;
    case 3: return k_strided_ct<V, L, 3, SGN>;
    case 4: return k_strided_ct<V, L, 4, SGN>;
    case 5: return k_strided_ct<V, L, 5, SGN>;
    case 6: return k_strided_ct<V, L, 6, SGN>;
    case 8: return k_strided_ct<V, L, 8, SGN>;
    default: return nullptr;
  }
}
template <typename V, int SGN> static StridedKernel<V> strided_ct_for(int L, int B) {
  if (L == 200) return strided_ct_width<V, 200, SGN>(B);
  if (L == 128) return strided_ct_width<V, 128, SGN>(B);
  if (L == 192) return strided_ct_width<V, 192, SGN>(B);
  return nullptr;
}

bool has_strided_ct(int L, int B, int prec, int sgn) {
  if (prec == 64)
    return (sgn > 0 ? strided_ct_for<double2, +1>(L, B) : strided_ct_for<double2, -1>(L, B)) != nullptr;
  return (sgn > 0 ? strided_ct_for<float2, +1>(L, B) : strided_ct_for<float2, -1>(L, B)) != nullptr;
}

template <typename V>
static bool launch_strided_ct_impl(const SArgs<V>& a, int L, int B, int sgn, dim3 grid,
                                   size_t lds, hipStream_t s) {
  StridedKernel<V> k = sgn > 0 ? strided_ct_for<V, +1>(L, B) : strided_ct_for<V, -1>(L, B);
  if (!k || !a.tws || lds > 65536) return false;
  const int nl = 3 * B;
  hipLaunchKernelGGL(k, grid, dim3((256 / nl) * nl), lds, s, a);
  return true;
}
bool launch_strided_ct(const SArgs<double2>& a, int L, int B, int sgn, dim3 grid, size_t lds,
                       hipStream_t s) {
  return launch_strided_ct_impl(a, L, B, sgn, grid, lds, s);
}
bool launch_strided_ct(const SArgs<float2>& a, int L, int B, int sgn, dim3 grid, size_t lds,
                       hipStream_t s) {
  return launch_strided_ct_impl(a, L, B, sgn, grid, lds, s);
}

}  // namespace ifdda
