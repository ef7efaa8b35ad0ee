// ifdda_farfield.hip -- outputs computed from solved fields, on the device:
//   cross sections Cext, Cabs, Csca       cdmlib.f:2725-2753
//   far field on the k-plane (quick)      diffractefft2d.f:1-183
//   Csca and asymmetry g from it          computegcfft2d.f:1-54
// for R right-hand sides at once (RHS r at offset r * 3 nd).
//
// The reference zero-pads every z-plane of the dipole moments to an
// nfft2d x nfft2d grid, takes a full 2D FFT and keeps the (2 imax + 1)^2
// bins inside the light cone.  Here only those bins are evaluated, as
// separable sums with phase tables that reproduce the reference's index
// mapping (tabfft2): x sums per (z-plane, y-row) (k_ff_rows), then y and z
// sums per (kx, ky) with the transverse projection and the lattice-origin
// phase (k_ff_cols).  Work O(N (2 imax + 1) + nz ny (2 imax + 1)^2) per RHS
// instead of nz nfft2d^2 log nfft2d.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ifdda.h"
#include "ifdda_internal.h"

namespace ifdda {

namespace {

constexpr int kFfThreads = 256;
constexpr int kXsBlocks = 256;

// Deterministic block sum of NR doubles (fixed shuffle and LDS order).
template <int NR>
__device__ __forceinline__ void block_sum(double (&v)[NR], double* out) {
  for (int off = 32; off > 0; off >>= 1)
    for (int q = 0; q < NR; ++q) v[q] += __shfl_down(v[q], off, 64);
  __shared__ double sh[NR][kFfThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < NR; ++q) sh[q][wid] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < NR; ++q) {
      double s = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[q][w];
      out[q] = s;
    }
}

// Per-block partial sums of Im(E0* . p), Im(p . Eloc*), |p|^2 (cdmlib.f:2735-2741).
__global__ void __launch_bounds__(kFfThreads) k_xsec_part(const double2* E0, const double2* FF,
                                                          const double2* FL, long long nd,
                                                          double* part) {
  const long long o = (long long)blockIdx.y * 3 * nd;
  double v[3] = {0.0, 0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < 3 * nd;
       i += (long long)gridDim.x * blockDim.x) {
    const double2 e = E0[o + i], p = FF[o + i], l = FL[o + i];
    v[0] += e.x * p.y - e.y * p.x;  // Im(conj(e) p)
    v[1] += p.y * l.x - p.x * l.y;  // Im(p conj(l))
    v[2] += p.x * p.x + p.y * p.y;
  }
  double s[3];
  block_sum<3>(v, s);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) part[((long long)blockIdx.y * 3 + q) * gridDim.x + blockIdx.x] = s[q];
}

// out[3 r + {0, 1, 2}] = Cext, Cabs, Csca of RHS r (cdmlib.f:2743-2753).
__global__ void __launch_bounds__(kFfThreads) k_xsec_fin(const double* part, int nblk, double k0,
                                                         const double* I0, double* out) {
  const int r = blockIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
    for (int q = 0; q < 3; ++q) v[q] += part[((long long)r * 3 + q) * nblk + b];
  double s[3];
  block_sum<3>(v, s);
  if (threadIdx.x == 0) {
    const double pi = 3.141592653589793238462643383279502884;
    const double cext = 4.0 * pi * k0 * s[0] / I0[r];
    double cabs = 4.0 * pi * k0 * (s[1] - 2.0 / 3.0 * k0 * k0 * k0 * s[2]) / I0[r];
    const double csca = cext - cabs;
    if (cabs <= cext * 1e-12) cabs = 0.0;
    out[3 * r] = cext;
    out[3 * r + 1] = cabs;
    out[3 * r + 2] = csca;
  }
}

// x sums: X[r][i][k][b][c] = sum_a FF(a, b, k, c) px[i][a] for one (k, b) row
// per block (blockIdx.x = k ny + b, blockIdx.y = r).  Moments come through
// the dipole map (null: dense box).
__global__ void __launch_bounds__(kFfThreads) k_ff_rows(const double2* FF, long long nd,
                                                        const int32_t* tab, int nx, int ny, int nz,
                                                        const double2* px, int n, double2* X) {
  extern __shared__ double2 row[];  // nx * 3
  const int kb = blockIdx.x, r = blockIdx.y;
  const double2* F = FF + (long long)r * 3 * nd;
  for (int e = threadIdx.x; e < nx * 3; e += blockDim.x) {
    const int a = e / 3, c = e - a * 3;
    const long long cell = (long long)kb * nx + a;
    const long long dip = tab ? (long long)tab[cell] - 1 : cell;
    row[e] = dip >= 0 ? F[dip * 3 + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const long long plane = (long long)nz * ny;
  for (int e = threadIdx.x; e < n * 3; e += blockDim.x) {
    const int i = e / 3, c = e - i * 3;
    const double2* p = px + (long long)i * nx;
    double2 s = make_double2(0.0, 0.0);
    for (int a = 0; a < nx; ++a) s = cadd(s, cmul(row[a * 3 + c], p[a]));
    X[(((long long)r * n + i) * plane + kb) * 3 + c] = s;
  }
}

// y and z sums for one (kx, ky) bin per wavefront (blockIdx.x = ii + n jj,
// blockIdx.y = r), then the transverse projection and the origin phase.
// MODE 0, far field (diffractefft2d.f:139-176): both half spaces, bins
// outside the light cone zero.  MODE 1, lens pupil (diffractefft2dlens.f:
// 137-178): the side dside only, bins with kx^2 + ky^2 < na2, weight
// k0^2 / (-2 pi i kz); pz then holds exp(-i dside kz z).
template <int MODE>
__global__ void __launch_bounds__(64) k_ff_cols(const double2* X, int ny, int nz, int n,
                                                const double2* py, const double2* pz,
                                                const double2* pv, double dk, double k0,
                                                double na2, double dside, double2* Epos,
                                                double2* Eneg) {
  const int ij = blockIdx.x, r = blockIdx.y;
  const int ii = ij % n, jj = ij / n;
  const int imax = (n - 1) / 2;
  const double kx = (ii - imax) * dk, ky = (jj - imax) * dk;
  const long long no = (long long)r * 3 * n * n;
  const double kz2 = k0 * k0 - kx * kx - ky * ky;
  const bool inside = MODE == 0 ? kz2 > 0.0 : kx * kx + ky * ky < na2;
  if (!inside) {
    if (threadIdx.x < 3) {
      Epos[no + (long long)threadIdx.x * n * n + ij] = make_double2(0.0, 0.0);
      if (MODE == 0) Eneg[no + (long long)threadIdx.x * n * n + ij] = make_double2(0.0, 0.0);
    }
    return;
  }
  const long long plane = (long long)nz * ny;
  const double2* Xi = X + ((long long)r * n + ii) * plane * 3;
  const double2* pyj = py + (long long)jj * ny;
  const double2* pzij = pz + (long long)ij * nz;
  double2 sp[3], sn[3];
  for (int c = 0; c < 3; ++c) sp[c] = sn[c] = make_double2(0.0, 0.0);
  for (long long p = threadIdx.x; p < plane; p += 64) {
    const int k = (int)(p / ny), b = (int)(p - (long long)k * ny);
    const double2 w = cmul(pyj[b], pzij[k]);     // e^{-i ky y} e^{-i kz z}
    const double2* x = Xi + p * 3;
    if (MODE == 0) {
      const double2 wn = cmul(pyj[b], make_double2(pzij[k].x, -pzij[k].y));  // e^{+i kz z}
      for (int c = 0; c < 3; ++c) {
        const double2 v = x[c];
        sp[c] = cadd(sp[c], cmul(v, w));
        sn[c] = cadd(sn[c], cmul(v, wn));
      }
    } else {
      for (int c = 0; c < 3; ++c) sp[c] = cadd(sp[c], cmul(x[c], w));
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int c = 0; c < 3; ++c) {
      sp[c].x += __shfl_down(sp[c].x, off, 64);
      sp[c].y += __shfl_down(sp[c].y, off, 64);
      sn[c].x += __shfl_down(sn[c].x, off, 64);
      sn[c].y += __shfl_down(sn[c].y, off, 64);
    }
  if (threadIdx.x != 0) return;
  const double kz = sqrt(kz2);
  const double2 v = pv[ij];
  if (MODE == 1) {
    // weight k0^2 / (-2 pi i kz) = i k0^2 / (2 pi kz)
    const double2 f = cmul(v, make_double2(0.0, k0 * k0 / (6.283185307179586476925286766559 * kz)));
    const double kk[3] = {kx, ky, dside * kz};
    double2 ct = make_double2(0.0, 0.0);
    for (int c = 0; c < 3; ++c) ct = cadd(ct, cscal(sp[c], kk[c] / k0));
    for (int c = 0; c < 3; ++c)
      Epos[no + (long long)c * n * n + ij] = cmul(csub(sp[c], cscal(ct, kk[c] / k0)), f);
    return;
  }
  const double kv[3] = {kx, ky, kz};
  // k = (kx, ky, +kz) for the upper half space, (kx, ky, -kz) for the lower
  for (int h = 0; h < 2; ++h) {
    const double2* s = h == 0 ? sp : sn;
    const double sg = h == 0 ? 1.0 : -1.0;
    const double kk[3] = {kv[0], kv[1], sg * kv[2]};
    double2 ct = make_double2(0.0, 0.0);
    for (int c = 0; c < 3; ++c) ct = cadd(ct, cscal(s[c], kk[c] / k0));
    double2* E = h == 0 ? Epos : Eneg;
    for (int c = 0; c < 3; ++c)
      E[no + (long long)c * n * n + ij] = cmul(csub(s[c], cscal(ct, kk[c] / k0)), v);
  }
}

// out[2 r + {0, 1}] = Csca, g (computegcfft2d.f:14-52).
__global__ void __launch_bounds__(kFfThreads) k_gc(const double2* Epos, const double2* Eneg, int n,
                                                   double dk, double k0, const double* I0,
                                                   const double* ux, double* out) {
  const int r = blockIdx.x;
  const int imax = (n - 1) / 2;
  const long long no = (long long)r * 3 * n * n;
  const double* u = ux + 3 * r;
  double v[2] = {0.0, 0.0};
  for (int ij = threadIdx.x; ij < n * n; ij += blockDim.x) {
    const int ii = ij % n, jj = ij / n;
    const double kx = (ii - imax) * dk, ky = (jj - imax) * dk;
    const double kz2 = k0 * k0 - kx * kx - ky * ky;
    if (!(kz2 > 0.0)) continue;
    const double kz = sqrt(kz2);
    for (int h = 0; h < 2; ++h) {
      const double2* E = h == 0 ? Epos : Eneg;
      double emod = 0.0;
      for (int c = 0; c < 3; ++c) {
        const double2 e = E[no + (long long)c * n * n + ij];
        emod += e.x * e.x + e.y * e.y;
      }
      const double w = dk * dk * emod / kz;
      v[0] += w;
      v[1] += w * (kx * u[0] + ky * u[1] + (h == 0 ? kz : -kz) * u[2]) / k0;
    }
  }
  double s[2];
  block_sum<2>(v, s);
  if (threadIdx.x == 0) {
    out[2 * r] = s[0] * k0 * k0 * k0 / I0[r];
    out[2 * r + 1] = s[1] / s[0];
  }
}

bool dev_ensure(ifdda_ctx* c, size_t bytes) {
  if (bytes <= c->ffcap) return true;
  if (c->ffbuf) (void)hipFree(c->ffbuf);
  c->ffbuf = nullptr;
  c->ffcap = 0;
  if (!hip_ok(hipMalloc(&c->ffbuf, bytes), "hipMalloc far-field workspace")) return false;
  c->ffcap = bytes;
  return true;
}

}  // namespace

int cross_sections_dev(ifdda_ctx* c, int R, const double2* E0, const double2* FF,
                       const double2* FL, double k0, const double* I0, double* out,
                       hipStream_t s) {
  const long long nd = c->nd;
  const size_t need = sizeof(double) * ((size_t)R * 3 * kXsBlocks + (size_t)R * 4);
  if (!dev_ensure(c, need)) return IFDDA_ERR_HIP;
  double* part = (double*)c->ffbuf;
  double* dI0 = part + (size_t)R * 3 * kXsBlocks;
  double* dout = dI0 + R;
  if (!hip_ok(hipMemcpyAsync(dI0, I0, sizeof(double) * R, hipMemcpyHostToDevice, s), "copy I0"))
    return IFDDA_ERR_HIP;
  hipLaunchKernelGGL(k_xsec_part, dim3(kXsBlocks, R), dim3(kFfThreads), 0, s, E0, FF, FL, nd, part);
  hipLaunchKernelGGL(k_xsec_fin, dim3(R), dim3(kFfThreads), 0, s, (const double*)part, kXsBlocks,
                     k0, (const double*)dI0, dout);
  if (!hip_ok(hipMemcpyAsync(out, dout, sizeof(double) * 3 * R, hipMemcpyDeviceToHost, s),
              "copy cross sections") ||
      !hip_ok(hipStreamSynchronize(s), "cross sections"))
    return IFDDA_ERR_HIP;
  return IFDDA_OK;
}

// imax and deltak as diffractefft2d.f:24-72; returns "" or the reference's message.
std::string farfield_geometry(int nfft2d, double k0, double d, int* imax, double* dk) {
  *dk = 2.0 * 3.141592653589793238462643383279502884 / ((double)nfft2d * d);
  *imax = 0;
  if (nfft2d > 16384) return "nfft2d too large";
  if (*dk >= k0) return "In FFT Poynting nfft2d too small";
  *imax = (int)std::lround(k0 / *dk) + 1;
  if (2 * *imax + 1 > nfft2d || 2 * *imax + 1 < 7) return "In FFT diffract nfft2d too small";
  return "";
}

// The same for the lens pupil, numaper in k units (diffractefft2dlens.f:24-72).
std::string lens_geometry(int nfft2d, double k0, double d, double numaper, int* imax, double* dk) {
  *dk = 2.0 * 3.141592653589793238462643383279502884 / ((double)nfft2d * d);
  *imax = 0;
  if (nfft2d > 16384) return "nfft2d too large";
  if (*dk >= numaper) return "In FFT lens nfft2d too small";
  *imax = (int)std::lround(k0 / *dk) + 1;
  if (2 * *imax + 1 > nfft2d) return "In FFT diffraction for lens nfft2d too small!";
  if (2 * *imax + 1 < 7) return "In FFT diffract nfft2d too small";
  return "";
}

int farfield_dev(ifdda_ctx* c, int R, const double2* FF, const int32_t* tab, long long nd,
                 int nfft2d, double k0, double d, double x0, double y0, double z0, double2* Epos,
                 double2* Eneg, int* imax_out, double* dk_out, hipStream_t s, int mode,
                 double numaper, int nside) {
  const GridPlan& g = c->g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (nx > nfft2d || ny > nfft2d) {
    set_error("ifdda_farfield: the box is wider than nfft2d");
    return IFDDA_ERR_ARG;
  }
  int imax;
  double dk;
  const std::string msg = mode == 0 ? farfield_geometry(nfft2d, k0, d, &imax, &dk)
                                    : lens_geometry(nfft2d, k0, d, numaper, &imax, &dk);
  const double dside = mode == 0 ? 1.0 : (double)nside;
  *imax_out = imax;
  *dk_out = dk;
  if (!msg.empty()) {
    set_error("ifdda_farfield: " + msg);
    return IFDDA_ERR_ARG;
  }
  const int n = 2 * imax + 1, h = nfft2d / 2;
  // Phase tables in FP64 on the host.  The reference puts cell a (1-based
  // a = 1..nx) at FFT position tabfft2(a) and reads frequency i at position
  // tabfft2(i + nfft2d/2 + 1) (diffractefft2d.f:49-55, 104, 146), so a bin is
  // sum_a F(a) exp(-2 pi i p(a) q(i) / N) with p, q those 0-based positions;
  // the origin phase exp(-i (var1 i + var2 j)) follows (diffractefft2d.f:27-28).
  const double tpi = 6.283185307179586476925286766559;
  std::vector<double2> px((size_t)n * nx), py((size_t)n * ny), pz((size_t)n * n * nz), pv((size_t)n * n);
  auto tabf = [&](int i1) { return i1 - h - 1 >= 0 ? i1 - h : h + i1; };
  auto bin_phase = [&](int i, int a) {
    const long long pa = tabf(a + 1) - 1, qi = tabf(i + h + 1) - 1;
    const double ang = -tpi * (double)((pa * qi) % nfft2d) / (double)nfft2d;
    return make_double2(std::cos(ang), std::sin(ang));
  };
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < nx; ++a) px[(size_t)i * nx + a] = bin_phase(i - imax, a);
    for (int b = 0; b < ny; ++b) py[(size_t)i * ny + b] = bin_phase(i - imax, b);
  }
  const double var1 = (x0 + h * d) * dk, var2 = (y0 + h * d) * dk;
  for (int jj = 0; jj < n; ++jj)
    for (int ii = 0; ii < n; ++ii) {
      const double kx = (ii - imax) * dk, ky = (jj - imax) * dk;
      const double kz2 = k0 * k0 - kx * kx - ky * ky;
      const double kz = kz2 > 0.0 ? std::sqrt(kz2) : 0.0;
      const size_t ij = (size_t)ii + (size_t)n * jj;
      for (int k = 0; k < nz; ++k) {
        const double ang = -dside * kz * (z0 + k * d);
        pz[ij * nz + k] = make_double2(std::cos(ang), std::sin(ang));
      }
      const double av = -(var1 * (ii - imax) + var2 * (jj - imax));
      pv[ij] = make_double2(std::cos(av), std::sin(av));
    }
  // workspace: tables, then X[r][i][k][b][c]
  const size_t nt = px.size() + py.size() + pz.size() + pv.size();
  const size_t nxr = (size_t)R * n * nz * ny * 3;
  if (!dev_ensure(c, sizeof(double2) * (nt + nxr))) return IFDDA_ERR_HIP;
  double2* dpx = (double2*)c->ffbuf;
  double2* dpy = dpx + px.size();
  double2* dpz = dpy + py.size();
  double2* dpv = dpz + pz.size();
  double2* X = dpv + pv.size();
  if (!hip_ok(hipMemcpyAsync(dpx, px.data(), sizeof(double2) * px.size(), hipMemcpyHostToDevice, s), "copy px") ||
      !hip_ok(hipMemcpyAsync(dpy, py.data(), sizeof(double2) * py.size(), hipMemcpyHostToDevice, s), "copy py") ||
      !hip_ok(hipMemcpyAsync(dpz, pz.data(), sizeof(double2) * pz.size(), hipMemcpyHostToDevice, s), "copy pz") ||
      !hip_ok(hipMemcpyAsync(dpv, pv.data(), sizeof(double2) * pv.size(), hipMemcpyHostToDevice, s), "copy pv"))
    return IFDDA_ERR_HIP;
  const size_t lds = sizeof(double2) * nx * 3;
  if (lds > 65536) {
    set_error("ifdda_farfield: box too wide for the row stage (nx > 1365)");
    return IFDDA_ERR_ARG;
  }
  hipLaunchKernelGGL(k_ff_rows, dim3(nz * ny, R), dim3(kFfThreads), lds, s, FF, nd, tab, nx, ny, nz,
                     (const double2*)dpx, n, X);
  if (mode == 0)
    hipLaunchKernelGGL(k_ff_cols<0>, dim3(n * n, R), dim3(64), 0, s, (const double2*)X, ny, nz, n,
                       (const double2*)dpy, (const double2*)dpz, (const double2*)dpv, dk, k0, 0.0,
                       1.0, Epos, Eneg);
  else
    hipLaunchKernelGGL(k_ff_cols<1>, dim3(n * n, R), dim3(64), 0, s, (const double2*)X, ny, nz, n,
                       (const double2*)dpy, (const double2*)dpz, (const double2*)dpv, dk, k0,
                       numaper * numaper, dside, Epos, Eneg);
  if (!hip_ok(hipGetLastError(), "far-field launch") || !hip_ok(hipStreamSynchronize(s), "far field"))
    return IFDDA_ERR_HIP;
  return IFDDA_OK;
}

int gc_dev(ifdda_ctx* c, int R, const double2* Epos, const double2* Eneg, int imax, double dk,
           double k0, const double* I0, const double* theta, const double* phi, double* out,
           hipStream_t s) {
  const int n = 2 * imax + 1;
  const double rad = 3.141592653589793238462643383279502884 / 180.0;
  std::vector<double> h((size_t)R * 4);
  for (int r = 0; r < R; ++r) {
    h[R + 3 * r] = std::sin(theta[r] * rad) * std::cos(phi[r] * rad);
    h[R + 3 * r + 1] = std::sin(theta[r] * rad) * std::sin(phi[r] * rad);
    h[R + 3 * r + 2] = std::cos(theta[r] * rad);
    h[r] = I0[r];
  }
  if (!dev_ensure(c, sizeof(double) * (h.size() + 2 * (size_t)R))) return IFDDA_ERR_HIP;
  double* d = (double*)c->ffbuf;
  if (!hip_ok(hipMemcpyAsync(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s), "copy gc"))
    return IFDDA_ERR_HIP;
  hipLaunchKernelGGL(k_gc, dim3(R), dim3(kFfThreads), 0, s, Epos, Eneg, n, dk, k0, (const double*)d,
                     (const double*)(d + R), d + 4 * (size_t)R);
  if (!hip_ok(hipMemcpyAsync(out, d + 4 * (size_t)R, sizeof(double) * 2 * R, hipMemcpyDeviceToHost, s),
              "copy gc out") ||
      !hip_ok(hipStreamSynchronize(s), "computegcfft2d"))
    return IFDDA_ERR_HIP;
  return IFDDA_OK;
}

}  // namespace ifdda

using namespace ifdda;

static hipStream_t pick(ifdda_ctx* c, void* stream) {
  return stream ? (hipStream_t)stream : c->stream;
}

extern "C" int ifdda_cross_sections_device(ifdda_ctx* c, int nrhs, const void* FF0, const void* FF,
                                           const void* FFloc, double k0, const double* I0,
                                           double* out, void* stream) {
  if (!c || nrhs < 1 || !FF0 || !FF || !FFloc || !I0 || !out || !(k0 > 0.0)) {
    set_error("ifdda_cross_sections_device: bad argument");
    return IFDDA_ERR_ARG;
  }
  if (!c->have_dipoles) {
    set_error("ifdda_cross_sections_device: set the dipoles first");
    return IFDDA_ERR_STATE;
  }
  return cross_sections_dev(c, nrhs, (const double2*)FF0, (const double2*)FF, (const double2*)FFloc,
                            k0, I0, out, pick(c, stream));
}

extern "C" int ifdda_farfield_device(ifdda_ctx* c, int nrhs, const void* FF, int nfft2d, double k0,
                                     double d, double x0, double y0, double z0, void* Epos,
                                     void* Eneg, int* imax, double* deltak, void* stream) {
  if (!c || nrhs < 1 || !FF || !Epos || !Eneg || !imax || !deltak || nfft2d < 1 || !(k0 > 0.0) ||
      !(d > 0.0)) {
    set_error("ifdda_farfield_device: bad argument");
    return IFDDA_ERR_ARG;
  }
  if (!c->have_dipoles) {
    set_error("ifdda_farfield_device: set the dipoles first");
    return IFDDA_ERR_STATE;
  }
  return farfield_dev(c, nrhs, (const double2*)FF, c->tab, c->nd, nfft2d, k0, d, x0, y0, z0,
                      (double2*)Epos, (double2*)Eneg, imax, deltak, pick(c, stream), 0, 0.0, 1);
}

extern "C" int ifdda_farfield_size(int nfft2d, double k0, double d, int* imax, double* deltak) {
  if (!imax || !deltak || nfft2d < 1 || !(k0 > 0.0) || !(d > 0.0)) {
    set_error("ifdda_farfield_size: bad argument");
    return IFDDA_ERR_ARG;
  }
  const std::string msg = farfield_geometry(nfft2d, k0, d, imax, deltak);
  if (!msg.empty()) {
    set_error("ifdda_farfield_size: " + msg);
    return IFDDA_ERR_ARG;
  }
  return IFDDA_OK;
}

extern "C" int ifdda_computegcfft2d_device(ifdda_ctx* c, int nrhs, const void* Epos,
                                           const void* Eneg, int imax, double deltak, double k0,
                                           const double* I0, const double* theta,
                                           const double* phi, double* out, void* stream) {
  if (!c || nrhs < 1 || !Epos || !Eneg || imax < 1 || !(deltak > 0.0) || !(k0 > 0.0) || !I0 ||
      !theta || !phi || !out) {
    set_error("ifdda_computegcfft2d_device: bad argument");
    return IFDDA_ERR_ARG;
  }
  return gc_dev(c, nrhs, (const double2*)Epos, (const double2*)Eneg, imax, deltak, k0, I0, theta,
                phi, out, pick(c, stream));
}
