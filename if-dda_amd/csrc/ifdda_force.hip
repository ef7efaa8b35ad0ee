// ifdda_force.hip -- optical force and torque on every dipole
// (cdmlib.f:3102-3560, nforce / ntorque) from a solved field:
//   F_d(i) = 1/2 Re sum_c conj(p_c) dE_c/dx_d (i) * 4 pi eps0,
//   dE/dx_d = i k_d E_inc + sum_j dT/dx_d(r_i - r_j) p_j       (derivativefield2.f)
//   tau(i)  = 1/2 Re conj(p) x (E_loc + 2/3 i k0^3 p) * 4 pi eps0 + (r_i - r_g) x F(i)
// with p = alpha E (FF), E_loc = E (FFloc), r_g the mean dipole position.
// The scattered derivative is one FFT convolution per direction through the
// matvec passes: the folded FFT of d T / d x_d (odd along x_d, so the z pass
// runs its k_zconv_der variant), no polarisability, a zero right-hand side:
// y = 0 - (1/M) conv, i.e. minus the field derivative.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/ifdda.h"
#include "ifdda_internal.h"

namespace ifdda {
namespace {

constexpr int kFThreads = 256;
constexpr int kFMaxBlocks = 1024;
constexpr int kFSums = 12;  // F (3), tau_loc (3), cell x F (3), cell (3)
constexpr double kC = 299792458.0;                    // cdmlib.f:478
constexpr double kQuatPiEps0 = 1.0 / (kC * kC * 1e-7);  // cdmlib.f:480

__device__ __forceinline__ double2 cm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double re_conj_mul(double2 a, double2 b) {  // Re(conj(a) b)
  return a.x * b.x + a.y * b.y;
}

struct ForceArgs {
  const double2* Einc;
  const double2* FF;
  const double2* Eloc;
  const double2* y[3];   // -(d E_scat / d x_d) at the dipoles
  const double* kvec;    // 3 per RHS
  const int32_t* dipcell;  // box cell of each dipole, or null (dense box)
  long long nd;
  int nx, ny;
  double k0, q;
  double* fdip;  // 3 per dipole per RHS, or null
  double* part;  // [RHS][kFSums][nblk]
  int nblk;
};

// Per-dipole force and local torque; block partial sums of F, tau_loc,
// cell x F and cell in a fixed order (cdmlib.f:3341-3370, 3529-3545).
__global__ void __launch_bounds__(kFThreads) k_force(ForceArgs a) {
  const int r = blockIdx.y;
  const long long off = (long long)r * 3 * a.nd;
  const double kx = a.kvec[3 * r], ky = a.kvec[3 * r + 1], kz = a.kvec[3 * r + 2];
  const double kd[3] = {kx, ky, kz};
  const double k3 = a.k0 * a.k0 * a.k0;
  double acc[kFSums];
  for (int q = 0; q < kFSums; ++q) acc[q] = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < a.nd;
       i += (long long)gridDim.x * blockDim.x) {
    const long long e = off + 3 * i;
    double2 p[3], E[3], L[3];
    for (int c = 0; c < 3; ++c) p[c] = a.FF[e + c], E[c] = a.Einc[e + c], L[c] = a.Eloc[e + c];
    double F[3];
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) {
        // i k_d E_inc - y_d: the incident derivative (ondeplaned.f:43-58)
        // plus the scattered one
        const double2 yi = a.y[d][e + c];
        const double2 der = make_double2(-kd[d] * E[c].y - yi.x, kd[d] * E[c].x - yi.y);
        s += re_conj_mul(p[c], der);
      }
      F[d] = 0.5 * s * a.q;
    }
    double2 Em[3];
    for (int c = 0; c < 3; ++c) {
      const double2 ip = make_double2(-p[c].y * (2.0 / 3.0) * k3, p[c].x * (2.0 / 3.0) * k3);
      Em[c] = make_double2(ip.x + L[c].x, ip.y + L[c].y);
    }
    const double tl[3] = {0.5 * (re_conj_mul(p[1], Em[2]) - re_conj_mul(p[2], Em[1])) * a.q,
                          0.5 * (-re_conj_mul(p[0], Em[2]) + re_conj_mul(p[2], Em[0])) * a.q,
                          0.5 * (re_conj_mul(p[0], Em[1]) - re_conj_mul(p[1], Em[0])) * a.q};
    const long long cell = a.dipcell ? (long long)a.dipcell[i] : i;
    const double cx = (double)(cell % a.nx), cy = (double)((cell / a.nx) % a.ny),
                 cz = (double)(cell / ((long long)a.nx * a.ny));
    if (a.fdip)
      for (int d = 0; d < 3; ++d) a.fdip[off + 3 * i + d] = F[d];
    for (int d = 0; d < 3; ++d) acc[d] += F[d], acc[3 + d] += tl[d];
    acc[6] += cy * F[2] - cz * F[1];
    acc[7] += cz * F[0] - cx * F[2];
    acc[8] += cx * F[1] - cy * F[0];
    acc[9] += cx;
    acc[10] += cy;
    acc[11] += cz;
  }
  __shared__ double sh[kFSums][kFThreads];
  for (int q = 0; q < kFSums; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = kFThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < kFSums; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kFSums)
    a.part[((long long)r * kFSums + threadIdx.x) * a.nblk + blockIdx.x] = sh[threadIdx.x][0];
}

// Sum the block partials of every RHS in block order.
__global__ void k_force_fin(const double* part, int nblk, double* out) {
  const int r = blockIdx.x, q = threadIdx.x;
  if (q >= kFSums) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[((long long)r * kFSums + q) * nblk + b];
  out[r * kFSums + q] = s;
}

// Per-dipole torque: tau_loc + (r_i - r_g) x F_i with r_g the mean cell
// (cdmlib.f:3529-3545); recomputes tau_loc from p and E_loc.
__global__ void __launch_bounds__(kFThreads) k_torque_dip(ForceArgs a, const double* mean,
                                                          double d, double* tdip) {
  const int r = blockIdx.y;
  const long long off = (long long)r * 3 * a.nd;
  const double k3 = a.k0 * a.k0 * a.k0;
  const double gx = mean[3 * r], gy = mean[3 * r + 1], gz = mean[3 * r + 2];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < a.nd;
       i += (long long)gridDim.x * blockDim.x) {
    const long long e = off + 3 * i;
    double2 p[3], Em[3];
    for (int c = 0; c < 3; ++c) {
      p[c] = a.FF[e + c];
      const double2 L = a.Eloc[e + c];
      Em[c] = make_double2(-p[c].y * (2.0 / 3.0) * k3 + L.x, p[c].x * (2.0 / 3.0) * k3 + L.y);
    }
    const long long cell = a.dipcell ? (long long)a.dipcell[i] : i;
    const double rx = ((double)(cell % a.nx) - gx) * d, ry = ((double)((cell / a.nx) % a.ny) - gy) * d,
                 rz = ((double)(cell / ((long long)a.nx * a.ny)) - gz) * d;
    const double* F = a.fdip + e;
    tdip[e] = 0.5 * (re_conj_mul(p[1], Em[2]) - re_conj_mul(p[2], Em[1])) * a.q + (ry * F[2] - rz * F[1]);
    tdip[e + 1] = 0.5 * (-re_conj_mul(p[0], Em[2]) + re_conj_mul(p[2], Em[0])) * a.q + (rz * F[0] - rx * F[2]);
    tdip[e + 2] = 0.5 * (re_conj_mul(p[0], Em[1]) - re_conj_mul(p[1], Em[0])) * a.q + (rx * F[1] - ry * F[0]);
  }
}

// The folded derivative tensors for (d, k0), built on first use.
bool ensure_dgreen(ifdda_ctx* c, double d, double k0, hipStream_t s) {
  if (c->gd_ready && c->gd_d == d && c->gd_k0 == k0) return true;
  const size_t nf = (size_t)c->g.F[0] * c->g.F[1] * c->g.F[2] * 6;
  if (!c->Gd64 && !hip_ok(hipMalloc(&c->Gd64, sizeof(double2) * nf * 3), "hipMalloc force Green"))
    return false;
  const size_t M = (size_t)c->g.L[0] * c->g.L[1] * c->g.L[2];
  double2* full = green_full(c, M);
  if (!full) return false;
  const double2* tw[3] = {c->tw64[0], c->tw64[1], c->tw64[2]};
  const int* pos[3] = {c->pos[0], c->pos[1], c->pos[2]};
  bool ok = true;
  for (int dd = 0; ok && dd < 3; ++dd)
    ok = green_dbuild_device(c->g, tw, pos, d, k0, dd, full, c->Gd64 + nf * dd, s);
  ok = ok && hip_ok(hipStreamSynchronize(s), "force Green build");
  if (!ok) return false;
  c->gd_ready = true;
  c->gd_d = d;
  c->gd_k0 = k0;
  return true;
}

}  // namespace

int derivative_field_box(ifdda_ctx* c, const double* mom, double d, double k0, int dd, double* out) {
  if (!(d > 0.0) || !(k0 > 0.0) || dd < 0 || dd > 2) {
    set_error("derivativefield2: need d > 0, k0 > 0 and test in 1..3");
    return IFDDA_ERR_ARG;
  }
  hipStream_t s = c->stream;
  if (!ensure_dgreen(c, d, k0, s)) return IFDDA_ERR_HIP;
  if (ensure_scratch(c, 1) != IFDDA_OK) return IFDDA_ERR_HIP;
  const long long n = 3LL * c->g.nx * c->g.ny * c->g.nz;
  const size_t nf = (size_t)c->g.F[0] * c->g.F[1] * c->g.F[2] * 6;
  double2* buf = nullptr;  // moments, zero, result
  bool ok = hip_ok(hipMalloc(&buf, sizeof(double2) * n * 3), "hipMalloc derivative field") &&
            hip_ok(hipMemcpyAsync(buf, mom, sizeof(double2) * n, hipMemcpyHostToDevice, s),
                   "copy moments") &&
            hip_ok(hipMemsetAsync(buf + n, 0, sizeof(double2) * n, s), "zero rhs");
  if (ok) {
    MatvecDev<double2> m = c->mv64();
    m.tab = nullptr;  // every box cell: the moments come in box order
    m.G = c->Gd64 + nf * dd;
    m.alpha_mode = kAlphaNone;
    m.dpar = dd + 1;
    m.nrhs = 1;
    m.vstride = n;
    m.timer = nullptr;
    ok = launch_matvec_f64(c->g, m, buf, buf + n, buf + 2 * n, s, nullptr) &&
         hip_ok(hipMemcpyAsync(out, buf + 2 * n, sizeof(double2) * n, hipMemcpyDeviceToHost, s),
                "copy derivative field") &&
         hip_ok(hipStreamSynchronize(s), "derivative field");
  }
  (void)hipFree(buf);
  if (!ok) return IFDDA_ERR_HIP;
  for (long long i = 0; i < 2 * n; ++i) out[i] = -out[i];  // y = 0 - conv / M
  return IFDDA_OK;
}

}  // namespace ifdda

using namespace ifdda;

extern "C" int ifdda_force_device(ifdda_ctx* c, int R, const void* E_inc, const void* FF,
                                  const void* FFloc, const double* kvec, double d, double k0,
                                  double* force, double* torque, void* fdip, void* tdip,
                                  void* stream) {
  if (!c || !E_inc || !FF || !FFloc || !kvec || !force || R < 1) {
    set_error("ifdda_force_device: NULL argument or nrhs < 1");
    return IFDDA_ERR_ARG;
  }
  if (!(d > 0.0) || !(k0 > 0.0)) {
    set_error("ifdda_force_device: need d > 0 and k0 > 0");
    return IFDDA_ERR_ARG;
  }
  if (tdip && !fdip) {
    set_error("ifdda_force_device: per-dipole torques need the per-dipole forces (fdip)");
    return IFDDA_ERR_ARG;
  }
  if (!c->have_dipoles || c->nd < 1) {
    set_error("ifdda_force_device: no dipoles set");
    return IFDDA_ERR_STATE;
  }
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  if (!ensure_dgreen(c, d, k0, s)) return IFDDA_ERR_HIP;
  if (ensure_scratch(c, R) != IFDDA_OK) return IFDDA_ERR_HIP;
  const long long n = 3 * c->nd, tot = n * R;
  const size_t nf = (size_t)c->g.F[0] * c->g.F[1] * c->g.F[2] * 6;
  const int nblk = (int)std::max<long long>(1, std::min<long long>(kFMaxBlocks, (c->nd + kFThreads - 1) / kFThreads));
  double2* buf = nullptr;  // zero, y_x, y_y, y_z
  double *kv = nullptr, *part = nullptr, *sums = nullptr, *mean = nullptr;
  bool ok = hip_ok(hipMalloc(&buf, sizeof(double2) * tot * 4), "hipMalloc force") &&
            hip_ok(hipMalloc(&kv, sizeof(double) * 3 * R), "hipMalloc force") &&
            hip_ok(hipMalloc(&part, sizeof(double) * kFSums * nblk * R), "hipMalloc force") &&
            hip_ok(hipMalloc(&sums, sizeof(double) * kFSums * R), "hipMalloc force") &&
            hip_ok(hipMalloc(&mean, sizeof(double) * 3 * R), "hipMalloc force") &&
            hip_ok(hipMemsetAsync(buf, 0, sizeof(double2) * tot, s), "zero rhs") &&
            hip_ok(hipMemcpyAsync(kv, kvec, sizeof(double) * 3 * R, hipMemcpyHostToDevice, s),
                   "copy k");
  for (int dd = 0; ok && dd < 3; ++dd) {
    MatvecDev<double2> m = c->mv64();
    m.G = c->Gd64 + nf * dd;
    m.alpha_mode = kAlphaNone;
    m.dpar = dd + 1;
    m.nrhs = R;
    m.vstride = n;
    m.timer = nullptr;
    ok = launch_matvec_f64(c->g, m, (const double2*)FF, buf, buf + tot * (1 + dd), s, nullptr);
  }
  std::vector<double> h(kFSums * R);
  ForceArgs a{};
  if (ok) {
    a.Einc = (const double2*)E_inc;
    a.FF = (const double2*)FF;
    a.Eloc = (const double2*)FFloc;
    for (int dd = 0; dd < 3; ++dd) a.y[dd] = buf + tot * (1 + dd);
    a.kvec = kv;
    a.dipcell = c->dipcell;
    a.nd = c->nd;
    a.nx = c->g.nx;
    a.ny = c->g.ny;
    a.k0 = k0;
    a.q = kQuatPiEps0;
    a.fdip = (double*)fdip;
    a.part = part;
    a.nblk = nblk;
    hipLaunchKernelGGL(k_force, dim3(nblk, R), dim3(kFThreads), 0, s, a);
    hipLaunchKernelGGL(k_force_fin, dim3(R), dim3(32), 0, s, (const double*)part, nblk, sums);
    ok = hip_ok(hipGetLastError(), "force launch") &&
         hip_ok(hipMemcpyAsync(h.data(), sums, sizeof(double) * kFSums * R, hipMemcpyDeviceToHost, s),
                "copy force") &&
         hip_ok(hipStreamSynchronize(s), "force");
  }
  std::vector<double> hm(3 * R);
  if (ok) {
    const double ndd = (double)c->nd;
    for (int r = 0; r < R; ++r) {
      const double* q = h.data() + kFSums * r;
      const double g[3] = {q[9] / ndd, q[10] / ndd, q[11] / ndd};  // mean cell
      for (int k = 0; k < 3; ++k) {
        force[3 * r + k] = q[k];
        hm[3 * r + k] = g[k];
      }
      if (torque) {
        // sum_i (r_i - r_g) x F_i = d (sum cell x F - g x F_tot)
        const double F[3] = {q[0], q[1], q[2]};
        const double gxF[3] = {g[1] * F[2] - g[2] * F[1], g[2] * F[0] - g[0] * F[2],
                               g[0] * F[1] - g[1] * F[0]};
        for (int k = 0; k < 3; ++k) torque[3 * r + k] = q[3 + k] + d * (q[6 + k] - gxF[k]);
      }
    }
  }
  if (ok && tdip) {
    ok = hip_ok(hipMemcpyAsync(mean, hm.data(), sizeof(double) * 3 * R, hipMemcpyHostToDevice, s),
                "copy mean") ;
    if (ok) {
      hipLaunchKernelGGL(k_torque_dip, dim3(nblk, R), dim3(kFThreads), 0, s, a,
                         (const double*)mean, d, (double*)tdip);
      ok = hip_ok(hipGetLastError(), "torque launch") && hip_ok(hipStreamSynchronize(s), "torque");
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(buf);
  (void)hipFree(kv);
  (void)hipFree(part);
  (void)hipFree(sums);
  (void)hipFree(mean);
  return ok ? IFDDA_OK : IFDDA_ERR_HIP;
}
