// ifdda_beams.hip -- incident plane waves and polarisabilities on the device.
//
// The step before the solve: the right-hand sides of a microscope's
// incidences or of a request are plane waves at the dipoles, and the
// polarisability follows from the permittivity per cell.  Both are formed
// here from a few parameters per wave / per method instead of host loops over
// the dipoles (ondeplane.f, ondecirce.f, ondeplanekxky.f, polarisabilite.f).
#include <cmath>
#include <string>
#include <vector>

#include "../../include/ifdda.h"
#include "ifdda_internal.h"

namespace ifdda {

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

struct Z {
  double x, y;
};
__host__ __device__ inline Z zm(Z a, Z b) { return Z{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline Z zs(Z a, Z b) { return Z{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline Z za(Z a, Z b) { return Z{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline Z zr(double r, Z a) { return Z{r * a.x, r * a.y}; }
// Smith's division
__host__ __device__ inline Z zd(Z a, Z b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return Z{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  }
  const double r = b.x / b.y, d = b.y + b.x * r;
  return Z{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}

// Polarisability of one cell of side d: poladiff with dddis = 1
// (polarisabilite.f:20-59): radius of the sphere of the cell's volume,
// Clausius-Mossotti, then the method's correction.
__host__ __device__ inline Z pola_value(Z eps, double d, double k0, double eps0, int method) {
  const Z one{1.0, 0.0}, ic{0.0, 1.0};
  const double indice = sqrt(eps0), k03 = k0 * k0 * k0 * indice * indice * indice;
  const double maille = d, rayon = d * pow(0.75 / kPi, 1.0 / 3.0), volume = rayon * rayon * rayon;
  const Z rapport = zd(Z{eps.x - eps0, eps.y}, Z{eps.x + 2.0 * eps0, eps.y});
  const Z pola = zr(volume, rapport);
  const Z rr = zm(zr(2.0 / 3.0 * k03, ic), pola);  // 2/3 i k03 pola
  switch (method) {
    case IFDDA_POLA_CM:
      return zr(eps0, pola);
    case IFDDA_POLA_RR:
      return zd(zr(eps0, pola), zs(one, rr));
    case IFDDA_POLA_GB:
      return zd(pola, zs(zs(one, rr), zr(k0 * k0 * indice * indice / rayon, pola)));
    case IFDDA_POLA_LA: {
      const double a = k0 * indice * rayon;
      const Z e{cos(a), sin(a)};
      const Z f = zs(zm(Z{1.0, -a}, e), one);  // (1 - i k0 n a) exp(i k0 n a) - 1
      return zd(pola, zs(one, zr(2.0, zm(rapport, f))));
    }
    default: {  // IFDDA_POLA_LR
      const double sldr = 1.0 / 5.0;
      const Z er{eps.x / eps0, eps.y / eps0};
      const Z c = za(Z{-1.8915316, 0.0}, zs(zr(0.1648469, er), zr(1.7700004 * sldr, er)));
      const Z t = zs(zr(k0 * k0 * indice * indice / maille, c),
                     zr(2.0 / 3.0 * k03 * indice * indice * indice, ic));
      return zd(zr(eps0, pola), za(one, zm(pola, t)));
    }
  }
}

__global__ void __launch_bounds__(256) k_pola(const double2* eps, long long nd, double d, double k0,
                                              double eps0, int method, double2* al64,
                                              float2* al32) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nd;
       i += (long long)gridDim.x * blockDim.x) {
    const Z a = pola_value(Z{eps[i].x, eps[i].y}, d, k0, eps0, method);
    al64[i] = make_double2(a.x, a.y);
    al32[i] = make_float2((float)a.x, (float)a.y);
  }
}

__device__ __forceinline__ double2 cm2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// E_r[3 p + c] = e_r[c] exp(i k_r . x_p) at dipole p's cell centre
// x = origin + d (i, j, k); RHS r starts at off[r] (or r * 3 nd).
__global__ void __launch_bounds__(256) k_waves(double2* E, const long long* off, long long nd,
                                               const int32_t* dipcell, int nx, int ny, double x0,
                                               double y0, double z0, double d, const double* kv,
                                               const double2* ev) {
  const int r = blockIdx.y;
  const double kx = kv[3 * r], ky = kv[3 * r + 1], kz = kv[3 * r + 2];
  const double2 e0 = ev[3 * r], e1 = ev[3 * r + 1], e2 = ev[3 * r + 2];
  double2* Er = E + (off ? off[r] : (long long)r * 3 * nd);
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < nd;
       p += (long long)gridDim.x * blockDim.x) {
    const long long cell = dipcell ? (long long)dipcell[p] : p;
    double2 ph = make_double2(0.0, 0.0);
    if (cell >= 0) {
      const long long i = cell % nx, j = (cell / nx) % ny, k = cell / ((long long)nx * ny);
      const double arg = kx * (x0 + i * d) + ky * (y0 + j * d) + kz * (z0 + k * d);
      double sn, cs;
      sincos(arg, &sn, &cs);
      ph = make_double2(cs, sn);
    }
    Er[3 * p] = cm2(e0, ph);
    Er[3 * p + 1] = cm2(e1, ph);
    Er[3 * p + 2] = cm2(e2, ph);
  }
}

}  // namespace

double2 poladiff_host(double2 eps, double d, double k0, double eps0, int method) {
  const Z a = pola_value(Z{eps.x, eps.y}, d, k0, eps0, method);
  return make_double2(a.x, a.y);
}

void plane_wave_vector(double k0, double2 E0, double pp, double kx, double ky, double2 e[3],
                       double* kz_out) {
  // ondeplanekxky.f: s along k x z, p completing (s, p, k); normal incidence
  // takes p along x, s along y
  const double p = std::sqrt(pp), s = std::sqrt(1.0 - pp);
  const double2 E0s = make_double2(E0.x * s, E0.y * s), E0p = make_double2(E0.x * p, E0.y * p);
  const double kz = std::sqrt(k0 * k0 - kx * kx - ky * ky);
  *kz_out = kz;
  if (kz >= 0.99999 * k0) {
    e[0] = E0p;
    e[1] = E0s;
    e[2] = make_double2(0.0, 0.0);
    return;
  }
  const double r = std::sqrt(kx * kx + ky * ky);
  const double nsx = -ky / r, nsy = kx / r;
  double npx = kz * nsy, npy = -kz * nsx, npz = -kx * nsy + ky * nsx;
  const double nn = std::sqrt(npx * npx + npy * npy + npz * npz);
  npx /= nn, npy /= nn, npz /= nn;
  e[0] = make_double2(E0s.x * nsx + E0p.x * npx, E0s.y * nsx + E0p.y * npx);
  e[1] = make_double2(E0s.x * nsy + E0p.x * npy, E0s.y * nsy + E0p.y * npy);
  e[2] = make_double2(E0p.x * npz, E0p.y * npz);
}

std::string wave_vectors(int kind, const double* prm, double2 E0, double k0, double k[3],
                         double2 e[3]) {
  if (kind == IFDDA_WAVE_KXKY) {
    const double pp = prm[2];
    if (!(pp >= 0.0 && pp <= 1.0)) return "problem in plane incident wave (pp outside [0, 1])";
    if (!(prm[0] * prm[0] + prm[1] * prm[1] <= k0 * k0)) return "kx^2 + ky^2 above k0^2";
    k[0] = prm[0], k[1] = prm[1];
    plane_wave_vector(k0, E0, pp, prm[0], prm[1], e, &k[2]);
    return "";
  }
  const double th = prm[0] * kPi / 180.0, ph = prm[1] * kPi / 180.0;
  k[0] = k0 * std::sin(th) * std::cos(ph);
  k[1] = k0 * std::sin(th) * std::sin(ph);
  k[2] = k0 * std::cos(th);
  const Z e0{E0.x, E0.y};
  Z s, p;
  if (kind == IFDDA_WAVE_PLANE) {  // ondeplane.f:26-45
    const double pp = prm[2];
    if (!(pp >= 0.0 && pp <= 1.0)) return "problem in plane incident wave (pp outside [0, 1])";
    const double sg = std::signbit(th) ? -1.0 : 1.0;
    s = zr(std::sqrt(1.0 - pp) * sg, e0);
    p = zr(std::sqrt(pp) * sg, e0);
  } else if (kind == IFDDA_WAVE_CIRCULAR) {  // ondecirce.f:23-51
    const double us = 1.0 / std::sqrt(2.0);
    s = zr(us, e0);
    p = zm(Z{0.0, std::signbit(prm[2]) ? -us : us}, e0);
  } else {
    return "unknown wave kind";
  }
  const Z ex = za(zr(-std::sin(ph), s), zr(std::cos(th) * std::cos(ph), p));
  const Z ey = za(zr(std::cos(ph), s), zr(std::cos(th) * std::sin(ph), p));
  const Z ez = zr(-std::sin(th), p);
  e[0] = make_double2(ex.x, ex.y);
  e[1] = make_double2(ey.x, ey.y);
  e[2] = make_double2(ez.x, ez.y);
  return "";
}

bool launch_waves(double2* E, const long long* off, int R, long long nd, const int32_t* dipcell,
                  int nx, int ny, const double origin[3], double d, const double* kv,
                  const double2* ev, hipStream_t s) {
  const unsigned gx = (unsigned)std::min<long long>(1024, (nd + 255) / 256);
  hipLaunchKernelGGL(k_waves, dim3(gx, R), dim3(256), 0, s, E, off, nd, dipcell, nx, ny,
                     origin[0], origin[1], origin[2], d, kv, ev);
  return hip_ok(hipGetLastError(), "incident waves");
}

bool launch_pola(const double2* eps, long long nd, double d, double k0, double eps0, int method,
                 double2* al64, float2* al32, hipStream_t s) {
  const unsigned gx = (unsigned)std::min<long long>(1024, (nd + 255) / 256);
  hipLaunchKernelGGL(k_pola, dim3(gx), dim3(256), 0, s, eps, nd, d, k0, eps0, method, al64, al32);
  return hip_ok(hipGetLastError(), "polarisability");
}

}  // namespace ifdda

using namespace ifdda;

extern "C" int ifdda_incident_device(ifdda_ctx* c, int kind, int R, const double* params,
                                     const double E0[2], double k0, double d,
                                     const double origin[3], void* E, void* stream) {
  if (!c || !params || !E0 || !origin || !E || R < 1 || !(k0 > 0.0) || !(d > 0.0)) {
    set_error("ifdda_incident_device: bad argument");
    return IFDDA_ERR_ARG;
  }
  if (!c->have_dipoles) {
    set_error("ifdda_incident_device: set the dipoles first");
    return IFDDA_ERR_STATE;
  }
  std::vector<double> kv(3 * (size_t)R);
  std::vector<double2> ev(3 * (size_t)R);
  for (int r = 0; r < R; ++r) {
    const std::string msg = wave_vectors(kind, params + 3 * r, make_double2(E0[0], E0[1]), k0,
                                         &kv[3 * r], &ev[3 * r]);
    if (!msg.empty()) {
      set_error("ifdda_incident_device: wave " + std::to_string(r) + ": " + msg);
      return IFDDA_ERR_ARG;
    }
  }
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  double* dk = nullptr;
  double2* de = nullptr;
  bool ok = hip_ok(hipMalloc(&dk, sizeof(double) * kv.size()), "hipMalloc waves") &&
            hip_ok(hipMalloc(&de, sizeof(double2) * ev.size()), "hipMalloc waves") &&
            hip_ok(hipMemcpyAsync(dk, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice, s), "copy waves") &&
            hip_ok(hipMemcpyAsync(de, ev.data(), sizeof(double2) * ev.size(), hipMemcpyHostToDevice, s), "copy waves") &&
            launch_waves((double2*)E, nullptr, R, c->nd, c->dipcell, c->g.nx, c->g.ny, origin, d, dk,
                         de, s) &&
            hip_ok(hipStreamSynchronize(s), "incident waves");
  (void)hipFree(dk);
  (void)hipFree(de);
  return ok ? IFDDA_OK : IFDDA_ERR_HIP;
}

extern "C" int ifdda_polarizability_device(ifdda_ctx* c, const double* eps, double d, double k0,
                                           double eps0, int method) {
  if (!c || !eps || !(d > 0.0) || !(k0 > 0.0) || !(eps0 > 0.0) || method < IFDDA_POLA_CM ||
      method > IFDDA_POLA_LR) {
    set_error("ifdda_polarizability_device: bad argument");
    return IFDDA_ERR_ARG;
  }
  if (!c->have_dipoles) {
    set_error("ifdda_polarizability_device: set the dipoles first");
    return IFDDA_ERR_STATE;
  }
  const long long nd = c->nd;
  double2* de = nullptr;
  if (!hip_ok(hipMalloc(&de, sizeof(double2) * nd), "hipMalloc eps")) return IFDDA_ERR_HIP;
  (void)hipFree(c->al64);  // a scalar per dipole, for the present dipoles
  (void)hipFree(c->al32);
  c->al64 = nullptr;
  c->al32 = nullptr;
  c->alpha_mode = kAlphaNone;
  if (!hip_ok(hipMalloc(&c->al64, sizeof(double2) * nd), "hipMalloc alpha") ||
      !hip_ok(hipMalloc(&c->al32, sizeof(float2) * nd), "hipMalloc alpha")) {
    (void)hipFree(de);
    return IFDDA_ERR_HIP;
  }
  const bool ok = hip_ok(hipMemcpyAsync(de, eps, sizeof(double2) * nd, hipMemcpyHostToDevice, c->stream), "copy eps") &&
                  launch_pola(de, nd, d, k0, eps0, method, c->al64, c->al32, c->stream) &&
                  hip_ok(hipStreamSynchronize(c->stream), "polarisability");
  (void)hipFree(de);
  if (!ok) return IFDDA_ERR_HIP;
  c->alpha_mode = kAlphaScalar;
  return IFDDA_OK;
}
