// ifdda_microscope.hip -- the brightfield microscope of microsbf.f on the
// device, many incidences per solve (quick lens path, nquicklens = 1; the
// direct-sum pupil of nquicklens = 0 with direct_lens):
//
//   condenser lattice of incidences        microsbf.f:90-172
//   incident plane waves at the dipoles    ondeplanekxky.f (k_onde)
//   local field, all incidences at once    ifdda_solve_multi_device
//   field in the lens pupil                diffractefft2dlens.f (far-field
//                                          kernels, lens mode), or the direct
//                                          sum of microsbf.f:629-668
//                                          (k_lens_direct)
//   pupil: focus phase, incident wave,     microsbf.f:676-790 (k_pupil)
//   lens rotation
//   image plane                            fouriertoimage.f (k_img_x, k_img_acc)
//   incoherent sum of |E|^2, square root   microsbf.f:806-845
//
// The pupil holds (2 imax + 1)^2 bins; for an even nfft2d the reference's
// forward FFT followed by its quadrant swap scaled by deltak^2 / gross is
// img(a, b) = deltak^2 / gross sum_{i,j} P(i, j) exp(-2 pi i ((a - h) i +
// (b - h) j) / N), h = N / 2, which k_img_x (over i) and k_img_acc (over j)
// evaluate directly; k_img_acc adds |img|^2 of every right-hand side in a
// fixed order per pixel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ifdda.h"
#include "ifdda_internal.h"

namespace ifdda {

int farfield_dev(ifdda_ctx* c, int R, const double2* FF, const int32_t* tab, long long nd,
                 int nfft2d, double k0, double d, double x0, double y0, double z0, double2* Epos,
                 double2* Eneg, int* imax_out, double* dk_out, hipStream_t s, int mode,
                 double numaper, int nside);
std::string lens_geometry(int nfft2d, double k0, double d, double numaper, int* imax, double* dk);

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

// P[r][0..2][bin] = M_bin Elens[r][.][bin]; P[r][3..5][bin] = the same plus
// M_bin einc_r at the incident bin of RHS r.  M_bin = focus phase x lens
// rotation x sqrt(u_z / v_z) (zero outside the objective aperture).
__global__ void __launch_bounds__(256) k_pupil(const double2* Elens, int nb, const double2* M,
                                               const int* incbin, const double2* einc,
                                               double2* P) {
  const int r = blockIdx.y;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double2* m = M + (long long)b * 9;
  const double2* e = Elens + (long long)r * 3 * nb;
  double2 s[3];
  for (int c = 0; c < 3; ++c)
    s[c] = cadd(cadd(cmul(m[3 * c], e[b]), cmul(m[3 * c + 1], e[nb + b])), cmul(m[3 * c + 2], e[2 * nb + b]));
  double2 w[3] = {s[0], s[1], s[2]};
  if (incbin[r] == b) {
    const double2* q = einc + 3 * r;
    for (int c = 0; c < 3; ++c)
      w[c] = cadd(w[c], cadd(cadd(cmul(m[3 * c], q[0]), cmul(m[3 * c + 1], q[1])), cmul(m[3 * c + 2], q[2])));
  }
  double2* o = P + (long long)r * 6 * nb;
  for (int c = 0; c < 3; ++c) {
    o[(long long)c * nb + b] = s[c];
    o[(long long)(3 + c) * nb + b] = w[c];
  }
}

// G[r][c][jj][a] = sum_ii P[r][c][jj n + ii] ex[a][ii]   (a < N)
__global__ void __launch_bounds__(256) k_img_x(const double2* P, int n, int N, const double2* ex,
                                               double2* G) {
  const long long row = blockIdx.y;  // (r * 6 + c) * n + jj
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= N) return;
  const double2* p = P + row * n;
  const double2* x = ex + (long long)a * n;
  double2 s = make_double2(0.0, 0.0);
  for (int ii = 0; ii < n; ++ii) s = cadd(s, cmul(p[ii], x[ii]));
  G[row * N + a] = s;
}

// I[c][b][a] += sum_r |scale sum_jj G[r][c][jj][a] ey[b][jj]|^2
__global__ void __launch_bounds__(256) k_img_acc(const double2* G, int R, int n, int N,
                                                 const double2* ey, double scale2, double* I) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y, c = blockIdx.z;
  if (a >= N) return;
  const double2* y = ey + (long long)b * n;
  double acc = 0.0;
  for (int r = 0; r < R; ++r) {
    const double2* g = G + ((long long)(r * 6 + c) * n) * N + a;
    double2 s = make_double2(0.0, 0.0);
    for (int jj = 0; jj < n; ++jj) s = cadd(s, cmul(g[(long long)jj * N], y[jj]));
    acc += (s.x * s.x + s.y * s.y) * scale2;
  }
  I[((long long)c * N + b) * N + a] += acc;
}

__global__ void __launch_bounds__(256) k_sqrt(double* I, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    I[i] = sqrt(I[i]);
}

__global__ void __launch_bounds__(256) k_zero(double* I, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    I[i] = 0.0;
}

}  // namespace

// Pupil field of the nquicklens = 0 path (microsbf.f:629-668,
// microssansinc.f:600-640): for every bin (ii, jj) of the (2 imax + 1)^2
// pupil with kx^2 + ky^2 <= (NA k0)^2, n = (kx, ky, nside kz) / k0 and
//   E = i k0^2 / (2 pi kz) sum_dipoles exp(-i k0 n.r) (p - (n.p) n)
// (zero elsewhere).  One block per (bin, right-hand side); the dipoles are
// summed per thread in a fixed stride order, then over threads in a fixed
// tree, so the result does not depend on timing.
__global__ void __launch_bounds__(256) k_lens_direct(const double2* FF, long long nd,
                                                     const int32_t* dipcell, int nx, int ny,
                                                     double ox, double oy, double oz, double d,
                                                     int imax, double dk, double k0, double nak,
                                                     double nside, double2* E) {
  const int n = 2 * imax + 1, nb = n * n;
  const int bin = blockIdx.x, r = blockIdx.y;
  const int ii = bin % n, jj = bin / n;
  const double kx = dk * (ii - imax), ky = dk * (jj - imax), t = kx * kx + ky * ky;
  double2* o = E + (long long)r * 3 * nb;
  if (t > nak * nak) {
    if (threadIdx.x < 3) o[(long long)threadIdx.x * nb + bin] = make_double2(0.0, 0.0);
    return;
  }
  const double kz = sqrt(k0 * k0 - t);
  const double n0 = kx / k0, n1 = ky / k0, n2 = nside * sqrt(1.0 - n0 * n0 - n1 * n1);
  const double2* p = FF + (long long)r * 3 * nd;
  double2 acc[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  for (long long i = threadIdx.x; i < nd; i += blockDim.x) {
    const long long cell = dipcell ? (long long)dipcell[i] : i;
    const long long a = cell % nx, b = (cell / nx) % ny, c = cell / ((long long)nx * ny);
    const double x = ox + d * a, y = oy + d * b, z = oz + d * c;
    double sn, cs;
    sincos(-k0 * (n0 * x + n1 * y + n2 * z), &sn, &cs);
    const double2 ph = make_double2(cs, sn);
    const double2 p0 = p[3 * i], p1 = p[3 * i + 1], p2 = p[3 * i + 2];
    const double2 np = make_double2(n0 * p0.x + n1 * p1.x + n2 * p2.x,
                                    n0 * p0.y + n1 * p1.y + n2 * p2.y);
    acc[0] = cadd(acc[0], cmul(ph, make_double2(p0.x - np.x * n0, p0.y - np.y * n0)));
    acc[1] = cadd(acc[1], cmul(ph, make_double2(p1.x - np.x * n1, p1.y - np.y * n1)));
    acc[2] = cadd(acc[2], cmul(ph, make_double2(p2.x - np.x * n2, p2.y - np.y * n2)));
  }
  __shared__ double2 red[3][256];
  for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = cadd(red[c][threadIdx.x], red[c][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 3) {
    // k0^2 / (-2 pi i kz) = i k0^2 / (2 pi kz)
    const double f = k0 * k0 / (2.0 * kPi * kz);
    const double2 v = red[threadIdx.x][0];
    o[(long long)threadIdx.x * nb + bin] = make_double2(-f * v.y, f * v.x);
  }
}

// Pupil step and size of the nquicklens = 0 path (microsbf.f:139-146,
// microsdf.f:140-147, microssansinc.f: the same): deltakx = 2 pi / (nfft2d
// d) / 2^k for the first k with imaxk0 = nint(k0 / deltakx) + 1 > 5.  The
// pupil must fit the image grid (2 imaxk0 + 1 <= nfft2d), which the
// reference does not check.
static std::string direct_geometry(int nfft2d, double k0, double d, int* imax, double* dk) {
  if (nfft2d > 16384) return "nfft2d too large";
  for (int k = 0; k < 60; ++k) {
    *dk = 2.0 * kPi / ((double)nfft2d * d) / std::ldexp(1.0, k);
    *imax = (int)std::lround(k0 / *dk) + 1;
    if (*imax > 5) break;
  }
  if (2 * *imax + 1 > nfft2d) return "In FFT diffraction for lens nfft2d too small!";
  return "";
}

// The condenser's incidences (microsbf.f:90-172), in the reference's order:
// polarisation outer, idelta, jdelta inner.
struct Incidence {
  int idelta, jdelta;
  double kx, ky, pp;
};

// dark: the darkfield microscope's ring of incidences (microsdf.f:159-221):
// ideltam = max(int(2 pi / (deltak NA_inc / k0)) + 1, 8) directions at
// phi = 2 pi idelta / ideltam on the condenser's aperture cone, both
// polarisations (npolainc is 0 there), polarisation outer.
static std::string incidence_list(const ifdda_ctx* c, const ifdda_microscope_opts& m,
                                  std::vector<Incidence>* out, double* deltak, bool dark) {
  const double k0 = m.k0, d = m.d;
  const double numaperk = k0 * m.numaperinc;
  const double xr = (c->cellbox[1] - c->cellbox[0]) * d, yr = (c->cellbox[3] - c->cellbox[2]) * d;
  const double deltakm = kPi / std::max(xr, yr);
  double dk;
  int imaxinc;
  if (m.direct_lens) {
    // nquicklens = 0 (microsbf.f:148-157, 179-183): deltak = deltakx
    // dnint(deltakm / deltakx) / 2^k for the first k with imaxinc > 2
    int imaxk0;
    double dkx;
    const std::string g = direct_geometry(m.nfft2d, k0, d, &imaxk0, &dkx);
    if (!g.empty()) return g;
    const double q = std::round(deltakm / dkx);
    if (q == 0.0) return "In FFT lens nfft2d too small";
    for (int k = 0;; ++k) {
      dk = dkx * q / std::ldexp(1.0, k);
      imaxinc = (int)std::lround(k0 / dk) + 1;
      if (imaxinc > 2) break;
    }
    if (dk >= numaperk) return "In FFT lens nfft2d too small";
  } else {
    const double dkx = 2.0 * kPi / (d * m.nfft2d);
    if (dkx >= numaperk) return "In FFT lens nfft2d too small:Increase size of FFT 1";
    int imul = (int)(deltakm / dkx);
    if (imul == 0) return "In FFT lens nfft2d too small:Increase size of FFT 2";
    for (;;) {
      dk = dkx * imul;
      imaxinc = (int)std::lround(k0 / dk) + 1;
      if (imaxinc > 3) break;
      if (--imul == 0) return "In FFT lens nfft2d too small:Increase size of FFT 3";
    }
  }
  *deltak = dk;
  if (dark) {
    const double dtheta = dk / k0 * m.numaperinc;
    const int ideltam = std::max((int)(2.0 * kPi / dtheta) + 1, 8);
    for (int ip = 0; ip < 2; ++ip)
      for (int id = 0; id < ideltam; ++id) {
        const double phi = (double)id * 2.0 * kPi / (double)ideltam;
        out->push_back({id, 0, std::cos(phi) * k0 * m.numaperinc,
                        std::sin(phi) * k0 * m.numaperinc, ip == 0 ? 0.0 : 1.0});
      }
    return "";
  }
  const int npol = m.npolainc == 0 ? 2 : 1;
  for (int ip = 0; ip < npol; ++ip) {
    const double pp = m.npolainc == 1 ? 0.0 : m.npolainc == 2 ? 1.0 : (ip == 0 ? 0.0 : 1.0);
    for (int id = -imaxinc; id <= imaxinc; ++id)
      for (int jd = -imaxinc; jd <= imaxinc; ++jd) {
        const double kx = id * dk, ky = jd * dk;
        if (kx * kx + ky * ky <= numaperk * numaperk) out->push_back({id, jd, kx, ky, pp});
      }
  }
  return "";
}

static bool mbuf_ensure(ifdda_ctx* c, size_t bytes) {
  if (bytes <= c->mcap) return true;
  if (c->mbuf) (void)hipFree(c->mbuf);
  c->mbuf = nullptr;
  c->mcap = 0;
  if (!hip_ok(hipMalloc(&c->mbuf, bytes), "hipMalloc microscope workspace")) return false;
  c->mcap = bytes;
  return true;
}

}  // namespace ifdda

using namespace ifdda;

static int incidences_api(ifdda_ctx* c, const ifdda_microscope_opts* m, int* ninc, double* kxy,
                          int cap, double* deltak, bool dark) {
  if (!c || !m || !ninc || !deltak) {
    set_error("ifdda_microscope_incidences: NULL argument");
    return IFDDA_ERR_ARG;
  }
  if (!c->have_dipoles) {
    set_error("ifdda_microscope_incidences: set the dipoles first");
    return IFDDA_ERR_STATE;
  }
  std::vector<Incidence> inc;
  const std::string msg = incidence_list(c, *m, &inc, deltak, dark);
  if (!msg.empty()) {
    set_error("ifdda_microscope_incidences: " + msg);
    return IFDDA_ERR_ARG;
  }
  *ninc = (int)inc.size();
  if (kxy)
    for (int i = 0; i < (int)inc.size() && i < cap; ++i) {
      kxy[3 * i] = inc[i].kx;
      kxy[3 * i + 1] = inc[i].ky;
      kxy[3 * i + 2] = inc[i].pp;
    }
  return IFDDA_OK;
}

extern "C" int ifdda_microscope_incidences(ifdda_ctx* c, const ifdda_microscope_opts* m, int* ninc,
                                           double* kxy, int cap, double* deltak) {
  return incidences_api(c, m, ninc, kxy, cap, deltak, false);
}

extern "C" int ifdda_microsdf_incidences(ifdda_ctx* c, const ifdda_microscope_opts* m, int* ninc,
                                         double* kxy, int cap, double* deltak) {
  return incidences_api(c, m, ninc, kxy, cap, deltak, true);
}

// The brightfield (microsbf.f) and darkfield (microsdf.f) microscopes: they
// differ in the incidences (lattice in the condenser disk / ring on its
// cone) and in how the direct wave enters the pupil (below).  nside = -1
// runs their reflection forms microsbfneg / microsdfneg (microssansinc.f,
// what cdmlib.f:3737-3796 calls for that side): the same incidences, the
// lens on z < 0, no incident wave, so images 3..5 equal 0..2.
static int microscope_run(ifdda_ctx* c, const ifdda_solve_opts* o, const ifdda_microscope_opts* m,
                          double* images, int* ninc, ifdda_solve_info* infos, int infos_cap,
                          volatile int* nstop, bool dark) {
  if (!c || !o || !m || !images || !ninc) {
    set_error(std::string(dark ? "ifdda_microsdf_device: " : "ifdda_microsbf_device: ") + "NULL argument");
    return IFDDA_ERR_ARG;
  }
  if (!(m->numaperinc > 0.0 && m->numaperinc < 1.0) || !(m->numaper > 0.0 && m->numaper <= 1.0)) {
    set_error(std::string(dark ? "ifdda_microsdf_device: " : "ifdda_microsbf_device: ") + "NA inc strictly between 0 and 1");
    return IFDDA_ERR_ARG;
  }
  if (m->nfft2d < 8 || m->nfft2d % 2 || !(m->k0 > 0.0) || !(m->d > 0.0) || !(m->gross > 0.0) ||
      (m->nside != 1 && m->nside != -1) || m->npolainc < 0 || m->npolainc > 2) {
    set_error(std::string(dark ? "ifdda_microsdf_device: " : "ifdda_microsbf_device: ") + "bad option (nfft2d even >= 8, k0, d, gross > 0, nside +-1, "
              "npolainc 0..2)");
    return IFDDA_ERR_ARG;
  }
  if (!c->have_dipoles || !c->green_ready || c->alpha_mode == kAlphaNone) {
    set_error(std::string(dark ? "ifdda_microsdf_device: " : "ifdda_microsbf_device: ") + "set the Green tensor, dipoles and polarisability first");
    return IFDDA_ERR_STATE;
  }
  const int N = m->nfft2d, h = N / 2;
  const double k0 = m->k0, d = m->d, nak = m->numaper * k0;
  std::vector<Incidence> inc;
  double dkinc = 0.0;
  std::string msg;
  const bool user = !dark && m->ninc_user > 0;
  if (user) {  // given plane waves inside the light cone
    if (!m->kxy_user) {
      set_error("ifdda_microsbf_device: ninc_user > 0 needs kxy_user");
      return IFDDA_ERR_ARG;
    }
    for (int i = 0; i < m->ninc_user; ++i) {
      const double kx = m->kxy_user[3 * i], ky = m->kxy_user[3 * i + 1], pp = m->kxy_user[3 * i + 2];
      if (!(kx * kx + ky * ky < k0 * k0) || !(pp >= 0.0 && pp <= 1.0)) {
        set_error("ifdda_microsbf_device: incidence " + std::to_string(i) +
                  " outside the light cone or pp outside [0, 1]");
        return IFDDA_ERR_ARG;
      }
      inc.push_back({0, 0, kx, ky, pp});
    }
  } else {
    msg = incidence_list(c, *m, &inc, &dkinc, dark);
  }
  int imax = 0;
  double dk = 0.0;
  if (msg.empty())
    msg = m->direct_lens ? direct_geometry(N, k0, d, &imax, &dk) : lens_geometry(N, k0, d, nak, &imax, &dk);
  if (!msg.empty()) {
    set_error(std::string(dark ? "ifdda_microsdf_device: " : "ifdda_microsbf_device: ") + "" + msg);
    return IFDDA_ERR_ARG;
  }
  const int ni = (int)inc.size();
  *ninc = ni;
  const int n = 2 * imax + 1, nb = n * n;
  const double2 E0 = make_double2(m->E0[0], m->E0[1]);

  // per-bin pupil matrices M_bin = zfocus sqrt(u_z / v_z) Rot(u -> v) (microsbf.f:710-790)
  std::vector<double2> M((size_t)nb * 9, make_double2(0.0, 0.0));
  for (int jj = 0; jj < n; ++jj)
    for (int ii = 0; ii < n; ++ii) {
      const double kx = dk * (ii - imax), ky = dk * (jj - imax), t = kx * kx + ky * ky;
      if (t > nak * nak) continue;
      const double kz = std::sqrt(k0 * k0 - t);
      const double2 zf = make_double2(std::cos(kz * m->zlens), std::sin(kz * m->zlens));
      double Rm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, sc = 1.0;
      if (t != 0.0) {
        // nside = -1 (microsbfneg / microsdfneg): u_z and v_z point to z < 0
        // (microssansinc.f:494-497)
        const double u0 = kx / k0, u1v = ky / k0,
                     u2 = std::sqrt(1.0 - u0 * u0 - u1v * u1v) * m->nside;
        const double v0 = -kx / k0 / m->gross, v1 = -ky / k0 / m->gross,
                     v2 = std::sqrt(1.0 - v0 * v0 - v1 * v1) * m->nside;
        double a1 = u1v * v2 - u2 * v1, a2 = -u0 * v2 + u2 * v0;
        const double co = u0 * v0 + u1v * v1 + u2 * v2, si = std::sqrt(a1 * a1 + a2 * a2);
        a1 /= si, a2 /= si;
        sc = std::sqrt(u2 / v2);
        const double R9[9] = {a1 * a1 + (1 - a1 * a1) * co, a1 * a2 * (1 - co), a2 * si,
                              a1 * a2 * (1 - co), a2 * a2 + (1 - a2 * a2) * co, -a1 * si,
                              -a2 * si, a1 * si, co};
        std::memcpy(Rm, R9, sizeof(Rm));
      }
      double2* mb = &M[((size_t)jj * n + ii) * 9];
      for (int q = 0; q < 9; ++q) mb[q] = make_double2(zf.x * Rm[q] * sc, zf.y * Rm[q] * sc);
    }
  // image-plane phase tables ex[a][ii] = exp(-2 pi i (a - h)(ii - imax) / N)
  std::vector<double2> ex((size_t)N * n);
  for (int a = 0; a < N; ++a)
    for (int ii = 0; ii < n; ++ii) {
      const long long q = ((long long)(a - h) * (ii - imax)) % N;
      const double ang = -2.0 * kPi * (double)q / (double)N;
      ex[(size_t)a * n + ii] = make_double2(std::cos(ang), std::sin(ang));
    }
  const double scale = dk * dk / m->gross;

  // batches of incidences
  const int B = m->max_batch > 0 ? std::min(m->max_batch, ni) : ni;
  const long long nd = c->nd;
  const size_t n3 = 3 * (size_t)nd;
  const size_t sz_E = (size_t)B * n3, sz_L = (size_t)B * 3 * nb, sz_P = (size_t)B * 6 * nb,
               sz_G = (size_t)B * 6 * n * N;
  // workspace slices, each rounded up to 256 bytes (the same rounding below)
  auto r256 = [](size_t nbytes) { return (nbytes + 255) & ~(size_t)255; };
  const size_t z2 = sizeof(double2);
  const size_t bytes = 3 * r256(z2 * sz_E) + r256(z2 * sz_L) + r256(z2 * sz_P) + r256(z2 * sz_G) +
                       r256(z2 * nb * 9) + r256(z2 * N * n) + r256(z2 * 3 * B) +
                       r256(sizeof(double) * 6 * N * N) + r256(sizeof(double) * 3 * B) +
                       r256(sizeof(int) * B);
  if (!mbuf_ensure(c, bytes)) return IFDDA_ERR_HIP;
  char* w = (char*)c->mbuf;
  char* const wend = w + bytes;
  auto take = [&](size_t nbytes) {
    char* p = w;
    w += r256(nbytes);
    return p;
  };
  double2* dEinc = (double2*)take(sizeof(double2) * sz_E);
  double2* dFF = (double2*)take(sizeof(double2) * sz_E);
  double2* dFL = (double2*)take(sizeof(double2) * sz_E);
  double2* dL = (double2*)take(sizeof(double2) * sz_L);
  double2* dP = (double2*)take(sizeof(double2) * sz_P);
  double2* dG = (double2*)take(sizeof(double2) * sz_G);
  double2* dM = (double2*)take(sizeof(double2) * nb * 9);
  double2* dex = (double2*)take(sizeof(double2) * N * n);
  double2* dev = (double2*)take(sizeof(double2) * 3 * B);
  double* dI = (double*)take(sizeof(double) * 6 * N * N);
  double* dkv = (double*)take(sizeof(double) * 3 * B);
  int* dbin = (int*)take(sizeof(int) * B);
  if (w > wend) {  // the slices above must match the size computed for them
    set_error(std::string(dark ? "ifdda_microsdf_device: " : "ifdda_microsbf_device: ") + "internal workspace layout error");
    return IFDDA_ERR_STATE;
  }
  hipStream_t s = c->stream;
  if (!hip_ok(hipMemcpyAsync(dM, M.data(), sizeof(double2) * M.size(), hipMemcpyHostToDevice, s), "copy M") ||
      !hip_ok(hipMemcpyAsync(dex, ex.data(), sizeof(double2) * ex.size(), hipMemcpyHostToDevice, s), "copy ex"))
    return IFDDA_ERR_HIP;
  hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, s, dI, (long long)6 * N * N);

  std::vector<ifdda_solve_info> inf(B);
  std::vector<double> hk(3 * B);
  std::vector<double2> he(3 * B), hinc(3 * B);
  std::vector<int> hbin(B);
  for (int i0 = 0; i0 < ni; i0 += B) {
    const int R = std::min(B, ni - i0);
    for (int r = 0; r < R; ++r) {
      const Incidence& q = inc[i0 + r];
      double kz;
      double2 e[3];
      plane_wave_vector(k0, E0, q.pp, q.kx, q.ky, e, &kz);
      hk[3 * r] = q.kx, hk[3 * r + 1] = q.ky, hk[3 * r + 2] = kz;
      for (int cc = 0; cc < 3; ++cc) he[3 * r + cc] = e[cc];
      // incident bin: deltakroutine(kxinc, kyinc, deltak, deltak, k0) with the
      // condenser step (microsbf.f:689-690), or with the pupil step
      // (microsdf.f:668-669)
      const double dkb = dark || user ? dk : dkinc;
      int ik = (int)std::lround(q.kx / dkb), jk = (int)std::lround(q.ky / dkb);
      if ((dkb * ik) * (dkb * ik) + (dkb * jk) * (dkb * jk) > k0 * k0)
        ik = (int)(q.kx / dkb), jk = (int)(q.ky / dkb);
      hbin[r] = -1;
      const double bx = dk * ik, by = dk * jk;
      // the objective on the far side (nside = -1) sees no incident wave:
      // microsbfneg / microsdfneg form the scattered images only
      if (m->nside == 1 && std::abs(ik) <= imax && std::abs(jk) <= imax &&
          bx * bx + by * by <= nak * nak) {
        hbin[r] = (jk + imax) * n + (ik + imax);
        double2 eb[3];
        double kzb;
        plane_wave_vector(k0, E0, q.pp, bx, by, eb, &kzb);  // at the bin's (kx, ky), the origin
        if (dark) {  // i E / (dkx dky) on every component (microsdf.f:712-717)
          for (int cc = 0; cc < 3; ++cc)
            hinc[3 * r + cc] = make_double2(-eb[cc].y / (dk * dk), eb[cc].x / (dk * dk));
        } else {
          hinc[3 * r] = make_double2(eb[0].x / (dk * dk), eb[0].y / (dk * dk));
          hinc[3 * r + 1] = make_double2(eb[1].x / (dk * dk), eb[1].y / (dk * dk));
          hinc[3 * r + 2] = make_double2(eb[2].x / dk, eb[2].y / dk);  // as microsbf.f:745-746
        }
      } else {
        hinc[3 * r] = hinc[3 * r + 1] = hinc[3 * r + 2] = make_double2(0.0, 0.0);
      }
    }
    if (!hip_ok(hipMemcpyAsync(dkv, hk.data(), sizeof(double) * 3 * R, hipMemcpyHostToDevice, s), "copy k") ||
        !hip_ok(hipMemcpyAsync(dev, he.data(), sizeof(double2) * 3 * R, hipMemcpyHostToDevice, s), "copy e"))
      return IFDDA_ERR_HIP;
    if (!launch_waves(dEinc, nullptr, R, nd, c->dipcell, c->g.nx, c->g.ny, m->origin, d, dkv, dev, s))
      return IFDDA_ERR_HIP;
    int rc = ifdda_solve_multi_device(c, o, R, dEinc, dFF, dFL, inf.data(), nstop, nullptr);
    if (rc != IFDDA_OK) return rc;
    for (int r = 0; r < R; ++r)
      if (infos && i0 + r < infos_cap) infos[i0 + r] = inf[r];
    if (nstop && *nstop == -1) {
      set_error(std::string(dark ? "ifdda_microsdf_device: " : "ifdda_microsbf_device: ") + "Calculation cancelled during iterative method");
      return IFDDA_OK;
    }
    if (m->direct_lens) {
      hipLaunchKernelGGL(k_lens_direct, dim3(nb, R), dim3(256), 0, s, (const double2*)dFF, nd,
                         (const int32_t*)c->dipcell, c->g.nx, c->g.ny, m->origin[0], m->origin[1],
                         m->origin[2], d, imax, dk, k0, nak, (double)m->nside, dL);
    } else {
      int im2;
      double dk2;
      rc = farfield_dev(c, R, dFF, c->tab, nd, N, k0, d, m->origin[0], m->origin[1], m->origin[2],
                        dL, nullptr, &im2, &dk2, s, 1, nak, m->nside);
      if (rc != IFDDA_OK) return rc;
    }
    // einc (double2) is staged in the ev slot after the waves kernel has consumed it
    if (!hip_ok(hipMemcpyAsync(dev, hinc.data(), sizeof(double2) * 3 * R, hipMemcpyHostToDevice, s), "copy einc") ||
        !hip_ok(hipMemcpyAsync(dbin, hbin.data(), sizeof(int) * R, hipMemcpyHostToDevice, s), "copy bins"))
      return IFDDA_ERR_HIP;
    hipLaunchKernelGGL(k_pupil, dim3((nb + 255) / 256, R), dim3(256), 0, s, (const double2*)dL, nb,
                       (const double2*)dM, (const int*)dbin, (const double2*)dev, dP);
    hipLaunchKernelGGL(k_img_x, dim3((N + 255) / 256, R * 6 * n), dim3(256), 0, s, (const double2*)dP,
                       n, N, (const double2*)dex, dG);
    hipLaunchKernelGGL(k_img_acc, dim3((N + 255) / 256, N, 6), dim3(256), 0, s, (const double2*)dG,
                       R, n, N, (const double2*)dex, scale * scale, dI);
    if (!hip_ok(hipGetLastError(), "microscope launch") || !hip_ok(hipStreamSynchronize(s), "microscope"))
      return IFDDA_ERR_HIP;
  }
  hipLaunchKernelGGL(k_sqrt, dim3(256), dim3(256), 0, s, dI, (long long)6 * N * N);
  if (!hip_ok(hipMemcpyAsync(images, dI, sizeof(double) * 6 * N * N, hipMemcpyDeviceToHost, s), "copy images") ||
      !hip_ok(hipStreamSynchronize(s), "microscope images"))
    return IFDDA_ERR_HIP;
  return IFDDA_OK;
}

extern "C" int ifdda_microsbf_device(ifdda_ctx* c, const ifdda_solve_opts* o,
                                     const ifdda_microscope_opts* m, double* images, int* ninc,
                                     ifdda_solve_info* infos, int infos_cap, volatile int* nstop) {
  return microscope_run(c, o, m, images, ninc, infos, infos_cap, nstop, false);
}

extern "C" int ifdda_microsdf_device(ifdda_ctx* c, const ifdda_solve_opts* o,
                                     const ifdda_microscope_opts* m, double* images, int* ninc,
                                     ifdda_solve_info* infos, int infos_cap, volatile int* nstop) {
  return microscope_run(c, o, m, images, ninc, infos, infos_cap, nstop, true);
}
