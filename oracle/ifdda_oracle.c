/*
 * ifdda_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * FP64 CPU restatement of the IF-DDA FFT-accelerated DDA solve, written
 * from the reference's documented algorithm; each function cites the
 * reference file:line it follows.  Loaded only by tests/, smoke() and
 * bench.py's cpu_baseline leg, as the checker.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   orc_propa_espace_libre  closed form limits, symmetry, compiled reference
 *   orc_propa_int_sim       compiled reference (propaespacelibreintsim_),
 *                           Simpson convergence against a finer rule
 *   orc_fft1d / orc_fft3d   numpy.fft, brute-force DFT
 *   orc_green_fft           compiled reference (greencalculfft_,
 *                           greencalculquadfft_) and golden fixtures
 *   orc_matvec              orc_matvec_direct (O(Nd^2) sum), compiled
 *                           reference (produitfftmatvect3_/..opt_), fixtures
 *   orc_inverserig          compiled reference (inverserig_) trajectories and
 *                           fields, dense numpy.linalg.solve on small systems
 */
#include "ifdda_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TWO_PI 6.283185307179586476925286766559

/* ------------------------------------------------------------------ */
/* Green tensor                                                        */
/* ------------------------------------------------------------------ */

/* espace_libre.f:2-43: T = e^{ikR}[(3RR/R^2 - I)(1/R - ik)/R^2
 *                                  + (I - RR/R^2) k^2/R], 0 at R = 0.
 * R = (x,y,z) - (x0,y0,z0); no 1/(4 pi eps0) factor (Gaussian units). */
void orc_propa_espace_libre(double x, double y, double z, double x0, double y0,
                            double z0, double k0, orc_zc T[9]) {
  double r[3] = {x - x0, y - y0, z - z0};
  double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  double R = sqrt(r2);
  if (R == 0.0) {
    for (int a = 0; a < 9; ++a) T[a] = 0.0;
    return;
  }
  orc_zc c1 = (1.0 / R - I * k0) / r2;
  orc_zc c2 = k0 * k0 / R;
  orc_zc ph = cexp(I * k0 * R);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double rr = r[a] * r[b] / r2;
      double id = (a == b) ? 1.0 : 0.0;
      T[3 * a + b] = ((3.0 * rr - id) * c1 + (id - rr) * c2) * ph;
    }
}

/* propaesplibreintsim.f:2-107.  Zero when R <= d/10; the point value when
 * R >= lim*lambda; otherwise a 129^3-node composite Simpson average of T over
 * the cube of side d centred on the observation point (x,y,z), with the
 * off-diagonal entries zeroed along axes whose offset is <= d/10. */
void orc_propa_int_sim(double x, double y, double z, double x0, double y0,
                       double z0, double k0, double d, double lim,
                       orc_zc T[9]) {
  const double lambda = TWO_PI / k0;
  const double dm = 0.1 * d;
  double R = sqrt((x - x0) * (x - x0) + (y - y0) * (y - y0) +
                  (z - z0) * (z - z0));
  if (R <= dm) {
    for (int a = 0; a < 9; ++a) T[a] = 0.0;
    return;
  }
  if (R >= lim * lambda) {
    orc_propa_espace_libre(x, y, z, x0, y0, z0, k0, T);
    return;
  }
  enum { NSTEP = 128 };
  double w[NSTEP + 1];
  for (int i = 0; i <= NSTEP; ++i)
    w[i] = (i == 0 || i == NSTEP) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
  const double h = d / NSTEP;
  const double ax = x - 0.5 * d, ay = y - 0.5 * d, az = z - 0.5 * d;
  const double fx = fabs(x - x0) <= dm ? 0.0 : 1.0;
  const double fy = fabs(y - y0) <= dm ? 0.0 : 1.0;
  const double fz = fabs(z - z0) <= dm ? 0.0 : 1.0;
  orc_zc s[6] = {0, 0, 0, 0, 0, 0};
  orc_zc P[9];
  for (int i = 0; i <= NSTEP; ++i)
    for (int j = 0; j <= NSTEP; ++j)
      for (int k = 0; k <= NSTEP; ++k) {
        double ww = w[i] * w[j] * w[k];
        orc_propa_espace_libre(ax + h * i, ay + h * j, az + h * k, x0, y0, z0,
                               k0, P);
        s[0] += P[0] * ww;
        s[1] += P[1] * ww;
        s[2] += P[2] * ww;
        s[3] += P[4] * ww;
        s[4] += P[5] * ww;
        s[5] += P[8] * ww;
      }
  double sc = (h / 3.0) * (h / 3.0) * (h / 3.0) / (d * d * d);
  T[0] = s[0] * sc;
  T[1] = T[3] = s[1] * sc * fx * fy;
  T[2] = T[6] = s[2] * sc * fx * fz;
  T[4] = s[3] * sc;
  T[5] = T[7] = s[4] * sc * fy * fz;
  T[8] = s[5] * sc;
}

/* ------------------------------------------------------------------ */
/* FFT: recursive mixed radix on the smallest prime factor             */
/* ------------------------------------------------------------------ */

static int smallest_factor(int n) {
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

/* out[0..n) = sum_t in[t*s] exp(sign 2 pi i t k / n) */
static void dft_rec(orc_zc *out, const orc_zc *in, int n, int s, int sign,
                    orc_zc *tmp, const orc_zc *tw, int N) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = smallest_factor(n), m = n / p;
  for (int r = 0; r < p; ++r)
    dft_rec(out + r * m, in + r * s, m, s * p, sign, tmp, tw, N);
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) tmp[r] = out[r * m + k];
    for (int q = 0; q < p; ++q) {
      orc_zc acc = 0;
      int t = k + q * m;
      for (int r = 0; r < p; ++r) {
        long e = ((long)r * t) % n; /* exp(sign 2 pi i e/n) = tw[e N/n] */
        acc += tmp[r] * tw[e * (N / n)];
      }
      out[t] = acc;
    }
  }
}

void orc_fft1d(orc_zc *x, int n, int sign) {
  if (n < 2) return;
  orc_zc *out = malloc(sizeof(orc_zc) * n);
  orc_zc *tmp = malloc(sizeof(orc_zc) * n);
  orc_zc *tw = malloc(sizeof(orc_zc) * n);
  for (int e = 0; e < n; ++e) {
    double ang = sign * TWO_PI * (double)e / (double)n;
    tw[e] = cos(ang) + I * sin(ang);
  }
  dft_rec(out, x, n, 1, sign, tmp, tw, n);
  free(tw);
  memcpy(x, out, sizeof(orc_zc) * n);
  free(out);
  free(tmp);
}

/* fftsingletonz.f:21-43 transforms x, y, z in turn; the order does not
 * change the exact result. */
void orc_fft3d(orc_zc *v, int n1, int n2, int n3, int sign) {
  int nmax = n1 > n2 ? n1 : n2;
  if (n3 > nmax) nmax = n3;
  orc_zc *line = malloc(sizeof(orc_zc) * nmax);
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j) orc_fft1d(v + (size_t)n1 * (j + (size_t)n2 * k), n1, sign);
  for (int k = 0; k < n3; ++k)
    for (int i = 0; i < n1; ++i) {
      for (int j = 0; j < n2; ++j) line[j] = v[i + (size_t)n1 * (j + (size_t)n2 * k)];
      orc_fft1d(line, n2, sign);
      for (int j = 0; j < n2; ++j) v[i + (size_t)n1 * (j + (size_t)n2 * k)] = line[j];
    }
  for (int j = 0; j < n2; ++j)
    for (int i = 0; i < n1; ++i) {
      for (int k = 0; k < n3; ++k) line[k] = v[i + (size_t)n1 * (j + (size_t)n2 * k)];
      orc_fft1d(line, n3, sign);
      for (int k = 0; k < n3; ++k) v[i + (size_t)n1 * (j + (size_t)n2 * k)] = line[k];
    }
  free(line);
}

/* Circulant offset of padded index idx for a box of n cells on L points:
 * idx < n -> idx, idx > L-n -> idx-L, otherwise unused (zero).  For L = 2n
 * this is calculgreenfft.f:29-51 (zero plane at idx = n). */
static int circ_offset(int idx, int n, int L, int *ok) {
  if (idx < n) { *ok = 1; return idx; }
  if (idx > L - n) { *ok = 1; return idx - L; }
  *ok = 0;
  return 0;
}

/* calculgreenfft.f:1-104 / calculgreenquadfft.f:1-105 */
void orc_green_fft(int nx, int ny, int nz, int L1, int L2, int L3, double d,
                   double k0, int nquad, orc_zc *G) {
  size_t M = (size_t)L1 * L2 * L3;
  for (int kk = 0; kk < L3; ++kk)
    for (int jj = 0; jj < L2; ++jj)
      for (int ii = 0; ii < L1; ++ii) {
        size_t idx = ii + (size_t)L1 * (jj + (size_t)L2 * kk);
        int okx, oky, okz;
        int i = circ_offset(ii, nx, L1, &okx);
        int j = circ_offset(jj, ny, L2, &oky);
        int k = circ_offset(kk, nz, L3, &okz);
        orc_zc T[9];
        if (!(okx && oky && okz)) {
          for (int a = 0; a < 9; ++a) T[a] = 0.0;
        } else if (nquad != 0 && i * i + j * j + k * k <= nquad * nquad) {
          orc_propa_int_sim(0.0, 0.0, 0.0, (double)i * d, (double)j * d,
                            (double)k * d, k0, d, 1.0, T);
        } else {
          orc_propa_espace_libre(0.0, 0.0, 0.0, (double)i * d, (double)j * d,
                                 (double)k * d, k0, T);
        }
        G[0 * M + idx] = T[0];
        G[1 * M + idx] = T[1];
        G[2 * M + idx] = T[2];
        G[3 * M + idx] = T[4];
        G[4 * M + idx] = T[5];
        G[5 * M + idx] = T[8];
      }
  for (int c = 0; c < 6; ++c) orc_fft3d(G + c * M, L1, L2, L3, +1);
}

/* ------------------------------------------------------------------ */
/* Matvec                                                              */
/* ------------------------------------------------------------------ */

/* produitfftmatvect3.f:2-129 (tabdip) and produitfftmatvectopt.f:3-122
 * (tabdip == NULL, dipole index = box index). */
void orc_matvec(const orc_zc *G, const int32_t *tabdip, int nx, int ny, int nz,
                int L1, int L2, int L3, const orc_zc *FFX, orc_zc *FFB) {
  size_t M = (size_t)L1 * L2 * L3;
  orc_zc *v = calloc(3 * M, sizeof(orc_zc));
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t cell = i + (size_t)nx * (j + (size_t)ny * k);
        long dip;
        if (tabdip) {
          if (tabdip[cell] == 0) continue;
          dip = tabdip[cell] - 1;
        } else {
          dip = (long)cell;
        }
        size_t idx = i + (size_t)L1 * (j + (size_t)L2 * k);
        for (int c = 0; c < 3; ++c) v[c * M + idx] = FFX[3 * dip + c];
      }
  for (int c = 0; c < 3; ++c) orc_fft3d(v + c * M, L1, L2, L3, +1);
  for (size_t n = 0; n < M; ++n) {
    orc_zc ex = v[n], ey = v[M + n], ez = v[2 * M + n];
    orc_zc gxx = G[n], gxy = G[M + n], gxz = G[2 * M + n], gyy = G[3 * M + n],
           gyz = G[4 * M + n], gzz = G[5 * M + n];
    v[n] = gxx * ex + gxy * ey + gxz * ez;
    v[M + n] = gxy * ex + gyy * ey + gyz * ez;
    v[2 * M + n] = gxz * ex + gyz * ey + gzz * ez;
  }
  for (int c = 0; c < 3; ++c) orc_fft3d(v + c * M, L1, L2, L3, -1);
  double dM = (double)M;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t cell = i + (size_t)nx * (j + (size_t)ny * k);
        long dip;
        if (tabdip) {
          if (tabdip[cell] == 0) continue;
          dip = tabdip[cell] - 1;
        } else {
          dip = (long)cell;
        }
        size_t idx = i + (size_t)L1 * (j + (size_t)L2 * k);
        for (int c = 0; c < 3; ++c) FFB[3 * dip + c] -= v[c * M + idx] / dM;
      }
  free(v);
}

/* The same operator summed directly over dipole pairs (no FFT). */
void orc_matvec_direct(const int32_t *tabdip, int nx, int ny, int nz, double d,
                       double k0, const orc_zc *FFX, orc_zc *FFB) {
  size_t N = (size_t)nx * ny * nz;
  for (size_t a = 0; a < N; ++a) {
    long da = tabdip ? tabdip[a] - 1 : (long)a;
    if (da < 0) continue;
    int ia = a % nx, ja = (a / nx) % ny, ka = a / ((size_t)nx * ny);
    orc_zc acc[3] = {0, 0, 0};
    for (size_t b = 0; b < N; ++b) {
      long db = tabdip ? tabdip[b] - 1 : (long)b;
      if (db < 0 || b == a) continue;
      int ib = b % nx, jb = (b / nx) % ny, kb = b / ((size_t)nx * ny);
      orc_zc T[9];
      orc_propa_espace_libre(ia * d, ja * d, ka * d, ib * d, jb * d, kb * d, k0, T);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) acc[r] += T[3 * r + c] * FFX[3 * db + c];
    }
    for (int r = 0; r < 3; ++r) FFB[3 * da + r] -= acc[r];
  }
}

/* ------------------------------------------------------------------ */
/* Krylov solvers + driver                                             */
/* ------------------------------------------------------------------ */

typedef struct {
  const orc_zc *G;
  const int32_t *tab;
  int nx, ny, nz, L1, L2, L3, nb, nmax;
  const orc_zc *pola;
  orc_zc *ax;
  int ncompte;
} amul_ctx;

/* out = A v = v - T (alpha v): inverserig.f:54-87 (xr = xi; xi = alpha.xr;
 * produitfftmatvect3(XI -> XR)).  polarisa is (nmax,3,3) column-major. */
static void amul(amul_ctx *c, const orc_zc *v, orc_zc *out) {
  for (int i = 0; i < c->nb; ++i)
    for (int a = 0; a < 3; ++a) {
      orc_zc s = 0;
      for (int b = 0; b < 3; ++b)
        s += c->pola[i + (size_t)c->nmax * (a + 3 * b)] * v[3 * i + b];
      c->ax[3 * i + a] = s;
    }
  if (out != v) memcpy(out, v, sizeof(orc_zc) * 3 * c->nb);
  orc_matvec(c->G, c->tab, c->nx, c->ny, c->nz, c->L1, c->L2, c->L3, c->ax, out);
  c->ncompte++;
}

static orc_zc dotc(const orc_zc *a, const orc_zc *b, int n) { /* sum conj(a) b */
  orc_zc s = 0;
  for (int i = 0; i < n; ++i) s += conj(a[i]) * b[i];
  return s;
}

static void trace_push(orc_trace *t, orc_zc s0, orc_zc s1, orc_zc s2, orc_zc s3,
                       orc_zc s4, double res) {
  if (!t || t->n >= t->cap) return;
  orc_zc *p = t->s + 5 * t->n;
  p[0] = s0; p[1] = s1; p[2] = s2; p[3] = s3; p[4] = s4;
  t->res[t->n] = res;
  t->n++;
}

/* gpbicg.f:81-338.  Returns STATUS; x holds the iterate. */
static int gpbicg(amul_ctx *C, orc_zc *x, const orc_zc *b, int n, int maxit,
                  double tol, int *itno, int *steperr, orc_trace *tr) {
  orc_zc *W = calloc((size_t)n * 12, sizeof(orc_zc));
  orc_zc *r0s = W, *p = W + n, *r = W + 2 * n, *y = W + 3 * n, *t = W + 4 * n,
         *Ap = W + 5 * n, *At = W + 6 * n, *u = W + 7 * n, *w = W + 8 * n,
         *z = W + 9 * n, *told = W + 10 * n, *Ax = W + 11 * n;
  double norm = 0;
  for (int i = 0; i < n; ++i) norm += creal(b[i] * conj(b[i]));
  norm = sqrt(norm);
  amul(C, x, Ax);
  orc_zc r0rn = 0;
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] - Ax[i];
    r0s[i] = conj(r[i]);
    r0rn += r0s[i] * r[i];
  }
  orc_zc alpha = 0, beta = 0, eta = 0, dzeta = 0;
  int status = 0, it = -1;
  *steperr = -1;
  for (;;) {
    ++it;
    for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - u[i]);
    amul(C, p, Ap);
    orc_zc ct = 0;
    for (int i = 0; i < n; ++i) ct += r0s[i] * Ap[i];
    if (ct == 0) { status = -1; *steperr = 1; break; }
    alpha = r0rn / ct;
    for (int i = 0; i < n; ++i) {
      y[i] = t[i] - r[i] + alpha * (Ap[i] - w[i]);
      told[i] = t[i];
      t[i] = r[i] - alpha * Ap[i];
    }
    amul(C, t, At);
    if (it == 0) {
      eta = 0;
      orc_zc num = dotc(At, t, n), den = dotc(At, At, n);
      if (den == 0) { status = -1; *steperr = 2; break; }
      dzeta = num / den;
    } else {
      orc_zc c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
      for (int i = 0; i < n; ++i) {
        c1 += conj(At[i]) * At[i];
        c2 += conj(y[i]) * y[i];
        c3 += conj(At[i]) * y[i];
        c4 += conj(At[i]) * t[i];
        c5 += conj(y[i]) * t[i];
      }
      orc_zc den = c1 * c2 - c3 * conj(c3);
      if (den == 0) { status = -1; *steperr = 3; break; }
      dzeta = (c2 * c4 - c5 * c3) / den;
      eta = (c1 * c5 - conj(c3) * c4) / den;
    }
    double res = 0;
    for (int i = 0; i < n; ++i) {
      u[i] = dzeta * Ap[i] + eta * (told[i] - r[i] + beta * u[i]);
      z[i] = dzeta * r[i] + eta * z[i] - alpha * u[i];
      x[i] = x[i] + alpha * p[i] + z[i];
      r[i] = t[i] - eta * y[i] - dzeta * At[i];
      res += creal(r[i] * conj(r[i]));
    }
    res = sqrt(res) / norm;
    trace_push(tr, alpha, beta, eta, dzeta, r0rn, res);
    if (res <= tol) { status = 1; break; }
    orc_zc ct2 = 0;
    for (int i = 0; i < n; ++i) ct2 += r0s[i] * r[i];
    if (r0rn == 0) { status = -1; *steperr = 4; break; }
    beta = alpha * ct2 / dzeta / r0rn;
    r0rn = ct2;
    for (int i = 0; i < n; ++i) w[i] = At[i] + beta * Ap[i];
    if (it > maxit) { status = 1; *steperr = 0; break; }
  }
  *itno = it;
  free(W);
  return status;
}

/* qmrbicgstab.f:1-253 (methode 1 or 2). */
static int qmrbicgstab(amul_ctx *C, orc_zc *x, const orc_zc *b, int n, int maxit,
                       double tol, int methode, int *itno, int *steperr,
                       orc_trace *tr) {
  orc_zc *W = calloc((size_t)n * 10, sizeof(orc_zc));
  orc_zc *r = W, *p = W + n, *v = W + 2 * n, *dd = W + 3 * n, *rt = W + 4 * n,
         *s = W + 5 * n, *xt = W + 6 * n, *dt = W + 7 * n, *t = W + 8 * n,
         *Ax = W + 9 * n;
  double norm = 0;
  for (int i = 0; i < n; ++i) norm += creal(b[i] * conj(b[i]));
  norm = sqrt(norm);
  amul(C, x, Ax);
  orc_zc ct = 0;
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] - Ax[i];
    rt[i] = r[i];
    ct += r[i] * conj(r[i]);
  }
  orc_zc rho = 1, alpha = 1, omega = 1, tau = csqrt(ct), theta = 0, eta = 0;
  orc_zc thetat, taut, etat, c;
  int status = 0, it = 0;
  *steperr = -1;
  for (;;) {
    ++it;
    orc_zc rho0 = rho;
    rho = dotc(rt, r, n);
    orc_zc beta = rho * alpha / rho0 / omega;
    for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    amul(C, p, v);
    alpha = dotc(rt, v, n);
    if (cabs(alpha) == 0.0) { status = -1; *steperr = 1; break; }
    alpha = rho / alpha;
    thetat = 0;
    for (int i = 0; i < n; ++i) {
      s[i] = r[i] - alpha * v[i];
      thetat += conj(s[i]) * s[i];
    }
    if (cabs(tau) == 0.0) { status = -1; *steperr = 2; break; }
    thetat = csqrt(thetat) / tau;
    c = 1.0 / csqrt(1.0 + thetat * thetat);
    taut = tau * thetat * c;
    etat = c * c * alpha;
    for (int i = 0; i < n; ++i) {
      dt[i] = p[i] + theta * theta * eta / alpha * dd[i];
      xt[i] = x[i] + etat * dt[i];
    }
    amul(C, s, t);
    orc_zc om, cm;
    if (methode == 1) {
      om = dotc(t, t, n);
      cm = dotc(s, t, n);
    } else {
      om = dotc(s, t, n);
      cm = dotc(s, s, n);
    }
    if (cabs(om) == 0.0) { status = -1; *steperr = 3; break; }
    omega = cm / om;
    theta = 0;
    for (int i = 0; i < n; ++i) {
      r[i] = s[i] - omega * t[i];
      theta += conj(r[i]) * r[i];
    }
    if (cabs(taut) == 0.0) { status = -1; *steperr = 4; break; }
    theta = csqrt(theta) / taut;
    c = 1.0 / csqrt(1.0 + theta * theta);
    tau = taut * theta * c;
    eta = c * c * omega;
    if (cabs(omega) == 0.0) { status = -1; *steperr = 5; break; }
    for (int i = 0; i < n; ++i) {
      dd[i] = s[i] + thetat * thetat * etat / omega * dt[i];
      x[i] = xt[i] + eta * dd[i];
    }
    double res = sqrt((double)(it + 1)) * cabs(tau) / norm;
    trace_push(tr, alpha, beta, omega, tau, eta, res);
    if (res <= tol) { status = 1; break; }
    if (it > maxit) { status = 1; *steperr = 0; break; }
  }
  *itno = it;
  free(W);
  return status;
}

/* tfqmr.f:1-218 (PIM TFQMR: one matvec per half step). */
static int tfqmr(amul_ctx *C, orc_zc *x, const orc_zc *b, int n, int maxit,
                 double tol, int *itno, int *steperr, orc_trace *tr) {
  orc_zc *W = calloc((size_t)n * 9, sizeof(orc_zc));
  orc_zc *u = W, *w = W + n, *r = W + 2 * n, *v = W + 3 * n, *dd = W + 4 * n,
         *rt = W + 5 * n, *u1 = W + 6 * n, *Au = W + 7 * n, *Auo = W + 8 * n;
  double norm = 0;
  for (int i = 0; i < n; ++i) norm += creal(b[i] * conj(b[i]));
  norm = sqrt(norm);
  amul(C, x, Au);
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] - Au[i];
    w[i] = r[i];
    u[i] = r[i];
  }
  amul(C, u, v);
  orc_zc rho = 0;
  for (int i = 0; i < n; ++i) {
    dd[i] = 0;
    Au[i] = v[i];
    rt[i] = r[i];
    rho += conj(rt[i]) * r[i];
  }
  orc_zc tau = norm, theta = 0, eta = 0, alpha = 0, beta = 0;
  int status = 0, it = -1;
  *steperr = -1;
  for (;;) {
    ++it;
    if (it % 2 == 0) {
      alpha = rho / dotc(rt, v, n);
      for (int i = 0; i < n; ++i) u1[i] = u[i] - alpha * v[i];
    }
    orc_zc th = 0;
    for (int i = 0; i < n; ++i) {
      w[i] = w[i] - alpha * Au[i];
      dd[i] = u[i] + theta * theta / alpha * eta * dd[i];
      th += conj(w[i]) * w[i];
    }
    theta = csqrt(th) / tau;
    orc_zc c = 1.0 / csqrt(1.0 + theta * theta);
    tau = tau * theta * c;
    eta = c * c * alpha;
    for (int i = 0; i < n; ++i) x[i] = x[i] + eta * dd[i];
    double res = sqrt((double)(it + 1)) * cabs(tau) / norm;
    trace_push(tr, alpha, beta, rho, tau, eta, res);
    if (res <= tol) { status = 1; break; }
    if (it % 2 == 1) {
      orc_zc rho0 = rho;
      rho = dotc(rt, w, n);
      beta = rho / rho0;
      for (int i = 0; i < n; ++i) {
        u1[i] = w[i] + beta * u[i];
        Auo[i] = Au[i];
      }
    }
    memcpy(u, u1, sizeof(orc_zc) * n);
    amul(C, u, Au);
    if (it % 2 == 1)
      for (int i = 0; i < n; ++i) v[i] = Au[i] + beta * (Auo[i] + beta * v[i]);
    if (it > maxit) { status = 1; *steperr = 0; break; }
  }
  *itno = it;
  free(W);
  return status;
}

/* inverserig.f:1-1090 (branches GPBICG1 :44-109, TFQMR :607-678,
 * QMRBICGSTAB1/2 :816-947). */
void orc_inverserig(const orc_zc *G, const int32_t *tabdip, int nx, int ny,
                    int nz, int L1, int L2, int L3, int nbsphere,
                    const orc_zc *polarisa, int nmax, int method, double tol,
                    int maxit, const orc_zc *FF0, orc_zc *FF, orc_zc *FFloc,
                    orc_solve_info *info, orc_trace *trace) {
  int n = 3 * nbsphere;
  amul_ctx C = {G, tabdip, nx, ny, nz, L1, L2, L3, nbsphere, nmax, polarisa,
                malloc(sizeof(orc_zc) * n), 0};
  orc_zc *x = malloc(sizeof(orc_zc) * n), *Ax = malloc(sizeof(orc_zc) * n);
  memcpy(x, FF0, sizeof(orc_zc) * n); /* x0 = E_inc (inverserig.f:36-42) */
  int itno = 0, steperr = -1, status;
  memset(info, 0, sizeof(*info));
  if (method == ORC_GPBICG1)
    status = gpbicg(&C, x, FF0, n, maxit, tol, &itno, &steperr, trace);
  else if (method == ORC_QMRBICGSTAB1 || method == ORC_QMRBICGSTAB2)
    status = qmrbicgstab(&C, x, FF0, n, maxit, tol,
                         method == ORC_QMRBICGSTAB1 ? 1 : 2, &itno, &steperr, trace);
  else
    status = tfqmr(&C, x, FF0, n, maxit, tol, &itno, &steperr, trace);
  info->status = status;
  info->steperr = steperr;
  if (status < 0) {
    info->nstop = 1;
    info->ncompte = C.ncompte;
    info->nloop = itno;
  } else {
    memcpy(FFloc, x, sizeof(orc_zc) * n);
    amul(&C, x, Ax);
    double norm = 0, t1 = 0;
    for (int i = 0; i < n; ++i) {
      norm += creal(FF0[i] * conj(FF0[i]));
      orc_zc e = Ax[i] - FF0[i];
      t1 += creal(e * conj(e));
      FF[i] = C.ax[i];
    }
    info->tol1 = sqrt(t1) / sqrt(norm);
    info->ncompte = C.ncompte;
    /* GPBICG1 and TFQMR start at ITNO=-1 and the driver adds one;
     * QMRBICGSTAB counts from 1 and the driver adds nothing. */
    info->nloop = (method == ORC_GPBICG1 || method == ORC_TFQMR) ? itno + 1 : itno;
  }
  free(C.ax);
  free(x);
  free(Ax);
}
