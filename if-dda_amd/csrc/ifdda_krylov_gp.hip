// ifdda_krylov_gp.hip -- the other Krylov methods inverserig.f dispatches to:
//   GPBICGsafe    gpbicgsafe.f     (inverserig.f:176-240)
//   GPBICGplus    gpbicgplus.f     (inverserig.f:241-305)
//   GPBICGAR      gpbicgar.f       (inverserig.f:306-370; 'GPBICGAR2' lands here too,
//                                   the driver compares methodeit(1:8) first)
//   BICGSTARPLUS  bicgstarplus.f   (inverserig.f:435-502)
//   GPBICOR       gpbicor.f        (inverserig.f:948-1013)
//   CORS          cors.f           (inverserig.f:1014-1078)
//   QMRCLA        qmrpim.f         (inverserig.f:503-606; PIMZQMR, the
//                                   transposed operator from MvOp::mvt*)
// GPBICG2 (gpbicg2.f) shares the GPBICG kernels of ifdda_krylov.hip.
//
// Same device scheme as the other solvers: every scalar in KState (FP64),
// one-block finish kernels over fixed per-block partials, kernels that return
// at once when KState.done is set.  The element-wise updates follow the
// reference's expressions term by term.  Work columns the reference keeps
// only to carry a value from one loop to the next inside an iteration (a
// t = r - alpha Ap used once) are recomputed where they are used instead of
// stored; in FP32 the recomputed value is rounded as a stored one would be,
// so it matches what a store and a load return.  Columns the reference leaves
// uninitialised before their first (zero-weighted) use start at zero.
#include <algorithm>
#include <cstring>

#include "../../include/ifdda.h"
#include "ifdda_internal.h"
#include "ifdda_krylov.h"

namespace ifdda {

__device__ __forceinline__ void kbreak(KState* st, int steperr) {
  st->status = -1;
  st->steperr = steperr;
  st->done = 1;
}

// Vector kernels: the block's RHS (VN::rhs), its state, its slice of the vectors and of
// the partial sums; leave when the solve of this RHS has ended.
#define VK_ENTER()                                   \
  const int rh_ = n.rhs();                           \
  st += rh_;                                         \
  v = v.shifted(n.off(rh_));                         \
  part += (long long)rh_ * kMaxDots * nblk;          \
  if (st->done) return;
// Finish kernels: reduce NR partial sums into st->dots; thread 0 goes on.
#define FK_ENTER(NR)                                 \
  const int rh_ = blockIdx.y;                        \
  st += rh_;                                         \
  part += (long long)rh_ * kMaxDots * nblk;          \
  if (st->done) return;                              \
  finish_sum<NR>(st, part, nblk);                    \
  if (threadIdx.x != 0) return;

__device__ __forceinline__ double2 zcdot(double2 a, double2 b) { return zmul(zconj(a), b); }
__device__ __forceinline__ double2 zscale(double a, double2 b) { return make_double2(a * b.x, a * b.y); }

// The recurrence residual test and the iteration cap shared by the GPBiCG
// variants (gpbicgsafe.f:239-265 and the same lines of the others): dots[0]
// = |r|^2, dots[1] = conj(r0) r.
__device__ __forceinline__ void gp_tail(KState* st, double2* trace, int steperr_r0) {
  const double res = sqrt(st->dots[0].x) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->alpha, st->beta, st->eta, st->dzeta, st->r0rn, res);
  if (res <= st->tol) {
    st->status = 1;
    st->done = 1;
    return;
  }
  const double2 ct = st->dots[1];
  if (zzero(st->r0rn)) {
    kbreak(st, steperr_r0);
    return;
  }
  if (zzero(st->dzeta)) {  // gpbicgar.f:200-204 (STEPERR 6); the others divide by zero there
    kbreak(st, 6);
    return;
  }
  st->beta = zdiv(zdiv(zmul(st->alpha, ct), st->dzeta), st->r0rn);
  st->r0rn = ct;
  if (st->it > st->maxit) {
    st->status = 1;
    st->steperr = 0;
    st->done = 1;
    return;
  }
  st->it++;
}

// zeta and eta from five products (gpbicgsafe.f:139-170, gpbicgar.f:104-137):
// d1 = (y,y), d2 = (Ar,r), d3 = (y,r), d4 = (Ar,y), d5 = (Ar,Ar); the first
// iteration takes the steepest-descent step zeta = (Ar,r)/(Ar,Ar), eta = 0.
__device__ __forceinline__ void gp_zeta_eta(KState* st, double2 d1, double2 d2, double2 d3,
                                            double2 d4, double2 d5) {
  if (st->it == 0) {
    if (zzero(d5)) {
      kbreak(st, 2);
      return;
    }
    st->dzeta = zdiv(d2, d5);
    st->eta = make_double2(0.0, 0.0);
  } else {
    const double2 den = zsub(zmul(d5, d1), zmul(d4, zconj(d4)));
    if (zzero(den)) {
      kbreak(st, 3);
      return;
    }
    st->dzeta = zdiv(zsub(zmul(d1, d2), zmul(d3, d4)), den);
    st->eta = zdiv(zsub(zmul(d5, d3), zmul(zconj(d4), d2)), den);
  }
}

// r = b - A x0 (b when x0 = 0), r0 = r, the listed columns zero; partial
// sum r conj(r) (R0RN).  Shared by the GPBiCG variants.
template <typename V>
__device__ __forceinline__ double2 start_residual(const V* b, const V* Ax, int x0zero, long long i) {
  double2 r = LD(b, i);
  if (!x0zero) r = zsub(r, LD(Ax, i));
  return r;
}

// ============================================================== GPBICGsafe
// gpbicgsafe.f work columns: 1 r0, 2 p, 3 u, 4 r, 5 Au, 6 Ar, 7 Ap, 9 z,
// 10 y, 11 x (8 = t is recomputed).
template <typename V> struct SfV {
  V *x, *b, *r0, *p, *u, *r, *Au, *Ar, *Ap, *z, *y, *Ax;
  __device__ SfV shifted(long long o) const {
    return SfV{x + o, b + o, r0 + o, p + o, u + o, r + o, Au + o, Ar + o, Ap + o, z + o, y + o, Ax + o};
  }
};

template <typename V>
__global__ void __launch_bounds__(kThreads) sf_start(const KState* st, SfV<V> v, int x0zero, VN n,
                                                     double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  const double2 z0 = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = Vt<V>::rnd(start_residual(v.b, v.Ax, x0zero, i));
    ST(v.r0, i, r);
    ST(v.r, i, r);
    ST(v.p, i, z0); ST(v.u, i, z0); ST(v.z, i, z0); ST(v.y, i, z0); ST(v.Ap, i, z0); ST(v.Au, i, z0);
    acc[0] = zadd(acc[0], zmul(r, zconj(r)));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// R0RN, beta = 0, ITNO = 0 (gpbicgsafe.f:107-127)
__global__ void __launch_bounds__(kThreads) gp_fstart(KState* st, const double2* part, int nblk) {
  FK_ENTER(1);
  st->r0rn = st->dots[0];
  st->beta = make_double2(0.0, 0.0);
  st->it = 0;
}

// After Ar = A r: p = r + beta (p - u), Ap = Ar + beta (Ap - Au); partial
// (r0,Ap), (y,y), (Ar,r), (y,r), (Ar,y), (Ar,Ar) (gpbicgsafe.f:129-162).
template <typename V>
__global__ void __launch_bounds__(kThreads) sf_k2(KState* st, SfV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 be = st->beta;
  double2 acc[6];
  for (int q = 0; q < 6; ++q) acc[q] = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 6);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = LD(v.r, i), Ar = LD(v.Ar, i), y = LD(v.y, i);
    ST(v.p, i, zadd(r, zmul(be, zsub(LD(v.p, i), LD(v.u, i)))));
    const double2 Ap = Vt<V>::rnd(zadd(Ar, zmul(be, zsub(LD(v.Ap, i), LD(v.Au, i)))));
    ST(v.Ap, i, Ap);
    acc[0] = zadd(acc[0], zcdot(LD(v.r0, i), Ap));
    acc[1] = zadd(acc[1], zcdot(y, y));
    acc[2] = zadd(acc[2], zcdot(Ar, r));
    acc[3] = zadd(acc[3], zcdot(y, r));
    acc[4] = zadd(acc[4], zcdot(Ar, y));
    acc[5] = zadd(acc[5], zcdot(Ar, Ar));
  }
  block_partials<6>(acc, part, nblk, bx_);
}

// alpha = R0RN / (r0,Ap), then zeta and eta (gpbicgsafe.f:139-170);
// gpbicgplus.f orders the same products differently, see pl_f2.
__global__ void __launch_bounds__(kThreads) sf_f2(KState* st, const double2* part, int nblk) {
  FK_ENTER(6);
  const double2* d = st->dots;
  if (zzero(d[0])) {
    kbreak(st, 1);
    return;
  }
  st->alpha = zdiv(st->r0rn, d[0]);
  gp_zeta_eta(st, d[1], d[2], d[3], d[4], d[5]);
}

// u = zeta Ap + eta (y + beta u) (gpbicgsafe.f:172-176)
template <typename V>
__global__ void __launch_bounds__(kThreads) sf_k3(KState* st, SfV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 be = st->beta, et = st->eta, dz = st->dzeta;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_))
    ST(v.u, i, zadd(zmul(dz, LD(v.Ap, i)), zmul(et, zadd(LD(v.y, i), zmul(be, LD(v.u, i))))));
}

// After Au = A u: t = r - alpha Ap, z, y, x, r = t - y; partial |r|^2 and
// (r0,r) (gpbicgsafe.f:180-190, 230-236)
template <typename V>
__global__ void __launch_bounds__(kThreads) sf_k4(KState* st, SfV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 al = st->alpha, et = st->eta, dz = st->dzeta;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 2);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = LD(v.r, i), u = LD(v.u, i);
    const double2 t = Vt<V>::rnd(zsub(r, zmul(al, LD(v.Ap, i))));
    const double2 z = Vt<V>::rnd(zsub(zadd(zmul(dz, r), zmul(et, LD(v.z, i))), zmul(al, u)));
    const double2 y = Vt<V>::rnd(zsub(zadd(zmul(dz, LD(v.Ar, i)), zmul(et, LD(v.y, i))),
                                      zmul(al, LD(v.Au, i))));
    ST(v.z, i, z);
    ST(v.y, i, y);
    ST(v.x, i, zadd(zadd(LD(v.x, i), zmul(al, LD(v.p, i))), z));
    const double2 rn = Vt<V>::rnd(zsub(t, y));
    ST(v.r, i, rn);
    acc[0].x += rn.x * rn.x + rn.y * rn.y;
    acc[1] = zadd(acc[1], zcdot(LD(v.r0, i), rn));
  }
  block_partials<2>(acc, part, nblk, bx_);
}

// residual test, beta = alpha (r0,r) / zeta / R0RN (gpbicgsafe.f:192-219)
__global__ void __launch_bounds__(kThreads) gp_f4(KState* st, const double2* part, int nblk,
                                                  double2* trace) {
  FK_ENTER(2);
  gp_tail(st, trace, 4);
}

// ============================================================== GPBICGplus
// gpbicgplus.f work columns: 1 r0, 2 p, 3 r, 4 Ap, 6 Ar, 7 Au, 8 u, 9 y,
// 10 z, 11 x (5 = t is recomputed).
template <typename V> struct PlV {
  V *x, *b, *r0, *p, *r, *Ap, *Ar, *Au, *u, *y, *z, *Ax;
  __device__ PlV shifted(long long o) const {
    return PlV{x + o, b + o, r0 + o, p + o, r + o, Ap + o, Ar + o, Au + o, u + o, y + o, z + o, Ax + o};
  }
};

template <typename V>
__global__ void __launch_bounds__(kThreads) pl_start(const KState* st, PlV<V> v, int x0zero, VN n,
                                                     double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  const double2 z0 = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = Vt<V>::rnd(start_residual(v.b, v.Ax, x0zero, i));
    ST(v.r0, i, r);
    ST(v.r, i, r);
    ST(v.p, i, z0); ST(v.Ap, i, z0); ST(v.Au, i, z0); ST(v.u, i, z0); ST(v.z, i, z0); ST(v.y, i, z0);
    acc[0] = zadd(acc[0], zmul(r, zconj(r)));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// After Ar = A r: p = r + beta (p - u), Ap = Ar + beta (Ap - Au); partial
// (r0,Ap), (y,y), (Ar,Ar), (Ar,r), (y,r), (y,Ar) (gpbicgplus.f:108-140)
template <typename V>
__global__ void __launch_bounds__(kThreads) pl_k2(KState* st, PlV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 be = st->beta;
  double2 acc[6];
  for (int q = 0; q < 6; ++q) acc[q] = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 6);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = LD(v.r, i), Ar = LD(v.Ar, i), y = LD(v.y, i);
    ST(v.p, i, zadd(r, zmul(be, zsub(LD(v.p, i), LD(v.u, i)))));
    const double2 Ap = Vt<V>::rnd(zadd(Ar, zmul(be, zsub(LD(v.Ap, i), LD(v.Au, i)))));
    ST(v.Ap, i, Ap);
    acc[0] = zadd(acc[0], zcdot(LD(v.r0, i), Ap));
    acc[1] = zadd(acc[1], zcdot(y, y));
    acc[2] = zadd(acc[2], zcdot(Ar, Ar));
    acc[3] = zadd(acc[3], zcdot(Ar, r));
    acc[4] = zadd(acc[4], zcdot(y, r));
    acc[5] = zadd(acc[5], zcdot(y, Ar));
  }
  block_partials<6>(acc, part, nblk, bx_);
}

// gpbicgplus.f:118-163: c1 = (y,y), c2 = (Ar,Ar), c3 = (Ar,r), c4 = (y,r),
// c5 = (y,Ar); den = c1 c2 - c5 conj(c5)
__global__ void __launch_bounds__(kThreads) pl_f2(KState* st, const double2* part, int nblk) {
  FK_ENTER(6);
  const double2* d = st->dots;
  if (zzero(d[0])) {
    kbreak(st, 1);
    return;
  }
  st->alpha = zdiv(st->r0rn, d[0]);
  const double2 c1 = d[1], c2 = d[2], c3 = d[3], c4 = d[4], c5 = d[5];
  if (st->it == 0) {
    if (zzero(c2)) {
      kbreak(st, 2);
      return;
    }
    st->dzeta = zdiv(c3, c2);
    st->eta = make_double2(0.0, 0.0);
  } else {
    const double2 den = zsub(zmul(c1, c2), zmul(c5, zconj(c5)));
    if (zzero(den)) {
      kbreak(st, 3);
      return;
    }
    st->dzeta = zdiv(zsub(zmul(c1, c3), zmul(c4, zconj(c5))), den);
    st->eta = zdiv(zsub(zmul(c2, c4), zmul(c5, c3)), den);
  }
}

// u = zeta Ap + eta (y + beta u), z = zeta r + eta z - alpha u
// (gpbicgplus.f:165-171)
template <typename V>
__global__ void __launch_bounds__(kThreads) pl_k3(KState* st, PlV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 al = st->alpha, be = st->beta, et = st->eta, dz = st->dzeta;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 u = Vt<V>::rnd(zadd(zmul(dz, LD(v.Ap, i)), zmul(et, zadd(LD(v.y, i), zmul(be, LD(v.u, i))))));
    ST(v.u, i, u);
    ST(v.z, i, zsub(zadd(zmul(dz, LD(v.r, i)), zmul(et, LD(v.z, i))), zmul(al, u)));
  }
}

// After Au = A u: y, x, r = t - y with t = r - alpha Ap; partial |r|^2 and
// (r0,r) (gpbicgplus.f:175-183, 195-199)
template <typename V>
__global__ void __launch_bounds__(kThreads) pl_k4(KState* st, PlV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 al = st->alpha, et = st->eta, dz = st->dzeta;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 2);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 y = Vt<V>::rnd(zsub(zadd(zmul(dz, LD(v.Ar, i)), zmul(et, LD(v.y, i))),
                                      zmul(al, LD(v.Au, i))));
    ST(v.y, i, y);
    ST(v.x, i, zadd(zadd(LD(v.x, i), zmul(al, LD(v.p, i))), LD(v.z, i)));
    const double2 t = Vt<V>::rnd(zsub(LD(v.r, i), zmul(al, LD(v.Ap, i))));
    const double2 rn = Vt<V>::rnd(zsub(t, y));
    ST(v.r, i, rn);
    acc[0].x += rn.x * rn.x + rn.y * rn.y;
    acc[1] = zadd(acc[1], zcdot(LD(v.r0, i), rn));
  }
  block_partials<2>(acc, part, nblk, bx_);
}

// ================================================================ GPBICGAR
// gpbicgar.f work columns: 1 r0, 2 p, 3 r, 4 t, 5 z, 6 Ap, 7 Ar, 8 u, 9 y,
// 11 x, 12 Au.  A r is formed once before the loop and at the end of every
// iteration; Ap by the recurrence.
template <typename V> struct ArV {
  V *x, *b, *r0, *p, *r, *t, *z, *Ap, *Ar, *u, *y, *Au, *Ax;
  __device__ ArV shifted(long long o) const {
    return ArV{x + o, b + o, r0 + o, p + o, r + o, t + o, z + o, Ap + o, Ar + o, u + o, y + o,
               Au + o, Ax + o};
  }
};

template <typename V>
__global__ void __launch_bounds__(kThreads) ar_start(const KState* st, ArV<V> v, int x0zero, VN n,
                                                     double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  const double2 z0 = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = Vt<V>::rnd(start_residual(v.b, v.Ax, x0zero, i));
    ST(v.r0, i, r);
    ST(v.r, i, r);
    ST(v.p, i, z0); ST(v.t, i, z0); ST(v.z, i, z0); ST(v.Ap, i, z0); ST(v.u, i, z0); ST(v.y, i, z0);
    ST(v.Au, i, z0);
    acc[0] = zadd(acc[0], zmul(r, zconj(r)));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// p = r + beta (p - u), Ap = Ar + beta (Ap - Au); partial (r0,Ap), (y,y),
// (Ar,r), (y,r), (Ar,y), (Ar,Ar) (gpbicgar.f:95-137): the products of
// gpbicgsafe.f under other names, finished by sf_f2.
template <typename V>
__global__ void __launch_bounds__(kThreads) ar_k1(KState* st, ArV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 be = st->beta;
  double2 acc[6];
  for (int q = 0; q < 6; ++q) acc[q] = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 6);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = LD(v.r, i), Ar = LD(v.Ar, i), y = LD(v.y, i);
    ST(v.p, i, zadd(r, zmul(be, zsub(LD(v.p, i), LD(v.u, i)))));
    const double2 Ap = Vt<V>::rnd(zadd(Ar, zmul(be, zsub(LD(v.Ap, i), LD(v.Au, i)))));
    ST(v.Ap, i, Ap);
    acc[0] = zadd(acc[0], zcdot(LD(v.r0, i), Ap));
    acc[1] = zadd(acc[1], zcdot(y, y));
    acc[2] = zadd(acc[2], zcdot(Ar, r));
    acc[3] = zadd(acc[3], zcdot(y, r));
    acc[4] = zadd(acc[4], zcdot(Ar, y));
    acc[5] = zadd(acc[5], zcdot(Ar, Ar));
  }
  block_partials<6>(acc, part, nblk, bx_);
}

// u = zeta Ap + eta (t - r + beta u) (gpbicgar.f:156-160)
template <typename V>
__global__ void __launch_bounds__(kThreads) ar_k2(KState* st, ArV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 be = st->beta, et = st->eta, dz = st->dzeta;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_))
    ST(v.u, i, zadd(zmul(dz, LD(v.Ap, i)),
                    zmul(et, zadd(zsub(LD(v.t, i), LD(v.r, i)), zmul(be, LD(v.u, i))))));
}

// After Au = A u: t, z, y, x, r = t - y; partial |r|^2 and (r0,r)
// (gpbicgar.f:167-176, 194-197)
template <typename V>
__global__ void __launch_bounds__(kThreads) ar_k3(KState* st, ArV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 al = st->alpha, et = st->eta, dz = st->dzeta;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 2);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = LD(v.r, i), u = LD(v.u, i);
    const double2 t = Vt<V>::rnd(zsub(r, zmul(al, LD(v.Ap, i))));
    const double2 z = Vt<V>::rnd(zsub(zadd(zmul(dz, r), zmul(et, LD(v.z, i))), zmul(al, u)));
    const double2 y = Vt<V>::rnd(zsub(zadd(zmul(dz, LD(v.Ar, i)), zmul(et, LD(v.y, i))),
                                      zmul(al, LD(v.Au, i))));
    ST(v.t, i, t);
    ST(v.z, i, z);
    ST(v.y, i, y);
    ST(v.x, i, zadd(zadd(LD(v.x, i), zmul(al, LD(v.p, i))), z));
    const double2 rn = Vt<V>::rnd(zsub(t, y));
    ST(v.r, i, rn);
    acc[0].x += rn.x * rn.x + rn.y * rn.y;
    acc[1] = zadd(acc[1], zcdot(LD(v.r0, i), rn));
  }
  block_partials<2>(acc, part, nblk, bx_);
}

// residual test, then (after the reference's A r) beta (gpbicgar.f:178-213)
__global__ void __launch_bounds__(kThreads) ar_f3(KState* st, const double2* part, int nblk,
                                                  double2* trace) {
  FK_ENTER(2);
  gp_tail(st, trace, 5);
}

// ============================================================ BICGSTARPLUS
// bicgstarplus.f work columns: 1 r0* = conj(r0), 2 Ar, 3 r, 4 u, 5 s, 8 e,
// 9 g, 10 Au, 11 x, 12 h (6 and 7, formed from the others in one loop and
// used in the next, are recomputed).
template <typename V> struct BpV {
  V *x, *b, *r0s, *Ar, *r, *u, *s5, *e8, *g9, *Au, *h12, *Ax;
  __device__ BpV shifted(long long o) const {
    return BpV{x + o, b + o, r0s + o, Ar + o, r + o, u + o, s5 + o, e8 + o, g9 + o, Au + o, h12 + o,
               Ax + o};
  }
};

template <typename V>
__global__ void __launch_bounds__(kThreads) bp_start(const KState* st, BpV<V> v, int x0zero, VN n,
                                                     double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  const double2 z0 = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = Vt<V>::rnd(start_residual(v.b, v.Ax, x0zero, i));
    const double2 rs = zconj(r);
    ST(v.r, i, r);
    ST(v.r0s, i, rs);
    ST(v.e8, i, z0); ST(v.u, i, z0); ST(v.g9, i, z0); ST(v.s5, i, z0); ST(v.h12, i, z0);
    acc[0] = zadd(acc[0], zmul(rs, r));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// After Ar = A r: c1 = r0*.r, c2 = r0*.Ar, c3 = (s,s), c4 = (Ar,r),
// c5 = (Ar,s), c6 = (s,r), c7 = (Ar,Ar), c8 = r0*.h (bicgstarplus.f:110-132;
// the first iteration uses c1, c2, c4, c7, :86-108)
template <typename V>
__global__ void __launch_bounds__(kThreads) bp_k1(KState* st, BpV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  double2 acc[8];
  for (int q = 0; q < 8; ++q) acc[q] = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 8);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r0s = LD(v.r0s, i), r = LD(v.r, i), Ar = LD(v.Ar, i), s = LD(v.s5, i);
    acc[0] = zadd(acc[0], zmul(r0s, r));
    acc[1] = zadd(acc[1], zmul(r0s, Ar));
    acc[2] = zadd(acc[2], zcdot(s, s));
    acc[3] = zadd(acc[3], zcdot(Ar, r));
    acc[4] = zadd(acc[4], zcdot(Ar, s));
    acc[5] = zadd(acc[5], zcdot(s, r));
    acc[6] = zadd(acc[6], zcdot(Ar, Ar));
    acc[7] = zadd(acc[7], zmul(r0s, LD(v.h12, i)));
  }
  block_partials<8>(acc, part, nblk, bx_);
}

// alpha, beta, zeta, eta (bicgstarplus.f:86-160)
__global__ void __launch_bounds__(kThreads) bp_f1(KState* st, const double2* part, int nblk) {
  FK_ENTER(8);
  const double2 c1 = st->dots[0], c2 = st->dots[1], c3 = st->dots[2], c4 = st->dots[3],
                c5 = st->dots[4], c6 = st->dots[5], c7 = st->dots[6], c8 = st->dots[7];
  if (st->it == 0) {
    st->eta = st->beta = make_double2(0.0, 0.0);
    st->r0rn = c1;
    if (zzero(c7)) {
      kbreak(st, 2);
      return;
    }
    st->dzeta = zdiv(c4, c7);
    if (zzero(c2)) {
      kbreak(st, 3);
      return;
    }
    st->alpha = zdiv(c1, c2);
    return;
  }
  if (zzero(st->dzeta)) {
    kbreak(st, 4);
    return;
  }
  if (zzero(st->r0rn)) {
    kbreak(st, 5);
    return;
  }
  const double2 mal = make_double2(-st->alpha.x, -st->alpha.y);
  st->beta = zdiv(zmul(zdiv(mal, st->dzeta), c1), st->r0rn);
  const double2 den1 = zsub(c2, zmul(st->beta, c8));
  if (zzero(den1)) {
    kbreak(st, 6);
    return;
  }
  st->alpha = zdiv(c1, den1);
  const double2 den = zsub(zmul(c7, c3), zmul(c5, zconj(c5)));
  if (zzero(den)) {
    kbreak(st, 7);
    return;
  }
  st->dzeta = zdiv(zsub(zmul(c3, c4), zmul(c5, c6)), den);
  st->eta = zdiv(zsub(zmul(c7, c6), zmul(zconj(c5), c4)), den);
  st->r0rn = c1;
}

// the column 7 = Ar - beta h and u = zeta col7 + eta (s - beta u)
// (bicgstarplus.f:162-167)
template <typename V>
__device__ __forceinline__ double2 bp_col7(double2 Ar, double2 h, double2 be) {
  return Vt<V>::rnd(zsub(Ar, zmul(be, h)));
}
template <typename V>
__global__ void __launch_bounds__(kThreads) bp_k2(KState* st, BpV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 be = st->beta, et = st->eta, dz = st->dzeta;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 c7 = bp_col7<V>(LD(v.Ar, i), LD(v.h12, i), be);
    ST(v.u, i, zadd(zmul(dz, c7), zmul(et, zsub(LD(v.s5, i), zmul(be, LD(v.u, i))))));
  }
}

// After Au = A u (bicgstarplus.f:171-182): e = r - beta e - u, h = col7 - Au,
// g = col6 - alpha u, s = zeta Ar + eta s - alpha Au,
// x += col6 + alpha e, r = r - (zeta Ar + eta s) - alpha h; partial |r|^2.
template <typename V>
__global__ void __launch_bounds__(kThreads) bp_k3(KState* st, BpV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 al = st->alpha, be = st->beta, et = st->eta, dz = st->dzeta;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = LD(v.r, i), u = LD(v.u, i), Ar = LD(v.Ar, i), Au = LD(v.Au, i),
                  s = LD(v.s5, i);
    const double2 c7 = bp_col7<V>(Ar, LD(v.h12, i), be);
    const double2 c6 = Vt<V>::rnd(zadd(zmul(dz, r), zmul(et, LD(v.g9, i))));
    const double2 e = Vt<V>::rnd(zsub(zsub(r, zmul(be, LD(v.e8, i))), u));
    const double2 h = Vt<V>::rnd(zsub(c7, Au));
    ST(v.e8, i, e);
    ST(v.h12, i, h);
    ST(v.g9, i, zsub(c6, zmul(al, u)));
    const double2 w = Vt<V>::rnd(zadd(zmul(dz, Ar), zmul(et, s)));  // xi in the reference
    ST(v.s5, i, zsub(w, zmul(al, Au)));
    ST(v.x, i, zadd(zadd(LD(v.x, i), c6), zmul(al, e)));
    const double2 rn = Vt<V>::rnd(zsub(zsub(r, w), zmul(al, h)));
    ST(v.r, i, rn);
    acc[0].x += rn.x * rn.x + rn.y * rn.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// residual test and the iteration cap (bicgstarplus.f:184-198)
__global__ void __launch_bounds__(kThreads) bp_f3(KState* st, const double2* part, int nblk,
                                                  double2* trace) {
  FK_ENTER(1);
  const double res = sqrt(st->dots[0].x) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->alpha, st->beta, st->eta, st->dzeta, st->r0rn, res);
  if (res <= st->tol) {
    st->status = 1;
    st->done = 1;
    return;
  }
  if (st->it > st->maxit) {
    st->status = 1;
    st->steperr = 0;
    st->done = 1;
    return;
  }
  st->it++;
}

// ================================================================= GPBICOR
// gpbicor.f work columns: 1 r0, 2 p, 3 Ar, 4 Ap (q), 5 w, 6 At of the
// previous iteration, 8 u, 9 u~, 10 r, 11 z, 12 t of the previous
// iteration, x (FFloc).  The reference keeps y and At in its xi / xr
// arguments; here At lives in the matvec output (two buffers whose roles
// swap every iteration: column 6 is the other one) and y and t are
// recomputed where they are used.
template <typename V> struct CoV {
  V *x, *b, *r0, *p, *Ar, *Ap, *w, *Atp, *At, *tp, *u, *ut, *r, *z, *Ax;
  __device__ CoV shifted(long long o) const {
    return CoV{x + o, b + o, r0 + o, p + o, Ar + o, Ap + o, w + o, Atp + o, At + o, tp + o, u + o,
               ut + o, r + o, z + o, Ax + o};
  }
};

// r = r0 = b - A x0; the columns the first iteration reads as zero
// (gpbicor.f:127-140, 176-181)
template <typename V>
__global__ void __launch_bounds__(kThreads) co_start(const KState* st, CoV<V> v, int x0zero, VN n,
                                                     double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  const double2 z0 = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = Vt<V>::rnd(start_residual(v.b, v.Ax, x0zero, i));
    ST(v.r0, i, r);
    ST(v.r, i, r);
    ST(v.p, i, z0); ST(v.Ap, i, z0); ST(v.w, i, z0); ST(v.Atp, i, z0); ST(v.tp, i, z0);
    ST(v.u, i, z0); ST(v.ut, i, z0); ST(v.z, i, z0);
    acc[0].x += r.x * r.x + r.y * r.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) co_fstart(KState* st, const double2* part, int nblk) {
  FK_ENTER(1);
  st->beta = make_double2(0.0, 0.0);
  st->it = 0;
}

// After Ar = A r: partial rho = (r0,Ar) (gpbicor.f:160-170)
template <typename V>
__global__ void __launch_bounds__(kThreads) co_k1(KState* st, CoV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zcdot(LD(v.r0, i), LD(v.Ar, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

// rho, beta = rho / rhoa alpha / zeta (gpbicor.f:172-198)
__global__ void __launch_bounds__(kThreads) co_f1(KState* st, const double2* part, int nblk) {
  FK_ENTER(1);
  st->rho = st->dots[0];
  if (zzero(st->rho)) {
    kbreak(st, 1);
    return;
  }
  if (st->it == 0)
    st->beta = make_double2(0.0, 0.0);
  else
    st->beta = zdiv(zmul(zdiv(st->rho, st->rho0), st->alpha), st->dzeta);
}

// p, Ap (and w) (gpbicor.f:176-210): the first iteration starts p = r0,
// Ap = Ar; later w = At_prev + beta Ap, p = r + beta (p - u),
// Ap = Ar + beta (Ap - u~)
template <typename V>
__global__ void __launch_bounds__(kThreads) co_k2(KState* st, CoV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 be = st->beta;
  const bool first = st->it == 0;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    if (first) {
      ST(v.p, i, LD(v.r0, i));
      ST(v.Ap, i, LD(v.Ar, i));
    } else {
      const double2 Ap = LD(v.Ap, i);
      ST(v.w, i, zadd(LD(v.Atp, i), zmul(be, Ap)));
      ST(v.p, i, zadd(LD(v.r, i), zmul(be, zsub(LD(v.p, i), LD(v.u, i)))));
      ST(v.Ap, i, zadd(LD(v.Ar, i), zmul(be, zsub(Ap, LD(v.ut, i)))));
    }
  }
}

// After A Ap (into At): partial (r0, A Ap) (gpbicor.f:217-226)
template <typename V>
__global__ void __launch_bounds__(kThreads) co_k3(KState* st, CoV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zcdot(LD(v.r0, i), LD(v.At, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) co_f3(KState* st, const double2* part, int nblk) {
  FK_ENTER(1);
  if (zzero(st->dots[0])) {
    kbreak(st, 2);
    return;
  }
  st->alpha = zdiv(st->rho, st->dots[0]);
}

// y = t_prev - r + alpha (Ap - w) and t = r - alpha Ap (gpbicor.f:236-244)
template <typename V>
__device__ __forceinline__ void co_yt(const CoV<V>& v, long long i, double2 al, double2 Ap,
                                      double2 r, double2* y, double2* t) {
  *y = Vt<V>::rnd(zadd(zsub(LD(v.tp, i), r), zmul(al, zsub(Ap, LD(v.w, i)))));
  *t = Vt<V>::rnd(zsub(r, zmul(al, Ap)));
}

// At = Ar - alpha A Ap in place; partial (At,At), (y,t), (y,At), (At,t),
// (y,y) (gpbicor.f:236-290; the first iteration uses (At,t) and (At,At))
template <typename V>
__global__ void __launch_bounds__(kThreads) co_k4(KState* st, CoV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 al = st->alpha;
  double2 acc[5];
  for (int q = 0; q < 5; ++q) acc[q] = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 5);
  GRID_LOOP(i, n.len(rh_)) {
    double2 y, t;
    co_yt(v, i, al, LD(v.Ap, i), LD(v.r, i), &y, &t);
    const double2 At = Vt<V>::rnd(zsub(LD(v.Ar, i), zmul(al, LD(v.At, i))));
    ST(v.At, i, At);
    acc[0] = zadd(acc[0], zcdot(At, At));
    acc[1] = zadd(acc[1], zcdot(y, t));
    acc[2] = zadd(acc[2], zcdot(y, At));
    acc[3] = zadd(acc[3], zcdot(At, t));
    acc[4] = zadd(acc[4], zcdot(y, y));
  }
  block_partials<5>(acc, part, nblk, bx_);
}

// zeta, eta (gpbicor.f:250-299)
__global__ void __launch_bounds__(kThreads) co_f4(KState* st, const double2* part, int nblk) {
  FK_ENTER(5);
  const double2 c1 = st->dots[0], c2 = st->dots[1], c3 = st->dots[2], c4 = st->dots[3],
                c5 = st->dots[4];
  if (st->it == 0) {
    st->eta = make_double2(0.0, 0.0);
    if (zzero(c1)) {
      kbreak(st, 3);
      return;
    }
    st->dzeta = zdiv(c4, c1);
  } else {
    const double2 den = zsub(zmul(c1, c5), zmul(c3, zconj(c3)));
    if (zzero(den)) {
      kbreak(st, 4);
      return;
    }
    st->eta = zdiv(zsub(zmul(c1, c2), zmul(c3, c4)), den);
    st->dzeta = zdiv(zsub(zmul(c5, c4), zmul(c2, zconj(c3))), den);
  }
}

// u, u~, z, then r = t - eta y - zeta At, x += alpha p + z, t_prev = t
// (gpbicor.f:303-331); the At buffer becomes column 6 of the next iteration.
template <typename V>
__global__ void __launch_bounds__(kThreads) co_k5(KState* st, CoV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 al = st->alpha, be = st->beta, et = st->eta, dz = st->dzeta;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = LD(v.r, i), Ap = LD(v.Ap, i), Ar = LD(v.Ar, i), At = LD(v.At, i),
                  tp = LD(v.tp, i);
    double2 y, t;
    co_yt(v, i, al, Ap, r, &y, &t);
    const double2 u = Vt<V>::rnd(zadd(zmul(dz, Ap), zmul(et, zadd(zsub(tp, r), zmul(be, LD(v.u, i))))));
    ST(v.u, i, u);
    ST(v.ut, i, zadd(zdiv(zmul(dz, zsub(Ar, At)), al),
                     zmul(et, zadd(zsub(LD(v.Atp, i), Ar), zmul(be, LD(v.ut, i))))));
    const double2 z = Vt<V>::rnd(zsub(zadd(zmul(dz, r), zmul(et, LD(v.z, i))), zmul(al, u)));
    ST(v.z, i, z);
    const double2 rn = Vt<V>::rnd(zsub(zsub(t, zmul(et, y)), zmul(dz, At)));
    ST(v.r, i, rn);
    acc[0].x += rn.x * rn.x + rn.y * rn.y;
    ST(v.x, i, zadd(zadd(LD(v.x, i), zmul(al, LD(v.p, i))), z));
    ST(v.tp, i, t);
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// residual test, rhoa = rho, the iteration cap (gpbicor.f:333-354)
__global__ void __launch_bounds__(kThreads) co_f5(KState* st, const double2* part, int nblk,
                                                  double2* trace) {
  FK_ENTER(1);
  const double res = sqrt(st->dots[0].x) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->alpha, st->beta, st->eta, st->dzeta, st->rho, res);
  if (res <= st->tol) {
    st->status = 1;
    st->done = 1;
    return;
  }
  st->rho0 = st->rho;  // RHOA
  if (st->it + 1 > st->maxit) {  // ITNO = it + 1 here
    st->status = 1;
    st->steperr = 0;
    st->done = 1;
    return;
  }
  st->it++;
}

// ==================================================================== CORS
// cors.f work columns: 1 r0* = conj(A r0), 2 r, 3 Ar, 4 e, 5 q, 6 d, 7 h,
// 8 f, 9 A d, 10 x.
template <typename V> struct CsV {
  V *x, *b, *r0s, *r, *Ar, *e, *q, *d, *h, *f, *Ad, *Ax;
  __device__ CsV shifted(long long o) const {
    return CsV{x + o, b + o, r0s + o, r + o, Ar + o, e + o, q + o, d + o, h + o, f + o, Ad + o, Ax + o};
  }
};

template <typename V>
__global__ void __launch_bounds__(kThreads) cs_start(const KState* st, CsV<V> v, int x0zero, VN n,
                                                     double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = Vt<V>::rnd(start_residual(v.b, v.Ax, x0zero, i));
    ST(v.r, i, r);
    acc[0].x += r.x * r.x + r.y * r.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) cs_fstart(KState* st, const double2* part, int nblk) {
  FK_ENTER(1);
  st->it = 0;
}

// After Ar = A r: r0* = conj(Ar) in the first iteration; partial
// rho = r0*.Ar (cors.f:91-108)
template <typename V>
__global__ void __launch_bounds__(kThreads) cs_k1(KState* st, CsV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const bool first = st->it == 0;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 Ar = LD(v.Ar, i);
    double2 rs;
    if (first) {
      rs = zconj(Ar);
      ST(v.r0s, i, rs);
    } else {
      rs = LD(v.r0s, i);
    }
    acc[0] = zadd(acc[0], zmul(rs, Ar));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) cs_f1(KState* st, const double2* part, int nblk) {
  FK_ENTER(1);
  st->rho = st->dots[0];
  if (zzero(st->rho)) {
    kbreak(st, 1);
    return;
  }
  if (st->it > 0) st->beta = zdiv(st->rho, st->rho0);
}

// e, q, d (cors.f:116-138)
template <typename V>
__global__ void __launch_bounds__(kThreads) cs_k2(KState* st, CsV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 be = st->beta;
  const bool first = st->it == 0;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    if (first) {
      const double2 Ar = LD(v.Ar, i);
      ST(v.e, i, LD(v.r, i));
      ST(v.q, i, Ar);
      ST(v.d, i, Ar);
    } else {
      const double2 f = LD(v.f, i);
      ST(v.e, i, zadd(LD(v.r, i), zmul(be, LD(v.h, i))));
      const double2 q = Vt<V>::rnd(zadd(LD(v.Ar, i), zmul(be, f)));
      ST(v.q, i, q);
      ST(v.d, i, zadd(q, zmul(be, zadd(f, zmul(be, LD(v.d, i))))));
    }
  }
}

// After Ad = A d: partial r0*.Ad (cors.f:142-148)
template <typename V>
__global__ void __launch_bounds__(kThreads) cs_k3(KState* st, CsV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zmul(LD(v.r0s, i), LD(v.Ad, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) cs_f3(KState* st, const double2* part, int nblk) {
  FK_ENTER(1);
  if (zzero(st->dots[0])) {
    kbreak(st, 2);
    return;
  }
  st->alpha = zdiv(st->rho, st->dots[0]);
}

// h, f, x, r; partial |r|^2 (cors.f:155-166)
template <typename V>
__global__ void __launch_bounds__(kThreads) cs_k4(KState* st, CsV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 al = st->alpha;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 e = LD(v.e, i), d = LD(v.d, i), q = LD(v.q, i), Ad = LD(v.Ad, i);
    ST(v.h, i, zsub(e, zmul(al, d)));
    ST(v.f, i, zsub(q, zmul(al, Ad)));
    ST(v.x, i, zadd(LD(v.x, i), zmul(al, zsub(zscale(2.0, e), zmul(al, d)))));
    const double2 rn = Vt<V>::rnd(zsub(LD(v.r, i), zmul(al, zsub(zscale(2.0, q), zmul(al, Ad)))));
    ST(v.r, i, rn);
    acc[0].x += rn.x * rn.x + rn.y * rn.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// residual test, rhoa = rho, the iteration cap (cors.f:167-181)
__global__ void __launch_bounds__(kThreads) cs_f4(KState* st, const double2* part, int nblk,
                                                  double2* trace) {
  FK_ENTER(1);
  const double res = sqrt(st->dots[0].x) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->alpha, st->beta, st->rho, st->rho0, st->dots[0], res);
  if (res <= st->tol) {
    st->status = 1;
    st->done = 1;
    return;
  }
  st->rho0 = st->rho;  // RHOA
  if (st->it > st->maxit) {
    st->status = 1;
    st->steperr = 0;
    st->done = 1;
    return;
  }
  st->it++;
}

// ================================================================== QMRCLA
// qmrpim.f (PIMZQMR) work columns: 1 r, 2 p, 3 A p, 4 q, 5 d, 6 s, 7 v,
// 8 w, 9 A^T w; x is XS.  Products v.w are bilinear (no conjugate).
template <typename V> struct QcV {
  V *x, *b, *r, *p, *Ap, *q, *d, *s, *v, *w, *Atw, *zero, *Ax;
  __device__ QcV shifted(long long o) const {
    return QcV{x + o, b + o, r + o, p + o, Ap + o, q + o, d + o, s + o, v + o, w + o, Atw + o,
               zero + o, Ax + o};
  }
};

// r = b - A x0, v = w = r, p = q = d = s = 0 (and the zero vector the
// transposed matvec adds to); partial v.conj(v), w.conj(w), v.w
// (qmrpim.f:48-68)
template <typename V>
__global__ void __launch_bounds__(kThreads) qc_start(const KState* st, QcV<V> v, int x0zero, VN n,
                                                     double2* part, int nblk) {
  VK_ENTER();
  double2 acc[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  const double2 z0 = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 3);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = Vt<V>::rnd(start_residual(v.b, v.Ax, x0zero, i));
    ST(v.r, i, r);
    ST(v.v, i, r);
    ST(v.w, i, r);
    ST(v.p, i, z0); ST(v.q, i, z0); ST(v.d, i, z0); ST(v.s, i, z0); ST(v.zero, i, z0);
    acc[0] = zadd(acc[0], zmul(r, zconj(r)));
    acc[1] = zadd(acc[1], zmul(r, zconj(r)));
    acc[2] = zadd(acc[2], zmul(r, r));
  }
  block_partials<3>(acc, part, nblk, bx_);
}

// gamma = sqrt(v.conj(v)), ksi = sqrt(w.conj(w)), rho = v.w; the initial
// lambda, kappa, theta (qmrpim.f:33-36, 73-77)
__global__ void __launch_bounds__(kThreads) qc_fstart(KState* st, const double2* part, int nblk) {
  FK_ENTER(3);
  st->gam = zsqrt(st->dots[0]);
  st->ksi = zsqrt(st->dots[1]);
  st->rho = st->dots[2];
  st->lambda = make_double2(1.0, 0.0);
  st->kappa = st->theta = make_double2(-1.0, 0.0);
  st->mu = make_double2(0.0, 0.0);
  st->it = 0;
}

// partial v.(A^T w) (qmrpim.f:71-75, 139-143)
template <typename V>
__global__ void __launch_bounds__(kThreads) qc_kd4(KState* st, QcV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zmul(LD(v.v, i), LD(v.Atw, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

// The head of an iteration (qmrpim.f:92-104): ITNO + 1, gamma and ksi
// non-zero.
__device__ __forceinline__ void qc_head(KState* st) {
  st->it++;
  if (zabs(st->gam) == 0.0) {
    st->status = -3;
    st->steperr = 6;
    st->done = 1;
    return;
  }
  if (zabs(st->ksi) == 0.0) {
    st->status = -3;
    st->steperr = 7;
    st->done = 1;
  }
}

// epsilon = v.A^T w, tau = epsilon / rho (qmrpim.f:77-89), then the head of
// the first iteration
__global__ void __launch_bounds__(kThreads) qc_f0(KState* st, const double2* part, int nblk) {
  FK_ENTER(1);
  st->epsq = st->dots[0];
  if (zabs(st->rho) == 0.0) {
    st->status = -3;
    st->steperr = 5;
    st->done = 1;
    return;
  }
  st->tau = zdiv(st->epsq, st->rho);
  qc_head(st);
}

// p = v / gamma - mu p, q = (A^T w - gamma mu q) / ksi (qmrpim.f:105-112)
template <typename V>
__global__ void __launch_bounds__(kThreads) qc_k1(KState* st, QcV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 ga = st->gam, ks = st->ksi, mu = st->mu, gm = zmul(st->gam, st->mu);
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    ST(v.p, i, zsub(zdiv(LD(v.v, i), ga), zmul(mu, LD(v.p, i))));
    ST(v.q, i, zdiv(zsub(LD(v.Atw, i), zmul(gm, LD(v.q, i))), ks));
  }
}

// After A p: v = A p - tau / gamma v, w = q - tau / ksi w; partial
// v.conj(v), w.conj(w), v.w (qmrpim.f:116-133)
template <typename V>
__global__ void __launch_bounds__(kThreads) qc_k2(KState* st, QcV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 tg = zdiv(st->tau, st->gam), tk = zdiv(st->tau, st->ksi);
  double2 acc[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 3);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 vv = Vt<V>::rnd(zsub(LD(v.Ap, i), zmul(tg, LD(v.v, i))));
    const double2 ww = Vt<V>::rnd(zsub(LD(v.q, i), zmul(tk, LD(v.w, i))));
    ST(v.v, i, vv);
    ST(v.w, i, ww);
    acc[0] = zadd(acc[0], zmul(vv, zconj(vv)));
    acc[1] = zadd(acc[1], zmul(ww, zconj(ww)));
    acc[2] = zadd(acc[2], zmul(vv, ww));
  }
  block_partials<3>(acc, part, nblk, bx_);
}

// gamma, ksi, rho of the new v and w, keeping the previous ones
// (qmrpim.f:144-149)
__global__ void __launch_bounds__(kThreads) qc_f2(KState* st, const double2* part, int nblk) {
  FK_ENTER(3);
  st->omega = st->gam;  // GAMMA0
  st->gam = zsqrt(st->dots[0]);
  st->etat = st->ksi;  // KSI0
  st->ksi = zsqrt(st->dots[1]);
  st->rho0 = st->rho;
  st->rho = st->dots[2];
}

// epsilon, mu, tau, theta, kappa, lambda (qmrpim.f:150-183)
__global__ void __launch_bounds__(kThreads) qc_f3(KState* st, const double2* part, int nblk) {
  FK_ENTER(1);
  st->epsq = st->dots[0];
  const double2 gam0 = st->omega, ksi0 = st->etat;
  double2 den = zmul(zmul(st->gam, st->tau), st->rho0);
  if (zabs(den) == 0.0) {
    st->status = -3;
    st->steperr = 12;
    st->done = 1;
    return;
  }
  st->mu = zdiv(zmul(zmul(gam0, ksi0), st->rho), den);
  if (zabs(st->rho) == 0.0) {
    st->status = -3;
    st->steperr = 13;
    st->done = 1;
    return;
  }
  const double2 tau0 = st->tau;
  st->tau = zsub(zdiv(st->epsq, st->rho), zmul(st->gam, st->mu));
  const double at = zabs(tau0), ag = zabs(st->gam);
  const double2 abstau02 = make_double2(at * at, 0.0), absgam2 = make_double2(ag * ag, 0.0);
  den = zadd(zmul(st->lambda, abstau02), absgam2);
  if (zabs(den) == 0.0) {
    st->status = -3;
    st->steperr = 14;
    st->done = 1;
    return;
  }
  st->theta = zdiv(zmul(abstau02, zsub(make_double2(1.0, 0.0), st->lambda)), den);
  const double2 k = zdiv(zmul(zmul(gam0, zconj(tau0)), st->kappa), den);
  st->kappa = make_double2(-k.x, -k.y);
  st->lambda = zdiv(zmul(st->lambda, abstau02), den);
}

// d = theta d + kappa p, s = theta s + kappa A p, x += d, r -= s; partial
// r.conj(r) (qmrpim.f:184-193)
template <typename V>
__global__ void __launch_bounds__(kThreads) qc_k4(KState* st, QcV<V> v, VN n, double2* part, int nblk) {
  VK_ENTER();
  const double2 th = st->theta, ka = st->kappa;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 d = Vt<V>::rnd(zadd(zmul(th, LD(v.d, i)), zmul(ka, LD(v.p, i))));
    const double2 sv = Vt<V>::rnd(zadd(zmul(th, LD(v.s, i)), zmul(ka, LD(v.Ap, i))));
    ST(v.d, i, d);
    ST(v.s, i, sv);
    ST(v.x, i, zadd(LD(v.x, i), d));
    const double2 rn = Vt<V>::rnd(zsub(LD(v.r, i), sv));
    ST(v.r, i, rn);
    acc[0] = zadd(acc[0], zmul(rn, zconj(rn)));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// TOLE = sqrt|r.conj(r)| / |b| against tol; past MAXIT a failure
// (qmrpim.f:194-204); then the next iteration's head
__global__ void __launch_bounds__(kThreads) qc_f4(KState* st, const double2* part, int nblk,
                                                  double2* trace) {
  FK_ENTER(1);
  const double res = sqrt(zabs(st->dots[0])) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->gam, st->ksi, st->tau, st->theta, st->kappa, res);
  if (res <= st->tol) {
    st->status = 1;
    st->done = 1;
    return;
  }
  if (st->it > st->maxit) {
    st->status = -1;
    st->steperr = 0;
    st->done = 1;
    return;
  }
  qc_head(st);
}

#undef LD
#undef ST
#undef VK_ENTER
#undef FK_ENTER

// --------------------------------------------------------------- host loops
bool krylov_gp_method(int method) {
  return method == IFDDA_GPBICGSAFE || method == IFDDA_GPBICGPLUS || method == IFDDA_GPBICGAR ||
         method == IFDDA_BICGSTARPLUS || method == IFDDA_GPBICOR || method == IFDDA_CORS ||
         method == IFDDA_QMRCLA;
}

template <typename V>
bool krylov_gp(KLoop& L, V** vv, int method, KRun* out) {
  Solver* S = L.S;
  const VN vn = L.vn;
  const int nb = L.nb, R = L.R;
  const long long n = L.n;
  hipStream_t s = L.s;
  const dim3& G = *L.grid;  // shrinks as right-hand sides of a ragged layout finish
  const dim3 T(kThreads), F(1, R);
  KState* st = S->st;
  double2* part = S->part;
  const double2* cpart = part;
  auto mvv = [&](const V* in, V* out_) { return mv<V>(L.op, in, in, out_, s, L.done, R, n, L.dstride); };
  auto mvtv = [&](const V* in, const V* zero, V* out_) {
    return mvt<V>(L.op, in, zero, out_, s, L.done, R, n, L.dstride);
  };
  auto loop = [&](auto&& body, const char* what) -> bool {
    for (int it = 0;; ++it) {
      if (!body(it)) return false;
      if (!hip_ok(hipGetLastError(), what)) return false;
      const int stp = L.step_poll(it);
      if (stp < 0) return false;
      if (stp > 0) return true;
    }
  };
  // per-RHS outcome: nloop from ITNO, matvecs from where the run stopped
  auto finish = [&](auto&& nmv_of) {
    for (int r = 0; r < R; ++r) {
      const KState& h = S->st_host[r];
      KRun& o = out[r];
      o.status = h.status;
      o.steperr = h.steperr;
      if (L.active && !L.active[r]) continue;
      o.nloop = h.it + 1;
      o.nmv = L.nmv0 + nmv_of(h);
    }
  };
  V* x = vv[0];
  V* b = vv[1];
  V* Ax = vv[13];
  if (method == IFDDA_GPBICGSAFE) {
    SfV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], vv[8], vv[9], vv[10], Ax};
    hipLaunchKernelGGL(sf_start<V>, G, T, 0, s, (const KState*)st, v, L.x0zero, vn, part, nb);
    hipLaunchKernelGGL(gp_fstart, F, T, 0, s, st, cpart, nb);
    if (!loop([&](int) {
          if (!mvv(v.r, v.Ar)) return false;
          hipLaunchKernelGGL(sf_k2<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(sf_f2, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(sf_k3<V>, G, T, 0, s, st, v, vn, part, nb);
          if (!mvv(v.u, v.Au)) return false;
          hipLaunchKernelGGL(sf_k4<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(gp_f4, F, T, 0, s, st, cpart, nb, L.trace);
          return true;
        }, "gpbicgsafe launch"))
      return false;
    // A r, then A u; a breakdown in alpha, zeta or eta comes after A r alone
    finish([](const KState& h) {
      return 2 * h.it + (h.status == -1 && h.steperr >= 1 && h.steperr <= 3 ? 1 : 2);
    });
  } else if (method == IFDDA_GPBICGPLUS) {
    PlV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], vv[8], vv[9], vv[10], Ax};
    hipLaunchKernelGGL(pl_start<V>, G, T, 0, s, (const KState*)st, v, L.x0zero, vn, part, nb);
    hipLaunchKernelGGL(gp_fstart, F, T, 0, s, st, cpart, nb);
    if (!loop([&](int) {
          if (!mvv(v.r, v.Ar)) return false;
          hipLaunchKernelGGL(pl_k2<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(pl_f2, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(pl_k3<V>, G, T, 0, s, st, v, vn, part, nb);
          if (!mvv(v.u, v.Au)) return false;
          hipLaunchKernelGGL(pl_k4<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(gp_f4, F, T, 0, s, st, cpart, nb, L.trace);
          return true;
        }, "gpbicgplus launch"))
      return false;
    finish([](const KState& h) {
      return 2 * h.it + (h.status == -1 && h.steperr >= 1 && h.steperr <= 3 ? 1 : 2);
    });
  } else if (method == IFDDA_GPBICGAR) {
    ArV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], vv[8], vv[9], vv[10], vv[11], Ax};
    hipLaunchKernelGGL(ar_start<V>, G, T, 0, s, (const KState*)st, v, L.x0zero, vn, part, nb);
    hipLaunchKernelGGL(gp_fstart, F, T, 0, s, st, cpart, nb);
    if (!mvv(v.r, v.Ar)) return false;  // gpbicgar.f:83-93
    if (!loop([&](int) {
          hipLaunchKernelGGL(ar_k1<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(sf_f2, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(ar_k2<V>, G, T, 0, s, st, v, vn, part, nb);
          if (!mvv(v.u, v.Au)) return false;
          hipLaunchKernelGGL(ar_k3<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(ar_f3, F, T, 0, s, st, cpart, nb, L.trace);
          return mvv(v.r, v.Ar);
        }, "gpbicgar launch"))
      return false;
    // A r0, then A u and (unless converged) A r per iteration; STEPERR 5
    // and 6 come after that A r (gpbicgar.f:178-207)
    finish([](const KState& h) {
      const bool brk = h.status == -1;
      if (brk && h.steperr >= 1 && h.steperr <= 3) return 1 + 2 * h.it;
      if (brk) return 1 + 2 * h.it + 2;
      return 1 + 2 * h.it + 1;
    });
  } else if (method == IFDDA_BICGSTARPLUS) {
    BpV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], vv[8], vv[9], vv[10], Ax};
    hipLaunchKernelGGL(bp_start<V>, G, T, 0, s, (const KState*)st, v, L.x0zero, vn, part, nb);
    hipLaunchKernelGGL(gp_fstart, F, T, 0, s, st, cpart, nb);
    if (!loop([&](int) {
          if (!mvv(v.r, v.Ar)) return false;
          hipLaunchKernelGGL(bp_k1<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(bp_f1, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(bp_k2<V>, G, T, 0, s, st, v, vn, part, nb);
          if (!mvv(v.u, v.Au)) return false;
          hipLaunchKernelGGL(bp_k3<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(bp_f3, F, T, 0, s, st, cpart, nb, L.trace);
          return true;
        }, "bicgstarplus launch"))
      return false;
    finish([](const KState& h) { return 2 * h.it + (h.status == -1 ? 1 : 2); });
  } else if (method == IFDDA_GPBICOR) {
    V* const atb[2] = {vv[9], vv[10]};  // At / At of the previous iteration
    CoV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], atb[1], atb[0], vv[7], vv[8], vv[11], vv[12],
             Ax, Ax};
    // the Ax slot is read by co_start only, then holds z
    hipLaunchKernelGGL(co_start<V>, G, T, 0, s, (const KState*)st, v, L.x0zero, vn, part, nb);
    hipLaunchKernelGGL(co_fstart, F, T, 0, s, st, cpart, nb);
    if (!loop([&](int it) {
          v.At = atb[it & 1];
          v.Atp = atb[(it + 1) & 1];
          if (!mvv(v.r, v.Ar)) return false;
          hipLaunchKernelGGL(co_k1<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(co_f1, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(co_k2<V>, G, T, 0, s, st, v, vn, part, nb);
          if (!mvv(v.Ap, v.At)) return false;
          hipLaunchKernelGGL(co_k3<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(co_f3, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(co_k4<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(co_f4, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(co_k5<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(co_f5, F, T, 0, s, st, cpart, nb, L.trace);
          return true;
        }, "gpbicor launch"))
      return false;
    // ITNO counts from 1 and the driver adds nothing (inverserig.f:948-1013)
    finish([](const KState& h) { return 2 * h.it + (h.status == -1 && h.steperr == 1 ? 1 : 2); });
  } else if (method == IFDDA_CORS) {
    CsV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], vv[8], vv[9], vv[10], Ax};
    hipLaunchKernelGGL(cs_start<V>, G, T, 0, s, (const KState*)st, v, L.x0zero, vn, part, nb);
    hipLaunchKernelGGL(cs_fstart, F, T, 0, s, st, cpart, nb);
    if (!loop([&](int) {
          if (!mvv(v.r, v.Ar)) return false;
          hipLaunchKernelGGL(cs_k1<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(cs_f1, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(cs_k2<V>, G, T, 0, s, st, v, vn, part, nb);
          if (!mvv(v.d, v.Ad)) return false;
          hipLaunchKernelGGL(cs_k3<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(cs_f3, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(cs_k4<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(cs_f4, F, T, 0, s, st, cpart, nb, L.trace);
          return true;
        }, "cors launch"))
      return false;
    finish([](const KState& h) { return 2 * h.it + (h.status == -1 && h.steperr == 1 ? 1 : 2); });
  } else if (method == IFDDA_QMRCLA) {
    QcV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], vv[8], vv[9], vv[10], vv[11], Ax};
    hipLaunchKernelGGL(qc_start<V>, G, T, 0, s, (const KState*)st, v, L.x0zero, vn, part, nb);
    hipLaunchKernelGGL(qc_fstart, F, T, 0, s, st, cpart, nb);
    if (!mvtv(v.w, v.zero, v.Atw)) return false;  // inverserig.f:559-590
    hipLaunchKernelGGL(qc_kd4<V>, G, T, 0, s, st, v, vn, part, nb);
    hipLaunchKernelGGL(qc_f0, F, T, 0, s, st, cpart, nb);
    if (!loop([&](int) {
          hipLaunchKernelGGL(qc_k1<V>, G, T, 0, s, st, v, vn, part, nb);
          if (!mvv(v.p, v.Ap)) return false;
          hipLaunchKernelGGL(qc_k2<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(qc_f2, F, T, 0, s, st, cpart, nb);
          if (!mvtv(v.w, v.zero, v.Atw)) return false;
          hipLaunchKernelGGL(qc_kd4<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(qc_f3, F, T, 0, s, st, cpart, nb);
          hipLaunchKernelGGL(qc_k4<V>, G, T, 0, s, st, v, vn, part, nb);
          hipLaunchKernelGGL(qc_f4, F, T, 0, s, st, cpart, nb, L.trace);
          return true;
        }, "qmrcla launch"))
      return false;
    // NLOOP = ITNO (inverserig.f:503-606 adds nothing); A^T w0, then A p and
    // A^T w per iteration; breakdowns 6 and 7 are found at the head of
    // iteration ITNO, before its matvecs
    for (int r = 0; r < R; ++r) {
      const KState& h = S->st_host[r];
      KRun& o = out[r];
      o.status = h.status;
      o.steperr = h.steperr;
      if (L.active && !L.active[r]) continue;
      o.nloop = h.it;
      int k = h.it;
      if (h.status == -3 && (h.steperr == 6 || h.steperr == 7)) k = h.it - 1;
      if (h.status == -3 && h.steperr == 5) k = 0;
      o.nmv = L.nmv0 + 1 + 2 * k;
    }
  } else {
    set_error("unknown method");
    return false;
  }
  return true;
}

template bool krylov_gp<double2>(KLoop&, double2**, int, KRun*);
template bool krylov_gp<float2>(KLoop&, float2**, int, KRun*);

}  // namespace ifdda
