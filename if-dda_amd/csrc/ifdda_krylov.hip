// ifdda_krylov.hip -- device-resident Krylov solvers behind ifdda_solve():
//   GPBICG       gpbicg.f:81-338
//   QMRBICGSTAB  qmrbicgstab.f:1-253 (methode 1 and 2)
//   TFQMR        tfqmr.f:1-218
// and the inverserig.f driver around them (x0, true residual, FF = alpha.E).
//
// Every scalar of the recurrences lives on the device (KState) and is
// computed in FP64 by one-block "finish" kernels from per-block partial sums
// (fixed block count, fixed summation order: results do not depend on what
// else runs).  Vector kernels and matvec passes read KState.done and return at
// once when the solve has ended, so the host may queue whole iterations and
// poll the state only now and then.  Dots are sum conj(a) b in FP64 for both
// vector precisions.
//
// precision 32: iterative refinement.  The residual r = E_inc - A x is formed
// with the FP64 matvec; A e = r is solved with FP32 vectors and the FP32
// matvec to a relative tolerance; x += e; repeat until |r|/|E_inc| <= tol.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/ifdda.h"
#include "ifdda_internal.h"
#include "ifdda_krylov.h"

namespace ifdda {

// ======================================================= shared set-up kernels

// x = x0 (copy of src, or 0); partial |b|^2
template <typename V>
__global__ void __launch_bounds__(kThreads) k_init(const V* b, V* x, const V* src, VN n,
                                                  double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  part += (long long)rh_ * kMaxDots * nblk;
  b += n.off(rh_);
  x += n.off(rh_);
  if (src) src += n.off(rh_);
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 bi = LD(b, i);
    acc[0].x += bi.x * bi.x + bi.y * bi.y;
    ST(x, i, src ? LD(src, i) : make_double2(0.0, 0.0));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) f_init(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) st->norm = sqrt(st->dots[0].x);
}

// ============================================================ GPBICG kernels
// gpbicg.f:103-114 work columns.  t/told and Ap/Apold are two buffer pairs
// whose roles the host exchanges every iteration: this replaces the copy
// told = t (gpbicg.f:192-194) and the stored w = At + beta Ap
// (gpbicg.f:319-325), which is formed where it is used from At and the
// previous iteration's Ap.
template <typename V> struct GV {
  V *x, *b, *r, *r0s, *p, *Ap, *t, *At, *y, *u, *Apold, *z, *told, *Ax;
  __device__ GV shifted(long long o) const {
    return GV{x + o, b + o, r + o, r0s + o, p + o, Ap + o, t + o, At + o, y + o, u + o, Apold + o,
              z + o, told + o, Ax + o};
  }
};

// r = b - Ax (Ax = 0 when x0 = 0); r0* = conj(r); p = t = told = u = z = 0;
// partial sum r0* r  (gpbicg.f:133-148).  pinit: p = r instead, the first
// iteration's p = r + beta (p - u) with beta = 0, for the look-ahead loop,
// which forms p at the end of the iteration before.
template <typename V>
__global__ void __launch_bounds__(kThreads) g_start(GV<V> v, int x0zero, VN n,
                                                   double2* part, int nblk, int pinit = 0) {
  const int rh_ = n.rhs();  // right-hand side
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  const double2 z0 = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    double2 r = LD(v.b, i);
    if (!x0zero) r = zsub(r, LD(v.Ax, i));
    ST(v.r, i, r);
    ST(v.r0s, i, zconj(r));
    ST(v.p, i, pinit ? r : z0);
    ST(v.t, i, z0); ST(v.told, i, z0); ST(v.u, i, z0); ST(v.z, i, z0);
    ST(v.Ap, i, z0); ST(v.Apold, i, z0); ST(v.At, i, z0);
    const double2 rs = zconj(r);
    acc[0] = zadd(acc[0], zmul(rs, r));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) g_fstart(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    st->r0rn = st->dots[0];
    st->beta = make_double2(0.0, 0.0);
    st->it = 0;
  }
}

// p = r + beta (p - u) (gpbicg.f:161-169)
template <typename V>
__global__ void __launch_bounds__(kThreads) g_k1(const KState* st, GV<V> v, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (st->done) return;
  const double2 beta = st->beta;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) ST(v.p, i, zadd(LD(v.r, i), zmul(beta, zsub(LD(v.p, i), LD(v.u, i)))));
}

// partial sum r0* Ap (gpbicg.f:174-181)
template <typename V>
__global__ void __launch_bounds__(kThreads) g_k2(const KState* st, GV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zmul(LD(v.r0s, i), LD(v.Ap, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

// alpha = r0rn / <r0*, Ap>, or breakdown (gpbicg.f:183-187)
__device__ __forceinline__ void gp_alpha(KState* st) {
  const double2 c = st->dots[0];
  if (zzero(c)) {
    st->status = -1;
    st->steperr = 1;
    st->done = 1;
  } else {
    st->alpha = zdiv(st->r0rn, c);
  }
}

__global__ void __launch_bounds__(kThreads) g_f2(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) gp_alpha(st);
}

// g_f2 on the partials of a fused P5 pass (nblk per RHS)
__global__ void __launch_bounds__(kThreads) g_f2x(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;
  st += rh_;
  part += (long long)rh_ * nblk;
  if (st->done) return;
  finish_sum_any<1>(st, part, nblk);
  if (threadIdx.x == 0) gp_alpha(st);
}

// w = At + beta Apold (0 in the first iteration; gpbicg.f:319-325 of the
// previous iteration), y = told - r + alpha (Ap - w), t = r - alpha Ap
// (gpbicg.f:192-200); told holds the previous t.  gp2: y summed as
// gpbicg2.f:61 writes it, told - r - alpha w + alpha Ap.
template <typename V>
__global__ void __launch_bounds__(kThreads) g_k3(const KState* st, GV<V> v, VN n, int gp2) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (st->done) return;
  const double2 al = st->alpha, be = st->beta;
  const bool first = st->it == 0;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 t = LD(v.told, i), r = LD(v.r, i), Ap = LD(v.Ap, i);
    const double2 w = first ? make_double2(0.0, 0.0) : zadd(LD(v.At, i), zmul(be, LD(v.Apold, i)));
    ST(v.y, i, gp2 ? zadd(zsub(zsub(t, r), zmul(al, w)), zmul(al, Ap))
                   : zadd(zsub(t, r), zmul(al, zsub(Ap, w))));
    ST(v.t, i, zsub(r, zmul(al, Ap)));
  }
}

// <At,At>, <y,y>, <At,y>, <At,t>, <y,t> (gpbicg.f:222-253)
template <typename V>
__global__ void __launch_bounds__(kThreads) g_k4(const KState* st, GV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[5];
  for (int r = 0; r < 5; ++r) acc[r] = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 5);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 At = LD(v.At, i), y = LD(v.y, i), t = LD(v.t, i);
    const double2 cAt = zconj(At), cy = zconj(y);
    acc[0] = zadd(acc[0], zmul(cAt, At));
    acc[1] = zadd(acc[1], zmul(cy, y));
    acc[2] = zadd(acc[2], zmul(cAt, y));
    acc[3] = zadd(acc[3], zmul(cAt, t));
    acc[4] = zadd(acc[4], zmul(cy, t));
  }
  block_partials<5>(acc, part, nblk, bx_);
}

// zeta and eta from the five products, or breakdown (gpbicg.f:217-264);
// gp2: the steepest-descent step (eta = 0) on every even iteration
// (gpbicg2.f:75-88), not only the first
__device__ __forceinline__ void gp_zeta(KState* st, int gp2) {
  {
    const double2 c1 = st->dots[0], c2 = st->dots[1], c3 = st->dots[2], c4 = st->dots[3],
                  c5 = st->dots[4];
    if (gp2 ? st->it % 2 == 0 : st->it == 0) {  // gpbicg.f:217-235
      st->eta = make_double2(0.0, 0.0);
      if (zzero(c1)) {
        st->status = -1;
        st->steperr = 2;
        st->done = 1;
        return;
      }
      st->dzeta = zdiv(c4, c1);
    } else {  // gpbicg.f:255-264
      const double2 den = zsub(zmul(c1, c2), zmul(c3, zconj(c3)));
      if (zzero(den)) {
        st->status = -1;
        st->steperr = 3;
        st->done = 1;
        return;
      }
      st->dzeta = zdiv(zsub(zmul(c2, c4), zmul(c5, c3)), den);
      st->eta = zdiv(zsub(zmul(c1, c5), zmul(zconj(c3), c4)), den);
    }
  }
}

__global__ void __launch_bounds__(kThreads) g_f4(KState* st, const double2* part, int nblk, int gp2) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<5>(st, part, nblk);
  if (threadIdx.x == 0) gp_zeta(st, gp2);
}

// g_f4 on the partials of a fused P5 pass (5 x nblk per RHS)
__global__ void __launch_bounds__(kThreads) g_f4x(KState* st, const double2* part, int nblk, int gp2) {
  const int rh_ = blockIdx.y;
  st += rh_;
  part += (long long)rh_ * 5 * nblk;
  if (st->done) return;
  finish_sum_any<5>(st, part, nblk);
  if (threadIdx.x == 0) gp_zeta(st, gp2);
}

// u, z, x, r updates; partial |r|^2 and r0* r (gpbicg.f:270-281, :300-307)
template <typename V>
__global__ void __launch_bounds__(kThreads) g_k5(const KState* st, GV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  const double2 al = st->alpha, be = st->beta, et = st->eta, dz = st->dzeta;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 2);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 Ap = LD(v.Ap, i), r = LD(v.r, i), t = LD(v.t, i), y = LD(v.y, i),
                  At = LD(v.At, i);
    const double2 u = zadd(zmul(dz, Ap), zmul(et, zadd(zsub(LD(v.told, i), r), zmul(be, LD(v.u, i)))));
    const double2 z = zsub(zadd(zmul(dz, r), zmul(et, LD(v.z, i))), zmul(al, u));
    ST(v.u, i, u);
    ST(v.z, i, z);
    ST(v.x, i, zadd(zadd(LD(v.x, i), zmul(al, LD(v.p, i))), z));
    const double2 rn = zsub(zsub(t, zmul(et, y)), zmul(dz, At));
    ST(v.r, i, rn);
    const double2 rs = Vt<V>::rnd(rn);  // the stored value feeds the dots
    acc[0].x += rs.x * rs.x + rs.y * rs.y;
    acc[1] = zadd(acc[1], zmul(LD(v.r0s, i), rs));
  }
  block_partials<2>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) g_f5(KState* st, const double2* part, int nblk,
                                                double2* trace) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<2>(st, part, nblk);
  if (threadIdx.x != 0) return;
  const double res = sqrt(st->dots[0].x) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->alpha, st->beta, st->eta, st->dzeta, st->r0rn, res);
  if (res <= st->tol) {  // gpbicg.f:285-296
    st->status = 1;
    st->done = 1;
    return;
  }
  const double2 ct = st->dots[1];
  if (zzero(st->r0rn)) {  // gpbicg.f:308-312
    st->status = -1;
    st->steperr = 4;
    st->done = 1;
    return;
  }
  st->beta = zdiv(zdiv(zmul(st->alpha, ct), st->dzeta), st->r0rn);
  st->r0rn = ct;
  if (st->it > st->maxit) {  // gpbicg.f:327-330
    st->status = 1;
    st->steperr = 0;
    st->done = 1;
    return;
  }
  st->it++;
}

// ------------------------------------------------ GPBICG, look-ahead form
// The same recurrence with its last reduction moved ahead, so that the step
// that updates u, z, x and r also forms the next iteration's p and no kernel
// of its own does (g_k1):
//   r_new = t - eta y - zeta At, so <r0*, r_new> = <r0*, t> - eta <r0*, y>
//   - zeta <r0*, At>, whose three products the second matvec's gathering
//   pass sums with the five of gpbicg.f:222-253 (MODE 7);
//   beta_new = alpha <r0*, r_new> / (zeta <r0*, r>)   (gpbicg.f:314)
// is then known before the update, which writes p = r_new + beta_new
// (p - u_new) (gpbicg.f:161-169) over p once x has read the old p.  t is
// re-formed from r and Ap as g_k3 stored it, and r0* is not read.  Vector
// traffic per iteration falls from 29 to 25 streams of 3 Nd values.  The
// stopping test, the breakdown tests and their order are gpbicg.f's; only
// <r0*, r_new> is a sum of three products instead of one (equal in exact
// arithmetic, and in rounding to ~1e-16 of |r0*| |t| in FP64).

// zeta and eta (g_f4x), then <r0*, r_new> and the next beta
__global__ void __launch_bounds__(kThreads) g_f4xL(KState* st, const double2* part, int nblk,
                                                  int gp2) {
  const int rh_ = blockIdx.y;
  st += rh_;
  part += (long long)rh_ * 8 * nblk;
  if (st->done) return;
  finish_sum_any<8>(st, part, nblk);
  if (threadIdx.x != 0) return;
  gp_zeta(st, gp2);
  if (st->done) return;
  const double2* d = st->dots;
  const double2 ct = zsub(zsub(d[5], zmul(st->eta, d[6])), zmul(st->dzeta, d[7]));
  st->ct_la = ct;
  // gpbicg.f:308-316: with r0rn = 0 the iteration ends in a breakdown (the
  // check of g_f2xL)
  st->beta_la = zzero(st->r0rn) ? make_double2(0.0, 0.0)
                                : zdiv(zdiv(zmul(st->alpha, ct), st->dzeta), st->r0rn);
  st->lapend = 1;
}

// u, z, x, r updates and the next p; partial |r|^2 (gpbicg.f:270-281,
// :300-307, :161-169)
template <typename V>
__global__ void __launch_bounds__(kThreads) g_k5L(const KState* st, GV<V> v, VN n,
                                                 double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  const double2 al = st->alpha, be = st->beta, et = st->eta, dz = st->dzeta, bn = st->beta_la;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 Ap = LD(v.Ap, i), r = LD(v.r, i), y = LD(v.y, i), At = LD(v.At, i),
                  uo = LD(v.u, i), p = LD(v.p, i);
    const double2 t = Vt<V>::rnd(zsub(r, zmul(al, Ap)));  // as g_k3 stored it
    const double2 u = zadd(zmul(dz, Ap), zmul(et, zadd(zsub(LD(v.told, i), r), zmul(be, uo))));
    const double2 z = zsub(zadd(zmul(dz, r), zmul(et, LD(v.z, i))), zmul(al, u));
    ST(v.u, i, u);
    ST(v.z, i, z);
    ST(v.x, i, zadd(zadd(LD(v.x, i), zmul(al, p)), z));
    const double2 rn = zsub(zsub(t, zmul(et, y)), zmul(dz, At));
    ST(v.r, i, rn);
    // the next p from the stored r and u, as g_k1 would read them
    const double2 rs = Vt<V>::rnd(rn), us = Vt<V>::rnd(u);
    ST(v.p, i, zadd(rs, zmul(bn, zsub(p, us))));
    acc[0].x += rs.x * rs.x + rs.y * rs.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// The end of a look-ahead iteration (g_f5 with <r0*, r_new> and beta from
// g_f4xL), on the reduced |r|^2 in st->dots[0], by one thread.
__device__ __forceinline__ void gp_check_la(KState* st, double2* trace) {
  const double res = sqrt(st->dots[0].x) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->alpha, st->beta, st->eta, st->dzeta, st->r0rn, res);
  if (res <= st->tol) {  // gpbicg.f:285-296
    st->status = 1;
    st->done = 1;
    return;
  }
  if (zzero(st->r0rn)) {  // gpbicg.f:308-312
    st->status = -1;
    st->steperr = 4;
    st->done = 1;
    return;
  }
  st->beta = st->beta_la;
  st->r0rn = st->ct_la;
  if (st->it > st->maxit) {  // gpbicg.f:327-330
    st->status = 1;
    st->steperr = 0;
    st->done = 1;
    return;
  }
  st->it++;
}

// The previous iteration's end (gp_check_la on the |r|^2 partials of g_k5L,
// when g_f4xL left it due) and then alpha from the first gather's <r0*, Ap>
// partials (g_f2x): one finish kernel per iteration fewer than g_f5L +
// g_f2x, from iteration kLaSplit on (g_f5L ends the earlier ones).  The check
// of iteration k then runs after iteration k+1's first matvec, which is
// wasted once per solve that runs past kLaSplit; the iterations, matvec
// counts and results are those of the order g_f5L, g_f2x.
__global__ void __launch_bounds__(kThreads) g_f2xL(KState* st, const double2* part, int nblk,
                                                  const double2* part5, int nb5, double2* trace) {
  const int rh_ = blockIdx.y;
  st += rh_;
  if (st->done) return;
  if (st->lapend) {
    __shared__ int stop;
    finish_sum<1>(st, part + (long long)rh_ * kMaxDots * nblk, nblk);
    if (threadIdx.x == 0) {
      st->lapend = 0;
      gp_check_la(st, trace);
      stop = st->done;
    }
    __syncthreads();
    if (stop) return;
  }
  finish_sum_any<1>(st, part5 + (long long)rh_ * nb5, nb5);
  if (threadIdx.x == 0) gp_alpha(st);
}

// The same check as a kernel of its own, right after g_k5L: the first
// kLaSplit iterations of a look-ahead loop end this way, so a solve that
// converges within them runs no matvec past its last iteration (the merged
// check of g_f2xL wastes one per solve: 1 of ~15 on a 7-iteration solve, 1 of
// ~300 on a 146-iteration one).  Same sums, same order, same results.
__global__ void __launch_bounds__(kThreads) g_f5L(KState* st, const double2* part, int nblk,
                                                 double2* trace) {
  const int rh_ = blockIdx.y;
  st += rh_;
  if (st->done || !st->lapend) return;
  finish_sum<1>(st, part + (long long)rh_ * kMaxDots * nblk, nblk);
  if (threadIdx.x != 0) return;
  st->lapend = 0;
  gp_check_la(st, trace);
}
constexpr int kLaSplit = 16;

// ======================================================= QMRBICGSTAB kernels
template <typename V> struct QV {  // qmrbicgstab.f:23-31 work columns
  V *x, *b, *r, *p, *v, *d, *rt, *s, *xt, *dt, *t, *Ax;
  __device__ QV shifted(long long o) const {
    return QV{x + o, b + o, r + o, p + o, v + o, d + o, rt + o, s + o, xt + o, dt + o, t + o, Ax + o};
  }
};

// r = b - Ax; p = v = d = 0; rt = r; partial r.conj(r) (qmrbicgstab.f:50-62)
template <typename V>
__global__ void __launch_bounds__(kThreads) q_start(QV<V> v, int x0zero, VN n,
                                                   double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  const double2 z0 = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    double2 r = LD(v.b, i);
    if (!x0zero) r = zsub(r, LD(v.Ax, i));
    ST(v.r, i, r);
    ST(v.rt, i, r);
    ST(v.p, i, z0); ST(v.v, i, z0); ST(v.d, i, z0);
    acc[0] = zadd(acc[0], zmul(r, zconj(r)));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) q_fstart(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {  // qmrbicgstab.f:66-75
    st->rho = zone();
    st->alpha = zone();
    st->omega = zone();
    st->tau = zsqrt(st->dots[0]);
    st->theta = make_double2(0.0, 0.0);
    st->eta = make_double2(0.0, 0.0);
    st->it = 1;
  }
}

// partial rt* r (qmrbicgstab.f:79-87)
template <typename V>
__global__ void __launch_bounds__(kThreads) q_k1(const KState* st, QV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zmul(zconj(LD(v.rt, i)), LD(v.r, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) q_f1(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    const double2 rho0 = st->rho;
    st->rho = st->dots[0];
    st->beta = zdiv(zdiv(zmul(st->rho, st->alpha), rho0), st->omega);  // qmrbicgstab.f:88
  }
}

// p = r + beta (p - omega v) (qmrbicgstab.f:92-100)
template <typename V>
__global__ void __launch_bounds__(kThreads) q_k2(const KState* st, QV<V> v, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (st->done) return;
  const double2 be = st->beta, om = st->omega;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_))
  ST(v.p, i, zadd(LD(v.r, i), zmul(be, zsub(LD(v.p, i), zmul(om, LD(v.v, i))))));
}

// partial rt* v (qmrbicgstab.f:105-113)
template <typename V>
__global__ void __launch_bounds__(kThreads) q_k3(const KState* st, QV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zmul(zconj(LD(v.rt, i)), LD(v.v, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) q_f3(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    const double2 a = st->dots[0];
    if (zabs(a) == 0.0) {  // qmrbicgstab.f:114-118
      st->status = -1;
      st->steperr = 1;
      st->done = 1;
      return;
    }
    st->alpha = zdiv(st->rho, a);
  }
}

// s = r - alpha v; partial s* s (qmrbicgstab.f:123-131)
template <typename V>
__global__ void __launch_bounds__(kThreads) q_k4(const KState* st, QV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  const double2 al = st->alpha;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    ST(v.s, i, zsub(LD(v.r, i), zmul(al, LD(v.v, i))));
    const double2 s = LD(v.s, i);
    acc[0] = zadd(acc[0], zmul(zconj(s), s));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// thetat, c, taut, etat (qmrbicgstab.f:132-146)
__global__ void __launch_bounds__(kThreads) q_f4(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    if (zabs(st->tau) == 0.0) {
      st->status = -1;
      st->steperr = 2;
      st->done = 1;
      return;
    }
    st->thetat = zdiv(zsqrt(st->dots[0]), st->tau);
    const double2 c = zdiv(zone(), zsqrt(zadd(zone(), zmul(st->thetat, st->thetat))));
    st->taut = zmul(zmul(st->tau, st->thetat), c);
    st->etat = zmul(zmul(c, c), st->alpha);
  }
}

// dt = p + theta^2 eta/alpha d; xt = x + etat dt (qmrbicgstab.f:151-159)
template <typename V>
__global__ void __launch_bounds__(kThreads) q_k5(const KState* st, QV<V> v, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (st->done) return;
  const double2 f = zdiv(zmul(zmul(st->theta, st->theta), st->eta), st->alpha);
  const double2 et = st->etat;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 dt = zadd(LD(v.p, i), zmul(f, LD(v.d, i)));
    ST(v.dt, i, dt);
    ST(v.xt, i, zadd(LD(v.x, i), zmul(et, LD(v.dt, i))));
  }
}

// omega dots (qmrbicgstab.f:163-185)
template <typename V>
__global__ void __launch_bounds__(kThreads) q_k6(const KState* st, QV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  const bool m1 = st->methode == 1;
  RHS_BLOCKS(n, rh_, 2);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 t = LD(v.t, i), s = LD(v.s, i);
    if (m1) {
      acc[0] = zadd(acc[0], zmul(zconj(t), t));
      acc[1] = zadd(acc[1], zmul(zconj(s), t));
    } else {
      acc[0] = zadd(acc[0], zmul(zconj(s), t));
      acc[1] = zadd(acc[1], zmul(zconj(s), s));
    }
  }
  block_partials<2>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) q_f6(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<2>(st, part, nblk);
  if (threadIdx.x == 0) {
    const double2 om = st->dots[0];
    if (zabs(om) == 0.0) {  // qmrbicgstab.f:187-191
      st->status = -1;
      st->steperr = 3;
      st->done = 1;
      return;
    }
    st->omega = zdiv(st->dots[1], om);
  }
}

// r = s - omega t; partial r* r (qmrbicgstab.f:197-205)
template <typename V>
__global__ void __launch_bounds__(kThreads) q_k7(const KState* st, QV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  const double2 om = st->omega;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    ST(v.r, i, zsub(LD(v.s, i), zmul(om, LD(v.t, i))));
    const double2 r = LD(v.r, i);
    acc[0] = zadd(acc[0], zmul(zconj(r), r));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// theta, c, tau, eta (qmrbicgstab.f:206-229); the x update of this iteration
// still runs when the residual test below is met, so `done` is set by q_f8.
__global__ void __launch_bounds__(kThreads) q_f7(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    if (zabs(st->taut) == 0.0) {
      st->status = -1;
      st->steperr = 4;
      st->done = 1;
      return;
    }
    st->theta = zdiv(zsqrt(st->dots[0]), st->taut);
    const double2 c = zdiv(zone(), zsqrt(zadd(zone(), zmul(st->theta, st->theta))));
    st->tau = zmul(zmul(st->taut, st->theta), c);
    st->eta = zmul(zmul(c, c), st->omega);
    if (zabs(st->omega) == 0.0) {
      st->status = -1;
      st->steperr = 5;
      st->done = 1;
    }
  }
}

// d = s + thetat^2 etat/omega dt; x = xt + eta d (qmrbicgstab.f:231-238)
template <typename V>
__global__ void __launch_bounds__(kThreads) q_k8(const KState* st, QV<V> v, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (st->done) return;
  const double2 f = zdiv(zmul(zmul(st->thetat, st->thetat), st->etat), st->omega);
  const double2 et = st->eta;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    ST(v.d, i, zadd(LD(v.s, i), zmul(f, LD(v.dt, i))));
    ST(v.x, i, zadd(LD(v.xt, i), zmul(et, LD(v.d, i))));
  }
}

// QMR residual bound (qmrbicgstab.f:241-251)
__global__ void q_f8(KState* st, double2* trace) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  if (st->done) return;
  const double res = sqrt((double)(st->it + 1)) * zabs(st->tau) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->alpha, st->beta, st->omega, st->tau, st->eta, res);
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

// ============================================================== TFQMR kernels
template <typename V> struct TV {  // tfqmr.f:34-41 work columns
  V *x, *b, *r, *u, *w, *v, *d, *rt, *u1, *Au, *Auo, *Ax;
  __device__ TV shifted(long long o) const {
    return TV{x + o, b + o, r + o, u + o, w + o, v + o, d + o, rt + o, u1 + o, Au + o, Auo + o, Ax + o};
  }
};

// r = b - Ax; w = u = r (tfqmr.f:59-69)
template <typename V>
__global__ void __launch_bounds__(kThreads) t_start(TV<V> v, int x0zero, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  v = v.shifted(n.off(rh_));
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    double2 r = LD(v.b, i);
    if (!x0zero) r = zsub(r, LD(v.Ax, i));
    ST(v.r, i, r);
    ST(v.w, i, r);
    ST(v.u, i, r);
  }
}

// v = Au (given in v.v by the matvec); d = 0; Au = v; rt = r; partial rt* r
// (tfqmr.f:73-84)
template <typename V>
__global__ void __launch_bounds__(kThreads) t_start2(TV<V> v, VN n, double2* part,
                                                    int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    ST(v.d, i, make_double2(0.0, 0.0));
    ST(v.Au, i, LD(v.v, i));
    const double2 r = LD(v.r, i);
    ST(v.rt, i, r);
    acc[0] = zadd(acc[0], zmul(zconj(r), r));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) t_fstart(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {  // tfqmr.f:87-96
    st->rho = st->dots[0];
    st->tau = make_double2(st->norm, 0.0);
    st->theta = make_double2(0.0, 0.0);
    st->eta = make_double2(0.0, 0.0);
    st->alpha = make_double2(0.0, 0.0);
    st->beta = make_double2(0.0, 0.0);
    st->it = 0;
    // |b| = 0: tfqmr.f divides by NORM and never meets its test; report a
    // breakdown at once instead (the reference does not return).
    if (st->norm == 0.0 && !st->done) {
      st->status = -1;
      st->steperr = 1;
      st->done = 1;
    }
  }
}

// even step: partial rt* v (tfqmr.f:104-112)
template <typename V>
__global__ void __launch_bounds__(kThreads) t_k1(const KState* st, TV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zmul(zconj(LD(v.rt, i)), LD(v.v, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) t_f1(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) st->alpha = zdiv(st->rho, st->dots[0]);
}

// (even) u1 = u - alpha v; w = w - alpha Au; d = u + theta^2/alpha eta d;
// partial w* w (tfqmr.f:115-136)
template <typename V>
__global__ void __launch_bounds__(kThreads) t_k2(const KState* st, TV<V> v, int even, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  const double2 al = st->alpha;
  const double2 f = zmul(zdiv(zmul(st->theta, st->theta), al), st->eta);
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 u = LD(v.u, i);
    if (even) ST(v.u1, i, zsub(u, zmul(al, LD(v.v, i))));
    ST(v.w, i, zsub(LD(v.w, i), zmul(al, LD(v.Au, i))));
    ST(v.d, i, zadd(u, zmul(f, LD(v.d, i))));
    const double2 w = LD(v.w, i);
    acc[0] = zadd(acc[0], zmul(zconj(w), w));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// theta, c, tau, eta (tfqmr.f:139-144)
__global__ void __launch_bounds__(kThreads) t_f2(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    st->theta = zdiv(zsqrt(st->dots[0]), st->tau);
    const double2 c = zdiv(zone(), zsqrt(zadd(zone(), zmul(st->theta, st->theta))));
    st->tau = zmul(zmul(st->tau, st->theta), c);
    st->eta = zmul(zmul(c, c), st->alpha);
  }
}

// x = x + eta d (tfqmr.f:147-153)
template <typename V>
__global__ void __launch_bounds__(kThreads) t_k3(const KState* st, TV<V> v, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (st->done) return;
  const double2 et = st->eta;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) ST(v.x, i, zadd(LD(v.x, i), zmul(et, LD(v.d, i))));
}

// residual bound (tfqmr.f:156-160)
__global__ void t_f3(KState* st, double2* trace) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  if (st->done) return;
  const double res = sqrt((double)(st->it + 1)) * zabs(st->tau) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->alpha, st->beta, st->rho, st->tau, st->eta, res);
  if (res <= st->tol) {
    st->status = 1;
    st->done = 1;
  }
}

// (odd) partial rt* w (tfqmr.f:163-173)
template <typename V>
__global__ void __launch_bounds__(kThreads) t_k4(const KState* st, TV<V> v, VN n,
                                                double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zmul(zconj(LD(v.rt, i)), LD(v.w, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) t_f4(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    const double2 rho0 = st->rho;
    st->rho = st->dots[0];
    st->beta = zdiv(st->rho, rho0);
  }
}

// (odd) u1 = w + beta u; Auo = Au (tfqmr.f:176-183); then u = u1 (:187-194)
template <typename V>
__global__ void __launch_bounds__(kThreads) t_k5(const KState* st, TV<V> v, int odd, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (st->done) return;
  const double2 be = st->beta;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    if (odd) {
      ST(v.u, i, zadd(LD(v.w, i), zmul(be, LD(v.u, i))));
      ST(v.Auo, i, LD(v.Au, i));
    } else {
      ST(v.u, i, LD(v.u1, i));
    }
  }
}

// (odd) v = Au + beta (Auo + beta v) (tfqmr.f:206-214)
template <typename V>
__global__ void __launch_bounds__(kThreads) t_k6(const KState* st, TV<V> v, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (st->done) return;
  const double2 be = st->beta;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_))
  ST(v.v, i, zadd(LD(v.Au, i), zmul(be, zadd(LD(v.Auo, i), zmul(be, LD(v.v, i))))));
}

// end of a half step: cap check, ITNO + 1 (tfqmr.f:216)
__global__ void t_f5(KState* st) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  if (st->done) return;
  if (st->it > st->maxit) {
    st->status = 1;
    st->steperr = 0;
    st->done = 1;
    return;
  }
  st->it++;
}

// ===================================================== driver / refinement
// res = Ax - b, partial |res|^2 ; FF = alpha x (inverserig.f:95-105)
// ================================================================ CG kernels
// zcg.f: conjugate gradients for the complex-symmetric A = I - T alpha with
// bilinear products (r.r, p.w, r.s are not conjugated); the stopping test
// uses |r| / |b|.  Work columns: x (WRK1), r (WRK2), p (WRK3), s (WRK4),
// w (WRK5).
template <typename V> struct CV {
  V *x, *b, *r, *p, *w, *s, *Ax;
  __device__ CV shifted(long long o) const { return CV{x + o, b + o, r + o, p + o, w + o, s + o, Ax + o}; }
};

// r = b - Ax, p = r; partial |b|^2, r.r (zcg.f:53-74)
template <typename V>
__global__ void __launch_bounds__(kThreads) c_start(CV<V> v, int x0zero, VN n, double2* part,
                                                   int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 2);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 b = LD(v.b, i);
    double2 r = b;
    if (!x0zero) r = zsub(r, LD(v.Ax, i));
    ST(v.r, i, r);
    ST(v.p, i, r);
    const double2 rs = Vt<V>::rnd(r);
    acc[0].x += b.x * b.x + b.y * b.y;
    acc[1] = zadd(acc[1], zmul(rs, rs));
  }
  block_partials<2>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) c_fstart(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  finish_sum<2>(st, part, nblk);
  if (threadIdx.x == 0) {
    st->norm = sqrt(st->dots[0].x);
    st->rho0 = st->dots[1];
    st->it = 0;
  }
}

// partial p.w (zcg.f:78-84)
template <typename V>
__global__ void __launch_bounds__(kThreads) c_k1(const KState* st, CV<V> v, VN n, double2* part,
                                                int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zmul(LD(v.p, i), LD(v.w, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

// xi = p.w (zcg.f:85-88; the reference stops the program on xi = 0), then
// the first loop head: nloop = 1, the iteration cap, alpha = rho0 / xi
// (zcg.f:92-107)
__global__ void __launch_bounds__(kThreads) c_f1(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    st->xxi = st->dots[0];
    if (zzero(st->xxi)) {
      st->status = -1;
      st->steperr = 1;
      st->done = 1;
      return;
    }
    st->it = 1;
    if (st->it >= st->maxit) {  // NSTAT = -4
      st->status = -1;
      st->steperr = 0;
      st->done = 1;
      return;
    }
    st->alpha = zdiv(st->rho0, st->xxi);
  }
}

// (after the first iteration) p = r + beta p, w = s + beta w (zcg.f:159-165);
// x += alpha p, r -= alpha w; partial |r|^2 (zcg.f:110-120)
template <typename V>
__global__ void __launch_bounds__(kThreads) c_k2(const KState* st, CV<V> v, VN n, double2* part,
                                                int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  const double2 al = st->alpha, be = st->beta;
  const bool first = st->it == 1;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    double2 p = LD(v.p, i), w = LD(v.w, i);
    if (!first) {
      p = Vt<V>::rnd(zadd(LD(v.r, i), zmul(be, p)));
      w = Vt<V>::rnd(zadd(LD(v.s, i), zmul(be, w)));
      ST(v.p, i, p);
      ST(v.w, i, w);
    }
    ST(v.x, i, zadd(LD(v.x, i), zmul(al, p)));
    const double2 r = zsub(LD(v.r, i), zmul(al, w));
    ST(v.r, i, r);
    const double2 rs = Vt<V>::rnd(r);
    acc[0].x += rs.x * rs.x + rs.y * rs.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// tole = |r| / |b|; converged when tole <= tol (zcg.f:121-138)
__global__ void __launch_bounds__(kThreads) c_f2(KState* st, const double2* part, int nblk,
                                                double2* trace) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x != 0) return;
  const double res = sqrt(fabs(st->dots[0].x)) / st->norm;
  st->residu = res;
  push_trace(st, trace, st->alpha, st->beta, st->xxi, st->rho0, make_double2(0.0, 0.0), res);
  if (res <= st->tol) {
    st->status = 1;
    st->done = 1;
  }
}

// partial r.r, r.s (zcg.f:150-156)
template <typename V>
__global__ void __launch_bounds__(kThreads) c_k3(const KState* st, CV<V> v, VN n, double2* part,
                                                int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 2);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 r = LD(v.r, i);
    acc[0] = zadd(acc[0], zmul(r, r));
    acc[1] = zadd(acc[1], zmul(r, LD(v.s, i)));
  }
  block_partials<2>(acc, part, nblk, bx_);
}

// beta = rho / rho0, xi = delta - beta^2 xi (zcg.f:158-169), then the loop
// head: nloop += 1, the iteration cap, alpha = rho0 / xi (zcg.f:92-107)
__global__ void __launch_bounds__(kThreads) c_f3(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<2>(st, part, nblk);
  if (threadIdx.x == 0) {
    const double2 rho = st->dots[0], delta = st->dots[1];
    const double2 be = zdiv(rho, st->rho0);
    st->beta = be;
    st->rho0 = rho;
    st->xxi = zsub(delta, zmul(zmul(be, be), st->xxi));
    st->it++;
    if (st->it >= st->maxit) {  // NSTAT = -4
      st->status = -1;
      st->steperr = 0;
      st->done = 1;
      return;
    }
    if (zzero(st->xxi)) {  // (no test in the reference: a division by zero there)
      st->status = -1;
      st->steperr = 1;
      st->done = 1;
      return;
    }
    st->alpha = zdiv(st->rho0, st->xxi);
  }
}

// ========================================================== BICGSTAB kernels
// pimzbicgstab.f, as the reference runs it: rho = sum conj(r) r of the
// current residual (pimzbicgstab.f:71-74; the shadow residual only enters
// xi), stopping test |r|^2 / |r0|^2 <= tol / 10 (pimzbicgstab.f:168-181).
// Work columns: x, r (1), rt (2), p (3), s (5), t (6), v (7).
template <typename V> struct BV {
  V *x, *b, *r, *rt, *p, *v, *s, *t, *Ax;
  __device__ BV shifted(long long o) const {
    return BV{x + o, b + o, r + o, rt + o, p + o, v + o, s + o, t + o, Ax + o};
  }
};

// r = b - Ax, rt = r, p = v = 0; partial |r|^2 (pimzbicgstab.f:48-58)
template <typename V>
__global__ void __launch_bounds__(kThreads) s_start(BV<V> v, int x0zero, VN n, double2* part,
                                                   int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  const double2 z0 = make_double2(0.0, 0.0);
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    double2 r = LD(v.b, i);
    if (!x0zero) r = zsub(r, LD(v.Ax, i));
    ST(v.r, i, r);
    ST(v.rt, i, r);
    ST(v.p, i, z0);
    ST(v.v, i, z0);
    const double2 rs = Vt<V>::rnd(r);
    acc[0].x += rs.x * rs.x + rs.y * rs.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// The head of iteration it + 1 (pimzbicgstab.f:69-89): rho0 = rho, rho =
// conj(r).r (the latest residual's, already summed), kappa = rho0 omega,
// beta = rho alpha / kappa.
__device__ __forceinline__ void s_head(KState* st) {
  st->it++;
  const double2 rho0 = st->rho;
  st->rho0 = rho0;
  st->rho = make_double2(st->rr, 0.0);
  const double2 kappa = zmul(rho0, st->omega);
  if (zabs(kappa) == 0.0) {
    st->status = -1;
    st->steperr = 6;
    st->done = 1;
    return;
  }
  st->beta = zdiv(zmul(st->rho, st->alpha), kappa);
}

__global__ void __launch_bounds__(kThreads) s_fstart(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    st->norm = st->dots[0].x;  // the squared norm, as the reference keeps NORM
    st->rr = st->dots[0].x;
    st->rho = st->alpha = st->omega = make_double2(1.0, 0.0);
    st->it = 0;
    s_head(st);
  }
}

// p = r + beta (p - omega v) (pimzbicgstab.f:91-100)
template <typename V>
__global__ void __launch_bounds__(kThreads) s_k1(const KState* st, BV<V> v, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (st->done) return;
  const double2 be = st->beta, om = st->omega;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 q = Vt<V>::rnd(zsub(LD(v.p, i), zmul(om, LD(v.v, i))));
    ST(v.p, i, zadd(LD(v.r, i), zmul(be, q)));
  }
}

// partial conj(rt).v (pimzbicgstab.f:104-108)
template <typename V>
__global__ void __launch_bounds__(kThreads) s_k2(const KState* st, BV<V> v, VN n, double2* part,
                                                int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) acc[0] = zadd(acc[0], zmul(zconj(LD(v.rt, i)), LD(v.v, i)));
  block_partials<1>(acc, part, nblk, bx_);
}

// alpha = rho / xi (pimzbicgstab.f:111-117)
__global__ void __launch_bounds__(kThreads) s_f2(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0) {
    st->xxi = st->dots[0];
    if (zabs(st->xxi) == 0.0) {
      st->status = -1;
      st->steperr = 10;
      st->done = 1;
      return;
    }
    st->alpha = zdiv(st->rho, st->xxi);
  }
}

// s = r - alpha v; partial conj(s).s (pimzbicgstab.f:118-124)
template <typename V>
__global__ void __launch_bounds__(kThreads) s_k3(const KState* st, BV<V> v, VN n, double2* part,
                                                int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  const double2 al = st->alpha;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 sv = zsub(LD(v.r, i), zmul(al, LD(v.v, i)));
    ST(v.s, i, sv);
    const double2 ss = Vt<V>::rnd(sv);
    acc[0] = zadd(acc[0], zmul(zconj(ss), ss));
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// |s| < 1e-15: soft breakdown (pimzbicgstab.f:125-134)
__global__ void __launch_bounds__(kThreads) s_f3(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x == 0 && zabs(zsqrt(st->dots[0])) < 1.0e-15) {
    st->status = -1;
    st->steperr = 12;
    st->done = 1;
  }
}

// partial conj(t).t, conj(t).s (pimzbicgstab.f:144-148)
template <typename V>
__global__ void __launch_bounds__(kThreads) s_k4(const KState* st, BV<V> v, VN n, double2* part,
                                                int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 2);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 ct = zconj(LD(v.t, i));
    acc[0] = zadd(acc[0], zmul(ct, LD(v.t, i)));
    acc[1] = zadd(acc[1], zmul(ct, LD(v.s, i)));
  }
  block_partials<2>(acc, part, nblk, bx_);
}

// omega = conj(t).s / conj(t).t (pimzbicgstab.f:150-156)
__global__ void __launch_bounds__(kThreads) s_f4(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<2>(st, part, nblk);
  if (threadIdx.x == 0) {
    if (zabs(st->dots[0]) == 0.0) {
      st->status = -1;
      st->steperr = 14;
      st->done = 1;
      return;
    }
    st->omega = zdiv(st->dots[1], st->dots[0]);
  }
}

// x += alpha p + omega s, r = s - omega t; partial conj(r).r
// (pimzbicgstab.f:158-166)
template <typename V>
__global__ void __launch_bounds__(kThreads) s_k5(const KState* st, BV<V> v, VN n, double2* part,
                                                int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  const double2 al = st->alpha, om = st->omega;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 sv = LD(v.s, i);
    ST(v.x, i, zadd(zadd(LD(v.x, i), zmul(al, LD(v.p, i))), zmul(om, sv)));
    const double2 r = zsub(sv, zmul(om, LD(v.t, i)));
    ST(v.r, i, r);
    const double2 rs = Vt<V>::rnd(r);
    acc[0].x += rs.x * rs.x + rs.y * rs.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// |r|^2 / |r0|^2 against tol / 10, the iteration cap (pimzbicgstab.f:167-181),
// then the next iteration's head
__global__ void __launch_bounds__(kThreads) s_f5(KState* st, const double2* part, int nblk,
                                                double2* trace) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  if (st->done) return;
  finish_sum<1>(st, part, nblk);
  if (threadIdx.x != 0) return;
  st->rr = st->dots[0].x;
  const double c1 = st->rr / st->norm;
  st->residu = c1;
  push_trace(st, trace, st->alpha, st->beta, st->omega, st->rho, st->xxi, c1);
  if (st->it > st->maxit) {  // STATUS = -1 (the reference then reports ITNO = MAXIT)
    st->status = -1;
    st->steperr = 0;
    st->done = 1;
    return;
  }
  if (fabs(c1) <= st->tol / 10.0) {
    st->status = 1;
    st->done = 1;
    return;
  }
  s_head(st);
}

// x = p for the right-hand sides that ended in a soft breakdown
// (pimzbicgstab.f:126-133)
template <typename V>
__global__ void __launch_bounds__(kThreads) s_soft(const KState* st, BV<V> v, VN n) {
  const int rh_ = n.rhs();  // right-hand side
  st += rh_;
  v = v.shifted(n.off(rh_));
  if (!(st->status == -1 && st->steperr == 12)) return;
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) ST(v.x, i, LD(v.p, i));
}

template <typename V>
__global__ void __launch_bounds__(kThreads) k_resid(const V* Ax, const V* b, VN n,
                                                   double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  part += (long long)rh_ * kMaxDots * nblk;
  Ax += n.off(rh_);
  b += n.off(rh_);
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    const double2 e = zsub(LD(Ax, i), LD(b, i));
    acc[0].x += e.x * e.x + e.y * e.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

__global__ void __launch_bounds__(kThreads) f_resid(KState* st, const double2* part, int nblk) {
  const int rh_ = blockIdx.y;  // right-hand side
  st += rh_;
  part += (long long)rh_ * kMaxDots * nblk;
  finish_sum<1>(st, part, nblk);
}

// FF = alpha . E per dipole (inverserig.f:70-82 applied to the solution)
__global__ void __launch_bounds__(kThreads) k_alpha_apply(const double2* x, double2* out,
                                                         const double2* al, int mode,
                                                         long long nd) {
  const int rh_ = blockIdx.y;  // right-hand side
  x += rh_ * 3 * nd;
  out += rh_ * 3 * nd;
  const int nbr_ = gridDim.x, bx_ = blockIdx.x;
  GRID_LOOP(i, nd) {
    double2 e0 = x[3 * i], e1 = x[3 * i + 1], e2 = x[3 * i + 2];
    if (mode == kAlphaScalar) {
      const double2 a = al[i];
      e0 = zmul(a, e0);
      e1 = zmul(a, e1);
      e2 = zmul(a, e2);
    } else if (mode == kAlphaTensor) {
      const double2* m = al + 9 * i;
      const double2 f0 = zadd(zadd(zmul(m[0], e0), zmul(m[1], e1)), zmul(m[2], e2));
      const double2 f1 = zadd(zadd(zmul(m[3], e0), zmul(m[4], e1)), zmul(m[5], e2));
      const double2 f2 = zadd(zadd(zmul(m[6], e0), zmul(m[7], e1)), zmul(m[8], e2));
      e0 = f0;
      e1 = f1;
      e2 = f2;
    }
    out[3 * i] = e0;
    out[3 * i + 1] = e1;
    out[3 * i + 2] = e2;
  }
}

// y = alpha^T y + x per dipole, y holding -T x on entry: the transposed
// operator A^T x = x - alpha^T T x as inverserig.f:576-589 forms it (the
// transpose of the 3x3 alpha, summed over its first index in order).
// Right-hand sides whose skip flag is set are left alone.
template <typename V>
__global__ void __launch_bounds__(kThreads) k_alpha_t(const V* x, V* y, const V* al, int mode,
                                                     long long nd, const int* skip,
                                                     int skip_stride) {
  const int rh_ = blockIdx.y;  // right-hand side
  if (skip && skip[(long long)rh_ * skip_stride]) return;
  x += rh_ * 3 * nd;
  y += rh_ * 3 * nd;
  const int nbr_ = gridDim.x, bx_ = blockIdx.x;
  GRID_LOOP(i, nd) {
    double2 e[3], f[3];
    for (int a = 0; a < 3; ++a) e[a] = Vt<V>::ld(y, 3 * i + a);
    if (mode == kAlphaScalar) {
      const double2 s = Vt<V>::ld(al, i);
      for (int a = 0; a < 3; ++a) f[a] = zmul(s, e[a]);
    } else {
      const V* m = al + 9 * i;
      for (int a = 0; a < 3; ++a)
        f[a] = zadd(zadd(zmul(Vt<V>::ld(m, a), e[0]), zmul(Vt<V>::ld(m, 3 + a), e[1])),
                    zmul(Vt<V>::ld(m, 6 + a), e[2]));
    }
    for (int a = 0; a < 3; ++a) Vt<V>::st(y, 3 * i + a, zadd(f[a], Vt<V>::ld(x, 3 * i + a)));
  }
}

// r = b - Ax (FP64), cast to FP32 for the inner solve; partial |r|^2
__global__ void __launch_bounds__(kThreads) k_refine_resid(const double2* b, const double2* Ax,
                                                          int have_ax, double2* r64, float2* r32,
                                                          VN n, double2* part, int nblk) {
  const int rh_ = n.rhs();  // right-hand side
  part += (long long)rh_ * kMaxDots * nblk;
  b += n.off(rh_);
  Ax += n.off(rh_);
  r64 += n.off(rh_);
  r32 += n.off(rh_);
  double2 acc[1] = {make_double2(0.0, 0.0)};
  RHS_BLOCKS(n, rh_, 1);
  GRID_LOOP(i, n.len(rh_)) {
    double2 r = b[i];
    if (have_ax) r = zsub(r, Ax[i]);
    r64[i] = r;
    r32[i] = make_float2((float)r.x, (float)r.y);
    acc[0].x += r.x * r.x + r.y * r.y;
  }
  block_partials<1>(acc, part, nblk, bx_);
}

// x += e (FP32 correction into the FP64 iterate)
__global__ void __launch_bounds__(kThreads) k_refine_add(double2* x, const float2* e, VN n,
                                                        const int* skip) {
  const int rh_ = n.rhs();  // right-hand side
  if (skip[rh_]) return;
  x += n.off(rh_);
  e += n.off(rh_);
  RHS_BLOCKS_NOSUM(n, rh_);
  GRID_LOOP(i, n.len(rh_)) {
    const float2 v = e[i];
    x[i] = make_double2(x[i].x + (double)v.x, x[i].y + (double)v.y);
  }
}

#undef LD
#undef ST

// ------------------------------------------------------------- host driver
// p[0 .. n) = 0
__global__ void __launch_bounds__(kThreads) k_zero2(double2* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = make_double2(0.0, 0.0);
}

// Block offsets of the flat vector grid of a ragged layout: RHS r gets the
// blocks its own solve launches, none once it has finished (unless all).
__global__ void k_ragged_blocks(const long long* __restrict__ lens, const KState* __restrict__ st,
                                int R, int all, int* __restrict__ kb) {
  if (threadIdx.x != 0) return;
  int acc = 0;
  kb[0] = 0;
  for (int r = 0; r < R; ++r) {
    if (all || !st[r].done) acc += blocks_of(lens[r]);
    kb[r + 1] = acc;
  }
}

void solver_delete(Solver* S) {
  if (!S) return;
  for (int i = 0; i < 14; ++i) {
    (void)hipFree(S->v64[i]);
    (void)hipFree(S->v32[i]);
  }
  (void)hipFree(S->part);
  (void)hipFree(S->st);
  (void)hipHostFree(S->st_host);
  (void)hipFree(S->act);
  (void)hipFree(S->trace);
  (void)hipFree(S->rr64);
  (void)hipFree(S->lens);
  (void)hipFree(S->offs);
  (void)hipFree(S->part5);
  (void)hipFree(S->kb);
  delete S;
}

void solver_free(ifdda_ctx* c) {
  solver_delete((Solver*)c->solver);
  c->solver = nullptr;
}

// Buffers of a solver: vectors of `total` elements for R right-hand sides
// of at most n elements each (R * n when the RHS lie n apart), FP32 copies
// when need32.
static int blocks_for(long long n) {
  return (int)std::max<long long>(1, std::min<long long>(kMaxBlocks, (n + kThreads - 1) / kThreads));
}

static bool solver_alloc(Solver* S, long long n, int R, bool need32, long long total,
                         int part_blocks) {
  S->n = n;
  S->R = R;
  S->total = total;
  S->nblk = blocks_for(n);
  S->cap_total = total;
  S->cap_R = R;
  S->part_blocks = std::max(part_blocks, S->nblk);
  const size_t nv = (size_t)std::max<long long>(total, 1);
  bool ok = hip_ok(hipMalloc(&S->part, sizeof(double2) * kMaxDots * S->part_blocks * R), "hipMalloc partials") &&
            hip_ok(hipMalloc(&S->st, sizeof(KState) * R), "hipMalloc state") &&
            hip_ok(hipHostMalloc(&S->st_host, sizeof(KState) * R), "hipHostMalloc state") &&
            hip_ok(hipMalloc(&S->act, sizeof(int) * R), "hipMalloc flags");
  for (int i = 0; ok && i < 14; ++i)
    ok = hip_ok(hipMalloc(&S->v64[i], sizeof(double2) * nv), "hipMalloc krylov");
  if (ok && need32) {
    ok = hip_ok(hipMalloc(&S->rr64, sizeof(double2) * nv), "hipMalloc resid");
    for (int i = 0; ok && i < 14; ++i)
      ok = hip_ok(hipMalloc(&S->v32[i], sizeof(float2) * nv), "hipMalloc krylov32");
  }
  return ok;
}

Solver* solver_new(int R, bool need32, const long long* lens, const long long* offs,
                   long long min_total, int min_R) {
  long long n = 1, total = 0;
  for (int r = 0; r < R; ++r) {
    n = std::max(n, lens[r]);
    total = std::max(total, offs[r] + lens[r]);
  }
  Solver* S = new Solver();
  const int cap_R = std::max(R, min_R);
  bool ok = solver_alloc(S, n, cap_R, need32, std::max(total, min_total), kMaxBlocks) &&
            hip_ok(hipMalloc(&S->lens, sizeof(long long) * cap_R), "hipMalloc lengths") &&
            hip_ok(hipMalloc(&S->offs, sizeof(long long) * cap_R), "hipMalloc offsets") &&
            hip_ok(hipMemcpy(S->lens, lens, sizeof(long long) * R, hipMemcpyHostToDevice),
                   "copy lengths") &&
            hip_ok(hipMemcpy(S->offs, offs, sizeof(long long) * R, hipMemcpyHostToDevice),
                   "copy offsets") &&
            hip_ok(hipMalloc(&S->kb, sizeof(int) * (cap_R + 1)), "hipMalloc block offsets");
  if (!ok) {
    solver_delete(S);
    return nullptr;
  }
  S->R = R;  // this layout; the buffers hold cap_R right-hand sides
  S->total = total;
  S->lens_h.assign(lens, lens + R);
  S->kb_full = 0;
  for (int r = 0; r < R; ++r) S->kb_full += blocks_of(lens[r]);
  return S;
}

void solver_capacity(const Solver* S, long long* total, int* R) {
  *total = S ? S->cap_total : 0;
  *R = S ? S->cap_R : 0;
}

bool solver_refit(Solver* S, int R, bool need32, const long long* lens, const long long* offs,
                  hipStream_t s) {
  long long n = 1, total = 0;
  for (int r = 0; r < R; ++r) {
    n = std::max(n, lens[r]);
    total = std::max(total, offs[r] + lens[r]);
  }
  if (!S || total > S->cap_total || R > S->cap_R || (need32 && !S->v32[0]) ||
      blocks_for(n) > S->part_blocks || !S->lens || !S->kb)
    return false;
  S->n = n;
  S->R = R;
  S->total = total;
  S->nblk = blocks_for(n);
  S->lens_h.assign(lens, lens + R);
  S->kb_full = 0;
  for (int r = 0; r < R; ++r) S->kb_full += blocks_of(lens[r]);
  return hip_ok(hipMemcpyAsync(S->lens, lens, sizeof(long long) * R, hipMemcpyHostToDevice, s),
                "copy lengths") &&
         hip_ok(hipMemcpyAsync(S->offs, offs, sizeof(long long) * R, hipMemcpyHostToDevice, s),
                "copy offsets");
}

static Solver* solver_get(ifdda_ctx* c, int R, bool need32) {
  Solver* S = (Solver*)c->solver;
  const long long n = 3 * c->nd;
  if (S && (S->n != n || S->R < R)) {
    const int cap = S->tracecap;
    solver_free(c);
    S = nullptr;
    if (cap) {  // keep trace recording enabled across re-allocation
      S = new Solver();
      S->tracecap = cap;
      if (!hip_ok(hipMalloc(&S->trace, sizeof(double2) * 6 * cap), "hipMalloc trace")) {
        delete S;
        return nullptr;
      }
      c->solver = S;
    }
  }
  if (!S || !S->st) {
    if (!S) {
      S = new Solver();
      c->solver = S;
    }
    if (!solver_alloc(S, n, R, false, n * R, 0)) {
      solver_free(c);
      return nullptr;
    }
  }
  if (need32 && !S->v32[0]) {
    const size_t nv = (size_t)std::max<long long>(n, 1) * S->R;
    bool ok = hip_ok(hipMalloc(&S->rr64, sizeof(double2) * nv), "hipMalloc resid");
    for (int i = 0; ok && i < 14; ++i)
      ok = hip_ok(hipMalloc(&S->v32[i], sizeof(float2) * nv), "hipMalloc krylov32");
    if (!ok) {
      solver_free(c);
      return nullptr;
    }
  }
  S->total = n * R;  // this call's right-hand sides (the buffers may hold more)
  return S;
}

// The operator of one context: its box, Green tensor, dipoles and
// polarisability, shared by all R right-hand sides.
struct CtxOp : MvOp {
  ifdda_ctx* c;
  explicit CtxOp(ifdda_ctx* c_) : c(c_) {}
  bool mv64(const double2* x, const double2* b, double2* y, hipStream_t s, const int* skip, int R,
            long long n, int skip_stride) const override {
    MatvecDev<double2> m = c->mv64();
    m.nrhs = R;
    m.vstride = n;
    m.skip_stride = skip_stride;
    return launch_matvec_f64(c->g, m, x, b, y, s, skip);
  }
  bool mv32(const float2* x, const float2* b, float2* y, hipStream_t s, const int* skip, int R,
            long long n, int skip_stride) const override {
    MatvecDev<float2> m = c->mv32();
    m.nrhs = R;
    m.vstride = n;
    m.skip_stride = skip_stride;
    return launch_matvec_f32(c->g, m, x, b, y, s, skip);
  }
  bool alpha_apply(const double2* x, double2* FF, int R, long long n, hipStream_t s) const override {
    const int nb = (int)std::max<long long>(1, std::min<long long>(kMaxBlocks, (n + kThreads - 1) / kThreads));
    hipLaunchKernelGGL(k_alpha_apply, dim3(nb, R), dim3(kThreads), 0, s, x, FF,
                       (const double2*)c->al64, c->alpha_mode, c->nd);
    return hip_ok(hipGetLastError(), "alpha apply");
  }
  int fuse_blocks(int prec, int R) const override {
    if (!c->specialized || !((c->fuse_mask >> (prec == 64 ? 0 : 1)) & 1)) return 0;
    return matvec_p5_blocks(c->g, prec, R);
  }
  bool fuse_p1() const override { return (c->fuse_mask >> 2) & 1; }
  bool fuse_la(int prec) const override { return (c->fuse_mask >> (prec == 64 ? 4 : 3)) & 1; }
  bool mvf64(const double2* x, const double2* b, double2* y, hipStream_t s, const int* skip, int R,
             long long n, int skip_stride, const KFuse<double2>& f) const override {
    MatvecDev<double2> m = c->mv64();
    m.nrhs = R;
    m.vstride = n;
    m.skip_stride = skip_stride;
    return launch_matvec_fused_f64(c->g, m, x, b, y, s, skip, f);
  }
  bool mvf32(const float2* x, const float2* b, float2* y, hipStream_t s, const int* skip, int R,
             long long n, int skip_stride, const KFuse<float2>& f) const override {
    MatvecDev<float2> m = c->mv32();
    m.nrhs = R;
    m.vstride = n;
    m.skip_stride = skip_stride;
    return launch_matvec_fused_f32(c->g, m, x, b, y, s, skip, f);
  }
  // -T x by the matvec without alpha (b = 0), then alpha^T (-T x) + x
  template <typename V>
  bool transposed(MatvecDev<V> m, const V* al, const V* x, const V* zero, V* y, hipStream_t s,
                  const int* skip, int R, long long n, int skip_stride) const {
    m.nrhs = R;
    m.vstride = n;
    m.skip_stride = skip_stride;
    m.alpha_mode = kAlphaNone;
    bool ok;
    if constexpr (std::is_same<V, double2>::value)
      ok = launch_matvec_f64(c->g, m, x, zero, y, s, skip);
    else
      ok = launch_matvec_f32(c->g, m, x, zero, y, s, skip);
    if (!ok) return false;
    const int nb = (int)std::max<long long>(1, std::min<long long>(kMaxBlocks, (c->nd + kThreads - 1) / kThreads));
    hipLaunchKernelGGL(k_alpha_t<V>, dim3(nb, R), dim3(kThreads), 0, s, x, y, al, c->alpha_mode,
                       c->nd, skip, skip_stride);
    return hip_ok(hipGetLastError(), "transposed alpha");
  }
  bool mvt64(const double2* x, const double2* zero, double2* y, hipStream_t s, const int* skip,
             int R, long long n, int skip_stride) const override {
    return transposed<double2>(c->mv64(), (const double2*)c->al64, x, zero, y, s, skip, R, n,
                               skip_stride);
  }
  bool mvt32(const float2* x, const float2* zero, float2* y, hipStream_t s, const int* skip, int R,
             long long n, int skip_stride) const override {
    return transposed<float2>(c->mv32(), (const float2*)c->al32, x, zero, y, s, skip, R, n,
                              skip_stride);
  }
};

template <typename V>
bool alpha_t_launch(const V* x, V* y, const V* al, int mode, long long nd, hipStream_t s,
                    const int* skip) {
  const int nb = (int)std::max<long long>(1, std::min<long long>(kMaxBlocks, (nd + kThreads - 1) / kThreads));
  hipLaunchKernelGGL(k_alpha_t<V>, dim3(nb, 1), dim3(kThreads), 0, s, x, y, al, mode, nd, skip, 0);
  return hip_ok(hipGetLastError(), "transposed alpha");
}
template bool alpha_t_launch<double2>(const double2*, double2*, const double2*, int, long long,
                                      hipStream_t, const int*);
template bool alpha_t_launch<float2>(const float2*, float2*, const float2*, int, long long,
                                     hipStream_t, const int*);

bool MvOp::mvt64(const double2*, const double2*, double2*, hipStream_t, const int*, int, long long,
                 int) const {
  set_error("the transposed operator (QMRCLA) is not available for this solve");
  return false;
}
bool MvOp::mvt32(const float2*, const float2*, float2*, hipStream_t, const int*, int, long long,
                 int) const {
  set_error("the transposed operator (QMRCLA) is not available for this solve");
  return false;
}

bool alpha_apply_one(const double2* x, double2* FF, const double2* al, int mode, long long nd,
                     hipStream_t s) {
  const int nb = (int)std::max<long long>(1, std::min<long long>(kMaxBlocks, (3 * nd + kThreads - 1) / kThreads));
  hipLaunchKernelGGL(k_alpha_apply, dim3(nb, 1), dim3(kThreads), 0, s, x, FF, al, mode, nd);
  return hip_ok(hipGetLastError(), "alpha apply");
}


// Ragged layouts: point the flat vector grid (S->kb, grid G) and the
// operator's launches at the RHS still running (per S->st_host / S->st done
// flags), or at every RHS (all).  The block tables are rewritten on the
// device, in stream order.
static bool ragged_grid(const MvOp& op, Solver* S, int R, bool all, hipStream_t s, dim3& G) {
  std::vector<int> hd(R, 0);
  int total = 0;
  for (int r = 0; r < R; ++r) {
    hd[r] = all ? 0 : S->st_host[r].done;
    if (!hd[r]) total += blocks_of(S->lens_h[r]);
  }
  if (total == 0) return true;  // nothing runs: keep the tables as they are
  hipLaunchKernelGGL(k_ragged_blocks, dim3(1), dim3(64), 0, s, (const long long*)S->lens,
                     (const KState*)S->st, R, all ? 1 : 0, S->kb);
  if (!hip_ok(hipGetLastError(), "ragged grid") || !op.compact(S->st, hd.data(), R, all, s))
    return false;
  G = dim3(total);
  return true;
}

static bool poll(Solver* S, int R, hipStream_t s) {
  return hip_ok(hipMemcpyAsync(S->st_host, S->st, sizeof(KState) * R, hipMemcpyDeviceToHost, s),
                "copy state") &&
         hip_ok(hipStreamSynchronize(s), "solver step");
}

static bool all_done(const Solver* S, int R) {
  for (int r = 0; r < R; ++r)
    if (!S->st_host[r].done) return false;
  return true;
}

// Solves A x = b for R right-hand sides at once on vectors `vv` (14 slots of
// R*n values of type V; slot 0 = x, holding x0 unless x0zero; slot 1 = b).
// tols[r]: tolerance per RHS; active[r] == 0 starts RHS r as finished.
// Returns false on a HIP error; *cancelled when nstop became -1.
template <typename V>
static bool krylov(const MvOp& op, Solver* S, V** vv, int R, int method, const double* tols,
                   const char* active, int maxit, int x0zero, int check_every,
                   volatile int* nstop, hipStream_t s, KRun* out, bool* cancelled) {
  const long long n = S->n;
  const bool ragged = S->lens != nullptr;
  const VN vn{n, S->lens, S->offs, ragged ? S->kb : nullptr, R};
  const int nb = S->nblk;
  dim3 G = ragged ? dim3(S->kb_full) : dim3(nb, R);
  const dim3 T(kThreads), F(1, R), One(1);
  double2* trace = R == 1 ? S->trace : nullptr;
  for (int r = 0; r < R; ++r) {
    KState& k = S->st_host[r];
    std::memset(&k, 0, sizeof(KState));
    k.tol = tols[r];
    k.maxit = maxit;
    k.methode = method == IFDDA_QMRBICGSTAB2 ? 2 : 1;
    k.tracecap = S->tracecap;
    k.steperr = -1;
    if (active && !active[r]) {
      k.done = 1;
      k.status = 1;
    }
  }
  if (!hip_ok(hipMemcpyAsync(S->st, S->st_host, sizeof(KState) * R, hipMemcpyHostToDevice, s),
              "state init"))
    return false;
  // RHS that start finished (refinement: converged) launch nothing
  int nlive = 0;
  for (int r = 0; r < R; ++r) nlive += S->st_host[r].done ? 0 : 1;
  if (ragged && nlive < R && !ragged_grid(op, S, R, false, s, G)) return false;
  const int* done = &S->st->done;
  const int dstride = (int)(sizeof(KState) / sizeof(int));
  int nmv0 = 0;
  V* x = vv[0];
  V* b = vv[1];
  V* Ax = vv[13];
  hipLaunchKernelGGL(k_init<V>, G, T, 0, s, (const V*)b, x, x0zero ? (const V*)nullptr : (const V*)x,
                     vn, S->part, nb);
  hipLaunchKernelGGL(f_init, F, T, 0, s, S->st, (const double2*)S->part, nb);
  if (!x0zero) {
    if (!mv<V>(op, x, x, Ax, s, done, R, n, dstride)) return false;
    ++nmv0;
  }
  // Poll the device state every 8 iterations by default: work queued after
  // a RHS ended is dropped on the device (KState.done, matvec skip flags).
  if (check_every <= 0) check_every = 8;
  *cancelled = false;
  auto step_poll = [&](int it) -> int {  // 1: stop, 0: continue, -1: error
    if ((it + 1) % check_every == 0 || it >= maxit + 1) {
      if (!poll(S, R, s)) return -1;
      if (all_done(S, R)) return 1;
      if (nstop && *nstop == -1) {
        *cancelled = true;
        return 1;
      }
      if (ragged) {  // RHS that finished since the last poll leave the launches
        int live = 0;
        for (int r = 0; r < R; ++r) live += S->st_host[r].done ? 0 : 1;
        if (live < nlive) {
          nlive = live;
          if (!ragged_grid(op, S, R, false, s, G)) return -1;
        }
      }
    }
    return 0;
  };

  if (method == IFDDA_GPBICG1 || method == IFDDA_GPBICG2) {
    const int gp2 = method == IFDDA_GPBICG2;
    V* const tb[2] = {vv[6], vv[12]};  // t / told
    V* const ab[2] = {vv[5], vv[10]};  // Ap / Apold
    GV<V> v{x, b, vv[2], vv[3], vv[4], ab[0], tb[0], vv[7], vv[8], vv[9], ab[1], vv[11], tb[1], Ax};
    // Fused matvec ends (one context): <r0*, Ap> inside the first matvec's
    // last pass, the five products of step 4 inside the second's, and on
    // request p = r + beta (p - u) inside the first matvec's scatter (it
    // replaces g_k1 but reads r, u and p through the dipole map, slower than
    // g_k1's stream on the 100^3 cube).  The vectors are those of the
    // separate kernels; only the products are summed in another order.
    // Look-ahead (la): the second matvec's last pass also sums <r0*, t>,
    // <r0*, y> and <r0*, At>, and g_k5L forms the next p (no g_k1).
    const int prec = std::is_same<V, double2>::value ? 64 : 32;
    const int nb5 = S->lens ? 0 : op.fuse_blocks(prec, R);
    const bool la = nb5 > 0 && op.fuse_la(prec) && S->part5 && S->part5_cap >= 8LL * R * nb5;
    const bool fuse = la || (nb5 > 0 && S->part5 && S->part5_cap >= 5LL * R * nb5);
    const bool p1f = fuse && !la && op.fuse_p1();
    hipLaunchKernelGGL(g_start<V>, G, T, 0, s, v, x0zero, vn, S->part, nb, la ? 1 : 0);
    hipLaunchKernelGGL(g_fstart, F, T, 0, s, S->st, (const double2*)S->part, nb);
    for (int it = 0; la; ++it) {
      v.told = tb[it & 1];
      v.t = tb[(it + 1) & 1];
      v.Ap = ab[it & 1];
      v.Apold = ab[(it + 1) & 1];
      KFuse<V> f1;
      f1.p5 = 1;
      f1.r0s = v.r0s;
      f1.part = S->part5;
      f1.pstride = nb5;
      if (!mvf<V>(op, v.p, v.p, v.Ap, s, done, R, n, dstride, f1)) return false;
      hipLaunchKernelGGL(g_f2xL, F, T, 0, s, S->st, (const double2*)S->part, nb,
                         (const double2*)S->part5, nb5, trace);
      hipLaunchKernelGGL(g_k3<V>, G, T, 0, s, (const KState*)S->st, v, vn, gp2);
      KFuse<V> f2;
      f2.p5 = 3;
      f2.y = v.y;
      f2.r0s = v.r0s;
      f2.part = S->part5;
      f2.pstride = nb5;
      if (!mvf<V>(op, v.t, v.t, v.At, s, done, R, n, dstride, f2)) return false;
      hipLaunchKernelGGL(g_f4xL, F, T, 0, s, S->st, (const double2*)S->part5, nb5, gp2);
      hipLaunchKernelGGL(g_k5L<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      if (it < kLaSplit)
        hipLaunchKernelGGL(g_f5L, F, T, 0, s, S->st, (const double2*)S->part, nb, trace);
      if (!hip_ok(hipGetLastError(), "gpbicg launch")) return false;
      const int st = step_poll(it);
      if (st < 0) return false;
      if (st > 0) break;
    }
    for (int it = 0; fuse && !la; ++it) {
      v.told = tb[it & 1];
      v.t = tb[(it + 1) & 1];
      v.Ap = ab[it & 1];
      v.Apold = ab[(it + 1) & 1];
      if (!p1f) hipLaunchKernelGGL(g_k1<V>, G, T, 0, s, (const KState*)S->st, v, vn);
      KFuse<V> f1;
      f1.p1 = p1f ? 1 : 0;
      f1.p = v.p;
      f1.r = v.r;
      f1.u = v.u;
      f1.st = S->st;
      f1.p5 = 1;
      f1.r0s = v.r0s;
      f1.part = S->part5;
      f1.pstride = nb5;
      if (!mvf<V>(op, v.p, v.p, v.Ap, s, done, R, n, dstride, f1)) return false;
      hipLaunchKernelGGL(g_f2x, F, T, 0, s, S->st, (const double2*)S->part5, nb5);
      hipLaunchKernelGGL(g_k3<V>, G, T, 0, s, (const KState*)S->st, v, vn, gp2);
      KFuse<V> f2;
      f2.p5 = 2;
      f2.y = v.y;
      f2.part = S->part5;
      f2.pstride = nb5;
      if (!mvf<V>(op, v.t, v.t, v.At, s, done, R, n, dstride, f2)) return false;
      hipLaunchKernelGGL(g_f4x, F, T, 0, s, S->st, (const double2*)S->part5, nb5, gp2);
      hipLaunchKernelGGL(g_k5<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(g_f5, F, T, 0, s, S->st, (const double2*)S->part, nb, trace);
      if (!hip_ok(hipGetLastError(), "gpbicg launch")) return false;
      const int st = step_poll(it);
      if (st < 0) return false;
      if (st > 0) break;
    }
    for (int it = 0; !fuse; ++it) {
      v.told = tb[it & 1];
      v.t = tb[(it + 1) & 1];
      v.Ap = ab[it & 1];
      v.Apold = ab[(it + 1) & 1];
      hipLaunchKernelGGL(g_k1<V>, G, T, 0, s, (const KState*)S->st, v, vn);
      if (!mv<V>(op, v.p, v.p, v.Ap, s, done, R, n, dstride)) return false;
      hipLaunchKernelGGL(g_k2<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(g_f2, F, T, 0, s, S->st, (const double2*)S->part, nb);
      hipLaunchKernelGGL(g_k3<V>, G, T, 0, s, (const KState*)S->st, v, vn, gp2);
      if (!mv<V>(op, v.t, v.t, v.At, s, done, R, n, dstride)) return false;
      hipLaunchKernelGGL(g_k4<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(g_f4, F, T, 0, s, S->st, (const double2*)S->part, nb, gp2);
      hipLaunchKernelGGL(g_k5<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(g_f5, F, T, 0, s, S->st, (const double2*)S->part, nb, trace);
      if (!hip_ok(hipGetLastError(), "gpbicg launch")) return false;
      const int st = step_poll(it);
      if (st < 0) return false;
      if (st > 0) break;
    }
    for (int r = 0; r < R; ++r) {
      const KState& h = S->st_host[r];
      KRun& o = out[r];
      o.status = h.status;
      o.steperr = h.steperr;
      if (active && !active[r]) continue;
      o.nloop = h.it + 1;  // ITNO starts at -1 (gpbicg.f:156), the driver adds one
      // two matvecs per finished iteration; a breakdown at STEPERR 1 stops
      // after the first
      o.nmv = nmv0 + 2 * h.it + (h.status == -1 && h.steperr == 1 ? 1 : 2);
    }
  } else if (method == IFDDA_QMRBICGSTAB1 || method == IFDDA_QMRBICGSTAB2) {
    QV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], vv[8], vv[9], vv[10], Ax};
    hipLaunchKernelGGL(q_start<V>, G, T, 0, s, v, x0zero, vn, S->part, nb);
    hipLaunchKernelGGL(q_fstart, F, T, 0, s, S->st, (const double2*)S->part, nb);
    for (int it = 0;; ++it) {
      hipLaunchKernelGGL(q_k1<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(q_f1, F, T, 0, s, S->st, (const double2*)S->part, nb);
      hipLaunchKernelGGL(q_k2<V>, G, T, 0, s, (const KState*)S->st, v, vn);
      if (!mv<V>(op, v.p, v.p, v.v, s, done, R, n, dstride)) return false;
      hipLaunchKernelGGL(q_k3<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(q_f3, F, T, 0, s, S->st, (const double2*)S->part, nb);
      hipLaunchKernelGGL(q_k4<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(q_f4, F, T, 0, s, S->st, (const double2*)S->part, nb);
      hipLaunchKernelGGL(q_k5<V>, G, T, 0, s, (const KState*)S->st, v, vn);
      if (!mv<V>(op, v.s, v.s, v.t, s, done, R, n, dstride)) return false;
      hipLaunchKernelGGL(q_k6<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(q_f6, F, T, 0, s, S->st, (const double2*)S->part, nb);
      hipLaunchKernelGGL(q_k7<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(q_f7, F, T, 0, s, S->st, (const double2*)S->part, nb);
      hipLaunchKernelGGL(q_k8<V>, G, T, 0, s, (const KState*)S->st, v, vn);
      hipLaunchKernelGGL(q_f8, F, One, 0, s, S->st, trace);
      if (!hip_ok(hipGetLastError(), "qmrbicgstab launch")) return false;
      const int st = step_poll(it);
      if (st < 0) return false;
      if (st > 0) break;
    }
    for (int r = 0; r < R; ++r) {
      const KState& h = S->st_host[r];
      KRun& o = out[r];
      o.status = h.status;
      o.steperr = h.steperr;
      if (active && !active[r]) continue;
      o.nloop = h.it;  // ITNO counts from 1 (qmrbicgstab.f:75); no driver increment
      int last = 2;
      if (h.status == -1 && (h.steperr == 1 || h.steperr == 2)) last = 1;
      o.nmv = nmv0 + 2 * (h.it - 1) + last;
    }
  } else if (method == IFDDA_TFQMR) {
    TV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], vv[8], vv[9], vv[10], Ax};
    hipLaunchKernelGGL(t_start<V>, G, T, 0, s, v, x0zero, vn);
    if (!mv<V>(op, v.u, v.u, v.v, s, done, R, n, dstride)) return false;
    hipLaunchKernelGGL(t_start2<V>, G, T, 0, s, v, vn, S->part, nb);
    hipLaunchKernelGGL(t_fstart, F, T, 0, s, S->st, (const double2*)S->part, nb);
    for (int it = 0;; ++it) {
      const int even = (it % 2 == 0);
      if (even) {
        hipLaunchKernelGGL(t_k1<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
        hipLaunchKernelGGL(t_f1, F, T, 0, s, S->st, (const double2*)S->part, nb);
      }
      hipLaunchKernelGGL(t_k2<V>, G, T, 0, s, (const KState*)S->st, v, even, vn, S->part, nb);
      hipLaunchKernelGGL(t_f2, F, T, 0, s, S->st, (const double2*)S->part, nb);
      hipLaunchKernelGGL(t_k3<V>, G, T, 0, s, (const KState*)S->st, v, vn);
      hipLaunchKernelGGL(t_f3, F, One, 0, s, S->st, trace);
      if (!even) {
        hipLaunchKernelGGL(t_k4<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
        hipLaunchKernelGGL(t_f4, F, T, 0, s, S->st, (const double2*)S->part, nb);
      }
      hipLaunchKernelGGL(t_k5<V>, G, T, 0, s, (const KState*)S->st, v, !even, vn);
      if (!mv<V>(op, v.u, v.u, v.Au, s, done, R, n, dstride)) return false;
      if (!even) hipLaunchKernelGGL(t_k6<V>, G, T, 0, s, (const KState*)S->st, v, vn);
      hipLaunchKernelGGL(t_f5, F, One, 0, s, S->st);
      if (!hip_ok(hipGetLastError(), "tfqmr launch")) return false;
      const int st = step_poll(it);
      if (st < 0) return false;
      if (st > 0) break;
    }
    for (int r = 0; r < R; ++r) {
      const KState& h = S->st_host[r];
      KRun& o = out[r];
      o.status = h.status;
      o.steperr = h.steperr;
      if (active && !active[r]) continue;
      o.nloop = h.it + 1;        // ITNO starts at -1 (tfqmr.f:96), the driver adds one
      o.nmv = nmv0 + 1 + h.it;   // A u0, then one matvec per half step that did not converge
    }
  } else if (method == IFDDA_CG) {
    CV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], Ax};
    hipLaunchKernelGGL(c_start<V>, G, T, 0, s, v, x0zero, vn, S->part, nb);
    hipLaunchKernelGGL(c_fstart, F, T, 0, s, S->st, (const double2*)S->part, nb);
    if (!mv<V>(op, v.p, v.p, v.w, s, done, R, n, dstride)) return false;
    hipLaunchKernelGGL(c_k1<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
    hipLaunchKernelGGL(c_f1, F, T, 0, s, S->st, (const double2*)S->part, nb);
    for (int it = 0;; ++it) {
      hipLaunchKernelGGL(c_k2<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(c_f2, F, T, 0, s, S->st, (const double2*)S->part, nb, trace);
      if (!mv<V>(op, v.r, v.r, v.s, s, done, R, n, dstride)) return false;
      hipLaunchKernelGGL(c_k3<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(c_f3, F, T, 0, s, S->st, (const double2*)S->part, nb);
      if (!hip_ok(hipGetLastError(), "cg launch")) return false;
      const int st = step_poll(it);
      if (st < 0) return false;
      if (st > 0) break;
    }
    for (int r = 0; r < R; ++r) {
      const KState& h = S->st_host[r];
      KRun& o = out[r];
      o.status = h.status;
      o.steperr = h.steperr;
      if (active && !active[r]) continue;
      o.nloop = h.it;  // NLOOP (zcg.f:92)
      // A p0, then one A r per iteration that went on (zcg.f:140-148)
      o.nmv = nmv0 + std::max(h.it, 1);
    }
  } else if (method == IFDDA_BICGSTAB) {
    BV<V> v{x, b, vv[2], vv[3], vv[4], vv[5], vv[6], vv[7], Ax};
    hipLaunchKernelGGL(s_start<V>, G, T, 0, s, v, x0zero, vn, S->part, nb);
    hipLaunchKernelGGL(s_fstart, F, T, 0, s, S->st, (const double2*)S->part, nb);
    for (int it = 0;; ++it) {
      hipLaunchKernelGGL(s_k1<V>, G, T, 0, s, (const KState*)S->st, v, vn);
      if (!mv<V>(op, v.p, v.p, v.v, s, done, R, n, dstride)) return false;
      hipLaunchKernelGGL(s_k2<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(s_f2, F, T, 0, s, S->st, (const double2*)S->part, nb);
      hipLaunchKernelGGL(s_k3<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(s_f3, F, T, 0, s, S->st, (const double2*)S->part, nb);
      if (!mv<V>(op, v.s, v.s, v.t, s, done, R, n, dstride)) return false;
      hipLaunchKernelGGL(s_k4<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(s_f4, F, T, 0, s, S->st, (const double2*)S->part, nb);
      hipLaunchKernelGGL(s_k5<V>, G, T, 0, s, (const KState*)S->st, v, vn, S->part, nb);
      hipLaunchKernelGGL(s_f5, F, T, 0, s, S->st, (const double2*)S->part, nb, trace);
      if (!hip_ok(hipGetLastError(), "bicgstab launch")) return false;
      const int st = step_poll(it);
      if (st < 0) return false;
      if (st > 0) break;
    }
    hipLaunchKernelGGL(s_soft<V>, G, T, 0, s, (const KState*)S->st, v, vn);
    for (int r = 0; r < R; ++r) {
      const KState& h = S->st_host[r];
      KRun& o = out[r];
      o.status = h.status;
      o.steperr = h.steperr;
      if (active && !active[r]) continue;
      const int k = h.it;  // ITNO (pimzbicgstab.f:69)
      o.nloop = std::min(k, maxit);
      // A p and A s per iteration; a breakdown ends after the first (10,
      // 12) or before either (6) of them (pimzbicgstab.f:79-134)
      int nmv = 2 * k;
      if (h.status == -1 && (h.steperr == 10 || h.steperr == 12)) nmv = 2 * k - 1;
      if (h.status == -1 && h.steperr == 6) nmv = 2 * (k - 1);
      o.nmv = nmv0 + nmv;
    }
  } else if (krylov_gp_method(method)) {
    KLoop L{op, S, vn, &G, nb, R, n, s, trace, done, dstride, nmv0, x0zero, active, step_poll};
    if (!krylov_gp<V>(L, vv, method, out)) return false;
  } else {
    set_error("unknown method");
    return false;
  }
  // every RHS back in the launches (the caller's kernels cover them all)
  return !ragged || nlive == R || ragged_grid(op, S, R, true, s, G);
}

// Norms of R residual vectors whose partial sums k_resid / k_init /
// k_refine_resid left in S->part.
static bool resid_norms(Solver* S, int R, double* out, hipStream_t s) {
  hipLaunchKernelGGL(f_resid, dim3(1, R), dim3(kThreads), 0, s, S->st, (const double2*)S->part, S->nblk);
  if (!poll(S, R, s)) return false;
  for (int r = 0; r < R; ++r) out[r] = std::sqrt(S->st_host[r].dots[0].x);
  return true;
}

static void set_info_str(ifdda_solve_info* info, const char* m) {
  std::memset(info->infostr, ' ', sizeof(info->infostr));
  std::memcpy(info->infostr, m, std::min(std::strlen(m), sizeof(info->infostr)));
}

// Device solve of R right-hand sides: E_inc, FF, FFloc are device complex128
// arrays of R * 3*nd values (FF / FFloc may be null).
static int solve_dev(ifdda_ctx* c, const ifdda_solve_opts* o, int R, const double2* Einc,
                     double2* FF, double2* FFloc, ifdda_solve_info* infos, volatile int* nstop,
                     hipStream_t s) {
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < R; ++r) {
    std::memset(&infos[r], 0, sizeof(ifdda_solve_info));
    set_info_str(&infos[r], "");
  }
  if (o->method < IFDDA_GPBICG1 || o->method > IFDDA_QMRCLA) {
    set_error("ifdda_solve: unknown method");
    return IFDDA_ERR_ARG;
  }
  if (o->precision != 64 && o->precision != 32) {
    set_error("ifdda_solve: precision must be 64 or 32");
    return IFDDA_ERR_ARG;
  }
  if (R < 1) {
    set_error("ifdda_solve: nrhs must be >= 1");
    return IFDDA_ERR_ARG;
  }
  if (!c->green_ready || !c->have_dipoles || c->alpha_mode == kAlphaNone) {
    set_error("ifdda_solve: Green tensor, dipoles and polarisability must be set");
    return IFDDA_ERR_STATE;
  }
  if (ensure_scratch(c, R) != IFDDA_OK) return IFDDA_ERR_HIP;
  Solver* S = solver_get(c, R, o->precision == 32);
  if (!S) return IFDDA_ERR_HIP;
  if (o->method == IFDDA_GPBICG1 || o->method == IFDDA_GPBICG2) {
    // partial sums of the inner products formed in fused matvec ends
    const long long need = 8LL * R * CtxOp(c).fuse_blocks(o->precision, R);
    if (need > S->part5_cap) {
      (void)hipFree(S->part5);
      S->part5 = nullptr;
      S->part5_cap = 0;
      if (!hip_ok(hipMalloc(&S->part5, sizeof(double2) * need), "hipMalloc fused partials"))
        return IFDDA_ERR_HIP;
      S->part5_cap = need;
    }
  }
  return solve_core(CtxOp(c), S, o, R, Einc, FF, FFloc, infos, nstop, s, t0);
}

int solve_core(const MvOp& op, Solver* S, const ifdda_solve_opts* o, int R, const double2* Einc,
               double2* FF, double2* FFloc, ifdda_solve_info* infos, volatile int* nstop,
               hipStream_t s, std::chrono::steady_clock::time_point t0) {
  const int maxit = o->maxit > 0 ? o->maxit : 10000;
  const long long n = S->n;
  const bool ragged = S->lens != nullptr;
  const VN vn{n, S->lens, S->offs, ragged ? S->kb : nullptr, R};
  const int nb = S->nblk;
  const dim3 G = ragged ? dim3(S->kb_full) : dim3(nb, R), T(kThreads);
  if (ragged) {
    // the flat grid with every RHS in; partial slots past an RHS's block
    // count stay zero (no block writes them)
    hipLaunchKernelGGL(k_zero2, dim3(kMaxBlocks), T, 0, s, S->part,
                       (long long)kMaxDots * S->part_blocks * S->cap_R);
    hipLaunchKernelGGL(k_ragged_blocks, dim3(1), dim3(64), 0, s, (const long long*)S->lens,
                       (const KState*)S->st, R, 1, S->kb);
    if (!hip_ok(hipGetLastError(), "ragged layout")) return IFDDA_ERR_HIP;
  }
  double2* x = S->v64[0];
  double2* b = S->v64[1];
  double2* Ax = S->v64[13];
  if (!hip_ok(hipMemcpyAsync(b, Einc, sizeof(double2) * S->total, hipMemcpyDeviceToDevice, s),
              "copy E_inc"))
    return IFDDA_ERR_HIP;
  if (!o->x0_zero &&
      !hip_ok(hipMemcpyAsync(x, Einc, sizeof(double2) * S->total, hipMemcpyDeviceToDevice, s),
              "copy x0"))
    return IFDDA_ERR_HIP;
  if (nstop && *nstop == -1) {
    for (int r = 0; r < R; ++r) {
      infos[r].nstop = -1;
      set_info_str(&infos[r], "Calculation cancelled during iterative method");
    }
    return IFDDA_OK;
  }
  std::vector<KRun> run(R);
  std::vector<int> ncompte(R, 0), nloop(R, 0), outer(R, 0);
  // ax_ok[r]: Ax of RHS r holds the FP64 matvec of its current x (the last
  // refinement check), which is then also the final true residual's
  std::vector<char> ax_ok(R, 0);
  std::vector<double> e(R), bn(R);  // final |b - A x| and |b|
  bool cancelled = false;
  if (o->precision == 64) {
    std::vector<double> tols(R, o->tol);
    if (!krylov<double2>(op, S, S->v64, R, o->method, tols.data(), nullptr, maxit, o->x0_zero,
                         o->check_every, nstop, s, run.data(), &cancelled))
      return IFDDA_ERR_HIP;
    for (int r = 0; r < R; ++r) {
      ncompte[r] = run[r].nmv;
      nloop[r] = run[r].nloop;
    }
  } else {
    // Iterative refinement with FP32 inner solves, per RHS.
    double2* r64 = S->rr64;
    float2** v32 = S->v32;
    std::vector<double> bnorm(R), rn(R), rho(R), tin(R);
    std::vector<char> active(R, 1);
    std::vector<int> act(R, 0);
    hipLaunchKernelGGL(k_init<double2>, G, T, 0, s, (const double2*)b, x,
                       o->x0_zero ? (const double2*)nullptr : (const double2*)x, vn, S->part, nb);
    if (!resid_norms(S, R, bnorm.data(), s)) return IFDDA_ERR_HIP;
    bool have_ax = !o->x0_zero;
    if (!hip_ok(hipMemcpyAsync(S->act, act.data(), sizeof(int) * R, hipMemcpyHostToDevice, s),
                "copy flags"))
      return IFDDA_ERR_HIP;
    for (int sweep = 0;; ++sweep) {
      if (have_ax) {
        if (!mv<double2>(op, x, x, Ax, s, S->act, R, n, 1)) return IFDDA_ERR_HIP;
        for (int r = 0; r < R; ++r) {
          ncompte[r] += active[r] ? 1 : 0;
          if (active[r]) ax_ok[r] = 1;
        }
      }
      hipLaunchKernelGGL(k_refine_resid, G, T, 0, s, (const double2*)b, (const double2*)Ax,
                         (int)have_ax, r64, v32[1], vn, S->part, nb);
      if (!resid_norms(S, R, rn.data(), s)) return IFDDA_ERR_HIP;
      have_ax = true;
      int nact = 0;
      for (int r = 0; r < R; ++r) {
        if (!active[r]) continue;
        e[r] = rn[r];  // the final residual too, if x stays as it is
        bn[r] = bnorm[r];
        rho[r] = bnorm[r] > 0.0 ? rn[r] / bnorm[r] : 0.0;
        outer[r] = sweep;
        if (rho[r] <= o->tol || sweep >= 30) {
          active[r] = 0;
          act[r] = 1;
          run[r].status = 1;
          run[r].steperr = rho[r] <= o->tol ? -1 : 0;
        } else {
          ++nact;
          // Inner relative tolerance: 0.9 of the remaining gap (the FP32
          // recurrence residual tracks the FP64 true residual to ~1e-6), not
          // below what an FP32 recurrence resolves.
          tin[r] = std::max(0.9 * o->tol / rho[r], 1e-5);
        }
      }
      if (!hip_ok(hipMemcpyAsync(S->act, act.data(), sizeof(int) * R, hipMemcpyHostToDevice, s),
                  "copy flags"))
        return IFDDA_ERR_HIP;
      if (nact == 0) break;
      if (nstop && *nstop == -1) {
        cancelled = true;
        break;
      }
      std::vector<KRun> in(R);
      if (!krylov<float2>(op, S, v32, R, o->method, tin.data(), active.data(), maxit, 1,
                          o->check_every, nstop, s, in.data(), &cancelled))
        return IFDDA_ERR_HIP;
      bool broke = false;
      for (int r = 0; r < R; ++r) {
        if (!active[r]) continue;
        ncompte[r] += in[r].nmv;
        nloop[r] += in[r].nloop;
        if (in[r].status < 0) {  // breakdown in the inner solve ends this RHS
          run[r] = in[r];
          active[r] = 0;
          act[r] = 1;
          broke = true;
        }
      }
      if (cancelled) break;
      if (broke &&
          !hip_ok(hipMemcpyAsync(S->act, act.data(), sizeof(int) * R, hipMemcpyHostToDevice, s),
                  "copy flags"))
        return IFDDA_ERR_HIP;
      hipLaunchKernelGGL(k_refine_add, G, T, 0, s, x, (const float2*)v32[0], vn, (const int*)S->act);
      for (int r = 0; r < R; ++r)
        if (active[r]) ax_ok[r] = 0;  // x moved
    }
  }
  if (cancelled) {
    for (int r = 0; r < R; ++r) {
      infos[r].nstop = -1;
      infos[r].ncompte = ncompte[r];
      set_info_str(&infos[r], "Calculation cancelled during iterative method");
    }
    return IFDDA_OK;
  }
  // Final true residual with the FP64 matvec (inverserig.f:95-105); after
  // refinement the last check already holds it for every RHS.
  const bool final_mv = std::find(ax_ok.begin(), ax_ok.end(), 0) != ax_ok.end();
  // Mixed precision with x unmoved since every right-hand side's last
  // refinement check: that check's FP64 residual |b - A x| and |b| are the
  // final ones (same matvec, same sums).  Else the true residual is formed.
  if (o->precision != 32 || final_mv) {
    if (final_mv && !mv<double2>(op, x, x, Ax, s, nullptr, R, n, 0)) return IFDDA_ERR_HIP;
    hipLaunchKernelGGL(k_resid<double2>, G, T, 0, s, (const double2*)Ax, (const double2*)b, vn,
                       S->part, nb);
    if (!resid_norms(S, R, e.data(), s)) return IFDDA_ERR_HIP;
    hipLaunchKernelGGL(k_init<double2>, G, T, 0, s, (const double2*)b, Ax, (const double2*)Ax, vn,
                       S->part, nb);
    if (!resid_norms(S, R, bn.data(), s)) return IFDDA_ERR_HIP;
  }
  if (FFloc &&
      !hip_ok(hipMemcpyAsync(FFloc, x, sizeof(double2) * S->total, hipMemcpyDeviceToDevice, s),
              "copy E"))
    return IFDDA_ERR_HIP;
  if (FF && !op.alpha_apply(x, FF, R, n, s)) return IFDDA_ERR_HIP;
  if (!hip_ok(hipStreamSynchronize(s), "solve")) return IFDDA_ERR_HIP;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int r = 0; r < R; ++r) {
    ifdda_solve_info& inf = infos[r];
    inf.status = run[r].status;
    inf.steperr = run[r].steperr;
    inf.nloop = nloop[r];
    // the final matvec runs for every RHS once one needs it; it counts only
    // for the RHS whose last refinement check did not already give A x
    inf.ncompte = ncompte[r] + (ax_ok[r] ? 0 : 1);
    inf.outer = o->precision == 32 ? outer[r] : 0;
    inf.seconds = secs;
    inf.tol1 = bn[r] > 0.0 ? e[r] / bn[r] : 0.0;
    // inverserig.f:796-808 divides BICGSTAB's residual by the NORM the solver
    // returns, the squared initial residual norm (pimzbicgstab.f:50-53)
    if (o->method == IFDDA_BICGSTAB && o->precision == 64 && S->st_host[r].norm > 0.0)
      inf.tol1 = e[r] / S->st_host[r].norm;
    if (run[r].status < 0) {  // inverserig.f:47-52; the TFQMR branch has no '!' (:614)
      inf.nstop = 1;
      inf.ncompte = ncompte[r];
      set_info_str(&inf, o->method == IFDDA_TFQMR ? "Problem to solve Ax=b" : "Problem to solve Ax=b!");
    }
  }
  return IFDDA_OK;
}

}  // namespace ifdda

using namespace ifdda;

extern "C" int ifdda_solve_multi_device(ifdda_ctx* c, const ifdda_solve_opts* o, int nrhs,
                                        const void* E_inc, void* FF, void* FFloc,
                                        ifdda_solve_info* infos, volatile int* nstop,
                                        void* stream) {
  if (!c || !o || !E_inc || !infos) {
    set_error("ifdda_solve_multi_device: NULL argument");
    return IFDDA_ERR_ARG;
  }
  return solve_dev(c, o, nrhs, (const double2*)E_inc, (double2*)FF, (double2*)FFloc, infos, nstop,
                   stream ? (hipStream_t)stream : c->stream);
}

extern "C" int ifdda_solve_device(ifdda_ctx* c, const ifdda_solve_opts* o, const void* E_inc,
                                  void* FF, void* FFloc, ifdda_solve_info* info,
                                  volatile int* nstop, void* stream) {
  return ifdda_solve_multi_device(c, o, 1, E_inc, FF, FFloc, info, nstop, stream);
}

extern "C" int ifdda_solve_multi(ifdda_ctx* c, const ifdda_solve_opts* o, int nrhs,
                                 const double* E_inc, double* FF, double* FFloc,
                                 ifdda_solve_info* infos, volatile int* nstop) {
  if (!c || !o || !E_inc || !infos) {
    set_error("ifdda_solve_multi: NULL argument");
    return IFDDA_ERR_ARG;
  }
  if (!c->have_dipoles) {
    set_error("ifdda_solve: no dipoles");
    return IFDDA_ERR_STATE;
  }
  if (nrhs < 1) {
    set_error("ifdda_solve_multi: nrhs must be >= 1");
    return IFDDA_ERR_ARG;
  }
  const size_t n = 3 * (size_t)c->nd * nrhs;
  if (n == 0) {
    for (int r = 0; r < nrhs; ++r) std::memset(&infos[r], 0, sizeof(ifdda_solve_info));
    return IFDDA_OK;
  }
  // device copies of E_inc, FF, FFloc in the context's grow-only staging
  // (a solve per call would otherwise pay three allocations and three
  // device-synchronising frees)
  if (c->scap < 3 * n) {
    (void)hipFree(c->sstage);
    c->sstage = nullptr;
    c->scap = 0;
    if (!hip_ok(hipMalloc(&c->sstage, sizeof(double2) * 3 * n), "hipMalloc solve staging"))
      return IFDDA_ERR_HIP;
    c->scap = 3 * n;
  }
  double2 *dE = c->sstage, *dFF = dE + n, *dL = dFF + n;
  hipStream_t s = c->stream;
  int rc = IFDDA_OK;
  if (!hip_ok(hipMemcpyAsync(dE, E_inc, sizeof(double2) * n, hipMemcpyHostToDevice, s), "copy E_inc"))
    rc = IFDDA_ERR_HIP;
  if (!rc) rc = solve_dev(c, o, nrhs, dE, dFF, dL, infos, nstop, s);
  if (!rc) {
    if (FF && !hip_ok(hipMemcpyAsync(FF, dFF, sizeof(double2) * n, hipMemcpyDeviceToHost, s), "copy FF"))
      rc = IFDDA_ERR_HIP;
    if (!rc && FFloc &&
        !hip_ok(hipMemcpyAsync(FFloc, dL, sizeof(double2) * n, hipMemcpyDeviceToHost, s), "copy FFloc"))
      rc = IFDDA_ERR_HIP;
  }
  if (!hip_ok(hipStreamSynchronize(s), "ifdda_solve")) rc = rc ? rc : IFDDA_ERR_HIP;
  return rc;
}

extern "C" int ifdda_solve(ifdda_ctx* c, const ifdda_solve_opts* o, const double* E_inc,
                           double* FF, double* FFloc, ifdda_solve_info* info,
                           volatile int* nstop) {
  return ifdda_solve_multi(c, o, 1, E_inc, FF, FFloc, info, nstop);
}

// Trace access for parity tests: copies up to cap rows of
// (alpha, beta, s2, s3, s4, residual) recorded by the last single-RHS solve.
extern "C" int ifdda_solve_trace(ifdda_ctx* c, int cap, double* out, int* nrows) {
  Solver* S = c ? (Solver*)c->solver : nullptr;
  if (!S || !S->st || !out || !nrows) {
    set_error("ifdda_solve_trace: no solver state");
    return IFDDA_ERR_STATE;
  }
  if (!hip_ok(hipMemcpy(S->st_host, S->st, sizeof(KState), hipMemcpyDeviceToHost), "copy state"))
    return IFDDA_ERR_HIP;
  const int n = std::min(cap, S->st_host->ntrace);
  *nrows = n;
  if (n > 0 && !hip_ok(hipMemcpy(out, S->trace, sizeof(double2) * 6 * n, hipMemcpyDeviceToHost),
                       "copy trace"))
    return IFDDA_ERR_HIP;
  return IFDDA_OK;
}

// Enable per-iteration scalar recording for the next single-RHS solves.
extern "C" int ifdda_solve_trace_enable(ifdda_ctx* c, int cap) {
  if (!c || cap < 0) {
    set_error("ifdda_solve_trace_enable: bad argument");
    return IFDDA_ERR_ARG;
  }
  Solver* S = (Solver*)c->solver;
  if (!S) {
    S = new Solver();
    c->solver = S;
  }
  (void)hipFree(S->trace);
  S->trace = nullptr;
  S->tracecap = 0;
  if (cap > 0) {
    if (!hip_ok(hipMalloc(&S->trace, sizeof(double2) * 6 * cap), "hipMalloc trace"))
      return IFDDA_ERR_HIP;
    S->tracecap = cap;
  }
  return IFDDA_OK;
}

// ---------------------------------------------------------------------------
// ifdda_solve_device_noalloc: the device solve of one right-hand side with
// every Krylov work vector, partial sum and state taken from `scratch`, a
// fixed array of IFDDA_NOALLOC_BYTES in device memory that each GPU holds
// for the life of the process.  The context's own buffers (FFT scratch,
// Green tensor, map, polarisability) are made when the context is set up;
// after that this call makes no device allocation.  The host never uses the
// array's name as an address: hipGetSymbolAddress gives each device's copy
// once, and that pointer is what the kernels receive.
// ---------------------------------------------------------------------------
namespace ifdda {

static __device__ __align__(256) unsigned char scratch[IFDDA_NOALLOC_BYTES];

namespace {

constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu[kNoAllocDevices];   // one solve at a time per device
unsigned char* g_noalloc_ptr[kNoAllocDevices] = {};
std::mutex g_noalloc_init;

// The current device's copy of `scratch` (looked up on its first use).
int noalloc_base(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_noalloc_init);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("ifdda_solve_device_noalloc: hipGetSymbolAddress: ") +
                hipGetErrorString(e));
      return IFDDA_ERR_HIP;
    }
    g_noalloc_ptr[dev] = (unsigned char*)p;
  }
  *out = g_noalloc_ptr[dev];
  return IFDDA_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the work of one right-hand side of n elements; with base set,
// also points the solver's buffers into base.
size_t noalloc_layout(Solver* S, long long n, bool need32, long long part5, unsigned char* base) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> unsigned char* {
    unsigned char* p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  const int nblk = blocks_for(n);
  S->part = (double2*)take(sizeof(double2) * kMaxDots * nblk);
  S->st = (KState*)take(sizeof(KState));
  S->act = (int*)take(sizeof(int));
  S->part5 = part5 > 0 ? (double2*)take(sizeof(double2) * part5) : nullptr;
  S->part5_cap = part5;
  for (int i = 0; i < 14; ++i) S->v64[i] = (double2*)take(sizeof(double2) * n);
  if (need32) {
    S->rr64 = (double2*)take(sizeof(double2) * n);
    for (int i = 0; i < 14; ++i) S->v32[i] = (float2*)take(sizeof(float2) * n);
  }
  return off;
}

}  // namespace
}  // namespace ifdda

extern "C" int ifdda_solve_device_noalloc(ifdda_ctx* c, const ifdda_solve_opts* o,
                                          const void* E_inc, void* FF, void* FFloc,
                                          ifdda_solve_info* info, volatile int* nstop,
                                          void* stream) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!c || !o || !E_inc || !info) {
    set_error("ifdda_solve_device_noalloc: NULL argument");
    return IFDDA_ERR_ARG;
  }
  std::memset(info, 0, sizeof(*info));
  set_info_str(info, "");
  if (o->method < IFDDA_GPBICG1 || o->method > IFDDA_QMRCLA || (o->precision != 64 && o->precision != 32)) {
    set_error("ifdda_solve_device_noalloc: unknown method or precision");
    return IFDDA_ERR_ARG;
  }
  if (!c->green_ready || !c->have_dipoles || c->alpha_mode == kAlphaNone || c->scr_rhs < 1 ||
      !c->scrA || !c->scrB) {
    set_error("ifdda_solve_device_noalloc: Green tensor, dipoles and polarisability must be set");
    return IFDDA_ERR_STATE;
  }
  const long long n = 3 * (long long)c->nd;
  if (n == 0) return IFDDA_OK;
  Solver S;
  const bool need32 = o->precision == 32;
  const long long part5 = o->method == IFDDA_GPBICG1 || o->method == IFDDA_GPBICG2
                              ? 8LL * CtxOp(c).fuse_blocks(o->precision, 1)
                              : 0;
  const size_t need = noalloc_layout(&S, n, need32, part5, nullptr);
  if (need > (size_t)IFDDA_NOALLOC_BYTES) {
    set_error("ifdda_solve_device_noalloc: " + std::to_string(c->nd) + " dipoles need " +
              std::to_string(need) + " bytes of work at precision " +
              std::to_string(o->precision) + ", above the fixed " +
              std::to_string((long long)IFDDA_NOALLOC_BYTES) + " (use ifdda_solve_device)");
    return IFDDA_ERR_ARG;
  }
  int dev = 0;
  if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return IFDDA_ERR_HIP;
  if (dev < 0 || dev >= kNoAllocDevices) {
    set_error("ifdda_solve_device_noalloc: device index out of range");
    return IFDDA_ERR_ARG;
  }
  unsigned char* base = nullptr;
  int rc = noalloc_base(dev, &base);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_noalloc_mu[dev]);
  noalloc_layout(&S, n, need32, part5, base);
  KState host_state;  // polled copy of the device state (pageable)
  S.st_host = &host_state;
  S.n = n;
  S.R = 1;
  S.nblk = blocks_for(n);
  S.part_blocks = S.nblk;
  S.total = n;
  S.cap_total = n;
  S.cap_R = 1;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  // every buffer is written before it is read (init kernels, state upload),
  // so nothing a previous call left in the array leaks into this one
  rc = solve_core(CtxOp(c), &S, o, 1, (const double2*)E_inc, (double2*)FF, (double2*)FFloc, info,
                  nstop, s, t0);
  // S only borrows the array: clear its pointers so nothing frees them
  for (int i = 0; i < 14; ++i) {
    S.v64[i] = nullptr;
    S.v32[i] = nullptr;
  }
  S.part = nullptr;
  S.part5 = nullptr;
  S.st = nullptr;
  S.st_host = nullptr;
  S.act = nullptr;
  S.rr64 = nullptr;
  return rc;
}
