/*
 * ifdda.h -- C ABI of libifdda_hip.so, the gfx950 (MI355X) implementation of
 * IF-DDA's FFT-accelerated discrete-dipole solve.
 *
 * Two layers:
 *   1. a handle API (ifdda_ctx) holding the device-resident Green tensor,
 *      dipole map, polarisability and Krylov state of one box;
 *   2. entry points with the exact argument lists of the cdmlib Fortran
 *      routines they replace (every argument by reference, Fortran default
 *      INTEGER = int32, DOUBLE COMPLEX = two doubles, plans = int64).  The
 *      iso_c_binding module if-dda_amd/fortran/ifdda_dropin.f90 exposes them
 *      to Fortran callers under the reference's own subroutine names.
 *
 * Complex arrays are interleaved (re, im) doubles in the reference's layout:
 * dipole vectors AoS (Ex,Ey,Ez) per dipole; grids x-fastest; polarisa
 * (nmax,3,3) column-major.  All functions return IFDDA_OK (0) or an error
 * code; ifdda_last_error() gives the text.  One GPU (the current HIP device).
 */
#ifndef IFDDA_H
#define IFDDA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFDDA_OK 0
#define IFDDA_ERR_ARG 1     /* invalid argument or unsupported size          */
#define IFDDA_ERR_HIP 2     /* a HIP runtime call failed                       */
#define IFDDA_ERR_STATE 3   /* Green tensor / dipoles / alpha not set up       */

/* Solver selection: the methodeit strings of inverserig.f:44,607,816,880. */
#define IFDDA_GPBICG1 1
#define IFDDA_QMRBICGSTAB1 2
#define IFDDA_QMRBICGSTAB2 3
#define IFDDA_TFQMR 4
#define IFDDA_CG 5        /* zcg.f: conjugate gradients, bilinear products   */
#define IFDDA_BICGSTAB 6  /* pimzbicgstab.f                                 */
#define IFDDA_GPBICG2 7       /* gpbicg2.f: GPBICG, BiCGSTAB step every other iteration */
#define IFDDA_GPBICGSAFE 8    /* gpbicgsafe.f                                    */
#define IFDDA_GPBICGPLUS 9    /* gpbicgplus.f                                    */
#define IFDDA_GPBICGAR 10     /* gpbicgar.f (also what 'GPBICGAR2' runs)         */
#define IFDDA_BICGSTARPLUS 11 /* bicgstarplus.f                                  */
#define IFDDA_GPBICOR 12      /* gpbicor.f                                       */
#define IFDDA_CORS 13         /* cors.f                                          */
#define IFDDA_QMRCLA 14       /* qmrpim.f (PIMZQMR), with A^T = I - alpha^T T    */

typedef struct ifdda_ctx ifdda_ctx;

/* Text of the last error on this thread. */
const char *ifdda_last_error(void);

/* ---------------------------------------------------------------------- */
/* Handle API                                                             */
/* ---------------------------------------------------------------------- */

/* Context for a box of nx*ny*nz cells.  Replaces the grid sizing and FFTW
 * plan creation of cdmlib.f:1141-1165: the padded FFT lengths are the
 * smallest 2,3,5-smooth L >= 2n per axis (L = 2n whenever 2n is smooth). */
int ifdda_create(int nx, int ny, int nz, ifdda_ctx **out);
void ifdda_destroy(ifdda_ctx *ctx);
int ifdda_fft_lengths(const ifdda_ctx *ctx, int L[3]);

/* FFT'd Green tensor built on the device in FP64 (kept also as FP32):
 * greencalculfft (calculgreenfft.f:1-104) for nquad = 0, and
 * greencalculquadfft (calculgreenquadfft.f:1-105) for nquad > 0.
 * aretecube = lattice spacing d [m], k0 = 2 pi / lambda [1/m]. */
int ifdda_green_build(ifdda_ctx *ctx, double aretecube, double k0, int nquad, double tolinit);

/* Take a caller-computed FFT'd Green tensor in the reference layout: six
 * complex128 arrays (xx,xy,xz,yy,yz,zz) of (2nx)(2ny)(2nz) points, as filled
 * by the reference's greencalculfft.  Requires L = 2n on every axis.  The
 * even/odd symmetry of the tensor is checked; the device keeps one octant. */
int ifdda_green_set(ifdda_ctx *ctx, const double *const G[6]);

/* Unfold the device tensor into six full (L1 L2 L3) complex128 arrays. */
int ifdda_green_get(ifdda_ctx *ctx, double *const G[6]);

/* Dipole map: tabdip[i + nx*(j + ny*k)] = 1-based dipole number or 0
 * (the Tabdip array of produitfftmatvect3.f:37-56).  tabdip == NULL selects
 * the dense box, dipole index = box index (produitfftmatvectopt.f:37-50). */
int ifdda_set_dipoles(ifdda_ctx *ctx, const int32_t *tabdip, int ndipole);

/* Polarisability per dipole.  layout 0: the reference's polarisa(nmax,3,3)
 * column-major complex128 (inverserig.f:70-82); layout 1: one isotropic
 * complex128 per dipole. */
int ifdda_set_polarizability(ifdda_ctx *ctx, const double *pola, int nmax, int layout);

/* FFB <- FFB - (1/M) IFFT[G . FFT[scatter(FFX)]] on host arrays of
 * 3*ndipole complex128, FP64 (produitfftmatvect3.f:2-129 semantics: FFX is
 * already alpha-multiplied). */
int ifdda_matvec(ifdda_ctx *ctx, const double *FFX, double *FFB);

/* Device-resident matvec y = b - T(alpha.x) (alpha applied if apply_alpha,
 * else y = b - T x).  prec = 64: complex128 buffers; prec = 32: complex64.
 * x, b, y are device pointers to 3*ndipole complex values; y may alias b.
 * stream: a hipStream_t or NULL for the context's stream. */
int ifdda_matvec_device(ifdda_ctx *ctx, int prec, int apply_alpha, const void *x, const void *b,
                        void *y, void *stream);

/* Solve options (the reference's methodeit / tol / nlim, plus precision). */
typedef struct {
  int method;      /* IFDDA_GPBICG1 ... IFDDA_QMRCLA                       */
  int precision;   /* 64: plain FP64; 32: FP32 matvec + FP64 refinement    */
  int x0_zero;     /* 0: x0 = E_inc as inverserig.f:36-42; 1: x0 = 0        */
  int maxit;       /* iteration cap; <= 0 means 10000 (inverserig.f:31)     */
  double tol;      /* relative residual tolerance                          */
  int check_every; /* iterations between host convergence polls (0: auto) */
} ifdda_solve_opts;

typedef struct {
  double tol1;     /* true residual |A E - E_inc| / |E_inc| (FP64 matvec)  */
  int nloop;       /* iterations, counted as the reference reports them    */
  int ncompte;     /* matvecs performed (FP64 + FP32)                      */
  int nstop;       /* 0 ok, 1 failure (infostr), -1 cancelled              */
  int status;      /* solver STATUS: 1 done, -1 breakdown                  */
  int steperr;     /* solver STEPERR                                       */
  int outer;       /* refinement sweeps (precision 32), else 0             */
  double seconds;  /* wall time of the solve                               */
  char infostr[64];
} ifdda_solve_info;

/* inverserig.f:1-1090 on host arrays: solves (I - T alpha) E = E_inc for the
 * 3*ndipole complex128 vector E.  Outputs FFloc = E (local field) and
 * FF = alpha.E (dipole moments); either may be NULL.  nstop, if not NULL,
 * is polled between iterations: another thread setting it to -1 cancels. */
int ifdda_solve(ifdda_ctx *ctx, const ifdda_solve_opts *opts, const double *E_inc, double *FF,
                double *FFloc, ifdda_solve_info *info, volatile int *nstop);

/* Same with device pointers (complex128, 3*ndipole each). */
int ifdda_solve_device(ifdda_ctx *ctx, const ifdda_solve_opts *opts, const void *E_inc, void *FF,
                       void *FFloc, ifdda_solve_info *info, volatile int *nstop, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory inside their loop (tight serving loops, code that forbids
 * hipMalloc): ifdda_solve_device with the Krylov work vectors, partial sums
 * and solver state in a fixed array of IFDDA_NOALLOC_BYTES that each GPU
 * holds from the first call until the process ends.  The call itself
 * allocates no device memory; the context must be set up (Green tensor,
 * dipoles, polarisability) beforehand.  The work of one right-hand side is
 * about 224 bytes per vector element (3 per dipole) at precision 64 and 352
 * at precision 32, so up to ~99,000 dipoles (FP64) or ~63,000 (mixed) fit;
 * larger problems return IFDDA_ERR_ARG with the sizes in ifdda_last_error().
 * Results equal ifdda_solve_device's bit for bit.  Calls in one process on
 * one GPU share the array and run one at a time (a lock); the array keeps
 * the last call's vectors, and each call overwrites what it reads. */
#define IFDDA_NOALLOC_BYTES (64ull << 20)
int ifdda_solve_device_noalloc(ifdda_ctx *ctx, const ifdda_solve_opts *opts, const void *E_inc,
                               void *FF, void *FFloc, ifdda_solve_info *info, volatile int *nstop,
                               void *stream);

/* Several right-hand sides (incidences) of one object in one batched solve:
 * E_inc, FF, FFloc hold nrhs vectors of 3*ndipole complex128, one after the
 * other; infos has nrhs entries.  Each RHS runs the same recurrence as a
 * single solve with its own scalars and convergence test; RHS that converge
 * stop costing work while the others continue.  Replaces the sequential
 * incidence loop of microsbf.f:214-281 (one inverserig call per incidence). */
int ifdda_solve_multi(ifdda_ctx *ctx, const ifdda_solve_opts *opts, int nrhs, const double *E_inc,
                      double *FF, double *FFloc, ifdda_solve_info *infos, volatile int *nstop);
int ifdda_solve_multi_device(ifdda_ctx *ctx, const ifdda_solve_opts *opts, int nrhs,
                             const void *E_inc, void *FF, void *FFloc, ifdda_solve_info *infos,
                             volatile int *nstop, void *stream);

/* Instrumentation (no counterpart in the reference, which only prints wall
 * and CPU time, cdmlib.f:1956-1959): record HIP events around each matvec
 * pass on the stream it runs on, for up to cap passes; cap = 0 disables.
 * ifdda_timing_read returns the summed milliseconds and launch counts per
 * slot [precision (0: FP64, 1: FP32) * 5 + pass (0..4 = P1..P5)] and resets. */
int ifdda_timing_enable(ifdda_ctx *ctx, int cap);
int ifdda_timing_read(ifdda_ctx *ctx, double ms_sum[10], int count[10]);
/* Which passes are bracketed (bit p = pass p+1; default 31 = all), and only
 * every every-th launch of them (default 1).  Each event pair adds a few
 * microseconds between launches, so a timed solve samples only the pass it
 * reports. */
int ifdda_timing_passes(ifdda_ctx *ctx, int mask, int every);

/* Kernel selection (no counterpart in the reference): on = 1 (default)
 * runs the z pass with the kernels compiled for a fixed FFT length where the
 * library has one for this grid, 0 always runs the generic kernels.  Both
 * compute the same products; tests compare them. */
int ifdda_use_specialized(ifdda_ctx *ctx, int on);

/* GPBICG inner products fused into its matvecs (no counterpart in the
 * reference, whose gpbicg.f forms each product in a loop of its own): the
 * products of gpbicg.f:174-181 and :222-253 are summed by the pass that
 * gathers the matvec's result, as it writes it, instead of by separate
 * kernels that read the vectors again.  mask bit 0: FP64 solves, bit 1: FP32
 * (the inner solves of mixed precision; the default is 2); bit 2 also moves
 * the update p = r + beta (p - u) (gpbicg.f:161-169) into the matvec's
 * scatter pass, in place of its own kernel (off by default: slower on the
 * 100^3 cube, where the scatter reads r, u and p through the dipole map
 * less efficiently than the separate kernel streams them).  Bits 3 (FP32,
 * on by default: mask 10) and 4 (FP64) select the look-ahead form of the
 * fused loop: the second matvec's gathering pass also sums <r0*,t>, <r0*,y>
 * and <r0*,At>, which give <r0*, r_new> and the next beta (gpbicg.f:300-316)
 * before the u/z/x/r update, and that update writes the next p itself
 * (25 vector streams per iteration instead of 29; bit 2 is then ignored).
 * The vectors are the same; the products are summed in another fixed order,
 * so fused and unfused trajectories agree to rounding.  Single-context
 * solves with the specialized kernels on (ifdda_use_specialized) only; packs
 * always run the separate kernels. */
int ifdda_fuse_krylov(ifdda_ctx *ctx, int mask);

/* The z pass (FFT-z . G . IFFT-z, produitfftmatvect3.f:62-107) as the
 * register-staged kernel (no counterpart in the reference, whose FFTs run
 * per component through fftsingletonz): the length-L3 transforms split into
 * 8-point steps across threads and L3/8-point steps inside one thread's
 * registers, for L3 = 128, 192, 200 at the tile widths it is built for.
 * mask bit 0: FP64 matvecs, bit 1: FP32.  Same products as the other z
 * kernels to rounding; single-context matvecs with the specialized kernels
 * on only. */
int ifdda_zpass_registers(ifdda_ctx *ctx, int mask);
/* *has = 1 when this grid's z pass has a compile-time kernel for prec 64/32. */
int ifdda_specialized_z(ifdda_ctx *ctx, int prec, int *has);

/* Tile-width autotuner.  It replaces the reference's untimed FFT planning
 * (FFTW_ESTIMATE plans, cdmlib.f:1153-1158) with a measured choice of the
 * matvec passes' tile widths (box lines per workgroup) in launches of one
 * right-hand side (solves and matvecs of one vector; batched right-hand
 * sides and packs keep the widths of fixed LDS budgets).  ifdda_autotune
 * times P1..P5 of this context's grid at every candidate width (the widths
 * that LDS budgets of 6..64 KB give, plus the current one), `reps` timed
 * matvecs each (0: 5), for prec 64, 32 or 0 (both).  It keeps the fastest
 * width per pass in the context and records it in a process-wide table
 * keyed by box size and precision; contexts created later for the same box
 * start from that table.  widths (optional, 10 ints) receives
 * the widths now in use: P1..P5 in FP64, then P1..P5 in FP32.  Needs a Green
 * tensor and dipoles; synchronises the context's stream.  A width changes
 * how lines are grouped, not the arithmetic of a line: the generic kernels
 * give bit-identical results at every width, and a z-pass width without a
 * compile-time kernel runs the generic kernel (ifdda_use_specialized). */
int ifdda_autotune(ifdda_ctx *ctx, int prec, int reps, int widths[10]);
/* The tile widths of this context, laid out as in ifdda_autotune. */
int ifdda_tile_widths(ifdda_ctx *ctx, int widths[10]);
/* Read the table entry of a box (prec 64 or 32) into widths[5] (P1..P5):
 * returns 1 when there is one, 0 when not, or an error code.  Also usable
 * to save the table between runs. */
int ifdda_tune_get(int nx, int ny, int nz, int prec, int widths[5]);
/* Write the table entry of a box, e.g. one saved by a previous run;
 * IFDDA_ERR_ARG when a width does not fit that box.  Contexts created
 * afterwards use it. */
int ifdda_tune_set(int nx, int ny, int nz, int prec, const int widths[5]);

/* Per-iteration scalars of the last solve: rows of six complex values
 * (GPBICG: alpha, beta, eta, dzeta, R0RN; QMRBICGSTAB: alpha, beta, omega,
 * tau, eta; TFQMR: alpha, beta, rho, tau, eta; then the residual), recorded
 * when enabled with a row capacity cap > 0. */
int ifdda_solve_trace_enable(ifdda_ctx *ctx, int cap);
int ifdda_solve_trace(ifdda_ctx *ctx, int cap, double *out, int *nrows);

/* ---------------------------------------------------------------------- */
/* Reference-shaped entry points (all arguments by reference).            */
/* ---------------------------------------------------------------------- */

/* produitfftmatvect3.f:2-5 */
void ifdda_produitfftmatvect3(const double *Gxx, const double *Gxy, const double *Gxz,
                              const double *Gyy, const double *Gyz, const double *Gzz, double *vectx,
                              double *vecty, double *vectz, const int32_t *Tabdip,
                              const int32_t *ntotalm, const int32_t *ntotal, const int32_t *nmax,
                              const int32_t *ndipole, const int32_t *nxm, const int32_t *nym,
                              const int32_t *nzm, const int32_t *nx, const int32_t *ny,
                              const int32_t *nz, const int32_t *nx2, const int32_t *ny2,
                              const int32_t *nxy2, const int32_t *nz2, const double *FFX,
                              double *FFB, const int64_t *planb, const int64_t *planf);

/* produitfftmatvectopt.f:3-6 (note the plan order planf, planb) */
void ifdda_produitfftmatvectopt(const double *Gxx, const double *Gxy, const double *Gxz,
                                const double *Gyy, const double *Gyz, const double *Gzz,
                                double *vectx, double *vecty, double *vectz, const int32_t *ntotalm,
                                const int32_t *ntotal, const int32_t *nmax, const int32_t *nxm,
                                const int32_t *nym, const int32_t *nzm, const int32_t *nx,
                                const int32_t *ny, const int32_t *nz, const int32_t *nx2,
                                const int32_t *ny2, const int32_t *nxy2, const int32_t *nz2,
                                const double *FFX, double *FFB, const int64_t *planf,
                                const int64_t *planb);

/* calculgreenfft.f:1-3 */
void ifdda_greencalculfft(const int32_t *nx, const int32_t *ny, const int32_t *nz,
                          const int32_t *nx2, const int32_t *ny2, const int32_t *nz2,
                          const int32_t *nxy2, const int32_t *nxm, const int32_t *nym,
                          const int32_t *nzm, const double *aretecube, const double *k0,
                          double *Gxx, double *Gxy, double *Gxz, double *Gyy, double *Gyz,
                          double *Gzz, const int64_t *planb);

/* calculgreenquadfft.f:1-3 */
void ifdda_greencalculquadfft(const int32_t *nx, const int32_t *ny, const int32_t *nz,
                              const int32_t *nx2, const int32_t *ny2, const int32_t *nz2,
                              const int32_t *nxy2, const int32_t *nxm, const int32_t *nym,
                              const int32_t *nzm, const int32_t *nquad, const double *tolinit,
                              const double *aretecube, const double *k0, double *Gxx, double *Gxy,
                              double *Gxz, double *Gyy, double *Gyz, double *Gzz,
                              const int64_t *planb);

/* inverserig.f:1-6.  methodeit / infostr are passed with explicit lengths
 * (the hidden Fortran character lengths, made explicit by the shim). */
void ifdda_inverserig(const double *Gxx, const double *Gxy, const double *Gxz, const double *Gyy,
                      const double *Gyz, const double *Gzz, double *vectx, double *vecty,
                      double *vectz, const int32_t *Tabdip, const int32_t *ntotalm,
                      const int32_t *ntotal, const int32_t *ldabi, const int32_t *nlar,
                      const int32_t *nmax, const int32_t *ndipole, const int32_t *nxm,
                      const int32_t *nym, const int32_t *nzm, const int32_t *nx, const int32_t *ny,
                      const int32_t *nz, const int32_t *nx2, const int32_t *ny2,
                      const int32_t *nxy2, const int32_t *nz2, const int32_t *nbsphere,
                      const int32_t *nbsphere3, double *XI, double *XR, double *wrk, double *FF,
                      const double *FF0, double *FFloc, const double *polarisa,
                      const char *methodeit, const double *tol, double *tol1, int32_t *nloop,
                      int32_t *ncompte, const int64_t *planf, const int64_t *planb,
                      volatile int32_t *nstop, char *infostr, size_t len_methodeit,
                      size_t len_infostr);

/* inverserigopt.f:1-6: the dense-box driver (no Tabdip). */
void ifdda_inverserigopt(const double *Gxx, const double *Gxy, const double *Gxz,
                         const double *Gyy, const double *Gyz, const double *Gzz, double *vectx,
                         double *vecty, double *vectz, const int32_t *ntotalm,
                         const int32_t *ntotal, const int32_t *ldabi, const int32_t *nlar,
                         const int32_t *nmax, const int32_t *nxm, const int32_t *nym,
                         const int32_t *nzm, const int32_t *nx, const int32_t *ny,
                         const int32_t *nz, const int32_t *nx2, const int32_t *ny2,
                         const int32_t *nxy2, const int32_t *nz2, const int32_t *nbsphere,
                         const int32_t *nbsphere3, double *XI, double *XR, double *wrk,
                         double *FF, const double *FF0, double *FFloc, const double *polarisa,
                         const char *methodeit, const double *tol, double *tol1, int32_t *nloop,
                         int32_t *ncompte, const int64_t *planf, const int64_t *planb,
                         volatile int32_t *nstop, char *infostr, size_t len_methodeit,
                         size_t len_infostr);

/* ---------------------------------------------------------------------------
 * Outputs from solved fields (device vectors, complex128; R right-hand sides
 * at offsets r * 3 * nd as in ifdda_solve_multi_device).
 * ------------------------------------------------------------------------- */

/* Cross sections of cdmlib.f:2725-2753: Cext = 4 pi k0 sum Im(E0* . p) / I0,
 * Cabs = 4 pi k0 sum [Im(p . Eloc*) - 2/3 k0^3 |p|^2] / I0, Csca = Cext - Cabs
 * (Cabs then set to 0 when <= 1e-12 Cext).  FF0 = incident field at the
 * dipoles, FF = dipole moments, FFloc = local field; I0[r] = |E0|^2 of RHS r.
 * out[3 r + {0,1,2}] = Cext, Cabs, Csca (host). */
int ifdda_cross_sections_device(ifdda_ctx *ctx, int nrhs, const void *FF0, const void *FF,
                                const void *FFloc, double k0, const double *I0, double *out,
                                void *stream);

/* Reference-shaped entries of the optical-force FFTs (the Fortran drop-in
 * binds calculforcefft_, calculforcefftopt_ and derivativefield2_ to them):
 * calculforcefft.f:1-2 / calculforcefftopt.f:1-2 keep the moments FF (dipole
 * order with Tabdip, box order without) for the box; derivativefield2.f:1-4
 * then writes FF0 (3 nx ny nz complex, box order) = the x_test derivative of
 * their field at every box cell.  vect* and DFFTTENSOR* are accepted and left
 * alone (scratch in the reference). */
void ifdda_calculforcefft(const int32_t *nx, const int32_t *ny, const int32_t *nz,
                          const int32_t *nx2, const int32_t *ny2, const int32_t *nz2,
                          const int32_t *nxy2, const int32_t *nxm, const int32_t *nym,
                          const int32_t *nzm, const int32_t *Tabdip, double *vectx, double *vecty,
                          double *vectz, const double *FF, const int64_t *planb);
void ifdda_calculforcefftopt(const int32_t *nx, const int32_t *ny, const int32_t *nz,
                             const int32_t *nx2, const int32_t *ny2, const int32_t *nz2,
                             const int32_t *nxy2, const int32_t *nxm, const int32_t *nym,
                             const int32_t *nzm, double *vectx, double *vecty, double *vectz,
                             const double *FF, const int64_t *planb);
void ifdda_derivativefield2(const double *aretecube, const double *k0, const int32_t *nx,
                            const int32_t *ny, const int32_t *nz, const int32_t *nx2,
                            const int32_t *ny2, const int32_t *nz2, const int32_t *nxy2,
                            const int32_t *ntotal, const int32_t *ntotalm, const int32_t *nmax,
                            double *Dxx, double *Dxy, double *Dxz, double *Dyy, double *Dyz,
                            double *Dzz, double *vectx, double *vecty, double *vectz,
                            const int32_t *test, double *FF0, const int64_t *planb,
                            const int64_t *planf);

/* Optical force and torque on every dipole (cdmlib.f:3102-3560, nforce /
 * ntorque) for R solved right-hand sides in the layout of
 * ifdda_solve_multi_device (device complex128: E_inc, FF = alpha E, FFloc =
 * E), each a plane or circular wave of wave vector kvec[3r..3r+2] (1/m, host):
 * the incident derivative is i k E_inc (ondeplaned.f, ondecirced.f), the
 * scattered one the FFT convolution of derivativefield2.f with d T / dx
 * (built on first use for this d, k0 and kept with the context).
 * force[3r..3r+2], torque[3r..3r+2] (host): totals in N and N m, the torque
 * about the mean dipole position; fdip, tdip: optional device double arrays
 * of R * 3 * nd per-dipole values (tdip needs fdip), or NULL. */
int ifdda_force_device(ifdda_ctx *ctx, int nrhs, const void *E_inc, const void *FF,
                       const void *FFloc, const double *kvec, double d, double k0, double *force,
                       double *torque, void *fdip, void *tdip, void *stream);

/* Far field by the quick 2D-FFT method of diffractefft2d.f:1-183 for the
 * dipole moments FF (device, dipole order; the context's dipole map places
 * them in the box): lattice spacing d, (x0, y0, z0) = centre of box cell
 * (1,1,1), FFT size nfft2d.  Epos / Eneg (device): per RHS 3 x (2 imax + 1)^2
 * complex128, [c][jj][ii] with kx = (ii - imax) deltak, ky = (jj - imax)
 * deltak, the field radiated into z > 0 / z < 0 (zero outside the light
 * cone).  *imax, *deltak as the reference; its errors ('nfft2d too large',
 * 'In FFT Poynting nfft2d too small', 'In FFT diffract nfft2d too small')
 * return IFDDA_ERR_ARG with that text. */
int ifdda_farfield_device(ifdda_ctx *ctx, int nrhs, const void *FF, int nfft2d, double k0, double d,
                          double x0, double y0, double z0, void *Epos, void *Eneg, int *imax,
                          double *deltak, void *stream);
/* imax and deltak of that far field (to size Epos / Eneg before the call). */
int ifdda_farfield_size(int nfft2d, double k0, double d, int *imax, double *deltak);

/* Csca and asymmetry g (computegcfft2d.f:1-54) from that far field; theta,
 * phi = incidence direction of RHS r in degrees; out[2 r + {0,1}] = Csca, g. */
int ifdda_computegcfft2d_device(ifdda_ctx *ctx, int nrhs, const void *Epos, const void *Eneg,
                                int imax, double deltak, double k0, const double *I0,
                                const double *theta, const double *phi, double *out,
                                void *stream);

/* Reference-shaped (the drop-in binds these to diffractefft2d_ /
 * computegcfft2d_).  diffractefft2d.f:1-3: FF box-ordered (3 nx ny nz, the
 * dense box the reference uses for this method, cdmlib.f:630), xs/ys/zs per
 * cell; fills tabfft2 and the caller's (nfft2d, nfft2d, 3) Ediffkzpos /
 * Ediffkzneg; nstop = 1 and infostr on the reference's errors.  Eloinx/y/z
 * and the plans are accepted and unused. */
void ifdda_diffractefft2d(const int32_t *nx, const int32_t *ny, const int32_t *nz,
                          const int32_t *nxm, const int32_t *nym, const int32_t *nzm,
                          const int32_t *nfft2d, int32_t *tabfft2, const double *k0,
                          const double *xs, const double *ys, const double *zs,
                          const double *aretecube, double *Eloinx, double *Eloiny,
                          double *Eloinz, const double *FF, int32_t *imax, double *deltakx,
                          double *deltaky, double *Ediffkzpos, double *Ediffkzneg,
                          const int64_t *plan2f, const int64_t *plan2b, int32_t *nstop,
                          char *infostr, size_t len_infostr);
/* computegcfft2d.f:1-3 (argument order of the reference). */
void ifdda_computegcfft2d(const int32_t *imax, const double *deltakx, const double *deltaky,
                          const double *k0, const double *I0, const double *theta,
                          const double *phi, const int32_t *nfft2d, const double *Ediffkzpos,
                          const double *Ediffkzneg, double *gasym, double *Cscai);

/* ---------------------------------------------------------------------------
 * Brightfield microscope (microsbf.f:1-965): the condenser's incidences
 * solved together, lens, image.  Quick-lens path (nquicklens = 1) by
 * default; direct_lens = 1 selects the nquicklens = 0 path.
 * ------------------------------------------------------------------------- */
typedef struct {
  int nfft2d;         /* FFT size of the pupil / image grid (even)             */
  double numaper;     /* objective NA, 0 < NA <= 1                             */
  double numaperinc;  /* condenser NA, 0 < NA < 1                              */
  double gross;       /* magnification                                         */
  double zlens;       /* position of the focal plane [m]                       */
  int nside;          /* 1: objective on the z > 0 side (microsbf.f,
                         microsdf.f); -1: z < 0, the reflection microscopes
                         microsbfneg / microsdfneg of microssansinc.f
                         (scattered light only: images 3..5 equal 0..2)    */
  int npolainc;       /* 0: both polarisations (pp = 0 then 1), 1: pp = 0, 2: pp = 1 */
  double E0[2];       /* amplitude of each illumination (re, im), as microsbf.f:205-209 */
  double k0;          /* wave number [1/m]                                     */
  double d;           /* lattice spacing [m]                                   */
  double origin[3];   /* centre of box cell (1,1,1) [m]                         */
  int max_batch;      /* incidences per multi-RHS solve (0: all at once)        */
  int direct_lens;    /* 0: quick lens, the pupil field by FFT
                         (diffractefft2dlens.f; nquicklens = 1).  1: the
                         nquicklens = 0 path: pupil step and condenser step
                         of microsbf.f:139-157 and the pupil field as the
                         direct sum over dipoles of microsbf.f:629-668
                         (microssansinc.f:600-640 for nside = -1)          */
  int ninc_user;      /* > 0 (brightfield only): illuminate with these plane
                         waves instead of the condenser lattice of
                         microsbf.f:90-172 -- kxy_user holds (kx, ky, pp) per
                         incidence (host); the direct wave goes to its
                         nearest pupil bin                                   */
  const double *kxy_user;
} ifdda_microscope_opts;

/* The incidences microsbf.f:90-172 takes (polarisation outer, idelta,
 * jdelta inner): *ninc, and for the first cap of them (kx, ky, pp) in kxy;
 * *deltak = step of the condenser lattice. */
int ifdda_microscope_incidences(ifdda_ctx *ctx, const ifdda_microscope_opts *m, int *ninc,
                                double *kxy, int cap, double *deltak);

/* microsbf.f: for every incidence of the condenser (plane waves of
 * ondeplanekxky.f), solve (opts: method, tolerance, precision; all
 * incidences of a batch in one multi-RHS solve), take the pupil field
 * (diffractefft2dlens.f), add the incident wave at its pupil bin, apply the
 * lens rotation and focus, form the image (fouriertoimage.f) and sum |E|^2
 * over incidences.  images (host, 6 * nfft2d^2 doubles): square roots of
 * the sums for the scattered field x, y, z and the total field x, y, z,
 * [component][y][x] (x fastest) as the reference's Eimagex ... Eimageincz.
 * infos[i] (optional, first infos_cap incidences) report each solve.  The
 * context's Green tensor, dipoles and polarisability must be set. */
int ifdda_microsbf_device(ifdda_ctx *ctx, const ifdda_solve_opts *opts,
                          const ifdda_microscope_opts *m, double *images, int *ninc,
                          ifdda_solve_info *infos, int infos_cap, volatile int *nstop);

/* Darkfield and phase microscope (microsdf.f:1-915; quick lens, or
 * direct_lens for its nquicklens = 0 path): as
 * ifdda_microsbf_device with the incidences of microsdf.f:159-221 -- a ring
 * of ideltam = max(int(2 pi k0 / (deltak NA_inc)) + 1, 8) plane waves on the
 * condenser cone (kx, ky) = k0 NA_inc (cos phi, sin phi), both
 * polarisations (npolainc is ignored, as there) -- and the direct wave added
 * at its bin of the pupil grid as i E / (dkx dky) on every component
 * (microsdf.f:668-717).  With NA_inc above the objective's NA the direct
 * wave stays outside the pupil (dark field). */
int ifdda_microsdf_device(ifdda_ctx *ctx, const ifdda_solve_opts *opts,
                          const ifdda_microscope_opts *m, double *images, int *ninc,
                          ifdda_solve_info *infos, int infos_cap, volatile int *nstop);
/* The incidences of ifdda_microsdf_device (polarisation outer, phi inner);
 * *deltak as ifdda_microscope_incidences. */
int ifdda_microsdf_incidences(ifdda_ctx *ctx, const ifdda_microscope_opts *m, int *ninc,
                              double *kxy, int cap, double *deltak);

/* ---------------------------------------------------------------------------
 * Runner: many independent solves at once (no counterpart in the reference,
 * whose callers loop over objects, wavelengths and requests one cdmlib call
 * at a time).  A pool of worker threads, each with its own stream, takes
 * jobs in submission order: a free worker takes the head job and every
 * queued job with the same solve options and a box within a factor 2 in
 * cells that fits the pack (at most
 * max_members right-hand sides, 64 by default) and the device budget
 * mem_budget (bytes) next to the packs already running (a job larger than
 * the budget runs alone), and solves the group as one pack (as
 * ifdda_solve_batch).  Results do not depend on what else runs.
 * ------------------------------------------------------------------------- */
typedef struct {
  int nx, ny, nz;          /* box                                             */
  double d, k0;            /* lattice spacing [m], wave number [1/m]          */
  int nquad;               /* Green near-field quadrature, as ifdda_green_build */
  double tolinit;
  const int32_t *tabdip;   /* nx*ny*nz box map (1-based dipole or 0), NULL: dense box */
  int nd;                  /* dipoles                                         */
  const double *alpha;     /* nd complex polarisabilities (isotropic)         */
  int nrhs;                /* incidences                                      */
  const double *E_inc;     /* nrhs * 3 nd complex, RHS r at r * 3 nd          */
  ifdda_solve_opts opts;
  /* Optional, when E_inc is NULL: the nrhs incident plane waves formed on
   * the device (as ifdda_incident_device) */
  int wave_kind;           /* IFDDA_WAVE_*                                    */
  const double *waves;     /* nrhs * 3 parameters of that kind                */
  double E0[2];            /* amplitude (re, im)                              */
  double origin[3];        /* centre of box cell (1,1,1) [m]                  */
  /* Optional, when alpha is NULL: the polarisability from the permittivity
   * (as ifdda_polarizability_device) */
  const double *eps;       /* nd complex                                      */
  int pola;                /* IFDDA_POLA_*                                    */
  double eps0;             /* background permittivity (0 means 1)             */
} ifdda_job;

/* ---------------------------------------------------------------------------
 * Incident waves and polarisabilities on the device (the step before the
 * solve; the reference forms them per dipole in host loops).
 * ------------------------------------------------------------------------- */
enum {
  IFDDA_WAVE_PLANE = 1,    /* ondeplane.f: params (theta, phi [deg], pp), ss = 1 - pp  */
  IFDDA_WAVE_CIRCULAR = 2, /* ondecirce.f: params (theta, phi [deg], helicity sign)     */
  IFDDA_WAVE_KXKY = 3      /* ondeplanekxky.f: params (kx, ky [1/m], pp)              */
};
enum { IFDDA_POLA_CM = 0, IFDDA_POLA_RR = 1, IFDDA_POLA_GB = 2, IFDDA_POLA_LA = 3, IFDDA_POLA_LR = 4 };

/* R incident plane waves of amplitude E0 (re, im) at the context's dipoles,
 * whose cell (i, j, k) sits at origin + d (i, j, k) (origin: centre of cell
 * (1,1,1)): E (device complex128, R * 3 nd, RHS r at r * 3 nd, the layout of
 * ifdda_solve_multi_device); params R * 3 as the kind says.  Replaces the
 * loops over dipoles around ondeplane / ondecirce / ondeplanekxky
 * (cdmlib.f incident-field set-up, microsbf.f:226-230). */
int ifdda_incident_device(ifdda_ctx *ctx, int kind, int R, const double *params, const double E0[2],
                          double k0, double d, const double origin[3], void *E, void *stream);
/* The context's polarisability from the permittivity of each dipole's cell
 * (eps: host, nd complex): poladiff (polarisabilite.f:1-62, dddis = 1, cell
 * of side d) with method CM, RR (= PS), GB, LA or LR, on the device. */
int ifdda_polarizability_device(ifdda_ctx *ctx, const double *eps, double d, double k0, double eps0,
                                int method);

/* Independent problems solved together (no counterpart in the reference):
 * every right-hand side of every job shares the matvec launches (packed
 * launches over the members' own boxes, Green tensors, maps and
 * polarisabilities) and one lockstep Krylov solve with per-RHS state and
 * convergence, so each result equals the job's own solve.  All jobs use
 * jobs[0].opts.  FFloc[j] / FF[j] (nrhs * 3 nd complex, or FFloc / FF NULL)
 * and infos[j] (nrhs entries) receive job j's results.  Right-hand sides
 * that have converged drop out of every launch.  The workspace of the
 * current device (FFT tables, the Green tensors of recent (box, d, k0)
 * keys, grow-only buffers) is kept between calls; calls on one device run
 * one at a time. */
int ifdda_solve_batch(int njobs, const ifdda_job *jobs, double *const *FFloc, double *const *FF,
                      ifdda_solve_info *const *infos, volatile int *nstop);
/* Free the workspaces ifdda_solve_batch keeps (every device). */
int ifdda_solve_batch_release(void);

typedef struct ifdda_runner ifdda_runner;
int ifdda_runner_create(int nworkers, double mem_budget, ifdda_runner **out);
/* Copies the job's arrays; returns at once with a ticket. */
int ifdda_runner_submit(ifdda_runner *r, const ifdda_job *job, long long *ticket);
/* Blocks until the job has run; FFloc / FF (nrhs * 3 nd complex), infos
 * (nrhs) may be NULL; times[3] (optional) = seconds queued, running, total. */
int ifdda_runner_wait(ifdda_runner *r, long long ticket, double *FFloc, double *FF,
                      ifdda_solve_info *infos, double *times);
/* Right-hand sides per pack (1: every job alone, apart from its own RHS). */
int ifdda_runner_pack_limit(ifdda_runner *r, int max_members);
/* Runs what is queued, then stops the workers. */
int ifdda_runner_destroy(ifdda_runner *r);

#ifdef __cplusplus
}
#endif

#endif /* IFDDA_H */
