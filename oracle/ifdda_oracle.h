/*
 * ifdda_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, single-threaded restatement of the IF-DDA hot path
 * (the cdmlib Fortran sources of pchaumet/IF-DDA).  It is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product library never links it and never
 * falls back to it.
 *
 * Conventions (as in the reference):
 *   - complex numbers are double _Complex == Fortran DOUBLE COMPLEX;
 *   - grids are x-fastest: index i + L1*(j + L2*k), 0-based here;
 *   - dipole vectors are AoS (Ex,Ey,Ez) per dipole, 3*Nd entries;
 *   - tabdip[i + nx*(j + ny*k)] holds the 1-based dipole number or 0;
 *   - the six Green components are ordered xx, xy, xz, yy, yz, zz.
 */
#ifndef IFDDA_ORACLE_H
#define IFDDA_ORACLE_H

#include <complex.h>
#include <stdint.h>

typedef double _Complex orc_zc;

/* Free-space dyadic Green tensor, espace_libre.f:2-43.  T is 3x3 row-major. */
void orc_propa_espace_libre(double x, double y, double z, double x0, double y0,
                            double z0, double k0, orc_zc T[9]);

/* Cell-integrated tensor, propaesplibreintsim.f:2-107. */
void orc_propa_int_sim(double x, double y, double z, double x0, double y0,
                       double z0, double k0, double d, double lim, orc_zc T[9]);

/* Unnormalised in-place complex DFT, exp(sign*2*pi*i*j*k/n). */
void orc_fft1d(orc_zc *x, int n, int sign);
/* 3D transform of an x-fastest (n1,n2,n3) grid (fftsingletonz.f:21-43). */
void orc_fft3d(orc_zc *v, int n1, int n2, int n3, int sign);

/* Green tensor on the (L1,L2,L3) circulant grid, then FFT sign +1
 * (calculgreenfft.f:1-104; nquad>0: calculgreenquadfft.f:1-105).
 * G points to 6 arrays of L1*L2*L3 values (xx,xy,xz,yy,yz,zz). */
void orc_green_fft(int nx, int ny, int nz, int L1, int L2, int L3,
                   double d, double k0, int nquad, orc_zc *G);

/* FFB -= (1/M) IFFT[G . FFT[scatter(FFX)]] at dipole cells
 * (produitfftmatvect3.f:2-129).  tabdip==NULL means the dense box
 * (produitfftmatvectopt.f:3-122).  G as produced by orc_green_fft. */
void orc_matvec(const orc_zc *G, const int32_t *tabdip, int nx, int ny, int nz,
                int L1, int L2, int L3, const orc_zc *FFX, orc_zc *FFB);

/* Direct O(Nd^2) sum: FFB -= sum_{j!=i} T(r_i - r_j) FFX_j on the lattice. */
void orc_matvec_direct(const int32_t *tabdip, int nx, int ny, int nz, double d,
                       double k0, const orc_zc *FFX, orc_zc *FFB);

/* Solver selection, names as the reference's methodeit strings. */
enum { ORC_GPBICG1 = 1, ORC_QMRBICGSTAB1 = 2, ORC_QMRBICGSTAB2 = 3, ORC_TFQMR = 4 };

/* Per-iteration scalars recorded by the solvers (may be NULL). */
typedef struct {
  int cap, n;          /* capacity and count of recorded iterations       */
  orc_zc *s;           /* n x 5 scalars (method-specific, see .c)         */
  double *res;         /* n recurrence residuals                          */
} orc_trace;

typedef struct {
  double tol1;         /* true residual |A x - b| / |b| (inverserig.f:95-105) */
  int nloop;           /* iterations as the reference reports them            */
  int ncompte;         /* matvecs performed in this solve                     */
  int nstop;           /* 0 ok, 1 solver breakdown                             */
  int status, steperr; /* solver STATUS / STEPERR                              */
} orc_solve_info;

/* The driver inverserig.f:1-1090 for GPBICG1 / QMRBICGSTAB1/2 / TFQMR.
 * x0 = FF0 (E_inc).  polarisa is (nmax,3,3) column-major (Fortran).
 * Outputs FFloc = E, FF = alpha.E.  maxit as nlim (10000 in the reference). */
void orc_inverserig(const orc_zc *G, const int32_t *tabdip, int nx, int ny,
                    int nz, int L1, int L2, int L3, int nbsphere,
                    const orc_zc *polarisa, int nmax, int method, double tol,
                    int maxit, const orc_zc *FF0, orc_zc *FF, orc_zc *FFloc,
                    orc_solve_info *info, orc_trace *trace);

#endif
