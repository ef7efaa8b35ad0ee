! ifdda_dropin.f90 -- iso_c_binding drop-in for the cdmlib hot path.
!
! External subroutines with the names and argument lists of the cdmlib
! routines they replace, so that existing callers (cdmlib.f, microsbf.f,
! microsdf.f, microssansinc.f) link against libifdda_dropin.so unchanged:
!
!   produitfftmatvect3   produitfftmatvect3.f:2-5
!   produitfftmatvectopt produitfftmatvectopt.f:3-6
!   greencalculfft       calculgreenfft.f:1-3
!   greencalculquadfft   calculgreenquadfft.f:1-3
!   inverserig           inverserig.f:1-6
!   inverserigopt        inverserigopt.f:1-6
!   calculforcefft       calculforcefft.f:1-2     (optical force, cdmlib.f:3159)
!   calculforcefftopt    calculforcefftopt.f:1-2
!   derivativefield2     derivativefield2.f:1-4
!
! Each forwards its arguments by reference to the C entry point of the same
! name prefixed with ifdda_ (include/ifdda.h), which runs on the GPU.  The
! hidden character lengths of methodeit / infostr are passed explicitly.

module ifdda_capi
  use iso_c_binding
  implicit none
  interface
    subroutine c_matvect3(gxx, gxy, gxz, gyy, gyz, gzz, vx, vy, vz, tabdip, ntotalm, ntotal, &
         nmax, ndipole, nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, nz2, ffx, ffb, planb, planf) &
         bind(C, name="ifdda_produitfftmatvect3")
      import :: c_double_complex, c_int32_t, c_int64_t
      complex(c_double_complex) :: gxx(*), gxy(*), gxz(*), gyy(*), gyz(*), gzz(*)
      complex(c_double_complex) :: vx(*), vy(*), vz(*), ffx(*), ffb(*)
      integer(c_int32_t) :: tabdip(*), ntotalm, ntotal, nmax, ndipole, nxm, nym, nzm, nx, ny, &
           nz, nx2, ny2, nxy2, nz2
      integer(c_int64_t) :: planb, planf
    end subroutine

    subroutine c_matvectopt(gxx, gxy, gxz, gyy, gyz, gzz, vx, vy, vz, ntotalm, ntotal, nmax, &
         nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, nz2, ffx, ffb, planf, planb) &
         bind(C, name="ifdda_produitfftmatvectopt")
      import :: c_double_complex, c_int32_t, c_int64_t
      complex(c_double_complex) :: gxx(*), gxy(*), gxz(*), gyy(*), gyz(*), gzz(*)
      complex(c_double_complex) :: vx(*), vy(*), vz(*), ffx(*), ffb(*)
      integer(c_int32_t) :: ntotalm, ntotal, nmax, nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, &
           nz2
      integer(c_int64_t) :: planf, planb
    end subroutine

    subroutine c_forcefft(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, tabdip, vx, vy, vz, &
         ff, planb) bind(C, name="ifdda_calculforcefft")
      import :: c_double_complex, c_int32_t, c_int64_t
      integer(c_int32_t) :: nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, tabdip(*)
      complex(c_double_complex) :: vx(*), vy(*), vz(*), ff(*)
      integer(c_int64_t) :: planb
    end subroutine

    subroutine c_forcefftopt(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, vx, vy, vz, ff, &
         planb) bind(C, name="ifdda_calculforcefftopt")
      import :: c_double_complex, c_int32_t, c_int64_t
      integer(c_int32_t) :: nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm
      complex(c_double_complex) :: vx(*), vy(*), vz(*), ff(*)
      integer(c_int64_t) :: planb
    end subroutine

    subroutine c_dfield(d, k0, nx, ny, nz, nx2, ny2, nz2, nxy2, ntotal, ntotalm, nmax, dxx, dxy, &
         dxz, dyy, dyz, dzz, vx, vy, vz, test, ff0, planb, planf) &
         bind(C, name="ifdda_derivativefield2")
      import :: c_double, c_double_complex, c_int32_t, c_int64_t
      real(c_double) :: d, k0
      integer(c_int32_t) :: nx, ny, nz, nx2, ny2, nz2, nxy2, ntotal, ntotalm, nmax, test
      complex(c_double_complex) :: dxx(*), dxy(*), dxz(*), dyy(*), dyz(*), dzz(*)
      complex(c_double_complex) :: vx(*), vy(*), vz(*), ff0(*)
      integer(c_int64_t) :: planb, planf
    end subroutine

    subroutine c_green(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, d, k0, gxx, gxy, gxz, &
         gyy, gyz, gzz, planb) bind(C, name="ifdda_greencalculfft")
      import :: c_double, c_double_complex, c_int32_t, c_int64_t
      integer(c_int32_t) :: nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm
      real(c_double) :: d, k0
      complex(c_double_complex) :: gxx(*), gxy(*), gxz(*), gyy(*), gyz(*), gzz(*)
      integer(c_int64_t) :: planb
    end subroutine

    subroutine c_greenquad(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, nquad, tolinit, d, &
         k0, gxx, gxy, gxz, gyy, gyz, gzz, planb) bind(C, name="ifdda_greencalculquadfft")
      import :: c_double, c_double_complex, c_int32_t, c_int64_t
      integer(c_int32_t) :: nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, nquad
      real(c_double) :: tolinit, d, k0
      complex(c_double_complex) :: gxx(*), gxy(*), gxz(*), gyy(*), gyz(*), gzz(*)
      integer(c_int64_t) :: planb
    end subroutine

    subroutine c_inverserig(gxx, gxy, gxz, gyy, gyz, gzz, vx, vy, vz, tabdip, ntotalm, ntotal, &
         ldabi, nlar, nmax, ndipole, nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, nz2, nbsphere, &
         nbsphere3, xi, xr, wrk, ff, ff0, ffloc, polarisa, methodeit, tol, tol1, nloop, &
         ncompte, planf, planb, nstop, infostr, lmeth, linfo) bind(C, name="ifdda_inverserig")
      import :: c_double, c_double_complex, c_int32_t, c_int64_t, c_char, c_size_t
      complex(c_double_complex) :: gxx(*), gxy(*), gxz(*), gyy(*), gyz(*), gzz(*)
      complex(c_double_complex) :: vx(*), vy(*), vz(*), xi(*), xr(*), wrk(*), ff(*), ff0(*), &
           ffloc(*), polarisa(*)
      integer(c_int32_t) :: tabdip(*), ntotalm, ntotal, ldabi, nlar, nmax, ndipole, nxm, nym, &
           nzm, nx, ny, nz, nx2, ny2, nxy2, nz2, nbsphere, nbsphere3, nloop, ncompte, nstop
      character(kind=c_char) :: methodeit(*), infostr(*)
      real(c_double) :: tol, tol1
      integer(c_int64_t) :: planf, planb
      integer(c_size_t), value :: lmeth, linfo
    end subroutine

    subroutine c_inverserigopt(gxx, gxy, gxz, gyy, gyz, gzz, vx, vy, vz, ntotalm, ntotal, &
         ldabi, nlar, nmax, nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, nz2, nbsphere, &
         nbsphere3, xi, xr, wrk, ff, ff0, ffloc, polarisa, methodeit, tol, tol1, nloop, &
         ncompte, planf, planb, nstop, infostr, lmeth, linfo) bind(C, name="ifdda_inverserigopt")
      import :: c_double, c_double_complex, c_int32_t, c_int64_t, c_char, c_size_t
      complex(c_double_complex) :: gxx(*), gxy(*), gxz(*), gyy(*), gyz(*), gzz(*)
      complex(c_double_complex) :: vx(*), vy(*), vz(*), xi(*), xr(*), wrk(*), ff(*), ff0(*), &
           ffloc(*), polarisa(*)
      integer(c_int32_t) :: ntotalm, ntotal, ldabi, nlar, nmax, nxm, nym, nzm, nx, ny, nz, &
           nx2, ny2, nxy2, nz2, nbsphere, nbsphere3, nloop, ncompte, nstop
      character(kind=c_char) :: methodeit(*), infostr(*)
      real(c_double) :: tol, tol1
      integer(c_int64_t) :: planf, planb
      integer(c_size_t), value :: lmeth, linfo
    end subroutine
    subroutine c_diffractefft2d(nx, ny, nz, nxm, nym, nzm, nfft2d, tabfft2, k0, xs, ys, zs, &
         aretecube, eloinx, eloiny, eloinz, ff, imax, deltakx, deltaky, ediffkzpos, ediffkzneg, &
         plan2f, plan2b, nstop, infostr, linfo) bind(C, name="ifdda_diffractefft2d")
      import :: c_double, c_double_complex, c_int32_t, c_int64_t, c_char, c_size_t
      integer(c_int32_t) :: nx, ny, nz, nxm, nym, nzm, nfft2d, tabfft2(*), imax, nstop
      real(c_double) :: k0, xs(*), ys(*), zs(*), aretecube, deltakx, deltaky
      complex(c_double_complex) :: eloinx(*), eloiny(*), eloinz(*), ff(*), ediffkzpos(*), &
           ediffkzneg(*)
      integer(c_int64_t) :: plan2f, plan2b
      character(kind=c_char) :: infostr(*)
      integer(c_size_t), value :: linfo
    end subroutine
    subroutine c_computegcfft2d(imax, deltakx, deltaky, k0, i0, theta, phi, nfft2d, ediffkzpos, &
         ediffkzneg, gasym, cscai) bind(C, name="ifdda_computegcfft2d")
      import :: c_double, c_double_complex, c_int32_t
      integer(c_int32_t) :: imax, nfft2d
      real(c_double) :: deltakx, deltaky, k0, i0, theta, phi, gasym, cscai
      complex(c_double_complex) :: ediffkzpos(*), ediffkzneg(*)
    end subroutine
  end interface
end module ifdda_capi

subroutine produitfftmatvect3(FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, &
     FFTTENSORyz, FFTTENSORzz, vectx, vecty, vectz, Tabdip, ntotalm, ntotal, nmax, ndipole, &
     nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, nz2, FFX, FFB, planb, planf)
  use ifdda_capi
  implicit none
  integer :: ntotalm, ntotal, nmax, ndipole, nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, nz2
  integer :: Tabdip(nxm*nym*nzm)
  double complex :: FFTTENSORxx(ntotalm), FFTTENSORxy(ntotalm), FFTTENSORxz(ntotalm), &
       FFTTENSORyy(ntotalm), FFTTENSORyz(ntotalm), FFTTENSORzz(ntotalm), vectx(ntotalm), &
       vecty(ntotalm), vectz(ntotalm), FFX(3*nmax), FFB(3*nmax)
  integer*8 :: planb, planf
  call c_matvect3(FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, FFTTENSORzz, &
       vectx, vecty, vectz, Tabdip, ntotalm, ntotal, nmax, ndipole, nxm, nym, nzm, nx, ny, nz, &
       nx2, ny2, nxy2, nz2, FFX, FFB, planb, planf)
end subroutine produitfftmatvect3

subroutine produitfftmatvectopt(FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, &
     FFTTENSORyz, FFTTENSORzz, vectx, vecty, vectz, ntotalm, ntotal, nmax, nxm, nym, nzm, nx, &
     ny, nz, nx2, ny2, nxy2, nz2, FFX, FFB, planf, planb)
  use ifdda_capi
  implicit none
  integer :: ntotalm, ntotal, nmax, nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, nz2
  double complex :: FFTTENSORxx(ntotalm), FFTTENSORxy(ntotalm), FFTTENSORxz(ntotalm), &
       FFTTENSORyy(ntotalm), FFTTENSORyz(ntotalm), FFTTENSORzz(ntotalm), vectx(ntotalm), &
       vecty(ntotalm), vectz(ntotalm), FFX(3*nmax), FFB(3*nmax)
  integer*8 :: planf, planb
  call c_matvectopt(FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, &
       FFTTENSORzz, vectx, vecty, vectz, ntotalm, ntotal, nmax, nxm, nym, nzm, nx, ny, nz, nx2, &
       ny2, nxy2, nz2, FFX, FFB, planf, planb)
end subroutine produitfftmatvectopt

subroutine greencalculfft(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, aretecube, k0, &
     FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, FFTTENSORzz, planb)
  use ifdda_capi
  implicit none
  integer :: nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm
  double precision :: aretecube, k0
  double complex, dimension(8*nxm*nym*nzm) :: FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, &
       FFTTENSORyy, FFTTENSORyz, FFTTENSORzz
  integer*8 :: planb
  call c_green(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, aretecube, k0, FFTTENSORxx, &
       FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, FFTTENSORzz, planb)
end subroutine greencalculfft

subroutine greencalculquadfft(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, nquad, tolinit, &
     aretecube, k0, FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, &
     FFTTENSORzz, planb)
  use ifdda_capi
  implicit none
  integer :: nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, nquad
  double precision :: tolinit, aretecube, k0
  double complex, dimension(8*nxm*nym*nzm) :: FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, &
       FFTTENSORyy, FFTTENSORyz, FFTTENSORzz
  integer*8 :: planb
  call c_greenquad(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, nquad, tolinit, aretecube, &
       k0, FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, FFTTENSORzz, planb)
end subroutine greencalculquadfft

subroutine inverserig(FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, &
     FFTTENSORzz, vectx, vecty, vectz, Tabdip, ntotalm, ntotal, ldabi, nlar, nmax, ndipole, &
     nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, nz2, nbsphere, nbsphere3, XI, XR, wrk, FF, FF0, &
     FFloc, polarisa, methodeit, tol, tol1, nloop, ncompte, planf, planb, nstop, infostr)
  use ifdda_capi
  implicit none
  integer :: ntotalm, ntotal, ldabi, nlar, nmax, ndipole, nxm, nym, nzm, nx, ny, nz, nx2, ny2, &
       nxy2, nz2, nbsphere, nbsphere3, nloop, ncompte, nstop
  integer :: Tabdip(nxm*nym*nzm)
  double complex, dimension(8*nxm*nym*nzm) :: FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, &
       FFTTENSORyy, FFTTENSORyz, FFTTENSORzz, vectx, vecty, vectz
  double complex, dimension(3*nxm*nym*nzm) :: XI, XR, FF, FF0, FFloc
  double complex, dimension(3*nxm*nym*nzm, 12) :: wrk
  double complex, dimension(nxm*nym*nzm, 3, 3) :: polarisa
  character(12) :: methodeit
  character(64) :: infostr
  double precision :: tol, tol1
  integer*8 :: planf, planb
  call c_inverserig(FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, &
       FFTTENSORzz, vectx, vecty, vectz, Tabdip, ntotalm, ntotal, ldabi, nlar, nmax, ndipole, &
       nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, nz2, nbsphere, nbsphere3, XI, XR, wrk, FF, &
       FF0, FFloc, polarisa, methodeit, tol, tol1, nloop, ncompte, planf, planb, nstop, &
       infostr, int(len(methodeit), c_size_t), int(len(infostr), c_size_t))
end subroutine inverserig

subroutine inverserigopt(FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, &
     FFTTENSORzz, vectx, vecty, vectz, ntotalm, ntotal, ldabi, nlar, nmax, nxm, nym, nzm, nx, &
     ny, nz, nx2, ny2, nxy2, nz2, nbsphere, nbsphere3, XI, XR, wrk, FF, FF0, FFloc, polarisa, &
     methodeit, tol, tol1, nloop, ncompte, planf, planb, nstop, infostr)
  use ifdda_capi
  implicit none
  integer :: ntotalm, ntotal, ldabi, nlar, nmax, nxm, nym, nzm, nx, ny, nz, nx2, ny2, nxy2, &
       nz2, nbsphere, nbsphere3, nloop, ncompte, nstop
  double complex, dimension(8*nxm*nym*nzm) :: FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, &
       FFTTENSORyy, FFTTENSORyz, FFTTENSORzz, vectx, vecty, vectz
  double complex, dimension(3*nxm*nym*nzm) :: XI, XR, FF, FF0, FFloc
  double complex, dimension(3*nxm*nym*nzm, 12) :: wrk
  double complex, dimension(nxm*nym*nzm, 3, 3) :: polarisa
  character(12) :: methodeit
  character(64) :: infostr
  double precision :: tol, tol1
  integer*8 :: planf, planb
  call c_inverserigopt(FFTTENSORxx, FFTTENSORxy, FFTTENSORxz, FFTTENSORyy, FFTTENSORyz, &
       FFTTENSORzz, vectx, vecty, vectz, ntotalm, ntotal, ldabi, nlar, nmax, nxm, nym, nzm, nx, &
       ny, nz, nx2, ny2, nxy2, nz2, nbsphere, nbsphere3, XI, XR, wrk, FF, FF0, FFloc, polarisa, &
       methodeit, tol, tol1, nloop, ncompte, planf, planb, nstop, infostr, &
       int(len(methodeit), c_size_t), int(len(infostr), c_size_t))
end subroutine inverserigopt

subroutine diffractefft2d(nx, ny, nz, nxm, nym, nzm, nfft2d, tabfft2, k0, xs, ys, zs, aretecube, &
     Eloinx, Eloiny, Eloinz, FF, imax, deltakx, deltaky, Ediffkzpos, Ediffkzneg, plan2f, plan2b, &
     nstop, infostr)
  use ifdda_capi
  implicit none
  integer :: nx, ny, nz, nxm, nym, nzm, nfft2d, nstop, imax
  integer :: tabfft2(nfft2d)
  double precision :: k0, aretecube, deltakx, deltaky
  double precision, dimension(nxm*nym*nzm) :: xs, ys, zs
  double complex, dimension(nfft2d*nfft2d) :: Eloinx, Eloiny, Eloinz
  double complex :: FF(3*nxm*nym*nzm), Ediffkzpos(nfft2d, nfft2d, 3), Ediffkzneg(nfft2d, nfft2d, 3)
  integer*8 :: plan2f, plan2b
  character(64) :: infostr
  call c_diffractefft2d(nx, ny, nz, nxm, nym, nzm, nfft2d, tabfft2, k0, xs, ys, zs, aretecube, &
       Eloinx, Eloiny, Eloinz, FF, imax, deltakx, deltaky, Ediffkzpos, Ediffkzneg, plan2f, &
       plan2b, nstop, infostr, int(len(infostr), c_size_t))
end subroutine diffractefft2d

subroutine computegcfft2d(imax, deltakx, deltaky, k0, I0, theta, phi, nfft2d, Ediffkzpos, &
     Ediffkzneg, gasym, Cscai)
  use ifdda_capi
  implicit none
  integer :: imax, nfft2d
  double precision :: deltakx, deltaky, k0, I0, theta, phi, gasym, Cscai
  double complex :: Ediffkzpos(nfft2d, nfft2d, 3), Ediffkzneg(nfft2d, nfft2d, 3)
  call c_computegcfft2d(imax, deltakx, deltaky, k0, I0, theta, phi, nfft2d, Ediffkzpos, &
       Ediffkzneg, gasym, Cscai)
end subroutine computegcfft2d

subroutine calculforcefft(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, tabdip, vectx, vecty, &
     vectz, FF, planb)
  use ifdda_capi
  implicit none
  integer :: nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm
  integer, dimension(nxm*nym*nzm) :: tabdip
  double complex, dimension(8*nxm*nym*nzm) :: vectx, vecty, vectz
  double complex, dimension(3*nxm*nym*nzm) :: FF
  integer*8 :: planb
  call c_forcefft(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, tabdip, vectx, vecty, vectz, &
       FF, planb)
end subroutine calculforcefft

subroutine calculforcefftopt(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, vectx, vecty, &
     vectz, FF, planb)
  use ifdda_capi
  implicit none
  integer :: nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm
  double complex, dimension(8*nxm*nym*nzm) :: vectx, vecty, vectz
  double complex, dimension(3*nxm*nym*nzm) :: FF
  integer*8 :: planb
  call c_forcefftopt(nx, ny, nz, nx2, ny2, nz2, nxy2, nxm, nym, nzm, vectx, vecty, vectz, FF, &
       planb)
end subroutine calculforcefftopt

subroutine derivativefield2(aretecube, k0, nx, ny, nz, nx2, ny2, nz2, nxy2, ntotal, ntotalm, &
     nmax, DFFTTENSORxx, DFFTTENSORxy, DFFTTENSORxz, DFFTTENSORyy, DFFTTENSORyz, DFFTTENSORzz, &
     vectx, vecty, vectz, test, FF0, planb, planf)
  use ifdda_capi
  implicit none
  integer :: nx, ny, nz, nx2, ny2, nz2, nxy2, ntotal, ntotalm, nmax, test
  double precision :: aretecube, k0
  double complex :: DFFTTENSORxx(ntotalm), DFFTTENSORxy(ntotalm), DFFTTENSORxz(ntotalm), &
       DFFTTENSORyy(ntotalm), DFFTTENSORyz(ntotalm), DFFTTENSORzz(ntotalm), vectx(ntotalm), &
       vecty(ntotalm), vectz(ntotalm), FF0(3*nmax)
  integer*8 :: planb, planf
  call c_dfield(aretecube, k0, nx, ny, nz, nx2, ny2, nz2, nxy2, ntotal, ntotalm, nmax, &
       DFFTTENSORxx, DFFTTENSORxy, DFFTTENSORxz, DFFTTENSORyy, DFFTTENSORyz, DFFTTENSORzz, vectx, &
       vecty, vectz, test, FF0, planb, planf)
end subroutine derivativefield2
