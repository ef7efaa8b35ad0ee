"""TEST INFRASTRUCTURE ONLY: ctypes access to the oracle and the compiled reference.

``Oracle`` wraps ``liboracle.so`` (the plain-C FP64 restatement in
``ifdda_oracle.c``).  ``Ref`` wraps ``_ref/libcdmref.so``, the reference's own
cdmlib Fortran compiled in place by ``oracle/Makefile``; it exists only where
the reference sources were available at build time.  Only tests/, smoke() and
bench.py's cpu_baseline leg import this module, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcdmref.so")

METHODS = {"GPBICG1": 1, "QMRBICGSTAB1": 2, "QMRBICGSTAB2": 3, "TFQMR": 4}

_p = np.ctypeslib.ndpointer
_zc = _p(np.complex128, flags="C_CONTIGUOUS")
_i32 = _p(np.int32, flags="C_CONTIGUOUS")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class _Trace(C.Structure):
    _fields_ = [("cap", C.c_int), ("n", C.c_int), ("s", C.c_void_p), ("res", C.c_void_p)]


class _Info(C.Structure):
    _fields_ = [("tol1", C.c_double), ("nloop", C.c_int), ("ncompte", C.c_int),
                ("nstop", C.c_int), ("status", C.c_int), ("steperr", C.c_int)]


class Oracle:
    """The C restatement (FP64, single thread)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        self.lib = C.CDLL(path)
        L = self.lib
        L.orc_propa_espace_libre.argtypes = [C.c_double] * 7 + [_zc]
        L.orc_propa_int_sim.argtypes = [C.c_double] * 9 + [_zc]
        L.orc_fft3d.argtypes = [_zc, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_fft1d.argtypes = [_zc, C.c_int, C.c_int]
        L.orc_green_fft.argtypes = [C.c_int] * 6 + [C.c_double, C.c_double, C.c_int, _zc]
        L.orc_matvec.argtypes = [_zc, C.c_void_p] + [C.c_int] * 6 + [_zc, _zc]
        L.orc_matvec_direct.argtypes = [C.c_void_p] + [C.c_int] * 3 + [C.c_double, C.c_double, _zc, _zc]
        L.orc_inverserig.argtypes = ([_zc, C.c_void_p] + [C.c_int] * 7 + [_zc, C.c_int, C.c_int,
                                     C.c_double, C.c_int, _zc, _zc, _zc, C.POINTER(_Info), C.c_void_p])

    def propa(self, r_obs, r_src, k0):
        T = np.zeros(9, np.complex128)
        self.lib.orc_propa_espace_libre(*map(float, r_obs), *map(float, r_src), float(k0), T)
        return T.reshape(3, 3)

    def propa_int(self, r_obs, r_src, k0, d, lim=1.0):
        T = np.zeros(9, np.complex128)
        self.lib.orc_propa_int_sim(*map(float, r_obs), *map(float, r_src), float(k0), float(d),
                                   float(lim), T)
        return T.reshape(3, 3)

    def fft3d(self, v, sign):
        """v: array indexed [k][j][i] (x fastest)."""
        a = np.ascontiguousarray(v, np.complex128).copy()
        n3, n2, n1 = a.shape
        self.lib.orc_fft3d(a.reshape(-1), n1, n2, n3, int(sign))
        return a

    def green_fft(self, nx, ny, nz, L, d, k0, nquad=0):
        """Returns G[6, L3, L2, L1] (xx,xy,xz,yy,yz,zz), FFT sign +1."""
        L1, L2, L3 = L
        G = np.zeros(6 * L1 * L2 * L3, np.complex128)
        self.lib.orc_green_fft(nx, ny, nz, L1, L2, L3, float(d), float(k0), int(nquad), G)
        return G.reshape(6, L3, L2, L1)

    def matvec(self, G, tabdip, nx, ny, nz, FFX, FFB):
        """FFB - conv(FFX) at dipole cells; tabdip None = dense box."""
        L3, L2, L1 = G.shape[1:]
        out = np.ascontiguousarray(FFB, np.complex128).copy()
        tab = None if tabdip is None else np.ascontiguousarray(tabdip, np.int32)
        self.lib.orc_matvec(np.ascontiguousarray(G).reshape(-1), _ptr(tab), nx, ny, nz, L1, L2, L3,
                            np.ascontiguousarray(FFX, np.complex128), out)
        return out

    def matvec_direct(self, tabdip, nx, ny, nz, d, k0, FFX, FFB):
        out = np.ascontiguousarray(FFB, np.complex128).copy()
        tab = None if tabdip is None else np.ascontiguousarray(tabdip, np.int32)
        self.lib.orc_matvec_direct(_ptr(tab), nx, ny, nz, float(d), float(k0),
                                   np.ascontiguousarray(FFX, np.complex128), out)
        return out

    def inverserig(self, G, tabdip, nx, ny, nz, polarisa, FF0, method="GPBICG1", tol=1e-4,
                   maxit=10000, trace_cap=0):
        """polarisa: complex (nmax,3,3) array (Fortran meaning, [i,a,b])."""
        L3, L2, L1 = G.shape[1:]
        nb = FF0.size // 3
        nmax = polarisa.shape[0]
        pol_f = np.asfortranarray(polarisa, np.complex128).reshape(-1, order="F")
        pol_f = np.ascontiguousarray(pol_f)
        FF = np.zeros(3 * nb, np.complex128)
        FFloc = np.zeros(3 * nb, np.complex128)
        info = _Info()
        tr = None
        if trace_cap:
            s = np.zeros(5 * trace_cap, np.complex128)
            res = np.zeros(trace_cap, np.float64)
            tr = _Trace(trace_cap, 0, s.ctypes.data, res.ctypes.data)
        tab = None if tabdip is None else np.ascontiguousarray(tabdip, np.int32)
        self.lib.orc_inverserig(np.ascontiguousarray(G).reshape(-1), _ptr(tab), nx, ny, nz, L1, L2,
                                L3, nb, pol_f, nmax, METHODS[method], float(tol), int(maxit),
                                np.ascontiguousarray(FF0, np.complex128), FF, FFloc, C.byref(info),
                                None if tr is None else C.byref(tr))
        out = dict(FF=FF, FFloc=FFloc, tol1=info.tol1, nloop=info.nloop, ncompte=info.ncompte,
                   nstop=info.nstop, status=info.status, steperr=info.steperr)
        if tr is not None:
            out["trace"] = s.reshape(-1, 5)[: tr.n].copy()
            out["trace_res"] = res[: tr.n].copy()
        return out


# --------------------------------------------------------------------------
# The compiled reference (Fortran calling convention: everything by reference,
# hidden character lengths appended as size_t).
# --------------------------------------------------------------------------

def _ri(v):
    return C.byref(C.c_int(int(v)))


def _rd(v):
    return C.byref(C.c_double(float(v)))


def _ri8(v=0):
    return C.byref(C.c_int64(int(v)))


class Ref:
    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"reference library missing: {path}")
        self.lib = C.CDLL(path)

    @staticmethod
    def available(path: str = REF_SO) -> bool:
        return os.path.exists(path)

    def flush_stdout(self):
        """Flush the Fortran runtime's buffered unit 6 (oracle/ref_stdout.f90)."""
        f = getattr(self.lib, "ref_flush_stdout", None)
        if f is not None:
            f()

    def propa(self, r_obs, r_src, k0):
        T = np.zeros((3, 3), np.complex128, order="F")
        a = [C.c_double(float(v)) for v in (*r_obs, *r_src, k0)]
        self.lib.propa_espace_libre_(*[C.byref(x) for x in a], T.ctypes.data_as(C.c_void_p))
        return np.array(T)

    def propa_int(self, r_obs, r_src, k0, d, lim=1.0, relreq=1e-4):
        T = np.zeros((3, 3), np.complex128, order="F")
        a = [C.c_double(float(v)) for v in (*r_obs, *r_src, k0, d, lim, relreq)]
        self.lib.propaespacelibreintsim_(*[C.byref(x) for x in a], T.ctypes.data_as(C.c_void_p))
        return np.array(T)

    def green_fft(self, nx, ny, nz, d, k0, nquad=0, tolinit=1e-4):
        """greencalculfft_ / greencalculquadfft_ on the 2n grid; returns G[6,L3,L2,L1]."""
        nx2, ny2, nz2 = 2 * nx, 2 * ny, 2 * nz
        M = nx2 * ny2 * nz2
        G = [np.zeros(M, np.complex128) for _ in range(6)]
        gp = [g.ctypes.data_as(C.c_void_p) for g in G]
        common = [_ri(nx), _ri(ny), _ri(nz), _ri(nx2), _ri(ny2), _ri(nz2), _ri(nx2 * ny2),
                  _ri(nx), _ri(ny), _ri(nz)]
        if nquad:
            self.lib.greencalculquadfft_(*common, _ri(nquad), _rd(tolinit), _rd(d), _rd(k0), *gp,
                                         _ri8())
        else:
            self.lib.greencalculfft_(*common, _rd(d), _rd(k0), *gp, _ri8())
        return np.stack(G).reshape(6, nz2, ny2, nx2)

    def propder(self, r_obs, r_src, k0, test):
        """propesplibdermat_: d T / d x_test at r_obs for a source at r_src."""
        S = np.zeros((3, 3, 3), np.complex128, order="F")
        a = [C.c_double(float(v)) for v in (*r_obs, *r_src, k0)]
        self.lib.propesplibdermat_(*[C.byref(x) for x in a], _ri(test), S.ctypes.data_as(C.c_void_p))
        return np.array(S[:, :, test - 1])

    def dfield(self, FF, tabdip, nx, ny, nz, d, k0, test):
        """calculforcefft_ then derivativefield2_: the x_test derivative of the
        field radiated by the moments FF (dipole order), on every box cell
        (3 nx ny nz values, box order)."""
        nx2, ny2, nz2 = 2 * nx, 2 * ny, 2 * nz
        M = nx2 * ny2 * nz2
        nmax = nx * ny * nz
        V = self._vect(M)
        D = self._vect(M) + self._vect(M)
        FFp = np.zeros(3 * nmax, np.complex128)
        FFp[: FF.size] = FF
        tab = np.ascontiguousarray(tabdip, np.int32)
        P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        dims = [_ri(nx), _ri(ny), _ri(nz), _ri(nx2), _ri(ny2), _ri(nz2), _ri(nx2 * ny2)]
        self.lib.calculforcefft_(*dims, _ri(nx), _ri(ny), _ri(nz), P(tab), *[P(v) for v in V],
                                 P(FFp), _ri8())
        out = np.zeros(3 * nmax, np.complex128)
        self.lib.derivativefield2_(_rd(d), _rd(k0), _ri(nx), _ri(ny), _ri(nz), _ri(nx2), _ri(ny2),
                                   _ri(nz2), _ri(nx2 * ny2), _ri(M), _ri(M), _ri(nmax),
                                   *[P(x) for x in D], *[P(v) for v in V], _ri(test), P(out),
                                   _ri8(), _ri8())
        return out

    def ondeplaned(self, r, k0, E0, ss, pp, theta, phi):
        """ondeplaned_ (test 4): Eder[c, d] = d E_c / d x_d of the plane wave."""
        Eder = np.zeros((3, 3), np.complex128, order="F")
        e0 = (C.c_double * 2)(E0.real, E0.imag)
        self.lib.ondeplaned_(*[_rd(v) for v in r], _rd(k0), C.byref(e0), _rd(ss), _rd(pp),
                             _rd(theta), _rd(phi), _ri(4), Eder.ctypes.data_as(C.c_void_p))
        return np.array(Eder)

    def bhmie(self, x, m, nang=2):
        """BHMIE_ of module_mie.f: (qext, qsca, qback, gsca)."""
        S1 = np.zeros(2 * 1000 - 1, np.complex128)
        S2 = np.zeros(2 * 1000 - 1, np.complex128)
        mm = (C.c_double * 2)(m.real, m.imag)
        out = [C.c_double(0) for _ in range(4)]
        self.lib.bhmie_(_rd(x), C.byref(mm), _ri(nang), S1.ctypes.data_as(C.c_void_p),
                        S2.ctypes.data_as(C.c_void_p), *[C.byref(o) for o in out])
        return tuple(o.value for o in out)

    def _vect(self, M):
        return [np.zeros(M, np.complex128) for _ in range(3)]

    def matvec3(self, G, tabdip, nx, ny, nz, FFX, FFB):
        """produitfftmatvect3_: returns FFB - conv(FFX)."""
        nz2, ny2, nx2 = G.shape[1:]
        M = nx2 * ny2 * nz2
        Gs = [np.ascontiguousarray(G[c]).reshape(-1) for c in range(6)]
        V = self._vect(M)
        nd = FFX.size // 3
        out = np.ascontiguousarray(FFB, np.complex128).copy()
        tab = np.ascontiguousarray(tabdip, np.int32)
        self.lib.produitfftmatvect3_(*[g.ctypes.data_as(C.c_void_p) for g in Gs],
                                     *[v.ctypes.data_as(C.c_void_p) for v in V],
                                     tab.ctypes.data_as(C.c_void_p), _ri(M), _ri(M), _ri(nd),
                                     _ri(nx * ny * nz), _ri(nx), _ri(ny), _ri(nz), _ri(nx), _ri(ny),
                                     _ri(nz), _ri(nx2), _ri(ny2), _ri(nx2 * ny2), _ri(nz2),
                                     np.ascontiguousarray(FFX, np.complex128).ctypes.data_as(C.c_void_p),
                                     out.ctypes.data_as(C.c_void_p), _ri8(), _ri8())
        return out

    def matvec_opt(self, G, nx, ny, nz, FFX, FFB):
        """produitfftmatvectopt_ (dense box): returns FFB - conv(FFX)."""
        nz2, ny2, nx2 = G.shape[1:]
        M = nx2 * ny2 * nz2
        Gs = [np.ascontiguousarray(G[c]).reshape(-1) for c in range(6)]
        V = self._vect(M)
        nmax = nx * ny * nz
        out = np.ascontiguousarray(FFB, np.complex128).copy()
        self.lib.produitfftmatvectopt_(*[g.ctypes.data_as(C.c_void_p) for g in Gs],
                                       *[v.ctypes.data_as(C.c_void_p) for v in V],
                                       _ri(M), _ri(M), _ri(nmax), _ri(nx), _ri(ny), _ri(nz), _ri(nx),
                                       _ri(ny), _ri(nz), _ri(nx2), _ri(ny2), _ri(nx2 * ny2), _ri(nz2),
                                       np.ascontiguousarray(FFX, np.complex128).ctypes.data_as(C.c_void_p),
                                       out.ctypes.data_as(C.c_void_p), _ri8(), _ri8())
        return out

    def inverserig(self, G, tabdip, nx, ny, nz, polarisa, FF0, method="GPBICG1", tol=1e-4):
        """inverserig_ (x0 = E_inc).  polarisa: (nmax,3,3) complex."""
        nz2, ny2, nx2 = G.shape[1:]
        M = nx2 * ny2 * nz2
        nmax = nx * ny * nz
        nb = FF0.size // 3
        Gs = [np.ascontiguousarray(G[c]).reshape(-1) for c in range(6)]
        V = self._vect(M)
        tab = np.ascontiguousarray(tabdip, np.int32)
        XI = np.zeros(3 * nmax, np.complex128)
        XR = np.zeros(3 * nmax, np.complex128)
        wrk = np.zeros(3 * nmax * 12, np.complex128)
        FF = np.zeros(3 * nmax, np.complex128)
        FF0p = np.zeros(3 * nmax, np.complex128)
        FF0p[: 3 * nb] = FF0
        FFloc = np.zeros(3 * nmax, np.complex128)
        pol = np.zeros((nmax, 3, 3), np.complex128)
        pol[: polarisa.shape[0]] = polarisa
        pol_f = np.asfortranarray(pol)
        meth = method.ljust(12).encode()
        info = C.create_string_buffer(b" " * 64, 64)
        tol1 = C.c_double(0)
        nloop = C.c_int(0)
        ncompte = C.c_int(0)
        nstop = C.c_int(0)
        P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        self.lib.inverserig_(*[P(g) for g in Gs], *[P(v) for v in V], P(tab), _ri(M), _ri(M),
                             _ri(3 * nmax), _ri(12), _ri(nmax), _ri(nmax), _ri(nx), _ri(ny), _ri(nz),
                             _ri(nx), _ri(ny), _ri(nz), _ri(nx2), _ri(ny2), _ri(nx2 * ny2), _ri(nz2),
                             _ri(nb), _ri(3 * nb), P(XI), P(XR), P(wrk), P(FF), P(FF0p), P(FFloc),
                             P(pol_f), C.c_char_p(meth), _rd(tol), C.byref(tol1), C.byref(nloop),
                             C.byref(ncompte), _ri8(), _ri8(), C.byref(nstop), info,
                             C.c_size_t(12), C.c_size_t(64))
        return dict(FF=FF[: 3 * nb].copy(), FFloc=FFloc[: 3 * nb].copy(), tol1=tol1.value,
                    nloop=nloop.value, ncompte=ncompte.value, nstop=nstop.value,
                    infostr=info.raw.decode(errors="replace").rstrip())

    def inverserigopt(self, G, nx, ny, nz, polarisa, FF0, method="GPBICG1", tol=1e-4):
        """inverserigopt_ (inverserigopt.f:1-6): the dense-box driver, every
        box cell a dipole (nbsphere = nx ny nz), x0 = E_inc; its 39 arguments
        (planf before planb) and the two hidden character lengths."""
        nz2, ny2, nx2 = G.shape[1:]
        M = nx2 * ny2 * nz2
        nmax = nx * ny * nz
        Gs = [np.ascontiguousarray(G[c]).reshape(-1) for c in range(6)]
        V = self._vect(M)
        XI = np.zeros(3 * nmax, np.complex128)
        XR = np.zeros(3 * nmax, np.complex128)
        wrk = np.zeros(3 * nmax * 12, np.complex128)
        FF = np.zeros(3 * nmax, np.complex128)
        FF0p = np.ascontiguousarray(FF0, np.complex128).copy()
        FFloc = np.zeros(3 * nmax, np.complex128)
        pol_f = np.asfortranarray(np.asarray(polarisa, np.complex128))
        meth = method.ljust(12).encode()
        info = C.create_string_buffer(b" " * 64, 64)
        tol1 = C.c_double(0)
        nloop = C.c_int(0)
        ncompte = C.c_int(0)
        nstop = C.c_int(0)
        P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        self.lib.inverserigopt_(*[P(g) for g in Gs], *[P(v) for v in V], _ri(M), _ri(M),
                                _ri(3 * nmax), _ri(12), _ri(nmax), _ri(nx), _ri(ny), _ri(nz),
                                _ri(nx), _ri(ny), _ri(nz), _ri(nx2), _ri(ny2), _ri(nx2 * ny2),
                                _ri(nz2), _ri(nmax), _ri(3 * nmax), P(XI), P(XR), P(wrk), P(FF),
                                P(FF0p), P(FFloc), P(pol_f), C.c_char_p(meth), _rd(tol),
                                C.byref(tol1), C.byref(nloop), C.byref(ncompte), _ri8(), _ri8(),
                                C.byref(nstop), info, C.c_size_t(12), C.c_size_t(64))
        return dict(FF=FF, FFloc=FFloc, tol1=tol1.value, nloop=nloop.value,
                    ncompte=ncompte.value, nstop=nstop.value,
                    infostr=info.raw.decode(errors="replace").rstrip())

    def poladiff(self, d, eps, k0, method="RR", eps0=1.0):
        out = C.c_double * 2
        res = out()
        e = (C.c_double * 2)(eps.real, eps.imag)
        m = method.ljust(2).encode()
        self.lib.poladiff_(_rd(d), C.byref(e), _rd(eps0), _rd(k0), _ri(1), C.c_char_p(m), C.byref(res),
                           C.c_size_t(2))
        return complex(res[0], res[1])

    def diffractefft2d(self, FFbox, nx, ny, nz, nfft2d, k0, xs1, ys1, zs, d):
        """diffractefft2d_ (Singleton FFT build).  FFbox box-ordered moments;
        zs[k] the z of plane k.  Returns (Epos, Eneg, imax, deltak, nstop,
        infostr) with E[ii, jj, c] cut to ii, jj < 2 imax + 1."""
        N = nx * ny * nz
        cell = np.arange(N)
        xs = np.full(N, float(xs1))
        ys = np.full(N, float(ys1))
        zz = np.asarray(zs, np.float64)[cell // (nx * ny)].copy()
        xs = np.ascontiguousarray(xs)
        ys = np.ascontiguousarray(ys)
        tab = np.zeros(nfft2d, np.int32)
        scratch = [np.zeros(nfft2d * nfft2d, np.complex128) for _ in range(3)]
        F = np.ascontiguousarray(FFbox, np.complex128).reshape(-1)
        Ep = np.zeros((3, nfft2d, nfft2d), np.complex128)  # Fortran (nfft2d, nfft2d, 3)
        En = np.zeros((3, nfft2d, nfft2d), np.complex128)
        imax = C.c_int(0)
        dkx = C.c_double(0.0)
        dky = C.c_double(0.0)
        nstop = C.c_int(0)
        info = C.create_string_buffer(b" " * 64, 64)
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        self.lib.diffractefft2d_(_ri(nx), _ri(ny), _ri(nz), _ri(nx), _ri(ny), _ri(nz), _ri(nfft2d),
                                 P(tab), _rd(k0), P(xs), P(ys), P(zz), _rd(d), *[P(s) for s in scratch],
                                 P(F), C.byref(imax), C.byref(dkx), C.byref(dky), P(Ep), P(En), _ri8(),
                                 _ri8(), C.byref(nstop), info, C.c_size_t(64))
        n = 2 * imax.value + 1
        # Fortran E(ii, jj, c) is C-order [c][jj][ii]
        cut = lambda E: np.transpose(E[:, :n, :n], (2, 1, 0)).copy()
        return (cut(Ep), cut(En), imax.value, dkx.value, nstop.value,
                info.raw.decode(errors="replace").rstrip())

    def computegcfft2d(self, Epos, Eneg, imax, dk, k0, I0, theta, phi, nfft2d):
        """computegcfft2d_: (Csca, g) from E[ii, jj, c] (compact)."""
        n = 2 * imax + 1
        full = []
        for E in (Epos, Eneg):
            A = np.zeros((3, nfft2d, nfft2d), np.complex128)
            A[:, :n, :n] = np.transpose(E, (2, 1, 0))
            full.append(A)
        g = C.c_double(0.0)
        cs = C.c_double(0.0)
        self.lib.computegcfft2d_(_ri(imax), _rd(dk), _rd(dk), _rd(k0), _rd(I0), _rd(theta), _rd(phi),
                                 _ri(nfft2d), full[0].ctypes.data_as(C.c_void_p),
                                 full[1].ctypes.data_as(C.c_void_p), C.byref(g), C.byref(cs))
        return cs.value, g.value

    def ondeplanekxky(self, pos, k0, E0, pp, kx, ky):
        out = np.zeros((len(pos), 3), np.complex128)
        e0 = (C.c_double * 2)(E0.real, E0.imag)
        nstop = C.c_int(0)
        info = C.create_string_buffer(b" " * 64, 64)
        for n, (x, y, z) in enumerate(pos):
            ex, ey, ez = ((C.c_double * 2)() for _ in range(3))
            self.lib.ondeplanekxky_(_rd(x), _rd(y), _rd(z), _rd(k0), C.byref(e0), _rd(1.0 - pp), _rd(pp),
                                    _rd(kx), _rd(ky), C.byref(ex), C.byref(ey), C.byref(ez),
                                    C.byref(nstop), info, C.c_size_t(64))
            out[n] = [complex(ex[0], ex[1]), complex(ey[0], ey[1]), complex(ez[0], ez[1])]
        return out.reshape(-1)

    def ondeplane(self, pos, k0, E0, pp, theta_deg, phi_deg):
        """ondeplane_ (ondeplane.f:1-50) at each position; ss = 1 - pp."""
        out = np.zeros((len(pos), 3), np.complex128)
        e0 = (C.c_double * 2)(E0.real, E0.imag)
        nstop = C.c_int(0)
        info = C.create_string_buffer(b" " * 64, 64)
        for n, (x, y, z) in enumerate(pos):
            ex, ey, ez = ((C.c_double * 2)() for _ in range(3))
            self.lib.ondeplane_(_rd(x), _rd(y), _rd(z), _rd(k0), C.byref(e0), _rd(1.0 - pp), _rd(pp),
                                _rd(theta_deg), _rd(phi_deg), C.byref(ex), C.byref(ey), C.byref(ez),
                                C.byref(nstop), info, C.c_size_t(64))
            out[n] = [complex(ex[0], ex[1]), complex(ey[0], ey[1]), complex(ez[0], ez[1])]
        return out.reshape(-1)

    def ondecirce(self, pos, k0, E0, helicity, theta_deg, phi_deg):
        """ondecirce_ (ondecirce.f:1-55) at each position."""
        out = np.zeros((len(pos), 3), np.complex128)
        e0 = (C.c_double * 2)(E0.real, E0.imag)
        for n, (x, y, z) in enumerate(pos):
            ex, ey, ez = ((C.c_double * 2)() for _ in range(3))
            self.lib.ondecirce_(_rd(x), _rd(y), _rd(z), _rd(k0), C.byref(e0), _rd(helicity),
                                _rd(theta_deg), _rd(phi_deg), C.byref(ex), C.byref(ey), C.byref(ez))
            out[n] = [complex(ex[0], ex[1]), complex(ey[0], ey[1]), complex(ez[0], ez[1])]
        return out.reshape(-1)

    def deltakroutine(self, kxinc, kyinc, dkx, dky, numaper):
        i, j = C.c_int(0), C.c_int(0)
        self.lib.deltakroutine_(_rd(kxinc), _rd(kyinc), _rd(dkx), _rd(dky), _rd(numaper), C.byref(i),
                                C.byref(j))
        return i.value, j.value

    def irradiance(self, P0, w0):
        e0 = (C.c_double * 2)()
        irra = C.c_double(0.0)
        self.lib.irradiance_(_rd(P0), _rd(w0), C.byref(e0), C.byref(irra))
        return complex(e0[0], e0[1]), irra.value

    def diffractefft2dlens(self, FFbox, nx, ny, nz, nfft2d, k0, xs1, ys1, zs, d, numaper, nside):
        """diffractefft2dlens_: (E[ii, jj, c] compact, imax, deltak, nstop, infostr)."""
        N = nx * ny * nz
        xs = np.full(N, float(xs1))
        ys = np.full(N, float(ys1))
        zz = np.asarray(zs, np.float64)[np.arange(N) // (nx * ny)].copy()
        tab = np.zeros(nfft2d, np.int32)
        scratch = [np.zeros(nfft2d * nfft2d, np.complex128) for _ in range(3)]
        F = np.ascontiguousarray(FFbox, np.complex128).reshape(-1)
        Ep = np.zeros((3, nfft2d, nfft2d), np.complex128)
        imax, nstop = C.c_int(0), C.c_int(0)
        dkx, dky = C.c_double(0.0), C.c_double(0.0)
        info = C.create_string_buffer(b" " * 64, 64)
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        self.lib.diffractefft2dlens_(_ri(nx), _ri(ny), _ri(nz), _ri(nx), _ri(ny), _ri(nz), _ri(nfft2d),
                                     P(tab), _rd(k0), P(xs), P(ys), P(zz), _rd(d),
                                     *[P(s) for s in scratch], P(F), C.byref(imax), C.byref(dkx),
                                     C.byref(dky), P(Ep), _rd(numaper), _ri(nside), _ri8(), _ri8(),
                                     C.byref(nstop), info, C.c_size_t(64))
        n = 2 * imax.value + 1
        E = np.transpose(Ep[:, :n, :n], (2, 1, 0)).copy()
        return E, imax.value, dkx.value, nstop.value, info.raw.decode(errors="replace").rstrip()

    MICROSCOPES = ("microsbf", "microsdf", "microsbfneg", "microsdfneg")

    def microscope(self, kind, G, pos, polarisa, d, k0, nfft2d, numaper, numaperinc, P0, w0, gross,
                   zlens, nquicklens, method="GPBICG1", tol=1e-4):
        """One of the four whole microscopes, run as cdmlib.f:3674-3800 calls
        them (numaper scaled by k0, nrig = 0, nproche = -1 so each incidence
        goes through inverserig, nmat = 1 so no file is written, npolainc = 0:
        both polarisations): microsbf_ (microsbf.f:1-12), microsdf_
        (microsdf.f:1-12), and the nside = -1 microsbfneg_ / microsdfneg_
        (microssansinc.f:854-863, 1-10), which have no with-incident image
        arguments.  The box is dense: every cell a dipole (Tabdip(i) = i,
        nbsphere = nx ny nz), empty cells with zero polarisability, pos the
        box-ordered cell centres.  Returns the six images [scattered x, y, z,
        with-incident x, y, z][y][x] (the square roots of the summed
        intensities; the nside = -1 kinds repeat the scattered image), kxy,
        xy, P0 after the call (P0 / number of illuminations), irra, nstop,
        infostr."""
        assert kind in self.MICROSCOPES
        neg = kind.endswith("neg")
        nside = -1 if neg else 1
        nz2, ny2, nx2 = G.shape[1:]
        nx, ny, nz = nx2 // 2, ny2 // 2, nz2 // 2
        M = nx2 * ny2 * nz2
        nmax = nx * ny * nz
        assert pos.shape == (nmax, 3) and polarisa.shape == (nmax, 3, 3)
        Gs = [np.ascontiguousarray(G[c]).reshape(-1) for c in range(6)]
        V = self._vect(M)
        tabfft2 = np.zeros(nfft2d, np.int32)
        tab = np.arange(1, nmax + 1, dtype=np.int32)
        XI, XR, FF, FF0, FFloc = (np.zeros(3 * nmax, np.complex128) for _ in range(5))
        wrk = np.zeros(3 * nmax * 12, np.complex128)
        pol_f = np.asfortranarray(np.asarray(polarisa, np.complex128))
        eps_f = np.zeros((nmax, 3, 3), np.complex128, order="F")
        xs, ys, zs = (np.ascontiguousarray(pos[:, a], np.float64) for a in range(3))
        planes = [np.zeros(nfft2d * nfft2d, np.complex128) for _ in range(12)]
        Ekp = np.zeros(3 * nfft2d * nfft2d, np.complex128)
        Ekn = np.zeros(3 * nfft2d * nfft2d, np.complex128)
        kxy = np.zeros(nfft2d, np.float64)
        xy = np.zeros(nfft2d, np.float64)
        meth = method.ljust(12).encode()
        info = C.create_string_buffer(b" " * 64, 64)
        tol1, p0, irra = C.c_double(0), C.c_double(float(P0)), C.c_double(0)
        nloop, ncompte, nstop = C.c_int(0), C.c_int(0), C.c_int(0)
        file_id, group_id = C.c_int(0), C.c_int(0)  # integer(hid_t), hid_t = 4 without HDF5
        P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        # Eimagex..z, Efourierx..z, Eimageincx..z, Efourierincx..z in argument order
        Eim, Efo, Eiminc, Efoinc = planes[0:3], planes[3:6], planes[6:9], planes[9:12]
        inc_args = [] if neg else [*[P(a) for a in Eiminc], *[P(a) for a in Efoinc]]
        args = [*[P(g) for g in Gs], *[P(v) for v in V], _ri(M), _ri(M), _ri(3 * nmax), _ri(12),
                _ri(nmax), _ri(nx), _ri(ny), _ri(nz), _ri(nx), _ri(ny), _ri(nz), _ri(nx2), _ri(ny2),
                _ri(nx2 * ny2), _ri(nz2), _ri(nmax), _ri(nmax), _ri(3 * nmax), _ri(-1), _ri(0),
                _ri(nfft2d), P(tabfft2), P(tab), P(XI), P(XR), P(wrk), P(FF), P(FF0), P(FFloc),
                P(pol_f), P(eps_f), C.c_char_p(meth), _rd(tol), C.byref(tol1), C.byref(nloop),
                C.byref(ncompte), P(xs), P(ys), P(zs), _rd(d), _rd(numaper * k0),
                _rd(numaperinc), _ri(0), _ri(nquicklens), _rd(1.0), _rd(k0), C.byref(p0),
                C.byref(irra), _rd(w0), _rd(gross), _rd(zlens), *[P(a) for a in Eim],
                *[P(a) for a in Efo], *inc_args, P(Ekp), P(Ekn), P(kxy), P(xy), _ri(nside),
                _ri8(), _ri8(), _ri8(), _ri8(), _ri(1), C.byref(file_id), C.byref(group_id),
                C.byref(nstop), info, C.c_size_t(12), C.c_size_t(64)]
        getattr(self.lib, kind + "_")(*args)
        self.flush_stdout()
        img = np.stack([a.reshape(nfft2d, nfft2d) for a in (*Eim, *(Eim if neg else Eiminc))])
        return dict(images=img, kxy=kxy, xy=xy, P0=p0.value, irra=irra.value, nstop=nstop.value,
                    infostr=info.raw.decode(errors="replace").rstrip())

    def fouriertoimage(self, dk, gross, planes):
        """fouriertoimage_ on six N x N planes ([y][x]): scattered x, y, z then
        with-incident x, y, z."""
        N = planes[0].shape[0]
        A = [np.ascontiguousarray(p, np.complex128).copy() for p in planes]
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        self.lib.fouriertoimage_(_rd(dk), _rd(dk), _rd(gross), *[P(a) for a in A], _ri(N), _ri(N // 2),
                                 _ri8(), _ri8())
        return A
