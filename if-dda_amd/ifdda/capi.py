"""ctypes binding of libifdda_hip.so (the C ABI declared in include/ifdda.h).

``Solver`` wraps one ``ifdda_ctx``: a box of nx*ny*nz cells with its
device-resident Green tensor, dipole map, polarisability and Krylov state.
Host arrays are numpy complex128 in the reference's layouts (dipole vectors
AoS (Ex,Ey,Ez); grids x-fastest; polarisa (nmax,3,3) with Fortran meaning).
Device arrays may be passed as integer device pointers (e.g. from torch
tensors, used here only as memory).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
lib_path = os.path.join(PKG, "libifdda_hip.so")

GPBICG1, QMRBICGSTAB1, QMRBICGSTAB2, TFQMR = 1, 2, 3, 4
METHODS = {"GPBICG1": GPBICG1, "QMRBICGSTAB1": QMRBICGSTAB1, "QMRBICGSTAB2": QMRBICGSTAB2,
           "TFQMR": TFQMR, "CG": 5, "BICGSTAB": 6, "GPBICG2": 7, "GPBICGsafe": 8, "GPBICGplus": 9,
           "GPBICGAR": 10, "GPBICGAR2": 10, "BICGSTARPLUS": 11, "GPBICOR": 12, "CORS": 13,
           "QMRCLA": 14}


class IfddaError(RuntimeError):
    pass


class SolveOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("precision", C.c_int), ("x0_zero", C.c_int),
                ("maxit", C.c_int), ("tol", C.c_double), ("check_every", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("tol1", C.c_double), ("nloop", C.c_int), ("ncompte", C.c_int), ("nstop", C.c_int),
                ("status", C.c_int), ("steperr", C.c_int), ("outer", C.c_int),
                ("seconds", C.c_double), ("infostr", C.c_char * 64)]

    def as_dict(self):
        return dict(tol1=self.tol1, nloop=self.nloop, ncompte=self.ncompte, nstop=self.nstop,
                    status=self.status, steperr=self.steperr, outer=self.outer,
                    seconds=self.seconds, infostr=self.infostr.decode(errors="replace").rstrip())


_lib = None


class MicroscopeOpts(C.Structure):
    """ifdda_microscope_opts (include/ifdda.h)."""
    _fields_ = [("nfft2d", C.c_int), ("numaper", C.c_double), ("numaperinc", C.c_double),
                ("gross", C.c_double), ("zlens", C.c_double), ("nside", C.c_int),
                ("npolainc", C.c_int), ("E0", C.c_double * 2), ("k0", C.c_double),
                ("d", C.c_double), ("origin", C.c_double * 3), ("max_batch", C.c_int),
                ("direct_lens", C.c_int), ("ninc_user", C.c_int), ("kxy_user", C.c_void_p)]


class Job(C.Structure):
    """ifdda_job (include/ifdda.h)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("d", C.c_double),
                ("k0", C.c_double), ("nquad", C.c_int), ("tolinit", C.c_double),
                ("tabdip", C.c_void_p), ("nd", C.c_int), ("alpha", C.c_void_p),
                ("nrhs", C.c_int), ("E_inc", C.c_void_p), ("opts", SolveOpts),
                ("wave_kind", C.c_int), ("waves", C.c_void_p), ("E0", C.c_double * 2),
                ("origin", C.c_double * 3), ("eps", C.c_void_p), ("pola", C.c_int),
                ("eps0", C.c_double)]


WAVES = {"plane": 1, "circular": 2, "kxky": 3}
POLA = {"CM": 0, "RR": 1, "PS": 1, "GB": 2, "LA": 3, "LR": 4}


def lib():
    """Load libifdda_hip.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(lib_path):
        raise IfddaError(f"{lib_path} is missing: run __graft_entry__.build() first")
    L = C.CDLL(lib_path)
    vp, ip, dp = C.c_void_p, C.c_int, C.c_double
    L.ifdda_last_error.restype = C.c_char_p
    L.ifdda_create.argtypes = [ip, ip, ip, C.POINTER(vp)]
    L.ifdda_destroy.argtypes = [vp]
    L.ifdda_destroy.restype = None
    L.ifdda_fft_lengths.argtypes = [vp, C.POINTER(C.c_int)]
    L.ifdda_green_build.argtypes = [vp, dp, dp, ip, dp]
    L.ifdda_green_set.argtypes = [vp, C.POINTER(vp)]
    L.ifdda_green_get.argtypes = [vp, C.POINTER(vp)]
    L.ifdda_set_dipoles.argtypes = [vp, vp, ip]
    L.ifdda_set_polarizability.argtypes = [vp, vp, ip, ip]
    L.ifdda_matvec.argtypes = [vp, vp, vp]
    L.ifdda_matvec_device.argtypes = [vp, ip, ip, vp, vp, vp, vp]
    L.ifdda_solve.argtypes = [vp, C.POINTER(SolveOpts), vp, vp, vp, C.POINTER(SolveInfo),
                              C.POINTER(C.c_int)]
    L.ifdda_solve_device.argtypes = [vp, C.POINTER(SolveOpts), vp, vp, vp, C.POINTER(SolveInfo),
                                     C.POINTER(C.c_int), vp]
    L.ifdda_solve_device_noalloc.argtypes = [vp, C.POINTER(SolveOpts), vp, vp, vp,
                                             C.POINTER(SolveInfo), C.POINTER(C.c_int), vp]
    L.ifdda_solve_multi.argtypes = [vp, C.POINTER(SolveOpts), ip, vp, vp, vp, C.POINTER(SolveInfo),
                                    C.POINTER(C.c_int)]
    L.ifdda_solve_multi_device.argtypes = [vp, C.POINTER(SolveOpts), ip, vp, vp, vp,
                                           C.POINTER(SolveInfo), C.POINTER(C.c_int), vp]
    L.ifdda_solve_trace.argtypes = [vp, ip, vp, C.POINTER(C.c_int)]
    L.ifdda_solve_trace_enable.argtypes = [vp, ip]
    L.ifdda_timing_enable.argtypes = [vp, ip]
    L.ifdda_timing_passes.argtypes = [vp, ip, ip]
    L.ifdda_use_specialized.argtypes = [vp, ip]
    L.ifdda_fuse_krylov.argtypes = [vp, ip]
    L.ifdda_zpass_registers.argtypes = [vp, ip]
    pd = C.POINTER(C.c_double)
    L.ifdda_cross_sections_device.argtypes = [vp, ip, vp, vp, vp, dp, pd, pd, vp]
    L.ifdda_farfield_device.argtypes = [vp, ip, vp, ip, dp, dp, dp, dp, dp, vp, vp,
                                        C.POINTER(C.c_int), pd, vp]
    L.ifdda_farfield_size.argtypes = [ip, dp, dp, C.POINTER(C.c_int), pd]
    L.ifdda_computegcfft2d_device.argtypes = [vp, ip, vp, vp, ip, dp, dp, pd, pd, pd, pd, vp]
    L.ifdda_runner_create.argtypes = [ip, dp, C.POINTER(vp)]
    L.ifdda_runner_submit.argtypes = [vp, C.POINTER(Job), C.POINTER(C.c_longlong)]
    L.ifdda_runner_wait.argtypes = [vp, C.c_longlong, pd, pd, C.POINTER(SolveInfo), pd]
    L.ifdda_runner_destroy.argtypes = [vp]
    L.ifdda_runner_pack_limit.argtypes = [vp, ip]
    L.ifdda_incident_device.argtypes = [vp, ip, ip, pd, pd, dp, dp, pd, vp, vp]
    L.ifdda_polarizability_device.argtypes = [vp, vp, dp, dp, dp, ip]
    L.ifdda_solve_batch.argtypes = [ip, C.POINTER(Job), C.POINTER(pd), C.POINTER(pd),
                                    C.POINTER(C.POINTER(SolveInfo)), C.POINTER(C.c_int)]
    L.ifdda_solve_batch_release.argtypes = []
    L.ifdda_microscope_incidences.argtypes = [vp, C.POINTER(MicroscopeOpts), C.POINTER(C.c_int), pd,
                                              ip, pd]
    L.ifdda_microsdf_incidences.argtypes = [vp, C.POINTER(MicroscopeOpts), C.POINTER(C.c_int), pd,
                                              ip, pd]
    L.ifdda_microsbf_device.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(MicroscopeOpts), pd,
                                        C.POINTER(C.c_int), C.POINTER(SolveInfo), ip, C.POINTER(C.c_int)]
    L.ifdda_microsdf_device.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(MicroscopeOpts), pd,
                                        C.POINTER(C.c_int), C.POINTER(SolveInfo), ip, C.POINTER(C.c_int)]
    L.ifdda_specialized_z.argtypes = [vp, ip, C.POINTER(C.c_int)]
    L.ifdda_timing_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.ifdda_force_device.argtypes = [vp, ip, vp, vp, vp, pd, dp, dp, pd, pd, vp, vp, vp]
    pi = C.POINTER(C.c_int)
    L.ifdda_autotune.argtypes = [vp, ip, ip, pi]
    L.ifdda_tile_widths.argtypes = [vp, pi]
    L.ifdda_tune_get.argtypes = [ip, ip, ip, ip, pi]
    L.ifdda_tune_set.argtypes = [ip, ip, ip, ip, pi]
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise IfddaError(f"{what} failed ({rc}): {lib().ifdda_last_error().decode()}")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class Solver:
    """One box (nx, ny, nz) on the GPU; mirrors the cdmlib hot-path routines."""

    def __init__(self, nx: int, ny: int, nz: int):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        h = C.c_void_p()
        _check(lib().ifdda_create(self.nx, self.ny, self.nz, C.byref(h)), "ifdda_create")
        self.h = h
        L = (C.c_int * 3)()
        _check(lib().ifdda_fft_lengths(self.h, L), "ifdda_fft_lengths")
        self.L = tuple(L)
        self.nd = None

    def close(self):
        if getattr(self, "h", None):
            lib().ifdda_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- Green tensor (calculgreenfft.f / calculgreenquadfft.f) -------------
    def greencalculfft(self, aretecube: float, k0: float, nquad: int = 0, tolinit: float = 1e-4):
        _check(lib().ifdda_green_build(self.h, float(aretecube), float(k0), int(nquad),
                                       float(tolinit)), "ifdda_green_build")

    def green_set(self, G):
        """G: complex128 [6, 2nz, 2ny, 2nx] in the reference layout."""
        comps = [_c128(G[c]).reshape(-1) for c in range(6)]
        arr = (C.c_void_p * 6)(*[_p(x) for x in comps])
        _check(lib().ifdda_green_set(self.h, arr), "ifdda_green_set")

    def green_get(self):
        L1, L2, L3 = self.L
        G = np.zeros((6, L3, L2, L1), np.complex128)
        arr = (C.c_void_p * 6)(*[_p(G[c]) for c in range(6)])
        _check(lib().ifdda_green_get(self.h, arr), "ifdda_green_get")
        return G

    # -- object ---------------------------------------------------------------
    def set_dipoles(self, tabdip=None, ndipole: int | None = None):
        """tabdip: int32 box map (1-based dipole or 0), None = dense box."""
        if tabdip is None:
            _check(lib().ifdda_set_dipoles(self.h, None, 0), "ifdda_set_dipoles")
            self.nd = self.nx * self.ny * self.nz
            return
        t = np.ascontiguousarray(tabdip, np.int32).reshape(-1)
        if t.size != self.nx * self.ny * self.nz:
            raise IfddaError("tabdip must hold nx*ny*nz entries")
        nd = int(t.max()) if ndipole is None else int(ndipole)
        self._tab = t
        _check(lib().ifdda_set_dipoles(self.h, _p(t), nd), "ifdda_set_dipoles")
        self.nd = nd

    def set_polarizability(self, pola):
        """pola: (nd,) isotropic complex, or (nmax,3,3) tensor (Fortran meaning [i,a,b])."""
        a = np.asarray(pola)
        if a.ndim == 1:
            a = _c128(a)
            _check(lib().ifdda_set_polarizability(self.h, _p(a), a.size, 1),
                   "ifdda_set_polarizability")
        else:
            f = np.ascontiguousarray(np.asfortranarray(a, np.complex128).reshape(-1, order="F"))
            _check(lib().ifdda_set_polarizability(self.h, _p(f), a.shape[0], 0),
                   "ifdda_set_polarizability")

    # -- matvec (produitfftmatvect3.f) ---------------------------------------
    def produitfftmatvect3(self, FFX, FFB):
        """Returns FFB - conv(FFX) (FFX already alpha-multiplied), FP64."""
        x = _c128(FFX).reshape(-1)
        y = _c128(FFB).reshape(-1).copy()
        _check(lib().ifdda_matvec(self.h, _p(x), _p(y)), "ifdda_matvec")
        return y

    def matvec_device(self, x_ptr, b_ptr, y_ptr, prec=64, apply_alpha=True, stream=None):
        _check(lib().ifdda_matvec_device(self.h, int(prec), int(apply_alpha), C.c_void_p(x_ptr),
                                         C.c_void_p(b_ptr), C.c_void_p(y_ptr),
                                         C.c_void_p(stream) if stream else None),
               "ifdda_matvec_device")

    # -- solve (inverserig.f) -------------------------------------------------
    @staticmethod
    def _opts(method, tol, precision, x0_zero, maxit, check_every):
        return SolveOpts(METHODS[method] if isinstance(method, str) else int(method),
                         int(precision), int(bool(x0_zero)), int(maxit), float(tol),
                         int(check_every))

    def inverserig(self, E_inc, method="GPBICG1", tol=1e-4, precision=64, x0_zero=False,
                   maxit=10000, check_every=0, nstop=None):
        """Solve (I - T alpha) E = E_inc.  Returns dict(FF, FFloc, tol1, nloop, ...)."""
        b = _c128(E_inc).reshape(-1)
        FF = np.zeros_like(b)
        FFloc = np.zeros_like(b)
        o = self._opts(method, tol, precision, x0_zero, maxit, check_every)
        info = SolveInfo()
        stop = C.c_int(0) if nstop is None else nstop
        _check(lib().ifdda_solve(self.h, C.byref(o), _p(b), _p(FF), _p(FFloc), C.byref(info),
                                 C.byref(stop)), "ifdda_solve")
        out = info.as_dict()
        out.update(FF=FF, FFloc=FFloc)
        return out

    def inverserig_multi(self, E_inc, method="GPBICG1", tol=1e-4, precision=64, x0_zero=False,
                         maxit=10000, check_every=0):
        """R incidences at once: E_inc complex128 [R, 3*nd].  Returns (FF, FFloc, infos)."""
        b = _c128(E_inc)
        R = b.shape[0]
        FF = np.zeros_like(b)
        FFloc = np.zeros_like(b)
        o = self._opts(method, tol, precision, x0_zero, maxit, check_every)
        infos = (SolveInfo * R)()
        stop = C.c_int(0)
        _check(lib().ifdda_solve_multi(self.h, C.byref(o), R, _p(b), _p(FF), _p(FFloc), infos,
                                       C.byref(stop)), "ifdda_solve_multi")
        return FF, FFloc, [i.as_dict() for i in infos]

    def solve_multi_device(self, R, Einc_ptr, FF_ptr, FFloc_ptr, method="GPBICG1", tol=1e-4,
                           precision=64, x0_zero=False, maxit=10000, check_every=0, stream=None):
        o = self._opts(method, tol, precision, x0_zero, maxit, check_every)
        infos = (SolveInfo * R)()
        stop = C.c_int(0)
        _check(lib().ifdda_solve_multi_device(self.h, C.byref(o), int(R), C.c_void_p(Einc_ptr),
                                              C.c_void_p(FF_ptr) if FF_ptr else None,
                                              C.c_void_p(FFloc_ptr) if FFloc_ptr else None,
                                              infos, C.byref(stop),
                                              C.c_void_p(stream) if stream else None),
               "ifdda_solve_multi_device")
        return [i.as_dict() for i in infos]

    def solve_device(self, Einc_ptr, FF_ptr, FFloc_ptr, method="GPBICG1", tol=1e-4, precision=64,
                     x0_zero=False, maxit=10000, check_every=0, stream=None):
        o = self._opts(method, tol, precision, x0_zero, maxit, check_every)
        info = SolveInfo()
        stop = C.c_int(0)
        _check(lib().ifdda_solve_device(self.h, C.byref(o), C.c_void_p(Einc_ptr),
                                        C.c_void_p(FF_ptr) if FF_ptr else None,
                                        C.c_void_p(FFloc_ptr) if FFloc_ptr else None,
                                        C.byref(info), C.byref(stop),
                                        C.c_void_p(stream) if stream else None),
               "ifdda_solve_device")
        return info.as_dict()

    def solve_device_noalloc(self, Einc_ptr, FF_ptr, FFloc_ptr, method="GPBICG1", tol=1e-4,
                             precision=64, x0_zero=False, maxit=10000, check_every=0, stream=None):
        """solve_device with the work in the library's fixed device array: no
        device allocation in the call (one RHS, up to IFDDA_NOALLOC_BYTES of work)."""
        o = self._opts(method, tol, precision, x0_zero, maxit, check_every)
        info = SolveInfo()
        stop = C.c_int(0)
        _check(lib().ifdda_solve_device_noalloc(self.h, C.byref(o), C.c_void_p(Einc_ptr),
                                                C.c_void_p(FF_ptr) if FF_ptr else None,
                                                C.c_void_p(FFloc_ptr) if FFloc_ptr else None,
                                                C.byref(info), C.byref(stop),
                                                C.c_void_p(stream) if stream else None),
               "ifdda_solve_device_noalloc")
        return info.as_dict()

    def timing_enable(self, cap: int):
        """Record HIP events around each matvec pass (up to cap passes)."""
        _check(lib().ifdda_timing_enable(self.h, int(cap)), "ifdda_timing_enable")

    # -- outputs from solved fields (cdmlib.f cross sections, diffractefft2d) --
    def cross_sections_device(self, R, E0_ptr, FF_ptr, FFloc_ptr, k0, I0, stream=None):
        """Returns [R, 3] (Cext, Cabs, Csca); I0 per RHS (|E0|^2)."""
        i0 = np.ascontiguousarray(np.broadcast_to(np.asarray(I0, np.float64), (R,)))
        out = np.zeros(3 * R)
        D = C.POINTER(C.c_double)
        _check(lib().ifdda_cross_sections_device(self.h, int(R), C.c_void_p(E0_ptr), C.c_void_p(FF_ptr),
                                                 C.c_void_p(FFloc_ptr), float(k0),
                                                 i0.ctypes.data_as(D), out.ctypes.data_as(D),
                                                 C.c_void_p(stream) if stream else None),
               "ifdda_cross_sections_device")
        return out.reshape(R, 3)

    def force_device(self, R, Einc_ptr, FF_ptr, FFloc_ptr, kvec, d, k0, fdip_ptr=None, tdip_ptr=None,
                     stream=None):
        """ifdda_force_device on device arrays; returns (force [R, 3], torque [R, 3])."""
        kv = np.ascontiguousarray(np.broadcast_to(np.asarray(kvec, np.float64).reshape(-1, 3), (R, 3)))
        fo, to = np.zeros(3 * R), np.zeros(3 * R)
        D = C.POINTER(C.c_double)
        _check(lib().ifdda_force_device(self.h, int(R), C.c_void_p(Einc_ptr), C.c_void_p(FF_ptr),
                                        C.c_void_p(FFloc_ptr), kv.ctypes.data_as(D), float(d), float(k0),
                                        fo.ctypes.data_as(D), to.ctypes.data_as(D),
                                        C.c_void_p(fdip_ptr) if fdip_ptr else None,
                                        C.c_void_p(tdip_ptr) if tdip_ptr else None,
                                        C.c_void_p(stream) if stream else None),
               "ifdda_force_device")
        return fo.reshape(R, 3), to.reshape(R, 3)

    def force(self, E_inc, FF, FFloc, kvec, d, k0, per_dipole=False):
        """Optical force and torque (ifdda_force_device) for R solved RHS given
        as host arrays [R, 3 nd] (or [3 nd]); kvec [R, 3].  Returns dict force
        [R, 3] (N), torque [R, 3] (N m) and, with per_dipole, fdip / tdip
        [R, nd, 3]."""
        import torch
        E = np.atleast_2d(_c128(E_inc))
        R = E.shape[0]
        dev = [torch.from_numpy(np.ascontiguousarray(np.atleast_2d(_c128(a))).view(np.float64).copy()).cuda()
               for a in (E_inc, FF, FFloc)]
        kv = np.ascontiguousarray(np.broadcast_to(np.asarray(kvec, np.float64).reshape(-1, 3), (R, 3)))
        fo, to = np.zeros(3 * R), np.zeros(3 * R)
        fd = td = None
        if per_dipole:
            fd = torch.zeros(R * 3 * self.nd, dtype=torch.float64, device="cuda")
            td = torch.zeros_like(fd)
        D = C.POINTER(C.c_double)
        _check(lib().ifdda_force_device(self.h, R, *[C.c_void_p(t.data_ptr()) for t in dev],
                                        kv.ctypes.data_as(D), float(d), float(k0), fo.ctypes.data_as(D),
                                        to.ctypes.data_as(D),
                                        C.c_void_p(fd.data_ptr()) if fd is not None else None,
                                        C.c_void_p(td.data_ptr()) if td is not None else None, None),
               "ifdda_force_device")
        out = dict(force=fo.reshape(R, 3), torque=to.reshape(R, 3))
        if per_dipole:
            torch.cuda.synchronize()
            out["fdip"] = fd.cpu().numpy().reshape(R, self.nd, 3)
            out["tdip"] = td.cpu().numpy().reshape(R, self.nd, 3)
        return out

    @staticmethod
    def farfield_size(nfft2d, k0, d):
        im = C.c_int(0)
        dk = C.c_double(0.0)
        _check(lib().ifdda_farfield_size(int(nfft2d), float(k0), float(d), C.byref(im), C.byref(dk)),
               "ifdda_farfield_size")
        return im.value, dk.value

    def farfield_device(self, R, FF_ptr, nfft2d, k0, d, origin, Epos_ptr, Eneg_ptr, stream=None):
        """Far field into device buffers of R * 3 * (2 imax + 1)^2 complex128
        each; returns (imax, deltak)."""
        im = C.c_int(0)
        dk = C.c_double(0.0)
        x0, y0, z0 = (float(v) for v in origin)
        _check(lib().ifdda_farfield_device(self.h, int(R), C.c_void_p(FF_ptr), int(nfft2d), float(k0),
                                           float(d), x0, y0, z0, C.c_void_p(Epos_ptr),
                                           C.c_void_p(Eneg_ptr), C.byref(im), C.byref(dk),
                                           C.c_void_p(stream) if stream else None),
               "ifdda_farfield_device")
        return im.value, dk.value

    def computegcfft2d_device(self, R, Epos_ptr, Eneg_ptr, imax, dk, k0, I0, theta, phi, stream=None):
        """Returns [R, 2] (Csca, g)."""
        D = C.POINTER(C.c_double)
        arr = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (R,)))
        i0, th, ph = arr(I0), arr(theta), arr(phi)
        out = np.zeros(2 * R)
        _check(lib().ifdda_computegcfft2d_device(self.h, int(R), C.c_void_p(Epos_ptr), C.c_void_p(Eneg_ptr),
                                                 int(imax), float(dk), float(k0), i0.ctypes.data_as(D),
                                                 th.ctypes.data_as(D), ph.ctypes.data_as(D),
                                                 out.ctypes.data_as(D),
                                                 C.c_void_p(stream) if stream else None),
               "ifdda_computegcfft2d_device")
        return out.reshape(R, 2)

    # -- brightfield microscope (microsbf.f, quick lens) ---------------------
    @staticmethod
    def microscope_opts(nfft2d, numaper, numaperinc, k0, d, origin, gross=100.0, zlens=0.0,
                        nside=1, npolainc=0, E0=1.0, max_batch=0, direct_lens=False,
                        incidences=None):
        """direct_lens: the nquicklens = 0 path (pupil field as a direct sum
        over dipoles, with its own pupil and condenser steps).  incidences:
        [(kx, ky, pp)] plane waves to illuminate with instead of the
        condenser lattice (brightfield)."""
        m = MicroscopeOpts()
        if incidences is not None:
            kxy = np.ascontiguousarray(incidences, np.float64).reshape(-1, 3)
            m._kxy = kxy  # kept alive with the options
            m.ninc_user = kxy.shape[0]
            m.kxy_user = kxy.ctypes.data
        m.nfft2d, m.numaper, m.numaperinc, m.gross, m.zlens = int(nfft2d), numaper, numaperinc, gross, zlens
        m.nside, m.npolainc, m.k0, m.d, m.max_batch = int(nside), int(npolainc), k0, d, int(max_batch)
        m.direct_lens = int(bool(direct_lens))
        m.E0[0], m.E0[1] = complex(E0).real, complex(E0).imag
        for i in range(3):
            m.origin[i] = float(origin[i])
        return m

    def microscope_incidences(self, m, dark=False):
        """[(kx, ky, pp)] of the condenser (brightfield lattice, or the
        darkfield ring), and the condenser step deltak."""
        name = "ifdda_microsdf_incidences" if dark else "ifdda_microscope_incidences"
        fn = getattr(lib(), name)
        n = C.c_int(0)
        dk = C.c_double(0.0)
        _check(fn(self.h, C.byref(m), C.byref(n), None, 0, C.byref(dk)), name)
        kxy = np.zeros(3 * n.value)
        D = C.POINTER(C.c_double)
        _check(fn(self.h, C.byref(m), C.byref(n), kxy.ctypes.data_as(D), n.value, C.byref(dk)), name)
        return kxy.reshape(-1, 3), dk.value

    def microsbf(self, m, method="GPBICG1", tol=1e-4, precision=64, x0_zero=False, maxit=10000,
                 check_every=0, nstop=None, dark=False):
        """Brightfield (microsbf.f) or, dark=True, darkfield (microsdf.f)
        microscope.  Returns (images [6, N, N]: scattered x, y, z, total x,
        y, z; infos)."""
        o = self._opts(method, tol, precision, x0_zero, maxit, check_every)
        count = m.ninc_user if (m.ninc_user > 0 and not dark) else \
            self.microscope_incidences(m, dark)[0].shape[0]
        infos = (SolveInfo * max(1, count))()
        N = m.nfft2d
        img = np.zeros(6 * N * N)
        n = C.c_int(0)
        stop = nstop if nstop is not None else C.c_int(0)
        name = "ifdda_microsdf_device" if dark else "ifdda_microsbf_device"
        _check(getattr(lib(), name)(self.h, C.byref(o), C.byref(m),
                                    img.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n),
                                    infos, len(infos), C.byref(stop)), name)
        return img.reshape(6, N, N), [infos[i].as_dict() for i in range(n.value)]

    def microsdf(self, m, **kw):
        """Darkfield microscope (microsdf.f), as microsbf."""
        return self.microsbf(m, dark=True, **kw)

    def incident_device(self, waves, k0, d, origin, wave="plane", E0=1.0, out=None):
        """R plane waves at the dipoles, formed on the device
        (ifdda_incident_device); returns a torch complex128 device tensor
        [R, 3 nd] viewed as float64 pairs (or fills `out`)."""
        import torch
        W = np.ascontiguousarray(waves, np.float64).reshape(-1, 3)
        R = W.shape[0]
        if out is None:
            out = torch.empty(R * 3 * self.nd * 2, dtype=torch.float64, device="cuda")
        e0 = (C.c_double * 2)(complex(E0).real, complex(E0).imag)
        org = (C.c_double * 3)(*[float(v) for v in origin])
        _check(lib().ifdda_incident_device(self.h, WAVES[wave], R, W.ctypes.data_as(C.POINTER(C.c_double)),
                                           e0, float(k0), float(d), org, C.c_void_p(out.data_ptr()), None),
               "ifdda_incident_device")
        return out

    def polarizability_device(self, eps, d, k0, method="RR", eps0=1.0):
        """alpha from eps per dipole on the device (ifdda_polarizability_device)."""
        e = _c128(eps).reshape(-1)
        _check(lib().ifdda_polarizability_device(self.h, _p(e), float(d), float(k0), float(eps0),
                                                 POLA[method]), "ifdda_polarizability_device")

    def use_specialized(self, on: bool):
        """Select compile-time (default) or generic z-pass kernels."""
        _check(lib().ifdda_use_specialized(self.h, int(bool(on))), "ifdda_use_specialized")

    def fuse_krylov(self, mask: int):
        """GPBICG work fused into the matvec ends: bit 0 FP64, bit 1 FP32; bit 2
        also forms p = r + beta (p - u) inside the scatter pass; bits 3 (FP32) and
        4 (FP64) the look-ahead loop (default 10: FP32 fused, look-ahead)."""
        _check(lib().ifdda_fuse_krylov(self.h, int(mask)), "ifdda_fuse_krylov")

    def zpass_registers(self, mask: int):
        """The register-staged z pass: bit 0 FP64 matvecs, bit 1 FP32."""
        _check(lib().ifdda_zpass_registers(self.h, int(mask)), "ifdda_zpass_registers")

    def specialized_z(self, prec: int = 64) -> bool:
        """Whether this grid's z pass has a compile-time kernel."""
        has = C.c_int(0)
        _check(lib().ifdda_specialized_z(self.h, int(prec), C.byref(has)), "ifdda_specialized_z")
        return bool(has.value)

    def autotune(self, prec: int = 0, reps: int = 0):
        """Time the matvec passes at every candidate tile width and keep the
        fastest per pass (ifdda_autotune).  Returns the widths now in use,
        {64: [P1..P5], 32: [P1..P5]}."""
        w = (C.c_int * 10)()
        _check(lib().ifdda_autotune(self.h, int(prec), int(reps), w), "ifdda_autotune")
        return {64: list(w[:5]), 32: list(w[5:])}

    def tile_widths(self):
        """The tile widths of this context, {64: [P1..P5], 32: [P1..P5]}."""
        w = (C.c_int * 10)()
        _check(lib().ifdda_tile_widths(self.h, w), "ifdda_tile_widths")
        return {64: list(w[:5]), 32: list(w[5:])}

    def timing_passes(self, mask: int, every: int = 1):
        """Bracket only the passes whose bit is set (bit p = P(p+1)), and of
        those every every-th launch."""
        _check(lib().ifdda_timing_passes(self.h, int(mask), int(every)), "ifdda_timing_passes")

    def timing_read(self):
        """Returns (ms_sum[2,5], count[2,5]): [FP64/FP32][P1..P5]; resets."""
        ms = (C.c_double * 10)()
        cnt = (C.c_int * 10)()
        _check(lib().ifdda_timing_read(self.h, ms, cnt), "ifdda_timing_read")
        return np.array(ms[:]).reshape(2, 5), np.array(cnt[:]).reshape(2, 5)

    def trace_enable(self, cap: int):
        _check(lib().ifdda_solve_trace_enable(self.h, int(cap)), "ifdda_solve_trace_enable")

    def trace(self, cap: int = 100000):
        out = np.zeros(6 * cap, np.complex128)
        n = C.c_int(0)
        _check(lib().ifdda_solve_trace(self.h, int(cap), _p(out), C.byref(n)), "ifdda_solve_trace")
        return out.reshape(-1, 6)[: n.value].copy()


def tune_get(nx, ny, nz, prec):
    """The autotuner's widths (P1..P5) recorded for a box and precision, or
    None."""
    w = (C.c_int * 5)()
    rc = lib().ifdda_tune_get(int(nx), int(ny), int(nz), int(prec), w)
    if rc not in (0, 1):
        _check(rc, "ifdda_tune_get")
    return list(w[:]) if rc == 1 else None


def tune_set(nx, ny, nz, prec, widths):
    """Record tile widths (P1..P5) for a box and precision, e.g. saved from an
    earlier run; contexts created afterwards use them."""
    w = (C.c_int * 5)(*[int(v) for v in widths])
    _check(lib().ifdda_tune_set(int(nx), int(ny), int(nz), int(prec), w), "ifdda_tune_set")


class Runner:
    """Many independent solves at once (ifdda_runner_*): a pool of workers
    with their own contexts and streams, admission by device footprint."""

    def __init__(self, nworkers: int = 8, mem_budget: float = 240e9, max_members: int = 64):
        self.h = C.c_void_p()
        _check(lib().ifdda_runner_create(int(nworkers), float(mem_budget), C.byref(self.h)),
               "ifdda_runner_create")
        _check(lib().ifdda_runner_pack_limit(self.h, int(max_members)), "ifdda_runner_pack_limit")
        self._shape = {}

    def close(self):
        if self.h:
            lib().ifdda_runner_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, nx, ny, nz, d, k0, alpha=None, E_inc=None, tabdip=None, method="GPBICG1",
               tol=1e-4, precision=32, x0_zero=True, maxit=10000, nquad=0, tolinit=1e-4, **kw):
        """E_inc: [R, 3 nd] or [3 nd] complex, or waves=[R, 3] parameters
        (wave="plane" | "circular" | "kxky", E0, origin) formed on the
        device; alpha: nd complex, or eps=nd complex (pola="RR", eps0).
        Returns a ticket (the arrays are copied before this returns)."""
        j, keep, R, nd = _job(nx, ny, nz, d, k0, Solver._opts(method, tol, precision, x0_zero, maxit, 0),
                              alpha, E_inc, tabdip, nquad, tolinit, **kw)
        t = C.c_longlong(0)
        _check(lib().ifdda_runner_submit(self.h, C.byref(j), C.byref(t)), "ifdda_runner_submit")
        self._shape[t.value] = (R, nd)
        return t.value

    def wait(self, ticket):
        R, nd = self._shape.pop(ticket)
        FL = np.zeros((R, 3 * nd), np.complex128)
        FF = np.zeros_like(FL)
        infos = (SolveInfo * R)()
        times = np.zeros(3)
        D = C.POINTER(C.c_double)
        _check(lib().ifdda_runner_wait(self.h, C.c_longlong(ticket), FL.ctypes.data_as(D),
                                       FF.ctypes.data_as(D), infos, times.ctypes.data_as(D)),
               "ifdda_runner_wait")
        return {"FFloc": FL, "FF": FF, "infos": [i.as_dict() for i in infos],
                "queued_s": times[0], "run_s": times[1], "latency_s": times[2]}


def _job(nx, ny, nz, d, k0, opts, alpha=None, E_inc=None, tabdip=None, nquad=0, tolinit=1e-4,
         eps=None, pola="RR", eps0=1.0, waves=None, wave="plane", E0=1.0, origin=(0.0, 0.0, 0.0)):
    """An ifdda_job and the arrays it points at (keep them alive): alpha or
    eps (+ pola, eps0); E_inc or waves (+ wave, E0, origin)."""
    j = Job()
    j.nx, j.ny, j.nz, j.d, j.k0 = int(nx), int(ny), int(nz), float(d), float(k0)
    j.nquad, j.tolinit = int(nquad), float(tolinit)
    keep = []
    if alpha is not None:
        al = _c128(alpha).reshape(-1)
        j.alpha = al.ctypes.data
    else:
        al = _c128(eps).reshape(-1)
        j.eps = al.ctypes.data
        j.pola, j.eps0 = POLA[pola], float(eps0)
    keep.append(al)
    if E_inc is not None:
        E = _c128(E_inc)
        E = E.reshape(1, -1) if E.ndim == 1 else E.reshape(E.shape[0], -1)
        j.E_inc = E.ctypes.data
    else:
        E = np.ascontiguousarray(waves, np.float64).reshape(-1, 3)
        j.waves = E.ctypes.data
        j.wave_kind = WAVES[wave]
        E0 = complex(E0)
        j.E0[0], j.E0[1] = E0.real, E0.imag
        for a in range(3):
            j.origin[a] = float(origin[a])
    keep.append(E)
    j.nd, j.nrhs = al.size, E.shape[0]
    if tabdip is not None:
        tab = np.ascontiguousarray(tabdip, np.int32).reshape(-1)
        j.tabdip = tab.ctypes.data
        keep.append(tab)
    j.opts = opts
    return j, keep, E.shape[0], al.size


def solve_batch_release():
    """Free the workspace solve_batch keeps between calls."""
    _check(lib().ifdda_solve_batch_release(), "ifdda_solve_batch_release")


def solve_batch(problems, method="GPBICG1", tol=1e-4, precision=32, x0_zero=True, maxit=10000,
                check_every=0):
    """Independent problems solved together (ifdda_solve_batch): every
    right-hand side of every problem shares the matvec launches and one
    lockstep Krylov solve.  problems: iterable of dicts with keys nx, ny, nz,
    d, k0, alpha (or eps, pola, eps0), E_inc ([R, 3 nd] or [3 nd]; or waves
    [R, 3], wave, E0, origin) and optionally tabdip, nquad.
    Returns one dict per problem: FFloc, FF ([R, 3 nd]) and infos."""
    opts = Solver._opts(method, tol, precision, x0_zero, maxit, check_every)
    jobs, keep, outs = [], [], []
    for p in problems:
        extra = {k: p[k] for k in ("eps", "pola", "eps0", "waves", "wave", "E0", "origin") if k in p}
        j, k, R, nd = _job(p["nx"], p["ny"], p["nz"], p["d"], p["k0"], opts, p.get("alpha"),
                           p.get("E_inc"), p.get("tabdip"), p.get("nquad", 0), **extra)
        jobs.append(j)
        keep.append(k)
        outs.append({"FFloc": np.empty((R, 3 * nd), np.complex128),
                     "FF": np.empty((R, 3 * nd), np.complex128), "_info": (SolveInfo * R)()})
    n = len(jobs)
    D = C.POINTER(C.c_double)
    arr = (Job * max(n, 1))(*jobs)
    FL = (D * max(n, 1))(*[o["FFloc"].ctypes.data_as(D) for o in outs])
    FF = (D * max(n, 1))(*[o["FF"].ctypes.data_as(D) for o in outs])
    IN = (C.POINTER(SolveInfo) * max(n, 1))(*[C.cast(o["_info"], C.POINTER(SolveInfo)) for o in outs])
    stop = C.c_int(0)
    _check(lib().ifdda_solve_batch(n, arr, FL, FF, IN, C.byref(stop)), "ifdda_solve_batch")
    for o in outs:
        o["infos"] = [i.as_dict() for i in o.pop("_info")]
    return outs
