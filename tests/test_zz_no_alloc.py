"""ifdda_solve_device_noalloc (work in the library's fixed device array, no
device allocation in the call) against ifdda_solve_device on the same
inputs: fields, moments, iteration and matvec counts bit-equal, at both
precisions, on two calls in a row whose second problem is smaller; a
problem larger than the array is refused with the sizes in the message."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(gpu, s, k0, alpha):
    sv = gpu.Solver(s.nx, s.ny, s.nz)
    sv.greencalculfft(s.d, k0)
    sv.set_dipoles(s.tabdip)
    sv.set_polarizability(alpha)
    return sv


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.complex128).view(np.float64).copy()).cuda()


def _host(t):
    return t.cpu().numpy().view(np.complex128)


def _raw_noalloc(gpu, sv, E, FF, FL, precision, tol):
    """The entry point through the C ABI (ctypes on libifdda_hip.so)."""
    o = gpu.Solver._opts("GPBICG1", tol, precision, True, 10000, 0)
    info = gpu.capi.SolveInfo()
    stop = C.c_int(0)
    rc = gpu.lib().ifdda_solve_device_noalloc(sv.h, C.byref(o), C.c_void_p(E.data_ptr()),
                                              C.c_void_p(FF.data_ptr()), C.c_void_p(FL.data_ptr()),
                                              C.byref(info), C.byref(stop), None)
    return rc, info.as_dict()


def _compare(gpu, sv, E_host, precision, tol):
    import torch
    E = _dev(E_host)
    FF0, FL0 = torch.zeros_like(E), torch.zeros_like(E)
    ref = sv.solve_device(E.data_ptr(), FF0.data_ptr(), FL0.data_ptr(), tol=tol, precision=precision,
                          x0_zero=True)
    FF1, FL1 = torch.full_like(E, 7.0), torch.full_like(E, -3.0)
    rc, got = _raw_noalloc(gpu, sv, E, FF1, FL1, precision, tol)
    torch.cuda.synchronize()
    assert rc == 0, gpu.lib().ifdda_last_error()
    assert ref["nstop"] == 0 and got["nstop"] == 0
    assert got["nloop"] == ref["nloop"] and got["ncompte"] == ref["ncompte"]
    assert got["tol1"] == ref["tol1"] and got["tol1"] <= tol
    assert np.array_equal(_host(FL1), _host(FL0))
    assert np.array_equal(_host(FF1), _host(FF0))


@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_equals_solve_device_two_calls(gpu, precision):
    from ifdda import scene
    lam = 632.8e-9
    k0 = 2 * np.pi / lam
    s, k0, alpha, E = scene.config1()  # 4224 dipoles
    small = scene.sphere(8, 80e-9, 3.0 + 0.2j)  # 280 dipoles: the second, smaller call
    a_small = scene.polarizability(small.eps, small.d, k0, "RR")
    E_small = scene.plane_wave(small.pos, k0, 30.0, 10.0, pp=0.5)
    print(f"noalloc precision {precision}: {s.nd} then {small.nd} dipoles", flush=True)
    with _setup(gpu, s, k0, alpha) as sv1, _setup(gpu, small, k0, a_small) as sv2:
        _compare(gpu, sv1, E, precision, 1e-6)
        _compare(gpu, sv2, E_small, precision, 1e-6)


def test_noalloc_refuses_a_problem_above_its_array(gpu):
    """A 48^3 dense cube (110,592 dipoles) needs ~74 MB of FP64 work."""
    import torch
    from ifdda import scene
    lam = 632.8e-9
    k0 = 2 * np.pi / lam
    s = scene.cube(48, 48 * 20e-9, 2.25 + 0j)
    alpha = scene.polarizability(s.eps, s.d, k0, "RR")
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.greencalculfft(s.d, k0)
        sv.set_dipoles(None)
        sv.set_polarizability(alpha)
        E = torch.zeros(2 * 3 * s.nd, dtype=torch.float64, device="cuda")
        rc, _ = _raw_noalloc(gpu, sv, E, torch.zeros_like(E), torch.zeros_like(E), 64, 1e-4)
        assert rc == 1
        assert b"above the fixed" in gpu.lib().ifdda_last_error()
