"""GPU parity of the Green tensor and the matvec against the oracle and the
reference's golden vectors (FP64: 1e-12 relative; FP32: 2e-5 relative)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT, golden, relerr

pytestmark = pytest.mark.gpu

LAM = 632.8e-9
K0 = 2 * np.pi / LAM
D = 20e-9


def _case(nx, ny, nz, seed, frac=0.6):
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    mask = rng.random(N) < frac
    nd = int(mask.sum())
    tab = np.zeros(N, np.int32)
    tab[mask] = np.arange(1, nd + 1, dtype=np.int32)
    x = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    b = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    return tab, x, b


def test_green_matches_reference_fixture(gpu):
    g = golden("green_654.npz")
    for nq in (0, 1, 2):
        with gpu.Solver(6, 5, 4) as sv:
            print(f"green 6x5x4 nquad={nq}", flush=True)
            sv.greencalculfft(D, K0, nquad=nq)
            G = sv.green_get()
            ref = g[f"G_nquad{nq}"]
            assert G.shape == ref.shape
            assert relerr(G, ref) < 1e-12


def test_matvec_matches_reference_fixture(gpu):
    f = golden("matvec_654.npz")
    with gpu.Solver(6, 5, 4) as sv:
        sv.greencalculfft(D, K0)
        sv.set_dipoles(f["tabdip"])
        y = sv.produitfftmatvect3(f["FFX"], f["FFB"])
        assert relerr(y, f["out3"], np.abs(f["out3"] - f["FFB"]).max()) < 1e-12
        sv.set_dipoles(None)
        y = sv.produitfftmatvect3(f["FFX_dense"], f["FFB_dense"])
        assert relerr(y, f["out_opt"], np.abs(f["out_opt"] - f["FFB_dense"]).max()) < 1e-12


def test_matvec_mixed_radix_fixture(gpu):
    f = golden("green_91012.npz")
    with gpu.Solver(9, 10, 12) as sv:
        assert sv.L == (18, 20, 24)
        sv.greencalculfft(D, K0)
        assert relerr(sv.green_get(), f["G"]) < 1e-12
        sv.set_dipoles(f["tabdip"])
        y = sv.produitfftmatvect3(f["FFX"], f["FFB"])
        assert relerr(y, f["out3"], np.abs(f["out3"] - f["FFB"]).max()) < 1e-12


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 1), (7, 5, 3), (8, 8, 8), (13, 4, 11),
                                   (20, 20, 20), (25, 3, 16)])
def test_matvec_vs_oracle_shapes(gpu, oracle, shape):
    """Box sizes whose 2n is not 2,3,5-smooth (7, 13, 11) use a larger padded
    length; the product on the box is unchanged."""
    nx, ny, nz = shape
    print(f"matvec shape {shape}", flush=True)
    tab, x, b = _case(nx, ny, nz, sum(shape))
    G = oracle.green_fft(nx, ny, nz, (2 * nx, 2 * ny, 2 * nz), D, K0)
    ref = oracle.matvec(G, tab, nx, ny, nz, x, b)
    with gpu.Solver(nx, ny, nz) as sv:
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        y = sv.produitfftmatvect3(x, b)
    assert relerr(y, ref, max(np.abs(ref - b).max(), 1e-300)) < 1e-12


@pytest.mark.parametrize("shape", [(9, 7, 6), (20, 20, 20)])
def test_identity_map_runs_as_dense_box(gpu, oracle, shape):
    """A Tabdip mapping every cell to dipole cell+1 is the dense box of
    produitfftmatvectopt.f: the device drops the map; the product equals the
    oracle's with that map, and the dense (map-less) call."""
    nx, ny, nz = shape
    tab, x, b = _case(nx, ny, nz, 21, frac=1.0)
    assert (tab == np.arange(1, tab.size + 1)).all()
    G = oracle.green_fft(nx, ny, nz, (2 * nx, 2 * ny, 2 * nz), D, K0)
    ref = oracle.matvec(G, tab, nx, ny, nz, x, b)
    with gpu.Solver(nx, ny, nz) as sv:
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        y = sv.produitfftmatvect3(x, b)
        sv.set_dipoles(None)
        yd = sv.produitfftmatvect3(x, b)
    scale = np.abs(ref - b).max()
    assert relerr(y, ref, scale) < 1e-12
    assert relerr(yd, y, scale) < 1e-15


def test_matvec_direct_sum(gpu, oracle):
    """The FFT convolution equals the O(Nd^2) dipole sum."""
    nx, ny, nz = 7, 6, 5
    tab, x, b = _case(nx, ny, nz, 7)
    ref = oracle.matvec_direct(tab, nx, ny, nz, D, K0, x, b)
    with gpu.Solver(nx, ny, nz) as sv:
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        y = sv.produitfftmatvect3(x, b)
    assert relerr(y, ref, np.abs(ref - b).max()) < 1e-12


def test_green_set_roundtrip_and_symmetry_check(gpu):
    g = golden("green_654.npz")
    G = g["G_nquad2"]
    with gpu.Solver(6, 5, 4) as sv:
        sv.green_set(G)
        assert relerr(sv.green_get(), G) < 1e-12
        bad = G.copy()
        bad[1] = np.roll(bad[1], 1, axis=2)  # breaks the odd parity of xy
        with pytest.raises(gpu.IfddaError):
            sv.green_set(bad)


def _dev(t):
    return t.data_ptr()


def test_matvec_device_fp32_and_alpha(gpu, oracle):
    """Device entry: y = x - T(alpha x); FP64 to 1e-12, FP32 to 2e-5."""
    import torch
    nx, ny, nz = 12, 10, 9
    tab, x, _ = _case(nx, ny, nz, 3, frac=0.7)
    nd = int(tab.max())
    rng = np.random.default_rng(4)
    alpha = (1 + rng.random(nd)) * 1e-24 + 1j * rng.random(nd) * 1e-26
    G = oracle.green_fft(nx, ny, nz, (2 * nx, 2 * ny, 2 * nz), D, K0)
    ax = (alpha[:, None] * x.reshape(-1, 3)).reshape(-1)
    ref = oracle.matvec(G, tab, nx, ny, nz, ax, x)
    with gpu.Solver(nx, ny, nz) as sv:
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        sv.set_polarizability(alpha)
        xd = torch.from_numpy(x.view(np.float64).copy()).cuda()
        yd = torch.zeros_like(xd)
        sv.matvec_device(_dev(xd), _dev(xd), _dev(yd), prec=64)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(np.complex128)
        assert relerr(y, ref, np.abs(ref - x).max()) < 1e-12
        x32 = torch.from_numpy(x.astype(np.complex64).view(np.float32).copy()).cuda()
        y32 = torch.zeros_like(x32)
        sv.matvec_device(_dev(x32), _dev(x32), _dev(y32), prec=32)
        torch.cuda.synchronize()
        y = y32.cpu().numpy().view(np.complex64).astype(np.complex128)
        assert relerr(y, ref, np.abs(ref - x).max()) < 2e-5


def test_matvec_tensor_alpha(gpu, oracle):
    """Anisotropic 3x3 polarisability (the full polarisa(nmax,3,3) path)."""
    import torch
    nx, ny, nz = 6, 7, 5
    tab, x, _ = _case(nx, ny, nz, 11, frac=0.8)
    nd = int(tab.max())
    rng = np.random.default_rng(5)
    pol = (rng.standard_normal((nd, 3, 3)) + 1j * rng.standard_normal((nd, 3, 3))) * 1e-24
    G = oracle.green_fft(nx, ny, nz, (2 * nx, 2 * ny, 2 * nz), D, K0)
    ax = np.einsum("iab,ib->ia", pol, x.reshape(-1, 3)).reshape(-1)
    ref = oracle.matvec(G, tab, nx, ny, nz, ax, x)
    with gpu.Solver(nx, ny, nz) as sv:
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        sv.set_polarizability(pol)
        xd = torch.from_numpy(x.view(np.float64).copy()).cuda()
        yd = torch.zeros_like(xd)
        sv.matvec_device(_dev(xd), _dev(xd), _dev(yd), prec=64)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(np.complex128)
    assert relerr(y, ref, np.abs(ref - x).max()) < 1e-12


def test_fortran_dropin_matvec(gpu):
    """The reference's own call shape, through libifdda_dropin.so."""
    f = golden("matvec_654.npz")
    g = golden("green_654.npz")
    L = C.CDLL(os.path.join(ROOT, "if-dda_amd", "libifdda_dropin.so"))
    nx, ny, nz = 6, 5, 4
    nx2, ny2, nz2 = 12, 10, 8
    M = nx2 * ny2 * nz2
    Gs = [np.ascontiguousarray(g["G_nquad0"][c]).reshape(-1) for c in range(6)]
    V = [np.zeros(M, np.complex128) for _ in range(3)]
    tab = np.ascontiguousarray(f["tabdip"], np.int32)
    out = f["FFB"].copy()
    ri = lambda v: C.byref(C.c_int(v))  # noqa: E731
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    nd = int(tab.max())
    L.produitfftmatvect3_(*[P(x) for x in Gs], *[P(v) for v in V], P(tab), ri(M), ri(M), ri(nd),
                          ri(nx * ny * nz), ri(nx), ri(ny), ri(nz), ri(nx), ri(ny), ri(nz),
                          ri(nx2), ri(ny2), ri(nx2 * ny2), ri(nz2), P(f["FFX"].copy()), P(out),
                          C.byref(C.c_int64(0)), C.byref(C.c_int64(0)))
    assert relerr(out, f["out3"], np.abs(f["out3"] - f["FFB"]).max()) < 1e-12


@pytest.mark.parametrize("n", [20, 64, 96, 100])
def test_specialized_z_pass_equals_generic(gpu, n):
    """The kernels compiled for a fixed FFT length (z pass: the 40^3, 128^3,
    192^3 and 200^3 grids; FP64 x passes and y passes: 128, 192, 200) give the generic
    kernels' product: FP64 to 1e-14, FP32 to 1e-6 relative to |T alpha x|."""
    import torch
    tab, x, _ = _case(n, n, n, 10 + n, frac=0.5)
    nd = int(tab.max())
    rng = np.random.default_rng(n)
    alpha = (1 + rng.random(nd)) * 1e-24 + 1j * rng.random(nd) * 1e-26
    with gpu.Solver(n, n, n) as sv:
        print(f"specialized z pass {n}^3 L={sv.L}", flush=True)
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        sv.set_polarizability(alpha)
        for prec, tol in ((64, 1e-14), (32, 1e-6)):
            assert sv.specialized_z(prec)
            cdt, rdt = (np.complex128, np.float64) if prec == 64 else (np.complex64, np.float32)
            xd = torch.from_numpy(x.astype(cdt).view(rdt).copy()).cuda()
            outs = []
            for on in (True, False):
                sv.use_specialized(on)
                yd = torch.zeros_like(xd)
                sv.matvec_device(_dev(xd), _dev(xd), _dev(yd), prec=prec)
                torch.cuda.synchronize()
                y = yd.cpu().numpy().view(cdt).astype(np.complex128)
                outs.append(y - x.astype(cdt).astype(np.complex128))  # -T alpha x
            sv.use_specialized(True)
            assert relerr(outs[0], outs[1]) < tol, (prec, relerr(outs[0], outs[1]))


@pytest.mark.parametrize("n", [64, 96, 100])
def test_register_z_pass_equals_lds_z_pass(gpu, n):
    """The register-staged z pass (8-point steps across threads, L/8-point
    steps in one thread's registers) gives the LDS-staged compile-time z
    pass's product on the 128^3, 192^3 and 200^3 grids: FP64 to 1e-13, FP32
    to 1e-6 relative to |T alpha x|."""
    import torch
    tab, x, _ = _case(n, n, n, 20 + n, frac=0.5)
    nd = int(tab.max())
    rng = np.random.default_rng(n + 1)
    alpha = (1 + rng.random(nd)) * 1e-24 + 1j * rng.random(nd) * 1e-26
    with gpu.Solver(n, n, n) as sv:
        print(f"register z pass {n}^3 L={sv.L}", flush=True)
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        sv.set_polarizability(alpha)
        for prec, tol in ((64, 1e-13), (32, 1e-6)):
            cdt, rdt = (np.complex128, np.float64) if prec == 64 else (np.complex64, np.float32)
            xd = torch.from_numpy(x.astype(cdt).view(rdt).copy()).cuda()
            outs = []
            for mask in (3, 0):
                sv.zpass_registers(mask)
                yd = torch.zeros_like(xd)
                sv.matvec_device(_dev(xd), _dev(xd), _dev(yd), prec=prec)
                torch.cuda.synchronize()
                y = yd.cpu().numpy().view(cdt).astype(np.complex128)
                outs.append(y - x.astype(cdt).astype(np.complex128))  # -T alpha x
            sv.zpass_registers(0)
            err = relerr(outs[0], outs[1])
            assert err < tol, (prec, err)


def test_register_z_pass_batch_solve(gpu):
    """Four right-hand sides of the 20-sphere brightfield object solved at
    once on its 128^3 grid (the batched z pass, right-hand sides in groups per
    tile) with the register-staged z pass: FP64 iteration counts and fields
    as with the LDS-staged one."""
    from ifdda import scene
    s, k0, alpha, E = scene.config3(2)
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.greencalculfft(s.d, k0)
        sv.set_dipoles(s.tabdip)
        sv.set_polarizability(alpha)
        res = {}
        for mask in (3, 0):
            sv.zpass_registers(mask)
            print(f"register z pass batch solve mask {mask}", flush=True)
            res[mask] = sv.inverserig_multi(E, "GPBICG1", 1e-6, precision=64, x0_zero=True)
        sv.zpass_registers(0)
        _, FLa, ia = res[3]
        _, FLb, ib = res[0]
        for r in range(E.shape[0]):
            assert ia[r]["nstop"] == 0 and ia[r]["tol1"] <= 1e-6
            assert abs(ia[r]["nloop"] - ib[r]["nloop"]) <= 1
            assert relerr(FLa[r], FLb[r]) < 1e-9
