"""Sparse maps whose box lines and whole planes hold no dipole: the matvec
passes then skip those lines of the x-transformed grid (the scatter writes
none, the y transform reads none back, the inverse y transform writes none,
the gather reads none) and gathering tiles without a dipole only write their
zero inner-product partials.  The results must not move: FP64 matvec within
1e-12 of the oracle (produitfftmatvect3.f:37-127), FP32 within 2e-5, GPBICG
iteration counts equal to the oracle's and fields within 1e-10 (FP64), mixed
precision within the refinement tolerance (gpbicg.f, inverserig.f:36-105)."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

LAM = 632.8e-9
K0 = 2 * np.pi / LAM
D = 20e-9


def _two_spheres(nx=24, ny=24, nz=40, r=5.0):
    """Two spheres, one near each z end: planes 12..27 and most lines empty."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    inside = ((i - 7.5) ** 2 + (j - 8.0) ** 2 + (k - 6.0) ** 2 <= r * r) | \
             ((i - 15.5) ** 2 + (j - 15.0) ** 2 + (k - 33.0) ** 2 <= r * r)
    inside = inside.reshape(-1)
    tab = np.zeros(inside.size, np.int32)
    nd = int(inside.sum())
    tab[inside] = np.arange(1, nd + 1, dtype=np.int32)
    return tab, nd


def _line_stats(tab, nx, ny, nz):
    occ = (tab.reshape(nz, ny, nx) > 0).any(axis=2)
    return float(occ.mean()), int((~occ.any(axis=1)).sum())


def test_matvec_with_empty_lines_and_planes(gpu, oracle):
    nx, ny, nz = 24, 24, 40
    tab, nd = _two_spheres(nx, ny, nz)
    frac, empty_planes = _line_stats(tab, nx, ny, nz)
    assert frac < 0.3 and empty_planes >= 10  # the masks have work to skip
    rng = np.random.default_rng(5)
    x = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    b = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    with gpu.Solver(nx, ny, nz) as sv:
        print(f"sparse lines {nx}x{ny}x{nz}: {nd} dipoles", flush=True)
        sv.greencalculfft(D, K0)
        G = oracle.green_fft(nx, ny, nz, sv.L, D, K0)
        sv.set_dipoles(tab)
        ref = oracle.matvec(G, tab, nx, ny, nz, x, b)
        scale = np.abs(ref - b).max()
        y = sv.produitfftmatvect3(x, b)
        assert relerr(y, ref, scale) < 1e-12
        # the same product again (the scratch grids now hold the first call's
        # values on every line the masks skip)
        y2 = sv.produitfftmatvect3(x, b)
        assert np.array_equal(y, y2)
        import torch
        x32 = torch.from_numpy(x.astype(np.complex64).view(np.float32).copy()).cuda()
        b32 = torch.from_numpy(b.astype(np.complex64).view(np.float32).copy()).cuda()
        y32 = torch.zeros_like(x32)
        sv.matvec_device(x32.data_ptr(), b32.data_ptr(), y32.data_ptr(), prec=32, apply_alpha=False)
        torch.cuda.synchronize()
        y = y32.cpu().numpy().view(np.complex64).astype(np.complex128)
        assert relerr(y, ref, scale) < 2e-5
        del x32, b32, y32


def test_random_sparse_map_matvec(gpu, oracle):
    """A random map with ~80% empty x lines, some planes empty."""
    nx, ny, nz = 12, 16, 10
    rng = np.random.default_rng(11)
    N = nx * ny * nz
    occ = rng.random(N) < 0.02
    occ[0] = True
    nd = int(occ.sum())
    tab = np.zeros(N, np.int32)
    tab[occ] = np.arange(1, nd + 1, dtype=np.int32)
    frac, _ = _line_stats(tab, nx, ny, nz)
    assert frac < 0.4
    x = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    b = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    with gpu.Solver(nx, ny, nz) as sv:
        sv.greencalculfft(D, K0)
        G = oracle.green_fft(nx, ny, nz, sv.L, D, K0)
        sv.set_dipoles(tab)
        ref = oracle.matvec(G, tab, nx, ny, nz, x, b)
        y = sv.produitfftmatvect3(x, b)
        assert relerr(y, ref, np.abs(ref - b).max()) < 1e-12


def test_solves_with_empty_lines(gpu, oracle):
    from ifdda import scene
    nx, ny, nz = 24, 24, 40
    tab, nd = _two_spheres(nx, ny, nz)
    eps = np.full(nd, 2.25 + 0.01j)
    alpha = scene.polarizability(eps, D, K0, "RR")
    cells = np.nonzero(tab)[0]
    pos = np.stack([(cells % nx) * D, ((cells // nx) % ny) * D, (cells // (nx * ny)) * D], axis=1)
    E = scene.plane_wave(pos, K0, 30.0, 0.0, pp=1.0)
    with gpu.Solver(nx, ny, nz) as sv:
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        sv.set_polarizability(alpha)
        G = oracle.green_fft(nx, ny, nz, sv.L, D, K0)
        pol = np.zeros((nx * ny * nz, 3, 3), np.complex128)
        for a in range(3):
            pol[:nd, a, a] = alpha
        ref = oracle.inverserig(G, tab, nx, ny, nz, pol, E, "GPBICG1", 1e-4)
        r = sv.inverserig(E, "GPBICG1", 1e-4)
        assert r["nstop"] == 0 and r["nloop"] == ref["nloop"]
        assert relerr(r["FFloc"], ref["FFloc"]) < 1e-10
        # mixed precision: the FP32 inner solves run the fused gathers, whose
        # tiles without a dipole write zero partials
        acc = sv.inverserig(E, "GPBICG1", 1e-10)
        for tol in (1e-4, 1e-8):
            m = sv.inverserig(E, "GPBICG1", tol, precision=32, x0_zero=True)
            assert m["nstop"] == 0 and m["tol1"] <= tol
            assert relerr(m["FFloc"], acc["FFloc"]) < 10 * tol
