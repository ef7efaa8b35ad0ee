"""GPU checks of the tile-width autotuner: tuning keeps the matvec's results
(bit for bit with the generic kernels, to rounding where a z-pass width
switches between the compile-time and the generic kernel), reports widths
that the context then uses, and later contexts of the same box start from
the recorded widths."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

LAM = 632.8e-9
K0 = 2 * np.pi / LAM
D = 20e-9


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    N = n[0] * n[1] * n[2]
    mask = rng.random(N) < 0.6
    nd = int(mask.sum())
    tab = np.zeros(N, np.int32)
    tab[mask] = np.arange(1, nd + 1, dtype=np.int32)
    x = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    alpha = (1 + rng.random(nd)) * 1e-24 + 1j * rng.random(nd) * 1e-26
    return tab, x, alpha


def _matvecs(sv, x):
    """y = x - T(alpha x) in FP64 and FP32 through the device entry."""
    import torch
    out = {}
    for prec, cdt, rdt in ((64, np.complex128, np.float64), (32, np.complex64, np.float32)):
        xd = torch.from_numpy(x.astype(cdt).view(rdt).copy()).cuda()
        yd = torch.zeros_like(xd)
        sv.matvec_device(xd.data_ptr(), xd.data_ptr(), yd.data_ptr(), prec=prec)
        torch.cuda.synchronize()
        out[prec] = yd.cpu().numpy().view(cdt).copy()
    return out


@pytest.mark.parametrize("n,specialized", [((20, 20, 20), False), ((20, 20, 20), True),
                                           ((24, 18, 30), False)])
def test_autotune_keeps_results(gpu, n, specialized):
    tab, x, alpha = _problem(n, sum(n))
    with gpu.Solver(*n) as sv:
        print(f"autotune {n} specialized={specialized}", flush=True)
        sv.use_specialized(specialized)
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        sv.set_polarizability(alpha)
        before = sv.tile_widths()
        y0 = _matvecs(sv, x)
        chosen = sv.autotune(reps=3)
        assert chosen == sv.tile_widths()
        assert all(b >= 1 for p in (64, 32) for b in chosen[p])
        print(f"  widths {before} -> {chosen}", flush=True)
        y1 = _matvecs(sv, x)
        for prec, tol in ((64, 1e-14), (32, 1e-6)):
            if not specialized:
                assert np.array_equal(y0[prec], y1[prec]), prec
            else:
                t0 = y0[prec].astype(np.complex128) - x.astype(y0[prec].dtype)
                t1 = y1[prec].astype(np.complex128) - x.astype(y1[prec].dtype)
                assert relerr(t1, t0) < tol, (prec, relerr(t1, t0))
        # the table now holds the widths; a new context of the box uses them
        from ifdda import capi
        try:
            assert capi.tune_get(*n, 64) == chosen[64]
            assert capi.tune_get(*n, 32) == chosen[32]
            with gpu.Solver(*n) as sv2:
                assert sv2.tile_widths() == chosen
        finally:  # later tests of this box run the default plan
            for prec in (64, 32):
                capi.tune_set(*n, prec, before[prec])


def test_recorded_widths_give_the_same_products(gpu):
    """Widths set through the table (every pass one line per tile) give the
    default plan's products bit for bit with the generic kernels."""
    from ifdda import capi
    n = (14, 12, 10)
    tab, x, alpha = _problem(n, 7)
    with gpu.Solver(*n) as sv:
        sv.use_specialized(False)
        sv.greencalculfft(D, K0)
        sv.set_dipoles(tab)
        sv.set_polarizability(alpha)
        y0 = _matvecs(sv, x)
        default = sv.tile_widths()
    capi.tune_set(*n, 64, [1, 1, 1, 1, 1])
    capi.tune_set(*n, 32, [2, 3, 1, 3, 2])
    try:
        with gpu.Solver(*n) as sv:
            assert sv.tile_widths() == {64: [1, 1, 1, 1, 1], 32: [2, 3, 1, 3, 2]}
            assert sv.tile_widths() != default
            sv.use_specialized(False)
            sv.greencalculfft(D, K0)
            sv.set_dipoles(tab)
            sv.set_polarizability(alpha)
            y1 = _matvecs(sv, x)
    finally:
        for prec in (64, 32):
            capi.tune_set(*n, prec, default[prec])
    for prec in (64, 32):
        assert np.array_equal(y0[prec], y1[prec]), prec


def test_autotune_needs_a_green_tensor(gpu):
    with gpu.Solver(8, 8, 8) as sv:
        with pytest.raises(gpu.IfddaError):
            sv.autotune()
