"""GPBICG with its vector work fused into the matvec ends (ifdda_fuse_krylov):
<r0*, Ap> inside the first P5 gather, the five step-4 products inside the
second, (mask bit 2) p = r + beta (p - u) inside P1's scatter
(gpbicg.f:161-253), and (bits 3, 4) the look-ahead loop, whose second gather
also sums <r0*, t>, <r0*, y>, <r0*, At> for <r0*, r_new> and the next beta
(gpbicg.f:300-316) and whose update kernel forms the next p.  The fused and
the separate kernels form the same vectors; only the inner products are
summed in another order, so the loops take the same iterations and give the
same field to round-off, for one and for several right-hand sides, in FP64
and in mixed precision."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sph(gpu):
    from ifdda import scene
    s, k0, alpha, _ = scene.config1()
    th = [0.0, 25.0, 50.0]
    ph = [0.0, 40.0, 200.0]
    E = np.stack([scene.plane_wave(s.pos, k0, t, p, pp=pp) for t, p, pp in zip(th, ph, (1.0, 0.3, 0.0))])
    sv = gpu.Solver(s.nx, s.ny, s.nz)
    sv.greencalculfft(s.d, k0)
    sv.set_dipoles(s.tabdip)
    sv.set_polarizability(alpha)
    yield s, E, sv
    sv.fuse_krylov(10)
    sv.close()


@pytest.mark.parametrize("prec,tol,ftol", [(64, 1e-6, 1e-9), (32, 1e-6, 1e-5)])
def test_fused_equals_separate_single(sph, prec, tol, ftol):
    s, E, sv = sph
    out = {}
    for mask in (0, 3, 7, 27):
        sv.fuse_krylov(mask)
        print(f"fused krylov: mask {mask} precision {prec}", flush=True)
        out[mask] = sv.inverserig(E[1], "GPBICG1", tol, precision=prec, x0_zero=True)
    ref = sv.inverserig(E[1], "GPBICG1", 1e-10, precision=64)
    a = out[0]
    for m in (3, 7, 27):
        b = out[m]
        assert a["nstop"] == 0 and b["nstop"] == 0
        assert a["tol1"] <= tol and b["tol1"] <= tol
        if prec == 64 and m == 27:  # <r0*, r_new> as three products: rounding only
            assert abs(a["nloop"] - b["nloop"]) <= 1
        elif prec == 64:
            assert a["nloop"] == b["nloop"] and a["ncompte"] == b["ncompte"]
            assert abs(a["tol1"] / b["tol1"] - 1) < 1e-6
        else:  # FP32 inner solves: another rounding of the products may move an iteration
            assert abs(a["nloop"] - b["nloop"]) <= 2
        assert relerr(b["FFloc"], a["FFloc"]) < ftol
        assert relerr(b["FFloc"], ref["FFloc"]) < 10 * tol
    # the p update inside P1 forms the same p as its own kernel
    if prec == 64:
        assert out[7]["nloop"] == out[3]["nloop"]
        assert relerr(out[7]["FFloc"], out[3]["FFloc"]) < 1e-12


@pytest.mark.parametrize("prec,ftol", [(64, 1e-9), (32, 1e-5)])
def test_fused_equals_separate_multi(sph, prec, ftol):
    s, E, sv = sph
    res = {}
    for mask in (0, 3, 7, 27):
        sv.fuse_krylov(mask)
        print(f"fused krylov multi: mask {mask} precision {prec}", flush=True)
        res[mask] = sv.inverserig_multi(E, "GPBICG1", 1e-6, precision=prec, x0_zero=True)
    _, FLa, ia = res[0]
    for m in (3, 7, 27):
        _, FLb, ib = res[m]
        for r in range(E.shape[0]):
            assert ia[r]["nstop"] == 0 and ib[r]["nstop"] == 0
            assert ib[r]["tol1"] <= 1e-6
            if prec == 64:
                assert abs(ia[r]["nloop"] - ib[r]["nloop"]) <= (1 if m == 27 else 0)
            assert relerr(FLb[r], FLa[r]) < ftol


@pytest.mark.parametrize("prec", [64, 32])
def test_lookahead_scalars(sph, prec):
    """The look-ahead loop's alpha, beta, eta, zeta, <r0*, r> and residual per
    iteration follow the separate kernels' (gpbicg.f:161-316): FP64 to 1e-9 of
    each column's scale, FP32 inner solves to 1e-3 (their vectors are rounded
    to FP32 and the products summed in another order)."""
    s, E, sv = sph
    tr = {}
    sv.trace_enable(64)
    try:
        for mask in (0, 27):
            sv.fuse_krylov(mask)
            print(f"look-ahead trace: mask {mask} precision {prec}", flush=True)
            # mixed precision: tol 1e-4 is one refinement sweep, whose inner
            # solve the trace records
            sv.inverserig(E[0], "GPBICG1", 1e-6 if prec == 64 else 1e-4, precision=prec,
                          x0_zero=True)
            tr[mask] = sv.trace()
    finally:
        sv.trace_enable(0)
    a, b = tr[0], tr[27]
    k = min(len(a), len(b), 12)
    assert k >= 3
    tol = 1e-9 if prec == 64 else 1e-3
    scale = np.abs(a[:k]).max(axis=0).clip(1e-300)
    err = (np.abs(a[:k] - b[:k]) / scale).max(axis=0)
    assert (err < tol).all(), err
