"""Multi-RHS solves: each right-hand side of a batch reproduces its own
single solve (same recurrence, iteration count and field to round-off); a
zero RHS breaks down alone; mixed precision meets the tolerance per RHS."""
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
    E = np.concatenate([E, np.zeros((1, E.shape[1]), complex)])  # a zero RHS last
    sv = gpu.Solver(s.nx, s.ny, s.nz)
    sv.greencalculfft(s.d, k0)
    sv.set_dipoles(s.tabdip)
    sv.set_polarizability(alpha)
    yield s, E, sv
    sv.close()


@pytest.mark.parametrize("method", ["GPBICG1", "QMRBICGSTAB1", "TFQMR"])
def test_batch_equals_single(sph, method):
    s, E, sv = sph
    FF, FL, infos = sv.inverserig_multi(E, method, 1e-6)
    for r in range(3):
        one = sv.inverserig(E[r], method, 1e-6)
        assert infos[r]["nstop"] == 0 and one["nstop"] == 0
        assert infos[r]["nloop"] == one["nloop"]
        assert infos[r]["ncompte"] == one["ncompte"]
        assert relerr(FL[r], one["FFloc"]) < 1e-13
        assert relerr(FF[r], one["FF"]) < 1e-13
        assert abs(infos[r]["tol1"] - one["tol1"]) < 1e-12
    assert infos[3]["nstop"] == 1  # the zero RHS: breakdown, the others unaffected


def test_batch_mixed_precision(sph):
    s, E, sv = sph
    FF, FL, infos = sv.inverserig_multi(E[:3], "GPBICG1", 1e-6, precision=32, x0_zero=True)
    for r in range(3):
        one = sv.inverserig(E[r], "GPBICG1", 1e-10)
        assert infos[r]["nstop"] == 0 and infos[r]["tol1"] <= 1e-6
        assert relerr(FL[r], one["FFloc"]) < 1e-5

