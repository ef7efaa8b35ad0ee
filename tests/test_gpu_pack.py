"""Independent problems solved together (ifdda_solve_batch): every result,
iteration count and matvec count equals the problem's own solve bit for bit
(same tiles, same reductions in the same order), whatever else is in the
pack; dense boxes, mapped boxes and multi-incidence problems mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _direct(gpu, p, precision, tol, method="GPBICG1"):
    """The problem's own solve through a context, with the generic z pass
    (the one packed launches use)."""
    out = []
    with gpu.Solver(p["nx"], p["ny"], p["nz"]) as sv:
        sv.use_specialized(False)
        sv.greencalculfft(p["d"], p["k0"])
        sv.set_dipoles(p.get("tabdip"))
        sv.set_polarizability(p["alpha"])
        E = np.atleast_2d(p["E_inc"])
        for r in range(E.shape[0]):
            out.append(sv.inverserig(E[r], method, tol, precision=precision, x0_zero=True))
    return out


def _problems(count, seed):
    from ifdda import scene
    probs = []
    for (s, k0, alpha, E) in scene.config4(count, seed=seed):
        probs.append(dict(nx=s.nx, ny=s.ny, nz=s.nz, d=s.d, k0=k0, alpha=alpha, E_inc=E,
                          tabdip=s.tabdip))
    return probs


def _check_equal(batch_out, direct_out, converged=True):
    for r, ref in enumerate(direct_out):
        info = batch_out["infos"][r]
        assert info["nstop"] == ref["nstop"]
        if converged:
            assert info["nstop"] == 0
        assert info["nloop"] == ref["nloop"] and info["ncompte"] == ref["ncompte"]
        # a stalled solve may end in NaN, the same NaN alone and in the pack
        assert np.array_equal(info["tol1"], ref["tol1"], equal_nan=True)
        assert np.array_equal(batch_out["FFloc"][r], ref["FFloc"], equal_nan=True)
        assert np.array_equal(batch_out["FF"][r], ref["FF"], equal_nan=True)


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_equals_own_solves(gpu, precision):
    from ifdda import scene
    tol = 1e-6 if precision == 64 else 1e-4
    probs = _problems(6, seed=21)
    # a dense box (no map) and a mapped problem with three incidences
    lam = 600e-9
    k0 = 2 * np.pi / lam
    c = scene.cube(9, 9 * 30e-9, 2.0 + 0.1j)
    probs.append(dict(nx=c.nx, ny=c.ny, nz=c.nz, d=c.d, k0=k0,
                      alpha=scene.polarizability(c.eps, c.d, k0), tabdip=None,
                      E_inc=scene.plane_wave(c.pos, k0, 20.0, 30.0)))
    s = scene.sphere(12, 150e-9, 2.25 + 0.0j)
    probs.append(dict(nx=s.nx, ny=s.ny, nz=s.nz, d=s.d, k0=k0, tabdip=s.tabdip,
                      alpha=scene.polarizability(s.eps, s.d, k0),
                      E_inc=np.stack([scene.plane_wave(s.pos, k0, t, 0.0) for t in (0.0, 30.0, 60.0)])))
    outs = gpu.solve_batch(probs, "GPBICG1", tol, precision=precision)
    for p, o in zip(probs, outs):
        print(f"pack member {p['nx']}x{p['ny']}x{p['nz']} R={np.atleast_2d(p['E_inc']).shape[0]} "
              f"it={[i['nloop'] for i in o['infos']]}", flush=True)
        _check_equal(o, _direct(gpu, p, precision, tol))


def test_result_same_alone_and_in_a_pack(gpu):
    probs = _problems(10, seed=33)
    alone = gpu.solve_batch([probs[4]], "GPBICG1", 1e-4, precision=32)[0]
    mixed = gpu.solve_batch(probs, "GPBICG1", 1e-4, precision=32)[4]
    assert np.array_equal(alone["FFloc"], mixed["FFloc"])
    assert np.array_equal(alone["FF"], mixed["FF"])
    assert [i["nloop"] for i in alone["infos"]] == [i["nloop"] for i in mixed["infos"]]


def test_batch_rejects_bad_jobs(gpu):
    p = _problems(2, seed=5)
    bad = dict(p[1])
    tab = np.array(bad["tabdip"], np.int32).copy()
    tab[tab > 0] = 1  # dipole 1 mapped many times
    bad["tabdip"] = tab
    with pytest.raises(gpu.IfddaError, match="Tabdip"):
        gpu.solve_batch([p[0], bad])


@pytest.mark.parametrize("method", ["GPBICG2", "GPBICGsafe", "GPBICGplus", "GPBICGAR", "BICGSTARPLUS",
                                    "GPBICOR", "CORS", "CG", "BICGSTAB", "TFQMR", "QMRBICGSTAB2",
                                    "QMRCLA"])
def test_pack_methods_equal_own_solves(gpu, method):
    """Every Krylov method gives each member of a pack the result of its own
    context solve, bit for bit, in both precisions (QMRCLA's transposed
    operator included: the pack's launches without alpha, then alpha^T per
    member); a
    member whose method stalls (BICGSTAB's stopping test of
    pimzbicgstab.f:168-181 can run to the iteration cap) stalls the same way
    alone, with the same counts and fields."""
    probs = _problems(4, seed=33)
    for precision, tol in ((64, 1e-6), (32, 1e-4)):
        print(f"pack {method} fp{precision}", flush=True)
        outs = gpu.solve_batch(probs, method, tol, precision=precision, maxit=2000)
        for p, o in zip(probs, outs):
            ref = []
            with gpu.Solver(p["nx"], p["ny"], p["nz"]) as sv:
                sv.use_specialized(False)
                sv.greencalculfft(p["d"], p["k0"])
                sv.set_dipoles(p.get("tabdip"))
                sv.set_polarizability(p["alpha"])
                ref.append(sv.inverserig(p["E_inc"], method, tol, precision=precision, x0_zero=True,
                                         maxit=2000))
            _check_equal(o, ref, converged=method != "BICGSTAB")
