"""The runner (many independent solves at once, grouped into packs): every job's result equals the
same job solved alone through the handle API, bit for bit (all kernels are
deterministic), whatever runs beside it; a budget too small for two jobs
runs them one at a time."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _direct(gpu, s, k0, alpha, E, precision, tol):
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.use_specialized(False)  # the generic z pass, as the runner's packed launches
        sv.greencalculfft(s.d, k0)
        sv.set_dipoles(s.tabdip)
        sv.set_polarizability(alpha)
        return sv.inverserig(E, "GPBICG1", tol, precision=precision, x0_zero=True)


@pytest.mark.parametrize("precision", [64, 32])
def test_runner_jobs_equal_direct_solves(gpu, precision):
    from ifdda import scene
    probs = scene.config4(10, seed=7)
    tol = 1e-6 if precision == 64 else 1e-4
    with gpu.Runner(nworkers=4, mem_budget=64e9) as run:
        tickets = [run.submit(s.nx, s.ny, s.nz, s.d, k0, alpha, E, tabdip=s.tabdip, tol=tol,
                              precision=precision) for (s, k0, alpha, E) in probs]
        outs = [run.wait(t) for t in tickets]
    for (s, k0, alpha, E), o in zip(probs, outs):
        print(f"job {s.nx}x{s.ny}x{s.nz} nd={s.nd} it={o['infos'][0]['nloop']}", flush=True)
        ref = _direct(gpu, s, k0, alpha, E, precision, tol)
        assert o["infos"][0]["nstop"] == 0 and o["infos"][0]["tol1"] <= tol
        assert o["infos"][0]["nloop"] == ref["nloop"]
        assert np.array_equal(o["FFloc"][0], ref["FFloc"])
        assert np.array_equal(o["FF"][0], ref["FF"])


def test_same_job_alone_and_in_a_mix(gpu):
    from ifdda import scene
    probs = scene.config4(12, seed=9)
    s, k0, alpha, E = probs[0]
    E2 = np.stack([E, np.conj(E)])  # two incidences in one job
    with gpu.Runner(nworkers=1, mem_budget=64e9) as run:
        alone = run.wait(run.submit(s.nx, s.ny, s.nz, s.d, k0, alpha, E2, tabdip=s.tabdip))
    with gpu.Runner(nworkers=6, mem_budget=64e9) as run:
        tk = [run.submit(p.nx, p.ny, p.nz, p.d, kk, al, ee, tabdip=p.tabdip) for (p, kk, al, ee) in probs[1:6]]
        mine = run.submit(s.nx, s.ny, s.nz, s.d, k0, alpha, E2, tabdip=s.tabdip)
        tk += [run.submit(p.nx, p.ny, p.nz, p.d, kk, al, ee, tabdip=p.tabdip) for (p, kk, al, ee) in probs[6:]]
        mixed = run.wait(mine)
        for t in tk:
            assert run.wait(t)["infos"][0]["nstop"] == 0
    assert np.array_equal(alone["FFloc"], mixed["FFloc"])
    assert [i["nloop"] for i in alone["infos"]] == [i["nloop"] for i in mixed["infos"]]


def test_budget_serialises_jobs(gpu):
    from ifdda import scene
    probs = scene.config4(4, seed=11)
    with gpu.Runner(nworkers=4, mem_budget=1.0) as run:
        tk = [run.submit(p.nx, p.ny, p.nz, p.d, kk, al, ee, tabdip=p.tabdip) for (p, kk, al, ee) in probs]
        outs = [run.wait(t) for t in tk]
    # with room for one job at a time the runs do not overlap: the last job
    # waited for all the others
    assert all(o["infos"][0]["nstop"] == 0 for o in outs)
    assert max(o["latency_s"] for o in outs) >= 0.95 * sum(o["run_s"] for o in outs)


def test_runner_rejects_bad_jobs(gpu):
    with gpu.Runner(nworkers=1, mem_budget=1e9) as run:
        with pytest.raises(gpu.IfddaError, match="bad job"):
            run.submit(4, 4, 4, -1.0, 1e7, np.ones(8), np.ones(24))
