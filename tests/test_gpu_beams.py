"""Incident plane waves and polarisabilities formed on the device
(ifdda_beams.hip) against the compiled reference (ondeplane.f, ondecirce.f,
ondeplanekxky.f, polarisabilite.f), and jobs that carry wave parameters and
permittivities instead of arrays: their solves agree with those of the same
jobs given explicit arrays to the solver tolerance, and with each other
(batch, runner) bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAM = 632.8e-9
K0 = 2 * np.pi / LAM


def _sphere():
    from ifdda import scene
    return scene.sphere(10, 120e-9, 2.25 + 0.2j)


@pytest.mark.parametrize("wave", ["plane", "circular", "kxky"])
def test_incident_waves_match_reference(gpu, ref, wave):
    s = _sphere()
    E0 = 0.8 + 0.3j
    if wave == "kxky":
        params = np.array([[0.3 * K0, -0.2 * K0, 0.4], [0.0, 0.0, 1.0], [-0.6 * K0, 0.5 * K0, 0.0]])
    else:
        params = np.array([[30.0, 10.0, 0.5 if wave == "plane" else 1.0],
                           [-20.0, 200.0, 0.0 if wave == "plane" else -1.0],
                           [0.0, 45.0, 1.0]])
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.set_dipoles(s.tabdip)
        out = sv.incident_device(params, K0, s.d, s.origin, wave=wave, E0=E0)
        got = out.cpu().numpy().view(np.complex128).reshape(len(params), -1)
    for r, (a, b, c) in enumerate(params):
        if wave == "plane":
            exp = ref.ondeplane(s.pos, K0, E0, c, a, b)
        elif wave == "circular":
            exp = ref.ondecirce(s.pos, K0, E0, c, a, b)
        else:
            exp = ref.ondeplanekxky(s.pos, K0, E0, c, a, b)
        err = np.abs(got[r] - exp).max() / np.abs(exp).max()
        print(f"{wave} wave {r}: {err:.2e}", flush=True)
        assert err < 1e-13


@pytest.mark.parametrize("method", ["CM", "RR", "GB", "LA", "LR"])
def test_polarizability_device_matches_reference(gpu, ref, method):
    """The device polarisability enters the matvec as the reference's
    poladiff values do: y = x - T alpha x agrees to rounding."""
    import torch
    s = _sphere()
    rng = np.random.default_rng(7)
    eps = 2.0 + rng.random(s.nd) + 1j * 0.3 * rng.random(s.nd)
    al_ref = np.array([ref.poladiff(s.d, e, K0, method) for e in eps])
    x = rng.standard_normal(3 * s.nd) + 1j * rng.standard_normal(3 * s.nd)
    ys = []
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.greencalculfft(s.d, K0)
        sv.set_dipoles(s.tabdip)
        for dev in (False, True):
            if dev:
                sv.polarizability_device(eps, s.d, K0, method)
            else:
                sv.set_polarizability(al_ref)
            xd = torch.from_numpy(x.view(np.float64).copy()).cuda()
            yd = torch.zeros_like(xd)
            sv.matvec_device(xd.data_ptr(), xd.data_ptr(), yd.data_ptr(), prec=64)
            torch.cuda.synchronize()
            ys.append(yd.cpu().numpy().view(np.complex128) - x)
    err = np.abs(ys[1] - ys[0]).max() / np.abs(ys[0]).max()
    print(f"poladiff {method}: {err:.2e}", flush=True)
    assert err < 1e-13


def test_jobs_with_descriptors_equal_explicit_arrays(gpu):
    """Batch and runner jobs given (waves, eps) solve to the results of the
    same jobs given (E_inc, alpha) formed on the host: same iteration counts,
    fields to rounding of the inputs."""
    from ifdda import scene
    probs = scene.config4(6, seed=13)
    explicit, described = [], []
    rng = np.random.default_rng(2)
    for (s, k0, alpha, _) in probs:
        th, ph, pp = rng.uniform(0, 80, 2).tolist() + [float(rng.random())]
        E = scene.plane_wave(s.pos, k0, th, ph, pp=pp)[None]
        explicit.append(dict(nx=s.nx, ny=s.ny, nz=s.nz, d=s.d, k0=k0, tabdip=s.tabdip,
                             alpha=scene.polarizability(s.eps, s.d, k0, "RR"), E_inc=E))
        described.append(dict(nx=s.nx, ny=s.ny, nz=s.nz, d=s.d, k0=k0, tabdip=s.tabdip, eps=s.eps,
                              pola="RR", waves=[[th, ph, pp]], wave="plane", origin=s.origin))
    a = gpu.solve_batch(explicit, "GPBICG1", 1e-6, precision=64)
    b = gpu.solve_batch(described, "GPBICG1", 1e-6, precision=64)
    with gpu.Runner(nworkers=2, mem_budget=64e9) as run:
        tk = [run.submit(p["nx"], p["ny"], p["nz"], p["d"], p["k0"], tabdip=p["tabdip"], eps=p["eps"],
                         pola="RR", waves=p["waves"], wave="plane", origin=p["origin"], tol=1e-6,
                         precision=64) for p in described]
        c = [run.wait(t) for t in tk]
    for x, y, z in zip(a, b, c):
        # inputs equal to rounding (device sincos vs numpy exp, C++ vs numpy
        # poladiff): both converge to the tolerance, to the same fields within
        # it; a long ill-conditioned solve (70-200+ GPBiCG iterations) may end
        # a few iterations apart on rounding-level input differences (70 vs 73
        # seen when the build's multiply-add contraction changed the device
        # waves' last bits)
        assert x["infos"][0]["tol1"] <= 1e-6 and y["infos"][0]["tol1"] <= 1e-6
        nx_, ny_ = x["infos"][0]["nloop"], y["infos"][0]["nloop"]
        assert abs(nx_ - ny_) <= max(4, 0.05 * nx_), (nx_, ny_)
        ref = np.abs(x["FFloc"]).max()
        assert np.abs(x["FFloc"] - y["FFloc"]).max() < 1e-4 * ref
        # the same descriptors through the runner: bit for bit
        assert y["infos"][0]["nloop"] == z["infos"][0]["nloop"]
        assert np.array_equal(y["FFloc"], z["FFloc"])
