"""Brightfield and darkfield microscopes (microsbf.f, microsdf.f,
microssansinc.f; quick and direct-sum lens) on the GPU against the compiled
reference's own whole microscopes (golden/microscopes.npz) and against an
independent CPU pipeline: the oracle's inverserig per incidence, the numpy
restatements of diffractefft2dlens / ondeplanekxky / the image assembly
(tests/micro_pipeline.py, itself checked against the same fixture in
test_microscopy_oracle.py).  Images agree to the solver tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAM = 632.8e-9
K0 = 2 * np.pi / LAM


@pytest.fixture(scope="module")
def small(gpu):
    from ifdda import scene
    s = scene.sphere(6, 60e-9, 2.25 + 0.1j)
    alpha = scene.polarizability(s.eps, s.d, K0, "RR")
    sv = gpu.Solver(s.nx, s.ny, s.nz)
    sv.greencalculfft(s.d, K0)
    sv.set_dipoles(s.tabdip)
    sv.set_polarizability(alpha)
    m = sv.microscope_opts(128, 0.8, 0.5, K0, s.d, s.origin, gross=100.0, zlens=50e-9, nside=1,
                           npolainc=0, E0=1.0 + 0.5j)
    yield s, alpha, sv, m
    sv.close()


def test_incidence_lattice(small):
    import microscopy as mi
    s, alpha, sv, m = small
    kxy, dk = sv.microscope_incidences(m)
    inc, dkinc, _ = mi.incidences(K0, 0.5, s.d, 128, s.pos[:, 0], s.pos[:, 1])
    assert abs(dk / dkinc - 1) < 1e-15
    ref = [(kx, ky, pp) for pp in (0.0, 1.0) for (_, _, kx, ky) in inc]
    assert kxy.shape == (len(ref), 3)
    assert np.allclose(kxy, np.array(ref), rtol=0, atol=1e-9 * K0)


def _cpu_images(s, alpha, oracle, m, tol, dark=False):
    from micro_pipeline import cpu_images
    return cpu_images(s, alpha, oracle, m, tol, dark=dark)


def test_brightfield_images_match_cpu_pipeline(small, oracle):
    s, alpha, sv, m = small
    print(f"brightfield {s.nd} dipoles nfft2d {m.nfft2d}", flush=True)
    img, infos = sv.microsbf(m, "GPBICG1", 1e-10, precision=64)
    assert all(i["nstop"] == 0 and i["tol1"] <= 1e-10 for i in infos)
    ref = _cpu_images(s, alpha, oracle, m, 1e-10)
    for c in range(6):
        scale = np.abs(ref[c]).max()
        assert scale > 0
        assert np.abs(img[c] - ref[c]).max() < 1e-7 * scale, (c, np.abs(img[c] - ref[c]).max() / scale)
    # the incident light dominates the total image; the scattered image is weaker
    assert np.abs(ref[3:]).max() > np.abs(ref[:3]).max()


def test_brightfield_batches_and_precision(small):
    """Incidences solved in batches of 5 give the all-at-once images; the mixed
    precision images agree to the solve tolerance."""
    s, alpha, sv, m = small
    full, _ = sv.microsbf(m, "GPBICG1", 1e-8, precision=64)
    m5 = sv.microscope_opts(128, 0.8, 0.5, K0, s.d, s.origin, gross=100.0, zlens=50e-9, nside=1,
                            npolainc=0, E0=1.0 + 0.5j, max_batch=5)
    part, infos = sv.microsbf(m5, "GPBICG1", 1e-8, precision=64)
    assert np.abs(part - full).max() < 1e-12 * np.abs(full).max()
    mixed, infos = sv.microsbf(m, "GPBICG1", 1e-6, precision=32, x0_zero=True)
    assert all(i["tol1"] <= 1e-6 for i in infos)
    assert np.abs(mixed - full).max() < 1e-4 * np.abs(full).max()


def test_brightfield_argument_errors(small, gpu):
    s, alpha, sv, m = small
    bad = sv.microscope_opts(128, 0.8, 1.0, K0, s.d, s.origin)
    with pytest.raises(gpu.IfddaError, match="NA inc strictly between 0 and 1"):
        sv.microsbf(bad)
    tiny = sv.microscope_opts(16, 0.8, 0.5, K0, s.d, s.origin)
    with pytest.raises(gpu.IfddaError, match="nfft2d too small"):
        sv.microsbf(tiny)


@pytest.mark.parametrize("na,nainc", [(0.5, 0.8), (0.8, 0.5)])
def test_darkfield_images_match_cpu_pipeline(small, oracle, na, nainc):
    """Condenser NA above the objective's (dark field: the direct wave misses
    the pupil, so total = scattered image) and below it (the direct wave
    enters, phase-microscope case)."""
    import microscopy as mi
    s, alpha, sv, _ = small
    m = sv.microscope_opts(128, na, nainc, K0, s.d, s.origin, gross=100.0, zlens=50e-9, nside=1,
                           npolainc=0, E0=1.0 + 0.5j)
    kxy, dk = sv.microscope_incidences(m, dark=True)
    ring, dkinc, _ = mi.incidences_df(K0, nainc, s.d, 128, s.pos[:, 0], s.pos[:, 1])
    assert abs(dk / dkinc - 1) < 1e-15
    assert np.allclose(kxy, np.array([(kx, ky, pp) for (_, pp, kx, ky) in ring]), rtol=0,
                       atol=1e-9 * K0)
    print(f"darkfield NA {na} NA_inc {nainc}: {len(ring)} incidences", flush=True)
    img, infos = sv.microsdf(m, method="GPBICG1", tol=1e-10, precision=64)
    assert len(infos) == len(ring) and all(i["nstop"] == 0 and i["tol1"] <= 1e-10 for i in infos)
    ref = _cpu_images(s, alpha, oracle, m, 1e-10, dark=True)
    for c in range(6):
        scale = np.abs(ref[c]).max()
        assert scale > 0
        assert np.abs(img[c] - ref[c]).max() < 1e-7 * scale, (c, np.abs(img[c] - ref[c]).max() / scale)
    if nainc > na:
        assert np.array_equal(img[:3], img[3:])
    else:
        assert np.abs(img[3:] - img[:3]).max() > 0.1 * np.abs(img[:3]).max()


@pytest.mark.parametrize("dark", [False, True])
def test_reflection_microscopes_match_cpu_pipeline(small, oracle, dark):
    """nside = -1: microsbfneg / microsdfneg (microssansinc.f), the objective
    on the z < 0 side: the lens on that side, u_z and v_z of the lens
    rotation negative, no incident wave (total image = scattered image)."""
    s, alpha, sv, _ = small
    m = sv.microscope_opts(128, 0.8, 0.5, K0, s.d, s.origin, gross=100.0, zlens=-50e-9, nside=-1,
                           npolainc=0, E0=1.0 + 0.5j)
    print(f"reflection microscope dark={dark}", flush=True)
    img, infos = (sv.microsdf(m, method="GPBICG1", tol=1e-10, precision=64) if dark else
                  sv.microsbf(m, "GPBICG1", 1e-10, precision=64))
    assert all(i["nstop"] == 0 for i in infos)
    ref = _cpu_images(s, alpha, oracle, m, 1e-10, dark=dark)
    for c in range(3):
        scale = np.abs(ref[c]).max()
        assert scale > 0
        assert np.abs(img[c] - ref[c]).max() < 1e-7 * scale, (c, np.abs(img[c] - ref[c]).max() / scale)
    assert np.array_equal(img[:3], img[3:])
    # the light scattered back differs from the light scattered forward
    fwd = _cpu_images(s, alpha, oracle, sv.microscope_opts(128, 0.8, 0.5, K0, s.d, s.origin, gross=100.0,
                                                           zlens=50e-9, nside=1, npolainc=0,
                                                           E0=1.0 + 0.5j), 1e-10, dark=dark)
    assert np.abs(fwd[0] - ref[0]).max() > 0.01 * np.abs(fwd[0]).max()


@pytest.mark.parametrize("dark,nside", [(False, 1), (True, 1), (False, -1)])
def test_direct_lens_matches_cpu_pipeline(small, oracle, dark, nside):
    """nquicklens = 0 (direct_lens): the pupil and condenser steps of
    microsbf.f:139-157 and the pupil field as the direct sum over dipoles
    (microsbf.f:629-668; its restatement equals the reference's FFT lens at
    common bins, test_microscopy_oracle.py), then the same pupil and image
    steps as the quick lens."""
    import microscopy as mi
    s, alpha, sv, _ = small
    m = sv.microscope_opts(128, 0.8, 0.5, K0, s.d, s.origin, gross=100.0, zlens=50e-9, nside=nside,
                           npolainc=0, E0=1.0 + 0.5j, direct_lens=True)
    print(f"direct lens dark={dark} nside={nside}", flush=True)
    kxy, dk = sv.microscope_incidences(m, dark=dark)
    if dark:
        ring, dkinc, _ = mi.incidences_df(K0, 0.5, s.d, 128, s.pos[:, 0], s.pos[:, 1], direct=True)
        ref_inc = [(kx, ky, pp) for (_, pp, kx, ky) in ring]
    else:
        inc, dkinc, _ = mi.incidences(K0, 0.5, s.d, 128, s.pos[:, 0], s.pos[:, 1], direct=True)
        ref_inc = [(kx, ky, pp) for pp in (0.0, 1.0) for (_, _, kx, ky) in inc]
    assert abs(dk / dkinc - 1) < 1e-15
    assert kxy.shape == (len(ref_inc), 3)
    assert np.allclose(kxy, np.array(ref_inc), rtol=0, atol=1e-9 * K0)
    img, infos = sv.microsbf(m, "GPBICG1", 1e-10, precision=64, dark=dark)
    assert all(i["nstop"] == 0 and i["tol1"] <= 1e-10 for i in infos)
    ref = _cpu_images(s, alpha, oracle, m, 1e-10, dark=dark)
    for c in range(6):
        scale = np.abs(ref[c]).max()
        assert scale > 0
        assert np.abs(img[c] - ref[c]).max() < 1e-7 * scale, (c, np.abs(img[c] - ref[c]).max() / scale)


def test_given_incidences_equal_the_condenser_lattice(small):
    """The condenser's own incidences passed as an explicit list give the
    same scattered images (same solves, pupil fields and sums); only the
    direct wave's pupil bin follows the given-incidence rule (nearest pupil
    bin), so the total images are compared only where no incident wave
    enters: an off-lattice list with the direct waves outside the objective
    NA equals the scattered image."""
    s, alpha, sv, m = small
    kxy, _ = sv.microscope_incidences(m)
    img, infos = sv.microsbf(m, tol=1e-8)
    mu = sv.microscope_opts(128, 0.8, 0.5, K0, s.d, s.origin, gross=100.0, zlens=50e-9, nside=1,
                            npolainc=0, E0=1.0 + 0.5j, incidences=kxy)
    imu, infu = sv.microsbf(mu, tol=1e-8)
    assert len(infu) == len(infos) == kxy.shape[0]
    assert np.array_equal(img[:3], imu[:3])
    # oblique waves (sin theta = 0.87) far outside an NA 0.3 objective: even
    # rounded to the pupil grid (step 0.25 k0 here) no direct wave enters
    th = np.radians(60.0)
    far = np.array([[K0 * np.sin(th) * np.cos(p), K0 * np.sin(th) * np.sin(p), pp]
                    for p in (0.3, 2.0) for pp in (0.0, 1.0)])
    mf = sv.microscope_opts(128, 0.3, 0.5, K0, s.d, s.origin, gross=100.0, zlens=50e-9, nside=1,
                            npolainc=0, E0=1.0, incidences=far)
    imf, inff = sv.microsbf(mf, tol=1e-8)
    assert len(inff) == 4 and all(i["nstop"] == 0 for i in inff)
    assert np.allclose(imf[3:], imf[:3], rtol=1e-12, atol=0)


@pytest.fixture(scope="module")
def golden_sphere(gpu):
    from conftest import golden
    from micro_pipeline import micro_scene
    fx = golden("microscopes.npz")
    s, _ = micro_scene()
    sv = gpu.Solver(s.nx, s.ny, s.nz)
    sv.greencalculfft(s.d, K0)
    sv.set_dipoles(s.tabdip)
    sv.set_polarizability(np.full(s.nd, complex(fx["alpha"])))
    yield s, sv, fx
    sv.close()


def _golden_cases():
    from micro_pipeline import MICRO_CASES, micro_key
    return [pytest.param(c, id=micro_key(*c)) for c in MICRO_CASES]


@pytest.mark.parametrize("case", _golden_cases())
def test_microscopes_match_reference_fixture(golden_sphere, case):
    """The device microscope against the compiled reference's microsbf_,
    microsdf_, microsbfneg_, microsdfneg_ (quick lens nquicklens = 1 and the
    direct sum nquicklens = 0) on the same sphere and incident power: the
    number of illuminations, and each image plane at 64 seeded pixels, its
    projections, max and sum to 1e-8 of the plane's max (both solves at tol
    1e-10)."""
    from micro_pipeline import MICRO_GROSS, MICRO_NFFT, MICRO_TOL, case_geometry, check_images, micro_key
    s, sv, fx = golden_sphere
    kind, na, nainc, nq = case
    key = micro_key(*case)
    dark, nside, zlens, direct = case_geometry(*case)
    E0 = complex(fx[f"{key}_E0"])
    m = sv.microscope_opts(MICRO_NFFT, na, nainc, K0, s.d, s.origin, gross=MICRO_GROSS, zlens=zlens,
                           nside=nside, npolainc=0, E0=E0, direct_lens=direct)
    print(f"reference microscope {key}", flush=True)
    kxy, _ = sv.microscope_incidences(m, dark=dark)
    assert kxy.shape[0] == int(fx[f"{key}_nillum"])  # both polarisations of every incidence
    img, infos = sv.microsbf(m, "GPBICG1", MICRO_TOL, precision=64, dark=dark)
    assert all(i["nstop"] == 0 and i["tol1"] <= MICRO_TOL for i in infos)
    check_images(img, fx, key, 1e-8)
