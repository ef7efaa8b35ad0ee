"""Pins (CPU) of the numpy restatement of the brightfield microscope pieces
(oracle/microscopy.py) against the reference's fixtures and, where it is
built here, the compiled reference: ondeplanekxky, deltakroutine,
irradiance, diffractefft2dlens, fouriertoimage (literal quadrant swap and
its even-N form)."""
import numpy as np
import pytest

from conftest import golden, relerr

LAM = 632.8e-9
K0 = 2 * np.pi / LAM
D = 20e-9


@pytest.mark.parametrize("tag,N,na,side", [("a", 64, 0.9, 1), ("b", 96, 0.5, -1)])
def test_lens_matches_reference_fixture(tag, N, na, side):
    import microscopy as mi
    f = golden("microscope.npz")
    E, imax, dk = mi.diffractefft2dlens(f["box_FF"], 6, 5, 4, N, K0, -50e-9, -40e-9, f["box_zs"], D,
                                        na * K0, side)
    assert imax == int(f[f"lens_{tag}_imax"]) and abs(dk / float(f[f"lens_{tag}_dk"]) - 1) < 1e-15
    assert relerr(E, f[f"lens_{tag}"]) < 1e-12


@pytest.mark.parametrize("tag,N,na,side", [("a", 64, 0.9, 1), ("b", 96, 0.5, -1)])
def test_direct_lens_sum_matches_reference_lens_fixture(tag, N, na, side):
    """The nquicklens = 0 pupil field (microsbf.f:629-668, a direct sum over
    the dipoles; microssansinc.f:600-640 for nside = -1) samples the sum that
    the reference's FFT lens (diffractefft2dlens.f) evaluates: at the same
    step and size it equals the reference's own output at every bin inside
    the aperture."""
    import microscopy as mi
    f = golden("microscope.npz")
    imax, dk = int(f[f"lens_{tag}_imax"]), float(f[f"lens_{tag}_dk"])
    nx, ny, nz = 6, 5, 4
    F = np.asarray(f["box_FF"]).reshape(nz, ny, nx, 3)
    c, b, a = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([-50e-9 + a.ravel() * D, -40e-9 + b.ravel() * D,
                    np.asarray(f["box_zs"])[c.ravel()]], 1)
    E = mi.lens_field_direct(pos, F.reshape(-1, 3), K0, dk, imax, na * K0, side)
    ref = f[f"lens_{tag}"]
    k = np.arange(-imax, imax + 1) * dk
    inside = k[:, None] ** 2 + k[None, :] ** 2 < (na * K0) ** 2
    assert inside.sum() >= 9
    assert relerr(E[inside], ref[inside]) < 1e-12, relerr(E[inside], ref[inside])


def test_fouriertoimage_matches_reference_fixture():
    import microscopy as mi
    f = golden("microscope.npz")
    planes = list(f["f2i_in"])
    for out in (mi.fouriertoimage(3e5, 100.0, planes), mi.fouriertoimage_even(3e5, 100.0, planes)):
        assert relerr(np.stack(out), f["f2i_out"]) < 1e-12


def test_pieces_match_compiled_reference(ref):
    import microscopy as mi
    rng = np.random.default_rng(4)
    pos = rng.standard_normal((7, 3)) * 1e-7
    for pp, kx, ky in ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.4, 3e6, -2e6), (1.0, -6e6, 2e6)):
        assert relerr(mi.ondeplanekxky(pos, K0, 0.7 - 0.3j, pp, kx, ky),
                      ref.ondeplanekxky(pos, K0, 0.7 - 0.3j, pp, kx, ky)) < 1e-14
    for kx, ky, dk in ((3.1e6, -2.2e6, 1e6), (9.8e6, 0.5e6, 1e6), (-4.5e6, 4.5e6, 1e6)):
        assert mi.deltakroutine(kx, ky, dk, dk, K0) == ref.deltakroutine(kx, ky, dk, dk, K0)
    a, b = mi.irradiance(2e-3, 5e-7), ref.irradiance(2e-3, 5e-7)
    assert abs(a[0] / b[0] - 1) < 1e-14 and abs(a[1] / b[1] - 1) < 1e-14
    F = rng.standard_normal(3 * 60) + 1j * rng.standard_normal(3 * 60)
    zs = 5e-9 + D * np.arange(3)
    E1, i1, _ = mi.diffractefft2dlens(F, 5, 4, 3, 72, K0, -30e-9, 10e-9, zs, D, 0.7 * K0, -1)
    E2, i2, _, nstop, _ = ref.diffractefft2dlens(F, 5, 4, 3, 72, K0, -30e-9, 10e-9, zs, D, 0.7 * K0, -1)
    assert nstop == 0 and i1 == i2 and relerr(E1, E2) < 1e-12
    planes = [rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16)) for _ in range(6)]
    got, want = mi.fouriertoimage(2e5, 40.0, planes), ref.fouriertoimage(2e5, 40.0, planes)
    assert relerr(np.stack(got), np.stack(want)) < 1e-12


@pytest.mark.parametrize("nainc,N", [(0.8, 128), (0.3, 256), (0.95, 256)])
def test_darkfield_ring(nainc, N):
    """microsdf.f:159-221: the darkfield incidences lie on the cone of the
    condenser NA, equally spaced in azimuth from phi = 0, ideltam =
    max(int(2 pi / (deltak NA_inc / k0)) + 1, 8) of them per polarisation
    (pp = 0 first), deltak the brightfield condenser step."""
    import microscopy as mi
    k0 = 2 * np.pi / 632.8e-9
    d = 20e-9
    xs = np.arange(6) * d
    ring, dkinc, dkx = mi.incidences_df(k0, nainc, d, N, xs, xs)
    _, dk_bf, _ = mi.incidences(k0, nainc, d, N, xs, xs)
    assert dkinc == dk_bf and abs(dkinc / dkx - round(dkinc / dkx)) < 1e-12
    m = len(ring) // 2
    assert m == max(int(2 * np.pi / (dkinc / k0 * nainc)) + 1, 8) and len(ring) == 2 * m
    kx = np.array([r[2] for r in ring])
    ky = np.array([r[3] for r in ring])
    assert np.allclose(np.hypot(kx, ky), k0 * nainc, rtol=1e-14)
    phi = np.mod(np.arctan2(ky[:m], kx[:m]), 2 * np.pi)
    assert np.allclose(phi, 2 * np.pi * np.arange(m) / m, atol=1e-12)
    assert [r[1] for r in ring] == [0.0] * m + [1.0] * m
    assert np.array_equal(kx[:m], kx[m:]) and np.array_equal(ky[:m], ky[m:])


def test_reflection_image_mirror_symmetry():
    """A dipole in the z = 0 plane radiates mirror images into z > 0 and
    z < 0 (E_x, E_y equal, E_z opposite).  The reflection microscope
    (nside = -1, lens rotation with u_z, v_z < 0) must then give the
    transmission microscope's scattered images exactly, component by
    component, with the focal plane on the dipole (zlens = 0: the focus
    phase exp(i kz zlens) is not mirrored, microssansinc.f:487); a lens
    rotation that ignored the side would not."""
    import microscopy as mi
    N, d = 64, 20e-9
    p = np.array([1.0 + 0.3j, -0.4 + 0.2j, 0.0])
    FFbox = p.copy()  # one cell at the origin
    imgs = {}
    for side in (1, -1):
        El, imax, dk = mi.diffractefft2dlens(FFbox, 1, 1, 1, N, K0, 0.0, 0.0, np.array([0.0]), d,
                                             0.9 * K0, side)
        inc = [(0, 0, 0.0, 0.0, 0.0)]
        imgs[side] = mi.image_assembly([El], inc, K0, 0.0, 0.9 * K0, 0.0, 50.0, dk, imax, dk, N,
                                       nside=side)
    for c in range(3):
        scale = np.abs(imgs[1][c]).max()
        assert scale > 0
        assert np.abs(imgs[1][c] - imgs[-1][c]).max() < 1e-12 * scale
    # without the incident wave (E0 = 0) both keep total = scattered
    assert np.array_equal(imgs[-1][0], imgs[-1][3])


@pytest.mark.parametrize("dark", [False, True])
def test_image_assembly_of_the_bare_incident_wave(dark):
    """No scattered light, one normal incidence: the pupil holds the incident
    wave alone at bin (0, 0) and the image is flat.  Brightfield adds E/dk^2
    (microsbf.f:736-746), darkfield i E/dk^2 (microsdf.f:712-717); the image
    plane scales by dk^2/gross, so every pixel reads |E_c|/gross."""
    import microscopy as mi
    N, imax, dk, gross = 64, 5, 3e5, 40.0
    E0 = 0.7 - 0.2j
    El = np.zeros((2 * imax + 1, 2 * imax + 1, 3), np.complex128)
    for pp in (0.0, 1.0):
        img = mi.image_assembly([El], [(0, 0, 0.0, 0.0, pp)], K0, E0, 0.5 * K0, 0.0, gross, dk, imax,
                                dk, N, dark=dark)
        e = mi.ondeplanekxky(np.zeros((1, 3)), K0, E0, pp, 0.0, 0.0)
        for c in range(3):
            assert np.abs(img[c]).max() == 0.0
            assert np.allclose(img[3 + c], abs(e[c]) / gross, rtol=1e-12, atol=1e-300)


def _fixture_case(case, oracle, tol=1e-10):
    from types import SimpleNamespace

    import micro_pipeline as mp
    fx = golden("microscopes.npz")
    s, _ = mp.micro_scene()
    key = mp.micro_key(*case)
    dark, nside, zlens, direct = mp.case_geometry(*case)
    E0 = complex(fx[f"{key}_E0"])
    m = SimpleNamespace(direct_lens=direct, numaperinc=case[2], numaper=case[1], nfft2d=mp.MICRO_NFFT,
                        E0=(E0.real, E0.imag), zlens=zlens, gross=mp.MICRO_GROSS, nside=nside)
    img = mp.cpu_images(s, np.full(s.nd, complex(fx["alpha"])), oracle, m, tol, dark=dark)
    return img, fx, key


def _cpu_fixture_cases():
    from micro_pipeline import MICRO_CASES, micro_key
    # the direct-sum lens (nquicklens = 0) in every microscope, and the quick
    # lens in the two brightfield ones (the darkfield quick-lens cases run on
    # the GPU side only, to keep this suite short)
    keep = [c for c in MICRO_CASES if c[3] == 0 or c[0].startswith("microsbf")]
    return [pytest.param(c, id=micro_key(*c)) for c in keep]


@pytest.mark.parametrize("case", _cpu_fixture_cases())
def test_pipeline_matches_reference_microscopes(oracle, case):
    """The whole CPU pipeline (oracle inverserig, incidences / incidences_df,
    direct_geometry, lens_field_direct or diffractefft2dlens, image_assembly)
    against the compiled reference's microsbf_ / microsdf_ / microsbfneg_ /
    microsdfneg_ (golden/microscopes.npz): images to 1e-8 of each plane's
    max, both solved to 1e-10."""
    from micro_pipeline import check_images
    img, fx, key = _fixture_case(case, oracle)
    check_images(img, fx, key, 1e-8)


def test_fixture_catches_planted_mistakes(oracle, monkeypatch):
    """Mistakes planted in the oracle's direct-lens steps (CPU only) make the
    direct-lens brightfield case fail the fixture: the pupil step's
    threshold (imaxk0 > 4 instead of > 5), the phase sign of the dipole sum,
    and the nside = +1 lens written with n_z of the other side."""
    import microscopy as mi
    from micro_pipeline import check_images
    case = ("microsbf", 0.8, 0.5, 0)
    geo, lens = mi.direct_geometry, mi.lens_field_direct

    def geo_threshold(nfft2d, k0, d):
        k = 0
        while True:
            dkx = 2.0 * np.pi / (nfft2d * d) / 2.0 ** k
            if mi.nint(k0 / dkx) + 1 > 4:
                return mi.nint(k0 / dkx) + 1, dkx
            k += 1

    def lens_phase(pos, FF, k0, dk, imax, numaper, nside=1):
        return np.conj(lens(pos, np.conj(FF), k0, dk, imax, numaper, nside))

    def lens_side(pos, FF, k0, dk, imax, numaper, nside=1):
        return lens(pos, FF, k0, dk, imax, numaper, -nside)

    for name, fn in (("direct_geometry", geo_threshold), ("lens_field_direct", lens_phase),
                     ("lens_field_direct", lens_side)):
        with monkeypatch.context() as mp:
            mp.setattr(mi, name, fn)
            img, fx, key = _fixture_case(case, oracle)
        with pytest.raises(AssertionError):
            check_images(img, fx, key, 1e-8)
    assert mi.direct_geometry is geo and mi.lens_field_direct is lens
