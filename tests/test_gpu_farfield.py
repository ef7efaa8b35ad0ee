"""Outputs from solved fields on the GPU against the reference's fixtures and
the oracle: far field (diffractefft2d), Csca and g (computegcfft2d), cross
sections (cdmlib.f:2725-2753); FP64, 1e-12 relative.  End to end: the GPU
solve of config 1 gives Cext within 2% of Mie."""
import ctypes as C

import numpy as np
import pytest

from conftest import golden, relerr

pytestmark = pytest.mark.gpu

LAM = 632.8e-9
K0 = 2 * np.pi / LAM
D = 20e-9


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.complex128).view(np.float64).copy()).cuda()


def _host(t):
    return t.cpu().numpy().view(np.complex128)


def _farfield(sv, FF, R, N, k0, d, origin):
    """Device far field of R RHS; returns E[r, ii, jj, c] (pos, neg), imax, dk."""
    import torch
    imax, dk = sv.farfield_size(N, k0, d)
    n = 2 * imax + 1
    Ep = torch.zeros(R * 3 * n * n * 2, dtype=torch.float64, device="cuda")
    En = torch.zeros_like(Ep)
    Fd = _dev(FF)
    im2, dk2 = sv.farfield_device(R, Fd.data_ptr(), N, k0, d, origin, Ep.data_ptr(), En.data_ptr())
    torch.cuda.synchronize()
    assert (im2, dk2) == (imax, dk)
    cut = lambda t: np.transpose(_host(t).reshape(R, 3, n, n), (0, 3, 2, 1))  # [r][c][jj][ii] -> [r][ii][jj][c]
    return cut(Ep), cut(En), imax, dk, (Ep, En)


@pytest.mark.parametrize("N", [64, 65])
def test_farfield_dense_box_matches_reference_fixture(gpu, N):
    f = golden("farfield.npz")
    with gpu.Solver(6, 5, 4) as sv:
        print(f"far field 6x5x4 nfft2d={N}", flush=True)
        sv.set_dipoles(None)
        origin = (float(f["box_x0"]), float(f["box_y0"]), float(f["box_zs"][0]))
        Ep, En, imax, dk, dev = _farfield(sv, f["box_FF"], 1, N, K0, D, origin)
        assert imax == int(f[f"box{N}_imax"])
        scale = np.abs(f[f"box{N}_Epos"]).max()
        assert relerr(Ep[0], f[f"box{N}_Epos"], scale) < 1e-12
        assert relerr(En[0], f[f"box{N}_Eneg"], scale) < 1e-12
        cs = sv.computegcfft2d_device(1, dev[0].data_ptr(), dev[1].data_ptr(), imax, dk, K0, 1.0,
                                      20.0, 30.0)
        assert abs(cs[0, 0] / float(f[f"box{N}_csca"]) - 1) < 1e-12
        assert abs(cs[0, 1] - float(f[f"box{N}_g"])) < 1e-12


def test_farfield_sparse_config1_matches_reference_fixture(gpu):
    """Dipole-ordered moments placed through Tabdip reproduce the reference's
    box-ordered far field of the config-1 solution (nfft2d 128)."""
    from ifdda import scene
    s, k0, alpha, E = scene.config1()
    g = golden("solve_cfg1.npz")
    f = golden("farfield.npz")
    p = (alpha[:, None] * g["GPBICG1_E"].reshape(-1, 3)).reshape(-1)
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.set_dipoles(s.tabdip)
        Ep, En, imax, dk, dev = _farfield(sv, p, 1, 128, k0, s.d, s.origin)
        scale = np.abs(f["cfg1_Epos"]).max()
        assert relerr(Ep[0], f["cfg1_Epos"], scale) < 1e-12
        assert relerr(En[0], f["cfg1_Eneg"], scale) < 1e-12
        cs = sv.computegcfft2d_device(1, dev[0].data_ptr(), dev[1].data_ptr(), imax, dk, k0, 1.0,
                                      0.0, 0.0)
        assert abs(cs[0, 0] / float(f["cfg1_csca"]) - 1) < 1e-12


def test_cross_sections_and_multi_rhs_far_field(gpu):
    """Three incidences at once: each RHS's cross sections and far field equal
    the oracle's restatement of that RHS alone."""
    import farfield as ff
    from ifdda import scene
    s, k0, alpha, _ = scene.config1()
    rng = np.random.default_rng(2)
    R = 3
    E0 = np.stack([scene.plane_wave(s.pos, k0, t, ph, pp=pp)
                   for t, ph, pp in ((0.0, 0.0, 1.0), (30.0, 45.0, 0.5), (60.0, 200.0, 0.0))])
    El = E0 * (1.0 + 0.2 * rng.standard_normal(E0.shape)) + 0.1j * rng.standard_normal(E0.shape)
    FF = (np.repeat(alpha, 3)[None, :] * El)
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.set_dipoles(s.tabdip)
        dE0, dFF, dEl = _dev(E0), _dev(FF), _dev(El)
        out = sv.cross_sections_device(R, dE0.data_ptr(), dFF.data_ptr(), dEl.data_ptr(), k0, [1.0, 2.0, 0.5])
        for r, i0 in zip(range(R), (1.0, 2.0, 0.5)):
            ref = ff.cross_sections(E0[r], FF[r], El[r], k0, i0)
            # Cabs cancels (lossless RR dipoles absorb nothing): compare on the
            # scale of the summed terms, not of the result
            scale = 4 * np.pi * k0 / i0 * (np.abs(E0[r] * FF[r]).sum() + np.abs(FF[r] * El[r]).sum()
                                           + 2.0 / 3.0 * k0 ** 3 * (np.abs(FF[r]) ** 2).sum())
            assert np.abs(np.asarray(out[r]) - np.asarray(ref)).max() < 1e-13 * scale, (out[r], ref)
        Ep, En, imax, dk, _ = _farfield(sv, FF, R, 96, k0, s.d, s.origin)
        zs = s.origin[2] + s.d * np.arange(s.nz)
        for r in range(R):
            FFbox = np.zeros((s.nx * s.ny * s.nz, 3), np.complex128)
            cells = np.nonzero(s.tabdip)[0]
            FFbox[cells] = FF[r].reshape(-1, 3)[s.tabdip[cells] - 1]
            a, b, im, _ = ff.diffractefft2d(FFbox.reshape(-1), s.nx, s.ny, s.nz, 96, k0, s.origin[0],
                                            s.origin[1], zs, s.d)
            scale = np.abs(a).max()
            assert im == imax and relerr(Ep[r], a, scale) < 1e-12 and relerr(En[r], b, scale) < 1e-12


def test_farfield_errors_follow_the_reference(gpu):
    with gpu.Solver(6, 5, 4) as sv:
        sv.set_dipoles(None)
        with pytest.raises(gpu.IfddaError, match="too small"):
            sv.farfield_size(8, K0, D)
        with pytest.raises(gpu.IfddaError, match="too large"):
            sv.farfield_size(20000, K0, D)


def test_reference_shaped_diffractefft2d(gpu):
    """ifdda_diffractefft2d / ifdda_computegcfft2d with the reference's host
    arrays (Fortran (nfft2d, nfft2d, 3) layout) against the fixtures."""
    f = golden("farfield.npz")
    L = gpu.lib()
    nx, ny, nz, N = 6, 5, 4, 64
    ncell = nx * ny * nz
    xs = np.full(ncell, float(f["box_x0"]))
    ys = np.full(ncell, float(f["box_y0"]))
    zs = np.asarray(f["box_zs"])[np.arange(ncell) // (nx * ny)].copy()
    F = np.ascontiguousarray(f["box_FF"], np.complex128)
    tab = np.zeros(N, np.int32)
    Ep = np.zeros((3, N, N), np.complex128)
    En = np.zeros((3, N, N), np.complex128)
    scr = np.zeros(N * N, np.complex128)
    im, nstop = C.c_int(0), C.c_int(0)
    dkx, dky = C.c_double(0.0), C.c_double(0.0)
    info = C.create_string_buffer(b" " * 64, 64)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    ri = lambda v: C.byref(C.c_int(v))
    rd = lambda v: C.byref(C.c_double(v))
    L.ifdda_diffractefft2d(ri(nx), ri(ny), ri(nz), ri(nx), ri(ny), ri(nz), ri(N), P(tab), rd(K0),
                           P(xs), P(ys), P(zs), rd(D), P(scr), P(scr), P(scr), P(F), C.byref(im),
                           C.byref(dkx), C.byref(dky), P(Ep), P(En), None, None, C.byref(nstop), info,
                           C.c_size_t(64))
    assert nstop.value == 0 and im.value == int(f["box64_imax"])
    n = 2 * im.value + 1
    cut = lambda A: np.transpose(A[:, :n, :n], (2, 1, 0))
    scale = np.abs(f["box64_Epos"]).max()
    assert relerr(cut(Ep), f["box64_Epos"], scale) < 1e-12
    assert relerr(cut(En), f["box64_Eneg"], scale) < 1e-12
    assert np.abs(Ep[:, n:, :]).max() == 0.0  # outside the corner: zero
    g, cs = C.c_double(0.0), C.c_double(0.0)
    L.ifdda_computegcfft2d(C.byref(im), C.byref(dkx), C.byref(dky), rd(K0), rd(1.0), rd(20.0), rd(30.0),
                           ri(N), P(Ep), P(En), C.byref(g), C.byref(cs))
    assert abs(cs.value / float(f["box64_csca"]) - 1) < 1e-12
    assert abs(g.value - float(f["box64_g"])) < 1e-12
    # an error of the reference: nfft2d too small
    L.ifdda_diffractefft2d(ri(nx), ri(ny), ri(nz), ri(nx), ri(ny), ri(nz), ri(8), P(tab), rd(K0),
                           P(xs), P(ys), P(zs), rd(D), P(scr), P(scr), P(scr), P(F), C.byref(im),
                           C.byref(dkx), C.byref(dky), P(Ep), P(En), None, None, C.byref(nstop), info,
                           C.c_size_t(64))
    assert nstop.value == 1 and info.raw.decode().startswith("In FFT Poynting nfft2d too small")


def test_config1_end_to_end_vs_mie(gpu):
    """GPU solve (GPBICG, tol 1e-4, x0 = 0), GPU cross sections and far field:
    Cext within 2% of Mie, Cabs = 0, far-field Csca within 3% of Cext
    (nfft2d 512), g within 2% of Mie."""
    import torch
    from ifdda import scene
    from mie import bhmie
    s, k0, alpha, E = scene.config1()
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.greencalculfft(s.d, k0)
        sv.set_dipoles(s.tabdip)
        sv.set_polarizability(alpha)
        dE = _dev(E)
        FF = torch.empty_like(dE)
        FL = torch.empty_like(dE)
        info = sv.solve_device(dE.data_ptr(), FF.data_ptr(), FL.data_ptr(), "GPBICG1", 1e-4, x0_zero=True)
        assert info["nstop"] == 0 and info["tol1"] <= 1e-4
        cext, cabs, csca = sv.cross_sections_device(1, dE.data_ptr(), FF.data_ptr(), FL.data_ptr(), k0, 1.0)[0]
        imax, dk = sv.farfield_size(512, k0, s.d)
        n = 2 * imax + 1
        Ep = torch.zeros(3 * n * n * 2, dtype=torch.float64, device="cuda")
        En = torch.zeros_like(Ep)
        sv.farfield_device(1, FF.data_ptr(), 512, k0, s.d, s.origin, Ep.data_ptr(), En.data_ptr())
        cs, gg = sv.computegcfft2d_device(1, Ep.data_ptr(), En.data_ptr(), imax, dk, k0, 1.0, 0.0, 0.0)[0]
    qext, qsca, gmie = bhmie(k0 * 200e-9, np.sqrt(2.25 + 0j), 200)
    area = np.pi * (200e-9) ** 2
    assert abs(cext / (qext * area) - 1) < 0.02
    assert cabs == 0.0 and csca == cext
    assert abs(cs / cext - 1) < 0.03
    assert abs(gg / gmie - 1) < 0.02
