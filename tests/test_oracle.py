"""Pins of the oracle (CPU): each check is chosen so that a dropped term, a
wrong sign or index, or a transposed operand in the restatement fails one.

  T(r)          far-field limit, static near-field limit, symmetry, parity,
                the compiled reference (when built here)
  FFT           numpy.fft, brute-force DFT, Parseval
  Green G       golden fixtures from the reference, brute-force DFT of the
                spatial fill, even/odd parity
  matvec        golden fixtures, the O(Nd^2) direct sum, linearity,
                reciprocity A^T = A for scalar alpha
  quadrature    golden fixtures; Simpson limit vs the point value far away
  solvers       golden reference solves (iteration counts, fields), dense
                numpy.linalg.solve, zero-RHS breakdown
  Mie           BHMIE cross sections vs DDA on config 1 (a few %)
  far field     the numpy restatement of diffractefft2d / computegcfft2d
                vs the reference's golden fixtures (even and odd nfft2d,
                config 1) and the compiled reference; physics: the far-field
                Csca of config 1 converges to Cext (optical theorem for a
                lossless sphere) and g to Mie's
"""
import numpy as np
import pytest

from conftest import golden, relerr

LAM = 632.8e-9
K0 = 2 * np.pi / LAM
D = 20e-9


# ------------------------------------------------------------------ T(r)
def test_propa_far_field_limit(oracle):
    """|R| >> lambda: T -> k^2 e^{ikR}/R (I - RR), transverse."""
    R = np.array([3.0, -4.0, 12.0]) * 1e-3  # 13 mm
    T = oracle.propa((0, 0, 0), -R, K0)  # R = obs - src
    r = np.linalg.norm(R)
    u = R / r
    far = K0 ** 2 * np.exp(1j * K0 * r) / r * (np.eye(3) - np.outer(u, u))
    assert relerr(T, far) < 1e-4
    assert np.abs(T @ u).max() < 1e-4 * np.abs(T).max()


def test_propa_static_limit(oracle):
    """k R -> 0: T -> (3 RR - I)/R^3 (dipole field)."""
    R = np.array([1.0, 2.0, -2.0]) * 1e-12
    T = oracle.propa(R, (0, 0, 0), K0)
    r = np.linalg.norm(R)
    u = R / r
    stat = (3 * np.outer(u, u) - np.eye(3)) / r ** 3
    assert relerr(T, stat) < 1e-9


def test_propa_symmetry_and_parity(oracle):
    R = np.array([0.3, -0.7, 1.1]) * 1e-7
    T = oracle.propa(R, (0, 0, 0), K0)
    assert np.allclose(T, T.T, rtol=0, atol=1e-15 * np.abs(T).max())
    assert relerr(oracle.propa(-R, (0, 0, 0), K0), T) < 1e-15
    assert np.all(oracle.propa((0, 0, 0), (0, 0, 0), K0) == 0)


def test_propa_matches_reference(oracle, ref):
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, b = rng.standard_normal(3) * 1e-7, rng.standard_normal(3) * 1e-7
        assert relerr(oracle.propa(a, b, K0), ref.propa(a, b, K0)) < 1e-14


def test_propa_int_matches_reference(oracle, ref):
    for off in [(1, 0, 0), (1, 1, 0), (2, 1, 1), (0, 0, 1)]:
        o = np.array(off) * D
        assert relerr(oracle.propa_int((0, 0, 0), o, K0, D), ref.propa_int((0, 0, 0), o, K0, D)) < 1e-12


def test_propa_int_far_cell_is_point_value(oracle):
    """Beyond lim*lambda the integrated tensor is the point value; just
    inside, the Simpson average differs from it by O((d/R)^2)."""
    o = np.array([40.0, 0, 0]) * D  # 800 nm > lambda
    assert relerr(oracle.propa_int((0, 0, 0), o, K0, D), oracle.propa((0, 0, 0), o, K0)) == 0.0
    o = np.array([20.0, 0, 0]) * D  # 400 nm < lambda
    Ti = oracle.propa_int((0, 0, 0), o, K0, D)
    Tp = oracle.propa((0, 0, 0), o, K0)
    assert 1e-6 < relerr(Ti, Tp) < 1e-2


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(4, 6, 10), (5, 9, 8), (7, 3, 2)])
def test_fft3d_vs_numpy(oracle, shape):
    rng = np.random.default_rng(1)
    v = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    # sign -1 is numpy's forward transform; sign +1 is n * ifft
    assert relerr(oracle.fft3d(v, -1), np.fft.fftn(v)) < 1e-13
    assert relerr(oracle.fft3d(v, +1), np.fft.ifftn(v) * v.size) < 1e-13


def test_fft_brute_force_and_parseval(oracle):
    n = 15
    rng = np.random.default_rng(2)
    v = rng.standard_normal((1, 1, n)) + 1j * rng.standard_normal((1, 1, n))
    k = np.arange(n)
    W = np.exp(2j * np.pi * np.outer(k, k) / n)
    assert relerr(oracle.fft3d(v, +1)[0, 0], W @ v[0, 0]) < 1e-13
    F = oracle.fft3d(v, -1)
    assert abs(np.sum(np.abs(F) ** 2) / n - np.sum(np.abs(v) ** 2)) < 1e-12 * np.sum(np.abs(v) ** 2)


# ----------------------------------------------------------------- Green
def test_green_matches_golden(oracle):
    g = golden("green_654.npz")
    for nq in (0, 1, 2):
        G = oracle.green_fft(6, 5, 4, (12, 10, 8), D, K0, nquad=nq)
        assert relerr(G, g[f"G_nquad{nq}"]) < 1e-12, nq


def test_green_is_dft_of_the_fill(oracle):
    """Brute-force: G_hat(k) = sum_r T(offset(r) d) e^{+2 pi i k.r/L} on a tiny grid."""
    nx, ny, nz = 3, 2, 2
    L = (6, 4, 4)
    G = oracle.green_fft(nx, ny, nz, L, D, K0)
    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    fill = np.zeros((6, L[2], L[1], L[0]), complex)

    def off(i, n, Ln):
        return i if i < n else (i - Ln if i > Ln - n else None)
    for k in range(L[2]):
        for j in range(L[1]):
            for i in range(L[0]):
                o = (off(i, nx, L[0]), off(j, ny, L[1]), off(k, nz, L[2]))
                if None in o:
                    continue
                T = oracle.propa((0, 0, 0), np.array(o) * D, K0)
                for c, (a, b) in enumerate(comps):
                    fill[c, k, j, i] = T[a, b]
    ref = np.fft.ifftn(fill, axes=(1, 2, 3)) * np.prod(L)
    assert relerr(G, ref) < 1e-13


def test_green_parity(oracle):
    G = oracle.green_fft(6, 5, 4, (12, 10, 8), D, K0)
    mirror = lambda a, ax: np.roll(np.flip(a, axis=ax), 1, axis=ax)  # noqa: E731  k -> L - k
    # axes of G[c]: 0 = z, 1 = y, 2 = x
    par = {0: (1, 1, 1), 1: (1, -1, -1), 2: (-1, 1, -1), 3: (1, 1, 1), 4: (-1, -1, 1), 5: (1, 1, 1)}
    for c, (pz, py, px) in par.items():
        for ax, p in ((0, pz), (1, py), (2, px)):
            assert relerr(mirror(G[c], ax), p * G[c], np.abs(G[c]).max()) < 1e-13


# ---------------------------------------------------------------- matvec
def test_matvec_matches_golden(oracle):
    f = golden("matvec_654.npz")
    G = oracle.green_fft(6, 5, 4, (12, 10, 8), D, K0)
    y = oracle.matvec(G, f["tabdip"], 6, 5, 4, f["FFX"], f["FFB"])
    assert relerr(y, f["out3"], np.abs(f["out3"] - f["FFB"]).max()) < 1e-13
    y = oracle.matvec(G, None, 6, 5, 4, f["FFX_dense"], f["FFB_dense"])
    assert relerr(y, f["out_opt"], np.abs(f["out_opt"] - f["FFB_dense"]).max()) < 1e-13
    g = golden("green_91012.npz")
    G = oracle.green_fft(9, 10, 12, (18, 20, 24), D, K0)
    assert relerr(G, g["G"]) < 1e-12
    y = oracle.matvec(G, g["tabdip"], 9, 10, 12, g["FFX"], g["FFB"])
    assert relerr(y, g["out3"], np.abs(g["out3"] - g["FFB"]).max()) < 1e-13


def test_matvec_direct_sum_and_padding(oracle):
    """FFT convolution = direct dipole sum, on 2n and on larger padded grids."""
    nx, ny, nz = 5, 4, 3
    f = golden("matvec_654.npz")
    rng = np.random.default_rng(3)
    N = nx * ny * nz
    mask = rng.random(N) < 0.7
    nd = int(mask.sum())
    tab = np.zeros(N, np.int32)
    tab[mask] = np.arange(1, nd + 1)
    x = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    b = np.zeros(3 * nd, complex)
    ref = oracle.matvec_direct(tab, nx, ny, nz, D, K0, x, b)
    for L in [(10, 8, 6), (12, 9, 8), (15, 10, 6)]:
        G = oracle.green_fft(nx, ny, nz, L, D, K0)
        assert relerr(oracle.matvec(G, tab, nx, ny, nz, x, b), ref) < 1e-13
    assert f["out3"].size  # fixture present


def test_matvec_linear_and_reciprocal(oracle):
    """A = I - T alpha with scalar alpha: T is symmetric (x^T T y = y^T T x)."""
    nx, ny, nz = 4, 4, 3
    G = oracle.green_fft(nx, ny, nz, (8, 8, 6), D, K0)
    rng = np.random.default_rng(4)
    N = nx * ny * nz
    x, y = (rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N) for _ in range(2))
    z = np.zeros(3 * N, complex)
    Tx = -oracle.matvec(G, None, nx, ny, nz, x, z)
    Ty = -oracle.matvec(G, None, nx, ny, nz, y, z)
    assert abs(y @ Tx - x @ Ty) < 1e-12 * abs(y @ Tx)
    T2 = -oracle.matvec(G, None, nx, ny, nz, 2 * x + 3j * y, z)
    assert relerr(T2, 2 * Tx + 3j * Ty) < 1e-13


# ---------------------------------------------------------------- solvers
@pytest.mark.parametrize("method", ["GPBICG1", "QMRBICGSTAB1", "QMRBICGSTAB2", "TFQMR"])
def test_solvers_match_reference_golden(oracle, method):
    import ifdda.scene as scene
    g = golden("solve_cfg1.npz")
    s, k0, alpha, E = scene.config1()
    assert np.array_equal(s.tabdip, g["tabdip"]) and relerr(alpha, g["alpha"]) < 1e-15
    G = oracle.green_fft(20, 20, 20, (40, 40, 40), s.d, k0)
    pol = np.zeros((8000, 3, 3), complex)
    for a in range(3):
        pol[: s.nd, a, a] = alpha
    o = oracle.inverserig(G, s.tabdip, 20, 20, 20, pol, E, method, 1e-4)
    assert o["nloop"] == int(g[f"{method}_nloop"])
    assert o["ncompte"] == int(g[f"{method}_ncompte"])
    assert abs(o["tol1"] - float(g[f"{method}_tol1"])) < 1e-9
    assert relerr(o["FFloc"], g[f"{method}_E"]) < 1e-11


@pytest.mark.parametrize("method", ["GPBICG1", "QMRBICGSTAB1", "TFQMR"])
@pytest.mark.parametrize("tag,tol", [("4", 1e-4), ("6", 1e-6)])
def test_dense_box_solvers_match_inverserigopt_golden(oracle, method, tag, tol):
    """inverserigopt.f: every box cell is a dipole (no map), alpha = 0 outside
    the object; the oracle's dense-box driver reproduces the compiled
    reference's iteration and matvec counts, tol1 and field."""
    g = golden("solve_opt.npz")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    G = oracle.green_fft(nx, ny, nz, (2 * nx, 2 * ny, 2 * nz), float(g["d"]), float(g["k0"]))
    pol = np.zeros((nx * ny * nz, 3, 3), complex)
    for a in range(3):
        pol[:, a, a] = g["alpha"]
    o = oracle.inverserig(G, None, nx, ny, nz, pol, g["E_inc"], method, tol)
    key = f"{method}_{tag}"
    assert o["nstop"] == 0
    assert o["nloop"] == int(g[f"{key}_nloop"])
    assert o["ncompte"] == int(g[f"{key}_ncompte"])
    assert abs(o["tol1"] / float(g[f"{key}_tol1"]) - 1) < 1e-8
    assert relerr(o["FFloc"], g[f"{key}_E"]) < 1e-11
    assert relerr(o["FF"], g[f"{key}_FF"]) < 1e-11


def test_solver_vs_dense_solve(oracle):
    import ifdda.scene as scene
    s = scene.sphere(4, 40e-9, 4.0 + 1.0j)
    alpha = scene.polarizability(s.eps, s.d, K0, "RR")
    E = scene.plane_wave(s.pos, K0, 30.0, 10.0, pp=0.5)
    G = oracle.green_fft(4, 4, 4, (8, 8, 8), s.d, K0)
    n = 3 * s.nd
    A = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1
        A[:, j] = oracle.matvec(G, s.tabdip, 4, 4, 4, np.repeat(alpha, 3) * e, e)
    xs = np.linalg.solve(A, E)
    pol = np.zeros((64, 3, 3), complex)
    for a in range(3):
        pol[: s.nd, a, a] = alpha
    for m in ["GPBICG1", "QMRBICGSTAB1", "QMRBICGSTAB2", "TFQMR"]:
        o = oracle.inverserig(G, s.tabdip, 4, 4, 4, pol, E, m, 1e-11)
        assert o["tol1"] < 1e-10 and relerr(o["FFloc"], xs) < 1e-9, m


def test_zero_rhs_breakdown(oracle):
    import ifdda.scene as scene
    s = scene.sphere(4, 40e-9, 2.0)
    G = oracle.green_fft(4, 4, 4, (8, 8, 8), s.d, K0)
    pol = np.zeros((64, 3, 3), complex)
    o = oracle.inverserig(G, s.tabdip, 4, 4, 4, pol, np.zeros(3 * s.nd, complex), "GPBICG1", 1e-4)
    assert o["nstop"] == 1 and o["status"] == -1 and o["steperr"] == 1


# ------------------------------------------------------------------- Mie
def test_mie_cross_sections_vs_dda(oracle):
    """BHMIE (Bohren-Huffman) vs DDA Cext on the 400 nm sphere: the DDA at
    20 cells per diameter is within a few percent; Cabs = 0 for real eps."""
    import ifdda.scene as scene
    from mie import bhmie
    g = golden("solve_cfg1.npz")
    s, k0, alpha, E = scene.config1()
    Eloc = g["GPBICG1_E"].reshape(-1, 3)
    p = alpha[:, None] * Eloc
    E0 = E.reshape(-1, 3)
    # cdmlib.f:2730-2746
    cext = 4 * np.pi * k0 * np.sum(np.imag(np.sum(np.conj(E0) * p, axis=1)))
    cabs = 4 * np.pi * k0 * np.sum(np.imag(np.sum(p * np.conj(Eloc), axis=1))
                                   - 2.0 / 3.0 * k0 ** 3 * np.sum(np.abs(p) ** 2, axis=1))
    qext, qsca, gsca = bhmie(k0 * 200e-9, np.sqrt(2.25 + 0j), 200)
    cext_mie = qext * np.pi * (200e-9) ** 2
    assert abs(cext / cext_mie - 1) < 0.05
    assert abs(cabs) < 1e-3 * cext
    assert abs(qext - qsca) < 1e-10  # lossless sphere: extinction = scattering


# ------------------------------------------------------------- far field
def _farfield_box_case():
    f = golden("farfield.npz")
    return f, 6, 5, 4


@pytest.mark.parametrize("N", [64, 65])
def test_farfield_restatement_matches_reference_fixture(N):
    import farfield as ff
    f, nx, ny, nz = _farfield_box_case()
    Ep, En, imax, dk = ff.diffractefft2d(f["box_FF"], nx, ny, nz, N, K0, float(f["box_x0"]),
                                         float(f["box_y0"]), f["box_zs"], D)
    assert imax == int(f[f"box{N}_imax"]) and abs(dk / float(f[f"box{N}_dk"]) - 1) < 1e-15
    scale = np.abs(f[f"box{N}_Epos"]).max()
    assert relerr(Ep, f[f"box{N}_Epos"], scale) < 1e-12
    assert relerr(En, f[f"box{N}_Eneg"], scale) < 1e-12
    cs, g = ff.computegcfft2d(Ep, En, imax, dk, K0, 1.0, 20.0, 30.0)
    assert abs(cs / float(f[f"box{N}_csca"]) - 1) < 1e-12
    assert abs(g - float(f[f"box{N}_g"])) < 1e-12


def test_farfield_restatement_matches_compiled_reference(ref):
    import farfield as ff
    rng = np.random.default_rng(5)
    nx, ny, nz = 7, 4, 3
    F = rng.standard_normal(3 * nx * ny * nz) + 1j * rng.standard_normal(3 * nx * ny * nz)
    zs = 10e-9 + D * np.arange(nz)
    for N in (50, 96):
        a = ff.diffractefft2d(F, nx, ny, nz, N, K0, -70e-9, 5e-9, zs, D)
        b = ref.diffractefft2d(F, nx, ny, nz, N, K0, -70e-9, 5e-9, zs, D)
        assert b[4] == 0 and a[2] == b[2]
        scale = np.abs(b[0]).max()
        assert relerr(a[0], b[0], scale) < 1e-12 and relerr(a[1], b[1], scale) < 1e-12
        ca = ff.computegcfft2d(a[0], a[1], a[2], a[3], K0, 2.0, 35.0, 120.0)
        cb = ref.computegcfft2d(b[0], b[1], b[2], b[3], K0, 2.0, 35.0, 120.0, N)
        assert abs(ca[0] / cb[0] - 1) < 1e-12 and abs(ca[1] - cb[1]) < 1e-12
    # the reference's error paths
    with pytest.raises(ValueError, match="too small"):
        ff.diffractefft2d(F, nx, ny, nz, 8, K0, 0.0, 0.0, zs, D)
    assert ref.diffractefft2d(F, nx, ny, nz, 8, K0, 0.0, 0.0, zs, D)[4] == 1


def test_farfield_config1_physics():
    """Config 1 (lossless sphere): the far-field Csca integrates to Cext within
    the k-grid quadrature (2% at nfft2d = 512), g within 2% of Mie's, and the
    restatement reproduces the reference's nfft2d = 128 fixture."""
    import farfield as ff
    import ifdda.scene as scene
    from mie import bhmie
    s, k0, alpha, E = scene.config1()
    g = golden("solve_cfg1.npz")
    f = golden("farfield.npz")
    El = g["GPBICG1_E"]
    p = (alpha[:, None] * El.reshape(-1, 3)).reshape(-1)
    Cext, Cabs, Csca = ff.cross_sections(E, p, El, k0, 1.0)
    assert Cabs == 0.0 and Csca == Cext
    FFbox = np.zeros((s.nx * s.ny * s.nz, 3), np.complex128)
    cells = np.nonzero(s.tabdip)[0]
    FFbox[cells] = p.reshape(-1, 3)[s.tabdip[cells] - 1]
    zs = s.origin[2] + s.d * np.arange(s.nz)
    Ep, En, im, dk = ff.diffractefft2d(FFbox.reshape(-1), s.nx, s.ny, s.nz, 128, k0, s.origin[0],
                                       s.origin[1], zs, s.d)
    scale = np.abs(f["cfg1_Epos"]).max()
    assert relerr(Ep, f["cfg1_Epos"], scale) < 1e-12 and relerr(En, f["cfg1_Eneg"], scale) < 1e-12
    Ep, En, im, dk = ff.diffractefft2d(FFbox.reshape(-1), s.nx, s.ny, s.nz, 512, k0, s.origin[0],
                                       s.origin[1], zs, s.d)
    cs, gg = ff.computegcfft2d(Ep, En, im, dk, k0, 1.0, 0.0, 0.0)
    assert abs(cs / Cext - 1) < 0.03
    _, _, gmie = bhmie(k0 * 200e-9, np.sqrt(2.25 + 0j), 200)
    assert abs(gg / gmie - 1) < 0.02


def test_inputs_match_compiled_reference(ref):
    """The restatements that form every solve's inputs here -- scene.plane_wave
    (ondeplane.f), scene.circular_wave (ondecirce.f), scene.polarizability
    (poladiff, polarisabilite.f) -- against the compiled reference."""
    from ifdda import scene
    k0 = 2 * np.pi / 632.8e-9
    pos = np.random.default_rng(3).uniform(-1e-6, 1e-6, (6, 3))
    E0 = 0.7 - 0.2j
    for th, ph, pp in ((30.0, 10.0, 0.5), (-20.0, 200.0, 0.0), (0.0, 45.0, 1.0), (75.0, -30.0, 0.3)):
        a = scene.plane_wave(pos, k0, th, ph, pp=pp, E0=E0)
        b = ref.ondeplane(pos, k0, E0, pp, th, ph)
        assert np.abs(a - b).max() < 1e-14 * np.abs(b).max(), (th, ph, pp)
    for th, ph, h in ((30.0, 10.0, 1.0), (-20.0, 200.0, -1.0), (0.0, 45.0, 1.0)):
        a = scene.circular_wave(pos, k0, th, ph, helicity=h, E0=E0)
        b = ref.ondecirce(pos, k0, E0, h, th, ph)
        assert np.abs(a - b).max() < 1e-14 * np.abs(b).max(), (th, ph, h)
    for method in ("CM", "RR"):
        for eps in (2.25 + 0j, 4.0 - 0.5j, 1.5 + 0.3j):
            a = scene.polarizability(np.array([eps]), 20e-9, k0, method)[0]
            b = ref.poladiff(20e-9, eps, k0, method)
            assert abs(a - b) <= 1e-15 * abs(b), (method, eps)


@pytest.mark.parametrize("x,m", [(0.5, 1.5 + 0j), (1.99, 1.5 + 0j), (4.0, 1.33 + 0.01j),
                                 (7.5, 2.0 + 0.5j)])
def test_mie_matches_compiled_bhmie(ref, x, m):
    """oracle/mie.py against module_mie.f's BHMIE, compiled: Qext, Qsca, g."""
    from mie import bhmie
    qext, qsca, g = bhmie(x, m)
    rqext, rqsca, _, rg = ref.bhmie(x, m)
    assert abs(qext / rqext - 1) < 1e-12 and abs(qsca / rqsca - 1) < 1e-12
    assert abs(g - rg) < 1e-12 * max(1.0, abs(rg))


def test_config3_scene():
    """The brightfield batch object: 20 spheres on a 64^3 lattice, 2 RHS per
    incidence (p then s)."""
    from ifdda import scene
    s, k0, alpha, E = scene.config3(n_inc=4)
    assert E.shape == (8, 3 * s.nd)
    assert 30000 < s.nd < 40000  # 20 spheres of ~1760 cells
