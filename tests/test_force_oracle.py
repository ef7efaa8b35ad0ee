"""The optical-force checker (oracle/force.py) pinned to the compiled
reference and to physics: the Cartesian d T / dx of the restatement against
propesplibdermat at random and on-axis points; derivativefield2's FFT
convolution against the direct sum; the plane-wave derivative against
ondeplaned; the config-1 force of the reference's own fields against the Mie
radiation pressure (Cext - g Csca) I0 / (8 pi) 4 pi eps0, and the zero torque
and transverse force that the sphere's symmetry fixes."""
import numpy as np
import pytest

from conftest import golden

LAM = 632.8e-9
K0 = 2 * np.pi / LAM
D = 20e-9


def test_dT_matches_propesplibdermat(ref):
    import force as fo
    rng = np.random.default_rng(1)
    worst = 0.0
    for t in range(60):
        ro, rs = rng.normal(size=3) * 3 * D, rng.normal(size=3) * 3 * D
        if t % 6 == 0:
            rs[:2] = ro[:2]  # on the z axis through the source (the reference's a = 0 branch)
        for test in (1, 2, 3):
            S = ref.propder(ro, rs, K0, test)
            worst = max(worst, np.abs(fo.dT(ro - rs, K0, test - 1) - S).max() / np.abs(S).max())
    assert worst < 1e-13, worst


def test_derivativefield2_is_the_direct_sum(ref):
    import force as fo
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(3)
    N = nx * ny * nz
    mask = rng.random(N) < 0.6
    nd = int(mask.sum())
    tab = np.zeros(N, np.int32)
    tab[mask] = np.arange(1, nd + 1)
    FF = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    cells = np.nonzero(mask)[0]
    allc = np.arange(N)
    grid = lambda c: np.stack([c % nx, (c // nx) % ny, c // (nx * ny)], 1) * D  # noqa: E731
    for test in (1, 2, 3):
        out = ref.dfield(FF, tab, nx, ny, nz, D, K0, test).reshape(N, 3)
        direct = fo.dfield_direct(grid(allc), grid(cells), FF, K0, test - 1)
        assert np.abs(out - direct).max() < 1e-12 * np.abs(direct).max()


def test_plane_wave_derivative_is_ik_E(ref):
    r = np.array([1e-8, 2e-8, -3e-8])
    th, ph, pp, E0 = 20.0, 40.0, 0.3, 1.0 + 0.5j
    Eder = ref.ondeplaned(r, K0, E0, 0.0, pp, th, ph)
    t, p = np.radians(th), np.radians(ph)
    k = K0 * np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])
    s = np.sqrt(1 - pp) * E0
    pa = np.sqrt(pp) * E0
    E = np.array([-s * np.sin(p) + pa * np.cos(t) * np.cos(p), s * np.cos(p) + pa * np.cos(t) * np.sin(p),
                  -pa * np.sin(t)]) * np.exp(1j * k @ r)
    assert np.abs(Eder - 1j * np.outer(E, k)).max() < 1e-12 * np.abs(Eder).max()


def test_config1_force_near_mie_radiation_pressure():
    import force as fo
    import mie
    g = golden("force.npz")
    a = 200e-9
    qext, qsca, gg = mie.bhmie(K0 * a, 1.5 + 0j)
    Fm = (qext - gg * qsca) * np.pi * a * a / (8 * np.pi) * fo.QUATPIEPS0
    F = g["cfg1_F"]
    assert abs(F[2] / Fm - 1) < 0.04  # DDA at 20 cells per diameter: 2.4 %
    assert np.abs(F[:2]).max() < 1e-10 * F[2]
    assert np.abs(g["cfg1_tau"]).max() < 1e-10 * F[2] * a


def test_force_restatement_reproduces_golden_box():
    """The fixture's per-dipole forces are force_torque of its own inputs
    (guards the fixture against a later change of the restatement)."""
    import force as fo
    g = golden("force.npz")
    tab = g["box_tab"]
    nx, ny = 6, 5
    ci = np.nonzero(tab)[0][np.argsort(tab[tab > 0])]
    cells = np.stack([ci % nx, (ci // nx) % ny, ci // (nx * ny)], 1)
    fdip, tdip, F, tau = fo.force_torque(list(g["box_dE"]), g["box_Einc"], g["box_FF"], g["box_FFloc"],
                                         g["box_kvec"], K0, cells, D)
    assert np.array_equal(fdip, g["box_fdip"]) and np.array_equal(tdip, g["box_tdip"])
