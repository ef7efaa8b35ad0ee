"""Optical force and torque on the device (ifdda_force_device) against the
compiled reference's derivative fields with the force of cdmlib.f restated in
oracle/force.py: a seeded 6x5x4 box (per-dipole forces and torques), config 1
solved on the device (totals), a dense box against the direct sum, and R
right-hand sides against single calls."""
import numpy as np
import pytest

from conftest import golden, relerr

pytestmark = pytest.mark.gpu

LAM = 632.8e-9
K0 = 2 * np.pi / LAM
D = 20e-9


def test_force_box_matches_reference(gpu):
    g = golden("force.npz")
    tab = g["box_tab"]
    with gpu.Solver(6, 5, 4) as sv:
        print("force 6x5x4 seeded", flush=True)
        sv.set_dipoles(tab)
        o = sv.force(g["box_Einc"], g["box_FF"], g["box_FFloc"], g["box_kvec"], D, K0, per_dipole=True)
    assert relerr(o["fdip"][0], g["box_fdip"]) < 1e-11
    assert relerr(o["tdip"][0], g["box_tdip"]) < 1e-11
    assert relerr(o["force"][0], g["box_F"], np.abs(g["box_fdip"]).sum()) < 1e-11
    assert relerr(o["torque"][0], g["box_tau"], np.abs(g["box_tdip"]).sum()) < 1e-11


def test_force_config1_matches_reference(gpu):
    from ifdda import scene
    g = golden("force.npz")
    s, k0, alpha, E = scene.config1()
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.greencalculfft(s.d, k0)
        sv.set_dipoles(s.tabdip)
        sv.set_polarizability(alpha)
        r = sv.inverserig(E, "GPBICG1", 1e-8)
        assert r["nloop"] == int(g["cfg1_nloop"])
        print("force config1", flush=True)
        o = sv.force(E, r["FF"], r["FFloc"], [0.0, 0.0, k0], s.d, k0)
        # the same call twice: the derivative tensor is kept, results equal
        o2 = sv.force(E, r["FF"], r["FFloc"], [0.0, 0.0, k0], s.d, k0)
    F, Fr = o["force"][0], g["cfg1_F"]
    assert abs(F[2] / Fr[2] - 1) < 1e-8
    assert np.abs(F[:2]).max() < 1e-10 * Fr[2]
    assert np.abs(o["torque"][0]).max() < 1e-10 * Fr[2] * 200e-9
    assert np.array_equal(o["force"], o2["force"]) and np.array_equal(o["torque"], o2["torque"])


def test_force_dense_box_vs_direct_sum(gpu):
    """A dense box (every cell a dipole, no map): the direct-sum derivative
    field and the restated force (oracle/force.py)."""
    import force as fo
    nx, ny, nz = 4, 3, 5
    N = nx * ny * nz
    rng = np.random.default_rng(8)
    FF = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    FL = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    Ei = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    k = K0 * np.array([0.3, -0.4, np.sqrt(1 - 0.25)])
    c = np.arange(N)
    cells = np.stack([c % nx, (c // nx) % ny, c // (nx * ny)], 1)
    dEs = [fo.dfield_direct(cells * D, cells * D, FF, K0, dd) for dd in range(3)]
    fdip, tdip, F, tau = fo.force_torque(dEs, Ei, FF, FL, k, K0, cells, D)
    with gpu.Solver(nx, ny, nz) as sv:
        print("force dense 4x3x5", flush=True)
        sv.set_dipoles(None)
        o = sv.force(Ei, FF, FL, k, D, K0, per_dipole=True)
    assert relerr(o["fdip"][0], fdip) < 1e-11
    assert relerr(o["tdip"][0], tdip) < 1e-11


def test_force_many_rhs_equal_single_calls(gpu):
    g = golden("force.npz")
    rng = np.random.default_rng(4)
    nd = int(g["box_tab"].max())
    extra = [rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd) for _ in range(3)]
    Es = np.stack([g["box_Einc"], extra[0]])
    FFs = np.stack([g["box_FF"], extra[1]])
    FLs = np.stack([g["box_FFloc"], extra[2]])
    ks = np.stack([g["box_kvec"], g["box_kvec"][::-1]])
    with gpu.Solver(6, 5, 4) as sv:
        print("force two right-hand sides", flush=True)
        sv.set_dipoles(g["box_tab"])
        both = sv.force(Es, FFs, FLs, ks, D, K0, per_dipole=True)
        for r in range(2):
            one = sv.force(Es[r], FFs[r], FLs[r], ks[r], D, K0, per_dipole=True)
            assert np.array_equal(one["fdip"][0], both["fdip"][r])
            assert np.array_equal(one["force"][0], both["force"][r])
            assert np.array_equal(one["torque"][0], both["torque"][r])


def test_dropin_derivativefield2_matches_reference(gpu):
    """calculforcefft_ then derivativefield2_ through libifdda_dropin.so, as
    cdmlib.f:3159-3300 calls them: the derivative field at the dipole cells
    equals the compiled reference's (force.npz), and at the other cells the
    direct sum of oracle/force.py."""
    import ctypes as C
    import os
    import force as fo
    g = golden("force.npz")
    tab = np.ascontiguousarray(g["box_tab"], np.int32)
    nx, ny, nz = 6, 5, 4
    N = nx * ny * nz
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    L = C.CDLL(os.path.join(root, "if-dda_amd", "libifdda_dropin.so"))
    ri = lambda v: C.byref(C.c_int(int(v)))  # noqa: E731
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    M = 8 * N
    V = [np.zeros(M, np.complex128) for _ in range(3)]
    Dg = [np.zeros(M, np.complex128) for _ in range(6)]
    FF = np.zeros(3 * N, np.complex128)
    FF[: g["box_FF"].size] = g["box_FF"]
    plan = C.c_int64(0)
    dims = [ri(nx), ri(ny), ri(nz), ri(2 * nx), ri(2 * ny), ri(2 * nz), ri(4 * nx * ny)]
    L.calculforcefft_(*dims, ri(nx), ri(ny), ri(nz), P(tab), *[P(v) for v in V], P(FF), C.byref(plan))
    ci = np.nonzero(tab)[0][np.argsort(tab[tab > 0])]
    cells = np.stack([np.arange(N) % nx, (np.arange(N) // nx) % ny, np.arange(N) // (nx * ny)], 1)
    for t in (1, 2, 3):
        print(f"drop-in derivativefield2 test={t}", flush=True)
        out = np.zeros(3 * N, np.complex128)
        L.derivativefield2_(C.byref(C.c_double(D)), C.byref(C.c_double(K0)), ri(nx), ri(ny), ri(nz),
                            ri(2 * nx), ri(2 * ny), ri(2 * nz), ri(4 * nx * ny), ri(M), ri(M), ri(N),
                            *[P(x) for x in Dg], *[P(v) for v in V], ri(t), P(out), C.byref(plan),
                            C.byref(plan))
        out = out.reshape(N, 3)
        assert relerr(out[ci], g["box_dE"][t - 1]) < 1e-12
        direct = fo.dfield_direct(cells * D, cells[ci] * D, g["box_FF"], K0, t - 1)
        assert relerr(out, direct) < 1e-12
