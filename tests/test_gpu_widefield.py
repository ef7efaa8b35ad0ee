"""The wide-field near field of cdmlib (cdmlib.f:1967-2303): after a solve,
cdmlib builds the Green tensor of a box enlarged by nxmp, nymp, nzmp cells
on each side, puts -FF at the object's cells (0 elsewhere) and E_inc at
every cell, and calls produitfftmatvectopt on the dense enlarged box, which
gives E = E_inc + T.FF everywhere.  Both calls are drop-ins here.  Checks:
at the object's cells the result reproduces the solved local field (the
solve's own equation), and everywhere it equals the direct sum."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

LAM = 632.8e-9
K0 = 2 * np.pi / LAM


def test_wide_field_on_the_enlarged_box(gpu, oracle):
    from ifdda import scene
    s = scene.sphere(8, 80e-9, 2.25 + 0.1j)
    alpha = scene.polarizability(s.eps, s.d, K0, "RR")
    E = scene.plane_wave(s.pos, K0, 20.0, 10.0, pp=1.0)
    with gpu.Solver(s.nx, s.ny, s.nz) as sv:
        sv.greencalculfft(s.d, K0)
        sv.set_dipoles(s.tabdip)
        sv.set_polarizability(alpha)
        out = sv.inverserig(E, "GPBICG1", 1e-12, precision=64, x0_zero=True)
    assert out["nstop"] == 0 and out["tol1"] <= 1e-12
    m = (2, 3, 1)  # nxmp, nymp, nzmp
    nxp, nyp, nzp = s.nx + 2 * m[0], s.ny + 2 * m[1], s.nz + 2 * m[2]
    # cell centres of the enlarged box (cdmlib.f: x = xmin + (k - nxmp - 1) d)
    ox, oy, oz = (s.origin[a] - m[a] * s.d for a in range(3))
    k, j, i = np.meshgrid(np.arange(nzp), np.arange(nyp), np.arange(nxp), indexing="ij")
    pos = np.stack([ox + i.ravel() * s.d, oy + j.ravel() * s.d, oz + k.ravel() * s.d], 1)
    xr = scene.plane_wave(pos, K0, 20.0, 10.0, pp=1.0).reshape(-1, 3)
    xi = np.zeros((nxp * nyp * nzp, 3), np.complex128)
    # object cell (a, b, c) of the small box sits at (a + nxmp, b + nymp, c + nzmp)
    cells = np.nonzero(s.tabdip)[0]
    a, b, c = cells % s.nx, (cells // s.nx) % s.ny, cells // (s.nx * s.ny)
    big = (a + m[0]) + nxp * ((b + m[1]) + nyp * (c + m[2]))
    dip = s.tabdip[cells] - 1
    FF = out["FF"].reshape(-1, 3)
    xi[big] = -FF[dip]
    xr[big] = E.reshape(-1, 3)[dip]  # inside the object cdmlib takes FF0 = E_inc
    with gpu.Solver(nxp, nyp, nzp) as wide:
        print(f"wide field {nxp}x{nyp}x{nzp}", flush=True)
        wide.greencalculfft(s.d, K0)
        wide.set_dipoles(None)  # the dense box of produitfftmatvectopt
        Ew = wide.produitfftmatvect3(xi.reshape(-1), xr.reshape(-1)).reshape(-1, 3)
    # the solved equation: E_loc = E_inc + T.FF at the dipoles
    Eloc = out["FFloc"].reshape(-1, 3)
    assert relerr(Ew[big], Eloc[dip]) < 1e-10, relerr(Ew[big], Eloc[dip])
    # everywhere: the direct sum over the object's dipoles
    ref = oracle.matvec_direct(None, nxp, nyp, nzp, s.d, K0, xi.reshape(-1), xr.reshape(-1))
    assert relerr(Ew.reshape(-1), ref) < 1e-12, relerr(Ew.reshape(-1), ref)
