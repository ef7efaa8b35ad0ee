"""Test helper (not collected): the microscope cases of golden/microscopes.npz
and the independent CPU pipeline that forms their images.

The pipeline solves each incidence with the oracle's inverserig, forms the
pupil field with the numpy restatements of diffractefft2dlens.f (quick lens)
or microsbf.f:629-668 (direct sum), and assembles the images as microsbf.f /
microsdf.f / microssansinc.f do (oracle/microscopy.py).  The fixture holds
the compiled reference's own microsbf_ / microsdf_ / microsbfneg_ /
microsdfneg_ on the same sphere, run on a dense box (every cell a dipole,
alpha = 0 outside the sphere), which is the same linear system as the
sparse map used here and on the device.
"""
from __future__ import annotations

import numpy as np

LAM = 632.8e-9
K0 = 2 * np.pi / LAM

MICRO_CASES = [(kind, na, nainc, nq)
               for kind in ("microsbf", "microsdf", "microsbfneg", "microsdfneg")
               for na, nainc in ((0.8, 0.5), (0.5, 0.8))
               if not (kind.startswith("microsbf") and na < nainc)
               for nq in (1, 0)]
MICRO_NFFT, MICRO_W0, MICRO_GROSS, MICRO_ZLENS, MICRO_TOL = 128, 1e-6, 100.0, 50e-9, 1e-10


def micro_key(kind, na, nainc, nq):
    return f"{kind}_na{round(na * 10)}_inc{round(nainc * 10)}_q{nq}"


def micro_weights(n, k=16, seed=2025):
    """k seeded real weight rows of length n: an image plane I is kept as W @ I."""
    return np.random.default_rng(seed).standard_normal((k, n))


def micro_pixels(n, k=64, seed=2026):
    return np.sort(np.random.default_rng(seed).choice(n, k, replace=False))


def micro_scene():
    """The sphere of microscopes.npz: radius 60 nm on 6^3 cells of 20 nm,
    eps 2.25+0.1i; returns (scene, box-ordered cell centres)."""
    from ifdda import scene
    s = scene.sphere(6, 60e-9, 2.25 + 0.1j)
    c = s.origin[0] + s.d * np.arange(s.nx)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    return s, np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], 1)


def case_geometry(kind, na, nainc, nq):
    """(dark, nside, zlens, direct) of a fixture case."""
    neg = kind.endswith("neg")
    return "df" in kind, -1 if neg else 1, -MICRO_ZLENS if neg else MICRO_ZLENS, nq == 0


def check_images(img, fx, key, rtol):
    """img: six image planes [nfft2d, nfft2d]; compared with the fixture's
    values at its seeded pixels (to rtol of the plane's max), its projections
    (to the bound those pixel errors imply), max and sum."""
    N2 = MICRO_NFFT * MICRO_NFFT
    a = np.asarray(img).reshape(6, N2)
    if np.iscomplexobj(a):
        assert np.abs(a.imag).max() <= rtol * np.abs(a).max()
        a = a.real
    W = micro_weights(N2)
    mx = fx[f"{key}_max"]
    assert (mx > 0).all()
    pix = a[:, fx["pixels"]]
    err = np.abs(pix - fx[f"{key}_pix"]).max(axis=1) / mx
    assert (err < rtol).all(), (key, "pixels", err)
    assert (np.abs(a.max(axis=1) / mx - 1) < rtol).all(), (key, "max")
    assert (np.abs(a.sum(axis=1) - fx[f"{key}_sum"]) < rtol * N2 * mx).all(), (key, "sum")
    bound = rtol * mx[:, None] * np.abs(W).sum(axis=1)[None, :]
    perr = np.abs(a @ W.T - fx[f"{key}_proj"])
    assert (perr < bound).all(), (key, "projections", (perr / bound).max())


def cpu_images(s, alpha, oracle, m, tol, dark=False):
    """The six images of the microscope described by m (its nfft2d,
    numaper, numaperinc, E0, zlens, gross, nside, direct_lens) for the scene
    s with per-dipole polarisability alpha, every incidence solved by the
    oracle's inverserig to tol."""
    import microscopy as mi
    direct = bool(m.direct_lens)
    if dark:
        ring, dkinc, _ = mi.incidences_df(K0, m.numaperinc, s.d, m.nfft2d, s.pos[:, 0], s.pos[:, 1],
                                          direct=direct)
        waves = [(idl, 0, kx, ky, pp) for (idl, pp, kx, ky) in ring]
    else:
        inc, dkinc, _ = mi.incidences(K0, m.numaperinc, s.d, m.nfft2d, s.pos[:, 0], s.pos[:, 1],
                                      direct=direct)
        waves = [(idl, jdl, kx, ky, pp) for pp in (0.0, 1.0) for (idl, jdl, kx, ky) in inc]
    G = oracle.green_fft(s.nx, s.ny, s.nz, (2 * s.nx, 2 * s.ny, 2 * s.nz), s.d, K0)
    pol = np.zeros((s.nx * s.ny * s.nz, 3, 3), np.complex128)
    for a in range(3):
        pol[: s.nd, a, a] = alpha
    E0 = complex(m.E0[0], m.E0[1])
    zs = s.origin[2] + s.d * np.arange(s.nz)
    cells = np.nonzero(s.tabdip)[0]
    lens = []
    for (idl, jdl, kx, ky, pp) in waves:
        Einc = mi.ondeplanekxky(s.pos, K0, E0, pp, kx, ky)
        o = oracle.inverserig(G, s.tabdip, s.nx, s.ny, s.nz, pol, Einc, "GPBICG1", tol)
        p = (alpha[:, None] * o["FFloc"].reshape(-1, 3))
        FFbox = np.zeros((s.nx * s.ny * s.nz, 3), np.complex128)
        FFbox[cells] = p[s.tabdip[cells] - 1]
        if direct:  # nquicklens = 0: the pupil field as the direct sum over dipoles
            imax, dk = mi.direct_geometry(m.nfft2d, K0, s.d)
            El = mi.lens_field_direct(s.pos, p, K0, dk, imax, m.numaper * K0, m.nside)
        else:
            El, imax, dk = mi.diffractefft2dlens(FFbox.reshape(-1), s.nx, s.ny, s.nz, m.nfft2d, K0,
                                                 s.origin[0], s.origin[1], zs, s.d, m.numaper * K0,
                                                 m.nside)
        lens.append(El)
    return mi.image_assembly(lens, waves, K0, E0, m.numaper * K0, m.zlens, m.gross, dkinc, imax, dk,
                             m.nfft2d, dark=dark, nside=m.nside)
