"""Write golden vectors from the compiled reference (oracle/_ref/libcdmref.so).

Run in a container where the reference sources were available when
``make -C oracle ref`` ran:

    python tests/golden/make_golden.py

Fixtures (inputs and the reference's outputs, nothing else):
  green_654.npz    greencalculfft_ / greencalculquadfft_ (nquad 0,1,2) on a
                   6x5x4 box, d = 20 nm, lambda = 632.8 nm
  matvec_654.npz   produitfftmatvect3_ (seeded 60% mask) and
                   produitfftmatvectopt_ on that box, seeded FFX / FFB
  green_91012.npz  greencalculfft_ and produitfftmatvect3_ on a 9x10x12 box
                   (FFT lengths 18, 20, 24: radices 2, 3, 5 mixed)
  solve_cfg1.npz   inverserig_ on the 400 nm sphere (20^3 lattice, 4224
                   dipoles, eps 2.25, RR, x-polarised plane wave), tol 1e-4,
                   for GPBICG1, QMRBICGSTAB1, QMRBICGSTAB2, TFQMR: E (FFloc),
                   tol1, nloop, ncompte
  solve_more.npz   the same for CG (zcg.f) and BICGSTAB (pimzbicgstab.f), at
                   tol 1e-4 and 1e-6 (written by `make_golden.py more`, which
                   leaves the other fixtures as they are)
  solve_gp.npz     the same for GPBICG2, GPBICGsafe, GPBICGplus, GPBICGAR,
                   BICGSTARPLUS, GPBICOR, CORS and QMRCLA at tol 1e-4 and 1e-6, the
                   fields as 32 projections on seeded complex weights
                   (projection_weights) to keep the file small
                   (`make_golden.py gp`; run with OMP_NUM_THREADS=1: the
                   reference's GPBICOR gives run-to-run differences of ~1e-5
                   on several OpenMP threads, deterministic on one)
  force.npz        optical force and torque: calculforcefft_ + derivativefield2_
                   (d/dx, d/dy, d/dz of the scattered field) on a seeded 6x5x4
                   box with seeded moments, incident and local fields, and on
                   config 1 solved by inverserig_ (GPBICG1, tol 1e-8); the
                   force and torque of cdmlib.f:3102-3560 restated in
                   oracle/force.py on those derivatives (`make_golden.py force`)
  microscope.npz   diffractefft2dlens_ (nside +1 and -1, two NA) on the seeded
                   6x5x4 box and fouriertoimage_ on seeded 32 x 32 planes
  solve_opt.npz    inverserigopt_ (the dense-box driver, every box cell a
                   dipole, alpha = 0 outside a sphere) on a 12 x 10 x 9 box for
                   GPBICG1, QMRBICGSTAB1 and TFQMR at tol 1e-4 and 1e-6: E
                   (FFloc), FF, tol1, nloop, ncompte (`make_golden.py opt`)
  microscopes.npz  the four whole microscopes microsbf_, microsdf_,
                   microsbfneg_, microsdfneg_ (quick lens nquicklens = 1 and
                   direct sum nquicklens = 0; darkfield at NA 0.8 / condenser
                   0.5 and 0.5 / 0.8) on a 120 nm sphere (6^3 cells,
                   eps 2.25+0.1i, RR) at nfft2d 128, GPBICG1 tol 1e-10: per case
                   the number of illuminations, the E0 they carry, and each
                   image plane as 16 projections on seeded real weights
                   (micro_weights), its values at 64 seeded pixels, its max
                   and its sum (`make_golden.py micro`)
  farfield.npz     diffractefft2d_ and computegcfft2d_: a seeded box-ordered
                   FF on 6x5x4 at nfft2d 64 and 65 (odd), and the config-1
                   GPBICG1 moments (alpha E) at nfft2d 128; Epos / Eneg cut to
                   the (2 imax + 1)^2 corner, imax, deltak, Csca, g
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "if-dda_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import Ref  # noqa: E402
from ifdda import scene  # noqa: E402

LAM = 632.8e-9
K0 = 2 * np.pi / LAM
D = 20e-9


def seeded_case(nx, ny, nz, seed, frac=0.6):
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    mask = rng.random(N) < frac
    nd = int(mask.sum())
    tab = np.zeros(N, np.int32)
    tab[mask] = np.arange(1, nd + 1, dtype=np.int32)
    ffx = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    ffb = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    ffx_d = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    ffb_d = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    return tab, ffx, ffb, ffx_d, ffb_d


def main():
    r = Ref()
    nx, ny, nz = 6, 5, 4
    G0 = r.green_fft(nx, ny, nz, D, K0)
    G1 = r.green_fft(nx, ny, nz, D, K0, nquad=1)
    G2 = r.green_fft(nx, ny, nz, D, K0, nquad=2)
    np.savez_compressed(os.path.join(HERE, "green_654.npz"), nx=nx, ny=ny, nz=nz, d=D, k0=K0,
                        G_nquad0=G0, G_nquad1=G1, G_nquad2=G2)
    tab, ffx, ffb, ffx_d, ffb_d = seeded_case(nx, ny, nz, 654)
    out3 = r.matvec3(G0, tab, nx, ny, nz, ffx, ffb)
    outo = r.matvec_opt(G0, nx, ny, nz, ffx_d, ffb_d)
    np.savez_compressed(os.path.join(HERE, "matvec_654.npz"), nx=nx, ny=ny, nz=nz, d=D, k0=K0,
                        tabdip=tab, FFX=ffx, FFB=ffb, out3=out3, FFX_dense=ffx_d, FFB_dense=ffb_d,
                        out_opt=outo)

    nx, ny, nz = 9, 10, 12
    G = r.green_fft(nx, ny, nz, D, K0)
    tab, ffx, ffb, _, _ = seeded_case(nx, ny, nz, 91012, frac=0.5)
    out3 = r.matvec3(G, tab, nx, ny, nz, ffx, ffb)
    np.savez_compressed(os.path.join(HERE, "green_91012.npz"), nx=nx, ny=ny, nz=nz, d=D, k0=K0,
                        G=G, tabdip=tab, FFX=ffx, FFB=ffb, out3=out3)

    s, k0, alpha, E = scene.config1()
    Gs = r.green_fft(s.nx, s.ny, s.nz, s.d, k0)
    pol = np.zeros((s.nx * s.ny * s.nz, 3, 3), np.complex128)
    for a in range(3):
        pol[: s.nd, a, a] = alpha
    res = {}
    for m in ["GPBICG1", "QMRBICGSTAB1", "QMRBICGSTAB2", "TFQMR"]:
        o = r.inverserig(Gs, s.tabdip, s.nx, s.ny, s.nz, pol, E, m, 1e-4)
        assert o["nstop"] == 0, o
        res[f"{m}_E"] = o["FFloc"]
        res[f"{m}_tol1"] = o["tol1"]
        res[f"{m}_nloop"] = o["nloop"]
        res[f"{m}_ncompte"] = o["ncompte"]
        print(m, o["nloop"], o["ncompte"], o["tol1"])
    np.savez_compressed(os.path.join(HERE, "solve_cfg1.npz"), tabdip=s.tabdip, d=s.d, k0=k0,
                        alpha=alpha, E_inc=E, **res)
    farfield(r, s, k0, alpha, res["GPBICG1_E"])
    microscope(r)


def microscope(r):
    out = {}
    rng = np.random.default_rng(12)
    nx, ny, nz = 6, 5, 4
    F = rng.standard_normal(3 * nx * ny * nz) + 1j * rng.standard_normal(3 * nx * ny * nz)
    zs = -30e-9 + D * np.arange(nz)
    out.update(box_FF=F, box_zs=zs)
    for tag, N, na, side in (("a", 64, 0.9, 1), ("b", 96, 0.5, -1)):
        E, imax, dk, nstop, info = r.diffractefft2dlens(F, nx, ny, nz, N, K0, -50e-9, -40e-9, zs, D,
                                                        na * K0, side)
        assert nstop == 0, info
        out.update({f"lens_{tag}": E, f"lens_{tag}_imax": imax, f"lens_{tag}_dk": dk})
    planes = [rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32)) for _ in range(6)]
    img = r.fouriertoimage(3e5, 100.0, planes)
    out.update(f2i_in=np.stack(planes), f2i_out=np.stack(img))
    np.savez_compressed(os.path.join(HERE, "microscope.npz"), **out)


def farfield(r, s, k0, alpha, Eloc):
    out = {}
    rng = np.random.default_rng(11)
    nx, ny, nz = 6, 5, 4
    F = rng.standard_normal(3 * nx * ny * nz) + 1j * rng.standard_normal(3 * nx * ny * nz)
    zs = -30e-9 + D * np.arange(nz)
    out.update(box_FF=F, box_zs=zs, box_x0=-50e-9, box_y0=-40e-9)
    for N in (64, 65):
        Ep, En, imax, dk, nstop, info = r.diffractefft2d(F, nx, ny, nz, N, K0, -50e-9, -40e-9, zs, D)
        assert nstop == 0, info
        cs, g = r.computegcfft2d(Ep, En, imax, dk, K0, 1.0, 20.0, 30.0, N)
        out.update({f"box{N}_Epos": Ep, f"box{N}_Eneg": En, f"box{N}_imax": imax, f"box{N}_dk": dk,
                    f"box{N}_csca": cs, f"box{N}_g": g})
    # config 1: moments alpha E in box order (zeros outside the sphere)
    FFbox = np.zeros((s.nx * s.ny * s.nz, 3), np.complex128)
    cells = np.nonzero(s.tabdip)[0]
    FFbox[cells] = (alpha[:, None] * Eloc.reshape(-1, 3))[s.tabdip[cells] - 1]
    zs1 = s.origin[2] + s.d * np.arange(s.nz)
    N = 128
    Ep, En, imax, dk, nstop, info = r.diffractefft2d(FFbox.reshape(-1), s.nx, s.ny, s.nz, N, k0,
                                                     s.origin[0], s.origin[1], zs1, s.d)
    assert nstop == 0, info
    cs, g = r.computegcfft2d(Ep, En, imax, dk, k0, 1.0, 0.0, 0.0, N)
    print("config 1 far field: Csca", cs, "g", g, "imax", imax)
    out.update(cfg1_Epos=Ep, cfg1_Eneg=En, cfg1_imax=imax, cfg1_dk=dk, cfg1_csca=cs, cfg1_g=g,
               cfg1_nfft2d=N)
    np.savez_compressed(os.path.join(HERE, "farfield.npz"), **out)


def projection_weights(n, k=32, seed=2024):
    """k seeded complex weight rows of length n: a field E is kept as W @ E."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))


def solve_more(r, methods=("CG", "BICGSTAB"), name="solve_more.npz", project=False):
    s, k0, alpha, E = scene.config1()
    Gs = r.green_fft(s.nx, s.ny, s.nz, s.d, k0)
    pol = np.zeros((s.nx * s.ny * s.nz, 3, 3), np.complex128)
    for a in range(3):
        pol[: s.nd, a, a] = alpha
    res = {}
    for m in methods:
        for tag, tol in (("4", 1e-4), ("6", 1e-6)):
            o = r.inverserig(Gs, s.tabdip, s.nx, s.ny, s.nz, pol, E, m, tol)
            assert o["nstop"] == 0, o
            key = f"{m}_{tag}"
            if project:
                res[f"{key}_Eproj"] = projection_weights(E.size) @ o["FFloc"]
            else:
                res[f"{key}_E"] = o["FFloc"]
                res[f"{key}_FF"] = o["FF"]
            res[f"{key}_tol1"] = o["tol1"]
            res[f"{key}_nloop"] = o["nloop"]
            res[f"{key}_ncompte"] = o["ncompte"]
            print(key, o["nloop"], o["ncompte"], o["tol1"])
    if project:  # the inputs are those of solve_more.npz; keep a checksum of them
        res.update(E_inc_proj=projection_weights(E.size) @ E, tabdip_sum=int(s.tabdip.sum()))
        np.savez_compressed(os.path.join(HERE, name), **res)
    else:
        np.savez_compressed(os.path.join(HERE, name), tabdip=s.tabdip, E_inc=E, **res)


def force(r):
    import force as fo
    out = {}
    nx, ny, nz = 6, 5, 4
    tab, ffx, ffb, _, _ = seeded_case(nx, ny, nz, 21)
    nd = int(tab.max())
    rng = np.random.default_rng(22)
    Einc = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    th, ph = np.radians(20.0), np.radians(40.0)
    kvec = K0 * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    cellidx = np.nonzero(tab)[0][np.argsort(tab[tab > 0])]
    cells = np.stack([cellidx % nx, (cellidx // nx) % ny, cellidx // (nx * ny)], 1)
    dEs = [r.dfield(ffx, tab, nx, ny, nz, D, K0, t + 1).reshape(-1, 3)[cellidx] for t in range(3)]
    fdip, tdip, F, tau = fo.force_torque(dEs, Einc, ffx, ffb, kvec, K0, cells, D)
    out.update(box_tab=tab, box_FF=ffx, box_FFloc=ffb, box_Einc=Einc, box_kvec=kvec,
               box_dE=np.stack(dEs), box_fdip=fdip, box_tdip=tdip, box_F=F, box_tau=tau)
    # config 1, solved by the reference
    s, k0, alpha, E = scene.config1()
    Gs = r.green_fft(s.nx, s.ny, s.nz, s.d, k0)
    pol = np.zeros((s.nx * s.ny * s.nz, 3, 3), np.complex128)
    for a in range(3):
        pol[: s.nd, a, a] = alpha
    o = r.inverserig(Gs, s.tabdip, s.nx, s.ny, s.nz, pol, E, "GPBICG1", 1e-8)
    assert o["nstop"] == 0, o
    ci = np.nonzero(s.tabdip)[0][np.argsort(s.tabdip[s.tabdip > 0])]
    cells = np.stack([ci % s.nx, (ci // s.nx) % s.ny, ci // (s.nx * s.ny)], 1)
    dEs = [r.dfield(o["FF"], s.tabdip, s.nx, s.ny, s.nz, s.d, k0, t + 1).reshape(-1, 3)[ci]
           for t in range(3)]
    fdip, tdip, F, tau = fo.force_torque(dEs, E, o["FF"], o["FFloc"], np.array([0.0, 0.0, k0]), k0,
                                         cells, s.d)
    print("config 1 force", F, "torque", tau, "nloop", o["nloop"])
    out.update(cfg1_F=F, cfg1_tau=tau, cfg1_fz_dip_sum_abs=np.abs(fdip[:, 2]).sum(),
               cfg1_nloop=o["nloop"], cfg1_tol1=o["tol1"])
    np.savez_compressed(os.path.join(HERE, "force.npz"), **out)


def microscopes(r):
    from micro_pipeline import (MICRO_CASES, MICRO_GROSS, MICRO_NFFT, MICRO_TOL, MICRO_W0,
                                MICRO_ZLENS, micro_key, micro_pixels, micro_scene, micro_weights)
    s, box = micro_scene()
    alpha = r.poladiff(s.d, 2.25 + 0.1j, K0, "RR")
    nmax = s.nx * s.ny * s.nz
    pol = np.zeros((nmax, 3, 3), np.complex128)
    for a in range(3):
        pol[s.tabdip > 0, a, a] = alpha
    G = r.green_fft(s.nx, s.ny, s.nz, s.d, K0)
    N2 = MICRO_NFFT * MICRO_NFFT
    W = micro_weights(N2)
    pix = micro_pixels(N2)
    out = dict(alpha=alpha, pixels=pix, nfft2d=MICRO_NFFT, w0=MICRO_W0, gross=MICRO_GROSS,
               zlens=MICRO_ZLENS, tol=MICRO_TOL, k0=K0)
    for kind, na, nainc, nq in MICRO_CASES:
        zl = -MICRO_ZLENS if kind.endswith("neg") else MICRO_ZLENS
        o = r.microscope(kind, G, box, pol, s.d, K0, MICRO_NFFT, na, nainc, 1.0, MICRO_W0,
                         MICRO_GROSS, zl, nq, "GPBICG1", MICRO_TOL)
        assert o["nstop"] == 0, o["infostr"]
        img = o["images"].reshape(6, N2)
        assert np.abs(img.imag).max() == 0.0
        img = img.real
        E0, _ = r.irradiance(o["P0"], MICRO_W0)
        k = micro_key(kind, na, nainc, nq)
        out.update({f"{k}_nillum": int(round(1.0 / o["P0"])), f"{k}_E0": E0,
                    f"{k}_proj": img @ W.T, f"{k}_pix": img[:, pix], f"{k}_max": img.max(axis=1),
                    f"{k}_sum": img.sum(axis=1)})
        print(k, "illuminations", out[f"{k}_nillum"], "max", img.max(axis=1))
    np.savez_compressed(os.path.join(HERE, "microscopes.npz"), **out)


def opt_case():
    """The dense-box problem of solve_opt.npz: a sphere of radius 100 nm
    (eps 4+0.5i, RR) centred in a 12 x 10 x 9 box of 20 nm cells; every box
    cell is a dipole (inverserigopt.f), alpha = 0 outside the sphere; a plane
    wave at theta 20, phi 35, pp 0.4 over every cell."""
    nx, ny, nz = 12, 10, 9
    d = D
    cx = (np.arange(nx) - (nx - 1) / 2) * d
    cy = (np.arange(ny) - (ny - 1) / 2) * d
    cz = (np.arange(nz) - (nz - 1) / 2) * d
    Z, Y, X = np.meshgrid(cz, cy, cx, indexing="ij")
    pos = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], 1)
    inside = np.linalg.norm(pos, axis=1) < 100e-9
    alpha = np.where(inside, scene.polarizability(np.full(pos.shape[0], 4.0 + 0.5j), d, K0, "RR"), 0)
    E = scene.plane_wave(pos, K0, 20.0, 35.0, pp=0.4)
    return nx, ny, nz, d, alpha.astype(np.complex128), E


def solve_opt(r):
    nx, ny, nz, d, alpha, E = opt_case()
    G = r.green_fft(nx, ny, nz, d, K0)
    pol = np.zeros((nx * ny * nz, 3, 3), np.complex128)
    for a in range(3):
        pol[:, a, a] = alpha
    res = {}
    for m in ["GPBICG1", "QMRBICGSTAB1", "TFQMR"]:
        for tag, tol in (("4", 1e-4), ("6", 1e-6)):
            o = r.inverserigopt(G, nx, ny, nz, pol, E, m, tol)
            assert o["nstop"] == 0, o
            k = f"{m}_{tag}"
            res.update({f"{k}_E": o["FFloc"], f"{k}_FF": o["FF"], f"{k}_tol1": o["tol1"],
                        f"{k}_nloop": o["nloop"], f"{k}_ncompte": o["ncompte"]})
            print("inverserigopt", k, o["nloop"], o["ncompte"], o["tol1"])
    np.savez_compressed(os.path.join(HERE, "solve_opt.npz"), nx=nx, ny=ny, nz=nz, d=d, k0=K0,
                        alpha=alpha, E_inc=E, **res)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "opt":
        solve_opt(Ref())
    elif len(sys.argv) > 1 and sys.argv[1] == "more":
        solve_more(Ref())
    elif len(sys.argv) > 1 and sys.argv[1] == "micro":
        microscopes(Ref())
    elif len(sys.argv) > 1 and sys.argv[1] == "force":
        force(Ref())
    elif len(sys.argv) > 1 and sys.argv[1] == "gp":
        solve_more(Ref(), ("GPBICG2", "GPBICGsafe", "GPBICGplus", "GPBICGAR", "BICGSTARPLUS",
                           "GPBICOR", "CORS", "QMRCLA"), "solve_gp.npz", project=True)
    else:
        main()
