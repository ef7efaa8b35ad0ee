"""Host-side producers of solve inputs: lattices, polarisability, incident waves.

These restate the reference's callers of the hot path (object, beam and
polarisability routines, which stay on the host and call the solve
unchanged) so that benchmarks and tests can build the workloads:

  sphere lattice        objectsphere.f:98-113  (centres -r + d(k - 1/2), |r| < R)
  cube lattice          objectcube.f:95-106    (centres k d - (side + d)/2)
  polarisability        polarisabilite.f:20-44 (CM, RR)
  linear plane wave     ondeplane.f:32-49
  circular plane wave   ondecirce.f:23-52
  box lattice           nx x ny x nz cells of side d, centres per axis as
                        objectcube.f:95-106; sphere / ellipsoid / cylinder
                        are the cells whose centre lies inside
  random permittivity   a seeded Gaussian random field with correlation
                        sigma^2 exp(-r^2/l^2) (the statistics of
                        objectparainhomo.f:150-205; not its realisation)

Units are SI (metres); the reference converts its nm inputs the same way.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Scene:
    nx: int
    ny: int
    nz: int
    d: float                 # lattice spacing [m]
    tabdip: np.ndarray       # int32 [nx*ny*nz], 1-based dipole or 0
    pos: np.ndarray          # float64 [nd, 3] dipole positions [m]
    eps: np.ndarray          # complex128 [nd] permittivity
    origin: tuple = (0.0, 0.0, 0.0)  # centre of box cell (1,1,1) [m]

    @property
    def nd(self) -> int:
        return int(self.pos.shape[0])


def _lattice(centres_x, centres_y, centres_z, inside):
    """Box order x fastest; inside: bool [nz, ny, nx]."""
    nz, ny, nx = inside.shape
    flat = inside.reshape(-1)
    nd = int(flat.sum())
    tab = np.zeros(flat.size, np.int32)
    tab[flat] = np.arange(1, nd + 1, dtype=np.int32)
    Z, Y, X = np.meshgrid(centres_z, centres_y, centres_x, indexing="ij")
    pos = np.stack([X.reshape(-1)[flat], Y.reshape(-1)[flat], Z.reshape(-1)[flat]], axis=1)
    return tab, pos


def sphere(nnnr: int, radius: float, eps: complex) -> Scene:
    d = 2.0 * radius / nnnr
    c = -radius + d * (np.arange(1, nnnr + 1) - 0.5)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    inside = np.sqrt(X * X + Y * Y + Z * Z) < radius
    tab, pos = _lattice(c, c, c, inside)
    return Scene(nnnr, nnnr, nnnr, d, tab, pos, np.full(pos.shape[0], eps, np.complex128),
                 (c[0], c[0], c[0]))


def cube(nnnr: int, side: float, eps) -> Scene:
    d = side / nnnr
    c = np.arange(1, nnnr + 1) * d - (side + d) / 2.0
    inside = np.ones((nnnr, nnnr, nnnr), bool)
    tab, pos = _lattice(c, c, c, inside)
    e = np.broadcast_to(np.asarray(eps, np.complex128), (pos.shape[0],)).copy()
    return Scene(nnnr, nnnr, nnnr, d, tab, pos, e, (c[0], c[0], c[0]))


def gaussian_random_field(shape, d: float, sigma: float, corr: float, seed: int) -> np.ndarray:
    """Real field on a (nz,ny,nx) lattice, covariance sigma^2 exp(-r^2/corr^2).

    White noise filtered in k-space by the square root of the spectrum
    exp(-k^2 corr^2 / 4), then rescaled to the exact sample std sigma."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    ks = [2 * np.pi * np.fft.fftfreq(n, d) for n in shape]
    K2 = ks[0][:, None, None] ** 2 + ks[1][None, :, None] ** 2 + ks[2][None, None, :] ** 2
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * np.exp(-K2 * corr * corr / 8.0)))
    f -= f.mean()
    return f * (sigma / f.std())


def inhomogeneous_cube(nnnr: int, side: float, eps_mean: complex, sigma: float, corr: float,
                       seed: int) -> Scene:
    s = cube(nnnr, side, eps_mean)
    field = gaussian_random_field((nnnr, nnnr, nnnr), s.d, sigma, corr, seed)
    s.eps = (eps_mean + field.reshape(-1)).astype(np.complex128)
    return s


def polarizability(eps, d: float, k0: float, method: str = "RR", eps0: float = 1.0):
    """poladiff with dddis = 1 (polarisabilite.f:20-44): cell of side d."""
    eps = np.asarray(eps, np.complex128)
    a3 = (d * (0.75 / np.pi) ** (1.0 / 3.0)) ** 3
    cm = (eps - eps0) / (eps + 2.0 * eps0) * a3
    if method == "CM":
        return cm * eps0
    if method == "RR":
        k03 = k0 ** 3 * np.sqrt(eps0) ** 3
        return eps0 * cm / (1.0 - 2.0 / 3.0 * 1j * k03 * cm)
    raise ValueError(f"polarizability method {method!r} not provided")


def plane_wave(pos, k0: float, theta_deg: float = 0.0, phi_deg: float = 0.0, pp: float = 1.0,
               E0: complex = 1.0) -> np.ndarray:
    """ondeplane.f:32-49; returns AoS (Ex,Ey,Ez) per dipole, complex128 [3*nd]."""
    th, ph = np.radians(theta_deg), np.radians(phi_deg)
    p, s = np.sqrt(pp), np.sqrt(1.0 - pp)
    sg = np.sign(th) if th != 0 else 1.0
    E0s, E0p = E0 * s * sg, E0 * p * sg
    k = k0 * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    e = np.array([-E0s * np.sin(ph) + E0p * np.cos(th) * np.cos(ph),
                  E0s * np.cos(ph) + E0p * np.cos(th) * np.sin(ph),
                  -E0p * np.sin(th)], np.complex128)
    ph_ = np.exp(1j * (pos @ k))
    return (ph_[:, None] * e[None, :]).reshape(-1)


def circular_wave(pos, k0: float, theta_deg: float = 0.0, phi_deg: float = 0.0,
                  helicity: float = 1.0, E0: complex = 1.0) -> np.ndarray:
    """ondecirce.f:23-52: s part E0/sqrt2 plus p part i sign(helicity) E0/sqrt2."""
    th, ph = np.radians(theta_deg), np.radians(phi_deg)
    E0s = E0 / np.sqrt(2.0)
    E0p = E0 * 1j * np.sign(helicity) / np.sqrt(2.0)
    k = k0 * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    e = np.array([-E0s * np.sin(ph) + E0p * np.cos(th) * np.cos(ph),
                  E0s * np.cos(ph) + E0p * np.cos(th) * np.sin(ph),
                  -E0p * np.sin(th)], np.complex128)
    ph_ = np.exp(1j * (pos @ k))
    return (ph_[:, None] * e[None, :]).reshape(-1)


def config1():
    """Sphere of diameter 400 nm, eps 2.25, 20^3 lattice, lambda 632.8 nm, RR,
    x-polarised plane wave along +z."""
    lam = 632.8e-9
    k0 = 2 * np.pi / lam
    s = sphere(20, 200e-9, 2.25 + 0j)
    alpha = polarizability(s.eps, s.d, k0, "RR")
    E = plane_wave(s.pos, k0, 0.0, 0.0, pp=1.0)
    return s, k0, alpha, E


def config2(n: int = 100, seed: int = 2):
    """n^3 cube of side n*20 nm, eps 2.25+0.01i with a Gaussian random real
    perturbation (std 0.1, correlation 100 nm), lambda 632.8 nm, circular
    plane wave at theta = 30 deg, phi = 0."""
    lam = 632.8e-9
    k0 = 2 * np.pi / lam
    d = 20e-9
    s = inhomogeneous_cube(n, n * d, 2.25 + 0.01j, 0.1, 100e-9, seed)
    alpha = polarizability(s.eps, s.d, k0, "RR")
    E = circular_wave(s.pos, k0, 30.0, 0.0, 1.0)
    return s, k0, alpha, E


def random_spheres(nnnr: int, box: float, nsph: int, radius: float, eps: complex, seed: int) -> Scene:
    """nsph non-overlapping spheres (seeded uniform centres, rejection) on an
    nnnr^3 lattice filling a cube of side `box`; cell centres as
    objectcube.f:95-106, a cell is inside when its centre is in a sphere."""
    rng = np.random.default_rng(seed)
    centres = []
    tries = 0
    while len(centres) < nsph:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place the spheres")
        c = rng.uniform(-box / 2 + radius, box / 2 - radius, 3)
        if all(np.linalg.norm(c - o) >= 2 * radius for o in centres):
            centres.append(c)
    d = box / nnnr
    g = np.arange(1, nnnr + 1) * d - (box + d) / 2.0
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    inside = np.zeros(X.shape, bool)
    for c in centres:
        inside |= (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 < radius ** 2
    tab, pos = _lattice(g, g, g, inside)
    return Scene(nnnr, nnnr, nnnr, d, tab, pos, np.full(pos.shape[0], eps, np.complex128),
                 (g[0], g[0], g[0]))


def pupil_incidences(n_inc: int, na: float, seed: int):
    """(theta, phi) in degrees of n_inc directions uniform over the pupil disk
    |sin theta| <= na (the illumination cone of a brightfield microscope)."""
    rng = np.random.default_rng(seed)
    rr = na * np.sqrt(rng.uniform(0.0, 1.0, n_inc))
    ph = rng.uniform(0.0, 2 * np.pi, n_inc)
    return np.degrees(np.arcsin(rr)), np.degrees(ph)


def config3(n_inc: int = 128, seed: int = 3):
    """20 spheres r = 150 nm, eps 2.0+0.001i, in a 1.28 um box on a 64^3
    lattice, lambda 632.8 nm; n_inc incidences in an NA 0.95 cone x 2 linear
    polarisations (p then s): E_inc is [2*n_inc, 3*nd]."""
    lam = 632.8e-9
    k0 = 2 * np.pi / lam
    s = random_spheres(64, 1.28e-6, 20, 150e-9, 2.0 + 0.001j, seed)
    alpha = polarizability(s.eps, s.d, k0, "RR")
    th, ph = pupil_incidences(n_inc, 0.95, seed + 1)
    E = np.stack([plane_wave(s.pos, k0, t, p, pp=pp) for pp in (1.0, 0.0) for t, p in zip(th, ph)])
    return s, k0, alpha, E


def _box_centres(n: int, d: float) -> np.ndarray:
    return np.arange(1, n + 1) * d - (n * d + d) / 2.0


def shape_box(kind: str, nx: int, ny: int, nz: int, d: float, eps: complex) -> Scene:
    """kind: 'sphere' (diameter = smallest side), 'cube' (every cell),
    'ellipsoid' (semi-axes = half sides), 'cylinder' (axis z, diameter =
    smaller of the x, y sides, full height)."""
    gx, gy, gz = _box_centres(nx, d), _box_centres(ny, d), _box_centres(nz, d)
    Z, Y, X = np.meshgrid(gz, gy, gx, indexing="ij")
    a, b, c = nx * d / 2, ny * d / 2, nz * d / 2
    if kind == "sphere":
        r = min(a, b, c)
        inside = X * X + Y * Y + Z * Z < r * r
    elif kind == "cube":
        inside = np.ones(X.shape, bool)
    elif kind == "ellipsoid":
        inside = (X / a) ** 2 + (Y / b) ** 2 + (Z / c) ** 2 < 1.0
    elif kind == "cylinder":
        r = min(a, b)
        inside = X * X + Y * Y < r * r
    else:
        raise ValueError(f"unknown shape {kind!r}")
    tab, pos = _lattice(gx, gy, gz, inside)
    return Scene(nx, ny, nz, d, tab, pos, np.full(pos.shape[0], eps, np.complex128),
                 (gx[0], gy[0], gz[0]))


def config4(count: int = 512, seed: int = 4):
    """Uneven batch: per solve a shape from {sphere, cube, ellipsoid,
    cylinder}, nx, ny, nz uniform in [8, 48], Re eps in [1.5, 4], Im eps in
    [0, 0.5], lambda in [400, 800] nm, d = lambda / (10 |n|), a plane wave of
    random direction (uniform on the sphere) and polarisation (pp in [0, 1]).
    Returns a list of (Scene, k0, alpha, E_inc)."""
    rng = np.random.default_rng(seed)
    kinds = ["sphere", "cube", "ellipsoid", "cylinder"]
    out = []
    while len(out) < count:
        kind = kinds[rng.integers(4)]
        nx, ny, nz = (int(v) for v in rng.integers(8, 49, 3))
        eps = complex(rng.uniform(1.5, 4.0), rng.uniform(0.0, 0.5))
        lam = rng.uniform(400e-9, 800e-9)
        d = lam / (10.0 * abs(np.sqrt(eps)))
        k0 = 2 * np.pi / lam
        th = np.degrees(np.arccos(rng.uniform(-1.0, 1.0)))
        ph = rng.uniform(0.0, 360.0)
        pp = rng.uniform(0.0, 1.0)
        s = shape_box(kind, nx, ny, nz, d, eps)
        if s.nd == 0:
            continue
        alpha = polarizability(s.eps, d, k0, "RR")
        out.append((s, k0, alpha, plane_wave(s.pos, k0, th, ph, pp=pp)))
    return out
