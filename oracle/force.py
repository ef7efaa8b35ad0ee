"""TEST INFRASTRUCTURE ONLY: numpy restatement of the optical force and
torque of cdmlib (the checker for ifdda_force_device; imported by tests/ and
tests/golden/make_golden.py only).

  dT(r, k0, dd)       d T / d r_dd of the free-space tensor at r = observation
                      minus source (espace_libre_derive_mat.f, propesplibdermat,
                      restated in Cartesian form)
  dfield_direct(...)  sum_j dT(r_i - r_j) p_j, the convolution of
                      derivativefield2.f:22-110 as a direct sum
  force_torque(...)   the per-dipole force and torque and their totals,
                      cdmlib.f:3102-3560
"""
from __future__ import annotations

import numpy as np

C_LIGHT = 299792458.0                       # cdmlib.f:478
QUATPIEPS0 = 1.0 / (C_LIGHT * C_LIGHT * 1e-7)  # cdmlib.f:480


def dT(r, k0, dd):
    """3x3 complex d T / d r_dd at r (zero at r = 0).  T = A I + B r r^T with
    A = e^{ikR}(k^2/R - 1/R^3 + ik/R^2), B = e^{ikR}(3/R^5 - 3ik/R^4 - k^2/R^3)."""
    r = np.asarray(r, np.float64)
    R = float(np.linalg.norm(r))
    if R == 0.0:
        return np.zeros((3, 3), np.complex128)
    e = np.exp(1j * k0 * R)
    B = e * (3 / R**5 - 3j * k0 / R**4 - k0 * k0 / R**3)
    Ap = e * (1j * k0**3 / R - 2 * k0 * k0 / R**2 - 3j * k0 / R**3 + 3 / R**4)
    Bp = e * (-15 / R**6 + 15j * k0 / R**5 + 6 * k0 * k0 / R**4 - 1j * k0**3 / R**3)
    I = np.eye(3)
    ed = I[dd]
    return (I * Ap * r[dd] / R + (np.outer(ed, r) + np.outer(r, ed)) * B
            + np.outer(r, r) * Bp * r[dd] / R)


def dfield_direct(pos_obs, pos_src, FF, k0, dd):
    """sum_j dT(pos_obs_i - pos_src_j) p_j (O(n m); small cases only)."""
    p = np.asarray(FF).reshape(-1, 3)
    out = np.zeros((len(pos_obs), 3), np.complex128)
    for i, ro in enumerate(pos_obs):
        for j, rs in enumerate(pos_src):
            out[i] += dT(ro - rs, k0, dd) @ p[j]
    return out


def force_torque(dEs, Einc, FF, FFloc, kvec, k0, cells, d):
    """dEs: 3 arrays (nd, 3) of the scattered field derivative at the dipoles
    (derivativefield2.f output); Einc, FF (= alpha E), FFloc (= E): 3 nd
    complex; kvec: the incident wave vector (the incident derivative is
    i k_d E_inc, ondeplaned.f:43-58); cells: (nd, 3) integer cell indices.
    Returns fdip (nd, 3), tdip (nd, 3), F (3), tau (3) (cdmlib.f:3341-3370
    for the force with 4 pi eps0 = 1 / (c^2 1e-7), :3502-3545 for the torque
    about the mean dipole position)."""
    p = np.asarray(FF).reshape(-1, 3)
    E = np.asarray(Einc).reshape(-1, 3)
    L = np.asarray(FFloc).reshape(-1, 3)
    pc = np.conj(p)
    fdip = np.zeros((len(p), 3))
    for dd in range(3):
        der = 1j * kvec[dd] * E + dEs[dd]
        fdip[:, dd] = 0.5 * np.real(np.sum(pc * der, axis=1)) * QUATPIEPS0
    Em = 2.0 / 3.0 * 1j * k0**3 * p + L
    tloc = 0.5 * np.real(np.stack([pc[:, 1] * Em[:, 2] - pc[:, 2] * Em[:, 1],
                                   -pc[:, 0] * Em[:, 2] + pc[:, 2] * Em[:, 0],
                                   pc[:, 0] * Em[:, 1] - pc[:, 1] * Em[:, 0]], 1)) * QUATPIEPS0
    rc = (np.asarray(cells, np.float64) - np.mean(np.asarray(cells, np.float64), axis=0)) * d
    tdip = tloc + np.cross(rc, fdip)
    return fdip, tdip, fdip.sum(axis=0), tdip.sum(axis=0)
