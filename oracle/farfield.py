"""TEST INFRASTRUCTURE ONLY: numpy restatement of the outputs computed from a
solved field, as cdmlib does them:

  cross_sections   Cext, Cabs, Csca            cdmlib.f:2725-2753
  diffractefft2d   far field on the k-plane    diffractefft2d.f:1-183
                   (quick method, 2D FFT per z-plane)
  computegcfft2d   Csca and g from that field  computegcfft2d.f:1-54

Only tests/ and bench.py's checks import this module.
"""
from __future__ import annotations

import numpy as np


def nint(x):
    """Fortran NINT / IDNINT: nearest integer, halves away from zero."""
    return int(np.copysign(np.floor(abs(x) + 0.5), x))


def cross_sections(FF0, FF, FFloc, k0, I0):
    """(Cext, Cabs, Csca): cdmlib.f:2730-2753.  FF0 incident field, FF dipole
    moments, FFloc local field, AoS per dipole; I0 = |E0|^2."""
    e0 = np.asarray(FF0).reshape(-1, 3)
    p = np.asarray(FF).reshape(-1, 3)
    el = np.asarray(FFloc).reshape(-1, 3)
    k03 = k0 ** 3
    cext = np.sum(np.imag(np.conj(e0) * p))
    cabs = np.sum(np.imag(p * np.conj(el))) - 2.0 / 3.0 * k03 * np.sum(np.real(p * np.conj(p)))
    Cext = 4.0 * np.pi * k0 * cext / I0
    Cabs = 4.0 * np.pi * k0 * cabs / I0
    Csca = Cext - Cabs
    if Cabs <= Cext * 1e-12:
        Cabs = 0.0
    return Cext, Cabs, Csca


def tabfft2(nfft2d):
    """diffractefft2d.f:49-55 (1-based values)."""
    h = nfft2d // 2
    return np.array([i - h if i - h - 1 >= 0 else h + i for i in range(1, nfft2d + 1)])


def diffractefft2d(FFbox, nx, ny, nz, nfft2d, k0, xs1, ys1, zs, d):
    """diffractefft2d.f:1-183.  FFbox: dipole moments in box order (cell
    i + nx j + nx ny k, three components each, zeros for empty cells); zs[k]
    the z of plane k.  Returns (Epos, Eneg, imax, deltak) with E[ii, jj, c]
    for ii, jj = 0..2 imax (kx = (ii - imax) deltak), or raises ValueError
    with the reference's message."""
    N = nfft2d
    h = N // 2
    dk = 2.0 * np.pi / (N * d)
    var1 = (xs1 + h * d) * dk
    var2 = (ys1 + h * d) * dk
    if N > 16384:
        raise ValueError("nfft2d too large")
    if dk >= k0:
        raise ValueError("In FFT Poynting nfft2d too small")
    tab = tabfft2(N)
    imax = nint((k0 / dk)) + 1
    if 2 * imax + 1 > N or 2 * imax + 1 < 7:
        raise ValueError("In FFT diffract nfft2d too small")
    F = np.asarray(FFbox).reshape(nz, ny, nx, 3)
    n = 2 * imax + 1
    Epos = np.zeros((n, n, 3), np.complex128)
    Eneg = np.zeros((n, n, 3), np.complex128)
    iv = np.arange(-imax, imax + 1)
    KX, KY = np.meshgrid(iv * dk, iv * dk, indexing="ij")  # [ii, jj]
    disk = k0 * k0 - KX * KX - KY * KY > 0.0
    KZ = np.sqrt(np.where(disk, k0 * k0 - KX * KX - KY * KY, 0.0))
    bins = tab[iv + h] - 1  # 0-based FFT bin of frequency index i (tab(i + nfft2d2 + 1))
    vph = np.exp(-1j * (var1 * iv[:, None] + var2 * iv[None, :]))
    xpos = tab[:nx] - 1
    ypos = tab[:ny] - 1
    for k in range(nz):
        P = np.zeros((3, N, N), np.complex128)  # [c][y][x], x fastest
        P[:, ypos[:, None], xpos[None, :]] = np.moveaxis(F[k], 2, 0)
        P = np.fft.fft2(P, axes=(1, 2))  # fftsingletonz2d with ISIGN = -1
        E = P[:, bins[None, :], bins[:, None]]  # [c][ii][jj]: x bin of ii, y bin of jj
        Ex, Ey, Ez = E
        for sgn, out in ((1.0, Epos), (-1.0, Eneg)):
            ctmp = (KX * Ex + KY * Ey + sgn * KZ * Ez) / k0
            ph = np.exp(-1j * sgn * KZ * zs[k]) * vph
            out[..., 0] += np.where(disk, (Ex - KX * ctmp / k0) * ph, 0.0)
            out[..., 1] += np.where(disk, (Ey - KY * ctmp / k0) * ph, 0.0)
            out[..., 2] += np.where(disk, (Ez - sgn * KZ * ctmp / k0) * ph, 0.0)
    return Epos, Eneg, imax, dk


def computegcfft2d(Epos, Eneg, imax, dk, k0, I0, theta, phi):
    """computegcfft2d.f:1-54: (Csca, g); theta, phi of the incidence in degrees."""
    iv = np.arange(-imax, imax + 1)
    KX, KY = np.meshgrid(iv * dk, iv * dk, indexing="ij")
    disk = k0 * k0 - KX * KX - KY * KY > 0.0
    KZ = np.sqrt(np.where(disk, k0 * k0 - KX * KX - KY * KY, 1.0))
    th, ph = np.radians(theta), np.radians(phi)
    ux, uy, uz = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)
    csca = 0.0
    g = 0.0
    for E, s in ((Epos, 1.0), (Eneg, -1.0)):
        emod = np.sum(np.abs(E) ** 2, axis=2)
        w = np.where(disk, dk * dk * emod / KZ, 0.0)
        csca += w.sum()
        g += (w * (KX * ux + KY * uy + s * KZ * uz) / k0).sum()
    g /= csca
    return csca * k0 ** 3 / I0, g
