"""TEST INFRASTRUCTURE ONLY: Mie series for a homogeneous sphere.

The Bohren & Huffman (1983, appendix A) algorithm that module_mie.f:99-307
(BHMIE) implements: logarithmic derivative D_n(mx) by downward recurrence,
Riccati-Bessel psi_n, chi_n by upward recurrence, coefficients a_n, b_n.
Returns the efficiencies Qext, Qsca and the asymmetry parameter g.
"""
from __future__ import annotations

import numpy as np


def bhmie(x: float, m: complex, nang: int = 2):
    """x = k a (size parameter), m = relative refractive index."""
    del nang  # angular functions are not needed for the cross sections
    y = x * m
    xstop = x + 4.0 * x ** (1.0 / 3.0) + 2.0
    nstop = int(xstop)
    nmx = int(max(xstop, abs(y))) + 15
    D = np.zeros(nmx + 1, complex)
    for n in range(nmx, 0, -1):
        D[n - 1] = n / y - 1.0 / (D[n] + n / y)
    psi0, psi1 = np.cos(x), np.sin(x)
    chi0, chi1 = -np.sin(x), np.cos(x)
    xi1 = psi1 - 1j * chi1
    qsca = qext = gsum = 0.0
    an1 = bn1 = 0.0
    for n in range(1, nstop + 1):
        psi = (2.0 * n - 1.0) * psi1 / x - psi0
        chi = (2.0 * n - 1.0) * chi1 / x - chi0
        xi = psi - 1j * chi
        an = ((D[n] / m + n / x) * psi - psi1) / ((D[n] / m + n / x) * xi - xi1)
        bn = ((m * D[n] + n / x) * psi - psi1) / ((m * D[n] + n / x) * xi - xi1)
        qsca += (2 * n + 1) * (abs(an) ** 2 + abs(bn) ** 2)
        qext += (2 * n + 1) * (an.real + bn.real)
        gsum += (2.0 * n + 1.0) / (n * (n + 1.0)) * (an * np.conj(bn)).real
        if n > 1:
            gsum += (n - 1.0) * (n + 1.0) / n * (an1 * np.conj(an) + bn1 * np.conj(bn)).real
        psi0, psi1 = psi1, psi
        chi0, chi1 = chi1, chi
        xi1 = psi1 - 1j * chi1
        an1, bn1 = an, bn
    qsca *= 2.0 / x ** 2
    qext *= 2.0 / x ** 2
    g = 4.0 / (x ** 2 * qsca) * gsum
    return qext, qsca, g
