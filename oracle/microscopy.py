"""TEST INFRASTRUCTURE ONLY: numpy restatement of the brightfield and
darkfield microscopes of cdmlib (microsbf.f, microsdf.f, quick-lens path
nquicklens = 1) and their pieces:

  ondeplanekxky       plane wave of wave vector (kx, ky, kz)   ondeplanekxky.f
  deltakroutine       bin of the incident wave                 deltakroutine.f
  irradiance          E0 from power and waist                  irradiance.f
  diffractefft2dlens  field in the lens pupil                  diffractefft2dlens.f
  fouriertoimage      pupil -> image plane                     fouriertoimage.f
  incidences          the condenser's (kx, ky) lattice         microsbf.f:90-172
  incidences_df       the darkfield ring of incidences         microsdf.f:91-221
  image_assembly      pupil fill, lens rotation, |E|^2 sums    microsbf.f:676-845,
                      (dark=True)                               microsdf.f:655-790
  direct_geometry     pupil step of nquicklens = 0             microsbf.f:139-146
  lens_field_direct   pupil field as a direct sum (nquicklens  microsbf.f:629-668,
                      = 0)                                      microssansinc.f:600-640

The pieces are pinned one by one to the compiled reference and its
fixtures, and the whole pipeline (these functions around the oracle's
inverserig) to the reference's own microsbf_, microsdf_, microsbfneg_ and
microsdfneg_, quick and direct-sum lens (tests/golden/microscopes.npz).

Only tests/ imports this module.
"""
from __future__ import annotations

import numpy as np

from farfield import nint, tabfft2


def ondeplanekxky(pos, k0, E0, pp, kx, ky):
    """ondeplanekxky.f: AoS (Ex, Ey, Ez) at positions pos [n, 3]; the
    reference's ss argument is unused (s = sqrt(1 - pp))."""
    if pp < 0.0 or pp > 1.0:
        raise ValueError("problem in plance incident wave")
    p, s = np.sqrt(pp), np.sqrt(1.0 - pp)
    E0s, E0p = E0 * s, E0 * p
    kz = np.sqrt(k0 * k0 - kx * kx - ky * ky)
    if kz >= 0.99999 * k0:
        e = np.array([E0p, E0s, 0.0], np.complex128)
    else:
        r = np.sqrt(kx * kx + ky * ky)
        nsx, nsy = -ky / r, kx / r
        npx, npy, npz = kz * nsy, -kz * nsx, -kx * nsy + ky * nsx
        nn = np.sqrt(npx * npx + npy * npy + npz * npz)
        npx, npy, npz = npx / nn, npy / nn, npz / nn
        e = np.array([E0s * nsx + E0p * npx, E0s * nsy + E0p * npy, E0p * npz], np.complex128)
    ph = np.exp(1j * (np.asarray(pos) @ np.array([kx, ky, kz])))
    return (ph[:, None] * e[None, :]).reshape(-1)


def deltakroutine(kxinc, kyinc, dkx, dky, numaper):
    i = nint((kxinc / dkx))
    j = nint((kyinc / dky))
    if (dkx * i) ** 2 + (dky * j) ** 2 <= numaper * numaper:
        return i, j
    return int(kxinc / dkx), int(kyinc / dky)


def irradiance(P0, w0):
    c = 299792458.0
    eps0 = 1.0 / (c * c * 4.0 * np.pi * 1e-7)
    irra = P0 / (w0 * w0 * np.pi)
    return np.sqrt(2.0 * irra / c / eps0) + 0j, irra


def diffractefft2dlens(FFbox, nx, ny, nz, nfft2d, k0, xs1, ys1, zs, d, numaper, nside):
    """diffractefft2dlens.f: E[ii, jj, c], imax, deltak (numaper in k units)."""
    N = nfft2d
    h = N // 2
    dk = 2.0 * np.pi / (N * d)
    if N > 16384:
        raise ValueError("nfft2d too large")
    if dk >= numaper:
        raise ValueError("In FFT lens nfft2d too small")
    var1 = (xs1 + h * d) * dk
    var2 = (ys1 + h * d) * dk
    tab = tabfft2(N)
    imax = nint((k0 / dk)) + 1
    if 2 * imax + 1 > N:
        raise ValueError("In FFT diffraction for lens nfft2d too small!")
    if 2 * imax + 1 < 7:
        raise ValueError("In FFT diffract nfft2d too small")
    ds = float(nside)
    F = np.asarray(FFbox).reshape(nz, ny, nx, 3)
    n = 2 * imax + 1
    E = np.zeros((n, n, 3), np.complex128)
    iv = np.arange(-imax, imax + 1)
    KX, KY = np.meshgrid(iv * dk, iv * dk, indexing="ij")
    ap = KX * KX + KY * KY < numaper * numaper
    KZ = np.sqrt(np.where(ap, k0 * k0 - KX * KX - KY * KY, 1.0))
    bins = tab[iv + h] - 1
    vph = np.exp(-1j * (var1 * iv[:, None] + var2 * iv[None, :]))
    xpos, ypos = tab[:nx] - 1, tab[:ny] - 1
    for k in range(nz):
        P = np.zeros((3, N, N), np.complex128)
        P[:, ypos[:, None], xpos[None, :]] = np.moveaxis(F[k], 2, 0)
        P = np.fft.fft2(P, axes=(1, 2))
        Ex, Ey, Ez = P[:, bins[None, :], bins[:, None]]
        ct = (KX * Ex + KY * Ey + KZ * ds * Ez) / k0
        ph = np.exp(-1j * ds * KZ * zs[k]) * vph / (-2.0 * np.pi * 1j * KZ)
        E[..., 0] += np.where(ap, (Ex - KX * ct / k0) * ph, 0.0)
        E[..., 1] += np.where(ap, (Ey - KY * ct / k0) * ph, 0.0)
        E[..., 2] += np.where(ap, (Ez - KZ * ds * ct / k0) * ph, 0.0)
    return E * k0 * k0, imax, dk


def fouriertoimage(dk, gross, planes):
    """fouriertoimage.f on N x N planes ([y][x], x fastest, in place): forward
    FFT, then the quadrant swap scaled by deltak^2 / gross."""
    N = planes[0].shape[0]
    h = N // 2
    tmp = dk * dk / gross
    out = []
    for P in planes:
        A = np.fft.fft2(P).reshape(-1).copy()
        for i in range(-h, h):
            for j in range(-h, 0):
                ix = i + 1 if i >= 0 else N + i + 1
                iy = j + 1 if j >= 0 else N + j + 1
                ind = ix + N * (iy - 1) - 1
                kk = i + h + 1 + N * (j + h) - 1
                A[kk], A[ind] = A[ind] * tmp, A[kk] * tmp
        out.append(A.reshape(N, N))
    return out


def fouriertoimage_even(dk, gross, planes):
    """fouriertoimage for an even N: the quadrant swap is fftshift, so the
    result is fftshift(fft2(P)) * deltak^2 / gross (tests check it against
    the literal swap and the compiled reference)."""
    tmp = dk * dk / gross
    return [np.fft.fftshift(np.fft.fft2(P)) * tmp for P in planes]


def direct_geometry(nfft2d, k0, d):
    """microsbf.f:139-146 (nquicklens = 0): deltakx = 2 pi / (nfft2d d) / 2^k
    for the first k with imaxk0 = nint(k0 / deltakx) + 1 > 5.  Returns
    imaxk0, deltakx."""
    k = 0
    while True:
        dkx = 2.0 * np.pi / (nfft2d * d) / 2.0 ** k
        imax = nint(k0 / dkx) + 1
        if imax > 5:
            return imax, dkx
        k += 1


def lens_field_direct(pos, FF, k0, dk, imax, numaper, nside=1):
    """microsbf.f:629-668 (microssansinc.f:600-640 for nside = -1): at every
    pupil bin (i, j) with |k| <= numaper (k units), n = (kx, ky, nside kz) /
    k0 and E[i, j] = k0^2 / (-2 pi i kz) sum_dipoles exp(-i k0 n.r) (p -
    (n.p) n); zero elsewhere.  pos [nd, 3], FF [3 nd] (p per dipole)."""
    P = np.asarray(FF).reshape(-1, 3)
    n = 2 * imax + 1
    E = np.zeros((n, n, 3), np.complex128)
    for a, i in enumerate(range(-imax, imax + 1)):
        for b, j in enumerate(range(-imax, imax + 1)):
            kx, ky = dk * i, dk * j
            if np.sqrt(kx * kx + ky * ky) > numaper:
                continue
            kz = np.sqrt(k0 * k0 - kx * kx - ky * ky)
            nv = np.array([kx / k0, ky / k0,
                           nside * np.sqrt(1.0 - (kx / k0) ** 2 - (ky / k0) ** 2)])
            ph = np.exp(-1j * k0 * (pos @ nv))
            pn = P @ nv
            Em = (ph[:, None] * (P - pn[:, None] * nv[None, :])).sum(0)
            E[a, b] = Em * k0 * k0 / (-2.0 * np.pi * 1j * kz)
    return E


def incidences(k0, numaperinc, d, nfft2d, xs, ys, direct=False):
    """microsbf.f:90-172: the condenser's (kx, ky) list (idelta outer,
    jdelta inner), deltak, deltakx.  nquicklens = 1 by default; direct: the
    nquicklens = 0 steps (microsbf.f:139-157, 179-183)."""
    deltakm = np.pi / max(np.max(xs) - np.min(xs), np.max(ys) - np.min(ys))
    numaperk = k0 * numaperinc
    if direct:
        _, dkx = direct_geometry(nfft2d, k0, d)
        q = float(nint(deltakm / dkx))  # dnint
        k = 0
        while True:
            dkinc = dkx * q / 2.0 ** k
            imaxinc = nint(k0 / dkinc) + 1
            if imaxinc > 2:
                break
            k += 1
        if dkinc >= numaperk:
            raise ValueError("In FFT lens nfft2d too small")
    else:
        dkx = 2.0 * np.pi / (d * nfft2d)
        if dkx >= numaperk:
            raise ValueError("In FFT lens nfft2d too small:Increase size of FFT 1")
        imul = int(deltakm / dkx)
        if imul == 0:
            raise ValueError("In FFT lens nfft2d too small:Increase size of FFT 2")
        while True:
            dkinc = dkx * imul
            imaxinc = nint((k0 / dkinc)) + 1
            if imaxinc > 3:
                break
            imul -= 1
            if imul == 0:
                raise ValueError("In FFT lens nfft2d too small:Increase size of FFT 3")
    inc = []
    for idel in range(-imaxinc, imaxinc + 1):
        for jdel in range(-imaxinc, imaxinc + 1):
            kx, ky = idel * dkinc, jdel * dkinc
            if kx * kx + ky * ky <= numaperk * numaperk:
                inc.append((idel, jdel, kx, ky))
    return inc, dkinc, dkx


def incidences_df(k0, numaperinc, d, nfft2d, xs, ys, direct=False):
    """microsdf.f:91-221 with nquicklens = 1: deltak as the brightfield's
    (microsdf.f:91-137), then ideltam = max(int(2 pi / (deltak NA_inc / k0))
    + 1, 8) incidences at phi = 2 pi idelta / ideltam on the cone of NA_inc
    (kx, ky) = k0 NA_inc (cos phi, sin phi), for pp = 0 then pp = 1
    (npolainc = 0 there).  Returns [(idelta, pp, kx, ky)], deltak, deltakx."""
    _, dkinc, dkx = incidences(k0, numaperinc, d, nfft2d, xs, ys, direct=direct)
    dtheta = dkinc / k0 * numaperinc
    ideltam = max(int(2.0 * np.pi / dtheta) + 1, 8)
    out = []
    for pp in (0.0, 1.0):
        for idel in range(ideltam):
            phi = idel * 2.0 * np.pi / ideltam
            out.append((idel, pp, np.cos(phi) * k0 * numaperinc, np.sin(phi) * k0 * numaperinc))
    return out, dkinc, dkx


def image_assembly(Elens_list, inc_list, k0, E0, numaper, zlens, gross, dkinc, imax, dk, nfft2d,
                   dark=False, nside=1):
    """microsbf.f:676-845: for each (pupil field, (idelta, jdelta, kxinc, kyinc,
    pp)) fill the pupil (focus phase; the incident wave at its bin), rotate by
    the lens, fouriertoimage, add |E|^2; returns the six sqrt-images
    (scattered x, y, z, with incident x, y, z) as [y][x] arrays.  dark: the
    darkfield's incident wave (microsdf.f:668-717): its bin from the pupil
    step dk, added as i E / dk^2 on every component.  nside = -1: the
    objective on the z < 0 side (microssansinc.f, microsbfneg / microsdfneg:
    no incident wave in the pupil, so the with-incident images equal the
    scattered ones; the lens rotation takes u_z, v_z times nside,
    microssansinc.f:494-497, 1397-1400)."""
    N = nfft2d
    acc = [np.zeros((N, N)) for _ in range(6)]
    iv = np.arange(-imax, imax + 1)
    for El, (idel, jdel, kxinc, kyinc, pp) in zip(Elens_list, inc_list):
        if dark:
            ikx, jky = deltakroutine(kxinc, kyinc, dk, dk, k0)
        else:
            ikx, jky = deltakroutine(kxinc, kyinc, dkinc, dkinc, k0)
        planes = [np.zeros((N, N), np.complex128) for _ in range(6)]
        for a, i in enumerate(iv):
            for b, j in enumerate(iv):
                kx, ky = dk * i, dk * j
                t = kx * kx + ky * ky
                if t > numaper * numaper:
                    continue
                kz = np.sqrt(k0 * k0 - t)
                zf = np.exp(1j * kz * zlens)
                ix, iy = i % N, j % N
                s = El[a, b] * zf
                w = s.copy()
                if i == ikx and j == jky and nside == 1:
                    e = ondeplanekxky(np.zeros((1, 3)), k0, E0, pp, kx, ky)
                    if dark:
                        w = w + np.array([e[0], e[1], e[2]]) / (dk * dk) * zf * 1j
                    else:
                        w = w + np.array([e[0] / (dk * dk), e[1] / (dk * dk), e[2] / dk]) * zf
                if t != 0.0:
                    u = np.array([kx / k0, ky / k0, 0.0])
                    u[2] = np.sqrt(1.0 - u[0] ** 2 - u[1] ** 2) * nside
                    v = np.array([-kx / k0 / gross, -ky / k0 / gross, 0.0])
                    v[2] = np.sqrt(1.0 - v[0] ** 2 - v[1] ** 2) * nside
                    u1 = u[1] * v[2] - u[2] * v[1]
                    u2 = -u[0] * v[2] + u[2] * v[0]
                    co = u @ v
                    si = np.sqrt(u1 * u1 + u2 * u2)
                    u1, u2 = u1 / si, u2 / si
                    sc = np.sqrt(u[2] / v[2])
                    Rm = np.array([[u1 * u1 + (1 - u1 * u1) * co, u1 * u2 * (1 - co), u2 * si],
                                   [u1 * u2 * (1 - co), u2 * u2 + (1 - u2 * u2) * co, -u1 * si],
                                   [-u2 * si, u1 * si, co]])
                    s = (Rm @ s) * sc
                    w = (Rm @ w) * sc
                for c in range(3):
                    planes[c][iy, ix] = s[c]
                    planes[3 + c][iy, ix] = w[c]
        img = fouriertoimage_even(dk, gross, planes) if N % 2 == 0 else fouriertoimage(dk, gross, planes)
        for c in range(6):
            acc[c] += np.abs(img[c]) ** 2
    return [np.sqrt(a) for a in acc]
