"""Host FP64 setup of the spectral filters (one-time, per screen geometry).

The PSD amplitude filter, the near-DC cell integration and the subharmonic
amplitudes are computed once in float64 on the host and uploaded; the HIP
kernels apply them.  Same arithmetic as src/pyturb/fourier.py:131-228 of the
reference, vectorised over cells.
"""

from __future__ import annotations

import math

import numpy as np


def psd(f_squared, r0, L0, power_law=11.0 / 3.0, inner_scale=0.0):
    """Modified von Karman phase PSD [rad^2 m^2] (fourier.py:131-149)."""
    f2 = np.asarray(f_squared, dtype=np.float64)
    floor = 0.0 if math.isinf(L0) else 1.0 / (L0 * L0)
    out = (0.023 * r0 ** (-5.0 / 3.0)) * np.power(f2 + floor, -0.5 * power_law)
    if inner_scale > 0.0:
        fm = 5.92 / (2.0 * math.pi * inner_scale)
        out = out * np.exp(-f2 / (fm * fm))
    return out


def band_amplitudes(fx, fy, width, r0, L0, power_law=11.0 / 3.0, inner_scale=0.0,
                    samples=33):
    """Cell-integrated mode amplitudes for arrays of cell centres.

    33x33 midpoint quadrature per cell (fourier.py:151-172):
    ``amp^2 = mean(PSD) w^2 * sum(PSD f^2)/sum(PSD) / |f_c|^2``.
    """
    fx = np.atleast_1d(np.asarray(fx, dtype=np.float64))
    fy = np.atleast_1d(np.asarray(fy, dtype=np.float64))
    u = ((np.arange(samples) + 0.5) / samples - 0.5) * width
    gx = fx[:, None, None] + u[None, :, None]
    gy = fy[:, None, None] + u[None, None, :]
    f2 = gx * gx + gy * gy
    p = psd(f2, r0, L0, power_law, inner_scale)
    power = p.mean(axis=(1, 2)) * width * width
    f2_eff = (p * f2).sum(axis=(1, 2)) / p.sum(axis=(1, 2))
    return np.sqrt(power * f2_eff / (fx * fx + fy * fy))


def build_filters(n, dx, r0, L0, subharmonics=8, power_law=11.0 / 3.0, inner_scale=0.0):
    """Main-grid amplitude (n, n) and per-level 3x3 subharmonic amplitudes.

    Returns ``(amplitude, f, sh_amp)`` in float64: ``amplitude`` is
    ``sqrt(PSD) * df`` on the ``fftfreq`` grid with DC zeroed and the cells
    with ``|k|, |l| <= min(8, (n-1)//2)`` re-integrated; ``sh_amp`` is
    (levels, 3, 3) with zero centres (fourier.py:174-228).
    """
    n = int(n)
    df = 1.0 / (n * dx)
    f = np.fft.fftfreq(n, d=dx)
    f2 = np.add.outer(f * f, f * f)
    f2[0, 0] = 1.0
    amp = np.sqrt(psd(f2, r0, L0, power_law, inner_scale)) * df
    amp[0, 0] = 0.0
    kmax = min(8, (n - 1) // 2)
    if kmax > 0:
        ks = np.arange(-kmax, kmax + 1)
        ki, kj = np.meshgrid(ks, ks, indexing="ij")
        keep = (ki != 0) | (kj != 0)
        ki, kj = ki[keep], kj[keep]
        amp[ki, kj] = band_amplitudes(f[ki], f[kj], df, r0, L0, power_law, inner_scale)
    levels = int(subharmonics)
    sh_amp = np.zeros((levels, 3, 3))
    if levels:
        off = np.array([-1.0, 0.0, 1.0])
        ia, ib = np.meshgrid(np.arange(3), np.arange(3), indexing="ij")
        keep = (ia != 1) | (ib != 1)
        for p in range(levels):
            dfp = df / 3.0 ** (p + 1)
            vals = band_amplitudes(off[ia[keep]] * dfp, off[ib[keep]] * dfp, dfp, r0, L0,
                                   power_law, inner_scale)
            sh_amp[p][keep] = vals
    return amp, f, sh_amp


def subharmonic_freqs(n, dx, levels):
    """(levels, 3) mode frequencies (-1, 0, 1) * df / 3^p [1/m]."""
    df = 1.0 / (n * dx)
    return np.array([[-1.0, 0.0, 1.0]]) * (df / 3.0 ** np.arange(1, levels + 1))[:, None]


def dlnpsd_dL0(f_squared, L0, power_law=11.0 / 3.0):
    """d ln(PSD) / d L0 = power_law * L0^-3 / (f^2 + L0^-2) (zero for infinite L0)."""
    f2 = np.asarray(f_squared, dtype=np.float64)
    if math.isinf(L0):
        return np.zeros_like(f2)
    return power_law * L0 ** -3.0 / (f2 + 1.0 / (L0 * L0))


def band_dlnamp_dL0(fx, fy, width, L0, power_law=11.0 / 3.0, inner_scale=0.0, samples=33):
    """d ln(band amplitude) / d L0 of :func:`band_amplitudes` cells.

    With amp^2 = mean(p) w^2 (sum p f^2 / sum p) / f_c^2 the mean(p) and
    sum(p) terms cancel: d ln amp = 0.5 * sum(p q f^2) / sum(p f^2), with
    q = d ln p / d L0 per quadrature sample.
    """
    fx = np.atleast_1d(np.asarray(fx, dtype=np.float64))
    fy = np.atleast_1d(np.asarray(fy, dtype=np.float64))
    u = ((np.arange(samples) + 0.5) / samples - 0.5) * width
    gx = fx[:, None, None] + u[None, :, None]
    gy = fy[:, None, None] + u[None, None, :]
    f2 = gx * gx + gy * gy
    p = psd(f2, 1.0, L0, power_law, inner_scale)
    q = dlnpsd_dL0(f2, L0, power_law)
    return 0.5 * (p * q * f2).sum(axis=(1, 2)) / (p * f2).sum(axis=(1, 2))


def build_filters_dL0(n, dx, L0, subharmonics=8, power_law=11.0 / 3.0, inner_scale=0.0):
    """Unit-r0 filters and their logarithmic L0 derivatives (FP64).

    Returns ``(amplitude, dlnamp, sh_amp, sh_dlnamp)``: the output of
    :func:`build_filters` at r0 = 1 and d ln(.)/d L0 of each entry (zero
    where the amplitude is zero).  The r0 dependence is the exact factor
    r0^(-5/6), so these are all the filter derivatives a profile fit needs.
    """
    amp, f, sh_amp = build_filters(n, dx, 1.0, L0, subharmonics, power_law, inner_scale)
    n = int(n)
    df = 1.0 / (n * dx)
    f2 = np.add.outer(f * f, f * f)
    dln = 0.5 * dlnpsd_dL0(f2, L0, power_law)
    dln[0, 0] = 0.0
    kmax = min(8, (n - 1) // 2)
    if kmax > 0:
        ks = np.arange(-kmax, kmax + 1)
        ki, kj = np.meshgrid(ks, ks, indexing="ij")
        keep = (ki != 0) | (kj != 0)
        ki, kj = ki[keep], kj[keep]
        dln[ki, kj] = band_dlnamp_dL0(f[ki], f[kj], df, L0, power_law, inner_scale)
    levels = int(subharmonics)
    sh_dln = np.zeros((levels, 3, 3))
    if levels:
        off = np.array([-1.0, 0.0, 1.0])
        ia, ib = np.meshgrid(np.arange(3), np.arange(3), indexing="ij")
        keep = (ia != 1) | (ib != 1)
        for p in range(levels):
            dfp = df / 3.0 ** (p + 1)
            sh_dln[p][keep] = band_dlnamp_dL0(off[ia[keep]] * dfp, off[ib[keep]] * dfp, dfp, L0,
                                              power_law, inner_scale)
    return amp, dln, sh_amp, sh_dln
