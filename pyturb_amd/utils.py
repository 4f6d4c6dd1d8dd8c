"""Unit conversions and the host-side structure-function estimator."""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

_RAD_TO_ARCSEC = 180.0 * 3600.0 / np.pi


def to_numpy(array) -> np.ndarray:
    """Copy a device tensor (or any array) to a host ``numpy.ndarray``."""
    if hasattr(array, "detach") and hasattr(array, "cpu"):
        return array.detach().cpu().numpy()
    if hasattr(array, "get"):
        return array.get()
    return np.asarray(array)


def r0_from_seeing(seeing: float, wavelength: float = 500e-9) -> float:
    """Fried parameter [m] from seeing FWHM [arcsec]: FWHM = 0.98 lambda / r0."""
    return 0.98 * wavelength / (seeing / _RAD_TO_ARCSEC)


def seeing_from_r0(r0: float, wavelength: float = 500e-9) -> float:
    return 0.98 * wavelength / r0 * _RAD_TO_ARCSEC


def r0_at_wavelength(r0: float, wavelength_in: float, wavelength_out: float) -> float:
    return r0 * (wavelength_out / wavelength_in) ** 1.2


def opd_to_phase(opd, wavelength: float):
    return opd * (2.0 * np.pi / wavelength)


def phase_to_opd(phase, wavelength: float):
    return phase * (wavelength / (2.0 * np.pi))


def structure_function(phase, pixel_scale: float = 1.0,
                       max_separation: Optional[int] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Axis-averaged phase structure function ``D(r)`` (host float64).

    Accepts a screen ``(n, n)`` or stack ``(count, n, n)``; device tensors
    are copied to the host.  ``D(r) = 0.5 * (<d_rows^2> + <d_cols^2>)`` at
    pixel separations ``1..max_separation`` (default ``n // 4``).
    """
    ph = to_numpy(phase).astype(np.float64)
    if ph.ndim == 2:
        ph = ph[None]
    n = min(ph.shape[-2], ph.shape[-1])
    smax = n // 4 if max_separation is None else int(max_separation)
    if not 1 <= smax < n:
        raise ValueError("max_separation must be in [1, n)")
    out = np.empty(smax)
    for k in range(smax):
        s = k + 1
        a = ph[..., s:, :] - ph[..., :-s, :]
        b = ph[..., :, s:] - ph[..., :, :-s]
        out[k] = 0.5 * (np.mean(a * a) + np.mean(b * b))
    return np.arange(1, smax + 1) * pixel_scale, out
