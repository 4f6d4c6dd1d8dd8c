"""Single frozen-flow layer, spectral method: ``FourierFlowScreen``.

Drop-in for src/pyturb/flow.py:34-139.  The fixed spectrum and per-level
subharmonic coefficients are drawn on the host from the same generator
stream as the reference (flow.py:71-94) and uploaded once (complex64
spectrum, complex128 coefficients); ``translate`` is one
``pt_spectral_frames`` call with one layer.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from . import _hip, ops
from .fourier import PhaseScreen

__all__ = ["FourierFlowScreen"]


def draw_layer(rng, template: PhaseScreen):
    """Host draw of one layer's spectrum and subharmonic 3x3 blocks (float64)."""
    n = template.n
    w = rng.standard_normal((2, n, n), dtype=template.dtype)
    spectrum = (w[0].astype(np.float64) + 1j * w[1].astype(np.float64)) * template._amplitude_host
    coeffs = np.zeros((template._n_sh, 3, 3), dtype=np.complex128)
    for p in range(template._n_sh):
        v = rng.standard_normal((2, 3, 3), dtype=template.dtype).astype(np.float64)
        coeffs[p] = (v[0] + 1j * v[1]) * template._sh_amp_host[p]
    return spectrum, coeffs


class FourierFlowScreen:
    """One frozen layer evaluated at arbitrary continuous offsets (metres)."""

    def __init__(self, template: PhaseScreen, seed: Any = None) -> None:
        self.template = template
        self.n = template.n
        self.pixel_scale = template.pixel_scale
        self.r0 = template.r0
        self.L0 = template.L0
        self.device = template.device
        self.dtype = template.dtype
        self._rng = np.random.default_rng(seed)
        self.reseed()

    def reseed(self, seed: Any = None) -> "FourierFlowScreen":
        if seed is not None:
            self._rng = np.random.default_rng(seed)
        spec, coeffs = draw_layer(self._rng, self.template)
        self._spectrum_host = spec
        self._sh_coeffs_host = coeffs
        dev = self.template._dev
        self._spectrum = _hip.to_device(spec.astype(np.complex64)[None], None, dev)
        self._sh_coeffs = (_hip.to_device(coeffs[:, None], np.complex128, dev)
                           if self.template._n_sh else None)
        return self

    def translate(self, sx: float, sy: float) -> Any:
        """Screen blown by ``(sx, sy)`` metres (axis 0, axis 1), radians."""
        disp = _hip.to_device(np.array([[[float(sx), float(sy)]]]), np.float64, self.template._dev)
        out = ops.spectral_frames(self._spectrum, disp, self._sh_coeffs, self.pixel_scale)[0]
        return out.double() if self.dtype == np.float64 else out

    @property
    def _f(self):
        return self.template._f_host

    def __repr__(self) -> str:
        return (f"FourierFlowScreen(n={self.n}, pixel_scale={self.pixel_scale}, "
                f"r0={self.r0}, L0={self.L0}, device={self.device!r})")
