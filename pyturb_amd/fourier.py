"""FFT phase screens on the MI355X: ``PhaseScreen`` (Monte-Carlo draws).

Drop-in for src/pyturb/fourier.py:31-294.  Construction validates exactly
like the reference and builds the PSD filters on the host in float64
(:mod:`pyturb_amd.spectrum`).  ``generate`` draws the white noise with the
same host ``numpy.random.Generator`` stream, shapes and order as the
reference (so a seed gives the reference's realisation), uploads it and runs
one fused HIP call: noise colouring, 2-D inverse FFT, real/imag split into
screen pairs, subharmonic field with exact mean removal.  The arithmetic is
float32 on the device; ``dtype`` selects the host draw (and hence the
realisation) and the returned tensor's dtype.

``generate(..., noise=, sh_noise=)`` accepts injected device noise instead
(the parity and benchmark path).
"""

from __future__ import annotations

from typing import Any, Optional, Union

import numpy as np

from . import _hip, ops
from .backend import resolve_device
from .config import ScreenConfig
from .spectrum import build_filters, subharmonic_freqs

__all__ = ["PhaseScreen"]


class PhaseScreen:
    """Independent von Karman / Kolmogorov phase screens (radians).

    Parameters mirror the reference: ``n``, ``pixel_scale`` [m], ``r0`` [m],
    ``L0`` [m] (``numpy.inf`` = Kolmogorov), ``subharmonics`` (levels),
    ``power_law``, ``inner_scale``, ``seed``, ``device`` (``"mi355x"``) and
    ``dtype``.  Screens are device tensors.
    """

    def __init__(self, n: int, pixel_scale: float, r0: float, L0: float = 25.0,
                 subharmonics: int = 8, power_law: float = 11.0 / 3.0,
                 inner_scale: float = 0.0, seed: Any = None, device: str = "mi355x",
                 dtype: Union[str, np.dtype] = "float32") -> None:
        config = ScreenConfig.create(n, pixel_scale, r0, L0, device, dtype,
                                     finite_outer_scale=False)
        if subharmonics < 0:
            raise ValueError("subharmonics must be >= 0")
        if not np.isfinite(power_law) or power_law <= 2.0:
            raise ValueError("power_law must be > 2 and finite (Kolmogorov is 11/3)")
        if not np.isfinite(inner_scale) or inner_scale < 0:
            raise ValueError("inner_scale must be >= 0 and finite (0 disables it)")
        self.n = config.n
        self.pixel_scale = config.pixel_scale
        self.r0 = config.r0
        self.L0 = config.L0
        self.subharmonics = int(subharmonics)
        self.power_law = float(power_law)
        self.inner_scale = float(inner_scale)
        self.device = config.device
        self.dtype = config.dtype
        self._device_index = resolve_device(config.device)
        self._rng = np.random.default_rng(seed)
        amp, f, sh_amp = build_filters(self.n, self.pixel_scale, self.r0, self.L0,
                                       self.subharmonics, self.power_law, self.inner_scale)
        self._amplitude_host = amp
        self._sh_amp_host = sh_amp
        self._f_host = f
        self._sh_freqs_host = subharmonic_freqs(self.n, self.pixel_scale, self.subharmonics)
        self._dev = _hip.device(self._device_index)
        self._amplitude = _hip.to_device(amp, np.float32, self._dev)
        self._sh_amp = (_hip.to_device(sh_amp, np.float64, self._dev)
                        if self.subharmonics else None)
        self._n_sh = self.subharmonics

    # ------------------------------------------------------------------
    def draw_noise(self, count: int):
        """Host white noise in the reference's order (fourier.py:256-273)."""
        n_fft = (count + 1) // 2
        main = self._rng.standard_normal((2, n_fft, self.n, self.n), dtype=self.dtype)
        sh = [self._rng.standard_normal((2, count, 3, 3), dtype=self.dtype)
              for _ in range(self._n_sh)]
        return main, (np.stack(sh) if sh else None)

    def _upload_noise(self, main, sh):
        t = _hip.torch()
        noise = _hip.to_device(main, np.float32, self._dev)
        shn = _hip.to_device(sh, np.float32, self._dev) if sh is not None else None
        return noise, shn

    def generate(self, count: Optional[int] = None, *, noise=None, sh_noise=None,
                 screen_scale=None, out=None) -> Any:
        """Generate independent screens: ``(n, n)`` or ``(count, n, n)`` radians."""
        squeeze = count is None
        count = 1 if squeeze else int(count)
        if count < 1:
            raise ValueError("count must be >= 1")
        if noise is None:
            main, sh = self.draw_noise(count)
            noise, sh_noise = self._upload_noise(main, sh)
        screens = ops.generate(noise, sh_noise, self._amplitude, self._sh_amp, count,
                               screen_scale=screen_scale, out=out)
        if self.dtype == np.float64:
            screens = screens.double()
        return screens[0] if squeeze else screens

    def __repr__(self) -> str:
        extra = ""
        if self.power_law != 11.0 / 3.0:
            extra += f", power_law={self.power_law}"
        if self.inner_scale > 0.0:
            extra += f", inner_scale={self.inner_scale}"
        return (f"PhaseScreen(n={self.n}, pixel_scale={self.pixel_scale}, r0={self.r0}, "
                f"L0={self.L0}, subharmonics={self.subharmonics}{extra}, "
                f"device={self.device!r}, dtype={self.dtype.name!r})")
