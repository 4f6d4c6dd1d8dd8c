"""Validated immutable configuration (the reference's ValueError contract).

Restates the input checks of src/pyturb/config.py (GridConfig :21-45,
ScreenConfig :69-116, AtmosphereConfig :119-293, ExtrusionConfig
:296-416): same conditions, same exception type and the same key words in
the messages, so callers that match on them keep working.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Optional, Sequence, Tuple, Union

import numpy as np

from .profiles import Layer, _fractions
from .utils import r0_from_seeing

__all__ = ["GridConfig", "ScreenConfig", "AtmosphereConfig", "ExtrusionConfig", "LayerConfig"]


@dataclass(frozen=True)
class GridConfig:
    n: int
    pixel_scale: float
    device: str
    dtype: np.dtype

    @classmethod
    def create(cls, n, pixel_scale, device, dtype) -> "GridConfig":
        if not isinstance(n, (int, np.integer)) or isinstance(n, bool) or n < 2:
            raise ValueError("n must be an integer of at least 2")
        if not np.isfinite(pixel_scale) or pixel_scale <= 0:
            raise ValueError("pixel_scale must be positive and finite")
        try:
            dt = np.dtype(dtype)
        except TypeError as exc:
            raise ValueError("dtype must be float32 or float64") from exc
        if dt not in (np.dtype("float32"), np.dtype("float64")):
            raise ValueError("dtype must be float32 or float64")
        return cls(int(n), float(pixel_scale), device, dt)


@dataclass(frozen=True)
class LayerConfig:
    altitude: float
    cn2_fraction: float
    wind_speed: float
    wind_direction: float
    L0: float

    def to_layer(self) -> Layer:
        return Layer(self.altitude, self.cn2_fraction, self.wind_speed,
                     self.wind_direction, self.L0)


@dataclass(frozen=True)
class ScreenConfig:
    n: int
    pixel_scale: float
    r0: float
    L0: float
    device: str
    dtype: np.dtype

    @classmethod
    def create(cls, n, pixel_scale, r0, L0, device, dtype, *, finite_outer_scale: bool):
        grid = GridConfig.create(n, pixel_scale, device, dtype)
        if not np.isfinite(r0) or r0 <= 0:
            raise ValueError("r0 must be positive and finite")
        outer = np.inf if L0 is None else L0
        if np.isnan(outer) or outer <= 0:
            raise ValueError("L0 must be positive (use numpy.inf for Kolmogorov)")
        if finite_outer_scale and not np.isfinite(outer):
            raise ValueError("InfinitePhaseScreen requires a finite positive outer scale L0")
        return cls(grid.n, grid.pixel_scale, float(r0), float(outer), grid.device, grid.dtype)


@dataclass(frozen=True)
class AtmosphereConfig:
    layers: Tuple[LayerConfig, ...]
    r0_zenith: float
    wavelength: float
    zenith_angle: float
    diameter: float
    grid: GridConfig
    L0_override: Optional[float]
    power_law: float
    inner_scale: float
    subharmonics: int
    field_of_view: float
    tau_boil: Tuple[float, ...]
    has_tau_boil: bool
    engine: str
    interp: str
    lgs_altitude: Optional[float]
    dispersion: Optional[str]
    wet_fraction: float
    seed: Optional[int]

    @classmethod
    def create(cls, layers, r0, seeing, wavelength, zenith_angle, diameter, n, L0,
               power_law, inner_scale, subharmonics, field_of_view, tau_boil, engine,
               interp, lgs_altitude, dispersion, wet_fraction, device, dtype, seed):
        model = list(layers)
        if not model:
            raise ValueError("at least one layer is required")
        if (r0 is None) == (seeing is None):
            raise ValueError("give exactly one of r0 or seeing")
        if not np.isfinite(wavelength) or wavelength <= 0:
            raise ValueError("wavelength must be positive and finite [m]")
        if not np.isfinite(zenith_angle) or not 0.0 <= zenith_angle < 90.0:
            raise ValueError("zenith_angle must be in [0, 90) degrees")
        if not np.isfinite(diameter) or diameter <= 0:
            raise ValueError("diameter must be positive and finite")
        if not isinstance(n, (int, np.integer)) or n < 2:
            raise ValueError("n must be an integer of at least 2")
        grid = GridConfig.create(n, float(diameter) / n, device, dtype)
        if not np.isfinite(field_of_view) or field_of_view < 0:
            raise ValueError("field_of_view must be finite and >= 0 arcsec")
        if engine not in ("spectral", "extrude"):
            raise ValueError("engine must be 'spectral' or 'extrude'")
        if interp not in ("cubic", "linear", "lanczos"):
            raise ValueError("interp must be 'cubic', 'linear', or 'lanczos'")
        if not np.isfinite(power_law) or power_law <= 2.0:
            raise ValueError("power_law must be > 2 and finite (Kolmogorov is 11/3)")
        if not np.isfinite(inner_scale) or inner_scale < 0:
            raise ValueError("inner_scale must be >= 0 and finite (0 disables it)")
        if not isinstance(subharmonics, (int, np.integer)) or subharmonics < 0:
            raise ValueError("subharmonics must be an integer >= 0")
        if L0 is not None and (np.isnan(L0) or L0 <= 0):
            raise ValueError("L0 override must be positive (use numpy.inf for Kolmogorov)")
        if engine == "extrude" and power_law != 11.0 / 3.0:
            raise ValueError("non-Kolmogorov power_law requires engine='spectral'")
        if engine == "extrude" and inner_scale > 0:
            raise ValueError("inner_scale requires engine='spectral'")
        if lgs_altitude is not None and (not np.isfinite(lgs_altitude) or lgs_altitude <= 0):
            raise ValueError("lgs_altitude must be positive and finite [m]")
        if dispersion not in (None, "edlen", "ciddor"):
            raise ValueError("dispersion must be None, 'edlen', or 'ciddor'")
        if not np.isfinite(wet_fraction) or not 0.0 <= wet_fraction <= 1.0:
            raise ValueError("wet_fraction must be finite and in [0, 1]")
        if wet_fraction > 0.0 and dispersion != "ciddor":
            raise ValueError("wet_fraction > 0 requires dispersion='ciddor'")
        if seeing is not None:
            if not np.isfinite(seeing) or seeing <= 0:
                raise ValueError("seeing must be positive and finite [arcsec]")
            r0 = r0_from_seeing(seeing, wavelength)
        if not np.isfinite(r0) or r0 <= 0:
            raise ValueError("r0 (or the r0 implied by seeing) must be positive and finite [m]")
        fr = _fractions(model)
        cfg_layers = tuple(
            LayerConfig(ly.altitude, float(f), ly.wind_speed, ly.wind_direction,
                        ly.L0 if L0 is None else float(L0))
            for ly, f in zip(model, fr))
        if engine == "extrude" and any(not np.isfinite(ly.L0) for ly in cfg_layers):
            raise ValueError("engine='extrude' requires a finite outer scale L0 for every layer")
        if tau_boil is None:
            tau = np.full(len(cfg_layers), np.inf)
        else:
            tau = np.broadcast_to(np.asarray(tau_boil, dtype=np.float64),
                                  (len(cfg_layers),)).astype(np.float64)
        if np.any(~np.isfinite(tau) & ~np.isinf(tau)) or np.any(tau <= 0):
            raise ValueError("tau_boil must be positive (or None for frozen flow)")
        return cls(cfg_layers, float(r0), float(wavelength), float(zenith_angle),
                   float(diameter), grid, None if L0 is None else float(L0),
                   float(power_law), float(inner_scale), int(subharmonics),
                   float(field_of_view), tuple(float(v) for v in tau),
                   tau_boil is not None, engine, interp,
                   None if lgs_altitude is None else float(lgs_altitude), dispersion,
                   float(wet_fraction), seed)


@dataclass(frozen=True)
class ExtrusionConfig:
    grid: GridConfig
    layer_r0: Tuple[float, ...]
    layer_L0: Tuple[float, ...]
    layer_wind: Tuple[Tuple[float, float], ...]
    layer_altitude_los: Tuple[float, ...]
    field_of_view_pix: Tuple[float, ...]
    stencil_rows: int
    interp: str
    lgs_altitude_los: Optional[float]
    seeds: Tuple[Any, ...]
    far_points: int = 0

    @classmethod
    def create(cls, n, pixel_scale, layer_r0, layer_L0, layer_wind, layer_altitude_los,
               field_of_view_pix, stencil_rows, interp, lgs_altitude_los, device, dtype,
               seeds, far_points=0):
        grid = GridConfig.create(n, pixel_scale, device, dtype)
        r0s = tuple(float(v) for v in layer_r0)
        L0s = tuple(float(v) for v in layer_L0)
        winds = tuple((float(a), float(b)) for a, b in layer_wind)
        alts = tuple(float(v) for v in layer_altitude_los)
        L = len(r0s)
        if L == 0:
            raise ValueError("at least one extrusion layer is required")
        if not (len(L0s) == len(winds) == len(alts) == L):
            raise ValueError("layer_r0, layer_L0, layer_wind and layer_altitude_los must "
                             "have equal length")
        if any(not np.isfinite(v) or v <= 0 for v in r0s):
            raise ValueError("every layer_r0 value must be positive and finite")
        if any(not np.isfinite(v) or v <= 0 for v in L0s):
            raise ValueError("every layer_L0 value must be positive and finite")
        if any(not np.isfinite(v) for w in winds for v in w):
            raise ValueError("every layer_wind component must be finite")
        if any(not np.isfinite(v) for v in alts):
            raise ValueError("every layer_altitude_los value must be finite")
        if np.ndim(field_of_view_pix) == 0:
            fov = (float(field_of_view_pix),) * L
        else:
            fov = tuple(float(v) for v in field_of_view_pix)
        if len(fov) != L:
            raise ValueError("field_of_view_pix must be scalar or one value per layer")
        if any(not np.isfinite(v) or v < 0 for v in fov):
            raise ValueError("field_of_view_pix values must be finite and >= 0")
        if not isinstance(stencil_rows, (int, np.integer)) or not 1 <= stencil_rows < grid.n:
            raise ValueError("stencil_rows must be in [1, n)")
        if interp not in ("cubic", "linear", "lanczos"):
            raise ValueError("interp must be 'cubic', 'linear', or 'lanczos'")
        if lgs_altitude_los is not None and (not np.isfinite(lgs_altitude_los)
                                             or lgs_altitude_los <= 0):
            raise ValueError("lgs_altitude_los must be positive and finite")
        if lgs_altitude_los is not None and any(a >= lgs_altitude_los for a in alts):
            raise ValueError("lgs_altitude_los must exceed every layer altitude")
        if seeds is None:
            sd = (None,) * L
        else:
            sd = tuple(seeds)
            if len(sd) != L:
                raise ValueError("seeds must contain one value per layer")
        # far stencil points (no reference counterpart): 8 per far row, <= 4 rows
        if (not isinstance(far_points, (int, np.integer)) or far_points < 0
                or far_points % 8 or far_points > 32):
            raise ValueError("far_points must be 0, 8, 16, 24 or 32")
        return cls(grid, r0s, L0s, winds, alts, fov, int(stencil_rows), interp,
                   None if lgs_altitude_los is None else float(lgs_altitude_los), sd,
                   int(far_points))
