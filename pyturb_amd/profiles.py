"""Turbulence-profile inputs: the ``Layer`` value and a few named profiles.

Host-side model inputs only (the profile library itself is outside the
accelerated path).  ``Layer`` keeps the reference's conventions
(src/pyturb/profiles.py:55-85): altitude [m] at zenith, a relative Cn2
weight, wind speed [m/s] and wind direction in **degrees** measured from
axis 0 toward axis 1, and a per-layer outer scale.  The named profiles
carry the reference's published tables (altitudes, weights, winds), which
are data, not code.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Sequence, Tuple

import numpy as np


@dataclass
class Layer:
    altitude: float
    cn2_fraction: float
    wind_speed: float = 10.0
    wind_direction: float = 0.0
    L0: float = 25.0

    @property
    def wind_vector(self) -> Tuple[float, float]:
        th = np.deg2rad(self.wind_direction)
        return self.wind_speed * float(np.cos(th)), self.wind_speed * float(np.sin(th))


def _table(alt, frac, wind, direc, L0=25.0) -> List[Layer]:
    return [Layer(float(a), float(f), float(w), float(d), float(L0))
            for a, f, w, d in zip(alt, frac, wind, direc)]


_PROFILES: Dict[str, callable] = {
    "single-layer": lambda: _table([0.0], [1.0], [10.0], [0.0]),
    "two-layer": lambda: _table([0.0, 10000.0], [0.7, 0.3], [8.0, 30.0], [0.0, 90.0]),
    "paranal-median": lambda: _table(
        [30, 140, 281, 562, 1125, 2250, 4500, 9000, 18000],
        [0.242, 0.12, 0.098, 0.059, 0.043, 0.061, 0.13, 0.13, 0.117],
        [5.5, 6.6, 6.7, 8.0, 9.9, 15.0, 25.0, 32.0, 14.0],
        [0, 20, 40, 60, 90, 120, 150, 180, 210]),
    "mauna-kea": lambda: _table(
        [500, 1000, 2000, 4000, 8000, 16000],
        [0.2283, 0.0883, 0.0666, 0.1458, 0.3350, 0.1350],
        [10.0] * 6, [0, 30, 60, 90, 120, 150], L0=10.0),
    "keck": lambda: _table(
        [0, 2100, 4100, 6500, 9000, 12000, 14800],
        [0.369, 0.219, 0.127, 0.101, 0.046, 0.111, 0.027],
        [6.7, 13.9, 20.8, 29.0, 29.0, 29.0, 29.0],
        [0, 30, 60, 90, 120, 150, 180], L0=20.0),
}


def list_profiles() -> List[str]:
    return sorted(_PROFILES)


def get_profile(name: str) -> List[Layer]:
    key = str(name).lower()
    if key not in _PROFILES:
        raise ValueError(f"Unknown profile {name!r}. Available: {list_profiles()}.")
    return _PROFILES[key]()


def _fractions(layers: Sequence[Layer]) -> np.ndarray:
    frac = np.asarray([ly.cn2_fraction for ly in layers], dtype=np.float64)
    if np.any(frac < 0):
        raise ValueError("cn2_fraction values must be non-negative")
    total = frac.sum()
    if total <= 0:
        raise ValueError("cn2_fraction values must sum to a positive number")
    return frac / total


def _weighted_53(values, layers):
    f = _fractions(layers)
    return float(np.sum(f * np.asarray(values, dtype=np.float64) ** (5.0 / 3.0)) ** 0.6)


def mean_turbulence_height(layers) -> float:
    return _weighted_53([ly.altitude for ly in layers], layers)


def effective_wind_speed(layers) -> float:
    return _weighted_53([ly.wind_speed for ly in layers], layers)


def isoplanatic_angle(layers, r0: float) -> float:
    h = mean_turbulence_height(layers)
    return np.inf if h == 0 else 0.314 * r0 / h


def coherence_time(layers, r0: float) -> float:
    v = effective_wind_speed(layers)
    return np.inf if v == 0 else 0.314 * r0 / v


def greenwood_frequency(layers, r0: float) -> float:
    return 0.134 / coherence_time(layers, r0)
