"""pyturb_amd -- MI355X (gfx950) backend of pyturb's phase-screen hot path.

Public names follow the reference package (src/pyturb/__init__.py) for the
accelerated path: ``PhaseScreen``, ``InfinitePhaseScreen``, ``Atmosphere``
plus the ``Layer`` profile value and helpers.  Every array result is a
float32 ROCm tensor computed by the HIP library ``libpyturb_hip.so``.
"""

from __future__ import annotations

__version__ = "0.1.0"

from .atmosphere import Atmosphere, ExtrudeBoilingPerformanceWarning, PeriodicWrapWarning
from .backend import resolve_device
from .flow import FourierFlowScreen
from .fourier import PhaseScreen
from .infinite import InfinitePhaseScreen, phase_covariance
from .io import load, save
from .profiles import (
    Layer,
    coherence_time,
    effective_wind_speed,
    get_profile,
    greenwood_frequency,
    isoplanatic_angle,
    list_profiles,
    mean_turbulence_height,
)
from .utils import (
    opd_to_phase,
    phase_to_opd,
    r0_at_wavelength,
    r0_from_seeing,
    seeing_from_r0,
    structure_function,
    to_numpy,
)

__all__ = [
    "Atmosphere",
    "ExtrudeBoilingPerformanceWarning",
    "FourierFlowScreen",
    "InfinitePhaseScreen",
    "Layer",
    "PeriodicWrapWarning",
    "PhaseScreen",
    "coherence_time",
    "effective_wind_speed",
    "get_profile",
    "greenwood_frequency",
    "isoplanatic_angle",
    "list_profiles",
    "load",
    "mean_turbulence_height",
    "opd_to_phase",
    "phase_covariance",
    "phase_to_opd",
    "r0_at_wavelength",
    "r0_from_seeing",
    "resolve_device",
    "save",
    "seeing_from_r0",
    "structure_function",
    "to_numpy",
]
