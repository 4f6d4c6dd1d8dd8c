"""Device dispatch: the ``device`` strings this backend accepts.

The reference resolves ``device`` to NumPy or CuPy (src/pyturb/backend.py:
16-45).  This package is the MI355X backend: ``"mi355x"`` (aliases
``"hip"``, ``"rocm"``, and ``"gpu"``/``"cuda"`` so that GPU code written for
the reference keeps working) selects the HIP library on ROCm device 0, or
``"mi355x:1"`` etc. for another device.  The reference's CPU names are
rejected with a ValueError: this package has no CPU compute path.
"""

from __future__ import annotations

from .utils import to_numpy  # noqa: F401  (re-exported like the reference)

HIP_NAMES = frozenset({"mi355x", "hip", "rocm", "gpu", "cuda"})
CPU_NAMES = frozenset({"cpu", "numpy"})


def resolve_device(device: str) -> int:
    """Validate a device string; return the ROCm device index."""
    name = str(device).lower()
    base, _, idx = name.partition(":")
    if base in HIP_NAMES:
        if idx and not idx.isdigit():
            raise ValueError(f"Unknown device {device!r}.")
        return int(idx) if idx else 0
    if base in CPU_NAMES:
        raise ValueError(
            f"device={device!r}: pyturb_amd is the MI355X backend and has no CPU "
            "compute path; use device='mi355x' (or the reference pyturb for NumPy)."
        )
    raise ValueError(f"Unknown device {device!r}. Expected one of {sorted(HIP_NAMES)}.")
