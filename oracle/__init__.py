"""CPU oracle for the phase-screen hot path (test infrastructure only).

This package restates, in plain FP64 NumPy, the arithmetic of the reference
pyturb (jacotay7/pyturb) on the path that the HIP library accelerates:
PSD filters and subharmonic bases, batched screen generation, the L0-grouped
Monte-Carlo sample, the spectral frozen-flow frame (on- and off-axis), the
von Karman covariance and extrusion matrices, the row-ring extruder and its
multi-layer readout.

It is the checker, never the product: only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` may import it.  The product
(``pyturb_amd``) never calls into it and fails loudly when its HIP library is
missing.
"""

from .pyturb_oracle import *  # noqa: F401,F403
