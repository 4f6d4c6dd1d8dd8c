"""Row-ring bookkeeping for the extruders (host integers only).

The reference keeps ``buffer[i]`` = virtual row ``base + i`` and, when the
buffer is full, copies the still-needed rows down (src/pyturb/ring.py:9-29).
Here the device ring never moves data: local row ``i`` lives at physical row
``(start + i) % cap``, and compaction advances ``start`` by the number of
rows the reference would have dropped.  ``base`` and ``fill`` evolve exactly
as in the reference, so readout coordinates (``travel - base``) and the
row-clamping window are the reference's.
"""

from __future__ import annotations

from math import floor
from typing import List, Tuple


class Ring:
    """``far`` (optional): distances behind the newest stencil row of far
    stencil rows; compaction then also keeps the farthest of them, and
    ``plan`` lists their physical rows."""

    __slots__ = ("cap", "m", "base", "fill", "start", "far")

    def __init__(self, cap: int, stencil_rows: int, base: int, fill: int,
                 far: Tuple[int, ...] = ()) -> None:
        self.cap = int(cap)
        self.m = int(stencil_rows)
        self.base = int(base)
        self.fill = int(fill)
        self.start = 0
        self.far = tuple(int(d) for d in far)

    def phys(self, local: int) -> int:
        return (self.start + local) % self.cap

    def compact(self, first_readable_row: float) -> None:
        """ring.py:9-29: keep rows from min(floor(first)-1, last m) on."""
        target = int(floor(first_readable_row)) - 1
        stencil = self.base + self.fill - self.m
        if self.far:  # the farthest far row of the next extrusion
            stencil = min(stencil, self.base + self.fill - 1 - self.far[-1])
        keep_from = max(1, min(target, stencil) - self.base)
        self.start = (self.start + keep_from) % self.cap
        self.base += keep_from
        self.fill -= keep_from

    def plan(self, top_virtual: int, first_readable_row: float) -> List[Tuple[int, Tuple[int, ...]]]:
        """Rows to extrude until virtual row ``top_virtual`` exists.

        Returns ``[(target_phys, (z_phys...)), ...]`` in order -- with far
        rows, ``(target_phys, (z_phys...), (far_phys...))`` -- and advances
        the bookkeeping (compacting when full, as the reference does before
        each extrusion).
        """
        rows = []
        while self.base + self.fill - 1 < top_virtual:
            if self.fill == self.cap:
                self.compact(first_readable_row)
            f = self.fill
            z = tuple(self.phys(f - self.m + k) for k in range(self.m))
            if self.far:
                if f - 1 - self.far[-1] < 0:
                    raise RuntimeError("far stencil row behind the ring's oldest row")
                rows.append((self.phys(f), z, tuple(self.phys(f - 1 - d) for d in self.far)))
            else:
                rows.append((self.phys(f), z))
            self.fill = f + 1
        return rows
