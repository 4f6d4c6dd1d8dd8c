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
        first, count = self.extend(top_virtual, first_readable_row)
        return self.rows(first, count)

    def extend(self, top_virtual: int, first_readable_row: float) -> Tuple[int, int]:
        """The bookkeeping of :meth:`plan` without the per-row lists.

        Returns ``(first_target_phys, count)``: the new rows are physical
        rows ``(first + i) % cap``, i < count.  They are consecutive modulo
        ``cap`` because a compaction moves ``start`` and ``fill`` by the same
        amount, which leaves the physical row of the next new row in place.
        """
        need = int(top_virtual) - (self.base + self.fill - 1)
        first = (self.start + self.fill) % self.cap
        count = 0
        while need > 0:
            if self.fill == self.cap:
                self.compact(first_readable_row)
            if self.far and self.fill - 1 - self.far[-1] < 0:
                raise RuntimeError("far stencil row behind the ring's oldest row")
            k = min(need, self.cap - self.fill)
            self.fill += k
            count += k
            need -= k
        return first, count

    def rows(self, first: int, count: int) -> List[Tuple[int, Tuple[int, ...]]]:
        """Per-row ``(target_phys, (z_phys...)[, (far_phys...)])`` of ``count``
        consecutive new rows from physical row ``first`` (what :meth:`plan`
        returns): a row's stencil rows are the m physical rows before it and
        its far rows lie ``d + 1`` rows behind it, modulo ``cap``."""
        cap, m = self.cap, self.m
        out = []
        for i in range(count):
            t = (first + i) % cap
            z = tuple((t - m + k) % cap for k in range(m))
            if self.far:
                out.append((t, z, tuple((t - 1 - d) % cap for d in self.far)))
            else:
                out.append((t, z))
        return out
