"""Multi-layer extrusion engine on the MI355X (src/pyturb/extrude.py:307-915).

Geometry, buffer sizing and the ring bookkeeping follow the reference
exactly (host integers and float64); the data path is on the device:

* one (layers, cap, W) float32 ring-buffer array in HBM; a compaction is an
  index shift (:class:`pyturb_amd.ring.Ring`), never a copy;
* ``set_time`` plans every layer's new rows on the host, then one
  launch (pt_extrude_step) computes all their innovation terms
  ``B_unit beta`` and runs the sequential ``A z`` recurrence round by round
  with A held in registers and a grid barrier between rounds; shapes it does
  not cover run the innovation GEMM and one launch per round instead (the
  same rows, bit for bit);
* ``integrate``/``integrate_dirs`` is one fused readout launch for all
  layers and up to 8 directions, each direction with its own per-layer
  magnification (so NGS directions and an LGS cone share a launch).

Innovations default to the reference's host stream (each layer's generator,
after its seed screen); ``innovations=callable(layer, k) -> (k, W) device
tensor`` injects them instead.

``far_points`` (0, 8, 16, 24 or 32; no reference counterpart) adds single
pixels further back to the m full stencil rows: 8 per far row, far rows at
(m - 1) + round(H^(j/R)) rows behind the newest stencil row (H = W // 2),
columns (2 i + 1) W / 16.  They condition each new row on the large scales
that two rows cannot see.  The ring keeps the farthest of them; with
far_points = 0 everything is the reference's algorithm, bit for bit the
output of a build without the option.
"""

from __future__ import annotations

import ctypes as _ct
import math
from typing import Any, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _hip
from .backend import resolve_device
from .config import ExtrusionConfig
from .fourier import PhaseScreen
from .infinite import (StepState, device_extrusion, device_seed_factor, extrude_rounds,
                       extrude_step_or_rounds, far_rows_back,
                       readout, row_pitch)
from .ring import Ring

__all__ = ["ExtrudedAtmosphere", "along_extent_and_span"]


def along_extent_and_span(n, vx, vy, magnification, fov_margin_pix) -> Tuple[float, int]:
    """Along-wind half-extent and ring span of one layer (extrude.py:222-247)."""
    speed = math.hypot(vx, vy)
    c = vx / speed if speed > 0 else 1.0
    s = vy / speed if speed > 0 else 0.0
    ext = (n - 1) / 2.0 * float(magnification) * (abs(c) + abs(s))
    span = int(math.ceil(ext + fov_margin_pix)) - int(math.floor(-ext - fov_margin_pix)) + 3
    return ext, span


class _Layer:
    __slots__ = ("speed", "cos", "sin", "alt_pix", "along_min", "along_max", "fov", "mag",
                 "ring", "travel", "rng", "bscale", "A", "B", "L0", "tau_boil")


class ExtrudedAtmosphere:
    """Frozen-flow multi-layer extruder (non-periodic, unbounded)."""

    def __init__(self, n: int, pixel_scale: float, layer_r0: Sequence[float],
                 layer_L0: Sequence[float], layer_wind: Sequence[Tuple[float, float]],
                 layer_altitude_los: Sequence[float],
                 field_of_view_pix: Union[float, Sequence[float]] = 0.0,
                 stencil_rows: int = 2, interp: str = "cubic",
                 lgs_altitude_los: Optional[float] = None, device: str = "mi355x",
                 dtype: Union[str, np.dtype] = "float32", seeds=None, innovations=None,
                 seed_screens=None, tau_boil: Optional[Sequence[float]] = None,
                 boil_seed: Optional[int] = None, far_points: int = 0) -> None:
        config = ExtrusionConfig.create(n, pixel_scale, layer_r0, layer_L0, layer_wind,
                                        layer_altitude_los, field_of_view_pix, stencil_rows,
                                        interp, lgs_altitude_los, device, dtype, seeds,
                                        far_points)
        if config.far_points and tau_boil is not None and any(
                np.isfinite(float(v)) for v in tau_boil):
            raise ValueError("far_points cannot be combined with boiling (tau_boil): fresh "
                             "boiling windows are extruded from full stencil rows only")
        self.config = config
        self.n = config.grid.n
        self.dx = config.grid.pixel_scale
        self.device = config.grid.device
        self.dtype = config.grid.dtype
        self.interp = config.interp
        self.m = config.stencil_rows
        self._dev = _hip.device(resolve_device(self.device))
        L = len(config.layer_r0)
        self.n_layers = L
        if config.lgs_altitude_los is not None:
            mags = [1.0 - a / config.lgs_altitude_los for a in config.layer_altitude_los]
        else:
            mags = [1.0] * L
        fovs = list(config.field_of_view_pix)
        ext_span = [along_extent_and_span(self.n, vx, vy, mg, fv)
                    for (vx, vy), mg, fv in zip(config.layer_wind, mags, fovs)]
        self.width = int(np.ceil(2.0 * max(e for e, _ in ext_span) + 2.0 * max(fovs))) + 4
        self.far_points = config.far_points
        far = far_rows_back(self.width, self.m, self.far_points) if self.far_points else ()
        self.capacity = max(s for _, s in ext_span) + self.n + self.m + 32 + (far[-1] if far else 0)
        W, cap = self.width, self.capacity
        t = _hip.torch()
        self._buf = t.zeros((L, cap, row_pitch(W)), dtype=t.float32, device=self._dev)
        self._innovations = innovations
        self.layers: List[_Layer] = []
        g = (np.arange(self.n, dtype=np.float64) - (self.n - 1) / 2.0)
        for i in range(L):
            r0 = config.layer_r0[i]
            L0 = config.layer_L0[i]
            vx, vy = config.layer_wind[i]
            ly = _Layer()
            ly.speed = math.hypot(vx, vy)
            ly.cos = vx / ly.speed if ly.speed > 0 else 1.0
            ly.sin = vy / ly.speed if ly.speed > 0 else 0.0
            ly.alt_pix = config.layer_altitude_los[i] / self.dx
            ly.mag = mags[i]
            ly.fov = fovs[i]
            gm = g * float(mags[i])
            along = gm[:, None] * ly.cos + gm[None, :] * ly.sin
            ly.along_min = float(along.min())
            ly.along_max = float(along.max())
            span = (int(np.ceil(ly.along_max + ly.fov)) - int(np.floor(ly.along_min - ly.fov)) + 3)
            ly.A, B = device_extrusion(W, self.m, self.dx, L0, 1.0, self._dev,
                                       far_points=self.far_points)
            ly.B = B
            ly.bscale = float(r0) ** (-5.0 / 6.0)
            ly.L0 = float(L0)
            ly.tau_boil = float("inf") if tau_boil is None else float(tau_boil[i])
            ly.rng = np.random.default_rng(config.seeds[i])
            if seed_screens is not None:
                init = seed_screens[i]
            else:
                gen = PhaseScreen(W, self.dx, float(r0), float(L0), seed=ly.rng,
                                  device=self.device, dtype=self.dtype.name)
                init = gen.generate()
            # far rows need history behind the first new row: seed `extra`
            # more rows of the seed screen below the reference's window
            extra = max(0, far[-1] + 1 - span) if far else 0
            self._buf[i, :span + extra, :W] = init[:span + extra].float()
            ly.ring = Ring(cap, self.m, int(np.floor(ly.along_min - ly.fov)) - 2 - extra,
                           span + extra, far)
            ly.travel = 0.0
            self.layers.append(ly)
        self._wind_cs = [(ly.cos, ly.sin) for ly in self.layers]
        self._As = [ly.A for ly in self.layers]
        self._Bs = [ly.B for ly in self.layers]
        self._bscales = [ly.bscale for ly in self.layers]
        # readout arguments kept as arrays (per-layer constants; ring state and
        # direction geometry refreshed per call)
        self._alt_pix = np.array([ly.alt_pix for ly in self.layers])
        self._cos = np.array([ly.cos for ly in self.layers])
        self._sin = np.array([ly.sin for ly in self.layers])
        self._mag = np.array([ly.mag for ly in self.layers])
        self._dir_cache = None
        self._wind_cs_np = np.array(self._wind_cs, dtype=np.float64)
        # each set_time is one pt_extrude_step launch where its shape allows
        # (pt_extrude_step), else the innovation GEMM and a launch per round
        self._step = StepState(self._dev, far)
        # boiling (extrude.py:586-654): one stream for every layer; the device
        # draws it from Philox keyed by boil_seed, counted by (step, layer, kind)
        self.boiling = any(np.isfinite(ly.tau_boil) for ly in self.layers)
        self._boil_seed = 0 if boil_seed is None else int(boil_seed)
        self._boil_index = 0
        # injected boiling draws (tests): (step, layer, window rows h) ->
        # (seed (m W,), rows (h - m, W)) host arrays
        self._boil_draws = None

    # ------------------------------------------------------------------
    def _betas(self, counts):
        if self._innovations is not None:
            return [self._innovations(i, k) if k else None for i, k in enumerate(counts)]
        # seeded draws: every layer's rows in one pitched array (a pitch of a
        # multiple of 4 floats, as the one-launch step reads them), one copy
        W, pitch = self.width, (self.width + 3) // 4 * 4
        host = np.zeros((sum(counts), pitch), np.float32)
        spans, off = [], 0
        for i, k in enumerate(counts):
            if k:
                host[off:off + k, :W] = self.layers[i].rng.standard_normal((k, W), dtype=self.dtype)
            spans.append((off, k))
            off += k
        dev = _hip.to_device(host, None, self._dev)
        return [dev[o:o + k, :W] if k else None for o, k in spans]

    def set_time(self, t: float) -> None:
        """Advance every layer to time ``t`` [s] (monotonic)."""
        spans = []
        dx = self.dx
        for ly in self.layers:
            travel = float(ly.speed * t / dx)
            if travel < ly.travel:
                raise ValueError("wind travel is monotonic; travel cannot decrease")
            ly.travel = travel
            top = math.ceil(ly.along_max + ly.fov + travel) + 2
            spans.append(ly.ring.extend(top, ly.along_min - ly.fov + travel))
        counts = [k for _, k in spans]
        if any(counts):
            betas = self._betas(counts)
            layers = self.layers
            extrude_step_or_rounds(self._As, self._Bs, self._buf, self.capacity, self.width, self.m,
                                   spans, lambda l: layers[l].ring.rows(*spans[l]), betas,
                                   self._bscales, far_points=self.far_points, step=self._step)

    def boil_step(self, dt: float) -> None:
        """One AR(1) blend of every boiling layer's readable window toward a
        fresh independent window (extrude.py:472-496, 660-752).

        A fresh window of h rows is S w on its first m rows (S S^T = C_zz,
        scaled by r0^(-5/6)) and the layer's own recurrence A z + B beta on
        the rest: the seed product and the recurrence run on the matrix
        cores (pt_innovation_gemm, pt_extrude_rows), the blend
        buf = a buf + sqrt(1 - a^2) fresh in pt_rows_blend.  A no-op unless
        some layer has a finite tau_boil.
        """
        if not self.boiling:
            return
        t = _hip.torch()
        lib = _hip.lib()
        st = _hip.stream_ptr()
        W, m, ld = self.width, self.m, int(self._buf.shape[-1])
        regions = []
        for l, ly in enumerate(self.layers):
            if not np.isfinite(ly.tau_boil):
                continue
            a = math.exp(-dt / ly.tau_boil)
            b = math.sqrt(max(0.0, 1.0 - a * a))
            lo = max(0, int(math.floor(ly.along_min + ly.travel)) - ly.ring.base - 3)
            hi = ly.ring.fill
            if hi > lo:
                regions.append((l, lo, hi, a, b))
        if not regions:
            return
        H = max(hi - lo for _, lo, hi, _, _ in regions)
        # zero padding past W: the recurrence's 16-byte loads read it
        fresh = t.zeros((len(regions), H, ld), dtype=t.float32, device=self._dev)
        As, Bs, plans, betas, scales = [], [], [], [], []
        mw = m * W
        for j, (l, lo, hi, a, b) in enumerate(regions):
            ly, h = self.layers[l], hi - lo
            S = device_seed_factor(W, m, self.dx, ly.L0, self._dev)
            ldS = int(S.shape[-1])
            inj = None if self._boil_draws is None else self._boil_draws(self._boil_index, l, h)
            if inj is None:
                w = t.empty((mw,), dtype=t.float32, device=self._dev)
                _hip.check(lib.pt_normals(_hip.ptr(w), mw, self._boil_seed,
                                          self._boil_stream(l, 0), ly.bscale, st), "pt_normals")
                beta = t.empty((max(h - m, 0), W), dtype=t.float32, device=self._dev)
                if h > m:
                    _hip.check(lib.pt_normals(_hip.ptr(beta), (h - m) * W, self._boil_seed,
                                              self._boil_stream(l, 1), 1.0, st), "pt_normals")
            else:
                w = _hip.to_device(np.asarray(inj[0], np.float64) * ly.bscale, np.float32, self._dev)
                beta = _hip.to_device(np.asarray(inj[1], np.float64).reshape(-1, W)[:max(h - m, 0)],
                                      np.float32, self._dev)
            seed = t.empty((1, ldS), dtype=t.float32, device=self._dev)
            ptrs = (_ct.c_void_p * 1)(_hip.ptr(w))
            _hip.check(lib.pt_innovation_gemm(ptrs, _hip.ptr(S), 1, mw, ldS, _hip.ptr(seed), st),
                       "pt_innovation_gemm")
            _hip.check(lib.pt_rows_blend(_hip.ptr(fresh[j]), ld, 0, H, _hip.ptr(seed), W,
                                         min(m, h), W, 0.0, 1.0, st), "pt_rows_blend")
            A, B = device_extrusion(W, m, self.dx, ly.L0, 1.0, self._dev)
            As.append(A)
            Bs.append(B)
            plans.append([(k, tuple(range(k - m, k))) for k in range(m, h)])
            betas.append(beta)
            scales.append(ly.bscale)
        extrude_rounds(As, Bs, fresh, H, W, m, plans, betas, scales)
        for j, (l, lo, hi, a, b) in enumerate(regions):
            ring = self.layers[l].ring
            _hip.check(lib.pt_rows_blend(_hip.ptr(self._buf[l]), ld, ring.phys(lo), self.capacity,
                                         _hip.ptr(fresh[j]), ld, hi - lo, W, float(a), float(b), st),
                       "pt_rows_blend")
        self._boil_index += 1

    def _boil_stream(self, layer: int, kind: int) -> int:
        return (self._boil_index * 64 + layer) * 2 + kind

    def dir_geometry(self, thx, thy, mags=None) -> np.ndarray:
        """(D, L, 3) readout geometry: along/perp shifts and magnification.

        The direction terms depend only on the directions, so the last set is
        kept; each call adds the layers' current travel - base."""
        thx = np.atleast_1d(np.asarray(thx, dtype=np.float64))
        thy = np.atleast_1d(np.asarray(thy, dtype=np.float64))
        key = (thx.tobytes(), thy.tobytes(),
               None if mags is None else np.asarray(mags, dtype=np.float64).tobytes())
        c = self._dir_cache
        if c is None or c[0] != key:
            dxp = self._alt_pix[None, :] * thx[:, None]
            dyp = self._alt_pix[None, :] * thy[:, None]
            base = np.empty((thx.size, self.n_layers, 3))
            base[:, :, 0] = dxp * self._cos + dyp * self._sin
            base[:, :, 1] = -dxp * self._sin + dyp * self._cos
            base[:, :, 2] = self._mag
            if mags is not None:
                base[:, :, 2] = np.broadcast_to(np.asarray(mags, dtype=np.float64), base.shape[:2])
            c = self._dir_cache = (key, base)
        geom = c[1].copy()
        geom[:, :, 0] = np.array([ly.travel - ly.ring.base for ly in self.layers]) + c[1][:, :, 0]
        return geom

    def integrate_dirs(self, thx, thy, scale: float = 1.0, mags=None, out=None) -> Any:
        """Summed phase for D directions (tangents) -> (D, n, n) float32."""
        geom = self.dir_geometry(thx, thy, mags)
        return readout(self._buf, self.capacity, self.width, [ly.ring for ly in self.layers],
                       self._wind_cs_np, geom, self.n, self.interp, scale=scale, out=out)

    def integrate(self, thx: float = 0.0, thy: float = 0.0) -> Any:
        out = self.integrate_dirs([thx], [thy])[0]
        return out.double() if self.dtype == np.float64 else out
