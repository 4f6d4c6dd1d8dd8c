"""Infinite (extruded) screens: covariance, extrusion matrices and the
single-layer ``InfinitePhaseScreen`` (src/pyturb/infinite.py:54-395).

The von Karman covariance and the ``A``/``B`` matrices are one-time host
float64 work with the same numpy/scipy routines as the reference
(``cho_factor``/``cho_solve`` with a least-squares fallback, ``eigh``), so
``B``'s eigenvector conventions match.  The ring buffer, the row recurrence
``A z + B beta`` and the sub-pixel row interpolation run on the device
(``pt_innovation_gemm``, ``pt_extrude_rows``, ``pt_readout``).
"""

from __future__ import annotations

from typing import Any, Dict, Tuple, Union

import numpy as np
from scipy import linalg
from scipy.special import gamma, kv

from . import _hip
from .backend import resolve_device
from .config import ScreenConfig
from .fourier import PhaseScreen
from .ring import Ring

__all__ = ["InfinitePhaseScreen", "phase_covariance", "build_extrusion"]

_INTERP_CODE = {"linear": 0, "cubic": 1, "lanczos": 2}


def phase_covariance(r, r0: float, L0: float) -> np.ndarray:
    """Von Karman phase covariance C(r) [rad^2] (infinite.py:88-117).

    ``C(0) = 0.0863 (L0/r0)^(5/3)``; needs a finite ``L0``.
    """
    if not np.isfinite(L0):
        raise ValueError("phase_covariance requires a finite outer scale L0")
    r = np.asarray(r, dtype=np.float64)
    pref = ((L0 / r0) ** (5.0 / 3.0) * 2.0 ** (-5.0 / 6.0) * gamma(11.0 / 6.0)
            / np.pi ** (8.0 / 3.0) * (4.8 * gamma(1.2)) ** (5.0 / 6.0))
    x = 2.0 * np.pi * r / L0
    core = np.full_like(x, 2.0 ** (-1.0 / 6.0) * gamma(5.0 / 6.0))
    pos = x > 0
    core[pos] = x[pos] ** (5.0 / 6.0) * kv(5.0 / 6.0, x[pos])
    return pref * core


def _spd_solve(spd, rhs):
    try:
        return linalg.cho_solve(linalg.cho_factor(spd), rhs)
    except np.linalg.LinAlgError:
        return np.linalg.lstsq(spd, rhs, rcond=None)[0]


FAR_COLUMNS = 8  # far points per far row


def far_rows_back(width: int, stencil_rows: int, far_points: int) -> Tuple[int, ...]:
    """Distances (in rows, behind the newest stencil row) of the far rows.

    ``far_points`` = 8 R single pixels in R rows: row j = 1..R lies
    ``(m - 1) + round(H^(j/R))`` rows behind the newest stencil row, with the
    horizon H = width // 2 (strictly increasing), so the first far row is just
    behind the full-row stencil and the last half a width back.  Each far row
    holds the pixels at columns ``far_columns(width)``.
    """
    R = int(far_points) // FAR_COLUMNS
    H = max(2, int(width) // 2)
    m = int(stencil_rows)
    out, prev = [], m - 1
    for j in range(1, R + 1):
        d = max(prev + 1, (m - 1) + int(round(H ** (j / R))))
        out.append(d)
        prev = d
    return tuple(out)


def far_columns(width: int) -> Tuple[int, ...]:
    """Columns of the far points of a row: (2 i + 1) width // 16, i < 8."""
    return tuple(((2 * i + 1) * int(width)) // (2 * FAR_COLUMNS) for i in range(FAR_COLUMNS))


def stencil_points(width: int, stencil_rows: int, far_points: int = 0):
    """(x, y) of the stencil: m full rows at y = 0..m-1 in row-major order,
    then the far points, far row by far row (nearest first), 8 columns each;
    the new row is at y = m."""
    m, W = int(stencil_rows), int(width)
    zy, zx = np.divmod(np.arange(m * W, dtype=np.float64), W)
    if far_points:
        cols = np.array(far_columns(W), dtype=np.float64)
        ys = [(m - 1) - d for d in far_rows_back(W, m, far_points)]
        zx = np.concatenate([zx, np.tile(cols, len(ys))])
        zy = np.concatenate([zy, np.repeat(np.array(ys, dtype=np.float64), len(cols))])
    return zx, zy


def build_extrusion(width: int, stencil_rows: int, pixel_scale: float, L0: float,
                    r0: float = 1.0, far_points: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """``(A, B)`` float64 for extruding one ``width``-pixel row.

    Stencil rows at y = 0..m-1 (plus ``far_points`` single pixels further
    back, see ``stencil_points``), new row at y = m; ``A = C_xz C_zz^-1``,
    ``B = V sqrt(max(lambda, 0))`` of the symmetrised conditional covariance
    (extrude.py:250-304, infinite.py:231-269; far points have no reference
    counterpart and leave the construction unchanged).
    """
    m, W = int(stencil_rows), int(width)
    zx, zy = stencil_points(W, m, far_points)
    nx = np.arange(W, dtype=np.float64)
    ny = np.full(W, float(m))

    def cov(ax, ay, bx, by):
        sep = np.hypot(ax[:, None] - bx[None, :], ay[:, None] - by[None, :])
        return phase_covariance(sep * pixel_scale, r0, L0)

    c_zz = cov(zx, zy, zx, zy)
    c_xz = cov(nx, ny, zx, zy)
    c_xx = cov(nx, ny, nx, ny)
    a_mat = _spd_solve(c_zz, c_xz.T).T
    res = c_xx - a_mat @ c_xz.T
    res = 0.5 * (res + res.T)
    lam, vec = np.linalg.eigh(res)
    return a_mat, vec * np.sqrt(np.clip(lam, 0.0, None))


_MATRIX_CACHE: Dict[tuple, Tuple[Any, Any]] = {}


def row_pitch(width: int) -> int:
    """Device row pitch in floats: a multiple of 64 (256-byte rows)."""
    return (int(width) + 63) // 64 * 64


def far_pitch(far_points: int) -> int:
    """Columns of A's far-point segment: far_points rounded up to 16."""
    return (int(far_points) + 15) // 16 * 16


def device_extrusion(width, stencil_rows, pixel_scale, L0, r0, dev, far_points=0):
    """(A, B) as float32 device tensors with pitched, zero-padded rows.

    A is (W, m*ld + far_pitch(F)) with stencil row s in columns
    [s*ld, s*ld + W) and the F far points in [m*ld, m*ld + F); B is
    (W, ld).  Cached per geometry and device.
    """
    key = (int(width), int(stencil_rows), float(pixel_scale), round(float(L0), 9),
           float(r0), str(dev), int(far_points))
    hit = _MATRIX_CACHE.get(key)
    if hit is None:
        W, m, F = int(width), int(stencil_rows), int(far_points)
        ld = row_pitch(W)
        a, b = build_extrusion(W, m, pixel_scale, L0, r0, far_points=F)
        ap = np.zeros((W, m * ld + far_pitch(F)), dtype=np.float32)
        for s_ in range(m):
            ap[:, s_ * ld:s_ * ld + W] = a[:, s_ * W:(s_ + 1) * W]
        if F:
            ap[:, m * ld:m * ld + F] = a[:, m * W:]
        bp = np.zeros((W, ld), dtype=np.float32)
        bp[:, :W] = b
        hit = (_hip.to_device(ap, None, dev), _hip.to_device(bp, None, dev))
        _MATRIX_CACHE[key] = hit
    return hit


_SEED_CACHE: Dict[tuple, Any] = {}


def device_seed_factor(width, stencil_rows, pixel_scale, L0, dev):
    """S_unit (m W, round_up(m W, 64)) float32 on the device, S S^T = C_zz at
    r0 = 1: Cholesky of the stencil-block covariance, else its symmetric
    square root (extrude.py:284-303).  Cached per geometry and device."""
    key = (int(width), int(stencil_rows), float(pixel_scale), round(float(L0), 9), str(dev))
    hit = _SEED_CACHE.get(key)
    if hit is None:
        m, W = int(stencil_rows), int(width)
        zy, zx = np.divmod(np.arange(m * W, dtype=np.float64), W)
        sep = np.hypot(zx[:, None] - zx[None, :], zy[:, None] - zy[None, :])
        c_zz = phase_covariance(sep * pixel_scale, 1.0, L0)
        c_zz = 0.5 * (c_zz + c_zz.T)
        try:
            s_unit = np.linalg.cholesky(c_zz)
        except np.linalg.LinAlgError:
            w, v = np.linalg.eigh(c_zz)
            s_unit = v * np.sqrt(np.clip(w, 0.0, None))
        sp = np.zeros((m * W, row_pitch(m * W)), dtype=np.float32)
        sp[:, :m * W] = s_unit
        hit = _hip.to_device(sp, None, dev)
        _SEED_CACHE[key] = hit
    return hit


class StepState:
    """Device state of the one-launch extrusion step (pt_extrude_step): the
    grid barrier's words (zeroed once, then carried from call to call) and a
    pinned status word the kernel sets if a barrier wait times out."""

    SPIN_LIMIT = 1 << 22  # barrier polls before a workgroup gives up (seconds)
    SYNC_BYTES = 2048     # PT_EXTRUDE_SYNC_BYTES

    LMAX = 16             # layers per launch

    def __init__(self, dev, far=()) -> None:
        t = _hip.torch()
        self.sync = t.zeros((self.SYNC_BYTES // 8,), dtype=t.int64, device=dev)
        self.status = t.zeros((1,), dtype=t.int32, pin_memory=True)
        self._status = self.status.numpy()  # host view of the pinned word
        self.far = tuple(int(d) for d in far)
        R = max(1, len(self.far))
        self.far_arr = (ctypes_int * R)(*(list(self.far) or [0]))
        self.enabled = True
        self.spin_limit = self.SPIN_LIMIT
        self.launches = 0
        # launch arguments, filled in place per call
        L = self.LMAX
        self.first = (ctypes_int * L)()
        self.rows = (ctypes_int * L)()
        self.ptrs = (_ct.c_void_p * L)()
        self.pitch = (ctypes_ll * L)()
        self.scale = (ctypes_float * L)()

    def check(self) -> None:
        """Raise if an earlier step's barrier timed out (the rows it extruded
        are then incomplete); the barrier words are reset first."""
        if self._status[0] != 0:
            self.status.zero_()
            self.sync.zero_()
            raise _hip.HipError("pt_extrude_step: a grid barrier timed out or was aborted; the extrusion state "
                                "is incomplete")


def _extrude_step(step, A, B, buf, cap, width, m, spans, betas, bscales, far_points):
    """One pt_extrude_step launch for every layer's new rows, or False when
    the call shape is outside what the one-launch step implements (then the
    caller runs the round-by-round path).

    ``spans[l]`` = (first_target_phys, count): layer l's new rows are the
    physical rows (first + i) % cap, i < count."""
    t = _hip.torch()
    L = len(spans)
    if L > StepState.LMAX or m > 2 or any(a is not A[0] for a in A) or any(b is not B[0] for b in B):
        return False
    total = sum(k for _, k in spans)
    if total == 0:
        return True
    if total > 256:
        return False
    f32 = t.float32
    dev = buf.get_device()
    keep = []
    for l in range(L):
        first, k = spans[l]
        step.first[l] = first
        step.rows[l] = k
        step.scale[l] = bscales[l]
        if not k:
            step.ptrs[l] = None
            step.pitch[l] = width
            continue
        beta = betas[l]
        st = beta.stride()
        if beta.dtype is not f32 or beta.get_device() != dev or beta.dim() != 2 \
                or beta.shape[1] != width or st[1] != 1:
            beta = beta.to(device=buf.device, dtype=f32).reshape(-1, width).contiguous()
            st = beta.stride()
        if beta.shape[0] != k:
            raise ValueError(f"layer {l}: {beta.shape[0]} innovation rows for {k} new rows")
        p = beta.data_ptr()
        if p % 16 or st[0] % 4:
            # the kernel reads rows as float4: a 16-byte aligned copy with a
            # pitch of a multiple of 4 (callers that want no copy supply rows
            # laid out so, e.g. views of a (k, 4 ceil(W / 4)) tensor)
            padded = t.zeros((k, (width + 3) // 4 * 4), dtype=f32, device=buf.device)
            padded[:, :width] = beta
            beta = padded
            p, st = beta.data_ptr(), beta.stride()
        keep.append(beta)
        step.ptrs[l] = p
        step.pitch[l] = st[0]
    step.check()
    rc = _hip.lib().pt_extrude_step(_hip.ptr(A[0]), _hip.ptr(B[0]), width, int(buf.shape[-1]), m,
                                    int(far_points), step.far_arr, _hip.ptr(buf), cap, L, step.first,
                                    step.rows, step.ptrs, step.pitch, step.scale, _hip.ptr(step.sync),
                                    step.status.data_ptr(), int(step.spin_limit), _hip.stream_ptr())
    if rc == _hip.PT_ERR_UNSUPPORTED:
        _hip.lib().pt_last_error()
        return False
    _hip.check(rc, "pt_extrude_step")
    step.launches += 1
    return True


def extrude_step_or_rounds(A, B, buf, cap, width, m, spans, rows_of, betas, bscales, far_points=0,
                           step=None):
    """Layer l's new rows are ``spans[l]`` = (first_target_phys, count):
    one pt_extrude_step launch when ``step`` is enabled and the shape allows,
    else the per-row plans ``rows_of(l)`` through :func:`extrude_rounds`."""
    As = A if isinstance(A, (list, tuple)) else [A] * len(spans)
    Bs = B if isinstance(B, (list, tuple)) else [B] * len(spans)
    if step is not None and step.enabled and _extrude_step(step, As, Bs, buf, cap, width, m, spans,
                                                           betas, bscales, far_points):
        return
    plans = [rows_of(l) if spans[l][1] else [] for l in range(len(spans))]
    extrude_rounds(As, Bs, buf, cap, width, m, plans, betas, bscales, far_points=far_points)


def extrude_rounds(A, B, buf, cap, width, m, plans, betas, bscales, far_points=0, step=None):
    """Run the planned rows of several layers on the device.

    With ``step`` (a StepState) the whole step is one launch
    (pt_extrude_step) when its shape allows; otherwise, or without it, the
    innovation GEMM and one launch per round.

    ``plans[l]``: list of (target_phys, z_phys) -- or (target_phys, z_phys,
    far_phys) with ``far_points`` > 0 -- for layer l, in order;
    ``betas[l]``: (k_l, width) float32 device innovations; ``bscales[l]``:
    r0^(-5/6) factor.  ``A``/``B`` may be per-layer lists (outer-scale
    groups) or shared tensors.
    """
    t = _hip.torch()
    L = len(plans)
    ld = int(buf.shape[-1])
    lib = _hip.lib()
    st = _hip.stream_ptr()
    As = A if isinstance(A, (list, tuple)) else [A] * L
    Bs = B if isinstance(B, (list, tuple)) else [B] * L
    if step is not None and step.enabled:
        spans = []
        for p in plans:
            if any(p[i][0] != (p[0][0] + i) % cap for i in range(len(p))):
                break  # rows of a step are consecutive modulo cap
            spans.append((p[0][0] if p else 0, len(p)))
        else:
            if _extrude_step(step, As, Bs, buf, cap, width, m, spans, betas, bscales, far_points):
                return
    # innovation GEMM per distinct B (one per outer-scale group); the kernel
    # reads every innovation row in place through a host array of pointers
    bb_of = {}
    bbrow = [0] * L
    for key in dict.fromkeys(id(b) for b in Bs):
        members = [l for l in range(L) if id(Bs[l]) == key and len(plans[l])]
        if not members:
            continue
        rows = []
        keep = []
        for l in members:
            beta = betas[l]
            if beta.dtype != t.float32 or beta.device != buf.device or beta.dim() != 2 \
                    or beta.shape[1] != width or beta.stride(1) != 1:
                beta = beta.to(device=buf.device, dtype=t.float32).reshape(-1, width).contiguous()
            if beta.shape[0] != len(plans[l]):
                raise ValueError(f"layer {l}: {beta.shape[0]} innovation rows for "
                                 f"{len(plans[l])} new rows")
            keep.append(beta)
            base, pitch = beta.data_ptr(), beta.stride(0) * 4
            rows.extend(base + i * pitch for i in range(beta.shape[0]))
        total = len(rows)
        bb = t.empty((total, ld), dtype=t.float32, device=buf.device)
        ptrs = (_ct.c_void_p * total)(*rows)
        rc = lib.pt_innovation_gemm(ptrs, _hip.ptr(Bs[members[0]]), total, width, ld,
                                    _hip.ptr(bb), st)
        _hip.check(rc, "pt_innovation_gemm")
        off = 0
        for l in members:
            bb_of[l] = bb
            bbrow[l] = off
            off += len(plans[l])
    rounds = max((len(p) for p in plans), default=0)
    for r in range(rounds):
        active = [l for l in range(L) if len(plans[l]) > r]
        for key in dict.fromkeys((id(As[l]), id(bb_of[l])) for l in active):
            grp = [l for l in active if (id(As[l]), id(bb_of[l])) == key]
            n = len(grp)
            layer = (ctypes_int * n)(*grp)
            zr = (ctypes_int * (n * m))(*[z for l in grp for z in plans[l][r][1]])
            tr = (ctypes_int * n)(*[plans[l][r][0] for l in grp])
            br = (ctypes_int * n)(*[bbrow[l] + r for l in grp])
            sc = (ctypes_float * n)(*[float(bscales[l]) for l in grp])
            if far_points:
                R = len(plans[grp[0]][r][2])
                fr = (ctypes_int * (n * R))(*[z for l in grp for z in plans[l][r][2]])
                rc = lib.pt_extrude_rows_far(_hip.ptr(As[grp[0]]), width, ld, m, int(far_points),
                                             _hip.ptr(bb_of[grp[0]]), _hip.ptr(buf), cap, n,
                                             layer, zr, fr, tr, br, sc, st)
                _hip.check(rc, "pt_extrude_rows_far")
                continue
            rc = lib.pt_extrude_rows(_hip.ptr(As[grp[0]]), width, ld, m, _hip.ptr(bb_of[grp[0]]),
                                     _hip.ptr(buf), cap, n, layer, zr, tr, br, sc, st)
            _hip.check(rc, "pt_extrude_rows")


import ctypes as _ct  # noqa: E402

ctypes_int = _ct.c_int
ctypes_float = _ct.c_float
ctypes_double = _ct.c_double
ctypes_ll = _ct.c_longlong


def readout(buf, cap, width, rings, wind_cs, dir_geom, n, interp, scale=1.0, out=None,
            accumulate=False):
    """Fused readout of ring buffers (see pt_readout); returns (D, n, n)."""
    t = _hip.torch()
    L = len(rings)
    D = len(dir_geom)
    if out is None:
        out = t.empty((D, n, n), dtype=t.float32, device=buf.device)
    start = np.array([r.start for r in rings], dtype=np.int32)
    fill = np.array([r.fill for r in rings], dtype=np.int64)
    cs = np.ascontiguousarray(wind_cs, dtype=np.float64)
    geom = np.ascontiguousarray(dir_geom, dtype=np.float64)
    if cs.size != 2 * L or geom.size != 3 * D * L:
        raise ValueError("readout: wind_cs must be (layers, 2) and dir_geom (directions, layers, 3)")
    rc = _hip.lib().pt_readout(_hip.ptr(buf), L, cap, width, int(buf.shape[-1]), start.ctypes.data,
                               fill.ctypes.data, cs.ctypes.data, geom.ctypes.data, D, n,
                               _INTERP_CODE[interp], float(scale), int(bool(accumulate)),
                               _hip.ptr(out), _hip.stream_ptr())
    _hip.check(rc, "pt_readout")
    return out


def readout_adjoint(buf, cap, width, rings, wind_cs, dir_geom, n, interp, grad, scale=1.0):
    """dLoss/d(sa, sp, mag) per (direction, layer) for dLoss/d(out) = grad (see
    pt_readout_adjoint); returns a (D, L, 3) float64 device tensor."""
    t = _hip.torch()
    L = len(rings)
    D = len(dir_geom)
    if tuple(grad.shape) != (D, n, n) or grad.dtype != t.float32:
        raise ValueError(f"grad must be ({D}, {n}, {n}) float32")
    grad = grad.contiguous()
    out = t.empty((D, L, 3), dtype=t.float64, device=buf.device)
    work = t.empty((int(_hip.lib().pt_readout_adjoint_work_elems(n)),), dtype=t.float32,
                   device=buf.device)
    start = (ctypes_int * L)(*[r.start for r in rings])
    fill = (ctypes_ll * L)(*[r.fill for r in rings])
    cs = (ctypes_double * (2 * L))(*[v for c in wind_cs for v in c])
    geom = (ctypes_double * (3 * D * L))(*np.asarray(dir_geom, dtype=np.float64).ravel())
    rc = _hip.lib().pt_readout_adjoint(_hip.ptr(buf), L, cap, width, int(buf.shape[-1]), start,
                                       fill, cs, geom, D, n, _INTERP_CODE[interp], float(scale),
                                       _hip.ptr(grad), _hip.ptr(out), _hip.ptr(work),
                                       _hip.stream_ptr())
    _hip.check(rc, "pt_readout_adjoint")
    return out


class InfinitePhaseScreen:
    """Endless frozen-flow screen extruded row by row along axis 0.

    Same parameters and semantics as the reference (infinite.py:120-395):
    ``step(k)``/``advance(pixels)`` move the wind monotonically, the pupil is
    interpolated at the exact sub-pixel offset; results are device tensors.
    Innovations are drawn from the same host stream as the reference unless
    ``innovations`` (a callable ``k -> (k, n) device tensor``) is given.
    """

    def __init__(self, n: int, pixel_scale: float, r0: float, L0: float = 25.0,
                 stencil_rows: int = 2, interp: str = "cubic", seed: Any = None,
                 device: str = "mi355x", dtype: Union[str, np.dtype] = "float32",
                 innovations=None) -> None:
        config = ScreenConfig.create(n, pixel_scale, r0, L0, device, dtype,
                                     finite_outer_scale=True)
        if (not isinstance(stencil_rows, (int, np.integer)) or not 1 <= stencil_rows < config.n):
            raise ValueError("stencil_rows must be in [1, n)")
        if interp not in ("cubic", "linear", "lanczos"):
            raise ValueError("interp must be 'cubic', 'linear', or 'lanczos'")
        self.n = config.n
        self.pixel_scale = config.pixel_scale
        self.r0 = config.r0
        self.L0 = config.L0
        self.stencil_rows = int(stencil_rows)
        self.interp = interp
        self.device = config.device
        self.dtype = config.dtype
        self._dev = _hip.device(resolve_device(config.device))
        self._A, self._B = device_extrusion(self.n, self.stencil_rows, self.pixel_scale,
                                            self.L0, self.r0, self._dev)
        gen = PhaseScreen(self.n, self.pixel_scale, self.r0, self.L0, seed=seed,
                          device=device, dtype=dtype)
        self._rng = gen._rng
        self._innovations = innovations
        self._capacity = self.n + self.stencil_rows + 6
        t = _hip.torch()
        self._buf = t.zeros((1, self._capacity, row_pitch(self.n)), dtype=t.float32,
                            device=self._dev)
        self._buf[0, :self.n, :self.n] = gen.generate().float()
        self._ring = Ring(self._capacity, self.stencil_rows, 0, self.n)
        self._travel = 0.0
        self._count = 0
        self._advance_to(0.0)

    def _draw(self, k: int):
        if self._innovations is not None:
            beta = self._innovations(k)
        else:
            beta = _hip.to_device(self._rng.standard_normal((k, self.n), dtype=self.dtype),
                                  np.float32, self._dev)
        self._count += k
        return beta

    def _advance_to(self, travel: float) -> None:
        if travel < self._travel:
            raise ValueError("wind travel is monotonic; travel cannot decrease")
        self._travel = float(travel)
        top = int(np.floor(self._travel)) + self.n + 2
        plan = self._ring.plan(top, self._travel)
        if plan:
            extrude_rounds(self._A, self._B, self._buf, self._capacity, self.n,
                           self.stencil_rows, [plan], [self._draw(len(plan))], [1.0])
        half = 0.5 * (self.n - 1)
        geom = [[[(self._travel - self._ring.base) + half, -0.5, 1.0]]]
        out = readout(self._buf, self._capacity, self.n, [self._ring], [(1.0, 0.0)], geom,
                      self.n, self.interp)
        self._current = out[0].double() if self.dtype == np.float64 else out[0]

    @property
    def screen(self) -> Any:
        return self._current

    def step(self, steps: int = 1) -> Any:
        steps = int(steps)
        if steps < 1:
            raise ValueError("steps must be >= 1")
        self._advance_to(self._travel + steps)
        return self._current

    def advance(self, pixels: float) -> Any:
        pixels = float(pixels)
        if pixels < 0:
            raise ValueError("pixels must be >= 0 (wind travel is monotonic)")
        self._advance_to(self._travel + pixels)
        return self._current

    @property
    def travel(self) -> float:
        return self._travel

    def __repr__(self) -> str:
        return (f"InfinitePhaseScreen(n={self.n}, pixel_scale={self.pixel_scale}, "
                f"r0={self.r0}, L0={self.L0}, stencil_rows={self.stencil_rows}, "
                f"interp={self.interp!r}, device={self.device!r}, dtype={self.dtype.name!r})")
