"""HIP-graph replay of fixed-shape hot-path calls.

The library's entry points launch kernels only: no memset, copy or
allocation happens inside a call, so a sequence of calls over static
buffers records into a HIP graph (``torch.cuda.CUDAGraph``) and replays with
one launch from the host.

* ``Captured(fn)`` records any such sequence, e.g. a Monte-Carlo ensemble of
  mixed screen sizes (one ``ops.generate_groups`` call per size class over
  noise that the caller refreshes in place between replays).
* ``CapturedFrames(atm, frames, directions)`` is ``atm.opd_batch`` for a
  fixed number of times.  A closed-loop simulation asks for a few frames at
  a time, over and over, and each eager call is then dominated by fixed
  costs: five launches through ctypes, a displacement upload and the fold
  of the unchanged layer stack.  The stack is folded once at construction
  (``ops.spectral_fold``); each call writes the displacement table into
  pinned host memory, which the table kernels read in place (no upload),
  and replays the per-call kernels (three for calls of up to 8 frames).
  Two pinned tables and two graphs alternate, so the host fills one while
  the GPU may still read the other.
* ``FoldedFrames(atm, frames, directions)`` is the same call pattern without
  a graph: one ``pt_spectral_frames_folded`` call per step through ctypes.
  On ROCm a graph launch costs the host about 23 us, three kernel launches
  about 9 us, so for calls of a few frames this is the faster loop.
"""

from __future__ import annotations

from typing import Any, Optional, Sequence

import numpy as np

from . import _hip, ops


class Captured:
    """A HIP graph of the launches ``fn()`` makes on the current stream.

    ``fn`` runs once eagerly on a side stream (so that lazily built state
    such as LDS limits is set up outside the capture), then once under
    capture.  Every buffer it touches must stay allocated and at the same
    address while the graph is replayed.
    """

    def __init__(self, fn, device=None) -> None:
        t = _hip.torch()
        dev = device if device is not None else t.cuda.current_device()
        side = t.cuda.Stream(device=dev)
        side.wait_stream(t.cuda.current_stream(dev))
        with t.cuda.stream(side):
            fn()
        t.cuda.current_stream(dev).wait_stream(side)
        self.graph = t.cuda.CUDAGraph()
        with t.cuda.graph(self.graph):
            fn()

    def replay(self) -> None:
        """Launch the recorded work on the current stream."""
        self.graph.replay()


class CapturedFrames:
    """``atm.opd_batch(times, directions)`` for a fixed number of times and
    directions, replayed from a HIP graph.  The returned tensor is the graph's
    static output: copy it before the next call if it must be kept."""

    def __init__(self, atm, frames: int, directions: Optional[Sequence] = None,
                 wavelength: Optional[float] = None) -> None:
        t = _hip.torch()
        if atm.engine != "spectral" or atm._lgs_mag is not None or atm._boiling:
            raise ValueError("CapturedFrames needs a frozen-flow spectral atmosphere "
                             "(no LGS cone, no boiling)")
        self.atm = atm
        self.frames = int(frames)
        if directions is None:
            self.ox, self.oy = np.zeros(1), np.zeros(1)
        else:
            self.ox, self.oy = atm._dirs_tan(directions)
        self.D = len(self.ox)
        L = len(atm._layers)
        self.scale = atm._scale(wavelength)
        F = self.frames * self.D
        dev = atm._dev
        # disp[t, d, l] = v_l * t + h_l * theta_d, in the order opd_batch uses:
        # the per-call host work is one multiply and one add into the pinned table
        self._vel = np.stack([atm._vx, atm._vy], axis=-1)  # (L, 2)
        self._base = np.stack([atm._alt[None, :] * self.ox[:, None],
                               atm._alt[None, :] * self.oy[:, None]], axis=-1)  # (D, L, 2)
        self.out = t.empty((F, atm.n, atm.n), dtype=t.float32, device=dev)
        P = 0 if atm._sh_coeffs is None else int(atm._sh_coeffs.shape[0])
        self.work = t.empty((ops.frames_work_elems(atm.n_screen, F, L, P),), dtype=t.complex64,
                            device=dev)
        ns = atm.n_screen
        # the fused engine's fold depends only on the (fixed) spectra: fold
        # once here and record only the per-call kernels
        self.folded = (ns & (ns - 1)) == 0 and 32 <= ns <= 2048 and L <= 16 and P <= 12
        if self.folded:
            ops.spectral_fold(atm._spectra, F, self.work)
            # the table kernels read the pinned tables in place
            self._disp = [t.zeros((F, L, 2), dtype=t.float64, pin_memory=True) for _ in range(2)]
            self._host = self._disp
        else:
            # the general path reads its displacements on the device: one
            # upload per call from the pinned table
            self._host = [t.zeros((F, L, 2), dtype=t.float64, pin_memory=True) for _ in range(2)]
            self._disp = [t.zeros((F, L, 2), dtype=t.float64, device=dev) for _ in range(2)]
        self._views = [h.numpy().reshape(self.frames, self.D, L, 2) for h in self._host]
        self._done = [t.cuda.Event(), t.cuda.Event()]  # table k is free once this completes
        self._caps = [Captured(lambda k=k: self._launch(k), dev) for k in range(2)]
        self._next = 0

    def _launch(self, k: int) -> None:
        a = self.atm
        ops.spectral_frames(a._spectra, self._disp[k], a._sh_coeffs, a.pixel_scale,
                            crop_offset=a.margin_pix, n_out=a.n, scale=self.scale,
                            out=self.out, work=self.work, folded=self.folded)

    def __call__(self, times: Sequence[float]) -> Any:
        a = self.atm
        times = np.asarray(times, dtype=np.float64).ravel()
        if times.size != self.frames:
            raise ValueError(f"expected {self.frames} times, got {times.size}")
        k = self._next
        self._next ^= 1
        self._done[k].synchronize()  # the replay that last read table k has finished
        v = self._views[k]
        np.multiply(times[:, None, None, None], self._vel[None, None], out=v)
        np.add(v, self._base[None], out=v)
        if not a._wrap_warned:
            a._check_wrap(v[..., 0], v[..., 1])
        if not self.folded:
            self._disp[k].copy_(self._host[k], non_blocking=True)
        self._caps[k].replay()
        self._done[k].record()
        out = a._cast(self.out)
        if self.D > 1:
            out = out.view(self.frames, self.D, a.n, a.n)
        return out


class FoldedFrames:
    """``atm.opd_batch(times, directions)`` for a fixed number of times and
    directions, as one ``pt_spectral_frames_folded`` call per step on a stack
    folded once.  The displacement tables stay in pinned host memory: a ring
    of 2 B tables, an event recorded after every B calls, and a table
    rewritten only once the batch that last read it has finished, so the
    host waits on an event once per B calls instead of every call.  The
    returned tensor is reused by the next call: copy it if it must be kept."""

    BATCH = 16  # calls per event

    def __init__(self, atm, frames: int, directions: Optional[Sequence] = None,
                 wavelength: Optional[float] = None) -> None:
        import ctypes

        t = _hip.torch()
        if atm.engine != "spectral" or atm._lgs_mag is not None or atm._boiling:
            raise ValueError("FoldedFrames needs a frozen-flow spectral atmosphere "
                             "(no LGS cone, no boiling)")
        ns = atm.n_screen
        L = len(atm._layers)
        P = 0 if atm._sh_coeffs is None else int(atm._sh_coeffs.shape[0])
        if not ((ns & (ns - 1)) == 0 and 32 <= ns <= 2048 and L <= 16 and P <= 12):
            raise ValueError("FoldedFrames needs a power-of-two screen of 32..2048, <= 16 layers "
                             "and <= 12 subharmonic levels")
        self.atm = atm
        self.frames = int(frames)
        if directions is None:
            ox, oy = np.zeros(1), np.zeros(1)
        else:
            ox, oy = atm._dirs_tan(directions)
        self.D = len(ox)
        F = self.frames * self.D
        dev = atm._dev
        self._vel = np.stack([atm._vx, atm._vy], axis=-1)
        self._base = np.stack([atm._alt[None, :] * ox[:, None], atm._alt[None, :] * oy[:, None]],
                              axis=-1)
        nt = 2 * self.BATCH
        self._host = [t.zeros((F, L, 2), dtype=t.float64, pin_memory=True) for _ in range(nt)]
        self._views = [h.numpy().reshape(self.frames, self.D, L, 2) for h in self._host]
        self._done = [t.cuda.Event(), t.cuda.Event()]  # end of the last batch on each half of the ring
        self._armed = [False, False]
        self.out = t.empty((F, atm.n, atm.n), dtype=t.float32, device=dev)
        self.work = t.empty((ops.frames_work_elems(ns, F, L, P),), dtype=t.complex64, device=dev)
        ops.spectral_fold(atm._spectra, F, self.work)
        lib = _hip.lib()
        self._fn = lib.pt_spectral_frames_folded
        sh = _hip.ptr(atm._sh_coeffs) if P else None
        scale = float(atm._scale(wavelength))
        # everything but the displacement table is fixed: prebuilt arguments
        self._args = [(L, ns, float(atm.pixel_scale), ctypes.c_void_p(h.data_ptr()), F,
                       sh, P, int(atm.margin_pix), int(atm.n), scale, _hip.ptr(self.out),
                       _hip.ptr(self.work)) for h in self._host]
        self._next = 0

    def __call__(self, times: Sequence[float]) -> Any:
        a = self.atm
        times = np.asarray(times, dtype=np.float64).ravel()
        if times.size != self.frames:
            raise ValueError(f"expected {self.frames} times, got {times.size}")
        k = self._next
        self._next = (k + 1) % len(self._host)
        half, pos = divmod(k, self.BATCH)
        if pos == 0 and self._armed[half]:
            self._done[half].synchronize()  # the batch that last read these tables has finished
        v = self._views[k]
        np.multiply(times[:, None, None, None], self._vel[None, None], out=v)
        np.add(v, self._base[None], out=v)
        if not a._wrap_warned:
            a._check_wrap(v[..., 0], v[..., 1])
        _hip.check(self._fn(*self._args[k], _hip.stream_ptr()), "pt_spectral_frames_folded")
        if pos == self.BATCH - 1:
            self._done[half].record()
            self._armed[half] = True
        out = a._cast(self.out)
        if self.D > 1:
            out = out.view(self.frames, self.D, a.n, a.n)
        return out
