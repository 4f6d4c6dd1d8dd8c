"""Thin typed wrappers over the C entry points (device tensors in and out).

Each wrapper checks shapes/dtypes on the host before launching (the
kernels assume them), allocates the output and workspace with the torch
caching allocator when the caller does not pass them, and launches on the
current torch stream.  No wrapper computes anything on the host.
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from . import _hip


def _t():
    return _hip.torch()


def fast_size(n: int) -> bool:
    """Power-of-two grids 32..2048 run the fused LDS Stockham kernels."""
    return 32 <= n <= 2048 and (n & (n - 1)) == 0


def work_elems(kind: int, n: int, sets: int, levels: int) -> int:
    """Workspace in complex64 elements (pt_work_elems): 0 generate, 1 frames."""
    v = int(_hip.lib().pt_work_elems(int(kind), int(n), int(sets), int(levels)))
    if v < 0:
        raise ValueError("pt_work_elems: bad arguments")
    return v


def frames_work_elems(n: int, frames: int, layers: int, levels: int) -> int:
    """Exact workspace of spectral_frames in complex64 elements."""
    v = int(_hip.lib().pt_frames_work_elems(int(n), int(frames), int(layers), int(levels)))
    if v < 0:
        raise ValueError("pt_frames_work_elems: bad arguments")
    return v


def _require(t, dtype, shape, name):
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: expected {dtype}, got {t.dtype}")
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous device tensor")


def fft2(x, sign: int = 1, out=None):
    """Unnormalised 2-D DFT over the last two axes of a complex64 tensor.

    ``sign=+1`` is ``ifft2(x) * n0 * n1`` (the reference's convention),
    ``sign=-1`` the forward DFT.
    """
    t = _t()
    if x.dtype != t.complex64:
        raise ValueError("fft2: complex64 input required")
    x = x.contiguous()
    n0, n1 = x.shape[-2:]
    batch = int(np.prod(x.shape[:-2])) if x.dim() > 2 else 1
    if out is None:
        out = t.empty_like(x)
    work = None if fast_size(n0) and n0 == n1 else t.empty_like(x)
    rc = _hip.lib().pt_fft2(_hip.ptr(x), _hip.ptr(out), _hip.ptr(work), batch, n0, n1, int(sign),
                            _hip.stream_ptr())
    _hip.check(rc, "pt_fft2")
    return out


def generate(noise, sh_noise, amplitude, sh_amp, count: int, *, crop_offset: int = 0,
             n_out: Optional[int] = None, scale: float = 1.0, screen_scale=None,
             out=None, accumulate: bool = False, work=None):
    """Fused ``PhaseScreen.generate`` on injected noise (see pt_generate).

    ``noise`` (2, ceil(count/2), n, n) float32; ``sh_noise`` (P, 2, count, 3, 3)
    float32 or None; ``amplitude`` (n, n) float32; ``sh_amp`` (P, 3, 3)
    float64.  Returns ``out`` (count, n_out, n_out) float32.
    """
    t = _t()
    n = int(amplitude.shape[-1])
    n_fft = (count + 1) // 2
    _require(noise, t.float32, (2, n_fft, n, n), "noise")
    _require(amplitude, t.float32, (n, n), "amplitude")
    P = 0 if sh_amp is None else int(sh_amp.shape[0])
    if P:
        _require(sh_amp, t.float64, (P, 3, 3), "sh_amp")
        _require(sh_noise, t.float32, (P, 2, count, 3, 3), "sh_noise")
    n_out = n if n_out is None else int(n_out)
    if out is None:
        out = t.zeros((count, n_out, n_out), dtype=t.float32, device=noise.device) if accumulate \
            else t.empty((count, n_out, n_out), dtype=t.float32, device=noise.device)
    _require(out, t.float32, (count, n_out, n_out), "out")
    if screen_scale is not None:
        _require(screen_scale, t.float32, (count,), "screen_scale")
    need = work_elems(0, n, count, P)
    if work is None:
        work = t.empty((need,), dtype=t.complex64, device=noise.device)
    elif work.numel() < need or work.dtype != t.complex64:
        raise ValueError(f"work: need {need} complex64 elements")
    plane = n_fft * n * n * 4
    base = _hip.ptr(noise)
    rc = _hip.lib().pt_generate(base, base + plane, _hip.ptr(amplitude),
                                _hip.ptr(sh_noise) if P else None,
                                _hip.ptr(sh_amp) if P else None, P, int(count), n,
                                int(crop_offset), n_out, float(scale),
                                _hip.ptr(screen_scale), int(bool(accumulate)),
                                _hip.ptr(out), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "pt_generate")
    return out


def generate_groups(noise, sh_noise, amplitude, sh_amp, count: int, *, crop_offset: int = 0,
                    n_out: Optional[int] = None, scale: float = 1.0, screen_scale=None,
                    out=None, accumulate: bool = False, work=None):
    """Summed screens of G outer-scale groups in one pass (pt_generate_groups).

    ``noise`` (G, 2, ceil(count/2), n, n) float32; ``sh_noise``
    (G, P, 2, count, 3, 3) float32 or None; ``amplitude`` (G, n, n) float32;
    ``sh_amp`` (G, P, 3, 3) float64.  Returns ``out`` (count, n_out, n_out).
    """
    t = _t()
    G, n = int(amplitude.shape[0]), int(amplitude.shape[-1])
    n_fft = (count + 1) // 2
    _require(noise, t.float32, (G, 2, n_fft, n, n), "noise")
    _require(amplitude, t.float32, (G, n, n), "amplitude")
    P = 0 if sh_amp is None else int(sh_amp.shape[1])
    if P:
        _require(sh_amp, t.float64, (G, P, 3, 3), "sh_amp")
        _require(sh_noise, t.float32, (G, P, 2, count, 3, 3), "sh_noise")
    n_out = n if n_out is None else int(n_out)
    if out is None:
        out = t.zeros((count, n_out, n_out), dtype=t.float32, device=noise.device) if accumulate \
            else t.empty((count, n_out, n_out), dtype=t.float32, device=noise.device)
    _require(out, t.float32, (count, n_out, n_out), "out")
    if screen_scale is not None:
        _require(screen_scale, t.float32, (count,), "screen_scale")
    need = work_elems(0, n, count, P)
    if work is None:
        work = t.empty((need,), dtype=t.complex64, device=noise.device)
    elif work.numel() < need or work.dtype != t.complex64:
        raise ValueError(f"work: need {need} complex64 elements")
    rc = _hip.lib().pt_generate_groups(_hip.ptr(noise), _hip.ptr(amplitude),
                                       _hip.ptr(sh_noise) if P else None,
                                       _hip.ptr(sh_amp) if P else None, G, P, int(count), n,
                                       int(crop_offset), n_out, float(scale),
                                       _hip.ptr(screen_scale), int(bool(accumulate)),
                                       _hip.ptr(out), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "pt_generate_groups")
    return out


class EnsembleClass:
    """One size class of :func:`generate_ensemble`: the arrays of a
    :func:`generate_groups` call without crop (``screen_scale``, ``out`` and
    ``work`` optional)."""

    __slots__ = ("noise", "sh_noise", "amplitude", "sh_amp", "count", "screen_scale", "out", "work")

    def __init__(self, noise, sh_noise, amplitude, sh_amp, count: int, screen_scale=None,
                 out=None, work=None) -> None:
        self.noise, self.sh_noise, self.amplitude, self.sh_amp = noise, sh_noise, amplitude, sh_amp
        self.count, self.screen_scale, self.out, self.work = int(count), screen_scale, out, work


def generate_ensemble(classes, *, scale: float = 1.0, persistent_cols: bool = False):
    """Screens of several grid sizes in three launches shared by every class
    (pt_generate_ensemble): class k gives the screens of
    ``generate_groups(noise, sh_noise, amplitude, sh_amp, count,
    screen_scale=..., scale=scale)`` bit for bit.  ``classes`` is a list of
    :class:`EnsembleClass` with the same group count and subharmonic depth;
    n must be a power of two in 128..2048.  The row pass and the subharmonic
    prep are persistent launches over every class; the column pass runs per
    class, or, with ``persistent_cols``, as one persistent launch too.
    Returns the list of outputs."""
    import ctypes

    t = _t()
    K = len(classes)
    if not 1 <= K <= 8:
        raise ValueError("generate_ensemble: 1..8 size classes")
    G = int(classes[0].amplitude.shape[0])
    P = 0 if classes[0].sh_amp is None else int(classes[0].sh_amp.shape[1])
    ns, outs, works = [], [], []
    for c in classes:
        n, count = int(c.amplitude.shape[-1]), c.count
        if int(c.amplitude.shape[0]) != G or (0 if c.sh_amp is None else int(c.sh_amp.shape[1])) != P:
            raise ValueError("generate_ensemble: every class needs the same groups and levels")
        _require(c.noise, t.float32, (G, 2, (count + 1) // 2, n, n), "noise")
        _require(c.amplitude, t.float32, (G, n, n), "amplitude")
        if P:
            _require(c.sh_amp, t.float64, (G, P, 3, 3), "sh_amp")
            _require(c.sh_noise, t.float32, (G, P, 2, count, 3, 3), "sh_noise")
        if c.screen_scale is not None:
            _require(c.screen_scale, t.float32, (count,), "screen_scale")
        out = c.out if c.out is not None else t.empty((count, n, n), dtype=t.float32,
                                                      device=c.noise.device)
        _require(out, t.float32, (count, n, n), "out")
        need = work_elems(0, n, count, P)
        work = c.work if c.work is not None else t.empty((need,), dtype=t.complex64,
                                                         device=c.noise.device)
        if work.numel() < need or work.dtype != t.complex64:
            raise ValueError(f"work: need {need} complex64 elements")
        ns.append(n)
        outs.append(out)
        works.append(work)
    vp = ctypes.c_void_p * K
    ci = ctypes.c_int * K
    rc = _hip.lib().pt_generate_ensemble(
        K, ci(*ns), ci(*[c.count for c in classes]), vp(*[_hip.ptr(c.noise) for c in classes]),
        vp(*[_hip.ptr(c.amplitude) for c in classes]),
        vp(*[_hip.ptr(c.sh_noise) for c in classes]) if P else None,
        vp(*[_hip.ptr(c.sh_amp) for c in classes]) if P else None, G, P,
        vp(*[_hip.ptr(c.screen_scale) for c in classes]), float(scale),
        vp(*[_hip.ptr(o) for o in outs]), vp(*[_hip.ptr(w) for w in works]), int(bool(persistent_cols)),
        _hip.stream_ptr())
    _hip.check(rc, "pt_generate_ensemble")
    return outs


def spectral_fold(spectra, frames: int, work):
    """Fold the layer stack into ``work`` for calls of ``frames`` frames
    (pt_spectral_fold); then ``spectral_frames(..., folded=True)`` skips it."""
    t = _t()
    L, n = int(spectra.shape[0]), int(spectra.shape[-1])
    _require(spectra, t.complex64, (L, n, n), "spectra")
    rc = _hip.lib().pt_spectral_fold(_hip.ptr(spectra), L, n, int(frames), _hip.ptr(work),
                                     _hip.stream_ptr())
    _hip.check(rc, "pt_spectral_fold")


def spectral_frames(spectra, disp, sh_coeffs, pixel_scale: float, *, crop_offset: int = 0,
                    n_out: Optional[int] = None, scale: float = 1.0, out=None, work=None,
                    folded: bool = False):
    """Frozen-flow frames (see pt_spectral_frames).

    ``spectra`` (L, n, n) complex64; ``disp`` (F, L, 2) float64 metres;
    ``sh_coeffs`` (P, L, 3, 3) complex128 or None.  Returns (F, n_out, n_out).
    With ``folded``, ``work`` already holds the fold of these spectra for
    this frame count (``spectral_fold``): pt_spectral_frames_folded
    (power-of-two n, <= 16 layers, <= 12 levels); ``disp`` may then be a
    pinned host tensor, read by the table kernels in place.
    """
    t = _t()
    L, n = int(spectra.shape[0]), int(spectra.shape[-1])
    _require(spectra, t.complex64, (L, n, n), "spectra")
    F = int(disp.shape[0])
    if folded and not disp.is_cuda:
        # a reused fold reads the displacements in its table kernels only, so
        # they may stay in pinned host memory (no upload per call)
        if not disp.is_pinned() or disp.dtype != t.float64 or tuple(disp.shape) != (F, L, 2) \
                or not disp.is_contiguous():
            raise ValueError(f"disp: expected a contiguous ({F}, {L}, 2) float64 device or "
                             "pinned host tensor")
    else:
        _require(disp, t.float64, (F, L, 2), "disp")
    P = 0 if sh_coeffs is None else int(sh_coeffs.shape[0])
    if P:
        _require(sh_coeffs, t.complex128, (P, L, 3, 3), "sh_coeffs")
    n_out = n if n_out is None else int(n_out)
    if out is None:
        out = t.empty((F, n_out, n_out), dtype=t.float32, device=spectra.device)
    _require(out, t.float32, (F, n_out, n_out), "out")
    need = frames_work_elems(n, F, L, P)
    if folded:
        if work is None or work.numel() < need:
            raise ValueError("folded: pass the workspace that spectral_fold filled")
        dptr = disp.data_ptr() if not disp.is_cuda else _hip.ptr(disp)
        rc = _hip.lib().pt_spectral_frames_folded(L, n, float(pixel_scale), dptr, F,
                                                  _hip.ptr(sh_coeffs) if P else None, P,
                                                  int(crop_offset), n_out, float(scale),
                                                  _hip.ptr(out), _hip.ptr(work), _hip.stream_ptr())
        _hip.check(rc, "pt_spectral_frames_folded")
        return out
    if work is None or work.numel() < need:
        work = t.empty((need,), dtype=t.complex64, device=spectra.device)
    rc = _hip.lib().pt_spectral_frames(_hip.ptr(spectra), L, n, float(pixel_scale),
                                       _hip.ptr(disp), F, _hip.ptr(sh_coeffs) if P else None,
                                       P, int(crop_offset), n_out, float(scale),
                                       _hip.ptr(out), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "pt_spectral_frames")
    return out


def layer_spectra(noise, amplitude, group, gain, out=None):
    """Layer spectra ``gain_l * amplitude[group_l] * (w_re + i w_im)`` (pt_layer_spectra).

    ``noise`` (L, 2, n, n) float32, ``amplitude`` (G, n, n) float32; ``group``
    and ``gain`` are host sequences of length L.  Returns (L, n, n) complex64.
    """
    t = _t()
    L, n = int(noise.shape[0]), int(noise.shape[-1])
    G = int(amplitude.shape[0])
    _require(noise, t.float32, (L, 2, n, n), "noise")
    _require(amplitude, t.float32, (G, n, n), "amplitude")
    grp = np.ascontiguousarray(np.asarray(group, dtype=np.int32))
    gn = np.ascontiguousarray(np.asarray(gain, dtype=np.float64))
    if grp.shape != (L,) or gn.shape != (L,) or grp.min() < 0 or grp.max() >= G:
        raise ValueError("layer_spectra: group/gain must have one entry per layer, groups < G")
    if out is None:
        out = t.empty((L, n, n), dtype=t.complex64, device=noise.device)
    _require(out, t.complex64, (L, n, n), "out")
    rc = _hip.lib().pt_layer_spectra(_hip.ptr(noise), _hip.ptr(amplitude), grp.ctypes.data,
                                     gn.ctypes.data, L, n, _hip.ptr(out), _hip.stream_ptr())
    _hip.check(rc, "pt_layer_spectra")
    return out


def spectral_frames_adjoint(noise, amplitude, dlnamp, group, gain, disp, grad, sh_unit, sh_rho,
                            pixel_scale: float, *, crop_offset: int = 0, scale: float = 1.0,
                            out=None, work=None):
    """Adjoint of layer_spectra + spectral_frames (pt_spectral_frames_adjoint).

    ``grad`` (F, n_out, n_out) float32 is dLoss/dframes.  Returns (F, L, 4)
    float64: dLoss/dgain_l, dLoss/ds_x, dLoss/ds_y, dLoss/dL0_l per frame.
    """
    t = _t()
    L, n = int(noise.shape[0]), int(noise.shape[-1])
    G = int(amplitude.shape[0])
    _require(noise, t.float32, (L, 2, n, n), "noise")
    _require(amplitude, t.float32, (G, n, n), "amplitude")
    _require(dlnamp, t.float32, (G, n, n), "dlnamp")
    F = int(disp.shape[0])
    _require(disp, t.float64, (F, L, 2), "disp")
    n_out = int(grad.shape[-1])
    _require(grad, t.float32, (F, n_out, n_out), "grad")
    P = 0 if sh_unit is None else int(sh_unit.shape[0])
    if P:
        _require(sh_unit, t.complex128, (P, L, 3, 3), "sh_unit")
        _require(sh_rho, t.float64, (P, L, 3, 3), "sh_rho")
    grp = np.ascontiguousarray(np.asarray(group, dtype=np.int32))
    gn = np.ascontiguousarray(np.asarray(gain, dtype=np.float64))
    if grp.shape != (L,) or gn.shape != (L,) or grp.min() < 0 or grp.max() >= G:
        raise ValueError("spectral_frames_adjoint: group/gain must have one entry per layer")
    if out is None:
        out = t.empty((F, L, 4), dtype=t.float64, device=noise.device)
    _require(out, t.float64, (F, L, 4), "out")
    need = int(_hip.lib().pt_adjoint_work_elems(n, F, P))
    if need < 0:
        raise ValueError("pt_adjoint_work_elems: bad arguments")
    if work is None or work.numel() < need:
        work = t.empty((need,), dtype=t.complex64, device=noise.device)
    rc = _hip.lib().pt_spectral_frames_adjoint(
        _hip.ptr(noise), _hip.ptr(amplitude), _hip.ptr(dlnamp), grp.ctypes.data, gn.ctypes.data,
        L, n, float(pixel_scale), _hip.ptr(disp), F, _hip.ptr(grad), int(crop_offset), n_out,
        _hip.ptr(sh_unit) if P else None, _hip.ptr(sh_rho) if P else None, P, float(scale),
        _hip.ptr(out), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "pt_spectral_frames_adjoint")
    return out


def boiled_frames(spectra, amplitude, inv_tau, f_ref, disp, dt: float, pixel_scale: float, *,
                  sh_coeffs=None, sh_amp=None, steps: Optional[int] = None, noise=None,
                  sh_noise=None, seed: int = 0, step0: int = 0, crop_offset: int = 0,
                  n_out: Optional[int] = None, scale: float = 1.0, out=None, work=None):
    """Frames with spectral boiling (pt_boiled_frames); updates the state.

    ``spectra`` (L, n, n) complex64 and ``sh_coeffs`` (P, L, 3, 3)
    complex128 are the boiling state, advanced in place by ``steps`` steps
    (default: one per frame); ``amplitude`` (L, n, n) float32 and ``sh_amp``
    (P, L, 3, 3) float64 the layers' PSD amplitudes; ``inv_tau`` and
    ``f_ref`` (L,) float64 device tensors.  ``disp`` (F, L, 2) float64
    metres, or None for boiling only.  ``noise`` (steps, 2, L, n, n) float32
    and ``sh_noise`` (steps, 2, P, L, 3, 3) float64 inject the normals;
    otherwise the device stream keyed by ``seed`` from step ``step0``.
    Returns (F, n_out, n_out) float32 (None when F = 0).
    """
    t = _t()
    L, n = int(spectra.shape[0]), int(spectra.shape[-1])
    _require(spectra, t.complex64, (L, n, n), "spectra")
    _require(amplitude, t.float32, (L, n, n), "amplitude")
    _require(inv_tau, t.float64, (L,), "inv_tau")
    _require(f_ref, t.float64, (L,), "f_ref")
    F = 0 if disp is None else int(disp.shape[0])
    if F:
        _require(disp, t.float64, (F, L, 2), "disp")
    steps = F if steps is None else int(steps)
    P = 0 if sh_coeffs is None else int(sh_coeffs.shape[0])
    if P:
        _require(sh_coeffs, t.complex128, (P, L, 3, 3), "sh_coeffs")
        _require(sh_amp, t.float64, (P, L, 3, 3), "sh_amp")
    if noise is not None:
        _require(noise, t.float32, (steps, 2, L, n, n), "noise")
        if P:
            _require(sh_noise, t.float64, (steps, 2, P, L, 3, 3), "sh_noise")
    n_out = n if n_out is None else int(n_out)
    if F:
        if out is None:
            out = t.empty((F, n_out, n_out), dtype=t.float32, device=spectra.device)
        _require(out, t.float32, (F, n_out, n_out), "out")
        need = int(_hip.lib().pt_boil_work_elems(n, F, L, P))
        if need < 0:
            raise ValueError("pt_boil_work_elems: bad arguments")
        if work is None or work.numel() < need:
            work = t.empty((need,), dtype=t.complex64, device=spectra.device)
    rc = _hip.lib().pt_boiled_frames(
        _hip.ptr(spectra), _hip.ptr(amplitude), _hip.ptr(inv_tau), _hip.ptr(f_ref), L, n,
        float(pixel_scale), _hip.ptr(disp) if F else None, F, steps, float(dt),
        _hip.ptr(sh_coeffs) if P else None, _hip.ptr(sh_amp) if P else None, P,
        _hip.ptr(noise) if noise is not None else None,
        _hip.ptr(sh_noise) if (noise is not None and P) else None,
        int(seed) & 0xFFFFFFFFFFFFFFFF, int(step0), int(crop_offset), n_out, float(scale),
        _hip.ptr(out) if F else None, _hip.ptr(work) if F else None, _hip.stream_ptr())
    _hip.check(rc, "pt_boiled_frames")
    return out if F else None


def lgs_zoom(screens, idx, w, n: int, *, scale: float = 1.0, out=None):
    """Cone-zoomed layer sum (pt_lgs_zoom).

    ``screens`` (L, F, ns, ns) float32; ``idx`` (L, T, n) int32 and ``w``
    (L, T, n) float32 device tap tables.  Returns (F, n, n) float32.
    """
    t = _t()
    L, F, ns = int(screens.shape[0]), int(screens.shape[1]), int(screens.shape[-1])
    _require(screens, t.float32, (L, F, ns, ns), "screens")
    T = int(idx.shape[1])
    _require(idx, t.int32, (L, T, n), "idx")
    _require(w, t.float32, (L, T, n), "w")
    if out is None:
        out = t.empty((F, n, n), dtype=t.float32, device=screens.device)
    _require(out, t.float32, (F, n, n), "out")
    rc = _hip.lib().pt_lgs_zoom(_hip.ptr(screens), L, F, ns, _hip.ptr(idx), _hip.ptr(w), T, n,
                                float(scale), _hip.ptr(out), _hip.stream_ptr())
    _hip.check(rc, "pt_lgs_zoom")
    return out
