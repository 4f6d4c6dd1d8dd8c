"""Differentiable spectral frame model: fit layer profiles and winds to OPD.

The reference computes no derivatives.  This module writes its spectral
frozen-flow frames (src/pyturb/atmosphere.py:882-938, layer spectra from
:532-579) as a function of the layer parameters, so that per-layer Cn2
fractions, r0, outer scales and wind vectors can be fitted to OPD telemetry:

    r0_l   = r0 * frac_l^(-3/5)                     (atmosphere.py:401-402)
    gain_l = r0_l^(-5/6) = r0^(-5/6) * frac_l^(1/2)  (the PSD is linear in
                                                     r0^(-5/3), fourier.py:143)
    S_l    = gain_l * amp_1(L0_l) * (w_l,re + i w_l,im)    white noise w fixed
    c_pl   = gain_l * sh_amp_1(L0_l) * (v_pl,re + i v_pl,im)
    s_l(t) = wind_l * t + altitude_l * tan(theta)
    frames = scale * spectral_frames(S, s, c)

``amp_1``/``sh_amp_1`` are the unit-r0 filters (FP64 host, exact r0
scaling).  The forward runs ``pt_layer_spectra`` + ``pt_spectral_frames``;
the backward runs ``pt_spectral_frames_adjoint``, which recomputes the layer
spectra from the noise instead of storing them and returns, per frame and
layer, dLoss/dgain, dLoss/ds and dLoss/dL0.  torch chains those to the
fractions, r0 and winds (a handful of scalars per layer).
"""

from __future__ import annotations

import math
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import _hip, ops
from .backend import resolve_device
from .spectrum import build_filters_dL0

__all__ = ["ProfileFrames", "readout_frames"]


class ProfileFrames:
    """Fixed white noise plus the differentiable frame model over it.

    ``noise`` (L, 2, n, n) and ``sh_noise`` (P, L, 2, 3, 3) default to
    ``numpy.random.default_rng(seed)`` draws in that order.
    """

    def __init__(self, n: int, pixel_scale: float, n_layers: int, subharmonics: int = 8,
                 wavelength: float = 500e-9, noise=None, sh_noise=None, seed=None,
                 power_law: float = 11.0 / 3.0, device: str = "mi355x") -> None:
        if n_layers < 1 or n_layers > 16:
            raise ValueError("ProfileFrames: 1..16 layers")
        if subharmonics < 0 or subharmonics > 12:
            raise ValueError("ProfileFrames: 0..12 subharmonic levels")
        if not (pixel_scale > 0):
            raise ValueError("pixel_scale must be > 0")
        self.n = int(n)
        self.pixel_scale = float(pixel_scale)
        self.n_layers = int(n_layers)
        self.subharmonics = int(subharmonics)
        self.power_law = float(power_law)
        self.scale = float(wavelength) / (2.0 * math.pi)
        self._dev = _hip.device(resolve_device(device))
        rng = np.random.default_rng(seed)
        L, P = self.n_layers, self.subharmonics
        if noise is None:
            noise = rng.standard_normal((L, 2, n, n), dtype=np.float32)
        if sh_noise is None:
            sh_noise = rng.standard_normal((P, L, 2, 3, 3))
        noise = np.asarray(noise)
        sh_noise = np.asarray(sh_noise, dtype=np.float64)
        if noise.shape != (L, 2, n, n) or sh_noise.shape != (P, L, 2, 3, 3):
            raise ValueError("noise must be (L, 2, n, n) and sh_noise (P, L, 2, 3, 3)")
        self.noise = _hip.to_device(noise, np.float32, self._dev)
        self.sh_noise = sh_noise[:, :, 0] + 1j * sh_noise[:, :, 1]  # (P, L, 3, 3)
        self._filters: Dict[float, Tuple[np.ndarray, ...]] = {}
        self._stacks: Dict[Tuple[float, ...], tuple] = {}

    # ------------------------------------------------------------------
    def _filter(self, L0: float):
        key = round(float(L0), 12)
        if key not in self._filters:
            self._filters[key] = build_filters_dL0(self.n, self.pixel_scale, key,
                                                   self.subharmonics, self.power_law)
        return self._filters[key]

    def filters_for(self, L0: Sequence[float]):
        """Device amplitude / d ln amplitude stacks and host subharmonic terms."""
        vals = [round(float(v), 12) for v in L0]
        if any(not (v > 0) for v in vals):
            raise ValueError("outer scales must be > 0")
        uniq = sorted(set(vals))
        key = tuple(uniq)
        if key not in self._stacks:
            filt = [self._filter(v) for v in uniq]
            amp = _hip.to_device(np.stack([f[0] for f in filt]), np.float32, self._dev)
            dln = _hip.to_device(np.stack([f[1] for f in filt]), np.float32, self._dev)
            if len(self._stacks) > 8:
                self._stacks.clear()
            self._stacks[key] = (amp, dln)
        amp, dln = self._stacks[key]
        group = [uniq.index(v) for v in vals]
        P = self.subharmonics
        sh_unit = np.zeros((P, self.n_layers, 3, 3), dtype=np.complex128)
        sh_rho = np.zeros((P, self.n_layers, 3, 3))
        for l, v in enumerate(vals):
            f = self._filter(v)
            sh_unit[:, l] = f[2] * self.sh_noise[:, l]
            sh_rho[:, l] = f[3]
        return amp, dln, group, sh_unit, sh_rho

    # ------------------------------------------------------------------
    def displacements(self, wind, times, altitudes=None, theta=(0.0, 0.0)):
        """(T, L, 2) float64 device displacements s = wind t + h tan(theta)."""
        t = _hip.torch()
        wind = t.as_tensor(wind, dtype=t.float64, device=self._dev)
        times = t.as_tensor(times, dtype=t.float64, device=self._dev)
        disp = times[:, None, None] * wind[None, :, :]
        if altitudes is not None and (theta[0] != 0.0 or theta[1] != 0.0):
            h = t.as_tensor(altitudes, dtype=t.float64, device=self._dev)
            tan = t.tensor([math.tan(theta[0]), math.tan(theta[1])], dtype=t.float64,
                           device=self._dev)
            disp = disp + h[None, :, None] * tan[None, None, :]
        return disp

    def frames(self, fractions, r0, L0, wind, times, altitudes=None, theta=(0.0, 0.0)):
        """OPD frames (T, n, n) float32 [m], differentiable in every argument.

        ``fractions`` (L,) are used as given (normalise them, e.g. softmax),
        ``r0`` scalar [m], ``L0`` scalar or (L,) [m], ``wind`` (L, 2) [m/s],
        ``times`` (T,) [s].
        """
        t = _hip.torch()
        L = self.n_layers
        fr = t.as_tensor(fractions, dtype=t.float64, device=self._dev)
        r0 = t.as_tensor(r0, dtype=t.float64, device=self._dev)
        L0 = t.as_tensor(L0, dtype=t.float64, device=self._dev)
        if L0.dim() == 0:
            L0 = L0.expand(L)
        if fr.shape != (L,) or L0.shape != (L,):
            raise ValueError("fractions and L0 must have one entry per layer")
        gain = r0.pow(-5.0 / 6.0) * fr.sqrt()
        disp = self.displacements(wind, times, altitudes, theta)
        return _SpectralFramesFn.apply(gain, L0, disp, self)


def _make_fn():
    t = _hip.torch()

    class SpectralFramesFn(t.autograd.Function):
        """frames(gain, L0, disp) with the HIP adjoint as its backward."""

        @staticmethod
        def forward(ctx, gain, L0, disp, model):
            amp, dln, group, sh_unit, sh_rho = model.filters_for(L0.detach().cpu().numpy())
            g = gain.detach().cpu().numpy().astype(np.float64)
            spectra = ops.layer_spectra(model.noise, amp, group, g)
            P = model.subharmonics
            coeffs = None
            if P:
                coeffs = _hip.to_device(sh_unit * g[None, :, None, None], np.complex128, model._dev)
            d = disp.detach().contiguous()
            out = ops.spectral_frames(spectra, d, coeffs, model.pixel_scale, scale=model.scale)
            ctx.model = model
            ctx.state = (amp, dln, group, g, sh_unit, sh_rho)
            ctx.in_dev = (gain.device, L0.device, disp.device)
            ctx.save_for_backward(d)
            return out

        @staticmethod
        def backward(ctx, grad_out):
            (d,) = ctx.saved_tensors
            model = ctx.model
            amp, dln, group, g, sh_unit, sh_rho = ctx.state
            P = model.subharmonics
            su = _hip.to_device(sh_unit, np.complex128, model._dev) if P else None
            sr = _hip.to_device(sh_rho, np.float64, model._dev) if P else None
            res = ops.spectral_frames_adjoint(model.noise, amp, dln, group, g, d,
                                              grad_out.float().contiguous(), su, sr,
                                              model.pixel_scale, scale=model.scale)
            dg = res[:, :, 0].sum(0).to(ctx.in_dev[0])
            dL0 = res[:, :, 3].sum(0).to(ctx.in_dev[1])
            dd = res[:, :, 1:3].contiguous().to(ctx.in_dev[2])
            return dg, dL0, dd, None

    return SpectralFramesFn


class _LazyFn:
    def __init__(self, make):
        self._make = make
        self._fn = None

    def apply(self, *args):
        if self._fn is None:
            self._fn = self._make()
        return self._fn.apply(*args)


_SpectralFramesFn = _LazyFn(_make_fn)


# ---------------------------------------------------------------------------
# Extruder readout as a differentiable function of its geometry
# ---------------------------------------------------------------------------
def readout_frames(ext, thx, thy, mags=None, travel_offset=None, scale: float = 1.0):
    """Summed phase of an ExtrudedAtmosphere for D directions, differentiable in
    the viewing tangents ``thx``/``thy`` (D,), the per-layer extra along-wind
    travel ``travel_offset`` (L,) [pixels] and the magnifications ``mags``
    (D, L), with the ring buffers held fixed.

    The geometry is the reference's (src/pyturb/extrude.py:458-469,807-822):
    ``sa = travel - base + h (thx cos + thy sin)``, ``sp = h (-thx sin + thy cos)``
    with h the layer altitude in pixels; the backward pass is
    ``pt_readout_adjoint``.
    """
    t = _hip.torch()
    dev = ext._buf.device
    L = ext.n_layers
    thx = t.as_tensor(thx, dtype=t.float64, device=dev).reshape(-1)
    thy = t.as_tensor(thy, dtype=t.float64, device=dev).reshape(-1)
    D = thx.numel()
    alt = t.tensor([ly.alt_pix for ly in ext.layers], dtype=t.float64, device=dev)
    cw = t.tensor([ly.cos for ly in ext.layers], dtype=t.float64, device=dev)
    sw = t.tensor([ly.sin for ly in ext.layers], dtype=t.float64, device=dev)
    base = t.tensor([ly.travel - ly.ring.base for ly in ext.layers], dtype=t.float64, device=dev)
    if travel_offset is not None:
        base = base + t.as_tensor(travel_offset, dtype=t.float64, device=dev)
    dxp = alt[None, :] * thx[:, None]
    dyp = alt[None, :] * thy[:, None]
    sa = base[None, :] + dxp * cw + dyp * sw
    sp = -dxp * sw + dyp * cw
    if mags is None:
        mg = t.tensor([ly.mag for ly in ext.layers], dtype=t.float64, device=dev).expand(D, L)
    else:
        mg = t.as_tensor(mags, dtype=t.float64, device=dev).expand(D, L)
    geom = t.stack([sa, sp, mg], dim=-1)
    return _ReadoutFn.apply(geom, ext, float(scale))


def _make_readout_fn():
    t = _hip.torch()
    from .infinite import readout, readout_adjoint

    class ReadoutFn(t.autograd.Function):
        """Extruder readout with pt_readout_adjoint as its backward."""

        @staticmethod
        def forward(ctx, geom, ext, scale):
            g = geom.detach().cpu().numpy()
            ctx.ext, ctx.scale, ctx.geom = ext, scale, g
            ctx.dev = geom.device
            return readout(ext._buf, ext.capacity, ext.width, [ly.ring for ly in ext.layers],
                           ext._wind_cs, g, ext.n, ext.interp, scale=scale)

        @staticmethod
        def backward(ctx, grad_out):
            ext = ctx.ext
            res = readout_adjoint(ext._buf, ext.capacity, ext.width, [ly.ring for ly in ext.layers],
                                  ext._wind_cs, ctx.geom, ext.n, ext.interp,
                                  grad_out.float().contiguous(), scale=ctx.scale)
            return res.to(ctx.dev), None, None

    return ReadoutFn


_ReadoutFn = _LazyFn(_make_readout_fn)
