"""Layered atmosphere on the MI355X: ``Atmosphere`` (spectral + extrude).

Drop-in for src/pyturb/atmosphere.py:108-1318 on the accelerated path:
construction and validation, the SeedSequence tree that fixes every layer's
realisation (atmosphere.py:393-470), ``opd(t, directions)``, ``frames``,
``evolve``, ``sample``, ``reset``, ``time_to_wrap`` and the wrap warning,
and the integrated quantities.

Device design:

* spectral engine -- the (L, n, n) complex64 layer spectra and the
  (P, L, 3, 3) complex128 subharmonic coefficients live in HBM; a frame, a
  set of directions or a whole batch of frame times is one
  ``pt_spectral_frames`` call (fused layer sum + 2-D inverse FFT +
  subharmonics + crop + OPD scale).  ``frames()`` computes frames in chunks
  and yields them in order; :meth:`opd_batch` exposes the batched call.
* extrude engine -- :class:`pyturb_amd.extrude.ExtrudedAtmosphere`.
* ``sample`` -- one fused ``pt_generate`` call per outer-scale group,
  accumulating into the cropped OPD output.

Spectral boiling (``tau_boil`` on ``engine="spectral"``, pt_boiled_frames),
extruder boiling (fresh windows on the matrix cores, pt_rows_blend) and the
spectral LGS cone (``lgs_altitude``, per-layer transforms and pt_lgs_zoom)
run on the device.  Chromatic dispersion models are outside this backend
and raise ``NotImplementedError``.
"""

from __future__ import annotations

import warnings
from typing import Any, Dict, Iterator, List, Optional, Sequence, Tuple

import numpy as np

from . import _hip, ops
from . import profiles as _profiles
from .backend import resolve_device
from .config import AtmosphereConfig
from .flow import draw_layer
from .fourier import PhaseScreen
from .profiles import Layer
from .utils import r0_at_wavelength, seeing_from_r0

__all__ = ["Atmosphere", "PeriodicWrapWarning", "ExtrudeBoilingPerformanceWarning"]

_ARCSEC_TO_RAD = np.pi / (180.0 * 3600.0)
_FRAME_CHUNK = 64  # first frozen-flow chunk of frames(); later chunks double
_FRAME_BUDGET = 1 << 30  # bytes of output frames per chunk at most (and 512 frames)
# per-layer screens of one LGS-cone batch (L x frames x ns^2 float32): a
# few GiB of the 288 GB HBM, so each batch amortises its per-layer launches
_LGS_BUDGET = 4 << 30


def _lgs_zoom_taps(n: int, n_screen: int, mag, interp: str):
    """Per-layer cone-zoom taps (atmosphere.py:581-616), FP64 on the host.

    The pupil grid scaled about the screen centre by each layer's
    magnification: positions ``(ns - 1)/2 + (i - (n - 1)/2) mag``, screen
    indices clipped to the screen, weights of ``interp`` (linear; Catmull-Rom
    of extrude.py:74-83; normalised Lanczos-3 of infinite.py:67-85).
    Returns idx (L, T, n) int32 and w (L, T, n) float32.
    """
    centre = 0.5 * (n_screen - 1)
    g = np.arange(n, dtype=np.float64) - 0.5 * (n - 1)
    idx, wts = [], []
    for m in mag:
        pos = centre + g * float(m)
        p0 = np.floor(pos)
        t = pos - p0
        if interp == "linear":
            offs, w = (0, 1), (1.0 - t, t)
        elif interp == "lanczos":
            offs = tuple(range(-2, 4))
            raw = [np.sinc(t - d) * np.sinc((t - d) / 3.0) for d in offs]
            tot = sum(raw[1:], raw[0])
            w = tuple(r / tot for r in raw)
        else:
            t2, t3 = t * t, t * t * t
            offs = (-1, 0, 1, 2)
            w = (0.5 * (-t + 2.0 * t2 - t3), 0.5 * (2.0 - 5.0 * t2 + 3.0 * t3),
                 0.5 * (t + 4.0 * t2 - 3.0 * t3), 0.5 * (t3 - t2))
        idx.append(np.stack([np.clip(p0 + o, 0, n_screen - 1) for o in offs]))
        wts.append(np.stack(w))
    idx = np.stack(idx).astype(np.int32)
    # the device zoom stages windows [idx[:, 0, i0], idx[:, -1, i1]]: both
    # axes must be non-decreasing
    if np.any(np.diff(idx, axis=1) < 0) or np.any(np.diff(idx, axis=2) < 0):
        raise ValueError("LGS zoom taps are not monotone")
    return idx, np.stack(wts).astype(np.float32)
_BOIL_CHUNK = 128  # boiled frames per device call


class PeriodicWrapWarning(UserWarning):
    """The spectral engine's periodic screen has wrapped for some layer."""


class ExtrudeBoilingPerformanceWarning(UserWarning):
    """``tau_boil`` on ``engine="extrude"`` is much slower than on "spectral":
    every step re-extrudes a fresh window row by row (atmosphere.py:69-86)."""


class _LayerState:
    __slots__ = ("generator", "vx", "vy", "altitude_los")

    def __init__(self, generator, vx, vy, altitude_los):
        self.generator = generator
        self.vx = vx
        self.vy = vy
        self.altitude_los = altitude_los


class Atmosphere:
    """A multi-layer turbulent atmosphere producing pupil OPD (metres).

    Same parameters as the reference ``pyturb.Atmosphere``; ``device``
    defaults to ``"mi355x"``.  Outputs are float32 device tensors (float64
    when ``dtype="float64"``, computed in float32).
    """

    def __init__(self, layers: Sequence[Layer], r0: Optional[float] = None,
                 seeing: Optional[float] = None, wavelength: float = 500e-9,
                 zenith_angle: float = 0.0, diameter: float = 8.0, n: int = 512,
                 L0: Optional[float] = None, power_law: float = 11.0 / 3.0,
                 inner_scale: float = 0.0, subharmonics: int = 8,
                 field_of_view: float = 0.0, tau_boil=None, engine: str = "spectral",
                 interp: str = "cubic", lgs_altitude: Optional[float] = None,
                 dispersion: Optional[str] = None, wet_fraction: float = 0.0,
                 device: str = "mi355x", dtype: str = "float32",
                 seed: Optional[int] = None, far_points: int = 0) -> None:
        if far_points and engine != "extrude":
            raise ValueError("far_points applies to engine='extrude' only")
        config = AtmosphereConfig.create(
            layers, r0, seeing, wavelength, zenith_angle, diameter, n, L0, power_law,
            inner_scale, subharmonics, field_of_view, tau_boil, engine, interp,
            lgs_altitude, dispersion, wet_fraction, device, dtype, seed)
        if config.engine == "extrude" and config.has_tau_boil:
            warnings.warn(
                "boiling (tau_boil) on engine='extrude' re-extrudes an independent fresh screen "
                "every step and is markedly slower than the spectral engine's per-mode AR(1) "
                "boiling. Prefer engine='spectral' for boiling-heavy runs unless genuine "
                "non-periodicity is required.", ExtrudeBoilingPerformanceWarning, stacklevel=2)
        self.config = config
        self.engine = config.engine
        self.interp = config.interp
        self.lgs_altitude = config.lgs_altitude
        self.dispersion = config.dispersion
        self.wet_fraction = config.wet_fraction
        self.wavelength = config.wavelength
        self.r0_zenith = config.r0_zenith
        self.zenith_angle = config.zenith_angle
        self.diameter = config.diameter
        self.n = config.grid.n
        self.subharmonics = config.subharmonics
        self.power_law = config.power_law
        self.inner_scale = config.inner_scale
        self.field_of_view = config.field_of_view
        self.device = config.grid.device
        self.dtype = config.grid.dtype.name
        self.pixel_scale = config.grid.pixel_scale
        self._device_index = resolve_device(self.device)

        cos_z = np.cos(np.deg2rad(self.zenith_angle))
        self.airmass = 1.0 / cos_z
        self.r0_los = self.r0_zenith * cos_z ** (3.0 / 5.0)
        self.layers = [ly.to_layer() for ly in config.layers]
        max_alt = max(ly.altitude for ly in self.layers) * self.airmass
        margin_pix = int(np.ceil(max_alt * np.tan(self.field_of_view * _ARCSEC_TO_RAD)
                                 / self.pixel_scale))
        self.margin_pix = margin_pix
        self.n_screen = self.n + 2 * margin_pix
        self._crop = slice(margin_pix, margin_pix + self.n)
        self._screen_period_m = self.n_screen * self.pixel_scale
        self._wrap_warned = False
        self.tau_boil = np.asarray(config.tau_boil, dtype=np.float64)
        self.seed = config.seed
        self._profile_name: Optional[str] = None

        if self.lgs_altitude is not None:
            lgs_los = self.lgs_altitude * self.airmass
            mag = 1.0 - np.array([ly.altitude * self.airmass for ly in self.layers]) / lgs_los
            if np.any(mag <= 0.0):
                raise ValueError(
                    f"lgs_altitude ({self.lgs_altitude:.0f} m) must exceed every layer "
                    "altitude (projected to the line of sight)")
        # per-layer cone magnification of the spectral engine (atmosphere.py:484-498)
        self._lgs_mag = mag if (self.lgs_altitude is not None and self.engine == "spectral") \
            else None

        # Seed tree (atmosphere.py:393-437): master -> per-layer children,
        # two boiling streams, then (gen, flow, ext) per layer.
        master = np.random.SeedSequence(self.seed)
        children = master.spawn(len(self.layers))
        # spectral boiling stream, then the extruder's (atmosphere.py:395-397)
        self._boil_seed = int(master.spawn(1)[0].generate_state(1)[0])
        self._ext_boil_seed = int(master.spawn(1)[0].generate_state(1)[0])
        self._boiling = self.engine == "spectral" and bool(np.any(np.isfinite(self.tau_boil)))
        # injected boiling draws (tests): step index -> (main (2, L, ns, ns),
        # subharmonic (2, P, L, 3, 3)) host arrays, replacing the device stream
        self._boil_noise = None
        self._epoch = 0
        self._layers: List[_LayerState] = []
        ext_r0, ext_L0, ext_wind, ext_alt, ext_seeds, flow_seeds = [], [], [], [], [], []
        for layer, child in zip(self.layers, children):
            r0_i = self.r0_los * layer.cn2_fraction ** (-3.0 / 5.0)
            gen_seed, flow_seed, ext_seed = child.spawn(3)
            gen = None
            if self.engine == "spectral":
                gen = PhaseScreen(self.n_screen, self.pixel_scale, r0_i, layer.L0,
                                  self.subharmonics, self.power_law, self.inner_scale,
                                  seed=int(gen_seed.generate_state(1)[0]),
                                  device=self.device, dtype=self.dtype)
            vx, vy = layer.wind_vector
            alt = layer.altitude * self.airmass
            self._layers.append(_LayerState(gen, vx, vy, alt))
            ext_r0.append(r0_i)
            ext_L0.append(layer.L0)
            ext_wind.append((vx, vy))
            ext_alt.append(alt)
            ext_seeds.append(int(ext_seed.generate_state(1)[0]))
            flow_seeds.append(int(flow_seed.generate_state(1)[0]))
        self._layer_r0 = np.array(ext_r0)

        # sample() generators: one per distinct round(L0, 9) (atmosphere.py:439-470)
        groups: Dict[float, List[int]] = {}
        for i, ly in enumerate(self.layers):
            groups.setdefault(round(float(ly.L0), 9), []).append(i)
        self._sample_specs = []
        for idx, child in zip(groups.values(), master.spawn(len(groups))):
            if len(idx) == 1:
                self._sample_specs.append(("layer", idx[0], None, None))
            else:
                r0_agg = sum(ext_r0[i] ** (-5.0 / 3.0) for i in idx) ** (-3.0 / 5.0)
                self._sample_specs.append(
                    ("group", self.layers[idx[0]].L0, r0_agg, int(child.generate_state(1)[0])))
        self._sample_generators: Optional[List[PhaseScreen]] = None
        self._group_filter_cache = None
        self._flow_seeds = flow_seeds

        self._t = 0.0
        self._vx = np.array([s.vx for s in self._layers])
        self._vy = np.array([s.vy for s in self._layers])
        self._alt = np.array([s.altitude_los for s in self._layers])
        if self.engine == "spectral":
            self._build_batched()
        else:
            from .extrude import ExtrudedAtmosphere

            fov_pix = [a * np.tan(self.field_of_view * _ARCSEC_TO_RAD) / self.pixel_scale
                       for a in ext_alt]
            self._ext_kwargs = dict(
                n=self.n, pixel_scale=self.pixel_scale, layer_r0=ext_r0, layer_L0=ext_L0,
                layer_wind=ext_wind, layer_altitude_los=ext_alt, field_of_view_pix=fov_pix,
                interp=self.interp,
                lgs_altitude_los=(None if self.lgs_altitude is None
                                  else self.lgs_altitude * self.airmass),
                device=self.device, dtype=self.dtype, seeds=ext_seeds,
                tau_boil=(list(self.tau_boil) if config.has_tau_boil else None),
                boil_seed=self._ext_boil_seed, far_points=far_points)
            self._ext = ExtrudedAtmosphere(**self._ext_kwargs)

    # ------------------------------------------------------------------
    def _build_batched(self) -> None:
        """Upload the stacked layer spectra and subharmonic coefficients."""
        L = len(self._layers)
        ns = self.n_screen
        P = self.subharmonics
        spectra = np.empty((L, ns, ns), dtype=np.complex64)
        coeffs = np.zeros((P, L, 3, 3), dtype=np.complex128)
        for l, (state, fseed) in enumerate(zip(self._layers, self._flow_seeds)):
            spec, cf = draw_layer(np.random.default_rng(fseed), state.generator)
            spectra[l] = spec
            coeffs[:, l] = cf
        dev = _hip.device(self._device_index)
        self._dev = dev
        self._spectra = _hip.to_device(spectra, None, dev)
        self._sh_coeffs = _hip.to_device(coeffs, None, dev) if P else None
        self._spectra_host_c64 = None
        self._boil_index = 0
        if self._lgs_mag is not None:
            idx, w = _lgs_zoom_taps(self.n, ns, self._lgs_mag, self.interp)
            self._zoom_idx = _hip.to_device(idx, None, dev)
            self._zoom_w = _hip.to_device(w, None, dev)
        if self._boiling:
            t = _hip.torch()
            gens = [st.generator for st in self._layers]
            self._boil_amp = t.stack([g._amplitude for g in gens]).contiguous()
            self._boil_sh_amp = (t.stack([g._sh_amp for g in gens], dim=1).contiguous()
                                 if P else None)
            df = 1.0 / (ns * self.pixel_scale)
            L0s = np.array([float(ly.L0) for ly in self.layers])
            f_ref = np.where(np.isfinite(L0s) & (L0s > 0), 1.0 / np.where(L0s > 0, L0s, 1.0), df)
            with np.errstate(divide="ignore"):
                inv_tau = np.where(np.isfinite(self.tau_boil), 1.0 / self.tau_boil, 0.0)
            self._boil_f_ref = _hip.to_device(f_ref, np.float64, dev)
            self._boil_inv_tau = _hip.to_device(inv_tau, np.float64, dev)

    @classmethod
    def from_profile(cls, name: str, **kwargs) -> "Atmosphere":
        atm = cls(_profiles.get_profile(name), **kwargs)
        atm._profile_name = str(name).lower()
        return atm

    # ------------------------------------------------------------------ properties
    def r0_at(self, wavelength: float) -> float:
        return r0_at_wavelength(self.r0_los, self.wavelength, wavelength)

    @property
    def r0(self) -> float:
        return self.r0_los

    @property
    def seeing(self) -> float:
        return seeing_from_r0(self.r0_los, self.wavelength)

    @property
    def theta0(self) -> float:
        los = [Layer(ly.altitude * self.airmass, ly.cn2_fraction, ly.wind_speed,
                     ly.wind_direction, ly.L0) for ly in self.layers]
        return _profiles.isoplanatic_angle(los, self.r0_los) / _ARCSEC_TO_RAD

    @property
    def tau0(self) -> float:
        return _profiles.coherence_time(self.layers, self.r0_los)

    @property
    def greenwood_frequency(self) -> float:
        return _profiles.greenwood_frequency(self.layers, self.r0_los)

    @property
    def time_to_wrap(self) -> float:
        if self.engine != "spectral":
            return float("inf")
        speeds = np.array([ly.wind_speed for ly in self.layers])
        with np.errstate(divide="ignore"):
            periods = self._screen_period_m / speeds
        finite = periods[np.isfinite(periods)]
        return float(finite.min()) if finite.size else float("inf")

    @property
    def time(self) -> float:
        return self._t

    def _check_wrap(self, disp_x, disp_y, stacklevel: int = 3) -> None:
        """Warn once per run the first time a layer's travel passes the
        screen period (atmosphere.py:732-751); reset() re-arms it."""
        disp = np.hypot(np.asarray(disp_x), np.asarray(disp_y))
        wrapped = disp > self._screen_period_m
        if not np.any(wrapped):
            return
        self._wrap_warned = True
        per_layer = disp.reshape(-1, len(self.layers))
        idx = int(np.argmax((per_layer > self._screen_period_m).any(axis=0)))
        layer = self.layers[idx]
        travel = float(per_layer[:, idx].max())
        warnings.warn(
            f"engine='spectral' layer {idx} (altitude={layer.altitude:.0f} m, "
            f"wind_speed={layer.wind_speed:.1f} m/s) has wrapped: its cumulative travel "
            f"({travel:.2f} m) exceeds the screen's period ({self._screen_period_m:.2f} m). "
            "Its 'new' turbulence is a repeat of previous turbulence, biasing temporal "
            "statistics over this run. Use engine='extrude' for non-periodic frozen flow, "
            "or keep run duration below `atm.time_to_wrap` seconds.",
            PeriodicWrapWarning, stacklevel=stacklevel)

    def _wrap_check_at(self, t: float) -> None:
        """The on-axis wrap check of one frame at time ``t``, made when that
        frame is handed out (atmosphere.py:1146-1149 checks inside
        ``_phase``, frame by frame)."""
        if not self._wrap_warned and self.engine == "spectral":
            self._check_wrap(self._vx * t, self._vy * t, stacklevel=4)

    # ------------------------------------------------------------------ output
    def _scale(self, wavelength: Optional[float]) -> float:
        s = self.wavelength / (2.0 * np.pi)
        if wavelength is None:
            return s
        if self.dispersion is not None:
            raise NotImplementedError("dispersion models are not provided by the MI355X backend")
        return s * (2.0 * np.pi / float(wavelength))

    def _cast(self, x):
        return x.double() if self.dtype == "float64" else x

    def _dirs_tan(self, directions) -> Tuple[np.ndarray, np.ndarray]:
        oxs, oys = [], []
        for thx, thy in directions:
            radius = float(np.hypot(thx, thy))
            if radius > self.field_of_view:
                raise ValueError(
                    f"direction ({thx}, {thy}) arcsec has radius {radius:.3g} arcsec, "
                    f"exceeding the declared field_of_view={self.field_of_view} arcsec")
            oxs.append(np.tan(thx * _ARCSEC_TO_RAD))
            oys.append(np.tan(thy * _ARCSEC_TO_RAD))
        return np.array(oxs), np.array(oys)

    def _spectral(self, times, oxs, oys, scale, check_wrap=True):
        """(T*D) frames for every (time, direction) pair, one device call.

        ``check_wrap=False`` leaves the wrap warning to the caller, which
        checks each frame as it hands it out (frames())."""
        times = np.atleast_1d(np.asarray(times, dtype=np.float64))
        dx = self._vx[None, None, :] * times[:, None, None] + self._alt * oxs[None, :, None]
        dy = self._vy[None, None, :] * times[:, None, None] + self._alt * oys[None, :, None]
        if check_wrap and not self._wrap_warned:
            self._check_wrap(dx, dy, stacklevel=4)
        disp = np.stack([dx, dy], axis=-1).reshape(-1, len(self._layers), 2)
        if self._lgs_mag is not None:
            return self._lgs_frames(disp, scale)
        disp_dev = _hip.to_device(disp, np.float64, self._dev)
        return ops.spectral_frames(self._spectra, disp_dev, self._sh_coeffs, self.pixel_scale,
                                   crop_offset=self.margin_pix, n_out=self.n, scale=scale)

    def _lgs_frames(self, disp, scale):
        """LGS-cone frames (atmosphere.py:1000-1090): each layer transformed
        on its own over a chunk of frames (pt_spectral_frames with one
        layer, subharmonics and their mean included), then the cone zoom and
        the layer sum (pt_lgs_zoom)."""
        t = _hip.torch()
        F, L, ns = disp.shape[0], len(self._layers), self.n_screen
        out = t.empty((F, self.n, self.n), dtype=t.float32, device=self._dev)
        k = max(1, min(F, 65535, _LGS_BUDGET // (4 * L * ns * ns)))
        screens = t.empty((L, k, ns, ns), dtype=t.float32, device=self._dev)
        for f0 in range(0, F, k):
            m = min(k, F - f0)
            for l in range(L):
                d = _hip.to_device(np.ascontiguousarray(disp[f0:f0 + m, l:l + 1]), np.float64,
                                   self._dev)
                cf = (self._sh_coeffs[:, l:l + 1].contiguous() if self._sh_coeffs is not None
                      else None)
                ops.spectral_frames(self._spectra[l:l + 1], d, cf, self.pixel_scale,
                                    n_out=ns, out=screens[l, :m])
            ops.lgs_zoom(screens[:, :m] if m == k else screens[:, :m].contiguous(),
                         self._zoom_idx, self._zoom_w, self.n, scale=scale, out=out[f0:f0 + m])
        return out

    def opd(self, t: float = 0.0, directions=None, wavelength: Optional[float] = None) -> Any:
        """OPD [m] (or phase at ``wavelength``) at time ``t`` [s]."""
        scale = self._scale(wavelength)
        if directions is None:
            return self._cast(self._phase_scaled(float(t), 0.0, 0.0, scale))
        oxs, oys = self._dirs_tan(directions)
        if self.engine == "spectral":
            return self._cast(self._spectral([float(t)], oxs, oys, scale))
        self._ext.set_time(float(t))
        return self._cast(self._ext.integrate_dirs(oxs, oys, scale=scale))

    def _phase_scaled(self, t, ox, oy, scale):
        if self.engine == "spectral":
            return self._spectral([t], np.array([ox]), np.array([oy]), scale)[0]
        self._ext.set_time(t)
        return self._ext.integrate_dirs(np.array([ox]), np.array([oy]), scale=scale)[0]

    def _phase(self, t: float, ox: float, oy: float):
        """Reference-wavelength phase [rad] (atmosphere.py:871-880)."""
        return self._phase_scaled(t, ox, oy, 1.0)

    def opd_batch(self, times: Sequence[float], directions=None,
                  wavelength: Optional[float] = None) -> Any:
        """Spectral frames for many times in one call: (T, n, n) or (T, D, n, n).

        Equivalent to stacking ``opd(t, directions)`` over ``times``; random
        access in time is what the spectral engine allows.
        """
        if self.engine != "spectral":
            raise ValueError("opd_batch needs engine='spectral' (the extruder streams)")
        scale = self._scale(wavelength)
        times = np.asarray(times, dtype=np.float64).ravel()
        if directions is None:
            return self._cast(self._spectral(times, np.zeros(1), np.zeros(1), scale))
        oxs, oys = self._dirs_tan(directions)
        out = self._spectral(times, oxs, oys, scale)
        return self._cast(out.view(len(times), len(oxs), self.n, self.n))

    def frames(self, dt: float, steps: int,
               wavelength: Optional[float] = None) -> Iterator[Tuple[float, Any]]:
        """Yield ``(t, opd)`` for ``steps`` frames spaced by ``dt`` seconds."""
        if dt <= 0:
            raise ValueError("dt must be positive")
        if steps < 1:
            raise ValueError("steps must be >= 1")
        dt = float(dt)
        scale = self._scale(wavelength)
        if self._boiling:
            yield from self._boiled_frames(dt, int(steps), scale)
            return
        if self.engine == "extrude":
            # frame, then advance the clock and boil once (atmosphere.py:1146-1154)
            for _ in range(int(steps)):
                t0 = self._t
                yield t0, self._cast(self._phase_scaled(t0, 0.0, 0.0, scale))
                self._t += dt
                self._ext.boil_step(dt)  # no-op unless tau_boil is finite
            return
        left = int(steps)
        # chunks of 64, 128, 256, ... frames up to a memory-sized cap: the
        # first frame comes quickly, long runs amortise each call's fixed
        # work (fold, tables, launches) over up to 512 frames
        cap = max(_FRAME_CHUNK, min(512, _FRAME_BUDGET // (4 * self.n * self.n)))
        size = _FRAME_CHUNK
        while left > 0:
            k = min(left, size)
            size = min(2 * size, cap)
            times = []
            t = self._t
            for _ in range(k):
                times.append(t)
                t += dt
            chunk = self._cast(self._spectral(times, np.zeros(1), np.zeros(1), scale,
                                              check_wrap=False))
            for i in range(k):
                if self._t != times[i]:  # clock moved under us (reset): recompute
                    break
                self._wrap_check_at(times[i])
                yield times[i], chunk[i]
                self._t += dt
                left -= 1
            else:
                continue

    def evolve(self, dt: float, wavelength: Optional[float] = None) -> Any:
        if dt <= 0:
            raise ValueError("dt must be positive")
        self._t += float(dt)
        self._boil_step(float(dt))  # no-op unless tau_boil is finite
        if self.engine == "extrude":
            self._ext.boil_step(float(dt))
        return self._cast(self._phase_scaled(self._t, 0.0, 0.0, self._scale(wavelength)))

    # ------------------------------------------------------------------ boiling
    def _boil_call(self, times, dt: float, steps: int, scale: float = 1.0):
        """Frames at ``times`` (each followed by a boiling step), then
        ``steps - len(times)`` more steps; advances the boiling state."""
        t = _hip.torch()
        F = len(times)
        disp = None
        if F:
            times = np.asarray(times, dtype=np.float64)
            dx = self._vx[None, :] * times[:, None]
            dy = self._vy[None, :] * times[:, None]
            disp = _hip.to_device(np.stack([dx, dy], axis=-1), np.float64, self._dev)
        noise = sh_noise = None
        if self._boil_noise is not None:
            draws = [self._boil_noise(self._boil_index + k) for k in range(steps)]
            noise = _hip.to_device(np.stack([np.asarray(d[0]) for d in draws]), np.float32,
                                   self._dev)
            if self._sh_coeffs is not None:
                sh_noise = _hip.to_device(np.stack([np.asarray(d[1]) for d in draws]),
                                          np.float64, self._dev)
        out = ops.boiled_frames(
            self._spectra, self._boil_amp, self._boil_inv_tau, self._boil_f_ref, disp, dt,
            self.pixel_scale, sh_coeffs=self._sh_coeffs, sh_amp=self._boil_sh_amp, steps=steps,
            noise=noise, sh_noise=sh_noise, seed=self._boil_seed, step0=self._boil_index,
            crop_offset=self.margin_pix, n_out=self.n, scale=scale)
        self._boil_index += steps
        del t
        return out

    def _boil_step(self, dt: float) -> None:
        """One AR(1) boiling step of every mode (atmosphere.py:1092-1126)."""
        if self._boiling:
            self._boil_call([], dt, 1)

    def _boiled_frames(self, dt: float, steps: int, scale: float):
        """frames() with boiling: chunks of frames from one device call each.

        A chunk advances the state by one step per frame.  If the caller
        stops after fewer frames, the state is put back to the chunk's start
        and advanced by the steps actually taken; the device stream is
        counted by step, so this reproduces the state exactly.  A reset()
        while a frame is out takes effect as in the reference: the clock
        then advances and the fresh state boils once.
        """
        if self._lgs_mag is not None:
            # the cone zooms every layer on its own, so each frame is an LGS
            # frame of the current state followed by one boiling step
            for _ in range(steps):
                t0 = self._t
                yield t0, self._cast(self._spectral([t0], np.zeros(1), np.zeros(1), scale)[0])
                self._t += dt
                self._boil_step(dt)
            return
        left = steps
        while left > 0:
            k = 1 if self._boil_noise is not None else min(left, _BOIL_CHUNK)
            epoch = self._epoch
            snap = (self._spectra.clone(),
                    None if self._sh_coeffs is None else self._sh_coeffs.clone(),
                    self._boil_index)
            times = [self._t + i * dt for i in range(k)]
            chunk = self._cast(self._boil_call(times, dt, k, scale))
            used = 0
            try:
                for i in range(k):
                    self._wrap_check_at(times[i])
                    yield times[i], chunk[i]
                    if self._epoch != epoch:  # reset while the frame was out
                        break
                    used += 1
                    self._t += dt
                    left -= 1
            finally:
                if used < k and self._epoch == epoch:
                    self._spectra.copy_(snap[0])
                    if snap[1] is not None:
                        self._sh_coeffs.copy_(snap[1])
                    self._boil_index = snap[2]
                    if used:
                        self._boil_call([], dt, used)
            if self._epoch != epoch:
                self._t += dt
                self._boil_step(dt)
                left -= 1

    # ------------------------------------------------------------------ Monte-Carlo
    def _generators(self) -> List[PhaseScreen]:
        if self._sample_generators is None:
            gens = []
            for kind, a, r0_agg, seed in self._sample_specs:
                if kind == "layer":
                    st = self._layers[a]
                    if st.generator is None:  # extrude engine: build lazily, same seed tree
                        st.generator = self._layer_generator(a)
                    gens.append(st.generator)
                else:
                    gens.append(PhaseScreen(self.n_screen, self.pixel_scale, r0_agg, a,
                                            self.subharmonics, self.power_law,
                                            self.inner_scale, seed=seed, device=self.device,
                                            dtype=self.dtype))
            self._sample_generators = gens
        return self._sample_generators

    def _layer_generator(self, i: int) -> PhaseScreen:
        master = np.random.SeedSequence(self.seed)
        child = master.spawn(len(self.layers))[i]
        gen_seed = child.spawn(3)[0]
        return PhaseScreen(self.n_screen, self.pixel_scale, self._layer_r0[i],
                           self.layers[i].L0, self.subharmonics, self.power_law,
                           self.inner_scale, seed=int(gen_seed.generate_state(1)[0]),
                           device=self.device, dtype=self.dtype)

    def sample(self, count: Optional[int] = None, wavelength: Optional[float] = None,
               *, noises=None) -> Any:
        """Independent integrated OPDs: one fused generate per L0 group.

        ``noises``: optional list of ``(noise, sh_noise)`` device tensors per
        group (injected noise) instead of the host draw.
        """
        squeeze = count is None
        cnt = 1 if squeeze else int(count)
        if cnt < 1:
            raise ValueError("count must be >= 1")
        scale = self._scale(wavelength)
        gens = self._generators()
        if len(gens) == 1:
            gen = gens[0]
            if noises is None:
                main, sh = gen.draw_noise(cnt)
                noise, shn = gen._upload_noise(main, sh)
            else:
                noise, shn = noises[0]
            out = ops.generate(noise, shn, gen._amplitude, gen._sh_amp, cnt,
                               crop_offset=self.margin_pix, n_out=self.n, scale=scale)
        else:
            # all L0 groups in one pass: their filtered noise is summed before a
            # single transform (the groups' draws keep their own streams)
            t = _hip.torch()
            dev = _hip.device(self._device_index)
            G, n, P = len(gens), self.n_screen, self.subharmonics
            n_fft = (cnt + 1) // 2
            noise = t.empty((G, 2, n_fft, n, n), dtype=t.float32, device=dev)
            shn = t.empty((G, P, 2, cnt, 3, 3), dtype=t.float32, device=dev) if P else None
            for g, gen in enumerate(gens):
                if noises is None:
                    main, sh = gen.draw_noise(cnt)
                    noise[g].copy_(t.from_numpy(np.ascontiguousarray(main, dtype=np.float32)))
                    if P:
                        shn[g].copy_(t.from_numpy(np.ascontiguousarray(sh, dtype=np.float32)))
                else:
                    noise[g].copy_(noises[g][0])
                    if P:
                        shn[g].copy_(noises[g][1])
            amp, sh_amp = self._group_filters()
            out = ops.generate_groups(noise, shn, amp, sh_amp, cnt, crop_offset=self.margin_pix,
                                      n_out=self.n, scale=scale)
        out = self._cast(out)
        return out[0] if squeeze else out

    def _group_filters(self):
        """Stacked device filters of the Monte-Carlo groups (built once)."""
        if self._group_filter_cache is None:
            t = _hip.torch()
            gens = self._generators()
            amp = t.stack([g._amplitude for g in gens]).contiguous()
            sh = t.stack([g._sh_amp for g in gens]).contiguous() if self.subharmonics else None
            self._group_filter_cache = (amp, sh)
        return self._group_filter_cache

    def reset(self) -> "Atmosphere":
        self._t = 0.0
        self._wrap_warned = False
        self._epoch += 1
        if self.engine == "extrude":
            from .extrude import ExtrudedAtmosphere

            self._ext = ExtrudedAtmosphere(**self._ext_kwargs)
        elif self._boiling:
            # boiling changed the stored turbulence: rebuild it and rewind the stream
            self._build_batched()
        return self

    @property
    def metadata(self) -> Dict[str, Any]:
        l0s = {round(float(ly.L0), 9) for ly in self.layers}
        return {
            "units": "metres", "pixel_scale": self.pixel_scale, "diameter": self.diameter,
            "n": self.n, "r0": self.r0_los,
            "L0": float(self.layers[0].L0) if len(l0s) == 1 else None,
            "wavelength": self.wavelength, "seeing": self.seeing, "theta0": self.theta0,
            "tau0": self.tau0, "zenith_angle": self.zenith_angle,
            "n_layers": len(self.layers), "engine": self.engine, "interp": self.interp,
            "power_law": self.power_law, "inner_scale": self.inner_scale,
            "subharmonics": self.subharmonics, "field_of_view": self.field_of_view,
            # one rate for every layer, else None (atmosphere.py:1270-1274)
            "tau_boil": (float(self.tau_boil[0]) if np.all(np.isfinite(self.tau_boil))
                         and np.all(self.tau_boil == self.tau_boil[0]) else None),
            "dispersion": self.dispersion,
            "wet_fraction": self.wet_fraction,
            "lgs_altitude": 0.0 if self.lgs_altitude is None else float(self.lgs_altitude),
            "device": self.device, "dtype": self.dtype, "seed": self.seed, "time": self._t,
            "profile": self._profile_name,
        }

    def __repr__(self) -> str:
        return (f"Atmosphere(layers={len(self.layers)}, r0={self.r0_los:.3f} m, "
                f"seeing={self.seeing:.2f}\", n={self.n}, diameter={self.diameter} m, "
                f"device={self.device!r})")
