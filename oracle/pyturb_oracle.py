"""FP64 NumPy restatement of the reference phase-screen hot path.

TEST INFRASTRUCTURE ONLY.  This module is the oracle the HIP library is
checked against; it is imported by ``tests/``, ``__graft_entry__.smoke`` and
the ``cpu_baseline`` leg of ``bench.py`` and by nothing in ``pyturb_amd``.

Every function names the reference file:line it restates (paths relative to
the reference checkout, ``src/pyturb/...``).  All arithmetic is float64 /
complex128; noise and innovations are *injected* as arrays so the GPU and
the oracle consume identical inputs.

Pinning (how each function is checked, see ``tests/test_oracle.py``):

* ``psd``, ``band_amplitude``, ``build_filters``: golden vectors written from
  the reference (``tests/golden/make_golden.py``) plus the closed form of the
  PSD and the flat-PSD limit of the band amplitude.
* ``ifft2_unnormalised``: brute-force DFT on tiny grids.
* ``generate``: reference golden screens (seed=1, n=256 subset + checksums),
  and the real/imag pairing identity.
* ``integrate`` / ``integrate_dirs`` / ``flow_translate``: reference golden
  frames (9-layer n=64 atmosphere), the integer-pixel ``np.roll`` identity and
  the layer-sum == sum of single-layer translates identity.
* ``phase_covariance``: C(0) = 0.0863 (L0/r0)^(5/3) closed form and golden.
* ``build_extrusion``: ``A C_zz = C_xz`` and ``B B^T = C_xx - A C_xz^T``
  identities and golden A/B at W=32.
* ``compact_row_ring`` / ``OracleExtruder`` / ``readout``: golden readouts
  from a 2-layer n=32 extruder, exactness at integer travel and
  "one big jump == many small steps".
* ``noll_to_zernike`` / ``zernike_basis`` / ``zernike_decompose`` /
  ``noll_residual`` / ``noll_variance``: the reference's known-answer index
  table, orthonormality over the pupil, exact recovery of a synthetic
  combination of modes, the radial polynomials' closed forms, and Noll's
  table values with the large-J asymptote.
"""

from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
from scipy import linalg as _sla
from scipy.special import gamma as _gamma
from scipy.special import kv as _kv

ARCSEC = np.pi / (180.0 * 3600.0)

__all__ = [
    "ARCSEC",
    "psd",
    "band_amplitude",
    "build_filters",
    "draw_generate_noise",
    "ifft2_unnormalised",
    "set_fft_workers",
    "generate",
    "subharmonic_field",
    "draw_flow",
    "flow_translate",
    "integrate",
    "integrate_dirs",
    "lgs_zoom_taps",
    "integrate_lgs",
    "OracleAtmosphere",
    "phase_covariance",
    "build_extrusion",
    "compact_row_ring",
    "far_stencil",
    "catmull_rom_weights",
    "lanczos_weights",
    "readout",
    "pupil_grids",
    "along_extent_and_span",
    "OracleExtruder",
    "extrusion_seed_factor",
    "OracleInfinite",
    "structure_function",
    "r0_from_seeing",
    "rel_rms",
    "noll_to_zernike",
    "zernike_basis",
    "zernike_decompose",
    "noll_residual",
    "noll_variance",
    "device_boil_noise",
]


# ---------------------------------------------------------------------------
# small helpers
# ---------------------------------------------------------------------------
def rel_rms(test, ref) -> float:
    """Relative RMS error ``||test-ref||_2 / ||ref||_2`` in float64."""
    test = np.asarray(test, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = float(np.sqrt(np.mean(ref * ref)))
    return float(np.sqrt(np.mean((test - ref) ** 2))) / (den if den > 0 else 1.0)


def r0_from_seeing(seeing: float, wavelength: float = 500e-9) -> float:
    """utils.py:137-142 -- FWHM = 0.98 lambda / r0."""
    return 0.98 * wavelength / (seeing * ARCSEC)


def structure_function(phase, pixel_scale=1.0, max_separation=None):
    """utils.py:159-201 -- axis-averaged structure function estimate."""
    ph = np.asarray(phase, dtype=np.float64)
    if ph.ndim == 2:
        ph = ph[None]
    n = min(ph.shape[-2:])
    smax = n // 4 if max_separation is None else int(max_separation)
    out = np.empty(smax)
    for k, s in enumerate(range(1, smax + 1)):
        d0 = ph[..., s:, :] - ph[..., :-s, :]
        d1 = ph[..., :, s:] - ph[..., :, :-s]
        out[k] = 0.5 * (np.mean(d0 * d0) + np.mean(d1 * d1))
    return np.arange(1, smax + 1) * pixel_scale, out


# ---------------------------------------------------------------------------
# PSD, band amplitudes, filters  (fourier.py:131-228)
# ---------------------------------------------------------------------------
def psd(f_squared, r0, L0, power_law=11.0 / 3.0, inner_scale=0.0):
    """fourier.py:131-149 -- modified von Karman phase PSD [rad^2 m^2]."""
    f2 = np.asarray(f_squared, dtype=np.float64)
    k0 = 0.0 if np.isinf(L0) else L0 ** -2.0
    val = 0.023 * r0 ** (-5.0 / 3.0) * (f2 + k0) ** (-0.5 * power_law)
    if inner_scale > 0.0:
        fm = 5.92 / (2.0 * np.pi * inner_scale)
        val = val * np.exp(-f2 / (fm * fm))
    return val


def band_amplitude(fx, fy, width, r0, L0, power_law=11.0 / 3.0,
                   inner_scale=0.0, samples=33):
    """fourier.py:151-172 -- cell-integrated mode amplitude.

    Midpoint quadrature of the PSD over a ``width``-wide square cell centred
    on ``(fx, fy)``: ``amp^2 = mean(PSD) w^2 * <f^2>_PSD / |f_c|^2``.
    """
    u = (np.arange(samples) + 0.5) / samples - 0.5
    ax = fx + u * width
    ay = fy + u * width
    f2 = ax[:, None] ** 2 + ay[None, :] ** 2
    p = psd(f2, r0, L0, power_law, inner_scale)
    power = p.mean() * width * width
    f2_eff = float(np.sum(p * f2) / np.sum(p))
    return math.sqrt(power * f2_eff / (fx * fx + fy * fy))


def build_filters(n, dx, r0, L0, subharmonics=8, power_law=11.0 / 3.0,
                  inner_scale=0.0):
    """fourier.py:174-228 -- main-grid amplitude and subharmonic bases.

    Returns a dict with ``amplitude`` (n,n), ``f`` (n,), ``sh_amp`` (P,3,3),
    ``sh_freq`` (P,3) and ``sh_basis`` (P,3,n) complex, all float64.
    """
    df = 1.0 / (n * dx)
    f = np.fft.fftfreq(n, d=dx)
    f2 = f[:, None] ** 2 + f[None, :] ** 2
    f2[0, 0] = 1.0
    amp = np.sqrt(psd(f2, r0, L0, power_law, inner_scale)) * df
    amp[0, 0] = 0.0
    kmax = min(8, (n - 1) // 2)
    for a in range(-kmax, kmax + 1):
        for b in range(-kmax, kmax + 1):
            if a or b:
                amp[a, b] = band_amplitude(f[a], f[b], df, r0, L0, power_law,
                                           inner_scale)
    x = (np.arange(n) - 0.5 * n) * dx
    P = int(subharmonics)
    sh_amp = np.zeros((P, 3, 3))
    sh_freq = np.zeros((P, 3))
    sh_basis = np.zeros((P, 3, n), dtype=np.complex128)
    for p in range(P):
        dfp = df / 3.0 ** (p + 1)
        fp = np.array([-dfp, 0.0, dfp])
        sh_freq[p] = fp
        for a in range(3):
            for b in range(3):
                if a != 1 or b != 1:
                    sh_amp[p, a, b] = band_amplitude(fp[a], fp[b], dfp, r0, L0,
                                                     power_law, inner_scale)
        sh_basis[p] = np.exp(2j * np.pi * fp[:, None] * x[None, :])
    return dict(amplitude=amp, f=f, sh_amp=sh_amp, sh_freq=sh_freq,
                sh_basis=sh_basis, n=n, dx=dx)


# ---------------------------------------------------------------------------
# generate  (fourier.py:233-282)
# ---------------------------------------------------------------------------
def draw_generate_noise(rng, n, count, levels, dtype="float64"):
    """The reference's draw order for one ``generate(count)`` call.

    fourier.py:256-257 (main noise (2, ceil(count/2), n, n)) then
    fourier.py:271-272 (one (2, count, 3, 3) block per subharmonic level).
    """
    n_fft = (count + 1) // 2
    main = rng.standard_normal((2, n_fft, n, n), dtype=dtype)
    sh = np.stack([rng.standard_normal((2, count, 3, 3), dtype=dtype)
                   for _ in range(levels)]) if levels else np.zeros((0, 2, count, 3, 3))
    return main, sh


_FFT_WORKERS: Optional[int] = None


def set_fft_workers(workers: Optional[int]) -> None:
    """Threads of the oracle's 2-D transforms: None is NumPy's single-thread
    pocketfft (the reference's default, ``backend.py:48-50``), an integer
    runs ``scipy.fft.ifft2(workers=...)``, as the reference does when its
    global worker count is set (``backend.py:92-113``).  Used by the CPU
    baseline's all-cores row."""
    global _FFT_WORKERS
    _FFT_WORKERS = None if workers is None else int(workers)


def ifft2_unnormalised(x):
    """``ifft2 * n0*n1``: the +i, unnormalised 2-D DFT over the last two axes."""
    x = np.asarray(x, dtype=np.complex128)
    scale = x.shape[-1] * x.shape[-2]
    if _FFT_WORKERS is not None:
        import scipy.fft

        return scipy.fft.ifft2(x, axes=(-2, -1), workers=_FFT_WORKERS) * scale
    return np.fft.ifft2(x, axes=(-2, -1)) * scale


def subharmonic_field(coeffs, basis):
    """Real low-frequency field of stacked 3x3 coefficients, mean removed.

    ``coeffs`` (..., P, 3, 3) complex, ``basis`` (P, 3, n) complex ->
    (..., n, n) real: ``Re(sum_p basis_p^T c_p basis_p)`` minus its mean over
    the grid (fourier.py:274-278; atmosphere.py:931-934).
    """
    coeffs = np.asarray(coeffs)
    P, n = basis.shape[0], basis.shape[-1]
    lead = coeffs.shape[:-3]
    m = np.matmul(coeffs, basis)  # (..., P, 3, n): two batched products, as
    m = m.reshape(lead + (3 * P, n))  # fourier.py:274-277 forms them
    low = np.matmul(basis.reshape(3 * P, n).T, m).real
    return low - low.mean(axis=(-2, -1), keepdims=True)


def generate(filters, noise, sh_noise, count=None):
    """fourier.py:233-282 -- screens from injected noise, float64.

    ``noise`` (2, n_fft, n, n); ``sh_noise`` (P, 2, count, 3, 3).
    Returns (count, n, n) radians.
    """
    noise = np.asarray(noise, dtype=np.float64)
    if count is None:
        count = 2 * noise.shape[1]
    field = ifft2_unnormalised((noise[0] + 1j * noise[1]) * filters["amplitude"])
    screens = np.concatenate((field.real, field.imag))[:count]
    P = filters["sh_amp"].shape[0]
    if P:
        sh = np.asarray(sh_noise, dtype=np.float64)
        cn = (sh[:, 0] + 1j * sh[:, 1]) * filters["sh_amp"][:, None]  # (P,count,3,3)
        screens = screens + subharmonic_field(np.moveaxis(cn, 0, 1),
                                              filters["sh_basis"])
    return screens


# ---------------------------------------------------------------------------
# spectral frozen flow  (flow.py:71-129, atmosphere.py:882-998)
# ---------------------------------------------------------------------------
def draw_flow(rng, filters, dtype="float64"):
    """flow.py:71-94 -- fixed spectrum and per-level 3x3 coefficients."""
    n = filters["n"]
    w = rng.standard_normal((2, n, n), dtype=dtype).astype(np.float64)
    spectrum = (w[0] + 1j * w[1]) * filters["amplitude"]
    coeffs = []
    for p in range(filters["sh_amp"].shape[0]):
        v = rng.standard_normal((2, 3, 3), dtype=dtype).astype(np.float64)
        coeffs.append((v[0] + 1j * v[1]) * filters["sh_amp"][p])
    P = len(coeffs)
    return spectrum, (np.stack(coeffs) if P else np.zeros((0, 3, 3), complex))


def _shift_phasor(s, freqs):
    return np.exp(2j * np.pi * np.multiply.outer(np.asarray(s, float), freqs))


def integrate(spectra, sh_coeffs, f, sh_freq, sh_basis, disp_x, disp_y,
              crop=None):
    """atmosphere.py:882-938 -- one on-axis frame, float64.

    ``spectra`` (L,n,n); ``sh_coeffs`` (P,L,3,3); ``disp_x/y`` (L,) metres.
    """
    px = _shift_phasor(disp_x, f)  # (L, n)
    py = _shift_phasor(disp_y, f)
    spec = np.einsum("lij,li,lj->ij", spectra, px, py)
    total = ifft2_unnormalised(spec).real
    P = sh_coeffs.shape[0]
    if P:
        qx = np.exp(2j * np.pi * np.asarray(disp_x)[None, :, None] * sh_freq[:, None, :])
        qy = np.exp(2j * np.pi * np.asarray(disp_y)[None, :, None] * sh_freq[:, None, :])
        shifted = np.einsum("plab,pla,plb->pab", sh_coeffs, qx, qy)
        total = total + subharmonic_field(shifted, sh_basis)
    if crop is not None:
        total = total[crop, crop]
    return total


def lgs_zoom_taps(n, n_screen, mag, interp="cubic"):
    """atmosphere.py:581-616 -- per-layer cone-zoom taps: screen index
    (clipped to the screen) and weight, (L, T, n) each, for the pupil grid
    scaled about the screen centre by each layer's magnification."""
    centre = (n_screen - 1) / 2.0
    base = np.arange(n, dtype=np.float64) - (n - 1) / 2.0
    idx, wts = [], []
    for m in np.atleast_1d(mag):
        pos = centre + base * float(m)
        p0 = np.floor(pos).astype(np.int64)
        offs, w = _taps(interp, pos - p0)
        idx.append(np.stack([np.clip(p0 + o, 0, n_screen - 1) for o in offs]))
        wts.append(np.stack([np.asarray(v, float) for v in w]))
    return np.stack(idx), np.stack(wts)


def integrate_lgs(spectra, sh_coeffs, f, sh_freq, sh_basis, disp_x, disp_y,
                  zoom_idx, zoom_w):
    """atmosphere.py:1000-1090 -- LGS-cone frame, float64.

    Each layer's shifted spectrum is transformed on its own, its own
    subharmonic field (mean removed per layer) added, and the layer screen
    sampled by the separable cone-zoom taps; the layers are then summed.
    """
    px = _shift_phasor(disp_x, f)  # (L, n)
    py = _shift_phasor(disp_y, f)
    screens = ifft2_unnormalised(spectra * px[:, :, None] * py[:, None, :]).real
    P = sh_coeffs.shape[0]
    if P:
        qx = np.exp(2j * np.pi * np.asarray(disp_x)[None, :, None] * sh_freq[:, None, :])
        qy = np.exp(2j * np.pi * np.asarray(disp_y)[None, :, None] * sh_freq[:, None, :])
        shifted = sh_coeffs * qx[:, :, :, None] * qy[:, :, None, :]  # (P, L, 3, 3)
        screens = screens + subharmonic_field(np.moveaxis(shifted, 0, 1), sh_basis)
    total = 0.0
    for l in range(screens.shape[0]):
        rows = np.einsum("ti,tic->ic", zoom_w[l], screens[l][zoom_idx[l]])  # (n, ns)
        total = total + np.einsum("tj,itj->ij", zoom_w[l], rows[:, zoom_idx[l]])
    return total


def integrate_dirs(spectra, sh_coeffs, f, sh_freq, sh_basis, disp_x, disp_y,
                   crop=None):
    """atmosphere.py:940-998 -- (D,L) displacements -> (D, n, n)."""
    return np.stack([
        integrate(spectra, sh_coeffs, f, sh_freq, sh_basis, dxs, dys, crop)
        for dxs, dys in zip(np.asarray(disp_x), np.asarray(disp_y))
    ])


def flow_translate(spectrum, coeffs, filters, sx, sy):
    """flow.py:96-129 -- single-layer screen blown by (sx, sy) metres."""
    return integrate(spectrum[None], coeffs[:, None], filters["f"],
                     filters["sh_freq"], filters["sh_basis"], [sx], [sy])


class OracleAtmosphere:
    """atmosphere.py:305-579 + 1191-1219 -- spectral-engine state in FP64.

    Reproduces the reference's SeedSequence tree (atmosphere.py:393-437,
    448-470) so that, for a given ``seed`` and ``dtype``, the layer spectra,
    subharmonic coefficients and sample generators are the reference's own.
    ``layers`` is a sequence of (altitude, cn2_fraction, wind_speed,
    wind_direction_deg, L0) tuples.
    """

    def __init__(self, layers, r0, wavelength=500e-9, zenith_angle=0.0,
                 diameter=8.0, n=512, L0=None, power_law=11.0 / 3.0,
                 inner_scale=0.0, subharmonics=8, field_of_view=0.0,
                 dtype="float32", seed=None, tau_boil=None, lgs_altitude=None,
                 interp="cubic"):
        lay = [tuple(float(v) for v in ly) for ly in layers]
        frac = np.array([ly[1] for ly in lay])
        frac = frac / frac.sum()
        if L0 is not None:
            lay = [(a, f, s, d, float(L0)) for (a, f, s, d, _) in lay]
        self.layers = lay
        self.n = int(n)
        self.dx = diameter / n
        self.wavelength = wavelength
        cz = math.cos(math.radians(zenith_angle))
        self.airmass = 1.0 / cz
        self.r0_los = r0 * cz ** 0.6
        alt = np.array([ly[0] for ly in lay]) * self.airmass
        margin = int(np.ceil(alt.max() * np.tan(field_of_view * ARCSEC) / self.dx))
        self.margin = margin
        self.ns = self.n + 2 * margin
        self.crop = slice(margin, margin + self.n)
        self.alt = alt
        theta = np.deg2rad([ly[3] for ly in lay])
        speed = np.array([ly[2] for ly in lay])
        self.vx = speed * np.cos(theta)
        self.vy = speed * np.sin(theta)
        self.speed = speed
        self.r0_layer = self.r0_los * frac ** (-0.6)
        self.period = self.ns * self.dx
        master = np.random.SeedSequence(seed)
        children = master.spawn(len(lay))
        # spectral boiling stream (atmosphere.py:395-396), then the
        # extruder's boiling stream (atmosphere.py:397)
        self.boil_seed = int(master.spawn(1)[0].generate_state(1)[0])
        master.spawn(1)
        self.dtype = dtype
        if tau_boil is None:
            self.tau_boil = np.full(len(lay), np.inf)
        else:
            self.tau_boil = np.broadcast_to(np.asarray(tau_boil, float), (len(lay),)).copy()
        self.filters, spectra, coeffs = [], [], []
        self.gen_seeds, self.ext_seeds = [], []
        kw = dict(subharmonics=subharmonics, power_law=power_law,
                  inner_scale=inner_scale)
        for ly, r0_i, child in zip(lay, self.r0_layer, children):
            gseed, fseed, eseed = child.spawn(3)
            filt = build_filters(self.ns, self.dx, r0_i, ly[4], **kw)
            self.filters.append(filt)
            self.gen_seeds.append(int(gseed.generate_state(1)[0]))
            self.ext_seeds.append(int(eseed.generate_state(1)[0]))
            rng = np.random.default_rng(int(fseed.generate_state(1)[0]))
            spec, cf = draw_flow(rng, filt, dtype)
            spectra.append(spec)
            coeffs.append(cf)
        self.spectra = np.stack(spectra)
        self.sh_coeffs = np.stack(coeffs, axis=1)  # (P, L, 3, 3)
        self._initial = (self.spectra.copy(), self.sh_coeffs.copy())
        self.f = self.filters[0]["f"]
        self.sh_freq = self.filters[0]["sh_freq"]
        self.sh_basis = self.filters[0]["sh_basis"]
        self.L0 = np.array([ly[4] for ly in lay])
        self.t = 0.0
        self.lgs_mag = None
        if lgs_altitude is not None:  # atmosphere.py:484-498, 578-579
            self.lgs_mag = 1.0 - self.alt / (lgs_altitude * self.airmass)
            if np.any(self.lgs_mag <= 0.0):
                raise ValueError("lgs_altitude must exceed every layer altitude")
            self.zoom_idx, self.zoom_w = lgs_zoom_taps(self.n, self.ns, self.lgs_mag, interp)
        self.boil_rng = np.random.default_rng(self.boil_seed)
        # sample() generators, one per distinct round(L0, 9)
        groups: Dict[float, List[int]] = {}
        for i, ly in enumerate(lay):
            groups.setdefault(round(ly[4], 9), []).append(i)
        self.sample_groups = []
        for idx, child in zip(groups.values(), master.spawn(len(groups))):
            if len(idx) == 1:
                i = idx[0]
                self.sample_groups.append((self.filters[i], self.gen_seeds[i]))
                continue
            r0_agg = float(np.sum(self.r0_layer[idx] ** (-5.0 / 3.0)) ** -0.6)
            filt = build_filters(self.ns, self.dx, r0_agg, lay[idx[0]][4], **kw)
            self.sample_groups.append((filt, int(child.generate_state(1)[0])))

    def displacements(self, t, ox=0.0, oy=0.0):
        return self.vx * t + self.alt * ox, self.vy * t + self.alt * oy

    def phase(self, t, ox=0.0, oy=0.0):
        sx, sy = self.displacements(t, ox, oy)
        if self.lgs_mag is not None:  # atmosphere.py:894-895
            return integrate_lgs(self.spectra, self.sh_coeffs, self.f, self.sh_freq,
                                 self.sh_basis, sx, sy, self.zoom_idx, self.zoom_w)
        return integrate(self.spectra, self.sh_coeffs, self.f, self.sh_freq,
                         self.sh_basis, sx, sy, self.crop)

    def opd(self, t=0.0, directions=None):
        scale = self.wavelength / (2.0 * np.pi)
        if directions is None:
            return self.phase(t) * scale
        return np.stack([self.phase(t, np.tan(a * ARCSEC), np.tan(b * ARCSEC))
                         for a, b in directions]) * scale

    # ------------------------------------------------------------- boiling
    def boil_tau_maps(self):
        """atmosphere.py:618-658 -- per-mode boiling time constants.

        tau(f) = tau_boil * max(f / f_ref, 1)^(-2/3) with f_ref = 1/L0, or
        the grid fundamental 1/(n_screen dx) for L0 = inf; (L, ns, ns) for
        the grid and (P, L, 3, 3) for the subharmonic cells."""
        df = 1.0 / (self.ns * self.dx)
        f_ref = np.array([1.0 / L0 if np.isfinite(L0) and L0 > 0 else df for L0 in self.L0])
        fr = np.hypot(self.f[:, None], self.f[None, :])
        ratio = np.maximum(fr[None] / f_ref[:, None, None], 1.0)
        main = self.tau_boil[:, None, None] * ratio ** (-2.0 / 3.0)
        fp = self.sh_freq  # (P, 3)
        fr_sh = np.hypot(fp[:, :, None], fp[:, None, :])  # (P, 3, 3)
        ratio_sh = np.maximum(fr_sh[:, None] / f_ref[None, :, None, None], 1.0)
        sh = self.tau_boil[None, :, None, None] * ratio_sh ** (-2.0 / 3.0)
        return main, sh

    def draw_boil_noise(self):
        """The draws of one boiling step, in the reference's order and dtype
        (atmosphere.py:1114-1116, 1122-1124): main (2, L, ns, ns), then the
        subharmonic (2, P, L, 3, 3)."""
        L, P = self.spectra.shape[0], self.sh_coeffs.shape[0]
        main = self.boil_rng.standard_normal((2, L, self.ns, self.ns), dtype=self.dtype)
        sh = (self.boil_rng.standard_normal((2, P, L, 3, 3), dtype=self.dtype) if P else None)
        return main, sh

    def boil_step(self, dt, noise=None):
        """atmosphere.py:1092-1126 -- one AR(1) update per mode.

        S <- a S + sqrt(1 - a^2) (n0 + i n1) A with a = exp(-dt / tau(f)); the
        same for the subharmonic coefficients.  Nothing is drawn when every
        layer is frozen (infinite tau_boil).  ``noise``: an injected
        ``(main, sh)`` pair instead of the boiling stream's draw."""
        with np.errstate(divide="ignore"):
            alpha = np.where(np.isfinite(self.tau_boil), np.exp(-dt / self.tau_boil), 1.0)
        if np.all(alpha >= 1.0):
            return
        main_tau, sh_tau = self.boil_tau_maps()
        main, sh = self.draw_boil_noise() if noise is None else noise
        main = np.asarray(main, np.float64)
        a = np.exp(-dt / main_tau)
        b = np.sqrt(np.clip(1.0 - a * a, 0.0, None))
        amp = np.stack([fl["amplitude"] for fl in self.filters])
        self.spectra = a * self.spectra + b * ((main[0] + 1j * main[1]) * amp)
        if self.sh_coeffs.shape[0]:
            sh = np.asarray(sh, np.float64)
            a_sh = np.exp(-dt / sh_tau)
            b_sh = np.sqrt(np.clip(1.0 - a_sh * a_sh, 0.0, None))
            amps = np.stack([fl["sh_amp"] for fl in self.filters], axis=1)  # (P, L, 3, 3)
            self.sh_coeffs = a_sh * self.sh_coeffs + b_sh * ((sh[0] + 1j * sh[1]) * amps)

    def frames(self, dt, steps, noises=None):
        """atmosphere.py:1128-1154 -- yield (t, opd), then advance the clock
        and boil once; ``noises``: optional per-step injected draws."""
        for k in range(int(steps)):
            t = self.t
            yield t, self.opd(t)
            self.t += float(dt)
            self.boil_step(float(dt), None if noises is None else noises[k])

    def evolve(self, dt, noise=None):
        """atmosphere.py:1156-1189 -- advance, boil once, return the OPD."""
        self.t += float(dt)
        self.boil_step(float(dt), noise)
        return self.opd(self.t)

    def reset(self):
        """atmosphere.py:1221-1243 -- clock to 0, initial turbulence and
        boiling stream restored."""
        self.t = 0.0
        self.spectra, self.sh_coeffs = (a.copy() for a in self._initial)
        self.boil_rng = np.random.default_rng(self.boil_seed)

    def sample(self, count, noises=None, dtype="float64"):
        """atmosphere.py:1191-1219 -- summed group screens, OPD metres.

        ``noises``: optional list of ``(main, sh)`` per group; otherwise drawn
        from each group's own seed (fresh generator, first call).
        """
        total = 0.0
        for g, (filt, gseed) in enumerate(self.sample_groups):
            if noises is None:
                rng = np.random.default_rng(gseed)
                main, sh = draw_generate_noise(rng, self.ns, count,
                                               filt["sh_amp"].shape[0], dtype)
            else:
                main, sh = noises[g]
            total = total + generate(filt, main, sh, count)
        return total[:, self.crop, self.crop] * (self.wavelength / (2 * np.pi))


# ---------------------------------------------------------------------------
# covariance and extrusion matrices  (infinite.py:54-117, extrude.py:250-304)
# ---------------------------------------------------------------------------
def phase_covariance(r, r0, L0):
    """infinite.py:88-117 -- von Karman phase covariance [rad^2]."""
    r = np.asarray(r, dtype=np.float64)
    c = ((L0 / r0) ** (5.0 / 3.0) * 2.0 ** (-5.0 / 6.0) * _gamma(11.0 / 6.0)
         / np.pi ** (8.0 / 3.0) * (4.8 * _gamma(1.2)) ** (5.0 / 6.0))
    x = 2.0 * np.pi * r / L0
    core = np.empty_like(x)
    zero = x <= 0
    core[zero] = 2.0 ** (-1.0 / 6.0) * _gamma(5.0 / 6.0)
    xs = x[~zero]
    core[~zero] = xs ** (5.0 / 6.0) * _kv(5.0 / 6.0, xs)
    return c * core


def _spd_solve(spd, rhs):
    """infinite.py:54-64 -- Cholesky solve with least-squares fallback."""
    try:
        return _sla.cho_solve(_sla.cho_factor(spd), rhs)
    except np.linalg.LinAlgError:
        return np.linalg.lstsq(spd, rhs, rcond=None)[0]


def far_stencil(width, stencil_rows, far_points):
    """The optional far points of the extrusion stencil (no reference
    counterpart; the reference's stencil is full rows, extrude.py:272-275).

    ``far_points`` = 8 R: R far rows, row j (1..R) lying
    ``(m - 1) + round((W // 2) ** (j / R))`` rows behind the newest stencil row
    (raised by one where needed so the distances strictly increase), each with
    8 pixels at columns ``(2 i + 1) W // 16``.  Returns (distances, columns).
    """
    m, W, R = int(stencil_rows), int(width), int(far_points) // 8
    horizon = max(2, W // 2)
    dist = []
    for j in range(1, R + 1):
        d = (m - 1) + int(round(horizon ** (j / R)))
        if dist and d <= dist[-1]:
            d = dist[-1] + 1
        dist.append(max(d, m))
    cols = [((2 * i + 1) * W) // 16 for i in range(8)]
    return dist, cols


def build_extrusion(width, stencil_rows, dx, L0, r0=1.0, far_points=0):
    """extrude.py:250-304 / infinite.py:231-269 -- (A, B) in float64.

    Stencil rows at y = 0..m-1, new row at y = m, columns x = 0..W-1;
    ``A = C_xz C_zz^-1`` and ``B = V sqrt(max(lambda, 0))`` of the
    symmetrised residual covariance.  With ``far_points`` the stencil also
    holds the far points of ``far_stencil`` (rows y = m - 1 - d), after the
    full rows, far row by far row.
    """
    m, W = int(stencil_rows), int(width)
    yy, xx = np.meshgrid(np.arange(m, dtype=float), np.arange(W, dtype=float),
                         indexing="ij")
    zx, zy = xx.ravel(), yy.ravel()
    if far_points:
        dist, cols = far_stencil(W, m, far_points)
        fy, fx = np.meshgrid([m - 1.0 - d for d in dist], np.asarray(cols, float),
                             indexing="ij")
        zx, zy = np.concatenate([zx, fx.ravel()]), np.concatenate([zy, fy.ravel()])
    nx, ny = np.arange(W, dtype=float), np.full(W, float(m))

    def cov(ax, ay, bx, by):
        d = np.hypot(ax[:, None] - bx[None, :], ay[:, None] - by[None, :])
        return phase_covariance(d * dx, r0, L0)

    czz = cov(zx, zy, zx, zy)
    cxz = cov(nx, ny, zx, zy)
    cxx = cov(nx, ny, nx, ny)
    A = _spd_solve(czz, cxz.T).T
    res = cxx - A @ cxz.T
    res = 0.5 * (res + res.T)
    lam, vec = np.linalg.eigh(res)
    B = vec * np.sqrt(np.clip(lam, 0.0, None))
    return A, B


# ---------------------------------------------------------------------------
# ring buffer, interpolation, readout  (ring.py, extrude.py, infinite.py)
# ---------------------------------------------------------------------------
def compact_row_ring(buf, base, fill, stencil_rows, first_readable, far_back=None):
    """ring.py:9-29 -- drop consumed rows, keep the stencil tail (and, with
    far points, the farthest far row ``far_back`` rows behind the newest)."""
    keep_from = min(int(math.floor(first_readable)) - 1,
                    base + fill - stencil_rows) - base
    if far_back is not None:
        keep_from = min(keep_from, fill - 1 - int(far_back))
    keep_from = max(1, keep_from)
    keep = fill - keep_from
    buf[:keep] = buf[keep_from:fill].copy()
    return base + keep_from, keep


def catmull_rom_weights(t):
    """extrude.py:74-83 -- taps at offsets (-1, 0, 1, 2)."""
    t2 = t * t
    t3 = t2 * t
    return (0.5 * (-t + 2 * t2 - t3), 0.5 * (2 - 5 * t2 + 3 * t3),
            0.5 * (t + 4 * t2 - 3 * t3), 0.5 * (t3 - t2))


def lanczos_weights(t, a=3):
    """infinite.py:67-85 -- normalised Lanczos-a taps at offsets -a+1..a."""
    offs = tuple(range(-a + 1, a + 1))
    raw = [np.sinc(t - d) * np.sinc((t - d) / a) for d in offs]
    tot = sum(raw[1:], raw[0])
    return offs, tuple(r / tot for r in raw)


def _taps(interp, frac):
    if interp == "linear":
        return (0, 1), (1.0 - frac, frac)
    if interp == "lanczos":
        return lanczos_weights(frac)
    return (-1, 0, 1, 2), catmull_rom_weights(frac)


def readout(buf, fill, along, perp, shift_along, shift_perp, interp="cubic"):
    """extrude.py:883-915 -- multi-layer 2-D interpolated gather, summed.

    ``buf`` (L, cap, W); ``fill`` (L,) valid rows; ``along``/``perp``
    (L, n, n) grids; shifts (L,).  Float64 accumulation.
    """
    L, _, W = buf.shape
    total = 0.0
    for l in range(L):
        row = along[l] + shift_along[l]
        col = perp[l] + shift_perp[l]
        r0 = np.floor(row).astype(np.int64)
        c0 = np.floor(col).astype(np.int64)
        ro, rw = _taps(interp, row - r0)
        co, cw = _taps(interp, col - c0)
        b = buf[l].astype(np.float64)
        acc = 0.0
        for dr, wr in zip(ro, rw):
            rr = np.clip(r0 + dr, 0, int(fill[l]) - 1)
            racc = 0.0
            for dc, wc in zip(co, cw):
                cc = np.clip(c0 + dc, 0, W - 1)
                racc = racc + wc * b[rr, cc]
            acc = acc + wr * racc
        total = total + acc
    return total


def pupil_grids(n, cos_w, sin_w, W, magnification=1.0):
    """extrude.py:376-383 -- rotated pupil grid in wind-frame pixels."""
    g = (np.arange(n, dtype=np.float64) - 0.5 * (n - 1)) * magnification
    along = g[:, None] * cos_w + g[None, :] * sin_w
    perp = -g[:, None] * sin_w + g[None, :] * cos_w + 0.5 * W
    return along, perp


def along_extent_and_span(n, vx, vy, mag, fov):
    """extrude.py:222-247."""
    sp = math.hypot(vx, vy)
    c = vx / sp if sp > 0 else 1.0
    s = vy / sp if sp > 0 else 0.0
    ext = 0.5 * (n - 1) * mag * (abs(c) + abs(s))
    span = int(math.ceil(ext + fov)) - int(math.floor(-ext - fov)) + 3
    return ext, span


def extrusion_seed_factor(width, stencil_rows, dx, L0, r0=1.0):
    """extrude.py:284-303 -- S with S S^T = C_zz (Cholesky, else the
    symmetric square root), the seed of a fresh stencil block."""
    m, W = int(stencil_rows), int(width)
    zy, zx = np.divmod(np.arange(m * W, dtype=np.float64), W)
    sep = np.hypot(zx[:, None] - zx[None, :], zy[:, None] - zy[None, :])
    c_zz = phase_covariance(sep * dx, r0, L0)
    c_zz = 0.5 * (c_zz + c_zz.T)
    try:
        return np.linalg.cholesky(c_zz)
    except np.linalg.LinAlgError:
        w, v = np.linalg.eigh(c_zz)
        return v * np.sqrt(np.clip(w, 0.0, None))


class OracleExtruder:
    """extrude.py:307-469, 520-659 -- frozen-flow multi-layer extruder, FP64.

    Innovations are injected: ``innovation(layer, k)`` must return the k-th
    (W,) white-noise row for that layer (the reference draws it from the
    layer's generator after the seed screen, extrude.py:428).  ``seed_rows``
    (L, span_l, W) gives each layer's initial rows (extrude.py:408-420).
    """

    def __init__(self, n, dx, layer_r0, layer_L0, layer_wind, layer_alt,
                 fov_pix=0.0, stencil_rows=2, lgs_alt=None, A=None, B_unit=None,
                 tau_boil=None, boil_seed=None, far_points=0):
        self.n, self.dx, self.m = int(n), float(dx), int(stencil_rows)
        self.far_points = int(far_points)
        L = len(layer_r0)
        self.tau_boil = [np.inf] * L if tau_boil is None else [float(t) for t in tau_boil]
        fov = [float(fov_pix)] * L if np.ndim(fov_pix) == 0 else list(fov_pix)
        mags = [1.0 - a / lgs_alt for a in layer_alt] if lgs_alt else [1.0] * L
        ext, spans = zip(*[along_extent_and_span(self.n, vx, vy, mg, fv)
                           for (vx, vy), mg, fv in zip(layer_wind, mags, fov)])
        self.W = int(math.ceil(2.0 * max(ext) + 2.0 * max(fov))) + 4
        self.far_dist, self.far_cols = (far_stencil(self.W, self.m, far_points) if far_points
                                        else ([], []))
        reach = self.far_dist[-1] if self.far_dist else 0
        self.cap = max(spans) + self.n + self.m + 32 + reach
        # far rows need history behind the first extruded row: extra seed rows
        self.extra = [max(0, reach + 1 - s) if reach else 0 for s in spans]
        self.spans = [s + e for s, e in zip(spans, self.extra)]
        self.L = L
        if A is None:
            A, B_unit = build_extrusion(self.W, self.m, self.dx, layer_L0[0],
                                        far_points=far_points)
        self.A = np.asarray(A, np.float64)
        self.B = [np.asarray(B_unit, np.float64) * r ** (-5.0 / 6.0) for r in layer_r0]
        self.buf = np.zeros((L, self.cap, self.W))
        self.base = np.zeros(L, np.int64)
        self.fill = np.zeros(L, np.int64)
        self.travel = np.zeros(L)
        self.speed, self.cos, self.sin = [], [], []
        self.along, self.perp = [], []
        self.along_min, self.along_max, self.fov = [], [], fov
        self.alt_pix = [a / self.dx for a in layer_alt]
        for l, ((vx, vy), mg) in enumerate(zip(layer_wind, mags)):
            sp = math.hypot(vx, vy)
            c = vx / sp if sp > 0 else 1.0
            s = vy / sp if sp > 0 else 0.0
            self.speed.append(sp)
            self.cos.append(c)
            self.sin.append(s)
            a, p = pupil_grids(self.n, c, s, self.W, mg)
            self.along.append(a)
            self.perp.append(p)
            self.along_min.append(float(a.min()))
            self.along_max.append(float(a.max()))
            self.base[l] = int(math.floor(a.min() - fov[l])) - 2 - self.extra[l]
        self.innov_count = np.zeros(L, np.int64)
        # boiling (extrude.py:586-654): S = S_unit r0^(-5/6), one stream
        self.S = None
        if any(np.isfinite(t) for t in self.tau_boil):
            s_unit = extrusion_seed_factor(self.W, self.m, self.dx, layer_L0[0])
            self.S = [s_unit * r ** (-5.0 / 6.0) for r in layer_r0]
            self.boil_rng = np.random.default_rng(boil_seed)

    def seed(self, seed_rows):
        for l in range(self.L):
            k = self.spans[l]
            self.buf[l, :k] = np.asarray(seed_rows[l], np.float64)[:k]
            self.fill[l] = k

    def _extrude_one(self, l, innovation):
        if self.fill[l] == self.cap:
            self.base[l], self.fill[l] = compact_row_ring(
                self.buf[l], int(self.base[l]), int(self.fill[l]), self.m,
                self.along_min[l] - self.fov[l] + self.travel[l],
                self.far_dist[-1] if self.far_dist else None)
        f = int(self.fill[l])
        z = self.buf[l, f - self.m:f].ravel()
        if self.far_dist:
            rows = [f - 1 - d for d in self.far_dist]
            assert min(rows) >= 0, "far row behind the oldest ring row"
            z = np.concatenate([z, self.buf[l][np.ix_(rows, self.far_cols)].ravel()])
        beta = np.asarray(innovation(l, int(self.innov_count[l])), np.float64)
        self.innov_count[l] += 1
        self.buf[l, f] = self.A @ z + self.B[l] @ beta
        self.fill[l] = f + 1

    def rows_needed(self, l, travel):
        top = int(math.ceil(self.along_max[l] + self.fov[l] + travel)) + 2
        return max(0, top - int(self.base[l] + self.fill[l] - 1))

    def set_time(self, t, innovation):
        for l in range(self.L):
            tr = self.speed[l] * t / self.dx
            if tr < self.travel[l]:
                raise ValueError("wind travel is monotonic; travel cannot decrease")
            self.travel[l] = tr
            top = int(math.ceil(self.along_max[l] + self.fov[l] + tr)) + 2
            while self.base[l] + self.fill[l] - 1 < top:
                self._extrude_one(l, innovation)

    def shifts(self, thx=0.0, thy=0.0):
        sa = np.empty(self.L)
        sp = np.empty(self.L)
        for l in range(self.L):
            dxp, dyp = self.alt_pix[l] * thx, self.alt_pix[l] * thy
            oa = dxp * self.cos[l] + dyp * self.sin[l]
            op = -dxp * self.sin[l] + dyp * self.cos[l]
            sa[l] = (self.travel[l] - self.base[l]) + oa
            sp[l] = op
        return sa, sp

    def integrate(self, thx=0.0, thy=0.0, interp="cubic"):
        sa, sp = self.shifts(thx, thy)
        return readout(self.buf, self.fill, np.stack(self.along),
                       np.stack(self.perp), sa, sp, interp)

    def boil_region(self, l, dt):
        """extrude.py:472-496 -- (lo, hi, a, b) of layer l's readable rows."""
        if not np.isfinite(self.tau_boil[l]):
            return None
        a = math.exp(-dt / self.tau_boil[l])
        b = math.sqrt(max(0.0, 1.0 - a * a))
        lo = max(0, int(math.floor(self.along_min[l] + self.travel[l])) - int(self.base[l]) - 3)
        hi = int(self.fill[l])
        return None if hi <= lo else (lo, hi, a, b)

    def fresh_window(self, l, height, draw=None):
        """extrude.py:695-716 -- an independent (height, W) window: S w
        seeds m rows, A/B extrude the rest.  ``draw(k)`` returns the k-th
        white-noise vector (k = 0: the seed, m W values; then one row of W
        per extruded row); by default the boiling stream, in that order."""
        draw = draw or (lambda k, size: self.boil_rng.standard_normal(size))
        W, m = self.W, self.m
        out = np.empty((height, W))
        seed = (self.S[l] @ np.asarray(draw(0, m * W), float)).reshape(m, W)
        rows = min(m, height)
        out[:rows] = seed[:rows]
        for k in range(m, height):
            out[k] = self.A @ out[k - m:k].ravel() + self.B[l] @ np.asarray(draw(k - m + 1, W), float)
        return out

    def boil_step(self, dt, draws=None):
        """extrude.py:718-752 (NumPy path) -- every boiling layer's readable
        window blended toward a fresh window, in layer order.  ``draws``:
        optional per-layer ``draw`` callables (see fresh_window)."""
        if self.S is None:
            return
        for l in range(self.L):
            region = self.boil_region(l, dt)
            if region is None:
                continue
            lo, hi, a, b = region
            fresh = self.fresh_window(l, hi - lo, None if draws is None else draws[l])
            self.buf[l, lo:hi] = a * self.buf[l, lo:hi] + b * fresh


class OracleInfinite:
    """infinite.py:120-348 -- single-layer extruder along axis 0, FP64."""

    def __init__(self, n, dx, r0, L0, seed_screen, stencil_rows=2,
                 interp="cubic", A=None, B=None):
        self.n, self.m, self.interp = int(n), int(stencil_rows), interp
        if A is None:
            A, B = build_extrusion(self.n, self.m, dx, L0, r0)
        self.A, self.B = np.asarray(A, float), np.asarray(B, float)
        self.cap = self.n + self.m + 6
        self.buf = np.zeros((self.cap, self.n))
        self.buf[:self.n] = seed_screen
        self.base, self.fill, self.travel = 0, self.n, 0.0
        self.count = 0

    def _advance_to(self, travel, innovation):
        if travel < self.travel:
            raise ValueError("wind travel is monotonic; travel cannot decrease")
        self.travel = float(travel)
        top = int(np.floor(self.travel)) + self.n + 2
        while self.base + self.fill - 1 < top:
            if self.fill == self.cap:
                self.base, self.fill = compact_row_ring(
                    self.buf, self.base, self.fill, self.m, self.travel)
            z = self.buf[self.fill - self.m:self.fill].ravel()
            beta = np.asarray(innovation(self.count), float)
            self.count += 1
            self.buf[self.fill] = self.A @ z + self.B @ beta
            self.fill += 1
        return self.sample()

    def sample(self):
        pos = (self.travel - self.base) + np.arange(self.n)
        i0 = np.floor(pos).astype(np.int64)
        t = (pos - i0)[:, None]
        offs, ws = _taps(self.interp, t)
        out = 0.0
        for d, w in zip(offs, ws):
            out = out + w * self.buf[np.clip(i0 + d, 0, self.fill - 1)]
        return out


# ---------------------------------------------------------------------------
# Zernike modes and Noll (1976) variances, for the validation statistics
# ---------------------------------------------------------------------------
# Noll (1976) Table IV: residual variance after removing the first J modes, in
# (D/r0)^(5/3) rad^2 (analysis.py:40-49); beyond the table the asymptote
# 0.2944 J^(-sqrt(3)/2) (analysis.py:51-58).
_NOLL_TABLE = (1.0299, 0.582, 0.134, 0.111, 0.0880, 0.0648, 0.0587, 0.0525, 0.0463, 0.0401,
               0.0377, 0.0352, 0.0328, 0.0304, 0.0279, 0.0267, 0.0255, 0.0243, 0.0232,
               0.0220, 0.0208)


def noll_to_zernike(j: int) -> Tuple[int, int]:
    """analysis.py:61-72 -- Noll index j >= 1 to (radial n, azimuthal m).

    Modes are counted row by row of the radial order n (n + 1 modes each);
    inside a row |m| rises in steps of 2 from n mod 2, even j takes +|m|
    (cosine), odd j takes -|m| (sine)."""
    if j < 1:
        raise ValueError("Noll index j must be >= 1")
    n, first = 0, 1  # first Noll index of radial order n
    while first + n + 1 <= j:
        first += n + 1
        n += 1
    k = j - first  # position inside the row
    mag = (n % 2) + 2 * ((k + (n + 1) % 2) // 2)
    return n, (mag if j % 2 == 0 else -mag)


def _zernike_radial(n: int, m: int, rho):
    """analysis.py:75-86 -- R_n^|m|(rho) by its finite sum."""
    m = abs(m)
    out = np.zeros_like(rho)
    for k in range((n - m) // 2 + 1):
        c = (-1) ** k * math.factorial(n - k) / (
            math.factorial(k) * math.factorial((n + m) // 2 - k) * math.factorial((n - m) // 2 - k))
        out = out + c * rho ** (n - 2 * k)
    return out


def zernike_basis(n_modes: int, n_pixels: int, diameter_pixels: Optional[float] = None):
    """analysis.py:89-136 -- (n_modes, n, n) Noll-normalised modes on the
    inscribed disc (zero outside); pixel centres at (i - (n-1)/2) / radius."""
    if n_modes < 1 or n_pixels < 2:
        raise ValueError("n_modes >= 1 and n_pixels >= 2 required")
    radius = 0.5 * (n_pixels if diameter_pixels is None else diameter_pixels)
    c = (np.arange(n_pixels) - 0.5 * (n_pixels - 1)) / radius
    x, y = c[:, None], c[None, :]
    rho = np.sqrt(x * x + y * y)
    theta = np.arctan2(np.broadcast_to(y, rho.shape), np.broadcast_to(x, rho.shape))
    inside = rho <= 1.0
    out = np.zeros((n_modes, n_pixels, n_pixels))
    for i in range(n_modes):
        n, m = noll_to_zernike(i + 1)
        ang = np.cos(m * theta) if m >= 0 else np.sin(-m * theta)
        scale = math.sqrt(n + 1.0) * (math.sqrt(2.0) if m != 0 else 1.0)
        out[i] = np.where(inside, scale * _zernike_radial(n, m, rho) * ang, 0.0)
    return out


def zernike_decompose(phase, n_modes: int, basis=None):
    """analysis.py:139-162 -- least-squares coefficients over the pupil
    (the support of piston); (count, n_modes) for a stack."""
    phase = np.asarray(phase, dtype=np.float64)
    single = phase.ndim == 2
    stack = phase[None] if single else phase
    if basis is None:
        basis = zernike_basis(n_modes, stack.shape[-1])
    pupil = basis[0] != 0
    design = basis[:, pupil].T
    coeffs = np.linalg.lstsq(design, stack[:, pupil].T, rcond=None)[0].T
    return coeffs[0] if single else coeffs


def noll_residual(j: int) -> float:
    """analysis.py:51-58 -- Delta_J in (D/r0)^(5/3) units."""
    if j < 1:
        raise ValueError("Noll index j must be >= 1")
    if j <= len(_NOLL_TABLE):
        return _NOLL_TABLE[j - 1]
    return 0.2944 * j ** (-math.sqrt(3.0) / 2.0)


def noll_variance(j: int, diameter: float, r0: float) -> float:
    """analysis.py:165-174 -- Kolmogorov variance of mode j >= 2 [rad^2]."""
    if j < 2:
        raise ValueError("piston (j=1) has no finite variance; use j >= 2")
    return (noll_residual(j - 1) - noll_residual(j)) * (diameter / r0) ** (5.0 / 3.0)


# ---------------------------------------------------------------------------
# The device boiling stream: Philox4x32-10 (Salmon, Moraes, Dror & Shaw,
# "Parallel random numbers: as easy as 1, 2, 3", SC 2011) + Box-Muller
# ---------------------------------------------------------------------------
_PHILOX_M = (np.uint64(0xD2511F53), np.uint64(0xCD9E8D57))
_PHILOX_W = (np.uint32(0x9E3779B9), np.uint32(0xBB67AE85))


def philox4x32_10(ctr, key):
    """Philox4x32 with 10 rounds on uint32 arrays; ``ctr`` (..., 4), ``key`` (2,)."""
    c = [np.asarray(ctr, dtype=np.uint32)[..., i].astype(np.uint64) for i in range(4)]
    k0, k1 = np.uint32(key[0]), np.uint32(key[1])
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = _PHILOX_M[0] * c[0]
        p1 = _PHILOX_M[1] * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0 = np.uint32((int(k0) + int(_PHILOX_W[0])) & 0xFFFFFFFF)
        k1 = np.uint32((int(k1) + int(_PHILOX_W[1])) & 0xFFFFFFFF)
    return np.stack([v.astype(np.uint32) for v in c], axis=-1)


def box_muller(a, b):
    """Two standard normals from 32-bit words: u1 = (a + 1) 2^-32 in (0, 1],
    u2 = (b >> 8) 2^-24 in [0, 1) (the angle takes the top 24 bits)."""
    u1 = (np.asarray(a, np.float64) + 1.0) * 2.0 ** -32
    u2 = (np.asarray(b, np.uint32) >> np.uint32(8)).astype(np.float64) * 2.0 ** -24
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)


def device_boil_noise(seed, step, L, n, P):
    """The normals the device stream uses at one boiling step, in the
    reference's layout: main (2, L, n, n) and subharmonic (2, P, L, 3, 3).

    Mode e and its mirror m = -e mod n (e <= m) share one Philox call with
    counter (e, layer, step low, step high): words 0-1 give the normals of
    e, words 2-3 those of m; subharmonic cell q uses (q, 2^31, step)."""
    key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    k0, k1 = np.divmod(np.arange(n * n), n)
    mirror = ((n - k0) % n) * n + (n - k1) % n
    e = np.arange(n * n)
    own = e <= mirror
    main = np.zeros((2, L, n * n))
    for l in range(L):
        ctr = np.zeros((int(own.sum()), 4), dtype=np.uint32)
        ctr[:, 0] = e[own]
        ctr[:, 1] = l
        ctr[:, 2] = step & 0xFFFFFFFF
        ctr[:, 3] = (step >> 32) & 0xFFFFFFFF
        r = philox4x32_10(ctr, key)
        re_e, im_e = box_muller(r[:, 0], r[:, 1])
        re_m, im_m = box_muller(r[:, 2], r[:, 3])
        main[0, l, mirror[own]] = re_m
        main[1, l, mirror[own]] = im_m
        main[0, l, e[own]] = re_e  # self-conjugate modes take words 0-1
        main[1, l, e[own]] = im_e
    sh = None
    if P:
        count = P * L * 9
        ctr = np.zeros((count, 4), dtype=np.uint32)
        ctr[:, 0] = np.arange(count)
        ctr[:, 1] = 0x80000000
        ctr[:, 2] = step & 0xFFFFFFFF
        ctr[:, 3] = (step >> 32) & 0xFFFFFFFF
        r = philox4x32_10(ctr, key)
        sh = np.stack(box_muller(r[:, 0], r[:, 1])).reshape(2, P, L, 3, 3)
    return main.reshape(2, L, n, n), sh
