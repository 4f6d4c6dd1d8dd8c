"""Ensemble statistics of the device screens against analytic turbulence theory.

These run the product path (HIP kernels through the public API) and compare
with closed forms, as the reference's statistical tests and validation do
(tests/test_fourier.py:44-87, tests/test_infinite.py:169-239,
tests/test_analysis.py:83-98, benchmarks/bench_compare.py:178-225):

* Monte-Carlo structure function: fractional RMS error against von Karman
  D(r) = 2 [C(0) - C(r)] over r in [4 dx, D/4] below 2 % (2048 screens, so
  the Monte-Carlo noise is ~0.3 % and the bound tests the systematic bias;
  the FP64 oracle of the same algorithm gives 1.7 % on 1024 screens, a
  deficit that grows to -3 % at D/4 -- the method's own bias);
* the same for the two-group (L0 = 25 m / 50 m) ensemble of pt_generate_groups;
* frozen-flow temporal PSD slope of the spectral engine in the -8/3 band;
* extruded structure function along the wind and isotropy after burn-in;
* Kolmogorov Zernike-mode variances against Noll (1976), with the bands of
  tests/test_analysis.py:42-61.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _vk(r, r0, L0):
    return 2.0 * (O.phase_covariance(0.0, r0, L0) - O.phase_covariance(r, r0, L0))


def _sf_device(x, smax):
    """Axis-averaged structure function of a (B, n, n) device stack (float64)."""
    import torch

    x = x.to(torch.float64)
    out = []
    for s in range(1, smax + 1):
        a = x[:, s:, :] - x[:, :-s, :]
        b = x[:, :, s:] - x[:, :, :-s]
        out.append(0.5 * (float((a * a).mean()) + float((b * b).mean())))
    return np.array(out)


def _frac_rms(d, target, r, dx, D):
    mid = (r >= 4 * dx) & (r <= D / 4)
    rel = (d[mid] - target[mid]) / target[mid]
    return float(np.sqrt(np.mean(rel ** 2)))


def test_monte_carlo_structure_function_within_two_percent(hip):
    import pyturb_amd as pa

    print("structure function 2048 x 256^2", flush=True)
    n, D, r0, L0 = 256, 8.0, 0.15, 25.0
    dx = D / n
    gen = pa.PhaseScreen(n, dx, r0, L0, seed=1, device="mi355x")
    acc, m = 0.0, 0
    for _ in range(4):
        acc = acc + _sf_device(gen.generate(512), n // 4)
        m += 1
    d = acc / m
    r = np.arange(1, n // 4 + 1) * dx
    err = _frac_rms(d, _vk(r, r0, L0), r, dx, D)
    print("fractional rms error", err)
    assert err < 0.02


def test_two_group_ensemble_structure_function_within_two_percent(hip):
    """Two outer-scale groups, each half the Cn2, summed in one transform."""
    import torch
    from pyturb_amd import ops
    from pyturb_amd.spectrum import build_filters

    print("two-group structure function", flush=True)
    n, D, r0 = 256, 8.0, 0.15
    dx = D / n
    r0g = r0 * 0.5 ** -0.6
    dev = torch.device("cuda")
    filt = [build_filters(n, dx, r0g, L0, 8) for L0 in (25.0, 50.0)]
    amp = torch.tensor(np.stack([f[0] for f in filt]), dtype=torch.float32, device=dev)
    sha = torch.tensor(np.stack([f[2] for f in filt]), dtype=torch.float64, device=dev)
    g = torch.Generator(device=dev).manual_seed(4)
    acc = 0.0
    for _ in range(4):
        noise = torch.randn((2, 2, 256, n, n), generator=g, device=dev)
        shn = torch.randn((2, 8, 2, 512, 3, 3), generator=g, device=dev)
        acc = acc + _sf_device(ops.generate_groups(noise, shn, amp, sha, 512), n // 4)
    d = acc / 4
    r = np.arange(1, n // 4 + 1) * dx
    target = _vk(r, r0g, 25.0) + _vk(r, r0g, 50.0)
    err = _frac_rms(d, target, r, dx, D)
    print("fractional rms error", err)
    assert err < 0.02


def test_spectral_frames_temporal_psd_slope(hip):
    """A pupil point under frozen flow shows the ~ -8/3 temporal power law."""
    import warnings

    import pyturb_amd as pa

    layers = [pa.Layer(0.0, 1.0, wind_speed=10.0, wind_direction=0.0, L0=100.0)]
    atm = pa.Atmosphere(layers, r0=0.15, n=64, diameter=4.0, seed=3, subharmonics=6,
                        device="mi355x")
    times = np.arange(2048) * 1e-3
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        frames = atm.opd_batch(times).cpu().numpy()
    series = frames[:, 16:48:8, 16:48:8].reshape(2048, -1).T  # 16 pupil points
    series = series - series.mean(axis=-1, keepdims=True)
    psd = (np.abs(np.fft.rfft(series, axis=-1)) ** 2).mean(axis=0)[1:]
    freq = np.fft.rfftfreq(2048, d=1e-3)[1:]
    band = (freq >= 5) & (freq <= 60)
    slope = np.polyfit(np.log(freq[band]), np.log(psd[band]), 1)[0]
    print("temporal PSD slope", slope)
    assert -3.3 < slope < -2.2


def test_extruded_structure_function_matches_theory(hip):
    """Rows extruded along the wind follow D(r) = 2 [C(0) - C(r)]."""
    import pyturb_amd as pa

    print("extruded rows D(r)", flush=True)
    r0, L0, dx = 0.12, 4.0, 0.04
    layer = pa.InfinitePhaseScreen(n=32, pixel_scale=dx, r0=r0, L0=L0, seed=11, device="mi355x")
    rows = []
    for _ in range(4000):
        rows.append(layer.step(1)[-1])
    import torch

    history = torch.stack(rows).double()
    seps = np.arange(1, 13)
    measured = np.array([float(((history[s:] - history[:-s]) ** 2).mean()) for s in seps])
    ratio = measured / _vk(seps * dx, r0, L0)
    print("ratio", ratio)
    assert np.all(ratio > 0.7) and np.all(ratio < 1.3)
    assert abs(ratio.mean() - 1.0) < 0.15


def test_extruded_screen_is_isotropic_after_burn_in(hip):
    import pyturb_amd as pa

    print("extruded isotropy", flush=True)
    n, dx, r0, L0 = 48, 0.05, 0.15, 25.0
    seps = [8, 12, 16]
    snaps = [120 + 20 * k for k in range(8)]
    along = np.zeros(len(seps))
    cross = np.zeros(len(seps))
    for seed in range(50):
        scr = pa.InfinitePhaseScreen(n=n, pixel_scale=dx, r0=r0, L0=L0, seed=seed,
                                     device="mi355x")
        prev = 0
        for travel in snaps:
            s = scr.advance(travel - prev).double().cpu().numpy()
            prev = travel
            for i, d in enumerate(seps):
                along[i] += np.mean((s[d:, :] - s[:-d, :]) ** 2)
                cross[i] += np.mean((s[:, d:] - s[:, :-d]) ** 2)
    along /= 50 * len(snaps)
    cross /= 50 * len(snaps)
    theory = _vk(np.array(seps) * dx, r0, L0)
    assert np.all((0.85 < along / theory) & (along / theory < 1.15)), along / theory
    assert np.all((0.85 < cross / theory) & (cross / theory < 1.15)), cross / theory
    assert 0.85 < along.mean() / cross.mean() < 1.15


def test_zernike_variances_match_noll(hip):
    """Device screens at L0 = inf decomposed on Noll modes 1..15 over the
    inscribed pupil: modes 4..15 together and the astigmatism pair against
    Noll's variances (tip/tilt are under-sampled by a finite screen and are
    not banded, as in the reference's own check)."""
    import pyturb_amd as pa

    print("Noll variances 1600 x 128^2", flush=True)
    D, n, r0 = 4.0, 128, 0.4
    gen = pa.PhaseScreen(n, D / n, r0, np.inf, seed=7, device="mi355x")
    basis = O.zernike_basis(15, n)
    coeffs = np.concatenate([O.zernike_decompose(gen.generate(400).double().cpu().numpy(), 15,
                                                 basis=basis) for _ in range(4)])
    measured = coeffs.var(axis=0)
    mid = measured[3:15].sum() / sum(O.noll_variance(j, D, r0) for j in range(4, 16))
    astig = 0.5 * (measured[4] + measured[5]) / O.noll_variance(5, D, r0)
    print("modes 4-15 ratio", mid, "astigmatism ratio", astig)
    assert 0.85 < mid < 1.2
    assert 0.8 < astig < 1.25
