"""Spectral boiling on the device (pt_boiled_frames) against the oracle.

* the reference's own boiled frames (golden, atmosphere.py:1092-1189 with
  its seed tree and draw order), with the reference's draws injected;
* the device's Philox stream against the oracle fed the same normals
  (oracle.device_boil_noise), 9 layers, 8 subharmonic levels;
* chunking: a frames() generator abandoned mid-chunk leaves the state the
  reference would have, so later frames continue exactly;
* statistics of the AR(1) process: the spectra stay stationary, and the
  frame autocorrelation follows sum A^2 exp(-lag dt / tau(f)) / sum A^2.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from conftest import config2_layer_tuples, load_golden

pytestmark = pytest.mark.gpu

BOIL_LAYERS = [(0.0, 0.5, 8.0, 30.0, 25.0), (3000.0, 0.3, 14.0, 110.0, 25.0),
               (9000.0, 0.2, 21.0, 250.0, 25.0)]
BOIL_TAU = [0.02, np.inf, 0.05]


def _np(x):
    return x.double().cpu().numpy()


def test_boiled_frames_match_reference_golden(hip):
    """frames(), evolve(), opd() and reset() with the reference's draws."""
    import pyturb_amd as pa

    print("boiling golden n=32", flush=True)
    g = load_golden("boil_n32_seed6.npz")
    ora = O.OracleAtmosphere(BOIL_LAYERS, 0.15, diameter=4.0, n=32, subharmonics=3,
                             dtype="float64", seed=6, tau_boil=BOIL_TAU)
    draws = [ora.draw_boil_noise() for _ in range(8)]
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in BOIL_LAYERS], r0=0.15, diameter=4.0, n=32,
                        subharmonics=3, tau_boil=BOIL_TAU, dtype="float64", seed=6,
                        device="mi355x")
    atm._boil_noise = lambda k: draws[k]
    seq = [_np(f) for _, f in atm.frames(0.004, 5)]
    seq.append(_np(atm.evolve(0.004)))
    seq.append(_np(atm.opd(0.1)))
    err = O.rel_rms(np.stack(seq), g["frames"])
    print("boiled frames rel rms", err)
    assert err < 1e-5
    atm.reset()
    again = np.stack([_np(f) for _, f in atm.frames(0.004, 2)])
    assert O.rel_rms(again, g["after_reset"]) < 1e-5


def _config2(pa, n, tau, subharmonics=8, seed=2):
    return pa.Atmosphere([pa.Layer(*ly) for ly in config2_layer_tuples()], r0=0.15, diameter=8.0,
                         n=n, L0=25.0, subharmonics=subharmonics, tau_boil=tau, seed=seed,
                         device="mi355x")


def test_device_stream_matches_oracle(hip):
    """Philox normals drawn on the device == the oracle's restatement."""
    import pyturb_amd as pa

    print("boiling device stream n=64 9 layers", flush=True)
    n, dt = 64, 0.002
    atm = _config2(pa, n, 0.03)
    ora = O.OracleAtmosphere(config2_layer_tuples(), 0.15, diameter=8.0, n=n, L0=25.0,
                             subharmonics=8, dtype="float32", seed=2, tau_boil=0.03)
    assert ora.boil_seed == atm._boil_seed
    noises = [O.device_boil_noise(ora.boil_seed, k, 9, n, 8) for k in range(8)]
    dev = np.stack([_np(f) for _, f in atm.frames(dt, 6)] + [_np(atm.evolve(dt))])
    ref = [f for _, f in ora.frames(dt, 6, noises=noises)]
    ref.append(ora.evolve(dt, noise=noises[6]))
    err = O.rel_rms(dev, np.stack(ref))
    print("device stream rel rms", err)
    assert err < 1e-5


def test_abandoned_frames_generator_keeps_exact_state(hip):
    """As in the reference's generator (atmosphere.py:1146-1154), the clock
    advances and the state boils only when a yield resumes: three next()
    calls and close() leave t = 2 dt after two boiling steps, so the next
    frames() starts with the frame at 2 dt again."""
    import pyturb_amd as pa

    print("boiling chunk restore n=64", flush=True)
    dt = 0.003
    whole = _config2(pa, 64, 0.02)
    ref = np.stack([_np(f) for _, f in whole.frames(dt, 10)])
    atm = _config2(pa, 64, 0.02)
    gen = atm.frames(dt, 10)
    first = np.stack([_np(next(gen)[1]) for _ in range(3)])
    gen.close()  # the chunk computed 10 frames; 3 were taken
    assert np.isclose(atm.time, 2 * dt)
    rest = np.stack([_np(f) for _, f in atm.frames(dt, 7)])
    assert np.isclose(atm.time, 9 * dt)
    last = _np(next(atm.frames(dt, 1))[1])
    assert np.array_equal(first, ref[:3])
    # frames share a transform in pairs, so a frame paired differently in
    # the two runs (8 dt: alone here, with 9 dt there) differs by rounding
    assert O.rel_rms(rest, ref[2:9]) < 1e-6
    assert O.rel_rms(last, ref[9]) < 1e-6


def test_boiled_spectra_stay_stationary(hip):
    """After many AR(1) steps every mode keeps E|S|^2 = 2 A^2."""
    import pyturb_amd as pa

    print("boiling stationarity n=512", flush=True)
    atm = _config2(pa, 512, 0.02)
    for _, f in atm.frames(0.002, 300):
        pass
    s = atm._spectra.abs().double() ** 2
    a = atm._boil_amp.double() ** 2
    live = a > 0
    ratio = float((s[live] / a[live]).mean())
    print("E|S|^2 / A^2", ratio)
    assert abs(ratio - 2.0) < 0.05


def test_frame_autocorrelation_matches_theory(hip):
    """No wind, no subharmonics: corr(lag) = sum A^2 a^lag / sum A^2."""
    import torch

    import pyturb_amd as pa

    print("boiling autocorrelation n=128", flush=True)
    n, dt, tau = 128, 0.001, 0.01
    layers = [(0.0, 0.6, 0.0, 0.0, 25.0), (5000.0, 0.4, 0.0, 0.0, 25.0)]
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in layers], r0=0.15, diameter=4.0, n=n,
                        subharmonics=0, tau_boil=tau, seed=8, device="mi355x")
    frames = torch.stack([f for _, f in atm.frames(dt, 4000)]).double()
    frames = frames - frames.mean(dim=0)
    var = float((frames * frames).mean())
    ora = O.OracleAtmosphere(layers, 0.15, diameter=4.0, n=n, subharmonics=0, seed=8,
                             tau_boil=tau)
    main_tau, _ = ora.boil_tau_maps()
    w = np.stack([fl["amplitude"] for fl in ora.filters]) ** 2
    for lag in (1, 3, 8, 20):
        meas = float((frames[lag:] * frames[:-lag]).mean()) / var
        theory = float((w * np.exp(-lag * dt / main_tau)).sum() / w.sum())
        print("lag", lag, "measured", meas, "theory", theory)
        assert abs(meas - theory) < 0.06
