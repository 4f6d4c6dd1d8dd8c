"""Extruder boiling on the device against the reference and theory.

* the reference's boiled extruder readouts (golden, extrude.py:472-496,
  660-752, NumPy path), with the reference's boiling draws injected;
* with no wind the readout is a pure AR(1) in time: frames k steps apart
  correlate as a^k = exp(-k dt / tau_boil), and the variance is kept.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from conftest import load_golden

pytestmark = pytest.mark.gpu


def _np(x):
    return x.double().cpu().numpy()


def test_extruder_boiling_matches_reference_golden(hip):
    from pyturb_amd.extrude import ExtrudedAtmosphere

    print("extruder boiling golden n=32", flush=True)
    g = load_golden("extrude_boil_n32.npz")["frames"]
    ext = ExtrudedAtmosphere(n=32, pixel_scale=0.25, layer_r0=[0.2, 0.35], layer_L0=[25.0, 25.0],
                             layer_wind=[(6.0, 2.5), (-3.0, 11.0)],
                             layer_altitude_los=[0.0, 8000.0], field_of_view_pix=[0.0, 3.0],
                             interp="cubic", dtype="float64", seeds=[11, 12],
                             tau_boil=[0.03, 0.08], boil_seed=77)
    rng = np.random.default_rng(77)  # the reference's boiling stream, its order
    m, W = ext.m, ext.width
    ext._boil_draws = lambda step, l, h: (rng.standard_normal(m * W),
                                          rng.standard_normal((max(h - m, 0), W)))
    got = []
    for k in range(5):
        ext.set_time(k * 0.01)
        got.append(_np(ext.integrate(0.0, 0.0)))
        got.append(_np(ext.integrate(2e-5, -1e-5)))
        ext.boil_step(0.01)
    err = O.rel_rms(np.stack(got), g)
    print("rel rms", err)
    assert err < 1e-5


def test_extruder_boiling_is_ar1_in_time(hip):
    import torch

    import pyturb_amd as pa

    print("extruder boiling AR(1) n=32", flush=True)
    tau, dt = 0.02, 0.004
    atm = pa.Atmosphere([pa.Layer(0.0, 1.0, 0.0, 0.0, 25.0)], r0=0.15, diameter=2.0, n=32,
                        engine="extrude", tau_boil=tau, seed=9, device="mi355x")
    frames = torch.stack([f for _, f in atm.frames(dt, 1500)]).double()
    frames = frames - frames.mean(dim=0)
    var = (frames * frames).mean(dim=(1, 2))
    assert float(var[750:].mean() / var[:750].mean()) == pytest.approx(1.0, abs=0.15)
    v = float((frames * frames).mean())
    for lag in (1, 3, 10):
        meas = float((frames[lag:] * frames[:-lag]).mean()) / v
        print("lag", lag, "measured", meas, "theory", np.exp(-lag * dt / tau))
        assert abs(meas - np.exp(-lag * dt / tau)) < 0.08
