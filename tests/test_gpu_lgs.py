"""Spectral LGS cone on the device (per-layer frames + pt_lgs_zoom).

* the reference's own LGS frames (golden, atmosphere.py:581-616,
  1000-1090) for linear, cubic and Lanczos taps, on- and off-axis, on a
  field-of-view grid that is not a power of two;
* the 9-layer 512^2 atmosphere with a 90 km beacon against the FP64 oracle;
* many frames in one call (several transform batches, the last one short)
  against the oracle;
* the cone with boiling, reference draws injected.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from conftest import config2_layer_tuples, load_golden

pytestmark = pytest.mark.gpu

LAYERS3 = [(0.0, 0.5, 8.0, 30.0, 25.0), (3000.0, 0.3, 14.0, 110.0, 25.0),
           (9000.0, 0.2, 21.0, 250.0, 25.0)]
LGS_DIRS = [(0.0, 0.0), (6.0, -4.0)]


def _np(x):
    return x.double().cpu().numpy()


@pytest.mark.parametrize("interp", ["linear", "cubic", "lanczos"])
def test_lgs_frames_match_reference_golden(hip, interp):
    import pyturb_amd as pa

    print("LGS golden", interp, flush=True)
    g = load_golden("lgs_n32_seed7.npz")
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in LAYERS3], r0=0.15, diameter=4.0, n=32,
                        subharmonics=3, field_of_view=10.0, lgs_altitude=20e3, interp=interp,
                        dtype="float64", seed=7, device="mi355x")
    assert atm.n_screen != 32
    got = np.stack([_np(atm.opd(t, directions=LGS_DIRS)) for t in (0.0, 0.013)])
    err = O.rel_rms(got, g[interp])
    print("rel rms", err)
    assert err < 1e-5


def test_lgs_config2_512_matches_oracle(hip):
    import pyturb_amd as pa

    print("LGS 512^2 9 layers", flush=True)
    kw = dict(diameter=8.0, n=512, L0=25.0, subharmonics=8, seed=2)
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in config2_layer_tuples()], r0=0.15,
                        lgs_altitude=90e3, device="mi355x", **kw)
    ora = O.OracleAtmosphere(config2_layer_tuples(), 0.15, lgs_altitude=90e3, **kw)
    times = [0.0, 0.0371, 0.5]
    got = np.stack([_np(f) for f in atm.opd_batch(times)])
    ref = np.stack([ora.opd(t) for t in times])
    err = O.rel_rms(got, ref)
    print("rel rms", err)
    assert err < 1e-5


def test_lgs_many_frames_match_oracle(hip, monkeypatch):
    import pyturb_amd as pa
    import pyturb_amd.atmosphere as atmosphere_mod

    print("LGS 150 frames n=64 in batches of 40", flush=True)
    monkeypatch.setattr(atmosphere_mod, "_LGS_BUDGET", 4 * 9 * 64 * 64 * 40)
    kw = dict(diameter=8.0, n=64, L0=25.0, subharmonics=4, seed=3)
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in config2_layer_tuples()], r0=0.15,
                        lgs_altitude=90e3, interp="linear", device="mi355x", **kw)
    ora = O.OracleAtmosphere(config2_layer_tuples(), 0.15, lgs_altitude=90e3, interp="linear",
                             **kw)
    times = np.arange(150) * 1e-3
    got = _np(atm.opd_batch(times))
    ref = np.stack([ora.opd(t) for t in times])
    assert O.rel_rms(got, ref) < 1e-5


def test_lgs_with_boiling_matches_oracle(hip):
    import pyturb_amd as pa

    print("LGS + boiling n=32", flush=True)
    kw = dict(diameter=4.0, n=32, subharmonics=3, dtype="float64", seed=6)
    ora = O.OracleAtmosphere(LAYERS3, 0.15, tau_boil=[0.02, np.inf, 0.05], lgs_altitude=20e3,
                             **kw)
    draws = [ora.draw_boil_noise() for _ in range(4)]
    ora.reset()
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in LAYERS3], r0=0.15,
                        tau_boil=[0.02, np.inf, 0.05], lgs_altitude=20e3, device="mi355x", **kw)
    atm._boil_noise = lambda k: draws[k]
    got = np.stack([_np(f) for _, f in atm.frames(0.004, 3)] + [_np(atm.evolve(0.004))])
    ref = [f for _, f in ora.frames(0.004, 3, noises=draws)]
    ref.append(ora.evolve(0.004, noise=draws[3]))
    err = O.rel_rms(got, np.stack(ref))
    print("rel rms", err)
    assert err < 1e-5


@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_lgs_zoom_wide_windows_match_numpy(hip, interp):
    """pt_lgs_zoom with magnification 1.4 on two of three layers: their
    windows exceed the LDS box and take the direct path, the third (0.9)
    is staged; against the separable sum evaluated in NumPy."""
    import torch

    from pyturb_amd import ops
    from pyturb_amd.atmosphere import _lgs_zoom_taps

    print("LGS zoom wide windows", interp, flush=True)
    n, ns, F = 48, 96, 3
    mags = [1.4, 0.9, 1.4]
    idx, w = _lgs_zoom_taps(n, ns, mags, interp)
    rng = np.random.default_rng(12)
    scr = rng.standard_normal((len(mags), F, ns, ns)).astype(np.float32)
    dev = torch.device("cuda", 0)
    got = ops.lgs_zoom(torch.tensor(scr, device=dev), torch.tensor(idx, device=dev),
                       torch.tensor(w, device=dev), n, scale=0.7).double().cpu().numpy()
    ref = np.zeros((F, n, n))
    s64, w64 = scr.astype(np.float64), w.astype(np.float64)
    for l in range(len(mags)):
        for f in range(F):
            # rows then columns: sum_a sum_b w_a(i) w_b(j) S[idx_a(i), idx_b(j)]
            rows = np.einsum("ai,aic->ic", w64[l], s64[l, f][idx[l]])
            ref[f] += np.einsum("bj,ibj->ij", w64[l], rows[:, idx[l]])
    ref *= 0.7
    assert O.rel_rms(got, ref) < 1e-6
