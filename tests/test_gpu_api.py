"""GPU tests of the public API's semantics that the reference's own tests pin:

* ``FourierFlowScreen.translate`` against the oracle's restatement of
  flow.py:96-129 on the same host draw, and the exact integer-pixel roll
  (reference tests/test_atmosphere.py:328-334);
* ``time_to_wrap`` = screen period / wind speed, infinite for the extruder
  (tests/test_atmosphere.py:74-82);
* the wrap warning: once per run, re-armed by ``reset``, silent within a
  period (tests/test_atmosphere.py:85-109), and raised by ``frames`` at the
  frame that wraps (atmosphere.py:1146-1149 checks frame by frame);
* ``reset`` followed by ``frames`` replays the frozen-flow run bit for bit
  (atmosphere.py:1221-1243), on both engines;
* the reuse guard of ``pt_spectral_rows_folded``: a fold left by a
  uniform-step ``pt_spectral_frames`` call gives NaN frames, the fold of
  ``pt_spectral_fold`` gives the frames of the full call.
"""

from __future__ import annotations

import warnings

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-5


def _np(t):
    return t.detach().cpu().numpy()


def test_flow_translate_matches_oracle(hip):
    import pyturb_amd as pa
    from pyturb_amd.flow import FourierFlowScreen

    print("flow translate n=128, 3 levels", flush=True)
    ps = pa.PhaseScreen(128, 0.02, 0.15, 25.0, subharmonics=3, seed=3, device="mi355x")
    fl = FourierFlowScreen(ps, seed=5)
    filt = O.build_filters(128, 0.02, 0.15, 25.0, subharmonics=3)
    # sub-pixel, negative and many-period offsets (exact phase reduction)
    for sx, sy in [(0.0, 0.0), (0.0137, -0.291), (-3.77, 1.05), (57.3, -211.9)]:
        got = _np(fl.translate(sx, sy))
        want = O.flow_translate(fl._spectrum_host, fl._sh_coeffs_host, filt, sx, sy)
        err = O.rel_rms(got, want)
        assert err < REL, (sx, sy, err)


def test_flow_translate_integer_pixels_is_a_roll(hip):
    import pyturb_amd as pa
    from pyturb_amd.flow import FourierFlowScreen

    ps = pa.PhaseScreen(128, 0.02, 0.15, 25.0, subharmonics=0, seed=3, device="mi355x")
    fl = FourierFlowScreen(ps, seed=5)
    base = _np(fl.translate(0.0, 0.0))
    shifted = _np(fl.translate(7 * 0.02, 0.0))
    # a +s shift reads the screen at x + s: a roll by -7 along axis 0
    scale = np.sqrt(np.mean(base ** 2))
    assert np.max(np.abs(shifted - np.roll(base, -7, axis=0))) < 2e-5 * scale
    both = _np(fl.translate(-3 * 0.02, 5 * 0.02))
    assert np.max(np.abs(both - np.roll(np.roll(base, 3, axis=0), -5, axis=1))) < 2e-5 * scale


def _one_layer(**kw):
    import pyturb_amd as pa

    layer = pa.Layer(0.0, 1.0, wind_speed=10.0, wind_direction=0.0)
    return pa.Atmosphere([layer], r0=0.15, diameter=8.0, n=64, seed=0, device="mi355x", **kw)


def test_time_to_wrap_is_period_over_speed(hip):
    atm = _one_layer()
    assert atm.time_to_wrap == pytest.approx(8.0 / 10.0)
    assert _one_layer(engine="extrude").time_to_wrap == float("inf")


def test_wrap_warning_fires_once_and_reset_rearms(hip):
    import pyturb_amd as pa

    atm = _one_layer()
    wrap_t = atm.time_to_wrap
    with warnings.catch_warnings():
        warnings.simplefilter("error", pa.PeriodicWrapWarning)
        atm.opd(wrap_t * 0.5)  # within one period: silent
    with pytest.warns(pa.PeriodicWrapWarning, match=r"cumulative travel \(12\.00 m\)"):
        atm.opd(wrap_t * 1.5)
    with warnings.catch_warnings():
        warnings.simplefilter("error", pa.PeriodicWrapWarning)
        atm.opd(wrap_t * 3.0)  # already warned in this run
    atm.reset()
    with pytest.warns(pa.PeriodicWrapWarning):
        atm.opd(wrap_t * 1.5)


def test_frames_warn_at_the_wrapping_frame(hip):
    import pyturb_amd as pa

    for kw in ({}, {"tau_boil": 0.5}):
        atm = _one_layer(**kw)
        seen = []
        with warnings.catch_warnings(record=True) as rec:
            warnings.simplefilter("always", pa.PeriodicWrapWarning)
            for k, (t, _) in enumerate(atm.frames(0.1, 12)):
                hits = [w for w in rec if issubclass(w.category, pa.PeriodicWrapWarning)]
                seen.append(len(hits))
        # t = 0.9 s is the first frame past the 8 m period at 10 m/s: the
        # warning comes with that frame (index 9), not when a chunk is built
        assert seen == [0] * 9 + [1] * 3, (kw, seen)


def test_reset_replays_frozen_flow_frames(hip):
    import torch

    for engine in ("spectral", "extrude"):
        print("reset replay", engine, flush=True)
        atm = _one_layer(engine=engine, subharmonics=3)
        first = torch.stack([f.clone() for _, f in atm.frames(2e-3, 6)])
        assert atm.time == pytest.approx(12e-3)
        atm.reset()
        assert atm.time == 0.0
        again = torch.stack([f.clone() for _, f in atm.frames(2e-3, 6)])
        assert torch.equal(first, again), engine


def test_folded_reuse_guard(hip):
    """pt_spectral_rows_folded on a stepped fold gives NaN frames; on
    pt_spectral_fold's fold it gives the full call's frames."""
    import torch

    from pyturb_amd import ops

    dev = torch.device("cuda")
    rng = np.random.default_rng(8)
    L, n, F = 3, 64, 6
    spectra = torch.from_numpy((rng.standard_normal((L, n, n)) + 1j * rng.standard_normal((L, n, n)))
                               .astype(np.complex64)).to(dev)
    times = np.arange(F) * 1e-3  # equally spaced: the full call folds the step in
    v = rng.uniform(-20, 20, (L, 2))
    disp = torch.from_numpy(times[:, None, None] * v[None]).to(dev)
    work = torch.empty((ops.frames_work_elems(n, F, L, 0),), dtype=torch.complex64, device=dev)
    full = ops.spectral_frames(spectra, disp, None, 0.125, work=work).clone()
    stale = ops.spectral_frames(spectra, disp, None, 0.125, work=work, folded=True)
    assert torch.isnan(stale).all()
    ops.spectral_fold(spectra, F, work)
    reused = ops.spectral_frames(spectra, disp, None, 0.125, work=work, folded=True)
    assert O.rel_rms(_np(reused), _np(full)) < 1e-6
