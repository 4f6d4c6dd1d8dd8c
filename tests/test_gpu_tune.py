"""Autotuned launch shapes (pyturb_amd.tune): every radix-8 shape of the row
and column passes reproduces the default shape's output bit for bit, and
every radix-4 plan agrees with it within 1e-6 relative RMS, for the
Monte-Carlo generate path and for the frame engine's pass B; the frame
engine's two transform orders agree with each other and with the FP64
oracle at every size; the tuner selects, stores and re-applies its
choices."""

from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import config2_layer_tuples

pytestmark = pytest.mark.gpu


@pytest.fixture
def tuned_reset():
    from pyturb_amd import tune

    tune.reset()
    yield tune
    tune.reset()


@pytest.mark.parametrize("n,count", [(64, 24), (256, 7), (1024, 3)])
def test_every_generate_shape_is_bit_identical(hip, tuned_reset, n, count):
    import torch

    from pyturb_amd.tune import _ensemble

    tune = tuned_reset
    print(f"tune shapes generate n={n} count={count}", flush=True)
    run, out = _ensemble(n, count, 2, 3, 5, torch.device("cuda", 0))
    run()
    ref = out.clone()
    for knob in ("gen_rows", "gen_cols"):
        for var in tune.variants(knob, n):
            tune.set_variant(knob, n, var["variant"])
            out.fill_(float("nan"))
            run()
            torch.cuda.synchronize()
            if var["radix"] == 8:
                assert torch.equal(out, ref), (knob, var)
            else:  # another plan: other rounding
                assert _rel_rms(out, ref) < 1e-6, (knob, var, _rel_rms(out, ref))
        tune.set_variant(knob, n, 0)


def _rel_rms(a, b):
    a, b = a.double(), b.double()
    return float(((a - b).norm() / b.norm()).item())


def test_every_column_shape_is_bit_identical_for_frames(hip, tuned_reset):
    import pyturb_amd as pa

    tune = tuned_reset
    print("tune shapes frames n=128 (row-first order)", flush=True)
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in config2_layer_tuples()], r0=0.15, n=128,
                        seed=4, device="mi355x")
    tune.set_variant("frames_order", atm.n_screen, 1)
    times = np.arange(9) * 1e-3
    ref = atm.opd_batch(times).clone()
    for var in tune.variants("gen_cols", atm.n_screen):
        tune.set_variant("gen_cols", atm.n_screen, var["variant"])
        got = atm.opd_batch(times)
        if var["radix"] == 8:
            assert got.equal(ref), var
        else:
            assert _rel_rms(got, ref) < 1e-6, var


@pytest.mark.parametrize("n,fov,frames", [(32, 0.0, 5), (64, 8.0, 7), (128, 0.0, 4),
                                          (256, 5.0, 3), (512, 0.0, 5), (1024, 0.0, 3),
                                          (2048, 0.0, 2)])
def test_frame_orders_agree_with_each_other_and_the_oracle(hip, tuned_reset, n, fov, frames):
    """Column-first and row-first frame engines on the same atmosphere:
    within 1e-6 relative RMS of each other, and within 1e-5 of the FP64
    oracle (crop margins, odd frame counts, 1..9 layers)."""
    import oracle as O
    import pyturb_amd as pa

    tune = tuned_reset
    print(f"frame orders n={n} fov={fov} frames={frames}", flush=True)
    layers = config2_layer_tuples()[: (9 if n <= 512 else 3)]
    kw = dict(diameter=8.0, n=n, L0=25.0, subharmonics=6, seed=5, field_of_view=fov)
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in layers], r0=0.15, device="mi355x", **kw)
    ns = atm.n_screen
    times = np.arange(frames) * 7e-3
    dirs = [(0.0, 0.0), (fov * 0.6, -fov * 0.3)] if fov else None
    got = {}
    for order in (0, 1):
        if ns in (32, 64, 128, 256, 512, 1024, 2048):
            tune.set_variant("frames_order", ns, order)
        got[order] = atm.opd_batch(times, directions=dirs).double().cpu().numpy()
    assert O.rel_rms(got[0], got[1]) < 1e-6
    if n <= 512:
        ora = O.OracleAtmosphere(layers, 0.15, **kw)
        if dirs is None:
            ref = np.stack([ora.opd(t) for t in times])
        else:
            ref = np.stack([ora.opd(t, directions=dirs) for t in times])
        assert O.rel_rms(got[0], ref) < 1e-5


def test_autotune_selects_stores_and_reloads(hip, tuned_reset, tmp_path):
    tune = tuned_reset
    print("autotune n=128,256", flush=True)
    path = str(tmp_path / "tune.json")
    res = tune.autotune((128, 256), screens=16, repeats=2, levels=3, cache=path)
    for n in (128, 256):
        for knob in tune.KNOBS:
            valid = {v["variant"] for v in tune.variants(knob, n)}
            assert res[n][knob] in valid
            assert tune.get(knob, n) == res[n][knob]
            assert set(map(int, res[n]["ms"][knob])) == valid
    data = json.load(open(path))
    assert list(data) == [tune.device_key()]
    tune.reset()
    applied = tune.load_cache(path)
    assert applied == {n: {k: res[n][k] for k in tune.KNOBS} for n in (128, 256)}
    for n in (128, 256):
        for knob in tune.KNOBS:
            assert tune.get(knob, n) == res[n][knob]


@pytest.mark.parametrize("n,off,n_out,frames,layers,levels", [(128, 9, 100, 5, 4, 4),
                                                               (512, 40, 400, 3, 16, 12),
                                                               (32, 0, 32, 1, 1, 0)])
def test_frame_orders_agree_on_cropped_odd_calls(hip, tuned_reset, n, off, n_out, frames, layers,
                                                 levels):
    """ops.spectral_frames with a crop window, an odd frame count, up to 16
    layers and 12 subharmonic levels: both transform orders agree."""
    import torch

    from pyturb_amd import ops

    tune = tuned_reset
    print(f"frame orders cropped n={n} off={off} F={frames} L={layers} P={levels}", flush=True)
    rng = np.random.default_rng(n + frames)
    dev = torch.device("cuda", 0)
    spec = (rng.standard_normal((layers, n, n)) + 1j * rng.standard_normal((layers, n, n))) / n
    spectra = torch.tensor(spec.astype(np.complex64), device=dev)
    disp = torch.tensor(rng.uniform(-3.0, 3.0, (frames, layers, 2)), dtype=torch.float64,
                        device=dev)
    cf = None
    if levels:
        c = rng.standard_normal((levels, layers, 3, 3)) + 1j * rng.standard_normal((levels, layers, 3, 3))
        cf = torch.tensor(c * 0.05, dtype=torch.complex128, device=dev)
    got = []
    for order in (0, 1):
        tune.set_variant("frames_order", n, order)
        got.append(ops.spectral_frames(spectra, disp, cf, 0.02, crop_offset=off, n_out=n_out,
                                       scale=1.7).double().cpu().numpy())
    err = np.sqrt(np.mean((got[0] - got[1]) ** 2) / np.mean(got[1] ** 2))
    print("rel rms", err)
    assert got[0].shape == (frames, n_out, n_out)
    assert err < 1e-6


def test_tune_cache_applied_when_the_library_loads(hip, tuned_reset, tmp_path, monkeypatch):
    """PYTURB_AMD_TUNE_CACHE names a cache file whose choices for this
    device are applied when the library is (re)loaded; other devices'
    entries are ignored."""
    from pyturb_amd import _hip

    tune = tuned_reset
    path = tmp_path / "cache.json"
    with open(path, "w") as fh:
        json.dump({tune.device_key(): {"128": {"gen_rows": 2, "gen_cols": 4, "frames_order": 1}},
                   "gfx000:1CU": {"256": {"gen_cols": 5}}}, fh)
    monkeypatch.setenv("PYTURB_AMD_TUNE_CACHE", str(path))
    monkeypatch.setattr(_hip, "_lib", None)
    _hip.lib()  # the same shared object again: its process-wide selection is updated
    assert (tune.get("gen_rows", 128), tune.get("gen_cols", 128), tune.get("frames_order", 128)) \
        == (2, 4, 1)
    assert tune.get("gen_cols", 256) == 0
