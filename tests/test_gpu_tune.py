"""Autotuned launch shapes (pyturb_amd.tune): every shape of the row and
column passes reproduces the default shape's output bit for bit, for the
Monte-Carlo generate path and for the frame engine's pass B, and the tuner
selects, stores and re-applies its choices."""

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
            assert torch.equal(out, ref), (knob, var)
        tune.set_variant(knob, n, 0)


def test_every_column_shape_is_bit_identical_for_frames(hip, tuned_reset):
    import pyturb_amd as pa

    tune = tuned_reset
    print("tune shapes frames n=128", flush=True)
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in config2_layer_tuples()], r0=0.15, n=128,
                        seed=4, device="mi355x")
    times = np.arange(9) * 1e-3
    ref = atm.opd_batch(times).clone()
    for var in tune.variants("gen_cols", atm.n_screen):
        tune.set_variant("gen_cols", atm.n_screen, var["variant"])
        got = atm.opd_batch(times)
        assert got.equal(ref), var


def test_autotune_selects_stores_and_reloads(hip, tuned_reset, tmp_path):
    tune = tuned_reset
    print("autotune n=128,256", flush=True)
    path = str(tmp_path / "tune.json")
    res = tune.autotune((128, 256), screens=16, repeats=2, levels=3, cache=path)
    for n in (128, 256):
        for knob in ("gen_rows", "gen_cols"):
            valid = {v["variant"] for v in tune.variants(knob, n)}
            assert res[n][knob] in valid
            assert tune.get(knob, n) == res[n][knob]
            assert set(map(int, res[n]["ms"][knob])) == valid
    data = json.load(open(path))
    assert list(data) == [tune.device_key()]
    tune.reset()
    applied = tune.load_cache(path)
    assert applied == {n: {k: res[n][k] for k in ("gen_rows", "gen_cols")} for n in (128, 256)}
    for n in (128, 256):
        assert tune.get("gen_cols", n) == res[n]["gen_cols"]
