"""CPU tests of the product's host logic and the C-ABI boundary (no GPU)."""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, load_golden


def test_host_filters_match_reference_golden_and_oracle():
    from pyturb_amd.spectrum import build_filters

    g = load_golden("generate_n256_seed1.npz")
    amp, f, sh_amp = build_filters(256, 0.03125, 0.15, 25.0, 3)
    idx = np.r_[0:9, 248:256]
    np.testing.assert_allclose(amp[np.ix_(idx, idx)], g["amp_block"], rtol=1e-12)
    np.testing.assert_allclose(amp[5], g["amp_row5"], rtol=1e-12)
    np.testing.assert_allclose(sh_amp, g["sh_amp"], rtol=1e-12)
    ref = O.build_filters(100, 0.07, 0.2, np.inf, subharmonics=5, power_law=3.4,
                          inner_scale=0.05)
    amp2, _, sh2 = build_filters(100, 0.07, 0.2, np.inf, 5, 3.4, 0.05)
    np.testing.assert_allclose(amp2, ref["amplitude"], rtol=1e-12)
    np.testing.assert_allclose(sh2, ref["sh_amp"], rtol=1e-12)


def test_host_flow_draw_matches_oracle_seed_tree():
    """The product's host draw of layer spectra follows the reference stream."""
    from pyturb_amd.flow import draw_layer

    class _T:  # minimal template carrying what draw_layer reads
        pass

    filt = O.build_filters(32, 0.25, 0.1, 25.0, subharmonics=2)
    t = _T()
    t.n, t.dtype, t._n_sh = 32, np.dtype("float64"), 2
    t._amplitude_host, t._sh_amp_host = filt["amplitude"], filt["sh_amp"]
    spec, cf = draw_layer(np.random.default_rng(9), t)
    spec_o, cf_o = O.draw_flow(np.random.default_rng(9), filt)
    np.testing.assert_allclose(spec, spec_o, rtol=1e-14)
    np.testing.assert_allclose(cf, cf_o, rtol=1e-14)


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "pyturb_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|long long|const char\*)\s+(pt_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_header_symbol():
    from pyturb_amd import _hip

    so = ctypes.CDLL(_hip.LIB_PATH)
    syms = _header_symbols()
    assert "pt_generate" in syms and "pt_spectral_frames" in syms
    for name in syms:
        assert hasattr(so, name), name
        assert name in _hip.SIGNATURES, name
    assert _hip.lib().pt_abi_version() == _hip.ABI_VERSION == 6
    assert "pt_work_elems" in syms
    # host-only sizing: power-of-two fused path vs general path
    # frames: pairs of frames share one transform, plus the folded stack, its
    # Nyquist side table, the pair tables and Nyquist sums, and the uniform-step
    # words (16 check flags + 2 * 16 steps, uint32; each section rounded to 32 elements)
    r32 = lambda v: (v + 31) // 32 * 32
    assert _hip.lib().pt_frames_work_elems(512, 10, 9, 0) == (
        5 * 512 * 512 + 9 * 512 * 512 + 9 * 1024 + r32(5 * 2 * 9 * (512 // 8 + 9)) + 5 * 1024 + 32)
    assert _hip.lib().pt_frames_work_elems(100, 10, 9, 0) == 2 * 10 * 100 * 100
    assert _hip.lib().pt_frames_work_elems(512, 10, 17, 0) == 2 * 10 * 512 * 512
    for L in (1, 9, 16, 40):
        assert _hip.lib().pt_work_elems(1, 512, 10, 3) >= _hip.lib().pt_frames_work_elems(512, 10, L, 3)
    assert _hip.lib().pt_work_elems(1, 100, 10, 0) == 2 * 10 * 100 * 100
    assert _hip.lib().pt_work_elems(0, 64, 5, 2) == 3 * 64 * 64 + 2 * 64 + 5 * 2 * 64 + (5 * 64 + 1) // 2


def test_c_abi_rejects_bad_arguments_without_launching():
    """Argument checks run on the host and return PT_ERR_ARG (no GPU needed)."""
    from pyturb_amd import _hip

    L = _hip.lib()
    rc = L.pt_generate(None, None, None, None, None, 0, 4, 64, 0, 64, 1.0, None, 0,
                       None, None, None)
    assert rc == _hip.PT_ERR_ARG
    assert b"NULL" in L.pt_last_error()
    rc = L.pt_spectral_frames(None, 1, 64, 0.1, None, 1, None, 0, 0, 64, 1.0, None, None, None)
    assert rc == _hip.PT_ERR_ARG
    rc = L.pt_fft2(None, None, None, 1, 64, 64, 1, None)
    assert rc == _hip.PT_ERR_ARG
    with pytest.raises(ValueError, match="pt_fft2"):
        _hip.check(rc, "pt_fft2")


def test_device_names():
    from pyturb_amd.backend import resolve_device

    assert resolve_device("mi355x") == 0
    assert resolve_device("hip:1") == 1
    assert resolve_device("gpu") == 0
    with pytest.raises(ValueError, match="no CPU"):
        resolve_device("cpu")
    with pytest.raises(ValueError, match="Unknown device"):
        resolve_device("tpu")


def test_config_validation_contract():
    """The reference's ValueError contract is kept before any device work."""
    import pyturb_amd as pa

    with pytest.raises(ValueError, match="integer"):
        pa.PhaseScreen(n=32.0, pixel_scale=0.01, r0=0.1)
    with pytest.raises(ValueError, match="dtype"):
        pa.PhaseScreen(n=32, pixel_scale=0.01, r0=0.1, dtype="float16")
    with pytest.raises(ValueError):
        pa.PhaseScreen(n=1, pixel_scale=0.01, r0=0.1)
    with pytest.raises(ValueError):
        pa.PhaseScreen(n=32, pixel_scale=-1, r0=0.1)
    with pytest.raises(ValueError):
        pa.PhaseScreen(n=32, pixel_scale=0.01, r0=0.1, L0=-5)
    with pytest.raises(ValueError):
        pa.PhaseScreen(n=32, pixel_scale=0.01, r0=0.1, power_law=2.0)
    layers = pa.get_profile("single-layer")
    with pytest.raises(ValueError):
        pa.Atmosphere(layers)
    with pytest.raises(ValueError):
        pa.Atmosphere(layers, r0=0.1, seeing=0.8)
    with pytest.raises(ValueError, match="engine"):
        pa.Atmosphere(layers, r0=0.1, engine="fft")
    with pytest.raises(ValueError, match="interp"):
        pa.Atmosphere(layers, r0=0.1, interp="quintic")
    with pytest.raises(ValueError, match="finite outer scale"):
        pa.Atmosphere(layers, r0=0.1, engine="extrude", L0=np.inf)


def test_structure_function_host_estimator():
    import pyturb_amd as pa

    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 16, 16))
    r, d = pa.structure_function(x, 0.5, max_separation=4)
    r2, d2 = O.structure_function(x, 0.5, max_separation=4)
    np.testing.assert_allclose(r, r2)
    np.testing.assert_allclose(d, d2)


def test_tune_state_is_host_side_and_validated():
    """pt_tune_* select launch shapes per size; arguments are checked on the
    host and the effective shape clamps to [n/8, n^2/8] threads."""
    from pyturb_amd import _hip, tune

    lib = _hip.lib()
    assert lib.pt_tune_variants(_hip.PT_TUNE_GEN_ROWS) == 5
    assert lib.pt_tune_variants(_hip.PT_TUNE_GEN_COLS) == 8
    assert lib.pt_tune_variants(_hip.PT_TUNE_FRAMES) == 2
    assert [(v["threads"], v["tile"]) for v in tune.variants("frames_order", 512)] == [(1024, 16),
                                                                                      (256, 4)]
    assert [(v["threads"], v["tile"]) for v in tune.variants("frames_order", 1024)] == [(1024, 8),
                                                                                       (256, 2)]
    assert lib.pt_tune_variants(7) == -_hip.PT_ERR_ARG
    assert lib.pt_tune_set(7, 512, 0) == _hip.PT_ERR_ARG
    assert lib.pt_tune_set(0, 500, 0) == _hip.PT_ERR_ARG      # not a power of two
    assert lib.pt_tune_set(0, 4096, 0) == _hip.PT_ERR_ARG     # above 2048
    assert lib.pt_tune_set(0, 512, 5) == _hip.PT_ERR_ARG      # rows has variants 0..4
    assert lib.pt_tune_get(1, 100) == -_hip.PT_ERR_ARG
    # shapes: (threads, tile, xcd order, largest radix)
    shapes = [(v["threads"], v["tile"], v["xcd"], v["radix"]) for v in tune.variants("gen_cols", 512)]
    assert shapes == [(1024, 16, 1, 8), (512, 8, 1, 8), (256, 4, 1, 8), (1024, 16, 0, 8),
                      (512, 8, 0, 8), (256, 4, 0, 8), (1024, 16, 1, 4), (1024, 16, 0, 4)]
    assert [(v["threads"], v["tile"], v["radix"]) for v in tune.variants("gen_cols", 2048)
            if v["xcd"]] == [(1024, 4, 8), (512, 2, 8), (256, 1, 8), (1024, 4, 4)]
    assert [(v["threads"], v["tile"], v["radix"]) for v in tune.variants("gen_rows", 32)] == [
        (128, 32, 8), (128, 32, 4)]
    assert [(v["threads"], v["tile"], v["radix"]) for v in tune.variants("gen_rows", 2048)] == [
        (256, 1, 8), (512, 2, 8), (256, 1, 4), (512, 2, 4)]
    assert tune.describe("gen_cols", 512, 6) == "1024 threads x 16 cols xcd, radix 4"
    assert [v["radix"] for v in tune.variants("frames_order", 512)] == [8, 8]
    try:
        tune.set_variant("gen_cols", 256, 4)
        assert tune.get("gen_cols", 256) == 4 and tune.get("gen_cols", 512) == 0
    finally:
        tune.reset()
    assert tune.get("gen_cols", 256) == 0
    with pytest.raises(ValueError):
        tune.set_variant("gen_rows", 256, 9)
    with pytest.raises(ValueError):
        tune.variants("fft", 256)


def test_ring_rows_hold_their_stencil_and_far_rows():
    """The index-shift ring (pyturb_amd.ring) against a brute-force model of
    the reference's copying ring (src/pyturb/ring.py:9-29, extrude.py:424-449):
    every planned row's stencil and far rows must hold the virtual rows the
    reference would read, through compactions and wraps; extend() and rows()
    must give plan()'s rows; and base/fill must follow the reference rule."""
    import math
    import random

    from pyturb_amd.ring import Ring

    rng = random.Random(5)
    for trial in range(200):
        cap = rng.randint(12, 60)
        m = rng.choice([1, 2, 3])
        far = tuple(sorted(rng.sample(range(m, m + 8), rng.randint(0, 3))))
        base = rng.randint(-20, 5)
        fill = rng.randint(max(m, far[-1] + 2 if far else m), min(cap, 30))
        ring = Ring(cap, m, base, fill, far)
        twin = Ring(cap, m, base, fill, far)
        mem = [None] * cap  # virtual row held by each physical row
        for i in range(fill):
            mem[ring.phys(i)] = base + i
        ref_base, ref_fill = base, fill  # the reference's copying ring
        top, first = base + fill - 1, float(base)
        for _ in range(40):
            top += rng.randint(0, 7)
            first += rng.uniform(0, 5)
            rows = ring.plan(top, first)
            f0, k = twin.extend(top, first)
            assert twin.rows(f0, k) == rows
            for row in rows:
                if ref_fill == cap:  # ring.py:23-29, before the extrusion
                    stencil = ref_base + ref_fill - m
                    if far:
                        stencil = min(stencil, ref_base + ref_fill - 1 - far[-1])
                    keep = max(1, min(math.floor(first) - 1, stencil) - ref_base)
                    ref_base, ref_fill = ref_base + keep, ref_fill - keep
                v = ref_base + ref_fill  # virtual row being extruded
                assert [mem[z] for z in row[1]] == [v - m + q for q in range(m)]
                if far:
                    assert [mem[z] for z in row[2]] == [v - 1 - d for d in far]
                mem[row[0]] = v
                ref_fill += 1
            assert (ring.base, ring.fill) == (ref_base, ref_fill) == (twin.base, twin.fill)
