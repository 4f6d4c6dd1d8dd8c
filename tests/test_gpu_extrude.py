"""GPU parity of the extrusion engine: ring recurrence, innovation GEMM and
the fused readout, against the oracle and the reference's golden frames.

Bar: relative RMS <= 1e-5 against the FP64 oracle on identical innovations.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from conftest import config2_layer_tuples, load_golden

pytestmark = pytest.mark.gpu

REL = 1e-5


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("interp", ["linear", "cubic", "lanczos"])
@pytest.mark.parametrize("n,W,cap", [(16, 30, 40), (80, 150, 160)])
def test_readout_matches_oracle_on_random_ring(hip, interp, n, W, cap):
    """Pure-function check of pt_readout: ring wrap, row/column clamping,
    per-direction magnification; n=16 runs the direct-gather kernel, n=80
    the LDS box kernel with partial tiles.  The last direction magnifies
    above 1, so its layers' footprints exceed the box (sized for
    magnification <= 1) and take the box kernel's direct-gather fallback,
    one layer fitting and two not."""
    import torch
    from pyturb_amd.infinite import readout
    from pyturb_amd.ring import Ring

    print(f"readout {interp} n={n}", flush=True)
    rng = np.random.default_rng(3)
    L = 3
    buf = rng.standard_normal((L, cap, W))
    rings = []
    for start, fill in [(0, cap), (cap // 3, cap - 5), (cap - 3, cap // 2)]:
        r = Ring(cap, 2, 0, fill)
        r.start = start
        rings.append(r)
    cs = [(np.cos(a), np.sin(a)) for a in (0.3, 2.0, -1.1)]
    c = 0.4 * n  # centre the footprints in the buffer; some taps still clamp
    geom = np.array([[[c + 0.3, 0.7, 1.0], [c + 2.9, -1.3, 1.0], [c - 0.5, 0.2, 1.0]],
                     [[c + 1.1, 2.2, 0.8], [c + 4.4, 0.0, 0.6], [c - 2.0, -3.0, 0.9]],
                     [[2.5, -0.4, 1.0], [float(cap), 1.5, 1.0], [1.0, float(W), 0.7]],
                     [[c + 0.6, -0.9, 1.3], [c + 1.7, 0.4, 1.0], [c - 1.0, 1.1, 1.5]]])
    dev = torch.device("cuda")
    got = _np(readout(torch.from_numpy(buf.astype(np.float32)).to(dev), cap, W, rings, cs,
                      geom, n, interp, scale=0.5))
    local = np.stack([np.roll(buf[l], -rings[l].start, axis=0) for l in range(L)])
    fill = np.array([r.fill for r in rings])
    for d in range(len(geom)):
        along, perp = zip(*[O.pupil_grids(n, cc, ss, W, geom[d, l, 2])
                            for l, (cc, ss) in enumerate(cs)])
        ref = O.readout(local.astype(np.float32).astype(np.float64), fill, np.stack(along),
                        np.stack(perp), geom[d, :, 0], geom[d, :, 1], interp) * 0.5
        assert O.rel_rms(got[d], ref) < REL, (d, O.rel_rms(got[d], ref))


@pytest.mark.parametrize("interp", ["linear", "cubic", "lanczos"])
def test_extruder_matches_reference_golden(hip, interp):
    from pyturb_amd.extrude import ExtrudedAtmosphere

    print(f"extruder golden {interp}", flush=True)
    g = load_golden("extrude_n32.npz")[interp]
    ext = ExtrudedAtmosphere(n=32, pixel_scale=0.25, layer_r0=[0.2, 0.35],
                             layer_L0=[25.0, 25.0], layer_wind=[(6.0, 2.5), (-3.0, 11.0)],
                             layer_altitude_los=[0.0, 8000.0], field_of_view_pix=[0.0, 3.0],
                             interp=interp, dtype="float64", seeds=[11, 12])
    k = 0
    for t in (0.0, 0.01, 0.037, 0.4):
        ext.set_time(t)
        for th in ((0.0, 0.0), (2e-5, -1e-5)):
            got = _np(ext.integrate(*th))
            assert O.rel_rms(got, g[k]) < REL, (t, th, O.rel_rms(got, g[k]))
            k += 1


@pytest.mark.parametrize("interp", ["cubic", "linear"])
def test_infinite_matches_reference_golden(hip, interp):
    import pyturb_amd as pa
    from pyturb_amd.infinite import InfinitePhaseScreen

    g = load_golden("infinite_n32_seed5.npz")[interp]
    scr = InfinitePhaseScreen(n=32, pixel_scale=0.1, r0=0.15, L0=25.0, interp=interp, seed=5,
                              dtype="float64")
    seq = [scr.screen, scr.advance(0.3), scr.step(1), scr.advance(2.7), scr.advance(40.0)]
    for k, s in enumerate(seq):
        assert O.rel_rms(_np(s), g[k]) < REL, k


def _config3_small(n=64, steps=120, interp="linear", seed=3):
    """9-layer profile at small n with injected innovations (seed 3)."""
    layers = config2_layer_tuples()
    dx = 8.0 / n
    r0_los = 0.15
    frac = np.array([ly[1] for ly in layers])
    frac = frac / frac.sum()
    r0s = r0_los * frac ** (-0.6)
    winds = [(s * np.cos(np.radians(d)), s * np.sin(np.radians(d))) for _, _, s, d, _ in layers]
    alts = [ly[0] for ly in layers]
    fov = [a * np.tan(30 * O.ARCSEC) / dx for a in alts]
    return dict(n=n, pixel_scale=dx, layer_r0=list(r0s), layer_L0=[25.0] * 9,
                layer_wind=winds, layer_altitude_los=alts, field_of_view_pix=fov,
                interp=interp)


def test_extruder_multilayer_run_matches_oracle(hip):
    """9 layers, 30" field, 6 ring directions + on-axis, 120 steps of 1 ms."""
    import torch
    from pyturb_amd.extrude import ExtrudedAtmosphere

    print("extruder 9-layer run", flush=True)
    kw = _config3_small()
    rng = np.random.default_rng(3)
    pool = {}

    def innov_host(l, k):
        key = l
        arr = pool.setdefault(key, [])
        new = rng.standard_normal((k, W)).astype(np.float32)
        arr.append(new)
        return new

    # oracle first (fixes W and draws), then the device with the same rows
    ref = O.OracleExtruder(kw["n"], kw["pixel_scale"], kw["layer_r0"], kw["layer_L0"],
                           kw["layer_wind"], kw["layer_altitude_los"], kw["field_of_view_pix"])
    W = ref.W
    seeds = [rng.standard_normal((s, W)).astype(np.float32) for s in ref.spans]
    ref.seed([s.astype(np.float64) for s in seeds])
    streams = [np.random.default_rng(100 + l) for l in range(9)]
    drawn = [[] for _ in range(9)]

    def o_innov(l, k):
        row = streams[l].standard_normal(W).astype(np.float32)
        drawn[l].append(row)
        return row.astype(np.float64)

    dev = torch.device("cuda")
    used = [0] * 9

    def d_innov(l, k):
        rows = np.stack(drawn[l][used[l]:used[l] + k])
        used[l] += k
        return torch.from_numpy(rows).to(dev)

    ext = ExtrudedAtmosphere(**kw, innovations=d_innov,
                             seed_screens=[torch.from_numpy(np.pad(s, ((0, 0), (0, 0)))).to(dev)
                                           for s in seeds])
    assert ext.width == W
    ang = np.radians(np.arange(6) * 60.0)
    thx = np.concatenate([[0.0], np.tan(30 * O.ARCSEC) * np.cos(ang)])
    thy = np.concatenate([[0.0], np.tan(30 * O.ARCSEC) * np.sin(ang)])
    worst = 0.0
    for step in range(121):
        t = step * 1e-3
        ref.set_time(t, o_innov)
        ext.set_time(t)
        if step % 40 == 0:
            got = _np(ext.integrate_dirs(thx, thy))
            for d in range(7):
                want = ref.integrate(thx[d], thy[d], interp="linear")
                worst = max(worst, O.rel_rms(got[d], want))
    print("worst rel rms", worst)
    assert worst < REL


def test_extruder_big_jump_equals_small_steps(hip):
    import torch
    from pyturb_amd.extrude import ExtrudedAtmosphere

    kw = _config3_small(n=32, interp="cubic")
    dev = torch.device("cuda")

    def make():
        streams = [np.random.default_rng(50 + l) for l in range(9)]

        def innov(l, k):
            return torch.from_numpy(streams[l].standard_normal((k, big.width if False else W))
                                    .astype(np.float32)).to(dev)
        return innov

    probe = ExtrudedAtmosphere(**kw, seeds=list(range(9)))
    W = probe.width
    a = ExtrudedAtmosphere(**kw, seeds=list(range(9)), innovations=make())
    a.set_time(0.7)
    jumped = _np(a.integrate())
    b = ExtrudedAtmosphere(**kw, seeds=list(range(9)), innovations=make())
    for k in range(1, 71):
        b.set_time(k * 0.01)
    stepped = _np(b.integrate())
    assert O.rel_rms(jumped, stepped) < 1e-6
