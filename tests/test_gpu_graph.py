"""HIP-graph replay of the frame engine (pyturb_amd.graph.CapturedFrames):
every replay equals the eager call on the same times, with eager work and
synchronisations between replays."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import config2_layer_tuples

pytestmark = pytest.mark.gpu


def test_captured_frames_replays_match_eager(hip):
    import torch

    import pyturb_amd as pa
    from pyturb_amd.graph import CapturedFrames

    print("graph replay n=128, 8 frames x 3 directions", flush=True)
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in config2_layer_tuples()], r0=0.15, n=128,
                        field_of_view=20.0, seed=2, device="mi355x")
    dirs = [(0.0, 0.0), (10.0, -5.0), (-12.0, 3.0)]
    cap = CapturedFrames(atm, 8, directions=dirs)
    rng = np.random.default_rng(0)
    for k in range(5):
        times = np.sort(rng.uniform(0.0, 0.2, 8))
        got = cap(times).clone()
        torch.cuda.synchronize()
        want = atm.opd_batch(times, directions=dirs)  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(got, want), k


@pytest.mark.parametrize("frames,ndirs", [(1, 1), (3, 1), (2, 3)])
def test_captured_few_frames_replays_match_eager(hip, frames, ndirs):
    """Calls of at most 8 frames build both table sets in one launch and read
    the displacements from pinned host memory; every replay (the two graphs
    alternate) equals the eager call, with eager work between replays."""
    import torch

    import pyturb_amd as pa
    from pyturb_amd.graph import CapturedFrames

    print(f"graph replay n=256, {frames} frames x {ndirs} directions", flush=True)
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in config2_layer_tuples()], r0=0.15, n=256,
                        field_of_view=20.0, seed=2, device="mi355x")
    dirs = [(0.0, 0.0), (10.0, -5.0), (-12.0, 3.0)][:ndirs]
    cap = CapturedFrames(atm, frames, directions=dirs if ndirs > 1 else None)
    rng = np.random.default_rng(frames)
    for k in range(6):
        times = np.sort(rng.uniform(0.0, 0.2, frames))
        got = cap(times).clone()
        torch.cuda.synchronize()
        want = (atm.opd_batch(times, directions=dirs) if ndirs > 1 else atm.opd_batch(times))
        torch.cuda.synchronize()
        assert torch.equal(got, want), k


@pytest.mark.parametrize("frames,ndirs", [(1, 1), (2, 3), (12, 1)])
def test_folded_frames_loop_matches_eager(hip, frames, ndirs):
    """FoldedFrames (one pt_spectral_frames_folded call per step on a fold
    made once, pinned displacement tables) equals the eager call."""
    import torch

    import pyturb_amd as pa
    from pyturb_amd.graph import FoldedFrames

    # a 41.2" field widens the 256-pixel pupil to a 512-pixel screen (the
    # fused engine needs a power of two); directions up to 41"
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in config2_layer_tuples()], r0=0.15, n=256,
                        field_of_view=41.2 if ndirs > 1 else 0.0, seed=2, device="mi355x")
    assert atm.n_screen in (256, 512)
    dirs = [(0.0, 0.0), (30.0, -15.0), (-22.0, 33.0)][:ndirs]
    loop = FoldedFrames(atm, frames, directions=dirs if ndirs > 1 else None)
    rng = np.random.default_rng(frames + 10)
    for k in range(6):
        times = np.sort(rng.uniform(0.0, 0.2, frames))
        got = loop(times).clone()
        want = (atm.opd_batch(times, directions=dirs) if ndirs > 1 else atm.opd_batch(times))
        torch.cuda.synchronize()
        assert torch.equal(got, want), k


def test_captured_mixed_ensemble_replays_match_eager(hip):
    """Two size classes and two outer-scale groups recorded once; the noise is
    redrawn in place before each replay and the result must equal the eager
    call on the same noise."""
    import torch

    from pyturb_amd import _hip, ops
    from pyturb_amd.graph import Captured
    from pyturb_amd.spectrum import build_filters

    print("graph replay ensemble 64x32^2 + 8x128^2, 2 groups", flush=True)
    dev = _hip.device(0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(11)
    classes = []
    for n, count in ((32, 64), (128, 8)):
        amps, shs = [], []
        for L0 in (25.0, 50.0):
            amp, _, sh = build_filters(n, 8.0 / n, 0.15, L0, 3)
            amps.append(torch.tensor(amp, dtype=torch.float32, device=dev))
            shs.append(torch.tensor(sh, dtype=torch.float64, device=dev))
        c = dict(n=n, count=count, amp=torch.stack(amps), sh=torch.stack(shs),
                 noise=torch.empty((2, 2, (count + 1) // 2, n, n), device=dev),
                 shn=torch.empty((2, 3, 2, count, 3, 3), device=dev),
                 scale=torch.rand(count, generator=gen, device=dev) + 0.5,
                 out=torch.empty((count, n, n), device=dev),
                 work=torch.empty((ops.work_elems(0, n, count, 3),), dtype=torch.complex64,
                                  device=dev))
        classes.append(c)

    def run():
        for c in classes:
            ops.generate_groups(c["noise"], c["shn"], c["amp"], c["sh"], c["count"],
                                screen_scale=c["scale"], out=c["out"], work=c["work"])

    for c in classes:
        c["noise"].normal_(generator=gen)
        c["shn"].normal_(generator=gen)
    cap = Captured(run)
    for k in range(4):
        for c in classes:
            c["noise"].normal_(generator=gen)
            c["shn"].normal_(generator=gen)
        cap.replay()
        got = [c["out"].clone() for c in classes]
        torch.cuda.synchronize()
        for c in classes:
            c["out"].fill_(float("nan"))
        run()  # eager, same noise
        torch.cuda.synchronize()
        for g, c in zip(got, classes):
            assert torch.isfinite(g).all()
            assert torch.equal(g, c["out"]), (k, c["n"])


def test_folded_loop_reuses_its_tables_safely(hip):
    """FoldedFrames over 40 calls (the ring of 32 pinned displacement tables
    wraps, and the host waits on an event once per 16 calls): every call's
    frame equals the eager call bit for bit, so no table was rewritten while
    a call could still read it."""
    import torch

    import pyturb_amd as pa
    from pyturb_amd.graph import FoldedFrames

    print("folded loop ring", flush=True)
    atm = pa.Atmosphere([pa.Layer(*ly) for ly in config2_layer_tuples()], r0=0.15, n=512, seed=4,
                        device="mi355x")
    loop = FoldedFrames(atm, 1)
    rng = np.random.default_rng(21)
    times = np.sort(rng.uniform(0.0, 0.3, 40))
    got = []
    for t in times:  # no synchronisation between calls
        got.append(loop([t]).clone())
    for k, t in enumerate(times):
        want = atm.opd_batch([t])
        assert torch.equal(got[k], want), k


# (five frames: two complete pairs, so random displacements fail the
# uniform-step check; with one pair it passes trivially)
@pytest.mark.parametrize("n,L,F,P", [(128, 5, 5, 4), (1024, 3, 5, 4)])
@pytest.mark.parametrize("order", [0, 1])
def test_folded_frames_equal_unfolded(hip, order, n, L, F, P):
    """spectral_fold once, then spectral_frames(folded=True) for two calls
    with different displacements: bit-identical to the unfolded call, for
    both transform orders.  Random displacements keep the unfolded call on
    the plain fold; with equally spaced displacements the unfolded call folds
    the uniform step in and the two agree to 1e-6."""
    import torch

    from pyturb_amd import ops, tune

    print(f"folded frames order {order} n {n}", flush=True)
    rng = np.random.default_rng(40 + order)
    dev = torch.device("cuda", 0)
    spec = (rng.standard_normal((L, n, n)) + 1j * rng.standard_normal((L, n, n))) / n
    spectra = torch.tensor(spec.astype(np.complex64), device=dev)
    cf = torch.tensor(rng.standard_normal((P, L, 3, 3)) * 0.1 + 0j, dtype=torch.complex128,
                      device=dev)
    work = torch.empty((ops.frames_work_elems(n, F, L, P),), dtype=torch.complex64, device=dev)
    try:
        tune.set_variant("frames_order", n, order)
        ops.spectral_fold(spectra, F, work)
        for k in range(2):
            disp = torch.tensor(rng.uniform(-2.0, 2.0, (F, L, 2)), dtype=torch.float64, device=dev)
            got = ops.spectral_frames(spectra, disp, cf, 0.05, work=work, folded=True).clone()
            ref = ops.spectral_frames(spectra, disp, cf, 0.05)
            assert torch.equal(got, ref), k
        ramp = torch.arange(F, dtype=torch.float64, device=dev)[:, None, None]
        disp = disp[:1] + 0.013 * ramp * torch.tensor(rng.uniform(-1.0, 1.0, (1, L, 2)), device=dev)
        got = ops.spectral_frames(spectra, disp, cf, 0.05, work=work, folded=True).clone()
        ref = ops.spectral_frames(spectra, disp, cf, 0.05)
        err = ((got - ref).square().mean() / ref.square().mean()).sqrt().item()
        assert err < 1e-6, err
    finally:
        tune.reset()


@pytest.mark.parametrize("order", [0, 1])
def test_uniform_step_decision_boundary(hip, order):
    """The uniform-step path is taken when every frame pair's fixed-point step
    matches pair 0's to within one unit of 2^-32 cycles per axis, and not
    otherwise.  With n = 64 and a 1/64 m pixel the period is 1 m, so
    displacements of m / 2^32 metres are exact phase units.  The plain path
    is recognised by being bit-identical to the folded call (which always
    takes it); the uniform path differs from it by rounding, within 1e-6."""
    import torch

    from pyturb_amd import ops, tune

    print(f"uniform-step decision order {order}", flush=True)
    n, L, F = 64, 2, 4
    rng = np.random.default_rng(70 + order)
    dev = torch.device("cuda", 0)
    spec = (rng.standard_normal((L, n, n)) + 1j * rng.standard_normal((L, n, n))) / n
    spectra = torch.tensor(spec.astype(np.complex64), device=dev)
    base = rng.integers(0, 1 << 31, (L, 2)).astype(np.float64)
    step = rng.integers(1 << 20, 1 << 24, (L, 2)).astype(np.float64)
    work = torch.empty((ops.frames_work_elems(n, F, L, 0),), dtype=torch.complex64, device=dev)
    try:
        tune.set_variant("frames_order", n, order)
        ops.spectral_fold(spectra, F, work)
        for extra, uniform in ((0, True), (1, True), (-1, True), (2, False), (-2, False)):
            units = base[None] + np.arange(F)[:, None, None] * step[None]
            units[3, 0, 0] += extra  # the second frame of pair 1
            disp = torch.tensor(units / 2.0**32, dtype=torch.float64, device=dev)
            got = ops.spectral_frames(spectra, disp, None, 1.0 / n)
            plain = ops.spectral_frames(spectra, disp, None, 1.0 / n, work=work, folded=True)
            if uniform:
                assert not torch.equal(got, plain), extra
                err = ((got - plain).square().mean() / plain.square().mean()).sqrt().item()
                assert err < 1e-6, (extra, err)
            else:
                assert torch.equal(got, plain), extra
    finally:
        tune.reset()
