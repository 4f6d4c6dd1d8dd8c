"""The one-launch extrusion step (pt_extrude_step): the innovation terms and
every recurrence round of a set_time call in one launch, with a
bounded grid barrier between rounds.  It replaces the innovation GEMM plus
one launch per round (src/pyturb/extrude.py:424-430 for every layer) and
must give the same rows bit for bit:

* on hand-made plans with ragged row counts per layer (0 to 9 new rows),
  rows that wrap around the ring's end, with and without far points and with
  one or two stencil rows;
* through an ExtrudedAtmosphere run (both paths from the same seeds);
* the barrier's failure path, forced from the host: with the abort word
  set before the launch every workgroup leaves at its first barrier, the
  status word is set and the host raises; after the reset a clean call is
  exact again.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(W, m, F, L, cap, seed):
    import torch

    from pyturb_amd.infinite import far_pitch, far_rows_back, row_pitch

    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    ld = row_pitch(W)
    ldm = m * ld + far_pitch(F)
    A = 0.7 * torch.randn((W, ldm), generator=g, device=dev) / np.sqrt(m * W + F)
    for s in range(m):
        A[:, s * ld + W:(s + 1) * ld] = 0
    A[:, m * ld + F:] = 0
    B = torch.randn((W, ld), generator=g, device=dev) / np.sqrt(W)
    B[:, W:] = 0
    buf = torch.randn((L, cap, ld), generator=g, device=dev)
    buf[..., W:] = 0
    far = far_rows_back(W, m, F) if F else ()
    return dev, g, A, B, buf, far


def _plans(counts, starts, cap, m, far):
    plans = []
    for k, t0 in zip(counts, starts):
        rows = []
        for i in range(k):
            tg = (t0 + i) % cap
            z = tuple((tg - m + s) % cap for s in range(m))
            rows.append((tg, z, tuple((tg - 1 - d) % cap for d in far)) if far else (tg, z))
        plans.append(rows)
    return plans


@pytest.mark.parametrize("W,m,F", [(100, 2, 0), (300, 2, 32), (1000, 2, 16), (257, 1, 8),
                                   (1100, 2, 32)])
def test_step_matches_round_path(hip, W, m, F):
    import torch

    from pyturb_amd.infinite import StepState, extrude_rounds

    L, cap = 5, 160
    print(f"step vs rounds W={W} m={m} F={F}", flush=True)
    dev, g, A, B, buf, far = _setup(W, m, F, L, cap, 11 + W)
    counts = [4, 0, 9, 1, 7]
    starts = [20, 3, cap - 4, 77, cap - 1]  # layers 2 and 4 wrap past the ring's end
    plans = _plans(counts, starts, cap, m, far)
    betas = [torch.randn((k, W), generator=g, device=dev) for k in counts]
    scales = [0.5, 1.0, 1.7, 0.9, 2.2]
    ref, got = buf.clone(), buf.clone()
    extrude_rounds(A, B, ref, cap, W, m, plans, betas, scales, far_points=F)
    step = StepState(dev, far)
    extrude_rounds(A, B, got, cap, W, m, plans, betas, scales, far_points=F, step=step)
    torch.cuda.synchronize()
    assert step.launches == 1 and int(step.status[0]) == 0
    assert not torch.equal(ref, buf)
    diff = (ref - got).abs().max().item()
    assert torch.equal(ref, got), diff
    # a second call on the same barrier words (its generation carries over)
    plans2 = _plans([3, 2, 0, 5, 1], [24, 3, 5, 78, 6], cap, m, far)
    betas2 = [torch.randn((len(p), W), generator=g, device=dev) for p in plans2]
    extrude_rounds(A, B, ref, cap, W, m, plans2, betas2, scales, far_points=F)
    extrude_rounds(A, B, got, cap, W, m, plans2, betas2, scales, far_points=F, step=step)
    torch.cuda.synchronize()
    assert step.launches == 2 and torch.equal(ref, got)
    # the barrier count carried over both calls; no abort
    assert int(step.sync[0]) == (max(counts) - 1) + (5 - 1) and int(step.sync[1]) == 0


def test_step_declines_what_it_does_not_cover(hip):
    import torch

    from pyturb_amd.infinite import StepState, extrude_rounds

    W, m, F, L, cap = 120, 2, 0, 2, 400
    dev, g, A, B, buf, far = _setup(W, m, F, L, cap, 5)
    plans = _plans([200, 100], [0, 150], cap, m, far)  # 300 new rows: beyond the one-launch step
    betas = [torch.randn((len(p), W), generator=g, device=dev) for p in plans]
    ref, got = buf.clone(), buf.clone()
    extrude_rounds(A, B, ref, cap, W, m, plans, betas, [1.0, 1.0])
    step = StepState(dev, far)
    extrude_rounds(A, B, got, cap, W, m, plans, betas, [1.0, 1.0], step=step)
    torch.cuda.synchronize()
    assert step.launches == 0 and torch.equal(ref, got)


@pytest.mark.parametrize("F", [0, 32])
def test_extruder_step_path_equals_round_path(hip, F):
    import torch

    from pyturb_amd.extrude import ExtrudedAtmosphere

    kw = dict(n=64, pixel_scale=0.125, layer_r0=[0.15, 0.3, 0.5], layer_L0=[25.0] * 3,
              layer_wind=[(10.0, 0.0), (-4.0, 7.0), (2.0, -15.0)],
              layer_altitude_los=[0.0, 4000.0, 10000.0], field_of_view_pix=2.0,
              seeds=[7, 8, 9], far_points=F)
    print(f"extruder step path vs rounds, far {F}", flush=True)
    a = ExtrudedAtmosphere(**kw)
    b = ExtrudedAtmosphere(**kw)
    b._step.enabled = False
    thx = np.array([0.0, 2.0]) * 0.125 / 10000.0
    thy = np.array([0.0, -1.0]) * 0.125 / 10000.0
    for k in range(1, 121):
        t = k * 0.011
        a.set_time(t)
        b.set_time(t)
        if k % 30 == 0:
            assert torch.equal(a.integrate_dirs(thx, thy), b.integrate_dirs(thx, thy))
    torch.cuda.synchronize()
    assert a._step.launches >= 100 and b._step.launches == 0
    assert torch.equal(a._buf, b._buf)


def test_barrier_timeout_is_reported_and_recoverable(hip):
    import torch

    from pyturb_amd import _hip
    from pyturb_amd.infinite import StepState, extrude_rounds

    W, m, F, L, cap = 400, 2, 32, 2, 200
    dev, g, A, B, buf, far = _setup(W, m, F, L, cap, 99)
    plans = _plans([5, 3], [10, 60], cap, m, far)
    betas = [torch.randn((len(p), W), generator=g, device=dev) for p in plans]
    step = StepState(dev, far)
    step.spin_limit = 64
    # the abort word, set by the host: every workgroup leaves at its first barrier
    step.sync[1] = 1
    got = buf.clone()
    print("barrier abort forced from the host (poll limit 64)", flush=True)
    extrude_rounds(A, B, got, cap, W, m, plans, betas, [1.0, 1.0], far_points=F, step=step)
    torch.cuda.synchronize()
    assert step.launches == 1 and int(step.status[0]) == 1
    with pytest.raises(_hip.HipError, match="timed out"):
        step.check()
    assert int(step.status[0]) == 0 and int(step.sync.abs().sum()) == 0
    # after the reset a clean call is exact
    step.spin_limit = StepState.SPIN_LIMIT
    ref, got = buf.clone(), buf.clone()
    extrude_rounds(A, B, ref, cap, W, m, plans, betas, [1.0, 1.0], far_points=F)
    extrude_rounds(A, B, got, cap, W, m, plans, betas, [1.0, 1.0], far_points=F, step=step)
    torch.cuda.synchronize()
    assert int(step.status[0]) == 0 and torch.equal(ref, got)


def test_step_argument_errors(hip):
    import torch

    from pyturb_amd import _hip

    dev = torch.device("cuda")
    W, ld, cap = 64, 64, 32
    A = torch.zeros((W, 2 * ld), device=dev)
    B = torch.zeros((W, ld), device=dev)
    buf = torch.zeros((1, cap, ld), device=dev)
    beta = torch.zeros((2, W), device=dev)
    sync = torch.zeros((256,), dtype=torch.int64, device=dev)
    status = torch.zeros((1,), dtype=torch.int32, pin_memory=True)
    one = lambda v, T=ctypes.c_int: (T * 1)(v)  # noqa: E731
    args = dict(first=one(5), rows=one(2), beta=(ctypes.c_void_p * 1)(beta.data_ptr()),
                pitch=one(W, ctypes.c_longlong), sc=one(1.0, ctypes.c_float))

    def call(m=2, F=0, far=None, layers=1, rows=None, first=None):
        return hip.pt_extrude_step(_hip.ptr(A), _hip.ptr(B), W, ld, m, F, far, _hip.ptr(buf), cap,
                                   layers, first or args["first"], rows or args["rows"], args["beta"],
                                   args["pitch"], args["sc"], _hip.ptr(sync), status.data_ptr(), 1000,
                                   _hip.stream_ptr())

    assert call(F=8, far=None) == _hip.PT_ERR_ARG          # far points without far rows
    assert call(F=12, far=one(3)) == _hip.PT_ERR_ARG       # not a multiple of 8
    assert call(first=one(cap)) == _hip.PT_ERR_ARG         # first row outside the ring
    assert call(rows=one(300)) == _hip.PT_ERR_UNSUPPORTED  # more than 256 new rows
    assert call(m=3) == _hip.PT_ERR_UNSUPPORTED            # three stencil rows
    assert call(layers=17) == _hip.PT_ERR_UNSUPPORTED
    assert call(rows=one(0)) == 0                          # nothing to do
    torch.cuda.synchronize()
    assert int(status[0]) == 0
