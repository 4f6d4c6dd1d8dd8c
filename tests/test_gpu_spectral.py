"""GPU parity: FFT passes, fused generate and spectral frames vs the oracle.

Tolerances: FP32 device arithmetic against the FP64 oracle on identical
noise, relative RMS <= 1e-5 (the stated bar), max-abs checks where noted.
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


@pytest.mark.parametrize("n", [32, 64, 128, 256, 512, 1024, 2048])
@pytest.mark.parametrize("sign", [1, -1])
def test_fft2_pow2_matches_numpy(hip, n, sign):
    import torch
    from pyturb_amd import ops

    print(f"fft2 n={n} sign={sign}", flush=True)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    got = _np(ops.fft2(torch.from_numpy(x.astype(np.complex64)).cuda(), sign))
    ref = np.fft.ifft2(x) * n * n if sign > 0 else np.fft.fft2(x)
    err = np.sqrt(np.mean(np.abs(got - ref) ** 2) / np.mean(np.abs(ref) ** 2))
    assert err < 2e-6, err


def test_fft2_inplace(hip):
    import torch
    from pyturb_amd import ops

    rng = np.random.default_rng(1)
    x = (rng.standard_normal((3, 256, 256)) + 1j * rng.standard_normal((3, 256, 256)))
    t = torch.from_numpy(x.astype(np.complex64)).cuda()
    ops.fft2(t, 1, out=t)
    ref = np.fft.ifft2(x) * 256 * 256
    got = _np(t)
    assert np.sqrt(np.mean(np.abs(got - ref) ** 2) / np.mean(np.abs(ref) ** 2)) < 2e-6


def _gen_case(n, dx, r0, L0, P, count, seed):
    filt = O.build_filters(n, dx, r0, L0, subharmonics=P)
    rng = np.random.default_rng(seed)
    main, sh = O.draw_generate_noise(rng, n, count, P)
    return filt, main, sh


def _run_generate(filt, main, sh, count, **kw):
    import torch
    from pyturb_amd import ops

    P = filt["sh_amp"].shape[0]
    dev = torch.device("cuda")
    noise = torch.from_numpy(main.astype(np.float32)).to(dev)
    shn = torch.from_numpy(sh.astype(np.float32)).to(dev) if P else None
    amp = torch.from_numpy(filt["amplitude"].astype(np.float32)).to(dev)
    sha = torch.from_numpy(filt["sh_amp"]).to(dev) if P else None
    return ops.generate(noise, shn, amp, sha, count, **kw)


def test_generate_config1_matches_oracle(hip):
    """n=256, dx=1/32, r0=0.15, L0=25, 3 levels, 64 screens, seed=1 noise."""
    print("generate config1", flush=True)
    filt, main, sh = _gen_case(256, 0.03125, 0.15, 25.0, 3, 64, 1)
    ref = O.generate(filt, main.astype(np.float32).astype(np.float64),
                     sh.astype(np.float32).astype(np.float64), 64)
    got = _np(_run_generate(filt, main, sh, 64))
    assert O.rel_rms(got, ref) < REL
    # and against the reference's own float64 realisation (golden)
    g = load_golden("generate_n256_seed1.npz")
    assert O.rel_rms(got[:2], g["screens01"]) < REL


@pytest.mark.parametrize("n,count,P", [(64, 5, 0), (64, 7, 8), (128, 1, 2), (512, 4, 8),
                                       (1024, 2, 8), (32, 3, 1), (2048, 2, 8)])
def test_generate_shapes_levels_odd_counts(hip, n, count, P):
    print(f"generate n={n} count={count} P={P}", flush=True)
    filt, main, sh = _gen_case(n, 8.0 / n, 0.12, 30.0, P, count, 7)
    ref = O.generate(filt, main.astype(np.float32).astype(np.float64),
                     sh.astype(np.float32).astype(np.float64) if P else sh, count)
    got = _np(_run_generate(filt, main, sh, count))
    assert got.shape == (count, n, n)
    assert O.rel_rms(got, ref) < REL


def test_generate_crop_scale_accumulate(hip):
    import torch

    n, count, P = 128, 6, 4
    filt, main, sh = _gen_case(n, 0.0625, 0.1, 20.0, P, count, 3)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    ref = O.generate(filt, f32(main), f32(sh), count)[:, 10:110, 10:110]
    ss = np.linspace(0.5, 2.0, count).astype(np.float32)
    out = _run_generate(filt, main, sh, count, crop_offset=10, n_out=100, scale=0.25,
                        screen_scale=torch.from_numpy(ss).cuda())
    got = _np(out)
    want = ref * 0.25 * ss[:, None, None]
    assert O.rel_rms(got, want) < REL
    _run_generate(filt, main, sh, count, crop_offset=10, n_out=100, scale=0.25,
                  screen_scale=torch.from_numpy(ss).cuda(), out=out, accumulate=True)
    assert O.rel_rms(_np(out), 2 * want) < REL


def _frames_case(n, times, P=8, seed=2):
    atm = O.OracleAtmosphere(config2_layer_tuples(), r0=0.15, diameter=8.0, n=n, L0=25.0,
                             subharmonics=P, dtype="float64", seed=seed)
    return atm


def _run_frames(atm, disp, scale=1.0):
    import torch
    from pyturb_amd import ops

    dev = torch.device("cuda")
    spectra = torch.from_numpy(atm.spectra.astype(np.complex64)).to(dev)
    P = atm.sh_coeffs.shape[0]
    coeffs = torch.from_numpy(np.ascontiguousarray(atm.sh_coeffs)).to(dev) if P else None
    d = torch.from_numpy(np.ascontiguousarray(disp, dtype=np.float64)).to(dev)
    return _np(ops.spectral_frames(spectra, d, coeffs, atm.dx, crop_offset=atm.margin,
                                   n_out=atm.n, scale=scale))


@pytest.mark.parametrize("n", [64, 512])
def test_spectral_frames_config2_match_oracle(hip, n):
    """9-layer profile, winds from seed 2, up to 2 s of travel (wrapped)."""
    print(f"frames n={n}", flush=True)
    atm = _frames_case(n, None)
    times = [0.0, 0.001, 0.0137, 0.5, 1.999]
    disp = np.stack([np.stack(atm.displacements(t), axis=-1) for t in times])
    got = _run_frames(atm, disp)
    for k, t in enumerate(times):
        ref = atm.phase(t)
        assert O.rel_rms(got[k], ref) < REL, (t, O.rel_rms(got[k], ref))


@pytest.mark.parametrize("n,frames", [(32, 3), (64, 4), (256, 5)])
def test_spectral_frames_white_spectra_all_modes(hip, n, frames):
    """Flat random layer spectra put full weight on every mode, including the
    Nyquist row and column, where the paired (Hermitian) transform needs its
    exact correction; odd frame counts leave a single frame in the last pair."""
    import torch
    from pyturb_amd import ops

    print(f"white frames n={n} frames={frames}", flush=True)
    rng = np.random.default_rng(n + frames)
    L, dx = 3, 0.05
    spectra = (rng.standard_normal((L, n, n)) + 1j * rng.standard_normal((L, n, n))).astype(np.complex64)
    disp = rng.uniform(-3.0, 3.0, (frames, L, 2))
    disp[0] = 0.0
    f = np.fft.fftfreq(n, d=dx)
    sp64 = spectra.astype(np.complex128)
    dev = torch.device("cuda")
    got = _np(ops.spectral_frames(torch.from_numpy(spectra).to(dev),
                                  torch.from_numpy(disp).to(dev), None, dx))
    for t in range(frames):
        ref = O.integrate(sp64, np.zeros((0, L, 3, 3), complex), f, None, None,
                          disp[t, :, 0], disp[t, :, 1])
        assert O.rel_rms(got[t], ref) < REL, (t, O.rel_rms(got[t], ref))


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("n,frames", [(64, 9), (256, 8)])
def test_uniform_step_frames_match_oracle(hip, order, n, frames):
    """Equally spaced times move every frame pair by one step per layer: the
    fold then carries the step's phasor and pass A forms one phasor product
    per layer and element.  The same times with one frame moved off the grid
    take the two-product path.  Both against the oracle frame by frame, on
    flat spectra (full weight on the Nyquist lines), for both transform
    orders, and the two paths agree to 1e-6 on the frames they share."""
    import torch
    from pyturb_amd import ops, tune

    print(f"uniform step n={n} frames={frames} order={order}", flush=True)
    rng = np.random.default_rng(7 * n + frames)
    L, dx, dt = 3, 0.05, 1e-3
    spectra = (rng.standard_normal((L, n, n)) + 1j * rng.standard_normal((L, n, n))).astype(np.complex64)
    wind = rng.uniform(-40.0, 40.0, (L, 2))
    base = rng.uniform(-5.0, 5.0, (L, 2))
    times = 3.7 + dt * np.arange(frames)
    disp = base[None] + times[:, None, None] * wind[None]
    off = disp.copy()
    off[frames - 3] += 0.37 * dx  # breaks the uniform step of one pair
    f = np.fft.fftfreq(n, d=dx)
    sp64 = spectra.astype(np.complex128)
    dev = torch.device("cuda")
    sp = torch.from_numpy(spectra).to(dev)
    try:
        tune.set_variant("frames_order", n, order)
        got = {k: _np(ops.spectral_frames(sp, torch.from_numpy(d).to(dev), None, dx))
               for k, d in (("uniform", disp), ("off", off))}
    finally:
        tune.reset()
    for k, d in (("uniform", disp), ("off", off)):
        for t in range(frames):
            ref = O.integrate(sp64, np.zeros((0, L, 3, 3), complex), f, None, None, d[t, :, 0], d[t, :, 1])
            assert O.rel_rms(got[k][t], ref) < REL, (k, t, O.rel_rms(got[k][t], ref))
    for t in range(frames):
        if t != frames - 3:
            assert O.rel_rms(got["uniform"][t], got["off"][t]) < 1e-6, t


@pytest.mark.parametrize("n", [64, 256, 512, 1024])
def test_few_frame_calls_equal_batched_calls(hip, n):
    """Frames do not depend on how many other frames share their call: a call
    of 1 to 4 frame pairs gives, bit for bit, the same frames as a larger
    call (column-first order, 9 layers), with equally spaced times (uniform
    step) and with jittered ones, 2 to 8 frames."""
    import torch
    from pyturb_amd import ops, tune

    print(f"few-frame calls n={n}", flush=True)
    rng = np.random.default_rng(n + 5)
    L, dx = 9, 8.0 / n
    spectra = (rng.standard_normal((L, n, n)) + 1j * rng.standard_normal((L, n, n))).astype(np.complex64)
    wind = rng.uniform(-30.0, 30.0, (L, 2))
    times = 0.4 + 1e-3 * np.arange(12)
    dev = torch.device("cuda")
    sp = torch.from_numpy(spectra).to(dev)
    try:
        tune.set_variant("frames_order", n, 0)
        for jitter in (0.0, 0.3):
            t = times + jitter * 1e-3 * rng.uniform(-1.0, 1.0, times.size)
            disp = torch.from_numpy(t[:, None, None] * wind[None]).to(dev)
            big = ops.spectral_frames(sp, disp, None, dx)  # 6 pairs: the batched pass A
            # 1..4 pairs: the few-pair pass A.  The two calls must take the same
            # path: a single pair always passes the uniform-step check, so it is
            # compared on equally spaced times only; and a frame alone in its
            # pair shares its transform with no partner (unless the uniform step
            # supplies one), so jittered calls compare their complete pairs
            for F in ((7, 8) if jitter else (2, 7, 8)):
                few = ops.spectral_frames(sp, disp[:F].contiguous(), None, dx)
                m = F if not jitter else F - F % 2
                assert torch.equal(few[:m], big[:m]), (jitter, F)
    finally:
        tune.reset()


@pytest.mark.parametrize("order", [0, 1])
def test_spectral_frames_many_frames_match_oracle(hip, order):
    """20 equally spaced frames with 8 subharmonic levels: a call of 16 or
    more frames reads the subharmonic basis from its table instead of
    evaluating it per frame, and the frame pairs share one uniform step."""
    from pyturb_amd import tune

    print(f"20 frames order={order}", flush=True)
    atm = _frames_case(64, None)
    times = 0.25 + 1e-3 * np.arange(20)
    disp = np.stack([np.stack(atm.displacements(t), axis=-1) for t in times])
    try:
        tune.set_variant("frames_order", atm.ns, order)
        got = _run_frames(atm, disp)
    finally:
        tune.reset()
    for k, t in enumerate(times):
        assert O.rel_rms(got[k], atm.phase(t)) < REL, (t, O.rel_rms(got[k], atm.phase(t)))


def test_spectral_frame_matches_reference_golden(hip):
    g = load_golden("atmosphere_n64_seed2.npz")
    atm = _frames_case(64, None)
    disp = np.stack(atm.displacements(0.02), axis=-1)[None]
    got = _run_frames(atm, disp, scale=500e-9 / (2 * np.pi))
    assert O.rel_rms(got[0], g["frame_nofov_t002"]) < REL


def test_spectral_integer_shift_is_roll(hip):
    atm = _frames_case(128, None, P=0)
    dx = atm.dx
    disp = np.zeros((2, 9, 2))
    disp[1, :, 0] = 5 * dx
    disp[1, :, 1] = -3 * dx
    got = _run_frames(atm, disp)
    np.testing.assert_allclose(got[1], np.roll(np.roll(got[0], -5, 0), 3, 1),
                               atol=2e-5 * np.abs(got[0]).max())


def test_phase_reduction_at_large_travel(hip):
    """Exact fixed-point argument reduction: a shift by whole periods is exact."""
    atm = _frames_case(256, None, P=0)
    period = atm.ns * atm.dx
    base = np.stack(atm.displacements(0.01), axis=-1)
    disp = np.stack([base, base + 1000 * period])
    got = _run_frames(atm, disp)
    assert O.rel_rms(got[1], got[0]) < 1e-6


# ------------------------------------------------------------------ drop-in API
def test_atmosphere_api_matches_reference_golden(hip):
    import pyturb_amd as pa

    g = load_golden("atmosphere_n64_seed2.npz")
    layers = [pa.Layer(*ly) for ly in config2_layer_tuples()]
    atm = pa.Atmosphere(layers, r0=0.15, diameter=8.0, n=64, L0=25.0, dtype="float64",
                        seed=2)
    got = _np(atm.opd(0.02))
    assert got.dtype == np.float64
    assert O.rel_rms(got, g["frame_nofov_t002"]) < REL
    ts = [t for t, _ in atm.frames(dt=1e-3, steps=3)]
    assert ts == [0.0, 1e-3, 2e-3]


def test_sample_api_matches_reference_golden(hip):
    import pyturb_amd as pa

    g = load_golden("sample_n64_seed4.npz")
    layers = [pa.Layer(0, 0.4, 10.0, 0.0, 25.0), pa.Layer(1000, 0.2, 10.0, 0.0, 50.0),
              pa.Layer(5000, 0.3, 10.0, 0.0, 25.0), pa.Layer(10000, 0.1, 10.0, 0.0, 50.0)]
    atm = pa.Atmosphere(layers, r0=0.15, diameter=8.0, n=64, dtype="float64", seed=4)
    got = _np(atm.sample(4))
    assert O.rel_rms(got, g["sample"]) < REL


@pytest.mark.parametrize("n,count", [(128, 7), (96, 4)])
def test_generate_groups_equals_sum_of_groups(hip, n, count):
    """pt_generate_groups (one transform for all outer-scale groups) against
    the oracle's per-group screens summed; n = 96 runs the general-size path."""
    import torch
    from pyturb_amd import ops

    print(f"generate groups n={n} count={count}", flush=True)
    rng = np.random.default_rng(n)
    dx, P = 8.0 / n, 3
    L0s = [20.0, 35.0, 60.0]
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    n_fft = (count + 1) // 2
    noise = rng.standard_normal((3, 2, n_fft, n, n)).astype(np.float32)
    shn = rng.standard_normal((3, P, 2, count, 3, 3)).astype(np.float32)
    filts = [O.build_filters(n, dx, 0.12, L0, subharmonics=P) for L0 in L0s]
    ref = sum(O.generate(f, f32(noise[g]), f32(shn[g]), count) for g, f in enumerate(filts))
    ss = np.linspace(0.7, 1.3, count).astype(np.float32)
    off, n_out, scale = 5, n - 10, 0.5
    want = ref[:, off:off + n_out, off:off + n_out] * scale * ss[:, None, None]
    dev = torch.device("cuda")
    amp = torch.from_numpy(np.stack([f["amplitude"] for f in filts]).astype(np.float32)).to(dev)
    sha = torch.from_numpy(np.stack([f["sh_amp"] for f in filts])).to(dev)
    out = ops.generate_groups(torch.from_numpy(noise).to(dev), torch.from_numpy(shn).to(dev), amp,
                              sha, count, crop_offset=off, n_out=n_out, scale=scale,
                              screen_scale=torch.from_numpy(ss).to(dev))
    assert O.rel_rms(_np(out), want) < REL
    ops.generate_groups(torch.from_numpy(noise).to(dev), torch.from_numpy(shn).to(dev), amp, sha,
                        count, crop_offset=off, n_out=n_out, scale=scale,
                        screen_scale=torch.from_numpy(ss).to(dev), out=out, accumulate=True)
    assert O.rel_rms(_np(out), 2 * want) < REL


def test_phase_screen_api_matches_reference_golden(hip):
    import pyturb_amd as pa

    g = load_golden("generate_n256_seed1.npz")
    gen = pa.PhaseScreen(n=256, pixel_scale=0.03125, r0=0.15, L0=25.0, subharmonics=3,
                         seed=1, dtype="float64")
    got = _np(gen.generate(64))
    assert O.rel_rms(got[:2], g["screens01"]) < REL
    np.testing.assert_allclose(got.sum(axis=(1, 2)), g["sums"],
                               atol=REL * np.abs(got).sum() / 64)


# ------------------------------------------------------------------ general-size path
@pytest.mark.parametrize("n,count,P", [(96, 3, 2), (100, 4, 8), (64, 4, 14)])
def test_generate_general_sizes_and_deep_levels(hip, n, count, P):
    """Non-power-of-two grids (direct-DFT path) and > 12 levels."""
    print(f"generate general n={n} P={P}", flush=True)
    filt, main, sh = _gen_case(n, 8.0 / n, 0.12, 30.0, P, count, 11)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    ref = O.generate(filt, f32(main), f32(sh), count)
    got = _np(_run_generate(filt, main, sh, count))
    assert O.rel_rms(got, ref) < REL


def test_fov_atmosphere_frames_and_directions_match_reference_golden(hip):
    """30" field of view: 112^2 oversized screens, crop, three directions."""
    import pyturb_amd as pa

    g = load_golden("atmosphere_n64_seed2.npz")
    layers = [pa.Layer(*ly) for ly in config2_layer_tuples()]
    atm = pa.Atmosphere(layers, r0=0.15, diameter=8.0, n=64, L0=25.0, field_of_view=30.0,
                        dtype="float64", seed=2)
    assert atm.n_screen == int(g["n_screen"])
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for t, want in zip(g["times"], g["frames"]):
            assert O.rel_rms(_np(atm.opd(float(t))), want) < REL
    got = _np(atm.opd(0.01, directions=[tuple(d) for d in g["dirs"]]))
    assert got.shape == (3, 64, 64)
    for d in range(3):
        assert O.rel_rms(got[d], g["dframes"][d]) < REL
    batch = _np(atm.opd_batch([0.01, 0.01], directions=[tuple(d) for d in g["dirs"]]))
    np.testing.assert_array_equal(batch[1], got)


def test_fft2_general_size(hip):
    import torch
    from pyturb_amd import ops

    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 72, 72)) + 1j * rng.standard_normal((2, 72, 72))
    got = _np(ops.fft2(torch.from_numpy(x.astype(np.complex64)).cuda(), 1))
    ref = np.fft.ifft2(x) * 72 * 72
    assert np.sqrt(np.mean(np.abs(got - ref) ** 2) / np.mean(np.abs(ref) ** 2)) < 2e-6
