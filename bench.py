"""Benchmark of the MI355X phase-screen hot path.

Headline workload (the 9-layer spectral frozen-flow atmosphere): 512^2
pupil, D = 8 m, r0 = 0.15 m, L0 = 25 m, heights [0, .5, 1, 2, 4, 8, 12, 16,
20] km, Cn2 fractions [.5, .1, .08, .07, .06, .06, .05, .05, .03], winds
U(5, 35) m/s and directions U(0, 2 pi) from seed 2, 8 subharmonic levels,
dt = 1 ms, 2000 frames, FP32.  One step = all 2000 frames (two fused HIP
launches); inputs (layer spectra, displacement table) are resident in HBM.

Also reported on the same line:
  * the Monte-Carlo ensemble: 2048 x 256^2 + 1536 x 512^2 + 512 x 1024^2
    screens, two outer-scale groups (L0 = 25 m / 50 m, half the Cn2 each),
    per-screen r0 ~ U(0.08, 0.25) m (seed 4), white noise resident in HBM,
    through per-class pt_generate_groups calls and through the one-call
    pt_generate_ensemble (the faster is mc_screens_per_s, both are listed),
    with per-pass times and HBM fractions in roofline.passes;
  * the 1024^2 9-layer extrusion with 7 readout directions (one extrusion
    step and one tiled readout per set_time), with and without the
    32-point far stencil;
  * boiled frames (tau_boil = 0.2 s) and LGS-cone frames (90 km beacon);
  * the plain GPU form of both (torch ops, one kernel chain per layer, FP32
    everywhere -- the reference's array-module algorithm on ROCm);
  * roofline of the dominant kernel (pass A of the frame engine, in the
    transform order the autotuner picked) from HIP events on the launch
    stream;
  * the CPU baseline: the FP64 NumPy oracle timed on this host (1 thread,
    and every available core in cpu_baseline.rows);
  * a closed-loop call pattern, one frame per call, eager against a
    HIP-graph replay (pyturb_amd.graph.CapturedFrames), and the ensemble
    replayed from a HIP graph;
  * the launch shapes the autotuner picked (pyturb_amd.tune) for the
    row and column passes at each size;
  * 16 atmospheres called in turn, 125 frames each, so that their 302 MB of
    layer stacks stream from HBM instead of staying in the MALL.

``python bench.py [--gpus 1] [--steps K] [--warmup W]`` prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "9-layer 512\u00b2 OPD frames/s + Monte-Carlo screens/s, fused vs plain, %HBM peak"
HBM_PEAK_GBS = 8000.0
# wave64 VALU instruction issue peak, non-packed FP32: 1024 SIMDs at 2.4 GHz,
# one wave64 instruction per SIMD every 4 cycles (G instructions/s)
VALU_PEAK_GIPS = 1024 * 2.4e9 / 4 / 1e9
# Per-launch counter figures (HBM bytes from rocprofv3 FETCH_SIZE / WRITE_SIZE
# scaled against a plain copy, wave64 VALU instructions from SQ_INSTS_VALU) of
# the frame engine and the Monte-Carlo passes; each summary states its runs.
FRAMES_PROFILE = os.path.join("profiles", "spectral_frames_512_9layer.md")
MC_PROFILE = os.path.join("profiles", "mc_ensemble_mixed_sizes.md")
N_FRAMES = 2000
DT = 1e-3
N = 512
HEIGHTS = [0.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0, 12000.0, 16000.0, 20000.0]
FRACS = [0.5, 0.1, 0.08, 0.07, 0.06, 0.06, 0.05, 0.05, 0.03]
MC_SIZES = [(256, 2048), (512, 1536), (1024, 512)]
MC_L0 = [25.0, 50.0]


def profile_layers():
    import pyturb_amd as pa

    rng = np.random.default_rng(2)
    speeds = rng.uniform(5.0, 35.0, 9)
    dirs = rng.uniform(0.0, 2.0 * np.pi, 9)
    return [pa.Layer(h, f, float(v), float(np.degrees(d)), 25.0)
            for h, f, v, d in zip(HEIGHTS, FRACS, speeds, dirs)]


def build_frames_workload(dev):
    import pyturb_amd as pa
    from pyturb_amd import _hip

    atm = pa.Atmosphere(profile_layers(), r0=0.15, diameter=8.0, n=N, L0=25.0, subharmonics=8,
                        seed=2, device="mi355x")
    times = np.arange(N_FRAMES) * DT
    dx = atm._vx[None, :] * times[:, None]
    dy = atm._vy[None, :] * times[:, None]
    disp = _hip.to_device(np.stack([dx, dy], axis=-1), np.float64, dev)
    return atm, times, disp


def run_frames(atm, disp, out, work, events=None, frames=N_FRAMES):
    from pyturb_amd import _hip

    lib = _hip.lib()
    st = _hip.stream_ptr()
    L = len(atm.layers)
    P = atm.subharmonics
    scale = atm.wavelength / (2.0 * np.pi)
    if events:
        events[0].record()
    rc = lib.pt_spectral_rows(_hip.ptr(atm._spectra), L, N, atm.pixel_scale, _hip.ptr(disp),
                              frames, _hip.ptr(work), st)
    _hip.check(rc, "pt_spectral_rows")
    if events:
        events[1].record()
    rc = lib.pt_spectral_cols(_hip.ptr(work), L, N, atm.pixel_scale, _hip.ptr(disp), frames,
                              _hip.ptr(atm._sh_coeffs), P, 0, N, float(scale), _hip.ptr(out), st)
    _hip.check(rc, "pt_spectral_cols")
    if events:
        events[2].record()


BOIL_TAU = 0.2  # s, every layer: the reference benchmark's boiling case (bench_suite.py:157)


def build_boil(dev):
    """The same 9-layer 512^2 atmosphere, boiling with tau_boil = 0.2 s."""
    import pyturb_amd as pa
    from pyturb_amd import _hip

    atm = pa.Atmosphere(profile_layers(), r0=0.15, diameter=8.0, n=N, L0=25.0, subharmonics=8,
                        seed=2, tau_boil=BOIL_TAU, device="mi355x")
    times = np.arange(N_FRAMES) * DT
    dx = atm._vx[None, :] * times[:, None]
    dy = atm._vy[None, :] * times[:, None]
    disp = _hip.to_device(np.stack([dx, dy], axis=-1), np.float64, dev)
    return atm, disp


def run_boil(atm, disp, out, work, step0):
    """2000 boiled frames (a boiling step after each) from the device stream."""
    from pyturb_amd import ops

    ops.boiled_frames(atm._spectra, atm._boil_amp, atm._boil_inv_tau, atm._boil_f_ref, disp, DT,
                      atm.pixel_scale, sh_coeffs=atm._sh_coeffs, sh_amp=atm._boil_sh_amp,
                      seed=atm._boil_seed, step0=step0, scale=atm.wavelength / (2.0 * np.pi),
                      out=out, work=work)


def plain_frame(atm, t, ctx):
    """The reference's per-frame algorithm as plain torch ops, per layer."""
    import torch

    spec, f, fp, basis, coeffs = ctx
    L = spec.shape[0]
    sx = torch.tensor(atm._vx * t, dtype=torch.float32, device=spec.device)
    sy = torch.tensor(atm._vy * t, dtype=torch.float32, device=spec.device)
    S = torch.zeros_like(spec[0])
    for l in range(L):
        phx = torch.exp((2j * np.pi) * sx[l] * f)
        phy = torch.exp((2j * np.pi) * sy[l] * f)
        S = S + spec[l] * phx[:, None] * phy[None, :]
    total = (torch.fft.ifft2(S) * (N * N)).real
    px = torch.exp((2j * np.pi) * sx[None, :, None] * fp[:, None, :])
    py = torch.exp((2j * np.pi) * sy[None, :, None] * fp[:, None, :])
    shifted = (coeffs * px[:, :, :, None] * py[:, :, None, :]).sum(dim=1)
    m = torch.matmul(shifted, basis).reshape(-1, N)
    low = torch.matmul(basis.reshape(-1, N).transpose(0, 1).contiguous(), m).real
    low = low - low.mean()
    return (total + low) * (atm.wavelength / (2.0 * np.pi))


def plain_frames_batched(atm, times, ctx):
    """The same plain algorithm for a batch of frames: one elementwise kernel
    chain per layer over the whole batch, one batched transform (FP32)."""
    import torch

    spec, f, fp, basis, coeffs = ctx
    L = spec.shape[0]
    tt = torch.tensor(np.asarray(times), dtype=torch.float64, device=spec.device)
    sx = (tt[:, None] * torch.tensor(atm._vx, device=spec.device)).float()  # (F, L)
    sy = (tt[:, None] * torch.tensor(atm._vy, device=spec.device)).float()
    S = None
    for l in range(L):
        phx = torch.exp((2j * np.pi) * sx[:, l, None] * f[None, :])
        phy = torch.exp((2j * np.pi) * sy[:, l, None] * f[None, :])
        term = spec[l][None] * phx[:, :, None] * phy[:, None, :]
        S = term if S is None else S + term
    total = (torch.fft.ifft2(S) * (N * N)).real
    px = torch.exp((2j * np.pi) * sx[:, None, :, None] * fp[None, :, None, :])  # (F, P, L, 3)
    py = torch.exp((2j * np.pi) * sy[:, None, :, None] * fp[None, :, None, :])
    shifted = (coeffs[None] * px[..., :, None] * py[..., None, :]).sum(dim=2)  # (F, P, 3, 3)
    m = torch.matmul(shifted, basis[None])  # (F, P, 3, n)
    low = torch.einsum("pax,fpay->fxy", basis, m).real
    low = low - low.mean(dim=(-2, -1), keepdim=True)
    return (total + low) * (atm.wavelength / (2.0 * np.pi))


def plain_context(atm):
    import torch

    dev = atm._spectra.device
    n = atm.n_screen
    f = torch.tensor(np.fft.fftfreq(n, d=atm.pixel_scale), dtype=torch.float32, device=dev)
    df = 1.0 / (n * atm.pixel_scale)
    P = atm.subharmonics
    fpn = np.array([-1.0, 0.0, 1.0])[None, :] * (df / 3.0 ** np.arange(1, P + 1))[:, None]
    x = (np.arange(n) - n / 2.0) * atm.pixel_scale
    basis = np.exp(2j * np.pi * fpn[:, :, None] * x[None, None, :])
    return (atm._spectra, f, torch.tensor(fpn, dtype=torch.float32, device=dev),
            torch.tensor(basis, dtype=torch.complex64, device=dev),
            atm._sh_coeffs.to(torch.complex64))


def build_mc(dev):
    """Mixed-size ensemble state: amplitudes, HBM-resident noise, scales."""
    import torch

    from pyturb_amd import _hip
    from pyturb_amd.spectrum import build_filters

    rng = np.random.default_rng(4)
    total = sum(c for _, c in MC_SIZES)
    r0 = rng.uniform(0.08, 0.25, total)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4)
    classes = []
    first = 0
    for n, count in MC_SIZES:
        dx = 8.0 / n
        groups = []
        # each L0 group carries half the Cn2: r0_g^(-5/3) = 0.5 r0^(-5/3)
        scale = torch.tensor((r0[first:first + count] * 0.5 ** -0.6) ** (-5.0 / 6.0),
                             dtype=torch.float32, device=dev)
        G = len(MC_L0)
        n_fft = (count + 1) // 2
        noise = torch.randn((G, 2, n_fft, n, n), generator=gen, device=dev, dtype=torch.float32)
        shn = torch.randn((G, 8, 2, count, 3, 3), generator=gen, device=dev, dtype=torch.float32)
        for g, L0 in enumerate(MC_L0):
            amp, _, sh_amp = build_filters(n, dx, 1.0, L0, 8)
            groups.append(dict(amp=_hip.to_device(amp, np.float32, dev),
                               sh_amp=_hip.to_device(sh_amp, np.float64, dev),
                               amp_host=amp, sh_amp_host=sh_amp, noise=noise[g], shn=shn[g]))
        out = torch.empty((count, n, n), dtype=torch.float32, device=dev)
        from pyturb_amd.ops import work_elems

        work = torch.empty((work_elems(0, n, count, 8),), dtype=torch.complex64, device=dev)
        classes.append(dict(n=n, count=count, scale=scale, groups=groups, out=out, work=work,
                            r0=r0[first:first + count], noise=noise, shn=shn,
                            amp=torch.stack([grp["amp"] for grp in groups]),
                            sh_amp=torch.stack([grp["sh_amp"] for grp in groups])))
        first += count
    return classes


def ops_generate_groups(c):
    """One size class, all its L0 groups, one pt_generate_groups call."""
    from pyturb_amd import ops

    ops.generate_groups(c["noise"], c["shn"], c["amp"], c["sh_amp"], c["count"],
                        screen_scale=c["scale"], out=c["out"], work=c["work"])


def run_mc(classes):
    """Every size class in one pass over all its L0 groups (pt_generate_groups)."""
    for c in classes:
        ops_generate_groups(c)


def run_mc_persistent(classes, persistent_cols=False):
    """Every size class through pt_generate_ensemble: one persistent row pass
    and one prep over all classes, then the column passes per class (or one
    persistent column pass)."""
    from pyturb_amd import ops

    ens = c0 = classes[0].get("ensemble")
    if c0 is None:
        ens = [ops.EnsembleClass(c["noise"], c["shn"], c["amp"], c["sh_amp"], c["count"],
                                 screen_scale=c["scale"], out=c["out"], work=c["work"])
               for c in classes]
        classes[0]["ensemble"] = ens
    ops.generate_ensemble(ens, persistent_cols=persistent_cols)


def run_mc_plain(classes):
    """Plain form: reference generate() algorithm as torch ops per group."""
    import torch

    for c in classes:
        n, count = c["n"], c["count"]
        total = None
        for grp in c["groups"]:
            ctx = grp.setdefault("plain", _plain_gen_ctx(grp, n))
            amp, basis, shamp = ctx
            nz = grp["noise"]
            spec = torch.complex(nz[0], nz[1]) * amp
            field = torch.fft.ifft2(spec) * (n * n)
            scr = torch.cat((field.real, field.imag))[:count]
            shn = grp["shn"]
            cn = torch.complex(shn[:, 0], shn[:, 1]) * shamp  # (P, count, 3, 3)
            m = torch.matmul(cn, basis[:, None]).transpose(0, 1).reshape(count, -1, n)
            low = torch.matmul(basis.reshape(-1, n).transpose(0, 1).contiguous()[None], m).real
            low = low - low.mean(dim=(-2, -1), keepdim=True)
            scr = (scr + low) * c["scale"][:, None, None]
            total = scr if total is None else total + scr
        c["plain_out"] = total


def _plain_gen_ctx(grp, n):
    import torch

    dev = grp["amp"].device
    P = grp["sh_amp_host"].shape[0]
    fpn = np.array([-1.0, 0.0, 1.0])[None, :] / (n * 3.0 ** np.arange(1, P + 1))[:, None]
    x = np.arange(n) - n / 2.0
    basis = torch.tensor(np.exp(2j * np.pi * fpn[:, :, None] * x[None, None, :]),
                         dtype=torch.complex64, device=dev)
    return (grp["amp"], basis, torch.tensor(grp["sh_amp_host"], dtype=torch.complex64,
                                            device=dev)[:, None])


EXT_N = 1024
EXT_STEPS = 1000
LGS_ALT = 90e3
LGS_FRAMES = 512  # spectral LGS-cone frames per timed call (9 per-layer transforms each)


def build_extrude(dev, far_points=32, steps=EXT_STEPS):
    """9-layer 1024^2 extruder, 30" field, innovations from seed 3 in HBM;
    stencil of 2 full rows plus ``far_points`` far pixels."""
    import torch

    from pyturb_amd.extrude import ExtrudedAtmosphere

    layers = profile_layers()
    dx = 8.0 / EXT_N
    frac = np.array(FRACS) / np.sum(FRACS)
    r0s = 0.15 * frac ** -0.6
    winds = [ly.wind_vector for ly in layers]
    arcsec = np.pi / (180 * 3600)
    fov = [h * np.tan(30 * arcsec) / dx for h in HEIGHTS]
    gen = torch.Generator(device=dev)
    gen.manual_seed(3)
    # rows each layer extrudes over the run: travel + lookahead, with margin
    speeds = np.array([ly.wind_speed for ly in layers])
    rows = (speeds * steps * 1e-3 / dx + 8).astype(int)
    pool = {}
    used = [0] * len(layers)
    ext = None

    def innov(l, k):
        if l not in pool:  # rows with a pitch of a multiple of 4 floats (16-byte aligned)
            pitch = (ext.width + 3) // 4 * 4
            pool[l] = torch.randn((int(rows[l]) + 64, pitch), generator=gen, device=dev)[:, :ext.width]
        out = pool[l][used[l]:used[l] + k]
        used[l] += k
        return out

    ext = ExtrudedAtmosphere(n=EXT_N, pixel_scale=dx, layer_r0=list(r0s), layer_L0=[25.0] * 9,
                             layer_wind=winds, layer_altitude_los=HEIGHTS,
                             field_of_view_pix=fov, interp="linear", seeds=list(range(30, 39)),
                             innovations=innov, far_points=far_points)
    ang = np.radians(np.arange(6) * 60.0)
    thx = np.concatenate([np.tan(30 * arcsec) * np.cos(ang), [0.0]])
    thy = np.concatenate([np.tan(30 * arcsec) * np.sin(ang), [0.0]])
    mags = np.ones((7, 9))
    mags[6] = 1.0 - np.array(HEIGHTS) / LGS_ALT
    out = torch.empty((7, EXT_N, EXT_N), dtype=torch.float32, device=dev)
    return ext, thx, thy, mags, out


def run_extrude(ext, thx, thy, mags, out, steps, t0=0):
    scale = 500e-9 / (2 * np.pi)
    for k in range(steps):
        ext.set_time((t0 + k) * 1e-3)
        ext.integrate_dirs(thx, thy, scale=scale, mags=mags, out=out)


def profile_figures(path, prefix):
    """``{name: number}`` from the ``prefix<name>: <number>`` lines of a
    committed counter summary (empty if the file or the lines are absent)."""
    out = {}
    try:
        with open(os.path.join(ROOT, path)) as fh:
            for line in fh:
                if line.startswith(prefix) and ":" in line:
                    name, val = line[len(prefix):].split(":", 1)
                    try:
                        out[name.strip()] = float(val.strip().split()[0])
                    except (ValueError, IndexError):
                        pass
    except OSError:
        pass
    return out


def timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def host_description():
    """CPU model, core counts and library versions of this host."""
    import platform

    import scipy

    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        import numba  # noqa: F401

        numba_v = numba.__version__
    except ImportError:
        numba_v = None
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(),
            "cores_available": len(os.sched_getaffinity(0)),
            "numpy": np.__version__, "scipy": scipy.__version__, "numba": numba_v}


def cpu_frames(atm, budget_s):
    """FP64 oracle frames of the headline workload for about budget_s."""
    import oracle as O

    spectra = atm._spectra.cpu().numpy().astype(np.complex128)
    coeffs = atm._sh_coeffs.cpu().numpy()
    f = np.fft.fftfreq(N, d=atm.pixel_scale)
    filt = O.build_filters(N, atm.pixel_scale, 0.15, 25.0, subharmonics=atm.subharmonics)
    done, t0 = 0, time.perf_counter()
    while True:
        t = done * DT
        O.integrate(spectra, coeffs, f, filt["sh_freq"], filt["sh_basis"], atm._vx * t,
                    atm._vy * t)
        done += 1
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    return done / el, done, el


def cpu_mc(budget_s):
    """FP64 oracle screens of the mixed ensemble (both L0 groups summed, per-
    screen r0), in the classes' 4:3:1 proportion (a round is 8 screens at 256^2,
    6 at 512^2 and 2 at 1024^2), for about budget_s."""
    import oracle as O

    rng = np.random.default_rng(4)
    rounds = []
    for n, count in ((256, 8), (512, 6), (1024, 2)):
        filts = [O.build_filters(n, 8.0 / n, 1.0, L0, subharmonics=8) for L0 in MC_L0]
        noise = [O.draw_generate_noise(rng, n, count, 8) for _ in MC_L0]
        gain = (rng.uniform(0.08, 0.25, count) * 0.5 ** -0.6) ** (-5.0 / 6.0)
        rounds.append((filts, noise, gain, count))
    done, t0 = 0, time.perf_counter()
    while True:
        for filts, noise, gain, count in rounds:
            tot = sum(O.generate(fl, nz[0], nz[1], count) for fl, nz in zip(filts, noise))
            tot *= gain[:, None, None]
            done += count
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    return done / el, done, el


def cpu_baseline(atm, budget_s=7.0):
    """The FP64 NumPy oracle on this host: frames and Monte-Carlo screens, each
    on 1 thread (NumPy pocketfft, BLAS limited to one thread) and on every
    available core (scipy.fft workers, all BLAS threads)."""
    import oracle as O
    from threadpoolctl import threadpool_limits

    host = host_description()
    cores = host["cores_available"]
    rows = []
    for threads in (1, cores):
        O.set_fft_workers(None if threads == 1 else threads)
        try:
            with threadpool_limits(limits=threads):
                fr, fd, fe = cpu_frames(atm, budget_s)
                mr, md, me = cpu_mc(budget_s)
        finally:
            O.set_fft_workers(None)
        rows.append({"workload": "frames", "threads": threads, "value": round(fr, 2),
                     "unit": "frames/s", "sample": f"{fd} frames in {fe:.1f} s"})
        rows.append({"workload": "monte_carlo", "threads": threads, "value": round(mr, 2),
                     "unit": "screens/s", "sample": f"{md} mixed screens in {me:.1f} s"})
    return rows, host


def parity(atm, out, frames=(0, 1000, 1999)):
    import oracle as O

    spectra = atm._spectra.cpu().numpy().astype(np.complex128)
    coeffs = atm._sh_coeffs.cpu().numpy()
    f = np.fft.fftfreq(N, d=atm.pixel_scale)
    filt = O.build_filters(N, atm.pixel_scale, 0.15, 25.0, subharmonics=atm.subharmonics)
    worst = 0.0
    scale = atm.wavelength / (2 * np.pi)
    for k in frames:
        t = k * DT
        ref = O.integrate(spectra, coeffs, f, filt["sh_freq"], filt["sh_basis"], atm._vx * t,
                          atm._vy * t) * scale
        worst = max(worst, O.rel_rms(out[k].cpu().numpy(), ref))
    return worst


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the engine clock settles over the first ~10 calls of 2000 frames
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain GPU form")
    ap.add_argument("--no-mc", action="store_true", help="skip the Monte-Carlo ensemble")
    ap.add_argument("--no-extrude", action="store_true", help="skip the extrusion workload")
    ap.add_argument("--no-boil", action="store_true", help="skip the boiling-frames workload")
    ap.add_argument("--no-lgs", action="store_true", help="skip the LGS-cone frames workload")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the built-in FFT launch shapes (skip the autotuner)")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the 16-atmosphere HBM-streamed frames variant")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the one-frame-per-call loop (eager vs HIP-graph replay)")
    ap.add_argument("--extrude-far", choices=("both", "32", "0"), default="both",
                    help="extrusion stencils to run: 2 rows + 32 far points, 2 rows, or both")
    ap.add_argument("--extrude-steps", type=int, default=1000,
                    help="extrusion steps per stencil (20 of them untimed warm-up)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy 3 times (counter scaling)")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        import torch

        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": a.numel() * 4, "copies": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the frame and Monte-Carlo workloads shard by frames/screens with no "
              "exchange; see DESIGN.md)", file=sys.stderr)
        return 2
    import torch

    from pyturb_amd import _hip

    dev = _hip.device(0)
    tuned = None
    if not args.no_tune:
        from pyturb_amd import tune

        print("bench: autotuning the FFT launch shapes", file=sys.stderr, flush=True)
        sizes = sorted({N} | {n for n, _ in MC_SIZES})
        tune.autotune(sizes)
        tuned = {str(n): {k: tune.describe(k, n) for k in tune.KNOBS} for n in sizes}
        torch.cuda.empty_cache()
    print("bench: building 9-layer 512^2 spectral workload", file=sys.stderr, flush=True)
    atm, times, disp = build_frames_workload(dev)
    out = torch.empty((N_FRAMES, N, N), dtype=torch.float32, device=dev)
    from pyturb_amd.ops import frames_work_elems

    L = len(atm.layers)
    work = torch.empty((frames_work_elems(N, N_FRAMES, L, atm.subharmonics),),
                       dtype=torch.complex64, device=dev)

    for _ in range(args.warmup):
        run_frames(atm, disp, out, work)
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        run_frames(atm, disp, out, work, evs[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fps = args.steps * N_FRAMES / elapsed
    t_a = float(np.mean([e[0].elapsed_time(e[1]) for e in evs])) * 1e-3
    t_b = float(np.mean([e[1].elapsed_time(e[2]) for e in evs])) * 1e-3
    # per-frame pass figures: pass A reads the 8L n^2 B stack and writes an
    # 8 n^2 B intermediate; pass B reads that intermediate (one c64 grid per
    # frame pair, 4 n^2 B per frame) and writes the 4 n^2 B frame
    bytes_a = (8.0 * L + 8.0) * N * N * N_FRAMES
    bytes_b = 8.0 * N * N * N_FRAMES
    ach_a = bytes_a / t_a / 1e9
    frame_bytes = (8.0 * L + 4.0) * N * N
    print(f"bench: frames {fps:.0f}/s pass A {t_a*1e3:.3f} ms pass B {t_b*1e3:.3f} ms",
          file=sys.stderr, flush=True)
    par = parity(atm, out)

    extra = {}
    if not args.no_extrude:
        # the stated stencil (2 rows + 32 far points), then the reference's
        # full-row stencil alone
        legs = [(32, "extrude"), (0, "extrude_rows_only")]
        if args.extrude_far != "both":
            legs = [lg for lg in legs if str(lg[0]) == args.extrude_far]
        ext_steps = max(21, args.extrude_steps)
        for far, key in legs:
            print(f"bench: building 9-layer 1024^2 extruder, {far} far points (host FP64 A/B)",
                  file=sys.stderr, flush=True)
            tb = time.perf_counter()
            ext, thx, thy, mags, eout = build_extrude(dev, far, ext_steps)
            extra[f"{key}_setup_s"] = round(time.perf_counter() - tb, 1)
            run_extrude(ext, thx, thy, mags, eout, 20)
            torch.cuda.synchronize()
            te0 = time.perf_counter()
            run_extrude(ext, thx, thy, mags, eout, ext_steps - 20, t0=20)
            torch.cuda.synchronize()
            te = time.perf_counter() - te0
            extra[f"{key}_steps_per_s"] = round((ext_steps - 20) / te, 1)
            extra[f"{key}_frames_per_s_7dirs"] = round(7 * (ext_steps - 20) / te, 1)
            extra["extrude_width"] = ext.width
            extra["extrude_steps"] = ext_steps
            extra["extrude_far_points"] = 32
            del ext, eout
            torch.cuda.empty_cache()
    if not args.no_graph:
        print("bench: one frame per call, eager vs HIP-graph replay", file=sys.stderr, flush=True)
        from pyturb_amd.graph import CapturedFrames

        from pyturb_amd.graph import FoldedFrames

        cap = CapturedFrames(atm, 1)
        loop = FoldedFrames(atm, 1)
        ncall = 500
        tg = timed(lambda: [cap([float(t)]) for t in times[:ncall]], 1, 1)
        tf = timed(lambda: [loop([float(t)]) for t in times[:ncall]], 1, 1)
        te = timed(lambda: [atm.opd(float(t)) for t in times[:ncall]], 1, 1)
        extra["loop_eager_frames_per_s"] = round(ncall / te, 1)
        extra["loop_graph_frames_per_s"] = round(ncall / tg, 1)
        extra["loop_folded_frames_per_s"] = round(ncall / tf, 1)
        del cap, loop
    # the public generator (chunked device calls, one yield per frame)
    def frames_api():
        for _ in atm.frames(DT, N_FRAMES):
            pass

    atm.reset()
    tapi = timed(frames_api, 1, 1)
    extra["frames_api_frames_per_s"] = round(N_FRAMES / tapi, 1)
    if not args.no_plain:
        # plain vs fused on matched call patterns: a batch of frames per call
        # (the headline's 2000-frame call against the plain form in batches of
        # 50), and one frame per call (eager against eager)
        ctx = plain_context(atm)
        nplain = 200
        tp = timed(lambda: [plain_frame(atm, float(t), ctx) for t in times[:nplain]], 1, 1)
        extra["plain_frames_per_s"] = round(nplain / tp, 1)
        chunk = 50
        tpb = timed(lambda: [plain_frames_batched(atm, times[k:k + chunk], ctx)
                             for k in range(0, N_FRAMES, chunk)], 1, 1)
        extra["plain_batched_frames_per_s"] = round(N_FRAMES / tpb, 1)
        extra["fused_vs_plain_frames_batched"] = round(fps / (N_FRAMES / tpb), 1)
        if "loop_eager_frames_per_s" in extra:
            extra["fused_vs_plain_frames_one_per_call"] = round(
                extra["loop_eager_frames_per_s"] / (nplain / tp), 1)
    if not args.no_mc:
        print("bench: building Monte-Carlo ensemble", file=sys.stderr, flush=True)
        classes = build_mc(dev)
        nscr = sum(c for _, c in MC_SIZES)
        reps_mc = max(1, args.steps // 2)
        # two library paths for the mixed ensemble: per-class
        # pt_generate_groups calls (the tuned launch shapes) and
        # pt_generate_ensemble (one persistent row pass and prep over every
        # class at the default shapes, per-class column passes); the faster
        # is the figure, both are reported, and the persistent column pass
        # beside them
        tpe = timed(lambda: run_mc_persistent(classes), reps_mc, 1)
        extra["mc_ensemble_screens_per_s"] = round(nscr * reps_mc / tpe, 1)
        tpg = timed(lambda: run_mc(classes), reps_mc, 1)
        extra["mc_per_class_screens_per_s"] = round(nscr * reps_mc / tpg, 1)
        tpc = timed(lambda: run_mc_persistent(classes, True), reps_mc, 1)
        extra["mc_persistent_cols_screens_per_s"] = round(nscr * reps_mc / tpc, 1)
        mc_fn, tm, mc_path = (run_mc_persistent, tpe, "pt_generate_ensemble (persistent rows + prep, "
                              "per-class columns)") if tpe <= tpg else (
            run_mc, tpg, "per-class pt_generate_groups calls")
        mc_rate = nscr * reps_mc / tm
        # each class's pt_generate_groups call on its own, HIP events on the
        # launch stream (row pass + subharmonic prep + column pass)
        cls_ms = {}
        for c in classes:
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            reps = 5
            ev[0].record()
            for _ in range(reps):
                ops_generate_groups(c)
            ev[1].record()
            torch.cuda.synchronize()
            ms = ev[0].elapsed_time(ev[1]) / reps
            alg = (4.0 * len(MC_L0) + 4.0) * c["n"] * c["n"] * c["count"]
            cls_ms[str(c["n"])] = {"screens": c["count"], "call_ms": round(ms, 4),
                                   "algorithmic_GBs": round(alg / ms / 1e6, 1),
                                   "frac_hbm": round(alg / ms / 1e6 / HBM_PEAK_GBS, 3)}
        extra["mc_classes"] = cls_ms
        px = sum(n * n * c for n, c in MC_SIZES)
        mc_bytes = (4.0 * len(MC_L0) + 4.0) * px
        extra["mc_screens_per_s"] = round(mc_rate, 1)
        extra["mc_ms_per_ensemble"] = round(1e3 * tm / reps_mc, 3)
        extra["mc_algorithmic_GBs"] = round(mc_bytes * mc_rate / nscr / 1e9, 1)
        extra["mc_frac_hbm_peak"] = round(mc_bytes * mc_rate / nscr / 1e9 / HBM_PEAK_GBS, 4)
        extra["mc_path"] = mc_path
        if not args.no_graph:
            from pyturb_amd.graph import Captured

            cap = Captured(lambda: mc_fn(classes))
            tg = timed(cap.replay, reps_mc, 1)
            extra["mc_graph_screens_per_s"] = round(nscr * reps_mc / tg, 1)
            del cap
        if not args.no_plain:
            tpm = timed(lambda: run_mc_plain(classes), 1, 1)
            extra["mc_plain_screens_per_s"] = round(nscr / tpm, 1)
            extra["fused_vs_plain_mc"] = round(mc_rate / (nscr / tpm), 1)
        del classes
        torch.cuda.empty_cache()
    if not args.no_stream:
        # The 18.9 MB layer stack of one atmosphere stays in the 256 MB MALL
        # across steps.  Sixteen atmospheres (302 MB of stacks) called in
        # turn, 125 frames each, make every call's fold read its stack from
        # HBM; pass A then reads the folded copy that the call just wrote.
        print("bench: 16 atmospheres in turn (HBM-streamed stacks)", file=sys.stderr, flush=True)
        import pyturb_amd as pa

        K = 16
        per = N_FRAMES // K
        ws = []
        for k in range(K):
            a_k = pa.Atmosphere(profile_layers(), r0=0.15, diameter=8.0, n=N, L0=25.0,
                                subharmonics=8, seed=100 + k, device="mi355x")
            ws.append(a_k)
        sdisp = disp[:per].contiguous()
        sout = out[:per]

        def stream_once():
            for a_k in ws:
                run_frames(a_k, sdisp, sout, work, frames=per)

        ts = timed(stream_once, max(1, args.steps // 2), 1)
        extra["stream16_frames_per_s"] = round(K * per * max(1, args.steps // 2) / ts, 1)
        extra["stream16_stack_MB"] = round(K * 8.0 * L * N * N / 1e6, 1)
        del ws
        torch.cuda.empty_cache()
    if not args.no_boil:
        print("bench: boiling frames", file=sys.stderr, flush=True)
        batm, bdisp = build_boil(dev)
        need = int(_hip.lib().pt_boil_work_elems(N, N_FRAMES, L, batm.subharmonics))
        bwork = torch.empty((need,), dtype=torch.complex64, device=dev)
        steps_done = [0]

        def boil_once():
            run_boil(batm, bdisp, out, bwork, steps_done[0])
            steps_done[0] += N_FRAMES

        tb = timed(boil_once, max(1, args.steps // 2), 1)
        extra["boil_frames_per_s"] = round(N_FRAMES * max(1, args.steps // 2) / tb, 1)

        def boil_api():  # the public iterator, chunked device calls
            for _ in batm.frames(DT, N_FRAMES):
                pass

        ta = timed(boil_api, 1, 1)
        extra["boil_api_frames_per_s"] = round(N_FRAMES / ta, 1)
        extra["boil_tau_s"] = BOIL_TAU
        extra["reference_published_rtx5090_boil_fps_512"] = 2134
        del batm, bdisp, bwork
        torch.cuda.empty_cache()
    if not args.no_lgs:
        print("bench: LGS-cone frames", file=sys.stderr, flush=True)
        import pyturb_amd as pa

        latm = pa.Atmosphere(profile_layers(), r0=0.15, diameter=8.0, n=N, L0=25.0,
                             subharmonics=8, seed=2, lgs_altitude=LGS_ALT, device="mi355x")
        ltimes = np.arange(LGS_FRAMES) * DT
        tl = timed(lambda: latm.opd_batch(ltimes), 1, 1)
        extra["lgs_frames_per_s"] = round(LGS_FRAMES / tl, 1)
        extra["reference_published_rtx5090_lgs_fps_512"] = 793
        del latm
        torch.cuda.empty_cache()
    from pyturb_amd import tune as _tune

    order = _tune.get("frames_order", N)
    tag = "column_first" if order == 0 else "row_first"
    fprof = profile_figures(FRAMES_PROFILE, "")
    traffic_a = fprof.get(f"traffic_bytes_per_launch_{tag}")
    traffic_b = fprof.get(f"passb_traffic_bytes_per_launch_{tag}")
    valu_a = fprof.get(f"valu_insts_per_launch_{tag}")
    src = (f"{FRAMES_PROFILE}: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a 1 GiB copy) "
           "and SQ_INSTS_VALU over the kernels of each call, bench.py --steps 2 --warmup 1")
    passes = [{
        "pass": "B: prep + " + ("k_frame_rows_epi<512>" if order == 0 else "k_gen_cols_set<512>"),
        "bound": "hbm", "call_ms": round(t_b * 1e3, 4),
        "algorithmic_bytes": bytes_b, "algorithmic_GBs": round(bytes_b / t_b / 1e9, 1),
        "frac": round(bytes_b / t_b / 1e9 / HBM_PEAK_GBS, 3),
        "traffic": traffic_b,
        "traffic_GBs": round(traffic_b / t_b / 1e9, 1) if traffic_b else None,
        "traffic_frac": round(traffic_b / t_b / 1e9 / HBM_PEAK_GBS, 3) if traffic_b else None}]
    mprof = profile_figures(MC_PROFILE, "mc_pass ")
    for name in sorted({k.rsplit("_", 1)[0] for k in mprof}):
        us, alg, tr = (mprof.get(f"{name}_us"), mprof.get(f"{name}_algorithmic"),
                       mprof.get(f"{name}_traffic"))
        if not us:
            continue
        passes.append({"pass": f"monte-carlo {name}", "bound": "hbm", "call_ms": us * 1e-3,
                       "source": f"{MC_PROFILE} (kernel trace and counters of the profiled run)",
                       "algorithmic_GBs": round(alg / us / 1e3, 1) if alg else None,
                       "frac": round(alg / us / 1e3 / HBM_PEAK_GBS, 3) if alg else None,
                       "traffic": tr,
                       "traffic_GBs": round(tr / us / 1e3, 1) if tr else None,
                       "traffic_frac": round(tr / us / 1e3 / HBM_PEAK_GBS, 3) if tr else None})
    cpu = None
    if not args.no_cpu:
        print("bench: CPU baseline (FP64 oracle: 1 thread, then every core)", file=sys.stderr,
              flush=True)
        rows, host = cpu_baseline(atm)
        first = rows[0]
        cpu = {"value": first["value"], "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"{first['sample']} of the same 9-layer 512^2 workload, FP64 NumPy "
                         "oracle (pocketfft), 1 thread",
               "rows": rows, "host": host}
    line = {
        "metric": METRIC,
        "value": round(fps, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded host-drawn layer spectra and winds; device-drawn MC noise)",
        "config": {"workload": "9-layer 512^2 spectral frozen flow, 2000 frames, dt=1 ms",
                   "layers": L, "n": N, "frames_per_step": N_FRAMES, "subharmonics": 8},
        # pass A (pt_spectral_rows: k_frame_step + fold + pair tables + pass A),
        # timed with HIP events on its stream.  Its stack is served from L2, so
        # the algorithmic bytes exceed the HBM peak and do not bound it: it is
        # bound by VALU issue (wave64 VALU instructions per launch from the
        # profile over the live launch time, against the non-packed peak)
        "roofline": {
            "kernel": ("pass A: k_frame_pair_cols<512> (column-first)" if order == 0
                       else "pass A: k_frame_pair_rows<512> (row-first)"),
            "bound": "valu",
            "achieved": round(valu_a / t_a / 1e9, 1) if valu_a else None,
            "peak": VALU_PEAK_GIPS, "unit": "G wave64 VALU instructions/s",
            "frac": round(valu_a / t_a / 1e9 / VALU_PEAK_GIPS, 3) if valu_a else None,
            "traffic": traffic_a, "traffic_source": src,
            "traffic_GBs": round(traffic_a / t_a / 1e9, 1) if traffic_a else None,
            "hbm_frac": round(traffic_a / t_a / 1e9 / HBM_PEAK_GBS, 3) if traffic_a else None,
            "call_ms": round(t_a * 1e3, 4),
            "algorithmic_bytes": bytes_a,
            "algorithmic_GBs_L2_served": round(ach_a, 1),
            "frame_algorithmic_GBs": round(frame_bytes * fps / 1e9, 1),
            "passes": passes},
        "cpu_baseline": cpu,
        "autotuned_shapes": tuned,
        "parity_rel_rms_vs_fp64": float(f"{par:.3g}"),
        "reference_published_rtx5090_fps_512": 3133,
    }
    line.update(extra)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
