"""GPU parity at the workloads' own sizes.

* The mixed Monte-Carlo ensemble (2048 x 256^2 + 1536 x 512^2 + 512 x
  1024^2 screens, two outer-scale groups L0 = 25 / 50 m each carrying half
  the Cn2, per-screen r0 ~ U(0.08, 0.25) m from seed 4, device-drawn noise):
  every class is generated whole by one pt_generate_groups call, and screen
  pairs at the start, middle and end of each class are recomputed by the FP64
  oracle from the same noise (summed per-group generate, fourier.py:233-282,
  atmosphere.py:439-470).
* The 1024^2 extrusion (9 layers, W = 2194): 20 steps of 1 ms against the
  FP64 oracle extruder on the same seed rows and innovations, read out for 6
  directions on a 30" ring plus a 90 km LGS cone in one pt_readout call
  (extrude.py:424-469, 754-915).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-5
MC_SIZES = [(256, 2048), (512, 1536), (1024, 512)]
MC_L0 = [25.0, 50.0]
HEIGHTS = [0.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0, 12000.0, 16000.0, 20000.0]
FRACS = [0.5, 0.1, 0.08, 0.07, 0.06, 0.06, 0.05, 0.05, 0.03]


def test_mixed_ensemble_full_size_matches_oracle(hip):
    import torch

    from pyturb_amd import ops
    from pyturb_amd.spectrum import build_filters

    dev = torch.device("cuda")
    rng = np.random.default_rng(4)
    r0 = rng.uniform(0.08, 0.25, sum(c for _, c in MC_SIZES))
    gen = torch.Generator(device=dev)
    gen.manual_seed(4)
    first = 0
    for n, count in MC_SIZES:
        print(f"ensemble class {count} x {n}^2, 2 groups", flush=True)
        dx, P, G = 8.0 / n, 8, len(MC_L0)
        n_fft = (count + 1) // 2
        r0c = r0[first:first + count]
        first += count
        # each group carries half the Cn2: r0_g^(-5/3) = 0.5 r0^(-5/3)
        gain = (r0c * 0.5 ** -0.6) ** (-5.0 / 6.0)
        noise = torch.randn((G, 2, n_fft, n, n), generator=gen, device=dev)
        shn = torch.randn((G, P, 2, count, 3, 3), generator=gen, device=dev)
        filts = [build_filters(n, dx, 1.0, L0, P) for L0 in MC_L0]
        amp = torch.tensor(np.stack([f[0] for f in filts]), dtype=torch.float32, device=dev)
        sha = torch.tensor(np.stack([f[2] for f in filts]), dtype=torch.float64, device=dev)
        out = ops.generate_groups(noise, shn, amp, sha, count,
                                  screen_scale=torch.tensor(gain, dtype=torch.float32, device=dev))
        oracle_filts = [O.build_filters(n, dx, 1.0, L0, subharmonics=P) for L0 in MC_L0]
        for k in (0, n_fft // 2, n_fft - 1):
            pair = [k, n_fft + k]  # real and imaginary part of transform k
            ref = 0.0
            for g, f in enumerate(oracle_filts):
                nz = noise[g][:, k:k + 1].double().cpu().numpy()
                sh = shn[g][:, :, pair].double().cpu().numpy()
                ref = ref + O.generate(f, nz, sh, 2)
            ref = ref * np.float32(gain[pair])[:, None, None]
            got = out[pair].cpu().numpy()
            err = O.rel_rms(got, ref)
            assert err < REL, (n, k, err)
        del noise, shn, out
        torch.cuda.empty_cache()


def test_extrusion_1024_ring_and_lgs_match_oracle(hip):
    import torch

    from pyturb_amd.extrude import ExtrudedAtmosphere

    n = 1024
    dx = 8.0 / n
    rng = np.random.default_rng(2)
    speeds = rng.uniform(5.0, 35.0, 9)
    dirs = rng.uniform(0.0, 2.0 * np.pi, 9)
    winds = [(float(s * np.cos(d)), float(s * np.sin(d))) for s, d in zip(speeds, dirs)]
    frac = np.array(FRACS) / np.sum(FRACS)
    r0s = list(0.15 * frac ** -0.6)
    fov = [h * np.tan(30 * O.ARCSEC) / dx for h in HEIGHTS]
    print("oracle extruder 1024^2, 9 layers (FP64 A/B)", flush=True)
    ref = O.OracleExtruder(n, dx, r0s, [25.0] * 9, winds, HEIGHTS, fov, stencil_rows=2)
    W = ref.W
    # seed rows: von Karman screens of W x W at each layer's r0, as the
    # extruder seeds itself (extrude.py:408-420), drawn on the host
    filt = O.build_filters(W, dx, 1.0, 25.0, subharmonics=8)
    main, sh = O.draw_generate_noise(np.random.default_rng(3), W, 10, 8)
    unit = O.generate(filt, main, sh, 10)
    seeds = [(unit[l, :s] * r0s[l] ** (-5.0 / 6.0)).astype(np.float32)
             for l, s in enumerate(ref.spans)]
    del main, unit
    ref.seed([s.astype(np.float64) for s in seeds])
    streams = [np.random.default_rng(300 + l) for l in range(9)]
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

    print("device extruder 1024^2", flush=True)
    ext = ExtrudedAtmosphere(n=n, pixel_scale=dx, layer_r0=r0s, layer_L0=[25.0] * 9,
                             layer_wind=winds, layer_altitude_los=HEIGHTS, field_of_view_pix=fov,
                             interp="linear", innovations=d_innov,
                             seed_screens=[torch.from_numpy(s).to(dev) for s in seeds])
    assert ext.width == W == 2194
    ang = np.radians(np.arange(6) * 60.0)
    thx = np.concatenate([np.tan(30 * O.ARCSEC) * np.cos(ang), [0.0]])
    thy = np.concatenate([np.tan(30 * O.ARCSEC) * np.sin(ang), [0.0]])
    lgs_mag = 1.0 - np.array(HEIGHTS) / 90e3
    mags = np.ones((7, 9))
    mags[6] = lgs_mag
    worst = 0.0
    for step in range(21):
        t = step * 1e-3
        ref.set_time(t, o_innov)
        ext.set_time(t)
        if step in (0, 20):
            print(f"readout step {step}", flush=True)
            got = ext.integrate_dirs(thx, thy, mags=mags).cpu().numpy()
            for d in range(6):
                want = ref.integrate(thx[d], thy[d], interp="linear")
                err = O.rel_rms(got[d], want)
                print(f"  direction {d}: {err:.2e}", flush=True)
                worst = max(worst, err)
            sa, sp = ref.shifts(0.0, 0.0)
            grids = [O.pupil_grids(n, ref.cos[l], ref.sin[l], W, lgs_mag[l]) for l in range(9)]
            want = O.readout(ref.buf, ref.fill, np.stack([g[0] for g in grids]),
                             np.stack([g[1] for g in grids]), sa, sp, "linear")
            err = O.rel_rms(got[6], want)
            print(f"  LGS cone: {err:.2e}", flush=True)
            worst = max(worst, err)
    assert sum(used) == sum(len(d) for d in drawn) > 0
    print("worst rel rms", worst)
    assert worst < REL
