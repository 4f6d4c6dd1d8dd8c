"""Write the golden fixtures in this folder from the reference pyturb.

Run once, in the development container that has the reference checkout:

    python tests/golden/make_golden.py /root/reference/src

It imports the reference (pure Python; NumPy/SciPy only, float64 paths) and
stores small ``.npz`` files of inputs and outputs for the parts of the hot
path that closed forms and identities do not fix on their own:

* the near-DC integrated band amplitudes and subharmonic amplitudes/bases
  (fourier.py:151-228) and the exact ``generate`` realisation for a seed
  (fourier.py:233-282, including the RNG draw order);
* the Atmosphere seed tree and spectral frames (atmosphere.py:393-470,
  882-998) on a 9-layer n=64 grid, on- and off-axis, and ``sample``;
* the extrusion matrices (extrude.py:250-304) at W=32;
* extruder readouts (extrude.py:754-915) for linear/cubic/lanczos on a
  2-layer n=32 atmosphere, and InfinitePhaseScreen frames (infinite.py);
* spectral boiling (atmosphere.py:618-658, 1092-1189, 1221-1243): the
  per-mode time constants and the boiled frames of a 3-layer n=32
  atmosphere through frames(), evolve() and reset(), which fix the boiling
  stream's seed and draw order.

* the spectral LGS cone (atmosphere.py:581-616, 1000-1090): frames of a
  3-layer n=32 atmosphere for a 20 km beacon, on- and off-axis, with each
  interpolation.

* extruder boiling (extrude.py:472-496, 660-752, NumPy path): readouts of a
  2-layer n=32 extruder boiling at two rates, between set_time steps.

``python tests/golden/make_golden.py /root/reference/src boil`` (or ``lgs``,
``extboil``) writes only that fixture.

Only data is written here; nothing from the reference is copied into the
repository, and nothing at test time reads the reference.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def config2_layers(pyturb):
    """The 9-layer profile used by the spectral-frame workload (seed=2)."""
    rng = np.random.default_rng(2)
    speeds = rng.uniform(5.0, 35.0, 9)
    dirs = rng.uniform(0.0, 2.0 * np.pi, 9)
    heights = [0.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0, 12000.0, 16000.0, 20000.0]
    fracs = [0.5, 0.1, 0.08, 0.07, 0.06, 0.06, 0.05, 0.05, 0.03]
    return [pyturb.Layer(h, f, float(v), float(np.degrees(d)), 25.0)
            for h, f, v, d in zip(heights, fracs, speeds, dirs)]


BOIL_LAYERS = [(0.0, 0.5, 8.0, 30.0, 25.0), (3000.0, 0.3, 14.0, 110.0, 25.0),
               (9000.0, 0.2, 21.0, 250.0, 25.0)]  # altitude, Cn2, speed, direction deg, L0
BOIL_TAU = [0.02, float("inf"), 0.05]


def write_boil(pyturb) -> None:
    layers = [pyturb.Layer(*ly) for ly in BOIL_LAYERS]
    atm = pyturb.Atmosphere(layers, r0=0.15, diameter=4.0, n=32, subharmonics=3,
                            tau_boil=BOIL_TAU, dtype="float64", seed=6)
    main_tau, sh_tau = atm._boil_main_tau, atm._boil_sh_tau
    seq = [f.copy() for _, f in atm.frames(0.004, 5)]
    seq.append(atm.evolve(0.004).copy())
    seq.append(atm.opd(0.1).copy())  # opd does not boil
    atm.reset()
    again = [f.copy() for _, f in atm.frames(0.004, 2)]
    np.savez_compressed(os.path.join(HERE, "boil_n32_seed6.npz"),
                        main_tau=main_tau, sh_tau=sh_tau, frames=np.stack(seq),
                        after_reset=np.stack(again))


LGS_DIRS = [(0.0, 0.0), (6.0, -4.0)]  # arcsec, inside the 10" field


def write_lgs(pyturb) -> None:
    out = {}
    for interp in ("linear", "cubic", "lanczos"):
        atm = pyturb.Atmosphere([pyturb.Layer(*ly) for ly in BOIL_LAYERS], r0=0.15,
                                diameter=4.0, n=32, subharmonics=3, field_of_view=10.0,
                                lgs_altitude=20e3, interp=interp, dtype="float64", seed=7)
        out[interp] = np.stack([atm.opd(t, directions=LGS_DIRS) for t in (0.0, 0.013)])
        out[interp + "_mag"] = atm._lgs_mag
    np.savez_compressed(os.path.join(HERE, "lgs_n32_seed7.npz"), **out)


def write_extboil(pyturb) -> None:
    from pyturb.extrude import ExtrudedAtmosphere

    ext = ExtrudedAtmosphere(
        n=32, pixel_scale=0.25, layer_r0=[0.2, 0.35], layer_L0=[25.0, 25.0],
        layer_wind=[(6.0, 2.5), (-3.0, 11.0)], layer_altitude_los=[0.0, 8000.0],
        field_of_view_pix=[0.0, 3.0], interp="cubic", dtype="float64", seeds=[11, 12],
        tau_boil=[0.03, 0.08], boil_seed=77)
    frames = []
    for k in range(5):
        ext.set_time(k * 0.01)
        frames.append(ext.integrate(0.0, 0.0))
        frames.append(ext.integrate(2e-5, -1e-5))
        ext.boil_step(0.01)
    np.savez_compressed(os.path.join(HERE, "extrude_boil_n32.npz"), frames=np.stack(frames))


def main(ref_src: str, only: str = "") -> None:
    sys.path.insert(0, ref_src)
    import pyturb  # noqa: E402  (reference, this container only)
    from pyturb.extrude import ExtrudedAtmosphere, build_extrusion

    writers = {"boil": write_boil, "lgs": write_lgs, "extboil": write_extboil}
    if only in writers:
        writers[only](pyturb)
        print(only, "fixture written to", HERE)
        return

    # 1. generate: n=256, dx=1/32, r0=0.15, L0=25, 3 subharmonic levels, seed 1.
    gen = pyturb.PhaseScreen(n=256, pixel_scale=0.03125, r0=0.15, L0=25.0,
                             subharmonics=3, seed=1, dtype="float64")
    scr = gen.generate(64)
    idx = np.r_[0:9, 248:256]
    np.savez_compressed(
        os.path.join(HERE, "generate_n256_seed1.npz"),
        amp_block=gen._amplitude[np.ix_(idx, idx)],
        amp_row5=gen._amplitude[5],
        sh_amp=np.stack([a for a, _ in gen._sh_bases]),
        sh_basis=np.stack([b for _, b in gen._sh_bases]),
        screens01=scr[:2],
        screen63_patch=scr[63, 100:132, 40:72],
        sums=scr.sum(axis=(1, 2)),
        sumsq=(scr * scr).sum(axis=(1, 2)),
    )

    # 2. 9-layer spectral atmosphere, n=64, float64, seed=2, 30" field.
    layers = config2_layers(pyturb)
    atm = pyturb.Atmosphere(layers, r0=0.15, diameter=8.0, n=64, L0=25.0,
                            field_of_view=30.0, dtype="float64", seed=2)
    times = np.array([0.0, 0.0137, 0.5])
    frames = np.stack([atm.opd(float(t)) for t in times])
    dirs = np.array([[0.0, 0.0], [30.0, 0.0], [-10.0, 20.0]])
    dframes = atm.opd(0.01, directions=[tuple(d) for d in dirs])
    atm_nofov = pyturb.Atmosphere(layers, r0=0.15, diameter=8.0, n=64, L0=25.0,
                                  dtype="float64", seed=2)
    f0 = atm_nofov.opd(0.02)
    np.savez_compressed(
        os.path.join(HERE, "atmosphere_n64_seed2.npz"),
        times=times, frames=frames, dirs=dirs, dframes=dframes,
        frame_nofov_t002=f0,
        spectra_abs_sum=np.abs(atm._spectra).sum(axis=(1, 2)),
        sh_coeffs=atm._sh_coeffs, n_screen=atm.n_screen,
    )

    # 3. sample() with two outer-scale groups (L0 25 / 50), n=64, float64.
    two = [pyturb.Layer(h, f, 10.0, 0.0, L0)
           for h, f, L0 in zip([0, 1000, 5000, 10000], [0.4, 0.2, 0.3, 0.1],
                               [25.0, 50.0, 25.0, 50.0])]
    atm_s = pyturb.Atmosphere(two, r0=0.15, diameter=8.0, n=64,
                              dtype="float64", seed=4)
    np.savez_compressed(os.path.join(HERE, "sample_n64_seed4.npz"),
                        sample=atm_s.sample(4))

    # 4. extrusion matrices at W=32.
    a_mat, b_unit, _ = build_extrusion(32, 2, 0.125, 25.0, np, np.float64)
    np.savez_compressed(os.path.join(HERE, "extrusion_W32.npz"), A=a_mat, B=b_unit)

    # 5. extruder readouts on a 2-layer n=32 atmosphere.
    out = {}
    for interp in ("linear", "cubic", "lanczos"):
        ext = ExtrudedAtmosphere(
            n=32, pixel_scale=0.25, layer_r0=[0.2, 0.35], layer_L0=[25.0, 25.0],
            layer_wind=[(6.0, 2.5), (-3.0, 11.0)],
            layer_altitude_los=[0.0, 8000.0], field_of_view_pix=[0.0, 3.0],
            interp=interp, dtype="float64", seeds=[11, 12])
        frames_i = []
        for t in (0.0, 0.01, 0.037, 0.4):
            ext.set_time(t)
            frames_i.append(ext.integrate(0.0, 0.0))
            frames_i.append(ext.integrate(2e-5, -1e-5))
        out[interp] = np.stack(frames_i)
    np.savez_compressed(os.path.join(HERE, "extrude_n32.npz"), **out)

    # 6. InfinitePhaseScreen frames (cubic and linear).
    inf = {}
    for interp in ("cubic", "linear"):
        scr_inf = pyturb.InfinitePhaseScreen(n=32, pixel_scale=0.1, r0=0.15,
                                             L0=25.0, interp=interp, seed=5,
                                             dtype="float64")
        seq = [scr_inf.screen.copy(), scr_inf.advance(0.3).copy(),
               scr_inf.step(1).copy(), scr_inf.advance(2.7).copy(),
               scr_inf.advance(40.0).copy()]
        inf[interp] = np.stack(seq)
    np.savez_compressed(os.path.join(HERE, "infinite_n32_seed5.npz"), **inf)

    # 7. spectral boiling; 8. spectral LGS cone; 9. extruder boiling.
    write_boil(pyturb)
    write_lgs(pyturb)
    write_extboil(pyturb)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src",
         sys.argv[2] if len(sys.argv) > 2 else "")
