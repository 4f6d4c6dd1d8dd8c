"""Pin the FP64 oracle: golden vectors from the reference, closed forms,
identities and brute force.  CPU only."""

from __future__ import annotations

import math
import warnings

import numpy as np
import pytest

import oracle as O
from conftest import config2_layer_tuples, load_golden


# ---------------------------------------------------------------- filters
def test_psd_closed_form_and_limits():
    # 0.023 r0^-5/3 (f^2 + L0^-2)^(-11/6): check two hand-evaluated points.
    r0, L0 = 0.15, 25.0
    f2 = np.array([0.0, 4.0])
    want = 0.023 * r0 ** (-5 / 3) * (f2 + 1 / 625.0) ** (-11 / 6)
    np.testing.assert_allclose(O.psd(f2, r0, L0), want, rtol=1e-14)
    # Kolmogorov: pure power law, slope -11/3 in |f|
    p1, p2 = O.psd(np.array([1.0, 4.0]), 1.0, np.inf)
    assert math.isclose(math.log(p2 / p1) / math.log(2.0), -11 / 3, rel_tol=1e-12)
    # inner scale multiplies by exp(-f^2/fm^2), fm = 5.92/(2 pi l0)
    fm = 5.92 / (2 * np.pi * 0.2)
    assert math.isclose(O.psd(1.0, 1.0, np.inf, inner_scale=0.2),
                        O.psd(1.0, 1.0, np.inf) * math.exp(-1 / fm ** 2), rel_tol=1e-13)


def test_band_amplitude_flat_limit_and_centre_rule():
    # A cell far from DC has a nearly flat PSD: amp -> sqrt(PSD(fc) * w^2).
    fx, fy, w = 40.0, 13.0, 0.01
    amp = O.band_amplitude(fx, fy, w, 0.15, 25.0)
    flat = math.sqrt(O.psd(fx * fx + fy * fy, 0.15, 25.0)) * w
    assert abs(amp / flat - 1) < 1e-5
    # Near DC the integrated amplitude differs strongly from the centre rule.
    amp0 = O.band_amplitude(0.125, 0.0, 0.125, 0.15, 25.0)
    centre = math.sqrt(O.psd(0.125 ** 2, 0.15, 25.0)) * 0.125
    assert abs(amp0 / centre - 1) > 0.01


def test_filters_match_reference_golden():
    g = load_golden("generate_n256_seed1.npz")
    filt = O.build_filters(256, 0.03125, 0.15, 25.0, subharmonics=3)
    idx = np.r_[0:9, 248:256]
    np.testing.assert_allclose(filt["amplitude"][np.ix_(idx, idx)], g["amp_block"],
                               rtol=1e-13, atol=0)
    np.testing.assert_allclose(filt["amplitude"][5], g["amp_row5"], rtol=1e-13)
    np.testing.assert_allclose(filt["sh_amp"], g["sh_amp"], rtol=1e-13)
    np.testing.assert_allclose(filt["sh_basis"], g["sh_basis"], rtol=0, atol=1e-14)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(4, 4), (8, 2), (3, 5), (16, 16)])
def test_ifft2_unnormalised_equals_brute_force(shape):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    n0, n1 = shape
    j0 = np.arange(n0)
    j1 = np.arange(n1)
    F0 = np.exp(2j * np.pi * np.outer(j0, j0) / n0)
    F1 = np.exp(2j * np.pi * np.outer(j1, j1) / n1)
    np.testing.assert_allclose(O.ifft2_unnormalised(x), F0 @ x @ F1, atol=1e-11)


# ---------------------------------------------------------------- generate
def test_generate_matches_reference_golden():
    g = load_golden("generate_n256_seed1.npz")
    filt = O.build_filters(256, 0.03125, 0.15, 25.0, subharmonics=3)
    rng = np.random.default_rng(1)
    main, sh = O.draw_generate_noise(rng, 256, 64, 3)
    scr = O.generate(filt, main, sh, 64)
    np.testing.assert_allclose(scr[:2], g["screens01"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(scr[63, 100:132, 40:72], g["screen63_patch"], atol=1e-10)
    np.testing.assert_allclose(scr.sum(axis=(1, 2)), g["sums"], rtol=1e-9, atol=1e-7)
    np.testing.assert_allclose((scr * scr).sum(axis=(1, 2)), g["sumsq"], rtol=1e-11)


def test_generate_pairs_real_and_imag_parts():
    # screens[k] and screens[n_fft + k] are Re/Im of one coloured FFT
    filt = O.build_filters(16, 0.1, 0.2, 10.0, subharmonics=0)
    rng = np.random.default_rng(3)
    main, sh = O.draw_generate_noise(rng, 16, 6, 0)
    scr = O.generate(filt, main, sh, 6)
    field = O.ifft2_unnormalised((main[0] + 1j * main[1]) * filt["amplitude"])
    np.testing.assert_allclose(scr[:3], field.real, atol=1e-12)
    np.testing.assert_allclose(scr[3:], field.imag, atol=1e-12)


def test_subharmonic_field_is_zero_mean_and_matches_direct_sum():
    rng = np.random.default_rng(5)
    filt = O.build_filters(12, 0.2, 0.1, 30.0, subharmonics=2)
    c = rng.standard_normal((2, 3, 3)) + 1j * rng.standard_normal((2, 3, 3))
    low = O.subharmonic_field(c, filt["sh_basis"])
    assert abs(low.mean()) < 1e-12
    x = (np.arange(12) - 6.0) * 0.2
    direct = np.zeros((12, 12))
    for p in range(2):
        fp = filt["sh_freq"][p]
        for a in range(3):
            for b in range(3):
                direct += (c[p, a, b] * np.exp(2j * np.pi * (fp[a] * x[:, None]
                                                              + fp[b] * x[None, :]))).real
    np.testing.assert_allclose(low, direct - direct.mean(), atol=1e-11)


# ---------------------------------------------------------------- spectral
@pytest.fixture(scope="module")
def atm64():
    return O.OracleAtmosphere(config2_layer_tuples(), r0=0.15, diameter=8.0, n=64,
                              L0=25.0, field_of_view=30.0, dtype="float64", seed=2)


def test_atmosphere_seed_tree_matches_reference(atm64):
    g = load_golden("atmosphere_n64_seed2.npz")
    assert atm64.ns == int(g["n_screen"])
    np.testing.assert_allclose(np.abs(atm64.spectra).sum(axis=(1, 2)),
                               g["spectra_abs_sum"], rtol=1e-12)
    np.testing.assert_allclose(atm64.sh_coeffs, g["sh_coeffs"], rtol=1e-12, atol=0)


def test_spectral_frames_match_reference_golden(atm64):
    g = load_golden("atmosphere_n64_seed2.npz")
    for t, want in zip(g["times"], g["frames"]):
        np.testing.assert_allclose(atm64.opd(float(t)), want, rtol=0,
                                   atol=1e-12 * np.abs(want).max())
    got = atm64.opd(0.01, directions=[tuple(d) for d in g["dirs"]])
    np.testing.assert_allclose(got, g["dframes"], atol=1e-12 * np.abs(got).max())
    nofov = O.OracleAtmosphere(config2_layer_tuples(), r0=0.15, diameter=8.0, n=64,
                               L0=25.0, dtype="float64", seed=2)
    want = g["frame_nofov_t002"]
    np.testing.assert_allclose(nofov.opd(0.02), want, atol=1e-12 * np.abs(want).max())


def test_integer_pixel_shift_is_a_roll():
    # atmosphere test_frozen_flow_shifts_screen_by_wind: +s rolls by -k.
    filt = O.build_filters(32, 0.25, 0.15, 25.0, subharmonics=0)
    rng = np.random.default_rng(9)
    spec, cf = O.draw_flow(rng, filt)
    base = O.flow_translate(spec, cf, filt, 0.0, 0.0)
    moved = O.flow_translate(spec, cf, filt, 5 * 0.25, -3 * 0.25)
    np.testing.assert_allclose(moved, np.roll(np.roll(base, -5, 0), 3, 1), atol=1e-10)


def test_layer_sum_equals_sum_of_translates(atm64):
    sx, sy = atm64.displacements(0.3)
    total = O.integrate(atm64.spectra, atm64.sh_coeffs, atm64.f, atm64.sh_freq,
                        atm64.sh_basis, sx, sy)
    parts = sum(O.flow_translate(atm64.spectra[l], atm64.sh_coeffs[:, l],
                                 atm64.filters[l], sx[l], sy[l])
                for l in range(len(sx)))
    np.testing.assert_allclose(total, parts, rtol=1e-10, atol=1e-10)


def test_sample_groups_match_reference_golden():
    g = load_golden("sample_n64_seed4.npz")
    layers = [(0, 0.4, 10.0, 0.0, 25.0), (1000, 0.2, 10.0, 0.0, 50.0),
              (5000, 0.3, 10.0, 0.0, 25.0), (10000, 0.1, 10.0, 0.0, 50.0)]
    atm = O.OracleAtmosphere(layers, r0=0.15, diameter=8.0, n=64, dtype="float64", seed=4)
    assert len(atm.sample_groups) == 2
    got = atm.sample(4)
    np.testing.assert_allclose(got, g["sample"], rtol=0, atol=1e-12 * np.abs(got).max())


def test_r0_scaling_is_exact_power():
    # amplitude ~ r0^(-5/6) for every cell, integrated or not
    a = O.build_filters(32, 0.1, 0.1, 25.0, subharmonics=2)
    b = O.build_filters(32, 0.1, 0.3, 25.0, subharmonics=2)
    k = (0.3 / 0.1) ** (5 / 6)
    np.testing.assert_allclose(a["amplitude"], b["amplitude"] * k, rtol=1e-12)
    np.testing.assert_allclose(a["sh_amp"], b["sh_amp"] * k, rtol=1e-12)


# ---------------------------------------------------------------- covariance
def test_phase_covariance_variance_and_decay():
    r0, L0 = 0.15, 25.0
    c0 = O.phase_covariance(np.array([0.0]), r0, L0)[0]
    assert c0 == pytest.approx(0.0863 * (L0 / r0) ** (5 / 3), rel=2e-3)
    c = O.phase_covariance(np.linspace(0, 50, 200), r0, L0)
    assert np.all(np.diff(c) < 0)
    # continuity at r -> 0
    assert O.phase_covariance(np.array([1e-9]), r0, L0)[0] == pytest.approx(c0, rel=1e-5)


def test_extrusion_matrices_identities_and_golden():
    g = load_golden("extrusion_W32.npz")
    A, B = O.build_extrusion(32, 2, 0.125, 25.0)
    np.testing.assert_allclose(A, g["A"], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(B, g["B"], rtol=1e-8, atol=1e-10)
    # A C_zz = C_xz and B B^T = C_xx - A C_xz^T
    yy, xx = np.meshgrid(np.arange(2.0), np.arange(32.0), indexing="ij")
    zx, zy = xx.ravel(), yy.ravel()
    nx, ny = np.arange(32.0), np.full(32, 2.0)
    cov = lambda ax, ay, bx, by: O.phase_covariance(
        np.hypot(ax[:, None] - bx[None], ay[:, None] - by[None]) * 0.125, 1.0, 25.0)
    czz, cxz, cxx = cov(zx, zy, zx, zy), cov(nx, ny, zx, zy), cov(nx, ny, nx, ny)
    np.testing.assert_allclose(A @ czz, cxz, atol=1e-6 * np.abs(cxz).max())
    np.testing.assert_allclose(B @ B.T, cxx - A @ cxz.T, atol=1e-8 * np.abs(cxx).max())


def test_far_stencil_geometry_known_values():
    """The far points' definition: R = F / 8 far rows at (m - 1) +
    round((W // 2)^(j / R)) rows behind the newest stencil row, 8 columns
    (2 i + 1) W // 16 each; values worked by hand from that definition."""
    dist, cols = O.far_stencil(2194, 2, 32)  # H = 1097: 1097^(1/4) = 5.76, ^(1/2) = 33.1, ^(3/4) = 190.7
    assert dist == [7, 34, 192, 1098]
    assert cols == [137, 411, 685, 959, 1234, 1508, 1782, 2056]
    assert O.far_stencil(40, 2, 32)[0] == [3, 5, 10, 21]  # H = 20: 2.11, 4.47, 9.46, 20
    assert O.far_stencil(40, 3, 16)[0] == [6, 22]  # m - 1 = 2: 2 + round(4.47), 2 + 20
    assert O.far_stencil(10, 2, 32)[0] == [2, 3, 4, 6]  # H = 5: 1.50 -> 1, 2.24 -> 2, 3.34 -> 3, 5
    assert O.far_stencil(40, 2, 0)[0] == []
    # the product's host geometry is the same definition
    from pyturb_amd.infinite import far_columns, far_rows_back
    for W, m, F in ((2194, 2, 32), (40, 2, 32), (40, 3, 16), (10, 2, 32), (68, 2, 24)):
        d, c = O.far_stencil(W, m, F)
        assert list(far_rows_back(W, m, F)) == d
        assert list(far_columns(W)) == c


def test_far_point_extrusion_identities():
    """With far points, A C_zz = C_xz and B B^T = C_xx - A C_xz^T over the
    full stencil (m rows, then the far points row by row); the extra data can
    only lower the conditional variance; far_points = 0 is the plain build;
    the product's FP64 build is the same."""
    W, m, F, dx, L0 = 24, 2, 16, 0.1, 25.0
    A, B = O.build_extrusion(W, m, dx, L0, far_points=F)
    assert A.shape == (W, m * W + F) and B.shape == (W, W)
    dist, cols = O.far_stencil(W, m, F)
    pts = [(x, y) for y in range(m) for x in range(W)]
    pts += [(x, m - 1 - d) for d in dist for x in cols]
    zx, zy = np.array(pts, float).T
    nx, ny = np.arange(W, dtype=float), np.full(W, float(m))
    cov = lambda ax, ay, bx, by: O.phase_covariance(
        np.hypot(ax[:, None] - bx[None], ay[:, None] - by[None]) * dx, 1.0, L0)
    czz, cxz, cxx = cov(zx, zy, zx, zy), cov(nx, ny, zx, zy), cov(nx, ny, nx, ny)
    np.testing.assert_allclose(A @ czz, cxz, atol=1e-6 * np.abs(cxz).max())
    np.testing.assert_allclose(B @ B.T, cxx - A @ cxz.T, atol=1e-8 * np.abs(cxx).max())
    A0, B0 = O.build_extrusion(W, m, dx, L0)
    assert np.trace(B @ B.T) < np.trace(B0 @ B0.T)
    A00, B00 = O.build_extrusion(W, m, dx, L0, far_points=0)
    assert np.array_equal(A00, A0) and np.array_equal(B00, B0)
    from pyturb_amd.infinite import build_extrusion as product_build
    Ap, Bp = product_build(W, m, dx, L0, far_points=F)
    np.testing.assert_allclose(Ap, A, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(Bp @ Bp.T, B @ B.T, rtol=1e-9, atol=1e-12)


def test_far_point_extruder_keeps_far_rows_and_big_jump_equals_small_steps():
    """The far rows stay in the ring through compactions (the oracle asserts
    each one is there), and one jump equals many small steps."""
    def run(times):
        ext = O.OracleExtruder(n=16, dx=0.5, layer_r0=[0.2, 0.3], layer_L0=[20.0, 20.0],
                               layer_wind=[(7.0, 3.0), (-2.0, -9.0)], layer_alt=[0.0, 3000.0],
                               far_points=32)
        rng = np.random.default_rng(4)
        ext.seed([rng.standard_normal((s, ext.W)) for s in ext.spans])
        streams = [np.random.default_rng(100 + l) for l in range(2)]
        innov = lambda l, k: streams[l].standard_normal(ext.W)
        for t in times:
            ext.set_time(t, innov)
        return ext.integrate(), ext
    big, ext = run([20.0])
    small, _ = run(np.linspace(0, 20.0, 101))
    assert ext.innov_count.min() > 3 * ext.cap  # several compactions
    np.testing.assert_allclose(big, small, rtol=0, atol=1e-12 * 40)


# ---------------------------------------------------------------- extruder
def _reference_like_extruder(interp_seeds=(11, 12)):
    """Oracle extruder fed with the reference's own seed screens/innovations."""
    kw = dict(n=32, dx=0.25, layer_r0=[0.2, 0.35], layer_L0=[25.0, 25.0],
              layer_wind=[(6.0, 2.5), (-3.0, 11.0)], layer_alt=[0.0, 8000.0],
              fov_pix=[0.0, 3.0])
    ext = O.OracleExtruder(**kw)
    rngs = [np.random.default_rng(s) for s in interp_seeds]
    seeds = []
    for l, rng in enumerate(rngs):
        filt = O.build_filters(ext.W, 0.25, kw["layer_r0"][l], 25.0, subharmonics=8)
        main, sh = O.draw_generate_noise(rng, ext.W, 1, 8)
        seeds.append(O.generate(filt, main, sh, 1)[0])
    ext.seed(seeds)
    innov = lambda l, k: rngs[l].standard_normal(ext.W)
    return ext, innov


@pytest.mark.parametrize("interp", ["linear", "cubic", "lanczos"])
def test_extruder_readout_matches_reference_golden(interp):
    g = load_golden("extrude_n32.npz")[interp]
    ext, innov = _reference_like_extruder()
    k = 0
    for t in (0.0, 0.01, 0.037, 0.4):
        ext.set_time(t, innov)
        for th in ((0.0, 0.0), (2e-5, -1e-5)):
            got = ext.integrate(*th, interp=interp)
            np.testing.assert_allclose(got, g[k], rtol=0, atol=1e-9 * np.abs(g[k]).max())
            k += 1


def test_extruder_integer_travel_is_exact_translation():
    # Along-wind integer travel shifts the readout by whole rows (wind along axis 0).
    ext = O.OracleExtruder(n=16, dx=0.5, layer_r0=[0.2], layer_L0=[20.0],
                           layer_wind=[(10.0, 0.0)], layer_alt=[0.0])
    rng = np.random.default_rng(1)
    ext.seed([rng.standard_normal((ext.spans[0], ext.W))])
    innov = lambda l, k: rng.standard_normal(ext.W)
    ext.set_time(0.0, innov)
    a = ext.integrate()
    ext.set_time(3 * 0.5 / 10.0, innov)  # exactly 3 pixels
    b = ext.integrate()
    np.testing.assert_allclose(b[:-3], a[3:], atol=1e-12)


def test_extruder_big_jump_equals_small_steps():
    def run(times):
        ext = O.OracleExtruder(n=16, dx=0.5, layer_r0=[0.2, 0.3], layer_L0=[20.0, 20.0],
                               layer_wind=[(7.0, 3.0), (-2.0, -9.0)], layer_alt=[0.0, 3000.0])
        rng = np.random.default_rng(4)
        ext.seed([rng.standard_normal((s, ext.W)) for s in ext.spans])
        streams = [np.random.default_rng(100 + l) for l in range(2)]
        innov = lambda l, k: streams[l].standard_normal(ext.W)
        for t in times:
            ext.set_time(t, innov)
        return ext.integrate()
    # The recurrence is identical; only compaction bookkeeping (the ring base)
    # differs, which moves the last bits of the fractional readout position.
    np.testing.assert_allclose(run([3.0]), run(np.linspace(0, 3.0, 37)),
                               rtol=0, atol=1e-12 * 40)


def test_readout_weights_partition_unity():
    t = np.linspace(0, 1, 17, endpoint=False)
    assert np.allclose(sum(O.catmull_rom_weights(t)), 1.0)
    assert np.allclose(sum(O.lanczos_weights(t)[1]), 1.0)


@pytest.mark.parametrize("interp", ["cubic", "linear"])
def test_infinite_matches_reference_golden(interp):
    g = load_golden("infinite_n32_seed5.npz")[interp]
    rng = np.random.default_rng(5)
    filt = O.build_filters(32, 0.1, 0.15, 25.0, subharmonics=8)
    main, sh = O.draw_generate_noise(rng, 32, 1, 8)
    seed = O.generate(filt, main, sh, 1)[0]
    inf = O.OracleInfinite(32, 0.1, 0.15, 25.0, seed, interp=interp)
    innov = lambda k: rng.standard_normal(32)
    seq = [inf._advance_to(0.0, innov)]
    for tr in (0.3, 1.3, 4.0, 44.0):
        seq.append(inf._advance_to(tr, innov))
    np.testing.assert_allclose(np.stack(seq), g, atol=1e-10 * np.abs(g).max())


def test_structure_function_of_oracle_screens_matches_theory():
    # Kolmogorov ensemble, D(r) = 6.88 (r/r0)^(5/3) within 15% (fourier tests)
    filt = O.build_filters(128, 0.02, 0.1, np.inf, subharmonics=8)
    rng = np.random.default_rng(7)
    main, sh = O.draw_generate_noise(rng, 128, 16, 8)
    scr = O.generate(filt, main, sh, 16)
    r, d = O.structure_function(scr, 0.02, max_separation=20)
    ratio = d[3:] / (6.88 * (r[3:] / 0.1) ** (5 / 3))
    assert np.all((ratio > 0.8) & (ratio < 1.2))


# ---------------------------------------------------------------------------
# Zernike modes and Noll variances (validation statistics)
# ---------------------------------------------------------------------------
def test_noll_index_table():
    """Known answers of tests/test_analysis.py:10-14, plus the sine/cosine
    parity rule: even j carries cos (m >= 0), odd j sin (m < 0), m = 0 apart."""
    expected = {1: (0, 0), 2: (1, 1), 3: (1, -1), 4: (2, 0), 5: (2, -2),
                6: (2, 2), 7: (3, -1), 8: (3, 1), 11: (4, 0)}
    for j, nm in expected.items():
        assert O.noll_to_zernike(j) == nm
    seen = set()
    for j in range(1, 67):
        n, m = O.noll_to_zernike(j)
        assert (n - abs(m)) % 2 == 0 and abs(m) <= n
        if m:
            assert (m > 0) == (j % 2 == 0)
        seen.add((n, m))
    assert len(seen) == 66  # radial orders 0..10, every (n, m) exactly once
    with pytest.raises(ValueError):
        O.noll_to_zernike(0)


def test_zernike_radial_closed_forms():
    rho = np.linspace(0.0, 1.0, 11)
    from oracle.pyturb_oracle import _zernike_radial as R

    assert np.allclose(R(2, 0, rho), 2 * rho ** 2 - 1)
    assert np.allclose(R(3, 1, rho), 3 * rho ** 3 - 2 * rho)
    assert np.allclose(R(4, 0, rho), 6 * rho ** 4 - 6 * rho ** 2 + 1)
    assert np.allclose(R(4, 2, rho), 4 * rho ** 4 - 3 * rho ** 2)
    for n in range(8):
        for m in range(n % 2, n + 1, 2):
            assert R(n, m, np.array([1.0]))[0] == pytest.approx(1.0)  # R_n^m(1) = 1


def test_zernike_basis_orthonormal_and_decomposition_exact():
    basis = O.zernike_basis(15, 256)
    pupil = basis[0] != 0
    vecs = basis[:, pupil]
    gram = vecs @ vecs.T / pupil.sum()
    assert np.allclose(np.diag(gram), 1.0, atol=0.02)
    assert np.max(np.abs(gram - np.diag(np.diag(gram)))) < 0.02
    rng = np.random.default_rng(0)
    truth = rng.standard_normal((3, 15))
    stack = np.einsum("cj,jxy->cxy", truth, basis)
    assert np.allclose(O.zernike_decompose(stack, 15, basis=basis), truth, atol=1e-10)
    assert np.allclose(O.zernike_decompose(stack[0], 15), truth[0], atol=1e-10)


def test_noll_residuals_table_and_asymptote():
    r = [O.noll_residual(j) for j in range(1, 40)]
    assert all(a > b for a, b in zip(r, r[1:]))
    assert O.noll_residual(1) == pytest.approx(1.0299)
    assert O.noll_residual(3) == pytest.approx(0.134)
    assert O.noll_residual(22) == pytest.approx(0.2944 * 22 ** (-np.sqrt(3) / 2))
    # tip + tilt carry 0.448 (D/r0)^(5/3) together (Noll 1976, Delta_1 - Delta_3)
    assert O.noll_variance(2, 8.0, 0.2) + O.noll_variance(3, 8.0, 0.2) == pytest.approx(
        (1.0299 - 0.134) * 40.0 ** (5 / 3))
    with pytest.raises(ValueError):
        O.noll_variance(1, 8.0, 0.2)


# ---------------------------------------------------------------------------
# spectral boiling (atmosphere.py:618-658, 1092-1189, 1221-1243)
# ---------------------------------------------------------------------------
_BOIL_LAYERS = [(0.0, 0.5, 8.0, 30.0, 25.0), (3000.0, 0.3, 14.0, 110.0, 25.0),
                (9000.0, 0.2, 21.0, 250.0, 25.0)]
_BOIL_TAU = [0.02, np.inf, 0.05]


def _boil_oracle():
    return O.OracleAtmosphere(_BOIL_LAYERS, 0.15, diameter=4.0, n=32, subharmonics=3,
                              dtype="float64", seed=6, tau_boil=_BOIL_TAU)


def test_boiling_matches_reference_golden():
    """Seed tree, draw order, AR(1) update, frames/evolve/opd/reset."""
    g = load_golden("boil_n32_seed6.npz")
    atm = _boil_oracle()
    main_tau, sh_tau = atm.boil_tau_maps()
    assert np.allclose(main_tau, g["main_tau"], rtol=1e-12)
    assert np.allclose(sh_tau, g["sh_tau"], rtol=1e-12)
    seq = [f for _, f in atm.frames(0.004, 5)]
    seq.append(atm.evolve(0.004))
    seq.append(atm.opd(0.1))
    assert O.rel_rms(np.stack(seq), g["frames"]) < 1e-12
    atm.reset()
    again = np.stack([f for _, f in atm.frames(0.004, 2)])
    assert O.rel_rms(again, g["after_reset"]) < 1e-12


def test_boil_tau_maps_closed_form():
    """tau = tau_boil below f_ref, tau_boil (f/f_ref)^(-2/3) above; frozen
    layers keep tau = inf (alpha = 1, nothing changes)."""
    atm = _boil_oracle()
    main_tau, _ = atm.boil_tau_maps()
    f_ref = 1.0 / 25.0
    fr = np.hypot(atm.f[:, None], atm.f[None, :])
    expect = 0.02 * np.where(fr > f_ref, (fr / f_ref) ** (-2.0 / 3.0), 1.0)
    assert np.allclose(main_tau[0], expect, rtol=1e-13)
    assert np.all(np.isinf(main_tau[1]))
    before = atm.spectra[1].copy()
    atm.boil_step(0.01)
    assert np.array_equal(atm.spectra[1], before)


def test_boiling_is_stationary():
    """The AR(1) step keeps each mode's variance: a stationary ensemble of
    spectra stays stationary (mean |S|^2 / A^2 ~ 2 after many steps)."""
    atm = _boil_oracle()
    amp = np.stack([fl["amplitude"] for fl in atm.filters])
    for _ in range(200):
        atm.boil_step(0.01)
    live = amp[0] > 0
    ratio = np.mean(np.abs(atm.spectra[0][live]) ** 2 / amp[0][live] ** 2)
    assert abs(ratio - 2.0) < 0.15


def test_philox_known_answers_and_normals():
    """Philox4x32-10 against the Random123 known-answer vectors (Salmon et
    al. 2011), and the Box-Muller normals' first two moments."""
    from oracle.pyturb_oracle import box_muller, philox4x32_10

    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF),
            (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, expect in kat:
        got = philox4x32_10(np.array(ctr, dtype=np.uint32), key)
        assert [int(v) for v in got] == list(expect)
    ctr = np.zeros((200000, 4), dtype=np.uint32)
    ctr[:, 0] = np.arange(200000)
    r = philox4x32_10(ctr, (7, 9))
    z = np.concatenate(box_muller(r[:, 0], r[:, 1]) + box_muller(r[:, 2], r[:, 3]))
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.01


def test_device_boil_noise_layout():
    """Each mode and its mirror get distinct normals from one counter; the
    stream depends only on (seed, step)."""
    a, sa = O.device_boil_noise(99, 3, 2, 16, 2)
    b, _ = O.device_boil_noise(99, 3, 2, 16, 2)
    c, _ = O.device_boil_noise(99, 4, 2, 16, 2)
    assert np.array_equal(a, b) and not np.allclose(a, c)
    assert a.shape == (2, 2, 16, 16) and sa.shape == (2, 2, 2, 3, 3)
    assert len(np.unique(a[0, 0])) == 256  # no mode shares a value with its mirror


# ---------------------------------------------------------------------------
# spectral LGS cone (atmosphere.py:581-616, 1000-1090)
# ---------------------------------------------------------------------------
LGS_DIRS = [(0.0, 0.0), (6.0, -4.0)]


def test_lgs_frames_match_reference_golden():
    g = load_golden("lgs_n32_seed7.npz")
    for interp in ("linear", "cubic", "lanczos"):
        atm = O.OracleAtmosphere(_BOIL_LAYERS, 0.15, diameter=4.0, n=32, subharmonics=3,
                                 field_of_view=10.0, dtype="float64", seed=7,
                                 lgs_altitude=20e3, interp=interp)
        assert np.allclose(atm.lgs_mag, g[interp + "_mag"], rtol=1e-14)
        got = np.stack([atm.opd(t, directions=LGS_DIRS) for t in (0.0, 0.013)])
        assert O.rel_rms(got, g[interp]) < 1e-12, interp


def test_lgs_zoom_taps_identities():
    """mag = 1 on an unpadded grid samples the screen exactly (identity
    taps); the weights of every output pixel sum to 1."""
    for interp in ("linear", "cubic", "lanczos"):
        idx, w = O.lgs_zoom_taps(16, 16, [1.0, 0.7], interp)
        assert np.allclose(w.sum(axis=1), 1.0)
        screen = np.random.default_rng(0).standard_normal((16, 16))
        rows = np.einsum("ti,tic->ic", w[0], screen[idx[0]])
        out = np.einsum("tj,itj->ij", w[0], rows[:, idx[0]])
        assert np.allclose(out, screen, atol=1e-12)


# ---------------------------------------------------------------------------
# extruder boiling (extrude.py:472-496, 660-752)
# ---------------------------------------------------------------------------
def test_extruder_boiling_matches_reference_golden():
    g = load_golden("extrude_boil_n32.npz")["frames"]
    kw = dict(n=32, dx=0.25, layer_r0=[0.2, 0.35], layer_L0=[25.0, 25.0],
              layer_wind=[(6.0, 2.5), (-3.0, 11.0)], layer_alt=[0.0, 8000.0],
              fov_pix=[0.0, 3.0], tau_boil=[0.03, 0.08], boil_seed=77)
    ext = O.OracleExtruder(**kw)
    rngs = [np.random.default_rng(s) for s in (11, 12)]
    seeds = []
    for l, rng in enumerate(rngs):
        filt = O.build_filters(ext.W, 0.25, kw["layer_r0"][l], 25.0, subharmonics=8)
        main, sh = O.draw_generate_noise(rng, ext.W, 1, 8)
        seeds.append(O.generate(filt, main, sh, 1)[0])
    ext.seed(seeds)
    innov = lambda l, k: rngs[l].standard_normal(ext.W)
    got = []
    for k in range(5):
        ext.set_time(k * 0.01, innov)
        got.append(ext.integrate(0.0, 0.0))
        got.append(ext.integrate(2e-5, -1e-5))
        ext.boil_step(0.01)
    assert O.rel_rms(np.stack(got), g) < 1e-9


def test_extruder_boiling_preserves_covariance():
    """A fresh window has the stencil covariance: S S^T = C_zz (seed) and
    every blend keeps a^2 + b^2 = 1."""
    S = O.extrusion_seed_factor(24, 2, 0.25, 25.0)
    zy, zx = np.divmod(np.arange(48.0), 24)
    czz = O.phase_covariance(np.hypot(zx[:, None] - zx, zy[:, None] - zy) * 0.25, 1.0, 25.0)
    assert np.allclose(S @ S.T, czz, atol=1e-9 * czz.max())
    ext = O.OracleExtruder(n=16, dx=0.25, layer_r0=[0.2], layer_L0=[25.0],
                           layer_wind=[(5.0, 0.0)], layer_alt=[0.0], tau_boil=[0.05])
    rng = np.random.default_rng(2)
    ext.seed([rng.standard_normal((ext.spans[0], ext.W))])
    ext.set_time(0.3, lambda l, k: rng.standard_normal(ext.W))
    lo, hi, a, b = ext.boil_region(0, 0.01)
    assert np.isclose(a * a + b * b, 1.0) and np.isclose(a, np.exp(-0.01 / 0.05))
    # the window starts 3 rows below the lowest row the pupil reads
    assert lo == int(np.floor(ext.along_min[0] + ext.travel[0])) - ext.base[0] - 3
    assert hi == ext.fill[0]
