"""pt_spectral_frames_noalloc (static device scratch, no workspace argument)
against pt_spectral_frames with a caller workspace, through the C-ABI.

Two calls in a row on the shared scratch, the second on a smaller input
(fewer layers, levels and frames, a smaller grid, a crop), then a general-size
grid; each must equal pt_spectral_frames bit for bit on the same inputs.
The reference workspaces are filled with NaN first, so neither side can lean
on memory that happens to be zero.  An oversize call is refused on the host.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, L, P, F, seed, uniform):
    import torch

    rng = np.random.default_rng(seed)
    dev = torch.device("cuda")
    spectra = torch.from_numpy((rng.standard_normal((L, n, n)) + 1j * rng.standard_normal((L, n, n)))
                               .astype(np.complex64)).to(dev)
    times = np.arange(F) * 1e-3 if uniform else np.sort(rng.uniform(0, 0.3, F))
    v = rng.uniform(-30, 30, (L, 2))
    disp = torch.from_numpy(np.ascontiguousarray(times[:, None, None] * v[None])).to(dev)
    coeffs = (torch.from_numpy((rng.standard_normal((P, L, 3, 3)) + 1j * rng.standard_normal((P, L, 3, 3))))
              .to(dev) if P else None)
    return spectra, disp, coeffs


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _run(hip, spectra, disp, coeffs, dx, off, n_out, scale, noalloc):
    import torch

    from pyturb_amd import _hip

    L, n = spectra.shape[0], spectra.shape[-1]
    F = disp.shape[0]
    P = 0 if coeffs is None else coeffs.shape[0]
    out = torch.full((F, n_out, n_out), float("nan"), dtype=torch.float32, device=spectra.device)
    st = ctypes.c_void_p(_hip.stream_ptr())
    if noalloc:
        rc = hip.pt_spectral_frames_noalloc(_ptr(spectra), L, n, dx, _ptr(disp), F, _ptr(coeffs), P, off,
                                            n_out, scale, _ptr(out), st)
    else:
        work = torch.full((int(hip.pt_frames_work_elems(n, F, L, P)),), complex("nan+nanj"),
                          dtype=torch.complex64, device=spectra.device)
        rc = hip.pt_spectral_frames(_ptr(spectra), L, n, dx, _ptr(disp), F, _ptr(coeffs), P, off, n_out,
                                    scale, _ptr(out), _ptr(work), st)
    assert rc == 0, hip.pt_last_error()
    torch.cuda.synchronize()
    return out


def test_noalloc_frames_equal_workspace_frames(hip):
    import torch

    cases = [  # (n, layers, levels, frames, uniform times, crop offset, n_out, scale)
        (512, 9, 8, 12, True, 0, 512, 7.9e-8),
        (128, 3, 3, 5, False, 10, 100, 1.0),
        (96, 2, 2, 3, False, 0, 96, 0.5),
    ]
    for k, (n, L, P, F, uni, off, n_out, scale) in enumerate(cases):
        print(f"noalloc call {k}: n={n} layers={L} levels={P} frames={F}", flush=True)
        spectra, disp, coeffs = _case(n, L, P, F, 40 + k, uni)
        got = _run(hip, spectra, disp, coeffs, 8.0 / n, off, n_out, scale, True)
        want = _run(hip, spectra, disp, coeffs, 8.0 / n, off, n_out, scale, False)
        assert torch.isfinite(want).all()
        assert torch.equal(got, want), k


def test_noalloc_refuses_oversize_on_the_host(hip):
    from pyturb_amd import _hip

    # 512^2, 9 layers, 8 levels: 44 frames fit the 64 MiB scratch, 45 do not
    assert hip.pt_frames_work_elems(512, 44, 9, 8) * 8 <= 64 << 20
    assert hip.pt_frames_work_elems(512, 45, 9, 8) * 8 > 64 << 20
    rc = hip.pt_spectral_frames_noalloc(None, 9, 512, 1.0 / 64, None, 45, None, 8, 0, 512, 1.0, None, None)
    assert rc == _hip.PT_ERR_ARG
    assert b"static scratch" in hip.pt_last_error()
