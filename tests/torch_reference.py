"""Plain-torch autograd form of the differentiable frame model (test reference).

The same model as ``pyturb_amd.fit.ProfileFrames`` written with plain torch
ops (complex128 by default, CPU), so torch's own autograd supplies the
reference gradient the HIP adjoint is checked against.  Filters are rebuilt
from the outer scale with torch ops (regular cells sqrt(PSD) df, the
near-DC cells and the subharmonic cells by the 33 x 33 midpoint quadrature
of src/pyturb/fourier.py:151-228), so dL0 flows through them too.
"""

from __future__ import annotations

import math

import numpy as np
import torch


def _psd(f2, L0, power_law):
    return 0.023 * (f2 + L0 ** -2.0) ** (-0.5 * power_law)


def _band(fx, fy, width, L0, power_law, samples=33):
    u = ((torch.arange(samples, dtype=torch.float64) + 0.5) / samples - 0.5) * width
    gx = fx[:, None, None] + u[None, :, None]
    gy = fy[:, None, None] + u[None, None, :]
    f2 = gx * gx + gy * gy
    p = _psd(f2, L0, power_law)
    power = p.mean(dim=(1, 2)) * width * width
    f2e = (p * f2).sum(dim=(1, 2)) / p.sum(dim=(1, 2))
    return torch.sqrt(power * f2e / (fx * fx + fy * fy))


def filters(n, dx, L0, levels, power_law=11.0 / 3.0):
    """Unit-r0 (amplitude (n, n), sh_amp (P, 3, 3)) as differentiable functions of L0."""
    df = 1.0 / (n * dx)
    f = torch.tensor(np.fft.fftfreq(n, d=dx), dtype=torch.float64)
    f2 = f[:, None] ** 2 + f[None, :] ** 2
    f2 = torch.where(f2 == 0, torch.ones_like(f2), f2)
    amp = torch.sqrt(_psd(f2, L0, power_law)) * df
    mask = torch.ones((n, n), dtype=torch.float64)
    mask[0, 0] = 0.0
    amp = amp * mask
    kmax = min(8, (n - 1) // 2)
    if kmax > 0:
        ks = np.arange(-kmax, kmax + 1)
        ki, kj = np.meshgrid(ks, ks, indexing="ij")
        keep = (ki != 0) | (kj != 0)
        ki, kj = ki[keep] % n, kj[keep] % n
        vals = _band(f[ki], f[kj], df, L0, power_law)
        amp = amp.index_put((torch.tensor(ki), torch.tensor(kj)), vals)
    off = torch.tensor([-1.0, 0.0, 1.0], dtype=torch.float64)
    ia, ib = np.meshgrid(np.arange(3), np.arange(3), indexing="ij")
    keep = ((ia != 1) | (ib != 1)).ravel()
    sh = []
    for p in range(levels):
        dfp = df / 3.0 ** (p + 1)
        vals = _band(off[ia.ravel()[keep]] * dfp, off[ib.ravel()[keep]] * dfp, dfp, L0, power_law)
        blk = torch.zeros(9, dtype=torch.float64).index_put((torch.tensor(np.nonzero(keep)[0]),), vals)
        sh.append(blk.reshape(3, 3))
    sh_amp = torch.stack(sh) if sh else torch.zeros((0, 3, 3), dtype=torch.float64)
    return amp, sh_amp


def frames(noise, sh_noise, gain, L0, disp, dx, levels, scale, power_law=11.0 / 3.0):
    """(T, n, n) frames; noise (L, 2, n, n), sh_noise (P, L, 2, 3, 3), disp (T, L, 2)."""
    L, n = noise.shape[0], noise.shape[-1]
    f = torch.tensor(np.fft.fftfreq(n, d=dx), dtype=torch.float64)
    w = torch.complex(noise[:, 0].double(), noise[:, 1].double())
    spec, coeffs = [], []
    v = torch.complex(sh_noise[:, :, 0].double(), sh_noise[:, :, 1].double())  # (P, L, 3, 3)
    for l in range(L):
        amp, sh_amp = filters(n, dx, L0[l], levels, power_law)
        spec.append(gain[l] * amp * w[l])
        coeffs.append(gain[l] * sh_amp * v[:, l])
    S = torch.stack(spec)
    two_pi_i = 2j * math.pi
    px = torch.exp(two_pi_i * disp[:, :, 0, None] * f)  # (T, L, n)
    py = torch.exp(two_pi_i * disp[:, :, 1, None] * f)
    X = (S[None] * px[:, :, :, None] * py[:, :, None, :]).sum(dim=1)
    total = (torch.fft.ifft2(X) * (n * n)).real
    if levels:
        C = torch.stack(coeffs, dim=1)  # (P, L, 3, 3)
        df = 1.0 / (n * dx)
        fp = torch.tensor([-1.0, 0.0, 1.0], dtype=torch.float64)[None, :] * \
            (df / 3.0 ** torch.arange(1, levels + 1, dtype=torch.float64))[:, None]  # (P, 3)
        x = (torch.arange(n, dtype=torch.float64) - n / 2.0) * dx
        basis = torch.exp(two_pi_i * fp[:, :, None] * x[None, None, :])  # (P, 3, n)
        qx = torch.exp(two_pi_i * disp[:, None, :, 0, None] * fp[None, :, None, :])  # (T, P, L, 3)
        qy = torch.exp(two_pi_i * disp[:, None, :, 1, None] * fp[None, :, None, :])
        Q = (C[None] * qx[..., :, None] * qy[..., None, :]).sum(dim=2)  # (T, P, 3, 3)
        m = torch.einsum("tpab,pby->tpay", Q, basis)
        low = torch.einsum("pax,tpay->txy", basis, m).real
        low = low - low.mean(dim=(-2, -1), keepdim=True)
        total = total + low
    return total * scale
