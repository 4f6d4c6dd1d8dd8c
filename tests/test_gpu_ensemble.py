"""The mixed-size Monte-Carlo launch (pt_generate_ensemble: persistent row
pass and prep, then per-class or persistent column passes): every size class
of one call gives the screens of its own pt_generate_groups call bit for bit (src/pyturb/atmosphere.py:1191-1219 and fourier.py:233-282 once
per size), with and without subharmonics and per-screen scales, for the
sizes the persistent kernels cover (128..2048) in any order and for odd
screen counts; the full config-sized ensemble (2048 x 256^2 + 1536 x 512^2 +
512 x 1024^2, two outer-scale groups) equals the per-class calls too; sizes
outside the kernels' range are refused on the host."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _class(n, count, G, P, seed, dev, scaled):
    import torch

    from pyturb_amd import ops
    from pyturb_amd.spectrum import build_filters

    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    noise = torch.randn((G, 2, (count + 1) // 2, n, n), generator=gen, device=dev)
    shn = torch.randn((G, P, 2, count, 3, 3), generator=gen, device=dev) if P else None
    filts = [build_filters(n, 8.0 / n, 1.0, 25.0 * (g + 1), P) for g in range(G)]
    amp = torch.tensor(np.stack([f[0] for f in filts]), dtype=torch.float32, device=dev)
    sha = torch.tensor(np.stack([f[2] for f in filts]), dtype=torch.float64, device=dev) if P else None
    scale = None
    if scaled:
        scale = torch.rand((count,), generator=gen, device=dev) + 0.5
    return ops.EnsembleClass(noise, shn, amp, sha, count, screen_scale=scale)


def _per_class(c, scale):
    from pyturb_amd import ops

    return ops.generate_groups(c.noise, c.sh_noise, c.amplitude, c.sh_amp, c.count,
                               screen_scale=c.screen_scale, scale=scale)


@pytest.mark.parametrize("persistent_cols", [False, True])
@pytest.mark.parametrize("sizes,G,P,scaled", [
    ([(256, 7), (128, 20), (1024, 3), (512, 16)], 2, 8, True),
    ([(2048, 2), (128, 1)], 1, 3, False),
    ([(512, 33), (256, 18)], 3, 0, True),
])
def test_ensemble_equals_per_class_calls(hip, sizes, G, P, scaled, persistent_cols):
    import torch

    from pyturb_amd import ops

    dev = torch.device("cuda", 0)
    print(f"ensemble {sizes} groups={G} levels={P} persistent_cols={persistent_cols}", flush=True)
    classes = [_class(n, c, G, P, 10 + k, dev, scaled) for k, (n, c) in enumerate(sizes)]
    scale = 0.75
    got = ops.generate_ensemble(classes, scale=scale, persistent_cols=persistent_cols)
    for (n, count), c, out in zip(sizes, classes, got):
        ref = _per_class(c, scale)
        assert out.shape == (count, n, n)
        assert torch.isfinite(out).all()
        assert torch.equal(out, ref), (n, count)
    # a second call on the same classes: same screens
    again = ops.generate_ensemble(classes, scale=scale, persistent_cols=persistent_cols)
    assert all(torch.equal(a, b) for a, b in zip(got, again))


def test_config_sized_ensemble_equals_per_class_calls(hip):
    import torch

    from pyturb_amd import ops

    dev = torch.device("cuda", 0)
    print("ensemble 2048 x 256^2 + 1536 x 512^2 + 512 x 1024^2, 2 groups, 8 levels", flush=True)
    sizes = [(256, 2048), (512, 1536), (1024, 512)]
    classes = [_class(n, c, 2, 8, 4 + k, dev, True) for k, (n, c) in enumerate(sizes)]
    for persistent_cols in (False, True):
        got = ops.generate_ensemble(classes, persistent_cols=persistent_cols)
        for (n, count), c, out in zip(sizes, classes, got):
            assert torch.equal(out, _per_class(c, 1.0)), (n, persistent_cols)


def test_ensemble_refuses_sizes_outside_the_kernels_range(hip):
    import torch

    from pyturb_amd import ops

    dev = torch.device("cuda", 0)
    bad = [_class(64, 4, 1, 2, 1, dev, False)]
    with pytest.raises(NotImplementedError, match="n=64"):
        ops.generate_ensemble(bad)
    with pytest.raises(ValueError, match="1..8"):
        ops.generate_ensemble([])
