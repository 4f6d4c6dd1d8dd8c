"""Saving device results (tests/test_io.py:17-32 of the reference, on the
device path): one DtoH copy, exact values and the atmosphere's metadata."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_save_load_device_opd(hip, tmp_path):
    import pyturb_amd as pa

    atm = pa.Atmosphere.from_profile("two-layer", seeing=0.8, n=48, seed=1, device="mi355x")
    opd = atm.opd(0.0)
    path = tmp_path / "screen.npz"
    pa.save(path, opd, **atm.metadata)
    data, meta = pa.load(path)
    np.testing.assert_array_equal(data, opd.cpu().numpy())
    assert meta["units"] == "metres"
    assert meta["r0"] == pytest.approx(atm.r0, rel=1e-4)
    assert meta["pixel_scale"] == pytest.approx(atm.pixel_scale, rel=1e-6)
