"""Screen I/O (pyturb_amd.io; the reference's tests/test_io.py:14-46 on host
arrays): the reference's file formats, metadata rules and FITS gating."""

from __future__ import annotations

import json

import numpy as np
import pytest

import pyturb_amd as pa


def test_npz_roundtrip_and_none_dropped(tmp_path):
    screens = np.random.default_rng(0).standard_normal((4, 32, 32)).astype(np.float32)
    path = tmp_path / "cube.npz"
    pa.save(path, screens, pixel_scale=0.02, r0=0.15, seed=None, units="radians")
    data, meta = pa.load(path)
    np.testing.assert_array_equal(data, screens)
    assert data.dtype == np.float32 and data.shape == (4, 32, 32)
    assert "seed" not in meta
    assert meta["r0"] == 0.15 and meta["units"] == "radians"
    assert meta["pyturb_version"] == pa.__version__


def test_npz_layout_is_the_references(tmp_path):
    """``data`` plus a JSON string under ``__pyturb_meta__`` (io.py:71-75),
    so a file written by either package loads in the other."""
    path = tmp_path / "x.npz"
    pa.save(path, np.arange(6.0).reshape(2, 3), pixel_scale=0.5)
    with np.load(path, allow_pickle=False) as h:
        assert sorted(h.files) == ["__pyturb_meta__", "data"]
        assert json.loads(str(h["__pyturb_meta__"]))["pixel_scale"] == 0.5
    other = tmp_path / "y.npz"
    np.savez(other, data=np.ones(3), __pyturb_meta__=json.dumps({"r0": 0.2}))
    data, meta = pa.load(other)
    assert meta == {"r0": 0.2} and np.array_equal(data, np.ones(3))


def test_fits_needs_astropy(tmp_path):
    try:
        import astropy  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError):
            pa.save(tmp_path / "s.fits", np.zeros((2, 2)))
        return
    pa.save(tmp_path / "s.fits", np.ones((3, 3), np.float32), units="metres", r0=0.1)
    data, meta = pa.load(tmp_path / "s.fits")
    assert data.dtype.byteorder in ("=", "|") and meta["units"] == "metres"
