"""Save and load phase/OPD screens with their metadata (src/pyturb/io.py:39-145).

Device tensors are copied to the host once (one DtoH copy of the contiguous
tensor); the file formats are the reference's, so files written here load in
pyturb and back:

* ``.npz``: the array under ``data`` and the metadata as a JSON string under
  ``__pyturb_meta__`` (io.py:71-75, 86-90);
* ``.fits``/``.fit``: a primary HDU with ``BUNIT`` from ``units`` and one
  header card per metadata entry (io.py:112-128), through astropy when it is
  installed, else ImportError as in the reference (io.py:94-104).
"""

from __future__ import annotations

import json
import os
from typing import Any, Dict, Tuple, Union

import numpy as np

from .utils import to_numpy

__all__ = ["save", "load"]

_META_KEY = "__pyturb_meta__"


def _version() -> str:
    from . import __version__

    return __version__


def save(path: Union[str, os.PathLike], data: Any, **metadata: Any) -> None:
    """Write ``data`` (a phase/OPD array or device tensor) with ``metadata``;
    ``None`` values are dropped and ``pyturb_version`` is added."""
    array = np.ascontiguousarray(to_numpy(data))
    meta = {str(k): v for k, v in metadata.items() if v is not None}
    meta.setdefault("pyturb_version", _version())
    path = str(path)
    if path.lower().endswith((".fits", ".fit")):
        _save_fits(path, array, meta)
    else:
        np.savez(path, data=array, **{_META_KEY: json.dumps(meta)})


def load(path: Union[str, os.PathLike]) -> Tuple[np.ndarray, Dict[str, Any]]:
    """The host array and metadata dictionary written by :func:`save`."""
    path = str(path)
    if path.lower().endswith((".fits", ".fit")):
        return _load_fits(path)
    with np.load(path, allow_pickle=False) as handle:
        data = handle["data"]
        meta = json.loads(str(handle[_META_KEY])) if _META_KEY in handle else {}
    return data, meta


def _astropy_fits():
    try:
        from astropy.io import fits
    except ImportError as exc:
        raise ImportError("FITS support needs astropy; use a .npz path instead.") from exc
    return fits


def _save_fits(path: str, array: np.ndarray, meta: Dict[str, Any]) -> None:
    import warnings

    fits = _astropy_fits()
    from astropy.io.fits.verify import VerifyWarning

    hdu = fits.PrimaryHDU(data=array)
    hdu.header["BUNIT"] = str(meta.get("units", "radians"))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", VerifyWarning)  # long keys become HIERARCH cards
        for key, value in meta.items():
            hdu.header[key.upper()] = (value if isinstance(value, (int, float, bool, str))
                                       else json.dumps(value))
        hdu.header.add_comment("Written by pyturb")
        hdu.writeto(path, overwrite=True)


def _load_fits(path: str) -> Tuple[np.ndarray, Dict[str, Any]]:
    fits = _astropy_fits()
    skip = {"SIMPLE", "BITPIX", "EXTEND", "COMMENT", "HISTORY", "BUNIT", "NAXIS"}
    skip |= {f"NAXIS{i}" for i in range(1, 10)}
    with fits.open(path) as hdul:
        hdu = hdul[0]
        data = np.asarray(hdu.data)
        data = data.astype(data.dtype.newbyteorder("="), copy=False)  # FITS is big-endian
        meta = {key.lower(): value for key, value in hdu.header.items()
                if key and key not in skip}
    return data, meta
