"""ctypes binding of ``libpyturb_hip.so`` and device-array plumbing.

The library is the only compute path of this package: every hot call goes
through one of the ``pt_*`` C entry points declared in
``include/pyturb_hip.h``.  If the shared object is missing or no ROCm device
is visible, the calls raise -- there is no CPU fallback.

PyTorch supplies device memory and the current HIP stream; tensors cross the
boundary as raw device pointers.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyturb_hip.so")

PT_OK = 0
PT_ERR_ARG = 1
PT_ERR_HIP = 2
PT_ERR_UNSUPPORTED = 3
PT_TUNE_GEN_ROWS = 0
PT_TUNE_GEN_COLS = 1
PT_TUNE_FRAMES = 2
ABI_VERSION = 6  # PT_ABI_VERSION of include/pyturb_hip.h

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_f = ctypes.c_float

# name -> (restype, argtypes); mirrors include/pyturb_hip.h
SIGNATURES = {
    "pt_last_error": (ctypes.c_char_p, []),
    "pt_abi_version": (_i, []),
    "pt_work_elems": (ctypes.c_longlong, [_i, _i, _i, _i]),
    "pt_frames_work_elems": (ctypes.c_longlong, [_i, _i, _i, _i]),
    "pt_fft2": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "pt_generate": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _f, _vp, _i, _vp, _vp, _vp]),
    "pt_generate_groups": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _f, _vp, _i, _vp, _vp,
                                _vp]),
    "pt_generate_ensemble": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _f, _vp, _vp, _i,
                                  _vp]),
    "pt_spectral_frames": (_i, [_vp, _i, _i, _d, _vp, _i, _vp, _i, _i, _i, _f, _vp, _vp, _vp]),
    "pt_spectral_frames_noalloc": (_i, [_vp, _i, _i, _d, _vp, _i, _vp, _i, _i, _i, _f, _vp, _vp]),
    "pt_spectral_rows": (_i, [_vp, _i, _i, _d, _vp, _i, _vp, _vp]),
    "pt_spectral_cols": (_i, [_vp, _i, _i, _d, _vp, _i, _vp, _i, _i, _i, _f, _vp, _vp]),
    "pt_spectral_fold": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "pt_spectral_rows_folded": (_i, [_i, _i, _d, _vp, _i, _vp, _vp]),
    "pt_spectral_frames_folded": (_i, [_i, _i, _d, _vp, _i, _vp, _i, _i, _i, _f, _vp, _vp, _vp]),
    "pt_innovation_gemm": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "pt_extrude_rows": (_i, [_vp, _i, _i, _i, _vp, _vp, ctypes.c_longlong, _i, _vp, _vp, _vp,
                             _vp, _vp, _vp]),
    "pt_extrude_rows_far": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, ctypes.c_longlong, _i, _vp, _vp,
                                 _vp, _vp, _vp, _vp, _vp]),
    "pt_extrude_step": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _vp, ctypes.c_longlong, _i, _vp, _vp, _vp,
                             _vp, _vp, _vp, _vp, ctypes.c_uint, _vp]),
    "pt_readout": (_i, [_vp, _i, ctypes.c_longlong, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _f,
                        _i, _vp, _vp]),
    "pt_readout_adjoint_work_elems": (ctypes.c_longlong, [_i]),
    "pt_readout_adjoint": (_i, [_vp, _i, ctypes.c_longlong, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _i,
                                _f, _vp, _vp, _vp, _vp]),
    "pt_layer_spectra": (_i, [_vp, _vp, _vp, _vp, _i, _i, _vp, _vp]),
    "pt_adjoint_work_elems": (ctypes.c_longlong, [_i, _i, _i]),
    "pt_spectral_frames_adjoint": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _d, _vp, _i, _vp, _i, _i,
                                        _vp, _vp, _i, _d, _vp, _vp, _vp]),
    "pt_lgs_zoom": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _i, _f, _vp, _vp]),
    "pt_normals": (_i, [_vp, ctypes.c_longlong, ctypes.c_ulonglong, ctypes.c_ulonglong, _f, _vp]),
    "pt_rows_blend": (_i, [_vp, ctypes.c_longlong, _i, _i, _vp, ctypes.c_longlong, _i, _i, _f, _f,
                           _vp]),
    "pt_boil_work_elems": (ctypes.c_longlong, [_i, _i, _i, _i]),
    "pt_boiled_frames": (_i, [_vp, _vp, _vp, _vp, _i, _i, _d, _vp, _i, _i, _d, _vp, _vp, _i, _vp,
                              _vp, ctypes.c_ulonglong, ctypes.c_ulonglong, _i, _i, _f, _vp, _vp,
                              _vp]),
    "pt_tune_variants": (_i, [_i]),
    "pt_tune_info": (_i, [_i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                          ctypes.POINTER(_i)]),
    "pt_tune_set": (_i, [_i, _i, _i]),
    "pt_tune_get": (_i, [_i, _i]),
}

_lib: Optional[ctypes.CDLL] = None


class HipError(RuntimeError):
    """A HIP-library call failed (launch error, unsupported size, ...)."""


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Open the shared library and bind every exported entry point."""
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -m pyturb_amd.build` "
            "(hipcc, --offload-arch=gfx950). This package has no CPU fallback."
        )
    so = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(so, name)
        fn.restype = res
        fn.argtypes = args
    if so.pt_abi_version() != ABI_VERSION:
        raise ImportError(f"{path} has ABI version {so.pt_abi_version()}, this package needs "
                          f"{ABI_VERSION}: rebuild it with `python -m pyturb_amd.build`")
    return so


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
        cache = os.environ.get("PYTURB_AMD_TUNE_CACHE")
        if cache and os.path.exists(cache):  # launch shapes chosen by pyturb_amd.tune
            from . import tune

            tune.load_cache(cache)
    return _lib


def check(rc: int, what: str) -> None:
    if rc == PT_OK:
        return
    msg = lib().pt_last_error().decode(errors="replace")
    if rc == PT_ERR_ARG:
        raise ValueError(f"{what}: {msg}")
    if rc == PT_ERR_UNSUPPORTED:
        raise NotImplementedError(f"{what}: {msg}")
    raise HipError(f"{what}: {msg}")


# ---------------------------------------------------------------------------
# torch plumbing
# ---------------------------------------------------------------------------
def torch():
    import torch as _t

    return _t


def device(index: int = 0):
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError(
            "no ROCm GPU is visible: the MI355X backend needs a gfx950 device "
            "(this package has no CPU compute path)"
        )
    return t.device("cuda", index)


def stream_ptr() -> int:
    """The current stream of the current device (torch's raw-stream query,
    a fraction of the cost of building a torch.cuda.Stream object)."""
    C = torch()._C
    return C._cuda_getCurrentRawStream(C._cuda_getDevice())


def ptr(tensor) -> int:
    """Device pointer of a contiguous device tensor (None -> NULL)."""
    if tensor is None:
        return None
    if not tensor.is_cuda:
        raise ValueError("expected a device tensor")
    if not tensor.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return tensor.data_ptr()


def to_device(array, dtype=None, dev=None):
    """Host array -> contiguous device tensor (one H2D copy)."""
    t = torch()
    a = np.ascontiguousarray(array if dtype is None else np.asarray(array, dtype=dtype))
    return t.from_numpy(a).to(dev or device(), non_blocking=False)


def empty(shape, dtype, dev=None):
    t = torch()
    return t.empty(tuple(int(s) for s in shape), dtype=dtype, device=dev or device())
