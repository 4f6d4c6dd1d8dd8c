"""Autotuner for the launch shapes of the fused FFT passes.

Three knobs, set per power-of-two grid size n (32..2048):

* ``gen_rows``: threads per workgroup of the row pass of ``ops.generate*``
  (128, 256 or 512; 8 T / n rows per workgroup) and the FFT plan's largest
  radix (8, or 4 at 256 and 512 threads);
* ``gen_cols``: threads per workgroup (256, 512 or 1024; a tile of
  8 T / n adjacent columns), workgroup order (XCD-contiguous or flat) and
  largest radix (8, or 4 at 1024 threads) of the column pass shared by
  ``ops.generate*`` and the row-first frame engine's pass B;
* ``frames_order``: the transform order of the fused frame engine
  (``ops.spectral_frames``): column-first (variant 0: layer sum and column
  transform in 16-column tiles, then rows with the epilogue, whole rows
  written) or row-first (variant 1: layer sum and row transform, then
  column tiles with the epilogue).

``autotune`` times every distinct shape of each knob on a synthetic
workload of the given size with HIP events on the launch stream, checks
the output against the default's (bit for bit for the launch shapes, which
run the same per-element arithmetic; within 1e-6 relative RMS for the
radix-4 plans and the two transform orders, which round differently), keeps
the fastest and, given a path,
stores the choice in a JSON cache keyed by the device (architecture and
compute-unit count).  ``load_cache`` applies a stored choice; setting the
environment variable ``PYTURB_AMD_TUNE_CACHE`` to such a file applies it
when the library is loaded.  Without tuning every size runs its built-in
shape (variant 0), the measured best on MI355X for the headline sizes.
"""

from __future__ import annotations

import ctypes
import json
import os
import warnings
from typing import Dict, Iterable, List, Optional

import numpy as np

from . import _hip

KNOBS = {"gen_rows": _hip.PT_TUNE_GEN_ROWS, "gen_cols": _hip.PT_TUNE_GEN_COLS,
         "frames_order": _hip.PT_TUNE_FRAMES}


def _knob(name: str) -> int:
    if name not in KNOBS:
        raise ValueError(f"unknown knob {name!r}; expected one of {sorted(KNOBS)}")
    return KNOBS[name]


def variants(knob: str, n: int) -> List[Dict[str, int]]:
    """Distinct launch shapes of ``knob`` at size ``n``: one dict per shape
    (``variant``, ``threads``, ``tile``, ``xcd``, ``radix``), lowest variant
    first."""
    lib = _hip.lib()
    k = _knob(knob)
    count = lib.pt_tune_variants(k)
    seen, out = set(), []
    for v in range(count):
        t, m, x, r = _info(k, n, v)
        key = (t, m, x, r)
        if key in seen:
            continue
        seen.add(key)
        out.append({"variant": v, "threads": t, "tile": m, "xcd": x, "radix": r})
    return out


def _info(k: int, n: int, v: int):
    t, m, x, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _hip.check(_hip.lib().pt_tune_info(k, int(n), int(v), ctypes.byref(t), ctypes.byref(m),
                                       ctypes.byref(x), ctypes.byref(r)), "pt_tune_info")
    return t.value, m.value, x.value, r.value


def describe(knob: str, n: int, variant: Optional[int] = None) -> str:
    """Readable shape of a variant (the selected one by default)."""
    v = get(knob, n) if variant is None else int(variant)
    t, m, x, r = _info(_knob(knob), n, v)
    if knob == "frames_order":
        return ("column-first" if v == 0 else "row-first") + f", pass A {t} threads"
    unit = "rows" if knob == "gen_rows" else "cols"
    order = "" if knob == "gen_rows" else (" xcd" if x else " flat")
    return f"{t} threads x {m} {unit}{order}, radix {r}"


def get(knob: str, n: int) -> int:
    """Variant selected for ``knob`` at size ``n``."""
    v = _hip.lib().pt_tune_get(_knob(knob), int(n))
    if v < 0:
        _hip.check(-v, "pt_tune_get")
    return v


def set_variant(knob: str, n: int, variant: int) -> None:
    """Select ``variant`` of ``knob`` for size ``n`` (0 = built-in shape)."""
    _hip.check(_hip.lib().pt_tune_set(_knob(knob), int(n), int(variant)), "pt_tune_set")


def reset(sizes: Iterable[int] = (32, 64, 128, 256, 512, 1024, 2048)) -> None:
    """Return every knob to its built-in shape."""
    for n in sizes:
        for knob in KNOBS:
            set_variant(knob, n, 0)


def device_key(device=None) -> str:
    """Cache key of a device: architecture and compute-unit count."""
    t = _hip.torch()
    p = t.cuda.get_device_properties(device if device is not None else t.cuda.current_device())
    return f"{getattr(p, 'gcnArchName', p.name)}:{p.multi_processor_count}CU"


def _ensemble(n: int, count: int, groups: int, levels: int, seed: int, dev):
    """Synthetic mixed-outer-scale ensemble of ``count`` n x n screens."""
    from . import ops
    from .spectrum import build_filters

    t = _hip.torch()
    gen = t.Generator(device=dev)
    gen.manual_seed(seed)
    amps, shs = [], []
    for g in range(groups):
        amp, _, sh = build_filters(n, 8.0 / n, 0.15, 25.0 * (g + 1), levels)
        amps.append(t.tensor(amp, dtype=t.float32, device=dev))
        shs.append(t.tensor(sh, dtype=t.float64, device=dev))
    noise = t.randn((groups, 2, (count + 1) // 2, n, n), generator=gen, device=dev)
    shn = t.randn((groups, levels, 2, count, 3, 3), generator=gen, device=dev) if levels else None
    amp = t.stack(amps)
    sh_amp = t.stack(shs) if levels else None
    out = t.empty((count, n, n), dtype=t.float32, device=dev)
    work = t.empty((ops.work_elems(0, n, count, levels),), dtype=t.complex64, device=dev)

    def run():
        ops.generate_groups(noise, shn, amp, sh_amp, count, out=out, work=work)

    return run, out


def _frames(n: int, frames: int, layers: int, levels: int, seed: int, dev):
    """Synthetic frame-engine call: random layer spectra, subharmonic
    coefficients and displacements (up to a quarter period)."""
    from . import ops

    t = _hip.torch()
    rng = np.random.default_rng(seed)
    spec = (rng.standard_normal((layers, n, n)) + 1j * rng.standard_normal((layers, n, n))) / n
    spectra = t.tensor(spec.astype(np.complex64), device=dev)
    disp = t.tensor(rng.uniform(0.0, 0.25 * n * 0.01, (frames, layers, 2)), dtype=t.float64,
                    device=dev)
    coeffs = None
    if levels:
        cf = rng.standard_normal((levels, layers, 3, 3)) + 1j * rng.standard_normal((levels, layers, 3, 3))
        coeffs = t.tensor(cf * 1e-2, dtype=t.complex128, device=dev)
    out = t.empty((frames, n, n), dtype=t.float32, device=dev)
    work = t.empty((ops.frames_work_elems(n, frames, layers, levels),), dtype=t.complex64, device=dev)

    def run():
        ops.spectral_frames(spectra, disp, coeffs, 0.01, out=out, work=work)

    return run, out


def _time_ms(run, repeats: int) -> float:
    t = _hip.torch()
    run()  # first launch of a shape loads its code object
    evs = [(t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True))
           for _ in range(repeats)]
    for a, b in evs:
        a.record()
        run()
        b.record()
    t.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) for a, b in evs]))


def autotune(sizes: Iterable[int] = (256, 512, 1024), *, screens: Optional[int] = None,
             groups: int = 2, levels: int = 8, repeats: int = 5, seed: int = 0,
             cache: Optional[str] = None, verbose: bool = False) -> Dict[int, dict]:
    """Pick the fastest launch shape of each knob for each size.

    The workload per size is ``screens`` screens (default: about 2^26
    pixels, at least 16) of ``groups`` outer-scale groups with ``levels``
    subharmonic levels, and as many frames of a 9-layer atmosphere.  The
    knobs are tuned one after the other (rows with the default column
    shape, then columns with the chosen rows, then the frame engine's
    order).  A variant whose output differs from the default's is reported
    and skipped.  The choices stay selected; returns ``{n: {knob: variant,
    "ms": {knob: {variant: ms}}}}``.
    """
    t = _hip.torch()
    dev = _hip.device()
    result: Dict[int, dict] = {}

    def tune_knob(knob, n, run, out, exact, entry):
        set_variant(knob, n, 0)
        run()
        ref = out.clone()
        times: Dict[int, float] = {}
        for var in variants(knob, n):
            set_variant(knob, n, var["variant"])
            try:
                ms = _time_ms(run, repeats)
            except (ValueError, NotImplementedError, _hip.HipError) as exc:
                warnings.warn(f"autotune: {knob} variant {var} at n={n} failed ({exc}); skipped",
                              RuntimeWarning, stacklevel=3)
                continue
            if exact and var.get("radix", 8) == 8:
                ok = bool(t.equal(out, ref))
            else:
                ok = float(((out.double() - ref.double()).norm() / ref.double().norm()).item()) < 1e-6
            if not ok:
                warnings.warn(f"autotune: {knob} variant {var} at n={n} changed the output; "
                              "skipped", RuntimeWarning, stacklevel=3)
                continue
            times[var["variant"]] = ms
            if verbose:
                print(f"autotune n={n} {knob} {var}: {ms:.3f} ms")
        best = min(times, key=times.get) if times else 0
        set_variant(knob, n, best)
        entry[knob] = best
        entry["ms"][knob] = {str(v): round(ms, 4) for v, ms in times.items()}

    for n in sizes:
        n = int(n)
        count = int(screens) if screens else max(16, (1 << 26) // (n * n))
        for knob in KNOBS:
            set_variant(knob, n, 0)
        entry: dict = {"ms": {}}
        run, out = _ensemble(n, count, groups, levels, seed, dev)
        tune_knob("gen_rows", n, run, out, True, entry)
        tune_knob("gen_cols", n, run, out, True, entry)
        del run, out
        run, out = _frames(n, count, 9, min(levels, 8), seed, dev)
        tune_knob("frames_order", n, run, out, False, entry)
        del run, out
        result[n] = entry
    if cache:
        save_cache(cache, result)
    return result


def save_cache(path: str, result: Dict[int, dict]) -> None:
    """Merge ``autotune``'s choices into the JSON cache at ``path`` under
    this device's key."""
    data = {}
    if os.path.exists(path):
        with open(path) as fh:
            data = json.load(fh)
    dk = device_key()
    ent = data.setdefault(dk, {})
    for n, e in result.items():
        ent[str(int(n))] = {k: int(e[k]) for k in KNOBS}
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def load_cache(path: str) -> Dict[int, dict]:
    """Apply the choices stored for this device in ``path``; returns them
    (empty when the file holds none for this device)."""
    with open(path) as fh:
        data = json.load(fh)
    ent = data.get(device_key(), {})
    applied = {}
    for n, choice in ent.items():
        for knob, v in choice.items():
            if knob in KNOBS:
                set_variant(knob, int(n), int(v))
        applied[int(n)] = {k: int(v) for k, v in choice.items() if k in KNOBS}
    return applied
