"""Autotuner of the tracker kernel's build and launch shape.

For a homotopy and a representative sample of its batch, times every candidate
(waves per SIMD the kernel is compiled for = its register budget, x workgroups per CU of
the persistent launch) and keeps the fastest.  Choices are cached in a JSON file next to
libhcgpu.so, keyed by the program's generated source and the device, so a workload is
tuned once.  Tuning never changes results: every candidate runs the same arithmetic; the
tuner checks that the sample's return codes agree across candidates.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
from pathlib import Path
from typing import Dict, Iterable, Optional, Sequence, Tuple

import numpy as np

from . import abi
from .gpu import GpuContext, track_batch

DEFAULT_CANDIDATES: Tuple[Tuple[int, int], ...] = ((1, 4), (1, 8), (2, 8), (2, 16), (4, 16), (4, 32), (8, 32))


def _cache_path() -> Path:
    return Path(__file__).resolve().parent.parent / "_build" / "autotune.json"


def _key(ctx: GpuContext, homotopies: Sequence) -> str:
    h = hashlib.sha256()
    L = abi.lib()
    for H in homotopies:
        need = ctypes.c_size_t()
        abi.check(L.hcg_generate(ctypes.byref(H.desc()), None, 0, ctypes.byref(need), None))
        buf = ctypes.create_string_buffer(need.value)
        abi.check(L.hcg_generate(ctypes.byref(H.desc()), buf, need.value, ctypes.byref(need), None))
        h.update(buf.value)
    return f"dev{ctx.device}:{h.hexdigest()[:24]}"


def _load() -> Dict[str, dict]:
    try:
        return json.loads(_cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store(key: str, entry: dict) -> None:
    data = _load()
    data[key] = entry
    p = _cache_path()
    p.parent.mkdir(parents=True, exist_ok=True)
    tmp = p.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    tmp.replace(p)


def tune(ctx: GpuContext, homotopy, starts, opts: Optional[abi.hcg_options] = None, group_of_path=None,
         candidates: Iterable[Tuple[int, int]] = DEFAULT_CANDIDATES, force: bool = False) -> dict:
    """Best {waves_per_eu, blocks_per_cu} for tracking `starts` (a representative sample)
    with `homotopy`; cached unless force=True."""
    key = _key(ctx, [homotopy])
    if not force:
        hit = _load().get(key)
        if hit is not None:
            return hit
    opts = opts or abi.default_options()
    table = []
    ref_codes = None
    for w, bpc in candidates:
        prog = ctx.program(homotopy, waves_per_eu=w)
        o = abi.hcg_options.from_buffer_copy(opts)
        o.blocks_per_cu = bpc
        track_batch(prog, starts, o, group_of_path=group_of_path)  # warm
        out = track_batch(prog, starts, o, group_of_path=group_of_path)
        ms = ctx.last_kernel_ms()
        if ref_codes is None:
            ref_codes = out.code.copy()
        elif not np.array_equal(ref_codes, out.code):
            raise RuntimeError(f"return codes differ between launch shapes (waves_per_eu={w}, blocks_per_cu={bpc})")
        table.append({"waves_per_eu": w, "blocks_per_cu": bpc, "kernel_ms": round(ms, 3)})
    best = min(table, key=lambda r: r["kernel_ms"])
    entry = {"waves_per_eu": best["waves_per_eu"], "blocks_per_cu": best["blocks_per_cu"], "sample_paths": len(starts),
             "table": table}
    _store(key, entry)
    return entry


def tuned_choice(ctx: GpuContext, homotopy) -> Optional[dict]:
    """The cached choice for `homotopy` on this device, if it was tuned."""
    return _load().get(_key(ctx, [homotopy]))
