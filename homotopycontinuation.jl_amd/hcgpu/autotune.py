"""Autotuner of the tracker's layout, build and launch shape.

For a homotopy and a representative sample of its batch, times every candidate and keeps the
fastest.  A candidate is a layout (one path per lane, or one path per 16-lane group for
systems of up to 16 variables, per 32-lane group up to 24), the waves per SIMD the per-lane
kernel is compiled for (its register budget), and the workgroups per CU of the persistent
launch.  Choices are cached in a
JSON file next to libhcgpu.so, keyed by the program's generated source and the device, so a
workload is tuned once; GpuContext.program() and track_batch() then use the cached choice by
default.  Tuning never changes results: every candidate runs the same arithmetic (both layouts
equal the oracle bit for bit), and the tuner checks that the sample's return codes and
endpoints agree across candidates.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import abi
from .gpu import GpuContext, track_batch

DEFAULT_CANDIDATES: Tuple[Tuple[int, int], ...] = ((1, 4), (1, 8), (2, 8), (2, 16), (4, 16), (4, 32), (8, 32))
GROUP_CANDIDATES: Tuple[int, ...] = (1, 2)  # workgroups per CU of the lane-group kernel
GROUP_MAX_N = 24


def candidates_for(n: int, lane=DEFAULT_CANDIDATES, group=GROUP_CANDIDATES) -> List[dict]:
    """Every (layout, waves_per_eu, blocks_per_cu) candidate for a system of n variables."""
    out = [dict(layout="lane", waves_per_eu=w, blocks_per_cu=b) for w, b in lane]
    if n <= GROUP_MAX_N:
        out += [dict(layout="group", waves_per_eu=0, blocks_per_cu=b) for b in group]
    return out


def _as_candidate(c) -> dict:
    if isinstance(c, dict):
        return dict(layout=c.get("layout", "lane"), waves_per_eu=int(c.get("waves_per_eu", 0)),
                    blocks_per_cu=int(c.get("blocks_per_cu", 0)))
    w, b = c
    return dict(layout="lane", waves_per_eu=int(w), blocks_per_cu=int(b))


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
         candidates: Optional[Iterable] = None, force: bool = False) -> dict:
    """Best {layout, waves_per_eu, blocks_per_cu} for tracking `starts` (a representative
    sample) with `homotopy`; cached unless force=True.  candidates: dicts as candidates_for()
    returns, or (waves_per_eu, blocks_per_cu) pairs of the per-lane layout; default: every
    candidate for the homotopy's size."""
    key = _key(ctx, [homotopy])
    if not force:
        hit = _load().get(key)
        if hit is not None:
            return _normalized(hit)
    opts = opts or abi.default_options()
    cands = [_as_candidate(c) for c in (candidates if candidates is not None else candidates_for(homotopy.n))]
    table = []
    ref = None
    for c in cands:
        prog = ctx.program(homotopy, waves_per_eu=c["waves_per_eu"], layout=c["layout"])
        o = abi.hcg_options.from_buffer_copy(opts)
        o.blocks_per_cu = c["blocks_per_cu"]
        track_batch(prog, starts, o, group_of_path=group_of_path)  # warm
        out = track_batch(prog, starts, o, group_of_path=group_of_path)
        ms = ctx.last_kernel_ms()
        if ref is None:
            ref = out
        elif not (np.array_equal(ref.code, out.code) and np.array_equal(ref.solution, out.solution, equal_nan=True)):
            raise RuntimeError(f"results differ between launch shapes ({c})")
        table.append(dict(c, kernel_ms=round(ms, 3)))
    best = min(table, key=lambda r: r["kernel_ms"])
    entry = {"layout": best["layout"], "waves_per_eu": best["waves_per_eu"], "blocks_per_cu": best["blocks_per_cu"],
             "sample_paths": len(starts), "table": table}
    _store(key, entry)
    ctx._tuned[id(homotopy)] = entry  # later ctx.program(homotopy) calls build the tuned choice
    return entry


def _normalized(entry: dict) -> dict:
    e = dict(entry)
    e.setdefault("layout", "lane")
    return e


def tuned_choice(ctx: GpuContext, homotopy) -> Optional[dict]:
    """The cached choice for `homotopy` on this device, if it was tuned."""
    hit = _load().get(_key(ctx, [homotopy]))
    return None if hit is None else _normalized(hit)
