"""Tile-shape autotuner with a JSON cache.

The generated kernels stage a tile plus halo in LDS; the best tile depends on the grid sizes,
the halo and the number of variables.  ``tune_batch`` times one batched right-hand side per
candidate tile on the device and keeps the fastest; results are cached in a JSON file keyed by
the batch's structure and sizes (``$MOL_TUNE_CACHE``, default ``~/.cache/molgpu/tune.json``;
a cache that cannot be written is skipped).
"""
from __future__ import annotations

import hashlib
import json
import os
from typing import Dict, List, Optional, Sequence, Tuple

from .discretization import Lowered

CANDIDATES = {
    1: [(256, 1, 1), (128, 1, 1), (512, 1, 1)],
    # 2D: tiles of whole rows of the instances (256 wide; a tile is trimmed to its instance)
    # first, then square-ish tiles
    2: [(256, 16, 1), (256, 12, 1), (256, 8, 1), (128, 16, 1), (64, 16, 1), (32, 16, 1)],
    3: [(32, 4, 2), (16, 8, 2), (16, 4, 4), (32, 8, 1)],
}


def _cache_path() -> str:
    return os.environ.get("MOL_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "molgpu", "tune.json"))


def _load(path: str) -> Dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, table: Dict) -> None:
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def batch_key(lows: Sequence[Lowered]) -> str:
    """Structure and sizes of a batch (instances' grids, boxes, operators, precision)."""
    desc = [(tuple(l.n), [tuple(x) for x in l.lo], [tuple(x) for x in l.hi], l.dtype,
             [(s["kind"], s["axis"]) for s in l.slots], len(l.closures)) for l in lows]
    return hashlib.sha1(json.dumps(desc).encode()).hexdigest()


def time_batch_rhs(batch, reps: int = 20) -> float:
    """Average microseconds of one batched right-hand side (back-to-back launches)."""
    import torch
    u = torch.tensor(batch.u0, dtype=batch.torch_dtype(), device="cuda")
    du = torch.empty_like(u)
    for _ in range(3):
        batch.f(du, u, 0.0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        batch.f(du, u, 0.0)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3 / reps


def tune_batch(lows: Sequence[Lowered], candidates: Optional[List[Tuple[int, int, int]]] = None,
               quick: bool = False, use_cache: bool = True) -> Tuple[Tuple[int, int, int], Dict[str, float]]:
    """Fastest tile for a batch, and the timing table {"tx,ty,tz": microseconds}."""
    from .batch import BatchProblem
    key = batch_key(lows)
    path = _cache_path()
    cache = _load(path) if use_cache else {}
    if key in cache:
        ent = cache[key]
        return tuple(ent["tile"]), ent["table"]
    cands = candidates or CANDIDATES[lows[0].ndim]
    if quick:
        cands = cands[:2]
    table: Dict[str, float] = {}
    best, best_us = None, float("inf")
    for tile in cands:
        b = BatchProblem(lows, tile=tile)
        us = time_batch_rhs(b)
        table["%d,%d,%d" % tuple(tile)] = us
        if us < best_us:
            best, best_us = tuple(tile), us
        del b
    if use_cache:
        cache[key] = {"tile": list(best), "table": table}
        _store(path, cache)
    return best, table
