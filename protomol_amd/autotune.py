"""Launch-shape autotuner for device-resident MD.

Times short MD runs of the actual workload (the same systems, the same
force model) over a grid of launch shapes and returns the fastest:

  skin        pair-list buffer (A): a longer list per rebuild vs more rebuilds
  cell_scale  x-y column width (multiple of (8/density)^(1/3)): the shape of
              the 8-atom clusters, i.e. the cell tiles the pair kernel walks
  pair_split  waves sharing one super-cluster's pair list (work per wave
              against waves in flight)
  pair_waves  waves per workgroup of the pair kernel (1, 2, 4, 8)

Results are kept in a JSON cache keyed by a signature of the workload, so a
second run of the same workload skips the search.
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os
import time

import numpy as np

from . import _lib as L

DEFAULT_GRID = {
    "skin": (1.0, 1.3),
    "cell_scale": (1.0, 1.25),
    "pair_split": (8, 12),
    "pair_waves": (2, 4, 8),
}


def signature(force: L.ForceDesc, systems, grid) -> str:
    f = {name: getattr(force, name) for name, _ in L.ForceDesc._fields_ if name not in ("switch_a", "switch_b")}
    for sw in ("switch_a", "switch_b"):
        s = getattr(force, sw)
        f[sw] = [s.kind, s.switchon, s.cutoff, s.switchoff, s.order]
    sizes = sorted(int(s.n) for s in systems)
    boxes = sorted(round(float(s.box[0]), 2) for s in systems)
    blob = json.dumps({"force": f, "sizes": sizes, "boxes": boxes, "grid": grid}, sort_keys=True, default=str)
    return hashlib.sha1(blob.encode()).hexdigest()


def _time_md(force, systems, v0, steps, dt_fs, warm):
    with L.Context(force, systems) as ctx:
        ctx.md_set_state(np.concatenate([s.x for s in systems]), np.concatenate(v0))
        ctx.md_run(warm, dt_fs, energies=False)
        t0 = time.perf_counter()
        ctx.md_run(steps, dt_fs, energies=False)
        return (time.perf_counter() - t0) * 1e3 / steps


def tune(force: L.ForceDesc, systems, v0, dt_fs: float = 1.0, steps: int = 30, warm: int = 10, grid=None,
         cache_path: str | None = None, log=None):
    """Return (best parameter dict, list of (params, ms_per_step)) for MD of
    ``systems`` (one context, all systems batched) with start velocities v0."""
    grid = dict(grid or DEFAULT_GRID)
    key = signature(force, systems, grid)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cache = json.load(fh)
        if key in cache:
            return cache[key]["best"], cache[key]["table"]
    names = list(grid)
    table = []
    for values in itertools.product(*(grid[n] for n in names)):
        params = dict(zip(names, values))
        f = apply(force, params)
        try:
            ms = _time_md(f, systems, v0, steps, dt_fs, warm)
        except L.PMError as err:  # e.g. a skin too long for the smallest box
            ms = float("inf")
            if log:
                log(f"autotune {params}: {err}")
        table.append((params, ms))
        if log:
            log(f"autotune {params}: {ms:.3f} ms/step")
    best = min(table, key=lambda t: t[1])[0]
    if cache_path:
        cache[key] = {"best": best, "table": table}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as fh:
            json.dump(cache, fh, indent=1)
    return best, table


def apply(force: L.ForceDesc, params: dict) -> L.ForceDesc:
    f = L.ForceDesc.from_buffer_copy(force)
    for n, v in params.items():
        setattr(f, n, v)
    return f
