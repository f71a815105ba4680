"""Autotuner of the patch kernels.

The one launch-shape parameter of the patch assembly is the number of cells
per patch workgroup (64, 128 or 256, capped at each kernel's largest): fewer
cells per patch mean smaller LDS footprints and more resident workgroups but
more nodes shared between patches (summed by the fix-up pass). `tune` times
each candidate on the caller's own mesh and state with HIP events
(PDRD_OPT_TIMING: patch kernels plus fix-up per iteration) and returns the
fastest; results are cached in a JSON file under a key made of the physics,
scheme, precision, element types and a power-of-two size class of the mesh,
so a mesh of the same kind and size reuses the choice.
"""
from __future__ import annotations

import json
import math
import os

from . import _lib as L

CANDIDATES = (64, 128, 256)
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "pdekit_amd", "patch_cells.json")


def cache_key(mesh, physics: int, scheme: int, precision: int, quad_order) -> str:
    kinds = []
    for b in mesh.blocks:
        q = quad_order[b.shape] if isinstance(quad_order, dict) else quad_order
        kinds.append(f"s{b.shape}p{mesh.order}q{q}")
    size = int(math.log2(max(1, mesh.nb_cells)))
    return f"ph{physics}/sc{scheme}/pr{precision}/{'+'.join(sorted(kinds))}/e2^{size}"


def load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store_cache(path: str, key: str, entry: dict) -> None:
    data = load_cache(path)
    data[key] = entry
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def measure(mesh, physics, scheme, params, quad_order, state, cfl, cells: int, precision: int = 0,
            iters: int = 10):
    """Mean time (ms) of one iteration with `cells` cells per patch, and the
    cells per patch actually used."""
    from .rdm import PGRDMethodExplicit

    m = PGRDMethodExplicit(physics, scheme, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_order, patch_cells=cells)
    try:
        if m.get_option(L.OPT_SCATTER) != 2:
            return math.inf, cells
        used = m.get_option(L.OPT_PATCH_CELLS)
        m.set_option(L.OPT_PRECISION, precision)
        m.set_option(L.OPT_TIMING, 1)
        m.set_vec_function("solution", state)
        m.iterate(cfl, 2)  # code-object load, first-touch
        m.iterate(cfl, iters)
        c, n = m.last_timing()
        return c + n, used
    finally:
        m.destroy()


def tune(mesh, physics, scheme, params, quad_order, state, cfl, precision: int = 0, candidates=CANDIDATES,
         iters: int = 10, cache_path: str | None = DEFAULT_CACHE, refresh: bool = False) -> int:
    """Cells per patch for this kind of problem: from the cache, or measured."""
    key = cache_key(mesh, physics, scheme, precision, quad_order)
    if cache_path and not refresh:
        hit = load_cache(cache_path).get(key)
        if hit and int(hit.get("cells", 0)) in candidates:
            return int(hit["cells"])
    times = {}
    for c in candidates:
        t, used = measure(mesh, physics, scheme, params, quad_order, state, cfl, c, precision, iters)
        if used not in times or t < times[used]:
            times[used] = t
    best = min(times, key=times.get)
    if cache_path:
        store_cache(cache_path, key, {"cells": best, "ms_per_iteration": {str(k): v for k, v in times.items()},
                                      "cells_in_mesh": mesh.nb_cells})
    return best
