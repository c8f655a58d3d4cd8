"""Autotuner of the step's launch shape.

Knobs:
  nsub        y sub-cells per unit cell (the cell height, 1/nsub): fewer
              candidate pairs per bead vs. more cells to count and scan.
              A physics-neutral knob in the reference too (its `nsub` key,
              cel/engine/oercel.c:264), so it is a per-replica parameter.
  pair_block  workgroup size of the pair kernel (64 / 128 / 256 lanes).
  tile_beads  target i-beads per pair tile (rows per tile follow from it).

The tuner times a few steps of a synthetic system of the requested size class
on the device for each candidate and keeps the fastest.  Results are cached per
size class (log2 N, density) in a JSON file next to this module, so a later
engine of the same class reuses them without timing.
"""
import json
import math
import os
import sys
import time

from . import engine as _engine
from .params import Params

# the cache file: OC_AUTOTUNE_CACHE, or a JSON file next to this module
CACHE = os.environ.get("OC_AUTOTUNE_CACHE") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                             "autotune_cache.json")
NSUB = (1, 2, 3, 4, 6)
BLOCKS = (64, 128, 256)
TILE_FACTORS = (0.6, 0.8, 1.2)


def size_class(N, rho=3.0):
    return f"log2N={int(round(math.log2(max(N, 2))))},rho={rho:g}"


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(c):
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(c, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def _time_steps(params_list, pair_block, tile_beads, steps):
    with _engine.Engine(params_list, pair_block=pair_block, tile_beads=tile_beads) as e:
        e.start()
        e.step(4)
        e.step_count()
        t0 = time.perf_counter()
        e.step(steps)
        return (time.perf_counter() - t0) / steps


def tune(params, steps=20, use_cache=True, verbose=False):
    """Best (nsub, pair_block, tile_beads) for the size class of `params` (a
    finalized Params of one replica).  Timing is done on that system."""
    key = size_class(params.N, params.rho)
    cache = _load() if use_cache else {}
    if key in cache:
        c = cache[key]
        return c["nsub"], c["pair_block"], c["tile_beads"]
    best = None
    # 1) cell height with the default launch shape
    for ns in NSUB:
        p = params.copy()
        p.nsub = ns
        t = _time_steps([p], 256, 0, steps)
        if verbose:
            print(f"  {key} nsub={ns}: {t * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
        if best is None or t < best[0]:
            best = (t, ns, 256, 0)
    # 2) pair workgroup and tile size at that cell height
    p = params.copy()
    p.nsub = best[1]
    for b in BLOCKS:
        for fct in TILE_FACTORS:
            tb = int(b * fct)
            t = _time_steps([p], b, tb, steps)
            if verbose:
                print(f"  {key} block={b} tile={tb}: {t * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
            if t < best[0]:
                best = (t, best[1], b, tb)
    t, ns, b, tb = best
    cache = _load()
    cache[key] = {"nsub": ns, "pair_block": b, "tile_beads": tb, "ms_per_step": t * 1e3, "N": params.N}
    _save(cache)
    return ns, b, tb


def tune_ensemble(params_list, steps=10, verbose=False):
    """Per-replica nsub from the tuned size classes (tuned on demand), then the
    pair launch shape for the whole batch.  Returns (params_list, block, tile)."""
    out = []
    classes = {}
    for p in params_list:
        k = size_class(p.N, p.rho)
        if k not in classes:
            classes[k] = tune(p, steps=steps, verbose=verbose)[0]
        q = p.copy()
        q.nsub = classes[k]
        out.append(q)
    best = None
    for b in BLOCKS:
        for fct in TILE_FACTORS:
            tb = int(b * fct)
            t = _time_steps(out, b, tb, steps)
            if verbose:
                print(f"  ensemble block={b} tile={tb}: {t * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
            if best is None or t < best[0]:
                best = (t, b, tb)
    return out, best[1], best[2]


__all__ = ["tune", "tune_ensemble", "size_class", "Params"]
