"""Launch-shape autotuner for a batch arena: the tile, then the flux threads per cell once few
slots run.

The main launch-shape choice of the time step is the tile: threads per block of every cell,
face and wall kernel (64, 128 or 256; each slot's ranges are padded to it, so a block never
spans two meshes).  Small tiles spread a latency-bound arena over more compute units; large
tiles cut padding and block scheduling overhead once the arena fills the chip.  The tuner
builds the same arena at every candidate tile, times `trials` trial replays after `warmup`
and keeps the fastest, except that a larger tile within 1% of the fastest wins the tie: it
has fewer blocks, less padding and fewer faces evaluated by two tiles, and its batches end
sooner once few slots run (256 LF2 meshes, FP32: tiles 64 and 128 within 0.3% per full-batch
step, but 99.6 s against 97.0 s for the whole steady solve).  The choice changes only the block partition: per-slot results are
bitwise identical between runs at the same tile, and within the FP64/FP32 tolerances across
tiles (the monitor's block partial sums are the only tile-dependent reduction order).

The second choice is the few-slot launch shape (`Batch.set_launch_shape`): when few slots of a
closed batch are still running, each trial is bound by latency and two threads per cell in the
flux launches shorten it; with many running it is slower (fewer blocks per CU).  The tuner times
trials with k = 32, 16, 8 and 2 running slots at one and two threads per cell and keeps the
largest k at and below which two threads win every probe (0: never).  Results are bitwise
independent of both choices at a given tile.

Results are cached per (precision, slot count, total cells rounded to 1%, feature set) for
the life of the process.
"""
from __future__ import annotations

import time

from . import batch as B

TILES = (64, 128, 256)
TIE = 0.01  # relative step-time margin within which the larger tile is kept
_CACHE: dict = {}


def choose(timings: dict) -> int:
    """The tile to use from {tile: seconds per trial}: the largest within TIE of the fastest."""
    fastest = min(timings.values())
    return max(t for t, s in timings.items() if s <= fastest * (1 + TIE))


def _key(meshes, precision, features):
    cells = sum(int(m.n_cells) for m in meshes)
    return (precision, len(meshes), round(cells, -max(0, len(str(cells)) - 2)),
            tuple(getattr(features, k) for k in ("laminar", "sa", "small", "visc_group")))


def time_tile(meshes, params, states, precision: int, tile: int, trials: int = 50, warmup: int = 10) -> float:
    """Seconds per trial of the arena of these meshes at this tile (graph replays, synchronised)."""
    with B.Batch.of_meshes(meshes, params, states, precision=precision, tile=tile) as b:
        b.set_targets(1e30)
        b.run(warmup)
        t0 = time.perf_counter()
        b.run(trials)
        return (time.perf_counter() - t0) / trials


def pick_tile(meshes, params, states, precision: int = 32, tiles=TILES, trials: int = 50,
              warmup: int = 10) -> tuple[int, dict]:
    """(best tile, {tile: seconds per trial}) for this set of meshes."""
    feats = B.union([B.capacity(m)[1] for m in meshes])
    key = _key(meshes, precision, feats)
    if key in _CACHE:
        return _CACHE[key]
    timings = {t: time_tile(meshes, params, states, precision, t, trials, warmup) for t in tiles}
    best = choose(timings)
    _CACHE[key] = (best, timings)
    return best, timings


FEW_PROBES = (32, 16, 8, 2)


def pick_few_slots(meshes, params, states, precision: int, tile: int, probes=FEW_PROBES, seconds: float = 0.25,
                   warmup: int = 10) -> tuple[int, dict]:
    """(few_slots for two flux threads per cell, {k: (us per trial at 1, at 2)}) at this tile;
    (0, {}) when the tile has no room for two threads per cell (tile 256)."""
    if tile * 2 > 256 or len(meshes) < 2:
        return 0, {}
    feats = B.union([B.capacity(m)[1] for m in meshes])
    key = ("few",) + _key(meshes, precision, feats) + (tile,)
    if key in _CACHE:
        return _CACHE[key]
    n = len(meshes)
    out = {}
    with B.Batch.of_meshes(meshes, params, states, precision=precision, tile=tile) as b:
        b.set_targets(1e30)
        b.run(warmup)
        for k in sorted({min(k, n) for k in probes}, reverse=True):
            per = []
            for spread in (1, 2):
                b.set_launch_shape(spread, 0, spread)
                b.set_targets([1e30] * k + [0.0] * (n - k))
                b.run(warmup)
                before = sum(s["steps"] + s["rejected"] for s in b.states()[:k])
                t0 = time.perf_counter()
                b.advance("all", budget=seconds)
                elapsed = time.perf_counter() - t0
                trials = (sum(s["steps"] + s["rejected"] for s in b.states()[:k]) - before) / k
                per.append(elapsed / max(trials, 1.0) * 1e6)
            out[k] = (round(per[0], 1), round(per[1], 1))
    few = choose_few(out)
    _CACHE[key] = (few, out)
    return few, out


def choose_few(probes: dict) -> int:
    """From {running slots: (us per trial at one, at two threads per cell)}: the largest k such
    that two threads win at k and at every smaller probe (0 if they lose at the smallest)."""
    few = 0
    for k in sorted(probes):
        if probes[k][1] < probes[k][0]:
            few = k
        else:
            break
    return few
