"""Launch-tile autotuner for a batch arena.

The only launch-shape choice of the time step is the tile: threads per block of every cell,
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
