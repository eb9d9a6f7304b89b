"""Tile / march-chunk autotuner for the stage kernel.

For one grid shape it measures candidate (ty, tz, xc) settings of the marching kernel on
a batch of identical events large enough to keep the GPU busy, so that the figure is the
throughput the shape reaches inside an ensemble rather than the latency of one event
alone.  Results are kept per shape in memory and, when a path is given, in a JSON file.
Every candidate gives the same fields (the kernels are bit-identical across tiles); the
tuner only chooses speed.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import _lib

TILE_THREADS = 256
MAX_HALO_CELLS = 378
MAX_FACES = 256


def _valid(ny, nz, ty, tz):
    hy = 2 if ny > 1 else 0
    hz = 2 if nz > 1 else 0
    if ty < 1 or tz < 1 or ty * tz > TILE_THREADS:
        return False
    if (ty + 2 * hy) * (tz + 2 * hz) > MAX_HALO_CELLS:
        return False
    if (hy and (ty + 1) * tz > MAX_FACES) or (hz and ty * (tz + 1) > MAX_FACES):
        return False
    return True


def candidates(nx, ny, nz, default=None):
    """Tile shapes and chunk lengths worth timing for a grid; ``default`` (ty, tz), the
    library's own choice for the grid (its cost model of the wave passes), comes first."""
    tiles = []
    if nz == 1:
        for nt in range(1, 5):
            ty = -(-ny // nt)
            if ty <= TILE_THREADS:
                tiles.append((ty, 1))
        tiles += [(t, 1) for t in (128, 192, 256) if t < ny]
    else:
        for tz in range(4, 65):
            for ty in range(1, TILE_THREADS // tz + 1):
                if ty * tz >= 160 and _valid(ny, nz, ty, tz):
                    pad = (-(-ny // ty) * ty) * (-(-nz // tz) * tz) / float(ny * nz)
                    tiles.append((pad * TILE_THREADS / (ty * tz), ty, tz))
        tiles = [(ty, tz) for _, ty, tz in sorted(tiles)[:6]]
    if default:
        tiles = [tuple(int(x) for x in default)] + tiles
    tiles = [t for t in dict.fromkeys(tiles) if _valid(ny, nz, *t)]
    chunks = sorted(set(c for c in (2, 4, 8, 12, 16, 24, 32) if c <= nx) | {0})
    return [(ty, tz, xc) for (ty, tz) in tiles for xc in chunks]


class Autotuner(object):
    """Per-shape best (ty, tz, xc) for the marching stage kernel."""

    def __init__(self, params, eos_table, path=None, target_cells=8e6, steps=3, device=0):
        self.params = params
        self.eos_table = eos_table
        self.path = path
        self.target_cells = target_cells
        self.steps = steps
        self.device = device
        self.cache = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.cache = {tuple(json.loads(k)): tuple(v) for k, v in json.load(f).items()}

    def key(self, dims):
        bulk = bool(self.params.flags & _lib.CV_FLAG_BULK)
        return (int(dims[0]), int(dims[1]), int(dims[2]), int(bulk), int(self.eos_table.kind_id))

    def tune(self, dims, ev):
        """Best (ty, tz, xc, ms_per_cell_update) for grid dims with initial field ev."""
        k = self.key(dims)
        if k in self.cache:
            return self.cache[k]
        nx, ny, nz = (int(x) for x in dims)
        ncell = nx * ny * nz
        copies = int(max(1, min(64, np.ceil(self.target_cells / ncell))))
        batch = _lib.Batch(self.params, self.eos_table, [dims] * copies, device=self.device)
        try:
            for e in range(copies):
                h = batch.handle(e)
                h.visc_initialize(self.params.tau0)
                h.ideal_stage(1, np.float32(self.params.tau0))
                h.set('ev1', ev)
            best = None
            _, dty, dtz = batch.handle(0).stage_kernel()
            for ty, tz, xc in candidates(nx, ny, nz, default=(dty, dtz)):
                try:
                    for e in range(copies):
                        h = batch.handle(e)
                        h.set_stage_kernel(2, ty, tz)
                        h.set_march_chunk(xc)
                except _lib.CvError:
                    continue
                batch.steps(1)
                batch.steps(self.steps)
                ms = batch.last_elapsed_ms() / self.steps
                per = ms / (ncell * copies)
                if best is None or per < best[3]:
                    best = (ty, tz, xc, per)
        finally:
            batch.close()
        self.cache[k] = best
        if self.path:
            with open(self.path, 'w') as f:
                json.dump({json.dumps(list(kk)): list(v) for kk, v in self.cache.items()}, f, indent=1)
        return best

    def apply(self, batch, dims_list, evs):
        """Tune every distinct shape of a batch and set its events' tiles."""
        for e, d in enumerate(dims_list):
            ty, tz, xc, _ = self.tune(d, evs[e])
            h = batch.handle(e)
            h.set_stage_kernel(2, ty, tz)
            h.set_march_chunk(xc)
