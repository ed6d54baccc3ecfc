"""Autotuner for the step kernel: picks the block size, the lane grouping
(rows per wavefront vs shots per lane group), the kernel variant and the XCD
block remap by timing short runs of the actual problem on the GPU.

Each candidate runs ``steps`` steps (after ``warmup``) with the energy guard
only at the end, timed with HIP events on the run's stream; the fastest
candidate wins.  Results are cached per (nodes, nnz, shots, precision).
"""
import itertools

from . import gpu

_cache = {}


def candidates(nshots):
    """(shots per lane group, block, kernel variant, xcd remap) to try."""
    if nshots <= 1:
        sps = [1]
    elif nshots <= 4:
        sps = [4, 16]
    elif nshots <= 16:
        sps = [16, 64]
    else:
        sps = [16, 32, 64]
    return list(itertools.product(sps, [64, 128, 256], [1, 2, 3, 4], [True, False]))


def tune(op, md, src, rec, dt, f0, t0, precision, steps=40, cands=None, log=None):
    """Returns (best dict, table of (candidate, ms/step))."""
    nshots = len(src)
    key = (op.node_num, op.nnz, nshots, precision)
    if key in _cache:
        return _cache[key]
    table = []
    for sp, block, variant, xcd in (cands or candidates(nshots)):
        # two graph replays of `steps` launches: the first warms up, the second is timed
        run = gpu.Run(op, md, src, 2 * steps, rec, dt, f0, t0, precision=precision, block=block,
                      shots_per_lane_group=sp, kernel_variant=variant, xcd_remap=xcd, guard_every=steps)
        run.advance(steps)
        run.advance(steps)
        ms = run.last_ms_per_step
        run.close()
        cand = dict(shots_per_lane_group=sp, block=block, kernel_variant=variant, xcd_remap=xcd)
        table.append((cand, ms))
        if log:
            log(f"tune {cand}: {ms:.4f} ms/step")
    best = min(table, key=lambda c: c[1])[0]
    _cache[key] = (best, table)
    return best, table
