"""Autotuner for the step kernel: picks the block size, the lane grouping
(rows per wavefront vs shots per lane group), the kernel variant, the XCD
block map (contiguous ranges, round-robin, or stripes of Q row blocks) and the
row order (node order or a Hilbert walk over the node coordinates) by timing
short runs of the actual problem on the GPU.

Each candidate runs ``steps`` steps (after ``steps`` untimed ones) as one
captured graph with the energy guard at the end, timed with HIP events on the
run's stream; the fastest candidate wins.  Results are cached per (nodes, nnz,
shots, precision).  Every candidate gives the same results bit for bit.
"""
import itertools
import math

from . import gpu, host

_cache = {}


def stripe_widths(nodes, sp, block):
    """XCD stripe widths to try: about (row blocks per grid line) / 8, the line
    length taken as sqrt(nodes) (a roughly square mesh), halved and doubled."""
    rows_per_block = (block // 64) * (1 if sp >= 64 else 64 // sp)
    q = math.sqrt(max(nodes, 1)) / (rows_per_block * 8.0)
    return sorted({max(1, int(round(q * f))) for f in (0.5, 1.0, 2.0)})


def candidates(nshots, nodes=160801):
    """(shots per lane group, block, kernel variant, xcd remap, row order) to try."""
    if nshots <= 1:
        sps, blocks = [1], [64, 128]
    elif nshots <= 4:
        sps, blocks = [4, 16], [256]
    elif nshots <= 16:
        sps, blocks = [16, 64], [256]
    else:
        sps, blocks = [32, 64], [256]
    # variant 7 (LDS-staged gathers, lane groups of 16+ shots) is left out: it
    # measured about 2x slower than 5 on Q4 400x400 x 128 shots (the block's
    # staging and its compute do not overlap at 1-3 blocks per CU)
    variants = [1, 3, 4, 5, 6]
    cands = list(itertools.product(sps, blocks, variants, [True, False], ["node", "hilbert"]))
    if nshots > 1:
        # XCD stripes of Q row blocks (node order): all XCDs sweep the rows together;
        # Q near (row blocks per grid line) / 8 keeps vertical neighbours in one L2
        for sp, block in itertools.product(sps, blocks):
            qs = [q for q in stripe_widths(nodes, sp, block) if q > 1]
            cands += list(itertools.product([sp], [block], variants, qs, ["node"]))
    return cands


def tune(op, md, src, rec, dt, f0, t0, precision, steps=40, cands=None, log=None, coords=None):
    """Returns (best dict of gpu.Run keyword arguments, table of (candidate, ms/step)).
    coords = (x, y) of the nodes enables the Hilbert row order candidates."""
    nshots = len(src)
    key = (op.node_num, op.nnz, nshots, precision, coords is not None)
    if key in _cache:
        return _cache[key]
    orders = {"node": None}
    if coords is not None:
        orders["hilbert"] = host.locality_order(*coords)
    def timed(sp, block, variant, xcd, oname, k):
        # two graph replays of k launches: the first warms up, the second is timed
        run = gpu.Run(op, md, src, 2 * k, rec, dt, f0, t0, precision=precision, block=block,
                      shots_per_lane_group=sp, kernel_variant=variant, xcd_remap=xcd, guard_every=k,
                      row_order=orders[oname])
        run.advance(k)
        run.advance(k)
        ms = run.last_ms_per_step
        run.close()
        return ms

    table = []
    for sp, block, variant, xcd, oname in (cands or candidates(nshots, op.node_num)):
        if oname not in orders:
            continue
        try:
            ms = timed(sp, block, variant, xcd, oname, steps)
        except gpu.NativeError as e:
            if e.code != -1:  # only an argument the shape cannot take (e.g. a halo beyond LDS) is skipped
                raise
            if log:
                log(f"tune skip sp={sp} block={block} variant={variant} order={oname}: {e}")
            continue
        cand = dict(shots_per_lane_group=sp, block=block, kernel_variant=variant, xcd_remap=xcd, row_order=oname)
        table.append((cand, ms))
        if log:
            log(f"tune {cand}: {ms:.4f} ms/step")
    # the few fastest are often within a per cent: time them again over twice the
    # steps and keep the better of the two figures of each
    table.sort(key=lambda c: c[1])
    for i in range(min(3, len(table))):
        c, ms = table[i]
        ms2 = timed(c["shots_per_lane_group"], c["block"], c["kernel_variant"], c["xcd_remap"], c["row_order"],
                    2 * steps)
        table[i] = (c, min(ms, ms2))
        if log:
            log(f"tune again {c}: {ms2:.4f} ms/step")
    best = dict(min(table, key=lambda c: c[1])[0])
    best["row_order"] = orders[best["row_order"]]
    _cache[key] = (best, table)
    return best, table


def describe(best):
    """JSON-friendly form of a tuned launch (the row order by name)."""
    d = dict(best)
    d["row_order"] = "node" if d.get("row_order") is None else "hilbert"
    return d
