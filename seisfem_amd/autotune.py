"""Autotuner for the step kernel: picks the block size, the lane grouping
(rows per wavefront vs shots per lane group), the kernel variant (with the
sliced layouts' sorting window), the XCD block map (contiguous ranges,
round-robin, or stripes of Q row blocks) and the row order (node order, a
Hilbert walk over the node coordinates, or patches of a structured mesh) by
timing short runs of the actual problem on the GPU.

Each candidate runs ``steps`` steps (after ``steps`` untimed ones) as one
captured graph with the energy guard at the end, timed with HIP events on the
run's stream; the fastest candidate wins.  Results are cached per (nodes, nnz,
shots, precision).  Every candidate gives the same results bit for bit.
"""
import math

from . import gpu, host

_cache = {}


def stripe_widths(nodes, sp, block):
    """XCD stripe widths to try: about (row blocks per grid line) / 8, the line
    length taken as sqrt(nodes) (a roughly square mesh), halved and doubled."""
    rows_per_block = (block // 64) * (1 if sp >= 64 else 64 // sp)
    q = math.sqrt(max(nodes, 1)) / (rows_per_block * 8.0)
    return sorted({max(1, int(round(q * f))) for f in (0.5, 1.0, 2.0)})


def candidates(nshots, nodes=160801, patches=False):
    """Launch shapes to try, as gpu.Run keyword dicts (row_order by name: "node",
    "hilbert", or "patchWxH" for patch_order(W, H))."""
    def c(sp, block, variant, xcd, order, sigma=0):
        return dict(shots_per_lane_group=sp, block=block, kernel_variant=variant, xcd_remap=xcd, row_order=order,
                    sell_sigma=sigma)

    if nshots <= 1:
        # one row per lane: per-row layouts, the sliced layout (8) and, on meshes
        # with a line structure, the sliced layout with LDS-staged halos over
        # patches of the mesh (9)
        cands = [c(1, b, v, x, o) for b in (64, 128) for v in (1, 5) for x in (True, False)
                 for o in ("node", "hilbert")]
        cands += [c(1, b, 8, x, "node", sg) for b in (128, 256) for x in (True, False) for sg in (0, 128)]
        if patches:
            cands += [c(1, b, 9, x, o) for b, o in ((256, "patch16x16"), (256, "patch32x8"), (64, "patch16x4"))
                      for x in (True, False)]
        return cands
    if nshots <= 4:
        sps, blocks = [4, 16], [256]
    elif nshots <= 16:
        sps, blocks = [16, 64], [256]
    else:
        sps, blocks = [32, 64], [256]
    # variant 7 (LDS-staged gathers, lane groups of 16+ shots) is left out: it
    # measured about 2x slower than 5 on Q4 400x400 x 128 shots (the block's
    # staging and its compute do not overlap at 1-3 blocks per CU)
    variants = [1, 3, 4, 5, 6]
    cands = [c(sp, b, v, x, o) for sp in sps for b in blocks for v in variants for x in (True, False)
             for o in ("node", "hilbert")]
    # XCD stripes of Q row blocks (node order): all XCDs sweep the rows together;
    # Q near (row blocks per grid line) / 8 keeps vertical neighbours in one L2
    for sp in sps:
        for b in blocks:
            cands += [c(sp, b, v, q, "node") for q in stripe_widths(nodes, sp, b) if q > 1 for v in variants]
    return cands


def tune(op, md, src, rec, dt, f0, t0, precision, steps=40, cands=None, log=None, coords=None):
    """Returns (best dict of gpu.Run keyword arguments, table of (candidate, ms/step)).
    coords = (x, y) of the nodes enables the Hilbert and patch row orders."""
    nshots = len(src)
    key = (op.node_num, op.nnz, nshots, precision, coords is not None)
    if key in _cache:
        return _cache[key]
    orders = {"node": None}
    patches = False
    if coords is not None:
        orders["hilbert"] = host.locality_order(*coords)
        if host.mesh_line_length(*coords) > 0:
            patches = True
            for w, h in ((16, 16), (32, 8), (16, 4)):
                orders[f"patch{w}x{h}"] = host.patch_order(coords[0], coords[1], w, h)

    def timed(cand, k):
        # two graph replays of k launches: the first warms up, the second is timed
        kw = dict(cand, row_order=orders[cand["row_order"]])
        run = gpu.Run(op, md, src, 2 * k, rec, dt, f0, t0, precision=precision, guard_every=k, **kw)
        run.advance(k)
        run.advance(k)
        ms = run.last_ms_per_step
        run.close()
        return ms

    table = []
    for cand in (cands or candidates(nshots, op.node_num, patches)):
        if cand["row_order"] not in orders:
            continue
        try:
            ms = timed(cand, steps)
        except gpu.NativeError as e:
            if e.code != -1:  # only an argument the shape cannot take (e.g. a halo beyond LDS) is skipped
                raise
            if log:
                log(f"tune skip {cand}: {e}")
            continue
        table.append((cand, ms))
        if log:
            log(f"tune {cand}: {ms:.4f} ms/step")
    # the few fastest are often within a per cent: time them again over twice the
    # steps and keep the better of the two figures of each
    table.sort(key=lambda c: c[1])
    for i in range(min(3, len(table))):
        c, ms = table[i]
        ms2 = timed(c, 2 * steps)
        table[i] = (c, min(ms, ms2))
        if log:
            log(f"tune again {c}: {ms2:.4f} ms/step")
    best = dict(min(table, key=lambda c: c[1])[0])
    best["row_order_name"] = best["row_order"]
    best["row_order"] = orders[best["row_order"]]
    _cache[key] = (best, table)
    return best, table


def describe(best):
    """JSON-friendly form of a tuned launch (the row order by name)."""
    d = dict(best)
    name = d.pop("row_order_name", None)
    d["row_order"] = name or ("node" if d.get("row_order") is None else "given")
    return d


def run_kwargs(best):
    """gpu.Run keyword arguments of a tuned launch."""
    d = dict(best)
    d.pop("row_order_name", None)
    return d


def tune_multi_schedule(make_runs, steps=200, blocks_per_cu=(1, 2, 3), log=None):
    """Schedule of a multi-run: graph-replayed per-step launches (0) or the
    persistent cooperative launch with k resident blocks per CU (the knob).
    make_runs() returns fresh gpu.Run objects of the job (at least 2 * steps
    steps long); each candidate steps ``steps`` steps after as many untimed ones.  Returns (best k, table of
    (k, ms/step)); a k whose blocks cannot all be resident is skipped."""
    table = []
    for k in (0,) + tuple(blocks_per_cu):
        runs = make_runs()
        multi = gpu.Multi(runs)
        try:
            multi.set_persistent(k)
        except gpu.NativeError as e:
            if e.code != -1:
                raise
            if log:
                log(f"schedule {k} blocks/CU skipped: {e}")
            k = None
        if k is not None:
            multi.advance(steps)
            multi.advance(steps)
            table.append((k, multi.last_ms_per_step))
            if log:
                log(f"schedule {'per-step launches' if k == 0 else f'persistent {k} blocks/CU'}: "
                    f"{table[-1][1]:.4f} ms/step")
        multi.close()
        for r in runs:
            r.close()
    best = min(table, key=lambda c: c[1])[0]
    return best, table
