"""sfg_run_snapshot_noalloc (no device allocation; staging in a fixed device
array) against sfg_run_snapshot, through the C ABI: bit-identical on two
calls in a row, the second on a smaller run, then again on the first."""
import ctypes as C

import numpy as np
import pytest

from seisfem_amd import gpu, host, workloads
from seisfem_amd._native import check, lib

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-case]", msg, flush=True)


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def snap_noalloc(run, shot):
    n = run.op.node_num
    U = np.full(n, np.nan)
    W = np.full(n, np.nan)
    check(lib().sfg_run_snapshot_noalloc(run.h, shot, _pd(U), _pd(W)), "sfg_run_snapshot_noalloc")
    return U, W


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def _setup(nelem, nshots):
    case = workloads.small_q4(nelem=nelem, steps=40, nshots=nshots)
    prob = case["problem"]
    op = gpu.Operator(prob["node_num"], prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    return case, op, md


def test_snapshot_noalloc_matches_snapshot():
    # larger run first: FP64, 3 shots with uneven step counts
    ca, opa, mda = _setup(40, 3)
    say(f"snapshot_noalloc: Q4 40x40 ({opa.node_num} nodes), 3 shots FP64, uneven steps")
    ra = gpu.Run(opa, mda, ca["src"], 40, ca["rec"], ca["dt"], ca["f0"], ca["t0"], precision=gpu.FP64,
                 nsteps=np.array([40, 25, 33], np.int32), shots_per_lane_group=4)
    ra.advance(30)
    for s in range(3):
        U0, W0 = ra.snapshot(s)
        U1, W1 = snap_noalloc(ra, s)
        assert np.max(np.abs(U0)) > 0
        assert same_bits(U0, U1) and same_bits(W0, W1), s

    # second call on a smaller run: FP32 with bfloat16 Kahan terms, 64-shot groups
    cb, opb, mdb = _setup(12, 70)
    assert opb.node_num < opa.node_num
    say(f"snapshot_noalloc: Q4 12x12 ({opb.node_num} nodes), 70 shots FP32+Kahan16")
    rb = gpu.Run(opb, mdb, cb["src"], 40, cb["rec"], cb["dt"], cb["f0"], cb["t0"], precision=gpu.FP32_KAHAN16,
                 shots_per_lane_group=64, kernel_variant=3)
    rb.advance(17)
    for s in (0, 35, 69):
        U0, W0 = rb.snapshot(s)
        U1, W1 = snap_noalloc(rb, s)
        assert np.max(np.abs(U0)) > 0
        assert same_bits(U0, U1) and same_bits(W0, W1), s

    # back on the larger run after more steps: the shared array holds no stale values
    ra.advance(10)
    for s in range(3):
        U0, W0 = ra.snapshot(s)
        U1, W1 = snap_noalloc(ra, s)
        assert same_bits(U0, U1) and same_bits(W0, W1), s

    # bad shot index: an argument error, nothing launched
    n = opb.node_num
    U = np.zeros(n)
    W = np.zeros(n)
    assert lib().sfg_run_snapshot_noalloc(rb.h, 70, _pd(U), _pd(W)) == -1
    for h in (rb, mdb, opb, ra, mda, opa):
        h.close()


def test_snapshot_noalloc_limit_is_declared():
    """The staging limit is the header's SFG_NOALLOC_MAX_NODES (4 Mi nodes)."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "include", "seisfem_gpu.h")).read()
    m = re.search(r"#define SFG_NOALLOC_MAX_NODES \(1 << (\d+)\)", text)
    assert m and int(m.group(1)) == 22
    # configuration 2 (Q9 400x400, 801 x 801 nodes) fits
    assert int(lib().sfh_mesh_node_num(host.Q9, 400, 400)) == 801 * 801 <= 1 << 22
