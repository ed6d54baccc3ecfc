"""Full-size configuration 1 (Q4 200x200, 40 401 nodes, 1998 steps) on the GPU,
checked through properties that hold at any size:

* linearity: doubling the source amplitude doubles every FP64 seismogram
  sample exactly (every operation of the loop is linear in the state and a
  factor 2 commutes with rounding), except where values are subnormal;
* mirror symmetry: homogeneous medium, source on the mesh's mirror line y =
  1000 m, receivers symmetric about it: u even and w odd in y up to rounding
  (the reference's force has a -3.5e-15 horizontal part, cos(90 deg) with its
  truncated pi);
* the FP32 + Kahan loop stays within 1e-4 relative L2 of the FP64 loop (u and
  w concatenated), the accuracy bar of the batched forms.

The full-size meshes of configurations 1-3 (Q4 200x200, Q9 400x400, T6 with
topography) also run their first 100 FP64 steps against the oracle, bit for
bit, at every node within 60 m of the source (a point force; the oracle has no
distributed source).
"""
import numpy as np
import pytest

import oracle
from seisfem_amd import gpu, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-case]", msg, flush=True)


@pytest.fixture(scope="module")
def cfg1():
    case = workloads.config1()
    prob = case["problem"]
    op = gpu.Operator(prob["node_num"], prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    yield case, op, md
    md.close()
    op.close()


def _run(cfg, precision, amplitude=1.0e10):
    case, op, md = cfg
    steps = case["steps_exec"]
    run = gpu.Run(op, md, case["src"][:1], steps, case["rec"], case["dt"], case["f0"], case["t0"],
                  amplitude=amplitude, precision=precision)
    run.advance(steps)
    u, w = run.seismograms()
    run.close()
    return u[0], w[0]


def test_full_size_linearity_in_the_amplitude(cfg1):
    say("config 1 full size, FP64, amplitude 1e10 and 2e10")
    u1, w1 = _run(cfg1, gpu.FP64, 1.0e10)
    u2, w2 = _run(cfg1, gpu.FP64, 2.0e10)
    assert np.max(np.abs(u1)) > 1e-3
    for a, b in ((u1, u2), (w1, w2)):
        normal = np.abs(a) > 1e-290
        assert np.array_equal(2.0 * a[normal], b[normal])
        assert np.max(np.abs(2.0 * a[~normal] - b[~normal]), initial=0.0) < 1e-290


def test_full_size_mirror_symmetry(cfg1):
    case = cfg1[0]
    rec_y = case["problem"]["y"][case["rec"]]
    assert np.allclose(rec_y + rec_y[::-1], 2000.0)
    say("config 1 full size, FP64, mirror symmetry about y = 1000 m")
    u, w = _run(cfg1, gpu.FP64)
    su = np.max(np.abs(u))
    assert np.max(np.abs(u - u[:, ::-1])) < 1e-9 * su
    assert np.max(np.abs(w + w[:, ::-1])) < 1e-9 * su


@pytest.mark.parametrize("prec", [gpu.FP32_KAHAN, gpu.FP32_KAHAN16])
def test_full_size_fp32_kahan_within_1e4_of_fp64(cfg1, prec):
    say(f"config 1 full size, FP32+Kahan (precision {prec}) vs FP64")
    u64, w64 = _run(cfg1, gpu.FP64)
    u32, w32 = _run(cfg1, prec)
    a = np.concatenate([u32.ravel(), w32.ravel()])
    b = np.concatenate([u64.ravel(), w64.ravel()])
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-4


def _bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.mark.parametrize("name", ["config1", "config2", "config3"])
def test_full_size_fp64_first_steps_bit_identical_to_oracle(name):
    case = getattr(workloads, name)()
    prob = case["problem"]
    s = int(case["src"][0])
    x, y = prob["x"], prob["y"]
    rec = np.where(np.hypot(x - x[s], y - y[s]) <= 60.0)[0].astype(np.int32)
    steps = 100
    say(f"{name} full size ({prob['node_num']} nodes), FP64, {steps} steps, {rec.size} receivers around node {s}")
    res = gpu.simulate(prob, np.array([s], np.int32), steps, rec, case["dt"], case["f0"], case["t0"],
                       precision=gpu.FP64)
    ref = oracle.oracle_masslump(prob, s, steps + 2, rec, case["dt"], case["f0"], case["t0"])
    assert ref["status"] == steps
    assert np.max(np.abs(ref["seis_u"])) > 0 and np.max(np.abs(ref["seis_w"])) > 0
    assert _bits(res["seis_u"][0], ref["seis_u"]) and _bits(res["seis_w"][0], ref["seis_w"])


def test_batched_fp64_bench_shape_bit_identical_to_oracle():
    """The batched workload's shape (128 shots on Q4 400x400, 64-shot lane
    groups, kernel variant 5, XCD stripes of 12 row blocks, uneven step counts)
    in FP64: three shots, 60 steps, every node within 60 m of each source,
    bit-identical to the oracle's single-shot runs."""
    case = workloads.config4a()
    prob = case["problem"]
    x, y = prob["x"], prob["y"]
    steps = 60
    nsteps = np.clip(case["nsteps"] - 1510, 20, steps).astype(np.int32)   # uneven, 20..60 steps
    nsteps[[0, 64, 127]] = steps
    near = lambda s: np.hypot(x - x[s], y - y[s]) <= 60.0
    rec = np.where(near(int(case["src"][0])) | near(int(case["src"][64])) | near(int(case["src"][127])))[0]
    rec = rec.astype(np.int32)
    say(f"config 4a batched FP64: 128 shots, steps {int(nsteps.min())}..{steps}, {rec.size} receivers")
    op = gpu.Operator(prob["node_num"], prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    run = gpu.Run(op, md, case["src"], steps, rec, case["dt"], case["f0"], case["t0"], precision=gpu.FP64,
                  nsteps=nsteps, shots_per_lane_group=64, block=256, kernel_variant=5, xcd_remap=12)
    run.advance(steps)
    u, w = run.seismograms()
    run.close()
    md.close()
    op.close()
    for k in (0, 64, 127):
        ref = oracle.oracle_masslump(prob, int(case["src"][k]), steps + 2, rec, case["dt"], case["f0"], case["t0"])
        assert np.max(np.abs(ref["seis_u"])) > 0
        assert _bits(u[k], ref["seis_u"]) and _bits(w[k], ref["seis_w"]), k
