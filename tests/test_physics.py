"""Travel times of the oracle's loop (the reference's masslump algorithm,
restated in C): a vertical point force in a homogeneous isotropic medium
radiates P along the force axis and S across it.  The moveout between two
receivers at different distances from the source, taken as the lag that best
aligns their vertical-displacement traces, gives the P speed (from c11, c33)
on the axis and the S speed (from c44) across it.  A swapped stiffness, a
wrong sign in the split-field update or a mis-scaled mass would move both
far outside the tolerances, which allow for the mesh's dispersion.
"""
import numpy as np
import pytest

import oracle
from seisfem_amd import host, workloads

VP, VS, RHO = 3000.0, 1732.0, 2000.0
EDGE, DT, F0 = 10.0, 5.0e-4, 15.0
T0 = 1.2 / F0


def _model(depth, y):
    one = np.ones(depth.size)
    return workloads.isotropic_model(RHO * one, VP * one, VS * one)


@pytest.fixture(scope="module")
def shot():
    # 1000 m high (x), 800 m wide (y); free surface on top, M-PML 150 m on the other sides;
    # source 200 m below the surface in the middle
    case = workloads.structured_case(host.Q4, 100, 80, EDGE, _model, 15, DT, F0, T0, 700, 200.0, [400.0], 2,
                                     0.0, 800.0, mpml_fn=oracle.oracle_abc_mpml)
    prob = case["problem"]
    x, y = prob["x"], prob["y"]
    xs, ys = 800.0, 400.0
    assert abs(x[case["src"][0]] - xs) < 1e-9 and abs(y[case["src"][0]] - ys) < 1e-9

    def node(px, py):
        return host.node_location(EDGE, x, y, px, py)

    axis = [node(xs - d, ys) for d in (200.0, 400.0)]       # below the source, on the force axis
    across = [node(xs, ys + d) for d in (100.0, 200.0)]     # same depth, across the axis
    rec = np.array(axis + across, np.int32)
    steps = 700
    ref = oracle.oracle_masslump(prob, int(case["src"][0]), steps + 2, rec, DT, F0, T0)
    assert ref["status"] == steps
    return ref["seis_u"]                                    # [steps][receivers]


def _lag(a, b, max_lag):
    """Delay of trace b behind trace a (seconds) maximising their correlation."""
    best, arg = -np.inf, 0
    for k in range(max_lag + 1):
        c = float(np.dot(a[: a.size - k], b[k:]))
        if c > best:
            best, arg = c, k
    return arg * DT


def test_p_moveout_along_the_force_axis(shot):
    u = shot
    assert np.max(np.abs(u[:, 0])) > 0
    dt_p = _lag(u[:, 0], u[:, 1], 400)
    assert abs(dt_p - 200.0 / VP) < 0.03 * 200.0 / VP, dt_p      # 66.7 ms for 200 m (measured 66.0)


def test_s_moveout_across_the_force_axis(shot):
    u = shot
    dt_s = _lag(u[:, 2], u[:, 3], 400)
    assert abs(dt_s - 100.0 / VS) < 0.03 * 100.0 / VS, dt_s      # 57.7 ms for 100 m (measured 58.0)
