"""dt estimate (no counterpart in the reference; dt is user input there).

Pins: (1) lambda_max of M_L^{-1} K against a dense eigen-solve (numpy) of the
same operator on small meshes; (2) the actual FP64 loop without M-PML stays
bounded at 0.97 dt_max and blows up (energy guard) at 1.05 dt_max, which pins
dt_max = 2/sqrt(lambda_max) for the reference's Newmark update.
"""
import numpy as np
import pytest

from seisfem_amd import NativeError, gpu, host, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-case]", msg, flush=True)


def dense(n, csr):
    p, j, x = csr
    d = np.zeros((n, n))
    for i in range(n):
        d[i, j[p[i]:p[i + 1]]] = x[p[i]:p[i + 1]]
    return d


def dense_lambda(prob):
    n = prob["node_num"]
    S = [dense(n, prob["csr"][k]) for k in range(1, 5)]
    c11, c13, c33, c44 = prob["c"]
    free = (prob["dirichlet"] != 1)[:, None]
    m = prob["mass_lump"][:, None]
    Kuu = c11[:, None] * S[0] + c44[:, None] * S[1]
    Kuw = c13[:, None] * S[2] + c44[:, None] * S[3]
    Kwu = c44[:, None] * S[2] + c13[:, None] * S[3]
    Kww = c44[:, None] * S[0] + c33[:, None] * S[1]
    A = np.block([[free * Kuu / m, free * Kuw / m], [free * Kwu / m, free * Kww / m]])
    return float(np.max(np.abs(np.linalg.eigvals(A))))


def no_pml_case(tc, nelem, heterogeneous=False):
    def model(depth, y):
        n = depth.size
        vp = np.full(n, 3000.0) + (700.0 * np.sin(depth / 37.0) if heterogeneous else 0.0)
        return workloads.isotropic_model(np.full(n, 2000.0) + (depth if heterogeneous else 0.0), vp, vp / 1.73)

    zero = lambda *a, **k: [np.zeros(a[2].size) for _ in range(6)]  # noqa: E731
    return workloads.structured_case(tc, nelem, nelem, 10.0, model, 2, 5e-4, 25.0, 0.04, 100, 20.0,
                                     5.0 * nelem, 5, 0.0, 10.0 * nelem, mpml_fn=zero)


@pytest.mark.parametrize("tc,nelem,het", [(host.Q4, 10, False), (host.Q9, 5, True), (host.T6, 6, False),
                                          (host.T3, 9, True)])
def test_power_iteration_matches_dense_eigensolve(tc, nelem, het):
    case = no_pml_case(tc, nelem, het)
    prob = case["problem"]
    say(f"dt estimate {host.TYPE_NAMES[tc]} {nelem}x{nelem} heterogeneous={het}")
    op = gpu.Operator(prob["node_num"], prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    lam, dtm, it = gpu.estimate_dt(op, md, max_iter=20000, tol=1e-12)
    ref = dense_lambda(prob)
    assert abs(lam - ref) <= 1e-6 * ref, (lam, ref, it)
    assert abs(dtm - 2.0 / np.sqrt(ref)) <= 1e-6 * dtm


def test_loop_stability_limit():
    case = no_pml_case(host.Q4, 10)
    prob = case["problem"]
    op = gpu.Operator(prob["node_num"], prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    lam, dtm, _ = gpu.estimate_dt(op, md, max_iter=20000, tol=1e-12)
    say(f"stability: dt_max {dtm:.6e}; runs at 0.97 and 1.05 dt_max")
    ok = gpu.Run(op, md, case["src"], 600, case["rec"], 0.97 * dtm, case["f0"], case["t0"], precision=gpu.FP64,
                 guard_every=20)
    ok.advance(600)
    e = ok.energy()
    assert np.all(np.isfinite(e)) and np.max(e) < 1e7
    ok.close()
    bad = gpu.Run(op, md, case["src"], 600, case["rec"], 1.05 * dtm, case["f0"], case["t0"], precision=gpu.FP64,
                  guard_every=20)
    with pytest.raises(NativeError) as ei:
        bad.advance(600)
    assert ei.value.code == -6
    bad.close()


def test_stream_copy_yardstick():
    say("stream copy 256 MiB")
    gbs, ms = gpu.stream_copy(256 << 20, 5)
    # 2 x 256 MiB moved per copy: a device copy is well above PCIe rates and below 2x the HBM peak
    assert 500.0 < gbs < 16000.0 and ms > 0.0
