"""The benchmark configurations at their full sizes and lengths on the GPU.

The FP64 GPU loop is bit-identical to the oracle (test_gpu_parity.py, and the
first 100 full-size steps in test_gpu_full_size.py), so at full length it is
the FP64 reference each FP32 form is held to: rel-L2 <= 1e-4 of the
seismograms (u and w concatenated, every executed step), the metric's bar.
Reference: the update whose long-run drift the Kahan terms hold is
time_evolution/elastic_wave.c:754-786.

* configuration 4 part 1: 128 shots on Q4 400x400, U[1500,3000] steps, the
  tuned batched launch (64-shot lane groups, summed rows, kernel variant 3,
  XCD stripes of 13) in both FP32+Kahan forms, against the batched FP64 run;
* configuration 4 part 2: the 16 models of uneven size with part 1 in one
  launch per step, each model against its own FP64 run over its full length,
  part 1 bit-identical to its separate run;
* configuration 2: Q9 400x400 VTI over its 4000 steps, both Kahan forms;
* configuration 3: T6 with topography and the explosive source over its
  first 1500 steps (the M-PML on T6 grows without bound later, in the FP64
  reference loop as well);
* the explosive source radiates P waves only: in a homogeneous medium the
  displacement around it is radial, with no tangential (S) motion at the S time;
* configuration 5's network (13 -> 1024 x 3 -> 32000): gradients of every
  layer and one Adam step against the MLP oracle.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from seisfem_amd import gpu, host, surrogate, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-case]", msg, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def _shot_errors(u, w, u64, w64, nsteps):
    errs = []
    for s in range(len(nsteps)):
        k = int(nsteps[s])
        errs.append(rel_l2(np.concatenate([u[s, :k].ravel(), w[s, :k].ravel()]),
                           np.concatenate([u64[s, :k].ravel(), w64[s, :k].ravel()])))
    return np.array(errs)


TUNED = dict(shots_per_lane_group=64, block=256, kernel_variant=3, xcd_remap=13)


@pytest.fixture(scope="module")
def cfg4a():
    case = workloads.config4a()
    prob = case["problem"]
    op = gpu.Operator(prob["node_num"], prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    base = dict(dt=case["dt"], f0=case["f0"], t0=case["t0"])
    L = int(case["nsteps"].max())
    say(f"config 4a batched FP64: 128 shots, steps {int(case['nsteps'].min())}..{L}")
    r = gpu.Run(op, md, case["src"], L, case["rec"], precision=gpu.FP64, nsteps=case["nsteps"], **base, **TUNED)
    assert r.advance(L) == L
    u64, w64 = r.seismograms()
    r.close()
    yield case, op, md, base, L, u64, w64
    md.close()
    op.close()


@pytest.mark.parametrize("prec", [gpu.FP32_KAHAN16, gpu.FP32_KAHAN])
def test_config4a_batched_fp32_full_length_within_1e4(cfg4a, prec):
    case, op, md, base, L, u64, w64 = cfg4a
    nsteps = case["nsteps"]
    assert nsteps.min() >= 1500 and nsteps.max() <= 3000 and len(nsteps) == 128
    say(f"config 4a batched precision {prec}: 128 shots x {L} launches, tuned launch {TUNED}")
    r = gpu.Run(op, md, case["src"], L, case["rec"], precision=prec, nsteps=nsteps, **base, **TUNED)
    assert r.advance(L) == L
    u, w = r.seismograms()
    r.close()
    errs = _shot_errors(u, w, u64, w64, nsteps)
    say(f"  rel-L2 vs FP64: max {errs.max():.2e}, median {np.median(errs):.2e}")
    assert np.all(np.isfinite(errs)) and errs.max() < 1e-4
    # every shot moved, and stopped at its own step count
    for s in (0, 63, 127):
        k = int(nsteps[s])
        assert np.max(np.abs(u64[s, :k])) > 0 and not np.any(u[s, k:]) and not np.any(u64[s, k:])


def test_config4b_multi_launch_full_length(cfg4a):
    """Configuration 4 part 2 with part 1 in the same multi-run (FP32 + bfloat16
    Kahan terms; part 1 in its own step launch, the models in the sliced
    layout in a launch of their own, as the benchmark runs them): each of the 16
    models within 1e-4 of its own FP64 run over its full step count; part 1
    bit-identical to its separate run."""
    case, op, md, base, L, u64, w64 = cfg4a
    models = workloads.config4b()
    sizes = [c["size"] for c in models]
    assert set(sizes) <= {(100, 100), (200, 150), (300, 200), (500, 300)} and len(models) == 16
    say(f"config 4b: 128 shots + 16 models {sizes} in one launch, FP32+Kahan16")
    single = gpu.Run(op, md, case["src"], L, case["rec"], precision=gpu.FP32_KAHAN16, nsteps=case["nsteps"],
                     **base, **TUNED)
    single.advance(L)
    us, ws = single.seismograms()
    single.close()
    part1 = gpu.Run(op, md, case["src"], L, case["rec"], precision=gpu.FP32_KAHAN16, nsteps=case["nsteps"],
                    **base, **TUNED)
    keep, runs32, runs64 = [], [part1], []
    for c in models:
        p = c["problem"]
        mop = gpu.Operator(p["node_num"], p["csr"])
        mmd = gpu.Medium(mop, p["c"], p["mass_lump"], p["dirichlet"], p["mpml"])
        keep += [mop, mmd]
        runs32.append(gpu.Run(mop, mmd, c["src"], L, c["rec"], c["dt"], c["f0"], c["t0"],
                              precision=gpu.FP32_KAHAN16, nsteps=c["nsteps"], block=256, shots_per_lane_group=1,
                              kernel_variant=8, xcd_remap=False))
        runs64.append(gpu.Run(mop, mmd, c["src"], L, c["rec"], c["dt"], c["f0"], c["t0"], precision=gpu.FP64,
                              nsteps=c["nsteps"]))
    g32 = gpu.Multi(runs32)
    assert g32.advance(L) == L
    g64 = gpu.Multi(runs64)
    assert g64.advance(L) == L
    u1, w1 = part1.seismograms()
    assert _bits(u1, us) and _bits(w1, ws)
    errs = []
    for c, r32, r64 in zip(models, runs32[1:], runs64):
        a, b = r32.seismograms(), r64.seismograms()
        k = int(c["nsteps"][0])
        assert 1500 <= k <= 3000 and np.max(np.abs(b[0][0, :k])) > 0
        errs.append(rel_l2(np.concatenate([a[0][0, :k].ravel(), a[1][0, :k].ravel()]),
                           np.concatenate([b[0][0, :k].ravel(), b[1][0, :k].ravel()])))
    say(f"  per-model rel-L2 vs FP64: max {max(errs):.2e}")
    assert max(errs) < 1e-4
    g32.close()
    g64.close()
    for r in runs32 + runs64:
        r.close()
    for h in keep:
        h.close()


def _single(case, op, md, prec, steps, sources=None):
    r = gpu.Run(op, md, case["src"][:1], steps, case["rec"], case["dt"], case["f0"], case["t0"], precision=prec,
                sources=sources)
    assert r.advance(steps) == steps
    u, w = r.seismograms()
    r.close()
    return np.concatenate([u[0].ravel(), w[0].ravel()])


def test_config2_fp32_kahan_4000_steps_within_1e4():
    case = workloads.config2()
    prob = case["problem"]
    op = gpu.Operator(prob["node_num"], prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    steps = case["steps_exec"]
    assert steps == 4000 and prob["node_num"] == 801 * 801
    say(f"config 2 Q9 400x400 VTI: FP64 {steps} steps")
    ref = _single(case, op, md, gpu.FP64, steps)
    assert np.max(np.abs(ref)) > 0
    for prec in (gpu.FP32_KAHAN, gpu.FP32_KAHAN16):
        say(f"config 2 precision {prec}: {steps} steps")
        err = rel_l2(_single(case, op, md, prec, steps), ref)
        say(f"  rel-L2 vs FP64 {err:.2e}")
        assert err < 1e-4
    md.close()
    op.close()


def test_config3_explosive_1500_steps_within_1e4():
    case = workloads.config3(steps_exec=1500)
    prob = case["problem"]
    op = gpu.Operator(prob["node_num"], prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    say(f"config 3 T6 topography, explosive source ({len(case['sources'][0][0])} nodes): FP64 1500 steps")
    ref = _single(case, op, md, gpu.FP64, 1500, case["sources"])
    assert np.max(np.abs(ref)) > 0
    for prec in (gpu.FP32_KAHAN, gpu.FP32_KAHAN16):
        say(f"config 3 precision {prec}: 1500 steps")
        err = rel_l2(_single(case, op, md, prec, 1500, case["sources"]), ref)
        say(f"  rel-L2 vs FP64 {err:.2e}")
        assert err < 1e-4
    md.close()
    op.close()


def test_explosive_source_radiates_p_waves_only():
    """Homogeneous Q4 medium (vp 3000, vs 1732 m/s; 5 m elements, 20 per P
    wavelength at 30 Hz), explosive source at the centre of a 2.4 km square,
    receivers on a ring of radius 400 m (no reflection from the M-PML or the
    free surface reaches them in the window): the displacement is radial
    (tangential part < 3 % of it over the record; the discrete source is four
    axial forces and the mesh is not isotropic, so not 0), peaks at the P time
    r/vp, and around the S time r/vs the tangential motion stays below 3 % of
    the P peak.  A point force at the same node in the same medium puts
    tangential (S) motion there above 30 % of its P peak, so the window can
    see one."""
    nel, edge = 480, 5.0
    H = nel * edge

    def model(depth, y):
        n = depth.size
        return workloads.isotropic_model(np.full(n, 2000.0), np.full(n, 3000.0), np.full(n, 1732.0))

    steps, dt = 1600, 2.5e-4
    case = workloads.structured_case(host.Q4, nel, nel, edge, model, 20, dt, 10.0, 0.12, steps, 0.0,
                                     H / 2, 1, 0.0, H, None)
    prob = case["problem"]
    x, y = prob["x"], prob["y"]
    xc = yc = H / 2
    R = 400.0
    ang = np.linspace(0.0, 2 * np.pi, 24, endpoint=False)
    rec = np.array([host.node_location(edge, x, y, xc + R * np.cos(a), yc + R * np.sin(a)) for a in ang], np.int32)
    dx, dy = x[rec] - xc, y[rec] - yc
    r = np.hypot(dx, dy)
    ex = workloads.explosive_source(host.Q4, prob["element_node"], x, y, xc, yc, moment=1.0)
    centre = np.array([host.node_location(edge, x, y, xc, yc)], np.int32)
    assert x[centre[0]] == xc and y[centre[0]] == yc
    op = gpu.Operator(prob["node_num"], prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    say("explosive source in a homogeneous Q4 480x480 medium: radial, P only")
    out = {}
    for name, src in (("explosive", [ex]), ("force", None)):
        run = gpu.Run(op, md, centre, steps, rec, dt, 10.0, 0.12, precision=gpu.FP64, sources=src)
        run.advance(steps)
        out[name] = run.seismograms()
        run.close()
    md.close()
    op.close()
    t = (np.arange(steps) + 3) * dt - 0.12                       # time after the wavelet's peak
    p_win = np.abs(t - R / 3000.0) < 0.05
    s_win = np.abs(t - R / 1732.0) < 0.03

    def parts(name):
        u, w = out[name][0][0], out[name][1][0]                  # [steps][rec]: u along x, w along y
        return (u * dx + w * dy) / r, (-u * dy + w * dx) / r

    radial, tangential = parts("explosive")
    f_rad, f_tan = parts("force")
    peak_p = np.max(np.abs(radial[p_win]))
    ratio = np.linalg.norm(tangential) / np.linalg.norm(radial)
    s_ratio = np.max(np.abs(tangential[s_win])) / peak_p
    f_ratio = np.max(np.abs(f_tan[s_win])) / np.max(np.abs(f_rad[p_win]))
    say(f"  tangential / radial {ratio:.3e}; S-window tangential / P peak {s_ratio:.3e}; point force {f_ratio:.3f}")
    assert np.argmax(np.max(np.abs(radial), axis=1)) in np.flatnonzero(p_win)
    assert ratio < 0.03 and s_ratio < 0.03
    assert f_ratio > 0.3


def test_config5_network_gradients_and_adam_step():
    """13 -> 1024 -> 1024 -> 1024 -> 32000 (configuration 5's shape), batch 32:
    loss and every layer's gradients against the MLP oracle (same bf16 rounding
    points, float64 arithmetic), then one Adam step."""
    B = 32
    net = surrogate.MLP(13, 1024, 3, 32000, batch=B, seed=7)
    assert net.params == 13 * 1024 + 1024 + 2 * (1024 * 1024 + 1024) + 1024 * 32000 + 32000
    rng = np.random.default_rng(5)
    X = rng.uniform(0, 1, (B, 13)).astype(np.float32)
    T = rng.standard_normal((B, 32000)).astype(np.float32)
    W0, b0 = zip(*[net.layer(l) for l in range(net.nlayers)])
    loss_ref, gW, gb = mo.loss_and_grads([w.astype(np.float64) for w in W0], [b.astype(np.float64) for b in b0],
                                         X, T, round_bf16=True)
    say("config 5 network 13-1024x3-32000: gradients of 4 layers vs the oracle, batch 32")
    for l in range(net.nlayers):
        loss, g, b = net.gradients(X, T, l)
        assert abs(loss - loss_ref) <= 1e-4 * loss_ref
        assert rel_l2(g.astype(np.float64), gW[l]) < 2e-3, l
        assert rel_l2(b.astype(np.float64), gb[l]) < 2e-3, l
    net.set_data(X, T)
    say("config 5 network: one Adam step")
    net.train_step(np.arange(B), 1e-3)
    for l in range(net.nlayers):
        W1, b1 = net.layer(l)
        wr, _, _ = mo.adam_step(W0[l].astype(np.float64), gW[l], 0.0, 0.0, 1, 1e-3)
        dw, dr = W1 - W0[l], wr - W0[l]
        big = np.abs(gW[l]) > 1e-3 * np.max(np.abs(gW[l]))
        assert np.mean(np.sign(dw[big]) == np.sign(dr[big])) > 0.99, l
        assert rel_l2(dw, dr) < 5e-2, l
    net.close()
