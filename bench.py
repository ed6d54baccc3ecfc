"""Benchmark: DOF-steps/s of the explicit elastic M-PML masslump loop on one MI355X.

Workload (default): 128 shots sharing one Q4 400x400 mesh (dx = 10 m,
160 801 nodes, 321 602 DOF per shot), 6-layer isotropic model, M-PML 30,
dt = 0.5 ms, executed steps drawn U[1500, 3000] per shot (seeded), FP32 state
with a Kahan-compensated displacement update whose compensation terms are kept
as bfloat16 -- the batched multi-shot configuration.  Inputs are resident in
HBM before the timed region.

One bench step is one pass of the time loop over the whole job: all 128 shots
from rest to their own last step (one step launch per time step, 2 995 with
the seeded step counts).
--warmup W untimed passes, then --steps K timed passes; the run is reset to
rest (sfg_run_reset) at the start of every pass.

  value        whole-job DOF-steps/s over the K timed passes (wall clock, device synchronised)
  roofline     the step kernel: algorithmic bytes per launch / average launch time
               (HIP events on the kernel's stream), against 8 TB/s
  fp32_kahan_terms  the same job with the compensation terms in FP32
  fp64_single  the plain single-shot FP64 loop (bit-identical to the reference) on the same mesh
  fp64_batched the same job in FP64 (bit-identical per shot), the accuracy check's reference
  accuracy     max rel-L2 (u and w concatenated, every executed step) of the FP32 seismograms of
               the last timed pass vs the FP64 loop, over all 128 shots at full length
  multi_model  configuration 4 part 2: the 128 shots plus 16 Q4 models of uneven size in one
               launch per step, each model checked against its FP64 run
  config2      Q9 400x400 VTI single shot (configuration 2): FP64 and FP32+Kahan step kernels
  surrogate_network  configuration 5's network: Adam step time, inference rate, output-layer GEMM
               TFLOP/s against the bf16 MFMA peak
  cpu_baseline the C restatement of the reference loop (oracle) on the host's physical cores,
               step loop timed alone, with and without the reference's %f text output

python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STEP_RANGE = (1500, 3000)   # executed steps per shot, drawn uniformly (configuration 4)
# HBM bytes per step-kernel launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled by a copy kernel of known size); written by the profiling runs
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_kernel_traffic.json")


def measured_traffic(workload_key):
    """(bytes per launch, launch shape it was measured with) or (None, None)."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None, None
    return d.get(workload_key), d.get("_about", {}).get(workload_key, {}).get("launch")


def log(msg):
    print("[bench]", msg, file=sys.stderr, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cpu_share():
    """Physical cores this process may use, capped at the 16-core share of one GPU."""
    from oracle.cpu_baseline import physical_cores
    phys, _ = physical_cores()
    return max(1, min(phys, 16))


def cpu_baseline(nelem, seconds):
    """The oracle's loop in a child process (no GPU), OpenMP bound to physical cores."""
    threads = cpu_share()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--nelem", str(nelem), "--seconds", str(seconds)]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=max(120.0, 8 * seconds))
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def pinned_launch(spec, x, y):
    """gpu.Run keywords of a pinned launch 'shots_per_lane_group,block,variant,xcd[,order[,window]]'
    (xcd: 1 contiguous ranges, 0 round-robin, Q > 1 stripes of Q row blocks; order: 0 node,
    1 Hilbert, or patchWxH; window: the sliced layouts' sorting window)."""
    from seisfem_amd import host
    v = spec.split(",")
    order, name = None, "node"
    if len(v) > 4 and v[4] not in ("", "0"):
        if v[4].startswith("patch"):
            w, h = (int(q) for q in v[4][5:].split("x"))
            order, name = host.patch_order(x, y, w, h), v[4]
        else:
            order, name = host.locality_order(x, y), "hilbert"
    xcd = int(v[3])
    return dict(shots_per_lane_group=int(v[0]), block=int(v[1]), kernel_variant=int(v[2]),
                xcd_remap=xcd if xcd > 1 else bool(xcd), row_order=order, row_order_name=name,
                sell_sigma=int(v[5]) if len(v) > 5 else 0)


def config2_leg(args):
    """Configuration 2 (Q9 400x400 VTI, one shot): FP64 (reference arithmetic) and
    FP32+Kahan step kernels, each at its tuned (or --launch2) launch shape."""
    from seisfem_amd import autotune, gpu, workloads
    t_c = time.time()
    c2 = workloads.config2(nelem=args.nelem, steps_exec=args.config2_steps)
    p2 = c2["problem"]
    n2 = p2["node_num"]
    op2 = gpu.Operator(n2, p2["csr"])
    md2 = gpu.Medium(op2, p2["c"], p2["mass_lump"], p2["dirichlet"], p2["mpml"])
    log(f"config 2 setup {time.time() - t_c:.1f}s: Q9 {args.nelem}x{args.nelem}, N={n2}, nnz={op2.nnz}")
    out = {"workload": f"Q9 {args.nelem}x{args.nelem} VTI, 1 shot, first {args.config2_steps} steps",
           "nodes": n2, "nnz_union": op2.nnz}
    b2 = dict(dt=c2["dt"], f0=c2["f0"], t0=c2["t0"])
    for name, pr in (("fp64", gpu.FP64), ("fp32_kahan", gpu.FP32_KAHAN)):
        if args.launch2:
            best2 = pinned_launch(args.launch2, p2["x"], p2["y"])
        else:
            best2, _ = autotune.tune(op2, md2, c2["src"], c2["rec"], precision=pr, steps=20, log=log,
                                     coords=(p2["x"], p2["y"]), **b2)
        r2 = gpu.Run(op2, md2, c2["src"], args.config2_steps, c2["rec"], precision=pr,
                     **dict(b2, **autotune.run_kwargs(best2)))
        r2.advance(args.config2_steps)
        ms2 = r2.last_ms_per_step
        by2 = r2.step_bytes(1)
        r2.close()
        out[name] = {"ms_per_launch": ms2, "dof_steps_per_s": 2.0 * n2 / (ms2 * 1e-3),
                     "bytes_per_launch": by2, "achieved_GBs": by2 / (ms2 * 1e-3) / 1e9,
                     "frac": by2 / (ms2 * 1e-3) / 1e9 / HBM_PEAK_GBS, "tuned": autotune.describe(best2)}
        log(f"config 2 {name}: {ms2:.4f} ms/launch, {out[name]['dof_steps_per_s'] / 1e9:.2f} G DOF-steps/s, "
            f"{out[name]['achieved_GBs']:.0f} GB/s ({out[name]['frac']:.2f} of peak)")
    md2.close()
    op2.close()
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16


def net_params(widths):
    return sum(a * b + b for a, b in zip(widths[:-1], widths[1:]))


def mlp_leg():
    """Configuration 5's network (13 -> 1024 x 3 -> 32000, batch 256) on synthetic
    data: Adam training-step time, bf16 inference rate, and the output layer's
    forward GEMM (256 x 32000 x 1024) timed alone against the bf16 MFMA peak."""
    from seisfem_amd import surrogate
    rng = np.random.default_rng(0)
    nin, hid, nh, nout, B = surrogate.NIN, 1024, 3, 32000, 256
    X = rng.uniform(0.0, 1.0, (16 * B, nin)).astype(np.float32)
    Y = rng.standard_normal((16 * B, nout)).astype(np.float32)
    net = surrogate.MLP(nin, hid, nh, nout, batch=B, seed=1)
    net.set_data(X, Y)
    net.train(1, lr=1e-3)
    _, ms = net.train(4, lr=1e-3)
    _, pms = net.predict(X)
    net.close()
    widths = [nin] + [hid] * nh + [nout]
    macs = sum(a * b for a, b in zip(widths[:-1], widths[1:]))
    # forward 2 flops per MAC; backward: dW for every layer and dX for all but the first
    train_flops = 2.0 * B * (3 * macs - widths[0] * widths[1])
    # the GEMM alone, back to back: 100 untimed launches (the clock settles over the
    # first few hundred), then the average of 500
    surrogate.gemm_bf16_time(B, nout, hid, reps=100)
    g_ms = surrogate.gemm_bf16_time(B, nout, hid, reps=500)
    g_tf = 2.0 * B * nout * hid / (g_ms * 1e-3) / 1e12
    out = {"network": f"{nin}-{hid}x{nh}-{nout}, batch {B}, synthetic data", "params": net_params(widths),
           "train_ms_per_step": ms, "train_TFLOPs": train_flops / (ms * 1e-3) / 1e12,
           "predict_samples_per_s": X.shape[0] / (pms * 1e-3),
           "output_gemm": {"shape": f"{B}x{nout}x{hid}", "ms": g_ms, "TFLOPs": g_tf,
                           "frac": g_tf / MFMA_BF16_PEAK_TFLOPS}}
    log(f"network: {ms:.3f} ms per Adam step, {out['predict_samples_per_s']:.3g} samples/s inference, "
        f"output GEMM {g_ms * 1e3:.1f} us = {g_tf:.0f} TFLOP/s ({out['output_gemm']['frac']:.2f} of bf16 peak)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed passes over the whole job")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes over the whole job")
    ap.add_argument("--nelem", type=int, default=400)
    ap.add_argument("--shots", type=int, default=128)
    ap.add_argument("--precision", default="fp32_kahan16", choices=["fp64", "fp32", "fp32_kahan", "fp32_kahan16"],
                    help="fp32_kahan16: FP32 with the Kahan terms kept as bfloat16 (the default); fp32_kahan: "
                         "the terms in FP32")
    ap.add_argument("--no-kahan32", action="store_true",
                    help="skip the pass with FP32 Kahan terms (run after an fp32_kahan16 job)")
    ap.add_argument("--fp64-steps", type=int, default=500, help="steps of the single-shot FP64 rate")
    ap.add_argument("--acc-shots", type=int, default=0, help="shots checked against FP64 (0 = all)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="host time of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-acc", action="store_true")
    ap.add_argument("--no-multi", action="store_true", help="skip configuration 4 part 2 (16 models, same launch)")
    ap.add_argument("--no-config2", action="store_true", help="skip the Q9 400x400 single-shot leg")
    ap.add_argument("--config2-steps", type=int, default=400)
    ap.add_argument("--config2-only", action="store_true",
                    help="run only the Q9 400x400 leg and print its JSON (profiling runs)")
    ap.add_argument("--launch2", default="",
                    help="pin the config-2 launch of both forms (same form as --launch; no tuning)")
    ap.add_argument("--quick", action="store_true", help="small mesh and step counts (smoke of the bench itself)")
    ap.add_argument("--no-mlp", action="store_true", help="skip the configuration-5 network leg")
    ap.add_argument("--mlp-only", action="store_true", help="only the configuration-5 network leg")
    ap.add_argument("--launch", default="",
                    help="pin the batched launch 'shots_per_lane_group,block,variant,xcd[,order[,window]]' (no "
                         "tuning); xcd: 1 contiguous ranges, 0 round-robin, Q > 1 stripes of Q row blocks; order: "
                         "0 node, 1 Hilbert, patchWxH patches of W nodes x H grid lines")
    ap.add_argument("--launch64", default="", help="pin the FP64 single-shot launch shape (same form)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {args.gpus} requested); "
              "shots would shard across GPUs with no communication, but no multi-GPU figure is produced.",
              file=sys.stderr)
        sys.exit(2)
    if args.mlp_only:
        print(json.dumps({"surrogate_network": mlp_leg()}), flush=True)
        return
    if args.config2_only:
        from seisfem_amd import gpu
        # the copy kernel whose byte counters calibrate the profiler's (FETCH_SIZE, WRITE_SIZE)
        copy_gbs, _ = gpu.stream_copy(1 << 30, 5)
        print(json.dumps({"config2": config2_leg(args), "measured_copy_GBs": copy_gbs}), flush=True)
        return
    step_range = STEP_RANGE
    if args.quick:
        args.nelem, args.fp64_steps, args.cpu_seconds, args.config2_steps = 100, 100, 2.0, 50
        step_range = (100, 200)
    K, W = max(args.steps, 1), max(args.warmup, 0)

    from seisfem_amd import autotune, gpu, host, workloads

    # measured copy bandwidth of this device (1 GiB buffers), reported beside the nominal peak
    copy_gbs, copy_ms = gpu.stream_copy(1 << 30, 20)
    log(f"stream copy 1 GiB: {copy_gbs:.0f} GB/s ({copy_ms:.3f} ms per copy)")

    t_setup = time.time()
    case = workloads.config4a(nshots=args.shots, nelem=args.nelem, step_range=step_range)
    prob = case["problem"]
    N = prob["node_num"]
    nsteps = case["nsteps"]
    L = int(nsteps.max())                       # launches of one pass over the job
    setup_s = time.time() - t_setup
    log(f"setup {setup_s:.1f}s: Q4 {args.nelem}x{args.nelem}, N={N}, shots={args.shots}, "
        f"steps {nsteps.min()}..{L} (drawn U[{step_range[0]},{step_range[1]}])")
    prec = gpu.PRECISION[args.precision]
    t_up = time.time()
    op = gpu.Operator(N, prob["csr"])
    md = gpu.Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    upload_s = time.time() - t_up
    base = dict(dt=case["dt"], f0=case["f0"], t0=case["t0"])
    # launch shape chosen by measurement on this problem (untimed)
    coords = (prob["x"], prob["y"])

    if args.launch:
        best = pinned_launch(args.launch, *coords)
    else:
        best, _ = autotune.tune(op, md, case["src"], case["rec"], precision=prec, steps=30, log=log, coords=coords,
                                **base)
    if args.launch64:
        best64 = pinned_launch(args.launch64, *coords)
    else:
        best64, _ = autotune.tune(op, md, case["src"][:1], case["rec"], precision=gpu.FP64, steps=30, log=log,
                                  coords=coords, **base)
    log(f"tuned batched {autotune.describe(best)}; tuned fp64 single {autotune.describe(best64)}")
    kw = dict(base, **autotune.run_kwargs(best))
    kw64 = dict(base, **autotune.run_kwargs(best64))

    # the step kernel with every shot active (untimed for the figure): 2 x 20 launches
    warm = gpu.Run(op, md, case["src"], 40, case["rec"], precision=prec, **dict(kw, guard_every=20))
    warm.advance(20)
    warm.advance(20)
    ms_all_active = warm.last_ms_per_step
    warm.close()

    # the job: W untimed passes, then K timed passes, each from rest to every shot's last step
    run = gpu.Run(op, md, case["src"], L, case["rec"], precision=prec, nsteps=nsteps, **kw)
    for _ in range(W):
        run.reset()
        run.advance(L)
    run.sync()
    launch_ms = []
    t0 = time.perf_counter()
    for _ in range(K):
        run.reset()
        run.advance(L)                          # returns after the stream has drained
        launch_ms.append(run.last_ms_per_step)  # HIP events around the pass's L launches
    run.sync()
    wall = time.perf_counter() - t0
    dof_job = 2.0 * N * float(nsteps.sum())
    value = K * dof_job / wall
    ms_launch = float(np.mean(launch_ms))
    log(f"batched {args.precision}: {K} passes of {L} launches in {wall:.2f}s ({1e3 * wall / K:.1f} ms per pass), "
        f"{value / 1e9:.2f} G DOF-steps/s; step kernel {ms_launch:.4f} ms per launch (device)")
    # roofline of the step kernel over the same launches: algorithmic bytes of launch t are
    # those of the shots still active at t; the guard kernel (one per guard_every launches,
    # ~0.1% of the time) is inside the event interval
    active = [(nsteps > t).sum() for t in range(L)]
    bytes_launch = float(np.mean([run.step_bytes(a) for a in active]))
    achieved = bytes_launch / (ms_launch * 1e-3) / 1e9
    bytes_full = run.step_bytes(args.shots)
    seis_u, seis_w = run.seismograms()
    run.close()
    log(f"step kernel: {bytes_launch / 1e9:.3f} GB algorithmic per launch -> {achieved:.0f} GB/s; "
        f"all shots active: {ms_all_active:.4f} ms, {bytes_full / (ms_all_active * 1e-3) / 1e9:.0f} GB/s")

    # the same job with the Kahan terms in FP32 (same launch shape), one pass
    k32 = None
    if args.precision == "fp32_kahan16" and not args.no_kahan32:
        rk = gpu.Run(op, md, case["src"], L, case["rec"], precision=gpu.FP32_KAHAN, nsteps=nsteps, **kw)
        rk.advance(L)
        ms_k = rk.last_ms_per_step
        k32_u, k32_w = rk.seismograms()
        bytes_k = float(np.mean([rk.step_bytes(a) for a in active]))
        rk.close()
        k32 = {"precision": "fp32_kahan", "value": dof_job / (ms_k * L * 1e-3), "ms_per_launch": ms_k,
               "roofline_achieved_GBs": bytes_k / (ms_k * 1e-3) / 1e9, "bytes_per_launch": bytes_k,
               "timing": "device time of one pass (HIP events)"}
        log(f"batched fp32_kahan (FP32 Kahan terms): {ms_k:.3f} ms/launch, {k32['value'] / 1e9:.2f} G DOF-steps/s")

    # plain single-shot FP64 loop on the same mesh (the reference-exact path)
    r64 = gpu.Run(op, md, case["src"][:1], args.fp64_steps, case["rec"], precision=gpu.FP64, **kw64)
    r64.advance(args.fp64_steps)
    ms64 = r64.last_ms_per_step
    r64.close()
    fp64_rate = 2.0 * N / (ms64 * 1e-3)
    log(f"fp64 single shot: {ms64:.4f} ms/step -> {fp64_rate / 1e9:.2f} G DOF-steps/s")

    # accuracy gate: FP32(+Kahan) seismograms vs the FP64 loop, every shot at its full length
    acc = None
    acc_median = None
    fp64_batched = None
    if not args.no_acc:
        nacc = args.acc_shots if 0 < args.acc_shots < args.shots else args.shots
        idx = np.unique(np.linspace(0, args.shots - 1, nacc).astype(int))
        # the reference's arithmetic for all these shots in one run, same launch shape as the
        # batched job (every variant and tile map gives the same bits)
        ra = gpu.Run(op, md, case["src"][idx], L, case["rec"], precision=gpu.FP64, nsteps=nsteps[idx], **kw)
        ra.advance(L)
        ms_a = ra.last_ms_per_step
        u64, w64 = ra.seismograms()
        ra.close()
        dof_a = 2.0 * N * float(nsteps[idx].sum())
        fp64_batched = {"value": dof_a / (ms_a * L * 1e-3), "ms_per_launch": ms_a, "shots": int(len(idx))}
        log(f"batched fp64 (reference arithmetic, {len(idx)} shots): {ms_a:.3f} ms/launch, "
            f"{fp64_batched['value'] / 1e9:.2f} G DOF-steps/s")

        def shot_errors(su, sw):
            errs = []
            for q, s in enumerate(idx):
                k = int(nsteps[s])
                a = np.concatenate([su[s, :k].ravel(), sw[s, :k].ravel()])
                b = np.concatenate([u64[q, :k].ravel(), w64[q, :k].ravel()])
                errs.append(rel_l2(a, b))
            return errs

        errs = shot_errors(seis_u, seis_w)
        acc = max(errs)
        acc_median = float(np.median(errs))
        log(f"accuracy: max rel-L2 vs FP64 over {len(idx)} shots, steps {int(nsteps[idx].min())}..."
            f"{int(nsteps[idx].max())}: {acc:.3e} (median {acc_median:.3e})")
        if k32 is not None:
            e32 = shot_errors(k32_u, k32_w)
            k32["accuracy_max_rel_l2"] = max(e32)
            log(f"accuracy fp32_kahan: max rel-L2 {max(e32):.3e} (median {np.median(e32):.3e})")
    if k32 is not None:
        del k32_u, k32_w

    # configuration 4 part 2 in the same launch as part 1: 16 Q4 models of uneven size
    multi = None
    if not args.no_multi:
        t_m = time.time()
        models = workloads.config4b(step_range=step_range)
        log(f"multi-model setup {time.time() - t_m:.1f}s: sizes {[c['size'] for c in models]}")
        keep, runs = [], []
        # the multi-run kernel takes lane groups of 1, 32 or 64 shots; a lane's arithmetic does
        # not depend on the grouping, so part 1 stays bit-identical to the single run
        sp1 = kw.get("shots_per_lane_group", 64)
        part1 = gpu.Run(op, md, case["src"], L, case["rec"], precision=prec, nsteps=nsteps,
                        **dict(kw, shots_per_lane_group=sp1 if sp1 in (32, 64) else 64))
        runs.append(part1)
        for c in models:
            p = c["problem"]
            mop = gpu.Operator(p["node_num"], p["csr"])
            mmd = gpu.Medium(mop, p["c"], p["mass_lump"], p["dirichlet"], p["mpml"])
            keep += [mop, mmd]
            # single-shot models in the sliced layout (round-robin tiles), stepped by a
            # launch of single-shot forms of their own, after part 1's own step launch
            runs.append(gpu.Run(mop, mmd, c["src"], L, c["rec"], c["dt"], c["f0"], c["t0"], precision=prec,
                                nsteps=c["nsteps"], block=kw.get("block", 256), shots_per_lane_group=1,
                                kernel_variant=8, xcd_remap=False))
        grp = gpu.Multi(runs)
        grp.advance(L)
        ms_m = grp.last_ms_per_step
        dof_m = dof_job + sum(2.0 * c["problem"]["node_num"] * int(c["nsteps"][0]) for c in models)
        # every part-2 model against its own FP64 run (the 16 FP64 runs share one launch too)
        r64s = [gpu.Run(keep[2 * qi], keep[2 * qi + 1], c["src"], L, c["rec"], c["dt"], c["f0"], c["t0"],
                        precision=gpu.FP64, nsteps=c["nsteps"]) for qi, c in enumerate(models)]
        g64 = gpu.Multi(r64s)
        g64.advance(L)
        errs = []
        for qi, c in enumerate(models):
            u32, w32 = runs[1 + qi].seismograms()
            u64m, w64m = r64s[qi].seismograms()
            k = int(c["nsteps"][0])
            errs.append(rel_l2(np.concatenate([u32[0, :k].ravel(), w32[0, :k].ravel()]),
                               np.concatenate([u64m[0, :k].ravel(), w64m[0, :k].ravel()])))
        g64.close()
        for r in r64s:
            r.close()
        # part-1 shots of the multi launch against the single-run result of the same steps
        u1, w1 = part1.seismograms()
        p1_same = bool(np.array_equal(u1, seis_u) and np.array_equal(w1, seis_w))
        bytes_m = float(np.mean([grp.step_bytes(t) for t in range(L)]))
        multi = {"value": dof_m / (ms_m * L * 1e-3), "ms_per_launch": ms_m, "models": len(models),
                 "achieved_GBs": bytes_m / (ms_m * 1e-3) / 1e9, "frac": bytes_m / (ms_m * 1e-3) / 1e9 / HBM_PEAK_GBS,
                 "sizes": [list(c["size"]) for c in models], "max_rel_l2_vs_fp64": max(errs),
                 "part1_identical_to_single_run": p1_same,
                 "launch": "per step part 1's own step launch (128 shots) and one launch of the 16 "
                           "single-shot models (sliced layout), both in one hipGraph"}
        log(f"multi-model launch: {ms_m:.3f} ms/launch, {multi['value'] / 1e9:.2f} G DOF-steps/s, "
            f"{multi['achieved_GBs']:.0f} GB/s, rel-L2 {max(errs):.2e}")
        grp.close()
        for r in runs:
            r.close()
        # part 2 alone: the graph-replayed per-step launch against the persistent
        # schedule (cooperative launch per guard interval, tile queues, grid barrier)
        def part2_runs():
            return [gpu.Run(keep[2 * qi], keep[2 * qi + 1], c["src"], L, c["rec"], c["dt"], c["f0"], c["t0"],
                            precision=prec, nsteps=c["nsteps"], block=kw.get("block", 256), shots_per_lane_group=1,
                            kernel_variant=8, xcd_remap=False) for qi, c in enumerate(models)]
        best_s, tab_s = autotune.tune_multi_schedule(part2_runs, steps=min(300, L // 2), log=log)
        multi["part2_schedule"] = {"ms_per_step": {("per_step" if k == 0 else f"persistent_{k}_per_cu"): ms
                                                   for k, ms in tab_s},
                                   "best": "per_step" if best_s == 0 else f"persistent_{best_s}_per_cu"}
        for h in keep:
            h.close()

    # configuration 2: Q9 400x400 VTI, one shot, FP64 (reference arithmetic) and FP32+Kahan
    config2 = None if args.no_config2 else config2_leg(args)
    network = None if (args.no_mlp or args.quick) else mlp_leg()

    # CPU baseline: the oracle (C restatement of the reference loop), bounded sample, child process
    cpu = None
    if not args.no_cpu:
        cb = cpu_baseline(args.nelem, args.cpu_seconds)
        cpu = {"value": cb["dof_steps_per_s_no_io"], "unit": "DOF-steps/s", "cores": cb["threads"], "kind": "port",
               "cpu_model": cb["cpu_model"], "physical_cores_visible": cb["physical_cores"],
               "omp": f"OMP_PROC_BIND={cb['omp_proc_bind']} OMP_PLACES={cb['omp_places']}",
               "with_text_io": {"value": cb["dof_steps_per_s_with_io"], "ms_per_step": cb["ms_per_step_with_io"],
                                "bytes_written": cb["io_bytes"]},
               "reference_form": {"value": cb["dof_steps_per_s_reference_form"],
                                  "ms_per_step": cb["ms_per_step_reference_form"],
                                  "what": "with the text output and the reference's shared, unreduced energy "
                                          "sums in the parallel update loop (elastic_wave.c:776-784)"},
               "ms_per_step": cb["ms_per_step_no_io"],
               "sample": f"FP64, one shot of the same Q4 {args.nelem}x{args.nelem} mesh and model, first "
                         f"{cb['steps']} steps; step loop alone timed with omp_get_wtime (arrays allocated and "
                         f"touched before it), 44 csr_matvec calls per step as the reference issues them, "
                         f"OpenMP over rows; 'value' without output, 'with_text_io' with the reference's %f "
                         f"seismogram ({cb['rec_num']} receivers) and wavefield text files"}
        log(f"cpu baseline: {cpu['value'] / 1e6:.1f} M DOF-steps/s on {cb['threads']} threads "
            f"({cb['ms_per_step_no_io']:.2f} ms/step; with %f output {cb['ms_per_step_with_io']:.2f} ms/step)")

    # L2-miss bytes per step-kernel launch from the committed rocprofv3 runs, and the
    # launch shape ("shots_per_lane_group,block,variant,xcd") they were measured with
    traffic, traffic_launch = measured_traffic(f"{args.precision}_{args.shots}x_q4_{args.nelem}")
    out = {
        "metric": "DOF-steps/s + %HBM peak, batched FP32 vs plain FP64, at <=1e-4 rel-L2 seismograms",
        "value": value,
        "unit": "DOF-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp32_kahan": "f32", "fp32_kahan16": "f32"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"{args.shots} shots x Q4 {args.nelem}x{args.nelem} (dx=10 m), uneven executed steps "
                               f"U[{step_range[0]},{step_range[1]}] per shot (drawn {int(nsteps.min())}..{L}), "
                               f"{args.precision}; one step = one pass over the whole job ({L} launches)",
                   "nodes": N, "dof_per_shot": 2 * N, "shots": args.shots, "precision": args.precision,
                   "launches_per_step": L, "dof_steps_per_step": dof_job,
                   "parallelism": "1 GPU", "host_setup_s": setup_s, "upload_s": upload_s,
                   "tuned": autotune.describe(best),
                   "tuned_fp64": autotune.describe(best64)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_launch": traffic_launch,
                     "kernel": "step_kernel (fused SpMM + M-PML RHS + lumped solve + Newmark)",
                     "bytes": "algorithmic bytes of the reference form (26 state fields read and written per "
                              "node and shot, plus the Kahan terms of the form run); 'traffic' is what the "
                              "kernel moves across L2 per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE)",
                     "ms_per_launch": ms_launch, "bytes_per_launch": bytes_launch,
                     "measured_copy_GBs": copy_gbs, "frac_of_measured_copy": achieved / copy_gbs,
                     "traffic_GBs": None if traffic is None else traffic / (ms_launch * 1e-3) / 1e9,
                     "all_shots_active": {"ms_per_launch": ms_all_active, "bytes_per_launch": bytes_full,
                                          "achieved": bytes_full / (ms_all_active * 1e-3) / 1e9}},
        "fp64_single": {"value": fp64_rate, "ms_per_launch": ms64, "speedup_batched_vs_fp64": value / fp64_rate},
        "fp32_kahan_terms": None if k32 is None else dict(k32, speedup_vs_fp64=k32["value"] / fp64_rate),
        "fp64_batched": fp64_batched,
        "accuracy_max_rel_l2": acc,
        "accuracy_median_rel_l2": acc_median,
        "accuracy_covers": None if acc is None else f"every executed step of every shot ({int(nsteps.min())}..{L})",
        "multi_model": multi,
        "config2": config2,
        "surrogate_network": network,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    md.close()
    op.close()


if __name__ == "__main__":
    main()
