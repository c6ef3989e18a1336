#!/usr/bin/env python3
"""PD iterations/s of ICTGV2 at 256x256 x 30 frames x 16 coils, batched vs plain.

One step = one primal-dual iteration (dual update + projections, data-term
z update with K and K^H, primal update + extragradient, and the step-size rule
on its schedule k < 10 || k % 50 == 0) over every resident reconstruction.

* plain:   one reconstruction, the task-list kernels launched eagerly per
           iteration (no graph, no batching).
* batched: --batch reconstructions resident, advanced together by one
           task-list launch sequence per iteration, replayed from a hipGraph.

value = batched aggregate PD iterations/s (iterations x reconstructions per
second); inputs are resident in HBM before the timed region.  The roofline
entry is the dominant kernel's algorithmic bytes per launch over its average
duration, timed with HIP events on the engine's stream.  cpu_baseline is the
FP64 oracle (C, OpenMP) on the host cores for a few iterations of the same
reconstruction.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
                       [--dual fp32|bf16|q16] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_per_voxel(kernel, method, bf16):
    """Compulsory HBM bytes per voxel of the regulariser kernels.
    dual:   read ext (P), read y (D), write y (D)
    primal: read y (D), read K^H z (1), read x (P), write x (P), write ext (P)"""
    P = {"tv": 1, "tgv2": 4, "ictgv2": 8}[method]
    D = {"tv": 3, "tgv2": 9, "ictgv2": 18}[method]
    bp, bd = 8, (4 if bf16 else 8)
    if kernel == "dual":
        return P * bp + 2 * D * bd
    if kernel == "primal":
        return D * bd + bp + 3 * P * bp
    return None


def measured_traffic(workload, batch, dual, kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs recorded in profiles/traffic.json (counters scaled by a
    plain copy of known size), when that record is for this workload, batch
    and dual storage; otherwise None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if (rec.get("workload"), rec.get("batch"), rec.get("dual_storage")) != (workload, batch, dual):
        return None
    return rec.get("hbm_bytes_per_launch", {}).get(kernel)


def serving_leg(A, synth, torch, dev, rate, seconds, iterations=500):
    """Continuous batching under Poisson arrivals: mixed TV / TGV2 / ICTGV2
    Cartesian requests (n in 128..256, T 8..30, C 8..20, 500 iterations each).
    Request inputs are built on the device before the clock starts; a
    request's latency runs from its scheduled arrival to the moment its
    result has been copied out."""
    stream = synth.serving_stream(rate, seconds)
    pool = synth.serving_pool()
    reqs = []
    gen = torch.Generator(device=dev)
    gen.manual_seed(11)
    for (t_arr, method, n, T, C) in stream:
        P = pool[n]
        rho = torch.from_numpy(np.ascontiguousarray(P["rho"][:T], np.complex64)).to(dev)
        b1 = torch.from_numpy(np.ascontiguousarray(P["b1"][:C], np.complex64)).to(dev)
        mask = torch.from_numpy(np.ascontiguousarray(P["masks"][method][:T], np.float32)).to(dev)
        op = A.CartesianOperator(n, n, C, T, mask=mask, centered=False)
        data = op.BackwardOperation(rho, b1)
        sig = 0.01 * float(data.abs().max().item())
        noise = torch.randn(data.shape, dtype=torch.complex64, device=dev, generator=gen)
        data = data + sig * noise * mask[:, None]
        x0 = op.ForwardOperation(data, b1)
        m = {"tv": 0, "tgv2": 1, "ictgv2": 2}[method]
        reqs.append(dict(t=t_arr, method=method, n=n, T=T, C=C,
                         req=A.Request(m, data, b1, mask=mask, x0=x0,
                                       params={"maxIt": iterations})))
        del op
    torch.cuda.synchronize()
    eng = A.Engine(use_graphs=True)
    outs = {}
    lat, waiting, running = [], list(range(len(reqs))), {}
    t0 = time.perf_counter()
    steps = 0
    while waiting or running:
        now = time.perf_counter() - t0
        while waiting and reqs[waiting[0]]["t"] <= now:
            i = waiting[0]
            rid = eng.submit(reqs[i]["req"])
            if rid is None:
                break  # HBM budget full: retry after completions
            running[rid] = i
            waiting.pop(0)
        if not running:
            time.sleep(max(0.0, reqs[waiting[0]]["t"] - (time.perf_counter() - t0)))
            continue
        eng.step()
        steps += 1
        for rid, i in list(running.items()):
            r = reqs[i]
            if i not in outs:
                outs[i] = torch.empty((r["T"], r["n"], r["n"]), dtype=torch.complex64, device=dev)
            if eng.poll(rid, outs[i]):
                lat.append(time.perf_counter() - t0 - r["t"])
                del running[rid]
    makespan = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3
    total_its = iterations * len(reqs)
    vox_its = sum(r["n"] * r["n"] * r["T"] for r in reqs) * iterations
    finite = all(bool(torch.isfinite(torch.view_as_real(o)).all()) for o in outs.values())
    return {"requests": len(reqs), "rate_per_s": rate, "window_s": seconds,
            "iterations_each": iterations, "engine_steps": steps,
            "makespan_s": round(makespan, 3),
            "p50_latency_ms": round(float(np.percentile(lat_ms, 50)), 1),
            "p99_latency_ms": round(float(np.percentile(lat_ms, 99)), 1),
            "iters_per_s": round(total_its / makespan, 1),
            "voxel_iters_per_s": round(vox_its / makespan, 1),
            "all_finite": finite}


def radial_leg(A, synth, torch, dev, iters, dump_dir=None):
    """Non-Cartesian TGV2: golden-angle radial, 512 samples x 21 spokes per
    frame, 30 frames, 12 coils, 256x256, Kaiser-Bessel width 3 on a 1.5x grid,
    density |k|.  One reconstruction through the PDRecon API (eager launches);
    k-space = K rho + complex Gaussian noise (1% of max |K rho|), formed with
    the operator itself; u0 = K^H d."""
    wl = synth.RADIAL
    h = synth.make_radial_host(wl)
    f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.complex64)).to(dev)
    n, T, C = wl.n, wl.frames, wl.coils
    t0 = time.perf_counter()
    op = A.NoncartesianOperator(n, n, C, T, wl.n_fe, wl.spokes, f32(h["traj"]), f32(h["dcf"]),
                                kernel_width=wl.kernel_width, osf=wl.osf)
    t_setup = time.perf_counter() - t0
    rho, b1 = c64(h["rho"]), c64(h["b1"])
    data = op.BackwardOperation(rho, b1)
    gen = torch.Generator(device=dev)
    gen.manual_seed(wl.seed + 3000)
    sig = 0.01 * float(data.abs().max().item())
    data = data + sig * torch.randn(data.shape, dtype=torch.complex64, device=dev, generator=gen)
    x0 = op.ForwardOperation(data, b1)
    rec = A.TGV2(n, n, C, T, op, None)
    p = rec.GetParams()
    p.update(alpha0=2.0, alpha1=1.0, maxIt=10)
    rec.SetParams(p)
    x = x0.clone()
    rec.IterativeReconstruction(data, x, b1)  # warm-up
    p["maxIt"] = iters
    rec.SetParams(p)
    x = x0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec.IterativeReconstruction(data, x, b1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if dump_dir:
        os.makedirs(dump_dir, exist_ok=True)
        np.save(os.path.join(dump_dir, f"{wl.name}_{iters}_fp32.npy"), x.cpu().numpy())
        np.save(os.path.join(dump_dir, f"{wl.name}_data.npy"), data.cpu().numpy())
    return {"workload": wl.name, "method": "tgv2", "iterations": iters,
            "iters_per_s": round(iters / dt, 1), "ms_per_iter": round(1e3 * dt / iters, 4),
            "operator_setup_s": round(t_setup, 3), "lambda": round(p["lambda"], 6),
            "points_per_frame": wl.n_fe * wl.spokes, "grid": [round(wl.osf * n)] * 2}


def small_leg(A, synth, torch, dev, B, steps, warmup):
    """Batching where it matters: the TV 128x128x12 frames x 8 coils workload
    is too small to fill the GPU alone.  One reconstruction with eager
    launches against B of them advanced together from a replayed graph."""
    wl = synth.CONFIGS["tv_128x12x8"]
    inp = synth.make(wl)
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.complex64)).to(dev)
    data, b1, x0 = c64(inp["data"]), c64(inp["b1"]), c64(inp["x0"])
    mask = torch.from_numpy(np.ascontiguousarray(inp["mask"], np.float32)).to(dev)
    params = dict(wl.params, maxIt=warmup + steps + 10, adaptLambda=0)

    def run(use_graphs, n):
        eng = A.Engine(use_graphs=use_graphs, max_batch=n)
        for _ in range(n):
            eng.submit(A.Request(0, data, b1, mask=mask, x0=x0, params=params))
        for _ in range(warmup):
            eng.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.step()
        torch.cuda.synchronize()
        return n * steps / (time.perf_counter() - t0)

    plain = run(False, 1)
    batched = run(True, B)
    return {"workload": wl.name, "batch": B, "plain_iters_per_s": round(plain, 1),
            "batched_iters_per_s": round(batched, 1), "batched_over_plain": round(batched / plain, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--dual", choices=["fp32", "bf16", "q16"], default="q16",
                    help="dual storage: float32, bfloat16, or 16-bit fixed point scaled by the "
                         "projection radius (4 bytes per complex like bf16; the one of the two "
                         "4-byte formats that meets NRMSE <= 1e-4 at full size)")
    ap.add_argument("--workload", default="ictgv2_256x30x16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="arrival window of the serving stream (0: skip the serving leg)")
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--small-batch", type=int, default=64,
                    help="reconstructions of the TV 128x128x12x8 workload batched in the "
                         "small-request leg (0: skip)")
    ap.add_argument("--radial-iters", type=int, default=200,
                    help="iterations of the radial TGV2 leg (0: skip)")
    ap.add_argument("--dump-radial", default=None,
                    help="directory: save the radial leg's x and k-space data as .npy")
    ap.add_argument("--dump-x", default=None,
                    help="directory: also run one full reconstruction (the workload's iteration "
                         "count) with fp32 and with bf16 duals and save x as .npy, for the NRMSE "
                         "check against the FP64 oracle (python -m oracle.nrmse)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested). Reconstructions shard by request; run one "
              f"process per GPU instead.", file=sys.stderr)
        sys.exit(2)

    import torch

    import avionic_amd as A
    from avionic_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    wl = synth.CONFIGS[args.workload]
    log(f"generating {wl.name} inputs")
    inp = synth.make(wl)
    T, C, H, W = inp["data"].shape
    method = {"tv": 0, "tgv2": 1, "ictgv2": 2}[wl.method]
    dual_kind = {"fp32": 0, "bf16": 1, "q16": 2}[args.dual]
    bf16 = dual_kind != 0  # 4-byte duals

    def c64(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.complex64)).to(dev)

    data, b1, x0 = c64(inp["data"]), c64(inp["b1"]), c64(inp["x0"])
    mask = torch.from_numpy(np.ascontiguousarray(inp["mask"], np.float32)).to(dev)
    total_its = args.warmup + args.steps + 10
    params = dict(wl.params)
    params.update(maxIt=total_its, dualStorage=dual_kind)
    if "lambda" in wl.params:
        params["adaptLambda"] = 0

    def request():
        return A.Request(method, data, b1, mask=mask, x0=x0, centered=False, params=params)

    def timed(engine, n_req):
        for _ in range(args.warmup):
            engine.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            engine.step()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # ---- plain: one reconstruction, eager launches
    log("plain path")
    eng = A.Engine(use_graphs=False, max_batch=1)
    rid = eng.submit(request())
    t_plain = timed(eng, 1)
    plain_its = args.steps / t_plain
    del eng

    # ---- batched: B reconstructions, one graph-replayed launch sequence
    log(f"batched path (B={args.batch})")
    eng = A.Engine(use_graphs=True, max_batch=args.batch)
    ids = [eng.submit(request()) for _ in range(args.batch)]
    assert all(i is not None for i in ids)
    t_batch = timed(eng, args.batch)
    batched_its = args.batch * args.steps / t_batch
    ms_per_step = 1e3 * t_batch / args.steps
    del eng

    # ---- same result alone and in the mix (60 iterations covers the k<10
    #      and k%50 step-rule iterations)
    log("alone vs in-mix check")
    p60 = dict(params, maxIt=60)
    alone = A.Engine(use_graphs=False, max_batch=1)
    ra = alone.submit(A.Request(method, data, b1, mask=mask, x0=x0, params=p60))
    xa = torch.empty_like(x0)
    while not alone.poll(ra, xa):
        alone.step()
    mix = A.Engine(use_graphs=True, max_batch=4)
    other = dict(params, maxIt=90)
    mix.submit(A.Request(method, data, b1, mask=mask, x0=torch.zeros_like(x0), params=other))
    rm = mix.submit(A.Request(method, data, b1, mask=mask, x0=x0, params=p60))
    xm = torch.empty_like(x0)
    while not mix.poll(rm, xm):
        mix.step()
    torch.cuda.synchronize()
    mix_diff = float((xa - xm).abs().max().item())
    del alone, mix

    # ---- a plain device copy of known size, the reference point that scales
    #      rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes
    calib = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)  # 256 MiB
    calib_out = torch.empty_like(calib)
    calib_out.copy_(calib)
    torch.cuda.synchronize()
    del calib, calib_out

    # ---- per-kernel times (HIP events on the engine stream), fresh engine
    log("kernel timing")
    eng = A.Engine(use_graphs=False, max_batch=args.batch)
    for _ in range(args.batch):
        eng.submit(request())
    ktimes = A.engine_time_kernels(eng, reps=20)
    del eng
    N = W * H * T
    roof = None
    per_voxel = {k: algorithmic_bytes_per_voxel(k, wl.method, bf16) for k in ("dual", "primal")}
    dom = max(per_voxel, key=lambda k: ktimes[k])
    bytes_per_launch = per_voxel[dom] * N * args.batch
    achieved = bytes_per_launch / (ktimes[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(wl.name, args.batch, args.dual, dom),
            "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(ktimes[dom], 4)}

    small = None
    if args.small_batch > 0:
        log(f"small-request leg (tv_128x12x8, B={args.small_batch})")
        small = small_leg(A, synth, torch, dev, args.small_batch, args.steps, args.warmup)

    if args.dump_x:
        os.makedirs(args.dump_x, exist_ok=True)
        for ds in ("fp32", "bf16", "q16"):
            p_full = dict(params, maxIt=wl.iterations,
                          dualStorage={"fp32": 0, "bf16": 1, "q16": 2}[ds])
            e1 = A.Engine(use_graphs=True, max_batch=1)
            rid = e1.submit(A.Request(method, data, b1, mask=mask, x0=x0, params=p_full))
            xo = torch.empty_like(x0)
            while not e1.poll(rid, xo):
                e1.step()
            torch.cuda.synchronize()
            path = os.path.join(args.dump_x, f"{wl.name}_{wl.iterations}_{ds}.npy")
            np.save(path, xo.cpu().numpy())
            log(f"saved {path}")
            del e1

    radial = None
    if args.radial_iters > 0:
        log(f"radial TGV2 leg ({synth.RADIAL.name}, {args.radial_iters} iterations)")
        radial = radial_leg(A, synth, torch, dev, args.radial_iters, args.dump_radial)

    serving = None
    if args.serve_seconds > 0:
        log(f"serving stream ({args.serve_rate}/s for {args.serve_seconds}s)")
        serving = serving_leg(A, synth, torch, dev, args.serve_rate, args.serve_seconds)

    # ---- CPU baseline: FP64 oracle on the host cores, a few iterations
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (FP64 oracle)")
        import oracle as O
        cores = min(16, os.cpu_count() or 1)
        p = O.default_params(wl.method)
        k, d = O.adapt_lambda_defaults(wl.method)
        lam = wl.params.get("lambda") or O.adapt_lambda_cart(inp["mask"], W, H, T, k, d)
        p.update(maxIt=args.cpu_iters, lambda_=lam, **{kk: v for kk, v in wl.params.items()
                                                      if kk in ("alpha0", "alpha1")})
        t0 = time.perf_counter()
        O.recon(p, inp["data"], inp["b1"], inp["mask"], inp["x0"], nthreads=cores)
        dt = time.perf_counter() - t0
        cpu = {"value": round(args.cpu_iters / dt, 4), "unit": "PD iterations/s", "cores": cores,
               "kind": "port",
               "sample": f"{args.cpu_iters} PD iterations of one {wl.name} reconstruction "
                         f"(FP64 C oracle, OpenMP {cores} threads)"}

    out = {
        "metric": "PD iters/s ICTGV2 256\u00b2\u00d730\u00d716c batched vs plain @NRMSE\u22641e-4; "
                  "p50/p99 latency",
        "value": round(batched_its, 2),
        "unit": "PD iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" + ({0: "", 1: "+bf16-duals", 2: "+q16-duals"}[dual_kind]),
        "data": "synthetic",
        "config": {"workload": wl.name, "method": wl.method, "matrix": [H, W], "frames": T,
                   "coils": C, "accel": wl.accel, "batch": args.batch,
                   "dual_storage": args.dual, "parallelism": "single-gpu"},
        "plain_iters_per_s": round(plain_its, 2),
        "batched_over_plain": round(batched_its / plain_its, 3),
        "alone_vs_mix_max_abs_diff": mix_diff,
        "kernel_ms": {k: round(v, 4) for k, v in ktimes.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "small_requests": small,
        "radial": radial,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
