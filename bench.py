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
second); inputs are resident in HBM before the timed region.  The batched
requests then run to 1000 iterations and the first result's NRMSE against the
FP64 oracle's reconstruction of the same inputs (tests/golden/fullsize) is
reported with the whole-reconstruction rate.  The roofline entry is the
dominant kernel's algorithmic bytes per launch over its average duration,
timed with HIP events on the engine's stream, against the 8 TB/s peak and
against a 4 GiB device copy.  Legs: TGV2 256x256x25x12 plain and batched
with fp32 duals, the headline workload with fp32 duals, TV 128x128x12x8
batched, radial TGV2, and the serving stream.  cpu_baseline is the FP64
oracle (C, OpenMP) timed on the host cores over steady-state iterations.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
                       [--dual fp32|bf16|q16] [--no-cpu] [--no-legs]
"""
from __future__ import annotations

import argparse
import hashlib
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


def serving_leg(A, synth, torch, dev, rate, seconds, iterations=500, cap_bytes=0,
                max_queue=-1):
    """Continuous batching under Poisson arrivals: mixed TV / TGV2 / ICTGV2
    Cartesian requests (n in 128..256, T 8..30, C 8..20, 500 iterations each).
    Request inputs are built on the device before the clock starts.  Each
    request is submitted at its arrival time; the engine holds it, admits it
    in FIFO order while the running requests' bytes stay within cap_bytes,
    and queues up to max_queue more (submit refuses beyond: load shedding,
    counted as rejected).  A request's latency runs from its arrival to the
    moment its result has been copied out."""
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
    eng = A.Engine(use_graphs=True, resident_bytes_cap=int(cap_bytes), max_queue=max_queue)
    outs = {}
    lat, held, rejected = [], {}, 0
    nxt, steps, max_queued = 0, 0, 0
    t0 = time.perf_counter()
    while nxt < len(reqs) or held:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt]["t"] <= now:
            rid = eng.submit(reqs[nxt]["req"])
            if rid is None:
                rejected += 1
            else:
                held[rid] = nxt
            nxt += 1
        if not held:
            time.sleep(max(0.0, reqs[nxt]["t"] - (time.perf_counter() - t0)))
            continue
        eng.step()
        steps += 1
        max_queued = max(max_queued, eng.counters()["queued"])
        for rid, i in list(held.items()):
            r = reqs[i]
            if i not in outs:
                outs[i] = torch.empty((r["T"], r["n"], r["n"]), dtype=torch.complex64, device=dev)
            if eng.poll(rid, outs[i]):
                lat.append(time.perf_counter() - t0 - r["t"])
                del held[rid]
    makespan = time.perf_counter() - t0
    c = eng.counters()
    done = list(outs)
    lat_ms = np.array(lat) * 1e3 if lat else np.zeros(1)
    total_its = iterations * len(lat)
    vox_its = sum(reqs[i]["n"] ** 2 * reqs[i]["T"] for i in done) * iterations
    finite = all(bool(torch.isfinite(torch.view_as_real(o)).all()) for o in outs.values())
    return {"rate_per_s": rate, "window_s": seconds, "requests": len(reqs),
            "completed": len(lat), "rejected": rejected,
            "resident_bytes_cap": int(cap_bytes), "max_queue": max_queue,
            "max_queued": max_queued, "iterations_each": iterations, "engine_steps": steps,
            "makespan_s": round(makespan, 3),
            "p50_latency_ms": round(float(np.percentile(lat_ms, 50)), 1),
            "p99_latency_ms": round(float(np.percentile(lat_ms, 99)), 1),
            "iters_per_s": round(total_its / makespan, 1),
            "voxel_iters_per_s": round(vox_its / makespan, 1),
            "graph_captures": c["graph_captures"], "capture_ms": round(c["capture_ms"], 1),
            "all_finite": finite}


def radial_leg(A, synth, torch, dev, iters, batch=4):
    """Non-Cartesian TGV2: golden-angle radial, 512 samples x 21 spokes per
    frame, 30 frames, 12 coils, 256x256, Kaiser-Bessel width 3 on a 1.5x grid,
    density |k|.  k-space = K rho + complex Gaussian noise (1% of max |K rho|),
    formed with the operator itself; u0 = K^H d.  Plain: one reconstruction
    through the PDRecon API (eager launches).  Batched: `batch` reconstructions
    of the same geometry (one shared operator) through the engine, the
    regulariser as task-list launches and each request's gridding passes
    inside the replayed iteration graph."""
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
    # batched: the same requests through the engine
    prm = {"alpha0": 2.0, "alpha1": 1.0, "lambda": p["lambda"], "adaptLambda": 0,
           "maxIt": iters + 20}
    eng = A.Engine(use_graphs=True, max_batch=batch)
    for _ in range(batch):
        assert eng.submit(A.Request(1, data, b1, x0=x0, params=prm, op=op)) is not None
    for _ in range(10):
        eng.step()
    torch.cuda.synchronize()
    tb = time.perf_counter()
    for _ in range(iters):
        eng.step()
    torch.cuda.synchronize()
    dtb = time.perf_counter() - tb
    del eng
    return {"workload": wl.name, "method": "tgv2", "iterations": iters,
            "iters_per_s": round(iters / dt, 1), "ms_per_iter": round(1e3 * dt / iters, 4),
            "batch": batch, "batched_iters_per_s": round(batch * iters / dtb, 1),
            "batched_ms_per_step": round(1e3 * dtb / iters, 4),
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
    untimed = STEP_RULE_PHASE + TUNE_PHASE + warmup
    params = dict(wl.params, maxIt=untimed + steps + 10, adaptLambda=0)

    def run(use_graphs, n):
        eng = A.Engine(use_graphs=use_graphs, max_batch=n)
        for _ in range(n):
            eng.submit(A.Request(0, data, b1, mask=mask, x0=x0, params=params))
        for _ in range(untimed):
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


FIXTURE_DIR = os.path.join(ROOT, "tests", "golden", "fullsize")
FIXTURE_ITERATIONS = {"tv_128x12x8": 500, "tgv2_256x25x12": 1000, "ictgv2_256x30x16": 1000}
# AdaptLambda constants of InitLambda (tv.cpp / tgv2.cpp, ictgv2.cpp)
LAMBDA_KD = {"tv": (0.4 * 0.2991, 10.0 * 0.2991), "tgv2": (0.4 * 0.2991, 10.0 * 0.2991),
             "ictgv2": (0.4 / (5.1191 * 0.35), 10.0 / (5.1191 * 0.35))}
STEP_RULE_PHASE = 10  # iterations k < 10 adapt sigma / tau every iteration
# The engine times its launch variants once a launch set has run 8 plain
# iterations unchanged, over 3 plain iterations per variant (at most 4): 20
# more untimed iterations keep that out of the timed region.
TUNE_PHASE = 20


def input_digest(data, b1, x0, mask):
    """sha256 prefix of the inputs as the GPU receives them (complex64 data,
    coil maps and u0, float32 mask), widened to 128 / 64 bits: the key the
    FP64 fixtures are stored under."""
    h = hashlib.sha256()
    for a in (data, b1, x0):
        h.update(np.ascontiguousarray(np.asarray(a, np.complex64).astype(np.complex128)).tobytes())
    h.update(np.ascontiguousarray(np.asarray(mask, np.float32).astype(np.float64)).tobytes())
    return h.hexdigest()[:16]


def load_fixture(name):
    path = os.path.join(FIXTURE_DIR, f"{name}_{FIXTURE_ITERATIONS.get(name, 0)}.npz")
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as f:
        return dict(x=f["x"].astype(np.complex128), digest=str(f["digest"]),
                    iterations=int(f["iterations"]), lam=float(f["lam"]))


def nrmse(x, ref):
    """||x - x_ref|| / ||x_ref|| over the complex image (matlab/myerror.m:68-70)."""
    x, ref = np.asarray(x, np.complex128).ravel(), np.asarray(ref, np.complex128).ravel()
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def cartesian_leg(A, synth, torch, dev, name, batch, dual, steps, warmup, plain=True,
                  check_nrmse=True, kernel_times=False):
    """Batched (and plain) PD iterations/s of one Cartesian workload.

    Protocol: every request first runs the k < 10 step-rule phase untimed
    (setup), then `warmup` untimed steps, then exactly `steps` timed steps
    (one step = one PD iteration of every resident reconstruction, the
    k % 50 == 0 step-rule iterations included as they come).  With
    check_nrmse the batched requests then run to the fixture's iteration
    count, the whole reconstruction is timed too, and the first result's
    NRMSE against the FP64 fixture is returned."""
    wl = synth.CONFIGS[name]
    inp = synth.make(wl)
    T, C, H, W = inp["data"].shape
    method = {"tv": 0, "tgv2": 1, "ictgv2": 2}[wl.method]
    dual_kind = {"fp32": 0, "bf16": 1, "q16": 2}[dual]
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.complex64)).to(dev)
    data, b1, x0 = c64(inp["data"]), c64(inp["b1"]), c64(inp["x0"])
    mask = torch.from_numpy(np.ascontiguousarray(inp["mask"], np.float32)).to(dev)
    if "lambda" in wl.params:
        lam = float(wl.params["lambda"])
    else:  # the Cartesian AdaptLambda (cartesian_operator.cpp:47-62)
        m = np.asarray(inp["mask"], np.float32).astype(np.float64)
        k, d = LAMBDA_KD[wl.method]
        lam = float(m.size / float((m ** 2).sum()) * k + d)
    fx = load_fixture(name) if check_nrmse else None
    fixture_ok = (fx is not None and fx["digest"] == input_digest(inp["data"], inp["b1"],
                                                                   inp["x0"], inp["mask"]))
    if fixture_ok:
        lam = fx["lam"]
    run_its = STEP_RULE_PHASE + TUNE_PHASE + warmup + steps
    full_its = fx["iterations"] if fixture_ok else 0
    params = {k: v for k, v in wl.params.items() if k in ("alpha0", "alpha1")}
    params.update({"lambda": lam, "adaptLambda": 0, "dualStorage": dual_kind,
                   "maxIt": max(full_its, run_its + 10)})

    def request():
        return A.Request(method, data, b1, mask=mask, x0=x0, centered=False, params=params)

    def run(engine):
        """(setup s, warmup s, timed s)."""
        out = []
        for n in (STEP_RULE_PHASE + TUNE_PHASE, warmup, steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(n):
                engine.step()
            torch.cuda.synchronize()
            out.append(time.perf_counter() - t0)
        return out

    res = {"workload": name, "method": wl.method, "matrix": [H, W], "frames": T, "coils": C,
           "batch": batch, "dual_storage": dual}
    if plain:
        log(f"{name}: plain path ({dual} duals)")
        eng = A.Engine(use_graphs=False, max_batch=1)
        eng.submit(request())
        t_plain = run(eng)[2]
        res["plain_iters_per_s"] = round(steps / t_plain, 2)
        del eng
    log(f"{name}: batched path (B={batch}, {dual} duals)")
    eng = A.Engine(use_graphs=True, max_batch=batch)
    ids = [eng.submit(request()) for _ in range(batch)]
    assert all(i is not None for i in ids)
    t_setup, t_warm, t_batch = run(eng)
    res["batched_iters_per_s"] = round(batch * steps / t_batch, 2)
    res["ms_per_step"] = round(1e3 * t_batch / steps, 4)
    if plain:
        res["batched_over_plain"] = round(res["batched_iters_per_s"] / res["plain_iters_per_s"], 3)
    if fixture_ok:
        x = torch.empty_like(x0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while not eng.poll(ids[0], x):
            eng.step()
        torch.cuda.synchronize()
        t_rest = time.perf_counter() - t0
        t_full = t_setup + t_warm + t_batch + t_rest
        res["full_reconstruction"] = {
            "iterations": full_its,
            "seconds": round(t_full, 3),
            "iters_per_s": round(batch * full_its / t_full, 2),
            "note": "the batch's whole reconstructions, every step-rule iteration included"}
        res["nrmse"] = float(f"{nrmse(x.cpu().numpy(), fx['x']):.4g}")
        res["nrmse_vs"] = f"FP64 oracle, {full_its} iterations (tests/golden/fullsize)"
    elif check_nrmse:
        res["nrmse"] = None
        res["nrmse_vs"] = "fixture missing or inputs differ"
    del eng
    if kernel_times:
        log(f"{name}: kernel timing")
        eng = A.Engine(use_graphs=False, max_batch=batch)
        for _ in range(batch):
            eng.submit(request())
        res["_kernel_ms"] = A.engine_time_kernels(eng, reps=20)
        res["_voxels"] = W * H * T
        del eng
    return res


def copy_calibration(torch, dev, gib=4):
    """A plain device copy of `gib` GiB into another `gib` GiB (a working set
    many times the 256 MB MALL): the achievable streaming rate, timed with
    HIP events, and the reference point that scales rocprofv3 FETCH_SIZE /
    WRITE_SIZE counters to bytes."""
    n = gib * (1 << 30) // 4
    src = torch.empty(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    src.fill_(1.0)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 5
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return {"bytes_per_copy": 2 * n * 4, "ms": round(ms, 4),
            "gbs": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)}


def cpu_baseline(synth, name, seconds):
    """The FP64 C oracle (OpenMP) on the host cores: steady-state PD iterations
    (fixed sigma, tau: no step-rule reductions) of the full-size workload,
    after one untimed warm-up iteration; all cores for about `seconds`, then
    one iteration on one core."""
    import oracle as O
    wl = synth.CONFIGS[name]
    inp = synth.make(wl)
    T, C, H, W = inp["data"].shape
    p = O.default_params(wl.method)
    m = np.asarray(inp["mask"], np.float64)
    k, d = LAMBDA_KD[wl.method]
    lam = wl.params.get("lambda") or float(m.size / float((m ** 2).sum()) * k + d)
    p.update(lambda_=lam, schedule=2, sigma=0.25, tau=0.25,
             **{kk: v for kk, v in wl.params.items() if kk in ("alpha0", "alpha1")})
    args = (inp["data"], inp["b1"], inp["mask"], inp["x0"])
    cores = min(16, os.cpu_count() or 1)

    def timed(its, threads):
        t0 = time.perf_counter()
        O.recon(dict(p, maxIt=its), *args, nthreads=threads)
        return time.perf_counter() - t0

    t1 = timed(1, cores)  # warm-up (also sizes the sample)
    its = int(min(20, max(3, seconds / max(t1, 1e-3))))
    dt = timed(its, cores)
    dt1 = timed(1, 1)
    return {"value": round(its / dt, 4), "unit": "PD iterations/s", "cores": cores,
            "kind": "port",
            "sample": f"{its} steady-state PD iterations (fixed sigma/tau) of one full-size "
                      f"{name} reconstruction after 1 warm-up iteration, FP64 C oracle, "
                      f"OpenMP {cores} threads",
            "single_core": {"value": round(1.0 / dt1, 5), "cores": 1,
                            "sample": "1 steady-state PD iteration, 1 thread"}}


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
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target length of the all-cores CPU baseline sample")
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="arrival window of each serving run (0: skip the serving leg)")
    ap.add_argument("--serve-rates", default="10,15,20,50",
                    help="Poisson arrival rates (requests/s) of the serving load sweep")
    ap.add_argument("--serve-cap-gib", type=float, default=4.0,
                    help="running requests' bytes admitted at once (engine resident_bytes_cap)")
    ap.add_argument("--serve-queue", type=int, default=16,
                    help="waiting requests the engine queues before refusing (load shedding)")
    ap.add_argument("--small-batch", type=int, default=64,
                    help="reconstructions of the TV 128x128x12x8 workload batched in the "
                         "small-request leg (0: skip)")
    ap.add_argument("--radial-iters", type=int, default=200,
                    help="iterations of the radial TGV2 leg (0: skip)")
    ap.add_argument("--radial-osf", type=float, default=None,
                    help="grid oversampling of the radial leg (default: the workload's 1.5; the "
                         "reference's default configuration uses 2)")
    ap.add_argument("--no-nrmse", action="store_true",
                    help="skip running the batched requests to 1000 iterations for the NRMSE "
                         "(shorter profiler runs)")
    ap.add_argument("--no-legs", action="store_true",
                    help="only the headline workload (no cfg1 / fp32-dual / small / radial / "
                         "serving legs)")
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
    bf16 = args.dual != "fp32"  # 4-byte duals

    head = cartesian_leg(A, synth, torch, dev, wl.name, args.batch, args.dual, args.steps,
                         args.warmup, plain=True, check_nrmse=not args.no_nrmse,
                         kernel_times=True)
    ktimes = head.pop("_kernel_ms")
    N = head.pop("_voxels")

    # ---- same result alone and in the mix (60 iterations covers the k<10
    #      and k%50 step-rule iterations)
    log("alone vs in-mix check")
    inp = synth.make(wl)
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.complex64)).to(dev)
    data, b1, x0 = c64(inp["data"]), c64(inp["b1"]), c64(inp["x0"])
    mask = torch.from_numpy(np.ascontiguousarray(inp["mask"], np.float32)).to(dev)
    method = {"tv": 0, "tgv2": 1, "ictgv2": 2}[wl.method]
    p60 = dict(wl.params, maxIt=60, dualStorage={"fp32": 0, "bf16": 1, "q16": 2}[args.dual])
    alone = A.Engine(use_graphs=False, max_batch=1)
    ra = alone.submit(A.Request(method, data, b1, mask=mask, x0=x0, params=p60))
    xa = torch.empty_like(x0)
    while not alone.poll(ra, xa):
        alone.step()
    mix = A.Engine(use_graphs=True, max_batch=4)
    mix.submit(A.Request(method, data, b1, mask=mask, x0=torch.zeros_like(x0),
                         params=dict(p60, maxIt=90)))
    rm = mix.submit(A.Request(method, data, b1, mask=mask, x0=x0, params=p60))
    xm = torch.empty_like(x0)
    while not mix.poll(rm, xm):
        mix.step()
    torch.cuda.synchronize()
    mix_diff = float((xa - xm).abs().max().item())
    del alone, mix, data, b1, x0, mask

    log("copy calibration")
    calib = copy_calibration(torch, dev)

    per_voxel = {k: algorithmic_bytes_per_voxel(k, wl.method, bf16) for k in ("dual", "primal")}
    dom = max(per_voxel, key=lambda k: ktimes[k])
    bytes_per_launch = per_voxel[dom] * N * args.batch
    achieved = bytes_per_launch / (ktimes[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(wl.name, args.batch, args.dual, dom),
            "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(ktimes[dom], 4),
            "copy_gbs": calib["gbs"], "frac_of_copy": round(achieved / calib["gbs"], 4),
            "copy_calibration": calib}

    legs = {}
    if not args.no_legs:
        # the TGV2 workload as the plain fp32 path it is quoted on, and batched
        legs["tgv2_256x25x12"] = cartesian_leg(A, synth, torch, dev, "tgv2_256x25x12",
                                               args.batch, "fp32", args.steps, args.warmup)
        # the headline workload with 8-byte (fp32) duals, the reference's storage
        legs["fp32_duals"] = cartesian_leg(A, synth, torch, dev, wl.name, args.batch, "fp32",
                                           args.steps, args.warmup, plain=False,
                                           check_nrmse=False)
    small = None
    if args.small_batch > 0 and not args.no_legs:
        log(f"small-request leg (tv_128x12x8, B={args.small_batch})")
        small = small_leg(A, synth, torch, dev, args.small_batch, args.steps, args.warmup)

    radial = None
    if args.radial_iters > 0 and not args.no_legs:
        if args.radial_osf is not None:
            import dataclasses
            synth.RADIAL = dataclasses.replace(synth.RADIAL, osf=float(args.radial_osf))
        log(f"radial TGV2 leg ({synth.RADIAL.name}, osf {synth.RADIAL.osf}, "
            f"{args.radial_iters} iterations)")
        radial = radial_leg(A, synth, torch, dev, args.radial_iters)

    serving = None
    if args.serve_seconds > 0 and not args.no_legs:
        serving = []
        for rate in [float(r) for r in args.serve_rates.split(",") if r]:
            log(f"serving stream ({rate}/s for {args.serve_seconds}s)")
            serving.append(serving_leg(A, synth, torch, dev, rate, args.serve_seconds,
                                       cap_bytes=args.serve_cap_gib * (1 << 30),
                                       max_queue=args.serve_queue))

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (FP64 oracle)")
        cpu = cpu_baseline(synth, wl.name, args.cpu_seconds)

    out = {
        "metric": "PD iters/s ICTGV2 256\u00b2\u00d730\u00d716c batched vs plain @NRMSE\u22641e-4; "
                  "p50/p99 latency",
        "value": head["batched_iters_per_s"],
        "unit": "PD iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" + {"fp32": "", "bf16": "+bf16-duals", "q16": "+q16-duals"}[args.dual],
        "data": "synthetic",
        "config": {"workload": wl.name, "method": wl.method, "matrix": head["matrix"],
                   "frames": head["frames"], "coils": head["coils"], "accel": wl.accel,
                   "batch": args.batch, "dual_storage": args.dual, "parallelism": "single-gpu"},
        "protocol": f"each request runs its k<{STEP_RULE_PHASE} step-rule phase and "
                    f"{TUNE_PHASE} launch-variant timing iterations untimed, then "
                    f"{args.warmup} warmup + {args.steps} timed steps (k%50==0 step-rule "
                    f"iterations included); full_reconstruction times the whole run",
        "plain_iters_per_s": head["plain_iters_per_s"],
        "batched_over_plain": head["batched_over_plain"],
        "nrmse": head.get("nrmse"),
        "nrmse_vs": head.get("nrmse_vs"),
        "full_reconstruction": head.get("full_reconstruction"),
        "alone_vs_mix_max_abs_diff": mix_diff,
        "kernel_ms": {k: round(v, 4) for k, v in ktimes.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "legs": legs,
        "small_requests": small,
        "radial": radial,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
