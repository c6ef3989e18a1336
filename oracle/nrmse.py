"""NRMSE of full-size GPU reconstructions against the FP64 oracle.

TEST INFRASTRUCTURE ONLY (the checker).  Usage:

    python bench.py --dump-x DIR ...            # on the GPU: saves x per dual storage
    python -m oracle.nrmse --workload ictgv2_256x30x16 DIR/*.npy [--ref REF.npy]

Runs the oracle PD loop for the workload's full iteration count on the same
synthetic inputs (or loads a previously saved oracle result with --ref; with
--save-ref it stores the one it computed) and prints, per GPU result,
NRMSE = ||x - x_ref||_2 / ||x_ref||_2 over the complex image (matlab
myerror.m:68-70).
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from avionic_amd import synth  # noqa: E402


def oracle_reference(name, iterations=None, nthreads=0):
    wl = synth.CONFIGS[name]
    inp = synth.make(wl)
    T, C, H, W = inp["data"].shape
    p = O.default_params(wl.method)
    k, d = O.adapt_lambda_defaults(wl.method)
    lam = wl.params.get("lambda") or O.adapt_lambda_cart(inp["mask"], W, H, T, k, d)
    p.update(maxIt=iterations or wl.iterations, lambda_=lam,
             **{kk: v for kk, v in wl.params.items() if kk in ("alpha0", "alpha1")})
    return O.recon(p, inp["data"], inp["b1"], inp["mask"], inp["x0"], nthreads=nthreads)


def nrmse(x, ref):
    return float(np.linalg.norm((x - ref).ravel()) / np.linalg.norm(ref.ravel()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", required=True)
    ap.add_argument("--ref", default=None, help="saved oracle result (.npy)")
    ap.add_argument("--save-ref", default=None)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("gpu_results", nargs="+")
    a = ap.parse_args()
    if a.ref:
        ref = np.load(a.ref)
    else:
        t0 = time.time()
        ref = oracle_reference(a.workload, nthreads=a.threads)
        print(f"oracle: {time.time() - t0:.0f} s", flush=True)
        if a.save_ref:
            np.save(a.save_ref, ref)
    for path in a.gpu_results:
        x = np.load(path).astype(np.complex128)
        print(f"{os.path.basename(path)}: NRMSE {nrmse(x, ref):.3e}")


if __name__ == "__main__":
    main()
