#!/usr/bin/env python3
"""Write the full-size FP64-oracle fixtures the GPU NRMSE tests compare with.

For each benchmark workload (tests/fullsize.py) this runs the FP64 C oracle
(oracle/avionic_oracle.c, oracle_nufft.c) for the workload's full iteration
count on its rounded inputs and stores the reconstructed image as complex64
in tests/golden/fullsize/<workload>_<iterations>.npz, together with a digest
of the inputs, lambda and the iteration count.  The inputs are regenerated
from their seeds by tests/fullsize.py; only the result is stored.

Usage: python tests/golden/make_fullsize_fixtures.py [--threads N] [names...]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import fullsize as FS  # noqa: E402
import oracle as O  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("names", nargs="*", default=list(FS.CARTESIAN) + [FS.RADIAL])
    a = ap.parse_args()
    os.makedirs(FS.FIXTURE_DIR, exist_ok=True)
    for name in a.names:
        t0 = time.time()
        if name == FS.RADIAL:
            case = FS.radial_case(O)
            p = FS.oracle_params(O, case)
            x = O.recon_nc(p, case["data"], case["b1"], case["traj"], case["dcf"], case["x0"],
                           J=case["J"], osf=case["osf"], nthreads=a.threads)
        else:
            case = FS.cartesian_case(name)
            p = FS.oracle_params(O, case)
            x = O.recon(p, case["data"], case["b1"], case["mask"], case["x0"],
                        nthreads=a.threads)
        dt = time.time() - t0
        np.savez_compressed(FS.fixture_path(name), x=x.astype(np.complex64),
                            digest=np.array(case["digest"]),
                            iterations=np.array(case["iterations"]),
                            lam=np.array(case["lam"]), oracle_seconds=np.array(dt),
                            threads=np.array(a.threads))
        print(f"{name}: {case['iterations']} iterations in {dt:.0f} s "
              f"(lambda {case['lam']:.6g}, inputs {case['digest']})", flush=True)


if __name__ == "__main__":
    main()
