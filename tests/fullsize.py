"""Inputs and parameters of the benchmark workloads at their full sizes, shared
by the fixture script (tests/golden/make_fullsize_fixtures.py) and the GPU
tests that compare the HIP path against those fixtures.

Every array is rounded to the precision the GPU receives (complex64 data,
coil maps and initial image, float32 masks, trajectories and densities) before
the FP64 oracle sees it, so a fixture measures the solver's arithmetic and
nothing of the input rounding.

Workloads (sizes, sampling, phantom, noise: avionic_amd/synth.py):
  tv_128x12x8            TV, lambda = 0.01, 500 iterations
  tgv2_256x25x12         TGV2, alpha0 / alpha1 = 2 / 1, lambda adapted from the
                         mask (cartesian_operator.cpp:47-62), 1000 iterations
  ictgv2_256x30x16       ICTGV2, CUDA InitParams defaults (ictgv2.cpp:37-64),
                         lambda adapted from the mask, 1000 iterations
  tgv2_radial_256x30x12  TGV2 on golden-angle radial data, alpha0 / alpha1 =
                         2 / 1, lambda adapted (noncartesian_operator.cpp:88-95),
                         Kaiser-Bessel width 3 on a 1.5x grid, density |k|;
                         k-space = K rho (the oracle's fp64 gridding) + 1%
                         complex Gaussian noise, u0 = K^H d
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

from avionic_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE_DIR = os.path.join(HERE, "golden", "fullsize")

CARTESIAN = ("tv_128x12x8", "tgv2_256x25x12", "ictgv2_256x30x16")
RADIAL = "tgv2_radial_256x30x12"
# Iterations of each fixture (the radial one is shorter: the fp64 gridding
# oracle runs about 7 s per iteration on 8 host cores).
ITERATIONS = {"tv_128x12x8": 500, "tgv2_256x25x12": 1000, "ictgv2_256x30x16": 1000,
              RADIAL: 500}
TV_LAMBDA_K, TV_LAMBDA_D = 0.4 * 0.2991, 10.0 * 0.2991  # tv.cpp / tgv2.cpp InitLambda
ICTGV2_LAMBDA_K, ICTGV2_LAMBDA_D = 0.4 / (5.1191 * 0.35), 10.0 / (5.1191 * 0.35)


def _c64(a):
    return np.ascontiguousarray(a, np.complex64).astype(np.complex128)


def _f32(a):
    return np.ascontiguousarray(a, np.float32).astype(np.float64)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def cartesian_lambda(mask, k, d):
    """AdaptLambda of the Cartesian operator: k W H T / ||mask||^2 + d."""
    return float(mask.size / float((mask ** 2).sum()) * k + d)


def cartesian_case(name: str) -> dict:
    """Rounded inputs plus the oracle's and the product's parameter sets."""
    wl = synth.CONFIGS[name]
    inp = synth.make(wl)
    data, b1, x0 = _c64(inp["data"]), _c64(inp["b1"]), _c64(inp["x0"])
    mask = _f32(inp["mask"])
    its = ITERATIONS[name]
    method = {"tv": 0, "tgv2": 1, "ictgv2": 2}[wl.method]
    if "lambda" in wl.params:
        lam = float(wl.params["lambda"])
    elif wl.method == "ictgv2":
        lam = cartesian_lambda(mask, ICTGV2_LAMBDA_K, ICTGV2_LAMBDA_D)
    else:
        lam = cartesian_lambda(mask, TV_LAMBDA_K, TV_LAMBDA_D)
    # product parameters: InitParams defaults, lambda as above
    prod = {"maxIt": its, "lambda": lam, "adaptLambda": 0}
    prod.update({k: v for k, v in wl.params.items() if k in ("alpha0", "alpha1")})
    return dict(name=name, method=method, method_name=wl.method, data=data, b1=b1, x0=x0,
                mask=mask, iterations=its, lam=lam, product_params=prod,
                extra={k: v for k, v in wl.params.items() if k in ("alpha0", "alpha1")},
                digest=digest(data, b1, x0, mask))


def oracle_params(O, case: dict, iterations=None) -> dict:
    p = O.default_params(case["method_name"])
    p.update(maxIt=int(iterations or case["iterations"]), lambda_=case["lam"], **case["extra"])
    return p


def radial_case(O) -> dict:
    """The radial workload with k-space formed by the oracle's gridding."""
    wl = synth.RADIAL
    h = synth.make_radial_host(wl)
    rho, b1 = _c64(h["rho"]), _c64(h["b1"])
    traj, dcf = _f32(h["traj"]), _f32(h["dcf"])
    k = O.nufft_K(rho, b1, traj, dcf, wl.kernel_width, wl.osf)
    rng = np.random.default_rng(wl.seed + 3000)
    sig = 0.01 * np.abs(k).max()
    noise = rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)
    data = _c64(k + sig * noise)
    x0 = _c64(O.nufft_KH(data, b1, traj, dcf, wl.kernel_width, wl.osf))
    lam = float(wl.n / (wl.spokes * np.pi / 2.0) * TV_LAMBDA_K + TV_LAMBDA_D)
    its = ITERATIONS[RADIAL]
    return dict(name=RADIAL, method=1, method_name="tgv2", data=data, b1=b1, x0=x0, traj=traj,
                dcf=dcf, J=wl.kernel_width, osf=wl.osf, n_fe=wl.n_fe, spokes=wl.spokes,
                iterations=its, lam=lam,
                product_params={"maxIt": its, "lambda": lam, "adaptLambda": 0, "alpha0": 2.0,
                                "alpha1": 1.0},
                extra={"alpha0": 2.0, "alpha1": 1.0},
                digest=digest(data, b1, x0, traj, dcf))


def fixture_path(name: str) -> str:
    return os.path.join(FIXTURE_DIR, f"{name}_{ITERATIONS[name]}.npz")


def load_fixture(name: str) -> dict:
    """x_ref (complex128 from the stored complex64) and its metadata."""
    with np.load(fixture_path(name), allow_pickle=False) as f:
        return dict(x=f["x"].astype(np.complex128), digest=str(f["digest"]),
                    iterations=int(f["iterations"]), lam=float(f["lam"]))


def nrmse(x, ref) -> float:
    """||x - x_ref||_2 / ||x_ref||_2 over the complex image (matlab/myerror.m:68-70)."""
    x = np.asarray(x, np.complex128).ravel()
    ref = np.asarray(ref, np.complex128).ravel()
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))
