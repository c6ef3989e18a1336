"""HIP path vs the FP64 oracle on seeded random inputs: the operator pair at
the configuration sizes (and odd / KAT sizes), every stencil, the
projections, the standalone FFT, and short PD loops of all three methods."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(20261202)


def crandn(*shape):
    return RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def line_mask(T, H, W, frac=0.3, seed=0):
    r = np.random.default_rng(seed)
    m = (r.random((T, H, 1)) < frac).astype(np.float64)
    m[:, H // 2 - 1:H // 2 + 1] = 1.0
    return m * np.ones((1, 1, W))


SIZES = [(5, 5, 2, 2), (6, 6, 2, 3), (7, 12, 3, 2), (32, 32, 3, 4), (128, 128, 2, 3),
         (160, 160, 2, 2), (192, 192, 1, 2), (224, 224, 1, 2), (256, 256, 2, 2), (384, 384, 1, 1),
         (96, 80, 2, 3),
         # radix-11 / -13 stages and lengths above 512 (the wide kernel variants): the
         # reference's cine geometry 168 x 416 (cartesian_operator.cc:152-153,
         # demo_avionic_cine.sh:73-74), 11 x 13 products, 768 / 1024 long rows and columns
         (168, 416, 1, 3), (176, 143, 2, 2), (88, 208, 2, 2), (768, 96, 1, 1), (96, 1024, 1, 1)]


@pytest.mark.parametrize("W,H,T,C", SIZES)
@pytest.mark.parametrize("centered", [False, True])
def test_operator_pair(gpu, W, H, T, C, centered):
    print(f"[gpu op] W={W} H={H} T={T} C={C} centered={centered}", flush=True)
    x, b1, z = crandn(T, H, W), crandn(C, H, W), crandn(T, C, H, W)
    for mask in (None, line_mask(T, H, W)):
        ref = O.cart_K(x, b1, mask, centered)
        got = gpu.cart_K(x, b1, mask, centered)
        assert rel(got, ref) < 2e-6
        ref = O.cart_KH(z, b1, mask, centered)
        got = gpu.cart_KH(z, b1, mask, centered)
        assert rel(got, ref) < 2e-6


@pytest.mark.parametrize("W,H", [(5, 5), (6, 6), (10, 10), (12, 20), (128, 128), (256, 256), (384, 384), (160, 224),
                                 (168, 416), (143, 11), (1024, 768)])
@pytest.mark.parametrize("sign", [-1, 1])
@pytest.mark.parametrize("centered", [False, True])
def test_fft2(gpu, W, H, sign, centered):
    a = crandn(3, H, W)
    got = gpu.fft2(a, sign, centered)
    ref = np.stack([O.fft2(a[i], sign, centered) for i in range(3)])
    assert rel(got, ref) < 2e-6


@pytest.mark.parametrize("dims", [(3, 5, 6), (4, 16, 16), (1, 4, 7), (2, 2, 2), (5, 33, 40)])
def test_stencils(gpu, dims):
    T, H, W = dims
    d = (0.7, 1.3, 0.45)
    u, p, v, q = crandn(T, H, W), crandn(3, T, H, W), crandn(3, T, H, W), crandn(6, T, H, W)
    assert rel(gpu.grad(u, d), O.grad(u, *d)) < 1e-6
    assert rel(gpu.symgrad(v, d), O.symgrad(v, *d)) < 1e-6
    assert rel(gpu.div(p, d), O.div(p, *d)) < 1e-6
    assert rel(gpu.symdiv(q, d), O.symdiv(q, *d)) < 1e-6
    y = crandn(3, T, H, W) * 2
    assert rel(gpu.prox3(y, 1.3), O.prox3(y, 1.3)) < 1e-6
    y6 = crandn(6, T, H, W) * 2
    assert rel(gpu.prox6(y6, 0.7), O.prox6(y6, 0.7)) < 1e-6


def _problem(T=4, H=32, W=32, C=4, seed=3):
    r = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.zeros((T, H, W), complex)
    for t in range(T):
        img[t] += 100.0 * (((yy - H / 2 - t) ** 2 + (xx - W / 2) ** 2) < (H / 4) ** 2)
        img[t] += 50.0 * (((yy - H / 3) ** 2 + (xx - W / 3 - t) ** 2) < (H / 8) ** 2)
    b1 = r.standard_normal((C, H, W)) * 0.1 + 1.0 + 1j * r.standard_normal((C, H, W)) * 0.1
    b1 /= np.sqrt((np.abs(b1) ** 2).sum(0))
    mask = line_mask(T, H, W, frac=0.35, seed=seed)
    data = O.cart_K(img, b1, mask)
    data += 0.5 * (r.standard_normal(data.shape) + 1j * r.standard_normal(data.shape)) * mask[:, None]
    return data, b1, mask


@pytest.mark.parametrize("method", ["tv", "tgv2", "ictgv2", "tvtemp", "ictv"])
@pytest.mark.parametrize("schedule", [0, 1])
def test_recon_loop_matches_oracle(gpu, method, schedule):
    print(f"[gpu recon] {method} schedule={schedule}", flush=True)
    data, b1, mask = _problem()
    p = O.default_params(method)
    T, C, H, W = data.shape
    k, d = O.adapt_lambda_defaults(method)
    p.update(maxIt=200, schedule=schedule, lambda_=O.adapt_lambda_cart(mask, W, H, T, k, d))
    if method == "tvtemp":
        p.update(dt=0.7)  # used as given, no time/space weights
    if method == "ictv":
        # ICTV's step rule never shrinks the steps (its nKx is the data term
        # alone, ictv.cpp:96-131): start from steps the loop converges with
        p.update(sigma=0.1, tau=0.1)
    ref = O.recon(p, data, b1, mask, logs=True)
    got = gpu.recon_logs(p, data, b1, mask, None)
    assert rel(got["x"], ref["x"]) < 1e-4
    np.testing.assert_allclose(got["sigma"], ref["sigma"], rtol=1e-5)
    np.testing.assert_allclose(got["tau"], ref["tau"], rtol=1e-5)


def test_recon_bf16_duals_runs(gpu):
    data, b1, mask = _problem()
    p = O.default_params("ictgv2")
    T, C, H, W = data.shape
    k, d = O.adapt_lambda_defaults("ictgv2")
    p.update(maxIt=200, lambda_=O.adapt_lambda_cart(mask, W, H, T, k, d))
    ref = O.recon(p, data, b1, mask)
    q = dict(p, dualStorage=1)
    got = gpu.recon_logs(q, data, b1, mask, None)
    err = rel(got["x"], ref)
    print(f"[gpu recon bf16] NRMSE {err:.3e}", flush=True)
    assert err < 5e-2


@pytest.mark.parametrize("method", ["tv", "tgv2", "ictgv2", "tvtemp", "ictv"])
def test_recon_q16_duals_match_oracle(gpu, method):
    """16-bit fixed-point duals scaled by the projection radius (dualStorage
    2): 4 bytes per complex dual like bf16, with an absolute resolution of
    radius / 32767, which keeps the reconstruction within 1e-4 of the FP64
    oracle where bf16 (8-bit mantissa) does not."""
    data, b1, mask = _problem()
    p = O.default_params(method)
    T, C, H, W = data.shape
    k, d = O.adapt_lambda_defaults(method)
    p.update(maxIt=200, lambda_=O.adapt_lambda_cart(mask, W, H, T, k, d))
    if method == "ictv":
        p.update(sigma=0.1, tau=0.1)
    ref = O.recon(p, data, b1, mask)
    got = gpu.recon_logs(dict(p, dualStorage=2), data, b1, mask, None)
    err = rel(got["x"], ref)
    print(f"[gpu recon q16] {method} NRMSE {err:.3e}", flush=True)
    assert err < 1e-4
