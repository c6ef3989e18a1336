"""FP64 operator-norm power iteration on the GPU (avionic_recon_operator_norm)
and the fixed-step mode adaptSchedule = 2 (sigma = tau = 0.99 / ||K_full||)
against the FP64 oracle (orc_opnorm*, itself pinned by a closed form and a
dense SVD in test_oracle_opnorm.py): the same deterministic start vector and
step count give the same estimate to ~1e-12; a full mask with unit-norm coils
gives the closed form; a reconstruction in the fixed-step mode matches the
oracle's, eager and through the engine."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

API = {"tv": 0, "tgv2": 1, "ictgv2": 2, "tvtemp": 3, "ictv": 4}


def weights(method, dt_temporal=1.0):
    p = O.default_params(method)
    ds, dt = O.time_space_weights(p["tsw"])
    ds2, dt2 = O.time_space_weights(p["tsw2"])
    if method == "tvtemp":
        return [np.inf, np.inf, dt_temporal, ds2, ds2, dt2]
    return [ds, ds, dt, ds2, ds2, dt2]


def solver(gpu, method, W, H, C, T, mask, centered=False, params=None):
    A = gpu.A
    op = A.CartesianOperator(W, H, C, T, None if mask is None else gpu._r(mask), centered)
    cls = {"tv": A.TV, "tgv2": A.TGV2, "ictgv2": A.ICTGV2, "tvtemp": A.TVTEMP, "ictv": A.ICTV}
    s = cls[method](W, H, C, T, op)
    p = s.GetParams()
    p["lambda"] = 1.0
    p["adaptLambda"] = 0
    if params:
        p.update(params)
    s.SetParams(p)
    return s, op


@pytest.mark.parametrize("method", list(API))
@pytest.mark.parametrize("centered", [False, True])
def test_norm_matches_oracle(gpu, method, centered):
    print(f"[gpu opnorm] {method} centered={centered}", flush=True)
    W, H, T, C = 40, 32, 5, 3
    r = np.random.default_rng(5)
    b1 = r.standard_normal((C, H, W)) + 1j * r.standard_normal((C, H, W))
    b1 = b1.astype(np.complex64).astype(np.complex128)
    mask = (r.random((T, H, 1)) < 0.4) * np.ones((1, 1, W))
    s, _ = solver(gpu, method, W, H, C, T, mask, centered, {"dt": 1.0})
    got = s.OperatorNorm(gpu._t(b1), 30)
    ref = O.opnorm(API[method], b1, mask, W, H, T, weights(method), 30, centered=centered)
    assert abs(got - ref) <= 1e-11 * ref, (got, ref)


def test_norm_closed_form(gpu):
    W, H, T, C = 7, 6, 4, 3
    r = np.random.default_rng(2)
    b1 = r.standard_normal((C, H, W)) + 1j * r.standard_normal((C, H, W))
    b1 /= np.sqrt((np.abs(b1) ** 2).sum(0))
    b1 = b1.astype(np.complex64).astype(np.complex128)
    # time/space weight 1 (ds = dt = 1): the top two eigenvalues of M differ
    # by 5%, so 500 steps converge far below the tolerance
    s, _ = solver(gpu, "tv", W, H, C, T, None, params={"timeSpaceWeight": 1.0})
    ds, dt = O.time_space_weights(1.0)
    lap = lambda n, d: 4.0 * np.sin(np.pi * (n - 1) / (2 * n)) ** 2 / d ** 2
    # the float coil maps are unit-norm only to float precision (K^H K is
    # diag(sum |b1|^2) = 1 + O(1e-7))
    exact = np.sqrt(1.0 + lap(W, ds) + lap(H, ds) + lap(T, dt))
    got = s.OperatorNorm(gpu._t(b1), 500)
    assert abs(got - exact) <= 1e-6 * exact, (got, exact)


def test_norm_volume_matches_oracle(gpu):
    A = gpu.A
    W, H, D, C = 24, 20, 10, 3
    r = np.random.default_rng(8)
    b1 = (r.standard_normal((C, D, H, W)) + 1j * r.standard_normal((C, D, H, W)))
    b1 = b1.astype(np.complex64).astype(np.complex128)
    mask = (r.random((D, H, W)) < 0.5).astype(np.float64)
    op = A.CartesianOperator3D(W, H, D, C, gpu._r(mask))
    s = A.TGV2_3D(W, H, C, D, op)
    p = s.GetParams()
    p.update(dx=1.0, dy=1.3, dz=0.9, adaptLambda=0)
    p["lambda"] = 1.0
    s.SetParams(p)
    got = s.OperatorNorm(gpu._t(b1), 25)
    ref = O.opnorm3(b1, mask, W, H, D, [1.0, 1.3, 0.9, 1, 1, 1], 25)
    assert abs(got - ref) <= 1e-11 * ref, (got, ref)


@pytest.mark.parametrize("W,osf", [(48, 1.5), (40, 2.0)])
def test_norm_radial_matches_oracle(gpu, W, osf):
    A = gpu.A
    H, C, T, nfe, spf = W, 2, 3, 2 * W, 11
    tr = O.radial_trajectory(nfe, spf, T).astype(np.float32).astype(np.float64)
    dcf = np.sqrt(tr[:, 0] ** 2 + tr[:, 1] ** 2).astype(np.float32).astype(np.float64)
    r = np.random.default_rng(4)
    b1 = (r.standard_normal((C, H, W)) + 1j * r.standard_normal((C, H, W)))
    b1 = b1.astype(np.complex64).astype(np.complex128)
    op = A.NoncartesianOperator(W, H, C, T, nfe, spf, gpu._r(tr), gpu._r(dcf), kernel_width=3,
                                osf=osf)
    s = A.TGV2(W, H, C, T, op)
    p = s.GetParams()
    p["lambda"] = 1.0
    p["adaptLambda"] = 0
    s.SetParams(p)
    got = s.OperatorNorm(gpu._t(b1), 20)
    ref = O.opnorm_nc(1, b1, tr, dcf, T, weights("tgv2"), 20, J=3, osf=osf)
    assert abs(got - ref) <= 1e-10 * ref, (got, ref)


def test_norm_full_size_tv_workload(gpu):
    import fullsize as FS
    case = FS.cartesian_case("tv_128x12x8")
    T, C, H, W = case["data"].shape
    s, _ = solver(gpu, "tv", W, H, C, T, case["mask"])
    got = s.OperatorNorm(gpu._t(case["b1"]), 8)
    ref = O.opnorm(0, case["b1"], case["mask"], W, H, T, weights("tv"), 8)
    assert abs(got - ref) <= 1e-11 * ref, (got, ref)


@pytest.mark.parametrize("method", ["tv", "tgv2", "ictgv2"])
def test_fixed_step_reconstruction_matches_oracle(gpu, method):
    print(f"[gpu opnorm] fixed-step {method}", flush=True)
    A = gpu.A
    W, H, T, C = 32, 32, 4, 3
    r = np.random.default_rng(12)
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.stack([60.0 * (((yy - 15 - t) ** 2 + (xx - 16) ** 2) < 70) for t in range(T)]) + 0j
    b1 = r.standard_normal((C, H, W)) * 0.2 + 1.0 + 0.1j
    b1 /= np.sqrt((np.abs(b1) ** 2).sum(0))
    b1 = b1.astype(np.complex64).astype(np.complex128)
    mask = (r.random((T, H, 1)) < 0.5) * np.ones((1, 1, W))
    data = O.cart_K(img, b1, mask).astype(np.complex64).astype(np.complex128)
    its, lam, npi = 200, 5.0, 40
    nk = O.opnorm(API[method], b1, mask, W, H, T, weights(method), npi)
    p = O.default_params(method)
    p.update(maxIt=its, lambda_=lam, schedule=2, sigma=0.99 / nk, tau=0.99 / nk)
    ref = O.recon(p, data, b1, mask, None, logs=True)
    s, _ = solver(gpu, method, W, H, C, T, mask,
                  params={"maxIt": its, "adaptSchedule": 2, "opnormIterations": npi, "lambda": lam})
    x = gpu._t(np.zeros((T, H, W)))
    s.IterativeReconstruction(gpu._t(data), x, gpu._t(b1))
    got = gpu._n(x)
    assert np.allclose(s.sigma_log(), 0.99 / nk, rtol=1e-10, atol=0)
    assert np.allclose(s.tau_log(), 0.99 / nk, rtol=1e-10, atol=0)
    assert np.linalg.norm(got - ref["x"]) / np.linalg.norm(ref["x"]) < 1e-4
    # the engine in the same mode: the same fixed steps and result
    eng = A.Engine(use_graphs=True)
    q = A.default_params(API[method])
    q.update(maxIt=its, adaptSchedule=2, opnormIterations=npi, adaptLambda=0)
    q["lambda"] = lam
    rid = eng.submit(A.Request(API[method], gpu._t(data), gpu._t(b1), mask=gpu._r(mask), params=q))
    xe = gpu._t(np.zeros((T, H, W)))
    steps = 0
    while not eng.poll(rid, xe):
        eng.step()
        steps += 1
        if steps == its:
            sg, tu = eng.step_log(rid)
            assert np.allclose(sg, 0.99 / nk, rtol=1e-10, atol=0)
    ge = gpu._n(xe)
    assert np.linalg.norm(ge - ref["x"]) / np.linalg.norm(ref["x"]) < 1e-4
