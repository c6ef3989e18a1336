"""The in-loop primal-dual gap and the debug exports on the GPU
(ictgv2.cpp:462-516, tv.cpp:176-192): the eager solver's PDGap,
ICTGVnorm, DATAfidelity and x3_component exports against the FP64 oracle
(oracle gap log: orc_recon with gapEvery; norms and fidelity from the
oracle's functionals on its own iterate), the engine's gap log against the
eager one, 4-byte duals read on the device with each component's radius (the
ICTV y3 group uses r3), and the stopPDGap early exit in both paths."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

T, H, W, C = 4, 32, 32, 3
GAP_EVERY = 20


def problem(seed=5):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.stack([80.0 * (((yy - 14 - t) ** 2 + (xx - 17) ** 2) < 60)
                    + 20.0 * (((yy - 8) ** 2 + (xx - 9 + t) ** 2) < 12) for t in range(T)])
    img = img.astype(complex) * np.exp(0.2j)
    b1 = rng.standard_normal((C, H, W)) * 0.1 + 1.0 + 0.1j * rng.standard_normal((C, H, W))
    b1 /= np.sqrt((np.abs(b1) ** 2).sum(0))
    mask = (rng.random((T, H, 1)) < 0.45) * np.ones((1, 1, W))
    data = O.cart_K(img, b1, mask)
    data += mask[:, None] * 0.5 * (rng.standard_normal(data.shape)
                                   + 1j * rng.standard_normal(data.shape))
    c = lambda a: a.astype(np.complex64).astype(np.complex128)
    return c(data), c(b1), mask.astype(np.float32).astype(np.float64)


METHODS = {"tv": 0, "tgv2": 1, "ictgv2": 2, "ictv": 4}


def oracle_params(method, maxit, lam):
    p = O.default_params(method)
    p.update(maxIt=maxit, lambda_=lam, gapEvery=GAP_EVERY)
    if method == "ictv":  # ICTV stalls from sigma = tau = 1/3 (its step rule never shrinks)
        p.update(sigma=0.1, tau=0.1)
    return p


def gpu_params(A, method, maxit, lam, ds):
    p = A.default_params(METHODS[method])
    p.update(maxIt=maxit, adaptLambda=0, dualStorage=ds)
    p["lambda"] = lam
    if method == "ictv":
        p.update(sigma=0.1, tau=0.1)
    return p


def d6_of(method):
    p = O.default_params(method)
    ds, dt = O.time_space_weights(p["tsw"])
    ds2, dt2 = O.time_space_weights(p["tsw2"])
    return [ds, ds, dt, ds2, ds2, dt2]


def eager(gpu, method, data, b1, mask, maxit, lam, ds, stop=0.0):
    A = gpu.A
    op = A.CartesianOperator(W, H, C, T, gpu._r(mask), False)
    s = {"tv": A.TV, "tgv2": A.TGV2, "ictgv2": A.ICTGV2, "ictv": A.ICTV}[method](W, H, C, T, op)
    p = gpu_params(A, method, maxit, lam, ds)
    p["stopPDGap"] = stop
    s.SetParams(p)
    s.SetDebug(True, GAP_EVERY)
    x = gpu._t(np.zeros((T, H, W)))
    s.IterativeReconstruction(gpu._t(data), x, gpu._t(b1))
    out = {}
    s.ExportAdditionalResults("", lambda _d, name, arr: out.__setitem__(name, arr.copy()))
    out["x"] = gpu._n(x)
    return out


LAM = 4.0


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("ds", [0, 2])
def test_eager_exports_match_oracle(gpu, method, ds):
    print(f"[gpu gap] eager exports {method} dualStorage={ds}", flush=True)
    data, b1, mask = problem()
    maxit = 4 * GAP_EVERY + 1  # gaps at k = 0, 20, 40, 60, 80; final state = iterate 80
    ref = O.recon(oracle_params(method, maxit, LAM), data, b1, mask, None, logs=True)
    got = eager(gpu, method, data, b1, mask, maxit, LAM, ds)
    gap = got["PDGap"].real.astype(np.float64)
    assert gap.shape == ref["gap"].shape
    tol = 2e-4 if ds == 0 else 2e-3
    np.testing.assert_allclose(gap, ref["gap"], rtol=tol)
    u = ref["primal"][0]
    assert np.linalg.norm(got["x"] - u) / np.linalg.norm(u) < (1e-4 if ds == 0 else 1e-3)
    if method in ("ictgv2", "ictv"):
        d6 = d6_of(method)
        p = O.default_params(method)
        if method == "ictgv2":
            pr = ref["primal"]
            norm = O.ictgv2_norm(pr[0], pr[1:4], pr[4], pr[5:8], p["alpha0"], p["alpha1"],
                                 p["alpha"], d6)
            w = pr[4]
        else:
            pr = ref["primal"]
            norm = O.ictv_norm(pr[0], pr[1], p["alpha1"], p["alpha"], d6)
            w = pr[1]
        fid = LAM / 2.0 * np.linalg.norm(O.cart_K(u, b1, mask) - data)
        # ICTGV2 logs the initial norm and fidelity before the loop (ictgv2.cpp:318-326)
        gnorm = got["ICTGVnorm" if method == "ictgv2" else "ictvnorm"].real[-1]
        gfid = got["DATAfidelity"].real[-1]
        assert abs(gnorm - norm) <= tol * abs(norm)
        assert abs(gfid - fid) <= tol * abs(fid)
        x3 = got["x3_component"].astype(np.complex128).reshape(T, H, W)
        assert np.linalg.norm(x3 - w) / np.linalg.norm(w) < (1e-4 if ds == 0 else 1e-3)


def test_ictv_q16_gap_equals_f32_gap(gpu):
    """The ICTV y3 group is stored relative to r3 = alpha1 (1 - alpha) / m; a
    decode with r2 (the TGV2 grouping) would scale it by r2 / r3."""
    data, b1, mask = problem(seed=9)
    maxit = 3 * GAP_EVERY + 1
    f32 = eager(gpu, "ictv", data, b1, mask, maxit, LAM, 0)["PDGap"].real
    q16 = eager(gpu, "ictv", data, b1, mask, maxit, LAM, 2)["PDGap"].real
    np.testing.assert_allclose(q16, f32, rtol=2e-3)


@pytest.mark.parametrize("method", list(METHODS))
def test_engine_gap_log_and_stop(gpu, method):
    print(f"[gpu gap] engine {method}", flush=True)
    A = gpu.A
    data, b1, mask = problem(seed=3)
    maxit = 6 * GAP_EVERY + 1
    ref = eager(gpu, method, data, b1, mask, maxit, LAM, 0)
    gap = ref["PDGap"].real.astype(np.float64)
    eng = A.Engine(use_graphs=True)
    rid = eng.submit(A.Request(METHODS[method], gpu._t(data), gpu._t(b1), mask=gpu._r(mask),
                               params=gpu_params(A, method, maxit, LAM, 0), debug=True,
                               debugstep=GAP_EVERY))
    x = gpu._t(np.zeros((T, H, W)))
    while True:
        eng.step()
        g, r, f = eng.gap_log(rid)
        if eng.poll(rid, x):
            break
    assert g.shape == gap.shape
    np.testing.assert_allclose(g, gap, rtol=1e-4)
    # stop below a threshold between gap j (the first gap from k = 40 on that
    # is below every earlier one) and the smallest earlier gap: both paths run
    # iterations 0 .. 20 j and return that iterate (the ICTGV2 gap is not
    # monotone)
    j = next(i for i in range(2, len(gap)) if gap[i] < gap[:i].min())
    thr = float(np.sqrt(gap[j] * gap[:j].min()))
    stop_e = eager(gpu, method, data, b1, mask, maxit, LAM, 0, stop=thr)
    assert stop_e["PDGap"].shape[0] == j + 1
    eng2 = A.Engine(use_graphs=True)
    p = gpu_params(A, method, maxit, LAM, 0)
    p["stopPDGap"] = thr
    rid = eng2.submit(A.Request(METHODS[method], gpu._t(data), gpu._t(b1), mask=gpu._r(mask),
                                params=p))
    xs = gpu._t(np.zeros((T, H, W)))
    steps = 0
    while not eng2.poll(rid, xs):
        eng2.step()
        steps += 1
    assert steps == j * GAP_EVERY + 1, (steps, j)
    o = O.recon(oracle_params(method, j * GAP_EVERY + 1, LAM), data, b1, mask, None)
    got = gpu._n(xs)
    assert np.linalg.norm(got - stop_e["x"]) / np.linalg.norm(stop_e["x"]) < 1e-5
    assert np.linalg.norm(got - o) / np.linalg.norm(o) < 1e-4
