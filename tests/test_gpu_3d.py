"""Volumetric Cartesian operator and TGV2_3D on the GPU against the fp64
oracle (cartesian_operator3d.cpp, tgv2_3d.cpp)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("D,H,W,C,masked", [(8, 16, 12, 3, True), (5, 24, 20, 2, False),
                                            (16, 32, 32, 4, True)])
def test_operator3d_matches_oracle(gpu, D, H, W, C, masked):
    import avionic_amd as A
    print(f"[gpu 3d op] D={D} H={H} W={W} C={C} masked={masked}", flush=True)
    rng = np.random.default_rng(D + H)
    x, b1, z = crandn(rng, D, H, W), crandn(rng, C, D, H, W), crandn(rng, C, D, H, W)
    mask = (rng.random((D, H, W)) < 0.4).astype(float) if masked else None
    op = A.CartesianOperator3D(W, H, D, C, gpu._r(mask))
    x32 = x.astype(np.complex64).astype(np.complex128)
    b32 = b1.astype(np.complex64).astype(np.complex128)
    z32 = z.astype(np.complex64).astype(np.complex128)
    zt = gpu._t(z)
    got_k = gpu._n(op.BackwardOperation(gpu._t(x), gpu._t(b1)))
    assert rel(got_k, O.cart3_K(x32, b32, mask)) < 2e-6
    got_kh = gpu._n(op.ForwardOperation(zt, gpu._t(b1)))
    assert rel(got_kh, O.cart3_KH(z32, b32, mask)) < 2e-6
    # K^H works on a masked copy: the caller's k-space is unchanged
    assert np.array_equal(gpu._n(zt), z32)
    if masked:
        lam = op.AdaptLambda(0.5, 2.0)
        assert lam == pytest.approx(0.5 * D * H * W / mask.sum() + 2.0, rel=1e-12)


def test_tgv2_3d_recon_matches_oracle(gpu):
    import avionic_amd as A
    rng = np.random.default_rng(7)
    D, H, W, C = 6, 24, 20, 3
    zz, yy, xx = np.mgrid[0:D, 0:H, 0:W]
    img = 50.0 * (((yy - H / 2) ** 2 + (xx - W / 2) ** 2 + 4 * (zz - D / 2) ** 2) < 40) + 0j
    b1 = 1.0 + 0.2 * crandn(rng, C, D, H, W)
    mask = (rng.random((D, H, W)) < 0.5).astype(float)
    data = O.cart3_K(img, b1, mask) + 0.3 * crandn(rng, C, D, H, W) * mask[None]
    p = O.default_params("tgv2_3d")
    p.update(maxIt=150, lambda_=5.0, dx=1.0, dy=1.0, dz=2.0)
    ref = O.recon3(p, data, b1, mask)
    op = A.CartesianOperator3D(W, H, D, C, gpu._r(mask))
    rec = A.TGV2_3D(W, H, C, D, op, None)
    q = rec.GetParams()
    assert q["alpha0"] == pytest.approx(np.sqrt(3.0))
    q.update(maxIt=150, adaptLambda=0, dx=1.0, dy=1.0, dz=2.0)
    q["lambda"] = 5.0
    rec.SetParams(q)
    x = gpu._t(np.zeros((D, H, W)))
    rec.IterativeReconstruction(gpu._t(data), x, gpu._t(b1))
    err = rel(gpu._n(x), ref["x"])
    print(f"[gpu tgv2_3d] NRMSE vs oracle {err:.3e}", flush=True)
    assert err < 1e-4
    np.testing.assert_allclose(rec.sigma_log(), ref["sigma"], rtol=1e-5)


def test_method_and_operator_must_match(gpu):
    import avionic_amd as A
    op3 = A.CartesianOperator3D(8, 8, 4, 2, None)
    with pytest.raises(Exception, match="volumetric"):
        A.TGV2(8, 8, 2, 4, op3, None)
    op2 = A.CartesianOperator(8, 8, 2, 4, None, centered=False)
    with pytest.raises(Exception, match="volumetric"):
        A.TGV2_3D(8, 8, 2, 4, op2, None)
