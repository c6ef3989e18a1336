"""The oracle's volumetric operator and TGV2_3D loop (cartesian_operator3d.cpp,
tgv2_3d.cpp), pinned without the reference: the 3D transform against numpy's
fftn / ifftn, K and K^H against their closed forms and each other
(adjointness), and the whole TGV2_3D loop against the 2D+t TGV2 loop (itself
pinned by the reference's KATs).  With one coil, b1 = 1 and a full mask both
operators are unitary, so both problems reduce to
lambda/2 ||u - f||^2 + TGV(u) and the iterates agree step by step when the
volume's (dx, dy, dz) equal the time/space weights (ds, ds, dt) of TGV2."""
import numpy as np
import pytest

import oracle as O


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("shape", [(3, 5, 6), (4, 8, 8), (7, 4, 9)])
def test_fft3_matches_numpy(shape):
    a = crandn(np.random.default_rng(0), *shape)
    np.testing.assert_allclose(O.fft3(a, -1), np.fft.fftn(a), rtol=0, atol=1e-11)
    np.testing.assert_allclose(O.fft3(a, +1), np.fft.ifftn(a) * a.size, rtol=0, atol=1e-11)


def test_cart3_closed_form_and_adjoint():
    rng = np.random.default_rng(1)
    D, H, W, C = 5, 6, 7, 3
    x, b1 = crandn(rng, D, H, W), crandn(rng, C, D, H, W)
    mask = (rng.random((D, H, W)) < 0.4).astype(float)
    z = crandn(rng, C, D, H, W)
    n = D * H * W
    k = O.cart3_K(x, b1, mask)
    want = mask[None] * np.fft.ifftn(b1 * x[None], axes=(1, 2, 3)) * np.sqrt(n)
    np.testing.assert_allclose(k, want, rtol=0, atol=1e-12)
    kh = O.cart3_KH(z, b1, mask)
    want = (np.conj(b1) * np.fft.fftn(mask[None] * z, axes=(1, 2, 3))).sum(0) / np.sqrt(n)
    np.testing.assert_allclose(kh, want, rtol=0, atol=1e-12)
    lhs, rhs = np.vdot(k, z), np.vdot(x, kh)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_tgv2_3d_equals_tgv2_on_unitary_operators():
    rng = np.random.default_rng(2)
    D, H, W = 4, 8, 6
    f = crandn(rng, D, H, W)
    tsw = 3.0
    ds, dt = O.time_space_weights(tsw)
    p2 = O.default_params("tgv2")
    p2.update(maxIt=150, lambda_=4.0, tsw=tsw)
    x2 = O.recon(p2, O.cart_K(f, np.ones((1, H, W))), np.ones((1, H, W)))
    p3 = O.default_params("tgv2_3d")
    p3.update(maxIt=150, lambda_=4.0, dx=ds, dy=ds, dz=dt, alpha0=p2["alpha0"])
    ones = np.ones((1, D, H, W))
    r3 = O.recon3(p3, O.cart3_K(f, ones), ones)
    np.testing.assert_allclose(r3["x"], x2, rtol=0, atol=1e-9 * np.abs(x2).max())
