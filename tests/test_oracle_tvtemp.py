"""The oracle's temporal TV (TVTEMP, tv_temp.cpp:94-188) pinned without the
reference: its temporal difference against the t component of the full
gradient (a separate code path), exact adjointness of the difference and its
transpose, and the converged loop against the closed-form minimiser of the
two-frame problem

    min_u  lambda/2 ||K u - d||^2 + sum_px |u_1 - u_0| / dt

with one coil, b1 = 1 and a full mask, where K is a unitary FFT per frame, so
the problem splits into independent pixels: with delta = f_1 - f_0 and
s = 1 / (lambda dt), u = f + (s, -s) delta/|delta| when |delta| > 2 s, and both
frames equal (f_0 + f_1) / 2 otherwise.  A spatial term, a wrong dt scaling,
prox radius or lambda placement all move this solution."""
import numpy as np
import pytest

import oracle as O


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_tgrad_is_t_component_of_gradient_and_adjoint():
    rng = np.random.default_rng(1)
    x, y = crandn(rng, 5, 4, 6), crandn(rng, 5, 4, 6)
    for dt in (1.0, 0.37):
        g = O.tgrad(x, dt)
        np.testing.assert_allclose(g, O.grad(x, 1.0, 1.0, dt)[2], rtol=0, atol=1e-14)
        lhs = np.vdot(g, y)
        rhs = np.vdot(x, O.tgrad_T(y, dt))
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    assert O.tvtemp_norm(x, 0.5) == pytest.approx(np.abs(O.tgrad(x, 0.5)).sum(), rel=1e-14)


def test_two_frame_denoising_closed_form():
    rng = np.random.default_rng(2)
    H, W, lam, dt = 4, 5, 2.0, 1.5
    f = 0.5 * crandn(rng, 2, H, W)
    b1 = np.ones((1, H, W), complex)
    d = O.cart_K(f, b1)
    p = O.default_params("tvtemp")
    p.update(maxIt=4000, lambda_=lam, dt=dt, gapEvery=500)
    r = O.recon(p, d, b1, logs=True)
    s = 1.0 / (lam * dt)
    delta = f[1] - f[0]
    big = np.abs(delta) > 2 * s
    assert big.any() and (~big).any(), "fixture covers both regimes"
    unit = delta / np.abs(delta)
    u0 = np.where(big, f[0] + s * unit, 0.5 * (f[0] + f[1]))
    u1 = np.where(big, f[1] - s * unit, 0.5 * (f[0] + f[1]))
    np.testing.assert_allclose(r["x"][0], u0, rtol=0, atol=1e-6)
    np.testing.assert_allclose(r["x"][1], u1, rtol=0, atol=1e-6)
    # the primal-dual gap of the loop closes at the minimiser
    assert r["gap"][-1] < 1e-6 * r["gap"][0]
