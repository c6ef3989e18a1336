"""The oracle's ICTV (ictv.cpp:227-368) pinned without the reference.  With
equal time/space weights for both TV terms the infimal convolution of
r1 TV(u - w) and r3 TV(w) is min(r1, r3) TV(u) = alpha1 TV(u)
(r1 = alpha1 alpha / m, r3 = alpha1 (1 - alpha) / m, m = min(alpha,
1 - alpha)), so the converged ICTV image equals the converged TV image with
lambda / alpha1.  That fixes the regulariser operator, its adjoint, both
projection radii and the lambda placement; a wrong sign in the w update or a
swapped radius moves the solution.

ICTV's step rule keeps only the data term in nKx (ictv.cpp:96-131: the
regulariser sum is overwritten before use), so it never shrinks the steps; the
loop converges here from sigma = tau = 0.1."""
import numpy as np

import oracle as O


def test_equal_weights_reduce_to_tv():
    rng = np.random.default_rng(4)
    T, H, W, C = 3, 8, 8, 2
    img = rng.standard_normal((T, H, W)) + 1j * rng.standard_normal((T, H, W))
    b1 = np.ones((C, H, W)) / np.sqrt(C) + 0j
    d = O.cart_K(img, b1)
    lam, a1 = 3.0, 2.0
    p = O.default_params("ictv")
    p.update(maxIt=12000, lambda_=lam, tsw=3.0, tsw2=3.0, alpha=0.3, alpha1=a1, sigma=0.1,
             tau=0.1, gapEvery=3000)
    r = O.recon(p, d, b1, logs=True)
    q = O.default_params("tv")
    q.update(maxIt=6000, lambda_=lam / a1, tsw=3.0)
    ref = O.recon(q, d, b1)
    err = np.abs(r["x"] - ref).max() / np.abs(ref).max()
    assert err < 5e-6
    # the PD gap closes
    assert r["gap"][-1] < 1e-6 * r["gap"][0]
    # the step sizes are never reduced (data-only nKx)
    assert np.all(r["sigma"] == 0.1) and np.all(r["tau"] == 0.1)


def test_ictv_norm_matches_its_definition():
    rng = np.random.default_rng(5)
    T, H, W = 4, 6, 7
    u = rng.standard_normal((T, H, W)) + 1j * rng.standard_normal((T, H, W))
    w = rng.standard_normal((T, H, W)) + 1j * rng.standard_normal((T, H, W))
    d6 = np.array([0.8, 0.8, 1.3, 1.1, 1.1, 0.6])
    a1, al = 1.7, 0.35
    m = min(al, 1 - al)
    got = O.ictv_norm(u, w, a1, al, d6)
    want = a1 * (al / m) * O.tv_norm(u - w, *d6[:3]) + a1 * ((1 - al) / m) * O.tv_norm(w, *d6[3:])
    assert abs(got - want) < 1e-12 * abs(want)
