"""The oracle's power iteration for ||K_full|| (orc_opnorm*, the step-size
estimate of matlab/getopnorm.m:1-16 on the stacked regulariser + data
operator), pinned two ways:

* closed form: with a full mask and coil maps of unit sum of squares, K^H K is
  the identity and the gradient's normal operator is the Kronecker sum of the
  1D Neumann Laplacians, so ||[grad; K]||^2 = 1 + sum over axes of
  4 sin^2(pi (n-1) / (2 n)) / d^2;
* brute force: on tiny grids the full matrix of K_full is assembled column by
  column from the oracle's KAT-pinned primitives (grad, symgrad with the
  off-diagonal entries scaled by sqrt(2) for the |.|_6 pairing, the data
  operator) and its largest singular value comes from LAPACK.  A wrong adjoint,
  weight or sign in the iteration's K_full^* K_full moves the estimate away.
"""
import numpy as np
import pytest

import oracle as O

TV, TGV2, ICTGV2, TVTEMP, ICTV = 0, 1, 2, 3, 4


def unit_coils(C, H, W, seed):
    r = np.random.default_rng(seed)
    b1 = r.standard_normal((C, H, W)) + 1j * r.standard_normal((C, H, W))
    return b1 / np.sqrt((np.abs(b1) ** 2).sum(0))


@pytest.mark.parametrize("W,H,T,ds,dt", [(6, 5, 4, 1.0, 1.0), (7, 4, 3, 1.3, 0.7), (5, 5, 2, 0.8, 2.1)])
def test_tv_closed_form(W, H, T, ds, dt):
    b1 = unit_coils(3, H, W, 1)
    lap = lambda n, d: 4.0 * np.sin(np.pi * (n - 1) / (2 * n)) ** 2 / d ** 2
    exact = np.sqrt(1.0 + lap(W, ds) + lap(H, ds) + lap(T, dt))
    got = O.opnorm(TV, b1, None, W, H, T, [ds, ds, dt, 1, 1, 1], 400)
    assert abs(got - exact) <= 1e-9 * exact


def dense_kfull(method, W, H, T, d6, kop):
    """Columns K_full e_j for every complex primal basis vector e_j (the
    imaginary unit basis gives i times the same column: one pass suffices)."""
    N = W * H * T
    P = O.primal_components(method)
    s2 = np.sqrt(2.0)
    cols = []
    for j in range(P * N):
        e = np.zeros(P * N, complex)
        e[j] = 1.0
        x = e.reshape(P, T, H, W)
        parts = []
        if method == TVTEMP:
            parts.append(O.tgrad(x[0], d6[2]))
        elif method == TV:
            parts.append(O.grad(x[0], d6[0], d6[1], d6[2]))
        elif method == ICTV:
            parts.append(O.grad(x[0] - x[1], d6[0], d6[1], d6[2]))
            parts.append(O.grad(x[1], d6[3], d6[4], d6[5]))
        else:
            u = x[0] - (x[4] if method == ICTGV2 else 0)
            parts.append(O.grad(u, d6[0], d6[1], d6[2]) - x[1:4])
            e6 = O.symgrad(x[1:4], d6[0], d6[1], d6[2])
            e6[3:] *= s2
            parts.append(e6)
            if method == ICTGV2:
                parts.append(O.grad(x[4], d6[3], d6[4], d6[5]) - x[5:8])
                e6 = O.symgrad(x[5:8], d6[3], d6[4], d6[5])
                e6[3:] *= s2
                parts.append(e6)
        parts.append(kop(x[0]))
        cols.append(np.concatenate([p.ravel() for p in parts]))
    return np.stack(cols, axis=1)


METHODS = [TV, TGV2, ICTGV2, TVTEMP, ICTV]


def weights(method):
    p = O.default_params(method)
    ds, dt = O.time_space_weights(p["tsw"])
    ds2, dt2 = O.time_space_weights(p["tsw2"])
    if method == TVTEMP:
        return [np.inf, np.inf, 1.7, ds2, ds2, dt2]
    return [ds, ds, dt, ds2, ds2, dt2]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("centered", [False, True])
def test_cartesian_matches_dense_svd(method, centered):
    W, H, T, C = 4, 3, 3, 2
    r = np.random.default_rng(7 + method)
    b1 = r.standard_normal((C, H, W)) + 1j * r.standard_normal((C, H, W))
    mask = (r.random((T, H, 1)) < 0.6) * np.ones((1, 1, W))
    d6 = weights(method)
    A = dense_kfull(method, W, H, T, d6, lambda u: O.cart_K(u, b1, mask, centered))
    exact = np.linalg.svd(A, compute_uv=False)[0]
    got = O.opnorm(method, b1, mask, W, H, T, d6, 3000, centered=centered)
    assert abs(got - exact) <= 1e-6 * exact, (got, exact)


def test_radial_matches_dense_svd():
    W = H = 6
    T, C, nfe, spf = 2, 2, 8, 3
    traj = O.radial_trajectory(nfe, spf, T)
    dcf = np.sqrt(traj[:, 0] ** 2 + traj[:, 1] ** 2)
    r = np.random.default_rng(3)
    b1 = r.standard_normal((C, H, W)) + 1j * r.standard_normal((C, H, W))
    d6 = weights(TGV2)
    A = dense_kfull(TGV2, W, H, T, d6, lambda u: O.nufft_K(u, b1, traj, dcf, 3, 2.0))
    exact = np.linalg.svd(A, compute_uv=False)[0]
    got = O.opnorm_nc(TGV2, b1, traj, dcf, T, d6, 3000, J=3, osf=2.0)
    assert abs(got - exact) <= 1e-6 * exact, (got, exact)


def test_volume_matches_dense_svd():
    W, H, D, C = 3, 4, 3, 2
    r = np.random.default_rng(11)
    b1 = r.standard_normal((C, D, H, W)) + 1j * r.standard_normal((C, D, H, W))
    mask = (r.random((D, H, W)) < 0.7).astype(float)
    d6 = [1.0, 1.3, 0.9, 1, 1, 1]
    A = dense_kfull(TGV2, W, H, D, d6, lambda u: O.cart3_K(u, b1, mask))
    exact = np.linalg.svd(A, compute_uv=False)[0]
    got = O.opnorm3(b1, mask, W, H, D, d6, 3000)
    assert abs(got - exact) <= 1e-6 * exact, (got, exact)
