"""Checks that pin the FP64 oracle where the reference KATs leave room:
textbook identities (naive DFT, Parseval, adjointness of every operator pair),
closed forms, and mistakes planted in a copy of the oracle's results that these
checks must catch."""
import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(1234)


def crandn(*shape):
    return RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)


def naive_dft(a, sign):
    n = len(a)
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ a


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 10, 11, 13, 16, 30, 64, 97, 128, 160, 192, 224, 256, 384])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1_matches_naive_dft(n, sign):
    a = crandn(n)
    np.testing.assert_allclose(O.fft1(a, sign), naive_dft(a, sign), rtol=0, atol=1e-9 * max(1, n))


@pytest.mark.parametrize("shape", [(5, 5), (6, 6), (10, 10), (12, 20), (7, 9)])
@pytest.mark.parametrize("centered", [False, True])
def test_fft2_matches_numpy(shape, centered):
    a = crandn(*shape)
    ref = np.fft.fft2(np.fft.ifftshift(a)) if centered else np.fft.fft2(a)
    if centered:
        ref = np.fft.fftshift(ref)
    np.testing.assert_allclose(O.fft2(a, -1, centered), ref, atol=1e-9)
    inv = np.fft.ifft2(np.fft.ifftshift(a)) if centered else np.fft.ifft2(a)
    if centered:
        inv = np.fft.fftshift(inv)
    np.testing.assert_allclose(O.fft2(a, +1, centered), inv * a.size, atol=1e-9)


def _inner(a, b):
    return np.vdot(a.ravel(), b.ravel())


@pytest.mark.parametrize("dims", [(2, 5, 6), (3, 8, 8), (1, 4, 7)])
@pytest.mark.parametrize("centered", [False, True])
def test_cart_operator_adjoint_and_closed_form(dims, centered):
    T, H, W = dims
    C = 3
    x, z, b1 = crandn(T, H, W), crandn(T, C, H, W), crandn(C, H, W)
    mask = (RNG.random((T, H, W)) < 0.5).astype(np.float64)
    Kx = O.cart_K(x, b1, mask, centered)
    KHz = O.cart_KH(z, b1, mask, centered)
    assert abs(_inner(Kx, z) - _inner(x, KHz)) < 1e-9 * abs(_inner(Kx, z))
    # closed form with numpy's FFT: K = M . ifft2_ortho(b1 . x) (backward_opt.m)
    sh = (lambda a: np.fft.fftshift(a, axes=(-2, -1))) if centered else (lambda a: a)
    ish = (lambda a: np.fft.ifftshift(a, axes=(-2, -1))) if centered else (lambda a: a)
    ref = sh(np.fft.ifft2(ish(b1[None] * x[:, None]), norm="ortho")) * mask[:, None]
    np.testing.assert_allclose(Kx, ref, atol=1e-12)
    # unit-norm coils and full mask: K^H K = identity (Parseval)
    b1n = b1 / np.sqrt((np.abs(b1) ** 2).sum(0))
    np.testing.assert_allclose(O.cart_KH(O.cart_K(x, b1n, None, centered), b1n, None, centered), x,
                               atol=1e-10)


@pytest.mark.parametrize("dims", [(3, 5, 6), (1, 4, 4), (4, 1, 7), (2, 2, 2)])
def test_stencil_adjointness(dims):
    T, H, W = dims
    d = (0.7, 1.3, 0.45)
    u, p = crandn(T, H, W), crandn(3, T, H, W)
    v, q = crandn(3, T, H, W), crandn(6, T, H, W)
    # <grad u, p> = -<u, div p>
    a = _inner(O.grad(u, *d), p)
    b = -_inner(u, O.div(p, *d))
    assert abs(a - b) < 1e-10 * max(1, abs(a))
    # <E v, q>_w = -<v, divsym q>, off-diagonals weighted twice
    w = np.array([1, 1, 1, 2, 2, 2])[:, None, None, None]
    a = _inner(O.symgrad(v, *d), w * q)
    b = -_inner(v, O.symdiv(q, *d))
    assert abs(a - b) < 1e-10 * max(1, abs(a))


def test_stencil_closed_forms():
    T, H, W = 3, 4, 5
    u = crandn(T, H, W)
    g = O.grad(u, 2.0, 3.0, 0.5)
    np.testing.assert_allclose(g[0][:, :, :-1], np.diff(u, axis=2) / 2.0)
    np.testing.assert_allclose(g[1][:, :-1], np.diff(u, axis=1) / 3.0)
    np.testing.assert_allclose(g[2][:-1], np.diff(u, axis=0) / 0.5)
    assert not g[0][:, :, -1].any() and not g[1][:, -1].any() and not g[2][-1].any()
    # constant image: zero gradient; linear ramp in x: constant inside
    assert not O.grad(np.ones((T, H, W), complex)).any()


def test_prox_closed_form():
    y = crandn(3, 2, 3, 4) * 3
    r = 1.7
    n = np.sqrt((np.abs(y) ** 2).sum(0))
    np.testing.assert_allclose(O.prox3(y, r), y / np.maximum(1, n / r))
    q = crandn(6, 2, 3, 4) * 3
    w = np.array([1, 1, 1, 2, 2, 2])[:, None, None, None]
    n6 = np.sqrt((w * np.abs(q) ** 2).sum(0))
    np.testing.assert_allclose(O.prox6(q, r), q / np.maximum(1, n6 / r))


def test_ellipke_and_weights_closed_form():
    from scipy.special import ellipe
    for m in (0.0, 0.1, 0.5, 0.75, 0.9, 0.999):
        assert abs(O.ellipke(m) - ellipe(m)) < 1e-12
    ds, dt = O.time_space_weights(1.0)
    assert abs(ds - 1) < 1e-12 and abs(dt - 1) < 1e-12
    table = {6: (3.962971, 0.660495), 4: (2.730596, 0.682649), 0.5: (0.770982, 1.541964),
             6.5: (4.274053, 0.657547), 2.5: (1.831325, 0.732530)}
    for tsw, (eds, edt) in table.items():
        ds, dt = O.time_space_weights(tsw)
        assert abs(ds - eds) < 1e-6 and abs(dt - edt) < 1e-6
        assert abs(dt / ds * tsw - 1.0) < 1e-12  # ratio w2/w1 = tsw
    with pytest.raises(ValueError):
        O.time_space_weights(-1.0)


def test_adapt_step_branches():
    # sigma*tau <= tmp^2: unchanged
    assert O.adapt_step(1.0, 1.0, 1.0, 0.5, 0.5) == (0.5, 0.5)
    # small violation: scale both by 0.95
    s, t = O.adapt_step(1.0, 0.5, 1.0, 0.51, 0.51)
    assert abs(s - 0.51 * 0.95) < 1e-15 and abs(t - 0.51 * 0.95) < 1e-15
    # large violation: set to tmp * ratio, tmp / ratio
    s, t = O.adapt_step(2.0, 1.0, 2.0, 1.0, 1.0)
    assert abs(s - 1.0) < 1e-15 and abs(t - 0.25) < 1e-15


def _small_problem(T=3, H=8, W=8, C=2, seed=7):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((T, H, W)) + 1j * rng.standard_normal((T, H, W))
    b1 = rng.standard_normal((C, H, W)) + 1j * rng.standard_normal((C, H, W))
    b1 /= np.sqrt((np.abs(b1) ** 2).sum(0))
    mask = (rng.random((T, H, 1)) < 0.6).astype(float) * np.ones((1, 1, W))
    data = O.cart_K(x * 50, b1, mask)
    return data, b1, mask


@pytest.mark.parametrize("method", ["tv", "tgv2", "ictgv2"])
def test_loop_pd_gap_decreases(method):
    data, b1, mask = _small_problem()
    p = O.default_params(method)
    p.update(maxIt=300, lambda_=5.0, gapEvery=50)
    r = O.recon(p, data, b1, mask, logs=True)
    assert r["gap"][-1] < 0.2 * r["gap"][0]
    # the step rule only ever shrinks sigma*tau
    st = r["sigma"] * r["tau"]
    assert np.all(np.diff(st) <= 1e-15)


def test_planted_mistakes_are_caught():
    """Perturbations of the oracle's outputs that a plausible bug would cause
    (wrong FFT sign, transposed stencil, dropped off-diagonal weight) break
    the identities above."""
    T, H, W = 2, 5, 6
    x, b1 = crandn(T, H, W), crandn(2, H, W)
    mask = np.ones((T, H, W))
    z = crandn(T, 2, H, W)
    Kx_wrong_sign = np.conj(O.cart_K(np.conj(x), np.conj(b1), mask))  # -i exponent
    KHz = O.cart_KH(z, b1, mask)
    assert abs(_inner(Kx_wrong_sign, z) - _inner(x, KHz)) > 1e-3
    v, q = crandn(3, T, H, W), crandn(6, T, H, W)
    a = _inner(O.symgrad(v), q)  # unweighted: off-diagonals counted once
    assert abs(a + _inner(v, O.symdiv(q))) > 1e-3
    u, p = crandn(T, H, W), crandn(3, T, H, W)
    g_t = O.grad(u.transpose(0, 2, 1).copy()).transpose(0, 1, 3, 2)  # x and y swapped
    assert abs(_inner(g_t, p) + _inner(u, O.div(p))) > 1e-3
