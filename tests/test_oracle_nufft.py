"""FP64 Kaiser-Bessel gridding oracle of the non-Cartesian operator: exact
adjointness, accuracy against the exact non-uniform DFT, the published beta
formula, and the reference's AdaptLambda KAT.  Gridding values themselves are
not pinned by any reference vector (the reference's NUFFT lives in the
un-vendored gpuNUFFT / IRT libraries): parity unpinned beyond these checks."""
import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(77)


def crandn(*s):
    return RNG.standard_normal(s) + 1j * RNG.standard_normal(s)


def test_kb_beta_closed_form():
    # Beatty et al. 2005: beta = pi sqrt((J/osf)^2 (osf - 1/2)^2 - 0.8)
    assert abs(O.kb_beta(3, 1.5) - np.pi * np.sqrt(3.2)) < 1e-12
    assert abs(O.kb_beta(4, 2.0) - np.pi * np.sqrt(4.0 * 2.25 - 0.8)) < 1e-12


@pytest.mark.parametrize("W,H,T,C,nfe,spf", [(16, 16, 2, 2, 16, 5), (20, 12, 1, 3, 24, 4)])
def test_nufft_adjoint(W, H, T, C, nfe, spf):
    traj = O.radial_trajectory(nfe, spf, T)
    dcf = np.abs(traj).sum(1) + 0.1
    x, b1 = crandn(T, H, W), crandn(C, H, W)
    z = crandn(T, C, nfe * spf)
    a = np.vdot(O.nufft_K(x, b1, traj, dcf).ravel(), z.ravel())
    b = np.vdot(x.ravel(), O.nufft_KH(z, b1, traj, dcf).ravel())
    assert abs(a - b) < 1e-10 * abs(a)


@pytest.mark.parametrize("J,osf,tol", [(3, 1.5, 2e-2), (4, 2.0, 2e-3), (6, 2.0, 2e-5)])
def test_nufft_matches_exact_dft(J, osf, tol):
    T, H, W, C = 1, 24, 24, 2
    traj = O.radial_trajectory(48, 6, T)
    yy, xx = np.mgrid[0:H, 0:W]
    x = (((yy - 12) ** 2 + (xx - 11) ** 2) < 49).astype(complex)[None] * (1 + 0.3j)
    b1 = np.ones((C, H, W), complex)
    b1[1] = np.exp(1j * 0.1 * xx)
    exact = O.nudft_K(x, b1, traj)
    approx = O.nufft_K(x, b1, traj, J=J, osf=osf)
    err = np.linalg.norm(approx - exact) / np.linalg.norm(exact)
    assert err < tol, err


def test_nudft_cartesian_grid_reduces_to_fft():
    """On the Cartesian grid k = j/N the exact sum is the centred unitary DFT
    with the -i sign: a closed form the NUFFT definition must reduce to."""
    N = 8
    ky, kx = np.meshgrid((np.arange(N) - N // 2) / N, (np.arange(N) - N // 2) / N, indexing="ij")
    traj = np.stack([kx.ravel(), ky.ravel()])[None]
    x = crandn(1, N, N)
    b1 = np.ones((1, N, N), complex)
    got = O.nudft_K(x, b1, traj).reshape(N, N)
    ref = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x[0]), norm="ortho"))
    np.testing.assert_allclose(got, ref, atol=1e-10)


def test_recon_nc_runs_and_decreases_residual():
    T, H, W, C = 2, 16, 16, 2
    traj = O.radial_trajectory(32, 8, T)
    dcf = np.sqrt((traj ** 2).sum(1)) + 1.0 / 32
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.stack([50.0 * (((yy - 8) ** 2 + (xx - 8 - t) ** 2) < 16) for t in range(T)]).astype(complex)
    b1 = np.ones((C, H, W), complex) / np.sqrt(C)
    data = O.nufft_K(img, b1, traj, dcf)
    p = O.default_params("tgv2")
    p.update(maxIt=100, lambda_=20.0)
    x = O.recon_nc(p, data, b1, traj, dcf)
    r0 = np.linalg.norm(data)
    r1 = np.linalg.norm(O.nufft_K(x, b1, traj, dcf) - data)
    assert r1 < 0.5 * r0


def test_coil_tar_nc_sums_frames():
    """The all-frames time average equals the sum of each frame's own adjoint
    gridding (the adjoint is a sum over points), so a frame or axis mix-up in
    the rearrangement of noncartesian_coil_construction.cpp:38-60 fails;
    u is the root-sum-of-squares with the phase of the coil sum."""
    rng = np.random.default_rng(11)
    H, W, C, T, nfe, spf = 12, 14, 3, 4, 16, 3
    M = nfe * spf
    tr = O.radial_trajectory(nfe, spf, T)
    dcf = rng.random((T, M)) + 0.1
    k = rng.standard_normal((T, C, M)) + 1j * rng.standard_normal((T, C, M))
    u, crec = O.coil_tar_nc(k, tr, dcf, (H, W))
    ones = np.ones((1, H, W))
    for c in range(C):
        ref = sum(O.nufft_KH(k[t:t + 1, c:c + 1], ones, tr[t:t + 1], dcf[t:t + 1], J=3, osf=2.0)[0]
                  for t in range(T))
        np.testing.assert_allclose(crec[c], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    s = crec.sum(0)
    np.testing.assert_allclose(np.abs(u), np.sqrt((np.abs(crec) ** 2).sum(0)), rtol=1e-12)
    np.testing.assert_allclose(u * np.abs(s), np.abs(u) * s, rtol=0, atol=1e-10 * np.abs(u * s).max())
    u2, _ = O.coil_tar_nc(k, tr, dcf, (H, W), apply_phase=False)
    np.testing.assert_allclose(u2, np.abs(u), rtol=1e-12)
