"""Coil-sensitivity estimation on the GPU (coil.hip) against the FP64 oracle
(oracle/oracle_coil.c, pinned by the coil_construction.cc KATs in
test_kats.py): every step (time-averaged reconstruction, H1 magnitudes, coil
order, TGV2 image estimate, B1Recon) and the whole PerformCoilConstruction on
a masked dynamic phantom with smooth coil maps, uncentred and centred."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def problem(W=48, H=40, C=4, T=6, seed=3):
    from avionic_amd import synth
    rho = synth.phantom(max(W, H), T, 5, seed)[:, :H, :W]
    b1 = synth.coil_maps(max(W, H), C, seed)[:, :H, :W]
    mask = np.fft.ifftshift(synth.vd_line_mask(H, T, 3, 8, seed), axes=(1, 2))[:, :, :1] * np.ones((1, 1, W))
    k = np.stack([np.fft.ifft2(b1 * rho[t][None], norm="ortho") * mask[t][None] for t in range(T)])
    c = lambda a: np.asarray(a, np.complex64).astype(np.complex128)
    return c(k), mask.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("centered", [False, True])
def test_steps_match_oracle(gpu, centered):
    print(f"[gpu coil] steps centered={centered}", flush=True)
    k, mask = problem()
    T, C, H, W = k.shape
    u, crec = gpu.coil_tar(k, mask, centered)
    ur, crecr = O.coil_tar(k, mask, centered)
    assert rel(crec, crecr) < 2e-6 and rel(u, ur) < 2e-6
    au = np.abs(ur).astype(np.complex128)
    b1 = gpu.coil_b1_from_uh1(au, crecr)
    b1r = O.coil_b1_from_uh1(au, crecr)
    assert rel(b1, b1r) < 1e-4, rel(b1, b1r)
    assert gpu.coil_find_max_sum(crecr) == O.coil_find_max_sum(crecr)
    inds = [O.coil_find_max_sum(crecr), 1]
    x0 = np.zeros((H, W), np.complex128)
    ug = gpu.coil_utgv2(b1r, crecr, inds, x0)
    uo = O.coil_utgv2(b1r, crecr, inds, x0)
    assert rel(ug, uo) < 1e-4, rel(ug, uo)
    xg = gpu.coil_b1_recon(uo, crecr, 300, 2.0)
    xo = O.coil_b1_recon(uo, crecr, 300, 2.0)
    assert rel(xg, xo) < 1e-4, rel(xg, xo)


def test_construction_matches_oracle(gpu):
    print("[gpu coil] PerformCoilConstruction", flush=True)
    k, mask = problem(C=5)
    u, b1, order = gpu.coil_construction(k, mask)
    ur, b1r, order_r = O.coil_construction(k, mask)
    assert list(order) == list(order_r)
    # unit sum of squares per pixel (B1Recon's final normalisation)
    np.testing.assert_allclose((np.abs(b1) ** 2).sum(0), 1.0, rtol=1e-5)
    assert rel(b1, b1r) < 1e-3, rel(b1, b1r)
    assert rel(u, ur) < 1e-3, rel(u, ur)


def radial_problem(W=32, H=28, C=3, T=4, nfe=48, spf=9, seed=4):
    from avionic_amd import synth
    rho = synth.phantom(max(W, H), T, 5, seed)[:, :H, :W]
    b1 = synth.coil_maps(max(W, H), C, seed)[:, :H, :W]
    tr = O.radial_trajectory(nfe, spf, T).astype(np.float32).astype(np.float64)
    dcf = (np.sqrt(tr[:, 0] ** 2 + tr[:, 1] ** 2) + 0.02).astype(np.float32).astype(np.float64)
    k = O.nufft_K(rho, b1, tr, dcf, J=3, osf=2.0)
    k = np.asarray(k, np.complex64).astype(np.complex128)
    return k, tr, dcf


def test_radial_time_average_and_construction(gpu):
    """NoncartesianCoilConstruction: the all-frames time average of the
    library against the oracle's, then the whole construction from it."""
    import avionic_amd as A
    print("[gpu coil] radial time average + construction", flush=True)
    k, tr, dcf = radial_problem()
    T, C, M = k.shape
    H, W = 28, 32
    # the operator's own kernel width / oversampling differ from the time
    # average's fixed 3 / 2
    op = A.NoncartesianOperator(W, H, C, T, 48, 9, gpu._r(tr), gpu._r(dcf), kernel_width=2, osf=1.5)
    cc = A.CoilConstruction(W, H, C, T, op)
    for phase in (True, False):
        u, crec = cc.TimeAveragedReconstruction(gpu._t(k), applyPhase=phase)
        ur, crecr = O.coil_tar_nc(k, tr, dcf, (H, W), apply_phase=phase)
        assert rel(gpu._n(crec), crecr) < 2e-6, rel(gpu._n(crec), crecr)
        assert rel(gpu._n(u), ur) < 2e-6, rel(gpu._n(u), ur)
    u, b1, order = cc.PerformCoilConstruction(gpu._t(k))
    ur, crecr = O.coil_tar_nc(k, tr, dcf, (H, W))
    uo, b1o, order_o = O.coil_construction_from(crecr, ur)
    assert list(order) == list(order_o)
    assert rel(gpu._n(b1), b1o) < 1e-3, rel(gpu._n(b1), b1o)
    assert rel(gpu._n(u), uo) < 1e-3, rel(gpu._n(u), uo)
