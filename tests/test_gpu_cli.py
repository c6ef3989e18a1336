"""The `avionic` command line end to end on the GPU: its result file equals
the library's reconstruction of the same prepared inputs (Cartesian: masked
data, main.cpp:509-519; radial: sqrt(density)-weighted data,
main.cpp:546-570), and its extra-result files follow the export names of
ictgv2.cpp:494-516."""
import os
import subprocess

import numpy as np
import pytest

import avionic_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(A.LIB_PATH), "avionic")


def _run(args):
    print("[gpu cli]", " ".join(args[:6]), flush=True)
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("W,H,C,T", [(32, 24, 3, 4),
                                     # the reference's cine geometry (demo_avionic_cine.sh:73-74:
                                     # nX = 168 encodings, nY = 416 readouts): radix-13 columns
                                     (168, 416, 3, 2)])
def test_cli_cartesian_matches_library(gpu, tmp_path, W, H, C, T):
    rng = np.random.default_rng(5)
    k = (rng.standard_normal((T, C, H, W)) + 1j * rng.standard_normal((T, C, H, W))) * 10
    mask = (rng.random((T, H, 1)) < 0.4) * np.ones((1, 1, W))
    b1 = rng.standard_normal((C, H, W)) * 0.2 + 1.0 + 0.1j
    u0 = rng.standard_normal((H, W)) + 0j
    for name, v in (("k", k), ("m", mask.astype(float)), ("b1", b1), ("u0", u0)):
        A.write_bin(tmp_path / f"{name}.bin", v)
    (tmp_path / "p.cfg").write_text("[ictgv2]\nlambda = 4.0\n")
    _run(["-m", "ICTGV2", "-i", "60", "-p", str(tmp_path / "p.cfg"), "-d",
          f"{W}:{H}:1:{W}:{H}:1:{C}:{T}", "-s", str(tmp_path / "b1.bin"), "-u",
          str(tmp_path / "u0.bin"), "-e", str(tmp_path / "k.bin"), str(tmp_path / "m.bin"),
          str(tmp_path / "x.bin")])
    got = A.read_bin(tmp_path / "x.bin").reshape(T, H, W)
    # the library on the same prepared inputs
    op = A.CartesianOperator(W, H, C, T, gpu._r(mask), centered=False)
    rec = A.ICTGV2(W, H, C, T, op, None)
    p = rec.GetParams()
    p.update(maxIt=60, adaptLambda=0)
    p["lambda"] = 4.0
    rec.SetParams(p)
    x = gpu._t(np.repeat(u0[None], T, 0))
    rec.IterativeReconstruction(gpu._t(k * mask[:, None]), x, gpu._t(b1))
    ref = gpu._n(x)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    # -e writes the solver's extra results next to the output: x3 (the w
    # component) equals the library's export of the same run
    assert os.path.exists(tmp_path / "x3_component")
    x3 = A.read_bin(tmp_path / "x3_component")
    assert x3.size == W * H * T
    ex = {}
    rec.ExportAdditionalResults("", lambda _d, name, arr: ex.__setitem__(name, arr.copy()))
    ref3 = ex["x3_component"].astype(np.complex128).ravel()
    np.testing.assert_allclose(x3.ravel(), ref3, rtol=0, atol=1e-5 * np.abs(ref3).max())


def test_cli_radial_matches_library(gpu, tmp_path):
    rng = np.random.default_rng(6)
    W = H = 32
    C, T, nfe, spf = 2, 3, 48, 7
    M = nfe * spf
    tr = O.radial_trajectory(nfe, spf, T).astype(np.float32).astype(np.float64)
    dcf = np.sqrt(tr[:, 0] ** 2 + tr[:, 1] ** 2) + 0.01
    k = rng.standard_normal((T, C, M)) + 1j * rng.standard_normal((T, C, M))
    b1 = np.ones((C, H, W), complex)
    for name, v in (("k", k), ("traj", tr), ("w", dcf), ("b1", b1)):
        A.write_bin(tmp_path / f"{name}.bin", v)
    (tmp_path / "p.cfg").write_text("[tgv2]\nlambda = 2.0\n[gpunufft]\nkernelWidth = 3\nosf = 1.5\n")
    _run(["-n", "-m", "TGV2", "-i", "40", "-p", str(tmp_path / "p.cfg"), "-d",
          f"{W}:{H}:1:{nfe}:{spf}:1:{C}:{T}", "-s", str(tmp_path / "b1.bin"), "-w",
          str(tmp_path / "w.bin"), str(tmp_path / "k.bin"), str(tmp_path / "traj.bin"),
          str(tmp_path / "x.bin")])
    got = A.read_bin(tmp_path / "x.bin").reshape(T, H, W)
    op = A.NoncartesianOperator(W, H, C, T, nfe, spf, gpu._r(tr), gpu._r(dcf), kernel_width=3,
                                osf=1.5)
    rec = A.TGV2(W, H, C, T, op, None)
    p = rec.GetParams()
    p.update(maxIt=40, adaptLambda=0)
    p["lambda"] = 2.0
    rec.SetParams(p)
    data = (k * np.sqrt(dcf.astype(np.float32))[:, None, :]).astype(np.complex64)
    # -s without -u: u0 = (|crec| + sum_c conj(b1_c) crec_c) / 2 from the
    # all-frames time average (PerformInitalizationGivenB1, main.cpp:250-287)
    tu, crec = A.CoilConstruction(W, H, C, T, op).TimeAveragedReconstruction(gpu._t(data), False)
    tu, crec = gpu._n(tu), gpu._n(crec)
    u0 = 0.5 * (tu + (np.conj(b1) * crec).sum(0))
    x = gpu._t(np.repeat(u0[None], T, 0))
    rec.IterativeReconstruction(gpu._t(data), x, gpu._t(b1))
    ref = gpu._n(x)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def _median_kept(mag):
    """Median of the magnitudes >= 0.9 * max from the first maximum onwards
    (raw_data_preparation.cpp:661-686), in float64, with the distance of the
    nearest magnitude to the threshold (a float32 computation agrees with this
    selection when that distance is well above float32 rounding)."""
    v = np.asarray(mag, np.float64).ravel()
    i = int(np.argmax(v))
    thr = v[i] * 0.9
    kept = np.sort(v[i:][v[i:] >= thr])
    k = kept.size
    med = kept[k // 2] if k % 2 else 0.5 * (kept[k // 2 - 1] + kept[k // 2])
    return med, np.min(np.abs(v[i:] - thr)) / thr


def _disk(rng, T, H, W, r):
    """Frames of a disk of magnitude about 10: a plateau, so that many pixels
    lie within 10 % of the maximum and the selection matters."""
    yy, xx = np.mgrid[0:H, 0:W]
    d = ((yy - H / 2) ** 2 + (xx - W / 2) ** 2) < r * r
    return np.repeat((d * (10 + 0.3 * rng.standard_normal((H, W))))[None], T, 0) + 0j


def test_cli_normalization_cartesian(gpu, tmp_path):
    """-o: the data is scaled by 255 / median of the time-averaged
    root-sum-of-squares image of a centred operator
    (raw_data_preparation.cpp:821-844, cartesian_coil_construction.cpp:23-76);
    the factor is written to datanorm_factor.bin (main.cpp:792-796) and the
    result is the reconstruction of the scaled data."""
    rng = np.random.default_rng(11)
    W, H, C, T = 32, 24, 3, 4
    img = _disk(rng, T, H, W, 8)
    b1 = rng.standard_normal((C, H, W)) * 0.05 + 1.0 + 0.1j
    k = O.cart_K(img, b1, None, centered=True)
    k = k + 0.05 * (rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape))
    mask = (rng.random((T, H, 1)) < 0.5) * np.ones((1, 1, W))
    u0 = rng.standard_normal((H, W)) + 0j
    for name, v in (("k", k), ("m", mask.astype(float)), ("b1", b1), ("u0", u0)):
        A.write_bin(tmp_path / f"{name}.bin", v)
    (tmp_path / "p.cfg").write_text("[tgv2]\nlambda = 3.0\n")
    _run(["-o", "-m", "TGV2", "-i", "30", "-p", str(tmp_path / "p.cfg"), "-d",
          f"{W}:{H}:1:{W}:{H}:1:{C}:{T}", "-s", str(tmp_path / "b1.bin"), "-u",
          str(tmp_path / "u0.bin"), str(tmp_path / "k.bin"), str(tmp_path / "m.bin"),
          str(tmp_path / "x.bin")])
    fac = A.read_bin(tmp_path / "datanorm_factor.bin")
    assert fac.size == 1 and np.iscomplexobj(fac) and fac[0].imag == 0
    # the factor from the fp64 oracle: centred unitary transform of each coil's
    # time average, root-sum-of-squares over coils
    km = k * mask[:, None]
    avg = km.sum(0) / np.maximum(mask.sum(0), 1.0)[None]
    ones = np.ones((1, H, W), complex)
    rss = np.sqrt(sum(np.abs(O.cart_KH(avg[c][None, None], ones, None, centered=True)[0]) ** 2
                      for c in range(C)))
    med, margin = _median_kept(rss)
    assert margin > 1e-4, "fixture too close to the selection threshold"
    # 22 pixels lie above 0.9 * max, 3 of them from the first maximum on
    v = rss.ravel()
    assert (v >= 0.9 * v.max()).sum() > 10 and (v[np.argmax(v):] >= 0.9 * v.max()).sum() < 10
    print(f"[gpu cli] datanorm {fac[0].real:.7g} vs oracle {255 / med:.7g}", flush=True)
    assert fac[0].real == pytest.approx(255.0 / med, rel=1e-5)
    # the result: the library's reconstruction of the scaled data
    got = A.read_bin(tmp_path / "x.bin").reshape(T, H, W)
    op = A.CartesianOperator(W, H, C, T, gpu._r(mask), centered=False)
    rec = A.TGV2(W, H, C, T, op, None)
    p = rec.GetParams()
    p.update(maxIt=30, adaptLambda=0)
    p["lambda"] = 3.0
    rec.SetParams(p)
    x = gpu._t(np.repeat(u0[None], T, 0))
    data = (km.astype(np.complex64) * np.float32(fac[0].real)).astype(np.complex64)
    rec.IterativeReconstruction(gpu._t(data), x, gpu._t(b1))
    ref = gpu._n(x)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def test_cli_normalization_radial(gpu, tmp_path):
    """-o on radial data: 255 * frames / median of the root-sum-of-squares of
    the adjoint gridding of all frames' points together, kernel width 3 and
    grid oversampling 2 (raw_data_preparation.cpp:736-793,
    noncartesian_coil_construction.cpp:21-123), against the fp64 gridding."""
    rng = np.random.default_rng(12)
    W = H = 32
    C, T, nfe, spf = 2, 3, 48, 7
    M = nfe * spf
    tr = O.radial_trajectory(nfe, spf, T).astype(np.float32).astype(np.float64)
    dcf = (np.sqrt(tr[:, 0] ** 2 + tr[:, 1] ** 2) + 0.01).astype(np.float32).astype(np.float64)
    b1 = np.ones((C, H, W), complex)
    k = O.nufft_K(_disk(rng, T, H, W, 9), b1, tr, None, J=3, osf=2.0)
    k = k + 0.05 * (rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape))
    for name, v in (("k", k), ("traj", tr), ("w", dcf), ("b1", b1)):
        A.write_bin(tmp_path / f"{name}.bin", v)
    (tmp_path / "p.cfg").write_text("[tv]\nlambda = 2.0\n")
    _run(["-o", "-n", "-m", "TV", "-i", "10", "-p", str(tmp_path / "p.cfg"), "-d",
          f"{W}:{H}:1:{nfe}:{spf}:1:{C}:{T}", "-s", str(tmp_path / "b1.bin"), "-w",
          str(tmp_path / "w.bin"), str(tmp_path / "k.bin"), str(tmp_path / "traj.bin"),
          str(tmp_path / "x.bin")])
    fac = A.read_bin(tmp_path / "datanorm_factor.bin")
    # all frames as one: trajectory [1][2][T*M], density [1][T*M]; the data
    # carries sqrt(density) as prepared (main.cpp:546-570)
    trc = np.concatenate([tr[:, 0].ravel()[None], tr[:, 1].ravel()[None]])[None]
    dc = dcf.reshape(1, T * M)
    kw = k * np.sqrt(dcf)[:, None, :]
    ones = np.ones((1, H, W), complex)
    rss = np.sqrt(sum(np.abs(O.nufft_KH(kw[:, c].reshape(1, 1, T * M), ones, trc, dc, J=3,
                                        osf=2.0)[0]) ** 2 for c in range(C)))
    med, margin = _median_kept(rss)
    assert margin > 1e-4, "fixture too close to the selection threshold"
    print(f"[gpu cli] radial datanorm {fac[0].real:.7g} vs oracle {255 * T / med:.7g}", flush=True)
    assert fac[0].real == pytest.approx(255.0 * T / med, rel=1e-4)


@pytest.mark.parametrize("method", ["TVTEMP", "ICTV"])
def test_cli_other_regularisers_match_library(gpu, tmp_path, method):
    """-m TVTEMP / ICTV with their [tvtemp] / [ictv] sections
    (options_parser.cpp:204-210, 240-258) equal the library's solver on the
    same prepared inputs."""
    rng = np.random.default_rng(13)
    W, H, C, T = 24, 20, 2, 4
    k = (rng.standard_normal((T, C, H, W)) + 1j * rng.standard_normal((T, C, H, W))) * 10
    mask = (rng.random((T, H, 1)) < 0.5) * np.ones((1, 1, W))
    b1 = rng.standard_normal((C, H, W)) * 0.2 + 1.0 + 0.1j
    u0 = rng.standard_normal((H, W)) + 0j
    for name, v in (("k", k), ("m", mask.astype(float)), ("b1", b1), ("u0", u0)):
        A.write_bin(tmp_path / f"{name}.bin", v)
    if method == "TVTEMP":
        cfg, extra = "[tvtemp]\nlambda = 2.0\ndt = 0.8\n", dict(dt=0.8)
    else:
        cfg = ("[ictv]\nlambda = 2.0\nsigma = 0.1\ntau = 0.1\ntimeSpaceWeight = 3\n"
               "timeSpaceWeight2 = 0.7\nalpha = 0.4\nalpha1 = 1.2\n")
        extra = dict(sigma=0.1, tau=0.1, timeSpaceWeight=3.0, timeSpaceWeight2=0.7, alpha=0.4,
                     alpha1=1.2)
    (tmp_path / "p.cfg").write_text(cfg)
    _run(["-m", method, "-i", "40", "-p", str(tmp_path / "p.cfg"), "-d",
          f"{W}:{H}:1:{W}:{H}:1:{C}:{T}", "-s", str(tmp_path / "b1.bin"), "-u",
          str(tmp_path / "u0.bin"), str(tmp_path / "k.bin"), str(tmp_path / "m.bin"),
          str(tmp_path / "x.bin")])
    got = A.read_bin(tmp_path / "x.bin").reshape(T, H, W)
    op = A.CartesianOperator(W, H, C, T, gpu._r(mask), centered=False)
    rec = getattr(A, method)(W, H, C, T, op, None)
    p = rec.GetParams()
    p.update(maxIt=40, adaptLambda=0, **extra)
    p["lambda"] = 2.0
    rec.SetParams(p)
    x = gpu._t(np.repeat(u0[None], T, 0))
    rec.IterativeReconstruction(gpu._t(k * mask[:, None]), x, gpu._t(b1))
    ref = gpu._n(x)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def test_cli_tgv2_3d_matches_library(gpu, tmp_path):
    """-m TGV2_3D on one volume (-d w:h:depth:...:c:1) with the [tgv2_3D]
    section (options_parser.cpp:227-237): coil-major k-space masked per coil
    (main.cpp:466-474), equal to the library's volumetric solver."""
    rng = np.random.default_rng(14)
    W, H, D, C = 16, 12, 6, 3
    k = (rng.standard_normal((C, D, H, W)) + 1j * rng.standard_normal((C, D, H, W))) * 10
    mask = (rng.random((D, H, W)) < 0.5).astype(float)
    b1 = rng.standard_normal((C, D, H, W)) * 0.2 + 1.0 + 0.1j
    u0 = rng.standard_normal((D, H, W)) + 0j
    for name, v in (("k", k), ("m", mask), ("b1", b1), ("u0", u0)):
        A.write_bin(tmp_path / f"{name}.bin", v)
    (tmp_path / "p.cfg").write_text("[tgv2_3D]\nlambda = 3.0\ndz = 2.0\n")
    _run(["-m", "TGV2_3D", "-i", "40", "-p", str(tmp_path / "p.cfg"), "-d",
          f"{W}:{H}:{D}:{W}:{H}:{D}:{C}:1", "-s", str(tmp_path / "b1.bin"), "-u",
          str(tmp_path / "u0.bin"), str(tmp_path / "k.bin"), str(tmp_path / "m.bin"),
          str(tmp_path / "x.bin")])
    got = A.read_bin(tmp_path / "x.bin").reshape(D, H, W)
    op = A.CartesianOperator3D(W, H, D, C, gpu._r(mask))
    rec = A.TGV2_3D(W, H, C, D, op, None)
    p = rec.GetParams()
    p.update(maxIt=40, adaptLambda=0, dz=2.0)
    p["lambda"] = 3.0
    rec.SetParams(p)
    x = gpu._t(u0)
    rec.IterativeReconstruction(gpu._t(k * mask[None]), x, gpu._t(b1))
    ref = gpu._n(x)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def _coil_phantom(rng, T, C, H, W):
    from avionic_amd import synth
    rho = synth.phantom(max(W, H), T, 5, 7)[:, :H, :W]
    b1 = synth.coil_maps(max(W, H), C, 7)[:, :H, :W]
    return rho, b1


def test_cli_coil_construction_cartesian(gpu, tmp_path):
    """No -s: the coil maps and u0 come from PerformCoilConstruction on the
    uncentred masked operator (main.cpp:618-649); b1_reconstructed.bin and
    u0_reconstructed.bin hold them and the reconstruction starts from u0."""
    rng = np.random.default_rng(8)
    W, H, C, T = 40, 32, 4, 5
    rho, b1t = _coil_phantom(rng, T, C, H, W)
    mask = (rng.random((T, H, 1)) < 0.5) * np.ones((1, 1, W))
    mask[:, :4] = 1
    k = np.stack([np.fft.ifft2(b1t * rho[t][None], norm="ortho") for t in range(T)]) * 100
    for name, v in (("k", k), ("m", mask.astype(float))):
        A.write_bin(tmp_path / f"{name}.bin", v)
    (tmp_path / "p.cfg").write_text("[tgv2]\nlambda = 4.0\n")
    _run(["-m", "TGV2", "-i", "20", "-p", str(tmp_path / "p.cfg"), "-d",
          f"{W}:{H}:1:{W}:{H}:1:{C}:{T}", str(tmp_path / "k.bin"), str(tmp_path / "m.bin"),
          str(tmp_path / "x.bin")])
    op = A.CartesianOperator(W, H, C, T, gpu._r(mask), centered=False)
    data = gpu._t(k * mask[:, None])
    u, b1, _ = A.CoilConstruction(W, H, C, T, op).PerformCoilConstruction(data)
    u0, b1 = gpu._n(u)[C - 1], gpu._n(b1)
    got_b1 = A.read_bin(tmp_path / "b1_reconstructed.bin").reshape(C, H, W)
    got_u0 = A.read_bin(tmp_path / "u0_reconstructed.bin").reshape(H, W)
    np.testing.assert_allclose(got_b1, b1, rtol=0, atol=1e-5 * np.abs(b1).max())
    np.testing.assert_allclose(got_u0, u0, rtol=0, atol=1e-5 * np.abs(u0).max())
    rec = A.TGV2(W, H, C, T, op, None)
    p = rec.GetParams()
    p.update(maxIt=20, adaptLambda=0)
    p["lambda"] = 4.0
    rec.SetParams(p)
    x = gpu._t(np.repeat(u0[None], T, 0))
    rec.IterativeReconstruction(data, x, gpu._t(b1))
    ref = gpu._n(x)
    got = A.read_bin(tmp_path / "x.bin").reshape(T, H, W)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def test_cli_coil_construction_radial(gpu, tmp_path):
    """No -s on radial data: PerformNonCartesianCoilConstruction
    (main.cpp:225-247), u0 divided by the frame count (:641-644)."""
    rng = np.random.default_rng(9)
    W = H = 32
    C, T, nfe, spf = 3, 4, 48, 9
    M = nfe * spf
    rho, b1t = _coil_phantom(rng, T, C, H, W)
    tr = O.radial_trajectory(nfe, spf, T).astype(np.float32).astype(np.float64)
    dcf = np.sqrt(tr[:, 0] ** 2 + tr[:, 1] ** 2) + 0.01
    k = O.nufft_K(rho, b1t, tr, None, J=3, osf=2.0) * 100
    for name, v in (("k", k), ("traj", tr), ("w", dcf)):
        A.write_bin(tmp_path / f"{name}.bin", v)
    (tmp_path / "p.cfg").write_text("[tv]\nlambda = 2.0\n[gpunufft]\nkernelWidth = 3\nosf = 1.5\n")
    _run(["-n", "-m", "TV", "-i", "20", "-p", str(tmp_path / "p.cfg"), "-d",
          f"{W}:{H}:1:{nfe}:{spf}:1:{C}:{T}", "-w", str(tmp_path / "w.bin"), str(tmp_path / "k.bin"),
          str(tmp_path / "traj.bin"), str(tmp_path / "x.bin")])
    op = A.NoncartesianOperator(W, H, C, T, nfe, spf, gpu._r(tr), gpu._r(dcf), kernel_width=3,
                                osf=1.5)
    data = gpu._t((k * np.sqrt(dcf.astype(np.float32))[:, None, :]).astype(np.complex64))
    u, b1, _ = A.CoilConstruction(W, H, C, T, op).PerformCoilConstruction(data)
    u0, b1 = gpu._n(u)[C - 1] / T, gpu._n(b1)
    got_b1 = A.read_bin(tmp_path / "b1_reconstructed.bin").reshape(C, H, W)
    got_u0 = A.read_bin(tmp_path / "u0_reconstructed.bin").reshape(H, W)
    np.testing.assert_allclose(got_b1, b1, rtol=0, atol=1e-5 * np.abs(b1).max())
    np.testing.assert_allclose(got_u0, u0, rtol=0, atol=1e-5 * np.abs(u0).max())
    rec = A.TV(W, H, C, T, op, None)
    p = rec.GetParams()
    p.update(maxIt=20, adaptLambda=0)
    p["lambda"] = 2.0
    rec.SetParams(p)
    x = gpu._t(np.repeat(u0[None], T, 0))
    rec.IterativeReconstruction(data, x, gpu._t(b1))
    ref = gpu._n(x)
    got = A.read_bin(tmp_path / "x.bin").reshape(T, H, W)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def test_cli_given_b1_initial_image_cartesian(gpu, tmp_path):
    """-s without -u on Cartesian data: u0 = (|crec| + sum_c conj(b1_c)
    crec_c) / 2 from the operator's own time average."""
    rng = np.random.default_rng(10)
    W, H, C, T = 24, 20, 3, 4
    rho, b1 = _coil_phantom(rng, T, C, H, W)
    mask = (rng.random((T, H, 1)) < 0.5) * np.ones((1, 1, W))
    k = np.stack([np.fft.ifft2(b1 * rho[t][None], norm="ortho") for t in range(T)]) * 100
    for name, v in (("k", k), ("m", mask.astype(float)), ("b1", b1)):
        A.write_bin(tmp_path / f"{name}.bin", v)
    (tmp_path / "p.cfg").write_text("[tv]\nlambda = 3.0\n")
    _run(["-m", "TV", "-i", "15", "-p", str(tmp_path / "p.cfg"), "-d",
          f"{W}:{H}:1:{W}:{H}:1:{C}:{T}", "-s", str(tmp_path / "b1.bin"), str(tmp_path / "k.bin"),
          str(tmp_path / "m.bin"), str(tmp_path / "x.bin")])
    assert not os.path.exists(tmp_path / "b1_reconstructed.bin")
    op = A.CartesianOperator(W, H, C, T, gpu._r(mask), centered=False)
    data = gpu._t(k * mask[:, None])
    tu, crec = A.CoilConstruction(W, H, C, T, op).TimeAveragedReconstruction(data, False)
    tu, crec = gpu._n(tu), gpu._n(crec)
    b1c = b1.astype(np.complex64).astype(np.complex128)
    u0 = 0.5 * (tu + (np.conj(b1c) * crec).sum(0))
    rec = A.TV(W, H, C, T, op, None)
    p = rec.GetParams()
    p.update(maxIt=15, adaptLambda=0)
    p["lambda"] = 3.0
    rec.SetParams(p)
    x = gpu._t(np.repeat(u0[None], T, 0))
    rec.IterativeReconstruction(data, x, gpu._t(b1))
    ref = gpu._n(x)
    got = A.read_bin(tmp_path / "x.bin").reshape(T, H, W)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
