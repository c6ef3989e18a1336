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


def test_cli_cartesian_matches_library(gpu, tmp_path):
    rng = np.random.default_rng(5)
    W, H, C, T = 32, 24, 3, 4
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
    # -e writes the solver's extra results next to the output
    assert os.path.exists(tmp_path / "x3_component")
    assert A.read_bin(tmp_path / "x3_component").size == W * H * T


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
    x = gpu._t(np.zeros((T, H, W)))
    data = (k * np.sqrt(dcf.astype(np.float32))[:, None, :]).astype(np.complex64)
    rec.IterativeReconstruction(gpu._t(data), x, gpu._t(b1))
    ref = gpu._n(x)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
