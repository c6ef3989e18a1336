"""The reference's .bin vector format (matlab/writebin_vector.m:35-44,
readbin_vector.m:30-40) through the C-ABI, and the `avionic` command line's
argument, parameter-file and input checking (--dry-run: no GPU needed)."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import avionic_amd as A

CLI = os.path.join(os.path.dirname(A.LIB_PATH), "avionic")


def matlab_style_write(path, v):
    """Byte layout written by writebin_vector.m, restated with struct."""
    v = np.asarray(v).ravel()
    cplx = np.iscomplexobj(v)
    with open(path, "wb") as f:
        f.write(struct.pack("<BI", 1 if cplx else 0, v.size))
        f.write(np.ascontiguousarray(v.real, "<f8").tobytes())
        if cplx:
            f.write(np.ascontiguousarray(v.imag, "<f8").tobytes())


def test_bin_roundtrip_and_matlab_layout(tmp_path):
    rng = np.random.default_rng(0)
    z = rng.standard_normal(37) + 1j * rng.standard_normal(37)
    r = rng.standard_normal(11)
    A.write_bin(tmp_path / "z.bin", z)
    A.write_bin(tmp_path / "r.bin", r)
    np.testing.assert_array_equal(A.read_bin(tmp_path / "z.bin"), z)
    np.testing.assert_array_equal(A.read_bin(tmp_path / "r.bin"), r)
    # byte-identical to the MATLAB writer's layout, both directions
    matlab_style_write(tmp_path / "m.bin", z)
    assert (tmp_path / "m.bin").read_bytes() == (tmp_path / "z.bin").read_bytes()
    np.testing.assert_array_equal(A.read_bin(tmp_path / "m.bin"), z)
    raw = (tmp_path / "r.bin").read_bytes()
    assert raw[0] == 0 and struct.unpack("<I", raw[1:5])[0] == 11 and len(raw) == 5 + 88


def test_bin_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(struct.pack("<BI", 1, 10) + b"\0" * 80)  # complex header, real-sized payload
    with pytest.raises(A.AvionicError, match="size mismatch"):
        A.read_bin(p)
    (tmp_path / "short.bin").write_bytes(b"\1\2")
    with pytest.raises(A.AvionicError, match="truncated"):
        A.read_bin(tmp_path / "short.bin")
    with pytest.raises(A.AvionicError, match="cannot open"):
        A.read_bin(tmp_path / "missing.bin")


def _problem(tmp_path, W=8, H=6, C=2, T=3):
    rng = np.random.default_rng(1)
    A.write_bin(tmp_path / "k.bin", rng.standard_normal(T * C * H * W) + 0j)
    A.write_bin(tmp_path / "m.bin", (rng.random(T * H * W) < 0.5).astype(float))
    A.write_bin(tmp_path / "b1.bin", np.ones(C * H * W, complex))
    (tmp_path / "p.cfg").write_text(
        "# test parameters\nmethod = TGV2\nmaxIt = 77\n[tgv2]\nlambda = 2.5\n"
        "alpha0 = 1.7 # inline comment\ntimeSpaceWeight = 3\n[adaptlambda]\nk = 0.5\nd = 2\n")
    return f"{W}:{H}:1:{W}:{H}:1:{C}:{T}"


def _run(args, ok=True):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=60)
    if ok:
        assert r.returncode == 0, r.stderr
    return r


@pytest.mark.skipif(not os.path.exists(CLI), reason="avionic CLI not built")
def test_cli_resolves_config_and_overrides(tmp_path):
    dims = _problem(tmp_path)
    base = ["-d", dims, "-s", str(tmp_path / "b1.bin"), "-p", str(tmp_path / "p.cfg"),
            str(tmp_path / "k.bin"), str(tmp_path / "m.bin"), str(tmp_path / "out.bin"), "--dry-run"]
    cfg = json.loads(_run(base).stdout)
    assert cfg["method"] == "TGV2" and cfg["maxIt"] == 77
    assert cfg["lambda"] == 2.5 and cfg["alpha0"] == 1.7 and cfg["timeSpaceWeight"] == 3
    assert cfg["adaptLambda"] == 0 and cfg["initial_image"] == "time-averaged"
    cfg = json.loads(_run(["-m", "tv", "-i", "5", "-a"] + base).stdout)
    assert cfg["method"] == "TV" and cfg["maxIt"] == 5
    assert cfg["adaptLambda"] == 1 and cfg["adaptLambdaK"] == 0.5 and cfg["adaptLambdaD"] == 2
    assert cfg["normalize"] is False
    assert json.loads(_run(["-o"] + base).stdout)["normalize"] is True


@pytest.mark.skipif(not os.path.exists(CLI), reason="avionic CLI not built")
def test_cli_rejects_what_it_cannot_do(tmp_path):
    dims = _problem(tmp_path)
    files = [str(tmp_path / "k.bin"), str(tmp_path / "m.bin"), str(tmp_path / "out.bin")]
    # without -s the coil maps are estimated (2D+t); a volume needs them given
    cfg = json.loads(_run(["-d", dims, "-p", str(tmp_path / "p.cfg")] + files + ["--dry-run"]).stdout)
    assert cfg["coil_maps"] == "coil-construction" and cfg["initial_image"] == "coil-construction"
    r = _run(["-m", "TGV2_3D", "-a", "-d", "8:6:3:8:6:3:3:1", "-p", str(tmp_path / "p.cfg")] + files +
             ["--dry-run"], ok=False)
    assert r.returncode == 3 and "coil maps" in r.stderr
    # wrong sizes
    r = _run(["-d", "8:6:1:8:6:1:3:3", "-s", str(tmp_path / "b1.bin"), "-p",
              str(tmp_path / "p.cfg")] + files + ["--dry-run"], ok=False)
    assert r.returncode == 2 and "k-space data has" in r.stderr
    # a volume needs -m TGV2_3D
    r = _run(["-d", "8:6:3:8:6:1:3:1", "-s", str(tmp_path / "b1.bin"), "-p",
              str(tmp_path / "p.cfg")] + files + ["--dry-run"], ok=False)
    assert r.returncode == 3 and "TGV2_3D" in r.stderr
    for flag, code in (("-r", 3), ("--bogus", 2)):
        r = _run([flag, "-d", dims] + files, ok=False)
        assert r.returncode == code, (flag, r.stderr)
    r = _run(["-m", "ASLTGV2RECON4D", "-d", dims, "-s", str(tmp_path / "b1.bin"), "-p",
              str(tmp_path / "p.cfg")] + files + ["--dry-run"], ok=False)
    assert r.returncode == 3 and "not part of this build" in r.stderr
    # no lambda and no adaptation
    r = _run(["-d", dims, "-s", str(tmp_path / "b1.bin")] + files + ["--dry-run"], ok=False)
    assert r.returncode == 2 and "lambda" in r.stderr
    # a key the reference's parameter file does not define (boost's
    # parse_config_file throws on it, options_parser.cpp:393)
    for bad in ("[tgv2]\nlamda = 2.5\n", "[tgv3]\nlambda = 2.5\n", "maxIter = 5\n"):
        (tmp_path / "bad.cfg").write_text(bad)
        r = _run(["-d", dims, "-s", str(tmp_path / "b1.bin"), "-p", str(tmp_path / "bad.cfg")] +
                 files + ["--dry-run"], ok=False)
        assert r.returncode != 0 and "unrecognised option" in r.stderr, bad
