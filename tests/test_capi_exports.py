"""The C-ABI library loads without a GPU and exports every function that
include/avionic.h declares; the Python mirror exposes the reference API."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "avionic.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(avionic_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(n for n in names if not n.endswith("_cb")))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 40
    for must in ("avionic_recon_run", "avionic_operator_forward", "avionic_operator_backward",
                 "avionic_recon_create", "avionic_engine_submit"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import avionic_amd as A
    L = ctypes.CDLL(A.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_host_only_entry_points_work_without_gpu():
    import avionic_amd as A
    assert "gfx950" in A.version()
    assert abs(A.ellipke(0.5) - 1.3506438810476753) < 1e-12
    ds, dt = A.time_space_weights(4.0)
    assert abs(ds - 2.730596) < 1e-6 and abs(dt - 0.682649) < 1e-6
    with pytest.raises(A.AvionicError) as e:
        A.time_space_weights(-1.0)
    assert e.value.code == -1 and "Invalid ratio" in str(e.value)
    for m, tsw in ((0, 5.0), (1, 5.0), (2, 6.5)):
        p = A.default_params(m)
        assert p["maxIt"] == 500 and abs(p["sigma"] - 1 / 3) < 1e-15
        assert abs(p["timeSpaceWeight"] - tsw) < 1e-15


def test_python_mirror_has_reference_methods():
    import avionic_amd as A
    for cls in (A.TV, A.TGV2, A.ICTGV2):
        for m in ("IterativeReconstruction", "GetParams", "AdaptLambda", "SetVerbose", "SetDebug",
                  "AdaptStepSize", "ComputeGStar", "ComputePDGap", "ExportAdditionalResults"):
            assert hasattr(cls, m)
    for m in ("ForwardOperation", "BackwardOperation", "AdaptLambda"):
        assert hasattr(A.CartesianOperator, m) and hasattr(A.NoncartesianOperator, m)
