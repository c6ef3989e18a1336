"""avionic_operator_backward_noalloc (K without device allocation) against
avionic_operator_backward, called through the C-ABI: bitwise equal on two
calls in a row (the second on a smaller operator, so stale scratch contents
from the first must not leak into it), plus the size and operator-kind
rejections."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

AVIONIC_E_INVALID, AVIONIC_E_UNSUPPORTED = -1, -7


def _lib():
    import avionic_amd as A
    L = A.lib()
    P = ctypes.c_void_p
    L.avionic_operator_backward_noalloc.restype = ctypes.c_int
    L.avionic_operator_backward_noalloc.argtypes = [P, P, P, P, P]
    return A, L


def _operator(A, L, W, H, C, T, mask):
    op = ctypes.c_void_p()
    mp = None if mask is None else ctypes.c_void_p(mask.data_ptr())
    rc = L.avionic_cartesian_operator_create(W, H, C, T, mp, 0, ctypes.byref(op))
    assert rc == 0, L.avionic_last_error()
    return op


def _run_pair(A, L, torch, W, H, C, T, seed):
    print(f"[no-alloc] K {W}x{H} C={C} T={T} seed={seed}", flush=True)
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.randn(T, H, W, dtype=torch.complex64, generator=g).to(dev)
    b1 = torch.randn(C, H, W, dtype=torch.complex64, generator=g).to(dev)
    mask = (torch.rand(T, H, 1, generator=g) < 0.3).float().expand(T, H, W).contiguous().to(dev)
    op = _operator(A, L, W, H, C, T, mask)
    try:
        ref = torch.empty(T, C, H, W, dtype=torch.complex64, device=dev)
        got = torch.full((T, C, H, W), 7.0 + 3.0j, dtype=torch.complex64, device=dev)
        P = ctypes.c_void_p
        rc = L.avionic_operator_backward(op, P(x.data_ptr()), P(ref.data_ptr()),
                                         P(b1.data_ptr()), None)
        assert rc == 0, L.avionic_last_error()
        rc = L.avionic_operator_backward_noalloc(op, P(x.data_ptr()), P(got.data_ptr()),
                                                 P(b1.data_ptr()), None)
        assert rc == 0, L.avionic_last_error()
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(got), torch.view_as_real(ref))
        assert float(ref.abs().max()) > 0.0
    finally:
        L.avionic_operator_destroy(op)


def test_noalloc_matches_backward_twice_second_smaller():
    import torch
    A, L = _lib()
    _run_pair(A, L, torch, 256, 192, 12, 10, seed=1)
    _run_pair(A, L, torch, 96, 80, 3, 4, seed=2)


def test_noalloc_rejects_oversized_and_noncartesian():
    import torch
    A, L = _lib()
    dev = torch.device("cuda:0")
    P = ctypes.c_void_p
    # fully sampled 512 x 512 x 8 coils x 5 frames: 10.5 M compact values > 8 M
    W, H, C, T = 512, 512, 8, 5
    op = _operator(A, L, W, H, C, T, None)
    try:
        x = torch.zeros(T, H, W, dtype=torch.complex64, device=dev)
        b1 = torch.zeros(C, H, W, dtype=torch.complex64, device=dev)
        k = torch.zeros(T, C, H, W, dtype=torch.complex64, device=dev)
        rc = L.avionic_operator_backward_noalloc(op, P(x.data_ptr()), P(k.data_ptr()),
                                                 P(b1.data_ptr()), None)
        assert rc == AVIONIC_E_INVALID
        assert b"exceeds the no-allocation scratch" in L.avionic_last_error()
    finally:
        L.avionic_operator_destroy(op)
    # a non-Cartesian operator is refused before any launch
    nfe, spf, T, C, W = 32, 5, 2, 2, 24
    r = (np.arange(nfe) - nfe / 2) / nfe
    traj = np.zeros((T, 2, spf * nfe), np.float32)
    for t in range(T):
        a = (t * spf + np.arange(spf)) * 2.39996
        traj[t, 0] = (r[None] * np.cos(a)[:, None]).ravel()
        traj[t, 1] = (r[None] * np.sin(a)[:, None]).ravel()
    tr = torch.from_numpy(traj).to(dev)
    op = ctypes.c_void_p()
    rc = L.avionic_noncartesian_operator_create(W, W, C, T, nfe, spf, P(tr.data_ptr()), None, 3,
                                                ctypes.c_double(1.5), ctypes.byref(op))
    assert rc == 0, L.avionic_last_error()
    try:
        x = torch.zeros(T, W, W, dtype=torch.complex64, device=dev)
        b1 = torch.zeros(C, W, W, dtype=torch.complex64, device=dev)
        k = torch.zeros(T, C, spf * nfe, dtype=torch.complex64, device=dev)
        rc = L.avionic_operator_backward_noalloc(op, P(x.data_ptr()), P(k.data_ptr()),
                                                 P(b1.data_ptr()), None)
        assert rc == AVIONIC_E_UNSUPPORTED
    finally:
        L.avionic_operator_destroy(op)
