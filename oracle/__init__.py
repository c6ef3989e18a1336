"""FP64 CPU oracle for the AVIONIC primal-dual loop (ctypes front end).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the timed
CPU baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package.  The product (``avionic_amd``)
never imports it.

All arrays are numpy ``complex128`` (images, k-space, coil maps) or
``float64`` (masks, trajectories), in the reference's CUDA layout:
image ``[T, H, W]``, k-space ``[T, C, H, W]``, mask ``[T, H, W]``,
coil maps ``[C, H, W]``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

TV, TGV2, ICTGV2, TVTEMP, ICTV, TGV2_3D = 0, 1, 2, 3, 4, 5
METHODS = {"tv": TV, "tgv2": TGV2, "ictgv2": ICTGV2, "tvtemp": TVTEMP, "ictv": ICTV,
           "tgv2_3d": TGV2_3D}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int),
        ("maxIt", ctypes.c_int),
        ("sigma", ctypes.c_double),
        ("tau", ctypes.c_double),
        ("sigmaTauRatio", ctypes.c_double),
        ("tsw", ctypes.c_double),
        ("tsw2", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("alpha0", ctypes.c_double),
        ("alpha1", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("schedule", ctypes.c_int),
        ("gapEvery", ctypes.c_int),
        ("centered", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("dz", ctypes.c_double),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def _declare(L):
    L.orc_fft1.argtypes = [_P, _I, _I]
    L.orc_fft2.argtypes = [_P, _I, _I, _I, _I]
    L.orc_cart_K.argtypes = [_P, _P, _P, _I, _I, _I, _I, _I, _P]
    L.orc_cart_KH.argtypes = [_P, _P, _P, _I, _I, _I, _I, _I, _P]
    for f in ("orc_grad", "orc_symgrad", "orc_div", "orc_symdiv"):
        getattr(L, f).argtypes = [_P, _I, _I, _I, _D, _D, _D, _P]
    L.orc_prox3.argtypes = [_P, ctypes.c_long, _D]
    L.orc_prox6.argtypes = [_P, ctypes.c_long, _D]
    L.orc_ellipke.argtypes = [_D]
    L.orc_ellipke.restype = _D
    L.orc_time_space_weights.argtypes = [_D, _P, _P]
    L.orc_time_space_weights.restype = _I
    L.orc_adapt_lambda_cart.argtypes = [_P, _I, _I, _I, _D, _D]
    L.orc_adapt_lambda_cart.restype = _D
    L.orc_adapt_step.argtypes = [_D, _D, _D, _P, _P]
    L.orc_tv_norm.argtypes = [_P, _I, _I, _I, _D, _D, _D]
    L.orc_tv_norm.restype = _D
    for f in ("orc_tgrad", "orc_tgrad_T"):
        getattr(L, f).argtypes = [_P, _I, _I, _I, _D, _P]
        getattr(L, f).restype = None
    L.orc_tvtemp_norm.argtypes = [_P, _I, _I, _I, _D]
    L.orc_tvtemp_norm.restype = _D
    L.orc_ictv_norm.argtypes = [_P, _P, _I, _I, _I, _D, _D, _P]
    L.orc_ictv_norm.restype = _D
    L.orc_fft3.argtypes = [_P, _I, _I, _I, _I]
    L.orc_fft3.restype = None
    for f in ("orc_cart3_K", "orc_cart3_KH"):
        getattr(L, f).argtypes = [_P, _P, _P, _I, _I, _I, _I, _P]
        getattr(L, f).restype = None
    L.orc_recon3.argtypes = [ctypes.c_void_p, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I]
    L.orc_recon3.restype = _I
    L.orc_tgv2_norm.argtypes = [_P, _P, _I, _I, _I, _D, _D, _D, _D, _D]
    L.orc_tgv2_norm.restype = _D
    L.orc_ictgv2_norm.argtypes = [_P, _P, _P, _P, _I, _I, _I] + [_D] * 9
    L.orc_ictgv2_norm.restype = _D
    L.orc_gstar.argtypes = [_I, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _D, _P]
    L.orc_gstar.restype = _D
    L.orc_pdgap.argtypes = [_I, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I,
                            _D, _D, _D, _D, _P]
    L.orc_pdgap.restype = _D
    L.orc_solver_adapt_step.argtypes = [_I, _P, _P, _P, _I, _I, _I, _I, _I, _P,
                                        _D, _P, _P, _P, _P]
    L.orc_recon.argtypes = [ctypes.POINTER(OrcParams), _I, _I, _I, _I, _P, _P,
                            _P, _P, _P, _P, _P, _P, _P, _I]
    L.orc_recon.restype = _I
    L.orc_primal_components.argtypes = [_I]
    L.orc_primal_components.restype = _I
    L.orc_dual_components.argtypes = [_I]
    L.orc_dual_components.restype = _I
    if hasattr(L, "orc_nufft_K"):
        L.orc_kb_beta.argtypes = [_I, _D]
        L.orc_kb_beta.restype = _D
        for f in ("orc_nufft_K", "orc_nufft_KH"):
            getattr(L, f).argtypes = [_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _D, _P]
        L.orc_nudft_K.argtypes = [_P, _P, _P, _P, _I, _I, _I, _I, _I, _P]


def _c(a, dtype=np.complex128):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------- FFT
def fft1(a, sign):
    a = _c(a).copy()
    lib().orc_fft1(_ptr(a), a.shape[0], int(sign))
    return a


def fft2(a, sign, centered=False):
    a = _c(a).copy()
    H, W = a.shape
    lib().orc_fft2(_ptr(a), H, W, int(sign), int(bool(centered)))
    return a


# ---------------------------------------------------------------------- operator
def cart_K(x, b1, mask=None, centered=False):
    x, b1 = _c(x), _c(b1)
    T, H, W = x.shape
    C = b1.shape[0]
    m = None if mask is None else _c(mask, np.float64)
    z = np.zeros((T, C, H, W), np.complex128)
    lib().orc_cart_K(_ptr(x), _ptr(b1), _ptr(m), W, H, C, T, int(bool(centered)), _ptr(z))
    return z


def cart_KH(z, b1, mask=None, centered=False):
    z, b1 = _c(z), _c(b1)
    T, C, H, W = z.shape
    m = None if mask is None else _c(mask, np.float64)
    x = np.zeros((T, H, W), np.complex128)
    lib().orc_cart_KH(_ptr(z), _ptr(b1), _ptr(m), W, H, C, T, int(bool(centered)), _ptr(x))
    return x


# ---------------------------------------------------------------------- stencils
def _stencil(name, inp, ncomp_out, dims, dx, dy, dt):
    T, H, W = dims
    out = np.zeros((ncomp_out, T, H, W), np.complex128)
    getattr(lib(), name)(_ptr(inp), W, H, T, float(dx), float(dy), float(dt), _ptr(out))
    return out


def grad(u, dx=1.0, dy=1.0, dt=1.0):
    u = _c(u)
    return _stencil("orc_grad", u, 3, u.shape, dx, dy, dt)


def symgrad(v, dx=1.0, dy=1.0, dt=1.0):
    v = _c(v)
    return _stencil("orc_symgrad", v, 6, v.shape[1:], dx, dy, dt)


def div(p, dx=1.0, dy=1.0, dt=1.0):
    p = _c(p)
    return _stencil("orc_div", p, 1, p.shape[1:], dx, dy, dt)[0]


def symdiv(q, dx=1.0, dy=1.0, dt=1.0):
    q = _c(q)
    return _stencil("orc_symdiv", q, 3, q.shape[1:], dx, dy, dt)


def prox3(y, radius):
    y = _c(y).copy()
    lib().orc_prox3(_ptr(y), y[0].size, float(radius))
    return y


def prox6(q, radius):
    q = _c(q).copy()
    lib().orc_prox6(_ptr(q), q[0].size, float(radius))
    return q


# ----------------------------------------------------------------------- scalars
def ellipke(m):
    return lib().orc_ellipke(float(m))


def time_space_weights(tsw):
    ds, dt = ctypes.c_double(), ctypes.c_double()
    rc = lib().orc_time_space_weights(float(tsw), ctypes.byref(ds), ctypes.byref(dt))
    if rc != 0:
        raise ValueError("ComputeTimeSpaceWeights: Invalid ratio.")
    return ds.value, dt.value


def adapt_lambda_cart(mask, W, H, T, k, d):
    m = None if mask is None else _c(mask, np.float64)
    return lib().orc_adapt_lambda_cart(_ptr(m), W, H, T, float(k), float(d))


def adapt_step(nKx, nx, ratio, sigma, tau):
    s, t = ctypes.c_double(sigma), ctypes.c_double(tau)
    lib().orc_adapt_step(float(nKx), float(nx), float(ratio), ctypes.byref(s), ctypes.byref(t))
    return s.value, t.value


def solver_adapt_step(method, delta, b1, mask, centered, d6, ratio, sigma, tau):
    """delta: [P, T, H, W].  Returns (sigma, tau, nKx, nx)."""
    delta, b1 = _c(delta), _c(b1)
    P, T, H, W = delta.shape
    C = b1.shape[0]
    m = None if mask is None else _c(mask, np.float64)
    d6 = np.ascontiguousarray(d6, np.float64)
    s, t = ctypes.c_double(sigma), ctypes.c_double(tau)
    nk, nx = ctypes.c_double(), ctypes.c_double()
    lib().orc_solver_adapt_step(int(method), _ptr(delta), _ptr(b1), _ptr(m), W, H, C, T,
                                int(bool(centered)), _ptr(d6), float(ratio),
                                ctypes.byref(s), ctypes.byref(t), ctypes.byref(nk),
                                ctypes.byref(nx))
    return s.value, t.value, nk.value, nx.value


# ------------------------------------------------------------------- functionals
def tgrad(x, dt=1.0):
    """TVTEMP's temporal forward difference (utils.cpp:134-143)."""
    x = _c(x)
    T, H, W = x.shape
    out = np.zeros_like(x)
    lib().orc_tgrad(_ptr(x), W, H, T, float(dt), _ptr(out))
    return out


def tgrad_T(y, dt=1.0):
    """Adjoint of tgrad, i.e. minus the temporal divergence (utils.cpp:795-805)."""
    y = _c(y)
    T, H, W = y.shape
    out = np.zeros_like(y)
    lib().orc_tgrad_T(_ptr(y), W, H, T, float(dt), _ptr(out))
    return out


def tvtemp_norm(u, dt=1.0):
    u = _c(u)
    T, H, W = u.shape
    return lib().orc_tvtemp_norm(_ptr(u), W, H, T, float(dt))


def ictv_norm(u, w, alpha1, alpha, d6):
    u, w = _c(u), _c(w)
    T, H, W = u.shape
    d = np.ascontiguousarray(d6, np.float64)
    return lib().orc_ictv_norm(_ptr(u), _ptr(w), W, H, T, float(alpha1), float(alpha), _ptr(d))


def tv_norm(u, dx, dy, dt):
    u = _c(u)
    T, H, W = u.shape
    return lib().orc_tv_norm(_ptr(u), W, H, T, dx, dy, dt)


def tgv2_norm(u, v, alpha0, alpha1, dx, dy, dt):
    u, v = _c(u), _c(v)
    T, H, W = u.shape
    return lib().orc_tgv2_norm(_ptr(u), _ptr(v), W, H, T, alpha0, alpha1, dx, dy, dt)


def ictgv2_norm(u, v, w, v2, alpha0, alpha1, alpha, d6):
    u, v, w, v2 = _c(u), _c(v), _c(w), _c(v2)
    T, H, W = u.shape
    return lib().orc_ictgv2_norm(_ptr(u), _ptr(v), _ptr(w), _ptr(v2), W, H, T,
                                 alpha0, alpha1, alpha, *[float(d) for d in d6])


def gstar(method, x1, y, z, data, b1, mask, centered, lam, d6):
    x1, y, z, data, b1 = _c(x1), _c(y), _c(z), _c(data), _c(b1)
    T, C, H, W = data.shape
    m = None if mask is None else _c(mask, np.float64)
    d6 = np.ascontiguousarray(d6, np.float64)
    return lib().orc_gstar(int(method), _ptr(x1), _ptr(y), _ptr(z), _ptr(data), _ptr(b1),
                           _ptr(m), W, H, C, T, int(bool(centered)), float(lam), _ptr(d6))


def pdgap(method, x, y, z, data, b1, mask, centered, lam, alpha0, alpha1, alpha, d6):
    x, y, z, data, b1 = _c(x), _c(y), _c(z), _c(data), _c(b1)
    T, C, H, W = data.shape
    m = None if mask is None else _c(mask, np.float64)
    d6 = np.ascontiguousarray(d6, np.float64)
    return lib().orc_pdgap(int(method), _ptr(x), _ptr(y), _ptr(z), _ptr(data), _ptr(b1),
                           _ptr(m), W, H, C, T, int(bool(centered)), float(lam),
                           float(alpha0), float(alpha1), float(alpha), _ptr(d6))


# ------------------------------------------------------------------------ solver
def default_params(method):
    """The CUDA InitParams defaults (tv.cpp:39-57, tgv2.cpp:36-56,
    ictgv2.cpp:37-64); lambda is left for the caller (adaptLambda)."""
    m = METHODS[method] if isinstance(method, str) else int(method)
    p = dict(method=m, maxIt=500, sigma=1.0 / 3.0, tau=1.0 / 3.0, sigmaTauRatio=1.0,
             tsw=5.0, tsw2=1.0, lambda_=1.0, alpha0=float(np.sqrt(2.0)), alpha1=1.0,
             alpha=0.5, schedule=0, gapEvery=0, centered=0, dt=1.0, dx=1.0, dy=1.0, dz=1.0)
    if m == ICTGV2:
        p.update(tsw=6.5, tsw2=2.5, alpha=0.6423)
    if m == ICTV:  # ictv.cpp:37-62
        p.update(tsw=4.0, tsw2=0.5, alpha=0.5, alpha1=1.0)
    if m == TGV2_3D:  # tgv2_3d.cpp InitParams
        p.update(alpha0=float(np.sqrt(3.0)))
    return p


def adapt_lambda_defaults(method):
    m = METHODS[method] if isinstance(method, str) else int(method)
    if m in (ICTGV2, ICTV):
        return 0.4 / (5.1191 * 0.35), 10.0 / (5.1191 * 0.35)
    return 0.4 * 0.2991, 10.0 * 0.2991


def recon(params: dict, data, b1, mask=None, x0=None, nthreads=0, logs=False):
    """Run the PD loop.  Returns x (u) or, with logs=True, a dict with
    x, all primal components, sigma/tau logs and gap log."""
    pr = OrcParams(**params)
    data, b1 = _c(data), _c(b1)
    T, C, H, W = data.shape
    P = lib().orc_primal_components(pr.method)
    m = None if mask is None else _c(mask, np.float64)
    x = np.zeros((T, H, W), np.complex128) if x0 is None else _c(x0).copy()
    xall = np.zeros((P, T, H, W), np.complex128)
    sig = np.zeros(max(pr.maxIt, 1))
    tau = np.zeros(max(pr.maxIt, 1))
    ng_cap = pr.maxIt // max(pr.gapEvery, 1) + 2
    gaps = np.zeros(ng_cap)
    ng = ctypes.c_int(0)
    rc = lib().orc_recon(ctypes.byref(pr), W, H, C, T, _ptr(data), _ptr(b1), _ptr(m),
                         _ptr(x), _ptr(xall), _ptr(sig), _ptr(tau), _ptr(gaps),
                         ctypes.byref(ng), int(nthreads))
    if rc < 0:
        raise ValueError("invalid parameters")
    if not logs:
        return x
    return dict(x=x, primal=xall, sigma=sig[: pr.maxIt], tau=tau[: pr.maxIt],
                gap=gaps[: ng.value])


def opnorm(method, b1, mask, W, H, T, d6, iters, centered=False):
    """||K_full|| after `iters` power-iteration steps (Cartesian 2D+t)."""
    L = lib()
    L.orc_opnorm.argtypes = [_I, _P, _P, _I, _I, _I, _I, _I, _P, _I]
    L.orc_opnorm.restype = _D
    b1 = _c(b1)
    m = None if mask is None else _c(mask, np.float64)
    d = np.ascontiguousarray(d6, np.float64)
    return L.orc_opnorm(int(method), _ptr(b1), _ptr(m), W, H, b1.shape[0], T, int(bool(centered)),
                        _ptr(d), int(iters))


def opnorm3(b1, mask, W, H, D, d6, iters):
    """||K_full|| of TGV2_3D on a volume (b1 [C][D][H][W])."""
    L = lib()
    L.orc_opnorm3.argtypes = [_P, _P, _I, _I, _I, _I, _P, _I]
    L.orc_opnorm3.restype = _D
    b1 = _c(b1)
    m = None if mask is None else _c(mask, np.float64)
    d = np.ascontiguousarray(d6, np.float64)
    return L.orc_opnorm3(_ptr(b1), _ptr(m), W, H, D, b1.shape[0], _ptr(d), int(iters))


def opnorm_nc(method, b1, traj, dcf, T, d6, iters, J=3, osf=1.5):
    """||K_full|| with the non-Cartesian (gridding) data operator."""
    L = lib()
    L.orc_opnorm_nc.argtypes = [_I, _P, _I, _I, _I, _I, _I, _I, _D, _P, _P, _P, _I]
    L.orc_opnorm_nc.restype = _D
    b1 = _c(b1)
    traj = _c(traj, np.float64)
    w = None if dcf is None else _c(dcf, np.float64)
    C, H, W = b1.shape
    d = np.ascontiguousarray(d6, np.float64)
    return L.orc_opnorm_nc(int(method), _ptr(b1), W, H, C, T, traj.shape[2], int(J), float(osf),
                           _ptr(traj), _ptr(w), _ptr(d), int(iters))


# ------------------------------------------------------- coil construction
COIL_DEFAULTS = dict(uH1mu=1e-5, uReg=0.2, uNrIt=100, uSigma=np.sqrt(1.0 / 12.0),
                     uTau=np.sqrt(1.0 / 12.0), uAlpha0=np.sqrt(2.0), uAlpha1=1.0, b1Reg=2.0,
                     b1NrIt=500, b1Sigma=np.sqrt(1.0 / 8.0), b1Tau=np.sqrt(1.0 / 8.0),
                     b1FinalReg=0.1, b1FinalNrIt=1000)
_COIL_ORDER = ["uH1mu", "uReg", "uNrIt", "uSigma", "uTau", "uAlpha0", "uAlpha1", "b1Reg",
               "b1NrIt", "b1Sigma", "b1Tau", "b1FinalReg", "b1FinalNrIt"]


def _coil_lib():
    L = lib()
    if not hasattr(L, "_coil_declared"):
        L.orc_laplace_periodic.argtypes = [_P, _I, _I, _P]
        L.orc_coil_tar.argtypes = [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P]
        L.orc_coil_b1_from_uh1.argtypes = [_P, _P, _I, _I, _I, _D, _P]
        L.orc_coil_b1_from_uh1.restype = _I
        L.orc_coil_find_max_sum.argtypes = [_P, _I, _I, _I]
        L.orc_coil_find_max_sum.restype = _I
        L.orc_coil_utgv2.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P]
        L.orc_coil_b1_recon.argtypes = [_P, _P, _I, _I, _I, _I, _D, _P, _P]
        L.orc_coil_construction.argtypes = [_P, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P]
        L.orc_coil_construction_from.argtypes = [_P, _P, _I, _I, _I, _P, _P, _P, _P]
        L._coil_declared = True
    return L


def laplace_periodic(u):
    u = _c(u)
    out = np.zeros_like(u)
    _coil_lib().orc_laplace_periodic(_ptr(u), u.shape[1], u.shape[0], _ptr(out))
    return out


def coil_tar(kdata, mask=None, centered=False, apply_phase=True):
    """TimeAveragedReconstruction: kdata [T][C][H][W] -> (u [H][W], crec [C][H][W])."""
    kdata = _c(kdata)
    T, C, H, W = kdata.shape
    m = None if mask is None else _c(mask, np.float64)
    u = np.zeros((H, W), np.complex128)
    crec = np.zeros((C, H, W), np.complex128)
    _coil_lib().orc_coil_tar(_ptr(kdata), _ptr(m), W, H, C, T, int(bool(centered)),
                             int(bool(apply_phase)), _ptr(u), _ptr(crec))
    return u, crec


def coil_tar_nc(kdata, traj, dcf, shape, apply_phase=True):
    """NoncartesianCoilConstruction::TimeAveragedReconstruction
    (noncartesian_coil_construction.cpp:21-123): kdata [T][C][M]
    (sqrt(density)-weighted, as the operator takes it), traj [T][2][M],
    dcf [T][M] or None, shape (H, W).  Every frame's points of a coil are
    gridded together by the adjoint NUFFT (nufft_KH on one frame holding all
    T*M points, kernel width 3, grid oversampling 2, :26-86) into crec_c;
    u = sqrt(sum_c |crec_c|^2), with the phase of sum_c crec_c when
    apply_phase (:108-123).  Returns (u [H][W], crec [C][H][W])."""
    kdata = _c(kdata)
    traj = _c(traj, np.float64)
    T, C, M = kdata.shape
    H, W = shape
    tr = np.stack([traj[:, 0, :].ravel(), traj[:, 1, :].ravel()])[None]
    w = None if dcf is None else _c(dcf, np.float64).reshape(1, T * M)
    ones = np.ones((1, H, W), np.complex128)
    crec = np.stack([nufft_KH(kdata[:, c, :].reshape(1, 1, T * M), ones, tr, w, J=3, osf=2.0)[0]
                     for c in range(C)])
    a = np.sqrt((np.abs(crec) ** 2).sum(0))
    u = a * np.exp(1j * np.angle(crec.sum(0))) if apply_phase else a.astype(np.complex128)
    return u, crec


def coil_construction_from(crec, u0, **params):
    """CoilConstruction::PerformCoilConstruction (coil_construction.cpp:447-557)
    from time-averaged coil images crec [C][H][W] and u0 [H][W]: returns
    (u [C][H][W], b1 [C][H][W], order)."""
    crec, u0 = _c(crec), _c(u0)
    C, H, W = crec.shape
    prm = dict(COIL_DEFAULTS, **params)
    pa = np.array([prm[k] for k in _COIL_ORDER], np.float64)
    u = np.zeros((C, H, W), np.complex128)
    b1 = np.zeros((C, H, W), np.complex128)
    order = np.zeros(C, np.int32)
    _coil_lib().orc_coil_construction_from(_ptr(crec), _ptr(u0), W, H, C, _ptr(pa), _ptr(u),
                                           _ptr(b1), _ptr(order))
    return u, b1, order


def coil_b1_from_uh1(u, crec, mu=1e-5):
    u, crec = _c(u), _c(crec)
    C, H, W = crec.shape
    b1 = np.zeros_like(crec)
    _coil_lib().orc_coil_b1_from_uh1(_ptr(u), _ptr(crec), W, H, C, float(mu), _ptr(b1))
    return b1


def coil_find_max_sum(crec):
    crec = _c(crec)
    C, H, W = crec.shape
    return _coil_lib().orc_coil_find_max_sum(_ptr(crec), W, H, C)


def coil_utgv2(b1, crec, inds, x0, nu=0.2, alpha0=np.sqrt(2.0), alpha1=1.0,
               sigma=np.sqrt(1.0 / 12.0), tau=np.sqrt(1.0 / 12.0), iters=100):
    b1, crec = _c(b1), _c(crec)
    C, H, W = crec.shape
    x = _c(x0).copy()
    ind = np.ascontiguousarray(inds, np.int32)
    p = np.array([nu, alpha0, alpha1, sigma, tau, iters], np.float64)
    _coil_lib().orc_coil_utgv2(_ptr(b1), _ptr(crec), _ptr(ind), len(ind), W, H, _ptr(p), _ptr(x))
    return x


def coil_b1_recon(u0, crec, iters=500, mu=2.0, sigma=np.sqrt(1.0 / 8.0), tau=np.sqrt(1.0 / 8.0)):
    u0, crec = _c(u0), _c(crec)
    C, H, W = crec.shape
    x = np.zeros_like(crec)
    p = np.array([sigma, tau], np.float64)
    _coil_lib().orc_coil_b1_recon(_ptr(u0), _ptr(crec), C, W, H, int(iters), float(mu), _ptr(p),
                                  _ptr(x))
    return x


def coil_construction(kdata, mask=None, centered=False, **params):
    """PerformCoilConstruction: returns (u [C][H][W], b1 [C][H][W], order)."""
    kdata = _c(kdata)
    T, C, H, W = kdata.shape
    m = None if mask is None else _c(mask, np.float64)
    prm = dict(COIL_DEFAULTS, **params)
    pa = np.array([prm[k] for k in _COIL_ORDER], np.float64)
    u = np.zeros((C, H, W), np.complex128)
    b1 = np.zeros((C, H, W), np.complex128)
    order = np.zeros(C, np.int32)
    _coil_lib().orc_coil_construction(_ptr(kdata), _ptr(m), W, H, C, T, int(bool(centered)),
                                      _ptr(pa), _ptr(u), _ptr(b1), _ptr(order))
    return u, b1, order


def fft3(a, sign):
    """Unnormalised 3D DFT of a [D][H][W] array (sign -1: e^{-i})."""
    a = _c(a).copy()
    D, H, W = a.shape
    lib().orc_fft3(_ptr(a), D, H, W, int(sign))
    return a


def cart3_K(x, b1, mask=None):
    """Volumetric K: x [D][H][W], b1 [C][D][H][W] -> [C][D][H][W]."""
    x, b1 = _c(x), _c(b1)
    D, H, W = x.shape
    C = b1.shape[0]
    m = None if mask is None else _c(mask, np.float64)
    out = np.zeros((C, D, H, W), np.complex128)
    lib().orc_cart3_K(_ptr(x), _ptr(b1), _ptr(m), W, H, D, C, _ptr(out))
    return out


def cart3_KH(z, b1, mask=None):
    z, b1 = _c(z), _c(b1)
    C, D, H, W = z.shape
    m = None if mask is None else _c(mask, np.float64)
    out = np.zeros((D, H, W), np.complex128)
    lib().orc_cart3_KH(_ptr(z), _ptr(b1), _ptr(m), W, H, D, C, _ptr(out))
    return out


def recon3(params: dict, data, b1, mask=None, x0=None, nthreads=0):
    """TGV2_3D loop on a volume: data, b1 [C][D][H][W]; returns x [D][H][W]
    and the sigma/tau logs."""
    pr = OrcParams(**params)
    data, b1 = _c(data), _c(b1)
    C, D, H, W = data.shape
    m = None if mask is None else _c(mask, np.float64)
    x = np.zeros((D, H, W), np.complex128) if x0 is None else _c(x0).copy()
    xall = np.zeros((lib().orc_primal_components(TGV2), D, H, W), np.complex128)
    sig = np.zeros(max(pr.maxIt, 1))
    tau = np.zeros(max(pr.maxIt, 1))
    rc = lib().orc_recon3(ctypes.byref(pr), W, H, D, C, _ptr(data), _ptr(b1), _ptr(m), _ptr(x),
                          _ptr(xall), _ptr(sig), _ptr(tau), int(nthreads))
    if rc < 0:
        raise ValueError("invalid parameters")
    return dict(x=x, sigma=sig[: pr.maxIt], tau=tau[: pr.maxIt])


def primal_components(method):
    return lib().orc_primal_components(int(method))


def dual_components(method):
    return lib().orc_dual_components(int(method))


# -------------------------------------------------------------- non-Cartesian
def kb_beta(J, osf):
    return lib().orc_kb_beta(int(J), float(osf))


def _nc_dims(traj):
    T, two, M = traj.shape
    assert two == 2
    return T, M


def nufft_K(x, b1, traj, dcf=None, J=3, osf=1.5):
    x, b1 = _c(x), _c(b1)
    traj = _c(traj, np.float64)
    T, H, W = x.shape
    C = b1.shape[0]
    _, M = _nc_dims(traj)
    w = None if dcf is None else _c(dcf, np.float64)
    out = np.zeros((T, C, M), np.complex128)
    lib().orc_nufft_K(_ptr(x), _ptr(b1), _ptr(traj), _ptr(w), W, H, C, T, M, int(J), float(osf),
                      _ptr(out))
    return out


def nufft_KH(k, b1, traj, dcf=None, J=3, osf=1.5, shape=None):
    k, b1 = _c(k), _c(b1)
    traj = _c(traj, np.float64)
    T, C, M = k.shape
    H, W = b1.shape[1:]
    w = None if dcf is None else _c(dcf, np.float64)
    out = np.zeros((T, H, W), np.complex128)
    lib().orc_nufft_KH(_ptr(k), _ptr(b1), _ptr(traj), _ptr(w), W, H, C, T, M, int(J), float(osf),
                       _ptr(out))
    return out


def nudft_K(x, b1, traj, dcf=None):
    x, b1 = _c(x), _c(b1)
    traj = _c(traj, np.float64)
    T, H, W = x.shape
    C = b1.shape[0]
    _, M = _nc_dims(traj)
    w = None if dcf is None else _c(dcf, np.float64)
    out = np.zeros((T, C, M), np.complex128)
    lib().orc_nudft_K(_ptr(x), _ptr(b1), _ptr(traj), _ptr(w), W, H, C, T, M, _ptr(out))
    return out


def recon_nc(params: dict, data, b1, traj, dcf=None, x0=None, J=3, osf=1.5, nthreads=0,
             logs=False):
    L = lib()
    if not hasattr(L, "_nc_declared"):
        L.orc_recon_nc.argtypes = [ctypes.POINTER(OrcParams), _I, _I, _I, _I, _I, _I, _D, _P, _P,
                                   _P, _P, _P, _P, _P, _P, _I]
        L.orc_recon_nc.restype = _I
        L._nc_declared = True
    pr = OrcParams(**params)
    data, b1 = _c(data), _c(b1)
    traj = _c(traj, np.float64)
    T, C, M = data.shape
    H, W = b1.shape[1:]
    w = None if dcf is None else _c(dcf, np.float64)
    P = L.orc_primal_components(pr.method)
    x = np.zeros((T, H, W), np.complex128) if x0 is None else _c(x0).copy()
    xall = np.zeros((P, T, H, W), np.complex128)
    sig, tau = np.zeros(max(pr.maxIt, 1)), np.zeros(max(pr.maxIt, 1))
    rc = L.orc_recon_nc(ctypes.byref(pr), W, H, C, T, M, int(J), float(osf), _ptr(traj), _ptr(w),
                        _ptr(data), _ptr(b1), _ptr(x), _ptr(xall), _ptr(sig), _ptr(tau),
                        int(nthreads))
    if rc < 0:
        raise ValueError("invalid parameters")
    if not logs:
        return x
    return dict(x=x, primal=xall, sigma=sig[: pr.maxIt], tau=tau[: pr.maxIt])


def radial_trajectory(n_fe, spokes_per_frame, frames, golden=True):
    """Golden-angle radial spokes: [frames, 2, spokes*n_fe] in cycles/pixel,
    k in [-0.5, 0.5), angle increment 111.246 degrees (golden angle)."""
    ga = np.pi * (np.sqrt(5.0) - 1.0) / 2.0 if golden else np.pi / spokes_per_frame
    r = (np.arange(n_fe) - n_fe / 2.0) / n_fe
    out = np.zeros((frames, 2, spokes_per_frame * n_fe))
    for t in range(frames):
        for s in range(spokes_per_frame):
            a = (t * spokes_per_frame + s) * ga
            out[t, 0, s * n_fe:(s + 1) * n_fe] = r * np.cos(a)
            out[t, 1, s * n_fe:(s + 1) * n_fe] = r * np.sin(a)
    return out
