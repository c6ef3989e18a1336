"""MI355X-native AVIONIC primal-dual reconstruction (Python host mirror).

Python front end over ``libavionic.so`` (HIP kernels for gfx950) through its
C-ABI (``include/avionic.h``).  The classes mirror the reference's C++ API:

* ``CartesianOperator`` / ``NoncartesianOperator``  — ``BaseOperator``
  (CUDA/include/base_operator.h:24-112): ``ForwardOperation`` (k-space ->
  image, K^H), ``BackwardOperation`` (image -> k-space, K), ``AdaptLambda``.
* ``TV`` / ``TGV2`` / ``ICTGV2``  — ``PDRecon`` (CUDA/include/pd_recon.h:88-185):
  ``IterativeReconstruction``, ``GetParams``, ``AdaptLambda``, ``SetVerbose``,
  ``SetDebug``, ``AdaptStepSize``, ``ComputeGStar``, ``ComputePDGap``,
  ``ExportAdditionalResults``.
* ``gradient``, ``symmetric_gradient``, ``divergence``, ... — ``utils::``
  (CUDA/include/utils.h).

Device memory is handled with torch tensors (complex64 / float32 on
``cuda``); torch is plumbing only, every computation runs in libavionic.
There is no CPU fallback: without the built library or a HIP device the calls
raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libavionic.so")

TV_METHOD, TGV2_METHOD, ICTGV2_METHOD, TVTEMP_METHOD, ICTV_METHOD, TGV2_3D_METHOD = 0, 1, 2, 3, 4, 5

_P = ctypes.c_void_p
_I = ctypes.c_int
_U = ctypes.c_uint
_D = ctypes.c_double
_F = ctypes.c_float
_SZ = ctypes.c_size_t


class AvionicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"avionic error {code}: {msg}")
        self.code = code


class _Params(ctypes.Structure):
    _fields_ = [
        ("maxIt", ctypes.c_uint),
        ("dx", _D), ("dy", _D), ("dz", _D),
        ("stopPDGap", _D),
        ("ds", _D), ("dt", _D),
        ("sigma", _D), ("tau", _D), ("sigmaTauRatio", _D),
        ("timeSpaceWeight", _D),
        ("lambda", _D),
        ("adaptLambda", _I),
        ("adaptLambdaK", _D), ("adaptLambdaD", _D),
        ("alpha0", _D), ("alpha1", _D),
        ("alpha", _D),
        ("timeSpaceWeight2", _D),
        ("dx2", _D), ("dy2", _D), ("dt2", _D),
        ("adaptSchedule", _I),
        ("dualStorage", _I),
        ("opnormIterations", ctypes.c_uint),
    ]


def _to_dict(p: _Params) -> dict:
    return {name: getattr(p, name) for name, _ in _Params._fields_}


def _from_dict(d: dict) -> _Params:
    p = _Params()
    for name, typ in _Params._fields_:
        if name in d:
            setattr(p, name, typ(d[name]).value if typ in (_I, ctypes.c_uint) else float(d[name]))
    return p


_EXPORT_CB = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.POINTER(ctypes.c_float), _SZ, _P)

_lib = None


def lib():
    """Load libavionic.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m avionic_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    sig = {
        "avionic_last_error": (ctypes.c_char_p, []),
        "avionic_version": (ctypes.c_char_p, []),
        "avionic_bin_read": (_I, [ctypes.c_char_p, _P, _P, _SZ, ctypes.POINTER(_SZ),
                                  ctypes.POINTER(_I)]),
        "avionic_bin_write": (_I, [ctypes.c_char_p, _P, _P, _SZ]),
        "avionic_normalization_factor": (_I, [_P, _SZ, ctypes.POINTER(ctypes.c_float)]),
        "avionic_cartesian_operator_create": (_I, [_U, _U, _U, _U, _P, _I, ctypes.POINTER(_P)]),
        "avionic_cartesian3d_operator_create": (_I, [_U, _U, _U, _U, _P, ctypes.POINTER(_P)]),
        "avionic_noncartesian_operator_create": (_I, [_U, _U, _U, _U, _U, _U, _P, _P, _I, _D,
                                                      ctypes.POINTER(_P)]),
        "avionic_operator_destroy": (_I, [_P]),
        "avionic_operator_forward": (_I, [_P, _P, _P, _P, _P]),
        "avionic_operator_backward": (_I, [_P, _P, _P, _P, _P]),
        "avionic_operator_adapt_lambda": (_I, [_P, _D, _D, ctypes.POINTER(_D)]),
        "avionic_operator_kspace_size": (_SZ, [_P]),
        "avionic_default_params": (_I, [_I, ctypes.POINTER(_Params)]),
        "avionic_recon_create": (_I, [_I, _U, _U, _U, _U, ctypes.POINTER(_Params), _P,
                                      ctypes.POINTER(_P)]),
        "avionic_recon_destroy": (_I, [_P]),
        "avionic_recon_get_params": (_I, [_P, ctypes.POINTER(_Params)]),
        "avionic_recon_set_params": (_I, [_P, ctypes.POINTER(_Params)]),
        "avionic_recon_set_verbose": (_I, [_P, _I]),
        "avionic_recon_operator_norm": (_I, [_P, _P, _U, ctypes.POINTER(_D), _P]),
        "avionic_recon_set_debug": (_I, [_P, _I, _I]),
        "avionic_recon_adapt_lambda": (_I, [_P, _D, _D, ctypes.POINTER(_D)]),
        "avionic_recon_run": (_I, [_P, _P, _P, _P, _P]),
        "avionic_recon_adapt_step_size": (_I, [_P, _P, _P, _P]),
        "avionic_recon_compute_gstar": (_I, [_P, _P, _P, _P, _P, _P, ctypes.POINTER(_D), _P]),
        "avionic_recon_compute_pdgap": (_I, [_P, _P, _P, _P, _P, _P, ctypes.POINTER(_D), _P]),
        "avionic_recon_get_step_log": (_I, [_P, _P, _P, ctypes.POINTER(_SZ)]),
        "avionic_recon_export": (_I, [_P, _EXPORT_CB, _P]),
        "avionic_gradient": (_I, [_P, _P, _U, _U, _U, _F, _F, _F, _P]),
        "avionic_symmetric_gradient": (_I, [_P, _P, _U, _U, _U, _F, _F, _F, _P]),
        "avionic_divergence": (_I, [_P, _P, _U, _U, _U, _F, _F, _F, _P]),
        "avionic_symmetric_divergence": (_I, [_P, _P, _U, _U, _U, _F, _F, _F, _P]),
        "avionic_proximal_map3": (_I, [_P, _SZ, _F, _P]),
        "avionic_proximal_map6": (_I, [_P, _SZ, _F, _P]),
        "avionic_tv_norm": (_I, [_P, _U, _U, _U, _F, _F, _F, ctypes.POINTER(_D), _P]),
        "avionic_tgv2_norm": (_I, [_P, _P, _U, _U, _U, _D, _D, _F, _F, _F, ctypes.POINTER(_D), _P]),
        "avionic_ictgv2_norm": (_I, [_P, _P, _P, _P, _U, _U, _U, _D, _D, _D, _F, _F, _F, _F, _F, _F,
                                     ctypes.POINTER(_D), _P]),
        "avionic_ellipke": (_D, [_D]),
        "avionic_time_space_weights": (_I, [_D, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
        "avionic_fft2": (_I, [_P, _U, _U, _U, _I, _I, _I, _P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _check(rc):
    if rc != 0:
        raise AvionicError(rc, lib().avionic_last_error().decode())


def version() -> str:
    return lib().avionic_version().decode()


# ------------------------------------------------------------------ plumbing
def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t, dtype=None):
    """Device pointer of a contiguous cuda tensor (checked)."""
    if t is None:
        return None
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a cuda tensor")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _c64(t):
    return _dptr(t, _torch().complex64)


def _f32(t):
    return _dptr(t, _torch().float32)


def _empty_c64(*shape):
    torch = _torch()
    return torch.empty(*shape, dtype=torch.complex64, device="cuda")


# ----------------------------------------------------------------- operators
class BaseOperator:
    """BaseOperator (CUDA/include/base_operator.h:24-112)."""

    def __init__(self, handle, width, height, coils, frames):
        self._h = handle
        self.width, self.height, self.coils, self.frames = width, height, coils, frames

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().avionic_operator_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def kspace_shape(self):
        n = lib().avionic_operator_kspace_size(self._h)
        if isinstance(self, CartesianOperator):
            return (self.frames, self.coils, self.height, self.width)
        return (self.frames, self.coils, n)

    def ForwardOperation(self, kdata, b1, out=None):
        """k-space -> coil-combined image (K^H)."""
        if out is None:
            out = _empty_c64(self.frames, self.height, self.width)
        _check(lib().avionic_operator_forward(self._h, _c64(kdata), _c64(out), _c64(b1), _stream()))
        return out

    def BackwardOperation(self, x, b1, out=None):
        """image -> coil-wise k-space (K)."""
        if out is None:
            out = _empty_c64(*self.kspace_shape())
        _check(lib().avionic_operator_backward(self._h, _c64(x), _c64(out), _c64(b1), _stream()))
        return out

    def AdaptLambda(self, k, d):
        lam = ctypes.c_double()
        _check(lib().avionic_operator_adapt_lambda(self._h, float(k), float(d), ctypes.byref(lam)))
        return lam.value


class CartesianOperator(BaseOperator):
    """CartesianOperator(width, height, coils, frames[, mask][, centered])
    (CUDA/include/cartesian_operator.h:18-25).  Without a mask the operator is
    fully sampled (ZeroMask); ``centered`` defaults to True as in the
    reference's (W, H, C, T[, mask]) constructors."""

    def __init__(self, width, height, coils, frames, mask=None, centered=True):
        h = ctypes.c_void_p()
        _check(lib().avionic_cartesian_operator_create(width, height, coils, frames, _f32(mask),
                                                       int(bool(centered)), ctypes.byref(h)))
        super().__init__(h, width, height, coils, frames)
        self.centered = bool(centered)


class CartesianOperator3D(BaseOperator):
    """CartesianOperator3D(width, height, depth, coils[, mask])
    (CUDA/include/cartesian_operator3d.h): unitary 3D transforms; k-space and
    coil maps are [coils][depth][height][width], the image [depth][height][width]
    (``frames`` holds the depth)."""

    def __init__(self, width, height, depth, coils, mask=None):
        h = ctypes.c_void_p()
        _check(lib().avionic_cartesian3d_operator_create(width, height, depth, coils, _f32(mask),
                                                         ctypes.byref(h)))
        super().__init__(h, width, height, coils, depth)

    def kspace_shape(self):
        return (self.coils, self.frames, self.height, self.width)


class NoncartesianOperator(BaseOperator):
    """NoncartesianOperator (CUDA/include/noncartesian_operator.h): radial /
    arbitrary trajectories by Kaiser-Bessel gridding.  traj: float32
    [frames, 2, spokesPerFrame*nFE] in [-0.5, 0.5); dens: float32
    [frames, spokesPerFrame*nFE] or None."""

    def __init__(self, width, height, coils, frames, nFE, spokesPerFrame, traj, dens=None,
                 kernel_width=3, osf=1.5):
        h = ctypes.c_void_p()
        _check(lib().avionic_noncartesian_operator_create(width, height, coils, frames, nFE,
                                                          spokesPerFrame, _f32(traj), _f32(dens),
                                                          int(kernel_width), float(osf),
                                                          ctypes.byref(h)))
        super().__init__(h, width, height, coils, frames)
        self.nFE, self.spokesPerFrame = nFE, spokesPerFrame


def noncartesian_adapt_lambda(width, spokes_per_frame, k, d):
    """NoncartesianOperator::AdaptLambda (noncartesian_operator.cpp:88-95)
    evaluated by the library on a minimal operator of that geometry."""
    torch = _torch()
    nFE = width
    traj = torch.zeros(1, 2, spokes_per_frame * nFE, dtype=torch.float32, device="cuda")
    op = NoncartesianOperator(width, width, 1, 1, nFE, spokes_per_frame, traj)
    return op.AdaptLambda(k, d)


# ------------------------------------------------------------------- solvers
class PDRecon:
    """PDRecon (CUDA/include/pd_recon.h:88-185)."""

    METHOD = None

    def __init__(self, width, height, coils, frames, mrOp, params=None):
        self.width, self.height, self.coils, self.frames = width, height, coils, frames
        self.op = mrOp
        h = ctypes.c_void_p()
        pp = None
        if params is not None:
            base = _Params()
            _check(lib().avionic_default_params(self.METHOD, ctypes.byref(base)))
            merged = _to_dict(base)
            merged.update(params)
            pp = ctypes.byref(_from_dict(merged))
        _check(lib().avionic_recon_create(self.METHOD, width, height, coils, frames, pp,
                                          mrOp.handle, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().avionic_recon_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def N(self):
        return self.width * self.height * self.frames

    def GetParams(self) -> dict:
        p = _Params()
        _check(lib().avionic_recon_get_params(self._h, ctypes.byref(p)))
        return _to_dict(p)

    def SetParams(self, params: dict):
        cur = self.GetParams()
        cur.update(params)
        _check(lib().avionic_recon_set_params(self._h, ctypes.byref(_from_dict(cur))))

    def AdaptLambda(self, k, d):
        lam = ctypes.c_double()
        _check(lib().avionic_recon_adapt_lambda(self._h, float(k), float(d), ctypes.byref(lam)))
        return lam.value

    def SetVerbose(self, verbose):
        _check(lib().avionic_recon_set_verbose(self._h, int(bool(verbose))))

    def SetDebug(self, debug, debugstep=1):
        _check(lib().avionic_recon_set_debug(self._h, int(bool(debug)), int(debugstep)))

    def IterativeReconstruction(self, data, x, b1):
        """x is updated in place (initial image in, reconstruction out)."""
        _check(lib().avionic_recon_run(self._h, _c64(data), _c64(x), _c64(b1), _stream()))
        return x

    def OperatorNorm(self, b1, iterations=50):
        """||K_full|| of this solver's stacked regulariser + data operator by
        an FP64 power iteration on the GPU (matlab/getopnorm.m); the fixed-step
        mode adaptSchedule = 2 sets sigma = tau = 0.99 / ||K_full||."""
        n = ctypes.c_double()
        _check(lib().avionic_recon_operator_norm(self._h, _c64(b1), int(iterations),
                                                 ctypes.byref(n), _stream()))
        return n.value

    def AdaptStepSize(self, delta, b1):
        _check(lib().avionic_recon_adapt_step_size(self._h, _c64(delta), _c64(b1), _stream()))

    def ComputeGStar(self, x1, y, z, data, b1):
        out = ctypes.c_double()
        _check(lib().avionic_recon_compute_gstar(self._h, _c64(x1), _c64(y), _c64(z), _c64(data),
                                                 _c64(b1), ctypes.byref(out), _stream()))
        return out.value

    def ComputePDGap(self, x, y, z, data, b1):
        out = ctypes.c_double()
        _check(lib().avionic_recon_compute_pdgap(self._h, _c64(x), _c64(y), _c64(z), _c64(data),
                                                 _c64(b1), ctypes.byref(out), _stream()))
        return out.value

    def _step_log(self):
        n = ctypes.c_size_t(1 << 30)
        _check(lib().avionic_recon_get_step_log(self._h, None, None, ctypes.byref(n)))
        s = np.zeros(max(n.value, 1))
        t = np.zeros(max(n.value, 1))
        m = ctypes.c_size_t(n.value)
        _check(lib().avionic_recon_get_step_log(self._h, s.ctypes.data_as(_P), t.ctypes.data_as(_P),
                                                ctypes.byref(m)))
        return s[: m.value], t[: m.value]

    def sigma_log(self):
        return self._step_log()[0]

    def tau_log(self):
        return self._step_log()[1]

    def ExportAdditionalResults(self, outputDir, callback):
        """callback(outputDir, name, numpy complex64 array)."""
        def _cb(name, data, count, _user):
            arr = np.ctypeslib.as_array(data, shape=(2 * count,)).copy()
            callback(outputDir, name.decode(), arr.view(np.complex64))
        cb = _EXPORT_CB(_cb)
        _check(lib().avionic_recon_export(self._h, cb, None))


class TV(PDRecon):
    """TV (CUDA/include/tv.h)."""
    METHOD = TV_METHOD


class TGV2(PDRecon):
    """TGV2 (CUDA/include/tgv2.h)."""
    METHOD = TGV2_METHOD


class ICTGV2(PDRecon):
    """ICTGV2 (CUDA/include/ictgv2.h)."""
    METHOD = ICTGV2_METHOD


class TVTEMP(PDRecon):
    """TVTEMP, temporal TV (CUDA/include/tv_temp.h): parameter dt is used as
    given; dx, dy and timeSpaceWeight do not apply."""
    METHOD = TVTEMP_METHOD


class ICTV(PDRecon):
    """ICTV, infimal convolution of two TV terms (CUDA/include/ictv.h)."""
    METHOD = ICTV_METHOD


class TGV2_3D(PDRecon):
    """TGV2_3D (CUDA/include/tgv2_3d.h): TGV2 on a volume with a
    CartesianOperator3D; ``frames`` is the depth, dx, dy, dz apply."""
    METHOD = TGV2_3D_METHOD


def default_params(method) -> dict:
    p = _Params()
    _check(lib().avionic_default_params(int(method), ctypes.byref(p)))
    return _to_dict(p)


# ------------------------------------------------------------------- utils::
def _dims(t, lead):
    s = tuple(t.shape)
    T, H, W = s[lead:]
    return W, H, T


def gradient(u, dx=1.0, dy=1.0, dt=1.0):
    W, H, T = _dims(u, 0)
    out = _empty_c64(3, T, H, W)
    _check(lib().avionic_gradient(_c64(u), _c64(out), W, H, T, dx, dy, dt, _stream()))
    return out


def symmetric_gradient(v, dx=1.0, dy=1.0, dt=1.0):
    W, H, T = _dims(v, 1)
    out = _empty_c64(6, T, H, W)
    _check(lib().avionic_symmetric_gradient(_c64(v), _c64(out), W, H, T, dx, dy, dt, _stream()))
    return out


def divergence(p, dx=1.0, dy=1.0, dt=1.0):
    W, H, T = _dims(p, 1)
    out = _empty_c64(T, H, W)
    _check(lib().avionic_divergence(_c64(p), _c64(out), W, H, T, dx, dy, dt, _stream()))
    return out


def symmetric_divergence(q, dx=1.0, dy=1.0, dt=1.0):
    W, H, T = _dims(q, 1)
    out = _empty_c64(3, T, H, W)
    _check(lib().avionic_symmetric_divergence(_c64(q), _c64(out), W, H, T, dx, dy, dt, _stream()))
    return out


def proximal_map3(y, scale):
    """In place: y /= max(1, |y|_3 * scale) (utils::ProximalMap3)."""
    n = y[0].numel()
    _check(lib().avionic_proximal_map3(_c64(y), n, float(scale), _stream()))
    return y


def proximal_map6(q, scale):
    n = q[0].numel()
    _check(lib().avionic_proximal_map6(_c64(q), n, float(scale), _stream()))
    return q


def tv_norm(u, dx, dy, dt):
    W, H, T = _dims(u, 0)
    out = ctypes.c_double()
    _check(lib().avionic_tv_norm(_c64(u), W, H, T, dx, dy, dt, ctypes.byref(out), _stream()))
    return out.value


def tgv2_norm(u, v, alpha0, alpha1, dx, dy, dt):
    W, H, T = _dims(u, 0)
    out = ctypes.c_double()
    _check(lib().avionic_tgv2_norm(_c64(u), _c64(v), W, H, T, float(alpha0), float(alpha1), dx, dy,
                                   dt, ctypes.byref(out), _stream()))
    return out.value


def ictgv2_norm(u, v, w, v2, alpha0, alpha1, alpha, dx, dy, dt, dx2, dy2, dt2):
    W, H, T = _dims(u, 0)
    out = ctypes.c_double()
    _check(lib().avionic_ictgv2_norm(_c64(u), _c64(v), _c64(w), _c64(v2), W, H, T, float(alpha0),
                                     float(alpha1), float(alpha), dx, dy, dt, dx2, dy2, dt2,
                                     ctypes.byref(out), _stream()))
    return out.value


def read_bin(path):
    """Read a reference .bin vector file (uint8 iscomplex, uint32 n, n float64
    real parts, [n float64 imaginary parts]; matlab/readbin_vector.m:30-40)
    into a float64 / complex128 numpy array."""
    L = lib()
    n, cplx = _SZ(), ctypes.c_int()
    _check(L.avionic_bin_read(os.fsencode(path), None, None, 0, ctypes.byref(n), ctypes.byref(cplx)))
    re_ = np.zeros(n.value, np.float64)
    im_ = np.zeros(n.value, np.float64)
    _check(L.avionic_bin_read(os.fsencode(path), re_.ctypes.data_as(_P), im_.ctypes.data_as(_P),
                              n.value, ctypes.byref(n), ctypes.byref(cplx)))
    return re_ + 1j * im_ if cplx.value else re_


def write_bin(path, v):
    """Write a vector as a reference .bin file (complex if v is complex;
    matlab/writebin_vector.m:35-44)."""
    v = np.asarray(v).ravel()
    re_ = np.ascontiguousarray(v.real, np.float64)
    im_ = np.ascontiguousarray(v.imag, np.float64) if np.iscomplexobj(v) else None
    _check(lib().avionic_bin_write(os.fsencode(path), re_.ctypes.data_as(_P),
                                   None if im_ is None else im_.ctypes.data_as(_P), re_.size))


def normalization_factor(magnitude):
    """RawDataPreparation::FindNormalizationFactor (raw_data_preparation.cpp
    :661-686): median of the float32 magnitudes >= 0.9 * max, counted from the
    first maximum onwards."""
    v = np.ascontiguousarray(np.asarray(magnitude, np.float32).ravel())
    out = ctypes.c_float()
    _check(lib().avionic_normalization_factor(v.ctypes.data_as(_P), v.size, ctypes.byref(out)))
    return out.value


def ellipke(m):
    return lib().avionic_ellipke(float(m))


def time_space_weights(tsw):
    ds, dt = ctypes.c_double(), ctypes.c_double()
    _check(lib().avionic_time_space_weights(float(tsw), ctypes.byref(ds), ctypes.byref(dt)))
    return ds.value, dt.value


def fft2(a, sign=-1, centered=False, unitary=False):
    """In-place 2D FFT over the last two axes of a complex64 cuda tensor."""
    H, W = a.shape[-2], a.shape[-1]
    batch = int(a.numel() // (H * W))
    _check(lib().avionic_fft2(_c64(a), H, W, batch, int(sign), int(bool(centered)),
                              int(bool(unitary)), _stream()))
    return a


# ------------------------------------------------------- batching / serving
class _EngineConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", _SZ), ("use_graphs", _I), ("max_batch", _I),
                ("iterations_per_tick", _I), ("resident_bytes_cap", _SZ), ("max_queue", _I),
                ("voxels_per_lane", _I), ("fixed_launch", _I)]


class _EngineCounts(ctypes.Structure):
    _fields_ = [("held", _I), ("running", _I), ("queued", _I), ("bytes_held", _SZ),
                ("bytes_running", _SZ), ("hbm_budget_bytes", _SZ),
                ("graph_captures", ctypes.c_longlong), ("capture_ms", _D),
                ("launch_tunings", ctypes.c_longlong)]


class _ReqDesc(ctypes.Structure):
    _fields_ = [("method", _I), ("width", _U), ("height", _U), ("coils", _U), ("frames", _U),
                ("params", ctypes.POINTER(_Params)), ("mask", _P), ("centered", _I),
                ("data", _P), ("b1", _P), ("x0", _P), ("stream", _P), ("verbose", _I),
                ("debug", _I), ("debugstep", _I), ("op", _P)]


_ENGINE_SIGS = {
    "avionic_engine_create": (_I, [ctypes.POINTER(_EngineConfig), ctypes.POINTER(_P)]),
    "avionic_engine_destroy": (_I, [_P]),
    "avionic_request_footprint": (_SZ, [ctypes.POINTER(_ReqDesc)]),
    "avionic_engine_submit": (_I, [_P, ctypes.POINTER(_ReqDesc), ctypes.POINTER(ctypes.c_longlong)]),
    "avionic_engine_step": (_I, [_P, _P]),
    "avionic_engine_poll": (_I, [_P, ctypes.c_longlong, _P, _P]),
    "avionic_engine_cancel": (_I, [_P, ctypes.c_longlong]),
    "avionic_engine_stats": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_SZ)]),
    "avionic_engine_counters": (_I, [_P, ctypes.POINTER(_EngineCounts)]),
    "avionic_engine_get_step_log": (_I, [_P, ctypes.c_longlong, _P, _P, ctypes.POINTER(_SZ)]),
    "avionic_engine_get_gap_log": (_I, [_P, ctypes.c_longlong, _P, _P, _P, ctypes.POINTER(_SZ)]),
}

AVIONIC_E_CAPACITY = -5
AVIONIC_E_NOTFOUND = -6


def _engine_lib():
    L = lib()
    if not getattr(L, "_engine_declared", False):
        for name, (res, args) in _ENGINE_SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._engine_declared = True
    return L


class Request:
    """Inputs of one reconstruction request (device tensors).  Without op the
    engine builds a Cartesian operator from mask / centered (data [T, C, H, W]);
    with op (a CartesianOperator or NoncartesianOperator that outlives the
    request) data is that operator's k-space, e.g. [T, C, M] for radial data."""

    def __init__(self, method, data, b1, mask=None, x0=None, centered=False, params=None,
                 verbose=False, debug=False, debugstep=1, op=None):
        self.method = int(method)
        self.verbose, self.debug, self.debugstep = bool(verbose), bool(debug), int(debugstep)
        self.op = op
        if op is not None:
            C, H, W = b1.shape
            T = data.shape[0]
        else:
            T, C, H, W = data.shape
        self.shape = (W, H, C, T)
        self.data, self.b1, self.mask, self.x0 = data, b1, mask, x0
        self.centered = bool(centered)
        self._params = None
        if params is not None:
            base = default_params(self.method)
            base.update(params)
            self._params = _from_dict(base)

    def desc(self):
        W, H, C, T = self.shape
        d = _ReqDesc()
        d.method = self.method
        d.width, d.height, d.coils, d.frames = W, H, C, T
        d.params = ctypes.pointer(self._params) if self._params is not None else None
        d.mask = _f32(self.mask)
        d.centered = int(self.centered)
        d.data = _c64(self.data)
        d.b1 = _c64(self.b1)
        d.x0 = _c64(self.x0)
        d.stream = _stream()
        d.verbose, d.debug, d.debugstep = int(self.verbose), int(self.debug), self.debugstep
        d.op = self.op.handle if self.op is not None else None
        return d


class Engine:
    """Continuous-batching executor (avionic_engine_*): requests are held
    against an HBM budget (default 90% of the free device memory), admitted in
    FIFO order while the running requests stay within max_batch and
    resident_bytes_cap, and advance together in task-list launches
    (hipGraph-replayed when use_graphs).  max_queue: waiting requests beyond
    which submit returns None (default 0: no queue; < 0: unbounded).
    voxels_per_lane: 1 runs the dual / primal updates with one voxel per lane;
    0 (default) with a pair of voxels per lane for requests of even width.
    fixed_launch: keep the default launch variant; by default a launch set
    that has run 8 plain iterations unchanged times its variants (one voxel
    or a pair per lane, column pass before or after the dual update; all give
    the same bits) and keeps the fastest."""

    def __init__(self, hbm_budget_bytes=0, use_graphs=True, max_batch=0, iterations_per_tick=1,
                 resident_bytes_cap=0, max_queue=0, voxels_per_lane=0, fixed_launch=False):
        L = _engine_lib()
        cfg = _EngineConfig(int(hbm_budget_bytes), int(bool(use_graphs)), int(max_batch),
                            int(iterations_per_tick), int(resident_bytes_cap), int(max_queue),
                            int(voxels_per_lane), int(bool(fixed_launch)))
        h = ctypes.c_void_p()
        _check(L.avionic_engine_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _engine_lib().avionic_engine_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @staticmethod
    def footprint(req: Request) -> int:
        d = req.desc()
        return int(_engine_lib().avionic_request_footprint(ctypes.byref(d)))

    def submit(self, req: Request):
        """Returns the request id, or None when it does not fit right now."""
        d = req.desc()
        rid = ctypes.c_longlong()
        rc = _engine_lib().avionic_engine_submit(self._h, ctypes.byref(d), ctypes.byref(rid))
        if rc == AVIONIC_E_CAPACITY:
            return None
        _check(rc)
        return rid.value

    def step(self):
        rc = _engine_lib().avionic_engine_step(self._h, _stream())
        if rc < 0:
            _check(rc)
        return rc

    def poll(self, rid, out):
        rc = _engine_lib().avionic_engine_poll(self._h, int(rid), _c64(out), _stream())
        if rc < 0:
            _check(rc)
        return rc == 1

    def cancel(self, rid):
        _check(_engine_lib().avionic_engine_cancel(self._h, int(rid)))

    def stats(self):
        r, b = ctypes.c_int(), ctypes.c_size_t()
        _check(_engine_lib().avionic_engine_stats(self._h, ctypes.byref(r), ctypes.byref(b)))
        return r.value, b.value

    def counters(self) -> dict:
        """held / running / queued requests, their bytes, the HBM budget, and
        the iteration graphs captured so far with the host time they took."""
        c = _EngineCounts()
        _check(_engine_lib().avionic_engine_counters(self._h, ctypes.byref(c)))
        return {name: getattr(c, name) for name, _ in _EngineCounts._fields_}

    def step_log(self, rid):
        n = ctypes.c_size_t(1 << 30)
        _check(_engine_lib().avionic_engine_get_step_log(self._h, int(rid), None, None,
                                                         ctypes.byref(n)))
        s, t = np.zeros(max(1, n.value)), np.zeros(max(1, n.value))
        m = ctypes.c_size_t(n.value)
        _check(_engine_lib().avionic_engine_get_step_log(self._h, int(rid), s.ctypes.data_as(_P),
                                                         t.ctypes.data_as(_P), ctypes.byref(m)))
        return s[: m.value], t[: m.value]

    def gap_log(self, rid):
        """(gap, regulariser norm, data fidelity) arrays of the in-loop PD gap
        evaluations so far (SetVerbose / SetDebug / stopPDGap)."""
        L = _engine_lib()
        n = ctypes.c_size_t(1 << 30)
        _check(L.avionic_engine_get_gap_log(self._h, int(rid), None, None, None, ctypes.byref(n)))
        g, r, f = (np.zeros(max(1, n.value)) for _ in range(3))
        m = ctypes.c_size_t(n.value)
        _check(L.avionic_engine_get_gap_log(self._h, int(rid), g.ctypes.data_as(_P),
                                            r.ctypes.data_as(_P), f.ctypes.data_as(_P),
                                            ctypes.byref(m)))
        return g[: m.value], r[: m.value], f[: m.value]


def engine_time_kernels(engine: Engine, reps=20):
    """{kernel name: average ms per launch} (see avionic_engine_time_kernels)."""
    L = _engine_lib()
    if not hasattr(L, "_tk"):
        L.avionic_engine_time_kernels.restype = _I
        L.avionic_engine_time_kernels.argtypes = [_P, _I, _P, _P, ctypes.POINTER(_SZ)]
        L._tk = True
    n = ctypes.c_size_t(16)
    ms = (ctypes.c_double * 16)()
    names = (ctypes.c_char_p * 16)()
    _check(L.avionic_engine_time_kernels(engine._h, int(reps), ms, names, ctypes.byref(n)))
    return {names[i].decode(): ms[i] for i in range(n.value)}


# --------------------------------------------------------- coil construction
class _CoilParams(ctypes.Structure):
    _fields_ = [("uH1mu", _D), ("uReg", _D), ("uNrIt", _U), ("uSigma", _D), ("uTau", _D),
                ("uAlpha0", _D), ("uAlpha1", _D), ("b1Reg", _D), ("b1NrIt", _U),
                ("b1Sigma", _D), ("b1Tau", _D), ("b1FinalReg", _D), ("b1FinalNrIt", _U)]


_COIL_SIGS = {
    "avionic_default_coil_params": (_I, [ctypes.POINTER(_CoilParams)]),
    "avionic_coil_time_averaged": (_I, [_P, _P, _P, _P, _I, _P]),
    "avionic_coil_b1_from_uh1": (_I, [_U, _U, _U, _P, _P, _D, _P, _P]),
    "avionic_coil_find_max_sum": (_I, [_U, _U, _U, _P, ctypes.POINTER(_I), _P]),
    "avionic_coil_utgv2": (_I, [_U, _U, _U, _P, _P, ctypes.POINTER(_I), _I,
                                ctypes.POINTER(_CoilParams), _P, _P]),
    "avionic_coil_b1_recon": (_I, [_U, _U, _U, _P, _P, _U, _D, ctypes.POINTER(_CoilParams), _P,
                                   _P]),
    "avionic_coil_construction": (_I, [_P, _P, ctypes.POINTER(_CoilParams), _P, _P,
                                       ctypes.POINTER(_I), _P]),
}


def _coil_lib():
    L = lib()
    if not getattr(L, "_coil_declared", False):
        for name, (res, args) in _COIL_SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._coil_declared = True
    return L


def default_coil_params() -> dict:
    """CoilConstruction::InitParams (coil_construction.cpp:22-46)."""
    p = _CoilParams()
    _check(_coil_lib().avionic_default_coil_params(ctypes.byref(p)))
    return {name: getattr(p, name) for name, _ in _CoilParams._fields_}


class CoilConstruction:
    """CoilConstruction / CartesianCoilConstruction (CUDA/include/
    coil_construction.h:58-141, cartesian_coil_construction.h): coil
    sensitivities and an image estimate from k-space data.  Device tensors:
    k-space [T, C, H, W], images [C, H, W] (one image [H, W])."""

    def __init__(self, width, height, coils, frames, mrOp=None, params=None):
        self.W, self.H, self.C, self.T = int(width), int(height), int(coils), int(frames)
        self.op = mrOp
        base = default_coil_params()
        if params:
            base.update(params)
        self._p = _CoilParams(**{k: int(v) if k.endswith("NrIt") else float(v)
                                 for k, v in base.items()})

    def GetParams(self):
        return {name: getattr(self._p, name) for name, _ in _CoilParams._fields_}

    def _img(self, n):
        import torch
        return torch.zeros((n, self.H, self.W) if n > 1 else (self.H, self.W),
                           dtype=torch.complex64, device="cuda")

    def TimeAveragedReconstruction(self, kdata, applyPhase=True):
        """(u [H, W], crec [C, H, W])."""
        u, crec = self._img(1), self._img(self.C)
        _check(_coil_lib().avionic_coil_time_averaged(self.op._h, _c64(kdata), _c64(u),
                                                      _c64(crec), int(bool(applyPhase)),
                                                      _stream()))
        return u, crec

    def B1FromUH1(self, u, crec):
        b1 = self._img(crec.shape[0] if crec.dim() == 3 else 1)
        C = crec.shape[0] if crec.dim() == 3 else 1
        _check(_coil_lib().avionic_coil_b1_from_uh1(self.W, self.H, C, _c64(u), _c64(crec),
                                                    float(self._p.uH1mu), _c64(b1), _stream()))
        return b1

    def FindMaximumSum(self, crec):
        i = ctypes.c_int()
        C = crec.shape[0] if crec.dim() == 3 else 1
        _check(_coil_lib().avionic_coil_find_max_sum(self.W, self.H, C, _c64(crec),
                                                     ctypes.byref(i), _stream()))
        return i.value

    def UTGV2Recon(self, b1, crec, x, inds):
        """x [H, W] is updated in place."""
        C = crec.shape[0] if crec.dim() == 3 else 1
        arr = (ctypes.c_int * len(inds))(*[int(i) for i in inds])
        _check(_coil_lib().avionic_coil_utgv2(self.W, self.H, C, _c64(b1), _c64(crec), arr,
                                              len(inds), ctypes.byref(self._p), _c64(x),
                                              _stream()))
        return x

    def B1Recon(self, u0, crec, coils, maxIt=500, mu=2.0):
        b1 = self._img(int(coils))
        _check(_coil_lib().avionic_coil_b1_recon(self.W, self.H, int(coils), _c64(u0), _c64(crec),
                                                 int(maxIt), float(mu), ctypes.byref(self._p),
                                                 _c64(b1), _stream()))
        return b1

    def PerformCoilConstruction(self, kdata):
        """(u [C, H, W], b1 [C, H, W], coil order)."""
        u, b1 = self._img(self.C), self._img(self.C)
        order = (ctypes.c_int * self.C)()
        _check(_coil_lib().avionic_coil_construction(self.op._h, _c64(kdata), ctypes.byref(self._p),
                                                     _c64(u), _c64(b1), order, _stream()))
        return u, b1, list(order)
