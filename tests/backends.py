"""Two implementations of the same numpy-in/numpy-out interface.

``OracleBackend`` wraps the FP64 CPU oracle (``oracle/``); ``GpuBackend``
drives the HIP library through its C-ABI (``avionic_amd``).  The known-answer
and parity tests run the same cases on both.
"""
import math

import numpy as np

import oracle as O


def _d3(d):
    return [float(d[0]), float(d[1]), float(d[2])]


class OracleBackend:
    name = "oracle"

    def grad(self, u, d):
        return O.grad(u, *_d3(d))

    def symgrad(self, v, d):
        return O.symgrad(v, *_d3(d))

    def div(self, p, d):
        return O.div(p, *_d3(d))

    def symdiv(self, q, d):
        return O.symdiv(q, *_d3(d))

    def prox3(self, y, radius):
        return O.prox3(y, radius)

    def prox6(self, q, radius):
        return O.prox6(q, radius)

    def cart_K(self, x, b1, mask, centered):
        return O.cart_K(x, b1, mask, centered)

    def cart_KH(self, z, b1, mask, centered):
        return O.cart_KH(z, b1, mask, centered)

    def fft2(self, a, sign, centered):
        return O.fft2(a, sign, centered)

    def solver_adapt_step(self, method, delta, b1, mask, centered, d6, ratio, sigma, tau):
        s, t, _, _ = O.solver_adapt_step(method, delta, b1, mask, centered, d6, ratio, sigma, tau)
        return s, t

    def tv_norm(self, u, d):
        return O.tv_norm(u, *_d3(d))

    def tgv2_norm(self, u, v, a0, a1, d):
        return O.tgv2_norm(u, v, a0, a1, *_d3(d))

    def ictgv2_norm(self, u, v, w, v2, a0, a1, alpha, d6):
        return O.ictgv2_norm(u, v, w, v2, a0, a1, alpha, d6)

    def gstar(self, method, x1, y, z, data, b1, mask, centered, lam, d6):
        return O.gstar(method, x1, y, z, data, b1, mask, centered, lam, d6)

    def pdgap(self, method, x, y, z, data, b1, mask, centered, lam, a0, a1, alpha, d6):
        return O.pdgap(method, x, y, z, data, b1, mask, centered, lam, a0, a1, alpha, d6)

    def default_params(self, method):
        return O.default_params(method)

    def adapt_lambda_defaults(self, method):
        return O.adapt_lambda_defaults(method)

    def adapt_lambda_cart(self, mask, W, H, T, k, d):
        return O.adapt_lambda_cart(mask, W, H, T, k, d)

    def adapt_lambda_noncart(self, W, spf, k, d):
        # noncartesian_operator.cpp:88-95: lambda = k * W / (spf * pi / 2) + d
        return k * W / (spf * math.pi / 2.0) + d

    def ellipke(self, m):
        return O.ellipke(m)

    def recon(self, params, data, b1, mask, x0):
        return O.recon(params, data, b1, mask, x0)

    def recon_logs(self, params, data, b1, mask, x0):
        return O.recon(params, data, b1, mask, x0, logs=True)

    # -- coil construction (coil_construction.cpp)
    def laplace_periodic(self, u):
        return O.laplace_periodic(u)

    def coil_tar(self, kdata, mask=None, centered=False):
        return O.coil_tar(kdata, mask, centered)

    def coil_b1_from_uh1(self, u, crec):
        return O.coil_b1_from_uh1(u, crec)

    def coil_find_max_sum(self, crec):
        return O.coil_find_max_sum(crec)

    def coil_utgv2(self, b1, crec, inds, x0):
        return O.coil_utgv2(b1, crec, inds, x0)

    def coil_b1_recon(self, u0, crec, iters=500, mu=2.0):
        return O.coil_b1_recon(u0, crec, iters, mu)

    def coil_construction(self, kdata, mask=None, centered=False):
        return O.coil_construction(kdata, mask, centered)


class GpuBackend:
    """HIP product path through the C-ABI.  Inputs are converted to
    complex64 on the device; outputs come back as complex128 numpy."""
    name = "gpu"

    def __init__(self):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("GpuBackend needs a HIP device")
        import avionic_amd as A
        self.A = A
        self.torch = torch
        self.dev = torch.device("cuda:0")
        A.lib()  # load libavionic.so or fail loudly

    # -- helpers
    def _t(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, np.complex64)).to(self.dev)

    def _r(self, a):
        if a is None:
            return None
        return self.torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(self.dev)

    def _n(self, t):
        self.torch.cuda.synchronize()
        return t.cpu().numpy().astype(np.complex128)

    def grad(self, u, d):
        return self._n(self.A.gradient(self._t(u), *_d3(d)))

    def symgrad(self, v, d):
        return self._n(self.A.symmetric_gradient(self._t(v), *_d3(d)))

    def div(self, p, d):
        return self._n(self.A.divergence(self._t(p), *_d3(d)))

    def symdiv(self, q, d):
        return self._n(self.A.symmetric_divergence(self._t(q), *_d3(d)))

    def prox3(self, y, radius):
        t = self._t(y)
        self.A.proximal_map3(t, 1.0 / radius)
        return self._n(t)

    def prox6(self, q, radius):
        t = self._t(q)
        self.A.proximal_map6(t, 1.0 / radius)
        return self._n(t)

    def cart_K(self, x, b1, mask, centered):
        T, H, W = x.shape
        C = b1.shape[0]
        op = self.A.CartesianOperator(W, H, C, T, self._r(mask), bool(centered))
        return self._n(op.BackwardOperation(self._t(x), self._t(b1)))

    def cart_KH(self, z, b1, mask, centered):
        T, C, H, W = z.shape
        op = self.A.CartesianOperator(W, H, C, T, self._r(mask), bool(centered))
        return self._n(op.ForwardOperation(self._t(z), self._t(b1)))

    def fft2(self, a, sign, centered):
        t = self._t(a)
        self.A.fft2(t, sign=sign, centered=centered, unitary=False)
        return self._n(t)

    def _solver(self, method, dims, C, mask, centered, params=None):
        T, H, W = dims
        op = self.A.CartesianOperator(W, H, C, T, self._r(mask), bool(centered))
        cls = {0: self.A.TV, 1: self.A.TGV2, 2: self.A.ICTGV2, 3: self.A.TVTEMP,
               4: self.A.ICTV}[method]
        return cls(W, H, C, T, op, params)

    def solver_adapt_step(self, method, delta, b1, mask, centered, d6, ratio, sigma, tau):
        P, T, H, W = delta.shape
        s = self._solver(method, (T, H, W), b1.shape[0], mask, centered)
        p = s.GetParams()
        p.update(sigma=sigma, tau=tau, sigmaTauRatio=ratio, dx=d6[0], dy=d6[1], dt=d6[2],
                 dx2=d6[3], dy2=d6[4], dt2=d6[5])
        s.SetParams(p)
        s.AdaptStepSize(self._t(delta), self._t(b1))
        p = s.GetParams()
        return p["sigma"], p["tau"]

    def tv_norm(self, u, d):
        return self.A.tv_norm(self._t(u), *_d3(d))

    def tgv2_norm(self, u, v, a0, a1, d):
        return self.A.tgv2_norm(self._t(u), self._t(v), a0, a1, *_d3(d))

    def ictgv2_norm(self, u, v, w, v2, a0, a1, alpha, d6):
        return self.A.ictgv2_norm(self._t(u), self._t(v), self._t(w), self._t(v2), a0, a1,
                                  alpha, *[float(x) for x in d6])

    def _gap_solver(self, method, data, b1, mask, centered, lam, d6, a0=None, a1=None,
                    alpha=None):
        T, C, H, W = data.shape
        s = self._solver(method, (T, H, W), C, mask, centered)
        p = s.GetParams()
        p.update(**{"lambda": lam, "dx": d6[0], "dy": d6[1], "dt": d6[2], "dx2": d6[3],
                    "dy2": d6[4], "dt2": d6[5]})
        if a0 is not None:
            p.update(alpha0=a0, alpha1=a1, alpha=alpha)
        s.SetParams(p)
        return s

    def gstar(self, method, x1, y, z, data, b1, mask, centered, lam, d6):
        s = self._gap_solver(method, data, b1, mask, centered, lam, d6)
        return s.ComputeGStar(self._t(x1), self._t(y), self._t(z), self._t(data), self._t(b1))

    def pdgap(self, method, x, y, z, data, b1, mask, centered, lam, a0, a1, alpha, d6):
        s = self._gap_solver(method, data, b1, mask, centered, lam, d6, a0, a1, alpha)
        return s.ComputePDGap(self._t(x), self._t(y), self._t(z), self._t(data), self._t(b1))

    def default_params(self, method):
        return O.default_params(method)

    def adapt_lambda_defaults(self, method):
        return O.adapt_lambda_defaults(method)

    def adapt_lambda_cart(self, mask, W, H, T, k, d):
        op = self.A.CartesianOperator(W, H, 1, T, self._r(mask), False)
        return op.AdaptLambda(k, d)

    def adapt_lambda_noncart(self, W, spf, k, d):
        return self.A.noncartesian_adapt_lambda(W, spf, k, d)

    def ellipke(self, m):
        return self.A.ellipke(m)

    def recon(self, params, data, b1, mask, x0):
        return self.recon_logs(params, data, b1, mask, x0)["x"]

    def recon_logs(self, params, data, b1, mask, x0):
        T, C, H, W = data.shape
        method = int(params["method"])
        s = self._solver(method, (T, H, W), C, mask, bool(params.get("centered", 0)))
        p = s.GetParams()
        p.update(maxIt=int(params["maxIt"]), sigma=params["sigma"], tau=params["tau"],
                 sigmaTauRatio=params["sigmaTauRatio"], timeSpaceWeight=params["tsw"],
                 timeSpaceWeight2=params["tsw2"], alpha0=params["alpha0"],
                 alpha1=params["alpha1"], alpha=params["alpha"],
                 adaptSchedule=int(params.get("schedule", 0)))
        if method == 3:
            p["dt"] = params["dt"]
        p["lambda"] = params["lambda_"]
        if "dualStorage" in params:
            p["dualStorage"] = int(params["dualStorage"])
        s.SetParams(p)
        x = self._t(np.zeros((T, H, W)) if x0 is None else x0)
        s.IterativeReconstruction(self._t(data), x, self._t(b1))
        return dict(x=self._n(x), sigma=np.array(s.sigma_log()), tau=np.array(s.tau_log()),
                    primal=None)

    # -- coil construction (coil_construction.cpp) through the C-ABI
    def _coil(self, W, H, C, T, mask=None, centered=False):
        op = self.A.CartesianOperator(W, H, C, T, self._r(mask), bool(centered))
        return self.A.CoilConstruction(W, H, C, T, op), op

    def coil_tar(self, kdata, mask=None, centered=False):
        T, C, H, W = kdata.shape
        cc, _op = self._coil(W, H, C, T, mask, centered)
        u, crec = cc.TimeAveragedReconstruction(self._t(kdata))
        return self._n(u), self._n(crec)

    def coil_b1_from_uh1(self, u, crec):
        C, H, W = crec.shape
        cc = self.A.CoilConstruction(W, H, C, 1)
        return self._n(cc.B1FromUH1(self._t(u), self._t(crec)))

    def coil_find_max_sum(self, crec):
        C, H, W = crec.shape
        return self.A.CoilConstruction(W, H, C, 1).FindMaximumSum(self._t(crec))

    def coil_utgv2(self, b1, crec, inds, x0):
        C, H, W = crec.shape
        x = self._t(x0)
        self.A.CoilConstruction(W, H, C, 1).UTGV2Recon(self._t(b1), self._t(crec), x, inds)
        return self._n(x)

    def coil_b1_recon(self, u0, crec, iters=500, mu=2.0):
        C, H, W = crec.shape
        cc = self.A.CoilConstruction(W, H, C, 1)
        return self._n(cc.B1Recon(self._t(u0), self._t(crec), C, iters, mu))

    def coil_construction(self, kdata, mask=None, centered=False):
        T, C, H, W = kdata.shape
        cc, _op = self._coil(W, H, C, T, mask, centered)
        u, b1, order = cc.PerformCoilConstruction(self._t(kdata))
        return self._n(u), self._n(b1), order
