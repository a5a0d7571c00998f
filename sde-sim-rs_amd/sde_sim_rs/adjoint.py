"""Pathwise sensitivities of a correlated GBM basket (no reference counterpart).

The price kernel runs the Euler step of the DSL form
    dX_i = (mu_i * X_i) * dt + sum_k (sigma_i * L_ik * X_i) * dW_k
(src/sim/euler.rs:5-37 on the reference's equation language, with the Cholesky factor L of the
correlation written into the coefficients) and a reverse-sweep adjoint kernel returns the
path-averaged derivatives with respect to X0, mu, sigma and L. This module maps dL to the
correlation matrix (reverse mode of the Cholesky factorisation) and wraps the pair in a
torch.autograd.Function.
"""
import ctypes

import numpy as np

from . import _native as nat
from .api import _stream_handle


def basket_config(seed=3, assets=8, factors=3):
    """The 8-asset basket workload: loadings B (assets x factors) ~ N(0, 1) plus a unit diagonal
    idiosyncratic variance, Sigma = B B^T + I normalised to unit diagonal; X0 ~ U(80, 120),
    mu ~ U(0, 0.08), sigma ~ U(0.1, 0.5), all from numpy default_rng(seed) in that order."""
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((assets, factors))
    cov = B @ B.T + np.eye(assets)
    d = np.sqrt(np.diag(cov))
    rho = cov / np.outer(d, d)
    np.fill_diagonal(rho, 1.0)
    x0 = rng.uniform(80.0, 120.0, assets)
    mu = rng.uniform(0.0, 0.08, assets)
    sigma = rng.uniform(0.1, 0.5, assets)
    return {"x0": x0, "mu": mu, "sigma": sigma, "rho": rho}


def cholesky_backward(L, Lbar):
    """Reverse mode of Sigma = L L^T (Murray, "Differentiation of the Cholesky decomposition",
    2016): Sigmabar = sym(L^-T Phi(L^T Lbar) L^-1), Phi = lower triangle with halved diagonal,
    sym(A) = (A + A^T) / 2. Returns the symmetric Sigmabar (df/dSigma_ij with Sigma_ij and
    Sigma_ji treated as one entry counted on both sides)."""
    L = np.asarray(L, dtype=np.float64)
    Lbar = np.tril(np.asarray(Lbar, dtype=np.float64))
    P = np.tril(L.T @ Lbar)
    P[np.diag_indices_from(P)] *= 0.5
    Linv = np.linalg.inv(L)
    S = Linv.T @ P @ Linv
    return 0.5 * (S + S.T)


def basket_dsl(x0, mu, sigma, rho):
    """The basket's DSL form (one Levy process per asset, one dW per factor of L)."""
    L = np.linalg.cholesky(np.asarray(rho, dtype=np.float64))
    D = len(x0)
    eqs = []
    for i in range(D):
        terms = ["(%r * X%d) * dt" % (float(mu[i]), i)]
        for k in range(i + 1):
            terms.append("(%r * %r * X%d) * dW%d" % (float(sigma[i]), float(L[i, k]), i, k + 1))
        eqs.append("dX%d = " % i + " + ".join(terms))
    init = {"X%d" % i: float(x0[i]) for i in range(D)}
    return eqs, init


_STORAGE = {"auto": nat.STORAGE_AUTO, "tape": nat.STORAGE_TAPE, "checkpoints": nat.STORAGE_CHECKPOINTS}


class Basket:
    """Price and adjoint of a weighted basket call on D <= 8 correlated GBM assets.

    `storage` picks how the adjoint keeps the forward states: "tape" (every state in device
    memory, steps * 8 * paths values for D > 1), "checkpoints" (every 19th state, the rest
    recomputed during the reverse sweep) or "auto" (the tape when it fits in half of the free
    device memory). `self.storage` is the form in use."""

    def __init__(self, assets, steps, paths, horizon=1.0, strike=100.0, width=0.5, seed=0, precision="f32",
                 storage="auto"):
        if storage not in _STORAGE:
            raise ValueError("storage must be 'auto', 'tape' or 'checkpoints'")
        spec = nat.BasketSpec()
        spec.assets = assets
        spec.steps = steps
        spec.paths = paths
        spec.horizon = horizon
        spec.strike = strike
        spec.width = width
        spec.seed = int(seed) & (2 ** 64 - 1)
        spec.precision = nat.F32 if precision in ("f32", "float32") else nat.F64
        spec.storage = _STORAGE[storage]
        self.assets = assets
        self.steps, self.paths, self.horizon, self.strike, self.width = steps, paths, horizon, strike, width
        self.precision = precision
        h = ctypes.c_void_p()
        nat.check(nat.lib.sdesim_basket_create(ctypes.byref(spec), ctypes.byref(h)))
        self._h = h
        self.storage = "tape" if nat.lib.sdesim_basket_storage(h) == nat.STORAGE_TAPE else "checkpoints"

    def evaluate(self, x0, mu, sigma, rho, weights=None, adjoint=True, stream=None):
        D = self.assets

        def vec(v, name):
            v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
            if v.size != D:
                raise ValueError("%s must hold %d values (one per asset), got %d" % (name, D, v.size))
            return v

        x0, mu, sigma = vec(x0, "x0"), vec(mu, "mu"), vec(sigma, "sigma")
        w = np.full(D, 1.0 / D) if weights is None else vec(weights, "weights")
        rho = np.asarray(rho, dtype=np.float64)
        if rho.shape != (D, D):
            raise ValueError("rho must be a %d x %d correlation matrix, got shape %s" % (D, D, rho.shape))
        L = np.ascontiguousarray(np.linalg.cholesky(rho))
        nl = D * (D + 1) // 2
        out = np.zeros(2 + 3 * D + nl, dtype=np.float64)
        nat.check(nat.lib.sdesim_basket_eval(self._h, nat.dptr(x0), nat.dptr(mu), nat.dptr(sigma), nat.dptr(w),
                                             nat.dptr(L), 1 if adjoint else 0, nat.dptr(out), _stream_handle(stream)))
        res = {"price": out[0], "variance": out[1], "L": L}
        if adjoint:
            Lbar = np.zeros((D, D))
            Lbar[np.tril_indices(D)] = out[2 + 3 * D:]
            res.update({"x0": out[2:2 + D].copy(), "mu": out[2 + D:2 + 2 * D].copy(),
                        "sigma": out[2 + 2 * D:2 + 3 * D].copy(), "chol": Lbar,
                        "rho": cholesky_backward(L, Lbar)})
        return res

    def __del__(self):
        # At interpreter exit the module globals may already be cleared (the process releases
        # the device memory itself then).
        h = getattr(self, "_h", None)
        lib = getattr(nat, "lib", None)
        if h is not None and h.value and lib is not None:
            lib.sdesim_basket_free(h)
            self._h = None


def _autograd_function():
    import torch

    class BasketPrice(torch.autograd.Function):
        """E[payoff](x0, mu, sigma, rho) with the adjoint kernel's pathwise gradients.
        rho's gradient is the symmetric dE/dSigma (both triangles)."""

        @staticmethod
        def forward(ctx, x0, mu, sigma, rho, basket):
            r = basket.evaluate(x0.detach().cpu().numpy(), mu.detach().cpu().numpy(),
                                sigma.detach().cpu().numpy(), rho.detach().cpu().numpy(), adjoint=True)
            f = lambda a, like: torch.as_tensor(a, dtype=like.dtype, device=like.device)  # noqa: E731
            ctx.save_for_backward(f(r["x0"], x0), f(r["mu"], mu), f(r["sigma"], sigma), f(r["rho"], rho))
            return torch.as_tensor(r["price"], dtype=x0.dtype, device=x0.device)

        @staticmethod
        def backward(ctx, g):
            gx0, gmu, gsig, grho = ctx.saved_tensors
            return g * gx0, g * gmu, g * gsig, g * grho, None

    return BasketPrice


_BASKET_PRICE = None


def basket_price(x0, mu, sigma, rho, basket):
    """Differentiable basket price (torch tensors in, scalar tensor out)."""
    global _BASKET_PRICE
    if _BASKET_PRICE is None:
        _BASKET_PRICE = _autograd_function()
    return _BASKET_PRICE.apply(x0, mu, sigma, rho, basket)
