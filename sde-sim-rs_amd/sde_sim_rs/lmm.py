"""Multi-factor forward-rate (LIBOR market) model: swaption price and the 40 curve sensitivities.

No reference counterpart. The kernel (lmm_kernels.hip) steps 40 forwards with spot-measure
log-Euler, the drift evaluated as a prefix sum over forwards, uniforms from the scrambled Sobol
sequence of dimension steps x 3 and the A&S inverse normal; the reverse sweep returns
d price / d L_i(0).
"""
import ctypes

import numpy as np

from . import _native as nat
from .api import _stream_handle

NF, NFACT = 40, 3


def forward_rate_config(seed=5, tenor=0.25):
    """The 40-forward workload: T_i = i * tenor; initial curve 0.02 + 0.01 (1 - exp(-T_i / 5)) plus
    N(0, 1e-4^2) noise from numpy default_rng(seed); volatilities 0.15 + 0.1 exp(-0.3 T_i);
    correlation exp(-0.1 |T_i - T_j|) reduced to rank 3 (top eigenpairs, rows renormalised)."""
    T = tenor * np.arange(NF)
    rng = np.random.default_rng(seed)
    L0 = 0.02 + 0.01 * (1.0 - np.exp(-T / 5.0)) + rng.normal(0.0, 1e-4, NF)
    vol = 0.15 + 0.1 * np.exp(-0.3 * T)
    rho = np.exp(-0.1 * np.abs(T[:, None] - T[None, :]))
    w, V = np.linalg.eigh(rho)
    top = np.argsort(w)[::-1][:NFACT]
    B = V[:, top] * np.sqrt(w[top])
    B = B / np.linalg.norm(B, axis=1, keepdims=True)
    return {"L0": L0, "vol": vol, "loadings": np.ascontiguousarray(B), "times": T}


class ForwardRateModel:
    """Receiver swaption on the swap [T_e, T_40] (e = expiry_index), fixed rate `strike`,
    priced under the spot measure with `steps` log-Euler steps to T_e.

    `storage` picks how the adjoint keeps the forward states: "tape" (all steps + 1 states of the
    40 forwards in device memory), "checkpoints" (every 16th state, each segment recomputed once
    during the reverse sweep) or "auto" (the tape when it fits in half of the free device
    memory). `self.storage` is the form in use."""

    def __init__(self, steps=160, paths=1 << 20, expiry_index=4, strike=0.03, tenor=0.25, scramble_seed=5,
                 precision="f32", storage="auto"):
        if steps % expiry_index:
            raise ValueError("steps must be a multiple of expiry_index")
        storages = {"auto": nat.STORAGE_AUTO, "tape": nat.STORAGE_TAPE, "checkpoints": nat.STORAGE_CHECKPOINTS}
        if storage not in storages:
            raise ValueError("storage must be 'auto', 'tape' or 'checkpoints'")
        spec = nat.LmmSpec()
        spec.forwards, spec.factors = NF, NFACT
        spec.steps = steps
        spec.steps_per_period = steps // expiry_index
        spec.expiry_index = expiry_index
        spec.precision = nat.F32 if precision in ("f32", "float32") else nat.F64
        spec.paths = paths
        spec.tenor = tenor
        spec.strike = strike
        spec.scramble_seed = int(scramble_seed) & (2 ** 64 - 1)
        spec.storage = storages[storage]
        self.steps, self.paths, self.expiry_index, self.strike, self.tenor = steps, paths, expiry_index, strike, tenor
        self.scramble_seed = scramble_seed
        h = ctypes.c_void_p()
        nat.check(nat.lib.sdesim_lmm_create(ctypes.byref(spec), ctypes.byref(h)))
        self._h = h
        self.storage = "tape" if nat.lib.sdesim_lmm_storage(h) == nat.STORAGE_TAPE else "checkpoints"

    def evaluate(self, L0, vol, loadings, adjoint=True, stream=None):
        L0 = np.ascontiguousarray(L0, dtype=np.float64).reshape(-1)
        vol = np.ascontiguousarray(vol, dtype=np.float64).reshape(-1)
        B = np.ascontiguousarray(loadings, dtype=np.float64)
        if L0.size != NF or vol.size != NF or B.size != NF * NFACT:
            raise ValueError("L0 and vol must hold %d values and loadings %d x %d" % (NF, NF, NFACT))
        B = B.reshape(NF, NFACT)
        out = np.zeros(2 + NF, dtype=np.float64)
        nat.check(nat.lib.sdesim_lmm_eval(self._h, nat.dptr(L0), nat.dptr(vol), nat.dptr(B), 1 if adjoint else 0,
                                          nat.dptr(out), _stream_handle(stream)))
        r = {"price": out[0], "variance": out[1]}
        if adjoint:
            r["dL0"] = out[2:].copy()
        return r

    def __del__(self):
        # At interpreter exit the module globals may already be cleared (the process releases
        # the device memory itself then).
        h = getattr(self, "_h", None)
        lib = getattr(nat, "lib", None)
        if h is not None and h.value and lib is not None:
            lib.sdesim_lmm_free(h)
            self._h = None


def _autograd_function():
    import torch

    class ForwardRatePrice(torch.autograd.Function):
        """E[swaption payoff](L0) with the adjoint kernel's 40 curve sensitivities as the gradient.
        The volatilities and loadings are inputs without gradients."""

        @staticmethod
        def forward(ctx, L0, vol, loadings, model):
            r = model.evaluate(L0.detach().cpu().numpy(), vol.detach().cpu().numpy(),
                               loadings.detach().cpu().numpy(), adjoint=True)
            ctx.save_for_backward(torch.as_tensor(r["dL0"], dtype=L0.dtype, device=L0.device))
            return torch.as_tensor(r["price"], dtype=L0.dtype, device=L0.device)

        @staticmethod
        def backward(ctx, g):
            (dL0,) = ctx.saved_tensors
            return g * dL0, None, None, None

    return ForwardRatePrice


_FORWARD_RATE_PRICE = None


def forward_rate_price(L0, vol, loadings, model):
    """Differentiable swaption price (torch tensors in, scalar tensor out; gradient for L0)."""
    global _FORWARD_RATE_PRICE
    if _FORWARD_RATE_PRICE is None:
        _FORWARD_RATE_PRICE = _autograd_function()
    return _FORWARD_RATE_PRICE.apply(L0, vol, loadings, model)
