"""Forward-rate model on the GPU: price and curve sensitivities against reverse-mode autodiff of
the same paths (torch, CPU float64) and against central finite differences (common random
numbers)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _sobol_normals(oracle, paths, steps, seed):
    """eps[s, n, f] from Sobol point s, dimension 3 n + f, XOR mask word, u = (y + 0.5) 2^-32."""
    from sde_sim_rs import _native as nat
    dims = steps * 3
    pts = np.zeros((paths, dims), dtype=np.uint32)
    nat.check(nat.lib.sdesim_sobol_points(dims, 0, paths, pts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
    y = pts ^ oracle.xor_masks(seed, dims)[None, :]
    u = (y.astype(np.float64) + 0.5) * 2.0 ** -32
    eps = np.vectorize(oracle.inv_normal)(u)
    return eps.reshape(paths, steps, 3)


def test_lmm_f64_matches_torch_autodiff(oracle):
    from sde_sim_rs.lmm import ForwardRateModel, forward_rate_config
    c = forward_rate_config(seed=5)
    S, N, e, K, tau = 64, 16, 4, 0.03, 0.25
    print("case: lmm f64 price + adjoint vs torch autodiff S=%d N=%d" % (S, N), flush=True)
    m = ForwardRateModel(steps=N, paths=S, expiry_index=e, strike=K, scramble_seed=5, precision="f64")
    r = m.evaluate(c["L0"], c["vol"], c["loadings"], adjoint=True)
    eps = torch.tensor(_sobol_normals(oracle, S, N, 5))
    L0 = torch.tensor(c["L0"], requires_grad=True)
    vol = torch.tensor(c["vol"])
    B = torch.tensor(c["loadings"])
    dt = tau * e / N
    spp = N // e
    L = L0.expand(S, 40).clone()
    for n in range(N):
        eta = n // spp + 1
        a = tau * vol[eta:] * L[:, eta:] / (1 + tau * L[:, eta:])
        C = torch.cumsum(a[:, :, None] * B[None, eta:, :], dim=1)
        drift = vol[eta:] * (B[None, eta:, :] * C).sum(-1) - 0.5 * vol[eta:] ** 2
        w = (B[None, eta:, :] * eps[:, n, None, :]).sum(-1)
        L = torch.cat([L[:, :eta], L[:, eta:] * torch.exp(drift * dt + vol[eta:] * np.sqrt(dt) * w)], dim=1)
    numer = torch.prod(1 + tau * L[:, :e], dim=1)
    P = torch.cumprod(1 / (1 + tau * L[:, e:]), dim=1)
    A = tau * P.sum(1)
    U = K * A - (1 - P[:, -1])
    V = torch.clamp(U, min=0) / numer
    V.mean().backward()
    assert abs(r["price"] - float(V.mean())) <= 1e-11 * abs(float(V.mean())), (r["price"], float(V.mean()))
    np.testing.assert_allclose(r["dL0"], L0.grad.numpy(), rtol=1e-8, atol=1e-10)


def test_lmm_f32_sensitivities_against_central_differences():
    from sde_sim_rs.lmm import ForwardRateModel, forward_rate_config
    c = forward_rate_config(seed=5)
    S, N = 1 << 16, 160
    print("case: lmm f32 adjoint vs f64 central FD (CRN) S=%d N=%d" % (S, N), flush=True)
    adj = ForwardRateModel(steps=N, paths=S, precision="f32").evaluate(c["L0"], c["vol"], c["loadings"])
    fwd = ForwardRateModel(steps=N, paths=S, precision="f64")
    worst = 0.0
    scale = np.max(np.abs(adj["dL0"]))
    for i in range(40):
        h = 1e-3 * c["L0"][i]
        lp, lm = c["L0"].copy(), c["L0"].copy()
        lp[i] += h
        lm[i] -= h
        fd = (fwd.evaluate(lp, c["vol"], c["loadings"], adjoint=False)["price"]
              - fwd.evaluate(lm, c["vol"], c["loadings"], adjoint=False)["price"]) / (2 * h)
        err = abs(adj["dL0"][i] - fd) / max(abs(fd), 1e-3 * scale)
        worst = max(worst, err)
        assert err <= 1e-4, (i, adj["dL0"][i], fd)
    print("max relative error of 40 curve sensitivities vs FD: %.3g" % worst)


def test_forward_rate_torch_autograd_function():
    from sde_sim_rs.lmm import ForwardRateModel, forward_rate_config, forward_rate_price
    c = forward_rate_config(seed=5)
    print("case: lmm autograd function S=4096 N=16", flush=True)
    m = ForwardRateModel(steps=16, paths=4096, precision="f64")
    r = m.evaluate(c["L0"], c["vol"], c["loadings"], adjoint=True)
    L0 = torch.tensor(c["L0"], requires_grad=True)
    p = forward_rate_price(L0, torch.tensor(c["vol"]), torch.tensor(c["loadings"]), m)
    assert abs(float(p.detach()) - r["price"]) <= 1e-15 * abs(r["price"])
    (2.0 * p).backward()
    np.testing.assert_allclose(L0.grad.numpy(), 2.0 * r["dL0"], rtol=1e-12)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_lmm_tape_equals_checkpoints(precision):
    """The two state storages of the adjoint (every state on a tape; checkpoints every 16 steps,
    each segment recomputed into the segment tape) reverse the same paths: price and the 40
    sensitivities agree to rounding. 2^12 + 37 paths (a partial last block), 40 steps to expiry
    index 4 (two full segments and a partial one)."""
    from sde_sim_rs.lmm import ForwardRateModel, forward_rate_config
    c = forward_rate_config(seed=5)
    S, N = (1 << 12) + 37, 40
    print("case: lmm %s tape vs checkpoints S=%d N=%d" % (precision, S, N), flush=True)
    out = {}
    for storage in ("tape", "checkpoints"):
        m = ForwardRateModel(steps=N, paths=S, precision=precision, storage=storage)
        assert m.storage == storage
        out[storage] = m.evaluate(c["L0"], c["vol"], c["loadings"], adjoint=True)
    rtol = 1e-12 if precision == "f64" else 1e-5
    t, k = out["tape"], out["checkpoints"]
    assert abs(t["price"] - k["price"]) <= rtol * abs(k["price"])
    np.testing.assert_allclose(t["dL0"], k["dL0"], rtol=0, atol=rtol * np.max(np.abs(k["dL0"])))


def test_forward_rate_model_refuses_mismatched_parameters():
    """A curve, volatility vector or loading matrix of the wrong size raises ValueError before
    the library reads it."""
    from sde_sim_rs.lmm import ForwardRateModel, forward_rate_config
    c = forward_rate_config(seed=5)
    print("case: forward-rate parameter checks", flush=True)
    m = ForwardRateModel(steps=16, paths=256, expiry_index=4, precision="f32")
    for k, v in [("L0", c["L0"][:39]), ("vol", np.append(c["vol"], 0.2)), ("loadings", c["loadings"][:, :2])]:
        args = dict(L0=c["L0"], vol=c["vol"], loadings=c["loadings"])
        args[k] = v
        with pytest.raises(ValueError):
            m.evaluate(**args)
    r = m.evaluate(c["L0"], c["vol"], c["loadings"])
    assert np.isfinite(r["price"]) and r["dL0"].shape == (40,)
