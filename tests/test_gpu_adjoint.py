"""Basket price and pathwise adjoint on the GPU: against the DSL form, against reverse-mode
autodiff of the same paths (torch, CPU float64), against central finite differences with
common random numbers, and against Black-Scholes delta and vega for one asset."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _cfg():
    from sde_sim_rs.adjoint import basket_config
    return basket_config(seed=3)


def test_basket_price_matches_dsl_form():
    import sde_sim_rs
    from sde_sim_rs.adjoint import Basket, basket_dsl
    c = _cfg()
    S, N = 1 << 14, 73
    print("case: basket f64 price vs DSL form S=%d N=%d" % (S, N), flush=True)
    b = Basket(8, N, S, horizon=1.0, strike=100.0, width=0.0, seed=17, precision="f64")
    r = b.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=False)
    eqs, init = basket_dsl(c["x0"], c["mu"], c["sigma"], c["rho"])
    times = np.arange(N + 1) * (1.0 / N)
    plan = sde_sim_rs.Plan(eqs, times, S, init, "pseudo", "euler", seed=17, outputs=("terminal",),
                           payoff="max((X0 + X1 + X2 + X3 + X4 + X5 + X6 + X7) / 8 - 100, 0)")
    plan.run(None)
    m = plan.moments()["payoff"]
    assert abs(r["price"] - m[0]) <= 1e-9 * abs(m[0]), (r["price"], m[0])


@pytest.mark.parametrize("assets", [3, 5])
def test_basket_fewer_than_eight_assets_matches_dsl_form(assets):
    """Baskets of 2-7 assets run the eight-asset kernels padded: asset k of step n still draws
    number n*A + k of the path's stream, as the A-increment DSL form does, so the price equals the
    DSL form's (FP64, plain call)."""
    import sde_sim_rs
    from sde_sim_rs.adjoint import Basket, basket_dsl
    c = _cfg()
    A = assets
    x0, mu, sig, rho = c["x0"][:A], c["mu"][:A], c["sigma"][:A], c["rho"][:A, :A]
    S, N = 1 << 13, 29
    print("case: basket f64 price vs DSL form A=%d S=%d N=%d" % (A, S, N), flush=True)
    b = Basket(A, N, S, horizon=1.0, strike=100.0, width=0.0, seed=23, precision="f64")
    r = b.evaluate(x0, mu, sig, rho, adjoint=False)
    eqs, init = basket_dsl(x0, mu, sig, rho)
    times = np.arange(N + 1) * (1.0 / N)
    plan = sde_sim_rs.Plan(eqs, times, S, init, "pseudo", "euler", seed=23, outputs=("terminal",),
                           payoff="max((%s) / %d - 100, 0)" % (" + ".join("X%d" % i for i in range(A)), A))
    plan.run(None)
    m = plan.moments()["payoff"]
    assert abs(r["price"] - m[0]) <= 1e-9 * abs(m[0]), (r["price"], m[0])


@pytest.mark.parametrize("D", [8, 3])
def test_adjoint_f64_matches_torch_autodiff(oracle, D):
    from sde_sim_rs.adjoint import Basket
    c = _cfg()
    c = {"x0": c["x0"][:D], "mu": c["mu"][:D], "sigma": c["sigma"][:D], "rho": c["rho"][:D, :D]}
    S, N, seed = 32, 45, 1234
    print("case: basket f64 adjoint vs torch autodiff D=%d S=%d N=%d" % (D, S, N), flush=True)
    b = Basket(D, N, S, horizon=1.0, strike=100.0, width=0.5, seed=seed, precision="f64")
    r = b.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=True)
    # Reference: the same paths in torch float64 on the CPU, differentiated by autograd.
    L = torch.tensor(np.linalg.cholesky(c["rho"]), dtype=torch.float64, requires_grad=True)
    x0 = torch.tensor(c["x0"], requires_grad=True)
    mu = torch.tensor(c["mu"], requires_grad=True)
    sg = torch.tensor(c["sigma"], requires_grad=True)
    eps = np.zeros((S, N, D))
    for s in range(S):
        u = oracle.pseudo_f64(seed + s, N * D)
        eps[s] = np.array([oracle.inv_normal(v) for v in u]).reshape(N, D)
    eps = torch.tensor(eps)
    dt = 1.0 / N
    X = x0.expand(S, D)
    for n in range(N):
        Z = eps[:, n, :] @ torch.tril(L).T
        X = X + X * (mu * dt + sg * np.sqrt(dt) * Z)
    B = X.mean(dim=1)
    pay = 0.5 * torch.nn.functional.softplus((B - 100.0) / 0.5)
    pay.mean().backward()
    assert abs(r["price"] - float(pay.mean())) <= 1e-12 * abs(r["price"])
    for name, t in (("x0", x0), ("mu", mu), ("sigma", sg)):
        np.testing.assert_allclose(r[name], t.grad.numpy(), rtol=1e-9, atol=1e-12, err_msg=name)
    np.testing.assert_allclose(r["chol"], np.tril(L.grad.numpy()), rtol=1e-9, atol=1e-12)


def test_adjoint_f32_against_central_differences():
    from sde_sim_rs.adjoint import Basket
    c = _cfg()
    S, N = 1 << 16, 365
    print("case: basket f32 adjoint vs f64 central FD (CRN) S=%d N=%d" % (S, N), flush=True)
    adj = Basket(8, N, S, strike=100.0, width=0.5, seed=99, precision="f32").evaluate(
        c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=True)
    fwd = Basket(8, N, S, strike=100.0, width=0.5, seed=99, precision="f64")
    price = lambda **kw: fwd.evaluate(**{**c, **kw}, adjoint=False)["price"]  # noqa: E731
    worst = 0.0
    for name in ("x0", "mu", "sigma"):
        for i in range(8):
            v = c[name].copy()
            h = 1e-3 * abs(v[i])
            vp, vm = v.copy(), v.copy()
            vp[i] += h
            vm[i] -= h
            fd = (price(**{name: vp}) - price(**{name: vm})) / (2 * h)
            scale = max(abs(fd), 1e-3 * np.max(np.abs(adj[name])))
            err = abs(adj[name][i] - fd) / scale
            worst = max(worst, err)
            assert err <= 1e-4, (name, i, adj[name][i], fd)
    rho = c["rho"]
    for i in range(8):
        for j in range(i):
            h = 1e-3 * max(abs(rho[i, j]), 0.1)
            rp, rm = rho.copy(), rho.copy()
            rp[i, j] = rp[j, i] = rho[i, j] + h
            rm[i, j] = rm[j, i] = rho[i, j] - h
            fd = (price(rho=rp) - price(rho=rm)) / (2 * h)
            an = adj["rho"][i, j] + adj["rho"][j, i]
            scale = max(abs(fd), 1e-3 * np.max(np.abs(adj["rho"])))
            err = abs(an - fd) / scale
            worst = max(worst, err)
            assert err <= 1e-4, ("rho", i, j, an, fd)
    print("max relative error of 52 adjoint Greeks vs FD: %.3g" % worst)


def test_black_scholes_delta_and_vega_one_asset():
    from scipy.stats import norm
    from sde_sim_rs.adjoint import Basket
    x0, mu, sig, K, T = 100.0, 0.03, 0.25, 100.0, 1.0
    S, N = 1 << 20, 365
    print("case: one-asset call delta/vega vs Black-Scholes S=%d N=%d" % (S, N), flush=True)
    r = Basket(1, N, S, horizon=T, strike=K, width=0.0, seed=5, precision="f32").evaluate(
        [x0], [mu], [sig], [[1.0]], adjoint=True)
    F = x0 * np.exp(mu * T)
    d1 = (np.log(x0 / K) + (mu + 0.5 * sig * sig) * T) / (sig * np.sqrt(T))
    price = F * norm.cdf(d1) - K * norm.cdf(d1 - sig * np.sqrt(T))
    delta = np.exp(mu * T) * norm.cdf(d1)
    vega = F * norm.pdf(d1) * np.sqrt(T)
    se = np.sqrt(r["variance"] / S)
    assert abs(r["price"] - price) < 4 * se + 2e-3 * price, (r["price"], price)
    assert abs(r["x0"][0] - delta) < 4e-3 * delta, (r["x0"][0], delta)
    assert abs(r["sigma"][0] - vega) < 1.5e-2 * vega, (r["sigma"][0], vega)


def test_torch_autograd_function():
    from sde_sim_rs.adjoint import Basket, basket_price
    c = _cfg()
    b = Basket(8, 50, 1 << 14, seed=3, precision="f32")
    x0 = torch.tensor(c["x0"], requires_grad=True)
    mu = torch.tensor(c["mu"], requires_grad=True)
    sg = torch.tensor(c["sigma"], requires_grad=True)
    rho = torch.tensor(c["rho"], requires_grad=True)
    p = basket_price(x0, mu, sg, rho, b)
    (2.0 * p).backward()
    r = b.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"])
    np.testing.assert_allclose(x0.grad.numpy(), 2 * r["x0"], rtol=1e-12)
    np.testing.assert_allclose(rho.grad.numpy(), 2 * r["rho"], rtol=1e-12)


@pytest.mark.parametrize("precision,D", [("f64", 8), ("f32", 8), ("f64", 1)])
def test_adjoint_tape_equals_checkpoints(precision, D):
    """The two state storages of the adjoint (every state on a tape; checkpoints every 19 steps
    with recomputation) reverse the same paths: price and gradients agree to rounding. 2^12 + 37
    paths (a partial last block), 45 steps (two full checkpoint segments and a partial one)."""
    from sde_sim_rs.adjoint import Basket
    c = _cfg()
    x0, mu, sg, rho = c["x0"][:D], c["mu"][:D], c["sigma"][:D], c["rho"][:D, :D]
    S, N = (1 << 12) + 37, 45
    print("case: basket %s D=%d tape vs checkpoints S=%d N=%d" % (precision, D, S, N), flush=True)
    out = {}
    for storage in ("tape", "checkpoints"):
        b = Basket(D, N, S, strike=100.0, width=0.5, seed=11, precision=precision, storage=storage)
        assert b.storage == storage
        out[storage] = b.evaluate(x0, mu, sg, rho, adjoint=True)
    rtol = 1e-12 if precision == "f64" else 1e-5
    t, k = out["tape"], out["checkpoints"]
    assert abs(t["price"] - k["price"]) <= rtol * abs(k["price"])
    for name in ("x0", "mu", "sigma", "chol"):
        scale = np.max(np.abs(k[name]))
        np.testing.assert_allclose(t[name], k[name], rtol=0, atol=rtol * scale, err_msg=name)


def test_handles_release_device_memory():
    """Creating and freeing adjoint handles returns their device memory (the basket's state tape
    here is 3.4 GB): after one cycle (which also lets the runtime make its one-time allocations,
    such as loading the kernels' code), nine more create/evaluate/free cycles leave free memory
    where it was."""
    import gc
    from sde_sim_rs.adjoint import Basket
    from sde_sim_rs.lmm import ForwardRateModel, forward_rate_config
    c, lc = _cfg(), forward_rate_config(seed=5)
    print("case: handle memory release", flush=True)

    def cycle():
        b = Basket(8, 100, 1 << 20, seed=1, precision="f32", storage="tape")
        b.evaluate(c["x0"], c["mu"], c["sigma"], c["rho"], adjoint=True)
        m = ForwardRateModel(steps=40, paths=1 << 14, precision="f32", storage="checkpoints")
        m.evaluate(lc["L0"], lc["vol"], lc["loadings"], adjoint=True)
        del b, m
        gc.collect()
        torch.cuda.synchronize()

    cycle()
    before = torch.cuda.mem_get_info()[0]
    for _ in range(9):
        cycle()
    after = torch.cuda.mem_get_info()[0]
    assert after >= before - (16 << 20), (before, after)


def test_basket_refuses_mismatched_parameters():
    """Parameter vectors of the wrong length and a correlation matrix of the wrong shape raise
    ValueError before the library reads them."""
    from sde_sim_rs.adjoint import Basket
    print("case: basket parameter checks", flush=True)
    b = Basket(3, 4, 256, seed=1, precision="f32")
    ok = dict(x0=[100.0] * 3, mu=[0.01] * 3, sigma=[0.2] * 3, rho=np.eye(3))
    for k, v in [("x0", [100.0] * 2), ("mu", [0.01] * 4), ("sigma", [0.2]), ("rho", np.eye(2))]:
        with pytest.raises(ValueError):
            b.evaluate(**dict(ok, **{k: v}))
    with pytest.raises(ValueError):
        b.evaluate(**ok, weights=[0.5, 0.5])
    r = b.evaluate(**ok)
    assert np.isfinite(r["price"]) and r["x0"].shape == (3,)
