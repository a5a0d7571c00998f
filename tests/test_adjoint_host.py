"""CPU checks of the adjoint host logic: Cholesky reverse mode and the basket workload."""
import numpy as np

from sde_sim_rs.adjoint import basket_config, basket_dsl, cholesky_backward


def test_cholesky_backward_against_finite_differences():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((6, 6))
    S = A @ A.T + 6 * np.eye(6)
    Lbar = np.tril(rng.standard_normal((6, 6)))
    f = lambda M: float(np.sum(Lbar * np.linalg.cholesky(M)))  # noqa: E731
    G = cholesky_backward(np.linalg.cholesky(S), Lbar)
    h = 1e-6
    for i in range(6):
        for j in range(i + 1):
            E = np.zeros((6, 6))
            E[i, j] = E[j, i] = 1.0
            fd = (f(S + h * E) - f(S - h * E)) / (2 * h)
            an = G[i, j] + G[j, i] if i != j else G[i, i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (i, j, fd, an)


def test_basket_config_is_a_correlation_matrix():
    c = basket_config(seed=3)
    rho = c["rho"]
    assert rho.shape == (8, 8)
    assert np.allclose(np.diag(rho), 1.0) and np.allclose(rho, rho.T)
    assert np.all(np.linalg.eigvalsh(rho) > 0)
    assert np.all((c["x0"] >= 80) & (c["x0"] <= 120)) and np.all((c["sigma"] >= 0.1) & (c["sigma"] <= 0.5))


def test_basket_dsl_form_parses():
    import sde_sim_rs
    c = basket_config(seed=3)
    eqs, init = basket_dsl(c["x0"], c["mu"], c["sigma"], c["rho"])
    u = sde_sim_rs.Universe(eqs)
    assert u.num_processes == 8 and u.num_increments == 8
