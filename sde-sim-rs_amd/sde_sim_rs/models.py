"""Built-in models: canonical DSL forms and the batched-instance generator.

The built-in GBM, Ornstein-Uhlenbeck and CIR models of the batched kernel (batch_kernels.hip)
and Heston are written here in the reference's equation DSL (src/proc/util.rs:68-134). A
built-in instance run by the batched kernel and its DSL form run by the generic kernel produce
the same f64 paths on the same grid and seeds. Parameters are printed with repr(), which
round-trips every double exactly.
"""
import numpy as np

GBM, OU, CIR = 0, 1, 2
KIND_NAMES = {GBM: "gbm", OU: "ou", CIR: "cir"}


def gbm(mu, sigma, name="X"):
    return ["d%s = (%r * %s) * dt + (%r * %s) * dW1" % (name, float(mu), name, float(sigma), name)]


def ou(kappa, theta, sigma, name="X"):
    return ["d%s = (%r * (%r - %s)) * dt + (%r) * dW1" % (name, float(kappa), float(theta), name, float(sigma))]


def cir(kappa, theta, sigma, name="X"):
    return ["d%s = (%r * (%r - max(%s, 0))) * dt + (%r * max(%s, 0)^0.5) * dW1"
            % (name, float(kappa), float(theta), name, float(sigma), name)]


def heston(r=0.02, kappa=2.0, theta=0.04, xi=0.5, rho=-0.7):
    """Heston (S, v) with full truncation; the second Brownian factor carries sqrt(1 - rho^2)."""
    rs = float(np.sqrt(1.0 - rho * rho))
    return ["dS = (%r * S) * dt + (max(v, 0)^0.5 * S) * dW1" % float(r),
            "dv = (%r * (%r - max(v, 0))) * dt + (%r * %r * max(v, 0)^0.5) * dW1 + (%r * %r * max(v, 0)^0.5) * dW2"
            % (float(kappa), float(theta), float(rho), float(xi), rs, float(xi))]


def instance_equations(kind, p0, p1, p2):
    if kind == GBM:
        return gbm(p0, p1)
    if kind == OU:
        return ou(p0, p1, p2)
    if kind == CIR:
        return cir(p0, p1, p2)
    raise ValueError("unknown model kind %r" % (kind,))


def batched_instances(n=4096, seed=11):
    """The batched workload: n independent 1-D SDEs drawn from `seed` (numpy default_rng).

    type uniform over {GBM, OU, CIR}; mu, theta ~ U(0, 0.1); sigma ~ U(0.05, 0.6);
    kappa ~ U(0.5, 5); X0 ~ U(50, 150) (CIR: U(0.01, 0.1)); paths log-uniform in [2^8, 2^16];
    steps uniform integer in [50, 2000]; horizon ~ U(0.25, 5); each instance's own seed is a
    63-bit integer drawn from the same generator (path s uses seed_i + s).
    Returns a list of dicts with keys kind, steps, paths, x0, horizon, p0, p1, p2, seed.
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kind = int(rng.integers(0, 3))
        mu_or_theta = float(rng.uniform(0.0, 0.1))
        sigma = float(rng.uniform(0.05, 0.6))
        kappa = float(rng.uniform(0.5, 5.0))
        x0 = float(rng.uniform(0.01, 0.1)) if kind == CIR else float(rng.uniform(50.0, 150.0))
        paths = int(np.floor(np.exp(rng.uniform(np.log(2.0 ** 8), np.log(2.0 ** 16)))))
        steps = int(rng.integers(50, 2001))
        horizon = float(rng.uniform(0.25, 5.0))
        iseed = int(rng.integers(0, 2 ** 63))
        if kind == GBM:
            p = (mu_or_theta, sigma, 0.0)
        else:
            p = (kappa, mu_or_theta, sigma)
        out.append(dict(kind=kind, steps=steps, paths=paths, x0=x0, horizon=horizon, p0=p[0], p1=p[1], p2=p[2],
                        seed=iseed))
    return out


def discrete_euler_mean(inst):
    """E[X_N] of the Euler chain for GBM and OU (exact for the discrete scheme):
    GBM X0 (1 + mu dt)^N; OU theta + (X0 - theta)(1 - kappa dt)^N. CIR: the continuous-time
    mean theta + (X0 - theta) exp(-kappa T) (the truncated chain differs by O(dt))."""
    dt = inst["horizon"] / inst["steps"]
    if inst["kind"] == GBM:
        return inst["x0"] * (1.0 + inst["p0"] * dt) ** inst["steps"]
    if inst["kind"] == OU:
        return inst["p1"] + (inst["x0"] - inst["p1"]) * (1.0 - inst["p0"] * dt) ** inst["steps"]
    return inst["p1"] + (inst["x0"] - inst["p1"]) * np.exp(-inst["p0"] * inst["horizon"])
