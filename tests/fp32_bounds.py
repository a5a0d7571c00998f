"""Rounding-error bound for FP32 Heston paths against the FP64 paths (test helper).

The FP32 kernel of the Heston "Runge-Kutta" scheme (src/sim/runge_kutta.rs:5-107 with the
reference's stale probe cache) evaluates the same recursion as the FP64 oracle on the same
Sobol-XOR uniforms, with every operation rounded to f32. `heston_rk_bound` runs the FP64
recursion along each path and, beside it, a running bound on |FP32 value - FP64 value| for the
state (S, v) and for the probe row that the next step's stage 1 reads:

* every f32 operation rounds by at most u = 2^-24 relative; a sum of products is charged 2u per
  term (a product and an addition, whether or not they are fused); constants and the grid's dt
  and sqrt(dt) are rounded once to f32 (u);
* v_log_f32, v_sqrt_f32 and v_rcp_f32 are charged 2 ulp (4u; the ISA documents 1 ulp);
* the inverse normal: q = min(u, 1 - u) is formed from the integer point by one conversion and
  one fused multiply-add (2u q); a relative error e in q moves -log2 q by e / ln 2, t' = sqrt(L)
  then carries L's error halved plus its own, and x' = t' - num'(t') / den'(t') carries t''s
  error times |dx'/dt'| plus the rational's own rounding (24u of num'/den': 7 roundings of the
  numerator, 9 of the denominator, the reciprocal and the product) and the subtraction's (u |x'|);
  the host-tabulated tail (qi < 1233, |z| > 5) is rounded once (u |x'|);
* the kink: for a state error e, max(v, 0)^0.5 moves by at most
  sup over |w - v| <= e of |sqrt(max(w, 0)) - sqrt(max(v, 0))|, which is e / (2 sqrt(v)) to first
  order away from zero and sqrt(e) at zero;
* errors propagate through every coefficient, increment and stage to first order in each error
  and with their products (error x error) added, so a larger incoming error cannot hide.

The bound makes no use of any FP32 result. tests/test_fp32_bounds.py checks it against an f32
evaluation of the step in numpy (which must stay inside it) and against planted mistakes (which
must leave it), for the Sobol-XOR uniforms and for the pseudo-random draws.
"""
import numpy as np

U = 2.0 ** -24
C_PSI = float(np.sqrt(2.0 * np.log(2.0)))  # sqrt(2 ln 2): t = C t'
# A&S 26.2.23 constants (src/proc/increment.rs:170-175) rescaled to t' = t / C, x' = x / C.
A0, A1, A2 = 2.515517 / C_PSI, 0.802853, 0.010328 * C_PSI
B1, B2, B3 = 1.432788 * C_PSI, 0.189269 * C_PSI ** 2, 0.001308 * C_PSI ** 3


def inv_normal(p):
    """The reference's f64 inverse normal (A&S 26.2.23) of an array."""
    p = np.asarray(p, dtype=np.float64)
    q = np.where(p < 0.5, p, 1.0 - p)
    t = np.sqrt(-2.0 * np.log(q))
    x = t - ((0.010328 * t + 0.802853) * t + 2.515517) / (((0.001308 * t + 0.189269) * t + 1.432788) * t + 1.0)
    return np.where(p < 0.5, -x, x)


def psi_error(y):
    """Bound on |Psi_f32(u) - Psi_f64(u)| / C (units of x' = |z| / C) for the XOR-mode uniform
    u = (y + 0.5) 2^-32 of the 32-bit point word y."""
    y = np.asarray(y, dtype=np.uint32)
    qi = np.minimum(y, ~y)
    q = (qi.astype(np.float64) + 0.5) * 2.0 ** -32
    L = -np.log2(q)
    tp = np.sqrt(L)
    num = (A2 * tp + A1) * tp + A0
    den = ((B3 * tp + B2) * tp + B1) * tp + 1.0
    R = num / den
    xp = tp - R
    dnum = 2.0 * A2 * tp + A1
    dden = (3.0 * B3 * tp + 2.0 * B2) * tp + B1
    dxdt = np.abs(1.0 - (dnum * den - num * dden) / (den * den))
    # q: 2u q; L: 4u L + 2u / ln 2 (+1%); t': dL / (2 t') + 4u t'.
    dL = 4.0 * U * L + 2.01 * U / np.log(2.0)
    dt = dL / (2.0 * tp) + 4.0 * U * tp
    err = dxdt * dt + 24.0 * U * R + U * np.abs(xp)
    return np.where(qi < 1233, U * np.abs(xp), err)


def psi_error_words(u):
    """Bound on |Psi_f32(u) - Psi_f64(u)| (units of z) for the pseudo-random (and FP32 Sobol-shift)
    form, which takes q = min(u, 1 - u) from the draw's two 32-bit words: q' = fma(f32(hi),
    2^-32, f32(lo) 2^-64) is within 2u q + 2^-53 of the f64 uniform's q; t = sqrt(-2 ln 2 *
    log2 q') with the constant rounded to f32; x = t - num(t) / den(t) with the A&S constants
    rounded to f32. q' < 2.87e-7 (|z| > 5) takes the f64 formula on the exact q, rounded once."""
    u = np.asarray(u, dtype=np.float64)
    q = np.where(u < 0.5, u, 1.0 - u)
    L = -np.log2(q)
    t = np.sqrt(2.0 * np.log(2.0) * L)
    num = (0.010328 * t + 0.802853) * t + 2.515517
    den = ((0.001308 * t + 0.189269) * t + 1.432788) * t + 1.0
    R = num / den
    x = t - R
    dnum = 2.0 * 0.010328 * t + 0.802853
    dden = (3.0 * 0.001308 * t + 2.0 * 0.189269) * t + 1.432788
    dxdt = np.abs(1.0 - (dnum * den - num * dden) / (den * den))
    # q: 2u q + 2^-53 (relative 2.01u for q >= 2.87e-7); L: 4u L + that / ln 2; the argument of
    # the square root: L's error + 2u (the constant and the product); t: half of it + 4u.
    dq = 2.0 * U + 2.0 ** -53 / q
    dL = 4.0 * U * L + dq / np.log(2.0)
    dt = t * (0.5 * (dL / L + 2.0 * U) + 4.0 * U)
    err = dxdt * dt + 24.0 * U * R + U * np.abs(x)
    return np.where(q < 2.87e-7, U * np.abs(x), err)


def _kink(a, e):
    """sup over |w - a| <= e of |sqrt(max(w, 0)) - sqrt(max(a, 0))|."""
    ap = np.maximum(a, 0.0)
    hi = np.maximum(a + e, 0.0)
    lo = np.maximum(a - e, 0.0)
    sa, sh, sl = np.sqrt(ap), np.sqrt(hi), np.sqrt(lo)
    with np.errstate(invalid="ignore", divide="ignore"):
        up = np.where(sh + sa > 0, (hi - ap) / (sh + sa), 0.0)
        dn = np.where(sa + sl > 0, (ap - lo) / (sa + sl), 0.0)
    return np.maximum(up, dn)


def heston_rk_bound(y, times, x0=(100.0, 0.04), r=0.02, kappa=2.0, theta=0.04, c1=-0.7 * 0.5,
                    c2=0.7141428428542850 * 0.5, states=None, perturb=None, draws="xor"):
    """FP64 paths and FP32 error bounds of the Heston RK scheme with the reference's cache.

    y: the uniforms of each path, [S][N * 2] (draw or dimension 2 n + k): with draws="xor" the
    XOR-scrambled Sobol words (uint32; u = (y + 0.5) 2^-32), with draws="words" the f64 uniforms
    of the pseudo-random stream, whose FP32 inverse normal is formed from the draw's two words.
    Model: dS = r S dt + sqrt(v+) S dW1, dv = kappa (theta - v+) dt + c1 sqrt(v+) dW1
    + c2 sqrt(v+) dW2 (v+ = max(v, 0)). Returns (values, bound), each [S][T][2] f64.

    states=None: the FP64 paths from x0 and a bound on |FP32 path - FP64 path| propagated over
    all steps (absolute values of every sensitivity, so it compounds with the number of steps).

    states=[S][T][2] (the FP32 kernel's filtration, as f64): one step at a time. Row n + 1 of
    the result is the FP64 step applied to the kernel's own row n (exact inputs), with stage 1 on
    the probe row rebuilt in f64 from the kernel's rows, and the bound is what one FP32 step
    adds: the step's roundings, the increments' and the difference between the kernel's f32
    probe and the rebuilt one. That difference follows its own recursion, which contracts (the
    probe depends on the previous probe only through the coefficients), so nothing compounds.

    `perturb` (tests of the bound only) replaces the scheme: "textbook" evaluates stage 1 on the
    settled row instead of the previous probe.
    """
    y = np.asarray(y, dtype=np.uint32 if draws == "xor" else np.float64)
    S = y.shape[0]
    times = np.asarray(times, dtype=np.float64)
    N = len(times) - 1
    X = np.empty((S, N + 1, 2))
    E = np.empty((S, N + 1, 2))
    s = np.full(S, float(x0[0]))
    v = np.full(S, float(x0[1]))
    es = np.abs(np.float64(np.float32(x0[0])) - x0[0]) * np.ones(S)
    ev = np.abs(np.float64(np.float32(x0[1])) - x0[1]) * np.ones(S)
    cs, cv, ecs, ecv = s.copy(), v.copy(), es.copy(), ev.copy()
    X[:, 0, 0], X[:, 0, 1], E[:, 0, 0], E[:, 0, 1] = s, v, es, ev
    a = np.abs

    def coeffs(ps, pv, eps_, epv):
        rt = np.sqrt(np.maximum(pv, 0.0))
        ert = _kink(pv, epv) + 4.0 * U * rt
        k1 = theta - np.maximum(pv, 0.0)
        ek1 = epv + U * (theta + np.maximum(pv, 0.0)) + U * a(k1)
        k2 = rt * ps
        ek2 = ert * a(ps) + rt * eps_ + ert * eps_ + U * a(k2)
        return ps, eps_, k1, ek1, k2, ek2, rt, ert

    for n in range(N):
        if states is not None:
            s = np.asarray(states[:, n, 0], dtype=np.float64)
            v = np.asarray(states[:, n, 1], dtype=np.float64)
            es = np.zeros(S)
            ev = np.zeros(S)
        dt = times[n + 1] - times[n]
        sdt = np.sqrt(dt)
        y1, y2 = y[:, 2 * n], y[:, 2 * n + 1]
        if draws == "xor":
            u1 = (y1.astype(np.float64) + 0.5) * 2.0 ** -32
            u2 = (y2.astype(np.float64) + 0.5) * 2.0 ** -32
            e1, e2 = C_PSI * psi_error(y1), C_PSI * psi_error(y2)
        else:
            u1, u2 = y1, y2
            e1, e2 = psi_error_words(u1), psi_error_words(u2)
        dw1 = sdt * inv_normal(u1)
        dw2 = sdt * inv_normal(u2)
        edw1 = sdt * e1 + 4.0 * U * a(dw1)
        edw2 = sdt * e2 + 6.0 * U * a(dw2)  # (c2 sqrt(dt)) times the increment in f32
        sk = np.where(u1 > 0.5, 1.0, -1.0)
        adt, kdt = r * dt, kappa * dt
        g = c1 * dw1 + c2 * dw2
        eg = a(c1) * edw1 + a(c2) * edw2 + U * a(c1 * dw1) + 2.0 * U * (a(c1 * dw1) + a(c2 * dw2))

        if perturb == "textbook":
            cs, cv = s, v
        K0, EK0, K1, EK1, K2, EK2, K3, EK3 = coeffs(cs, cv, ecs, ecv)
        k1s = K0 * adt + K2 * dw1
        ek1s = (EK0 * adt + a(K0) * 3.0 * U * adt + EK2 * a(dw1) + a(K2) * edw1 + EK2 * edw1
                + 2.0 * U * (a(K0 * adt) + a(K2 * dw1)))
        k1v = K1 * kdt + K3 * g
        ek1v = (EK1 * kdt + a(K1) * U * kdt + EK3 * a(g) + K3 * eg + EK3 * eg
                + 2.0 * U * (a(K1 * kdt) + a(K3 * g)))
        # Probe: X + k1 + sk sqrt(dt) (sum of the Wiener coefficients).
        pe_s = K2 * sk * sdt
        epe_s = EK2 * sdt + a(K2) * U * sdt + U * a(pe_s)
        pe_v = (c1 * K3) * sk * sdt + (c2 * K3) * sk * sdt
        epe_v = EK3 * a(c1 + c2) * sdt + K3 * a(c1 + c2) * sdt * 3.0 * U + 4e-16 * (a(c1) + a(c2)) * K3 * sdt \
            + U * a(pe_v)
        ps = s + k1s + pe_s
        eps_ = es + ek1s + epe_s + 2.0 * U * (a(s) + a(k1s) + a(pe_s))
        pv = v + k1v + pe_v
        epv = ev + ek1v + epe_v + 2.0 * U * (a(v) + a(k1v) + a(pe_v))
        # Stage 2 on the probe.
        Q0, EQ0, Q1, EQ1, Q2, EQ2, Q3, EQ3 = coeffs(ps, pv, eps_, epv)
        k2s = Q0 * adt + Q2 * dw1
        ek2s = (EQ0 * adt + a(Q0) * 3.0 * U * adt + EQ2 * a(dw1) + a(Q2) * edw1 + EQ2 * edw1
                + 2.0 * U * (a(Q0 * adt) + a(Q2 * dw1)))
        k2v = Q1 * kdt + Q3 * g
        ek2v = (EQ1 * kdt + a(Q1) * U * kdt + EQ3 * a(g) + Q3 * eg + EQ3 * eg
                + 2.0 * U * (a(Q1 * kdt) + a(Q3 * g)))
        s_new = s + 0.5 * (k1s + k2s)
        v_new = v + 0.5 * (k1v + k2v)
        es = es + 0.5 * (ek1s + ek2s) + U * (a(k1s) + a(k2s)) + U * a(s_new)
        ev = ev + 0.5 * (ek1v + ek2v) + U * (a(k1v) + a(k2v)) + U * a(v_new)
        s, v = s_new, v_new
        cs, cv, ecs, ecv = ps, pv, eps_, epv
        X[:, n + 1, 0], X[:, n + 1, 1] = s, v
        E[:, n + 1, 0], E[:, n + 1, 1] = es, ev
    return X, E


def emulate_f32(y, times, x0=(100.0, 0.04), r=0.02, kappa=2.0, theta=0.04, c1=-0.7 * 0.5,
                c2=0.7141428428542850 * 0.5, draws="xor"):
    """The same Heston RK recursion evaluated in numpy float32 (each operation rounded; the
    increments from the rescaled rational in t' = sqrt(-log2 q) and the f64-tabulated tail),
    [S][T][2] as float64. An independent FP32 evaluation for checking the bound on the CPU."""
    f = np.float32
    y = np.asarray(y, dtype=np.uint32 if draws == "xor" else np.float64)
    S = y.shape[0]
    times = np.asarray(times, dtype=np.float64)
    N = len(times) - 1
    out = np.empty((S, N + 1, 2))
    s = np.full(S, f(x0[0]), dtype=f)
    v = np.full(S, f(x0[1]), dtype=f)
    out[:, 0, 0], out[:, 0, 1] = s, v

    def xprime(yw):
        yc = ~yw
        q = np.abs(yc.view(np.int32).astype(f) * f(2.0 ** -32) + f(2.0 ** -33))
        tp = np.sqrt(-np.log2(q))
        num = (f(A2) * tp + f(A1)) * tp + f(A0)
        den = ((f(B3) * tp + f(B2)) * tp + f(B1)) * tp + f(1.0)
        xp = tp - num * (f(1.0) / den)
        qi = np.minimum(yw, ~yw)
        tail = qi < 1233
        if tail.any():
            qq = (qi[tail].astype(np.float64) + 0.5) * 2.0 ** -32
            xp[tail] = f(-inv_normal(qq) / C_PSI)
        return xp, yw < np.uint32(1 << 31)  # x' >= 0 and whether u < 1/2

    def zwords(uw):
        # |z| and whether u < 1/2, from q rounded to f32 (the words form's q' up to 2^-53).
        q = np.where(uw < 0.5, uw, 1.0 - uw)
        q32 = q.astype(f)
        t = np.sqrt(f(1.3862943611198906) * -np.log2(q32))
        num = (f(0.010328) * t + f(0.802853)) * t + f(2.515517)
        den = ((f(0.001308) * t + f(0.189269)) * t + f(1.432788)) * t + f(1.0)
        x = t - num * (f(1.0) / den)
        tail = q < 2.87e-7
        if tail.any():
            x[tail] = f(-inv_normal(q[tail]))
        return x, uw < 0.5

    cs, cv = s.copy(), v.copy()
    for n in range(N):
        dt = f(times[n + 1] - times[n])
        sdt = f(np.sqrt(times[n + 1] - times[n]))
        if draws == "xor":
            xp1, neg1 = xprime(y[:, 2 * n])
            xp2, neg2 = xprime(y[:, 2 * n + 1])
            m1 = sdt * f(C_PSI)
            m2 = (f(c2) * sdt) * f(C_PSI)
        else:
            xp1, neg1 = zwords(y[:, 2 * n])
            xp2, neg2 = zwords(y[:, 2 * n + 1])
            m1 = sdt
            m2 = f(c2) * sdt
        w0 = np.where(neg1, -m1, m1) * xp1
        w1 = np.where(neg2, -m2, m2) * xp2
        sks = np.where(neg1, -sdt, sdt) if draws == "xor" else np.where(y[:, 2 * n] > 0.5, sdt, -sdt)
        g0 = f(c1) * w0 + w1

        def co(ps, pv):
            rt = np.sqrt(np.maximum(pv, f(0.0)))
            return ps, f(theta) - np.maximum(pv, f(0.0)), rt * ps, rt

        K0, K1, K2, K3 = co(cs, cv)
        k1s = K0 * (f(r) * dt) + K2 * w0
        k1v = K1 * (f(kappa) * dt) + K3 * g0
        ps = s + k1s + K2 * sks
        pv = v + k1v + K3 * (f(c1 + c2) * sks)
        Q0, Q1, Q2, Q3 = co(ps, pv)
        k2s = Q0 * (f(r) * dt) + Q2 * w0
        k2v = Q1 * (f(kappa) * dt) + Q3 * g0
        s = s + f(0.5) * (k1s + k2s)
        v = v + f(0.5) * (k1v + k2v)
        cs, cv = ps, pv
        out[:, n + 1, 0], out[:, n + 1, 1] = s, v
    return out
