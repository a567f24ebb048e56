"""Pin the oracle's step (FP32 mirror and FP64) with closed forms and
invariants, chosen so that a dropped term, a wrong sign or a swapped operand
fails one of them."""
import math

import numpy as np
import pytest

import oracle


def single_robot(pos, pairs=(), lbar=(), mat=(), mass=None):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    pairs = np.asarray(pairs, np.uint16).reshape(-1, 2)
    pop = oracle.Population([0, n], [0, len(pairs)], np.full(n, 1.0, np.float32) if mass is None
                            else np.asarray(mass, np.float32), pairs,
                            np.asarray(lbar, np.float32), np.asarray(mat, np.uint8))
    return pop, pos


def flat(*arrs):
    return tuple(None if a is None else np.asarray(a).reshape(-1) for a in arrs)


def table(rows):
    """rows: composite-id indexed (k, dL0, omega, phi); row 0 must be air."""
    r = np.asarray(rows, np.float32)
    return oracle.MaterialTable(r[:, 0], r[:, 1], r[:, 2], r[:, 3])


AIR = (0, 0, 0, 0)


def test_free_fall_is_exact_symplectic_euler():
    pop, pos = single_robot([[0.0, 1.0, 0.0]], mass=[0.001])
    mt = table([AIR])
    dt, g, n = 1e-3, 9.81, 200
    env = oracle.make_env(dt, gravity=g, damping=1.0)
    p64, v64 = flat(*oracle.force_f64(pop, mt, env, 0.0, n, pos, np.zeros_like(pos)))
    gd = np.float64(np.float32(g)) * np.float64(np.float32(dt)) ** 2
    assert p64[1] == pytest.approx(1.0 - gd * n * (n + 1) / 2, rel=1e-12)
    assert v64[1] == pytest.approx(-np.float64(np.float32(g)) * np.float64(np.float32(dt)) * n,
                                   rel=1e-12)
    p32, _ = flat(*oracle.force_f32(pop, mt, env, 0.0, n, pos, np.zeros_like(pos)))
    assert p32[1] == pytest.approx(p64[1], rel=1e-5)
    assert p32[0] == 0.0 and p32[2] == 0.0


def test_xpbd_single_spring_lambda():
    """One XPBD step of a stretched spring between two resting masses."""
    d0, L0, k, dt = 0.12, 0.1, 50.0, 1e-3
    pop, pos = single_robot([[0, 0, 0], [d0, 0, 0]], pairs=[[0, 1]], lbar=[L0], mat=[1])
    mt = table([AIR, (k, 0, 0, 0)])
    p, v, st = flat(*oracle.xpbd_f64(pop, mt, dt, 0.0, 1, pos, np.zeros_like(pos), np.zeros(1)))
    kf, dtf, d0f, L0f = (np.float64(np.float32(x)) for x in (k, dt, d0, L0))
    K = 2.0 + (1.0 / kf) / dtf / dtf
    lam = -(d0f - L0f) / K
    n_x = -d0f / (d0f + np.float64(np.float32(1e-12)))  # (xa - xb)/(d + eps), xa = 0
    assert p[0] == pytest.approx(lam * n_x, rel=1e-12)            # +dp on a
    assert p[3] == pytest.approx(d0f - lam * n_x, rel=1e-12)      # -dp on b
    assert st[0] == pytest.approx(lam / L0f, rel=1e-12)
    assert v[0] == pytest.approx(0.99 * lam * n_x / dtf, rel=1e-12)
    # the spring pulls the masses together
    assert p[0] > 0 and p[3] < d0f
    p32, v32, st32 = flat(*oracle.xpbd_f32(pop, mt, dt, 0.0, 1, pos, np.zeros_like(pos),
                                           np.zeros(1, np.float32)))
    assert p32[0] == pytest.approx(p[0], rel=1e-6)


def test_actuation_rest_length():
    """L = Lbar (1 + A sin(omega t + phi)) enters lambda at the step's time."""
    L0, A, w, ph, k, dt, t0 = 0.1, 0.14, 4 * math.pi, 0.3, 50.0, 1e-3, 0.37
    pop, pos = single_robot([[0, 0, 0], [L0, 0, 0]], pairs=[[0, 1]], lbar=[L0], mat=[1])
    mt = table([AIR, (k, A, w, ph)])
    p, _, _ = flat(*oracle.xpbd_f64(pop, mt, dt, t0, 1, pos, np.zeros_like(pos)))
    f = lambda x: np.float64(np.float32(x))  # noqa: E731
    L = f(L0) * (1 + f(A) * math.sin(f(w) * f(t0) + f(ph)))
    K = 2.0 + (1.0 / f(k)) / f(dt) / f(dt)
    lam = -(f(L0) - L) / K
    assert p[0] == pytest.approx(lam * -f(L0) / (f(L0) + f(1e-12)), rel=1e-9)


def test_air_spring_is_a_noop():
    pop, pos = single_robot([[0, 0, 0], [0.2, 0, 0]], pairs=[[0, 1]], lbar=[0.1], mat=[0])
    mt = table([AIR])
    p, v, _ = oracle.xpbd_f32(pop, mt, 1e-3, 0.0, 50, pos, np.zeros_like(pos))
    assert np.array_equal(p.reshape(-1, 3), pos)
    env = oracle.make_env(1e-3, damping=1.0)
    p, _ = oracle.force_f32(pop, mt, env, 0.0, 50, pos, np.zeros_like(pos))
    assert np.array_equal(p.reshape(-1, 3), pos)


def test_xpbd_momentum_invariant():
    rng = np.random.default_rng(0)
    n = 12
    pos = rng.random((n, 3)).astype(np.float32)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.4]
    lbar = [np.linalg.norm(pos[i] - pos[j]) * rng.uniform(0.8, 1.2) for i, j in pairs]
    pop, pos = single_robot(pos, pairs, lbar, [1] * len(pairs))
    mt = table([AIR, (50, 0.14, 2 * math.pi, 0.5)])
    p, _, _ = oracle.xpbd_f64(pop, mt, 5e-3, 0.0, 300, pos, np.zeros_like(pos))
    assert np.allclose(p.reshape(-1, 3).mean(0), pos.astype(np.float64).mean(0), atol=1e-12)
    assert np.abs(p.reshape(-1, 3) - pos).max() > 1e-3  # something moved


def ground_pair(stretch, k_spring):
    m, g, kg = 0.001, 9.81, 1e5
    y0 = -np.float32(m) * np.float32(g) / np.float32(kg)   # resting penetration
    L0 = 0.1
    pos = [[0.0, y0, 0.0], [L0 + stretch, y0, 0.0]]
    pop, pos = single_robot(pos, pairs=[[0, 1]], lbar=[L0], mat=[1], mass=[m, m])
    mt = table([AIR, (k_spring, 0, 0, 0)])
    env = oracle.make_env(1e-4, gravity=g, ground_k=kg, mu_s=1.0, mu_k=0.8, damping=1.0)
    return pop, pos, mt, env


def test_static_friction_holds():
    # spring force 10 * 1e-4 = 1e-3 N < mu_s * m g = 9.81e-3 N
    pop, pos, mt, env = ground_pair(1e-4, 10.0)
    p, v = oracle.force_f64(pop, mt, env, 0.0, 500, pos, np.zeros_like(pos))
    p = p.reshape(-1, 3)
    assert p[0, 0] == pos[0, 0] and p[1, 0] == pos[1, 0]
    assert np.all(v.reshape(-1, 3)[:, 0] == 0)


def test_kinetic_friction_first_step():
    # spring force 1000 * 1e-3 = 1 N > mu_s m g: kinetic, |F_h| reduced by mu_k F_n
    pop, pos, mt, env = ground_pair(1e-3, 1000.0)
    p, _ = flat(*oracle.force_f64(pop, mt, env, 0.0, 1, pos, np.zeros_like(pos)))
    f = lambda x: np.float64(np.float32(x))  # noqa: E731
    d = f(pos[1, 0]) - f(pos[0, 0])
    fs = f(1000.0) * (d - f(0.1)) * d / (d + f(1e-12))  # tension on a, toward +x
    fn = f(1e5) * -f(pos[0, 1])
    acc = (fs - f(0.8) * fn) / f(0.001)
    assert p[0] - f(pos[0, 0]) == pytest.approx(acc * f(1e-4) ** 2, rel=1e-6)
    assert p[3] - f(pos[1, 0]) == pytest.approx(-acc * f(1e-4) ** 2, rel=1e-6)


def test_linear_drag_decay():
    pop, pos = single_robot([[0, 0, 0]], mass=[0.001])
    mt = table([AIR])
    c, m, dt, n = 1.0, 0.001, 1e-4, 300
    env = oracle.make_env(dt, drag=c, damping=1.0)
    v0 = np.array([[0.5, -0.2, 0.1]], np.float32)
    _, v = flat(*oracle.force_f64(pop, mt, env, 0.0, n, pos, v0))
    f = lambda x: np.float64(np.float32(x))  # noqa: E731
    factor = (1 - f(c) * f(dt) / f(m)) ** n
    assert np.allclose(v, v0.astype(np.float64).reshape(-1) * factor, rtol=1e-10)


def test_two_mass_oscillation_frequency():
    k, m, L0, dt = 100.0, 0.01, 0.1, 1e-4
    pop, pos = single_robot([[0, 0, 0], [L0 * 1.05, 0, 0]], pairs=[[0, 1]], lbar=[L0], mat=[1],
                            mass=[m, m])
    mt = table([AIR, (k, 0, 0, 0)])
    env = oracle.make_env(dt, damping=1.0)
    xs = []
    p, v = pos.astype(np.float64), np.zeros((2, 3))
    for _ in range(400):
        p, v = oracle.force_f64(pop, mt, env, 0.0, 10, p, v)
        p, v = p.reshape(-1, 3), v.reshape(-1, 3)
        xs.append(p[1, 0] - p[0, 0] - L0)
    xs = np.array(xs)
    crossings = np.where(np.diff(np.sign(xs)) != 0)[0]
    period = 2 * np.mean(np.diff(crossings)) * 10 * dt
    assert period == pytest.approx(2 * math.pi / math.sqrt(2 * k / m), rel=0.02)
    # symplectic Euler keeps the oscillation amplitude (no drift in energy)
    assert np.abs(xs).max() == pytest.approx(0.05 * L0, rel=0.05)


def fma32(a, b, c):
    """Correctly rounded float32 fused multiply-add, by exact rationals."""
    from fractions import Fraction
    exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    x = np.float32(float(exact))
    best = None
    for cand in (np.nextafter(x, np.float32(-np.inf)), x, np.nextafter(x, np.float32(np.inf))):
        err = abs(Fraction(float(cand)) - exact)
        even = (int(np.float32(cand).view(np.uint32)) & 1) == 0
        if best is None or err < best[0] or (err == best[0] and even):
            best = (err, np.float32(cand))
    return best[1]


def numpy_xpbd_step_f32(pos, vel, pairs, lbar, mat, K, R, dt):
    """Independent float32 restatement of one XPBD step with an explicit
    per-mass gather over its springs in ascending index (the kernels'
    order) instead of the oracle's spring-order scatter."""
    f = np.float32
    xt = (pos + vel * f(dt)).astype(np.float32)
    n = len(pos)
    terms = {}
    for s, (a, b) in enumerate(pairs):
        if mat[s] == 0:
            continue
        L = fma32(lbar[s], R[mat[s]], lbar[s])
        d3 = (xt[a] - xt[b]).astype(np.float32)
        ss = fma32(d3[1], d3[1], f(d3[0] * d3[0]))
        d = f(np.sqrt(fma32(d3[2], d3[2], ss)))
        den = f(d + f(1e-12))
        nrm = (d3 / den).astype(np.float32)
        lam = f(f(-(d - L)) / K[mat[s]])
        terms[s] = (lam * nrm).astype(np.float32)
    out_p, out_v = pos.copy(), vel.copy()
    for i in range(n):
        acc = np.zeros(3, np.float32)
        for s, (a, b) in enumerate(pairs):      # ascending spring index
            if s in terms and a == i:           # first endpoint: + term
                acc = (acc + terms[s]).astype(np.float32)
            if s in terms and b == i:           # second endpoint: - term
                acc = (acc - terms[s]).astype(np.float32)
        xn = (xt[i] + acc).astype(np.float32)
        out_v[i] = ((0.99 * (xn - pos[i]).astype(np.float64)) / np.float64(f(dt))).astype(np.float32)
        out_p[i] = xn
    return out_p, out_v


def test_mirror_matches_independent_gather_restatement():
    rng = np.random.default_rng(3)
    n = 9
    pos = (rng.random((n, 3)) * 0.3).astype(np.float32)
    vel = (rng.normal(0, 0.05, (n, 3))).astype(np.float32)
    pairs = [(int(a), int(b)) for a, b in rng.integers(0, n, (20, 2))]
    lbar = (rng.random(20) * 0.2 + 0.05).astype(np.float32)
    mat = rng.integers(0, 3, 20).astype(np.uint8)
    rows = [AIR, (50, 0.14, 6.2831855, 0.0), (100, 0.0, 6.2831855, 0.0)]
    mt = table(rows)
    dt, t0 = np.float32(1e-3), np.float32(0.25)
    K = np.array([1.0] + [2.0 + ((1.0 / np.float32(r[0])) / dt) / dt for r in rows[1:]],
                 np.float32)
    R = np.array([np.float32(r[1]) * oracle.sinf(fma32(np.float32(r[2]), t0, np.float32(r[3])))
                  for r in rows], np.float32)
    pop, _ = single_robot(pos, pairs, lbar, mat)
    p, v, _ = oracle.xpbd_f32(pop, mt, float(dt), float(t0), 1, pos, vel)
    ep, ev = numpy_xpbd_step_f32(pos, vel, pairs, lbar, mat, K, R, dt)
    assert np.array_equal(p.reshape(-1, 3).view(np.uint32), ep.view(np.uint32))
    assert np.array_equal(v.reshape(-1, 3).view(np.uint32), ev.view(np.uint32))


def test_step_times_match_float_loop():
    for t0, dt, dur in ((0.0, 1e-4, 1.0), (0.37, 1e-3, 0.05), (0.0, 5e-3, 10.0)):
        ts = oracle.step_times(t0, dt, dur)
        r, t, exp = np.float32(dur), np.float32(t0), []
        while r > np.float32(0):
            exp.append(t)
            t = np.float32(t + np.float32(dt))
            r = np.float32(r - np.float32(dt))
        assert np.array_equal(ts, np.array(exp, np.float32))


def numpy_force_step_f32(pos, vel, mass, pairs, lbar, mat, k, R, env):
    """Independent float32 restatement of one force-solver step: per-mass
    gather of the f * delta terms as fused multiply-adds, gravity, drag,
    ground contact and friction."""
    f = np.float32
    dt, g, kg, mus, muk, c, damp = (f(env[key]) for key in
                                    ("dt", "gravity", "ground_k", "mu_s", "mu_k", "drag", "damping"))
    terms = {}
    for s, (a, b) in enumerate(pairs):
        if mat[s] == 0:
            continue
        L = fma32(lbar[s], R[mat[s]], lbar[s])
        d3 = (pos[a] - pos[b]).astype(np.float32)
        ss = fma32(d3[1], d3[1], f(d3[0] * d3[0]))
        d = f(np.sqrt(fma32(d3[2], d3[2], ss)))
        fs = f(f(k[mat[s]] * f(L - d)) * f(f(1) / f(d + f(1e-12))))
        terms[s] = (fs, d3)
    out_p, out_v = pos.copy(), vel.copy()
    for i in range(len(pos)):
        F = np.zeros(3, np.float32)
        for s, (a, b) in enumerate(pairs):      # ascending spring index
            if s in terms and a == i:           # first endpoint: fma(+f, delta, sum)
                F = np.array([fma32(terms[s][0], terms[s][1][c], F[c]) for c in range(3)], np.float32)
            if s in terms and b == i:           # second endpoint: fma(-f, delta, sum)
                F = np.array([fma32(-terms[s][0], terms[s][1][c], F[c]) for c in range(3)], np.float32)
        v = vel[i].copy()
        F[1] = f(F[1] - f(mass[i] * g))
        if c != 0:
            F = (F - (c * v).astype(np.float32)).astype(np.float32)
        y = pos[i, 1]
        if kg != 0 and y < 0:
            fn = f(kg * f(-y))
            F[1] = f(F[1] + fn)
            fh = f(np.sqrt(f(f(F[0] * F[0]) + f(F[2] * F[2]))))
            if fh <= f(mus * fn):
                F[0] = F[2] = v[0] = v[2] = f(0)
            else:
                sc = f(f(1) - f(f(muk * fn) / fh))
                F[0], F[2] = f(F[0] * sc), f(F[2] * sc)
        inv = f(f(1) / mass[i])
        u = (v + ((F * inv).astype(np.float32) * dt).astype(np.float32)).astype(np.float32)
        out_p[i] = (pos[i] + (u * dt).astype(np.float32)).astype(np.float32)
        out_v[i] = (damp * u).astype(np.float32)
    return out_p, out_v


@pytest.mark.parametrize("drag", [0.0, 1.0])
def test_force_mirror_matches_independent_restatement(drag):
    rng = np.random.default_rng(11)
    n = 10
    pos = (rng.random((n, 3)) * 0.3).astype(np.float32)
    pos[:4, 1] = -rng.random(4).astype(np.float32) * 1e-4   # four masses in ground contact
    vel = (rng.normal(0, 0.05, (n, 3))).astype(np.float32)
    pairs = [(int(a), int(b)) for a, b in rng.integers(0, n, (24, 2)) if a != b]
    m = len(pairs)
    lbar = (rng.random(m) * 0.2 + 0.05).astype(np.float32)
    mat = rng.integers(0, 3, m).astype(np.uint8)
    rows = [AIR, (5e3, 0.14, 12.566371, 0.0), (1e4, 0.0, 12.566371, 0.0)]
    mt = table(rows)
    env = dict(dt=1e-4, gravity=9.81, ground_k=1e5, mu_s=1.0, mu_k=0.8, drag=drag, damping=0.99)
    t0 = np.float32(0.3)
    k = np.array([0.0] + [r[0] for r in rows[1:]], np.float32)
    R = np.array([np.float32(r[1]) * oracle.sinf(fma32(np.float32(r[2]), t0, np.float32(r[3])))
                  for r in rows], np.float32)
    mass = np.full(n, 1e-3, np.float32)
    pop, _ = single_robot(pos, pairs, lbar, mat, mass=mass)
    e = oracle.make_env(**env)
    p, v = oracle.force_f32(pop, mt, e, float(t0), 1, pos, vel)
    ep, ev = numpy_force_step_f32(pos, vel, mass, pairs, lbar, mat, k, R, env)
    assert np.array_equal(np.asarray(p).reshape(-1, 3).view(np.uint32), ep.view(np.uint32))
    assert np.array_equal(np.asarray(v).reshape(-1, 3).view(np.uint32), ev.view(np.uint32))


def test_fp64_one_second_is_sensitive_to_one_ulp():
    """Why FP32 and FP64 displacement rankings over a 1 s land evaluation
    cannot be identical: the FP64 step itself, with one coordinate of one
    mass per robot moved by one float ulp (about 3e-9 m), ends up to
    0.46 mm away and orders robot pairs differently (64 robots of 8^3
    voxels, 10000 steps, measured 2 pairs)."""
    import os
    from helpers import LAND, workload_case
    case = workload_case(n_robots=64, voxels=8, p_empty=0.2, seed=0)
    env = oracle.make_env(1e-4, **LAND)
    mo, mm = case.a["mass_offset"], case.a["mass_material"]
    th = max(1, min(16, len(os.sched_getaffinity(0))))

    def disp(p):
        p = p.reshape(-1, 3)
        out = []
        for r in range(len(mo) - 1):
            sl = slice(mo[r], mo[r + 1])
            air = mm[sl] == 0
            out.append(p[sl][~air].mean(0)[0] - case.pos0[sl][~air].astype(np.float64).mean(0)[0])
        return np.array(out)

    pos = case.pos0.copy()
    rng = np.random.default_rng(0)
    for r in range(len(mo) - 1):
        i = mo[r] + rng.integers(0, mo[r + 1] - mo[r])
        pos[i, 0] = np.nextafter(pos[i, 0], np.float32(1))
    a = disp(oracle.force_f64(case.pop, case.mt, env, 0.0, 10000, case.pos0, case.vel0, nthreads=th)[0])
    b = disp(oracle.force_f64(case.pop, case.mt, env, 0.0, 10000, pos, case.vel0, nthreads=th)[0])
    i, j = np.triu_indices(len(a), 1)
    assert np.abs(a - b).max() > 1e-4
    assert (np.sign(a[i] - a[j]) != np.sign(b[i] - b[j])).any()
