"""Synthetic tuning problems of the benchmark configurations, in the reference's file format.

  c1  mass-spring-damper (m=1, c=0.4, k=4), 2 states, forward Euler dt=1e-3, clamp(u,-2,2) as a U
      SAT instruction, 1024 candidates K in [0.1,100]^2, L in [1,1000]^2 1/s (log-uniform, applied as
      dt*L in the discrete observer), seed 0, T=2000, F=200.
  c2  two-mass flexible drive (m1=1, m2=0.5, k=50 N/m, c1=0.05, c2=0.2 N s/m), 4 states, dt=5e-4,
      Coulomb friction 0.3 sign(v1) and drag 0.02 v2|v2| as NL chains, LQR/LQR-dual seeds, K and L
      log-uniform over +-2 decades around them (one scale per candidate, +-0.3 decade elementwise jitter), 65536 candidates, seed 1, T=5000, F=400.
  c3  random stable 8-state plant (poles Re in [-20,-0.5], Im in [0,30], conjugate pairs), ZOH at
      dt=1e-3, Kalman gains (Q=I, R=0.01) as the observer seed, four nonlinearities (saturation,
      sign, x1*x3, x2/(1+|x2|)), NL/U/DC lengths 8/4/4, 262144 candidates, seed 2, T=4000, F=512.

Every generator is deterministic in its seed.  N = 1/DC gain is computed in double and stored as
float32; A_I (= A - I) is formed in double before rounding, not as float(A) - 1.
"""
import numpy as np

from .problem import Problem, encode, info_width

OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_SAT, OP_SGN = 1, 2, 3, 4, 5, 6
STATE, PREV = 0, 35

SPEC = dict(damping=0.7, overshoot=0.1, steady_state_error=0.02, crossover_coeff=1.0, target_gm=6.0,
            target_pm=45.0, target_b=0.2, eig_fail_err=1e18, unstable_err=1e6)


class Params:
    """BASE_PARAMS builder: const(v) returns the base-36 digit (1..34) of value v."""

    def __init__(self):
        self.vals = []

    def const(self, v):
        v = float(np.float32(v))
        for i, x in enumerate(self.vals):
            if x == v:
                return i + 1
        if len(self.vals) >= 34:
            raise ValueError("more than 34 BASE_PARAMS")
        self.vals.append(v)
        return len(self.vals)

    def array(self):
        return np.array(self.vals if self.vals else [0.0], dtype=np.float32)


def row(width, col, op, a, b):
    r = np.zeros(width, dtype=np.int32)
    r[col] = encode(op, a, b)
    return r


def dc_gain_N(A, B, C, K):
    n = A.shape[0]
    M = np.eye(n) - A + np.outer(B, K)
    g = C @ np.linalg.solve(M, B)
    return 1.0 / g


def around(rng, seed_gain, n, decades=2.0, jitter=0.3):
    """n gain vectors log-uniform within +-decades of seed_gain: one scale per candidate, times an
    elementwise +-jitter decade perturbation (independent per-element scaling over the full range
    leaves most loops unstable, which would make the workload mostly trivially rejected rows)."""
    k = seed_gain.size
    scale = 10.0 ** rng.uniform(-decades, decades, size=(n, 1))
    jit = 10.0 ** rng.uniform(-jitter, jitter, size=(n, k))
    return seed_gain[None, :] * scale * jit


def dlqr(A, B, Q, R):
    from scipy.linalg import solve_discrete_are
    B2 = B.reshape(-1, 1)
    P = solve_discrete_are(A, B2, Q, np.atleast_2d(R))
    K = np.linalg.solve(np.atleast_2d(R) + B2.T @ P @ B2, B2.T @ P @ A)
    return K.ravel()


def kalman_predictor(A, C, Q, R):
    from scipy.linalg import solve_discrete_are
    C2 = C.reshape(1, -1)
    P = solve_discrete_are(A.T, C2.T, Q, np.atleast_2d(R))
    L = A @ P @ C2.T @ np.linalg.inv(C2 @ P @ C2.T + np.atleast_2d(R))
    return L.ravel()


def _assemble(name, cfg, ns, nd, A, A_I, BU, BW, C, params, dc, u, nl, K, L, seed_meta, rounds=1,
              pad_to_threads=False):
    A32 = np.float32(A)
    gains = np.concatenate([K, L], axis=1)
    # N = 1 / (C (I - A + B K^T)^-1 B) for every candidate, as one batched solve
    ns_ = A.shape[0]
    M = (np.eye(ns_) - A)[None] + BU[None, :, None] * K[:, None, :]
    N = 1.0 / (np.linalg.solve(M, np.broadcast_to(BU, (K.shape[0], ns_))[..., None])[..., 0] @ C)
    gains = np.concatenate([gains, N[:, None]], axis=1).astype(np.float32)
    n = gains.shape[0]
    if pad_to_threads:
        nthr = (int(cfg["n_cores"]) // 128) * 1536
        reps = -(-n // nthr) * nthr
        gains = np.concatenate([gains, gains[: reps - n]], axis=0) if reps > n else gains
    info = np.zeros((gains.shape[0], info_width(cfg)), dtype=np.float32)
    nw = int(cfg["nonlin_dim"])
    return Problem(
        config=cfg,
        ss_a=A32.reshape(ns, ns).copy(), ss_a_i=np.float32(A_I).reshape(ns, ns).copy(),
        ss_b=np.float32(BU).copy(), ss_c=np.float32(C).copy(), ss_d=np.zeros(1, np.float32),
        d_a=A32.reshape(1, nd, nd).copy(), d_a_i=np.float32(A_I).reshape(1, nd, nd).copy(),
        d_bu=np.float32(BU).reshape(1, nd).copy(), d_bw=np.float32(BW).reshape(1, nd, nw).copy(),
        d_c=np.float32(C).reshape(1, nd).copy(), base_params=params.array().reshape(1, -1),
        dc_info=np.array(dc, dtype=np.int32).reshape(-1, nd), u_info=np.array(u, dtype=np.int32).reshape(-1, ns),
        nl_info=np.array(nl, dtype=np.int32).reshape(-1, nd + 1) if len(nl) else np.zeros((0, nd + 1), np.int32),
        gains=gains, info=info, name=name, meta=seed_meta)


def c1(n=1024, seed=0, pad_to_threads=False, T=2000, F=200):
    rng = np.random.default_rng(seed)
    m, c, k, dt = 1.0, 0.4, 4.0, 1e-3
    Ac = np.array([[0.0, 1.0], [-k / m, -c / m]])
    Bc = np.array([0.0, 1.0 / m])
    C = np.array([1.0, 0.0])
    A_I = dt * Ac
    A = np.eye(2) + A_I
    BU = dt * Bc
    P = Params()
    lo, hi, one = P.const(-2.0), P.const(2.0), P.const(1.0)
    u = [row(2, 0, OP_SAT, lo, hi)]                       # u = clamp(u_raw, -2, 2)
    dc = [row(2, 0, OP_MUL, STATE, one)]                  # y = x0 * 1
    K = 10.0 ** rng.uniform(-1.0, 2.0, size=(n, 2))
    # observer gains drawn as continuous-time gains in [1, 1000] 1/s, applied as dt*L in the
    # discrete observer x^+ = x + A_I x + B u + L (y - yhat)
    L = dt * 10.0 ** rng.uniform(0.0, 3.0, size=(n, 2))
    duration = T * dt
    cfg = dict(statespace_dim=2, discrete_dim=2, nonlin_dim=0, n_params=len(P.vals), input_size=0,
               dc_instructions=1, u_instructions=1, nl_instructions=[], n_cores=128, n_rounds=1, mode="batch",
               time=T, duration=duration, ref_step_time=0.05, ref_amplitude=1.0, fit_start=0.05, fit_end=0.75,
               steady_start=0.75, bandwidth=0.5, freq=F, freq_min=1e-2, dump_diagnostics=False, **SPEC)
    return _assemble("c1", cfg, 2, 2, A, A_I, BU, np.zeros((2, 0)), C, P, dc, u, [], K, L,
                     dict(seed=seed), pad_to_threads=pad_to_threads)


def c2(n=65536, seed=1, pad_to_threads=False, T=5000, F=400):
    rng = np.random.default_rng(seed)
    m1, m2, ks, c1d, c2d, dt = 1.0, 0.5, 50.0, 0.05, 0.2, 5e-4
    # states x1, v1, x2, v2; c1d motor-side viscous damping, c2d damper across the spring
    Ac = np.array([
        [0.0, 1.0, 0.0, 0.0],
        [-ks / m1, -(c1d + c2d) / m1, ks / m1, c2d / m1],
        [0.0, 0.0, 0.0, 1.0],
        [ks / m2, c2d / m2, -ks / m2, -c2d / m2],
    ])
    Bc = np.array([0.0, 1.0 / m1, 0.0, 0.0])
    C = np.array([1.0, 0.0, 0.0, 0.0])
    A_I = dt * Ac
    A = np.eye(4) + A_I
    BU = dt * Bc
    BW = dt * np.array([[0.0, 0.0], [-1.0 / m1, 0.0], [0.0, 0.0], [0.0, -1.0 / m2]])
    P = Params()
    one, cc, cd = P.const(1.0), P.const(0.3), P.const(0.02)
    u = [row(4, 0, OP_MUL, PREV, one)]                    # u = u_raw
    dc = [row(4, 0, OP_MUL, STATE, one)]                  # y = x1 (motor position)
    nl = [row(5, 1, OP_SGN, STATE, STATE),                # W0 = 0.3 sign(v1)
          row(5, 0, OP_MUL, PREV, cc),
          row(5, 3, OP_SGN, STATE, STATE),                # W1 = 0.02 v2 |v2|
          row(5, 3, OP_MUL, PREV, STATE),
          row(5, 3, OP_MUL, PREV, STATE),
          row(5, 0, OP_MUL, PREV, cd)]
    K0 = dlqr(A, BU, np.eye(4), 1.0)
    L0 = kalman_predictor(A, C, np.eye(4), 1.0)
    K = around(rng, K0, n)
    L = around(rng, L0, n)
    cfg = dict(statespace_dim=4, discrete_dim=4, nonlin_dim=2, n_params=len(P.vals), input_size=0,
               dc_instructions=1, u_instructions=1, nl_instructions=[2, 4], n_cores=128, n_rounds=1, mode="batch",
               time=T, duration=T * dt, ref_step_time=0.02, ref_amplitude=1.0, fit_start=0.02, fit_end=0.75,
               steady_start=0.75, bandwidth=1.0, freq=F, freq_min=1e-2, dump_diagnostics=False, **SPEC)
    return _assemble("c2", cfg, 4, 4, A, A_I, BU, BW, C, P, dc, u, nl, K, L, dict(seed=seed, K0=K0, L0=L0),
                     pad_to_threads=pad_to_threads)


def random_stable_plant(rng, n, dt, re_range=(-20.0, -0.5), im_range=(0.0, 30.0)):
    """n-state (n even) continuous plant with conjugate-pair poles, ZOH-discretised."""
    from scipy.linalg import expm
    J = np.zeros((n, n))
    for p in range(n // 2):
        s = rng.uniform(*re_range)
        w = rng.uniform(*im_range)
        J[2 * p:2 * p + 2, 2 * p:2 * p + 2] = [[s, w], [-w, s]]
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Ac = V @ J @ V.T
    Bc = rng.standard_normal(n)
    M = np.zeros((n + 1, n + 1))
    M[:n, :n] = Ac * dt
    M[:n, n] = Bc * dt
    E = expm(M)
    A = E[:n, :n]
    B = E[:n, n]
    # A - I from the series of expm(Ac dt) - I, not from the rounded A
    A_I = A - np.eye(n)
    return Ac, A, A_I, B


def c3(n=262144, seed=2, pad_to_threads=False, T=4000, F=512):
    rng = np.random.default_rng(seed)
    ns, dt = 8, 1e-3
    Ac, A, A_I, BU = random_stable_plant(rng, ns, dt)
    C = np.array([0.5, 0.0, 0.5, 0.0, 0.5, 0.0, 0.0, 0.0])
    BW = dt * 0.5 * rng.standard_normal((ns, 4))
    P = Params()
    one, zero, mone, half = P.const(1.0), P.const(0.0), P.const(-1.0), P.const(0.5)
    umax, umin, ten = P.const(50.0), P.const(-50.0), P.const(10.0)
    # U chain (4): u = clamp(u_raw + 0, -50, 50) * 1 + 0
    u = [row(ns, 0, OP_ADD, PREV, zero), row(ns, 0, OP_SAT, umin, umax), row(ns, 0, OP_MUL, PREV, one),
         row(ns, 0, OP_ADD, PREV, zero)]
    # DC chain (4): y = 0.5 (x0 + x2 + x4)
    dc = [row(ns, 0, OP_MUL, STATE, one), row(ns, 2, OP_ADD, PREV, STATE), row(ns, 4, OP_ADD, PREV, STATE),
          row(ns, 0, OP_MUL, PREV, half)]
    w = ns + 1
    nl = [row(w, 0, OP_SAT, mone, one),                   # W0 = clamp(x0, -1, 1)
          row(w, 1, OP_SGN, STATE, STATE),                # W1 = sign(x1)
          row(w, 1, OP_ADD, STATE, zero),                 # W2 = x1 * x3
          row(w, 3, OP_MUL, PREV, STATE),
          row(w, 2, OP_SGN, STATE, STATE),                # W3 = x2 / (1 + |x2|)
          row(w, 2, OP_MUL, PREV, STATE),
          row(w, 0, OP_ADD, PREV, one),
          row(w, 2, OP_DIV, STATE, PREV)]
    del ten
    K0 = dlqr(A, BU, np.eye(ns), 1.0)
    L0 = kalman_predictor(A, C, np.eye(ns), 0.01)
    K = around(rng, K0, n)
    L = around(rng, L0, n)
    cfg = dict(statespace_dim=ns, discrete_dim=ns, nonlin_dim=4, n_params=len(P.vals), input_size=0,
               dc_instructions=4, u_instructions=4, nl_instructions=[1, 1, 2, 4], n_cores=128, n_rounds=1,
               mode="batch", time=T, duration=T * dt, ref_step_time=0.025, ref_amplitude=1.0, fit_start=0.025,
               fit_end=0.75, steady_start=0.75, bandwidth=2.0, freq=F, freq_min=1e-2, dump_diagnostics=False,
               **SPEC)
    return _assemble("c3", cfg, ns, ns, A, A_I, BU, BW, C, P, dc, u, nl, K, L, dict(seed=seed, K0=K0, L0=L0),
                     pad_to_threads=pad_to_threads)


def chain_drive(masses, springs, dampers, c_motor, dt):
    """Chain of len(masses) masses (states x1, v1, x2, v2, ...): spring/damper i joins mass i and
    i+1, motor damping on mass 1, force input on mass 1, output x1; forward Euler at dt.
    Returns Ac, A, A_I, BU, BW (BW: Coulomb friction on v1, drag on the last mass's velocity)."""
    m = len(masses)
    ns = 2 * m
    Ac = np.zeros((ns, ns))
    for i in range(m):
        Ac[2 * i, 2 * i + 1] = 1.0
    Ac[1, 1] -= c_motor / masses[0]
    for i in range(m - 1):
        k, c = springs[i], dampers[i]
        a, b = i, i + 1
        for (p_, q_, sgn) in ((a, b, 1.0), (b, a, 1.0)):
            Ac[2 * p_ + 1, 2 * p_] -= sgn * k / masses[p_]
            Ac[2 * p_ + 1, 2 * q_] += sgn * k / masses[p_]
            Ac[2 * p_ + 1, 2 * p_ + 1] -= sgn * c / masses[p_]
            Ac[2 * p_ + 1, 2 * q_ + 1] += sgn * c / masses[p_]
    Bc = np.zeros(ns)
    Bc[1] = 1.0 / masses[0]
    BWc = np.zeros((ns, 2))
    BWc[1, 0] = -1.0 / masses[0]
    BWc[2 * m - 1, 1] = -1.0 / masses[-1]
    A_I = dt * Ac
    return Ac, np.eye(ns) + A_I, A_I, dt * Bc, dt * BWc


class C4:
    """Batched uneven Monte Carlo: n_plants flexible-drive chains of 1-4 masses (2, 4, 6 or 8
    states, as the two-mass drive of c2: Coulomb friction on the motor mass and quadratic drag on
    the load mass as NL chains), masses / springs / dampers drawn around the c2 values.  Each plant
    is one serial-mode problem over `rounds` rounds: every round perturbs the physical parameters
    and the NL constants (BASE_PARAMS) by a uniform +-10 % and re-discretises the plant
    (DISCRETE_*); the observer model stays nominal.  Candidates per plant log-uniform in
    [nmin, nmax] around LQR / LQR-dual seeds (as c2), horizons uniform in [tmin, tmax] steps,
    dt = 5e-4, F = 400."""

    def __init__(self, seed=3, n_plants=512, rounds=64, nmin=256, nmax=16384, tmin=1000, tmax=8000, F=400,
                 dt=5e-4):
        rng = np.random.default_rng(seed)
        self.rounds = rounds
        self.plants = []
        for j in range(n_plants):
            m = int(rng.integers(1, 5))
            masses = np.concatenate([[1.0], rng.uniform(0.3, 1.0, m - 1)])
            springs = rng.uniform(30.0, 70.0, m - 1)
            dampers = rng.uniform(0.1, 0.3, m - 1)
            c_motor = 0.05 if m > 1 else 0.4
            n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
            T = int(rng.integers(tmin, tmax + 1))
            # nominal model: the observer, and the seeds of the candidate gains
            Ac, A, A_I, BU, BW = chain_drive(masses, springs, dampers, c_motor, dt)
            ns = 2 * m
            C = np.zeros(ns)
            C[0] = 1.0
            P = Params()
            one, cc, cd = P.const(1.0), P.const(0.3), P.const(0.02)
            u = [row(ns, 0, OP_MUL, PREV, one)]
            dc = [row(ns, 0, OP_MUL, STATE, one)]
            vl = 2 * m - 1
            nl = [row(ns + 1, 1, OP_SGN, STATE, STATE), row(ns + 1, 0, OP_MUL, PREV, cc),
                  row(ns + 1, vl, OP_SGN, STATE, STATE), row(ns + 1, vl, OP_MUL, PREV, STATE),
                  row(ns + 1, vl, OP_MUL, PREV, STATE), row(ns + 1, 0, OP_MUL, PREV, cd)]
            K0 = dlqr(A, BU, np.eye(ns), 1.0)
            L0 = kalman_predictor(A, C, np.eye(ns), 1.0)
            K = around(rng, K0, n)
            L = around(rng, L0, n)
            cfg = dict(statespace_dim=ns, discrete_dim=ns, nonlin_dim=2, n_params=len(P.vals), input_size=0,
                       dc_instructions=1, u_instructions=1, nl_instructions=[2, 4], n_cores=128, n_rounds=rounds,
                       mode="serial", time=T, duration=T * dt, ref_step_time=0.02, ref_amplitude=1.0,
                       fit_start=0.02, fit_end=0.75, steady_start=0.75, bandwidth=1.0, freq=F, freq_min=1e-2,
                       dump_diagnostics=False, **SPEC)
            prob = _assemble("c4_%d" % j, cfg, ns, ns, A, A_I, BU, BW, C, P, dc, u, nl, K, L,
                             dict(seed=seed, plant=j))
            # per-round perturbed plants (serial-mode slices) and BASE_PARAMS
            da, dai, dbu, dbw, dcs, bps = [], [], [], [], [], []
            base = P.array()
            for r in range(rounds):
                f = 1.0 + rng.uniform(-0.1, 0.1, 2 * m + 2)
                _, Ar, Ar_I, BUr, BWr = chain_drive(masses * f[:m], springs * f[m:2 * m - 1] if m > 1 else springs,
                                                     dampers * f[m:2 * m - 1] if m > 1 else dampers,
                                                     c_motor * f[2 * m - 1], dt)
                da.append(np.float32(Ar))
                dai.append(np.float32(Ar_I))
                dbu.append(np.float32(BUr))
                dbw.append(np.float32(BWr))
                dcs.append(np.float32(C))
                bp = base.copy()
                bp[1] = np.float32(0.3 * f[2 * m])
                bp[2] = np.float32(0.02 * f[2 * m + 1])
                bps.append(bp)
            prob.d_a, prob.d_a_i = np.stack(da), np.stack(dai)
            prob.d_bu, prob.d_bw, prob.d_c = np.stack(dbu), np.stack(dbw), np.stack(dcs)
            prob.base_params = np.stack(bps)
            self.plants.append(prob)

    def candidates(self):
        return sum(p.gains.shape[0] for p in self.plants)


def c5_job(rng, nmin=256, nmax=65536, tmin=500, tmax=5000, fmin=128, fmax=512, n=None):
    """One serving job: a random stable plant as in c3 with dimension in 2..8 (conjugate pairs plus
    one real pole for odd dimensions), 0-4 nonlinear terms chosen from saturation, sign, product
    and x/(1+|x|), candidates log-uniform in [nmin, nmax] around LQR / Kalman seeds, horizon
    uniform in [tmin, tmax] steps at dt = 1e-3, F uniform in [fmin, fmax]."""
    from scipy.linalg import expm
    ns = int(rng.integers(2, 9))
    dt = 1e-3
    J = np.zeros((ns, ns))
    for p in range(ns // 2):
        sr, w = rng.uniform(-20.0, -0.5), rng.uniform(0.0, 30.0)
        J[2 * p:2 * p + 2, 2 * p:2 * p + 2] = [[sr, w], [-w, sr]]
    if ns % 2:
        J[ns - 1, ns - 1] = rng.uniform(-20.0, -0.5)
    V, _ = np.linalg.qr(rng.standard_normal((ns, ns)))
    Ac = V @ J @ V.T
    Bc = rng.standard_normal(ns)
    M = np.zeros((ns + 1, ns + 1))
    M[:ns, :ns] = Ac * dt
    M[:ns, ns] = Bc * dt
    E = expm(M)
    A, BU = E[:ns, :ns], E[:ns, ns]
    A_I = A - np.eye(ns)
    C = np.zeros(ns)
    C[0] = 1.0
    nw = int(rng.integers(0, 5))
    P = Params()
    one, zero, mone = P.const(1.0), P.const(0.0), P.const(-1.0)
    umin, umax = P.const(-50.0), P.const(50.0)
    u = [row(ns, 0, OP_SAT, umin, umax)]
    dc = [row(ns, 0, OP_MUL, STATE, one)]
    w = ns + 1
    last = ns - 1
    chains = [[row(w, 0, OP_SAT, mone, one)],                                          # clamp(x0, -1, 1)
              [row(w, min(1, last), OP_SGN, STATE, STATE)],                            # sign(x1)
              [row(w, 0, OP_ADD, STATE, zero), row(w, last, OP_MUL, PREV, STATE)],     # x0 * x_last
              [row(w, last, OP_SGN, STATE, STATE), row(w, last, OP_MUL, PREV, STATE),  # x/(1+|x|)
               row(w, 0, OP_ADD, PREV, one), row(w, last, OP_DIV, STATE, PREV)]]
    pick = sorted(rng.choice(4, size=nw, replace=False)) if nw else []
    nl = [r for q in pick for r in chains[q]]
    BW = dt * 0.5 * rng.standard_normal((ns, nw))
    K0 = dlqr(A, BU, np.eye(ns), 1.0)
    L0 = kalman_predictor(A, C, np.eye(ns), 0.01)
    n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))) if n is None else n
    K = around(rng, K0, n)
    L = around(rng, L0, n)
    T = int(rng.integers(tmin, tmax + 1))
    F = int(rng.integers(fmin, fmax + 1))
    cfg = dict(statespace_dim=ns, discrete_dim=ns, nonlin_dim=nw, n_params=len(P.vals), input_size=0,
               dc_instructions=1, u_instructions=1, nl_instructions=[len(chains[q]) for q in pick], n_cores=128,
               n_rounds=1, mode="batch", time=T, duration=T * dt, ref_step_time=0.025, ref_amplitude=1.0,
               fit_start=0.025, fit_end=0.75, steady_start=0.75, bandwidth=2.0, freq=F, freq_min=1e-2,
               dump_diagnostics=False, **SPEC)
    return _assemble("c5", cfg, ns, ns, A, A_I, BU, BW, C, P, dc, u, nl, K, L, dict(ns=ns, nw=nw))


CONFIGS = {"c1": c1, "c2": c2, "c3": c3}
