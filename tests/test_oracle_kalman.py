"""The oracle's Kalman-filter simulator (reference src/discrete_solves.cu:83-185; gains rows
[K, N, R, Q], 2 ns + 2 floats) pinned by

  * a matrix-form textbook Kalman filter iterated in numpy with the same step order (predict the
    state and P, gain, update with the step's measurement, Joseph-form covariance update), on
    linear plants with pass-through U and DC chains;
  * the closed-loop DC value: on a noise-free linear plant a stable filter drives xhat to the plant
    state, so y settles at C (I - A + B K)^-1 B N amp whatever Q and R are."""
import numpy as np

from oracle import oracle
from problems import f32, general_problem


def mats(p):
    ns, nd = p.ns, p.nd
    return dict(SA=f32(p.ss_a).reshape(ns, ns), SAI=f32(p.ss_a_i).reshape(ns, ns), SB=f32(p.ss_b), SC=f32(p.ss_c),
                DAI=f32(p.d_a_i[0]).reshape(nd, nd), DBU=f32(p.d_bu[0]), DC=f32(p.d_c[0]))


def kalman_numpy(m, row, T, k0, amp):
    ns = m["SA"].shape[0]
    K, N, R, Q = row[:ns], row[ns], row[ns + 1], row[ns + 2:]
    xh = np.zeros(ns)
    xp = np.zeros(m["DAI"].shape[0])
    P = np.zeros((ns, ns))
    ys = np.empty(T)
    for t in range(T):
        y = m["DC"] @ xp
        ys[t] = y
        u = N * (amp if t >= k0 else 0.0) - K @ xh
        xh = xh + (m["SAI"] @ xh + m["SB"] * u)
        P = m["SA"] @ P @ m["SA"].T + np.diag(Q)
        kg = P @ m["SC"] / (m["SC"] @ P @ m["SC"] + R)
        xh = xh + kg * (y - m["SC"] @ xh)
        ikh = np.eye(ns) - np.outer(kg, m["SC"])
        P = ikh @ P @ ikh.T + R * np.outer(kg, kg)
        xp = xp + (m["DAI"] @ xp + m["DBU"] * u)
    return ys


def two_state(dt=1e-3):
    Ac = np.array([[0.0, 1.0], [-4.0, -0.4]])
    return dt * Ac, dt * np.array([0.0, 1.0])


def test_kalman_matches_textbook_filter():
    AI, B = two_state()
    p = general_problem(AI, B, [1.0, 0.0], 0.0, AI, B, 0, [[0.0] * 5], T=800, F=16)
    rng = np.random.default_rng(4)
    rows = np.float32(np.column_stack([rng.uniform(5, 60, (6, 2)), np.full(6, 4.0), 10 ** rng.uniform(-3, 0, 6),
                                       10 ** rng.uniform(-5, -1, (6, 2))]))
    y = oracle.simulate_kalman(p, rows)
    m = mats(p)
    k0 = next(t for t in range(800) if not t < 0.05 * 800)
    for i in range(rows.shape[0]):
        want = kalman_numpy(m, f32(rows[i]), 800, k0, 1.0)
        np.testing.assert_allclose(y[i], want, rtol=1e-10, atol=1e-13)


def test_kalman_three_by_two_observer_plant():
    """ns = 3 observer model over a 2-state plant (different matrices), dense gains."""
    rng = np.random.default_rng(8)
    sai = rng.uniform(-0.02, 0.02, (3, 3)) - 0.01 * np.eye(3)
    dai = rng.uniform(-0.02, 0.02, (2, 2)) - 0.01 * np.eye(2)
    p = general_problem(sai, rng.uniform(-0.1, 0.1, 3), rng.uniform(-1, 1, 3), 0.0, dai, rng.uniform(-0.1, 0.1, 2),
                        1, [[0.0] * 7], T=500, F=16)
    rows = np.float32(np.column_stack([rng.uniform(-2, 2, (5, 3)), np.full(5, 1.5), 10 ** rng.uniform(-2, 0, 5),
                                       10 ** rng.uniform(-4, -1, (5, 3))]))
    y = oracle.simulate_kalman(p, rows)
    m = mats(p)
    for i in range(5):
        np.testing.assert_allclose(y[i], kalman_numpy(m, f32(rows[i]), 500, 25, 1.0), rtol=1e-10, atol=1e-13)


def test_kalman_settles_at_the_state_feedback_dc_value():
    AI, B = two_state()
    T = 24000
    K = np.array([396.0, 11.6])
    p = general_problem(AI, B, [1.0, 0.0], 0.0, AI, B, 0, [[0.0] * 5], T=T, F=16, amplitude=2.0)
    m = mats(p)
    A = np.eye(2) + m["DAI"]
    g = m["DC"] @ np.linalg.solve(np.eye(2) - A + np.outer(m["DBU"], f32(K)), m["DBU"])
    for R, Q in [(1e-2, 1e-3), (1.0, 1e-5)]:
        row = np.float32(np.concatenate([K, [1.0 / g, R], [Q, Q]]))[None]
        y = oracle.simulate_kalman(p, row)[0]
        yfin = g * float(np.float32(1.0 / g)) * 2.0
        assert abs(y[-1] - yfin) <= 1e-9 * abs(yfin), (R, Q, y[-1], yfin)
