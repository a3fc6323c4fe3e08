"""Independent pins of the oracle cells that the reference's text alone would leave unchecked.

Every expected value below comes from a computation that shares no code with the oracle: numpy's
LAPACK eigenvalues of a LIN_A built here from the block formula, exact trajectories iterated in
numpy / mpmath, exact gain and phase crossovers by root finding on the analytic loop, and the
correctly rounded float32 arithmetic of mpmath.  The reference's scoring formulas are the
specification (src/modalsolve.cu:88-106, src/allmargin.cu:119-180); they are applied here to those
independent inputs.

  grid          src/main.cu:421-428, src/allmargin.cu:111      float32 derivation vs mpmath
  LIN_A         src/eigensetup.cu:16-56                        all four blocks, ns != nd, D != 0
  modes, dedupe src/modalsetup.cu:8-47                         wn, zeta, unique count vs numpy eig
  yss, overshoot src/modalsetup.cu:52-62                       settled closed-loop DC value, exact peak
  GM, DM, PM    src/allmargin.cu:126-151                       brentq crossings of the analytic loop
  behaviour     src/modalsolve.cu:88-106                       single-mode aggregate, hinges
  margin, total src/allmargin.cu:167-180                       active hinges on all three margins
  branches      src/modalsetup.cu:26, src/allmargin.cu:46-48,80-87,161-165, src/modalsolve.cu:106
"""
import math

import mpmath as mp
import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import oracle
from problems import OP_MUL, PREV, STATE, f32, general_problem, row

OP_DIV = 4
EPS30 = float(np.float32(np.float32(1.2958e-30)))
EPS6 = float(np.float32(1e-6))
FLTMAX = float(np.finfo(np.float32).max)


def cells(p, o64, i):
    d = p.derived()
    r = o64[i]
    nt = d.total_dim
    return dict(sc=r[d.info_scoring:d.info_scoring + 3], fr=r[d.info_freqresp:d.info_freqresp + 4],
                am=r[d.info_allmargin:d.info_allmargin + 5], md=r[d.info_modal:d.info_modal + 3 * nt + 1])


def mats(p):
    """The plant and observer values the engine reads (float32, promoted)."""
    ns, nd = p.ns, p.nd
    return dict(SA=f32(p.ss_a).reshape(ns, ns), SAI=f32(p.ss_a_i).reshape(ns, ns), SB=f32(p.ss_b), SC=f32(p.ss_c),
                SD=float(f32(p.ss_d)[0]), DA=f32(p.d_a[0]).reshape(nd, nd), DAI=f32(p.d_a_i[0]).reshape(nd, nd),
                DBU=f32(p.d_bu[0]), DC=f32(p.d_c[0]))


def lina_block(m, K, L):
    """LIN_A = [[SA - SB K' - L (SC - SD K)', L DC'], [-DBU K', DA]] (src/eigensetup.cu:16-56)."""
    top = np.hstack([m["SA"] - np.outer(m["SB"], K) - np.outer(L, m["SC"] - m["SD"] * K), np.outer(L, m["DC"])])
    bot = np.hstack([-np.outer(m["DBU"], K), m["DA"]])
    return np.vstack([top, bot])


def closed_loop(m, K, L, N):
    """Linear closed loop of the simulation, z = [xh; xp], z' = M z + b r (src/discrete_solves.cu:30-78):
    u = N r - K xh, y = DC xp, xh' = xh + SAI xh + SB u + L (y - SC xh - SD u), xp' = xp + DAI xp + DBU u."""
    ns, nd = len(K), m["DA"].shape[0]
    # u = N r - K xh; yhat = SC xh + SD u
    Mh = np.eye(ns) + m["SAI"] - np.outer(m["SB"], K) - np.outer(L, m["SC"] - m["SD"] * K)
    Mhp = np.outer(L, m["DC"])
    Mp = np.eye(nd) + m["DAI"]
    Mph = -np.outer(m["DBU"], K)
    M = np.block([[Mh, Mhp], [Mph, Mp]])
    b = np.concatenate([m["SB"] * N - L * m["SD"] * N, m["DBU"] * N])
    c = np.concatenate([np.zeros(ns), m["DC"]])
    return M, b, c


def trajectory(M, b, c, T, k0, amp):
    z = np.zeros(M.shape[0])
    y = np.empty(T)
    for t in range(T):
        y[t] = c @ z
        z = M @ z + b * (amp if t >= k0 else 0.0)
    return y


def step_index(p):
    T = int(p.config["time"])
    return next(t for t in range(T) if not t < float(p.config["ref_step_time"]) * T)


# ------------------------------------------------------------------------------------------------
def test_grid_is_the_float32_derivation():
    """z_i = (cosf, sinf)(expf(logf(fmin) + inc * i) * ts), inc = (logf(fmax) - logf(fmin)) / (F - 1),
    every step rounded to float32 (src/main.cu:423-427); theta_w = atan2f of the float point."""
    from problems import linear_problem
    for F, T, dur in [(200, 2000, 2.0), (401, 5000, 2.5), (37, 300, 0.3)]:
        p = linear_problem([[-0.02]], [0.05], [1.0], [[2.0, 0.3, 1.0]], T=T, F=F, dt=dur / T)
        d = p.derived()
        g = oracle.grid(p)
        mp.mp.prec = 120
        r = lambda v: np.float32(float(v))            # correctly rounded to double, then to float32
        ts = np.float32(d.ts)
        fmin, fmax = np.float32(p.config["freq_min"]), np.float32(d.freq_max)
        lmin, lmax = r(mp.log(mp.mpf(float(fmin)))), r(mp.log(mp.mpf(float(fmax))))
        inc = np.float32(np.float32(lmax - lmin) / np.float32(F - 1))
        exact = 0
        for w in range(F):
            th = np.float32(r(mp.exp(mp.mpf(float(np.float32(lmin + np.float32(inc * np.float32(w))))))) * ts)
            zr, zi = r(mp.cos(mp.mpf(float(th)))), r(mp.sin(mp.mpf(float(th))))
            tw = r(mp.atan2(mp.mpf(float(zi)), mp.mpf(float(zr))))
            got = (np.float32(g["z"][w].real), np.float32(g["z"][w].imag), np.float32(g["theta"][w]))
            same = (got[0] == zr, got[1] == zi, got[2] == tw)
            exact += all(same)
            for a, e in zip(got, (zr, zi, tw)):
                assert abs(float(a) - float(e)) <= float(np.spacing(np.float32(abs(e)))), (F, w)
            # the promoted values are exactly the float32 values
            assert g["z"][w].real == float(got[0]) and g["z"][w].imag == float(got[1])
        assert exact >= F - 1, (F, exact)


def lina_problem():
    rng = np.random.default_rng(5)
    ns, nd = 3, 2
    sa_i = rng.uniform(-0.05, 0.05, (ns, ns))
    sb, sc = rng.uniform(-1, 1, ns), rng.uniform(-1, 1, ns)
    da_i = rng.uniform(-0.05, 0.05, (nd, nd))
    dbu = rng.uniform(-1, 1, nd)
    gains = rng.uniform(-0.5, 0.5, (12, 2 * ns + 1))
    return general_problem(sa_i, sb, sc, 0.3, da_i, dbu, 1, gains, T=120, F=16)


def test_lina_all_blocks_and_eigenvalues():
    """LIN_A with ns != nd, nonzero D and random blocks, against the block formula; the eigenvalue
    dump against LAPACK's eigenvalues of that matrix (unordered)."""
    ns, nd = 3, 2
    p = lina_problem()
    _, _, dg = oracle.evaluate(p, diag=True)
    m = mats(p)
    nt = ns + nd
    g = f32(p.gains)
    for i in range(g.shape[0]):
        K, L = g[i, :ns], g[i, ns:2 * ns]
        M = lina_block(m, K, L)
        np.testing.assert_allclose(dg["lina"][i].reshape(nt, nt), M, rtol=0, atol=4e-16 * max(1.0, np.abs(M).max()))
        want = np.sort_complex(np.linalg.eigvals(M))
        e = dg["eig"][i]
        got = np.sort_complex(e[0::2] + 1j * e[1::2])
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


# A one-state observer model and a one-state plant whose linearised loop has one complex pair:
# observer a = 0.999, b = 0.005, c = 1; plant da = 0.999, dbu = 0.02 (differs from the observer's b).
OBS = dict(sa_i=[[-0.001]], sb=[0.005], sc=[1.0], sd=0.0, da_i=[[-0.001]], dbu=[0.02], dc_col=0)


def loop_tf(m, K, L, zs):
    """Analytic loop P(z) G(z) of the margins sweep for ns = 1 (batch mode: the observer model is
    the plant): P = SC SB / (z - SA) + SD, G = K L / (z - A1), A1 = LIN_A[0, 0]."""
    a1 = lina_block(m, K, L)[0, 0]
    P = m["SC"][0] * m["SB"][0] / (zs - m["SA"][0, 0]) + m["SD"]
    G = K[0] * L[0] / (zs - a1)
    return P, G


def exact_margins(m, K, L, ts, theta):
    """GM (dB) at the phase crossover, PM (deg) and delay margin (ms) at the gain crossover, by
    brentq on the analytic loop over the grid's angle range."""
    f = lambda t: loop_tf(m, K, L, np.exp(1j * t))
    lp = lambda t: f(t)[0] * f(t)[1]
    # the top grid point is Nyquist: its float angle lies just beyond pi, so its arg is near -pi
    grid = np.linspace(theta[0], np.pi, 200001)
    vals = lp(grid)
    mag = np.abs(vals)
    ph = np.unwrap(np.angle(vals))
    out = {}
    gi = np.where(np.diff(np.sign(mag - 1.0)) != 0)[0]
    if gi.size:
        tc = brentq(lambda t: abs(lp(t)) - 1.0, grid[gi[0]], grid[gi[0] + 1], xtol=1e-15)
        phc = np.interp(tc, grid, ph)
        pm = 180.0 + np.degrees(phc)
        pm -= 360.0 * math.floor((pm + 180.0) / 360.0)
        out.update(pm=pm, dm=np.radians(pm) * ts / tc * 1000.0, tc=tc)
    pi_ = np.where(np.diff(np.floor((ph + np.pi) / (2 * np.pi))) != 0)[0]
    if pi_.size:
        k = max(np.floor((ph[pi_[0]] + np.pi) / (2 * np.pi)), np.floor((ph[pi_[0] + 1] + np.pi) / (2 * np.pi)))
        target = 2 * np.pi * k - np.pi
        tp = brentq(lambda t: np.interp(t, grid, ph) - target, grid[pi_[0]], grid[pi_[0] + 1], xtol=1e-15)
        # refine on the analytic phase near the bracket
        tp = brentq(lambda t: np.angle(lp(t) * np.exp(-1j * target)), tp - 1e-4, tp + 1e-4, xtol=1e-15)
        out.update(gm=-20.0 * np.log10(abs(lp(tp))), tp=tp)
    return out


def single_mode_problem(spec=None, amplitude=1.0, n_scale=1.05):
    m0 = general_problem(**OBS, gains=[[2.0, 0.03, 1.0]], T=2000, F=3000, spec=spec, amplitude=amplitude)
    m = mats(m0)
    K, L = f32([2.0]), f32([0.03])
    M, b, c = closed_loop(m, K, L, 1.0)
    dc_gain = c @ np.linalg.solve(np.eye(M.shape[0]) - M, b)
    N = n_scale / dc_gain
    return general_problem(**OBS, gains=[[2.0, 0.03, N]], T=2000, F=3000, spec=spec, amplitude=amplitude)


@pytest.mark.parametrize("amplitude", [1.0, -2.0])
def test_single_mode_behaviour_and_margins(amplitude):
    """One complex pair: the aggregate wn / zeta equal that mode's |s| and -Re s / |s| whatever the
    fitted energies; yss and overshoot from the exact trajectory and DC value; GM / PM / DM from the
    analytic loop; behaviour, margin error and total from those values with every hinge active."""
    spec = dict(target_gm=60.0, target_pm=120.0, target_b=0.9, overshoot=0.01, steady_state_error=0.02)
    p = single_mode_problem(spec, amplitude)
    d = p.derived()
    ts = float(d.ts)
    _, o64, dg = oracle.evaluate(p, diag=True)
    c_ = cells(p, o64, 0)
    m = mats(p)
    g = f32(p.gains)[0]
    K, L, N = g[:1], g[1:2], g[2]
    lam = np.linalg.eigvals(lina_block(m, K, L))
    assert abs(lam[0].imag) > 1e-3 and np.all(np.abs(lam) < 1)
    s = np.log(lam[0]) / ts
    wn, zeta = abs(s), -s.real / abs(s)
    # modes, dedupe of the conjugate pair, aggregate over the one oscillatory mode
    assert c_["md"][3 * 2] == 1.0                                     # md[3 nt] = unique modes
    assert abs(c_["md"][0] - wn) <= 1e-9 * wn and abs(c_["md"][1] - zeta) <= 1e-9
    assert abs(c_["fr"][0] - wn) <= 1e-9 * wn and abs(c_["fr"][1] - zeta) <= 1e-9
    # yss and overshoot: settled DC value of the closed loop and the exact peak
    M, b, cvec = closed_loop(m, K, L, N)
    T = int(p.config["time"])
    y = trajectory(M, b, cvec, T, step_index(p), float(f32(amplitude)))
    yfin = float(cvec @ np.linalg.solve(np.eye(2) - M, b)) * float(f32(amplitude))
    peak = y.max() if amplitude > 0 else y.min()
    os_exact = (peak - yfin) / yfin
    assert abs(c_["fr"][3] - yfin) <= 1e-9 * abs(yfin)
    assert abs(c_["fr"][2] - os_exact) <= 1e-8
    assert dg["y"][0][T - 1] == pytest.approx(yfin, rel=1e-9)
    # behaviour with the overshoot and steady-state hinges active
    dmp, bw = float(f32(0.7)), float(f32(1.0))
    twn = 2 * np.pi * bw / math.sqrt((1 - 2 * dmp ** 2) + math.sqrt(4 * dmp ** 4 - 4 * dmp ** 2 + 2))
    os_spec, sse = float(f32(0.01)), float(f32(0.02))
    h_os = max(0.0, os_exact - os_spec) / os_spec
    h_ss = max(0.0, abs(yfin / float(f32(amplitude)) - 1.0) - sse) / sse
    assert h_os > 0 and h_ss > 0
    beh = math.log(wn / twn) ** 2 + ((zeta - dmp) / dmp) ** 2 + h_os ** 2 + h_ss ** 2
    assert abs(c_["sc"][1] - beh) <= 1e-6 * beh
    # margins
    gr = oracle.grid(p)
    ex = exact_margins(m, K, L, ts, gr["theta"])
    assert "gm" in ex and "pm" in ex
    assert c_["am"][3] == 1.0
    assert abs(c_["am"][0] - ex["gm"]) <= 2e-4, (c_["am"][0], ex["gm"])
    assert abs(c_["am"][1] - ex["pm"]) <= 2e-4, (c_["am"][1], ex["pm"])
    assert abs(c_["am"][2] - ex["dm"]) <= 1e-5 * ex["dm"], (c_["am"][2], ex["dm"])
    P, G = loop_tf(m, K, L, gr["z"])
    # alpha = 1 / |P| at z_c = polar(1, crossover ts), the angle and its cos / sin in float32
    mp.mp.prec = 80
    ang = np.float32(np.float32(d.crossover) * np.float32(d.ts))
    zc = complex(float(np.float32(float(mp.cos(float(ang))))), float(np.float32(float(mp.sin(float(ang))))))
    alpha = 1.0 / abs(loop_tf(m, K, L, np.array([zc]))[0][0])
    assert abs(gr["alpha"] - alpha) <= 1e-13 * alpha
    bwv = np.abs(1 + P * G) / np.sqrt((1 + (alpha * np.abs(P)) ** 2) * (1 + (np.abs(G) / alpha) ** 2))
    assert abs(c_["am"][4] - min(1.0, bwv.min())) <= 1e-12
    tgm, tpm, tb = float(f32(60.0)), float(f32(120.0)), float(f32(0.9))
    e_gm, e_pm, e_b = (tgm - ex["gm"]) / tgm, (tpm - ex["pm"]) / tpm, (tb - bwv.min()) / tb
    assert e_gm > 0 and e_pm > 0 and e_b > 0
    margin = e_gm ** 2 + e_pm ** 2 + e_b ** 2
    assert abs(c_["sc"][2] - margin) <= 5e-5 * margin
    assert abs(c_["sc"][0] - math.log(beh + margin)) <= 5e-5


def test_inactive_hinges_and_flat_spec():
    """The same loop with every spec met: the hinges contribute nothing, so the behaviour is the
    wn / zeta terms alone and the margin error is zero."""
    spec = dict(target_gm=1.0, target_pm=10.0, target_b=0.01, overshoot=5.0, steady_state_error=0.5)
    p = single_mode_problem(spec, 1.0, n_scale=1.0)
    d = p.derived()
    _, o64, _ = oracle.evaluate(p)
    c_ = cells(p, o64, 0)
    m = mats(p)
    g = f32(p.gains)[0]
    s = np.log(np.linalg.eigvals(lina_block(m, g[:1], g[1:2]))[0]) / float(d.ts)
    wn, zeta = abs(s), -s.real / abs(s)
    dmp = float(f32(0.7))
    twn = 2 * np.pi * 1.0 / math.sqrt((1 - 2 * dmp ** 2) + math.sqrt(4 * dmp ** 4 - 4 * dmp ** 2 + 2))
    beh = math.log(wn / twn) ** 2 + ((zeta - dmp) / dmp) ** 2
    assert abs(c_["sc"][1] - beh) <= 1e-8 * beh
    assert c_["sc"][2] == 0.0
    assert abs(c_["sc"][0] - math.log(beh + EPS30)) <= 1e-12


def overshoot_problem(amp):
    """Mass-spring plant, state feedback wn 20 rad/s, zeta 0.3, fast observer, N 10 % off the DC gain."""
    dt = 1e-3
    Ac = np.array([[0.0, 1.0], [-4.0, -0.4]])
    K = [396.0, 11.6]
    L = [0.1, 2.5]
    p = general_problem(dt * Ac, dt * np.array([0.0, 1.0]), [1.0, 0.0], 0.0, dt * Ac, dt * np.array([0.0, 1.0]),
                        0, [K + L + [1.0]], T=6000, F=32, amplitude=amp)
    m = mats(p)
    M, b, c = closed_loop(m, f32(K), f32(L), 1.0)
    p.gains[0, 4] = np.float32(1.1 / float(c @ np.linalg.solve(np.eye(4) - M, b)))
    return p


def test_overshoot_and_yss_two_state():
    """Mass-spring plant with an underdamped state-feedback loop (wn 20 rad/s, zeta 0.3) and a fast
    observer, N 10 % off the DC gain: fr[3] is the settled closed-loop DC value and fr[2] the
    relative overshoot of the exact trajectory, for a positive and a negative step."""
    for amp in (1.0, -0.5):
        p = overshoot_problem(amp)
        m = mats(p)
        g = f32(p.gains)[0]
        M, b, c = closed_loop(m, g[:2], g[2:4], g[4])
        y = trajectory(M, b, c, 6000, step_index(p), float(f32(amp)))
        yfin = float(c @ np.linalg.solve(np.eye(4) - M, b)) * float(f32(amp))
        peak = y.max() if amp > 0 else y.min()
        _, o64, _ = oracle.evaluate(p)
        fr = cells(p, o64, 0)["fr"]
        assert abs(fr[3] - yfin) <= 1e-9 * abs(yfin)
        assert (peak - yfin) / yfin > 0.2
        assert abs(fr[2] - (peak - yfin) / yfin) <= 1e-8


# ------------------------------------------------------------------------------------------------
# branches
def unstable_problem():
    p0 = general_problem(**OBS, gains=[[-40.0, 0.03, 1.0]], T=400, F=32)
    info = np.full(p0.info.shape, 7.0, np.float32)
    info[:, p0.derived().info_scoring] = 0.0
    return general_problem(**OBS, gains=[[-40.0, 0.03, 1.0]], T=400, F=32, info=info)


def test_linearly_unstable_branch():
    """max |lambda| > 1: sc = UE + log(max |lambda|) / ts in all three cells, margins zeroed, and the
    fr / md cells keep their input values (src/modalsetup.cu:26, src/allmargin.cu:46-47)."""
    p = unstable_problem()
    d = p.derived()
    _, o64, _ = oracle.evaluate(p)
    c_ = cells(p, o64, 0)
    m = mats(p)
    g = f32(p.gains)[0]
    lam = np.abs(np.linalg.eigvals(lina_block(m, g[:1], g[1:2]))).max()
    assert lam > 1.0
    want = float(f32(1e6)) + math.log(lam) / float(d.ts)
    assert abs(c_["sc"][0] - want) <= 1e-12 * want
    assert c_["sc"][1] == c_["sc"][0] and c_["sc"][2] == c_["sc"][0]
    assert np.all(c_["am"] == 0.0)
    assert np.all(c_["fr"] == 7.0) and np.all(c_["md"] == 7.0)


def nominal_problem():
    obs = dict(OBS)
    obs.update(sa_i=[[0.01]])
    return general_problem(**obs, gains=[[0.0, 0.5, 0.05]], T=400, F=32)


def test_nominally_unstable_branch():
    """Stable linearised loop, unstable nominal loop (the observer model, which the batch-mode
    margins use as the plant, has a pole at 1.01 and K = 0): sc[2] = UE, margins zeroed, total =
    log(behaviour + UE) (src/allmargin.cu:80-87)."""
    p = nominal_problem()
    m = mats(p)
    g = f32(p.gains)[0]
    assert np.abs(np.linalg.eigvals(lina_block(m, g[:1], g[1:2]))).max() < 1.0
    acl = np.array([[lina_block(m, g[:1], g[1:2])[0, 0], -g[1] * m["SC"][0]], [m["SB"][0] * g[0], m["SA"][0, 0]]])
    assert np.abs(np.linalg.eigvals(acl)).max() > 1.0
    _, o64, _ = oracle.evaluate(p)
    c_ = cells(p, o64, 0)
    ue = float(f32(1e6))
    assert c_["sc"][2] == ue
    assert np.isfinite(c_["sc"][1]) and 0.0 < c_["sc"][1] < ue
    assert abs(c_["sc"][0] - math.log(c_["sc"][1] + ue)) <= 1e-13
    assert np.all(c_["am"] == 0.0)


def sweepfail_problem():
    p0 = general_problem([[-0.1, 0.0], [0.0, -0.1]], [0.0, 0.0], [0.0, 0.0], 1.0, [[-0.1, 0.0], [0.0, -0.1]],
                         [0.1, 0.1], 0, [[0.0, 0.0, 0.3, 0.2, 1.0]], T=400, F=64)
    z = oracle.grid(p0)["z"]
    w = next(w for w in range(20, 64) if abs(z[w]) < 1.0 - 1e-9)
    zr, zi = z[w].real, z[w].imag
    sa = np.array([[zr, -zi], [zi, zr]])
    p = general_problem(sa - np.eye(2), [0.0, 0.0], [0.0, 0.0], 1.0, [[-0.1, 0.0], [0.0, -0.1]], [0.1, 0.1], 0,
                        [[0.0, 0.0, 0.3, 0.2, 1.0]], T=400, F=64, sa=sa)
    return p, sa


def test_sweep_failure_branch():
    """The compensator (z I - A1) is exactly singular at one grid point: A1 = [[zr, -zi], [zi, zr]]
    with z = zr + j zi a float grid point, K = 0 and SC = 0 so that A1 = SA.  The sweep stops:
    sc[0] = sc[2] = eig_fail_err, am = [0, 0, 0, 1, 0] (src/allmargin.cu:161-165)."""
    p, sa = sweepfail_problem()
    assert np.array_equal(mats(p)["SA"], sa)
    _, o64, _ = oracle.evaluate(p)
    c_ = cells(p, o64, 0)
    ef = float(f32(1e18))
    assert c_["sc"][0] == ef and c_["sc"][2] == ef
    assert list(c_["am"]) == [0.0, 0.0, 0.0, 1.0, 0.0]


def nonfinite_problem():
    p = general_problem(**OBS, gains=[[2.0, 0.03, 1.0]], T=300, F=32)
    cfg = dict(p.config)
    cfg.update(nonlin_dim=1, nl_instructions=[1])
    p.config = cfg
    p.nl_info = np.array([row(2, 0, OP_DIV, 1, STATE)], np.int32)   # digit 1 = BASE_PARAMS[0] = 1.0
    p.d_bw = np.float32([[[0.01]]])
    return p


def test_nonfinite_behaviour_branch():
    """A plant nonlinearity 1 / x0 is infinite at t = 0, so the trajectory, yss and the fit are not
    finite while the linearised loop is stable: the behaviour is not finite and becomes
    eig_fail_err (src/modalsolve.cu:106); the margins then treat the row as rejected."""
    p = nonfinite_problem()
    assert p.base_params[0, 0] == 1.0
    _, o64, dg = oracle.evaluate(p, diag=True)
    assert not np.isfinite(dg["y"][0][-1])
    c_ = cells(p, o64, 0)
    ef = float(f32(1e18))
    assert c_["sc"][0] == ef and c_["sc"][1] == ef and c_["sc"][2] == ef
    assert np.all(c_["am"] == 0.0)


def realmode_problem():
    da_i = np.array([[-0.001, 0.005], [0.0, -0.02]])
    return general_problem([[-0.001]], [0.005], [1.0], 0.0, da_i, [0.02, 0.2], 0, [[2.0, 0.03, 1.0]], T=2000, F=64)


def test_real_mode_excluded_from_aggregate():
    """A two-state plant adds a real closed-loop mode that carries fit energy: the aggregate wn /
    zeta are still exactly those of the one oscillatory mode (src/modalsolve.cu:85: modes with
    Im < 1e-6 are left out), and the unique count is 2."""
    p = realmode_problem()
    d = p.derived()
    m = mats(p)
    g = f32(p.gains)[0]
    lam = np.linalg.eigvals(lina_block(m, g[:1], g[1:2]))
    assert np.all(np.abs(lam) < 1.0)
    osc = lam[np.argmax(lam.imag)]
    s = np.log(osc) / float(d.ts)
    wn, zeta = abs(s), -s.real / abs(s)
    _, o64, _ = oracle.evaluate(p)
    c_ = cells(p, o64, 0)
    nt = 3
    assert c_["md"][3 * nt] == 2.0
    energies = c_["md"][2 * nt:2 * nt + 2]
    real_e = [e for k, e in enumerate(energies) if c_["md"][2 * k + 1] == 1.0]
    assert len(real_e) == 1 and real_e[0] > 1e-3
    assert abs(c_["fr"][0] - wn) <= 1e-9 * wn and abs(c_["fr"][1] - zeta) <= 1e-9


def pin_cases():
    """Every problem of this file, for the GPU chain's parity test."""
    active = dict(target_gm=60.0, target_pm=120.0, target_b=0.9, overshoot=0.01, steady_state_error=0.02)
    met = dict(target_gm=1.0, target_pm=10.0, target_b=0.01, overshoot=5.0, steady_state_error=0.5)
    return [("lina", lina_problem()), ("single_mode", single_mode_problem(active, 1.0)),
            ("single_mode_neg", single_mode_problem(active, -2.0)), ("spec_met", single_mode_problem(met, 1.0, 1.0)),
            ("overshoot", overshoot_problem(1.0)), ("overshoot_neg", overshoot_problem(-0.5)),
            ("unstable", unstable_problem()), ("nominal_unstable", nominal_problem()),
            ("sweep_failure", sweepfail_problem()[0]), ("nonfinite", nonfinite_problem()),
            ("real_mode", realmode_problem())]
