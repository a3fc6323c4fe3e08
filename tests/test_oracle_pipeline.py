"""The oracle's pipeline stages pinned by closed forms and independent formulations.

  simulation   (src/discrete_solves.cu:12-79): linear closed loop by matrix powers in mpmath;
               saturated loop against a plain Python step loop of the same equations.
  modal fit    (src/modalsetup.cu:64-96, src/modalsolve.cu:150-190): Gram and right-hand side from a
               vectorised numpy basis; coefficients vs numpy.linalg.lstsq; energies and wn/zeta.
  margins      (src/allmargin.cu:100-180): first-order plant and compensator, loop evaluated in
               numpy complex; GM/PM against the exact crossings (brentq) within grid-interpolation error;
               b(w) minimum in closed form.
"""
import mpmath as mp
import numpy as np
import pytest

from oracle import oracle
from problems import linear_problem


def two_state(dt=1e-3):
    Ac = np.array([[0.0, 1.0], [-4.0, -0.4]])
    return dt * Ac, dt * np.array([0.0, 1.0]), np.array([1.0, 0.0])


def test_linear_simulation_matches_matrix_powers():
    A_I, B, C = two_state()
    gains = [[3.0, 2.0, 0.02, 0.05, 4.0]]
    T = 150
    p = linear_problem(A_I, B, C, gains, T=T, fit=(0.05, 0.7), steady=0.7)
    _, _, dg = oracle.evaluate(p, diag=True)
    y = dg["y"][0]
    mp.mp.prec = 160
    f = lambda v: mp.mpf(float(np.float32(v)))
    A = mp.matrix(2, 2); Bv = mp.matrix(2, 1); Cv = mp.matrix(1, 2)
    for i in range(2):
        Bv[i] = f(B[i]); Cv[i] = f(C[i])
        for j in range(2):
            A[i, j] = (1 if i == j else 0) + f(A_I[i, j])
    K = [f(g) for g in gains[0][:2]]; L = [f(g) for g in gains[0][2:4]]; N = f(gains[0][4])
    # z = [xh; xp]: xh' = A xh + B u + L (y - C xh), xp' = A xp + B u, u = N r - K xh, y = C xp
    M = mp.matrix(4, 4); b = mp.matrix(4, 1)
    for i in range(2):
        for j in range(2):
            M[i, j] = A[i, j] - Bv[i] * K[j] - L[i] * Cv[j]
            M[i, 2 + j] = L[i] * Cv[j]
            M[2 + i, j] = -Bv[i] * K[j]
            M[2 + i, 2 + j] = A[i, j]
        b[i] = Bv[i] * N
        b[2 + i] = Bv[i] * N
    k0 = next(t for t in range(T) if not t < 0.05 * T)
    z = mp.matrix(4, 1)
    for t in range(T):
        yt = z[2]
        assert abs(float(yt) - y[t]) <= 1e-12 * max(1.0, abs(float(yt))), t
        z = M * z + b * (1 if t >= k0 else 0)


def test_saturated_simulation_matches_step_loop():
    A_I, B, C = two_state()
    gains = [[40.0, 9.0, 0.02, 0.05, 45.0]]
    T = 400
    p = linear_problem(A_I, B, C, gains, T=T, sat=2.0, fit=(0.05, 0.7), steady=0.7)
    _, _, dg = oracle.evaluate(p, diag=True)
    A32 = np.float32(A_I).astype(np.float64); B32 = np.float32(B).astype(np.float64)
    K, L, N = np.array(gains[0][:2]), np.array(gains[0][2:4]), gains[0][4]
    K = np.float32(K).astype(float); L = np.float32(L).astype(float); N = float(np.float32(N))
    xh = np.zeros(2); xp = np.zeros(2)
    ys = []
    saturated = False
    for t in range(T):
        y = xp[0]
        ys.append(y)
        u = min(max(N * (0.0 if t < 0.05 * T else 1.0) - K @ xh, -2.0), 2.0)
        saturated |= abs(u) == 2.0
        yerr = y - xh[0]
        xh = xh + (A32 @ xh + B32 * u + L * yerr)
        xp = xp + (A32 @ xp + B32 * u)
    assert saturated
    np.testing.assert_allclose(dg["y"][0], ys, rtol=1e-11, atol=1e-13)


def modes_from_eigs(eig, nt, ts):
    w = eig[0::2] + 1j * eig[1::2]
    s = np.log(np.maximum(np.abs(w), float(np.float32(1.2958e-30)))) / ts + 1j * np.angle(w) / ts
    out = []
    pre, pim = 1e10, 1e10
    eps = float(np.float32(1e-6))
    for z in s:
        dup = abs(z.real - pre) < eps and abs(abs(z.imag) - abs(pim)) < eps
        pre, pim = z.real, z.imag
        if not dup:
            out.append((z.real, abs(z.imag)))
    return out


def test_modal_fit_gram_rhs_and_solution():
    A_I, B, C = two_state()
    rng = np.random.default_rng(3)
    gains = []
    for _ in range(6):
        k = rng.uniform([1, 1], [30, 8])
        l = rng.uniform([0.01, 0.01], [0.1, 0.2])
        gains.append([k[0], k[1], l[0], l[1], 4.0 + k[0]])
    T = 2000
    p = linear_problem(A_I, B, C, gains, T=T)
    d = p.derived()
    _, o64, dg = oracle.evaluate(p, diag=True)
    nt = 4
    ts = float(d.ts)
    checked = 0
    for i in range(len(gains)):
        modes = modes_from_eigs(dg["eig"][i], nt, ts)
        u = len(modes)
        tval = (d.fit_start_id + np.arange(d.fit_count)) * ts
        cols = []
        for re, im in modes:
            sh = max(re * float(d.fit_t0), re * float(d.fit_t1))
            dec = np.exp(re * tval - sh)
            cols += [dec * np.cos(im * tval), dec * np.sin(im * tval)]
        Phi = np.array(cols).T
        y = dg["y"][i]
        yss = y[d.steady_start_id:].mean()
        G = Phi.T @ Phi
        rhs = Phi.T @ (y[d.fit_start_id:d.fit_start_id + d.fit_count] - yss)
        ata = dg["ata"][i].reshape(2 * nt, 2 * nt)
        atb = dg["atb"][i]
        nc = 2 * u
        lower = np.tril(G)
        np.testing.assert_allclose(np.tril(ata[:nc, :nc]), lower, rtol=1e-9, atol=1e-9 * np.abs(G).max())
        np.testing.assert_allclose(atb[:nc], rhs, rtol=1e-7, atol=1e-9 * np.abs(rhs).max())
        # solution and energies of well-conditioned fits
        keep = np.diag(G) > 0
        scale = np.sqrt(np.diag(G)[keep])
        kappa = np.linalg.cond(G[np.ix_(keep, keep)] / scale[:, None] / scale[None, :])
        if kappa > 1e8:
            continue
        coef = np.zeros(nc)
        coef[keep] = np.linalg.lstsq(Phi[:, keep], y[d.fit_start_id:d.fit_start_id + d.fit_count] - yss, rcond=None)[0]
        md = o64[i, d.info_modal:]
        for e in range(u):
            c0, c1 = coef[2 * e], coef[2 * e + 1]
            energy = c0 * c0 * G[2 * e, 2 * e] + 2 * c0 * c1 * G[2 * e + 1, 2 * e] + c1 * c1 * G[2 * e + 1, 2 * e + 1]
            assert abs(md[2 * nt + e] - energy) <= 1e-6 * max(abs(energy), 1e-12) + 1e-12 * np.abs(G).max()
            wn = max(np.hypot(*modes[e]), float(np.float32(1e-6)))
            assert abs(md[2 * e] - wn) <= 1e-12 * wn
        checked += 1
    assert checked >= 3


def test_margins_first_order_loop():
    # 1-state observer model x' = a x + b u, y = x; compensator A1 = a - bK - L, loop P(z) G(z)
    a_i, b = -0.02, 0.05
    K, L = 2.0, 0.3
    F = 4000
    p = linear_problem([[a_i]], [b], [1.0], [[K, L, 1.0]], T=200, F=F, fit=(0.05, 0.7), steady=0.7)
    g = oracle.grid(p)
    _, o64, _ = oracle.evaluate(p)
    d = p.derived()
    am = o64[0, d.info_allmargin:d.info_allmargin + 5]
    a = float(np.float32(1.0 + a_i))   # ss_a holds float32(I + A_I)
    bb = float(np.float32(b))
    K, L = float(np.float32(K)), float(np.float32(L))
    A1 = a - bb * K - L
    z = g["z"]
    loop = (bb / (z - a)) * (K * L / (z - A1))
    # exact gain crossover frequency by root finding on theta, compared within grid interpolation error
    from scipy.optimize import brentq
    th = np.angle(z)
    mag = np.abs(loop)
    idx = np.where(np.diff(np.sign(mag - 1.0)) != 0)[0]
    assert idx.size >= 1
    f = lambda t: abs((bb / (np.exp(1j * t) - a)) * (K * L / (np.exp(1j * t) - A1))) - 1.0
    tc = brentq(f, th[idx[0]], th[idx[0] + 1])
    lc = (bb / (np.exp(1j * tc) - a)) * (K * L / (np.exp(1j * tc) - A1))
    pm_exact = 180.0 + np.degrees(np.angle(lc))
    pm_exact -= 360.0 * np.floor((pm_exact + 180.0) / 360.0)
    assert am[3] == 1.0
    assert abs(am[1] - pm_exact) < 0.05
    # b(w) = |1 + PC| / sqrt((1 + |aP|^2)(1 + |C/a|^2)), minimum over the grid
    alpha = g["alpha"]
    Pw = bb / (z - a)
    Cw = K * L / (z - A1)
    bw = np.abs(1 + Pw * Cw) / np.sqrt((1 + (alpha * np.abs(Pw)) ** 2) * (1 + (np.abs(Cw) / alpha) ** 2))
    assert abs(am[4] - min(1.0, bw.min())) <= 1e-12
    np.testing.assert_allclose(g["presp"], Pw, rtol=1e-12)
