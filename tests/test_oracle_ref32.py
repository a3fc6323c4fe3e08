"""The oracle's ref-FP32 chain (the reference's own precision map) pinned against the plain FP64
chain, which the other oracle tests pin independently, and its float solvers against LAPACK.

ref-FP32 restates src/discrete_solves.cu:12-79 (float, hi/lo state), src/eigensetup.cu (float),
src/eigensolve.cu (double solve stored as float), src/modalsetup.cu / modalsolve.cu (float Gram
sums and LDLT) and src/allmargin.cu (complex<float> sweep).  On well-conditioned loops it must
land within float rounding of the FP64 chain, cell by cell, and take every branch the FP64 chain
takes; a dropped compensation term, a wrong sign or a wrong operand shows as an error far above
float rounding on at least one of the pinned problems.
"""
import numpy as np
import pytest

from controllerengine_amd import workloads
from oracle import oracle
from tests.test_oracle_pins import pin_cases


def test_ldlt32_matches_numpy():
    rng = np.random.default_rng(5)
    for n in (1, 2, 4, 8, 16, 32):
        for _ in range(20):
            q, _ = np.linalg.qr(rng.standard_normal((n, n)))
            m = (q * rng.uniform(1.0, 1e3, n)) @ q.T
            m32 = m.astype(np.float32)
            b = rng.standard_normal(n).astype(np.float32)
            x = oracle.ldlt_solve32(m32, b)
            ref = np.linalg.solve(m32.astype(np.float64), b.astype(np.float64))
            assert np.abs(x - ref).max() <= 2e-3 * np.abs(ref).max(), (n, np.abs(x - ref).max())
    # a zero pivot leaves its unknown at zero (Eigen's LDLT solve)
    m = np.diag([4.0, 0.0, 2.0]).astype(np.float32)
    x = oracle.ldlt_solve32(m, np.array([8.0, 5.0, 2.0], dtype=np.float32))
    assert list(x) == [2.0, 0.0, 1.0]


def test_shifted_solve32_matches_numpy():
    rng = np.random.default_rng(6)
    for n in range(1, 9):
        for _ in range(20):
            a = (rng.standard_normal((n, n)) * 0.3).astype(np.float32)
            b = rng.standard_normal(n).astype(np.float32)
            th = rng.uniform(0.01, 3.0)
            z = np.complex64(np.cos(th) + 1j * np.sin(th))
            ok, x = oracle.shifted_solve32(a, z, b)
            ref = np.linalg.solve(np.complex128(z) * np.eye(n) - a.astype(np.float64), b.astype(np.complex128))
            assert ok
            assert np.abs(x - ref).max() <= 1e-4 * max(1.0, np.abs(ref).max()), (n, np.abs(x - ref).max())


def stable_rows(p, o64):
    d = p.derived()
    return o64[:, d.info_scoring] < float(p.config["unstable_err"])


@pytest.mark.parametrize("cfg,tol", [("c1", 1e-7), ("c3", 1e-6)])
def test_trajectory_is_compensated(cfg, tol):
    """Float state with the hi/lo pairs follows the FP64 trajectory to a few float ulps of its
    scale over the whole horizon (plain float accumulation drifts by 1e-6 to 1e-5)."""
    p = workloads.CONFIGS[cfg]().subset(np.arange(192))
    _, o64, dg = oracle.evaluate(p, diag=True)
    _, y32 = oracle.evaluate_ref32(p, want_y=True)
    st = stable_rows(p, o64)
    assert st.sum() > 50
    scale = np.abs(dg["y"]).max(axis=1)
    err = np.abs(y32.astype(np.float64) - dg["y"]).max(axis=1) / scale
    assert err[st].max() <= tol, err[st].max()


def cell_check(p, o64, r32, name):
    d = p.derived()
    ns, nt = int(p.config["statespace_dim"]), d.total_dim
    for i in range(o64.shape[0]):
        a, b = o64[i], r32[i].astype(np.float64)
        sc, fr, am = d.info_scoring, d.info_freqresp, d.info_allmargin
        # branch taken: unstable flag, margin-stable flag, nominal penalty
        assert (a[sc] >= p.config["unstable_err"]) == (b[sc] >= p.config["unstable_err"]), (name, i)
        assert a[am + 3] == b[am + 3], (name, i)
        for k in list(range(sc, sc + 3)) + list(range(fr, fr + 4)) + list(range(am, am + 5)):
            x, y = a[k], b[k]
            if not np.isfinite(x) or abs(x) >= 1e38:
                assert (not np.isfinite(y) and np.isnan(x) == np.isnan(y)) or abs(y) >= 1e38 or x == y, (name, i, k, x, y)
                continue
            assert abs(x - y) <= 2e-3 * max(1.0, abs(x)), (name, i, k, x, y)


@pytest.mark.parametrize("k", range(11))
def test_pinned_problems_match_fp64_cell_by_cell(k):
    name, p = pin_cases()[k]
    _, o64, _ = oracle.evaluate(p)
    r32, _ = oracle.evaluate_ref32(p)
    cell_check(p, o64, r32, name)


def test_c1_block_matches_fp64():
    p = workloads.CONFIGS["c1"]()
    _, o64, _ = oracle.evaluate(p)
    r32, _ = oracle.evaluate_ref32(p)
    cell_check(p, o64, r32, "c1")
