"""The oracle's restated Eigen routines pinned against LAPACK (numpy) and closed forms.

  eigvals      (Eigen EigenSolver<double>; reference src/eigensolve.cu:11, src/allmargin.cu:72)
  ldlt_solve   (Eigen LDLT<Lower>; reference src/modalsolve.cu:164)
  shifted_solve (Eigen PartialPivLU<complex>; reference src/allmargin.cu:26,104)
"""
import numpy as np
import pytest

from oracle import oracle


def match_sets(a, b, tol):
    a = list(a)
    for z in b:
        k = int(np.argmin([abs(z - w) for w in a]))
        assert abs(z - a[k]) <= tol * max(1.0, abs(z)), (z, a[k])
        a.pop(k)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 12, 16])
def test_eigvals_match_lapack_as_sets(n):
    rng = np.random.default_rng(n)
    for trial in range(40):
        a = rng.standard_normal((n, n)) * 10.0 ** rng.uniform(-3, 3)
        st, w = oracle.eigvals(a)
        assert st == 0
        match_sets(w, np.linalg.eigvals(a), 1e-9 * max(1.0, np.abs(a).max()))
        # invariants: trace and determinant
        assert abs(w.sum().real - np.trace(a)) <= 1e-10 * max(1.0, np.abs(a).sum())


def test_eigvals_known_roots_companion_and_blocks():
    roots = np.array([0.9, -0.5, 0.3 + 0.4j, 0.3 - 0.4j, 0.95j, -0.95j, 0.1, -0.99])
    poly = np.real(np.poly(roots))
    comp = np.zeros((8, 8))
    comp[0, :] = -poly[1:]
    comp[1:, :-1] = np.eye(7)
    st, w = oracle.eigvals(comp)
    assert st == 0
    match_sets(w, roots, 1e-9)
    # conjugate pairs come out adjacent, positive imaginary part first
    for i in range(len(w)):
        if w[i].imag > 0:
            assert w[i + 1] == np.conj(w[i])
    # block diagonal rotation-scaling blocks
    blk = np.zeros((4, 4))
    blk[:2, :2] = [[0.5, 0.8], [-0.8, 0.5]]
    blk[2:, 2:] = [[0.1, 0.0], [0.0, -0.7]]
    st, w = oracle.eigvals(blk)
    match_sets(w, [0.5 + 0.8j, 0.5 - 0.8j, 0.1, -0.7], 1e-14)


def test_eigvals_zero_and_diagonal():
    st, w = oracle.eigvals(np.zeros((5, 5)))
    assert st == 0 and np.all(w == 0)
    d = np.diag([3.0, -1.0, 2.0])
    st, w = oracle.eigvals(d)
    assert sorted(w.real) == [-1.0, 2.0, 3.0]


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32])
def test_ldlt_matches_lstsq_on_spd_and_semidefinite(n):
    rng = np.random.default_rng(100 + n)
    for trial in range(20):
        m = rng.standard_normal((n + 3, n))
        g = m.T @ m
        b = rng.standard_normal(n)
        x = oracle.ldlt_solve(g, b)
        ref = np.linalg.solve(g, b)
        assert np.allclose(x, ref, rtol=1e-8, atol=1e-8 * np.abs(ref).max())
    # zero rows/columns (padded normal equations): those unknowns come out exactly 0
    g = np.zeros((n, n))
    k = max(1, n // 2)
    m = rng.standard_normal((k + 2, k))
    g[:k, :k] = m.T @ m
    b = np.zeros(n)
    b[:k] = rng.standard_normal(k)
    x = oracle.ldlt_solve(g, b)
    assert np.all(x[k:] == 0.0)
    assert np.allclose(x[:k], np.linalg.solve(g[:k, :k], b[:k]), rtol=1e-9)
    # all-zero system solves to zero
    assert np.all(oracle.ldlt_solve(np.zeros((n, n)), np.ones(n)) == 0.0)


def test_ldlt_symmetric_pivoting_on_indefinite():
    rng = np.random.default_rng(5)
    for trial in range(50):
        n = 6
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        g = q @ np.diag(rng.uniform(-3, 3, n)) @ q.T
        b = rng.standard_normal(n)
        assert np.allclose(oracle.ldlt_solve(g, b), np.linalg.solve(g, b), rtol=1e-8, atol=1e-9)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8])
def test_shifted_solve_matches_numpy(n):
    rng = np.random.default_rng(200 + n)
    for trial in range(50):
        a = rng.standard_normal((n, n))
        th = rng.uniform(0, np.pi)
        z = np.exp(1j * th)
        b = rng.standard_normal(n)
        ok, x = oracle.shifted_solve(a, z, b)
        assert ok
        ref = np.linalg.solve(z * np.eye(n) - a, b)
        assert np.allclose(x, ref, rtol=1e-9, atol=1e-12)


def test_shifted_solve_singular_is_not_finite():
    a = np.eye(3)
    ok, x = oracle.shifted_solve(a, 1.0 + 0j, np.ones(3))
    assert not ok


def _same_bits(a, b):
    a = np.nan_to_num(np.asarray(a), nan=7.0)
    b = np.nan_to_num(np.asarray(b), nan=7.0)
    return np.array_equal(a.real.view(np.uint64), b.real.view(np.uint64)) and \
        np.array_equal(a.imag.view(np.uint64), b.imag.view(np.uint64))


def test_sign_flip_equivariance_of_the_eigensolver():
    """With D = diag(I_k, -I), the solve of D A D returns the bits of the solve of A unless the solve
    of A reports a sign-flip event (a Householder vector with a zero leading entry and a nonzero
    tail, or a Francis start test at row k - 1 whose outcome depends on the sign of its right-hand
    side).  The GPU uses this to decide the nominal loop of the margins stage, which is D LIN_A D
    when the margins plant is the simulated plant and D = 0, from the LIN_A solve.  Checked on
    random dense, sparse and quantised matrices, where events do occur, and on matrices without."""
    rng = np.random.default_rng(11)
    events = checked = 0
    for n in [2, 3, 4, 6, 8, 12, 16]:
        k = n // 2
        d = np.concatenate([np.ones(k), -np.ones(n - k)])
        for trial in range(300):
            A = rng.standard_normal((n, n))
            if trial % 3 == 1:
                A[rng.random((n, n)) < 0.4] = 0.0
            if trial % 3 == 2:
                A = np.round(A * 4) / 4
            st1, e1, ev = oracle.eigvals_events(A, k)
            st2, e2 = oracle.eigvals(d[:, None] * A * d[None, :])
            events += ev
            if not ev:
                checked += 1
                assert st1 == st2 and _same_bits(e1, e2), (n, trial)
    assert checked > 1500 and events > 50
