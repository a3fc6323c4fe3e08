"""Kalman-filter simulation on the GPU (ce_simulate_kalman): FP64 trajectories equal the oracle's
bit for bit (NaN == NaN) on a linear 2-state loop, a 3-state observer over a 2-state plant and the
8-state plant with its four nonlinearities; the mixed (compensated float) trajectories stay within
1e-4 relative of them."""
import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
from oracle import oracle
from problems import general_problem

pytestmark = pytest.mark.gpu


def cases():
    rng = np.random.default_rng(12)
    dt = 1e-3
    AI = dt * np.array([[0.0, 1.0], [-4.0, -0.4]])
    B = dt * np.array([0.0, 1.0])
    p1 = general_problem(AI, B, [1.0, 0.0], 0.0, AI, B, 0, [[0.0] * 5], T=1200, F=16)
    r1 = np.column_stack([rng.uniform(5, 60, (256, 2)), np.full(256, 4.0), 10 ** rng.uniform(-3, 0, 256),
                          10 ** rng.uniform(-5, -1, (256, 2))])
    sai = rng.uniform(-0.02, 0.02, (3, 3)) - 0.01 * np.eye(3)
    dai = rng.uniform(-0.02, 0.02, (2, 2)) - 0.01 * np.eye(2)
    p2 = general_problem(sai, rng.uniform(-0.1, 0.1, 3), rng.uniform(-1, 1, 3), 0.0, dai, rng.uniform(-0.1, 0.1, 2),
                         1, [[0.0] * 7], T=700, F=16)
    r2 = np.column_stack([rng.uniform(-2, 2, (130, 3)), np.full(130, 1.5), 10 ** rng.uniform(-2, 0, 130),
                          10 ** rng.uniform(-4, -1, (130, 3))])
    p3 = workloads.c3(n=200, T=600, F=32)
    r3 = np.column_stack([p3.gains[:, :8], p3.gains[:, 16], 10 ** rng.uniform(-3, -1, 200),
                          10 ** rng.uniform(-4, -2, (200, 8))])
    return [("linear2", p1, np.float32(r1)), ("obs3_plant2", p2, np.float32(r2)), ("c3_nl8", p3, np.float32(r3))]


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_kalman_fp64_matches_oracle_and_mixed_close(engine, idx):
    name, p, rows = cases()[idx]
    print("kalman", name, rows.shape, flush=True)
    want = oracle.simulate_kalman(p, rows)
    got = engine.simulate_kalman(p, rows, PRECISION_FP64)
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    assert same.all(), (name, int((~same).sum()))
    mix = engine.simulate_kalman(p, rows, PRECISION_MIXED).astype(np.float64)
    fin = np.isfinite(want).all(axis=1) & (np.abs(want).max(axis=1) < 1e6)
    assert fin.sum() >= rows.shape[0] // 2
    scale = np.abs(want[fin]).max(axis=1, keepdims=True) + 1e-30
    assert np.max(np.abs(mix[fin] - want[fin]) / scale) <= 1e-4, name
