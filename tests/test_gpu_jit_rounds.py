"""Serial rounds through one engine: the plant constants change every round (reference serial mode,
src/main.cu:467-470), the instruction structure does not.  The engine compiles the first round's
simulation with its constants in the source and every later round with the structure's kernel
that reads them from the descriptor: one more compilation in all, not one per round.  The rows of
every round equal the oracle's (FP64) and the rows of a fresh engine's literal kernel (mixed)."""
import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def plant():
    w = workloads.C4(seed=21, n_plants=3, rounds=5, nmin=300, nmax=500, tmin=800, tmax=1200, F=64)
    return max(w.plants, key=lambda p: p.ns)


@pytest.mark.parametrize("prec", [PRECISION_FP64, PRECISION_MIXED])
def test_one_compile_per_structure_over_rounds(prec):
    from controllerengine_amd.engine import Engine
    p = plant()
    eng = Engine(0)
    try:
        counts = []
        for r in range(5):
            print("round", r, "prec", prec, flush=True)
            _, got64 = eng.evaluate(p, precision=prec, rnd=r)
            counts.append(eng.jit_compiles())
            if prec == PRECISION_FP64:
                _, ref64, _ = oracle.evaluate(p, rnd=r)
                same = (got64 == ref64) | (np.isnan(got64) & np.isnan(ref64))
                assert same.all(), (r, int((~same).sum()))
            else:
                fresh = Engine(0)
                try:
                    _, lit64 = fresh.evaluate(p, precision=prec, rnd=r)
                finally:
                    fresh.close()
                assert np.array_equal(got64.view(np.uint64), lit64.view(np.uint64)), r
        assert counts[0] >= 1
        assert counts[1] == counts[0] + 1 and counts[4] == counts[1], counts
    finally:
        eng.close()


def test_sim_history_ring():
    """The simulation-stage history keeps the last `calls` evaluations, oldest first, and recording
    it changes no row."""
    from controllerengine_amd.engine import Engine
    p = plant()
    eng = Engine(0)
    try:
        _, plain = eng.evaluate(p, precision=PRECISION_MIXED)
        eng.set_sim_history(2)
        assert eng.sim_history() == []
        for r in range(3):
            print("history call", r, flush=True)
            _, got = eng.evaluate(p, precision=PRECISION_MIXED)
            assert np.array_equal(got.view(np.uint64), plain.view(np.uint64))
        h = eng.sim_history()
        assert len(h) == 2 and all(ms > 0.0 for ms in h), h
        eng.set_sim_history(0)
        eng.evaluate(p, precision=PRECISION_MIXED)
        assert eng.sim_history() == []
    finally:
        eng.close()
