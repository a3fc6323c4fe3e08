"""Gain search through the engine (GPU): the search's best row is the chain's row for that
candidate, and the best score never rises across iterations."""
import dataclasses

import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
from controllerengine_amd.search import GainSearch
from oracle import oracle

pytestmark = pytest.mark.gpu


def test_search_fp64_best_row_matches_oracle(engine):
    p = workloads.c1(n=512)
    print("case c1 search fp64 n=2048 iters=4", flush=True)
    s = GainSearch(p, engine=engine, precision=PRECISION_FP64, n=2048, seed=5)
    res = s.run(iters=4)
    best = [h["best"] for h in res.history]
    assert all(b2 <= b1 for b1, b2 in zip(best, best[1:]))
    assert np.isfinite(res.best_score)
    one = dataclasses.replace(p, gains=res.best_gains.reshape(1, -1),
                              info=np.zeros((1, p.info.shape[1]), np.float32))
    ref32, _, _ = oracle.evaluate(one)
    same = (ref32[0] == res.best_row) | (np.isnan(ref32[0]) & np.isnan(res.best_row))
    assert same.all(), "search row differs from the oracle in %d cells" % int((~same).sum())


def test_search_mixed_improves_on_the_problem_block(engine):
    p = workloads.c2(n=4096)
    print("case c2 search mixed n=8192 iters=5", flush=True)
    _, rows = engine.evaluate(p, PRECISION_MIXED)
    sc = p.derived().info_scoring
    start = np.nanmin(rows[:, sc])
    s = GainSearch(p, engine=engine, precision=PRECISION_MIXED, n=8192, seed=1)
    res = s.run(iters=5)
    print("block best %.6g search best %.6g" % (start, res.best_score), flush=True)
    assert res.best_score <= start
