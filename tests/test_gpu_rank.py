"""Device ranking and selection (ce_rank) against numpy on the oracle's rows.

The order must equal a stable argsort of the float score (ties by index, NaN last) and the pass
flags the rule of metrics.pass_flags, exactly, on the FP64 rows (which equal the oracle's) of the
benchmark workloads and on rows with planted ties, signed zeros, infinities and NaN scores."""
import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, metrics, workloads
from controllerengine_amd.ranking import rank
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,n", [("c1", 1024), ("c2", 4096), ("c3", 2048)])
def test_rank_matches_numpy_on_oracle_rows(engine, name, n):
    p = workloads.CONFIGS[name](n=n)
    print("rank", name, n, flush=True)
    r32, _, _ = oracle.evaluate(p)
    g32, _ = engine.evaluate(p, PRECISION_FP64)
    assert np.array_equal(g32.view(np.uint32), r32.view(np.uint32))
    order, passf, npass = rank(p, g32)
    assert np.array_equal(order, metrics.ranking(p, r32))
    want = metrics.pass_flags(p, r32.astype(np.float64))
    assert np.array_equal(passf, want) and npass == int(want.sum())
    # top-k is the prefix of the full order; mixed rows rank the same way through the same call
    o10, _, _ = rank(p, g32, k=10)
    assert np.array_equal(o10, order[:10])
    m32, _ = engine.evaluate(p, PRECISION_MIXED)
    om, pm, _ = rank(p, m32)
    assert np.array_equal(om, metrics.ranking(p, m32))
    assert np.array_equal(pm, metrics.pass_flags(p, m32.astype(np.float64)))


def test_rank_edge_scores(engine):
    p = workloads.c1(n=700)
    rows, _ = engine.evaluate(p, PRECISION_FP64)
    d = p.derived()
    rng = np.random.default_rng(3)
    s = rows[:, d.info_scoring]
    s[::7] = 2.5                      # ties
    s[1::11] = -0.0
    s[2::11] = 0.0
    s[3::13] = np.inf
    s[4::13] = -np.inf
    s[5::17] = np.nan
    s[6::19] = -rng.random(len(s[6::19]), dtype=np.float32)
    rows[:, d.info_scoring] = s
    order, _, _ = rank(p, rows)
    assert np.array_equal(order, metrics.ranking(p, rows))
    assert np.isnan(s[order[-1]]) and s[order[0]] == -np.inf
