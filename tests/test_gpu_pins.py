"""The oracle's pinning problems through the GPU chain (every score branch, D != 0, ns != nd, a
non-finite trajectory, an exactly singular sweep point).

FP64: rows bit-identical to the oracle (NaN == NaN), with and without the specialised simulation
kernel.  Mixed: the same branch on every row (stable flag, rejected / nominally unstable / sweep
failure cells) and the margins within the matched-margin tolerances where both rows have them.
"""
import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED
from oracle import oracle
from test_gpu_parity import assert_rows_equal
from test_oracle_pins import pin_cases

pytestmark = pytest.mark.gpu

CASES = pin_cases()


@pytest.mark.parametrize("jit", ["1", "0"])
@pytest.mark.parametrize("name", [n for n, _ in CASES])
def test_fp64_pin_case_matches_oracle(engine, name, jit, monkeypatch):
    monkeypatch.setenv("CE_JIT", jit)
    p = dict(CASES)[name]
    print("case", name, "jit", jit, flush=True)
    _, ref64, _ = oracle.evaluate(p)
    _, got64 = engine.evaluate(p, PRECISION_FP64)
    assert_rows_equal(got64, ref64, name)


@pytest.mark.parametrize("name", [n for n, _ in CASES])
def test_mixed_pin_case_branches(engine, name):
    p = dict(CASES)[name]
    print("case", name, "mixed", flush=True)
    d = p.derived()
    _, ref64, _ = oracle.evaluate(p)
    _, got64 = engine.evaluate(p, PRECISION_MIXED)
    sc, am = d.info_scoring, d.info_allmargin
    ue, ef = float(np.float32(1e6)), float(np.float32(1e18))
    for r, g in zip(ref64, got64):
        assert r[am + 3] == g[am + 3]
        assert (r[sc + 2] == ue) == (g[sc + 2] == ue)
        assert (r[sc + 2] == ef) == (g[sc + 2] == ef)
        assert (r[sc] >= ue) == (g[sc] >= ue)
        if r[am + 3] == 1.0 and r[sc + 2] < ue:
            assert abs(r[am + 0] - g[am + 0]) <= 0.01
            assert abs(r[am + 1] - g[am + 1]) <= 0.01
            assert abs(r[am + 4] - g[am + 4]) <= 1e-4
