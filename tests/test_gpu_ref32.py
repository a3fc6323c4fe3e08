"""The mixed chain's float simulation is the reference's float simulation, bit for bit.

The product's float kernels (runtime-specialised and interpreted) keep the reference's hi/lo state
(src/discrete_solves.cu:12-79, hladdf src/main.cu:129-155); the oracle's ref-FP32 chain restates
the same arithmetic on the CPU.  Every simulated trajectory (the diagnostics dump of y) and the
overshoot and steady-state cells the simulation writes (src/modalsetup.cu:52-62, float) must equal
the oracle's, for candidates both chains simulate.
"""
import numpy as np
import pytest

from controllerengine_amd import PRECISION_MIXED, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("jit", ["1", "0"])
@pytest.mark.parametrize("name,n", [("c1", 512), ("c2", 512), ("c3", 256)])
def test_mixed_simulation_is_reference_float_simulation(engine, name, n, jit, monkeypatch):
    monkeypatch.setenv("CE_JIT", jit)
    p = workloads.CONFIGS[name](n=n)
    d = p.derived()
    print("case", name, n, "jit", jit, flush=True)
    r32, y32 = oracle.evaluate_ref32(p, want_y=True)
    diag = engine.set_diagnostics(n, p.cfg_struct())
    try:
        _, got64 = engine.evaluate(p, PRECISION_MIXED)
        y_gpu = diag["y"].cpu().numpy()
    finally:
        engine.set_diagnostics(n, p.cfg_struct(), enable=False)
    ue = float(p.config["unstable_err"])
    simulated = got64[:, d.info_scoring] < ue           # the GPU simulates linearly stable candidates
    assert simulated.sum() >= n // 8
    same = y_gpu[simulated].view(np.uint32) == y32[simulated].view(np.uint32)
    assert same.all(), "%d of %d trajectory samples differ" % ((~same).sum(), same.size)
    both = simulated & (r32[:, d.info_scoring] < ue)
    fr = d.info_freqresp
    for k in (2, 3):
        g = got64[both, fr + k].astype(np.float32)
        assert np.array_equal(g.view(np.uint32), r32[both, fr + k].view(np.uint32)), k
