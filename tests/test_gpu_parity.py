"""FP64 GPU chain vs the CPU oracle, through the C ABI.

The FP64 path is built to reproduce the oracle's operation order exactly, so the INFO rows are
compared bit for bit (NaN == NaN); the stated tolerance of the FP64 mode (1e-9 relative) is asserted
as well so that a failure reports its size.
"""
import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu

REL_TOL = 1e-9


def assert_rows_equal(got, ref, label):
    assert got.shape == ref.shape, label
    both_nan = np.isnan(got) & np.isnan(ref)
    same = (got == ref) | both_nan
    if not same.all():
        bad = np.argwhere(~same)
        r, c = bad[0]
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        rel[both_nan] = 0
        msg = "%s: %d cells differ (first row %d col %d: got %r ref %r); max rel %.3g" % (
            label, len(bad), r, c, got[r, c], ref[r, c], np.nanmax(rel))
        assert np.nanmax(rel) <= REL_TOL, msg
        pytest.fail(msg)


@pytest.mark.parametrize("jit", ["1", "0"])
@pytest.mark.parametrize("name,n", [("c1", 1024), ("c2", 1024), ("c3", 512)])
def test_fp64_chain_matches_oracle(engine, name, n, jit, monkeypatch):
    from controllerengine_amd import _abi
    monkeypatch.setenv("CE_JIT", jit)
    p = workloads.CONFIGS[name](n=n)
    print("case", name, n, "jit", jit, flush=True)
    _, ref64, _ = oracle.evaluate(p)
    got32, got64 = engine.evaluate(p, PRECISION_FP64)
    assert _abi.lib().ce_engine_last_sim_jit(engine.h) == int(jit)
    assert_rows_equal(got64, ref64, name)
    assert np.array_equal(got32.view(np.uint32), ref64.astype(np.float32).view(np.uint32))


def test_detmath_device_matches_host(engine):
    import ctypes
    import torch
    from controllerengine_amd import _abi
    rng = np.random.default_rng(7)
    n = 1 << 16
    x = np.concatenate([rng.uniform(-700, 700, n), rng.standard_normal(n) * 1e4, [0.0, -0.0, np.inf, -np.inf, np.nan]])
    y = np.concatenate([rng.uniform(-5, 5, n), rng.standard_normal(n), [0.0, 1.0, -1.0, np.inf, 2.0]])
    for fn in range(7):
        xx = np.abs(x) if fn in (1, 6) else x
        host = np.zeros_like(xx)
        _abi.lib().ce_detmath_host(fn, xx.ctypes.data, y.ctypes.data, host.ctypes.data, xx.size)
        dx = torch.from_numpy(xx).cuda()
        dy = torch.from_numpy(y).cuda()
        do = torch.empty_like(dx)
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert _abi.lib().ce_detmath_device(fn, ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()),
                                            ctypes.c_void_p(do.data_ptr()), xx.size, st) == 0
        dev = do.cpu().numpy()
        assert np.array_equal(dev.view(np.uint64), host.view(np.uint64)) or \
            np.array_equal(np.nan_to_num(dev, nan=7.0).view(np.uint64), np.nan_to_num(host, nan=7.0).view(np.uint64)), fn
