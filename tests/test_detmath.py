"""Deterministic elementary functions (controllerengine_amd/csrc/ce_detmath.h) against mpmath.

Both the oracle's copy (compiled by g++) and the product's host copy (compiled by hipcc's host
pass) are checked: accuracy within a few ulp of the correctly rounded value, C99 special cases,
and identical bits between the two builds."""
import ctypes

import mpmath as mp
import numpy as np
import pytest

from controllerengine_amd import _abi
from oracle import oracle

mp.mp.prec = 120
FNS = {0: mp.exp, 1: mp.log, 2: mp.sin, 3: mp.cos, 4: lambda y, x: mp.atan2(y, x),
       5: lambda x, y: mp.sqrt(x * x + y * y), 6: mp.log10}
BOUND = {0: 2, 1: 2, 2: 3, 3: 3, 4: 4, 5: 1, 6: 3}


def product_host(fn, x, y):
    out = np.zeros_like(x)
    _abi.lib().ce_detmath_host(fn, x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p),
                               out.ctypes.data_as(ctypes.c_void_p), x.size)
    return out


def ulp_err(got, x, y, fn):
    errs = []
    for g, a, b in zip(got, x, y):
        ref = FNS[fn](mp.mpf(a), mp.mpf(b)) if fn in (4, 5) else FNS[fn](mp.mpf(a))
        r = float(ref)
        if r == 0.0:
            errs.append(0.0 if g == 0.0 else np.inf)
            continue
        u = np.spacing(abs(r))
        errs.append(float(abs(mp.mpf(g) - ref) / u))
    return max(errs)


def samples(fn, rng, n=400):
    if fn == 0:
        x = np.concatenate([rng.uniform(-745, 709, n), rng.uniform(-1, 1, n // 4), rng.uniform(-1e-8, 1e-8, 20)])
    elif fn in (1, 6):
        x = np.concatenate([10.0 ** rng.uniform(-300, 300, n), rng.uniform(0.5, 2.0, n // 2),
                            1.0 + rng.uniform(-1e-6, 1e-6, 20), [5e-324, 2.2e-308]])
    elif fn in (2, 3):
        x = np.concatenate([rng.uniform(-20000, 20000, n), rng.uniform(-4, 4, n // 2), [1e-300, np.pi / 2, np.pi]])
    else:
        x = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 3, n)
    y = rng.standard_normal(x.size) * 10.0 ** rng.uniform(-3, 3, x.size)
    return x.astype(np.float64), y.astype(np.float64)


@pytest.mark.parametrize("fn", list(FNS))
def test_accuracy_and_identical_builds(fn):
    rng = np.random.default_rng(100 + fn)
    x, y = samples(fn, rng)
    a = oracle.detmath(fn, x, y)
    b = product_host(fn, x, y)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), "oracle and product builds differ"
    assert ulp_err(a, x, y, fn) <= BOUND[fn]


def test_special_values():
    inf, nan = np.inf, np.nan
    e = oracle.detmath(0, np.array([0.0, -inf, inf, nan, 710.0, -746.0]))
    assert e[0] == 1.0 and e[1] == 0.0 and e[2] == inf and np.isnan(e[3]) and e[4] == inf and e[5] == 0.0
    lg = oracle.detmath(1, np.array([0.0, -0.0, -1.0, inf, nan, 1.0]))
    assert lg[0] == -inf and lg[1] == -inf and np.isnan(lg[2]) and lg[3] == inf and np.isnan(lg[4]) and lg[5] == 0.0
    s = oracle.detmath(2, np.array([inf, nan, 0.0, -0.0]))
    assert np.isnan(s[0]) and np.isnan(s[1]) and s[2] == 0.0
    ys = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, inf, inf, -inf, 1.0, 1.0])
    xs = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, inf, -inf, 1.0, inf, -inf])
    got = oracle.detmath(4, ys, xs)
    ref = np.arctan2(ys, xs)
    assert np.allclose(got, ref, rtol=1e-15, atol=0) and np.array_equal(np.signbit(got), np.signbit(ref))
    h = oracle.detmath(5, np.array([inf, nan, 3e300, 3e-300, 0.0]), np.array([nan, 1.0, 4e300, 4e-300, 0.0]))
    assert h[0] == inf and np.isnan(h[1]) and abs(h[2] / 5e300 - 1) < 1e-15 and abs(h[3] / 5e-300 - 1) < 1e-15 and h[4] == 0
