"""The eigensolver's square root (grp_sqrt: the hardware reciprocal square root and its corrections
without range scaling, inside [2^-700, 2^700]) against the host's IEEE sqrt, bit for bit.

Inputs: every binade of positive doubles (random bit patterns), the guard's edges and their
neighbours, subnormals, signed zeros, infinities, NaN and negative numbers.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs():
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 0x7FF0000000000000, 1 << 22, dtype=np.int64)
    x = [bits.view(np.float64)]
    x.append(rng.uniform(1e-12, 1e4, 1 << 20))
    x.append(np.exp(rng.uniform(-700, 700, 1 << 20)))
    edges = []
    for e in (-700, 700, -767, -1022, 1023):
        v = np.float64(2.0) ** e
        edges += [v, np.nextafter(v, 0), np.nextafter(v, np.inf)]
        for k in range(1, 64):
            edges += [v * (1 + k * 2.0 ** -52), v * (1 - k * 2.0 ** -53)]
    x.append(np.array(edges))
    x.append(np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -1.0, 5e-324, 2.2250738585072014e-308, 1.0, 4.0]))
    return np.ascontiguousarray(np.concatenate(x))


def test_eigensolver_sqrt_matches_ieee(engine):
    import torch
    from controllerengine_amd import _abi
    x = _inputs()
    y = np.zeros_like(x)
    host = np.zeros_like(x)
    assert _abi.lib().ce_detmath_host(7, x.ctypes.data, y.ctypes.data, host.ctypes.data, x.size) == 0
    dx = torch.from_numpy(x).cuda()
    dy = torch.from_numpy(y).cuda()
    do = torch.empty_like(dx)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    print("grp_sqrt inputs", x.size, flush=True)
    assert _abi.lib().ce_detmath_device(7, ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()),
                                        ctypes.c_void_p(do.data_ptr()), x.size, st) == 0
    dev = do.cpu().numpy()
    same = (dev.view(np.uint64) == host.view(np.uint64)) | (np.isnan(dev) & np.isnan(host))
    assert same.all(), "grp_sqrt differs from sqrt at %d inputs, first %r: %r vs %r" % (
        int((~same).sum()), x[~same][0], dev[~same][0], host[~same][0])
