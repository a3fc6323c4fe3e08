"""ce_rank_noalloc (the ranking entry point that allocates no device memory) against ce_rank,
through the C ABI: identical order, pass flags and count on two calls in a row, the second with
fewer candidates (the static scratch keeps the first call's contents), and a clear error for a
candidate count above the maximum (checked on the host before anything is enqueued)."""
import ctypes

import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, _abi, workloads

pytestmark = pytest.mark.gpu


def _call(fn, cfg, d_info, n, k):
    import torch
    dev = d_info.device
    order = torch.full((max(k, 1),), -1, dtype=torch.int32, device=dev)
    passf = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    npass = torch.full((1,), -1, dtype=torch.int32, device=dev)
    err = _abi.errbuf()
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    rc = fn(ctypes.byref(cfg), ctypes.c_void_p(d_info.data_ptr()), n, k, ctypes.c_void_p(order.data_ptr()),
            ctypes.c_void_p(passf.data_ptr()), ctypes.c_void_p(npass.data_ptr()), st, err, len(err))
    assert rc == 0, err.value.decode()
    torch.cuda.synchronize()
    return order.cpu().numpy()[:k], passf.cpu().numpy(), int(npass.cpu().numpy()[0])


def test_noalloc_rank_equals_rank(engine):
    import torch
    lib = _abi.lib()
    for name, n, k in [("c2", 6000, 6000), ("c1", 700, 50)]:
        p = workloads.CONFIGS[name](n=n)
        print("noalloc", name, n, k, flush=True)
        rows, _ = engine.evaluate(p, PRECISION_FP64)
        d = torch.from_numpy(rows).cuda()
        cfg = p.cfg_struct()
        want = _call(lib.ce_rank, cfg, d, n, k)
        got = _call(lib.ce_rank_noalloc, cfg, d, n, k)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]) and got[2] == want[2]
        assert set(np.unique(got[1])) <= {0, 1}


def test_noalloc_rejects_too_many(engine):
    import torch
    p = workloads.c1(n=16)
    cfg = p.cfg_struct()
    d = torch.zeros((16, p.derived().info_mx_width), dtype=torch.float32, device="cuda")
    err = _abi.errbuf()
    n = (1 << 20) + 1
    rc = _abi.lib().ce_rank_noalloc(ctypes.byref(cfg), ctypes.c_void_p(d.data_ptr()), n, 0, None, None, None, None,
                                    err, len(err))
    assert rc == 1 and b"exceeds the maximum" in err.value
