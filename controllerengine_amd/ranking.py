"""Device ranking and selection of a candidate block's INFO rows, through ce_rank.

The reference leaves the selection to its caller (README.md:22-35); ce_rank orders candidates by
score on the device and flags those that meet every requirement the rows show (the rule is stated
in include/ControllerEngine.h).  torch provides device memory and the stream only.
"""
import ctypes

import numpy as np

from . import _abi


def rank_device(cfg, d_info, k=None, want_pass=True, stream=None):
    """d_info: n x info_w float32 device tensor.  Returns device tensors (order[k], pass[n] uint8,
    npass[1]) enqueued on torch's current stream (or `stream`)."""
    import torch
    n = int(d_info.shape[0])
    k = n if k is None else int(k)
    dev = d_info.device
    order = torch.empty(max(k, 1), dtype=torch.int32, device=dev)
    passf = torch.empty(n, dtype=torch.uint8, device=dev) if want_pass else None
    npass = torch.empty(1, dtype=torch.int32, device=dev)
    st = stream if stream is not None else ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    err = _abi.errbuf()
    rc = _abi.lib().ce_rank(ctypes.byref(cfg), ctypes.c_void_p(d_info.data_ptr()), n, k,
                            ctypes.c_void_p(order.data_ptr()), ctypes.c_void_p(passf.data_ptr()) if want_pass else None,
                            ctypes.c_void_p(npass.data_ptr()), st, err, len(err))
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return order[:k], passf, npass


def rank(problem, rows, k=None):
    """Host convenience: rows (n x info_w float32) -> (order[k], pass[n] bool, npass) as numpy."""
    import torch
    d = torch.from_numpy(np.ascontiguousarray(rows, dtype=np.float32)).cuda()
    order, passf, npass = rank_device(problem.cfg_struct(), d, k)
    torch.cuda.synchronize()
    return order.cpu().numpy(), passf.cpu().numpy().astype(bool), int(npass.cpu().numpy()[0])
