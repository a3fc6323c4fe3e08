"""Batched evaluation of many problems through ce_batch_* (one grouped launch per stage kernel
and template instance, optional hipGraph replay and grid autotuning).

Inputs of all problems are contiguous device arrays in the order the problems were added; the
host convenience `evaluate` assembles them from numpy rows.  torch provides device memory and the
stream only; there is no CPU path.
"""
import ctypes

import numpy as np

from . import _abi


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Batch:
    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("Batch needs a HIP device")
        self.torch = torch
        self.device = device
        err = _abi.errbuf()
        self.h = _abi.lib().ce_batch_create(int(device), err, len(err))
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.problems = []

    def close(self):
        if self.h:
            _abi.lib().ce_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        _abi.lib().ce_batch_clear(self.h)
        self.problems = []

    def add(self, problem, rnd=0, n=None):
        """Adds one problem (round `rnd` of its plant); n defaults to its candidate rows."""
        n = int(problem.gains.shape[0] if n is None else n)
        err = _abi.errbuf()
        cfg = problem.cfg_struct()
        pl = problem.plant_struct(rnd)
        idx = _abi.lib().ce_batch_add(self.h, ctypes.byref(cfg), ctypes.byref(pl), n, err, len(err))
        if idx < 0:
            raise ValueError(err.value.decode())
        self.problems.append((problem, n))
        return idx

    def update(self, index, problem, rnd=0):
        err = _abi.errbuf()
        cfg = problem.cfg_struct()
        pl = problem.plant_struct(rnd)
        if _abi.lib().ce_batch_update(self.h, int(index), ctypes.byref(cfg), ctypes.byref(pl), err, len(err)) != 0:
            raise ValueError(err.value.decode())

    def offsets(self, index):
        g, i = ctypes.c_longlong(), ctypes.c_longlong()
        _abi.lib().ce_batch_offsets(self.h, int(index), ctypes.byref(g), ctypes.byref(i))
        return g.value, i.value

    def set_options(self, use_graph=True, autotune=False, profiling=False):
        _abi.lib().ce_batch_set_options(self.h, int(use_graph), int(autotune), int(profiling))

    def set_plan(self, eig_lin=-1, eig_nom=-1, sim=-1):
        """Fix the stage plans (eigen stages: 0 per capacity, default lanes; 1 per capacity,
        alternative lanes; 2 fused; simulation: 0 per structure, 1 fused); -1 hands them back to
        the autotuner."""
        if _abi.lib().ce_batch_set_plan(self.h, int(eig_lin), int(eig_nom), int(sim)) != 0:
            raise ValueError("invalid batch plan")

    def plan(self):
        buf = (ctypes.c_int * 3)()
        _abi.lib().ce_batch_plan(self.h, buf)
        return dict(eig_lin=buf[0], eig_nom=buf[1], sim=buf[2])

    def info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _abi.lib().ce_batch_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return dict(launches=a.value, used_graph=bool(b.value), tuned=bool(c.value))

    def kernel_times(self):
        names = ctypes.create_string_buffer(8192)
        buf = (ctypes.c_double * 256)()
        k = _abi.lib().ce_batch_kernel_times(self.h, names, len(names), buf, 256)
        if k <= 0:
            return []
        return list(zip(names.value.decode().split(","), [buf[i] for i in range(k)]))

    def stream_handle(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def run_device(self, precision, d_gains, d_info_in, d_info_out, d_info64=None):
        err = _abi.errbuf()
        rc = _abi.lib().ce_batch_run(self.h, int(precision), _ptr(d_gains), _ptr(d_info_in), _ptr(d_info_out),
                                     _ptr(d_info64), self.stream_handle(), err, len(err))
        if rc != 0:
            raise RuntimeError(err.value.decode())

    def pack(self, rows_gains, rows_info):
        """Concatenates per-problem numpy rows into device arrays (problem order)."""
        torch = self.torch
        g = np.concatenate([np.ascontiguousarray(a, dtype=np.float32).ravel() for a in rows_gains])
        i = np.concatenate([np.ascontiguousarray(a, dtype=np.float32).ravel() for a in rows_info])
        dev = f"cuda:{self.device}"
        return torch.from_numpy(g).to(dev), torch.from_numpy(i).to(dev)

    def evaluate(self, precision, gains=None, info=None):
        """Host convenience: runs every problem on its own rows (default: the problem's first n
        rows) and returns per-problem (float rows, double rows)."""
        torch = self.torch
        gs = gains or [p.gains[:n] for p, n in self.problems]
        ins = info or [p.info[:n] for p, n in self.problems]
        d_g, d_i = self.pack(gs, ins)
        d_o = torch.empty_like(d_i)
        d_64 = torch.empty(d_i.shape, dtype=torch.float64, device=d_i.device)
        self.run_device(precision, d_g, d_i, d_o, d_64)
        torch.cuda.synchronize(self.device)
        o32, o64 = d_o.cpu().numpy(), d_64.cpu().numpy()
        out = []
        for k, (p, n) in enumerate(self.problems):
            _, io = self.offsets(k)
            w = ins[k].shape[1]
            out.append((o32[io:io + n * w].reshape(n, w), o64[io:io + n * w].reshape(n, w)))
        return out
