"""Serving loop through ce_server_* (footprint admission, continuous batching on several streams,
p50/p99 job latency).  Host numpy rows in, host numpy rows out; no CPU path."""
import ctypes

import numpy as np

from . import _abi

STATS = ["jobs", "p50_ms", "p99_ms", "mean_ms", "max_ms", "evals_per_s", "batches", "peak_jobs_in_flight",
         "peak_batches_in_flight", "peak_bytes", "budget_bytes"]


class Server:
    def __init__(self, precision, device=0, mem_budget=0, slots=2, max_jobs_per_batch=64):
        err = _abi.errbuf()
        self.h = _abi.lib().ce_server_create(int(device), int(precision), int(mem_budget), int(slots),
                                             int(max_jobs_per_batch), err, len(err))
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.widths = {}

    def close(self):
        if self.h:
            _abi.lib().ce_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, problem, rnd=0, gains=None, info=None):
        g = np.ascontiguousarray(problem.gains if gains is None else gains, dtype=np.float32)
        i = np.ascontiguousarray(problem.info if info is None else info, dtype=np.float32)
        cfg = problem.cfg_struct()
        pl = problem.plant_struct(rnd)
        err = _abi.errbuf()
        jid = _abi.lib().ce_server_submit(self.h, ctypes.byref(cfg), ctypes.byref(pl), g.ctypes.data, i.ctypes.data,
                                          int(g.shape[0]), err, len(err))
        if jid < 0:
            raise ValueError(err.value.decode())
        self.widths[jid] = i.shape
        return jid

    def poll(self, jid):
        """None while pending, else (float rows, double rows)."""
        shape = self.widths[jid]
        o32 = np.empty(shape, np.float32)
        o64 = np.empty(shape, np.float64)
        err = _abi.errbuf()
        rc = _abi.lib().ce_server_poll(self.h, int(jid), o32.ctypes.data, o64.ctypes.data, err, len(err))
        if rc == 0:
            return None
        del self.widths[jid]
        if rc == 1:
            return o32, o64
        raise RuntimeError("job %d: %s" % (jid, err.value.decode() if rc == -2 else "unknown job"))

    def drain(self, timeout_s=600.0):
        return _abi.lib().ce_server_drain(self.h, float(timeout_s)) == 0

    def stats(self):
        buf = (ctypes.c_double * len(STATS))()
        k = _abi.lib().ce_server_stats(self.h, buf, len(STATS))
        return {STATS[i]: buf[i] for i in range(k)}

    def reset_stats(self):
        _abi.lib().ce_server_reset_stats(self.h)
