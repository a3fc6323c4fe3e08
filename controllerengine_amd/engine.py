"""Python host mirror of the engine: device-resident evaluation through the C ABI.

torch provides device memory and the stream handle only; every stage runs in the HIP kernels of
libControllerEngine.so.  There is no CPU fallback: without the library or a GPU the calls raise.
"""
import ctypes

import numpy as np

from . import _abi


def _dev_ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Engine:
    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("Engine needs a HIP device")
        self.torch = torch
        self.device = device
        err = _abi.errbuf()
        self.h = _abi.lib().ce_engine_create(int(device), err, len(err))
        if not self.h:
            raise RuntimeError(err.value.decode())
        self._diag = None

    def close(self):
        if self.h:
            _abi.lib().ce_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream_handle(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def to_device(self, arr, dtype=None):
        t = self.torch.from_numpy(np.ascontiguousarray(arr))
        return t.to(f"cuda:{self.device}")

    def set_diagnostics(self, n, cfg, enable=True):
        torch = self.torch
        if not enable:
            _abi.lib().ce_engine_set_diagnostics(self.h, None, None, None, None, None)
            self._diag = None
            return None
        nt = cfg.statespace_dim + cfg.discrete_dim
        dev = f"cuda:{self.device}"
        d = dict(y=torch.zeros((n, cfg.time), dtype=torch.float32, device=dev),
                 lina=torch.zeros((n, nt * nt), dtype=torch.float32, device=dev),
                 eig=torch.zeros((n, 2 * nt), dtype=torch.float32, device=dev),
                 ata=torch.zeros((n, 4 * nt * nt), dtype=torch.float32, device=dev),
                 atb=torch.zeros((n, 2 * nt), dtype=torch.float32, device=dev))
        _abi.lib().ce_engine_set_diagnostics(self.h, _dev_ptr(d["y"]), _dev_ptr(d["lina"]), _dev_ptr(d["eig"]),
                                             _dev_ptr(d["ata"]), _dev_ptr(d["atb"]))
        self._diag = d
        return d

    def evaluate_device(self, cfg, plant, precision, d_gains, d_info_in, d_info_out, d_info64=None):
        """Enqueue the chain on torch's current stream; all tensors on this device."""
        n = int(d_gains.shape[0])
        err = _abi.errbuf()
        rc = _abi.lib().ce_evaluate(self.h, ctypes.byref(cfg), ctypes.byref(plant), int(precision),
                                    _dev_ptr(d_gains), _dev_ptr(d_info_in), _dev_ptr(d_info_out),
                                    _dev_ptr(d_info64), n, self.stream_handle(), err, len(err))
        if rc != 0:
            raise RuntimeError(err.value.decode())

    def evaluate(self, problem, precision=_abi.PRECISION_FP64, rnd=0, gains=None, info=None, want64=True):
        """Host convenience: copy rows in, run, copy rows out (float32 rows and, optionally, double rows)."""
        torch = self.torch
        g = self.to_device(np.asarray(problem.gains if gains is None else gains, dtype=np.float32))
        i_in = self.to_device(np.asarray(problem.info if info is None else info, dtype=np.float32))
        i_out = torch.empty_like(i_in)
        i64 = torch.empty(i_in.shape, dtype=torch.float64, device=i_in.device) if want64 else None
        self.evaluate_device(problem.cfg_struct(), problem.plant_struct(rnd), precision, g, i_in, i_out, i64)
        torch.cuda.synchronize(self.device)
        return i_out.cpu().numpy(), (i64.cpu().numpy() if want64 else None)

    def simulate_kalman(self, problem, gains, precision=_abi.PRECISION_FP64, rnd=0):
        """Kalman-filter simulation (ce_simulate_kalman) of gains rows [K, N, R, Q] (2 ns + 2
        floats); returns the trajectories, n x time (double in FP64, float in mixed)."""
        torch = self.torch
        g = self.to_device(np.ascontiguousarray(gains, dtype=np.float32))
        n = int(g.shape[0])
        T = int(problem.config["time"])
        mixed = precision == _abi.PRECISION_MIXED
        y = torch.empty((n, T), dtype=torch.float32 if mixed else torch.float64, device=g.device)
        err = _abi.errbuf()
        rc = _abi.lib().ce_simulate_kalman(self.h, ctypes.byref(problem.cfg_struct()),
                                           ctypes.byref(problem.plant_struct(rnd)), int(precision), _dev_ptr(g), n,
                                           None if mixed else _dev_ptr(y), _dev_ptr(y) if mixed else None,
                                           self.stream_handle(), err, len(err))
        if rc != 0:
            raise RuntimeError(err.value.decode())
        torch.cuda.synchronize(self.device)
        return y.cpu().numpy()

    def stage_times(self):
        buf = (ctypes.c_double * 8)()
        k = _abi.lib().ce_engine_stage_times(self.h, buf, 8)
        return [buf[i] for i in range(k)]

    def last_sim_kernel(self):
        """Kernel name of the last evaluation's simulation launch."""
        buf = ctypes.create_string_buffer(64)
        _abi.lib().ce_engine_last_sim_kernel(self.h, buf, 64)
        return buf.value.decode()

    def last_sim_split(self):
        """SIMD count the last call split its simulation by (0: not split); the one-lane kernel took
        the list slots below sim_split_slots(live, cap), a 2-lane kernel the rest."""
        return int(_abi.lib().ce_engine_last_sim_split(self.h))

    def set_sim_history(self, calls):
        """Record the simulation kernel's duration (HIP events on the launch stream) for the next
        `calls` evaluations (a ring; 0 turns it off)."""
        if _abi.lib().ce_engine_set_sim_history(self.h, int(calls)) != 0:
            raise RuntimeError("ce_engine_set_sim_history failed")

    def sim_history(self):
        """Durations (ms) of the simulation kernels of the recorded evaluations, oldest first."""
        buf = (ctypes.c_double * 4096)()
        k = _abi.lib().ce_engine_sim_history(self.h, buf, 4096)
        if k < 0:
            raise RuntimeError("ce_engine_sim_history failed")
        return [buf[i] for i in range(k)]

    def jit_compiles(self):
        """hipRTC compilations of this engine (simulation kernels)."""
        return int(_abi.lib().ce_engine_jit_compiles(self.h))

    def sweep_counts(self):
        """(FP64-refined points, candidates swept entirely in FP64) of the last mixed evaluate."""
        r, f = ctypes.c_int(0), ctypes.c_int(0)
        if _abi.lib().ce_engine_sweep_counts(self.h, ctypes.byref(r), ctypes.byref(f)) != 0:
            raise RuntimeError("ce_engine_sweep_counts: the last evaluation was not mixed")
        return r.value, f.value

    def set_profiling(self, on=True):
        _abi.lib().ce_engine_set_profiling(self.h, 1 if on else 0)

    def kernel_times(self):
        """[(kernel name, ms)] of the last evaluate with profiling on, in launch order."""
        names = ctypes.create_string_buffer(2048)
        buf = (ctypes.c_double * 64)()
        k = _abi.lib().ce_engine_kernel_times(self.h, names, len(names), buf, 64)
        if k <= 0:
            return []
        return list(zip(names.value.decode().split(","), [buf[i] for i in range(k)]))


def sim_split_slots(live, cap):
    """First list slot of the split simulation's remainder for `live` simulated candidates and `cap`
    SIMDs (`live` when the simulation is not split)."""
    return int(_abi.lib().ce_sim_split_slots(int(live), int(cap)))
