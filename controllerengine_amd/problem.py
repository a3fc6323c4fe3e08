"""A tuning problem in the reference's file format: ControllerEngine.config + the .bin matrices,
instruction tables, gains (CUDA_MX) and INFO rows.

Layouts follow reference README.md:39-91 and src/main.cu:341-400 (row-major float32 / int32,
lower-case .bin names).  A Problem can be written to a directory for RunControllerEngine(), or
turned into the ce_config / ce_plant structs of the extended C API.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import _abi

CONFIG_KEYS = [
    "statespace_dim", "discrete_dim", "nonlin_dim", "n_params", "input_size", "dc_instructions",
    "u_instructions", "nl_instructions", "n_cores", "n_rounds", "mode", "time", "duration", "ref_step_time",
    "ref_amplitude", "fit_start", "fit_end", "steady_start", "bandwidth", "damping", "overshoot",
    "steady_state_error", "crossover_coeff", "target_gm", "target_pm", "target_b", "eig_fail_err",
    "unstable_err", "freq", "freq_min", "dump_diagnostics",
]


def fmt(v):
    if isinstance(v, (list, tuple)):
        return ",".join(fmt(x) for x in v)
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    return str(v)


def encode(op, a, b):
    """Base-36 instruction entry (reference README.md:65-85): op*36^2 + a*36 + b.
    a/b: 0 = state under the entry's column, 35 = running value, k in 1..34 = BASE_PARAMS[k-1]."""
    return int(op) * 1296 + int(a) * 36 + int(b)


@dataclass
class Problem:
    config: dict
    ss_a: np.ndarray
    ss_a_i: np.ndarray
    ss_b: np.ndarray
    ss_c: np.ndarray
    ss_d: np.ndarray
    d_a: np.ndarray          # (R, nd, nd) in serial mode, else (1, nd, nd)
    d_a_i: np.ndarray
    d_bu: np.ndarray
    d_bw: np.ndarray         # (R, nd, nw)
    d_c: np.ndarray
    base_params: np.ndarray  # (R, n_params)
    dc_info: np.ndarray      # (dc_instructions, nd) int32
    u_info: np.ndarray       # (u_instructions, ns) int32
    nl_info: np.ndarray      # (sum nl, nd+1) int32
    gains: np.ndarray        # (n_rounds*n_threads or n, 2ns+1) float32
    info: np.ndarray         # (same rows, info_w) float32
    name: str = "problem"
    meta: dict = field(default_factory=dict)

    # ---------------------------------------------------------------------------------------
    @property
    def ns(self):
        return int(self.config["statespace_dim"])

    @property
    def nd(self):
        return int(self.config["discrete_dim"])

    def config_text(self):
        lines = []
        for k in CONFIG_KEYS:
            if k in self.config:
                lines.append("%s = %s" % (k, fmt(self.config[k])))
        return "\n".join(lines) + "\n"

    def cfg_struct(self):
        cfg = _abi.CeConfig()
        err = _abi.errbuf()
        rc = _abi.lib().ce_config_parse(self.config_text().encode(), b"ControllerEngine.config",
                                        ctypes.byref(cfg), err, len(err))
        if rc != 0:
            raise ValueError(err.value.decode())
        return cfg

    def derived(self):
        d = _abi.CeDerived()
        err = _abi.errbuf()
        cfg = self.cfg_struct()
        if _abi.lib().ce_config_derive(ctypes.byref(cfg), ctypes.byref(d), err, len(err)) != 0:
            raise ValueError(err.value.decode())
        return d

    def plant_struct(self, rnd=0):
        p = _abi.CePlant()
        ns, nd = self.ns, self.nd
        r = rnd if self.d_a.shape[0] > 1 else 0

        def put(dst, arr, dtype=np.float32):
            flat = np.ascontiguousarray(arr, dtype=dtype).ravel()
            np.frombuffer(dst, dtype=dtype)[:flat.size] = flat

        put(p.ss_a, self.ss_a.reshape(ns, ns))
        put(p.ss_a_i, self.ss_a_i.reshape(ns, ns))
        put(p.ss_b, self.ss_b)
        put(p.ss_c, self.ss_c)
        p.ss_d = float(np.float32(np.ravel(self.ss_d)[0]))
        put(p.d_a, self.d_a[r])
        put(p.d_a_i, self.d_a_i[r])
        put(p.d_bu, self.d_bu[r])
        put(p.d_bw, self.d_bw[r])
        put(p.d_c, self.d_c[r])
        put(p.base_params, self.base_params[r if self.base_params.shape[0] > 1 else 0])
        put(p.dc_info, self.dc_info, np.int32)
        put(p.u_info, self.u_info, np.int32)
        put(p.nl_info, self.nl_info, np.int32)
        return p

    def round_rows(self, rnd, n_per_round):
        sl = slice(rnd * n_per_round, (rnd + 1) * n_per_round)
        return self.gains[sl], self.info[sl]

    def write_dir(self, path):
        os.makedirs(path, exist_ok=True)
        with open(os.path.join(path, "ControllerEngine.config"), "w") as f:
            f.write(self.config_text())

        def w(name, arr, dtype=np.float32):
            np.ascontiguousarray(arr, dtype=dtype).tofile(os.path.join(path, name))

        w("STATESPACE_A.bin", self.ss_a)
        w("STATESPACE_A_I.bin", self.ss_a_i)
        w("STATESPACE_B.bin", self.ss_b)
        w("STATESPACE_C.bin", self.ss_c)
        w("STATESPACE_D.bin", np.ravel(self.ss_d)[:1])
        w("DISCRETE_A.bin", self.d_a)
        w("DISCRETE_A_I.bin", self.d_a_i)
        w("DISCRETE_BU.bin", self.d_bu)
        w("DISCRETE_BW.bin", self.d_bw)
        w("DISCRETE_C.bin", self.d_c)
        w("BASE_PARAMS.bin", self.base_params)
        w("DC_INFO.bin", self.dc_info, np.int32)
        w("U_INFO.bin", self.u_info, np.int32)
        w("NL_INFO.bin", self.nl_info, np.int32)
        w("CUDA_MX.bin", self.gains)
        w("INFO_MX.bin", self.info)
        return path

    def subset(self, idx):
        """A copy holding only candidate rows idx (same plant and config)."""
        import copy
        q = copy.copy(self)
        q.gains = np.ascontiguousarray(self.gains[idx])
        q.info = np.ascontiguousarray(self.info[idx])
        return q


def info_width(config):
    nt = int(config["statespace_dim"]) + int(config["discrete_dim"])
    return int(config.get("input_size", 0)) + 12 + 3 * nt + 1
