"""Writes the RunControllerEngine() fixtures under tests/golden/.

Each fixture is one compressed .npz holding every input file of the reference's file protocol
(ControllerEngine.config text plus the .bin files, reference README.md:39-91 and src/main.cu:341-462)
and the expected INFO_MX_OUT.bin rows.  The expected rows come from the CPU oracle (FP64 chain),
whose stages are pinned by the closed-form and known-answer tests in tests/.  The reference holds no
sample inputs or outputs of its own.

  c1_batch    mass-spring-damper, SAT clamp in the U chain, batch mode, 1 round of 1536 candidates
  c2_serial   two-mass drive with Coulomb/quadratic friction NL chains, serial mode, 2 rounds with a
              different plant per round, input_size=3 pass-through columns, INFO_MX input with
              stale values in the output columns

Run:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from controllerengine_amd import workloads  # noqa: E402
from controllerengine_amd.problem import info_width  # noqa: E402
from oracle import oracle  # noqa: E402

FILES = ["STATESPACE_A.bin", "STATESPACE_A_I.bin", "STATESPACE_B.bin", "STATESPACE_C.bin", "STATESPACE_D.bin",
         "DISCRETE_A.bin", "DISCRETE_A_I.bin", "DISCRETE_BU.bin", "DISCRETE_BW.bin", "DISCRETE_C.bin",
         "BASE_PARAMS.bin", "DC_INFO.bin", "U_INFO.bin", "NL_INFO.bin", "CUDA_MX.bin", "INFO_MX.bin"]


def c1_batch():
    return workloads.c1(n=1536, seed=11, T=2000, F=200)


def c2_serial():
    p = workloads.c2(n=2 * 1536, seed=12, T=1000, F=128)
    cfg = dict(p.config)
    cfg.update(mode="serial", n_rounds=2, input_size=3)
    p.config = cfg
    # round 1: heavier load mass (m2 = 0.7): rebuild the discrete model rows that depend on it
    dt = 5e-4
    m2, ks, c2d = 0.7, 50.0, 0.2
    a1 = p.d_a_i[0].astype(np.float64).copy()
    a1[3] = dt * np.array([ks / m2, c2d / m2, -ks / m2, -c2d / m2])
    bw1 = p.d_bw[0].astype(np.float64).copy()
    bw1[3, 1] = -dt / m2
    p.d_a_i = np.stack([p.d_a_i[0], np.float32(a1)])
    p.d_a = np.stack([p.d_a[0], np.float32(np.eye(4) + a1)])
    p.d_bu = np.stack([p.d_bu[0], p.d_bu[0]])
    p.d_bw = np.stack([p.d_bw[0], np.float32(bw1)])
    p.d_c = np.stack([p.d_c[0], p.d_c[0]])
    bp1 = p.base_params[0].copy()
    bp1[1] = np.float32(0.25)             # Coulomb level of round 1
    p.base_params = np.stack([p.base_params[0], bp1])
    rng = np.random.default_rng(13)
    p.info = np.zeros((p.gains.shape[0], info_width(cfg)), np.float32)
    p.info[:, :3] = rng.standard_normal((p.gains.shape[0], 3))      # pass-through input columns
    p.info[::7, 3:] = 5.0                                            # stale outputs to be overwritten
    return p


def files_of(p):
    tmp = os.path.join(HERE, "_tmp_fixture")
    p.write_dir(tmp)
    out = {"config": np.frombuffer(p.config_text().encode(), dtype=np.uint8)}
    for f in FILES:
        with open(os.path.join(tmp, f), "rb") as fh:
            out[f] = np.frombuffer(fh.read(), dtype=np.uint8)
        os.remove(os.path.join(tmp, f))
    os.remove(os.path.join(tmp, "ControllerEngine.config"))
    os.rmdir(tmp)
    return out


def expected_rows(p):
    d = p.derived()
    nthr = d.n_threads
    rows = []
    for r in range(int(p.config["n_rounds"])):
        g, info = p.round_rows(r, nthr)
        out32, _, _ = oracle.evaluate(p, rnd=r, gains=g, info=info)
        rows.append(out32)
    return np.concatenate(rows, axis=0)


def main():
    for name, make in [("c1_batch", c1_batch), ("c2_serial", c2_serial)]:
        p = make()
        data = files_of(p)
        data["INFO_MX_OUT.bin"] = np.frombuffer(expected_rows(p).tobytes(), dtype=np.uint8)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
        print(name, "written")


if __name__ == "__main__":
    main()
