"""Randomised parity over serving-job plants (GPU): random stable plants of order 2..8 with 0..4
nonlinear terms, random horizons and frequency counts.  For every job, the FP64 rows equal the
oracle's bit for bit, and the mixed chain is within the mixed-mode tolerances of the FP64 chain
(|dGM| <= 0.01 dB, |dPM| <= 0.01 deg, |db| <= 1e-4, identical pass/fail and stability flags).

CE_RANDOM_JOBS (default 24) and CE_RANDOM_JOB_N (default 512) scale the run."""
import os

import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, metrics, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu

JOBS = int(os.environ.get("CE_RANDOM_JOBS", "24"))
N = int(os.environ.get("CE_RANDOM_JOB_N", "512"))


def test_random_jobs(engine):
    bad = []
    for seed in range(JOBS):
        p = workloads.c5_job(np.random.default_rng(5000 + seed), n=N, tmin=300, tmax=2000, fmin=64, fmax=512)
        print("case random job seed %d ns %d nw %d T %d F %d" % (seed, p.ns, p.config["nonlin_dim"],
                                                               p.config["time"], p.config["freq"]), flush=True)
        _, o64, _ = oracle.evaluate(p)
        _, r64 = engine.evaluate(p, PRECISION_FP64)
        same = (r64 == o64) | (np.isnan(r64) & np.isnan(o64))
        _, m64 = engine.evaluate(p, PRECISION_MIXED)
        rep = metrics.compare(p, r64, m64)
        ok = (same.all() and rep["max_dgm_db"] <= 0.01 and rep["max_dpm_deg"] <= 0.01 and rep["max_db"] <= 1e-4
              and rep["passfail_mismatch"] == 0 and rep["stable_mismatch"] == 0)
        if not ok:
            bad.append((seed, int((~same).sum()), rep))
    assert not bad, bad
