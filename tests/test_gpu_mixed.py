"""Mixed-precision chain against the FP64 chain on the same inputs (GPU).

Tolerances (stated targets of the mixed mode): |dGM| <= 0.01 dB, |dPM| <= 0.01 deg, |db| <= 1e-4 on
every candidate whose margins both modes evaluate; the pass/fail flags (linearly stable, nominally
stable, every margin hinge at zero) identical; the behaviour score (modal fit of the float
trajectory) reported against FP64 and bounded loosely."""
import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
from controllerengine_amd import metrics
from oracle import oracle


def fit_condition(problem, n):
    """Condition number of the column-scaled normal equations of each candidate's modal fit
    (oracle's Gram dump), over the 2u basis columns actually used."""
    d = problem.derived()
    nt = d.total_dim
    _, o64, dg = oracle.evaluate(problem, diag=True)
    u = o64[:, d.info_modal + 3 * nt].astype(int)
    kappa = np.full(n, np.inf)
    for i in range(n):
        k = 2 * u[i]
        if k == 0:
            continue
        a = dg["ata"][i].reshape(2 * nt, 2 * nt)[:k, :k]
        g = np.tril(a) + np.tril(a, -1).T
        dd = np.sqrt(np.diag(g))
        dd[dd <= 1e-6] = 1.0
        sg = g / dd[:, None] / dd[None, :]
        keep = np.diag(g) > 0
        if keep.sum() == 0:
            continue
        kappa[i] = np.linalg.cond(sg[np.ix_(keep, keep)])
    return kappa

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,n", [("c1", 1024), ("c2", 2048), ("c3", 1024)])
def test_mixed_margins_match_fp64(engine, name, n):
    p = workloads.CONFIGS[name](n=n)
    print("case", name, n, flush=True)
    _, ref = engine.evaluate(p, PRECISION_FP64)
    _, got = engine.evaluate(p, PRECISION_MIXED)
    rep = metrics.compare(p, ref, got)
    print(rep, flush=True)
    assert rep["n_margin_rows"] > 0
    assert rep["max_dgm_db"] <= 0.01
    assert rep["max_dpm_deg"] <= 0.01
    assert rep["max_db"] <= 1e-4
    assert rep["passfail_mismatch"] == 0
    assert rep["stable_mismatch"] == 0
    # the margin part of the pass rule alone (the full rule passes few or no candidates of these
    # workloads): the same candidates meet GM, PM and b in both chains
    assert rep["margin_pass_mismatch"] == 0
    # behaviour score from the float trajectory: close wherever the least-squares fit is well posed
    kappa = fit_condition(p, n)
    d = p.derived()
    sc = d.info_scoring
    both = (ref[:, sc + 1] > 0) & (ref[:, sc + 1] < 1e6) & (got[:, sc + 1] < 1e6)
    rel = np.abs(got[:, sc + 1] - ref[:, sc + 1]) / np.maximum(np.abs(ref[:, sc + 1]), 1e-12)
    well = both & (kappa < 1e6)
    print("behaviour rel err: well-posed fits (kappa<1e6) n=%d max %.3g; kappa>=1e6 n=%d median %.3g"
          % (well.sum(), rel[well].max() if well.any() else 0.0, (both & ~well).sum(),
             np.median(rel[both & ~well]) if (both & ~well).any() else 0.0), flush=True)
    if well.any():
        assert rel[well].max() < 1e-2


def test_mixed_specialised_and_interpreted_agree(engine, monkeypatch):
    p = workloads.c2(n=512)
    monkeypatch.setenv("CE_JIT", "1")
    _, a = engine.evaluate(p, PRECISION_MIXED)
    monkeypatch.setenv("CE_JIT", "0")
    _, b = engine.evaluate(p, PRECISION_MIXED)
    # the specialised float kernel leaves out terms with a zero coefficient: identical results
    # for finite trajectories (a trajectory that overflows may end as inf instead of NaN)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    assert same.all(), "float simulation kernels differ in %d cells" % int((~same).sum())


@pytest.mark.parametrize("name,n,lanes", [("c2", 65536, 2), ("c1", 1024, 4), ("c3", 1024, 4)])
def test_mixed_lane_groups_equal_interpreted(engine, name, n, lanes, monkeypatch):
    """The float literal kernel splits a candidate's step over 2 or 4 lanes when the live
    candidates leave SIMDs idle (c2's full block: ~27k simulated candidates, 2 lanes; 1,024
    candidates: 4 lanes).  Its rows equal the interpreted one-lane kernel's bit for bit."""
    p = workloads.CONFIGS[name](n=n)
    print("case", name, n, "lanes", lanes, flush=True)
    monkeypatch.setenv("CE_JIT", "1")
    _, a = engine.evaluate(p, PRECISION_MIXED)
    monkeypatch.setenv("CE_JIT", "0")
    _, b = engine.evaluate(p, PRECISION_MIXED)
    sc = p.derived().info_scoring
    live = int((a[:, sc] < float(p.config["unstable_err"])).sum())
    cap = 1024  # SIMDs of an MI355X
    assert (live <= 16 * cap) == (lanes == 4) and (16 * cap < live <= 32 * cap) == (lanes == 2), live
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    assert same.all(), "lane-group and interpreted float kernels differ in %d cells" % int((~same).sum())


@pytest.mark.parametrize("name,n", [("c2", 4096), ("c3", 1024), ("c3", 262144)])
def test_mixed_two_streams_equal_one_stream(engine, name, n):
    """The default mixed chain runs the nominal eigensolve, the sweep and the refinement on a second
    stream beside a simulation that holds whole SIMDs and signals its start; with per-kernel timing on,
    every kernel runs in order on one stream with the plain simulation kernel.  Same rows, bit for bit.
    The full c3 block (149,427 simulated candidates: two full rounds of one-lane waves on the
    device's 1,024 SIMDs and a remainder of 287 waves) also splits the simulation: the remainder runs
    with 2 lanes per candidate on a third stream, with its own fit."""
    p = workloads.CONFIGS[name](n=n)
    print("case", name, n, flush=True)
    _, two = engine.evaluate(p, PRECISION_MIXED)
    engine.set_profiling(True)
    try:
        _, one = engine.evaluate(p, PRECISION_MIXED)
    finally:
        engine.set_profiling(False)
    _, again = engine.evaluate(p, PRECISION_MIXED)
    for other in (one, again):
        same = (two == other) | (np.isnan(two) & np.isnan(other))
        assert same.all(), "rows differ in %d cells" % int((~same).sum())
