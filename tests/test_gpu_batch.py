"""Batched evaluation of problems of uneven dimensions, horizons and frequency counts.

Every problem's rows from one batched run must equal the single-problem chain's rows: bit for bit
in FP64 (and therefore the oracle's), and in mixed precision as well (the batched structural
simulation performs the same float operations as the per-problem specialised one).  Graph replay
and autotuned grids must not change any result."""
import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def problems():
    return [workloads.c1(n=300, seed=5), workloads.c2(n=520, seed=6, T=1500, F=96),
            workloads.c3(n=200, seed=7, T=900, F=128), workloads.c1(n=64, seed=8, T=700, F=40),
            workloads.c2(n=130, seed=9, T=2100, F=200)]


def single(engine, p, prec):
    return engine.evaluate(p, precision=prec)


@pytest.fixture(scope="module")
def batch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from controllerengine_amd.batch import Batch
    b = Batch(0)
    ps = problems()
    for p in ps:
        b.add(p)
    yield b, ps
    b.close()


@pytest.mark.parametrize("prec", [PRECISION_FP64, PRECISION_MIXED])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_single(engine, batch, prec, graph):
    b, ps = batch
    b.set_options(use_graph=graph, autotune=False)
    print("batch run prec", prec, "graph", graph, flush=True)
    outs = b.evaluate(prec)
    outs2 = b.evaluate(prec)  # second run: graph replay when enabled
    if graph:
        assert b.info()["used_graph"]
    for p, (o32, o64), (q32, q64) in zip(ps, outs, outs2):
        s32, s64 = single(engine, p, prec)
        assert np.array_equal(o64.view(np.uint64), s64.view(np.uint64)), p.name
        assert np.array_equal(o32.view(np.uint32), s32.view(np.uint32)), p.name
        assert np.array_equal(q64.view(np.uint64), o64.view(np.uint64)), p.name


def test_batch_fp64_matches_oracle(batch):
    b, ps = batch
    b.set_options(use_graph=True)
    outs = b.evaluate(PRECISION_FP64)
    for p, (o32, o64) in zip(ps, outs):
        r32, r64, _ = oracle.evaluate(p)
        assert np.array_equal(o64.view(np.uint64), r64.view(np.uint64)), p.name


def test_batch_autotune_and_update(engine, batch):
    b, ps = batch
    b.set_options(use_graph=True, autotune=True)
    print("batch autotune", flush=True)
    outs = b.evaluate(PRECISION_MIXED)
    assert b.info()["tuned"]
    plan = b.plan()
    print("tuned plan", plan, flush=True)
    assert plan["eig_lin"] in (0, 1, 2) and plan["eig_nom"] in (0, 1, 2) and plan["sim"] in (0, 1)
    for p, (o32, o64) in zip(ps, outs):
        s32, s64 = single(engine, p, PRECISION_MIXED)
        assert np.array_equal(o64.view(np.uint64), s64.view(np.uint64)), p.name
    # new plant values for problem 1 (same shape): the graph is replayed with the new descriptor
    q = workloads.c2(n=520, seed=6, T=1500, F=96)
    q.d_a_i = q.d_a_i * np.float32(1.05)
    q.d_a = (np.eye(q.nd, dtype=np.float32)[None] + q.d_a_i).astype(np.float32)
    b.update(1, q)
    print("batch update", flush=True)
    outs = b.evaluate(PRECISION_MIXED)
    assert b.info()["used_graph"]
    s32, s64 = single(engine, q, PRECISION_MIXED)
    assert np.array_equal(outs[1][1].view(np.uint64), s64.view(np.uint64))
    b.update(1, ps[1])
    b.set_options(use_graph=True, autotune=False)


@pytest.mark.parametrize("plan", [(1, 1, 0), (2, 2, 1), (0, 2, 0), (2, 1, 1)])
@pytest.mark.parametrize("prec", [PRECISION_FP64, PRECISION_MIXED])
def test_batch_plans_equal_single(engine, prec, plan):
    """Every stage plan gives the single chain's rows bit for bit: the eigen stages with 8 lanes per
    candidate at orders 5-16 (plan 1) or as one fused launch over orders 4, 8 and 16 (plan 2), and
    the simulation as one fused launch over every structure (sim plan 1)."""
    from controllerengine_amd.batch import Batch
    b = Batch(0)
    try:
        ps = problems()
        for p in ps:
            b.add(p)
        b.set_plan(*plan)
        b.set_options(use_graph=True, autotune=False)
        print("batch plan", plan, "prec", prec, flush=True)
        outs = b.evaluate(prec)
        assert b.plan() == dict(eig_lin=plan[0], eig_nom=plan[1], sim=plan[2])
        for p, (o32, o64) in zip(ps, outs):
            s32, s64 = single(engine, p, prec)
            assert np.array_equal(o64.view(np.uint64), s64.view(np.uint64)), (p.name, plan)
    finally:
        b.close()


def order_jobs():
    # serving-job plants of every order 2..8 in one batch (the batched sweep, simulation and eigen
    # kernels of each order), each drawn from the smallest seed whose plant has that order
    out = []
    for ns in range(2, 9):
        s = 0
        while int(np.random.default_rng(s).integers(2, 9)) != ns:
            s += 1
        out.append(workloads.c5_job(np.random.default_rng(s), n=200, tmin=400, tmax=900, fmin=64, fmax=160))
    return out


@pytest.mark.parametrize("prec", [PRECISION_FP64, PRECISION_MIXED])
def test_batch_every_order_equals_single(engine, prec):
    from controllerengine_amd.batch import Batch
    ps = order_jobs()
    b = Batch(0)
    try:
        for p in ps:
            b.add(p)
        b.set_options(use_graph=True, autotune=False)
        print("batch of orders", [p.ns for p in ps], "prec", prec, flush=True)
        outs = b.evaluate(prec)
        for p, (o32, o64) in zip(ps, outs):
            s32, s64 = single(engine, p, prec)
            assert np.array_equal(o64.view(np.uint64), s64.view(np.uint64)), (p.ns, p.name)
            assert np.array_equal(o32.view(np.uint32), s32.view(np.uint32)), (p.ns, p.name)
    finally:
        b.close()


def test_c4_serial_rounds_in_one_batch(engine):
    """The batched Monte Carlo workload (BASELINE config 4 at reduced size): serial-mode chain-drive
    plants of 2-8 states with perturbed plants every round (reference serial rounds,
    src/main.cu:467-470, src/config.cu:186-192), all in one batch with graph replay and autotuned
    grids; each round's plants are uploaded with ce_batch_update.  FP64 rows equal the oracle's
    bit for bit every round; mixed rows keep every branch and meet the margin tolerances, and equal
    the single-problem chain's mixed rows."""
    from controllerengine_amd.batch import Batch
    from controllerengine_amd.metrics import compare
    w = workloads.C4(seed=13, n_plants=10, rounds=3, nmin=64, nmax=400, tmin=600, tmax=1600, F=96)
    assert len({p.ns for p in w.plants}) >= 3
    b = Batch(0)
    try:
        for p in w.plants:
            b.add(p, rnd=0)
        b.set_options(use_graph=True, autotune=True)
        for r in range(w.rounds):
            for j, p in enumerate(w.plants):
                b.update(j, p, rnd=r)
            print("c4 round", r, "fp64", flush=True)
            outs = b.evaluate(PRECISION_FP64)
            assert b.info()["used_graph"] or r == 0
            ref = []
            for p, (o32, o64) in zip(w.plants, outs):
                _, r64, _ = oracle.evaluate(p, rnd=r)
                ref.append(r64)
                same = (o64 == r64) | (np.isnan(o64) & np.isnan(r64))
                assert same.all(), (p.name, r, int((~same).sum()))
            print("c4 round", r, "mixed", flush=True)
            outs = b.evaluate(PRECISION_MIXED)
            for p, (o32, o64), r64 in zip(w.plants, outs, ref):
                d = p.derived()
                am, sc = d.info_allmargin, d.info_scoring
                assert np.array_equal(o64[:, am + 3], r64[:, am + 3]), p.name
                assert np.array_equal(o64[:, sc] >= 1e6, r64[:, sc] >= 1e6), p.name
                cmp = compare(p, r64, o64)
                assert cmp["max_dgm_db"] <= 0.01 and cmp["max_dpm_deg"] <= 0.01 and cmp["max_db"] <= 1e-4, (p.name, cmp)
                assert cmp["passfail_mismatch"] == 0, (p.name, cmp)
            if r == w.rounds - 1:
                for p, (o32, o64) in list(zip(w.plants, outs))[:4]:
                    _, s64 = engine.evaluate(p, precision=PRECISION_MIXED, rnd=r)
                    assert np.array_equal(o64.view(np.uint64), s64.view(np.uint64)), p.name
    finally:
        b.close()
