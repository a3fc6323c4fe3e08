"""Serving loop: jobs of uneven dimensions submitted while others run; every job's rows equal the
rows of the same job evaluated alone by the single-problem chain (bit for bit, both precisions),
under a memory budget that forces several batches and overlapping batches in flight."""
import threading
import time

import numpy as np
import pytest

from controllerengine_amd import PRECISION_FP64, PRECISION_MIXED, workloads

pytestmark = pytest.mark.gpu


def jobs(k=10, seed=11):
    rng = np.random.default_rng(seed)
    return [workloads.c5_job(rng, nmin=64, nmax=700, tmin=300, tmax=1200, fmin=64, fmax=200) for _ in range(k)]


@pytest.mark.parametrize("prec", [PRECISION_FP64, PRECISION_MIXED])
def test_server_results_equal_alone(engine, prec):
    from controllerengine_amd.server import Server
    js = jobs()
    alone = [engine.evaluate(j, precision=prec) for j in js]
    # a budget of about three jobs' footprint: admission has to split the stream into batches
    srv = Server(prec, mem_budget=3 * 700 * 40000, slots=2, max_jobs_per_batch=4)
    ids = []
    print("server submit prec", prec, flush=True)

    def producer():
        for j in js:
            ids.append(srv.submit(j))
            time.sleep(0.002)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    assert srv.drain(120.0)
    st = srv.stats()
    print("server stats", st, flush=True)
    assert st["jobs"] == len(js) and st["batches"] >= 2
    for jid, (a32, a64) in zip(ids, alone):
        r = srv.poll(jid)
        assert r is not None
        assert np.array_equal(r[1].view(np.uint64), a64.view(np.uint64))
        assert np.array_equal(r[0].view(np.uint32), a32.view(np.uint32))
    srv.close()


def test_server_poll_pending_and_unknown():
    from controllerengine_amd.server import Server
    srv = Server(PRECISION_MIXED, slots=1)
    j = jobs(1, seed=5)[0]
    jid = srv.submit(j)
    assert srv.drain(120.0)
    r = srv.poll(jid)
    assert r is not None and r[0].shape == j.info.shape
    with pytest.raises(KeyError):
        srv.poll(jid)
    srv.close()
