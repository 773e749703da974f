"""Docking service with continuous batching (needs a GPU).

Every job's results must equal, bit for bit, those of the same poses minimized
alone with att_minimize: the engine runs the same kernels on the job's own
buffers, whatever else is admitted beside it.  The tests drive jobs of
different pairs (rigid and with modes, different sizes) through one server
at once, under a budget that forces admission to wait and grids to be evicted,
and through the error paths."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    from gpuattract_amd import serve
    return serve


@pytest.fixture(scope="module")
def ctx():
    from gpuattract_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def pairs(W):
    return [
        W.make_workload("serve a", 91, 500, 120, 60),
        W.make_workload("serve b", 92, 800, 200, 150, 4, 2, with_lig_grid=False),
        W.make_workload("serve c", 93, 500, 90, 40, 0, 3, with_lig_grid=False),
        W.make_workload("serve d", 94, 1200, 300, 90),
    ]


def alone(ctx, wl, evals):
    hs = ctx.upload_workload(wl)
    try:
        return ctx.minimize(hs, wl.dofs, evals)
    finally:
        for k in ("rec", "lig", "grid", "params"):
            ctx.release(hs[k])


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[:4], b[:4]))


def test_serve_alone_and_in_the_mix(ctx, S, pairs):
    evals = [60, 40, 80, 30]
    ref = [alone(ctx, wl, e) for wl, e in zip(pairs, evals)]
    with S.Server(0) as srv:
        ids = [srv.add_workload(wl) for wl in pairs]
        # each job alone on the server
        for i, (wl, e) in enumerate(zip(pairs, evals)):
            r = srv.wait(srv.submit(ids[i], wl.dofs, e))
            assert same(r, ref[i]), f"pair {i} alone on the server"
        # all at once, twice over, submitted from two threads
        jobs, lock = [], threading.Lock()

        def submitter(order):
            for i in order:
                j = srv.submit(ids[i], pairs[i].dofs, evals[i])
                with lock:
                    jobs.append((i, j))

        ts = [threading.Thread(target=submitter, args=(o,)) for o in ([0, 1, 2, 3], [3, 2, 1, 0])]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        for i, j in jobs:
            r = srv.wait(j)
            assert same(r, ref[i]), f"pair {i} in the mix"
            assert r[4]["grid_cached"] == 1
        st = srv.stats()
        print(st)
        assert st["jobs_done"] == 12
        assert st["max_running"] >= 2  # jobs overlapped
        assert st["grid_misses"] == 4 and st["grid_hits"] == 8
        assert st["evictions"] == 0
        assert 0 < st["p50_ms"] <= st["p99_ms"]
        assert st["poses_done"] == 3 * sum(wl.dofs.shape[0] for wl in pairs)


def test_serve_budget_admission_and_eviction(ctx, S, pairs):
    wls = [pairs[0], pairs[2]]  # two receptors of 500 beads
    evals = [50, 50]
    ref = [alone(ctx, wl, e) for wl, e in zip(wls, evals)]
    # footprint of one job with its receptor and ligand
    with S.Server(0) as srv:
        ids = [srv.add_workload(wl) for wl in wls]
        peaks, held = [], []
        for i in range(2):
            held.append(srv.wait(srv.submit(ids[i], wls[i].dofs, evals[i]))[4]["bytes"])
            peaks.append(srv.stats()["peak_bytes"])
    # the second peak holds the first pair's uploads, not its job memory
    one = max(peaks[0], peaks[1] - (peaks[0] - held[0]))
    budget = int(one * 1.3)  # one pair at a time
    with S.Server(0, budget) as srv:
        ids = [srv.add_workload(wl) for wl in wls]
        jobs = [(k % 2, srv.submit(ids[k % 2], wls[k % 2].dofs, evals[k % 2])) for k in range(4)]
        res = [(i, srv.wait(j)) for i, j in jobs]
        for i, r in res:
            assert same(r, ref[i])
        st = srv.stats()
        print(st, [r[4]["admit_seq"] for _, r in res])
        # one receptor fits at a time: the second pair's first job waits for
        # memory, and the third job (first pair, receptor resident) may pass it
        assert st["peak_bytes"] <= budget
        assert st["evictions"] >= 1 and st["grid_misses"] >= 2
        assert sorted(r[4]["admit_seq"] for _, r in res) == [1, 2, 3, 4]


def test_serve_later_jobs_pass_a_job_waiting_for_memory(ctx, S, pairs, W):
    """Budget for one large pair plus one small one: while the first large
    job runs, a second large job (another receptor) waits for memory and a
    small job submitted after it is admitted first; all three equal their
    runs alone."""
    big = [pairs[3], W.make_workload("serve e", 95, 1200, 300, 90)]
    small = pairs[0]
    evals = 60
    peaks = []
    for wl in big + [small]:
        with S.Server(0) as srv:
            srv.wait(srv.submit(srv.add_workload(wl), wl.dofs, evals))
            peaks.append(srv.stats()["peak_bytes"])
    budget = max(peaks[0], peaks[1]) + int(peaks[2] * 1.2)
    assert budget < peaks[0] + peaks[1]
    ref = [alone(ctx, wl, evals) for wl in big + [small]]
    with S.Server(0, budget) as srv:
        ids = [srv.add_workload(wl) for wl in big + [small]]
        jobs = [srv.submit(i, wl.dofs, evals) for i, wl in zip(ids, big + [small])]
        res = [srv.wait(j) for j in jobs]
        seq = [r[4]["admit_seq"] for r in res]
        print(f"budget {budget} B, peaks {peaks}, admission order {seq}")
        for r, rr in zip(res, ref):
            assert same(r, rr)
        assert seq[0] == 1 and seq[2] == 2 and seq[1] == 3
        assert srv.stats()["peak_bytes"] <= budget


def test_serve_errors_and_timeouts(ctx, S, pairs):
    from gpuattract_amd import hip
    wl = pairs[3]
    with S.Server(0) as srv:
        ids = srv.add_workload(wl)
        with pytest.raises(hip.AttError):
            srv.submit(dict(ids, rec=12345), wl.dofs, 10)
        with pytest.raises(hip.AttError):
            srv.submit(ids, wl.dofs, -1)
        with pytest.raises(hip.AttError):
            srv.submit(ids, wl.dofs[:0], 10)
        big = np.repeat(wl.dofs, 40, axis=0)
        j = srv.submit(ids, big, 400)
        assert srv.wait(j, timeout_ms=0) is None  # still running
        r = srv.wait(j)
        assert r[0].shape == big.shape and np.all(r[3] <= 400)
    # a job larger than the whole budget fails alone, the next one runs
    with S.Server(0, 1 << 20) as srv:
        ids = srv.add_workload(wl)
        j = srv.submit(ids, wl.dofs, 10)
        with pytest.raises(hip.AttError, match="budget"):
            srv.wait(j)


def test_serve_device_built_grids(ctx, S, pairs):
    """A receptor registered as a body: its grids are built on the device at
    first use and equal the host generator's, so results equal att_minimize
    on the uploaded host grids bit for bit."""
    from gpuattract_amd import hip
    wl = pairs[1]
    ref = alone(ctx, wl, 40)
    with S.Server(0) as srv:
        ids = srv.add_workload(wl, device_grids=True)
        r = srv.wait(srv.submit(ids, wl.dofs, 40))
        assert same(r, ref) and r[4]["grid_cached"] == 0
        r2 = srv.wait(srv.submit(ids, wl.dofs, 40))
        assert same(r2, ref) and r2[4]["grid_cached"] == 1
        other = srv.add_params(wl.params)
        with pytest.raises(hip.AttError, match="another table"):
            srv.submit(dict(ids, params=other), wl.dofs, 10)
