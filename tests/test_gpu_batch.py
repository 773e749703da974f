"""Batched adjoint over many uneven pairs (needs a GPU).

- At the single-pair kernel's shape (1 wave per pose for ligands under 384
  beads, 2 above) a batch returns, pair by pair, exactly the single-pair
  adjoint (same device code, same order).
- Every launch shape (1, 2 or 4 waves per pose, any tile size) and the shape
  the autotuner picks return, pair by pair, the oracle's FP32 adjoint in the
  kernel's summation order for the bucket's waves per pose, bit for bit, and
  the oracle's sequential FP32 adjoint within 1e-5 of the condition scale.
- Replays of the captured graph are identical, with eager work between them.
"""
import numpy as np
import pytest

from parity import TOL, cond_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gpuattract_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def pairs(W):
    mixed = W.make_batch(7, 6, rec_range=(150, 500), lig_range=(40, 300), poses_range=(50, 300))
    rigid = W.make_batch(8, 2, rec_range=(150, 500), lig_range=(150, 700), poses_range=(50, 200),
                         modes_range=(0, 0))
    return mixed + rigid


@pytest.fixture(scope="module")
def staged(ctx, pairs):
    st = []
    for wl in pairs:
        hs = ctx.upload_workload(wl)
        n, ow = wl.dofs.shape[0], (27 if wl.modes else 7)
        st.append(dict(wl=wl, hs=hs, d_dofs=ctx.to_device(wl.dofs), d_out=ctx.dev_alloc(n * ow * 4),
                       n=n, ow=ow))
    return st


def _batch(ctx, staged):
    return ctx.batch([dict(hs=s["hs"], d_dofs=s["d_dofs"].ptr, d_out=s["d_out"].ptr, n=s["n"],
                           modes=s["wl"].modes) for s in staged])


def _results(ctx, staged):
    return [ctx.to_host(s["d_out"], (s["n"], s["ow"])) for s in staged]


def _single(ctx, s):
    out = ctx.dev_alloc(s["n"] * s["ow"] * 4)
    ctx.adjoint_dev(s["hs"], s["d_dofs"].ptr, s["wl"].modes, s["n"], None, out.ptr)
    return ctx.to_host(out, (s["n"], s["ow"]))


def test_buckets(ctx, staged):
    b = _batch(ctx, staged)
    info = b.buckets()
    assert sum(k["numPairs"] for k in info) == len(staged)
    assert {k["modeCapacity"] for k in info} == {0, 5, 10}
    b.close()


def test_batch_equals_single_pair_at_same_shape(ctx, staged):
    b = _batch(ctx, staged)
    # size classes 0 and 1 hold the ligands under 384 beads
    for k, info in enumerate(b.buckets()):
        b.set_shape(k, 1 if info["sizeClass"] <= 1 else 2, 16)
    b.run()
    ctx.sync()
    got = _results(ctx, staged)
    for s, g in zip(staged, got):
        assert np.array_equal(g, _single(ctx, s)), s["wl"].name
    b.close()


def _check_oracle(O, staged, got, waves):
    """waves[k]: the waves per pose pair k was scored with."""
    for s, g, wv in zip(staged, got, waves):
        wl = s["wl"]
        ref, sc = O.adjoint(wl, scale=True)
        r32 = O.adjoint(wl, f32=True)
        cols = [c for c in range(s["ow"]) if sc[:, c].max() > 0]
        e = cond_err(g[:, cols], r32[:, cols], sc[:, cols])
        assert e.max() <= TOL, (wl.name, e.max(axis=0))
        k = O.adjoint_kernel_order(wl, waves=wv)[:, :s["ow"]].astype(np.float32)
        assert np.array_equal(g.view(np.uint32), k.view(np.uint32)), (wl.name, wv)


@pytest.mark.parametrize("waves,tile", [(1, 4), (2, 64), (4, 1)])
def test_batch_shapes_match_oracle(ctx, O, staged, waves, tile):
    b = _batch(ctx, staged)
    for k in range(len(b.buckets())):
        b.set_shape(k, waves, tile)
    b.run()
    ctx.sync()
    _check_oracle(O, staged, _results(ctx, staged), [waves] * len(staged))
    b.close()


def test_tuned_batch_and_graph_replays(ctx, O, staged):
    b = _batch(ctx, staged)
    b.tune()
    info = b.buckets()
    print("tuned shapes:", [(k["modeCapacity"], k["sizeClass"], k["wavesPerPose"],
                             k["posesPerTile"], round(k["ms"], 3)) for k in info])
    assert all(k["ms"] >= 0 for k in info)
    b.run()
    ctx.sync()
    first = _results(ctx, staged)
    _check_oracle(O, staged, first, [info[b.pair_bucket(k)]["wavesPerPose"] for k in range(len(staged))])
    for _ in range(3):
        _single(ctx, staged[0])  # eager work between replays
        b.run()
        ctx.sync()
        again = _results(ctx, staged)
        for a, f in zip(again, first):
            assert np.array_equal(a, f)
    b.close()


def test_batch_errors(ctx, staged):
    from gpuattract_amd import hip
    b = _batch(ctx, staged)
    nb = len(b.buckets())
    with pytest.raises(hip.AttError):
        b.set_shape(nb, 4, 16)
    with pytest.raises(hip.AttError):
        b.set_shape(0, 3, 16)
    with pytest.raises(hip.AttError):
        b.set_shape(0, 4, 0)
    b.close()
    empty = ctx.batch([])
    empty.run()
    ctx.sync()
    empty.close()


def test_spatial_order_changes_no_result(ctx, staged):
    """Pairs of 256 poses or more are scored in their spatial order, dealt into
    per-XCD queues; every pose still lands in its own slot, so results with the
    order switched off are the same bits."""
    assert any(s["n"] >= 256 for s in staged)
    runs = []
    for enabled in (False, True):
        ctx.set_pose_order(enabled)
        b = _batch(ctx, staged)
        for k in range(len(b.buckets())):
            b.set_shape(k, 2, 4)
        for s in staged:  # different fills: a slot left unwritten by either run differs
            ctx.write(s["d_out"], np.full((s["n"], s["ow"]), np.nan if enabled else -7.0, np.float32))
        ctx.sync()
        b.run()
        ctx.sync()
        runs.append(_results(ctx, staged))
        b.close()
    ctx.set_pose_order(True)
    for a, c, s in zip(runs[0], runs[1], staged):
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32)), s["wl"].name
