"""The benchmark configurations at their full sizes (needs a GPU).

* flexible scoring 3000/800 beads, 5 + 5 modes, 100,000 poses: the forward
  kernel against the oracle's forward restatement (the reference's summation
  order) on 500 poses, the adjoint with parameter gradients against the
  oracle's FP32 adjoint in the kernel's order (bit for bit) and its parameter
  gradient against the FP64 oracle, then all 100,000 poses through
  size-independent properties: adjoint energy = forward energy bit for bit,
  results independent of the pose order and of the call's other poses,
  repeat runs identical, exact linearity in the upstream weight;
* 64 uneven pairs (make_batch(20261204, 64), receptor grids built on the
  device): the tuned graph replay against the kernel-order oracle on the
  first 160 poses of every pair, with the waves per pose of each pair's
  bucket, and replays identical;
* minimization 4000/1000 beads, 10 + 10 modes: 200 poses at 10 evaluations
  equal to the oracle's VA13 on the kernel-order adjoint (final DOFs, energies
  and evaluation counts), and the same poses minimized inside the full
  50,000-pose run equal to the 200-pose run;
* the serving stream (make_serve_stream(20261206, 40)): one job per ligand
  size class, served with its receptor's grids built on the device, against
  the oracle's VA13 on the same poses.

In every case the maximum plain relative energy error is printed and
asserted (zero: the energies are the same bits)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gpuattract_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def plain_rel(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.abs(x - y) / np.abs(y)
    r[x == y] = 0.0
    return r


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def threads():
    import os
    return os.cpu_count() or 1


# ---------------------------------------------------------------- flexible
@pytest.fixture(scope="module")
def flex(W, ctx):
    wl = W.make_workload("flexible scoring 3000/800, 5+5 modes", 20261203, 3000, 800, 100000,
                         5, 5, with_basis=True)
    hs = ctx.upload_workload(wl)
    ctx.upload_basis(hs["grid"], wl)
    return wl, hs


def test_flexible_3000_800_against_oracle(ctx, O, flex):
    wl, hs = flex
    k = 500
    d = wl.dofs[:k]
    fwd = ctx.score_modes(hs, d)
    ref = O.score_modes(wl, d, nthreads=threads())
    lig_cols = [0, 1, 2, 3, 4, 5, 6] + [17 + m for m in range(wl.lig.M)]
    e = plain_rel(fwd[:, 0], ref[:, 0]).max()
    print(f"forward, {k} poses: max plain relative energy error {e:.1e}")
    assert e == 0.0
    assert np.array_equal(bits(fwd[:, lig_cols]), bits(ref[:, lig_cols]))
    rc = slice(7, 7 + wl.rec.M)
    # (the receptor mode gradients are projected per pair rather than summed
    # per receptor bead first; overflowing clash poses are inf in both)
    assert np.array_equal(np.isfinite(fwd[:, rc]), np.isfinite(ref[:, rc]))
    fin = np.all(np.isfinite(ref[:, rc]), axis=1)
    g, r = fwd[fin, rc].astype(np.float64), ref[fin, rc].astype(np.float64)
    err = np.linalg.norm(g - r, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-30)
    print(f"forward receptor mode gradients ({fin.sum()} finite poses): normwise max {err.max():.1e}")
    assert err.max() <= 1e-5

    T = wl.params.T
    a, pg = ctx.adjoint(hs, d, pgrad=True, T=T)
    ko = O.adjoint_kernel_order(wl, d, nthreads=threads())[:, :27]
    assert np.array_equal(a.view(np.uint32), bits(ko))
    assert plain_rel(a[:, 0], ref[:, 0]).max() == 0.0
    _, pref = O.adjoint(wl, d, pgrad=True)
    scale = np.abs(pref).max()
    perr = np.abs(pg - pref).max() / scale
    print(f"adjoint, {k} poses: bit-exact with the kernel-order oracle; parameter gradient "
          f"max error {perr:.1e} of its largest entry")
    assert perr <= 1e-4


def test_flexible_3000_800_all_poses_properties(ctx, flex):
    wl, hs = flex
    n = wl.dofs.shape[0]
    T = wl.params.T
    fwd = ctx.score_modes(hs, wl.dofs)
    a, pg = ctx.adjoint(hs, wl.dofs, pgrad=True, T=T)
    assert np.all(np.isfinite(pg))
    # the adjoint's energy is the forward kernel's, whose energy is the
    # reference's, for every pose
    assert np.array_equal(a[:, 0].view(np.uint32), fwd[:, 0].view(np.uint32))
    # repeat: the same bits, parameter gradient included
    a2, pg2 = ctx.adjoint(hs, wl.dofs, pgrad=True, T=T)
    assert np.array_equal(a.view(np.uint32), a2.view(np.uint32))
    assert np.array_equal(pg.view(np.uint32), pg2.view(np.uint32))
    # pose order: per-pose results follow their poses
    perm = np.random.default_rng(11).permutation(n)
    ap, pgp = ctx.adjoint(hs, wl.dofs[perm], pgrad=True, T=T)
    assert np.array_equal(ap.view(np.uint32), a[perm].view(np.uint32))
    assert np.allclose(pgp, pg, rtol=1e-4, atol=1e-5 * np.abs(pg).max())
    fp = ctx.score_modes(hs, wl.dofs[perm])
    assert np.array_equal(fp.view(np.uint32), fwd[perm].view(np.uint32))
    # a call's other poses do not matter: two halves give the same rows
    h = n // 3
    a1 = ctx.adjoint(hs, wl.dofs[:h])
    b1 = ctx.adjoint(hs, wl.dofs[h:])
    assert np.array_equal(np.concatenate([a1, b1]).view(np.uint32), a.view(np.uint32))
    # upstream weight 2: every gradient exactly doubles (a power of two
    # scales every product and sum exactly), the energy does not
    w2 = np.full(n, 2.0, np.float32)
    aw, pgw = ctx.adjoint(hs, wl.dofs, w=w2, pgrad=True, T=T)
    assert np.array_equal(aw[:, 0].view(np.uint32), a[:, 0].view(np.uint32))
    two = (2.0 * a[:, 1:]).astype(np.float32)
    fin = np.isfinite(two)
    assert np.array_equal(np.isfinite(aw[:, 1:]), fin)
    assert np.array_equal(aw[:, 1:][fin].view(np.uint32), two[fin].view(np.uint32))
    assert np.array_equal(np.isnan(aw[:, 1:]), np.isnan(two))
    assert np.array_equal(pgw.view(np.uint32), (2.0 * pg).view(np.uint32))
    print(f"{n} poses: adjoint energy = forward energy, repeat, permutation, split and "
          f"weight properties hold")


def test_pair_counts_of_the_production_kernels(ctx, W, O, flex):
    """Per-pose {neighbour-list entries walked, pairs inside the plateau}
    written by k_rigid, k_modes and k_adjoint themselves (att_debug_pair_counts)
    equal the oracle's integer counts: rigid 1000/300 (2,000 of its 10,000
    poses) and flexible 3000/800 (500 poses); on all 100,000 flexible poses the
    forward and adjoint kernels count the same pairs."""
    rig = W.make_workload("rigid scoring 1000/300", 20261202, 1000, 300, 10000)
    hr = ctx.upload_workload(rig)
    m = 2000
    want = O.pair_counts(rig, rig.dofs[:m])
    for name, fn in (("k_rigid", lambda: ctx.score_rigid(hr, rig.dofs)),
                     ("k_adjoint", lambda: ctx.adjoint(hr, rig.dofs))):
        got = ctx.pair_counts(fn, rig.dofs.shape[0])
        assert np.array_equal(got[:m], want), name
        print(f"rigid 1000/300, {name}: {m} poses equal to the oracle; mean {got[:, 0].mean():.0f} "
              f"list entries, {got[:, 1].mean():.0f} pairs inside the plateau per pose")
    wl, hs = flex
    k = 500
    want = O.pair_counts(wl, wl.dofs[:k])
    n = wl.dofs.shape[0]
    fwd = ctx.pair_counts(lambda: ctx.score_modes(hs, wl.dofs), n)
    adj = ctx.pair_counts(lambda: ctx.adjoint(hs, wl.dofs, pgrad=True, T=wl.params.T), n)
    assert np.array_equal(fwd[:k], want) and np.array_equal(adj[:k], want)
    assert np.array_equal(fwd, adj)
    print(f"flexible 3000/800: k_modes and k_adjoint equal to the oracle on {k} poses and to each "
          f"other on {n}; mean {fwd[:, 0].mean():.0f} list entries, {fwd[:, 1].mean():.0f} pairs")


# ---------------------------------------------------------------- batched
@pytest.fixture(scope="module")
def batch(W, ctx):
    pairs = W.make_batch(20261204, 64, host_grids=False)
    staged = []
    for wl in pairs:
        hs = ctx.upload_workload(wl)
        n, ow = wl.dofs.shape[0], (27 if wl.modes else 7)
        staged.append(dict(wl=wl, hs=hs, d_dofs=ctx.to_device(wl.dofs), n=n, ow=ow,
                           d_out=ctx.dev_alloc(n * ow * 4)))
    b = ctx.batch([dict(hs=s["hs"], d_dofs=s["d_dofs"].ptr, d_out=s["d_out"].ptr, n=s["n"],
                        modes=s["wl"].modes) for s in staged])
    b.tune()
    b.run()
    ctx.sync()
    first = [ctx.to_host(s["d_out"], (s["n"], s["ow"])) for s in staged]
    yield b, staged, first
    b.close()
    for s in staged:
        for key in ("rec", "lig", "grid", "params"):
            ctx.release(s["hs"][key])
        s["d_dofs"].release()
        s["d_out"].release()


def test_batched_64_pairs_replays_identical(ctx, batch):
    b, staged, first = batch
    for _ in range(2):
        b.run()
        ctx.sync()
        for s, f in zip(staged, first):
            assert np.array_equal(ctx.to_host(s["d_out"], (s["n"], s["ow"])).view(np.uint32),
                                  f.view(np.uint32))
    print(f"{sum(s['n'] for s in staged)} poses in {len(staged)} pairs: three replays identical")


@pytest.mark.parametrize("part", range(4))
def test_batched_64_pairs_against_oracle(ctx, W, O, batch, part):
    b, staged, first = batch
    buckets = b.buckets()
    k = 160
    worst = 0.0
    for i in range(16 * part, 16 * part + 16):
        s = staged[i]
        wl = s["wl"]
        W.fill_grids_from_device(ctx, wl)
        waves = buckets[b.pair_bucket(i)]["wavesPerPose"]
        c = min(k, s["n"])
        ko = O.adjoint_kernel_order(wl, wl.dofs[:c], waves=waves, nthreads=threads())[:, :s["ow"]]
        got = first[i][:c]
        worst = max(worst, plain_rel(got[:, 0], ko[:, 0]).max())
        assert np.array_equal(got.view(np.uint32), bits(ko)), wl.name
        # host memory: one pair's grids at a time
        T = wl.params.T
        wl.rec_inner = W.IGrid(wl.rec_inner.geom, T + 1, None)
        wl.rec_outer = W.IGrid(wl.rec_outer.geom, T + 1, None)
        wl.rec_nl = W.NLGrid(wl.rec_nl.geom, wl.rec_nl.plateau, None, None)
    print(f"pairs {16 * part}-{16 * part + 15}: first {k} poses bit-exact with the kernel-order "
          f"oracle, max plain relative energy error {worst:.1e}")
    assert worst == 0.0


# ---------------------------------------------------------------- minimization
def test_minimization_4000_1000_against_oracle(ctx, W, O):
    wl = W.make_workload("minimization 4000/1000, 10+10 modes", 20261205, 4000, 1000, 50000,
                         10, 10, with_lig_grid=False)
    hs = ctx.upload_workload(wl)
    m = 200
    d, el, eb, nf = ctx.minimize(hs, wl.dofs[:m], 10)
    od, oel, oeb, onf = O.minimize_kernel_order(wl, wl.dofs[:m], ivmax=10, nthreads=threads())
    assert np.array_equal(nf, onf)
    assert np.array_equal(d.view(np.uint32), bits(od))
    assert np.array_equal(el.astype(np.float64), oel)
    assert np.array_equal(eb.astype(np.float64), oeb)
    print(f"{m} poses at 10 evaluations: equal to the oracle's VA13, max plain relative energy "
          f"error {plain_rel(eb, oeb).max():.1e}")
    # the same poses inside the full run
    D, EL, EB, NF = ctx.minimize(hs, wl.dofs, 10)
    assert np.array_equal(D[:m].view(np.uint32), d.view(np.uint32))
    assert np.array_equal(EB[:m], eb) and np.array_equal(NF[:m], nf)
    assert np.all(NF <= 10)
    for key in ("rec", "lig", "grid", "params"):
        ctx.release(hs[key])


# ---------------------------------------------------------------- serving
def test_serving_stream_one_job_per_size_class(ctx, W, O):
    from gpuattract_amd import serve
    params, recs, jobs = W.make_serve_stream(20261206, 40, max_fun_eval=200)
    bounds = [192, 384, 640, 1024, 1 << 30]
    picks = []
    for lo, hi in zip([0] + bounds[:-1], bounds):
        for k, j in enumerate(jobs):
            if lo <= j.lig.n < hi:
                picks.append(k)
                break
    assert len(picks) >= 4
    P = 32
    with serve.Server(0) as srv:
        pid = srv.add_params(params)
        rid = {}
        for k in picks:
            j = jobs[k]
            r = recs[j.receptor]
            if j.receptor not in rid:
                rid[j.receptor] = srv.add_receptor_body(r.body, pid, r.inner, r.outer, 10.0, 10.8)
            lid = srv.add_ligand(j.lig)
            res = srv.wait(srv.submit(dict(rec=rid[j.receptor], lig=lid, params=pid), j.dofs[:P],
                                      j.max_fun_eval))
            T = params.T
            wl = W.Workload(f"job {k}", r.body, j.lig, params, W.IGrid(r.inner, T + 1, None),
                            W.IGrid(r.outer, T + 1, None), W.NLGrid(r.inner, 10.0, None, None),
                            j.dofs[:P])
            wl.meta = dict(plateau=10.0, neighbourdis=10.8)
            W.fill_grids_from_device(ctx, wl)
            od, oel, oeb, onf = O.minimize_kernel_order(wl, ivmax=j.max_fun_eval, nthreads=threads())
            dd = res[0]
            assert np.array_equal(res[3], onf), k
            assert np.array_equal(dd.view(np.uint32), bits(od[:, :dd.shape[1]])), k
            assert np.array_equal(res[1].astype(np.float64), oel), k
            assert np.array_equal(res[2].astype(np.float64), oeb), k
            print(f"job {k} ({r.body.n}/{j.lig.n} beads, {r.body.M}+{j.lig.M} modes): {P} poses "
                  f"equal to the oracle's VA13 at <= {j.max_fun_eval} evaluations, max plain "
                  f"relative energy error {plain_rel(res[2], oeb).max():.1e}")
