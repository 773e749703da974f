"""HIP path vs the CPU oracle on the same seeded inputs (needs a GPU).

Bit-exact: rotation matrices, transformed coordinates, per-bead forces and
energies, grid cell indices, neighbour-list cells and pair counts, and the
per-pose energy, force, torque and ligand mode gradients of the forward
kernels (the oracle folds each bead's pairs in list order and the beads in
bead order, as the kernels and the reference do); the adjoint against the
oracle's kernel-order restatement.  The forward's receptor mode gradients
(the reaction projected per pair) within 1e-5 of their norm.  The
tolerance of tests/parity.py (within 1e-5 of the condition scale) is
asserted beside, with the FP64 comparison printed.
"""
import numpy as np
import pytest

from parity import TOL, cond_err, forcelim_safe, summarize

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gpuattract_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def _host_sincosf(x):
    """The host C library's sincosf values (the reference's sin/cos)."""
    import ctypes
    import os
    L = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_bin",
                                 "libtrig_host.so"))
    L.trig_host_eval.restype = None
    L.trig_host_eval.argtypes = [ctypes.c_void_p, ctypes.c_size_t] + [ctypes.c_void_p] * 4
    x = np.ascontiguousarray(x, np.float32)
    out = [np.zeros_like(x) for _ in range(4)]
    L.trig_host_eval(x.ctypes.data, x.size, *[o.ctypes.data for o in out])
    return out[0], out[1]


def test_device_sincosf_is_the_c_library(ctx):
    """The kernels' sin/cos (sincosf_libm on the device) equal the host C
    library's sincosf bit for bit: a sweep of 4M floats through [-8, 8]
    (every 256th float pattern below magnitude 8), 1M uniform angles in
    [-2pi, 2pi], the large-argument reduction, special values and every
    golden-fixture angle."""
    import os
    rng = np.random.default_rng(17)
    u = np.arange(0, 0x41000000, 256, dtype=np.uint32)
    sweep = np.concatenate([u, u | np.uint32(0x80000000)]).view(np.float32)
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    fixture = [np.load(os.path.join(gold, "rotmat.npz"))["ang"].ravel()]
    for f in sorted(os.listdir(gold)):
        if f.endswith(".npz"):
            z = np.load(os.path.join(gold, f))
            if "dofs" in z.files:
                fixture.append(z["dofs"][:, 3:6].ravel())
    x = np.concatenate([
        sweep,
        rng.uniform(-2 * np.pi, 2 * np.pi, 1000000).astype(np.float32),
        rng.uniform(-1e6, 1e6, 100000).astype(np.float32),
        np.array([0.0, -0.0, 1e-45, 2.0 ** -12, 0.75, 119.99, 120.0, 3e38, -3e38], np.float32),
    ] + fixture)
    s, c = ctx.sincosf(x)
    hs, hc = _host_sincosf(x)
    assert np.array_equal(s.view(np.uint32), hs.view(np.uint32))
    assert np.array_equal(c.view(np.uint32), hc.view(np.uint32))
    # infinities and NaN give NaN, as the library does
    s, c = ctx.sincosf(np.array([np.inf, -np.inf, np.nan], np.float32))
    assert np.all(np.isnan(s)) and np.all(np.isnan(c))


@pytest.mark.parametrize("plateau", [10.0, 4.25, 6.0, 7.5, 8.8, 10.8, 12.0, 15.3])
def test_plateau_ratio_is_the_table(ctx, O, plateau):
    """The plateau ratio the list walks evaluate per pair equals every entry of
    the table the reference builds (model/NLGrid.tpp:44-51): the oracle's
    table (pinned to the reference's golden values at plateau 10) and, where
    oracle/_ref is built, the reference NLGrid's own table."""
    lut = O.ratio_lut(plateau)
    P2 = np.float32(plateau) * np.float32(plateau)
    assert len(lut) == int(np.float32(10000) * P2) + 1
    dev = ctx.plateau_ratio(P2, len(lut))
    assert np.array_equal(dev.view(np.uint32), lut.view(np.uint32))
    if O.ref_available():
        ref = O.ref_ratio_lut(plateau)
        assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))


def test_rotmats_bit_exact(ctx, O):
    """Rotation matrices as the kernels build them equal the oracle's (the
    reference's formula on the C library's sin/cos) and the reference's own
    (golden rotmat.npz) bit for bit."""
    import os
    rng = np.random.default_rng(5)
    ang = np.concatenate([
        rng.uniform(-7, 7, (200000, 3)),
        rng.uniform(-1e-3, 1e-3, (1000, 3)),
        rng.uniform(-100, 100, (1000, 3)),
        np.array([[0, 0, 0], [np.pi, 0, 0], [0, np.pi / 2, 0], [-0.0, 0.0, -0.0]]),
    ]).astype(np.float32)
    R = ctx.rotmats(ang)
    ref = O.rotmats(ang)
    assert np.array_equal(R.view(np.uint32), ref.view(np.uint32))
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rotmat.npz"))
    assert np.array_equal(ctx.rotmats(f["ang"]).view(np.uint32), f["R"].view(np.uint32))


def _fp64_statement(name, g, o32, o64, sc):
    """The error against the FP64 recompute is stated, not bounded by 1e-5:
    the reference's own FP32 per-bead arithmetic already differs from FP64 by
    more.  Asserted: the GPU result is no further from FP64 than the
    reference's FP32 result (up to 10% and 1e-6 of the scale)."""
    e_gpu = cond_err(g, o64, sc)
    e_ref = cond_err(o32, o64, sc)
    print(summarize(f"{name} gpu vs fp64", g, o64, sc))
    print(summarize(f"{name} oracle fp32 vs fp64", o32, o64, sc))
    assert e_gpu.max() <= 1.1 * e_ref.max() + 1e-6


def _check_beads(ctx, hs, wl, O, dof):
    g = ctx.beads(hs, wl, dof)
    o = O.beads(wl, dof)
    assert np.array_equal(g["xyz"], o["xyz"])
    assert np.array_equal(g["cell"], o["cell"])
    assert np.array_equal(g["nl_pairs"], o["nl_pairs"])
    assert np.array_equal(g["pot"], o["pot"])
    if wl.modes:
        assert np.array_equal(g["rec_pot"], o["rec_pot"])
        assert np.array_equal(g["rec_xyz"], o["rec_xyz"])
    return o


def test_beads_rigid_bit_exact(ctx, wl_rigid, O):
    hs = ctx.upload_workload(wl_rigid)
    used_nl = 0
    for k in range(0, 300, 7):
        o = _check_beads(ctx, hs, wl_rigid, O, wl_rigid.dofs[k])
        used_nl += int((o["nl_pairs"] > 0).sum())
    assert used_nl > 0  # the neighbour-list term was exercised


def test_beads_modes_bit_exact(ctx, wl_modes, O):
    hs = ctx.upload_workload(wl_modes)
    for k in range(0, 200, 9):
        _check_beads(ctx, hs, wl_modes, O, wl_modes.dofs[k])


def _rigid_parity(ctx, wl, O, nthreads=8):
    hs = ctx.upload_workload(wl)
    g = ctx.score_rigid(hs, wl.dofs)
    o32 = O.score_rigid(wl, nthreads=nthreads)
    o64, sc = O.score_rigid(wl, nthreads=nthreads, f64=True)
    ok = forcelim_safe(o64)
    assert ok.mean() > 0.99
    e32 = cond_err(g[ok], o32[ok], sc[ok])
    print(summarize("rigid gpu vs oracle fp32", g[ok], o32[ok], sc[ok]))
    assert e32.max() <= TOL, e32.max(axis=0)
    # every pose, clamped forces included: the oracle sums in the kernels'
    # (the reference's) order, so the bits are equal
    assert np.array_equal(g.view(np.uint32), o32.view(np.uint32))
    _fp64_statement("rigid", g[ok], o32[ok], o64[ok], sc[ok])
    return g, o32


def _assert_modes_bits(g, o32, Mr):
    """Energy, force, torque and ligand mode gradients equal the oracle's
    bits; the receptor mode gradients (the reaction projected per pair, not
    per receptor bead as the reference sums it) within 1e-5 of their norm."""
    keep = list(range(7)) + list(range(17, 27))
    assert np.array_equal(g[:, keep].view(np.uint32), o32[:, keep].view(np.uint32))
    if Mr:
        r = g[:, 7:7 + Mr].astype(np.float64)
        q = o32[:, 7:7 + Mr].astype(np.float64)
        fin = np.all(np.isfinite(q), axis=1)
        assert np.array_equal(fin, np.all(np.isfinite(r), axis=1))
        nrm = np.linalg.norm(q[fin], axis=1)
        err = np.linalg.norm(r[fin] - q[fin], axis=1)
        assert np.all(err <= 1e-5 * nrm + 1e-30), (err / np.maximum(nrm, 1e-30)).max()


def test_score_rigid_small(ctx, wl_rigid, O):
    _rigid_parity(ctx, wl_rigid, O)


def test_score_rigid_1000_300(ctx, W, O):
    wl = W.config_rigid_1000_300(10000)
    g, _ = _rigid_parity(ctx, wl, O, nthreads=16)
    assert np.all(np.isfinite(g))


def test_score_modes_small(ctx, wl_modes, O):
    hs = ctx.upload_workload(wl_modes)
    g = ctx.score_modes(hs, wl_modes.dofs)
    o32 = O.score_modes(wl_modes, nthreads=8)
    o64, sc = O.score_modes(wl_modes, nthreads=8, f64=True)
    ok = forcelim_safe(o64)
    cols = [c for c in range(27) if sc[:, c].max() > 0]
    e = cond_err(g[ok][:, cols], o32[ok][:, cols], sc[ok][:, cols])
    print(summarize("modes gpu vs oracle fp32", g[ok][:, cols], o32[ok][:, cols], sc[ok][:, cols]))
    assert e.max() <= TOL, e.max(axis=0)
    _assert_modes_bits(g, o32, wl_modes.rec.M)
    _fp64_statement("modes", g[ok][:, cols], o32[ok][:, cols], o64[ok][:, cols], sc[ok][:, cols])
    # unused mode slots are exactly zero
    assert np.all(g[:, 7 + wl_modes.rec.M:17] == 0) and np.all(g[:, 17 + wl_modes.lig.M:] == 0)


def test_device_resident_matches_host_path(ctx, wl_modes, wl_rigid):
    for wl, out_w in ((wl_rigid, 7), (wl_modes, 27)):
        hs = ctx.upload_workload(wl)
        host = ctx.score_modes(hs, wl.dofs) if wl.modes else ctx.score_rigid(hs, wl.dofs)
        n = wl.dofs.shape[0]
        d_in = ctx.to_device(wl.dofs)
        d_out = ctx.dev_alloc(n * out_w * 4)
        f = ctx.score_modes_dev if wl.modes else ctx.score_rigid_dev
        e0, e1 = ctx.event(), ctx.event()
        ctx.record(e0)
        f(hs, d_in.ptr, n, d_out.ptr, None)
        ctx.record(e1)
        assert ctx.elapsed_ms(e0, e1) > 0
        assert np.array_equal(ctx.to_host(d_out, (n, out_w)), host)


def test_forward_spatial_order_changes_no_result(ctx, wl_modes, wl_rigid):
    """From 4096 poses on, the forward kernels take their poses in spatial
    order, the workgroups of each XCD one contiguous range of it; every pose
    is still written to its own slot, so the outputs are the same bits as in
    input order.  5003 poses: the ranges of the eight XCDs are uneven."""
    rng = np.random.default_rng(3)
    for wl, out_w in ((wl_rigid, 7), (wl_modes, 27)):
        hs = ctx.upload_workload(wl)
        n = 5003
        d = np.resize(wl.dofs, (n, wl.dofs.shape[1])).astype(np.float32)
        d[:, :3] += rng.normal(scale=2.0, size=(n, 3)).astype(np.float32)
        d_in = ctx.to_device(d)
        d_out = ctx.dev_alloc(n * out_w * 4)
        f = ctx.score_modes_dev if wl.modes else ctx.score_rigid_dev
        res = []
        # different fills before the two runs: a slot left unwritten by either
        # run differs (clashing poses may legitimately give inf or NaN)
        for enabled, fill in ((False, np.nan), (True, -7.0)):
            ctx.set_pose_order(enabled)
            ctx.write(d_out, np.full((n, out_w), fill, np.float32))
            f(hs, d_in.ptr, n, d_out.ptr, None)
            ctx.sync()
            res.append(ctx.to_host(d_out, (n, out_w)))
        ctx.set_pose_order(True)
        assert np.array_equal(res[0].view(np.uint32), res[1].view(np.uint32))


def test_split_and_permutation_invariance(ctx, wl_rigid):
    hs = ctx.upload_workload(wl_rigid)
    full = ctx.score_rigid(hs, wl_rigid.dofs)
    parts = np.concatenate([ctx.score_rigid(hs, wl_rigid.dofs[a:b])
                            for a, b in ((0, 1), (1, 77), (77, 300))])
    assert np.array_equal(full, parts)
    perm = np.random.default_rng(1).permutation(300)
    assert np.array_equal(ctx.score_rigid(hs, wl_rigid.dofs[perm]), full[perm])


def test_edge_cases(ctx, W, O):
    from gpuattract_amd import hip
    # type-0 beads, 12-6 potential, constant dielectric, one-bead ligand
    for kw in (dict(type0_frac=0.2), dict(shape=12), dict(dielec=0), dict(charged_frac=1.0)):
        wl = W.make_workload("edge", 21, 150, 40, 64, **kw)
        hs = ctx.upload_workload(wl)
        _check_beads(ctx, hs, wl, O, wl.dofs[3])
        g = ctx.score_rigid(hs, wl.dofs)
        o32 = O.score_rigid(wl)
        o64, sc = O.score_rigid(wl, f64=True)
        ok = forcelim_safe(o64)
        assert cond_err(g[ok], o32[ok], sc[ok]).max() <= TOL
        assert np.array_equal(g.view(np.uint32), o32.view(np.uint32))
    wl = W.make_workload("one bead", 22, 100, 1, 32)
    hs = ctx.upload_workload(wl)
    g = ctx.score_rigid(hs, wl.dofs)
    o64, sc = O.score_rigid(wl, f64=True)
    # one bead: its neighbour-list pairs folded in list order, as the oracle
    # (and the reference) fold them
    assert np.array_equal(g.view(np.uint32), O.score_rigid(wl).view(np.uint32))
    assert cond_err(g, o64, sc).max() <= 1e-3
    # poses far outside every grid score exactly zero
    far = wl.dofs.copy()
    far[:, :3] = 1e4
    assert np.all(ctx.score_rigid(hs, far) == 0)
    # NaN poses do not fault and stay outside the grids
    nan = wl.dofs[:4].copy()
    nan[:, 0] = np.nan
    ctx.score_rigid(hs, nan)
    # empty request
    assert ctx.score_rigid(hs, wl.dofs[:0]).shape == (0, 7)
    # error paths
    with pytest.raises(hip.AttError) as e:
        ctx.score_rigid(dict(hs, rec=987654), wl.dofs)
    assert e.value.status == 4
    with pytest.raises(hip.AttError) as e:
        ctx.score_modes(dict(hs, lig_grid=hs["params"]), np.zeros((2, 26), np.float32))
    assert e.value.status == 4
    # a ligand type beyond the parameter table is rejected before any launch
    small = W.make_params(3, T=4)
    hp = ctx.upload_params(small)
    with pytest.raises(hip.AttError) as e:
        ctx.score_rigid(dict(hs, params=hp), wl.dofs)
    assert e.value.status == 1


def test_many_types_table_in_global_memory(ctx, W, O):
    """40 bead types: the pair table (1600 entries) exceeds what the kernels
    stage in LDS (32x32), so they read it from global memory and their LDS
    slot arrays start at offset 0; forward and adjoint still equal the
    oracle's bits (the forward's receptor mode gradients to 1e-5 normwise)."""
    wl = W.make_workload("many types", 31, 300, 90, 64, num_types=40)
    hs = ctx.upload_workload(wl)
    g = ctx.score_rigid(hs, wl.dofs)
    o64, sc = O.score_rigid(wl, f64=True)
    ok = forcelim_safe(o64)
    assert cond_err(g[ok], O.score_rigid(wl)[ok], sc[ok]).max() <= TOL
    assert np.array_equal(g.view(np.uint32), O.score_rigid(wl).view(np.uint32))
    a = ctx.adjoint(hs, wl.dofs)
    r32 = O.adjoint(wl, f32=True)
    _, sca = O.adjoint(wl, scale=True)
    cols = [c for c in range(7) if sca[:, c].max() > 0]
    assert cond_err(a[:, cols], r32[:, cols], sca[:, cols]).max() <= TOL
    k = O.adjoint_kernel_order(wl)[:, :7].astype(np.float32)
    assert np.array_equal(a.view(np.uint32), k.view(np.uint32))
    wm = W.make_workload("many types modes", 32, 300, 90, 32, 2, 3, num_types=40)
    hm = ctx.upload_workload(wm)
    gm = ctx.score_modes(hm, wm.dofs)
    om = O.score_modes(wm)
    _, scm = O.score_modes(wm, f64=True)
    colsm = scm.max(axis=0) > 0
    assert cond_err(gm[:, colsm], om[:, colsm], scm[:, colsm]).max() <= TOL
    _assert_modes_bits(gm, om, wm.rec.M)
    am = ctx.adjoint(hm, wm.dofs)
    km = O.adjoint_kernel_order(wm)[:, :27].astype(np.float32)
    assert np.array_equal(am.view(np.uint32), km.view(np.uint32))


@pytest.mark.parametrize("n_lig", [63, 64, 65, 129])
def test_bead_chunk_boundaries(ctx, W, O, n_lig):
    """Ligands around the 64-bead chunks of the scoring kernels' lane-per-bead
    phases: forward results equal the oracle's bits (the receptor mode
    gradients to 1e-5 normwise)."""
    wl = W.make_workload(f"chunks {n_lig}", 40 + n_lig, 250, n_lig, 48)
    hs = ctx.upload_workload(wl)
    g = ctx.score_rigid(hs, wl.dofs)
    o64, sc = O.score_rigid(wl, f64=True)
    ok = forcelim_safe(o64)
    o32 = O.score_rigid(wl)
    assert cond_err(g[ok], o32[ok], sc[ok]).max() <= TOL
    assert np.array_equal(g.view(np.uint32), o32.view(np.uint32))
    wm = W.make_workload(f"chunks modes {n_lig}", 50 + n_lig, 250, n_lig, 24, 2, 2)
    hm = ctx.upload_workload(wm)
    gm = ctx.score_modes(hm, wm.dofs)
    om = O.score_modes(wm)
    _, scm = O.score_modes(wm, f64=True)
    cols = scm.max(axis=0) > 0
    assert cond_err(gm[:, cols], om[:, cols], scm[:, cols]).max() <= TOL
    _assert_modes_bits(gm, om, wm.rec.M)
