"""HIP adjoint vs the oracle's adjoint (needs a GPU).

Bit for bit: every output but the parameter gradient against the oracle's
FP32 adjoint in the kernel's summation order; the energy against the forward
path (the reference's energy).

Tolerance: per-pose gradients within 1e-5 of the condition scale (sum of
|terms|, oracle FP64) against the oracle's FP32 adjoint, which forms the bead
positions bit-identically; against FP64 the GPU error is bounded by 1.1x the
FP32 restatement's own error + 1e-6 (the FP32 rounding of positions in steep
grid regions, which no summation order removes).  Parameter gradients (per-wave LDS
slices, summed in a fixed order: repeatable bits, but a different order from
the oracle's) within 1e-4 of the largest entry.  The adjoint's energy equals the forward's
energy up to summation order.  The torch harness is checked against central
finite differences of the GPU energy and against the adjoint directly.
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
def wl_adj_rigid(W):
    return W.make_workload("adj rigid", 51, 400, 120, 300, with_basis=True)


@pytest.fixture(scope="module")
def wl_adj_modes(W):
    return W.make_workload("adj modes", 52, 360, 110, 200, 3, 4, with_basis=True)


def _upload(ctx, wl):
    hs = ctx.upload_workload(wl)
    ctx.upload_basis(hs["grid"], wl)
    return hs


@pytest.mark.parametrize("which", ["rigid", "modes"])
def test_adjoint_against_oracle(ctx, O, which, wl_adj_rigid, wl_adj_modes):
    wl = wl_adj_rigid if which == "rigid" else wl_adj_modes
    hs = _upload(ctx, wl)
    rng = np.random.default_rng(3)
    w = rng.uniform(-1.5, 2.0, wl.dofs.shape[0]).astype(np.float32)
    g, pg = ctx.adjoint(hs, wl.dofs, w=w, pgrad=True, T=wl.params.T)
    ref, sc, pref = O.adjoint(wl, w=w.astype(np.float64), pgrad=True, scale=True)
    r32 = O.adjoint(wl, w=w.astype(np.float64), f32=True)
    cols = list(range(7)) if which == "rigid" else [c for c in range(27) if sc[:, c].max() > 0]
    e32 = cond_err(g[:, cols], r32[:, cols], sc[:, cols])
    assert e32.max() <= TOL, e32.max(axis=0)
    e_gpu = cond_err(g[:, cols], ref[:, cols], sc[:, cols])
    e_ref = cond_err(r32[:, cols], ref[:, cols], sc[:, cols])
    print(f"{which}: GPU vs FP64 max {e_gpu.max():.2e}, FP32 oracle vs FP64 max {e_ref.max():.2e}")
    assert e_gpu.max() <= 1.1 * e_ref.max() + 1e-6
    scale = np.abs(pref).max()
    assert np.abs(pg - pref).max() <= 1e-4 * scale
    # the adjoint's energy is the forward's energy, bit for bit
    fwd = ctx.score_modes(hs, wl.dofs) if wl.modes else ctx.score_rigid(hs, wl.dofs)
    assert np.array_equal(g[:, 0], fwd[:, 0])


@pytest.mark.parametrize("which", ["rigid", "modes", "rigid500", "modes500"])
def test_adjoint_bit_exact_in_kernel_order(ctx, W, O, which, wl_adj_rigid, wl_adj_modes):
    """Every output but the parameter gradient equals the oracle's FP32
    adjoint evaluated in the kernel's summation order (one wave per pose under
    384 ligand beads, two above), weights included; the energy is the
    reference's (bead-order sums of list-order folds)."""
    wl = {"rigid": wl_adj_rigid, "modes": wl_adj_modes,
          "rigid500": W.make_workload("adj rigid 500", 55, 700, 500, 150, with_basis=True),
          "modes500": W.make_workload("adj modes 500", 56, 700, 500, 150, 5, 5, with_basis=True)}[which]
    hs = _upload(ctx, wl)
    w = np.random.default_rng(4).uniform(-1.5, 2.0, wl.dofs.shape[0]).astype(np.float32)
    g = ctx.adjoint(hs, wl.dofs, w=w)
    k = O.adjoint_kernel_order(wl, w=w)[:, :g.shape[1]].astype(np.float32)
    assert np.array_equal(g.view(np.uint32), k.view(np.uint32))
    fwd = ctx.score_modes(hs, wl.dofs) if wl.modes else ctx.score_rigid(hs, wl.dofs)
    assert np.array_equal(g[:, 0], fwd[:, 0])


def test_adjoint_without_parameters_and_device_path(ctx, wl_adj_modes):
    wl = wl_adj_modes
    hs = _upload(ctx, wl)
    host = ctx.adjoint(hs, wl.dofs)
    n = wl.dofs.shape[0]
    d_in = ctx.to_device(wl.dofs)
    d_out = ctx.dev_alloc(n * 27 * 4)
    ctx.adjoint_dev(hs, d_in.ptr, True, n, None, d_out.ptr)
    assert np.array_equal(ctx.to_host(d_out, (n, 27)), host)
    # parameter gradients need the basis grids: a grid without them is refused
    from gpuattract_amd import hip
    hs2 = ctx.upload_workload(wl)
    with pytest.raises(hip.AttError):
        ctx.adjoint(hs2, wl.dofs, pgrad=True, T=wl.params.T)


def test_adjoint_deterministic(ctx, wl_adj_rigid):
    hs = _upload(ctx, wl_adj_rigid)
    a = ctx.adjoint(hs, wl_adj_rigid.dofs)
    b = ctx.adjoint(hs, wl_adj_rigid.dofs)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n_lig", [110, 500])
def test_parameter_gradient_bits_repeat(ctx, W, n_lig):
    """The parameter gradient is summed in a fixed order (one LDS slice per
    wave, the slices and the workgroup rows added in order): repeated calls
    return the same bits, for one wave per pose (110 ligand beads) and two
    (500)."""
    wl = W.make_workload(f"pg repeat {n_lig}", 53, 600, n_lig, 3000, 3, 2, with_basis=True)
    hs = _upload(ctx, wl)
    T = wl.params.T
    runs = [ctx.adjoint(hs, wl.dofs, pgrad=True, T=T) for _ in range(3)]
    for g, pg in runs[1:]:
        assert np.array_equal(g.view(np.uint32), runs[0][0].view(np.uint32))
        assert np.array_equal(pg.view(np.uint32), runs[0][1].view(np.uint32))


def test_parameter_gradient_follows_the_bead_type_not_the_grid_slot(ctx, W, O):
    """A grid alphabet that orders the probe types differently (the ligand's
    mapped types then differ from its types): the adjoint returns the same
    bits, parameter gradient included, and matches the oracle."""
    from helpers import permute_alphabet
    wl = W.make_workload("alphabet", 54, 400, 120, 400, with_basis=True, repulsive_frac=0.5)
    T = wl.params.T
    wp = permute_alphabet(W, wl, np.roll(np.arange(1, T + 1), 5))
    assert np.mean(wp.lig.mtype != wl.lig.type) > 0.5
    hs, hp = _upload(ctx, wl), _upload(ctx, wp)
    g0, p0 = ctx.adjoint(hs, wl.dofs, pgrad=True, T=T)
    g1, p1 = ctx.adjoint(hp, wp.dofs, pgrad=True, T=T)
    assert np.array_equal(g0, g1) and np.array_equal(p0, p1)
    _, pref = O.adjoint(wp, pgrad=True)
    assert np.abs(p1 - pref).max() <= 1e-4 * np.abs(pref).max()


def test_torch_harness_finite_differences(ctx, wl_adj_modes):
    torch = pytest.importorskip("torch")
    import torch_harness as TH
    wl = wl_adj_modes
    hs = _upload(ctx, wl)
    T = wl.params.T
    dofs = torch.tensor(wl.dofs[:32], dtype=torch.float64, requires_grad=True)
    table = torch.zeros((T * T, 3), dtype=torch.float64, requires_grad=True)
    E = TH.energy(dofs, table, ctx, hs, T)
    wts = torch.linspace(0.5, 1.5, 32, dtype=torch.float64)
    (E * wts).sum().backward()
    g = dofs.grad.numpy()
    direct = ctx.adjoint(hs, wl.dofs[:32], w=wts.numpy().astype(np.float32))
    assert np.allclose(g[:, :6], direct[:, 1:7], rtol=1e-6, atol=0)
    assert np.allclose(g[:, 6:], direct[:, 7:], rtol=1e-6, atol=0)
    assert table.grad is not None and np.abs(table.grad.numpy()).max() > 0
    # central differences of the GPU's FP32 energy in translation and the
    # Euler angles (h = 1e-3) and the receptor and ligand mode amplitudes
    # (h = 1e-2 A: their gradients are smaller, so the energy's FP32 rounding
    # needs the wider step); the energy is piecewise smooth (trilinear cells,
    # plateau), so the median pins the derivative and a few poses may straddle
    # a kink
    base = wl.dofs[:32].copy()
    cols = [(0, 1e-3, 2e-3), (1, 1e-3, 2e-3), (2, 1e-3, 2e-3), (3, 1e-3, 2e-3), (4, 1e-3, 2e-3),
            (5, 1e-3, 2e-3)]
    cols += [(6 + m, 1e-2, 2e-3) for m in range(wl.rec.M)]
    cols += [(16 + m, 1e-2, 2e-3) for m in range(wl.lig.M)]
    for col, h, tol in cols:
        dp, dm = base.copy(), base.copy()
        dp[:, col] += h
        dm[:, col] -= h
        ep = ctx.adjoint(hs, dp)[:, 0].astype(np.float64)
        em = ctx.adjoint(hs, dm)[:, 0].astype(np.float64)
        fd = (ep - em) / (dp[:, col].astype(np.float64) - dm[:, col].astype(np.float64))
        an = direct[:, 1 + col] / wts.numpy()
        rel = np.abs(fd - an) / (np.abs(an) + np.abs(fd) + 1e-2)
        print(f"DOF column {col}: median relative difference {np.median(rel):.1e}")
        assert np.median(rel) < tol, col
        assert np.mean(rel < 10 * tol) > 0.8, col


def test_adjoint_vs_reference_forces(ctx, O, wl_adj_rigid):
    """The reference's reported force is the interpolated force channel plus
    the pair force minus the plateau force, scaled by 0.01 per step while
    |F|^2 > 1e18 (reducePotForce's clamp); the exact gradient differs from it
    by construction.  Stated over the poses in contact whose force is not
    clamped, with a loose bound on the translation part."""
    wl = wl_adj_rigid
    hs = _upload(ctx, wl)
    g = ctx.adjoint(hs, wl.dofs)
    f = ctx.score_rigid(hs, wl.dofs)
    dt, F = g[:, 1:4].astype(np.float64), -f[:, 1:4].astype(np.float64)
    nF, ng = np.linalg.norm(F, axis=1), np.linalg.norm(dt, axis=1)
    keep = (ng > 1e-3) & (ng < 1e8)
    assert keep.sum() >= 20
    rel = np.linalg.norm(dt - F, axis=1)[keep] / nF[keep]
    print(f"|exact gradient - reference force| / |force| over {keep.sum()} unclamped poses in "
          f"contact: median {np.median(rel):.3f}, p90 {np.percentile(rel, 90):.3f}")
    assert np.median(rel) < 0.5


def test_spatial_pose_order_changes_no_pose_result(W):
    """Poses scored in the spatial order (default) and in input order give the
    same per-pose results bit for bit; the parameter-gradient sums agree to
    summation-order rounding (the poses fall to other workgroups, whose rows
    are added in a fixed order); two runs in one order give the same bits."""
    from gpuattract_amd import hip
    # enough poses for the ordered path (at least four per resident workgroup)
    wl = W.make_workload("order", 61, 600, 150, 12000, 3, 2, with_basis=True)
    T = wl.params.T
    ctx = hip.Context(0)
    try:
        hs = ctx.upload_workload(wl)
        ctx.upload_basis(hs["grid"], wl)
        a, pa = ctx.adjoint(hs, wl.dofs, pgrad=True, T=T)
        ctx.set_pose_order(False)
        b, pb = ctx.adjoint(hs, wl.dofs, pgrad=True, T=T)
        assert np.array_equal(a, b)
        assert np.allclose(pa, pb, rtol=1e-4, atol=1e-5 * np.abs(pb).max())
        ctx.set_pose_order(True)
        c, pc = ctx.adjoint(hs, wl.dofs, pgrad=True, T=T)
        assert np.array_equal(a, c)
        assert np.array_equal(pa, pc)
    finally:
        ctx.close()


def test_concurrent_streams_take_turns_on_scratch(W):
    """Two device-resident calls in flight on two streams of one context use
    its scratch (pose order, parameter-gradient rows) in turn: each returns
    what it returns alone."""
    from gpuattract_amd import hip
    wl = W.make_workload("streams", 62, 600, 150, 12000, 3, 2, with_basis=True)
    T = wl.params.T
    n, cols = wl.dofs.shape[0], 3 * T * T
    ctx = hip.Context(0)
    try:
        hs = ctx.upload_workload(wl)
        ctx.upload_basis(hs["grid"], wl)
        sets = [wl.dofs, wl.dofs.copy()]
        sets[1][:, :3] += np.float32(1.5)
        d_in = [ctx.to_device(d) for d in sets]
        d_out = [ctx.dev_alloc(n * 27 * 4) for _ in sets]
        d_pg = [ctx.dev_alloc(cols * 4) for _ in sets]
        alone = []
        for k in range(2):
            ctx.adjoint_dev(hs, d_in[k].ptr, True, n, None, d_out[k].ptr, d_pg[k].ptr)
            ctx.sync()
            alone.append((ctx.to_host(d_out[k], (n, 27)), ctx.to_host(d_pg[k], (cols,))))
        streams = [ctx.new_stream(), ctx.new_stream()]
        for k in range(2):
            ctx.write(d_out[k], np.full((n, 27), -7.0, np.float32))
        for k in range(2):
            ctx.adjoint_dev(hs, d_in[k].ptr, True, n, None, d_out[k].ptr, d_pg[k].ptr, streams[k])
        for k in range(2):
            ctx.sync(streams[k])
        for k in range(2):
            out, pg = ctx.to_host(d_out[k], (n, 27)), ctx.to_host(d_pg[k], (cols,))
            assert np.array_equal(out.view(np.uint32), alone[k][0].view(np.uint32))
            assert np.array_equal(pg.view(np.uint32), alone[k][1].view(np.uint32))
    finally:
        ctx.close()


def test_parameter_gradient_records_skip_beads_without_grid_term(ctx, W, O):
    """The grid-cell records of beads without a grid term (type-0 beads; the
    beads of poses moved outside both grids) are skipped: the parameter
    gradient equals the FP64 oracle's to 1e-4 of its largest entry and the
    per-bead basis path (the allocation-free entry point) to FP32 rounding;
    the other outputs are the same bits on both paths."""
    wl = W.make_workload("pg edges", 57, 500, 200, 400, 2, 2, with_basis=True, type0_frac=0.1)
    wl.dofs[:50, :3] *= 4.0  # 50 poses far outside the grids
    hs = _upload(ctx, wl)
    T = wl.params.T
    n = wl.dofs.shape[0]
    g, pg = ctx.adjoint(hs, wl.dofs, pgrad=True, T=T)
    _, pref = O.adjoint(wl, pgrad=True)
    pg, pref = pg.reshape(-1), pref.reshape(-1)
    assert np.abs(pg - pref).max() <= 1e-4 * np.abs(pref).max()
    d_in = ctx.to_device(wl.dofs)
    d_out = ctx.dev_alloc(n * 27 * 4)
    d_pg = ctx.dev_alloc(3 * T * T * 4)
    ctx.adjoint_dev_noalloc(hs, d_in.ptr, True, n, None, d_out.ptr, d_pg.ptr)
    ctx.sync()
    pb = ctx.to_host(d_pg, (3 * T * T,))
    assert np.abs(pg - pb).max() <= 1e-5 * np.abs(pg).max()
    assert np.array_equal(ctx.to_host(d_out, (n, 27)).view(np.uint32), g.view(np.uint32))
    assert np.all(g[:50, 0] == 0.0)  # nothing in reach: no energy
