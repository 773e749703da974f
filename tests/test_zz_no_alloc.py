"""att_score_adjoint_dev_noalloc (needs a GPU): the allocation-free variant of
att_score_adjoint_dev returns the same bits as that entry point run in input
pose order for the energies and the DOF and mode gradients, on two calls in a
row (the second with fewer poses, reusing the per-GPU scratch), with and
without parameter gradients, for one and two waves per pose.  The parameter
gradient's grid part is summed in another fixed order (at every bead from the
basis grids, where att_score_adjoint_dev sums per grid cell after sorting
its records), so it agrees to FP32 rounding (1e-5 of its largest entry) and
repeats bit for bit.  Tables beyond its scratch are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("n_lig,modes", [(120, False), (450, True)])
def test_noalloc_adjoint_equals_allocating_entry_point(W, n_lig, modes):
    from gpuattract_amd import hip
    wl = W.make_workload(f"noalloc {n_lig}", 91 + n_lig, 600, n_lig, 3000,
                         *((3, 2) if modes else ()), with_basis=True)
    T = wl.params.T
    ow = 27 if modes else 7
    ctx = hip.Context(0)
    try:
        hs = ctx.upload_workload(wl)
        ctx.upload_basis(hs["grid"], wl)
        ctx.set_pose_order(False)
        w = np.random.default_rng(2).uniform(0.5, 1.5, wl.dofs.shape[0]).astype(np.float32)
        for n in (3000, 700):
            d_in = ctx.to_device(wl.dofs[:n])
            d_w = ctx.to_device(w[:n])
            outs = [ctx.dev_alloc(n * ow * 4) for _ in range(2)]
            pgs = [ctx.dev_alloc(3 * T * T * 4) for _ in range(2)]
            for k in range(2):
                ctx.write(outs[k], np.full((n, ow), np.nan, np.float32))
            ctx.adjoint_dev(hs, d_in.ptr, modes, n, d_w.ptr, outs[0].ptr, pgs[0].ptr)
            ctx.adjoint_dev_noalloc(hs, d_in.ptr, modes, n, d_w.ptr, outs[1].ptr, pgs[1].ptr)
            ctx.sync()
            a, b = (ctx.to_host(o, (n, ow)) for o in outs)
            pa, pb = (ctx.to_host(p, (3 * T * T,)) for p in pgs)
            assert np.array_equal(_bits(a), _bits(b)), n
            scale = np.abs(pa).max()
            assert scale > 0
            err = np.abs(pa - pb).max() / scale
            print(f"{n} poses: parameter gradients agree to {err:.1e} of the largest entry")
            assert err <= 1e-5, n
            ctx.adjoint_dev_noalloc(hs, d_in.ptr, modes, n, d_w.ptr, outs[1].ptr, pgs[1].ptr)
            ctx.sync()
            assert np.array_equal(_bits(ctx.to_host(pgs[1], (3 * T * T,))), _bits(pb)), n
            # without parameter gradients
            ctx.adjoint_dev_noalloc(hs, d_in.ptr, modes, n, None, outs[1].ptr)
            ctx.sync()
            ctx.adjoint_dev(hs, d_in.ptr, modes, n, None, outs[0].ptr)
            ctx.sync()
            assert np.array_equal(_bits(ctx.to_host(outs[0], (n, ow))), _bits(ctx.to_host(outs[1], (n, ow))))
    finally:
        ctx.close()


def test_noalloc_refuses_tables_beyond_its_scratch(W):
    from gpuattract_amd import hip
    wl = W.make_workload("noalloc 20 types", 93, 300, 90, 64, num_types=20, with_basis=True)
    ctx = hip.Context(0)
    try:
        hs = ctx.upload_workload(wl)
        ctx.upload_basis(hs["grid"], wl)
        n = wl.dofs.shape[0]
        d_in = ctx.to_device(wl.dofs)
        d_out = ctx.dev_alloc(n * 7 * 4)
        d_pg = ctx.dev_alloc(3 * 20 * 20 * 4)
        with pytest.raises(hip.AttError) as e:
            ctx.adjoint_dev_noalloc(hs, d_in.ptr, False, n, None, d_out.ptr, d_pg.ptr)
        assert e.value.status == 5  # ATT_ERR_UNSUPPORTED
        # without parameter gradients the scratch is not used
        ctx.adjoint_dev_noalloc(hs, d_in.ptr, False, n, None, d_out.ptr)
        ctx.sync()
        ref = ctx.adjoint(hs, wl.dofs)
        assert np.array_equal(_bits(ctx.to_host(d_out, (n, 7))), _bits(ref))
    finally:
        ctx.close()
