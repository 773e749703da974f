"""Grid generation on the device (needs a GPU).

The device generator sums in double in the host generator's order with
correctly rounded division and square root, so every grid value and every
neighbour list must equal the host generator's (csrc/gen/attract_gen.cpp,
whose semantics tests/test_gridgen_semantics.py pins) bit for bit; a grid
handle built on the device must score exactly like one uploaded from the host
grids."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gpuattract_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def cases(W):
    rec = W.make_body(31, 0, 700, W.scaled_radius(700, W.DENSITY_REF), 12, 0.2, 0.05)
    lig = W.make_body(31, 1, 150, W.scaled_radius(150, W.LIG_REF), 12, 0.2, 0.05)
    p8 = W.make_params(31, 12)
    p12 = W.make_params(32, 12, shape=12, dielec=0)
    p15 = W.make_params(33, 15)  # more probe types than one kernel pass holds
    big = W.make_body(34, 0, 500, W.scaled_radius(500, W.DENSITY_REF), 15, 0.3, 0.0)
    gi = W.inner_geom(rec, 1.3, 6.0)
    gl = W.inner_geom(lig, 0.9, 6.0)
    go = W.cube_geom((0.0, 0.0, 0.0), 35.0, 3.1)
    gb = W.inner_geom(big, 1.7, 4.0)
    return [
        ("receptor inner, 8-6", rec, p8, gi, False, 10.0),
        ("receptor outer, 8-6", rec, p8, go, False, 10.0),
        ("ligand grid", lig, p8, gl, True, 10.0),
        ("12-6, constant dielectric", rec, p12, gi, False, 10.0),
        ("plateau below rmin (full pair potential)", rec, p8, gi, False, 3.5),
        ("15 types", big, p15, gb, False, 10.0),
    ]


@pytest.mark.parametrize("k", range(6))
def test_device_grid_equals_host_generator(ctx, W, k):
    name, body, p, geom, is_lig, plateau = cases(W)[k]
    dev = ctx.generate_grid(body, p, geom, plateau, is_lig)
    host = W.build_grid(geom, body, p, is_lig, plateau).data
    bad = np.flatnonzero(dev.view(np.uint32) != host.view(np.uint32))
    print(f"{name}: {geom.nvox} points x {p.T + 1} grids, {bad.size} values differ")
    assert bad.size == 0, (name, bad[:8], dev[bad[:8]], host[bad[:8]])


@pytest.mark.parametrize("nd", [10.8, 6.0])
def test_device_neighbour_lists_equal_host(ctx, W, nd):
    for name, body, p, geom, is_lig, plateau in cases(W)[:3]:
        desc, lst = ctx.generate_nl(body, geom, nd)
        h = W.build_nl(geom, body, nd, plateau)
        assert np.array_equal(desc, h.desc), name
        assert np.array_equal(lst, h.list), name


def test_built_grid_scores_like_uploaded(ctx, W):
    wl = W.make_workload("gridgen scoring", 35, 900, 200, 300, 3, 2, with_lig_grid=True)
    hs = ctx.upload_workload(wl)
    gi, go = wl.rec_inner.geom, wl.rec_outer.geom
    g_dev = ctx.build_grid(wl.rec, wl.params, gi, go, 10.0, 10.8)
    hd = dict(hs, grid=g_dev)
    assert np.array_equal(ctx.score_modes(hd, wl.dofs), ctx.score_modes(hs, wl.dofs))
    assert np.array_equal(ctx.adjoint(hd, wl.dofs), ctx.adjoint(hs, wl.dofs))
    # ligand grids too
    lg = ctx.build_grid(wl.lig, wl.params, wl.lig_inner.geom, wl.lig_outer.geom, 10.0, 0.0,
                        src_is_ligand=True)
    hl = dict(hs, lig_grid=lg)
    assert np.array_equal(ctx.score_modes(hl, wl.dofs), ctx.score_modes(hs, wl.dofs))


def test_grid_generation_errors(ctx, W):
    from gpuattract_amd import hip
    rec = W.make_body(36, 0, 100, W.scaled_radius(100, W.DENSITY_REF))
    p = W.make_params(36)
    g = W.inner_geom(rec, 1.0, 5.0)
    bad = W.Geom(1, g.ny, g.nz, g.ox, g.oy, g.oz, g.spacing)
    with pytest.raises(hip.AttError):
        ctx.generate_grid(rec, p, bad)
    neg = W.Geom(g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, -1.0)
    with pytest.raises(hip.AttError):
        ctx.build_grid(rec, p, neg, g)
    with pytest.raises(hip.AttError):
        ctx.generate_grid(rec, p, g, plateau=0.0)
    with pytest.raises(hip.AttError):
        ctx.generate_nl(rec, g, -2.0)


def test_workload_without_host_grids(ctx, W):
    """make_workload(host_grids=False): upload_workload builds the receptor
    grids on the device and scores exactly like the host-built workload;
    fill_grids_from_device returns the host generator's arrays."""
    args = ("no host grids", 37, 800, 250, 200, 2, 3)
    wh = W.make_workload(*args, with_lig_grid=True)
    wd = W.make_workload(*args, with_lig_grid=True, host_grids=False)
    assert wd.rec_inner.data is None and wd.rec_nl.desc is None
    hh, hd = ctx.upload_workload(wh), ctx.upload_workload(wd)
    assert np.array_equal(ctx.score_modes(hd, wd.dofs), ctx.score_modes(hh, wh.dofs))
    assert np.array_equal(ctx.adjoint(hd, wd.dofs), ctx.adjoint(hh, wh.dofs))
    W.fill_grids_from_device(ctx, wd)
    for a, b in ((wd.rec_inner, wh.rec_inner), (wd.rec_outer, wh.rec_outer)):
        assert np.array_equal(a.data.view(np.uint32), b.data.view(np.uint32))
    assert np.array_equal(wd.rec_nl.desc, wh.rec_nl.desc)
    assert np.array_equal(wd.rec_nl.list, wh.rec_nl.list)
