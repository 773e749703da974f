"""Semantics of the host grid generator (csrc/gen/attract_gen.cpp).

It makes the grids that every parity test scores against, and the device
generator must reproduce it bit for bit, so it is pinned here by what the
mathematics fixes: closed forms for one bead, force = -grad E by central
differences beyond the plateau, the plateau force for near beads, Coulomb's
cap, superposition over beads, the table index convention for ligand grids,
and brute-force neighbour lists."""
import numpy as np
import pytest


def body(W, pts, types, charges):
    pts = np.asarray(pts, np.float64)
    xyz = np.ascontiguousarray(pts.T.astype(np.float32).ravel())
    t = np.asarray(types, np.uint32)
    return W.Body(xyz, t, t.copy(), np.asarray(charges, np.float32), np.zeros(0, np.float32),
                  np.zeros(0, np.float32))


def line_geom(W, n, spacing, x0=0.0):
    """n points along x at y = z = 0 (3 x 3 in y, z so the grid is valid)."""
    return W.Geom(n, 3, 3, x0, -spacing, -spacing, spacing)


def at(W, g, data, t, ix, iy=1, iz=1):
    k = ((t * g.nz + iz) * g.ny + iy) * g.nx + ix
    return data[4 * k:4 * k + 4].astype(np.float64)


def pair(p, a, t):
    k = (a - 1) * p.T + (t - 1)
    return (float(p.rc[k]), float(p.ac[k]), float(p.ipon[k]), float(p.rmin2[k]),
            float(p.emin[k]))


@pytest.mark.parametrize("shape", [8, 12])
def test_one_bead_energy_closed_form(W, shape):
    p = W.make_params(41, 4, shape=shape)
    b = body(W, [[0.0, 0.0, 0.0]], [2], [0.0])
    g = line_geom(W, 40, 0.5, 0.25)
    d = W.build_grid(g, b, p, False, 10.0).data
    s = 8 if shape == 8 else 12
    for t in range(1, 5):
        rc, ac, ipon, _, _ = pair(p, 2, t)
        for ix in range(g.nx):
            r = 0.25 + 0.5 * ix
            rp = max(r, 10.0)
            e = ipon * (rc / rp ** s - ac / rp ** 6)
            got = at(W, g, d, t, ix)[3]
            assert got == pytest.approx(e, rel=2e-6, abs=1e-30), (t, r)
        # no electrostatics from an uncharged bead
        assert np.all(at(W, g, d, 0, 5) == 0.0)


def test_force_is_minus_energy_gradient_beyond_plateau(W):
    p = W.make_params(42, 3)
    b = body(W, [[0.3, -0.2, 0.1], [-1.0, 0.5, 0.0], [0.4, 0.9, -0.7]], [1, 3, 2], [1, -1, 1])
    h = 1e-2
    for center in ([12.0, 1.0, -0.5], [7.0, 9.0, 3.0], [-4.0, -11.0, 2.0]):
        g = W.Geom(3, 3, 3, center[0] - h, center[1] - h, center[2] - h, h)
        d = W.build_grid(g, b, p, False, 10.0).data
        for t in range(0, 4):
            F = at(W, g, d, t, 1, 1, 1)[:3]
            fd = -np.array([(at(W, g, d, t, 2, 1, 1)[3] - at(W, g, d, t, 0, 1, 1)[3]),
                            (at(W, g, d, t, 1, 2, 1)[3] - at(W, g, d, t, 1, 0, 1)[3]),
                            (at(W, g, d, t, 1, 1, 2)[3] - at(W, g, d, t, 1, 1, 0)[3])]) / (2 * h)
            scale = np.abs(F).max() + 1e-30
            assert np.allclose(F, fd, atol=3e-3 * scale), (t, center, F, fd)


def test_near_bead_plateau_value_and_force(W):
    p = W.make_params(43, 2)
    b = body(W, [[0.0, 0.0, 0.0]], [1], [0.0])
    g = line_geom(W, 30, 0.5, 0.5)
    d = W.build_grid(g, b, p, False, 10.0).data
    for t in (1, 2):
        at_p = at(W, g, d, t, 19)          # r = 10.0 exactly
        for ix in range(19):               # r < 10: V(P) and the plateau force
            v = at(W, g, d, t, ix)
            assert v[3] == at_p[3]
            assert v[0] == pytest.approx(at_p[0], rel=1e-6) and v[1] == 0 and v[2] == 0


@pytest.mark.parametrize("dielec", [0, 1])
def test_coulomb_closed_form_and_cap(W, dielec):
    p = W.make_params(44, 2, dielec=dielec)
    b = body(W, [[0.0, 0.0, 0.0]], [1], [-1.0])
    g = line_geom(W, 70, 1.0, 0.5)
    d = W.build_grid(g, b, p, False, 10.0).data
    fe = float(np.float32(p.ffelec))
    for ix in range(g.nx):
        r = 0.5 + ix
        rp = max(r, 10.0)
        if dielec == 0:
            e = -fe * (1 / rp - 1 / 50.0)
        else:
            e = -fe * (1 / rp ** 2 - 1 / 2500.0)
        e = e if r < 50.0 else 0.0
        assert at(W, g, d, 0, ix)[3] == pytest.approx(e, rel=2e-6, abs=1e-12), r


def test_superposition_over_beads(W):
    rng = np.random.default_rng(45)
    p = W.make_params(45, 5)
    pts = rng.uniform(-8, 8, (12, 3))
    types = rng.integers(1, 6, 12)
    q = rng.choice([-1.0, 0.0, 1.0], 12)
    g = W.Geom(7, 6, 5, -14.0, -12.0, -9.0, 4.0)
    full = W.build_grid(g, body(W, pts, types, q), p, False, 10.0).data.astype(np.float64)
    parts = sum(W.build_grid(g, body(W, pts[k:k + 1], types[k:k + 1], q[k:k + 1]), p, False,
                             10.0).data.astype(np.float64) for k in range(12))
    assert np.allclose(full, parts, rtol=1e-5, atol=1e-6 * np.abs(full).max())


def test_ligand_grids_use_the_transposed_table(W):
    p = W.make_params(46, 4)
    pt = W.make_params(46, 4)
    # the generated table is symmetric: make it asymmetric
    up = np.triu(np.ones((4, 4), bool), 1).ravel()
    p.rc[up] *= np.float32(1.3)
    p.ac[up] *= np.float32(0.8)
    for name in ("rc", "ac", "ipon", "rmin2", "emin"):
        a = getattr(p, name).reshape(4, 4)
        setattr(pt, name, np.ascontiguousarray(a.T).ravel())
    assert not np.array_equal(p.rc, pt.rc)
    b = body(W, [[0.0, 0.0, 0.0], [2.0, 1.0, 0.5]], [1, 3], [0.0, 1.0])
    g = W.Geom(5, 4, 3, 9.0, -3.0, -2.0, 1.5)
    lig = W.build_grid(g, b, p, True, 10.0).data
    rec = W.build_grid(g, b, pt, False, 10.0).data
    assert np.array_equal(lig, rec)


def test_neighbour_lists_brute_force(W):
    rng = np.random.default_rng(47)
    pts = rng.uniform(-10, 10, (60, 3)).astype(np.float32)
    types = rng.integers(0, 4, 60)
    b = body(W, pts, types, np.zeros(60))
    g = W.Geom(9, 8, 7, -12.0, -11.0, -10.0, 3.0)
    nl = W.build_nl(g, b, 6.5, 10.0)
    P = pts.astype(np.float64)
    for iz in range(g.nz):
        for iy in range(g.ny):
            for ix in range(g.nx):
                v = (iz * g.ny + iy) * g.nx + ix
                c = np.array([g.ox + ix * np.float64(np.float32(g.spacing)),
                              g.oy + iy * np.float64(np.float32(g.spacing)),
                              g.oz + iz * np.float64(np.float32(g.spacing))])
                d2 = ((c - P) ** 2).sum(1)
                want = np.flatnonzero((d2 <= 6.5 ** 2) & (types != 0))
                cnt, start = nl.desc[2 * v], nl.desc[2 * v + 1]
                assert np.array_equal(nl.list[start:start + cnt], want), (ix, iy, iz)
