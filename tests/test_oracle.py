"""The CPU oracle, pinned (CPU only).

1. Against the reference's own compiled CPU code: tests/golden/*.npz were
   written by tests/golden/make_golden.py from oracle/_ref (the reference's
   CPUEnergyService6D item processor and model functions).  The oracle must be
   bit-identical on every pose: results, per-bead coordinates, forces and
   energies, grid cells and neighbour-list pair counts.
2. Against mathematics, independent of the reference: rotation and torque
   matrices, trilinear exactness on affine fields, the 8-6/12-6 pair potential
   with its plateau subtraction, Coulomb, and energy/gradient consistency.
"""
import json
import os

import numpy as np
import pytest

from helpers import affine_grid, body, workload, zero_grid


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RIGID = ["rigid_300_80", "rigid_1000_300", "rigid_type0", "rigid_12_6", "rigid_const_dielec"]
MODES = ["modes_300_90", "modes_3000_800", "modes_rec_only"]


def _regenerate(W, name):
    f = np.load(os.path.join(GOLD, name + ".npz"))
    spec = json.loads(str(f["spec"]))
    wl = W.make_workload(spec["name"], *spec["args"], **spec["kw"])
    from make_golden_sums import input_checksums
    assert input_checksums(W, wl) == json.loads(str(f["checksums"])), \
        "the generator no longer reproduces the golden inputs"
    assert np.array_equal(wl.dofs, f["dofs"])
    return wl, f


def _same_bits(a, b):
    """Equal floats bit for bit (NaNs compare by their bits too)."""
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("name", RIGID)
def test_rigid_against_reference(W, O, name):
    wl, f = _regenerate(W, name)
    o = O.score_rigid(wl, nthreads=4)
    assert _same_bits(o, f["ref"])
    for k in range(f["bead_xyz"].shape[0]):
        b = O.beads(wl, wl.dofs[k])
        assert _same_bits(b["xyz"], f["bead_xyz"][k])
        assert _same_bits(b["pot"], f["bead_pot"][k])
        assert np.array_equal(b["cell"], f["bead_cells"][k])
        assert np.array_equal(b["nl_pairs"], f["bead_npairs"][k])
    # the fixtures exercise every branch of the cell logic
    cells = f["bead_cells"]
    assert (cells[..., 0] >= 0).any() and (cells[..., 3] >= 0).any() and (cells[..., 6] >= 0).any()
    assert (f["bead_npairs"] > 0).any()


@pytest.mark.parametrize("name", RIGID)
def test_pair_counts_against_reference(W, O, name):
    """Per-pose pair counts {list entries walked, pairs inside the plateau}
    from the reference's own per-bead values (its neighbour-list cells, the
    list lengths of the generated grid, its per-bead pair counts), against the
    oracle's."""
    wl, f = _regenerate(W, name)
    k = f["bead_cells"].shape[0]
    got = O.pair_counts(wl, wl.dofs[:k])
    ge = wl.rec_nl.geom
    lengths = wl.rec_nl.desc.reshape(-1, 2)[:, 0]
    for j in range(k):
        c = f["bead_cells"][j][:, 6:9]
        m = c[:, 0] >= 0
        v = (c[m, 2].astype(np.int64) * ge.ny + c[m, 1]) * ge.nx + c[m, 0]
        assert got[j, 0] == lengths[v].sum()
        assert got[j, 1] == f["bead_npairs"][j].sum()
    assert (got[:, 1] > 0).any() and (got[:, 0] >= got[:, 1]).all()


@pytest.mark.parametrize("name", MODES)
def test_modes_against_reference(W, O, name):
    wl, f = _regenerate(W, name)
    o = O.score_modes(wl, nthreads=4)
    assert _same_bits(o, f["ref"])
    b = O.beads(wl, wl.dofs[0])
    assert _same_bits(b["rec_pot"], f["rec_pot"])
    assert _same_bits(b["pot"], f["lig_pot"])


def test_param_derivation_against_reference(W):
    """attg_derive_params (the product's restatement of readFile.cpp:479-513)."""
    f = np.load(os.path.join(GOLD, "params.npz"))
    for shape in (8, 12):
        p = W.make_params(301, T=12, shape=shape)
        assert np.array_equal(p.rbc, f[f"{shape}_rbc"]) and np.array_equal(p.abc, f[f"{shape}_abc"])
        for k in ("rc", "ac", "ipon", "rmin2", "emin"):
            assert np.array_equal(getattr(p, k), f[f"{shape}_{k}"]), (shape, k)


def test_ratio_table_against_reference(W, O):
    f = np.load(os.path.join(GOLD, "ratio.npz"))
    lut = O.ratio_lut(10.0)
    assert len(lut) == int(f["size"])
    assert np.array_equal(lut[f["idx"]], f["val"])
    assert str(W.checksum(lut)) == str(f["checksum"])


def test_rotmats_against_reference(O):
    f = np.load(os.path.join(GOLD, "rotmat.npz"))
    R = np.stack([O.rotmat(*a) for a in f["ang"]])
    TM = np.stack([O.torquemat(*a) for a in f["ang"]])
    assert _same_bits(R, f["R"]) and _same_bits(TM, f["TM"])


# ------------------------------------------------------------ closed forms

def test_rotation_matrix_identities(O):
    rng = np.random.default_rng(0)
    for a in rng.uniform(-7, 7, (200, 3)).astype(np.float32):
        R = O.rotmat(*a).astype(np.float64).reshape(3, 3)
        assert np.abs(R @ R.T - np.eye(3)).max() < 1e-6
        assert abs(np.linalg.det(R) - 1) < 1e-6
    assert np.array_equal(O.rotmat(0, 0, 0), np.eye(3, dtype=np.float32).ravel())
    # ssi = 0: a pure z rotation by rot - phi (model/matrixFunctions.h:64-74)
    for phi, rot in rng.uniform(-3, 3, (50, 2)):
        R = O.rotmat(phi, 0.0, rot).reshape(3, 3)
        c, s = np.cos(rot - phi), np.sin(rot - phi)
        assert np.allclose(R, [[c, s, 0], [-s, c, 0], [0, 0, 1]], atol=3e-7)


def test_torque_matrix_is_rotation_derivative(O):
    """tm[k][j][l] = dR[k][l] / d(angle_j), angles (phi, ssi, rot)."""
    rng = np.random.default_rng(1)
    h = 1e-2
    for a in rng.uniform(-3, 3, (40, 3)):
        tm = O.torquemat(*a.astype(np.float32)).astype(np.float64).reshape(3, 3, 3)
        for j in range(3):
            ap, am = a.copy(), a.copy()
            ap[j] += h
            am[j] -= h
            # rotation matrix in double for the finite difference
            fd = (_rot64(ap) - _rot64(am)) / (2 * h)
            assert np.abs(tm[:, j, :] - fd).max() < 1e-4


def _rot64(a):
    phi, ssi, rot = a
    cS, sS, cP, sP, cR, sR = np.cos(ssi), np.sin(ssi), np.cos(phi), np.sin(phi), np.cos(rot), np.sin(rot)
    return np.array([[cR * cS * cP + sR * sP, sR * cS * cP - cR * sP, sS * cP],
                     [cR * cS * sP - sR * cP, sR * cS * sP + cR * cP, sS * sP],
                     [-cR * sS, -sR * sS, cS]])


def _far_receptor(W):
    return body(W, [[0.0, 0.0, 0.0], [3.2, 0.0, 0.0]], [1, 2], [0.0, 0.0])


def test_trilinear_exact_on_affine_fields(W, O):
    """Interpolation reproduces an affine field up to rounding, and its force
    channel; a one-bead ligand placed at random points inside the grid."""
    params = W.make_params(5, T=3)
    rec = _far_receptor(W)
    lig = body(W, [[0.0, 0.0, 0.0]], [2], [1.0])
    g = np.array([0.3, -0.2, 0.5])
    gi = W.inner_geom(rec, 0.875, 10.0)
    inner = affine_grid(W, gi, params.T, 1.5, g)
    outer = affine_grid(W, W.cube_geom((0.0, 0.0, 0.0), 60.0, 2.6), params.T, 1.5, g)
    rng = np.random.default_rng(2)
    lo = np.array([gi.ox, gi.oy, gi.oz]) + 0.5
    hi = lo + (np.array([gi.nx, gi.ny, gi.nz]) - 2) * gi.spacing
    pts = rng.uniform(lo, hi, (64, 3))
    dofs = np.zeros((64, 6), np.float32)
    dofs[:, :3] = pts
    wl = workload(W, rec, lig, params, dofs, inner, outer, plateau=1.0, neighbourdis=1.1)
    # NL lists only reach 1.1 A from the receptor beads: keep points outside.
    far = np.min(np.linalg.norm(pts[:, None, :] - rec.xyz.reshape(3, -1).T[None], axis=2), axis=1) > 2.5
    res = O.score_rigid(wl)[far]
    p = dofs[far, :3].astype(np.float64)
    E = 2 * (1.5 + p @ g)  # type grid + charge(1) * electrostatic grid
    assert np.allclose(res[:, 0], E, rtol=2e-6, atol=1e-5)
    assert np.allclose(res[:, 1:4], -2 * g[None, :], rtol=1e-5, atol=1e-6)


def _pair_workload(W, params, r, direction, q_rec=0.0, q_lig=0.0, t_rec=1, t_lig=2):
    rec = body(W, [[0.0, 0.0, 0.0]], [t_rec], [q_rec])
    lig = body(W, [[0.0, 0.0, 0.0]], [t_lig], [q_lig])
    d = np.asarray(direction, np.float64)
    d /= np.linalg.norm(d)
    dofs = np.zeros((len(r), 6), np.float32)
    dofs[:, :3] = np.asarray(r)[:, None] * d[None, :]
    return workload(W, rec, lig, params, dofs), d


def _lj64(r2, p, k, shape):
    rc, ac, ivor, rmin2, emin = (float(p.rc[k]), float(p.ac[k]), float(p.ipon[k]),
                                 float(p.rmin2[k]), float(p.emin[k]))
    i2 = 1.0 / r2
    rr23 = i2 ** 3
    rep = rc * (i2 if shape == 8 else rr23)
    vlj = (rep - ac) * rr23
    f = 6 * vlj + (2 if shape == 8 else 6) * rep * rr23
    if r2 < rmin2:
        return vlj + (ivor - 1) * emin, f
    return ivor * vlj, ivor * f


@pytest.mark.parametrize("shape", [8, 12])
def test_pair_potential_closed_form(W, O, shape):
    """Neighbour-list term of one pair against the closed form: V(r) - V(P)
    with V the 8-6 (12-6) potential and its r < rmin branch, and force
    fb(r) d/r^2 minus the plateau force fb(P) d/(r P) (model/neighborlist.h:90-157)."""
    params = W.make_params(7, T=3, shape=shape, repulsive_frac=0.5)
    P = 10.0
    k = params.T * 0 + 1  # recType 1, ligType 2
    r = np.linspace(2.6, 9.99, 60)
    wl, d = _pair_workload(W, params, r, [0.3, -0.5, 0.8])
    res = O.score_rigid(wl)
    for i, ri in enumerate(r):
        e_r, f_r = _lj64(ri * ri, params, k, shape)
        e_p, f_p = _lj64(P * P, params, k, shape)
        E = e_r - e_p
        F = (f_r / ri - f_p / P) * d  # = fb(r) d/r^2 - fb(P) d/(rP), d unit
        scale = abs(e_r) + abs(e_p)
        assert abs(res[i, 0] - E) <= 1e-5 * scale + 1e-7
        assert np.abs(res[i, 1:4] - F).max() <= 1e-5 * (abs(f_r / ri) + abs(f_p / P)) + 1e-7
    # continuity: the pair term vanishes at the plateau
    wl2, _ = _pair_workload(W, params, [P - 1e-4], [1, 0, 0])
    assert abs(O.score_rigid(wl2)[0, 0]) < 1e-4 * abs(_lj64(P * P, params, k, shape)[0]) + 1e-7
    # energy minimum of the 8-6 potential: V(rmin) = emin for an attractive pair
    if shape == 8:
        kk = int(np.argmax(params.ipon == 1))
        e_min, f_min = _lj64(float(params.rmin2[kk]) * (1 + 1e-9), params, kk, 8)
        assert abs(e_min - float(params.emin[kk])) < 1e-5 * abs(float(params.emin[kk]))
        assert abs(f_min) < 1e-3 * abs(float(params.emin[kk]))


@pytest.mark.parametrize("dielec", [0, 1])
def test_coulomb_closed_form(W, O, dielec):
    params = W.make_params(9, T=3, dielec=dielec)
    # make the LJ part vanish: zero the pair parameters
    for a in ("rc", "ac", "emin", "rmin2"):
        getattr(params, a)[:] = 0
    P = 10.0
    r = np.linspace(3.0, 9.9, 30)
    wl, d = _pair_workload(W, params, r, [1, 2, 2], q_rec=1.0, q_lig=-1.0)
    res = O.score_rigid(wl)
    c = -1.0 * float(params.ffelec)
    if dielec == 1:
        E = c * ((1 / r ** 2 - 1 / 2500.0) - (1 / P ** 2 - 1 / 2500.0))
        Fm = 2 * c / r ** 3 - 2 * c / P ** 3
    else:
        E = c * ((1 / r - 1 / 50.0) - (1 / P - 1 / 50.0))
        Fm = c / r ** 2 - c / P ** 2
    assert np.allclose(res[:, 0], E, rtol=2e-5, atol=1e-6)
    assert np.allclose(res[:, 1:4], Fm[:, None] * d[None, :], rtol=2e-5, atol=1e-6)


def test_gradients_are_energy_derivatives_on_affine_field(W, O):
    """With an affine grid (force channel = -grad E) the reported pos/ang of a
    rigid pose are -dE/dt and -dE/d(angles) (model/reduction.h:39-104,
    matrixFunctions.h:79-127): checked by central differences."""
    params = W.make_params(11, T=4)
    rec = _far_receptor(W)
    rng = np.random.default_rng(3)
    lig = body(W, rng.uniform(-4, 4, (12, 3)), rng.integers(1, 5, 12), np.zeros(12))
    g = np.array([0.4, 0.1, -0.3])
    inner = affine_grid(W, W.inner_geom(rec, 0.875, 20.0), params.T, 0.7, g)
    outer = affine_grid(W, W.cube_geom((0.0, 0.0, 0.0), 60.0, 2.6), params.T, 0.7, g)
    base = np.array([[14.0, -3.0, 5.0, 0.7, 1.1, -0.4]])
    wl = workload(W, rec, lig, params, base, inner, outer, plateau=1.0, neighbourdis=1.1)
    r0 = O.score_rigid(wl)[0]
    h = [1e-2] * 3 + [2e-3] * 3
    for j in range(6):
        dp, dm = base.copy(), base.copy()
        dp[0, j] += h[j]
        dm[0, j] -= h[j]
        e = O.score_rigid(wl, np.vstack([dp, dm]).astype(np.float32), f64=True)[0][:, 0]
        fd = (e[0] - e[1]) / (2 * h[j])
        assert abs(-fd - r0[1 + j]) < 2e-3 * (abs(fd) + 1), (j, fd, r0[1 + j])


def test_ligand_mode_gradient_is_energy_derivative(W, O):
    """Flexible scoring: the ligand mode entries are +dE/da (before the
    correction, here k = 0), model/reduction_modes.h:28-57."""
    params = W.make_params(13, T=3)
    rng = np.random.default_rng(4)
    n = 10
    rec = body(W, [[0.0, 0.0, 0.0], [3.1, 0.0, 0.0]], [1, 2], [0.0, 0.0],
               modes=np.zeros(3 * 2 * 1), mode_force=[0.0])
    M = 2
    mv = rng.normal(size=(3, n, M))
    lig = body(W, rng.uniform(-4, 4, (n, 3)), rng.integers(1, 4, n), np.zeros(n),
               modes=mv.reshape(3, n * M), mode_force=np.zeros(M))
    g = np.array([0.2, -0.6, 0.3])
    inner = affine_grid(W, W.inner_geom(rec, 0.875, 20.0), params.T, 0.5, g)
    outer = affine_grid(W, W.cube_geom((0.0, 0.0, 0.0), 60.0, 2.6), params.T, 0.5, g)
    lgeo = W.inner_geom(lig, 0.875, 5.0)
    lig_inner = zero_grid(W, lgeo, params.T)
    lig_outer = zero_grid(W, W.cube_geom((0.0, 0.0, 0.0), 10.0, 2.6), params.T)
    base = np.zeros((1, 26))
    base[0, :6] = [15.0, 2.0, -4.0, 0.3, 0.9, 1.7]
    base[0, 16:18] = [0.4, -0.7]
    wl = workload(W, rec, lig, params, base, inner, outer, plateau=1.0, neighbourdis=1.1,
                  lig_inner=lig_inner, lig_outer=lig_outer)
    r0 = O.score_modes(wl)[0]
    for m in range(M):
        dp, dm = base.copy(), base.copy()
        dp[0, 16 + m] += 1e-2
        dm[0, 16 + m] -= 1e-2
        e = O.score_modes(wl, np.vstack([dp, dm]).astype(np.float32), f64=True)[0][:, 0]
        fd = (e[0] - e[1]) / 2e-2
        assert abs(fd - r0[17 + m]) < 2e-3 * (abs(fd) + 1)


def test_force_clamp(W, O):
    """reducePotForce scales |F| > 1e9 by 0.01 up to three times (model/reduction.h:60-72)."""
    params = W.make_params(15, T=2)
    rec = _far_receptor(W)
    lig = body(W, [[0.0, 0.0, 0.0]], [1], [0.0])
    gi = W.inner_geom(rec, 0.875, 10.0)
    out = []
    for gx in (1e8, 1e10, 1e12, 1e14, 1e16):
        g = np.array([gx, 0.0, 0.0])
        inner = affine_grid(W, gi, params.T, 0.0, g)
        outer = affine_grid(W, W.cube_geom((0.0, 0.0, 0.0), 60.0, 2.6), params.T, 0.0, g)
        wl = workload(W, rec, lig, params, np.array([[8.0, 8.0, 8.0, 0, 0, 0]]), inner, outer,
                      plateau=1.0, neighbourdis=1.1)
        out.append(-O.score_rigid(wl)[0, 1])
    assert np.allclose(out, [1e8, 1e8, 1e8, 1e8, 1e10], rtol=1e-5)
