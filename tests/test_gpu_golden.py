"""The HIP library against the reference's own outputs (needs a GPU).

tests/golden/*.npz hold what the reference's CPU code, compiled from its
sources (oracle/_ref), returned for seeded workloads (tests/golden/
make_golden.py).  Here the drop-in paths run on the GPU and are compared with
those arrays directly, with no restatement in between:

* per bead (rigid fixtures, the first 16 poses): transformed coordinates,
  forces and energies, grid cells and neighbour-list pair counts, bit for bit;
  (modes fixtures, pose 0) receptor and ligand forces bit for bit;
* per pose, through the C-ABI (att_score_rigid / att_score_modes) and
  through the C++ drop-in Server (MI355EnergyService6D, items of 100 poses):
  energy, force and torque bit for bit (the kernels sum every bead's pair
  terms in list order and the beads in bead order, as the reference does),
  and for the flexible path the ligand mode gradients bit for bit; the
  receptor mode gradients (the reaction forces are projected on the modes per
  pair instead of being summed per receptor bead first) within 1e-5 of the
  gradient's norm over the receptor modes, their plain relative errors
  printed;
* the adjoint kernel's energy bit for bit (its gradient is the exact one,
  which the reference does not compute).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
RIGID = ["rigid_300_80", "rigid_1000_300", "rigid_type0", "rigid_12_6", "rigid_const_dielec"]
MODES = ["modes_300_90", "modes_3000_800", "modes_rec_only"]
TOL_E = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from gpuattract_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def _load(W, name):
    f = np.load(os.path.join(GOLD, name + ".npz"))
    spec = json.loads(str(f["spec"]))
    wl = W.make_workload(spec["name"], *spec["args"], **spec["kw"])
    assert np.array_equal(wl.dofs, f["dofs"])
    return wl, f


def plain_rel(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.abs(x - y) / np.abs(y)
    r[x == y] = 0.0
    return r


def _report(name, got, ref, cols):
    parts = []
    for c in cols:
        r = plain_rel(got[:, c], ref[:, c])
        parts.append(f"c{c} max {r.max():.1e} exact {np.mean(got[:, c] == ref[:, c]):.2f}")
    print(name + ": " + "; ".join(parts))


def _same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("name", RIGID)
def test_rigid_drop_in_against_reference(ctx, W, name):
    wl, f = _load(W, name)
    ref = f["ref"]
    hs = ctx.upload_workload(wl)
    got = ctx.score_rigid(hs, wl.dofs)
    _report(name, got, ref, range(7))
    assert _same_bits(got, ref)
    for k in range(f["bead_xyz"].shape[0]):
        b = ctx.beads(hs, wl, wl.dofs[k])
        assert _same_bits(b["xyz"], f["bead_xyz"][k]), k
        assert _same_bits(b["pot"], f["bead_pot"][k]), k
        assert np.array_equal(b["cell"], f["bead_cells"][k]), k
        assert np.array_equal(b["nl_pairs"], f["bead_npairs"][k]), k
    # the adjoint kernel's energy (the energy of its exact gradient)
    a = ctx.adjoint(hs, wl.dofs)
    assert _same_bits(a[:, 0], ref[:, 0])


@pytest.mark.parametrize("name", MODES)
def test_modes_drop_in_against_reference(ctx, W, name):
    wl, f = _load(W, name)
    ref = f["ref"]
    hs = ctx.upload_workload(wl)
    got = ctx.score_modes(hs, wl.dofs)
    cols = [0, 1, 2, 3, 4, 5, 6] + [7 + m for m in range(wl.rec.M)] + [17 + m for m in range(wl.lig.M)]
    _report(name, got, ref, cols)
    lig_cols = [0, 1, 2, 3, 4, 5, 6] + [17 + m for m in range(wl.lig.M)]
    assert _same_bits(got[:, lig_cols], ref[:, lig_cols])
    if wl.rec.M:
        rc = slice(7, 7 + wl.rec.M)
        g, r = got[:, rc].astype(np.float64), ref[:, rc].astype(np.float64)
        fin = np.all(np.isfinite(r), axis=1)
        norm = np.linalg.norm(r[fin], axis=1)
        err = np.linalg.norm(g[fin] - r[fin], axis=1) / np.maximum(norm, 1e-30)
        print(f"{name} receptor modes: normwise max {err.max():.1e}")
        assert err.max() <= TOL_E
        assert np.array_equal(np.isfinite(got[:, rc]), np.isfinite(ref[:, rc]))
    b = ctx.beads(hs, wl, wl.dofs[0])
    assert _same_bits(b["rec_pot"], f["rec_pot"])
    assert _same_bits(b["pot"], f["lig_pot"])
    a = ctx.adjoint(hs, wl.dofs)
    assert _same_bits(a[:, 0], ref[:, 0])


def test_server_drop_in_against_reference(W):
    """as::Server with MI355EnergyService6D<float> (items of 100 poses, two
    workers, four requests in flight) on the rigid 1000/300 fixture."""
    from gpuattract_amd import hip
    L = C.CDLL(os.path.join(HERE, "cpp", "_bin", "libservice_driver.so"))
    L.drv_error.restype = C.c_char_p
    wl, f = _load(W, "rigid_1000_300")
    p = wl.params
    tab = hip.param_table(p)
    rec, lig = hip.protein_desc(wl.rec), hip.protein_desc(wl.lig)
    gi, go, nl = hip.grid_desc(wl.rec_inner), hip.grid_desc(wl.rec_outer), hip.nl_desc(wl.rec_nl)
    d = np.ascontiguousarray(wl.dofs, np.float32)
    out = np.full((d.shape[0], 7), np.nan, np.float32)
    st = L.drv_score_rigid(C.byref(rec), C.byref(lig), C.byref(gi), C.byref(go), C.byref(nl),
                           hip._p(tab), p.T, p.shape, p.dielec, C.c_float(p.epsilon),
                           C.c_float(p.ffelec), hip._p(d), C.c_size_t(d.shape[0]), hip._p(out),
                           2, 100, 4)
    assert st == 0, L.drv_error().decode()
    _report("server rigid_1000_300", out, f["ref"], range(7))
    assert _same_bits(out, f["ref"])
