"""Hand-built workloads for closed-form tests (small bodies, custom grids)."""
import numpy as np


def body(W, xyz, types, charges, modes=None, mode_force=None):
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    t = np.asarray(types, np.uint32).reshape(n)
    m = np.zeros(0, np.float32) if modes is None else np.asarray(modes, np.float32).ravel()
    mf = np.zeros(0, np.float32) if mode_force is None else np.asarray(mode_force, np.float32)
    return W.Body(np.ascontiguousarray(xyz.T.astype(np.float32).ravel()), t, t.copy(),
                  np.asarray(charges, np.float32).reshape(n), m, mf)


def zero_grid(W, geom, T):
    return W.IGrid(geom, T + 1, np.zeros((T + 1) * geom.nvox * 4, np.float32))


def affine_grid(W, geom, T, a, g):
    """E = a + g.p and force = -g at every grid point, for every grid."""
    ix = np.arange(geom.nx) * np.float64(geom.spacing) + geom.ox
    iy = np.arange(geom.ny) * np.float64(geom.spacing) + geom.oy
    iz = np.arange(geom.nz) * np.float64(geom.spacing) + geom.oz
    Z, Y, X = np.meshgrid(iz, iy, ix, indexing="ij")
    E = a + g[0] * X + g[1] * Y + g[2] * Z
    one = np.stack([np.full_like(E, -g[0]), np.full_like(E, -g[1]), np.full_like(E, -g[2]), E], axis=-1)
    data = np.broadcast_to(one, (T + 1,) + one.shape).astype(np.float32).ravel()
    return W.IGrid(geom, T + 1, np.ascontiguousarray(data))


def workload(W, rec, lig, params, dofs, inner=None, outer=None, plateau=10.0,
             neighbourdis=10.8, lig_inner=None, lig_outer=None):
    T = params.T
    gi = W.inner_geom(rec, 0.875, 10.0)
    go = W.cube_geom((0.0, 0.0, 0.0), 60.0, 2.6)
    inner = inner if inner is not None else zero_grid(W, gi, T)
    outer = outer if outer is not None else zero_grid(W, go, T)
    nl = W.build_nl(inner.geom, rec, neighbourdis, plateau)
    wl = W.Workload("hand-built", rec, lig, params, inner, outer, nl,
                    np.ascontiguousarray(dofs, np.float32))
    if lig_inner is not None:
        wl.lig_inner, wl.lig_outer = lig_inner, lig_outer
    return wl


def permute_alphabet(W, wl, perm):
    """The same physical workload with the receptor grids' probe types in the
    order of a grid alphabet: grid m (1-based) holds probe type perm[m-1]
    (grid 0, electrostatics, stays first) and the ligand's mapped types follow
    (mtype = position of its type in the alphabet + 1, as the reference's
    TypeMap does, model/TypeMap.h:52-60).  Scores, gradients and parameter
    gradients must not change."""
    import copy
    perm = np.asarray(perm, np.int64)
    T = wl.params.T
    assert sorted(perm.tolist()) == list(range(1, T + 1))
    pos = np.zeros(T + 1, np.uint32)
    pos[perm] = np.arange(1, T + 1, dtype=np.uint32)

    def regrid(g):
        d = g.data.reshape(g.ngrids, -1)
        slabs = np.concatenate([d[:1], d[perm]], axis=0)
        return W.IGrid(g.geom, g.ngrids, np.ascontiguousarray(slabs.ravel()))

    out = copy.copy(wl)
    lig = copy.copy(wl.lig)
    t = lig.type
    lig.mtype = np.where((t >= 1) & (t <= T), pos[np.minimum(t, T)], 0).astype(np.uint32)
    lig.mtype[wl.lig.mtype == 0] = 0
    out.lig = lig
    out.rec_inner, out.rec_outer = regrid(wl.rec_inner), regrid(wl.rec_outer)
    return out
