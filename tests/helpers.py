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
