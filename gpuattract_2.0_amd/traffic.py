"""Algorithmic bytes per pose of the scoring path (the roofline numerator).

Per bead touched by a pose:
    20                      bead attributes (x, y, z, type, charge)
  + 128 * (g + e)           8 float4 grid corners for the bead's type (g) and,
                            for charged beads, for the electrostatic grid (e),
                            when the bead lies in the inner or outer grid
  + 8 * n + 4 * K           neighbour-list descriptor when the bead is in the
                            neighbour-list box, plus 4 bytes per listed neighbour
  + 12 * M                  mode vectors of the bead (flexible scoring)
  + 64 * T                  the adjoint's parameter gradient: 2T basis values
                            (S_a, S6_a) at the 8 corners, for beads in a grid
Per pose: 24 bytes DOF in, 28 bytes result out, plus 4*(M_rec+M_lig) each way
for mode amplitudes and gradients.

The counts g, e, n, K come from the poses' geometry (double precision, a
seeded sample of poses); a bead within rounding distance of a box face may be
classified differently than in FP32, which does not move the average.
"""
import numpy as np


def _rotmats(ang):
    phi, ssi, rot = ang[:, 0], ang[:, 1], ang[:, 2]
    cS, sS, cP, sP, cR, sR = np.cos(ssi), np.sin(ssi), np.cos(phi), np.sin(phi), np.cos(rot), np.sin(rot)
    R = np.empty((len(ang), 3, 3))
    R[:, 0, 0] = cR * cS * cP + sR * sP
    R[:, 0, 1] = sR * cS * cP - cR * sP
    R[:, 0, 2] = sS * cP
    R[:, 1, 0] = cR * cS * sP - sR * cP
    R[:, 1, 1] = sR * cS * sP + cR * cP
    R[:, 1, 2] = sS * sP
    R[:, 2, 0] = -cR * sS
    R[:, 2, 1] = -sR * sS
    R[:, 2, 2] = cS
    return R


def _inside(g, p):
    ge = g.geom
    lo = np.array([ge.ox, ge.oy, ge.oz])
    hi = lo + (np.array([ge.nx, ge.ny, ge.nz]) - 1) * ge.spacing
    return np.all((p >= lo) & (p <= hi), axis=-1)


def _deformed(body, amps):
    """(P, n, 3) deformed coordinates for amplitudes (P, M)."""
    n, M = body.n, body.M
    x = body.xyz.reshape(3, n).T.astype(np.float64)
    if M == 0:
        return np.broadcast_to(x, (amps.shape[0], n, 3))
    modes = body.modes.reshape(3, n, M).transpose(1, 2, 0).astype(np.float64)  # (n, M, 3)
    return x[None] + np.einsum("pm,nmc->pnc", amps.astype(np.float64), modes)


def _grid_bytes(inner, outer, p, mtype, charge, basis_T=0):
    ing = _inside(inner, p)
    inany = ing | _inside(outer, p)
    g = inany & (mtype[None, :] != 0)
    e = g & (np.abs(charge)[None, :] > 0.001)
    return 128.0 * (g.sum(axis=1) + e.sum(axis=1)) + 64.0 * basis_T * g.sum(axis=1)


def _nl_bytes(nl, p, typ):
    ge = nl.geom
    lo = np.array([ge.ox, ge.oy, ge.oz])
    inb = _inside(nl, p) & (typ[None, :] != 0)
    idx = np.rint((p - lo) / ge.spacing).astype(np.int64)
    idx = np.clip(idx, 0, np.array([ge.nx, ge.ny, ge.nz]) - 1)
    vox = (idx[..., 2] * ge.ny + idx[..., 1]) * ge.nx + idx[..., 0]
    K = nl.desc.reshape(-1, 2)[vox, 0].astype(np.float64)
    return 8.0 * inb.sum(axis=1) + 4.0 * (K * inb).sum(axis=1)


def bytes_per_pose(wl, sample=256, seed=0, adjoint=False, pgrad=False):
    """Mean algorithmic bytes per pose over a sample of the workload's poses.
    adjoint=True counts the adjoint's reads: the receptor beads are deformed
    for the neighbour list but not interpolated on the ligand grid (that term
    is not part of the energy); pgrad=True adds the basis-grid reads of the
    pair-parameter gradient."""
    rng = np.random.default_rng(seed)
    n = wl.dofs.shape[0]
    pick = rng.choice(n, size=min(sample, n), replace=False)
    d = wl.dofs[pick].astype(np.float64)
    R = _rotmats(d[:, 3:6])
    t = d[:, :3]
    lig, rec = wl.lig, wl.rec
    if wl.modes:
        ampR, ampL = d[:, 6:6 + rec.M], d[:, 16:16 + lig.M]
    else:
        ampR = np.zeros((len(d), 0))
        ampL = np.zeros((len(d), 0))
    xl = _deformed(lig, ampL)
    pl = np.einsum("pij,pnj->pni", R, xl) + t[:, None, :]
    B = 20.0 * lig.n + 12.0 * lig.n * lig.M
    B = B + _grid_bytes(wl.rec_inner, wl.rec_outer, pl, lig.mtype, lig.charge,
                        wl.params.T if pgrad else 0)
    B = B + _nl_bytes(wl.rec_nl, pl, lig.type)
    per_pose_io = 24.0 + 28.0
    if wl.modes:
        xr = _deformed(rec, ampR)
        Rinv = _rotmats(-d[:, 3:6])
        pinv = -np.einsum("pji,pj->pi", R, t)  # R^T (-t)
        pr = np.einsum("pij,pnj->pni", Rinv, xr) + pinv[:, None, :]
        B = B + 20.0 * rec.n + 12.0 * rec.n * rec.M
        if not adjoint:
            B = B + _grid_bytes(wl.lig_inner, wl.lig_outer, pr, rec.mtype, rec.charge)
        per_pose_io += 2 * 4.0 * (rec.M + lig.M)
    B = B + per_pose_io
    return float(np.mean(B)), float(np.std(B))
