"""The oracle's exact adjoint, pinned by central finite differences (CPU only).

The oracle's FP64 energy is differentiated numerically with respect to the
pose DOFs, the mode amplitudes and the pair parameters; for the parameters the
whole pipeline is re-run (grids rebuilt from the perturbed table), so the
basis-grid route of the adjoint is checked against the grids it stands for.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def wl_adj(W):
    return W.make_workload("adjoint", 31, 300, 80, 40, 3, 4, with_basis=True)


def _fd(O, wl, col, h=1e-3):
    dp, dm = wl.dofs.copy(), wl.dofs.copy()
    dp[:, col] += h
    dm[:, col] -= h
    delta = dp[:, col].astype(np.float64) - dm[:, col].astype(np.float64)
    return (O.adjoint(wl, dp)[:, 0] - O.adjoint(wl, dm)[:, 0]) / delta


def test_energy_matches_scoring(O, wl_adj):
    res = O.adjoint(wl_adj)
    e64 = O.score_modes(wl_adj, f64=True)[0][:, 0]
    assert np.allclose(res[:, 0], e64, rtol=1e-12, atol=1e-9 * np.abs(e64).max())


@pytest.mark.parametrize("col,out,tol", [
    (0, 1, 1e-4), (1, 2, 1e-4), (2, 3, 1e-4),          # translation
    (3, 4, 3e-3), (4, 5, 3e-3), (5, 6, 3e-3),          # Euler angles
    (6, 7, 1e-5), (8, 9, 1e-5),                        # receptor modes
    (16, 17, 1e-5), (19, 20, 1e-5),                    # ligand modes
])
def test_dof_gradient_finite_differences(O, wl_adj, col, out, tol):
    an = O.adjoint(wl_adj)[:, out]
    fd = _fd(O, wl_adj, col)
    rel = np.abs(fd - an) / (np.abs(an) + np.abs(fd) + 1e-3)
    # the energy is only piecewise smooth (trilinear cells, plateau cut-off):
    # the median pins the derivative, a few poses may straddle a kink
    assert np.median(rel) < tol
    assert np.mean(rel < 10 * tol) > 0.85


def test_upstream_weights(O, wl_adj):
    w = np.linspace(-2.0, 3.0, wl_adj.dofs.shape[0])
    a1, p1 = O.adjoint(wl_adj, pgrad=True)
    aw, pw = O.adjoint(wl_adj, w=w, pgrad=True)
    assert np.allclose(aw[:, 0], a1[:, 0])
    assert np.allclose(aw[:, 1:], a1[:, 1:] * w[:, None])
    # per-pose parameter gradients: rows of single-pose calls
    rows = [O.adjoint(wl_adj, wl_adj.dofs[k:k + 1], pgrad=True)[1] for k in range(len(w))]
    assert np.allclose(pw, sum(wk * r for wk, r in zip(w, rows)), rtol=1e-9, atol=1e-12)


def test_parameter_gradient_finite_differences(W, O):
    """dE/d(rc, ac, emin) against differences of the full pipeline, the grids
    rebuilt from the perturbed table (a pose set where both the grid and the
    neighbour-list terms contribute)."""
    wl = W.make_workload("pgrad", 41, 120, 30, 16, with_basis=True, plateau=10.0)
    _, pg = O.adjoint(wl, pgrad=True)
    T = wl.params.T
    rng = np.random.default_rng(0)
    # the table entries with the largest gradient, plus random ones
    order = np.argsort(-np.abs(pg[:, 0]))
    picks = list(order[:4]) + list(rng.choice(T * T, 4, replace=False))

    def energy(p):
        wl2 = W.Workload(wl.name, wl.rec, wl.lig, p,
                         W.build_grid(wl.rec_inner.geom, wl.rec, p, False, 10.0),
                         W.build_grid(wl.rec_outer.geom, wl.rec, p, False, 10.0),
                         wl.rec_nl, wl.dofs)
        return O.score_rigid(wl2, f64=True)[0][:, 0].sum()

    import copy
    for k in picks:
        for j, name in enumerate(("rc", "ac")):
            base = getattr(wl.params, name)
            h = 1e-3 * abs(float(base[k])) + 1e-6
            pp, pm = copy.deepcopy(wl.params), copy.deepcopy(wl.params)
            getattr(pp, name)[k] = base[k] + h
            getattr(pm, name)[k] = base[k] - h
            hh = float(getattr(pp, name)[k]) - float(getattr(pm, name)[k])
            fd = (energy(pp) - energy(pm)) / hh
            assert abs(fd - pg[k, j]) <= 2e-3 * (abs(fd) + abs(pg[k, j])) + 1e-9, (k, name, fd, pg[k, j])
