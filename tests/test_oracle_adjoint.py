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


def _pipeline_energy(W, O, wl):
    """Sum of the FP64 energies with the grids rebuilt from wl.params."""
    def energy(p):
        wl2 = W.Workload(wl.name, wl.rec, wl.lig, p,
                         W.build_grid(wl.rec_inner.geom, wl.rec, p, False, 10.0),
                         W.build_grid(wl.rec_outer.geom, wl.rec, p, False, 10.0),
                         wl.rec_nl, wl.dofs)
        return O.score_rigid(wl2, f64=True)[0][:, 0].sum()
    return energy


@pytest.fixture(scope="module")
def wl_pgrad(W):
    """A pose set where the grid terms, the neighbour-list terms and the
    repulsive branch (ipon = -1 inside rmin, the only pairs with a dE/demin)
    all contribute."""
    return W.make_workload("pgrad", 41, 120, 30, 16, with_basis=True, plateau=10.0,
                           repulsive_frac=0.5)


def test_parameter_gradient_finite_differences(W, O, wl_pgrad):
    """dE/d(rc, ac, emin) against differences of the full pipeline, the grids
    rebuilt from the perturbed table."""
    wl = wl_pgrad
    _, pg = O.adjoint(wl, pgrad=True)
    T = wl.params.T
    rng = np.random.default_rng(0)
    energy = _pipeline_energy(W, O, wl)
    import copy
    for j, name in enumerate(("rc", "ac", "emin")):
        # the table entries with the largest gradient, plus random ones
        order = np.argsort(-np.abs(pg[:, j]))
        picks = list(order[:4]) + list(rng.choice(T * T, 3, replace=False))
        assert np.abs(pg[order[0], j]) > 0, name
        for k in picks:
            base = getattr(wl.params, name)
            h = 1e-3 * abs(float(base[k])) + 1e-6
            pp, pm = copy.deepcopy(wl.params), copy.deepcopy(wl.params)
            getattr(pp, name)[k] = base[k] + h
            getattr(pm, name)[k] = base[k] - h
            hh = float(getattr(pp, name)[k]) - float(getattr(pm, name)[k])
            fd = (energy(pp) - energy(pm)) / hh
            assert abs(fd - pg[k, j]) <= 2e-3 * (abs(fd) + abs(pg[k, j])) + 1e-9, (k, name, fd, pg[k, j])


def _rederive(W, params, k, rmin=None, emin=None):
    """The table re-derived (the product's restatement of readFile.cpp:479-513)
    from raw attract.par values whose entry k has the given (rmin, emin)."""
    import copy
    p = copy.deepcopy(params)
    r = np.sqrt(np.float64(p.rmin2[k])) if rmin is None else rmin
    e = np.float64(p.emin[k]) if emin is None else emin
    if p.shape == 8:
        p.rbc[k] = np.float32(r * np.sqrt(0.75))
        p.abc[k] = np.float32(-256.0 / 27.0 * e)
    else:
        p.rbc[k] = np.float32(r / np.sqrt(1.25992105))
        p.abc[k] = np.float32(-4.0 * e)
    L = W.gen_lib()
    L.attg_derive_params(p.T, p.shape, W._p(p.rbc), W._p(p.abc), W._p(p.ipon_raw), W._p(p.rc),
                         W._p(p.ac), W._p(p.ipon), W._p(p.rmin2), W._p(p.emin))
    return p


@pytest.mark.parametrize("shape", [8, 12])
def test_rmin_emin_gradient_finite_differences(W, O, shape):
    """att_param_grad_rmin_emin (the chain rule through the table derivation)
    on the oracle's parameter gradient, against central differences of the
    pipeline in (rmin, emin): the raw parameters of one entry moved, the table
    re-derived and the grids rebuilt."""
    from gpuattract_amd import hip
    wl = W.make_workload(f"pgrad {shape}", 42, 120, 30, 16, with_basis=True, plateau=10.0,
                         repulsive_frac=0.5, shape=shape)
    _, pg = O.adjoint(wl, pgrad=True)
    g = hip.param_grad_rmin_emin(wl.params, pg)
    energy = _pipeline_energy(W, O, wl)
    rng = np.random.default_rng(1)
    for j, name in enumerate(("rmin", "emin")):
        order = np.argsort(-np.abs(g[:, j]))
        for k in list(order[:3]) + list(rng.choice(wl.params.T ** 2, 2, replace=False)):
            r0 = float(np.sqrt(np.float64(wl.params.rmin2[k])))
            e0 = float(wl.params.emin[k])
            h = 1e-3 * (r0 if j == 0 else abs(e0))
            kw = (lambda v: dict(rmin=v)) if j == 0 else (lambda v: dict(emin=v))
            x0 = r0 if j == 0 else e0
            pp = _rederive(W, wl.params, k, **kw(x0 + h))
            pm = _rederive(W, wl.params, k, **kw(x0 - h))
            fd = (energy(pp) - energy(pm)) / (2 * h)
            assert abs(fd - g[k, j]) <= 3e-3 * (abs(fd) + abs(g[k, j])) + 1e-6, (shape, k, name, fd, g[k, j])


def test_parameter_gradient_follows_the_bead_type_not_the_grid_slot(W, O, wl_pgrad):
    """With a grid alphabet that orders the probe types differently, the grid
    slot of a bead (its mapped type) differs from its type; the energy, the
    DOF gradient and the parameter gradient (whose grid part reads the table
    column of the bead's type) stay the same bits."""
    from helpers import permute_alphabet
    wl = wl_pgrad
    T = wl.params.T
    perm = np.roll(np.arange(1, T + 1), 5)
    wp = permute_alphabet(W, wl, perm)
    assert np.mean(wp.lig.mtype != wl.lig.type) > 0.5
    a0, p0 = O.adjoint(wl, pgrad=True)
    a1, p1 = O.adjoint(wp, pgrad=True)
    assert np.array_equal(a0, a1) and np.array_equal(p0, p1)
    assert np.array_equal(O.score_rigid(wl), O.score_rigid(wp))


@pytest.mark.parametrize("modes", [False, True])
def test_kernel_order_adjoint(W, O, modes):
    """The FP32 adjoint in the HIP kernel's summation order: its energy is the
    reference's energy bit for bit (list-order folds per bead, bead-order sum:
    the forward restatement, which equals the compiled reference on the golden
    fixtures); its gradient is as close to FP64 as the sequential FP32
    restatement's; only the receptor mode sums depend on the waves per pose."""
    wl = W.make_workload("kadj", 71, 600, 450, 24, *((4, 3) if modes else ()), with_basis=True)
    ref64, sc = O.adjoint(wl, scale=True)
    r32 = O.adjoint(wl, f32=True)
    fwd = (O.score_modes if modes else O.score_rigid)(wl)
    ks = {wv: O.adjoint_kernel_order(wl, waves=wv) for wv in (1, 2, 4)}
    for wv, k in ks.items():
        assert np.array_equal(k[:, 0].astype(np.float32), fwd[:, 0])
        cols = [c for c in range(27) if sc[:, c].max() > 0]
        e_k = np.abs(k[:, cols] - ref64[:, cols]) / np.maximum(sc[:, cols], 1e-30)
        e_r = np.abs(r32[:, cols] - ref64[:, cols]) / np.maximum(sc[:, cols], 1e-30)
        assert e_k.max() <= 1.1 * e_r.max() + 1e-6
        keep = [c for c in range(27) if not 7 <= c < 17]
        assert np.array_equal(k[:, keep], ks[1][:, keep])
