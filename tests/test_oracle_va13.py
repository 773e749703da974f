"""The oracle's VA13 restatement (apps/em/minfor.f, mc11a.f), pinned by what
the mathematics fixes: it finds the exact minimum of strictly convex
quadratics (to the rounding of the data) within a bounded number of
evaluations, Rosenbrock's minimum, respects the evaluation limit, reports the
last evaluated point, and lowers docking energies (CPU only)."""
import numpy as np
import pytest


def _quadratic(n, cond, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = Q @ np.diag(np.geomspace(1.0, cond, n)) @ Q.T
    xs = rng.normal(size=n)
    return A, xs, (lambda x: (0.5 * (x - xs) @ A @ (x - xs), A @ (x - xs)))


@pytest.mark.parametrize("n,cond", [(6, 10.0), (16, 100.0), (26, 1000.0)])
def test_quadratic_minimum(O, n, cond):
    A, xs, fg = _quadratic(n, cond, n)
    x0 = xs + np.random.default_rng(1).normal(size=n)
    xl, fl, fb, nf, ev = O.va13(fg, x0, ivmax=2000)
    best = min(ev, key=lambda e: e[1])[0]
    assert fb == pytest.approx(min(e[1] for e in ev))
    assert np.linalg.norm(best - xs) < 1e-4 * (1 + np.linalg.norm(xs))
    assert fb < 1e-9
    assert nf < 40 * n  # quasi-Newton, not gradient descent
    # the reported point is the last one evaluated
    assert np.array_equal(xl, ev[-1][0]) and fl == ev[-1][1]


def test_rosenbrock(O):
    def fg(x):
        f = np.sum(100.0 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
        g = np.zeros_like(x)
        g[:-1] = -400.0 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2.0 * (1 - x[:-1])
        g[1:] += 200.0 * (x[1:] - x[:-1] ** 2)
        return f, g
    x0 = np.array([-1.2, 1.0, -0.5, 0.8, 1.3, 0.2])
    xl, fl, fb, nf, ev = O.va13(fg, x0, ivmax=3000)
    best = min(ev, key=lambda e: e[1])[0]
    assert fb < 1e-8
    assert np.allclose(best, 1.0, atol=1e-3)


def test_evaluation_limit(O):
    A, xs, fg = _quadratic(6, 1e4, 3)
    xl, fl, fb, nf, ev = O.va13(fg, xs + 1.0, ivmax=7)
    assert nf == 7 and len(ev) == 8  # the initial evaluation plus ivmax line-search steps


def test_docking_energies_decrease(W, O):
    wl = W.make_workload("min", 71, 300, 80, 24, 3, 2)
    e0 = O.adjoint(wl, f32=True)[:, 0]
    d, el, eb, nf = O.minimize(wl, ivmax=120)
    assert np.all(eb <= e0 + 1e-6 * np.abs(e0))
    assert np.mean(eb < e0 - 1e-3) > 0.8
    assert np.all(nf <= 120)
    # the reported DOFs reproduce the reported energy
    e_chk = O.adjoint(wl, d, f32=True)[:, 0]
    assert np.array_equal(e_chk, el)


def _unpack(a, n):
    """H = L D L^T from mc11a's packed factors (per column i: D_i, then
    L_{i+1..n-1, i})."""
    L = np.eye(n)
    D = np.zeros(n)
    k = 0
    for i in range(n):
        D[i] = a[k]
        L[i + 1:, i] = a[k + 1:k + n - i]
        k += n - i
    return L @ np.diag(D) @ L.T


@pytest.mark.parametrize("n", [1, 2, 5, 12, 26])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_factor_update_is_rank_one_update(O, n, sign):
    """mc11a's update of positive definite factors (full rank, no prior solve)
    yields the factors of H + sig z z^T, checked by dense recomputation."""
    import ctypes as C
    L = O.oracle_lib()
    L.orc_va13_update.restype = None
    L.orc_va13_update.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_void_p,
                                  C.POINTER(C.c_int), C.c_int, C.c_double]
    rng = np.random.default_rng(n)
    for trial in range(20):
        a = np.zeros(n * (n + 1) // 2)
        k = 0
        for i in range(n):
            a[k] = rng.uniform(0.5, 2.0)
            a[k + 1:k + n - i] = rng.normal(scale=0.5, size=n - 1 - i)
            k += n - i
        H = _unpack(a, n)
        z = rng.normal(size=n)
        # a negative weight small enough to keep the result positive definite
        lam = np.linalg.eigvalsh(H)[0]
        sig = rng.uniform(0.1, 2.0) if sign > 0 else -0.5 * lam / (z @ z)
        want = H + sig * np.outer(z, z)
        w = np.zeros(n)
        ir = C.c_int(n)
        zz = z.copy()
        L.orc_va13_update(a.ctypes.data, n, zz.ctypes.data, sig, w.ctypes.data, C.byref(ir), 0, 0.0)
        assert ir.value == n
        assert np.allclose(_unpack(a, n), want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
