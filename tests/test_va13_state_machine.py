"""The device minimizer's VA13 state machine (csrc/hip/va13.h, compiled for the
host by tests/cpp/va13_host.cpp) against the oracle's restatement
(oracle/oracle_va13.c): on the same objectives the two ask for exactly the same
sequence of points (bit for bit) and report the same results.  Objectives:
quadratics, Rosenbrock, a quartic with coupled terms, and real docking poses
(the oracle's FP32 adjoint).  These reach every branch of the line search and
the update branches minfor's calls reach; the remaining branches of the factor
update (positive rank on entry, negative weight without a prior solve, a zero
diagonal to re-seed) are compared directly on random factors (CPU only)."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def host():
    so = os.path.join(ROOT, "tests", "cpp", "_bin", "libva13_host.so")
    if not os.path.exists(so):
        pytest.fail(f"{so} is not built: run __graft_entry__.build()")
    L = C.CDLL(so)
    L.FG = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                       C.POINTER(C.c_double))
    L.va13h_minimize.restype = C.c_int
    L.va13h_minimize.argtypes = [C.c_int, C.c_void_p, C.c_int, L.FG, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def run_host(L, fg, x0, ivmax):
    n = len(x0)
    evals = []

    def cb(_, xp, fp, gp):
        x = np.ctypeslib.as_array(xp, (n,)).copy()
        f, g = fg(x)
        evals.append((x, float(f)))
        fp[0] = float(f)
        for i in range(n):
            gp[i] = float(g[i])

    cf = L.FG(cb)
    x = np.ascontiguousarray(x0, np.float64)
    xl = np.zeros(n)
    fl, fb, nf = C.c_double(), C.c_double(), C.c_int()
    assert L.va13h_minimize(n, x.ctypes.data, ivmax, cf, None, xl.ctypes.data, C.byref(fl),
                            C.byref(fb), C.byref(nf)) == 0
    return xl, fl.value, fb.value, nf.value, evals


def _same(O, L, fg, x0, ivmax):
    a = O.va13(fg, x0, ivmax)
    b = run_host(L, fg, x0, ivmax)
    assert len(a[4]) == len(b[4])
    for (xa, fa), (xb, fb) in zip(a[4], b[4]):
        assert np.array_equal(xa, xb) and fa == fb
    assert np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2] and a[3] == b[3]
    return len(a[4])


@pytest.mark.parametrize("n,seed", [(1, 0), (6, 1), (16, 2), (26, 3)])
def test_quadratics(O, host, n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = Q @ np.diag(np.geomspace(1.0, 300.0, n)) @ Q.T
    xs = rng.normal(size=n)
    _same(O, host, lambda x: (0.5 * (x - xs) @ A @ (x - xs), A @ (x - xs)), xs + 1.0, 500)


def test_rosenbrock_and_quartic(O, host):
    def rosen(x):
        f = np.sum(100.0 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
        g = np.zeros_like(x)
        g[:-1] = -400.0 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2.0 * (1 - x[:-1])
        g[1:] += 200.0 * (x[1:] - x[:-1] ** 2)
        return f, g

    def quartic(x):
        s = np.sum(x)
        return np.sum(x ** 4) + s ** 2 - 3.0 * x[0], 4 * x ** 3 + 2 * s - 3.0 * (np.arange(len(x)) == 0)

    assert _same(O, host, rosen, np.array([-1.2, 1.0, -0.5, 0.8, 1.3, 0.2]), 3000) > 40
    _same(O, host, quartic, np.linspace(-2, 2, 11), 500)
    _same(O, host, rosen, np.array([-1.2, 1.0, -0.5, 0.8, 1.3, 0.2]), 9)  # evaluation limit


def test_docking_objective(W, O, host):
    """The real objective: poses of a flexible pair scored by the FP32 adjoint."""
    wl = W.make_workload("va13 poses", 73, 300, 80, 6, 3, 2)
    Mr, Ml = wl.rec.M, wl.lig.M

    def fg(s):
        d = np.zeros((1, 26), np.float32)
        d[0, 3:6], d[0, 0:3] = s[0:3], s[3:6]
        d[0, 6:6 + Mr], d[0, 16:16 + Ml] = s[6:6 + Mr], s[6 + Mr:]
        r = O.adjoint(wl, d, f32=True)[0]
        g = np.concatenate([r[4:7], r[1:4], r[7:7 + Mr], r[17:17 + Ml]])
        return r[0], g

    total = 0
    for p in range(wl.dofs.shape[0]):
        d = wl.dofs[p].astype(np.float64)
        s0 = np.concatenate([d[3:6], d[0:3], d[6:6 + Mr], d[16:16 + Ml]])
        total += _same(O, host, fg, s0, 60)
    assert total > 100


def test_factor_update_all_branches(O, host):
    """mc11a's update on random packed factors, every combination of weight
    sign, rank on entry, prior solve (mk), aliasing, zero diagonals and eps."""
    L = O.oracle_lib()
    L.orc_va13_update.restype = None
    L.orc_va13_update.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_void_p,
                                  C.POINTER(C.c_int), C.c_int, C.c_double]
    host.va13h_update.restype = None
    host.va13h_update.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_void_p,
                                  C.POINTER(C.c_int), C.c_int, C.c_double, C.c_int]
    rng = np.random.default_rng(11)
    cases = 0
    for n in (1, 2, 3, 6, 11):
        m = n * (n + 1) // 2
        for trial in range(60):
            a = rng.normal(size=m)
            diag = np.cumsum([0] + [n - i for i in range(n - 1)])
            a[diag] = np.abs(a[diag]) + 0.1
            if trial % 3 == 0:
                a[diag[rng.integers(n)]] = 0.0
            z = rng.normal(size=n)
            if trial % 5 == 0:
                z[rng.integers(n)] = 0.0
            w = rng.normal(size=n)
            sig = rng.choice([-1.0, 1.0]) * rng.uniform(0.1, 3.0)
            ir0 = int(rng.choice([-n, -1, 0, 1, n]))
            mk = int(rng.integers(2))
            eps = float(rng.choice([0.0, 1e-3]))
            alias = int(rng.integers(2))
            a1, z1, w1 = a.copy(), z.copy(), (z.copy() if alias else w.copy())
            a2, z2, w2 = a.copy(), z.copy(), w.copy()
            ir1, ir2 = C.c_int(ir0), C.c_int(ir0)
            if alias:
                L.orc_va13_update(a1.ctypes.data, n, z1.ctypes.data, sig, z1.ctypes.data,
                                  C.byref(ir1), mk, eps)
                w1 = z1
                w2 = z.copy()
            else:
                L.orc_va13_update(a1.ctypes.data, n, z1.ctypes.data, sig, w1.ctypes.data,
                                  C.byref(ir1), mk, eps)
            host.va13h_update(a2.ctypes.data, n, z2.ctypes.data, sig, w2.ctypes.data, C.byref(ir2),
                              mk, eps, alias)
            assert np.array_equal(a1, a2) and np.array_equal(z1, z2) and ir1.value == ir2.value
            if not alias:
                assert np.array_equal(w1, w2)
            cases += 1
    assert cases == 300
