"""The drop-in path on the GPU: requests through as::Server with the MI355
scoring services (C++, tests/cpp/service_driver.cpp) return exactly what the
direct C-ABI calls return on the same workload, for several worker counts,
item sizes and numbers of requests in flight (needs a GPU)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def drv():
    L = C.CDLL(os.path.join(ROOT, "tests", "cpp", "_bin", "libservice_driver.so"))
    L.drv_error.restype = C.c_char_p
    return L


@pytest.fixture(scope="module")
def ctx():
    from gpuattract_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def _common_args(wl):
    from gpuattract_amd import hip
    p = wl.params
    tab = hip.param_table(p)
    keep = [tab]
    return keep, (hip._p(tab), p.T, p.shape, p.dielec, C.c_float(p.epsilon), C.c_float(p.ffelec))


def _run_rigid(drv, wl, workers, item, requests):
    from gpuattract_amd import hip
    rec, lig = hip.protein_desc(wl.rec), hip.protein_desc(wl.lig)
    gi, go, nl = hip.grid_desc(wl.rec_inner), hip.grid_desc(wl.rec_outer), hip.nl_desc(wl.rec_nl)
    keep, tail = _common_args(wl)
    d = np.ascontiguousarray(wl.dofs, np.float32)
    out = np.full((d.shape[0], 7), np.nan, np.float32)
    st = drv.drv_score_rigid(C.byref(rec), C.byref(lig), C.byref(gi), C.byref(go), C.byref(nl),
                             *tail, hip._p(d), C.c_size_t(d.shape[0]), hip._p(out),
                             workers, item, requests)
    assert st == 0, drv.drv_error().decode()
    return out


def _run_modes(drv, wl, workers, item, requests):
    from gpuattract_amd import hip
    rec, lig = hip.protein_desc(wl.rec), hip.protein_desc(wl.lig)
    gi, go, nl = hip.grid_desc(wl.rec_inner), hip.grid_desc(wl.rec_outer), hip.nl_desc(wl.rec_nl)
    li, lo = hip.grid_desc(wl.lig_inner), hip.grid_desc(wl.lig_outer)
    keep, tail = _common_args(wl)
    d = np.ascontiguousarray(wl.dofs, np.float32)
    out = np.full((d.shape[0], 27), np.nan, np.float32)
    st = drv.drv_score_modes(C.byref(rec), C.byref(lig), C.byref(gi), C.byref(go), C.byref(nl),
                             C.byref(li), C.byref(lo), *tail, hip._p(d), C.c_size_t(d.shape[0]),
                             hip._p(out), workers, item, requests)
    assert st == 0, drv.drv_error().decode()
    return out


@pytest.mark.parametrize("workers,item,requests", [(1, 100, 1), (1, 7, 3), (2, 33, 4), (3, 1000, 2)])
def test_server_rigid_equals_direct(drv, ctx, wl_rigid, workers, item, requests):
    hs = ctx.upload_workload(wl_rigid)
    direct = ctx.score_rigid(hs, wl_rigid.dofs)
    got = _run_rigid(drv, wl_rigid, workers, item, requests)
    assert np.array_equal(got, direct, equal_nan=True)


@pytest.mark.parametrize("workers,item,requests", [(1, 100, 1), (2, 17, 3)])
def test_server_modes_equals_direct(drv, ctx, wl_modes, workers, item, requests):
    hs = ctx.upload_workload(wl_modes)
    direct = ctx.score_modes(hs, wl_modes.dofs)
    got = _run_modes(drv, wl_modes, workers, item, requests)
    assert np.array_equal(got, direct, equal_nan=True)
