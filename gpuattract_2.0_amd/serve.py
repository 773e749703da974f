"""ctypes binding of the docking service (include/attract_serve.h, in
lib/libattract_service.so).

A Server owns one engine thread and device context.  Receptors, ligands and
tables are registered once (host copies); jobs are sets of starting poses
minimized on the device.  Loading fails loudly when the library is missing.
"""
import ctypes as C
import os

import numpy as np

from . import hip

LIBPATH = os.path.join(hip.PKG, "lib", "libattract_service.so")

P = C.c_void_p
H = C.c_uint64
ATT_NOT_READY = 6

_lib = None


class JobInfo(C.Structure):
    _fields_ = [("queued_ms", C.c_double), ("run_ms", C.c_double), ("latency_ms", C.c_double),
                ("grid_cached", C.c_int), ("rounds", C.c_uint), ("bytes", C.c_size_t),
                ("admit_seq", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("jobs_done", C.c_uint64), ("poses_done", C.c_uint64),
                ("evaluations", C.c_uint64), ("grid_hits", C.c_uint64),
                ("grid_misses", C.c_uint64), ("evictions", C.c_uint64),
                ("budget_bytes", C.c_size_t), ("peak_bytes", C.c_size_t),
                ("max_running", C.c_uint), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("busy_s", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} missing: build it first (__graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        L.att_server_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(P)]
        L.att_server_destroy.argtypes = [P]
        L.att_server_destroy.restype = None
        L.att_server_last_error.argtypes = [P]
        L.att_server_last_error.restype = C.c_char_p
        L.att_server_add_receptor.argtypes = [P, C.POINTER(hip.ProteinDesc),
                                              C.POINTER(hip.GridDesc), C.POINTER(hip.GridDesc),
                                              C.POINTER(hip.NLDesc), C.POINTER(H)]
        L.att_server_add_receptor_body.argtypes = [P, C.POINTER(hip.ProteinDesc), H,
                                                   C.POINTER(hip.GridGeom),
                                                   C.POINTER(hip.GridGeom), C.c_float, C.c_float,
                                                   C.POINTER(H)]
        L.att_server_add_ligand.argtypes = [P, C.POINTER(hip.ProteinDesc), C.POINTER(H)]
        L.att_server_add_params.argtypes = [P, P, C.c_uint, C.c_int, C.c_int, C.c_float,
                                            C.c_float, C.POINTER(H)]
        L.att_server_submit.argtypes = [P, H, H, H, P, C.c_int, C.c_size_t, C.c_int,
                                        C.POINTER(H)]
        L.att_server_wait.argtypes = [P, H, C.c_int, P, P, P, P, C.POINTER(JobInfo)]
        L.att_server_get_stats.argtypes = [P, C.POINTER(Stats)]
        _lib = L
    return _lib


class Server:
    """att_server: continuous-batching minimization service on one device."""

    def __init__(self, device=0, hbm_budget=0):
        self.L = lib()
        h = P()
        st = self.L.att_server_create(device, hbm_budget, C.byref(h))
        if st != hip.ATT_OK:
            raise hip.AttError(st, self.L.att_server_last_error(None).decode())
        self.h = h
        self._pending = {}  # job id -> (n, dof width)

    def close(self):
        if self.h:
            self.L.att_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, st):
        if st != hip.ATT_OK:
            raise hip.AttError(st, self.L.att_server_last_error(self.h).decode())

    def add_receptor(self, body, inner, outer, nl):
        d = hip.protein_desc(body)
        gi, go, nd = hip.grid_desc(inner), hip.grid_desc(outer), hip.nl_desc(nl)
        out = H()
        self._chk(self.L.att_server_add_receptor(self.h, C.byref(d), C.byref(gi), C.byref(go),
                                                 C.byref(nd), C.byref(out)))
        return out.value

    def add_receptor_body(self, body, params_id, inner, outer, plateau=10.0, neighbourdis=10.8):
        """A receptor whose grids the server builds on the device at first use."""
        d = hip.protein_desc(body)
        gi, go = hip.grid_geom(inner), hip.grid_geom(outer)
        out = H()
        self._chk(self.L.att_server_add_receptor_body(self.h, C.byref(d), params_id, C.byref(gi),
                                                      C.byref(go), plateau, neighbourdis,
                                                      C.byref(out)))
        return out.value

    def add_ligand(self, body):
        d = hip.protein_desc(body)
        out = H()
        self._chk(self.L.att_server_add_ligand(self.h, C.byref(d), C.byref(out)))
        return out.value

    def add_params(self, p):
        tab = hip.param_table(p)
        out = H()
        self._chk(self.L.att_server_add_params(self.h, hip._p(tab), p.T, p.shape, p.dielec,
                                               p.epsilon, p.ffelec, C.byref(out)))
        return out.value

    def add_workload(self, wl, device_grids=False):
        """Registers a workload's receptor (with its host grids, or as a body
        whose grids are built on the device), ligand and table."""
        par = self.add_params(wl.params)
        if device_grids:
            rec = self.add_receptor_body(wl.rec, par, wl.rec_inner.geom, wl.rec_outer.geom,
                                         wl.rec_nl.plateau, wl.meta["neighbourdis"])
        else:
            rec = self.add_receptor(wl.rec, wl.rec_inner, wl.rec_outer, wl.rec_nl)
        return dict(rec=rec, lig=self.add_ligand(wl.lig), params=par)

    def submit(self, ids, dofs, max_fun_eval=500):
        d = np.ascontiguousarray(dofs, np.float32)
        width = d.shape[1]
        if width not in (6, 26):
            raise ValueError("DOF records are 6 (rigid) or 26 (modes) floats")
        out = H()
        self._chk(self.L.att_server_submit(self.h, ids["rec"], ids["lig"], ids["params"],
                                           hip._p(d), 1 if width == 26 else 0, d.shape[0],
                                           max_fun_eval, C.byref(out)))
        self._pending[out.value] = (d.shape[0], width)
        return out.value

    def wait(self, job, timeout_ms=-1):
        """(dofs_out, e_last, e_best, nfun, info) of a job; None on time-out."""
        n, width = self._pending[job]
        dofs = np.zeros((n, width), np.float32)
        el, eb = np.zeros(n, np.float32), np.zeros(n, np.float32)
        nf = np.zeros(n, np.int32)
        info = JobInfo()
        st = self.L.att_server_wait(self.h, job, timeout_ms, hip._p(dofs), hip._p(el),
                                    hip._p(eb), hip._p(nf), C.byref(info))
        if st == ATT_NOT_READY:
            return None
        del self._pending[job]
        self._chk(st)
        return dofs, el, eb, nf, {k: getattr(info, k) for k, _ in JobInfo._fields_}

    def stats(self):
        s = Stats()
        self._chk(self.L.att_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}
