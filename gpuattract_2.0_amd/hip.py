"""ctypes binding of libattract_hip.so (include/attract_hip.h).

Host-buffer calls take numpy arrays; the *_dev calls take device pointers
from the library's own allocator (Context.dev_alloc / to_device) and a
hipStream_t (None = the context's stream), so the process needs no other GPU
runtime.  Loading fails loudly when the library is missing: there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "lib", "libattract_hip.so")

P = C.c_void_p
H = C.c_uint64

ATT_OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "HIP error", 3: "out of memory", 4: "bad handle",
          5: "unsupported"}


class ProteinDesc(C.Structure):
    _fields_ = [("n", C.c_uint), ("numModes", C.c_uint), ("xyz", P), ("type", P),
                ("mappedType", P), ("charge", P), ("modes", P), ("modeForce", P)]


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_uint), ("ny", C.c_uint), ("nz", C.c_uint), ("numGrids", C.c_uint),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("spacing", C.c_float),
                ("data", P)]


class NLDesc(C.Structure):
    _fields_ = [("nx", C.c_uint), ("ny", C.c_uint), ("nz", C.c_uint),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("spacing", C.c_float),
                ("plateau", C.c_float), ("desc", P), ("list", P), ("numEl", C.c_size_t)]


class GridGeom(C.Structure):
    _fields_ = [("nx", C.c_uint), ("ny", C.c_uint), ("nz", C.c_uint),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("spacing", C.c_float)]


def grid_geom(ge):
    """att_grid_geom of a workload Geom."""
    return GridGeom(ge.nx, ge.ny, ge.nz, ge.ox, ge.oy, ge.oz, ge.spacing)


class BatchPair(C.Structure):
    _fields_ = [("rec", C.c_uint64), ("lig", C.c_uint64), ("grid", C.c_uint64),
                ("params", C.c_uint64), ("d_dofs", P), ("d_out", P), ("d_w", P),
                ("n", C.c_size_t), ("modes", C.c_int)]


class BatchBucket(C.Structure):
    _fields_ = [("modeCapacity", C.c_uint), ("sizeClass", C.c_int), ("numPairs", C.c_uint),
                ("numTiles", C.c_uint), ("wavesPerPose", C.c_int), ("posesPerTile", C.c_uint),
                ("ms", C.c_float)]


PAIR_DTYPE = np.dtype([("rc", np.float32), ("ac", np.float32), ("ipon", np.int32),
                       ("rmin2", np.float32), ("emin", np.float32)])

_lib = None


def protein_desc(body):
    """att_protein_desc over a workload Body's arrays (which must outlive it)."""
    return ProteinDesc(body.n, body.M, _p(body.xyz), _p(body.type), _p(body.mtype),
                       _p(body.charge), _p(body.modes) if body.M else None,
                       _p(body.mode_force) if body.M else None)


def grid_desc(g):
    ge = g.geom
    return GridDesc(ge.nx, ge.ny, ge.nz, g.ngrids, ge.ox, ge.oy, ge.oz, ge.spacing, _p(g.data))


def nl_desc(nl):
    ge = nl.geom
    return NLDesc(ge.nx, ge.ny, ge.nz, ge.ox, ge.oy, ge.oz, ge.spacing, nl.plateau,
                  _p(nl.desc), _p(nl.list) if len(nl.list) else None, len(nl.list))


def param_table(p):
    """The T*T att_pair_params table of a Params object."""
    tab = np.zeros(p.T * p.T, PAIR_DTYPE)
    tab["rc"], tab["ac"], tab["ipon"] = p.rc, p.ac, p.ipon
    tab["rmin2"], tab["emin"] = p.rmin2, p.emin
    return tab


def param_grad_rmin_emin(params, pgrad):
    """(T*T, 2) {dE/drmin, dE/demin} from the (T*T, 3) {dE/drc, dE/dac,
    dE/demin} parameter gradient (att_param_grad_rmin_emin: the chain rule
    through the table's derivation, fileIO/readFile.cpp:479-513)."""
    tab = param_table(params)
    pg = np.ascontiguousarray(pgrad, np.float32).reshape(-1, 3)
    out = np.zeros((pg.shape[0], 2), np.float32)
    st = lib().att_param_grad_rmin_emin(_p(tab), params.T, params.shape, _p(pg), _p(out))
    if st != 0:
        raise AttError(st, "att_param_grad_rmin_emin: invalid arguments")
    return out


class AttError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} missing: the HIP extension must be built "
                               "(__graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIBPATH)
        L.att_ctx_create.argtypes = [C.c_int, C.POINTER(P)]
        L.att_set_pose_order.argtypes = [P, C.c_int]
        L.att_set_modes_waves.argtypes = [P, C.c_int]
        L.att_debug_pair_counts.argtypes = [P, P, C.c_size_t]
        L.att_ctx_destroy.argtypes = [P]
        L.att_ctx_destroy.restype = None
        L.att_last_error.argtypes = [P]
        L.att_last_error.restype = C.c_char_p
        L.att_upload_protein.argtypes = [P, C.POINTER(ProteinDesc), C.POINTER(H)]
        L.att_upload_grid.argtypes = [P, C.POINTER(GridDesc), C.POINTER(GridDesc),
                                      C.POINTER(NLDesc), C.POINTER(H)]
        L.att_upload_params.argtypes = [P, P, C.c_uint, C.c_int, C.c_int, C.c_float, C.c_float,
                                        C.POINTER(H)]
        L.att_release.argtypes = [P, H]
        L.att_score_rigid.argtypes = [P, H, H, H, H, P, C.c_size_t, P, P]
        L.att_score_rigid_dev.argtypes = [P, H, H, H, H, P, C.c_size_t, P, P]
        L.att_score_modes.argtypes = [P, H, H, H, H, H, P, C.c_size_t, P, P]
        L.att_score_modes_dev.argtypes = [P, H, H, H, H, H, P, C.c_size_t, P, P]
        L.att_debug_beads_rigid.argtypes = [P, H, H, H, H, P, P, P, P, P]
        L.att_debug_beads_modes.argtypes = [P, H, H, H, H, H, P, P, P, P, P, P, P]
        L.att_debug_rotmats.argtypes = [P, P, C.c_size_t, P]
        L.att_debug_sincosf.argtypes = [P, P, C.c_size_t, P, P]
        L.att_debug_plateau_ratio.argtypes = [P, C.c_float, C.c_size_t, P]
        L.att_param_grad_rmin_emin.argtypes = [P, C.c_uint, C.c_int, P, P]
        L.att_score_adjoint_dev_noalloc.argtypes = [P, H, H, H, H, P, C.c_int, C.c_size_t, P, P, P, P]
        L.att_sync.argtypes = [P, P]
        L.att_build_info.restype = C.c_char_p
        L.att_dev_alloc.argtypes = [P, C.c_size_t, C.POINTER(P)]
        L.att_dev_free.argtypes = [P, P]
        L.att_host_alloc.argtypes = [P, C.c_size_t, C.POINTER(P)]
        L.att_host_free.argtypes = [P, P]
        L.att_memcpy_h2d.argtypes = [P, P, P, C.c_size_t, P]
        L.att_memcpy_d2h.argtypes = [P, P, P, C.c_size_t, P]
        L.att_get_stream.argtypes = [P, C.POINTER(P)]
        L.att_event_create.argtypes = [P, C.POINTER(P)]
        L.att_event_destroy.argtypes = [P, P]
        L.att_event_record.argtypes = [P, P, P]
        L.att_event_elapsed_ms.argtypes = [P, P, P, C.POINTER(C.c_float)]
        L.att_debug_copy.argtypes = [P, P, P, C.c_size_t, P]
        L.att_upload_basis.argtypes = [P, H, C.c_uint, P, P]
        L.att_score_adjoint.argtypes = [P, H, H, H, H, P, C.c_int, C.c_size_t, P, P, P, P]
        L.att_score_adjoint_dev.argtypes = [P, H, H, H, H, P, C.c_int, C.c_size_t, P, P, P, P]
        L.att_stream_create.argtypes = [P, C.POINTER(P)]
        L.att_stream_destroy.argtypes = [P, P]
        L.att_memcpy_h2d_async.argtypes = [P, P, P, C.c_size_t, P]
        L.att_memcpy_d2h_async.argtypes = [P, P, P, C.c_size_t, P]
        L.att_event_query.argtypes = [P, P]
        L.att_event_sync.argtypes = [P, P]
        L.att_minimize.argtypes = [P, H, H, H, H, P, C.c_int, C.c_size_t, C.c_int, P, P, P, P, P]
        L.att_minimize_dev.argtypes = [P, H, H, H, H, P, C.c_int, C.c_size_t, C.c_int, P, P, P, P,
                                       P]
        L.att_generate_grid.argtypes = [P, C.POINTER(ProteinDesc), P, C.c_uint, C.c_int, C.c_int,
                                        C.c_float, C.c_float, C.c_int, C.POINTER(GridGeom), P]
        L.att_generate_nl.argtypes = [P, C.POINTER(ProteinDesc), C.POINTER(GridGeom), C.c_float,
                                      P, P, C.c_size_t, C.POINTER(C.c_uint64)]
        L.att_build_grid.argtypes = [P, C.POINTER(ProteinDesc), P, C.c_uint, C.c_int, C.c_int,
                                     C.c_float, C.c_float, C.c_int, C.POINTER(GridGeom),
                                     C.POINTER(GridGeom), C.c_float, C.POINTER(H)]
        L.att_batch_create.argtypes = [P, C.POINTER(BatchPair), C.c_uint, C.POINTER(P)]
        L.att_batch_destroy.argtypes = [P, P]
        L.att_batch_buckets.argtypes = [P]
        L.att_batch_buckets.restype = C.c_uint
        L.att_batch_bucket_info.argtypes = [P, P, C.c_uint, C.POINTER(BatchBucket)]
        L.att_batch_pair_bucket.argtypes = [P, P, C.c_uint, C.POINTER(C.c_uint)]
        L.att_batch_set_shape.argtypes = [P, P, C.c_uint, C.c_int, C.c_uint]
        L.att_batch_tune.argtypes = [P, P, P]
        L.att_batch_run.argtypes = [P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


class DeviceBuffer:
    """Device allocation owned by a Context (freed on release or GC)."""

    def __init__(self, ctx, ptr, nbytes):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def release(self):
        if self.ptr and self.ctx.h:
            self.ctx.L.att_dev_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Batch:
    """att_batch: many pairs scored by graph-captured persistent launches."""

    def __init__(self, ctx, pairs):
        self.ctx = ctx
        arr = (BatchPair * len(pairs))()
        for k, q in enumerate(pairs):
            hs = q["hs"]
            arr[k] = BatchPair(hs["rec"], hs["lig"], hs["grid"], hs["params"], q["d_dofs"],
                               q["d_out"], q.get("d_w"), q["n"], 1 if q["modes"] else 0)
        h = P()
        ctx._chk(ctx.L.att_batch_create(ctx.h, arr, len(pairs), C.byref(h)))
        self.h = h.value

    def buckets(self):
        out = []
        for b in range(self.ctx.L.att_batch_buckets(self.h)):
            info = BatchBucket()
            self.ctx._chk(self.ctx.L.att_batch_bucket_info(self.ctx.h, self.h, b, C.byref(info)))
            out.append({f: getattr(info, f) for f, _ in BatchBucket._fields_})
        return out

    def pair_bucket(self, pair):
        k = C.c_uint()
        self.ctx._chk(self.ctx.L.att_batch_pair_bucket(self.ctx.h, self.h, pair, C.byref(k)))
        return k.value

    def set_shape(self, bucket, waves, tile):
        self.ctx._chk(self.ctx.L.att_batch_set_shape(self.ctx.h, self.h, bucket, waves, tile))

    def tune(self, stream=None):
        self.ctx._chk(self.ctx.L.att_batch_tune(self.ctx.h, self.h, stream))

    def run(self, stream=None):
        self.ctx._chk(self.ctx.L.att_batch_run(self.ctx.h, self.h, stream))

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.att_batch_destroy(self.ctx.h, self.h)
        self.h = None


class Context:
    """One att_ctx on a HIP device, with helpers to upload a Workload."""

    def __init__(self, device=0):
        self.L = lib()
        h = P()
        st = self.L.att_ctx_create(device, C.byref(h))
        if st != ATT_OK:
            raise AttError(st, f"cannot create a context on HIP device {device}")
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            self.L.att_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st != ATT_OK:
            raise AttError(st, self.L.att_last_error(self.h).decode())

    def set_pose_order(self, enabled):
        """Spatial pose order of the single-pair adjoint (default on)."""
        self._chk(self.L.att_set_pose_order(self.h, 1 if enabled else 0))

    def pair_counts(self, fn, n):
        """Per-pose {listed neighbours walked, pairs inside the plateau} of the
        scoring call fn() over n poses, from the production kernel: (n, 2) uint32."""
        buf = self.dev_alloc(max(1, n) * 8)
        self._chk(self.L.att_debug_pair_counts(self.h, buf.ptr, n))
        try:
            fn()
            self.sync()
            return self.to_host(buf, (n, 2), np.uint32)
        finally:
            self._chk(self.L.att_debug_pair_counts(self.h, None, 0))
            buf.release()

    def set_modes_waves(self, waves):
        """Waves per pose of the flexible forward kernel (0: by ligand size)."""
        self._chk(self.L.att_set_modes_waves(self.h, waves))

    # -- uploads ---------------------------------------------------------
    def upload_protein(self, body):
        d = protein_desc(body)
        out = H()
        self._chk(self.L.att_upload_protein(self.h, C.byref(d), C.byref(out)))
        return out.value

    def upload_grid(self, inner, outer, nl=None):
        gi, go = grid_desc(inner), grid_desc(outer)
        nd = nl_desc(nl) if nl is not None else None
        out = H()
        self._chk(self.L.att_upload_grid(self.h, C.byref(gi), C.byref(go),
                                         C.byref(nd) if nd is not None else None, C.byref(out)))
        return out.value

    def upload_params(self, p):
        tab = param_table(p)
        out = H()
        self._chk(self.L.att_upload_params(self.h, _p(tab), p.T, p.shape, p.dielec, p.epsilon,
                                           p.ffelec, C.byref(out)))
        return out.value

    def release(self, handle):
        self._chk(self.L.att_release(self.h, handle))

    # -- grid generation on the device -----------------------------------
    def generate_grid(self, body, p, geom, plateau=10.0, src_is_ligand=False):
        """The (T+1) point grids of `body` over `geom` (a workload Geom), as a
        flat float32 array [grid][z][y][x][4]."""
        d, g, tab = protein_desc(body), grid_geom(geom), param_table(p)
        out = np.zeros((p.T + 1) * geom.nvox * 4, np.float32)
        self._chk(self.L.att_generate_grid(self.h, C.byref(d), _p(tab), p.T, p.shape, p.dielec,
                                           p.ffelec, plateau, 1 if src_is_ligand else 0,
                                           C.byref(g), _p(out)))
        return out

    def generate_nl(self, body, geom, neighbourdis=10.8):
        """(desc, list) of the neighbour lists of `body` over `geom`."""
        d, g = protein_desc(body), grid_geom(geom)
        desc = np.zeros(geom.nvox * 2, np.uint32)
        tot = C.c_uint64(0)
        self._chk(self.L.att_generate_nl(self.h, C.byref(d), C.byref(g), neighbourdis, _p(desc),
                                         None, 0, C.byref(tot)))
        lst = np.zeros(max(1, tot.value), np.uint32)
        self._chk(self.L.att_generate_nl(self.h, C.byref(d), C.byref(g), neighbourdis, _p(desc),
                                         _p(lst), len(lst), C.byref(tot)))
        return desc, lst[:tot.value]

    def build_grid(self, body, p, inner, outer, plateau=10.0, neighbourdis=10.8,
                   src_is_ligand=False):
        """Grid handle built on the device (neighbourdis <= 0: no neighbour list)."""
        d, tab = protein_desc(body), param_table(p)
        gi, go = grid_geom(inner), grid_geom(outer)
        out = H()
        self._chk(self.L.att_build_grid(self.h, C.byref(d), _p(tab), p.T, p.shape, p.dielec,
                                        p.ffelec, plateau, 1 if src_is_ligand else 0, C.byref(gi),
                                        C.byref(go), neighbourdis, C.byref(out)))
        return out.value

    def upload_workload(self, wl):
        """Upload everything a workload needs; returns a dict of handles.  A
        workload without host grids (make_workload(host_grids=False)) gets its
        receptor grids built on the device."""
        if wl.rec_inner.data is None:
            m = wl.meta
            grid = self.build_grid(wl.rec, wl.params, wl.rec_inner.geom, wl.rec_outer.geom,
                                   m["plateau"], m["neighbourdis"])
        else:
            grid = self.upload_grid(wl.rec_inner, wl.rec_outer, wl.rec_nl)
        hs = dict(rec=self.upload_protein(wl.rec), lig=self.upload_protein(wl.lig), grid=grid,
                  params=self.upload_params(wl.params))
        if wl.lig_inner is not None:
            hs["lig_grid"] = self.upload_grid(wl.lig_inner, wl.lig_outer, None)
        return hs

    # -- scoring ---------------------------------------------------------
    def score_rigid(self, hs, dofs):
        d = np.ascontiguousarray(dofs, np.float32).reshape(-1, 6)
        out = np.zeros((d.shape[0], 7), np.float32)
        self._chk(self.L.att_score_rigid(self.h, hs["rec"], hs["lig"], hs["grid"], hs["params"],
                                         _p(d), d.shape[0], _p(out), None))
        return out

    def score_rigid_dev(self, hs, d_dofs, n, d_out, stream=None):
        self._chk(self.L.att_score_rigid_dev(self.h, hs["rec"], hs["lig"], hs["grid"],
                                             hs["params"], d_dofs, n, d_out, stream))

    def score_modes(self, hs, dofs):
        d = np.ascontiguousarray(dofs, np.float32).reshape(-1, 26)
        out = np.zeros((d.shape[0], 27), np.float32)
        self._chk(self.L.att_score_modes(self.h, hs["rec"], hs["lig"], hs["grid"], hs["lig_grid"],
                                         hs["params"], _p(d), d.shape[0], _p(out), None))
        return out

    def upload_basis(self, grid_handle, wl):
        self._chk(self.L.att_upload_basis(self.h, grid_handle, wl.params.T,
                                          _p(wl.rec_inner_basis), _p(wl.rec_outer_basis)))

    def adjoint(self, hs, dofs, w=None, pgrad=False, T=None):
        """Exact gradient of the returned energy: (n, 7|27) {E, w*dE/dpos,
        w*dE/dang[, w*dE/dmodesRec[10], w*dE/dmodesLig[10]]} and optionally the
        (T*T, 3) parameter gradient."""
        d = np.ascontiguousarray(dofs, np.float32)
        modes = d.shape[1] == 26
        out = np.zeros((d.shape[0], 27 if modes else 7), np.float32)
        ww = None if w is None else np.ascontiguousarray(w, np.float32)
        pg = np.zeros((T * T, 3), np.float32) if pgrad else None
        self._chk(self.L.att_score_adjoint(self.h, hs["rec"], hs["lig"], hs["grid"], hs["params"],
                                           _p(d), 1 if modes else 0, d.shape[0], _p(ww), _p(out),
                                           _p(pg), None))
        return (out, pg) if pgrad else out

    def adjoint_dev(self, hs, d_dofs, modes, n, d_w, d_out, d_pgrad=None, stream=None):
        self._chk(self.L.att_score_adjoint_dev(self.h, hs["rec"], hs["lig"], hs["grid"],
                                               hs["params"], d_dofs, 1 if modes else 0, n, d_w,
                                               d_out, d_pgrad, stream))

    def adjoint_dev_noalloc(self, hs, d_dofs, modes, n, d_w, d_out, d_pgrad=None, stream=None):
        self._chk(self.L.att_score_adjoint_dev_noalloc(self.h, hs["rec"], hs["lig"], hs["grid"],
                                                       hs["params"], d_dofs, 1 if modes else 0, n,
                                                       d_w, d_out, d_pgrad, stream))

    def score_modes_dev(self, hs, d_dofs, n, d_out, stream=None):
        self._chk(self.L.att_score_modes_dev(self.h, hs["rec"], hs["lig"], hs["grid"],
                                             hs["lig_grid"], hs["params"], d_dofs, n, d_out,
                                             stream))

    def sync(self, stream=None):
        self._chk(self.L.att_sync(self.h, stream))

    def minimize(self, hs, dofs, max_fun_eval=500):
        """VA13 from every pose (host arrays): (dofs_out, e_last, e_best, nfun)."""
        d = np.ascontiguousarray(dofs, np.float32)
        n, modes = d.shape[0], d.shape[1] == 26
        out = np.zeros_like(d)
        el, eb = np.zeros(n, np.float32), np.zeros(n, np.float32)
        nf = np.zeros(n, np.int32)
        self._chk(self.L.att_minimize(self.h, hs["rec"], hs["lig"], hs["grid"], hs["params"],
                                      _p(d), 1 if modes else 0, n, max_fun_eval, _p(out), _p(el),
                                      _p(eb), _p(nf), None))
        return out, el, eb, nf

    def minimize_dev(self, hs, d_dofs, modes, n, max_fun_eval, d_out, d_energy, d_ebest=None,
                     d_nfun=None, stream=None):
        self._chk(self.L.att_minimize_dev(self.h, hs["rec"], hs["lig"], hs["grid"], hs["params"],
                                          d_dofs, 1 if modes else 0, n, max_fun_eval, d_out,
                                          d_energy, d_ebest, d_nfun, stream))

    def batch(self, pairs):
        """pairs: dicts {hs, d_dofs, d_out, n, modes[, d_w]} (device pointers)."""
        return Batch(self, pairs)

    # -- device buffers and timing ----------------------------------------
    def dev_alloc(self, nbytes):
        p = P()
        self._chk(self.L.att_dev_alloc(self.h, nbytes, C.byref(p)))
        return DeviceBuffer(self, p.value, nbytes)

    def to_device(self, arr):
        a = np.ascontiguousarray(arr)
        buf = self.dev_alloc(a.nbytes)
        self._chk(self.L.att_memcpy_h2d(self.h, buf.ptr, _p(a), a.nbytes, None))
        return buf

    def write(self, buf, arr):
        """Copy a host array into the start of an existing device buffer."""
        a = np.ascontiguousarray(arr)
        if a.nbytes > buf.nbytes:
            raise ValueError("array larger than the device buffer")
        self._chk(self.L.att_memcpy_h2d(self.h, buf.ptr, _p(a), a.nbytes, None))

    def to_host(self, buf, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        self._chk(self.L.att_memcpy_d2h(self.h, _p(out), buf.ptr, out.nbytes, None))
        return out

    def copy(self, dst, src, nbytes, stream=None):
        self._chk(self.L.att_debug_copy(self.h, dst.ptr, src.ptr, nbytes, stream))

    def stream(self):
        s = P()
        self._chk(self.L.att_get_stream(self.h, C.byref(s)))
        return s.value

    def new_stream(self):
        """A stream of the context's own (att_stream_create), destroyed with it."""
        s = P()
        self._chk(self.L.att_stream_create(self.h, C.byref(s)))
        return s.value

    def event(self):
        e = P()
        self._chk(self.L.att_event_create(self.h, C.byref(e)))
        return e.value

    def record(self, ev, stream=None):
        self._chk(self.L.att_event_record(self.h, ev, stream))

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        self._chk(self.L.att_event_elapsed_ms(self.h, a, b, C.byref(ms)))
        return ms.value

    # -- diagnostics -----------------------------------------------------
    def beads(self, hs, wl, dof):
        n, nr = wl.lig.n, wl.rec.n
        out = dict(xyz=np.zeros(3 * n, np.float32), pot=np.zeros(4 * n, np.float32),
                   cell=np.full(9 * n, -1, np.int32), nl_pairs=np.zeros(n, np.uint32))
        d = np.ascontiguousarray(dof, np.float32)
        if wl.modes:
            out["rec_pot"] = np.zeros(3 * nr, np.float32)
            out["rec_xyz"] = np.zeros(3 * nr, np.float32)
            self._chk(self.L.att_debug_beads_modes(
                self.h, hs["rec"], hs["lig"], hs["grid"], hs["lig_grid"], hs["params"], _p(d),
                _p(out["xyz"]), _p(out["pot"]), _p(out["cell"]), _p(out["nl_pairs"]),
                _p(out["rec_pot"]), _p(out["rec_xyz"])))
        else:
            self._chk(self.L.att_debug_beads_rigid(
                self.h, hs["rec"], hs["lig"], hs["grid"], hs["params"], _p(d), _p(out["xyz"]),
                _p(out["pot"]), _p(out["cell"]), _p(out["nl_pairs"])))
        out["cell"] = out["cell"].reshape(n, 9)
        return out

    def rotmats(self, angles):
        a = np.ascontiguousarray(angles, np.float32).reshape(-1, 3)
        R = np.zeros((a.shape[0], 9), np.float32)
        self._chk(self.L.att_debug_rotmats(self.h, _p(a), a.shape[0], _p(R)))
        return R

    def sincosf(self, x):
        """(sin, cos) of float32 values as the kernels evaluate them."""
        a = np.ascontiguousarray(x, np.float32).ravel()
        s = np.zeros_like(a)
        c = np.zeros_like(a)
        self._chk(self.L.att_debug_sincosf(self.h, _p(a), a.size, _p(s), _p(c)))
        return s, c

    def plateau_ratio(self, P2, n):
        """Entries 0..n-1 of the plateau-ratio table of plateau^2 P2 as the
        kernels evaluate them per pair."""
        r = np.zeros(n, np.float32)
        self._chk(self.L.att_debug_plateau_ratio(self.h, float(P2), n, _p(r)))
        return r
