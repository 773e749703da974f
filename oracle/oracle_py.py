"""ctypes bindings to the CPU oracle (liboracle.so) and the compiled reference
harness (_ref/libattract_ref.so).

TEST INFRASTRUCTURE ONLY — imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, never by product code.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class IGridC(C.Structure):
    _fields_ = [("nx", C.c_uint), ("ny", C.c_uint), ("nz", C.c_uint), ("ngrids", C.c_uint),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("spacing", C.c_float),
                ("data", P)]


class NLGridC(C.Structure):
    _fields_ = [("nx", C.c_uint), ("ny", C.c_uint), ("nz", C.c_uint),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("spacing", C.c_float),
                ("plateau", C.c_float), ("desc", P), ("list", P), ("nlist", C.c_uint64)]


class BodyC(C.Structure):
    _fields_ = [("n", C.c_uint), ("M", C.c_uint), ("xyz", P), ("type", P), ("mtype", P),
                ("charge", P), ("modes", P), ("modeForce", P)]


class OrcParamsC(C.Structure):
    _fields_ = [("T", C.c_uint), ("shape", C.c_int), ("dielec", C.c_int), ("ffelec", C.c_float),
                ("rc", P), ("ac", P), ("ipon", P), ("rmin2", P), ("emin", P)]


class RefTableC(C.Structure):
    _fields_ = [("T", C.c_uint), ("shape", C.c_int), ("dielec", C.c_int),
                ("epsilon", C.c_float), ("ffelec", C.c_float),
                ("rc", P), ("ac", P), ("ipon", P), ("rmin2", P), ("emin", P)]


class BeadsC(C.Structure):
    _fields_ = [("xyz", P), ("pot", P), ("cell", P), ("nl_pairs", P), ("rec_pot", P),
                ("rec_xyz", P)]


def igrid_c(g):
    ge = g.geom
    return IGridC(ge.nx, ge.ny, ge.nz, g.ngrids, ge.ox, ge.oy, ge.oz, ge.spacing, _p(g.data))


def nl_c(g):
    ge = g.geom
    return NLGridC(ge.nx, ge.ny, ge.nz, ge.ox, ge.oy, ge.oz, ge.spacing, g.plateau,
                   _p(g.desc), _p(g.list), len(g.list))


def body_c(b):
    return BodyC(b.n, b.M, _p(b.xyz), _p(b.type), _p(b.mtype), _p(b.charge),
                 _p(b.modes) if b.M else None, _p(b.mode_force) if b.M else None)


def params_c(p):
    return OrcParamsC(p.T, p.shape, p.dielec, p.ffelec, _p(p.rc), _p(p.ac), _p(p.ipon),
                      _p(p.rmin2), _p(p.emin))


def table_c(p):
    return RefTableC(p.T, p.shape, p.dielec, p.epsilon, p.ffelec, _p(p.rc), _p(p.ac),
                     _p(p.ipon), _p(p.rmin2), _p(p.emin))


_orc = None
_ref = None


def oracle_lib():
    global _orc
    if _orc is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: build the oracle (make -C oracle)")
        L = C.CDLL(path)
        rig = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC), C.POINTER(IGridC),
               C.POINTER(NLGridC), C.POINTER(OrcParamsC), P, C.c_size_t, P]
        L.orc_score_rigid.restype = C.c_int
        L.orc_score_rigid.argtypes = rig + [C.c_int]
        L.orc_score_rigid_f64.restype = C.c_int
        L.orc_score_rigid_f64.argtypes = rig + [P, C.c_int]
        L.orc_beads_rigid.restype = C.c_int
        L.orc_beads_rigid.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                      C.POINTER(IGridC), C.POINTER(NLGridC),
                                      C.POINTER(OrcParamsC), P, C.POINTER(BeadsC)]
        mod = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC), C.POINTER(IGridC),
               C.POINTER(NLGridC), C.POINTER(IGridC), C.POINTER(IGridC), C.POINTER(OrcParamsC),
               P, C.c_size_t, P]
        L.orc_score_modes.restype = C.c_int
        L.orc_score_modes.argtypes = mod + [C.c_int]
        L.orc_score_modes_f64.restype = C.c_int
        L.orc_score_modes_f64.argtypes = mod + [P, C.c_int]
        L.orc_beads_modes.restype = C.c_int
        L.orc_beads_modes.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                      C.POINTER(IGridC), C.POINTER(NLGridC), C.POINTER(IGridC),
                                      C.POINTER(IGridC), C.POINTER(OrcParamsC), P,
                                      C.POINTER(BeadsC)]
        FG = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                         C.POINTER(C.c_double))
        L.FG = FG
        L.orc_va13.restype = C.c_int
        L.orc_va13.argtypes = [C.c_int, P, C.c_int, FG, C.c_void_p, P, P, P, P]
        for fn in (L.orc_minimize_f32, L.orc_minimize_f64):
            fn.restype = C.c_int
            fn.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                           C.POINTER(IGridC), C.POINTER(NLGridC), C.POINTER(OrcParamsC), P,
                           C.c_size_t, C.c_int, P, P, P, P]
        for fn in (L.orc_adjoint_f64, L.orc_adjoint_f32):
            fn.restype = C.c_int
            fn.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                           C.POINTER(IGridC), C.POINTER(NLGridC),
                           C.POINTER(OrcParamsC), P, P, P, C.c_size_t, P, P, P, P]
        L.orc_ratio_size.restype = C.c_uint
        L.orc_ratio_size.argtypes = [C.c_float]
        L.orc_ratio_lut.restype = None
        L.orc_ratio_lut.argtypes = [C.c_float, P]
        L.orc_euler2rotmat.restype = None
        L.orc_euler2rotmat.argtypes = [C.c_float, C.c_float, C.c_float, P]
        L.orc_euler2torquemat.restype = None
        L.orc_euler2torquemat.argtypes = [C.c_float, C.c_float, C.c_float, P]
        _orc = L
    return _orc


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libattract_ref.so"))


def ref_lib():
    global _ref
    if _ref is None:
        L = C.CDLL(os.path.join(HERE, "_ref", "libattract_ref.so"))
        L.ref_score_rigid.restype = C.c_int
        L.ref_score_rigid.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                      C.POINTER(IGridC), C.POINTER(NLGridC),
                                      C.POINTER(RefTableC), P, C.c_size_t, P]
        L.ref_beads_rigid.restype = C.c_int
        L.ref_beads_rigid.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                      C.POINTER(IGridC), C.POINTER(NLGridC),
                                      C.POINTER(RefTableC), P, P, P]
        L.ref_cells_rigid.restype = C.c_int
        L.ref_cells_rigid.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                      C.POINTER(IGridC), C.POINTER(NLGridC), P, P, P]
        L.ref_score_modes.restype = C.c_int
        L.ref_score_modes.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                      C.POINTER(IGridC), C.POINTER(NLGridC), C.POINTER(IGridC),
                                      C.POINTER(IGridC), C.POINTER(NLGridC),
                                      C.POINTER(RefTableC), P, C.c_size_t, P, P, P]
        L.ref_score_rigid_mt.restype = C.c_int
        L.ref_score_rigid_mt.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                         C.POINTER(IGridC), C.POINTER(NLGridC),
                                         C.POINTER(RefTableC), P, C.c_size_t, P, C.c_int,
                                         C.POINTER(C.c_double)]
        L.ref_score_modes_mt.restype = C.c_int
        L.ref_score_modes_mt.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                         C.POINTER(IGridC), C.POINTER(NLGridC), C.POINTER(IGridC),
                                         C.POINTER(IGridC), C.POINTER(NLGridC),
                                         C.POINTER(RefTableC), P, C.c_size_t, P, C.c_int,
                                         C.POINTER(C.c_double)]
        L.ref_pdb.restype = C.c_int
        L.ref_pdb.argtypes = [C.c_char_p, C.c_uint, C.POINTER(C.c_uint), P, P, P]
        L.ref_modes.restype = C.c_int
        L.ref_modes.argtypes = [C.c_char_p, C.c_char_p, C.c_uint, P, P]
        L.ref_params_file.restype = C.c_int
        L.ref_params_file.argtypes = [C.c_char_p, C.c_uint, C.POINTER(C.c_uint), P, P, P, P, P]
        L.ref_dof.restype = C.c_int
        L.ref_dof.argtypes = [C.c_char_p, C.c_int, C.c_uint, C.c_uint, P, P, C.c_size_t,
                              C.POINTER(C.c_size_t), P]
        L.ref_format_results.restype = C.c_size_t
        L.ref_format_results.argtypes = [P, C.c_size_t, C.c_int, C.c_uint, C.c_uint, P, C.c_size_t]
        L.ref_alphabet.restype = C.c_int
        L.ref_alphabet.argtypes = [C.c_char_p, P, C.c_uint, C.POINTER(C.c_uint)]
        L.ref_score_files.restype = C.c_int
        L.ref_score_files.argtypes = [C.c_char_p] * 7 + [C.c_double, P, C.c_size_t,
                                                         C.POINTER(C.c_size_t)]
        L.ref_score_files_modes.restype = C.c_int
        L.ref_score_files_modes.argtypes = [C.c_char_p] * 10 + [C.c_int, C.c_char_p, C.c_double,
                                                                  P, C.c_size_t,
                                                                  C.POINTER(C.c_size_t)]
        L.ref_gridfile_load.restype = C.c_int
        L.ref_gridfile_load.argtypes = [C.c_char_p]
        L.ref_gridfile_info.restype = None
        L.ref_gridfile_info.argtypes = [P, P, C.POINTER(C.c_uint64)]
        L.ref_gridfile_copy.restype = None
        L.ref_gridfile_copy.argtypes = [P, P, P, P]
        L.ref_gridfile_free.restype = None
        L.ref_read_params.restype = C.c_int
        L.ref_read_params.argtypes = [C.c_uint, C.c_int, P, P, P, P, P, P, P, P]
        L.ref_ratio_lut.restype = C.c_uint
        L.ref_ratio_lut.argtypes = [C.c_float, P, C.c_uint]
        L.ref_euler2rotmat.restype = None
        L.ref_euler2rotmat.argtypes = [C.c_float, C.c_float, C.c_float, P]
        L.ref_euler2torquemat.restype = None
        L.ref_euler2torquemat.argtypes = [C.c_float, C.c_float, C.c_float, P]
        _ref = L
    return _ref


# ---------------------------------------------------------------- oracle API

def score_rigid(wl, dofs=None, nthreads=1, f64=False):
    """FP32 results (n, 7); with f64=True the FP64 recompute and its condition
    scale (sum of |terms| per component), both (n, 7) float64."""
    L = oracle_lib()
    d = np.ascontiguousarray(wl.dofs if dofs is None else dofs, np.float32)
    n = d.shape[0]
    args = (C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)), C.byref(igrid_c(wl.rec_inner)),
            C.byref(igrid_c(wl.rec_outer)), C.byref(nl_c(wl.rec_nl)),
            C.byref(params_c(wl.params)), _p(d), n)
    if f64:
        res = np.zeros((n, 7), np.float64)
        scale = np.zeros((n, 7), np.float64)
        L.orc_score_rigid_f64(*args, _p(res), _p(scale), nthreads)
        return res, scale
    res = np.zeros((n, 7), np.float32)
    L.orc_score_rigid(*args, _p(res), nthreads)
    return res


def score_modes(wl, dofs=None, nthreads=1, f64=False):
    """FP32 results (n, 27); with f64=True the FP64 recompute and its scale."""
    L = oracle_lib()
    d = np.ascontiguousarray(wl.dofs if dofs is None else dofs, np.float32)
    n = d.shape[0]
    args = (C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)), C.byref(igrid_c(wl.rec_inner)),
            C.byref(igrid_c(wl.rec_outer)), C.byref(nl_c(wl.rec_nl)),
            C.byref(igrid_c(wl.lig_inner)), C.byref(igrid_c(wl.lig_outer)),
            C.byref(params_c(wl.params)), _p(d), n)
    if f64:
        res = np.zeros((n, 27), np.float64)
        scale = np.zeros((n, 27), np.float64)
        L.orc_score_modes_f64(*args, _p(res), _p(scale), nthreads)
        return res, scale
    res = np.zeros((n, 27), np.float32)
    L.orc_score_modes(*args, _p(res), nthreads)
    return res


def as_modes_dofs(wl, dofs=None):
    """DOFs in the 26-float DOF_6D_Modes layout (rigid DOFs padded with zeros)."""
    d = np.asarray(wl.dofs if dofs is None else dofs, np.float32)
    if d.shape[1] == 26:
        return np.ascontiguousarray(d)
    out = np.zeros((d.shape[0], 26), np.float32)
    out[:, :6] = d
    return out


def adjoint(wl, dofs=None, w=None, pgrad=False, scale=False, f32=False):
    """Exact gradient of the returned energy: (n, 27) {E, dE/dpos, dE/dang,
    dE/dmodesRec[10], dE/dmodesLig[10]} times w, and optionally the (T*T, 3)
    parameter gradient {dE/drc, dE/dac, dE/demin} summed over poses.  FP64
    arithmetic, or FP32 with f32=True (same bead positions as the FP32
    forward); results are returned as float64 either way."""
    L = oracle_lib()
    d = as_modes_dofs(wl, dofs)
    n = d.shape[0]
    out = np.zeros((n, 27), np.float64)
    ww = None if w is None else np.ascontiguousarray(w, np.float64)
    pg = np.zeros((wl.params.T ** 2, 3), np.float64) if pgrad else None
    sc = np.zeros((n, 27), np.float64) if scale else None
    fn = L.orc_adjoint_f32 if f32 else L.orc_adjoint_f64
    fn(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                      C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                      C.byref(nl_c(wl.rec_nl)), C.byref(params_c(wl.params)),
                      _p(wl.rec_inner_basis), _p(wl.rec_outer_basis), _p(d), n, _p(ww), _p(out),
                      _p(sc), _p(pg))
    ret = (out,)
    if scale:
        ret += (sc,)
    if pgrad:
        ret += (pg,)
    return ret if len(ret) > 1 else out


def va13(fg, x0, ivmax=500):
    """VA13 on a Python objective fg(x) -> (f, g).  Returns (x_last, f_last,
    f_best, nfun, evaluations) where evaluations lists every (x, f) asked for."""
    L = oracle_lib()
    n = len(x0)
    evals = []

    def cb(_, xp, fp, gp):
        x = np.ctypeslib.as_array(xp, (n,)).copy()
        f, g = fg(x)
        evals.append((x, float(f)))
        fp[0] = float(f)
        for i in range(n):
            gp[i] = float(g[i])

    cfun = L.FG(cb)
    x = np.ascontiguousarray(x0, np.float64)
    xl = np.zeros(n)
    fl, fb, nf = C.c_double(), C.c_double(), C.c_int()
    L.orc_va13(n, _p(x), ivmax, cfun, None, _p(xl), C.byref(fl), C.byref(fb), C.byref(nf))
    return xl, fl.value, fb.value, nf.value, evals


def minimize(wl, dofs=None, ivmax=500, f64=False):
    """VA13 on the FP32 (or, with f64=True, FP64) adjoint for each pose:
    (dofs_last (n, 26) float32, e_last, e_best, nfun)."""
    L = oracle_lib()
    d = as_modes_dofs(wl, dofs)
    n = d.shape[0]
    out = np.zeros((n, 26), np.float32)
    el, eb = np.zeros(n), np.zeros(n)
    nf = np.zeros(n, np.int32)
    (L.orc_minimize_f64 if f64 else L.orc_minimize_f32)(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                       C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                       C.byref(nl_c(wl.rec_nl)), C.byref(params_c(wl.params)), _p(d), n, ivmax,
                       _p(out), _p(el), _p(eb), _p(nf))
    return out, el, eb, nf


def kernel_waves(wl):
    """Waves per pose the HIP library uses for a single pair (1 below 384
    ligand beads, else 2)."""
    return 1 if wl.lig.n < 384 else 2


def _chunks(n, parts):
    step = max(1, -(-n // parts))
    return [(a, min(n, a + step)) for a in range(0, n, step)]


def adjoint_kernel_order(wl, dofs=None, w=None, waves=None, nthreads=8):
    """The FP32 exact adjoint in the HIP kernel's summation order for `waves`
    waves per pose (default: the single-pair choice): (n, 27) float64 holding
    float32 values; equal to the kernel's output bit for bit."""
    from concurrent.futures import ThreadPoolExecutor
    L = oracle_lib()
    L.orc_adjoint_kernel_order.restype = C.c_int
    L.orc_adjoint_kernel_order.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                           C.POINTER(IGridC), C.POINTER(NLGridC),
                                           C.POINTER(OrcParamsC), P, C.c_size_t, P, C.c_int, P]
    d = as_modes_dofs(wl, dofs)
    n = d.shape[0]
    out = np.zeros((n, 27), np.float64)
    ww = None if w is None else np.ascontiguousarray(w, np.float32)
    W = kernel_waves(wl) if waves is None else waves
    args = (body_c(wl.rec), body_c(wl.lig), igrid_c(wl.rec_inner), igrid_c(wl.rec_outer),
            nl_c(wl.rec_nl), params_c(wl.params))

    def run(ab):
        a, b = ab
        if b > a:
            rc = L.orc_adjoint_kernel_order(*[C.byref(x) for x in args], _p(d[a:b]), b - a,
                                            None if ww is None else _p(ww[a:b]), W, _p(out[a:b]))
            assert rc == 0
    with ThreadPoolExecutor(nthreads) as ex:
        list(ex.map(run, _chunks(n, nthreads)))
    return out


def minimize_kernel_order(wl, dofs=None, ivmax=500, waves=None, nthreads=8):
    """VA13 on the kernel-order FP32 adjoint (the device minimizer's
    trajectories): (dofs_last (n, 26) float32, e_last, e_best, nfun)."""
    from concurrent.futures import ThreadPoolExecutor
    L = oracle_lib()
    L.orc_minimize_kernel_order.restype = C.c_int
    L.orc_minimize_kernel_order.argtypes = [C.POINTER(BodyC), C.POINTER(BodyC), C.POINTER(IGridC),
                                            C.POINTER(IGridC), C.POINTER(NLGridC),
                                            C.POINTER(OrcParamsC), P, C.c_size_t, C.c_int, C.c_int,
                                            P, P, P, P]
    d = as_modes_dofs(wl, dofs)
    n = d.shape[0]
    out = np.zeros((n, 26), np.float32)
    el, eb = np.zeros(n), np.zeros(n)
    nf = np.zeros(n, np.int32)
    W = kernel_waves(wl) if waves is None else waves
    args = (body_c(wl.rec), body_c(wl.lig), igrid_c(wl.rec_inner), igrid_c(wl.rec_outer),
            nl_c(wl.rec_nl), params_c(wl.params))

    def run(ab):
        a, b = ab
        if b > a:
            rc = L.orc_minimize_kernel_order(*[C.byref(x) for x in args], _p(d[a:b]), b - a, ivmax, W,
                                             _p(out[a:b]), _p(el[a:b]), _p(eb[a:b]), _p(nf[a:b]))
            assert rc == 0
    with ThreadPoolExecutor(nthreads) as ex:
        list(ex.map(run, _chunks(n, nthreads)))
    return out, el, eb, nf


def beads(wl, dof):
    """Per-bead diagnostics of one pose: transformed coordinates, forces/energy,
    cell indices (inner xyz, outer xyz, NL xyz; -1 unused) and NL pair counts."""
    L = oracle_lib()
    n, nr = wl.lig.n, wl.rec.n
    out = dict(xyz=np.zeros(3 * n, np.float32), pot=np.zeros(4 * n, np.float32),
               cell=np.full(9 * n, -1, np.int32), nl_pairs=np.zeros(n, np.uint32))
    modes = wl.modes
    if modes:
        out["rec_pot"] = np.zeros(3 * nr, np.float32)
        out["rec_xyz"] = np.zeros(3 * nr, np.float32)
    b = BeadsC(_p(out["xyz"]), _p(out["pot"]), _p(out["cell"]), _p(out["nl_pairs"]),
               _p(out.get("rec_pot")), _p(out.get("rec_xyz")))
    d = np.ascontiguousarray(dof, np.float32)
    if modes:
        L.orc_beads_modes(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                          C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                          C.byref(nl_c(wl.rec_nl)), C.byref(igrid_c(wl.lig_inner)),
                          C.byref(igrid_c(wl.lig_outer)), C.byref(params_c(wl.params)), _p(d),
                          C.byref(b))
    else:
        L.orc_beads_rigid(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                          C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                          C.byref(nl_c(wl.rec_nl)), C.byref(params_c(wl.params)), _p(d),
                          C.byref(b))
    out["cell"] = out["cell"].reshape(n, 9)
    return out


def pair_counts(wl, dofs=None):
    """Per pose (n, 2) uint32: {neighbour-list entries of the cells of the
    ligand beads inside the neighbour-list box, pairs inside the plateau}, from
    the per-bead diagnostics of the NLPotForce restatement (model/
    neighborlist.h:47-158: the list of the bead's cell is walked, pairs with
    dr2 <= plateau^2 are evaluated)."""
    d = wl.dofs if dofs is None else dofs
    ge = wl.rec_nl.geom
    cnt = wl.rec_nl.desc.reshape(-1, 2)[:, 0]
    out = np.zeros((d.shape[0], 2), np.uint32)
    for k in range(d.shape[0]):
        b = beads(wl, d[k])
        c = b["cell"][:, 6:9]
        m = c[:, 0] >= 0
        v = (c[m, 2].astype(np.int64) * ge.ny + c[m, 1]) * ge.nx + c[m, 0]
        out[k, 0] = cnt[v].sum()
        out[k, 1] = b["nl_pairs"].sum()
    return out


def rotmat(phi, ssi, rot):
    R = np.zeros(9, np.float32)
    oracle_lib().orc_euler2rotmat(phi, ssi, rot, _p(R))
    return R


def rotmats(ang):
    """Rotation matrices (n, 9) of n angle triples (phi, ssi, rot)."""
    L = oracle_lib()
    a = np.ascontiguousarray(ang, np.float32).reshape(-1, 3)
    R = np.zeros((a.shape[0], 9), np.float32)
    L.orc_rotmats.restype = None
    L.orc_rotmats.argtypes = [P, C.c_size_t, P]
    L.orc_rotmats(_p(a), a.shape[0], _p(R))
    return R


def torquemat(phi, ssi, rot):
    t = np.zeros(27, np.float32)
    oracle_lib().orc_euler2torquemat(phi, ssi, rot, _p(t))
    return t


def ratio_lut(plateau):
    L = oracle_lib()
    out = np.zeros(L.orc_ratio_size(plateau), np.float32)
    L.orc_ratio_lut(plateau, _p(out))
    return out


# ------------------------------------------------------------- reference API

def ref_score_rigid(wl, dofs=None):
    L = ref_lib()
    d = np.ascontiguousarray(wl.dofs if dofs is None else dofs, np.float32)
    res = np.zeros((d.shape[0], 7), np.float32)
    rc = L.ref_score_rigid(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                           C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                           C.byref(nl_c(wl.rec_nl)), C.byref(table_c(wl.params)), _p(d),
                           d.shape[0], _p(res))
    assert rc == 0
    return res


def ref_beads_rigid(wl, dof):
    L = ref_lib()
    n = wl.lig.n
    xyz = np.zeros(3 * n, np.float32)
    pot = np.zeros(4 * n, np.float32)
    L.ref_beads_rigid(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                      C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                      C.byref(nl_c(wl.rec_nl)), C.byref(table_c(wl.params)),
                      _p(np.ascontiguousarray(dof, np.float32)), _p(xyz), _p(pot))
    return xyz, pot


def ref_cells_rigid(wl, dof):
    """The reference's own grid cells (n, 9) and plateau pair counts (n,) of
    one rigid pose (layout of beads()["cell"] / ["nl_pairs"])."""
    L = ref_lib()
    n = wl.lig.n
    cell = np.full(9 * n, -1, np.int32)
    npairs = np.zeros(n, np.uint32)
    L.ref_cells_rigid(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                      C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                      C.byref(nl_c(wl.rec_nl)), _p(np.ascontiguousarray(dof, np.float32)),
                      _p(cell), _p(npairs))
    return cell.reshape(n, 9), npairs


def ref_score_modes(wl, dofs=None):
    L = ref_lib()
    d = np.ascontiguousarray(wl.dofs if dofs is None else dofs, np.float32)
    res = np.zeros((d.shape[0], 27), np.float32)
    rec_pot = np.zeros(3 * wl.rec.n, np.float32)
    lig_pot = np.zeros(4 * wl.lig.n, np.float32)
    rc = L.ref_score_modes(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                           C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                           C.byref(nl_c(wl.rec_nl)), C.byref(igrid_c(wl.lig_inner)),
                           C.byref(igrid_c(wl.lig_outer)), C.byref(nl_c(wl.lig_nl)),
                           C.byref(table_c(wl.params)), _p(d), d.shape[0], _p(res),
                           _p(rec_pot), _p(lig_pot))
    assert rc == 0
    return res, rec_pot, lig_pot


def ref_pdb(path, cap=100000):
    """Reduced PDB through the reference's reader: (xyz SoA, type, charge)."""
    L = ref_lib()
    n = C.c_uint(0)
    xyz = np.zeros(3 * cap, np.float32)
    t = np.zeros(cap, np.uint32)
    q = np.zeros(cap, np.float32)
    if L.ref_pdb(os.fsencode(path), cap, C.byref(n), _p(xyz), _p(t), _p(q)) != 0:
        raise RuntimeError("capacity")
    N = n.value
    return xyz[:3 * N].copy(), t[:N].copy(), q[:N].copy()


def ref_modes(pdb, modefile, M, n):
    """Modes ([x: n*M][y][z]) and force constants through readHMMode."""
    L = ref_lib()
    modes = np.zeros(3 * n * M, np.float32)
    mf = np.zeros(M, np.float32)
    L.ref_modes(os.fsencode(pdb), os.fsencode(modefile), M, _p(modes), _p(mf))
    return modes, mf


def ref_params_file(path, cap=10000):
    L = ref_lib()
    T = C.c_uint(0)
    arrs = [np.zeros(cap, np.float32), np.zeros(cap, np.float32), np.zeros(cap, np.int32),
            np.zeros(cap, np.float32), np.zeros(cap, np.float32)]
    if L.ref_params_file(os.fsencode(path), cap, C.byref(T), *[_p(a) for a in arrs]) != 0:
        raise RuntimeError("capacity")
    k = T.value * T.value
    return T.value, [a[:k].copy() for a in arrs]


def ref_dof(path, modes=False, Mr=0, Ml=0, auto_rec=None, auto_lig=None, cap=100000):
    L = ref_lib()
    out = np.zeros(cap * (26 if modes else 6), np.float32)
    n = C.c_size_t(0)
    ar = np.zeros(3, np.float32) if auto_rec is None else np.asarray(auto_rec, np.float32)
    al = np.zeros(3, np.float32) if auto_lig is None else np.asarray(auto_lig, np.float32)
    st = L.ref_dof(os.fsencode(path), 1 if modes else 0, Mr, Ml, _p(ar), _p(al), cap, C.byref(n),
                   _p(out))
    if st != 0:
        raise RuntimeError(f"reference DOF reader status {st}")
    return out[:n.value * (26 if modes else 6)].reshape(n.value, -1).copy()


def ref_format_results(res, modes=False, Mr=0, Ml=0):
    L = ref_lib()
    r = np.ascontiguousarray(res, np.float32)
    ln = L.ref_format_results(_p(r), r.shape[0], 1 if modes else 0, Mr, Ml, None, 0)
    buf = C.create_string_buffer(max(1, ln))
    L.ref_format_results(_p(r), r.shape[0], 1 if modes else 0, Mr, Ml, buf, ln)
    return buf.raw[:ln].decode()


def ref_alphabet(path, cap=1000):
    """A grid alphabet file through the reference's reader."""
    L = ref_lib()
    out = np.zeros(cap, np.uint32)
    n = C.c_uint(0)
    if L.ref_alphabet(os.fsencode(path), _p(out), cap, C.byref(n)) != 0:
        raise RuntimeError("alphabet larger than the buffer")
    return out[:n.value].copy()


def ref_score_files(dof, rec, lig, grid, par, alphabet, dielec="variable", epsilon=15.0,
                    cap=100000):
    """The reference's scoring app chain on files (Configurator_6D<float> +
    CPUEnergyService6D behind as::Server): (n, 7) results."""
    L = ref_lib()
    res = np.zeros((cap, 7), np.float32)
    n = C.c_size_t(0)
    args = [os.fsencode(x) for x in (dof, rec, lig, grid, par, alphabet, dielec)]
    if L.ref_score_files(*args, epsilon, _p(res), cap, C.byref(n)) != 0:
        raise RuntimeError("more poses than the result buffer")
    return res[:n.value].copy()


def ref_score_files_modes(dof, rec, lig, grid_rec, grid_lig, par, alphabet_rec, alphabet_lig,
                          modes_rec, modes_lig, num_modes, dielec="variable", epsilon=15.0,
                          cap=100000):
    """The reference's flexible scoring chain on files (Configurator_6D_Modes
    + CPUEnergyService6DModes behind as::Server): (n, 27) results."""
    L = ref_lib()
    res = np.zeros((cap, 27), np.float32)
    n = C.c_size_t(0)
    paths = [os.fsencode(x) for x in (dof, rec, lig, grid_rec, grid_lig, par, alphabet_rec,
                                      alphabet_lig, modes_rec, modes_lig)]
    if L.ref_score_files_modes(*paths, num_modes, os.fsencode(dielec), epsilon, _p(res), cap,
                               C.byref(n)) != 0:
        raise RuntimeError("more poses than the result buffer")
    return res[:n.value].copy()


def ref_read_gridfile(path):
    """An ATTRACT grid file through the reference's reader (valid files only:
    the reference exits on a malformed one).  Returns dims[11], geo[13]
    (see ref_harness.cpp), inner, outer, desc, list."""
    L = ref_lib()
    if L.ref_gridfile_load(os.fsencode(path)) != 0:
        raise RuntimeError("reference reader failed")
    dims = np.zeros(11, np.uint32)
    geo = np.zeros(13, np.float32)
    ne = C.c_uint64(0)
    L.ref_gridfile_info(_p(dims), _p(geo), C.byref(ne))
    inner = np.zeros(int(dims[0]) * int(dims[1]) * int(dims[2]) * int(dims[3]) * 4, np.float32)
    outer = np.zeros(int(dims[4]) * int(dims[5]) * int(dims[6]) * int(dims[7]) * 4, np.float32)
    desc = np.zeros(int(dims[8]) * int(dims[9]) * int(dims[10]) * 2, np.uint32)
    lst = np.zeros(max(1, ne.value), np.uint32)
    L.ref_gridfile_copy(_p(inner), _p(outer), _p(desc), _p(lst))
    L.ref_gridfile_free()
    return dims, geo, inner, outer, desc, lst[:ne.value]


def ref_score_mt(wl, dofs=None, nthreads=16):
    """The reference's CPU path on `nthreads` worker threads taking items of
    100 poses: (results, seconds of scoring alone)."""
    L = ref_lib()
    d = np.ascontiguousarray(wl.dofs if dofs is None else dofs, np.float32)
    n = d.shape[0]
    secs = C.c_double(0.0)
    if wl.modes:
        res = np.zeros((n, 27), np.float32)
        rc = L.ref_score_modes_mt(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                                  C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                                  C.byref(nl_c(wl.rec_nl)), C.byref(igrid_c(wl.lig_inner)),
                                  C.byref(igrid_c(wl.lig_outer)), C.byref(nl_c(wl.lig_nl)),
                                  C.byref(table_c(wl.params)), _p(d), n, _p(res), nthreads,
                                  C.byref(secs))
    else:
        res = np.zeros((n, 7), np.float32)
        rc = L.ref_score_rigid_mt(C.byref(body_c(wl.rec)), C.byref(body_c(wl.lig)),
                                  C.byref(igrid_c(wl.rec_inner)), C.byref(igrid_c(wl.rec_outer)),
                                  C.byref(nl_c(wl.rec_nl)), C.byref(table_c(wl.params)), _p(d), n,
                                  _p(res), nthreads, C.byref(secs))
    assert rc == 0
    return res, secs.value


def ref_read_params(p):
    L = ref_lib()
    T = p.T
    out = dict(rc=np.zeros(T * T, np.float32), ac=np.zeros(T * T, np.float32),
               ipon=np.zeros(T * T, np.int32), rmin2=np.zeros(T * T, np.float32),
               emin=np.zeros(T * T, np.float32))
    rc = L.ref_read_params(T, p.shape, _p(p.rbc), _p(p.abc), _p(p.ipon_raw), _p(out["rc"]),
                           _p(out["ac"]), _p(out["ipon"]), _p(out["rmin2"]), _p(out["emin"]))
    assert rc == 0
    return out


def ref_ratio_lut(plateau):
    L = ref_lib()
    size = L.ref_ratio_lut(plateau, None, 0)
    out = np.zeros(size, np.float32)
    L.ref_ratio_lut(plateau, _p(out), size)
    return out


def ref_rotmat(phi, ssi, rot):
    R = np.zeros(9, np.float32)
    ref_lib().ref_euler2rotmat(phi, ssi, rot, _p(R))
    return R


def ref_torquemat(phi, ssi, rot):
    t = np.zeros(27, np.float32)
    ref_lib().ref_euler2torquemat(phi, ssi, rot, _p(t))
    return t
