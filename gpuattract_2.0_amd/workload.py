"""Seeded synthetic workloads in the reference's data model (numpy, host side).

Wraps libattract_gen.so (include/attract_gen.h).  The same arrays feed the HIP
library, the CPU oracle and the compiled-reference harness, so every consumer
scores bit-identical inputs.

Conventions (reference citations):
  body.xyz      float32 SoA [x: n][y: n][z: n]                   model/Protein.h:55-65
  body.modes    float32 [x: n*M][y: n*M][z: n*M], index i*M+m     model/Protein.h:88-98
  grid.data     float32 (ngrids, nz, ny, nx, 4) {fx,fy,fz,E},
                grid 0 = electrostatics                          model/IntrplGrid.h:185-198
  nl.desc       uint32 (nz*ny*nx, 2) {numEl, idx}; nl.list uint32 model/NLGrid.h:24-28
  params        T*T entries indexed T*(recType-1)+(ligType-1)    model/ParamTable.h:80-82
  dofs          rigid: (n, 6) {pos, ang};                         model/Types_6D.h:31-37
                modes: (n, 26) {pos, ang, modesRec[10], modesLig[10]}  model/Types_6D_Modes.h:33-40
"""
import ctypes as C
import dataclasses
import math
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")

FFELEC = 332.053986  # model/SimParam.h:39
MODES_MAX = 10       # model/Types_6D_Config.h:11
DENSITY_REF = (1000, 25.0)  # receptor: 1000 beads in a 25 A sphere
LIG_REF = (300, 15.0)       # ligand: 300 beads in a 15 A sphere


class _Geom(C.Structure):
    _fields_ = [("nx", C.c_uint), ("ny", C.c_uint), ("nz", C.c_uint),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("spacing", C.c_float)]


_gen = None


def gen_lib():
    global _gen
    if _gen is None:
        path = os.path.join(LIBDIR, "libattract_gen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run __graft_entry__.build() first")
        L = C.CDLL(path)
        P = C.c_void_p
        L.attg_uniform.restype = C.c_double
        L.attg_uniform.argtypes = [C.c_uint64, C.c_uint, C.c_uint64]
        L.attg_make_body.restype = C.c_int
        L.attg_make_body.argtypes = [C.c_uint64, C.c_uint, C.c_uint, C.c_double, C.c_double,
                                     C.c_uint, C.c_double, C.c_double, P, P, P]
        L.attg_make_raw_params.restype = None
        L.attg_make_raw_params.argtypes = [C.c_uint64, C.c_uint, C.c_double, C.c_double,
                                           C.c_double, C.c_double, C.c_double, P, P, P]
        L.attg_derive_params.restype = C.c_int
        L.attg_derive_params.argtypes = [C.c_uint, C.c_int, P, P, P, P, P, P, P, P]
        L.attg_inner_geom.restype = None
        L.attg_inner_geom.argtypes = [P, C.c_uint, C.c_float, C.c_float, C.POINTER(_Geom)]
        L.attg_cube_geom.restype = None
        L.attg_cube_geom.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                     C.POINTER(_Geom)]
        L.attg_build_grid.restype = C.c_int
        L.attg_build_grid.argtypes = [C.POINTER(_Geom), P, P, P, C.c_uint, C.c_uint, C.c_int,
                                      P, P, P, P, P, C.c_int, C.c_float, C.c_int, C.c_float,
                                      P, C.c_int]
        L.attg_build_basis.restype = C.c_int
        L.attg_build_basis.argtypes = [C.POINTER(_Geom), P, P, C.c_uint, C.c_uint, C.c_int, P,
                                       C.c_float, P, C.c_int]
        L.attg_build_nl.restype = C.c_int
        L.attg_build_nl.argtypes = [C.POINTER(_Geom), P, P, C.c_uint, C.c_float, P, P,
                                    C.POINTER(C.c_uint64), C.c_int]
        L.attg_make_poses.restype = None
        L.attg_make_poses.argtypes = [C.c_uint64, C.c_uint, C.c_uint64, C.c_uint64, C.c_double,
                                      C.c_double, C.c_float, C.c_float, C.c_float, P]
        L.attg_make_amplitudes.restype = None
        L.attg_make_amplitudes.argtypes = [C.c_uint64, C.c_uint, C.c_uint64, C.c_uint64,
                                           C.c_uint, C.c_double, P]
        L.attg_make_modes.restype = None
        L.attg_make_modes.argtypes = [C.c_uint64, C.c_uint, C.c_uint, C.c_uint, C.c_double,
                                      C.c_double, P, P]
        _gen = L
    return _gen


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


@dataclasses.dataclass
class Body:
    xyz: np.ndarray        # (3*n,) float32 SoA
    type: np.ndarray       # (n,) uint32, raw types
    mtype: np.ndarray      # (n,) uint32, grid-mapped types
    charge: np.ndarray     # (n,) float32
    modes: np.ndarray      # (3*n*M,) float32
    mode_force: np.ndarray  # (M,) float32

    @property
    def n(self):
        return int(self.type.shape[0])

    @property
    def M(self):
        return int(self.mode_force.shape[0])


@dataclasses.dataclass
class Geom:
    nx: int
    ny: int
    nz: int
    ox: float
    oy: float
    oz: float
    spacing: float

    def c(self):
        return _Geom(self.nx, self.ny, self.nz, self.ox, self.oy, self.oz, self.spacing)

    @property
    def nvox(self):
        return self.nx * self.ny * self.nz


@dataclasses.dataclass
class IGrid:
    geom: Geom
    ngrids: int
    data: np.ndarray  # (ngrids*nz*ny*nx*4,) float32


@dataclasses.dataclass
class NLGrid:
    geom: Geom
    plateau: float
    desc: np.ndarray  # (nvox*2,) uint32
    list: np.ndarray  # (numEl,) uint32


@dataclasses.dataclass
class Params:
    T: int
    shape: int
    dielec: int       # 0 constant, 1 variable
    epsilon: float
    ffelec: float
    rc: np.ndarray
    ac: np.ndarray
    ipon: np.ndarray
    rmin2: np.ndarray
    emin: np.ndarray
    rbc: np.ndarray   # raw attract.par values the table was derived from
    abc: np.ndarray
    ipon_raw: np.ndarray


@dataclasses.dataclass
class Workload:
    name: str
    rec: Body
    lig: Body
    params: Params
    rec_inner: IGrid
    rec_outer: IGrid
    rec_nl: NLGrid
    dofs: np.ndarray                 # (n, 6) or (n, 26) float32
    lig_inner: IGrid = None
    lig_outer: IGrid = None
    lig_nl: NLGrid = None
    rec_inner_basis: np.ndarray = None  # (2T, nz, ny, nx) flattened, for parameter gradients
    rec_outer_basis: np.ndarray = None
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def modes(self):
        return self.dofs.shape[1] == 26


def make_body(seed, entity, n, radius, num_types=12, charged_frac=0.2, type0_frac=0.0,
              min_spacing=3.0):
    L = gen_lib()
    xyz = np.zeros(3 * n, np.float32)
    typ = np.zeros(n, np.uint32)
    chg = np.zeros(n, np.float32)
    rc = L.attg_make_body(seed, entity, n, radius, min_spacing, num_types, charged_frac,
                          type0_frac, _p(xyz), _p(typ), _p(chg))
    if rc != 0:
        raise ValueError(f"cannot place {n} beads in a {radius} A sphere")
    return Body(xyz, typ, typ.copy(), chg, np.zeros(0, np.float32), np.zeros(0, np.float32))


def add_modes(body, seed, entity, M, h_lo=0.05, h_hi=0.2):
    L = gen_lib()
    modes = np.zeros(3 * body.n * M, np.float32)
    fc = np.zeros(M, np.float32)
    if M > 0:
        L.attg_make_modes(seed, entity, body.n, M, h_lo, h_hi, _p(modes), _p(fc))
    body.modes, body.mode_force = modes, fc
    return body


def make_params(seed, T=12, rmin=(3.0, 6.0), emin=(0.1, 1.0), repulsive_frac=0.2, shape=8,
                dielec=1, epsilon=15.0):
    L = gen_lib()
    rbc = np.zeros(T * T, np.float32)
    abc = np.zeros(T * T, np.float32)
    ipr = np.zeros(T * T, np.int32)
    L.attg_make_raw_params(seed, T, rmin[0], rmin[1], emin[0], emin[1], repulsive_frac,
                           _p(rbc), _p(abc), _p(ipr))
    out = [np.zeros(T * T, np.float32) for _ in range(2)]
    ipon = np.zeros(T * T, np.int32)
    rmin2 = np.zeros(T * T, np.float32)
    em = np.zeros(T * T, np.float32)
    if L.attg_derive_params(T, shape, _p(rbc), _p(abc), _p(ipr), _p(out[0]), _p(out[1]),
                            _p(ipon), _p(rmin2), _p(em)) != 0:
        raise ValueError("bad potential shape")
    ffelec = float(np.float32(FFELEC / epsilon))  # SimParam.ffelec is REAL
    return Params(T, shape, dielec, epsilon, ffelec, out[0], out[1], ipon, rmin2, em,
                  rbc, abc, ipr)


def inner_geom(body, spacing=0.875, margin=10.0):
    g = _Geom()
    gen_lib().attg_inner_geom(_p(body.xyz), body.n, spacing, margin, C.byref(g))
    return Geom(g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g.spacing)


def cube_geom(center, halfwidth, spacing):
    g = _Geom()
    gen_lib().attg_cube_geom(center[0], center[1], center[2], halfwidth, spacing, C.byref(g))
    return Geom(g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g.spacing)


def build_grid(geom, src, params, src_is_ligand, plateau=10.0):
    T = params.T
    data = np.zeros((T + 1) * geom.nvox * 4, np.float32)
    g = geom.c()
    rc = gen_lib().attg_build_grid(C.byref(g), _p(src.xyz), _p(src.type), _p(src.charge), src.n,
                                   T, params.shape, _p(params.rc), _p(params.ac),
                                   _p(params.ipon), _p(params.rmin2), _p(params.emin),
                                   1 if src_is_ligand else 0, plateau, params.dielec,
                                   params.ffelec, _p(data), _threads())
    if rc != 0:
        raise RuntimeError("grid build failed")
    return IGrid(geom, T + 1, data)


def build_basis(geom, src, params, plateau=10.0):
    """Basis grids (2*T, nz, ny, nx) of the receptor type sums S_a, S6_a."""
    T = params.T
    out = np.zeros(2 * T * geom.nvox, np.float32)
    g = geom.c()
    rc = gen_lib().attg_build_basis(C.byref(g), _p(src.xyz), _p(src.type), src.n, T,
                                    params.shape, _p(params.rmin2), plateau, _p(out), _threads())
    if rc != 0:
        raise ValueError("basis grids need rmin^2 <= plateau^2 for every pair")
    return out


def build_nl(geom, src, neighbourdis=10.8, plateau=10.0):
    L = gen_lib()
    desc = np.zeros(geom.nvox * 2, np.uint32)
    tot = C.c_uint64(0)
    g = geom.c()
    if L.attg_build_nl(C.byref(g), _p(src.xyz), _p(src.type), src.n, neighbourdis, _p(desc),
                       None, C.byref(tot), _threads()) != 0:
        raise RuntimeError("neighbour list too large")
    lst = np.zeros(max(1, tot.value), np.uint32)
    L.attg_build_nl(C.byref(g), _p(src.xyz), _p(src.type), src.n, neighbourdis, _p(desc),
                    _p(lst), C.byref(tot), _threads())
    return NLGrid(geom, plateau, desc, lst[:tot.value] if tot.value else lst[:0])


def make_poses(seed, n, r_lo, r_hi, first=0, center=(0.0, 0.0, 0.0), entity=4):
    dof = np.zeros((n, 6), np.float32)
    gen_lib().attg_make_poses(seed, entity, first, n, r_lo, r_hi, center[0], center[1],
                              center[2], _p(dof))
    return dof


def make_amplitudes(seed, n, M, sigma=1.0, first=0, entity=5):
    a = np.zeros((n, M), np.float32)
    if M:
        gen_lib().attg_make_amplitudes(seed, entity, first, n, M, sigma, _p(a))
    return a


def scaled_radius(n, ref):
    return ref[1] * (n / ref[0]) ** (1.0 / 3.0)


def lig_grid_geoms(lig, R_rec, r_hi, spacing=0.875, margin=10.0, outer_spacing=2.6):
    """Inner and outer geometry of a ligand's grids (the receptor beads are
    looked up in them): the inner box spans the ligand plus the margin, the
    outer cube every receptor bead of a pose whose translation is at most r_hi."""
    return (inner_geom(lig, spacing, margin),
            cube_geom((0.0, 0.0, 0.0), r_hi + R_rec + 2.0, outer_spacing))


def pair_lig_grid_geoms(wl):
    """lig_grid_geoms of a pair built by make_workload with its default shell."""
    R_rec = scaled_radius(wl.rec.n, DENSITY_REF)
    R_lig = scaled_radius(wl.lig.n, LIG_REF)
    return lig_grid_geoms(wl.lig, R_rec, R_rec + R_lig)


def make_workload(name, seed, n_rec, n_lig, n_poses, M_rec=0, M_lig=0, *, num_types=12,
                  spacing=0.875, margin=10.0, outer_spacing=2.6, plateau=10.0,
                  neighbourdis=10.8, charged_frac=0.2, type0_frac=0.0, shell=None,
                  dielec=1, shape=8, with_lig_grid=None, repulsive_frac=0.2,
                  with_basis=False, host_grids=True):
    """A receptor-ligand pair with grids and poses, as described by the
    benchmark configurations (receptor density of 1000 beads per 25 A sphere,
    ligand 300 per 15 A sphere, 3 A minimum spacing, 12 types, 20% charged).
    host_grids=False leaves the receptor's grids and neighbour list unbuilt
    (geometry only, data None): Context.upload_workload then builds them on
    the device, bit-identical to the host generator (fill_grids_from_device
    fetches them for host-side checks)."""
    R_rec = scaled_radius(n_rec, DENSITY_REF)
    R_lig = scaled_radius(n_lig, LIG_REF)
    rec = make_body(seed, 0, n_rec, R_rec, num_types, charged_frac, type0_frac)
    lig = make_body(seed, 1, n_lig, R_lig, num_types, charged_frac, type0_frac)
    add_modes(rec, seed, 50, M_rec)
    add_modes(lig, seed, 60, M_lig)
    params = make_params(seed, num_types, shape=shape, dielec=dielec,
                         repulsive_frac=repulsive_frac)
    if shell is None:
        shell = (0.75 * (R_rec + R_lig), 1.0 * (R_rec + R_lig))
    r_lo, r_hi = shell
    gi = inner_geom(rec, spacing, margin)
    go = cube_geom((0.0, 0.0, 0.0), r_hi + R_lig + 2.0, outer_spacing)
    if host_grids:
        rec_inner = build_grid(gi, rec, params, False, plateau)
        rec_outer = build_grid(go, rec, params, False, plateau)
        rec_nl = build_nl(gi, rec, neighbourdis, plateau)
    else:
        rec_inner, rec_outer = IGrid(gi, num_types + 1, None), IGrid(go, num_types + 1, None)
        rec_nl = NLGrid(gi, plateau, None, None)
    modes = (M_rec > 0 or M_lig > 0) if with_lig_grid is None else with_lig_grid
    dofs6 = make_poses(seed, n_poses, r_lo, r_hi)
    wl = Workload(name, rec, lig, params, rec_inner, rec_outer, rec_nl, dofs6)
    if modes:
        gli, glo = lig_grid_geoms(lig, R_rec, r_hi, spacing, margin, outer_spacing)
        wl.lig_inner = build_grid(gli, lig, params, True, plateau)
        wl.lig_outer = build_grid(glo, lig, params, True, plateau)
        wl.lig_nl = build_nl(gli, lig, neighbourdis, plateau)
    if modes or M_rec > 0 or M_lig > 0:
        dof = np.zeros((n_poses, 26), np.float32)
        dof[:, :6] = dofs6
        dof[:, 6:6 + M_rec] = make_amplitudes(seed, n_poses, M_rec, entity=5)
        dof[:, 16:16 + M_lig] = make_amplitudes(seed, n_poses, M_lig, entity=6)
        wl.dofs = dof
    if with_basis:
        wl.rec_inner_basis = build_basis(gi, rec, params, plateau)
        wl.rec_outer_basis = build_basis(go, rec, params, plateau)
    wl.meta = dict(seed=seed, n_rec=n_rec, n_lig=n_lig, n_poses=n_poses, M_rec=M_rec,
                   M_lig=M_lig, R_rec=R_rec, R_lig=R_lig, shell=shell, spacing=spacing,
                   outer_spacing=outer_spacing, plateau=plateau, neighbourdis=neighbourdis)
    return wl


def fill_grids_from_device(ctx, wl):
    """The receptor grids and neighbour list of a workload made with
    host_grids=False, generated on the device (att_generate_grid /
    att_generate_nl, the host generator's values bit for bit) into host
    arrays, for the CPU oracle."""
    m = wl.meta
    if wl.rec_inner.data is None:
        wl.rec_inner = IGrid(wl.rec_inner.geom, wl.params.T + 1,
                             ctx.generate_grid(wl.rec, wl.params, wl.rec_inner.geom, m["plateau"]))
        wl.rec_outer = IGrid(wl.rec_outer.geom, wl.params.T + 1,
                             ctx.generate_grid(wl.rec, wl.params, wl.rec_outer.geom, m["plateau"]))
        desc, lst = ctx.generate_nl(wl.rec, wl.rec_nl.geom, m["neighbourdis"])
        wl.rec_nl = NLGrid(wl.rec_nl.geom, m["plateau"], desc, lst)
    return wl


# The benchmark configurations (master seed 20261202 + i for configuration i).
def config_rigid_1000_300(n_poses=10000):
    return make_workload("rigid 1000/300", 20261202, 1000, 300, n_poses)


def config_modes_3000_800(n_poses=100000):
    return make_workload("modes 3000/800 5+5", 20261203, 3000, 800, n_poses, 5, 5)


def make_batch(seed, n_pairs, rec_range=(500, 5000), lig_range=(100, 1500), modes_range=(0, 10),
               poses_range=(1000, 50000), progress=None, **kw):
    """Pairs of uneven sizes for batched scoring: receptor and ligand bead
    counts and pose counts log-uniform in their ranges, each body's number of
    modes uniform in modes_range; pair k uses the master seed plus 1000 * (k+1)
    for its own bodies, table, grids and poses.  Ligand grids are not built
    (the adjoint does not use them); DOFs carry the modes layout when either
    body has modes."""
    def logu(entity, k, lo, hi):
        u = gen_lib().attg_uniform(seed, entity, k)
        return int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))

    def uni(entity, k, lo, hi):
        u = gen_lib().attg_uniform(seed, entity, k)
        return min(hi, lo + int(u * (hi - lo + 1)))

    pairs = []
    for k in range(n_pairs):
        n_rec = logu(60, 4 * k, *rec_range)
        n_lig = logu(60, 4 * k + 1, *lig_range)
        n_poses = logu(60, 4 * k + 2, *poses_range)
        M_rec = uni(61, 2 * k, *modes_range)
        M_lig = uni(61, 2 * k + 1, *modes_range)
        pairs.append(make_workload(f"pair {k}: {n_rec}/{n_lig} beads, {M_rec}+{M_lig} modes, "
                                   f"{n_poses} poses", seed + 1000 * (k + 1), n_rec, n_lig,
                                   n_poses, M_rec, M_lig, with_lig_grid=False, **kw))
        if progress:
            progress(k + 1, pairs[-1])
    return pairs


@dataclasses.dataclass
class ServeReceptor:
    body: Body
    radius: float
    inner: Geom
    outer: Geom


@dataclasses.dataclass
class ServeJob:
    receptor: int          # index into the stream's receptors
    lig: Body
    dofs: np.ndarray       # (n, 6) rigid or (n, 26) with modes
    max_fun_eval: int


def make_serve_stream(seed, n_jobs, reuse=0.3, rec_range=(500, 8000), lig_range=(100, 2000),
                      modes_range=(0, 10), poses_range=(500, 20000), max_fun_eval=200,
                      outer_spacing=2.6):
    """Docking jobs of a serving stream: each job docks a new ligand (beads
    log-uniform in lig_range, modes uniform in modes_range) with poses
    log-uniform in poses_range against a receptor that is, with probability
    `reuse`, one of the earlier jobs' receptors and otherwise a new one (beads
    log-uniform in rec_range, its own modes).  One force field for the stream.
    Grids are not built: each receptor carries its inner geometry (bounding box
    + 10 A) and an outer cube that covers the largest ligand on the pose shell.
    Returns (params, receptors, jobs)."""
    L = gen_lib()

    def u(entity, k):
        return L.attg_uniform(seed, entity, k)

    def logu(entity, k, lo, hi):
        return int(round(math.exp(math.log(lo) + u(entity, k) * (math.log(hi) - math.log(lo)))))

    def uni(entity, k, lo, hi):
        return min(hi, lo + int(u(entity, k) * (hi - lo + 1)))

    params = make_params(seed)
    r_lig_max = scaled_radius(lig_range[1], LIG_REF)
    recs, jobs = [], []
    for k in range(n_jobs):
        if recs and u(70, 2 * k) < reuse:
            ri = min(int(u(70, 2 * k + 1) * len(recs)), len(recs) - 1)
        else:
            ri = len(recs)
            n = logu(71, ri, *rec_range)
            R = scaled_radius(n, DENSITY_REF)
            body = make_body(seed + 7919 * (ri + 1), 0, n, R)
            add_modes(body, seed + 7919 * (ri + 1), 50, uni(72, ri, *modes_range))
            recs.append(ServeReceptor(body, R, inner_geom(body),
                                      cube_geom((0.0, 0.0, 0.0), R + 2 * r_lig_max + 2.0,
                                                outer_spacing)))
        rec = recs[ri]
        n_l = logu(73, k, *lig_range)
        R_l = scaled_radius(n_l, LIG_REF)
        lig = make_body(seed + 104729 * (k + 1), 1, n_l, R_l)
        M_l = uni(74, k, *modes_range)
        add_modes(lig, seed + 104729 * (k + 1), 60, M_l)
        n_p = logu(75, k, *poses_range)
        d6 = make_poses(seed + 104729 * (k + 1), n_p, 0.75 * (rec.radius + R_l),
                        1.0 * (rec.radius + R_l))
        M_r = rec.body.M
        if M_r or M_l:
            d = np.zeros((n_p, 26), np.float32)
            d[:, :6] = d6
            d[:, 6:6 + M_r] = make_amplitudes(seed + 104729 * (k + 1), n_p, M_r, entity=5)
            d[:, 16:16 + M_l] = make_amplitudes(seed + 104729 * (k + 1), n_p, M_l, entity=6)
        else:
            d = d6
        jobs.append(ServeJob(ri, lig, d, max_fun_eval))
    return params, recs, jobs


def checksum(*arrays):
    """Order-sensitive 64-bit checksum of raw array bytes (FNV-1a over 8-byte words)."""
    h = np.uint64(0xCBF29CE484222325)
    prime = np.uint64(0x100000001B3)
    with np.errstate(over="ignore"):
        for a in arrays:
            b = np.ascontiguousarray(a).view(np.uint8).ravel()
            pad = (-len(b)) % 8
            if pad:
                b = np.concatenate([b, np.zeros(pad, np.uint8)])
            w = b.view(np.uint64)
            # fold in blocks to keep it vectorised but order-dependent
            acc = np.uint64(0)
            for blk in np.array_split(w, max(1, len(w) // 65536)):
                s = np.bitwise_xor.reduce(blk * (np.arange(len(blk), dtype=np.uint64) * prime + np.uint64(1)))
                acc = (acc ^ s) * prime + np.uint64(len(blk))
            h = (h ^ acc) * prime
    return int(h)
