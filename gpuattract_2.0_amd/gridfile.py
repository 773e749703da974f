"""ctypes binding of the ATTRACT grid-file reader/writer (include/attract_io.h,
in lib/libattract_service.so)."""
import ctypes as C
import dataclasses
import os

import numpy as np

from . import hip

LIBPATH = os.path.join(hip.PKG, "lib", "libattract_service.so")
P = C.c_void_p

_lib = None


class GridFileInfo(C.Structure):
    _fields_ = [("gridspacing", C.c_double), ("gridextension", C.c_int),
                ("plateaudis", C.c_double), ("neighbourdis", C.c_double), ("natoms", C.c_int),
                ("pivot", C.c_double * 3), ("alphabet", C.c_ubyte * 99),
                ("alphabetsize", C.c_uint)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} missing: build it first (__graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        L.att_gridfile_read.argtypes = [C.c_char_p, C.POINTER(P)]
        L.att_gridfile_free.argtypes = [P]
        L.att_gridfile_free.restype = None
        L.att_gridfile_last_error.restype = C.c_char_p
        L.att_gridfile_get.argtypes = [P, C.POINTER(GridFileInfo), C.POINTER(hip.GridDesc),
                                       C.POINTER(hip.GridDesc), C.POINTER(hip.NLDesc)]
        L.att_gridfile_write.argtypes = [C.c_char_p, C.POINTER(GridFileInfo),
                                         C.POINTER(hip.GridDesc), C.POINTER(hip.GridDesc),
                                         C.POINTER(hip.NLDesc), P]
        _lib = L
    return _lib


@dataclasses.dataclass
class GridFile:
    info: GridFileInfo
    inner: tuple   # (nx, ny, nz, numGrids, ox, oy, oz, spacing)
    outer: tuple
    nl: tuple      # (nx, ny, nz, ox, oy, oz, spacing, plateau)
    inner_data: np.ndarray
    outer_data: np.ndarray
    desc: np.ndarray
    list: np.ndarray


def _chk(st):
    if st != hip.ATT_OK:
        raise hip.AttError(st, lib().att_gridfile_last_error().decode())


def read(path):
    L = lib()
    h = P()
    _chk(L.att_gridfile_read(os.fsencode(path), C.byref(h)))
    try:
        info, gi, go, nd = GridFileInfo(), hip.GridDesc(), hip.GridDesc(), hip.NLDesc()
        _chk(L.att_gridfile_get(h, C.byref(info), C.byref(gi), C.byref(go), C.byref(nd)))

        def arr(ptr, n, t):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(t)), (n,)).copy() if n else \
                np.zeros(0, np.float32 if t is C.c_float else np.uint32)

        ni = gi.nx * gi.ny * gi.nz * gi.numGrids * 4
        no = go.nx * go.ny * go.nz * go.numGrids * 4
        return GridFile(info,
                        (gi.nx, gi.ny, gi.nz, gi.numGrids, gi.ox, gi.oy, gi.oz, gi.spacing),
                        (go.nx, go.ny, go.nz, go.numGrids, go.ox, go.oy, go.oz, go.spacing),
                        (nd.nx, nd.ny, nd.nz, nd.ox, nd.oy, nd.oz, nd.spacing, nd.plateau),
                        arr(gi.data, ni, C.c_float), arr(go.data, no, C.c_float),
                        arr(nd.desc, nd.nx * nd.ny * nd.nz * 2, C.c_uint32),
                        arr(nd.list, nd.numEl, C.c_uint32))
    finally:
        L.att_gridfile_free(h)


def make_info(gridspacing, gridextension, plateaudis, neighbourdis, types, natoms=0,
              pivot=(0.0, 0.0, 0.0)):
    """Header for a file whose grids cover the 1-based types in `types`."""
    info = GridFileInfo()
    info.gridspacing, info.gridextension = gridspacing, gridextension
    info.plateaudis, info.neighbourdis, info.natoms = plateaudis, neighbourdis, natoms
    for k in range(3):
        info.pivot[k] = pivot[k]
    for t in types:
        info.alphabet[t - 1] = 1
    return info


def write(path, info, inner, outer, nl, rec_xyz=None):
    """inner/outer: workload IGrid; nl: workload NLGrid (its geometry must be
    the inner grid's)."""
    gi, go, nd = hip.grid_desc(inner), hip.grid_desc(outer), hip.nl_desc(nl)
    xyz = None if rec_xyz is None else np.ascontiguousarray(rec_xyz, np.float32)
    _chk(lib().att_gridfile_write(os.fsencode(path), C.byref(info), C.byref(gi), C.byref(go),
                                  C.byref(nd), hip._p(xyz)))
