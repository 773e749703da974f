"""ctypes binding of the reference's file formats (include/attract_formats.h, in
lib/libattract_service.so): reduced PDB bodies, mode files, attract.par, DOF
files with their receptor-frame conversion, and the result text."""
import ctypes as C
import dataclasses
import os

import numpy as np

from . import hip

LIBPATH = os.path.join(hip.PKG, "lib", "libattract_service.so")
P = C.c_void_p

_lib = None


class DofInfo(C.Structure):
    _fields_ = [("numMolecules", C.c_uint), ("numPoses", C.c_size_t), ("autoPivot", C.c_int),
                ("numPivots", C.c_uint), ("pivots", (C.c_float * 3) * 2),
                ("centeredReceptor", C.c_int), ("centeredLigands", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} missing: build it first (__graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        L.att_formats_last_error.restype = C.c_char_p
        L.att_body_read_pdb.argtypes = [C.c_char_p, C.POINTER(P)]
        L.att_body_read_modes.argtypes = [P, C.c_char_p, C.c_uint]
        L.att_body_pivotize.argtypes = [P, P, P]
        L.att_body_map_types.argtypes = [P, P, C.c_uint]
        L.att_body_desc.argtypes = [P, C.POINTER(hip.ProteinDesc)]
        L.att_body_free.argtypes = [P]
        L.att_body_free.restype = None
        L.att_read_params.argtypes = [C.c_char_p, P, C.c_uint, C.POINTER(C.c_uint),
                                      C.POINTER(C.c_int), C.POINTER(C.c_float),
                                      C.POINTER(C.c_float)]
        L.att_read_alphabet.argtypes = [C.c_char_p, P, C.c_uint, C.POINTER(C.c_uint)]
        L.att_dof_read.argtypes = [C.c_char_p, C.POINTER(P)]
        L.att_dof_get_info.argtypes = [P, C.POINTER(DofInfo)]
        L.att_dof_values.argtypes = [P, C.c_uint, C.c_size_t, P, C.POINTER(C.c_uint)]
        L.att_dof_ligand.argtypes = [P, P, P, C.c_int, C.c_uint, C.c_uint, P]
        L.att_dof_free.argtypes = [P]
        L.att_dof_free.restype = None
        L.att_format_results.argtypes = [P, C.c_size_t, C.c_int, C.c_uint, C.c_uint, P, C.c_size_t,
                                         C.POINTER(C.c_size_t)]
        _lib = L
    return _lib


def _chk(st):
    if st != hip.ATT_OK:
        raise hip.AttError(st, lib().att_formats_last_error().decode())


@dataclasses.dataclass
class BodyArrays:
    n: int
    M: int
    xyz: np.ndarray
    type: np.ndarray
    mapped: np.ndarray
    charge: np.ndarray
    modes: np.ndarray
    mode_force: np.ndarray


class Body:
    """A body read from a reduced PDB file (att_body)."""

    def __init__(self, path):
        self.h = P()
        _chk(lib().att_body_read_pdb(os.fsencode(path), C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().att_body_free(self.h)
            self.h = None

    def read_modes(self, path, num_modes):
        _chk(lib().att_body_read_modes(self.h, os.fsencode(path), num_modes))

    def pivotize(self, pivot=None):
        """Pivot given, or the centroid when None; returns the pivot used."""
        out = np.zeros(3, np.float32)
        pv = None if pivot is None else np.asarray(pivot, np.float32)
        _chk(lib().att_body_pivotize(self.h, hip._p(pv), hip._p(out)))
        return out

    def map_types(self, alphabet=()):
        a = np.asarray(alphabet, np.uint32)
        _chk(lib().att_body_map_types(self.h, hip._p(a) if len(a) else None, len(a)))

    def arrays(self):
        d = hip.ProteinDesc()
        _chk(lib().att_body_desc(self.h, C.byref(d)))
        n, M = d.n, d.numModes

        def take(ptr, count, t, dt):
            if not count or not ptr:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(t)), (count,)).copy()

        return BodyArrays(n, M, take(d.xyz, 3 * n, C.c_float, np.float32),
                          take(d.type, n, C.c_uint, np.uint32),
                          take(d.mappedType, n, C.c_uint, np.uint32),
                          take(d.charge, n, C.c_float, np.float32),
                          take(d.modes, 3 * n * M, C.c_float, np.float32),
                          take(d.modeForce, M, C.c_float, np.float32))


def read_params(path):
    """(numTypes, potShape, swiOn, swiOff, table) of an attract.par file."""
    T, shape, on, off = C.c_uint(0), C.c_int(0), C.c_float(0), C.c_float(0)
    _chk(lib().att_read_params(os.fsencode(path), None, 0, C.byref(T), C.byref(shape), C.byref(on),
                               C.byref(off)))
    tab = np.zeros(T.value * T.value, hip.PAIR_DTYPE)
    _chk(lib().att_read_params(os.fsencode(path), hip._p(tab), len(tab), C.byref(T), C.byref(shape),
                               C.byref(on), C.byref(off)))
    return T.value, shape.value, on.value, off.value, tab


def read_alphabet(path):
    """The bead types of a grid alphabet file, in file order (uint32)."""
    n = C.c_uint(0)
    _chk(lib().att_read_alphabet(os.fsencode(path), None, 0, C.byref(n)))
    out = np.zeros(n.value, np.uint32)
    _chk(lib().att_read_alphabet(os.fsencode(path), hip._p(out), len(out), C.byref(n)))
    return out


class DofFile:
    """A DOF file (att_dof_file)."""

    def __init__(self, path):
        self.h = P()
        _chk(lib().att_dof_read(os.fsencode(path), C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().att_dof_free(self.h)
            self.h = None

    def info(self):
        i = DofInfo()
        _chk(lib().att_dof_get_info(self.h, C.byref(i)))
        return i

    def values(self, molecule, pose):
        """The values of one line as read (float64): phi, ssi, rot, x, y, z,
        then the mode amplitudes."""
        v = np.zeros(16, np.float64)
        n = C.c_uint(0)
        _chk(lib().att_dof_values(self.h, molecule, pose, hip._p(v), C.byref(n)))
        return v[:n.value].copy()

    def ligand(self, pivot_rec=None, pivot_lig=None, modes=False, num_modes_rec=0,
               num_modes_lig=0):
        """Ligand DOFs in the receptor frame: (n, 6) or (n, 26) float32."""
        n = self.info().numPoses
        out = np.zeros((n, 26 if modes else 6), np.float32)
        pr = None if pivot_rec is None else np.asarray(pivot_rec, np.float32)
        pl = None if pivot_lig is None else np.asarray(pivot_lig, np.float32)
        _chk(lib().att_dof_ligand(self.h, hip._p(pr), hip._p(pl), 1 if modes else 0,
                                  num_modes_rec, num_modes_lig, hip._p(out)))
        return out


def format_results(results, modes=False, num_modes_rec=0, num_modes_lig=0):
    """The apps' result text for (n, 7) or (n, 27) results."""
    r = np.ascontiguousarray(results, np.float32)
    n = r.shape[0]
    ln = C.c_size_t(0)
    _chk(lib().att_format_results(hip._p(r), n, 1 if modes else 0, num_modes_rec, num_modes_lig,
                                  None, 0, C.byref(ln)))
    buf = C.create_string_buffer(ln.value)
    _chk(lib().att_format_results(hip._p(r), n, 1 if modes else 0, num_modes_rec, num_modes_lig,
                                  buf, ln.value, C.byref(ln)))
    return buf.raw[:ln.value].decode()
