"""Scoring from files, end to end, as the reference's scoring app test does
it (test/app/ScoringTest.cpp:31-103): a receptor and a ligand in reduced-PDB
files, an attract.par parameter file, the receptor's ATTRACT grid file with
its grid alphabet, and an ATTRACT DOF file with explicit pivots.

* This library: the files through its readers (include/attract_formats.h,
  include/attract_io.h), the ligand's types mapped with the alphabet, both
  bodies pivotized with the DOF file's pivots, the grid displaced by the
  receptor pivot, the DOFs converted to the receptor frame (the steps of
  service/Configurator_6D.tpp:31-104), then the drop-in as::Server with
  MI355EnergyService6D<float> on the GPU (two workers, items of 100 poses),
  and the result text through att_format_results.
* The reference: the same files through its own Configurator_6D<float>::init
  (readers, TypeMap, pivots, transformDOF_glob2rec, GridUnion::translate) and
  CPUEnergyService6D behind its as::Server, compiled in oracle/_ref, and its
  result printer.

Results bit for bit (energy, force, torque of every pose); result text byte
for byte.  The input files are written here from the seeded generator (the
reference's test resources hold no grid file, and their parameter file has
98 types for beads of type 99)."""
import ctypes as C
import os

import numpy as np
import pytest

from test_formats import write_dof, write_pdb
from test_gridfile import attract_geoms

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FELEC = 332.053986  # model/SimParam.h:39
EPSILON = 15.0      # the scoring app's default (cli/parser_constants.h:42), dielectric "variable"


def _write_params(path, p):
    T = p.T
    with open(path, "w") as f:
        f.write(f"{p.shape} {T} 0 0\n")
        for arr, fmt in ((p.rbc, "{:.6f}"), (p.abc, "{:.6f}"), (p.ipon_raw, "{:d}")):
            for i in range(T):
                f.write(" ".join(fmt.format(v) for v in arr[i * T:(i + 1) * T]) + "\n")


def _shifted(W, body, c):
    xyz = body.xyz.reshape(3, -1) + np.float32(c)[:, None]
    return W.Body(np.ascontiguousarray(xyz.astype(np.float32).ravel()), body.type.copy(),
                  body.mtype.copy(), body.charge.copy(), body.modes, body.mode_force)


def _files(W, d, seed=20261207, n_rec=400, n_lig=120, n_poses=1000):
    """Write the five input files into directory d; returns their paths."""
    from gpuattract_amd import formats as F, gridfile as G
    T = 12
    R_rec = W.scaled_radius(n_rec, W.DENSITY_REF)
    R_lig = W.scaled_radius(n_lig, W.LIG_REF)
    rec = W.make_body(seed, 0, n_rec, R_rec, T, 0.2, 0.0)
    lig = W.make_body(seed, 1, n_lig, R_lig, T, 0.2, 0.0)
    # the bodies away from the origin, so that pivots and the grid
    # displacement matter
    c_rec, c_lig = (31.5, -12.25, 18.0), (-7.75, 40.5, 3.25)
    paths = {k: os.path.join(d, v) for k, v in
             dict(rec="receptorr.pdb", lig="ligandr.pdb", par="attract.par",
                  grid="receptorgrid.grid", alphabet="receptorgrid.alphabet",
                  dof="docking.dat").items()}
    write_pdb(paths["rec"], _shifted(W, rec, c_rec))
    write_pdb(paths["lig"], _shifted(W, lig, c_lig))
    p = W.make_params(seed, T)
    _write_params(paths["par"], p)
    # the grid alphabet: the ligand's types, in ascending order
    alphabet = sorted(int(t) for t in set(lig.type.tolist()) if t > 0)
    with open(paths["alphabet"], "w") as f:
        f.write("".join(f"{t}\n" for t in alphabet))
    # the receptor's grids from the receptor as its file holds it (global frame)
    a = F.Body(paths["rec"]).arrays()
    recg = W.Body(a.xyz, a.type, a.type.copy(), a.charge, np.zeros(0, np.float32),
                  np.zeros(0, np.float32))
    spacing, ext, plateau, nd = 0.9, 8, 10.0, 10.8
    gi, go = attract_geoms(W, recg, spacing, ext, 6.0)
    inner = W.build_grid(gi, recg, p, False, plateau)
    outer = W.build_grid(go, recg, p, False, plateau)
    nl = W.build_nl(gi, recg, nd, plateau)
    sel = [0] + alphabet
    for g in (inner, outer):
        per = g.geom.nvox * 4
        g.data = np.concatenate([g.data[t * per:(t + 1) * per] for t in sel])
        g.ngrids = len(sel)
    info = G.make_info(spacing, ext, plateau, nd, alphabet, natoms=recg.n, pivot=c_rec)
    G.write(paths["grid"], info, inner, outer, nl, recg.xyz)
    # poses: ligand centres on a shell around the receptor centre, near contact
    d6 = W.make_poses(seed, n_poses, 0.95 * (R_rec + R_lig), 1.2 * (R_rec + R_lig))
    shift = np.float64(c_rec) - np.float64(c_lig)
    lig_rows = [np.concatenate([q[3:6], q[0:3] + shift]) for q in d6.astype(np.float64)]
    rec_rows = [np.zeros(6) for _ in range(n_poses)]
    header = (f"#pivot 1 {c_rec[0]} {c_rec[1]} {c_rec[2]}\n#pivot 2 {c_lig[0]} {c_lig[1]} {c_lig[2]}\n"
              "#centered receptor: false\n#centered ligands: false\n")
    write_dof(paths["dof"], rec_rows, lig_rows, header)
    return paths


def _ours(W, paths):
    """The files through this library's readers and the drop-in Server on
    the GPU: (n, 7) results."""
    from gpuattract_amd import formats as F, gridfile as G, hip
    L = C.CDLL(os.path.join(HERE, "cpp", "_bin", "libservice_driver.so"))
    L.drv_error.restype = C.c_char_p
    rec, lig = F.Body(paths["rec"]), F.Body(paths["lig"])
    lig.map_types(F.read_alphabet(paths["alphabet"]))
    dof = F.DofFile(paths["dof"])
    info = dof.info()
    assert not info.autoPivot and info.numPivots == 2
    piv = [np.array(info.pivots[k][:], np.float32) for k in range(2)]
    rec.pivotize(piv[0])
    lig.pivotize(piv[1])
    dofs = dof.ligand()
    T, shape, _, _, tab = F.read_params(paths["par"])
    g = G.read(paths["grid"])
    # GridUnion::translate by -(receptor pivot): origin + displacement in float
    pr = piv[0]

    def moved(o):
        return [float(np.float32(o[k]) + (-pr[k])) for k in range(3)]

    gi, go, gn = g.inner, g.outer, g.nl
    inner = W.IGrid(W.Geom(gi[0], gi[1], gi[2], *moved(gi[4:7]), gi[7]), gi[3], g.inner_data)
    outer = W.IGrid(W.Geom(go[0], go[1], go[2], *moved(go[4:7]), go[7]), go[3], g.outer_data)
    nlg = W.NLGrid(W.Geom(gn[0], gn[1], gn[2], *moved(gn[3:6]), gn[6]), gn[7], g.desc, g.list)
    rd, ld = hip.ProteinDesc(), hip.ProteinDesc()
    F._chk(F.lib().att_body_desc(rec.h, C.byref(rd)))
    F._chk(F.lib().att_body_desc(lig.h, C.byref(ld)))
    di, do, dn = hip.grid_desc(inner), hip.grid_desc(outer), hip.nl_desc(nlg)
    d = np.ascontiguousarray(dofs, np.float32)
    out = np.full((d.shape[0], 7), np.nan, np.float32)
    ffelec = float(np.float32(FELEC / EPSILON))
    st = L.drv_score_rigid(C.byref(rd), C.byref(ld), C.byref(di), C.byref(do), C.byref(dn),
                           hip._p(tab), T, shape, 1, C.c_float(EPSILON), C.c_float(ffelec),
                           hip._p(d), C.c_size_t(d.shape[0]), hip._p(out), 2, 100, 4)
    assert st == 0, L.drv_error().decode()
    return out


def test_scoring_from_files_against_reference(W, O, tmp_path):
    if not O.ref_available():
        pytest.skip("reference tree not built (oracle/_ref)")
    from gpuattract_amd import formats as F
    paths = _files(W, str(tmp_path))
    print("scoring from files: " + ", ".join(os.path.basename(v) for v in paths.values()),
          flush=True)
    ref = O.ref_score_files(paths["dof"], paths["rec"], paths["lig"], paths["grid"],
                            paths["par"], paths["alphabet"], "variable", EPSILON)
    got = _ours(W, paths)
    assert got.shape == ref.shape == (1000, 7)
    # the poses interact: most have a grid or pair term
    assert np.mean(ref[:, 0] != 0) > 0.9
    diff = np.flatnonzero(np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1))
    assert diff.size == 0, f"{diff.size} poses differ, first {diff[:5]}: {got[diff[:2]]} vs {ref[diff[:2]]}"
    text = F.format_results(got)
    assert text == O.ref_format_results(ref)
    print(f"{got.shape[0]} poses from files: results and result text equal the reference's; "
          f"{np.mean(ref[:, 0] < 0):.2f} with negative energy")
