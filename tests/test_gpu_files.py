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
98 types for beads of type 99).

Flexible docking the same way: both bodies' mode files, a ligand grid file
with its alphabet, mode amplitudes in the DOF file, through
MI355EnergyService6DModes here; on the reference's side the input steps of
service/Configurator_6D_Modes.tpp, each through the reference's own function,
and its modes model chain (the order of CPUEnergyService6DModes.tpp:121-327;
the modes service object itself fails its WorkerBuffer index assertion, and
the configurator converts the DOFs before it sets the mode counts they are
converted with, so the harness sets those first).
Energy, force, torque and ligand mode gradients bit for bit; the receptor
mode gradients (the reaction projected on the modes per pair, where the
reference sums it per receptor bead first) within 1e-5 of their norm; the
result text byte for byte once those columns are the reference's."""
import ctypes as C
import os

import numpy as np
import pytest

from test_formats import write_dof, write_modes, write_pdb
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


def _grid_file(W, G, path, body_path, p, probe_types, pivot):
    """An ATTRACT grid file of the body in body_path (its file coordinates)
    for the given probe types."""
    from gpuattract_amd import formats as F
    a = F.Body(body_path).arrays()
    src = W.Body(a.xyz, a.type, a.type.copy(), a.charge, np.zeros(0, np.float32),
                 np.zeros(0, np.float32))
    spacing, ext, plateau, nd = 0.9, 8, 10.0, 10.8
    gi, go = attract_geoms(W, src, spacing, ext, 6.0)
    inner = W.build_grid(gi, src, p, False, plateau)
    outer = W.build_grid(go, src, p, False, plateau)
    nl = W.build_nl(gi, src, nd, plateau)
    sel = [0] + list(probe_types)
    for g in (inner, outer):
        per = g.geom.nvox * 4
        g.data = np.concatenate([g.data[t * per:(t + 1) * per] for t in sel])
        g.ngrids = len(sel)
    info = G.make_info(spacing, ext, plateau, nd, probe_types, natoms=src.n, pivot=pivot)
    G.write(path, info, inner, outer, nl, src.xyz)


def _files(W, d, seed=20261207, n_rec=400, n_lig=120, n_poses=1000, modes=0):
    """Write the input files into directory d; returns their paths.  With
    modes > 0, both bodies get a mode file of that many modes, the ligand a
    grid file with its alphabet (the receptor's types), and the DOF lines mode
    amplitudes."""
    from gpuattract_amd import gridfile as G
    T = 12
    R_rec = W.scaled_radius(n_rec, W.DENSITY_REF)
    R_lig = W.scaled_radius(n_lig, W.LIG_REF)
    rec = W.make_body(seed, 0, n_rec, R_rec, T, 0.2, 0.0)
    lig = W.make_body(seed, 1, n_lig, R_lig, T, 0.2, 0.0)
    # the bodies away from the origin, so that pivots and the grid
    # displacements matter
    c_rec, c_lig = (31.5, -12.25, 18.0), (-7.75, 40.5, 3.25)
    names = dict(rec="receptorr.pdb", lig="ligandr.pdb", par="attract.par",
                 grid="receptorgrid.grid", alphabet="receptorgrid.alphabet", dof="docking.dat")
    if modes:
        names.update(grid_lig="ligandgrid.grid", alphabet_lig="ligandgrid.alphabet",
                     modes_rec="hm-rec.dat", modes_lig="hm-lig.dat")
    paths = {k: os.path.join(d, v) for k, v in names.items()}
    write_pdb(paths["rec"], _shifted(W, rec, c_rec))
    write_pdb(paths["lig"], _shifted(W, lig, c_lig))
    p = W.make_params(seed, T)
    _write_params(paths["par"], p)

    def alphabet_of(body, path):  # a body's types, in ascending order
        types = sorted(int(t) for t in set(body.type.tolist()) if t > 0)
        with open(path, "w") as f:
            f.write("".join(f"{t}\n" for t in types))
        return types

    # the receptor's grids from the receptor as its file holds it (global
    # frame), for the ligand's types
    _grid_file(W, G, paths["grid"], paths["rec"], p, alphabet_of(lig, paths["alphabet"]), c_rec)
    if modes:
        _grid_file(W, G, paths["grid_lig"], paths["lig"], p,
                   alphabet_of(rec, paths["alphabet_lig"]), c_lig)
        write_modes(paths["modes_rec"], n_rec, modes, seed + 1)
        write_modes(paths["modes_lig"], n_lig, modes, seed + 2)
    # poses: ligand centres on a shell around the receptor centre, near contact
    d6 = W.make_poses(seed, n_poses, 0.95 * (R_rec + R_lig), 1.2 * (R_rec + R_lig))
    shift = np.float64(c_rec) - np.float64(c_lig)
    amp = np.random.default_rng(seed).normal(scale=2.0, size=(n_poses, 2, modes))
    lig_rows = [np.concatenate([q[3:6], q[0:3] + shift, amp[k, 1]])
                for k, q in enumerate(d6.astype(np.float64))]
    rec_rows = [np.concatenate([np.zeros(6), amp[k, 0]]) for k in range(n_poses)]
    header = (f"#pivot 1 {c_rec[0]} {c_rec[1]} {c_rec[2]}\n#pivot 2 {c_lig[0]} {c_lig[1]} {c_lig[2]}\n"
              "#centered receptor: false\n#centered ligands: false\n")
    write_dof(paths["dof"], rec_rows, lig_rows, header)
    return paths


def _grid_from_file(W, G, path, pivot):
    """An ATTRACT grid file through this library's reader, displaced by
    -pivot as GridUnion::translate does (origin + displacement in float)."""
    g = G.read(path)

    def moved(o):
        return [float(np.float32(o[k]) + (-pivot[k])) for k in range(3)]

    gi, go, gn = g.inner, g.outer, g.nl
    inner = W.IGrid(W.Geom(gi[0], gi[1], gi[2], *moved(gi[4:7]), gi[7]), gi[3], g.inner_data)
    outer = W.IGrid(W.Geom(go[0], go[1], go[2], *moved(go[4:7]), go[7]), go[3], g.outer_data)
    nl = W.NLGrid(W.Geom(gn[0], gn[1], gn[2], *moved(gn[3:6]), gn[6]), gn[7], g.desc, g.list)
    return inner, outer, nl


def _ours(W, paths, modes=0):
    """The files through this library's readers and the drop-in Server on
    the GPU: (n, 7) results, or (n, 27) with modes."""
    from gpuattract_amd import formats as F, gridfile as G, hip
    L = C.CDLL(os.path.join(HERE, "cpp", "_bin", "libservice_driver.so"))
    L.drv_error.restype = C.c_char_p
    rec, lig = F.Body(paths["rec"]), F.Body(paths["lig"])
    lig.map_types(F.read_alphabet(paths["alphabet"]))
    if modes:
        rec.read_modes(paths["modes_rec"], modes)
        lig.read_modes(paths["modes_lig"], modes)
        rec.map_types(F.read_alphabet(paths["alphabet_lig"]))
    dof = F.DofFile(paths["dof"])
    info = dof.info()
    assert not info.autoPivot and info.numPivots == 2
    piv = [np.array(info.pivots[k][:], np.float32) for k in range(2)]
    rec.pivotize(piv[0])
    lig.pivotize(piv[1])
    dofs = dof.ligand(modes=modes > 0, num_modes_rec=modes, num_modes_lig=modes)
    T, shape, _, _, tab = F.read_params(paths["par"])
    inner, outer, nl = _grid_from_file(W, G, paths["grid"], piv[0])
    rd, ld = hip.ProteinDesc(), hip.ProteinDesc()
    F._chk(F.lib().att_body_desc(rec.h, C.byref(rd)))
    F._chk(F.lib().att_body_desc(lig.h, C.byref(ld)))
    di, do, dn = hip.grid_desc(inner), hip.grid_desc(outer), hip.nl_desc(nl)
    d = np.ascontiguousarray(dofs, np.float32)
    out = np.full((d.shape[0], 27 if modes else 7), np.nan, np.float32)
    ffelec = float(np.float32(FELEC / EPSILON))
    tail = (hip._p(tab), T, shape, 1, C.c_float(EPSILON), C.c_float(ffelec), hip._p(d),
            C.c_size_t(d.shape[0]), hip._p(out), 2, 100, 4)
    if modes:
        li, lo, _ = _grid_from_file(W, G, paths["grid_lig"], piv[1])
        dli, dlo = hip.grid_desc(li), hip.grid_desc(lo)
        st = L.drv_score_modes(C.byref(rd), C.byref(ld), C.byref(di), C.byref(do), C.byref(dn),
                               C.byref(dli), C.byref(dlo), *tail)
    else:
        st = L.drv_score_rigid(C.byref(rd), C.byref(ld), C.byref(di), C.byref(do), C.byref(dn),
                               *tail)
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


def test_flexible_scoring_from_files_against_reference(W, O, tmp_path):
    if not O.ref_available():
        pytest.skip("reference tree not built (oracle/_ref)")
    from gpuattract_amd import formats as F
    M = 4
    paths = _files(W, str(tmp_path), seed=20261208, n_rec=300, n_lig=120, n_poses=500, modes=M)
    print("flexible scoring from files: " +
          ", ".join(os.path.basename(v) for v in paths.values()), flush=True)
    ref = O.ref_score_files_modes(paths["dof"], paths["rec"], paths["lig"], paths["grid"],
                                  paths["grid_lig"], paths["par"], paths["alphabet"],
                                  paths["alphabet_lig"], paths["modes_rec"], paths["modes_lig"], M,
                                  "variable", EPSILON)
    got = _ours(W, paths, modes=M)
    assert got.shape == ref.shape == (500, 27)
    assert np.mean(ref[:, 0] != 0) > 0.9
    keep = list(range(7)) + list(range(17, 27))
    diff = np.flatnonzero(np.any(got[:, keep].view(np.uint32) != ref[:, keep].view(np.uint32),
                                 axis=1))
    assert diff.size == 0, f"{diff.size} poses differ, first {diff[:5]}"
    r = got[:, 7:7 + M].astype(np.float64)
    q = ref[:, 7:7 + M].astype(np.float64)
    err = np.linalg.norm(r - q, axis=1)
    nrm = np.linalg.norm(q, axis=1)
    assert np.all(err <= 1e-5 * nrm + 1e-30), (err / np.maximum(nrm, 1e-30)).max()
    same = got.copy()
    same[:, 7:7 + M] = ref[:, 7:7 + M]
    assert F.format_results(same, True, M, M) == O.ref_format_results(ref, True, M, M)
    print(f"{got.shape[0]} flexible poses from files: energy, force, torque and ligand mode "
          f"gradients equal the reference's; receptor mode gradients within "
          f"{(err / np.maximum(nrm, 1e-30)).max():.1e} of their norm")
