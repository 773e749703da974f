"""The reference's file formats (include/attract_formats.h), checked against
the reference's own readers, DOF conversion and result printer compiled in
oracle/_ref (fileIO/readFile.cpp, model/DOFConverter.h,
model/DOFTransform.cpp, model/Types_6D*.tpp) on files written here from the
seeded generator: arrays bit for bit, text byte for byte.  Malformed input
that the reference answers with undefined behaviour or exit() is refused with
an error here (only our side sees it)."""
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def F():
    from gpuattract_amd import formats
    return formats


@pytest.fixture(scope="module")
def R(O):
    if not O.ref_available():
        pytest.skip("reference tree not built (oracle/_ref)")
    return O


def write_pdb(path, body, trailing_newline=True, blank_every=0):
    X = body.xyz.reshape(3, -1)
    lines = []
    for i in range(body.n):
        if blank_every and i and i % blank_every == 0:
            lines.append("")
        lines.append(f"ATOM  {i + 1:5d} {'CA':<4s} {'ALA'} A{i // 3 + 1:4d}    "
                     f"{X[0, i]:8.3f}{X[1, i]:8.3f}{X[2, i]:8.3f}{int(body.type[i]):5d}"
                     f"{float(body.charge[i]):8.3f} 0 1.00")
    open(path, "w").write("\n".join(lines) + ("\n" if trailing_newline else ""))


def write_modes(path, n, M, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(M, n, 3))
    eig = rng.uniform(0.5, 5.0, M)
    with open(path, "w") as f:
        for m in range(M):
            f.write(f"  {m + 1}  {eig[m]:.6f}\n")
            for i in range(n):
                f.write(f"  {v[m, i, 0]:.6f}  {v[m, i, 1]:.6f}  {v[m, i, 2]:.6f}\n")


def test_pdb_matches_reference_reader(W, F, R, tmp_path):
    body = W.make_body(71, 0, 400, W.scaled_radius(400, W.DENSITY_REF), 12, 0.2, 0.05)
    for trailing, blanks in ((True, 0), (True, 37), (False, 0)):
        path = str(tmp_path / f"r_{trailing}_{blanks}.pdb")
        write_pdb(path, body, trailing, blanks)
        a = F.Body(path).arrays()
        xyz, t, q = R.ref_pdb(path)
        assert a.n == len(t)
        assert np.array_equal(a.xyz.view(np.uint32), xyz.view(np.uint32))
        assert np.array_equal(a.type, t) and np.array_equal(a.charge.view(np.uint32), q.view(np.uint32))
        # the reference drops a last line without a newline
        assert a.n == (body.n if trailing else body.n - 1)
        assert np.allclose(a.xyz.reshape(3, -1), body.xyz.reshape(3, -1)[:, :a.n], atol=6e-4)


def test_modes_match_reference_reader(W, F, R, tmp_path):
    body = W.make_body(72, 1, 150, W.scaled_radius(150, W.LIG_REF))
    pdb = str(tmp_path / "l.pdb")
    write_pdb(pdb, body)
    for M, extra in ((5, 0), (3, 2)):  # the file may hold more modes than asked for
        mf = str(tmp_path / f"m{M}.dat")
        write_modes(mf, body.n, M + extra, 72 + M)
        b = F.Body(pdb)
        b.read_modes(mf, M)
        a = b.arrays()
        modes, force = R.ref_modes(pdb, mf, M, body.n)
        assert np.array_equal(a.modes.view(np.uint32), modes.view(np.uint32))
        assert np.array_equal(a.mode_force.view(np.uint32), force.view(np.uint32))
        # each mode is normalised over the body
        m = a.modes.reshape(3, body.n, M)
        assert np.allclose((m ** 2).sum(axis=(0, 1)), 1.0, atol=1e-5)


def test_params_match_reference_reader(W, F, R, tmp_path):
    for shape in (8, 12):
        p = W.make_params(73 + shape, 12, shape=shape)
        path = str(tmp_path / f"attract{shape}.par")
        T = p.T
        with open(path, "w") as f:
            f.write(f"{shape} {T} 0 0\n")
            for arr, fmt in ((p.rbc, "{:.6f}"), (p.abc, "{:.6f}"), (p.ipon_raw, "{:d}")):
                for i in range(T):
                    f.write(" ".join(fmt.format(v) for v in arr[i * T:(i + 1) * T]) + "\n")
        T2, shp, on, off, tab = F.read_params(path)
        assert (T2, shp) == (T, shape)
        rT, (rc, ac, ipon, rmin2, emin) = R.ref_params_file(path)
        assert rT == T
        for ours, ref in ((tab["rc"], rc), (tab["ac"], ac), (tab["rmin2"], rmin2), (tab["emin"], emin)):
            assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(tab["ipon"], ipon)


def write_dof(path, rec, lig, header):
    with open(path, "w") as f:
        f.write(header)
        for k in range(len(lig)):
            f.write(f"#{k + 1}\n")
            for d in (rec[k], lig[k]):
                f.write("  " + " ".join(f"{v:.8f}" for v in d) + "\n")


@pytest.mark.parametrize("case", ["explicit", "auto", "centered", "receptor moves", "modes"])
def test_dof_files_match_reference(W, F, R, tmp_path, case):
    rng = np.random.default_rng(abs(hash(case)) % 1000)
    n = 200
    d6 = W.make_poses(74, n, 20.0, 30.0)
    lig = [np.concatenate([d[3:6], d[0:3]]) for d in d6]           # ang, pos
    rec = [np.zeros(6) for _ in range(n)]
    Mr = Ml = 0
    if case == "receptor moves":
        rec = [np.concatenate([rng.uniform(-3, 3, 3), rng.normal(scale=2.0, size=3)]) for _ in range(n)]
        rec[0] = np.array([0.0, 0.0, 0.0, 1.0, 2.0, 3.0])      # translation only
        rec[1] = np.array([0.1, 0.0, 0.0, 0.0, 0.0, 0.0])      # rotation only
    if case == "modes":
        Mr, Ml = 4, 3
        rec = [np.concatenate([np.zeros(6), rng.normal(size=Mr)]) for _ in range(n)]
        lig = [np.concatenate([l, rng.normal(size=Ml)]) for l in lig]
        lig[5] = lig[5][:6 + Ml - 1]                            # a missing amplitude reads as 0
    piv = "#pivot 1 1.5 -2.25 3.0\n#pivot 2 -0.5 0.75 10.125\n"
    cen = "#centered receptor: false\n#centered ligands: false\n"
    if case == "auto":
        piv = "#pivot auto\n"
    if case == "centered":
        cen = "#centered receptor: true\n#centered ligand: true\n"
    path = str(tmp_path / "dofs.dat")
    write_dof(path, rec, lig, "## attract docking\n" + piv + cen)
    f = F.DofFile(path)
    info = f.info()
    assert info.numMolecules == 2 and info.numPoses == n
    auto = (np.float32([0.25, -1.5, 2.0]), np.float32([3.0, 0.5, -1.25]))
    pr, pl = auto if case == "auto" else (None, None)
    ours = f.ligand(pr, pl, modes=case == "modes", num_modes_rec=Mr, num_modes_lig=Ml)
    ref = R.ref_dof(path, modes=case == "modes", Mr=Mr, Ml=Ml, auto_rec=auto[0], auto_lig=auto[1])
    if case == "modes":
        # the reference reads an absent amplitude from uninitialised memory
        keep = np.ones(n, bool)
        keep[5] = False
        assert np.array_equal(ours[keep].view(np.uint32), ref[keep].view(np.uint32))
        assert ours[5, 16 + Ml - 1] == 0.0
    else:
        assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


def test_dof_transform_is_a_frame_change(W, F, O, tmp_path):
    """Pinned beside the reference: placing the ligand with the converted DOFs
    (receptor fixed) equals placing it with the original DOFs and then applying
    the receptor's inverse transform as the reference defines it, R(-angles)
    after subtracting the receptor translation (model/DOFTransform.cpp:62-77);
    this checks the Euler extraction and the composition order."""
    rng = np.random.default_rng(5)
    n = 40
    rec = [np.concatenate([rng.uniform(-2, 2, 3), rng.normal(scale=3.0, size=3)]) for _ in range(n)]
    lig = [np.concatenate([rng.uniform(-2, 2, 3), rng.normal(scale=20.0, size=3)]) for _ in range(n)]
    path = str(tmp_path / "frame.dat")
    write_dof(path, rec, lig, "#pivot 1 0 0 0\n#pivot 2 0 0 0\n#centered receptor: true\n"
                              "#centered ligands: true\n")
    out = F.DofFile(path).ligand()
    rot = lambda a: O.rotmat(*[float(v) for v in a]).reshape(3, 3).astype(np.float64)
    x = rng.normal(size=(5, 3)) * 5
    for k in range(n):
        M = rot(-np.float32(rec[k][:3]))
        Rl = rot(np.float32(lig[k][:3]))
        Rn = rot(out[k, 3:6])
        want = (M @ ((Rl @ x.T).T + lig[k][3:] - rec[k][3:]).T).T
        got = (Rn @ x.T).T + out[k, 0:3]
        assert np.allclose(got, want, atol=2e-3), k


@pytest.mark.parametrize("modes", [False, True])
def test_result_text_matches_reference(F, R, modes):
    rng = np.random.default_rng(9)
    n = 30
    res = rng.normal(size=(n, 27 if modes else 7)).astype(np.float32) * \
        np.float32(10.0) ** rng.integers(-6, 9, size=(n, 27 if modes else 7))
    res[0, 0] = 0.0
    res[1, 0] = -1234.5678
    Mr, Ml = (5, 3) if modes else (0, 0)
    assert F.format_results(res, modes, Mr, Ml) == R.ref_format_results(res, modes, Mr, Ml)


def test_alphabet_matches_reference_reader(F, R, tmp_path):
    path = tmp_path / "rec.alphabet"
    path.write_text(" 3\n1\n 12 \n2\n99\n")
    ours = F.read_alphabet(str(path))
    assert list(ours) == [3, 1, 12, 2, 99]
    assert np.array_equal(ours, R.ref_alphabet(str(path)))
    from gpuattract_amd import hip
    for bad in ("1\n-2\n", "1\nx\n"):
        path.write_text(bad)
        with pytest.raises(hip.AttError, match="not a non-negative bead type"):
            F.read_alphabet(str(path))


def test_type_mapping(W, F, tmp_path):
    body = W.make_body(75, 1, 60, W.scaled_radius(60, W.LIG_REF), 6, 0.2, 0.1)
    path = str(tmp_path / "map.pdb")
    write_pdb(path, body)
    b = F.Body(path)
    alphabet = [3, 1, 2, 5, 4, 6, 99]
    b.map_types(alphabet)
    a = b.arrays()
    want = [0 if t == 0 else alphabet.index(t) + 1 for t in body.type]
    assert list(a.mapped) == want
    from gpuattract_amd import hip
    with pytest.raises(hip.AttError, match="not in the grid alphabet"):
        b.map_types([1, 2])


def test_malformed_inputs_are_refused(F, tmp_path):
    from gpuattract_amd import hip
    p = tmp_path / "short.pdb"
    p.write_text("ATOM      1  CA  ALA A   1\nATOM\n")
    with pytest.raises(hip.AttError, match="ends before column"):
        F.Body(str(p))
    d = tmp_path / "bad.dat"
    d.write_text("#pivot 1 0 0 0\n#pivot 2 0 0 0\n#centered receptor: maybe\n")
    with pytest.raises(hip.AttError, match="centering"):
        F.DofFile(str(d))
    d.write_text("#pivot 1 0 0 0\n#pivot 2 0 0 0\n#centered receptor: false\n#centered ligands: false\n"
                 "#1\n 0 0 0 0 0 0\n 0 0 0 1 2 3\n#2\n 0 0 0 0 0 0\n")
    with pytest.raises(hip.AttError, match="incomplete"):
        F.DofFile(str(d))
    par = tmp_path / "bad.par"
    par.write_text("7 2 0 0\n1 1 1 1\n")
    with pytest.raises(hip.AttError, match="Unknown potential shape"):
        F.read_params(str(par))
