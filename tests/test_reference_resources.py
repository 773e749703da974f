"""The reference's own test resources (test/resources/, copied to
tests/golden/attract/ by tests/golden/make_attract_files.py): real ATTRACT
files read by this library's readers.

* readDof_test.dat: the known answers of the reference's reader test
  (test/fileIO/readDof_test.cpp:17-36): three molecules of three poses, every
  line phi ssi rot = 1 2 3, x y z = 4 5 6, eight values, amplitudes 7 and 8;
  "#pivot auto", centred ligands.
* receptorr.pdb, ligandr.pdb, attract.par, receptorgrid.alphabet and the first
  200 structures of out_docking_orig.dat (a docking run's output with its
  "##" comment and "### SEED" lines): equal, bit for bit, to what the
  reference's own readers, DOF converter and frame transform return for the
  same files (compiled in oracle/_ref; skipped where it is not built).
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RES = os.path.join(HERE, "golden", "attract")


def res(name):
    return os.path.join(RES, name)


@pytest.fixture(scope="module")
def F():
    from gpuattract_amd import formats
    return formats


@pytest.fixture(scope="module")
def R(O):
    if not O.ref_available():
        pytest.skip("reference tree not built (oracle/_ref)")
    return O


def test_read_dof_known_answers(F):
    f = F.DofFile(res("readDof_test.dat"))
    info = f.info()
    assert info.numMolecules == 3 and info.numPoses == 3
    assert info.autoPivot == 1 and info.numPivots == 0
    assert info.centeredReceptor == 0 and info.centeredLigands == 1
    for m in range(3):
        for k in range(3):
            assert list(f.values(m, k)) == [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0]
    from gpuattract_amd import hip
    with pytest.raises(hip.AttError, match="Failed to open file non-existing"):
        F.DofFile("non-existing filename")


def test_reference_pdbs(F, R):
    for name in ("receptorr.pdb", "ligandr.pdb"):
        a = F.Body(res(name)).arrays()
        xyz, t, q = R.ref_pdb(res(name))
        assert a.n == len(t) > 0
        assert np.array_equal(a.xyz.view(np.uint32), xyz.view(np.uint32))
        assert np.array_equal(a.type, t) and np.array_equal(a.charge.view(np.uint32), q.view(np.uint32))
        assert 99 in set(a.type.tolist())  # ATTRACT's dummy type


def test_reference_parameter_file(F, R):
    T, shape, on, off, tab = F.read_params(res("attract.par"))
    assert (T, shape) == (98, 8)
    rT, (rc, ac, ipon, rmin2, emin) = R.ref_params_file(res("attract.par"))
    assert rT == T
    for ours, ref in ((tab["rc"], rc), (tab["ac"], ac), (tab["rmin2"], rmin2), (tab["emin"], emin)):
        assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(tab["ipon"], ipon)


def test_reference_alphabet_and_type_map(F, R):
    alphabet = F.read_alphabet(res("receptorgrid.alphabet"))
    assert list(alphabet) == list(range(1, 32)) + [99]
    assert np.array_equal(alphabet, R.ref_alphabet(res("receptorgrid.alphabet")))
    # the ligand's types through the alphabet (model/TypeMap.h:52-60): type t
    # to its position + 1, the dummy type 99 to 0 (no grid)
    b = F.Body(res("ligandr.pdb"))
    b.map_types(alphabet)
    a = b.arrays()
    want = [0 if t == 99 else list(alphabet).index(t) + 1 for t in a.type]
    assert list(a.mapped) == want


def test_reference_docking_dofs(F, R):
    path = res("out_docking_200.dat")
    f = F.DofFile(path)
    info = f.info()
    assert info.numMolecules == 2 and info.numPoses == 200 and info.autoPivot == 0
    assert np.array_equal(np.float32(info.pivots[0][:]), np.float32([32.823, -6.395, 23.484]))
    assert np.array_equal(np.float32(info.pivots[1][:]), np.float32([51.143, 7.680, 38.110]))
    # the first structure's lines, as written in the file
    assert np.array_equal(f.values(0, 0), np.zeros(6))
    assert np.array_equal(f.values(1, 0), [9.109876, 4.076754, -2.675263, -34.8734, -29.0924,
                                           -32.6334])
    ours = f.ligand()
    ref = R.ref_dof(path)
    assert ours.shape == ref.shape == (200, 6)
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
