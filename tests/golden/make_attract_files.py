"""Copies the reference's own test resources (test/resources/, data files
of its file-reader and scoring tests) into tests/golden/attract/ as fixtures,
the docking DOF file cut to its first 200 structures.  Run once where the
reference tree is present; the test suite reads only the copies."""
import os
import shutil
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/resources"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "attract")


def main():
    os.makedirs(DST, exist_ok=True)
    for name in ("receptorr.pdb", "ligandr.pdb", "attract.par", "receptorgrid.alphabet",
                 "readDof_test.dat"):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
    # the header and the first 200 pose blocks ("#<k>" lines) of the docking output
    out, blocks = [], 0
    with open(os.path.join(SRC, "out_docking_orig.dat")) as f:
        for line in f:
            if line.startswith("#") and line[1:].strip().isdigit():
                blocks += 1
                if blocks > 200:
                    break
            out.append(line)
    with open(os.path.join(DST, "out_docking_200.dat"), "w") as f:
        f.writelines(out)


if __name__ == "__main__":
    main()
