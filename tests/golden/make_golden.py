"""Write the golden vectors of tests/golden/ from the reference's own CPU code.

Runs in the development container only: it needs oracle/_ref/libattract_ref.so,
compiled by oracle/Makefile from the reference sources.  Each fixture stores
the generator parameters, a checksum of every input array (so a test can
prove it regenerated the identical inputs), the DOFs and the reference's
outputs: per-pose results and, for the first poses, per-bead coordinates,
forces and energies, and (rigid) the grid cells and neighbour-list pair
counts from the reference's own index functions.

    python tests/golden/make_golden.py
"""
import importlib.util
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

from make_golden_sums import input_checksums  # noqa: E402


def _package():
    spec = importlib.util.spec_from_file_location(
        "gpuattract_amd", os.path.join(ROOT, "gpuattract_2.0_amd", "__init__.py"),
        submodule_search_locations=[os.path.join(ROOT, "gpuattract_2.0_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gpuattract_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


RIGID = [
    dict(name="rigid_300_80", args=[101, 300, 80, 200], kw={}),
    dict(name="rigid_1000_300", args=[102, 1000, 300, 200], kw={}),
    dict(name="rigid_type0", args=[103, 250, 70, 120], kw=dict(type0_frac=0.15)),
    dict(name="rigid_12_6", args=[104, 250, 70, 120], kw=dict(shape=12)),
    dict(name="rigid_const_dielec", args=[105, 250, 70, 120], kw=dict(dielec=0, charged_frac=0.5)),
]
MODES = [
    dict(name="modes_300_90", args=[201, 300, 90, 100, 3, 4], kw={}),
    dict(name="modes_3000_800", args=[202, 3000, 800, 40, 5, 5], kw={}),
    dict(name="modes_rec_only", args=[203, 260, 70, 60, 4, 0], kw={}),
]


BEAD_POSES = 16  # poses with per-bead reference values (rigid)


def main():
    pkg = _package()
    from gpuattract_amd import workload as W
    import oracle_py as O
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libattract_ref.so missing: make -C oracle ref")
    for spec in RIGID:
        wl = W.make_workload(spec["name"], *spec["args"], **spec["kw"])
        res = O.ref_score_rigid(wl)
        k = min(BEAD_POSES, wl.dofs.shape[0])
        beads = [O.ref_beads_rigid(wl, wl.dofs[i]) for i in range(k)]
        cells = [O.ref_cells_rigid(wl, wl.dofs[i]) for i in range(k)]
        np.savez_compressed(
            os.path.join(HERE, spec["name"] + ".npz"),
            spec=json.dumps(spec), checksums=json.dumps(input_checksums(W, wl)),
            dofs=wl.dofs, ref=res, bead_xyz=np.stack([b[0] for b in beads]),
            bead_pot=np.stack([b[1] for b in beads]), bead_cells=np.stack([c[0] for c in cells]),
            bead_npairs=np.stack([c[1] for c in cells]))
        print(spec["name"], res.shape)
    for spec in MODES:
        wl = W.make_workload(spec["name"], *spec["args"], **spec["kw"])
        res, rec_pot, lig_pot = O.ref_score_modes(wl)
        np.savez_compressed(
            os.path.join(HERE, spec["name"] + ".npz"),
            spec=json.dumps(spec), checksums=json.dumps(input_checksums(W, wl)),
            dofs=wl.dofs, ref=res, rec_pot=rec_pot, lig_pot=lig_pot)
        print(spec["name"], res.shape)
    # Parameter derivation (fileIO/readFile.cpp:439-527) for both shapes.
    out = {}
    for shape in (8, 12):
        p = W.make_params(301, T=12, shape=shape)
        d = O.ref_read_params(p)
        for k, v in d.items():
            out[f"{shape}_{k}"] = v
        out[f"{shape}_rbc"], out[f"{shape}_abc"], out[f"{shape}_ipon_raw"] = p.rbc, p.abc, p.ipon_raw
    np.savez_compressed(os.path.join(HERE, "params.npz"), **out)
    # Plateau ratio table (model/NLGrid.tpp:44-51): checksum and a sample.
    lut = O.ref_ratio_lut(10.0)
    idx = np.unique(np.concatenate([np.arange(0, 2000), np.linspace(0, len(lut) - 1, 5000).astype(int)]))
    np.savez_compressed(os.path.join(HERE, "ratio.npz"), size=len(lut), checksum=str(W.checksum(lut)),
                        idx=idx, val=lut[idx])
    # Rotation and torque matrices of the reference's float build.
    rng = np.random.default_rng(401)
    ang = np.concatenate([rng.uniform(-7, 7, (1200, 3)), rng.uniform(-0.01, 0.01, (100, 3))]).astype(np.float32)
    R = np.stack([O.ref_rotmat(*a) for a in ang])
    TM = np.stack([O.ref_torquemat(*a) for a in ang])
    np.savez_compressed(os.path.join(HERE, "rotmat.npz"), ang=ang, R=R, TM=TM)
    print("params, ratio, rotmat written")


if __name__ == "__main__":
    main()
