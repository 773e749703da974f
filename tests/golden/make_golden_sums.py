"""Checksums of every input array of a workload (shared by make_golden.py
and the tests that regenerate the golden inputs)."""


def input_checksums(W, wl):
    arrs = {
        "rec.xyz": wl.rec.xyz, "rec.type": wl.rec.type, "rec.charge": wl.rec.charge,
        "lig.xyz": wl.lig.xyz, "lig.type": wl.lig.type, "lig.charge": wl.lig.charge,
        "rec.modes": wl.rec.modes, "lig.modes": wl.lig.modes,
        "rec.mode_force": wl.rec.mode_force, "lig.mode_force": wl.lig.mode_force,
        "params.rc": wl.params.rc, "params.ac": wl.params.ac, "params.emin": wl.params.emin,
        "params.rmin2": wl.params.rmin2, "params.ipon": wl.params.ipon,
        "rec_inner": wl.rec_inner.data, "rec_outer": wl.rec_outer.data,
        "rec_nl.desc": wl.rec_nl.desc, "rec_nl.list": wl.rec_nl.list, "dofs": wl.dofs,
    }
    if wl.lig_inner is not None:
        arrs.update({"lig_inner": wl.lig_inner.data, "lig_outer": wl.lig_outer.data})
    return {k: str(W.checksum(v)) for k, v in arrs.items()}
