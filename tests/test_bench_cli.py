"""bench.py's command line on the CPU: more than one GPU is refused with a
message, and workloads made without host grids carry the same geometry,
bodies and poses as those with them."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    msg = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_workload_without_host_grids_matches(W):
    a = W.make_workload("g", 38, 400, 120, 50, 2, 2)
    b = W.make_workload("g", 38, 400, 120, 50, 2, 2, host_grids=False)
    assert b.rec_inner.data is None and b.rec_outer.data is None and b.rec_nl.list is None
    assert a.rec_inner.geom == b.rec_inner.geom and a.rec_outer.geom == b.rec_outer.geom
    assert a.rec_nl.geom == b.rec_nl.geom and a.rec_nl.plateau == b.rec_nl.plateau
    assert np.array_equal(a.dofs, b.dofs)
    assert np.array_equal(a.rec.xyz, b.rec.xyz) and np.array_equal(a.lig.xyz, b.lig.xyz)
    assert np.array_equal(a.lig_inner.data, b.lig_inner.data)
