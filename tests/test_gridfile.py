"""ATTRACT grid files (include/attract_io.h).

Grids made by the seeded generator in ATTRACT's geometry (outer spacing twice
the inner, outer origin gridextension spacings below the inner) are written
with att_gridfile_write, then read back by att_gridfile_read and by the
reference's own reader (fileIO/readFile.cpp:260-422 over grid_orig.cpp:85-301,
compiled into oracle/_ref): both must return the written grids bit for bit,
with the reference's geometry conversion.  Malformed files must be refused
with the reference's messages (only our reader sees those: the reference's
exits the process)."""
import os
import struct

import numpy as np
import pytest


def attract_geoms(W, body, spacing, ext, margin):
    gi = W.inner_geom(body, spacing, margin)
    sp2 = float(np.float32(2 * spacing))
    o = [float(np.float32(np.float64(np.float32(v)) - ext * spacing)) for v in (gi.ox, gi.oy, gi.oz)]
    n_out = [int(np.ceil((n - 1) * spacing / (2 * spacing))) + ext + 1 for n in (gi.nx, gi.ny, gi.nz)]
    go = W.Geom(n_out[0], n_out[1], n_out[2], o[0], o[1], o[2], sp2)
    return gi, go


@pytest.fixture(scope="module")
def G():
    from gpuattract_amd import gridfile
    return gridfile


def make(W, G, seed, types, spacing=0.9, ext=8, n=300, plateau=10.0, nd=10.8):
    T = 12
    body = W.make_body(seed, 0, n, W.scaled_radius(n, W.DENSITY_REF), T, 0.2, 0.0)
    # keep only the alphabet's types (others become type 0, no grid)
    keep = np.isin(body.type, types)
    body.type[~keep] = 0
    body.mtype[:] = body.type
    p = W.make_params(seed, T)
    gi, go = attract_geoms(W, body, spacing, ext, 6.0)
    inner = W.build_grid(gi, body, p, False, plateau)
    outer = W.build_grid(go, body, p, False, plateau)
    nl = W.build_nl(gi, body, nd, plateau)
    # the file holds the electrostatic grid and the alphabet's types only
    sel = [0] + list(types)
    for g in (inner, outer):
        per = g.geom.nvox * 4
        g.data = np.concatenate([g.data[t * per:(t + 1) * per] for t in sel])
        g.ngrids = len(sel)
    info = G.make_info(spacing, ext, plateau, nd, types, natoms=body.n, pivot=(1.5, -2.0, 0.25))
    return body, inner, outer, nl, info


@pytest.mark.parametrize("types", [list(range(1, 13)), [1, 3, 4, 7, 11]])
def test_write_read_round_trip_and_reference_reader(W, G, O, tmp_path, types):
    body, inner, outer, nl, info = make(W, G, 51, types)
    path = str(tmp_path / "rec.grid")
    G.write(path, info, inner, outer, nl, body.xyz)
    f = G.read(path)
    gi, go = inner.geom, outer.geom
    assert f.inner == (gi.nx, gi.ny, gi.nz, len(types) + 1, np.float32(gi.ox), np.float32(gi.oy),
                       np.float32(gi.oz), np.float32(gi.spacing))
    assert f.outer[:4] == (go.nx, go.ny, go.nz, len(types) + 1)
    assert np.array_equal(np.float32(f.outer[4:]), np.float32([go.ox, go.oy, go.oz, go.spacing]))
    assert np.array_equal(f.inner_data.view(np.uint32), inner.data.view(np.uint32))
    assert np.array_equal(f.outer_data.view(np.uint32), outer.data.view(np.uint32))
    # empty voxels: count 0 and start 0xFFFFFFFF, as the reference converts them
    empty = nl.desc[0::2] == 0
    want = nl.desc.copy()
    want[1::2][empty] = 0xFFFFFFFF
    assert np.array_equal(f.desc, want) and np.array_equal(f.list, nl.list)
    assert f.info.alphabetsize == len(types) and f.info.natoms == body.n
    assert list(f.info.pivot) == [1.5, -2.0, 0.25]
    if not O.ref_available():
        pytest.skip("reference tree not built (oracle/_ref)")
    dims, geo, r_in, r_out, r_desc, r_list = O.ref_read_gridfile(path)
    assert list(dims) == [gi.nx, gi.ny, gi.nz, len(types) + 1, go.nx, go.ny, go.nz,
                          len(types) + 1, gi.nx, gi.ny, gi.nz]
    assert np.array_equal(geo[:4], np.float32(f.inner[4:]))
    assert np.array_equal(geo[4:8], np.float32(f.outer[4:]))
    assert np.array_equal(geo[8:12], np.float32(f.nl[3:7]))
    assert geo[12] == np.float32(np.float32(f.nl[7]) * np.float32(f.nl[7]))
    assert np.array_equal(r_in.view(np.uint32), f.inner_data.view(np.uint32))
    assert np.array_equal(r_out.view(np.uint32), f.outer_data.view(np.uint32))
    assert np.array_equal(r_desc, f.desc) and np.array_equal(r_list, f.list)


def test_neighbour_types_follow_the_plateau(W, G, tmp_path):
    body, inner, outer, nl, info = make(W, G, 52, list(range(1, 13)), n=120)
    path = str(tmp_path / "t.grid")
    G.write(path, info, inner, outer, nl, body.xyz)
    raw = open(path, "rb").read()
    # header: 1 + 2 + 8 + 4 + 8 + 8 + 99 + 12 + 24 + 4 + 24 bytes
    off = 1 + 2 + 8 + 4 + 8 + 8 + 99 + 12 + 24 + 4 + 24
    nrE, = struct.unpack_from("<i", raw, off)
    off += 8 + 16 * nrE
    nrN, = struct.unpack_from("<i", raw, off)
    off += 8
    types = np.frombuffer(raw, np.uint8, nrN * 8, off)[0::8]
    gi = inner.geom
    X = body.xyz.reshape(3, -1).astype(np.float64)
    for v in range(0, gi.nvox, 97):
        c, s = nl.desc[2 * v], nl.desc[2 * v + 1]
        ix, iy, iz = v % gi.nx, (v // gi.nx) % gi.ny, v // (gi.nx * gi.ny)
        p = np.array([np.float32(gi.ox), np.float32(gi.oy), np.float32(gi.oz)], np.float64) + \
            np.array([ix, iy, iz]) * np.float64(np.float32(gi.spacing))
        for k in range(c):
            j = nl.list[s + k]
            d2 = ((p - X[:, j]) ** 2).sum()
            assert types[s + k] == (1 if d2 < 100.0 else 2)


def test_malformed_files_are_refused(W, G, tmp_path):
    from gpuattract_amd import hip
    body, inner, outer, nl, info = make(W, G, 53, [1, 2, 3], n=80)
    path = str(tmp_path / "ok.grid")
    G.write(path, info, inner, outer, nl)
    raw = bytearray(open(path, "rb").read())

    def refused(data, match):
        p = str(tmp_path / "bad.grid")
        open(p, "wb").write(bytes(data))
        with pytest.raises(hip.AttError, match=match):
            G.read(p)

    torque = bytearray(raw)
    torque[0] = 1
    refused(torque, "torque grid")
    arch = bytearray(raw)
    struct.pack_into("<H", arch, 1, 32)
    refused(arch, "computed on 32 bit")
    refused(raw[:len(raw) // 2], "Reading error in grid file")
    off_e = 1 + 2 + 8 + 4 + 8 + 8 + 99 + 12 + 24 + 4 + 24
    shm = bytearray(raw)
    struct.pack_into("<i", shm, off_e + 4, 7)
    refused(shm, "shared memory")
    # a voxel pointing past the records
    nrE, = struct.unpack_from("<i", raw, off_e)
    off_n = off_e + 8 + 16 * nrE
    nrN, = struct.unpack_from("<i", raw, off_n)
    off_v = off_n + 8 + 8 * nrN + 8
    bad = bytearray(raw)
    struct.pack_into("<I", bad, off_v + 4 * 99, nrE + 5)
    refused(bad, "innergrid voxel 1 atom type 100")
    with pytest.raises(hip.AttError, match="Reading error"):
        G.read(str(tmp_path / "missing.grid"))


def test_writer_refuses_non_attract_geometry(W, G, tmp_path):
    from gpuattract_amd import hip
    body, inner, outer, nl, info = make(W, G, 54, [1, 2], n=60)
    info.gridextension += 1  # outer origin no longer matches
    with pytest.raises(hip.AttError, match="outer origin"):
        G.write(str(tmp_path / "x.grid"), info, inner, outer, nl)
