// ATTRACT grid files <-> the scoring path's descriptors (include/attract_io.h).
//
// The byte layout is that of attract::Grid on LP64 (fileIO/grid_orig.h:44-
// 100): header fields written one by one, then the energy/gradient records
// (float energy, float grad[3]), the neighbours ({char type; unsigned index}
// = 8 bytes), the inner voxels ({unsigned potential[100]; int neighbourlist;
// short nr_neighbours} = 408 bytes) and the outer potentials (unsigned[100]).
// Indices in voxels are 1-based, 0 = none; slot 99 of a potential is the
// electrostatic term and marks a voxel that carries potentials.
#include "attract_io.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

constexpr int kTypes = 99;              // MAXATOMTYPES
constexpr size_t kVoxelBytes = 408;     // sizeof(attract::Voxel)
constexpr size_t kPotentialBytes = 400;
constexpr size_t kEnerGradBytes = 16;
constexpr size_t kNeighbourBytes = 8;

thread_local std::string t_err;

att_status err(att_status st, const std::string& msg) {
    t_err = msg;
    return st;
}

struct File {
    FILE* f = nullptr;
    explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
    bool get(void* p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }
    bool put(const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
    template <class T>
    bool get(T& v) { return get(&v, sizeof(T)); }
    template <class T>
    bool put(const T& v) { return put(&v, sizeof(T)); }
};

inline uint32_t u32_at(const unsigned char* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

// Grid slot order of the converted descriptors: the alphabet's types in
// ascending order, then electrostatics (rotated to the front afterwards).
std::vector<int> slot_types(const unsigned char* alphabet) {
    std::vector<int> s;
    for (int t = 0; t <= kTypes; ++t)
        if (t == kTypes || alphabet[t]) s.push_back(t);
    return s;
}

}  // namespace

struct att_gridfile {
    att_gridfile_info info{};
    unsigned nIn[3] = {0, 0, 0}, nOut[3] = {0, 0, 0};
    float oriIn[3] = {0, 0, 0}, oriOut[3] = {0, 0, 0};
    float spIn = 0, spOut = 0, plateau = 0;
    unsigned numGrids = 0;
    std::vector<float> inner, outer;
    std::vector<uint32_t> desc, list;
};

extern "C" {

const char* att_gridfile_last_error(void) { return t_err.c_str(); }

void att_gridfile_free(att_gridfile* g) { delete g; }

att_status att_gridfile_read(const char* path, att_gridfile** out) {
    if (!path || !out) return err(ATT_ERR_INVALID, "null argument");
    *out = nullptr;
    File f(path, "rb");
    const std::string name(path);
    const std::string rerr = "Reading error in grid file " + name;
    if (!f.f) return err(ATT_ERR_INVALID, rerr);
    unsigned char torque = 0;
    uint16_t arch = 0;
    if (!f.get(torque)) return err(ATT_ERR_INVALID, rerr);
    if (torque)
        return err(ATT_ERR_UNSUPPORTED, "Reading error in grid file " + name +
                                            ", grid was computed as a torque grid, but it was "
                                            "specified as a normal grid");
    if (!f.get(arch)) return err(ATT_ERR_INVALID, rerr);
    if (arch != 64)
        return err(ATT_ERR_UNSUPPORTED, "Reading error in grid file " + name + ", grid was computed on " +
                                            std::to_string(arch) + " bit, but we are on 64 bit");
    std::unique_ptr<att_gridfile> g(new att_gridfile);
    att_gridfile_info& I = g->info;
    float ori[3];
    int32_t dims[6], shm = 0, nrE = 0, nrN = 0;
    if (!f.get(I.gridspacing) || !f.get(I.gridextension) || !f.get(I.plateaudis) ||
        !f.get(I.neighbourdis) || !f.get(I.alphabet, kTypes) || !f.get(ori, sizeof(ori)) ||
        !f.get(dims, sizeof(dims)) || !f.get(I.natoms) || !f.get(I.pivot, sizeof(I.pivot)))
        return err(ATT_ERR_INVALID, rerr);
    I.alphabetsize = 0;
    for (int t = 0; t < kTypes; ++t) I.alphabetsize += I.alphabet[t] != 0;
    for (int k = 0; k < 6; ++k)
        if (dims[k] < 0) return err(ATT_ERR_INVALID, rerr + ": negative grid dimension");
    std::vector<unsigned char> energrads, neighbours, innergrid, biggrid;
    if (!f.get(nrE) || !f.get(shm)) return err(ATT_ERR_INVALID, rerr);
    if (nrE < 0) return err(ATT_ERR_INVALID, rerr);
    if (nrE) {
        if (shm != -1)
            return err(ATT_ERR_UNSUPPORTED, rerr + ": potentials in shared memory segment " +
                                                std::to_string(shm) + " are not supported");
        energrads.resize((size_t)nrE * kEnerGradBytes);
        if (!f.get(energrads.data(), energrads.size())) return err(ATT_ERR_INVALID, rerr);
    }
    if (!f.get(nrN) || !f.get(shm)) return err(ATT_ERR_INVALID, rerr);
    if (nrN < 0) return err(ATT_ERR_INVALID, rerr);
    if (shm != -1)
        return err(ATT_ERR_UNSUPPORTED, rerr + ": neighbour list in shared memory segment " +
                                            std::to_string(shm) + " is not supported");
    neighbours.resize((size_t)nrN * kNeighbourBytes);
    if (!f.get(neighbours.data(), neighbours.size())) return err(ATT_ERR_INVALID, rerr);
    int64_t innerSize = 0, bigSize = 0;
    if (!f.get(innerSize)) return err(ATT_ERR_INVALID, rerr);
    const int64_t nIn = (int64_t)dims[0] * dims[1] * dims[2];
    const int64_t nOut = (int64_t)dims[3] * dims[4] * dims[5];
    if (innerSize != nIn) return err(ATT_ERR_INVALID, rerr + ": inner grid size does not match its dimensions");
    innergrid.resize((size_t)innerSize * kVoxelBytes);
    if (!f.get(innergrid.data(), innergrid.size())) return err(ATT_ERR_INVALID, rerr);
    if (!f.get(bigSize)) return err(ATT_ERR_INVALID, rerr);
    if (bigSize != nOut) return err(ATT_ERR_INVALID, rerr + ": outer grid size does not match its dimensions");
    biggrid.resize((size_t)bigSize * kPotentialBytes);
    if (!f.get(biggrid.data(), biggrid.size())) return err(ATT_ERR_INVALID, rerr);

    // Index checks of grid_orig.cpp:245-289, plus bounds on every index used.
    auto check_pot = [&](const unsigned char* pot, int64_t n, const char* which) -> att_status {
        const bool has = u32_at(pot + 4 * kTypes) != 0;
        for (int i = 0; i <= kTypes; ++i) {
            if (i < kTypes && !I.alphabet[i]) continue;
            const int dif = (int)u32_at(pot + 4 * i) - 1;
            if ((has && (dif < 0 || dif >= nrE)) || (!has && dif >= nrE))
                return err(ATT_ERR_INVALID, "Reading error in " + name + ", " + which + " voxel " +
                                                std::to_string(n + 1) + " atom type " + std::to_string(i + 1) +
                                                ": " + std::to_string(dif) + " >= " + std::to_string(nrE));
        }
        return ATT_OK;
    };
    for (int64_t n = 0; n < innerSize; ++n) {
        const unsigned char* v = innergrid.data() + n * kVoxelBytes;
        att_status st = check_pot(v, n, "innergrid");
        if (st != ATT_OK) return st;
        int32_t nl;
        int16_t nr;
        std::memcpy(&nl, v + 400, 4);
        std::memcpy(&nr, v + 404, 2);
        const bool has = u32_at(v + 4 * kTypes) != 0;
        if ((has && ((nl == 0) != (nr == 0) || (int64_t)nl + nr - 1 > nrN)) || nr < 0 ||
            (nr > 0 && (nl < 1 || (int64_t)nl + nr - 1 > nrN)))
            return err(ATT_ERR_INVALID, "Reading error in " + name + ", innergrid voxel " + std::to_string(n + 1) +
                                            " neighbourlist: " + std::to_string(nl - 1) + " + " +
                                            std::to_string(nr) + " >= " + std::to_string(nrN));
    }
    for (int64_t n = 0; n < bigSize; ++n) {
        att_status st = check_pot(biggrid.data() + n * kPotentialBytes, n, "biggrid");
        if (st != ATT_OK) return st;
    }

    // Conversion of readFile.cpp:260-422.
    const std::vector<int> slots = slot_types(I.alphabet);
    g->numGrids = (unsigned)slots.size();
    for (int k = 0; k < 3; ++k) {
        g->nIn[k] = (unsigned)dims[k];
        g->nOut[k] = (unsigned)dims[3 + k];
        const double o = (double)ori[k];
        g->oriIn[k] = (float)o;
        g->oriOut[k] = (float)(o - I.gridextension * I.gridspacing);
    }
    g->spIn = (float)I.gridspacing;
    g->spOut = (float)(2 * I.gridspacing);
    g->plateau = (float)I.plateaudis;
    auto convert = [&](const std::vector<unsigned char>& vox, size_t stride, int64_t npts,
                       std::vector<float>& dst) {
        dst.assign((size_t)npts * slots.size() * 4, 0.f);
        for (size_t s = 0; s < slots.size(); ++s) {
            // the electrostatic grid (last slot) goes first, the others shift by one
            const size_t to = (s + 1) % slots.size();
            for (int64_t n = 0; n < npts; ++n) {
                const int e = (int)u32_at(vox.data() + n * stride + 4 * slots[s]);
                if (e <= 0) continue;
                const unsigned char* r = energrads.data() + (size_t)(e - 1) * kEnerGradBytes;
                float* d = dst.data() + (to * (size_t)npts + n) * 4;
                std::memcpy(d + 0, r + 4, 12);  // grad[3]
                std::memcpy(d + 3, r, 4);       // energy
            }
        }
    };
    convert(innergrid, kVoxelBytes, innerSize, g->inner);
    convert(biggrid, kPotentialBytes, bigSize, g->outer);
    g->list.resize(nrN);
    for (int32_t n = 0; n < nrN; ++n) g->list[n] = u32_at(neighbours.data() + (size_t)n * kNeighbourBytes + 4);
    g->desc.resize((size_t)innerSize * 2);
    for (int64_t n = 0; n < innerSize; ++n) {
        int32_t nl;
        int16_t nr;
        std::memcpy(&nl, innergrid.data() + n * kVoxelBytes + 400, 4);
        std::memcpy(&nr, innergrid.data() + n * kVoxelBytes + 404, 2);
        g->desc[2 * n] = (uint32_t)nr;
        g->desc[2 * n + 1] = (uint32_t)(nl - 1);  // 1-based in the file
    }
    *out = g.release();
    return ATT_OK;
}

att_status att_gridfile_get(const att_gridfile* g, att_gridfile_info* info, att_grid_desc* inner,
                            att_grid_desc* outer, att_nlgrid_desc* nl) {
    if (!g) return err(ATT_ERR_INVALID, "null grid file");
    if (info) *info = g->info;
    if (inner)
        *inner = att_grid_desc{g->nIn[0], g->nIn[1], g->nIn[2], g->numGrids, g->oriIn[0], g->oriIn[1],
                               g->oriIn[2], g->spIn, g->inner.data()};
    if (outer)
        *outer = att_grid_desc{g->nOut[0], g->nOut[1], g->nOut[2], g->numGrids, g->oriOut[0],
                               g->oriOut[1], g->oriOut[2], g->spOut, g->outer.data()};
    if (nl)
        *nl = att_nlgrid_desc{g->nIn[0], g->nIn[1], g->nIn[2], g->oriIn[0], g->oriIn[1], g->oriIn[2],
                              g->spIn, g->plateau, g->desc.data(),
                              g->list.empty() ? nullptr : g->list.data(), g->list.size()};
    return ATT_OK;
}

att_status att_gridfile_write(const char* path, const att_gridfile_info* info,
                              const att_grid_desc* inner, const att_grid_desc* outer,
                              const att_nlgrid_desc* nl, const float* rec_xyz) {
    if (!path || !info || !inner || !outer || !nl || !inner->data || !outer->data || !nl->desc)
        return err(ATT_ERR_INVALID, "null argument");
    const std::vector<int> slots = slot_types(info->alphabet);
    const unsigned G = (unsigned)slots.size();
    if (inner->numGrids != G || outer->numGrids != G)
        return err(ATT_ERR_INVALID, "numGrids must be the number of alphabet types + 1");
    const double sp = info->gridspacing;
    if (inner->spacing != (float)sp || outer->spacing != (float)(2 * sp))
        return err(ATT_ERR_INVALID, "spacings must be gridspacing (inner) and 2 x gridspacing (outer)");
    const float io[3] = {inner->ox, inner->oy, inner->oz};
    const float oo[3] = {outer->ox, outer->oy, outer->oz};
    for (int k = 0; k < 3; ++k)
        if (oo[k] != (float)((double)io[k] - info->gridextension * sp))
            return err(ATT_ERR_INVALID, "outer origin must be inner origin - gridextension x gridspacing");
    if (nl->nx != inner->nx || nl->ny != inner->ny || nl->nz != inner->nz || nl->ox != inner->ox ||
        nl->oy != inner->oy || nl->oz != inner->oz || nl->spacing != inner->spacing)
        return err(ATT_ERR_INVALID, "the neighbour grid must have the inner grid's geometry");
    const int64_t nIn = (int64_t)inner->nx * inner->ny * inner->nz;
    const int64_t nOut = (int64_t)outer->nx * outer->ny * outer->nz;
    const int64_t nrE64 = (nIn + nOut) * G;
    if (nrE64 >= (int64_t)1 << 31 || nl->numEl >= ((size_t)1 << 31))
        return err(ATT_ERR_UNSUPPORTED, "grid too large for the file format");
    if (nl->numEl && !nl->list) return err(ATT_ERR_INVALID, "neighbour list missing");
    for (int64_t n = 0; n < nIn; ++n) {
        const uint64_t c = nl->desc[2 * n], s = nl->desc[2 * n + 1];
        if (c > 32767) return err(ATT_ERR_UNSUPPORTED, "more than 32767 neighbours at one voxel");
        if (c && s + c > nl->numEl) return err(ATT_ERR_INVALID, "neighbour descriptor outside the list");
    }
    File f(path, "wb");
    if (!f.f) return err(ATT_ERR_INVALID, std::string("Grid::write error for ") + path +
                                              ": Cannot open file for writing");
    bool ok = true;
    const unsigned char torque = 0;
    const uint16_t arch = 64;
    ok = ok && f.put(torque) && f.put(arch) && f.put(info->gridspacing) && f.put(info->gridextension) &&
         f.put(info->plateaudis) && f.put(info->neighbourdis);
    unsigned char alpha[kTypes];
    for (int t = 0; t < kTypes; ++t) alpha[t] = info->alphabet[t] ? 1 : 0;
    const int32_t dims[6] = {(int32_t)inner->nx, (int32_t)inner->ny, (int32_t)inner->nz,
                             (int32_t)outer->nx, (int32_t)outer->ny, (int32_t)outer->nz};
    ok = ok && f.put(alpha, kTypes) && f.put(io, sizeof(io)) && f.put(dims, sizeof(dims)) &&
         f.put(info->natoms) && f.put(info->pivot, sizeof(info->pivot));
    // energy/gradient records in voxel order (inner, then outer), types in slot order
    const int32_t nrE = (int32_t)nrE64, noShm = -1;
    ok = ok && f.put(nrE) && f.put(noShm);
    auto records = [&](const att_grid_desc* g, int64_t npts) {
        std::vector<unsigned char> buf((size_t)G * kEnerGradBytes);
        for (int64_t n = 0; n < npts && ok; ++n) {
            for (unsigned s = 0; s < G; ++s) {
                const unsigned from = (s + 1) % G;  // slot s of the file = grid (s+1)%G here
                const float* v = g->data + ((size_t)from * npts + n) * 4;
                unsigned char* r = buf.data() + (size_t)s * kEnerGradBytes;
                std::memcpy(r, v + 3, 4);
                std::memcpy(r + 4, v, 12);
            }
            ok = f.put(buf.data(), buf.size());
        }
    };
    records(inner, nIn);
    records(outer, nOut);
    const int32_t nrN = (int32_t)nl->numEl;
    ok = ok && f.put(nrN) && f.put(noShm);
    {
        const double P2 = info->plateaudis * info->plateaudis;
        std::vector<unsigned char> nb((size_t)nrN * kNeighbourBytes, 0);
        for (int64_t n = 0; n < nIn; ++n) {
            const uint32_t c = nl->desc[2 * n], s = nl->desc[2 * n + 1];
            if (!c) continue;
            const int64_t ix = n % inner->nx, iy = (n / inner->nx) % inner->ny, iz = n / ((int64_t)inner->nx * inner->ny);
            const double p[3] = {(double)io[0] + ix * (double)inner->spacing,
                                 (double)io[1] + iy * (double)inner->spacing,
                                 (double)io[2] + iz * (double)inner->spacing};
            for (uint32_t k = 0; k < c; ++k) {
                const uint32_t j = nl->list[s + k];
                unsigned char type = 2;
                if (rec_xyz) {
                    if (info->natoms <= 0 || j >= (uint32_t)info->natoms)
                        return err(ATT_ERR_INVALID, "neighbour index beyond natoms");
                    const size_t na = (size_t)info->natoms;
                    const double dx = p[0] - rec_xyz[j], dy = p[1] - rec_xyz[na + j],
                                 dz = p[2] - rec_xyz[2 * na + j];
                    type = dx * dx + dy * dy + dz * dz < P2 ? 1 : 2;
                }
                unsigned char* e = nb.data() + (size_t)(s + k) * kNeighbourBytes;
                e[0] = type;
                std::memcpy(e + 4, &j, 4);
            }
        }
        // entries no voxel refers to keep their index, type 2
        for (int32_t k = 0; k < nrN; ++k) {
            unsigned char* e = nb.data() + (size_t)k * kNeighbourBytes;
            if (!e[0]) {
                e[0] = 2;
                std::memcpy(e + 4, &nl->list[k], 4);
            }
        }
        ok = ok && f.put(nb.data(), nb.size());
    }
    // potentials: 1-based record indices in voxel order
    int64_t rec = 1;
    auto potential = [&](unsigned char* pot) {
        std::memset(pot, 0, kPotentialBytes);
        for (unsigned s = 0; s < G; ++s) {
            const uint32_t v = (uint32_t)rec++;
            std::memcpy(pot + 4 * slots[s], &v, 4);
        }
    };
    const int64_t innerSize = nIn, bigSize = nOut;
    ok = ok && f.put(innerSize);
    {
        std::vector<unsigned char> v(kVoxelBytes);
        for (int64_t n = 0; n < nIn && ok; ++n) {
            std::memset(v.data(), 0, kVoxelBytes);
            potential(v.data());
            const uint32_t c = nl->desc[2 * n];
            const int32_t start = c ? (int32_t)nl->desc[2 * n + 1] + 1 : 0;
            const int16_t cnt = (int16_t)c;
            std::memcpy(v.data() + 400, &start, 4);
            std::memcpy(v.data() + 404, &cnt, 2);
            ok = f.put(v.data(), kVoxelBytes);
        }
    }
    ok = ok && f.put(bigSize);
    {
        std::vector<unsigned char> v(kPotentialBytes);
        for (int64_t n = 0; n < nOut && ok; ++n) {
            potential(v.data());
            ok = f.put(v.data(), kPotentialBytes);
        }
    }
    if (!ok) return err(ATT_ERR_INVALID, std::string("Grid::write error for ") + path);
    return ATT_OK;
}

}  // extern "C"
