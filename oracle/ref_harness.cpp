// ref_harness.cpp — TEST INFRASTRUCTURE ONLY.
//
// A driver, written for this project, that runs the reference's own CPU code
// (compiled from its sources under the reference tree by oracle/Makefile into
// oracle/_ref/libattract_ref.so) on in-memory inputs, so that the CPU
// restatement in oracle.c can be checked against it.  Nothing here is shipped
// or called by the product; tests/golden/make_golden.py uses it to write the
// committed golden vectors.
//
//  rigid scoring  : the reference's CPUEnergyService6D<float> item processor
//                   (service/CPUEnergyService6D.tpp:62-204) on a WorkItem
//  modes scoring  : the model functions in the order of
//                   service/CPUEnergyService6DModes.tpp:121-327 (the service
//                   object itself sizes its buffers crosswise and cannot run
//                   with n_rec != n_lig)
//  param table    : fileIO/readFile.cpp:439-527 on a written attract.par
//  plateau ratios : model/NLGrid.tpp:44-51
//  file formats   : reduced PDB, mode files, attract.par, DOF files through
//                   readFile.cpp's readers, DOFConverter and
//                   transformDOF_glob2rec, and the result operator<<
//  grid files     : fileIO/readFile.cpp:260-422 (createGridFromGridFile over
//                   attract::Grid::read_std) on a file written by this
//                   project's att_gridfile_write
//  scoring from files: the reference's own chain of the scoring app
//                   (service/Configurator_6D.tpp:31-141: readers, type
//                   mapping from the grid alphabet, pivots, DOF transform,
//                   grid displacement, CPUEnergyService6D behind as::Server),
//                   as test/app/ScoringTest.cpp drives it; for flexible
//                   scoring the input steps of Configurator_6D_Modes.tpp and
//                   the modes model chain
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <type_traits>
#include <thread>
#include <unistd.h>
#include <vector>

#include <sstream>

#include "CPUEnergyService6D.h"
#include "CmdArgs.h"
#include "Configurator_6D.h"
#include "DOFConverter.h"
#include "DOFTransform.h"
#include "DataManager.h"
#include "GridUnion.h"
#include "IntrplGrid.h"
#include "NLGrid.h"
#include "ParamTable.h"
#include "Protein.h"
#include "SimParam.h"
#include "Types_6D.h"
#include "Types_6D_Modes.h"
#include "WorkItem.h"
#include "interpolation.h"
#include "matrixFunctions.h"
#include "neighborlist_modes.h"
#include "readFile.h"
#include "Request.h"
#include "Server.h"
#include "reduction_modes.h"
#include "transform_modes.h"

using namespace as;

namespace {

struct GridIn {
    unsigned nx, ny, nz, ngrids;
    float ox, oy, oz, spacing;
    const float* data;
};

struct NLIn {
    unsigned nx, ny, nz;
    float ox, oy, oz, spacing, plateau;
    const uint32_t* desc;
    const uint32_t* list;
    uint64_t nlist;
};

struct BodyIn {
    unsigned n, M;
    const float* xyz;
    const unsigned* type;
    const unsigned* mtype;
    const float* charge;
    const float* modes;
    const float* modeForce;
};

struct TableIn {
    unsigned T;
    int shape, dielec;
    float epsilon, ffelec;
    const float* rc;
    const float* ac;
    const int* ipon;
    const float* rmin2;
    const float* emin;
};

std::shared_ptr<Protein<float>> make_protein(const BodyIn& b) {
    auto p = std::make_shared<Protein<float>>();
    p->setNumAtoms(b.n);
    std::copy(b.xyz, b.xyz + 3 * (size_t)b.n, p->getOrCreatePosPtr());
    std::copy(b.type, b.type + b.n, p->getOrCreateTypePtr());
    p->setNumMappedTypes(1);
    std::copy(b.mtype, b.mtype + b.n, p->getOrCreateMappedPtr());
    std::copy(b.charge, b.charge + b.n, p->getOrCreateChargePtr());
    if (b.M > 0) {
        p->setNumModes(b.M);
        std::copy(b.modes, b.modes + 3 * (size_t)b.n * b.M, p->getOrCreateModePtr());
        std::copy(b.modeForce, b.modeForce + b.M, p->getOrCreateModeForcePtr());
    }
    return p;
}

std::shared_ptr<IntrplGrid<float>> make_igrid(const GridIn& g) {
    IntrplGrid<float>::Desc d;
    std::memset(d.typemask, 0, sizeof(d.typemask));
    d.numGrids = g.ngrids;
    d.dimN.x = g.nx; d.dimN.y = g.ny; d.dimN.z = g.nz;
    d.gridSpacing = g.spacing;
    d.posMin.x = g.ox; d.posMin.y = g.oy; d.posMin.z = g.oz;
    const size_t cnt = (size_t)g.ngrids * g.nx * g.ny * g.nz;
    d.grid = new float4[cnt];
    std::memcpy(d.grid, g.data, cnt * sizeof(float4));
    return std::make_shared<IntrplGrid<float>>(d);
}

std::shared_ptr<NLGrid<float>> make_nlgrid(const NLIn& g) {
    NLGrid<float>::Desc d;
    d.dimN.x = g.nx; d.dimN.y = g.ny; d.dimN.z = g.nz;
    d.gridSpacing = g.spacing;
    d.dPlateau = g.plateau;
    d.posMin.x = g.ox; d.posMin.y = g.oy; d.posMin.z = g.oz;
    const size_t nv = (size_t)g.nx * g.ny * g.nz;
    d.grid = new NeighbourDesc[nv];
    for (size_t v = 0; v < nv; ++v) { d.grid[v].numEl = g.desc[2 * v]; d.grid[v].idx = g.desc[2 * v + 1]; }
    d.numEl = (unsigned)g.nlist;
    d.neighborArray = new unsigned[g.nlist > 0 ? g.nlist : 1];
    std::copy(g.list, g.list + g.nlist, d.neighborArray);
    return std::make_shared<NLGrid<float>>(d);
}

std::shared_ptr<ParamTable<float>> make_table(const TableIn& t) {
    auto p = std::make_shared<ParamTable<float>>();
    p->setNumTypes(t.T);
    p->setPotShape(t.shape == 8 ? PotShape::_8_6 : PotShape::_12_6);
    auto* tab = p->getOrCreateTable();
    for (unsigned k = 0; k < t.T * t.T; ++k) {
        tab[k].rc = t.rc[k];
        tab[k].ac = t.ac[k];
        tab[k].ipon = t.ipon[k];
        tab[k].rmin2 = t.rmin2[k];
        tab[k].emin = t.emin[k];
    }
    return p;
}

std::shared_ptr<SimParam<float>> make_sim(const TableIn& t) {
    auto s = std::make_shared<SimParam<float>>();
    s->dielec = t.dielec == 0 ? Dielec::constant : Dielec::variable;
    s->epsilon = t.epsilon;
    s->ffelec = t.ffelec;
    return s;
}

}  // namespace

// Overload resolution the reference's headers see for unqualified abs/sqrt on a
// float inside namespace as (model/neighborlist.h:115, model/forcefield.h:104).
namespace as {
inline int ref_abs_probe() { return std::is_same<decltype(abs(1.5f)), float>::value ? 1 : 0; }
inline int ref_sqrt_probe() { return std::is_same<decltype(sqrt(1.5f)), float>::value ? 1 : 0; }
}  // namespace as

extern "C" {

int ref_abs_is_float() { return as::ref_abs_probe(); }
int ref_sqrt_is_float() { return as::ref_sqrt_probe(); }

void ref_euler2rotmat(float phi, float ssi, float rot, float* R) {
    auto m = euler2rotmat<float>(phi, ssi, rot);
    for (int i = 0; i < 9; ++i) R[i] = m.mat[i];
}

void ref_euler2torquemat(float phi, float ssi, float rot, float* out) {
    auto t = euler2torquemat<float>(phi, ssi, rot);
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l) out[9 * k + 3 * j + l] = t.mat[k][j][l];
}

unsigned ref_ratio_lut(float plateau, float* out, unsigned cap) {
    NLIn g{2, 2, 2, 0.f, 0.f, 0.f, 1.f, plateau, nullptr, nullptr, 0};
    uint32_t desc[16] = {0};
    g.desc = desc;
    auto nl = make_nlgrid(g);
    const unsigned size = (unsigned)int(10000 * nl->dPlateau2()) + 1;
    for (unsigned n = 0; n < size && n < cap; ++n) out[n] = nl->_ratio[n];
    return size;
}

// readParamTableFromFile on a temporary attract.par holding (rbc, abc, ipon).
int ref_read_params(unsigned T, int shape, const float* rbc, const float* abc, const int* ipon,
                    float* rc, float* ac, int* ipon_out, float* rmin2, float* emin) {
    char path[] = "/tmp/attract_par_XXXXXX";
    int fd = mkstemp(path);
    if (fd < 0) return -1;
    FILE* f = fdopen(fd, "w");
    std::fprintf(f, "%d %u 0 0\n", shape, T);
    for (unsigned k = 0; k < T * T; ++k) std::fprintf(f, "%.9g%c", rbc[k], (k % T == T - 1) ? '\n' : ' ');
    for (unsigned k = 0; k < T * T; ++k) std::fprintf(f, "%.9g%c", abc[k], (k % T == T - 1) ? '\n' : ' ');
    for (unsigned k = 0; k < T * T; ++k) std::fprintf(f, "%d%c", ipon[k], (k % T == T - 1) ? '\n' : ' ');
    std::fclose(f);
    auto tab = std::make_shared<ParamTable<float>>();
    readParamTableFromFile<float>(tab, path);
    unlink(path);
    for (unsigned k = 0; k < T * T; ++k) {
        const auto& p = tab->table()[k];
        rc[k] = p.rc; ac[k] = p.ac; ipon_out[k] = p.ipon; rmin2[k] = p.rmin2; emin[k] = p.emin;
    }
    return 0;
}

// Rigid scoring through the reference's CPU item processor.
int ref_score_rigid(const BodyIn* rec, const BodyIn* lig, const GridIn* inner, const GridIn* outer,
                    const NLIn* nl, const TableIn* tab, const float* dof, size_t n, float* res) {
    auto dm = std::make_shared<DataManager>();
    auto gu = std::make_shared<GridUnion<float>>(make_igrid(*inner), make_igrid(*outer), make_nlgrid(*nl));
    Common common;
    common.gridId = dm->add(gu);
    common.ligId = dm->add(make_protein(*lig));
    common.recId = dm->add(make_protein(*rec));
    common.tableId = dm->add(make_table(*tab));
    common.paramsId = dm->add(make_sim(*tab));
    CPUEnergyService6D<float> service(dm);
    auto proc = service.createItemProcessor();
    std::vector<DOF_6D<float>> in(n);
    std::vector<Result_6D<float>> out(n);
    for (size_t i = 0; i < n; ++i) {
        in[i].pos = Vec3<float>(dof[6 * i], dof[6 * i + 1], dof[6 * i + 2]);
        in[i].ang = Vec3<float>(dof[6 * i + 3], dof[6 * i + 4], dof[6 * i + 5]);
    }
    for (size_t lo = 0; lo < n; lo += 100) {
        unsigned sz = (unsigned)std::min<size_t>(100, n - lo);
        WorkItem<DOF_6D<float>, Common, Result_6D<float>> item(in.data() + lo, sz, &common, out.data() + lo);
        proc(&item);
        if (!item.isProcessed()) return -1;
    }
    for (size_t i = 0; i < n; ++i) {
        res[7 * i] = out[i].E;
        res[7 * i + 1] = out[i].pos.x; res[7 * i + 2] = out[i].pos.y; res[7 * i + 3] = out[i].pos.z;
        res[7 * i + 4] = out[i].ang.x; res[7 * i + 5] = out[i].ang.y; res[7 * i + 6] = out[i].ang.z;
    }
    return 0;
}

// Per-bead forces of one rigid pose: rotate_translate -> potForce -> NLPotForce.
int ref_beads_rigid(const BodyIn* rec, const BodyIn* lig, const GridIn* inner, const GridIn* outer,
                    const NLIn* nl, const TableIn* tab, const float* dof, float* xyz, float* pot) {
    auto pr = make_protein(*rec);
    auto pl = make_protein(*lig);
    auto gi = make_igrid(*inner);
    auto go = make_igrid(*outer);
    auto gn = make_nlgrid(*nl);
    auto pt = make_table(*tab);
    auto sp = make_sim(*tab);
    const unsigned N = pl->numAtoms();
    Vec3<float> pos(dof[0], dof[1], dof[2]), ang(dof[3], dof[4], dof[5]);
    rotate_translate(pl->xPos(), pl->yPos(), pl->zPos(), pos, ang, N, xyz, xyz + N, xyz + 2 * N);
    potForce(gi.get(), go.get(), pl.get(), xyz, xyz + N, xyz + 2 * N, pot, pot + N, pot + 2 * N, pot + 3 * N);
    NLPotForce(gn.get(), pr.get(), pl.get(), sp.get(), pt.get(), xyz, xyz + N, xyz + 2 * N, pot,
               pot + N, pot + 2 * N, pot + 3 * N);
    return 0;
}

// Grid cells and neighbour-list pair counts of one rigid pose, from the
// reference's own index functions on its own transformed coordinates:
// per bead 9 ints {inner x,y,z; outer x,y,z; NL x,y,z} (-1 where the bead has
// no such lookup: mapped type 0 or out of bounds for the grids as potForce
// branches (interpolation.h:56-102), type 0 or out of bounds for the list as
// NLPotForce (neighborlist.h:47-63)), and the number of listed neighbours
// inside the plateau (the pairs NLPotForce evaluates, neighborlist.h:66-80).
int ref_cells_rigid(const BodyIn* rec, const BodyIn* lig, const GridIn* inner, const GridIn* outer,
                    const NLIn* nl, const float* dof, int* cell, unsigned* npairs) {
    auto pr = make_protein(*rec);
    auto pl = make_protein(*lig);
    auto gi = make_igrid(*inner);
    auto go = make_igrid(*outer);
    auto gn = make_nlgrid(*nl);
    const unsigned N = pl->numAtoms();
    std::vector<float> xyz(3 * N);
    Vec3<float> pos(dof[0], dof[1], dof[2]), ang(dof[3], dof[4], dof[5]);
    rotate_translate(pl->xPos(), pl->yPos(), pl->zPos(), pos, ang, N, xyz.data(), xyz.data() + N,
                     xyz.data() + 2 * N);
    for (unsigned i = 0; i < N; ++i) {
        int* c = cell + 9 * i;
        for (int k = 0; k < 9; ++k) c[k] = -1;
        npairs[i] = 0;
        const float x = xyz[i], y = xyz[N + i], z = xyz[2 * N + i];
        const float3 p = make_float3(x, y, z);
        if (pl->mappedType()[i] != 0) {
            if (gi->outOfBounds_byPos(p)) {
                if (!go->outOfBounds_byPos(p)) {
                    const int3 ix = go->getIndex(p);
                    c[3] = ix.x; c[4] = ix.y; c[5] = ix.z;
                }
            } else {
                const int3 ix = gi->getIndex(p);
                c[0] = ix.x; c[1] = ix.y; c[2] = ix.z;
            }
        }
        if (pl->type()[i] != 0 && !gn->outOfBounds(x, y, z)) {
            int ix, iy, iz;
            gn->getIndex(x, y, z, ix, iy, iz);
            c[6] = ix; c[7] = iy; c[8] = iz;
            const NeighbourDesc& d = gn->getNeighbourDesc(ix, iy, iz);
            unsigned np = 0;
            for (unsigned j = 0; j < d.numEl; ++j) {
                const unsigned r = gn->getNeighbor(d.idx + j);
                const float dx = x - pr->xPos()[r], dy = y - pr->yPos()[r], dz = z - pr->zPos()[r];
                if (!gn->outOfPlateau(dx * dx + dy * dy + dz * dz)) ++np;
            }
            npairs[i] = np;
        }
    }
    return 0;
}

// Modes scoring: model functions in the order of CPUEnergyService6DModes.tpp:121-327,
// buffers sized by their own protein.  res = 27 floats per pose.
struct ModesSetup {
    std::shared_ptr<Protein<float>> pr, pl;
    GridUnion<float> gridRec, gridLig;
    std::shared_ptr<ParamTable<float>> pt;
    std::shared_ptr<SimParam<float>> sp;
    ModesSetup(const BodyIn* rec, const BodyIn* lig, const GridIn* recInner, const GridIn* recOuter,
               const NLIn* recNL, const GridIn* ligInner, const GridIn* ligOuter, const NLIn* ligNL,
               const TableIn* tab)
        : pr(make_protein(*rec)), pl(make_protein(*lig)),
          gridRec(make_igrid(*recInner), make_igrid(*recOuter), make_nlgrid(*recNL)),
          gridLig(make_igrid(*ligInner), make_igrid(*ligOuter), make_nlgrid(*ligNL)),
          pt(make_table(*tab)), sp(make_sim(*tab)) {}
    ModesSetup(std::shared_ptr<Protein<float>> rec, std::shared_ptr<Protein<float>> lig,
               const GridUnion<float>& gRec, const GridUnion<float>& gLig,
               std::shared_ptr<ParamTable<float>> table, std::shared_ptr<SimParam<float>> sim)
        : pr(rec), pl(lig), gridRec(gRec.inner, gRec.outer, gRec.NL),
          gridLig(gLig.inner, gLig.outer, gLig.NL), pt(table), sp(sim) {}
};

static void score_modes_range(const ModesSetup& S, const float* dof, size_t lo, size_t hi, float* res,
                              float* recPotOut, float* ligPotOut) {
    Protein<float>* pr = S.pr.get();
    Protein<float>* pl = S.pl.get();
    const unsigned NR = pr->numAtoms(), NL = pl->numAtoms();
    std::vector<float> trafoRec(3 * NR), defoRec(3 * NR), potRec(4 * NR), trafoLig(3 * NL), potLig(4 * NL);
    for (size_t i = lo; i < hi; ++i) {
        DOF_6D_Modes<float> d;
        const float* s = dof + 26 * i;
        d._6D.pos = Vec3<float>(s[0], s[1], s[2]);
        d._6D.ang = Vec3<float>(s[3], s[4], s[5]);
        for (int m = 0; m < MODES_MAX_NUMBER; ++m) { d.modesRec[m] = s[6 + m]; d.modesLig[m] = s[16 + m]; }
        DOF_6D_Modes<float> inv = invertDOF(d);
        rotate_translate_deform(pr->xPos(), pr->yPos(), pr->zPos(), inv._6D.pos, inv._6D.ang, NR,
                                pr->numModes(), inv.modesRec, pr->xModes(), pr->yModes(), pr->zModes(),
                                &defoRec[0], &defoRec[NR], &defoRec[2 * NR], &trafoRec[0],
                                &trafoRec[NR], &trafoRec[2 * NR]);
        rotate_translate(pl->xPos(), pl->yPos(), pl->zPos(), d._6D.pos, d._6D.ang, NL, pl->numModes(),
                         d.modesLig, pl->xModes(), pl->yModes(), pl->zModes(), &trafoLig[0],
                         &trafoLig[NL], &trafoLig[2 * NL]);
        potForce(S.gridLig.inner.get(), S.gridLig.outer.get(), pr, &trafoRec[0], &trafoRec[NR],
                 &trafoRec[2 * NR], &potRec[0], &potRec[NR], &potRec[2 * NR], &potRec[3 * NR]);
        rotate_forces(inv._6D.ang.inv(), NR, &potRec[0], &potRec[NR], &potRec[2 * NR]);
        potForce(S.gridRec.inner.get(), S.gridRec.outer.get(), pl, &trafoLig[0], &trafoLig[NL],
                 &trafoLig[2 * NL], &potLig[0], &potLig[NL], &potLig[2 * NL], &potLig[3 * NL]);
        NLPotForce(S.gridRec.NL.get(), pr, pl, S.sp.get(), S.pt.get(), &defoRec[0], &defoRec[NR],
                   &defoRec[2 * NR], &trafoLig[0], &trafoLig[NL], &trafoLig[2 * NL], &potLig[0],
                   &potLig[NL], &potLig[2 * NL], &potLig[3 * NL], &potRec[0], &potRec[NR], &potRec[2 * NR]);
        PotForce_Modes<float> red = reducePotForce<float, PotForce_Modes<float>>(
            &potLig[0], &potLig[NL], &potLig[2 * NL], &potLig[3 * NL], NL);
        reduceModeForce(d._6D.ang, &potLig[0], &potLig[NL], &potLig[2 * NL], pl->xModes(), pl->yModes(),
                        pl->zModes(), NL, pl->numModes(), red.modesLig);
        correctModeForce(pl->modeForce(), pl->numModes(), red.modesLig);
        reduceModeForce(&potRec[0], &potRec[NR], &potRec[2 * NR], pr->xModes(), pr->yModes(),
                        pr->zModes(), NR, pr->numModes(), red.modesRec);
        correctModeForce(pr->modeForce(), pr->numModes(), red.modesRec);
        Vec3<float> a = reduceTorque(pl->xPos(), pl->yPos(), pl->zPos(), &potLig[0], &potLig[NL],
                                     &potLig[2 * NL], NL, d._6D.ang);
        float* o = res + 27 * i;
        o[0] = red.E;
        o[1] = red.pos.x; o[2] = red.pos.y; o[3] = red.pos.z;
        o[4] = a.x; o[5] = a.y; o[6] = a.z;
        for (int m = 0; m < MODES_MAX_NUMBER; ++m) {
            o[7 + m] = m < (int)pr->numModes() ? red.modesRec[m] : 0.f;
            o[17 + m] = m < (int)pl->numModes() ? red.modesLig[m] : 0.f;
        }
        if (recPotOut && i == 0) std::copy(potRec.begin(), potRec.begin() + 3 * NR, recPotOut);
        if (ligPotOut && i == 0) std::copy(potLig.begin(), potLig.end(), ligPotOut);
    }
}

// Worker threads taking items of 100 poses from a shared counter, as the
// reference's server feeds its workers (server/Server.h:39).  Returns the
// wall time of the scoring alone (set-up excluded) in *seconds.
typedef std::function<bool(size_t&, size_t&)> Take;

static void run_items(size_t n, int nthreads, const std::function<void(int, const Take&)>& fn,
                      double* seconds) {
    std::atomic<size_t> next{0};
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int t = 0; t < std::max(1, nthreads); ++t)
        th.emplace_back([&, t] {
            Take take = [&](size_t& lo, size_t& hi) {
                lo = next.fetch_add(100);
                if (lo >= n) return false;
                hi = std::min(n, lo + 100);
                return true;
            };
            fn(t, take);
        });
    for (auto& x : th) x.join();
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int ref_score_modes(const BodyIn* rec, const BodyIn* lig, const GridIn* recInner,
                    const GridIn* recOuter, const NLIn* recNL, const GridIn* ligInner,
                    const GridIn* ligOuter, const NLIn* ligNL, const TableIn* tab, const float* dof,
                    size_t n, float* res, float* recPotOut, float* ligPotOut) {
    ModesSetup S(rec, lig, recInner, recOuter, recNL, ligInner, ligOuter, ligNL, tab);
    score_modes_range(S, dof, 0, n, res, recPotOut, ligPotOut);
    return 0;
}

int ref_score_modes_mt(const BodyIn* rec, const BodyIn* lig, const GridIn* recInner,
                       const GridIn* recOuter, const NLIn* recNL, const GridIn* ligInner,
                       const GridIn* ligOuter, const NLIn* ligNL, const TableIn* tab,
                       const float* dof, size_t n, float* res, int nthreads, double* seconds) {
    ModesSetup S(rec, lig, recInner, recOuter, recNL, ligInner, ligOuter, ligNL, tab);
    run_items(n, nthreads, [&](int, const Take& take) {
        size_t lo, hi;
        while (take(lo, hi)) score_modes_range(S, dof, lo, hi, res, nullptr, nullptr);
    }, seconds);
    return 0;
}

// Rigid scoring with the reference's CPUEnergyService6D, one item processor
// per worker thread (server/Server.tpp creates one per worker).
int ref_score_rigid_mt(const BodyIn* rec, const BodyIn* lig, const GridIn* inner, const GridIn* outer,
                       const NLIn* nl, const TableIn* tab, const float* dof, size_t n, float* res,
                       int nthreads, double* seconds) {
    auto dm = std::make_shared<DataManager>();
    auto gu = std::make_shared<GridUnion<float>>(make_igrid(*inner), make_igrid(*outer), make_nlgrid(*nl));
    Common common;
    common.gridId = dm->add(gu);
    common.ligId = dm->add(make_protein(*lig));
    common.recId = dm->add(make_protein(*rec));
    common.tableId = dm->add(make_table(*tab));
    common.paramsId = dm->add(make_sim(*tab));
    CPUEnergyService6D<float> service(dm);
    std::vector<DOF_6D<float>> in(n);
    std::vector<Result_6D<float>> out(n);
    for (size_t i = 0; i < n; ++i) {
        in[i].pos = Vec3<float>(dof[6 * i], dof[6 * i + 1], dof[6 * i + 2]);
        in[i].ang = Vec3<float>(dof[6 * i + 3], dof[6 * i + 4], dof[6 * i + 5]);
    }
    std::vector<decltype(service.createItemProcessor())> procs;
    for (int t = 0; t < std::max(1, nthreads); ++t) procs.push_back(service.createItemProcessor());
    std::atomic<int> bad{0};
    run_items(n, nthreads, [&](int t, const Take& take) {
        size_t lo, hi;
        while (take(lo, hi)) {
            WorkItem<DOF_6D<float>, Common, Result_6D<float>> item(in.data() + lo, (unsigned)(hi - lo),
                                                                  &common, out.data() + lo);
            procs[t](&item);
            if (!item.isProcessed()) bad = 1;
        }
    }, seconds);
    for (size_t i = 0; i < n; ++i) {
        res[7 * i] = out[i].E;
        res[7 * i + 1] = out[i].pos.x; res[7 * i + 2] = out[i].pos.y; res[7 * i + 3] = out[i].pos.z;
        res[7 * i + 4] = out[i].ang.x; res[7 * i + 5] = out[i].ang.y; res[7 * i + 6] = out[i].ang.z;
    }
    return bad ? -1 : 0;
}


// ATTRACT grid file through the reference's reader.  Only valid files may be
// given: the reference's reader exits the process on a malformed one.
static std::shared_ptr<GridUnion<float>> g_file_grid;

int ref_gridfile_load(const char* path) {
    g_file_grid = createGridFromGridFile<float>(std::string(path));
    return g_file_grid ? 0 : -1;
}

// dims: inner nx ny nz numGrids, outer nx ny nz numGrids, neighbour nx ny nz;
// geo: inner x y z spacing, outer x y z spacing, neighbour x y z spacing plateau^2
void ref_gridfile_info(unsigned* dims, float* geo, uint64_t* numEl) {
    auto& in = *g_file_grid->inner;
    auto& out = *g_file_grid->outer;
    auto& nl = *g_file_grid->NL;
    const unsigned d[11] = {in.dimN().x, in.dimN().y, in.dimN().z, in.numTypes(),
                            out.dimN().x, out.dimN().y, out.dimN().z, out.numTypes(),
                            nl.dimN().x, nl.dimN().y, nl.dimN().z};
    std::copy(d, d + 11, dims);
    const float g[13] = {in.pos().x, in.pos().y, in.pos().z, in.dVox(),
                         out.pos().x, out.pos().y, out.pos().z, out.dVox(),
                         nl.pos().x, nl.pos().y, nl.pos().z, nl.dVox(), nl.dPlateau2()};
    std::copy(g, g + 13, geo);
    *numEl = nl.neighborListSize();
}

void ref_gridfile_copy(float* inner, float* outer, uint32_t* desc, uint32_t* list) {
    auto& in = *g_file_grid->inner;
    auto& out = *g_file_grid->outer;
    auto& nl = *g_file_grid->NL;
    const size_t ni = (size_t)in.dimN().x * in.dimN().y * in.dimN().z * in.numTypes();
    const size_t no = (size_t)out.dimN().x * out.dimN().y * out.dimN().z * out.numTypes();
    std::memcpy(inner, in.getHostGridPtr(0), 16 * ni);
    std::memcpy(outer, out.getHostGridPtr(0), 16 * no);
    const size_t nv = (size_t)nl.dimN().x * nl.dimN().y * nl.dimN().z;
    for (size_t v = 0; v < nv; ++v) {
        desc[2 * v] = nl.grid()[v].numEl;
        desc[2 * v + 1] = nl.grid()[v].idx;
    }
    std::memcpy(list, nl.neighborList(), 4 * (size_t)nl.neighborListSize());
}

void ref_gridfile_free() { g_file_grid.reset(); }

// Reduced PDB through createProteinFromPDB<float>.
int ref_pdb(const char* path, unsigned cap, unsigned* n, float* xyz, unsigned* type, float* charge) {
    auto p = createProteinFromPDB<float>(std::string(path));
    *n = p->numAtoms();
    if (*n > cap) return -1;
    const unsigned N = *n;
    std::copy(p->xPos(), p->xPos() + N, xyz);
    std::copy(p->yPos(), p->yPos() + N, xyz + N);
    std::copy(p->zPos(), p->zPos() + N, xyz + 2 * N);
    std::copy(p->type(), p->type() + N, type);
    std::copy(p->charge(), p->charge() + N, charge);
    return 0;
}

// Normal modes through readHMMode<float> on the body of `pdb`.
int ref_modes(const char* pdb, const char* modefile, unsigned M, float* modes, float* modeForce) {
    auto p = createProteinFromPDB<float>(std::string(pdb));
    p->setNumModes(M);
    readHMMode<float>(p, std::string(modefile));
    const size_t nm = (size_t)p->numAtoms() * M;
    std::copy(p->xModes(), p->xModes() + nm, modes);
    std::copy(p->yModes(), p->yModes() + nm, modes + nm);
    std::copy(p->zModes(), p->zModes() + nm, modes + 2 * nm);
    std::copy(p->modeForce(), p->modeForce() + M, modeForce);
    return 0;
}

// attract.par through createParamTableFromFile<float>: T*T x {rc, ac, ipon, rmin2, emin}.
int ref_params_file(const char* path, unsigned cap, unsigned* T, float* rc, float* ac, int* ipon,
                    float* rmin2, float* emin) {
    auto tab = createParamTableFromFile<float>(std::string(path));
    *T = tab->numTypes();
    if ((size_t)*T * *T > cap) return -1;
    for (unsigned k = 0; k < *T * *T; ++k) {
        const auto& q = tab->table()[k];
        rc[k] = q.rc; ac[k] = q.ac; ipon[k] = q.ipon; rmin2[k] = q.rmin2; emin[k] = q.emin;
    }
    return 0;
}

// DOF file -> ligand DOFs in the receptor frame, as Configurator_6D[_Modes]
// does (service/Configurator_6D.tpp:56-100): header, readDOF, DOFConverter,
// pivots (explicit, or the given auto pivots), transformDOF_glob2rec.
int ref_dof(const char* path, int modes, unsigned Mr, unsigned Ml, const float* autoRec,
            const float* autoLig, size_t cap, size_t* n, float* out) {
    DOFHeader<float> h = readDOFHeader<float>(std::string(path));
    auto raw = readDOF(std::string(path));
    if (raw.size() != 2) return -1;
    if (h.auto_pivot) {
        h.pivots.push_back(Vec3<float>(autoRec[0], autoRec[1], autoRec[2]));
        h.pivots.push_back(Vec3<float>(autoLig[0], autoLig[1], autoLig[2]));
    }
    if (h.pivots.size() != 2) return -2;
    if (!modes) {
        auto m = DOFConverter<float>(raw);
        transformDOF_glob2rec(m[0], m[1], h.pivots[0], h.pivots[1], h.centered_receptor, h.centered_ligands);
        *n = m[1].size();
        if (*n > cap) return -3;
        for (size_t i = 0; i < *n; ++i) {
            float* d = out + 6 * i;
            d[0] = m[1][i].pos.x; d[1] = m[1][i].pos.y; d[2] = m[1][i].pos.z;
            d[3] = m[1][i].ang.x; d[4] = m[1][i].ang.y; d[5] = m[1][i].ang.z;
        }
    } else {
        Common_Modes::numModesRec = Mr;
        Common_Modes::numModesLig = Ml;
        auto m = DOFConverter_Modes<float>(raw);
        transformDOF_glob2rec(m[0], m[1], h.pivots[0], h.pivots[1], h.centered_receptor, h.centered_ligands);
        *n = m[1].size();
        if (*n > cap) return -3;
        for (size_t i = 0; i < *n; ++i) {
            float* d = out + 26 * i;
            const auto& q = m[1][i];
            d[0] = q._6D.pos.x; d[1] = q._6D.pos.y; d[2] = q._6D.pos.z;
            d[3] = q._6D.ang.x; d[4] = q._6D.ang.y; d[5] = q._6D.ang.z;
            for (unsigned k = 0; k < 10; ++k) {
                d[6 + k] = k < Mr ? q.modesRec[k] : 0.f;
                d[16 + k] = k < Ml ? q.modesLig[k] : 0.f;
            }
        }
    }
    return 0;
}

// A grid alphabet file through readGridAlphabetFromFile (readFile.cpp:901-926).
int ref_alphabet(const char* path, unsigned* out, unsigned cap, unsigned* n) {
    const std::vector<unsigned> v = readGridAlphabetFromFile(path);
    *n = (unsigned)v.size();
    if (v.size() > cap) return -1;
    std::copy(v.begin(), v.end(), out);
    return 0;
}

// Rigid scoring from files, as the scoring app and test/app/ScoringTest.cpp:
// Configurator_6D<float>::init on the given files (one CPU worker,
// CPUEnergyService6D), one Request over every DOF of the file through
// as::Server, results {E, pos, ang} per pose.  dielec "variable" or
// "constant", epsilon as the command line's.
int ref_score_files(const char* dof, const char* rec, const char* lig, const char* grid,
                    const char* par, const char* alphabet, const char* dielec, double epsilon,
                    float* res, size_t cap, size_t* n) {
    CmdArgs args;
    args.app = AppType::SCORE;
    args.dofName = dof;
    args.recName = rec;
    args.ligName = lig;
    args.gridRecName = grid;
    args.paramsName = par;
    args.alphabetRecName = alphabet;
    args.numModes = 0;
    args.numCPUs = 1;
    args.chunkSize = 1000;
    args.precision = "single";
    args.dielec = dielec;
    args.epsilon = epsilon;
    Configurator_6D<float> conf;
    conf.init(args);
    auto& dofs = conf.dofs();
    *n = dofs.size();
    if (dofs.size() > cap) return -1;
    Request<DOF_6D<float>, Common> request(dofs.data(), dofs.size(), conf.common());
    conf.server()->submit(request);
    std::vector<Result_6D<float>> out(dofs.size());
    conf.server()->wait(request, out.data());
    for (size_t i = 0; i < out.size(); ++i) {
        res[7 * i] = out[i].E;
        res[7 * i + 1] = out[i].pos.x; res[7 * i + 2] = out[i].pos.y; res[7 * i + 3] = out[i].pos.z;
        res[7 * i + 4] = out[i].ang.x; res[7 * i + 5] = out[i].ang.y; res[7 * i + 6] = out[i].ang.z;
    }
    return 0;
}

// Flexible scoring from files: the input steps of
// Configurator_6D_Modes<float>::init (service/Configurator_6D_Modes.tpp:
// readers, the ligand's types mapped with the receptor grid's alphabet,
// DOFConverter_Modes, pivots, transformDOF_glob2rec, the receptor grid
// displaced by the receptor pivot, numModes modes read for both bodies, the
// receptor's types mapped with the ligand grid's alphabet, the ligand grid
// displaced by the ligand pivot), each through the reference's own function,
// then the modes model chain above (the order of
// CPUEnergyService6DModes.tpp:121-327).  The service object itself is not
// used: it sizes its worker buffers crosswise and fails its WorkerBuffer
// index assertion.  One change of order: Common_Modes::numModesRec/Lig are
// set before DOFConverter_Modes, which reads them (the configurator sets them
// only after converting, so the app converts with their previous values, 0 in
// a fresh process, and reads every amplitude as 0).  Results {E, pos, ang,
// modesRec[10], modesLig[10]}.
int ref_score_files_modes(const char* dof, const char* rec, const char* lig, const char* gridRec,
                          const char* gridLig, const char* par, const char* alphabetRec,
                          const char* alphabetLig, const char* recModes, const char* ligModes,
                          int numModes, const char* dielec, double epsilon, float* res, size_t cap,
                          size_t* n) {
    auto receptor = createProteinFromPDB<float>(rec);
    auto ligand = createProteinFromPDB<float>(lig);
    auto paramTable = createParamTableFromFile<float>(par);
    auto gRec = createGridFromGridFile<float>(gridRec);
    auto simParam = std::make_shared<SimParam<float>>();
    simParam->dielec = std::string(dielec) == "variable" ? Dielec::variable : Dielec::constant;
    simParam->epsilon = static_cast<float>(epsilon);
    simParam->ffelec = static_cast<float>(FELEC / epsilon);
    TypeMap mapRec = createTypeMapFromVector(readGridAlphabetFromFile(alphabetRec));
    ligand->setNumMappedTypes(1);
    ligand->getOrCreateMappedPtr();
    applyDefaultMapping(ligand->numAtoms(), ligand->type(), ligand->type());
    applyMapping(mapRec, ligand->numAtoms(), ligand->type(), ligand->mappedType());
    DOFHeader<float> h = readDOFHeader<float>(dof);
    std::vector<std::vector<DOF>> raw = readDOF(dof);
    Common_Modes::numModesLig = numModes;
    Common_Modes::numModesRec = numModes;
    std::vector<std::vector<DOF_6D_Modes<float>>> mol = DOFConverter_Modes<float>(raw);
    if (mol.size() != 2) return -2;
    if (h.auto_pivot) {
        receptor->auto_pivotize();
        ligand->auto_pivotize();
        h.pivots.push_back(receptor->pivot());
        h.pivots.push_back(ligand->pivot());
    } else {
        if (h.pivots.size() != 2) return -2;
        receptor->pivotize(h.pivots[0]);
        ligand->pivotize(h.pivots[1]);
    }
    transformDOF_glob2rec(mol[0], mol[1], h.pivots[0], h.pivots[1], h.centered_receptor,
                          h.centered_ligands);
    gRec->translate(-make_real3(receptor->pivot().x, receptor->pivot().y, receptor->pivot().z));
    receptor->setNumModes(numModes);
    ligand->setNumModes(numModes);
    readHMMode<float>(receptor, recModes);
    readHMMode<float>(ligand, ligModes);
    TypeMap mapLig = createTypeMapFromVector(readGridAlphabetFromFile(alphabetLig));
    receptor->setNumMappedTypes(1);
    receptor->getOrCreateMappedPtr();
    applyDefaultMapping(receptor->numAtoms(), receptor->type(), receptor->type());
    applyMapping(mapLig, receptor->numAtoms(), receptor->type(), receptor->mappedType());
    auto gLig = createGridFromGridFile<float>(gridLig);
    gLig->translate(-make_real3(ligand->pivot().x, ligand->pivot().y, ligand->pivot().z));
    // the configurator's DOF buffer: pose, then the modes of each body
    const size_t np = mol[1].size();
    *n = np;
    if (np > cap) return -1;
    std::vector<float> d(26 * np, 0.f);
    for (size_t i = 0; i < np; ++i) {
        const DOF_6D_Modes<float>& q = mol[1][i];
        float* o = &d[26 * i];
        o[0] = q._6D.pos.x; o[1] = q._6D.pos.y; o[2] = q._6D.pos.z;
        o[3] = q._6D.ang.x; o[4] = q._6D.ang.y; o[5] = q._6D.ang.z;
        for (int k = 0; k < numModes; ++k) {
            o[6 + k] = q.modesRec[k];
            o[16 + k] = q.modesLig[k];
        }
    }
    ModesSetup S(receptor, ligand, *gRec, *gLig, paramTable, simParam);
    score_modes_range(S, d.data(), 0, np, res, nullptr, nullptr);
    return 0;
}

// The apps' result lines: `std::cout << res << std::endl` per pose.
size_t ref_format_results(const float* res, size_t n, int modes, unsigned Mr, unsigned Ml, char* buf,
                          size_t cap) {
    std::ostringstream s;
    for (size_t i = 0; i < n; ++i) {
        if (!modes) {
            const float* e = res + 7 * i;
            Result_6D<float> r;
            r.E = e[0];
            r.pos = Vec3<float>(e[1], e[2], e[3]);
            r.ang = Vec3<float>(e[4], e[5], e[6]);
            s << r << std::endl;
        } else {
            Common_Modes::numModesRec = Mr;
            Common_Modes::numModesLig = Ml;
            const float* e = res + 27 * i;
            Result_6D_Modes<float> r;
            r._6D.E = e[0];
            r._6D.pos = Vec3<float>(e[1], e[2], e[3]);
            r._6D.ang = Vec3<float>(e[4], e[5], e[6]);
            for (int k = 0; k < 10; ++k) {
                r.modesRec[k] = e[7 + k];
                r.modesLig[k] = e[17 + k];
            }
            s << r << std::endl;
        }
    }
    const std::string str = s.str();
    if (buf) std::memcpy(buf, str.data(), std::min(cap, str.size()));
    return str.size();
}

}  // extern "C"
