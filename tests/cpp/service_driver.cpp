// Test and benchmark driver (not part of the library): runs scoring requests
// through as::Server with the MI355 services, so the GPU tests can compare the
// drop-in path with direct C-ABI calls on the same workloads and the benchmark
// can time the drop-in path itself (drv_time_rigid / drv_time_modes).
//
// drv_score_rigid / drv_score_modes copy the descriptors' arrays into a
// DataManager, create the service on HIP device 0 with `workers` workers and
// the given item size, split the poses into `requests` requests submitted
// together, wait for them in reverse order and write the results in input
// order.  Return 0, or -1 with drv_error() describing the failure.
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "as/MI355EnergyService6D.h"
#include "as/Server.h"

namespace {

std::string g_error;

std::shared_ptr<as::ProteinData> protein(const att_protein_desc& d) {
    auto p = std::make_shared<as::ProteinData>();
    p->n = d.n;
    p->numModes = d.numModes;
    p->xyz.assign(d.xyz, d.xyz + 3 * (size_t)d.n);
    p->type.assign(d.type, d.type + d.n);
    p->mappedType.assign(d.mappedType, d.mappedType + d.n);
    p->charge.assign(d.charge, d.charge + d.n);
    if (d.numModes) {
        p->modes.assign(d.modes, d.modes + 3 * (size_t)d.n * d.numModes);
        p->modeForce.assign(d.modeForce, d.modeForce + d.numModes);
    }
    return p;
}

void grid(as::GridUnionData::Grid& g, const att_grid_desc& d) {
    g.nx = d.nx; g.ny = d.ny; g.nz = d.nz; g.numGrids = d.numGrids;
    g.ox = d.ox; g.oy = d.oy; g.oz = d.oz; g.spacing = d.spacing;
    g.data.assign(d.data, d.data + 4 * (size_t)d.numGrids * d.nx * d.ny * d.nz);
}

std::shared_ptr<as::GridUnionData> grid_union(const att_grid_desc& in, const att_grid_desc& out,
                                              const att_nlgrid_desc* nl) {
    auto g = std::make_shared<as::GridUnionData>();
    grid(g->inner, in);
    grid(g->outer, out);
    if (nl) {
        g->hasNL = true;
        g->nl.nx = nl->nx; g->nl.ny = nl->ny; g->nl.nz = nl->nz;
        g->nl.ox = nl->ox; g->nl.oy = nl->oy; g->nl.oz = nl->oz;
        g->nl.spacing = nl->spacing;
        g->nl.plateau = nl->plateau;
        g->nl.desc.assign(nl->desc, nl->desc + 2 * (size_t)nl->nx * nl->ny * nl->nz);
        if (nl->numEl) g->nl.list.assign(nl->list, nl->list + nl->numEl);
    }
    return g;
}

template <class Service, class Common>
int run(std::shared_ptr<Service> svc, Common common, const typename Service::input_t* dofs, size_t n,
        typename Service::result_t* out, unsigned workers, unsigned itemSize, unsigned requests) {
    using server_t = as::Server<typename Service::genericTypes_t>;
    server_t server(svc);
    server.createWorkers(workers);
    server.setItemSize(itemSize);
    std::vector<typename Service::input_t> in(dofs, dofs + n);
    std::vector<typename server_t::request_t> reqs(requests);
    std::vector<size_t> lo(requests + 1);
    for (unsigned r = 0; r <= requests; ++r) lo[r] = n * r / requests;
    for (unsigned r = 0; r < requests; ++r) {
        reqs[r].setDataAndSize(in.data() + lo[r], lo[r + 1] - lo[r]);
        reqs[r].setCommon(common);
        server.submit(reqs[r]);
    }
    for (unsigned r = requests; r-- > 0;) server.wait(reqs[r], out + lo[r]);
    return 0;
}

// One Server with `workers` workers and the given item size scores the n
// poses as `requests` requests submitted together, once untimed, then `reps`
// times; returns the mean wall time of one timed pass (submit of every
// request to the last wait).
template <class Service, class Common>
double time_server(std::shared_ptr<Service> svc, Common common, const typename Service::input_t* dofs,
                   size_t n, typename Service::result_t* out, unsigned workers, unsigned itemSize,
                   unsigned requests, unsigned reps) {
    using server_t = as::Server<typename Service::genericTypes_t>;
    server_t server(svc);
    server.createWorkers(workers);
    server.setItemSize(itemSize);
    std::vector<typename Service::input_t> in(dofs, dofs + n);
    std::vector<typename server_t::request_t> reqs(requests);
    std::vector<size_t> lo(requests + 1);
    for (unsigned r = 0; r <= requests; ++r) lo[r] = n * r / requests;
    auto pass = [&]() {
        for (unsigned r = 0; r < requests; ++r) {
            reqs[r].setDataAndSize(in.data() + lo[r], lo[r + 1] - lo[r]);
            reqs[r].setCommon(common);
            server.submit(reqs[r]);
        }
        for (unsigned r = 0; r < requests; ++r) server.wait(reqs[r], out + lo[r]);
    };
    pass();
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned k = 0; k < reps; ++k) pass();
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count() / (reps ? reps : 1);
}
std::shared_ptr<as::DataManager> tables(std::shared_ptr<as::DataManager> dm, const att_pair_params* tab,
                                        unsigned T, int shape, int dielec, float eps, float ffelec,
                                        as::id_t& tableId, as::id_t& paramsId) {
    auto t = std::make_shared<as::ParamTableData>();
    t->numTypes = T;
    t->potShape = shape;
    t->table.assign(tab, tab + (size_t)T * T);
    auto s = std::make_shared<as::SimParamData>();
    s->dielec = dielec;
    s->epsilon = eps;
    s->ffelec = ffelec;
    tableId = dm->add(t);
    paramsId = dm->add(s);
    return dm;
}

}  // namespace

extern "C" {

const char* drv_error() { return g_error.c_str(); }

int drv_score_rigid(const att_protein_desc* rec, const att_protein_desc* lig, const att_grid_desc* inner,
                    const att_grid_desc* outer, const att_nlgrid_desc* nl, const att_pair_params* tab,
                    unsigned T, int shape, int dielec, float eps, float ffelec, const float* dofs,
                    size_t n, float* out, unsigned workers, unsigned itemSize, unsigned requests) {
    try {
        auto dm = std::make_shared<as::DataManager>();
        as::Common c{};
        c.recId = dm->add(protein(*rec));
        c.ligId = dm->add(protein(*lig));
        c.gridId = dm->add(grid_union(*inner, *outer, nl));
        tables(dm, tab, T, shape, dielec, eps, ffelec, c.tableId, c.paramsId);
        auto svc = std::make_shared<as::MI355EnergyService6D<float>>(dm, std::vector<int>{0});
        return run(svc, c, reinterpret_cast<const as::DOF_6D<float>*>(dofs), n,
                   reinterpret_cast<as::Result_6D<float>*>(out), workers, itemSize, requests);
    } catch (std::exception const& e) {
        g_error = e.what();
        return -1;
    }
}

int drv_score_modes(const att_protein_desc* rec, const att_protein_desc* lig, const att_grid_desc* rInner,
                    const att_grid_desc* rOuter, const att_nlgrid_desc* nl, const att_grid_desc* lInner,
                    const att_grid_desc* lOuter, const att_pair_params* tab, unsigned T, int shape,
                    int dielec, float eps, float ffelec, const float* dofs, size_t n, float* out,
                    unsigned workers, unsigned itemSize, unsigned requests) {
    try {
        auto dm = std::make_shared<as::DataManager>();
        as::Common_Modes c{};
        c.recId = dm->add(protein(*rec));
        c.ligId = dm->add(protein(*lig));
        c.gridIdRec = dm->add(grid_union(*rInner, *rOuter, nl));
        c.gridIdLig = dm->add(grid_union(*lInner, *lOuter, nullptr));
        tables(dm, tab, T, shape, dielec, eps, ffelec, c.tableId, c.paramsId);
        auto svc = std::make_shared<as::MI355EnergyService6DModes<float>>(dm, std::vector<int>{0});
        return run(svc, c, reinterpret_cast<const as::DOF_6D_Modes<float>*>(dofs), n,
                   reinterpret_cast<as::Result_6D_Modes<float>*>(out), workers, itemSize, requests);
    } catch (std::exception const& e) {
        g_error = e.what();
        return -1;
    }
}

/* Throughput of the drop-in path (time_server above): *seconds receives the
 * mean wall time of one pass over the n poses. */
int drv_time_rigid(const att_protein_desc* rec, const att_protein_desc* lig, const att_grid_desc* inner,
                   const att_grid_desc* outer, const att_nlgrid_desc* nl, const att_pair_params* tab,
                   unsigned T, int shape, int dielec, float eps, float ffelec, const float* dofs,
                   size_t n, float* out, unsigned workers, unsigned itemSize, unsigned requests,
                   unsigned reps, double* seconds) {
    try {
        auto dm = std::make_shared<as::DataManager>();
        as::Common c{};
        c.recId = dm->add(protein(*rec));
        c.ligId = dm->add(protein(*lig));
        c.gridId = dm->add(grid_union(*inner, *outer, nl));
        tables(dm, tab, T, shape, dielec, eps, ffelec, c.tableId, c.paramsId);
        auto svc = std::make_shared<as::MI355EnergyService6D<float>>(dm, std::vector<int>{0});
        *seconds = time_server(svc, c, reinterpret_cast<const as::DOF_6D<float>*>(dofs), n,
                               reinterpret_cast<as::Result_6D<float>*>(out), workers, itemSize,
                               requests, reps);
        return 0;
    } catch (std::exception const& e) {
        g_error = e.what();
        return -1;
    }
}
int drv_time_modes(const att_protein_desc* rec, const att_protein_desc* lig, const att_grid_desc* rInner,
                   const att_grid_desc* rOuter, const att_nlgrid_desc* nl, const att_grid_desc* lInner,
                   const att_grid_desc* lOuter, const att_pair_params* tab, unsigned T, int shape,
                   int dielec, float eps, float ffelec, const float* dofs, size_t n, float* out,
                   unsigned workers, unsigned itemSize, unsigned requests, unsigned reps,
                   double* seconds) {
    try {
        auto dm = std::make_shared<as::DataManager>();
        as::Common_Modes c{};
        c.recId = dm->add(protein(*rec));
        c.ligId = dm->add(protein(*lig));
        c.gridIdRec = dm->add(grid_union(*rInner, *rOuter, nl));
        c.gridIdLig = dm->add(grid_union(*lInner, *lOuter, nullptr));
        tables(dm, tab, T, shape, dielec, eps, ffelec, c.tableId, c.paramsId);
        auto svc = std::make_shared<as::MI355EnergyService6DModes<float>>(dm, std::vector<int>{0});
        *seconds = time_server(svc, c, reinterpret_cast<const as::DOF_6D_Modes<float>*>(dofs), n,
                               reinterpret_cast<as::Result_6D_Modes<float>*>(out), workers, itemSize,
                               requests, reps);
        return 0;
    } catch (std::exception const& e) {
        g_error = e.what();
        return -1;
    }
}

/* DataManager error behaviour (CPU only). */
int drv_datamanager_errors(char* buf, size_t len) {
    std::string all;
    as::DataManager dm;
    try { dm.add(nullptr); } catch (std::exception const& e) { all += std::string(e.what()) + "\n"; }
    try { dm.get(3); } catch (std::exception const& e) { all += std::string(e.what()) + "\n"; }
    const as::id_t id = dm.add(std::make_shared<as::SimParamData>());
    dm.remove(id);
    try { dm.remove(id); } catch (std::exception const& e) { all += std::string(e.what()) + "\n"; }
    std::strncpy(buf, all.c_str(), len);
    if (len) buf[len - 1] = 0;
    return 0;
}

}  // extern "C"
