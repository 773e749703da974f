// Docking service with continuous batching (include/attract_serve.h).
//
// One engine thread owns the device context.  Each loop pass it
//   1. admits queued jobs in arrival order while their footprint fits the
//      budget (evicting least recently used receptors and ligands that no
//      running job uses), uploads what they need, starts their minimizers; a
//      job that has to wait for memory does not hold back later jobs that fit
//      now, up to kMaxBypass of them per waiting job, after which admission
//      waits for it (so a large job is not starved by a stream of small ones);
//   2. advances every running job whose last chunk of rounds has completed
//      (event query, no blocking): issues the next chunk on the job's stream,
//      or copies the results back and releases the job's memory.
// Jobs therefore enter and leave while others run; their kernels overlap on
// the device through their streams.  A job's kernels touch only its own
// buffers and the read-only receptor/ligand/table uploads, so its results do
// not depend on what else runs.
#include "attract_serve.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "as/DataManager.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Scoring rounds issued per chunk: the host learns the number of running
// poses once per chunk.
constexpr int kRoundsPerChunk = 4;

thread_local std::string t_err;

att_status set_err(att_status st, const std::string& msg) {
    t_err = msg;
    return st;
}

struct Receptor {
    as::ProteinData prot;
    as::GridUnionData grid;
    // a receptor registered as a body: grids built on the device on first use
    bool fromBody = false;
    uint64_t par = 0;
    att_grid_geom inner{}, outer{};
    float plateau = 0.f, nd = 0.f;
    size_t bytes = 0;
    att_handle hProt = 0, hGrid = 0;
    bool resident = false;
    int refs = 0;
    uint64_t lastUse = 0;
};

struct Ligand {
    as::ProteinData prot;
    size_t bytes = 0;
    att_handle h = 0;
    bool resident = false;
    int refs = 0;
    uint64_t lastUse = 0;
};

struct Params {
    std::vector<att_pair_params> table;
    unsigned T = 0;
    int shape = 8, dielec = 1;
    float eps = 1.f, ffelec = 0.f;
    att_handle h = 0;
};

struct Job {
    uint64_t id = 0, rec = 0, lig = 0, par = 0;
    std::vector<float> dofs;
    bool modes = false;
    size_t n = 0, ds = 6;
    int maxEval = 0;
    // set under the server lock
    bool done = false, failed = false;
    std::string error;
    // engine thread only
    bool finishing = false, chunkPending = false;
    int rounds = 0;
    Clock::time_point tSubmit, tAdmit, tDone;
    uint64_t admitSeq = 0;  // 1 for the first job admitted, 2 for the next, ...
    bool gridCached = false;
    void* stream = nullptr;
    void* ev = nullptr;
    att_minjob* mj = nullptr;
    void* dmem = nullptr;  // d_in | d_out | d_e | d_eb | d_nf
    size_t dbytes = 0, bytes = 0, held = 0;
    float *d_in = nullptr, *d_out = nullptr, *d_e = nullptr, *d_eb = nullptr;
    int* d_nf = nullptr;
    void* pinned = nullptr;  // running count (first 64 bytes), then the results
    std::vector<float> outDofs, outE, outEb;
    std::vector<int> outNf;
};

size_t grid_bytes(const as::GridUnionData::Grid& g) {
    const size_t cells = (size_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
    const size_t points = (size_t)g.nx * g.ny * g.nz;
    // voxel octets plus the point grid they are expanded from
    return g.numGrids * (cells * 128 + points * 16);
}

size_t protein_bytes(const as::ProteinData& p) {
    const size_t rowT = (3 * (size_t)p.numModes + 3) / 4;
    return p.n * (16 + 8 + 16 * rowT + 12 * (size_t)p.numModes) + 4096;
}

size_t receptor_bytes(const Receptor& r) {
    const auto& nl = r.grid.nl;
    const size_t lut = (size_t)(10000.0 * (double)nl.plateau * nl.plateau) + 1;
    return protein_bytes(r.prot) + grid_bytes(r.grid.inner) + grid_bytes(r.grid.outer) +
           4 * nl.desc.size() + 4 * nl.list.size() + 4 * lut;
}

// A receptor whose grids are built on the device: protein, octet grids, the
// neighbour-list descriptors, a bound on the list (every (bead, point) pair
// within the neighbour distance counted once), the ratio table, and the
// transient point grid of the generator.
size_t body_receptor_bytes(const Receptor& r, unsigned T) {
    auto octets = [&](const att_grid_geom& g) {
        return (size_t)(T + 1) * (g.nx - 1) * (g.ny - 1) * (g.nz - 1) * 128;
    };
    auto points = [&](const att_grid_geom& g) { return (size_t)(T + 1) * g.nx * g.ny * g.nz * 16; };
    const size_t npts = (size_t)r.inner.nx * r.inner.ny * r.inner.nz;
    const double sp = r.inner.spacing, nd = r.nd;
    const double perBead = 4.18879 * nd * nd * nd / (sp * sp * sp) + 1.0;
    const size_t listBound = std::min<size_t>((size_t)(perBead * r.prot.n), npts * (size_t)r.prot.n);
    const size_t lut = (size_t)(10000.0 * (double)r.plateau * r.plateau) + 1;
    return protein_bytes(r.prot) + octets(r.inner) + octets(r.outer) + 12 * npts + 4 * listBound +
           4 * lut + std::max(points(r.inner), points(r.outer));
}

void copy_protein(as::ProteinData& p, const att_protein_desc& d) {
    p.n = d.n;
    p.numModes = d.numModes;
    p.xyz.assign(d.xyz, d.xyz + 3 * (size_t)d.n);
    p.type.assign(d.type, d.type + d.n);
    p.mappedType.assign(d.mappedType, d.mappedType + d.n);
    p.charge.assign(d.charge, d.charge + d.n);
    if (d.numModes) {
        p.modes.assign(d.modes, d.modes + 3 * (size_t)d.n * d.numModes);
        p.modeForce.assign(d.modeForce, d.modeForce + d.numModes);
    }
}

bool protein_ok(const att_protein_desc* d) {
    return d && d->n && d->xyz && d->type && d->mappedType && d->charge &&
           d->numModes <= ATT_MODES_MAX && (!d->numModes || (d->modes && d->modeForce));
}

}  // namespace

// Later jobs admitted past one job that waits for memory, at most.
constexpr unsigned kMaxBypass = 32;

struct att_server {
    att_ctx* ctx = nullptr;
    size_t budget = 0, used = 0, peak = 0;
    std::mutex m;
    std::condition_variable cvWork, cvDone;
    bool stop = false;
    uint64_t nextId = 1, clock = 0;
    std::map<uint64_t, std::unique_ptr<Receptor>> receptors;
    std::map<uint64_t, std::unique_ptr<Ligand>> ligands;
    std::map<uint64_t, std::unique_ptr<Params>> params;
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<Job*> queue;
    uint64_t admitted = 0;
    const Job* waiting = nullptr;  // engine thread only: the oldest job waiting for memory
    unsigned bypassed = 0;         //   and how many later jobs were admitted past it
    std::vector<Job*> running;   // engine thread only
    std::vector<void*> streams;  // engine thread only: idle streams
    uint64_t jobsDone = 0, posesDone = 0, evaluations = 0, hits = 0, misses = 0, evictions = 0;
    unsigned maxRunning = 0;
    std::vector<double> latencies;
    double busy = 0.0;
    std::thread engine;

    void run();
    bool admit(Job* j);
    void finish(Job* j, const std::string& error);
    void release(Job* j);
    bool evict_for(size_t need, const Job* j);
    bool step(Job* j);
    std::string libErr() const { return att_last_error(ctx); }
};

// Frees a job's device resources and drops its references (engine thread).
void att_server::release(Job* j) {
    if (j->stream) att_sync(ctx, j->stream);
    if (j->mj) att_minjob_destroy(ctx, j->mj);
    if (j->dmem) att_dev_free(ctx, j->dmem);
    if (j->pinned) att_host_free(ctx, j->pinned);
    if (j->ev) att_event_destroy(ctx, j->ev);
    if (j->stream) streams.push_back(j->stream);
    j->mj = nullptr;
    j->dmem = j->pinned = j->ev = j->stream = nullptr;
    std::lock_guard<std::mutex> g(m);
    used -= j->bytes;
    j->bytes = 0;
    receptors[j->rec]->refs--;
    ligands[j->lig]->refs--;
}

// Ends an admitted job (engine thread): releases it, then publishes the
// outcome to waiters.
void att_server::finish(Job* j, const std::string& error) {
    release(j);
    j->tDone = Clock::now();
    std::lock_guard<std::mutex> g(m);
    if (error.empty()) {
        j->done = true;
        ++jobsDone;
        posesDone += j->n;
        for (int v : j->outNf) evaluations += (uint64_t)v + 1;
        latencies.push_back(ms_between(j->tSubmit, j->tDone));
    } else {
        j->failed = true;
        j->error = error;
    }
    cvDone.notify_all();
}

// Evicts resident receptors and ligands that no running job uses, least
// recently used first, until `need` more bytes fit (lock held by the caller).
bool att_server::evict_for(size_t need, const Job* j) {
    while (used + need > budget) {
        uint64_t best = 0, stamp = ~0ull;
        int kind = -1;
        for (auto& kv : receptors)
            if (kv.second->resident && kv.second->refs == 0 && kv.first != j->rec &&
                kv.second->lastUse < stamp) {
                best = kv.first; kind = 0; stamp = kv.second->lastUse;
            }
        for (auto& kv : ligands)
            if (kv.second->resident && kv.second->refs == 0 && kv.first != j->lig &&
                kv.second->lastUse < stamp) {
                best = kv.first; kind = 1; stamp = kv.second->lastUse;
            }
        if (kind < 0) return false;
        if (kind == 0) {
            Receptor& r = *receptors[best];
            att_release(ctx, r.hGrid);
            att_release(ctx, r.hProt);
            r.hGrid = r.hProt = 0;
            r.resident = false;
            used -= r.bytes;
        } else {
            Ligand& l = *ligands[best];
            att_release(ctx, l.h);
            l.h = 0;
            l.resident = false;
            used -= l.bytes;
        }
        ++evictions;
    }
    return true;
}

// Admits queued job j if its footprint fits (engine thread): removes it from
// the queue and starts it, or fails it when it can never fit.  Returns false
// when it has to wait for memory.
bool att_server::admit(Job* j) {
    Receptor* r;
    Ligand* l;
    Params* p;
    {
        std::lock_guard<std::mutex> g(m);
        r = receptors[j->rec].get();
        l = ligands[j->lig].get();
        p = params[j->par].get();
        const int dim = 6 + (j->modes ? (int)(r->prot.numModes + l->prot.numModes) : 0);
        j->dbytes = sizeof(float) * j->n * (2 * j->ds + 2) + sizeof(int) * j->n + 1024;
        j->bytes = j->held = att_minjob_bytes(j->n, dim, j->modes) + j->dbytes;
        const size_t need = j->bytes + (r->resident ? 0 : r->bytes) + (l->resident ? 0 : l->bytes);
        if (used + need > budget && !evict_for(need, j)) {
            if (!running.empty()) return false;
            // nothing will free memory: the job can never run
            queue.erase(std::find(queue.begin(), queue.end(), j));
            j->failed = true;
            j->error = "job needs " + std::to_string(need) + " bytes of device memory, budget " +
                       std::to_string(budget);
            j->tAdmit = j->tDone = Clock::now();
            cvDone.notify_all();
            return true;
        }
        queue.erase(std::find(queue.begin(), queue.end(), j));
        j->admitSeq = ++admitted;
        used += j->bytes;  // receptor and ligand bytes are added on upload
        r->refs++;
        l->refs++;
        r->lastUse = l->lastUse = ++clock;
        j->gridCached = r->resident;
        (r->resident ? hits : misses)++;
    }
    j->tAdmit = Clock::now();
    auto bad = [&](const char* what) {
        finish(j, std::string(what) + ": " + libErr());
        return true;
    };
    if (!r->resident) {
        const att_protein_desc pd = r->prot.desc();
        if (att_upload_protein(ctx, &pd, &r->hProt) != ATT_OK) return bad("receptor upload");
        att_status gs;
        if (r->fromBody) {
            gs = att_build_grid(ctx, &pd, p->table.data(), p->T, p->shape, p->dielec, p->ffelec,
                                r->plateau, 0, &r->inner, &r->outer, r->nd, &r->hGrid);
        } else {
            const att_grid_desc gi = r->grid.innerDesc(), go = r->grid.outerDesc();
            const att_nlgrid_desc nl = r->grid.nlDesc();
            gs = att_upload_grid(ctx, &gi, &go, &nl, &r->hGrid);
        }
        if (gs != ATT_OK) {
            att_release(ctx, r->hProt);
            r->hProt = 0;
            return bad("grid upload");
        }
        r->resident = true;
        std::lock_guard<std::mutex> g(m);
        used += r->bytes;
    }
    if (!l->resident) {
        const att_protein_desc pd = l->prot.desc();
        if (att_upload_protein(ctx, &pd, &l->h) != ATT_OK) return bad("ligand upload");
        l->resident = true;
        std::lock_guard<std::mutex> g(m);
        used += l->bytes;
    }
    {
        std::lock_guard<std::mutex> g(m);
        peak = std::max(peak, used);
    }
    if (!p->h && att_upload_params(ctx, p->table.data(), p->T, p->shape, p->dielec, p->eps,
                                   p->ffelec, &p->h) != ATT_OK)
        return bad("table upload");
    if (streams.empty()) {
        void* st = nullptr;
        if (att_stream_create(ctx, &st) != ATT_OK) return bad("stream");
        streams.push_back(st);
    }
    j->stream = streams.back();
    streams.pop_back();
    if (att_event_create(ctx, &j->ev) != ATT_OK) return bad("event");
    if (att_dev_alloc(ctx, j->dbytes, &j->dmem) != ATT_OK) return bad("device buffers");
    if (att_host_alloc(ctx, 64 + sizeof(float) * j->n * (j->ds + 2) + sizeof(int) * j->n, &j->pinned) != ATT_OK)
        return bad("pinned buffers");
    std::memset(j->pinned, 0, 64);
    j->d_in = static_cast<float*>(j->dmem);
    j->d_out = j->d_in + j->n * j->ds;
    j->d_e = j->d_out + j->n * j->ds;
    j->d_eb = j->d_e + j->n;
    j->d_nf = reinterpret_cast<int*>(j->d_eb + j->n);
    if (att_minjob_create(ctx, r->hProt, l->h, r->hGrid, p->h, j->modes ? 1 : 0, j->n, j->maxEval,
                          j->d_out, j->d_e, j->d_eb, j->d_nf, &j->mj) != ATT_OK)
        return bad("minimizer");
    if (att_memcpy_h2d_async(ctx, j->d_in, j->dofs.data(), sizeof(float) * j->n * j->ds, j->stream) != ATT_OK ||
        att_minjob_start(ctx, j->mj, j->d_in, j->stream) != ATT_OK)
        return bad("start");
    running.push_back(j);
    std::lock_guard<std::mutex> g(m);
    maxRunning = std::max<unsigned>(maxRunning, (unsigned)running.size());
    return true;
}

// Advances one running job whose previous chunk has completed.  Returns true
// when the job has ended (finished or failed).
bool att_server::step(Job* j) {
    float* h = reinterpret_cast<float*>(static_cast<char*>(j->pinned) + 64);
    if (j->finishing) {
        j->outDofs.assign(h, h + j->n * j->ds);
        j->outE.assign(h + j->n * j->ds, h + j->n * (j->ds + 1));
        j->outEb.assign(h + j->n * (j->ds + 1), h + j->n * (j->ds + 2));
        const int* nf = reinterpret_cast<const int*>(h + j->n * (j->ds + 2));
        j->outNf.assign(nf, nf + j->n);
        finish(j, "");
        return true;
    }
    const unsigned runningPoses = *static_cast<volatile unsigned*>(j->pinned);
    const int limit = att_minjob_round_limit(j->mj);
    if ((j->rounds > 0 && runningPoses == 0) || j->rounds >= limit) {
        const size_t n = j->n, ds = j->ds;
        const bool ok =
            att_memcpy_d2h_async(ctx, h, j->d_out, sizeof(float) * n * ds, j->stream) == ATT_OK &&
            att_memcpy_d2h_async(ctx, h + n * ds, j->d_e, sizeof(float) * n, j->stream) == ATT_OK &&
            att_memcpy_d2h_async(ctx, h + n * (ds + 1), j->d_eb, sizeof(float) * n, j->stream) == ATT_OK &&
            att_memcpy_d2h_async(ctx, h + n * (ds + 2), j->d_nf, sizeof(int) * n, j->stream) == ATT_OK &&
            att_event_record(ctx, j->ev, j->stream) == ATT_OK;
        if (!ok) {
            finish(j, "result copy: " + libErr());
            return true;
        }
        j->finishing = true;
    } else {
        const int r = std::min(kRoundsPerChunk, limit - j->rounds);
        if (att_minjob_rounds(ctx, j->mj, r, j->stream) != ATT_OK ||
            att_minjob_running(ctx, j->mj, static_cast<unsigned*>(j->pinned), j->stream) != ATT_OK ||
            att_event_record(ctx, j->ev, j->stream) != ATT_OK) {
            finish(j, "rounds: " + libErr());
            return true;
        }
        j->rounds += r;
    }
    j->chunkPending = true;
    return false;
}

void att_server::run() {
    Clock::time_point last = Clock::now();
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(m);
            if (running.empty() && queue.empty()) {
                cvWork.wait(lk, [&] { return stop || !queue.empty(); });
                if (queue.empty()) break;  // stop with nothing left to do
                last = Clock::now();
            }
        }
        bool progressed = false;
        // Admission in arrival order; past the oldest job waiting for memory
        // (while it has been bypassed fewer than kMaxBypass times) later jobs
        // are admitted when they fit.
        for (size_t idx = 0;;) {
            Job* j;
            {
                std::lock_guard<std::mutex> g(m);
                if (idx >= queue.size()) break;
                j = queue[idx];
            }
            if (admit(j)) {
                progressed = true;
                if (j == waiting) {
                    waiting = nullptr;
                    bypassed = 0;
                } else if (idx > 0) {
                    ++bypassed;
                }
                continue;  // the queue shrank at idx
            }
            if (idx == 0 && j != waiting) {
                waiting = j;
                bypassed = 0;
            }
            if (bypassed >= kMaxBypass) break;
            ++idx;
        }
        for (size_t k = 0; k < running.size();) {
            Job* j = running[k];
            if (j->chunkPending) {
                const att_status q = att_event_query(ctx, j->ev);
                if (q == ATT_NOT_READY) {
                    ++k;
                    continue;
                }
                j->chunkPending = false;
                if (q != ATT_OK) {
                    finish(j, "device: " + libErr());
                    running.erase(running.begin() + k);
                    progressed = true;
                    continue;
                }
            }
            progressed = true;
            if (step(j)) {
                running.erase(running.begin() + k);
                continue;
            }
            ++k;
        }
        const Clock::time_point now = Clock::now();
        {
            std::lock_guard<std::mutex> g(m);
            if (!running.empty()) busy += std::chrono::duration<double>(now - last).count();
        }
        last = now;
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(40));
    }
}

extern "C" {

att_status att_server_create(int device, size_t budget, att_server** out) {
    if (!out) return set_err(ATT_ERR_INVALID, "null output");
    std::unique_ptr<att_server> s(new att_server);
    if (att_ctx_create(device, &s->ctx) != ATT_OK) return set_err(ATT_ERR_HIP, "device context");
    size_t fr = 0, tot = 0;
    if (att_device_memory(s->ctx, &fr, &tot) != ATT_OK) {
        const std::string msg = s->libErr();
        att_ctx_destroy(s->ctx);
        return set_err(ATT_ERR_HIP, msg);
    }
    s->budget = budget ? std::min(budget, fr) : fr / 10 * 8;
    att_server* raw = s.get();
    s->engine = std::thread([raw] { raw->run(); });
    *out = s.release();
    return ATT_OK;
}

void att_server_destroy(att_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->m);
        s->stop = true;
        // jobs not admitted yet are dropped; running ones finish
        for (Job* j : s->queue) {
            j->failed = true;
            j->error = "server destroyed";
        }
        s->queue.clear();
        s->cvDone.notify_all();
    }
    s->cvWork.notify_all();
    if (s->engine.joinable()) s->engine.join();
    for (void* st : s->streams) att_stream_destroy(s->ctx, st);
    att_ctx_destroy(s->ctx);  // releases every upload
    delete s;
}

const char* att_server_last_error(const att_server*) { return t_err.c_str(); }

att_status att_server_add_receptor(att_server* s, const att_protein_desc* rec,
                                   const att_grid_desc* inner, const att_grid_desc* outer,
                                   const att_nlgrid_desc* nl, uint64_t* id) {
    if (!s || !id || !inner || !outer || !nl) return set_err(ATT_ERR_INVALID, "null argument");
    if (!protein_ok(rec) || !inner->data || !outer->data || !nl->desc || (nl->numEl && !nl->list))
        return set_err(ATT_ERR_INVALID, "incomplete receptor description");
    std::unique_ptr<Receptor> r(new Receptor);
    copy_protein(r->prot, *rec);
    auto copyGrid = [](as::GridUnionData::Grid& g, const att_grid_desc& d) {
        g.nx = d.nx; g.ny = d.ny; g.nz = d.nz; g.numGrids = d.numGrids;
        g.ox = d.ox; g.oy = d.oy; g.oz = d.oz; g.spacing = d.spacing;
        g.data.assign(d.data, d.data + 4 * (size_t)d.numGrids * d.nx * d.ny * d.nz);
    };
    copyGrid(r->grid.inner, *inner);
    copyGrid(r->grid.outer, *outer);
    r->grid.hasNL = true;
    auto& g = r->grid.nl;
    g.nx = nl->nx; g.ny = nl->ny; g.nz = nl->nz;
    g.ox = nl->ox; g.oy = nl->oy; g.oz = nl->oz;
    g.spacing = nl->spacing;
    g.plateau = nl->plateau;
    g.desc.assign(nl->desc, nl->desc + 2 * (size_t)nl->nx * nl->ny * nl->nz);
    if (nl->numEl) g.list.assign(nl->list, nl->list + nl->numEl);
    r->bytes = receptor_bytes(*r);
    std::lock_guard<std::mutex> lk(s->m);
    *id = s->nextId++;
    s->receptors[*id] = std::move(r);
    return ATT_OK;
}

att_status att_server_add_receptor_body(att_server* s, const att_protein_desc* rec, uint64_t params,
                                        const att_grid_geom* inner, const att_grid_geom* outer,
                                        float plateau, float neighbourdis, uint64_t* id) {
    if (!s || !id || !inner || !outer) return set_err(ATT_ERR_INVALID, "null argument");
    if (!protein_ok(rec)) return set_err(ATT_ERR_INVALID, "incomplete receptor description");
    if (!(plateau > 0.f) || !(neighbourdis > 0.f))
        return set_err(ATT_ERR_INVALID, "plateau and neighbour distance must be positive");
    std::unique_ptr<Receptor> r(new Receptor);
    copy_protein(r->prot, *rec);
    r->fromBody = true;
    r->par = params;
    r->inner = *inner;
    r->outer = *outer;
    r->plateau = plateau;
    r->nd = neighbourdis;
    std::lock_guard<std::mutex> lk(s->m);
    auto pit = s->params.find(params);
    if (pit == s->params.end()) return set_err(ATT_ERR_HANDLE, "unknown table id");
    r->bytes = body_receptor_bytes(*r, pit->second->T);
    *id = s->nextId++;
    s->receptors[*id] = std::move(r);
    return ATT_OK;
}

att_status att_server_add_ligand(att_server* s, const att_protein_desc* lig, uint64_t* id) {
    if (!s || !id) return set_err(ATT_ERR_INVALID, "null argument");
    if (!protein_ok(lig)) return set_err(ATT_ERR_INVALID, "incomplete ligand description");
    std::unique_ptr<Ligand> l(new Ligand);
    copy_protein(l->prot, *lig);
    l->bytes = protein_bytes(l->prot);
    std::lock_guard<std::mutex> lk(s->m);
    *id = s->nextId++;
    s->ligands[*id] = std::move(l);
    return ATT_OK;
}

att_status att_server_add_params(att_server* s, const att_pair_params* table, unsigned numTypes,
                                 int potShape, int dielec, float epsilon, float ffelec, uint64_t* id) {
    if (!s || !table || !id || numTypes == 0) return set_err(ATT_ERR_INVALID, "null argument or no types");
    std::unique_ptr<Params> p(new Params);
    p->table.assign(table, table + (size_t)numTypes * numTypes);
    p->T = numTypes;
    p->shape = potShape;
    p->dielec = dielec;
    p->eps = epsilon;
    p->ffelec = ffelec;
    std::lock_guard<std::mutex> lk(s->m);
    *id = s->nextId++;
    s->params[*id] = std::move(p);
    return ATT_OK;
}

att_status att_server_submit(att_server* s, uint64_t rec, uint64_t lig, uint64_t par,
                             const float* dofs, int modes, size_t n, int maxFunEval, uint64_t* job) {
    if (!s || !job || !dofs) return set_err(ATT_ERR_INVALID, "null argument");
    if (n == 0 || n > 0x7FFFFFFFull || maxFunEval < 0)
        return set_err(ATT_ERR_INVALID, "pose count out of range or negative evaluation limit");
    std::unique_ptr<Job> j(new Job);
    j->rec = rec;
    j->lig = lig;
    j->par = par;
    j->modes = modes != 0;
    j->ds = modes ? 26 : 6;
    j->n = n;
    j->maxEval = maxFunEval;
    j->dofs.assign(dofs, dofs + n * j->ds);
    j->tSubmit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (!s->receptors.count(rec) || !s->ligands.count(lig) || !s->params.count(par))
            return set_err(ATT_ERR_HANDLE, "unknown receptor, ligand or table id");
        const Receptor& r = *s->receptors[rec];
        if (r.fromBody && r.par != par)
            return set_err(ATT_ERR_INVALID, "the receptor's grids were registered with another table");
        if (s->stop) return set_err(ATT_ERR_INVALID, "server is shutting down");
        j->id = s->nextId++;
        *job = j->id;
        s->queue.push_back(j.get());
        s->jobs[j->id] = std::move(j);
    }
    s->cvWork.notify_all();
    return ATT_OK;
}

att_status att_server_wait(att_server* s, uint64_t id, int timeout_ms, float* dofs_out,
                           float* energy, float* ebest, int* nfun, att_job_info* info) {
    if (!s) return set_err(ATT_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_err(ATT_ERR_HANDLE, "unknown job id");
    Job* j = it->second.get();
    auto ended = [&] { return j->done || j->failed; };
    if (timeout_ms < 0) s->cvDone.wait(lk, ended);
    else if (!s->cvDone.wait_for(lk, std::chrono::milliseconds(timeout_ms), ended))
        return set_err(ATT_NOT_READY, "job still running");
    if (j->failed) {
        const std::string msg = j->error;
        s->jobs.erase(it);
        return set_err(ATT_ERR_UNSUPPORTED, msg);
    }
    if (dofs_out) std::copy(j->outDofs.begin(), j->outDofs.end(), dofs_out);
    if (energy) std::copy(j->outE.begin(), j->outE.end(), energy);
    if (ebest) std::copy(j->outEb.begin(), j->outEb.end(), ebest);
    if (nfun) std::copy(j->outNf.begin(), j->outNf.end(), nfun);
    if (info) {
        info->queued_ms = ms_between(j->tSubmit, j->tAdmit);
        info->run_ms = ms_between(j->tAdmit, j->tDone);
        info->latency_ms = ms_between(j->tSubmit, j->tDone);
        info->grid_cached = j->gridCached ? 1 : 0;
        info->rounds = (unsigned)j->rounds;
        info->bytes = j->held;
        info->admit_seq = j->admitSeq;
    }
    s->jobs.erase(it);
    return ATT_OK;
}

att_status att_server_get_stats(att_server* s, att_server_stats* o) {
    if (!s || !o) return set_err(ATT_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->m);
    o->jobs_done = s->jobsDone;
    o->poses_done = s->posesDone;
    o->evaluations = s->evaluations;
    o->grid_hits = s->hits;
    o->grid_misses = s->misses;
    o->evictions = s->evictions;
    o->budget_bytes = s->budget;
    o->peak_bytes = s->peak;
    o->max_running = s->maxRunning;
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
        if (l.empty()) return 0.0;
        const size_t k = std::min(l.size() - 1, (size_t)(q * (double)(l.size() - 1) + 0.5));
        return l[k];
    };
    o->p50_ms = pct(0.50);
    o->p99_ms = pct(0.99);
    o->busy_s = s->busy;
    return ATT_OK;
}

}  // extern "C"
