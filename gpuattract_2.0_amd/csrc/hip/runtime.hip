// Host runtime behind include/attract_hip.h: contexts, device-resident data
// objects addressed by handles, input validation, and the scoring entry points.
//
// Device data layout (per object, all in HBM, uploaded once):
//   protein  float4 {x,y,z,charge}[n], u32 type[n], u32 mappedType[n],
//            modes [c][m][n] (coalesced over beads) and [n][m][c] (one run per
//            bead, for neighbour-list gathers), modeForce[M]
//   grid     inner/outer float4 [grid][z][y][x] exactly as the reference's
//            IntrplGrid, neighbour-list {numEl, idx}[z][y][x] + u32 list (the
//            plateau-ratio table of model/NLGrid.tpp:44-51 is evaluated per
//            pair by the kernels, not stored)
//   params   att_pair_params[T*T] + shape/dielectric/ffelec
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "attract_hip.h"
#include "internal.h"

using namespace att;

namespace {

enum class Kind { Protein, Grid, Params };

struct Obj {
    explicit Obj(Kind k) : kind(k) {}
    virtual ~Obj() {
        for (void* p : allocs) (void)hipFree(p);
    }
    Kind kind;
    std::vector<void*> allocs;
};

struct ProtObj : Obj {
    ProtObj() : Obj(Kind::Protein) {}
    DProt d{};
    unsigned maxType = 0, maxMType = 0;
};

struct GridObj : Obj {
    GridObj() : Obj(Kind::Grid) {}
    DGrid inner{}, outer{};
    DBasis bin{}, bout{};
    DNL nl{};
    bool hasNL = false;
    uint32_t nlMaxIndex = 0;
    bool nlEmpty = true;
};

struct ParObj : Obj {
    ParObj() : Obj(Kind::Params) {}
    DPar d{};
    float epsilon = 0.f;
};

}  // namespace

struct att_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint64_t next = 1;
    std::unordered_map<uint64_t, std::unique_ptr<Obj>> objs;
    void* d_in = nullptr;
    size_t d_in_cap = 0;
    void* d_out = nullptr;
    size_t d_out_cap = 0;
    size_t maxLdsPerBlock = 65536;
    unsigned numCU = 256;
    void* d_pg = nullptr;  // per-workgroup parameter-gradient rows of the adjoint
    size_t d_pg_cap = 0;
    void* d_rec = nullptr;  // grid-cell records of the parameter gradient and their cell pass
    size_t d_rec_cap = 0;
    void* d_w = nullptr;   // staging of host-buffer adjoint calls
    size_t d_w_cap = 0;
    void* d_g = nullptr;
    size_t d_g_cap = 0;
    void* d_ord = nullptr;  // spatial pose order: permutation + sort scratch
    size_t d_ord_cap = 0;
    bool poseOrder = true;
    int modesWaves = 0;  // waves per pose of the flexible forward kernel, 0: by ligand size
    uint2* pairCounts = nullptr;  // att_debug_pair_counts: per-pose counts, capacity in poses
    size_t pairCountsCap = 0;
    // Recorded after the last kernel that used d_pg, d_rec or d_ord: a call on
    // another stream waits for it before reusing them.
    hipEvent_t scratchEv = nullptr;
    // Device address of this GPU's `scratch` (att_score_adjoint_dev_noalloc).
    void* noallocScratch = nullptr;
};

// Scratch of att_score_adjoint_dev_noalloc: the adjoint's per-workgroup
// parameter-derivative rows (float) and their group sums (double), for at
// most kNoAllocBlocks workgroups and 16 pair types (3 * 16^2 columns).  One
// copy per GPU the library runs on, held until the process ends (13 MB).
constexpr size_t kNoAllocBlocks = 4096;
constexpr size_t kNoAllocCols = 3 * 16 * 16;
constexpr size_t kNoAllocRowBytes = (sizeof(float) * kNoAllocBlocks * kNoAllocCols + 255) & ~(size_t)255;
constexpr size_t kNoAllocBytes = kNoAllocRowBytes + sizeof(double) * att::kPgradGroups * kNoAllocCols;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// The device address of `scratch` on each GPU (the array's host-side name is
// no device address): looked up once per device, with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void* g_scratchPtr[kMaxDevices] = {};

hipError_t scratch_address(int device, void** ptr) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[device] = p;
    }
    *ptr = g_scratchPtr[device];
    return hipSuccess;
}

}  // namespace

namespace {

att_status fail(att_ctx* c, att_status code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define ATT_HIP(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(ctx, e_ == hipErrorOutOfMemory ? ATT_ERR_NOMEM : ATT_ERR_HIP,         \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

template <class T>
att_status dev_copy(att_ctx* c, Obj* o, const T* host, size_t count, T** out) {
    *out = nullptr;
    if (count == 0) return ATT_OK;
    void* p = nullptr;
    ATT_HIP(c, hipMalloc(&p, sizeof(T) * count));
    o->allocs.push_back(p);
    ATT_HIP(c, hipMemcpy(p, host, sizeof(T) * count, hipMemcpyHostToDevice));
    *out = static_cast<T*>(p);
    return ATT_OK;
}

template <class T>
T* lookup(att_ctx* c, att_handle h, Kind k) {
    auto it = c->objs.find(h);
    if (it == c->objs.end() || it->second->kind != k) return nullptr;
    return static_cast<T*>(it->second.get());
}

att_handle store(att_ctx* c, std::unique_ptr<Obj> o) {
    const att_handle h = c->next++;
    c->objs.emplace(h, std::move(o));
    return h;
}

// Grid constants as IntrplGrid's constructor computes them (float arithmetic).
DGrid make_dgrid(const att_grid_desc& g, const float4* oct) {
    DGrid d{};
    d.oct = oct;
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz;
    d.nxy = g.nx * g.ny;
    d.nvox = g.nx * g.ny * g.nz;
    d.ngrids = g.numGrids;
    d.dVox = g.spacing;
    const float vol = d.dVox * d.dVox * d.dVox;
    d.vol_inv = 1.f / vol;
    d.dVox_inv = 1.f / d.dVox;
    d.lox = g.ox; d.loy = g.oy; d.loz = g.oz;
    d.hix = g.ox + (float)(g.nx - 1) * d.dVox;
    d.hiy = g.oy + (float)(g.ny - 1) * d.dVox;
    d.hiz = g.oz + (float)(g.nz - 1) * d.dVox;
    return d;
}

// Upload a [grid][z][y][x] float4 grid and expand it on the device into voxel
// octets (8x the memory, one 128-byte line per lookup); the compact copy is
// freed once expanded.
att_status upload_octets(att_ctx* c, Obj* o, const att_grid_desc& g, float4** out) {
    const size_t n = (size_t)g.nx * g.ny * g.nz * g.numGrids;
    const size_t noct = (size_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1) * g.numGrids * 8;
    void* tmp = nullptr;
    void* oct = nullptr;
    ATT_HIP(c, hipMalloc(&oct, sizeof(float4) * noct));
    o->allocs.push_back(oct);
    hipError_t e = hipMalloc(&tmp, sizeof(float4) * n);
    if (e == hipSuccess) e = hipMemcpy(tmp, g.data, sizeof(float4) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_octets(static_cast<const float4*>(tmp), static_cast<float4*>(oct), g.nx, g.ny,
                          g.nz, g.numGrids, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess)
        return fail(c, e == hipErrorOutOfMemory ? ATT_ERR_NOMEM : ATT_ERR_HIP,
                    std::string("grid upload: ") + hipGetErrorString(e));
    *out = static_cast<float4*>(oct);
    return ATT_OK;
}

// Expands a device-resident point grid into voxel octets (owned by o).
att_status expand_octets(att_ctx* c, Obj* o, const att_grid_desc& g, const float4* d_points,
                         float4** out) {
    const size_t noct = (size_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1) * g.numGrids * 8;
    void* oct = nullptr;
    ATT_HIP(c, hipMalloc(&oct, sizeof(float4) * noct));
    o->allocs.push_back(oct);
    ATT_HIP(c, launch_octets(d_points, static_cast<float4*>(oct), g.nx, g.ny, g.nz, g.numGrids,
                             c->stream));
    *out = static_cast<float4*>(oct);
    return ATT_OK;
}

// Neighbour-list constants of a grid whose desc/list are already on the
// device.
att_status attach_nl(att_ctx* c, GridObj* o, unsigned nx, unsigned ny, unsigned nz, float ox,
                     float oy, float oz, float spacing, float plateau, const uint2* ddesc,
                     const uint32_t* dlist) {
    (void)c;  // nothing to allocate
    const float P2 = plateau * plateau;
    DNL& d = o->nl;
    d.desc = ddesc;
    d.list = dlist;
    d.nx = nx; d.ny = ny; d.nz = nz;
    const float dv = spacing;
    d.lox = ox; d.loy = oy; d.loz = oz;
    d.hix = ox + (float)(nx - 1) * dv;
    d.hiy = oy + (float)(ny - 1) * dv;
    d.hiz = oz + (float)(nz - 1) * dv;
    d.dVox_inv = 1.f / dv;
    d.P2 = P2;
    d.P2_inv = 1.f / P2;
    o->hasNL = true;
    return ATT_OK;
}

att_status check_grid_desc(att_ctx* c, const att_grid_desc* g, const char* what) {
    if (!g || !g->data) return fail(c, ATT_ERR_INVALID, std::string(what) + ": missing grid data");
    if (g->nx < 2 || g->ny < 2 || g->nz < 2 || g->numGrids < 1)
        return fail(c, ATT_ERR_INVALID, std::string(what) + ": grid needs >= 2 points per axis and >= 1 grid");
    if (!(g->spacing > 0.f) || !std::isfinite(g->spacing))
        return fail(c, ATT_ERR_INVALID, std::string(what) + ": grid spacing must be positive");
    const uint64_t total = (uint64_t)g->nx * g->ny * g->nz * g->numGrids * 8;
    if (total >= (1ull << 34))
        return fail(c, ATT_ERR_UNSUPPORTED, std::string(what) + ": grid exceeds 2^31 points");
    return ATT_OK;
}

// Checks shared by the scoring entry points: everything a kernel indexes with
// is validated here so that no input can make a kernel read out of bounds.
att_status check_pair(att_ctx* c, const ProtObj* rec, const ProtObj* lig, const GridObj* g,
                      const ParObj* p) {
    if (!g->hasNL) return fail(c, ATT_ERR_INVALID, "receptor grid has no neighbour list");
    if (!g->nlEmpty && g->nlMaxIndex >= rec->d.n)
        return fail(c, ATT_ERR_INVALID, "neighbour list refers to beads beyond the receptor");
    if (lig->maxType > p->d.T)
        return fail(c, ATT_ERR_INVALID, "ligand type exceeds the number of parameter-table types");
    if (lig->maxMType >= g->inner.ngrids || lig->maxMType >= g->outer.ngrids)
        return fail(c, ATT_ERR_INVALID, "ligand mapped type has no grid");
    return ATT_OK;
}

}  // namespace

namespace {
// Waves per pose of the single-pair adjoint and of the minimizer's scoring
// rounds: one for ligands under 384 beads, two above (timed: a 300-bead
// ligand 7.5 % faster on one wave than on two; an 800-bead ligand fastest on
// two of 1, 2, 4).  Both paths use it, so a point scored by either gives the
// same bits.
int pair_waves(unsigned ligN) { return ligN < 384 ? 1 : 2; }
// Parameter-gradient records per call up to which the cell pass takes the
// grid part (68 B of scratch per record without pose weights, 76 B with:
// 9-10 GB at the cap), and the workgroups of that pass (one parameter row
// each).
constexpr size_t kPgRecordCap = (size_t)1 << 27;
constexpr unsigned kPgCellBlocks = 2048;
// Waves per pose of the flexible forward kernel (k_modes): with 384 ligand
// beads or more, four (3000/800 beads, 5 + 5 modes: 100,000 poses 3% slower
// than with two, a call of 100 poses 1.6x faster).  A function of the pair
// only, so a pose scores the same bits whatever else is in its call.
int modes_waves(unsigned ligN) { return ligN < 384 ? 1 : 4; }

// Scoped device buffers for the diagnostic entry points.
struct Tmp {
    std::vector<void*> p;
    ~Tmp() {
        for (void* q : p) (void)hipFree(q);
    }
    template <class T>
    hipError_t alloc(T** out, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1));
        if (e == hipSuccess) p.push_back(q);
        *out = static_cast<T*>(q);
        return e;
    }
};
}  // namespace

extern "C" {

const char* att_build_info(void) {
    return "gpuattract_2.0_amd: gfx950, -O3 -ffp-contract=off, packed-fp32 disabled";
}

att_status att_set_pose_order(att_ctx* c, int enabled) {
    if (!c) return ATT_ERR_INVALID;
    c->poseOrder = enabled != 0;
    return ATT_OK;
}

att_status att_debug_pair_counts(att_ctx* c, void* d_counts, size_t capacity) {
    if (!c) return ATT_ERR_INVALID;
    c->pairCounts = static_cast<uint2*>(d_counts);
    c->pairCountsCap = d_counts ? capacity : 0;
    return ATT_OK;
}

// The per-pose pair-count buffer for a call of n poses (null when off).
static att_status pair_counts(att_ctx* c, size_t n, uint2** out) {
    *out = nullptr;
    if (!c->pairCounts) return ATT_OK;
    if (n > c->pairCountsCap)
        return fail(c, ATT_ERR_INVALID, "pair-count buffer holds fewer poses than the call");
    *out = c->pairCounts;
    return ATT_OK;
}

att_status att_set_modes_waves(att_ctx* c, int waves) {
    if (!c) return ATT_ERR_INVALID;
    if (waves != 0 && waves != 1 && waves != 2 && waves != 4 && waves != 8)
        return fail(c, ATT_ERR_INVALID, "waves per pose must be 0, 1, 2, 4 or 8");
    c->modesWaves = waves;
    return ATT_OK;
}

att_status att_ctx_create(int hip_device, att_ctx** out) {
    if (!out) return ATT_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return ATT_ERR_HIP;
    if (hip_device < 0 || hip_device >= count) return ATT_ERR_INVALID;
    if (hipSetDevice(hip_device) != hipSuccess) return ATT_ERR_HIP;
    auto c = std::make_unique<att_ctx>();
    c->device = hip_device;
    if (hipStreamCreate(&c->stream) != hipSuccess) return ATT_ERR_HIP;
    if (hipEventCreateWithFlags(&c->scratchEv, hipEventDisableTiming) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        return ATT_ERR_HIP;
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, hip_device) == hipSuccess && lds > 0)
        c->maxLdsPerBlock = (size_t)lds;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, hip_device) == hipSuccess && cus > 0)
        c->numCU = (unsigned)cus;
    if (scratch_address(hip_device, &c->noallocScratch) != hipSuccess) {
        (void)hipEventDestroy(c->scratchEv);
        (void)hipStreamDestroy(c->stream);
        return ATT_ERR_HIP;
    }
    *out = c.release();
    return ATT_OK;
}

void att_ctx_destroy(att_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->objs.clear();
    if (c->d_in) (void)hipFree(c->d_in);
    if (c->d_out) (void)hipFree(c->d_out);
    if (c->d_pg) (void)hipFree(c->d_pg);
    if (c->d_rec) (void)hipFree(c->d_rec);
    if (c->d_ord) (void)hipFree(c->d_ord);
    if (c->d_w) (void)hipFree(c->d_w);
    if (c->d_g) (void)hipFree(c->d_g);
    if (c->scratchEv) (void)hipEventDestroy(c->scratchEv);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* att_last_error(const att_ctx* c) { return c ? c->err.c_str() : "no context"; }

att_status att_upload_protein(att_ctx* c, const att_protein_desc* d, att_handle* out) {
    if (!c || !d || !out) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    if (d->n == 0) return fail(c, ATT_ERR_INVALID, "protein has no beads");
    if (d->numModes > ATT_MODES_MAX) return fail(c, ATT_ERR_INVALID, "more than 10 modes");
    if (!d->xyz || !d->type || !d->mappedType || !d->charge)
        return fail(c, ATT_ERR_INVALID, "protein arrays missing");
    if (d->numModes > 0 && (!d->modes || !d->modeForce))
        return fail(c, ATT_ERR_INVALID, "mode arrays missing");
    const unsigned n = d->n, M = d->numModes;
    auto o = std::make_unique<ProtObj>();
    std::vector<float4> pos(n);
    for (unsigned i = 0; i < n; ++i)
        pos[i] = make_float4(d->xyz[i], d->xyz[n + i], d->xyz[2 * (size_t)n + i], d->charge[i]);
    o->maxType = *std::max_element(d->type, d->type + n);
    if (o->maxType > 0xFFFFu) return fail(c, ATT_ERR_INVALID, "bead type exceeds 65535");
    o->maxMType = *std::max_element(d->mappedType, d->mappedType + n);
    float4* dpos;
    unsigned *dtype, *dmtype;
    att_status st;
    if ((st = dev_copy(c, o.get(), pos.data(), n, &dpos)) != ATT_OK) return st;
    if ((st = dev_copy(c, o.get(), d->type, n, &dtype)) != ATT_OK) return st;
    if ((st = dev_copy(c, o.get(), d->mappedType, n, &dmtype)) != ATT_OK) return st;
    o->d.pos = dpos;
    o->d.type = dtype;
    o->d.mtype = dmtype;
    o->d.n = n;
    o->d.M = M;
    if (M > 0) {
        const size_t nM = (size_t)n * M;
        const unsigned rowT = (3 * M + 3) / 4;  // float4 per padded row
        std::vector<float> soa(3 * nM), aos((size_t)n * rowT * 4, 0.f);
        for (unsigned cc = 0; cc < 3; ++cc)
            for (unsigned i = 0; i < n; ++i)
                for (unsigned m = 0; m < M; ++m) {
                    const float v = d->modes[cc * nM + (size_t)i * M + m];
                    soa[(cc * M + m) * (size_t)n + i] = v;
                    aos[(size_t)i * rowT * 4 + 3 * m + cc] = v;
                }
        float *dsoa, *dmf;
        float4* daos;
        if ((st = dev_copy(c, o.get(), soa.data(), soa.size(), &dsoa)) != ATT_OK) return st;
        if ((st = dev_copy(c, o.get(), reinterpret_cast<const float4*>(aos.data()),
                           (size_t)n * rowT, &daos)) != ATT_OK)
            return st;
        if ((st = dev_copy(c, o.get(), d->modeForce, M, &dmf)) != ATT_OK) return st;
        o->d.modes = dsoa;
        o->d.modesT = daos;
        o->d.modeForce = dmf;
        o->d.rowT = rowT;
    }
    *out = store(c, std::move(o));
    return ATT_OK;
}

att_status att_upload_grid(att_ctx* c, const att_grid_desc* inner, const att_grid_desc* outer,
                           const att_nlgrid_desc* nl, att_handle* out) {
    if (!c || !out) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    att_status st;
    if ((st = check_grid_desc(c, inner, "inner")) != ATT_OK) return st;
    if ((st = check_grid_desc(c, outer, "outer")) != ATT_OK) return st;
    if (inner->numGrids != outer->numGrids)
        return fail(c, ATT_ERR_INVALID, "inner and outer grids differ in the number of types");
    auto o = std::make_unique<GridObj>();
    float4 *din, *dout;
    if ((st = upload_octets(c, o.get(), *inner, &din)) != ATT_OK) return st;
    if ((st = upload_octets(c, o.get(), *outer, &dout)) != ATT_OK) return st;
    o->inner = make_dgrid(*inner, din);
    o->outer = make_dgrid(*outer, dout);
    if (nl) {
        if (!nl->desc) return fail(c, ATT_ERR_INVALID, "neighbour grid descriptors missing");
        if (nl->nx < 1 || nl->ny < 1 || nl->nz < 1) return fail(c, ATT_ERR_INVALID, "empty neighbour grid");
        if (!(nl->plateau > 0.f) || !(nl->spacing > 0.f))
            return fail(c, ATT_ERR_INVALID, "neighbour grid needs positive spacing and plateau");
        const size_t nv = (size_t)nl->nx * nl->ny * nl->nz;
        if (nv >= (1ull << 31)) return fail(c, ATT_ERR_UNSUPPORTED, "neighbour grid too large");
        if (nl->numEl > 0 && !nl->list) return fail(c, ATT_ERR_INVALID, "neighbour list missing");
        if (nl->numEl >= (1ull << 32)) return fail(c, ATT_ERR_UNSUPPORTED, "neighbour list exceeds 2^32 entries");
        // Every voxel's list must lie inside the concatenated array.
        for (size_t v = 0; v < nv; ++v) {
            const uint64_t cnt = nl->desc[2 * v], idx = nl->desc[2 * v + 1];
            if (cnt > 0 && idx + cnt > nl->numEl)
                return fail(c, ATT_ERR_INVALID, "neighbour list descriptor points outside the list");
        }
        uint32_t mx = 0;
        for (size_t k = 0; k < nl->numEl; ++k) mx = std::max(mx, nl->list[k]);
        o->nlMaxIndex = mx;
        o->nlEmpty = nl->numEl == 0;
        uint2* ddesc;
        uint32_t* dlist = nullptr;
        if ((st = dev_copy(c, o.get(), reinterpret_cast<const uint2*>(nl->desc), nv, &ddesc)) != ATT_OK) return st;
        if (nl->numEl > 0 && (st = dev_copy(c, o.get(), nl->list, nl->numEl, &dlist)) != ATT_OK) return st;
        if ((st = attach_nl(c, o.get(), nl->nx, nl->ny, nl->nz, nl->ox, nl->oy, nl->oz, nl->spacing,
                            nl->plateau, ddesc, dlist)) != ATT_OK)
            return st;
    }
    *out = store(c, std::move(o));
    return ATT_OK;
}

att_status att_upload_params(att_ctx* c, const att_pair_params* table, unsigned numTypes,
                             int potShape, int dielec, float epsilon, float ffelec,
                             att_handle* out) {
    if (!c || !table || !out) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    if (numTypes == 0 || numTypes > 99) return fail(c, ATT_ERR_INVALID, "number of types must be 1..99");
    if (potShape != 8 && potShape != 12) return fail(c, ATT_ERR_INVALID, "potential shape must be 8 or 12");
    if (dielec != 0 && dielec != 1) return fail(c, ATT_ERR_INVALID, "dielectric must be 0 (constant) or 1 (variable)");
    auto o = std::make_unique<ParObj>();
    att_pair_params* dt;
    att_status st;
    if ((st = dev_copy(c, o.get(), table, (size_t)numTypes * numTypes, &dt)) != ATT_OK) return st;
    o->d.tab = dt;
    o->d.T = numTypes;
    o->d.shape = potShape;
    o->d.dielec = dielec;
    o->d.ffelec = ffelec;
    o->epsilon = epsilon;
    *out = store(c, std::move(o));
    return ATT_OK;
}

att_status att_release(att_ctx* c, att_handle h) {
    if (!c) return ATT_ERR_INVALID;
    auto it = c->objs.find(h);
    if (it == c->objs.end()) return fail(c, ATT_ERR_HANDLE, "unknown handle");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->objs.erase(it);
    return ATT_OK;
}

static att_status resolve(att_ctx* c, att_handle hr, att_handle hl, att_handle hg, att_handle hp,
                          ProtObj** rec, ProtObj** lig, GridObj** g, ParObj** p) {
    *rec = lookup<ProtObj>(c, hr, Kind::Protein);
    *lig = lookup<ProtObj>(c, hl, Kind::Protein);
    *g = lookup<GridObj>(c, hg, Kind::Grid);
    *p = lookup<ParObj>(c, hp, Kind::Params);
    if (!*rec || !*lig || !*g || !*p) return fail(c, ATT_ERR_HANDLE, "unknown handle or wrong kind");
    return ATT_OK;
}

static att_status grow(att_ctx* c, void** p, size_t* cap, size_t bytes) {
    if (bytes <= *cap) return ATT_OK;
    if (*p) ATT_HIP(c, hipEventSynchronize(c->scratchEv));  // its last user has finished
    if (*p) ATT_HIP(c, hipFree(*p));
    *p = nullptr;
    *cap = 0;
    ATT_HIP(c, hipMalloc(p, bytes));
    *cap = bytes;
    return ATT_OK;
}

// Spatial order of n poses (DOF rows of `stride` floats) into the context's
// scratch on stream s, the stream having waited for the scratch's previous
// user; *order stays null when the order is off or n is below minPoses.  The
// caller records scratchEv on s after its last kernel that reads the order.
static att_status pose_order(att_ctx* c, const float* d_dofs, unsigned stride, size_t n,
                             size_t minPoses, hipStream_t s, const unsigned** order) {
    *order = nullptr;
    if (!c->poseOrder || n < minPoses || n == 0) return ATT_OK;
    const size_t scratch = pose_order_scratch((unsigned)n);
    att_status st;
    if ((st = grow(c, &c->d_ord, &c->d_ord_cap, scratch + sizeof(unsigned) * n)) != ATT_OK) return st;
    unsigned* o = reinterpret_cast<unsigned*>(static_cast<char*>(c->d_ord) + scratch);
    ATT_HIP(c, hipStreamWaitEvent(s, c->scratchEv, 0));
    ATT_HIP(c, launch_pose_order(d_dofs, stride, (unsigned)n, c->d_ord, scratch, o, s));
    *order = o;
    return ATT_OK;
}

// Forward kernels sort their poses spatially from this many poses on.
constexpr size_t kForwardOrderMin = 4096;

static att_status ensure_staging(att_ctx* c, size_t inBytes, size_t outBytes) {
    if (inBytes > c->d_in_cap) {
        if (c->d_in) ATT_HIP(c, hipFree(c->d_in));
        c->d_in = nullptr;
        c->d_in_cap = 0;
        ATT_HIP(c, hipMalloc(&c->d_in, inBytes));
        c->d_in_cap = inBytes;
    }
    if (outBytes > c->d_out_cap) {
        if (c->d_out) ATT_HIP(c, hipFree(c->d_out));
        c->d_out = nullptr;
        c->d_out_cap = 0;
        ATT_HIP(c, hipMalloc(&c->d_out, outBytes));
        c->d_out_cap = outBytes;
    }
    return ATT_OK;
}

att_status att_score_rigid_dev(att_ctx* c, att_handle hr, att_handle hl, att_handle hg,
                               att_handle hp, const att_dof6* d_dofs, size_t n,
                               att_result6* d_out, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (n == 0) return ATT_OK;
    if (!d_dofs || !d_out) return fail(c, ATT_ERR_INVALID, "null buffer");
    if (n > 0xFFFFFFFFull) return fail(c, ATT_ERR_UNSUPPORTED, "too many poses in one call");
    ATT_HIP(c, hipSetDevice(c->device));
    ProtObj *rec, *lig;
    GridObj* g;
    ParObj* p;
    att_status st;
    if ((st = resolve(c, hr, hl, hg, hp, &rec, &lig, &g, &p)) != ATT_OK) return st;
    if ((st = check_pair(c, rec, lig, g, p)) != ATT_OK) return st;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    uint2* counts;
    if ((st = pair_counts(c, n, &counts)) != ATT_OK) return st;
    const unsigned* order = nullptr;
    if ((st = pose_order(c, reinterpret_cast<const float*>(d_dofs), 6, n, kForwardOrderMin, s, &order)) != ATT_OK)
        return st;
    ATT_HIP(c, launch_rigid(rec->d, lig->d, g->inner, g->outer, g->nl, p->d,
                            reinterpret_cast<const float*>(d_dofs), reinterpret_cast<float*>(d_out),
                            (unsigned)n, order, counts, s));
    if (order) ATT_HIP(c, hipEventRecord(c->scratchEv, s));
    return ATT_OK;
}

att_status att_score_rigid(att_ctx* c, att_handle hr, att_handle hl, att_handle hg, att_handle hp,
                           const att_dof6* dofs, size_t n, att_result6* out, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (n == 0) return ATT_OK;
    if (!dofs || !out) return fail(c, ATT_ERR_INVALID, "null buffer");
    ATT_HIP(c, hipSetDevice(c->device));
    att_status st;
    if ((st = ensure_staging(c, n * sizeof(att_dof6), n * sizeof(att_result6))) != ATT_OK) return st;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    ATT_HIP(c, hipMemcpyAsync(c->d_in, dofs, n * sizeof(att_dof6), hipMemcpyHostToDevice, s));
    if ((st = att_score_rigid_dev(c, hr, hl, hg, hp, static_cast<const att_dof6*>(c->d_in), n,
                                  static_cast<att_result6*>(c->d_out), s)) != ATT_OK)
        return st;
    ATT_HIP(c, hipMemcpyAsync(out, c->d_out, n * sizeof(att_result6), hipMemcpyDeviceToHost, s));
    ATT_HIP(c, hipStreamSynchronize(s));
    return ATT_OK;
}

static att_status check_modes(att_ctx* c, ProtObj* rec, ProtObj* lig, GridObj* gr, GridObj* gl,
                              ParObj* p) {
    att_status st;
    if ((st = check_pair(c, rec, lig, gr, p)) != ATT_OK) return st;
    if (rec->maxMType >= gl->inner.ngrids || rec->maxMType >= gl->outer.ngrids)
        return fail(c, ATT_ERR_INVALID, "receptor mapped type has no ligand grid");
    return ATT_OK;
}

att_status att_score_modes_dev(att_ctx* c, att_handle hr, att_handle hl, att_handle hgr,
                               att_handle hgl, att_handle hp, const att_dof6m* d_dofs, size_t n,
                               att_result6m* d_out, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (n == 0) return ATT_OK;
    if (!d_dofs || !d_out) return fail(c, ATT_ERR_INVALID, "null buffer");
    if (n > 0x7FFFFFFFull) return fail(c, ATT_ERR_UNSUPPORTED, "too many poses in one call");
    ATT_HIP(c, hipSetDevice(c->device));
    ProtObj *rec, *lig;
    GridObj* gr;
    ParObj* p;
    att_status st;
    if ((st = resolve(c, hr, hl, hgr, hp, &rec, &lig, &gr, &p)) != ATT_OK) return st;
    GridObj* gl = lookup<GridObj>(c, hgl, Kind::Grid);
    if (!gl) return fail(c, ATT_ERR_HANDLE, "unknown ligand grid handle");
    if ((st = check_modes(c, rec, lig, gr, gl, p)) != ATT_OK) return st;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    uint2* counts;
    if ((st = pair_counts(c, n, &counts)) != ATT_OK) return st;
    const unsigned* order = nullptr;
    if ((st = pose_order(c, reinterpret_cast<const float*>(d_dofs), 26, n, kForwardOrderMin, s, &order)) != ATT_OK)
        return st;
    ATT_HIP(c, launch_modes(rec->d, lig->d, gr->inner, gr->outer, gr->nl, gl->inner, gl->outer,
                            p->d, reinterpret_cast<const float*>(d_dofs),
                            reinterpret_cast<float*>(d_out), (unsigned)n, order,
                            c->modesWaves ? c->modesWaves : modes_waves(lig->d.n), counts, s));
    if (order) ATT_HIP(c, hipEventRecord(c->scratchEv, s));
    return ATT_OK;
}

att_status att_score_modes(att_ctx* c, att_handle hr, att_handle hl, att_handle hgr, att_handle hgl,
                           att_handle hp, const att_dof6m* dofs, size_t n, att_result6m* out,
                           void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (n == 0) return ATT_OK;
    if (!dofs || !out) return fail(c, ATT_ERR_INVALID, "null buffer");
    ATT_HIP(c, hipSetDevice(c->device));
    att_status st;
    if ((st = ensure_staging(c, n * sizeof(att_dof6m), n * sizeof(att_result6m))) != ATT_OK) return st;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    ATT_HIP(c, hipMemcpyAsync(c->d_in, dofs, n * sizeof(att_dof6m), hipMemcpyHostToDevice, s));
    if ((st = att_score_modes_dev(c, hr, hl, hgr, hgl, hp, static_cast<const att_dof6m*>(c->d_in), n,
                                  static_cast<att_result6m*>(c->d_out), s)) != ATT_OK)
        return st;
    ATT_HIP(c, hipMemcpyAsync(out, c->d_out, n * sizeof(att_result6m), hipMemcpyDeviceToHost, s));
    ATT_HIP(c, hipStreamSynchronize(s));
    return ATT_OK;
}


att_status att_upload_basis(att_ctx* c, att_handle hg, unsigned T, const float* inner,
                            const float* outer) {
    if (!c || !inner || !outer) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    GridObj* g = lookup<GridObj>(c, hg, Kind::Grid);
    if (!g) return fail(c, ATT_ERR_HANDLE, "unknown grid handle");
    if (T == 0 || T > 99) return fail(c, ATT_ERR_INVALID, "number of types must be 1..99");
    const DGrid* gs[2] = {&g->inner, &g->outer};
    const float* src[2] = {inner, outer};
    DBasis* dst[2] = {&g->bin, &g->bout};
    for (int k = 0; k < 2; ++k) {
        const DGrid& G = *gs[k];
        const size_t nv = (size_t)G.nx * G.ny * G.nz;
        const size_t nvo = (size_t)(G.nx - 1) * (G.ny - 1) * (G.nz - 1);
        void* tmp = nullptr;
        void* oct = nullptr;
        ATT_HIP(c, hipMalloc(&oct, sizeof(float4) * 2 * nvo * 2 * T));
        g->allocs.push_back(oct);
        hipError_t e = hipMalloc(&tmp, sizeof(float) * nv * 2 * T);
        if (e == hipSuccess) e = hipMemcpy(tmp, src[k], sizeof(float) * nv * 2 * T, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = launch_basis_octets(static_cast<const float*>(tmp), static_cast<float4*>(oct), G.nx,
                                    G.ny, G.nz, 2 * T, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (tmp) (void)hipFree(tmp);
        if (e != hipSuccess) return fail(c, ATT_ERR_HIP, std::string("basis upload: ") + hipGetErrorString(e));
        dst[k]->oct = static_cast<const float4*>(oct);
        dst[k]->nb = 2 * T;
    }
    return ATT_OK;
}

att_status att_score_adjoint_dev(att_ctx* c, att_handle hr, att_handle hl, att_handle hg,
                                 att_handle hp, const float* d_dofs, int modes, size_t n,
                                 const float* d_w, float* d_out, float* d_pgrad, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (!d_dofs || !d_out) return fail(c, ATT_ERR_INVALID, "null buffer");
    if (n > 0x7FFFFFFFull) return fail(c, ATT_ERR_UNSUPPORTED, "too many poses in one call");
    ATT_HIP(c, hipSetDevice(c->device));
    ProtObj *rec, *lig;
    GridObj* g;
    ParObj* p;
    att_status st;
    if ((st = resolve(c, hr, hl, hg, hp, &rec, &lig, &g, &p)) != ATT_OK) return st;
    if ((st = check_pair(c, rec, lig, g, p)) != ATT_OK) return st;
    if (lig->d.n > 0xFFFFu) return fail(c, ATT_ERR_UNSUPPORTED, "ligand exceeds 65535 beads");
    const bool pgrad = d_pgrad != nullptr;
    if (pgrad && (!g->bin.oct || g->bin.nb != 2 * p->d.T))
        return fail(c, ATT_ERR_INVALID, "parameter gradients need basis grids for this table (att_upload_basis)");
    unsigned slots = 0;
    const int kWaves = pair_waves(lig->d.n);
    const size_t dyn = adjoint_lds_bytes(rec->d, lig->d, p->d, pgrad, kWaves, &slots);
    if (dyn + adjoint_static_lds(kWaves) > c->maxLdsPerBlock)
        return fail(c, ATT_ERR_UNSUPPORTED, "ligand too large for the LDS of the adjoint kernel");
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const unsigned cols = 3 * p->d.T * p->d.T;
    // The grid part of the parameter gradient through grid-cell records and
    // the cell pass, up to kPgRecordCap records (bead x pose) per call; larger
    // calls take it from the basis grids at every bead.
    const size_t records = pgrad ? n * (size_t)lig->d.n : 0;
    const size_t cellsIn = (size_t)(g->inner.nx - 1) * (g->inner.ny - 1) * (g->inner.nz - 1);
    const size_t cellsOut = (size_t)(g->outer.nx - 1) * (g->outer.ny - 1) * (g->outer.nz - 1);
    const size_t invalid = (cellsIn + cellsOut) * p->d.T;
    const bool useRec = records > 0 && records <= kPgRecordCap && invalid < 0xFFFFFFFFull &&
                        sizeof(float) * cols <= c->maxLdsPerBlock;
    const int pgMode = !pgrad ? kPgNone : useRec ? kPgRecords : kPgFused;
    // persistent grid: as many workgroups as the device holds at once
    const unsigned mm = modes ? std::max(rec->d.M, lig->d.M) : 0;
    const size_t resident = (size_t)c->numCU * adjoint_blocks_per_cu(mm, kWaves, false, dyn, pgMode);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(n, resident));
    const unsigned cellBlocks = useRec ? kPgCellBlocks : 0;
    // per-workgroup rows (the adjoint's, then the cell pass's), then the
    // reduction's group sums (8-byte aligned)
    const size_t rowBytes = (sizeof(float) * ((size_t)blocks + cellBlocks) * cols + 255) & ~(size_t)255;
    if (pgrad && (st = grow(c, &c->d_pg, &c->d_pg_cap, rowBytes + pgrad_group_bytes(cols))) != ATT_OK)
        return st;
    unsigned endBit = 0;
    while (endBit < 32 && (invalid >> endBit) != 0) ++endBit;
    const bool recW = d_w != nullptr;
    if (useRec &&
        (st = grow(c, &c->d_rec, &c->d_rec_cap, pg_cells_bytes(records, endBit, recW))) != ATT_OK)
        return st;
    if (n == 0) {
        if (pgrad) ATT_HIP(c, launch_fill_zero(d_pgrad, cols, s));
        return ATT_OK;
    }
    PairDesc P{};
    P.rec = rec->d; P.lig = lig->d;
    P.inner = g->inner; P.outer = g->outer; P.nl = g->nl;
    P.par = p->d;
    P.bin = g->bin; P.bout = g->bout;
    P.dofs = d_dofs; P.wts = d_w; P.out = d_out;
    P.dofStride = modes ? 26 : 6; P.outStride = modes ? 27 : 7;
    P.n = (unsigned)n; P.tabSlots = slots;
    if ((st = pair_counts(c, n, &P.counts)) != ATT_OK) return st;
    PgCellBufs B{};
    if (useRec) {
        B = pg_cells_layout(c->d_rec, records, endBit, recW);
        P.recKey = B.key;
        if (recW)
            P.recPos4 = static_cast<float4*>(B.pos);
        else
            P.recPos3 = static_cast<float3*>(B.pos);
        P.recCellsIn = (unsigned)cellsIn;
        P.recInvalid = (unsigned)invalid;
    }
    if (blocks >= 8 && (st = pose_order(c, d_dofs, P.dofStride, n, 4 * (size_t)blocks, s, &P.order)) != ATT_OK)
        return st;
    const bool scratch = pgrad || P.order;
    if (pgrad && !P.order) ATT_HIP(c, hipStreamWaitEvent(s, c->scratchEv, 0));
    float* rows = pgrad ? static_cast<float*>(c->d_pg) : nullptr;
    ATT_HIP(c, launch_adjoint(P, modes != 0, rows, blocks, kWaves, dyn, s));
    if (useRec)
        ATT_HIP(c, launch_pg_cells(P, B, records, endBit, rows + (size_t)blocks * cols, cellBlocks, s));
    if (pgrad)
        ATT_HIP(c, launch_pgrad_reduce(rows, blocks + cellBlocks, cols,
                                       reinterpret_cast<double*>(static_cast<char*>(c->d_pg) + rowBytes),
                                       d_pgrad, s));
    if (scratch) ATT_HIP(c, hipEventRecord(c->scratchEv, s));
    return ATT_OK;
}

att_status att_score_adjoint_dev_noalloc(att_ctx* c, att_handle hr, att_handle hl, att_handle hg,
                                         att_handle hp, const float* d_dofs, int modes, size_t n,
                                         const float* d_w, float* d_out, float* d_pgrad,
                                         void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (!d_dofs || !d_out) return fail(c, ATT_ERR_INVALID, "null buffer");
    if (n > 0x7FFFFFFFull) return fail(c, ATT_ERR_UNSUPPORTED, "too many poses in one call");
    if (!c->noallocScratch) return fail(c, ATT_ERR_HIP, "no scratch address for this device");
    ATT_HIP(c, hipSetDevice(c->device));
    ProtObj *rec, *lig;
    GridObj* g;
    ParObj* p;
    att_status st;
    if ((st = resolve(c, hr, hl, hg, hp, &rec, &lig, &g, &p)) != ATT_OK) return st;
    if ((st = check_pair(c, rec, lig, g, p)) != ATT_OK) return st;
    const bool pgrad = d_pgrad != nullptr;
    if (pgrad && (!g->bin.oct || g->bin.nb != 2 * p->d.T))
        return fail(c, ATT_ERR_INVALID, "parameter gradients need basis grids for this table (att_upload_basis)");
    const unsigned cols = 3 * p->d.T * p->d.T;
    if (pgrad && cols > kNoAllocCols)
        return fail(c, ATT_ERR_UNSUPPORTED, "parameter gradients without allocation: at most 16 pair types");
    unsigned slots = 0;
    const int kWaves = pair_waves(lig->d.n);
    const size_t dyn = adjoint_lds_bytes(rec->d, lig->d, p->d, pgrad, kWaves, &slots);
    if (dyn + adjoint_static_lds(kWaves) > c->maxLdsPerBlock)
        return fail(c, ATT_ERR_UNSUPPORTED, "pair table too large for the LDS of the adjoint kernel");
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const unsigned mm = modes ? std::max(rec->d.M, lig->d.M) : 0;
    const size_t resident =
        (size_t)c->numCU * adjoint_blocks_per_cu(mm, kWaves, false, dyn, pgrad ? kPgFused : kPgNone);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(n, resident));
    if (pgrad && blocks > kNoAllocBlocks)
        return fail(c, ATT_ERR_UNSUPPORTED, "more resident workgroups than the no-allocation scratch holds");
    if (n == 0) {
        if (pgrad) ATT_HIP(c, launch_fill_zero(d_pgrad, cols, s));
        return ATT_OK;
    }
    PairDesc P{};
    P.rec = rec->d; P.lig = lig->d;
    P.inner = g->inner; P.outer = g->outer; P.nl = g->nl;
    P.par = p->d;
    P.bin = g->bin; P.bout = g->bout;
    P.dofs = d_dofs; P.wts = d_w; P.out = d_out;
    P.dofStride = modes ? 26 : 6; P.outStride = modes ? 27 : 7;
    P.n = (unsigned)n; P.tabSlots = slots;
    if ((st = pair_counts(c, n, &P.counts)) != ATT_OK) return st;
    P.order = nullptr;  // input order: the spatial order needs sort scratch
    // Every workgroup writes its whole row and the reduction every group sum,
    // so the scratch needs no clearing between calls.
    float* rows = static_cast<float*>(c->noallocScratch);
    double* groups = reinterpret_cast<double*>(static_cast<char*>(c->noallocScratch) + kNoAllocRowBytes);
    ATT_HIP(c, launch_adjoint(P, modes != 0, pgrad ? rows : nullptr, blocks, kWaves, dyn, s));
    if (pgrad) ATT_HIP(c, launch_pgrad_reduce(rows, blocks, cols, groups, d_pgrad, s));
    return ATT_OK;
}

att_status att_param_grad_rmin_emin(const att_pair_params* table, unsigned numTypes, int potShape,
                                    const float* pgrad, float* out) {
    if (!table || !pgrad || !out || (potShape != 8 && potShape != 12)) return ATT_ERR_INVALID;
    // rc = -kr emin rmin^s, ac = -ka emin rmin^6
    const double q = 1.25992105;  // the float constant of the 12-6 derivation, as a double
    const double s = potShape == 8 ? 8.0 : 12.0;
    const double kr = potShape == 8 ? 3.0 : 4.0 / (q * q * q * q * q * q);
    const double ka = potShape == 8 ? 4.0 : 4.0 / (q * q * q);
    for (size_t k = 0; k < (size_t)numTypes * numTypes; ++k) {
        const double r2 = table[k].rmin2 > 0.f ? (double)table[k].rmin2 : 0.0;
        const double r = std::sqrt(r2), em = (double)table[k].emin;
        const double gr = pgrad[3 * k], ga = pgrad[3 * k + 1], ge = pgrad[3 * k + 2];
        const double rs1 = std::pow(r, s - 1.0), r5 = r2 * r2 * r;
        out[2 * k] = (float)(gr * (-kr * em * s * rs1) + ga * (-ka * em * 6.0 * r5));
        out[2 * k + 1] = (float)(gr * (-kr * rs1 * r) + ga * (-ka * r5 * r) + ge);
    }
    return ATT_OK;
}

att_status att_score_adjoint(att_ctx* c, att_handle hr, att_handle hl, att_handle hg, att_handle hp,
                             const float* dofs, int modes, size_t n, const float* w, float* out,
                             float* pgrad, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (!dofs || !out) return fail(c, ATT_ERR_INVALID, "null buffer");
    ATT_HIP(c, hipSetDevice(c->device));
    ParObj* p = lookup<ParObj>(c, hp, Kind::Params);
    if (!p) return fail(c, ATT_ERR_HANDLE, "unknown parameter handle");
    const size_t ds = modes ? 26 : 6, os = modes ? 27 : 7;
    const size_t cols = 3 * (size_t)p->d.T * p->d.T;
    att_status st;
    if ((st = ensure_staging(c, std::max<size_t>(n, 1) * ds * sizeof(float),
                             std::max<size_t>(n, 1) * os * sizeof(float))) != ATT_OK)
        return st;
    if (w && (st = grow(c, &c->d_w, &c->d_w_cap, std::max<size_t>(n, 1) * sizeof(float))) != ATT_OK) return st;
    if (pgrad && (st = grow(c, &c->d_g, &c->d_g_cap, cols * sizeof(float))) != ATT_OK) return st;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (n) ATT_HIP(c, hipMemcpyAsync(c->d_in, dofs, n * ds * sizeof(float), hipMemcpyHostToDevice, s));
    if (w && n) ATT_HIP(c, hipMemcpyAsync(c->d_w, w, n * sizeof(float), hipMemcpyHostToDevice, s));
    if ((st = att_score_adjoint_dev(c, hr, hl, hg, hp, static_cast<const float*>(c->d_in), modes, n,
                                    w ? static_cast<const float*>(c->d_w) : nullptr,
                                    static_cast<float*>(c->d_out),
                                    pgrad ? static_cast<float*>(c->d_g) : nullptr, s)) != ATT_OK)
        return st;
    if (n) ATT_HIP(c, hipMemcpyAsync(out, c->d_out, n * os * sizeof(float), hipMemcpyDeviceToHost, s));
    if (pgrad) ATT_HIP(c, hipMemcpyAsync(pgrad, c->d_g, cols * sizeof(float), hipMemcpyDeviceToHost, s));
    ATT_HIP(c, hipStreamSynchronize(s));
    return ATT_OK;
}

att_status att_debug_beads_rigid(att_ctx* c, att_handle hr, att_handle hl, att_handle hg,
                                 att_handle hp, const att_dof6* dof, float* xyz, float* pot,
                                 int* cell, unsigned* nl_pairs) {
    if (!c || !dof || !xyz || !pot || !cell || !nl_pairs) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    ProtObj *rec, *lig;
    GridObj* g;
    ParObj* p;
    att_status st;
    if ((st = resolve(c, hr, hl, hg, hp, &rec, &lig, &g, &p)) != ATT_OK) return st;
    if ((st = check_pair(c, rec, lig, g, p)) != ATT_OK) return st;
    const unsigned n = lig->d.n;
    Tmp t;
    float *ddof, *dxyz, *dpot;
    int* dcell;
    unsigned* dnp;
    ATT_HIP(c, t.alloc(&ddof, 6));
    ATT_HIP(c, t.alloc(&dxyz, 3 * (size_t)n));
    ATT_HIP(c, t.alloc(&dpot, 4 * (size_t)n));
    ATT_HIP(c, t.alloc(&dcell, 9 * (size_t)n));
    ATT_HIP(c, t.alloc(&dnp, n));
    ATT_HIP(c, hipMemcpy(ddof, dof, sizeof(att_dof6), hipMemcpyHostToDevice));
    ATT_HIP(c, launch_debug_rigid(rec->d, lig->d, g->inner, g->outer, g->nl, p->d, ddof, dxyz, dpot,
                                  dcell, dnp, c->stream));
    ATT_HIP(c, hipStreamSynchronize(c->stream));
    ATT_HIP(c, hipMemcpy(xyz, dxyz, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
    ATT_HIP(c, hipMemcpy(pot, dpot, sizeof(float) * 4 * n, hipMemcpyDeviceToHost));
    ATT_HIP(c, hipMemcpy(cell, dcell, sizeof(int) * 9 * n, hipMemcpyDeviceToHost));
    ATT_HIP(c, hipMemcpy(nl_pairs, dnp, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
    return ATT_OK;
}

att_status att_debug_beads_modes(att_ctx* c, att_handle hr, att_handle hl, att_handle hgr,
                                 att_handle hgl, att_handle hp, const att_dof6m* dof, float* xyz,
                                 float* pot, int* cell, unsigned* nl_pairs, float* rec_pot,
                                 float* rec_xyz) {
    if (!c || !dof || !xyz || !pot || !cell || !nl_pairs || !rec_pot || !rec_xyz)
        return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    ProtObj *rec, *lig;
    GridObj* gr;
    ParObj* p;
    att_status st;
    if ((st = resolve(c, hr, hl, hgr, hp, &rec, &lig, &gr, &p)) != ATT_OK) return st;
    GridObj* gl = lookup<GridObj>(c, hgl, Kind::Grid);
    if (!gl) return fail(c, ATT_ERR_HANDLE, "unknown ligand grid handle");
    if ((st = check_pair(c, rec, lig, gr, p)) != ATT_OK) return st;
    if (rec->maxMType >= gl->inner.ngrids || rec->maxMType >= gl->outer.ngrids)
        return fail(c, ATT_ERR_INVALID, "receptor mapped type has no ligand grid");
    const unsigned n = lig->d.n, nr = rec->d.n;
    Tmp t;
    float *ddof, *dxyz, *dpot, *drpot, *drxyz, *drdefo;
    int* dcell;
    unsigned* dnp;
    ATT_HIP(c, t.alloc(&ddof, 26));
    ATT_HIP(c, t.alloc(&dxyz, 3 * (size_t)n));
    ATT_HIP(c, t.alloc(&dpot, 4 * (size_t)n));
    ATT_HIP(c, t.alloc(&dcell, 9 * (size_t)n));
    ATT_HIP(c, t.alloc(&dnp, n));
    ATT_HIP(c, t.alloc(&drpot, 3 * (size_t)nr));
    ATT_HIP(c, t.alloc(&drxyz, 3 * (size_t)nr));
    ATT_HIP(c, t.alloc(&drdefo, 3 * (size_t)nr));
    ATT_HIP(c, hipMemcpy(ddof, dof, sizeof(att_dof6m), hipMemcpyHostToDevice));
    ATT_HIP(c, launch_debug_modes(rec->d, lig->d, gr->inner, gr->outer, gr->nl, gl->inner, gl->outer,
                                  p->d, ddof, dxyz, dpot, dcell, dnp, drpot, drxyz, drdefo, c->stream));
    ATT_HIP(c, hipStreamSynchronize(c->stream));
    ATT_HIP(c, hipMemcpy(xyz, dxyz, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
    ATT_HIP(c, hipMemcpy(pot, dpot, sizeof(float) * 4 * n, hipMemcpyDeviceToHost));
    ATT_HIP(c, hipMemcpy(cell, dcell, sizeof(int) * 9 * n, hipMemcpyDeviceToHost));
    ATT_HIP(c, hipMemcpy(nl_pairs, dnp, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
    ATT_HIP(c, hipMemcpy(rec_pot, drpot, sizeof(float) * 3 * nr, hipMemcpyDeviceToHost));
    ATT_HIP(c, hipMemcpy(rec_xyz, drxyz, sizeof(float) * 3 * nr, hipMemcpyDeviceToHost));
    return ATT_OK;
}

att_status att_debug_rotmats(att_ctx* c, const float* angles, size_t n, float* R) {
    if (!c || (n && (!angles || !R))) return fail(c, ATT_ERR_INVALID, "null argument");
    if (n == 0) return ATT_OK;
    if (n > (1u << 30)) return fail(c, ATT_ERR_UNSUPPORTED, "too many angle triples");
    ATT_HIP(c, hipSetDevice(c->device));
    Tmp t;
    float *da, *dR;
    ATT_HIP(c, t.alloc(&da, 3 * n));
    ATT_HIP(c, t.alloc(&dR, 9 * n));
    ATT_HIP(c, hipMemcpy(da, angles, sizeof(float) * 3 * n, hipMemcpyHostToDevice));
    ATT_HIP(c, launch_debug_rotmats(da, (unsigned)n, dR, c->stream));
    ATT_HIP(c, hipStreamSynchronize(c->stream));
    ATT_HIP(c, hipMemcpy(R, dR, sizeof(float) * 9 * n, hipMemcpyDeviceToHost));
    return ATT_OK;
}

att_status att_debug_sincosf(att_ctx* c, const float* x, size_t n, float* sn, float* cs) {
    if (!c || (n && (!x || !sn || !cs))) return fail(c, ATT_ERR_INVALID, "null argument");
    if (n == 0) return ATT_OK;
    if (n > (1u << 30)) return fail(c, ATT_ERR_UNSUPPORTED, "too many arguments");
    ATT_HIP(c, hipSetDevice(c->device));
    Tmp t;
    float *dx, *ds, *dc;
    ATT_HIP(c, t.alloc(&dx, n));
    ATT_HIP(c, t.alloc(&ds, n));
    ATT_HIP(c, t.alloc(&dc, n));
    ATT_HIP(c, hipMemcpy(dx, x, sizeof(float) * n, hipMemcpyHostToDevice));
    ATT_HIP(c, launch_debug_sincosf(dx, (unsigned)n, ds, dc, c->stream));
    ATT_HIP(c, hipStreamSynchronize(c->stream));
    ATT_HIP(c, hipMemcpy(sn, ds, sizeof(float) * n, hipMemcpyDeviceToHost));
    ATT_HIP(c, hipMemcpy(cs, dc, sizeof(float) * n, hipMemcpyDeviceToHost));
    return ATT_OK;
}

att_status att_debug_plateau_ratio(att_ctx* c, float P2, size_t n, float* ratio) {
    if (!c || (n && !ratio)) return fail(c, ATT_ERR_INVALID, "null argument");
    if (n == 0) return ATT_OK;
    if (n > (1u << 30)) return fail(c, ATT_ERR_UNSUPPORTED, "too many entries");
    ATT_HIP(c, hipSetDevice(c->device));
    Tmp t;
    float* dr;
    ATT_HIP(c, t.alloc(&dr, n));
    ATT_HIP(c, launch_debug_plateau_ratio(P2, (unsigned)n, dr, c->stream));
    ATT_HIP(c, hipStreamSynchronize(c->stream));
    ATT_HIP(c, hipMemcpy(ratio, dr, sizeof(float) * n, hipMemcpyDeviceToHost));
    return ATT_OK;
}


att_status att_sync(att_ctx* c, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    ATT_HIP(c, hipSetDevice(c->device));
    ATT_HIP(c, hipStreamSynchronize(stream ? (hipStream_t)stream : c->stream));
    return ATT_OK;
}

att_status att_debug_copy(att_ctx* c, void* d_dst, const void* d_src, size_t bytes, void* stream) {
    if (!c || (bytes && (!d_dst || !d_src))) return fail(c, ATT_ERR_INVALID, "null argument");
    if (bytes % 16) return fail(c, ATT_ERR_INVALID, "copy size must be a multiple of 16 bytes");
    ATT_HIP(c, hipSetDevice(c->device));
    ATT_HIP(c, launch_copy(d_dst, d_src, bytes / 16, stream ? (hipStream_t)stream : c->stream));
    return ATT_OK;
}

att_status att_dev_alloc(att_ctx* c, size_t bytes, void** d_ptr) {
    if (!c || !d_ptr) return fail(c, ATT_ERR_INVALID, "null argument");
    *d_ptr = nullptr;
    ATT_HIP(c, hipSetDevice(c->device));
    ATT_HIP(c, hipMalloc(d_ptr, bytes ? bytes : 1));
    return ATT_OK;
}

att_status att_dev_free(att_ctx* c, void* d_ptr) {
    if (!c) return ATT_ERR_INVALID;
    ATT_HIP(c, hipSetDevice(c->device));
    ATT_HIP(c, hipFree(d_ptr));
    return ATT_OK;
}

att_status att_host_alloc(att_ctx* c, size_t bytes, void** h_ptr) {
    if (!c || !h_ptr) return fail(c, ATT_ERR_INVALID, "null argument");
    *h_ptr = nullptr;
    ATT_HIP(c, hipSetDevice(c->device));
    ATT_HIP(c, hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return ATT_OK;
}

att_status att_host_free(att_ctx* c, void* h_ptr) {
    if (!c) return ATT_ERR_INVALID;
    ATT_HIP(c, hipHostFree(h_ptr));
    return ATT_OK;
}

att_status att_memcpy_h2d(att_ctx* c, void* d_dst, const void* src, size_t bytes, void* stream) {
    if (!c || (bytes && (!d_dst || !src))) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    ATT_HIP(c, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, s));
    ATT_HIP(c, hipStreamSynchronize(s));
    return ATT_OK;
}

att_status att_memcpy_d2h(att_ctx* c, void* dst, const void* d_src, size_t bytes, void* stream) {
    if (!c || (bytes && (!dst || !d_src))) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    ATT_HIP(c, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, s));
    ATT_HIP(c, hipStreamSynchronize(s));
    return ATT_OK;
}

att_status att_get_stream(att_ctx* c, void** stream) {
    if (!c || !stream) return fail(c, ATT_ERR_INVALID, "null argument");
    *stream = (void*)c->stream;
    return ATT_OK;
}

att_status att_event_create(att_ctx* c, void** ev) {
    if (!c || !ev) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    hipEvent_t e;
    ATT_HIP(c, hipEventCreate(&e));
    *ev = (void*)e;
    return ATT_OK;
}

att_status att_event_destroy(att_ctx* c, void* ev) {
    if (!c) return ATT_ERR_INVALID;
    ATT_HIP(c, hipEventDestroy((hipEvent_t)ev));
    return ATT_OK;
}

att_status att_event_record(att_ctx* c, void* ev, void* stream) {
    if (!c || !ev) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipEventRecord((hipEvent_t)ev, stream ? (hipStream_t)stream : c->stream));
    return ATT_OK;
}

att_status att_event_elapsed_ms(att_ctx* c, void* a, void* b, float* ms) {
    if (!c || !a || !b || !ms) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipEventSynchronize((hipEvent_t)b));
    ATT_HIP(c, hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return ATT_OK;
}

att_status att_device_memory(att_ctx* c, size_t* free_bytes, size_t* total_bytes) {
    if (!c || !free_bytes || !total_bytes) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    ATT_HIP(c, hipMemGetInfo(free_bytes, total_bytes));
    return ATT_OK;
}

att_status att_stream_create(att_ctx* c, void** stream) {
    if (!c || !stream) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    hipStream_t s;
    ATT_HIP(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return ATT_OK;
}

att_status att_stream_destroy(att_ctx* c, void* stream) {
    if (!c || !stream) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    ATT_HIP(c, hipStreamDestroy((hipStream_t)stream));
    return ATT_OK;
}

att_status att_memcpy_h2d_async(att_ctx* c, void* d_dst, const void* src, size_t bytes, void* stream) {
    if (!c || (bytes && (!d_dst || !src))) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    ATT_HIP(c, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice,
                              stream ? (hipStream_t)stream : c->stream));
    return ATT_OK;
}

att_status att_memcpy_d2h_async(att_ctx* c, void* dst, const void* d_src, size_t bytes, void* stream) {
    if (!c || (bytes && (!dst || !d_src))) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    ATT_HIP(c, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost,
                              stream ? (hipStream_t)stream : c->stream));
    return ATT_OK;
}

att_status att_event_query(att_ctx* c, void* ev) {
    if (!c || !ev) return fail(c, ATT_ERR_INVALID, "null argument");
    const hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipErrorNotReady) return ATT_NOT_READY;
    ATT_HIP(c, e);
    return ATT_OK;
}

att_status att_event_sync(att_ctx* c, void* ev) {
    if (!c || !ev) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipEventSynchronize((hipEvent_t)ev));
    return ATT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches
// Many receptor-ligand pairs of uneven sizes scored by a few persistent
// launches.  Pairs are grouped into buckets by mode capacity (0, <=5, <=10)
// and ligand size (<192, <384, <640, <1024, larger beads); each bucket runs one
// k_adjoint_batch launch whose shape (waves per pose, poses per tile) the
// autotuner picks by timing the candidates on the bucket's own work.  The
// launches are captured into one hipGraph, replayed by att_batch_run.
// Each pair's poses are put in spatial order when the batch is created (from
// the DOFs its buffers hold then; the order changes no result, only locality),
// and each bucket's tiles are dealt into one queue per XCD: queue x holds the
// x-th eighth of every pair's poses.
struct att_batch {
    struct Bucket {
        unsigned mm = 0;
        int sizeClass = 0;
        std::vector<unsigned> pairs;  // indices into att_batch::pairs
        size_t lds = 0;               // dynamic LDS: the largest pair of the bucket
        int waves = 4;
        unsigned tile = 16;
        std::vector<BatchTile> tiles;  // queue by queue
        unsigned queueStart[kBatchQueues + 1] = {};
        BatchTile* d_tiles = nullptr;  // the tiles, then queueStart
        size_t d_tiles_cap = 0;
        float ms = -1.f;              // time of the chosen shape when tuned
    };
    std::vector<PairDesc> pairs;
    std::vector<double> cost;         // per pair, relative cost of one pose
    PairDesc* d_pairs = nullptr;
    std::vector<Bucket> buckets;
    unsigned* d_counters = nullptr;   // kBatchQueues per bucket
    unsigned* d_order = nullptr;      // spatial order of the pairs that have one
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
};

namespace {

// Ligand size classes: the bucket's LDS is its largest ligand's, so classes
// narrow enough that a pair runs near its own occupancy (timed: five classes
// beat three by 2.7% on 64 uneven pairs).
int size_class(unsigned ligN) {
    return ligN < 192 ? 0 : ligN < 384 ? 1 : ligN < 640 ? 2 : ligN < 1024 ? 3 : 4;
}

att_status batch_tiles(att_ctx* c, att_batch* b, att_batch::Bucket& k) {
    k.tiles.clear();
    std::vector<unsigned> order = k.pairs;
    // expensive pairs first, so the last tiles taken are the cheap ones
    std::stable_sort(order.begin(), order.end(),
                     [&](unsigned x, unsigned y) { return b->cost[x] > b->cost[y]; });
    for (unsigned x = 0; x < kBatchQueues; ++x) {
        k.queueStart[x] = (unsigned)k.tiles.size();
        for (unsigned p : order) {
            const unsigned n = b->pairs[p].n;
            const unsigned lo = (unsigned)((uint64_t)n * x / kBatchQueues);
            const unsigned hi = (unsigned)((uint64_t)n * (x + 1) / kBatchQueues);
            for (unsigned f = lo; f < hi; f += k.tile)
                k.tiles.push_back(BatchTile{p, f, std::min(k.tile, hi - f)});
        }
    }
    k.queueStart[kBatchQueues] = (unsigned)k.tiles.size();
    const size_t tileBytes = sizeof(BatchTile) * k.tiles.size();
    const size_t bytes = tileBytes + sizeof(k.queueStart);
    att_status st;
    if ((st = grow(c, reinterpret_cast<void**>(&k.d_tiles), &k.d_tiles_cap, bytes)) != ATT_OK) return st;
    if (tileBytes) ATT_HIP(c, hipMemcpy(k.d_tiles, k.tiles.data(), tileBytes, hipMemcpyHostToDevice));
    ATT_HIP(c, hipMemcpy(reinterpret_cast<char*>(k.d_tiles) + tileBytes, k.queueStart,
                         sizeof(k.queueStart), hipMemcpyHostToDevice));
    b->dirty = true;
    return ATT_OK;
}

hipError_t batch_launch(att_ctx* c, att_batch* b, size_t i, hipStream_t s) {
    att_batch::Bucket& k = b->buckets[i];
    const size_t resident =
        (size_t)c->numCU * adjoint_blocks_per_cu(k.mm, k.waves, true, k.lds);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(k.tiles.size(), resident));
    const unsigned* qs = reinterpret_cast<const unsigned*>(k.d_tiles + k.tiles.size());
    return launch_adjoint_batch(k.mm, k.waves, b->d_pairs, k.d_tiles, qs, (unsigned)k.tiles.size(),
                                b->d_counters + kBatchQueues * i, blocks, k.lds, s);
}

// Spatial order of every pair with enough poses to span several tiles per
// queue, sorted from the DOFs the pair's buffer holds now (in a scratch buffer
// of its own, freed before returning); sets PairDesc::order before the
// descriptors go to the device.
constexpr unsigned kBatchOrderMin = 256;

att_status batch_order(att_ctx* c, att_batch* b) {
    if (!c->poseOrder) return ATT_OK;
    size_t total = 0, scratch = 0;
    for (const PairDesc& P : b->pairs)
        if (P.n >= kBatchOrderMin) {
            total += P.n;
            scratch = std::max(scratch, pose_order_scratch(P.n));
        }
    if (total == 0) return ATT_OK;
    ATT_HIP(c, hipMalloc(&b->d_order, sizeof(unsigned) * total));
    void* tmp = nullptr;
    ATT_HIP(c, hipMalloc(&tmp, scratch));
    hipError_t e = hipSuccess;
    unsigned* o = b->d_order;
    for (PairDesc& P : b->pairs)
        if (P.n >= kBatchOrderMin && e == hipSuccess) {
            e = launch_pose_order(P.dofs, P.dofStride, P.n, tmp, scratch, o, c->stream);
            P.order = o;
            o += P.n;
        }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(c, ATT_ERR_HIP, std::string("pose order: ") + hipGetErrorString(e));
    return ATT_OK;
}

void batch_free_graph(att_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

}  // namespace

extern "C" {

att_status att_batch_destroy(att_ctx* c, att_batch* b);

}  // extern "C"

namespace {

att_status batch_fill(att_ctx* c, att_batch* b, const att_batch_pair* in, unsigned npairs) {
    std::map<std::pair<unsigned, int>, size_t> index;
    for (unsigned i = 0; i < npairs; ++i) {
        const att_batch_pair& q = in[i];
        ProtObj *rec, *lig;
        GridObj* g;
        ParObj* p;
        att_status st;
        if ((st = resolve(c, q.rec, q.lig, q.grid, q.params, &rec, &lig, &g, &p)) != ATT_OK) return st;
        if ((st = check_pair(c, rec, lig, g, p)) != ATT_OK) return st;
        if (lig->d.n > 0xFFFFu) return fail(c, ATT_ERR_UNSUPPORTED, "ligand exceeds 65535 beads");
        if (q.n > 0x7FFFFFFFull) return fail(c, ATT_ERR_UNSUPPORTED, "too many poses in one pair");
        if (q.n && (!q.d_dofs || !q.d_out)) return fail(c, ATT_ERR_INVALID, "null buffer");
        unsigned slots = 0;
        const size_t dyn = adjoint_lds_bytes(rec->d, lig->d, p->d, false, 1, &slots);
        if (dyn + adjoint_static_lds(4) > c->maxLdsPerBlock)
            return fail(c, ATT_ERR_UNSUPPORTED, "ligand too large for the LDS of the adjoint kernel");
        PairDesc P{};
        P.rec = rec->d; P.lig = lig->d;
        P.inner = g->inner; P.outer = g->outer; P.nl = g->nl;
        P.par = p->d;
        P.dofs = q.d_dofs; P.wts = q.d_w; P.out = q.d_out;
        P.dofStride = q.modes ? 26 : 6; P.outStride = q.modes ? 27 : 7;
        P.n = (unsigned)q.n; P.tabSlots = slots;
        const unsigned m = q.modes ? std::max(rec->d.M, lig->d.M) : 0;
        const unsigned mm = m == 0 ? 0 : m <= 5 ? 5 : ATT_MODES_MAX;
        const int sc = size_class(lig->d.n);
        auto key = std::make_pair(mm, sc);
        auto it = index.find(key);
        if (it == index.end()) {
            it = index.emplace(key, b->buckets.size()).first;
            b->buckets.emplace_back();
            b->buckets.back().mm = mm;
            b->buckets.back().sizeClass = sc;
        }
        att_batch::Bucket& k = b->buckets[it->second];
        k.pairs.push_back(i);
        k.lds = std::max(k.lds, dyn);
        b->pairs.push_back(P);
        b->cost.push_back((double)lig->d.n * (1.0 + 0.25 * m) + 0.05 * rec->d.n);
    }
    att_status st;
    if ((st = batch_order(c, b)) != ATT_OK) return st;
    if (npairs) {
        ATT_HIP(c, hipMalloc(&b->d_pairs, sizeof(PairDesc) * npairs));
        ATT_HIP(c, hipMemcpy(b->d_pairs, b->pairs.data(), sizeof(PairDesc) * npairs, hipMemcpyHostToDevice));
    }
    ATT_HIP(c, hipMalloc(&b->d_counters,
                         sizeof(unsigned) * kBatchQueues * std::max<size_t>(b->buckets.size(), 1)));
    for (auto& k : b->buckets)
        if ((st = batch_tiles(c, b, k)) != ATT_OK) return st;
    ATT_HIP(c, hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
    return ATT_OK;
}

}  // namespace

extern "C" {

att_status att_batch_create(att_ctx* c, const att_batch_pair* in, unsigned npairs, att_batch** out) {
    if (!c) return ATT_ERR_INVALID;
    if (!out || (npairs && !in)) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    att_batch* b = new att_batch;
    const att_status st = batch_fill(c, b, in, npairs);
    if (st != ATT_OK) {
        (void)att_batch_destroy(c, b);  // frees what was allocated before the failure
        return st;
    }
    *out = b;
    return ATT_OK;
}

att_status att_batch_destroy(att_ctx* c, att_batch* b) {
    if (!c || !b) return ATT_ERR_INVALID;
    (void)hipSetDevice(c->device);
    batch_free_graph(b);
    for (auto& k : b->buckets)
        if (k.d_tiles) (void)hipFree(k.d_tiles);
    if (b->d_pairs) (void)hipFree(b->d_pairs);
    if (b->d_counters) (void)hipFree(b->d_counters);
    if (b->d_order) (void)hipFree(b->d_order);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    delete b;
    return ATT_OK;
}

unsigned att_batch_buckets(const att_batch* b) { return b ? (unsigned)b->buckets.size() : 0u; }

att_status att_batch_bucket_info(att_ctx* c, const att_batch* b, unsigned bucket, att_batch_bucket* info) {
    if (!c) return ATT_ERR_INVALID;
    if (!b || !info || bucket >= b->buckets.size()) return fail(c, ATT_ERR_INVALID, "bad bucket");
    const att_batch::Bucket& k = b->buckets[bucket];
    info->modeCapacity = k.mm;
    info->sizeClass = k.sizeClass;
    info->numPairs = (unsigned)k.pairs.size();
    info->numTiles = (unsigned)k.tiles.size();
    info->wavesPerPose = k.waves;
    info->posesPerTile = k.tile;
    info->ms = k.ms;
    return ATT_OK;
}

att_status att_batch_pair_bucket(att_ctx* c, const att_batch* b, unsigned pair, unsigned* bucket) {
    if (!c) return ATT_ERR_INVALID;
    if (!b || !bucket || pair >= b->pairs.size()) return fail(c, ATT_ERR_INVALID, "bad pair");
    for (size_t k = 0; k < b->buckets.size(); ++k)
        for (unsigned q : b->buckets[k].pairs)
            if (q == pair) {
                *bucket = (unsigned)k;
                return ATT_OK;
            }
    return fail(c, ATT_ERR_INVALID, "pair in no bucket");
}

att_status att_batch_set_shape(att_ctx* c, att_batch* b, unsigned bucket, int wavesPerPose,
                               unsigned posesPerTile) {
    if (!c) return ATT_ERR_INVALID;
    if (!b || bucket >= b->buckets.size()) return fail(c, ATT_ERR_INVALID, "bad bucket");
    if ((wavesPerPose != 1 && wavesPerPose != 2 && wavesPerPose != 4) || posesPerTile == 0)
        return fail(c, ATT_ERR_INVALID, "waves per pose must be 1, 2 or 4 and tiles non-empty");
    ATT_HIP(c, hipSetDevice(c->device));
    att_batch::Bucket& k = b->buckets[bucket];
    if (k.lds + adjoint_static_lds(wavesPerPose) > c->maxLdsPerBlock)
        return fail(c, ATT_ERR_UNSUPPORTED, "shape does not fit the LDS");
    k.waves = wavesPerPose;
    k.ms = -1.f;
    if (k.tile != posesPerTile || k.tiles.empty()) {
        k.tile = posesPerTile;
        att_status st;
        if ((st = batch_tiles(c, b, k)) != ATT_OK) return st;
    }
    b->dirty = true;
    return ATT_OK;
}

att_status att_batch_tune(att_ctx* c, att_batch* b, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (!b) return fail(c, ATT_ERR_INVALID, "null batch");
    ATT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    hipEvent_t e0, e1;
    ATT_HIP(c, hipEventCreate(&e0));
    ATT_HIP(c, hipEventCreate(&e1));
    static const int kWaves[] = {1, 2, 4};
    static const unsigned kTiles[] = {1, 2, 4, 16};
    att_status st = ATT_OK;
    for (size_t i = 0; i < b->buckets.size() && st == ATT_OK; ++i) {
        att_batch::Bucket& k = b->buckets[i];
        float best = -1.f;
        int bw = 4;
        unsigned bt = 16;
        for (int w : kWaves) {
            if (k.lds + adjoint_static_lds(w) > c->maxLdsPerBlock) continue;
            for (unsigned tl : kTiles) {
                if ((st = att_batch_set_shape(c, b, (unsigned)i, w, tl)) != ATT_OK) break;
                hipError_t e = batch_launch(c, b, i, s);  // warm-up
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventRecord(e0, s);
                if (e == hipSuccess) e = batch_launch(c, b, i, s);
                if (e == hipSuccess) e = batch_launch(c, b, i, s);
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e != hipSuccess) {
                    st = fail(c, ATT_ERR_HIP, std::string("autotuner: ") + hipGetErrorString(e));
                    break;
                }
                ms *= 0.5f;
                if (best < 0.f || ms < best) { best = ms; bw = w; bt = tl; }
            }
            if (st != ATT_OK) break;
        }
        if (st == ATT_OK && (st = att_batch_set_shape(c, b, (unsigned)i, bw, bt)) == ATT_OK) k.ms = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

att_status att_batch_run(att_ctx* c, att_batch* b, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (!b) return fail(c, ATT_ERR_INVALID, "null batch");
    ATT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (b->buckets.empty()) return ATT_OK;
    if (b->dirty) {
        batch_free_graph(b);
        ATT_HIP(c, hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
        // one bucket after the other: buckets on parallel graph branches mix
        // pairs in the caches (timed 1.6% slower on two branches, 6% on four)
        hipError_t e = hipSuccess;
        for (size_t i = 0; i < b->buckets.size() && e == hipSuccess; ++i) e = batch_launch(c, b, i, b->capture);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(b->capture, &g);
        if (e != hipSuccess || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(c, ATT_ERR_HIP, std::string("batch capture: ") +
                                            hipGetErrorString(e != hipSuccess ? e : ee));
        }
        b->graph = g;
        ATT_HIP(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->dirty = false;
    }
    ATT_HIP(c, hipGraphLaunch(b->exec, s));
    return ATT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- minimizer
// A minimization job owns its device state (VA13 states, request lists and
// points, scores) and its scoring descriptor; rounds are issued on a stream
// without host synchronization, the number of running poses can be copied to
// host memory asynchronously.  att_minimize_dev runs one job to completion;
// the serving engine interleaves many.
struct att_minjob {
    MinRun R{};
    PairDesc P{};
    bool modes = false;
    size_t dyn = 0;
    unsigned blocks = 1;
    int waves = 2;
    void* mem = nullptr;
    size_t bytes = 0;
    unsigned* list[2] = {nullptr, nullptr};
    unsigned* count[2] = {nullptr, nullptr};
    float* req[2] = {nullptr, nullptr};
    float* res = nullptr;
    int cur = 0;
    int rounds = 0;
};

namespace {
size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

void minjob_layout(size_t n, int dim, unsigned ds, unsigned os, size_t* parts) {
    parts[0] = al256(sizeof(double) * n * va_elements(dim));
    parts[1] = al256(sizeof(VaState) * n);
    parts[2] = al256(sizeof(unsigned) * n);
    parts[3] = al256(2 * sizeof(unsigned));
    parts[4] = al256(sizeof(float) * n * ds);
    parts[5] = al256(sizeof(float) * n * os);
}
}  // namespace

extern "C" {

size_t att_minjob_bytes(size_t n, int dim, int modes) {
    size_t p[6];
    minjob_layout(n, dim, modes ? 26 : 6, modes ? 27 : 7, p);
    return p[0] + p[1] + 2 * p[2] + p[3] + 2 * p[4] + p[5];
}

att_status att_minjob_create(att_ctx* c, att_handle hr, att_handle hl, att_handle hg, att_handle hp,
                             int modes, size_t n, int maxFunEval, float* d_dofs_out,
                             float* d_energy, float* d_ebest, int* d_nfun, att_minjob** out) {
    if (!c) return ATT_ERR_INVALID;
    if (!out || (n && (!d_dofs_out || !d_energy)) || maxFunEval < 0)
        return fail(c, ATT_ERR_INVALID, "null buffer or negative evaluation limit");
    if (n == 0 || n > 0x7FFFFFFFull) return fail(c, ATT_ERR_UNSUPPORTED, "pose count out of range");
    ATT_HIP(c, hipSetDevice(c->device));
    ProtObj *rec, *lig;
    GridObj* g;
    ParObj* p;
    att_status st;
    if ((st = resolve(c, hr, hl, hg, hp, &rec, &lig, &g, &p)) != ATT_OK) return st;
    if ((st = check_pair(c, rec, lig, g, p)) != ATT_OK) return st;
    if (lig->d.n > 0xFFFFu) return fail(c, ATT_ERR_UNSUPPORTED, "ligand exceeds 65535 beads");
    std::unique_ptr<att_minjob> j(new att_minjob);
    unsigned slots = 0;
    j->dyn = adjoint_lds_bytes(rec->d, lig->d, p->d, false, 1, &slots);
    const int kWaves = pair_waves(lig->d.n);  // the single-pair adjoint's shape: same results
    j->waves = kWaves;
    if (j->dyn + adjoint_static_lds(kWaves) > c->maxLdsPerBlock)
        return fail(c, ATT_ERR_UNSUPPORTED, "ligand too large for the LDS of the adjoint kernel");
    MinRun& R = j->R;
    R.npose = (unsigned)n;
    R.Mr = modes ? (int)rec->d.M : 0;
    R.Ml = modes ? (int)lig->d.M : 0;
    R.dim = 6 + R.Mr + R.Ml;
    R.ivmax = maxFunEval;
    R.modes = modes != 0;
    R.ds = modes ? 26 : 6;
    R.os = modes ? 27 : 7;
    size_t pt[6];
    minjob_layout(n, R.dim, R.ds, R.os, pt);
    j->bytes = pt[0] + pt[1] + 2 * pt[2] + pt[3] + 2 * pt[4] + pt[5];
    ATT_HIP(c, hipMalloc(&j->mem, j->bytes));
    char* base = static_cast<char*>(j->mem);
    R.st = reinterpret_cast<double*>(base);
    R.sc = reinterpret_cast<VaState*>(base + pt[0]);
    j->list[0] = reinterpret_cast<unsigned*>(base + pt[0] + pt[1]);
    j->list[1] = reinterpret_cast<unsigned*>(base + pt[0] + pt[1] + pt[2]);
    j->count[0] = reinterpret_cast<unsigned*>(base + pt[0] + pt[1] + 2 * pt[2]);
    j->count[1] = j->count[0] + 1;
    j->req[0] = reinterpret_cast<float*>(base + pt[0] + pt[1] + 2 * pt[2] + pt[3]);
    j->req[1] = reinterpret_cast<float*>(base + pt[0] + pt[1] + 2 * pt[2] + pt[3] + pt[4]);
    j->res = reinterpret_cast<float*>(base + pt[0] + pt[1] + 2 * pt[2] + pt[3] + 2 * pt[4]);
    R.dofsOut = d_dofs_out;
    R.eLast = d_energy;
    R.eBest = d_ebest;
    R.nfun = d_nfun;
    PairDesc& P = j->P;
    P.rec = rec->d; P.lig = lig->d;
    P.inner = g->inner; P.outer = g->outer; P.nl = g->nl;
    P.par = p->d;
    P.bin = g->bin; P.bout = g->bout;
    P.out = j->res;
    P.dofStride = R.ds; P.outStride = R.os;
    P.n = (unsigned)n; P.tabSlots = slots;
    j->modes = modes != 0;
    const unsigned mm = modes ? std::max(rec->d.M, lig->d.M) : 0;
    const size_t resident = (size_t)c->numCU * adjoint_blocks_per_cu(mm, kWaves, false, j->dyn);
    j->blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(n, resident));
    *out = j.release();
    return ATT_OK;
}

att_status att_minjob_start(att_ctx* c, att_minjob* j, const float* d_dofs, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (!j || !d_dofs) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    j->cur = 0;
    j->rounds = 0;
    j->R.list = j->list[0]; j->R.count = j->count[0]; j->R.req = j->req[0];
    ATT_HIP(c, launch_min_start(j->R, d_dofs, s));
    return ATT_OK;
}

att_status att_minjob_rounds(att_ctx* c, att_minjob* j, int rounds, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (!j) return fail(c, ATT_ERR_INVALID, "null job");
    ATT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    for (int r = 0; r < rounds; ++r) {
        const int cur = j->cur, nxt = cur ^ 1;
        j->P.dofs = j->req[cur];
        j->P.nDev = j->count[cur];
        ATT_HIP(c, launch_adjoint(j->P, j->modes, nullptr, j->blocks, j->waves, j->dyn, s));
        j->R.list = j->list[nxt]; j->R.count = j->count[nxt]; j->R.req = j->req[nxt];
        ATT_HIP(c, launch_min_step(j->R, j->list[cur], j->count[cur], j->req[cur], j->res, s));
        j->cur = nxt;
        ++j->rounds;
    }
    return ATT_OK;
}

int att_minjob_round_limit(const att_minjob* j) { return j ? j->R.ivmax + 2 : 0; }

att_status att_minjob_running(att_ctx* c, att_minjob* j, unsigned* h_count, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (!j || !h_count) return fail(c, ATT_ERR_INVALID, "null argument");
    ATT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    ATT_HIP(c, hipMemcpyAsync(h_count, j->count[j->cur], sizeof(unsigned), hipMemcpyDeviceToHost, s));
    return ATT_OK;
}

att_status att_minjob_destroy(att_ctx* c, att_minjob* j) {
    if (!c || !j) return ATT_ERR_INVALID;
    (void)hipSetDevice(c->device);
    if (j->mem) (void)hipFree(j->mem);
    delete j;
    return ATT_OK;
}

att_status att_minimize_dev(att_ctx* c, att_handle hr, att_handle hl, att_handle hg, att_handle hp,
                            const float* d_dofs, int modes, size_t n, int maxFunEval,
                            float* d_dofs_out, float* d_energy, float* d_ebest, int* d_nfun,
                            void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if ((n && (!d_dofs || !d_dofs_out || !d_energy)) || maxFunEval < 0)
        return fail(c, ATT_ERR_INVALID, "null buffer or negative evaluation limit");
    if (n == 0) {
        // still validate the handles
        ProtObj *rec, *lig;
        GridObj* g;
        ParObj* p;
        att_status st;
        if ((st = resolve(c, hr, hl, hg, hp, &rec, &lig, &g, &p)) != ATT_OK) return st;
        return check_pair(c, rec, lig, g, p);
    }
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    att_minjob* j = nullptr;
    att_status st = att_minjob_create(c, hr, hl, hg, hp, modes, n, maxFunEval, d_dofs_out, d_energy,
                                      d_ebest, d_nfun, &j);
    if (st != ATT_OK) return st;
    unsigned* running = nullptr;
    if (hipHostMalloc(reinterpret_cast<void**>(&running), sizeof(unsigned)) != hipSuccess) {
        att_minjob_destroy(c, j);
        return fail(c, ATT_ERR_NOMEM, "pinned host word");
    }
    st = att_minjob_start(c, j, d_dofs, s);
    // every pose is evaluated at most maxFunEval + 1 times
    const int limit = att_minjob_round_limit(j);
    while (st == ATT_OK && j->rounds < limit) {
        st = att_minjob_rounds(c, j, std::min(8, limit - j->rounds), s);
        if (st == ATT_OK) st = att_minjob_running(c, j, running, s);
        if (st == ATT_OK && hipStreamSynchronize(s) != hipSuccess) st = fail(c, ATT_ERR_HIP, "synchronize");
        if (st == ATT_OK && *running == 0) break;
    }
    if (st == ATT_OK && hipStreamSynchronize(s) != hipSuccess) st = fail(c, ATT_ERR_HIP, "synchronize");
    (void)hipHostFree(running);
    att_minjob_destroy(c, j);
    return st;
}

att_status att_minimize(att_ctx* c, att_handle hr, att_handle hl, att_handle hg, att_handle hp,
                        const float* dofs, int modes, size_t n, int maxFunEval, float* dofs_out,
                        float* energy, float* ebest, int* nfun, void* stream) {
    if (!c) return ATT_ERR_INVALID;
    if (n && (!dofs || !dofs_out || !energy)) return fail(c, ATT_ERR_INVALID, "null buffer");
    ATT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const size_t ds = modes ? 26 : 6;
    Tmp tmp;
    float *d_in, *d_out, *d_e, *d_eb = nullptr;
    int* d_nf = nullptr;
    ATT_HIP(c, tmp.alloc(&d_in, n * ds));
    ATT_HIP(c, tmp.alloc(&d_out, n * ds));
    ATT_HIP(c, tmp.alloc(&d_e, n));
    if (ebest) ATT_HIP(c, tmp.alloc(&d_eb, n));
    if (nfun) ATT_HIP(c, tmp.alloc(&d_nf, n));
    ATT_HIP(c, hipMemcpyAsync(d_in, dofs, sizeof(float) * n * ds, hipMemcpyHostToDevice, s));
    att_status st = att_minimize_dev(c, hr, hl, hg, hp, d_in, modes, n, maxFunEval, d_out, d_e, d_eb,
                                     d_nf, s);
    if (st != ATT_OK) return st;
    ATT_HIP(c, hipMemcpyAsync(dofs_out, d_out, sizeof(float) * n * ds, hipMemcpyDeviceToHost, s));
    ATT_HIP(c, hipMemcpyAsync(energy, d_e, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    if (ebest) ATT_HIP(c, hipMemcpyAsync(ebest, d_eb, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    if (nfun) ATT_HIP(c, hipMemcpyAsync(nfun, d_nf, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    ATT_HIP(c, hipStreamSynchronize(s));
    return ATT_OK;
}

}  // extern "C"

// --- grid generation on the device -----------------------------------------

namespace {

// Device copies of a body's beads and of the pair table, arranged for the
// generator kernels (gridgen.hip).
struct GenInputs {
    Tmp tmp;
    GenSrc src{};
    float4* nlBeads = nullptr;
    unsigned* nlIdx = nullptr;
    unsigned nNl = 0;
    GenPair* tab = nullptr;
    bool linear = true;
    double P2 = 0.0;
};

att_status check_body(att_ctx* c, const att_protein_desc* b) {
    if (!b || !b->xyz || !b->type || !b->charge)
        return fail(c, ATT_ERR_INVALID, "body: missing coordinates, types or charges");
    if (b->n == 0) return fail(c, ATT_ERR_INVALID, "body has no beads");
    return ATT_OK;
}

att_status check_geom(att_ctx* c, const att_grid_geom* g, unsigned numGrids, const char* what) {
    if (!g) return fail(c, ATT_ERR_INVALID, std::string(what) + ": missing geometry");
    if (g->nx < 2 || g->ny < 2 || g->nz < 2)
        return fail(c, ATT_ERR_INVALID, std::string(what) + ": grid needs >= 2 points per axis");
    if (!(g->spacing > 0.f) || !std::isfinite(g->spacing) || !std::isfinite(g->ox) ||
        !std::isfinite(g->oy) || !std::isfinite(g->oz))
        return fail(c, ATT_ERR_INVALID, std::string(what) + ": spacing must be positive, origin finite");
    if ((uint64_t)g->nx * g->ny * g->nz * numGrids * 8 >= (1ull << 34))
        return fail(c, ATT_ERR_UNSUPPORTED, std::string(what) + ": grid exceeds 2^31 points");
    return ATT_OK;
}

GenGeom gen_geom(const att_grid_geom& g) {
    return GenGeom{g.nx, g.ny, g.nz, (double)g.ox, (double)g.oy, (double)g.oz, (double)g.spacing};
}

att_grid_desc grid_of(const att_grid_geom& g, unsigned numGrids) {
    att_grid_desc d{};
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz;
    d.numGrids = numGrids;
    d.ox = g.ox; d.oy = g.oy; d.oz = g.oz;
    d.spacing = g.spacing;
    return d;
}

// Beads with type 1..T feed the potential grids (sorted stably by type for the
// per-type sums, and in index order), the charged ones among them grid 0;
// beads with type != 0 feed the neighbour lists.  The table is seen from
// (source type a, probe type t): a receptor's probes are ligand types,
// table[(a-1)*T + (t-1)]; a ligand's probes are receptor types,
// table[(t-1)*T + (a-1)].
att_status gen_prepare(att_ctx* c, const att_protein_desc* b, const att_pair_params* table,
                       unsigned T, int srcIsLigand, float plateau, GenInputs& in) {
    const unsigned n = b->n;
    const float* X = b->xyz;
    const float* Y = b->xyz + n;
    const float* Z = b->xyz + 2 * (size_t)n;
    std::vector<float4> act, chg, nlb;
    std::vector<unsigned> actType, nlIdx;
    std::vector<unsigned> typeStart(T + 2, 0);
    for (unsigned j = 0; j < n; ++j) {
        const unsigned t = b->type[j];
        if (t != 0) {
            nlb.push_back(make_float4(X[j], Y[j], Z[j], 0.f));
            nlIdx.push_back(j);
        }
        if (T == 0 || t == 0 || t > T) continue;
        act.push_back(make_float4(X[j], Y[j], Z[j], 0.f));
        actType.push_back(t);
        typeStart[t + 1]++;
        if (b->charge[j] != 0.f) chg.push_back(make_float4(X[j], Y[j], Z[j], b->charge[j]));
    }
    for (unsigned a = 1; a <= T; ++a) typeStart[a + 1] += typeStart[a];
    std::vector<float4> sorted(act.size());
    {
        std::vector<unsigned> pos(typeStart.begin(), typeStart.end());
        for (size_t k = 0; k < act.size(); ++k) sorted[pos[actType[k]]++] = act[k];
    }
    in.P2 = (double)plateau * (double)plateau;
    std::vector<GenPair> tab((size_t)(T + 1) * (T + 1), GenPair{0, 0, 0, 0, 0});
    in.linear = true;
    for (unsigned a = 1; a <= T; ++a)
        for (unsigned t = 1; t <= T; ++t) {
            const att_pair_params& p = table[srcIsLigand ? (t - 1) * T + (a - 1) : (a - 1) * T + (t - 1)];
            tab[(size_t)a * (T + 1) + t] =
                GenPair{(double)p.rc, (double)p.ac, (double)p.ipon, (double)p.rmin2, (double)p.emin};
            if ((double)p.rmin2 > in.P2) in.linear = false;
        }
    auto up = [&](auto** dst, const auto& v) -> hipError_t {
        hipError_t e = in.tmp.alloc(dst, v.size());
        if (e == hipSuccess && !v.empty())
            e = hipMemcpy(*dst, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice);
        return e;
    };
    float4 *dsorted, *dact, *dchg;
    unsigned *dstart, *dtype;
    ATT_HIP(c, up(&dsorted, sorted));
    ATT_HIP(c, up(&dstart, typeStart));
    ATT_HIP(c, up(&dact, act));
    ATT_HIP(c, up(&dtype, actType));
    ATT_HIP(c, up(&dchg, chg));
    ATT_HIP(c, up(&in.nlBeads, nlb));
    ATT_HIP(c, up(&in.nlIdx, nlIdx));
    ATT_HIP(c, up(&in.tab, tab));
    in.src = GenSrc{dsorted, dstart, dact, dtype, (unsigned)act.size(), dchg, (unsigned)chg.size()};
    in.nNl = (unsigned)nlb.size();
    return ATT_OK;
}

att_status gen_grid(att_ctx* c, const GenInputs& in, const att_grid_geom& g, unsigned T, int shape,
                    int dielec, float ffelec, float4* d_out) {
    ATT_HIP(c, launch_gen_grid(gen_geom(g), in.src, in.tab, T, shape, in.linear, in.P2,
                               std::sqrt(in.P2), dielec, (double)ffelec, d_out, c->stream));
    return ATT_OK;
}

// Neighbour lists of the points of g: desc (npts uint2) and list (total
// entries) allocated by `alloc`, which keeps or frees them.
template <class Alloc>
att_status gen_nl(att_ctx* c, const GenInputs& in, const att_grid_geom& g, float neighbourdis,
                  Alloc&& alloc, uint2** d_desc, uint32_t** d_list, uint64_t* total) {
    const size_t npts = (size_t)g.nx * g.ny * g.nz;
    const double nd = neighbourdis;
    const double nd2 = nd * nd;
    const double cr = nd + 0.01;
    const float cullR2 = (float)(cr * cr);
    Tmp tmp;
    uint32_t* cnt;
    unsigned long long* scratch;
    ATT_HIP(c, tmp.alloc(&cnt, npts));
    ATT_HIP(c, tmp.alloc(&scratch, gen_scan_scratch(npts) / sizeof(unsigned long long)));
    ATT_HIP(c, alloc((void**)d_desc, sizeof(uint2) * npts));
    const GenGeom gg = gen_geom(g);
    ATT_HIP(c, launch_gen_nl_count(gg, in.nlBeads, in.nlIdx, in.nNl, nd2, cullR2, cnt, *d_desc,
                                   scratch, c->stream));
    unsigned long long tot = 0;
    const size_t nsb = (npts + 1023) / 1024;
    ATT_HIP(c, hipMemcpyAsync(&tot, scratch + nsb, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
    ATT_HIP(c, hipStreamSynchronize(c->stream));
    if (tot >= (1ull << 32)) return fail(c, ATT_ERR_UNSUPPORTED, "neighbour list exceeds 2^32 entries");
    *total = tot;
    *d_list = nullptr;
    if (tot > 0) {
        ATT_HIP(c, alloc((void**)d_list, sizeof(uint32_t) * tot));
        ATT_HIP(c, launch_gen_nl_fill(gg, in.nlBeads, in.nlIdx, in.nNl, nd2, cullR2, *d_desc,
                                      *d_list, c->stream));
    }
    ATT_HIP(c, hipStreamSynchronize(c->stream));
    return ATT_OK;
}

att_status check_gen_args(att_ctx* c, const att_protein_desc* body, const att_pair_params* table,
                          unsigned T, int shape, int dielec, float plateau) {
    att_status st;
    if ((st = check_body(c, body)) != ATT_OK) return st;
    if (!table) return fail(c, ATT_ERR_INVALID, "missing pair table");
    if (T == 0 || T > 99) return fail(c, ATT_ERR_INVALID, "number of types must be 1..99");
    if (shape != 8 && shape != 12) return fail(c, ATT_ERR_INVALID, "potential shape must be 8 or 12");
    if (dielec != 0 && dielec != 1) return fail(c, ATT_ERR_INVALID, "dielectric must be 0 (constant) or 1 (variable)");
    if (!(plateau > 0.f) || !std::isfinite(plateau)) return fail(c, ATT_ERR_INVALID, "plateau must be positive");
    return ATT_OK;
}

}  // namespace

extern "C" {

att_status att_generate_grid(att_ctx* c, const att_protein_desc* body, const att_pair_params* table,
                             unsigned numTypes, int potShape, int dielec, float ffelec,
                             float plateau, int srcIsLigand, const att_grid_geom* geom, float* out) {
    if (!c) return ATT_ERR_INVALID;
    if (!out) return fail(c, ATT_ERR_INVALID, "null output");
    ATT_HIP(c, hipSetDevice(c->device));
    att_status st;
    if ((st = check_gen_args(c, body, table, numTypes, potShape, dielec, plateau)) != ATT_OK) return st;
    if ((st = check_geom(c, geom, numTypes + 1, "grid")) != ATT_OK) return st;
    GenInputs in;
    if ((st = gen_prepare(c, body, table, numTypes, srcIsLigand, plateau, in)) != ATT_OK) return st;
    const size_t n4 = (size_t)geom->nx * geom->ny * geom->nz * (numTypes + 1);
    Tmp tmp;
    float4* d;
    ATT_HIP(c, tmp.alloc(&d, n4));
    if ((st = gen_grid(c, in, *geom, numTypes, potShape, dielec, ffelec, d)) != ATT_OK) return st;
    ATT_HIP(c, hipMemcpyAsync(out, d, sizeof(float4) * n4, hipMemcpyDeviceToHost, c->stream));
    ATT_HIP(c, hipStreamSynchronize(c->stream));
    return ATT_OK;
}

att_status att_generate_nl(att_ctx* c, const att_protein_desc* body, const att_grid_geom* geom,
                           float neighbourdis, uint32_t* desc, uint32_t* list,
                           size_t listCapacity, uint64_t* total) {
    if (!c) return ATT_ERR_INVALID;
    if (!desc || !total) return fail(c, ATT_ERR_INVALID, "null output");
    ATT_HIP(c, hipSetDevice(c->device));
    att_status st;
    if ((st = check_body(c, body)) != ATT_OK) return st;
    if ((st = check_geom(c, geom, 1, "neighbour grid")) != ATT_OK) return st;
    if (!(neighbourdis > 0.f) || !std::isfinite(neighbourdis))
        return fail(c, ATT_ERR_INVALID, "neighbour distance must be positive");
    GenInputs in;
    if ((st = gen_prepare(c, body, nullptr, 0, 0, 1.f, in)) != ATT_OK) return st;
    Tmp tmp;
    auto alloc = [&](void** p, size_t bytes) { return tmp.alloc(reinterpret_cast<char**>(p), bytes); };
    uint2* d_desc;
    uint32_t* d_list;
    uint64_t tot = 0;
    if ((st = gen_nl(c, in, *geom, neighbourdis, alloc, &d_desc, &d_list, &tot)) != ATT_OK) return st;
    *total = tot;
    const size_t npts = (size_t)geom->nx * geom->ny * geom->nz;
    ATT_HIP(c, hipMemcpy(desc, d_desc, sizeof(uint2) * npts, hipMemcpyDeviceToHost));
    if (list && tot > 0) {
        if (listCapacity < tot) return fail(c, ATT_ERR_INVALID, "list capacity below the total");
        ATT_HIP(c, hipMemcpy(list, d_list, sizeof(uint32_t) * tot, hipMemcpyDeviceToHost));
    }
    return ATT_OK;
}

att_status att_build_grid(att_ctx* c, const att_protein_desc* body, const att_pair_params* table,
                          unsigned numTypes, int potShape, int dielec, float ffelec, float plateau,
                          int srcIsLigand, const att_grid_geom* inner, const att_grid_geom* outer,
                          float neighbourdis, att_handle* out) {
    if (!c) return ATT_ERR_INVALID;
    if (!out) return fail(c, ATT_ERR_INVALID, "null output");
    ATT_HIP(c, hipSetDevice(c->device));
    att_status st;
    if ((st = check_gen_args(c, body, table, numTypes, potShape, dielec, plateau)) != ATT_OK) return st;
    if ((st = check_geom(c, inner, numTypes + 1, "inner")) != ATT_OK) return st;
    if ((st = check_geom(c, outer, numTypes + 1, "outer")) != ATT_OK) return st;
    if (neighbourdis > 0.f && !std::isfinite(neighbourdis))
        return fail(c, ATT_ERR_INVALID, "neighbour distance must be finite");
    GenInputs in;
    if ((st = gen_prepare(c, body, table, numTypes, srcIsLigand, plateau, in)) != ATT_OK) return st;
    auto o = std::make_unique<GridObj>();
    const att_grid_geom* geoms[2] = {inner, outer};
    float4* oct[2];
    for (int k = 0; k < 2; ++k) {
        const att_grid_desc gd = grid_of(*geoms[k], numTypes + 1);
        Tmp tmp;
        float4* d;
        ATT_HIP(c, tmp.alloc(&d, (size_t)gd.nx * gd.ny * gd.nz * gd.numGrids));
        if ((st = gen_grid(c, in, *geoms[k], numTypes, potShape, dielec, ffelec, d)) != ATT_OK) return st;
        if ((st = expand_octets(c, o.get(), gd, d, &oct[k])) != ATT_OK) return st;
        ATT_HIP(c, hipStreamSynchronize(c->stream));
    }
    o->inner = make_dgrid(grid_of(*inner, numTypes + 1), oct[0]);
    o->outer = make_dgrid(grid_of(*outer, numTypes + 1), oct[1]);
    if (neighbourdis > 0.f) {
        Obj* owner = o.get();
        auto alloc = [&](void** p, size_t bytes) {
            hipError_t e = hipMalloc(p, bytes);
            if (e == hipSuccess) owner->allocs.push_back(*p);
            return e;
        };
        uint2* d_desc;
        uint32_t* d_list;
        uint64_t tot = 0;
        if ((st = gen_nl(c, in, *inner, neighbourdis, alloc, &d_desc, &d_list, &tot)) != ATT_OK) return st;
        o->nlEmpty = tot == 0;
        o->nlMaxIndex = tot ? body->n - 1 : 0;  // bound: the list holds indices of this body
        if ((st = attach_nl(c, o.get(), inner->nx, inner->ny, inner->nz, inner->ox, inner->oy,
                            inner->oz, inner->spacing, plateau, d_desc, d_list)) != ATT_OK)
            return st;
    }
    *out = store(c, std::move(o));
    return ATT_OK;
}

}  // extern "C"
