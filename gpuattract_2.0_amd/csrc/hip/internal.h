// Launchers shared between the kernel translation units and the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"
#include "va13.h"

namespace att {

// `order`: a spatial permutation of the poses (or null); workgroups of one
// XCD then take a contiguous range of it (xcd_slot).
// `counts` (or null): per pose {listed neighbours of the beads walked, pairs
// inside the plateau}.
hipError_t launch_rigid(const DProt& rec, const DProt& lig, const DGrid& inner, const DGrid& outer,
                        const DNL& nl, const DPar& par, const float* dofs, float* res, unsigned n,
                        const unsigned* order, uint2* counts, hipStream_t s);

size_t modes_lds_bytes(const DProt& rec, const DProt& lig, const DPar& par, unsigned* tabSlots);

// `waves` waves per pose: 1, 2, 4 or 8.
hipError_t launch_modes(const DProt& rec, const DProt& lig, const DGrid& rInner,
                        const DGrid& rOuter, const DNL& nl, const DGrid& lInner,
                        const DGrid& lOuter, const DPar& par, const float* dofs, float* res,
                        unsigned n, const unsigned* order, int waves, uint2* counts, hipStream_t s);

hipError_t launch_debug_rigid(const DProt& rec, const DProt& lig, const DGrid& inner,
                              const DGrid& outer, const DNL& nl, const DPar& par,
                              const float* dof, float* xyz, float* pot, int* cell,
                              unsigned* npairs, hipStream_t s);

hipError_t launch_debug_modes(const DProt& rec, const DProt& lig, const DGrid& rInner,
                              const DGrid& rOuter, const DNL& nl, const DGrid& lInner,
                              const DGrid& lOuter, const DPar& par, const float* dof, float* xyz,
                              float* pot, int* cell, unsigned* npairs, float* rpot, float* rxyz,
                              float* rdefo, hipStream_t s);

hipError_t launch_debug_rotmats(const float* ang, unsigned n, float* R, hipStream_t s);
hipError_t launch_debug_sincosf(const float* x, unsigned n, float* sn, float* cs, hipStream_t s);
hipError_t launch_debug_plateau_ratio(float P2, unsigned n, float* ratio, hipStream_t s);

hipError_t launch_copy(void* dst, const void* src, size_t n16, hipStream_t s);

hipError_t launch_octets(const float4* grid, float4* oct, unsigned nx, unsigned ny, unsigned nz,
                         unsigned ngrids, hipStream_t s);

size_t adjoint_lds_bytes(const DProt& rec, const DProt& lig, const DPar& par, bool pgrad,
                         int waves, unsigned* tabSlots);
// Static LDS of the adjoint kernels with `waves` waves per pose, and how many
// of their workgroups one CU holds with `lds` bytes of dynamic LDS.
size_t adjoint_static_lds(int waves);
unsigned adjoint_blocks_per_cu(unsigned mm, int waves, bool batch, size_t lds, int pg = 0);

// How k_adjoint takes the grid part of the parameter gradient: not at all (no
// parameter gradient), from the basis grids at every bead (kPgFused), or as
// one record per bead {grid cell and bead type; position and weight} that the
// cell pass (pgcells.hip) contracts with the basis grids once per distinct
// (cell, type) after sorting (kPgRecords; PairDesc.recKey set).
constexpr int kPgNone = 0, kPgFused = 1, kPgRecords = 2;

// Scratch of the cell pass for `records` records: the adjoint's records
// (keys, then positions) at the front, sort and reduction buffers after.
struct PgCellBufs {
    unsigned* key;       // records as written by k_adjoint
    void* pos;           // float4 {x, y, z, w} per record, or float3 {x, y, z} without weights
    unsigned* keySorted;
    void* posSorted;
    bool weights;
    unsigned* cellKey;   // runs of equal keys (per group of 64 sorted records), in order
    float* cellW;        // their summed corner weights, 8 per run
    unsigned* groupRuns; // runs per group of 64 records
    unsigned* groupOffset;  // and their exclusive prefix sums
    void* temp;
    size_t tempBytes;
};
size_t pg_cells_bytes(size_t records, unsigned endBit, bool weights);
PgCellBufs pg_cells_layout(void* scratch, size_t records, unsigned endBit, bool weights);
// Cell pass over the records of P's k_adjoint launch: sort by key, sum the
// corner weights of each run of equal keys (fixed order), then `blocks`
// one-wave workgroups contract every run with its cell's basis values and
// write one parameter row each to rows[0 .. blocks).
hipError_t launch_pg_cells(const PairDesc& P, const PgCellBufs& B, size_t records, unsigned endBit,
                           float* rows, unsigned blocks, hipStream_t s);

// One pair: persistent grid of `blocks` workgroups of 64*waves threads.
hipError_t launch_adjoint(const PairDesc& P, bool modes, float* pgPartial, unsigned blocks,
                          int waves, size_t lds, hipStream_t s);
// Many pairs of one mode-capacity class: resets the kBatchQueues tile
// counters, then the persistent batch kernel over the tile queues
// [queueStart[x], queueStart[x+1]) (graph-capturable: no memset, no host sync).
hipError_t launch_adjoint_batch(unsigned mm, int waves, const PairDesc* pairs, const BatchTile* tiles,
                                const unsigned* queueStart, unsigned ntiles, unsigned* counters,
                                unsigned blocks, size_t lds, hipStream_t s);

// One device-resident minimization run (minimize.hip): per-pose VA13 states,
// the next round's request list, count and points, and the results.
struct MinRun {
    double* st;        // npose * va_elements(dim), element-major
    VaState* sc;       // npose
    unsigned npose;
    int dim, Mr, Ml, ivmax;
    bool modes;
    unsigned ds, os;   // DOF and result record widths (26/27 or 6/7)
    unsigned* list;    // next round: pose of each request slot
    unsigned* count;   // next round: number of requests
    float* req;        // next round: requested points as DOF records
    float* dofsOut;    // npose * ds: last evaluated DOFs
    float* eLast;      // npose: their energy
    float* eBest;      // npose or nullptr: lowest energy seen
    int* nfun;         // npose or nullptr: line-search evaluations
};
hipError_t launch_min_start(const MinRun& R, const float* dofs, hipStream_t s);
hipError_t launch_min_step(const MinRun& R, const unsigned* listIn, const unsigned* countIn,
                           const float* reqIn, const float* res, hipStream_t s);

// Zero-fills n floats with a kernel (graph-capturable, unlike hipMemsetAsync).
hipError_t launch_fill_zero(float* p, unsigned n, hipStream_t s);

// Sums rows x cols per-workgroup parameter rows into pg in a fixed order:
// kPgradGroups groups of consecutive rows in double (into `groups`, scratch of
// pgrad_group_bytes(cols)), then the group sums in order.
constexpr unsigned kPgradGroups = 32;
inline size_t pgrad_group_bytes(unsigned cols) { return sizeof(double) * kPgradGroups * cols; }
hipError_t launch_pgrad_reduce(const float* partial, unsigned rows, unsigned cols, double* groups,
                               float* pg, hipStream_t s);

hipError_t launch_basis_octets(const float* b, float4* oct, unsigned nx, unsigned ny, unsigned nz,
                               unsigned nb, hipStream_t s);

// Spatial pose order (order.hip): scratch bytes for n poses, and the launch
// writing the permutation to `order`.
size_t pose_order_scratch(unsigned n);
hipError_t launch_pose_order(const float* dofs, unsigned stride, unsigned n, void* scratch,
                             size_t scratchBytes, unsigned* order, hipStream_t s);

// Grid generation (gridgen.hip).  Geometry in double as the points are
// computed: p = origin + index * spacing.
struct GenGeom {
    unsigned nx, ny, nz;
    double ox, oy, oz, sp;
};
// Pair parameters seen from (source type a, probe type t), index a*(T+1)+t.
struct GenPair {
    double rc, ac, ipon, rmin2, emin;
};
struct GenSrc {
    const float4* sorted;       // beads with type 1..T, stably sorted by type {x,y,z,0}
    const unsigned* typeStart;  // T+2: type a occupies [typeStart[a], typeStart[a+1])
    const float4* act;          // the same beads in index order
    const unsigned* actType;
    unsigned nAct;
    const float4* chg;          // those with charge != 0, index order {x,y,z,q}
    unsigned nChg;
};
hipError_t launch_gen_grid(const GenGeom& g, const GenSrc& s, const GenPair* tab, unsigned T,
                           int shape, bool linear, double P2, double sqrtP2, int dielec, double fe,
                           float4* out, hipStream_t st);
size_t gen_scan_scratch(size_t npts);
// Counts per point (cnt), then desc = {count, exclusive offset}; the total
// lands in scratch[ceil(npts/1024)].
hipError_t launch_gen_nl_count(const GenGeom& g, const float4* beads, const unsigned* idx,
                               unsigned nb, double nd2, float cullR2, uint32_t* cnt, uint2* desc,
                               unsigned long long* scratch, hipStream_t st);
hipError_t launch_gen_nl_fill(const GenGeom& g, const float4* beads, const unsigned* idx,
                              unsigned nb, double nd2, float cullR2, const uint2* desc,
                              uint32_t* list, hipStream_t st);

}  // namespace att
