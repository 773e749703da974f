// Cell pass of the parameter gradient (grid part).
//
// The grid energy of a bead of type t at p is sum_a ipon(a,t) (rc(a,t) S_a(p)
// - ac(a,t) S6_a(p)) with S_a, S6_a the trilinear interpolants of the basis
// grids, so dE/drc(a,t) summed over every bead of type t of every pose is
// ipon(a,t) sum_cells sum_corners D_t(cell, corner) S_a(corner) / vol, where
// D_t(cell, corner) sums w_pose * (corner weight) over the beads of type t in
// that cell.  k_adjoint writes one record per bead (key = cell * T + t - 1,
// {x, y, z, w}, or {x, y, z} when the call has no pose weights); here the
// records are sorted by key (stable radix sort), the
// corner weights of each run of equal keys within a group of 64 sorted
// records summed by a segmented scan across the wave, the run sums compacted
// (integer scan of the per-group run counts), and each run sum contracted
// with its cell's basis octets.  The basis grids are thereby read about once
// per distinct (cell, type) instead of once per bead.  Every step has a fixed
// order: runs give the same bits.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "device.h"
#include "internal.h"

namespace att {

namespace {

// Summed corner weights of one (cell, type), corners in octet order
// 000 100 010 001 101 011 110 111.
struct CellW {
    float c[8];
};

// The geometry of one grid, as the adjoint's lookup uses it.
struct CellGeom {
    unsigned nx, ny;
    float lox, loy, loz, dVox;
};

__host__ __device__ inline CellGeom cell_geom(const DGrid& g) {
    return CellGeom{g.nx, g.ny, g.lox, g.loy, g.loz, g.dVox};
}

// A record's position and pose weight (1 for records without weights).
__device__ inline float4 record_pos(const float4* pos, size_t i) { return pos[i]; }
__device__ inline float4 record_pos(const float3* pos, size_t i) {
    const float3 p = pos[i];
    return make_float4(p.x, p.y, p.z, 1.f);
}

// Corner weights of a record {x, y, z, w} in cell `cell` (inner cells
// first), times its pose weight: the weights of grid_adj_one's trilinear
// interpolation (same cell, same differences).
__device__ inline CellW record_weights(const CellGeom& in, const CellGeom& out, unsigned cellsIn,
                                       unsigned cell, float4 p) {
    const bool inner = cell < cellsIn;
    const CellGeom& g = inner ? in : out;
    unsigned o = inner ? cell : cell - cellsIn;
    const unsigned ix = o % (g.nx - 1);
    o /= (g.nx - 1);
    const unsigned iy = o % (g.ny - 1), iz = o / (g.ny - 1);
    const float mnx = (float)(int)ix * g.dVox + g.lox;
    const float mny = (float)(int)iy * g.dVox + g.loy;
    const float mnz = (float)(int)iz * g.dVox + g.loz;
    const float ax = p.x - mnx, ay = p.y - mny, az = p.z - mnz;
    const float bx = (mnx + g.dVox) - p.x, by = (mny + g.dVox) - p.y, bz = (mnz + g.dVox) - p.z;
    CellW r;
    r.c[0] = bx * by * bz * p.w;
    r.c[1] = ax * by * bz * p.w;
    r.c[2] = bx * ay * bz * p.w;
    r.c[3] = bx * by * az * p.w;
    r.c[4] = ax * by * az * p.w;
    r.c[5] = bx * ay * az * p.w;
    r.c[6] = ax * ay * bz * p.w;
    r.c[7] = ax * ay * az * p.w;
    return r;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

__host__ __device__ inline size_t groups_of(size_t records) { return (records + 63) / 64; }

size_t temp_bytes(size_t records, unsigned endBit, bool weights) {
    size_t sortB = 0, scanB = 0;
    if (weights)
        (void)rocprim::radix_sort_pairs(nullptr, sortB, (unsigned*)nullptr, (unsigned*)nullptr,
                                        (float4*)nullptr, (float4*)nullptr, (unsigned)records, 0,
                                        endBit);
    else
        (void)rocprim::radix_sort_pairs(nullptr, sortB, (unsigned*)nullptr, (unsigned*)nullptr,
                                        (float3*)nullptr, (float3*)nullptr, (unsigned)records, 0,
                                        endBit);
    (void)rocprim::exclusive_scan(nullptr, scanB, (const unsigned*)nullptr, (unsigned*)nullptr, 0u,
                                  groups_of(records), rocprim::plus<unsigned>());
    return align256(sortB > scanB ? sortB : scanB);
}

}  // namespace

// Runs of equal keys among the sorted records, per group of 64 (one wave,
// lane = record): a segmented inclusive scan of the corner weights across the
// lanes (sorted keys: lanes l - d and l hold the same key only if every lane
// between does), so the last lane of each run holds the run's sum.  COUNT
// writes each group's number of runs (keys below `invalid`); otherwise the
// runs go to cellKey/cellW from the group's offset, in lane order.
template <bool COUNT, class Pos>
__global__ __launch_bounds__(256) void k_pg_runs(const unsigned* __restrict__ key,
                                                 const Pos* __restrict__ pos, size_t records,
                                                 CellGeom in, CellGeom out, unsigned T,
                                                 unsigned cellsIn, unsigned invalid,
                                                 unsigned* __restrict__ groupRuns,
                                                 const unsigned* __restrict__ groupOffset,
                                                 unsigned* __restrict__ cellKey,
                                                 CellW* __restrict__ cellW) {
    const unsigned lane = threadIdx.x & 63u;
    const size_t g = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (g >= groups_of(records)) return;  // uniform per wave
    const size_t i = g * 64 + lane;
    const unsigned k = i < records ? key[i] : 0xFFFFFFFFu;
    const bool live = k < invalid;
    CellW d{};
    if (!COUNT && live) d = record_weights(in, out, cellsIn, k / T, record_pos(pos, i));
#pragma unroll
    for (unsigned off = 1; off < 64; off <<= 1) {
        const unsigned ko = __shfl_up(k, off);
        const bool add = lane >= off && ko == k;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float v = __shfl_up(d.c[c], off);
            if (!COUNT && add) d.c[c] = d.c[c] + v;
        }
    }
    const unsigned kn = __shfl_down(k, 1);
    const bool tail = live && (lane == 63 || kn != k);
    const unsigned long long b = __ballot(tail);
    if (COUNT) {
        if (lane == 0) groupRuns[g] = (unsigned)__popcll(b);
        return;
    }
    if (tail) {
        const unsigned r = groupOffset[g] + (unsigned)__popcll(b & ((1ull << lane) - 1ull));
        cellKey[r] = k;
        cellW[r] = d;
    }
}

// One wave per workgroup: workgroup b contracts runs [b*per, ...) (per =
// ceil(runs / blocks)), lane by lane, into its LDS row (lane order within each
// atomic, program order across them), then writes the row.
__global__ __launch_bounds__(64) void k_pgrad_cells(const unsigned* __restrict__ cellKey,
                                                    const CellW* __restrict__ cellW,
                                                    const unsigned* __restrict__ groupRuns,
                                                    const unsigned* __restrict__ groupOffset,
                                                    unsigned groups,
                                                    DGrid inner, DGrid outer, DBasis bin,
                                                    DBasis bout, const att_pair_params* __restrict__ tab,
                                                    unsigned T, unsigned cellsIn, unsigned invalid,
                                                    float* __restrict__ rows) {
    extern __shared__ float s_row[];
    const unsigned npg = 3 * T * T, lane = threadIdx.x;
    for (unsigned k = lane; k < npg; k += 64) s_row[k] = 0.f;
    __syncthreads();
    const unsigned n = groupOffset[groups - 1] + groupRuns[groups - 1];
    const unsigned per = (n + gridDim.x - 1) / gridDim.x;
    const unsigned lo = min(n, blockIdx.x * per), hi = min(n, lo + per);
    for (unsigned i = lo + lane; i < hi; i += 64) {
        const unsigned key = cellKey[i];
        if (key >= invalid) continue;
        const unsigned t = key % T, cell = key / T;
        const bool in = cell < cellsIn;
        const unsigned o = in ? cell : cell - cellsIn;
        const float4* __restrict__ row = (in ? bin.oct : bout.oct) + (size_t)o * (in ? bin.nb : bout.nb) * 2;
        const float vol_inv = in ? inner.vol_inv : outer.vol_inv;
        const CellW W = cellW[i];
        for (unsigned a = 0; a < T; ++a) {
            const float4 s0 = row[4 * a], s1 = row[4 * a + 1];  // S_a at the 8 corners
            const float4 u0 = row[4 * a + 2], u1 = row[4 * a + 3];  // S6_a
            const float S = (s0.x * W.c[0] + s0.y * W.c[1] + s0.z * W.c[2] + s0.w * W.c[3] +
                             s1.x * W.c[4] + s1.y * W.c[5] + s1.z * W.c[6] + s1.w * W.c[7]) * vol_inv;
            const float S6 = (u0.x * W.c[0] + u0.y * W.c[1] + u0.z * W.c[2] + u0.w * W.c[3] +
                              u1.x * W.c[4] + u1.y * W.c[5] + u1.z * W.c[6] + u1.w * W.c[7]) * vol_inv;
            const unsigned k = a * T + t;
            const float iv = (float)tab[k].ipon;
            atomicAdd(&s_row[3 * k], iv * S);
            atomicAdd(&s_row[3 * k + 1], -iv * S6);
        }
    }
    __syncthreads();
    for (unsigned k = lane; k < npg; k += 64) rows[(size_t)blockIdx.x * npg + k] = s_row[k];
}

size_t pg_cells_bytes(size_t records, unsigned endBit, bool weights) {
    const size_t R = records, G = groups_of(records), pb = weights ? 16 : 12;
    return align256(4 * R) + align256(pb * R) + align256(4 * R) + align256(pb * R) +
           align256(4 * R) + align256(sizeof(CellW) * R) + 2 * align256(4 * G) +
           temp_bytes(records, endBit, weights);
}

PgCellBufs pg_cells_layout(void* scratch, size_t records, unsigned endBit, bool weights) {
    const size_t R = records, pb = weights ? 16 : 12;
    char* p = static_cast<char*>(scratch);
    PgCellBufs B{};
    B.weights = weights;
    B.key = reinterpret_cast<unsigned*>(p); p += align256(4 * R);
    B.pos = p; p += align256(pb * R);
    B.keySorted = reinterpret_cast<unsigned*>(p); p += align256(4 * R);
    B.posSorted = p; p += align256(pb * R);
    B.cellKey = reinterpret_cast<unsigned*>(p); p += align256(4 * R);
    B.cellW = reinterpret_cast<float*>(p); p += align256(sizeof(CellW) * R);
    B.groupRuns = reinterpret_cast<unsigned*>(p); p += align256(4 * groups_of(R));
    B.groupOffset = reinterpret_cast<unsigned*>(p); p += align256(4 * groups_of(R));
    B.temp = p;
    B.tempBytes = temp_bytes(records, endBit, weights);
    return B;
}

namespace {
template <class Pos>
hipError_t pg_cells_runs(const PairDesc& P, const PgCellBufs& B, size_t records, unsigned endBit,
                         float* rows, unsigned blocks, hipStream_t s) {
    size_t tb = B.tempBytes;
    Pos* pos = static_cast<Pos*>(B.pos);
    Pos* posSorted = static_cast<Pos*>(B.posSorted);
    hipError_t e = rocprim::radix_sort_pairs(B.temp, tb, B.key, B.keySorted, pos, posSorted,
                                             (unsigned)records, 0, endBit, s);
    if (e != hipSuccess) return e;
    const unsigned G = (unsigned)groups_of(records);
    const CellGeom in = cell_geom(P.inner), out = cell_geom(P.outer);
    const unsigned wgs = (G + 3) / 4;  // four groups (waves) per workgroup
    hipLaunchKernelGGL((k_pg_runs<true, Pos>), dim3(wgs), dim3(256), 0, s, B.keySorted,
                       (const Pos*)posSorted, records, in, out, P.par.T, P.recCellsIn, P.recInvalid,
                       B.groupRuns, nullptr, nullptr, nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    tb = B.tempBytes;
    e = rocprim::exclusive_scan(B.temp, tb, (const unsigned*)B.groupRuns, B.groupOffset, 0u, G,
                                rocprim::plus<unsigned>(), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_pg_runs<false, Pos>), dim3(wgs), dim3(256), 0, s, B.keySorted,
                       (const Pos*)posSorted, records, in, out, P.par.T, P.recCellsIn, P.recInvalid,
                       B.groupRuns, B.groupOffset, B.cellKey, reinterpret_cast<CellW*>(B.cellW));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const size_t lds = sizeof(float) * 3 * (size_t)P.par.T * P.par.T;
    hipLaunchKernelGGL(k_pgrad_cells, dim3(blocks), dim3(64), lds, s, B.cellKey,
                       reinterpret_cast<const CellW*>(B.cellW), B.groupRuns, B.groupOffset, G,
                       P.inner, P.outer, P.bin, P.bout, P.par.tab, P.par.T, P.recCellsIn,
                       P.recInvalid, rows);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_pg_cells(const PairDesc& P, const PgCellBufs& B, size_t records, unsigned endBit,
                           float* rows, unsigned blocks, hipStream_t s) {
    return B.weights ? pg_cells_runs<float4>(P, B, records, endBit, rows, blocks, s)
                     : pg_cells_runs<float3>(P, B, records, endBit, rows, blocks, s);
}

}  // namespace att
