// Scoring kernels for gfx950 and their launchers.
//
//  k_rigid  one 64-lane wavefront per pose, four poses per 256-thread workgroup.
//           Lanes stride over ligand beads (transform, grid interpolation),
//           then walk the neighbour-list pairs of each chunk of 64 beads one
//           pair per lane; per-lane partial sums of E, F and the torque
//           tensor; a butterfly reduction and a one-lane epilogue (force
//           clamp, torque matrix) write the 28-byte result.  Replaces
//           d_DOF2Pos, d_potForce, d_NLPotForce, d_reduce and h_finalReduce
//           (model/*.cu, model/reduction.h:159-207) with one launch and no
//           intermediate buffers.
//  k_modes  one 256-thread workgroup per pose.  Phase A moves the deformed
//           receptor into the ligand frame, scores it on the ligand grid and
//           keeps its deformed coordinates in LDS; phase B scores the ligand on
//           the receptor grid and runs the neighbour-list pairs, one pair per
//           lane, against those LDS coordinates, folding the receptor reaction
//           straight into the receptor mode gradients.  Semantics of
//           service/CPUEnergyService6DModes.tpp:121-327.
#include <algorithm>

#include "device.h"
#include "internal.h"

namespace att {

// Per-wave slot arrays of k_rigid's pair-parallel neighbour-list phase: bead
// position and charge, bead | type << 16, first list entry, inclusive pair
// prefix; 64 slots per wave, behind the pair table (16-byte aligned).
constexpr unsigned kRigidSlotBytes = 64 * (16 + 3 * sizeof(unsigned));
__host__ __device__ inline size_t rigid_slot_offset(unsigned tabSlots) {
    return (sizeof(att_pair_params) * (size_t)tabSlots + 15) & ~(size_t)15;
}

__global__ __launch_bounds__(256) void k_rigid(DProt rec, DProt lig, DGrid inner, DGrid outer,
                                               DNL nl, DPar par, const float* __restrict__ dofs,
                                               float* __restrict__ res, unsigned n,
                                               unsigned tabSlots,
                                               const unsigned* __restrict__ order) {
    extern __shared__ __align__(16) att_pair_params s_tab[];
    const att_pair_params* tab = stage_table(par, s_tab);
    __syncthreads();
    const unsigned lane = threadIdx.x & (kWave - 1);
    const unsigned wave = threadIdx.x / kWave;
    char* slots = reinterpret_cast<char*>(s_tab) + rigid_slot_offset(tabSlots) + wave * kRigidSlotBytes;
    float4* s_sp = reinterpret_cast<float4*>(slots);
    unsigned* s_smeta = reinterpret_cast<unsigned*>(s_sp + 64);
    unsigned* s_sstart = s_smeta + 64;
    unsigned* s_send = s_sstart + 64;
    const unsigned wg = order ? xcd_slot(blockIdx.x, gridDim.x) : blockIdx.x;
    const unsigned slot = wg * (blockDim.x / kWave) + wave;
    if (slot >= n) return;  // uniform per wave; no workgroup barrier follows
    const unsigned pose = order ? order[slot] : slot;
    const float* d = dofs + 6 * (size_t)pose;
    const float tx = d[0], ty = d[1], tz = d[2];
    float t[6];
    pose_trig(d[3], d[4], d[5], lane, t);
    float R[9];
    rotmat_from_trig(t[0], t[1], t[2], t[3], t[4], t[5], R);
    float acc[13];
#pragma unroll
    for (int k = 0; k < 13; ++k) acc[k] = 0.f;
    auto add_bead = [&](const float4& x, float fx, float fy, float fz, float e) {
        acc[0] += e;
        acc[1] += fx; acc[2] += fy; acc[3] += fz;
        acc[4] += x.x * fx; acc[5] += x.y * fx; acc[6] += x.z * fx;
        acc[7] += x.x * fy; acc[8] += x.y * fy; acc[9] += x.z * fy;
        acc[10] += x.x * fz; acc[11] += x.y * fz; acc[12] += x.z * fz;
    };
    const float4* __restrict__ rpos = rec.pos;
    // Chunks of 64 ligand beads, one per lane: grid term, neighbour-list cell.
    // The beads of the chunk that have a list become the chunk's slots (in
    // lane order); their pairs are flattened (prefix sum of the list lengths)
    // and walked 64 at a time, one pair per lane, each evaluated as nl_bead
    // evaluates it.  A bead's pair terms are summed over its run of lanes
    // (segmented sums, fixed order) and added by the run's first lane (the
    // pose sums are linear in the bead force, so the grid and the
    // neighbour-list parts of a bead are added separately).
    for (unsigned c0 = 0; c0 < lig.n; c0 += kWave) {
        const unsigned i = c0 + lane;
        unsigned cnt = 0, first = 0, lt = 0;
        float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
        V3 p{0.f, 0.f, 0.f};
        if (i < lig.n) {
            x = lig.pos[i];
            p = matvec(R, x.x, x.y, x.z);
            p.x = p.x + tx; p.y = p.y + ty; p.z = p.z + tz;
            const float4 pot = pot_bead(inner, outer, p.x, p.y, p.z, lig.mtype[i], x.w, nullptr);
            add_bead(x, pot.x, pot.y, pot.z, pot.w);
            lt = lig.type[i];
            // NLGrid bounds and getIndex as nl_bead
            if (lt != 0 && (p.x >= nl.lox && p.x <= nl.hix) && (p.y >= nl.loy && p.y <= nl.hiy) &&
                (p.z >= nl.loz && p.z <= nl.hiz)) {
                const int ix = (int)roundf((p.x - nl.lox) * nl.dVox_inv);
                const int iy = (int)roundf((p.y - nl.loy) * nl.dVox_inv);
                const int iz = (int)roundf((p.z - nl.loz) * nl.dVox_inv);
                const uint2 dsc = nl.desc[nl.nx * (iz * nl.ny + iy) + ix];
                cnt = dsc.x;
                first = dsc.y;
            }
        }
        const unsigned long long hasList = __ballot(cnt != 0u);
        const unsigned ns = (unsigned)__popcll(hasList);
        if (ns == 0) continue;
        // inclusive prefix of the list lengths over the lanes
        unsigned incl = cnt;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const unsigned y = __shfl_up(incl, o);
            if (lane >= (unsigned)o) incl += y;
        }
        const unsigned NP = __shfl(incl, kWave - 1);
        if (cnt) {
            const unsigned k = (unsigned)__popcll(hasList & ((1ull << lane) - 1ull));
            s_sp[k] = make_float4(p.x, p.y, p.z, x.w);
            s_smeta[k] = i | lt << 16;
            s_sstart[k] = first;
            s_send[k] = incl;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        unsigned b0 = 0;
        for (unsigned B = 0; B < NP; B += kWave) {
            const unsigned pr = B + lane;
            const bool valid = pr < NP;
            // the lane's slot: b0 + the number of slot ends at or below its pair
            const unsigned e = b0 + lane < ns ? s_send[b0 + lane] : 0xFFFFFFFFu;
            unsigned long long starts;
            const unsigned sl = min(b0 + ends_up_to(e, B, lane, starts), ns - 1u);
            b0 += (unsigned)__popcll(__ballot(e <= B + kWave));
            float fx = 0.f, fy = 0.f, fz = 0.f, en = 0.f;
            if (valid) {
                const float4 bp = s_sp[sl];
                const unsigned lt1 = (s_smeta[sl] >> 16) - 1u;
                const unsigned begin = sl ? s_send[sl - 1] : 0u;
                const unsigned nIdx = nl.list[s_sstart[sl] + (pr - begin)];
                const float4 rx = rpos[nIdx];
                const unsigned rt = rec.type[nIdx];
                float dx = bp.x - rx.x, dy = bp.y - rx.y, dz = bp.z - rx.z;
                const float dr2 = dx * dx + dy * dy + dz * dz;
                // outOfPlateau (NaN skipped); a receptor bead without a valid type
                // is skipped as in nl_bead
                if (dr2 <= nl.P2 && rt - 1u < par.T) {
                    const float ratio = nl.ratio[(int)(dr2 * 10000.f)];
                    const float dr2_inv = (float)(1.0 / (double)dr2);
                    dx *= dr2_inv; dy *= dr2_inv; dz *= dr2_inv;
                    const att_pair_params pp = tab[par.T * (rt - 1) + lt1];
                    float vx, vy, vz, ve;
                    lj(dr2, dr2_inv, dx, dy, dz, pp, par.shape, vx, vy, vz, ve);
                    fx += vx; fy += vy; fz += vz; en += ve;
                    const float rdx = ratio * dx, rdy = ratio * dy, rdz = ratio * dz;
                    lj(nl.P2, nl.P2_inv, rdx, rdy, rdz, pp, par.shape, vx, vy, vz, ve);
                    fx -= vx; fy -= vy; fz -= vz; en -= ve;
                    if (bp.w != 0.f) {
                        const float qlr = bp.w * rx.w * par.ffelec;
                        if (fabsf(qlr) > 0.001f) {
                            coulomb(dr2_inv, dx, dy, dz, qlr, par.dielec, vx, vy, vz, ve);
                            fx += vx; fy += vy; fz += vz; en += ve;
                            coulomb(nl.P2_inv, rdx, rdy, rdz, qlr, par.dielec, vx, vy, vz, ve);
                            fx -= vx; fy -= vy; fz -= vz; en -= ve;
                        }
                    }
                }
            }
            seg_run_sums(fx, fy, fz, en, valid ? sl : 0xFFFFFFFFu, lane);
            if (valid && (lane == 0 || ((starts >> lane) & 1ull)))
                add_bead(lig.pos[s_smeta[sl] & 0xFFFFu], fx, fy, fz, en);
        }
        // every lane has read the slots before the next chunk writes them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
#pragma unroll
    for (int k = 0; k < 13; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0) {
        float Fx = acc[1], Fy = acc[2], Fz = acc[3];
        force_clamp(Fx, Fy, Fz);
        float a3[3];
        torque_reduce(t[0], t[1], t[2], t[3], t[4], t[5], acc + 4, a3);
        float* o = res + 7 * (size_t)pose;
        o[0] = acc[0]; o[1] = Fx; o[2] = Fy; o[3] = Fz;
        o[4] = a3[0]; o[5] = a3[1]; o[6] = a3[2];
    }
}

constexpr int kModesVals = 13 + 2 * ATT_MODES_MAX;

// Beads per round of k_modes' pair-parallel neighbour-list phase.
constexpr unsigned kModesSlots = 128;

// Byte offset of that phase's slot arrays in k_modes' dynamic LDS: after the
// pair table, the deformed receptor (xyz floats, u16 types) and the queue.
__host__ __device__ inline size_t modes_slot_offset(const DProt& rec, const DProt& lig,
                                                    unsigned tabSlots) {
    const size_t q = sizeof(att_pair_params) * (size_t)tabSlots + sizeof(float) * 3 * (size_t)rec.n +
                     sizeof(unsigned short) * (((size_t)rec.n + 1) & ~(size_t)1);
    return (q + sizeof(unsigned) * (size_t)pow2_ceil(lig.n) + 15) & ~(size_t)15;
}

template <int MM>
__global__ __launch_bounds__(256) void k_modes(DProt rec, DProt lig, DGrid rInner, DGrid rOuter,
                                               DNL nl, DGrid lInner, DGrid lOuter, DPar par,
                                               const float* __restrict__ dofs,
                                               float* __restrict__ res, unsigned n,
                                               unsigned tabSlots,
                                               const unsigned* __restrict__ order) {
    extern __shared__ __align__(16) att_pair_params s_dyn[];
    __shared__ float s_red[4][kModesVals];
    const att_pair_params* tab = stage_table(par, s_dyn);
    float* s_rx = reinterpret_cast<float*>(s_dyn + tabSlots);
    float* s_ry = s_rx + rec.n;
    float* s_rz = s_ry + rec.n;
    unsigned short* s_rt = reinterpret_cast<unsigned short*>(s_rz + rec.n);
    unsigned* s_queue = reinterpret_cast<unsigned*>(s_rt + ((rec.n + 1) & ~1u));
    char* dynb = reinterpret_cast<char*>(s_dyn);
    const size_t so = modes_slot_offset(rec, lig, tabSlots);
    float4* s_sp = reinterpret_cast<float4*>(dynb + so);                 // bead position, charge
    unsigned* s_smeta = reinterpret_cast<unsigned*>(s_sp + kModesSlots);  // bead | type << 16
    unsigned* s_sstart = s_smeta + kModesSlots;                           // first list entry
    unsigned* s_send = s_sstart + kModesSlots;                            // inclusive pair prefix
    unsigned* s_bits = s_send + kModesSlots;  // one bit per ligand bead: has a neighbour list
    const unsigned nWords = (lig.n + 31) / 32;
    __shared__ unsigned s_scan[4];
    if (blockIdx.x >= n) return;  // uniform per workgroup
    const unsigned pose = order ? order[xcd_slot(blockIdx.x, gridDim.x)] : blockIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const float* d = dofs + 26 * (size_t)pose;
    const float tx = d[0], ty = d[1], tz = d[2];
    constexpr int MA = MM > 0 ? MM : 1;
    float ampR[MA], ampL[MA];
#pragma unroll
    for (int m = 0; m < MA; ++m) { ampR[m] = d[6 + m]; ampL[m] = d[16 + m]; }
    float t[6];
    pose_trig(d[3], d[4], d[5], lane, t);
    float R[9], RT[9], Rinv[9];
    rotmat_from_trig(t[0], t[1], t[2], t[3], t[4], t[5], R);
    RT[0] = R[0]; RT[1] = R[3]; RT[2] = R[6];
    RT[3] = R[1]; RT[4] = R[4]; RT[5] = R[7];
    RT[6] = R[2]; RT[7] = R[5]; RT[8] = R[8];
    // invertDOF (model/transform_modes.h:26-41): rotation R(-ang) (cos even,
    // sin odd), translation R^T (-pos).
    rotmat_from_trig(t[0], t[1], -t[2], -t[3], t[4], -t[5], Rinv);
    const V3 pinv = matvec(RT, -tx, -ty, -tz);

    float acc[13];
    float mR[MA], mL[MA];
#pragma unroll
    for (int k = 0; k < 13; ++k) acc[k] = 0.f;
#pragma unroll
    for (int m = 0; m < MA; ++m) { mR[m] = 0.f; mL[m] = 0.f; }

    for (unsigned w = threadIdx.x; w < nWords; w += blockDim.x) s_bits[w] = 0u;
    // Phase A: receptor beads.
    for (unsigned j = threadIdx.x; j < rec.n; j += blockDim.x) {
        const float4 x = rec.pos[j];
        BeadModes<MM> md;
        load_modes(rec, j, md);
        const V3 df = deform(md, rec.M, x.x, x.y, x.z, ampR);
        s_rx[j] = df.x; s_ry[j] = df.y; s_rz[j] = df.z;
        s_rt[j] = (unsigned short)rec.type[j];
        V3 p = matvec(Rinv, df.x, df.y, df.z);
        p.x = p.x + pinv.x; p.y = p.y + pinv.y; p.z = p.z + pinv.z;
        const float4 pot = pot_bead(lInner, lOuter, p.x, p.y, p.z, rec.mtype[j], x.w, nullptr);
        const V3 f = matvec(R, pot.x, pot.y, pot.z);  // rotate_forces (transform_modes.h:141-161)
        load_modes(rec, j, md);  // reloaded (cache hit) rather than held across the lookup
        mode_dots(md, rec.M, f.x, f.y, f.z, mR);
    }
    __syncthreads();

    // Phase B: ligand beads.
    const float4* __restrict__ rpos = rec.pos;
    const float4* __restrict__ rmodesT = rec.modesT;
    const unsigned Mr = rec.M, rowT = rec.rowT;
    auto recPos = [=](unsigned j) { return make_float3(s_rx[j], s_ry[j], s_rz[j]); };
    auto recType = [=](unsigned j) { return (unsigned)s_rt[j]; };
    auto reaction = [&](unsigned j, float fx, float fy, float fz) {
        const float4* row = rmodesT + (size_t)j * rowT;
        constexpr unsigned Q = (3 * MA + 3) / 4;
        float mt[4 * Q];
#pragma unroll
        for (unsigned q = 0; q < Q; ++q) {
            const float4 v = q < rowT ? row[q] : make_float4(0.f, 0.f, 0.f, 0.f);
            mt[4 * q] = v.x; mt[4 * q + 1] = v.y; mt[4 * q + 2] = v.z; mt[4 * q + 3] = v.w;
        }
#pragma unroll
        for (unsigned m = 0; m < (unsigned)MM; ++m)
            if (m < Mr) mR[m] -= fx * mt[3 * m] + fy * mt[3 * m + 1] + fz * mt[3 * m + 2];
    };
    // Per-bead sums are linear in the bead force, so the grid part and the
    // neighbour-list part of each ligand bead are added to them separately.
    auto add_bead = [&](const float4& x, unsigned i, float fx, float fy, float fz, float e) {
        acc[0] += e;
        acc[1] += fx; acc[2] += fy; acc[3] += fz;
        acc[4] += x.x * fx; acc[5] += x.y * fx; acc[6] += x.z * fx;
        acc[7] += x.x * fy; acc[8] += x.y * fy; acc[9] += x.z * fy;
        acc[10] += x.x * fz; acc[11] += x.y * fz; acc[12] += x.z * fz;
        // reduceModeForce, ligand (model/reduction_modes.h:28-57): R^T f . mode.
        const V3 fl = matvec(RT, fx, fy, fz);
        BeadModes<MM> md;
        load_modes(lig, i, md);
        mode_dots(md, lig.M, fl.x, fl.y, fl.z, mL);
    };
    auto place = [&](unsigned i, float4& x) {
        x = lig.pos[i];
        BeadModes<MM> md;
        load_modes(lig, i, md);
        const V3 df = deform(md, lig.M, x.x, x.y, x.z, ampL);
        V3 p = matvec(R, df.x, df.y, df.z);
        p.x = p.x + tx; p.y = p.y + ty; p.z = p.z + tz;
        return p;
    };
    // Phase B1: ligand beads on the receptor grid; beads with a neighbour list
    // are marked.
    for (unsigned i = threadIdx.x; i < lig.n; i += blockDim.x) {
        float4 x;
        const V3 p = place(i, x);
        const float4 pot = pot_bead(rInner, rOuter, p.x, p.y, p.z, lig.mtype[i], x.w, nullptr);
        add_bead(x, i, pot.x, pot.y, pot.z, pot.w);
        const unsigned cnt = nl_count(nl, p.x, p.y, p.z, lig.type[i]);
        if (cnt) atomicOr(&s_bits[i >> 5], 1u << (i & 31u));
    }
    __syncthreads();
    // Queue: the marked beads in bead order (each thread a contiguous range of
    // mark words; block scan of their counts).
    unsigned qn;
    {
        const unsigned per = (nWords + blockDim.x - 1) / blockDim.x;
        const unsigned lo = threadIdx.x * per, hi = min(nWords, lo + per);
        unsigned c = 0;
        for (unsigned w = lo; w < hi; ++w) c += (unsigned)__popc(s_bits[w]);
        unsigned o = block_scan(c, s_scan, qn);
        for (unsigned w = lo; w < hi; ++w)
            for (unsigned m = s_bits[w]; m; m &= m - 1u) s_queue[o++] = 32u * w + (unsigned)__ffs(m) - 1u;
    }
    __syncthreads();
    // Phase B2: the neighbour lists of the queued beads, pair-parallel: rounds
    // of up to kModesSlots beads (in bead order) flatten their
    // pairs (prefix sum of the list lengths) into one contiguous range per
    // wave, walked in chunks of 64 pairs, one pair per lane.  Each pair is
    // evaluated as nl_bead evaluates it (pair force minus the plateau force
    // along the ratio-scaled direction, Coulomb likewise, the reaction into
    // the receptor mode gradients); the terms of one bead are summed over its
    // run of lanes (segmented sums, fixed order) and added to the pose sums by
    // the run's first lane, once per wave that holds part of the bead (the
    // pose sums are linear in the bead force).
    const unsigned nw = blockDim.x / kWave;
    for (unsigned base = 0; base < qn; base += kModesSlots) {
        const unsigned ns = min(kModesSlots, qn - base);
        unsigned cnt = 0;
        if (threadIdx.x < ns) {
            const unsigned i = s_queue[base + threadIdx.x];
            float4 x;
            const V3 p = place(i, x);
            // NLGrid::getIndex as nl_bead (the bead was queued: it is in the box)
            const int ix = (int)roundf((p.x - nl.lox) * nl.dVox_inv);
            const int iy = (int)roundf((p.y - nl.loy) * nl.dVox_inv);
            const int iz = (int)roundf((p.z - nl.loz) * nl.dVox_inv);
            const uint2 dsc = nl.desc[nl.nx * (iz * nl.ny + iy) + ix];
            cnt = dsc.x;
            s_sp[threadIdx.x] = make_float4(p.x, p.y, p.z, x.w);
            s_smeta[threadIdx.x] = i | lig.type[i] << 16;
            s_sstart[threadIdx.x] = dsc.y;
        }
        unsigned NP;
        const unsigned ex = block_scan(cnt, s_scan, NP);
        if (threadIdx.x < ns) s_send[threadIdx.x] = ex + cnt;
        __syncthreads();
        const unsigned per = (NP + nw - 1) / nw;
        const unsigned rs = min(NP, (unsigned)wave * per), re = min(NP, rs + per);
        unsigned b0 = 0;
        {
            unsigned lo = 0, hi = ns;  // first slot whose end exceeds rs
            while (lo < hi) {
                const unsigned mid = (lo + hi) >> 1;
                if (s_send[mid] > rs) hi = mid; else lo = mid + 1;
            }
            b0 = lo;
        }
        for (unsigned B = rs; B < re; B += kWave) {
            const unsigned pr = B + lane;
            const bool valid = pr < re;
            // the lane's slot: b0 + the number of slot ends at or below its pair
            const unsigned e = b0 + lane < ns ? s_send[b0 + lane] : 0xFFFFFFFFu;
            unsigned long long starts;
            const unsigned sl = min(b0 + ends_up_to(e, B, lane, starts), ns - 1u);
            b0 += (unsigned)__popcll(__ballot(e <= B + kWave));
            float fx = 0.f, fy = 0.f, fz = 0.f, en = 0.f;
            if (valid) {
                const float4 bp = s_sp[sl];
                const unsigned lt1 = (s_smeta[sl] >> 16) - 1u;
                const unsigned begin = sl ? s_send[sl - 1] : 0u;
                const unsigned nIdx = nl.list[s_sstart[sl] + (pr - begin)];
                const float3 rp = recPos(nIdx);
                const unsigned rt = recType(nIdx);
                float dx = bp.x - rp.x, dy = bp.y - rp.y, dz = bp.z - rp.z;
                const float dr2 = dx * dx + dy * dy + dz * dz;
                // outOfPlateau (NaN skipped); a receptor bead without a valid type
                // is skipped as in nl_bead
                if (dr2 <= nl.P2 && rt - 1u < par.T) {
                    const float ratio = nl.ratio[(int)(dr2 * 10000.f)];
                    const float dr2_inv = (float)(1.0 / (double)dr2);
                    dx *= dr2_inv; dy *= dr2_inv; dz *= dr2_inv;
                    const att_pair_params pp = tab[par.T * (rt - 1) + lt1];
                    float vx, vy, vz, ve;
                    float rfx = 0.f, rfy = 0.f, rfz = 0.f;
                    lj(dr2, dr2_inv, dx, dy, dz, pp, par.shape, vx, vy, vz, ve);
                    fx += vx; fy += vy; fz += vz; en += ve;
                    rfx -= vx; rfy -= vy; rfz -= vz;
                    const float rdx = ratio * dx, rdy = ratio * dy, rdz = ratio * dz;
                    lj(nl.P2, nl.P2_inv, rdx, rdy, rdz, pp, par.shape, vx, vy, vz, ve);
                    fx -= vx; fy -= vy; fz -= vz; en -= ve;
                    rfx += vx; rfy += vy; rfz += vz;
                    if (bp.w != 0.f) {
                        const float qlr = bp.w * rpos[nIdx].w * par.ffelec;
                        if (fabsf(qlr) > 0.001f) {
                            coulomb(dr2_inv, dx, dy, dz, qlr, par.dielec, vx, vy, vz, ve);
                            fx += vx; fy += vy; fz += vz; en += ve;
                            rfx -= vx; rfy -= vy; rfz -= vz;
                            coulomb(nl.P2_inv, rdx, rdy, rdz, qlr, par.dielec, vx, vy, vz, ve);
                            fx -= vx; fy -= vy; fz -= vz; en -= ve;
                            rfx += vx; rfy += vy; rfz += vz;
                        }
                    }
                    reaction(nIdx, rfx, rfy, rfz);
                }
            }
            seg_run_sums(fx, fy, fz, en, valid ? sl : 0xFFFFFFFFu, lane);
            if (valid && (lane == 0 || ((starts >> lane) & 1ull))) {
                const unsigned i = s_smeta[sl] & 0xFFFFu;
                add_bead(lig.pos[i], i, fx, fy, fz, en);
            }
        }
        __syncthreads();  // slot arrays are reused by the next round
    }

    // Workgroup reduction: wave butterflies, then the four waves in order.
#pragma unroll
    for (int k = 0; k < 13; ++k) acc[k] = wave_sum(acc[k]);
#pragma unroll
    for (int m = 0; m < MM; ++m) { mR[m] = wave_sum(mR[m]); mL[m] = wave_sum(mL[m]); }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 13; ++k) s_red[wave][k] = acc[k];
#pragma unroll
        for (int m = 0; m < ATT_MODES_MAX; ++m) {
            s_red[wave][13 + m] = m < MM ? mR[m < MM ? m : 0] : 0.f;
            s_red[wave][13 + ATT_MODES_MAX + m] = m < MM ? mL[m < MM ? m : 0] : 0.f;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float v[kModesVals];
        const int nw = blockDim.x / kWave;
#pragma unroll
        for (int k = 0; k < kModesVals; ++k) {
            float s = s_red[0][k];
            for (int w = 1; w < nw; ++w) s += s_red[w][k];
            v[k] = s;
        }
        float Fx = v[1], Fy = v[2], Fz = v[3];
        force_clamp(Fx, Fy, Fz);
        float a3[3];
        torque_reduce(t[0], t[1], t[2], t[3], t[4], t[5], v + 4, a3);
        float* o = res + 27 * (size_t)pose;
        o[0] = v[0]; o[1] = Fx; o[2] = Fy; o[3] = Fz;
        o[4] = a3[0]; o[5] = a3[1]; o[6] = a3[2];
        for (unsigned m = 0; m < ATT_MODES_MAX; ++m) {
            o[7 + m] = m < rec.M ? correct_mode(v[13 + m], rec.modeForce[m]) : 0.f;
            o[17 + m] = m < lig.M ? correct_mode(v[13 + ATT_MODES_MAX + m], lig.modeForce[m]) : 0.f;
        }
    }
}

// ---------------------------------------------------------------- diagnostics

__global__ void k_debug_rigid(DProt rec, DProt lig, DGrid inner, DGrid outer, DNL nl, DPar par,
                              const float* dof, float* xyz, float* pot, int* cell,
                              unsigned* npairs) {
    const int lane = threadIdx.x;
    float t[6];
    pose_trig(dof[3], dof[4], dof[5], lane, t);
    float R[9];
    rotmat_from_trig(t[0], t[1], t[2], t[3], t[4], t[5], R);
    const unsigned n = lig.n;
    for (unsigned i = lane; i < n; i += kWave) {
        const float4 x = lig.pos[i];
        V3 p = matvec(R, x.x, x.y, x.z);
        p.x = p.x + dof[0]; p.y = p.y + dof[1]; p.z = p.z + dof[2];
        int* c = cell + 9 * i;
        for (int k = 0; k < 9; ++k) c[k] = -1;
        float4 pp = pot_bead(inner, outer, p.x, p.y, p.z, lig.mtype[i], x.w, c);
        float4 a;
        unsigned np = 0;
        if (nl_bead(nl, par, par.tab, RecPosG{rec.pos}, RecTypeG{rec.type}, RecChargeG{rec.pos},
                    NoReaction{}, p.x, p.y, p.z, lig.type[i], x.w, a, c + 6, &np)) {
            pp.x = pp.x + a.x; pp.y = pp.y + a.y; pp.z = pp.z + a.z; pp.w = pp.w + a.w;
        }
        xyz[i] = p.x; xyz[n + i] = p.y; xyz[2 * n + i] = p.z;
        pot[i] = pp.x; pot[n + i] = pp.y; pot[2 * n + i] = pp.z; pot[3 * n + i] = pp.w;
        npairs[i] = np;
    }
}

// One wave; the neighbour-list phase runs on lane 0 over the ligand beads in
// order so that receptor reaction forces accumulate in the CPU's order.
__global__ void k_debug_modes(DProt rec, DProt lig, DGrid rInner, DGrid rOuter, DNL nl,
                              DGrid lInner, DGrid lOuter, DPar par, const float* d, float* xyz,
                              float* pot, int* cell, unsigned* npairs, float* rpot, float* rxyz,
                              float* rdefo) {
    const int lane = threadIdx.x;
    float t[6];
    pose_trig(d[3], d[4], d[5], lane, t);
    float R[9], RT[9], Rinv[9];
    rotmat_from_trig(t[0], t[1], t[2], t[3], t[4], t[5], R);
    RT[0] = R[0]; RT[1] = R[3]; RT[2] = R[6];
    RT[3] = R[1]; RT[4] = R[4]; RT[5] = R[7];
    RT[6] = R[2]; RT[7] = R[5]; RT[8] = R[8];
    rotmat_from_trig(t[0], t[1], -t[2], -t[3], t[4], -t[5], Rinv);
    const V3 pinv = matvec(RT, -d[0], -d[1], -d[2]);
    const unsigned nr = rec.n, nL = lig.n;
    for (unsigned j = lane; j < nr; j += kWave) {
        const float4 x = rec.pos[j];
        BeadModes<ATT_MODES_MAX> md;
        load_modes(rec, j, md);
        const V3 df = deform(md, rec.M, x.x, x.y, x.z, d + 6);
        rdefo[j] = df.x; rdefo[nr + j] = df.y; rdefo[2 * nr + j] = df.z;
        V3 p = matvec(Rinv, df.x, df.y, df.z);
        p.x = p.x + pinv.x; p.y = p.y + pinv.y; p.z = p.z + pinv.z;
        rxyz[j] = p.x; rxyz[nr + j] = p.y; rxyz[2 * nr + j] = p.z;
        const float4 pp = pot_bead(lInner, lOuter, p.x, p.y, p.z, rec.mtype[j], x.w, nullptr);
        const V3 f = matvec(R, pp.x, pp.y, pp.z);
        rpot[j] = f.x; rpot[nr + j] = f.y; rpot[2 * nr + j] = f.z;
    }
    for (unsigned i = lane; i < nL; i += kWave) {
        const float4 x = lig.pos[i];
        BeadModes<ATT_MODES_MAX> md;
        load_modes(lig, i, md);
        const V3 df = deform(md, lig.M, x.x, x.y, x.z, d + 16);
        V3 p = matvec(R, df.x, df.y, df.z);
        p.x = p.x + d[0]; p.y = p.y + d[1]; p.z = p.z + d[2];
        xyz[i] = p.x; xyz[nL + i] = p.y; xyz[2 * nL + i] = p.z;
        int* c = cell + 9 * i;
        for (int k = 0; k < 9; ++k) c[k] = -1;
        const float4 pp = pot_bead(rInner, rOuter, p.x, p.y, p.z, lig.mtype[i], x.w, c);
        pot[i] = pp.x; pot[nL + i] = pp.y; pot[2 * nL + i] = pp.z; pot[3 * nL + i] = pp.w;
    }
    __syncthreads();
    if (lane != 0) return;
    auto recPos = [&](unsigned j) { return make_float3(rdefo[j], rdefo[nr + j], rdefo[2 * nr + j]); };
    auto reaction = [&](unsigned j, float fx, float fy, float fz) {
        rpot[j] += fx; rpot[nr + j] += fy; rpot[2 * nr + j] += fz;
    };
    for (unsigned i = 0; i < nL; ++i) {
        const float4 x = lig.pos[i];
        float4 a;
        unsigned np = 0;
        if (nl_bead(nl, par, par.tab, recPos, RecTypeG{rec.type}, RecChargeG{rec.pos}, reaction,
                    xyz[i], xyz[nL + i], xyz[2 * nL + i], lig.type[i], x.w, a, cell + 9 * i + 6,
                    &np)) {
            pot[i] += a.x; pot[nL + i] += a.y; pot[2 * nL + i] += a.z; pot[3 * nL + i] += a.w;
        }
        npairs[i] = np;
    }
}

__global__ void k_debug_rotmats(const float* ang, unsigned n, float* R) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // Per-lane evaluation (not the wave-cooperative path) of the same formula.
    double s0, c0, s1, c1, s2, c2;
    sincos((double)ang[3 * i + 1], &s0, &c0);  // ssi
    sincos((double)ang[3 * i + 0], &s1, &c1);  // phi
    sincos((double)ang[3 * i + 2], &s2, &c2);  // rot
    rotmat_from_trig((float)c0, (float)c1, (float)s0, (float)s1, (float)c2, (float)s2, R + 9 * i);
}

// Expands a [grid][z][y][x] float4 grid into voxel octets (see DGrid).
__global__ __launch_bounds__(256) void k_octets(const float4* __restrict__ g,
                                                float4* __restrict__ oct, unsigned nx,
                                                unsigned ny, unsigned nz, unsigned ngrids) {
    const size_t nvo = (size_t)(nx - 1) * (ny - 1) * (nz - 1);
    const size_t total = nvo * ngrids;
    for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total;
         o += (size_t)gridDim.x * blockDim.x) {
        size_t r = o;
        const unsigned ix = r % (nx - 1); r /= (nx - 1);
        const unsigned iy = r % (ny - 1); r /= (ny - 1);
        const unsigned iz = r % (nz - 1); r /= (nz - 1);
        const size_t t = r;
        const size_t nxy = (size_t)nx * ny;
        const float4* b = g + t * nxy * nz + nxy * iz + (size_t)nx * iy + ix;
        float4* d = oct + 8 * o;
        d[0] = b[0];
        d[1] = b[1];
        d[2] = b[nx];
        d[3] = b[nxy];
        d[4] = b[nxy + 1];
        d[5] = b[nxy + nx];
        d[6] = b[nx + 1];
        d[7] = b[nxy + nx + 1];
    }
}

__global__ __launch_bounds__(256) void k_copy(float4* __restrict__ dst,
                                              const float4* __restrict__ src, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------------ launchers

hipError_t launch_octets(const float4* grid, float4* oct, unsigned nx, unsigned ny, unsigned nz,
                         unsigned ngrids, hipStream_t s) {
    const size_t total = (size_t)(nx - 1) * (ny - 1) * (nz - 1) * ngrids;
    const size_t blocks = std::min<size_t>((total + 255) / 256, 256 * 256);
    hipLaunchKernelGGL(k_octets, dim3((unsigned)blocks), dim3(256), 0, s, grid, oct, nx, ny, nz,
                       ngrids);
    return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t n16, hipStream_t s) {
    if (n16 == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s,
                       static_cast<float4*>(dst), static_cast<const float4*>(src), n16);
    return hipGetLastError();
}

hipError_t launch_rigid(const DProt& rec, const DProt& lig, const DGrid& inner, const DGrid& outer,
                        const DNL& nl, const DPar& par, const float* dofs, float* res, unsigned n,
                        const unsigned* order, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned posesPerBlock = 4;
    const unsigned blocks = (n + posesPerBlock - 1) / posesPerBlock;
    const unsigned slots = par.T <= kMaxTypesLDS ? par.T * par.T : 0;
    const size_t lds = rigid_slot_offset(slots) + posesPerBlock * kRigidSlotBytes;
    hipLaunchKernelGGL(k_rigid, dim3(blocks), dim3(256), lds, s, rec, lig, inner, outer, nl, par,
                       dofs, res, n, slots, order);
    return hipGetLastError();
}

// Dynamic LDS of k_modes: pair table, deformed receptor xyz, receptor types
// (u16), the neighbour-list queue (one u32 per ligand bead, power of two),
// then (16-byte aligned) the slot arrays of the pair-parallel phase and one
// mark bit per ligand bead.
size_t modes_lds_bytes(const DProt& rec, const DProt& lig, const DPar& par, unsigned* tabSlots) {
    const unsigned slots = par.T <= kMaxTypesLDS ? par.T * par.T : 0;
    if (tabSlots) *tabSlots = slots;
    return modes_slot_offset(rec, lig, slots) + kModesSlots * (16 + 3 * sizeof(unsigned)) +
           sizeof(unsigned) * (((size_t)lig.n + 31) / 32);
}

hipError_t launch_modes(const DProt& rec, const DProt& lig, const DGrid& rInner,
                        const DGrid& rOuter, const DNL& nl, const DGrid& lInner,
                        const DGrid& lOuter, const DPar& par, const float* dofs, float* res,
                        unsigned n, const unsigned* order, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned slots = 0;
    const size_t lds = modes_lds_bytes(rec, lig, par, &slots);
    const unsigned mm = rec.M > lig.M ? rec.M : lig.M;
    if (mm == 0)
        hipLaunchKernelGGL(k_modes<0>, dim3(n), dim3(256), lds, s, rec, lig, rInner, rOuter, nl,
                           lInner, lOuter, par, dofs, res, n, slots, order);
    else if (mm <= 5)
        hipLaunchKernelGGL(k_modes<5>, dim3(n), dim3(256), lds, s, rec, lig, rInner, rOuter, nl,
                           lInner, lOuter, par, dofs, res, n, slots, order);
    else
        hipLaunchKernelGGL(k_modes<ATT_MODES_MAX>, dim3(n), dim3(256), lds, s, rec, lig, rInner,
                           rOuter, nl, lInner, lOuter, par, dofs, res, n, slots, order);
    return hipGetLastError();
}

hipError_t launch_debug_rigid(const DProt& rec, const DProt& lig, const DGrid& inner,
                              const DGrid& outer, const DNL& nl, const DPar& par,
                              const float* dof, float* xyz, float* pot, int* cell,
                              unsigned* npairs, hipStream_t s) {
    hipLaunchKernelGGL(k_debug_rigid, dim3(1), dim3(kWave), 0, s, rec, lig, inner, outer, nl, par,
                       dof, xyz, pot, cell, npairs);
    return hipGetLastError();
}

hipError_t launch_debug_modes(const DProt& rec, const DProt& lig, const DGrid& rInner,
                              const DGrid& rOuter, const DNL& nl, const DGrid& lInner,
                              const DGrid& lOuter, const DPar& par, const float* dof, float* xyz,
                              float* pot, int* cell, unsigned* npairs, float* rpot, float* rxyz,
                              float* rdefo, hipStream_t s) {
    hipLaunchKernelGGL(k_debug_modes, dim3(1), dim3(kWave), 0, s, rec, lig, rInner, rOuter, nl,
                       lInner, lOuter, par, dof, xyz, pot, cell, npairs, rpot, rxyz, rdefo);
    return hipGetLastError();
}

hipError_t launch_debug_rotmats(const float* ang, unsigned n, float* R, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_rotmats, dim3((n + 255) / 256), dim3(256), 0, s, ang, n, R);
    return hipGetLastError();
}

}  // namespace att
