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

// Beads per super-chunk of k_rigid (per wave) and their LDS: bead force and
// energy (float4), neighbour-list {count, start} (uint2), the class-ordered
// queue (u16) and its counters.
constexpr unsigned kRigidSuper = 256;
constexpr size_t kRigidWaveBytes =
    (kRigidSuper * (16 + 8 + 2) + sizeof(unsigned) * queue_counters(kRigidSuper) + 15) & ~(size_t)15;
__host__ __device__ inline size_t rigid_wave_offset(unsigned tabSlots) {
    return (sizeof(att_pair_params) * (size_t)tabSlots + 15) & ~(size_t)15;
}

__global__ __launch_bounds__(256) void k_rigid(DProt rec, DProt lig, DGrid inner, DGrid outer,
                                               DNL nl, DPar par, const float* __restrict__ dofs,
                                               float* __restrict__ res, unsigned n,
                                               unsigned tabSlots,
                                               const unsigned* __restrict__ order,
                                               uint2* __restrict__ counts) {
    extern __shared__ __align__(16) att_pair_params s_tab[];
    const PairTab tab = stage_table(par, s_tab);
    __syncthreads();
    const unsigned lane = threadIdx.x & (kWave - 1);
    const unsigned wave = threadIdx.x / kWave;
    char* wb = reinterpret_cast<char*>(s_tab) + rigid_wave_offset(tabSlots) + wave * kRigidWaveBytes;
    float4* s_f = reinterpret_cast<float4*>(wb);
    uint2* s_n = reinterpret_cast<uint2*>(s_f + kRigidSuper);
    unsigned short* s_q = reinterpret_cast<unsigned short*>(s_n + kRigidSuper);
    unsigned* s_cc = reinterpret_cast<unsigned*>(s_q + kRigidSuper);
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
    auto place = [&](unsigned i, float4& x) {
        x = lig.pos[i];
        V3 p = matvec(R, x.x, x.y, x.z);
        p.x = p.x + tx; p.y = p.y + ty; p.z = p.z + tz;
        return p;
    };
    float red = 0.f;  // the reference's reduction held by this lane (reduce_beads_ref)
    unsigned listed = 0, inside = 0;  // this lane's pair counts (counts != nullptr)
    const unsigned span = super_chunk(lig.n, kRigidSuper);
    for (unsigned sc = 0; sc < lig.n; sc += span) {
        const unsigned m = min(span, lig.n - sc);
        // P1: one bead per lane: rotate_translate, potForce, the list's cell.
        for (unsigned k = lane; k < m; k += kWave) {
            float4 x;
            const V3 p = place(sc + k, x);
            s_f[k] = pot_bead(inner, outer, p.x, p.y, p.z, lig.mtype[sc + k], x.w, nullptr);
            s_n[k] = nl_cell(nl, p.x, p.y, p.z, lig.type[sc + k]);
        }
        wave_sync();
        // P2: beads with a list, longest lists first.
        const unsigned qn = class_queue([&](unsigned k) { return s_n[k].x; }, m, s_q, s_cc, lane,
                                        kWave, wave_sync);
        // P3: one bead per lane walks its list in the reference's order
        // (NLPotForce); its sums are added to the bead's grid values as
        // outLig[i] += fAcc.
        for (unsigned q0 = 0; q0 < qn; q0 += kWave) {
            if (q0 + lane < qn) {
                const unsigned k = s_q[q0 + lane];
                float4 x;
                const V3 p = place(sc + k, x);
                unsigned np = 0;
                const float4 a = nl_walk(nl, par, tab, RecPosG{rec.pos}, RecTypeG{rec.type},
                                         RecChargeG{rec.pos}, NoReaction{}, p.x, p.y, p.z,
                                         lig.type[sc + k], x.w, s_n[k], &np);
                listed += s_n[k].x;
                inside += np;
                float4 f = s_f[k];
                f.x = f.x + a.x; f.y = f.y + a.y; f.z = f.z + a.z; f.w = f.w + a.w;
                s_f[k] = f;
            }
        }
        wave_sync();
        // P4: reducePotForce and reduceTorque, in bead order.
        red = reduce_beads_ref(red, lane, s_f, lig.pos, sc, m);
        wave_sync();  // the super-chunk's arrays are reused
    }
    float o[7];
    finish_ref(red, t, o);
    if (lane == 0) {
        float* r = res + 7 * (size_t)pose;
#pragma unroll
        for (int q = 0; q < 7; ++q) r[q] = o[q];
    }
    if (counts) {
        listed = wave_sum_u(listed);
        inside = wave_sum_u(inside);
        if (lane == 0) counts[pose] = make_uint2(listed, inside);
    }
}

// The deformed receptor position of bead j from its padded mode row (one or
// two 16-byte loads), in the order of deform(): bit-identical coordinates.
// The row is also left in `rowOut` (4 * ceil(3M / 4) floats) for the pair's
// reaction, which then does not load it again.
template <int MM>
struct RecPosDeform {
    const float4* pos;
    const float4* modesT;
    unsigned rowT, M;
    const float* amp;
    float* rowOut;
    __device__ float3 operator()(unsigned j) const {
        const float4 x = pos[j];
        float yx = x.x, yy = x.y, yz = x.z;
        if (MM > 0) {
            constexpr int MA = MM > 0 ? MM : 1;
            constexpr unsigned Q = (3 * MA + 3) / 4;
            float mt[4 * Q];
            const float4* row = modesT + (size_t)j * rowT;
#pragma unroll
            for (unsigned q = 0; q < Q; ++q) {
                const float4 v = q < rowT ? row[q] : make_float4(0.f, 0.f, 0.f, 0.f);
                mt[4 * q] = v.x; mt[4 * q + 1] = v.y; mt[4 * q + 2] = v.z; mt[4 * q + 3] = v.w;
            }
#pragma unroll
            for (unsigned i = 0; i < 4 * Q; ++i) rowOut[i] = mt[i];
#pragma unroll
            for (unsigned m = 0; m < (unsigned)MA; ++m)
                if (m < M) {
                    const float a = amp[m];
                    yx = yx + a * mt[3 * m];
                    yy = yy + a * mt[3 * m + 1];
                    yz = yz + a * mt[3 * m + 2];
                }
        }
        return make_float3(yx, yy, yz);
    }
};

// reduceModeForce of the ligand (model/reduction_modes.h:28-57) for mode mm
// over beads f[0..m) (LDS) of a super-chunk starting at bead sc, in bead
// order: red -= (R^T f) . mode.  The mode vectors of eight beads are loaded
// together.
__device__ __forceinline__ float reduce_modes_ref(float red, const float* RT, const float4* f,
                                                  const DProt& lig, unsigned mm, unsigned sc,
                                                  unsigned m) {
    const size_t nM = (size_t)lig.M * lig.n;
    const float* mx = lig.modes + (size_t)mm * lig.n + sc;
    unsigned k = 0;
    for (; k + 8 <= m; k += 8) {
        float ax[8], ay[8], az[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            ax[u] = mx[k + u];
            ay[u] = mx[nM + k + u];
            az[u] = mx[2 * nM + k + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float4 v = f[k + u];
            const V3 fa = matvec(RT, v.x, v.y, v.z);
            red -= fa.x * ax[u] + fa.y * ay[u] + fa.z * az[u];
        }
    }
    for (; k < m; ++k) {
        const float4 v = f[k];
        const V3 fa = matvec(RT, v.x, v.y, v.z);
        red -= fa.x * mx[k] + fa.y * mx[nM + k] + fa.z * mx[2 * nM + k];
    }
    return red;
}

// One workgroup of W waves per pose (W = 1 below 384 ligand beads, else 4;
// att_set_modes_waves overrides it).
// Phase A: every receptor bead deformed and moved into the ligand frame
// (invertDOF, rotate_translate_deform), scored on the ligand grid, its force
// rotated back (rotate_forces) and projected on the receptor modes.  Phase B:
// the ligand in super-chunks of at most 256 * W beads, as k_rigid: grid term and list
// cell per bead (P1), the class-ordered queue (P2), one bead per lane walking
// its list in the reference's order against the deformed receptor (each
// receptor bead deformed again from its mode row, so receptors of any size
// need no LDS) with the reaction projected on the receptor modes (P3), and
// the reference's bead-order reductions (P4): E, F and the torque tensor in
// lanes 0..12 of wave 0, the ligand's reduceModeForce in lanes of the last
// wave.  Semantics of service/CPUEnergyService6DModes.tpp:121-327.
// Occupancy: at least 4 waves per SIMD (3 with 10 modes), as the adjoint.
template <int MM, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(MM >= 10 ? 3 : 4))) void k_modes(DProt rec, DProt lig, DGrid rInner, DGrid rOuter,
                                                  DNL nl, DGrid lInner, DGrid lOuter, DPar par,
                                                  const float* __restrict__ dofs,
                                                  float* __restrict__ res, unsigned n,
                                                  const unsigned* __restrict__ order,
                                                  uint2* __restrict__ counts) {
    constexpr unsigned kSuper = 256u * W;
    extern __shared__ __align__(16) att_pair_params s_tab[];
    __shared__ float4 s_f[kSuper];
    __shared__ uint2 s_n[kSuper];
    __shared__ unsigned short s_q[kSuper];
    __shared__ unsigned s_cc[queue_counters(kSuper)];
    __shared__ float s_red[W][ATT_MODES_MAX];
    __shared__ float s_gl[ATT_MODES_MAX];
    __shared__ unsigned s_cnt[W][2];
    const PairTab tab = stage_table(par, s_tab);  // visible after phase B's first barrier
    if (blockIdx.x >= n) return;  // uniform per workgroup
    const unsigned pose = order ? order[xcd_slot(blockIdx.x, gridDim.x)] : blockIdx.x;
    const unsigned tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    constexpr unsigned nthr = 64u * W;
    constexpr unsigned modeWave = W - 1;  // runs the ligand's reduceModeForce
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

    float mR[MA];
#pragma unroll
    for (int m = 0; m < MA; ++m) mR[m] = 0.f;

    // Phase A: receptor beads, two lookups in flight per lane (pot_bead_flat).
#pragma unroll 2
    for (unsigned j = tid; j < rec.n; j += nthr) {
        const float4 x = rec.pos[j];
        BeadModes<MM> md;
        load_modes(rec, j, md);
        const V3 df = deform(md, rec.M, x.x, x.y, x.z, ampR);
        V3 p = matvec(Rinv, df.x, df.y, df.z);
        p.x = p.x + pinv.x; p.y = p.y + pinv.y; p.z = p.z + pinv.z;
        const float4 pot = pot_bead_flat(lInner, lOuter, p.x, p.y, p.z, rec.mtype[j], x.w);
        const V3 f = matvec(R, pot.x, pot.y, pot.z);  // rotate_forces (transform_modes.h:141-161)
        mode_dots(md, rec.M, f.x, f.y, f.z, mR);
    }

    // Phase B: ligand beads.
    const float4* __restrict__ rpos = rec.pos;
    const float4* __restrict__ rmodesT = rec.modesT;
    const unsigned Mr = rec.M, rowT = rec.rowT;
    // The walk takes one list entry at a time, so the reaction of a pair
    // follows the deformation of the same receptor bead: its mode row is the
    // one RecPosDeform left in rowBuf.
    constexpr int kWalkBatch = 1;
    static_assert(kWalkBatch == 1, "the reaction reads the row of the last deformed bead");
    float rowBuf[4 * ((3 * MA + 3) / 4)];
    auto reaction = [&](unsigned, float fx, float fy, float fz) {
        const float* mt = rowBuf;
#pragma unroll
        for (unsigned m = 0; m < (unsigned)MM; ++m)
            if (m < Mr) mR[m] -= fx * mt[3 * m] + fy * mt[3 * m + 1] + fz * mt[3 * m + 2];
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
    const unsigned Ml = lig.M, nL = lig.n;
    float red = 0.f;  // wave 0: E, F, torque tensor; modeWave: ligand modes
    unsigned listed = 0, inside = 0;  // this lane's pair counts (counts != nullptr)
    const unsigned span = super_chunk(nL, kSuper);
    for (unsigned sc = 0; sc < nL; sc += span) {
        const unsigned m = min(span, nL - sc);
        // P1
        for (unsigned k = tid; k < m; k += nthr) {
            float4 x;
            const V3 p = place(sc + k, x);
            s_f[k] = pot_bead(rInner, rOuter, p.x, p.y, p.z, lig.mtype[sc + k], x.w, nullptr);
            s_n[k] = nl_cell(nl, p.x, p.y, p.z, lig.type[sc + k]);
        }
        __syncthreads();
        // P2
        const unsigned qn = class_queue([&](unsigned k) { return s_n[k].x; }, m, s_q, s_cc, tid,
                                        nthr, [] { __syncthreads(); });
        // P3
        for (unsigned q0 = 0; q0 < qn; q0 += nthr) {
            if (q0 + tid < qn) {
                const unsigned k = s_q[q0 + tid];
                float4 x;
                const V3 p = place(sc + k, x);
                unsigned np = 0;
                const float4 a = nl_walk<kWalkBatch>(
                    nl, par, tab, RecPosDeform<MM>{rpos, rmodesT, rowT, Mr, ampR, rowBuf},
                    RecTypeG{rec.type}, RecChargeG{rpos}, reaction, p.x, p.y, p.z,
                    lig.type[sc + k], x.w, s_n[k], &np);
                listed += s_n[k].x;
                inside += np;
                float4 f = s_f[k];
                f.x = f.x + a.x; f.y = f.y + a.y; f.z = f.z + a.z; f.w = f.w + a.w;
                s_f[k] = f;
            }
        }
        __syncthreads();
        // P4: reducePotForce, reduceTorque and the ligand's reduceModeForce, in
        // bead order.
        if (wave == 0) red = reduce_beads_ref(red, lane, s_f, lig.pos, sc, m);
        const unsigned ml = W > 1 ? lane : lane - 13u;  // mode of this lane in modeWave
        if (wave == modeWave && ml < Ml) red = reduce_modes_ref(red, RT, s_f, lig, ml, sc, m);
        __syncthreads();  // the super-chunk's arrays are reused
    }

    // Receptor mode gradients: wave butterflies, then the waves in order.
#pragma unroll
    for (int m = 0; m < MM; ++m) mR[m] = wave_sum(mR[m]);
    if (lane == 0) {
#pragma unroll
        for (int m = 0; m < ATT_MODES_MAX; ++m) s_red[wave][m] = m < MM ? mR[m < MM ? m : 0] : 0.f;
    }
    if (wave == modeWave) {
        const unsigned ml = W > 1 ? lane : lane - 13u;
        if (ml < ATT_MODES_MAX) s_gl[ml] = red;
    }
    if (counts) {
        listed = wave_sum_u(listed);
        inside = wave_sum_u(inside);
        if (lane == 0) { s_cnt[wave][0] = listed; s_cnt[wave][1] = inside; }
    }
    __syncthreads();
    if (counts && threadIdx.x == 0) {
        uint2 c = make_uint2(0u, 0u);
        for (int w = 0; w < W; ++w) { c.x += s_cnt[w][0]; c.y += s_cnt[w][1]; }
        counts[pose] = c;
    }
    if (wave == 0) {
        float o[7];
        finish_ref(red, t, o);
        if (lane == 0) {
            float* r = res + 27 * (size_t)pose;
#pragma unroll
            for (int q = 0; q < 7; ++q) r[q] = o[q];
            for (unsigned m = 0; m < ATT_MODES_MAX; ++m) {
                float s = s_red[0][m];
                for (int w = 1; w < W; ++w) s += s_red[w][m];
                r[7 + m] = m < rec.M ? correct_mode(s, rec.modeForce[m]) : 0.f;
                r[17 + m] = m < lig.M ? correct_mode(s_gl[m], lig.modeForce[m]) : 0.f;
            }
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
        if (nl_bead(nl, par, PairTab{par.tab, false}, RecPosG{rec.pos}, RecTypeG{rec.type},
                    RecChargeG{rec.pos}, NoReaction{}, p.x, p.y, p.z, lig.type[i], x.w, a, c + 6,
                    &np)) {
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
        // the lookup k_modes' phase A makes, so its per-bead values are under test
        const float4 pp = pot_bead_flat(lInner, lOuter, p.x, p.y, p.z, rec.mtype[j], x.w);
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
        if (nl_bead(nl, par, PairTab{par.tab, false}, recPos, RecTypeG{rec.type},
                    RecChargeG{rec.pos}, reaction, xyz[i], xyz[nL + i], xyz[2 * nL + i],
                    lig.type[i], x.w, a, cell + 9 * i + 6, &np)) {
            pot[i] += a.x; pot[nL + i] += a.y; pot[2 * nL + i] += a.z; pot[3 * nL + i] += a.w;
        }
        npairs[i] = np;
    }
}

__global__ void k_debug_rotmats(const float* ang, unsigned n, float* R) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // Per-lane evaluation (not the wave-cooperative path) of the same formula.
    float s0, c0, s1, c1, s2, c2;
    sincosf_libm(ang[3 * i + 1], &s0, &c0);  // ssi
    sincosf_libm(ang[3 * i + 0], &s1, &c1);  // phi
    sincosf_libm(ang[3 * i + 2], &s2, &c2);  // rot
    rotmat_from_trig(c0, c1, s0, s1, c2, s2, R + 9 * i);
}

__global__ void k_debug_sincosf(const float* x, unsigned n, float* sn, float* cs) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sincosf_libm(x[i], sn + i, cs + i);
}

__global__ void k_debug_plateau_ratio(float P2, unsigned n, float* ratio) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ratio[i] = plateau_ratio(i, P2);
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
                        const unsigned* order, uint2* counts, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned posesPerBlock = 4;
    const unsigned blocks = (n + posesPerBlock - 1) / posesPerBlock;
    const unsigned slots = par.T <= kMaxTypesLDS ? par.T * par.T : 0;
    const size_t lds = rigid_wave_offset(slots) + posesPerBlock * kRigidWaveBytes;
    hipLaunchKernelGGL(k_rigid, dim3(blocks), dim3(256), lds, s, rec, lig, inner, outer, nl, par,
                       dofs, res, n, slots, order, counts);
    return hipGetLastError();
}

// Dynamic LDS of k_modes: the pair table (the super-chunk arrays are static).
size_t modes_lds_bytes(const DProt& rec, const DProt& lig, const DPar& par, unsigned* tabSlots) {
    (void)rec;
    (void)lig;
    const unsigned slots = par.T <= kMaxTypesLDS ? par.T * par.T : 0;
    if (tabSlots) *tabSlots = slots;
    return sizeof(att_pair_params) * (size_t)slots;
}

hipError_t launch_modes(const DProt& rec, const DProt& lig, const DGrid& rInner,
                        const DGrid& rOuter, const DNL& nl, const DGrid& lInner,
                        const DGrid& lOuter, const DPar& par, const float* dofs, float* res,
                        unsigned n, const unsigned* order, int waves, uint2* counts, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t lds = modes_lds_bytes(rec, lig, par, nullptr);
    const unsigned mm = rec.M > lig.M ? rec.M : lig.M;
#define ATT_LAUNCH_MODES(MMV, WV)                                                                 \
    hipLaunchKernelGGL((k_modes<MMV, WV>), dim3(n), dim3(64 * WV), lds, s, rec, lig, rInner, rOuter, \
                       nl, lInner, lOuter, par, dofs, res, n, order, counts)
#define ATT_LAUNCH_MODES_W(MMV)                                                                  \
    switch (waves) {                                                                              \
    case 8: ATT_LAUNCH_MODES(MMV, 8); break;                                                      \
    case 4: ATT_LAUNCH_MODES(MMV, 4); break;                                                      \
    case 2: ATT_LAUNCH_MODES(MMV, 2); break;                                                      \
    default: ATT_LAUNCH_MODES(MMV, 1); break;                                                     \
    }
    if (mm == 0) ATT_LAUNCH_MODES_W(0)
    else if (mm <= 5) ATT_LAUNCH_MODES_W(5)
    else ATT_LAUNCH_MODES_W(ATT_MODES_MAX)
#undef ATT_LAUNCH_MODES_W
#undef ATT_LAUNCH_MODES
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

hipError_t launch_debug_sincosf(const float* x, unsigned n, float* sn, float* cs, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_sincosf, dim3((n + 255) / 256), dim3(256), 0, s, x, n, sn, cs);
    return hipGetLastError();
}

hipError_t launch_debug_plateau_ratio(float P2, unsigned n, float* ratio, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_plateau_ratio, dim3((n + 255) / 256), dim3(256), 0, s, P2, n, ratio);
    return hipGetLastError();
}

hipError_t launch_debug_rotmats(const float* ang, unsigned n, float* R, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_rotmats, dim3((n + 255) / 256), dim3(256), 0, s, ang, n, R);
    return hipGetLastError();
}

}  // namespace att
