// Hand-written adjoint: exact gradient of the returned energy.
//
// E = sum over ligand beads of the interpolated grid energy (type grid plus
//     charge * electrostatic grid) + sum over listed pairs inside the plateau of
//     [LJ(r) - LJ(P)] + [Coulomb(r) - Coulomb(P)]
// (the energy of service/CPUEnergyService6D.tpp:100-195 and of the modes path,
// service/CPUEnergyService6DModes.tpp:121-327, whose E does not include the
// receptor's own interaction with the ligand grid).
//
// Gradients are recomputed, never stored per bead: the grid term by the
// derivative of the trilinear interpolant of the energy channel (not the
// interpolated force channel the reference reports), the pair term by the pair
// force (the plateau term is constant in position), the rotation by
// dR/d(angle) contracted with sum_i dE/dp_i (x) x~_i over the *deformed*
// ligand coordinates, the modes by the chain rule through the deformation.
// Parameter derivatives (rc, ac, emin of every type pair) come from the pair
// term directly and, for the grid term, from the per-receptor-type basis grids
// (the grid energy is sum_a ipon (rc S_a - ac S6_a) when the plateau exceeds
// every rmin).
//
// k_adjoint: persistent, one 256-thread workgroup per pose at a time; per-pose
// values go through wave butterflies, parameter derivatives accumulate in LDS
// across the workgroup's poses and are written once per workgroup, then
// k_pgrad_groups and k_pgrad_reduce add the per-workgroup rows in a fixed
// order.
#include <algorithm>

#include "device.h"
#include "internal.h"

namespace att {

// Energy and its position gradient at one point: pot as pot_bead computes it
// (bit-identical energy), plus d(E channel)/dp of the same voxel.
__device__ __forceinline__ void trilin_grad(const DGrid& g, float px, float py, float pz,
                                            unsigned ix, unsigned iy, unsigned iz, unsigned type,
                                            float4& V, float3& dE) {
    const float mnx = (float)(int)ix * g.dVox + g.lox;
    const float mny = (float)(int)iy * g.dVox + g.loy;
    const float mnz = (float)(int)iz * g.dVox + g.loz;
    const float ax = px - mnx, ay = py - mny, az = pz - mnz;
    const float bx = (mnx + g.dVox) - px, by = (mny + g.dVox) - py, bz = (mnz + g.dVox) - pz;
    V = trilin(g, px, py, pz, ix, iy, iz, type);
    const size_t o = (((size_t)type * (g.nz - 1) + iz) * (g.ny - 1) + iy) * (g.nx - 1) + ix;
    const float4* d = g.oct + 8 * o;  // same lines as trilin: cache hits
    const float E0 = d[0].w, E1 = d[1].w, E2 = d[2].w, E3 = d[3].w;
    const float E4 = d[4].w, E5 = d[5].w, E6 = d[6].w, E7 = d[7].w;
    // corners: 0 (000) 1 (100) 2 (010) 3 (001) 4 (101) 5 (011) 6 (110) 7 (111)
    dE.x = ((E1 - E0) * (by * bz) + (E4 - E3) * (by * az) + (E6 - E2) * (ay * bz) +
            (E7 - E5) * (ay * az)) * g.vol_inv;
    dE.y = ((E2 - E0) * (bx * bz) + (E5 - E3) * (bx * az) + (E6 - E1) * (ax * bz) +
            (E7 - E4) * (ax * az)) * g.vol_inv;
    dE.z = ((E3 - E0) * (bx * by) + (E4 - E1) * (ax * by) + (E5 - E2) * (bx * ay) +
            (E7 - E6) * (ax * ay)) * g.vol_inv;
}

struct PGrad {
    float* s;                     // LDS accumulator, T*T*3, or nullptr when not requested
    const att_pair_params* tab;   // the LDS copy of the pair table
    unsigned T;
    float w;                      // upstream weight of the current pose
};

// Basis channels read per step of the parameter-gradient loop: the loads of
// kBasisStep channels are issued together (64 B each) and their LDS updates
// follow.  Timed with 12 types: steps of 2 and 3 channels 1.1% faster than 4
// and 6 (same results).
constexpr unsigned kBasisStep = 2;

// Grid term of one ligand bead on one grid: energy (the forward's per-bead
// value), gradient, and the basis-grid parameter derivatives.
__device__ __forceinline__ float grid_adj_one(const DGrid& g, const DBasis& b, float px, float py,
                                              float pz, unsigned mtype, float charge, float3& G,
                                              const PGrad& pg) {
    const unsigned ix = grid_index(px, g.lox, g.dVox_inv, g.nx);
    const unsigned iy = grid_index(py, g.loy, g.dVox_inv, g.ny);
    const unsigned iz = grid_index(pz, g.loz, g.dVox_inv, g.nz);
    float4 pot;
    float3 dE;
    trilin_grad(g, px, py, pz, ix, iy, iz, mtype, pot, dE);
    G = dE;
    if (fabsf(charge) > 0.001f) {
        float4 el;
        float3 dEl;
        trilin_grad(g, px, py, pz, ix, iy, iz, 0, el, dEl);
        el.w = el.w * charge;
        pot.w = pot.w + el.w;
        G.x += charge * dEl.x; G.y += charge * dEl.y; G.z += charge * dEl.z;
    }
    if (pg.s && b.oct && mtype <= pg.T) {
        const float mnx = (float)(int)ix * g.dVox + g.lox;
        const float mny = (float)(int)iy * g.dVox + g.loy;
        const float mnz = (float)(int)iz * g.dVox + g.loz;
        const float ax = px - mnx, ay = py - mny, az = pz - mnz;
        const float bx = (mnx + g.dVox) - px, by = (mny + g.dVox) - py, bz = (mnz + g.dVox) - pz;
        const float w0 = bx * by * bz, w1 = ax * by * bz, w2 = bx * ay * bz, w3 = bx * by * az;
        const float w4 = ax * by * az, w5 = bx * ay * az, w6 = ax * ay * bz, w7 = ax * ay * az;
        const size_t o = ((size_t)iz * (g.ny - 1) + iy) * (g.nx - 1) + ix;
        const float4* row = b.oct + o * (size_t)b.nb * 2;
        for (unsigned a0 = 0; a0 < pg.T; a0 += kBasisStep) {
            float4 v[4 * kBasisStep];
#pragma unroll
            for (unsigned j = 0; j < kBasisStep; ++j) {
                const unsigned a = a0 + j < pg.T ? a0 + j : a0;  // past T: re-read a valid row
                v[4 * j] = row[4 * a];          // S_a at the 8 corners
                v[4 * j + 1] = row[4 * a + 1];
                v[4 * j + 2] = row[4 * a + 2];  // S6_a
                v[4 * j + 3] = row[4 * a + 3];
            }
#pragma unroll
            for (unsigned j = 0; j < kBasisStep; ++j) {
                if (a0 + j >= pg.T) break;
                const float4 s0 = v[4 * j], s1 = v[4 * j + 1], t0 = v[4 * j + 2], t1 = v[4 * j + 3];
                const float S = (s0.x * w0 + s0.y * w1 + s0.z * w2 + s0.w * w3 + s1.x * w4 +
                                 s1.y * w5 + s1.z * w6 + s1.w * w7) * g.vol_inv;
                const float S6 = (t0.x * w0 + t0.y * w1 + t0.z * w2 + t0.w * w3 + t1.x * w4 +
                                  t1.y * w5 + t1.z * w6 + t1.w * w7) * g.vol_inv;
                const unsigned k = (a0 + j) * pg.T + (mtype - 1);
                const float iv = (float)pg.tab[k].ipon * pg.w;
                atomicAdd(&pg.s[3 * k], iv * S);
                atomicAdd(&pg.s[3 * k + 1], -iv * S6);
            }
        }
    }
    return pot.w;
}

// Grid term of one ligand bead: inner grid, else outer grid, else nothing
// (model/interpolation.h:56-102).  Each lane selects its grid's geometry and
// runs the one lookup path, so a wave whose beads straddle the inner grid's
// boundary does not walk the lookup (and the basis loop) twice.
__device__ __forceinline__ float grid_adj(const DGrid& inner, const DGrid& outer,
                                          const DBasis& bin, const DBasis& bout, const DPar& par,
                                          float px, float py, float pz, unsigned mtype,
                                          float charge, float3& G, const PGrad& pg) {
    G = make_float3(0.f, 0.f, 0.f);
    if (mtype == 0) return 0.f;
    const bool in = !grid_out(inner, px, py, pz);
    if (!in && grid_out(outer, px, py, pz)) return 0.f;
    DGrid g{};
    g.oct = in ? inner.oct : outer.oct;
    g.nx = in ? inner.nx : outer.nx;
    g.ny = in ? inner.ny : outer.ny;
    g.nz = in ? inner.nz : outer.nz;
    g.lox = in ? inner.lox : outer.lox;
    g.loy = in ? inner.loy : outer.loy;
    g.loz = in ? inner.loz : outer.loz;
    g.dVox = in ? inner.dVox : outer.dVox;
    g.dVox_inv = in ? inner.dVox_inv : outer.dVox_inv;
    g.vol_inv = in ? inner.vol_inv : outer.vol_inv;
    DBasis b{};
    b.oct = in ? bin.oct : bout.oct;
    b.nb = in ? bin.nb : bout.nb;
    return grid_adj_one(g, b, px, py, pz, mtype, charge, G, pg);
}

// LJ derivatives with respect to (rc, ac, emin) at squared distance dr2.
__device__ __forceinline__ void lj_dparam(float dr2, float dr2_inv, const att_pair_params& p,
                                          int shape, float& drc, float& dac, float& dem) {
    const float rr23 = dr2_inv * dr2_inv * dr2_inv;
    const float rrd = shape == 8 ? dr2_inv : rr23;
    const float s = dr2 < p.rmin2 ? 1.f : (float)p.ipon;
    drc = s * (rrd * rr23);
    dac = -s * rr23;
    dem = dr2 < p.rmin2 ? (float)p.ipon - 1.f : 0.f;
}

// Neighbour-list cell of a ligand bead: {count, start} of its list, {0, 0}
// when the bead has no neighbour-list term (same bounds and index as nl_bead).
__device__ __forceinline__ uint2 nl_cell(const DNL& nl, float px, float py, float pz,
                                         unsigned ligType) {
    if (ligType == 0) return make_uint2(0u, 0u);
    if (!((px >= nl.lox && px <= nl.hix) && (py >= nl.loy && py <= nl.hiy) &&
          (pz >= nl.loz && pz <= nl.hiz)))
        return make_uint2(0u, 0u);
    const int ix = (int)roundf((px - nl.lox) * nl.dVox_inv);
    const int iy = (int)roundf((py - nl.loy) * nl.dVox_inv);
    const int iz = (int)roundf((pz - nl.loz) * nl.dVox_inv);
    return nl.desc[nl.nx * (iz * nl.ny + iy) + ix];
}

// Workgroup-shared state of one pose (static LDS); WAVES waves per pose,
// 64 * WAVES queued ligand beads per pair round.
template <int WAVES>
struct PairLDS {
    static constexpr unsigned kSlots = 64u * WAVES;
    float4 p[kSlots];         // transformed bead position, charge
    float4 x[kSlots];         // deformed bead coordinates x~, bead index | type << 16
    unsigned start[kSlots];   // first list entry
    unsigned end[kSlots];     // inclusive prefix of the pair counts
    float4 G[kSlots];         // sum of pair forces on the bead
    float4 head[WAVES];       // a wave's share of the bead it starts inside
    unsigned scan[WAVES];
    unsigned qn;
    float red[WAVES][13 + 2 * ATT_MODES_MAX];
};

// Adjoint of one pose by one workgroup of WAVES waves.
//  B1   every ligand bead: transform, grid term with its exact gradient (and
//       the basis-grid parameter derivatives), chain rule; beads with a
//       neighbour list are recorded with their {count, start}.
//  Q    the recorded beads are compacted in bead order (block scan).
//  B2   pair rounds of up to kSlots beads: the pairs of the round are
//       flattened (prefix sum of counts) and split into WAVES contiguous
//       ranges, one per wave, walked in chunks of 64 consecutive pairs: list
//       reads are coalesced and every lane has work until the range ends.
//       Receptor beads are deformed on the fly from their padded mode row
//       (one cache line).  Pair forces are summed per bead by a segmented
//       wave reduction; a bead split between two waves gets the later wave's
//       share through `head`, added in wave order, so results are
//       deterministic.  Then the chain rule for the NL force of each bead.
template <int MM, int WAVES>
__device__ __forceinline__ void adjoint_pose(const PairDesc& P, unsigned pose, PairLDS<WAVES>& S,
                                             char* dyn, const att_pair_params* tab, float* s_pg) {
    const DProt& rec = P.rec;
    const DProt& lig = P.lig;
    const DNL& nl = P.nl;
    const DPar& par = P.par;
    uint2* s_nl = reinterpret_cast<uint2*>(dyn + adjoint_nl_offset(P.tabSlots));
    unsigned short* s_queue = reinterpret_cast<unsigned short*>(dyn + adjoint_queue_offset(P.tabSlots, lig.n));
    constexpr unsigned kSlots = PairLDS<WAVES>::kSlots;
    const unsigned lane = threadIdx.x & (kWave - 1);
    const unsigned wave = threadIdx.x / kWave;
    constexpr int MA = MM > 0 ? MM : 1;
    const float4* __restrict__ rpos = rec.pos;
    const float4* __restrict__ rmodesT = rec.modesT;
    const unsigned Mr = MM > 0 ? rec.M : 0, Ml = MM > 0 ? lig.M : 0, rowT = rec.rowT;
    {
        __syncthreads();  // LDS of the previous pose is no longer read
        const float* d = P.dofs + (size_t)P.dofStride * pose;
        const float tx = d[0], ty = d[1], tz = d[2];
        const float wp = P.wts ? P.wts[pose] : 1.f;
        const PGrad pg{s_pg, tab, par.T, wp};
        float ampR[MA], ampL[MA];
#pragma unroll
        for (int m = 0; m < MA; ++m) {
            ampR[m] = MM > 0 ? d[6 + m] : 0.f;
            ampL[m] = MM > 0 ? d[16 + m] : 0.f;
        }
        float t[6];
        pose_trig(d[3], d[4], d[5], lane, t);
        float R[9];
        rotmat_from_trig(t[0], t[1], t[2], t[3], t[4], t[5], R);
        // accumulators: E, dE/dt (3), sum G (x) x~ (9), receptor and ligand modes
        float acc[13];
        float gR[MA], gL[MA];
#pragma unroll
        for (int k = 0; k < 13; ++k) acc[k] = 0.f;
#pragma unroll
        for (int m = 0; m < MA; ++m) { gR[m] = 0.f; gL[m] = 0.f; }

        auto place = [&](unsigned i, float4& x, V3& xt) __attribute__((always_inline)) {
            x = lig.pos[i];
            BeadModes<MM> md;
            load_modes(lig, i, md);
            xt = deform(md, Ml, x.x, x.y, x.z, ampL);
            V3 p = matvec(R, xt.x, xt.y, xt.z);
            p.x = p.x + tx; p.y = p.y + ty; p.z = p.z + tz;
            return p;
        };
        // Chain rule through p = R x~ + t for one bead's dE/dp = G.
        auto add_grad = [&](unsigned i, float xx, float xy, float xz, float Gx, float Gy, float Gz)
            __attribute__((always_inline)) {
            acc[1] += Gx; acc[2] += Gy; acc[3] += Gz;
            acc[4] += Gx * xx; acc[5] += Gx * xy; acc[6] += Gx * xz;
            acc[7] += Gy * xx; acc[8] += Gy * xy; acc[9] += Gy * xz;
            acc[10] += Gz * xx; acc[11] += Gz * xy; acc[12] += Gz * xz;
            if (MM > 0) {
                // G . R m = (R^T G) . m
                const float gx = R[0] * Gx + R[3] * Gy + R[6] * Gz;
                const float gy = R[1] * Gx + R[4] * Gy + R[7] * Gz;
                const float gz = R[2] * Gx + R[5] * Gy + R[8] * Gz;
                BeadModes<MM> md;
                load_modes(lig, i, md);
                mode_dots(md, Ml, -gx, -gy, -gz, gL);  // mode_dots subtracts
            }
        };

        // B1: grid term of every ligand bead; neighbour-list cells recorded.
        for (unsigned i = threadIdx.x; i < lig.n; i += blockDim.x) {
            float4 x;
            V3 xt;
            const V3 p = place(i, x, xt);
            float3 G;
            acc[0] += grid_adj(P.inner, P.outer, P.bin, P.bout, par, p.x, p.y, p.z, lig.mtype[i], x.w, G, pg);
            add_grad(i, xt.x, xt.y, xt.z, G.x, G.y, G.z);
            s_nl[i] = nl_cell(nl, p.x, p.y, p.z, lig.type[i]);
        }
        __syncthreads();

        // Q: beads with listed neighbours, in bead order.
        {
            const unsigned per = (lig.n + blockDim.x - 1) / blockDim.x;
            const unsigned lo = threadIdx.x * per, hi = min(lig.n, lo + per);
            unsigned cnt = 0;
            for (unsigned i = lo; i < hi; ++i) cnt += s_nl[i].x != 0u;
            unsigned total;
            unsigned o = block_scan(cnt, S.scan, total);
            for (unsigned i = lo; i < hi; ++i)
                if (s_nl[i].x != 0u) s_queue[o++] = (unsigned short)i;
            if (threadIdx.x == 0) S.qn = total;
        }
        __syncthreads();
        const unsigned qn = S.qn;

        // B2: pair rounds.
        for (unsigned base = 0; base < qn; base += kSlots) {
            const unsigned ns = min(kSlots, qn - base);
            unsigned cnt = 0;
            if (threadIdx.x < ns) {
                const unsigned i = s_queue[base + threadIdx.x];
                float4 x;
                V3 xt;
                const V3 p = place(i, x, xt);
                const uint2 c = s_nl[i];
                cnt = c.x;
                S.p[threadIdx.x] = make_float4(p.x, p.y, p.z, x.w);
                S.x[threadIdx.x] = make_float4(xt.x, xt.y, xt.z, __uint_as_float(i | lig.type[i] << 16));
                S.start[threadIdx.x] = c.y;
                S.G[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            unsigned NP;
            const unsigned ex = block_scan(cnt, S.scan, NP);
            if (threadIdx.x < ns) S.end[threadIdx.x] = ex + cnt;
            if (lane == 0) S.head[wave] = make_float4(0.f, 0.f, 0.f, 0.f);
            __syncthreads();

            // this wave's range of pairs and the slot it starts in
            const unsigned per = (NP + WAVES - 1) / WAVES;
            const unsigned rs = min(NP, wave * per), re = min(NP, rs + per);
            unsigned b0 = 0;
            {
                unsigned lo = 0, hi = ns;  // first slot whose end exceeds rs
                while (lo < hi) {
                    const unsigned mid = (lo + hi) >> 1;
                    if (S.end[mid] > rs) hi = mid; else lo = mid + 1;
                }
                b0 = lo;
            }
            const unsigned firstSlot = b0;
            const bool headSplit = rs < re && rs > (b0 ? S.end[b0 - 1] : 0u);
            for (unsigned B = rs; B < re; B += kWave) {
                const unsigned p = B + lane;
                const bool valid = p < re;
                // Slot of every lane: lane j holds the end of slot b0 + j (ends
                // increase strictly, the first exceeds B); the lane's slot is b0 +
                // the number of those ends at or below its pair (ends_up_to).
                const unsigned e = b0 + lane < ns ? S.end[b0 + lane] : 0xFFFFFFFFu;
                unsigned long long starts;
                const unsigned s = min(b0 + ends_up_to(e, B, lane, starts), ns - 1u);
                b0 += (unsigned)__popcll(__ballot(e <= B + kWave));

                float Fx = 0.f, Fy = 0.f, Fz = 0.f;
                if (valid) {
                    const float4 bp = S.p[s];
                    const unsigned meta = __float_as_uint(S.x[s].w);
                    const unsigned lt1 = (meta >> 16) - 1u;
                    const unsigned begin = s ? S.end[s - 1] : 0u;
                    const unsigned j = nl.list[S.start[s] + (p - begin)];
                    const float4 rx = rpos[j];
                    const unsigned rt = rec.type[j];
                    float mt[MM > 0 ? 4 * ((3 * MA + 3) / 4) : 4];
                    float yx = rx.x, yy = rx.y, yz = rx.z;
                    if (MM > 0) {
                        constexpr unsigned Q = (3 * MA + 3) / 4;
                        const float4* row = rmodesT + (size_t)j * rowT;
#pragma unroll
                        for (unsigned q = 0; q < Q; ++q) {
                            const float4 v = q < rowT ? row[q] : make_float4(0.f, 0.f, 0.f, 0.f);
                            mt[4 * q] = v.x; mt[4 * q + 1] = v.y; mt[4 * q + 2] = v.z; mt[4 * q + 3] = v.w;
                        }
                        // the deformation of deform(), same order of operations
#pragma unroll
                        for (unsigned m = 0; m < (unsigned)MA; ++m)
                            if (m < Mr) {
                                const float a = ampR[m];
                                yx = yx + a * mt[3 * m];
                                yy = yy + a * mt[3 * m + 1];
                                yz = yz + a * mt[3 * m + 2];
                            }
                    }
                    const float ddx = bp.x - yx, ddy = bp.y - yy, ddz = bp.z - yz;
                    const float dr2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (dr2 <= nl.P2 && rt - 1u < par.T) {
                        const float dr2_inv = (float)(1.0 / (double)dr2);
                        const float dx = ddx * dr2_inv, dy = ddy * dr2_inv, dz = ddz * dr2_inv;
                        const unsigned kk = par.T * (rt - 1) + lt1;
                        const att_pair_params pp = tab[kk];
                        float vx, vy, vz, ve;
                        lj(dr2, dr2_inv, dx, dy, dz, pp, par.shape, Fx, Fy, Fz, ve);
                        float e2 = ve;
                        lj(nl.P2, nl.P2_inv, 0.f, 0.f, 0.f, pp, par.shape, vx, vy, vz, ve);
                        e2 -= ve;
                        if (bp.w != 0.f) {
                            const float qlr = bp.w * rx.w * par.ffelec;
                            if (fabsf(qlr) > 0.001f) {
                                coulomb(dr2_inv, dx, dy, dz, qlr, par.dielec, vx, vy, vz, ve);
                                e2 += ve;
                                Fx += vx; Fy += vy; Fz += vz;
                                coulomb(nl.P2_inv, 0.f, 0.f, 0.f, qlr, par.dielec, vx, vy, vz, ve);
                                e2 -= ve;
                            }
                        }
                        acc[0] += e2;
                        if (MM > 0) {
#pragma unroll
                            for (unsigned m = 0; m < (unsigned)MA; ++m)
                                if (m < Mr) gR[m] += Fx * mt[3 * m] + Fy * mt[3 * m + 1] + Fz * mt[3 * m + 2];
                        }
                        if (pg.s) {
                            float drc, dac, dem, drcP, dacP, demP;
                            lj_dparam(dr2, dr2_inv, pp, par.shape, drc, dac, dem);
                            lj_dparam(nl.P2, nl.P2_inv, pp, par.shape, drcP, dacP, demP);
                            atomicAdd(&pg.s[3 * kk], pg.w * (drc - drcP));
                            atomicAdd(&pg.s[3 * kk + 1], pg.w * (dac - dacP));
                            if (dem != demP) atomicAdd(&pg.s[3 * kk + 2], pg.w * (dem - demP));
                        }
                    }
                }
                // sum of the forces of each run of lanes sharing a slot, in its first lane
                const unsigned seg = valid ? s : 0xFFFFFFFFu;
                seg_run_sums(Fx, Fy, Fz, seg, lane);
                // the first lane of a run: lane 0, or a slot starts at the lane
                if (valid && (lane == 0 || ((starts >> lane) & 1ull))) {
                    if (headSplit && seg == firstSlot) {
                        float4 h = S.head[wave];
                        h.x += Fx; h.y += Fy; h.z += Fz;
                        S.head[wave] = h;
                    } else {
                        float4 g = S.G[seg];
                        g.x += Fx; g.y += Fy; g.z += Fz;
                        S.G[seg] = g;
                    }
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                // shares of beads split between waves, in wave order
                for (unsigned w = 1; w < (unsigned)WAVES; ++w) {
                    const unsigned ws = min(NP, w * per);
                    if (ws >= NP) break;
                    unsigned lo = 0, hi = ns;
                    while (lo < hi) {
                        const unsigned mid = (lo + hi) >> 1;
                        if (S.end[mid] > ws) hi = mid; else lo = mid + 1;
                    }
                    if (ws > (lo ? S.end[lo - 1] : 0u)) {
                        float4 g = S.G[lo];
                        const float4 h = S.head[w];
                        g.x += h.x; g.y += h.y; g.z += h.z;
                        S.G[lo] = g;
                    }
                }
            }
            __syncthreads();
            // chain rule for the neighbour-list force of each bead of the round
            if (threadIdx.x < ns) {
                const float4 g = S.G[threadIdx.x], xs = S.x[threadIdx.x];
                add_grad(__float_as_uint(xs.w) & 0xFFFFu, xs.x, xs.y, xs.z, -g.x, -g.y, -g.z);
            }
            __syncthreads();  // slot arrays are reused by the next round
        }

        // Reduction and epilogue.
#pragma unroll
        for (int k = 0; k < 13; ++k) acc[k] = wave_sum(acc[k]);
#pragma unroll
        for (int m = 0; m < MA; ++m) { gR[m] = wave_sum(gR[m]); gL[m] = wave_sum(gL[m]); }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 13; ++k) S.red[wave][k] = acc[k];
#pragma unroll
            for (int m = 0; m < ATT_MODES_MAX; ++m) {
                S.red[wave][13 + m] = m < MA ? gR[m < MA ? m : 0] : 0.f;
                S.red[wave][13 + ATT_MODES_MAX + m] = m < MA ? gL[m < MA ? m : 0] : 0.f;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            float v[13 + 2 * ATT_MODES_MAX];
#pragma unroll
            for (int k = 0; k < 13 + 2 * ATT_MODES_MAX; ++k) {
                float s = S.red[0][k];
                for (int ww = 1; ww < WAVES; ++ww) s += S.red[ww][k];
                v[k] = s;
            }
            float da[3];
            torque_reduce(t[0], t[1], t[2], t[3], t[4], t[5], v + 4, da);
            float* o = P.out + (size_t)P.outStride * pose;
            o[0] = v[0];
            o[1] = wp * v[1]; o[2] = wp * v[2]; o[3] = wp * v[3];
            o[4] = wp * da[0]; o[5] = wp * da[1]; o[6] = wp * da[2];
            if (P.outStride >= 27)
#pragma unroll
                for (unsigned m = 0; m < ATT_MODES_MAX; ++m) {
                    o[7 + m] = m < Mr ? wp * v[13 + m] : 0.f;
                    o[17 + m] = m < Ml ? wp * v[13 + ATT_MODES_MAX + m] : 0.f;
                }
        }
    }
}

// Occupancy per mode capacity: the 10-mode kernels run at 3 waves per SIMD
// (168 VGPRs).  At 4 their per-lane mode amplitudes and gradients spill twice
// the scratch (252-332 against 84-168 B per lane); at 3 a 4000/1000-bead
// pair with 10 + 10 modes scores 10 % faster and the 64-pair batched replay
// takes 2.3 % less time.  The 0- and 5-mode kernels stay at 4: at 3 (and, for
// rigid bodies, at 5) they measured slower, e.g. a 1000/300-bead rigid pair
// 3.58 and 4.44 ms against 3.05 ms for 100k poses.  Occupancy changes no
// arithmetic: results are the same bits.
template <int MM>
constexpr int adjoint_waves_per_simd() { return MM >= 10 ? 3 : 4; }

template <int MM, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(adjoint_waves_per_simd<MM>()))) void k_adjoint(
    PairDesc P, float* __restrict__ pgPartial) {
    extern __shared__ att_pair_params s_dyn[];
    __shared__ PairLDS<WAVES> S;
    char* dyn = reinterpret_cast<char*>(s_dyn);
    const att_pair_params* tab = stage_table(P.par, s_dyn);
    float* s_pg = pgPartial ? reinterpret_cast<float*>(dyn + adjoint_pg_offset(P.tabSlots, P.lig.n)) : nullptr;
    const unsigned npg = 3 * P.par.T * P.par.T;
    if (s_pg)
        for (unsigned k = threadIdx.x; k < npg; k += blockDim.x) s_pg[k] = 0.f;
    const unsigned n = P.nDev ? min(P.nDev[0], P.n) : P.n;
    // With a spatial order (and at least 8 workgroups), workgroup b (on XCD
    // b % 8) walks the (b % 8)-th eighth of it; otherwise the poses in input
    // order, strided by the grid.
    unsigned first = blockIdx.x, end = n, step = gridDim.x;
    const bool ordered = P.order && gridDim.x >= 8;
    if (ordered) {
        const unsigned x = blockIdx.x & 7u;
        step = (gridDim.x - x + 7u) >> 3;
        first = (unsigned)(((unsigned long long)n * x) >> 3) + (blockIdx.x >> 3);
        end = (unsigned)(((unsigned long long)n * (x + 1)) >> 3);
    }
    for (unsigned i = first; i < end; i += step)
        adjoint_pose<MM, WAVES>(P, P.order ? P.order[i] : i, S, dyn, tab, s_pg);
    if (s_pg) {
        __syncthreads();
        for (unsigned k = threadIdx.x; k < npg; k += blockDim.x)
            pgPartial[(size_t)blockIdx.x * npg + k] = s_pg[k];
    }
}

// Batched adjoint over many pairs: persistent workgroups take tiles (pair,
// first pose, count) from eight queues, one per XCD: workgroup b (on XCD b % 8)
// drains queue b % 8, then helps with the others.  Queue x holds the x-th
// eighth of every pair's poses in that pair's spatial order, so one XCD's L2
// sees neighbouring poses.  The pair's descriptor is read with scalar loads
// and its table staged in LDS when the pair changes.  Occupancy as k_adjoint.
template <int MM, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(adjoint_waves_per_simd<MM>()))) void k_adjoint_batch(
    const PairDesc* __restrict__ pairs, const BatchTile* __restrict__ tiles,
    const unsigned* __restrict__ queueStart, unsigned* __restrict__ counters) {
    extern __shared__ att_pair_params s_dyn[];
    __shared__ PairLDS<WAVES> S;
    __shared__ unsigned s_tile;
    char* dyn = reinterpret_cast<char*>(s_dyn);
    unsigned cur = 0xFFFFFFFFu;
    const att_pair_params* tab = nullptr;
    for (unsigned q = 0; q < kBatchQueues; ++q) {
        const unsigned x = (blockIdx.x + q) % kBatchQueues;
        const unsigned lo = queueStart[x], hi = queueStart[x + 1];
        for (;;) {
            __syncthreads();  // s_tile and the table of the previous tile are no longer read
            if (threadIdx.x == 0) s_tile = lo + atomicAdd(&counters[x], 1u);
            __syncthreads();
            const unsigned t = __builtin_amdgcn_readfirstlane(s_tile);
            if (t >= hi) break;
            const BatchTile tile = tiles[t];
            const PairDesc& P = pairs[tile.pair];
            if (tile.pair != cur) {
                tab = stage_table(P.par, s_dyn);  // made visible by adjoint_pose's first barrier
                cur = tile.pair;
            }
            for (unsigned i = tile.first; i < tile.first + tile.count; ++i)
                adjoint_pose<MM, WAVES>(P, P.order ? P.order[i] : i, S, dyn, tab, nullptr);
        }
    }
}

// The wave primitives of the pair loop on given inputs (one wave, diagnostics).
__global__ __launch_bounds__(64) void k_debug_wave_ops(const float* __restrict__ F,
                                                       const unsigned* __restrict__ seg,
                                                       const unsigned* __restrict__ ends, unsigned B,
                                                       float* __restrict__ sums,
                                                       unsigned* __restrict__ counts,
                                                       unsigned long long* __restrict__ starts) {
    const unsigned lane = threadIdx.x;
    float Fx = F[lane], Fy = F[64 + lane], Fz = F[128 + lane];
    seg_run_sums(Fx, Fy, Fz, seg[lane], lane);
    sums[lane] = Fx; sums[64 + lane] = Fy; sums[128 + lane] = Fz;
    unsigned long long m;
    counts[lane] = ends_up_to(ends[lane], B, lane, m);
    if (lane == 0) starts[0] = m;
}

__global__ void k_reset_counters(unsigned* counters, unsigned n) {
    if (threadIdx.x < n) counters[threadIdx.x] = 0u;
}

__global__ void k_fill_zero(float* p, unsigned n) {
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.f;
}

hipError_t launch_fill_zero(float* p, unsigned n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero, dim3(std::min(1024u, (n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// Parameter-gradient rows of the workgroups, summed in a fixed order: group
// g (blockIdx.y) adds rows [g*rows/G, (g+1)*rows/G) in order, in double.
__global__ void k_pgrad_groups(const float* __restrict__ partial, unsigned rows, unsigned cols,
                               double* __restrict__ groups) {
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const unsigned g = blockIdx.y;
    const unsigned lo = (unsigned)((unsigned long long)rows * g / kPgradGroups);
    const unsigned hi = (unsigned)((unsigned long long)rows * (g + 1) / kPgradGroups);
    // the loads of 8 rows are issued together
    double s = 0.0;
    unsigned r = lo;
    for (; r + 8 <= hi; r += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = partial[(size_t)(r + k) * cols + c];
#pragma unroll
        for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; r < hi; ++r) s += partial[(size_t)r * cols + c];
    groups[(size_t)g * cols + c] = s;
}

// The group sums added in group order.
__global__ void k_pgrad_reduce(const double* __restrict__ groups, unsigned cols,
                               float* __restrict__ pg) {
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    double s = 0.0;
#pragma unroll 8
    for (unsigned g = 0; g < kPgradGroups; ++g) s += groups[(size_t)g * cols + c];
    pg[c] = (float)s;
}

// Expands basis grids [b][z][y][x] (nb floats per point) into per-voxel rows
// [b][corner] of 8 floats, corners in trilinear order.
__global__ __launch_bounds__(256) void k_basis_octets(const float* __restrict__ b,
                                                      float4* __restrict__ oct, unsigned nx,
                                                      unsigned ny, unsigned nz, unsigned nb) {
    const size_t nvo = (size_t)(nx - 1) * (ny - 1) * (nz - 1);
    const size_t nv = (size_t)nx * ny * nz, nxy = (size_t)nx * ny;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nvo * nb;
         t += (size_t)gridDim.x * blockDim.x) {
        const size_t o = t / nb;
        const unsigned k = (unsigned)(t % nb);
        size_t r = o;
        const unsigned ix = r % (nx - 1); r /= (nx - 1);
        const unsigned iy = r % (ny - 1); r /= (ny - 1);
        const unsigned iz = (unsigned)r;
        const float* s = b + (size_t)k * nv + nxy * iz + (size_t)nx * iy + ix;
        float4* dst = oct + (o * nb + k) * 2;
        dst[0] = make_float4(s[0], s[1], s[nx], s[nxy]);
        dst[1] = make_float4(s[nxy + 1], s[nxy + nx], s[nx + 1], s[nxy + nx + 1]);
    }
}

size_t adjoint_lds_bytes(const DProt& rec, const DProt& lig, const DPar& par, bool pgrad,
                         unsigned* tabSlots) {
    (void)rec;
    const unsigned slots = par.T <= kMaxTypesLDS ? par.T * par.T : 0;
    if (tabSlots) *tabSlots = slots;
    return adjoint_pg_offset(slots, lig.n) + (pgrad ? sizeof(float) * 3 * (size_t)par.T * par.T : 0);
}

size_t adjoint_static_lds(int waves) {
    switch (waves) {
        case 1: return sizeof(PairLDS<1>) + 16;
        case 2: return sizeof(PairLDS<2>) + 16;
        default: return sizeof(PairLDS<4>) + 16;
    }
}

namespace {
template <int MM>
const void* adjoint_kernel(int waves, bool batch) {
    switch (waves) {
        case 1: return batch ? (const void*)k_adjoint_batch<MM, 1> : (const void*)k_adjoint<MM, 1>;
        case 2: return batch ? (const void*)k_adjoint_batch<MM, 2> : (const void*)k_adjoint<MM, 2>;
        default: return batch ? (const void*)k_adjoint_batch<MM, 4> : (const void*)k_adjoint<MM, 4>;
    }
}
const void* adjoint_kernel_for(unsigned mm, int waves, bool batch) {
    if (mm == 0) return adjoint_kernel<0>(waves, batch);
    if (mm <= 5) return adjoint_kernel<5>(waves, batch);
    return adjoint_kernel<ATT_MODES_MAX>(waves, batch);
}
}  // namespace

unsigned adjoint_blocks_per_cu(unsigned mm, int waves, bool batch, size_t lds) {
    int b = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &b, adjoint_kernel_for(mm, waves, batch), 64 * waves, lds);
    return e == hipSuccess && b > 0 ? (unsigned)b : 1u;
}

hipError_t launch_adjoint(const PairDesc& P, bool modes, float* pgPartial, unsigned blocks,
                          int waves, size_t lds, hipStream_t s) {
    if (P.n == 0) return hipSuccess;
    const unsigned mm = modes ? (P.rec.M > P.lig.M ? P.rec.M : P.lig.M) : 0;
    void* args[] = {const_cast<PairDesc*>(&P), &pgPartial};
    return hipLaunchKernel(adjoint_kernel_for(mm, waves, false), dim3(blocks), dim3(64 * waves),
                           args, lds, s);
}

hipError_t launch_debug_wave_ops(const float* F, const unsigned* seg, const unsigned* ends,
                                 unsigned B, float* sums, unsigned* counts,
                                 unsigned long long* starts, hipStream_t s) {
    hipLaunchKernelGGL(k_debug_wave_ops, dim3(1), dim3(64), 0, s, F, seg, ends, B, sums, counts,
                       starts);
    return hipGetLastError();
}

hipError_t launch_adjoint_batch(unsigned mm, int waves, const PairDesc* pairs, const BatchTile* tiles,
                                const unsigned* queueStart, unsigned ntiles, unsigned* counters,
                                unsigned blocks, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, counters, kBatchQueues);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ntiles == 0) return e;
    void* args[] = {&pairs, &tiles, &queueStart, &counters};
    return hipLaunchKernel(adjoint_kernel_for(mm, waves, true), dim3(blocks), dim3(64 * waves), args,
                           lds, s);
}

hipError_t launch_pgrad_reduce(const float* partial, unsigned rows, unsigned cols, double* groups,
                               float* pg, hipStream_t s) {
    hipLaunchKernelGGL(k_pgrad_groups, dim3((cols + 255) / 256, kPgradGroups), dim3(256), 0, s,
                       partial, rows, cols, groups);
    hipLaunchKernelGGL(k_pgrad_reduce, dim3((cols + 255) / 256), dim3(256), 0, s,
                       (const double*)groups, cols, pg);
    return hipGetLastError();
}

hipError_t launch_basis_octets(const float* b, float4* oct, unsigned nx, unsigned ny, unsigned nz,
                               unsigned nb, hipStream_t s) {
    const size_t total = (size_t)(nx - 1) * (ny - 1) * (nz - 1) * nb;
    const size_t blocks = std::min<size_t>((total + 255) / 256, 256 * 256);
    hipLaunchKernelGGL(k_basis_octets, dim3((unsigned)blocks), dim3(256), 0, s, b, oct, nx, ny, nz,
                       nb);
    return hipGetLastError();
}

}  // namespace att
