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
// k_adjoint: persistent, one workgroup of 1, 2 or 4 waves per pose at a time;
// per-pose values go through wave butterflies, parameter derivatives
// accumulate in LDS (one slice per wave) across the workgroup's poses and are
// written once per workgroup, then k_pgrad_groups and k_pgrad_reduce add the
// per-workgroup rows in a fixed order.
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
    float* s;                     // this wave's LDS slice, T*T*3, or nullptr when not requested
    PairTab tab;                  // the pair table
    unsigned T;
    float w;                      // upstream weight of the current pose
    unsigned* key;                // kPgRecords: per (slot, bead) grid-cell record
    float4* pos4;                 //   {x, y, z, w}, or
    float3* pos3;                 //   {x, y, z} when the call has no weights
    unsigned cellsIn;             //   key = (cell, inner cells first) * T + type - 1
    unsigned invalid;             //   key of a bead without a grid term
};



// Basis channels read per step of the parameter-gradient loop: the loads of
// kBasisStep channels are issued together (64 B each) and their LDS updates
// follow.  Timed with 12 types: steps of 2 and 3 channels 1.1% faster than 4
// and 6 (same results).
constexpr unsigned kBasisStep = 2;

// Grid term of one ligand bead on one grid: energy (the forward's per-bead
// value), gradient, and the basis-grid parameter derivatives.
template <int PG>
__device__ __forceinline__ float grid_adj_one(const DGrid& g, const DBasis& b, float px, float py,
                                              float pz, unsigned mtype, unsigned type, float charge,
                                              float3& G, const PGrad& pg, unsigned cellBase,
                                              size_t rec) {
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
    // The grid of mapped type mtype holds the bead type's column of the pair
    // table (the grid alphabet maps type -> mtype), so the derivatives go to
    // the entries (a, type).
    if (PG == kPgRecords) {
        if (type - 1u < pg.T) {
            const unsigned o = (iz * (g.ny - 1) + iy) * (g.nx - 1) + ix;
            pg.key[rec] = (cellBase + o) * pg.T + (type - 1);
            if (pg.pos3)
                pg.pos3[rec] = make_float3(px, py, pz);
            else
                pg.pos4[rec] = make_float4(px, py, pz, pg.w);
        } else {
            pg.key[rec] = pg.invalid;
        }
    }
    if (PG == kPgFused && b.oct && type - 1u < pg.T) {
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
                const unsigned k = (a0 + j) * pg.T + (type - 1);
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
template <int PG>
__device__ __forceinline__ float grid_adj(const DGrid& inner, const DGrid& outer,
                                          const DBasis& bin, const DBasis& bout, const DPar& par,
                                          float px, float py, float pz, unsigned mtype,
                                          unsigned type, float charge, float3& G, const PGrad& pg,
                                          size_t rec) {
    G = make_float3(0.f, 0.f, 0.f);
    if (mtype == 0) {
        if (PG == kPgRecords) pg.key[rec] = pg.invalid;
        return 0.f;
    }
    const bool in = !grid_out(inner, px, py, pz);
    if (!in && grid_out(outer, px, py, pz)) {
        if (PG == kPgRecords) pg.key[rec] = pg.invalid;
        return 0.f;
    }
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
    return grid_adj_one<PG>(g, b, px, py, pz, mtype, type, charge, G, pg, in ? 0u : pg.cellsIn, rec);
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

// The neighbour-list pairs of one ligand bead at p (charge q, type lt) over
// its list d = {numEl, idx}, in list order:
//  eAcc  the pair energies folded as NLPotForce folds them (+LJ(r), -LJ(P),
//        +C(r), -C(P), model/neighborlist.h:66-158): the reference's bits;
//  gAcc  the sum of the pair forces (LJ and Coulomb at r; the plateau terms
//        are constant in position);
//  gR    += force . (receptor mode vector) per pair (dE/d receptor amplitude
//        through the receptor's deformation);
//  the parameter derivatives into this wave's LDS slice.
// Receptor beads are deformed from their padded mode row (one cache line).
template <int MM, int PG>
__device__ __forceinline__ void adj_walk(const PairDesc& P, const PairTab& tab,
                                         const PGrad& pg, float px, float py, float pz, float q,
                                         unsigned lt, uint2 d, const float* ampR, float* gR,
                                         float& eAcc, float3& gAcc, unsigned& inside) {
    constexpr int MA = MM > 0 ? MM : 1;
    const DNL& nl = P.nl;
    const DPar& par = P.par;
    const float4* __restrict__ rpos = P.rec.pos;
    const float4* __restrict__ rmodesT = P.rec.modesT;
    const unsigned Mr = MM > 0 ? P.rec.M : 0, rowT = P.rec.rowT;
    const unsigned lt1 = lt - 1;
    const uint32_t* lst = nl.list + d.y;
    eAcc = 0.f;
    gAcc = make_float3(0.f, 0.f, 0.f);
    unsigned nxt = d.x ? lst[0] : 0u;
    for (unsigned jj = 0; jj < d.x; ++jj) {
        const unsigned j = nxt;
        if (jj + 1 < d.x) nxt = lst[jj + 1];  // next entry in flight
        const float4 rx = rpos[j];
        const unsigned rt = P.rec.type[j];
        float mt[MM > 0 ? 4 * ((3 * MA + 3) / 4) : 4];
        float yx = rx.x, yy = rx.y, yz = rx.z;
        if (MM > 0) {
            constexpr unsigned Q = (3 * MA + 3) / 4;
            const float4* row = rmodesT + (size_t)j * rowT;
#pragma unroll
            for (unsigned qq = 0; qq < Q; ++qq) {
                const float4 v = qq < rowT ? row[qq] : make_float4(0.f, 0.f, 0.f, 0.f);
                mt[4 * qq] = v.x; mt[4 * qq + 1] = v.y; mt[4 * qq + 2] = v.z; mt[4 * qq + 3] = v.w;
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
        const float ddx = px - yx, ddy = py - yy, ddz = pz - yz;
        const float dr2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (!(dr2 <= nl.P2) || rt - 1u >= par.T) continue;
        ++inside;
        const float dr2_inv = (float)(1.0 / (double)dr2);
        const float dx = ddx * dr2_inv, dy = ddy * dr2_inv, dz = ddz * dr2_inv;
        const unsigned kk = par.T * (rt - 1) + lt1;
        const att_pair_params pp = tab[kk];
        float Fx, Fy, Fz, ve, vx, vy, vz;
        lj(dr2, dr2_inv, dx, dy, dz, pp, par.shape, Fx, Fy, Fz, ve);
        eAcc += ve;
        lj(nl.P2, nl.P2_inv, 0.f, 0.f, 0.f, pp, par.shape, vx, vy, vz, ve);
        eAcc -= ve;
        if (q != 0.f) {
            const float qlr = q * rx.w * par.ffelec;
            if (fabsf(qlr) > 0.001f) {
                coulomb(dr2_inv, dx, dy, dz, qlr, par.dielec, vx, vy, vz, ve);
                eAcc += ve;
                Fx += vx; Fy += vy; Fz += vz;
                coulomb(nl.P2_inv, 0.f, 0.f, 0.f, qlr, par.dielec, vx, vy, vz, ve);
                eAcc -= ve;
            }
        }
        gAcc.x += Fx; gAcc.y += Fy; gAcc.z += Fz;
        if (MM > 0) {
#pragma unroll
            for (unsigned m = 0; m < (unsigned)MA; ++m)
                if (m < Mr) gR[m] += Fx * mt[3 * m] + Fy * mt[3 * m + 1] + Fz * mt[3 * m + 2];
        }
        if (PG != kPgNone) {
            float drc, dac, dem, drcP, dacP, demP;
            lj_dparam(dr2, dr2_inv, pp, par.shape, drc, dac, dem);
            lj_dparam(nl.P2, nl.P2_inv, pp, par.shape, drcP, dacP, demP);
            atomicAdd(&pg.s[3 * kk], pg.w * (drc - drcP));
            atomicAdd(&pg.s[3 * kk + 1], pg.w * (dac - dacP));
            if (dem != demP) atomicAdd(&pg.s[3 * kk + 2], pg.w * (dem - demP));
        }
    }
}

// The workgroup-shared state of one pose (static LDS), WAVES waves per pose:
// the ligand is taken in super-chunks of 256 beads per wave (26 B of LDS per
// bead: four workgroups of two waves, or eight of one, per CU).
template <int WAVES>
struct PoseLDS {
    static constexpr unsigned kSuper = 256u * WAVES;
    float4 g[kSuper];     // per bead: dE/dp (grid, then minus the pair forces); energy
    uint2 nlc[kSuper];    // neighbour-list {count, start}
    unsigned short q[kSuper];
    unsigned cc[queue_counters(kSuper)];
    float red[WAVES][ATT_MODES_MAX];
    unsigned cnt[WAVES][2];
    float e;
};

// Adjoint of one pose by one workgroup of WAVES waves, the ligand in
// super-chunks of equal size, at most PoseLDS<WAVES>::kSuper beads (super_chunk):
//  P1  every bead: transform, grid term with its exact gradient (and the
//      basis-grid parameter derivatives), neighbour-list cell;
//  P2  beads with a list, longest lists first (class_queue);
//  P3  one bead per lane walks its list (adj_walk); the bead's energy becomes
//      grid + pair fold, the reference's per-bead value;
//  P4  the energy summed in bead order (reducePotForce's order, so the pose
//      energy is the reference's bits), and the chain rule through p = R x~ + t
//      by wave 0, bead k on lane k % 64, lane sums reduced by a butterfly.
// Every sum has a fixed order: results do not depend on timing.
template <int MM, int WAVES, int PG>
__device__ __forceinline__ void adjoint_pose(const PairDesc& P, unsigned pose, PoseLDS<WAVES>& S,
                                             const PairTab& tab, float* s_pg,
                                             size_t recBase) {
    const DProt& lig = P.lig;
    const DPar& par = P.par;
    const unsigned tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    constexpr unsigned nthr = 64u * WAVES;
    constexpr unsigned eThread = WAVES > 1 ? 64u : 0u;  // sums the energy
    constexpr int MA = MM > 0 ? MM : 1;
    const unsigned Ml = MM > 0 ? lig.M : 0, Mr = MM > 0 ? P.rec.M : 0;
    __syncthreads();  // LDS of the previous pose is no longer read
    const float* d = P.dofs + (size_t)P.dofStride * pose;
    const float tx = d[0], ty = d[1], tz = d[2];
    const float wp = P.wts ? P.wts[pose] : 1.f;
    const PGrad pg{PG != kPgNone ? s_pg + wave * 3 * par.T * par.T : nullptr, tab, par.T, wp,
                   P.recKey, P.recPos4, P.recPos3, P.recCellsIn, P.recInvalid};
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
    // wave 0: dE/dt (3), sum G (x) x~ (9), ligand modes; all waves: receptor modes
    float acc[12], gR[MA], gL[MA];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0.f;
#pragma unroll
    for (int m = 0; m < MA; ++m) { gR[m] = 0.f; gL[m] = 0.f; }
    float esum = 0.f;
    unsigned listed = 0, inside = 0;  // this lane's pair counts (P.counts != nullptr)

    auto deformed = [&](unsigned i, float4& x) __attribute__((always_inline)) {
        x = lig.pos[i];
        BeadModes<MM> md;
        load_modes(lig, i, md);
        return deform(md, Ml, x.x, x.y, x.z, ampL);
    };
    auto place = [&](unsigned i, float4& x) __attribute__((always_inline)) {
        const V3 xt = deformed(i, x);
        V3 p = matvec(R, xt.x, xt.y, xt.z);
        p.x = p.x + tx; p.y = p.y + ty; p.z = p.z + tz;
        return p;
    };

    const unsigned span = super_chunk(lig.n, PoseLDS<WAVES>::kSuper);
    for (unsigned sc = 0; sc < lig.n; sc += span) {
        const unsigned m = min(span, lig.n - sc);
        // P1
        for (unsigned k = tid; k < m; k += nthr) {
            float4 x;
            const V3 p = place(sc + k, x);
            float3 G;
            const float eg = grid_adj<PG>(P.inner, P.outer, P.bin, P.bout, par, p.x, p.y, p.z,
                                          lig.mtype[sc + k], lig.type[sc + k], x.w, G, pg,
                                          recBase + sc + k);
            S.g[k] = make_float4(G.x, G.y, G.z, eg);
            S.nlc[k] = nl_cell(P.nl, p.x, p.y, p.z, lig.type[sc + k]);
        }
        __syncthreads();
        // P2
        const unsigned qn = class_queue([&](unsigned k) { return S.nlc[k].x; }, m, S.q, S.cc, tid,
                                        nthr, [] { __syncthreads(); });
        // P3
        for (unsigned q0 = 0; q0 < qn; q0 += nthr) {
            if (q0 + tid < qn) {
                const unsigned k = S.q[q0 + tid];
                float4 x;
                const V3 p = place(sc + k, x);
                float eAcc;
                float3 gAcc;
                adj_walk<MM, PG>(P, tab, pg, p.x, p.y, p.z, x.w, lig.type[sc + k], S.nlc[k], ampR, gR,
                                 eAcc, gAcc, inside);
                listed += S.nlc[k].x;
                float4 g = S.g[k];
                g.x = g.x - gAcc.x; g.y = g.y - gAcc.y; g.z = g.z - gAcc.z;
                g.w = g.w + eAcc;  // outLig_E[i] += eAcc
                S.g[k] = g;
            }
        }
        __syncthreads();
        // P4
        if (tid == eThread) {
            unsigned k = 0;
            for (; k + 8 <= m; k += 8) {  // eight loads in flight, the sum sequential
                float e8[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) e8[u] = S.g[k + u].w;
#pragma unroll
                for (int u = 0; u < 8; ++u) esum += e8[u];
            }
            for (; k < m; ++k) esum += S.g[k].w;
        }
        if (wave == 0) {
            for (unsigned k = lane; k < m; k += kWave) {
                const float4 x = lig.pos[sc + k];
                BeadModes<MM> md;  // loaded once for the deformation and the mode gradient
                load_modes(lig, sc + k, md);
                const V3 xt = deform(md, Ml, x.x, x.y, x.z, ampL);
                const float4 G = S.g[k];
                acc[0] += G.x; acc[1] += G.y; acc[2] += G.z;
                acc[3] += G.x * xt.x; acc[4] += G.x * xt.y; acc[5] += G.x * xt.z;
                acc[6] += G.y * xt.x; acc[7] += G.y * xt.y; acc[8] += G.y * xt.z;
                acc[9] += G.z * xt.x; acc[10] += G.z * xt.y; acc[11] += G.z * xt.z;
                if (MM > 0) {
                    // dE/da_m = G . R m = (R^T G) . m
                    const float gx = R[0] * G.x + R[3] * G.y + R[6] * G.z;
                    const float gy = R[1] * G.x + R[4] * G.y + R[7] * G.z;
                    const float gz = R[2] * G.x + R[5] * G.y + R[8] * G.z;
                    mode_dots(md, Ml, -gx, -gy, -gz, gL);  // mode_dots subtracts
                }
            }
        }
        __syncthreads();  // the super-chunk's arrays are reused
    }

    // Reductions and epilogue.
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] = wave_sum(acc[k]);
#pragma unroll
        for (int m = 0; m < MA; ++m) gL[m] = wave_sum(gL[m]);
    }
#pragma unroll
    for (int m = 0; m < MA; ++m) gR[m] = wave_sum(gR[m]);
    if (lane == 0) {
#pragma unroll
        for (int m = 0; m < ATT_MODES_MAX; ++m) S.red[wave][m] = m < MA ? gR[m < MA ? m : 0] : 0.f;
    }
    if (tid == eThread) S.e = esum;
    if (P.counts) {
        listed = wave_sum_u(listed);
        inside = wave_sum_u(inside);
        if (lane == 0) { S.cnt[wave][0] = listed; S.cnt[wave][1] = inside; }
    }
    __syncthreads();
    if (P.counts && tid == 0) {
        uint2 c = make_uint2(0u, 0u);
        for (unsigned w = 0; w < (unsigned)WAVES; ++w) { c.x += S.cnt[w][0]; c.y += S.cnt[w][1]; }
        P.counts[pose] = c;
    }
    if (tid == 0) {
        float da[3];
        torque_reduce(t[0], t[1], t[2], t[3], t[4], t[5], acc + 3, da);
        float* o = P.out + (size_t)P.outStride * pose;
        o[0] = S.e;
        o[1] = wp * acc[0]; o[2] = wp * acc[1]; o[3] = wp * acc[2];
        o[4] = wp * da[0]; o[5] = wp * da[1]; o[6] = wp * da[2];
        if (P.outStride >= 27)
#pragma unroll
            for (unsigned m = 0; m < ATT_MODES_MAX; ++m) {
                float s = S.red[0][m];
                for (unsigned w = 1; w < (unsigned)WAVES; ++w) s += S.red[w][m];
                o[7 + m] = m < Mr ? wp * s : 0.f;
                o[17 + m] = m < Ml ? wp * gL[m < MA ? m : 0] : 0.f;
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

template <int MM, int WAVES, int PG>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(adjoint_waves_per_simd<MM>()))) void k_adjoint(
    PairDesc P, float* __restrict__ pgPartial) {
    extern __shared__ att_pair_params s_dyn[];
    __shared__ PoseLDS<WAVES> S;
    char* dyn = reinterpret_cast<char*>(s_dyn);
    const PairTab tab = stage_table(P.par, s_dyn);
    // Parameter derivatives: one LDS slice per wave, so that every slice is
    // updated by one wave only, in program order; the slices are added in
    // wave order at the end.  Runs therefore give the same bits.
    float* s_pg = PG != kPgNone ? reinterpret_cast<float*>(dyn + adjoint_pg_offset(P.tabSlots)) : nullptr;
    const unsigned npg = 3 * P.par.T * P.par.T;
    if (PG != kPgNone)
        for (unsigned k = threadIdx.x; k < npg * WAVES; k += blockDim.x) s_pg[k] = 0.f;
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
    // grid-cell records of slot i (kPgRecords) start at i * ligand beads
    for (unsigned i = first; i < end; i += step)
        adjoint_pose<MM, WAVES, PG>(P, P.order ? P.order[i] : i, S, tab, s_pg, (size_t)i * P.lig.n);
    if (PG != kPgNone) {
        __syncthreads();
        for (unsigned k = threadIdx.x; k < npg; k += blockDim.x) {
            float v = s_pg[k];
            for (unsigned w = 1; w < (unsigned)WAVES; ++w) v += s_pg[w * npg + k];
            pgPartial[(size_t)blockIdx.x * npg + k] = v;
        }
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
    __shared__ PoseLDS<WAVES> S;
    __shared__ unsigned s_tile;
    unsigned cur = 0xFFFFFFFFu;
    PairTab tab{nullptr, false};
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
                adjoint_pose<MM, WAVES, kPgNone>(P, P.order ? P.order[i] : i, S, tab, nullptr, 0);
        }
    }
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
                         int waves, unsigned* tabSlots) {
    (void)rec;
    (void)lig;
    const unsigned slots = par.T <= kMaxTypesLDS ? par.T * par.T : 0;
    if (tabSlots) *tabSlots = slots;
    return adjoint_pg_offset(slots) +
           (pgrad ? sizeof(float) * 3 * (size_t)par.T * par.T * (size_t)waves : 0);
}

size_t adjoint_static_lds(int waves) {
    switch (waves) {
        case 1: return sizeof(PoseLDS<1>) + 16;
        case 2: return sizeof(PoseLDS<2>) + 16;
        default: return sizeof(PoseLDS<4>) + 16;
    }
}

namespace {
template <int MM, int PG>
const void* adjoint_kernel(int waves, bool batch) {
    switch (waves) {
        case 1: return batch ? (const void*)k_adjoint_batch<MM, 1> : (const void*)k_adjoint<MM, 1, PG>;
        case 2: return batch ? (const void*)k_adjoint_batch<MM, 2> : (const void*)k_adjoint<MM, 2, PG>;
        default: return batch ? (const void*)k_adjoint_batch<MM, 4> : (const void*)k_adjoint<MM, 4, PG>;
    }
}
template <int PG>
const void* adjoint_kernel_pg(unsigned mm, int waves, bool batch) {
    if (mm == 0) return adjoint_kernel<0, PG>(waves, batch);
    if (mm <= 5) return adjoint_kernel<5, PG>(waves, batch);
    return adjoint_kernel<ATT_MODES_MAX, PG>(waves, batch);
}
const void* adjoint_kernel_for(unsigned mm, int waves, bool batch, int pg = kPgNone) {
    if (batch || pg == kPgNone) return adjoint_kernel_pg<kPgNone>(mm, waves, batch);
    if (pg == kPgFused) return adjoint_kernel_pg<kPgFused>(mm, waves, false);
    return adjoint_kernel_pg<kPgRecords>(mm, waves, false);
}
}  // namespace

unsigned adjoint_blocks_per_cu(unsigned mm, int waves, bool batch, size_t lds, int pg) {
    int b = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &b, adjoint_kernel_for(mm, waves, batch, pg), 64 * waves, lds);
    return e == hipSuccess && b > 0 ? (unsigned)b : 1u;
}

hipError_t launch_adjoint(const PairDesc& P, bool modes, float* pgPartial, unsigned blocks,
                          int waves, size_t lds, hipStream_t s) {
    if (P.n == 0) return hipSuccess;
    const unsigned mm = modes ? (P.rec.M > P.lig.M ? P.rec.M : P.lig.M) : 0;
    const int pg = !pgPartial ? kPgNone : P.recKey ? kPgRecords : kPgFused;
    void* args[] = {const_cast<PairDesc*>(&P), &pgPartial};
    return hipLaunchKernel(adjoint_kernel_for(mm, waves, false, pg), dim3(blocks), dim3(64 * waves),
                           args, lds, s);
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
