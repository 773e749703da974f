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
// k_pgrad_reduce adds the per-workgroup rows in a fixed order.
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
    float* s;      // LDS accumulator, T*T*3, or nullptr when not requested
    unsigned T;
    float w;       // upstream weight of the current pose
};

// Grid term of one ligand bead on one grid: energy (the forward's per-bead
// value), gradient, and the basis-grid parameter derivatives.
__device__ __forceinline__ float grid_adj_one(const DGrid& g, const DBasis& b, const DPar& par,
                                              float px, float py, float pz, unsigned mtype,
                                              float charge, float3& G, const PGrad& pg) {
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
        for (unsigned a = 0; a < pg.T; ++a) {
            const float4 s0 = row[4 * a], s1 = row[4 * a + 1];      // S_a at the 8 corners
            const float4 t0 = row[4 * a + 2], t1 = row[4 * a + 3];  // S6_a
            const float S = (s0.x * w0 + s0.y * w1 + s0.z * w2 + s0.w * w3 + s1.x * w4 +
                             s1.y * w5 + s1.z * w6 + s1.w * w7) * g.vol_inv;
            const float S6 = (t0.x * w0 + t0.y * w1 + t0.z * w2 + t0.w * w3 + t1.x * w4 +
                              t1.y * w5 + t1.z * w6 + t1.w * w7) * g.vol_inv;
            const unsigned k = a * pg.T + (mtype - 1);
            const float iv = (float)par.tab[k].ipon * pg.w;
            atomicAdd(&pg.s[3 * k], iv * S);
            atomicAdd(&pg.s[3 * k + 1], -iv * S6);
        }
    }
    return pot.w;
}

// Grid term of one ligand bead: inner grid, else outer grid, else nothing
// (model/interpolation.h:56-102).
__device__ __forceinline__ float grid_adj(const DGrid& inner, const DGrid& outer,
                                          const DBasis& bin, const DBasis& bout, const DPar& par,
                                          float px, float py, float pz, unsigned mtype,
                                          float charge, float3& G, const PGrad& pg) {
    G = make_float3(0.f, 0.f, 0.f);
    if (mtype == 0) return 0.f;
    if (!grid_out(inner, px, py, pz))
        return grid_adj_one(inner, bin, par, px, py, pz, mtype, charge, G, pg);
    if (!grid_out(outer, px, py, pz))
        return grid_adj_one(outer, bout, par, px, py, pz, mtype, charge, G, pg);
    return 0.f;
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

// Neighbour-list term of one ligand bead for the adjoint: energy (the
// forward's per-bead value), G -= pair force, receptor(j, F) gets the force
// on the receptor side, parameter derivatives into pg.
template <class RecPos, class RecType, class RecCharge, class Reaction>
__device__ __forceinline__ float nl_adj(const DNL& nl, const DPar& par, const att_pair_params* tab,
                                        RecPos recPos, RecType recType, RecCharge recCharge,
                                        Reaction reaction, float px, float py, float pz,
                                        unsigned ligType, float ligCharge, float3& G,
                                        const PGrad& pg) {
    if (ligType == 0) return 0.f;
    if (!((px >= nl.lox && px <= nl.hix) && (py >= nl.loy && py <= nl.hiy) &&
          (pz >= nl.loz && pz <= nl.hiz)))
        return 0.f;
    const int ix = (int)roundf((px - nl.lox) * nl.dVox_inv);
    const int iy = (int)roundf((py - nl.loy) * nl.dVox_inv);
    const int iz = (int)roundf((pz - nl.loz) * nl.dVox_inv);
    const uint2 d = nl.desc[nl.nx * (iz * nl.ny + iy) + ix];
    float e = 0.f;
    const unsigned lt1 = ligType - 1;
    const bool ligCharged = ligCharge != 0.f;
    const uint32_t* lst = nl.list + d.y;
    float drcP = 0.f, dacP = 0.f, demP = 0.f;
    for (unsigned j0 = 0; j0 < d.x; j0 += 4) {
        unsigned idx[4];
        float3 rp[4];
        unsigned rt[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) idx[k] = lst[j0 + k < d.x ? j0 + k : j0];
#pragma unroll
        for (int k = 0; k < 4; ++k) { rp[k] = recPos(idx[k]); rt[k] = recType(idx[k]); }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (j0 + k >= d.x) break;
            const unsigned nIdx = idx[k];
            const float ddx = px - rp[k].x, ddy = py - rp[k].y, ddz = pz - rp[k].z;
            const float dr2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if (!(dr2 <= nl.P2)) continue;
            if (rt[k] - 1u >= par.T) continue;
            const float dr2_inv = (float)(1.0 / (double)dr2);
            const float dx = ddx * dr2_inv, dy = ddy * dr2_inv, dz = ddz * dr2_inv;
            const unsigned kk = par.T * (rt[k] - 1) + lt1;
            const att_pair_params p = tab[kk];
            float vx, vy, vz, ve, Fx, Fy, Fz;
            lj(dr2, dr2_inv, dx, dy, dz, p, par.shape, Fx, Fy, Fz, ve);
            e += ve;
            float pe;
            lj(nl.P2, nl.P2_inv, 0.f, 0.f, 0.f, p, par.shape, vx, vy, vz, pe);
            e -= pe;
            if (ligCharged) {
                const float qlr = ligCharge * recCharge(nIdx) * par.ffelec;
                if (fabsf(qlr) > 0.001f) {
                    coulomb(dr2_inv, dx, dy, dz, qlr, par.dielec, vx, vy, vz, ve);
                    e += ve;
                    Fx += vx; Fy += vy; Fz += vz;
                    coulomb(nl.P2_inv, 0.f, 0.f, 0.f, qlr, par.dielec, vx, vy, vz, ve);
                    e -= ve;
                }
            }
            G.x -= Fx; G.y -= Fy; G.z -= Fz;
            reaction(nIdx, Fx, Fy, Fz);
            if (pg.s) {
                float drc, dac, dem;
                lj_dparam(dr2, dr2_inv, p, par.shape, drc, dac, dem);
                lj_dparam(nl.P2, nl.P2_inv, p, par.shape, drcP, dacP, demP);
                atomicAdd(&pg.s[3 * kk], pg.w * (drc - drcP));
                atomicAdd(&pg.s[3 * kk + 1], pg.w * (dac - dacP));
                if (dem != demP) atomicAdd(&pg.s[3 * kk + 2], pg.w * (dem - demP));
            }
        }
    }
    return e;
}

template <int MM>
__global__ __launch_bounds__(256) void k_adjoint(DProt rec, DProt lig, DGrid rInner, DGrid rOuter,
                                                 DNL nl, DPar par, DBasis bin, DBasis bout,
                                                 const float* __restrict__ dofs, unsigned dofStride,
                                                 const float* __restrict__ wts,
                                                 float* __restrict__ out, unsigned outStride,
                                                 float* __restrict__ pgPartial, unsigned n,
                                                 unsigned tabSlots) {
    extern __shared__ att_pair_params s_dyn[];
    __shared__ float s_red[4][13 + 2 * ATT_MODES_MAX];
    __shared__ unsigned s_qn;
    const att_pair_params* tab = stage_table(par, s_dyn);
    float* s_rx = reinterpret_cast<float*>(s_dyn + tabSlots);
    float* s_ry = s_rx + rec.n;
    float* s_rz = s_ry + rec.n;
    unsigned short* s_rt = reinterpret_cast<unsigned short*>(s_rz + rec.n);
    unsigned* s_queue = reinterpret_cast<unsigned*>(s_rt + ((rec.n + 1) & ~1u));
    float* s_pg = pgPartial ? reinterpret_cast<float*>(s_queue + pow2_ceil(lig.n)) : nullptr;
    const unsigned npg = 3 * par.T * par.T;
    if (s_pg)
        for (unsigned k = threadIdx.x; k < npg; k += blockDim.x) s_pg[k] = 0.f;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    constexpr int MA = MM > 0 ? MM : 1;
    const float4* __restrict__ rpos = rec.pos;
    const float4* __restrict__ rmodesT = rec.modesT;
    const unsigned Mr = MM > 0 ? rec.M : 0, Ml = MM > 0 ? lig.M : 0, rowT = rec.rowT;

    for (unsigned pose = blockIdx.x; pose < n; pose += gridDim.x) {
        __syncthreads();  // LDS of the previous pose is no longer read
        if (threadIdx.x == 0) s_qn = 0;
        const float* d = dofs + (size_t)dofStride * pose;
        const float tx = d[0], ty = d[1], tz = d[2];
        const float wp = wts ? wts[pose] : 1.f;
        const PGrad pg{s_pg, par.T, wp};
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

        // Phase A: deformed receptor into LDS (the neighbour list's positions).
        for (unsigned j = threadIdx.x; j < rec.n; j += blockDim.x) {
            const float4 x = rpos[j];
            BeadModes<MM> md;
            load_modes(rec, j, md);
            const V3 df = deform(md, Mr, x.x, x.y, x.z, ampR);
            s_rx[j] = df.x; s_ry[j] = df.y; s_rz[j] = df.z;
            s_rt[j] = (unsigned short)rec.type[j];
        }
        __syncthreads();

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
        auto add_grad = [&](unsigned i, const V3& xt, const float3& G) __attribute__((always_inline)) {
            acc[1] += G.x; acc[2] += G.y; acc[3] += G.z;
            acc[4] += G.x * xt.x; acc[5] += G.x * xt.y; acc[6] += G.x * xt.z;
            acc[7] += G.y * xt.x; acc[8] += G.y * xt.y; acc[9] += G.y * xt.z;
            acc[10] += G.z * xt.x; acc[11] += G.z * xt.y; acc[12] += G.z * xt.z;
            if (MM > 0) {
                // G . R m = (R^T G) . m
                const float gx = R[0] * G.x + R[3] * G.y + R[6] * G.z;
                const float gy = R[1] * G.x + R[4] * G.y + R[7] * G.z;
                const float gz = R[2] * G.x + R[5] * G.y + R[8] * G.z;
                BeadModes<MM> md;
                load_modes(lig, i, md);
                mode_dots(md, Ml, -gx, -gy, -gz, gL);  // mode_dots subtracts
            }
        };
        // Phase B1: grid term of every ligand bead; beads with neighbour lists queued.
        for (unsigned i = threadIdx.x; i < lig.n; i += blockDim.x) {
            float4 x;
            V3 xt;
            const V3 p = place(i, x, xt);
            float3 G;
            acc[0] += grid_adj(rInner, rOuter, bin, bout, par, p.x, p.y, p.z, lig.mtype[i], x.w, G, pg);
            add_grad(i, xt, G);
            const unsigned cnt = nl_count(nl, p.x, p.y, p.z, lig.type[i]);
            if (cnt) s_queue[atomicAdd(&s_qn, 1u)] = (cnt < 0xFFFFu ? cnt : 0xFFFFu) << 16 | i;
        }
        __syncthreads();
        const unsigned qn = s_qn;
        const unsigned qp = pow2_ceil(qn);
        for (unsigned q = qn + threadIdx.x; q < qp; q += blockDim.x) s_queue[q] = 0;
        __syncthreads();
        bitonic_sort_desc(s_queue, qp);
        // Phase B2: pair terms, longest lists first.
        auto recPos = [=](unsigned j) { return make_float3(s_rx[j], s_ry[j], s_rz[j]); };
        auto recType = [=](unsigned j) { return (unsigned)s_rt[j]; };
        auto reaction = [&](unsigned j, float fx, float fy, float fz) __attribute__((always_inline)) {
            if (MM > 0) {
                const float4* row = rmodesT + (size_t)j * rowT;
                constexpr unsigned Q = (3 * MA + 3) / 4;
                float mt[4 * Q];
#pragma unroll
                for (unsigned q = 0; q < Q; ++q) {
                    const float4 v = q < rowT ? row[q] : make_float4(0.f, 0.f, 0.f, 0.f);
                    mt[4 * q] = v.x; mt[4 * q + 1] = v.y; mt[4 * q + 2] = v.z; mt[4 * q + 3] = v.w;
                }
#pragma unroll
                for (unsigned m = 0; m < (unsigned)MA; ++m)
                    if (m < Mr) gR[m] += fx * mt[3 * m] + fy * mt[3 * m + 1] + fz * mt[3 * m + 2];
            }
        };
        for (unsigned q = threadIdx.x; q < qn; q += blockDim.x) {
            const unsigned i = s_queue[q] & 0xFFFFu;
            float4 x;
            V3 xt;
            const V3 p = place(i, x, xt);
            float3 G = make_float3(0.f, 0.f, 0.f);
            acc[0] += nl_adj(nl, par, tab, recPos, recType, RecChargeG{rpos}, reaction, p.x, p.y,
                             p.z, lig.type[i], x.w, G, pg);
            add_grad(i, xt, G);
        }

        // Reduction and epilogue.
#pragma unroll
        for (int k = 0; k < 13; ++k) acc[k] = wave_sum(acc[k]);
#pragma unroll
        for (int m = 0; m < MA; ++m) { gR[m] = wave_sum(gR[m]); gL[m] = wave_sum(gL[m]); }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 13; ++k) s_red[wave][k] = acc[k];
#pragma unroll
            for (int m = 0; m < ATT_MODES_MAX; ++m) {
                s_red[wave][13 + m] = m < MA ? gR[m < MA ? m : 0] : 0.f;
                s_red[wave][13 + ATT_MODES_MAX + m] = m < MA ? gL[m < MA ? m : 0] : 0.f;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            float v[13 + 2 * ATT_MODES_MAX];
            const int nw = blockDim.x / kWave;
#pragma unroll
            for (int k = 0; k < 13 + 2 * ATT_MODES_MAX; ++k) {
                float s = s_red[0][k];
                for (int ww = 1; ww < nw; ++ww) s += s_red[ww][k];
                v[k] = s;
            }
            float da[3];
            torque_reduce(t[0], t[1], t[2], t[3], t[4], t[5], v + 4, da);
            float* o = out + (size_t)outStride * pose;
            o[0] = v[0];
            o[1] = wp * v[1]; o[2] = wp * v[2]; o[3] = wp * v[3];
            o[4] = wp * da[0]; o[5] = wp * da[1]; o[6] = wp * da[2];
            if (outStride >= 27)
#pragma unroll
                for (unsigned m = 0; m < ATT_MODES_MAX; ++m) {
                    o[7 + m] = m < Mr ? wp * v[13 + m] : 0.f;
                    o[17 + m] = m < Ml ? wp * v[13 + ATT_MODES_MAX + m] : 0.f;
                }
        }
    }
    if (s_pg) {
        __syncthreads();
        for (unsigned k = threadIdx.x; k < npg; k += blockDim.x)
            pgPartial[(size_t)blockIdx.x * npg + k] = s_pg[k];
    }
}

// Sums the per-workgroup parameter-gradient rows in block order.
__global__ void k_pgrad_reduce(const float* __restrict__ partial, unsigned rows, unsigned cols,
                               float* __restrict__ pg) {
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    double s = 0.0;
    for (unsigned r = 0; r < rows; ++r) s += partial[(size_t)r * cols + c];
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
    return modes_lds_bytes(rec, lig, par, tabSlots) +
           (pgrad ? sizeof(float) * 3 * (size_t)par.T * par.T : 0);
}

hipError_t launch_adjoint(const DProt& rec, const DProt& lig, const DGrid& inner,
                          const DGrid& outer, const DNL& nl, const DPar& par, const DBasis& bin,
                          const DBasis& bout, const float* dofs, bool modes, const float* w,
                          float* out, float* pgPartial, unsigned blocks, unsigned n,
                          hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned slots = 0;
    const size_t lds = adjoint_lds_bytes(rec, lig, par, pgPartial != nullptr, &slots);
    const unsigned ds = modes ? 26 : 6, os = modes ? 27 : 7;
    const unsigned mm = modes ? (rec.M > lig.M ? rec.M : lig.M) : 0;
    if (mm == 0)
        hipLaunchKernelGGL(k_adjoint<0>, dim3(blocks), dim3(256), lds, s, rec, lig, inner, outer,
                           nl, par, bin, bout, dofs, ds, w, out, os, pgPartial, n, slots);
    else if (mm <= 5)
        hipLaunchKernelGGL(k_adjoint<5>, dim3(blocks), dim3(256), lds, s, rec, lig, inner, outer,
                           nl, par, bin, bout, dofs, ds, w, out, os, pgPartial, n, slots);
    else
        hipLaunchKernelGGL(k_adjoint<ATT_MODES_MAX>, dim3(blocks), dim3(256), lds, s, rec, lig,
                           inner, outer, nl, par, bin, bout, dofs, ds, w, out, os, pgPartial, n,
                           slots);
    return hipGetLastError();
}

hipError_t launch_pgrad_reduce(const float* partial, unsigned rows, unsigned cols, float* pg,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_pgrad_reduce, dim3((cols + 255) / 256), dim3(256), 0, s, partial, rows,
                       cols, pg);
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
