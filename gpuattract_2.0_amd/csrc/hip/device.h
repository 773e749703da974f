// Device-side data descriptors and the per-bead arithmetic of the scoring path.
//
// Every function reproduces the reference CPU path's FP32 operation order
// (no contraction: the library is built with -ffp-contract=off) so that
// per-bead values match the CPU restatement bit for bit; only the per-pose
// sums use a different (wave-tree) order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "attract_hip.h"

namespace att {

constexpr int kWave = 64;
constexpr unsigned kMaxTypesLDS = 32;  // pair table staged in LDS up to 32x32 types

// Interpolation grid (model/IntrplGrid.h:34-209) with the constants its
// constructor derives (model/IntrplGrid.tpp:30-41).  Stored as voxel octets:
// for every (grid, iz, iy, ix) with i < n-1 the eight corner float4 values, in
// the order trilinearInterpolation sums them (000, 100, 010, 001, 101, 011,
// 110, 111), contiguous and 128-byte aligned — one cache line per lookup.
struct DGrid {
    const float4* oct;
    unsigned nx, ny, nz, nxy, nvox, ngrids;
    float lox, loy, loz, hix, hiy, hiz;
    float dVox, dVox_inv, vol_inv;
};

// Neighbour-list grid (model/NLGrid.h:36-202, model/NLGrid.tpp:30-52).
struct DNL {
    const uint2* desc;
    const uint32_t* list;
    const float* ratio;
    unsigned nx, ny, nz;
    float lox, loy, loz, hix, hiy, hiz;
    float dVox_inv, P2, P2_inv;
};

// Protein (model/Protein.h:38-167): pos = {x, y, z, charge}.
struct DProt {
    const float4* pos;
    const unsigned* type;
    const unsigned* mtype;
    const float* modes;   // [c][m][n]: coalesced over beads
    const float4* modesT; // per bead one row of 3*M floats [m][c], padded to a float4 multiple
    const float* modeForce;
    unsigned n, M, rowT;  // rowT = float4 per modesT row
};

// Basis grids of the parameter derivative (attached to a DGrid of the same
// geometry): per voxel, per basis channel, the 8 corner values in trilinear
// order (two float4).  nb = 2T channels [S_1, S6_1, S_2, S6_2, ...].
struct DBasis {
    const float4* oct;
    unsigned nb;
};

struct DPar {
    const att_pair_params* tab;  // T*T, index T*(recType-1)+(ligType-1)
    unsigned T;
    int shape;   // 8 or 12
    int dielec;  // 0 constant, 1 variable
    float ffelec;
};

struct V3 { float x, y, z; };

// RotMat::operator*(Vec3) (commons/RotMat.h:122-129).
__device__ __forceinline__ V3 matvec(const float* m, float x, float y, float z) {
    V3 o;
    o.x = m[0] * x + m[1] * y + m[2] * z;
    o.y = m[3] * x + m[4] * y + m[5] * z;
    o.z = m[6] * x + m[7] * y + m[8] * z;
    return o;
}

// euler2rotmat (model/matrixFunctions.h:38-77) from the six trig values
// {cSSI, cPHI, sSSI, sPHI, cROT, sROT}.
__device__ __forceinline__ void rotmat_from_trig(float cSSI, float cPHI, float sSSI, float sPHI,
                                                 float cROT, float sROT, float* m) {
    const float cSSI_cPHI = cSSI * cPHI, cSSI_sPHI = cSSI * sPHI;
    const float sSSI_cPHI = sSSI * cPHI, sSSI_sPHI = sSSI * sPHI;
    m[0] = cROT * cSSI_cPHI + sROT * sPHI;
    m[1] = sROT * cSSI_cPHI - cROT * sPHI;
    m[2] = sSSI_cPHI;
    m[3] = cROT * cSSI_sPHI - sROT * cPHI;
    m[4] = sROT * cSSI_sPHI + cROT * cPHI;
    m[5] = sSSI_sPHI;
    m[6] = -cROT * sSSI;
    m[7] = -sROT * sSSI;
    m[8] = cSSI;
}

// euler2torquemat (model/matrixFunctions.h:79-127) contracted with the torque
// tensor T[k][l] as TorqueMat::rotateReduce does (commons/TorqueMat.h:27-37).
__device__ __forceinline__ void torque_reduce(float cSSI, float cPHI, float sSSI, float sPHI,
                                              float cROT, float sROT, const float* T, float* out) {
    const float cSSI_cPHI = cSSI * cPHI, cSSI_sPHI = cSSI * sPHI;
    const float sSSI_cPHI = sSSI * cPHI, sSSI_sPHI = sSSI * sPHI;
    float tm[27];
    tm[0] = -cROT * cSSI_sPHI + sROT * cPHI;
    tm[1] = -sROT * cSSI_sPHI - cROT * cPHI;
    tm[2] = -sSSI_sPHI;
    tm[9] = cROT * cSSI_cPHI + sROT * sPHI;
    tm[10] = sROT * cSSI_cPHI - cROT * sPHI;
    tm[11] = sSSI_cPHI;
    tm[18] = 0.f; tm[19] = 0.f; tm[20] = 0.f;
    tm[3] = -cROT * sSSI_cPHI;
    tm[4] = -sROT * sSSI_cPHI;
    tm[5] = cSSI_cPHI;
    tm[12] = -cROT * sSSI_sPHI;
    tm[13] = -sROT * sSSI_sPHI;
    tm[14] = cSSI_sPHI;
    tm[21] = -cROT * cSSI;
    tm[22] = -sROT * cSSI;
    tm[23] = -sSSI;
    tm[6] = -sROT * cSSI_cPHI + cROT * sPHI;
    tm[7] = cROT * cSSI_cPHI + sROT * sPHI;
    tm[8] = 0.f;
    tm[15] = -sROT * cSSI_sPHI - cROT * cPHI;
    tm[16] = cROT * cSSI_sPHI - sROT * cPHI;
    tm[17] = 0.f;
    tm[24] = sROT * sSSI;
    tm[25] = -cROT * sSSI;
    tm[26] = 0.f;
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            r0 += tm[9 * k + 0 + l] * T[3 * k + l];
            r1 += tm[9 * k + 3 + l] * T[3 * k + l];
            r2 += tm[9 * k + 6 + l] * T[3 * k + l];
        }
    out[0] = r0; out[1] = r1; out[2] = r2;
}

// The six trig values of a pose, evaluated once per wave: lanes 0..2 take
// sincos of (ssi, phi, rot) in double and round to float (the correctly
// rounded single-precision sin/cos), then every lane reads them back.
__device__ __forceinline__ void pose_trig(float phi, float ssi, float rot, int lane, float* t) {
    float c = 0.f, s = 0.f;
    if (lane < 3) {
        const double a = lane == 0 ? (double)ssi : lane == 1 ? (double)phi : (double)rot;
        double sd, cd;
        sincos(a, &sd, &cd);
        c = (float)cd;
        s = (float)sd;
    }
    t[0] = __shfl(c, 0);  // cSSI
    t[1] = __shfl(c, 1);  // cPHI
    t[2] = __shfl(s, 0);  // sSSI
    t[3] = __shfl(s, 1);  // sPHI
    t[4] = __shfl(c, 2);  // cROT
    t[5] = __shfl(s, 2);  // sROT
}

// IntrplGrid::outOfBounds_byPos (model/IntrplGrid.h:113-117).
// Written as "not inside" so that a NaN coordinate counts as outside (the
// reference's form would index the grid with it); identical for finite input.
__device__ __forceinline__ bool grid_out(const DGrid& g, float x, float y, float z) {
    return !((x >= g.lox && x <= g.hix) && (y >= g.loy && y <= g.hiy) && (z >= g.loz && z <= g.hiz));
}

__device__ __forceinline__ unsigned grid_index(float p, float lo, float inv, unsigned n) {
    const unsigned i = (unsigned)floorf((p - lo) * inv);
    return i < n - 2 ? i : n - 2;
}

// getVoxelByIndex + trilinearInterpolation (model/IntrplGrid.h:130-163,
// model/trilinIntrpl.h:28-65): eight float4 corners, weights and summation in
// the reference's order, then scaled by 1/voxelVolume.
__device__ __forceinline__ float4 trilin(const DGrid& g, float px, float py, float pz, unsigned ix,
                                         unsigned iy, unsigned iz, unsigned type) {
    const float mnx = (float)(int)ix * g.dVox + g.lox;
    const float mny = (float)(int)iy * g.dVox + g.loy;
    const float mnz = (float)(int)iz * g.dVox + g.loz;
    const float ax = px - mnx, ay = py - mny, az = pz - mnz;
    const float bx = (mnx + g.dVox) - px, by = (mny + g.dVox) - py, bz = (mnz + g.dVox) - pz;
    const float tmpMax_xy = bx * by;
    const float tmpMin_xy = ax * ay;
    const float w0 = tmpMax_xy * bz;
    const float w1 = (ax * by) * bz;
    const float w2 = (bx * ay) * bz;
    const float w3 = tmpMax_xy * az;
    const float w4 = (ax * by) * az;
    const float w5 = (bx * ay) * az;
    const float w6 = tmpMin_xy * bz;
    const float w7 = tmpMin_xy * az;
    const size_t o = (((size_t)type * (g.nz - 1) + iz) * (g.ny - 1) + iy) * (g.nx - 1) + ix;
    const float4* d = g.oct + 8 * o;
    const float4 c0 = d[0], c1 = d[1], c2 = d[2], c3 = d[3];
    const float4 c4 = d[4], c5 = d[5], c6 = d[6], c7 = d[7];
    float4 V;
    V.x = c0.x * w0; V.y = c0.y * w0; V.z = c0.z * w0; V.w = c0.w * w0;
    V.x = V.x + c1.x * w1; V.y = V.y + c1.y * w1; V.z = V.z + c1.z * w1; V.w = V.w + c1.w * w1;
    V.x = V.x + c2.x * w2; V.y = V.y + c2.y * w2; V.z = V.z + c2.z * w2; V.w = V.w + c2.w * w2;
    V.x = V.x + c3.x * w3; V.y = V.y + c3.y * w3; V.z = V.z + c3.z * w3; V.w = V.w + c3.w * w3;
    V.x = V.x + c4.x * w4; V.y = V.y + c4.y * w4; V.z = V.z + c4.z * w4; V.w = V.w + c4.w * w4;
    V.x = V.x + c5.x * w5; V.y = V.y + c5.y * w5; V.z = V.z + c5.z * w5; V.w = V.w + c5.w * w5;
    V.x = V.x + c6.x * w6; V.y = V.y + c6.y * w6; V.z = V.z + c6.z * w6; V.w = V.w + c6.w * w6;
    V.x = V.x + c7.x * w7; V.y = V.y + c7.y * w7; V.z = V.z + c7.z * w7; V.w = V.w + c7.w * w7;
    V.x = V.x * g.vol_inv; V.y = V.y * g.vol_inv; V.z = V.z * g.vol_inv; V.w = V.w * g.vol_inv;
    return V;
}

// interpolate (model/interpolation.h:26-53).
__device__ __forceinline__ float4 interp(const DGrid& g, float px, float py, float pz,
                                         unsigned type, float charge, int* cell) {
    const unsigned ix = grid_index(px, g.lox, g.dVox_inv, g.nx);
    const unsigned iy = grid_index(py, g.loy, g.dVox_inv, g.ny);
    const unsigned iz = grid_index(pz, g.loz, g.dVox_inv, g.nz);
    if (cell) { cell[0] = (int)ix; cell[1] = (int)iy; cell[2] = (int)iz; }
    float4 pot = trilin(g, px, py, pz, ix, iy, iz, type);
    if (fabsf(charge) > 0.001f) {
        float4 el = trilin(g, px, py, pz, ix, iy, iz, 0);
        el.x = el.x * charge; el.y = el.y * charge; el.z = el.z * charge; el.w = el.w * charge;
        pot.x = pot.x + el.x; pot.y = pot.y + el.y; pot.z = pot.z + el.z; pot.w = pot.w + el.w;
    }
    return pot;
}

// potForce for one bead (model/interpolation.h:56-102): inner grid, else outer
// grid, else zero; mapped type 0 contributes nothing.
__device__ __forceinline__ float4 pot_bead(const DGrid& inner, const DGrid& outer, float px,
                                           float py, float pz, unsigned mtype, float charge,
                                           int* cell) {
    float4 pot = make_float4(0.f, 0.f, 0.f, 0.f);
    if (mtype == 0) return pot;
    if (grid_out(inner, px, py, pz)) {
        if (!grid_out(outer, px, py, pz)) pot = interp(outer, px, py, pz, mtype, charge, cell ? cell + 3 : nullptr);
    } else {
        pot = interp(inner, px, py, pz, mtype, charge, cell);
    }
    return pot;
}

// LJPotForce (model/forcefield.h:28-83) with swi = 1: 6.0*vlj and the force
// factor are evaluated in double, as the reference's FP32 build does.
__device__ __forceinline__ void lj(float dr2, float dr2_inv, float dx, float dy, float dz,
                                   const att_pair_params& p, int shape, float& fx, float& fy,
                                   float& fz, float& e) {
    const float ivor = (float)p.ipon;
    const float rr23 = dr2_inv * dr2_inv * dr2_inv;
    const float rrd = shape == 8 ? dr2_inv : rr23;
    const float shapedelta = shape == 8 ? 2.f : 6.f;
    const float rep = p.rc * rrd;
    const float vlj = (rep - p.ac) * rr23;
    const double s = 6.0 * (double)vlj + (double)(shapedelta * (rep * rr23));
    float fb;
    if (dr2 < p.rmin2) {
        e = vlj + (ivor - 1.f) * p.emin;
        fb = (float)s;
    } else {
        e = ivor * vlj;
        fb = (float)((double)ivor * s);
    }
    fx = fb * dx; fy = fb * dy; fz = fb * dz;
}

// ChargePotForce (model/forcefield.h:85-146) with swi = 1.
__device__ __forceinline__ void coulomb(float dr2_inv, float dx, float dy, float dz, float q,
                                        int dielec, float& fx, float& fy, float& fz, float& e) {
    float dd;
    if (dielec == 0) dd = (float)((double)sqrtf(dr2_inv) - 1.0 / 50.0);
    else dd = (float)((double)dr2_inv - 1.0 / (50.0 * 50.0));
    if (dd < 0.f) dd = 0.f;
    e = q * dd;
    if (dd <= 0.f) { fx = 0.f; fy = 0.f; fz = 0.f; return; }
    if (dielec == 0) {
        const float et2 = q * sqrtf(dr2_inv);
        fx = et2 * dx; fy = et2 * dy; fz = et2 * dz;
    } else {
        const float et2 = q * dr2_inv;
        fx = (2.f * et2) * dx; fy = (2.f * et2) * dy; fz = (2.f * et2) * dz;
    }
}

// Neighbour-list term of one ligand bead (model/neighborlist.h:47-158 and the
// receptor reaction of model/neighborlist_modes.h:68-194).  Accessors:
//   recPos(j)    -> float3 receptor position used for the distance
//   recType(j)   -> receptor type
//   recCharge(j) -> receptor charge (only read when the ligand bead is charged)
//   reaction(j, fx, fy, fz) receives the pair's reaction force (no-op for rigid)
// List entries are fetched four at a time ahead of use; the pairs are still
// accumulated one after another in list order, as the reference does.
// Returns false when the bead is outside the neighbour-list grid.
template <class RecPos, class RecType, class RecCharge, class Reaction>
__device__ __forceinline__ bool nl_bead(const DNL& nl, const DPar& par,
                                        const att_pair_params* tab, RecPos recPos,
                                        RecType recType, RecCharge recCharge, Reaction reaction,
                                        float px, float py, float pz, unsigned ligType,
                                        float ligCharge, float4& acc, int* cell,
                                        unsigned* npairs) {
    acc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ligType == 0) return false;
    if (!((px >= nl.lox && px <= nl.hix) && (py >= nl.loy && py <= nl.hiy) &&
          (pz >= nl.loz && pz <= nl.hiz)))
        return false;
    // NLGrid::getIndex (model/NLGrid.h:100-108): round half away from zero.
    const int ix = (int)roundf((px - nl.lox) * nl.dVox_inv);
    const int iy = (int)roundf((py - nl.loy) * nl.dVox_inv);
    const int iz = (int)roundf((pz - nl.loz) * nl.dVox_inv);
    if (cell) { cell[0] = ix; cell[1] = iy; cell[2] = iz; }
    const uint2 d = nl.desc[nl.nx * (iz * nl.ny + iy) + ix];
    float fx = 0.f, fy = 0.f, fz = 0.f, e = 0.f;
    unsigned pairs = 0;
    const unsigned lt1 = ligType - 1;
    const bool ligCharged = ligCharge != 0.f;  // else q_lig*q_rec*ffelec = 0: no Coulomb term
    const uint32_t* lst = nl.list + d.y;
    // Indices of the next batch are loaded while the current batch computes.
    unsigned nidx[4];
    if (d.x != 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) nidx[k] = lst[(unsigned)k < d.x ? k : 0];
    }
    for (unsigned j0 = 0; j0 < d.x; j0 += 4) {
        unsigned idx[4];
        float3 rp[4];
        unsigned rt[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) idx[k] = nidx[k];
        if (j0 + 4 < d.x) {
#pragma unroll
            for (int k = 0; k < 4; ++k) nidx[k] = lst[j0 + 4 + k < d.x ? j0 + 4 + k : j0 + 4];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) { rp[k] = recPos(idx[k]); rt[k] = recType(idx[k]); }
        // Distances of the four pairs first, then their plateau ratios, so the
        // four table loads are in flight together.
        float ddx[4], ddy[4], ddz[4], dr2s[4], ratios[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ddx[k] = px - rp[k].x;
            ddy[k] = py - rp[k].y;
            ddz[k] = pz - rp[k].z;
            dr2s[k] = ddx[k] * ddx[k] + ddy[k] * ddy[k] + ddz[k] * ddz[k];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            ratios[k] = (j0 + k < d.x && dr2s[k] <= nl.P2) ? nl.ratio[(int)(dr2s[k] * 10000.f)] : 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (j0 + k >= d.x) break;
            const unsigned nIdx = idx[k];
            float dx = ddx[k], dy = ddy[k], dz = ddz[k];
            const float dr2 = dr2s[k];
            if (!(dr2 <= nl.P2)) continue;  // outOfPlateau (model/NLGrid.h:133-135); NaN skipped
            // A listed bead without a valid type would index outside the table
            // (the reference reads out of bounds there); such pairs are skipped.
            if (rt[k] - 1u >= par.T) continue;
            ++pairs;
            const float dr2_inv = (float)(1.0 / (double)dr2);
            dx *= dr2_inv; dy *= dr2_inv; dz *= dr2_inv;
            const att_pair_params p = tab[par.T * (rt[k] - 1) + lt1];
            float vx, vy, vz, ve;
            float rfx = 0.f, rfy = 0.f, rfz = 0.f;
            lj(dr2, dr2_inv, dx, dy, dz, p, par.shape, vx, vy, vz, ve);
            fx += vx; fy += vy; fz += vz; e += ve;
            rfx -= vx; rfy -= vy; rfz -= vz;
            const float ratio = ratios[k];
            const float rdx = ratio * dx, rdy = ratio * dy, rdz = ratio * dz;
            lj(nl.P2, nl.P2_inv, rdx, rdy, rdz, p, par.shape, vx, vy, vz, ve);
            fx -= vx; fy -= vy; fz -= vz; e -= ve;
            rfx += vx; rfy += vy; rfz += vz;
            if (ligCharged) {
                const float qlr = ligCharge * recCharge(nIdx) * par.ffelec;
                if (fabsf(qlr) > 0.001f) {
                    coulomb(dr2_inv, dx, dy, dz, qlr, par.dielec, vx, vy, vz, ve);
                    fx += vx; fy += vy; fz += vz; e += ve;
                    rfx -= vx; rfy -= vy; rfz -= vz;
                    coulomb(nl.P2_inv, rdx, rdy, rdz, qlr, par.dielec, vx, vy, vz, ve);
                    fx -= vx; fy -= vy; fz -= vz; e -= ve;
                    rfx += vx; rfy += vy; rfz += vz;
                }
            }
            reaction(nIdx, rfx, rfy, rfz);
        }
    }
    acc = make_float4(fx, fy, fz, e);
    if (npairs) *npairs = pairs;
    return true;
}

// Number of listed neighbours nl_bead will traverse for a bead (0 when the
// bead has no neighbour-list term); same bounds and index as nl_bead.
__device__ __forceinline__ unsigned nl_count(const DNL& nl, float px, float py, float pz,
                                             unsigned ligType) {
    if (ligType == 0) return 0;
    if (!((px >= nl.lox && px <= nl.hix) && (py >= nl.loy && py <= nl.hiy) &&
          (pz >= nl.loz && pz <= nl.hiz)))
        return 0;
    const int ix = (int)roundf((px - nl.lox) * nl.dVox_inv);
    const int iy = (int)roundf((py - nl.loy) * nl.dVox_inv);
    const int iz = (int)roundf((pz - nl.loz) * nl.dVox_inv);
    return nl.desc[nl.nx * (iz * nl.ny + iy) + ix].x;
}

// Bitonic sort of n (a power of two) keys in LDS into descending order, by the
// whole workgroup.
__device__ __forceinline__ void bitonic_sort_desc(unsigned* s, unsigned n) {
    for (unsigned k = 2; k <= n; k <<= 1)
        for (unsigned j = k >> 1; j > 0; j >>= 1) {
            for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
                const unsigned l = i ^ j;
                if (l > i) {
                    const unsigned a = s[i], b = s[l];
                    const bool swap = ((i & k) == 0) ? (a < b) : (a > b);
                    if (swap) { s[i] = b; s[l] = a; }
                }
            }
            __syncthreads();
        }
}

__host__ __device__ inline unsigned pow2_ceil(unsigned v) {
    unsigned p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Butterfly sum over the 64 lanes of a wave (same value on every lane).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// reducePotForce clamp (model/reduction.h:60-72); x *= 0.01 is a double multiply.
__device__ __forceinline__ void force_clamp(float& x, float& y, float& z) {
#pragma unroll
    for (int it = 0; it < 3; ++it) {
        const float magn2 = x * x + y * y + z * z;
        if (magn2 > 1.0e18f) {
            x = (float)((double)x * 0.01);
            y = (float)((double)y * 0.01);
            z = (float)((double)z * 0.01);
        }
    }
}

// correctModeForce (model/reduction_modes.h:100-115): d += 4 k d^4, the power in double.
__device__ __forceinline__ float correct_mode(float d, float k) {
    const double dd = (double)d, d2 = dd * dd;
    const float cf = (float)((double)(4.f * k) * (d2 * d2));
    return d + cf;
}

// Pair table into LDS when it fits (kMaxTypesLDS), else read from global.
__device__ __forceinline__ const att_pair_params* stage_table(const DPar& par,
                                                              att_pair_params* s_tab) {
    if (par.T <= kMaxTypesLDS) {
        for (unsigned k = threadIdx.x; k < par.T * par.T; k += blockDim.x) s_tab[k] = par.tab[k];
        return s_tab;
    }
    return par.tab;
}

// Byte offset of the per-bead neighbour-list cells behind the pair table.
__host__ __device__ constexpr size_t adjoint_nl_offset(unsigned tabSlots) {
    return (sizeof(att_pair_params) * (size_t)tabSlots + 15) & ~(size_t)15;
}

// Everything the adjoint of one pose needs about its receptor-ligand pair.
struct PairDesc {
    DProt rec, lig;
    DGrid inner, outer;
    DNL nl;
    DPar par;
    DBasis bin, bout;
    const float* dofs;   // n * dofStride floats
    const float* wts;    // n upstream weights, or nullptr (all 1)
    float* out;          // n * outStride floats
    unsigned dofStride, outStride, n, tabSlots;
    const unsigned* nDev;  // when set, the number of poses is read here at launch (<= n)
    const unsigned* order; // when set, a spatial permutation of the poses (XCD ranges)
};

// Dynamic LDS of a pair: pair table | per-bead {count, start} | queue of
// bead indices (u16) | parameter-gradient accumulators.
__host__ __device__ inline size_t adjoint_queue_offset(unsigned tabSlots, unsigned ligN) {
    return adjoint_nl_offset(tabSlots) + 8 * (size_t)ligN;
}
__host__ __device__ inline size_t adjoint_pg_offset(unsigned tabSlots, unsigned ligN) {
    return adjoint_queue_offset(tabSlots, ligN) + 2 * (((size_t)ligN + 1) & ~(size_t)1);
}

// Exclusive prefix sum of one value per thread over the workgroup (s_w holds
// one value per wave); `total` receives the sum.  Contains two barriers.
__device__ __forceinline__ unsigned block_scan(unsigned v, unsigned* s_w, unsigned& total) {
    const unsigned lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    unsigned x = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const unsigned y = __shfl_up(x, d);
        if (lane >= (unsigned)d) x += y;
    }
    if (lane == kWave - 1) s_w[wave] = x;
    __syncthreads();
    unsigned before = 0, tot = 0;
    const unsigned nw = blockDim.x / kWave;
    for (unsigned w = 0; w < nw; ++w) {
        const unsigned t = s_w[w];
        before += w < wave ? t : 0u;
        tot += t;
    }
    total = tot;
    __syncthreads();
    return before + x - v;
}

// OR of v over the lanes of each row of 16 (DPP row rotations, which need no
// shift direction), then over the four rows (row heads by readlane): uniform.
__device__ __forceinline__ unsigned wave_or(unsigned v) {
#define ATT_ROR_OR(O) v |= (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x120 + O, 0xF, 0xF, false);
    ATT_ROR_OR(1)
    ATT_ROR_OR(2)
    ATT_ROR_OR(4)
    ATT_ROR_OR(8)
#undef ATT_ROR_OR
    return (unsigned)(__builtin_amdgcn_readlane((int)v, 0) | __builtin_amdgcn_readlane((int)v, 16) |
                      __builtin_amdgcn_readlane((int)v, 32) | __builtin_amdgcn_readlane((int)v, 48));
}

// Lane j holds e_j, the end of the j-th slot of a chunk of 64 pairs starting
// at B (ends strictly increasing, all above B; past the last slot, ~0u).
// Returns, for the pair B + lane, the number of ends at or below it: the
// ends inside (B, B + 64) are marked in one 64-bit mask m (bit e_j - B: a
// slot starts at that lane), and the lane counts the marks at or below its
// position.
__device__ __forceinline__ unsigned ends_up_to(unsigned e, unsigned B, unsigned lane,
                                               unsigned long long& m) {
    const unsigned r = e - B;
    const unsigned mlo = wave_or(r < 32u ? 1u << r : 0u);
    const unsigned mhi = wave_or(r - 32u < 32u ? 1u << (r - 32u) : 0u);
    m = ((unsigned long long)mhi << 32) | mlo;
    const unsigned long long upto = lane == 63u ? ~0ull : (2ull << lane) - 1ull;
    return (unsigned)__popcll(m & upto);
}

// Lane i reads lane i + O of its row of 16 (DPP row shift; O = 1, 2, 4, 8):
// no LDS traffic, unlike a cross-lane shuffle.
template <int O>
__device__ __forceinline__ unsigned row_down(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x100 + O, 0xF, 0xF, false);
}
template <int O>
__device__ __forceinline__ float row_down(float v) {
    return __uint_as_float(row_down<O>(__float_as_uint(v)));
}

// Segmented sums over runs of lanes with equal `seg` (runs are contiguous:
// seg never decreases across the wave): the first lane of each run receives
// the sum of the run, in a fixed order.  First the suffix sums inside each
// row of 16 lanes (DPP), then the partial of every later row whose first lane
// continues the run, taken from those lanes in row order.
__device__ __forceinline__ void seg_run_sums(float& Fx, float& Fy, float& Fz, unsigned seg,
                                             unsigned lane) {
    const unsigned col = lane & 15u;
#define ATT_ROW_STEP(O)                                                              \
    {                                                                                \
        const float ox = row_down<O>(Fx), oy = row_down<O>(Fy), oz = row_down<O>(Fz); \
        const unsigned os = row_down<O>(seg);                                        \
        if (col + O < 16u && os == seg) { Fx += ox; Fy += oy; Fz += oz; }            \
    }
    ATT_ROW_STEP(1)
    ATT_ROW_STEP(2)
    ATT_ROW_STEP(4)
    ATT_ROW_STEP(8)
#undef ATT_ROW_STEP
    const unsigned row = lane >> 4;
#pragma unroll
    for (unsigned r = 1; r < 4; ++r) {
        const unsigned hs = (unsigned)__builtin_amdgcn_readlane((int)seg, 16 * r);
        const float hx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Fx), 16 * r));
        const float hy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Fy), 16 * r));
        const float hz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Fz), 16 * r));
        if (row < r && hs == seg) { Fx += hx; Fy += hy; Fz += hz; }
    }
}


// The same with a fourth summed value (the energy of a pair).
__device__ __forceinline__ void seg_run_sums(float& Fx, float& Fy, float& Fz, float& Fe,
                                             unsigned seg, unsigned lane) {
    const unsigned col = lane & 15u;
#define ATT_ROW_STEP4(O)                                                              \
    {                                                                                 \
        const float ox = row_down<O>(Fx), oy = row_down<O>(Fy), oz = row_down<O>(Fz);  \
        const float oe = row_down<O>(Fe);                                             \
        const unsigned os = row_down<O>(seg);                                         \
        if (col + O < 16u && os == seg) { Fx += ox; Fy += oy; Fz += oz; Fe += oe; }   \
    }
    ATT_ROW_STEP4(1)
    ATT_ROW_STEP4(2)
    ATT_ROW_STEP4(4)
    ATT_ROW_STEP4(8)
#undef ATT_ROW_STEP4
    const unsigned row = lane >> 4;
#pragma unroll
    for (unsigned r = 1; r < 4; ++r) {
        const unsigned hs = (unsigned)__builtin_amdgcn_readlane((int)seg, 16 * r);
        const float hx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Fx), 16 * r));
        const float hy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Fy), 16 * r));
        const float hz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Fz), 16 * r));
        const float he = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Fe), 16 * r));
        if (row < r && hs == seg) { Fx += hx; Fy += hy; Fz += hz; Fe += he; }
    }
}

// Slot of workgroup b among nb in a non-persistent launch such that the
// workgroups of one XCD (dispatched round-robin: b % 8) take one contiguous
// range of slots: XCD x holds ceil((nb - x) / 8) workgroups, and its range
// starts after those of XCDs 0..x-1.  A bijection of [0, nb).
__device__ __forceinline__ unsigned xcd_slot(unsigned b, unsigned nb) {
    const unsigned x = b & 7u, q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// One unit of batched work: `count` consecutive poses of pair `pair` (in the
// pair's spatial order when it has one).
struct BatchTile {
    unsigned pair, first, count;
};

// Tile queues of a batched launch: one per XCD.
constexpr unsigned kBatchQueues = 8;

struct NoReaction {
    __device__ void operator()(unsigned, float, float, float) const {}
};

// Receptor accessors reading the protein arrays in global memory.
struct RecPosG {
    const float4* pos;
    __device__ float3 operator()(unsigned j) const {
        const float4 v = pos[j];
        return make_float3(v.x, v.y, v.z);
    }
};
struct RecTypeG {
    const unsigned* type;
    __device__ unsigned operator()(unsigned j) const { return type[j]; }
};
struct RecChargeG {
    const float4* pos;
    __device__ float operator()(unsigned j) const { return pos[j].w; }
};

// Mode vectors of bead i (modes stored [c][m][n], coalesced over beads); MM
// is the compile-time capacity (>= the body's number of modes).
template <int MM>
struct BeadModes {
    float x[MM > 0 ? MM : 1], y[MM > 0 ? MM : 1], z[MM > 0 ? MM : 1];
};

template <int MM>
__device__ __forceinline__ void load_modes(const DProt& b, unsigned i, BeadModes<MM>& md) {
    const size_t nM = (size_t)b.M * b.n;
#pragma unroll
    for (unsigned m = 0; m < (unsigned)MM; ++m) {
        md.x[m] = m < b.M ? b.modes[(size_t)m * b.n + i] : 0.f;
        md.y[m] = m < b.M ? b.modes[nM + (size_t)m * b.n + i] : 0.f;
        md.z[m] = m < b.M ? b.modes[2 * nM + (size_t)m * b.n + i] : 0.f;
    }
}

// Mode deformation x + sum_m a_m * mode_m, sequential over modes
// (model/transform.h:69-73).
template <int MM>
__device__ __forceinline__ V3 deform(const BeadModes<MM>& md, unsigned M, float x, float y,
                                     float z, const float* amp) {
    V3 p{x, y, z};
#pragma unroll
    for (unsigned m = 0; m < (unsigned)MM; ++m) {
        if (m < M) {
            const float a = amp[m];
            p.x = p.x + a * md.x[m];
            p.y = p.y + a * md.y[m];
            p.z = p.z + a * md.z[m];
        }
    }
    return p;
}

// Accumulates -f . mode_m into g[m] (model/reduction_modes.h:49-55, 83-86).
template <int MM>
__device__ __forceinline__ void mode_dots(const BeadModes<MM>& md, unsigned M, float fx, float fy,
                                          float fz, float* g) {
#pragma unroll
    for (unsigned m = 0; m < (unsigned)MM; ++m)
        if (m < M) g[m] -= fx * md.x[m] + fy * md.y[m] + fz * md.z[m];
}

}  // namespace att
