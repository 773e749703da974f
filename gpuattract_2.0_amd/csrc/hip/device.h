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
#include "libm_sincosf.h"

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
// sin and cos of (ssi, phi, rot) with the C library's single-precision values
// (sincosf_libm, as the reference's std::sin/std::cos(float)), then every lane
// reads them back.
__device__ __forceinline__ void pose_trig(float phi, float ssi, float rot, int lane, float* t) {
    float c = 0.f, s = 0.f;
    if (lane < 3) sincosf_libm(lane == 0 ? ssi : lane == 1 ? phi : rot, &s, &c);
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

// pot_bead without branches around the lookup: the grid (inner, else outer)
// is selected per lane and its octet always read, the result replaced by zero
// where the bead is in neither grid or has mapped type 0.  The same values as
// pot_bead; a loop over beads can have several lookups in flight.  The read
// stays inside the grid for any coordinate (grid_index clamps to the last
// voxel, NaN included) and type (the scoring entry points refuse bodies with
// a mapped type >= the grid's number of grids).
__device__ __forceinline__ float4 pot_bead_flat(const DGrid& inner, const DGrid& outer, float px,
                                                float py, float pz, unsigned mtype, float charge) {
    const bool in = !grid_out(inner, px, py, pz);
    const bool valid = mtype != 0 && (in || !grid_out(outer, px, py, pz));
    DGrid g;
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
    const unsigned ix = grid_index(px, g.lox, g.dVox_inv, g.nx);
    const unsigned iy = grid_index(py, g.loy, g.dVox_inv, g.ny);
    const unsigned iz = grid_index(pz, g.loz, g.dVox_inv, g.nz);
    float4 pot = trilin(g, px, py, pz, ix, iy, iz, mtype);
    if (fabsf(charge) > 0.001f) {  // interp's electrostatic term
        float4 el = trilin(g, px, py, pz, ix, iy, iz, 0);
        el.x = el.x * charge; el.y = el.y * charge; el.z = el.z * charge; el.w = el.w * charge;
        pot.x = pot.x + el.x; pot.y = pot.y + el.y; pot.z = pot.z + el.z; pot.w = pot.w + el.w;
    }
    return valid ? pot : make_float4(0.f, 0.f, 0.f, 0.f);
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

// Neighbour-list cell of a ligand bead (NLGrid bounds and getIndex,
// model/NLGrid.h:100-131, as NLPotForce tests them, model/neighborlist.h:
// 47-63): {numEl, idx} of the cell's list, {0, 0} when the bead has no
// neighbour-list term (type 0 or outside the grid).  cell receives the index.
__device__ __forceinline__ uint2 nl_cell(const DNL& nl, float px, float py, float pz,
                                         unsigned ligType, int* cell = nullptr) {
    if (ligType == 0) return make_uint2(0u, 0u);
    if (!((px >= nl.lox && px <= nl.hix) && (py >= nl.loy && py <= nl.hiy) &&
          (pz >= nl.loz && pz <= nl.hiz)))
        return make_uint2(0u, 0u);
    // NLGrid::getIndex (model/NLGrid.h:100-108): round half away from zero.
    const int ix = (int)roundf((px - nl.lox) * nl.dVox_inv);
    const int iy = (int)roundf((py - nl.loy) * nl.dVox_inv);
    const int iz = (int)roundf((pz - nl.loz) * nl.dVox_inv);
    if (cell) { cell[0] = ix; cell[1] = iy; cell[2] = iz; }
    return nl.desc[nl.nx * (iz * nl.ny + iy) + ix];
}

// The pair table as the list walks read it: the workgroup's LDS copy
// (stage_table) or the table in global memory.  Entries of the LDS copy are
// read through an LDS-typed pointer, so they are LDS loads; through a generic
// pointer they would be flat loads, which also wait for the vector-memory
// loads of the receptor beads in flight.
struct PairTab {
    const att_pair_params* p;
    bool lds;
    __device__ __forceinline__ att_pair_params operator[](unsigned k) const {
        if (lds) {
            typedef const __attribute__((address_space(3))) float* LdsF;
            typedef const __attribute__((address_space(3))) int* LdsI;
            const LdsF f = (LdsF)(const float*)(p + k);
            att_pair_params e;
            e.rc = f[0];
            e.ac = f[1];
            e.ipon = ((LdsI)(const int*)(p + k))[2];
            e.rmin2 = f[3];
            e.emin = f[4];
            return e;
        }
        return p[k];
    }
};

// Pair table into LDS when it fits (kMaxTypesLDS), else read from global.
__device__ __forceinline__ PairTab stage_table(const DPar& par, att_pair_params* s_tab) {
    if (par.T <= kMaxTypesLDS) {
        for (unsigned k = threadIdx.x; k < par.T * par.T; k += blockDim.x) s_tab[k] = par.tab[k];
        return PairTab{s_tab, true};
    }
    return PairTab{par.tab, false};
}

// Entry n of the plateau-ratio table, sqrt(((n + 0.5) / 10000) / P2) in double
// rounded to float (model/NLGrid.tpp:44-51, read by getRatio at n =
// int(dr2 * 10000), model/NLGrid.h:141-143).  Evaluated per pair rather than
// loaded: a division, a division and a square root, each correctly rounded
// in double as on the host, so the value is the table's to the bit (checked
// over every entry), and the walk waits on no load after the distance.
__device__ __forceinline__ float plateau_ratio(unsigned n, float P2) {
    return (float)__builtin_sqrt((((double)n + 0.5) / 10000.0) / (double)P2);
}

// The pair loop of NLPotForce for one ligand bead (model/neighborlist.h:66-158
// and the receptor reaction of model/neighborlist_modes.h:68-194) over the
// list d = {numEl, idx}: every pair in list order, each term added to or
// subtracted from the bead's accumulators in the reference's order, so the
// bead's sums are the reference's bits.  Accessors:
//   recPos(j)    -> float3 receptor position used for the distance
//   recType(j)   -> receptor type
//   recCharge(j) -> receptor charge (only read when the ligand bead is charged)
//   reaction(j, fx, fy, fz) receives the pair's reaction force (no-op for rigid)
// List entries are fetched NB at a time ahead of use (NB = 1 where recPos
// itself loads much, as a receptor deformed from its mode row).
template <int NB = 4, class RecPos, class RecType, class RecCharge, class Reaction>
__device__ __forceinline__ float4 nl_walk(const DNL& nl, const DPar& par,
                                          const PairTab& tab, RecPos recPos,
                                          RecType recType, RecCharge recCharge,
                                          Reaction reaction, float px, float py, float pz,
                                          unsigned ligType, float ligCharge, uint2 d,
                                          unsigned* npairs) {
    float fx = 0.f, fy = 0.f, fz = 0.f, e = 0.f;
    unsigned pairs = 0;
    const unsigned lt1 = ligType - 1;
    const bool ligCharged = ligCharge != 0.f;  // else q_lig*q_rec*ffelec = 0: no Coulomb term
    const uint32_t* lst = nl.list + d.y;
    // Indices of the next batch are loaded while the current batch computes.
    unsigned nidx[NB];
    if (d.x != 0) {
#pragma unroll
        for (int k = 0; k < NB; ++k) nidx[k] = lst[(unsigned)k < d.x ? k : 0];
    }
    for (unsigned j0 = 0; j0 < d.x; j0 += NB) {
        unsigned idx[NB];
        float3 rp[NB];
        unsigned rt[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) idx[k] = nidx[k];
        if (j0 + NB < d.x) {
#pragma unroll
            for (int k = 0; k < NB; ++k) nidx[k] = lst[j0 + NB + k < d.x ? j0 + NB + k : j0 + NB];
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) { rp[k] = recPos(idx[k]); rt[k] = recType(idx[k]); }
        // Distances of the batch's pairs first, then their plateau ratios, so
        // the table loads are in flight together.
        float ddx[NB], ddy[NB], ddz[NB], dr2s[NB], ratios[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            ddx[k] = px - rp[k].x;
            ddy[k] = py - rp[k].y;
            ddz[k] = pz - rp[k].z;
            dr2s[k] = ddx[k] * ddx[k] + ddy[k] * ddy[k] + ddz[k] * ddz[k];
        }
#pragma unroll
        for (int k = 0; k < NB; ++k)
            ratios[k] = (j0 + k < d.x && dr2s[k] <= nl.P2)
                            ? plateau_ratio((unsigned)(int)(dr2s[k] * 10000.f), nl.P2)
                            : 0.f;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
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
    if (npairs) *npairs = pairs;
    return make_float4(fx, fy, fz, e);
}

// Neighbour-list term of one ligand bead: cell, then the pair loop.  Returns
// false when the bead has no neighbour-list term.
template <class RecPos, class RecType, class RecCharge, class Reaction>
__device__ __forceinline__ bool nl_bead(const DNL& nl, const DPar& par,
                                        const PairTab& tab, RecPos recPos,
                                        RecType recType, RecCharge recCharge, Reaction reaction,
                                        float px, float py, float pz, unsigned ligType,
                                        float ligCharge, float4& acc, int* cell,
                                        unsigned* npairs) {
    acc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (npairs) *npairs = 0;
    if (ligType == 0) return false;
    if (!((px >= nl.lox && px <= nl.hix) && (py >= nl.loy && py <= nl.hiy) &&
          (pz >= nl.loz && pz <= nl.hiz)))
        return false;
    const uint2 d = nl_cell(nl, px, py, pz, ligType, cell);
    acc = nl_walk(nl, par, tab, recPos, recType, recCharge, reaction, px, py, pz, ligType,
                  ligCharge, d, npairs);
    return true;
}

// Beads whose neighbour lists are walked one per lane are ordered by list
// length: nine classes, longest first (>= 64, 48, 36, 28, 22, 16, 11, 6, 1
// pairs), bead order inside a class, so that the lanes of a wave walk lists of
// similar length.  The class of a bead with cnt > 0 listed neighbours:
constexpr unsigned kQueueClasses = 9;
__device__ __forceinline__ unsigned queue_class(unsigned cnt) {
    return (unsigned)(cnt < 64u) + (cnt < 48u) + (cnt < 36u) + (cnt < 28u) + (cnt < 22u) +
           (cnt < 16u) + (cnt < 11u) + (cnt < 6u);
}
__host__ __device__ constexpr unsigned queue_counters(unsigned beads) {
    return kQueueClasses * ((beads + 63) / 64) + 1;
}

// The listed beads (cnt(k) > 0) among beads 0..m-1 of a super-chunk, in the
// class order above, into s_q; returns their number.  Run by nthr threads (a
// multiple of 64, thread tid) with sync() a barrier over them; s_cc holds
// queue_counters(m) counters: per (class, chunk of 64 beads) the number of
// listed beads, then their exclusive prefix in (class, chunk) order.
template <class Cnt, class Sync>
__device__ __forceinline__ unsigned class_queue(Cnt cnt, unsigned m, unsigned short* s_q,
                                                unsigned* s_cc, unsigned tid, unsigned nthr,
                                                Sync sync) {
    const unsigned lane = tid & (kWave - 1), wv = tid / kWave, nw = nthr / kWave;
    const unsigned chunks = (m + kWave - 1) / kWave;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (unsigned ch = wv; ch < chunks; ch += nw) {
        const unsigned k = ch * kWave + lane;
        const unsigned c = k < m ? cnt(k) : 0u;
        const unsigned cls = c ? queue_class(c) : kQueueClasses;
#pragma unroll
        for (unsigned q = 0; q < kQueueClasses; ++q) {
            const unsigned long long b = __ballot(cls == q);
            if (lane == 0) s_cc[q * chunks + ch] = (unsigned)__popcll(b);
        }
    }
    sync();
    if (tid == 0) {
        unsigned sum = 0;
        for (unsigned e = 0; e < kQueueClasses * chunks; ++e) {
            const unsigned v = s_cc[e];
            s_cc[e] = sum;
            sum += v;
        }
        s_cc[kQueueClasses * chunks] = sum;
    }
    sync();
    for (unsigned ch = wv; ch < chunks; ch += nw) {
        const unsigned k = ch * kWave + lane;
        const unsigned c = k < m ? cnt(k) : 0u;
        const unsigned cls = c ? queue_class(c) : kQueueClasses;
#pragma unroll
        for (unsigned q = 0; q < kQueueClasses; ++q) {
            const unsigned long long b = __ballot(cls == q);
            if (cls == q) s_q[s_cc[q * chunks + ch] + (unsigned)__popcll(b & below)] = (unsigned short)k;
        }
    }
    const unsigned total = s_cc[kQueueClasses * chunks];
    sync();
    return total;
}

// Butterfly sum over the 64 lanes of a wave (same value on every lane).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// The same over unsigned counters (exact: any order gives the same sum).
__device__ __forceinline__ unsigned wave_sum_u(unsigned v) {
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

// Byte offset of the adjoint's parameter-derivative slices behind the pair
// table in its dynamic LDS.
__host__ __device__ constexpr size_t adjoint_pg_offset(unsigned tabSlots) {
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
    // Grid part of the parameter gradient as records (when set): slot i's bead
    // k writes recKey/recPos[i * lig.n + k]; keys (cell * T + type - 1, inner
    // cells first) below recInvalid, recInvalid for beads without a grid term.
    // Positions as {x, y, z, w} (recPos4), or {x, y, z} when the call has no
    // weights (recPos3; w = 1).
    unsigned* recKey;
    float4* recPos4;
    float3* recPos3;
    unsigned recCellsIn, recInvalid;
    // when set, per pose {listed neighbours of the beads walked, pairs inside
    // the plateau}
    uint2* counts;
};

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

// Barrier over the lanes of one wave (LDS written by the wave is visible to it).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The reference's per-pose reductions of the ligand's bead forces, in bead
// order, one quantity per lane (lanes 0..12 of a wave): reducePotForce's E
// and F (model/reduction.h:39-75) and reduceTorque's tensor sum of x_l * f_k
// over the undeformed coordinates (:80-104) in lane 4 + 3k + l.  Each call
// adds beads f[0..m) (float4 {fx, fy, fz, E}, LDS), whose undeformed
// coordinates are pos[first..first+m), to the lane's running sum.
__device__ __forceinline__ float reduce_beads_ref(float red, unsigned lane, const float4* f,
                                                  const float4* pos, unsigned first, unsigned m) {
    if (lane >= 13u) return red;
    const unsigned fc = lane == 0 ? 3u : lane < 4u ? lane - 1u : (lane - 4u) / 3u;
    const unsigned xc = lane < 4u ? 3u : (lane - 4u) % 3u;  // 3: no coordinate factor
    const float* ff = reinterpret_cast<const float*>(f);
    // The coordinates are the same for every lane and read-only: read through
    // the constant address space, they are scalar loads (one per wave, from
    // the scalar cache) rather than vector loads of one address.
    typedef const __attribute__((address_space(4))) float* ConstF;
    const ConstF pc = (ConstF)(const float*)(pos + first);
    auto px = [pc](unsigned i) { return make_float4(pc[4 * i], pc[4 * i + 1], pc[4 * i + 2], pc[4 * i + 3]); };
    auto coord = [xc](const float4& x) { return xc == 0u ? x.x : xc == 1u ? x.y : xc == 2u ? x.z : 1.f; };
    unsigned k = 0;
    // eight beads' loads are issued together; the sum stays sequential
    for (; k + 8 <= m; k += 8) {
        float4 xs[8];
        float fv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            xs[u] = px(k + u);
            fv[u] = ff[4 * (k + u) + fc];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) red += coord(xs[u]) * fv[u];  // 1 * f = f exactly
    }
    for (; k < m; ++k) red += coord(px(k)) * ff[4 * k + fc];
    return red;
}

// Beads per super-chunk for n beads and at most smax (a multiple of 64) per
// super-chunk: the fewest super-chunks, of equal size rounded up to 64.
__host__ __device__ inline unsigned super_chunk(unsigned n, unsigned smax) {
    const unsigned parts = n ? (n + smax - 1) / smax : 1u;
    const unsigned per = (n + parts - 1) / parts;
    return (per + 63u) & ~63u;
}

// The epilogue of those reductions, by every lane of the wave that ran them:
// the 13 sums gathered, the force clamped (reducePotForce) and the torque
// tensor contracted with dR/d(angle) (reduceTorque, commons/TorqueMat.h:
// 27-37); out = {E, F[3], dE/dang[3]}.
__device__ __forceinline__ void finish_ref(float red, const float* t, float* out) {
    float v[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) v[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(red), q));
    float Fx = v[1], Fy = v[2], Fz = v[3];
    force_clamp(Fx, Fy, Fz);
    float a3[3];
    torque_reduce(t[0], t[1], t[2], t[3], t[4], t[5], v + 4, a3);
    out[0] = v[0]; out[1] = Fx; out[2] = Fy; out[3] = Fz;
    out[4] = a3[0]; out[5] = a3[1]; out[6] = a3[2];
}

}  // namespace att
