// Potential and neighbour-list grids of a body, built on the device.
//
// The reference reads grids made by ATTRACT's own tools (fileIO/grid_orig.cpp:
// 85-301); the semantics here are those the scoring path assumes: for a probe
// of type t at grid point p,
//   E_t(p) = sum_j V_{type(j),t}(max(|p - x_j|, P))
// over the source beads (LJ of model/forcefield.h:28-83 with r' = max(r, P),
// P the plateau), the force channel the matching -dV/dr' along (p - x_j)/r,
// grid 0 the Coulomb term of model/forcefield.h:85-146 at r'; and for the
// neighbour-list grid, the source beads with type != 0 within `neighbourdis` of
// p in ascending index order (model/NLGrid.h:100-143 reads them).  Near beads
// (r < P) contribute V(P), exactly what the neighbour-list term subtracts
// (model/neighborlist.h:121-157).
//
// Every sum runs in double in the order of the host generator
// (csrc/gen/attract_gen.cpp), with contraction off and correctly rounded
// division and square root, so the device grids equal the host generator's
// bit for bit.  One thread per grid point; bead coordinates are wave-uniform
// loads.
#include "device.h"
#include "internal.h"

namespace att {

namespace {

__device__ inline void point_of(const GenGeom& g, size_t pidx, double& px, double& py, double& pz) {
    size_t q = pidx;
    const unsigned ix = (unsigned)(q % g.nx);
    q /= g.nx;
    const unsigned iy = (unsigned)(q % g.ny);
    const unsigned iz = (unsigned)(q / g.ny);
    px = g.ox + (double)ix * g.sp;
    py = g.oy + (double)iy * g.sp;
    pz = g.oz + (double)iz * g.sp;
}

// LJ at squared distance dr2 (already max'ed with the plateau), in double:
// energy and fb with force = fb * d / (r * r').
__device__ inline void lj_eval(double dr2, const GenPair& p, int shape, double& E, double& fb) {
    const double dr2_inv = 1.0 / dr2;
    const double rr23 = dr2_inv * dr2_inv * dr2_inv;
    const double rrd = shape == 8 ? dr2_inv : rr23;
    const double sd = shape == 8 ? 2.0 : 6.0;
    const double rep = p.rc * rrd;
    const double vlj = (rep - p.ac) * rr23;
    const double f = 6.0 * vlj + sd * (rep * rr23);
    if (dr2 < p.rmin2) {
        E = vlj + (p.ipon - 1.0) * p.emin;
        fb = f;
    } else {
        E = p.ipon * vlj;
        fb = p.ipon * f;
    }
}

}  // namespace

// Every pair in the r >= rmin branch once r' >= P: per source type a, the sums
// S = sum 1/r'^s, S6 = sum 1/r'^6 and their force directions, then
// E_t += ipon (rc S - ac S6) for the probe types tFirst..tFirst+tCount-1, in
// ascending a.
template <int TCH>
__global__ __launch_bounds__(256) void k_gen_lin(GenGeom g, const float4* __restrict__ sorted,
                                                 const unsigned* __restrict__ typeStart,
                                                 const GenPair* __restrict__ tab, unsigned T,
                                                 int shape, double P2, double sqrtP2,
                                                 unsigned tFirst, unsigned tCount,
                                                 float4* __restrict__ out, size_t npts) {
    const size_t pidx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (pidx >= npts) return;
    double px, py, pz;
    point_of(g, pidx, px, py, pz);
    double Et[TCH], Fx[TCH], Fy[TCH], Fz[TCH];
#pragma unroll
    for (int k = 0; k < TCH; ++k) Et[k] = Fx[k] = Fy[k] = Fz[k] = 0.0;
    const double sd = shape == 8 ? 2.0 : 6.0;
    const unsigned TT = T + 1;
    for (unsigned a = 1; a <= T; ++a) {
        double A0 = 0, A1 = 0, A2 = 0, A3 = 0, A4 = 0, A5 = 0, A6 = 0, A7 = 0;
        const unsigned j1 = typeStart[a + 1];
        for (unsigned j = typeStart[a]; j < j1; ++j) {
            const float4 b = sorted[j];
            const double dx = px - b.x, dy = py - b.y, dz = pz - b.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double r = sqrt(r2);
            const double rp2 = r2 > P2 ? r2 : P2;
            const double rp = r2 > P2 ? r : sqrtP2;
            const double i2 = 1.0 / rp2;
            const double i6 = i2 * i2 * i2;
            const double is = shape == 8 ? i6 * i2 : i6 * i6;
            A0 += is;
            A1 += i6;
            if (r > 0) {
                const double w = 1.0 / (r * rp);
                A2 += is * w * dx;
                A3 += is * w * dy;
                A4 += is * w * dz;
                A5 += i6 * w * dx;
                A6 += i6 * w * dy;
                A7 += i6 * w * dz;
            }
        }
#pragma unroll
        for (int k = 0; k < TCH; ++k) {
            if ((unsigned)k < tCount) {
                const GenPair p = tab[a * TT + tFirst + k];
                Et[k] += p.ipon * (p.rc * A0 - p.ac * A1);
                Fx[k] += p.ipon * ((6.0 + sd) * p.rc * A2 - 6.0 * p.ac * A5);
                Fy[k] += p.ipon * ((6.0 + sd) * p.rc * A3 - 6.0 * p.ac * A6);
                Fz[k] += p.ipon * ((6.0 + sd) * p.rc * A4 - 6.0 * p.ac * A7);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < TCH; ++k)
        if ((unsigned)k < tCount)
            out[(size_t)(tFirst + k) * npts + pidx] =
                make_float4((float)Fx[k], (float)Fy[k], (float)Fz[k], (float)Et[k]);
}

// Some pair can reach its r < rmin branch: the full LJ per bead and probe type,
// beads in ascending index order.
template <int TCH>
__global__ __launch_bounds__(256) void k_gen_full(GenGeom g, const float4* __restrict__ act,
                                                  const unsigned* __restrict__ actType, unsigned nAct,
                                                  const GenPair* __restrict__ tab, unsigned T,
                                                  int shape, double P2, double sqrtP2,
                                                  unsigned tFirst, unsigned tCount,
                                                  float4* __restrict__ out, size_t npts) {
    const size_t pidx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (pidx >= npts) return;
    double px, py, pz;
    point_of(g, pidx, px, py, pz);
    double Et[TCH], Fx[TCH], Fy[TCH], Fz[TCH];
#pragma unroll
    for (int k = 0; k < TCH; ++k) Et[k] = Fx[k] = Fy[k] = Fz[k] = 0.0;
    const unsigned TT = T + 1;
    for (unsigned j = 0; j < nAct; ++j) {
        const float4 b = act[j];
        const unsigned a = actType[j];
        const double dx = px - b.x, dy = py - b.y, dz = pz - b.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(r2);
        const double rp2 = r2 > P2 ? r2 : P2;
        const double rp = r2 > P2 ? r : sqrtP2;
        const double w = r > 0 ? 1.0 / (r * rp) : 0.0;
#pragma unroll
        for (int k = 0; k < TCH; ++k) {
            if ((unsigned)k < tCount) {
                double E, fb;
                lj_eval(rp2, tab[a * TT + tFirst + k], shape, E, fb);
                Et[k] += E;
                Fx[k] += fb * w * dx;
                Fy[k] += fb * w * dy;
                Fz[k] += fb * w * dz;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < TCH; ++k)
        if ((unsigned)k < tCount)
            out[(size_t)(tFirst + k) * npts + pidx] =
                make_float4((float)Fx[k], (float)Fy[k], (float)Fz[k], (float)Et[k]);
}

// Grid 0: Coulomb at r' = max(r, P), charged beads in ascending index order.
__global__ __launch_bounds__(256) void k_gen_elec(GenGeom g, const float4* __restrict__ chg,
                                                  unsigned nChg, double P2, double sqrtP2,
                                                  int dielec, double fe, float4* __restrict__ out,
                                                  size_t npts) {
    const size_t pidx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (pidx >= npts) return;
    double px, py, pz;
    point_of(g, pidx, px, py, pz);
    double E = 0, Fx = 0, Fy = 0, Fz = 0;
    for (unsigned j = 0; j < nChg; ++j) {
        const float4 b = chg[j];
        const double dx = px - b.x, dy = py - b.y, dz = pz - b.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(r2);
        const double rp2 = r2 > P2 ? r2 : P2;
        const double rp = r2 > P2 ? r : sqrtP2;
        const double c = (double)b.w * fe;
        double dd, fmag;
        if (dielec == 0) {
            dd = 1.0 / rp - 1.0 / 50.0;
            fmag = c / rp;
        } else {
            dd = 1.0 / rp2 - 1.0 / 2500.0;
            fmag = 2.0 * c / rp2;
        }
        if (dd <= 0) continue;
        E += c * dd;
        if (r > 0) {
            const double w = fmag / (r * rp);
            Fx += w * dx;
            Fy += w * dy;
            Fz += w * dz;
        }
    }
    out[pidx] = make_float4((float)Fx, (float)Fy, (float)Fz, (float)E);
}

// Neighbour lists: count (fill == false) or write (fill == true) the beads
// within sqrt(nd2) of each grid point.  A wave skips a bead, all lanes
// together, when it is farther than sqrt(cullR2) > sqrt(nd2) from the bounding
// box of the wave's points; the exact test decides the rest.
__global__ __launch_bounds__(256) void k_gen_nl(GenGeom g, const float4* __restrict__ beads,
                                                const unsigned* __restrict__ idx, unsigned nb,
                                                double nd2, float cullR2, size_t npts, bool fill,
                                                uint32_t* __restrict__ cnt,
                                                const uint2* __restrict__ desc,
                                                uint32_t* __restrict__ list) {
    const size_t pidx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const bool live = pidx < npts;
    double px = 0, py = 0, pz = 0;
    if (live) point_of(g, pidx, px, py, pz);
    float lx = live ? (float)px : 3e38f, ly = live ? (float)py : 3e38f, lz = live ? (float)pz : 3e38f;
    float hx = live ? (float)px : -3e38f, hy = live ? (float)py : -3e38f, hz = live ? (float)pz : -3e38f;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        lx = fminf(lx, __shfl_xor(lx, d));
        ly = fminf(ly, __shfl_xor(ly, d));
        lz = fminf(lz, __shfl_xor(lz, d));
        hx = fmaxf(hx, __shfl_xor(hx, d));
        hy = fmaxf(hy, __shfl_xor(hy, d));
        hz = fmaxf(hz, __shfl_xor(hz, d));
    }
    if (!(lx <= hx)) return;  // no live lane in this wave
    uint32_t* dst = (fill && live) ? list + desc[pidx].y : nullptr;
    unsigned c = 0;
    for (unsigned k = 0; k < nb; ++k) {
        const float4 b = beads[k];
        const float ex = fmaxf(fmaxf(b.x - hx, lx - b.x), 0.f);
        const float ey = fmaxf(fmaxf(b.y - hy, ly - b.y), 0.f);
        const float ez = fmaxf(fmaxf(b.z - hz, lz - b.z), 0.f);
        if (ex * ex + ey * ey + ez * ez > cullR2) continue;
        const double dx = px - b.x, dy = py - b.y, dz = pz - b.z;
        if (live && dx * dx + dy * dy + dz * dz <= nd2) {
            if (dst) dst[c] = idx[k];
            ++c;
        }
    }
    if (live && !fill) cnt[pidx] = c;
}

namespace {

constexpr unsigned kScanPer = 1024;  // elements per block: 256 threads x 4

__device__ inline unsigned long long wave_incl_scan(unsigned long long x) {
    const unsigned lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long y = __shfl_up(x, d);
        if (lane >= (unsigned)d) x += y;
    }
    return x;
}

}  // namespace

// Exclusive scan of the counts into desc {count, offset} (offsets local to
// blocks of kScanPer elements) and the block totals.
__global__ __launch_bounds__(256) void k_scan_local(const uint32_t* __restrict__ cnt,
                                                    uint2* __restrict__ desc,
                                                    unsigned long long* __restrict__ bsum,
                                                    size_t n) {
    __shared__ unsigned long long ws[4];
    const size_t base = blockIdx.x * (size_t)kScanPer + threadIdx.x * 4;
    uint32_t v[4];
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? cnt[base + k] : 0u;
        s += v[k];
    }
    const unsigned long long x = wave_incl_scan(s);
    const unsigned w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) ws[w] = x;
    __syncthreads();
    unsigned long long off = 0;
    for (unsigned i = 0; i < w; ++i) off += ws[i];
    unsigned long long e = off + x - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) desc[base + k] = make_uint2(v[k], (uint32_t)e);
        e += v[k];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = off + x;
}

// One block: block totals -> exclusive block offsets, and the grand total.
__global__ __launch_bounds__(256) void k_scan_top(unsigned long long* __restrict__ bsum,
                                                  unsigned nb, unsigned long long* __restrict__ total) {
    __shared__ unsigned long long ws[4];
    __shared__ unsigned long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const unsigned w = threadIdx.x >> 6;
    for (unsigned base = 0; base < nb; base += 256) {
        const unsigned i = base + threadIdx.x;
        const unsigned long long s = i < nb ? bsum[i] : 0ull;
        const unsigned long long x = wave_incl_scan(s);
        if ((threadIdx.x & 63) == 63) ws[w] = x;
        __syncthreads();
        unsigned long long off = carry;
        for (unsigned k = 0; k < w; ++k) off += ws[k];
        __syncthreads();
        if (i < nb) bsum[i] = off + x - s;
        if (threadIdx.x == 255) carry = off + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(uint2* __restrict__ desc,
                                                  const unsigned long long* __restrict__ boff,
                                                  size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) desc[i].y += (uint32_t)boff[i / kScanPer];
}

namespace {

constexpr int kTypeChunk = 12;

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_gen_grid(const GenGeom& g, const GenSrc& s, const GenPair* tab, unsigned T,
                           int shape, bool linear, double P2, double sqrtP2, int dielec, double fe,
                           float4* out, hipStream_t st) {
    const size_t npts = (size_t)g.nx * g.ny * g.nz;
    const unsigned nbk = blocks_for(npts);
    hipLaunchKernelGGL(k_gen_elec, dim3(nbk), dim3(256), 0, st, g, s.chg, s.nChg, P2, sqrtP2,
                       dielec, fe, out, npts);
    for (unsigned t0 = 1; t0 <= T; t0 += kTypeChunk) {
        const unsigned tc = std::min<unsigned>(kTypeChunk, T + 1 - t0);
        if (linear)
            hipLaunchKernelGGL(k_gen_lin<kTypeChunk>, dim3(nbk), dim3(256), 0, st, g, s.sorted,
                               s.typeStart, tab, T, shape, P2, sqrtP2, t0, tc, out, npts);
        else
            hipLaunchKernelGGL(k_gen_full<kTypeChunk>, dim3(nbk), dim3(256), 0, st, g, s.act,
                               s.actType, s.nAct, tab, T, shape, P2, sqrtP2, t0, tc, out, npts);
    }
    return hipGetLastError();
}

size_t gen_scan_scratch(size_t npts) {
    return sizeof(unsigned long long) * ((npts + kScanPer - 1) / kScanPer + 1);
}

hipError_t launch_gen_nl_count(const GenGeom& g, const float4* beads, const unsigned* idx,
                               unsigned nb, double nd2, float cullR2, uint32_t* cnt, uint2* desc,
                               unsigned long long* scratch, hipStream_t st) {
    const size_t npts = (size_t)g.nx * g.ny * g.nz;
    const unsigned nbk = blocks_for(npts);
    hipLaunchKernelGGL(k_gen_nl, dim3(nbk), dim3(256), 0, st, g, beads, idx, nb, nd2, cullR2, npts,
                       false, cnt, static_cast<const uint2*>(nullptr),
                       static_cast<uint32_t*>(nullptr));
    const unsigned nsb = (unsigned)((npts + kScanPer - 1) / kScanPer);
    hipLaunchKernelGGL(k_scan_local, dim3(nsb), dim3(256), 0, st, cnt, desc, scratch, npts);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, scratch, nsb, scratch + nsb);
    hipLaunchKernelGGL(k_scan_add, dim3(nbk), dim3(256), 0, st, desc, scratch, npts);
    return hipGetLastError();
}

hipError_t launch_gen_nl_fill(const GenGeom& g, const float4* beads, const unsigned* idx,
                              unsigned nb, double nd2, float cullR2, const uint2* desc,
                              uint32_t* list, hipStream_t st) {
    const size_t npts = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_gen_nl, dim3(blocks_for(npts)), dim3(256), 0, st, g, beads, idx, nb, nd2,
                       cullR2, npts, true, static_cast<uint32_t*>(nullptr), desc, list);
    return hipGetLastError();
}

}  // namespace att
