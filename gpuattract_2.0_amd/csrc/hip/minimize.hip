// Device-resident minimization of many poses with VA13 (va13.h), driven by
// the adjoint: each round scores the points the running minimizers asked
// for (the adjoint reads the number of requests from device memory), then
// k_min_step resumes every minimizer with its value and gradient and appends
// those still running, with their next point, to the next round's list.
// Minimizer states stay in HBM for the whole run; the host only reads the
// number of running poses every few rounds.
//
// The minimizer's variables are (angles, position, receptor modes, ligand
// modes) in double (apps/em/meta.h:90-170); a requested point is rounded to a
// float DOF record before scoring, as the reference builds DOF_6D<float>
// requests; the gradient is the exact dE/dvariables of the adjoint.
#include "device.h"
#include "internal.h"
#include "va13.h"

namespace att {

namespace {

// DOF record (pos, ang, modesRec[10], modesLig[10]; or pos, ang) <-> variables.
__device__ inline void dof_to_vars(const float* d, int Mr, int Ml, double* x) {
    for (int k = 0; k < 3; ++k) {
        x[k] = d[3 + k];
        x[3 + k] = d[k];
    }
    for (int m = 0; m < Mr; ++m) x[6 + m] = d[6 + m];
    for (int m = 0; m < Ml; ++m) x[6 + Mr + m] = d[16 + m];
}

__device__ inline void vars_to_dof(const double* x, int Mr, int Ml, bool modes, float* d) {
    for (int k = 0; k < 3; ++k) {
        d[3 + k] = (float)x[k];
        d[k] = (float)x[3 + k];
    }
    if (modes) {
        for (int m = 0; m < ATT_MODES_MAX; ++m) {
            d[6 + m] = m < Mr ? (float)x[6 + m] : 0.f;
            d[16 + m] = m < Ml ? (float)x[6 + Mr + m] : 0.f;
        }
    }
}

__device__ inline void result_to_grad(const float* r, int Mr, int Ml, bool modes, double* g) {
    for (int k = 0; k < 3; ++k) {
        g[k] = r[4 + k];      // dE/dangles
        g[3 + k] = r[1 + k];  // dE/dposition
    }
    if (modes) {
        for (int m = 0; m < Mr; ++m) g[6 + m] = r[7 + m];
        for (int m = 0; m < Ml; ++m) g[6 + Mr + m] = r[17 + m];
    }
}

}  // namespace

__global__ void k_min_start(MinRun R, const float* __restrict__ dofs) {
    const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) R.count[0] = R.npose;
    if (p >= R.npose) return;
    const VaView V{R.st + p, R.npose, R.dim};
    double x[kVaMaxDim], req[kVaMaxDim];
    dof_to_vars(dofs + (size_t)R.ds * p, R.Mr, R.Ml, x);
    for (int i = 0; i < R.dim; ++i) V.v(kVaX, i) = x[i];
    VaState s{};
    s.pc = kVaFresh;
    va_resume(s, V, R.dim, R.ivmax, 0.0, nullptr, req);
    R.sc[p] = s;
    R.list[p] = p;
    vars_to_dof(req, R.Mr, R.Ml, R.modes, R.req + (size_t)R.ds * p);
}

__global__ void k_min_step(MinRun R, const unsigned* __restrict__ listIn,
                           const unsigned* __restrict__ countIn, const float* __restrict__ reqIn,
                           const float* __restrict__ res) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= countIn[0]) return;
    const unsigned p = listIn[t];
    const VaView V{R.st + p, R.npose, R.dim};
    const float* r = res + (size_t)R.os * t;
    double g[kVaMaxDim], req[kVaMaxDim];
    result_to_grad(r, R.Mr, R.Ml, R.modes, g);
    VaState s = R.sc[p];
    const bool more = va_resume(s, V, R.dim, R.ivmax, (double)r[0], g, req);
    R.sc[p] = s;
    if (more) {
        const unsigned slot = atomicAdd(R.count, 1u);
        R.list[slot] = p;
        vars_to_dof(req, R.Mr, R.Ml, R.modes, R.req + (size_t)R.ds * slot);
    } else {
        // the point just scored is the result (the reference reports its
        // solver's last evaluated state and objective)
        const float* q = reqIn + (size_t)R.ds * t;
        float* o = R.dofsOut + (size_t)R.ds * p;
        for (unsigned k = 0; k < R.ds; ++k) o[k] = q[k];
        R.eLast[p] = r[0];
        if (R.eBest) R.eBest[p] = (float)s.gesa;
        if (R.nfun) R.nfun[p] = s.nfun;
    }
}

__global__ void k_min_reset(unsigned* count) {
    if (threadIdx.x == 0) count[0] = 0u;
}

hipError_t launch_min_start(const MinRun& R, const float* dofs, hipStream_t s) {
    hipLaunchKernelGGL(k_min_start, dim3((R.npose + 127) / 128), dim3(128), 0, s, R, dofs);
    return hipGetLastError();
}

hipError_t launch_min_step(const MinRun& R, const unsigned* listIn, const unsigned* countIn,
                           const float* reqIn, const float* res, hipStream_t s) {
    hipLaunchKernelGGL(k_min_reset, dim3(1), dim3(64), 0, s, R.count);
    hipLaunchKernelGGL(k_min_step, dim3((R.npose + 127) / 128), dim3(128), 0, s, R, listIn, countIn,
                       reqIn, res);
    return hipGetLastError();
}

}  // namespace att
