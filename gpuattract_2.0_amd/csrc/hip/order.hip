// Spatial pose order for the persistent scoring kernels.
//
// Poses are sorted by the Morton code of their translation inside the poses'
// bounding box (stable radix sort: a deterministic order).  The adjoint kernel
// gives each XCD a contiguous range of that order, so the poses one XCD holds
// at a time lie close together and share the grid lines in its L2; results are
// written to each pose's own slot, so the order changes no per-pose result.
#include <rocprim/device/device_radix_sort.hpp>

#include "device.h"
#include "internal.h"

namespace att {

namespace {

__device__ inline unsigned spread10(unsigned v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

}  // namespace

constexpr unsigned kBoxBlocks = 64;  // workgroups of the bounding-box pass
constexpr size_t kBoxBytes = 2048;   // their partial boxes, at the front of the scratch
static_assert(kBoxBlocks <= 64 && kBoxBlocks * 6 * sizeof(float) <= kBoxBytes, "box scratch");

// Partial bounding boxes of the translations (dofs[i*stride + 0..2]): block b
// writes the min and max over its grid-strided share to part[b*6 .. b*6+5].
__global__ __launch_bounds__(256) void k_pose_bbox(const float* __restrict__ dofs, unsigned stride,
                                                   unsigned n, float* __restrict__ part) {
    __shared__ float red[6][4];
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k) {
            const float v = dofs[(size_t)i * stride + k];
            lo[k] = fminf(lo[k], v);
            hi[k] = fmaxf(hi[k], v);
        }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], d));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], d));
        }
    const unsigned w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) {
            red[k][w] = lo[k];
            red[3 + k][w] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const unsigned k = threadIdx.x;
        float v = red[k][0];
        for (unsigned j = 1; j < 4; ++j) v = k < 3 ? fminf(v, red[k][j]) : fmaxf(v, red[k][j]);
        part[blockIdx.x * 6 + k] = v;
    }
}

// Morton key of each translation in the bounding box; every block first
// reduces the kBoxBlocks partial boxes (wave 0).
__global__ __launch_bounds__(256) void k_pose_keys(const float* __restrict__ dofs, unsigned stride,
                                                   unsigned n, const float* __restrict__ part,
                                                   unsigned* __restrict__ keys,
                                                   unsigned* __restrict__ idx) {
    __shared__ float box[6];
    if (threadIdx.x < 64) {
        float v[6];
        for (int k = 0; k < 6; ++k) v[k] = threadIdx.x < kBoxBlocks ? part[threadIdx.x * 6 + k]
                                                                    : (k < 3 ? 3e38f : -3e38f);
#pragma unroll
        for (int d = 1; d < 64; d <<= 1)
            for (int k = 0; k < 6; ++k) {
                const float o = __shfl_xor(v[k], d);
                v[k] = k < 3 ? fminf(v[k], o) : fmaxf(v[k], o);
            }
        if (threadIdx.x == 0)
            for (int k = 0; k < 6; ++k) box[k] = v[k];
    }
    __syncthreads();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned q[3];
    for (int k = 0; k < 3; ++k) {
        const float ext = box[3 + k] - box[k];
        const float t = ext > 0.f ? (dofs[(size_t)i * stride + k] - box[k]) / ext : 0.f;
        q[k] = (unsigned)fminf(fmaxf(t * 1024.f, 0.f), 1023.f);
    }
    keys[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
    idx[i] = i;
}

size_t pose_order_scratch(unsigned n) {
    size_t temp = 0;
    (void)rocprim::radix_sort_pairs(nullptr, temp, (unsigned*)nullptr, (unsigned*)nullptr,
                                    (unsigned*)nullptr, (unsigned*)nullptr, n, 0, 30);
    return kBoxBytes + 3 * sizeof(unsigned) * (size_t)n + ((temp + 255) & ~(size_t)255);
}

hipError_t launch_pose_order(const float* dofs, unsigned stride, unsigned n, void* scratch,
                             size_t scratchBytes, unsigned* order, hipStream_t s) {
    if (n == 0) return hipSuccess;
    char* p = static_cast<char*>(scratch);
    float* part = reinterpret_cast<float*>(p);
    unsigned* keysIn = reinterpret_cast<unsigned*>(p + kBoxBytes);
    unsigned* keysOut = keysIn + n;
    unsigned* idxIn = keysOut + n;
    void* temp = p + kBoxBytes + 3 * sizeof(unsigned) * (size_t)n;
    size_t tempBytes = scratchBytes - (kBoxBytes + 3 * sizeof(unsigned) * (size_t)n);
    hipLaunchKernelGGL(k_pose_bbox, dim3(kBoxBlocks), dim3(256), 0, s, dofs, stride, n, part);
    hipLaunchKernelGGL(k_pose_keys, dim3((n + 255) / 256), dim3(256), 0, s, dofs, stride, n, part,
                       keysIn, idxIn);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return rocprim::radix_sort_pairs(temp, tempBytes, keysIn, keysOut, idxIn, order, n, 0, 30, s);
}

}  // namespace att
