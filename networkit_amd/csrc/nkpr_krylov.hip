// Vector kernels of the FP32 GMRES correction solves used by the mixed-precision path.
//
// The Krylov basis V is one buffer of (m + 1) FP32 vectors of length n.  Dot products
// accumulate in FP64 and are reduced in a fixed order (block partials, then one block), so
// a run is reproducible.  Up to kDotGroup vectors are handled per launch: the vector w is
// read once per group instead of once per basis vector.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "nkpr_internal.hpp"

namespace nkpr {
namespace {

constexpr int kThreads = 256;
constexpr int kDotBlocks = 1024;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// partials[b * kDotGroup + i] = sum over this block's rows of w[u] * V[first + i][u]
__global__ __launch_bounds__(kThreads) void multidot_kernel(const float *__restrict__ w,
                                                            const float *__restrict__ V,
                                                            uint64_t stride, uint32_t count,
                                                            uint32_t n, double *partials) {
    __shared__ double s_red[kDotGroup][kThreads / 64];
    double acc[kDotGroup];
#pragma unroll
    for (int i = 0; i < kDotGroup; ++i) acc[i] = 0.0;
    for (uint64_t u = (uint64_t)blockIdx.x * kThreads + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * kThreads) {
        const double wu = (double)__builtin_nontemporal_load(w + u);
#pragma unroll
        for (int i = 0; i < kDotGroup; ++i)
            if ((uint32_t)i < count) acc[i] += wu * (double)__builtin_nontemporal_load(V + i * stride + u);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < kDotGroup; ++i) {
        const double v = wave_sum_d(acc[i]);
        if (lane == 0) s_red[i][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < kDotGroup) {
        double r = 0.0;
        for (int k = 0; k < kThreads / 64; ++k) r += s_red[threadIdx.x][k];
        partials[(uint64_t)blockIdx.x * kDotGroup + threadIdx.x] = r;
    }
}

__global__ __launch_bounds__(kThreads) void dot_final_kernel(const double *partials, int blocks,
                                                             uint32_t count, double *out) {
    __shared__ double s_red[kThreads / 64];
    for (uint32_t i = 0; i < count; ++i) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < blocks; b += kThreads) acc += partials[(uint64_t)b * kDotGroup + i];
        acc = wave_sum_d(acc);
        if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0.0;
            for (int k = 0; k < kThreads / 64; ++k) r += s_red[k];
            out[i] = r;
        }
        __syncthreads();
    }
}

struct Coefs {
    float c[kDotGroup];
};

// mode 0: w[u] -= sum_i c[i] V[i][u];  mode 1: w[u] = sum_i ...;  mode 2: w[u] += sum_i ...
__global__ __launch_bounds__(kThreads) void multiaxpy_kernel(float *__restrict__ w,
                                                             const float *__restrict__ V,
                                                             uint64_t stride, uint32_t count,
                                                             uint32_t n, Coefs c, int mode) {
    for (uint64_t u = (uint64_t)blockIdx.x * kThreads + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * kThreads) {
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < kDotGroup; ++i)
            if ((uint32_t)i < count) acc += c.c[i] * __builtin_nontemporal_load(V + i * stride + u);
        if (mode == 1) w[u] = acc;
        else if (mode == 2) w[u] += acc;
        else w[u] -= acc;
    }
}

// out[u] = a * x[u] + b * y[u]
__global__ __launch_bounds__(kThreads) void axpby_kernel(const float *__restrict__ x,
                                                         const float *__restrict__ y,
                                                         float *__restrict__ out, uint32_t n,
                                                         float a, float b) {
    for (uint64_t u = (uint64_t)blockIdx.x * kThreads + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * kThreads)
        out[u] = a * x[u] + (y ? b * y[u] : 0.f);
}

unsigned grid_for(uint32_t n) {
    return (unsigned)std::min<uint64_t>(((uint64_t)n + kThreads - 1) / kThreads, 8192);
}

}  // namespace

void launch_multidot(const float *w, const float *V, uint64_t stride, uint32_t count, uint32_t n,
                     double *partials, double *out, hipStream_t s) {
    for (uint32_t first = 0; first < count; first += kDotGroup) {
        const uint32_t c = std::min<uint32_t>(kDotGroup, count - first);
        const unsigned blocks = std::min<unsigned>(kDotBlocks, grid_for(n));
        hipLaunchKernelGGL(multidot_kernel, dim3(blocks), dim3(kThreads), 0, s, w,
                           V + first * stride, stride, c, n, partials);
        NKPR_HIP(hipGetLastError());
        hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, partials,
                           (int)blocks, c, out + first);
        NKPR_HIP(hipGetLastError());
    }
}

void launch_multiaxpy(float *w, const float *V, uint64_t stride, uint32_t count, uint32_t n,
                      const double *coef, bool assign, hipStream_t s) {
    for (uint32_t first = 0; first < count; first += kDotGroup) {
        const uint32_t c = std::min<uint32_t>(kDotGroup, count - first);
        Coefs cf{};
        for (uint32_t i = 0; i < c; ++i) cf.c[i] = (float)coef[first + i];
        hipLaunchKernelGGL(multiaxpy_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, w,
                           V + first * stride, stride, c, n, cf, assign ? (first == 0 ? 1 : 2) : 0);
        NKPR_HIP(hipGetLastError());
    }
}

void launch_axpby(const float *x, const float *y, float *out, uint32_t n, float a, float b,
                  hipStream_t s) {
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, x, y, out, n, a, b);
    NKPR_HIP(hipGetLastError());
}

}  // namespace nkpr
