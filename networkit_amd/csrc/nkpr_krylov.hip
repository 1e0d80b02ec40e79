// Vector kernels of the FP32 GMRES correction solves used by the mixed-precision path.
//
// The Krylov basis V is one buffer of (m + 1) FP32 vectors of length n at a stride padded to
// 64 floats, so every vector starts on a 256-byte boundary and the kernels stream it with
// 16-byte loads.  Dot products accumulate in FP64 and are reduced in a fixed order (block
// partials, then one block), so a run is reproducible.  Up to kDotGroup vectors are handled
// per launch (the vector count is a template parameter): the vector w is read once per group
// instead of once per basis vector.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "nkpr_internal.hpp"

namespace nkpr {
namespace {

constexpr int kThreads = 256;
constexpr int kDotBlocks = 1024;

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ f32x4 ld4(const float *p, uint32_t q) {
    return __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p) + q);
}

__device__ __forceinline__ double dot4(f32x4 a, f32x4 b) {
    return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
}

// Block partials of K accumulators: partials[block * kDotStride + i].
template <int K>
__device__ __forceinline__ void block_partials(const double (&acc)[K], double *partials) {
    __shared__ double s_red[K][kThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const double v = wave_sum_d(acc[i]);
        if (lane == 0) s_red[i][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double r = 0.0;
        for (int k = 0; k < kThreads / 64; ++k) r += s_red[threadIdx.x][k];
        partials[(uint64_t)blockIdx.x * kDotStride + threadIdx.x] = r;
    }
}

// out[i] = sum over blocks of partials[b][i], i < count, in a fixed order (a strided sum per
// thread, then the waves in order), so that every run gives the same bits.
__device__ void final_sums(const double *partials, int blocks, uint32_t count, double *out) {
    __shared__ double s_red[kThreads / 64];
    for (uint32_t i = 0; i < count; ++i) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < blocks; b += kThreads) acc += partials[(uint64_t)b * kDotStride + i];
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

__global__ __launch_bounds__(kThreads) void dot_final_kernel(const double *partials, int blocks,
                                                             uint32_t count, double *out) {
    final_sums(partials, blocks, count, out);
}

// partials[b][i] = sum over this block's elements of w[u] * V[i][u], i < C.  n4 = n / 4 whole
// quads; block 0 adds the n % 4 tail elements.
// SELF: V[C - 1] is w itself, whose square is taken from the loaded w (one pass fewer).
template <int C, bool SELF>
__global__ __launch_bounds__(kThreads) void multidot_kernel(const float *__restrict__ w,
                                                            const float *__restrict__ V,
                                                            uint64_t ld, uint32_t n4, uint32_t n,
                                                            double *partials) {
    constexpr int kRead = SELF ? C - 1 : C;
    double acc[C];
#pragma unroll
    for (int i = 0; i < C; ++i) acc[i] = 0.0;
    for (uint32_t q = blockIdx.x * kThreads + threadIdx.x; q < n4; q += gridDim.x * kThreads) {
        const f32x4 a = ld4(w, q);
#pragma unroll
        for (int i = 0; i < kRead; ++i) acc[i] += dot4(a, ld4(V + i * ld, q));
        if (SELF) acc[C - 1] += dot4(a, a);
    }
    if (blockIdx.x == 0)
        for (uint32_t u = 4 * n4 + threadIdx.x; u < n; u += kThreads)
#pragma unroll
            for (int i = 0; i < C; ++i) acc[i] += (double)w[u] * (double)V[i * ld + u];
    block_partials<C>(acc, partials);
}

// w -= sum_i c[i] V[i], then partials of V[i] . w (i < C) and of w . w (entry C): the
// projection step of classical Gram-Schmidt fused with the dots of the next pass.
template <int C>
__global__ __launch_bounds__(kThreads) void axpy_dot_kernel(float *__restrict__ w,
                                                            const float *__restrict__ V,
                                                            uint64_t ld, uint32_t n4, uint32_t n,
                                                            Coefs c, double *partials) {
    double acc[C + 1];
#pragma unroll
    for (int i = 0; i <= C; ++i) acc[i] = 0.0;
    for (uint32_t q = blockIdx.x * kThreads + threadIdx.x; q < n4; q += gridDim.x * kThreads) {
        f32x4 v[C];
        f32x4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < C; ++i) {
            v[i] = ld4(V + i * ld, q);
            s += c.c[i] * v[i];
        }
        const f32x4 x = ld4(w, q) - s;
        reinterpret_cast<f32x4 *>(w)[q] = x;
#pragma unroll
        for (int i = 0; i < C; ++i) acc[i] += dot4(v[i], x);
        acc[C] += dot4(x, x);
    }
    if (blockIdx.x == 0)
        for (uint32_t u = 4 * n4 + threadIdx.x; u < n; u += kThreads) {
            float s = 0.f;
#pragma unroll
            for (int i = 0; i < C; ++i) s += c.c[i] * V[i * ld + u];
            const float x = w[u] - s;
            w[u] = x;
#pragma unroll
            for (int i = 0; i < C; ++i) acc[i] += (double)V[i * ld + u] * (double)x;
            acc[C] += (double)x * (double)x;
        }
    block_partials<C + 1>(acc, partials);
}

// mode 0: w -= sum_i c[i] V[i];  mode 1: w = sum_i ...;  mode 2: w += sum_i ...
template <int C>
__global__ __launch_bounds__(kThreads) void multiaxpy_kernel(float *__restrict__ w,
                                                             const float *__restrict__ V,
                                                             uint64_t ld, uint32_t n4, uint32_t n,
                                                             Coefs c, int mode) {
    for (uint32_t q = blockIdx.x * kThreads + threadIdx.x; q < n4; q += gridDim.x * kThreads) {
        f32x4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < C; ++i) s += c.c[i] * ld4(V + i * ld, q);
        f32x4 *w4 = reinterpret_cast<f32x4 *>(w) + q;
        if (mode == 1) *w4 = s;
        else if (mode == 2) *w4 = *w4 + s;
        else *w4 = *w4 - s;
    }
    if (blockIdx.x == 0)
        for (uint32_t u = 4 * n4 + threadIdx.x; u < n; u += kThreads) {
            float s = 0.f;
#pragma unroll
            for (int i = 0; i < C; ++i) s += c.c[i] * V[i * ld + u];
            if (mode == 1) w[u] = s;
            else if (mode == 2) w[u] += s;
            else w[u] -= s;
        }
}

// w -= sum_i c_i V[i] with the projection coefficients of a Gram-Schmidt step formed on the
// device: c_i = sig_i (sig_i d[i]) from the dots d in device memory (the host's FP64 operations,
// rounded to FP32 as the host rounds them), so that the projection needs no host read of the
// dots.  sig holds the scales of the whole basis; this launch covers vectors first .. first + C.
template <int C>
__global__ __launch_bounds__(kThreads) void multiaxpy_dev_kernel(float *__restrict__ w,
                                                                 const float *__restrict__ V,
                                                                 uint64_t ld, uint32_t n4, uint32_t n,
                                                                 const double *__restrict__ d,
                                                                 Scales sig, uint32_t first) {
    float c[C];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const double si = sig.s[first + i];
        c[i] = (float)(si * (si * d[first + i]));
    }
    for (uint32_t q = blockIdx.x * kThreads + threadIdx.x; q < n4; q += gridDim.x * kThreads) {
        f32x4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < C; ++i) s += c[i] * ld4(V + i * ld, q);
        f32x4 *w4 = reinterpret_cast<f32x4 *>(w) + q;
        *w4 = *w4 - s;
    }
    if (blockIdx.x == 0)
        for (uint32_t u = 4 * n4 + threadIdx.x; u < n; u += kThreads) {
            float s = 0.f;
#pragma unroll
            for (int i = 0; i < C; ++i) s += c[i] * V[i * ld + u];
            w[u] -= s;
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

unsigned grid_for(uint64_t n) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kThreads - 1) / kThreads, 8192));
}

void check_aligned(const void *p, uint64_t ld) {
    NKPR_CHECK(((uintptr_t)p & 15) == 0 && ld % 4 == 0, "Krylov vectors must be 16-byte aligned");
}

// Calls f(std::integral_constant<int, C>) for C = count (1 <= count <= kDotGroup).
template <typename F>
void with_count(uint32_t count, F &&f) {
    switch (count) {
#define NKPR_CASE(C) \
    case C: f(std::integral_constant<int, C>{}); break;
        NKPR_CASE(1) NKPR_CASE(2) NKPR_CASE(3) NKPR_CASE(4) NKPR_CASE(5) NKPR_CASE(6)
        NKPR_CASE(7) NKPR_CASE(8) NKPR_CASE(9) NKPR_CASE(10) NKPR_CASE(11) NKPR_CASE(12)
        NKPR_CASE(13) NKPR_CASE(14) NKPR_CASE(15) NKPR_CASE(16) NKPR_CASE(17) NKPR_CASE(18)
        NKPR_CASE(19) NKPR_CASE(20) NKPR_CASE(21) NKPR_CASE(22) NKPR_CASE(23) NKPR_CASE(24)
#undef NKPR_CASE
        default: throw Error(NKPR_ERR_INVALID, "Krylov group size out of range");
    }
}
static_assert(kDotGroup == 24, "with_count covers 1..24");

}  // namespace

void launch_multidot(const float *w, const float *V, uint64_t ld, uint32_t count, uint32_t n,
                     double *partials, double *out, hipStream_t s) {
    check_aligned(w, ld);
    check_aligned(V, ld);
    const uint32_t n4 = n / 4;
    const unsigned blocks = std::min<unsigned>(kDotBlocks, grid_for(n4));
    for (uint32_t first = 0; first < count; first += kDotGroup) {
        const uint32_t c = std::min<uint32_t>(kDotGroup, count - first);
        // the group's last vector is w itself (the norm of a new Krylov vector)
        const bool self = V + (uint64_t)(first + c - 1) * ld == w;
        with_count(c, [&](auto C) {
            if (self)
                hipLaunchKernelGGL((multidot_kernel<decltype(C)::value, true>), dim3(blocks), dim3(kThreads),
                                   0, s, w, V + first * ld, ld, n4, n, partials);
            else
                hipLaunchKernelGGL((multidot_kernel<decltype(C)::value, false>), dim3(blocks), dim3(kThreads),
                                   0, s, w, V + first * ld, ld, n4, n, partials);
        });
        NKPR_HIP(hipGetLastError());
        hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, partials,
                           (int)blocks, c, out + first);
        NKPR_HIP(hipGetLastError());
    }
}

void launch_multiaxpy(float *w, const float *V, uint64_t ld, uint32_t count, uint32_t n,
                      const double *coef, bool assign, hipStream_t s) {
    check_aligned(w, ld);
    check_aligned(V, ld);
    const uint32_t n4 = n / 4;
    for (uint32_t first = 0; first < count; first += kDotGroup) {
        const uint32_t c = std::min<uint32_t>(kDotGroup, count - first);
        Coefs cf{};
        for (uint32_t i = 0; i < c; ++i) cf.c[i] = (float)coef[first + i];
        const int mode = assign ? (first == 0 ? 1 : 2) : 0;
        with_count(c, [&](auto C) {
            hipLaunchKernelGGL(multiaxpy_kernel<decltype(C)::value>, dim3(grid_for(n4)), dim3(kThreads), 0,
                               s, w, V + first * ld, ld, n4, n, cf, mode);
        });
        NKPR_HIP(hipGetLastError());
    }
}

void launch_gs_project(float *w, const float *V, uint64_t ld, uint32_t count, uint32_t n,
                       const double *d, const double *sig, hipStream_t s) {
    NKPR_CHECK(count >= 1 && count <= (uint32_t)kMaxKrylov, "Gram-Schmidt step size out of range");
    check_aligned(w, ld);
    check_aligned(V, ld);
    Scales sc{};
    for (uint32_t i = 0; i < count; ++i) sc.s[i] = sig[i];
    const uint32_t n4 = n / 4;
    for (uint32_t first = 0; first < count; first += kDotGroup) {
        const uint32_t c = std::min<uint32_t>(kDotGroup, count - first);
        with_count(c, [&](auto C) {
            hipLaunchKernelGGL(multiaxpy_dev_kernel<decltype(C)::value>, dim3(grid_for(n4)), dim3(kThreads),
                               0, s, w, V + first * ld, ld, n4, n, d, sc, first);
        });
        NKPR_HIP(hipGetLastError());
    }
}

void launch_axpy_dot(float *w, const float *V, uint64_t ld, uint32_t count, uint32_t n,
                     const double *coef, double *partials, double *out, hipStream_t s) {
    if (count > (uint32_t)kDotGroup) {
        // more vectors than one launch holds: the projection, then the dots
        launch_multiaxpy(w, V, ld, count, n, coef, false, s);
        launch_multidot(w, V, ld, count, n, partials, out, s);
        launch_multidot(w, w, ld, 1, n, partials, out + count, s);
        return;
    }
    check_aligned(w, ld);
    check_aligned(V, ld);
    const uint32_t n4 = n / 4;
    const unsigned blocks = std::min<unsigned>(kDotBlocks, grid_for(n4));
    Coefs cf{};
    for (uint32_t i = 0; i < count; ++i) cf.c[i] = (float)coef[i];
    with_count(count, [&](auto C) {
        hipLaunchKernelGGL(axpy_dot_kernel<decltype(C)::value>, dim3(blocks), dim3(kThreads), 0, s, w, V,
                           ld, n4, n, cf, partials);
    });
    NKPR_HIP(hipGetLastError());
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, partials, (int)blocks,
                       count + 1, out);
    NKPR_HIP(hipGetLastError());
}

void launch_axpby(const float *x, const float *y, float *out, uint32_t n, float a, float b,
                  hipStream_t s) {
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, x, y, out, n, a, b);
    NKPR_HIP(hipGetLastError());
}

}  // namespace nkpr
