// Pull-sweep kernels of the PageRank iteration for gfx950 (CDNA4, wave64).
//
// One sweep computes, for every existing node u (compact, relabelled id),
//     x'[u] = d * sum_{v in in(u)} c_vu * g[v]  +  rhs[u]
// which is the loop body of networkit/cpp/centrality/PageRank.cpp:56-66
// (pr[u] += scoreData[v] * w / deg[v]; pr[u] = pr[u] * damp + teleportProb) with
//   folded (weighted graphs):      c_vu = w_vu / deg_v, g = x
//   prescaled (unweighted graphs): c_vu = 1,            g = x * (1/deg)  (written by the
//                                                        previous sweep's epilogue)
// and rhs = (1-d)/n (+ d * sinkmass / n when sinks are distributed) or a per-node vector
// (correction solves of the mixed-precision refinement).
//
// The same kernel produces the convergence inputs of PageRank.cpp:34-51 (sum |x'-x| and
// sum (x'-x)^2 in FP64), the score sum used by the normalisation at :75-80 and the sink mass
// used at :81-92, as per-block partials.  A one-block finalize kernel then adds the
// partials in a fixed order, finishes the rows that were split across blocks, applies the
// stopping rule and raises a device-side `done` flag that turns later sweeps of the batch
// into no-ops, so the host reads the state back only once per batch of sweeps.
#include <hip/hip_runtime.h>

#include <cmath>

#include "nkpr_internal.hpp"

namespace nkpr {

namespace {

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum of one value per thread over a block of BLOCK threads; result valid in thread 0.
template <int BLOCK, typename T> __device__ __forceinline__ T block_sum(T v, T *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < BLOCK / 64; ++i) r += scratch[i];
    }
    return r;
}

struct Acc4 {
    double l1 = 0.0, l2 = 0.0, sum = 0.0, sink = 0.0;
};

// Row epilogue: the fused "pr[u] = pr[u]*damp + teleport" plus everything the loop needs
// from the new value (residual terms, score sum, sink mass, prescaled copy, FP32 residual).
template <typename T, bool BASEVEC, bool RESID>
__device__ __forceinline__ void row_epilogue(const SweepArgs<T> &a, uint32_t u, T acc, T damp,
                                             T c, Acc4 &s) {
    T y = damp * acc + c;
    if (BASEVEC) y += a.base_vec[u];
    const T xo = a.xold[u];
    a.xnew[u] = y;
    if (a.xhat_new != nullptr && u < a.sink_begin) a.xhat_new[u] = y * a.inv[u];
    const double dif = (double)y - (double)xo;
    if (RESID) a.r32[u] = (float)dif;
    s.l1 += fabs(dif);
    s.l2 += dif * dif;
    s.sum += (double)y;
    if (u >= a.sink_begin) s.sink += (double)y;
}

template <typename T> __device__ __forceinline__ T rhs_const(const SweepArgs<T> &a) {
    double c = a.tele;
    if (a.distribute) c += a.damp * a.ctl->sink_in / (double)a.n;
    return (T)c;
}

constexpr int kPer = kCap / kBlock;  // staged in-edges per thread in a stream item

// Sum of the products of in-edges [e0, e0+cnt) by the whole block (vector rows and split
// chunks).  Four independent gathers per thread per step keep enough loads in flight.
template <typename T, bool FOLDED>
__device__ __forceinline__ T block_row_sum(const SweepArgs<T> &a, uint64_t e0, uint32_t cnt,
                                           T *red) {
    T acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
    const uint32_t tid = threadIdx.x;
    uint32_t i = tid;
    for (; i + 3 * kBlock < cnt; i += 4 * kBlock) {
        const int32_t s0 = a.src[e0 + i], s1 = a.src[e0 + i + kBlock];
        const int32_t s2 = a.src[e0 + i + 2 * kBlock], s3 = a.src[e0 + i + 3 * kBlock];
        T v0 = a.xg[s0], v1 = a.xg[s1], v2 = a.xg[s2], v3 = a.xg[s3];
        if (FOLDED) {
            v0 *= a.coef[e0 + i];
            v1 *= a.coef[e0 + i + kBlock];
            v2 *= a.coef[e0 + i + 2 * kBlock];
            v3 *= a.coef[e0 + i + 3 * kBlock];
        }
        acc0 += v0;
        acc1 += v1;
        acc2 += v2;
        acc3 += v3;
    }
    for (; i < cnt; i += kBlock) {
        T v = a.xg[a.src[e0 + i]];
        if (FOLDED) v *= a.coef[e0 + i];
        acc0 += v;
    }
    T acc = (acc0 + acc1) + (acc2 + acc3);
    return block_sum<kBlock>(acc, red);
}

template <typename T, bool FOLDED, bool BASEVEC, bool RESID>
__global__ __launch_bounds__(kBlock) void sweep_kernel(SweepArgs<T> a) {
    if (a.ctl->done) return;  // converged earlier in this batch: no-op

    __shared__ T s_prod[kCap];
    __shared__ uint32_t s_ptr[kRowsMax + 1];
    __shared__ T s_sum[kRowsMax];
    __shared__ T s_redT[kBlock / 64];
    __shared__ double s_red[kBlock / 64];

    const T damp = (T)a.damp;
    const T c = rhs_const(a);
    const uint32_t tid = threadIdx.x;
    Acc4 st;

    for (uint32_t it = blockIdx.x; it < a.n_items; it += gridDim.x) {
        const uint2 item = a.items[it];
        if (item.y & kSplitFlag) {
            // one chunk of a row that is too long for one block
            const uint32_t row = item.x, chunk = item.y & ~kSplitFlag;
            const uint64_t rb = a.colptr[row], re = a.colptr[row + 1];
            const uint64_t e0 = rb + (uint64_t)chunk * kChunk;
            const uint64_t e1 = (e0 + kChunk < re) ? e0 + kChunk : re;
            const T part = block_row_sum<T, FOLDED>(a, e0, (uint32_t)(e1 - e0), s_redT);
            if (tid == 0) a.split_partial[it] = part;
            continue;
        }
        const uint32_t r0 = item.x, r1 = item.y, R = r1 - r0;
        const uint64_t e0 = a.colptr[r0];
        const uint32_t cnt = (uint32_t)(a.colptr[r1] - e0);
        if (R == 1 && cnt > (uint32_t)kCap) {
            // a heavy row: the whole block reduces it
            const T acc = block_row_sum<T, FOLDED>(a, e0, cnt, s_redT);
            if (tid == 0) row_epilogue<T, BASEVEC, RESID>(a, r0, acc, damp, c, st);
            continue;
        }
        // stream item: stage the products of all in-edges of rows [r0, r1) in LDS
        for (uint32_t i = tid; i <= R; i += kBlock) s_ptr[i] = (uint32_t)(a.colptr[r0 + i] - e0);
        if (cnt > 0) {
            int32_t sidx[kPer];
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const uint32_t i = tid + k * kBlock;
                sidx[k] = a.src[e0 + (i < cnt ? i : cnt - 1)];
            }
            T v[kPer];
#pragma unroll
            for (int k = 0; k < kPer; ++k) v[k] = a.xg[sidx[k]];
            if (FOLDED) {
#pragma unroll
                for (int k = 0; k < kPer; ++k) {
                    const uint32_t i = tid + k * kBlock;
                    v[k] *= a.coef[e0 + (i < cnt ? i : cnt - 1)];
                }
            }
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const uint32_t i = tid + k * kBlock;
                if (i < cnt) s_prod[i] = v[k];
            }
        }
        __syncthreads();
        // row sums: G consecutive lanes per row, G ~ average row length (power of two <= 64)
        const uint32_t avg = R ? cnt / R : 0;
        const int lg = avg >= 64 ? 6 : (avg >= 32 ? 5 : (avg >= 16 ? 4 : (avg >= 8 ? 3 : (avg >= 4 ? 2 : (avg >= 2 ? 1 : 0)))));
        const uint32_t G = 1u << lg;
        const uint32_t per_pass = kBlock >> lg;
        const uint32_t lane = tid & (G - 1);
        for (uint32_t base = 0; base < R; base += per_pass) {
            const uint32_t r = base + (tid >> lg);
            T acc = 0;
            if (r < R) {
                const uint32_t b = s_ptr[r], e = s_ptr[r + 1];
                for (uint32_t i = b + lane; i < e; i += G) acc += s_prod[i];
            }
            for (uint32_t off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
            if (r < R && lane == 0) s_sum[r] = acc;
        }
        __syncthreads();
        for (uint32_t r = tid; r < R; r += kBlock)
            row_epilogue<T, BASEVEC, RESID>(a, r0 + r, s_sum[r], damp, c, st);
        __syncthreads();
    }

    // per-block partials (l1, l2, sum, sink) in FP64
    double v;
    v = block_sum<kBlock>(st.l1, s_red);
    if (tid == 0) a.partials[blockIdx.x * 4 + 0] = v;
    v = block_sum<kBlock>(st.l2, s_red);
    if (tid == 0) a.partials[blockIdx.x * 4 + 1] = v;
    v = block_sum<kBlock>(st.sum, s_red);
    if (tid == 0) a.partials[blockIdx.x * 4 + 2] = v;
    v = block_sum<kBlock>(st.sink, s_red);
    if (tid == 0) a.partials[blockIdx.x * 4 + 3] = v;
}

// One block: finish split rows, add all partials in a fixed order, apply the stopping rule.
template <typename T, bool BASEVEC, bool RESID>
__global__ __launch_bounds__(kFinalizeThreads) void finalize_kernel(SweepArgs<T> a,
                                                                    FinalizeArgs f) {
    PhaseCtl *ctl = a.ctl;
    if (ctl->done) return;
    __shared__ double s_red[kFinalizeThreads / 64];
    const T damp = (T)a.damp;
    const T c = rhs_const(a);
    Acc4 st;
    for (uint32_t j = threadIdx.x; j < f.n_split_rows; j += kFinalizeThreads) {
        T acc = 0;
        const uint32_t first = f.split_first[j], count = f.split_count[j];
        for (uint32_t k = 0; k < count; ++k) acc += a.split_partial[first + k];
        row_epilogue<T, BASEVEC, RESID>(a, f.split_rows[j], acc, damp, c, st);
    }
    // split-row terms first (in row order), then block partials in block order
    double l1 = st.l1, l2 = st.l2, sum = st.sum, sink = st.sink;
    double b1 = 0, b2 = 0, b3 = 0, b4 = 0;
    for (uint32_t b = threadIdx.x; b < f.n_blocks; b += kFinalizeThreads) {
        b1 += a.partials[b * 4 + 0];
        b2 += a.partials[b * 4 + 1];
        b3 += a.partials[b * 4 + 2];
        b4 += a.partials[b * 4 + 3];
    }
    l1 = block_sum<kFinalizeThreads>(l1 + b1, s_red);
    l2 = block_sum<kFinalizeThreads>(l2 + b2, s_red);
    sum = block_sum<kFinalizeThreads>(sum + b3, s_red);
    sink = block_sum<kFinalizeThreads>(sink + b4, s_red);
    if (threadIdx.x == 0) {
        const PhaseRule &r = f.rule;
        const uint64_t it = ctl->iterations + 1;
        ctl->iterations = it;
        const double prev = ctl->l1;
        ctl->prev_l1 = prev;
        ctl->l1 = l1;
        ctl->l2 = l2;
        ctl->sum = sum;
        ctl->sink_out = sink;
        ctl->sink_in = sink;  // the next sweep reads the vector just written
        const double bound = (sum > 0.0) ? 2.0 * r.damp / (1.0 - r.damp) * l1 / sum : INFINITY;
        ctl->bound = bound;
        bool stop = it >= r.max_iterations;
        if (r.kind == kRuleStep) {
            const double step = (r.norm == NKPR_L2) ? sqrt(l2) : l1;
            stop = stop || step <= r.tol;
        } else if (r.kind == kRuleErr) {
            stop = stop || bound <= r.tol;
        } else {
            stop = stop || l1 <= r.tol_abs || bound <= r.tol;
            if (it >= 4 && l1 > r.stall_ratio * prev) stop = true;
        }
        ctl->done = stop ? 1u : 0u;
    }
}

__global__ void init_ctl_kernel(PhaseCtl *ctl, const double *sink_sum) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->sink_in = sink_sum ? *sink_sum : 0.0;
        ctl->l1 = INFINITY;
        ctl->l2 = INFINITY;
        ctl->sum = 0.0;
        ctl->sink_out = 0.0;
        ctl->prev_l1 = INFINITY;
        ctl->bound = INFINITY;
        ctl->iterations = 0;
        ctl->done = 0;
        ctl->pad = 0;
    }
}

__global__ void gather_start_kernel(const double *scores_z, const int32_t *perm, uint32_t n,
                                    double *x64, float *x32) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = scores_z[perm[i]];
    if (x64) x64[i] = v;
    if (x32) x32[i] = (float)v;
}

__global__ void prescale_kernel(const double *x64, const double *inv64, double *xhat64,
                                const float *x32, const float *inv32, float *xhat32,
                                uint32_t n, uint32_t sink_begin) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sink_begin || i >= n) return;
    if (xhat64) xhat64[i] = x64[i] * inv64[i];
    if (xhat32) xhat32[i] = x32[i] * inv32[i];
}

__global__ void to_double_kernel(const float *x32, double *x64, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x64[i] = (double)x32[i];
}

__global__ void add_correction_kernel(double *x64, const float *e32, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x64[i] += (double)e32[i];
}

__global__ void copy_f32_kernel(const float *in, float *out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

__global__ void scatter_result_kernel(const double *y, const int32_t *perm, uint32_t n,
                                      double factor, double *scores_z) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) scores_z[perm[i]] = y[i] / factor;
}

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

template <typename T>
__global__ __launch_bounds__(kRedThreads) void range_sum_kernel(const T *x, uint32_t begin,
                                                               uint32_t end, double *partials) {
    __shared__ double s_red[kRedThreads / 64];
    double acc = 0.0;
    for (uint64_t i = begin + (uint64_t)blockIdx.x * kRedThreads + threadIdx.x; i < end;
         i += (uint64_t)gridDim.x * kRedThreads)
        acc += (double)x[i];
    acc = block_sum<kRedThreads>(acc, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kRedThreads) void range_sum_final_kernel(const double *partials,
                                                                      int np, double *out) {
    __shared__ double s_red[kRedThreads / 64];
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedThreads) acc += partials[i];
    acc = block_sum<kRedThreads>(acc, s_red);
    if (threadIdx.x == 0) *out = acc;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__global__ __launch_bounds__(kRedThreads) void range_max_kernel(const double *x, uint32_t n,
                                                               double *partials) {
    __shared__ double s_red[kRedThreads / 64];
    double acc = -INFINITY;
    for (uint64_t i = (uint64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kRedThreads)
        acc = fmax(acc, x[i]);
    acc = wave_max(acc);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = s_red[0];
        for (int i = 1; i < kRedThreads / 64; ++i) r = fmax(r, s_red[i]);
        partials[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(kRedThreads) void range_max_final_kernel(const double *partials,
                                                                      int np, double *out) {
    __shared__ double s_red[kRedThreads / 64];
    double acc = -INFINITY;
    for (int i = threadIdx.x; i < np; i += kRedThreads) acc = fmax(acc, partials[i]);
    acc = wave_max(acc);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = s_red[0];
        for (int i = 1; i < kRedThreads / 64; ++i) r = fmax(r, s_red[i]);
        *out = r;
    }
}

inline unsigned grid1d(uint64_t n, int threads = 256) {
    return (unsigned)((n + threads - 1) / threads);
}

template <typename T, bool FOLDED, bool BASEVEC, bool RESID>
void launch_pair(const SweepArgs<T> &a, const FinalizeArgs &f, int grid, hipStream_t s,
                 hipEvent_t ev0, hipEvent_t ev1) {
    if (ev0) NKPR_HIP(hipEventRecord(ev0, s));
    hipLaunchKernelGGL((sweep_kernel<T, FOLDED, BASEVEC, RESID>), dim3(grid), dim3(kBlock), 0, s,
                       a);
    NKPR_HIP(hipGetLastError());
    if (ev1) NKPR_HIP(hipEventRecord(ev1, s));
    hipLaunchKernelGGL((finalize_kernel<T, BASEVEC, RESID>), dim3(1), dim3(kFinalizeThreads), 0,
                       s, a, f);
    NKPR_HIP(hipGetLastError());
}

template <typename T, bool FOLDED, bool BASEVEC, bool RESID>
const void *sweep_fn() {
    return reinterpret_cast<const void *>(&sweep_kernel<T, FOLDED, BASEVEC, RESID>);
}

}  // namespace

template <typename T>
void launch_sweep(const SweepArgs<T> &a, const FinalizeArgs &f, int kind, bool folded, int grid,
                  hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    if (kind == kSweepPlain) {
        if (folded) launch_pair<T, true, false, false>(a, f, grid, s, ev0, ev1);
        else launch_pair<T, false, false, false>(a, f, grid, s, ev0, ev1);
    } else if (kind == kSweepBaseVec) {
        if (folded) launch_pair<T, true, true, false>(a, f, grid, s, ev0, ev1);
        else launch_pair<T, false, true, false>(a, f, grid, s, ev0, ev1);
    } else {
        if (folded) launch_pair<T, true, false, true>(a, f, grid, s, ev0, ev1);
        else launch_pair<T, false, false, true>(a, f, grid, s, ev0, ev1);
    }
}

template void launch_sweep<float>(const SweepArgs<float> &, const FinalizeArgs &, int, bool,
                                  int, hipStream_t, hipEvent_t, hipEvent_t);
template void launch_sweep<double>(const SweepArgs<double> &, const FinalizeArgs &, int, bool,
                                   int, hipStream_t, hipEvent_t, hipEvent_t);

int sweep_grid_size(int device, bool dbl, bool folded, int kind, uint32_t n_items) {
    const void *fn;
    if (dbl) {
        if (kind == kSweepResid) fn = folded ? sweep_fn<double, true, false, true>() : sweep_fn<double, false, false, true>();
        else if (kind == kSweepBaseVec) fn = folded ? sweep_fn<double, true, true, false>() : sweep_fn<double, false, true, false>();
        else fn = folded ? sweep_fn<double, true, false, false>() : sweep_fn<double, false, false, false>();
    } else {
        if (kind == kSweepBaseVec) fn = folded ? sweep_fn<float, true, true, false>() : sweep_fn<float, false, true, false>();
        else fn = folded ? sweep_fn<float, true, false, false>() : sweep_fn<float, false, false, false>();
    }
    int per_cu = 0;
    NKPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0));
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (per_cu < 1) per_cu = 1;
    uint64_t g = (uint64_t)per_cu * (uint64_t)cus;
    if (g > n_items) g = n_items;
    if (g < 1) g = 1;
    return (int)g;
}

void launch_gather_start(const double *scores_z, const int32_t *perm, uint32_t n, double *x64,
                         float *x32, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(gather_start_kernel, dim3(grid1d(n)), dim3(256), 0, s, scores_z, perm, n,
                       x64, x32);
    NKPR_HIP(hipGetLastError());
}

void launch_prescale(const double *x64, const double *inv64, double *xhat64, const float *x32,
                     const float *inv32, float *xhat32, uint32_t n, uint32_t sink_begin,
                     hipStream_t s) {
    if (!n || !sink_begin) return;
    hipLaunchKernelGGL(prescale_kernel, dim3(grid1d(n)), dim3(256), 0, s, x64, inv64, xhat64,
                       x32, inv32, xhat32, n, sink_begin);
    NKPR_HIP(hipGetLastError());
}

void launch_to_double(const float *x32, double *x64, uint32_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(to_double_kernel, dim3(grid1d(n)), dim3(256), 0, s, x32, x64, n);
    NKPR_HIP(hipGetLastError());
}

void launch_add_correction(double *x64, const float *e32, uint32_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(add_correction_kernel, dim3(grid1d(n)), dim3(256), 0, s, x64, e32, n);
    NKPR_HIP(hipGetLastError());
}

void launch_copy_f32(const float *in, float *out, uint32_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(copy_f32_kernel, dim3(grid1d(n)), dim3(256), 0, s, in, out, n);
    NKPR_HIP(hipGetLastError());
}

void launch_scatter_result(const double *y, const int32_t *perm, uint32_t n, double factor,
                           double *scores_z, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(scatter_result_kernel, dim3(grid1d(n)), dim3(256), 0, s, y, perm, n,
                       factor, scores_z);
    NKPR_HIP(hipGetLastError());
}

template <typename T>
void launch_range_sum(const T *x, uint32_t begin, uint32_t end, double *partials, double *out,
                      hipStream_t s) {
    hipLaunchKernelGGL(range_sum_kernel<T>, dim3(kRedBlocks), dim3(kRedThreads), 0, s, x, begin,
                       end, partials);
    NKPR_HIP(hipGetLastError());
    hipLaunchKernelGGL(range_sum_final_kernel, dim3(1), dim3(kRedThreads), 0, s, partials,
                       kRedBlocks, out);
    NKPR_HIP(hipGetLastError());
}
template void launch_range_sum<float>(const float *, uint32_t, uint32_t, double *, double *,
                                      hipStream_t);
template void launch_range_sum<double>(const double *, uint32_t, uint32_t, double *, double *,
                                       hipStream_t);

void launch_range_max(const double *x, uint32_t n, double *partials, double *out,
                      hipStream_t s) {
    hipLaunchKernelGGL(range_max_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, x, n,
                       partials);
    NKPR_HIP(hipGetLastError());
    hipLaunchKernelGGL(range_max_final_kernel, dim3(1), dim3(kRedThreads), 0, s, partials,
                       kRedBlocks, out);
    NKPR_HIP(hipGetLastError());
}

void launch_init_ctl(PhaseCtl *ctl, const double *sink_sum, hipStream_t s) {
    hipLaunchKernelGGL(init_ctl_kernel, dim3(1), dim3(64), 0, s, ctl, sink_sum);
    NKPR_HIP(hipGetLastError());
}

}  // namespace nkpr
