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

#include <algorithm>
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

// Stores of a finished row value y (old value xo already loaded) and its reductions: the
// fused "pr[u] = pr[u]*damp + teleport" result plus everything the loop needs from it
// (residual terms, score sum, sink mass, prescaled copy, FP32 residual).
template <typename T, bool RESID>
__device__ __forceinline__ void epilogue_store(const SweepArgs<T> &a, uint32_t u, T y, T xo,
                                               Acc4 &s) {
    __builtin_nontemporal_store(y, a.xnew + u);
    if (a.xhat_new != nullptr && u < a.sink_begin)
        __builtin_nontemporal_store(y * __builtin_nontemporal_load(a.inv + u), a.xhat_new + u);
    const double dif = (double)y - (double)xo;
    if (RESID) __builtin_nontemporal_store((float)dif, a.r32 + u);
    s.l1 += fabs(dif);
    s.l2 += dif * dif;
    s.sum += (double)y;
    if (u >= a.sink_begin) s.sink += (double)y;
}

template <typename T, bool BASEVEC, bool RESID>
__device__ __forceinline__ void row_epilogue(const SweepArgs<T> &a, uint32_t u, T acc, T damp,
                                             T c, Acc4 &s) {
    T y = damp * acc + c;
    if (BASEVEC) y += a.base_vec[u];
    epilogue_store<T, RESID>(a, u, y, a.xold[u], s);
}

template <typename T> __device__ __forceinline__ T rhs_const(const SweepArgs<T> &a) {
    double c = a.tele;
    if (a.distribute) c += a.damp * a.ctl->sink_in / (double)a.n;
    return (T)c;
}

constexpr int kPer = kCap / kBlock;  // staged in-edges per thread in a stream item

// Streaming accesses (read once per sweep) are non-temporal so that they do not evict the
// cache lines of the gathered vector from L2; gathers use ordinary cached loads.
template <typename T> __device__ __forceinline__ T ld_stream(const T *p) {
    return __builtin_nontemporal_load(p);
}

// XCD (accelerator complex die) the calling wave runs on: hardware register XCC_ID, bits 3:0
__device__ __forceinline__ uint32_t xcc_id() {
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xFu;
}

// Sum of the products of in-edges [e0, e0+cnt) by the whole block (vector rows and split
// chunks).  Four independent gathers per thread per step keep enough loads in flight.
// Gathered value of source v: from the LDS hub table when v is one of the first nh ids.
template <typename T>
__device__ __forceinline__ T gather(const SweepArgs<T> &a, const T *hub, uint32_t nh, int32_t v) {
    T g;
    if ((uint32_t)v < nh) g = hub[v];
    else g = a.xg[v];
    return g;
}

template <typename T, bool FOLDED>
__device__ __forceinline__ T block_row_sum(const SweepArgs<T> &a, uint32_t e0, uint32_t cnt,
                                           T *red, const T *hub, uint32_t nh) {
    T acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
    const uint32_t tid = threadIdx.x;
    const int32_t *__restrict__ srcb = a.sc.src + e0;
    const T *__restrict__ cfb = FOLDED ? a.coef + e0 : nullptr;
    uint32_t i = tid;
    for (; i + 3 * kBlock < cnt; i += 4 * kBlock) {
        const int32_t s0 = ld_stream(srcb + i), s1 = ld_stream(srcb + i + kBlock);
        const int32_t s2 = ld_stream(srcb + i + 2 * kBlock), s3 = ld_stream(srcb + i + 3 * kBlock);
        T c0 = 1, c1 = 1, c2 = 1, c3 = 1;
        if (FOLDED) {
            c0 = ld_stream(cfb + i);
            c1 = ld_stream(cfb + i + kBlock);
            c2 = ld_stream(cfb + i + 2 * kBlock);
            c3 = ld_stream(cfb + i + 3 * kBlock);
        }
        acc0 += gather(a, hub, nh, s0) * c0;
        acc1 += gather(a, hub, nh, s1) * c1;
        acc2 += gather(a, hub, nh, s2) * c2;
        acc3 += gather(a, hub, nh, s3) * c3;
    }
    for (; i < cnt; i += kBlock) {
        T v = gather(a, hub, nh, ld_stream(srcb + i));
        if (FOLDED) v *= ld_stream(cfb + i);
        acc0 += v;
    }
    T acc = (acc0 + acc1) + (acc2 + acc3);
    return block_sum<kBlock>(acc, red);
}

// One stream item: rows [r0, r0+R) of a compressed list with absolute in-edge offsets `ptr`,
// in-edges [e0, e0+cnt) (cnt <= kCap).  DIRECT rows finish with the epilogue (row ids from
// sc.drow); sliced rows store their row sums into `out` (one slice's partial vector).
template <typename T, bool FOLDED, bool BASEVEC, bool RESID, bool DIRECT>
__device__ __forceinline__ void stream_item(const SweepArgs<T> &a, const uint4 item,
                                            const uint32_t *__restrict__ ptr, T *__restrict__ out,
                                            T *s_prod, uint32_t *s_ptr, T *s_sum, const T *hub,
                                            uint32_t nh, T damp, T c, Acc4 &st) {
    const uint32_t tid = threadIdx.x;
    const uint32_t r0 = item.x, R = item.y - item.x, e0 = item.z, cnt = item.w;
    // issue every load that does not depend on another load first
    const uint32_t *__restrict__ pb = ptr + r0;
    uint32_t p0 = 0, p1 = 0, p2 = 0;
    if (tid <= R) p0 = ld_stream(pb + tid) - e0;
    if (tid + kBlock <= R) p1 = ld_stream(pb + tid + kBlock) - e0;
    if (tid == 0 && R == 2 * kBlock) p2 = ld_stream(pb + R) - e0;
    uint32_t u0 = 0, u1 = 0;
    if (DIRECT) {
        if (tid < R) u0 = ld_stream(a.sc.drow + r0 + tid);
        if (tid + kBlock < R) u1 = ld_stream(a.sc.drow + r0 + tid + kBlock);
    }
    int32_t sidx[kPer];
    T cf[kPer];
    {
        // lane-predicated loads share one base address (immediate chunk offsets)
        const int32_t *__restrict__ srcb = a.sc.src + e0;
        const T *__restrict__ cfb = FOLDED ? a.coef + e0 : nullptr;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            sidx[k] = 0;
            cf[k] = 0;
            if (tid + k * kBlock < cnt) {
                sidx[k] = ld_stream(srcb + tid + k * kBlock);
                cf[k] = FOLDED ? ld_stream(cfb + tid + k * kBlock) : (T)1;
            }
        }
    }
    T xo0 = 0, xo1 = 0, bv0 = 0, bv1 = 0;
    if (DIRECT) {
        // folded: xold is also the gathered vector, keep its lines cacheable
        if (tid < R) {
            xo0 = FOLDED ? a.xold[u0] : ld_stream(a.xold + u0);
            if (BASEVEC) bv0 = ld_stream(a.base_vec + u0);
        }
        if (tid + kBlock < R) {
            xo1 = FOLDED ? a.xold[u1] : ld_stream(a.xold + u1);
            if (BASEVEC) bv1 = ld_stream(a.base_vec + u1);
        }
    }
    T v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) v[k] = gather(a, hub, nh, sidx[k]) * cf[k];
    if (tid <= R) s_ptr[tid] = p0;
    if (tid + kBlock <= R) s_ptr[tid + kBlock] = p1;
    if (tid == 0 && R == 2 * kBlock) s_ptr[R] = p2;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t i = tid + k * kBlock;
        if (i < cnt) s_prod[i] = v[k];
    }
    __syncthreads();
    // row sums: G consecutive lanes per row, G ~ average row length (power of two)
    const uint32_t avg = cnt / R;
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
    // one thread per row; the LDS buffers are rewritten only after the next item's first
    // barrier, which every thread reaches after this point
    if (DIRECT) {
        if (tid < R) {
            T y = damp * s_sum[tid] + c;
            if (BASEVEC) y += bv0;
            epilogue_store<T, RESID>(a, u0, y, xo0, st);
        }
        if (tid + kBlock < R) {
            T y = damp * s_sum[tid + kBlock] + c;
            if (BASEVEC) y += bv1;
            epilogue_store<T, RESID>(a, u1, y, xo1, st);
        }
    } else {
        if (tid < R) __builtin_nontemporal_store(s_sum[tid], out + r0 + tid);
        if (tid + kBlock < R) __builtin_nontemporal_store(s_sum[tid + kBlock], out + r0 + tid + kBlock);
    }
}

template <typename T, bool FOLDED, bool BASEVEC, bool RESID>
__global__ __launch_bounds__(kBlock) void sweep_kernel(SweepArgs<T> a) {
    if (a.ctl->done) return;  // converged earlier in this batch: no-op

    __shared__ T s_prod[kCap];
    __shared__ uint32_t s_ptr[kRowsMax + 1];
    __shared__ T s_sum[kRowsMax];
    __shared__ T s_redT[kBlock / 64];
    __shared__ double s_red[kBlock / 64];
    __shared__ uint32_t s_worker;
    __shared__ T s_hub[HubTable<T>::size];

    const T damp = (T)a.damp;
    const T c = rhs_const(a);
    const uint32_t tid = threadIdx.x;
    const uint32_t S = a.sc.n_slices, K = a.sc.workers;
    Acc4 st;

    // claim a worker (slice, rank), trying the slice of this block's XCD first.  Each slice
    // hands out exactly K ranks and the grid has S*K blocks, so every block gets one.
    if (tid == 0) {
        const uint32_t first = xcc_id() % S;
        uint32_t w = 0;
        for (uint32_t k = 0; k < S; ++k) {
            const uint32_t sl = (first + k) % S;
            const uint32_t r = atomicAdd(a.sc.claim + sl, 1u);
            if (r < K) {
                w = sl * K + r;
                break;
            }
        }
        s_worker = w;
    }
    // hub table: the gathered values of the most gathered sources
    const uint32_t nh = min(min((uint32_t)HubTable<T>::size, a.sink_begin), a.n);
    for (uint32_t i = tid; i < nh; i += kBlock) s_hub[i] = a.xg[i];
    __syncthreads();
    const uint32_t w = s_worker;
    const uint32_t sl = w / K, rank = w - sl * K;

    // 1) this slice's items (row sums into the slice's partial vector)
    {
        T *__restrict__ part = a.partial + (size_t)sl * a.sc.n_s;
        const uint32_t *__restrict__ ptr = a.sc.sptr + (size_t)sl * a.sc.n_s;
        const uint32_t ie = a.sc.sitem_off[sl + 1];
        uint32_t it = a.sc.sitem_off[sl] + rank;
        uint4 item = make_uint4(0, 0, 0, 0);
        if (it < ie) item = a.sc.sitems[it];
        while (it < ie) {
            const uint32_t nit = it + K;
            uint4 next = make_uint4(0, 0, 0, 0);
            if (nit < ie) next = a.sc.sitems[nit];
            if (item.y & kSplitFlag) {
                const T p = block_row_sum<T, FOLDED>(a, item.z, item.w, s_redT, s_hub, nh);
                if (tid == 0) a.split_partial[item.y & ~kSplitFlag] = p;
            } else if (item.y - item.x == 1 && item.w > (uint32_t)kCap) {
                const T p = block_row_sum<T, FOLDED>(a, item.z, item.w, s_redT, s_hub, nh);
                if (tid == 0) part[item.x] = p;
            } else {
                stream_item<T, FOLDED, BASEVEC, RESID, false>(a, item, ptr, part, s_prod, s_ptr,
                                                              s_sum, s_hub, nh, damp, c, st);
            }
            item = next;
            it = nit;
        }
    }
    // 2) direct rows, spread over all workers
    {
        const uint32_t nw = S * K;
        uint32_t it = w;
        uint4 item = make_uint4(0, 0, 0, 0);
        if (it < a.sc.n_ditems) item = a.sc.ditems[it];
        while (it < a.sc.n_ditems) {
            const uint32_t nit = it + nw;
            uint4 next = make_uint4(0, 0, 0, 0);
            if (nit < a.sc.n_ditems) next = a.sc.ditems[nit];
            stream_item<T, FOLDED, BASEVEC, RESID, true>(a, item, a.sc.dptr, nullptr, s_prod,
                                                         s_ptr, s_sum, s_hub, nh, damp, c, st);
            item = next;
            it = nit;
        }
    }

    // per-worker partials (l1, l2, sum, sink) in FP64: indexed by the claimed worker, whose
    // rows are fixed, not by the block, whose worker depends on the order of the claims, so
    // that the finalize kernel adds the same values in the same order on every run
    double v;
    v = block_sum<kBlock>(st.l1, s_red);
    if (tid == 0) a.partials[w * 4 + 0] = v;
    v = block_sum<kBlock>(st.l2, s_red);
    if (tid == 0) a.partials[w * 4 + 1] = v;
    v = block_sum<kBlock>(st.sum, s_red);
    if (tid == 0) a.partials[w * 4 + 2] = v;
    v = block_sum<kBlock>(st.sink, s_red);
    if (tid == 0) a.partials[w * 4 + 3] = v;
}

// Sliced rows: add the per-slice row sums in slice order and run the row epilogue.  Rows
// with split segments are left to the finalize kernel.
template <typename T, bool BASEVEC, bool RESID>
__global__ __launch_bounds__(kCombineThreads) void combine_kernel(SweepArgs<T> a,
                                                                  uint32_t block_base) {
    if (a.ctl->done) return;
    __shared__ double s_red[kCombineThreads / 64];
    const T damp = (T)a.damp;
    const T c = rhs_const(a);
    const uint32_t S = a.sc.n_slices, ns = a.sc.n_s;
    Acc4 st;
    for (uint32_t j = blockIdx.x * kCombineThreads + threadIdx.x; j < ns;
         j += gridDim.x * kCombineThreads) {
        if (a.sc.split_flag[j]) continue;
        T acc = 0;
        for (uint32_t s = 0; s < S; ++s) acc += ld_stream(a.partial + (size_t)s * ns + j);
        const uint32_t u = ld_stream(a.sc.srow + j);
        row_epilogue<T, BASEVEC, RESID>(a, u, acc, damp, c, st);
    }
    double v;
    const uint32_t b = block_base + blockIdx.x;
    v = block_sum<kCombineThreads>(st.l1, s_red);
    if (threadIdx.x == 0) a.partials[b * 4 + 0] = v;
    v = block_sum<kCombineThreads>(st.l2, s_red);
    if (threadIdx.x == 0) a.partials[b * 4 + 1] = v;
    v = block_sum<kCombineThreads>(st.sum, s_red);
    if (threadIdx.x == 0) a.partials[b * 4 + 2] = v;
    v = block_sum<kCombineThreads>(st.sink, s_red);
    if (threadIdx.x == 0) a.partials[b * 4 + 3] = v;
}

// One block: finish rows with split segments, add all partials in a fixed order, apply the
// stopping rule, reset the worker claim counters for the next sweep.
template <typename T, bool BASEVEC, bool RESID>
__global__ __launch_bounds__(kFinalizeThreads) void finalize_kernel(SweepArgs<T> a,
                                                                    FinalizeArgs f) {
    PhaseCtl *ctl = a.ctl;
    if (ctl->done) return;
    __shared__ double s_red[kFinalizeThreads / 64];
    const T damp = (T)a.damp;
    const T c = rhs_const(a);
    const uint32_t S = a.sc.n_slices, ns = a.sc.n_s;
    Acc4 st;
    for (uint32_t t = threadIdx.x; t < f.n_split_rows; t += kFinalizeThreads) {
        const uint32_t j = f.split_rows[t];
        T acc = 0;
        for (uint32_t s = 0; s < S; ++s) {
            const uint32_t cnt = f.split_count[t * S + s];
            if (cnt == 0) {
                acc += a.partial[(size_t)s * ns + j];
            } else {
                const uint32_t first = f.split_first[t * S + s];
                for (uint32_t k = 0; k < cnt; ++k) acc += a.split_partial[first + k];
            }
        }
        row_epilogue<T, BASEVEC, RESID>(a, a.sc.srow[j], acc, damp, c, st);
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
        if (a.sc.claim)
            for (int s = 0; s < kSlices; ++s) a.sc.claim[s] = 0u;
    }
}

__global__ void init_ctl_kernel(PhaseCtl *ctl, const double *sink_sum, uint32_t *claim) {
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
        for (int s = 0; s < kSlices; ++s) claim[s] = 0u;
    }
}

__global__ void gather_scaled_kernel(const double *in_z, const int32_t *perm, uint32_t n,
                                     double f, double *out64, float *out32) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = f * in_z[perm[i]];
    out64[i] = v;
    if (out32) out32[i] = (float)v;
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

// x64 = c + e32: the correction of a constant iterate that was never written out
__global__ void set_correction_kernel(double *x64, double c, const float *e32, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x64[i] = c + (double)e32[i];
}

// rowsum[u] = the FP64 sum of row u's in-coefficients, c_vu = w_vu / deg_v (coef) or 1 / deg_v
// (inv[v]): one wave per row, lanes strided over the row, then a fixed shuffle tree.
__global__ void rowsum_kernel(const uint32_t *rowptr, uint32_t n, const int32_t *src,
                              const double *coef, const double *inv, double *rowsum) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t u = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; u < n; u += nw) {
        double sum = 0.0;
        for (uint32_t e = rowptr[u] + lane; e < rowptr[u + 1]; e += 64)
            sum += coef ? coef[e] : inv[src[e]];
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
        if (lane == 0) rowsum[u] = sum;
    }
}

// flag = 1 when some x[u], u < n, differs from x[0]
__global__ void constant_check_kernel(const double *x, uint32_t n, uint32_t *flag) {
    const double x0 = x[0];
    int diff = 0;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x)
        diff |= x[u] != x0;
    if (__syncthreads_or(diff) && threadIdx.x == 0) atomicOr(flag, 1u);
}

// The residual sweep from the constant vector x = c in closed form: y = d c rowsum + rhs
// (+ base), ynew = y, r32 = float(y - c), and the sweep's block partials (l1, l2, sum, sink).
__global__ __launch_bounds__(256) void const_resid_kernel(const double *rowsum, uint32_t n,
                                                          double dc, double rhs,
                                                          const double *base, double c,
                                                          uint32_t sink_begin, double *ynew,
                                                          float *r32, double *partials) {
    __shared__ double s_red[256 / 64];
    double l1 = 0.0, l2 = 0.0, sum = 0.0, sink = 0.0;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x) {
        double y = dc * rowsum[u] + rhs;
        if (base) y += base[u];
        ynew[u] = y;
        const double dif = y - c;
        r32[u] = (float)dif;
        l1 += fabs(dif);
        l2 += dif * dif;
        sum += y;
        if (u >= sink_begin) sink += y;
    }
    l1 = block_sum<256>(l1, s_red);
    l2 = block_sum<256>(l2, s_red);
    sum = block_sum<256>(sum, s_red);
    sink = block_sum<256>(sink, s_red);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 4 + 0] = l1;
        partials[blockIdx.x * 4 + 1] = l2;
        partials[blockIdx.x * 4 + 2] = sum;
        partials[blockIdx.x * 4 + 3] = sink;
    }
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

// slot_u[slot] = compact id of the node in that slot, -1 for slots without one
__global__ void slot_clear_kernel(uint64_t z, int32_t *slot_u) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < z;
         i += (uint64_t)gridDim.x * blockDim.x)
        slot_u[i] = -1;
}
__global__ void slot_set_kernel(const int32_t *perm, uint32_t n, int32_t *slot_u) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
        slot_u[perm[u]] = (int32_t)u;
}

// The scatter of scatter_result_kernel as a gather in slot order: coalesced writes, and the
// reads of y that are not in order (the nodes with out-edges, y's head) stay in the caches.
// Slots without a node keep their value.
__global__ void gather_result_kernel(const double *y, const int32_t *slot_u, uint64_t z,
                                     double factor, double *scores_z) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= z) return;
    const int32_t i = slot_u[j];
    if (i >= 0) scores_z[j] = y[i] / factor;
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
void launch_pair(const SweepArgs<T> &a, const FinalizeArgs &f, int combine_grid, hipStream_t s,
                 hipEvent_t ev0, hipEvent_t ev1) {
    const unsigned grid = a.sc.n_slices * a.sc.workers;
    if (ev0) NKPR_HIP(hipEventRecord(ev0, s));
    hipLaunchKernelGGL((sweep_kernel<T, FOLDED, BASEVEC, RESID>), dim3(grid), dim3(kBlock), 0, s,
                       a);
    NKPR_HIP(hipGetLastError());
    if (a.sc.n_s > 0 && combine_grid > 0) {
        hipLaunchKernelGGL((combine_kernel<T, BASEVEC, RESID>), dim3(combine_grid),
                           dim3(kCombineThreads), 0, s, a, grid);
        NKPR_HIP(hipGetLastError());
    }
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
void launch_sweep(const SweepArgs<T> &a, const FinalizeArgs &f, int kind, bool folded,
                  int combine_grid, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    if (kind == kSweepPlain) {
        if (folded) launch_pair<T, true, false, false>(a, f, combine_grid, s, ev0, ev1);
        else launch_pair<T, false, false, false>(a, f, combine_grid, s, ev0, ev1);
    } else if (kind == kSweepBaseVec) {
        if (folded) launch_pair<T, true, true, false>(a, f, combine_grid, s, ev0, ev1);
        else launch_pair<T, false, true, false>(a, f, combine_grid, s, ev0, ev1);
    } else if (kind == kSweepResid) {
        if (folded) launch_pair<T, true, false, true>(a, f, combine_grid, s, ev0, ev1);
        else launch_pair<T, false, false, true>(a, f, combine_grid, s, ev0, ev1);
    } else {
        if (folded) launch_pair<T, true, true, true>(a, f, combine_grid, s, ev0, ev1);
        else launch_pair<T, false, true, true>(a, f, combine_grid, s, ev0, ev1);
    }
}

template void launch_sweep<float>(const SweepArgs<float> &, const FinalizeArgs &, int, bool,
                                  int, hipStream_t, hipEvent_t, hipEvent_t);
template void launch_sweep<double>(const SweepArgs<double> &, const FinalizeArgs &, int, bool,
                                   int, hipStream_t, hipEvent_t, hipEvent_t);

void launch_finalize64(const SweepArgs<double> &a, const FinalizeArgs &f, int kind, hipStream_t s) {
    // the template flags only matter for split rows, which a binned sweep does not have
    const bool basevec = kind == kSweepBaseVec || kind == kSweepBaseResid;
    const bool resid = kind == kSweepResid || kind == kSweepBaseResid;
    if (basevec && resid)
        hipLaunchKernelGGL((finalize_kernel<double, true, true>), dim3(1), dim3(kFinalizeThreads), 0, s, a, f);
    else if (basevec)
        hipLaunchKernelGGL((finalize_kernel<double, true, false>), dim3(1), dim3(kFinalizeThreads), 0, s, a, f);
    else if (resid)
        hipLaunchKernelGGL((finalize_kernel<double, false, true>), dim3(1), dim3(kFinalizeThreads), 0, s, a, f);
    else
        hipLaunchKernelGGL((finalize_kernel<double, false, false>), dim3(1), dim3(kFinalizeThreads), 0, s, a, f);
    NKPR_HIP(hipGetLastError());
}

int sweep_resident_blocks(int device, bool dbl, bool folded, int kind) {
    const void *fn;
    if (dbl) {
        if (kind == kSweepResid) fn = folded ? sweep_fn<double, true, false, true>() : sweep_fn<double, false, false, true>();
        else if (kind == kSweepBaseResid) fn = folded ? sweep_fn<double, true, true, true>() : sweep_fn<double, false, true, true>();
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
    return per_cu * std::max(cus, 1);
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

void launch_gather_scaled(const double *in_z, const int32_t *perm, uint32_t n, double f,
                          double *out64, float *out32, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(gather_scaled_kernel, dim3(grid1d(n)), dim3(256), 0, s, in_z, perm, n, f,
                       out64, out32);
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

void launch_set_correction(double *x64, double c, const float *e32, uint32_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(set_correction_kernel, dim3(grid1d(n)), dim3(256), 0, s, x64, c, e32, n);
    NKPR_HIP(hipGetLastError());
}

void launch_rowsum64(const uint32_t *rowptr, uint32_t n, const int32_t *src, const double *coef,
                     const double *inv, double *rowsum, hipStream_t s) {
    if (!n) return;
    const unsigned grid = (unsigned)std::min<uint64_t>(((uint64_t)n * 64 + 255) / 256, 1u << 16);
    hipLaunchKernelGGL(rowsum_kernel, dim3(grid), dim3(256), 0, s, rowptr, n, src, coef, inv, rowsum);
    NKPR_HIP(hipGetLastError());
}

void launch_constant_check(const double *x, uint32_t n, uint32_t *flag, hipStream_t s) {
    NKPR_HIP(hipMemsetAsync(flag, 0, 4, s));
    if (!n) return;
    hipLaunchKernelGGL(constant_check_kernel, dim3(std::min<unsigned>(grid1d(n), 2048)), dim3(256), 0, s,
                       x, n, flag);
    NKPR_HIP(hipGetLastError());
}

unsigned launch_const_resid(const double *rowsum, uint32_t n, double dc, double rhs,
                            const double *base, double c, uint32_t sink_begin, double *ynew,
                            float *r32, double *partials, hipStream_t s) {
    const unsigned grid = std::max(1u, std::min<unsigned>(grid1d(n), 2048));
    hipLaunchKernelGGL(const_resid_kernel, dim3(grid), dim3(256), 0, s, rowsum, n, dc, rhs, base, c,
                       sink_begin, ynew, r32, partials);
    NKPR_HIP(hipGetLastError());
    return grid;
}

void launch_copy_f32(const float *in, float *out, uint32_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(copy_f32_kernel, dim3(grid1d(n)), dim3(256), 0, s, in, out, n);
    NKPR_HIP(hipGetLastError());
}

void launch_slot_index(const int32_t *perm, uint32_t n, uint64_t z, int32_t *slot_u, hipStream_t s) {
    if (!z) return;
    hipLaunchKernelGGL(slot_clear_kernel, dim3((unsigned)std::min<uint64_t>((z + 255) / 256, 1u << 16)),
                       dim3(256), 0, s, z, slot_u);
    if (n)
        hipLaunchKernelGGL(slot_set_kernel, dim3((unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16)),
                           dim3(256), 0, s, perm, n, slot_u);
    NKPR_HIP(hipGetLastError());
}

void launch_gather_result(const double *y, const int32_t *slot_u, uint64_t z, double factor,
                          double *scores_z, hipStream_t s) {
    if (!z) return;
    hipLaunchKernelGGL(gather_result_kernel, dim3((unsigned)((z + 255) / 256)), dim3(256), 0, s, y,
                       slot_u, z, factor, scores_z);
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

void launch_init_ctl(PhaseCtl *ctl, const double *sink_sum, uint32_t *claim, hipStream_t s) {
    hipLaunchKernelGGL(init_ctl_kernel, dim3(1), dim3(64), 0, s, ctl, sink_sum, claim);
    NKPR_HIP(hipGetLastError());
}

}  // namespace nkpr
