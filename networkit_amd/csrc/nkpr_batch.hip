// Batched personalised PageRank: k teleport distributions on one graph in one SpMM sweep.
//
// Column j solves x_j = d P x_j + (1-d) t_j (+ d S_j / n with distributed sinks), the loop of
// networkit/cpp/centrality/PageRank.cpp:53-71 with the uniform teleport (1-d)/n replaced by a
// per-column distribution.  All columns advance together: one sweep reads every in-edge once
// and gathers, per edge, a 16-byte slice of the source's row of the iterate matrix.
//
// Layout (per graph, built on first use): a plain CSC in relabelled ids (uint32 row offsets,
// int32 sources, folded FP32/FP64 coefficients for weighted graphs) copied out of the sliced
// sweep layout.  Iterates are node-major matrices X[n][ld] (ld = k rounded up to 16 bytes),
// so the k values of one source are contiguous: a gather is a run of full cache lines.
//
// Work split.  A "row group" of LPR lanes handles one row; each lane owns 16 bytes of columns
// (4 FP32 or 2 FP64), so a group covers a column tile of LPR * 16 bytes and gridDim.y walks
// the tiles.  Rows with more than `heavy` in-edges are cut into pieces of at most `heavy`
// edges; piece sums go to a buffer that a combine kernel adds in piece order.  Per-column
// reductions (L1 / L2 step, sum, sink mass) are FP64 partials per block, reduced in a fixed
// order by the finalize kernel: results do not depend on scheduling.  The column tile width
// and `heavy` are picked by timing candidate configurations on the first run (autotuner).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "nkpr_internal.hpp"

namespace nkpr {

void ensure_buf(DevBuf &b, size_t bytes) { b.alloc(bytes); }

namespace {

uint32_t env_u32_batch(const char *name, uint32_t dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    return (uint32_t)std::strtoul(v, nullptr, 10);
}

constexpr int kBThreads = 512;
constexpr int kUnroll = 8;
constexpr int kFinThreads = 256;

enum BatchKind : int { kBPlain = 0, kBBase = 1, kBResid = 2 };

template <typename T> struct V16;
template <> struct V16<float> {
    using type = float4;
    static constexpr int n = 4;
};
template <> struct V16<double> {
    using type = double2;
    static constexpr int n = 2;
};

template <typename T>
__device__ __forceinline__ void load16(const T *p, T (&v)[V16<T>::n]) {
    const typename V16<T>::type q = *reinterpret_cast<const typename V16<T>::type *>(p);
    if constexpr (V16<T>::n == 4) {
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
        v[0] = q.x; v[1] = q.y;
    }
}

template <typename T>
__device__ __forceinline__ void store16(T *p, const T (&v)[V16<T>::n]) {
    typename V16<T>::type q;
    if constexpr (V16<T>::n == 4) {
        q.x = v[0]; q.y = v[1]; q.z = v[2]; q.w = v[3];
    } else {
        q.x = v[0]; q.y = v[1];
    }
    *reinterpret_cast<typename V16<T>::type *>(p) = q;
}

template <typename T> struct BatchArgs {
    const uint32_t *rowptr;   // [n + 1]
    const int32_t *src;       // [m]
    const T *coef;            // [m] folded w / deg(src) (weighted) or nullptr
    const T *inv;             // [n] 1 / deg (unweighted)
    const T *xold;            // [n][ld]
    T *xnew;                  // [n][ld]
    const float *base;        // [n][ld] right-hand side of a correction solve (kBBase)
    float *r32;               // [n][ld] float(x' - x) (kBResid)
    const uint32_t *tptr;     // [n + 1] row-sparse teleport: entries of row u
    const uint32_t *tcol;     // column of each entry
    const T *tval;            // (1 - d) t_c[u]
    const uint32_t *light;    // [n_light] rows with in-degree <= heavy
    const uint2 *pieces;      // [n_pieces] {first edge, count}
    const uint32_t *hrow;     // [n_heavy] rows cut into pieces
    const uint32_t *hpiece;   // [n_heavy + 1] piece range of each heavy row
    T *piece_part;            // [n_pieces][ld]
    double *partials;         // [(grid + combine grid)][ld][4]
    const BatchCol *col;      // [ld]
    const BatchCtl *ctl;
    uint32_t n, ld, n_light, n_pieces, n_heavy, sink_begin;
    double damp;
    int32_t distribute;
};

// Sum over the row's edges [e0, e1) of coef * X[src][c0 .. c0 + V) for one row group.
template <typename T, int LPR, bool FOLDED>
__device__ __forceinline__ void row_dot(const BatchArgs<T> &a, uint32_t e0, uint32_t e1,
                                        uint32_t c0, bool on, int l, T (&acc)[V16<T>::n]) {
    constexpr int V = V16<T>::n;
    for (uint32_t b = e0; b < e1; b += LPR) {
        const uint32_t cnt = min((uint32_t)LPR, e1 - b);
        int32_t s = 0;
        T cf = T(0);
        if ((uint32_t)l < cnt) {
            s = a.src[b + l];
            cf = FOLDED ? a.coef[b + l] : a.inv[s];
        }
        uint32_t q = 0;
        for (; q + kUnroll <= cnt; q += kUnroll) {
            int32_t sq[kUnroll];
            T cq[kUnroll];
#pragma unroll
            for (int i = 0; i < kUnroll; ++i) {
                sq[i] = __shfl(s, (int)q + i, LPR);
                cq[i] = __shfl(cf, (int)q + i, LPR);
            }
            if (on) {
                T v[kUnroll][V];
#pragma unroll
                for (int i = 0; i < kUnroll; ++i) load16(a.xold + (size_t)sq[i] * a.ld + c0, v[i]);
#pragma unroll
                for (int i = 0; i < kUnroll; ++i)
#pragma unroll
                    for (int j = 0; j < V; ++j) acc[j] += cq[i] * v[i][j];
            }
        }
        for (; q < cnt; ++q) {
            const int32_t sq = __shfl(s, (int)q, LPR);
            const T cq = __shfl(cf, (int)q, LPR);
            if (on) {
                T v[V];
                load16(a.xold + (size_t)sq * a.ld + c0, v);
#pragma unroll
                for (int j = 0; j < V; ++j) acc[j] += cq * v[j];
            }
        }
    }
}

// x'[u][c] for the lane's columns, stores, and the per-column reductions.
template <typename T, int KIND>
__device__ __forceinline__ void row_epilogue(const BatchArgs<T> &a, uint32_t u, uint32_t c0,
                                             const T (&acc)[V16<T>::n],
                                             const bool (&frozen)[V16<T>::n],
                                             const T (&sinkterm)[V16<T>::n],
                                             double (&red)[4][V16<T>::n]) {
    constexpr int V = V16<T>::n;
    const size_t off = (size_t)u * a.ld + c0;
    T old[V], rhs[V], nv[V];
    load16(a.xold + off, old);
    if constexpr (KIND == kBBase) {
        float b[V];
        if constexpr (V == 4) {
            const float4 q = *reinterpret_cast<const float4 *>(a.base + off);
            b[0] = q.x; b[1] = q.y; b[2] = q.z; b[3] = q.w;
        } else {
            const float2 q = *reinterpret_cast<const float2 *>(a.base + off);
            b[0] = q.x; b[1] = q.y;
        }
#pragma unroll
        for (int j = 0; j < V; ++j) rhs[j] = (T)b[j];
    } else {
#pragma unroll
        for (int j = 0; j < V; ++j) rhs[j] = T(0);
        const uint32_t t0 = a.tptr[u], t1 = a.tptr[u + 1];
        for (uint32_t t = t0; t < t1; ++t) {
            const uint32_t c = a.tcol[t] - c0;
            const T tv = a.tval[t];
#pragma unroll
            for (int j = 0; j < V; ++j)
                if (c == (uint32_t)j) rhs[j] += tv;
        }
    }
    const T d = (T)a.damp;
    const bool is_sink = u >= a.sink_begin;
#pragma unroll
    for (int j = 0; j < V; ++j) {
        nv[j] = frozen[j] ? old[j] : d * acc[j] + rhs[j] + sinkterm[j];
        const double diff = (double)nv[j] - (double)old[j];
        red[0][j] += fabs(diff);
        red[1][j] += diff * diff;
        red[2][j] += (double)nv[j];
        if (is_sink) red[3][j] += (double)nv[j];
    }
    store16(a.xnew + off, nv);
    if constexpr (KIND == kBResid) {
        if constexpr (V == 2) {
            *reinterpret_cast<float2 *>(a.r32 + off) =
                make_float2((float)((double)nv[0] - (double)old[0]),
                            (float)((double)nv[1] - (double)old[1]));
        } else {
            *reinterpret_cast<float4 *>(a.r32 + off) = make_float4(
                (float)(nv[0] - old[0]), (float)(nv[1] - old[1]), (float)(nv[2] - old[2]),
                (float)(nv[3] - old[3]));
        }
    }
}

// Column state of the lane: frozen flags and the distributed sink term d S_c / n.
template <typename T>
__device__ __forceinline__ void lane_columns(const BatchArgs<T> &a, uint32_t c0, bool on,
                                             bool (&frozen)[V16<T>::n], T (&sinkterm)[V16<T>::n]) {
    constexpr int V = V16<T>::n;
#pragma unroll
    for (int j = 0; j < V; ++j) {
        frozen[j] = true;
        sinkterm[j] = T(0);
        if (on) {
            const BatchCol &cc = a.col[c0 + j];
            frozen[j] = cc.done != 0u;
            if (a.distribute) sinkterm[j] = (T)(a.damp * cc.sink_in / (double)a.n);
        }
    }
}

// Adds red over the row groups of the block and writes partials[(blk * ld + col) * 4 + q].
template <typename T, int LPR>
__device__ __forceinline__ void block_partials(const BatchArgs<T> &a, uint32_t blk,
                                               uint32_t tile, double (&red)[4][V16<T>::n]) {
    constexpr int V = V16<T>::n;
    constexpr int KT = LPR * V;
    __shared__ double s_red[kBThreads / 64][LPR * V];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l = lane % LPR;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int j = 0; j < V; ++j) {
            double v = red[q][j];
#pragma unroll
            for (int off = LPR; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
            red[q][j] = v;
        }
        if (lane < LPR) {
#pragma unroll
            for (int j = 0; j < V; ++j) s_red[wave][l * V + j] = red[q][j];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < KT; t += kBThreads) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kBThreads / 64; ++w) s += s_red[w][t];
            const uint32_t c = tile * KT + (uint32_t)t;
            if (c < a.ld) a.partials[((size_t)blk * a.ld + c) * 4 + q] = s;
        }
        __syncthreads();
    }
}

template <typename T, int LPR, bool FOLDED, int KIND>
__global__ __launch_bounds__(kBThreads) void spmm_kernel(BatchArgs<T> a) {
    if (a.ctl->done) return;
    constexpr int V = V16<T>::n;
    constexpr int KT = LPR * V;
    constexpr int GROUPS = kBThreads / LPR;
    const int l = threadIdx.x % LPR;
    const uint32_t gidx = threadIdx.x / LPR;
    const uint32_t tile = blockIdx.y;
    const uint32_t c0 = tile * KT + (uint32_t)l * V;
    const bool on = c0 < a.ld;
    bool frozen[V];
    T sinkterm[V];
    lane_columns(a, c0, on, frozen, sinkterm);
    double red[4][V];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < V; ++j) red[q][j] = 0.0;

    const uint32_t stride = gridDim.x * GROUPS;
    for (uint32_t p = blockIdx.x * GROUPS + gidx; p < a.n_pieces; p += stride) {
        const uint2 pc = a.pieces[p];
        T acc[V];
#pragma unroll
        for (int j = 0; j < V; ++j) acc[j] = T(0);
        row_dot<T, LPR, FOLDED>(a, pc.x, pc.x + pc.y, c0, on, l, acc);
        if (on) store16(a.piece_part + (size_t)p * a.ld + c0, acc);
    }
    for (uint32_t i = blockIdx.x * GROUPS + gidx; i < a.n_light; i += stride) {
        const uint32_t u = a.light[i];
        T acc[V];
#pragma unroll
        for (int j = 0; j < V; ++j) acc[j] = T(0);
        row_dot<T, LPR, FOLDED>(a, a.rowptr[u], a.rowptr[u + 1], c0, on, l, acc);
        if (on) row_epilogue<T, KIND>(a, u, c0, acc, frozen, sinkterm, red);
    }
    block_partials<T, LPR>(a, blockIdx.x, tile, red);
}

// Heavy rows: add the piece sums in piece order, then the epilogue.
template <typename T, int LPR, int KIND>
__global__ __launch_bounds__(kBThreads) void heavy_combine_kernel(BatchArgs<T> a,
                                                                  uint32_t blk_base) {
    if (a.ctl->done) return;
    constexpr int V = V16<T>::n;
    constexpr int KT = LPR * V;
    constexpr int GROUPS = kBThreads / LPR;
    const int l = threadIdx.x % LPR;
    const uint32_t gidx = threadIdx.x / LPR;
    const uint32_t tile = blockIdx.y;
    const uint32_t c0 = tile * KT + (uint32_t)l * V;
    const bool on = c0 < a.ld;
    bool frozen[V];
    T sinkterm[V];
    lane_columns(a, c0, on, frozen, sinkterm);
    double red[4][V];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < V; ++j) red[q][j] = 0.0;
    for (uint32_t h = blockIdx.x * GROUPS + gidx; h < a.n_heavy; h += gridDim.x * GROUPS) {
        if (!on) continue;
        T acc[V];
#pragma unroll
        for (int j = 0; j < V; ++j) acc[j] = T(0);
        for (uint32_t p = a.hpiece[h]; p < a.hpiece[h + 1]; ++p) {
            T v[V];
            load16(a.piece_part + (size_t)p * a.ld + c0, v);
#pragma unroll
            for (int j = 0; j < V; ++j) acc[j] += v[j];
        }
        row_epilogue<T, KIND>(a, a.hrow[h], c0, acc, frozen, sinkterm, red);
    }
    block_partials<T, LPR>(a, blk_base + blockIdx.x, tile, red);
}

// One block per real column: reduce the block partials in a fixed order and apply the rule.
__global__ __launch_bounds__(kFinThreads) void batch_finalize_kernel(const double *partials,
                                                                     uint32_t nblocks,
                                                                     uint32_t ld, BatchCol *col,
                                                                     const BatchCtl *ctl,
                                                                     PhaseRule r) {
    if (ctl->done) return;
    const uint32_t c = blockIdx.x;
    if (col[c].done) return;
    __shared__ double s_red[4][kFinThreads / 64];
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < nblocks; b += kFinThreads) {
        const double *p = partials + ((size_t)b * ld + c) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) t[q] += p[q];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = t[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) s_red[q][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot[4];
        for (int q = 0; q < 4; ++q) {
            double s = 0.0;
            for (int w = 0; w < kFinThreads / 64; ++w) s += s_red[q][w];
            tot[q] = s;
        }
        BatchCol &cc = col[c];
        const uint64_t it = cc.phase_iter + 1;
        cc.phase_iter = it;
        cc.iterations += 1;
        const double prev = cc.l1;
        cc.prev_l1 = prev;
        cc.l1 = tot[0];
        cc.l2 = tot[1];
        cc.sum = tot[2];
        cc.sink_out = tot[3];
        cc.sink_in = tot[3];
        const double bound =
            (tot[2] > 0.0) ? 2.0 * r.damp / (1.0 - r.damp) * tot[0] / tot[2] : INFINITY;
        cc.bound = bound;
        // tolerance of a correction solve that would let the next residual certify
        cc.target = 0.5 * r.tol * tot[2] * (1.0 - r.damp) / (2.0 * r.damp);
        bool stop = it >= r.max_iterations;
        if (r.kind == kRuleStep) {
            const double step = (r.norm == NKPR_L2) ? sqrt(tot[1]) : tot[0];
            stop = stop || step <= r.tol;
        } else if (r.kind == kRuleErr) {
            stop = stop || bound <= r.tol;
        } else {
            // correction solve (tol_abs < 0: per-column target) or FP32 phase
            const double ta = r.tol_abs < 0.0 ? cc.solve_tol : r.tol_abs;
            stop = stop || tot[0] <= ta || bound <= r.tol;
            if (it >= 4 && tot[0] > r.stall_ratio * prev) stop = true;
        }
        cc.done = stop ? 1u : 0u;
        if (stop && r.kind == kRuleErr && bound <= r.tol) cc.final_ = 1u;
    }
}

// Counts the columns still running; sets ctl->done when none is (or the sweep cap is hit).
__global__ void batch_count_kernel(const BatchCol *col, uint32_t k, BatchCtl *ctl,
                                   uint64_t max_sweeps) {
    if (ctl->done) return;
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    uint32_t cnt = 0;
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) cnt += col[c].done ? 0u : 1u;
    atomicAdd(&s_cnt, cnt);
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl->sweeps += 1;
        ctl->active = s_cnt;
        ctl->done = (s_cnt == 0 || ctl->sweeps >= max_sweeps) ? 1u : 0u;
    }
}

// Phase start: done = final (certified columns stay frozen), counters reset, sink mass of the
// start vector from per-block column sums (sinks are rows [sink_begin, n)).
__global__ void batch_init_kernel(BatchCol *col, uint32_t k, uint32_t ld, BatchCtl *ctl,
                                  const double *sink_part, uint32_t sink_blocks) {
    for (uint32_t c = threadIdx.x + blockIdx.x * blockDim.x; c < ld; c += blockDim.x * gridDim.x) {
        BatchCol &cc = col[c];
        if (c >= k) cc.final_ = 1u;
        cc.done = cc.final_;
        cc.phase_iter = 0;
        cc.l1 = INFINITY;
        cc.prev_l1 = INFINITY;
        double s = 0.0;
        for (uint32_t b = 0; b < sink_blocks; ++b) s += sink_part[(size_t)b * ld + c];
        cc.sink_in = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->done = 0;
        ctl->active = k;
        ctl->sweeps = 0;
    }
}

// Per-block column sums of X[begin, end) (FP64): part[b * ld + c].
template <typename T>
__global__ void col_range_sum_kernel(const T *x, uint32_t ld, uint32_t begin, uint32_t end,
                                     uint32_t rows_per_block, double *part) {
    const uint32_t r0 = begin + blockIdx.x * rows_per_block;
    const uint32_t r1 = min(end, r0 + rows_per_block);
    for (uint32_t c = threadIdx.x; c < ld; c += blockDim.x) {
        double s = 0.0;
        for (uint32_t r = r0; r < r1; ++r) s += (double)x[(size_t)r * ld + c];
        part[(size_t)blockIdx.x * ld + c] = s;
    }
}

// X[u][j] = scores[j * z + perm[u]] (j < k), 0 in the padding columns.
// The caller's scores are column-major in slot order (scores[j * z + slot]); the iterates are
// node-major in compact order (x[u * ld + j]).  Both conversions go through 64 x 64 tiles in
// LDS, indexed by slot: every global access is a run of 64 consecutive doubles on either side
// (slot_u: compact id of a slot, or -1 for a slot without a node).
constexpr uint32_t kTile = 64;

// x64[u * ld + j] = scores[j * z + slot(u)] (0 for the padding columns j >= k)
__global__ __launch_bounds__(256) void batch_gather_kernel(const double *scores, uint64_t z,
                                                           const int32_t *slot_u, uint32_t k,
                                                           uint32_t ld, double *x64) {
    __shared__ double tile[kTile][kTile + 1];
    const uint32_t lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const uint64_t st = (z + kTile - 1) / kTile, ct = (ld + kTile - 1) / kTile;
    for (uint64_t t = blockIdx.x; t < st * ct; t += gridDim.x) {
        const uint64_t s0 = (t / ct) * kTile;
        const uint32_t j0 = (uint32_t)(t % ct) * kTile;
        const uint64_t slot = s0 + lane;
        for (uint32_t jj = row; jj < kTile; jj += 4) {
            const uint32_t j = j0 + jj;
            tile[jj][lane] = (j < k && slot < z) ? scores[(uint64_t)j * z + slot] : 0.0;
        }
        __syncthreads();
        for (uint32_t ss = row; ss < kTile; ss += 4) {
            if (s0 + ss >= z) break;
            const int32_t u = slot_u[s0 + ss];
            const uint32_t j = j0 + lane;
            if (u >= 0 && j < ld) x64[(uint64_t)u * ld + j] = tile[lane][ss];
        }
        __syncthreads();
    }
}

// scores[j * z + slot(u)] = x64[u * ld + j] / factor[j] for j < k; slots without a node keep
// their values
__global__ __launch_bounds__(256) void batch_scatter_kernel(const double *x64, const int32_t *slot_u,
                                                            uint32_t k, uint32_t ld, uint64_t z,
                                                            const double *factor, double *scores) {
    __shared__ double tile[kTile][kTile + 1];
    __shared__ int32_t su[kTile];
    const uint32_t lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const uint64_t st = (z + kTile - 1) / kTile, ct = (k + kTile - 1) / kTile;
    for (uint64_t t = blockIdx.x; t < st * ct; t += gridDim.x) {
        const uint64_t s0 = (t / ct) * kTile;
        const uint32_t j0 = (uint32_t)(t % ct) * kTile;
        if (threadIdx.x < kTile) su[threadIdx.x] = (s0 + threadIdx.x < z) ? slot_u[s0 + threadIdx.x] : -1;
        __syncthreads();
        for (uint32_t ss = row; ss < kTile; ss += 4) {
            const int32_t u = su[ss];
            const uint32_t j = j0 + lane;
            tile[lane][ss] = (u >= 0 && j < k) ? x64[(uint64_t)u * ld + j] / factor[j] : 0.0;
        }
        __syncthreads();
        for (uint32_t jj = row; jj < kTile; jj += 4) {
            const uint32_t j = j0 + jj;
            if (j < k && su[lane] >= 0) scores[(uint64_t)j * z + s0 + lane] = tile[jj][lane];
        }
        __syncthreads();
    }
}

// tptr[r] = first teleport entry of row r, r <= n: the offset of the first listed row >= r
// (urow ascending, nu rows; uoff[nu] = the entry count)
__global__ void tele_rowptr_kernel(const uint32_t *urow, const uint32_t *uoff, uint32_t nu,
                                   uint32_t n, uint32_t *tptr) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= n;
         r += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nu;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (urow[mid] < r) lo = mid + 1;
            else hi = mid;
        }
        tptr[r] = uoff[lo];
    }
}

__global__ void f64_to_f32_kernel(const double *in, float *out, uint64_t cnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) out[i] = (float)in[i];
}

__global__ void f32_to_f64_kernel(const float *in, double *out, uint64_t cnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) out[i] = (double)in[i];
}

// After a correction solve: x = y for certified columns, x += e for the others.
__global__ void batch_update_kernel(double *x, const double *y, const float *e, uint32_t ld,
                                    uint64_t cnt, const BatchCol *col) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = (uint32_t)(i % ld);
        x[i] = col[c].final_ ? y[i] : x[i] + (double)e[i];
    }
}

// A correction solve stops each column at the target its last residual sweep set.
__global__ void batch_set_solve_tol_kernel(BatchCol *col, uint32_t ld) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ld) col[c].solve_tol = col[c].target;
}

// Column totals of per-block column sums: out[c] = sum_b part[b * ld + c].
__global__ void col_total_kernel(const double *part, uint32_t nblocks, uint32_t ld, double *out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ld) return;
    double t = 0.0;
    for (uint32_t b = 0; b < nblocks; ++b) t += part[(size_t)b * ld + c];
    out[c] = t;
}

// e0[u][c] += shift * total[c]: with a column-stochastic operator (no sinks, or distributed
// sinks) 1^T (I - dP) = (1-d) 1^T, so the solution of e = dPe + r sums to 1^T r / (1-d).
// Starting there leaves no error along the eigenvalue-d mode, which would otherwise decay only
// as d^k.
__global__ void col_shift_kernel(float *e, uint32_t ld, uint64_t cnt, const double *total,
                                 double shift) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
         i += (uint64_t)gridDim.x * blockDim.x)
        e[i] += (float)(shift * total[i % ld]);
}

// Correction solves start from e = r.
__global__ void copy_f32_kernel(const float *in, float *out, uint64_t cnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) out[i] = in[i];
}

// Plain CSC out of the sliced sweep layout: row r's in-edges in slice order.
__global__ void plain_layout_kernel(const uint64_t *colptr, const uint32_t *drow,
                                    const uint32_t *dptr, uint32_t n_d, const uint32_t *srow,
                                    const uint32_t *sptr, uint32_t n_s, uint32_t S,
                                    const int32_t *src, const float *c32, const double *c64,
                                    int32_t *osrc, float *o32, double *o64) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
         w < (uint64_t)n_d + n_s; w += nw) {
    uint32_t row, nseg;
    if (w < n_d) {
        row = drow[w];
        nseg = 1;
    } else {
        row = srow[w - n_d];
        nseg = S;
    }
    uint64_t out = colptr[row];
    for (uint32_t q = 0; q < nseg; ++q) {
        uint32_t b, e;
        if (w < n_d) {
            b = dptr[w];
            e = dptr[w + 1];
        } else {
            const uint64_t j = w - n_d;
            b = sptr[(uint64_t)q * n_s + j];
            e = sptr[(uint64_t)q * n_s + j + 1];
        }
        for (uint32_t i = b + lane; i < e; i += 64) {
            const uint64_t o = out + (i - b);
            osrc[o] = src[i];
            if (o32) o32[o] = c32[i];
            if (o64) o64[o] = c64[i];
        }
        out += e - b;
    }
    }
}

__global__ void rowptr32_kernel(const uint64_t *colptr, uint32_t n, uint32_t *rowptr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) rowptr[i] = (uint32_t)colptr[i];
}

// capped grid of a grid-stride kernel
unsigned grid_for(uint64_t cnt, int threads = 256) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((cnt + threads - 1) / threads, 1u << 18));
}

unsigned blocks_for(uint64_t cnt, int threads = 256) {
    return (unsigned)((cnt + threads - 1) / threads);
}

// ---- launch dispatch over (T, LPR, FOLDED, KIND) -----------------------------------------
template <typename T, int LPR, int KIND>
void launch_lpr(const BatchArgs<T> &a, bool folded, unsigned grid, unsigned cgrid,
                unsigned tiles, hipStream_t s) {
    if (folded)
        hipLaunchKernelGGL((spmm_kernel<T, LPR, true, KIND>), dim3(grid, tiles), dim3(kBThreads), 0, s, a);
    else
        hipLaunchKernelGGL((spmm_kernel<T, LPR, false, KIND>), dim3(grid, tiles), dim3(kBThreads), 0, s, a);
    NKPR_HIP(hipGetLastError());
    if (a.n_heavy > 0) {
        hipLaunchKernelGGL((heavy_combine_kernel<T, LPR, KIND>), dim3(cgrid, tiles), dim3(kBThreads), 0, s, a, grid);
        NKPR_HIP(hipGetLastError());
    }
}

template <typename T, int KIND>
void launch_kind(const BatchArgs<T> &a, int lpr, bool folded, unsigned grid, unsigned cgrid,
                 hipStream_t s) {
    constexpr int V = V16<T>::n;
    const unsigned tiles = (a.ld + (unsigned)(lpr * V) - 1) / (unsigned)(lpr * V);
    switch (lpr) {
        case 4: launch_lpr<T, 4, KIND>(a, folded, grid, cgrid, tiles, s); break;
        case 16: launch_lpr<T, 16, KIND>(a, folded, grid, cgrid, tiles, s); break;
        default: launch_lpr<T, 64, KIND>(a, folded, grid, cgrid, tiles, s); break;
    }
}

template <typename T>
void launch_spmm(const BatchArgs<T> &a, int kind, int lpr, bool folded, unsigned grid,
                 unsigned cgrid, hipStream_t s) {
    if constexpr (sizeof(T) == 4) {
        if (kind == kBBase) launch_kind<T, kBBase>(a, lpr, folded, grid, cgrid, s);
        else launch_kind<T, kBPlain>(a, lpr, folded, grid, cgrid, s);
    } else {
        if (kind == kBResid) launch_kind<T, kBResid>(a, lpr, folded, grid, cgrid, s);
        else launch_kind<T, kBPlain>(a, lpr, folded, grid, cgrid, s);
    }
}

template <typename T> int occupancy_blocks(int device, int lpr, bool folded) {
    const void *fn;
    if (lpr == 4) fn = folded ? (const void *)&spmm_kernel<T, 4, true, kBPlain> : (const void *)&spmm_kernel<T, 4, false, kBPlain>;
    else if (lpr == 16) fn = folded ? (const void *)&spmm_kernel<T, 16, true, kBPlain> : (const void *)&spmm_kernel<T, 16, false, kBPlain>;
    else fn = folded ? (const void *)&spmm_kernel<T, 64, true, kBPlain> : (const void *)&spmm_kernel<T, 64, false, kBPlain>;
    int per_cu = 0, cus = 0;
    NKPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBThreads, 0));
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    return std::max(1, per_cu) * std::max(1, cus);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------

// Builds the plain CSC copy used by the batched sweeps (once per graph; after new weights
// only the coefficients are copied again).
void batch_layout(nkpr_graph *g) {
    if (g->b_ready && g->b_coef_ready) return;
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    if (g->b_ready) {
        const uint64_t rows = (uint64_t)g->n_d + g->n_s;
        if (rows > 0 && g->weighted) {
            hipLaunchKernelGGL(plain_layout_kernel, dim3((unsigned)std::min<uint64_t>((rows * 64 + 255) / 256, 1u << 18)), dim3(256), 0, s,
                               g->colptr.as<uint64_t>(), g->drow.as<uint32_t>(),
                               g->dptr.as<uint32_t>(), g->n_d, g->srow.as<uint32_t>(),
                               g->sptr.as<uint32_t>(), g->n_s, g->n_slices, g->src.as<int32_t>(),
                               g->coef32.as<float>(), g->coef64.as<double>(), g->b_src.as<int32_t>(),
                               g->b_coef32.as<float>(), g->b_coef64.as<double>());
            NKPR_HIP(hipGetLastError());
        }
        g->b_coef_ready = true;
        return;
    }
    g->b_rowptr.alloc(4 * ((size_t)n + 1));
    hipLaunchKernelGGL(rowptr32_kernel, dim3(blocks_for((uint64_t)n + 1)), dim3(256), 0, s,
                       g->colptr.as<uint64_t>(), n, g->b_rowptr.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    g->b_src.alloc(4 * std::max<uint64_t>(m, 1));
    if (g->weighted) {
        g->b_coef32.alloc(4 * std::max<uint64_t>(m, 1));
        g->b_coef64.alloc(8 * std::max<uint64_t>(m, 1));
    }
    const uint64_t rows = (uint64_t)g->n_d + g->n_s;
    if (rows > 0) {
        const unsigned blocks = (unsigned)std::min<uint64_t>((rows * 64 + 255) / 256, 1u << 18);
        hipLaunchKernelGGL(plain_layout_kernel, dim3(blocks), dim3(256), 0, s,
                           g->colptr.as<uint64_t>(), g->drow.as<uint32_t>(),
                           g->dptr.as<uint32_t>(), g->n_d, g->srow.as<uint32_t>(),
                           g->sptr.as<uint32_t>(), g->n_s, g->n_slices, g->src.as<int32_t>(),
                           g->weighted ? g->coef32.as<float>() : nullptr,
                           g->weighted ? g->coef64.as<double>() : nullptr,
                           g->b_src.as<int32_t>(),
                           g->weighted ? g->b_coef32.as<float>() : nullptr,
                           g->weighted ? g->b_coef64.as<double>() : nullptr);
        NKPR_HIP(hipGetLastError());
    }
    g->b_rowptr_host.resize((size_t)n + 1);
    g->b_perm_host.resize(n);
    read_back(s, {{g->b_rowptr_host.data(), g->b_rowptr.p, 4 * ((size_t)n + 1)},
                  {g->b_perm_host.data(), g->perm.p, 4 * (size_t)n}});
    g->b_iperm_host.assign(g->z, -1);
    for (uint32_t u = 0; u < n; ++u) g->b_iperm_host[(uint32_t)g->b_perm_host[u]] = (int32_t)u;
    g->b_ready = true;
    g->b_coef_ready = true;
}

// Light-row list and heavy-row pieces for one threshold (cached per threshold).
void batch_split(nkpr_graph *g, uint32_t heavy) {
    if (g->b_heavy == heavy && g->b_split_ready) return;
    const uint32_t n = g->n;
    const std::vector<uint32_t> &rp = g->b_rowptr_host;
    std::vector<uint32_t> light, hrow, hpiece(1, 0);
    std::vector<uint2> pieces;
    light.reserve(n);
    for (uint32_t u = 0; u < n; ++u) {
        const uint32_t b = rp[u], e = rp[u + 1];
        if (e - b <= heavy) {
            light.push_back(u);
            continue;
        }
        hrow.push_back(u);
        for (uint32_t q = b; q < e; q += heavy) pieces.push_back(make_uint2(q, std::min(heavy, e - q)));
        hpiece.push_back((uint32_t)pieces.size());
    }
    auto up = [&](DevBuf &d, const void *h, size_t bytes) {
        d.alloc(std::max<size_t>(bytes, 16));
        if (bytes) NKPR_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, g->stream));
    };
    up(g->b_light, light.data(), 4 * light.size());
    up(g->b_hrow, hrow.data(), 4 * hrow.size());
    up(g->b_hpiece, hpiece.data(), 4 * hpiece.size());
    up(g->b_pieces, pieces.data(), sizeof(uint2) * pieces.size());
    NKPR_HIP(hipStreamSynchronize(g->stream));
    g->b_n_light = (uint32_t)light.size();
    g->b_n_heavy = (uint32_t)hrow.size();
    g->b_n_pieces = (uint32_t)pieces.size();
    g->b_heavy = heavy;
    g->b_split_ready = true;
}

namespace {

struct BatchRun {
    nkpr_graph *g;
    const nkpr_params *p;
    uint32_t k, ld;
    int lpr[2];          // [dbl] autotuned lanes per row
    uint32_t heavy[2];
    unsigned grid[2], cgrid[2];
    BatchCol *col;
    BatchCtl *ctl;
    BatchCtl *ctl_host;
    double *sink_part;
    nkpr_batch_stats *st;
};

template <typename T> BatchArgs<T> make_args(BatchRun &br) {
    nkpr_graph *g = br.g;
    BatchArgs<T> a{};
    a.rowptr = g->b_rowptr.as<uint32_t>();
    a.src = g->b_src.as<int32_t>();
    if constexpr (sizeof(T) == 4) {
        a.coef = g->weighted ? g->b_coef32.as<float>() : nullptr;
        a.inv = g->weighted ? nullptr : g->inv32.as<float>();
        a.tval = g->b_tval32.as<float>();
        a.piece_part = g->b_piece32.as<float>();
    } else {
        a.coef = g->weighted ? g->b_coef64.as<double>() : nullptr;
        a.inv = g->weighted ? nullptr : g->inv64.as<double>();
        a.tval = g->b_tval64.as<double>();
        a.piece_part = g->b_piece64.as<double>();
    }
    a.tptr = g->b_tptr.as<uint32_t>();
    a.tcol = g->b_tcol.as<uint32_t>();
    a.light = g->b_light.as<uint32_t>();
    a.pieces = g->b_pieces.as<uint2>();
    a.hrow = g->b_hrow.as<uint32_t>();
    a.hpiece = g->b_hpiece.as<uint32_t>();
    a.partials = g->b_partials.as<double>();
    a.col = br.col;
    a.ctl = br.ctl;
    a.n = g->n;
    a.ld = br.ld;
    a.n_light = g->b_n_light;
    a.n_pieces = g->b_n_pieces;
    a.n_heavy = g->b_n_heavy;
    a.sink_begin = g->sink_begin;
    a.damp = br.p->damp;
    a.distribute = br.p->sink_handling == NKPR_SINK_DISTRIBUTE ? 1 : 0;
    return a;
}

void size_partials(BatchRun &br, unsigned grid, unsigned cgrid) {
    ensure_buf(br.g->b_partials, 8 * 4 * (size_t)(grid + cgrid) * br.ld);
}

template <typename T>
void size_pieces(BatchRun &br) {
    nkpr_graph *g = br.g;
    DevBuf &pp = sizeof(T) == 4 ? g->b_piece32 : g->b_piece64;
    ensure_buf(pp, sizeof(T) * std::max<size_t>((size_t)g->b_n_pieces * br.ld, 4));
}

template <typename T> void phase_init(BatchRun &br, const T *x) {
    nkpr_graph *g = br.g;
    const hipStream_t s = g->stream;
    const uint32_t n = g->n, ld = br.ld;
    const bool dist = br.p->sink_handling == NKPR_SINK_DISTRIBUTE;
    uint32_t sink_blocks = 0;
    if (dist && g->sink_begin < n) {
        const uint32_t rows = n - g->sink_begin;
        const uint32_t rpb = std::max<uint32_t>(256, (rows + 1023) / 1024);
        sink_blocks = (rows + rpb - 1) / rpb;
        hipLaunchKernelGGL(col_range_sum_kernel<T>, dim3(sink_blocks), dim3(256), 0, s, x, ld,
                           g->sink_begin, n, rpb, br.sink_part);
        NKPR_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(batch_init_kernel, dim3((ld + 255) / 256), dim3(256), 0, s, br.col, br.k,
                       ld, br.ctl, br.sink_part, sink_blocks);
    NKPR_HIP(hipGetLastError());
}

// Sweeps of one phase until every column is done or max_launch sweeps.
template <typename T>
uint64_t batch_phase(BatchRun &br, T *x0, T *x1, int kind, PhaseRule rule, uint64_t max_launch,
                     const float *base, float *r32, double *ms_acc) {
    nkpr_graph *g = br.g;
    const bool dbl = sizeof(T) == 8;
    const hipStream_t s = g->stream;
    batch_split(g, br.heavy[dbl]);
    size_pieces<T>(br);
    const unsigned grid = br.grid[dbl], cgrid = br.cgrid[dbl];
    size_partials(br, grid, cgrid);
    BatchArgs<T> a = make_args<T>(br);
    a.base = base;
    a.r32 = r32;
    const uint32_t nblocks = grid + (a.n_heavy ? cgrid : 0);
    const bool timing = br.p->timing != 0;
    if (timing) {
        while (g->events.size() < 128) {
            hipEvent_t e;
            NKPR_HIP(hipEventCreate(&e));
            g->events.push_back(e);
        }
    }
    T *xs[2] = {x0, x1};
    uint64_t launched = 0, done_sweeps = 0;
    uint64_t batch = (max_launch <= 2) ? max_launch : 4;
    while (launched < max_launch) {
        if (br.p->interrupt && *br.p->interrupt) throw Error(NKPR_ERR_INTERRUPTED, "interrupted");
        const uint64_t nb = std::min<uint64_t>(std::min<uint64_t>(batch, 64), max_launch - launched);
        for (uint64_t j = 0; j < nb; ++j) {
            const uint64_t q = launched + j;
            a.xold = xs[q & 1];
            a.xnew = xs[(q & 1) ^ 1];
            if (timing) NKPR_HIP(hipEventRecord(g->events[2 * j], s));
            launch_spmm<T>(a, kind, br.lpr[dbl], g->weighted, grid, cgrid, s);
            if (timing) NKPR_HIP(hipEventRecord(g->events[2 * j + 1], s));
            hipLaunchKernelGGL(batch_finalize_kernel, dim3(br.k), dim3(kFinThreads), 0, s,
                               g->b_partials.as<double>(), nblocks, br.ld, br.col,
                               (const BatchCtl *)br.ctl, rule);
            NKPR_HIP(hipGetLastError());
            hipLaunchKernelGGL(batch_count_kernel, dim3(1), dim3(256), 0, s, (const BatchCol *)br.col,
                               br.k, br.ctl, (uint64_t)max_launch);
            NKPR_HIP(hipGetLastError());
        }
        read_back(s, {{br.ctl_host, br.ctl, sizeof(BatchCtl)}});
        const uint64_t it = br.ctl_host->sweeps;
        if (timing) {
            for (uint64_t j = 0; j < it - done_sweeps && j < nb; ++j) {
                float ms = 0.f;
                NKPR_HIP(hipEventElapsedTime(&ms, g->events[2 * j], g->events[2 * j + 1]));
                *ms_acc += ms;
            }
        }
        done_sweeps = it;
        launched += nb;
        if (br.ctl_host->done) break;
        batch *= 2;
    }
    br.st->sweeps += done_sweeps;
    return done_sweeps;
}

// Times candidate (lanes per row, heavy threshold) pairs on the start vector; keeps the best.
template <typename T>
void autotune(BatchRun &br, T *x0, T *x1) {
    nkpr_graph *g = br.g;
    const bool dbl = sizeof(T) == 8;
    const hipStream_t s = g->stream;
    constexpr int V = V16<T>::n;
    const uint32_t key = br.ld * 2 + (dbl ? 1 : 0);
    // NKPR_PPR_LPR (4, 16 or 64 lanes per row) and NKPR_PPR_HEAVY (a split threshold >= 64)
    // restrict the candidates to the given value: a forced configuration is neither looked up
    // nor remembered
    const uint32_t force_lpr = env_u32_batch("NKPR_PPR_LPR", 0);
    const uint32_t force_heavy = env_u32_batch("NKPR_PPR_HEAVY", 0);
    NKPR_CHECK(force_lpr == 0 || force_lpr == 4 || force_lpr == 16 || force_lpr == 64,
               "NKPR_PPR_LPR must be 4, 16 or 64");
    NKPR_CHECK(force_heavy == 0 || force_heavy >= 64, "NKPR_PPR_HEAVY must be >= 64");
    const bool forced = force_lpr != 0 || force_heavy != 0;
    for (const auto &t : g->b_tuned) {
        if (t.key == key && !forced) {
            br.lpr[dbl] = t.lpr;
            br.heavy[dbl] = t.heavy;
            return;
        }
    }
    std::vector<int> lprs;
    if (force_lpr) {
        lprs.push_back((int)force_lpr);
    } else {
        for (int l : {64, 16, 4})  // keep tiles at least half full
            if (l == 4 || (uint32_t)(l * V) <= 2 * br.ld) lprs.push_back(l);
    }
    // heavier rows than this leave one row group running alone for too long
    std::vector<uint32_t> heavies = {256, 1024, 4096};
    if (force_heavy) heavies = {force_heavy};
    hipEvent_t e0, e1;
    NKPR_HIP(hipEventCreate(&e0));
    NKPR_HIP(hipEventCreate(&e1));
    float best = std::numeric_limits<float>::infinity();
    int best_l = lprs.front();
    uint32_t best_h = heavies.front();
    phase_init<T>(br, x0);
    for (uint32_t h : heavies) {
        batch_split(g, h);
        size_pieces<T>(br);
        for (int l : lprs) {
            br.lpr[dbl] = l;
            const unsigned grid = (unsigned)occupancy_blocks<T>(g->device, l, g->weighted);
            const unsigned cgrid = std::max<unsigned>(1, std::min<unsigned>(grid, (g->b_n_heavy * l + kBThreads - 1) / kBThreads));
            size_partials(br, grid, cgrid);
            BatchArgs<T> a = make_args<T>(br);
            a.xold = x0;
            a.xnew = x1;
            launch_spmm<T>(a, kBPlain, l, g->weighted, grid, cgrid, s);  // warm-up
            NKPR_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < 2; ++r) launch_spmm<T>(a, kBPlain, l, g->weighted, grid, cgrid, s);
            NKPR_HIP(hipEventRecord(e1, s));
            NKPR_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            NKPR_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                best_l = l;
                best_h = h;
            }
        }
        // with no heavy row the threshold does not matter
        if (g->b_n_heavy == 0) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    br.lpr[dbl] = best_l;
    br.heavy[dbl] = best_h;
    if (!forced) g->b_tuned.push_back({key, best_l, best_h});
}

template <typename T> void configure(BatchRun &br, T *x0, T *x1) {
    nkpr_graph *g = br.g;
    const bool dbl = sizeof(T) == 8;
    autotune<T>(br, x0, x1);
    batch_split(g, br.heavy[dbl]);
    br.grid[dbl] = (unsigned)occupancy_blocks<T>(g->device, br.lpr[dbl], g->weighted);
    br.cgrid[dbl] = std::max<unsigned>(
        1, std::min<unsigned>(br.grid[dbl], (g->b_n_heavy * br.lpr[dbl] + kBThreads - 1) / kBThreads));
}

}  // namespace

void ppr_batch_impl(nkpr_graph *g, const nkpr_params *p, uint32_t k, const uint64_t *seed_off,
                    const int32_t *seeds, const double *seed_w, double *scores,
                    nkpr_batch_stats *st, uint64_t *col_iterations, double *col_bound) {
    const auto t0 = std::chrono::steady_clock::now();
    NKPR_CHECK(g != nullptr && p != nullptr, "graph or params is NULL");
    NKPR_CHECK(p->damp > 0.0 && p->damp < 1.0, "damp must be in (0, 1)");
    NKPR_CHECK(p->norm == NKPR_L1 || p->norm == NKPR_L2, "norm must be NKPR_L1 or NKPR_L2");
    NKPR_CHECK(p->sink_handling == NKPR_SINK_NONE || p->sink_handling == NKPR_SINK_DISTRIBUTE,
               "unknown sink handling");
    NKPR_CHECK(p->precision >= NKPR_FP64 && p->precision <= NKPR_FP32, "unknown precision");
    NKPR_CHECK(p->stop == NKPR_STOP_STEP || p->stop == NKPR_STOP_L1_ERROR, "unknown stop rule");
    NKPR_CHECK(!(p->tol < 0.0) && !std::isnan(p->tol), "tol must be >= 0");
    NKPR_CHECK(!p->normalized, "batched personalised PageRank supports normalized = 0 only");
    NKPR_CHECK(p->teleport == nullptr, "batched runs take seed sets, not params->teleport");
    NKPR_CHECK(k == 0 || (seed_off != nullptr && scores != nullptr), "seed offsets or scores is NULL");
    std::memset(st, 0, sizeof(*st));
    if (k == 0 || g->n == 0) return;
    const uint32_t n = g->n;
    const uint64_t z = g->z;
    NKPR_CHECK(seed_off[0] == 0, "seed_offsets[0] must be 0");
    for (uint32_t j = 0; j < k; ++j)
        NKPR_CHECK(seed_off[j + 1] > seed_off[j], "every teleport column needs at least one seed");
    const uint64_t ns = seed_off[k];
    NKPR_CHECK(seeds != nullptr, "seed_nodes is NULL");
    const hipStream_t s = g->stream;
    batch_layout(g);

    // ---- row-sparse teleport in relabelled ids ----------------------------------------------
    // Entries (row, column, weight) in column order, sorted stably by row (so columns ascend
    // within a row and repeated seeds of a column stay in seed order), repeated (row, column)
    // pairs merged in that order; the rows with entries and their offsets go to the device,
    // which expands them into the row pointer.
    const std::vector<int32_t> &iperm = g->b_iperm_host;
    struct Ent {
        uint32_t row, col;
        double w;
    };
    std::vector<Ent> ent;
    ent.reserve(ns);
    const double d = p->damp;
    for (uint32_t j = 0; j < k; ++j) {
        double wsum = 0.0;
        for (uint64_t i = seed_off[j]; i < seed_off[j + 1]; ++i) {
            const int32_t v = seeds[i];
            NKPR_CHECK(v >= 0 && (uint64_t)v < z && iperm[v] >= 0,
                       "seed node " + std::to_string(v) + " is not a node of the graph");
            const double w = seed_w ? seed_w[i] : 1.0;
            NKPR_CHECK(w >= 0.0 && std::isfinite(w), "seed weights must be finite and >= 0");
            wsum += w;
        }
        NKPR_CHECK(wsum > 0.0, "seed weights of a column sum to 0");
        for (uint64_t i = seed_off[j]; i < seed_off[j + 1]; ++i)
            ent.push_back({(uint32_t)iperm[seeds[i]], j, (1.0 - d) * (seed_w ? seed_w[i] : 1.0) / wsum});
    }
    {
        // stable LSD radix sort by row, 11 bits per pass
        std::vector<Ent> tmp(ent.size());
        for (int shift = 0; shift < 32; shift += 11) {
            std::vector<uint32_t> cnt(2049, 0);
            for (const Ent &e : ent) ++cnt[((e.row >> shift) & 2047u) + 1];
            if (cnt[((ent.empty() ? 0u : (ent[0].row >> shift) & 2047u)) + 1] == ent.size()) continue;
            for (int b = 0; b < 2048; ++b) cnt[b + 1] += cnt[b];
            for (const Ent &e : ent) tmp[cnt[(e.row >> shift) & 2047u]++] = e;
            ent.swap(tmp);
        }
    }
    std::vector<uint32_t> urow, uoff, tcol;
    std::vector<double> tv64;
    for (size_t i = 0; i < ent.size(); ++i) {
        if (i > 0 && ent[i].row == ent[i - 1].row && ent[i].col == ent[i - 1].col) {
            tv64.back() += ent[i].w;
            continue;
        }
        if (urow.empty() || urow.back() != ent[i].row) {
            urow.push_back(ent[i].row);
            uoff.push_back((uint32_t)tcol.size());
        }
        tcol.push_back(ent[i].col);
        tv64.push_back(ent[i].w);
    }
    uoff.push_back((uint32_t)tcol.size());
    std::vector<float> tv32(tv64.size());
    for (size_t i = 0; i < tv64.size(); ++i) tv32[i] = (float)tv64[i];
    auto up = [&](DevBuf &b, const void *h, size_t bytes) {
        b.alloc(std::max<size_t>(bytes, 16));
        if (bytes) NKPR_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
    };
    up(g->b_tcol, tcol.data(), 4 * tcol.size());
    up(g->b_tval32, tv32.data(), 4 * tv32.size());
    up(g->b_tval64, tv64.data(), 8 * tv64.size());
    {
        DevBuf urow_d, uoff_d;
        up(urow_d, urow.data(), 4 * urow.size());
        up(uoff_d, uoff.data(), 4 * uoff.size());
        g->b_tptr.alloc(4 * ((size_t)n + 1));
        hipLaunchKernelGGL(tele_rowptr_kernel, dim3(grid_for((uint64_t)n + 1)), dim3(256), 0, s,
                           (const uint32_t *)urow_d.as<uint32_t>(), (const uint32_t *)uoff_d.as<uint32_t>(),
                           (uint32_t)urow.size(), n, g->b_tptr.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        // the staging buffers are freed when this block ends: the expansion must be done first
        NKPR_HIP(hipStreamSynchronize(s));
    }

    // ---- buffers -------------------------------------------------------------------------------
    const uint32_t ld = (k + 3u) & ~3u;  // 16 bytes of FP32 columns; FP64 uses 2 per lane
    const size_t cells = (size_t)n * ld;
    BatchRun br{};
    br.g = g;
    br.p = p;
    br.k = k;
    br.ld = ld;
    br.st = st;
    ensure_buf(g->b_col, sizeof(BatchCol) * ld);
    ensure_buf(g->b_ctl, sizeof(BatchCtl));
    if (!g->b_ctl_host) NKPR_HIP(hipHostMalloc((void **)&g->b_ctl_host, sizeof(BatchCtl), hipHostMallocDefault));
    br.col = g->b_col.as<BatchCol>();
    br.ctl = g->b_ctl.as<BatchCtl>();
    br.ctl_host = g->b_ctl_host;
    ensure_buf(g->b_sink_part, 8 * (size_t)1025 * ld);
    br.sink_part = g->b_sink_part.as<double>();
    NKPR_HIP(hipMemsetAsync(br.col, 0, sizeof(BatchCol) * ld, s));
    ensure_buf(g->b_x64a, 8 * cells);
    ensure_buf(g->b_x64b, 8 * cells);
    double *x64[2] = {g->b_x64a.as<double>(), g->b_x64b.as<double>()};

    double *sz = scores;
    if (!p->scores_on_device) {
        ensure_buf(g->b_zbuf, 8 * (size_t)z * k);
        sz = g->b_zbuf.as<double>();
        NKPR_HIP(hipMemcpyAsync(sz, scores, 8 * (size_t)z * k, hipMemcpyHostToDevice, s));
    }
    const int32_t *slot_u = slot_index(g);
    const uint64_t gather_tiles = ((z + kTile - 1) / kTile) * ((ld + kTile - 1) / kTile);
    hipLaunchKernelGGL(batch_gather_kernel, dim3((unsigned)std::min<uint64_t>(std::max<uint64_t>(gather_tiles, 1), 1u << 16)),
                       dim3(256), 0, s, (const double *)sz, z, slot_u, k, ld,
                       x64[0]);
    NKPR_HIP(hipGetLastError());

    const uint64_t max_it = std::max<uint64_t>(p->max_iterations, 1);
    double *result = nullptr;
    const bool debug = std::getenv("NKPR_DEBUG") != nullptr;
    auto debug_phase = [&](const char *name, uint64_t sweeps) {
        if (!debug) return;
        std::vector<BatchCol> c(k);
        read_back(s, {{c.data(), br.col, sizeof(BatchCol) * k}});
        double mb = 0.0, ml1 = 0.0;
        uint32_t fin = 0;
        for (const auto &x : c) {
            mb = std::max(mb, x.bound);
            ml1 = std::max(ml1, x.l1);
            fin += x.final_;
        }
        std::fprintf(stderr, "[nkpr batch] %s: %llu sweeps, max bound %.3e, max l1 %.3e, certified %u/%u\n",
                     name, (unsigned long long)sweeps, mb, ml1, fin, k);
    };
    if (p->precision == NKPR_FP64) {
        configure<double>(br, x64[0], x64[1]);
        PhaseRule r{};
        r.kind = p->stop == NKPR_STOP_STEP ? kRuleStep : kRuleErr;
        r.norm = p->norm;
        r.tol = p->tol;
        r.max_iterations = max_it;
        r.damp = d;
        phase_init<double>(br, x64[0]);
        const uint64_t K = batch_phase<double>(br, x64[0], x64[1], kBPlain, r, max_it, nullptr, nullptr, &st->ms_sweeps_fp64);
        st->sweeps_fp64 = K;
        result = x64[K & 1];
    } else {
        ensure_buf(g->b_x32a, 4 * cells);
        ensure_buf(g->b_x32b, 4 * cells);
        float *x32[2] = {g->b_x32a.as<float>(), g->b_x32b.as<float>()};
        hipLaunchKernelGGL(f64_to_f32_kernel, dim3(grid_for(cells)), dim3(256), 0, s,
                           (const double *)x64[0], x32[0], (uint64_t)cells);
        NKPR_HIP(hipGetLastError());
        configure<float>(br, x32[0], x32[1]);
        if (p->precision == NKPR_FP32) {
            PhaseRule r{};
            r.kind = p->stop == NKPR_STOP_STEP ? kRuleStep : kRuleErr;
            r.norm = p->norm;
            r.tol = p->tol;
            r.max_iterations = max_it;
            r.damp = d;
            phase_init<float>(br, x32[0]);
            const uint64_t K = batch_phase<float>(br, x32[0], x32[1], kBPlain, r, max_it, nullptr, nullptr, &st->ms_sweeps_fp32);
            st->sweeps_fp32 = K;
            hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(cells)), dim3(256), 0, s,
                               (const float *)x32[K & 1], x64[0], (uint64_t)cells);
            NKPR_HIP(hipGetLastError());
            result = x64[0];
        } else {
            // FP32 power sweeps to ~1e-5, then FP64 residual sweeps certify every column and
            // FP32 correction solves (e = d P e + r, power iteration) refine the rest.
            ensure_buf(g->b_r32, 4 * cells);
            float *r32 = g->b_r32.as<float>();
            PhaseRule ra{};
            ra.kind = kRuleAbsStall;
            ra.tol = std::max(p->tol, 2e-5);
            ra.tol_abs = 0.0;
            ra.stall_ratio = 0.95;
            ra.max_iterations = max_it;
            ra.damp = d;
            phase_init<float>(br, x32[0]);
            uint64_t total = batch_phase<float>(br, x32[0], x32[1], kBPlain, ra, max_it, nullptr, nullptr, &st->ms_sweeps_fp32);
            debug_phase("fp32 power", total);
            st->sweeps_fp32 += total;
            hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(cells)), dim3(256), 0, s,
                               (const float *)x32[total & 1], x64[0], (uint64_t)cells);
            NKPR_HIP(hipGetLastError());
            configure<double>(br, x64[0], x64[1]);
            for (int round = 0;; ++round) {
                PhaseRule rr{};
                rr.kind = kRuleErr;
                rr.norm = NKPR_L1;
                rr.tol = p->tol;
                rr.max_iterations = 1;
                rr.damp = d;
                phase_init<double>(br, x64[0]);
                total += batch_phase<double>(br, x64[0], x64[1], kBResid, rr, 1, nullptr, r32, &st->ms_sweeps_fp64);
                debug_phase("fp64 residual", 1);
                st->sweeps_fp64 += 1;
                result = x64[1];
                std::vector<BatchCol> cols(k);
                read_back(s, {{br.ctl_host, br.ctl, sizeof(BatchCtl)},
                              {cols.data(), br.col, sizeof(BatchCol) * k}});
                uint32_t open = 0;
                for (uint32_t j = 0; j < k; ++j) open += cols[j].final_ ? 0u : 1u;
                if (open == 0 || total >= max_it || round >= 30) break;
                ++st->refinements;
                // correction: e0 = r; per column stop at its own target step
                hipLaunchKernelGGL(copy_f32_kernel, dim3(grid_for(cells)), dim3(256), 0, s,
                                   (const float *)r32, x32[0], (uint64_t)cells);
                NKPR_HIP(hipGetLastError());
                if (g->sink_begin == n || p->sink_handling == NKPR_SINK_DISTRIBUTE) {
                    const uint32_t rpb = std::max<uint32_t>(256, (n + 1023) / 1024);
                    const uint32_t blocks = (n + rpb - 1) / rpb;
                    hipLaunchKernelGGL(col_range_sum_kernel<float>, dim3(blocks), dim3(256), 0, s,
                                       (const float *)r32, ld, 0u, n, rpb, br.sink_part);
                    double *tot = br.sink_part + (size_t)1024 * ld;
                    hipLaunchKernelGGL(col_total_kernel, dim3((ld + 255) / 256), dim3(256), 0, s,
                                       (const double *)br.sink_part, blocks, ld, tot);
                    hipLaunchKernelGGL(col_shift_kernel, dim3(grid_for(cells)), dim3(256), 0, s,
                                       x32[0], ld, (uint64_t)cells, (const double *)tot,
                                       d / (1.0 - d) / (double)n);
                    NKPR_HIP(hipGetLastError());
                }
                hipLaunchKernelGGL(batch_set_solve_tol_kernel, dim3((ld + 255) / 256), dim3(256), 0, s,
                                   br.col, ld);
                NKPR_HIP(hipGetLastError());
                PhaseRule rc{};
                rc.kind = kRuleAbsStall;
                rc.tol = -1.0;
                rc.tol_abs = -1.0;  // per-column target
                rc.stall_ratio = 0.95;
                rc.max_iterations = (max_it > total + 1) ? max_it - total - 1 : 1;
                rc.damp = d;
                phase_init<float>(br, x32[0]);
                const uint64_t kc = batch_phase<float>(br, x32[0], x32[1], kBBase, rc, rc.max_iterations, r32, nullptr, &st->ms_sweeps_fp32);
                debug_phase("fp32 correction", kc);
                st->sweeps_fp32 += kc;
                total += kc;
                hipLaunchKernelGGL(batch_update_kernel, dim3(grid_for(cells)), dim3(256), 0, s,
                                   x64[0], (const double *)x64[1], (const float *)x32[kc & 1], ld,
                                   (uint64_t)cells, (const BatchCol *)br.col);
                NKPR_HIP(hipGetLastError());
            }
        }
    }

    // ---- per-column normalisation by the column sum, scatter to [k][z] ------------------------
    std::vector<BatchCol> cols(k);
    read_back(s, {{cols.data(), br.col, sizeof(BatchCol) * k}});
    std::vector<double> factor(k);
    for (uint32_t j = 0; j < k; ++j) {
        NKPR_CHECK(cols[j].sum > 0.0, "a column's scores sum to zero: cannot normalise");
        factor[j] = cols[j].sum;
        st->max_iterations_col = std::max<uint64_t>(st->max_iterations_col, cols[j].iterations);
        // an FP32 run certifies nothing (its iterates carry rounding the step does not see)
        const double b = (p->precision == NKPR_MIXED || p->stop == NKPR_STOP_L1_ERROR)
                             ? (p->precision == NKPR_FP32 ? INFINITY : cols[j].bound)
                             : (p->norm == NKPR_L2 ? std::sqrt(cols[j].l2) : cols[j].l1);
        st->max_error_bound = std::max(st->max_error_bound, b);
        if (col_iterations) col_iterations[j] = cols[j].iterations;
        if (col_bound) col_bound[j] = b;
    }
    ensure_buf(g->b_factor, 8 * (size_t)k);
    NKPR_HIP(hipMemcpyAsync(g->b_factor.p, factor.data(), 8 * (size_t)k, hipMemcpyHostToDevice, s));
    const uint64_t scatter_tiles = ((z + kTile - 1) / kTile) * ((k + kTile - 1) / kTile);
    hipLaunchKernelGGL(batch_scatter_kernel, dim3((unsigned)std::min<uint64_t>(std::max<uint64_t>(scatter_tiles, 1), 1u << 16)),
                       dim3(256), 0, s, (const double *)result, slot_u, k, ld, z,
                       g->b_factor.as<double>(), sz);
    NKPR_HIP(hipGetLastError());
    read_back(s, {{scores, sz, p->scores_on_device ? 0 : 8 * (size_t)z * k}});
    st->tile_columns = (uint32_t)(br.lpr[p->precision == NKPR_FP64 ? 1 : 0] * (p->precision == NKPR_FP64 ? 2 : 4));
    st->heavy_degree = br.heavy[p->precision == NKPR_FP64 ? 1 : 0];
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace nkpr
