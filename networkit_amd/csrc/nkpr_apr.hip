// Approximate personalised PageRank by residual pushes (ACL), the kernel of NetworKit's
// ApproximatePageRank (networkit/cpp/scd/ApproximatePageRank.cpp:18-61) and PageRankNibble.
//
// The reference keeps a FIFO queue and pushes one node at a time:
//   push(u): res = r[u]; vol = weightedDegree(u, countSelfLoopsTwice = true);
//            every neighbour v (out-list order) gets mass = (1 - alpha) res w / (2 vol) and is
//            queued when its residual crosses eps vol(v); p[u] += alpha res;
//            r[u] = (1 - alpha) res / 2, queued again when r[u] / vol >= eps.
// Here all queued nodes push at once, round after round (a frontier), a wave per node:
//   a first kernel takes every frontier node's residual (so that mass arriving during the
//   round stays for the next one, whatever the schedule), then a wave per node adds alpha res
//   to p, sends the masses with atomic adds and keeps (1 - alpha) res / 2; a node whose
//   residual crosses its threshold through an add is appended to the next frontier once (a
//   flag bit per node).
// Residuals are 64-bit fixed point (2^-62 units; the total mass is at most 1), so the atomic
// sums are exact and independent of their order: the frontier of every round, p and r are the
// same on every run, bit for bit (only the order of a frontier's entries varies).  Each sent
// and kept mass is rounded to the nearest unit (at most 2^-63 off, the errors of a run add up
// like a random walk, about sqrt(edge pushes) 2^-63); thresholds are ceil(eps vol 2^62), so
// r < eps vol holds exactly for every residual below 2^-9.
// Every push moves mass as the reference's does, so the ACL identity
// p + pr_alpha(r) = pr_alpha(seeds) and the stopping rule (every residual below eps vol when the
// frontier is empty) hold as in the reference; the push order differs, so p differs from the
// reference's within the bound that rule gives.  The share a self-loop sends back to u is
// dropped, as in the reference (its r[u] = (1 - alpha) res / 2 is an assignment after the
// loop), so on graphs with self-loops the identity holds for the walk without them,
// (I + D^-1 A_offdiag) / 2 with D counting each self-loop twice.
//
// Layout: the graph's out-lists (forNeighborsOf order) as CSR with int32 ids and optional FP64
// weights, vol[z] computed once; p[z] (FP64), r[z] (fixed point), flags[z] (bit 0 touched, bit 1
// queued), two frontier arrays and the taken residuals.  Only the touched entries are cleared
// after a run, so a run costs O(z) only for the scan that lists the touched nodes.
// Algorithmic bytes per push: 4 + 8 (id, weight) and one 8-byte atomic per out-edge, 8 + 8 + 8
// per pushed node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nkpr_internal.hpp"

struct nkpr_apr {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t z = 0, m = 0;
    nkpr::DevBuf ptr, dst, w, vol;
    nkpr::DevBuf p, r, flags, f0, f1, taken, chunks, cnt, out_ids, out_p, out_r, out_cnt, seeds;
    uint64_t chunk_cap = 0;  // entries per heavy chunk list (two lists)
    bool dirty = true;  // p, r, flags hold a run that was not cleared
    ~nkpr_apr() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace nkpr {
namespace {

constexpr uint32_t kTouched = 1, kQueued = 2;
constexpr int kPushThreads = 256;
// rounds per host read of the counters: rounds after the frontier empties return at once, so a
// longer batch trades a few no-op launches for fewer host reads (RHG n = 10M, 128 rounds:
// 8.91 ms per run with 8, 8.67 ms with 16, 8.51 ms with 32)
constexpr int kBatch = 32;
constexpr double kUnit = 4611686018427387904.0;  // 2^62 fixed-point units per unit of mass

// ceil(eps vol 2^62): r < thr  <=>  r 2^-62 < eps vol (the product rounded as the reference's).
// A node without out-edges (vol 0) follows the reference's two tests: a neighbour's add never
// queues it (r[v] < vol eps is false), and as a pushed node (only a seed can be one) it is
// pushed again while any residual is left (r / vol >= eps holds for every r > 0).
__device__ __forceinline__ uint64_t threshold_of(double eps, double vol, bool self) {
    const double t = eps * vol * kUnit;
    if (!(t > 0.0)) return self ? 1ull : ~0ull;
    return t >= 1.8e19 ? ~0ull : (uint64_t)ceil(t);
}

// offsets rising from 0 (ptr[z] = m by construction): bad[0] counts the others
__global__ void apr_check_kernel(const uint64_t *ptr, uint64_t z, uint32_t *bad) {
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < z;
         u += (uint64_t)gridDim.x * blockDim.x)
        if (ptr[u] > ptr[u + 1] || (u == 0 && ptr[0] != 0)) atomicAdd(bad, 1u);
}

__global__ void apr_vol_kernel(const uint64_t *ptr, const int32_t *dst, const double *w,
                               uint64_t z, double *vol, uint32_t *bad) {
    // Graph::weightedDegree(u, countSelfLoopsTwice = true): a wave per node; bad[0] counts
    // weights that are negative or not numbers and ids outside [0, z)
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t u = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; u < z; u += nw) {
        double s = 0.0;
        for (uint64_t e = ptr[u] + lane; e < ptr[u + 1]; e += 64) {
            const double x = w ? w[e] : 1.0;
            const int32_t v = dst[e];
            if (!(x >= 0.0) || v < 0 || (uint64_t)v >= z) atomicAdd(bad, 1u);
            s += (uint64_t)v == u ? 2.0 * x : x;
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) vol[u] = s;
    }
}

__global__ void apr_clear_kernel(double *p, uint64_t *r, uint32_t *flags, uint64_t z) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < z;
         i += (uint64_t)gridDim.x * blockDim.x) {
        p[i] = 0.0;
        r[i] = 0ull;
        flags[i] = 0u;
    }
}

// after a run: clear the entries it touched (listed by the compaction)
__global__ void apr_reset_kernel(const uint32_t *ids, const uint32_t *count, double *p,
                                 uint64_t *r, uint32_t *flags) {
    const uint32_t k = count[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        const uint32_t u = ids[i];
        p[u] = 0.0;
        r[u] = 0ull;
        flags[u] = 0u;
    }
}

// Counters (cnt, a 64-byte device buffer of uint32).  Round k pushes frontier k & 1 (size
// cnt[k % 3]) and appends to the other frontier array (size cnt[(k + 1) % 3]); its take kernel
// clears cnt[(k + 2) % 3], the size round k - 1 pushed, for round k + 1's appends.  The heavy
// chunk lists alternate the same way (sizes cnt[kCntChunks + (k & 1)]).  [4..5] the out-edges
// of the pushed nodes, [6..7] the pushes, [8] the rounds with a non-empty frontier.
constexpr int kCntEdges = 4, kCntPushes = 6, kCntRounds = 8, kCntChunks = 10;
// A node with more out-edges than kPushChunk is pushed by several waves, kPushChunk edges each.
constexpr uint32_t kPushChunk = 256;

// the seeds: residual 1/|seeds|, touched, queued, in frontier 0 (ids ascending)
__global__ void apr_seed_kernel(const uint32_t *seeds, uint32_t k, uint64_t *r, uint32_t *flags,
                                uint32_t *front, uint32_t *cnt) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        const uint32_t s = seeds[i];
        r[s] = (uint64_t)(kUnit / (double)k);
        flags[s] = kTouched | kQueued;
        front[i] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x < 16) cnt[threadIdx.x] = threadIdx.x == 0 ? k : 0u;
}

// Appends the nodes of the lanes with `append` set at next[*ncnt ...], one counter update per
// wave (every lane of the wave calls it).
__device__ __forceinline__ void append_wave(bool append, uint32_t v, uint32_t *next,
                                            uint32_t *ncnt) {
    const uint64_t ball = __ballot(append);
    if (!ball) return;
    const int lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(ncnt, (uint32_t)__popcll(ball));
    base = __shfl(base, 0, 64);
    if (append) next[base + (uint32_t)__popcll(ball & ((1ull << lane) - 1))] = v;
}

// Round k, first kernel: a thread per frontier node u takes its residual res (kept in taken[]),
// adds alpha res to p[u] and leaves (1 - alpha) res / 2 in r[u], appending u to the next
// frontier when that alone reaches its threshold; a node with more than kPushChunk out-edges
// lists its chunks for the push kernel.  The push kernel then only adds to residuals.
__global__ void apr_take_kernel(const uint64_t *ptr, const double *vol, double alpha, double eps,
                                const uint32_t *front, uint32_t *next, uint32_t *cnt, int k3,
                                int k2, double *p, uint64_t *r, uint32_t *flags, uint64_t *taken,
                                uint2 *chunks) {
    __shared__ unsigned long long bedges;
    const uint32_t nf = cnt[k3];
    if (threadIdx.x == 0) bedges = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt[(k3 + 2) % 3] = 0u;
        cnt[kCntChunks + (k2 ^ 1)] = 0u;
        if (nf) {
            *reinterpret_cast<unsigned long long *>(cnt + kCntPushes) += nf;
            cnt[kCntRounds] += 1u;
        }
    }
    __syncthreads();
    uint32_t *ncnt = cnt + (k3 + 1) % 3;
    uint32_t *hcnt = cnt + kCntChunks + k2;
    const uint32_t stride = gridDim.x * blockDim.x;
    // whole waves step together (append_wave is a wave operation)
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < nf; i0 += stride) {
        const uint32_t i = i0 + threadIdx.x;
        bool again = false;
        uint32_t u = 0;
        if (i < nf) {
            u = front[i];
            const uint64_t res = r[u];
            taken[i] = res;
            const double dres = (double)res;
            p[u] += alpha * dres / kUnit;
            const uint64_t keep = __double2ull_rn((1.0 - alpha) * dres / 2);
            r[u] = keep;
            again = keep >= threshold_of(eps, vol[u], true);
            flags[u] = again ? (kTouched | kQueued) : kTouched;
            const uint64_t deg = ptr[u + 1] - ptr[u];
            atomicAdd(&bedges, (unsigned long long)deg);
            if (deg > kPushChunk) {
                const uint32_t nc = (uint32_t)((deg + kPushChunk - 1) / kPushChunk);
                const uint32_t at = atomicAdd(hcnt, nc);
                for (uint32_t c = 0; c < nc; ++c) chunks[at + c] = make_uint2(i, c);
            }
        }
        append_wave(again, u, next, ncnt);
    }
    __syncthreads();
    if (threadIdx.x == 0 && bedges)
        atomicAdd(reinterpret_cast<unsigned long long *>(cnt + kCntEdges), bedges);
}

// Round k, second kernel: a wave per frontier node with at most kPushChunk out-edges and a wave
// per chunk of the others sends the masses (1 - alpha) res w / (2 vol).  At the start of the
// kernel every residual lies below its threshold unless its node is already queued, and all
// adds are positive, so the add that crosses a threshold is the one that finds the residual
// below it and leaves it at or above it; that add sets the node's queued bit.
__global__ __launch_bounds__(kPushThreads) void apr_push_kernel(
    const uint64_t *ptr, const int32_t *dst, const double *w, const double *vol, double alpha,
    double eps, uint64_t *r, uint32_t *flags, const uint32_t *front, const uint64_t *taken,
    const uint2 *chunks, uint32_t *next, uint32_t *cnt, int k3, int k2) {
    const uint32_t nf = cnt[k3], nh = cnt[kCntChunks + k2];
    uint32_t *ncnt = cnt + (k3 + 1) % 3;
    const int lane = threadIdx.x & 63;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < nf + nh; t += nw) {
        uint32_t i;
        uint64_t b, end;
        if (t < nf) {
            i = t;
            const uint32_t u = front[i];
            b = ptr[u];
            end = ptr[u + 1];
            if (end - b > kPushChunk) continue;  // pushed by its chunks
        } else {
            const uint2 c = chunks[t - nf];
            i = c.x;
            const uint32_t u = front[i];
            b = ptr[u] + (uint64_t)c.y * kPushChunk;
            end = std::min<uint64_t>(ptr[u + 1], b + kPushChunk);
        }
        const uint32_t u = front[i];
        const double vu = vol[u];
        const double share = vu > 0.0 ? (1.0 - alpha) * (double)taken[i] / (2.0 * vu) : 0.0;
        for (uint64_t e0 = b; e0 < end; e0 += 64) {
            const uint64_t e = e0 + lane;
            bool append = false;
            uint32_t v = 0;
            // a self-loop sends nothing: the reference overwrites r[u] after its loop
            if (e < end && (v = (uint32_t)dst[e]) != u) {
                const uint64_t mass = __double2ull_rn(share * (w ? w[e] : 1.0));
                const uint64_t thr = threshold_of(eps, vol[v], false);
                const uint64_t old = atomicAdd(reinterpret_cast<unsigned long long *>(&r[v]),
                                               (unsigned long long)mass);
                if (old < thr && old + mass >= thr)
                    append = !(atomicOr(&flags[v], kQueued | kTouched) & kQueued);
                else if (!(flags[v] & kTouched))
                    atomicOr(&flags[v], kTouched);
            }
            append_wave(append, v, next, ncnt);
        }
    }
}

// touched nodes in ascending id order: per-block counts, then a one-block scan, then writes
__global__ void apr_count_kernel(const uint32_t *flags, uint64_t z, uint32_t *bc) {
    __shared__ uint32_t s[256];
    const uint64_t per = (z + gridDim.x - 1) / gridDim.x;
    const uint64_t b = (uint64_t)blockIdx.x * per, e = std::min<uint64_t>(z, b + per);
    uint32_t c = 0;
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) c += (flags[i] & kTouched) ? 1u : 0u;
    s[threadIdx.x] = c;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) bc[blockIdx.x] = s[0];
}

// exclusive sums of the nb (= blockDim.x, a multiple of 64) block counts, one thread each
__global__ void apr_scan_kernel(uint32_t *bc, uint32_t nb, uint32_t *total) {
    __shared__ uint32_t wsum[16];
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const uint32_t c = t < nb ? bc[t] : 0u;
    uint32_t x = c;  // inclusive scan inside the wave
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if ((int)lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (uint32_t i = 0; i < blockDim.x / 64; ++i) {
            const uint32_t v = wsum[i];
            wsum[i] = acc;
            acc += v;
        }
        total[0] = acc;
    }
    __syncthreads();
    if (t < nb) bc[t] = wsum[wv] + x - c;
}

__global__ void apr_compact_kernel(const uint32_t *flags, const double *p, const uint64_t *r,
                                   uint64_t z, const uint32_t *bc, uint32_t *ids, double *op,
                                   double *orr) {
    // one block per range; the range is walked in order 256 slots at a time
    __shared__ uint32_t s[256];
    __shared__ uint32_t base;
    const uint64_t per = (z + gridDim.x - 1) / gridDim.x;
    const uint64_t b = (uint64_t)blockIdx.x * per, e = std::min<uint64_t>(z, b + per);
    if (threadIdx.x == 0) base = bc[blockIdx.x];
    __syncthreads();
    for (uint64_t c0 = b; c0 < e; c0 += blockDim.x) {
        const uint64_t i = c0 + threadIdx.x;
        const uint32_t t = (i < e && (flags[i] & kTouched)) ? 1u : 0u;
        s[threadIdx.x] = t;
        __syncthreads();
        // inclusive scan in LDS (Hillis-Steele)
        for (int o = 1; o < 256; o <<= 1) {
            const uint32_t add = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (t) {
            const uint32_t at = base + s[threadIdx.x] - 1;
            ids[at] = (uint32_t)i;
            op[at] = p[i];
            orr[at] = (double)r[i] / kUnit;
        }
        __syncthreads();
        if (threadIdx.x == 0) base += s[255];
        __syncthreads();
    }
}

}  // namespace

void apr_create_impl(uint64_t z, const uint64_t *ptr, const int32_t *dst, const double *w,
                     int32_t device_ptrs, nkpr_apr **out) {
    NKPR_CHECK(out != nullptr, "out is NULL");
    NKPR_CHECK(z > 0 && z < 0x7FFFFFFFull, "the graph needs 1 .. 2^31 - 2 slots");
    NKPR_CHECK(ptr != nullptr, "ptr is NULL");
    std::unique_ptr<nkpr_apr> a(new nkpr_apr());
    NKPR_HIP(hipGetDevice(&a->device));
    NKPR_HIP(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
    const hipStream_t s = a->stream;
    const hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    uint64_t m = 0;
    if (device_ptrs) read_back(s, {{&m, ptr + z, 8}});
    else m = ptr[z];
    NKPR_CHECK(m == 0 || dst != nullptr, "dst is NULL");
    a->z = z;
    a->m = m;
    a->ptr.alloc(8 * (z + 1));
    a->dst.alloc(4 * std::max<uint64_t>(m, 1));
    NKPR_HIP(hipMemcpyAsync(a->ptr.p, ptr, 8 * (z + 1), kind, s));
    if (m) NKPR_HIP(hipMemcpyAsync(a->dst.p, dst, 4 * m, kind, s));
    if (w && m) {
        a->w.alloc(8 * m);
        NKPR_HIP(hipMemcpyAsync(a->w.p, w, 8 * m, kind, s));
    }
    a->vol.alloc(8 * z);
    a->p.alloc(8 * z);
    a->r.alloc(8 * z);
    a->flags.alloc(4 * z);
    a->f0.alloc(4 * z);
    a->f1.alloc(4 * z);
    a->taken.alloc(8 * z);
    // chunks of the nodes with more than kPushChunk out-edges: at most 2 deg / kPushChunk per node
    a->chunk_cap = 2 * m / kPushChunk + 64;
    a->chunks.alloc(2 * 8 * a->chunk_cap);
    a->cnt.alloc(64);
    a->out_cnt.alloc(4 * 1025);
    a->out_ids.alloc(4 * z);
    a->out_p.alloc(8 * z);
    a->out_r.alloc(8 * z);
    NKPR_HIP(hipMemsetAsync(a->cnt.p, 0, 64, s));
    hipLaunchKernelGGL(apr_check_kernel, dim3((unsigned)std::min<uint64_t>((z + 255) / 256, 65536)),
                       dim3(256), 0, s, a->ptr.as<uint64_t>(), z, a->cnt.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    uint32_t bad = 0;
    read_back(s, {{&bad, a->cnt.p, 4}});
    NKPR_CHECK(bad == 0, "offsets must rise from 0 to the number of entries");
    hipLaunchKernelGGL(apr_vol_kernel, dim3((unsigned)std::min<uint64_t>((z * 64 + 255) / 256, 65536)),
                       dim3(256), 0, s, a->ptr.as<uint64_t>(), a->dst.as<int32_t>(),
                       a->w.as<double>(), z, a->vol.as<double>(), a->cnt.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    // the host arrays may be released when this returns
    read_back(s, {{&bad, a->cnt.p, 4}});
    NKPR_CHECK(bad == 0, "neighbour ids must lie in [0, z) and weights must be non-negative");
    *out = a.release();
}

void apr_destroy_impl(nkpr_apr *a) { delete a; }

int apr_device(const nkpr_apr *a) { return a->device; }

void apr_run_impl(nkpr_apr *a, double alpha, double eps, const uint64_t *seeds, uint64_t nseeds,
                  uint64_t *count, int64_t *ids_out, double *p_out, double *r_out,
                  nkpr_apr_stats *st) {
    const auto t0 = std::chrono::steady_clock::now();
    NKPR_CHECK(a != nullptr, "handle is NULL");
    NKPR_CHECK(seeds != nullptr && nseeds > 0, "at least one seed is needed");
    NKPR_CHECK(count != nullptr, "count is NULL");
    NKPR_CHECK(alpha > 0.0 && alpha <= 1.0, "alpha must lie in (0, 1]");
    NKPR_CHECK(eps > 0.0, "epsilon must be > 0");
    // std::set<node>: ids ascending, no repeats
    std::vector<uint32_t> sd;
    sd.reserve(nseeds);
    for (uint64_t i = 0; i < nseeds; ++i) {
        NKPR_CHECK(seeds[i] < a->z, "seed id out of range");
        sd.push_back((uint32_t)seeds[i]);
    }
    std::sort(sd.begin(), sd.end());
    sd.erase(std::unique(sd.begin(), sd.end()), sd.end());
    const hipStream_t s = a->stream;
    const uint64_t z = a->z;
    a->seeds.alloc(4 * sd.size());
    NKPR_HIP(hipMemcpyAsync(a->seeds.p, sd.data(), 4 * sd.size(), hipMemcpyHostToDevice, s));
    if (a->dirty) {
        const unsigned gz = (unsigned)std::min<uint64_t>((z + 255) / 256, 65536);
        hipLaunchKernelGGL(apr_clear_kernel, dim3(gz), dim3(256), 0, s, a->p.as<double>(),
                           a->r.as<uint64_t>(), a->flags.as<uint32_t>(), z);
    }
    a->dirty = true;
    hipLaunchKernelGGL(apr_seed_kernel, dim3(1), dim3(256), 0, s, a->seeds.as<uint32_t>(),
                       (uint32_t)sd.size(), a->r.as<uint64_t>(), a->flags.as<uint32_t>(),
                       a->f0.as<uint32_t>(), a->cnt.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    // Rounds run in batches without a host read; a round whose frontier is empty returns at
    // once, and the host reads the counters after each batch.
    uint32_t *front[2] = {a->f0.as<uint32_t>(), a->f1.as<uint32_t>()};
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, a->device));
    const unsigned grid = (unsigned)std::max(cus, 1) * 8, tgrid = 128;
    uint32_t host_cnt[16] = {0};
    uint64_t k_round = 0;
    // rounds per host read of the counters (NKPR_APR_BATCH overrides)
    int batch = kBatch;
    if (const char *e = std::getenv("NKPR_APR_BATCH")) batch = std::max(1, std::min(256, std::atoi(e)));
    for (;;) {
        for (int b = 0; b < batch; ++b, ++k_round) {
            const int k2 = (int)(k_round & 1), k3 = (int)(k_round % 3);
            hipLaunchKernelGGL(apr_take_kernel, dim3(tgrid), dim3(256), 0, s,
                               a->ptr.as<uint64_t>(), a->vol.as<double>(), alpha, eps,
                               front[k2], front[k2 ^ 1], a->cnt.as<uint32_t>(), k3, k2,
                               a->p.as<double>(), a->r.as<uint64_t>(), a->flags.as<uint32_t>(),
                               a->taken.as<uint64_t>(), a->chunks.as<uint2>() + (k2 ? a->chunk_cap : 0));
            hipLaunchKernelGGL(apr_push_kernel, dim3(grid), dim3(kPushThreads), 0, s,
                               a->ptr.as<uint64_t>(), a->dst.as<int32_t>(), a->w.as<double>(),
                               a->vol.as<double>(), alpha, eps, a->r.as<uint64_t>(),
                               a->flags.as<uint32_t>(), front[k2], a->taken.as<uint64_t>(),
                               a->chunks.as<uint2>() + (k2 ? a->chunk_cap : 0), front[k2 ^ 1],
                               a->cnt.as<uint32_t>(), k3, k2);
        }
        NKPR_HIP(hipGetLastError());
        read_back(s, {{host_cnt, a->cnt.p, sizeof(host_cnt)}});
        if (host_cnt[k_round % 3] == 0) break;  // the frontier of the next round is empty
    }
    uint64_t rounds = host_cnt[kCntRounds], pushes = 0;
    std::memcpy(&pushes, host_cnt + kCntPushes, 8);
    // the touched nodes, ascending
    const unsigned nbk = 1024;
    hipLaunchKernelGGL(apr_count_kernel, dim3(nbk), dim3(256), 0, s, a->flags.as<uint32_t>(), z,
                       a->out_cnt.as<uint32_t>());
    hipLaunchKernelGGL(apr_scan_kernel, dim3(1), dim3(nbk), 0, s, a->out_cnt.as<uint32_t>(), nbk,
                       a->out_cnt.as<uint32_t>() + nbk);
    hipLaunchKernelGGL(apr_compact_kernel, dim3(nbk), dim3(256), 0, s, a->flags.as<uint32_t>(),
                       a->p.as<double>(), a->r.as<uint64_t>(), z, a->out_cnt.as<uint32_t>(),
                       a->out_ids.as<uint32_t>(), a->out_p.as<double>(), a->out_r.as<double>());
    hipLaunchKernelGGL(apr_reset_kernel, dim3(1024), dim3(256), 0, s, a->out_ids.as<uint32_t>(),
                       a->out_cnt.as<uint32_t>() + nbk, a->p.as<double>(), a->r.as<uint64_t>(),
                       a->flags.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    uint32_t k = 0;
    read_back(s, {{&k, a->out_cnt.as<uint32_t>() + nbk, 4}});
    a->dirty = false;
    *count = k;
    if (ids_out || p_out || r_out) {
        std::vector<uint32_t> ids(k);
        read_back(s, {{ids.data(), a->out_ids.p, ids_out ? 4 * (size_t)k : 0},
                      {p_out, a->out_p.p, p_out ? 8 * (size_t)k : 0},
                      {r_out, a->out_r.p, r_out ? 8 * (size_t)k : 0}});
        if (ids_out)
            for (uint32_t i = 0; i < k; ++i) ids_out[i] = (int64_t)ids[i];
    }
    if (st) {
        uint64_t ep = 0;
        std::memcpy(&ep, host_cnt + kCntEdges, 8);
        st->edge_pushes = ep;
        st->rounds = rounds;
        st->pushes = pushes;
        st->touched = k;
        st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

}  // namespace nkpr
