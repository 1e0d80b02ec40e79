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
#include <cstring>
#include <vector>

#include "nkpr_internal.hpp"

struct nkpr_apr {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t z = 0, m = 0;
    nkpr::DevBuf ptr, dst, w, vol;
    nkpr::DevBuf p, r, flags, f0, f1, taken, cnt, out_ids, out_p, out_r, out_cnt, seeds;
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
constexpr double kUnit = 4611686018427387904.0;  // 2^62 fixed-point units per unit of mass

// ceil(eps vol 2^62): r < thr  <=>  r 2^-62 < eps vol (the product rounded as the reference's).
// A node without neighbours (vol 0) is pushed again while any residual is left: the reference's
// r / vol >= eps holds for every r > 0 there.
__device__ __forceinline__ uint64_t threshold(double eps, double vol) {
    const double t = eps * vol * kUnit;
    if (!(t > 0.0)) return 1ull;
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

// the seeds: residual 1/|seeds|, touched, queued, in the first frontier (ids ascending)
__global__ void apr_seed_kernel(const uint32_t *seeds, uint32_t k, uint64_t *r, uint32_t *flags,
                                uint32_t *front, uint32_t *cnt) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        const uint32_t s = seeds[i];
        r[s] = (uint64_t)(kUnit / (double)k);
        flags[s] = kTouched | kQueued;
        front[i] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt[0] = k;
        cnt[1] = 0u;
        cnt[2] = 0u;  // cnt[2..3]: out-edges of the pushed nodes (64-bit)
        cnt[3] = 0u;
    }
}

__device__ __forceinline__ void enqueue(uint32_t v, uint32_t *flags, uint32_t *next,
                                        uint32_t *ncnt) {
    const uint32_t old = atomicOr(&flags[v], kQueued | kTouched);
    if (!(old & kQueued)) next[atomicAdd(ncnt, 1u)] = v;
}

// Before a round: every frontier node's residual is taken (and its queued bit cleared), so
// that the round pushes the residuals of its start.
__global__ void apr_take_kernel(const uint32_t *front, const uint32_t *cnt, uint64_t *r,
                                uint32_t *flags, uint64_t *taken) {
    const uint32_t nf = cnt[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x) {
        const uint32_t u = front[i];
        taken[i] = r[u];
        r[u] = 0ull;
        flags[u] &= ~kQueued;
    }
}

// One round: a wave per frontier node.  cnt[0] = this frontier's size, cnt[1] the next one's,
// cnt[2..3] the out-edges pushed so far (8-byte aligned: cnt is a 64-byte buffer).  At the
// start of a round every residual outside the frontier lies below its threshold and every
// frontier residual is 0, and all adds are positive, so the add that crosses a threshold is
// the one that finds the residual below it and leaves it at or above it.
__global__ __launch_bounds__(kPushThreads) void apr_push_kernel(
    const uint64_t *ptr, const int32_t *dst, const double *w, const double *vol, double alpha,
    double eps, double *p, uint64_t *r, uint32_t *flags, const uint32_t *front,
    const uint64_t *taken, uint32_t *next, uint32_t *cnt) {
    const uint32_t nf = cnt[0];
    const int lane = threadIdx.x & 63;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < nf; i += nw) {
        const uint32_t u = front[i];
        const double res = (double)taken[i];  // in units
        const double vu = vol[u];
        const double share = vu > 0.0 ? (1.0 - alpha) * res / (2.0 * vu) : 0.0;
        for (uint64_t e = ptr[u] + lane; e < ptr[u + 1]; e += 64) {
            const uint32_t v = (uint32_t)dst[e];
            if (v == u) continue;  // the reference overwrites r[u] after the loop: dropped
            const uint64_t mass = __double2ull_rn(share * (w ? w[e] : 1.0));
            const uint64_t thr = threshold(eps, vol[v]);
            const uint64_t old = atomicAdd(reinterpret_cast<unsigned long long *>(&r[v]),
                                           (unsigned long long)mass);
            if (old < thr && old + mass >= thr) enqueue(v, flags, next, cnt + 1);
            else atomicOr(&flags[v], kTouched);
        }
        if (lane == 0) {
            p[u] += alpha * res / kUnit;
            const uint64_t keep = __double2ull_rn((1.0 - alpha) * res / 2);
            const uint64_t thr = threshold(eps, vu);
            const uint64_t old = atomicAdd(reinterpret_cast<unsigned long long *>(&r[u]),
                                           (unsigned long long)keep);
            if (old < thr && old + keep >= thr) enqueue(u, flags, next, cnt + 1);
            atomicAdd(reinterpret_cast<unsigned long long *>(cnt + 2),
                      (unsigned long long)(ptr[u + 1] - ptr[u]));
        }
    }
}

// cnt[0] = cnt[1] (the next frontier becomes this one), cnt[1] = 0
__global__ void apr_swap_kernel(uint32_t *cnt) {
    cnt[0] = cnt[1];
    cnt[1] = 0u;
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

__global__ void apr_scan_kernel(uint32_t *bc, uint32_t nb, uint32_t *total) {
    if (threadIdx.x) return;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < nb; ++i) {
        const uint32_t c = bc[i];
        bc[i] = acc;
        acc += c;
    }
    total[0] = acc;
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
    uint32_t *cur = a->f0.as<uint32_t>(), *nxt = a->f1.as<uint32_t>();
    uint32_t nf = (uint32_t)sd.size();
    uint64_t rounds = 0, pushes = 0;
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, a->device));
    while (nf > 0) {
        const unsigned grid = (unsigned)std::max<uint64_t>(
            1, std::min<uint64_t>(((uint64_t)nf * 64 + kPushThreads - 1) / kPushThreads,
                                  (uint64_t)std::max(cus, 1) * 32));
        const unsigned tgrid = (unsigned)std::min<uint64_t>(((uint64_t)nf + 255) / 256, 4096);
        hipLaunchKernelGGL(apr_take_kernel, dim3(tgrid), dim3(256), 0, s, cur,
                           a->cnt.as<uint32_t>(), a->r.as<uint64_t>(), a->flags.as<uint32_t>(),
                           a->taken.as<uint64_t>());
        hipLaunchKernelGGL(apr_push_kernel, dim3(grid), dim3(kPushThreads), 0, s,
                           a->ptr.as<uint64_t>(), a->dst.as<int32_t>(), a->w.as<double>(),
                           a->vol.as<double>(), alpha, eps, a->p.as<double>(),
                           a->r.as<uint64_t>(), a->flags.as<uint32_t>(), cur,
                           a->taken.as<uint64_t>(), nxt, a->cnt.as<uint32_t>());
        hipLaunchKernelGGL(apr_swap_kernel, dim3(1), dim3(1), 0, s, a->cnt.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        pushes += nf;
        ++rounds;
        read_back(s, {{&nf, a->cnt.p, 4}});
        std::swap(cur, nxt);
    }
    // the touched nodes, ascending
    const unsigned nbk = 1024;
    hipLaunchKernelGGL(apr_count_kernel, dim3(nbk), dim3(256), 0, s, a->flags.as<uint32_t>(), z,
                       a->out_cnt.as<uint32_t>());
    hipLaunchKernelGGL(apr_scan_kernel, dim3(1), dim3(64), 0, s, a->out_cnt.as<uint32_t>(), nbk,
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
        read_back(s, {{&ep, a->cnt.as<uint32_t>() + 2, 8}});
        st->edge_pushes = ep;
        st->rounds = rounds;
        st->pushes = pushes;
        st->touched = k;
        st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

}  // namespace nkpr
