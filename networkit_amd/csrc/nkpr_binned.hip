// Binned application of the FP32 operator y = d P v for the GMRES correction solves of the
// mixed-precision path (the product of networkit/cpp/centrality/PageRank.cpp:56-66 without
// the teleport term).
//
// The pull sweep gathers v[src] once per in-edge: a random 4-byte access that costs a 128-byte
// line when it misses L2, which bounds the sweep by the gather request rate.  Here the product
// is split in two streaming passes over a second copy of the edges, ordered by (source chunk,
// destination bin, destination):
//   pass 1  per source chunk of 32768 ids: v of the chunk is staged in LDS, and every edge of
//           the chunk writes its contribution val[e] = v[src] * w/deg(src) in edge order;
//   pass 2  per destination bin of 32768 ids: the contributions of every (chunk, bin) segment
//           are added into an LDS accumulator, which is written (or, for bins split over
//           several work items, added) to y.
// Every gather is an LDS access and every HBM access is a stream: about 16 bytes per edge
// (source offset 2 + coefficient 4 + contribution 4 + 4, destination offset 2) instead of one
// random line per gather.  The LDS and global float atomics make the summation order, and so
// the last bits of y, vary from run to run; this operator only feeds the FP32 correction
// solves, whose result is certified by the deterministic FP64 residual sweeps.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nkpr_internal.hpp"
#include "nkpr_run.hpp"

namespace nkpr {
namespace {

constexpr uint32_t kSpan = 32768;   // ids per source chunk and per destination bin
constexpr int kSpanShift = 15;
constexpr int kPassThreads = 1024;
constexpr uint64_t kPass1Edges = 1u << 19;  // edges per pass-1 work item
constexpr uint64_t kPass2Edges = 1u << 20;  // edges per pass-2 work item (at least)
// A (chunk, bin) segment this long is cut into edge ranges that whole blocks stream; shorter
// segments are handled one wave per segment.
constexpr uint32_t kBigSegment = 1u << 16;
constexpr uint32_t kItemSole = 1, kItemRange = 2;

inline unsigned nb(uint64_t n, int t = 256) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + t - 1) / t, 65536));
}

__global__ void bin_row_kernel(const uint32_t *rowptr, uint32_t n, uint32_t *row) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < n; w += nw)
        for (uint32_t e = rowptr[w] + lane; e < rowptr[w + 1]; e += 64) row[e] = (uint32_t)w;
}

// key = ((chunk * nbins + bin) << 15) | destination offset: inside a segment the edges are
// sorted by destination, so equal destinations form runs that pass 2 adds up in registers
__global__ void bin_key_kernel(const int32_t *src, const uint32_t *row, uint64_t m, uint32_t nbins,
                               uint64_t *key, uint32_t *val) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)src[e];
        const uint64_t seg = (uint64_t)(v >> kSpanShift) * nbins + (row[e] >> kSpanShift);
        key[e] = (seg << kSpanShift) | (row[e] & (kSpan - 1));
        val[e] = (uint32_t)e;
    }
}

// seg[s] = first sorted edge of segment s (binary search over the sorted keys)
__global__ void bin_seg_kernel(const uint64_t *key, uint64_t m, uint64_t nseg, uint32_t *seg) {
    for (uint64_t sidx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; sidx <= nseg;
         sidx += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = sidx << kSpanShift;
        uint64_t lo = 0, hi = m;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (key[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        seg[sidx] = (uint32_t)lo;
    }
}

// Adds x into acc[d] for one window of 64 consecutive (destination-sorted) edges of a wave:
// a segmented inclusive scan over equal destinations, then one LDS atomic per run.
__device__ __forceinline__ void wave_run_add(float *acc, uint32_t d, float x, bool valid,
                                             int lane) {
    if (!valid) d = 0x10000u + (uint32_t)lane;  // never equal to a real offset or another lane
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const float xo = __shfl_up(x, off, 64);
        const uint32_t dq = __shfl_up(d, off, 64);
        if (lane >= off && dq == d) x += xo;
    }
    const uint32_t dn = __shfl_down(d, 1, 64);
    if (valid && (lane == 63 || dn != d)) atomicAdd(&acc[d], x);
}

__global__ void bin_emit_kernel(const uint32_t *order, uint64_t m, const int32_t *src,
                                const uint32_t *row, const float *coef, uint16_t *bsrc,
                                uint16_t *bdst, float *bcoef) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = order[j];
        bsrc[j] = (uint16_t)((uint32_t)src[e] & (kSpan - 1));
        bdst[j] = (uint16_t)(row[e] & (kSpan - 1));
        if (bcoef) bcoef[j] = coef[e];
    }
}

// Pass 1: item = {chunk, first edge, end edge}.  val[e] = v[src] * coef[e] (weighted) or
// v[src] / deg(src) (unweighted: inv).
__global__ __launch_bounds__(kPassThreads) void bin_scatter_kernel(
    const uint4 *items, uint32_t nitems, const float *v, const float *inv, uint32_t n,
    const uint16_t *bsrc, const float *bcoef, float *bval) {
    __shared__ float tab[kSpan];
    for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const uint4 item = items[it];
        const uint32_t base = item.x << kSpanShift;
        for (uint32_t i = threadIdx.x; i < kSpan; i += kPassThreads) {
            const uint32_t u = base + i;
            float x = 0.f;
            if (u < n) x = inv ? v[u] * inv[u] : v[u];
            tab[i] = x;
        }
        __syncthreads();
        const uint32_t e1 = item.z;
        uint32_t e = item.y + threadIdx.x;
        for (; e + 3 * kPassThreads < e1; e += 4 * kPassThreads) {
            uint16_t s[4];
            float c[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                s[k] = __builtin_nontemporal_load(bsrc + e + k * kPassThreads);
                c[k] = bcoef ? __builtin_nontemporal_load(bcoef + e + k * kPassThreads) : 1.f;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                __builtin_nontemporal_store(tab[s[k]] * c[k], bval + e + k * kPassThreads);
        }
        for (; e < e1; e += kPassThreads) {
            const float c = bcoef ? bcoef[e] : 1.f;
            bval[e] = tab[bsrc[e]] * c;
        }
        __syncthreads();
    }
}

// Pass 2: item = {bin, first chunk, end chunk, flags}: one wave per (chunk, bin) segment adds
// the segment's contributions into the LDS accumulator of the bin (segments of kBigSegment
// edges or more are skipped: they come as range items); or {bin, first edge, end edge,
// kItemRange}: the whole block streams one piece of a long segment.
__global__ __launch_bounds__(kPassThreads) void bin_gather_kernel(
    const uint4 *items, uint32_t nitems, const uint32_t *seg, uint32_t nbins,
    const uint16_t *bdst, const float *bval, float *y, uint32_t n, uint32_t big) {
    __shared__ float acc[kSpan];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kWaves = kPassThreads / 64;
    for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const uint4 item = items[it];
        for (uint32_t i = threadIdx.x; i < kSpan; i += kPassThreads) acc[i] = 0.f;
        __syncthreads();
        if (item.w & kItemRange) {
            // windows of 64 consecutive edges per wave; every lane of a wave takes part in
            // the window scan, so the loop bound is uniform over the wave
            const uint32_t wbase = item.y + (uint32_t)wave * 64;
            for (uint32_t w0 = wbase; w0 < item.z; w0 += 4 * kPassThreads) {
                uint32_t d[4];
                float x[4];
                bool ok[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t e = w0 + k * kPassThreads + lane;
                    ok[k] = e < item.z;
                    d[k] = ok[k] ? (uint32_t)__builtin_nontemporal_load(bdst + e) : 0u;
                    x[k] = ok[k] ? __builtin_nontemporal_load(bval + e) : 0.f;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) wave_run_add(acc, d[k], x[k], ok[k], lane);
            }
        }
        for (uint32_t c = item.y + wave; !(item.w & kItemRange) && c < item.z; c += kWaves) {
            const uint64_t sidx = (uint64_t)c * nbins + item.x;
            const uint32_t s0 = seg[sidx], s1 = seg[sidx + 1];
            if (s1 - s0 >= big) continue;
            for (uint32_t w0 = s0; w0 < s1; w0 += 256) {
                uint32_t d[4];
                float x[4];
                bool ok[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t e = w0 + 64 * k + lane;
                    ok[k] = e < s1;
                    d[k] = ok[k] ? (uint32_t)__builtin_nontemporal_load(bdst + e) : 0u;
                    x[k] = ok[k] ? __builtin_nontemporal_load(bval + e) : 0.f;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) wave_run_add(acc, d[k], x[k], ok[k], lane);
            }
        }
        __syncthreads();
        const uint32_t base = item.x << kSpanShift;
        for (uint32_t i = threadIdx.x; i < kSpan; i += kPassThreads) {
            const uint32_t u = base + i;
            if (u >= n) break;
            if (item.w & kItemSole) y[u] = acc[i];
            else if (acc[i] != 0.f) atomicAdd(&y[u], acc[i]);
        }
        __syncthreads();
    }
}

__global__ void zero_kernel(float *y, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] = 0.f;
}

// w = v - d y - d S / n   (S = sink mass of v when sinks are distributed)
__global__ void bin_apply_kernel(const float *v, const float *y, uint32_t n, float d,
                                 const double *sink_sum, float *w) {
    const float st = sink_sum ? (float)(d * *sink_sum / (double)n) : 0.f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        w[i] = v[i] - (d * y[i] + st);
}

}  // namespace

bool binned_enabled(const nkpr_graph *g) {
    // opt-in (NKPR_BINNED=1): on R-MAT scale 26 the two passes measure 2.9 + 9.8 ms against
    // 8.4 ms for the pull sweep, so the pull sweep stays the default operator
    const char *e = std::getenv("NKPR_BINNED");
    (void)g;
    return e && *e && std::atoi(e) != 0;
}

// Builds the binned edge copy (once per graph; coefficients again after new weights).
void binned_layout(nkpr_graph *g) {
    batch_layout(g);
    if (g->bn_ready && g->bn_coef_ready) return;
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const uint32_t nchunks = std::max<uint32_t>(1, (g->sink_begin + kSpan - 1) / kSpan);
    const uint32_t nbins = (n + kSpan - 1) / kSpan;
    const uint64_t nseg = (uint64_t)nchunks * nbins;
    DevBuf row, k1, k2, v1, v2, tmp;
    row.alloc(4 * std::max<uint64_t>(m, 1));
    if (n)
        hipLaunchKernelGGL(bin_row_kernel, dim3((unsigned)std::min<uint64_t>(((uint64_t)n * 64 + 255) / 256, 1u << 18)), dim3(256), 0, s,
                           g->b_rowptr.as<uint32_t>(), n, row.as<uint32_t>());
    if (g->bn_ready) {
        // new coefficients, same order: re-emit from the kept order
        if (m && g->weighted)
            hipLaunchKernelGGL(bin_emit_kernel, dim3(nb(m)), dim3(256), 0, s, g->bn_order.as<uint32_t>(), m,
                               g->b_src.as<int32_t>(), row.as<uint32_t>(), g->b_coef32.as<float>(),
                               g->bn_src.as<uint16_t>(), g->bn_dst.as<uint16_t>(), g->bn_coef.as<float>());
        NKPR_HIP(hipGetLastError());
        NKPR_HIP(hipStreamSynchronize(s));
        g->bn_coef_ready = true;
        return;
    }
    k1.alloc(8 * std::max<uint64_t>(m, 1));
    k2.alloc(8 * std::max<uint64_t>(m, 1));
    v1.alloc(4 * std::max<uint64_t>(m, 1));
    v2.alloc(4 * std::max<uint64_t>(m, 1));
    g->bn_seg.alloc(4 * (nseg + 1));
    if (!m) NKPR_HIP(hipMemsetAsync(g->bn_seg.p, 0, 4 * (nseg + 1), s));
    if (m) {
        hipLaunchKernelGGL(bin_key_kernel, dim3(nb(m)), dim3(256), 0, s, g->b_src.as<int32_t>(),
                           row.as<uint32_t>(), m, nbins, k1.as<uint64_t>(), v1.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        hipcub::DoubleBuffer<uint64_t> kb(k1.as<uint64_t>(), k2.as<uint64_t>());
        hipcub::DoubleBuffer<uint32_t> vb(v1.as<uint32_t>(), v2.as<uint32_t>());
        int bits = kSpanShift + 1;
        while (bits < 64 && (1ull << (bits - kSpanShift)) < nseg) ++bits;
        size_t tb = 0;
        NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb, vb, (size_t)m, 0, bits, s));
        tmp.alloc(tb);
        NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kb, vb, (size_t)m, 0, bits, s));
        g->bn_order.alloc(4 * m);
        NKPR_HIP(hipMemcpyAsync(g->bn_order.p, vb.Current(), 4 * m, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(bin_seg_kernel, dim3(nb(nseg + 1)), dim3(256), 0, s,
                           (const uint64_t *)kb.Current(), m, nseg, g->bn_seg.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
    }
    g->bn_src.alloc(2 * std::max<uint64_t>(m, 1) + 64);
    g->bn_dst.alloc(2 * std::max<uint64_t>(m, 1) + 64);
    g->bn_val.alloc(4 * std::max<uint64_t>(m, 1) + 64);
    if (g->weighted) g->bn_coef.alloc(4 * std::max<uint64_t>(m, 1) + 64);
    if (m)
        hipLaunchKernelGGL(bin_emit_kernel, dim3(nb(m)), dim3(256), 0, s, g->bn_order.as<uint32_t>(), m,
                           g->b_src.as<int32_t>(), row.as<uint32_t>(),
                           g->weighted ? g->b_coef32.as<float>() : nullptr, g->bn_src.as<uint16_t>(),
                           g->bn_dst.as<uint16_t>(), g->weighted ? g->bn_coef.as<float>() : nullptr);
    NKPR_HIP(hipGetLastError());
    std::vector<uint32_t> seg(nseg + 1);
    NKPR_HIP(hipMemcpyAsync(seg.data(), g->bn_seg.p, 4 * (nseg + 1), hipMemcpyDeviceToHost, s));
    NKPR_HIP(hipStreamSynchronize(s));

    // work items (NKPR_BINNED_ITEM overrides both sizes, e.g. to exercise split bins)
    uint64_t e1max = kPass1Edges, e2max = kPass2Edges;
    if (const char *ev = std::getenv("NKPR_BINNED_ITEM")) {
        const uint64_t v = std::strtoull(ev, nullptr, 10);
        if (v > 0) e1max = e2max = v;
    }
    std::vector<uint4> p1, p2;
    g->bn_big = (uint32_t)std::min<uint64_t>(kBigSegment, e2max);
    for (uint32_t c = 0; c < nchunks; ++c) {
        const uint32_t b0 = seg[(uint64_t)c * nbins], b1 = seg[(uint64_t)(c + 1) * nbins];
        for (uint64_t e = b0; e < b1; e += e1max)
            p1.push_back(make_uint4(c, (uint32_t)e, (uint32_t)std::min<uint64_t>(e + e1max, b1), 0));
    }
    for (uint32_t b = 0; b < nbins; ++b) {
        const size_t first = p2.size();
        uint32_t c0 = 0;
        uint64_t acc = 0;
        for (uint32_t c = 0; c < nchunks; ++c) {
            const uint32_t s0 = seg[(uint64_t)c * nbins + b], s1 = seg[(uint64_t)c * nbins + b + 1];
            const uint64_t len = s1 - s0;
            if (len >= g->bn_big) {
                // a long segment: block-wide pieces, skipped by the chunk-range items
                for (uint64_t e = s0; e < s1; e += e2max)
                    p2.push_back(make_uint4(b, (uint32_t)e, (uint32_t)std::min<uint64_t>(e + e2max, s1),
                                            kItemRange));
                continue;
            }
            if (acc > 0 && acc + len > e2max) {
                p2.push_back(make_uint4(b, c0, c, 0));
                c0 = c;
                acc = 0;
            }
            acc += len;
        }
        p2.push_back(make_uint4(b, c0, nchunks, 0));
        if (p2.size() - first == 1) p2.back().w = kItemSole;  // the bin's only item stores directly
    }
    g->bn_p1.alloc(sizeof(uint4) * std::max<size_t>(p1.size(), 1));
    g->bn_p2.alloc(sizeof(uint4) * std::max<size_t>(p2.size(), 1));
    if (!p1.empty()) NKPR_HIP(hipMemcpyAsync(g->bn_p1.p, p1.data(), sizeof(uint4) * p1.size(), hipMemcpyHostToDevice, s));
    NKPR_HIP(hipMemcpyAsync(g->bn_p2.p, p2.data(), sizeof(uint4) * p2.size(), hipMemcpyHostToDevice, s));
    g->bn_np1 = (uint32_t)p1.size();
    g->bn_np2 = (uint32_t)p2.size();
    g->bn_nbins = nbins;
    g->bn_multi = 0;
    for (const auto &it : p2) g->bn_multi += (it.w & kItemSole) ? 0u : 1u;
    g->bn_y.alloc(4 * std::max<uint32_t>(n, 1));
    NKPR_HIP(hipStreamSynchronize(s));
    if (std::getenv("NKPR_DEBUG"))
        std::fprintf(stderr, "[nkpr] binned layout: %u chunks x %u bins, %u + %u items (%u split)\n",
                     nchunks, nbins, g->bn_np1, g->bn_np2, g->bn_multi);
    g->bn_ready = true;
    g->bn_coef_ready = true;
}

// w = v - d P v (- d S(v)/n): the GMRES operator through the binned passes.
void binned_apply(nkpr_graph *g, const float *v, float *w, double damp, bool distribute,
                  hipEvent_t ev0, hipEvent_t ev1) {
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    const double *sink_sum = nullptr;
    if (distribute && g->sink_begin < n) {
        launch_range_sum<float>(v, g->sink_begin, n, g->red_partials.as<double>(), g->red_out.as<double>(), s);
        sink_sum = g->red_out.as<double>();
    }
    if (ev0) NKPR_HIP(hipEventRecord(ev0, s));
    float *y = g->bn_y.as<float>();
    if (g->bn_multi) hipLaunchKernelGGL(zero_kernel, dim3(nb(n)), dim3(256), 0, s, y, n);
    hipLaunchKernelGGL(bin_scatter_kernel, dim3(std::min<uint32_t>(g->bn_np1, (uint32_t)cus * 4)),
                       dim3(kPassThreads), 0, s, g->bn_p1.as<uint4>(), g->bn_np1, v,
                       g->weighted ? nullptr : g->inv32.as<float>(), n, g->bn_src.as<uint16_t>(),
                       g->weighted ? g->bn_coef.as<float>() : nullptr, g->bn_val.as<float>());
    hipLaunchKernelGGL(bin_gather_kernel, dim3(std::min<uint32_t>(g->bn_np2, (uint32_t)cus * 4)),
                       dim3(kPassThreads), 0, s, g->bn_p2.as<uint4>(), g->bn_np2,
                       g->bn_seg.as<uint32_t>(), g->bn_nbins, g->bn_dst.as<uint16_t>(),
                       g->bn_val.as<float>(), y, n, g->bn_big);
    if (ev1) NKPR_HIP(hipEventRecord(ev1, s));
    hipLaunchKernelGGL(bin_apply_kernel, dim3(nb(n)), dim3(256), 0, s, v, (const float *)y, n,
                       (float)damp, sink_sum, w);
    NKPR_HIP(hipGetLastError());
}

}  // namespace nkpr
