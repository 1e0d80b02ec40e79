// Graph upload and preprocessing: the device-side half of "export the Graph to CSC".
//
// Input: the in-lists NetworKit's PageRank iterates (Graph::forInEdgesOf,
// include/networkit/graph/Graph.hpp:1918-1938) as CSC, optional weights, optional exists[]
// and optional weighted out-degrees (Graph::weightedDegree, networkit/cpp/graph/Graph.cpp:
// 427-455,539).  Output layout in HBM (compact ids 0..n-1 over existing nodes):
//   - nodes relabelled by out-degree, descending (ties by slot id): the sources gathered most
//     often form a dense prefix of the gathered vector (cache reuse in L2 / Infinity Cache)
//     and the nodes without out-edges form the tail [sink_begin, n), so the sink mass is a
//     range sum and needs no per-node flag;
//   - colptr (uint64) over the relabelled rows, every in-list sorted by the new source id;
//   - weighted graphs: folded coefficients w_vu / deg_v in FP32 and FP64 (one stream per
//     edge instead of a weight stream plus a degree gather);
//   - unweighted graphs: 1/deg per node in FP32 and FP64 (the sweep prescales x instead);
//   - the sweep schedule (work items, see nkpr_internal.hpp).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "nkpr_internal.hpp"

namespace nkpr {
namespace {

inline unsigned blocks_for(uint64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// validation flags
constexpr uint32_t kBadSrc = 1, kBadPtr = 2, kBadExists = 4;

__global__ void validate_ptr_kernel(const uint64_t *colptr, uint64_t z, uint64_t m,
                                    const uint8_t *exists, uint32_t *flags) {
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= z) return;
    const uint64_t b = colptr[u], e = colptr[u + 1];
    if (e < b || e > m) atomicOr(flags, kBadPtr);
    if (exists && !exists[u] && e != b) atomicOr(flags, kBadExists);
}

__global__ void out_degree_kernel(const int32_t *src, uint64_t m, uint64_t z,
                                  const uint8_t *exists, const float *w, const double *w64,
                                  uint32_t *outcnt, double *wdeg, uint32_t *flags) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t v = src[e];
        if (v < 0 || (uint64_t)v >= z || (exists && !exists[v])) {
            atomicOr(flags, kBadSrc);
            continue;
        }
        atomicAdd(&outcnt[v], 1u);
        if (wdeg) atomicAdd(&wdeg[v], w64 ? w64[e] : (w ? (double)w[e] : 1.0));
    }
}

__global__ void relabel_keys_kernel(const uint32_t *outcnt, const uint8_t *exists, uint64_t z,
                                    uint32_t *keys, int32_t *vals) {
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= z) return;
    const bool ex = !exists || exists[u];
    // ascending sort of keys == descending out-degree; deleted slots last
    keys[u] = ex ? (0xFFFFFFFEu - outcnt[u]) : 0xFFFFFFFFu;
    vals[u] = (int32_t)u;
}

__global__ void count_kernel(const uint32_t *outcnt, const uint8_t *exists, uint64_t z,
                             unsigned long long *n_exist, unsigned long long *n_nonsink) {
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= z) return;
    const bool ex = !exists || exists[u];
    if (ex) {
        atomicAdd(n_exist, 1ull);
        if (outcnt[u] > 0) atomicAdd(n_nonsink, 1ull);
    }
}

__global__ void iperm_kernel(const int32_t *perm, uint32_t n, int32_t *iperm) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) iperm[perm[r]] = (int32_t)r;
}

__global__ void new_lengths_kernel(const int32_t *perm, uint32_t n, const uint64_t *colptr_in,
                                   uint64_t *len) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) {
        const int32_t o = perm[r];
        len[r] = colptr_in[o + 1] - colptr_in[o];
    } else if (r == n) {
        len[r] = 0;
    }
}

// One wave64 per old row: key = (new row << 32) | new source, value = old edge index.
__global__ void edge_keys_kernel(const uint64_t *colptr_in, const int32_t *src_in,
                                 const int32_t *iperm, uint64_t z, uint64_t *keys,
                                 uint32_t *vals) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; wave < z;
         wave += nwaves) {
        const int32_t r = iperm[wave];
        const uint64_t b = colptr_in[wave], e = colptr_in[wave + 1];
        for (uint64_t i = b + lane; i < e; i += 64) {
            keys[i] = ((uint64_t)(uint32_t)r << 32) | (uint32_t)iperm[src_in[i]];
            vals[i] = (uint32_t)i;
        }
    }
}

__global__ void edge_output_kernel(const uint64_t *keys, const uint32_t *vals, uint64_t m,
                                   const float *w_in, const double *w64_in, const int32_t *src_in,
                                   const double *wdeg, int32_t *src_out, float *coef32,
                                   double *coef64) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x) {
        src_out[i] = (int32_t)(uint32_t)(keys[i] & 0xFFFFFFFFull);
        if (coef32) {
            const uint32_t o = vals[i];
            const double w = w64_in ? w64_in[o] : (double)w_in[o];
            const double c = w / wdeg[src_in[o]];
            coef64[i] = c;
            coef32[i] = (float)c;
        }
    }
}

__global__ void inv_degree_kernel(const int32_t *perm, uint32_t n, uint32_t sink_begin,
                                  const double *wdeg, double *inv64, float *inv32) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double v = (r < sink_begin) ? 1.0 / wdeg[perm[r]] : 0.0;
    inv64[r] = v;
    inv32[r] = (float)v;
}

void build_schedule(nkpr_graph *g, const std::vector<uint64_t> &cp) {
    std::vector<uint2> split, vec, stream;
    std::vector<uint32_t> srow, sfirst, scount;
    const uint32_t n = g->n;
    uint64_t maxdeg = 0;
    uint32_t r = 0;
    while (r < n) {
        const uint64_t len = cp[r + 1] - cp[r];
        maxdeg = std::max(maxdeg, len);
        if (len > (uint64_t)kCap) {
            if (len > kVectorMax) {
                const uint32_t chunks = (uint32_t)((len + kChunk - 1) / kChunk);
                srow.push_back(r);
                sfirst.push_back((uint32_t)split.size());
                scount.push_back(chunks);
                for (uint32_t c = 0; c < chunks; ++c) split.push_back(make_uint2(r, c | kSplitFlag));
            } else {
                vec.push_back(make_uint2(r, r + 1));
            }
            ++r;
            continue;
        }
        const uint32_t r0 = r;
        uint64_t nnz = 0;
        while (r < n && r - r0 < (uint32_t)kRowsMax) {
            const uint64_t l = cp[r + 1] - cp[r];
            if (l > (uint64_t)kCap || nnz + l > (uint64_t)kCap) break;
            maxdeg = std::max(maxdeg, l);
            nnz += l;
            ++r;
        }
        stream.push_back(make_uint2(r0, r));
    }
    std::vector<uint2> items;
    items.reserve(split.size() + vec.size() + stream.size());
    items.insert(items.end(), split.begin(), split.end());
    items.insert(items.end(), vec.begin(), vec.end());
    items.insert(items.end(), stream.begin(), stream.end());
    NKPR_CHECK(items.size() < 0x7FFFFFFFull, "too many work items");
    g->n_items = (uint32_t)items.size();
    g->n_split_items = (uint32_t)split.size();
    g->n_split_rows = (uint32_t)srow.size();
    g->max_in_degree = maxdeg;
    g->items.alloc(sizeof(uint2) * std::max<size_t>(items.size(), 1));
    if (!items.empty())
        NKPR_HIP(hipMemcpyAsync(g->items.p, items.data(), sizeof(uint2) * items.size(),
                                hipMemcpyHostToDevice, g->stream));
    g->split_rows.alloc(sizeof(uint32_t) * std::max<size_t>(srow.size(), 1));
    g->split_first.alloc(sizeof(uint32_t) * std::max<size_t>(srow.size(), 1));
    g->split_count.alloc(sizeof(uint32_t) * std::max<size_t>(srow.size(), 1));
    if (!srow.empty()) {
        NKPR_HIP(hipMemcpyAsync(g->split_rows.p, srow.data(), 4 * srow.size(),
                                hipMemcpyHostToDevice, g->stream));
        NKPR_HIP(hipMemcpyAsync(g->split_first.p, sfirst.data(), 4 * srow.size(),
                                hipMemcpyHostToDevice, g->stream));
        NKPR_HIP(hipMemcpyAsync(g->split_count.p, scount.data(), 4 * srow.size(),
                                hipMemcpyHostToDevice, g->stream));
    }
    NKPR_HIP(hipStreamSynchronize(g->stream));
}

int bits_for(uint64_t v) {
    int b = 1;
    while (b < 32 && (1ull << b) < v) ++b;
    return b;
}

}  // namespace

void create_graph(const nkpr_csc *c, nkpr_graph *g) {
    const auto t0 = std::chrono::steady_clock::now();
    NKPR_CHECK(c != nullptr, "csc is NULL");
    NKPR_CHECK(c->z < 0x7FFFFFFFull, "node-id bound must be < 2^31");
    NKPR_CHECK(c->colptr != nullptr, "colptr is NULL");
    const uint64_t z = c->z;
    const hipStream_t s = g->stream;
    const bool dev = c->device_ptrs != 0;
    g->z = z;

    // --- inputs on the device -------------------------------------------------------------
    uint64_t m = 0;
    if (dev) NKPR_HIP(hipMemcpy(&m, c->colptr + z, sizeof(uint64_t), hipMemcpyDeviceToHost));
    else m = c->colptr[z];
    NKPR_CHECK(m < 0xFFFFFFFFull, "more than 2^32-1 in-list entries");
    NKPR_CHECK(m == 0 || c->src != nullptr, "src is NULL");
    g->m = m;
    g->weighted = c->weight != nullptr;

    DevBuf colptr_in_b, src_in_b, w_in_b, exists_b, wdeg_b;
    const uint64_t *colptr_in = c->colptr;
    const int32_t *src_in = c->src;
    const float *w_in = c->weight;
    const double *w64_in = c->weight64;
    NKPR_CHECK(c->weight64 == nullptr || c->weight != nullptr, "weight64 given without weight");
    DevBuf w64_in_b;
    const uint8_t *exists = c->exists;
    auto h2d = [&](DevBuf &b, const void *p, size_t bytes) {
        b.alloc(bytes);
        if (bytes) NKPR_HIP(hipMemcpyAsync(b.p, p, bytes, hipMemcpyHostToDevice, s));
        return b.p;
    };
    if (!dev) {
        colptr_in = (const uint64_t *)h2d(colptr_in_b, c->colptr, 8 * (z + 1));
        src_in = (const int32_t *)h2d(src_in_b, c->src, 4 * m);
        if (c->weight) w_in = (const float *)h2d(w_in_b, c->weight, 4 * m);
        if (c->weight64) w64_in = (const double *)h2d(w64_in_b, c->weight64, 8 * m);
        if (c->exists) exists = (const uint8_t *)h2d(exists_b, c->exists, z);
    }
    g->exists_host.clear();
    if (c->exists) {
        g->exists_host.resize(z);
        if (dev) NKPR_HIP(hipMemcpy(g->exists_host.data(), c->exists, z, hipMemcpyDeviceToHost));
        else std::memcpy(g->exists_host.data(), c->exists, z);
    }

    DevBuf flags_b, outcnt_b, wdeg_own_b, counts_b;
    flags_b.alloc(4);
    NKPR_HIP(hipMemsetAsync(flags_b.p, 0, 4, s));
    uint32_t *flags = flags_b.as<uint32_t>();
    if (z)
        hipLaunchKernelGGL(validate_ptr_kernel, dim3(blocks_for(z)), dim3(256), 0, s, colptr_in, z,
                           m, exists, flags);
    outcnt_b.alloc(4 * (z + 1));
    NKPR_HIP(hipMemsetAsync(outcnt_b.p, 0, 4 * (z + 1), s));
    const double *wdeg = nullptr;
    double *wdeg_own = nullptr;
    if (c->out_wdeg) {
        wdeg = dev ? c->out_wdeg : (const double *)h2d(wdeg_b, c->out_wdeg, 8 * z);
    } else {
        wdeg_own_b.alloc(8 * (z + 1));
        NKPR_HIP(hipMemsetAsync(wdeg_own_b.p, 0, 8 * (z + 1), s));
        wdeg_own = wdeg_own_b.as<double>();
        wdeg = wdeg_own;
    }
    if (m) {
        const unsigned gb = (unsigned)std::min<uint64_t>(blocks_for(m), 65536);
        hipLaunchKernelGGL(out_degree_kernel, dim3(gb), dim3(256), 0, s, src_in, m, z, exists, w_in,
                           w64_in, outcnt_b.as<uint32_t>(), wdeg_own, flags);
    }
    NKPR_HIP(hipGetLastError());
    uint32_t hflags = 0;
    NKPR_HIP(hipMemcpyAsync(&hflags, flags, 4, hipMemcpyDeviceToHost, s));
    NKPR_HIP(hipStreamSynchronize(s));
    NKPR_CHECK(!(hflags & kBadPtr), "colptr is not a non-decreasing offset array ending at m");
    NKPR_CHECK(!(hflags & kBadSrc), "an in-list names a source outside [0, z) or a deleted node");
    NKPR_CHECK(!(hflags & kBadExists), "a deleted slot has in-edges");

    // --- relabel: out-degree descending, deleted slots dropped ------------------------------
    counts_b.alloc(16);
    NKPR_HIP(hipMemsetAsync(counts_b.p, 0, 16, s));
    auto *counts = counts_b.as<unsigned long long>();
    if (z)
        hipLaunchKernelGGL(count_kernel, dim3(blocks_for(z)), dim3(256), 0, s,
                           outcnt_b.as<uint32_t>(), exists, z, counts, counts + 1);
    unsigned long long hc[2] = {0, 0};
    NKPR_HIP(hipMemcpyAsync(hc, counts, 16, hipMemcpyDeviceToHost, s));
    NKPR_HIP(hipStreamSynchronize(s));
    const uint32_t n = (uint32_t)hc[0];
    g->n = n;
    g->sink_begin = (uint32_t)hc[1];

    DevBuf k1, k2, v1, v2, tmp;
    g->perm.alloc(4 * std::max<uint64_t>(n, 1));
    if (z) {
        k1.alloc(4 * z);
        k2.alloc(4 * z);
        v1.alloc(4 * z);
        v2.alloc(4 * z);
        hipLaunchKernelGGL(relabel_keys_kernel, dim3(blocks_for(z)), dim3(256), 0, s,
                           outcnt_b.as<uint32_t>(), exists, z, k1.as<uint32_t>(), v1.as<int32_t>());
        hipcub::DoubleBuffer<uint32_t> kb(k1.as<uint32_t>(), k2.as<uint32_t>());
        hipcub::DoubleBuffer<int32_t> vb(v1.as<int32_t>(), v2.as<int32_t>());
        size_t tb = 0;
        NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb, vb, (size_t)z, 0, 32, s));
        tmp.alloc(tb);
        NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kb, vb, (size_t)z, 0, 32, s));
        if (n)
            NKPR_HIP(hipMemcpyAsync(g->perm.p, vb.Current(), 4 * (size_t)n,
                                    hipMemcpyDeviceToDevice, s));
    }
    // iperm (slot -> compact id), reuse k1 storage
    DevBuf iperm_b;
    iperm_b.alloc(4 * std::max<uint64_t>(z, 1));
    int32_t *iperm = iperm_b.as<int32_t>();
    NKPR_HIP(hipMemsetAsync(iperm, 0xFF, 4 * std::max<uint64_t>(z, 1), s));
    if (n)
        hipLaunchKernelGGL(iperm_kernel, dim3(blocks_for(n)), dim3(256), 0, s, g->perm.as<int32_t>(),
                           n, iperm);
    NKPR_HIP(hipStreamSynchronize(s));  // the sort buffers are freed below
    k1.reset(); k2.reset(); v1.reset(); v2.reset(); tmp.reset();

    // --- new colptr -------------------------------------------------------------------------
    DevBuf len_b;
    len_b.alloc(8 * ((uint64_t)n + 1));
    hipLaunchKernelGGL(new_lengths_kernel, dim3(blocks_for((uint64_t)n + 1)), dim3(256), 0, s,
                       g->perm.as<int32_t>(), n, colptr_in, len_b.as<uint64_t>());
    g->colptr.alloc(8 * ((uint64_t)n + 1));
    {
        size_t tb = 0;
        NKPR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, len_b.as<uint64_t>(),
                                                  g->colptr.as<uint64_t>(), (size_t)n + 1, s));
        tmp.alloc(tb);
        NKPR_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, len_b.as<uint64_t>(),
                                                  g->colptr.as<uint64_t>(), (size_t)n + 1, s));
        NKPR_HIP(hipStreamSynchronize(s));
        tmp.reset();
    }
    len_b.reset();

    // --- edges: sort by (new row, new source) ------------------------------------------------
    g->src.alloc(4 * std::max<uint64_t>(m, 1));
    if (g->weighted) {
        g->coef32.alloc(4 * std::max<uint64_t>(m, 1));
        g->coef64.alloc(8 * std::max<uint64_t>(m, 1));
    }
    if (m) {
        DevBuf ka, kb2, va, vb2;
        ka.alloc(8 * m);
        kb2.alloc(8 * m);
        va.alloc(4 * m);
        vb2.alloc(4 * m);
        const unsigned kblocks = (unsigned)std::min<uint64_t>(blocks_for(z * 64), 1u << 18);
        hipLaunchKernelGGL(edge_keys_kernel, dim3(kblocks), dim3(256), 0, s, colptr_in,
                           src_in, iperm, z, ka.as<uint64_t>(), va.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        hipcub::DoubleBuffer<uint64_t> kd(ka.as<uint64_t>(), kb2.as<uint64_t>());
        hipcub::DoubleBuffer<uint32_t> vd(va.as<uint32_t>(), vb2.as<uint32_t>());
        const int end_bit = 32 + bits_for(n);
        size_t tb = 0;
        NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kd, vd, (size_t)m, 0, end_bit, s));
        tmp.alloc(tb);
        NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kd, vd, (size_t)m, 0, end_bit, s));
        NKPR_HIP(hipStreamSynchronize(s));
        tmp.reset();
        const unsigned gb = (unsigned)std::min<uint64_t>(blocks_for(m), 65536);
        hipLaunchKernelGGL(edge_output_kernel, dim3(gb), dim3(256), 0, s, kd.Current(), vd.Current(),
                           m, g->weighted ? w_in : nullptr, g->weighted ? w64_in : nullptr, src_in,
                           wdeg, g->src.as<int32_t>(),
                           g->weighted ? g->coef32.as<float>() : nullptr,
                           g->weighted ? g->coef64.as<double>() : nullptr);
        NKPR_HIP(hipGetLastError());
        NKPR_HIP(hipStreamSynchronize(s));
    }
    if (!g->weighted) {
        g->inv32.alloc(4 * std::max<uint64_t>(n, 1));
        g->inv64.alloc(8 * std::max<uint64_t>(n, 1));
        if (n)
            hipLaunchKernelGGL(inv_degree_kernel, dim3(blocks_for(n)), dim3(256), 0, s,
                               g->perm.as<int32_t>(), n, g->sink_begin, wdeg, g->inv64.as<double>(),
                               g->inv32.as<float>());
        NKPR_HIP(hipGetLastError());
    }

    // --- schedule ---------------------------------------------------------------------------
    std::vector<uint64_t> cp((size_t)n + 1);
    NKPR_HIP(hipMemcpyAsync(cp.data(), g->colptr.p, 8 * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    NKPR_HIP(hipStreamSynchronize(s));
    build_schedule(g, cp);
    NKPR_HIP(hipStreamSynchronize(s));
    g->prep_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace nkpr
