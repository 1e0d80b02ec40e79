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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

#include "nkpr_internal.hpp"

namespace nkpr {
namespace {

inline unsigned blocks_for(uint64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// validation flags
constexpr uint32_t kBadSrc = 1, kBadPtr = 2, kBadExists = 4, kBadGroup = 8;
constexpr uint32_t kNotUnitDegree = 16;  // not an error: some node's given out-degree is not 1

// Sets kNotUnitDegree in *flags when deg[i] != 1 for an existing i < count (exists may be NULL).
__global__ void unit_degree_kernel(const double *deg, uint64_t count, const uint8_t *exists,
                                   uint32_t *flags) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        if ((!exists || exists[i]) && deg[i] != 1.0) atomicOr(flags, kNotUnitDegree);
}

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

// Graph of every slot of a batch: the group whose slot range [off[g], off[g+1]) holds it.
__global__ void group_id_kernel(const uint64_t *off, uint32_t groups, uint64_t z, uint32_t *gid) {
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= z) return;
    uint32_t lo = 0, hi = groups;  // last g with off[g] <= u
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (off[mid] <= u) lo = mid;
        else hi = mid;
    }
    gid[u] = lo;
}

// Batch relabelling: graph id first, then out-degree descending, deleted slots last.
__global__ void group_keys_kernel(const uint32_t *outcnt, const uint8_t *exists,
                                  const uint32_t *gid, uint32_t groups, uint64_t z,
                                  uint64_t *keys, int32_t *vals, uint32_t *cnt_exist,
                                  uint32_t *cnt_nonsink) {
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= z) return;
    const bool ex = !exists || exists[u];
    keys[u] = ex ? (((uint64_t)gid[u] << 32) | (0xFFFFFFFEu - outcnt[u]))
                 : (((uint64_t)groups << 32) | 0xFFFFFFFFu);
    vals[u] = (int32_t)u;
    if (ex) {
        atomicAdd(&cnt_exist[gid[u]], 1u);
        if (outcnt[u] > 0) atomicAdd(&cnt_nonsink[gid[u]], 1u);
    }
}

// Every in-edge must stay inside its graph.
__global__ void group_edges_kernel(const uint64_t *colptr, const int32_t *src, const uint32_t *gid,
                                   uint64_t z, uint32_t *flags) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < z; r += nwaves) {
        const uint32_t g = gid[r];
        for (uint64_t i = colptr[r] + lane; i < colptr[r + 1]; i += 64)
            if (gid[src[i]] != g) atomicOr(flags, kBadGroup);
    }
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

__global__ void inv_degree_kernel(const int32_t *perm, uint32_t n, const uint32_t *outcnt,
                                  const double *wdeg, double *inv64, float *inv32) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double v = (outcnt[perm[r]] > 0) ? 1.0 / wdeg[perm[r]] : 0.0;
    inv64[r] = v;
    if (inv32) inv32[r] = (float)v;
}

// flag = 1 when some FP64 weight is not the FP32 weight given beside it.  Each thread keeps its
// own result over a grid-stride loop; one atomic per block at most (the grid is capped by the
// launcher), so a graph whose weights all differ does not serialise on one address.
__global__ void fp32_exact_kernel(const float *w, const double *w64, uint64_t m, uint32_t *flag) {
    int diff = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x)
        diff |= (double)w[i] != w64[i];
    if (__syncthreads_or(diff) && threadIdx.x == 0) atomicOr(flag, 1u);
}

// invw64[u] = 1 / deg[u] (compact order; 0 for nodes without out-weight)
__global__ void inv_of_kernel(const double *deg, uint32_t n, double *invw64) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) invw64[r] = deg[r] > 0.0 ? 1.0 / deg[r] : 0.0;
}

// Per sliced row j: number of in-edges whose source falls in each slice (slice-major output).
// Slice of an in-edge of sliced row j from source v.  Sources ranked below `rep` (the most
// gathered ones; relabelled ids are out-degree ranks) are cached by every XCD and their edges
// are spread over the slices by row, which keeps the slices balanced and keeps the requests
// for the hottest lines off a single L2 channel; every other source belongs to one slice by
// its 32-id line group.
__device__ __forceinline__ uint32_t slice_of(uint32_t v, uint64_t j, uint32_t rep, uint32_t S) {
    return (v < rep) ? (uint32_t)(j % S) : ((v >> kSliceShift) % S);
}

__global__ void slice_count_kernel(const uint64_t *colptr, const int32_t *src,
                                   const uint32_t *srow, uint32_t n_s, uint32_t S, uint32_t rep,
                                   uint32_t *counts) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t j = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < n_s;
         j += nwaves) {
        const uint32_t r = srow[j];
        const uint64_t b = colptr[r], e = colptr[r + 1];
        uint32_t cnt[kSlices] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint64_t i = b; i < e; i += 64) {
            const bool valid = i + lane < e;
            const uint32_t sl = valid ? slice_of((uint32_t)src[i + lane], j, rep, S) : 0u;
            for (uint32_t q = 0; q < S; ++q) cnt[q] += __popcll(__ballot(valid && sl == q));
        }
        if (lane == 0)
            for (uint32_t q = 0; q < S; ++q) counts[(uint64_t)q * n_s + j] = cnt[q];
    }
}

__global__ void add_base_kernel(uint32_t *p, uint64_t n, uint32_t base) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] += base;
}

// Copy the in-edges of every row to its place in the sweep order: direct rows keep their
// (source-sorted) order; sliced rows are split slice by slice, source order kept within a
// slice (stable: ballot ranks).
__global__ void reorder_edges_kernel(const uint64_t *colptr, const int32_t *src,
                                     const float *c32, const double *c64, const uint32_t *rows,
                                     uint32_t nrows, const uint32_t *dptr, const uint32_t *sptr,
                                     uint32_t n_s, uint32_t S, uint32_t rep, int32_t *src2,
                                     float *c32_2, double *c64_2, const uint32_t *em,
                                     uint32_t *em2) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t j = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < nrows;
         j += nwaves) {
        const uint32_t r = rows[j];
        const uint64_t b = colptr[r], e = colptr[r + 1];
        if (dptr) {
            const uint32_t d = dptr[j];
            for (uint64_t i = b + lane; i < e; i += 64) {
                src2[d + (i - b)] = src[i];
                if (c32) {
                    c32_2[d + (i - b)] = c32[i];
                    c64_2[d + (i - b)] = c64[i];
                }
                if (em) em2[d + (i - b)] = em[i];
            }
            continue;
        }
        uint32_t run[kSlices] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint64_t i = b; i < e; i += 64) {
            const bool valid = i + lane < e;
            const int32_t v = valid ? src[i + lane] : 0;
            const uint32_t sl = valid ? slice_of((uint32_t)v, j, rep, S) : 0u;
            uint32_t dest = 0;
            for (uint32_t q = 0; q < S; ++q) {
                const uint64_t mask = __ballot(valid && sl == q);
                if (valid && sl == q)
                    dest = sptr[(uint64_t)q * n_s + j] + run[q] + (uint32_t)__popcll(mask & below);
                run[q] += (uint32_t)__popcll(mask);
            }
            if (valid) {
                src2[dest] = v;
                if (c32) {
                    c32_2[dest] = c32[i + lane];
                    c64_2[dest] = c64[i + lane];
                }
                if (em) em2[dest] = em[i + lane];
            }
        }
    }
}

struct SplitRec {
    uint32_t first[kSlices];
    uint32_t count[kSlices];
};

// Greedy work items over rows [0, nrows) of a list with absolute in-edge offsets ptr[0..nrows]:
// stream blocks of consecutive rows with at most kCap in-edges and kRowsMax rows; a longer
// row segment becomes a vector item, or split chunks (when allow_split) recorded in `splits`.
void build_items(const uint32_t *ptr, uint32_t nrows, uint32_t slice, std::vector<uint4> &out,
                 uint32_t &split_slots, std::map<uint32_t, SplitRec> *splits,
                 uint64_t &maxlen) {
    std::vector<uint4> split, vec, stream;
    uint32_t r = 0;
    while (r < nrows) {
        const uint64_t len = (uint64_t)ptr[r + 1] - ptr[r];
        maxlen = std::max(maxlen, len);
        if (len > (uint64_t)kCap) {
            if (len > kVectorMax && splits) {
                const uint32_t chunks = (uint32_t)((len + kChunk - 1) / kChunk);
                SplitRec &rec = (*splits)[r];
                rec.first[slice] = split_slots;
                rec.count[slice] = chunks;
                for (uint32_t c = 0; c < chunks; ++c) {
                    const uint64_t b = ptr[r] + (uint64_t)c * kChunk;
                    const uint64_t e = std::min<uint64_t>(b + kChunk, ptr[r + 1]);
                    split.push_back(make_uint4(r, split_slots++ | kSplitFlag, (uint32_t)b,
                                               (uint32_t)(e - b)));
                }
            } else {
                vec.push_back(make_uint4(r, r + 1, ptr[r], (uint32_t)len));
            }
            ++r;
            continue;
        }
        const uint32_t r0 = r;
        uint64_t nnz = 0;
        while (r < nrows && r - r0 < (uint32_t)kRowsMax) {
            const uint64_t l = (uint64_t)ptr[r + 1] - ptr[r];
            if (l > (uint64_t)kCap || nnz + l > (uint64_t)kCap) break;
            maxlen = std::max(maxlen, l);
            nnz += l;
            ++r;
        }
        stream.push_back(make_uint4(r0, r, ptr[r0], (uint32_t)nnz));
    }
    out.insert(out.end(), split.begin(), split.end());
    out.insert(out.end(), vec.begin(), vec.end());
    out.insert(out.end(), stream.begin(), stream.end());
}

template <typename V> void upload(nkpr_graph *g, nkpr::DevBuf &b, const std::vector<V> &h) {
    b.alloc(sizeof(V) * std::max<size_t>(h.size(), 1));
    if (!h.empty())
        NKPR_HIP(hipMemcpyAsync(b.p, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice,
                                g->stream));
}

uint32_t env_u32(const char *name, uint32_t dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    return (uint32_t)std::strtoul(v, nullptr, 10);
}

// Splits rows into direct and sliced ones, reorders the in-edges into sweep order and builds
// the work items.  Input: g->colptr / g->src / g->coef* in (row, source) order.
void build_schedule(nkpr_graph *g, const std::vector<uint64_t> &cp) {
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    // slicing pays off once the gathered vector outgrows one L2; below that everything is
    // direct.  NKPR_SLICES (1 or 8) and NKPR_DIRECT_MAX_DEG override the defaults.
    uint32_t S = (n >= (1u << 20)) ? (uint32_t)kSlices : 1u;
    S = env_u32("NKPR_SLICES", S);
    NKPR_CHECK(S == 1 || S == (uint32_t)kSlices, "NKPR_SLICES must be 1 or 8");
    uint32_t dmax = (S == 1) ? (uint32_t)kCap : kDefaultDirectMaxDeg;
    dmax = std::min<uint32_t>(env_u32("NKPR_DIRECT_MAX_DEG", dmax), (uint32_t)kCap);
    const uint32_t rep = env_u32("NKPR_SLICE_REPLICATE", kDefaultSliceReplicate);
    g->n_slices = S;
    g->direct_max_deg = dmax;

    std::vector<uint32_t> drow, srow, dptr;
    drow.reserve(n);
    uint64_t md = 0, maxdeg = 0;
    for (uint32_t r = 0; r < n; ++r) {
        const uint64_t len = cp[r + 1] - cp[r];
        maxdeg = std::max(maxdeg, len);
        if (len <= dmax) drow.push_back(r);
        else srow.push_back(r);
    }
    g->max_in_degree = maxdeg;
    dptr.resize(drow.size() + 1);
    for (size_t k = 0; k < drow.size(); ++k) {
        dptr[k] = (uint32_t)md;
        md += cp[drow[k] + 1] - cp[drow[k]];
    }
    dptr[drow.size()] = (uint32_t)md;
    g->n_d = (uint32_t)drow.size();
    g->n_s = (uint32_t)srow.size();
    const uint32_t n_s = g->n_s;
    upload(g, g->drow, drow);
    upload(g, g->dptr, dptr);
    upload(g, g->srow, srow);

    // per-slice offsets of the sliced rows: exclusive scan of the slice-major counts
    const uint64_t ns_tot = (uint64_t)S * n_s;
    NKPR_CHECK(ns_tot + 1 < 0xFFFFFFFFull, "too many sliced rows");
    g->sptr.alloc(4 * (ns_tot + 1));
    {
        DevBuf counts, tmp;
        counts.alloc(4 * (ns_tot + 1));
        NKPR_HIP(hipMemsetAsync(counts.p, 0, 4 * (ns_tot + 1), s));
        if (n_s) {
            const unsigned blocks = (unsigned)std::min<uint64_t>(blocks_for((uint64_t)n_s * 64), 1u << 18);
            hipLaunchKernelGGL(slice_count_kernel, dim3(blocks), dim3(256), 0, s,
                               g->colptr.as<uint64_t>(), g->src.as<int32_t>(),
                               g->srow.as<uint32_t>(), n_s, S, rep, counts.as<uint32_t>());
            NKPR_HIP(hipGetLastError());
        }
        size_t tb = 0;
        NKPR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts.as<uint32_t>(),
                                                  g->sptr.as<uint32_t>(), (size_t)ns_tot + 1, s));
        tmp.alloc(tb);
        NKPR_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, counts.as<uint32_t>(),
                                                  g->sptr.as<uint32_t>(), (size_t)ns_tot + 1, s));
        hipLaunchKernelGGL(add_base_kernel, dim3((unsigned)std::min<uint64_t>(blocks_for(ns_tot + 1), 65536)),
                           dim3(256), 0, s, g->sptr.as<uint32_t>(), ns_tot + 1, (uint32_t)md);
        NKPR_HIP(hipGetLastError());
        NKPR_HIP(hipStreamSynchronize(s));
    }

    // edges into sweep order (kCap zero entries of padding after the last one)
    {
        DevBuf src2, c32, c64;
        src2.alloc(4 * (m + kCap));
        NKPR_HIP(hipMemsetAsync(src2.as<int32_t>() + m, 0, 4 * kCap, s));
        if (g->weighted) {
            c32.alloc(4 * (m + kCap));
            c64.alloc(8 * (m + kCap));
            NKPR_HIP(hipMemsetAsync(c32.as<float>() + m, 0, 4 * kCap, s));
            NKPR_HIP(hipMemsetAsync(c64.as<double>() + m, 0, 8 * kCap, s));
        }
        const float *o32 = g->weighted ? g->coef32.as<float>() : nullptr;
        const double *o64 = g->weighted ? g->coef64.as<double>() : nullptr;
        float *n32 = g->weighted ? c32.as<float>() : nullptr;
        double *n64 = g->weighted ? c64.as<double>() : nullptr;
        DevBuf em2;
        const uint32_t *em = g->weighted ? g->emap.as<uint32_t>() : nullptr;
        if (g->weighted) em2.alloc(4 * std::max<uint64_t>(m, 1));
        uint32_t *emo = g->weighted ? em2.as<uint32_t>() : nullptr;
        if (g->n_d) {
            const unsigned blocks = (unsigned)std::min<uint64_t>(blocks_for((uint64_t)g->n_d * 64), 1u << 18);
            hipLaunchKernelGGL(reorder_edges_kernel, dim3(blocks), dim3(256), 0, s,
                               g->colptr.as<uint64_t>(), g->src.as<int32_t>(), o32, o64,
                               g->drow.as<uint32_t>(), g->n_d, g->dptr.as<uint32_t>(), nullptr,
                               0u, S, rep, src2.as<int32_t>(), n32, n64, em, emo);
            NKPR_HIP(hipGetLastError());
        }
        if (n_s) {
            const unsigned blocks = (unsigned)std::min<uint64_t>(blocks_for((uint64_t)n_s * 64), 1u << 18);
            hipLaunchKernelGGL(reorder_edges_kernel, dim3(blocks), dim3(256), 0, s,
                               g->colptr.as<uint64_t>(), g->src.as<int32_t>(), o32, o64,
                               g->srow.as<uint32_t>(), n_s, nullptr, g->sptr.as<uint32_t>(),
                               n_s, S, rep, src2.as<int32_t>(), n32, n64, em, emo);
            NKPR_HIP(hipGetLastError());
        }
        NKPR_HIP(hipStreamSynchronize(s));
        std::swap(g->src.p, src2.p);
        std::swap(g->src.bytes, src2.bytes);
        if (g->weighted) {
            std::swap(g->coef32.p, c32.p);
            std::swap(g->coef32.bytes, c32.bytes);
            std::swap(g->coef64.p, c64.p);
            std::swap(g->coef64.bytes, c64.bytes);
            std::swap(g->emap.p, em2.p);
            std::swap(g->emap.bytes, em2.bytes);
        }
    }

    // work items
    std::vector<uint4> ditems, sitems;
    uint32_t split_slots = 0;
    uint64_t maxlen = 0;
    build_items(dptr.data(), g->n_d, 0, ditems, split_slots, nullptr, maxlen);
    std::map<uint32_t, SplitRec> splits;
    std::vector<uint32_t> hs((size_t)ns_tot + 1);
    read_back(s, {{hs.data(), g->sptr.p, 4 * ((size_t)ns_tot + 1)}});
    for (uint32_t q = 0; q < S; ++q) {
        g->sitem_off[q] = (uint32_t)sitems.size();
        build_items(hs.data() + (size_t)q * n_s, n_s, q, sitems, split_slots, &splits, maxlen);
    }
    for (uint32_t q = S; q <= (uint32_t)kSlices; ++q) g->sitem_off[q] = (uint32_t)sitems.size();
    NKPR_CHECK(sitems.size() + ditems.size() < 0x7FFFFFFFull, "too many work items");
    g->n_ditems = (uint32_t)ditems.size();
    g->n_items = (uint32_t)(ditems.size() + sitems.size());
    g->n_split_items = split_slots;
    upload(g, g->ditems, ditems);
    upload(g, g->sitems, sitems);
    std::vector<uint8_t> flag(std::max<uint32_t>(n_s, 1), 0);
    std::vector<uint32_t> srows, sfirst, scount;
    for (auto &kv : splits) {
        flag[kv.first] = 1;
        srows.push_back(kv.first);
        for (uint32_t q = 0; q < S; ++q) {
            sfirst.push_back(kv.second.first[q]);
            scount.push_back(kv.second.count[q]);
        }
    }
    g->n_split_rows = (uint32_t)srows.size();
    upload(g, g->split_flag, flag);
    upload(g, g->split_rows, srows);
    upload(g, g->split_first, sfirst);
    upload(g, g->split_count, scount);
    g->claim.alloc(4 * kSlices);
    NKPR_HIP(hipMemsetAsync(g->claim.p, 0, 4 * kSlices, s));
    NKPR_HIP(hipStreamSynchronize(s));
    if (env_u32("NKPR_DEBUG", 0)) {
        std::fprintf(stderr, "[nkpr] n=%u direct rows %u (%llu edges, %u items), sliced rows %u, "
                             "slices %u, split rows %u\n", n, g->n_d, (unsigned long long)md,
                     g->n_ditems, n_s, S, g->n_split_rows);
        for (uint32_t q = 0; q < S; ++q)
            std::fprintf(stderr, "[nkpr]   slice %u: %llu edges, %u items\n", q,
                         (unsigned long long)(hs[(size_t)(q + 1) * n_s] - hs[(size_t)q * n_s]),
                         g->sitem_off[q + 1] - g->sitem_off[q]);
    }
}

int bits_for(uint64_t v) {
    int b = 1;
    while (b < 32 && (1ull << b) < v) ++b;
    return b;
}

}  // namespace

void create_graph(const nkpr_csc *c, nkpr_graph *g, const uint64_t *group_off,
                  uint32_t n_groups) {
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
    if (c->out_wdeg && z)
        hipLaunchKernelGGL(unit_degree_kernel, dim3((unsigned)std::min<uint64_t>(blocks_for(z), 65536)),
                           dim3(256), 0, s, wdeg, z, exists, flags);
    NKPR_HIP(hipGetLastError());
    uint32_t hflags = 0;
    read_back(s, {{&hflags, flags, 4}});
    g->unit_degrees = c->out_wdeg != nullptr && !(hflags & kNotUnitDegree);
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
    read_back(s, {{hc, counts, 16}});
    const uint32_t n = (uint32_t)hc[0];
    g->n = n;
    g->sink_begin = (uint32_t)hc[1];

    DevBuf k1, k2, v1, v2, tmp;
    g->perm.alloc(4 * std::max<uint64_t>(n, 1));
    if (group_off) {
        // batch of graphs: relabel inside every graph, graphs stay contiguous in id order
        NKPR_CHECK(n_groups >= 1, "a batch needs at least one graph");
        NKPR_CHECK(group_off[0] == 0 && group_off[n_groups] == z, "graph offsets must span [0, z)");
        for (uint32_t q = 0; q < n_groups; ++q)
            NKPR_CHECK(group_off[q] <= group_off[q + 1], "graph offsets must be non-decreasing");
        DevBuf offb, gidb, ce, cn, fl;
        offb.alloc(8 * ((uint64_t)n_groups + 1));
        NKPR_HIP(hipMemcpyAsync(offb.p, group_off, 8 * ((uint64_t)n_groups + 1), hipMemcpyHostToDevice, s));
        gidb.alloc(4 * std::max<uint64_t>(z, 1));
        ce.alloc(4 * (uint64_t)n_groups);
        cn.alloc(4 * (uint64_t)n_groups);
        fl.alloc(4);
        NKPR_HIP(hipMemsetAsync(ce.p, 0, 4 * (uint64_t)n_groups, s));
        NKPR_HIP(hipMemsetAsync(cn.p, 0, 4 * (uint64_t)n_groups, s));
        NKPR_HIP(hipMemsetAsync(fl.p, 0, 4, s));
        DevBuf k64a, k64b;
        k64a.alloc(8 * std::max<uint64_t>(z, 1));
        k64b.alloc(8 * std::max<uint64_t>(z, 1));
        v1.alloc(4 * std::max<uint64_t>(z, 1));
        v2.alloc(4 * std::max<uint64_t>(z, 1));
        if (z) {
            hipLaunchKernelGGL(group_id_kernel, dim3(blocks_for(z)), dim3(256), 0, s,
                               offb.as<uint64_t>(), n_groups, z, gidb.as<uint32_t>());
            hipLaunchKernelGGL(group_keys_kernel, dim3(blocks_for(z)), dim3(256), 0, s,
                               outcnt_b.as<uint32_t>(), exists, gidb.as<uint32_t>(), n_groups, z,
                               k64a.as<uint64_t>(), v1.as<int32_t>(), ce.as<uint32_t>(), cn.as<uint32_t>());
            const unsigned eb = (unsigned)std::min<uint64_t>(blocks_for(z * 64), 1u << 18);
            hipLaunchKernelGGL(group_edges_kernel, dim3(eb), dim3(256), 0, s, colptr_in, src_in,
                               gidb.as<uint32_t>(), z, fl.as<uint32_t>());
            NKPR_HIP(hipGetLastError());
            hipcub::DoubleBuffer<uint64_t> kb(k64a.as<uint64_t>(), k64b.as<uint64_t>());
            hipcub::DoubleBuffer<int32_t> vb(v1.as<int32_t>(), v2.as<int32_t>());
            const int end_bit = 32 + bits_for((uint64_t)n_groups + 1);
            size_t tb = 0;
            NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb, vb, (size_t)z, 0, end_bit, s));
            tmp.alloc(tb);
            NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kb, vb, (size_t)z, 0, end_bit, s));
            if (n)
                NKPR_HIP(hipMemcpyAsync(g->perm.p, vb.Current(), 4 * (size_t)n,
                                        hipMemcpyDeviceToDevice, s));
        }
        std::vector<uint32_t> he(n_groups), hn(n_groups);
        uint32_t hf = 0;
        read_back(s, {{he.data(), ce.p, 4 * (uint64_t)n_groups},
                      {hn.data(), cn.p, 4 * (uint64_t)n_groups},
                      {&hf, fl.p, 4}});
        NKPR_CHECK(!(hf & kBadGroup), "an edge joins two graphs of the batch");
        g->n_groups = n_groups;
        g->goff.assign((size_t)n_groups + 1, 0);
        g->gsink.assign(n_groups, 0);
        g->gslot.assign(group_off, group_off + n_groups + 1);
        for (uint32_t q = 0; q < n_groups; ++q) {
            g->goff[q + 1] = g->goff[q] + he[q];
            g->gsink[q] = g->goff[q] + hn[q];
        }
        // no global sink range: sinks are the tail of every graph's range
        g->sink_begin = n;
    } else if (z) {
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
        if (g->weighted) {
            // input position of every edge, kept so that new weights can be folded in place
            g->emap.alloc(4 * m);
            NKPR_HIP(hipMemcpyAsync(g->emap.p, vd.Current(), 4 * m, hipMemcpyDeviceToDevice, s));
        }
        NKPR_HIP(hipStreamSynchronize(s));
    }
    if (!g->weighted) {
        g->inv32.alloc(4 * std::max<uint64_t>(n, 1));
        g->inv64.alloc(8 * std::max<uint64_t>(n, 1));
        if (n)
            hipLaunchKernelGGL(inv_degree_kernel, dim3(blocks_for(n)), dim3(256), 0, s,
                               g->perm.as<int32_t>(), n, outcnt_b.as<uint32_t>(), wdeg, g->inv64.as<double>(),
                               g->inv32.as<float>());
        NKPR_HIP(hipGetLastError());
    } else {
        // 1 / weighted out-degree, and whether every weight is an FP32 number (then the binned
        // FP64 sweep reads 4-byte weights instead of 8-byte coefficients)
        g->invw64.alloc(8 * std::max<uint64_t>(n, 1));
        if (n)
            hipLaunchKernelGGL(inv_degree_kernel, dim3(blocks_for(n)), dim3(256), 0, s,
                               g->perm.as<int32_t>(), n, outcnt_b.as<uint32_t>(), wdeg, g->invw64.as<double>(),
                               (float *)nullptr);
        NKPR_HIP(hipGetLastError());
        g->w32_exact = true;
        if (w64_in && m) {
            NKPR_HIP(hipMemsetAsync(flags_b.p, 0, 4, s));
            hipLaunchKernelGGL(fp32_exact_kernel, dim3((unsigned)std::min<uint64_t>(blocks_for(m), 2048)), dim3(256), 0, s,
                               w_in, w64_in, m, flags_b.as<uint32_t>());
            NKPR_HIP(hipGetLastError());
            uint32_t hf = 0;
            read_back(s, {{&hf, flags_b.p, 4}});
            g->w32_exact = hf == 0;
        }
    }

    // --- schedule ---------------------------------------------------------------------------
    std::vector<uint64_t> cp((size_t)n + 1);
    read_back(s, {{cp.data(), g->colptr.p, 8 * ((size_t)n + 1)}});
    build_schedule(g, cp);
    NKPR_HIP(hipStreamSynchronize(s));
    g->prep_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

namespace {

// Weighted out-degree of every compact source from the edges in sweep order.
__global__ void sweep_wdeg_kernel(const int32_t *src, const uint32_t *emap, uint64_t m,
                                  const float *w, const double *w64, double *wdeg) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = emap[i];
        atomicAdd(&wdeg[src[i]], w64 ? w64[o] : (double)w[o]);
    }
}

__global__ void gather_wdeg_kernel(const double *wdeg_z, const int32_t *perm, uint32_t n,
                                   double *wdeg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wdeg[i] = wdeg_z[perm[i]];
}

// coef = w / deg(src) for every edge in sweep order (the fold of graph creation).
__global__ void refold_kernel(const int32_t *src, const uint32_t *emap, uint64_t m,
                              const float *w, const double *w64, const double *wdeg,
                              float *c32, double *c64) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = emap[i];
        const double c = (w64 ? w64[o] : (double)w[o]) / wdeg[src[i]];
        c64[i] = c;
        c32[i] = (float)c;
    }
}

}  // namespace

// New edge weights for the same structure (nkpr_graph_set_weights).
void set_weights(nkpr_graph *g, const float *w, const double *w64, const double *out_wdeg,
                 int device_ptrs) {
    NKPR_CHECK(g != nullptr, "graph is NULL");
    NKPR_CHECK(g->weighted, "the graph was created without weights");
    NKPR_CHECK(w != nullptr || w64 != nullptr, "weights are NULL");
    const hipStream_t s = g->stream;
    const uint64_t m = g->m, z = g->z;
    const uint32_t n = g->n;
    DevBuf wb, w64b, degz, deg;
    if (!device_ptrs) {
        if (w) {
            wb.alloc(4 * std::max<uint64_t>(m, 1));
            NKPR_HIP(hipMemcpyAsync(wb.p, w, 4 * m, hipMemcpyHostToDevice, s));
            w = wb.as<float>();
        }
        if (w64) {
            w64b.alloc(8 * std::max<uint64_t>(m, 1));
            NKPR_HIP(hipMemcpyAsync(w64b.p, w64, 8 * m, hipMemcpyHostToDevice, s));
            w64 = w64b.as<double>();
        }
    }
    deg.alloc(8 * std::max<uint32_t>(n, 1));
    if (out_wdeg) {
        const double *dz = out_wdeg;
        if (!device_ptrs) {
            degz.alloc(8 * std::max<uint64_t>(z, 1));
            NKPR_HIP(hipMemcpyAsync(degz.p, out_wdeg, 8 * z, hipMemcpyHostToDevice, s));
            dz = degz.as<double>();
        }
        if (n)
            hipLaunchKernelGGL(gather_wdeg_kernel, dim3(blocks_for(n)), dim3(256), 0, s, dz,
                               g->perm.as<int32_t>(), n, deg.as<double>());
        DevBuf uf;
        uf.alloc(4);
        NKPR_HIP(hipMemsetAsync(uf.p, 0, 4, s));
        if (n)
            hipLaunchKernelGGL(unit_degree_kernel, dim3((unsigned)std::min<uint64_t>(blocks_for(n), 65536)),
                               dim3(256), 0, s, (const double *)deg.as<double>(), (uint64_t)n,
                               (const uint8_t *)nullptr, uf.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        uint32_t hf = 0;
        read_back(s, {{&hf, uf.p, 4}});
        g->unit_degrees = !(hf & kNotUnitDegree);
    } else {
        g->unit_degrees = false;
        NKPR_HIP(hipMemsetAsync(deg.p, 0, 8 * std::max<uint32_t>(n, 1), s));
        if (m)
            hipLaunchKernelGGL(sweep_wdeg_kernel, dim3((unsigned)std::min<uint64_t>(blocks_for(m), 65536)),
                               dim3(256), 0, s, g->src.as<int32_t>(), g->emap.as<uint32_t>(), m, w,
                               w64, deg.as<double>());
    }
    NKPR_HIP(hipGetLastError());
    if (m)
        hipLaunchKernelGGL(refold_kernel, dim3((unsigned)std::min<uint64_t>(blocks_for(m), 65536)),
                           dim3(256), 0, s, g->src.as<int32_t>(), g->emap.as<uint32_t>(), m, w, w64,
                           deg.as<double>(), g->coef32.as<float>(), g->coef64.as<double>());
    NKPR_HIP(hipGetLastError());
    if (n)
        hipLaunchKernelGGL(inv_of_kernel, dim3(blocks_for(n)), dim3(256), 0, s, (const double *)deg.as<double>(), n,
                           g->invw64.as<double>());
    NKPR_HIP(hipGetLastError());
    bool exact = true;
    if (w && w64 && m) {
        DevBuf flag;
        flag.alloc(4);
        NKPR_HIP(hipMemsetAsync(flag.p, 0, 4, s));
        hipLaunchKernelGGL(fp32_exact_kernel, dim3((unsigned)std::min<uint64_t>(blocks_for(m), 2048)), dim3(256), 0, s,
                           w, w64, m, flag.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        uint32_t hf = 0;
        read_back(s, {{&hf, flag.p, 4}});
        exact = hf == 0;
    } else if (!w) {
        exact = false;  // FP64 weights alone: not checked
    }
    if (exact != g->w32_exact) g->bn64.ready = false;  // the FP64 copy changes its coefficient type
    g->w32_exact = exact;
    NKPR_HIP(hipStreamSynchronize(s));
    // derived copies follow the coefficients
    g->b_coef_ready = false;
    g->rowsum_ready = false;
    g->bn32.coef_ready = false;
    g->bn64.coef_ready = false;
    g->adj_stale = true;
}

}  // namespace nkpr
