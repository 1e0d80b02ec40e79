// Binned application of the FP32 operator y = d P v for the GMRES correction solves of the
// mixed-precision path (the product of networkit/cpp/centrality/PageRank.cpp:56-66 without
// the teleport term).
//
// The pull sweep gathers v[src] once per in-edge: a random 4-byte access that costs a 128-byte
// line when it misses L2, which bounds the sweep by the gather request rate.  Here the product
// is split in two streaming passes (propagation blocking: Beamer, Asanovic and Patterson,
// "Reducing PageRank Communication via Propagation Blocking", IPDPS 2017) over a second copy
// of the edges, cut into segments by (source chunk of 32768 ids, destination bin of 8192 ids)
// and sorted by source inside a segment:
//   pass 1  per source chunk, segments in chunk-major order: v of the chunk is staged in LDS and
//           every edge writes its contribution val = v[src] * w/deg(src), each segment's run to
//           the segment's place in bin-major order;
//   pass 2  per destination bin, segments in bin-major order, so one contiguous stream per bin:
//           the contributions are added into an LDS accumulator of 64-bit fixed-point
//           integers, which is written out as w = v - (d y + st) (a bin split over several work
//           items: the items' accumulators go to slots of their own, added in order).
// Every gather is an LDS access and every HBM access streams: about 17 bytes per edge (pass 1
// reads the source offset 2, the coefficient 4 and the bin-major place 1 and writes the
// contribution 4; pass 2 reads the contribution 4 and the destination offset 2) instead of one
// random line per gather.  Integer accumulation makes the result independent of the order of
// the atomics (the operator is deterministic) and is finer than FP32: the scale puts the
// largest possible row sum just below 2^61 units.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nkpr_internal.hpp"
#include "nkpr_run.hpp"

namespace nkpr {
namespace {

#ifndef NKPR_SPAN_SHIFT
#define NKPR_SPAN_SHIFT 15
#endif
#ifndef NKPR_BIN_SHIFT
#define NKPR_BIN_SHIFT 13
#endif
constexpr int kSpanShift = NKPR_SPAN_SHIFT;
constexpr uint32_t kSpan = 1u << kSpanShift;  // ids per source chunk (FP32 values staged in LDS)
constexpr int kBinShift = NKPR_BIN_SHIFT;
constexpr uint32_t kBin = 1u << kBinShift;    // ids per destination bin (64-bit accumulators in LDS)
static_assert(kSpanShift <= 15 && kBinShift <= 16, "offsets are stored in 16 bits");
constexpr int kPassThreads = 1024;
constexpr uint64_t kPass1Edges = 1u << 19;  // edges per pass-1 work item (at most)
constexpr uint64_t kPass2Edges = 1u << 20;  // edges per pass-2 work item (at most)
constexpr uint32_t kItemSole = 1;           // the bin's only pass-2 item: writes w directly
// Every segment is padded to a multiple of kQuad edges (padding slots contribute 0), so both
// passes read and write whole quads with 4-, 8- and 16-byte accesses.
constexpr uint32_t kQuad = 4;
constexpr uint32_t kPadEdge = 0xFFFFFFFFu;
// padded edge positions (and the quad loop counters) are 32-bit
constexpr uint64_t kMaxPadded = (1ull << 32) - (1ull << 16);
// Quads per lane and step of the two passes: the next step's loads are issued before this
// step's quads are used, so up to 2 * k quads per lane are in flight.  Two pass-2 quads keep
// the pass-2 kernel within 64 VGPRs, so that with 8192-row bins (64 KB of accumulators) two
// blocks share a CU; measured on R-MAT scale 26 this beats 16384-row bins with deeper
// prefetch.  A 16384-id chunk (two pass-1 blocks per CU) measured slower than 32768.
#ifndef NKPR_P1_QUADS
#define NKPR_P1_QUADS 1
#endif
#ifndef NKPR_P2_QUADS
#define NKPR_P2_QUADS 2
#endif
constexpr int kP1Quads = NKPR_P1_QUADS, kP2Quads = NKPR_P2_QUADS;

// Pass 2 adds the contributions as 64-bit fixed-point integers: integer addition is exact and
// associative (the result does not depend on the order of the atomics) and LDS integer
// atomics run at the rate of plain LDS stores.  The scale is 2^e with every partial sum below
// 2^61 units: |sum| <= max|v| * max_u sum_in coef(u) = bound.
__device__ __forceinline__ int fixed_exponent(const uint32_t *vmax_bits, double insum) {
    const double bound = (double)__uint_as_float(*vmax_bits) * insum;
    return bound > 0.0 ? 60 - ilogb(bound) : 0;
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// four contributions of pass 2: destination offsets (u16 pairs) and values
struct Quad {
    u32x2 d;
    f32x4 a;
};

__device__ __forceinline__ Quad load_quad(const uint16_t *bdst, const float *bval, uint32_t e) {
    Quad o;
    o.d = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(bdst + e));
    o.a = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(bval + e));
    return o;
}

__device__ __forceinline__ void fx_add(unsigned long long *acc, uint32_t d, float x, float scale) {
    atomicAdd(&acc[d], (unsigned long long)__float2ll_rn(x * scale));
}

__device__ __forceinline__ void add_quad(unsigned long long *acc, const Quad &o, float scale) {
    fx_add(acc, o.d.x & 0xFFFFu, o.a.x, scale);
    fx_add(acc, o.d.x >> 16, o.a.y, scale);
    fx_add(acc, o.d.y & 0xFFFFu, o.a.z, scale);
    fx_add(acc, o.d.y >> 16, o.a.w, scale);
}

inline unsigned nb(uint64_t n, int t = 256) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + t - 1) / t, 65536));
}

__global__ void bin_row_kernel(const uint32_t *rowptr, uint32_t n, uint32_t *row) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < n; w += nw)
        for (uint32_t e = rowptr[w] + lane; e < rowptr[w + 1]; e += 64) row[e] = (uint32_t)w;
}

// key = ((chunk * nbins + bin) << 15) | source offset: inside a segment the edges are sorted
// by source, which spreads the destinations of a wave's atomics over the accumulator
// (chunks of kSpan sources, bins of kBin destinations)
__global__ void bin_key_kernel(const int32_t *src, const uint32_t *row, uint64_t m, uint32_t nbins,
                               uint64_t *key, uint32_t *val) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)src[e];
        const uint64_t seg = (uint64_t)(v >> kSpanShift) * nbins + (row[e] >> kBinShift);
        key[e] = (seg << kSpanShift) | (v & (kSpan - 1));
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

// order[p] = edge at chunk-major padded position p: the sorted edge j of segment s goes to
// cseg[s] + (j - rseg[s]) (rseg: unpadded segment starts, cseg: padded ones)
__global__ void bin_place_kernel(const uint64_t *key, const uint32_t *ids, uint64_t m,
                                 const uint32_t *rseg, const uint32_t *cseg, uint32_t *order) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t sidx = key[j] >> kSpanShift;
        order[cseg[sidx] + ((uint32_t)j - rseg[sidx])] = ids[j];
    }
}

// pos[k] = bin-major position of chunk-major quad k; one wave per segment s, whose padded
// chunk-major run [cseg[s], cseg[s+1]) goes to bseg[s] onwards
__global__ void bin_pos_kernel(const uint32_t *cseg, const uint32_t *bseg, uint64_t nseg,
                               uint32_t *pos) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t s = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nseg; s += nw) {
        const uint32_t c0 = cseg[s], c1 = cseg[s + 1], b0 = bseg[s];
        for (uint32_t e = c0 + kQuad * lane; e < c1; e += kQuad * 64) pos[e / kQuad] = b0 + (e - c0);
    }
}

// Source offsets and coefficients in chunk-major order, destination offsets in bin-major
// order.  A padding slot reads the zero entry tab[kSpan] of pass 1, so it contributes 0.
__global__ void bin_emit_kernel(const uint32_t *order, const uint32_t *pos, uint64_t mp,
                                const int32_t *src, const uint32_t *row, const float *coef,
                                uint16_t *bsrc, uint16_t *bdst, float *bcoef) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < mp;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = order[j];
        const uint32_t jb = pos[j / kQuad] + (uint32_t)(j % kQuad);
        if (e == kPadEdge) {
            bsrc[j] = (uint16_t)kSpan;
            bdst[jb] = 0;
            if (bcoef) bcoef[j] = 0.f;
            continue;
        }
        bsrc[j] = (uint16_t)((uint32_t)src[e] & (kSpan - 1));
        bdst[jb] = (uint16_t)(row[e] & (kBin - 1));
        if (bcoef) bcoef[j] = coef[e];
    }
}

// The passes run one persistent block per LDS-sized slot of the device; a block takes the next
// work item from a queue counter until the list is done (the host sorts the items by
// decreasing size, so the largest start first and the short ones fill in at the end).  Every
// item body ends with a barrier, after which thread 0 may overwrite the shared slot.
__device__ __forceinline__ uint32_t next_item(uint32_t *queue, uint32_t *slot) {
    if (threadIdx.x == 0) *slot = atomicAdd(queue, 1u);
    __syncthreads();
    return *slot;
}

// Pass 1: item = {chunk, first edge, end edge} (chunk-major, quad-aligned).  The contribution
// of an edge is v[src] * coef (weighted) or v[src] / deg(src) (unweighted: inv), stored at the
// edge's bin-major place.  Also finds the largest |v| of the sources (the scale of pass 2;
// sinks contribute nothing), as the bits of a non-negative float, which order like unsigned
// integers.
__global__ __launch_bounds__(kPassThreads) void bin_scatter_kernel(
    const uint4 *items, uint32_t nitems, const float *v, const float *inv, uint32_t n,
    const uint16_t *bsrc, const float *bcoef, const uint32_t *pos, float *bval,
    uint32_t *vmax_bits, uint32_t *queue) {
    __shared__ float tab[kSpan + 1];
    __shared__ uint32_t slot;
    if (threadIdx.x == 0) tab[kSpan] = 0.f;  // read by the padding slots
    float vmax = 0.f;
    for (uint32_t it = next_item(queue, &slot); it < nitems; it = next_item(queue, &slot)) {
        const uint4 item = items[it];
        const uint32_t base = item.x << kSpanShift;
        for (uint32_t i = threadIdx.x; i < kSpan; i += kPassThreads) {
            const uint32_t u = base + i;
            float x = 0.f;
            if (u < n) {
                const float vu = v[u];
                vmax = fmaxf(vmax, fabsf(vu));
                x = inv ? vu * inv[u] : vu;
            }
            tab[i] = x;
        }
        __syncthreads();
        // thread t handles quads q0 + t + j * kPassThreads, so every load of a wave is
        // contiguous (8-byte source offsets, 16-byte coefficients, 4-byte places); kP1Quads
        // quads per step, loaded one step ahead
        const uint32_t q1 = item.z / kQuad, qt = item.y / kQuad + threadIdx.x;
        constexpr uint32_t kStep = kP1Quads * kPassThreads;
        const f32x4 one = {1.f, 1.f, 1.f, 1.f};
        u32x2 sp[kP1Quads];
        f32x4 c[kP1Quads];
        uint32_t at[kP1Quads];
#pragma unroll
        for (int k = 0; k < kP1Quads; ++k) {
            const uint32_t q = qt + k * kPassThreads;
            if (q < q1) {
                sp[k] = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(bsrc + q * kQuad));
                c[k] = bcoef ? __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(bcoef + q * kQuad)) : one;
                at[k] = __builtin_nontemporal_load(pos + q);
            }
        }
        for (uint32_t b = qt; b < q1; b += kStep) {
            u32x2 spn[kP1Quads];
            f32x4 cn[kP1Quads];
            uint32_t atn[kP1Quads];
#pragma unroll
            for (int k = 0; k < kP1Quads; ++k) {
                const uint32_t q = b + kStep + k * kPassThreads;
                if (q < q1) {
                    spn[k] = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(bsrc + q * kQuad));
                    cn[k] = bcoef ? __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(bcoef + q * kQuad)) : one;
                    atn[k] = __builtin_nontemporal_load(pos + q);
                }
            }
#pragma unroll
            for (int k = 0; k < kP1Quads; ++k) {
                const uint32_t q = b + k * kPassThreads;
                if (q < q1) {
                    f32x4 r;
                    r.x = tab[sp[k].x & 0xFFFFu] * c[k].x;
                    r.y = tab[sp[k].x >> 16] * c[k].y;
                    r.z = tab[sp[k].y & 0xFFFFu] * c[k].z;
                    r.w = tab[sp[k].y >> 16] * c[k].w;
                    __builtin_nontemporal_store(r, reinterpret_cast<f32x4 *>(bval + at[k]));
                }
                sp[k] = spn[k];
                c[k] = cn[k];
                at[k] = atn[k];
            }
        }
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(vmax_bits, __float_as_uint(vmax));
}

// Adds the quads f0, f0 + 64, f0 + 128, ... below `end` (quads from the start of bval) into
// acc, kP2Quads per step with the next step's loads issued before this step's adds.
__device__ __forceinline__ void stream_quads(unsigned long long *acc, const uint16_t *bdst,
                                             const float *bval, float scale, uint32_t f0,
                                             uint32_t end) {
    if (f0 >= end) return;
    const Quad none = {{0u, 0u}, {0.f, 0.f, 0.f, 0.f}};
    Quad cur[kP2Quads];
    bool has[kP2Quads];
#pragma unroll
    for (int k = 0; k < kP2Quads; ++k) {
        const uint32_t f = f0 + 64u * k;
        has[k] = f < end;
        cur[k] = has[k] ? load_quad(bdst, bval, f * kQuad) : none;
    }
    for (uint32_t f = f0 + 64u * kP2Quads; f < end; f += 64u * kP2Quads) {
        Quad nxt[kP2Quads];
        bool hn[kP2Quads];
#pragma unroll
        for (int k = 0; k < kP2Quads; ++k) {
            const uint32_t fk = f + 64u * k;
            hn[k] = fk < end;
            nxt[k] = hn[k] ? load_quad(bdst, bval, fk * kQuad) : none;
        }
#pragma unroll
        for (int k = 0; k < kP2Quads; ++k)
            if (has[k]) add_quad(acc, cur[k], scale);
#pragma unroll
        for (int k = 0; k < kP2Quads; ++k) {
            cur[k] = nxt[k];
            has[k] = hn[k];
        }
    }
#pragma unroll
    for (int k = 0; k < kP2Quads; ++k)
        if (has[k]) add_quad(acc, cur[k], scale);
}

// Pass 2: item = {bin, first edge, end edge, flags | slot << 2} (bin-major, quad-aligned).  Each
// wave streams an equal, contiguous share of the item's quads (64 consecutive quads per load,
// one per lane) into the bin's LDS accumulator.  A bin held by one item is written out as
// w = v - (d y + st) at once; a split bin's items write their accumulators to slots of their
// own.  Without `minus` the output is B v = d y + st instead (the operator whose Krylov space
// GMRES builds).
__global__ __launch_bounds__(kPassThreads) void bin_gather_kernel(
    const uint4 *items, uint32_t nitems, const uint16_t *bdst, const float *bval, long long *part,
    uint32_t n, const uint32_t *vmax_bits, double insum, const float *v, bool minus, float d,
    const double *sink_sum, float *w, uint32_t *queue) {
    __shared__ unsigned long long acc[kBin];
    __shared__ uint32_t slot;
    const int ex = fixed_exponent(vmax_bits, insum);
    const float scale = ldexpf(1.f, ex);
    const double unscale = ldexp(1.0, -ex);
    const float st = sink_sum ? (float)(d * *sink_sum / (double)n) : 0.f;
    constexpr uint32_t kWaves = kPassThreads / 64;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t it = next_item(queue, &slot); it < nitems; it = next_item(queue, &slot)) {
        const uint4 item = items[it];
        for (uint32_t i = threadIdx.x; i < kBin; i += kPassThreads) acc[i] = 0ull;
        __syncthreads();
        const uint32_t q0 = item.y / kQuad, total = item.z / kQuad - q0;
        const uint32_t share = ((total + kWaves - 1) / kWaves + 63) & ~63u;
        const uint32_t wb = min(total, wave * share), we = min(total, wb + share);
        stream_quads(acc, bdst, bval, scale, q0 + wb + lane, q0 + we);
        __syncthreads();
        const uint32_t base = item.x << kBinShift;
        for (uint32_t i = threadIdx.x; i < kBin; i += kPassThreads) {
            const uint32_t u = base + i;
            if (u >= n) break;
            if (item.w & kItemSole) {
                const float y = (float)((double)(long long)acc[i] * unscale);
                w[u] = minus ? v[u] - (d * y + st) : d * y + st;
            } else {
                part[(size_t)(item.w >> 2) * kBin + i] = (long long)acc[i];
            }
        }
        __syncthreads();
    }
}

// Bins split over several items: y = the items' partial accumulators added in item order, then
// w = v - (d y + st) (minus) or w = d y + st.  mb[j] = {bin, first slot, slot count}.
__global__ void bin_reduce_kernel(const uint4 *mb, uint32_t nmb, const long long *part, uint32_t n,
                                  const uint32_t *vmax_bits, double insum, const float *v, bool minus,
                                  float d, const double *sink_sum, float *w) {
    const double unscale = ldexp(1.0, -fixed_exponent(vmax_bits, insum));
    const float st = sink_sum ? (float)(d * *sink_sum / (double)n) : 0.f;
    const uint64_t total = (uint64_t)nmb * kBin;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 b = mb[t >> kBinShift];
        const uint32_t i = (uint32_t)(t & (kBin - 1));
        const uint32_t u = (b.x << kBinShift) + i;
        if (u >= n) continue;
        long long acc = 0;
        for (uint32_t q = 0; q < b.z; ++q) acc += part[(size_t)(b.y + q) * kBin + i];
        const float y = (float)((double)acc * unscale);
        w[u] = minus ? v[u] - (d * y + st) : d * y + st;
    }
}

// ctl[0] = largest |v| (bits), ctl[1], ctl[2] = queue counters of the two passes
__global__ void bin_clear_kernel(uint32_t *ctl) {
    if (threadIdx.x < 3) ctl[threadIdx.x] = 0u;
}

// largest in-coefficient sum of a row: sum_in w/deg(src) (weighted) or sum_in 1/deg(src); one
// wave per row, the block's largest sum goes to `best` with one atomic
__global__ void bin_insum_kernel(const uint32_t *rowptr, uint32_t n, const int32_t *src,
                                 const float *coef, const float *inv, unsigned long long *best) {
    __shared__ double wbest[4];
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    double mine = 0.0;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < n; w += nw) {
        double sum = 0.0;
        for (uint32_t e = rowptr[w] + lane; e < rowptr[w + 1]; e += 64)
            sum += coef ? (double)coef[e] : (double)inv[src[e]];
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
        mine = fmax(mine, sum);
    }
    if (lane == 0) wbest[threadIdx.x >> 6] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wbest[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) b = fmax(b, wbest[q]);
        atomicMax(best, (unsigned long long)__double_as_longlong(b));
    }
}

// Bound of the fixed-point sums: the largest in-coefficient sum of a row (x 1.01 for the FP32
// rounding of the coefficients).
void binned_insum(nkpr_graph *g) {
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    DevBuf best;
    best.alloc(8);
    NKPR_HIP(hipMemsetAsync(best.p, 0, 8, s));
    if (n)
        hipLaunchKernelGGL(bin_insum_kernel, dim3((unsigned)std::min<uint64_t>(((uint64_t)n * 64 + 255) / 256, 4096)),
                           dim3(256), 0, s, g->b_rowptr.as<uint32_t>(), n, g->b_src.as<int32_t>(),
                           g->weighted ? g->b_coef32.as<float>() : nullptr,
                           g->weighted ? nullptr : g->inv32.as<float>(), best.as<unsigned long long>());
    NKPR_HIP(hipGetLastError());
    unsigned long long hb = 0;
    NKPR_HIP(hipMemcpyAsync(&hb, best.p, 8, hipMemcpyDeviceToHost, s));
    NKPR_HIP(hipStreamSynchronize(s));
    double v;
    std::memcpy(&v, &hb, 8);
    g->bn_insum = v * 1.01;
}

// Items sorted by decreasing cost (stable: equal costs keep their order).
void sort_by_cost(std::vector<uint4> &items, const std::vector<uint64_t> &cost) {
    std::vector<uint32_t> idx(items.size());
    for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    std::vector<uint4> out(items.size());
    for (size_t i = 0; i < idx.size(); ++i) out[i] = items[idx[i]];
    items.swap(out);
}

}  // namespace

// The binned operator replaces the pull sweep in the GMRES corrections of graphs with at least
// kBinnedMinEdges edges (on R-MAT scale 26: 3.8 ms per application against 8.4 ms; on graphs
// whose gathers stay in the caches the pull sweep is as fast and needs no second edge copy).
// NKPR_BINNED=0 / 1 forces either operator.
constexpr uint64_t kBinnedMinEdges = 1ull << 26;

bool binned_enabled(const nkpr_graph *g) {
    if (g->bn_off) return false;
    if (const char *e = std::getenv("NKPR_BINNED"); e && *e) return std::atoi(e) != 0;
    return g->m >= kBinnedMinEdges;
}

// Builds the binned edge copy (once per graph; coefficients again after new weights).  Returns
// false when the padded copy would not fit 32-bit positions (the pull sweep is used instead).
bool binned_layout(nkpr_graph *g) {
    batch_layout(g);
    if (g->bn_off) return false;
    if (g->bn_ready && g->bn_coef_ready) return true;
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const uint32_t nchunks = std::max<uint32_t>(1, (g->sink_begin + kSpan - 1) / kSpan);
    const uint32_t nbins = (n + kBin - 1) / kBin;
    const uint64_t nseg = (uint64_t)nchunks * nbins;
    DevBuf row, k1, k2, v1, v2, tmp, rseg, cseg, bseg;
    row.alloc(4 * std::max<uint64_t>(m, 1));
    if (n)
        hipLaunchKernelGGL(bin_row_kernel, dim3((unsigned)std::min<uint64_t>(((uint64_t)n * 64 + 255) / 256, 1u << 18)), dim3(256), 0, s,
                           g->b_rowptr.as<uint32_t>(), n, row.as<uint32_t>());
    if (g->bn_ready) {
        // new coefficients, same order: re-emit from the kept order
        if (g->bn_mp && g->weighted)
            hipLaunchKernelGGL(bin_emit_kernel, dim3(nb(g->bn_mp)), dim3(256), 0, s, g->bn_order.as<uint32_t>(),
                               g->bn_pos.as<uint32_t>(), g->bn_mp, g->b_src.as<int32_t>(), row.as<uint32_t>(),
                               g->b_coef32.as<float>(), g->bn_src.as<uint16_t>(), g->bn_dst.as<uint16_t>(),
                               g->bn_coef.as<float>());
        NKPR_HIP(hipGetLastError());
        NKPR_HIP(hipStreamSynchronize(s));
        binned_insum(g);
        g->bn_coef_ready = true;
        return true;
    }
    // sort the edges by key and find the unpadded segment starts (chunk-major)
    std::vector<uint32_t> seg(nseg + 1, 0u);
    rseg.alloc(4 * (nseg + 1));
    if (m) {
        k1.alloc(8 * m);
        k2.alloc(8 * m);
        v1.alloc(4 * m);
        v2.alloc(4 * m);
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
        // keep the sorted halves in k1 / v1 and free the others
        if (kb.Current() != k1.as<uint64_t>()) std::swap(k1.p, k2.p), std::swap(k1.bytes, k2.bytes);
        if (vb.Current() != v1.as<uint32_t>()) std::swap(v1.p, v2.p), std::swap(v1.bytes, v2.bytes);
        NKPR_HIP(hipStreamSynchronize(s));
        k2.reset();
        v2.reset();
        tmp.reset();
        hipLaunchKernelGGL(bin_seg_kernel, dim3(nb(nseg + 1)), dim3(256), 0, s,
                           (const uint64_t *)k1.as<uint64_t>(), m, nseg, rseg.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        NKPR_HIP(hipMemcpyAsync(seg.data(), rseg.p, 4 * (nseg + 1), hipMemcpyDeviceToHost, s));
        NKPR_HIP(hipStreamSynchronize(s));
    }
    // padded segment starts: every segment rounded up to whole quads, laid out chunk-major
    // (pass 1) and bin-major (pass 2); both indexed by the chunk-major segment number
    auto padded = [&](uint64_t q) { return ((uint64_t)(seg[q + 1] - seg[q]) + kQuad - 1) / kQuad * kQuad; };
    std::vector<uint32_t> chost(nseg + 1), bhost(nseg + 1);
    uint64_t mp = 0;
    for (uint64_t q = 0; q < nseg; ++q) {
        chost[q] = (uint32_t)mp;
        mp += padded(q);
        if (mp > kMaxPadded) {
            g->bn_off = true;
            if (std::getenv("NKPR_DEBUG"))
                std::fprintf(stderr, "[nkpr] binned layout: padded copy exceeds 32-bit positions\n");
            return false;
        }
    }
    chost[nseg] = (uint32_t)mp;
    std::vector<uint32_t> bin_start(nbins + 1);
    uint64_t at = 0;
    for (uint32_t b = 0; b < nbins; ++b) {
        bin_start[b] = (uint32_t)at;
        for (uint32_t c = 0; c < nchunks; ++c) {
            const uint64_t q = (uint64_t)c * nbins + b;
            bhost[q] = (uint32_t)at;
            at += padded(q);
        }
    }
    bin_start[nbins] = (uint32_t)at;
    bhost[nseg] = (uint32_t)mp;
    g->bn_mp = mp;
    cseg.alloc(4 * (nseg + 1));
    bseg.alloc(4 * (nseg + 1));
    NKPR_HIP(hipMemcpyAsync(cseg.p, chost.data(), 4 * (nseg + 1), hipMemcpyHostToDevice, s));
    NKPR_HIP(hipMemcpyAsync(bseg.p, bhost.data(), 4 * (nseg + 1), hipMemcpyHostToDevice, s));
    g->bn_order.alloc(4 * std::max<uint64_t>(mp, 1));
    NKPR_HIP(hipMemsetAsync(g->bn_order.p, 0xFF, 4 * std::max<uint64_t>(mp, 1), s));
    if (m)
        hipLaunchKernelGGL(bin_place_kernel, dim3(nb(m)), dim3(256), 0, s, (const uint64_t *)k1.as<uint64_t>(),
                           (const uint32_t *)v1.as<uint32_t>(), m, (const uint32_t *)rseg.as<uint32_t>(),
                           (const uint32_t *)cseg.as<uint32_t>(), g->bn_order.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    g->bn_pos.alloc(4 * std::max<uint64_t>(mp / kQuad, 1));
    if (mp)
        hipLaunchKernelGGL(bin_pos_kernel, dim3((unsigned)std::min<uint64_t>((nseg * 64 + 255) / 256, 1u << 18)),
                           dim3(256), 0, s, (const uint32_t *)cseg.as<uint32_t>(),
                           (const uint32_t *)bseg.as<uint32_t>(), nseg, g->bn_pos.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    g->bn_src.alloc(2 * std::max<uint64_t>(mp, 1) + 64);
    g->bn_dst.alloc(2 * std::max<uint64_t>(mp, 1) + 64);
    g->bn_val.alloc(4 * std::max<uint64_t>(mp, 1) + 64);
    if (g->weighted) g->bn_coef.alloc(4 * std::max<uint64_t>(mp, 1) + 64);
    if (mp)
        hipLaunchKernelGGL(bin_emit_kernel, dim3(nb(mp)), dim3(256), 0, s, g->bn_order.as<uint32_t>(),
                           g->bn_pos.as<uint32_t>(), mp, g->b_src.as<int32_t>(), row.as<uint32_t>(),
                           g->weighted ? g->b_coef32.as<float>() : nullptr, g->bn_src.as<uint16_t>(),
                           g->bn_dst.as<uint16_t>(), g->weighted ? g->bn_coef.as<float>() : nullptr);
    NKPR_HIP(hipGetLastError());
    NKPR_HIP(hipStreamSynchronize(s));

    // work items (NKPR_BINNED_ITEM overrides both sizes, e.g. to exercise split bins); sizes in
    // whole quads so that every item boundary stays quad-aligned
    uint64_t e1max = kPass1Edges, e2max = kPass2Edges;
    if (const char *ev = std::getenv("NKPR_BINNED_ITEM")) {
        const uint64_t v = std::strtoull(ev, nullptr, 10);
        if (v > 0) e1max = e2max = (v + kQuad - 1) / kQuad * kQuad;
    }
    std::vector<uint4> p1, p2, mb;  // mb: {bin, first slot, slot count} of the split bins
    std::vector<uint64_t> cost1, cost2;
    for (uint32_t c = 0; c < nchunks; ++c) {
        const uint32_t b0 = chost[(uint64_t)c * nbins], b1 = chost[(uint64_t)(c + 1) * nbins];
        for (uint64_t e = b0; e < b1; e += e1max) {
            const uint64_t e1 = std::min<uint64_t>(e + e1max, b1);
            p1.push_back(make_uint4(c, (uint32_t)e, (uint32_t)e1, 0));
            cost1.push_back(e1 - e + kSpan);  // edges + the LDS table fill
        }
    }
    uint32_t slots = 0;
    for (uint32_t b = 0; b < nbins; ++b) {
        const uint64_t s0 = bin_start[b], s1 = bin_start[b + 1];
        const uint32_t pieces = (uint32_t)std::max<uint64_t>(1, (s1 - s0 + e2max - 1) / e2max);
        for (uint32_t i = 0; i < pieces; ++i) {
            const uint64_t e0 = s0 + i * e2max, e1 = std::min<uint64_t>(e0 + e2max, s1);
            const uint32_t flags = pieces == 1 ? kItemSole : (slots + i) << 2;
            p2.push_back(make_uint4(b, (uint32_t)e0, (uint32_t)e1, flags));
            cost2.push_back(e1 - e0 + kBin);  // edges + the accumulator clear and flush
        }
        if (pieces > 1) {
            // every item of a split bin writes its accumulator to a slot of its own
            mb.push_back(make_uint4(b, slots, pieces, 0));
            slots += pieces;
        }
    }
    sort_by_cost(p1, cost1);
    sort_by_cost(p2, cost2);
    g->bn_p1.alloc(sizeof(uint4) * std::max<size_t>(p1.size(), 1));
    g->bn_p2.alloc(sizeof(uint4) * std::max<size_t>(p2.size(), 1));
    if (!p1.empty()) NKPR_HIP(hipMemcpyAsync(g->bn_p1.p, p1.data(), sizeof(uint4) * p1.size(), hipMemcpyHostToDevice, s));
    if (!p2.empty()) NKPR_HIP(hipMemcpyAsync(g->bn_p2.p, p2.data(), sizeof(uint4) * p2.size(), hipMemcpyHostToDevice, s));
    g->bn_np1 = (uint32_t)p1.size();
    g->bn_np2 = (uint32_t)p2.size();
    g->bn_multi = slots;
    g->bn_nmb = (uint32_t)mb.size();
    g->bn_mb.alloc(sizeof(uint4) * std::max<size_t>(mb.size(), 1));
    if (!mb.empty()) NKPR_HIP(hipMemcpyAsync(g->bn_mb.p, mb.data(), sizeof(uint4) * mb.size(), hipMemcpyHostToDevice, s));
    g->bn_part.alloc(8 * (size_t)std::max<uint32_t>(slots, 1) * kBin);
    binned_insum(g);
    g->bn_vmax.alloc(16);
    NKPR_HIP(hipStreamSynchronize(s));
    if (std::getenv("NKPR_DEBUG"))
        std::fprintf(stderr, "[nkpr] binned layout: %u chunks x %u bins, %llu padded edges, %u + %u items (%u split)\n",
                     nchunks, nbins, (unsigned long long)mp, g->bn_np1, g->bn_np2, g->bn_multi);
    g->bn_ready = true;
    g->bn_coef_ready = true;
    return true;
}

// w = B v = d P v (+ d S(v)/n) through the binned passes, or w = A v = v - B v (minus).
void binned_apply(nkpr_graph *g, const float *v, float *w, double damp, bool distribute,
                  bool minus, hipEvent_t ev0, hipEvent_t ev1) {
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    if (!n) return;
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    const double *sink_sum = nullptr;
    if (distribute && g->sink_begin < n) {
        launch_range_sum<float>(v, g->sink_begin, n, g->red_partials.as<double>(), g->red_out.as<double>(), s);
        sink_sum = g->red_out.as<double>();
    }
    if (!g->bn_grid1) {
        // persistent grids: every block that fits on the device at once
        int b1 = 0, b2 = 0;
        NKPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, bin_scatter_kernel, kPassThreads, 0));
        NKPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b2, bin_gather_kernel, kPassThreads, 0));
        g->bn_grid1 = (uint32_t)std::max(1, b1 * cus);
        g->bn_grid2 = (uint32_t)std::max(1, b2 * cus);
    }
    if (ev0) NKPR_HIP(hipEventRecord(ev0, s));
    uint32_t *ctl = g->bn_vmax.as<uint32_t>();
    hipLaunchKernelGGL(bin_clear_kernel, dim3(1), dim3(64), 0, s, ctl);
    hipLaunchKernelGGL(bin_scatter_kernel, dim3(std::min<uint32_t>(std::max<uint32_t>(g->bn_np1, 1), g->bn_grid1)),
                       dim3(kPassThreads), 0, s, g->bn_p1.as<uint4>(), g->bn_np1, v,
                       g->weighted ? nullptr : g->inv32.as<float>(), n, g->bn_src.as<uint16_t>(),
                       g->weighted ? g->bn_coef.as<float>() : nullptr, g->bn_pos.as<uint32_t>(),
                       g->bn_val.as<float>(), ctl, ctl + 1);
    hipLaunchKernelGGL(bin_gather_kernel, dim3(std::min<uint32_t>(std::max<uint32_t>(g->bn_np2, 1), g->bn_grid2)),
                       dim3(kPassThreads), 0, s, g->bn_p2.as<uint4>(), g->bn_np2,
                       g->bn_dst.as<uint16_t>(), g->bn_val.as<float>(), g->bn_part.as<long long>(), n,
                       (const uint32_t *)ctl, g->bn_insum, v, minus, (float)damp, sink_sum, w, ctl + 2);
    if (g->bn_nmb)
        hipLaunchKernelGGL(bin_reduce_kernel, dim3(nb((uint64_t)g->bn_nmb * kBin)), dim3(256), 0, s,
                           g->bn_mb.as<uint4>(), g->bn_nmb, (const long long *)g->bn_part.as<long long>(), n,
                           (const uint32_t *)ctl, g->bn_insum, v, minus, (float)damp, sink_sum, w);
    if (ev1) NKPR_HIP(hipEventRecord(ev1, s));
    NKPR_HIP(hipGetLastError());
}

}  // namespace nkpr
