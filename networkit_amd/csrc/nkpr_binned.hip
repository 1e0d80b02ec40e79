// Binned applications of the PageRank product P x: the FP32 operator y = d P v of the GMRES
// correction solves of the mixed-precision path, and the FP64 sweep x' = d P x + b with the
// reference's epilogue (the product of networkit/cpp/centrality/PageRank.cpp:56-66).
//
// The pull sweep gathers x[src] once per in-edge: a random access that costs a 128-byte line
// when it misses L2, which bounds the sweep by the gather request rate.  Here the product is
// split in two streaming passes (propagation blocking: Beamer, Asanovic and Patterson,
// "Reducing PageRank Communication via Propagation Blocking", IPDPS 2017) over a second copy
// of the edges, cut into segments by (source chunk, destination bin of 8192 ids) and sorted by
// source inside a segment:
//   pass 1  per source chunk, segments in chunk-major order: x of the chunk is staged in LDS
//           and every edge writes its contribution x[src] * w/deg(src), each segment's run to
//           the segment's place in bin-major order (derived for every quad of 4 edges from a
//           segment-start mask and the first segment of its group of 64 quads, and one offset
//           per segment: about 0.1 byte per edge instead of a 1-byte stored place);
//   pass 2  per destination bin, segments in bin-major order, so one contiguous stream per
//           bin: the contributions are added into LDS accumulators of fixed-point integers,
//           from which the rows are finished (a bin split over several work items: the items'
//           accumulators go to slots of their own, added in order by the reduce kernel).
// Every gather is an LDS access and every HBM access streams.  Integer accumulation makes the
// result independent of the order of the atomics (deterministic), and exact: each row is the
// correctly scaled sum of its rounded products, rounded once.
//   FP32 copy (chunks of 32768 ids, FP32 table 128 KB): 64-bit accumulators, the largest
//       possible row sum just below 2^61 units; about 16 bytes per edge moved.
//   FP64 copy (chunks of 16384 ids, FP64 table 128 KB): 128-bit accumulators as two 64-bit
//       integers (whole units and 2^-lo_shift fractions of a unit), so the rows carry more
//       than FP64 precision; about 28 bytes per edge moved (24 when every weight is an FP32
//       number and the copy keeps 4-byte weights).  The row epilogue, the reductions and the
//       stopping rule are those of the pull sweep (nkpr_sweep.hip).
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
constexpr int kSpan32Shift = NKPR_SPAN_SHIFT;  // ids per source chunk of the FP32 copy
constexpr int kSpan64Shift = 14;               // ids per source chunk of the FP64 copy
constexpr int kBinShift = NKPR_BIN_SHIFT;
constexpr uint32_t kBin = 1u << kBinShift;     // ids per destination bin
static_assert(kSpan32Shift <= 15 && kBinShift <= 16, "offsets are stored in 16 bits");
constexpr int kPassThreads = 1024;
constexpr uint64_t kPass1Edges = 1u << 19;  // edges per pass-1 work item (at most)
constexpr uint64_t kPass2Edges = 1u << 20;  // edges per pass-2 work item (at most)
constexpr uint32_t kItemSole = 1;           // the bin's only pass-2 item: finishes the rows
// Every segment is padded to a multiple of kQuad edges (padding slots read the zero entry
// tab[span] of pass 1, so they contribute 0): both passes read and write whole quads.
constexpr uint32_t kQuad = 4;
constexpr uint32_t kPadEdge = 0xFFFFFFFFu;
// padded edge positions (and the quad loop counters) are 32-bit
constexpr uint64_t kMaxPadded = (1ull << 32) - (1ull << 16);
// Quads per lane and step of the two FP32 passes: the next step's loads are issued before this
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
// Hot sources of the FP32 operator (optional, NKPR_BINNED_HOT): the first `hot` relabelled ids
// (the sources with the most out-edges), at most kHot.  Their edges skip pass 1 and the
// contribution array: pass 2 keeps their values in an LDS table next to the accumulators and
// gathers from it, so such an edge moves 8 bytes (packed source and destination offsets,
// coefficient) instead of 16.  A hot entry packs (source << kBinShift) | destination offset;
// padding points at the zero entry tab[kHot].  With the table (64 KB) beside the accumulators
// (64 KB) one pass-2 block fits a CU, so that kernel keeps more quads in flight per lane.
constexpr int kHotShift = 14;
constexpr uint32_t kHot = 1u << kHotShift;
static_assert(kHotShift + 1 + kBinShift <= 32, "hot entries pack into 32 bits");
#ifndef NKPR_P2H_QUADS
#define NKPR_P2H_QUADS 4
#endif
constexpr int kP2HotQuads = NKPR_P2H_QUADS;

// Control words of one application: the largest |x| of the sources (the bits of a
// non-negative double order like unsigned integers) and the queue counters of the passes.
struct BinCtl {
    unsigned long long vmax;
    uint32_t q1, q2;
};

// The fixed-point scale is 2^e with every row sum below 2^61 units:
// |sum| <= max|x| * max_u sum_in coef(u) = bound.  e stays within [-126, 126] so that 2^e is a
// normal FP32 number (a smaller e only keeps fewer fraction bits of very small sums).
__device__ __forceinline__ int fixed_exponent(const BinCtl *ctl, double insum) {
    const double bound = __longlong_as_double((long long)ctl->vmax) * insum;
    return bound > 0.0 ? max(-126, min(126, 60 - ilogb(bound))) : 0;
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

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

// ---- layout construction ------------------------------------------------------------------

__global__ void bin_row_kernel(const uint32_t *rowptr, uint32_t n, uint32_t *row) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < n; w += nw)
        for (uint32_t e = rowptr[w] + lane; e < rowptr[w + 1]; e += 64) row[e] = (uint32_t)w;
}

// key = ((chunk * nbins + bin) << ss) | source offset: inside a segment the edges are sorted
// by source, which spreads the destinations of a wave's atomics over the accumulator
// (chunks of 2^ss sources, bins of kBin destinations)
// Hot edges (source below `hot`) go to the segments of a chunk past the last one, one per bin.
__global__ void bin_key_kernel(const int32_t *src, const uint32_t *row, uint64_t m, uint32_t nbins,
                               int ss, uint32_t nchunks, uint32_t hot, uint64_t *key, uint32_t *val) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)src[e];
        const uint32_t chunk = v < hot ? nchunks : (v >> ss);
        const uint64_t seg = (uint64_t)chunk * nbins + (row[e] >> kBinShift);
        key[e] = (seg << ss) | (v & ((1u << ss) - 1));
        val[e] = (uint32_t)e;
    }
}

// seg[s] = first sorted edge of segment s (binary search over the sorted keys)
__global__ void bin_seg_kernel(const uint64_t *key, uint64_t m, uint64_t nseg, int ss, uint32_t *seg) {
    for (uint64_t sidx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; sidx <= nseg;
         sidx += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = sidx << ss;
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
__global__ void bin_place_kernel(const uint64_t *key, const uint32_t *ids, uint64_t m, int ss,
                                 uint64_t nseg, const uint32_t *rseg, const uint32_t *cseg,
                                 uint32_t *order) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t sidx = key[j] >> ss;
        if (sidx >= nseg) continue;  // a hot edge
        order[cseg[sidx] + ((uint32_t)j - rseg[sidx])] = ids[j];
    }
}

// horder[p] = edge at padded hot position p: the sorted hot edge j of bin b (segment nseg + b)
// goes to hseg[b] + (j - rseg[nseg + b])
__global__ void bin_hot_place_kernel(const uint64_t *key, const uint32_t *ids, uint64_t j0,
                                     uint64_t m, int ss, uint64_t nseg, const uint32_t *rseg,
                                     const uint32_t *hseg, uint32_t *horder) {
    for (uint64_t j = j0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t sidx = key[j] >> ss;
        horder[hseg[sidx - nseg] + ((uint32_t)j - rseg[sidx])] = ids[j];
    }
}

// Hot entries: packed (source << kBinShift) | destination offset, and the coefficient; a
// padding slot points at the zero table entry kHot with coefficient 0.
__global__ void bin_hot_emit_kernel(const uint32_t *horder, uint64_t mh, const int32_t *src,
                                    const uint32_t *row, const float *coef, uint32_t *hidx,
                                    float *hcoef) {
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < mh;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = horder[p];
        if (e == kPadEdge) {
            hidx[p] = kHot << kBinShift;
            if (hcoef) hcoef[p] = 0.f;
            continue;
        }
        hidx[p] = ((uint32_t)src[e] << kBinShift) | (row[e] & (kBin - 1));
        if (hcoef) hcoef[p] = coef[e];
    }
}

// count of the edges whose source is below `bound` (one atomic per block)
__global__ void bin_count_below_kernel(const int32_t *src, uint64_t m, uint32_t bound,
                                       unsigned long long *count) {
    unsigned long long c = 0;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x)
        c += (uint32_t)src[e] < bound;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
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

// Bin-major place of chunk-major quad q from the group data: every segment starts on a quad,
// so the quad's segment is the group's first one plus the segment starts among lanes 1..l of
// the group, and its place is the quad's own position plus that segment's offset (modulo
// 2^32).  The segment index is clamped to the offset table.
__device__ __forceinline__ uint32_t derived_place(const unsigned long long *gmask,
                                                  const uint32_t *gfirst, const uint32_t *sdelta,
                                                  uint32_t nsd, uint32_t q) {
    const uint32_t grp = q >> 6, l = q & 63u;
    const unsigned long long below = (gmask[grp] & ~1ull) << (63u - l);
    const uint32_t sidx = min(gfirst[grp] + (uint32_t)__popcll(below), nsd - 1u);
    return q * kQuad + sdelta[sidx];
}

// flag = 1 when a derived place differs from pos (checked once per layout)
__global__ void bin_place_check_kernel(const uint32_t *pos, uint64_t nq,
                                       const unsigned long long *gmask, const uint32_t *gfirst,
                                       const uint32_t *sdelta, uint32_t nsd, uint32_t *flag) {
    int diff = 0;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq;
         q += (uint64_t)gridDim.x * blockDim.x)
        diff |= derived_place(gmask, gfirst, sdelta, nsd, (uint32_t)q) != pos[q];
    if (__syncthreads_or(diff) && threadIdx.x == 0) atomicOr(flag, 1u);
}

// Source offsets and coefficients in chunk-major order, destination offsets in bin-major
// order.  A padding slot gets the source offset 2^ss (the zero table entry).
template <typename C>
__global__ void bin_emit_kernel(const uint32_t *order, const uint32_t *pos, uint64_t mp, int ss,
                                const int32_t *src, const uint32_t *row, const C *coef,
                                uint16_t *bsrc, uint16_t *bdst, C *bcoef) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < mp;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = order[j];
        const uint32_t jb = pos[j / kQuad] + (uint32_t)(j % kQuad);
        if (e == kPadEdge) {
            bsrc[j] = (uint16_t)(1u << ss);
            bdst[jb] = 0;
            if (bcoef) bcoef[j] = (C)0;
            continue;
        }
        bsrc[j] = (uint16_t)((uint32_t)src[e] & ((1u << ss) - 1));
        bdst[jb] = (uint16_t)(row[e] & (kBin - 1));
        if (bcoef) bcoef[j] = coef[e];
    }
}

// The FP64 copy of a graph whose weights are all FP32 numbers keeps the weights instead of the
// folded coefficients: w = coef64 / (1/deg) recovers the weight exactly (its FP64 rounding
// error is far below half an FP32 unit), and pass 1 stages x / deg in its table.
__global__ void bin_emit_w32_kernel(const uint32_t *order, const uint32_t *pos, uint64_t mp, int ss,
                                    const int32_t *src, const uint32_t *row, const double *coef64,
                                    const double *invw, uint16_t *bsrc, uint16_t *bdst, float *bw) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < mp;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = order[j];
        const uint32_t jb = pos[j / kQuad] + (uint32_t)(j % kQuad);
        if (e == kPadEdge) {
            bsrc[j] = (uint16_t)(1u << ss);
            bdst[jb] = 0;
            bw[j] = 0.f;
            continue;
        }
        const uint32_t v = (uint32_t)src[e];
        bsrc[j] = (uint16_t)(v & ((1u << ss) - 1));
        bdst[jb] = (uint16_t)(row[e] & (kBin - 1));
        bw[j] = (float)(coef64[e] / invw[v]);
    }
}

// largest in-coefficient sum of a row: sum_in w/deg(src) (weighted) or sum_in 1/deg(src); one
// wave per row, the block's largest sum goes to `best` with one atomic
template <typename C>
__global__ void bin_insum_kernel(const uint32_t *rowptr, uint32_t n, const int32_t *src,
                                 const C *coef, const C *inv, unsigned long long *best) {
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

__global__ void bin_clear_kernel(BinCtl *ctl) {
    if (threadIdx.x == 0) {
        ctl->vmax = 0ull;
        ctl->q1 = 0u;
        ctl->q2 = 0u;
    }
}

// The passes run persistent blocks, as many as fit on the device; a block takes the next work
// item from a queue counter until the list is done (the host sorts the items by decreasing
// size, so the largest start first and the short ones fill in at the end).  Every item body
// ends with a barrier, after which thread 0 may overwrite the shared slot.
__device__ __forceinline__ uint32_t next_item(uint32_t *queue, uint32_t *slot) {
    if (threadIdx.x == 0) *slot = atomicAdd(queue, 1u);
    __syncthreads();
    return *slot;
}

__device__ __forceinline__ void note_vmax(BinCtl *ctl, double vmax) {
    for (int off = 32; off > 0; off >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&ctl->vmax, (unsigned long long)__double_as_longlong(vmax));
}

// ---- pass-1 tables --------------------------------------------------------------------------

// tab[i] = v[base + i] (* inv[base + i]) for i < kSpan (0 past n), and vmax = max |v|.  Every
// load of a thread is issued before its first LDS store, so a work item waits for one memory
// round trip here instead of one per element: 16-byte loads when the whole chunk lies below n
// and both vectors are 16-byte aligned, guarded element loads otherwise.
template <typename T, uint32_t kSpan>
__device__ __forceinline__ void fill_table(T *tab, const T *v, const T *inv, uint32_t base,
                                           uint32_t n, T &vmax) {
    typedef T t2 __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr uint32_t kLanes = 16 / sizeof(T);
    constexpr uint32_t kVec = kSpan / kLanes / kPassThreads;  // 16-byte loads per thread
    constexpr uint32_t kPer = kSpan / kPassThreads;           // elements per thread
    static_assert(kVec * kLanes * kPassThreads == kSpan, "table fill covers the chunk");
    const bool aligned = (((uintptr_t)v | (uintptr_t)inv) & 15) == 0;
    if (base + kSpan <= n && aligned) {
        const t2 *v2 = reinterpret_cast<const t2 *>(v + base);
        const t2 *i2 = reinterpret_cast<const t2 *>(inv + base);
        t2 a[kVec], b[kVec];
#pragma unroll
        for (uint32_t k = 0; k < kVec; ++k) a[k] = v2[threadIdx.x + k * kPassThreads];
        if (inv) {
#pragma unroll
            for (uint32_t k = 0; k < kVec; ++k) b[k] = i2[threadIdx.x + k * kPassThreads];
        }
#pragma unroll
        for (uint32_t k = 0; k < kVec; ++k) {
            t2 t = a[k];
#pragma unroll
            for (uint32_t l = 0; l < kLanes; ++l) vmax = fmax(vmax, fabs(a[k][l]));
            if (inv) t = a[k] * b[k];
            reinterpret_cast<t2 *>(tab)[threadIdx.x + k * kPassThreads] = t;
        }
    } else {
        T a[kPer], b[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t u = base + threadIdx.x + k * kPassThreads;
            a[k] = u < n ? v[u] : T(0);
            b[k] = (inv && u < n) ? inv[u] : T(1);
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            vmax = fmax(vmax, fabs(a[k]));
            tab[threadIdx.x + k * kPassThreads] = inv ? a[k] * b[k] : a[k];
        }
    }
}

// ---- FP32 operator ------------------------------------------------------------------------

// Pass 1: item = {chunk, first edge, end edge} (chunk-major, quad-aligned).  The contribution
// of an edge is v[src] * coef (weighted) or v[src] / deg(src) (unweighted: inv), stored at the
// edge's bin-major place.  Also finds the largest |v| of the sources (sinks contribute nothing).
// DERIVED: the places come from the group data (derived_place) instead of pos, and a store
// whose place is not inside the copy is dropped.
template <bool DERIVED>
__global__ __launch_bounds__(kPassThreads) void bin_scatter_kernel(
    const uint4 *items, uint32_t nitems, const float *v, const float *inv, uint32_t n,
    const uint16_t *bsrc, const float *bcoef, const uint32_t *pos, const unsigned long long *gmask,
    const uint32_t *gfirst, const uint32_t *sdelta, uint32_t nsd, uint32_t mp, float *bval,
    BinCtl *ctl) {
    constexpr uint32_t kSpan = 1u << kSpan32Shift;
    __shared__ __align__(16) float tab[kSpan + 1];
    __shared__ uint32_t slot;
    if (threadIdx.x == 0) tab[kSpan] = 0.f;  // read by the padding slots
    float vmax = 0.f;
    for (uint32_t it = next_item(&ctl->q1, &slot); it < nitems; it = next_item(&ctl->q1, &slot)) {
        const uint4 item = items[it];
        const uint32_t base = item.x << kSpan32Shift;
        fill_table<float, kSpan>(tab, v, inv, base, n, vmax);
        __syncthreads();
        // thread t handles quads q0 + t + j * kPassThreads, so every load of a wave is
        // contiguous (8-byte source offsets, 16-byte coefficients, 4-byte places or the group
        // data); kP1Quads quads per step, loaded one step ahead
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
                at[k] = DERIVED ? derived_place(gmask, gfirst, sdelta, nsd, q)
                                : __builtin_nontemporal_load(pos + q);
            }
        }
        for (uint32_t b = qt; b < q1; b += kStep) {
            u32x2 spn[kP1Quads];
            f32x4 cn[kP1Quads];
            uint32_t atn[kP1Quads];
            unsigned long long mkn[kP1Quads];
            uint32_t fn[kP1Quads];
#pragma unroll
            for (int k = 0; k < kP1Quads; ++k) {
                const uint32_t q = b + kStep + k * kPassThreads;
                if (q < q1) {
                    // group data first, so that its wait does not include the streams below
                    if (DERIVED) {
                        mkn[k] = gmask[q >> 6];
                        fn[k] = gfirst[q >> 6];
                    }
                    spn[k] = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(bsrc + q * kQuad));
                    cn[k] = bcoef ? __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(bcoef + q * kQuad)) : one;
                    if (!DERIVED) atn[k] = __builtin_nontemporal_load(pos + q);
                }
            }
#pragma unroll
            for (int k = 0; k < kP1Quads; ++k) {
                const uint32_t q = b + k * kPassThreads;
                if (q < q1 && (!DERIVED || at[k] <= mp - kQuad)) {
                    f32x4 r;
                    r.x = tab[sp[k].x & 0xFFFFu] * c[k].x;
                    r.y = tab[sp[k].x >> 16] * c[k].y;
                    r.z = tab[sp[k].y & 0xFFFFu] * c[k].z;
                    r.w = tab[sp[k].y >> 16] * c[k].w;
                    __builtin_nontemporal_store(r, reinterpret_cast<f32x4 *>(bval + at[k]));
                }
                if (DERIVED) {
                    const uint32_t qn = b + kStep + k * kPassThreads;
                    if (qn < q1) {
                        const unsigned long long below = (mkn[k] & ~1ull) << (63u - (qn & 63u));
                        const uint32_t sidx = min(fn[k] + (uint32_t)__popcll(below), nsd - 1u);
                        atn[k] = qn * kQuad + sdelta[sidx];
                    }
                }
                sp[k] = spn[k];
                c[k] = cn[k];
                at[k] = atn[k];
            }
        }
        __syncthreads();
    }
    note_vmax(ctl, (double)vmax);
}

// Adds the quads f0, f0 + 64, f0 + 128, ... below `end` (quads from the start of bval) into
// acc, kP2Quads per step with the next step's loads issued before this step's adds.
template <int kQ>
__device__ __forceinline__ void stream_quads(unsigned long long *acc, const uint16_t *bdst,
                                             const float *bval, float scale, uint32_t f0,
                                             uint32_t end) {
    if (f0 >= end) return;
    const Quad none = {{0u, 0u}, {0.f, 0.f, 0.f, 0.f}};
    Quad cur[kQ];
    bool has[kQ];
#pragma unroll
    for (int k = 0; k < kQ; ++k) {
        const uint32_t f = f0 + 64u * k;
        has[k] = f < end;
        cur[k] = has[k] ? load_quad(bdst, bval, f * kQuad) : none;
    }
    for (uint32_t f = f0 + 64u * kQ; f < end; f += 64u * kQ) {
        Quad nxt[kQ];
        bool hn[kQ];
#pragma unroll
        for (int k = 0; k < kQ; ++k) {
            const uint32_t fk = f + 64u * k;
            hn[k] = fk < end;
            nxt[k] = hn[k] ? load_quad(bdst, bval, fk * kQuad) : none;
        }
#pragma unroll
        for (int k = 0; k < kQ; ++k)
            if (has[k]) add_quad(acc, cur[k], scale);
#pragma unroll
        for (int k = 0; k < kQ; ++k) {
            cur[k] = nxt[k];
            has[k] = hn[k];
        }
    }
#pragma unroll
    for (int k = 0; k < kQ; ++k)
        if (has[k]) add_quad(acc, cur[k], scale);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// four hot entries of pass 2: packed source and destination offsets, coefficients
struct HotQuad {
    u32x4 id;
    f32x4 c;
};

__device__ __forceinline__ HotQuad load_hot(const uint32_t *hidx, const float *hcoef, uint32_t e) {
    HotQuad o;
    o.id = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(hidx + e));
    o.c = hcoef ? __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(hcoef + e))
                : f32x4{1.f, 1.f, 1.f, 1.f};
    return o;
}

// The contribution of a hot entry is computed as pass 1 computes it (table value times the
// coefficient in FP32), so the sums are bit-identical to those of the two-pass route.
__device__ __forceinline__ void add_hot(unsigned long long *acc, const float *tab, uint32_t id,
                                        float c, float scale) {
    fx_add(acc, id & (kBin - 1), tab[id >> kBinShift] * c, scale);
}

// Adds the hot quads f0, f0 + 64, ... below `end`, kQ per step, loads one step ahead.
template <int kQ>
__device__ __forceinline__ void stream_hot(unsigned long long *acc, const float *tab,
                                           const uint32_t *hidx, const float *hcoef, float scale,
                                           uint32_t f0, uint32_t end) {
    if (f0 >= end) return;
    HotQuad cur[kQ];
    bool has[kQ];
#pragma unroll
    for (int k = 0; k < kQ; ++k) {
        const uint32_t f = f0 + 64u * k;
        has[k] = f < end;
        if (has[k]) cur[k] = load_hot(hidx, hcoef, f * kQuad);
    }
    for (uint32_t f = f0 + 64u * kQ; f < end; f += 64u * kQ) {
        HotQuad nxt[kQ];
        bool hn[kQ];
#pragma unroll
        for (int k = 0; k < kQ; ++k) {
            const uint32_t fk = f + 64u * k;
            hn[k] = fk < end;
            if (hn[k]) nxt[k] = load_hot(hidx, hcoef, fk * kQuad);
        }
#pragma unroll
        for (int k = 0; k < kQ; ++k)
            if (has[k]) {
                add_hot(acc, tab, cur[k].id.x, cur[k].c.x, scale);
                add_hot(acc, tab, cur[k].id.y, cur[k].c.y, scale);
                add_hot(acc, tab, cur[k].id.z, cur[k].c.z, scale);
                add_hot(acc, tab, cur[k].id.w, cur[k].c.w, scale);
            }
#pragma unroll
        for (int k = 0; k < kQ; ++k) {
            cur[k] = nxt[k];
            has[k] = hn[k];
        }
    }
#pragma unroll
    for (int k = 0; k < kQ; ++k)
        if (has[k]) {
            add_hot(acc, tab, cur[k].id.x, cur[k].c.x, scale);
            add_hot(acc, tab, cur[k].id.y, cur[k].c.y, scale);
            add_hot(acc, tab, cur[k].id.z, cur[k].c.z, scale);
            add_hot(acc, tab, cur[k].id.w, cur[k].c.w, scale);
        }
}

// Pass 2: item = {bin, first edge, end edge, flags | slot << 2} (bin-major, quad-aligned).  Each
// wave streams an equal, contiguous share of the item's quads (64 consecutive quads per load,
// one per lane) into the bin's LDS accumulator.  A bin held by one item is written out as
// w = v - (d y + st) (minus) or w = B v = d y + st at once; a split bin's items write their
// accumulators to slots of their own.
// HOT: hitems[it] = {first, end} hot entry of the item (quad-aligned); an item with hot entries
// first stages the table tab[i] = v[i] (* inv[i]) of the `hot` hot sources (0 beyond), the
// value pass 1 would have staged, and adds them after its contributions.
template <bool HOT>
__global__ __launch_bounds__(kPassThreads) void bin_gather_kernel(
    const uint4 *items, uint32_t nitems, const uint16_t *bdst, const float *bval, long long *part,
    uint32_t n, BinCtl *ctl, double insum, const float *v, bool minus, float d,
    const double *sink_sum, float *w, const uint2 *hitems, const uint32_t *hidx,
    const float *hcoef, const float *inv, uint32_t hot) {
    __shared__ unsigned long long acc[kBin];
    __shared__ __align__(16) float htab[HOT ? kHot + 1 : 1];
    __shared__ uint32_t slot;
    const int ex = fixed_exponent(ctl, insum);
    const float scale = ldexpf(1.f, ex);
    const double unscale = ldexp(1.0, -ex);
    const float st = sink_sum ? (float)(d * *sink_sum / (double)n) : 0.f;
    constexpr uint32_t kWaves = kPassThreads / 64;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t it = next_item(&ctl->q2, &slot); it < nitems; it = next_item(&ctl->q2, &slot)) {
        const uint4 item = items[it];
        const uint2 hi = HOT ? hitems[it] : uint2{0u, 0u};
        for (uint32_t i = threadIdx.x; i < kBin; i += kPassThreads) acc[i] = 0ull;
        if (HOT && hi.y > hi.x) {
            const uint32_t lim = min(hot, n);
            for (uint32_t i = threadIdx.x; i <= kHot; i += kPassThreads) {
                float t = 0.f;
                if (i < lim) t = inv ? v[i] * inv[i] : v[i];
                htab[i] = t;
            }
        }
        __syncthreads();
        const uint32_t q0 = item.y / kQuad, total = item.z / kQuad - q0;
        const uint32_t share = ((total + kWaves - 1) / kWaves + 63) & ~63u;
        const uint32_t wb = min(total, wave * share), we = min(total, wb + share);
        stream_quads<HOT ? kP2HotQuads : kP2Quads>(acc, bdst, bval, scale, q0 + wb + lane, q0 + we);
        if (HOT && hi.y > hi.x) {
            const uint32_t h0 = hi.x / kQuad, htot = hi.y / kQuad - h0;
            const uint32_t hs = ((htot + kWaves - 1) / kWaves + 63) & ~63u;
            const uint32_t hb = min(htot, wave * hs), he = min(htot, hb + hs);
            stream_hot<kP2HotQuads>(acc, htab, hidx, hcoef, scale, h0 + hb + lane, h0 + he);
        }
        __syncthreads();
        const uint32_t base = item.x << kBinShift;
        if ((item.w & kItemSole) && minus) {
            // v of all rows of this thread loaded first
            constexpr uint32_t kRows = kBin / kPassThreads;
            float vv[kRows];
#pragma unroll
            for (uint32_t k = 0; k < kRows; ++k) {
                const uint32_t u = base + threadIdx.x + k * kPassThreads;
                vv[k] = u < n ? v[u] : 0.f;
            }
#pragma unroll
            for (uint32_t k = 0; k < kRows; ++k) {
                const uint32_t i = threadIdx.x + k * kPassThreads;
                const uint32_t u = base + i;
                if (u >= n) break;
                const float y = (float)((double)(long long)acc[i] * unscale);
                w[u] = vv[k] - (d * y + st);
            }
        } else {
            for (uint32_t i = threadIdx.x; i < kBin; i += kPassThreads) {
                const uint32_t u = base + i;
                if (u >= n) break;
                if (item.w & kItemSole) {
                    const float y = (float)((double)(long long)acc[i] * unscale);
                    w[u] = d * y + st;
                } else {
                    part[(size_t)(item.w >> 2) * kBin + i] = (long long)acc[i];
                }
            }
        }
        __syncthreads();
    }
}

// Bins split over several items: y = the items' partial accumulators added in item order, then
// w = v - (d y + st) (minus) or w = d y + st.  mb[j] = {bin, first slot, slot count, -}.
__global__ void bin_reduce_kernel(const uint4 *mb, uint32_t nmb, const long long *part, uint32_t n,
                                  const BinCtl *ctl, double insum, const float *v, bool minus,
                                  float d, const double *sink_sum, float *w) {
    const double unscale = ldexp(1.0, -fixed_exponent(ctl, insum));
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

// ---- FP64 sweep ---------------------------------------------------------------------------

// What the row epilogue of a sweep needs (the fields of SweepArgs<double> it reads).
struct Epi64 {
    const double *xold;
    double *xnew;
    double *xhat_new;        // prescaled new iterate (unweighted) or nullptr
    const double *inv;       // 1/deg (unweighted) or nullptr
    const double *base_vec;  // per-node right-hand side or nullptr
    float *r32;              // FP32 residual or nullptr
    double *partials;        // [slots * 4] (l1, l2, sum, sink)
    const PhaseCtl *pctl;
    double damp, tele;
    uint32_t n, sink_begin;
    int32_t distribute;
    int lo_shift;            // fraction bits of the low accumulator word
};

// Inputs of row u's epilogue, loaded ahead of the row (xold, the right-hand side entry and
// 1/deg where the prescaled copy is written).
struct Row64 {
    double xo, bv, iv;
};
__device__ __forceinline__ Row64 load_row64(const Epi64 &e, uint32_t u) {
    Row64 r{0.0, 0.0, 0.0};
    if (u < e.n) {
        r.xo = e.xold[u];
        if (e.base_vec) r.bv = e.base_vec[u];
        if (e.xhat_new != nullptr && u < e.sink_begin) r.iv = e.inv[u];
    }
    return r;
}

// The finished row u from its accumulated product y (the pull sweep's epilogue_store).
__device__ __forceinline__ void finish_row64(const Epi64 &e, uint32_t u, double y, double c,
                                             const Row64 &in, double (&s)[4]) {
    double x = e.damp * y + c;
    if (e.base_vec) x += in.bv;
    __builtin_nontemporal_store(x, e.xnew + u);
    if (e.xhat_new != nullptr && u < e.sink_begin) __builtin_nontemporal_store(x * in.iv, e.xhat_new + u);
    const double dif = x - in.xo;
    if (e.r32) __builtin_nontemporal_store((float)dif, e.r32 + u);
    s[0] += fabs(dif);
    s[1] += dif * dif;
    s[2] += x;
    if (u >= e.sink_begin) s[3] += x;
}

// Block sums of the four reductions into partials[slot * 4 + k] (fixed order: deterministic).
template <int kThreads>
__device__ __forceinline__ void block_partials(double (&s)[4], double *partials, uint32_t slot) {
    __shared__ double red[4][kThreads / 64];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double v = s[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0.0;
        for (int q = 0; q < kThreads / 64; ++q) v += red[threadIdx.x][q];
        partials[(size_t)slot * 4 + threadIdx.x] = v;
    }
    __syncthreads();
}

__device__ __forceinline__ double rhs_const64(const Epi64 &e) {
    double c = e.tele;
    if (e.distribute) c += e.damp * e.pctl->sink_in / (double)e.n;
    return c;
}

// A contribution c in 6 bytes: head h = float(c) and tail q = rint((c - h) * 2^(165 - E)),
// E the biased exponent of h (1 for zero and subnormal heads).  |c - h| <= ulp(h) / 2 =
// 2^(E - 151), so |q| <= 2^14; c is recovered as h + q * 2^(E - 165) to within 2^(E - 166),
// i.e. 2^-39 of |c| for a normal head.
__device__ __forceinline__ double exp2_of(int e) {
    return __longlong_as_double((long long)(1023 + e) << 52);
}
__device__ __forceinline__ int head_exponent(float h) {
    return max(1, (int)((__float_as_uint(h) >> 23) & 0xFFu));
}
__device__ __forceinline__ void pack6(double c, float &h, short &q) {
    h = (float)c;
    q = (short)__double2int_rn((c - (double)h) * exp2_of(165 - head_exponent(h)));
}
__device__ __forceinline__ double unpack6(float h, short q) {
    return (double)h + (double)q * exp2_of(head_exponent(h) - 165);
}

// four coefficients of an FP64 pass-1 quad, from FP64 coefficients or FP32 weights
__device__ __forceinline__ void load_coef4(const double *c, uint32_t e, f64x2 &c0, f64x2 &c1) {
    c0 = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(c + e));
    c1 = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(c + e + 2));
}
__device__ __forceinline__ void load_coef4(const float *c, uint32_t e, f64x2 &c0, f64x2 &c1) {
    const f32x4 w = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(c + e));
    c0 = f64x2{(double)w.x, (double)w.y};
    c1 = f64x2{(double)w.z, (double)w.w};
}

// Pass 1 (FP64): as the FP32 pass with an FP64 table of 16384 sources and FP64 contributions.
// inv scales the table (1/deg: unweighted graphs, and weighted ones whose copy keeps FP32
// weights); bcoef is the per-edge factor (FP64 coefficient or FP32 weight) or nullptr.
// COMPACT: the contributions go to 6-byte places (FP32 heads at bval as float[mp], 16-bit
// tails after them) instead of FP64 values.
// DERIVED: as in the FP32 pass 1, the places come from the group data and a store whose
// place is not inside the copy is dropped.
template <typename CT, bool COMPACT, bool DERIVED>
__global__ __launch_bounds__(kPassThreads) void bin_scatter64_kernel(
    const uint4 *items, uint32_t nitems, const double *x, const double *inv, uint32_t n,
    const uint16_t *bsrc, const CT *bcoef, const uint32_t *pos, const unsigned long long *gmask,
    const uint32_t *gfirst, const uint32_t *sdelta, uint32_t nsd, double *bval, uint64_t mp,
    BinCtl *ctl, const PhaseCtl *pctl) {
    constexpr uint32_t kSpan = 1u << kSpan64Shift;
    if (pctl->done) return;
    __shared__ __align__(16) double tab[kSpan + 1];
    __shared__ uint32_t slot;
    if (threadIdx.x == 0) tab[kSpan] = 0.0;  // read by the padding slots
    double vmax = 0.0;
    for (uint32_t it = next_item(&ctl->q1, &slot); it < nitems; it = next_item(&ctl->q1, &slot)) {
        const uint4 item = items[it];
        const uint32_t base = item.x << kSpan64Shift;
        fill_table<double, kSpan>(tab, x, inv, base, n, vmax);
        __syncthreads();
        // one quad per thread and step, the next step's loads issued before this step's stores
        const uint32_t q1 = item.z / kQuad;
        const f64x2 one = {1.0, 1.0};
        uint32_t q = item.y / kQuad + threadIdx.x;
        u32x2 sp = {kSpan | (kSpan << 16), kSpan | (kSpan << 16)};
        uint32_t at = 0;
        f64x2 c0 = one, c1 = one;
        if (q < q1) {
            sp = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(bsrc + q * kQuad));
            at = DERIVED ? derived_place(gmask, gfirst, sdelta, nsd, q)
                         : __builtin_nontemporal_load(pos + q);
            if (bcoef) load_coef4(bcoef, q * kQuad, c0, c1);
        }
        for (; q < q1; q += kPassThreads) {
            const uint32_t qn = q + kPassThreads;
            u32x2 spn = sp;
            uint32_t atn = 0, fn = 0;
            unsigned long long mkn = 0;
            f64x2 c0n = one, c1n = one;
            if (qn < q1) {
                // group data first, so that its wait does not include the streams below
                if (DERIVED) {
                    mkn = gmask[qn >> 6];
                    fn = gfirst[qn >> 6];
                }
                spn = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(bsrc + qn * kQuad));
                if (!DERIVED) atn = __builtin_nontemporal_load(pos + qn);
                if (bcoef) load_coef4(bcoef, qn * kQuad, c0n, c1n);
            }
            const bool store = !DERIVED || (uint64_t)at + kQuad <= mp;
            f64x2 r0 = {tab[sp.x & 0xFFFFu], tab[sp.x >> 16]};
            f64x2 r1 = {tab[sp.y & 0xFFFFu], tab[sp.y >> 16]};
            r0 *= c0;
            r1 *= c1;
            if constexpr (COMPACT) {
                float h0, h1, h2, h3;
                short q0, q1, q2, q3;
                pack6(r0.x, h0, q0);
                pack6(r0.y, h1, q1);
                pack6(r1.x, h2, q2);
                pack6(r1.y, h3, q3);
                const f32x4 h = {h0, h1, h2, h3};
                float *heads = reinterpret_cast<float *>(bval);
                unsigned short *tails = reinterpret_cast<unsigned short *>(heads + mp);
                const u32x2 t = {(uint32_t)(unsigned short)q0 | ((uint32_t)(unsigned short)q1 << 16),
                                 (uint32_t)(unsigned short)q2 | ((uint32_t)(unsigned short)q3 << 16)};
                if (store) {
                    __builtin_nontemporal_store(h, reinterpret_cast<f32x4 *>(heads + at));
                    __builtin_nontemporal_store(t, reinterpret_cast<u32x2 *>(tails + at));
                }
            } else if (store) {
                __builtin_nontemporal_store(r0, reinterpret_cast<f64x2 *>(bval + at));
                __builtin_nontemporal_store(r1, reinterpret_cast<f64x2 *>(bval + at + 2));
            }
            if (DERIVED && qn < q1) {
                const unsigned long long below = (mkn & ~1ull) << (63u - (qn & 63u));
                atn = qn * kQuad + sdelta[min(fn + (uint32_t)__popcll(below), nsd - 1u)];
            }
            sp = spn;
            at = atn;
            c0 = c0n;
            c1 = c1n;
        }
        __syncthreads();
    }
    note_vmax(ctl, vmax);
}

// One contribution t (already scaled by 2^e) into the two accumulator words: the whole units
// hi = trunc(t) and the fraction t - hi (exact) in units of 2^-lo_shift.
__device__ __forceinline__ void fx_add64(unsigned long long *hi, unsigned long long *lo, uint32_t d,
                                         double t, double lscale) {
    const double h = trunc(t);
    atomicAdd(&hi[d], (unsigned long long)(long long)h);
    atomicAdd(&lo[d], (unsigned long long)__double2ll_rn((t - h) * lscale));
}

// Pass 2 (FP64): the bin's contributions into 128-bit accumulators; a bin held by one item is
// finished here (the sweep epilogue, the block's reductions into partials slot w >> 2), a split
// bin's items write both words to part slots of their own.
template <bool COMPACT>
__global__ __launch_bounds__(kPassThreads) void bin_gather64_kernel(
    const uint4 *items, uint32_t nitems, const uint16_t *bdst, const double *bval, uint64_t mp,
    long long *part, BinCtl *ctl, double insum, Epi64 e) {
    if (e.pctl->done) return;
    __shared__ unsigned long long acc_hi[kBin];
    __shared__ unsigned long long acc_lo[kBin];
    __shared__ uint32_t slot;
    const int ex = fixed_exponent(ctl, insum);
    const double scale = ldexp(1.0, ex), unscale = ldexp(1.0, -ex);
    const double lscale = ldexp(1.0, e.lo_shift), lunscale = ldexp(1.0, -e.lo_shift);
    const double c = rhs_const64(e);
    const uint32_t n = e.n;
    for (uint32_t it = next_item(&ctl->q2, &slot); it < nitems; it = next_item(&ctl->q2, &slot)) {
        const uint4 item = items[it];
        for (uint32_t i = threadIdx.x; i < kBin; i += kPassThreads) acc_hi[i] = acc_lo[i] = 0ull;
        __syncthreads();
        for (uint32_t q = item.y / kQuad + threadIdx.x; q < item.z / kQuad; q += kPassThreads) {
            const u32x2 dd = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(bdst + q * kQuad));
            f64x2 a0, a1;
            if constexpr (COMPACT) {
                const float *heads = reinterpret_cast<const float *>(bval);
                const unsigned short *tails = reinterpret_cast<const unsigned short *>(heads + mp);
                const f32x4 h = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(heads + q * kQuad));
                const u32x2 t = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(tails + q * kQuad));
                a0 = f64x2{unpack6(h.x, (short)(t.x & 0xFFFFu)), unpack6(h.y, (short)(t.x >> 16))};
                a1 = f64x2{unpack6(h.z, (short)(t.y & 0xFFFFu)), unpack6(h.w, (short)(t.y >> 16))};
            } else {
                a0 = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(bval + q * kQuad));
                a1 = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(bval + q * kQuad + 2));
            }
            fx_add64(acc_hi, acc_lo, dd.x & 0xFFFFu, a0.x * scale, lscale);
            fx_add64(acc_hi, acc_lo, dd.x >> 16, a0.y * scale, lscale);
            fx_add64(acc_hi, acc_lo, dd.y & 0xFFFFu, a1.x * scale, lscale);
            fx_add64(acc_hi, acc_lo, dd.y >> 16, a1.y * scale, lscale);
        }
        __syncthreads();
        const uint32_t base = item.x << kBinShift;
        if (item.w & kItemSole) {
            // rows threadIdx.x + k * kPassThreads: all inputs loaded first, rows finished in order
            constexpr uint32_t kRows = kBin / kPassThreads;
            Row64 in[kRows];
#pragma unroll
            for (uint32_t k = 0; k < kRows; ++k) in[k] = load_row64(e, base + threadIdx.x + k * kPassThreads);
            double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (uint32_t k = 0; k < kRows; ++k) {
                const uint32_t i = threadIdx.x + k * kPassThreads;
                const uint32_t u = base + i;
                if (u >= n) break;
                const double y = ((double)(long long)acc_hi[i] + (double)(long long)acc_lo[i] * lunscale) * unscale;
                finish_row64(e, u, y, c, in[k], s);
            }
            block_partials<kPassThreads>(s, e.partials, item.w >> 2);
        } else {
            for (uint32_t i = threadIdx.x; i < kBin; i += kPassThreads) {
                if (base + i >= n) break;
                long long *p = part + ((size_t)(item.w >> 2) * kBin + i) * 2;
                p[0] = (long long)acc_hi[i];
                p[1] = (long long)acc_lo[i];
            }
        }
        __syncthreads();
    }
}

// Split bins (FP64): kReduce64Blocks blocks per bin, one row per thread, add the items' words in
// item order and finish the rows; block j of a bin writes partials slot mb.w + j.
// mb[i] = {bin, first slot, slot count, first partials slot}.
constexpr int kReduce64Threads = 1024;
constexpr uint32_t kReduce64Blocks = kBin / kReduce64Threads;
__global__ __launch_bounds__(kReduce64Threads) void bin_reduce64_kernel(
    const uint4 *mb, const long long *part, const BinCtl *ctl, double insum, Epi64 e) {
    if (e.pctl->done) return;
    const uint4 b = mb[blockIdx.x / kReduce64Blocks];
    const uint32_t j = blockIdx.x % kReduce64Blocks;
    const double unscale = ldexp(1.0, -fixed_exponent(ctl, insum));
    const double lunscale = ldexp(1.0, -e.lo_shift);
    const double c = rhs_const64(e);
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const uint32_t i = j * kReduce64Threads + threadIdx.x;
    const uint32_t u = (b.x << kBinShift) + i;
    if (u < e.n) {
        long long hi = 0, lo = 0;
        for (uint32_t q = 0; q < b.z; ++q) {
            const long long *p = part + ((size_t)(b.y + q) * kBin + i) * 2;
            hi += p[0];
            lo += p[1];
        }
        finish_row64(e, u, ((double)hi + (double)lo * lunscale) * unscale, c, load_row64(e, u), s);
    }
    block_partials<kReduce64Threads>(s, e.partials, b.w + j);
}

// ---- host side ----------------------------------------------------------------------------

// Bound of the fixed-point sums: the largest in-coefficient sum of a row (x 1.01 for the FP32
// rounding of the coefficients; FP64 products round by at most 2^-52).
void binned_insum(nkpr_graph *g, BinLayout &L) {
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    DevBuf best;
    best.alloc(8);
    NKPR_HIP(hipMemsetAsync(best.p, 0, 8, s));
    const unsigned grid = (unsigned)std::min<uint64_t>(((uint64_t)n * 64 + 255) / 256, 4096);
    if (n && L.fp64)
        hipLaunchKernelGGL(bin_insum_kernel<double>, dim3(grid), dim3(256), 0, s, g->b_rowptr.as<uint32_t>(), n,
                           g->b_src.as<int32_t>(), g->weighted ? g->b_coef64.as<double>() : nullptr,
                           g->weighted ? nullptr : g->inv64.as<double>(), best.as<unsigned long long>());
    else if (n)
        hipLaunchKernelGGL(bin_insum_kernel<float>, dim3(grid), dim3(256), 0, s, g->b_rowptr.as<uint32_t>(), n,
                           g->b_src.as<int32_t>(), g->weighted ? g->b_coef32.as<float>() : nullptr,
                           g->weighted ? nullptr : g->inv32.as<float>(), best.as<unsigned long long>());
    NKPR_HIP(hipGetLastError());
    unsigned long long hb = 0;
    read_back(s, {{&hb, best.p, 8}});
    double v;
    std::memcpy(&v, &hb, 8);
    L.insum = v * (L.fp64 ? 1.0 + 1e-9 : 1.01);
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

void emit(nkpr_graph *g, BinLayout &L, const uint32_t *row) {
    const hipStream_t s = g->stream;
    if (!L.mp) return;
    if (L.fp64 && L.w32)
        hipLaunchKernelGGL(bin_emit_w32_kernel, dim3(nb(L.mp)), dim3(256), 0, s, L.order.as<uint32_t>(),
                           L.pos.as<uint32_t>(), L.mp, L.span_shift, g->b_src.as<int32_t>(), row,
                           g->b_coef64.as<double>(), g->invw64.as<double>(), L.src.as<uint16_t>(),
                           L.dst.as<uint16_t>(), L.coef.as<float>());
    else if (L.fp64)
        hipLaunchKernelGGL(bin_emit_kernel<double>, dim3(nb(L.mp)), dim3(256), 0, s, L.order.as<uint32_t>(),
                           L.pos.as<uint32_t>(), L.mp, L.span_shift, g->b_src.as<int32_t>(), row,
                           g->weighted ? g->b_coef64.as<double>() : nullptr, L.src.as<uint16_t>(),
                           L.dst.as<uint16_t>(), g->weighted ? L.coef.as<double>() : nullptr);
    else
        hipLaunchKernelGGL(bin_emit_kernel<float>, dim3(nb(L.mp)), dim3(256), 0, s, L.order.as<uint32_t>(),
                           L.pos.as<uint32_t>(), L.mp, L.span_shift, g->b_src.as<int32_t>(), row,
                           g->weighted ? g->b_coef32.as<float>() : nullptr, L.src.as<uint16_t>(),
                           L.dst.as<uint16_t>(), g->weighted ? L.coef.as<float>() : nullptr);
    NKPR_HIP(hipGetLastError());
}

void emit_hot(nkpr_graph *g, BinLayout &L, const uint32_t *row) {
    if (!L.mh) return;
    hipLaunchKernelGGL(bin_hot_emit_kernel, dim3(nb(L.mh)), dim3(256), 0, g->stream,
                       L.horder.as<uint32_t>(), L.mh, g->b_src.as<int32_t>(), row,
                       g->weighted ? g->b_coef32.as<float>() : nullptr, L.hidx.as<uint32_t>(),
                       g->weighted ? L.hcoef.as<float>() : nullptr);
    NKPR_HIP(hipGetLastError());
}

// The number of hot sources of the FP32 copy.  The bypass is off unless NKPR_BINNED_HOT asks
// for it: a value k >= 1 takes min(k, kHot) ids, "auto" takes kHot ids when their edges are
// at least 5% of all.  Measured on R-MAT scale 26 (20% of the edges hot) in one process with
// both copies built: 3.33 ms per application with the bypass against 3.03 ms without, same
// results; the table halves pass 2's occupancy (one block per CU) and that costs more than
// the bytes it saves (16.4 instead of 18.2 GB of traffic per application).
uint32_t choose_hot(nkpr_graph *g, const BinLayout &L) {
    if (L.fp64 || g->n == 0 || g->m == 0) return 0;
    const char *e = std::getenv("NKPR_BINNED_HOT");
    if (!e || !*e) return 0;
    if (std::strcmp(e, "auto") != 0) {
        const long long k = std::atoll(e);
        return k <= 0 ? 0u : (uint32_t)std::min<long long>(k, kHot);
    }
    DevBuf cnt;
    cnt.alloc(8);
    NKPR_HIP(hipMemsetAsync(cnt.p, 0, 8, g->stream));
    hipLaunchKernelGGL(bin_count_below_kernel, dim3(nb(g->m)), dim3(256), 0, g->stream,
                       g->b_src.as<int32_t>(), g->m, kHot, cnt.as<unsigned long long>());
    NKPR_HIP(hipGetLastError());
    unsigned long long c = 0;
    read_back(g->stream, {{&c, cnt.p, 8}});
    return c * 20 >= g->m ? kHot : 0u;
}

// Builds the binned copy L (once per graph; coefficients again after new weights).  Returns
// false when the padded copy would not fit 32-bit positions.
bool build_layout(nkpr_graph *g, BinLayout &L) {
    if (L.ready && L.coef_ready) return true;
    // test hook: the allocation failure of a device that is too small for the copy
    if (const char *t = std::getenv("NKPR_TEST_BINNED_NOMEM"); t && std::atoi(t) != 0)
        throw Error(NKPR_ERR_NOMEM, "binned copy: allocation refused (NKPR_TEST_BINNED_NOMEM)");
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const int ss = L.span_shift;
    const uint32_t span = 1u << ss;
    const uint32_t nchunks = std::max<uint32_t>(1, (g->sink_begin + span - 1) / span);
    const uint32_t nbins = (n + kBin - 1) / kBin;
    const uint64_t nseg = (uint64_t)nchunks * nbins;
    DevBuf row, k1, k2, v1, v2, tmp, rseg, cseg, bseg;
    row.alloc(4 * std::max<uint64_t>(m, 1));
    if (n)
        hipLaunchKernelGGL(bin_row_kernel, dim3((unsigned)std::min<uint64_t>(((uint64_t)n * 64 + 255) / 256, 1u << 18)), dim3(256), 0, s,
                           g->b_rowptr.as<uint32_t>(), n, row.as<uint32_t>());
    if (L.ready) {
        // new coefficients, same order: re-emit from the kept order
        if (g->weighted) {
            emit(g, L, row.as<uint32_t>());
            emit_hot(g, L, row.as<uint32_t>());
        }
        NKPR_HIP(hipStreamSynchronize(s));
        binned_insum(g, L);
        L.coef_ready = true;
        return true;
    }
    // hot sources skip pass 1: their edges sort into one extra segment per bin (a chunk past
    // the last one)
    L.hot = choose_hot(g, L);
    const uint64_t nseg_all = nseg + (L.hot ? nbins : 0);
    // sort the edges by key and find the unpadded segment starts (chunk-major)
    std::vector<uint32_t> seg(nseg_all + 1, 0u);
    rseg.alloc(4 * (nseg_all + 1));
    if (m) {
        k1.alloc(8 * m);
        k2.alloc(8 * m);
        v1.alloc(4 * m);
        v2.alloc(4 * m);
        hipLaunchKernelGGL(bin_key_kernel, dim3(nb(m)), dim3(256), 0, s, g->b_src.as<int32_t>(),
                           row.as<uint32_t>(), m, nbins, ss, nchunks, L.hot, k1.as<uint64_t>(),
                           v1.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        hipcub::DoubleBuffer<uint64_t> kb(k1.as<uint64_t>(), k2.as<uint64_t>());
        hipcub::DoubleBuffer<uint32_t> vb(v1.as<uint32_t>(), v2.as<uint32_t>());
        int bits = ss + 1;
        while (bits < 64 && (1ull << (bits - ss)) < nseg_all) ++bits;
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
        hipLaunchKernelGGL(bin_seg_kernel, dim3(nb(nseg_all + 1)), dim3(256), 0, s,
                           (const uint64_t *)k1.as<uint64_t>(), m, nseg_all, ss, rseg.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        read_back(s, {{seg.data(), rseg.p, 4 * (nseg_all + 1)}});
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
            L.off = true;
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
    L.mp = mp;
    cseg.alloc(4 * (nseg + 1));
    bseg.alloc(4 * (nseg + 1));
    NKPR_HIP(hipMemcpyAsync(cseg.p, chost.data(), 4 * (nseg + 1), hipMemcpyHostToDevice, s));
    NKPR_HIP(hipMemcpyAsync(bseg.p, bhost.data(), 4 * (nseg + 1), hipMemcpyHostToDevice, s));
    L.order.alloc(4 * std::max<uint64_t>(mp, 1));
    NKPR_HIP(hipMemsetAsync(L.order.p, 0xFF, 4 * std::max<uint64_t>(mp, 1), s));
    if (m)
        hipLaunchKernelGGL(bin_place_kernel, dim3(nb(m)), dim3(256), 0, s, (const uint64_t *)k1.as<uint64_t>(),
                           (const uint32_t *)v1.as<uint32_t>(), m, ss, nseg, (const uint32_t *)rseg.as<uint32_t>(),
                           (const uint32_t *)cseg.as<uint32_t>(), L.order.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    // hot entries: per bin, sorted by source, padded to whole quads
    std::vector<uint32_t> hseg(nbins + 1, 0u);
    if (L.hot) {
        uint64_t h = 0;
        for (uint32_t b = 0; b < nbins; ++b) {
            hseg[b] = (uint32_t)h;
            h += ((uint64_t)(seg[nseg + b + 1] - seg[nseg + b]) + kQuad - 1) / kQuad * kQuad;
        }
        NKPR_CHECK(h <= kMaxPadded, "binned layout: hot entries exceed 32-bit positions");
        hseg[nbins] = (uint32_t)h;
        L.mh = h;
        DevBuf hs;
        hs.alloc(4 * (nbins + 1));
        NKPR_HIP(hipMemcpyAsync(hs.p, hseg.data(), 4 * (nbins + 1), hipMemcpyHostToDevice, s));
        L.horder.alloc(4 * std::max<uint64_t>(h, 1));
        NKPR_HIP(hipMemsetAsync(L.horder.p, 0xFF, 4 * std::max<uint64_t>(h, 1), s));
        if (m > seg[nseg])
            hipLaunchKernelGGL(bin_hot_place_kernel, dim3(nb(m - seg[nseg])), dim3(256), 0, s,
                               (const uint64_t *)k1.as<uint64_t>(), (const uint32_t *)v1.as<uint32_t>(),
                               (uint64_t)seg[nseg], m, ss, nseg, (const uint32_t *)rseg.as<uint32_t>(),
                               (const uint32_t *)hs.as<uint32_t>(), L.horder.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        L.hidx.alloc(4 * std::max<uint64_t>(h, 1) + 64);
        if (g->weighted) L.hcoef.alloc(4 * std::max<uint64_t>(h, 1) + 64);
        emit_hot(g, L, row.as<uint32_t>());
        NKPR_HIP(hipStreamSynchronize(s));  // hs goes out of scope
    }
    L.pos.alloc(4 * std::max<uint64_t>(mp / kQuad, 1));
    if (mp)
        hipLaunchKernelGGL(bin_pos_kernel, dim3((unsigned)std::min<uint64_t>((nseg * 64 + 255) / 256, 1u << 18)),
                           dim3(256), 0, s, (const uint32_t *)cseg.as<uint32_t>(),
                           (const uint32_t *)bseg.as<uint32_t>(), nseg, L.pos.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    L.derived = false;
    if (mp) {
        // group data of pass 1: about 0.1 byte per edge instead of the 1-byte place
        const uint64_t nq = mp / kQuad, ngrp = (nq + 63) / 64;
        std::vector<unsigned long long> gm(ngrp, 0ull);
        std::vector<uint32_t> gf(ngrp), sd;
        for (uint64_t q = 0; q < nseg; ++q) {
            if (!padded(q)) continue;
            const uint64_t qs = chost[q] / kQuad;
            gm[qs >> 6] |= 1ull << (qs & 63);
            sd.push_back(bhost[q] - chost[q]);
        }
        // quad 0 starts the first non-empty segment, so group 0 has bit 0 set
        uint32_t cum = 0;
        for (uint64_t i = 0; i < ngrp; ++i) {
            gf[i] = cum + (uint32_t)(gm[i] & 1ull) - 1u;
            cum += (uint32_t)__builtin_popcountll(gm[i]);
        }
        L.gmask.alloc(8 * ngrp);
        L.gfirst.alloc(4 * ngrp);
        L.sdelta.alloc(4 * sd.size());
        L.nsdelta = (uint32_t)sd.size();
        NKPR_HIP(hipMemcpyAsync(L.gmask.p, gm.data(), 8 * ngrp, hipMemcpyHostToDevice, s));
        NKPR_HIP(hipMemcpyAsync(L.gfirst.p, gf.data(), 4 * ngrp, hipMemcpyHostToDevice, s));
        NKPR_HIP(hipMemcpyAsync(L.sdelta.p, sd.data(), 4 * sd.size(), hipMemcpyHostToDevice, s));
        DevBuf chk;
        chk.alloc(4);
        NKPR_HIP(hipMemsetAsync(chk.p, 0, 4, s));
        hipLaunchKernelGGL(bin_place_check_kernel, dim3((unsigned)std::min<uint64_t>((nq + 255) / 256, 2048)),
                           dim3(256), 0, s, (const uint32_t *)L.pos.as<uint32_t>(), nq,
                           (const unsigned long long *)L.gmask.as<unsigned long long>(),
                           (const uint32_t *)L.gfirst.as<uint32_t>(), (const uint32_t *)L.sdelta.as<uint32_t>(),
                           L.nsdelta, chk.as<uint32_t>());
        NKPR_HIP(hipGetLastError());
        uint32_t bad = 1;
        read_back(s, {{&bad, chk.p, 4}});
        L.derived = bad == 0;
        if (!L.derived) {
            L.gmask.reset();
            L.gfirst.reset();
            L.sdelta.reset();
        }
        if (std::getenv("NKPR_DEBUG"))
            std::fprintf(stderr, "[nkpr] binned %s places: %llu groups, %u segments, derived %s\n",
                         L.fp64 ? "FP64" : "FP32", (unsigned long long)ngrp, L.nsdelta,
                         L.derived ? "yes" : "no (places read)");
    }
    const size_t vb = L.fp64 ? 8 : 4;
    L.w32 = L.fp64 && g->weighted && g->w32_exact;
    L.src.alloc(2 * std::max<uint64_t>(mp, 1) + 64);
    L.dst.alloc(2 * std::max<uint64_t>(mp, 1) + 64);
    L.val.alloc(vb * std::max<uint64_t>(mp, 1) + 64);
    if (g->weighted) L.coef.alloc((L.w32 ? 4 : vb) * std::max<uint64_t>(mp, 1) + 64);
    emit(g, L, row.as<uint32_t>());
    NKPR_HIP(hipStreamSynchronize(s));

    // work items (NKPR_BINNED_ITEM overrides both sizes, e.g. to exercise split bins); sizes in
    // whole quads so that every item boundary stays quad-aligned.  Items shrink on mid-size
    // graphs so that each persistent block still draws a few from the queue: about m / (2 CUs)
    // edges, a power of two within [2^16, 2^19] (pass 1) and [2^18, 2^20] (pass 2).  Measured:
    // R-MAT scale 26 (1.06e9 edges) keeps 2^19 / 2^20 (2^18 for both: 3.68 instead of 3.38 ms
    // per FP32 application); the RHG graph of 2.0e8 edges takes 2^18 / 2^18 (0.61 instead of
    // 0.66 ms).
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    const uint64_t per_block = mp / (2 * (uint64_t)std::max(cus, 1));
    auto pow2_within = [](uint64_t v, uint64_t lo, uint64_t hi) {
        uint64_t p = lo;
        while (p * 2 <= v && p * 2 <= hi) p *= 2;
        return p;
    };
    uint64_t e1max = pow2_within(per_block, 1ull << 16, kPass1Edges);
    uint64_t e2max = pow2_within(per_block, 1ull << 18, kPass2Edges);
    if (const char *ev = std::getenv("NKPR_BINNED_ITEM")) {
        const uint64_t v = std::strtoull(ev, nullptr, 10);
        if (v > 0) e1max = e2max = (v + kQuad - 1) / kQuad * kQuad;
    }
    std::vector<uint4> p1, p2, mb;  // mb: {bin, first slot, slot count, partials slot}
    std::vector<uint2> p2h;          // hot entries of each pass-2 item
    std::vector<uint64_t> cost1, cost2;
    for (uint32_t c = 0; c < nchunks; ++c) {
        const uint32_t b0 = chost[(uint64_t)c * nbins], b1 = chost[(uint64_t)(c + 1) * nbins];
        for (uint64_t e = b0; e < b1; e += e1max) {
            const uint64_t e1 = std::min<uint64_t>(e + e1max, b1);
            p1.push_back(make_uint4(c, (uint32_t)e, (uint32_t)e1, 0));
            cost1.push_back(e1 - e + span);  // edges + the LDS table fill
        }
        // the fixed-point scale needs max |v| of the hot sources too: their chunk keeps an
        // item (an empty one when all its edges are hot), whose table fill finds it
        if (c == 0 && L.hot && b0 == b1) {
            p1.push_back(make_uint4(0, b0, b0, 0));
            cost1.push_back(span);
        }
    }
    // every bin finished in one place (a sole item or the reduce kernel) has a partials slot
    uint32_t slots = 0, pslots = 0;
    for (uint32_t b = 0; b < nbins; ++b) {
        // the bin's work: its contributions, then its hot entries; pieces of at most e2max
        const uint64_t s0 = bin_start[b], s1 = bin_start[b + 1];
        const uint64_t h0 = L.hot ? hseg[b] : 0, h1 = L.hot ? hseg[b + 1] : 0;
        const uint64_t work = (s1 - s0) + (h1 - h0);
        const uint32_t pieces = (uint32_t)std::max<uint64_t>(1, (work + e2max - 1) / e2max);
        for (uint32_t i = 0; i < pieces; ++i) {
            const uint64_t a = i * e2max, z = std::min<uint64_t>(a + e2max, work);
            const uint64_t e0 = s0 + std::min(a, s1 - s0), e1 = s0 + std::min(z, s1 - s0);
            const uint64_t f0 = h0 + (a > s1 - s0 ? a - (s1 - s0) : 0);
            const uint64_t f1 = h0 + (z > s1 - s0 ? z - (s1 - s0) : 0);
            const uint32_t flags = pieces == 1 ? kItemSole | (pslots << 2) : (slots + i) << 2;
            p2.push_back(make_uint4(b, (uint32_t)e0, (uint32_t)e1, flags));
            p2h.push_back(make_uint2((uint32_t)f0, (uint32_t)f1));
            // entries + the accumulator clear and flush (+ the hot table when staged)
            cost2.push_back((e1 - e0) + (f1 - f0) + kBin + (f1 > f0 ? kHot : 0));
        }
        if (pieces > 1) {
            // every item of a split bin writes its accumulator to a slot of its own; the FP64
            // reduce kernel finishes the bin in kReduce64Blocks blocks with a partials slot each
            mb.push_back(make_uint4(b, slots, pieces, pslots));
            slots += pieces;
            pslots += L.fp64 ? kReduce64Blocks : 1;
        } else {
            ++pslots;
        }
    }
    sort_by_cost(p1, cost1);
    {
        // p2 and p2h in the same order
        std::vector<uint4> idx(p2.size());
        for (size_t i = 0; i < p2.size(); ++i) idx[i] = make_uint4((uint32_t)i, 0, 0, 0);
        sort_by_cost(idx, cost2);
        std::vector<uint4> q2(p2.size());
        std::vector<uint2> qh(p2.size());
        for (size_t i = 0; i < idx.size(); ++i) {
            q2[i] = p2[idx[i].x];
            qh[i] = p2h[idx[i].x];
        }
        p2.swap(q2);
        p2h.swap(qh);
    }
    L.p2h.alloc(sizeof(uint2) * std::max<size_t>(p2h.size(), 1));
    if (!p2h.empty()) NKPR_HIP(hipMemcpyAsync(L.p2h.p, p2h.data(), sizeof(uint2) * p2h.size(), hipMemcpyHostToDevice, s));
    L.p1.alloc(sizeof(uint4) * std::max<size_t>(p1.size(), 1));
    L.p2.alloc(sizeof(uint4) * std::max<size_t>(p2.size(), 1));
    if (!p1.empty()) NKPR_HIP(hipMemcpyAsync(L.p1.p, p1.data(), sizeof(uint4) * p1.size(), hipMemcpyHostToDevice, s));
    if (!p2.empty()) NKPR_HIP(hipMemcpyAsync(L.p2.p, p2.data(), sizeof(uint4) * p2.size(), hipMemcpyHostToDevice, s));
    L.np1 = (uint32_t)p1.size();
    L.np2 = (uint32_t)p2.size();
    L.multi = slots;
    L.nmb = (uint32_t)mb.size();
    L.nslots = pslots;
    L.mb.alloc(sizeof(uint4) * std::max<size_t>(mb.size(), 1));
    if (!mb.empty()) NKPR_HIP(hipMemcpyAsync(L.mb.p, mb.data(), sizeof(uint4) * mb.size(), hipMemcpyHostToDevice, s));
    L.part.alloc((L.fp64 ? 16 : 8) * (size_t)std::max<uint32_t>(slots, 1) * kBin);
    if (L.fp64) L.partials.alloc(32 * (size_t)std::max<uint32_t>(pslots, 1));
    binned_insum(g, L);
    L.ctl.alloc(sizeof(BinCtl));
    NKPR_HIP(hipStreamSynchronize(s));
    if (std::getenv("NKPR_DEBUG"))
        std::fprintf(stderr, "[nkpr] binned %s layout: %u chunks x %u bins, %llu padded edges, %llu hot entries (%u hot sources), %u + %u items (%u split)\n",
                     L.fp64 ? "FP64" : "FP32", nchunks, nbins, (unsigned long long)mp,
                     (unsigned long long)L.mh, L.hot, L.np1, L.np2, L.multi);
    L.ready = true;
    L.coef_ready = true;
    return true;
}

void release_layout(BinLayout &L) {
    for (DevBuf *b : {&L.order, &L.pos, &L.src, &L.dst, &L.coef, &L.val, &L.p1, &L.p2, &L.mb,
                      &L.part, &L.ctl, &L.partials, &L.gmask, &L.gfirst, &L.sdelta, &L.hidx,
                      &L.hcoef, &L.horder, &L.p2h})
        b->reset();
    L.hot = 0;
    L.mh = 0;
    L.ready = L.coef_ready = L.derived = false;
}

// Builds L if needed.  False when it cannot be built (too many edges for 32-bit positions, or
// not enough device memory for the copy and its build: about 17 + 26 bytes per edge for the
// FP32 copy, 25 + 26 for the FP64 one); the pull sweep is used instead.
bool layout_ready(nkpr_graph *g, BinLayout &L) {
    batch_layout(g);
    if (L.off) return false;
    try {
        return build_layout(g, L);
    } catch (const Error &e) {
        if (e.code != NKPR_ERR_NOMEM) throw;
        NKPR_HIP(hipStreamSynchronize(g->stream));
        release_layout(L);
        L.off = true;
        if (std::getenv("NKPR_DEBUG"))
            std::fprintf(stderr, "[nkpr] binned layout: %s; the pull sweep stays\n", e.what());
        return false;
    }
}

// persistent grids: every block that fits on the device at once
template <typename K1, typename K2>
void persistent_grids(nkpr_graph *g, BinLayout &L, K1 k1, K2 k2) {
    if (L.grid1) return;
    int cus = 0, b1 = 0, b2 = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    NKPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, k1, kPassThreads, 0));
    NKPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b2, k2, kPassThreads, 0));
    L.grid1 = (uint32_t)std::max(1, b1 * cus);
    L.grid2 = (uint32_t)std::max(1, b2 * cus);
}

// The binned copies serve graphs with at least kBinnedMinEdges edges (on R-MAT scale 26 the
// FP32 operator takes 3.2-3.6 ms per application against 8.4 ms for the pull sweep; on graphs
// whose gathers stay in the caches the pull sweep is as fast and needs no second copy).
// NKPR_BINNED=0 / 1 (FP32 operator) and NKPR_BINNED64=0 / 1 (FP64 sweep) force either way.
constexpr uint64_t kBinnedMinEdges = 1ull << 26;

bool policy(const nkpr_graph *g, const char *env) {
    if (const char *e = std::getenv(env); e && *e) return std::atoi(e) != 0;
    return g->m >= kBinnedMinEdges;
}

}  // namespace

bool binned_enabled(const nkpr_graph *g) { return !g->bn32.off && policy(g, "NKPR_BINNED"); }

bool binned_layout(nkpr_graph *g) {
    g->bn32.span_shift = kSpan32Shift;
    g->bn32.fp64 = false;
    return layout_ready(g, g->bn32);
}

// w = B v = d P v (+ d S(v)/n) through the binned passes, or w = A v = v - B v (minus).
void binned_apply(nkpr_graph *g, const float *v, float *w, double damp, bool distribute,
                  bool minus, hipEvent_t ev0, hipEvent_t ev1) {
    BinLayout &L = g->bn32;
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    if (!n) return;
    const double *sink_sum = nullptr;
    if (distribute && g->sink_begin < n) {
        launch_range_sum<float>(v, g->sink_begin, n, g->red_partials.as<double>(), g->red_out.as<double>(), s);
        sink_sum = g->red_out.as<double>();
    }
    // NKPR_BINNED_POS=1 reads the stored places instead of deriving them
    const char *pe = std::getenv("NKPR_BINNED_POS");
    const bool derived = L.derived && !(pe && *pe && std::atoi(pe) != 0);
    const auto gather = L.hot ? bin_gather_kernel<true> : bin_gather_kernel<false>;
    persistent_grids(g, L, bin_scatter_kernel<true>, gather);
    if (ev0) NKPR_HIP(hipEventRecord(ev0, s));
    BinCtl *ctl = L.ctl.as<BinCtl>();
    hipLaunchKernelGGL(bin_clear_kernel, dim3(1), dim3(64), 0, s, ctl);
    const auto scatter = derived ? bin_scatter_kernel<true> : bin_scatter_kernel<false>;
    // the sources at or past sink_begin have no out-edges: pass 1 reads them as 0, so v may hold
    // anything there (GMRES keeps other data past its head) and the scale ignores it
    hipLaunchKernelGGL(scatter, dim3(std::min<uint32_t>(std::max<uint32_t>(L.np1, 1), L.grid1)),
                       dim3(kPassThreads), 0, s, L.p1.as<uint4>(), L.np1, v,
                       g->weighted ? nullptr : g->inv32.as<float>(), g->sink_begin, L.src.as<uint16_t>(),
                       g->weighted ? L.coef.as<float>() : nullptr, L.pos.as<uint32_t>(),
                       L.gmask.as<unsigned long long>(), L.gfirst.as<uint32_t>(), L.sdelta.as<uint32_t>(),
                       L.nsdelta, (uint32_t)L.mp, L.val.as<float>(), ctl);
    hipLaunchKernelGGL(gather, dim3(std::min<uint32_t>(std::max<uint32_t>(L.np2, 1), L.grid2)),
                       dim3(kPassThreads), 0, s, L.p2.as<uint4>(), L.np2, L.dst.as<uint16_t>(),
                       L.val.as<float>(), L.part.as<long long>(), n, ctl, L.insum, v, minus, (float)damp,
                       sink_sum, w, L.p2h.as<uint2>(), L.hidx.as<uint32_t>(),
                       g->weighted ? L.hcoef.as<float>() : nullptr,
                       g->weighted ? nullptr : g->inv32.as<float>(), L.hot);
    if (L.nmb)
        hipLaunchKernelGGL(bin_reduce_kernel, dim3(nb((uint64_t)L.nmb * kBin)), dim3(256), 0, s,
                           L.mb.as<uint4>(), L.nmb, (const long long *)L.part.as<long long>(), n,
                           (const BinCtl *)ctl, L.insum, v, minus, (float)damp, sink_sum, w);
    if (ev1) NKPR_HIP(hipEventRecord(ev1, s));
    NKPR_HIP(hipGetLastError());
}

bool binned64_active(nkpr_graph *g) {
    if (g->bn64.off || !policy(g, "NKPR_BINNED64")) return false;
    g->bn64.span_shift = kSpan64Shift;
    g->bn64.fp64 = true;
    return layout_ready(g, g->bn64);
}

// One FP64 sweep x' = d P x + b through the binned passes, with the epilogue, reductions and
// stopping rule of launch_sweep<double> (the finalize kernel adds the per-bin partials in bin
// order).  Every kernel returns at once when the phase has stopped.
void binned_sweep64(nkpr_graph *g, const SweepArgs<double> &a, const FinalizeArgs &f, int kind,
                    hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    BinLayout &L = g->bn64;
    const uint32_t n = g->n;
    if (L.w32) persistent_grids(g, L, bin_scatter64_kernel<float, false, true>, bin_gather64_kernel<false>);
    else persistent_grids(g, L, bin_scatter64_kernel<double, false, true>, bin_gather64_kernel<false>);
    const char *pe = std::getenv("NKPR_BINNED_POS");
    const bool derived = L.derived && !(pe && *pe && std::atoi(pe) != 0);
    const bool basevec = kind == kSweepBaseVec || kind == kSweepBaseResid;
    const bool resid = kind == kSweepResid || kind == kSweepBaseResid;
    const bool compact = resid && a.compact64 != 0;
    Epi64 e{};
    e.xold = a.xold;
    e.xnew = a.xnew;
    e.xhat_new = a.xhat_new;
    e.inv = g->weighted ? nullptr : g->inv64.as<double>();
    e.base_vec = basevec ? a.base_vec : nullptr;
    e.r32 = resid ? a.r32 : nullptr;
    e.partials = L.partials.as<double>();
    e.pctl = a.ctl;
    e.damp = a.damp;
    e.tele = a.tele;
    e.n = n;
    e.sink_begin = a.sink_begin;
    e.distribute = a.distribute;
    // fraction bits of the low word: a row's fractions (each below 2^lo_shift in magnitude)
    // must sum below 2^63
    int deg_bits = 0;
    while (deg_bits < 40 && (1ull << deg_bits) <= (uint64_t)g->max_in_degree) ++deg_bits;
    e.lo_shift = std::min(40, 62 - deg_bits);
    if (ev0) NKPR_HIP(hipEventRecord(ev0, s));
    BinCtl *ctl = L.ctl.as<BinCtl>();
    hipLaunchKernelGGL(bin_clear_kernel, dim3(1), dim3(64), 0, s, ctl);
    if (n) {
        const dim3 g1(std::min<uint32_t>(std::max<uint32_t>(L.np1, 1), L.grid1));
        const dim3 g2(std::min<uint32_t>(std::max<uint32_t>(L.np2, 1), L.grid2));
        const double *inv = L.w32 ? (const double *)g->invw64.as<double>()
                                  : (g->weighted ? nullptr : (const double *)g->inv64.as<double>());
        auto pass1 = [&](auto kernel, const auto *coef) {
            hipLaunchKernelGGL(kernel, g1, dim3(kPassThreads), 0, s, L.p1.as<uint4>(), L.np1, a.xold, inv, n,
                               L.src.as<uint16_t>(), coef, L.pos.as<uint32_t>(),
                               (const unsigned long long *)L.gmask.as<unsigned long long>(),
                               (const uint32_t *)L.gfirst.as<uint32_t>(), (const uint32_t *)L.sdelta.as<uint32_t>(),
                               L.nsdelta, L.val.as<double>(), L.mp, ctl, (const PhaseCtl *)a.ctl);
        };
        const float *w32 = (const float *)L.coef.as<float>();
        const double *c64 = g->weighted ? (const double *)L.coef.as<double>() : nullptr;
        if (derived) {
            if (L.w32 && compact) pass1(bin_scatter64_kernel<float, true, true>, w32);
            else if (L.w32) pass1(bin_scatter64_kernel<float, false, true>, w32);
            else if (compact) pass1(bin_scatter64_kernel<double, true, true>, c64);
            else pass1(bin_scatter64_kernel<double, false, true>, c64);
        } else {
            if (L.w32 && compact) pass1(bin_scatter64_kernel<float, true, false>, w32);
            else if (L.w32) pass1(bin_scatter64_kernel<float, false, false>, w32);
            else if (compact) pass1(bin_scatter64_kernel<double, true, false>, c64);
            else pass1(bin_scatter64_kernel<double, false, false>, c64);
        }
        if (compact)
            hipLaunchKernelGGL(bin_gather64_kernel<true>, g2, dim3(kPassThreads), 0, s, L.p2.as<uint4>(), L.np2,
                               L.dst.as<uint16_t>(), L.val.as<double>(), L.mp, L.part.as<long long>(), ctl, L.insum, e);
        else
            hipLaunchKernelGGL(bin_gather64_kernel<false>, g2, dim3(kPassThreads), 0, s, L.p2.as<uint4>(), L.np2,
                               L.dst.as<uint16_t>(), L.val.as<double>(), L.mp, L.part.as<long long>(), ctl, L.insum, e);
        if (L.nmb)
            hipLaunchKernelGGL(bin_reduce64_kernel, dim3(L.nmb * kReduce64Blocks), dim3(kReduce64Threads), 0, s, L.mb.as<uint4>(),
                               (const long long *)L.part.as<long long>(), (const BinCtl *)ctl, L.insum, e);
    }
    SweepArgs<double> fa = a;
    fa.partials = L.partials.as<double>();
    FinalizeArgs ff = f;
    ff.n_split_rows = 0;
    ff.n_blocks = n ? L.nslots : 0;
    launch_finalize64(fa, ff, kind, s);
    if (ev1) NKPR_HIP(hipEventRecord(ev1, s));
    NKPR_HIP(hipGetLastError());
}

}  // namespace nkpr
