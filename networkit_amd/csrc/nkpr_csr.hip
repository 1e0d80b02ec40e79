// CSR matrix-vector products in NetworKit's CSR format, and AlgebraicPageRank on them.
//
// The matrix is a CSRGeneralMatrix<double> as NetworKit stores it
// (include/networkit/algebraic/CSRGeneralMatrix.hpp): row_idx[rows + 1] and col_idx[nnz]
// are 64-bit `index` values, values[nnz] are doubles, and a row keeps its entries in storage
// order (duplicates stay separate).  All arrays live in device memory.
//
//   nkpr_csr_spmv           CSRGeneralMatrix::operator*(const Vector&)
//                           (CSRGeneralMatrix.hpp:563-578): y[i] = sum_k values[k] x[col_idx[k]]
//   nkpr_csr_pagerank_step  one iteration of AlgebraicPageRank::run
//                           (AlgebraicPageRank.hpp:97-101): y = M x + shift and ||y - x||_2^2
//   nkpr_algebraic_pagerank AlgebraicPageRank::run (AlgebraicPageRank.hpp:94-119): the loop to
//                           ||y - x||_2 <= tol from 1/n, then y / sum(y)
//
// Kernels.  Every sum is added in an order fixed by the matrix alone (deterministic):
//   csr_short_kernel<G>   rows of at most 32 G entries: a group of G lanes (4..64, chosen from
//       the average row length) owns a row, strides over it and reduces with shuffles; at most
//       kShortGrid blocks step over row tiles.
//   long rows             the rest, cut into pieces of kPiece entries so that a few huge rows
//       (power-law matrices) spread over the whole device:
//       csr_long_scan_kernel    kLongGrid blocks count the long rows and pieces of their row
//                               ranges;
//       csr_long_offsets_kernel one block: exclusive sums of those counts;
//       csr_long_list_kernel    the same blocks write {row, first piece} of their long rows,
//                               in row order;
//       csr_piece_kernel        a block per piece (grid-stride): the piece's partial sum;
//       csr_long_fix_kernel     per long row: its pieces added in order, the shift, y, and
//                               for a step (y - x)^2.
// The step's (y - x)^2 is added per block (short kernel and fix kernel), and the block sums in
// block order by a one-block kernel into a device scalar.
//
// Scratch of a product: the long-row list (at most nnz / (32 G + 1) rows) and piece sums (at
// most nnz / kPiece + list rows), the block counts, and the step's block sums; a plain product
// and a step take it from the stream's memory pool for the duration of the call, the
// allocation-free step from a static device array (which caps the matrices it accepts).
//
// Algorithmic bytes per product (HBM-bound, random gathers of x): per entry col 8 + value 8 +
// gathered x 8; per row row_idx 8 + y 8 (+ x 8 for the step's norm).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "nkpr_internal.hpp"

// Scratch of nkpr_csr_pagerank_step_noalloc (48.1 MiB): the long-row list (2^21 rows), the
// piece sums (2^22), the block counts and offsets, and the step's block sums.  Every GPU the
// library runs on holds it until the process ends.  Host code never uses this name as an
// address: it takes the array's address per device with hipGetSymbolAddress
// (csr_noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[(1u << 21) * 8 + (1u << 22) * 8 + 65536];

namespace nkpr {
namespace {

constexpr int kShortThreads = 256;
constexpr uint32_t kShortGrid = 2048;
constexpr int kLongThreads = 1024;   // scan and list blocks
constexpr uint32_t kLongGrid = 2048;
constexpr int kPieceThreads = 256;
constexpr uint32_t kPiece = 4096;    // entries per piece of a long row
constexpr uint32_t kPieceGrid = 4096;
constexpr int kFixThreads = 256;
constexpr uint32_t kFixGrid = 256;
constexpr int kFinalThreads = 256;
constexpr uint32_t kNoallocList = 1u << 21, kNoallocPieces = 1u << 22;

struct CsrView {
    uint64_t rows;
    const uint64_t *row_idx;
    const uint64_t *col_idx;
    const double *val;
};

// The scratch of one product, carved from one buffer.
struct LongWork {
    uint32_t *bcount;   // [kLongGrid] long rows per scan block
    uint32_t *bpieces;  // [kLongGrid] pieces per scan block
    uint32_t *boff;     // [kLongGrid + 1] first list slot per scan block
    uint32_t *bpoff;    // [kLongGrid + 1] first piece per scan block
    uint32_t *total;    // [2] long rows, pieces
    uint2 *list;        // [list cap] {row, first piece}
    double *piece;      // [piece cap]
    double *partials;   // [kShortGrid + kFixGrid] the step's block sums
};

size_t work_bytes(uint64_t list_cap, uint64_t piece_cap) {
    return 8 * list_cap + 8 * piece_cap + 4 * (4 * (size_t)kLongGrid + 4) +
           8 * (size_t)(kShortGrid + kFixGrid) + 256;
}

LongWork carve(void *base, uint64_t list_cap, uint64_t piece_cap) {
    char *p = static_cast<char *>(base);
    LongWork w;
    w.list = reinterpret_cast<uint2 *>(p);
    p += 8 * list_cap;
    w.piece = reinterpret_cast<double *>(p);
    p += 8 * piece_cap;
    w.partials = reinterpret_cast<double *>(p);
    p += 8 * (size_t)(kShortGrid + kFixGrid);
    uint32_t *u = reinterpret_cast<uint32_t *>(p);
    w.bcount = u;
    w.bpieces = u + kLongGrid;
    w.boff = u + 2 * kLongGrid;
    w.bpoff = u + 3 * kLongGrid + 1;
    w.total = u + 4 * kLongGrid + 2;
    return w;
}

__device__ __forceinline__ uint32_t pieces_of(uint64_t len) {
    return (uint32_t)((len + kPiece - 1) / kPiece);
}

// Sum of the products of row entries [b, e) taken by one lane in steps of `stride`; two
// accumulators keep two gathers in flight.
__device__ __forceinline__ double row_part(const CsrView &A, const double *__restrict__ x,
                                           uint64_t b, uint64_t e, uint32_t stride) {
    double s0 = 0.0, s1 = 0.0;
    uint64_t k = b;
    for (; k + stride < e; k += 2 * (uint64_t)stride) {
        const uint64_t c0 = A.col_idx[k], c1 = A.col_idx[k + stride];
        const double v0 = A.val[k], v1 = A.val[k + stride];
        s0 += v0 * x[c0];
        s1 += v1 * x[c1];
    }
    if (k < e) s0 += A.val[k] * x[A.col_idx[k]];
    return s0 + s1;
}

// Block sum (fixed order: shuffles inside each wave, then thread 0 adds the waves); the
// result is valid in thread 0.
template <int kThreads>
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double wsum[kThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    double tot = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) tot += wsum[w];
    __syncthreads();
    return tot;
}

template <int G, bool STEP>
__global__ __launch_bounds__(kShortThreads) void csr_short_kernel(CsrView A,
                                                                  const double *__restrict__ x,
                                                                  double shift, double *y,
                                                                  double *partials,
                                                                  uint64_t tiles) {
    constexpr uint32_t groups = kShortThreads / G;
    constexpr uint64_t kLong = 32u * G;
    const uint32_t lane = threadIdx.x % G;
    const uint32_t group = threadIdx.x / G;
    double sq = 0.0;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t r = t * groups + group;
        double s = 0.0;
        bool own = false;
        if (r < A.rows) {
            const uint64_t b = A.row_idx[r], e = A.row_idx[r + 1];
            own = e - b <= kLong;
            if (own) s = row_part(A, x, b + lane, e, G);
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
        if (own && lane == 0) {
            const double v = s + shift;
            y[r] = v;
            if (STEP) {
                const double d = v - x[r];
                sq += d * d;
            }
        }
    }
    if (STEP) {
        const double tot = block_sum<kShortThreads>(sq);
        if (threadIdx.x == 0) partials[blockIdx.x] = tot;
    }
}

// Long rows of block b's row range [b rpb, (b + 1) rpb): count and pieces.
__global__ __launch_bounds__(kLongThreads) void csr_long_scan_kernel(CsrView A, uint64_t rpb,
                                                                     uint64_t lim, LongWork w) {
    const uint64_t r0 = (uint64_t)blockIdx.x * rpb;
    const uint64_t r1 = std::min<uint64_t>(A.rows, r0 + rpb);
    uint32_t cnt = 0, pcs = 0;
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += kLongThreads) {
        const uint64_t len = A.row_idx[r + 1] - A.row_idx[r];
        if (len > lim) {
            ++cnt;
            pcs += pieces_of(len);
        }
    }
    __shared__ uint32_t sc[kLongThreads / 64], sp[kLongThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o, 64);
        pcs += __shfl_xor(pcs, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sc[threadIdx.x >> 6] = cnt;
        sp[threadIdx.x >> 6] = pcs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, q = 0;
        for (int i = 0; i < kLongThreads / 64; ++i) {
            c += sc[i];
            q += sp[i];
        }
        w.bcount[blockIdx.x] = c;
        w.bpieces[blockIdx.x] = q;
    }
}

// One block: exclusive sums of the scan blocks' counts; total = {long rows, pieces}.
__global__ void csr_long_offsets_kernel(LongWork w, uint32_t nblocks) {
    if (threadIdx.x != 0) return;
    uint32_t c = 0, q = 0;
    for (uint32_t b = 0; b < nblocks; ++b) {
        w.boff[b] = c;
        w.bpoff[b] = q;
        c += w.bcount[b];
        q += w.bpieces[b];
    }
    w.boff[nblocks] = c;
    w.bpoff[nblocks] = q;
    w.total[0] = c;
    w.total[1] = q;
}

// The long rows of block b's range in row order: list[boff[b] + i] = {row, first piece}.
__global__ __launch_bounds__(kLongThreads) void csr_long_list_kernel(CsrView A, uint64_t rpb,
                                                                     uint64_t lim, LongWork w) {
    __shared__ uint32_t wc[kLongThreads / 64], wp[kLongThreads / 64];
    __shared__ uint32_t base_c, base_p;
    const uint32_t tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const uint64_t r0 = (uint64_t)blockIdx.x * rpb;
    const uint64_t r1 = std::min<uint64_t>(A.rows, r0 + rpb);
    if (tid == 0) {
        base_c = w.boff[blockIdx.x];
        base_p = w.bpoff[blockIdx.x];
    }
    __syncthreads();
    for (uint64_t base = r0; base < r1; base += kLongThreads) {
        const uint64_t r = base + tid;
        uint32_t np = 0;
        if (r < r1) {
            const uint64_t len = A.row_idx[r + 1] - A.row_idx[r];
            if (len > lim) np = pieces_of(len);
        }
        // in-wave inclusive prefix sums of the long-row count and of the pieces
        uint32_t ic = np ? 1u : 0u, ip = np;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t tc = __shfl_up(ic, o, 64), tp = __shfl_up(ip, o, 64);
            if (lane >= (uint32_t)o) {
                ic += tc;
                ip += tp;
            }
        }
        if (lane == 63) {
            wc[wave] = ic;
            wp[wave] = ip;
        }
        __syncthreads();
        uint32_t pc = base_c, pp = base_p;
        for (uint32_t k = 0; k < wave; ++k) {
            pc += wc[k];
            pp += wp[k];
        }
        if (np) w.list[pc + ic - 1] = make_uint2((uint32_t)r, pp + ip - np);
        __syncthreads();
        if (tid == 0) {
            for (int k = 0; k < kLongThreads / 64; ++k) {
                base_c += wc[k];
                base_p += wp[k];
            }
        }
        __syncthreads();
    }
}

// Partial sum of every piece (a block per piece, grid-stride over total[1] pieces).
__global__ __launch_bounds__(kPieceThreads) void csr_piece_kernel(CsrView A,
                                                                  const double *__restrict__ x,
                                                                  LongWork w) {
    const uint32_t nrows = w.total[0], npieces = w.total[1];
    for (uint32_t p = blockIdx.x; p < npieces; p += gridDim.x) {
        // the long row holding piece p: the last list entry whose first piece is <= p
        uint32_t lo = 0, hi = nrows;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (w.list[mid].y <= p) lo = mid;
            else hi = mid;
        }
        const uint2 e = w.list[lo];
        const uint64_t rb = A.row_idx[e.x], re = A.row_idx[e.x + 1];
        const uint64_t b = rb + (uint64_t)(p - e.y) * kPiece;
        const uint64_t end = std::min<uint64_t>(re, b + kPiece);
        const double s = block_sum<kPieceThreads>(row_part(A, x, b + threadIdx.x, end, kPieceThreads));
        if (threadIdx.x == 0) w.piece[p] = s;
    }
}

// Every long row from its pieces in order; for a step, (y - x)^2 per block into
// partials[first + block].
template <bool STEP>
__global__ __launch_bounds__(kFixThreads) void csr_long_fix_kernel(CsrView A,
                                                                   const double *__restrict__ x,
                                                                   double shift, double *y,
                                                                   LongWork w, uint32_t first) {
    const uint32_t nrows = w.total[0], npieces = w.total[1];
    double sq = 0.0;
    for (uint32_t i = blockIdx.x * kFixThreads + threadIdx.x; i < nrows;
         i += gridDim.x * kFixThreads) {
        const uint2 e = w.list[i];
        const uint32_t pend = i + 1 < nrows ? w.list[i + 1].y : npieces;
        double s = 0.0;
        for (uint32_t p = e.y; p < pend; ++p) s += w.piece[p];
        const double v = s + shift;
        y[e.x] = v;
        if (STEP) {
            const double d = v - x[e.x];
            sq += d * d;
        }
    }
    if (STEP) {
        const double tot = block_sum<kFixThreads>(sq);
        if (threadIdx.x == 0) w.partials[first + blockIdx.x] = tot;
    }
}

// out[0] = sum of partials[0, count) in a fixed order.
__global__ __launch_bounds__(kFinalThreads) void csr_final_kernel(const double *partials,
                                                                  uint32_t count, double *out) {
    __shared__ double part[kFinalThreads];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < count; i += kFinalThreads) s += partials[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = kFinalThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

// Block partials of sum(x) over fixed row tiles (deterministic).
__global__ __launch_bounds__(kFinalThreads) void csr_sum_kernel(const double *x, uint64_t n,
                                                                double *partials) {
    __shared__ double part[kFinalThreads];
    double s = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kFinalThreads + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kFinalThreads)
        s += x[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = kFinalThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = part[0];
}

__global__ void csr_fill_kernel(double *x, uint64_t n, double v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

__global__ void csr_scale_kernel(double *x, uint64_t n, const double *sum) {
    const double s = sum[0];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        x[i] /= s;
}

// Structure check of a matrix before a run: row_idx non-decreasing from 0, every column
// index below cols.  flag[0] is set when something is out of place.
__global__ void csr_check_kernel(CsrView A, uint64_t cols, uint64_t nnz, uint32_t *flag) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    bool bad = false;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < A.rows; i += stride)
        bad |= A.row_idx[i] > A.row_idx[i + 1] || A.row_idx[i + 1] > nnz;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += stride)
        bad |= A.col_idx[k] >= cols;
    if (bad) atomicOr(flag, 1u);
}

// Lanes per row from the average row length.
uint32_t lanes_for(uint64_t rows, uint64_t nnz) {
    const double avg = rows ? (double)nnz / (double)rows : 0.0;
    if (avg <= 6.0) return 4;
    if (avg <= 12.0) return 8;
    if (avg <= 24.0) return 16;
    if (avg <= 48.0) return 32;
    return 64;
}

uint32_t short_grid(uint64_t rows, uint32_t lanes) {
    const uint64_t tiles = (rows + kShortThreads / lanes - 1) / (kShortThreads / lanes);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, kShortGrid));
}

uint32_t long_grid(uint64_t rows) {
    const uint64_t chunks = (rows + kLongThreads - 1) / kLongThreads;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(chunks, kLongGrid));
}

// Upper bounds of the long-row list and of the pieces of a matrix with this shape.
void long_caps(uint64_t rows, uint64_t nnz, uint64_t &list_cap, uint64_t &piece_cap) {
    const uint64_t lim = 32ull * lanes_for(rows, nnz);
    list_cap = std::min<uint64_t>(rows, nnz / (lim + 1)) + 1;
    piece_cap = nnz / kPiece + list_cap + 1;
}

CsrView view_of(const nkpr_csr *a) {
    NKPR_CHECK(a != nullptr, "matrix is NULL");
    NKPR_CHECK(a->rows == 0 || (a->row_idx && a->col_idx && a->values),
               "matrix arrays must not be NULL");
    NKPR_CHECK(a->rows < 0xFFFFFFFFull, "more than 2^32 - 2 rows");
    CsrView v;
    v.rows = a->rows;
    v.row_idx = a->row_idx;
    v.col_idx = a->col_idx;
    v.val = a->values;
    return v;
}

// The product kernels (and, for a step, the final sum of their block sums into *sq_out).
template <bool STEP>
void launch_product(const CsrView &A, uint64_t nnz, const double *x, double shift, double *y,
                    const LongWork &w, double *sq_out, hipStream_t s) {
    if (A.rows == 0) {
        if (STEP) hipLaunchKernelGGL(csr_fill_kernel, dim3(1), dim3(64), 0, s, sq_out, 1, 0.0);
        return;
    }
    const uint32_t lanes = lanes_for(A.rows, nnz);
    const uint32_t gs = short_grid(A.rows, lanes);
    const uint64_t tiles = (A.rows + kShortThreads / lanes - 1) / (kShortThreads / lanes);
    switch (lanes) {
    case 4: hipLaunchKernelGGL((csr_short_kernel<4, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, w.partials, tiles); break;
    case 8: hipLaunchKernelGGL((csr_short_kernel<8, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, w.partials, tiles); break;
    case 16: hipLaunchKernelGGL((csr_short_kernel<16, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, w.partials, tiles); break;
    case 32: hipLaunchKernelGGL((csr_short_kernel<32, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, w.partials, tiles); break;
    default: hipLaunchKernelGGL((csr_short_kernel<64, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, w.partials, tiles); break;
    }
    NKPR_HIP(hipGetLastError());
    const uint64_t lim = 32ull * lanes;
    const uint32_t gl = long_grid(A.rows);
    const uint64_t rpb = (A.rows + gl - 1) / gl;
    hipLaunchKernelGGL(csr_long_scan_kernel, dim3(gl), dim3(kLongThreads), 0, s, A, rpb, lim, w);
    hipLaunchKernelGGL(csr_long_offsets_kernel, dim3(1), dim3(64), 0, s, w, gl);
    hipLaunchKernelGGL(csr_long_list_kernel, dim3(gl), dim3(kLongThreads), 0, s, A, rpb, lim, w);
    hipLaunchKernelGGL(csr_piece_kernel, dim3(kPieceGrid), dim3(kPieceThreads), 0, s, A, x, w);
    hipLaunchKernelGGL((csr_long_fix_kernel<STEP>), dim3(kFixGrid), dim3(kFixThreads), 0, s, A, x,
                       shift, y, w, gs);
    NKPR_HIP(hipGetLastError());
    if (STEP) {
        hipLaunchKernelGGL(csr_final_kernel, dim3(1), dim3(kFinalThreads), 0, s, w.partials,
                           gs + kFixGrid, sq_out);
        NKPR_HIP(hipGetLastError());
    }
}

// Device scratch of one call, taken from the stream's memory pool.
struct PoolBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    PoolBuf(size_t bytes, hipStream_t st) : s(st) {
        const hipError_t e = hipMallocAsync(&p, bytes, st);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(NKPR_ERR_NOMEM, std::string("hipMallocAsync failed: ") + hipGetErrorString(e));
        }
    }
    ~PoolBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
};

// The address of `scratch` on the current device, taken once per device (each GPU holds its
// own copy of the array at its own address).
void *csr_noalloc_scratch() {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    int dev = -1;
    NKPR_HIP(hipGetDevice(&dev));
    NKPR_CHECK(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        NKPR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        NKPR_CHECK(p != nullptr, "hipGetSymbolAddress returned NULL");
        addr[dev] = p;
    }
    return addr[dev];
}

}  // namespace

void csr_spmv_impl(const nkpr_csr *a, const double *x, double *y, hipStream_t s) {
    const CsrView A = view_of(a);
    NKPR_CHECK(A.rows == 0 || (x && y), "x and y must not be NULL");
    NKPR_CHECK(x != y || A.rows == 0, "x and y must be different vectors");
    if (A.rows == 0) return;
    uint64_t lc, pc;
    long_caps(A.rows, a->nnz, lc, pc);
    PoolBuf buf(work_bytes(lc, pc), s);
    launch_product<false>(A, a->nnz, x, 0.0, y, carve(buf.p, lc, pc), nullptr, s);
}

void csr_step_impl(const nkpr_csr *m, double shift, const double *x, double *y, double *step_sq,
                   hipStream_t s) {
    const CsrView A = view_of(m);
    NKPR_CHECK(m->rows == m->cols, "a PageRank step needs a square matrix");
    NKPR_CHECK(step_sq != nullptr, "step_sq must not be NULL");
    NKPR_CHECK(A.rows == 0 || (x && y), "x and y must not be NULL");
    NKPR_CHECK(x != y || A.rows == 0, "x and y must be different vectors");
    uint64_t lc, pc;
    long_caps(A.rows, m->nnz, lc, pc);
    PoolBuf buf(work_bytes(lc, pc), s);
    launch_product<true>(A, m->nnz, x, shift, y, carve(buf.p, lc, pc), step_sq, s);
}

void csr_step_noalloc_impl(const nkpr_csr *m, double shift, const double *x, double *y,
                           double *step_sq, hipStream_t s) {
    static_assert(sizeof(scratch) >= 8 * (size_t)kNoallocList + 8 * (size_t)kNoallocPieces +
                                          4 * (4 * (size_t)kLongGrid + 4) +
                                          8 * (size_t)(kShortGrid + kFixGrid) + 256,
                  "the scratch holds the capped work");
    const CsrView A = view_of(m);
    NKPR_CHECK(m->rows == m->cols, "a PageRank step needs a square matrix");
    NKPR_CHECK(step_sq != nullptr, "step_sq must not be NULL");
    NKPR_CHECK(A.rows == 0 || (x && y), "x and y must not be NULL");
    NKPR_CHECK(x != y || A.rows == 0, "x and y must be different vectors");
    uint64_t lc, pc;
    long_caps(A.rows, m->nnz, lc, pc);
    NKPR_CHECK(lc <= kNoallocList && pc <= kNoallocPieces,
               "matrix too large for the allocation-free step: it could have more than 2^21 rows "
               "longer than 32 x the lanes per row, or 2^22 pieces of 4096 entries (at most "
               "nnz / (32 lanes + 1) such rows: up to about 2.7e8 entries at an average row "
               "length of 6 or less, 1.07e9 at 24 or less, 4.3e9 above 48)");
    launch_product<true>(A, m->nnz, x, shift, y, carve(csr_noalloc_scratch(), kNoallocList,
                                                       kNoallocPieces), step_sq, s);
}

void csr_validate(const nkpr_csr *m, hipStream_t s) {
    const CsrView A = view_of(m);
    DevBuf flag;
    flag.alloc(4);
    NKPR_HIP(hipMemsetAsync(flag.p, 0, 4, s));
    hipLaunchKernelGGL(csr_check_kernel, dim3(2048), dim3(256), 0, s, A, m->cols, m->nnz,
                       flag.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    uint32_t bad = 0;
    uint64_t first = 0, last = 0;
    read_back(s, {{&bad, flag.p, 4}, {&first, A.row_idx, 8}, {&last, A.row_idx + A.rows, 8}});
    NKPR_CHECK(bad == 0 && first == 0 && last == m->nnz,
               "malformed CSR matrix (row_idx must rise from 0 to nnz, column indices < cols)");
}

void algebraic_pagerank_impl(const nkpr_csr *m, double damp, double tol, uint64_t max_iterations,
                             double *scores, int32_t scores_on_device, uint64_t *iterations) {
    const CsrView A = view_of(m);
    NKPR_CHECK(m->rows == m->cols, "AlgebraicPageRank needs a square matrix");
    NKPR_CHECK(m->rows > 0, "AlgebraicPageRank needs at least one row");
    NKPR_CHECK(scores != nullptr, "scores must not be NULL");
    NKPR_CHECK(damp >= 0.0 && damp <= 1.0, "damp must lie in [0, 1]");
    NKPR_CHECK(tol >= 0.0, "tol must be >= 0");
    NKPR_CHECK(max_iterations >= 1, "max_iterations must be >= 1");
    const uint64_t n = A.rows, nnz = m->nnz;
    uint64_t lc, pc;
    long_caps(n, nnz, lc, pc);
    // buffers first: the stream owner below is destroyed (and synchronised) before them
    DevBuf xa, xb, work, out, flag;
    xa.alloc(8 * n);
    xb.alloc(8 * n);
    work.alloc(std::max<size_t>(work_bytes(lc, pc), 8 * 1024));
    out.alloc(16);
    flag.alloc(4);
    hipStream_t s;
    NKPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamOwner {
        hipStream_t s;
        ~StreamOwner() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } owner{s};
    // the reference reads whatever the indices point at; here a malformed matrix is an error
    hipLaunchKernelGGL(csr_fill_kernel, dim3(1), dim3(64), 0, s, out.as<double>(), 2, 0.0);
    NKPR_HIP(hipMemsetAsync(flag.p, 0, 4, s));
    hipLaunchKernelGGL(csr_check_kernel, dim3(2048), dim3(256), 0, s, A, m->cols, nnz,
                       flag.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    uint32_t bad = 0;
    uint64_t last = 0;
    read_back(s, {{&bad, flag.p, 4}, {&last, A.row_idx + n, 8}});
    NKPR_CHECK(bad == 0 && last == nnz,
               "malformed CSR matrix (row_idx must rise from 0 to nnz, column indices < cols)");
    const LongWork w = carve(work.p, lc, pc);
    const double tele = (1.0 - damp) / (double)n;
    hipLaunchKernelGGL(csr_fill_kernel, dim3(1024), dim3(256), 0, s, xa.as<double>(), n,
                       1.0 / (double)n);
    NKPR_HIP(hipGetLastError());
    double *cur = xa.as<double>(), *nxt = xb.as<double>();
    uint64_t it = 0;
    double sq = 0.0;
    do {
        launch_product<true>(A, nnz, cur, tele, nxt, w, out.as<double>(), s);
        read_back(s, {{&sq, out.p, 8}});
        std::swap(cur, nxt);
        ++it;
    } while (std::sqrt(sq) > tol && it < max_iterations);
    hipLaunchKernelGGL(csr_sum_kernel, dim3(1024), dim3(kFinalThreads), 0, s, cur, n, w.partials);
    hipLaunchKernelGGL(csr_final_kernel, dim3(1), dim3(kFinalThreads), 0, s, w.partials, 1024u,
                       out.as<double>());
    hipLaunchKernelGGL(csr_scale_kernel, dim3(1024), dim3(256), 0, s, cur, n, out.as<double>());
    NKPR_HIP(hipGetLastError());
    if (scores_on_device) {
        NKPR_HIP(hipMemcpyAsync(scores, cur, 8 * n, hipMemcpyDeviceToDevice, s));
        NKPR_HIP(hipStreamSynchronize(s));
    } else {
        read_back(s, {{scores, cur, 8 * n}});
    }
    if (iterations) *iterations = it;
}

}  // namespace nkpr
