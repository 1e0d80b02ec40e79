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
// Kernels.  Rows are processed in two kernels, both of them deterministic (every sum is added
// in an order fixed by the matrix alone):
//   csr_short_kernel<G>  rows with at most kLongRow entries; a group of G lanes (4..64, chosen
//       from the average row length) owns a row, strides over its entries and reduces with
//       shuffles.  The grid is capped at kShortGrid blocks that step over row tiles.
//   csr_long_kernel      rows with more entries; each of at most kLongGrid blocks of 1024
//       threads scans a contiguous row range for them and reduces each with the whole block.
// The step variant fuses the shift and (y - x)^2: each block writes one FP64 partial, and a
// one-block kernel adds the partials in block order.  Its scratch is therefore at most
// kShortGrid + kLongGrid doubles whatever the size of the matrix.
//
// Algorithmic bytes per product (HBM-bound, random gathers of x): per entry col 8 + value 8 +
// gathered x 8; per row row_idx 8 + y 8 (+ x 8 for the step's norm).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "nkpr_internal.hpp"

#include <mutex>

// Scratch of nkpr_csr_pagerank_step_noalloc: the step's per-block partials (at most
// kShortGrid + kLongGrid = 4096 doubles, 32 KiB) in a device array that every GPU the library
// runs on holds until the process ends.  Host code never uses this name as an address: it
// takes the array's address per device with hipGetSymbolAddress (csr_noalloc_scratch).
static __device__ __align__(256) double scratch[4096];

namespace nkpr {
namespace {

constexpr uint32_t kLongRow = 2048;
constexpr int kShortThreads = 256;
constexpr uint32_t kShortGrid = 2048;
constexpr int kLongThreads = 1024;
constexpr uint32_t kLongGrid = 2048;
constexpr int kFinalThreads = 256;

struct CsrView {
    uint64_t rows;
    const uint64_t *row_idx;
    const uint64_t *col_idx;
    const double *val;
};

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

template <int G, bool STEP>
__global__ __launch_bounds__(kShortThreads) void csr_short_kernel(CsrView A,
                                                                  const double *__restrict__ x,
                                                                  double shift, double *y,
                                                                  double *partials,
                                                                  uint64_t tiles) {
    constexpr uint32_t groups = kShortThreads / G;
    const uint32_t lane = threadIdx.x % G;
    const uint32_t group = threadIdx.x / G;
    double sq = 0.0;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t r = t * groups + group;
        double s = 0.0;
        bool own = false;
        if (r < A.rows) {
            const uint64_t b = A.row_idx[r], e = A.row_idx[r + 1];
            own = e - b <= kLongRow;
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
        // block sum in a fixed order: within each wave by shuffles, then wave 0 adds the waves
        __shared__ double wsum[kShortThreads / 64];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
        if (threadIdx.x % 64 == 0) wsum[threadIdx.x / 64] = sq;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot = 0.0;
            for (int w = 0; w < kShortThreads / 64; ++w) tot += wsum[w];
            partials[blockIdx.x] = tot;
        }
    }
}

template <bool STEP>
__global__ __launch_bounds__(kLongThreads) void csr_long_kernel(CsrView A,
                                                                const double *__restrict__ x,
                                                                double shift, double *y,
                                                                double *partials,
                                                                uint64_t rows_per_block) {
    __shared__ uint64_t list[kLongThreads];
    __shared__ uint32_t wcount[kLongThreads / 64];
    __shared__ double wsum[kLongThreads / 64];
    __shared__ uint32_t nlist;
    const uint32_t tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per_block;
    const uint64_t r1 = std::min<uint64_t>(A.rows, r0 + rows_per_block);
    double sq = 0.0;  // thread 0 only
    for (uint64_t base = r0; base < r1; base += kLongThreads) {
        // the long rows of this chunk, in row order
        const uint64_t r = base + tid;
        const bool is_long = r < r1 && A.row_idx[r + 1] - A.row_idx[r] > kLongRow;
        const uint64_t mask = __ballot(is_long);
        if (lane == 0) wcount[wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = 0;
            for (int w = 0; w < kLongThreads / 64; ++w) {
                const uint32_t c = wcount[w];
                wcount[w] = acc;
                acc += c;
            }
            nlist = acc;
        }
        __syncthreads();
        if (is_long) {
            const uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
            list[wcount[wave] + (uint32_t)__popcll(below)] = r;
        }
        __syncthreads();
        const uint32_t cnt = nlist;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint64_t row = list[i];
            double s = row_part(A, x, A.row_idx[row] + tid, A.row_idx[row + 1], kLongThreads);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) wsum[wave] = s;
            __syncthreads();
            if (tid == 0) {
                double tot = 0.0;
                for (int w = 0; w < kLongThreads / 64; ++w) tot += wsum[w];
                const double v = tot + shift;
                y[row] = v;
                if (STEP) {
                    const double d = v - x[row];
                    sq += d * d;
                }
            }
            __syncthreads();
        }
    }
    if (STEP && tid == 0) partials[blockIdx.x] = sq;
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
    // at least one 1024-row chunk per block
    const uint64_t chunks = (rows + kLongThreads - 1) / kLongThreads;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(chunks, kLongGrid));
}

CsrView view_of(const nkpr_csr *a) {
    NKPR_CHECK(a != nullptr, "matrix is NULL");
    NKPR_CHECK(a->rows == 0 || (a->row_idx && a->col_idx && a->values),
               "matrix arrays must not be NULL");
    CsrView v;
    v.rows = a->rows;
    v.row_idx = a->row_idx;
    v.col_idx = a->col_idx;
    v.val = a->values;
    return v;
}

// The two product kernels (and, for a step, the final sum of their partials into *sq_out).
template <bool STEP>
void launch_product(const CsrView &A, uint64_t nnz_hint, const double *x, double shift, double *y,
                    double *partials, double *sq_out, hipStream_t s) {
    if (A.rows == 0) {
        if (STEP) hipLaunchKernelGGL(csr_fill_kernel, dim3(1), dim3(64), 0, s, sq_out, 1, 0.0);
        return;
    }
    const uint32_t lanes = lanes_for(A.rows, nnz_hint);
    const uint32_t gs = short_grid(A.rows, lanes);
    const uint64_t tiles = (A.rows + kShortThreads / lanes - 1) / (kShortThreads / lanes);
    switch (lanes) {
    case 4: hipLaunchKernelGGL((csr_short_kernel<4, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, partials, tiles); break;
    case 8: hipLaunchKernelGGL((csr_short_kernel<8, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, partials, tiles); break;
    case 16: hipLaunchKernelGGL((csr_short_kernel<16, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, partials, tiles); break;
    case 32: hipLaunchKernelGGL((csr_short_kernel<32, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, partials, tiles); break;
    default: hipLaunchKernelGGL((csr_short_kernel<64, STEP>), dim3(gs), dim3(kShortThreads), 0, s, A, x, shift, y, partials, tiles); break;
    }
    NKPR_HIP(hipGetLastError());
    const uint32_t gl = long_grid(A.rows);
    const uint64_t rpb = (A.rows + gl - 1) / gl;
    hipLaunchKernelGGL((csr_long_kernel<STEP>), dim3(gl), dim3(kLongThreads), 0, s, A, x, shift,
                       y, STEP ? partials + gs : nullptr, rpb);
    NKPR_HIP(hipGetLastError());
    if (STEP) {
        hipLaunchKernelGGL(csr_final_kernel, dim3(1), dim3(kFinalThreads), 0, s, partials,
                           gs + gl, sq_out);
        NKPR_HIP(hipGetLastError());
    }
}

// Device scratch of the step's partials, taken from the stream's memory pool for one call.
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

}  // namespace

uint32_t csr_step_partials(uint64_t rows, uint64_t nnz) {
    return rows ? short_grid(rows, lanes_for(rows, nnz)) + long_grid(rows) : 0;
}

void csr_spmv_impl(const nkpr_csr *a, const double *x, double *y, hipStream_t s) {
    const CsrView A = view_of(a);
    NKPR_CHECK(A.rows == 0 || (x && y), "x and y must not be NULL");
    NKPR_CHECK(x != y || A.rows == 0, "x and y must be different vectors");
    launch_product<false>(A, a->nnz, x, 0.0, y, nullptr, nullptr, s);
}

void csr_step_impl(const nkpr_csr *m, double shift, const double *x, double *y, double *step_sq,
                   double *scratch, uint64_t scratch_count, hipStream_t s) {
    const CsrView A = view_of(m);
    NKPR_CHECK(m->rows == m->cols, "a PageRank step needs a square matrix");
    NKPR_CHECK(step_sq != nullptr, "step_sq must not be NULL");
    NKPR_CHECK(A.rows == 0 || (x && y), "x and y must not be NULL");
    NKPR_CHECK(x != y || A.rows == 0, "x and y must be different vectors");
    const uint32_t need = csr_step_partials(A.rows, m->nnz);
    if (scratch) {
        NKPR_CHECK(need <= scratch_count, "step scratch too small");
        launch_product<true>(A, m->nnz, x, shift, y, scratch, step_sq, s);
    } else {
        PoolBuf buf(8 * (size_t)std::max<uint32_t>(need, 1), s);
        launch_product<true>(A, m->nnz, x, shift, y, static_cast<double *>(buf.p), step_sq, s);
    }
}

// The address of `scratch` on the current device, taken once per device (each GPU holds its
// own copy of the array at its own address).
double *csr_noalloc_scratch() {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    int dev = -1;
    NKPR_HIP(hipGetDevice(&dev));
    NKPR_CHECK(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        NKPR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        NKPR_CHECK(p != nullptr, "hipGetSymbolAddress returned NULL");
        addr[dev] = static_cast<double *>(p);
    }
    return addr[dev];
}

void csr_step_noalloc_impl(const nkpr_csr *m, double shift, const double *x, double *y,
                           double *step_sq, hipStream_t s) {
    constexpr uint64_t kScratch = sizeof(scratch) / sizeof(double);
    static_assert(kScratch >= kShortGrid + kLongGrid, "the step's grids fit the scratch");
    NKPR_CHECK(m != nullptr, "matrix is NULL");
    NKPR_CHECK(csr_step_partials(m->rows, m->nnz) <= kScratch,
               "the step needs more partials than the 4096 the scratch holds");
    csr_step_impl(m, shift, x, y, step_sq, csr_noalloc_scratch(), kScratch, s);
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
    const uint32_t np = csr_step_partials(n, nnz);
    // buffers first: the stream owner below is destroyed (and synchronised) before them
    DevBuf xa, xb, part, out, flag;
    xa.alloc(8 * n);
    xb.alloc(8 * n);
    part.alloc(8 * (size_t)std::max<uint32_t>(np, 1024));
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
    const double tele = (1.0 - damp) / (double)n;
    hipLaunchKernelGGL(csr_fill_kernel, dim3(1024), dim3(256), 0, s, xa.as<double>(), n,
                       1.0 / (double)n);
    NKPR_HIP(hipGetLastError());
    double *cur = xa.as<double>(), *nxt = xb.as<double>();
    uint64_t it = 0;
    double sq = 0.0;
    do {
        launch_product<true>(A, nnz, cur, tele, nxt, part.as<double>(), out.as<double>(), s);
        read_back(s, {{&sq, out.p, 8}});
        std::swap(cur, nxt);
        ++it;
    } while (std::sqrt(sq) > tol && it < max_iterations);
    hipLaunchKernelGGL(csr_sum_kernel, dim3(1024), dim3(kFinalThreads), 0, s, cur, n,
                       part.as<double>());
    hipLaunchKernelGGL(csr_final_kernel, dim3(1), dim3(kFinalThreads), 0, s, part.as<double>(),
                       1024u, out.as<double>());
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
