// AlgebraicPageRank (include/networkit/algebraic/algorithms/AlgebraicPageRank.hpp:94-119) on a
// PageRank handle built from NetworKit's CSR matrix.
//
// The reference iterates x <- M x + (1 - damp) / n from 1/n until ||x' - x||_2 <= tol and
// divides by the sum, with M = damp A^T D^-1 a CSRGeneralMatrix<double>.  Row i of M lists the
// sources j of i's in-edges with the value M(i, j): exactly the in-lists of a graph whose
// coefficients are M's values.  A handle created from them with unit out-degrees (as
// Eigenvector/Katz centrality use, nkpr_power.hip) folds those values unchanged, so the
// PageRank sweep with damp 1 and the teleport term (1 - damp) / n computes the reference's
// step, and its L2 step rule is the reference's loop condition.  The sweep then runs on the
// handle's layout: relabelled by out-degree, source-sliced, and through the binned FP64 copy
// from 2^26 entries on, instead of the CSR kernels' random gathers in the natural order.
// Every row sum is still the FP64 sum of the row's products (the binned copy adds them in
// 128-bit fixed point, more accurately); the parity tests hold it to the oracle's iteration
// count and 1e-12 in L1.
//
// The handle is built once per matrix (about as long as a few dozen CSR steps at R-MAT scale
// 26) and kept, so repeated runs pay for the sweeps alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "nkpr_internal.hpp"
#include "nkpr_run.hpp"

struct nkpr_algebraic {
    nkpr_graph *g = nullptr;
    uint64_t rows = 0;
    int device = 0;
    ~nkpr_algebraic() {
        if (g) nkpr_graph_destroy(g);
    }
};

namespace nkpr {
namespace {

// int32 column ids and FP32 values for the handle's lists, unit out-degrees
__global__ void csr_lists_kernel(const uint64_t *col, const double *val, uint64_t nnz,
                                 int32_t *src, float *w32) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
         k += (uint64_t)gridDim.x * blockDim.x) {
        src[k] = (int32_t)col[k];
        w32[k] = (float)val[k];
    }
}

__global__ void ones_kernel(double *d, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        d[i] = 1.0;
}

__global__ void alg_fill_kernel(double *x, uint64_t n, double v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

unsigned grid_for_n(uint64_t n) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 1u << 16));
}

}  // namespace

void algebraic_create_impl(const nkpr_csr *m, nkpr_algebraic **out) {
    NKPR_CHECK(out != nullptr, "out is NULL");
    NKPR_CHECK(m != nullptr, "matrix is NULL");
    NKPR_CHECK(m->rows == m->cols, "AlgebraicPageRank needs a square matrix");
    NKPR_CHECK(m->rows > 0 && m->rows < 0x7FFFFFFFull, "the matrix needs 1 .. 2^31 - 2 rows");
    NKPR_CHECK(m->nnz < 0xFFFFFFFFull, "the matrix needs fewer than 2^32 - 1 entries");
    NKPR_CHECK(m->row_idx && (m->nnz == 0 || (m->col_idx && m->values)),
               "matrix arrays must not be NULL");
    std::unique_ptr<nkpr_algebraic> h(new nkpr_algebraic());
    NKPR_HIP(hipGetDevice(&h->device));
    h->rows = m->rows;
    hipStream_t s;
    NKPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamOwner {
        hipStream_t s;
        ~StreamOwner() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    };
    // the temporaries outlive the stream owner (it synchronises first)
    DevBuf src, w32, ones;
    {
        StreamOwner owner{s};
        csr_validate(m, s);
        src.alloc(4 * std::max<uint64_t>(m->nnz, 1));
        w32.alloc(4 * std::max<uint64_t>(m->nnz, 1));
        ones.alloc(8 * m->rows);
        if (m->nnz)
            hipLaunchKernelGGL(csr_lists_kernel, dim3(grid_for_n(m->nnz)), dim3(256), 0, s,
                               m->col_idx, m->values, m->nnz, src.as<int32_t>(), w32.as<float>());
        hipLaunchKernelGGL(ones_kernel, dim3(grid_for_n(m->rows)), dim3(256), 0, s,
                           ones.as<double>(), m->rows);
        NKPR_HIP(hipGetLastError());
        NKPR_HIP(hipStreamSynchronize(s));
    }
    nkpr_csc c{};
    c.z = m->rows;
    c.colptr = m->row_idx;
    c.src = src.as<int32_t>();
    c.weight = w32.as<float>();
    c.weight64 = m->values;
    c.out_wdeg = ones.as<double>();
    c.exists = nullptr;
    c.device_ptrs = 1;
    const int rc = nkpr_graph_create(&c, &h->g);
    if (rc != NKPR_OK) {
        h->g = nullptr;
        throw Error(rc, nkpr_last_error());
    }
    *out = h.release();
}

void algebraic_destroy_impl(nkpr_algebraic *h) { delete h; }

int algebraic_device(const nkpr_algebraic *h) { return h->device; }

void algebraic_run_impl(nkpr_algebraic *h, double damp, double tol, uint64_t max_iterations,
                        double *scores, int32_t scores_on_device, uint64_t *iterations,
                        nkpr_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    NKPR_CHECK(h != nullptr && h->g != nullptr, "handle is NULL");
    NKPR_CHECK(scores != nullptr, "scores must not be NULL");
    NKPR_CHECK(damp >= 0.0 && damp <= 1.0, "damp must lie in [0, 1]");
    NKPR_CHECK(tol >= 0.0, "tol must be >= 0");
    NKPR_CHECK(max_iterations >= 1, "max_iterations must be >= 1");
    nkpr_graph *g = h->g;
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;  // every row exists
    NKPR_CHECK((uint64_t)n == h->rows, "the handle lost rows");
    prepare_run(g);
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    nkpr_params pp{};
    nkpr_params_default(&pp);
    pp.max_iterations = max_iterations;
    nkpr_stats st{};
    RunCtx rc{g, &pp, &st, max_iterations, {}, 0};
    rc.combine_grid = (int)std::min<uint64_t>(((uint64_t)g->n_s + kCombineThreads - 1) / kCombineThreads,
                                              (uint64_t)std::max(cus, 1) * 8);
    double *x0 = g->x64a.as<double>(), *x1 = g->x64b.as<double>();
    // the start 1/n is the same in every order
    hipLaunchKernelGGL(alg_fill_kernel, dim3(grid_for_n(n)), dim3(256), 0, s, x0, (uint64_t)n,
                       1.0 / (double)n);
    NKPR_HIP(hipGetLastError());
    launch_init_ctl(g->ctl.as<PhaseCtl>(), nullptr, g->claim.as<uint32_t>(), s);
    SweepArgs<double> a{};
    a.coef = g->coef64.as<double>();  // M's values (unit out-degrees)
    a.partial = g->part64.as<double>();
    a.split_partial = g->split_part64.as<double>();
    a.damp = 1.0;                            // M already holds damp
    a.tele = (1.0 - damp) / (double)n;       // AlgebraicPageRank.hpp:96
    a.n = n;
    a.sink_begin = g->sink_begin;
    a.distribute = 0;
    PhaseRule rule{};
    rule.kind = kRuleStep;                   // while ||x' - x||_2 > tol (AlgebraicPageRank.hpp:101)
    rule.norm = NKPR_L2;
    rule.tol = tol;
    rule.max_iterations = max_iterations;
    rule.damp = damp;
    Vecs<double> v{{x0, x1}, {nullptr, nullptr}};
    const bool timing = stats != nullptr;
    const uint64_t total = run_phase<double>(rc, a, v, kSweepPlain, rule, max_iterations, timing,
                                             &st.ms_sweeps_fp64);
    const double sum = g->ctl_host->sum;
    NKPR_CHECK(std::isfinite(sum) && sum != 0.0, "the scores do not sum to a finite non-zero value");
    // scores / sum in row order (AlgebraicPageRank.hpp:103-116)
    double *sz = scores;
    if (!scores_on_device) {
        ensure(g->zbuf, 8 * (size_t)n);
        sz = g->zbuf.as<double>();
    }
    launch_gather_result(v.x[total & 1], slot_index(g), n, sum, sz, s);
    read_back(s, {{scores, sz, scores_on_device ? 0 : 8 * (size_t)n}});
    if (iterations) *iterations = total;
    if (stats) {
        st.iterations = total;
        st.sweeps_fp64 = total;
        st.score_sum = sum;
        st.last_step = std::sqrt(g->ctl_host->l2);
        st.last_step_l1 = g->ctl_host->l1;
        st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        *stats = st;
    }
}

}  // namespace nkpr
