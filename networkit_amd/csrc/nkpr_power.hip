// Eigenvector and Katz centrality on the PageRank sweep machinery (nkpr_power_run).
//
// Both are power iterations over the weighted adjacency matrix A, pulled over in-lists like
// PageRank:
//   EigenvectorCentrality::run (networkit/cpp/centrality/EigenvectorCentrality.cpp:21-66):
//       y = A x;  len = ||y||_2;  x = y / len;  until |len - previous len| <= tol;
//       finally |x| when x at slot 0 is negative.
//   KatzCentrality::run (networkit/cpp/centrality/KatzCentrality.cpp:28-76):
//       y[u] = sum over in-edges (v, w) of w alpha (1 + x[v]) + beta;  len = ||y||_2;  x = y;
//       until |len - previous len| <= tol;  finally x / len.
// Start vector 1 everywhere; the norms are over existing nodes; a deleted slot keeps 1.
//
// A graph handle whose out-degrees are all 1 (created with out_wdeg = 1) has the weights
// themselves as its folded coefficients, so the pull sweep with damp d and right-hand side b
// computes y = d A x + b: d = 1, b = 0 for the eigenvector; for Katz d = alpha and
// b[u] = alpha (sum of u's in-weights) + beta, which is alpha A (1 + x) + beta with the terms
// grouped differently.  After every sweep a hook adds ||y||^2 in a fixed order, compares the
// norm with the previous one on the device (it sets the phase's stop flag, so the host reads
// back once per batch of sweeps, as in nkpr_run) and, for the eigenvector, divides y by it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>

#include "nkpr_internal.hpp"
#include "nkpr_run.hpp"

namespace nkpr {
namespace {

constexpr int kSqThreads = 256;
constexpr int kSqBlocks = 1024;

// State of the norm test, kept on the device across the sweeps of a run.
struct PowerState {
    double len;       // ||y||_2 of the last real sweep
    double prev;      // the norm of the sweep before
    double inv;       // 1 / len (the eigenvector's rescaling)
    uint64_t it;      // real sweeps
    uint32_t apply;   // the last hook belonged to a real sweep
    uint32_t bad;     // the norm is not finite (Katz with alpha above 1 / spectral radius)
};

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename T>
__global__ __launch_bounds__(kSqThreads) void sumsq_kernel(const T *y, uint32_t n,
                                                            const PhaseCtl *ctl, double *part) {
    if (ctl->done) return;
    __shared__ double s_red[kSqThreads / 64];
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kSqThreads + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kSqThreads) {
        const double v = (double)y[i];
        acc += v * v;
    }
    acc = wave_sum64(acc);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kSqThreads / 64; ++w) t += s_red[w];
        part[blockIdx.x] = t;
    }
}

// One block: the norm of this sweep's y and the stop test |len - prev| <= tol, written in
// the form of Aux::NumericTools::equal (len <= prev + tol and len >= prev - tol).
__global__ __launch_bounds__(kSqThreads) void power_decide_kernel(const double *part, int np,
                                                                   PowerState *ps, PhaseCtl *ctl,
                                                                   double tol) {
    if (ctl->done) {
        if (threadIdx.x == 0) ps->apply = 0u;  // a sweep after the stop: nothing to rescale
        return;
    }
    __shared__ double s_red[kSqThreads / 64];
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += kSqThreads) acc += part[i];
    acc = wave_sum64(acc);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double ss = 0.0;
        for (int w = 0; w < kSqThreads / 64; ++w) ss += s_red[w];
        const double len = sqrt(ss);
        const double prev = ps->len;
        ps->prev = prev;
        ps->len = len;
        ps->inv = 1.0 / len;
        ps->it += 1;
        ps->apply = 1u;
        bool stop = (len <= prev + tol) && (len >= prev - tol);
        if (!isfinite(len)) {
            ps->bad = 1u;
            stop = true;
        }
        if (stop) ctl->done = 1u;
    }
}

// y *= 1 / len (and the prescaled copy of an unweighted graph) after a real sweep.
template <typename T>
__global__ void power_scale_kernel(T *y, T *yh, uint32_t n, const PowerState *ps) {
    if (!ps->apply) return;
    const T f = (T)ps->inv;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        y[i] *= f;
        if (yh) yh[i] *= f;
    }
}

__global__ void fill_kernel(double *x, uint64_t count, double v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

__global__ void abs_kernel(double *x, uint64_t count) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = fabs(x[i]);
}

// b[u] = alpha * rowsum[u] + beta, rowsum = A 1 (a sweep of the all-ones vector)
__global__ void katz_base_kernel(const double *rowsum, uint32_t n, double alpha, double beta,
                                 double *b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = alpha * rowsum[i] + beta;
}

unsigned grid_of(uint64_t count, unsigned cap = 65536) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((count + 255) / 256, cap));
}

}  // namespace

void power_run_impl(nkpr_graph *g, const nkpr_power_params *p, double *scores, nkpr_stats *st) {
    const auto t0 = std::chrono::steady_clock::now();
    NKPR_CHECK(g != nullptr && p != nullptr, "graph or params is NULL");
    NKPR_CHECK(g->n_groups == 0, "a batch handle runs PageRank only (nkpr_run_batch)");
    NKPR_CHECK(p->kind == NKPR_EIGENVECTOR || p->kind == NKPR_KATZ, "unknown centrality kind");
    NKPR_CHECK(g->z == 0 || scores != nullptr, "scores is NULL");
    NKPR_CHECK(!(p->tol < 0.0) && !std::isnan(p->tol), "tol must be >= 0");
    NKPR_CHECK(p->kind != NKPR_KATZ || (p->alpha >= 0.0 && std::isfinite(p->alpha)),
               "Alpha must be non negative.");
    NKPR_CHECK(p->kind != NKPR_KATZ || std::isfinite(p->beta), "beta must be finite");
    NKPR_CHECK(g->n == 0 || g->unit_degrees,
               "the graph handle must be created with out_wdeg = 1 for every node (its "
               "coefficients are then the edge weights)");
    std::memset(st, 0, sizeof(*st));
    const uint64_t z = g->z;
    const uint32_t n = g->n;
    const hipStream_t s = g->stream;
    if (z == 0) return;
    double *sz = scores;
    if (!p->scores_on_device) {
        ensure(g->zbuf, 8 * z);
        sz = g->zbuf.as<double>();
    }
    // every slot starts at 1; a deleted slot keeps it (both reference loops leave it alone)
    hipLaunchKernelGGL(fill_kernel, dim3(grid_of(z)), dim3(256), 0, s, sz, z, 1.0);
    NKPR_HIP(hipGetLastError());
    if (n > 0) {
        prepare_run(g);
        const bool katz = p->kind == NKPR_KATZ;
        const bool folded = g->weighted;
        int cus = 0;
        NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
        nkpr_params pp{};
        nkpr_params_default(&pp);
        pp.interrupt = p->interrupt;
        pp.max_iterations = p->max_iterations;
        RunCtx rc{g, &pp, st, p->max_iterations, {}, 0};
        rc.combine_grid = (int)std::min<uint64_t>(((uint64_t)g->n_s + kCombineThreads - 1) / kCombineThreads,
                                                  (uint64_t)std::max(cus, 1) * 8);
        double *x0 = g->x64a.as<double>(), *x1 = g->x64b.as<double>();
        double *xh0 = folded ? nullptr : g->xh64a.as<double>();
        double *xh1 = folded ? nullptr : g->xh64b.as<double>();
        ensure(g->base64, 8 * (size_t)n);
        ensure(g->red_partials, 8 * (size_t)std::max(kSqBlocks, 1024));
        DevBuf state;
        state.alloc(sizeof(PowerState));
        PowerState *ps = state.as<PowerState>();
        NKPR_HIP(hipMemsetAsync(ps, 0, sizeof(PowerState), s));

        SweepArgs<double> a{};
        a.coef = folded ? g->coef64.as<double>() : nullptr;
        a.inv = folded ? nullptr : g->inv64.as<double>();
        a.partial = g->part64.as<double>();
        a.split_partial = g->split_part64.as<double>();
        a.tele = 0.0;
        a.n = n;
        a.sink_begin = g->sink_begin;
        a.distribute = 0;
        PhaseRule never{};  // the hook decides; the sweep's own rule never stops
        never.kind = kRuleStep;
        never.norm = NKPR_L1;
        never.tol = -1.0;
        never.max_iterations = std::numeric_limits<uint64_t>::max();
        never.damp = 0.5;
        double unused = 0.0;
        auto start_ones = [&]() {
            hipLaunchKernelGGL(fill_kernel, dim3(grid_of(n)), dim3(256), 0, s, x0, (uint64_t)n, 1.0);
            NKPR_HIP(hipGetLastError());
            if (!folded)
                launch_prescale(x0, g->inv64.as<double>(), xh0, nullptr, nullptr, nullptr, n, g->sink_begin, s);
            launch_init_ctl(g->ctl.as<PhaseCtl>(), nullptr, g->claim.as<uint32_t>(), s);
        };
        if (katz) {
            // b = alpha A 1 + beta, A 1 from one sweep of the all-ones vector
            start_ones();
            SweepArgs<double> r = a;
            r.damp = 1.0;
            Vecs<double> vr{{x0, x1}, {xh0, xh1}};
            run_phase<double>(rc, r, vr, kSweepPlain, never, 1, false, &unused);
            hipLaunchKernelGGL(katz_base_kernel, dim3(grid_of(n, 1u << 30)), dim3(256), 0, s,
                               (const double *)x1, n, p->alpha, p->beta, g->base64.as<double>());
            NKPR_HIP(hipGetLastError());
        }
        a.damp = katz ? p->alpha : 1.0;
        a.base_vec = katz ? g->base64.as<double>() : nullptr;
        const int kind = katz ? kSweepBaseVec : kSweepPlain;
        double *part = g->red_partials.as<double>();
        PhaseCtl *ctl = g->ctl.as<PhaseCtl>();
        auto hook = [&](const double *, double *y) {
            hipLaunchKernelGGL(sumsq_kernel<double>, dim3(std::min<unsigned>(grid_of(n), kSqBlocks)),
                               dim3(kSqThreads), 0, s, (const double *)y, n, (const PhaseCtl *)ctl, part);
            hipLaunchKernelGGL(power_decide_kernel, dim3(1), dim3(kSqThreads), 0, s,
                               (const double *)part, (int)std::min<unsigned>(grid_of(n), kSqBlocks), ps, ctl,
                               p->tol);
            if (!katz) {
                double *yh = folded ? nullptr : (y == x0 ? xh0 : xh1);
                hipLaunchKernelGGL(power_scale_kernel<double>, dim3(grid_of(n)), dim3(256), 0, s, y, yh,
                                   n, (const PowerState *)ps);
            }
            NKPR_HIP(hipGetLastError());
        };
        start_ones();
        Vecs<double> v{{x0, x1}, {xh0, xh1}};
        // the sweeps are timed with HIP events (ms_sweeps_fp64), the norm hooks are not
        const uint64_t total = run_phase<double>(rc, a, v, kind, never,
                                                 std::max<uint64_t>(p->max_iterations, 1), true,
                                                 &st->ms_sweeps_fp64, hook);
        PowerState hs{};
        read_back(s, {{&hs, ps, sizeof(PowerState)}});
        NKPR_CHECK(!hs.bad, "the iteration diverged (the norm is not finite): alpha must be "
                            "below 1 / the spectral radius of the adjacency matrix");
        const double *y = v.x[total & 1];
        st->iterations = total;
        st->sweeps_fp64 = total;
        st->last_step = std::fabs(hs.len - hs.prev);
        st->last_step_l1 = st->last_step;
        st->score_sum = hs.len;
        // eigenvector: y is already normalised; Katz: divided by the last norm
        launch_gather_result(y, slot_index(g), z, katz ? hs.len : 1.0, sz, s);
        if (!katz) {
            // sign correction on slot 0 (EigenvectorCentrality.cpp:58-63); deleted slots hold 1
            double s0 = 0.0;
            read_back(s, {{&s0, sz, 8}});
            if (s0 < 0.0) {
                hipLaunchKernelGGL(abs_kernel, dim3(grid_of(z)), dim3(256), 0, s, sz, z);
                NKPR_HIP(hipGetLastError());
            }
        }
    }
    read_back(s, {{scores, sz, p->scores_on_device ? 0 : 8 * z}});
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace nkpr
