// Adjoint of PageRank: gradients of a loss on the scores with respect to the edge weights and
// the teleport vector (no counterpart in the reference; the forward is the loop of
// networkit/cpp/centrality/PageRank.cpp:53-80 with its sum normalisation).
//
// Forward: y = d P(w) y + (1-d) t,  P[u][v] = w_vu / deg_v,  x = y / sum(y).
// For a loss L(x) with c = dL/dx:
//   dL/dy   g = (c - (c . x) 1) / sum(y)
//   adjoint lambda = d P^T lambda + g                  (solved on the transposed operator)
//   dL/dt   (1-d) lambda
//   dL/dw_vu = d y_v / deg_v (lambda_u - mu_v),  mu_v = sum_k w_vk lambda_k / deg_v = (P^T lambda)_v
// (w_vu enters P[u][v] directly and every P[k][v] through deg_v.)
//
// The transposed operator is built once per set of coefficients as a graph of its own:
// its in-lists are the out-lists of the forward graph and its edge coefficients are the
// forward's folded w / deg, so the forward's sweep kernels, schedule and FP32 GMRES serve it
// unchanged.  Distributed sinks are not supported (the reference has no sink handling).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "nkpr_internal.hpp"
#include "nkpr_run.hpp"

namespace nkpr {

void create_graph(const nkpr_csc *c, nkpr_graph *g, const uint64_t *group_off = nullptr,
                  uint32_t n_groups = 0);

namespace {

inline unsigned nblocks(uint64_t n, int t = 256) {
    return (unsigned)std::min<uint64_t>((n + t - 1) / t, 65536);
}

__global__ void row_of_edge_kernel(const uint32_t *rowptr, uint32_t n, uint32_t *row) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < n; w += nw)
        for (uint32_t e = rowptr[w] + lane; e < rowptr[w + 1]; e += 64) row[e] = (uint32_t)w;
}

__global__ void key_val_kernel(const int32_t *src, uint64_t m, uint32_t *key, uint32_t *val,
                               uint32_t *cnt) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        key[e] = (uint32_t)src[e];
        val[e] = (uint32_t)e;
        atomicAdd(&cnt[src[e]], 1u);
    }
}

// Transposed edge j (sorted by forward source): forward edge e = order[j] from v to row u.
__global__ void transpose_edges_kernel(const uint32_t *order, uint64_t m, const uint32_t *row,
                                       const int32_t *src, const float *c32, const double *c64,
                                       const float *inv32, const double *inv64, int32_t *tdst,
                                       float *tw32, double *tw64) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = order[j];
        tdst[j] = (int32_t)row[e];
        if (c32) {
            tw32[j] = c32[e];
            tw64[j] = c64[e];
        } else {
            tw32[j] = inv32[src[e]];
            tw64[j] = inv64[src[e]];
        }
    }
}

// Coefficients of the transposed entries after new forward weights.
__global__ void transpose_coef_kernel(const uint32_t *order, uint64_t m, const float *c32,
                                      const double *c64, float *tw32, double *tw64) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = order[j];
        tw32[j] = c32[e];
        tw64[j] = c64[e];
    }
}

__global__ void widen_kernel(const uint32_t *in, uint64_t n, uint64_t *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

__global__ void fill_kernel(double *x, uint64_t n, double v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// Deterministic FP64 sums over [0, n): per-block partials of a * b (b may be null: sum of |a|).
__global__ void dot_partial_kernel(const double *a, const double *b, uint64_t n, double *part) {
    __shared__ double s[256 / 64];
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        acc += b ? a[i] * b[i] : fabs(a[i]);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 256 / 64; ++w) t += s[w];
        part[blockIdx.x] = t;
    }
}

__global__ void dot_final_kernel(const double *part, int nb, double *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double t = 0.0;
        for (int b = 0; b < nb; ++b) t += part[b];
        *out = t;
    }
}

// g_f[u] = (c[perm[u]] - dot) / S ;  x_f[u] = x[perm[u]]
__global__ void seed_kernel(const double *c_z, const int32_t *perm, uint32_t n, const double *dot,
                            double inv_sum, double *gf) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) gf[u] = (c_z[perm[u]] - *dot) * inv_sum;
}

// out_z[perm_f[perm_a[i]]] = f * v[i]
__global__ void scatter2_kernel(const double *v, const int32_t *perm_a, const int32_t *perm_f,
                                uint32_t n, double f, double *out_z) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out_z[perm_f[perm_a[i]]] = f * v[i];
}

__global__ void wdeg_slot_kernel(const int32_t *src, uint64_t m, const float *w,
                                 const double *w64, double *deg) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&deg[src[e]], w64 ? w64[e] : (w ? (double)w[e] : 1.0));
}

// dL/dw for every in-edge of the caller's CSC: one wave per row.
__global__ void grad_weight_kernel(const uint64_t *colptr, const int32_t *src, uint64_t z,
                                   const double *x, double score_sum, const double *deg,
                                   const double *lam, const double *mu, double damp,
                                   double *grad) {
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const int lane = threadIdx.x & 63;
    for (uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < z; r += nw) {
        const uint64_t b = colptr[r], e = colptr[r + 1];
        if (b == e) continue;
        const double lr = lam[r];
        for (uint64_t i = b + lane; i < e; i += 64) {
            const int32_t v = src[i];
            grad[i] = damp * x[v] * score_sum / deg[v] * (lr - mu[v]);
        }
    }
}

__global__ void scale_kernel(const double *in, uint64_t n, double f, double *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f * in[i];
}

double device_dot(const double *a, const double *b, uint64_t n, double *part, double *out,
                  hipStream_t s) {
    const int nb = (int)std::max<unsigned>(1, std::min<unsigned>(nblocks(n), 1024));
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(256), 0, s, a, b, n, part);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(64), 0, s, (const double *)part, nb, out);
    NKPR_HIP(hipGetLastError());
    double h = 0.0;
    read_back(s, {{&h, out, 8}});
    return h;
}

// The graph of d P^T: rows are forward sources, entries their out-edges with the forward's
// folded coefficients.
void build_adjoint_graph(nkpr_graph *g) {
    batch_layout(g);
    if (g->adj && g->adj_stale) {
        // same structure, new coefficients: fold them into the transposed graph in place
        const uint64_t m = g->m;
        DevBuf tw32, tw64;
        tw32.alloc(4 * std::max<uint64_t>(m, 1));
        tw64.alloc(8 * std::max<uint64_t>(m, 1));
        if (m)
            hipLaunchKernelGGL(transpose_coef_kernel, dim3(nblocks(m)), dim3(256), 0, g->stream,
                               g->adj_order.as<uint32_t>(), m, g->b_coef32.as<float>(),
                               g->b_coef64.as<double>(), tw32.as<float>(), tw64.as<double>());
        NKPR_HIP(hipGetLastError());
        set_weights(g->adj.get(), tw32.as<float>(), tw64.as<double>(), g->adj_ones.as<double>(), 1);
        g->adj_stale = false;
        return;
    }
    if (g->adj) return;
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    DevBuf row, key1, key2, val1, val2, cnt, ptr32, ptr64, tdst, tw32, tw64, ones, tmp;
    row.alloc(4 * std::max<uint64_t>(m, 1));
    key1.alloc(4 * std::max<uint64_t>(m, 1));
    key2.alloc(4 * std::max<uint64_t>(m, 1));
    val1.alloc(4 * std::max<uint64_t>(m, 1));
    val2.alloc(4 * std::max<uint64_t>(m, 1));
    cnt.alloc(4 * ((uint64_t)n + 1));
    ptr32.alloc(4 * ((uint64_t)n + 1));
    NKPR_HIP(hipMemsetAsync(cnt.p, 0, 4 * ((uint64_t)n + 1), s));
    if (n)
        hipLaunchKernelGGL(row_of_edge_kernel, dim3((unsigned)std::min<uint64_t>(((uint64_t)n * 64 + 255) / 256, 1u << 18)),
                           dim3(256), 0, s, g->b_rowptr.as<uint32_t>(), n, row.as<uint32_t>());
    if (m)
        hipLaunchKernelGGL(key_val_kernel, dim3(nblocks(m)), dim3(256), 0, s,
                           g->b_src.as<int32_t>(), m, key1.as<uint32_t>(), val1.as<uint32_t>(),
                           cnt.as<uint32_t>());
    NKPR_HIP(hipGetLastError());
    hipcub::DoubleBuffer<uint32_t> kb(key1.as<uint32_t>(), key2.as<uint32_t>());
    hipcub::DoubleBuffer<uint32_t> vb(val1.as<uint32_t>(), val2.as<uint32_t>());
    int bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)n) ++bits;
    size_t tb = 0;
    if (m) {
        NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb, vb, (size_t)m, 0, bits, s));
        tmp.alloc(tb);
        NKPR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kb, vb, (size_t)m, 0, bits, s));
    }
    tb = 0;
    NKPR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.as<uint32_t>(), ptr32.as<uint32_t>(),
                                              (size_t)n + 1, s));
    DevBuf tmp2;
    tmp2.alloc(tb);
    NKPR_HIP(hipcub::DeviceScan::ExclusiveSum(tmp2.p, tb, cnt.as<uint32_t>(), ptr32.as<uint32_t>(),
                                              (size_t)n + 1, s));
    ptr64.alloc(8 * ((uint64_t)n + 1));
    hipLaunchKernelGGL(widen_kernel, dim3(nblocks((uint64_t)n + 1)), dim3(256), 0, s,
                       ptr32.as<uint32_t>(), (uint64_t)n + 1, ptr64.as<uint64_t>());
    tdst.alloc(4 * std::max<uint64_t>(m, 1));
    tw32.alloc(4 * std::max<uint64_t>(m, 1));
    tw64.alloc(8 * std::max<uint64_t>(m, 1));
    if (m)
        hipLaunchKernelGGL(transpose_edges_kernel, dim3(nblocks(m)), dim3(256), 0, s,
                           (const uint32_t *)vb.Current(), m, row.as<uint32_t>(),
                           g->b_src.as<int32_t>(), g->weighted ? g->b_coef32.as<float>() : nullptr,
                           g->weighted ? g->b_coef64.as<double>() : nullptr,
                           g->weighted ? nullptr : g->inv32.as<float>(),
                           g->weighted ? nullptr : g->inv64.as<double>(), tdst.as<int32_t>(),
                           tw32.as<float>(), tw64.as<double>());
    ones.alloc(8 * std::max<uint32_t>(n, 1));
    if (n) hipLaunchKernelGGL(fill_kernel, dim3(nblocks(n)), dim3(256), 0, s, ones.as<double>(), (uint64_t)n, 1.0);
    NKPR_HIP(hipGetLastError());
    NKPR_HIP(hipStreamSynchronize(s));
    std::unique_ptr<nkpr_graph> a(new nkpr_graph());
    a->device = g->device;
    a->stream = g->stream;
    nkpr_csc c{};
    c.z = n;
    c.colptr = ptr64.as<uint64_t>();
    c.src = tdst.as<int32_t>();
    c.weight = tw32.as<float>();
    c.weight64 = tw64.as<double>();
    c.out_wdeg = ones.as<double>();
    c.exists = nullptr;
    c.device_ptrs = 1;
    create_graph(&c, a.get());
    g->adj = std::move(a);
    g->adj_stale = false;
    if (g->weighted) {
        // kept to refresh the coefficients when the forward weights change
        g->adj_order.alloc(4 * std::max<uint64_t>(m, 1));
        if (m) NKPR_HIP(hipMemcpyAsync(g->adj_order.p, vb.Current(), 4 * m, hipMemcpyDeviceToDevice, s));
        std::swap(g->adj_ones.p, ones.p);
        std::swap(g->adj_ones.bytes, ones.bytes);
        NKPR_HIP(hipStreamSynchronize(s));
    }
}

}  // namespace

void adjoint_impl(nkpr_graph *g, const nkpr_params *p, const nkpr_csc *csc,
                  const double *scores, double score_sum, const double *grad_scores,
                  double *grad_weight, double *grad_teleport, nkpr_stats *st) {
    const auto t0 = std::chrono::steady_clock::now();
    NKPR_CHECK(g != nullptr && p != nullptr && csc != nullptr, "graph, params or csc is NULL");
    NKPR_CHECK(scores != nullptr && grad_scores != nullptr, "scores or grad_scores is NULL");
    NKPR_CHECK(p->damp > 0.0 && p->damp < 1.0, "damp must be in (0, 1)");
    NKPR_CHECK(p->sink_handling == NKPR_SINK_NONE,
               "the adjoint supports NKPR_SINK_NONE (the reference's sink behaviour) only");
    NKPR_CHECK(!p->normalized, "the adjoint differentiates the sum-normalised scores (normalized = 0)");
    NKPR_CHECK(p->precision >= NKPR_FP64 && p->precision <= NKPR_FP32, "unknown precision");
    NKPR_CHECK(score_sum > 0.0 && std::isfinite(score_sum), "score_sum must be the forward's positive sum");
    NKPR_CHECK(csc->z == g->z && csc->colptr != nullptr, "csc does not describe this graph");
    NKPR_CHECK(!(p->tol < 0.0) && !std::isnan(p->tol), "tol must be >= 0");
    std::memset(st, 0, sizeof(*st));
    const hipStream_t s = g->stream;
    const uint64_t z = g->z, m = g->m;
    const uint32_t n = g->n;
    const double d = p->damp;
    if (n == 0) return;

    // ---- inputs on the device -------------------------------------------------------------
    DevBuf hz[4];
    auto to_dev = [&](DevBuf &b, const void *h, size_t bytes, bool on_dev) -> const void * {
        if (on_dev || h == nullptr) return h;
        b.alloc(std::max<size_t>(bytes, 16));
        NKPR_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
        return b.p;
    };
    const bool vdev = p->scores_on_device != 0;
    const double *x_z = (const double *)to_dev(hz[0], scores, 8 * z, vdev);
    const double *c_z = (const double *)to_dev(hz[1], grad_scores, 8 * z, vdev);
    const bool edev = csc->device_ptrs != 0;
    DevBuf eb[4];
    const uint64_t *colptr = (const uint64_t *)to_dev(eb[0], csc->colptr, 8 * (z + 1), edev);
    const int32_t *src = (const int32_t *)to_dev(eb[1], csc->src, 4 * m, edev);
    const float *w32 = (const float *)to_dev(eb[2], csc->weight, 4 * m, edev);
    const double *w64 = (const double *)to_dev(eb[3], csc->weight64, 8 * m, edev);
    NKPR_CHECK(m == 0 || src != nullptr, "csc src is NULL");

    build_adjoint_graph(g);
    nkpr_graph *a = g->adj.get();
    for (auto &b : g->a_vec) ensure(b, 8 * std::max<uint64_t>(z, (uint64_t)n) + 64);
    double *gf = g->a_vec[0].as<double>();     // dL/dy, forward compact
    double *lam_z = g->a_vec[1].as<double>();  // lambda, slots
    double *mu_z = g->a_vec[2].as<double>();   // (P^T lambda), slots
    double *deg = g->a_vec[3].as<double>();    // weighted out-degree, slots
    double *part = g->a_vec[4].as<double>();
    double *scal = g->a_vec[5].as<double>();

    // ---- g = (c - c.x) / S in the transposed graph's order ---------------------------------
    // c . x over existing slots: deleted slots carry no score and no gradient
    prepare_run(a);
    launch_gather_scaled(x_z, g->perm.as<int32_t>(), n, 1.0, a->x64b.as<double>(), nullptr, s);
    launch_gather_scaled(c_z, g->perm.as<int32_t>(), n, 1.0, a->x64a.as<double>(), nullptr, s);
    device_dot(a->x64a.as<double>(), a->x64b.as<double>(), n, part, scal, s);
    hipLaunchKernelGGL(seed_kernel, dim3(nblocks(n)), dim3(256), 0, s, c_z, g->perm.as<int32_t>(),
                       n, (const double *)scal, 1.0 / score_sum, gf);
    NKPR_HIP(hipGetLastError());
    ensure(a->base64, 8 * (size_t)n);
    ensure(a->base32, 4 * (size_t)n);
    launch_gather_scaled(gf, a->perm.as<int32_t>(), n, 1.0, a->base64.as<double>(),
                         a->base32.as<float>(), s);
    const double gnorm = device_dot(a->base64.as<double>(), nullptr, n, part, scal, s);

    // ---- lambda = d P^T lambda + g ---------------------------------------------------------
    nkpr_params pa = *p;
    pa.sink_handling = NKPR_SINK_NONE;
    pa.teleport = nullptr;
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, a->device));
    RunCtx rc{a, &pa, st, pa.max_iterations, {}, 0};
    rc.combine_grid = (int)std::min<uint64_t>(((uint64_t)a->n_s + kCombineThreads - 1) / kCombineThreads,
                                              (uint64_t)std::max(cus, 1) * 8);
    const bool timing = p->timing != 0;
    Krylov kr = krylov_setup(a);
    a->op_pending = false;
    // w = A v (residual updates) and w = B v (the GMRES Krylov operator), A = I - B
    auto apply = [&](const float *v, float *w) {
        apply_operator(rc, v, w, kr, timing, &st->ms_sweeps_fp32, true);
    };
    auto apply_b = [&](const float *v, float *w) {
        apply_operator(rc, v, w, kr, timing, &st->ms_sweeps_fp32, false);
    };
    auto sweep_args = [&]() {
        SweepArgs<double> sa{};
        sa.coef = a->coef64.as<double>();
        sa.partial = a->part64.as<double>();
        sa.split_partial = a->split_part64.as<double>();
        sa.damp = d;
        sa.tele = 0.0;
        sa.n = n;
        sa.sink_begin = a->sink_begin;
        sa.distribute = 0;
        return sa;
    };
    const double target = p->tol * gnorm;  // relative L1 residual
    double *lam = a->x64a.as<double>();
    double *ybuf = a->x64b.as<double>();
    uint64_t total = 0;
    if (gnorm == 0.0) {
        hipLaunchKernelGGL(fill_kernel, dim3(nblocks(n)), dim3(256), 0, s, lam, (uint64_t)n, 0.0);
    } else if (p->precision == NKPR_FP32) {
        // restarted FP32 GMRES on A lambda = g from lambda = 0; each restart's residual is
        // r - A e, one more operator application
        const uint64_t cap = std::max<uint64_t>(1, p->max_iterations);
        hipLaunchKernelGGL(fill_kernel, dim3(nblocks(n)), dim3(256), 0, s, lam, (uint64_t)n, 0.0);
        NKPR_HIP(hipGetLastError());
        float *r = a->r32.as<float>();
        launch_copy_f32(a->base32.as<float>(), r, n, s);
        auto norm2 = [&](const float *v) {
            double h = 0.0;
            launch_multidot(v, v, kr.ld, 1, n, kr.dpart, kr.dout, s);
            read_back(s, {{&h, kr.dout, 8}});
            return std::sqrt(h);
        };
        const double g2 = norm2(r);
        double r2 = g2;
        for (int round = 0; round < kMaxRefinements && total < cap && r2 > p->tol * g2; ++round) {
            const double rtol = std::min(0.5, std::max(1e-7, p->tol * g2 / r2));
            total += gmres_f32(kr, apply_b, r, kr.e, rtol, cap - total, s);
            launch_add_correction(lam, kr.e, n, s);
            float *w = kr.V + kr.ld;  // free once the solve has returned
            apply(kr.e, w);
            ++total;
            launch_axpby(r, w, r, n, 1.0f, -1.0f, s);
            r2 = norm2(r);
        }
        st->sweeps_fp32 = total;
        st->last_step = r2 / std::max(g2, 1e-300);
        st->l1_error_bound = st->last_step;
    } else {
        // FP64 residual sweeps with FP32 GMRES corrections, from lambda = g
        launch_gather_scaled(gf, a->perm.as<int32_t>(), n, 1.0, lam, nullptr, s);
        for (int round = 0;; ++round) {
            launch_init_ctl(a->ctl.as<PhaseCtl>(), nullptr, a->claim.as<uint32_t>(), s);
            SweepArgs<double> sa = sweep_args();
            sa.base_vec = a->base64.as<double>();
            sa.r32 = kr.V;  // the residual is the first Krylov vector
            PhaseRule rr{};
            rr.kind = kRuleStep;
            rr.norm = NKPR_L1;
            rr.tol = target;
            rr.max_iterations = 1;
            rr.damp = d;
            Vecs<double> vr{{lam, ybuf}, {nullptr, nullptr}};
            total += run_phase<double>(rc, sa, vr, kSweepBaseResid, rr, 1, timing, &st->ms_sweeps_fp64);
            st->sweeps_fp64 += 1;
            const double l1 = a->ctl_host->l1;  // ||r||_1, r = (d P^T lambda + g) - lambda
            st->last_step_l1 = l1;
            st->last_step = l1 / gnorm;
            st->l1_error_bound = l1 / gnorm;
            if (l1 <= target || total >= p->max_iterations || round >= kMaxRefinements) {
                std::swap(lam, ybuf);  // lambda + r: one more fixed-point step
                break;
            }
            ++st->refinements;
            const double rtol = std::min(0.5, std::max(1e-7, 0.5 * target / l1));
            const uint64_t left = p->max_iterations > total + 1 ? p->max_iterations - total - 1 : 0;
            const uint64_t kc = gmres_f32(kr, apply_b, kr.V, kr.e, rtol, left, s);
            st->sweeps_fp32 += kc;
            st->sweeps_fp32_correction += kc;
            total += kc;
            launch_add_correction(lam, kr.e, n, s);
        }
    }
    collect_op_timing(a, &st->ms_sweeps_fp32);
    st->iterations = total;

    // ---- mu = P^T lambda (one FP64 sweep), back to slots ----------------------------------
    {
        launch_init_ctl(a->ctl.as<PhaseCtl>(), nullptr, a->claim.as<uint32_t>(), s);
        SweepArgs<double> sa = sweep_args();
        PhaseRule r1{};
        r1.kind = kRuleStep;
        r1.norm = NKPR_L1;
        r1.tol = -1.0;
        r1.max_iterations = 1;
        r1.damp = d;
        double *out = (lam == a->x64a.as<double>()) ? a->x64b.as<double>() : a->x64a.as<double>();
        Vecs<double> vm{{lam, out}, {nullptr, nullptr}};
        run_phase<double>(rc, sa, vm, kSweepPlain, r1, 1, timing, &st->ms_sweeps_fp64);
        NKPR_HIP(hipMemsetAsync(lam_z, 0, 8 * z, s));
        NKPR_HIP(hipMemsetAsync(mu_z, 0, 8 * z, s));
        hipLaunchKernelGGL(scatter2_kernel, dim3(nblocks(n)), dim3(256), 0, s, (const double *)lam,
                           a->perm.as<int32_t>(), g->perm.as<int32_t>(), n, 1.0, lam_z);
        hipLaunchKernelGGL(scatter2_kernel, dim3(nblocks(n)), dim3(256), 0, s, (const double *)out,
                           a->perm.as<int32_t>(), g->perm.as<int32_t>(), n, 1.0 / d, mu_z);
        NKPR_HIP(hipGetLastError());
    }

    // ---- dL/dt = (1 - d) lambda -------------------------------------------------------------
    if (grad_teleport) {
        if (vdev) {
            hipLaunchKernelGGL(scale_kernel, dim3(nblocks(z)), dim3(256), 0, s, (const double *)lam_z, z,
                               1.0 - d, grad_teleport);
        } else {
            hipLaunchKernelGGL(scale_kernel, dim3(nblocks(z)), dim3(256), 0, s, (const double *)lam_z, z,
                               1.0 - d, part);
            NKPR_HIP(hipGetLastError());
            read_back(s, {{grad_teleport, part, 8 * z}});
        }
        NKPR_HIP(hipGetLastError());
    }

    // ---- dL/dw over the caller's in-lists --------------------------------------------------
    if (grad_weight && m) {
        if (csc->out_wdeg) {
            const double *dz = (const double *)csc->out_wdeg;
            if (!edev) NKPR_HIP(hipMemcpyAsync(deg, dz, 8 * z, hipMemcpyHostToDevice, s));
            else NKPR_HIP(hipMemcpyAsync(deg, dz, 8 * z, hipMemcpyDeviceToDevice, s));
        } else {
            NKPR_HIP(hipMemsetAsync(deg, 0, 8 * z, s));
            hipLaunchKernelGGL(wdeg_slot_kernel, dim3(nblocks(m)), dim3(256), 0, s, src, m, w32, w64, deg);
        }
        DevBuf gw;
        double *gout = grad_weight;
        if (!edev) {
            gw.alloc(8 * m);
            gout = gw.as<double>();
        }
        hipLaunchKernelGGL(grad_weight_kernel, dim3((unsigned)std::min<uint64_t>((z * 64 + 255) / 256, 1u << 18)),
                           dim3(256), 0, s, colptr, src, z, x_z, score_sum, (const double *)deg,
                           (const double *)lam_z, (const double *)mu_z, d, gout);
        NKPR_HIP(hipGetLastError());
        // the staging buffer gw is freed at the end of this block: wait for the copy first
        if (!edev) read_back(s, {{grad_weight, gout, 8 * m}});
    }
    NKPR_HIP(hipStreamSynchronize(s));
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace nkpr
