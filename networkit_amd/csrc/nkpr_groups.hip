// PageRank on a batch of graphs at once (nkpr_run_batch; no counterpart in the reference,
// which runs PageRank::run once per graph, networkit/cpp/centrality/PageRank.cpp:20-104).
//
// The batch is one graph handle over the block-diagonal union of the graphs: graph q owns the
// compact ids [goff[q], goff[q+1]) (relabelled inside the graph, its sinks last), and no edge
// joins two graphs.  Every sweep is the ordinary sweep of the union with a per-node right-hand
// side b[u] = (1-d)/n_q (+ d S_q / n_q with distributed sinks).  After it, per-graph FP64
// reductions over node chunks (added in chunk order: deterministic) apply the reference's
// stopping rule to every graph on its own; a graph that has stopped is frozen (its new values
// are replaced by the old ones), so each graph ends with exactly the iterate and iteration
// count it would have alone.  Mixed precision certifies every graph on its own; the FP32
// correction solve runs on the union.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "nkpr_internal.hpp"
#include "nkpr_run.hpp"

namespace nkpr {
namespace {

constexpr uint32_t kChunkNodes = 16384;
constexpr int kRedThreads = 256;

struct Chunk {
    uint32_t begin, end, group, sink;  // compact node range, its graph, the graph's sink begin
};

struct GroupCtl {
    double l1, l2, sum, sink, prev_l1, bound;
    uint64_t iterations;
    uint32_t done;
    uint32_t pad;
};

inline unsigned nblk(uint64_t n, int t = 256) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + t - 1) / t, 65536));
}

// Per chunk: FP64 partial sums of |x'-x|, (x'-x)^2, x', x' over sinks; restores frozen graphs.
template <typename T>
__global__ __launch_bounds__(kRedThreads) void group_reduce_kernel(
    const Chunk *ch, T *xnew, const T *xold, const GroupCtl *gc, double *part,
    const uint32_t *flag, int restore) {
    if (*flag) return;
    const Chunk c = ch[blockIdx.x];
    const bool frozen = restore && gc[c.group].done;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t u = c.begin + threadIdx.x; u < c.end; u += kRedThreads) {
        if (frozen) {
            xnew[u] = xold[u];
            continue;
        }
        const double y = (double)xnew[u];
        const double d = y - (double)xold[u];
        a[0] += fabs(d);
        a[1] += d * d;
        a[2] += y;
        if (u >= c.sink) a[3] += y;
    }
    __shared__ double s[4][kRedThreads / 64];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = a[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0) s[q][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int w = 0; w < kRedThreads / 64; ++w) t += s[threadIdx.x][w];
        part[(size_t)blockIdx.x * 4 + threadIdx.x] = t;
    }
}

// One block: every graph adds its chunk partials in order and applies the rule; the next
// right-hand side per graph; all graphs done -> the phase stops.
__global__ __launch_bounds__(1024) void group_rule_kernel(
    const uint32_t *cfirst, const uint32_t *gn, uint32_t groups, const double *part,
    GroupCtl *gc, double *add, PhaseRule r, int dist, int base_only, PhaseCtl *ctl,
    uint32_t *flag) {
    if (*flag) return;
    __shared__ uint32_t s_active;
    if (threadIdx.x == 0) s_active = 0;
    __syncthreads();
    uint32_t active = 0;
    for (uint32_t q = threadIdx.x; q < groups; q += blockDim.x) {
        GroupCtl &g = gc[q];
        const double nq = (double)gn[q];
        if (g.done || gn[q] == 0) continue;
        double t[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t c = cfirst[q]; c < cfirst[q + 1]; ++c)
            for (int k = 0; k < 4; ++k) t[k] += part[(size_t)c * 4 + k];
        if (!base_only) {
            const uint64_t it = g.iterations + 1;
            g.iterations = it;
            g.prev_l1 = g.l1;
            g.l1 = t[0];
            g.l2 = t[1];
            g.sum = t[2];
            g.sink = t[3];
            const double bound = t[2] > 0.0 ? 2.0 * r.damp / (1.0 - r.damp) * t[0] / t[2] : INFINITY;
            g.bound = bound;
            bool stop = it >= r.max_iterations;
            if (r.kind == kRuleStep) stop = stop || ((r.norm == NKPR_L2) ? sqrt(t[1]) : t[0]) <= r.tol;
            else stop = stop || bound <= r.tol;
            g.done = stop ? 1u : 0u;
        } else {
            g.sink = t[3];
        }
        add[q] = (1.0 - r.damp) / nq + (dist ? r.damp * g.sink / nq : 0.0);
        if (!g.done) ++active;
    }
    atomicAdd(&s_active, active);
    __syncthreads();
    if (threadIdx.x == 0 && !base_only && s_active == 0) {
        *flag = 1u;
        ctl->done = 1u;
    }
}

// b[u] = add[graph of u]
template <typename T>
__global__ void group_base_kernel(const Chunk *ch, const double *add, T *b) {
    const Chunk c = ch[blockIdx.x];
    const T v = (T)add[c.group];
    for (uint32_t u = c.begin + threadIdx.x; u < c.end; u += blockDim.x) b[u] = v;
}

// term[q] = d S_q / n_q from the sink masses the last reduction left in gc
__global__ void group_sink_term_kernel(const GroupCtl *gc, const uint32_t *gn, uint32_t groups,
                                       double damp, double *term) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < groups) term[q] = gn[q] ? damp * gc[q].sink / (double)gn[q] : 0.0;
}

// w[u] = y[u] + term[q]: the union operator B v = d P v with per-graph sink terms
__global__ void group_apply_kernel(const Chunk *ch, const float *y, const double *term, float *w) {
    const Chunk c = ch[blockIdx.x];
    const float t = (float)term[c.group];
    for (uint32_t u = c.begin + threadIdx.x; u < c.end; u += blockDim.x) w[u] = y[u] + t;
}

__global__ void group_scatter_kernel(const Chunk *ch, const double *x, const int32_t *perm,
                                     const double *factor, double *scores, double *mx) {
    const Chunk c = ch[blockIdx.x];
    const double f = factor[c.group];
    double m = -INFINITY;
    for (uint32_t u = c.begin + threadIdx.x; u < c.end; u += blockDim.x) {
        const double v = x[u] / f;
        scores[perm[u]] = v;
        m = fmax(m, v);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    __shared__ double s[1024 / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = -INFINITY;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) t = fmax(t, s[w]);
        mx[blockIdx.x] = t;
    }
}

// first[k] = the score in graph k's first slot (0 for a graph without slots)
__global__ void group_first_kernel(const double *scores, const uint64_t *gslot, uint32_t groups,
                                   double *first) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < groups) first[k] = gslot[k] < gslot[k + 1] ? scores[gslot[k]] : 0.0;
}

__global__ void group_init_kernel(GroupCtl *gc, uint32_t groups, uint32_t *flag) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < groups) {
        gc[q].l1 = INFINITY;
        gc[q].prev_l1 = INFINITY;
        gc[q].bound = INFINITY;
        gc[q].done = 0;
        gc[q].iterations = 0;
    }
    if (q == 0) *flag = 0u;
}

struct GroupRun {
    nkpr_graph *g;
    const nkpr_params *p;
    Chunk *chunks;
    uint32_t nch;
    uint32_t *cfirst, *gn;
    GroupCtl *gc;
    double *part, *add;
    uint32_t *flag;
};

template <typename T>
void post_sweep(GroupRun &q, const T *xold, T *xnew, PhaseRule r, int restore, T *base) {
    nkpr_graph *g = q.g;
    const hipStream_t s = g->stream;
    const int dist = q.p->sink_handling == NKPR_SINK_DISTRIBUTE;
    hipLaunchKernelGGL(group_reduce_kernel<T>, dim3(q.nch), dim3(kRedThreads), 0, s,
                       (const Chunk *)q.chunks, xnew, xold, (const GroupCtl *)q.gc, q.part,
                       (const uint32_t *)q.flag, restore);
    hipLaunchKernelGGL(group_rule_kernel, dim3(1), dim3(1024), 0, s, (const uint32_t *)q.cfirst,
                       (const uint32_t *)q.gn, g->n_groups, (const double *)q.part, q.gc, q.add, r,
                       dist, 0, g->ctl.as<PhaseCtl>(), q.flag);
    if (dist && base)
        hipLaunchKernelGGL(group_base_kernel<T>, dim3(q.nch), dim3(256), 0, s, (const Chunk *)q.chunks,
                           (const double *)q.add, base);
    NKPR_HIP(hipGetLastError());
}

// flag = 1 when x is not constant on some graph (every node against its graph's first node)
__global__ void group_constant_check_kernel(const Chunk *chunks, const uint32_t *cfirst,
                                            const double *x, uint32_t *flag) {
    const Chunk c = chunks[blockIdx.x];
    const double x0 = x[chunks[cfirst[c.group]].begin];
    int diff = 0;
    for (uint32_t u = c.begin + threadIdx.x; u < c.end; u += blockDim.x) diff |= x[u] != x0;
    if (__syncthreads_or(diff) && threadIdx.x == 0) atomicOr(flag, 1u);
}

// The residual sweep from a start that is constant on every graph, in closed form: the
// in-neighbours of u lie in u's graph, so (P x)[u] = x[u] rowsum[u] and y = d x rowsum + base.
__global__ void group_const_resid_kernel(const double *x, const double *rowsum, const double *base,
                                         uint32_t n, double d, double *ynew, float *r32) {
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const double y = d * x[u] * rowsum[u] + base[u];
        ynew[u] = y;
        r32[u] = (float)(y - x[u]);
    }
}

// The right-hand side of the first sweep: sink masses of the start vector.
template <typename T> void start_base(GroupRun &q, const T *x, T *base, double damp) {
    nkpr_graph *g = q.g;
    const hipStream_t s = g->stream;
    hipLaunchKernelGGL(group_init_kernel, dim3(nblk(g->n_groups)), dim3(256), 0, s, q.gc,
                       g->n_groups, q.flag);
    hipLaunchKernelGGL(group_reduce_kernel<T>, dim3(q.nch), dim3(kRedThreads), 0, s,
                       (const Chunk *)q.chunks, const_cast<T *>(x), x, (const GroupCtl *)q.gc,
                       q.part, (const uint32_t *)q.flag, 0);
    PhaseRule r{};
    r.damp = damp;
    hipLaunchKernelGGL(group_rule_kernel, dim3(1), dim3(1024), 0, s, (const uint32_t *)q.cfirst,
                       (const uint32_t *)q.gn, g->n_groups, (const double *)q.part, q.gc, q.add, r,
                       q.p->sink_handling == NKPR_SINK_DISTRIBUTE ? 1 : 0, 1,
                       g->ctl.as<PhaseCtl>(), q.flag);
    if (base)
        hipLaunchKernelGGL(group_base_kernel<T>, dim3(q.nch), dim3(256), 0, s, (const Chunk *)q.chunks,
                           (const double *)q.add, base);
    NKPR_HIP(hipGetLastError());
}

}  // namespace

void run_batch_impl(nkpr_graph *g, const nkpr_params *p, double *scores, uint64_t *iterations,
                    double *maximum, nkpr_stats *st) {
    const auto t0 = std::chrono::steady_clock::now();
    NKPR_CHECK(g != nullptr && p != nullptr, "graph or params is NULL");
    NKPR_CHECK(g->n_groups > 0, "not a batch handle: use nkpr_graph_create_batch");
    NKPR_CHECK(g->z == 0 || scores != nullptr, "scores is NULL");
    NKPR_CHECK(p->damp >= 0.0 && p->damp < 1.0, "damp must be in [0, 1)");
    NKPR_CHECK(p->norm == NKPR_L1 || p->norm == NKPR_L2, "norm must be NKPR_L1 or NKPR_L2");
    NKPR_CHECK(p->sink_handling == NKPR_SINK_NONE || p->sink_handling == NKPR_SINK_DISTRIBUTE,
               "unknown sink handling");
    NKPR_CHECK(p->precision >= NKPR_FP64 && p->precision <= NKPR_FP32, "unknown precision");
    NKPR_CHECK(p->stop == NKPR_STOP_STEP || p->stop == NKPR_STOP_L1_ERROR, "unknown stop rule");
    NKPR_CHECK(!(p->tol < 0.0) && !std::isnan(p->tol), "tol must be >= 0");
    NKPR_CHECK(p->teleport == nullptr, "a batch of graphs takes no teleport vector");
    std::memset(st, 0, sizeof(*st));
    const uint64_t z = g->z;
    const uint32_t n = g->n, G = g->n_groups;
    const hipStream_t s = g->stream;
    const double damp = p->damp;
    if (z == 0) return;

    // ---- chunks and per-graph tables ------------------------------------------------------
    if (!g->q_nchunks && n) {
        std::vector<Chunk> ch;
        std::vector<uint32_t> cfirst(G + 1, 0), gn(G, 0);
        for (uint32_t q = 0; q < G; ++q) {
            cfirst[q] = (uint32_t)ch.size();
            gn[q] = g->goff[q + 1] - g->goff[q];
            for (uint32_t b = g->goff[q]; b < g->goff[q + 1]; b += kChunkNodes)
                ch.push_back({b, std::min(b + kChunkNodes, g->goff[q + 1]), q, g->gsink[q]});
        }
        cfirst[G] = (uint32_t)ch.size();
        g->q_nchunks = (uint32_t)ch.size();
        g->q_chunks.alloc(sizeof(Chunk) * ch.size() + 4 * (2 * (size_t)G + 1));
        NKPR_HIP(hipMemcpyAsync(g->q_chunks.p, ch.data(), sizeof(Chunk) * ch.size(), hipMemcpyHostToDevice, s));
        uint32_t *tail = (uint32_t *)((char *)g->q_chunks.p + sizeof(Chunk) * ch.size());
        NKPR_HIP(hipMemcpyAsync(tail, cfirst.data(), 4 * ((size_t)G + 1), hipMemcpyHostToDevice, s));
        NKPR_HIP(hipMemcpyAsync(tail + G + 1, gn.data(), 4 * (size_t)G, hipMemcpyHostToDevice, s));
        g->q_gslot.alloc(8 * ((size_t)G + 1));
        NKPR_HIP(hipMemcpyAsync(g->q_gslot.p, g->gslot.data(), 8 * ((size_t)G + 1), hipMemcpyHostToDevice, s));
        NKPR_HIP(hipStreamSynchronize(s));
    }
    ensure(g->q_part, 8 * 4 * std::max<size_t>(g->q_nchunks, 1));
    ensure(g->q_ctl, sizeof(GroupCtl) * G);
    ensure(g->q_add, 8 * (size_t)G);
    ensure(g->q_factor, 16 * (size_t)G + 8 * std::max<size_t>(g->q_nchunks, 1));
    ensure(g->q_flag, 16);
    GroupRun q{};
    q.g = g;
    q.p = p;
    q.chunks = g->q_chunks.as<Chunk>();
    q.nch = g->q_nchunks;
    q.cfirst = (uint32_t *)((char *)g->q_chunks.p + sizeof(Chunk) * q.nch);
    q.gn = q.cfirst + G + 1;
    q.gc = g->q_ctl.as<GroupCtl>();
    q.part = g->q_part.as<double>();
    q.add = g->q_add.as<double>();
    q.flag = g->q_flag.as<uint32_t>();

    double *sz = scores;
    if (!p->scores_on_device) {
        ensure(g->zbuf, 8 * z);
        sz = g->zbuf.as<double>();
        NKPR_HIP(hipMemcpyAsync(sz, scores, 8 * z, hipMemcpyHostToDevice, s));
    }
    std::vector<GroupCtl> hc(G);
    if (n) {
        prepare_run(g);
        ensure(g->base64, 8 * (size_t)n);
        ensure(g->base32, 4 * (size_t)n);
        launch_gather_start(sz, g->perm.as<int32_t>(), n, g->x64a.as<double>(),
                            p->precision == NKPR_FP32 ? g->x32a.as<float>() : nullptr, s);
        int cus = 0;
        NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
        RunCtx rc{g, p, st, p->max_iterations, {}, 0};
        rc.combine_grid = (int)std::min<uint64_t>(((uint64_t)g->n_s + kCombineThreads - 1) / kCombineThreads,
                                                  (uint64_t)std::max(cus, 1) * 8);
        const bool folded = g->weighted;
        const bool timing = p->timing != 0;
        const uint64_t max_it = std::max<uint64_t>(p->max_iterations, 1);
        PhaseCtl *ctl = g->ctl.as<PhaseCtl>();
        auto args64 = [&]() {
            SweepArgs<double> a{};
            a.coef = folded ? g->coef64.as<double>() : nullptr;
            a.partial = g->part64.as<double>();
            a.inv = folded ? nullptr : g->inv64.as<double>();
            a.split_partial = g->split_part64.as<double>();
            a.base_vec = g->base64.as<double>();
            a.damp = damp;
            a.n = n;
            a.sink_begin = n;
            return a;
        };
        auto args32 = [&]() {
            SweepArgs<float> a{};
            a.coef = folded ? g->coef32.as<float>() : nullptr;
            a.partial = g->part32.as<float>();
            a.inv = folded ? nullptr : g->inv32.as<float>();
            a.split_partial = g->split_part32.as<float>();
            a.base_vec = g->base32.as<float>();
            a.damp = damp;
            a.n = n;
            a.sink_begin = n;
            return a;
        };
        // the sweep's own finalize only counts sweeps; the per-graph rule stops the phase
        auto counting_rule = [&](uint64_t cap) {
            PhaseRule r{};
            r.kind = kRuleStep;
            r.norm = NKPR_L1;
            r.tol = -1.0;
            r.max_iterations = cap;
            r.damp = damp;
            return r;
        };
        PhaseRule gr{};
        gr.kind = (p->stop == NKPR_STOP_STEP) ? kRuleStep : kRuleErr;
        gr.norm = p->norm;
        gr.tol = p->tol;
        gr.max_iterations = max_it;
        gr.damp = damp;
        const double *result = nullptr;
        uint64_t total = 0;
        if (p->precision != NKPR_MIXED) {
            const bool dbl = p->precision == NKPR_FP64;
            if (dbl) {
                Vecs<double> v{{g->x64a.as<double>(), g->x64b.as<double>()}, {g->xh64a.as<double>(), g->xh64b.as<double>()}};
                start_base<double>(q, v.x[0], g->base64.as<double>(), damp);
                if (!folded) launch_prescale(v.x[0], g->inv64.as<double>(), v.xh[0], nullptr, nullptr, nullptr, n, n, s);
                launch_init_ctl(ctl, nullptr, g->claim.as<uint32_t>(), s);
                std::function<void(const double *, double *)> post = [&](const double *xo, double *xn) {
                    post_sweep<double>(q, xo, xn, gr, 1, g->base64.as<double>());
                };
                total = run_phase<double>(rc, args64(), v, kSweepBaseVec, counting_rule(max_it), max_it, timing, &st->ms_sweeps_fp64, post);
                st->sweeps_fp64 = total;
                result = v.x[total & 1];
            } else {
                Vecs<float> v{{g->x32a.as<float>(), g->x32b.as<float>()}, {g->xh32a.as<float>(), g->xh32b.as<float>()}};
                start_base<float>(q, v.x[0], g->base32.as<float>(), damp);
                if (!folded) launch_prescale(nullptr, nullptr, nullptr, v.x[0], g->inv32.as<float>(), v.xh[0], n, n, s);
                launch_init_ctl(ctl, nullptr, g->claim.as<uint32_t>(), s);
                std::function<void(const float *, float *)> post = [&](const float *xo, float *xn) {
                    post_sweep<float>(q, xo, xn, gr, 1, g->base32.as<float>());
                };
                total = run_phase<float>(rc, args32(), v, kSweepBaseVec, counting_rule(max_it), max_it, timing, &st->ms_sweeps_fp32, post);
                st->sweeps_fp32 = total;
                launch_to_double(v.x[total & 1], g->x64a.as<double>(), n, s);
                result = g->x64a.as<double>();
            }
        } else {
            // FP64 residual sweeps certify every graph; FP32 GMRES on the union corrects
            double *xcur = g->x64a.as<double>();
            double *ynew = g->x64b.as<double>();
            Krylov kr = krylov_setup(g);
            g->op_pending = false;
            const int dist = p->sink_handling == NKPR_SINK_DISTRIBUTE;
            // A v = v - d P v - (per graph) d S_q(v) / n_q
            // w = B v = d P v (+ per-graph sink terms), the GMRES Krylov operator (A = I - B)
            auto apply_b = [&](const float *v, float *w) {
                // d P v through the binned copy of the union on large batches, else one pull sweep
                if (!apply_binned(g, v, dist ? kr.y : w, damp, false, false, timing, &st->ms_sweeps_fp32)) {
                    if (!folded) launch_prescale(nullptr, nullptr, nullptr, v, g->inv32.as<float>(), kr.vh, n, n, s);
                    launch_init_ctl(ctl, nullptr, g->claim.as<uint32_t>(), s);
                    SweepArgs<float> a = args32();
                    a.base_vec = nullptr;
                    a.tele = 0.0;
                    Vecs<float> vv{{const_cast<float *>(v), dist ? kr.y : w}, {kr.vh, kr.yh}};
                    run_phase<float>(rc, a, vv, kSweepPlain, counting_rule(1), 1, timing, &st->ms_sweeps_fp32);
                }
                if (dist) {
                    PhaseRule r0{};
                    r0.damp = damp;
                    hipLaunchKernelGGL(group_init_kernel, dim3(nblk(G)), dim3(256), 0, s, q.gc, G, q.flag);
                    hipLaunchKernelGGL(group_reduce_kernel<float>, dim3(q.nch), dim3(kRedThreads), 0, s,
                                       (const Chunk *)q.chunks, const_cast<float *>(v), v, (const GroupCtl *)q.gc,
                                       q.part, (const uint32_t *)q.flag, 0);
                    hipLaunchKernelGGL(group_rule_kernel, dim3(1), dim3(1024), 0, s, (const uint32_t *)q.cfirst,
                                       (const uint32_t *)q.gn, G, (const double *)q.part, q.gc, q.add, r0, 1, 1, ctl, q.flag);
                    hipLaunchKernelGGL(group_sink_term_kernel, dim3(nblk(G)), dim3(256), 0, s,
                                       (const GroupCtl *)q.gc, (const uint32_t *)q.gn, G, damp, q.add);
                    hipLaunchKernelGGL(group_apply_kernel, dim3(q.nch), dim3(256), 0, s, (const Chunk *)q.chunks,
                                       (const float *)kr.y, (const double *)q.add, w);
                }
                NKPR_HIP(hipGetLastError());
            };
            PhaseRule cr = gr;
            cr.kind = kRuleErr;
            cr.max_iterations = UINT64_MAX;
            // A start that is constant on every graph (every first run: 1/n_q) takes its first
            // residual in closed form from the cached FP64 row sums of P: not counted as a sweep,
            // and it never certifies (the next residual is a full FP64 sweep).
            const char *ccenv = std::getenv("NKPR_CLOSED_START");
            bool closed = !(ccenv && ccenv[0] == '0') && n > 0 && q.nch > 0;
            if (closed) {
                ensure(g->flag_buf, 4);
                uint32_t *flag = g->flag_buf.as<uint32_t>();
                NKPR_HIP(hipMemsetAsync(flag, 0, 4, s));
                hipLaunchKernelGGL(group_constant_check_kernel, dim3(q.nch), dim3(256), 0, s, (const Chunk *)q.chunks,
                                   (const uint32_t *)q.cfirst, (const double *)xcur, flag);
                NKPR_HIP(hipGetLastError());
                uint32_t hf = 1;
                read_back(s, {{&hf, flag, 4}});
                closed = hf == 0;
            }
            if (closed) ensure_rowsum64(g);
            for (int round = 0;; ++round) {
                const bool closed_round = closed && round == 0;
                start_base<double>(q, xcur, g->base64.as<double>(), damp);
                if (!folded && !closed_round)
                    launch_prescale(xcur, g->inv64.as<double>(), g->xh64a.as<double>(), nullptr, nullptr, nullptr, n, n, s);
                launch_init_ctl(ctl, nullptr, g->claim.as<uint32_t>(), s);
                SweepArgs<double> a64 = args64();
                a64.r32 = kr.V;  // the residual is the first Krylov vector
                Vecs<double> vr{{xcur, ynew}, {g->xh64a.as<double>(), g->xh64b.as<double>()}};
                std::function<void(const double *, double *)> post = [&](const double *xo, double *xn) {
                    post_sweep<double>(q, xo, xn, cr, 0, nullptr);
                };
                if (closed_round) {
                    hipLaunchKernelGGL(group_const_resid_kernel, dim3(nblk(n)), dim3(256), 0, s, (const double *)xcur,
                                       (const double *)g->rowsum64.as<double>(), (const double *)g->base64.as<double>(),
                                       n, damp, ynew, kr.V);
                    NKPR_HIP(hipGetLastError());
                    post(xcur, ynew);
                } else {
                    total += run_phase<double>(rc, a64, vr, kSweepBaseResid, counting_rule(1), 1, timing, &st->ms_sweeps_fp64, post);
                    st->sweeps_fp64 += 1;
                }
                result = ynew;
                read_back(s, {{hc.data(), q.gc, sizeof(GroupCtl) * G}});
                // the correction must bring every graph's residual to its own target: the ratio
                // of the worst graph sets the relative tolerance of the union's solve
                double worst = 0.0;
                uint32_t open = 0;
                for (uint32_t k = 0; k < G; ++k) {
                    if (g->goff[k + 1] == g->goff[k]) continue;
                    const double target = 0.5 * p->tol * hc[k].sum * (1.0 - damp) / (2.0 * damp);
                    worst = std::max(worst, hc[k].l1 / std::max(target, 1e-300));
                    open += hc[k].bound <= p->tol ? 0u : 1u;
                }
                if ((open == 0 && !closed_round) || total >= max_it || round >= kMaxRefinements) break;
                ++st->refinements;
                const double rtol = std::min(0.5, std::max(1e-7, 0.5 / std::max(worst, 1e-300)));
                const uint64_t left = (max_it > total + 1) ? max_it - total - 1 : 0;
                const uint64_t kc = gmres_f32(kr, apply_b, kr.V, kr.e, rtol, left, s);
                collect_op_timing(g, &st->ms_sweeps_fp32);
                st->sweeps_fp32 += kc;
                st->sweeps_fp32_correction += kc;
                total += kc;
                launch_add_correction(xcur, kr.e, n, s);
            }
        }
        st->iterations = total;
        read_back(s, {{hc.data(), q.gc, sizeof(GroupCtl) * G}});

        // ---- per-graph normalisation (PageRank.cpp:73-92) and scatter --------------------------
        std::vector<double> factor(G, 1.0);
        for (uint32_t k = 0; k < G; ++k) {
            const uint32_t nq = g->goff[k + 1] - g->goff[k];
            if (!nq) continue;
            factor[k] = p->normalized ? (1.0 / nq) * ((1.0 - damp) + damp * hc[k].sink) : hc[k].sum;
            const double b = (p->precision == NKPR_MIXED || p->stop == NKPR_STOP_L1_ERROR)
                                 ? hc[k].bound
                                 : (p->norm == NKPR_L2 ? std::sqrt(hc[k].l2) : hc[k].l1);
            // an FP32 run certifies nothing (its iterates carry rounding the step does not see)
            st->l1_error_bound = p->precision == NKPR_FP32 ? INFINITY : std::max(st->l1_error_bound, hc[k].bound);
            st->last_step = std::max(st->last_step, b);
        }
        double *fdev = g->q_factor.as<double>();
        double *mdev = fdev + G;
        NKPR_HIP(hipMemcpyAsync(fdev, factor.data(), 8 * (size_t)G, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(group_scatter_kernel, dim3(q.nch), dim3(1024), 0, s, (const Chunk *)q.chunks,
                           result, g->perm.as<int32_t>(), (const double *)fdev, sz, mdev);
        NKPR_HIP(hipGetLastError());
        std::vector<double> cmax(q.nch);
        if (maximum) {
            // maximum(): 1 when not normalised; otherwise the largest score, seeded with the
            // graph's first slot (PageRank.cpp:95-102); the first slots are gathered on the
            // device and read with the chunk maxima in one copy
            double *first_dev = mdev + std::max<size_t>(q.nch, 1);
            hipLaunchKernelGGL(group_first_kernel, dim3((G + 255) / 256), dim3(256), 0, s, (const double *)sz,
                               (const uint64_t *)g->q_gslot.as<uint64_t>(), G, first_dev);
            NKPR_HIP(hipGetLastError());
            std::vector<Chunk> ch(q.nch);
            std::vector<double> first(G, 0.0);
            read_back(s, {{cmax.data(), mdev, 8 * (size_t)q.nch},
                          {ch.data(), q.chunks, sizeof(Chunk) * q.nch},
                          {first.data(), first_dev, 8 * (size_t)G}});
            for (uint32_t k = 0; k < G; ++k) maximum[k] = p->normalized ? first[k] : 1.0;
            if (p->normalized)
                for (uint32_t c = 0; c < q.nch; ++c)
                    maximum[ch[c].group] = std::max(maximum[ch[c].group], cmax[c]);
        }
    } else if (maximum) {
        for (uint32_t k = 0; k < G; ++k) maximum[k] = p->normalized ? 0.0 : 1.0;
    }
    if (iterations)
        for (uint32_t k = 0; k < G; ++k)
            iterations[k] = p->precision == NKPR_MIXED ? st->iterations : hc[k].iterations;
    read_back(s, {{scores, sz, p->scores_on_device ? 0 : 8 * z}});
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace nkpr
