// C ABI of libnkpr.so and the host orchestration of a PageRank run.
//
// nkpr_run restates the control flow of networkit/cpp/centrality/PageRank.cpp:20-104 on the
// device: start vector (:22-27, warm start included), sweeps until the stopping rule holds
// (:53-71), normalisation (:73-92) and maximum() (:95-102).  The sweeps are issued in
// batches; a device-side flag set by the finalize kernel turns the sweeps after convergence
// into no-ops, so the host synchronises once per batch, not once per sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "nkpr_internal.hpp"

namespace nkpr {
void create_graph(const nkpr_csc *c, nkpr_graph *g);
}

using namespace nkpr;

namespace {

thread_local std::string g_last_error;

template <typename F> int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return NKPR_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return NKPR_ERR_NOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return NKPR_ERR_INVALID;
    }
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        NKPR_HIP(hipGetDevice(&prev));
        if (prev != dev) NKPR_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

// Buffers of one precision for a ping-pong phase.
template <typename T> struct Vecs {
    T *x[2];
    T *xh[2];
};

void ensure(DevBuf &b, size_t bytes) { b.alloc(bytes); }

constexpr int kMaxRefinements = 30;

uint32_t env_u32(const char *name, uint32_t dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    return (uint32_t)std::strtoul(v, nullptr, 10);
}

struct RunCtx {
    nkpr_graph *g;
    const nkpr_params *p;
    nkpr_stats *st;
    uint64_t sweeps_left;
    int workers[2][3][2];  // [dbl][kind][folded] workers per slice
    int combine_grid;
};

Schedule make_schedule(nkpr_graph *g, uint32_t workers) {
    Schedule sc{};
    sc.src = g->src.as<int32_t>();
    sc.n_slices = g->n_slices;
    sc.drow = g->drow.as<uint32_t>();
    sc.dptr = g->dptr.as<uint32_t>();
    sc.ditems = g->ditems.as<uint4>();
    sc.n_d = g->n_d;
    sc.n_ditems = g->n_ditems;
    sc.srow = g->srow.as<uint32_t>();
    sc.sptr = g->sptr.as<uint32_t>();
    sc.sitems = g->sitems.as<uint4>();
    for (int q = 0; q <= kSlices; ++q) sc.sitem_off[q] = g->sitem_off[q];
    sc.n_s = g->n_s;
    sc.split_flag = g->split_flag.as<uint8_t>();
    sc.claim = g->claim.as<uint32_t>();
    sc.workers = workers;
    return sc;
}

// Launch sweeps of one phase until `done` or `max_launch` sweeps.  Returns completed sweeps.
template <typename T>
uint64_t run_phase(RunCtx &rc, SweepArgs<T> a, Vecs<T> v, int kind, PhaseRule rule,
                   uint64_t max_launch, bool timing, double *ms_acc) {
    nkpr_graph *g = rc.g;
    const bool folded = g->weighted;
    const bool dbl = sizeof(T) == 8;
    int &workers = rc.workers[dbl][kind][folded];
    if (workers <= 0)
        workers = std::max(1, sweep_resident_blocks(g->device, dbl, folded, kind) / (int)g->n_slices);
    FinalizeArgs f{};
    f.split_rows = g->split_rows.as<uint32_t>();
    f.split_first = g->split_first.as<uint32_t>();
    f.split_count = g->split_count.as<uint32_t>();
    f.n_split_rows = g->n_split_rows;
    f.n_blocks = (uint32_t)(workers * g->n_slices + (g->n_s ? rc.combine_grid : 0));
    f.rule = rule;
    ensure(g->partials, 8 * 4 * (size_t)f.n_blocks);
    a.sc = make_schedule(g, (uint32_t)workers);
    a.partials = g->partials.as<double>();
    a.ctl = g->ctl.as<PhaseCtl>();

    uint64_t launched = 0, done_sweeps = 0;
    uint64_t batch = (max_launch <= 2) ? max_launch : 4;
    const size_t need_ev = 2 * 64;
    if (timing && g->events.size() < need_ev) {
        while (g->events.size() < need_ev) {
            hipEvent_t e;
            NKPR_HIP(hipEventCreate(&e));
            g->events.push_back(e);
        }
    }
    while (launched < max_launch) {
        if (rc.p->interrupt && *rc.p->interrupt)
            throw Error(NKPR_ERR_INTERRUPTED, "interrupted");
        const uint64_t nb = std::min<uint64_t>(std::min<uint64_t>(batch, 64), max_launch - launched);
        for (uint64_t j = 0; j < nb; ++j) {
            const uint64_t k = launched + j;
            const int cur = (int)(k & 1), nxt = cur ^ 1;
            a.xold = v.x[cur];
            a.xnew = v.x[nxt];
            a.xg = folded ? v.x[cur] : v.xh[cur];
            a.xhat_new = folded ? nullptr : v.xh[nxt];
            hipEvent_t e0 = timing ? g->events[2 * j] : nullptr;
            hipEvent_t e1 = timing ? g->events[2 * j + 1] : nullptr;
            launch_sweep<T>(a, f, kind, folded, rc.combine_grid, g->stream, e0, e1);
        }
        NKPR_HIP(hipMemcpyAsync(g->ctl_host, g->ctl.p, sizeof(PhaseCtl), hipMemcpyDeviceToHost,
                                g->stream));
        NKPR_HIP(hipStreamSynchronize(g->stream));
        const uint64_t it = g->ctl_host->iterations;
        const uint64_t executed = it - done_sweeps;
        if (timing) {
            for (uint64_t j = 0; j < executed && j < nb; ++j) {
                float ms = 0.f;
                NKPR_HIP(hipEventElapsedTime(&ms, g->events[2 * j], g->events[2 * j + 1]));
                *ms_acc += ms;
            }
        }
        done_sweeps = it;
        launched += nb;
        if (g->ctl_host->done) break;
        batch *= 2;
    }
    return done_sweeps;
}

double rel_error_bound(double damp, double l1, double sum) {
    return sum > 0.0 ? 2.0 * damp / (1.0 - damp) * l1 / sum : INFINITY;
}

void run_impl(nkpr_graph *g, const nkpr_params *p, double *scores, nkpr_stats *st) {
    const auto t0 = std::chrono::steady_clock::now();
    NKPR_CHECK(g != nullptr && p != nullptr, "graph or params is NULL");
    NKPR_CHECK(g->z == 0 || scores != nullptr, "scores is NULL");
    NKPR_CHECK(p->damp >= 0.0 && p->damp < 1.0, "damp must be in [0, 1)");
    NKPR_CHECK(p->norm == NKPR_L1 || p->norm == NKPR_L2, "norm must be NKPR_L1 or NKPR_L2");
    NKPR_CHECK(p->sink_handling == NKPR_SINK_NONE || p->sink_handling == NKPR_SINK_DISTRIBUTE,
               "unknown sink handling");
    NKPR_CHECK(p->precision >= NKPR_FP64 && p->precision <= NKPR_FP32, "unknown precision");
    NKPR_CHECK(p->stop == NKPR_STOP_STEP || p->stop == NKPR_STOP_L1_ERROR, "unknown stop rule");
    NKPR_CHECK(!(p->tol < 0.0) && !std::isnan(p->tol), "tol must be >= 0");
    NKPR_CHECK(p->teleport == nullptr, "params->teleport is not supported by this build");
    DeviceGuard dg(g->device);
    std::memset(st, 0, sizeof(*st));
    const uint64_t z = g->z;
    const uint32_t n = g->n;
    const hipStream_t s = g->stream;
    if (z == 0) {
        st->maximum = p->normalized ? 0.0 : 1.0;
        return;
    }

    // start vector in slot order on the device
    double *sz;
    if (p->scores_on_device) {
        sz = scores;
    } else {
        ensure(g->zbuf, 8 * z);
        sz = g->zbuf.as<double>();
        NKPR_HIP(hipMemcpyAsync(sz, scores, 8 * z, hipMemcpyHostToDevice, s));
    }
    if (n == 0) {
        // no node: nothing to iterate; the reference's loop would not run meaningfully
        st->maximum = p->normalized ? (p->scores_on_device ? 0.0 : scores[0]) : 1.0;
        return;
    }

    const double damp = p->damp;
    const double tele = (1.0 - damp) / (double)n;
    const bool folded = g->weighted;
    const uint64_t max_it = p->max_iterations;
    ensure(g->ctl, sizeof(PhaseCtl));
    if (!g->ctl_host) NKPR_HIP(hipHostMalloc((void **)&g->ctl_host, sizeof(PhaseCtl), hipHostMallocDefault));
    ensure(g->red_partials, 8 * 1024);
    ensure(g->red_out, 8 * 4);
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    ensure(g->split_part32, 4 * std::max<size_t>(g->n_split_items, 1));
    ensure(g->split_part64, 8 * std::max<size_t>(g->n_split_items, 1));

    ensure(g->part32, 4 * std::max<size_t>((size_t)g->n_slices * g->n_s, 1));
    ensure(g->part64, 8 * std::max<size_t>((size_t)g->n_slices * g->n_s, 1));
    RunCtx rc{g, p, st, max_it, {}, 0};
    rc.combine_grid = (int)std::min<uint64_t>(((uint64_t)g->n_s + kCombineThreads - 1) / kCombineThreads,
                                              (uint64_t)std::max(cus, 1) * 8);
    const bool dist = p->sink_handling == NKPR_SINK_DISTRIBUTE;
    PhaseCtl *ctl = g->ctl.as<PhaseCtl>();

    auto init_phase = [&](const void *x, bool dbl) {
        const double *sink_sum = nullptr;
        if (dist && g->sink_begin < n) {
            if (dbl) launch_range_sum<double>((const double *)x, g->sink_begin, n, g->red_partials.as<double>(), g->red_out.as<double>(), s);
            else launch_range_sum<float>((const float *)x, g->sink_begin, n, g->red_partials.as<double>(), g->red_out.as<double>(), s);
            sink_sum = g->red_out.as<double>();
        }
        launch_init_ctl(ctl, sink_sum, g->claim.as<uint32_t>(), s);
    };

    auto base_args64 = [&]() {
        SweepArgs<double> a{};
        a.coef = folded ? g->coef64.as<double>() : nullptr;
        a.partial = g->part64.as<double>();
        a.inv = folded ? nullptr : g->inv64.as<double>();
        a.split_partial = g->split_part64.as<double>();
        a.damp = damp;
        a.tele = tele;
        a.n = n;
        a.sink_begin = g->sink_begin;
        a.distribute = dist ? 1 : 0;
        return a;
    };
    auto base_args32 = [&]() {
        SweepArgs<float> a{};
        a.coef = folded ? g->coef32.as<float>() : nullptr;
        a.partial = g->part32.as<float>();
        a.inv = folded ? nullptr : g->inv32.as<float>();
        a.split_partial = g->split_part32.as<float>();
        a.damp = damp;
        a.tele = tele;
        a.n = n;
        a.sink_begin = g->sink_begin;
        a.distribute = dist ? 1 : 0;
        return a;
    };

    const double *result = nullptr;  // compact FP64 result
    double res_sum = 0.0, res_sink = 0.0;
    uint64_t total = 0;
    const bool timing = p->timing != 0;

    auto alloc64 = [&]() {
        ensure(g->x64a, 8 * (size_t)n);
        ensure(g->x64b, 8 * (size_t)n);
        if (!folded) {
            ensure(g->xh64a, 8 * (size_t)n);
            ensure(g->xh64b, 8 * (size_t)n);
        }
    };
    auto alloc32 = [&]() {
        ensure(g->x32a, 4 * (size_t)n);
        ensure(g->x32b, 4 * (size_t)n);
        if (!folded) {
            ensure(g->xh32a, 4 * (size_t)n);
            ensure(g->xh32b, 4 * (size_t)n);
        }
    };
    Vecs<double> v64{};
    Vecs<float> v32{};

    if (p->precision == NKPR_FP64 || p->precision == NKPR_FP32) {
        const bool dbl = p->precision == NKPR_FP64;
        PhaseRule rule{};
        rule.kind = (p->stop == NKPR_STOP_STEP) ? kRuleStep : kRuleErr;
        rule.norm = p->norm;
        rule.tol = p->tol;
        rule.max_iterations = max_it;
        rule.damp = damp;
        if (dbl) {
            alloc64();
            v64 = {{g->x64a.as<double>(), g->x64b.as<double>()}, {g->xh64a.as<double>(), g->xh64b.as<double>()}};
            launch_gather_start(sz, g->perm.as<int32_t>(), n, v64.x[0], nullptr, s);
            if (!folded) launch_prescale(v64.x[0], g->inv64.as<double>(), v64.xh[0], nullptr, nullptr, nullptr, n, g->sink_begin, s);
            init_phase(v64.x[0], true);
            total = run_phase<double>(rc, base_args64(), v64, kSweepPlain, rule, std::max<uint64_t>(max_it, 1), timing, &st->ms_sweeps_fp64);
            st->sweeps_fp64 = total;
            result = v64.x[total & 1];
        } else {
            alloc32();
            ensure(g->x64a, 8 * (size_t)n);
            v32 = {{g->x32a.as<float>(), g->x32b.as<float>()}, {g->xh32a.as<float>(), g->xh32b.as<float>()}};
            launch_gather_start(sz, g->perm.as<int32_t>(), n, nullptr, v32.x[0], s);
            if (!folded) launch_prescale(nullptr, nullptr, nullptr, v32.x[0], g->inv32.as<float>(), v32.xh[0], n, g->sink_begin, s);
            init_phase(v32.x[0], false);
            total = run_phase<float>(rc, base_args32(), v32, kSweepPlain, rule, std::max<uint64_t>(max_it, 1), timing, &st->ms_sweeps_fp32);
            st->sweeps_fp32 = total;
            launch_to_double(v32.x[total & 1], g->x64a.as<double>(), n, s);
            result = g->x64a.as<double>();
        }
        res_sum = g->ctl_host->sum;
        res_sink = g->ctl_host->sink_out;
        st->last_step_l1 = g->ctl_host->l1;
        st->last_step = (p->norm == NKPR_L2) ? std::sqrt(g->ctl_host->l2) : g->ctl_host->l1;
        st->l1_error_bound = rel_error_bound(damp, g->ctl_host->l1, g->ctl_host->sum);
    } else {
        // ---- mixed precision: FP64 residual sweeps certify, FP32 GMRES solves corrections ----
        // With r = b - A x from an FP64 sweep (A = I - d P, P with the sink term when sinks
        // are distributed), A e = r is solved approximately in FP32 by GMRES(m) and x += e in
        // FP64.  Only FP64 quantities enter the certificate 2 d/(1-d) ||r||_1 / sum(x + r).
        alloc64();
        ensure(g->r32, 4 * (size_t)n);
        const int mres = (int)std::min<uint32_t>(std::max<uint32_t>(env_u32("NKPR_GMRES_RESTART", 20), 2), 64);
        ensure(g->kbasis, 4 * (size_t)(mres + 1) * n);
        ensure(g->kwork, 4 * (size_t)4 * n);
        ensure(g->kdot, 8 * ((size_t)kDotGroup * 1024 + 128));
        ensure(g->x32b, 4 * (size_t)n);
        double *xcur = g->x64a.as<double>();   // current FP64 iterate
        double *ynew = g->x64b.as<double>();   // FP64 sweep output
        double *xh = g->xh64a.as<double>();
        float *V = g->kbasis.as<float>();
        float *ky = g->kwork.as<float>();            // y = d P v
        float *kvh = ky + n;                         // prescaled v (unweighted)
        float *kyh = ky + 2 * (size_t)n;             // prescaled y (written, unused)
        float *ke = ky + 3 * (size_t)n;              // correction e
        double *dpart = g->kdot.as<double>();
        double *dout = dpart + (size_t)kDotGroup * 1024;
        launch_gather_start(sz, g->perm.as<int32_t>(), n, xcur, nullptr, s);

        auto dots = [&](const float *w, const float *Vb, uint32_t cnt, double *host) {
            launch_multidot(w, Vb, n, cnt, n, dpart, dout, s);
            NKPR_HIP(hipMemcpyAsync(host, dout, 8 * cnt, hipMemcpyDeviceToHost, s));
            NKPR_HIP(hipStreamSynchronize(s));
        };
        // w = A v = v - d P v  (one FP32 sweep with no teleport term)
        auto apply = [&](const float *v, float *w) {
            if (!folded) launch_prescale(nullptr, nullptr, nullptr, v, g->inv32.as<float>(), kvh, n, g->sink_begin, s);
            init_phase(v, false);
            SweepArgs<float> a = base_args32();
            a.tele = 0.0;
            Vecs<float> vv{{const_cast<float *>(v), ky}, {kvh, kyh}};
            PhaseRule r{};
            r.kind = kRuleStep;
            r.norm = NKPR_L1;
            r.tol = -1.0;
            r.max_iterations = 1;
            r.damp = damp;
            run_phase<float>(rc, a, vv, kSweepPlain, r, 1, timing, &st->ms_sweeps_fp32);
            launch_axpby(v, ky, w, n, 1.0f, -1.0f, s);
        };
        // GMRES(mres) from e = 0 on A e = rhs until ||A e - rhs||_2 <= rtol ||rhs||_2
        auto gmres = [&](const float *rhs, float *e, double rtol, uint64_t cap) -> uint64_t {
            double beta2 = 0.0;
            dots(rhs, rhs, 1, &beta2);
            const double beta = std::sqrt(beta2);
            if (!(beta > 0.0) || cap == 0) {
                launch_axpby(rhs, nullptr, e, n, 0.0f, 0.0f, s);
                return 0;
            }
            launch_axpby(rhs, nullptr, V, n, (float)(1.0 / beta), 0.0f, s);
            std::vector<double> R((size_t)mres * mres, 0.0), cs(mres), sn(mres), gv(mres + 1, 0.0);
            gv[0] = beta;
            uint64_t applies = 0;
            int k = 0;
            for (int j = 0; j < mres && applies < cap; ++j) {
                float *vj = V + (size_t)j * n;
                float *w = V + (size_t)(j + 1) * n;
                apply(vj, w);
                ++applies;
                std::vector<double> h(j + 2), h2(j + 1);
                // classical Gram-Schmidt, applied twice
                dots(w, V, j + 1, h.data());
                launch_multiaxpy(w, V, n, j + 1, n, h.data(), false, s);
                dots(w, V, j + 1, h2.data());
                launch_multiaxpy(w, V, n, j + 1, n, h2.data(), false, s);
                for (int i = 0; i <= j; ++i) h[i] += h2[i];
                double hn2 = 0.0;
                dots(w, w, 1, &hn2);
                const double hn = std::sqrt(hn2);
                h[j + 1] = hn;
                for (int i = 0; i < j; ++i) {
                    const double t = cs[i] * h[i] + sn[i] * h[i + 1];
                    h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                    h[i] = t;
                }
                const double den = std::hypot(h[j], h[j + 1]);
                cs[j] = den > 0.0 ? h[j] / den : 1.0;
                sn[j] = den > 0.0 ? h[j + 1] / den : 0.0;
                h[j] = den;
                gv[j + 1] = -sn[j] * gv[j];
                gv[j] = cs[j] * gv[j];
                for (int i = 0; i <= j; ++i) R[(size_t)i * mres + j] = h[i];
                k = j + 1;
                if (hn > 0.0) launch_axpby(w, nullptr, w, n, (float)(1.0 / hn), 0.0f, s);
                if (std::fabs(gv[j + 1]) <= rtol * beta || !(hn > 0.0)) break;
            }
            std::vector<double> y(k, 0.0);
            for (int i = k - 1; i >= 0; --i) {
                double t = gv[i];
                for (int q = i + 1; q < k; ++q) t -= R[(size_t)i * mres + q] * y[q];
                y[i] = R[(size_t)i * mres + i] != 0.0 ? t / R[(size_t)i * mres + i] : 0.0;
            }
            launch_multiaxpy(e, V, n, k, n, y.data(), true, s);
            return applies;
        };

        for (int round = 0;; ++round) {
            // FP64 residual sweep: ynew = d P xcur + b, r32 = float(ynew - xcur)
            if (!folded) launch_prescale(xcur, g->inv64.as<double>(), xh, nullptr, nullptr, nullptr, n, g->sink_begin, s);
            init_phase(xcur, true);
            PhaseRule rr{};
            rr.kind = kRuleErr;
            rr.norm = NKPR_L1;
            rr.tol = p->tol;
            rr.max_iterations = 1;
            rr.damp = damp;
            SweepArgs<double> a64 = base_args64();
            a64.r32 = g->r32.as<float>();
            Vecs<double> vr{{xcur, ynew}, {xh, g->xh64b.as<double>()}};
            const uint64_t k = run_phase<double>(rc, a64, vr, kSweepResid, rr, 1, timing, &st->ms_sweeps_fp64);
            st->sweeps_fp64 += k;
            total += k;
            const PhaseCtl c = *g->ctl_host;
            st->last_step_l1 = c.l1;
            st->last_step = c.l1;
            st->l1_error_bound = c.bound;
            res_sum = c.sum;
            res_sink = c.sink_out;
            result = ynew;
            if (c.bound <= p->tol || total >= max_it || round >= kMaxRefinements) break;
            // FP32 correction: A e = r, tightened so that the next residual certifies
            ++st->refinements;
            const double target = 0.5 * p->tol * c.sum * (1.0 - damp) / (2.0 * damp);
            const double rtol = std::min(0.5, std::max(1e-7, 0.5 * target / std::max(c.l1, 1e-300)));
            const uint64_t left = (max_it > total + 1) ? max_it - total - 1 : 0;  // keep one FP64 sweep
            const uint64_t kc = gmres(g->r32.as<float>(), ke, rtol, left);
            st->sweeps_fp32 += kc;
            st->sweeps_fp32_correction += kc;
            total += kc;
            launch_add_correction(xcur, ke, n, s);
        }
    }
    st->iterations = total;
    st->score_sum = res_sum;

    // ---- normalisation (PageRank.cpp:73-92) and scatter to slot order -----------------------
    double factor;
    if (!p->normalized) factor = res_sum;
    else factor = (1.0 / (double)n) * ((1.0 - damp) + damp * res_sink);
    launch_scatter_result(result, g->perm.as<int32_t>(), n, factor, sz, s);
    double maximum = 1.0;
    if (p->normalized) {
        launch_range_max(result, n, g->red_partials.as<double>(), g->red_out.as<double>(), s);
        double mx = 0.0, s0 = 0.0;
        NKPR_HIP(hipMemcpyAsync(&mx, g->red_out.p, 8, hipMemcpyDeviceToHost, s));
        NKPR_HIP(hipMemcpyAsync(&s0, sz, 8, hipMemcpyDeviceToHost, s));
        NKPR_HIP(hipStreamSynchronize(s));
        maximum = std::max(mx / factor, s0);
    }
    st->maximum = maximum;
    if (!p->scores_on_device) NKPR_HIP(hipMemcpyAsync(scores, sz, 8 * z, hipMemcpyDeviceToHost, s));
    NKPR_HIP(hipStreamSynchronize(s));
    st->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int nkpr_abi_version(void) { return NKPR_ABI_VERSION; }

const char *nkpr_last_error(void) { return g_last_error.c_str(); }

int nkpr_params_default(nkpr_params *p) {
    return guarded([&] {
        NKPR_CHECK(p != nullptr, "params is NULL");
        std::memset(p, 0, sizeof(*p));
        p->damp = 0.85;
        p->tol = 1e-8;
        p->max_iterations = std::numeric_limits<uint64_t>::max();
        p->norm = NKPR_L2;
        p->sink_handling = NKPR_SINK_NONE;
        p->normalized = 0;
        p->precision = NKPR_FP64;
        p->stop = NKPR_STOP_STEP;
    });
}

int nkpr_graph_create(const nkpr_csc *csc, nkpr_graph **out) {
    return guarded([&] {
        NKPR_CHECK(out != nullptr, "out is NULL");
        *out = nullptr;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw Error(NKPR_ERR_NODEVICE, "no HIP device is visible");
        int dev = 0;
        NKPR_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        NKPR_HIP(hipGetDeviceProperties(&prop, dev));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw Error(NKPR_ERR_NODEVICE, std::string("libnkpr is built for gfx950, device is ") + prop.gcnArchName);
        auto *g = new nkpr_graph();
        g->device = dev;
        try {
            NKPR_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
            create_graph(csc, g);
        } catch (...) {
            nkpr_graph_destroy(g);
            throw;
        }
        *out = g;
    });
}

int nkpr_graph_destroy(nkpr_graph *g) {
    if (!g) return NKPR_OK;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (hipEvent_t e : g->events) (void)hipEventDestroy(e);
    if (g->ctl_host) (void)hipHostFree(g->ctl_host);
    if (g->b_ctl_host) (void)hipHostFree(g->b_ctl_host);
    hipStream_t s = g->stream;
    delete g;  // frees device buffers
    if (s) (void)hipStreamDestroy(s);
    if (prev >= 0) (void)hipSetDevice(prev);
    return NKPR_OK;
}

int nkpr_graph_get_info(const nkpr_graph *g, nkpr_graph_info *info) {
    return guarded([&] {
        NKPR_CHECK(g != nullptr && info != nullptr, "graph or info is NULL");
        std::memset(info, 0, sizeof(*info));
        info->z = g->z;
        info->n = g->n;
        info->m = g->m;
        info->sinks = g->n - g->sink_begin;
        info->max_in_degree = g->max_in_degree;
        info->work_items = g->n_items;
        info->device_bytes = g->perm.bytes + g->colptr.bytes + g->src.bytes + g->coef32.bytes +
                             g->coef64.bytes + g->inv32.bytes + g->inv64.bytes + g->ditems.bytes +
                             g->sitems.bytes + g->sptr.bytes + g->drow.bytes + g->dptr.bytes +
                             g->srow.bytes;
        info->weighted = g->weighted ? 1 : 0;
        info->device = g->device;
        info->prep_ms = g->prep_ms;
    });
}

void *nkpr_graph_stream(const nkpr_graph *g) { return g ? (void *)g->stream : nullptr; }

int nkpr_run(nkpr_graph *g, const nkpr_params *p, double *scores, nkpr_stats *stats) {
    nkpr_stats local;
    return guarded([&] { run_impl(g, p, scores, stats ? stats : &local); });
}

int nkpr_ppr_batch(nkpr_graph *g, const nkpr_params *p, uint32_t k,
                   const uint64_t *seed_offsets, const int32_t *seed_nodes,
                   const double *seed_weights, double *scores, nkpr_batch_stats *stats,
                   uint64_t *col_iterations, double *col_error_bound) {
    nkpr_batch_stats local;
    return guarded([&] {
        NKPR_CHECK(g != nullptr, "graph is NULL");
        DeviceGuard dg(g->device);
        ppr_batch_impl(g, p, k, seed_offsets, seed_nodes, seed_weights, scores,
                       stats ? stats : &local, col_iterations, col_error_bound);
    });
}

int nkpr_ranking(uint64_t z, const double *scores, const uint8_t *exists, int64_t *ids_out,
                 double *scores_out, uint64_t *count) {
    return guarded([&] {
        NKPR_CHECK(z == 0 || (scores != nullptr && ids_out != nullptr), "NULL argument");
        std::vector<int64_t> ids;
        ids.reserve(z);
        for (uint64_t u = 0; u < z; ++u)
            if (!exists || exists[u]) ids.push_back((int64_t)u);
        std::stable_sort(ids.begin(), ids.end(), [&](int64_t a, int64_t b) {
            if (scores[a] == scores[b]) return a < b;
            return scores[a] > scores[b];
        });
        std::copy(ids.begin(), ids.end(), ids_out);
        if (scores_out)
            for (size_t i = 0; i < ids.size(); ++i) scores_out[i] = scores[ids[i]];
        if (count) *count = ids.size();
    });
}

}  // extern "C"
