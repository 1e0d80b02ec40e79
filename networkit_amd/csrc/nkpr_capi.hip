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
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "nkpr_internal.hpp"
#include "nkpr_run.hpp"

namespace nkpr {
void create_graph(const nkpr_csc *c, nkpr_graph *g, const uint64_t *group_off = nullptr,
                  uint32_t n_groups = 0);
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

}  // namespace

namespace nkpr {

void prepare_run(nkpr_graph *g) {
    ensure(g->ctl, sizeof(PhaseCtl));
    if (!g->ctl_host) NKPR_HIP(hipHostMalloc((void **)&g->ctl_host, sizeof(PhaseCtl), hipHostMallocDefault));
    ensure(g->red_partials, 8 * 1024);
    ensure(g->red_out, 8 * 4);
    ensure(g->split_part32, 4 * std::max<size_t>(g->n_split_items, 1));
    ensure(g->split_part64, 8 * std::max<size_t>(g->n_split_items, 1));
    ensure(g->part32, 4 * std::max<size_t>((size_t)g->n_slices * g->n_s, 1));
    ensure(g->part64, 8 * std::max<size_t>((size_t)g->n_slices * g->n_s, 1));
    const size_t n = g->n;
    ensure(g->x64a, 8 * n);
    ensure(g->x64b, 8 * n);
    ensure(g->x32a, 4 * n);
    ensure(g->x32b, 4 * n);
    if (!g->weighted) {
        ensure(g->xh64a, 8 * n);
        ensure(g->xh64b, 8 * n);
        ensure(g->xh32a, 4 * n);
        ensure(g->xh32b, 4 * n);
    }
}

const int32_t *slot_index(nkpr_graph *g) {
    if (!g->slot_ready) {
        g->slot_u.alloc(4 * std::max<uint64_t>(g->z, 1));
        launch_slot_index(g->perm.as<int32_t>(), g->n, g->z, g->slot_u.as<int32_t>(), g->stream);
        g->slot_ready = true;
    }
    return g->slot_u.as<int32_t>();
}

bool closed_start_enabled() {
    const char *e = std::getenv("NKPR_CLOSED_START");
    return !(e && e[0] == '0');
}

// One solve on compact vectors.  On entry g->x64a holds the FP64 start vector (and g->x32a its
// FP32 copy for NKPR_FP32), unless `start_const` is given: then the start is that constant on
// every node and x64a is not read (mixed precision without distributed sinks, closed-form start
// enabled).  base64 / base32 (compact, optional) replace the constant teleport (1-d)/n by a
// per-node right-hand side.  Returns the unnormalised FP64 result and the FP64 sums of its last
// sweep.
SolveResult solve_core(nkpr_graph *g, const nkpr_params *p, const double *base64,
                       const float *base32, nkpr_stats *st, const double *start_const) {
    const uint32_t n = g->n;
    const hipStream_t s = g->stream;
    const double damp = p->damp;
    const bool has_base = base64 != nullptr;
    const double tele = has_base ? 0.0 : (1.0 - damp) / (double)n;
    const bool folded = g->weighted;
    const uint64_t max_it = p->max_iterations;
    prepare_run(g);
    int cus = 0;
    NKPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    RunCtx rc{g, p, st, max_it, {}, 0};
    rc.combine_grid = (int)std::min<uint64_t>(((uint64_t)g->n_s + kCombineThreads - 1) / kCombineThreads,
                                              (uint64_t)std::max(cus, 1) * 8);
    const bool dist = p->sink_handling == NKPR_SINK_DISTRIBUTE;
    PhaseCtl *ctl = g->ctl.as<PhaseCtl>();
    const int kind_plain = has_base ? kSweepBaseVec : kSweepPlain;
    const int kind_resid = has_base ? kSweepBaseResid : kSweepResid;

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
        a.base_vec = base64;
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
        a.base_vec = base32;
        a.damp = damp;
        a.tele = tele;
        a.n = n;
        a.sink_begin = g->sink_begin;
        a.distribute = dist ? 1 : 0;
        return a;
    };

    SolveResult res{};
    uint64_t total = 0;
    const bool timing = p->timing != 0;
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
            v64 = {{g->x64a.as<double>(), g->x64b.as<double>()}, {g->xh64a.as<double>(), g->xh64b.as<double>()}};
            if (!folded) launch_prescale(v64.x[0], g->inv64.as<double>(), v64.xh[0], nullptr, nullptr, nullptr, n, g->sink_begin, s);
            init_phase(v64.x[0], true);
            total = run_phase<double>(rc, base_args64(), v64, kind_plain, rule, std::max<uint64_t>(max_it, 1), timing, &st->ms_sweeps_fp64);
            st->sweeps_fp64 = total;
            res.x = v64.x[total & 1];
        } else {
            v32 = {{g->x32a.as<float>(), g->x32b.as<float>()}, {g->xh32a.as<float>(), g->xh32b.as<float>()}};
            if (!folded) launch_prescale(nullptr, nullptr, nullptr, v32.x[0], g->inv32.as<float>(), v32.xh[0], n, g->sink_begin, s);
            init_phase(v32.x[0], false);
            total = run_phase<float>(rc, base_args32(), v32, kind_plain, rule, std::max<uint64_t>(max_it, 1), timing, &st->ms_sweeps_fp32);
            st->sweeps_fp32 = total;
            launch_to_double(v32.x[total & 1], g->x64a.as<double>(), n, s);
            res.x = g->x64a.as<double>();
        }
        res.sum = g->ctl_host->sum;
        res.sink = g->ctl_host->sink_out;
        st->last_step_l1 = g->ctl_host->l1;
        st->last_step = (p->norm == NKPR_L2) ? std::sqrt(g->ctl_host->l2) : g->ctl_host->l1;
        // FP32 iterates carry rounding errors that the step does not see: an FP32 run may stop
        // on the same formula, but it certifies nothing
        st->l1_error_bound = dbl ? rel_error_bound(damp, g->ctl_host->l1, g->ctl_host->sum) : INFINITY;
    } else {
        // ---- mixed precision: FP64 residual sweeps certify, FP32 GMRES solves corrections ----
        // With r = b - A x from an FP64 sweep (A = I - d P, P with the sink term when sinks
        // are distributed), A e = r is solved approximately in FP32 by GMRES(m) and x += e in
        // FP64.  Only FP64 quantities enter the certificate 2 d/(1-d) ||r||_1 / sum(x + r).
        double *xcur = g->x64a.as<double>();   // current FP64 iterate
        double *ynew = g->x64b.as<double>();   // FP64 sweep output
        double *xh = g->xh64a.as<double>();
        Krylov kr = krylov_setup(g);
        g->op_pending = false;
        auto apply_b = [&](const float *v, float *w) {
            apply_operator(rc, v, w, kr, timing, &st->ms_sweeps_fp32, false);
        };
        // Without sink handling the sinks (ids [sink_begin, n), no out-edges) are zero columns
        // of P: the other rows form a closed system, and each sink's correction follows from
        // them in the products GMRES computes anyway.  The Krylov vectors then span the head
        // [0, nh) only, nh = sink_begin rounded up to 64 entries (a few sink rows in the head
        // are harmless: their columns are zero too), which shrinks every Gram-Schmidt pass.
        // NKPR_SINK_SPLIT=0 keeps full-length vectors.
        Krylov krh = kr;
        {
            const char *se = std::getenv("NKPR_SINK_SPLIT");
            const uint64_t nh = ((uint64_t)g->sink_begin + 63) / 64 * 64;
            if (!dist && !(se && se[0] == '0') && nh < n) krh.n = (uint32_t)nh;
            if (std::getenv("NKPR_DEBUG"))
                std::fprintf(stderr, "[nkpr] Krylov vectors: %u of %u entries\n", krh.n, n);
        }
        // Through the binned FP64 copy, the residual sweeps keep each contribution c = coef * x
        // to within kCompact64Rel |c|, so y = d P x + b is off by at most d kCompact64Rel
        // ||x||_1 (the column sums of P are at most 1) with ||x||_1 <= sum(y) + ||r||_1.  That
        // error enters the certificate twice, through r and through the returned y: the bound
        // grows by (2 d / (1 - d) + 1) d kCompact64Rel (1 + ||r||_1 / sum(y)), about 2e-11 at
        // d = 0.85 (1% added for the rounding of these FP64 formulas).
        // NKPR_COMPACT64=0 keeps 8-byte contributions
        const char *cenv = std::getenv("NKPR_COMPACT64");
        const bool compact = binned64_active(g) && !(cenv && cenv[0] == '0');
        const double slack = compact ? (2.0 * damp / (1.0 - damp) + 1.0) * damp * kCompact64Rel * 1.01 : 0.0;
        // A run that starts from a constant vector x = c (every first run does: 1/n) begins
        // with y = d P x + b = d c rowsum + b in closed form, from the cached FP64 row sums of
        // P: an O(n) pass instead of an O(m) sweep.  It is not counted as a sweep, and it never
        // certifies (the next residual is a full FP64 sweep).
        NKPR_CHECK(!start_const || (!dist && n > 0 && closed_start_enabled()),
                   "a constant start needs the closed-form first residual");
        bool closed = closed_start_enabled() && n > 0;
        double cval = start_const ? *start_const : 0.0;
        if (closed && !start_const) {
            ensure(g->flag_buf, 4);
            uint32_t *flag = g->flag_buf.as<uint32_t>();
            launch_constant_check(xcur, n, flag, s);
            uint32_t hf = 1;
            read_back(s, {{&hf, flag, 4}, {&cval, xcur, 8}});
            closed = hf == 0;
        }
        if (closed) ensure_rowsum64(g);
        for (int round = 0;; ++round) {
            // FP64 residual sweep: ynew = d P xcur + b, r32 = float(ynew - xcur)
            PhaseRule rr{};
            rr.kind = kRuleErr;
            rr.norm = NKPR_L1;
            rr.tol = p->tol;
            rr.max_iterations = 1;
            rr.damp = damp;
            SweepArgs<double> a64 = base_args64();
            a64.r32 = kr.V;  // the residual is the first Krylov vector
            a64.compact64 = compact ? 1 : 0;
            const bool closed_round = closed && round == 0;
            if (closed_round) {
                init_phase(xcur, true);
                const double sinks = (double)(n - g->sink_begin);
                const double rhs = tele + (dist ? damp * cval * sinks / (double)n : 0.0);
                ensure(g->partials, 8 * 4 * 2048);
                const unsigned blocks = launch_const_resid(g->rowsum64.as<double>(), n, damp * cval, rhs,
                                                           base64, cval, g->sink_begin, ynew, kr.V,
                                                           g->partials.as<double>(), s);
                SweepArgs<double> fa = a64;
                fa.partials = g->partials.as<double>();
                fa.ctl = ctl;
                FinalizeArgs ff{};
                ff.n_blocks = blocks;
                ff.rule = rr;
                launch_finalize64(fa, ff, kind_resid, s);
                read_back(s, {{g->ctl_host, g->ctl.p, sizeof(PhaseCtl)}});
            } else {
                if (!folded) launch_prescale(xcur, g->inv64.as<double>(), xh, nullptr, nullptr, nullptr, n, g->sink_begin, s);
                init_phase(xcur, true);
                Vecs<double> vr{{xcur, ynew}, {xh, g->xh64b.as<double>()}};
                const uint64_t k = run_phase<double>(rc, a64, vr, kind_resid, rr, 1, timing, &st->ms_sweeps_fp64);
                st->sweeps_fp64 += k;
                total += k;
            }
            const PhaseCtl c = *g->ctl_host;
            const double bound = c.bound + (c.sum > 0.0 ? slack * (1.0 + c.l1 / c.sum) : 0.0);
            st->last_step_l1 = c.l1;
            st->last_step = c.l1;
            st->l1_error_bound = bound;
            res.sum = c.sum;
            res.sink = c.sink_out;
            res.x = ynew;
            if ((bound <= p->tol && !closed_round) || total >= max_it || round >= kMaxRefinements) break;
            // FP32 correction: A e = r, tightened so that the next residual certifies
            ++st->refinements;
            const double target = 0.5 * p->tol * c.sum * (1.0 - damp) / (2.0 * damp);
            const double rtol = std::min(0.5, std::max(1e-7, 0.5 * target / std::max(c.l1, 1e-300)));
            const uint64_t left = (max_it > total + 1) ? max_it - total - 1 : 0;  // keep one FP64 sweep
            const uint64_t kc = gmres_f32(krh, apply_b, krh.V, krh.e, rtol, left, s, n);
            collect_op_timing(g, &st->ms_sweeps_fp32);
            if (std::getenv("NKPR_DEBUG"))
                std::fprintf(stderr, "[nkpr] refinement %d: residual l1 %.3e, bound %.3e, rtol %.3e, %llu applications\n",
                             round, c.l1, c.bound, rtol, (unsigned long long)kc);
            st->sweeps_fp32 += kc;
            st->sweeps_fp32_correction += kc;
            total += kc;
            if (start_const && round == 0) launch_set_correction(xcur, cval, kr.e, n, s);
            else launch_add_correction(xcur, kr.e, n, s);
        }
    }
    res.total = total;
    st->iterations = total;
    st->score_sum = res.sum;
    return res;
}

// The FP64 row sums of P (sum of each row's in-coefficients), kept per graph until new weights.
void ensure_rowsum64(nkpr_graph *g) {
    if (g->rowsum_ready) return;
    const uint32_t n = g->n;
    batch_layout(g);
    g->rowsum64.alloc(8 * std::max<uint64_t>(n, 1));
    launch_rowsum64(g->b_rowptr.as<uint32_t>(), n, g->b_src.as<int32_t>(),
                    g->weighted ? g->b_coef64.as<double>() : nullptr,
                    g->weighted ? nullptr : g->inv64.as<double>(), g->rowsum64.as<double>(), g->stream);
    NKPR_HIP(hipStreamSynchronize(g->stream));
    g->rowsum_ready = true;
}

Krylov krylov_setup(nkpr_graph *g) {
    const size_t n = g->n;
    ensure(g->r32, 4 * n);
    Krylov k{};
    // restart length: the Gram-Schmidt cost of a cycle grows with its square, so long cycles
    // only pay where convergence needs them.  Measured in alternating processes: a slowly mixing
    // graph (random hyperbolic, n = 10M) 24.6 ms per solve with 10, 24.7-24.8 with 15, 26.2-26.4
    // with 12 and 27.2 with 20; a batch of 512 Erdos-Renyi graphs 47.4-47.5 ms with 10 and
    // 48.4-48.7 with 12 or 15; the R-MAT benchmark's refinements end within 10 steps
    k.m = (int)std::min<uint32_t>(std::max<uint32_t>(env_u32("NKPR_GMRES_RESTART", 10), 2), kMaxKrylov);
    const size_t ld = (n + 63) / 64 * 64;  // every vector on a 256-byte boundary
    ensure(g->kbasis, 4 * (size_t)(k.m + 1) * ld);
    ensure(g->kwork, 4 * (size_t)4 * ld);
    ensure(g->kdot, 8 * ((size_t)kDotPartials + kDotResults));
    k.n = g->n;
    k.ld = ld;
    k.V = g->kbasis.as<float>();
    k.y = g->kwork.as<float>();
    k.vh = k.y + ld;
    k.yh = k.y + 2 * ld;
    k.e = k.y + 3 * ld;
    k.dpart = g->kdot.as<double>();
    k.dout = k.dpart + (size_t)kDotPartials;
    if (!g->kdot_host) NKPR_HIP(hipHostMalloc((void **)&g->kdot_host, 8 * kDotResults, hipHostMallocDefault));
    if (!g->kdot_ev) NKPR_HIP(hipEventCreateWithFlags(&g->kdot_ev, hipEventDisableTiming));
    k.dhost = g->kdot_host;
    k.ev = g->kdot_ev;
    return k;
}

// w = B v (+ d S(v)/n when `distribute`) or w = A v = v - B v (minus) through the binned copy,
// when the graph uses one (false otherwise: nothing launched).
bool apply_binned(nkpr_graph *g, const float *v, float *w, double damp, bool distribute,
                  bool minus, bool timing, double *ms_acc) {
    if (!(binned_enabled(g) && binned_layout(g))) return false;
    if (timing) {
        // the previous application's time first (its events are done by now: the host has
        // read the dots computed from its output), so that no application waits on the host
        collect_op_timing(g, ms_acc);
        for (hipEvent_t &e : g->op_ev)
            if (!e) NKPR_HIP(hipEventCreate(&e));
    }
    binned_apply(g, v, w, damp, distribute, minus, timing ? g->op_ev[0] : nullptr,
                 timing ? g->op_ev[1] : nullptr);
    g->op_pending = timing;
    return true;
}

void collect_op_timing(nkpr_graph *g, double *ms_acc) {
    if (!g->op_pending) return;
    g->op_pending = false;
    float ms = 0.f;
    NKPR_HIP(hipEventSynchronize(g->op_ev[1]));
    NKPR_HIP(hipEventElapsedTime(&ms, g->op_ev[0], g->op_ev[1]));
    *ms_acc += ms;
}

// w = A v = v - B v (minus) or w = B v, B v = d P v: one FP32 sweep with no teleport term
// (sinks per the run's rule).
void apply_operator(RunCtx &rc, const float *v, float *w, const Krylov &kr, bool timing,
                    double *ms_acc, bool minus) {
    nkpr_graph *g = rc.g;
    const hipStream_t s = g->stream;
    const uint32_t n = g->n;
    const bool folded = g->weighted;
    const bool dist = rc.p->sink_handling == NKPR_SINK_DISTRIBUTE;
    if (apply_binned(g, v, w, rc.p->damp, dist, minus, timing, ms_acc)) return;
    if (!folded) launch_prescale(nullptr, nullptr, nullptr, v, g->inv32.as<float>(), kr.vh, n, g->sink_begin, s);
    const double *sink_sum = nullptr;
    if (dist && g->sink_begin < n) {
        launch_range_sum<float>(v, g->sink_begin, n, g->red_partials.as<double>(), g->red_out.as<double>(), s);
        sink_sum = g->red_out.as<double>();
    }
    launch_init_ctl(g->ctl.as<PhaseCtl>(), sink_sum, g->claim.as<uint32_t>(), s);
    SweepArgs<float> a{};
    a.coef = folded ? g->coef32.as<float>() : nullptr;
    a.partial = g->part32.as<float>();
    a.inv = folded ? nullptr : g->inv32.as<float>();
    a.split_partial = g->split_part32.as<float>();
    a.damp = rc.p->damp;
    a.tele = 0.0;
    a.n = n;
    a.sink_begin = g->sink_begin;
    a.distribute = dist ? 1 : 0;
    // the sweep writes B v to kr.y (A v follows) or straight to w
    Vecs<float> vv{{const_cast<float *>(v), minus ? kr.y : w}, {kr.vh, kr.yh}};
    PhaseRule r{};
    r.kind = kRuleStep;
    r.norm = NKPR_L1;
    r.tol = -1.0;
    r.max_iterations = 1;
    r.damp = rc.p->damp;
    run_phase<float>(rc, a, vv, kSweepPlain, r, 1, timing, ms_acc);
    if (minus) launch_axpby(v, kr.y, w, n, 1.0f, -1.0f, s);
}

}  // namespace nkpr

namespace {

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
    prepare_run(g);
    // A mixed run from a constant start (every first run: 1/n) needs no start vector in compact
    // order: its first residual is taken in closed form and its first correction is written as
    // c + e.  The check reads the slot-order scores when every slot holds a node (then they are
    // the values the gather would move), and runs only without distributed sinks (their sums
    // read the start vector).
    double cval = 0.0;
    bool start_const = false;
    if (p->precision == NKPR_MIXED && z == n && p->sink_handling != NKPR_SINK_DISTRIBUTE &&
        closed_start_enabled()) {
        ensure(g->flag_buf, 4);
        uint32_t *flag = g->flag_buf.as<uint32_t>();
        launch_constant_check(sz, n, flag, s);
        uint32_t hf = 1;
        read_back(s, {{&hf, flag, 4}, {&cval, sz, 8}});
        start_const = hf == 0;
    }
    if (!start_const)
        launch_gather_start(sz, g->perm.as<int32_t>(), n, g->x64a.as<double>(),
                            p->precision == NKPR_FP32 ? g->x32a.as<float>() : nullptr, s);

    // personalised teleport: b[u] = (1 - d) t[slot of u]
    const double *b64 = nullptr;
    const float *b32 = nullptr;
    if (p->teleport) {
        const double *tz = p->teleport;
        if (!p->scores_on_device) {
            ensure(g->tbuf, 8 * z);
            NKPR_HIP(hipMemcpyAsync(g->tbuf.p, p->teleport, 8 * z, hipMemcpyHostToDevice, s));
            tz = g->tbuf.as<double>();
        }
        ensure(g->base64, 8 * (size_t)n);
        ensure(g->base32, 4 * (size_t)n);
        launch_gather_scaled(tz, g->perm.as<int32_t>(), n, 1.0 - damp, g->base64.as<double>(),
                             g->base32.as<float>(), s);
        b64 = g->base64.as<double>();
        b32 = g->base32.as<float>();
    }
    const SolveResult res = solve_core(g, p, b64, b32, st, start_const ? &cval : nullptr);

    // ---- normalisation (PageRank.cpp:73-92) and scatter to slot order -----------------------
    double factor;
    if (!p->normalized) factor = res.sum;
    else factor = (1.0 / (double)n) * ((1.0 - damp) + damp * res.sink);
    launch_gather_result(res.x, slot_index(g), z, factor, sz, s);
    double maximum = 1.0;
    if (p->normalized) {
        launch_range_max(res.x, n, g->red_partials.as<double>(), g->red_out.as<double>(), s);
        double mx = 0.0, s0 = 0.0;
        read_back(s, {{&mx, g->red_out.p, 8}, {&s0, sz, 8}});
        maximum = std::max(mx / factor, s0);
    }
    st->maximum = maximum;
    read_back(s, {{scores, sz, p->scores_on_device ? 0 : 8 * z}});
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

namespace {
int create_handle(const nkpr_csc *csc, nkpr_graph **out, const uint64_t *group_off,
                  uint32_t n_groups) {
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
            create_graph(csc, g, group_off, n_groups);
        } catch (...) {
            nkpr_graph_destroy(g);
            throw;
        }
        *out = g;
    });
}
}  // namespace

int nkpr_graph_create(const nkpr_csc *csc, nkpr_graph **out) {
    return create_handle(csc, out, nullptr, 0);
}

int nkpr_graph_create_batch(const nkpr_csc *csc, uint32_t num_graphs,
                            const uint64_t *graph_offsets, nkpr_graph **out) {
    if (graph_offsets == nullptr || num_graphs == 0) {
        g_last_error = "a batch needs num_graphs >= 1 and graph_offsets";
        return NKPR_ERR_INVALID;
    }
    return create_handle(csc, out, graph_offsets, num_graphs);
}

int nkpr_run_batch(nkpr_graph *g, const nkpr_params *p, double *scores, uint64_t *iterations,
                   double *maximum, nkpr_stats *stats) {
    nkpr_stats local;
    return guarded([&] {
        NKPR_CHECK(g != nullptr, "graph is NULL");
        DeviceGuard dg(g->device);
        run_batch_impl(g, p, scores, iterations, maximum, stats ? stats : &local);
    });
}

int nkpr_graph_destroy(nkpr_graph *g) {
    if (!g) return NKPR_OK;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
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
        for (uint32_t q = 0; q < g->n_groups; ++q) info->sinks += g->goff[q + 1] - g->gsink[q];
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
    return guarded([&] {
        NKPR_CHECK(g == nullptr || g->n_groups == 0, "a batch handle runs with nkpr_run_batch");
        run_impl(g, p, scores, stats ? stats : &local);
    });
}

int nkpr_power_params_default(nkpr_power_params *p) {
    return guarded([&] {
        NKPR_CHECK(p != nullptr, "params is NULL");
        std::memset(p, 0, sizeof(*p));
        p->kind = NKPR_EIGENVECTOR;
        p->alpha = 0.0;
        p->beta = 0.1;
        p->tol = 1e-8;
        p->max_iterations = std::numeric_limits<uint64_t>::max();
    });
}

int nkpr_power_run(nkpr_graph *g, const nkpr_power_params *p, double *scores, nkpr_stats *stats) {
    nkpr_stats local;
    return guarded([&] {
        NKPR_CHECK(g != nullptr, "graph is NULL");
        DeviceGuard dg(g->device);
        power_run_impl(g, p, scores, stats ? stats : &local);
    });
}

int nkpr_graph_set_weights(nkpr_graph *g, const float *weight, const double *weight64,
                           const double *out_wdeg, int32_t device_ptrs) {
    return guarded([&] {
        NKPR_CHECK(g != nullptr, "graph is NULL");
        DeviceGuard dg(g->device);
        set_weights(g, weight, weight64, out_wdeg, device_ptrs);
    });
}

int nkpr_adjoint(nkpr_graph *g, const nkpr_params *p, const nkpr_csc *csc,
                 const double *scores, double score_sum, const double *grad_scores,
                 double *grad_weight, double *grad_teleport, nkpr_stats *stats) {
    nkpr_stats local;
    return guarded([&] {
        NKPR_CHECK(g != nullptr, "graph is NULL");
        NKPR_CHECK(g->n_groups == 0, "the adjoint runs on single graphs");
        DeviceGuard dg(g->device);
        adjoint_impl(g, p, csc, scores, score_sum, grad_scores, grad_weight, grad_teleport,
                     stats ? stats : &local);
    });
}

int nkpr_ppr_batch(nkpr_graph *g, const nkpr_params *p, uint32_t k,
                   const uint64_t *seed_offsets, const int32_t *seed_nodes,
                   const double *seed_weights, double *scores, nkpr_batch_stats *stats,
                   uint64_t *col_iterations, double *col_error_bound) {
    nkpr_batch_stats local;
    return guarded([&] {
        NKPR_CHECK(g != nullptr, "graph is NULL");
        NKPR_CHECK(g->n_groups == 0, "batched personalised PageRank runs on single graphs");
        DeviceGuard dg(g->device);
        ppr_batch_impl(g, p, k, seed_offsets, seed_nodes, seed_weights, scores,
                       stats ? stats : &local, col_iterations, col_error_bound);
    });
}

int nkpr_csr_spmv(const nkpr_csr *a, const double *x, double *y, void *stream) {
    return guarded([&] { csr_spmv_impl(a, x, y, (hipStream_t)stream); });
}

int nkpr_csr_pagerank_step(const nkpr_csr *m, double shift, const double *x, double *y,
                           double *step_sq, void *stream) {
    return guarded([&] {
        csr_step_impl(m, shift, x, y, step_sq, (hipStream_t)stream);
    });
}

int nkpr_csr_pagerank_step_noalloc(const nkpr_csr *m, double shift, const double *x, double *y,
                                   double *step_sq, void *stream) {
    return guarded([&] { csr_step_noalloc_impl(m, shift, x, y, step_sq, (hipStream_t)stream); });
}

int nkpr_algebraic_pagerank(const nkpr_csr *m, double damp, double tol, uint64_t max_iterations,
                            double *scores, int32_t scores_on_device, uint64_t *iterations) {
    return guarded([&] {
        algebraic_pagerank_impl(m, damp, tol, max_iterations, scores, scores_on_device,
                                iterations);
    });
}

int nkpr_algebraic_create(const nkpr_csr *m, nkpr_algebraic **out) {
    return guarded([&] { algebraic_create_impl(m, out); });
}

int nkpr_algebraic_run(nkpr_algebraic *h, double damp, double tol, uint64_t max_iterations,
                       double *scores, int32_t scores_on_device, uint64_t *iterations,
                       nkpr_stats *stats) {
    return guarded([&] {
        NKPR_CHECK(h != nullptr, "handle is NULL");
        DeviceGuard dg(algebraic_device(h));
        algebraic_run_impl(h, damp, tol, max_iterations, scores, scores_on_device, iterations,
                           stats);
    });
}

int nkpr_algebraic_destroy(nkpr_algebraic *h) {
    return guarded([&] {
        if (!h) return;
        DeviceGuard dg(algebraic_device(h));
        algebraic_destroy_impl(h);
    });
}

int nkpr_nkbg_read(const char *path, int32_t with_out_lists, nkpr_nkbg *out) {
    return guarded([&] {
        try {
            nkbg_read_impl(path, with_out_lists, out);
        } catch (...) {
            nkbg_free_impl(out);
            throw;
        }
    });
}

void nkpr_nkbg_free(nkpr_nkbg *g) { nkbg_free_impl(g); }

int nkpr_nkbg_write(const char *path, const nkpr_nkbg *g, uint64_t chunks, int32_t weight_format) {
    return guarded([&] { nkbg_write_impl(path, g, chunks, weight_format); });
}

int nkpr_apr_create(uint64_t z, const uint64_t *ptr, const int32_t *dst, const double *w,
                    int32_t device_ptrs, nkpr_apr **out) {
    return guarded([&] { apr_create_impl(z, ptr, dst, w, device_ptrs, out); });
}

int nkpr_apr_destroy(nkpr_apr *a) {
    return guarded([&] {
        if (!a) return;
        DeviceGuard dg(apr_device(a));
        apr_destroy_impl(a);
    });
}

int nkpr_apr_run(nkpr_apr *a, double alpha, double eps, const uint64_t *seeds, uint64_t nseeds,
                 uint64_t *count, int64_t *ids_out, double *p_out, double *r_out,
                 nkpr_apr_stats *stats) {
    return guarded([&] {
        NKPR_CHECK(a != nullptr, "handle is NULL");
        DeviceGuard dg(apr_device(a));
        apr_run_impl(a, alpha, eps, seeds, nseeds, count, ids_out, p_out, r_out, stats);
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
        // Centrality.cpp:33-39 orders by score descending, ties by id ascending.  Its comparator
        // is not a strict weak order when a score is NaN, which lets a sort read outside the
        // range; here NaN scores rank after every number (ties by id), so the order is total.
        std::sort(ids.begin(), ids.end(), [&](int64_t a, int64_t b) {
            const double sa = scores[a], sb = scores[b];
            const bool na = std::isnan(sa), nb = std::isnan(sb);
            if (na || nb) return na != nb ? nb : a < b;
            if (sa == sb) return a < b;
            return sa > sb;
        });
        std::copy(ids.begin(), ids.end(), ids_out);
        if (scores_out)
            for (size_t i = 0; i < ids.size(); ++i) scores_out[i] = scores[ids[i]];
        if (count) *count = ids.size();
    });
}

}  // extern "C"
