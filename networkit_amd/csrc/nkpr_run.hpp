// Host orchestration shared by the entry points of libnkpr.so: batches of sweeps with a
// device-side stop flag (run_phase), the FP32 GMRES used for correction solves, and one
// complete solve on compact vectors (solve_core).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#include "nkpr_internal.hpp"

namespace nkpr {

// Buffers of one precision for a ping-pong phase.
template <typename T> struct Vecs {
    T *x[2];
    T *xh[2];
};

inline void ensure(DevBuf &b, size_t bytes) { b.alloc(bytes); }

constexpr int kMaxRefinements = 30;

inline uint32_t env_u32(const char *name, uint32_t dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    return (uint32_t)std::strtoul(v, nullptr, 10);
}

struct RunCtx {
    nkpr_graph *g;
    const nkpr_params *p;
    nkpr_stats *st;
    uint64_t sweeps_left;
    int workers[2][4][2];  // [dbl][kind][folded] workers per slice
    int combine_grid;
};

inline Schedule make_schedule(nkpr_graph *g, uint32_t workers) {
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
                   uint64_t max_launch, bool timing, double *ms_acc,
                   const std::function<void(const T *, T *)> &post = {}) {
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

    // FP64 sweeps of large graphs run through the binned copy (per-graph hooks of a batch
    // handle run after either sweep)
    const bool binned64 = dbl && binned64_active(g);
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
            if constexpr (sizeof(T) == 8) {
                if (binned64) binned_sweep64(g, a, f, kind, g->stream, e0, e1);
                else launch_sweep<T>(a, f, kind, folded, rc.combine_grid, g->stream, e0, e1);
            } else {
                launch_sweep<T>(a, f, kind, folded, rc.combine_grid, g->stream, e0, e1);
            }
            if (post) post(v.x[cur], v.x[nxt]);
        }
        read_back(g->stream, {{g->ctl_host, g->ctl.p, sizeof(PhaseCtl)}});
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

inline double rel_error_bound(double damp, double l1, double sum) {
    return sum > 0.0 ? 2.0 * damp / (1.0 - damp) * l1 / sum : INFINITY;
}



// Result of solve_core: the unnormalised FP64 iterate (compact order) and its sums.
struct SolveResult {
    const double *x;
    double sum;
    double sink;
    uint64_t total;
};

// FP32 Krylov workspace of a graph (basis of m + 1 vectors and four work vectors).
struct Krylov {
    int m;
    uint32_t n;
    uint64_t ld;    // stride of the basis vectors: n rounded up to 64 floats
    float *V;       // basis [(m + 1) * ld]
    float *y;       // d P v
    float *vh, *yh; // prescaled copies (unweighted graphs)
    float *e;       // solution of the last solve
    double *dpart, *dout;
    double *dhost;  // pinned host copy of dout (kDotResults entries)
    hipEvent_t ev;  // recorded after a copy of dout to dhost
};
constexpr uint32_t kDotResults = 128;

void prepare_run(nkpr_graph *g);
bool closed_start_enabled();
SolveResult solve_core(nkpr_graph *g, const nkpr_params *p, const double *base64,
                       const float *base32, nkpr_stats *st, const double *start_const = nullptr);
Krylov krylov_setup(nkpr_graph *g);
void ensure_rowsum64(nkpr_graph *g);
void apply_operator(RunCtx &rc, const float *v, float *w, const Krylov &kr, bool timing,
                    double *ms_acc, bool minus);
// With `timing`, a binned application's time reaches *ms_acc when the next application starts
// or when collect_op_timing runs; every run that times applications calls it at its end (and
// clears g->op_pending at its start).
bool apply_binned(nkpr_graph *g, const float *v, float *w, double damp, bool distribute,
                  bool minus, bool timing, double *ms_acc);
void collect_op_timing(nkpr_graph *g, double *ms_acc);

// GMRES(m) from e = 0 on A e = rhs, A = I - B, until ||A e - rhs||_2 <= rtol ||rhs||_2 or `cap`
// operator applications.  `apply(v, w)` computes w = B v.  The Krylov spaces of A and B are
// the same, so the Arnoldi process runs on B, B V_j = V_{j+1} H_B, and the least-squares
// problem uses A V_j = V_{j+1} (I~ - H_B).  On B the Gram-Schmidt pass cancels little: A v_j
// is mostly v_j itself (B is a contraction), so Arnoldi on A would cancel most of every new
// vector and need the second pass each step.
//
// The basis is kept unnormalised: slot i holds u_i and the orthonormal vector is
// v_i = sig[i] u_i, with the scales on the host (no normalisation pass over the vector).  A
// step is one operator application, one multi-dot <u_i, w>, <w, w> (the only host read), and
// one projection.  The projected norm comes from ||w'||^2 = ||w||^2 - sum_i <v_i, w>^2
// (accumulated in FP64 from the stored FP32 vectors); when that cancels to eta^2 ||w||^2 or
// less, the projection is repeated with measured dots and norms (selective reorthogonalisation,
// the criterion of Daniel, Gragg, Kaufman and Stewart, with eta = 0.01: only a projection
// that cancels all but a hundredth of the norm is redone; the FP64 residual sweeps check
// every correction, and on the R-MAT scale-26, RHG and Erdos-Renyi workloads eta = 0.1
// repeated projections without changing the applications or the result).  A vector whose
// norm leaves [2^-50, 2^50] is rescaled on the device, so the FP32 entries stay far from the
// ends of the range.  `rhs` may be basis slot 0 itself; it is overwritten then.  The Hessenberg
// least-squares problem is solved on the host with Givens rotations.  Returns the number of
// applications.
//
// Head and tail (n_full > k.n): the Krylov vectors span the first k.n entries only, a closed
// system when no row below k.n draws from a source at or past k.n (the sources there have no
// out-edges).  `apply` still writes all n_full rows, and no projection touches the tail, so
// slot i + 1 keeps (B u_i) on the tail; the tail rows of A e = rhs then follow exactly:
// e_t = rhs_t + sum_i c_i (B u_i)_t with e_h = sum_i c_i u_i, which is what is written there.
template <typename Apply>
uint64_t gmres_f32(const Krylov &k, Apply &&apply, const float *rhs, float *e, double rtol,
                   uint64_t cap, hipStream_t s, uint32_t n_full = 0) {
    const uint32_t n = k.n;
    const uint32_t n_tail = n_full > n ? n_full - n : 0;
    const int mres = k.m;
    const uint64_t ld = k.ld;
    float *const V = k.V;
    // the results land in pinned memory (a direct copy, no staging through pageable memory)
    auto fetch = [&](double *host, uint32_t cnt) {
        read_back(s, {{k.dhost, k.dout, 8 * cnt}});
        std::memcpy(host, k.dhost, 8 * cnt);
    };
    auto dots = [&](const float *w, const float *Vb, uint32_t cnt, double *host) {
        launch_multidot(w, Vb, ld, cnt, n, k.dpart, k.dout, s);
        fetch(host, cnt);
    };
    // w -= V c (cnt vectors), then host[i] = <u_i, w> for i < cnt and host[cnt] = <w, w>
    auto project = [&](float *w, uint32_t cnt, const double *c, double *host) {
        launch_axpy_dot(w, V, ld, cnt, n, c, k.dpart, k.dout, s);
        fetch(host, cnt + 1);
    };
    const auto in_range = [](double nrm) { return nrm >= 0x1p-50 && nrm <= 0x1p50; };
    double beta2 = 0.0;
    dots(rhs, rhs, 1, &beta2);
    const double beta = std::sqrt(beta2);
    if (!(beta > 0.0) || cap == 0) {
        launch_axpby(rhs, nullptr, e, n, 0.0f, 0.0f, s);
        if (n_tail) launch_axpby(rhs + n, nullptr, e + n, n_tail, 1.0f, 0.0f, s);
        return 0;
    }
    std::vector<double> sig(mres + 1, 1.0);
    if (rhs == V && in_range(beta)) {
        sig[0] = 1.0 / beta;
    } else {
        launch_axpby(rhs, nullptr, V, n, (float)(1.0 / beta), 0.0f, s);
    }
    std::vector<double> R((size_t)mres * mres, 0.0), cs(mres), sn(mres), gv(mres + 1, 0.0);
    gv[0] = beta;
    const bool debug = std::getenv("NKPR_DEBUG") != nullptr;
    // second pass when |w'|^2 <= eta2 |w|^2 (NKPR_GMRES_REORTH overrides eta2)
    double reorth_eta2 = 1e-4;
    if (const char *ev = std::getenv("NKPR_GMRES_REORTH")) reorth_eta2 = std::atof(ev);
    uint64_t applies = 0;
    int kk = 0;
    std::vector<double> h(mres + 1), d(mres + 2), c(mres + 1);
    for (int j = 0; j < mres && applies < cap; ++j) {
        float *uj = V + (size_t)j * ld;
        float *w = V + (size_t)(j + 1) * ld;
        apply(uj, w);  // w = B u_j, so B v_j = sig[j] w
        ++applies;
        // <u_i, w> for i <= j and <w, w> (w is slot j + 1).  The projection w -= sum_i c_i u_i
        // with c_i = sig_i <u_i, w> sig_i is formed on the device and runs while the host reads
        // the dots, so that the device does not wait for the host between the steps.
        launch_multidot(w, V, ld, j + 2, n, k.dpart, k.dout, s);
        NKPR_HIP(hipMemcpyAsync(k.dhost, k.dout, 8 * (size_t)(j + 2), hipMemcpyDeviceToHost, s));
        NKPR_HIP(hipEventRecord(k.ev, s));
        launch_gs_project(w, V, ld, j + 1, n, k.dout, sig.data(), s);
        NKPR_HIP(hipEventSynchronize(k.ev));
        std::memcpy(d.data(), k.dhost, 8 * (size_t)(j + 2));
        const double ww = d[j + 1];
        double proj2 = 0.0;
        for (int i = 0; i <= j; ++i) {
            h[i] = sig[i] * d[i];  // <v_i, w>
            c[i] = sig[i] * h[i];  // w -= sum_i <v_i, w> v_i = sum_i c_i u_i
            proj2 += h[i] * h[i];
        }
        double hn2 = ww - proj2;
        const bool reorth = !(hn2 > reorth_eta2 * ww);
        if (reorth) {
            // the first pass (done above) measured, then the second pass
            dots(w, V, j + 2, d.data());
            for (int i = 0; i <= j; ++i) {
                const double h2 = sig[i] * d[i];
                h[i] += h2;
                c[i] = sig[i] * h2;
            }
            project(w, j + 1, c.data(), d.data());
            hn2 = d[j + 1];
        }
        double hn = std::sqrt(std::max(hn2, 0.0));
        // u_{j+1} = w (projected), v_{j+1} = w / hn
        sig[j + 1] = hn > 0.0 ? 1.0 / hn : 1.0;
        if (hn > 0.0 && !in_range(hn)) {
            launch_axpby(w, nullptr, w, n, (float)(1.0 / hn), 0.0f, s);
            sig[j + 1] = 1.0;
        }
        // column j of the Hessenberg matrix of A: I~ - H_B, with H_B's column sig[j] (h, hn)
        for (int i = 0; i <= j; ++i) h[i] = -sig[j] * h[i];
        h[j] += 1.0;
        h[j + 1] = -sig[j] * hn;
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
        kk = j + 1;
        if (debug)
            std::fprintf(stderr, "[nkpr]   gmres %d: relative residual %.3e%s\n", j,
                         std::fabs(gv[j + 1]) / beta, reorth ? " (reorthogonalised)" : "");
        if (std::fabs(gv[j + 1]) <= rtol * beta || !(hn > 0.0)) break;
    }
    std::vector<double> y(kk, 0.0);
    for (int i = kk - 1; i >= 0; --i) {
        double t = gv[i];
        for (int q = i + 1; q < kk; ++q) t -= R[(size_t)i * mres + q] * y[q];
        y[i] = R[(size_t)i * mres + i] != 0.0 ? t / R[(size_t)i * mres + i] : 0.0;
    }
    for (int i = 0; i < kk; ++i) y[i] *= sig[i];  // e = sum_i y_i v_i = sum_i (y_i sig_i) u_i
    launch_multiaxpy(e, V, ld, kk, n, y.data(), true, s);
    if (n_tail) {
        // e_t = rhs_t + sum_i c_i (B u_i)_t: rhs is slot 0 (its tail untouched), B u_i slot i + 1
        NKPR_CHECK(rhs == V, "a tail needs the right-hand side in slot 0");
        std::vector<double> ct(kk + 1);
        ct[0] = 1.0;
        for (int i = 0; i < kk; ++i) ct[i + 1] = y[i];
        launch_multiaxpy(e + n, V + n, ld, kk + 1, n_tail, ct.data(), true, s);
    }
    return applies;
}

}  // namespace nkpr
