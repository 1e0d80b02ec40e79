// Internal declarations shared by the translation units of libnkpr.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nkpr.h"

namespace nkpr {

// ---------------------------------------------------------------------------------------
// Sweep schedule.  A sweep block is 512 threads (8 wave64s).  A "stream" work item is a run
// of consecutive rows whose in-edges (at most kCap) are staged as products in LDS; a row
// segment with more than kCap in-edges is a "vector" item reduced by the whole block; one
// with more than kVectorMax in-edges is cut into kChunk-sized "split" items whose partial
// sums the finalize kernel adds in chunk order (deterministic).
//
// Source slicing.  Random gathers of the iterate are the cost of a sweep: an L2 miss moves a
// 128-byte line for a 4- or 8-byte value.  The gathered node ids are therefore cut into
// kSlices slices, interleaved in groups of 32 ids (one 128-byte line of FP32 values), and the
// in-edges of every row with more than `direct_max_deg` in-edges are stored slice by slice.
// Work on slice s is done by blocks running on XCD s (read from the hardware register
// XCC_ID), so each XCD's L2 caches only the lines of its own slice: the eight L2s hold
// distinct data instead of eight copies of the same hot lines.  The per-slice row sums are
// added in slice order by a combine kernel, which also runs the row epilogue.  Rows with few
// in-edges are processed unsliced ("direct") with the epilogue fused.  Which block serves
// which slice affects only speed: any assignment gives the same result.
// ---------------------------------------------------------------------------------------
constexpr int kBlock = 512;
constexpr int kCap = 4096;
constexpr int kRowsMax = 512;
constexpr uint32_t kVectorMax = 16384;
constexpr uint32_t kChunk = 8192;
constexpr uint32_t kSplitFlag = 0x80000000u;
constexpr int kFinalizeThreads = 1024;
constexpr int kSlices = 8;
constexpr int kSliceShift = 5;
constexpr uint32_t kDefaultDirectMaxDeg = 8;
constexpr uint32_t kDefaultSliceReplicate = 32768;  // top-ranked sources cached by every XCD
constexpr int kCombineThreads = 256;
// The most gathered sources (relabelled ids [0, hub count)) are copied into LDS by every sweep
// block: an LDS gather costs a fraction of an L2 hit.  32 KiB of FP32 values, 8 KiB of FP64.
template <typename T> struct HubTable { static constexpr uint32_t size = 8192; };
template <> struct HubTable<double> { static constexpr uint32_t size = 1024; };

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define NKPR_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t nkpr_e_ = (call);                                                         \
        if (nkpr_e_ != hipSuccess)                                                           \
            throw ::nkpr::Error(NKPR_ERR_HIP, std::string(#call) + " failed: " +              \
                                                  hipGetErrorString(nkpr_e_));               \
    } while (0)

#define NKPR_CHECK(cond, msg)                                                                \
    do {                                                                                     \
        if (!(cond)) throw ::nkpr::Error(NKPR_ERR_INVALID, (msg));                           \
    } while (0)

// Device-to-host reads.  The targets are often stack variables or vectors of the calling frame,
// so no copy may still be in flight when an exception unwinds that frame: every copy is
// issued, then the stream is synchronised, and only then is a failure reported.
struct HostCopy {
    void *dst;
    const void *src;
    size_t bytes;
};
inline void read_back(hipStream_t s, const HostCopy *copies, size_t count) {
    hipError_t first = hipSuccess;
    const char *what = "";
    for (size_t i = 0; i < count; ++i) {
        const HostCopy &c = copies[i];
        if (c.bytes == 0) continue;
        const hipError_t e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) {
            first = e;
            what = "hipMemcpyAsync (device to host)";
            break;
        }
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (first == hipSuccess && e != hipSuccess) {
        first = e;
        what = "hipStreamSynchronize";
    }
    if (first != hipSuccess)
        throw Error(NKPR_ERR_HIP, std::string(what) + " failed: " + hipGetErrorString(first));
}
inline void read_back(hipStream_t s, std::initializer_list<HostCopy> copies) {
    read_back(s, copies.begin(), copies.size());
}

// Owning device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        if (b <= bytes && p) return;
        reset();
        if (b == 0) b = 16;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(NKPR_ERR_NOMEM, "hipMalloc(" + std::to_string(b) +
                                            " bytes) failed: " + hipGetErrorString(e));
        }
        bytes = b;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Device-resident control block of one sweep phase; written by the finalize kernel.
struct PhaseCtl {
    double sink_in;   // sum of sink scores of the vector the next sweep reads
    double l1;        // ||x' - x||_1 of the last sweep
    double l2;        // ||x' - x||_2^2 of the last sweep
    double sum;       // sum of x' over existing nodes
    double sink_out;  // sum of x' over sinks
    double prev_l1;   // l1 of the sweep before
    double bound;     // certified L1 error bound (ERR rule)
    uint64_t iterations;
    uint32_t done;
    uint32_t pad;
};

// Stopping rules evaluated by the finalize kernel.
enum RuleKind : int32_t {
    kRuleStep = 0,       // reference: step in `norm` <= tol
    kRuleErr = 1,        // 2 d/(1-d) l1 / sum <= tol
    kRuleAbsStall = 2,   // l1 <= tol_abs, or the step stopped shrinking (rounding floor)
};

struct PhaseRule {
    int32_t kind;
    int32_t norm;
    double tol;
    double tol_abs;
    double stall_ratio;
    uint64_t max_iterations;
    double damp;
};

// Device view of the sweep schedule (built once per graph).
struct Schedule {
    const int32_t *src;          // [m + kCap] sources: direct region, then slice regions
    uint32_t n_slices;           // 1 or kSlices
    // direct rows (in-degree <= direct_max_deg): epilogue fused into the sweep
    const uint32_t *drow;        // [n_d] row ids
    const uint32_t *dptr;        // [n_d + 1] absolute in-edge offsets
    const uint4 *ditems;         // {k0, k1, e0, cnt} over direct index k
    uint32_t n_d, n_ditems;
    // sliced rows: per-slice partial sums, epilogue in the combine kernel
    const uint32_t *srow;        // [n_s] row ids
    const uint32_t *sptr;        // [n_slices * n_s + 1] slice-major in-edge offsets
    const uint4 *sitems;         // per slice: split chunks, vector rows, stream blocks
    uint32_t sitem_off[kSlices + 1];
    uint32_t n_s;
    const uint8_t *split_flag;   // [n_s] 1: the finalize kernel finishes this row
    uint32_t *claim;             // [kSlices] worker claim counters (reset every sweep)
    uint32_t workers;            // workers per slice; the sweep grid is n_slices * workers
};

template <typename T> struct SweepArgs {
    Schedule sc;
    const T *coef;           // [m + kCap] folded w/deg(src) (weighted) or nullptr
    const T *inv;            // [n] 1/deg (unweighted: prescale the new vector) or nullptr
    const T *xg;             // gather vector (x, or x*inv when prescaled)
    const T *xold;           // [n] previous iterate
    T *xnew;                 // [n] new iterate
    T *xhat_new;             // [n] prescaled new iterate (unweighted) or nullptr
    const T *base_vec;       // [n] per-node right-hand side (correction solves) or nullptr
    float *r32;              // [n] FP32 copy of x' - x (FP64 residual sweeps) or nullptr
    T *partial;              // [n_slices * n_s] per-slice row sums of sliced rows
    T *split_partial;        // [split chunks]
    double *partials;        // [(sweep grid + combine grid) * 4]  (l1, l2, sum, sink)
    PhaseCtl *ctl;
    double damp;
    double tele;             // constant right-hand side per node ((1-d)/n, or 0)
    uint32_t n;
    uint32_t sink_begin;     // nodes [sink_begin, n) have no out-edges
    int32_t distribute;      // add damp * sink_in / n to every node
    // FP64 residual sweeps through the binned copy: store each contribution in 6 bytes (an
    // FP32 head and a 16-bit tail, about 40 significant bits) instead of 8
    int32_t compact64;
};

// Relative error of a contribution stored as an FP32 head and a 16-bit tail (at most
// 2^-39 of its magnitude); a sweep that uses them may be off by this much times sum |c x|.
constexpr double kCompact64Rel = 0x1p-39;

struct FinalizeArgs {
    const uint32_t *split_rows;   // [n_split_rows] sliced-row index j
    const uint32_t *split_first;  // [n_split_rows * n_slices] first chunk (count 0: partial)
    const uint32_t *split_count;  // [n_split_rows * n_slices]
    uint32_t n_split_rows;
    uint32_t n_blocks;            // sweep grid + combine grid
    PhaseRule rule;
};

// Kernel launchers (nkpr_sweep.hip).
enum SweepKind : int {
    kSweepPlain = 0,      // x' = d P x + tele (+ sinks)
    kSweepBaseVec = 1,    // x' = d P x + base_vec (+ sinks): correction solve
    kSweepResid = 2,      // plain + r32 = float(x' - x)  (FP64 only)
    kSweepBaseResid = 3,  // base vector + r32  (FP64 only)
};

// sweep (grid = n_slices * workers) + combine (combine_grid blocks, if n_s > 0) + finalize
template <typename T>
void launch_sweep(const SweepArgs<T> &a, const FinalizeArgs &f, int kind, bool folded,
                  int combine_grid, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1);
// blocks of the sweep kernel that fit the device at once
int sweep_resident_blocks(int device, bool dbl, bool folded, int kind);

// Small vector kernels (nkpr_sweep.hip).
void launch_gather_start(const double *scores_z, const int32_t *perm, uint32_t n,
                         double *x64, float *x32, hipStream_t s);
void launch_prescale(const double *x64, const double *inv64, double *xhat64, const float *x32,
                     const float *inv32, float *xhat32, uint32_t n, uint32_t sink_begin,
                     hipStream_t s);
// out64[u] = f * in[perm[u]] and out32 = (float)out64 (a slot-order vector to compact order)
void launch_gather_scaled(const double *in_z, const int32_t *perm, uint32_t n, double f,
                          double *out64, float *out32, hipStream_t s);
void launch_to_double(const float *x32, double *x64, uint32_t n, hipStream_t s);
void launch_add_correction(double *x64, const float *e32, uint32_t n, hipStream_t s);
void launch_set_correction(double *x64, double c, const float *e32, uint32_t n, hipStream_t s);
void launch_copy_f32(const float *in, float *out, uint32_t n, hipStream_t s);
// rowsum[u] = sum of the in-coefficients of u (FP64), from the plain CSC copy
void launch_rowsum64(const uint32_t *rowptr, uint32_t n, const int32_t *src, const double *coef,
                     const double *inv, double *rowsum, hipStream_t s);
// flag (device) = 1 when x[0, n) is not constant
void launch_constant_check(const double *x, uint32_t n, uint32_t *flag, hipStream_t s);
// residual sweep from a constant vector c in closed form (ynew, r32 and the block partials
// (l1, l2, sum, sink) of a sweep); returns the number of partial blocks
unsigned launch_const_resid(const double *rowsum, uint32_t n, double dc, double rhs,
                            const double *base, double c, uint32_t sink_begin, double *ynew,
                            float *r32, double *partials, hipStream_t s);
// slot_u[slot] = compact id of the node there, -1 for slots without one (z entries)
void launch_slot_index(const int32_t *perm, uint32_t n, uint64_t z, int32_t *slot_u, hipStream_t s);
// scores_z[slot] = y[slot_u[slot]] / factor for slots with a node (the others keep their value)
void launch_gather_result(const double *y, const int32_t *slot_u, uint64_t z, double factor,
                          double *scores_z, hipStream_t s);
void launch_scatter_result(const double *y, const int32_t *perm, uint32_t n, double factor,
                           double *scores_z, hipStream_t s);
// Deterministic two-stage sums / max over x[begin, end).
template <typename T>
void launch_range_sum(const T *x, uint32_t begin, uint32_t end, double *partials,
                      double *out, hipStream_t s);
void launch_range_max(const double *x, uint32_t n, double *partials, double *out,
                      hipStream_t s);
void launch_init_ctl(PhaseCtl *ctl, const double *sink_sum, uint32_t *claim, hipStream_t s);

// FP32 Krylov vector kernels (nkpr_krylov.hip).  V holds vectors of length n at stride ld (a
// multiple of 4 floats; every vector 16-byte aligned).
constexpr int kDotGroup = 24;   // vectors per launch
constexpr int kDotStride = 32;  // doubles per block in the partials buffer (>= kDotGroup + 1)
constexpr int kDotPartials = kDotStride * 1024;
struct Coefs {
    float c[kDotGroup];
};
constexpr int kMaxKrylov = 64;  // largest GMRES restart length
struct Scales {
    double s[kMaxKrylov];
};
// out[i] = <w, V[i]> (FP64 accumulation) for i < count; partials: kDotPartials doubles
void launch_multidot(const float *w, const float *V, uint64_t ld, uint32_t count, uint32_t n,
                     double *partials, double *out, hipStream_t s);
// assign: w = sum_i coef[i] V[i];  otherwise w -= sum_i coef[i] V[i]
void launch_multiaxpy(float *w, const float *V, uint64_t ld, uint32_t count, uint32_t n,
                      const double *coef, bool assign, hipStream_t s);
// w -= sum_i c_i V[i] with c_i = sig[i] (sig[i] d[i]) formed on the device from the dots d
// (device memory) and the host's scales sig
void launch_gs_project(float *w, const float *V, uint64_t ld, uint32_t count, uint32_t n,
                       const double *d, const double *sig, hipStream_t s);
// w -= sum_i coef[i] V[i], then out[i] = <V[i], w> for i < count and out[count] = <w, w>
void launch_axpy_dot(float *w, const float *V, uint64_t ld, uint32_t count, uint32_t n,
                     const double *coef, double *partials, double *out, hipStream_t s);
// out = a x + b y  (y may be nullptr)
void launch_axpby(const float *x, const float *y, float *out, uint32_t n, float a, float b,
                  hipStream_t s);

// Batched personalised PageRank (nkpr_batch.hip).  Per-column control block.
struct BatchCol {
    double sink_in;     // sink mass of the vector the next sweep reads
    double l1, l2, sum, sink_out, prev_l1;
    double bound;       // certified L1 bound 2 d/(1-d) l1 / sum
    double target;      // step a correction solve must reach so the next residual certifies
    double solve_tol;   // target of the running correction solve
    uint64_t iterations;  // sweeps in which the column was active (all phases)
    uint64_t phase_iter;  // sweeps of the current phase
    uint32_t done;      // frozen for the rest of the phase
    uint32_t final_;    // certified (or padding): frozen for the rest of the run
};
struct BatchCtl {
    uint64_t sweeps;    // executed sweeps of the phase
    uint32_t active;
    uint32_t done;
};
struct BatchTune {
    uint32_t key;       // ld * 2 + (FP64)
    int lpr;
    uint32_t heavy;
};
void batch_layout(nkpr_graph *g);
void batch_split(nkpr_graph *g, uint32_t heavy);
void ensure_buf(DevBuf &b, size_t bytes);
// Binned FP32 operator of the GMRES corrections (nkpr_binned.hip).
bool binned_enabled(const nkpr_graph *g);
bool binned_layout(nkpr_graph *g);
void binned_apply(nkpr_graph *g, const float *v, float *w, double damp, bool distribute, bool minus,
                  hipEvent_t ev0, hipEvent_t ev1);
// Binned FP64 sweep: true when FP64 sweeps of this graph run binned (builds the copy on first
// use); binned_sweep64 is one sweep with the epilogue, reductions and stop rule of
// launch_sweep<double>.
bool binned64_active(nkpr_graph *g);
// slot -> compact id of the graph (z entries, -1 for slots without a node), built on first use
const int32_t *slot_index(nkpr_graph *g);
void binned_sweep64(nkpr_graph *g, const SweepArgs<double> &a, const FinalizeArgs &f, int kind,
                    hipStream_t s, hipEvent_t ev0, hipEvent_t ev1);
// The finalize step of launch_sweep<double> alone (nkpr_sweep.hip).
void launch_finalize64(const SweepArgs<double> &a, const FinalizeArgs &f, int kind, hipStream_t s);
void set_weights(nkpr_graph *g, const float *w, const double *w64, const double *out_wdeg,
                 int device_ptrs);
void adjoint_impl(nkpr_graph *g, const nkpr_params *p, const nkpr_csc *csc,
                  const double *scores, double score_sum, const double *grad_scores,
                  double *grad_weight, double *grad_teleport, nkpr_stats *st);
void run_batch_impl(nkpr_graph *g, const nkpr_params *p, double *scores, uint64_t *iterations,
                    double *maximum, nkpr_stats *st);
// nkpr_power_run (nkpr_power.hip): eigenvector and Katz centrality
void power_run_impl(nkpr_graph *g, const nkpr_power_params *p, double *scores, nkpr_stats *st);

// CSR products and AlgebraicPageRank (nkpr_csr.hip).  The plain product and the step take
// their scratch from the stream's memory pool, the allocation-free step from a static array.
void csr_spmv_impl(const nkpr_csr *a, const double *x, double *y, hipStream_t s);
// throws NKPR_ERR_INVALID unless row_idx rises from 0 to nnz and every column index is < cols
void csr_validate(const nkpr_csr *m, hipStream_t s);
void csr_step_impl(const nkpr_csr *m, double shift, const double *x, double *y, double *step_sq,
                   hipStream_t s);
void csr_step_noalloc_impl(const nkpr_csr *m, double shift, const double *x, double *y,
                           double *step_sq, hipStream_t s);
void algebraic_pagerank_impl(const nkpr_csr *m, double damp, double tol, uint64_t max_iterations,
                             double *scores, int32_t scores_on_device, uint64_t *iterations);

// AlgebraicPageRank through a PageRank handle built from the matrix (nkpr_algebraic.hip)
void algebraic_create_impl(const nkpr_csr *m, nkpr_algebraic **out);
void algebraic_run_impl(nkpr_algebraic *h, double damp, double tol, uint64_t max_iterations,
                        double *scores, int32_t scores_on_device, uint64_t *iterations,
                        nkpr_stats *stats);
void algebraic_destroy_impl(nkpr_algebraic *h);
int algebraic_device(const nkpr_algebraic *h);

// Approximate personalised PageRank (nkpr_apr.hip)
void apr_create_impl(uint64_t z, const uint64_t *ptr, const int32_t *dst, const double *w,
                     int32_t device_ptrs, nkpr_apr **out);
void apr_run_impl(nkpr_apr *a, double alpha, double eps, const uint64_t *seeds, uint64_t nseeds,
                  uint64_t *count, int64_t *ids_out, double *p_out, double *r_out,
                  nkpr_apr_stats *st);
void apr_destroy_impl(nkpr_apr *a);
int apr_device(const nkpr_apr *a);

// NetworKit binary graph files (nkpr_nkbg.hip)
void nkbg_read_impl(const char *path, int32_t with_out_lists, nkpr_nkbg *out);
void nkbg_free_impl(nkpr_nkbg *g);
void nkbg_write_impl(const char *path, const nkpr_nkbg *g, uint64_t chunks, int32_t weight_format);
void ppr_batch_impl(nkpr_graph *g, const nkpr_params *p, uint32_t k, const uint64_t *seed_off,
                    const int32_t *seeds, const double *seed_w, double *scores,
                    nkpr_batch_stats *st, uint64_t *col_iterations, double *col_bound);

}  // namespace nkpr

// The graph handle.
namespace nkpr {
// One binned copy of the edges (nkpr_binned.hip): segments of (source chunk, destination bin),
// padded to quads, in chunk-major order for pass 1 and bin-major order for pass 2.
struct BinLayout {
    bool ready = false, coef_ready = false, off = false;
    int span_shift = 15;        // ids per source chunk = 2^span_shift
    bool fp64 = false;          // FP64 coefficients and contributions
    bool w32 = false;           // FP64 copy: 4-byte weights (x / deg staged in the table)
    uint64_t mp = 0;            // padded edges
    DevBuf order, pos, src, dst, coef, val, p1, p2, mb, part, ctl, partials;
    // pass-1 places derived per group of 64 quads: segment-start mask, first segment,
    // chunk-major to bin-major offset of every non-empty segment; checked against pos
    DevBuf gmask, gfirst, sdelta;
    uint32_t nsdelta = 0;
    bool derived = false;
    double insum = 0.0;         // largest in-coefficient sum of a row (fixed-point bound)
    uint32_t np1 = 0, np2 = 0, multi = 0, nmb = 0, nslots = 0, grid1 = 0, grid2 = 0;
    // hot sources (FP32 copy): ids [0, hot) bypass pass 1; their entries per bin, padded to
    // quads (hidx packed offsets, hcoef, horder the edge of each position), the hot range of
    // every pass-2 item (p2h)
    uint32_t hot = 0;
    uint64_t mh = 0;
    DevBuf hidx, hcoef, horder, p2h;
};
}  // namespace nkpr

struct nkpr_graph {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t z = 0, m = 0;
    uint32_t n = 0, sink_begin = 0;
    uint64_t max_in_degree = 0;
    bool weighted = false;
    double prep_ms = 0.0;

    nkpr::DevBuf perm;      // int32 [n]   compact id -> slot
    nkpr::DevBuf colptr;    // uint64 [n+1] in-lists in (row, source) order (before slicing)
    nkpr::DevBuf src;       // int32 [m + kCap]   sweep order (direct region, slice regions)
    nkpr::DevBuf coef32;    // float [m + kCap]   (weighted)
    nkpr::DevBuf coef64;    // double [m + kCap]  (weighted)
    nkpr::DevBuf inv32;     // float [n]   (unweighted)
    nkpr::DevBuf inv64;     // double [n]  (unweighted)
    nkpr::DevBuf invw64;    // double [n]  1 / weighted out-degree (weighted)
    nkpr::DevBuf rowsum64;  // double [n]  FP64 sum of each row's in-coefficients (built on use)
    bool rowsum_ready = false;
    nkpr::DevBuf flag_buf;  // uint32 [1]  device-side yes/no answers read back by the host
    bool w32_exact = false; // every weight is an FP32 number (weighted)
    bool unit_degrees = false;  // created with out_wdeg = 1 for every node: coefficients = weights
    nkpr::DevBuf emap;      // uint32 [m]  input position of every edge in sweep order (weighted)
    std::vector<uint8_t> exists_host;  // empty: every slot exists

    // schedule
    uint32_t n_slices = 1, direct_max_deg = 0;
    nkpr::DevBuf drow, dptr, ditems;
    uint32_t n_d = 0, n_ditems = 0;
    nkpr::DevBuf srow, sptr, sitems, split_flag, claim;
    uint32_t sitem_off[nkpr::kSlices + 1] = {0};
    uint32_t n_s = 0;
    uint32_t n_items = 0, n_split_items = 0;
    nkpr::DevBuf split_rows, split_first, split_count;
    uint32_t n_split_rows = 0;

    // run workspace (lazily sized)
    nkpr::DevBuf zbuf;                  // double [z] staging of host scores
    nkpr::DevBuf x64a, x64b, xh64a, xh64b;
    nkpr::DevBuf x32a, x32b, xh32a, xh32b;
    nkpr::DevBuf r32, base64, base32, tbuf;
    nkpr::DevBuf partials;              // double [grid*4]
    nkpr::DevBuf red_partials;          // double [1024]
    nkpr::DevBuf red_out;               // double [4]
    nkpr::DevBuf split_part32, split_part64;
    nkpr::DevBuf part32, part64;        // per-slice row sums of sliced rows
    nkpr::DevBuf kbasis;                // FP32 Krylov basis [(m+1) * n]
    nkpr::DevBuf kwork;                 // FP32 scratch [2 * n]
    nkpr::DevBuf kdot;                  // FP64 dot partials and results
    double *kdot_host = nullptr;        // pinned landing buffer of the dot results
    hipEvent_t kdot_ev = nullptr;       // marks the landing of the dot results
    nkpr::DevBuf ctl;                   // PhaseCtl
    nkpr::PhaseCtl *ctl_host = nullptr; // pinned mirror
    std::vector<hipEvent_t> events;
    nkpr::DevBuf slot_u;                // int32 [z]: compact id of a slot, -1 without a node
    bool slot_ready = false;            // (built on first use, see slot_index)
    // timing of binned operator applications: the last application's events, whose elapsed
    // time is added by the next application or by collect_op_timing (no host wait per
    // application)
    hipEvent_t op_ev[2] = {nullptr, nullptr};
    bool op_pending = false;

    // batched personalised PageRank: plain CSC copy and its row split (built on first use)
    bool b_ready = false, b_split_ready = false, b_coef_ready = false;
    nkpr::DevBuf b_rowptr, b_src, b_coef32, b_coef64;
    std::vector<uint32_t> b_rowptr_host;
    std::vector<int32_t> b_perm_host;
    std::vector<int32_t> b_iperm_host;  // [z]: compact id of a slot, -1 without a node
    uint32_t b_heavy = 0, b_n_light = 0, b_n_heavy = 0, b_n_pieces = 0;
    nkpr::DevBuf b_light, b_hrow, b_hpiece, b_pieces;
    std::vector<nkpr::BatchTune> b_tuned;
    nkpr::DevBuf b_tptr, b_tcol, b_tval32, b_tval64;
    nkpr::DevBuf b_piece32, b_piece64, b_partials, b_col, b_ctl, b_sink_part, b_factor;
    nkpr::DevBuf b_x64a, b_x64b, b_x32a, b_x32b, b_r32, b_zbuf;
    nkpr::BatchCtl *b_ctl_host = nullptr;

    // batch of graphs (nkpr_graph_create_batch): graph q owns compact ids [goff[q], goff[q+1]),
    // its sinks are [gsink[q], goff[q+1]); slots [gslot[q], gslot[q+1])
    uint32_t n_groups = 0;
    std::vector<uint32_t> goff, gsink;
    std::vector<uint64_t> gslot;
    nkpr::DevBuf q_chunks, q_part, q_ctl, q_add, q_factor, q_flag;
    nkpr::DevBuf q_gslot;           // uint64 [n_groups + 1]: gslot on the device
    uint32_t q_nchunks = 0;

    // binned copies of the edges (nkpr_binned.hip, built on first use): the FP32 operator of
    // the GMRES corrections and the FP64 sweep
    nkpr::BinLayout bn32, bn64;

    // adjoint: the transposed operator d P^T as a graph of its own (built on first use; it
    // shares this graph's stream and device)
    std::unique_ptr<nkpr_graph> adj;
    bool adj_stale = false;          // forward coefficients changed since adj was built
    nkpr::DevBuf adj_order;          // uint32 [m] plain-copy edge of every transposed entry
    nkpr::DevBuf adj_ones;           // double [n] unit out-degrees of the transposed graph
    nkpr::DevBuf a_vec[6];

    nkpr_graph() = default;
    nkpr_graph(const nkpr_graph &) = delete;
    nkpr_graph &operator=(const nkpr_graph &) = delete;
    ~nkpr_graph() {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (hipEvent_t e : op_ev)
            if (e) (void)hipEventDestroy(e);
        if (ctl_host) (void)hipHostFree(ctl_host);
        if (kdot_host) (void)hipHostFree(kdot_host);
        if (kdot_ev) (void)hipEventDestroy(kdot_ev);
        if (b_ctl_host) (void)hipHostFree(b_ctl_host);
    }
};
