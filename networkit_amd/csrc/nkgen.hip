// Seeded synthetic-graph generators for the benchmark and test workloads (libnkgen.so).
//
// Not part of the PageRank boundary: these produce the input graphs of the workloads
// (directed R-MAT, Erdos-Renyi, Barabasi-Albert, threshold random hyperbolic graphs) with a
// counter-based random stream, so a
// graph depends only on its parameters and seed -- not on the thread count, the device, or
// whether it was made on the host or on the GPU.  NetworKit's own generators cannot serve
// here: its RmatGenerator always builds undirected graphs and scans sequentially
// (networkit/cpp/generators/RmatGenerator.cpp:33,112-117) and its Erdos-Renyi output depends
// on the OpenMP thread count (include/networkit/generators/ErdosRenyiEnumerator.hpp:87-136).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

extern "C" {
int nkgen_rmat(uint64_t seed, int scale, uint64_t first, uint64_t count, double a, double b,
               double c, int32_t *src, int32_t *dst, int on_device);
int nkgen_edge_weights(uint64_t seed, const int32_t *src, const int32_t *dst, uint64_t m,
                       float *w, int on_device);
int64_t nkgen_erdos_renyi(uint64_t seed, uint64_t n, double p, int directed, int64_t *u,
                          int64_t *v, int64_t cap);
int64_t nkgen_barabasi_albert(uint64_t seed, uint64_t n, uint64_t k, int64_t *u, int64_t *v,
                              int64_t cap);
int nkgen_rhg_points(uint64_t seed, uint64_t n, double avg_degree, double gamma, double *r,
                     double *theta, double *radius);
int64_t nkgen_rhg(uint64_t seed, uint64_t n, double avg_degree, double gamma, uint64_t *colptr,
                  int32_t *src, int64_t cap);
}

namespace {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// independent 64-bit value for (seed, stream, counter)
__host__ __device__ inline uint64_t rng_at(uint64_t seed, uint64_t stream, uint64_t counter) {
    return mix64(mix64(seed * 0x9e3779b97f4a7c15ull + stream) ^ (counter * 0xd1b54a32d192ed03ull));
}

__host__ __device__ inline double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

// one R-MAT draw: `scale` quadrant choices, most significant bit first
__host__ __device__ inline void rmat_draw(uint64_t seed, int scale, uint64_t i, double a,
                                          double ab, double abc, int32_t &s, int32_t &d) {
    uint32_t su = 0, du = 0;
    for (int lv = 0; lv < scale; ++lv) {
        const double r = u01(rng_at(seed, i, (uint64_t)lv));
        uint32_t bs, bd;
        if (r < a) { bs = 0; bd = 0; }
        else if (r < ab) { bs = 0; bd = 1; }
        else if (r < abc) { bs = 1; bd = 0; }
        else { bs = 1; bd = 1; }
        su = (su << 1) | bs;
        du = (du << 1) | bd;
    }
    s = (int32_t)su;
    d = (int32_t)du;
}

// U[0.5, 1.5) with 24 random bits: exactly representable in FP32
__host__ __device__ inline float edge_weight(uint64_t seed, int32_t s, int32_t d) {
    const uint64_t h = rng_at(seed ^ 0x5851f42d4c957f2dull, (uint64_t)(uint32_t)s,
                              (uint64_t)(uint32_t)d);
    return 0.5f + (float)(h >> 40) * 0x1.0p-24f;
}

__global__ void rmat_kernel(uint64_t seed, int scale, uint64_t first, uint64_t count, double a,
                            double ab, double abc, int32_t *src, int32_t *dst) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (uint64_t)gridDim.x * blockDim.x) {
        int32_t s, d;
        rmat_draw(seed, scale, first + k, a, ab, abc, s, d);
        src[k] = s;
        dst[k] = d;
    }
}

__global__ void weight_kernel(uint64_t seed, const int32_t *src, const int32_t *dst, uint64_t m,
                              float *w) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (uint64_t)gridDim.x * blockDim.x)
        w[k] = edge_weight(seed, src[k], dst[k]);
}

// ---- threshold random hyperbolic graph (temperature 0) -------------------------------------
// n points in a hyperbolic disk of radius R: angle uniform, radius with density
// alpha sinh(alpha r) / (cosh(alpha R) - 1), alpha = (gamma - 1) / 2; u ~ v iff their
// hyperbolic distance is <= R.  R follows the asymptotic average degree
// k = (2/pi) xi^2 n exp(-R/2), xi = alpha / (alpha - 1/2) (Krioukov et al., Phys. Rev. E 82,
// 036106, 2010).  Candidates come from radial bands sorted by angle: for a point and a band,
// the largest angular distance any point of the band can have while adjacent bounds a window
// found by binary search; every candidate is then tested exactly.
constexpr uint32_t kRhgBands = 64;
constexpr uint64_t kRhgStream = 0x52484731ull;

__host__ __device__ inline double rhg_radius(double alpha, uint64_t n, double k) {
    const double xi = alpha / (alpha - 0.5);
    return 2.0 * log(2.0 * xi * xi * (double)n / (M_PI * k));
}

__host__ __device__ inline void rhg_point(uint64_t seed, uint64_t i, double R, double alpha,
                                          double &r, double &theta) {
    theta = 2.0 * M_PI * u01(rng_at(seed, kRhgStream, 2 * i));
    const double u = u01(rng_at(seed, kRhgStream, 2 * i + 1));
    r = acosh(1.0 + u * (cosh(alpha * R) - 1.0)) / alpha;
}

// cosh d(a, b) <= cosh R in the cancellation-free form cosh(ra - rb) + 2 sinh ra sinh rb
// sin^2((ta - tb)/2); evaluated in the order of the point ids so that it is symmetric bit for bit
__host__ __device__ inline bool rhg_adjacent(uint32_t u, double ru, double tu, uint32_t v,
                                             double rv, double tv, double cosh_R) {
    const double ra = u < v ? ru : rv, rb = u < v ? rv : ru;
    const double ta = u < v ? tu : tv, tb = u < v ? tv : tu;
    const double s = sin(0.5 * (ta - tb));
    return cosh(ra - rb) + 2.0 * sinh(ra) * sinh(rb) * s * s <= cosh_R;
}

__global__ void rhg_points_kernel(uint64_t seed, uint64_t n, double R, double alpha, double *r,
                                  double *theta, uint64_t *key, uint32_t *val, uint32_t *bcount) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ri, ti;
    rhg_point(seed, i, R, alpha, ri, ti);
    r[i] = ri;
    theta[i] = ti;
    const uint32_t b = min(kRhgBands - 1, (uint32_t)(ri / R * kRhgBands));
    key[i] = ((uint64_t)b << 40) | (uint64_t)(ti / (2.0 * M_PI) * 1099511627776.0);
    val[i] = (uint32_t)i;
    atomicAdd(&bcount[b], 1u);
}

__global__ void rhg_sorted_kernel(const uint32_t *order, uint64_t n, const double *r,
                                  const double *theta, double *sr, double *st, uint32_t *sid) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = order[j];
    sr[j] = r[i];
    st[j] = theta[i];
    sid[j] = i;
}

__device__ inline uint32_t lower_bound_d(const double *a, uint32_t lo, uint32_t hi, double x) {
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// count (out == nullptr) or write the neighbours of every point
__global__ void rhg_adj_kernel(uint64_t n, double R, const double *r, const double *theta,
                               const uint32_t *bstart, const double *sr, const double *st,
                               const uint32_t *sid, uint32_t *deg, const uint64_t *colptr,
                               int32_t *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t u = (uint32_t)i;
    const double ru = r[u], tu = theta[u];
    const double cosh_R = cosh(R);
    const double sh_u = sinh(ru);
    uint64_t w = out ? colptr[u] : 0;
    uint32_t cnt = 0;
    for (uint32_t b = 0; b < kRhgBands; ++b) {
        const uint32_t s0 = bstart[b], s1 = bstart[b + 1];
        if (s0 == s1) continue;
        const double lo = R * b / kRhgBands, hi = R * (b + 1) / kRhgBands;
        const double rc = fmin(fmax(ru, lo), hi);
        const double num = cosh_R - cosh(ru - rc);
        const double den = 2.0 * sh_u * sinh(lo);
        double dth = 4.0;  // > pi: the whole band
        if (den > 0.0 && num < den) {
            const double q = fmin(1.0, num / den * (1.0 + 1e-9));
            dth = 2.0 * asin(sqrt(q)) * (1.0 + 1e-12) + 1e-15;
        }
        uint32_t ranges[2][2];
        int nr = 0;
        if (dth >= M_PI) {
            ranges[nr][0] = s0; ranges[nr][1] = s1; ++nr;
        } else {
            const double a = tu - dth, c = tu + dth;
            if (a < 0.0) {
                ranges[nr][0] = lower_bound_d(st, s0, s1, a + 2.0 * M_PI); ranges[nr][1] = s1; ++nr;
                ranges[nr][0] = s0; ranges[nr][1] = lower_bound_d(st, s0, s1, nextafter(c, 10.0)); ++nr;
            } else if (c >= 2.0 * M_PI) {
                ranges[nr][0] = lower_bound_d(st, s0, s1, a); ranges[nr][1] = s1; ++nr;
                ranges[nr][0] = s0; ranges[nr][1] = lower_bound_d(st, s0, s1, nextafter(c - 2.0 * M_PI, 10.0)); ++nr;
            } else {
                ranges[nr][0] = lower_bound_d(st, s0, s1, a);
                ranges[nr][1] = lower_bound_d(st, s0, s1, nextafter(c, 10.0)); ++nr;
            }
        }
        for (int q = 0; q < nr; ++q) {
            for (uint32_t j = ranges[q][0]; j < ranges[q][1]; ++j) {
                const uint32_t v = sid[j];
                if (v == u) continue;
                if (rhg_adjacent(u, ru, tu, v, sr[j], st[j], cosh_R)) {
                    if (out) out[w++] = (int32_t)v;
                    ++cnt;
                }
            }
        }
    }
    if (deg) deg[u] = cnt;
}

}  // namespace

extern "C" {

int nkgen_rmat(uint64_t seed, int scale, uint64_t first, uint64_t count, double a, double b,
               double c, int32_t *src, int32_t *dst, int on_device) {
    if (scale < 1 || scale > 30) return 1;
    const double ab = a + b, abc = a + b + c;
    if (on_device) {
        const uint64_t blocks = std::min<uint64_t>((count + 255) / 256, 1u << 20);
        if (count == 0) return 0;
        hipLaunchKernelGGL(rmat_kernel, dim3((unsigned)blocks), dim3(256), 0, 0, seed, scale,
                           first, count, a, ab, abc, src, dst);
        if (hipGetLastError() != hipSuccess) return 2;
        return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
    }
    for (uint64_t k = 0; k < count; ++k) rmat_draw(seed, scale, first + k, a, ab, abc, src[k], dst[k]);
    return 0;
}

int nkgen_edge_weights(uint64_t seed, const int32_t *src, const int32_t *dst, uint64_t m,
                       float *w, int on_device) {
    if (on_device) {
        if (m == 0) return 0;
        const uint64_t blocks = std::min<uint64_t>((m + 255) / 256, 1u << 20);
        hipLaunchKernelGGL(weight_kernel, dim3((unsigned)blocks), dim3(256), 0, 0, seed, src, dst,
                           m, w);
        if (hipGetLastError() != hipSuccess) return 2;
        return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
    }
    for (uint64_t k = 0; k < m; ++k) w[k] = edge_weight(seed, src[k], dst[k]);
    return 0;
}

// G(n, p) by geometric skipping over the candidate pairs of each row: (u, v), v > u for an
// undirected graph, v != u for a directed one.  Each row has its own random stream.  Returns
// the number of edges, or -1 if `cap` is too small (call with cap = 0 to count).
int64_t nkgen_erdos_renyi(uint64_t seed, uint64_t n, double p, int directed, int64_t *u,
                          int64_t *v, int64_t cap) {
    if (p <= 0.0 || n < 2) return 0;
    const double lq = std::log1p(-std::min(p, 1.0 - 1e-16));
    int64_t k = 0;
    for (uint64_t row = 0; row < n; ++row) {
        uint64_t ctr = 0;
        const uint64_t first = directed ? 0 : row + 1;
        int64_t cand = (int64_t)first - 1;
        for (;;) {
            const double r = u01(rng_at(seed, row, ctr++));
            const double skip = std::floor(std::log1p(-r) / lq);
            cand += 1 + (int64_t)skip;
            if (cand >= (int64_t)n || skip > 1e18) break;
            if (directed && (uint64_t)cand == row) continue;
            if (u && k < cap) {
                u[k] = (int64_t)row;
                v[k] = cand;
            } else if (u) {
                return -1;
            }
            ++k;
        }
    }
    return k;
}

// Barabasi-Albert preferential attachment (Batagelj-Brandes edge-list sampling): start from a
// clique-free seed of k isolated nodes; node i >= k attaches k edges to targets drawn from the
// endpoint list (degree-proportional), without repeating a target for the same node.
int64_t nkgen_barabasi_albert(uint64_t seed, uint64_t n, uint64_t k, int64_t *u, int64_t *v,
                              int64_t cap) {
    if (n <= k || k == 0) return 0;
    const int64_t m = (int64_t)((n - k) * k);
    if (!u) return m;
    if (cap < m) return -1;
    std::vector<int64_t> ends;
    ends.reserve(2 * (size_t)m);
    int64_t e = 0;
    std::vector<int64_t> chosen;
    for (uint64_t i = k; i < n; ++i) {
        chosen.clear();
        uint64_t ctr = 0;
        while (chosen.size() < k) {
            int64_t t;
            if (ends.empty()) {
                t = (int64_t)chosen.size();  // first new node links to the k seed nodes
            } else {
                const uint64_t h = rng_at(seed, i, ctr++);
                t = ends[(size_t)(h % ends.size())];
            }
            bool dup = false;
            for (int64_t c : chosen) dup = dup || (c == t);
            if (!dup) chosen.push_back(t);
        }
        for (int64_t t : chosen) {
            u[e] = (int64_t)i;
            v[e] = t;
            ++e;
            ends.push_back((int64_t)i);
            ends.push_back(t);
        }
    }
    return e;
}

int nkgen_rhg_points(uint64_t seed, uint64_t n, double avg_degree, double gamma, double *r,
                     double *theta, double *radius) {
    if (!(gamma > 2.0) || !(avg_degree > 0.0)) return 1;
    const double alpha = 0.5 * (gamma - 1.0);
    const double R = rhg_radius(alpha, n, avg_degree);
    if (radius) *radius = R;
    for (uint64_t i = 0; i < n && r && theta; ++i) rhg_point(seed, i, R, alpha, r[i], theta[i]);
    return 0;
}

// Threshold RHG on the device as in-lists (each edge in both endpoints' rows, no self-loops).
// colptr[n+1] (uint64) and src[cap] (int32) are device pointers; with colptr == NULL only the
// number of in-list entries is returned.  Returns -1 on bad parameters, -2 on a HIP error, -3
// if cap is too small.
int64_t nkgen_rhg(uint64_t seed, uint64_t n, double avg_degree, double gamma, uint64_t *colptr,
                  int32_t *src, int64_t cap) {
    if (!(gamma > 2.0) || !(avg_degree > 0.0) || n == 0 || n >= (1ull << 31)) return -1;
    const double alpha = 0.5 * (gamma - 1.0);
    const double R = rhg_radius(alpha, n, avg_degree);
    double *r = nullptr, *th = nullptr, *sr = nullptr, *st = nullptr;
    uint64_t *k1 = nullptr, *k2 = nullptr;
    uint32_t *v1 = nullptr, *v2 = nullptr, *sid = nullptr, *bcount = nullptr, *bstart = nullptr,
             *deg = nullptr;
    void *tmp = nullptr;
    int64_t result = -2;
    auto ok = [](hipError_t e) { return e == hipSuccess; };
    do {
        if (!ok(hipMalloc(&r, 8 * n)) || !ok(hipMalloc(&th, 8 * n)) || !ok(hipMalloc(&sr, 8 * n)) ||
            !ok(hipMalloc(&st, 8 * n)) || !ok(hipMalloc(&k1, 8 * n)) || !ok(hipMalloc(&k2, 8 * n)) ||
            !ok(hipMalloc(&v1, 4 * n)) || !ok(hipMalloc(&v2, 4 * n)) || !ok(hipMalloc(&sid, 4 * n)) ||
            !ok(hipMalloc(&bcount, 4 * (kRhgBands + 1))) || !ok(hipMalloc(&bstart, 4 * (kRhgBands + 1))) ||
            !ok(hipMalloc(&deg, 4 * (n + 1))))
            break;
        if (!ok(hipMemset(bcount, 0, 4 * (kRhgBands + 1)))) break;
        const unsigned blocks = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(rhg_points_kernel, dim3(blocks), dim3(256), 0, 0, seed, n, R, alpha, r, th,
                           k1, v1, bcount);
        hipcub::DoubleBuffer<uint64_t> kb(k1, k2);
        hipcub::DoubleBuffer<uint32_t> vb(v1, v2);
        size_t tb = 0;
        if (!ok(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb, vb, (size_t)n, 0, 46))) break;
        if (!ok(hipMalloc(&tmp, tb))) break;
        if (!ok(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kb, vb, (size_t)n, 0, 46))) break;
        hipLaunchKernelGGL(rhg_sorted_kernel, dim3(blocks), dim3(256), 0, 0, vb.Current(), n, r, th,
                           sr, st, sid);
        std::vector<uint32_t> hc(kRhgBands + 1, 0), hs(kRhgBands + 1, 0);
        if (!ok(hipMemcpy(hc.data(), bcount, 4 * kRhgBands, hipMemcpyDeviceToHost))) break;
        for (uint32_t b = 0; b < kRhgBands; ++b) hs[b + 1] = hs[b] + hc[b];
        if (!ok(hipMemcpy(bstart, hs.data(), 4 * (kRhgBands + 1), hipMemcpyHostToDevice))) break;
        hipLaunchKernelGGL(rhg_adj_kernel, dim3(blocks), dim3(256), 0, 0, n, R, r, th, bstart, sr, st,
                           sid, deg, nullptr, nullptr);
        if (!ok(hipGetLastError()) || !ok(hipDeviceSynchronize())) break;
        std::vector<uint32_t> hd(n);
        if (!ok(hipMemcpy(hd.data(), deg, 4 * n, hipMemcpyDeviceToHost))) break;
        uint64_t m = 0;
        for (uint64_t i = 0; i < n; ++i) m += hd[i];
        if (!colptr) {
            result = (int64_t)m;
            break;
        }
        if ((int64_t)m > cap) {
            result = -3;
            break;
        }
        std::vector<uint64_t> cp(n + 1, 0);
        for (uint64_t i = 0; i < n; ++i) cp[i + 1] = cp[i] + hd[i];
        if (!ok(hipMemcpy(colptr, cp.data(), 8 * (n + 1), hipMemcpyHostToDevice))) break;
        hipLaunchKernelGGL(rhg_adj_kernel, dim3(blocks), dim3(256), 0, 0, n, R, r, th, bstart, sr, st,
                           sid, nullptr, colptr, src);
        if (!ok(hipGetLastError()) || !ok(hipDeviceSynchronize())) break;
        result = (int64_t)m;
    } while (false);
    for (void *p : {(void *)r, (void *)th, (void *)sr, (void *)st, (void *)k1, (void *)k2, (void *)v1,
                    (void *)v2, (void *)sid, (void *)bcount, (void *)bstart, (void *)deg, tmp})
        if (p) (void)hipFree(p);
    return result;
}

}  // extern "C"
