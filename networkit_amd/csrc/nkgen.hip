// Seeded synthetic-graph generators for the benchmark and test workloads (libnkgen.so).
//
// Not part of the PageRank boundary: these produce the input graphs of the workloads
// (directed R-MAT, Erdos-Renyi, Barabasi-Albert) with a counter-based random stream, so a
// graph depends only on its parameters and seed -- not on the thread count, the device, or
// whether it was made on the host or on the GPU.  NetworKit's own generators cannot serve
// here: its RmatGenerator always builds undirected graphs and scans sequentially
// (networkit/cpp/generators/RmatGenerator.cpp:33,112-117) and its Erdos-Renyi output depends
// on the OpenMP thread count (include/networkit/generators/ErdosRenyiEnumerator.hpp:87-136).
#include <hip/hip_runtime.h>

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

}  // extern "C"
