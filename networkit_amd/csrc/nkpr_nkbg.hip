// NetworKit binary graph files (nkbg002) straight to the flat lists the device library takes.
//
// Format (include/networkit/io/NetworkitBinaryGraph.hpp:20-141, networkit/cpp/io/
// NetworkitBinaryGraph.md): a header of ten 64-bit words, one flag byte per node (bit 0: the
// node exists), the first node of every chunk but the first, then four sections, each opened
// by the byte offsets of chunks 1 .. chunks-1 inside it: the adjacency lists (plus their
// entry count) and the transposed lists (plus theirs), both as per-node varint counts followed
// by varint neighbour ids, and the weights of both.  An undirected graph stores at node u the
// neighbours v <= u in the adjacency lists and v >= u in the transposed lists.
//
// nkpr_nkbg_read restates NetworkitBinaryReader::read (NetworkitBinaryReader.cpp:19-237)
// without building NetworKit's per-node vectors: the file is memory-mapped and decoded chunk
// by chunk on host threads, once, into per-chunk buffers in final order, which are then moved
// into CSR arrays at the summed offsets:
//   pull lists   in-lists (directed: the transposed lists) or the adjacency (undirected: u's
//                adjacency list followed by its transposed list without self-loops, the order
//                the reference's reader appends them in), ready for nkpr_graph_create;
//   out-lists    optionally, for a directed graph (the adjacency lists).
// Every read is bounds-checked against the file; a malformed file is an error, not a crash.
//
// nkpr_nkbg_write restates NetworkitBinaryWriter::write (NetworkitBinaryWriter.cpp:19-350)
// from the same arrays (out-lists, plus in-lists when directed), including the weight format
// detection and the chunking (nodes / chunks nodes per chunk, the last takes the rest).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "nkpr_internal.hpp"

namespace nkpr {
namespace {

constexpr char kMagic[8] = {'n', 'k', 'b', 'g', '0', '0', '2', '\0'};
enum WeightFormat { kNone = 0, kVarint = 1, kSignedVarint = 2, kDouble = 3, kFloat = 4 };

struct MappedFile {
    const uint8_t *p = nullptr;
    size_t size = 0;
    int fd = -1;
    explicit MappedFile(const char *path) {
        fd = open(path, O_RDONLY);
        NKPR_CHECK(fd >= 0, std::string("cannot open ") + path);
        struct stat st;
        if (fstat(fd, &st) != 0) {
            close(fd);
            fd = -1;
            throw Error(NKPR_ERR_INVALID, std::string("cannot stat ") + path);
        }
        size = (size_t)st.st_size;
        if (size) {
            void *m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m == MAP_FAILED) {
                close(fd);
                fd = -1;
                throw Error(NKPR_ERR_INVALID, std::string("cannot map ") + path);
            }
            p = static_cast<const uint8_t *>(m);
        }
    }
    ~MappedFile() {
        if (p) munmap(const_cast<uint8_t *>(p), size);
        if (fd >= 0) close(fd);
    }
    uint64_t u64(size_t at) const {
        NKPR_CHECK(at <= size && size - at >= 8, "nkbg file truncated");
        uint64_t v;
        std::memcpy(&v, p + at, 8);
        return v;
    }
};

// A bounded byte stream over [pos, end) of the file.
struct Stream {
    const uint8_t *p;
    size_t pos, end;
    // NetworkitBinaryGraph.hpp:108-126: the lowest set bit of the first byte counts the data
    // bytes that follow (a first byte of 0: eight); the first byte's upper bits are the lowest
    // bits of the value, the data bytes the rest, little-endian
    uint64_t varint() {
        NKPR_CHECK(pos < end, "nkbg file truncated (varint)");
        const uint8_t head = p[pos];
        const int n = head ? __builtin_ctz(head) : 8;
        NKPR_CHECK(end - pos > (size_t)n, "nkbg file truncated (varint)");
        // pos < end, so end - pos does not wrap
        uint64_t v = head ? (uint64_t)(head >> (n + 1)) : 0;
        int bits = head ? 7 - n : 0;
        for (int i = 0; i < n; ++i, bits += 8) v |= (uint64_t)p[pos + 1 + i] << bits;
        pos += (size_t)n + 1;
        return v;
    }
    // steps over n varints without assembling them (the first byte gives each one's length)
    void skip_varints(uint64_t n) {
        for (uint64_t i = 0; i < n; ++i) {
            NKPR_CHECK(pos < end, "nkbg file truncated (varint)");
            const uint8_t head = p[pos];
            const size_t len = (size_t)(head ? __builtin_ctz(head) : 8) + 1;
            NKPR_CHECK(end - pos >= len, "nkbg file truncated (varint)");
            pos += len;
        }
    }
    void skip_weights(int fmt, uint64_t n) {
        switch (fmt) {
        case kVarint:
        case kSignedVarint: skip_varints(n); break;
        case kDouble:
        case kFloat: {
            const uint64_t size = fmt == kDouble ? 8 : 4;
            NKPR_CHECK(pos <= end && n <= (end - pos) / size, "nkbg file truncated (weights)");
            pos += (size_t)(n * size);
            break;
        }
        default: break;
        }
    }
    double weight(int fmt) {
        switch (fmt) {
        case kVarint: return (double)varint();
        case kSignedVarint: {
            const uint64_t z = varint();
            return (double)(int64_t)((z >> 1) ^ (~(z & 1) + 1));
        }
        case kDouble: {
            NKPR_CHECK(pos <= end && end - pos >= 8, "nkbg file truncated (weights)");
            double w;
            std::memcpy(&w, p + pos, 8);
            pos += 8;
            return w;
        }
        case kFloat: {
            NKPR_CHECK(pos <= end && end - pos >= 4, "nkbg file truncated (weights)");
            float w;
            std::memcpy(&w, p + pos, 4);
            pos += 4;
            return (double)w;
        }
        default: return 1.0;
        }
    }
};

// Runs f(chunk) for every chunk on up to hardware_concurrency host threads; the first error
// is rethrown after all threads have finished.
template <typename F> void for_chunks(uint64_t chunks, F &&f) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = (unsigned)std::min<uint64_t>(chunks, hw);
    std::atomic<uint64_t> next{0};
    std::vector<std::string> err(nt);
    std::vector<int> code(nt, 0);
    auto body = [&](unsigned t) {
        try {
            for (uint64_t c = next++; c < chunks; c = next++) f(c);
        } catch (const Error &e) {
            err[t] = e.what();
            code[t] = e.code;
            next = chunks;
        } catch (const std::exception &e) {
            err[t] = e.what();
            code[t] = NKPR_ERR_INVALID;
            next = chunks;
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(body, t);
    if (nt) body(0);
    for (auto &x : th) x.join();
    for (unsigned t = 0; t < nt; ++t)
        if (code[t]) throw Error(code[t], err[t]);
}

template <typename T> T *alloc_array(uint64_t count) {
    T *a = static_cast<T *>(std::malloc(std::max<uint64_t>(count, 1) * sizeof(T)));
    if (!a) throw std::bad_alloc();
    return a;
}

}  // namespace

void nkbg_read_impl(const char *path, int32_t with_out_lists, nkpr_nkbg *out) {
    NKPR_CHECK(path != nullptr && out != nullptr, "path or out is NULL");
    std::memset(out, 0, sizeof(*out));
    MappedFile f(path);
    NKPR_CHECK(f.size >= 80 && std::memcmp(f.p, kMagic, 8) == 0,
               "Reader expected another magic value");
    const uint64_t features = f.u64(16), nodes = f.u64(24), chunks = f.u64(32);
    const uint64_t o_base = f.u64(40), o_adj = f.u64(48), o_tr = f.u64(56), o_w = f.u64(64),
                   o_tw = f.u64(72);
    const bool directed = (features & 1) != 0;
    const int fmt = (int)((features & 0xE) >> 1);
    NKPR_CHECK(fmt <= kFloat, "unknown weight format");
    NKPR_CHECK(nodes < 0x7FFFFFFFull, "graph too large for int32 node ids");
    NKPR_CHECK(nodes == 0 || (chunks >= 1 && chunks <= nodes), "bad chunk count");
    NKPR_CHECK(o_base <= f.size && f.size - o_base >= nodes + 8 * (chunks ? chunks - 1 : 0),
               "nkbg file truncated (base data)");
    out->nodes = nodes;
    out->directed = directed;
    out->weighted = fmt != kNone;
    out->weight_format = fmt;
    out->exists = alloc_array<uint8_t>(nodes);
    for (uint64_t i = 0; i < nodes; ++i) out->exists[i] = f.p[o_base + i] & 1;
    out->colptr = alloc_array<uint64_t>(nodes + 1);
    out->colptr[0] = 0;
    if (nodes == 0) {
        out->src = alloc_array<int32_t>(0);
        return;
    }
    std::vector<uint64_t> first(chunks + 1);
    first[0] = 0;
    for (uint64_t c = 1; c < chunks; ++c) first[c] = f.u64(o_base + nodes + 8 * (c - 1));
    first[chunks] = nodes;
    for (uint64_t c = 0; c < chunks; ++c)
        NKPR_CHECK(first[c] <= first[c + 1], "chunk starts out of order");
    const uint64_t table = 8 * (chunks - 1);
    out->edges = f.u64(o_adj + table);
    // stream of chunk c in a section: offset table, then (adjacency and transposed lists) the
    // entry count, then the data; the chunk's data starts at the table's entry for c
    auto stream = [&](uint64_t sec, uint64_t extra, uint64_t c) {
        const uint64_t off = first[c] ? f.u64(sec + 8 * (c - 1)) : 0;
        Stream s{f.p, (size_t)(sec + table + extra + off), f.size};
        NKPR_CHECK(s.pos <= f.size, "nkbg section offset outside the file");
        return s;
    };
    const bool want_out = directed && with_out_lists;
    const bool weighted = fmt != kNone;
    // one pass: every chunk decodes its nodes' pull lists (and out-lists) into buffers of its
    // own, in final order (u's adjacency entries, then its transposed entries without the
    // self-loop when undirected); the offsets are then summed and the buffers moved into place
    struct ChunkLists {
        std::vector<int32_t> ids, out_ids;
        std::vector<double> w, out_w;
    };
    std::vector<ChunkLists> cl(chunks);
    std::vector<uint64_t> pull(nodes, 0), outc(want_out ? nodes : 0, 0);
    std::vector<uint64_t> loops(chunks, 0);
    for_chunks(chunks, [&](uint64_t c) {
        Stream a = stream(o_adj, 8, c), t = stream(o_tr, 8, c);
        Stream aw{f.p, 0, 0}, tw{f.p, 0, 0};
        if (weighted) {
            aw = stream(o_w, 0, c);
            tw = stream(o_tw, 0, c);
        }
        ChunkLists &L = cl[c];
        // a directed graph read without out-lists needs only the in-lists: its out-lists are
        // stepped over (their self-loops are counted in the in-lists, where each appears too)
        const bool skip_out = directed && !want_out;
        for (uint64_t u = first[c]; u < first[c + 1]; ++u) {
            const uint64_t nout = a.varint(), nin = t.varint();
            NKPR_CHECK(nout <= f.size && nin <= f.size, "nkbg list length outside the file");
            const size_t before = L.ids.size();
            if (skip_out) {
                a.skip_varints(nout);
                if (weighted) aw.skip_weights(fmt, nout);
            }
            for (uint64_t j = 0; j < (skip_out ? 0 : nout); ++j) {
                const uint64_t v = a.varint();
                NKPR_CHECK(v < nodes, "nkbg neighbour id out of range");
                const double w = weighted ? aw.weight(fmt) : 1.0;
                loops[c] += v == u;
                if (directed) {
                    if (want_out) {
                        L.out_ids.push_back((int32_t)v);
                        if (weighted) L.out_w.push_back(w);
                    }
                } else {
                    L.ids.push_back((int32_t)v);
                    if (weighted) L.w.push_back(w);
                }
            }
            for (uint64_t j = 0; j < nin; ++j) {
                const uint64_t v = t.varint();
                NKPR_CHECK(v < nodes, "nkbg neighbour id out of range");
                const double w = weighted ? tw.weight(fmt) : 1.0;
                if (v == u) {
                    if (!directed) continue;
                    if (skip_out) ++loops[c];
                }
                L.ids.push_back((int32_t)v);
                if (weighted) L.w.push_back(w);
            }
            pull[u] = L.ids.size() - before;
            if (want_out) outc[u] = nout;
        }
    });
    for (uint64_t u = 0; u < nodes; ++u) out->colptr[u + 1] = out->colptr[u] + pull[u];
    const uint64_t m = out->colptr[nodes];
    out->src = alloc_array<int32_t>(m);
    if (weighted) out->weight = alloc_array<double>(m);
    if (want_out) {
        out->out_ptr = alloc_array<uint64_t>(nodes + 1);
        out->out_ptr[0] = 0;
        for (uint64_t u = 0; u < nodes; ++u) out->out_ptr[u + 1] = out->out_ptr[u] + outc[u];
        out->out_dst = alloc_array<int32_t>(out->out_ptr[nodes]);
        if (weighted) out->out_weight = alloc_array<double>(out->out_ptr[nodes]);
    }
    for_chunks(chunks, [&](uint64_t c) {
        ChunkLists &L = cl[c];
        const uint64_t at = out->colptr[first[c]];
        NKPR_CHECK(at + L.ids.size() == out->colptr[first[c + 1]], "nkbg chunk lists do not add up");
        if (!L.ids.empty()) std::memcpy(out->src + at, L.ids.data(), 4 * L.ids.size());
        if (weighted && !L.w.empty()) std::memcpy(out->weight + at, L.w.data(), 8 * L.w.size());
        if (want_out) {
            const uint64_t ao = out->out_ptr[first[c]];
            if (!L.out_ids.empty()) std::memcpy(out->out_dst + ao, L.out_ids.data(), 4 * L.out_ids.size());
            if (weighted && !L.out_w.empty())
                std::memcpy(out->out_weight + ao, L.out_w.data(), 8 * L.out_w.size());
        }
        std::vector<int32_t>().swap(L.ids);  // released as soon as copied
        std::vector<double>().swap(L.w);
        std::vector<int32_t>().swap(L.out_ids);
        std::vector<double>().swap(L.out_w);
    });
    for (uint64_t c = 0; c < chunks; ++c) out->self_loops += loops[c];
}

void nkbg_free_impl(nkpr_nkbg *g) {
    if (!g) return;
    std::free(g->colptr);
    std::free(g->src);
    std::free(g->weight);
    std::free(g->out_ptr);
    std::free(g->out_dst);
    std::free(g->out_weight);
    std::free(g->exists);
    std::memset(g, 0, sizeof(*g));
}

namespace {

size_t varint_put(uint64_t value, uint8_t *b) {
    // NetworkitBinaryGraph.hpp:75-105
    if (value == 0) {
        b[0] = 1;
        return 1;
    }
    int nb;
    if (value >= (1ull << 56)) {
        b[0] = 0;
        nb = 8;
    } else {
        const int bits = 64 - __builtin_clzll(value);
        nb = (bits - 1) / 7;
        b[0] = (uint8_t)((1u << nb) | (uint8_t)(value << (nb + 1)));
        value >>= 7 - nb;
    }
    for (int i = 0; i < nb; ++i, value >>= 8) b[i + 1] = (uint8_t)(value & 0xFF);
    return (size_t)nb + 1;
}

size_t weight_put(double w, int fmt, uint8_t *b) {
    switch (fmt) {
    case kVarint: return varint_put((uint64_t)w, b);
    case kSignedVarint: {
        const int64_t x = (int64_t)w;
        return varint_put(((uint64_t)x << 1) ^ (x < 0 ? ~0ull : 0ull), b);
    }
    case kDouble: std::memcpy(b, &w, 8); return 8;
    case kFloat: {
        const float x = (float)w;
        std::memcpy(b, &x, 4);
        return 4;
    }
    default: return 0;
    }
}

}  // namespace

void nkbg_write_impl(const char *path, const nkpr_nkbg *g, uint64_t chunks, int32_t weight_format) {
    NKPR_CHECK(path != nullptr && g != nullptr, "path or graph is NULL");
    const uint64_t nodes = g->nodes;
    NKPR_CHECK(nodes == 0 || (g->out_ptr && g->out_dst), "out-lists are required");
    NKPR_CHECK(!g->directed || nodes == 0 || (g->colptr && g->src),
               "a directed graph needs its in-lists (colptr, src)");
    NKPR_CHECK(nodes < 0x7FFFFFFFull, "graph too large for int32 node ids");
    NKPR_CHECK(chunks >= 1, "chunks must be >= 1");
    const bool weighted = g->out_weight != nullptr;
    int fmt = weight_format;
    if (fmt < 0) {
        // NetworkitBinaryWriter.cpp:28-52 (autoDetect) over every stored weight
        fmt = kNone;
        if (weighted) {
            bool uns = true, ints = true, flt = true;
            for (uint64_t e = 0; e < g->out_ptr[nodes]; ++e) {
                const double w = g->out_weight[e];
                if (w < 0) uns = false;
                if (!(std::fabs(w) < 9.2e18) || w != (double)(int64_t)w) ints = false;
                if (w != (double)(float)w) flt = false;
            }
            fmt = ints ? (uns ? kVarint : kSignedVarint) : (flt ? kFloat : kDouble);
        }
    }
    NKPR_CHECK(fmt >= kNone && fmt <= kFloat, "unknown weight format");
    NKPR_CHECK(fmt == kNone || weighted, "a weight format needs weights");
    if (chunks > nodes) chunks = nodes;  // NetworkitBinaryWriter.cpp:133-137
    const bool directed = g->directed != 0;
    auto out_list = [&](uint64_t u, auto &&f) {
        for (uint64_t e = g->out_ptr[u]; e < g->out_ptr[u + 1]; ++e)
            f((uint64_t)g->out_dst[e], weighted ? g->out_weight[e] : 1.0);
    };
    auto in_list = [&](uint64_t u, auto &&f) {
        for (uint64_t e = g->colptr[u]; e < g->colptr[u + 1]; ++e)
            f((uint64_t)g->src[e], (g->weight && weighted) ? g->weight[e] : 1.0);
    };
    // the two lists of node u as the writer stores them
    auto lists = [&](uint64_t u, auto &&fa, auto &&ft) {
        if (directed) {
            out_list(u, fa);
            in_list(u, ft);
        } else {
            out_list(u, [&](uint64_t v, double w) {
                if (v <= u) fa(v, w);
            });
            out_list(u, [&](uint64_t v, double w) {
                if (v >= u) ft(v, w);
            });
        }
    };
    std::vector<uint64_t> first(chunks + 1, 0);
    for (uint64_t c = 1; c < chunks; ++c) first[c] = first[c - 1] + nodes / chunks;
    first[chunks] = nodes;
    std::vector<std::vector<uint8_t>> adj(chunks), tr(chunks), aw(chunks), tw(chunks);
    std::vector<uint64_t> na(chunks, 0), nt(chunks, 0);
    for_chunks(chunks, [&](uint64_t c) {
        uint8_t b[10];
        auto put = [&](std::vector<uint8_t> &s, const uint8_t *x, size_t k) { s.insert(s.end(), x, x + k); };
        for (uint64_t u = first[c]; u < first[c + 1]; ++u) {
            uint64_t co = 0, ci = 0;
            lists(u, [&](uint64_t, double) { ++co; }, [&](uint64_t, double) { ++ci; });
            put(adj[c], b, varint_put(co, b));
            put(tr[c], b, varint_put(ci, b));
            lists(u,
                  [&](uint64_t v, double w) {
                      put(adj[c], b, varint_put(v, b));
                      put(aw[c], b, weight_put(w, fmt, b));
                  },
                  [&](uint64_t v, double w) {
                      put(tr[c], b, varint_put(v, b));
                      put(tw[c], b, weight_put(w, fmt, b));
                  });
            na[c] += co;
            nt[c] += ci;
        }
    });
    auto prefix = [&](const std::vector<std::vector<uint8_t>> &s) {
        std::vector<uint64_t> p(chunks, 0);
        uint64_t acc = 0;
        for (uint64_t c = 0; c < chunks; ++c) p[c] = (acc += s[c].size());
        return p;
    };
    const auto pa = prefix(adj), pt = prefix(tr), paw = prefix(aw), ptw = prefix(tw);
    uint64_t a_size = 0, t_size = 0;
    for (uint64_t c = 0; c < chunks; ++c) {
        a_size += na[c];
        t_size += nt[c];
    }
    const uint64_t table = chunks ? 8 * (chunks - 1) : 0;
    uint64_t hdr[10];
    std::memcpy(&hdr[0], kMagic, 8);
    hdr[1] = 0;
    hdr[2] = (directed ? 1u : 0u) | (((uint64_t)fmt << 1) & 0xE);
    hdr[3] = nodes;
    hdr[4] = chunks;
    hdr[5] = 80;
    hdr[6] = hdr[5] + nodes + table;
    hdr[7] = hdr[6] + table + 8 + (chunks ? pa[chunks - 1] : 0);
    if (fmt != kNone) {
        hdr[8] = hdr[7] + table + 8 + (chunks ? pt[chunks - 1] : 0);
        hdr[9] = hdr[8] + table + (chunks ? paw[chunks - 1] : 0);
    } else {
        hdr[8] = hdr[9] = 0;
    }
    FILE *fp = std::fopen(path, "wb");
    NKPR_CHECK(fp != nullptr, std::string("cannot open ") + path + " for writing");
    bool ok = true;
    auto w = [&](const void *p, size_t k) {
        if (k && std::fwrite(p, 1, k, fp) != k) ok = false;
    };
    w(hdr, sizeof(hdr));
    std::vector<uint8_t> flags(nodes);
    for (uint64_t u = 0; u < nodes; ++u) flags[u] = (!g->exists || g->exists[u]) ? 1 : 0;
    w(flags.data(), nodes);
    for (uint64_t c = 1; c < chunks; ++c) w(&first[c], 8);
    auto section = [&](const std::vector<uint64_t> &p, const uint64_t *count,
                       const std::vector<std::vector<uint8_t>> &s) {
        for (uint64_t c = 1; c < chunks; ++c) w(&p[c - 1], 8);
        if (count) w(count, 8);
        for (uint64_t c = 0; c < chunks; ++c) w(s[c].data(), s[c].size());
    };
    section(pa, &a_size, adj);
    section(pt, &t_size, tr);
    section(paw, nullptr, aw);
    section(ptw, nullptr, tw);
    ok = (std::fclose(fp) == 0) && ok;
    NKPR_CHECK(ok, std::string("writing ") + path + " failed");
}

}  // namespace nkpr
