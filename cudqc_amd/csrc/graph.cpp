// Host preprocessing: text loader, directed two-hop lists, degree tables, level-0
// peel / cover / condense. Semantics follow the reference files cited per function;
// the code is organised for this build (flat CSR arrays, per-thread epoch stamps).
#include "graph.hpp"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

namespace dqc {

// Graphs up to this many vertices are preprocessed on one thread: on small inputs the
// cost of waking a thread team exceeds the work (and a capped CPU share makes it worse).
constexpr int32_t PAR_MIN = 16384;

// ---------------------------------------------------------------------------
// reference src/common.cpp:15-108. n = number of lines; ids must lie in [0, n);
// lists sorted, duplicates and self loops removed; m counted after the cleanup.
int load_text(const std::string& path, Graph& g, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) { err = "cannot open graph file " + path; return 1; }
    std::vector<char> buf;
    {
        std::fseek(f, 0, SEEK_END);
        long sz = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        buf.resize(sz > 0 ? (size_t)sz : 0);
        if (sz > 0 && std::fread(buf.data(), 1, buf.size(), f) != buf.size()) {
            std::fclose(f); err = "short read on " + path; return 1;
        }
        std::fclose(f);
    }
    const size_t len = buf.size();
    // line starts
    std::vector<size_t> starts;
    starts.reserve(len / 8 + 1);
    if (len > 0) starts.push_back(0);
    for (size_t i = 0; i < len; ++i)
        if (buf[i] == '\n' && i + 1 < len) starts.push_back(i + 1);
    const int32_t n = (int32_t)starts.size();
    // pass 1: parse arcs per line (parallel over lines), collect (u, v) pairs
    std::vector<std::vector<int32_t>> outl(n);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : bad) if (n > PAR_MIN)
    for (int32_t u = 0; u < n; ++u) {
        size_t p = starts[u];
        size_t e = (u + 1 < n) ? starts[u + 1] : len;
        std::vector<int32_t>& o = outl[u];
        while (p < e) {
            char c = buf[p];
            if (c == ' ' || c == '\t' || c == '\r' || c == '\n') { ++p; continue; }
            bool neg = false;
            if (c == '-' || c == '+') { neg = (c == '-'); ++p; }
            if (p >= e || buf[p] < '0' || buf[p] > '9') break;  // stream extraction stops
            int64_t v = 0;
            while (p < e && buf[p] >= '0' && buf[p] <= '9') { v = v * 10 + (buf[p] - '0'); if (v > (1ll << 40)) v = 1ll << 40; ++p; }
            if (neg) v = -v;
            if (v < 0 || v >= n) { bad = 1; break; }
            if ((int32_t)v != u) o.push_back((int32_t)v);
        }
        std::sort(o.begin(), o.end());
        o.erase(std::unique(o.begin(), o.end()), o.end());
    }
    if (bad) { err = "graph file holds a vertex id outside [0, n)"; return 2; }
    g.n = n;
    g.out.off.assign((size_t)n + 1, 0);
    for (int32_t u = 0; u < n; ++u) g.out.off[u + 1] = g.out.off[u] + outl[u].size();
    g.m = (int64_t)g.out.off[n];
    g.out.nbr.resize(g.m);
    std::vector<uint64_t> indeg((size_t)n + 1, 0);
    for (int32_t u = 0; u < n; ++u) {
        std::copy(outl[u].begin(), outl[u].end(), g.out.nbr.begin() + g.out.off[u]);
        for (int32_t v : outl[u]) indeg[v + 1]++;
        std::vector<int32_t>().swap(outl[u]);
    }
    // in-lists by counting sort; scanning u ascending keeps them sorted
    g.in.off.assign((size_t)n + 1, 0);
    for (int32_t v = 0; v < n; ++v) g.in.off[v + 1] = g.in.off[v] + indeg[v + 1];
    g.in.nbr.resize(g.m);
    std::vector<uint64_t> cur(g.in.off.begin(), g.in.off.end() - 1);
    for (int32_t u = 0; u < n; ++u)
        for (uint64_t k = g.out.off[u]; k < g.out.off[u + 1]; ++k) g.in.nbr[cur[g.out.nbr[k]]++] = u;
    g.two.off.assign((size_t)n + 1, 0);
    g.two.nbr.clear();
    return 0;
}

// ---------------------------------------------------------------------------
// Binary graph cache (no reference counterpart: the reference parses the text file on
// every run, src/common.cpp:28-57).
namespace {
constexpr char BIN_MAGIC[8] = {'D', 'Q', 'C', 'G', 'R', 'P', 'H', '1'};

// Word-wise 64-bit hash over a byte stream (multiply-xorshift per 8-byte word; the
// tail is zero-padded). Detects truncation and corruption of a cache file.
struct StreamHash {
    uint64_t h = 0x9E3779B97F4A7C15ull, pending = 0;
    int npend = 0;
    void mix(uint64_t w) {
        h ^= w + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        h *= 0xFF51AFD7ED558CCDull;
        h ^= h >> 33;
    }
    void add(const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        while (n && npend) {
            pending |= (uint64_t)*b++ << (8 * npend);
            --n;
            if (++npend == 8) { mix(pending); pending = 0; npend = 0; }
        }
        for (; n >= 8; n -= 8, b += 8) {
            uint64_t w;
            std::memcpy(&w, b, 8);
            mix(w);
        }
        while (n--) { pending |= (uint64_t)*b++ << (8 * npend); ++npend; }
    }
    uint64_t final() {
        if (npend) mix(pending);
        mix(0x5A17ull);
        return h;
    }
};

bool put(FILE* f, StreamHash& hs, const void* p, size_t n) {
    hs.add(p, n);
    return n == 0 || std::fwrite(p, 1, n, f) == n;
}
bool get(FILE* f, StreamHash& hs, void* p, size_t n) {
    if (n && std::fread(p, 1, n, f) != n) return false;
    hs.add(p, n);
    return true;
}
bool put_csr(FILE* f, StreamHash& hs, const Csr& c, int32_t n) {
    return put(f, hs, c.off.data(), sizeof(uint64_t) * ((size_t)n + 1)) &&
           put(f, hs, c.nbr.data(), sizeof(int32_t) * c.nbr.size());
}
// reads one CSR whose entry count is given by its own last offset (checked against max_len)
bool get_csr(FILE* f, StreamHash& hs, Csr& c, int32_t n, uint64_t max_len, std::string& err) {
    c.off.resize((size_t)n + 1);
    if (!get(f, hs, c.off.data(), sizeof(uint64_t) * ((size_t)n + 1))) { err = "truncated binary graph"; return false; }
    if (c.off[0] != 0) { err = "binary graph: offsets do not start at 0"; return false; }
    for (int32_t v = 0; v < n; ++v)
        if (c.off[v + 1] < c.off[v]) { err = "binary graph: decreasing offsets"; return false; }
    if (c.off[n] > max_len) { err = "binary graph: list longer than the file allows"; return false; }
    c.nbr.resize(c.off[n]);
    if (!get(f, hs, c.nbr.data(), sizeof(int32_t) * c.nbr.size())) { err = "truncated binary graph"; return false; }
    for (int32_t x : c.nbr)
        if (x < 0 || x >= n) { err = "binary graph: id outside [0, n)"; return false; }
    return true;
}
}  // namespace

bool save_binary(const std::string& path, const Graph& g, std::string& err) {
    const bool two = g.two.off.size() == (size_t)g.n + 1 && g.two.off[g.n] > 0;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = "cannot write " + path; return false; }
    StreamHash hs;
    const int32_t flags = two ? 1 : 0;
    const int64_t m = g.m;
    bool ok = put(f, hs, BIN_MAGIC, 8) && put(f, hs, &g.n, 4) && put(f, hs, &flags, 4) && put(f, hs, &m, 8) &&
              put_csr(f, hs, g.out, g.n) && put_csr(f, hs, g.in, g.n) && (!two || put_csr(f, hs, g.two, g.n));
    const uint64_t sum = hs.final();
    ok = ok && std::fwrite(&sum, 8, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) err = "short write on " + path;
    return ok;
}

bool is_binary_graph(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char m[8] = {0};
    const bool r = std::fread(m, 1, 8, f) == 8 && std::memcmp(m, BIN_MAGIC, 8) == 0;
    std::fclose(f);
    return r;
}

bool load_binary(const std::string& path, Graph& g, std::string& err, bool* has_twohop) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) { err = "cannot open graph file " + path; return false; }
    std::fseek(f, 0, SEEK_END);
    const long long fsz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    StreamHash hs;
    char magic[8];
    int32_t n = 0, flags = 0;
    int64_t m = 0;
    bool ok = get(f, hs, magic, 8) && std::memcmp(magic, BIN_MAGIC, 8) == 0 && get(f, hs, &n, 4) &&
              get(f, hs, &flags, 4) && get(f, hs, &m, 8) && n >= 0 && m >= 0 && (flags & ~1) == 0;
    if (!ok) { std::fclose(f); err = "not a binary graph file: " + path; return false; }
    // every list is bounded by the file size (an offset table cannot ask for more)
    const uint64_t max_len = fsz > 0 ? (uint64_t)fsz / 4 : 0;
    Graph r;
    r.n = n;
    r.m = m;
    ok = get_csr(f, hs, r.out, n, max_len, err) && get_csr(f, hs, r.in, n, max_len, err);
    if (ok && (r.out.off[n] != (uint64_t)m || r.in.off[n] != (uint64_t)m)) { err = "binary graph: arc counts differ"; ok = false; }
    if (ok && (flags & 1)) ok = get_csr(f, hs, r.two, n, max_len, err);
    uint64_t sum = 0;
    if (ok && (std::fread(&sum, 8, 1, f) != 1 || sum != hs.final())) { err = "binary graph: checksum mismatch"; ok = false; }
    std::fclose(f);
    if (!ok) return false;
    if (!(flags & 1)) { r.two.off.assign((size_t)n + 1, 0); r.two.nbr.clear(); }
    if (has_twohop) *has_twohop = (flags & 1) != 0;
    g = std::move(r);
    return true;
}

int load_graph_file(const std::string& path, Graph& g, std::string& err, int threads) {
    if (is_binary_graph(path)) {
        bool two = false;
        if (!load_binary(path, g, err, &two)) return 2;
        if (!two) build_twohop(g, threads);
        return 0;
    }
    const int rc = load_text(path, g, err);
    if (rc) return rc;
    build_twohop(g, threads);
    return 0;
}

// ---------------------------------------------------------------------------
// reference src/common.cpp:111-509 (GenLevel2NBs). For a vertex v with both out-
// and in-arcs: B = mutual neighbours, O = out-only, I = in-only. The one-direction
// sets are pruned to a greatest fixed point (w stays in I while an arc reaches it
// from O u B; w stays in O while it has an arc into I u B; both sides use the
// previous round's sets). The list is B u O* u I* plus every other x != v that has
// an arc from O*uB, an arc into O*uB, an arc from I*uB and an arc into I*uB.
namespace {
void twohop_impl(Graph& g, int threads, const std::vector<uint8_t>* src, const std::vector<uint8_t>* dst,
                 std::vector<uint32_t>* full_len);
}

void build_twohop(Graph& g, int threads) { twohop_impl(g, threads, nullptr, nullptr, nullptr); }

void build_twohop_restricted(Graph& g, int threads, const std::vector<uint8_t>& src, const std::vector<uint8_t>& dst,
                             std::vector<uint32_t>& full_len) {
    twohop_impl(g, threads, &src, &dst, &full_len);
}

namespace {
void twohop_impl(Graph& g, int threads, const std::vector<uint8_t>* src, const std::vector<uint8_t>* dst,
                 std::vector<uint32_t>* full_len) {
    const int32_t n = g.n;
    if (full_len) full_len->assign((size_t)n, 0);
    const uint64_t* oo = g.out.off.data();
    const int32_t* on = g.out.nbr.data();
    const uint64_t* io = g.in.off.data();
    const int32_t* in = g.in.nbr.data();
    if (threads < 1) threads = 1;
    // per-chunk outputs, stitched in vertex order afterwards
    std::vector<uint32_t> cnt((size_t)n, 0);
    const int32_t chunk = 1024;
    const int32_t nchunks = (n + chunk - 1) / chunk;
    std::vector<std::vector<int32_t>> chunk_lists(nchunks);
    // DQC_PROGRESS: one line per tenth of the chunks done (long builds on large graphs)
    const bool progress = std::getenv("DQC_PROGRESS") != nullptr;
    std::atomic<int32_t> chunks_done{0};
#pragma omp parallel num_threads(threads) if (n > PAR_MIN)
    {
        // state[x] bit layout, per vertex being processed (reset after each vertex):
        //   bit0 out-neighbour, bit1 in-neighbour, bit2 in the list, bits 8.. epoch stamp
        std::vector<uint32_t> rin((size_t)n, 0), rout((size_t)n, 0);  // survival round stamps
        std::vector<uint8_t> nb((size_t)n, 0), inl((size_t)n, 0), c4((size_t)n, 0);
        std::vector<int32_t> B, O, I, nO, nI, lst, touched;
#pragma omp for schedule(dynamic, 1)
        for (int32_t ch = 0; ch < nchunks; ++ch) {
            if (progress) {
                const int32_t d = chunks_done.fetch_add(1);
                if (nchunks >= 10 && d % (nchunks / 10) == 0)
                    std::fprintf(stderr, "dqc: two-hop lists %d/%d chunks\n", d, nchunks);
            }
            std::vector<int32_t>& out = chunk_lists[ch];
            int32_t v0 = ch * chunk, v1 = std::min(n, v0 + chunk);
            for (int32_t v = v0; v < v1; ++v) {
                if (oo[v + 1] == oo[v] || io[v + 1] == io[v]) continue;
                if (src && !(*src)[v]) continue;
                B.clear(); O.clear(); I.clear();
                for (uint64_t k = oo[v]; k < oo[v + 1]; ++k) nb[on[k]] |= 1;
                for (uint64_t k = io[v]; k < io[v + 1]; ++k) nb[in[k]] |= 2;
                for (uint64_t k = oo[v]; k < oo[v + 1]; ++k) {
                    int32_t w = on[k];
                    if (nb[w] == 3) B.push_back(w); else { O.push_back(w); rout[w] = 1; }
                }
                for (uint64_t k = io[v]; k < io[v + 1]; ++k) {
                    int32_t w = in[k];
                    if (nb[w] == 2) { I.push_back(w); rin[w] = 1; }
                }
                for (uint32_t round = 1;; ++round) {
                    size_t so = O.size(), si = I.size();
                    nO.clear(); nI.clear();
                    auto reach_in = [&](int32_t u) {
                        for (uint64_t k = oo[u]; k < oo[u + 1]; ++k) {
                            int32_t w = on[k];
                            if (rin[w] == round) { rin[w] = round + 1; nI.push_back(w); }
                        }
                    };
                    auto reach_out = [&](int32_t u) {
                        for (uint64_t k = io[u]; k < io[u + 1]; ++k) {
                            int32_t w = in[k];
                            if (rout[w] == round) { rout[w] = round + 1; nO.push_back(w); }
                        }
                    };
                    for (int32_t u : O) reach_in(u);
                    for (int32_t u : B) reach_in(u);
                    for (int32_t u : I) reach_out(u);
                    for (int32_t u : B) reach_out(u);
                    O.swap(nO);
                    I.swap(nI);
                    if (!(O.size() < so || I.size() < si)) break;
                }
                for (uint64_t k = oo[v]; k < oo[v + 1]; ++k) { nb[on[k]] = 0; rout[on[k]] = 0; }
                for (uint64_t k = io[v]; k < io[v + 1]; ++k) { nb[in[k]] = 0; rin[in[k]] = 0; }
                lst.clear();
                for (int32_t w : B) { lst.push_back(w); inl[w] = 1; }
                for (int32_t w : O) { lst.push_back(w); inl[w] = 1; }
                for (int32_t w : I) { lst.push_back(w); inl[w] = 1; }
                // the four arc conditions, tracked as bits of c4[x]
                touched.clear();
                auto scan = [&](const std::vector<int32_t>& src, bool outdir, uint8_t bit) {
                    const uint64_t* off = outdir ? oo : io;
                    const int32_t* nbr = outdir ? on : in;
                    for (int32_t u : src)
                        for (uint64_t k = off[u]; k < off[u + 1]; ++k) {
                            int32_t x = nbr[k];
                            if (x == v || inl[x]) continue;
                            if (c4[x] == 0) touched.push_back(x);
                            c4[x] |= bit;
                        }
                };
                scan(O, true, 1); scan(B, true, 1);    // u1 -> x, u1 in O* u B
                scan(O, false, 2); scan(B, false, 2);  // x -> u2, u2 in O* u B
                scan(I, true, 4); scan(B, true, 4);    // u3 -> x, u3 in I* u B
                scan(I, false, 8); scan(B, false, 8);  // x -> u4, u4 in I* u B
                for (int32_t x : touched) {
                    if (c4[x] == 15) lst.push_back(x);
                    c4[x] = 0;
                }
                for (int32_t w : B) inl[w] = 0;
                for (int32_t w : O) inl[w] = 0;
                for (int32_t w : I) inl[w] = 0;
                if (full_len) {  // the whole length, then only the flagged entries are kept
                    (*full_len)[v] = (uint32_t)lst.size();
                    lst.erase(std::remove_if(lst.begin(), lst.end(), [&](int32_t x) { return !(*dst)[x]; }), lst.end());
                }
                std::sort(lst.begin(), lst.end());
                cnt[v] = (uint32_t)lst.size();
                out.insert(out.end(), lst.begin(), lst.end());
            }
        }
    }
    g.two.off.assign((size_t)n + 1, 0);
    for (int32_t v = 0; v < n; ++v) g.two.off[v + 1] = g.two.off[v] + cnt[v];
    g.two.nbr.resize(g.two.off[n]);
#pragma omp parallel for schedule(static) num_threads(threads) if (n > PAR_MIN)
    for (int32_t ch = 0; ch < nchunks; ++ch) {
        if (chunk_lists[ch].empty()) continue;
        int32_t v0 = ch * chunk;
        std::copy(chunk_lists[ch].begin(), chunk_lists[ch].end(), g.two.nbr.begin() + g.two.off[v0]);
    }
}
}  // namespace

std::vector<int32_t> min_degree_table(double gamma, int len) {
    std::vector<int32_t> md((size_t)len, 0);
    for (int k = 1; k < len; ++k) md[k] = (int32_t)std::ceil(gamma * (double)(k - 1));
    return md;
}

std::vector<int32_t> bound_div_table(double gamma, int len) {
    std::vector<int32_t> ub((size_t)len, 0);
    for (int d = 0; d < len; ++d) ub[d] = (int32_t)((double)d / gamma);
    return ub;
}

// ---------------------------------------------------------------------------
namespace {
struct L0V {
    uint32_t id;
    int8_t label;  // 0 candidate, 2 covered, 3 cover, -1 pruned
    uint32_t oc, ic, l2;
};
inline int class_of(int8_t l) { return l == 2 ? 1 : l == 0 ? 2 : l == 3 ? 3 : 4; }
// enumeration order of reference inc/host_functions.hpp:74-128 at level 0
// (member degrees are all zero there).
inline bool l0_before(const L0V& a, const L0V& b) {
    int ca = class_of(a.label), cb = class_of(b.label);
    if (ca != cb) return ca < cb;
    int da = (int)std::min(a.oc, a.ic), db = (int)std::min(b.oc, b.ic);
    if (da != db) return da > db;
    if (a.l2 != b.l2) return a.l2 > b.l2;
    return a.id > b.id;
}
}  // namespace

// reference src/host_functions.cu:461-686 (h_initialize_tasks) and :690-934
// (h_condense_graph).
void level0_from_loaded(Graph& g, const int32_t* md_out, const int32_t* md_in, int min_size, int threads,
                        Level0& out) {
    const uint32_t tho = (uint32_t)md_out[min_size], thi = (uint32_t)md_in[min_size];
    std::vector<uint8_t> flag((size_t)g.n, 0);
    for (int32_t v = 0; v < g.n; ++v)
        flag[v] = (g.out.off[v + 1] - g.out.off[v] >= tho) && (g.in.off[v + 1] - g.in.off[v] >= thi);
    std::vector<uint32_t> full;
    build_twohop_restricted(g, threads, flag, flag, full);
    level0_condense(g, md_out, md_in, min_size, out, &full);
}

void level0_condense(Graph& g, const int32_t* md_out, const int32_t* md_in, int min_size,
                     Level0& res, const std::vector<uint32_t>* l2_full) {
    const int32_t n = g.n;
    const uint32_t tho = (uint32_t)md_out[min_size], thi = (uint32_t)md_in[min_size];
    const bool progress = std::getenv("DQC_PROGRESS") != nullptr;  // phase lines for large graphs
    auto note = [&](const char* what, long long x) {
        if (progress) std::fprintf(stderr, "dqc: level 0: %s %lld\n", what, x);
    };
    std::vector<L0V> v((size_t)n);
    std::vector<int32_t> alive, dropped;
    for (int32_t i = 0; i < n; ++i) {
        v[i].id = (uint32_t)i;
        v[i].oc = (uint32_t)(g.out.off[i + 1] - g.out.off[i]);
        v[i].ic = (uint32_t)(g.in.off[i + 1] - g.in.off[i]);
        v[i].l2 = l2_full ? (*l2_full)[i] : (uint32_t)(g.two.off[i + 1] - g.two.off[i]);
        bool ok = v[i].oc >= tho && v[i].ic >= thi && v[i].l2 >= (uint32_t)(min_size - 1);
        v[i].label = ok ? 0 : -1;
        (ok ? alive : dropped).push_back(i);
    }
    // full recount rounds while fewer than half of the previous candidates survive
    int32_t prev = n;
    while ((int32_t)alive.size() < prev / 2) {
        prev = (int32_t)alive.size();
        for (int32_t x : alive) { v[x].oc = 0; v[x].ic = 0; }
        for (int32_t x : alive) {
            for (uint64_t k = g.out.off[x]; k < g.out.off[x + 1]; ++k) if (v[g.out.nbr[k]].label == 0) v[g.out.nbr[k]].ic++;
            for (uint64_t k = g.in.off[x]; k < g.in.off[x + 1]; ++k) if (v[g.in.nbr[k]].label == 0) v[g.in.nbr[k]].oc++;
        }
        std::vector<int32_t> next;
        dropped.clear();
        for (int32_t x : alive) {
            if (v[x].oc >= tho && v[x].ic >= thi) next.push_back(x);
            else { v[x].label = -1; dropped.push_back(x); }
        }
        alive.swap(next);
        note("peel round, survivors", (long long)alive.size());
    }
    note("initial filter and peeling done, survivors", (long long)alive.size());
    // cascade the last dropped set through the survivors
    int32_t k = (int32_t)alive.size();
    for (size_t h = 0; h < dropped.size(); ++h) {
        int32_t x = dropped[h];
        for (uint64_t e = g.out.off[x]; e < g.out.off[x + 1]; ++e) {
            L0V& y = v[g.out.nbr[e]];
            if (y.label == 0 && --y.ic < thi) { y.label = -1; --k; dropped.push_back((int32_t)y.id); }
        }
        for (uint64_t e = g.in.off[x]; e < g.in.off[x + 1]; ++e) {
            L0V& y = v[g.in.nbr[e]];
            if (y.label == 0 && --y.oc < tho) { y.label = -1; --k; dropped.push_back((int32_t)y.id); }
        }
    }
    // cover vertex: first vertex with the largest min(out, in) among candidates
    int32_t cover = 0, best = 0;
    for (int32_t i = 0; i < n; ++i)
        if (v[i].label == 0 && (int32_t)std::min(v[i].oc, v[i].ic) > best) { best = (int32_t)std::min(v[i].oc, v[i].ic); cover = i; }
    v[cover].label = 3;
    {
        std::vector<uint8_t> out_of_cover((size_t)n, 0);
        for (uint64_t e = g.out.off[cover]; e < g.out.off[cover + 1]; ++e)
            if (v[g.out.nbr[e]].label == 0) out_of_cover[g.out.nbr[e]] = 1;
        for (uint64_t e = g.in.off[cover]; e < g.in.off[cover + 1]; ++e)
            if (out_of_cover[g.in.nbr[e]]) v[g.in.nbr[e]].label = 2;
    }
    // enumeration order; the survivors (labels 2, 0, 3) come first
    std::vector<int32_t> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return l0_before(v[a], v[b]); });
    std::vector<int32_t> rank((size_t)n, -1);
    for (int32_t r = 0; r < k; ++r) rank[order[r]] = r;
    res.orig_of.assign(order.begin(), order.begin() + k);
    res.ncov = 0;
    while (res.ncov < k && v[order[res.ncov]].label == 2) ++res.ncov;
    // condensed CSR: each list restricted to survivors, relabelled, re-sorted
    Graph c;
    c.n = k;
    auto restrict_csr = [&](const Csr& src, Csr& dst, std::vector<uint32_t>* len_out) {
        dst.off.assign((size_t)k + 1, 0);
        std::vector<std::vector<int32_t>> rows((size_t)k);
#pragma omp parallel for schedule(dynamic, 256) if (k > PAR_MIN)
        for (int32_t r = 0; r < k; ++r) {
            int32_t x = order[r];
            std::vector<int32_t>& row = rows[r];
            for (uint64_t e = src.off[x]; e < src.off[x + 1]; ++e) {
                int32_t y = rank[src.nbr[e]];
                if (y >= 0) row.push_back(y);
            }
            std::sort(row.begin(), row.end());
        }
        for (int32_t r = 0; r < k; ++r) dst.off[r + 1] = dst.off[r] + rows[r].size();
        dst.nbr.resize(dst.off[k]);
#pragma omp parallel for schedule(static) if (k > PAR_MIN)
        for (int32_t r = 0; r < k; ++r) std::copy(rows[r].begin(), rows[r].end(), dst.nbr.begin() + dst.off[r]);
        if (len_out) {
            len_out->resize((size_t)k);
            for (int32_t r = 0; r < k; ++r) (*len_out)[r] = (uint32_t)rows[r].size();
        }
    };
    note("cascade, cover and order done, condensed vertices", k);
    std::vector<uint32_t> l2c, odc, idc;
    restrict_csr(g.two, c.two, &l2c);
    note("condensed two-hop entries", (long long)c.two.nbr.size());
    restrict_csr(g.out, c.out, &odc);
    restrict_csr(g.in, c.in, &idc);
    note("condensed arcs", (long long)c.out.nbr.size());
    c.m = (int64_t)c.out.nbr.size();
    res.task.resize((size_t)k);
    for (int32_t r = 0; r < k; ++r) {
        const L0V& x = v[order[r]];
        if (odc[r] != x.oc || idc[r] != x.ic)
            std::fprintf(stderr, "dqc: inconsistent candidate degree at condensed vertex %d\n", r);
        TaskVertex& t = res.task[r];
        t.id = (uint32_t)r;
        t.out_mem = 0; t.in_mem = 0;
        t.out_can = x.oc; t.in_can = x.ic;
        t.lvl2adj = l2c[r];
    }
    g = std::move(c);
}

}  // namespace dqc
