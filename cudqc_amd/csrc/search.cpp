// Drop-in enumeration path (reference program.cu:93-170 + src/host_functions.cu:9-302):
// graph (text file, binary cache file or caller's in-memory CSR) -> two-hop lists ->
// tables -> level 0 + condense (host) -> GPU expansion of every task through the C-ABI
// -> raw clique file -> RemoveNonMax -> maximal file / caller's sink.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/dqc.h"
#include "graph.hpp"

namespace {

// Table length covering every index the expansion reads: k + n + 2 for sizes up to the
// input vertex count, with room for the min-size clamp (the reference sizes its tables
// n + 1 and reads beyond them only in degenerate cases; here the formula continues).
int table_len(int n, int min_size) { return 2 * (std::max(n, min_size) + 2) + 2; }

// The caller's minimum-degree table md[0 .. md_len) verbatim (the reference's h_search
// takes these arrays from its caller), continued beyond by the closed form of
// h_calculate_minimum_degrees (src/host_functions.cu:9-15) with the same gamma.
std::vector<int32_t> extend_table(const int32_t* md, int32_t md_len, double gamma, int len) {
    std::vector<int32_t> t = dqc::min_degree_table(gamma, len);
    if (md)
        for (int k = 0; k < std::min(md_len, len); ++k) t[k] = md[k];
    return t;
}

// Writes sets in the reference's raw / RemoveNonMax line format: "<len> v1 ... vk \n"
// (src/host_functions.cu:1222-1237, src/Quick_rmnonmax.cpp:66-69).
bool write_sets(const char* path, const std::vector<std::vector<int32_t>>& sets) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    std::string line;
    char tmp[16];
    for (auto& s : sets) {
        line.clear();
        int l = std::snprintf(tmp, sizeof tmp, "%d ", (int)s.size());
        line.append(tmp, l);
        for (int32_t x : s) { l = std::snprintf(tmp, sizeof tmp, "%d ", x); line.append(tmp, l); }
        line.push_back('\n');
        std::fwrite(line.data(), 1, line.size(), f);
    }
    return std::fclose(f) == 0;
}

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

bool gamma_ok(double g) { return g >= 0.5 && g <= 1.0; }

// Validates a caller's CSR against the invariants the reference's loader guarantees
// (src/common.cpp:60-78): offsets from 0, non-decreasing; ids in [0, n); each list
// strictly ascending (sorted, no duplicates) with no self loop; the in-lists the exact
// transpose of the out-lists. Copies it into g.
bool adopt_csr(int32_t n, const uint64_t* out_off, const int32_t* out_nbr, const uint64_t* in_off,
               const int32_t* in_nbr, const uint64_t* two_off, const int32_t* two_nbr, dqc::Graph& g,
               std::string& err) {
    auto copy = [&](const uint64_t* off, const int32_t* nbr, dqc::Csr& c, const char* what) {
        if (off[0] != 0) { err = std::string(what) + ": offsets must start at 0"; return false; }
        for (int32_t v = 0; v < n; ++v)
            if (off[v + 1] < off[v]) { err = std::string(what) + ": offsets decrease"; return false; }
        if (off[n] > 0 && !nbr) { err = std::string(what) + ": missing id array"; return false; }
        c.off.assign(off, off + (size_t)n + 1);
        if (off[n] > 0) c.nbr.assign(nbr, nbr + off[n]);
        else c.nbr.clear();
        for (int32_t v = 0; v < n; ++v)
            for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
                const int32_t x = nbr[e];
                if (x < 0 || x >= n || x == v || (e > off[v] && nbr[e - 1] >= x)) {
                    err = std::string(what) + ": lists must hold ascending distinct ids in [0, n), no self loop";
                    return false;
                }
            }
        return true;
    };
    g = dqc::Graph();
    g.n = n;
    if (!copy(out_off, out_nbr, g.out, "out") || !copy(in_off, in_nbr, g.in, "in")) return false;
    g.m = (int64_t)g.out.off[n];
    if (g.in.off[n] != g.out.off[n]) { err = "in-lists hold a different arc count than out-lists"; return false; }
    {  // in == transpose(out): scanning u ascending fills each in-list in ascending order
        std::vector<uint64_t> cur(g.in.off.begin(), g.in.off.end() - 1);
        for (int32_t u = 0; u < n; ++u)
            for (uint64_t e = g.out.off[u]; e < g.out.off[u + 1]; ++e) {
                const int32_t v = g.out.nbr[e];
                if (cur[v] >= g.in.off[v + 1] || g.in.nbr[cur[v]] != u) {
                    err = "in-lists are not the transpose of the out-lists";
                    return false;
                }
                ++cur[v];
            }
    }
    if (two_off) {
        if (!copy(two_off, two_nbr, g.two, "two-hop")) return false;
    } else {
        g.two.off.assign((size_t)n + 1, 0);  // built by level0_impl, for the candidates only
        g.two.nbr.clear();
    }
    return true;
}

// Text or binary graph for the search path; has_two: the file held two-hop lists.
int load_for_search(const char* path, dqc::Graph& g, std::string& err, bool& has_two) {
    has_two = false;
    if (dqc::is_binary_graph(path)) return dqc::load_binary(path, g, err, &has_two) ? 0 : 2;
    return dqc::load_text(path, g, err);
}

// Tables and level 0 + condense on g, which becomes the condensed graph (as h_search's
// hg does through h_condense_graph).
struct Prepared {
    int len = 0;
    std::vector<int32_t> mdo, mdi, ubo, ubi;
    dqc::Level0 l0;
};
// has_two: g holds complete two-hop lists; otherwise they are built here for the
// vertices level 0 can keep (dqc::level0_from_loaded), with identical results.
void level0_impl(dqc::Graph& g, bool has_two, const int32_t* md_out, const int32_t* md_in, int32_t md_len, double go,
                 double gi, int32_t min_size, Prepared& p, int threads) {
    p.len = table_len(g.n, min_size);
    p.mdo = extend_table(md_out, md_len, go, p.len);
    p.mdi = extend_table(md_in, md_len, gi, p.len);
    p.ubo = dqc::bound_div_table(go, p.len);
    p.ubi = dqc::bound_div_table(gi, p.len);
    if (has_two) dqc::level0_condense(g, p.mdo.data(), p.mdi.data(), min_size, p.l0);
    else dqc::level0_from_loaded(g, p.mdo.data(), p.mdi.data(), min_size, threads, p.l0);
}

// Upload of the condensed graph, the tables and the level-0 task into an open handle.
dqc_status upload_impl(dqc_handle* h, const dqc::Graph& g, const Prepared& p, int32_t min_size) {
    if (p.l0.task.empty()) return DQC_OK;
    dqc_status s = dqc_upload_graph(h, g.n, g.out.off.data(), g.out.nbr.data(), g.in.off.data(), g.in.nbr.data(),
                                    g.two.off.data(), g.two.nbr.data());
    if (s != DQC_OK) return s;
    s = dqc_set_tables(h, p.mdo.data(), p.mdi.data(), p.ubo.data(), p.ubi.data(), p.len, min_size);
    if (s != DQC_OK) return s;
    return dqc_push_root(h, (const uint32_t*)p.l0.task.data(), (int32_t)p.l0.task.size(), p.l0.ncov);
}

// The whole search over a loaded graph (with two-hop lists): level 0 on the host, then
// (when anything survives) one handle and the expansion to completion, then the raw
// file, the maximal filter, the maximal file and the sink.
dqc_status search_impl(dqc::Graph& g, bool has_two, double go, double gi, int32_t min_size, const int32_t* md_out,
                       const int32_t* md_in, int32_t md_len, const char* raw_path, const char* max_path,
                       dqc_set_sink sink, void* sink_ctx, const dqc_config* cfg, int64_t* max_cliques,
                       dqc_stats* st, int32_t* orig_of_out, std::chrono::steady_clock::time_point t_start) {
    Prepared p;
    level0_impl(g, has_two, md_out, md_in, md_len, go, gi, min_size, p, omp_get_max_threads());
    dqc_stats stats{};
    std::vector<std::vector<int32_t>> sets;
    if (!p.l0.task.empty()) {
        dqc_handle* h = nullptr;
        dqc_status s = dqc_open(cfg, &h);
        auto fail = [&](dqc_status code) {
            std::fprintf(stderr, "dqc: %s\n", h ? dqc_last_error(h) : "open failed");
            dqc_close(h);
            return code;
        };
        if (s != DQC_OK) return fail(s);
        s = upload_impl(h, g, p, min_size);
        if (s != DQC_OK) return fail(s);
        const double prep_ms = ms_since(t_start);
        s = dqc_run(h, DQC_SCHED_DEFAULT);
        if (s != DQC_OK) return fail(s);
        int64_t count = 0, total = 0;
        s = dqc_fetch_cliques(h, &count, &total, nullptr, nullptr);
        if (s != DQC_OK) return fail(s);
        std::vector<int32_t> sizes((size_t)count), ids((size_t)total);
        s = dqc_fetch_cliques(h, &count, &total, sizes.data(), ids.data());
        if (s != DQC_OK) return fail(s);
        dqc_get_stats(h, &stats);
        dqc_close(h);
        stats.preprocess_ms = prep_ms;
        sets.reserve((size_t)count);
        size_t q = 0;
        for (int64_t i = 0; i < count; ++i) {
            sets.emplace_back(ids.begin() + q, ids.begin() + q + sizes[i]);
            q += (size_t)sizes[i];
        }
    } else {
        stats.preprocess_ms = ms_since(t_start);
    }
    stats.n_condensed = (int32_t)p.l0.task.size();
    if (raw_path && !write_sets(raw_path, sets)) return DQC_E_IO;
    auto tr = std::chrono::steady_clock::now();
    auto mx = dqc::maximal_sets(sets);
    stats.rmnonmax_ms = ms_since(tr);
    if (max_path && !mx.empty() && !write_sets(max_path, mx)) return DQC_E_IO;
    if (sink)
        for (auto& m : mx) sink(sink_ctx, m.data(), (int32_t)m.size());
    if (orig_of_out && !p.l0.orig_of.empty())
        std::memcpy(orig_of_out, p.l0.orig_of.data(), sizeof(int32_t) * p.l0.orig_of.size());
    stats.maximal_cliques = (int64_t)mx.size();
    if (max_cliques) *max_cliques = (int64_t)mx.size();
    if (st) *st = stats;
    return DQC_OK;
}

}  // namespace

extern "C" {

void dqc_free(void* p) { std::free(p); }

dqc_status dqc_min_degrees(double gamma, int32_t n, int32_t* md) {
    if (!md || n < 0 || !(gamma >= 0.0 && gamma <= 1.0)) return DQC_E_ARG;
    auto t = dqc::min_degree_table(gamma, n + 1);
    std::memcpy(md, t.data(), sizeof(int32_t) * ((size_t)n + 1));
    return DQC_OK;
}

dqc_status dqc_load_graph(const char* graph_path, int32_t with_twohop, int32_t* n, int64_t* m, uint64_t** out_off,
                          int32_t** out_nbr, uint64_t** in_off, int32_t** in_nbr, uint64_t** two_off,
                          int32_t** two_nbr) {
    if (!graph_path || !n || !m) return DQC_E_ARG;
    dqc::Graph g;
    std::string err;
    int rc = 0;
    if (dqc::is_binary_graph(graph_path)) {
        bool two = false;
        rc = dqc::load_binary(graph_path, g, err, &two) ? 0 : 2;
        if (!rc && with_twohop && !two) dqc::build_twohop(g, omp_get_max_threads());
    } else {
        rc = dqc::load_text(graph_path, g, err);
        if (!rc && with_twohop) dqc::build_twohop(g, omp_get_max_threads());
    }
    if (rc) return rc == 1 ? DQC_E_IO : DQC_E_ARG;
    *n = g.n;
    *m = g.m;
    auto cp = [](const auto& v, auto** dst) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        if (!dst) return;
        *dst = (std::remove_pointer_t<decltype(dst)>)std::malloc(sizeof(T) * (v.size() + 1));
        if (!v.empty()) std::memcpy(*dst, v.data(), sizeof(T) * v.size());
    };
    cp(g.out.off, out_off); cp(g.out.nbr, out_nbr);
    cp(g.in.off, in_off); cp(g.in.nbr, in_nbr);
    cp(g.two.off, two_off); cp(g.two.nbr, two_nbr);
    return DQC_OK;
}

dqc_status dqc_save_graph(const char* graph_path, const char* bin_path, int32_t with_twohop) {
    if (!graph_path || !bin_path) return DQC_E_ARG;
    dqc::Graph g;
    std::string err;
    int rc = with_twohop ? dqc::load_graph_file(graph_path, g, err, omp_get_max_threads())
                         : (dqc::is_binary_graph(graph_path) ? (dqc::load_binary(graph_path, g, err) ? 0 : 2)
                                                             : dqc::load_text(graph_path, g, err));
    if (rc) {
        std::fprintf(stderr, "dqc: %s\n", err.c_str());
        return rc == 1 ? DQC_E_IO : DQC_E_ARG;
    }
    if (!with_twohop) { g.two.off.assign((size_t)g.n + 1, 0); g.two.nbr.clear(); }
    if (!dqc::save_binary(bin_path, g, err)) {
        std::fprintf(stderr, "dqc: %s\n", err.c_str());
        return DQC_E_IO;
    }
    return DQC_OK;
}

dqc_status dqc_level0(const char* graph_path, double out_gamma, double in_gamma, int32_t min_size,
                      int32_t* n_condensed, int32_t* ncov, uint32_t** records, int32_t** orig_of) {
    if (!graph_path || !n_condensed || !ncov || min_size <= 1) return DQC_E_ARG;
    dqc::Graph g;
    std::string err;
    bool has_two = false;
    int rc = load_for_search(graph_path, g, err, has_two);
    if (rc) return rc == 1 ? DQC_E_IO : DQC_E_ARG;
    Prepared p;
    level0_impl(g, has_two, nullptr, nullptr, 0, out_gamma, in_gamma, min_size, p, omp_get_max_threads());
    const dqc::Level0& l0 = p.l0;
    *n_condensed = (int32_t)l0.task.size();
    *ncov = l0.ncov;
    if (records) {
        *records = (uint32_t*)std::malloc(sizeof(dqc::TaskVertex) * (l0.task.size() + 1));
        if (!l0.task.empty()) std::memcpy(*records, l0.task.data(), sizeof(dqc::TaskVertex) * l0.task.size());
    }
    if (orig_of) {
        *orig_of = (int32_t*)std::malloc(sizeof(int32_t) * (l0.orig_of.size() + 1));
        if (!l0.orig_of.empty()) std::memcpy(*orig_of, l0.orig_of.data(), sizeof(int32_t) * l0.orig_of.size());
    }
    return DQC_OK;
}

dqc_status dqc_prepare_file(dqc_handle* h, const char* graph_path, double out_gamma, double in_gamma,
                            int32_t min_size, int32_t* n_condensed) {
    if (!h || !graph_path || !gamma_ok(out_gamma) || !gamma_ok(in_gamma) || min_size <= 1) return DQC_E_ARG;
    dqc::Graph g;
    std::string err;
    bool has_two = false;
    int rc = load_for_search(graph_path, g, err, has_two);
    if (rc) return rc == 1 ? DQC_E_IO : DQC_E_ARG;
    Prepared p;
    level0_impl(g, has_two, nullptr, nullptr, 0, out_gamma, in_gamma, min_size, p, omp_get_max_threads());
    if (n_condensed) *n_condensed = (int32_t)p.l0.task.size();
    return upload_impl(h, g, p, min_size);
}

dqc_status dqc_prepare_graph(dqc_handle* h, int32_t n, const uint64_t* out_off, const int32_t* out_nbr,
                             const uint64_t* in_off, const int32_t* in_nbr, const uint64_t* two_off,
                             const int32_t* two_nbr, const int32_t* md_out, const int32_t* md_in, int32_t md_len,
                             double out_gamma, double in_gamma, int32_t min_size, int32_t* n_condensed,
                             int32_t* orig_of) {
    if (!h || n < 0 || !out_off || !in_off || (two_nbr && !two_off) || !gamma_ok(out_gamma) || !gamma_ok(in_gamma) ||
        min_size <= 1 || (md_len > 0 && (!md_out || !md_in)) || md_len < 0)
        return DQC_E_ARG;
    dqc::Graph g;
    std::string err;
    if (!adopt_csr(n, out_off, out_nbr, in_off, in_nbr, two_off, two_nbr, g, err)) {
        std::fprintf(stderr, "dqc: %s\n", err.c_str());
        return DQC_E_ARG;
    }
    Prepared p;
    level0_impl(g, two_off != nullptr, md_out, md_in, md_len, out_gamma, in_gamma, min_size, p, omp_get_max_threads());
    if (n_condensed) *n_condensed = (int32_t)p.l0.task.size();
    if (orig_of && !p.l0.orig_of.empty())
        std::memcpy(orig_of, p.l0.orig_of.data(), sizeof(int32_t) * p.l0.orig_of.size());
    return upload_impl(h, g, p, min_size);
}

dqc_status dqc_prepare_batch(dqc_handle* h, const char* const* graph_paths, int32_t ngraphs, double out_gamma,
                             double in_gamma, int32_t min_size, int32_t* n_condensed) {
    if (!h || !graph_paths || ngraphs < 1 || ngraphs > 65535 || !gamma_ok(out_gamma) || !gamma_ok(in_gamma) ||
        min_size <= 1)
        return DQC_E_ARG;
    std::vector<dqc::Graph> g((size_t)ngraphs);
    std::vector<dqc::Level0> l0((size_t)ngraphs);
    std::vector<int> rcs((size_t)ngraphs, 0), lens((size_t)ngraphs, 0);
    // one graph per thread: load, two-hop lists, level 0 + condense (each graph alone)
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < ngraphs; ++i) {
        std::string err;
        if (!graph_paths[i]) { rcs[i] = 2; continue; }
        bool has_two = false;
        rcs[i] = load_for_search(graph_paths[i], g[i], err, has_two);
        if (rcs[i]) continue;
        Prepared p;
        level0_impl(g[i], has_two, nullptr, nullptr, 0, out_gamma, in_gamma, min_size, p, 1);
        lens[i] = p.len;
        l0[i] = std::move(p.l0);
    }
    int len = 2;
    for (int32_t i = 0; i < ngraphs; ++i) {
        if (rcs[i]) return rcs[i] == 1 ? DQC_E_IO : DQC_E_ARG;
        len = std::max(len, lens[i]);
        if (n_condensed) n_condensed[i] = (int32_t)l0[i].task.size();
    }
    std::vector<int32_t> nv((size_t)ngraphs);
    std::vector<const uint64_t*> oo(ngraphs), io(ngraphs), to(ngraphs);
    std::vector<const int32_t*> on(ngraphs), in(ngraphs), tn(ngraphs);
    for (int32_t i = 0; i < ngraphs; ++i) {
        // a graph with no level-0 survivors keeps no ids (its condensed graph is empty)
        nv[i] = l0[i].task.empty() ? 0 : g[i].n;
        oo[i] = g[i].out.off.data(); on[i] = g[i].out.nbr.data();
        io[i] = g[i].in.off.data(); in[i] = g[i].in.nbr.data();
        to[i] = g[i].two.off.data(); tn[i] = g[i].two.nbr.data();
    }
    dqc_status s = dqc_upload_graph_batch(h, ngraphs, nv.data(), oo.data(), on.data(), io.data(), in.data(), to.data(),
                                          tn.data());
    if (s != DQC_OK) return s;
    auto mdo = dqc::min_degree_table(out_gamma, len), mdi = dqc::min_degree_table(in_gamma, len);
    auto ubo = dqc::bound_div_table(out_gamma, len), ubi = dqc::bound_div_table(in_gamma, len);
    s = dqc_set_tables(h, mdo.data(), mdi.data(), ubo.data(), ubi.data(), len, min_size);
    if (s != DQC_OK) return s;
    for (int32_t i = 0; i < ngraphs; ++i) {
        if (l0[i].task.empty()) continue;
        s = dqc_push_graph_root(h, i, (const uint32_t*)l0[i].task.data(), (int32_t)l0[i].task.size(), l0[i].ncov);
        if (s != DQC_OK) return s;
    }
    return DQC_OK;
}

dqc_status dqc_finish_batch(dqc_handle* h, const char* const* max_paths, int64_t* max_cliques, dqc_stats* st) {
    if (!h) return DQC_E_ARG;
    int64_t count = 0, total = 0;
    dqc_status s = dqc_fetch_batch_cliques(h, &count, &total, nullptr, nullptr, nullptr);
    if (s != DQC_OK) return s;
    std::vector<int32_t> gid((size_t)count), sizes((size_t)count), ids((size_t)total);
    s = dqc_fetch_batch_cliques(h, &count, &total, gid.data(), sizes.data(), ids.data());
    if (s != DQC_OK) return s;
    dqc_stats stats{};
    dqc_get_stats(h, &stats);
    int32_t ngraphs = 0;
    s = dqc_batch_size(h, &ngraphs);
    if (s != DQC_OK) return s;
    std::vector<std::vector<std::vector<int32_t>>> sets((size_t)ngraphs);
    size_t p = 0;
    for (int64_t c = 0; c < count; ++c) {
        if (gid[c] < 0 || gid[c] >= ngraphs) return DQC_E_ARG;
        sets[gid[c]].emplace_back(ids.begin() + p, ids.begin() + p + sizes[c]);
        p += (size_t)sizes[c];
    }
    auto tr = std::chrono::steady_clock::now();
    std::vector<std::vector<std::vector<int32_t>>> mx((size_t)ngraphs);
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < ngraphs; ++i) mx[i] = dqc::maximal_sets(sets[i]);
    stats.rmnonmax_ms = ms_since(tr);
    int64_t all = 0;
    for (int32_t i = 0; i < ngraphs; ++i) {
        if (max_paths && max_paths[i] && !mx[i].empty() && !write_sets(max_paths[i], mx[i])) return DQC_E_IO;
        if (max_cliques) max_cliques[i] = (int64_t)mx[i].size();
        all += (int64_t)mx[i].size();
    }
    stats.maximal_cliques = all;
    if (st) *st = stats;
    return DQC_OK;
}

dqc_status dqc_search_batch(const char* const* graph_paths, int32_t ngraphs, double out_gamma, double in_gamma,
                            int32_t min_size, const char* const* max_paths, const dqc_config* cfg,
                            int64_t* max_cliques, dqc_stats* st) {
    dqc_handle* h = nullptr;
    dqc_status s = dqc_open(cfg, &h);
    auto done = [&](dqc_status code) {
        if (code != DQC_OK && h) std::fprintf(stderr, "dqc: %s\n", dqc_last_error(h));
        dqc_close(h);
        return code;
    };
    if (s != DQC_OK) return done(s);
    std::vector<int32_t> k((size_t)std::max(ngraphs, 1));
    auto tp = std::chrono::steady_clock::now();
    s = dqc_prepare_batch(h, graph_paths, ngraphs, out_gamma, in_gamma, min_size, k.data());
    if (s != DQC_OK) return done(s);
    const double prep_ms = ms_since(tp);
    s = dqc_run(h, DQC_SCHED_DEFAULT);
    if (s != DQC_OK) return done(s);
    dqc_stats stats{};
    s = dqc_finish_batch(h, max_paths, max_cliques, &stats);
    if (s != DQC_OK) return done(s);
    stats.preprocess_ms = prep_ms;
    if (st) *st = stats;
    return done(DQC_OK);
}

int64_t dqc_remove_non_max(const char* in_path, const char* out_path) {
    if (!in_path || !out_path) return -1;
    return dqc::remove_non_max(in_path, out_path);
}

dqc_status dqc_search_file(const char* graph_path, double out_gamma, double in_gamma, int32_t min_size,
                           const char* raw_path, const char* max_path, const dqc_config* cfg,
                           int64_t* max_cliques, dqc_stats* st) {
    // argument checks of reference program.cu:44-57
    if (!graph_path || !gamma_ok(out_gamma) || !gamma_ok(in_gamma) || min_size <= 1) return DQC_E_ARG;
    auto tp = std::chrono::steady_clock::now();
    dqc::Graph g;
    std::string err;
    bool has_two = false;
    int rc = load_for_search(graph_path, g, err, has_two);
    if (rc) {
        std::fprintf(stderr, "dqc: %s\n", err.c_str());
        return rc == 1 ? DQC_E_IO : DQC_E_ARG;
    }
    return search_impl(g, has_two, out_gamma, in_gamma, min_size, nullptr, nullptr, 0, raw_path, max_path, nullptr,
                       nullptr, cfg, max_cliques, st, nullptr, tp);
}

dqc_status dqc_search_graph(int32_t n, const uint64_t* out_off, const int32_t* out_nbr, const uint64_t* in_off,
                            const int32_t* in_nbr, const uint64_t* two_off, const int32_t* two_nbr,
                            const int32_t* md_out, const int32_t* md_in, int32_t md_len, double out_gamma,
                            double in_gamma, int32_t min_size, const char* raw_path, const char* max_path,
                            dqc_set_sink sink, void* sink_ctx, const dqc_config* cfg, int64_t* max_cliques,
                            dqc_stats* st, int32_t* orig_of) {
    if (n < 0 || !out_off || !in_off || (two_nbr && !two_off) || !gamma_ok(out_gamma) || !gamma_ok(in_gamma) ||
        min_size <= 1 || md_len < 0 || (md_len > 0 && (!md_out || !md_in)))
        return DQC_E_ARG;
    auto tp = std::chrono::steady_clock::now();
    dqc::Graph g;
    std::string err;
    if (!adopt_csr(n, out_off, out_nbr, in_off, in_nbr, two_off, two_nbr, g, err)) {
        std::fprintf(stderr, "dqc: %s\n", err.c_str());
        return DQC_E_ARG;
    }
    return search_impl(g, two_off != nullptr, out_gamma, in_gamma, min_size, md_out, md_in, md_len, raw_path, max_path,
                       sink, sink_ctx, cfg, max_cliques, st, orig_of, tp);
}

}  // extern "C"
