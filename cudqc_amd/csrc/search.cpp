// Drop-in enumeration path (reference program.cu:93-170 + src/host_functions.cu:17-302):
// text graph -> two-hop lists -> tables -> level 0 + condense (host) -> GPU expansion of
// every task through the C-ABI -> raw clique file -> RemoveNonMax -> maximal file.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/dqc.h"
#include "graph.hpp"

namespace {

// Table length covering every index the expansion reads: k + n + 2 for sizes up to the
// input vertex count, with room for the min-size clamp (the reference sizes its tables
// n + 1 and reads beyond them only in degenerate cases; here the formula continues).
int table_len(int n, int min_size) { return 2 * (std::max(n, min_size) + 2) + 2; }

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
    std::fclose(f);
    return true;
}

struct Timer {
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    double ms() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count(); }
};

}  // namespace

extern "C" {

void dqc_free(void* p) { std::free(p); }

dqc_status dqc_load_graph(const char* graph_path, int32_t with_twohop, int32_t* n, int64_t* m, uint64_t** out_off,
                          int32_t** out_nbr, uint64_t** in_off, int32_t** in_nbr, uint64_t** two_off,
                          int32_t** two_nbr) {
    if (!graph_path || !n || !m) return DQC_E_ARG;
    dqc::Graph g;
    std::string err;
    int rc = dqc::load_text(graph_path, g, err);
    if (rc) return rc == 1 ? DQC_E_IO : DQC_E_ARG;
    if (with_twohop) dqc::build_twohop(g, omp_get_max_threads());
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

dqc_status dqc_level0(const char* graph_path, double out_gamma, double in_gamma, int32_t min_size,
                      int32_t* n_condensed, int32_t* ncov, uint32_t** records, int32_t** orig_of) {
    if (!graph_path || !n_condensed || !ncov || min_size <= 1) return DQC_E_ARG;
    dqc::Graph g;
    std::string err;
    int rc = dqc::load_text(graph_path, g, err);
    if (rc) return rc == 1 ? DQC_E_IO : DQC_E_ARG;
    dqc::build_twohop(g, omp_get_max_threads());
    const int len = table_len(g.n, min_size);
    auto mdo = dqc::min_degree_table(out_gamma, len), mdi = dqc::min_degree_table(in_gamma, len);
    dqc::Level0 l0;
    dqc::level0_condense(g, mdo.data(), mdi.data(), min_size, l0);
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
    if (!h || !graph_path || !(out_gamma >= 0.5 && out_gamma <= 1.0) || !(in_gamma >= 0.5 && in_gamma <= 1.0) ||
        min_size <= 1)
        return DQC_E_ARG;
    dqc::Graph g;
    std::string err;
    int rc = dqc::load_text(graph_path, g, err);
    if (rc) return rc == 1 ? DQC_E_IO : DQC_E_ARG;
    dqc::build_twohop(g, omp_get_max_threads());
    const int len = table_len(g.n, min_size);
    auto mdo = dqc::min_degree_table(out_gamma, len), mdi = dqc::min_degree_table(in_gamma, len);
    auto ubo = dqc::bound_div_table(out_gamma, len), ubi = dqc::bound_div_table(in_gamma, len);
    dqc::Level0 l0;
    dqc::level0_condense(g, mdo.data(), mdi.data(), min_size, l0);
    if (n_condensed) *n_condensed = (int32_t)l0.task.size();
    if (l0.task.empty()) return DQC_OK;
    dqc_status s = dqc_upload_graph(h, g.n, g.out.off.data(), g.out.nbr.data(), g.in.off.data(), g.in.nbr.data(),
                                    g.two.off.data(), g.two.nbr.data());
    if (s != DQC_OK) return s;
    s = dqc_set_tables(h, mdo.data(), mdi.data(), ubo.data(), ubi.data(), len, min_size);
    if (s != DQC_OK) return s;
    return dqc_push_root(h, (const uint32_t*)l0.task.data(), (int32_t)l0.task.size(), l0.ncov);
}

dqc_status dqc_prepare_batch(dqc_handle* h, const char* const* graph_paths, int32_t ngraphs, double out_gamma,
                             double in_gamma, int32_t min_size, int32_t* n_condensed) {
    if (!h || !graph_paths || ngraphs < 1 || ngraphs > 65535 || !(out_gamma >= 0.5 && out_gamma <= 1.0) ||
        !(in_gamma >= 0.5 && in_gamma <= 1.0) || min_size <= 1)
        return DQC_E_ARG;
    std::vector<dqc::Graph> g((size_t)ngraphs);
    std::vector<dqc::Level0> l0((size_t)ngraphs);
    std::vector<int> rcs((size_t)ngraphs, 0), lens((size_t)ngraphs, 0);
    // one graph per thread: load, two-hop lists, level 0 + condense (each graph alone)
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < ngraphs; ++i) {
        std::string err;
        if (!graph_paths[i]) { rcs[i] = 2; continue; }
        rcs[i] = dqc::load_text(graph_paths[i], g[i], err);
        if (rcs[i]) continue;
        dqc::build_twohop(g[i], 1);
        lens[i] = table_len(g[i].n, min_size);
        auto mdo = dqc::min_degree_table(out_gamma, lens[i]), mdi = dqc::min_degree_table(in_gamma, lens[i]);
        dqc::level0_condense(g[i], mdo.data(), mdi.data(), min_size, l0[i]);
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
    s = dqc_prepare_batch(h, graph_paths, ngraphs, out_gamma, in_gamma, min_size, k.data());
    if (s != DQC_OK) return done(s);
    s = dqc_run(h, DQC_SCHED_DEFAULT);
    if (s != DQC_OK) return done(s);
    int64_t count = 0, total = 0;
    s = dqc_fetch_batch_cliques(h, &count, &total, nullptr, nullptr, nullptr);
    if (s != DQC_OK) return done(s);
    std::vector<int32_t> gid((size_t)count), sizes((size_t)count), ids((size_t)total);
    s = dqc_fetch_batch_cliques(h, &count, &total, gid.data(), sizes.data(), ids.data());
    if (s != DQC_OK) return done(s);
    dqc_stats stats{};
    dqc_get_stats(h, &stats);
    std::vector<std::vector<std::vector<int32_t>>> sets((size_t)ngraphs);
    size_t p = 0;
    for (int64_t c = 0; c < count; ++c) {
        sets[gid[c]].emplace_back(ids.begin() + p, ids.begin() + p + sizes[c]);
        p += (size_t)sizes[c];
    }
    int64_t all = 0;
    for (int32_t i = 0; i < ngraphs; ++i) {
        auto mx = dqc::maximal_sets(sets[i]);
        if (max_paths && max_paths[i] && !mx.empty() && !write_sets(max_paths[i], mx)) return done(DQC_E_IO);
        if (max_cliques) max_cliques[i] = (int64_t)mx.size();
        all += (int64_t)mx.size();
    }
    stats.maximal_cliques = all;
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
    if (!graph_path || !(out_gamma >= 0.5 && out_gamma <= 1.0) || !(in_gamma >= 0.5 && in_gamma <= 1.0) ||
        min_size <= 1)
        return DQC_E_ARG;
    dqc::Graph g;
    std::string err;
    int rc = dqc::load_text(graph_path, g, err);
    if (rc) {
        std::fprintf(stderr, "dqc: %s\n", err.c_str());
        return rc == 1 ? DQC_E_IO : DQC_E_ARG;
    }
    dqc::build_twohop(g, omp_get_max_threads());
    const int len = table_len(g.n, min_size);
    auto mdo = dqc::min_degree_table(out_gamma, len), mdi = dqc::min_degree_table(in_gamma, len);
    auto ubo = dqc::bound_div_table(out_gamma, len), ubi = dqc::bound_div_table(in_gamma, len);
    dqc::Level0 l0;
    dqc::level0_condense(g, mdo.data(), mdi.data(), min_size, l0);

    std::vector<std::vector<int32_t>> sets;
    dqc_stats stats{};
    if (!l0.task.empty()) {
        dqc_handle* h = nullptr;
        dqc_status s = dqc_open(cfg, &h);
        auto fail = [&](dqc_status code) {
            std::fprintf(stderr, "dqc: %s\n", dqc_last_error(h));
            dqc_close(h);
            return code;
        };
        if (s != DQC_OK) return fail(s);
        s = dqc_upload_graph(h, g.n, g.out.off.data(), g.out.nbr.data(), g.in.off.data(), g.in.nbr.data(),
                             g.two.off.data(), g.two.nbr.data());
        if (s != DQC_OK) return fail(s);
        s = dqc_set_tables(h, mdo.data(), mdi.data(), ubo.data(), ubi.data(), len, min_size);
        if (s != DQC_OK) return fail(s);
        s = dqc_push_root(h, (const uint32_t*)l0.task.data(), (int32_t)l0.task.size(), l0.ncov);
        if (s != DQC_OK) return fail(s);
        s = dqc_run(h, 0);
        if (s != DQC_OK) return fail(s);
        int64_t count = 0, total = 0;
        s = dqc_fetch_cliques(h, &count, &total, nullptr, nullptr);
        if (s != DQC_OK) return fail(s);
        std::vector<int32_t> sizes((size_t)count), ids((size_t)total);
        s = dqc_fetch_cliques(h, &count, &total, sizes.data(), ids.data());
        if (s != DQC_OK) return fail(s);
        dqc_get_stats(h, &stats);
        dqc_close(h);
        sets.reserve((size_t)count);
        size_t p = 0;
        for (int64_t i = 0; i < count; ++i) {
            sets.emplace_back(ids.begin() + p, ids.begin() + p + sizes[i]);
            p += (size_t)sizes[i];
        }
    }
    stats.n_condensed = (int32_t)l0.task.size();
    // raw file: "<len> ids... \n" per clique (reference src/host_functions.cu:1222-1237)
    if (raw_path) {
        FILE* f = std::fopen(raw_path, "wb");
        if (!f) return DQC_E_IO;
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
        std::fclose(f);
    }
    auto mx = dqc::maximal_sets(sets);
    if (max_path && !mx.empty()) {
        FILE* f = std::fopen(max_path, "wb");
        if (!f) return DQC_E_IO;
        std::string line;
        char tmp[16];
        for (auto& s : mx) {
            line.clear();
            int l = std::snprintf(tmp, sizeof tmp, "%d ", (int)s.size());
            line.append(tmp, l);
            for (int32_t x : s) { l = std::snprintf(tmp, sizeof tmp, "%d ", x); line.append(tmp, l); }
            line.push_back('\n');
            std::fwrite(line.data(), 1, line.size(), f);
        }
        std::fclose(f);
    }
    stats.maximal_cliques = (int64_t)mx.size();
    if (max_cliques) *max_cliques = (int64_t)mx.size();
    if (st) *st = stats;
    return DQC_OK;
}

}  // extern "C"
