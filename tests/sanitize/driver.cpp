// Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only, built with
// g++ by tests/test_sanitize.py; never shipped). Exercises the product's host
// preprocessing (text loader, OpenMP two-hop builder, degree tables, level 0 + condense,
// binary graph cache, maximal-set filter, generator) and the oracle's whole CPU path on
// one input, and prints a digest of each result so the test can compare them.
//
//   driver host   <graph> <out_gamma> <in_gamma> <min> <threads> <cache_path>
//   driver oracle <graph> <out_gamma> <in_gamma> <min> <threads> <max_out>
//   driver rmnonmax <raw_in> <max_out>
//   driver gen <config> <graph_index> <path>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../cudqc_amd/csrc/graph.hpp"

extern "C" {
int32_t dqc_generate(int32_t config, int32_t graph_index, const char* path);
struct OrcStats {
    int64_t v[11];
    double t[5];
};
int64_t orc_run(const char* graph_path, double gout, double gin, int min_size, int threads, int host_addfail_check,
                const char* raw_path, const char* max_path, double time_cap_s, OrcStats* st);
}

namespace {

uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

uint64_t digest(const dqc::Csr& c) {
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, c.off.data(), c.off.size() * sizeof(uint64_t));
    return fnv(h, c.nbr.data(), c.nbr.size() * sizeof(int32_t));
}

int host(int argc, char** argv) {
    if (argc < 8) return 2;
    const double go = std::atof(argv[3]), gi = std::atof(argv[4]);
    const int mn = std::atoi(argv[5]), threads = std::atoi(argv[6]);
    dqc::Graph g;
    std::string err;
    if (int rc = dqc::load_text(argv[2], g, err)) {
        std::fprintf(stderr, "load: %s\n", err.c_str());
        return 10 + rc;
    }
    dqc::build_twohop(g, threads);
    std::printf("graph n=%d m=%lld out=%016llx in=%016llx two=%016llx\n", g.n, (long long)g.m,
                (unsigned long long)digest(g.out), (unsigned long long)digest(g.in), (unsigned long long)digest(g.two));
    // binary cache round trip of the uncondensed graph
    if (!dqc::save_binary(argv[7], g, err)) {
        std::fprintf(stderr, "save: %s\n", err.c_str());
        return 20;
    }
    dqc::Graph r;
    bool has_two = false;
    if (!dqc::load_binary(argv[7], r, err, &has_two)) {
        std::fprintf(stderr, "load binary: %s\n", err.c_str());
        return 21;
    }
    if (r.n != g.n || r.m != g.m || digest(r.out) != digest(g.out) || digest(r.in) != digest(g.in) ||
        digest(r.two) != digest(g.two)) {
        std::fprintf(stderr, "binary round trip differs\n");
        return 22;
    }
    const int len = 2 * (std::max(g.n, mn) + 2) + 2;
    auto mdo = dqc::min_degree_table(go, len), mdi = dqc::min_degree_table(gi, len);
    dqc::Level0 l0;
    dqc::level0_condense(g, mdo.data(), mdi.data(), mn, l0);
    uint64_t h = fnv(1469598103934665603ull, l0.task.data(), l0.task.size() * sizeof(dqc::TaskVertex));
    h = fnv(h, l0.orig_of.data(), l0.orig_of.size() * sizeof(int32_t));
    std::printf("level0 k=%zu ncov=%d task=%016llx condensed two=%016llx\n", l0.task.size(), l0.ncov,
                (unsigned long long)h, (unsigned long long)digest(g.two));
    // the search path: two-hop lists built for the degree candidates only
    dqc::Level0 l1;
    dqc::level0_from_loaded(r, mdo.data(), mdi.data(), mn, threads, l1);
    uint64_t h1 = fnv(1469598103934665603ull, l1.task.data(), l1.task.size() * sizeof(dqc::TaskVertex));
    h1 = fnv(h1, l1.orig_of.data(), l1.orig_of.size() * sizeof(int32_t));
    if (h1 != h || digest(r.two) != digest(g.two) || l1.ncov != l0.ncov) {
        std::fprintf(stderr, "restricted two-hop path differs\n");
        return 23;
    }
    return 0;
}

int oracle(int argc, char** argv) {
    if (argc < 8) return 2;
    OrcStats st{};
    const int64_t r = orc_run(argv[2], std::atof(argv[3]), std::atof(argv[4]), std::atoi(argv[5]), std::atoi(argv[6]), 1,
                              nullptr, argv[7], 0.0, &st);
    std::printf("oracle maximal=%lld tasks=%lld branches=%lld\n", (long long)r, (long long)st.v[5], (long long)st.v[6]);
    return r < 0 ? 30 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "host") return host(argc, argv);
    if (mode == "oracle") return oracle(argc, argv);
    if (mode == "rmnonmax" && argc >= 4) {
        const int64_t r = dqc::remove_non_max(argv[2], argv[3]);
        std::printf("rmnonmax %lld\n", (long long)r);
        return r < 0 ? 40 : 0;
    }
    if (mode == "gen" && argc >= 5) {
        const int32_t n = dqc_generate(std::atoi(argv[2]), std::atoi(argv[3]), argv[4]);
        std::printf("gen n=%d\n", n);
        return n < 0 ? 50 : 0;
    }
    return 2;
}
