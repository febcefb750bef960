// Seeded synthetic directed graphs for the five benchmark workloads.
//
// Output is the adjacency-list text format the enumerator reads
// (line i = out-neighbours of vertex i, one line per vertex, empty lines kept),
// the same format CPU_Graph::CPU_Graph parses (reference src/common.cpp:15-57).
//
// RNG contract (fixed so every file is reproducible bit-for-bit):
//   * one std::mt19937_64 stream per graph, seeded with the workload number
//     (1..5); the batching workload seeds graph g with 3000 + g;
//   * uniform double  = (x >> 11) * 2^-53;
//   * bounded integer = Lemire multiply-shift ((x * range) >> 64, 128-bit);
//   * no std::*_distribution (their output is implementation-defined).
// Draw order per graph: background arcs, then planted-set sizes, then a
// Fisher-Yates permutation of all vertices, then the arcs of each planted set
// (its members are consecutive chunks of the permutation, so sets are
// disjoint), ordered pairs (u, v), u != v, in chunk order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

namespace {

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    uint64_t next() { return g(); }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    uint64_t bounded(uint64_t range) {
        return (uint64_t)(((unsigned __int128)next() * range) >> 64);
    }
};

using Arc = uint64_t;  // (u << 32) | v
inline Arc mk(uint32_t u, uint32_t v) { return ((uint64_t)u << 32) | v; }

void er_arcs(Rng& r, uint32_t n, double p, std::vector<Arc>& arcs) {
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t v = 0; v < n; ++v) {
            if (u == v) continue;
            if (r.uniform() < p) arcs.push_back(mk(u, v));
        }
}

// R-MAT: each arc descends `scale` quadrant choices; self loops are rejected and
// duplicates collapse. Arcs are drawn in batches until `target` unique arcs exist;
// the last batch is trimmed in draw order so the result does not depend on batch size.
void rmat_arcs(Rng& r, int scale, uint64_t target, double a, double b, double c,
               std::vector<Arc>& arcs) {
    std::vector<Arc> uniq;  // sorted unique arcs so far
    uniq.reserve(target);
    std::vector<Arc> batch;
    while (uniq.size() < target) {
        uint64_t want = target - uniq.size();
        batch.clear();
        batch.reserve(want);
        for (uint64_t k = 0; k < want; ++k) {
            uint32_t u = 0, v = 0;
            for (int l = 0; l < scale; ++l) {
                double x = r.uniform();
                uint32_t bu, bv;
                if (x < a) { bu = 0; bv = 0; }
                else if (x < a + b) { bu = 0; bv = 1; }
                else if (x < a + b + c) { bu = 1; bv = 0; }
                else { bu = 1; bv = 1; }
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            if (u == v) continue;
            batch.push_back(mk(u, v));
        }
        // keep, in draw order, the arcs not seen before (and not repeated in the batch)
        std::vector<Arc> fresh;
        fresh.reserve(batch.size());
        {
            std::vector<Arc> sorted_batch(batch);
            std::sort(sorted_batch.begin(), sorted_batch.end());
            // first occurrence in draw order: mark via a sorted (arc, index) pass
            std::vector<std::pair<Arc, uint32_t>> idx(batch.size());
            for (size_t k = 0; k < batch.size(); ++k) idx[k] = {batch[k], (uint32_t)k};
            std::sort(idx.begin(), idx.end());
            std::vector<uint8_t> keep(batch.size(), 0);
            for (size_t k = 0; k < idx.size(); ++k) {
                if (k > 0 && idx[k].first == idx[k - 1].first) continue;
                if (std::binary_search(uniq.begin(), uniq.end(), idx[k].first)) continue;
                keep[idx[k].second] = 1;
            }
            for (size_t k = 0; k < batch.size(); ++k)
                if (keep[k]) fresh.push_back(batch[k]);
        }
        if (fresh.size() > want) fresh.resize(want);
        std::sort(fresh.begin(), fresh.end());
        std::vector<Arc> merged;
        merged.reserve(uniq.size() + fresh.size());
        std::merge(uniq.begin(), uniq.end(), fresh.begin(), fresh.end(),
                   std::back_inserter(merged));
        uniq.swap(merged);
    }
    arcs.insert(arcs.end(), uniq.begin(), uniq.end());
}

void plant(Rng& r, uint32_t n, int count, int smin, int smax, double p,
           std::vector<Arc>& arcs) {
    std::vector<int> sizes(count);
    for (int s = 0; s < count; ++s) sizes[s] = smin + (int)r.bounded((uint64_t)(smax - smin + 1));
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    for (uint32_t i = n - 1; i > 0; --i) {
        uint32_t j = (uint32_t)r.bounded((uint64_t)i + 1);
        std::swap(perm[i], perm[j]);
    }
    size_t pos = 0;
    for (int s = 0; s < count; ++s) {
        if (pos + sizes[s] > n) break;  // cannot happen for the shipped workloads
        for (int i = 0; i < sizes[s]; ++i)
            for (int j = 0; j < sizes[s]; ++j) {
                if (i == j) continue;
                if (r.uniform() < p) arcs.push_back(mk(perm[pos + i], perm[pos + j]));
            }
        pos += sizes[s];
    }
}

int write_adjacency(const char* path, uint32_t n, std::vector<Arc>& arcs) {
    std::sort(arcs.begin(), arcs.end());
    arcs.erase(std::unique(arcs.begin(), arcs.end()), arcs.end());
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    std::vector<char> buf;
    buf.reserve(1 << 20);
    size_t k = 0;
    char tmp[16];
    for (uint32_t u = 0; u < n; ++u) {
        bool first = true;
        while (k < arcs.size() && (uint32_t)(arcs[k] >> 32) == u) {
            uint32_t v = (uint32_t)arcs[k];
            int len = std::snprintf(tmp, sizeof tmp, first ? "%u" : " %u", v);
            buf.insert(buf.end(), tmp, tmp + len);
            first = false;
            ++k;
        }
        buf.push_back('\n');
        if (buf.size() > (1u << 20)) {
            std::fwrite(buf.data(), 1, buf.size(), f);
            buf.clear();
        }
    }
    std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return 0;
}

}  // namespace

extern "C" {

// Writes workload `config` (1..5) to `path`. For workload 3, `graph_index` picks
// graph g in [0, 256); it is ignored otherwise. Returns the vertex count, or a
// negative value on error (-1 I/O, -2 bad argument).
int dqc_generate(int config, int graph_index, const char* path) {
    std::vector<Arc> arcs;
    uint32_t n = 0;
    switch (config) {
    case 1: {
        Rng r(1);
        n = 1000;
        er_arcs(r, n, 0.01, arcs);
        plant(r, n, 5, 12, 12, 0.95, arcs);
        break;
    }
    case 2: {
        Rng r(2);
        n = 1u << 16;
        rmat_arcs(r, 16, 1ull << 20, 0.57, 0.19, 0.19, arcs);
        plant(r, n, 50, 15, 30, 0.95, arcs);
        break;
    }
    case 3: {
        if (graph_index < 0 || graph_index >= 256) return -2;
        Rng r(3000 + (uint64_t)graph_index);
        double lo = std::log(200.0), hi = std::log(5000.0);
        n = (uint32_t)std::llround(std::exp(lo + (hi - lo) * r.uniform()));
        er_arcs(r, n, 8.0 / (double)(n - 1), arcs);
        int count = 1 + (int)r.bounded(10);
        plant(r, n, count, 10, 25, 0.92, arcs);
        break;
    }
    case 4: {
        Rng r(4);
        n = 1u << 20;
        rmat_arcs(r, 20, 1ull << 24, 0.57, 0.19, 0.19, arcs);
        plant(r, n, 500, 20, 40, 0.95, arcs);
        break;
    }
    case 5: {
        Rng r(5);
        n = 1u << 22;
        rmat_arcs(r, 22, 1ull << 26, 0.6, 0.15, 0.15, arcs);
        plant(r, n, 2000, 20, 60, 0.97, arcs);
        break;
    }
    default:
        return -2;
    }
    if (write_adjacency(path, n, arcs) != 0) return -1;
    return (int)n;
}

}  // extern "C"
