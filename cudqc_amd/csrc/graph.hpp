// Host-side graph structures and preprocessing for the MI355X enumerator.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dqc {

// One vertex of a task record, as stored in HBM (24 bytes, AoS). The label is not
// stored: in a task, positions [0, nm) are members, [nm, nm + ncov) are the covered
// vertices of the level-0 task, the rest are candidates (reference inc/common.hpp:64-73
// keeps an explicit int8 label; every stored task there has labels that follow
// positions this way).
struct TaskVertex {
    uint32_t id;
    uint32_t out_mem, out_can, in_mem, in_can;
    uint32_t lvl2adj;
};
static_assert(sizeof(TaskVertex) == 24, "task record layout");

struct Csr {
    std::vector<uint64_t> off;  // n + 1
    std::vector<int32_t> nbr;
};

// Mirrors CPU_Graph (reference inc/common.hpp:76-94): out / in / two-hop CSR.
struct Graph {
    int32_t n = 0;
    int64_t m = 0;
    Csr out, in, two;
};

// Text adjacency loader (reference src/common.cpp:15-108 semantics). Returns 0, or
// 1 on I/O failure, 2 on a malformed id (outside [0, n)); err gets a message.
int load_text(const std::string& path, Graph& g, std::string& err);

// Directed two-hop lists (reference src/common.cpp:111-509 semantics).
void build_twohop(Graph& g, int threads);

// Two-hop lists for the search path. Level 0 reads a vertex's two-hop list only if the
// vertex passes the degree part of its initial filter (src/host_functions.cu:504-505),
// and after condensing keeps only entries that survive it; so lists are built only for
// vertices with src[v] set, and only their entries x with dst[x] set are stored (both
// flags: out-degree >= md_out[min] and in-degree >= md_in[min], a superset of the
// survivors). full_len[v] gets the length of v's whole list (the filter and the
// enumeration order use it); vertices not built get 0.
void build_twohop_restricted(Graph& g, int threads, const std::vector<uint8_t>& src, const std::vector<uint8_t>& dst,
                             std::vector<uint32_t>& full_len);

// Binary graph file, a cache of what the text loader (and optionally the two-hop
// builder) produces, so that a large graph is parsed and its two-hop lists built once.
// Layout (little-endian): the 8 bytes "DQCGRPH1", int32 n, int32 flags (bit 0: two-hop
// lists present), int64 m, then out offsets (n + 1 uint64), out ids (m int32), in
// offsets, in ids, and, when flagged, two-hop offsets and ids; last, a 64-bit checksum
// of every byte before it. Loading checks the magic, the sizes, that offsets start at 0
// and never decrease, that ids lie in [0, n), and the checksum.
bool save_binary(const std::string& path, const Graph& g, std::string& err);
// Returns false with err set on any failure; g.two stays empty when the file holds no
// two-hop lists (two.off all zero).
bool load_binary(const std::string& path, Graph& g, std::string& err, bool* has_twohop = nullptr);
// True when the file starts with the binary magic.
bool is_binary_graph(const std::string& path);
// Text or binary graph, with its two-hop lists built when the file does not hold them.
// Returns 0, 1 on I/O failure, 2 on malformed content.
int load_graph_file(const std::string& path, Graph& g, std::string& err, int threads);

// md[0] = 0, md[k] = ceil(gamma (k-1)) in IEEE double (reference
// src/host_functions.cu:9-15), for k < len.
std::vector<int32_t> min_degree_table(double gamma, int len);
// ub[d] = (int)(d / gamma) in IEEE double, the upper-bound division of
// reference src/host_functions.cu:2165-2166 and src/device_kernels.cu:1402-1403.
std::vector<int32_t> bound_div_table(double gamma, int len);

// Level-0 peel, cover pruning and graph condensing (reference
// src/host_functions.cu:461-934). Replaces g by the condensed graph whose ids are the
// level-0 enumeration ranks, and returns the level-0 task in enumeration order.
struct Level0 {
    std::vector<TaskVertex> task;
    int32_t ncov = 0;              // covered vertices right after the (zero) members
    std::vector<int32_t> orig_of;  // input id of each condensed id
};
// l2_full (optional): whole two-hop list lengths when g.two holds restricted lists
// (build_twohop_restricted); null = the lengths of g.two's lists.
void level0_condense(Graph& g, const int32_t* md_out, const int32_t* md_in, int min_size,
                     Level0& out, const std::vector<uint32_t>* l2_full = nullptr);
// The search path's preprocessing of a loaded graph without two-hop lists: degree
// flags from the tables, restricted two-hop lists, then level 0 + condense.
void level0_from_loaded(Graph& g, const int32_t* md_out, const int32_t* md_in, int min_size, int threads,
                        Level0& out);

// The maximal-set filter of reference src/Quick_rmnonmax.cpp (RemoveNonMax):
// reads "<len> ids..." records, writes the maximal distinct sets, each ascending,
// in lexicographic order. Returns the number written, -1 on I/O error. The output
// file is written only when the input holds at least one set.
int64_t remove_non_max(const std::string& in_path, const std::string& out_path);
// Same filter over in-memory sets (each set is sorted in place).
std::vector<std::vector<int32_t>> maximal_sets(std::vector<std::vector<int32_t>>& sets);

}  // namespace dqc
