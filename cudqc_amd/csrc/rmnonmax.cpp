// Maximal-set filter with the contract of reference src/Quick_rmnonmax.cpp:93-377
// (RemoveNonMax): sets are sorted, stored in a prefix trie (children ascending), every
// leaf that is a proper subset of another stored set is marked non-maximal, and the
// remaining leaves are written in trie order (= lexicographic order) as
// "<len> <v1> ... <vk> \n". Sets that are proper prefixes of other sets end inside the
// trie and are never written; duplicate sets collapse onto one leaf.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "graph.hpp"

namespace dqc {
namespace {

struct Trie {
    struct Node {
        int32_t id;
        int32_t child = -1;  // first child (smallest id)
        int32_t next = -1;   // next sibling (larger id)
        bool maximal = true;
    };
    std::vector<Node> nodes;
    int32_t root_first = -1;  // first top-level node

    void insert(const int32_t* s, int len) {
        int32_t* link = &root_first;
        for (int i = 0; i < len; ++i) {
            // walk the sorted sibling list at this depth
            while (*link != -1 && nodes[*link].id < s[i]) link = &nodes[*link].next;
            if (*link != -1 && nodes[*link].id == s[i]) {
                link = &nodes[*link].child;
                continue;
            }
            Node nd;
            nd.id = s[i];
            nd.next = *link;
            int32_t idx = (int32_t)nodes.size();
            nodes.push_back(nd);
            *link = idx;
            link = &nodes[idx].child;
        }
    }

    // Marks every leaf whose path is a subsequence of s[pos..] under the sibling list
    // starting at `first`, except `self`.
    void mark_subsets(int32_t first, const int32_t* s, int len, int pos, int32_t self,
                      std::vector<std::pair<int32_t, int>>& stack) {
        stack.clear();
        stack.push_back({first, pos});
        while (!stack.empty()) {
            auto [node, p] = stack.back();
            stack.pop_back();
            // merge-walk siblings against s[p..]
            while (node != -1 && p < len) {
                if (nodes[node].id < s[p]) { node = nodes[node].next; continue; }
                if (nodes[node].id > s[p]) { ++p; continue; }
                // match
                if (nodes[node].child == -1) {
                    if (node != self) nodes[node].maximal = false;
                } else {
                    stack.push_back({nodes[node].child, p + 1});
                }
                node = nodes[node].next;
                ++p;
            }
        }
    }
};

}  // namespace

std::vector<std::vector<int32_t>> maximal_sets(std::vector<std::vector<int32_t>>& sets) {
    Trie t;
    size_t total = 0;
    for (auto& s : sets) {
        std::sort(s.begin(), s.end());
        total += s.size();
    }
    t.nodes.reserve(total + 1);
    for (auto& s : sets)
        if (!s.empty()) t.insert(s.data(), (int)s.size());
    // enumerate leaves in trie order with their paths
    std::vector<std::vector<int32_t>> leaves_path;
    std::vector<int32_t> leaves_node;
    {
        std::vector<int32_t> path, stk;
        // iterative DFS: stack of node ids, path tracks ids
        struct Frame { int32_t node; };
        std::vector<int32_t> nodestack;
        if (t.root_first != -1) nodestack.push_back(t.root_first);
        std::vector<int> depth_of;
        depth_of.push_back(0);
        std::vector<std::pair<int32_t, int>> work;
        if (t.root_first != -1) work.push_back({t.root_first, 0});
        while (!work.empty()) {
            auto [node, d] = work.back();
            work.pop_back();
            path.resize(d);
            path.push_back(t.nodes[node].id);
            if (t.nodes[node].next != -1) work.push_back({t.nodes[node].next, d});
            if (t.nodes[node].child == -1) {
                leaves_path.push_back(path);
                leaves_node.push_back(node);
            } else {
                work.push_back({t.nodes[node].child, d + 1});
            }
        }
    }
    std::vector<std::pair<int32_t, int>> stack;
    for (size_t i = 0; i < leaves_node.size(); ++i) {
        if (!t.nodes[leaves_node[i]].maximal) continue;
        const auto& p = leaves_path[i];
        t.mark_subsets(t.root_first, p.data(), (int)p.size(), 0, leaves_node[i], stack);
    }
    std::vector<std::vector<int32_t>> out;
    for (size_t i = 0; i < leaves_node.size(); ++i)
        if (t.nodes[leaves_node[i]].maximal) out.push_back(std::move(leaves_path[i]));
    return out;
}

int64_t remove_non_max(const std::string& in_path, const std::string& out_path) {
    FILE* f = std::fopen(in_path.c_str(), "rb");
    if (!f) return -1;
    std::vector<std::vector<int32_t>> sets;
    int len;
    while (std::fscanf(f, "%d", &len) == 1) {
        std::vector<int32_t> s((size_t)(len > 0 ? len : 0));
        bool ok = true;
        for (int i = 0; i < len; ++i)
            if (std::fscanf(f, "%d", &s[i]) != 1) { ok = false; break; }
        if (!ok) break;
        sets.push_back(std::move(s));
    }
    std::fclose(f);
    if (sets.empty()) return 0;
    auto mx = maximal_sets(sets);
    FILE* o = std::fopen(out_path.c_str(), "wb");
    if (!o) return -1;
    std::string line;
    char tmp[16];
    for (auto& s : mx) {
        line.clear();
        int l = std::snprintf(tmp, sizeof tmp, "%d ", (int)s.size());
        line.append(tmp, l);
        for (int32_t x : s) { l = std::snprintf(tmp, sizeof tmp, "%d ", x); line.append(tmp, l); }
        line.push_back('\n');
        std::fwrite(line.data(), 1, line.size(), o);
    }
    std::fclose(o);
    return (int64_t)mx.size();
}

}  // namespace dqc
