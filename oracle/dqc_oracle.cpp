// ORACLE — test infrastructure only. Never linked into, loaded by or called from
// the product library (cudqc_amd). Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, and only as the checker / CPU baseline.
//
// A plain C++/OpenMP restatement of the reference's CPU path for maximal directed
// quasi-clique enumeration (Mike12041204/cuDQC). Every function cites the reference
// file:line it follows. Integer types mirror the reference's on purpose: degree
// fields are uint32_t, so comparisons such as `deg_sum < int_threshold` are unsigned
// exactly where the reference's are.
//
// Pinning (see tests/): minimum-degree tables against the closed form; loader against
// hand-built files; two-hop lists against an independent set-builder restatement of
// the definition; RemoveNonMax against a brute-force subset filter; the whole method
// against exhaustive subset enumeration on small digraphs (gamma_in == gamma_out).
// The reference itself cannot be compiled here (it needs cuda_runtime.h and mpi.h),
// so there is no oracle/_ref build.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

namespace orc {

// Mirrors `struct Vertex` (reference inc/common.hpp:64-73).
struct Rec {
    uint32_t id;
    int8_t label;  // 0 cand, 1 member, 2 covered, 3 cover, 4 critical-adjacent, -1 pruned
    uint32_t om, oc, im, ic, l2;  // out_mem, out_can, in_mem, in_can, lvl2adj
};

struct Graph {
    int n = 0;
    int64_t m = 0;
    std::vector<uint64_t> oo, io, to;  // offsets
    std::vector<int> on, in, tn;       // neighbours
    // restricted two-hop build only: each source vertex's whole two-hop list length (the
    // lists themselves keep only the flagged entries); empty after a full build
    std::vector<uint32_t> l2full;
};

// ---------------------------------------------------------------------------
// Loader: reference src/common.cpp:15-108 (line count = n; line i lists out-neighbours
// of i; lists sorted, de-duplicated, self loops dropped; m counted after cleanup).
// Returns 0, or -1 on I/O error, -2 on an id outside [0, n).
int load_text(const char* path, Graph& g) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return -1;
    std::string data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    // count lines the way getline does: a final line without '\n' still counts
    int n = 0;
    for (size_t i = 0; i < data.size(); ++i)
        if (data[i] == '\n') ++n;
    if (!data.empty() && data.back() != '\n') ++n;
    std::vector<std::vector<int>> out(n), inn(n);
    size_t p = 0;
    for (int line = 0; line < n; ++line) {
        size_t e = data.find('\n', p);
        if (e == std::string::npos) e = data.size();
        size_t q = p;
        while (q < e) {
            while (q < e && (data[q] == ' ' || data[q] == '\t' || data[q] == '\r')) ++q;
            if (q >= e) break;
            bool neg = false;
            if (data[q] == '-' || data[q] == '+') { neg = data[q] == '-'; ++q; }
            if (q >= e || data[q] < '0' || data[q] > '9') break;  // `>>` stops at junk
            long long v = 0;
            while (q < e && data[q] >= '0' && data[q] <= '9') { v = v * 10 + (data[q] - '0'); ++q; }
            if (neg) v = -v;
            if (v < 0 || v >= n) return -2;
            out[line].push_back((int)v);
            inn[(int)v].push_back(line);
        }
        p = e + 1;
    }
    g.n = n;
    g.oo.assign(n + 1, 0);
    g.io.assign(n + 1, 0);
    g.m = 0;
    for (int i = 0; i < n; ++i) {
        for (auto* vec : {&out[i], &inn[i]}) {
            std::sort(vec->begin(), vec->end());
            vec->erase(std::unique(vec->begin(), vec->end()), vec->end());
            vec->erase(std::remove(vec->begin(), vec->end(), i), vec->end());
        }
        g.m += (int64_t)out[i].size();
        g.oo[i + 1] = g.oo[i] + out[i].size();
        g.io[i + 1] = g.io[i] + inn[i].size();
    }
    g.on.resize(g.oo[n]);
    g.in.resize(g.io[n]);
    for (int i = 0; i < n; ++i) {
        std::copy(out[i].begin(), out[i].end(), g.on.begin() + g.oo[i]);
        std::copy(inn[i].begin(), inn[i].end(), g.in.begin() + g.io[i]);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Directed two-hop lists: reference src/common.cpp:111-509 (GenLevel2NBs).
// For v with out>0 and in>0:
//   B = bidirectional neighbours, O/I = out-only / in-only neighbours;
//   greatest fixed point, both sides updated from the previous round's sets:
//     I keeps w if some u in O u B has u->w;  O keeps w if some u in I u B has w->u;
//   L(v) = B u O* u I* u C, where C = {x != v, x not yet in L, with
//     some u1 in O*uB: u1->x, some u2 in O*uB: x->u2, some u3 in I*uB: u3->x,
//     some u4 in I*uB: x->u4}; sorted ascending.
// Otherwise L(v) is empty.
// flag (restricted build): only flagged vertices get a list, it keeps only flagged
// entries, and g.l2full[v] records the whole list's length. Level 0 reads no list and no
// length of an unflagged vertex when the flags are its degree test (restricted_flags), so
// it gives the full build's result from far fewer entries.
void build_twohop(Graph& g, int threads, const std::vector<uint8_t>* flag = nullptr) {
    const int n = g.n;
    std::vector<std::vector<int>> lists(n);
    if (threads < 1) threads = 1;
    if (flag) g.l2full.assign(n, 0);
    else g.l2full.clear();
#pragma omp parallel num_threads(threads)
    {
        std::vector<int> st_out(n, 0), st_in(n, 0);  // round stamps of O / I members
        std::vector<uint8_t> isout(n, 0), isin(n, 0), inl(n, 0), mark(n, 0);
        std::vector<int> B, O, I, nO, nI, L, touched;
#pragma omp for schedule(dynamic, 64)
        for (int v = 0; v < n; ++v) {
            uint64_t ob = g.oo[v], oe = g.oo[v + 1], ib = g.io[v], ie = g.io[v + 1];
            if (oe == ob || ie == ib) continue;
            if (flag && !(*flag)[v]) continue;
            B.clear(); O.clear(); I.clear();
            for (uint64_t k = ob; k < oe; ++k) isout[g.on[k]] = 1;
            for (uint64_t k = ib; k < ie; ++k) isin[g.in[k]] = 1;
            for (uint64_t k = ob; k < oe; ++k) {
                int w = g.on[k];
                if (isin[w]) B.push_back(w);
                else { O.push_back(w); st_out[w] = 1; }
            }
            for (uint64_t k = ib; k < ie; ++k) {
                int w = g.in[k];
                if (!isout[w]) { I.push_back(w); st_in[w] = 1; }
            }
            int round = 1;
            while (true) {
                size_t so = O.size(), si = I.size();
                nO.clear(); nI.clear();
                // I-side survivors: reached by an arc from O u B
                for (int pass = 0; pass < 2; ++pass) {
                    const std::vector<int>& src = pass == 0 ? O : B;
                    for (int u : src)
                        for (uint64_t k = g.oo[u]; k < g.oo[u + 1]; ++k) {
                            int w = g.on[k];
                            if (st_in[w] == round) { st_in[w] = round + 1; nI.push_back(w); }
                        }
                }
                // O-side survivors: have an arc into I u B
                for (int pass = 0; pass < 2; ++pass) {
                    const std::vector<int>& src = pass == 0 ? I : B;
                    for (int u : src)
                        for (uint64_t k = g.io[u]; k < g.io[u + 1]; ++k) {
                            int w = g.in[k];
                            if (st_out[w] == round) { st_out[w] = round + 1; nO.push_back(w); }
                        }
                }
                O.swap(nO);
                I.swap(nI);
                ++round;
                if (!(O.size() < so || I.size() < si)) break;
            }
            for (uint64_t k = ob; k < oe; ++k) { st_out[g.on[k]] = 0; isout[g.on[k]] = 0; }
            for (uint64_t k = ib; k < ie; ++k) { st_in[g.in[k]] = 0; isin[g.in[k]] = 0; }
            L.clear();
            for (int w : B) { L.push_back(w); inl[w] = 1; }
            for (int w : O) { L.push_back(w); inl[w] = 1; }
            for (int w : I) { L.push_back(w); inl[w] = 1; }
            // C: four conditions, checked as successive marks 1 -> 2 -> 3 -> add
            touched.clear();
            auto OB = [&](auto&& f) { for (int u : O) f(u); for (int u : B) f(u); };
            auto IB = [&](auto&& f) { for (int u : I) f(u); for (int u : B) f(u); };
            OB([&](int u) {
                for (uint64_t k = g.oo[u]; k < g.oo[u + 1]; ++k) {
                    int x = g.on[k];
                    if (x != v && !inl[x] && mark[x] == 0) { mark[x] = 1; touched.push_back(x); }
                }
            });
            OB([&](int u) {
                for (uint64_t k = g.io[u]; k < g.io[u + 1]; ++k) {
                    int x = g.in[k];
                    if (mark[x] == 1) mark[x] = 2;
                }
            });
            IB([&](int u) {
                for (uint64_t k = g.oo[u]; k < g.oo[u + 1]; ++k) {
                    int x = g.on[k];
                    if (mark[x] == 2) mark[x] = 3;
                }
            });
            IB([&](int u) {
                for (uint64_t k = g.io[u]; k < g.io[u + 1]; ++k) {
                    int x = g.in[k];
                    if (mark[x] == 3 && !inl[x]) { inl[x] = 1; L.push_back(x); }
                }
            });
            for (int x : touched) mark[x] = 0;
            for (int x : L) inl[x] = 0;
            if (flag) {
                g.l2full[v] = (uint32_t)L.size();
                L.erase(std::remove_if(L.begin(), L.end(), [&](int x) { return !(*flag)[x]; }), L.end());
            }
            std::sort(L.begin(), L.end());
            lists[v] = L;
        }
    }
    g.to.assign(n + 1, 0);
    for (int v = 0; v < n; ++v) g.to[v + 1] = g.to[v] + lists[v].size();
    g.tn.resize(g.to[n]);
    for (int v = 0; v < n; ++v) std::copy(lists[v].begin(), lists[v].end(), g.tn.begin() + g.to[v]);
}

// Two-hop lists for level 0 (mo, mi: the min-degree thresholds at min_size): all of them,
// or with g_twohop_restricted only those of the vertices passing level 0's degree test,
// holding only such vertices (the product's host build does the same on its search path).
static bool g_twohop_restricted = false;
void build_twohop_for_level0(Graph& g, int threads, int mo, int mi) {
    if (!g_twohop_restricted) {
        build_twohop(g, threads);
        return;
    }
    std::vector<uint8_t> flag((size_t)g.n);
    for (int v = 0; v < g.n; ++v)
        flag[v] = (g.oo[v + 1] - g.oo[v] >= (uint64_t)mo) && (g.io[v + 1] - g.io[v] >= (uint64_t)mi);
    build_twohop(g, threads, &flag);
}

// ---------------------------------------------------------------------------
// md[0] = 0, md[k] = ceil(gamma * (k - 1)) in IEEE double
// (reference src/host_functions.cu:9-15). The reference allocates n+1 entries and
// reads past them only in degenerate cases; the table here continues the formula.
std::vector<int> mindeg_table(double gamma, int len) {
    std::vector<int> md(len);
    md[0] = 0;
    for (int k = 1; k < len; ++k) md[k] = (int)std::ceil(gamma * (k - 1));
    return md;
}

// ---------------------------------------------------------------------------
// Expansion. Mode selects the reference's host-level or device-level semantics.
enum Mode { HOST_L0 = 0, HOST_L1 = 1, DEVICE = 2 };

struct Params {
    const int* mdo;
    const int* mdi;
    double go, gi;
    int min;
    bool host_addfail_check;  // true: union of the 4 ranks (rank != 0 check), false: rank 0 only
};

struct Stats {
    uint64_t task_hist[32] = {0}, kid_hist[32] = {0};
    uint64_t tasks = 0, branches = 0, raw_cliques = 0, children = 0;
    uint64_t bytes = 0;  // algorithmic bytes, byte model of the measurement plan
};

struct Sink {
    std::vector<int> clq;         // concatenated: len, ids...
    std::vector<std::vector<Rec>> kids;
    Stats st;
};

// Q order (reference inc/host_functions.hpp:74-128, inc/device_kernels.hpp:31-80):
// label 1 < 2 < 0 < 3 < others, then min(mem) desc, min(can) desc, lvl2adj desc, id desc.
inline int label_rank(int8_t l) {
    switch (l) { case 1: return 0; case 2: return 1; case 0: return 2; case 3: return 3; default: return 4; }
}
inline bool q_less(const Rec& a, const Rec& b) {
    int ra = label_rank(a.label), rb = label_rank(b.label);
    if (ra != rb) return ra < rb;
    int am = (int)std::min(a.om, a.im), bm = (int)std::min(b.om, b.im);
    if (am != bm) return am > bm;
    int ac = (int)std::min(a.oc, a.ic), bc = (int)std::min(b.oc, b.ic);
    if (ac != bc) return ac > bc;
    if (a.l2 != b.l2) return a.l2 > b.l2;
    return a.id > b.id;
}

struct Expander {
    const Graph& g;
    const Params& P;
    std::vector<int> pos;  // id -> position, -1 absent (hd.vertex_order_map)
    std::vector<int> rem, rmv, com, cim, adj, tmpi;
    Sink* sink = nullptr;

    Expander(const Graph& gr, const Params& p) : g(gr), P(p), pos(gr.n, -1) {}

    int gmd(int k, const int* md) const { return k < P.min ? md[P.min] : md[k]; }
    void csr_bytes(uint64_t len) { sink->st.bytes += 4 * len + 16; }

    void map_set(const Rec* v, int n) { for (int i = 0; i < n; ++i) pos[v[i].id] = i; }
    void map_clear(const Rec* v, int n) { for (int i = 0; i < n; ++i) pos[v[i].id] = -1; }

    void emit_clique(const Rec* v, int n) {
        sink->clq.push_back(n);
        for (int i = 0; i < n; ++i) sink->clq.push_back((int)v[i].id);
        sink->st.raw_cliques++;
        sink->st.bytes += 4ull * n;
    }

    // reference inc/host_functions.hpp:167-193 (h_cand_isvalid); unsigned compares kept
    bool cand_valid(const Rec& v, int cs, int ub, int lb, int meo, int mei) const {
        const int *mdo = P.mdo, *mdi = P.mdi;
        if (v.om + v.oc < (unsigned)meo) return false;
        if (v.im + v.ic < (unsigned)mei) return false;
        if (v.om + v.oc < (unsigned)gmd((int)(cs + v.oc + 1), mdo)) return false;
        if (v.im + v.ic < (unsigned)gmd((int)(cs + v.ic + 1), mdi)) return false;
        if (v.om + (unsigned)ub - 1u < (unsigned)mdo[cs + ub]) return false;
        if (v.im + (unsigned)ub - 1u < (unsigned)mdi[cs + ub]) return false;
        if (v.om + v.oc < (unsigned)gmd(cs + lb, mdo)) return false;
        if (v.im + v.ic < (unsigned)gmd(cs + lb, mdi)) return false;
        return true;
    }
    // reference inc/host_functions.hpp:194-228 (h_vert_isextendable)
    bool vert_ext(const Rec& v, int cs, int ub, int lb, int meo, int mei) const {
        const int *mdo = P.mdo, *mdi = P.mdi;
        if (v.om + v.oc < (unsigned)meo) return false;
        if (v.im + v.ic < (unsigned)mei) return false;
        if (v.om + v.oc < (unsigned)gmd((int)(cs + v.oc), mdo)) return false;
        if (v.im + v.ic < (unsigned)gmd((int)(cs + v.ic), mdi)) return false;
        if (v.oc == 0 && v.om < (unsigned)gmd(cs + 1, mdo)) return false;
        if (v.ic == 0 && v.im < (unsigned)gmd(cs + 1, mdi)) return false;
        if (v.om + (unsigned)ub < (unsigned)mdo[cs + ub]) return false;
        if (v.im + (unsigned)ub < (unsigned)mdi[cs + ub]) return false;
        if (v.om + v.oc < (unsigned)gmd(cs + lb, mdo)) return false;
        if (v.im + v.ic < (unsigned)gmd(cs + lb, mdi)) return false;
        return true;
    }

    // reference src/host_functions.cu:2064-2252 / src/device_kernels.cu:1307-1486.
    // Returns false where either returns failure (the host version keeps computing
    // after a failure, but its caller discards the bounds in that case).
    bool lu_bounds(const Rec* v, int nm, int ncand, int& ub, int& lb, int& meo, int& mei) {
        const int *mdo = P.mdo, *mdi = P.mdi;
        int mino = (int)v[0].om, mino_c = (int)v[0].oc, sumo = (int)v[0].om, mint_o = (int)(v[0].om + v[0].oc);
        int mini = (int)v[0].im, mini_c = (int)v[0].ic, sumi = (int)v[0].im, mint_i = (int)(v[0].im + v[0].ic);
        for (int i = 1; i < nm; ++i) {
            sumo += (int)v[i].om;
            if ((unsigned)mino > v[i].om) { mino = (int)v[i].om; mino_c = (int)v[i].oc; }
            else if ((unsigned)mino == v[i].om) { if ((unsigned)mino_c > v[i].oc) mino_c = (int)v[i].oc; }
            if ((unsigned)mint_o > v[i].om + v[i].oc) mint_o = (int)(v[i].om + v[i].oc);
            sumi += (int)v[i].im;
            if ((unsigned)mini > v[i].im) { mini = (int)v[i].im; mini_c = (int)v[i].ic; }
            else if ((unsigned)mini == v[i].im) { if ((unsigned)mini_c > v[i].ic) mini_c = (int)v[i].ic; }
            if ((unsigned)mint_i > v[i].im + v[i].ic) mint_i = (int)(v[i].im + v[i].ic);
        }
        meo = gmd(nm + 1, mdo);
        mei = gmd(nm + 1, mdi);
        if (mino < mdo[nm] || mini < mdi[nm]) {
            int nmin = std::max(gmd(nm, mdo) - mino, gmd(nm, mdi) - mini);
            int no = nmin;
            while (no <= mino_c && mino + no < mdo[nm + no]) ++no;
            if (mino + no < mdo[nm + no]) return false;
            int ni = nmin;
            while (ni <= mini_c && mini + ni < mdi[nm + ni]) ++ni;
            if (mini + ni < mdi[nm + ni]) return false;
            lb = std::max(no, ni);
            ub = std::min((int)(mint_o / P.go), (int)(mint_i / P.gi)) + 1 - nm;
            if (ub > ncand) ub = ncand;
            if (lb < ub) {
                int so = 0, si = 0, i;
                for (i = 0; i < lb; ++i) { so += com[i]; si += cim[i]; }
                while (i < ub && sumo + si < nm * mdo[nm + i] && sumi + so < nm * mdi[nm + i]) {
                    so += com[i]; si += cim[i]; ++i;
                }
                if (sumo + so < nm * mdo[nm + i] && sumi + si < nm * mdi[nm + i]) return false;
                lb = i;
                int tight = i;
                while (i < ub) {
                    so += com[i]; si += cim[i]; ++i;
                    if (sumo + si >= nm * mdo[nm + i] && sumi + so >= nm * mdi[nm + i]) tight = i;
                }
                if (ub > tight) ub = tight;
                if (lb > 1) { meo = gmd(nm + lb, mdo); mei = gmd(nm + lb, mdi); }
            }
        } else {
            ub = ncand;
            lb = nm < P.min ? P.min - nm : 0;
        }
        if (nm + ub < P.min) return false;
        if (ub < 0 || ub < lb) return false;
        return true;
    }

    static void sort_desc(std::vector<int>& a, int n) { std::sort(a.begin(), a.begin() + n, std::greater<int>()); }

    // reference src/host_functions.cu:1841-2061 (h_degree_pruning). Expects the
    // position map set for v[0..tv) and com/cim[0..rc) filled; clears the map.
    bool degree_pruning(Rec* v, int& tv, int& nc, int nm, int rc, int& ub, int& lb, int& meo, int& mei) {
        sort_desc(com, rc);
        sort_desc(cim, rc);
        bool ok = lu_bounds(v, nm, rc, ub, lb, meo, mei);
        if (ok)
            for (int i = 0; i < nm; ++i)
                if (!vert_ext(v[i], nm, ub, lb, meo, mei)) { ok = false; break; }
        if (!ok) { map_clear(v, tv); return false; }
        int nr = 0, nd = 0;
        for (int i = nm; i < tv; ++i) {
            if (v[i].label == 0 && cand_valid(v[i], nm, ub, lb, meo, mei)) rem[nr++] = i;
            else rmv[nd++] = i;
        }
        while (nr > 0 && nd > 0) {
            if (nr < nd) {
                for (int i = 0; i < tv; ++i) { v[i].ic = 0; v[i].oc = 0; }
                for (int r = 0; r < nr; ++r) {
                    int id = (int)v[rem[r]].id;
                    csr_bytes(g.oo[id + 1] - g.oo[id]);
                    csr_bytes(g.io[id + 1] - g.io[id]);
                    for (uint64_t k = g.oo[id]; k < g.oo[id + 1]; ++k) { int p = pos[g.on[k]]; if (p > -1) v[p].ic++; }
                    for (uint64_t k = g.io[id]; k < g.io[id + 1]; ++k) { int p = pos[g.in[k]]; if (p > -1) v[p].oc++; }
                }
            } else {
                for (int r = 0; r < nd; ++r) {
                    int id = (int)v[rmv[r]].id;
                    csr_bytes(g.oo[id + 1] - g.oo[id]);
                    csr_bytes(g.io[id + 1] - g.io[id]);
                    for (uint64_t k = g.oo[id]; k < g.oo[id + 1]; ++k) { int p = pos[g.on[k]]; if (p > -1) v[p].ic--; }
                    for (uint64_t k = g.io[id]; k < g.io[id + 1]; ++k) { int p = pos[g.in[k]]; if (p > -1) v[p].oc--; }
                }
            }
            int nv = 0;
            for (int r = 0; r < nr; ++r) {
                const Rec& c = v[rem[r]];
                if (cand_valid(c, nm, ub, lb, meo, mei)) { com[nv] = (int)c.om; cim[nv] = (int)c.im; ++nv; }
            }
            sort_desc(com, nv);
            sort_desc(cim, nv);
            ok = lu_bounds(v, nm, nv, ub, lb, meo, mei);
            if (ok)
                for (int i = 0; i < nm; ++i)
                    if (!vert_ext(v[i], nm, ub, lb, meo, mei)) { ok = false; break; }
            if (!ok) { map_clear(v, tv); return false; }
            int keep = 0;
            nd = 0;
            for (int r = 0; r < nr; ++r) {
                if (cand_valid(v[rem[r]], nm, ub, lb, meo, mei)) rem[keep++] = rem[r];
                else rmv[nd++] = rem[r];
            }
            nr = keep;
        }
        map_clear(v, tv);
        for (int i = 0; i < nr; ++i) v[nm + i] = v[rem[i]];
        tv = tv - nc + nr;
        nc = nr;
        return true;
    }

    // reference src/host_functions.cu:1503-1571 + 1780-1818 (add, then diameter pruning)
    bool add_one(Rec* v, int& tv, int& nc, int& nm, int& ub, int& lb, int& meo, int& mei) {
        int mo = gmd(nm + 2, P.mdo), mi = gmd(nm + 2, P.mdi);
        map_set(v, tv);
        int p = (int)v[nm].id;
        v[nm].label = 1;
        ++nm;
        --nc;
        csr_bytes(g.oo[p + 1] - g.oo[p]);
        csr_bytes(g.io[p + 1] - g.io[p]);
        csr_bytes(g.to[p + 1] - g.to[p]);
        for (uint64_t k = g.oo[p]; k < g.oo[p + 1]; ++k) { int q = pos[g.on[k]]; if (q > -1) { v[q].im++; v[q].ic--; } }
        for (uint64_t k = g.io[p]; k < g.io[p + 1]; ++k) { int q = pos[g.in[k]]; if (q > -1) { v[q].om++; v[q].oc--; } }
        for (int i = nm; i < tv; ++i) v[i].label = -1;
        int rc = 0;
        for (uint64_t k = g.to[p]; k < g.to[p + 1]; ++k) {
            int q = pos[g.tn[k]];
            if (q >= nm && v[q].om + v[q].oc >= (unsigned)mo && v[q].im + v[q].ic >= (unsigned)mi) {
                v[q].label = 0;
                com[rc] = (int)v[q].om;
                cim[rc] = (int)v[q].im;
                ++rc;
            }
        }
        return degree_pruning(v, tv, nc, nm, rc, ub, lb, meo, mei);
    }

    // reference src/host_functions.cu:1574-1778 + 1820-1838. Returns 2 (critical
    // failure), 0 (failed vertex / invalid bounds) or 1.
    int critical(Rec* v, int& tv, int& nc, int& nm, int& ub, int& lb, int& meo, int& mei) {
        map_set(v, tv);
        for (int i = 0; i < nm; ++i) {
            int p = (int)v[i].id;
            if (v[i].om + v[i].oc == (unsigned)P.mdo[nm + lb] && v[i].oc > 0) {
                csr_bytes(g.oo[p + 1] - g.oo[p]);
                for (uint64_t k = g.oo[p]; k < g.oo[p + 1]; ++k) { int q = pos[g.on[k]]; if (q >= nm) v[q].label = 4; }
            }
            if (v[i].im + v[i].ic == (unsigned)P.mdi[nm + lb] && v[i].ic > 0) {
                csr_bytes(g.io[p + 1] - g.io[p]);
                for (uint64_t k = g.io[p]; k < g.io[p + 1]; ++k) { int q = pos[g.in[k]]; if (q >= nm) v[q].label = 4; }
            }
        }
        map_clear(v, tv);
        std::stable_partition(v + nm, v + tv, [](const Rec& r) { return r.label == 4; });
        int ncrit = 0;
        while (nm + ncrit < tv && v[nm + ncrit].label == 4) ++ncrit;
        if (ncrit == 0) return 1;
        map_set(v, tv);
        std::fill(adj.begin(), adj.begin() + tv, 0);
        for (int i = nm; i < nm + ncrit; ++i) {
            int p = (int)v[i].id;
            csr_bytes(g.to[p + 1] - g.to[p]);
            for (uint64_t k = g.to[p]; k < g.to[p + 1]; ++k) { int q = pos[g.tn[k]]; if (q > -1) adj[q]++; }
        }
        for (int i = 0; i < nm; ++i)
            if (adj[i] != ncrit) { map_clear(v, tv); return 2; }
        for (int i = nm; i < nm + ncrit; ++i)
            if (adj[i] < ncrit - 1) { map_clear(v, tv); return 2; }
        for (int i = nm; i < nm + ncrit; ++i) {
            int p = (int)v[i].id;
            csr_bytes(g.oo[p + 1] - g.oo[p]);
            csr_bytes(g.io[p + 1] - g.io[p]);
            for (uint64_t k = g.oo[p]; k < g.oo[p + 1]; ++k) { int q = pos[g.on[k]]; if (q > -1) { v[q].im++; v[q].ic--; } }
            for (uint64_t k = g.io[p]; k < g.io[p + 1]; ++k) { int q = pos[g.in[k]]; if (q > -1) { v[q].om++; v[q].oc--; } }
        }
        for (int i = nm; i < nm + ncrit; ++i) v[i].label = 1;
        nm += ncrit;
        nc -= ncrit;
        int rc = 0;
        for (int i = nm; i < tv; ++i) {
            if (adj[i] == ncrit) { com[rc] = (int)v[i].om; cim[rc] = (int)v[i].im; ++rc; }
            else v[i].label = -1;
        }
        return degree_pruning(v, tv, nc, nm, rc, ub, lb, meo, mei) ? 1 : 0;
    }

    // reference src/host_functions.cu:2254-2288 / src/device_kernels.cu:1488-1532
    void check_clique(const Rec* v, int nm) {
        if (nm < P.min) return;
        unsigned mo = (unsigned)P.mdo[nm], mi = (unsigned)P.mdi[nm];
        for (int i = 0; i < nm; ++i)
            if (v[i].om < mo || v[i].im < mi) return;
        emit_clique(v, nm);
    }

    // reference src/host_functions.cu:1386-1500 / src/device_kernels.cu:400-531.
    // Returns 0 (stop the task), 2 (skip this branch) or 1. A member falling short
    // wins over the next vertex falling short (the host's sequential order).
    int remove_one(Rec* v, int& tv, int& nc, int nm) {
        --nc;
        --tv;
        if (tv < P.min) return 0;
        int mo = gmd(nm + 1, P.mdo), mi = gmd(nm + 1, P.mdi);
        map_set(v, tv);
        int p = (int)v[tv].id;
        int res = 1;
        csr_bytes(g.oo[p + 1] - g.oo[p]);
        for (uint64_t k = g.oo[p]; k < g.oo[p + 1]; ++k) {
            int q = pos[g.on[k]];
            if (q > -1) {
                v[q].ic--;
                if (v[q].im + v[q].ic < (unsigned)mi) {
                    if (q < nm) { res = 0; break; }
                    else if (q == tv - 1) res = 2;
                }
            }
        }
        if (res == 0) { map_clear(v, tv); return 0; }
        csr_bytes(g.io[p + 1] - g.io[p]);
        for (uint64_t k = g.io[p]; k < g.io[p + 1]; ++k) {
            int q = pos[g.in[k]];
            if (q > -1) {
                v[q].oc--;
                if (v[q].om + v[q].oc < (unsigned)mo) {
                    if (q < nm) { res = 0; break; }
                    else if (q == tv - 1) res = 2;
                }
            }
        }
        if (res == 0) { map_clear(v, tv); return 0; }
        csr_bytes(g.to[p + 1] - g.to[p]);
        for (uint64_t k = g.to[p]; k < g.to[p + 1]; ++k) {
            int q = pos[g.tn[k]];
            if (q >= nm) v[q].l2--;
        }
        map_clear(v, tv);
        return res;
    }

    // reference src/host_functions.cu:1333-1383 (host: lvl2adj >= tot-1) and
    // src/device_kernels.cu:330-397 (device: lvl2adj >= num_cand-1)
    bool lookahead(const Rec* v, int tv, int nm, int nc, Mode mode) {
        unsigned mo = (unsigned)gmd(tv, P.mdo), mi = (unsigned)gmd(tv, P.mdi);
        for (int i = 0; i < nm; ++i)
            if (v[i].om + v[i].oc < mo || v[i].im + v[i].ic < mi) return false;
        unsigned l2min = mode == DEVICE ? (unsigned)(nc - 1) : (unsigned)(tv - 1);
        for (int i = nm; i < tv; ++i)
            if (v[i].om + v[i].oc < mo || v[i].im + v[i].ic < mi || v[i].l2 < l2min) return false;
        return true;
    }

    void ensure(int n) {
        if ((int)rem.size() < n + 4) {
            rem.resize(n + 4); rmv.resize(n + 4); com.resize(n + 4); cim.resize(n + 4); adj.resize(n + 4);
        }
    }

    // One task's branch loop: reference src/host_functions.cu:987-1134 (host levels)
    // and src/device_kernels.cu:25-173 (device levels).
    // max_steps > 0: only the loop steps j < ncov + first_step + max_steps run (a bounded
    // prefix of the same loop, the library's root_branches unit of work). first_step > 0:
    // the steps before ncov + first_step (a window's head) still remove and test
    // lookahead, but their branches run only until one of them produces a child, which
    // is then the first child (host level-1 role) and is dropped, as are the cliques,
    // children and branch counts of every head branch.
    void expand(const Rec* task, int tot, Mode mode, int max_steps = 0, int first_step = 0) {
        ensure(tot);
        sink->st.tasks++;
        sink->st.task_hist[31 - __builtin_clz((unsigned)tot | 1)]++;
        sink->st.bytes += 24ull * tot;
        std::vector<Rec> par(task, task + tot);
        Rec* rv = par.data();
        int nm = 0;
        while (nm < tot && rv[nm].label == 1) ++nm;
        int ncov = 0;
        if (mode != DEVICE)
            while (nm + ncov < tot && rv[nm + ncov].label == 2) ++ncov;
        int nc = tot - nm;
        const int expansions = max_steps > 0 ? std::min(nc, ncov + first_step + max_steps) : nc;
        std::vector<Rec> child(tot);
        bool head_child = false;  // a head branch produced a child
        Sink* const real_sink = sink;
        Sink scratch;
        for (int j = ncov; j < expansions; ++j) {
            if (j != ncov) {
                int r = remove_one(rv, tot, nc, nm);
                if (r == 0) break;
                if (r == 2) continue;
            }
            if (lookahead(rv, tot, nm, nc, mode)) {
                emit_clique(rv, tot);
                break;
            }
            const bool head = j - ncov < first_step;
            if (head && head_child) continue;
            // a head branch writes into a scratch sink: only whether it yields a child counts
            if (head) {
                scratch = Sink();
                sink = &scratch;
            }
            run_branch(rv, child.data(), tot, nm, nc, ncov, mode);
            if (head) {
                head_child = !scratch.kids.empty();
                sink = real_sink;
            }
        }
    }

    // The branch of the current step: child = members + the last vertex + the rest, then
    // add, pruning, critical, clique check and the child (reference
    // src/host_functions.cu:1046-1133, src/device_kernels.cu:92-165).
    void run_branch(Rec* rv, Rec* cv, int tot, int nm, int nc, int ncov, Mode mode) {
        int cm = nm, cc = nc, ct = tot;
        for (int i = 0; i < nm; ++i) cv[i] = rv[i];
        cv[nm] = rv[tot - 1];
        for (int i = nm; i < tot - 1; ++i) cv[i + 1] = rv[i];
        for (int i = nm + 1; i <= nm + ncov; ++i) cv[i].label = 0;
        int ub = 0, lb = 0, meo = 0, mei = 0;
        sink->st.branches++;
        if (!add_one(cv, ct, cc, cm, ub, lb, meo, mei)) {
            if (mode == DEVICE || P.host_addfail_check) check_clique(cv, cm);
            return;
        }
        int s = critical(cv, ct, cc, cm, ub, lb, meo, mei);
        if (s == 2) return;
        check_clique(cv, cm);
        if (s == 0) return;
        std::sort(cv, cv + ct, q_less);
        if (cc > 0) {
            sink->kids.emplace_back(cv, cv + ct);
            sink->st.children++;
            sink->st.kid_hist[31 - __builtin_clz((unsigned)ct | 1)]++;
            sink->st.bytes += 24ull * ct;
        }
    }
};

// ---------------------------------------------------------------------------
// Level 0: reference src/host_functions.cu:461-686 (h_initialize_tasks) and the
// condense step :690-934 (h_condense_graph). Produces the condensed graph (ids =
// level-0 Q-order rank) and the root task. `orig_of[i]` = input id of condensed i.
void level0(Graph& g, const Params& P, std::vector<Rec>& root, std::vector<int>& orig_of) {
    const int n = g.n;
    const int mo = P.mdo[P.min], mi = P.mdi[P.min];
    std::vector<Rec> v(n);
    std::vector<int> rem, rmv;
    for (int i = 0; i < n; ++i) {
        v[i].id = (uint32_t)i;
        v[i].om = 0; v[i].im = 0;
        v[i].oc = (uint32_t)(g.oo[i + 1] - g.oo[i]);
        v[i].ic = (uint32_t)(g.io[i + 1] - g.io[i]);
        v[i].l2 = g.l2full.empty() ? (uint32_t)(g.to[i + 1] - g.to[i]) : g.l2full[i];
        if (v[i].oc >= (unsigned)mo && v[i].ic >= (unsigned)mi && v[i].l2 >= (unsigned)(P.min - 1)) {
            v[i].label = 0; rem.push_back(i);
        } else {
            v[i].label = -1; rmv.push_back(i);
        }
    }
    int ncand = n;
    while ((int)rem.size() < ncand / 2) {  // peel rounds while fewer than half remain
        ncand = (int)rem.size();
        for (int x : rem) { v[x].oc = 0; v[x].ic = 0; }
        for (int x : rem) {
            for (uint64_t k = g.oo[x]; k < g.oo[x + 1]; ++k) if (v[g.on[k]].label == 0) v[g.on[k]].ic++;
            for (uint64_t k = g.io[x]; k < g.io[x + 1]; ++k) if (v[g.in[k]].label == 0) v[g.in[k]].oc++;
        }
        std::vector<int> keep;
        rmv.clear();
        for (int x : rem) {
            if (v[x].oc >= (unsigned)mo && v[x].ic >= (unsigned)mi) keep.push_back(x);
            else { v[x].label = -1; rmv.push_back(x); }
        }
        rem.swap(keep);
    }
    ncand = (int)rem.size();
    for (size_t h = 0; h < rmv.size(); ++h) {  // cascade
        int x = rmv[h];
        for (uint64_t k = g.oo[x]; k < g.oo[x + 1]; ++k) {
            int y = g.on[k];
            if (v[y].label == 0) {
                v[y].ic--;
                if (v[y].ic < (unsigned)mi) { v[y].label = -1; --ncand; rmv.push_back(y); }
            }
        }
        for (uint64_t k = g.io[x]; k < g.io[x + 1]; ++k) {
            int y = g.in[k];
            if (v[y].label == 0) {
                v[y].oc--;
                if (v[y].oc < (unsigned)mo) { v[y].label = -1; --ncand; rmv.push_back(y); }
            }
        }
    }
    // cover vertex: first index of max min(out_can, in_can), max starts at 0
    int best = 0, besti = 0;
    for (int i = 0; i < n; ++i)
        if (v[i].label == 0) {
            int d = (int)std::min(v[i].oc, v[i].ic);
            if (d > best) { best = d; besti = i; }
        }
    v[besti].label = 3;
    {
        std::vector<uint8_t> mark(n, 0);
        for (uint64_t k = g.oo[besti]; k < g.oo[besti + 1]; ++k) if (v[g.on[k]].label == 0) mark[g.on[k]] = 1;
        for (uint64_t k = g.io[besti]; k < g.io[besti + 1]; ++k) if (mark[g.in[k]]) v[g.in[k]].label = 2;
    }
    std::sort(v.begin(), v.end(), q_less);
    // condense: survivors are v[0..ncand); relabel to their position
    std::vector<int> newid(n, -1);
    orig_of.assign(ncand, 0);
    for (int i = 0; i < ncand; ++i) { newid[v[i].id] = i; orig_of[i] = (int)v[i].id; }
    Graph c;
    c.n = ncand;
    c.oo.assign(ncand + 1, 0); c.io.assign(ncand + 1, 0); c.to.assign(ncand + 1, 0);
    std::vector<int> tmp;
    auto restrict_list = [&](const std::vector<uint64_t>& off, const std::vector<int>& nb, int x,
                             std::vector<int>& dst, std::vector<uint64_t>& doff, int i) {
        tmp.clear();
        for (uint64_t k = off[x]; k < off[x + 1]; ++k) if (newid[nb[k]] >= 0) tmp.push_back(newid[nb[k]]);
        std::sort(tmp.begin(), tmp.end());
        dst.insert(dst.end(), tmp.begin(), tmp.end());
        doff[i + 1] = doff[i] + tmp.size();
        return (uint32_t)tmp.size();
    };
    for (int i = 0; i < ncand; ++i) {
        int x = orig_of[i];
        v[i].l2 = restrict_list(g.to, g.tn, x, c.tn, c.to, i);
        uint32_t od = restrict_list(g.oo, g.on, x, c.on, c.oo, i);
        uint32_t id = restrict_list(g.io, g.in, x, c.in, c.io, i);
        if (od != v[i].oc || id != v[i].ic) std::fprintf(stderr, "oracle: inconsistent candidate degree\n");
    }
    c.m = (int64_t)c.on.size();
    for (int i = 0; i < ncand; ++i) v[i].id = (uint32_t)i;
    root.assign(v.begin(), v.begin() + ncand);
    g = std::move(c);
}

// ---------------------------------------------------------------------------
// RemoveNonMax contract (reference src/Quick_rmnonmax.cpp:93-377): each set sorted
// ascending; duplicates and proper subsets of other sets dropped; the survivors
// written in lexicographic order as "<len> <v1> ... <vk> \n" ("%d " per number).
// Restated with an inverted index instead of the reference's trie.
std::vector<std::vector<int>> maximal_sets(std::vector<std::vector<int>> sets) {
    for (auto& s : sets) std::sort(s.begin(), s.end());
    std::sort(sets.begin(), sets.end());
    sets.erase(std::unique(sets.begin(), sets.end()), sets.end());
    int maxid = -1;
    for (auto& s : sets) for (int x : s) maxid = std::max(maxid, x);
    std::vector<std::vector<int>> post(maxid + 1);
    for (int i = 0; i < (int)sets.size(); ++i) for (int x : sets[i]) post[x].push_back(i);
    std::vector<std::vector<int>> out;
    for (int i = 0; i < (int)sets.size(); ++i) {
        const auto& s = sets[i];
        if (s.empty()) continue;
        int best = s[0];
        for (int x : s) if (post[x].size() < post[best].size()) best = x;
        bool dominated = false;
        for (int j : post[best]) {
            if (j == i || sets[j].size() <= s.size()) continue;
            if (std::includes(sets[j].begin(), sets[j].end(), s.begin(), s.end())) { dominated = true; break; }
        }
        if (!dominated) out.push_back(s);
    }
    return out;  // already lexicographic
}

}  // namespace orc

// ===========================================================================
// C API for the tests and the CPU baseline.
extern "C" {

struct orc_stats {
    int64_t n_input, m_input, n_condensed, m_condensed, twohop_len;
    int64_t tasks, branches, raw_cliques, max_cliques, children;
    int64_t bytes;
    double t_load, t_twohop, t_level0, t_expand, t_rmnonmax;
};

static double now_s() { return omp_get_wtime(); }

// Whole CPU path: load, two-hop, tables, level 0 + condense, expansion to completion
// (root with host level-0 semantics, its first child with host level-1 semantics,
// everything else with device semantics), then the maximal-set filter. Writes the raw
// clique file (if raw_path) and the maximal file (if max_path; not created when there
// are no cliques, as the reference does). Returns the maximal clique count or <0.
// time_cap_s > 0 stops the expansion early (throughput sample; result incomplete).
int64_t orc_run(const char* graph_path, double gout, double gin, int min_size, int threads,
                int host_addfail_check, const char* raw_path, const char* max_path,
                double time_cap_s, orc_stats* st) {
    using namespace orc;
    if (threads < 1) threads = omp_get_max_threads();
    orc_stats S;
    std::memset(&S, 0, sizeof S);
    double t0 = now_s();
    Graph g;
    int rc = load_text(graph_path, g);
    if (rc != 0) return rc == -1 ? -1 : -2;
    double t1 = now_s();
    S.n_input = g.n;
    S.m_input = g.m;
    if (getenv("ORC_VERBOSE")) std::fprintf(stderr, "oracle: load %.2fs n=%d m=%lld\n", t1 - t0, g.n, (long long)g.m);
    int len = 2 * (std::max(g.n, min_size) + 2) + 2;
    std::vector<int> mdo = mindeg_table(gout, len), mdi = mindeg_table(gin, len);
    build_twohop_for_level0(g, threads, mdo[min_size], mdi[min_size]);
    double t2 = now_s();
    if (getenv("ORC_VERBOSE")) std::fprintf(stderr, "oracle: twohop %.2fs len=%zu\n", t2 - t1, g.tn.size());
    Params P{mdo.data(), mdi.data(), gout, gin, min_size, host_addfail_check != 0};
    std::vector<Rec> root;
    std::vector<int> orig_of;
    level0(g, P, root, orig_of);
    double t3 = now_s();
    S.n_condensed = g.n;
    S.m_condensed = g.m;
    S.twohop_len = (int64_t)g.tn.size();
    if (getenv("ORC_VERBOSE")) std::fprintf(stderr, "oracle: level0 %.2fs k=%d\n", t3 - t2, g.n);

    std::vector<Sink> sinks(threads);
    std::atomic<bool> capped(false);
    double tcap_end = time_cap_s > 0 ? now_s() + time_cap_s : 0;
    std::vector<std::vector<Rec>> frontier;
    if (!root.empty()) {
        Expander ex(g, P);
        ex.sink = &sinks[0];
        ex.expand(root.data(), (int)root.size(), HOST_L0);
        std::vector<std::vector<Rec>> l0 = std::move(sinks[0].kids);
        sinks[0].kids.clear();
        if (!l0.empty()) {
            ex.expand(l0[0].data(), (int)l0[0].size(), HOST_L1);
            frontier.assign(std::make_move_iterator(l0.begin() + 1), std::make_move_iterator(l0.end()));
            for (auto& k : sinks[0].kids) frontier.push_back(std::move(k));
            sinks[0].kids.clear();
        }
    }
    if (getenv("ORC_VERBOSE")) std::fprintf(stderr, "oracle: host levels %.2fs frontier=%zu\n", now_s() - t3, frontier.size());
    // device-level tasks: depth-first per OpenMP task, dynamic balancing
    {
        std::vector<Expander*> exs(threads);
        for (int t = 0; t < threads; ++t) { exs[t] = new Expander(g, P); exs[t]->sink = &sinks[t]; }
        std::function<void(std::vector<Rec>)> run;
        std::atomic<int64_t> spawned((int64_t)frontier.size()), done(0);
        const bool verbose = getenv("ORC_VERBOSE") != nullptr;
        double next_report = now_s() + 10;
        run = [&](std::vector<Rec> task) {
            if (tcap_end > 0 && now_s() > tcap_end) { capped = true; return; }
            if (verbose && omp_get_thread_num() == 0 && now_s() > next_report) {
                next_report = now_s() + 10;
                std::fprintf(stderr, "oracle: t=%.0fs done=%lld pending=%lld\n", now_s() - t3,
                             (long long)done.load(), (long long)(spawned.load() - done.load()));
            }
            done++;
            int t = omp_get_thread_num();
            Expander* ex = exs[t];
            size_t before = ex->sink->kids.size();
            ex->expand(task.data(), (int)task.size(), DEVICE);
            std::vector<std::vector<Rec>> mine(std::make_move_iterator(ex->sink->kids.begin() + before),
                                               std::make_move_iterator(ex->sink->kids.end()));
            ex->sink->kids.resize(before);
            spawned += (int64_t)mine.size();
            for (auto& k : mine) {
#pragma omp task firstprivate(k) shared(run)
                run(std::move(k));
            }
        };
#pragma omp parallel num_threads(threads)
#pragma omp single
        {
            for (auto& k : frontier) {
#pragma omp task firstprivate(k) shared(run)
                run(std::move(k));
            }
        }
        for (int t = 0; t < threads; ++t) delete exs[t];
    }
    double t4 = now_s();
    std::vector<std::vector<int>> sets;
    for (auto& s : sinks) {
        S.tasks += s.st.tasks; S.branches += s.st.branches; S.raw_cliques += s.st.raw_cliques;
        S.children += s.st.children; S.bytes += s.st.bytes;
        size_t p = 0;
        while (p < s.clq.size()) {
            int l = s.clq[p];
            sets.emplace_back(s.clq.begin() + p + 1, s.clq.begin() + p + 1 + l);
            p += 1 + l;
        }
    }
    if (raw_path) {
        FILE* f = std::fopen(raw_path, "wb");
        if (!f) return -1;
        for (auto& s : sets) {
            std::fprintf(f, "%zu ", s.size());
            for (int x : s) std::fprintf(f, "%d ", x);
            std::fprintf(f, "\n");
        }
        std::fclose(f);
    }
    if (getenv("ORC_VERBOSE")) {
        uint64_t th[32] = {0}, kh[32] = {0};
        for (auto& s : sinks) for (int b = 0; b < 32; ++b) { th[b] += s.st.task_hist[b]; kh[b] += s.st.kid_hist[b]; }
        for (int b = 0; b < 32; ++b)
            if (th[b] || kh[b]) std::fprintf(stderr, "oracle: size [%d,%d): tasks %llu children %llu\n", 1 << b, 2 << b, (unsigned long long)th[b], (unsigned long long)kh[b]);
    }
    auto mx = maximal_sets(std::move(sets));
    S.max_cliques = (int64_t)mx.size();
    if (max_path && !mx.empty()) {
        FILE* f = std::fopen(max_path, "wb");
        if (!f) return -1;
        for (auto& s : mx) {
            std::fprintf(f, "%d ", (int)s.size());
            for (int x : s) std::fprintf(f, "%d ", x);
            std::fprintf(f, "\n");
        }
        std::fclose(f);
    }
    double t5 = now_s();
    S.t_load = t1 - t0; S.t_twohop = t2 - t1; S.t_level0 = t3 - t2;
    S.t_expand = t4 - t3; S.t_rmnonmax = t5 - t4;
    if (capped) S.max_cliques = -S.max_cliques - 1;  // flag: incomplete sample
    if (st) *st = S;
    return capped ? -3 : (int64_t)mx.size();
}

// Loader + two-hop only, exported as CSR for the unit tests. Arrays are malloc'd.
int orc_graph(const char* path, int with_twohop, int threads, int* n_out, int64_t* m_out,
              int64_t** oo, int** on, int64_t** io, int** in, int64_t** to, int** tn) {
    using namespace orc;
    Graph g;
    int rc = load_text(path, g);
    if (rc) return rc;
    if (with_twohop) build_twohop(g, threads < 1 ? 1 : threads);
    else g.to.assign(g.n + 1, 0);
    *n_out = g.n;
    *m_out = g.m;
    auto cp64 = [](const std::vector<uint64_t>& v) {
        int64_t* p = (int64_t*)std::malloc(sizeof(int64_t) * (v.size() + 1));
        for (size_t i = 0; i < v.size(); ++i) p[i] = (int64_t)v[i];
        return p;
    };
    auto cp32 = [](const std::vector<int>& v) {
        int* p = (int*)std::malloc(sizeof(int) * (v.size() + 1));
        if (!v.empty()) std::memcpy(p, v.data(), sizeof(int) * v.size());
        return p;
    };
    *oo = cp64(g.oo); *on = cp32(g.on); *io = cp64(g.io); *in = cp32(g.in);
    *to = cp64(g.to); *tn = cp32(g.tn);
    return 0;
}

// Restricted two-hop build (build_twohop with the degree flags out >= mo, in >= mi) as
// CSR plus each vertex's whole list length; arrays are malloc'd. Returns 0 or <0.
int orc_twohop_restricted(const char* path, int mo, int mi, int threads, int* n_out, int64_t** to, int** tn,
                          uint32_t** l2full) {
    using namespace orc;
    Graph g;
    int rc = load_text(path, g);
    if (rc) return rc;
    std::vector<uint8_t> flag((size_t)g.n);
    for (int v = 0; v < g.n; ++v)
        flag[v] = (g.oo[v + 1] - g.oo[v] >= (uint64_t)mo) && (g.io[v + 1] - g.io[v] >= (uint64_t)mi);
    build_twohop(g, threads < 1 ? 1 : threads, &flag);
    *n_out = g.n;
    *to = (int64_t*)std::malloc(sizeof(int64_t) * (g.to.size() + 1));
    for (size_t i = 0; i < g.to.size(); ++i) (*to)[i] = (int64_t)g.to[i];
    *tn = (int*)std::malloc(sizeof(int) * (g.tn.size() + 1));
    if (!g.tn.empty()) std::memcpy(*tn, g.tn.data(), sizeof(int) * g.tn.size());
    *l2full = (uint32_t*)std::malloc(sizeof(uint32_t) * (g.l2full.size() + 1));
    if (!g.l2full.empty()) std::memcpy(*l2full, g.l2full.data(), sizeof(uint32_t) * g.l2full.size());
    return 0;
}

// Level 0 + condense map: writes orig_of (input id of each condensed id) into a
// malloc'd array; returns the condensed count or <0.
int orc_condense_map(const char* path, double gout, double gin, int min_size, int threads, int** orig_of) {
    using namespace orc;
    Graph g;
    int rc = load_text(path, g);
    if (rc) return rc;
    int len = 2 * (std::max(g.n, min_size) + 2) + 2;
    std::vector<int> mdo = mindeg_table(gout, len), mdi = mindeg_table(gin, len);
    build_twohop_for_level0(g, threads < 1 ? 1 : threads, mdo[min_size], mdi[min_size]);
    Params P{mdo.data(), mdi.data(), gout, gin, min_size, true};
    std::vector<Rec> root;
    std::vector<int> oo;
    level0(g, P, root, oo);
    *orig_of = (int*)std::malloc(sizeof(int) * (oo.size() + 1));
    if (!oo.empty()) std::memcpy(*orig_of, oo.data(), sizeof(int) * oo.size());
    return (int)oo.size();
}

void orc_free(void* p) { std::free(p); }

// on != 0: later orc_run / orc_condense_map / orc_open build only the two-hop lists level 0
// can read (build_twohop_for_level0); the results are those of the full build.
void orc_set_twohop_restricted(int on) { orc::g_twohop_restricted = on != 0; }

int orc_mindeg(double gamma, int len, int* out) {
    auto md = orc::mindeg_table(gamma, len);
    std::memcpy(out, md.data(), sizeof(int) * len);
    return 0;
}

// Maximal-set filter on a raw clique file (the RemoveNonMax contract). Returns the
// number of sets written, 0 if the input has none (no output file then), <0 on error.
int orc_rmnonmax(const char* in_path, const char* out_path) {
    FILE* f = std::fopen(in_path, "rb");
    if (!f) return -1;
    std::vector<std::vector<int>> sets;
    int len;
    while (std::fscanf(f, "%d", &len) == 1) {
        std::vector<int> s(len);
        for (int i = 0; i < len; ++i)
            if (std::fscanf(f, "%d", &s[i]) != 1) { std::fclose(f); return -2; }
        sets.push_back(std::move(s));
    }
    std::fclose(f);
    auto mx = orc::maximal_sets(std::move(sets));
    if (mx.empty()) return 0;
    FILE* o = std::fopen(out_path, "wb");
    if (!o) return -1;
    for (auto& s : mx) {
        std::fprintf(o, "%d ", (int)s.size());
        for (int x : s) std::fprintf(o, "%d ", x);
        std::fprintf(o, "\n");
    }
    std::fclose(o);
    return (int)mx.size();
}

}  // extern "C"

// ===========================================================================
// Task-level API: one graph prepared once (load, two-hop, tables, level 0 + condense),
// then any task expanded one level with the chosen semantics. Records are 6 uint32
// {id, out_mem, out_can, in_mem, in_can, lvl2adj}; positions [0, nm) are members,
// [nm, nm + ncov) covered vertices (level 0 only), the rest candidates.
extern "C" {

struct orc_handle {
    orc::Graph g;
    std::vector<int> mdo, mdi;
    orc::Params P;
    std::vector<orc::Rec> root;
    int32_t ncov = 0;
};

void* orc_open(const char* path, double gout, double gin, int min_size, int threads, int host_addfail_check) {
    using namespace orc;
    auto* h = new orc_handle();
    if (load_text(path, h->g) != 0) { delete h; return nullptr; }
    int len = 2 * (std::max(h->g.n, min_size) + 2) + 2;
    h->mdo = mindeg_table(gout, len);
    h->mdi = mindeg_table(gin, len);
    build_twohop_for_level0(h->g, threads < 1 ? 1 : threads, h->mdo[min_size], h->mdi[min_size]);
    h->P = Params{h->mdo.data(), h->mdi.data(), gout, gin, min_size, host_addfail_check != 0};
    std::vector<int> orig;
    level0(h->g, h->P, h->root, orig);
    while (h->ncov < (int)h->root.size() && h->root[h->ncov].label == 2) ++h->ncov;
    return h;
}

void orc_close(void* hv) { delete (orc_handle*)hv; }

// Level-0 task: writes 6 * k uint32 when out != NULL; returns k; *ncov = covered count.
int orc_root(void* hv, uint32_t* out, int* ncov) {
    auto* h = (orc_handle*)hv;
    if (ncov) *ncov = h->ncov;
    if (out)
        for (size_t i = 0; i < h->root.size(); ++i) {
            const orc::Rec& r = h->root[i];
            uint32_t* o = out + 6 * i;
            o[0] = r.id; o[1] = r.om; o[2] = r.oc; o[3] = r.im; o[4] = r.ic; o[5] = r.l2;
        }
    return (int)h->root.size();
}

// Expands one task one level. Children: *kid_meta gets 2 int64 per child {record offset,
// size}, members counted by label; *kids the records; cliques as "len, ids..." int32.
// All outputs malloc'd (orc_free). Returns the number of children.
int orc_expand(void* hv, const uint32_t* rec, int nvert, int nm, int ncov, int mode, int max_steps, int first_step,
               uint32_t** kids,
               int64_t** kid_meta, int32_t** kid_nm, int32_t** clq, int64_t* clq_len, int64_t* branches) {
    using namespace orc;
    auto* h = (orc_handle*)hv;
    std::vector<Rec> t((size_t)nvert);
    for (int i = 0; i < nvert; ++i) {
        const uint32_t* r = rec + 6 * i;
        t[i].id = r[0]; t[i].om = r[1]; t[i].oc = r[2]; t[i].im = r[3]; t[i].ic = r[4]; t[i].l2 = r[5];
        t[i].label = i < nm ? 1 : (i < nm + ncov ? 2 : 0);
    }
    Sink sink;
    Expander ex(h->g, h->P);
    ex.sink = &sink;
    ex.expand(t.data(), nvert, (Mode)mode, max_steps, first_step);
    size_t tot = 0;
    for (auto& k : sink.kids) tot += k.size();
    *kids = (uint32_t*)std::malloc(sizeof(uint32_t) * 6 * (tot + 1));
    *kid_meta = (int64_t*)std::malloc(sizeof(int64_t) * 2 * (sink.kids.size() + 1));
    *kid_nm = (int32_t*)std::malloc(sizeof(int32_t) * (sink.kids.size() + 1));
    size_t p = 0;
    for (size_t c = 0; c < sink.kids.size(); ++c) {
        (*kid_meta)[2 * c] = (int64_t)p;
        (*kid_meta)[2 * c + 1] = (int64_t)sink.kids[c].size();
        int m = 0;
        while (m < (int)sink.kids[c].size() && sink.kids[c][m].label == 1) ++m;
        (*kid_nm)[c] = m;
        for (const Rec& r : sink.kids[c]) {
            uint32_t* o = *kids + 6 * p++;
            o[0] = r.id; o[1] = r.om; o[2] = r.oc; o[3] = r.im; o[4] = r.ic; o[5] = r.l2;
        }
    }
    *clq = (int32_t*)std::malloc(sizeof(int32_t) * (sink.clq.size() + 1));
    if (!sink.clq.empty()) std::memcpy(*clq, sink.clq.data(), sizeof(int32_t) * sink.clq.size());
    *clq_len = (int64_t)sink.clq.size();
    if (branches) *branches = (int64_t)sink.st.branches;
    return (int)sink.kids.size();
}

}  // extern "C"
