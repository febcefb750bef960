// MI355X (gfx950) expansion engine for maximal directed quasi-clique enumeration.
//
// One 256-thread workgroup owns one task at a time and runs its branch loop; the
// workgroups are persistent and pull tasks from the current batch with one atomic.
// All tasks live on one LIFO task stack in HBM: each launch expands a batch popped from
// the stack top, and children are pushed back onto the top with one combined 64-bit
// atomic per child (descriptor count and record count move together, so the records
// of consecutive descriptors are contiguous). Popping the newest tasks first keeps the
// stack close to depth-first order, which bounds its size the way the reference's LIFO
// buffer does (src/device_kernels.cu:300-326).
//
// Per workgroup state:
//   * LDS bitmaps over condensed vertex ids: parent-live, parent-member, child-member,
//     child-candidate, in-L(v), critical. Every "is this neighbour in the task" test
//     of the reference's vertex_order_map (src/device_kernels.cu:433-435) becomes one
//     LDS bit test;
//   * HBM arrays indexed by condensed id for the parent's and the child's degrees (one
//     16-byte record of four degrees per id for each), so scatters need no position
//     lookup;
//   * a bitonic sort of each child on a packed 64-bit enumeration-order key, in LDS for
//     children of up to SORT_N vertices and in the workgroup's HBM area above that (LDS
//     is kept small so that eight workgroups share a CU).
//
// Semantics follow the reference's per-task branch loop exactly (host level 0 / 1 or
// device levels, selected per task); the functions cite the reference lines they follow.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dqc.h"
#include "costmodel.hpp"
#include "graph.hpp"

namespace {

constexpr int NT = 256;            // threads per workgroup
constexpr int KW = 512;            // bitmap words: condensed ids < 16384
constexpr int KMAX = KW * 32;
#ifndef DQC_SORT_N
#define DQC_SORT_N 256
#endif
constexpr int SORT_N = DQC_SORT_N;  // LDS sort capacity (larger children sort in HBM)
constexpr int HMAX = 256;          // member-degree histogram bins (members < HMAX)
// Wavefront-per-task kernel: one 64-lane wavefront runs one task, with LDS id bitmaps of
// WAVE_KW words (condensed graphs of at most WAVE_KMAX vertices).
constexpr int WAVE_T = 64;
constexpr int WAVE_KW = 64;
constexpr int WAVE_KMAX = WAVE_KW * 32;
// Split passes (the removal and lookahead steps of a task cut into pieces, one step after
// another) run on workgroups of SPLIT_T threads: each step scans the removed vertex's
// lists, and four times the lanes cut the memory round trips of a long list fourfold.
constexpr int SPLIT_T = 1024;
#ifndef DQC_SCAN_UNR
#define DQC_SCAN_UNR 8
#endif
constexpr int SCAN_UNR = DQC_SCAN_UNR;  // list entries in flight per thread in list scans
constexpr int MODE_HOST_L0 = 0, MODE_HOST_L1 = 1, MODE_DEVICE = 2;
// Combined allocation counters: (count << 34) | size. Capacities are held below 2^30
// entries and 2^33 records / ids, so neither field can carry into the other even when
// a full buffer keeps receiving (failing) requests from every workgroup.
constexpr unsigned long long CNT_SHIFT = 34;
constexpr unsigned long long SIZE_MASK = (1ull << CNT_SHIFT) - 1;
constexpr unsigned long long MAX_ENTRIES = (1ull << 30) - 1;
constexpr unsigned long long MAX_RECORDS = 1ull << 33;

enum ErrBits : unsigned { ERR_SPILL = 1u, ERR_CLIQUES = 2u, ERR_MEMBERS = 4u, ERR_BATCH = 8u };

// flags of a task descriptor
constexpr int32_t TF_PIECE = 1;         // a piece of a split task: not counted, not split again
constexpr int32_t TF_REMOVE_FIRST = 2;  // the piece's first step starts with a removal
constexpr int32_t TF_SPLIT = 4;         // set by the batch scheduler: cut this task into pieces
constexpr int32_t TF_L1_DONE = 8;       // a level-0 piece whose task already wrote its first child

struct TaskDesc {
    unsigned long long off;
    int32_t tot, nm, ncov, mode;
    int32_t iters;  // loop steps to run (pieces); whole tasks run tot - nm - ncov
    int32_t flags;
    // identity and shape, for the task log and the cost model (pieces keep their task's)
    uint32_t uid, parent;
    uint16_t depth, gid;  // gid: graph of a batched run (0 for a single graph)
    uint16_t dmin_o, dmax_o, dmin_i, dmax_i;  // min / max of out / in degree inside the task
    int32_t step0;  // a piece: loop steps of its task before the piece's first (level-0 windows)
};
static_assert(sizeof(TaskDesc) == 56, "task descriptor layout");

// A task the schedulers may split: a whole task (not a piece) below level 0 with at
// least split_b loop steps.
__host__ __device__ inline bool split_eligible(const TaskDesc& d, int split_b) {
    const int steps = d.tot - d.nm - (d.mode == MODE_DEVICE ? 0 : d.ncov);
    return split_b > 0 && !(d.flags & TF_PIECE) && d.mode != MODE_HOST_L0 && steps >= split_b;
}
// Pieces for a split task of estimated cost c in a batch of total estimated cost C: the
// batch's work cut into about four pieces per workgroup of similar cost (C / (4 x
// workgroups) each), so a task gets pieces in proportion to its cost; at least 2, at most
// max_pieces.
__host__ __device__ inline int pieces_for(double c, double total, int workgroups, int max_pieces) {
    const double target = total / (4.0 * (double)(workgroups > 0 ? workgroups : 1));
    double p = target > 0.0 ? ceil(c / target) : (double)max_pieces;
    p = fmin(fmax(p, 2.0), (double)(max_pieces > 2 ? max_pieces : 2));
    return (int)p;
}
// Steps per piece when a task of `steps` loop steps is split: as many pieces as
// max_pieces allows, each of at least split_c steps (host and device schedulers both
// use this; the chunk travels in the descriptor's iters field of a TF_SPLIT task).
__host__ __device__ inline int split_chunk_for(int steps, int split_c, int max_pieces) {
    const int c = split_c > 0 ? split_c : 1;
    int pieces = (steps + c - 1) / c;
    if (max_pieces > 0 && pieces > max_pieces) pieces = max_pieces;
    if (pieces < 1) pieces = 1;
    return (steps + pieces - 1) / pieces;
}

// One task-log record per expanded task or piece (the dqc_task_log of the C-ABI).
struct TaskLog {
    uint32_t uid, parent;
    int32_t depth, nm, tot;
    uint16_t dmin_o, dmax_o, dmin_i, dmax_i;
    uint32_t branches, flags;
};
static_assert(sizeof(TaskLog) == 36, "task log layout");

// Per-queue counters (one per task stack / stream).
struct QCtr {
    unsigned long long next_task;  // next work item of the current launch
    unsigned long long sp_alloc;   // task stack top: (tasks << 34) | records
    unsigned long long root_first; // level-0 split: (branch << 32) | stack index of the first child
    unsigned long long ntask;      // device-resident steps: tasks in the current batch
    unsigned long long first;      //   its first stack index
    unsigned long long v0;         //   and its first record
    unsigned long long maxtot;     //   and its largest task (the stack reserve is sized by it)
    unsigned long long pend;       // batch tasks not yet finished: (tasks << 34) | records
    unsigned long long wg_limit;   // workgroups the stack has room for (device-resident steps)
    unsigned long long def_alloc;  // deferred region, growing down from the top of the stack
                                   // arrays: (pieces << 34) | records
    unsigned long long used;       // main + deferred: (tasks << 34) | records. Every push adds
                                   // to it first and tests the capacity on the value it returns,
                                   // so two regions can never be filled past each other
    unsigned long long split_budget;  // device-resident steps: piece records the order kernel
                                      // may mark for splitting in the current batch
};
// Counters shared by every launch of a handle.
struct GCtr {
    unsigned long long clq_alloc;  // clique store: (sets << 40) | ids
    unsigned long long tasks, branches, cliques, bytes, slow_sorts;
    unsigned long long phase[8];   // realtime ticks per phase (profiling runs)
    unsigned int error;
    unsigned int pad;
    unsigned long long next_uid;   // task ids handed out (the level-0 task is id 0)
    unsigned long long log_n;      // task-log records requested
    unsigned long long steps;      // device-resident steps that expanded a batch
    unsigned long long truncated;  // children not pushed: deeper than max_depth
    unsigned long long spill_state[4];  // first failed stack write: use, request, launch state, pending
    unsigned long long pushbacks;  // tasks whose rest went back to the stack for lack of room
    unsigned long long slices;     // tasks whose rest went back to the stack at a time-slice end
};

struct KParams {
    int32_t n;
    const unsigned long long *oo, *io, *to;
    const int32_t *on, *in, *tn;
    const int32_t *mdo, *mdi, *ubo, *ubi;
    int32_t min_size, host_addfail, profile;
    const TaskDesc* in_desc;
    const dqc::TaskVertex* in_vtx;
    unsigned long long in_ntask;  // ~0ull: the batch size is q->ntask (device-resident steps)
    TaskDesc* sp_desc;
    dqc::TaskVertex* sp_vtx;
    unsigned long long sp_cap_t, sp_cap_v;
    int32_t* clq_ids;
    unsigned long long* clq_meta;  // (start << 24) | len per set
    unsigned long long clq_cap_s, clq_cap_i;
    QCtr* q;
    GCtr* g;
    // level-0 split mode: work item t expands branch split_s[t] of the level-0 task
    const int32_t* split_s;
    const unsigned long long *lto;  // transposed two-hop CSR (who lists u)
    const int32_t* ltn;
    uint32_t* scratch;
    unsigned long long stride;  // uint32 words per workgroup
    TaskLog* log;               // task log (null = off)
    unsigned long long log_cap;
    int32_t kloc;               // largest vertex count of one graph (scratch arrays, bitmaps)
    int32_t sum_shift;          // least log2 ids per live-summary bit (HBM-bitmap variant)
    uint32_t max_depth;         // children deeper than this are counted, not pushed (0 = no bound)
    int32_t root_steps;         // > 0: a level-0 task runs only its loop steps s < root_steps
    int32_t root_first;         // children of level-0 steps s < root_first are counted, not pushed
    unsigned long long slice_ticks;  // > 0: time slice per launch in realtime-counter ticks (100 MHz)
    unsigned long long maxtot;  // largest task of the batch (~0ull: q->maxtot, device-resident steps)
    const int32_t* gbase;       // batched graphs: first id of each graph (multiples of 32); null = one graph
};

// Id bitmaps (parent-live, parent-member, child-member, child-candidate, in-L(v),
// critical): in LDS when the condensed graph has at most KMAX vertices, otherwise in the
// workgroup's HBM area. Either way they are all-zero between tasks: each branch and
// task clears the words of the vertices it set.
constexpr int NBITMAPS = 6;
// LDS words of the live-set summary of the HBM-bitmap variant (it takes the LDS the
// other variant gives its bitmaps): one bit per id up to 98,304 ids, per 2^s ids above
constexpr int SUMW = 3072;
struct Shared {
    unsigned long long sortk[SORT_N];
    int32_t ho[HMAX + 2], hi[HMAX + 2], so[HMAX + 2], si[HMAX + 2];
    int32_t cnt[12];
    int32_t red[12];
    int32_t lu[12];
    unsigned long long bcast;
    unsigned long long acc_bytes;
    unsigned long long clip[4];  // list ranges narrowed by clip_live (two pairs, alternating)
    int32_t idr[2];              // min / max live id of the task being loaded
};


__device__ __forceinline__ int bitw(unsigned long long x) { return x == 0 ? 0 : 64 - __clzll(x); }

// Number of entries < x in the sorted list a[lo, hi).
__device__ __forceinline__ uint32_t cnt_lt(const int32_t* a, unsigned long long lo, unsigned long long hi, int x) {
    unsigned long long l = lo, h = hi;
    while (l < h) {
        unsigned long long m = (l + h) >> 1;
        if (a[m] < x) l = m + 1; else h = m;
    }
    return (uint32_t)(l - lo);
}
// First index i of the sorted list a[b, e) with a[i] >= x (e if none), found by the 64
// lanes of one wavefront together: each round every lane tests one of 64 evenly spaced
// entries, which narrows the range 64-fold, so a list of 2^18 entries takes three
// dependent loads. Every lane of the wavefront calls it with the same arguments.
__device__ __forceinline__ unsigned long long wave_lower_bound(const int32_t* a, unsigned long long b,
                                                               unsigned long long e, int x, int lane) {
    while (true) {  // the answer lies in [b, e]
        const unsigned long long n = e - b;
        if (n == 0) return b;
        const unsigned long long step = (n + 63) / 64;
        const unsigned long long i = b + (unsigned long long)lane * step;
        const bool ge = i >= e || a[i] >= x;
        const unsigned long long m = __ballot(ge);
        const int f = m ? __builtin_ctzll(m) : 64;  // first sampled entry >= x
        if (step == 1 || f == 0) return b + (unsigned long long)f;
        const unsigned long long nb = b + (unsigned long long)(f - 1) * step + 1;
        e = f == 64 ? e : min(e, b + (unsigned long long)f * step);
        b = nb;
    }
}
__device__ __forceinline__ bool contains(const int32_t* a, unsigned long long lo, unsigned long long hi, int x) {
    unsigned long long l = lo + cnt_lt(a, lo, hi, x);
    return l < hi && a[l] == x;
}

// One field of the id-indexed degree records. A vertex's four degrees (out_mem,
// out_can, in_mem, in_can) share one 16-byte slot, so the accesses that read or write
// all four (copies into the child, validity tests, child records) touch one cache line
// per vertex instead of four.
struct Field {
    uint32_t* p;
    __device__ __forceinline__ uint32_t& operator[](long long x) const { return p[4 * x]; }
};

// A few zero words in device memory: the branch-free list scans load a bitmap word
// from here instead when the live-set summary rules an id out.
__device__ uint32_t zero_words[16];

// The kernel's parameter block, read in place from the kernarg segment (constant
// address space: scalar loads). Taking the address of the by-value kernel argument
// instead would copy the whole block into per-lane scratch memory.
typedef const __attribute__((address_space(4))) KParams* KPtr;

// All functions below are executed by every thread of the workgroup with uniform
// control flow; `W` bundles the per-workgroup pointers and uniform scalars. NT threads
// run one task: a 256-thread workgroup, or one wavefront (64) for small tasks.
template <int NT>
struct W {
    KPtr P;
    Shared* S;
    int tid;
    int kw;  // bitmap words in use
    uint32_t *plive, *pmem, *cmem, *ccand, *cL, *ccrit;  // id bitmaps (LDS or HBM)
    // Batched graphs share one id space; a task's ids all lie in its graph's range
    // [gb, gb + kloc), and the id-indexed arrays and bitmaps are addressed relative to gb.
    uint32_t* vbase;   // the workgroup's id-indexed value arrays (graph-relative)
    int gb;            // first id of the current task's graph (a multiple of 32)
    uint32_t cur_gid;
    // HBM-bitmap variant: a summary of the parent-live bitmap in LDS, one bit per 2^gsh
    // ids, set with the live bits and cleared with them at the end of the task (removals
    // leave it set). Every id in any of the six bitmaps is live in the parent, so an id
    // whose summary bit is clear is in none of them and bt answers from LDS without the
    // HBM read: the common case, since few neighbours of a vertex are in the task.
    uint32_t* lsum;  // null in the LDS-bitmap variant
    int gsh;
    // Every id live in the current task lies in [id_lo, id_hi) (the range of its ids when
    // it was loaded): the entries of a sorted list outside it are in no bitmap, so long
    // list scans skip them (clip_live).
    int id_lo, id_hi;
    int clip_pair;
    // v's out- and in-lists narrowed to [id_lo, id_hi) (branch setup)
    unsigned long long vob, voe, vib, vie;
    const uint32_t* zw;  // zero_words
    // Narrows the sorted list range [b, e) to its entries in [id_lo, id_hi); a list shorter
    // than 8 entries per thread is left whole. All threads call it with the same range.
    __device__ __forceinline__ void clip_live(const int32_t* a, unsigned long long& b, unsigned long long& e) {
        if (e - b < 8ull * NT || (id_lo == INT_MIN && id_hi == INT_MAX)) return;
        const int wave = tid >> 6, lane = tid & 63;
        unsigned long long* r = S->clip + clip_pair;
        if (wave == 0) {
            const unsigned long long x = wave_lower_bound(a, b, e, id_lo, lane);
            if (lane == 0) r[0] = x;
        }
        if (wave == (NT > 64 ? 1 : 0)) {
            const unsigned long long x = wave_lower_bound(a, b, e, id_hi, lane);
            if (lane == 0) r[1] = x;
        }
        __syncthreads();
        // (the other pair is written next: a thread that writes it has passed this barrier
        // again only after every thread read this pair)
        b = r[0];
        e = r[1] > b ? r[1] : b;
        clip_pair ^= 2;
    }
    __device__ __forceinline__ bool maybe_live(int id) const {
        const int g = (id - gb) >> gsh;
        return (lsum[g >> 5] >> (g & 31)) & 1u;
    }
    __device__ __forceinline__ void mark_live(int id) const {
        const int g = (id - gb) >> gsh;
        atomicOr(&lsum[g >> 5], 1u << (g & 31));
    }
    // graph-relative bit tests on the id bitmaps (gb is a multiple of 32)
    __device__ __forceinline__ bool bt(const uint32_t* bm, int id) const {
        if (lsum && !maybe_live(id)) return false;
        return (bm[(id - gb) >> 5] >> (id & 31)) & 1u;
    }
    __device__ __forceinline__ void bs(uint32_t* bm, int id) const { atomicOr(&bm[(id - gb) >> 5], 1u << (id & 31)); }
    __device__ __forceinline__ void bcl(uint32_t* bm, int id) const { atomicAnd(&bm[(id - gb) >> 5], ~(1u << (id & 31))); }
    // points the id-indexed arrays at graph `gid` of a batched run
    __device__ __forceinline__ void set_graph(uint32_t gid) {
        cur_gid = gid;
        gb = P->gbase ? P->gbase[gid] : 0;
        const long long k = P->kloc;
        uint32_t* pr = vbase - 4ll * gb;          // parent degree records
        uint32_t* cr = vbase + 4 * k - 4ll * gb;  // child degree records
        pom = Field{pr}; poc = Field{pr + 1}; pim = Field{pr + 2}; pic = Field{pr + 3};
        com = Field{cr}; coc = Field{cr + 1}; cim = Field{cr + 2}; cic = Field{cr + 3};
        pl2 = vbase + 8 * k - gb;
        adjc = vbase + 9 * k - gb;
    }
    int dirty_m, dirty_l;  // child bitmaps: members mem[0..dirty_m) and lst[0..dirty_l) are set
    bool crit_marks;       // critical() ran: members and ccrit bits among lst too
    bool cand_marks;       // ccand bits were set among lst
    // lvl2adj decrements of removals are applied when a value is next read: the removals
    // of positions [ptot, l2_tot) are pending (ptot: the parent's live count)
    int ptot, l2_tot;
    // Branch setup leaves the child degrees of the L(v) candidates implicit: the parent's,
    // with v's arcs moved from candidate to member (cL marks v's live in-neighbours, ccrit
    // its live out-neighbours). They are written only for a branch that passes its first
    // bound test (materialize). lazy_v = v while they are implicit, else -1.
    int lazy_v;
    __device__ __forceinline__ uint32_t in_v(int w) const { return bt(cL, w) ? 1u : 0u; }    // arc w -> v
    __device__ __forceinline__ uint32_t out_v(int w) const { return bt(ccrit, w) ? 1u : 0u; } // arc v -> w
    // writes the child degrees of the candidates R[0..n) (implicit until now) and sets their
    // candidate bits (the first bound test reads neither)
    __device__ __forceinline__ void materialize(const int32_t* R, int n) {
        for (int i = tid; i < n; i += NT) {
            const int w = R[i];
            const uint32_t a = in_v(w), b = out_v(w);
            com[w] = pom[w] + a; coc[w] = poc[w] - a; cim[w] = pim[w] + b; cic[w] = pic[w] - b;
            bs(ccand, w);
        }
        cand_marks = true;
        __syncthreads();
    }
    // clears the marks of v's neighbours once the implicit degrees are settled: until
    // critical() runs, every bit of cL and ccrit is such a mark, so the words of v's
    // neighbours are zeroed whole (plain stores, no read)
    __device__ __forceinline__ void clear_v_marks() {
        const long long n1 = (long long)(voe - vob), n = n1 + (long long)(vie - vib);
        for (long long k = tid; k < n; k += NT) {
            const bool in_list = k >= n1;
            const int w = in_list ? P->in[vib + (k - n1)] : P->on[vob + k];
            if (lsum && !maybe_live(w)) continue;  // (only live ids were marked)
            (in_list ? cL : ccrit)[(w - gb) >> 5] = 0u;
        }
        __syncthreads();
    }
    uint32_t cur_uid, cur_depth;  // the task being expanded (parent of the children written)
    uint32_t nbranch;             // branches run for the current task / piece
    uint32_t launch_steps;        // stack checks this workgroup passed in this launch
    unsigned long long t_start;   // thread 0: realtime counter at the launch's start (time slices)
    bool discard_child;           // the current level-0 branch lies before root_first: it only
                                  // tells whether it writes a child (no child, clique or count)
    Field pom, poc, pim, pic, com, coc, cim, cic;  // parent and child degrees (4 per id)
    uint32_t *pl2, *adjc;
    int32_t *pids, *mem, *la, *lb, *lc, *lst, *crit, *gat;
    unsigned long long* hsort;  // HBM sort keys for children above SORT_N (2k entries)
    unsigned long long bytes;  // per-thread traffic tally (CSR + records)
    unsigned long long ph[8];  // thread 0: ticks per phase
    unsigned long long ph_t;   // thread 0: start of the current phase
    __device__ __forceinline__ void tick(int k) {
        if (P->profile && tid == 0) {
            unsigned long long now = __builtin_amdgcn_s_memrealtime();
            ph[k] += now - ph_t;
            ph_t = now;
        }
    }

    __device__ __forceinline__ int gmd(int k, const int32_t* md) const { return k < P->min_size ? md[P->min_size] : md[k]; }

    // one broadcast of a value computed by thread 0 (two barriers: write, then read)
    __device__ __forceinline__ unsigned long long bcast0(unsigned long long v) {
        if (tid == 0) S->bcast = v;
        __syncthreads();
        unsigned long long r = S->bcast;
        __syncthreads();
        return r;
    }
    __device__ __forceinline__ int read_counter(int slot) {
        __syncthreads();
        int v = S->cnt[slot];
        __syncthreads();
        return v;
    }

    // Bitmap word holding id's bit, or 0 when the live-set summary rules the id out (the
    // HBM-bitmap variant); callers issue a whole batch of these before testing any, so a
    // batch of list entries costs one memory round trip for its bitmap words.
    __device__ __forceinline__ uint32_t bword(const uint32_t* bm, int id) const {
        if (!bm) return 0u;
        const bool m = !lsum || maybe_live(id);
        return m ? bm[(id - gb) >> 5] : 0u;
    }
    __device__ __forceinline__ static uint32_t bbit(uint32_t word, int id) { return (word >> (id & 31)) & 1u; }
    // Address of id's bitmap word, or of a zero word when the live-set summary rules the id
    // out: every lane issues its load, so a batch of them is counted exactly by the load
    // counters and the next batch of list entries can stay in flight meanwhile.
    __device__ __forceinline__ const uint32_t* bword_ptr(const uint32_t* bm, int id) const {
        const uint32_t* p = bm + ((id - gb) >> 5);
        return (!lsum || maybe_live(id)) ? p : zw;
    }

    // Appends to out[] (counted in S->cnt[slot], which the caller zeroed before a barrier)
    // every entry x of the list a[b, e) whose bmA bit is set and bmB bit clear, x != skip.
    // Software-pipelined and branch-free: a round's bitmap words are requested, then the
    // next round's list entries, then the words are tested, so the next round's entries
    // are in flight meanwhile; a wavefront appends its hits with one LDS atomic per
    // round and unrolled step. The order of out[] is not defined.
    __device__ __forceinline__ void wg_collect(const int32_t* __restrict__ a, unsigned long long b,
                                               unsigned long long e, const uint32_t* bmA, const uint32_t* bmB,
                                               int skip, int32_t* out, int slot) const {
        if (e <= b) return;
        const int32_t* ap = a + b;
        const int n = (int)(e - b);
        const int last = n - 1;
        constexpr int U = SCAN_UNR;
        const int lane = tid & 63;
        int k = tid;
        int v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = ap[min(k + u * NT, last)];
        while (k < n) {
            uint32_t wa[U], wb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                wa[u] = *bword_ptr(bmA, v[u]);
                wb[u] = *bword_ptr(bmB, v[u]);
            }
            const int k2 = k + U * NT;
            int v2[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v2[u] = ap[min(k2 + u * NT, last)];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int x = v[u];
                const bool hit = k + u * NT < n && bbit(wa[u], x) && !bbit(wb[u], x) && x != skip;
                const unsigned long long m = __ballot(hit);
                if (m) {
                    int base = 0;
                    if (lane == __builtin_ctzll(m)) base = atomicAdd(&S->cnt[slot], __popcll(m));
                    base = __shfl(base, __builtin_ctzll(m), 64);
                    if (hit) out[base + __popcll(m & ((1ull << lane) - 1ull))] = x;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = v2[u];
            k = k2;
        }
    }

    // As wg_scan_bits over the two lists a1[b1, e1) and a2[b2, e2) as one range (one round of
    // loads for both when they are short): f(id, bit in bmA, bit in bmB, entry of the second).
    template <class F>
    __device__ __forceinline__ void wg_scan_bits2(const int32_t* __restrict__ a1, unsigned long long b1,
                                                  unsigned long long e1, const int32_t* __restrict__ a2,
                                                  unsigned long long b2, unsigned long long e2,
                                                  const uint32_t* bmA, const uint32_t* bmB, F&& f) const {
        const long long n1 = (long long)(e1 - b1), n = n1 + (long long)(e2 - b2);
        for (long long k = tid; k < n; k += (long long)SCAN_UNR * NT) {
            int v[SCAN_UNR];
#pragma unroll
            for (int u = 0; u < SCAN_UNR; ++u) {
                const long long j = k + (long long)u * NT;
                v[u] = j < n1 ? a1[b1 + j] : (j < n ? a2[b2 + (j - n1)] : -1);
            }
            uint32_t wa[SCAN_UNR], wb[SCAN_UNR];
#pragma unroll
            for (int u = 0; u < SCAN_UNR; ++u) {
                wa[u] = v[u] >= 0 ? bword(bmA, v[u]) : 0u;
                wb[u] = v[u] >= 0 ? bword(bmB, v[u]) : 0u;
            }
#pragma unroll
            for (int u = 0; u < SCAN_UNR; ++u)
                if (v[u] >= 0) f(v[u], bbit(wa[u], v[u]), bbit(wb[u], v[u]), k + (long long)u * NT >= n1);
        }
    }

    // Visits every entry of the list a[b, e) once, spread over the workgroup, as
    // f(id, bit in bmA, bit in bmB) (bmB may be null: bit 0). Each thread issues SCAN_UNR
    // independent list loads, then the bitmap words of all of them, before it uses any,
    // so a long list costs one memory round trip per SCAN_UNR * NT entries for the
    // entries and one for their bitmap words, instead of one per NT entries and per test.
    // f must not change bmA or bmB.
    template <class F>
    __device__ __forceinline__ void wg_scan_bits(const int32_t* __restrict__ a, unsigned long long b,
                                                 unsigned long long e, const uint32_t* bmA, const uint32_t* bmB,
                                                 F&& f) const {
        unsigned long long i = b + (unsigned long long)tid;
        for (; i + (SCAN_UNR - 1) * NT < e; i += SCAN_UNR * NT) {
            int v[SCAN_UNR];
            uint32_t wa[SCAN_UNR], wb[SCAN_UNR];
#pragma unroll
            for (int u = 0; u < SCAN_UNR; ++u) v[u] = a[i + (unsigned long long)u * NT];
#pragma unroll
            for (int u = 0; u < SCAN_UNR; ++u) { wa[u] = bword(bmA, v[u]); wb[u] = bword(bmB, v[u]); }
#pragma unroll
            for (int u = 0; u < SCAN_UNR; ++u) f(v[u], bbit(wa[u], v[u]), bbit(wb[u], v[u]));
        }
        int v[SCAN_UNR];
        uint32_t wa[SCAN_UNR], wb[SCAN_UNR];
#pragma unroll
        for (int u = 0; u < SCAN_UNR; ++u) v[u] = i + (unsigned long long)u * NT < e ? a[i + (unsigned long long)u * NT] : -1;
#pragma unroll
        for (int u = 0; u < SCAN_UNR; ++u) {
            wa[u] = v[u] >= 0 ? bword(bmA, v[u]) : 0u;
            wb[u] = v[u] >= 0 ? bword(bmB, v[u]) : 0u;
        }
#pragma unroll
        for (int u = 0; u < SCAN_UNR; ++u)
            if (v[u] >= 0) f(v[u], bbit(wa[u], v[u]), bbit(wb[u], v[u]));
    }

    // Visits a[b, e) with the calling thread alone, 8 loads in flight.
    template <class F>
    __device__ __forceinline__ void thread_scan(const int32_t* __restrict__ a, unsigned long long b,
                                                unsigned long long e, F&& f) const {
        for (; b + 8 <= e; b += 8) {
            int v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = a[b + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) f(v[u]);
        }
        for (; b < e; ++b) f(a[b]);
    }
    // As thread_scan, with the bits of two bitmaps fetched per batch of 8 (as wg_scan_bits).
    template <class F>
    __device__ __forceinline__ void thread_scan_bits(const int32_t* __restrict__ a, unsigned long long b,
                                                     unsigned long long e, const uint32_t* bmA, const uint32_t* bmB,
                                                     F&& f) const {
        for (; b + 8 <= e; b += 8) {
            int v[8];
            uint32_t wa[8], wb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = a[b + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) { wa[u] = bword(bmA, v[u]); wb[u] = bword(bmB, v[u]); }
#pragma unroll
            for (int u = 0; u < 8; ++u) f(v[u], bbit(wa[u], v[u]), bbit(wb[u], v[u]));
        }
        for (; b < e; ++b) {
            const int w = a[b];
            f(w, bbit(bword(bmA, w), w), bbit(bword(bmB, w), w));
        }
    }

    // ---------------------------------------------------------------- predicates
    // reference inc/device_kernels.hpp:111-140 (d_cand_isvalid); unsigned compares kept
    __device__ __forceinline__ bool cand_valid(int w, int cs, int ub, int lb, int meo, int mei) const {
        const int32_t *mdo = P->mdo, *mdi = P->mdi;
        uint32_t om = com[w], oc = coc[w], im = cim[w], ic = cic[w];
        if (om + oc < (uint32_t)meo) return false;
        if (im + ic < (uint32_t)mei) return false;
        if (om + oc < (uint32_t)gmd((int)(cs + oc + 1), mdo)) return false;
        if (im + ic < (uint32_t)gmd((int)(cs + ic + 1), mdi)) return false;
        if (om + (uint32_t)ub - 1u < (uint32_t)mdo[cs + ub]) return false;
        if (im + (uint32_t)ub - 1u < (uint32_t)mdi[cs + ub]) return false;
        if (om + oc < (uint32_t)gmd(cs + lb, mdo)) return false;
        if (im + ic < (uint32_t)gmd(cs + lb, mdi)) return false;
        return true;
    }
    // reference inc/device_kernels.hpp:141-178 (d_vert_isextendable)
    __device__ __forceinline__ bool vert_ext(int w, int cs, int ub, int lb, int meo, int mei) const {
        const int32_t *mdo = P->mdo, *mdi = P->mdi;
        uint32_t om = com[w], oc = coc[w], im = cim[w], ic = cic[w];
        if (om + oc < (uint32_t)meo) return false;
        if (im + ic < (uint32_t)mei) return false;
        if (om + oc < (uint32_t)gmd((int)(cs + oc), mdo)) return false;
        if (im + ic < (uint32_t)gmd((int)(cs + ic), mdi)) return false;
        if (oc == 0 && om < (uint32_t)gmd(cs + 1, mdo)) return false;
        if (ic == 0 && im < (uint32_t)gmd(cs + 1, mdi)) return false;
        if (om + (uint32_t)ub < (uint32_t)mdo[cs + ub]) return false;
        if (im + (uint32_t)ub < (uint32_t)mdi[cs + ub]) return false;
        if (om + oc < (uint32_t)gmd(cs + lb, mdo)) return false;
        if (im + ic < (uint32_t)gmd(cs + lb, mdi)) return false;
        return true;
    }

    // ---------------------------------------------------------------- LU bounds
    // Sum of the i largest candidate member degrees, from the value tables built by
    // lu_tables (start index and sum-above per value, values 0..cm).
    __device__ __forceinline__ int psum(const int32_t* st, const int32_t* sa, int cm, int i) const {
        int lo = 0, hi = cm;  // smallest x with st[x] <= i
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (st[mid] <= i) hi = mid; else lo = mid + 1;
        }
        return sa[lo] + (i - st[lo]) * lo;
    }

    // smallest i in [a, b] with pred(i), else b + 1 (a > b gives a). Each round covers
    // 8 * NT indices; a thread stops at its own first hit (its indices ascend).
    template <class F>
    __device__ __forceinline__ int first_true(int a, int b, F pred) {
        for (int base = a; base <= b; base += 8 * NT) {
            int hit = INT_MAX;
            for (int u = 0; u < 8; ++u) {
                int i = base + u * NT + tid;
                if (i > b) break;
                if (pred(i)) { hit = i; break; }
            }
            if (tid == 0) S->red[0] = INT_MAX;
            __syncthreads();
            if (hit != INT_MAX) atomicMin(&S->red[0], hit);
            __syncthreads();
            int r = S->red[0];
            __syncthreads();
            if (r != INT_MAX) return r;
        }
        return a > b ? a : b + 1;
    }
    // largest i in [a, b] with pred(i), else a - 1 (rounds of 8 * NT from the top)
    template <class F>
    __device__ __forceinline__ int last_true(int a, int b, F pred) {
        for (int top = b; top >= a; top -= 8 * NT) {
            int hit = INT_MIN;
            for (int u = 0; u < 8; ++u) {
                int i = top - u * NT - tid;
                if (i < a) break;
                if (pred(i)) { hit = i; break; }
            }
            if (tid == 0) S->red[1] = INT_MIN;
            __syncthreads();
            if (hit != INT_MIN) atomicMax(&S->red[1], hit);
            __syncthreads();
            int r = S->red[1];
            __syncthreads();
            if (r != INT_MIN) return r;
        }
        return a - 1;
    }

    // reference src/device_kernels.cu:1307-1486 (d_calculate_LU_bounds), with the
    // member scan on thread 0 and the candidate-prefix loops as parallel searches over
    // the sorted candidate member degrees (held as histograms in S->ho / S->hi).
    __device__ __forceinline__ bool lu_bounds(int cm, int rc, int& ub, int& lb, int& meo, int& mei) {
        const int32_t *mdo = P->mdo, *mdi = P->mdi;
        if (tid == 0) {
            int x = mem[0];
            int mino = (int)com[x], mino_c = (int)coc[x], sumo = (int)com[x], mint_o = (int)(com[x] + coc[x]);
            int mini = (int)cim[x], mini_c = (int)cic[x], sumi = (int)cim[x], mint_i = (int)(cim[x] + cic[x]);
            for (int a = 1; a < cm; ++a) {
                int y = mem[a];
                uint32_t om = com[y], oc = coc[y], im = cim[y], ic = cic[y];
                sumo += (int)om;
                if ((uint32_t)mino > om) { mino = (int)om; mino_c = (int)oc; }
                else if ((uint32_t)mino == om) { if ((uint32_t)mino_c > oc) mino_c = (int)oc; }
                if ((uint32_t)mint_o > om + oc) mint_o = (int)(om + oc);
                sumi += (int)im;
                if ((uint32_t)mini > im) { mini = (int)im; mini_c = (int)ic; }
                else if ((uint32_t)mini == im) { if ((uint32_t)mini_c > ic) mini_c = (int)ic; }
                if ((uint32_t)mint_i > im + ic) mint_i = (int)(im + ic);
            }
            S->lu[0] = mino; S->lu[1] = mino_c; S->lu[2] = sumo; S->lu[3] = mint_o;
            S->lu[4] = mini; S->lu[5] = mini_c; S->lu[6] = sumi; S->lu[7] = mint_i;
            // value tables: st[x] = #values above x, sa[x] = sum of values above x
            int c = 0, s = 0;
            for (int v = cm; v >= 0; --v) { int h = S->ho[v]; S->ho[v] = c; S->so[v] = s; c += h; s += h * v; }
            c = 0; s = 0;
            for (int v = cm; v >= 0; --v) { int h = S->hi[v]; S->hi[v] = c; S->si[v] = s; c += h; s += h * v; }
        }
        __syncthreads();
        const int mino = S->lu[0], mino_c = S->lu[1], sumo = S->lu[2], mint_o = S->lu[3];
        const int mini = S->lu[4], mini_c = S->lu[5], sumi = S->lu[6], mint_i = S->lu[7];
        __syncthreads();
        meo = gmd(cm + 1, mdo);
        mei = gmd(cm + 1, mdi);
        if (mino < mdo[cm] || mini < mdi[cm]) {
            int nmin = max(gmd(cm, mdo) - mino, gmd(cm, mdi) - mini);
            int no = first_true(nmin, mino_c, [&](int i) { return !(mino + i < mdo[cm + i]); });
            if (mino + no < mdo[cm + no]) return false;
            int ni = first_true(nmin, mini_c, [&](int i) { return !(mini + i < mdi[cm + i]); });
            if (mini + ni < mdi[cm + ni]) return false;
            lb = max(no, ni);
            ub = min(P->ubo[mint_o], P->ubi[mint_i]) + 1 - cm;
            if (ub > rc) ub = rc;
            if (lb < ub) {
                const int32_t *sto = S->ho, *sao = S->so, *sti = S->hi, *sai = S->si;
                int i1 = first_true(lb, ub - 1, [&](int i) {
                    int so_ = psum(sto, sao, cm, i), si_ = psum(sti, sai, cm, i);
                    return !(sumo + si_ < cm * mdo[cm + i] && sumi + so_ < cm * mdi[cm + i]);
                });
                {
                    int so_ = psum(sto, sao, cm, i1), si_ = psum(sti, sai, cm, i1);
                    if (sumo + so_ < cm * mdo[cm + i1] && sumi + si_ < cm * mdi[cm + i1]) return false;
                }
                lb = i1;
                int t = last_true(i1 + 1, ub, [&](int i) {
                    int so_ = psum(sto, sao, cm, i), si_ = psum(sti, sai, cm, i);
                    return sumo + si_ >= cm * mdo[cm + i] && sumi + so_ >= cm * mdi[cm + i];
                });
                int tight = t >= i1 + 1 ? t : i1;
                if (ub > tight) ub = tight;
                if (lb > 1) { meo = gmd(cm + lb, mdo); mei = gmd(cm + lb, mdi); }
            }
        } else {
            ub = rc;
            lb = cm < P->min_size ? P->min_size - cm : 0;
        }
        if (cm + ub < P->min_size) return false;
        if (ub < 0 || ub < lb) return false;
        return true;
    }

    __device__ __forceinline__ void clear_hist(int cm) {
        for (int v = tid; v <= cm + 1 && v < HMAX + 2; v += NT) { S->ho[v] = 0; S->hi[v] = 0; }
    }

    __device__ __forceinline__ bool members_ext(int cm, int ub, int lb, int meo, int mei) {
        int bad = 0;
        for (int a = tid; a < cm; a += NT)
            if (!vert_ext(mem[a], cm, ub, lb, meo, mei)) bad = 1;
        return !__syncthreads_or(bad);
    }

    // Splits R[0..nr) by validity: valid ids to keep[], invalid ones to drop[] (their
    // candidate bit cleared). Returns the kept count; *nd gets the dropped count.
    __device__ __forceinline__ int filter(const int32_t* R, int nr, int32_t* keep, int32_t* drop, int* nd,
                          int cm, int ub, int lb, int meo, int mei) {
        if (tid == 0) { S->cnt[0] = 0; S->cnt[1] = 0; }
        __syncthreads();
        for (int i = tid; i < nr; i += NT) {
            int w = R[i];
            if (cand_valid(w, cm, ub, lb, meo, mei)) keep[atomicAdd(&S->cnt[0], 1)] = w;
            else { drop[atomicAdd(&S->cnt[1], 1)] = w; bcl(ccand, w); }
        }
        __syncthreads();
        int k = S->cnt[0];
        *nd = S->cnt[1];
        __syncthreads();
        return k;
    }

    // Number of entries of nbr[b, e) whose id is a current candidate (ccand bit), for one
    // thread: 8 list loads, then their 8 bitmap words, in flight per step.
    __device__ __forceinline__ uint32_t count_cand(const int32_t* nbr, unsigned long long b, unsigned long long e) const {
        uint32_t c = 0;
        for (; b + 8 <= e; b += 8) {
            int a[8];
            uint32_t w[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = nbr[b + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) w[u] = bword(ccand, a[u]);
#pragma unroll
            for (int u = 0; u < 8; ++u) c += bbit(w[u], a[u]);
        }
        for (; b < e; ++b) c += (uint32_t)bt(ccand, nbr[b]);
        return c;
    }
    // Same count for one wavefront (lanes stride the list), summed across the wave.
    __device__ __forceinline__ uint32_t count_cand_wave(const int32_t* nbr, unsigned long long b, unsigned long long e, int lane) const {
        uint32_t c = 0;
        unsigned long long i = b + lane;
        for (; i + 3 * 64 < e; i += 4 * 64) {
            const int a0 = nbr[i], a1 = nbr[i + 64], a2 = nbr[i + 128], a3 = nbr[i + 192];
            const uint32_t w0 = bword(ccand, a0), w1 = bword(ccand, a1), w2 = bword(ccand, a2), w3 = bword(ccand, a3);
            c += bbit(w0, a0) + bbit(w1, a1) + bbit(w2, a2) + bbit(w3, a3);
        }
        for (; i < e; i += 64) c += (uint32_t)bt(ccand, nbr[i]);
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
        return c;
    }

    // Number of entries of nbr[b, e) that are current candidates, for a group of G lanes
    // (lanes sub = 0..G-1 of one vertex's group): each lane takes every G-th entry, four
    // at a time, and the group's lanes read G consecutive entries per load, so a load
    // instruction touches few cache lines; the sum is formed in the group (G a power of 2,
    // groups aligned inside the wavefront).
    template <int G>
    __device__ __forceinline__ uint32_t count_cand_group(const int32_t* nbr, unsigned long long b, unsigned long long e,
                                                         int sub) const {
        uint32_t c = 0;
        unsigned long long i = b + (unsigned long long)sub;
        for (; i + 3 * G < e; i += 4 * G) {
            const int a0 = nbr[i], a1 = nbr[i + G], a2 = nbr[i + 2 * G], a3 = nbr[i + 3 * G];
            const uint32_t w0 = bword(ccand, a0), w1 = bword(ccand, a1), w2 = bword(ccand, a2), w3 = bword(ccand, a3);
            c += bbit(w0, a0) + bbit(w1, a1) + bbit(w2, a2) + bbit(w3, a3);
        }
        for (; i < e; i += G) c += (uint32_t)bt(ccand, nbr[i]);
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
        return c;
    }

    // can-degrees of live vertices recounted as |out(x) n R|, |in(x) n R| (R = ccand).
    // Each vertex is counted by a group of RG lanes (coalesced list reads, many vertices
    // in flight); vertices with long lists (hubs) are queued and counted one per wavefront.
    __device__ __forceinline__ void recount(int cm, const int32_t* R, int nr) {
        constexpr unsigned long long LONG_LIST = 512;
        constexpr int RG = 8;
        const unsigned long long *oo = P->oo, *io = P->io;
        const int N = cm + nr;
        if (tid == 0) S->cnt[8] = 0;
        __syncthreads();
        const int sub = tid & (RG - 1);
        for (int i = tid / RG; i < N; i += NT / RG) {
            const int x = i < cm ? mem[i] : R[i - cm];
            const unsigned long long ob = oo[x], oe = oo[x + 1], ib = io[x], ie = io[x + 1];
            if ((oe - ob) + (ie - ib) > LONG_LIST) {
                if (sub == 0) gat[atomicAdd(&S->cnt[8], 1)] = x;
                continue;
            }
            const uint32_t co = count_cand_group<RG>(P->on, ob, oe, sub);
            const uint32_t ci = count_cand_group<RG>(P->in, ib, ie, sub);
            if (sub == 0) {
                coc[x] = co;
                cic[x] = ci;
                bytes += 4ull * ((oe - ob) + (ie - ib)) + 32;
            }
        }
        __syncthreads();
        const int nlong = S->cnt[8];
        const int wave = tid >> 6, lane = tid & 63;
        for (int j = wave; j < nlong; j += NT / 64) {
            const int x = gat[j];
            const uint32_t co = count_cand_wave(P->on, oo[x], oo[x + 1], lane);
            const uint32_t ci = count_cand_wave(P->in, io[x], io[x + 1], lane);
            if (lane == 0) {
                coc[x] = co;
                cic[x] = ci;
                bytes += 4ull * ((oo[x + 1] - oo[x]) + (io[x + 1] - io[x])) + 32;
            }
        }
        __syncthreads();
    }
    // can-degrees of live vertices decremented by the arcs of dropped candidates
    __device__ __forceinline__ void decrement(const int32_t* D, int nd) {
        const unsigned long long *oo = P->oo, *io = P->io;
        for (int i = tid; i < nd; i += NT) {
            int u = D[i];
            thread_scan_bits(P->on, oo[u], oo[u + 1], cmem, ccand, [&](int w, uint32_t m, uint32_t c) {
                if (m | c) atomicSub(&cic[w], 1u);
            });
            thread_scan_bits(P->in, io[u], io[u + 1], cmem, ccand, [&](int w, uint32_t m, uint32_t c) {
                if (m | c) atomicSub(&coc[w], 1u);
            });
            bytes += 4ull * ((oo[u + 1] - oo[u]) + (io[u + 1] - io[u])) + 32;
        }
        __syncthreads();
    }
    // count valid candidates (old bounds) into the member-degree histograms
    __device__ __forceinline__ int hist_valid(const int32_t* R, int nr, int cm, int ub, int lb, int meo, int mei) {
        clear_hist(cm);
        __syncthreads();
        int c = 0;
        for (int i = tid; i < nr; i += NT) {
            int w = R[i];
            if (cand_valid(w, cm, ub, lb, meo, mei)) {
                atomicAdd(&S->ho[com[w]], 1);
                atomicAdd(&S->hi[cim[w]], 1);
                ++c;
            }
        }
        if (tid == 0) S->cnt[2] = 0;
        __syncthreads();
        if (c) atomicAdd(&S->cnt[2], c);
        return read_counter(2);
    }

    // reference src/device_kernels.cu:965-1305 (d_degree_pruning). On entry the
    // candidates that passed diameter pruning are R[0..rc) (bits in ccand, degrees in the
    // histograms); ccount is the child's candidate count before that pruning. Degrees are
    // recounted exactly over the remaining set (the reference's recount and decrement
    // branches produce the same values). On success *R / *nr hold the final candidates.
    __device__ __forceinline__ bool degree_prune(int cm, int32_t** R, int* nr, int ccount, int& ub, int& lb,
                                 int& meo, int& mei, int32_t* spare, int32_t* dropped) {
        int rc = *nr;
        if (!lu_bounds(cm, rc, ub, lb, meo, mei)) return false;
        if (!members_ext(cm, ub, lb, meo, mei)) return false;
        if (lazy_v >= 0) materialize(*R, rc);
        int nd = 0;
        int32_t* cur = *R;
        int k = filter(cur, rc, spare, dropped, &nd, cm, ub, lb, meo, mei);
        int32_t* tmp = cur; cur = spare; spare = tmp;
        int removed = ccount - k;
        bool all_mat = (removed == nd);
        while (k > 0 && removed > 0) {
            if (all_mat && nd < k + cm) decrement(dropped, nd);
            else recount(cm, cur, k);
            int nv = hist_valid(cur, k, cm, ub, lb, meo, mei);
            if (!lu_bounds(cm, nv, ub, lb, meo, mei)) { *R = cur; *nr = k; return false; }
            if (!members_ext(cm, ub, lb, meo, mei)) { *R = cur; *nr = k; return false; }
            int k2 = filter(cur, k, spare, dropped, &nd, cm, ub, lb, meo, mei);
            tmp = cur; cur = spare; spare = tmp;
            removed = nd;
            all_mat = true;
            k = k2;
        }
        *R = cur;
        *nr = k;
        return true;
    }

    // reference src/device_kernels.cu:1488-1532 (d_check_for_clique)
    __device__ __forceinline__ void check_clique(int cm) {
        if (cm < P->min_size || discard_child) return;  // a branch before a level-0 window emits nothing
        uint32_t mo = (uint32_t)P->mdo[cm], mi = (uint32_t)P->mdi[cm];
        int bad = 0;
        for (int a = tid; a < cm; a += NT) {
            int x = mem[a];
            if (com[x] < mo || cim[x] < mi) bad = 1;
        }
        if (__syncthreads_or(bad)) return;
        emit_clique(mem, cm);
    }

    __device__ __forceinline__ void emit_clique(const int32_t* ids, int len) {
        unsigned long long a = 0;
        if (tid == 0)
            a = __atomic_load_n(&P->g->error, __ATOMIC_RELAXED) ? ~0ull
                : atomicAdd(&P->g->clq_alloc, (1ull << CNT_SHIFT) | (unsigned long long)len);
        a = bcast0(a);
        if (a == ~0ull) return;
        unsigned long long idx = a >> CNT_SHIFT, off = a & SIZE_MASK;
        if (idx >= P->clq_cap_s || off + len > P->clq_cap_i) {
            if (tid == 0) atomicOr(&P->g->error, (unsigned)ERR_CLIQUES);
            return;
        }
        for (int i = tid; i < len; i += NT) P->clq_ids[off + i] = ids[i];
        if (tid == 0) {
            P->clq_meta[idx] = (off << 24) | (unsigned long long)len;
            atomicAdd(&P->g->cliques, 1ull);
        }
        bytes += 4ull * len / NT + 1;
    }

    // ---------------------------------------------------------------- sort + write
    // Orders members then candidates by the enumeration order of reference
    // inc/device_kernels.hpp:31-80 (min(mem) desc, min(can) desc, lvl2adj desc, id desc)
    // and pushes the child onto the task stack.
    __device__ __forceinline__ long long write_child(int cm, const int32_t* R, int nr, int mode) {
        const int N = cm + nr;
        if (discard_child) return -2;  // a level-0 branch before the bounded unit's window
        if (P->max_depth && cur_depth + 1u > P->max_depth) {  // depth-bounded enumeration
            if (tid == 0) atomicAdd(&P->g->truncated, 1ull);
            return -1;
        }
        flush_l2(ptot);  // the child's records carry the parent's lvl2adj
        // key field widths from this child's maxima
        uint32_t mmx = 0, mcx = 0, l2x = 0;
        uint32_t omin = UINT_MAX, omax = 0, imin = UINT_MAX, imax = 0;  // degree features
        for (int i = tid; i < N; i += NT) {
            int x = i < cm ? mem[i] : R[i - cm];
            gat[i] = x;
            mmx = max(mmx, min(com[x], cim[x]));
            mcx = max(mcx, min(coc[x], cic[x]));
            l2x = max(l2x, pl2[x]);
            const uint32_t od = com[x] + coc[x], id = cim[x] + cic[x];
            omin = min(omin, od); omax = max(omax, od); imin = min(imin, id); imax = max(imax, id);
        }
        if (tid == 0) {
            S->red[2] = 0; S->red[3] = 0; S->red[4] = 0;
            S->red[5] = -1; S->red[6] = 0; S->red[7] = -1; S->red[8] = 0;
        }
        __syncthreads();
        atomicMax((unsigned*)&S->red[2], mmx);
        atomicMax((unsigned*)&S->red[3], mcx);
        atomicMax((unsigned*)&S->red[4], l2x);
        atomicMin((unsigned*)&S->red[5], omin);
        atomicMax((unsigned*)&S->red[6], omax);
        atomicMin((unsigned*)&S->red[7], imin);
        atomicMax((unsigned*)&S->red[8], imax);
        __syncthreads();
        mmx = (uint32_t)S->red[2]; mcx = (uint32_t)S->red[3]; l2x = (uint32_t)S->red[4];
        omin = (uint32_t)S->red[5]; omax = (uint32_t)S->red[6]; imin = (uint32_t)S->red[7]; imax = (uint32_t)S->red[8];
        __syncthreads();
        const uint32_t idx_max = (uint32_t)(P->n - 1);
        const int bi = bitw(idx_max), bl = bitw(l2x), bc = bitw(mcx), bm = bitw(mmx);
        const int total = 1 + bm + bc + bl + bi;
        unsigned long long a = 0;
        if (tid == 0) a = reserve(false, (unsigned long long)N);
        a = bcast0(a);
        if (a == ~0ull) return -1;
        const unsigned long long t = a >> CNT_SHIFT, o = a & SIZE_MASK;
        TaskDesc* dd = P->sp_desc;
        dqc::TaskVertex* dv = P->sp_vtx + o;
        if (total <= 64) {
            int Np = 1;
            while (Np < N) Np <<= 1;
            unsigned long long* sk = Np <= SORT_N ? S->sortk : hsort;
            for (int i = tid; i < Np; i += NT) {
                unsigned long long key = ~0ull;
                if (i < N) {
                    int x = gat[i];
                    unsigned long long cls = i < cm ? 0ull : 1ull;
                    unsigned long long mm = mmx - min(com[x], cim[x]);
                    unsigned long long mc = mcx - min(coc[x], cic[x]);
                    unsigned long long l2 = l2x - pl2[x];
                    unsigned long long id = idx_max - (uint32_t)x;
                    key = cls;
                    key = (key << bm) | mm;
                    key = (key << bc) | mc;
                    key = (key << bl) | l2;
                    key = (key << bi) | id;
                }
                sk[i] = key;
            }
            __syncthreads();
            for (int k = 2; k <= Np; k <<= 1)
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int i = tid; i < Np; i += NT) {
                        int ixj = i ^ j;
                        if (ixj > i) {
                            unsigned long long x = sk[i], y = sk[ixj];
                            bool up = (i & k) == 0;
                            if ((x > y) == up) { sk[i] = y; sk[ixj] = x; }
                        }
                    }
                    __syncthreads();
                }
            const unsigned long long idmask = bi == 0 ? 0ull : ((1ull << bi) - 1);
            for (int i = tid; i < N; i += NT) {
                int x = (int)(idx_max - (uint32_t)(sk[i] & idmask));
                dv[i] = dqc::TaskVertex{(uint32_t)x, com[x], coc[x], cim[x], cic[x], pl2[x]};
            }
        } else {
            // fallback: rank by full comparison (keys wider than 64 bits)
            if (tid == 0) atomicAdd(&P->g->slow_sorts, 1ull);
            for (int i = tid; i < N; i += NT) {
                int x = gat[i];
                bool cx = i >= cm;
                uint32_t mx = min(com[x], cim[x]), cx2 = min(coc[x], cic[x]), lx = pl2[x];
                int rank = 0;
                for (int j2 = 0; j2 < N; ++j2) {
                    int y = gat[j2];
                    bool cy = j2 >= cm;
                    if (cy != cx) { rank += cy < cx; continue; }
                    uint32_t my = min(com[y], cim[y]);
                    if (my != mx) { rank += my > mx; continue; }
                    uint32_t cyc = min(coc[y], cic[y]);
                    if (cyc != cx2) { rank += cyc > cx2; continue; }
                    uint32_t ly = pl2[y];
                    if (ly != lx) { rank += ly > lx; continue; }
                    rank += (uint32_t)y > (uint32_t)x;
                }
                dv[rank] = dqc::TaskVertex{(uint32_t)x, com[x], coc[x], cim[x], cic[x], pl2[x]};
            }
        }
        if (tid == 0) {
            TaskDesc d;
            d.off = o; d.tot = N; d.nm = cm; d.ncov = 0; d.mode = mode;
            d.iters = N - cm; d.flags = 0; d.step0 = 0;
            d.uid = (uint32_t)(atomicAdd(&P->g->next_uid, 1ull) + 1);
            d.parent = cur_uid;
            d.depth = (uint16_t)min(cur_depth + 1u, 65535u); d.gid = (uint16_t)cur_gid;
            d.dmin_o = (uint16_t)min(omin, 65535u); d.dmax_o = (uint16_t)min(omax, 65535u);
            d.dmin_i = (uint16_t)min(imin, 65535u); d.dmax_i = (uint16_t)min(imax, 65535u);
            dd[t] = d;
        }
        bytes += 24ull * N / NT + 1;
        __syncthreads();
        return (long long)t;
    }

    // Thread 0: room for one push of `recs` records, onto the main stack or (deferred)
    // the deferred region. The capacity is tested on the combined main + deferred use
    // that the one atomic on q->used returns, so concurrent pushes to the two regions
    // cannot both pass on a stale view of the other; only then is the slot taken in its
    // region. Returns the region's previous (tasks << 34) | records, or ~0ull when full
    // or after an error.
    __device__ __forceinline__ unsigned long long reserve(bool deferred, unsigned long long recs) {
        if (__atomic_load_n(&P->g->error, __ATOMIC_RELAXED)) return ~0ull;
        const unsigned long long inc = (1ull << CNT_SHIFT) | recs;
        const unsigned long long u = atomicAdd(&P->q->used, inc);
        if ((u >> CNT_SHIFT) + 1 > P->sp_cap_t || (u & SIZE_MASK) + recs > P->sp_cap_v) {
            note_spill(u, recs | (deferred ? (1ull << 63) : 0ull));
            return ~0ull;
        }
        return atomicAdd(deferred ? &P->q->def_alloc : &P->q->sp_alloc, inc);
    }

    // Flags a stack overflow; the first one records the stack use, the request (bit 63:
    // a deferred piece), this workgroup's launch state and the pending batch share.
    __device__ __forceinline__ void note_spill(unsigned long long top, unsigned long long req) {
        if (atomicOr(&P->g->error, (unsigned)ERR_SPILL) & ERR_SPILL) return;
        P->g->spill_state[0] = top;
        P->g->spill_state[1] = req;
        P->g->spill_state[2] = gridDim.x;
        P->g->spill_state[3] = __atomic_load_n(&P->q->pend, __ATOMIC_RELAXED);
    }

    // Stack backpressure. After a step passes this check a workgroup writes at most one
    // child (no larger than the batch's largest task) before its next check, and every
    // batch task not yet finished may later be pushed back once as a piece of at most its
    // own size. The stack stays within capacity as long as each check sees room for one
    // child per resident workgroup plus all unfinished batch tasks; a workgroup that does
    // not see it pushes the rest of its task as a piece and stops it. At launch start the
    // pop leaves that room, so a workgroup's first check in a launch passes unless other
    // workgroups already wrote children (no launch can stall). Its later checks also keep
    // one task of the largest size (kloc) free, so that a launch leaves room for the next.
    // Returns 0 (go on), 1 (stack low: defer the rest) or 2 (the launch's time slice is
    // over: push the rest back onto the stack; never at a workgroup's first check).
    __device__ __forceinline__ int stack_low() {
        unsigned long long low = 0;
        if (tid == 0) {
            const unsigned long long a = __atomic_load_n(&P->q->used, __ATOMIC_RELAXED);
            const unsigned long long pend = __atomic_load_n(&P->q->pend, __ATOMIC_RELAXED);
            const unsigned long long mt = P->maxtot == ~0ull ? P->q->maxtot : P->maxtot;
            const unsigned long long wg = P->maxtot == ~0ull ? P->q->wg_limit : gridDim.x;
            const unsigned long long used_t = a >> CNT_SHIFT, used_v = a & SIZE_MASK;
            const unsigned long long margin = launch_steps ? (unsigned long long)P->kloc : 0ull;
            low = used_t + wg + (pend >> CNT_SHIFT) + (launch_steps ? 1 : 0) > P->sp_cap_t ||
                  used_v + wg * mt + (pend & SIZE_MASK) + margin > P->sp_cap_v;
            if (!low && P->slice_ticks && launch_steps &&
                __builtin_amdgcn_s_memrealtime() - t_start > P->slice_ticks)
                low = 2;
        }
        const int r = (int)bcast0(low);
        if (!r) ++launch_steps;
        return r;
    }

    // Pushes the current state of the task (live vertices in position order with their
    // current degrees and lvl2adj) as a piece that runs `iters` loop steps from here.
    // deferred: into the deferred region (backpressure), which is expanded only when the
    // main stack is empty, after the children written below the task have run.
    // step0: loop steps of the task before the piece's first; l1_done: a level-0 task whose
    // first child (the host level-1 task) is already written.
    __device__ __forceinline__ void write_piece(const TaskDesc& src, int tot, int nm, int mode, bool remove_first, int iters,
                                                bool deferred, int step0, bool l1_done) {
        flush_l2(tot);
        unsigned long long a = 0;
        if (tid == 0) a = reserve(deferred, (unsigned long long)tot);
        a = bcast0(a);
        if (a == ~0ull) return;
        // the deferred region fills the arrays from the top down: its newest piece has the
        // lowest descriptor index and the lowest records, so its newest pieces stay contiguous
        const unsigned long long t = deferred ? P->sp_cap_t - 1 - (a >> CNT_SHIFT) : a >> CNT_SHIFT;
        const unsigned long long o = deferred ? P->sp_cap_v - (a & SIZE_MASK) - (unsigned long long)tot : a & SIZE_MASK;
        dqc::TaskVertex* dv = P->sp_vtx + o;
        for (int i = tid; i < tot; i += NT) {
            const int x = pids[i];
            dv[i] = dqc::TaskVertex{(uint32_t)x, pom[x], poc[x], pim[x], pic[x], pl2[x]};
        }
        if (tid == 0) {
            TaskDesc d = src;
            d.off = o; d.tot = tot; d.nm = nm; d.ncov = 0; d.mode = mode;
            d.iters = iters;
            d.flags = TF_PIECE | (remove_first ? TF_REMOVE_FIRST : 0) | (l1_done ? TF_L1_DONE : 0);
            d.step0 = step0;
            P->sp_desc[t] = d;
        }
        bytes += 24ull * tot / NT + 1;
        __syncthreads();
    }

    // ---------------------------------------------------------------- branch steps
    // reference src/device_kernels.cu:400-531 / src/host_functions.cu:1386-1500.
    // Returns 0 (stop the task), 2 (skip this branch) or 1; a member falling short wins.
    __device__ __forceinline__ int remove_one(int& tot, int& nc, int nm) {
        --nc;
        --tot;
        if (tot < P->min_size) return 0;
        const uint32_t mo = (uint32_t)gmd(nm + 1, P->mdo), mi = (uint32_t)gmd(nm + 1, P->mdi);
        const int p = pids[tot], last = pids[tot - 1];
        if (tid == 0) bcl(plive, p);
        __syncthreads();
        // The out-, in- and two-hop lists touch disjoint fields (in_can, out_can, lvl2adj),
        // so they are walked in one phase. The reference stops after the out-list when a
        // member falls short; here the in- and two-hop decrements may also have run, which
        // is unobservable because a failure ends the task.
        int fail = 0, skip = 0;
        unsigned long long ob = P->oo[p], oe = P->oo[p + 1], ib = P->io[p], ie = P->io[p + 1];
        clip_live(P->on, ob, oe);
        clip_live(P->in, ib, ie);
        wg_scan_bits2(P->on, ob, oe, P->in, ib, ie, plive, pmem, [&](int w, uint32_t live, uint32_t memb, bool in_list) {
            if (!live) return;
            if (!in_list) {
                uint32_t ic = pic[w] - 1u;
                pic[w] = ic;
                if (pim[w] + ic < mi) { if (memb) fail = 1; else if (w == last) skip = 1; }
            } else {
                uint32_t oc = poc[w] - 1u;
                poc[w] = oc;
                if (pom[w] + oc < mo) { if (memb) fail = 1; else if (w == last) skip = 1; }
            }
        });
        // the two-hop list's lvl2adj decrements are left pending (flush_l2); its bytes are
        // counted here, as the reference's removal scans it
        if (tid == 0)
            bytes += 4ull * ((P->oo[p + 1] - P->oo[p]) + (P->io[p + 1] - P->io[p]) + (P->to[p + 1] - P->to[p])) + 48;
        if (__syncthreads_or(fail)) return 0;
        return __syncthreads_or(skip) ? 2 : 1;
    }

    // Applies the pending lvl2adj decrements of the removals of positions [tot, l2_tot)
    // (reference src/device_kernels.cu:513-524: a removal decrements lvl2adj of the live
    // non-members in the removed vertex's two-hop list). Only the values of vertices still
    // live are read again, and a vertex live now was live at each of those removals, so
    // applying them together gives the values of applying them one removal at a time.
    // Called before every read of lvl2adj: a task whose branches write no child and whose
    // lookahead fails on degrees never needs them.
    // The pending two-hop lists are shared out among the wavefronts (each takes the next
    // list when it is done with one), so the workgroup has the entries of several lists in
    // flight and no barrier between lists; a vertex in several of them takes atomic
    // decrements.
    __device__ __forceinline__ void flush_l2(int tot) {
        const int np = l2_tot - tot;
        if (np > 0) {
            if (tid == 0) S->cnt[10] = 0;
            __syncthreads();
            const int lane = tid & 63;
            const bool ranged = !(id_lo == INT_MIN && id_hi == INT_MAX);
            while (true) {
                int j = 0;
                if (lane == 0) j = atomicAdd(&S->cnt[10], 1);
                j = __shfl(j, 0, 64);
                if (j >= np) break;
                const int p = pids[tot + j];
                unsigned long long b = P->to[p], e = P->to[p + 1];
                if (ranged && e - b >= 512) {  // only ids in [id_lo, id_hi) can be live
                    b = wave_lower_bound(P->tn, b, e, id_lo, lane);
                    e = wave_lower_bound(P->tn, b, e, id_hi, lane);
                }
                wave_decrement_l2(P->tn + b, (int)(e - b), lane);
            }
            __syncthreads();
        }
        l2_tot = tot;
    }
    // One wavefront: lvl2adj -= 1 for every live non-member in the list a[0, n), with
    // SCAN_UNR entries per lane and their bitmap words in flight per round.
    __device__ __forceinline__ void wave_decrement_l2(const int32_t* __restrict__ a, int n, int lane) {
        if (n <= 0) return;
        constexpr int U = SCAN_UNR;
        const int last = n - 1;
        for (int k = lane; k < n; k += 64 * U) {
            int v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = a[min(k + u * 64, last)];
            uint32_t wa[U], wb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                wa[u] = *bword_ptr(plive, v[u]);
                wb[u] = *bword_ptr(pmem, v[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k + u * 64 < n && bbit(wa[u], v[u]) && !bbit(wb[u], v[u])) atomicSub(&pl2[v[u]], 1u);
        }
    }

    // reference src/host_functions.cu:1333-1383 (host: lvl2adj >= tot - 1) and
    // src/device_kernels.cu:330-397 (device: lvl2adj >= num_cand - 1). The degree tests
    // come first: lvl2adj is brought up to date only when all of them pass.
    __device__ __forceinline__ bool lookahead(int tot, int nm, int nc, int mode) {
        const uint32_t mo = (uint32_t)gmd(tot, P->mdo), mi = (uint32_t)gmd(tot, P->mdi);
        const uint32_t l2min = mode == MODE_DEVICE ? (uint32_t)(nc - 1) : (uint32_t)(tot - 1);
        for (int base = 0; base < tot; base += NT) {
            int i = base + tid, bad = 0;
            if (i < tot) {
                int x = pids[i];
                bad = (pom[x] + poc[x] < mo) || (pim[x] + pic[x] < mi);
            }
            if (__syncthreads_or(bad)) return false;
        }
        flush_l2(tot);
        for (int base = nm; base < tot; base += NT) {
            const int i = base + tid;
            const int bad = i < tot && pl2[pids[i]] < l2min;
            if (__syncthreads_or(bad)) return false;
        }
        return true;
    }

    // One branch (reference src/device_kernels.cu:92-165 / src/host_functions.cu:1046-1133):
    // add the last candidate v, diameter + degree pruning, critical-vertex pruning, clique
    // check, child write. Only members and the candidates in L(v) are materialised: all
    // others are dropped by diameter pruning before any of their values is read.
    // Returns true if a child was written.
    __device__ __forceinline__ bool branch(int tot, int nm, int mode, bool first_child_host) {
        ptot = tot;
        const int v = pids[tot - 1];
        const int cm = nm + 1;
        if (cm + 1 >= HMAX) {  // member histograms sized for fewer members
            if (tid == 0) atomicOr(&P->g->error, (unsigned)ERR_MEMBERS);
            return false;
        }
        begin_child(cm);
        // members: the parent's plus v
        dirty_m = nm + 1;
        for (int a = tid; a <= nm; a += NT) {
            int x = a < nm ? pids[a] : v;
            mem[a] = x;
            com[x] = pom[x]; coc[x] = poc[x]; cim[x] = pim[x]; cic[x] = pic[x];
            bs(cmem, x);
        }
        // candidates within two hops of v (the others fail diameter pruning), collected into
        // lst; their child degrees stay implicit (lazy_v)
        {
            unsigned long long b = P->to[v], e = P->to[v + 1];
            clip_live(P->tn, b, e);
            wg_collect(P->tn, b, e, plive, pmem, v, lst, 3);
        }
        if (tid == 0) bytes += 4ull * (P->to[v + 1] - P->to[v]) + 16;
        vob = P->oo[v]; voe = P->oo[v + 1]; vib = P->io[v]; vie = P->io[v + 1];
        clip_live(P->on, vob, voe);
        clip_live(P->in, vib, vie);
        __syncthreads();
        const int nl = S->cnt[3];
        dirty_l = nl;
        // reference src/device_kernels.cu:534-602 (d_add_one_vertex): v joins the members.
        // Members' degrees are updated; v's live candidate neighbours are marked.
        // (a vertex in both lists has its out- and in-fields updated by different threads)
        wg_scan_bits2(P->on, vob, voe, P->in, vib, vie, cmem, plive, [&](int w, uint32_t m, uint32_t l, bool in_list) {
            if (!in_list) {
                if (m) { cim[w] += 1u; cic[w] -= 1u; }
                else if (l) bs(ccrit, w);
            } else {
                if (m) { com[w] += 1u; coc[w] -= 1u; }
                else if (l) bs(cL, w);
            }
        });
        if (tid == 0) bytes += 4ull * ((P->oo[v + 1] - P->oo[v]) + (P->io[v + 1] - P->io[v])) + 32;
        __syncthreads();
        lazy_v = v;
        const long long t = branch_tail(nm, tot, nl, mode, first_child_host ? MODE_HOST_L1 : MODE_DEVICE, -1);
        lazy_v = -1;
        end_child();
        return t >= 0 || t == -2;  // -2: a child outside the bounded unit (still the first child)
    }

    __device__ __forceinline__ void begin_child(int cm) {
        if (tid == 0) { S->cnt[3] = 0; S->cnt[4] = 0; }
        clear_hist(cm);
        __syncthreads();
    }
    // Clears the child's bitmap words, so the bitmaps are all-zero between branches, also
    // across launches when they live in HBM. The marks of v's neighbours (cL, ccrit) are
    // already cleared when the branch's degree test ends; every other bit set in a branch
    // belongs to one of its members or L(v) candidates: member bits of mem[0..dirty_m),
    // candidate bits among lst, and the members and ccrit bits critical() sets among lst.
    __device__ __forceinline__ void end_child() {
        __syncthreads();
        const int nclear = dirty_m + (cand_marks || crit_marks ? dirty_l : 0);
        for (int i = tid; i < nclear; i += NT) {
            if (i < dirty_m) {
                cmem[(mem[i] - gb) >> 5] = 0;
            } else {
                const int wd = (lst[i - dirty_m] - gb) >> 5;
                if (cand_marks) ccand[wd] = 0;
                if (crit_marks) { cmem[wd] = 0; ccrit[wd] = 0; }
            }
        }
        __syncthreads();
        dirty_m = 0;
        dirty_l = 0;
        crit_marks = false;
        cand_marks = false;
    }

    // Branch s of the level-0 task, built without replaying the s earlier removals.
    // Level-0 positions are condensed ids and nothing there is a member, so after s
    // removals the live set is [0, k - s), v = k - s - 1, and every value the child
    // needs is a count on a sorted list: can-degrees count neighbours below v,
    // lvl2adj loses one per removed vertex (id >= k - s) whose two-hop list holds the
    // vertex (reference src/host_functions.cu:1386-1500, 1503-1559).
    __device__ __forceinline__ long long branch_root(int s, int k, const dqc::TaskVertex* root) {
        const int v = k - s - 1, tot = k - s;
        ptot = l2_tot = tot;  // the values below are computed as of step s
        id_lo = INT_MIN; id_hi = INT_MAX;  // (the live set is [0, tot): scans test ids themselves)
        begin_child(1);
        dirty_m = 1;
        const unsigned long long *oo = P->oo, *io = P->io;
        if (tid == 0) {
            mem[0] = v;
            com[v] = 0; cim[v] = 0;
            coc[v] = cnt_lt(P->on, oo[v], oo[v + 1], v);
            cic[v] = cnt_lt(P->in, io[v], io[v + 1], v);
            pl2[v] = root[v].lvl2adj - ((uint32_t)(P->lto[v + 1] - P->lto[v]) - cnt_lt(P->ltn, P->lto[v], P->lto[v + 1], tot));
            bs(cmem, v);
        }
        unsigned long long tb = P->to[v], te = P->to[v + 1];
        if (te - tb >= 8ull * NT) {  // only the entries below v are live
            if (tid < 64) {
                const unsigned long long x = wave_lower_bound(P->tn, tb, te, v, tid);
                if (tid == 0) S->clip[clip_pair] = x;
            }
            __syncthreads();
            te = S->clip[clip_pair];
            clip_pair ^= 2;
        }
        for (unsigned long long e = tb + tid; e < te; e += NT) {
            int u = P->tn[e];
            if (u >= v) continue;
            com[u] = contains(P->in, io[v], io[v + 1], u) ? 1u : 0u;   // u -> v
            cim[u] = contains(P->on, oo[v], oo[v + 1], u) ? 1u : 0u;   // v -> u
            coc[u] = cnt_lt(P->on, oo[u], oo[u + 1], v);
            cic[u] = cnt_lt(P->in, io[u], io[u + 1], v);
            const unsigned long long l0 = P->lto[u], l1 = P->lto[u + 1];
            pl2[u] = root[u].lvl2adj - ((uint32_t)(l1 - l0) - cnt_lt(P->ltn, l0, l1, tot));
            lst[atomicAdd(&S->cnt[3], 1)] = u;
        }
        if (tid == 0) bytes += 4ull * (P->to[v + 1] - P->to[v]) + 16;
        __syncthreads();
        const int nl = S->cnt[3];
        dirty_l = nl;
        __syncthreads();
        const long long t = branch_tail(0, tot, nl, MODE_HOST_L0, MODE_DEVICE, s);
        end_child();
        return t;
    }

    // Diameter test through child write (reference src/device_kernels.cu:122-165 /
    // src/host_functions.cu:1067-1133). Members are mem[0..nm] (v last), the L(v)
    // candidates lst[0..nl) with post-add degrees. Returns the child's stack index or -1.
    __device__ __forceinline__ long long branch_tail(int nm, int tot, int nl, int mode, int child_mode, int split_index) {
        int cm = nm + 1;
        // reference src/device_kernels.cu:830-906 (d_diameter_pruning): basic degree test
        {
            const uint32_t mo2 = (uint32_t)gmd(nm + 2, P->mdo), mi2 = (uint32_t)gmd(nm + 2, P->mdi);
            if (lazy_v < 0) cand_marks = true;
            for (int i = tid; i < nl; i += NT) {
                int w = lst[i];
                uint32_t om, oc, im, ic;
                if (lazy_v >= 0) {  // the add moves v's arcs from candidate to member
                    const uint32_t a = in_v(w), b = out_v(w);
                    om = pom[w] + a; oc = poc[w] - a; im = pim[w] + b; ic = pic[w] - b;
                } else {
                    om = com[w]; oc = coc[w]; im = cim[w]; ic = cic[w];
                }
                if (om + oc >= mo2 && im + ic >= mi2) {
                    if (lazy_v < 0) bs(ccand, w);  // (else when materialized)
                    la[atomicAdd(&S->cnt[4], 1)] = w;
                    atomicAdd(&S->ho[om], 1);
                    atomicAdd(&S->hi[im], 1);
                }
            }
        }
        int nr = read_counter(4);
        int ub = 0, lbd = 0, meo = 0, mei = 0;
        int32_t* R = la;
        if (tid == 0 && !discard_child) atomicAdd(&P->g->branches, 1ull);  // window branches only
        tick(3);
        bool ok = degree_prune(cm, &R, &nr, tot - 1 - nm, ub, lbd, meo, mei, lb, lc);
        if (lazy_v >= 0) {  // (degree_prune materialized the degrees if it got past its first test)
            clear_v_marks();
            lazy_v = -1;
        }
        tick(4);
        if (!ok) {
            // reference src/device_kernels.cu:131-132 (always) / src/host_functions.cu:1081
            if (mode == MODE_DEVICE || P->host_addfail) check_clique(cm);
            tick(7);
            return -1;
        }
        int s = critical(cm, R, nr, ub, lbd, meo, mei);
        tick(5);
        if (s == 2) return -1;
        check_clique(cm);
        tick(7);
        if (s == 0 || nr == 0) return -1;
        long long t = write_child(cm, R, nr, child_mode);
        // the first child of the level-0 task (a discarded one is marked 0xFFFFFFFF)
        if (split_index >= 0 && (t >= 0 || t == -2) && tid == 0)
            atomicMin(&P->q->root_first,
                      ((unsigned long long)split_index << 32) | (t >= 0 ? (unsigned long long)t : 0xFFFFFFFFull));
        tick(6);
        return t;
    }

    // reference src/device_kernels.cu:605-826 + 908-962 (critical vertex pruning,
    // diameter pruning by the critical set). Returns 2 (critical failure), 0 or 1.
    // cm / R / nr are updated in place.
    // On entry R is la or lb (degree_prune alternates those two and drops into lc).
    __device__ __forceinline__ int critical(int& cm, int32_t*& R, int& nr, int& ub, int& lbd, int& meo, int& mei) {
        crit_marks = true;
        const uint32_t tho = (uint32_t)P->mdo[cm + lbd], thi = (uint32_t)P->mdi[cm + lbd];
        for (int a = 0; a < cm; ++a) {
            int x = mem[a];
            if (com[x] + coc[x] == tho && coc[x] > 0u) {
                wg_scan_bits(P->on, P->oo[x], P->oo[x + 1], ccand, nullptr, [&](int w, uint32_t c, uint32_t) {
                    if (c) bs(ccrit, w);
                });
                if (tid == 0) bytes += 4ull * (P->oo[x + 1] - P->oo[x]) + 16;
            }
            if (cim[x] + cic[x] == thi && cic[x] > 0u) {
                wg_scan_bits(P->in, P->io[x], P->io[x + 1], ccand, nullptr, [&](int w, uint32_t c, uint32_t) {
                    if (c) bs(ccrit, w);
                });
                if (tid == 0) bytes += 4ull * (P->io[x + 1] - P->io[x]) + 16;
            }
        }
        if (tid == 0) { S->cnt[5] = 0; S->cnt[6] = 0; }
        __syncthreads();
        int32_t* K = crit;
        int32_t* rest = (R == la) ? lb : la;
        for (int i = tid; i < nr; i += NT) {
            int w = R[i];
            if (bt(ccrit, w)) K[atomicAdd(&S->cnt[5], 1)] = w;
            else rest[atomicAdd(&S->cnt[6], 1)] = w;
        }
        __syncthreads();
        const int nk = S->cnt[5], nrest = S->cnt[6];
        __syncthreads();
        if (nk == 0) return 1;
        if (cm + nk + 1 >= HMAX) {
            if (tid == 0) atomicOr(&P->g->error, (unsigned)ERR_MEMBERS);
            return 2;
        }
        for (int a = tid; a < cm; a += NT) adjc[mem[a]] = 0;
        for (int i = tid; i < nr; i += NT) adjc[R[i]] = 0;
        __syncthreads();
        // two-hop counts towards the critical set: one critical vertex at a time, its list
        // spread over the workgroup (entries of one list are distinct, so plain adds)
        for (int i = 0; i < nk; ++i) {
            const int c = K[i];
            wg_scan_bits(P->tn, P->to[c], P->to[c + 1], cmem, ccand, [&](int w, uint32_t m, uint32_t k) {
                if (m | k) adjc[w] += 1u;
            });
            if (tid == 0) bytes += 4ull * (P->to[c + 1] - P->to[c]) + 16;
            __syncthreads();
        }
        int bad = 0;
        for (int a = tid; a < cm; a += NT) if (adjc[mem[a]] != (uint32_t)nk) bad = 1;
        for (int i = tid; i < nk; i += NT) if ((int)adjc[K[i]] < nk - 1) bad = 1;
        if (__syncthreads_or(bad)) return 2;
        // the critical vertices join the members (reference src/device_kernels.cu:772-805)
        for (int i = 0; i < nk; ++i) {
            const int c = K[i];
            wg_scan_bits(P->on, P->oo[c], P->oo[c + 1], cmem, ccand, [&](int w, uint32_t m, uint32_t k) {
                if (m | k) { cim[w] += 1u; cic[w] -= 1u; }
            });
            wg_scan_bits(P->in, P->io[c], P->io[c + 1], cmem, ccand, [&](int w, uint32_t m, uint32_t k) {
                if (m | k) { com[w] += 1u; coc[w] -= 1u; }
            });
            if (tid == 0) bytes += 4ull * ((P->oo[c + 1] - P->oo[c]) + (P->io[c + 1] - P->io[c])) + 32;
            __syncthreads();
        }
        for (int i = tid; i < nk; i += NT) {
            int c = K[i];
            bs(cmem, c);
            bcl(ccand, c);
            mem[cm + i] = c;
        }
        cm += nk;
        dirty_m = cm;
        clear_hist(cm);
        if (tid == 0) S->cnt[7] = 0;
        __syncthreads();
        // reference src/device_kernels.cu:908-962: keep candidates two hops from every
        // critical vertex
        int32_t* keep = lc;
        for (int i = tid; i < nrest; i += NT) {
            int w = rest[i];
            if (adjc[w] == (uint32_t)nk) {
                keep[atomicAdd(&S->cnt[7], 1)] = w;
                atomicAdd(&S->ho[com[w]], 1);
                atomicAdd(&S->hi[cim[w]], 1);
            } else {
                bcl(ccand, w);
            }
        }
        int nkeep = read_counter(7);
        int32_t* old = R;  // no longer needed: becomes the spare list, rest the drop list
        R = keep;
        nr = nkeep;
        bool ok = degree_prune(cm, &R, &nr, nrest, ub, lbd, meo, mei, old, rest);
        return ok ? 1 : 0;
    }

    // ---------------------------------------------------------------- one task
    // reference src/device_kernels.cu:25-173 (device levels) and
    // src/host_functions.cu:987-1134 (host levels 0 and 1)
    //
    // A task with many branches is split instead: one pass runs only the removal and
    // lookahead steps, and every `chunk` steps (the piece length the scheduler chose,
    // carried in iters) pushes the state as a piece that runs the next `chunk` steps with
    // full branches (its first step removes unless it is the task's first). The pieces
    // run exactly the branches of the sequential loop, so the
    // branches, cliques and children are the same; the task is counted once, here.
    __device__ __forceinline__ void run_task(const TaskDesc& d, const dqc::TaskVertex* rv) {
        const int T = d.tot, nm = d.nm, mode = d.mode;
        const int ncov = mode == MODE_DEVICE ? 0 : d.ncov;
        const bool piece = (d.flags & TF_PIECE) != 0;
        set_graph(d.gid);
        cur_uid = d.uid;
        cur_depth = d.depth;
        nbranch = 0;
        if (tid == 0) { S->idr[0] = INT_MAX; S->idr[1] = INT_MIN; }
        __syncthreads();
        int lo = INT_MAX, hi = INT_MIN;
        for (int i = tid; i < T; i += NT) {
            dqc::TaskVertex r = rv[i];
            int x = (int)r.id;
            lo = min(lo, x); hi = max(hi, x);
            pids[i] = x;
            pom[x] = r.out_mem; poc[x] = r.out_can; pim[x] = r.in_mem; pic[x] = r.in_can; pl2[x] = r.lvl2adj;
            bs(plive, x);
            if (lsum) mark_live(x);
            if (i < nm) bs(pmem, x);
        }
        bytes += 24ull * T / NT + 1;
        ptot = l2_tot = T;  // no lvl2adj decrement pending
        if (tid == 0 && !piece) atomicAdd(&P->g->tasks, 1ull);
        if (lo <= hi) { atomicMin(&S->idr[0], lo); atomicMax(&S->idr[1], hi); }
        __syncthreads();
        id_lo = S->idr[0];
        id_hi = S->idr[1] + 1;
        int tot = T, nc = T - nm;
        int expansions = piece ? min(d.iters, nc) : nc;
        // a level-0 piece resumes its task's loop at step step0 (window bounds count from there)
        const int step0 = piece ? d.step0 : 0;
        if (mode == MODE_HOST_L0 && P->root_steps > 0)
            expansions = min(expansions, ncov + max(0, P->root_first + P->root_steps - step0));
        // the batch scheduler marks the tasks to split (host side, within a record budget)
        // a task marked TF_SPLIT carries its piece length (steps per piece) in iters
        const bool split = !piece && mode != MODE_HOST_L0 && (d.flags & TF_SPLIT);
        const int chunk = split ? max(1, d.iters) : 1;
        const bool remove_first = piece && (d.flags & TF_REMOVE_FIRST);
        bool first_host_child = (mode == MODE_HOST_L0) && !(d.flags & TF_L1_DONE);
        for (int j = ncov; j < expansions; ++j) {
            // The split pass writes only at chunk starts, so it checks the stack only there:
            // a deferred piece then takes the place of that chunk's piece and no step is
            // covered twice (a check inside a chunk would defer steps its piece already holds).
            const bool chunk_start = (j - ncov) % chunk == 0;
            const int low = (!split || chunk_start) ? stack_low() : 0;
            if (low) {
                // backpressure: the rest of the task becomes a deferred piece; at the end of
                // a time slice it goes back onto the main stack, to run in the next launch
                write_piece(d, tot, nm, mode, j != ncov || remove_first, expansions - j, low == 1, step0 + (j - ncov),
                            mode == MODE_HOST_L0 && !first_host_child);
                if (tid == 0) atomicAdd(low == 1 ? &P->g->pushbacks : &P->g->slices, 1ull);
                break;
            }
            if (split && chunk_start) write_piece(d, tot, nm, mode, j != ncov, min(chunk, expansions - j), false, j - ncov, false);
            if (j != ncov || remove_first) {
                tick(0);
                int r = remove_one(tot, nc, nm);
                tick(1);
                if (r == 0) break;
                if (r == 2) continue;
            }
            tick(0);
            bool la = lookahead(tot, nm, nc, mode);
            tick(2);
            if (la) {
                if (!split) emit_clique(pids, tot);  // the piece holding this step emits it
                break;
            }
            if (split) continue;
            // before a level-0 window only the first child matters (it takes the host
            // level-1 role); once it is found the branches there are skipped
            const bool before_window = mode == MODE_HOST_L0 && step0 + (j - ncov) < P->root_first;
            if (before_window && !first_host_child) continue;
            discard_child = before_window;
            bool wrote = branch(tot, nm, mode, first_host_child);
            discard_child = false;
            ++nbranch;
            tick(0);
            if (wrote) first_host_child = false;
        }
        if (P->log && tid == 0) {
            const unsigned long long li = atomicAdd(&P->g->log_n, 1ull);
            if (li < P->log_cap) {
                TaskLog r;
                r.uid = d.uid; r.parent = d.parent; r.depth = d.depth; r.nm = nm; r.tot = T;
                r.dmin_o = d.dmin_o; r.dmax_o = d.dmax_o; r.dmin_i = d.dmin_i; r.dmax_i = d.dmax_i;
                r.branches = nbranch; r.flags = (uint32_t)d.flags;
                P->log[li] = r;
            }
        }
        if (tid == 0) atomicSub(&P->q->pend, (1ull << CNT_SHIFT) | (unsigned long long)T);
        if (lsum) __syncthreads();  // every test of the summary is done before it clears
        for (int i = tid; i < T; i += NT) {
            const int wd = (pids[i] - gb) >> 5;
            plive[wd] = 0; pmem[wd] = 0;
            if (lsum) lsum[((pids[i] - gb) >> gsh) >> 5] = 0;  // every set bit came from a pid
        }
        __syncthreads();
    }
    // sets every summary word (all ones: any id may be live; zero: the all-clear state)
    __device__ __forceinline__ void fill_live_summary(uint32_t v) {
        __syncthreads();
        for (int i = tid; i < SUMW; i += NT) lsum[i] = v;
        __syncthreads();
    }
};

// Register budget: DQC_WAVES_PER_EU asks the compiler for that many waves per SIMD.
#ifndef DQC_WAVES_PER_EU
#define DQC_WAVES_PER_EU 8
#endif
// TPB threads per task (a 256-thread workgroup, or one wavefront) and LKW words per LDS
// id bitmap (the LDS-bitmap variant covers condensed ids below 32 * LKW).
// The wavefront kernel's LDS (8 KB per wavefront) allows 5 waves per SIMD, and its
// register budget is set for that many.
template <bool BIG, int TPB = NT, int LKW = KW>
__global__ __launch_bounds__(TPB)
__attribute__((amdgpu_waves_per_eu(TPB == WAVE_T ? 5 : TPB == SPLIT_T ? 4 : DQC_WAVES_PER_EU)))
void expand_kernel(KParams P) {
    if (P.in_ntask == ~0ull && blockIdx.x >= P.q->wg_limit) return;  // no stack room for them
    __shared__ Shared S;
    W<TPB> w;
    w.P = (KPtr)__builtin_amdgcn_kernarg_segment_ptr();  // the explicit arguments start the segment
    w.zw = zero_words;
    w.S = &S;
    w.tid = threadIdx.x;
    w.kw = (P.kloc + 31) >> 5;
    w.bytes = 0;
    for (int i = 0; i < 8; ++i) w.ph[i] = 0;
    w.ph_t = P.profile ? __builtin_amdgcn_s_memrealtime() : 0;
    const unsigned long long k = (unsigned long long)P.kloc;
    uint32_t* base = P.scratch + (unsigned long long)blockIdx.x * P.stride;
    w.vbase = base;
    w.set_graph(0);
    int32_t* lists = (int32_t*)(base + 10 * k);
    w.pids = lists; w.mem = lists + k; w.la = lists + 2 * k; w.lb = lists + 3 * k; w.lc = lists + 4 * k;
    w.lst = lists + 5 * k; w.crit = lists + 6 * k; w.gat = lists + 7 * k;
    w.hsort = (unsigned long long*)(lists + 8 * k);
    uint32_t* bm;
    if constexpr (BIG) {
        bm = base + 22 * k;  // kept all-zero between tasks; zeroed when the graph is uploaded
        __shared__ uint32_t lds_sum[SUMW];
        for (int i = threadIdx.x; i < SUMW; i += TPB) lds_sum[i] = 0;
        w.lsum = lds_sum;
        int gs = P.sum_shift;  // ids per summary bit: 2^gs, so that every graph-relative id has a bit
        while (((unsigned long long)(P.kloc > 0 ? P.kloc - 1 : 0) >> gs) >= 32ull * SUMW) ++gs;
        w.gsh = gs;
    } else {
        __shared__ uint32_t lds_bm[NBITMAPS * LKW];
        bm = lds_bm;
        for (int i = threadIdx.x; i < NBITMAPS * LKW; i += TPB) lds_bm[i] = 0;
        w.lsum = nullptr;
        w.gsh = 0;
    }
    const int bw = BIG ? w.kw : LKW;
    w.plive = bm; w.pmem = bm + bw; w.cmem = bm + 2 * bw; w.ccand = bm + 3 * bw; w.cL = bm + 4 * bw; w.ccrit = bm + 5 * bw;
    w.dirty_m = 0;
    w.dirty_l = 0;
    w.crit_marks = false;
    w.cand_marks = false;
    w.lazy_v = -1;
    w.id_lo = INT_MIN; w.id_hi = INT_MAX;
    w.clip_pair = 0;
    w.launch_steps = 0;
    w.t_start = __builtin_amdgcn_s_memrealtime();
    w.discard_child = false;
    if (threadIdx.x == 0) { S.acc_bytes = 0; S.cnt[9] = 0; }
    __syncthreads();
    while (true) {
        unsigned long long t = 0;
        if (threadIdx.x == 0) t = atomicAdd(&P.q->next_task, 1ull);
        t = w.bcast0(t);
        if (t >= (P.in_ntask == ~0ull ? P.q->ntask : P.in_ntask)) break;
        if (P.split_s) {
            const TaskDesc d = P.in_desc[0];
            w.set_graph(d.gid);
            w.cur_uid = d.uid;
            w.cur_depth = d.depth;
            // level-0 branches keep no live bitmap (the live set is an id range)
            if constexpr (BIG) w.fill_live_summary(~0u);
            w.discard_child = P.split_s[t] < P.root_first;
            w.branch_root(P.split_s[t], d.tot, P.in_vtx + d.off);
            w.discard_child = false;
            if constexpr (BIG) w.fill_live_summary(0u);
        } else {
            TaskDesc d = P.in_desc[t];
            w.run_task(d, P.in_vtx + d.off);
        }
    }
    atomicAdd(&S.acc_bytes, w.bytes);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&P.g->bytes, S.acc_bytes);
        if (P.profile) {
            w.tick(0);
            for (int i = 0; i < 8; ++i) atomicAdd(&P.g->phase[i], w.ph[i]);
        }
    }
}

// ---------------------------------------------------------------- device-resident steps
// One expansion step without host round trips: pop_kernel chooses the batch (the same
// rules as the host pop: at most pool_tasks newest tasks whose records fit the batch
// buffer and whose worst-case children fit the stack), the optional ordering kernels
// permute it, gather_kernel copies it into the batch buffer, expand_kernel runs it.

struct PopParams {
    TaskDesc* sdesc;
    QCtr* q;
    GCtr* g;
    unsigned long long pool_tasks, pool_vertices, spill_tasks, spill_vertices;
    unsigned long long workgroups;
};

constexpr int POP_T = 1024;

// Single workgroup. The batch is the newest tasks of the main stack, or of the deferred
// region when the main stack is empty: at most pool_tasks, as many as fit the batch
// buffer (the newest one always does, or the run stops with an error). The launch gets
// as many workgroups as the stack has room for one child each; with no room for the
// whole batch, its newest task alone is tried.
__global__ __launch_bounds__(POP_T) void pop_kernel(PopParams pp) {
    __shared__ int first_bad;
    __shared__ unsigned maxtot;
    const QCtr qv = *pp.q;
    const unsigned long long st_t = qv.sp_alloc >> CNT_SHIFT, st_v = qv.sp_alloc & SIZE_MASK;
    const unsigned long long df_t = qv.def_alloc >> CNT_SHIFT, df_v = qv.def_alloc & SIZE_MASK;
    const bool from_main = st_t > 0;
    const unsigned long long avail = from_main ? st_t : df_t;
    const unsigned long long m = avail < pp.pool_tasks ? avail : pp.pool_tasks;
    // the j-th newest task, and the records from the newest task through it
    auto idx = [&](unsigned long long j) { return from_main ? st_t - 1 - j : pp.spill_tasks - df_t + j; };
    auto span = [&](const TaskDesc& d) {
        return from_main ? st_v - d.off : d.off + (unsigned long long)d.tot - (pp.spill_vertices - df_v);
    };
    const int tid = threadIdx.x;
    if (tid == 0) { first_bad = -1; maxtot = 0; }
    __syncthreads();
    for (unsigned long long c0 = 0; c0 < m; c0 += POP_T) {
        const unsigned long long j = c0 + tid;
        if (j < m && j > 0 && span(pp.sdesc[idx(j)]) > pp.pool_vertices)  // records must fit the batch buffer
            atomicMin((unsigned*)&first_bad, (unsigned)j);  // -1 = none
        __syncthreads();
        const bool stop = first_bad >= 0;  // read by every thread before any next update
        __syncthreads();
        if (stop) break;
    }
    {
        const unsigned long long cnt = first_bad >= 0 ? (unsigned long long)first_bad : m;
        for (unsigned long long j = tid; j < cnt; j += POP_T) atomicMax(&maxtot, (unsigned)pp.sdesc[idx(j)].tot);
        __syncthreads();
    }
    if (tid == 0) {
        unsigned long long cnt = first_bad >= 0 ? (unsigned long long)first_bad : m;
        // nothing runs after an error, and a newest task larger than the batch buffer is one
        if (pp.g->error) cnt = 0;
        else if (m > 0 && span(pp.sdesc[idx(0)]) > pp.pool_vertices) {
            atomicOr(&pp.g->error, (unsigned)ERR_BATCH);
            cnt = 0;
        }
        unsigned long long wgl = pp.workgroups, mtot = maxtot;
        if (cnt > 0) {
            const unsigned long long room_t = pp.spill_tasks - st_t - df_t, free_v = pp.spill_vertices - st_v - df_v;
            wgl = min(wgl, min(free_v / (mtot ? mtot : 1), room_t));
            if (wgl == 0) {  // no room for the whole batch: the newest task alone, if it fits
                cnt = 1;
                mtot = (unsigned long long)pp.sdesc[idx(0)].tot;
                wgl = min(pp.workgroups, min(free_v / (mtot ? mtot : 1), room_t));
            }
            if (wgl == 0) {
                atomicOr(&pp.g->error, (unsigned)ERR_SPILL);
                cnt = 0;
            }
            // under stack pressure, two tasks per workgroup (a full batch would mostly be
            // deferred untouched)
            if (wgl < pp.workgroups && cnt > 2 * wgl) cnt = 2 * wgl;
        }
        if (cnt > 0) atomicAdd(&pp.g->steps, 1ull);
        QCtr& q = *pp.q;
        q.next_task = 0;
        q.ntask = cnt;
        q.maxtot = mtot;
        q.wg_limit = wgl;
        q.pend = 0;
        // piece-record budget of the split marks: a quarter of the free stack records, as
        // in the host pop (use counted before the pop)
        {
            const unsigned long long used_v = st_v + df_v;
            q.split_budget = pp.spill_vertices > used_v ? (pp.spill_vertices - used_v) / 4 : 0;
        }
        if (cnt > 0) {
            // first / v0: the batch's lowest descriptor index and record; both regions keep a
            // batch contiguous (gather_kernel copies [v0, last.off + last.tot))
            const unsigned long long first = from_main ? st_t - cnt : pp.spill_tasks - df_t;
            const unsigned long long v0 = pp.sdesc[first].off;
            const TaskDesc last = pp.sdesc[first + cnt - 1];
            const unsigned long long recs = last.off + (unsigned long long)last.tot - v0;
            q.first = first;
            q.v0 = v0;
            if (from_main) q.sp_alloc = (first << CNT_SHIFT) | v0;
            else q.def_alloc = ((df_t - cnt) << CNT_SHIFT) | (df_v - recs);
            q.pend = (cnt << CNT_SHIFT) | recs;
        }
        q.used = q.sp_alloc + q.def_alloc;  // field-wise sum: neither field carries
    }
}

// Steps a task or piece runs (a piece: its range; a task: all but the covered prefix).
__host__ __device__ inline int task_steps(const TaskDesc& d) {
    return (d.flags & TF_PIECE) ? d.iters : d.tot - d.nm - (d.mode == MODE_DEVICE ? 0 : d.ncov);
}
// Least steps of a task that SIZE / LEARNED order a batch for: the split threshold, or
// 1024 when splitting is off. A batch of shorter tasks keeps stack order, which keeps
// siblings and the tasks of one graph together; ordering it costs locality and gains
// no balance.
__host__ __device__ inline int long_task_steps(int split_b) { return split_b > 0 ? split_b : 1024; }
// Batch order for the SIZE / LEARNED schedules, one workgroup: tasks go into 64 buckets
// of descending log2 cost (cost: SIZE = steps x size, LEARNED = the model's prediction
// x size, the prediction already in pred[]), keeping stack order within a bucket, and in
// that order the eligible tasks are marked for splitting while the running total of
// their piece records fits q->split_budget (a prefix rule: the first task that does not
// fit ends the marking, as in the host pop). A marked task is cut into pieces_for(its
// cost, the batch's total cost) pieces. A batch without a task of long_task_steps
// steps keeps stack order and splits nothing. perm[k] = the batch index of the k-th
// task to run; split[k] = its piece length when it is to be split, else 0.
__global__ __launch_bounds__(POP_T) void order_kernel(const TaskDesc* __restrict__ sdesc, const QCtr* __restrict__ q,
                                                      const float* __restrict__ pred, int learned, int split_b,
                                                      int split_c, int max_pieces, int workgroups,
                                                      uint32_t* __restrict__ perm, uint32_t* __restrict__ split) {
    __shared__ unsigned hist[64], start[64];
    __shared__ unsigned wcnt[POP_T / 64][64];
    __shared__ unsigned long long scan[POP_T];
    __shared__ unsigned long long carry;
    __shared__ int any_long;
    __shared__ double csum[POP_T];
    const unsigned long long n = q->ntask, first = q->first, budget = q->split_budget;
    const int tid = threadIdx.x;
    if (tid < 64) hist[tid] = 0;
    if (tid == 0) { carry = 0; any_long = 0; }
    __syncthreads();
    auto cost = [&](unsigned long long i) {
        const TaskDesc d = sdesc[first + i];
        const double steps = (double)max(1, d.tot - d.nm);
        if (learned) {  // predicted branches x task size
            const double share = (d.flags & TF_PIECE) ? fmin(1.0, (double)d.iters / steps) : 1.0;
            return fmax(expm1((double)fminf(pred[i], 60.f)) * share * (double)d.tot, 0.0);
        }
        return ((d.flags & TF_PIECE) ? (double)d.iters : steps) * (double)d.tot;
    };
    auto bucket = [&](unsigned long long i) {
        int b = (int)(2.0 * log2(1.0 + cost(i)));
        return 63 - min(max(b, 0), 63);  // bucket 0 = costliest
    };
    double part = 0.0;
    for (unsigned long long i = tid; i < n; i += POP_T) {
        atomicAdd(&hist[bucket(i)], 1u);
        part += cost(i);
        const TaskDesc d = sdesc[first + i];
        if (task_steps(d) >= long_task_steps(split_b)) any_long = 1;
    }
    csum[tid] = part;
    __syncthreads();
    for (int off = POP_T / 2; off > 0; off >>= 1) {
        if (tid < off) csum[tid] += csum[tid + off];
        __syncthreads();
    }
    const double total = csum[0];
    if (!any_long) {  // no task long enough to split or to unbalance the launch: stack order
        for (unsigned long long i = tid; i < n; i += POP_T) { perm[i] = (uint32_t)i; split[i] = 0; }
        return;
    }
    if (tid == 0) {
        unsigned a = 0;
        for (int b = 0; b < 64; ++b) { start[b] = a; a += hist[b]; }
    }
    __syncthreads();
    // stable placement: within a bucket, tasks keep their stack order (chunks of POP_T in
    // index order; inside a chunk, each wave ranks its lanes by ballot, and the waves'
    // shares follow one another)
    const int wave = tid >> 6, lane = tid & 63;
    for (unsigned long long c0 = 0; c0 < n; c0 += POP_T) {
        const unsigned long long i = c0 + tid;
        const int b = i < n ? bucket(i) : 64;
        for (int j = tid; j < (POP_T / 64) * 64; j += POP_T) wcnt[j >> 6][j & 63] = 0;
        __syncthreads();
        unsigned long long rem = __ballot(b < 64);
        unsigned rank = 0;
        while (rem) {
            const int lb = __shfl(b, __builtin_ctzll(rem));
            const unsigned long long m = __ballot(b == lb);
            if (b == lb) rank = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) wcnt[wave][lb] = (unsigned)__popcll(m);
            rem &= ~m;
        }
        __syncthreads();
        if (tid < 64) {  // bucket tid: where each wave's share starts
            unsigned a = start[tid];
            for (int w = 0; w < POP_T / 64; ++w) {
                const unsigned c = wcnt[w][tid];
                wcnt[w][tid] = a;
                a += c;
            }
            start[tid] = a;
        }
        __syncthreads();
        if (b < 64) perm[wcnt[wave][b] + rank] = (uint32_t)i;
        __syncthreads();
    }
    // split marks in run order: eligible tasks while the running piece records fit
    for (unsigned long long c0 = 0; c0 < n; c0 += POP_T) {
        const unsigned long long k = c0 + tid;
        unsigned long long need = 0;
        int chunk = 0;
        if (k < n && split_b > 0) {
            const TaskDesc d = sdesc[first + perm[k]];
            const int steps = d.tot - d.nm - (d.mode == MODE_DEVICE ? 0 : d.ncov);
            if (split_eligible(d, split_b)) {
                chunk = split_chunk_for(steps, split_c, pieces_for(cost(perm[k]), total, workgroups, max_pieces));
                need = (unsigned long long)((steps + chunk - 1) / chunk) * (unsigned long long)d.tot;
            }
        }
        scan[tid] = need;
        __syncthreads();
        for (int off = 1; off < POP_T; off <<= 1) {
            unsigned long long a = tid >= off ? scan[tid - off] : 0;
            __syncthreads();
            scan[tid] += a;
            __syncthreads();
        }
        if (k < n) split[k] = (need > 0 && carry + scan[tid] <= budget) ? (uint32_t)chunk : 0u;
        __syncthreads();
        if (tid == POP_T - 1) carry += scan[tid];
        __syncthreads();
    }
}

// Copies the popped batch into the batch buffer: descriptors rebased (in perm order with
// split marks when perm is given) and the records.
__global__ void gather_kernel(const TaskDesc* __restrict__ sdesc, const dqc::TaskVertex* __restrict__ svtx,
                              const QCtr* __restrict__ q, const uint32_t* __restrict__ perm,
                              const uint32_t* __restrict__ split, TaskDesc* __restrict__ bdesc,
                              dqc::TaskVertex* __restrict__ bvtx) {
    const unsigned long long n = q->ntask;
    if (n == 0) return;
    const unsigned long long first = q->first, v0 = q->v0;
    const TaskDesc last = sdesc[first + n - 1];
    const unsigned long long vend = last.off + (unsigned long long)last.tot;
    const unsigned long long gid = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long gsz = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long k = gid; k < n; k += gsz) {
        TaskDesc d = sdesc[first + (perm ? perm[k] : k)];
        d.off -= v0;
        if (perm && split[k]) {
            d.flags |= TF_SPLIT;
            d.iters = (int32_t)split[k];  // piece length
        }
        bdesc[k] = d;
    }
    for (unsigned long long i = gid; i < vend - v0; i += gsz) bvtx[i] = svtx[v0 + i];
}

// Predicted log(1 + subtree branches) of tasks d[0..n) (the learned schedule's estimate);
// the model is staged in LDS, one task per thread.
// With q given, the tasks are the current device-resident batch on the stack.
__global__ __launch_bounds__(256) void desc_cost_kernel(const float* __restrict__ params, const TaskDesc* __restrict__ d,
                                                        unsigned long long n, float* __restrict__ y,
                                                        const QCtr* __restrict__ q) {
    __shared__ float p[dqc_cost::NPARAM];
    if (q) {
        n = q->ntask;
        d += q->first;
    }
    for (int i = threadIdx.x; i < dqc_cost::NPARAM; i += blockDim.x) p[i] = params[i];
    __syncthreads();
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const TaskDesc t = d[i];
        float x[dqc_cost::IN];
        dqc_cost::features(t.nm, t.tot, t.depth, t.dmin_o, t.dmax_o, t.dmin_i, t.dmax_i, x);
        y[i] = dqc_cost::forward(p, x);
    }
}

// Copies the records of tasks [first, first + m) of the task stack into the batch buffer
// (and their descriptors, rebased, when dst_desc is given).
__global__ void refill_kernel(const TaskDesc* sp_desc, const dqc::TaskVertex* sp_vtx,
                              unsigned long long first, unsigned long long m,
                              TaskDesc* dst_desc, dqc::TaskVertex* dst_vtx,
                              unsigned long long dst_t0, unsigned long long dst_v0) {
    const unsigned long long v0 = sp_desc[first].off;
    const unsigned long long vend = sp_desc[first + m - 1].off + (unsigned long long)sp_desc[first + m - 1].tot;
    const unsigned long long gid = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long gsz = (unsigned long long)gridDim.x * blockDim.x;
    if (dst_desc)
        for (unsigned long long i = gid; i < m; i += gsz) {
            TaskDesc d = sp_desc[first + i];
            d.off = d.off - v0 + dst_v0;
            dst_desc[dst_t0 + i] = d;
        }
    for (unsigned long long i = gid; i < vend - v0; i += gsz) dst_vtx[dst_v0 + i] = sp_vtx[v0 + i];
}

// Per-step outcome of the level-0 branch loop (reference src/host_functions.cu:1014-1044):
// bit0 = the removal at this step skips the branch (success 2), bit1 = lookahead holds,
// bit2 = fewer than min_size vertices remain (the loop stops). Uses the same counts as
// branch_root: after s removals the live set is [0, k - s).
__global__ void root_flags_kernel(KParams P, const dqc::TaskVertex* root, int k, int S, uint8_t* flags) {
    // S: the steps to classify, at most k - ncov (fewer in a bounded unit of work)
    const int mo1 = P.mdo[P.min_size], mi1 = P.mdi[P.min_size];  // gmd(nm + 1) with nm = 0
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S; s += gridDim.x * blockDim.x) {
        const int tot = k - s;
        uint8_t f = 0;
        if (s >= 1) {
            if (tot < P.min_size) { flags[s] = 4; continue; }
            const int p = k - s, q = k - s - 1;
            bool skip = false;
            if (contains(P.on, P.oo[p], P.oo[p + 1], q) && cnt_lt(P.in, P.io[q], P.io[q + 1], tot) < (uint32_t)mi1) skip = true;
            if (contains(P.in, P.io[p], P.io[p + 1], q) && cnt_lt(P.on, P.oo[q], P.oo[q + 1], tot) < (uint32_t)mo1) skip = true;
            if (skip) { flags[s] = 1; continue; }
        }
        const uint32_t mo = (uint32_t)(tot < P.min_size ? P.mdo[P.min_size] : P.mdo[tot]);
        const uint32_t mi = (uint32_t)(tot < P.min_size ? P.mdi[P.min_size] : P.mdi[tot]);
        const uint32_t l2min = (uint32_t)(tot - 1);
        bool la = true;
        for (int i = 0; i < tot && la; ++i) {
            if (cnt_lt(P.on, P.oo[i], P.oo[i + 1], tot) < mo || cnt_lt(P.in, P.io[i], P.io[i + 1], tot) < mi) { la = false; break; }
            const unsigned long long l0 = P.lto[i], l1 = P.lto[i + 1];
            uint32_t l2 = root[i].lvl2adj - ((uint32_t)(l1 - l0) - cnt_lt(P.ltn, l0, l1, tot));
            if (l2 < l2min) la = false;
        }
        if (la) f |= 2;
        flags[s] = f;
    }
}

}  // namespace

// =====================================================================================
// Host side
struct dqc_handle {
    int device = 0;
    int workgroups = 0;
    dqc_config cfg{};
    std::string err;
    // graph
    int32_t n = 0;
    int32_t kloc = 0;    // largest vertex count of one graph
    int ngraphs = 0;     // graphs in the id space (batched runs: more than one)
    int32_t* d_gbase = nullptr;
    std::vector<int32_t> gbase;
    bool big = false;    // condensed graph above KMAX vertices: HBM id bitmaps
    int launch_wg = 0;   // workgroups per launch (the scratch bounds it for large graphs)
    int wave_wg = 0;     // resident single-wavefront workgroups of the wavefront-per-task kernel
    int split_wg = 0;    // resident workgroups of the split-pass kernel
    unsigned long long batch_nsplit = 0;  // the popped batch's first batch_nsplit tasks are split passes
    unsigned long long batch_pend_rest = 0;  // (tasks << 34) | records of the batch after them
    bool batch_wave = false;  // the popped batch runs one task per wavefront
    int64_t wave_launches = 0;
    unsigned long long *d_oo = nullptr, *d_io = nullptr, *d_to = nullptr;
    int32_t *d_on = nullptr, *d_in = nullptr, *d_tn = nullptr;
    int32_t *d_mdo = nullptr, *d_mdi = nullptr, *d_ubo = nullptr, *d_ubi = nullptr;
    int32_t tlen = 0, min_size = 0;
    // batch buffer (tasks expanded by one launch) and the LIFO task stack
    TaskDesc* d_bdesc = nullptr;
    dqc::TaskVertex* d_bvtx = nullptr;
    TaskDesc* d_sdesc = nullptr;
    dqc::TaskVertex* d_svtx = nullptr;
    unsigned long long st_t = 0, st_v = 0;  // stack depth: tasks, records
    unsigned long long def_t = 0, def_v = 0;  // deferred region (from the top of the arrays down)
    unsigned long long peak_t = 0, peak_v = 0;  // largest stack use seen between launches
    // cliques
    int32_t* d_clq_ids = nullptr;
    unsigned long long* d_clq_meta = nullptr;
    QCtr* d_q = nullptr;
    GCtr* d_g = nullptr;
    unsigned long long* d_lto = nullptr;
    int32_t* d_ltn = nullptr;
    uint8_t* d_flags = nullptr;
    int32_t* d_split = nullptr;
    std::vector<std::vector<int32_t>> host_cliques;  // level-0 lookahead clique, if any
    uint32_t* d_scratch = nullptr;
    unsigned long long stride = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    dqc_stats st{};
    unsigned long long max_depth = 0;
    unsigned long long last_root_first = ~0ull;
    int64_t extra_tasks = 0;
    bool root_pending = false;
    int profile = 0;
    int sum_shift = 0;  // DQC_LIVE_SUMMARY_SHIFT: coarser live summary (tests of the HBM variant)
    unsigned long long phase[8] = {0};
    int64_t host_launches = 0;
    unsigned long long batch_maxtot = 0;  // largest task of the host-popped batch
    unsigned long long batch_pend = 0;    // its (tasks << 34) | records
    unsigned long long batch_grid = 0;    // workgroups the stack has room for
    std::vector<TaskDesc> hbatch;  // scheduled batch descriptors (host copy for the upload)
    std::vector<double> hcost;
    TaskLog* d_log = nullptr;      // task log (dqc_set_task_log)
    unsigned long long log_cap = 0;
    // device-resident steps: one captured HIP graph of resident_steps steps per schedule
    hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t* d_perm = nullptr;    // batch order (SIZE / LEARNED)
    uint32_t* d_mark = nullptr;    // split marks in that order (piece length, 0 = whole)
    float* d_model = nullptr;      // cost model (dqc_set_cost_model)
    float* d_cost = nullptr;       // per-task predictions of the current batch
    std::vector<float> hpred;
    std::vector<void*> allocs;
    // batch buffer and task stack capacities left to the library (dqc_config field 0):
    // sized from the graph when it is uploaded
    bool auto_pool_v = false, auto_spill_t = false, auto_spill_v = false;
    unsigned long long cap_pool_v = 0, cap_spill_t = 0, cap_spill_v = 0;  // allocated
};

namespace {

bool hip_ok(dqc_handle* h, hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    h->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

template <class T>
bool dev_alloc(dqc_handle* h, T** p, size_t count, const char* what) {
    void* q = nullptr;
    if (!hip_ok(h, hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)), what)) return false;
    h->allocs.push_back(q);
    *p = (T*)q;
    return true;
}

void drop_graphs(dqc_handle* h) {
    for (auto& e : h->gexec)
        if (e) {
            hipGraphExecDestroy(e);
            e = nullptr;
        }
}

void free_graph(dqc_handle* h) {
    void* gp[] = {h->d_oo, h->d_io, h->d_to, h->d_on, h->d_in, h->d_tn, h->d_scratch,
                  h->d_lto, h->d_ltn, h->d_flags, h->d_split, h->d_gbase};
    for (void* p : gp)
        if (p) {
            hipFree(p);
            h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), p), h->allocs.end());
        }
    h->d_oo = h->d_io = h->d_to = nullptr;
    h->d_on = h->d_in = h->d_tn = nullptr;
    h->d_scratch = nullptr;
    h->d_lto = nullptr; h->d_ltn = nullptr; h->d_flags = nullptr; h->d_split = nullptr;
    h->d_gbase = nullptr;
}

// Descriptor of a task pushed from the host, with the degree features the kernels
// compute for the children they write.
TaskDesc host_desc(const uint32_t* records, int32_t nvert, int32_t nm, int32_t ncov, int32_t mode,
                   unsigned long long off, uint32_t uid, uint32_t depth) {
    TaskDesc d{};
    d.off = off;
    d.tot = nvert; d.nm = nm; d.ncov = ncov; d.mode = mode;
    d.iters = nvert - nm - ncov; d.flags = 0;
    d.uid = uid; d.parent = UINT32_MAX; d.depth = (uint16_t)std::min<uint32_t>(depth, 65535u);
    uint32_t omin = UINT32_MAX, omax = 0, imin = UINT32_MAX, imax = 0;
    for (int32_t i = 0; i < nvert; ++i) {
        const uint32_t* r = records + 6 * i;  // id, out_mem, out_can, in_mem, in_can, lvl2adj
        const uint32_t od = r[1] + r[2], id = r[3] + r[4];
        omin = std::min(omin, od); omax = std::max(omax, od);
        imin = std::min(imin, id); imax = std::max(imax, id);
    }
    d.dmin_o = (uint16_t)std::min<uint32_t>(omin, 65535u); d.dmax_o = (uint16_t)std::min<uint32_t>(omax, 65535u);
    d.dmin_i = (uint16_t)std::min<uint32_t>(imin, 65535u); d.dmax_i = (uint16_t)std::min<uint32_t>(imax, 65535u);
    return d;
}

}  // namespace

extern "C" {

dqc_status dqc_open(const dqc_config* cfg, dqc_handle** out) {
    if (!out) return DQC_E_ARG;
    dqc_handle* h = new dqc_handle();
    if (cfg) h->cfg = *cfg;
    else h->cfg.host_addfail_check = 1;
    dqc_config& c = h->cfg;
    h->auto_pool_v = c.pool_vertices == 0;
    h->auto_spill_t = c.spill_tasks == 0;
    h->auto_spill_v = c.spill_vertices == 0;
    if (!c.pool_vertices) c.pool_vertices = 1ull << 27;
    if (!c.spill_tasks) c.spill_tasks = 1ull << 25;
    if (!c.spill_vertices) c.spill_vertices = 1ull << 30;
    if (!c.clique_sets) c.clique_sets = 1ull << 25;
    if (!c.clique_ids) c.clique_ids = 1ull << 29;
    if (!c.split_branches) c.split_branches = 1024;
    if (!c.wave_tasks) c.wave_tasks = DQC_WAVE_TASKS_DEFAULT;
    if (c.split_chunk <= 0) c.split_chunk = 32;
    if (c.schedule == DQC_SCHED_DEFAULT) c.schedule = DQC_SCHED_SIZE;
    h->device = c.device;
    *out = h;
    if (c.spill_tasks > MAX_ENTRIES || c.clique_sets > MAX_ENTRIES || c.spill_vertices > MAX_RECORDS ||
        c.clique_ids > MAX_RECORDS || c.pool_vertices > MAX_RECORDS || c.schedule < DQC_SCHED_FIFO ||
        c.schedule > DQC_SCHED_LEARNED) {
        h->err = "capacity above 2^30 tasks / cliques or 2^33 records / ids, or unknown schedule";
        return DQC_E_ARG;
    }
    h->profile = std::getenv("DQC_PHASES") != nullptr;
    if (const char* e = std::getenv("DQC_LIVE_SUMMARY_SHIFT")) h->sum_shift = std::max(0, std::min(20, std::atoi(e)));
    *out = h;
    if (!hip_ok(h, hipSetDevice(h->device), "hipSetDevice")) return DQC_E_HIP;
    int cus = 0;
    if (!hip_ok(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device), "device query"))
        return DQC_E_HIP;
    int per_cu = 0;
    if (!hip_ok(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, expand_kernel<false>, NT, 0), "occupancy"))
        return DQC_E_HIP;
    if (per_cu < 1) per_cu = 1;
    h->workgroups = c.workgroups > 0 ? c.workgroups : cus * per_cu;
    int per_cu_w = 0;
    if (!hip_ok(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_w, expand_kernel<false, WAVE_T, WAVE_KW>, WAVE_T, 0),
                "occupancy"))
        return DQC_E_HIP;
    h->wave_wg = cus * std::max(per_cu_w, 1);
    int per_cu_s = 0;
    if (!hip_ok(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_s, expand_kernel<true, SPLIT_T>, SPLIT_T, 0),
                "occupancy"))
        return DQC_E_HIP;
    h->split_wg = cus * std::max(per_cu_s, 1);
    if (!c.pool_tasks) c.pool_tasks = 32ull * (unsigned long long)h->workgroups;
    // non-blocking stream: every host read of its results synchronizes on it explicitly
    if (!hip_ok(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate")) return DQC_E_HIP;
    if (!hip_ok(h, hipEventCreate(&h->ev0), "hipEventCreate")) return DQC_E_HIP;
    if (!hip_ok(h, hipEventCreate(&h->ev1), "hipEventCreate")) return DQC_E_HIP;
    if (!dev_alloc(h, &h->d_bdesc, c.pool_tasks, "batch alloc")) return DQC_E_HIP;
    // the batch records and the task stack are allocated with the graph (size_buffers)
    if (!dev_alloc(h, &h->d_clq_ids, c.clique_ids, "clique alloc")) return DQC_E_HIP;
    if (!dev_alloc(h, &h->d_clq_meta, c.clique_sets, "clique alloc")) return DQC_E_HIP;
    if (!dev_alloc(h, &h->d_q, 1, "counter alloc") || !dev_alloc(h, &h->d_g, 1, "counter alloc"))
        return DQC_E_HIP;
    if (!hip_ok(h, hipMemset(h->d_q, 0, sizeof(QCtr)), "hipMemset") ||
        !hip_ok(h, hipMemset(h->d_g, 0, sizeof(GCtr)), "hipMemset"))
        return DQC_E_HIP;
    h->st.workgroups = h->workgroups;
    return DQC_OK;
}

dqc_status dqc_close(dqc_handle* h) {
    if (!h) return DQC_E_ARG;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    drop_graphs(h);
    for (void* p : h->allocs) hipFree(p);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return DQC_OK;
}

const char* dqc_last_error(const dqc_handle* h) { return h ? h->err.c_str() : "null handle"; }

}  // extern "C"

// The batch-record buffer and the task stack, sized from the largest graph (kloc) unless
// dqc_config set them: a stack of 8192 x kloc records (within [2^22, 2^30]; a task holds
// at most kloc records and the depth-first pop keeps the stack near the enumeration
// depth times the batch), tasks 1/16 of that, a batch buffer of 1024 x kloc records
// (within [2^20, 2^27]). Small graphs then take a few hundred MB instead of tens of GB.
static dqc_status size_buffers(dqc_handle* h) {
    dqc_config& c = h->cfg;
    const unsigned long long k = (unsigned long long)std::max(h->kloc, 1);
    auto clampu = [](unsigned long long x, unsigned long long lo, unsigned long long hi) {
        return std::max(lo, std::min(hi, x));
    };
    if (h->auto_spill_v) c.spill_vertices = clampu(8192ull * k, 1ull << 22, 1ull << 30);
    if (h->auto_spill_t) c.spill_tasks = clampu(c.spill_vertices / 16, 1ull << 16, 1ull << 25);
    if (h->auto_pool_v) c.pool_vertices = clampu(1024ull * k, 1ull << 20, 1ull << 27);
    auto realloc_buf = [&](auto** p, unsigned long long& cap, unsigned long long want, const char* what) {
        if (*p && cap == want) return true;
        if (*p) {
            hipFree(*p);
            h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), (void*)*p), h->allocs.end());
            *p = nullptr;
        }
        cap = 0;
        if (!dev_alloc(h, p, want, what)) return false;
        cap = want;
        return true;
    };
    const bool moved = h->cap_spill_t != c.spill_tasks || h->cap_spill_v != c.spill_vertices;
    if (!realloc_buf(&h->d_bvtx, h->cap_pool_v, c.pool_vertices, "batch alloc") ||
        !realloc_buf(&h->d_sdesc, h->cap_spill_t, c.spill_tasks, "stack alloc") ||
        !realloc_buf(&h->d_svtx, h->cap_spill_v, c.spill_vertices, "stack alloc"))
        return DQC_E_HIP;
    if (moved) h->st_t = h->st_v = h->def_t = h->def_v = 0;  // a new stack holds no task
    return DQC_OK;
}

// The wavefront-per-task kernel may run: enabled, and every id fits its LDS bitmaps.
static bool wave_ok(const dqc_handle* h) { return h->cfg.wave_tasks > 0 && !h->big && h->kloc <= WAVE_KMAX; }
// A batch whose largest task has maxtot vertices runs one task per wavefront.
static bool wave_batch(const dqc_handle* h, unsigned long long maxtot) {
    return wave_ok(h) && maxtot <= (unsigned long long)h->cfg.wave_tasks;
}
// Device-resident steps fix the kernel when they are captured: the wavefront kernel when
// every task of the graph (at most kloc vertices) is small enough for it.
static bool wave_resident(const dqc_handle* h) { return wave_ok(h) && h->kloc <= h->cfg.wave_tasks; }
// Per-workgroup scratch slots: one per workgroup of the larger of the two grids.
static unsigned long long scratch_slots(const dqc_handle* h) {
    return (unsigned long long)std::max(h->launch_wg, wave_ok(h) ? h->wave_wg : 0);
}

// Uploads n ids' CSR lists; kloc is the largest vertex count of one graph (n itself for
// a single graph; batched graphs share the id space with 32-aligned bases).
static dqc_status upload_impl(dqc_handle* h, int32_t n, int32_t kloc, const uint64_t* out_off, const int32_t* out_nbr,
                              const uint64_t* in_off, const int32_t* in_nbr, const uint64_t* two_off,
                              const int32_t* two_nbr) {
    hipSetDevice(h->device);
    drop_graphs(h);  // captured steps hold the old graph's pointers
    free_graph(h);
    h->n = n;
    h->kloc = std::max(kloc, 1);
    // id bitmaps no longer fit LDS: the kernel variant keeps them in HBM (DQC_HBM_BITMAPS=1
    // selects it for any size, which the tests use to run it on small graphs)
    h->big = kloc > KMAX || std::getenv("DQC_HBM_BITMAPS") != nullptr;
    size_t m_out = out_off[n], m_in = in_off[n], m_two = two_off[n];
    if (!dev_alloc(h, &h->d_oo, (size_t)n + 1, "graph alloc") || !dev_alloc(h, &h->d_io, (size_t)n + 1, "graph alloc") ||
        !dev_alloc(h, &h->d_to, (size_t)n + 1, "graph alloc") || !dev_alloc(h, &h->d_on, m_out, "graph alloc") ||
        !dev_alloc(h, &h->d_in, m_in, "graph alloc") || !dev_alloc(h, &h->d_tn, m_two, "graph alloc"))
        return DQC_E_HIP;
    bool ok = hip_ok(h, hipMemcpy(h->d_oo, out_off, sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice), "upload") &&
              hip_ok(h, hipMemcpy(h->d_io, in_off, sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice), "upload") &&
              hip_ok(h, hipMemcpy(h->d_to, two_off, sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice), "upload") &&
              (m_out == 0 || hip_ok(h, hipMemcpy(h->d_on, out_nbr, sizeof(int32_t) * m_out, hipMemcpyHostToDevice), "upload")) &&
              (m_in == 0 || hip_ok(h, hipMemcpy(h->d_in, in_nbr, sizeof(int32_t) * m_in, hipMemcpyHostToDevice), "upload")) &&
              (m_two == 0 || hip_ok(h, hipMemcpy(h->d_tn, two_nbr, sizeof(int32_t) * m_two, hipMemcpyHostToDevice), "upload"));
    if (!ok) return DQC_E_HIP;
    {
        // transposed two-hop lists (for each u, the vertices whose list holds u), sorted
        std::vector<uint64_t> lto((size_t)n + 1, 0);
        for (size_t e = 0; e < m_two; ++e) lto[two_nbr[e] + 1]++;
        for (int32_t u = 0; u < n; ++u) lto[u + 1] += lto[u];
        std::vector<int32_t> ltn(m_two);
        std::vector<uint64_t> cur(lto.begin(), lto.end() - 1);
        for (int32_t v = 0; v < n; ++v)
            for (uint64_t e = two_off[v]; e < two_off[v + 1]; ++e) ltn[cur[two_nbr[e]]++] = v;
        if (!dev_alloc(h, &h->d_lto, (size_t)n + 1, "graph alloc") || !dev_alloc(h, &h->d_ltn, m_two, "graph alloc") ||
            !dev_alloc(h, &h->d_flags, (size_t)n + 1, "graph alloc") || !dev_alloc(h, &h->d_split, (size_t)n + 1, "graph alloc"))
            return DQC_E_HIP;
        if (!hip_ok(h, hipMemcpy(h->d_lto, lto.data(), sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice), "upload") ||
            (m_two && !hip_ok(h, hipMemcpy(h->d_ltn, ltn.data(), sizeof(int32_t) * m_two, hipMemcpyHostToDevice), "upload")))
            return DQC_E_HIP;
    }
    // per-workgroup scratch: 10k words of id-indexed values (parent and child degree
    // records of 4 words per id, lvl2adj, critical counts) + 8 lists of k entries, the
    // HBM sort area (2k 64-bit keys: a child has at most k vertices, padded to a power of
    // 2) and, for large graphs, the six id bitmaps (all words even, so 8-byte aligned)
    const unsigned long long kk = (unsigned long long)h->kloc, kwords = (kk + 31) / 32;
    h->stride = 22ull * kk + (h->big ? (unsigned long long)NBITMAPS * kwords : 0ull);
    h->stride += h->stride & 1ull;
    // keep the scratch within 96 GiB (fewer resident workgroups for very large graphs)
    const unsigned long long cap_wg = (96ull << 30) / (h->stride * sizeof(uint32_t));
    h->launch_wg = (int)std::max<unsigned long long>(1, std::min<unsigned long long>(h->workgroups, cap_wg > 1 ? cap_wg - 1 : 1));
    h->st.workgroups = h->launch_wg;
    if (!dev_alloc(h, &h->d_scratch, h->stride * (size_t)(scratch_slots(h) + 1), "scratch alloc")) return DQC_E_HIP;
    if (h->big && !hip_ok(h, hipMemset(h->d_scratch, 0, sizeof(uint32_t) * h->stride * (size_t)(h->launch_wg + 1)), "scratch clear"))
        return DQC_E_HIP;
    h->st.n_condensed = n;
    return size_buffers(h);
}

extern "C" {

dqc_status dqc_upload_graph(dqc_handle* h, int32_t n, const uint64_t* out_off, const int32_t* out_nbr,
                            const uint64_t* in_off, const int32_t* in_nbr, const uint64_t* two_off,
                            const int32_t* two_nbr) {
    if (!h || n < 0 || !out_off || !in_off || !two_off) return DQC_E_ARG;
    h->ngraphs = 1;
    h->gbase.assign(1, 0);
    return upload_impl(h, n, n, out_off, out_nbr, in_off, in_nbr, two_off, two_nbr);
}

dqc_status dqc_upload_graph_batch(dqc_handle* h, int32_t ngraphs, const int32_t* nverts,
                                  const uint64_t* const* out_off, const int32_t* const* out_nbr,
                                  const uint64_t* const* in_off, const int32_t* const* in_nbr,
                                  const uint64_t* const* two_off, const int32_t* const* two_nbr) {
    if (!h || ngraphs < 1 || ngraphs > 65535 || !nverts || !out_off || !out_nbr || !in_off || !in_nbr || !two_off ||
        !two_nbr)
        return DQC_E_ARG;
    std::vector<int32_t> base((size_t)ngraphs);
    long long total = 0;
    int32_t kloc = 1;
    for (int32_t g = 0; g < ngraphs; ++g) {
        if (nverts[g] < 0) return DQC_E_ARG;
        base[g] = (int32_t)total;
        total += ((long long)nverts[g] + 31) / 32 * 32;
        kloc = std::max(kloc, nverts[g]);
    }
    if (total >= INT32_MAX) return DQC_E_UNSUPPORTED;
    const size_t N = (size_t)total;
    // union CSR: graph g's vertex v is id base[g] + v; padding ids have empty lists
    std::vector<uint64_t> oo(N + 1, 0), io(N + 1, 0), to(N + 1, 0);
    std::vector<int32_t> on, in, tn;
    auto append = [&](const uint64_t* off, const int32_t* nbr, int32_t nv, int32_t b, std::vector<uint64_t>& o,
                      std::vector<int32_t>& out) {
        for (int32_t v = 0; v < nv; ++v) {
            for (uint64_t e = off[v]; e < off[v + 1]; ++e) out.push_back(nbr[e] + b);
            o[(size_t)b + v + 1] = out.size();
        }
    };
    for (int32_t g = 0; g < ngraphs; ++g) {
        append(out_off[g], out_nbr[g], nverts[g], base[g], oo, on);
        append(in_off[g], in_nbr[g], nverts[g], base[g], io, in);
        append(two_off[g], two_nbr[g], nverts[g], base[g], to, tn);
    }
    for (size_t i = 1; i <= N; ++i) {  // padding ids continue the previous offset
        oo[i] = std::max(oo[i], oo[i - 1]);
        io[i] = std::max(io[i], io[i - 1]);
        to[i] = std::max(to[i], to[i - 1]);
    }
    dqc_status rc = upload_impl(h, (int32_t)N, kloc, oo.data(), on.data(), io.data(), in.data(), to.data(), tn.data());
    if (rc != DQC_OK) return rc;
    h->ngraphs = ngraphs;
    h->gbase = base;
    if (!dev_alloc(h, &h->d_gbase, (size_t)ngraphs, "graph alloc") ||
        !hip_ok(h, hipMemcpy(h->d_gbase, base.data(), sizeof(int32_t) * ngraphs, hipMemcpyHostToDevice), "upload"))
        return DQC_E_HIP;
    return DQC_OK;
}

dqc_status dqc_set_tables(dqc_handle* h, const int32_t* md_out, const int32_t* md_in, const int32_t* ub_out,
                          const int32_t* ub_in, int32_t len, int32_t min_size) {
    if (!h || len < 2 || min_size < 2 || min_size >= len) return DQC_E_ARG;
    hipSetDevice(h->device);
    drop_graphs(h);
    if (h->d_mdo) {
        for (void* p : {(void*)h->d_mdo, (void*)h->d_mdi, (void*)h->d_ubo, (void*)h->d_ubi}) {
            hipFree(p);
            h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), p), h->allocs.end());
        }
    }
    if (!dev_alloc(h, &h->d_mdo, len, "table alloc") || !dev_alloc(h, &h->d_mdi, len, "table alloc") ||
        !dev_alloc(h, &h->d_ubo, len, "table alloc") || !dev_alloc(h, &h->d_ubi, len, "table alloc"))
        return DQC_E_HIP;
    bool ok = hip_ok(h, hipMemcpy(h->d_mdo, md_out, sizeof(int32_t) * len, hipMemcpyHostToDevice), "tables") &&
              hip_ok(h, hipMemcpy(h->d_mdi, md_in, sizeof(int32_t) * len, hipMemcpyHostToDevice), "tables") &&
              hip_ok(h, hipMemcpy(h->d_ubo, ub_out, sizeof(int32_t) * len, hipMemcpyHostToDevice), "tables") &&
              hip_ok(h, hipMemcpy(h->d_ubi, ub_in, sizeof(int32_t) * len, hipMemcpyHostToDevice), "tables");
    if (!ok) return DQC_E_HIP;
    h->tlen = len;
    h->min_size = min_size;
    return DQC_OK;
}

static dqc_status push_root_impl(dqc_handle* h, int32_t gid, const uint32_t* records, int32_t nvert, int32_t ncov);

dqc_status dqc_push_root(dqc_handle* h, const uint32_t* records, int32_t nvert, int32_t ncov) {
    if (!h || h->ngraphs > 1) return DQC_E_ARG;
    return push_root_impl(h, 0, records, nvert, ncov);
}

dqc_status dqc_push_graph_root(dqc_handle* h, int32_t gid, const uint32_t* records, int32_t nvert, int32_t ncov) {
    if (!h || gid < 0 || gid >= h->ngraphs) return DQC_E_ARG;
    return push_root_impl(h, gid, records, nvert, ncov);
}

static dqc_status push_root_impl(dqc_handle* h, int32_t gid, const uint32_t* records, int32_t nvert, int32_t ncov) {
    if (nvert < 0 || ncov < 0 || ncov > nvert || (nvert > 0 && !records)) return DQC_E_ARG;
    if (nvert == 0) return DQC_OK;
    const int32_t b = h->gbase[gid];
    const int32_t nloc = (gid + 1 < h->ngraphs ? h->gbase[gid + 1] : h->n) - b;
    if (nvert > nloc) return DQC_E_ARG;
    std::vector<uint32_t> shifted;
    if (b != 0) {  // local ids to the shared id space
        shifted.assign(records, records + 6 * (size_t)nvert);
        for (int32_t i = 0; i < nvert; ++i) shifted[6 * (size_t)i] += (uint32_t)b;
        records = shifted.data();
    }
    for (int32_t i = 0; i < nvert; ++i)
        if (records[6 * (size_t)i] < (uint32_t)b || records[6 * (size_t)i] >= (uint32_t)(b + nloc)) return DQC_E_ARG;
    if ((unsigned long long)nvert > h->cfg.pool_vertices || h->st_t + h->def_t + 1 > h->cfg.spill_tasks ||
        h->st_v + h->def_v + nvert > h->cfg.spill_vertices) {
        h->err = "level-0 task exceeds the batch or stack capacity";
        return DQC_E_CAPACITY;
    }
    hipSetDevice(h->device);
    TaskDesc d = host_desc(records, nvert, 0, ncov, MODE_HOST_L0, h->st_v, 0, 0);
    d.gid = (uint16_t)gid;
    if (!hip_ok(h, hipMemcpy(h->d_svtx + h->st_v, records, sizeof(dqc::TaskVertex) * nvert, hipMemcpyHostToDevice), "push") ||
        !hip_ok(h, hipMemcpy(h->d_sdesc + h->st_t, &d, sizeof d, hipMemcpyHostToDevice), "push"))
        return DQC_E_HIP;
    h->st_t += 1;
    h->st_v += nvert;
    // one graph: its level-0 task runs branch-parallel; batched graphs run their level-0
    // tasks side by side, one workgroup each
    h->root_pending = h->ngraphs == 1 && std::getenv("DQC_SERIAL_ROOT") == nullptr;
    return DQC_OK;
}

dqc_status dqc_push_task(dqc_handle* h, const uint32_t* records, int32_t nvert, int32_t nm, int32_t ncov,
                         int32_t mode) {
    if (!h || !records || nvert < 1 || nm < 0 || ncov < 0 || nm + ncov > nvert || mode < 0 || mode > MODE_DEVICE)
        return DQC_E_ARG;
    if (nvert > h->n) return DQC_E_ARG;
    for (int32_t i = 0; i < nvert; ++i)
        if (records[6 * i] >= (uint32_t)h->n) return DQC_E_ARG;
    if ((unsigned long long)nvert > h->cfg.pool_vertices || h->st_t + h->def_t + 1 > h->cfg.spill_tasks ||
        h->st_v + h->def_v + nvert > h->cfg.spill_vertices) {
        h->err = "task exceeds the batch or stack capacity";
        return DQC_E_CAPACITY;
    }
    hipSetDevice(h->device);
    const TaskDesc d = host_desc(records, nvert, nm, ncov, mode, h->st_v, 0, 0);
    if (!hip_ok(h, hipMemcpy(h->d_svtx + h->st_v, records, sizeof(dqc::TaskVertex) * nvert, hipMemcpyHostToDevice), "push") ||
        !hip_ok(h, hipMemcpy(h->d_sdesc + h->st_t, &d, sizeof d, hipMemcpyHostToDevice), "push"))
        return DQC_E_HIP;
    h->st_t += 1;
    h->st_v += nvert;
    return DQC_OK;
}

dqc_status dqc_fetch_tasks(dqc_handle* h, int64_t* ntasks, int64_t* nrecords, int32_t* meta, uint32_t* records) {
    if (!h || !ntasks || !nrecords) return DQC_E_ARG;
    hipSetDevice(h->device);
    // the main stack, then the deferred region (records follow the main stack's)
    *ntasks = (int64_t)(h->st_t + h->def_t);
    *nrecords = (int64_t)(h->st_v + h->def_v);
    if (!meta || !records || *ntasks == 0) return DQC_OK;
    std::vector<TaskDesc> d((size_t)*ntasks);
    const unsigned long long dt0 = h->cfg.spill_tasks - h->def_t, dv0 = h->cfg.spill_vertices - h->def_v;
    if ((h->st_t && !hip_ok(h, hipMemcpy(d.data(), h->d_sdesc, sizeof(TaskDesc) * h->st_t, hipMemcpyDeviceToHost), "stack read")) ||
        (h->st_v && !hip_ok(h, hipMemcpy(records, h->d_svtx, sizeof(dqc::TaskVertex) * h->st_v, hipMemcpyDeviceToHost), "stack read")) ||
        (h->def_t && !hip_ok(h, hipMemcpy(d.data() + h->st_t, h->d_sdesc + dt0, sizeof(TaskDesc) * h->def_t, hipMemcpyDeviceToHost), "stack read")) ||
        (h->def_v && !hip_ok(h, hipMemcpy(records + 6 * h->st_v, h->d_svtx + dv0, sizeof(dqc::TaskVertex) * h->def_v, hipMemcpyDeviceToHost), "stack read")))
        return DQC_E_HIP;
    for (unsigned long long i = 0; i < h->st_t + h->def_t; ++i) {
        int32_t* m = meta + 9 * i;
        const unsigned long long off = i < h->st_t ? d[i].off : d[i].off - dv0 + h->st_v;
        m[0] = (int32_t)off; m[1] = d[i].tot; m[2] = d[i].nm; m[3] = d[i].mode;
        m[4] = (int32_t)d[i].uid; m[5] = (int32_t)d[i].parent; m[6] = d[i].depth; m[7] = d[i].flags;
        m[8] = d[i].ncov;
    }
    return DQC_OK;
}

dqc_status dqc_reset(dqc_handle* h) {
    if (!h) return DQC_E_ARG;
    hipSetDevice(h->device);
    h->st_t = 0;
    h->st_v = 0;
    h->def_t = 0;
    h->def_v = 0;
    h->peak_t = 0;
    h->peak_v = 0;
    h->root_pending = false;
    h->host_cliques.clear();
    GCtr g{};
    if (!hip_ok(h, hipMemset(h->d_g, 0, sizeof g), "reset") || !hip_ok(h, hipMemset(h->d_q, 0, sizeof(QCtr)), "reset"))
        return DQC_E_HIP;
    h->extra_tasks = 0;
    h->st.tasks_expanded = h->st.branches = h->st.raw_cliques = 0;
    // a failed launch can leave id-bitmap bits set in the HBM variant
    if (h->big && h->d_scratch &&
        !hip_ok(h, hipMemset(h->d_scratch, 0, sizeof(uint32_t) * h->stride * (size_t)(h->launch_wg + 1)), "reset"))
        return DQC_E_HIP;
    return DQC_OK;
}

static dqc_status batch_costs(dqc_handle* h, const TaskDesc* b, size_t n, unsigned long long first, int32_t sched,
                              double* cost);

// Pops the newest tasks off the stack into the batch buffer (as many as the batch
// holds); returns the number popped (0 with an error set when none fits).
//
// FIFO keeps stack order. SIZE and LEARNED order the batch by an estimate of each
// task's cost, largest first, so that the longest tasks start first, and mark the
// costliest tasks with at least split_branches branches for splitting, in that order
// while the running total of their piece records fits a budget of pool_vertices / 16
// per batch (half the free stack records when less is free; pieces copy the task's
// records and run all of its branches at once, so the budget bounds the stack growth).
// The first eligible task that does not fit ends the marking; the device-resident
// order kernel follows the same prefix rule with the same budget.
// Most pieces a split task is cut into: dqc_config.split_pieces, or one per workgroup.
static int max_pieces(const dqc_handle* h) { return h->cfg.split_pieces > 0 ? h->cfg.split_pieces : h->launch_wg; }

// Workgroups of a full launch for the popped batch's kernel.
static unsigned long long grid_cap(const dqc_handle* h) {
    return (unsigned long long)(h->batch_wave ? h->wave_wg : h->launch_wg);
}

static bool batch_has_long_task(const TaskDesc* b, unsigned long long n, int split_b) {
    for (unsigned long long i = 0; i < n; ++i)
        if (task_steps(b[i]) >= long_task_steps(split_b)) return true;
    return false;
}

static dqc_status pop_batch(dqc_handle* h, int32_t sched, unsigned long long* popped) {
    // newest tasks of the main stack, or of the deferred region when the main stack is empty
    const bool from_main = h->st_t > 0;
    const unsigned long long avail = from_main ? h->st_t : h->def_t;
    const unsigned long long m = std::min<unsigned long long>(avail, h->cfg.pool_tasks);
    const unsigned long long lo = from_main ? h->st_t - m : h->cfg.spill_tasks - h->def_t;  // lowest index read
    std::vector<TaskDesc> rd(m);
    if (!hip_ok(h, hipMemcpy(rd.data(), h->d_sdesc + lo, sizeof(TaskDesc) * m, hipMemcpyDeviceToHost), "stack read"))
        return DQC_E_HIP;
    // the j-th newest task, and the records from the newest task through it
    auto at = [&](unsigned long long j) -> const TaskDesc& { return from_main ? rd[m - 1 - j] : rd[j]; };
    auto span = [&](const TaskDesc& d) {
        return from_main ? h->st_v - d.off : d.off + (unsigned long long)d.tot - (h->cfg.spill_vertices - h->def_v);
    };
    if (span(at(0)) > h->cfg.pool_vertices) {
        h->err = "a task does not fit the batch buffer (raise pool_vertices)";
        return DQC_E_CAPACITY;
    }
    unsigned long long cnt = 1;  // keep the newest tasks whose records fit the batch buffer
    while (cnt < m && span(at(cnt)) <= h->cfg.pool_vertices) ++cnt;
    // (the stack cannot overflow from the batch's children: a workgroup that finds the
    // stack nearly full defers the rest of its task; see stack_low). The launch runs as
    // many workgroups as the stack has room for one child each; with no room for the
    // batch, the newest task alone is tried.
    auto room = [&](unsigned long long c) {
        unsigned long long mt = 0;
        for (unsigned long long j = 0; j < c; ++j) mt = std::max<unsigned long long>(mt, (unsigned long long)at(j).tot);
        h->batch_maxtot = std::max<unsigned long long>(mt, 1);
        h->batch_wave = wave_batch(h, h->batch_maxtot);
        const unsigned long long room_v = (h->cfg.spill_vertices - h->st_v - h->def_v) / h->batch_maxtot;
        const unsigned long long room_t = h->cfg.spill_tasks - h->st_t - h->def_t;
        h->batch_grid = std::min<unsigned long long>(grid_cap(h), std::min(room_v, room_t));
    };
    room(cnt);
    if (h->batch_grid == 0) {
        cnt = 1;
        room(cnt);
    }
    if (h->batch_grid == 0) {
        char st[256];
        std::snprintf(st, sizeof st, " [main %llu tasks / %llu records, deferred %llu / %llu, capacity %llu / %llu, "
                      "newest task %d records]", h->st_t, h->st_v, h->def_t, h->def_v,
                      (unsigned long long)h->cfg.spill_tasks, (unsigned long long)h->cfg.spill_vertices, at(0).tot);
        h->err = std::string("task stack full: no room left for the newest task's children (raise spill_tasks / "
                             "spill_vertices)") + st;
        return DQC_E_CAPACITY;
    }
    // under stack pressure (fewer workgroups than resident), two tasks per workgroup: a
    // full batch would mostly be deferred untouched
    if (h->batch_grid < grid_cap(h)) cnt = std::min(cnt, 2 * h->batch_grid);
    // the batch: descriptor indices [first, first + cnt), records [v0, v0 + recs)
    const unsigned long long first = from_main ? h->st_t - cnt : lo;
    const TaskDesc* b = from_main ? rd.data() + (m - cnt) : rd.data();
    const unsigned long long v0 = b[0].off, recs = b[cnt - 1].off + (unsigned long long)b[cnt - 1].tot - v0;
    h->batch_pend = (cnt << CNT_SHIFT) | recs;
    h->batch_nsplit = 0;
    h->batch_pend_rest = h->batch_pend;
    if (sched != DQC_SCHED_FIFO && !batch_has_long_task(b, cnt, h->cfg.split_branches))
        sched = DQC_SCHED_FIFO;  // nothing to split or to balance: stack order keeps locality
    if (sched == DQC_SCHED_FIFO) {
        refill_kernel<<<1024, 256, 0, h->stream>>>(h->d_sdesc, h->d_svtx, first, cnt, h->d_bdesc, h->d_bvtx, 0, 0);
    } else {
        h->hcost.resize(cnt);
        dqc_status rc = batch_costs(h, b, cnt, first, sched, h->hcost.data());
        if (rc != DQC_OK) return rc;
        std::vector<uint32_t> order(cnt);
        for (uint32_t i = 0; i < cnt; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return h->hcost[x] > h->hcost[y]; });
        h->hbatch.resize(cnt);
        const unsigned long long used = h->st_v + h->def_v;
        unsigned long long budget = h->cfg.spill_vertices > used ? (h->cfg.spill_vertices - used) / 4 : 0;
        double total = 0.0;
        for (unsigned long long i = 0; i < cnt; ++i) total += h->hcost[i];
        for (unsigned long long i = 0; i < cnt; ++i) {
            TaskDesc d = b[order[i]];
            d.off -= v0;
            const int steps = d.tot - d.nm - (d.mode == MODE_DEVICE ? 0 : d.ncov);
            if (split_eligible(d, h->cfg.split_branches)) {
                const int pieces = pieces_for(h->hcost[order[i]], total, h->launch_wg, max_pieces(h));
                const int chunk = split_chunk_for(steps, h->cfg.split_chunk, pieces);
                const unsigned long long need = (unsigned long long)((steps + chunk - 1) / chunk) * (unsigned long long)d.tot;
                if (need <= budget) {
                    budget -= need;
                    d.flags |= TF_SPLIT;
                    d.iters = chunk;  // piece length
                } else {
                    budget = 0;  // prefix rule: no later task is split in this batch
                }
            }
            h->hbatch[i] = d;
        }
        // the tasks to split run first, in a launch of their own on the split-pass kernel
        std::stable_partition(h->hbatch.begin(), h->hbatch.end(), [](const TaskDesc& d) { return (d.flags & TF_SPLIT) != 0; });
        unsigned long long rest_v = 0;
        for (unsigned long long i = 0; i < cnt; ++i) {
            if (h->hbatch[i].flags & TF_SPLIT) ++h->batch_nsplit;
            else rest_v += (unsigned long long)h->hbatch[i].tot;
        }
        h->batch_pend_rest = ((cnt - h->batch_nsplit) << CNT_SHIFT) | rest_v;
        if (!hip_ok(h, hipMemcpyAsync(h->d_bdesc, h->hbatch.data(), sizeof(TaskDesc) * cnt, hipMemcpyHostToDevice, h->stream),
                    "batch upload"))
            return DQC_E_HIP;
        refill_kernel<<<1024, 256, 0, h->stream>>>(h->d_sdesc, h->d_svtx, first, cnt, nullptr, h->d_bvtx, 0, 0);
    }
    if (!hip_ok(h, hipGetLastError(), "pop launch")) return DQC_E_HIP;
    if (from_main) {
        h->st_t = first;
        h->st_v = v0;
    } else {
        h->def_t -= cnt;
        h->def_v -= recs;
    }
    *popped = cnt;
    return DQC_OK;
}

// Cost estimates of a batch (b = host copy of the stack descriptors first .. first+n):
// SIZE uses branches x task size (every branch scans lists of vertices of the task);
// LEARNED evaluates the cost model on the device over the same descriptors. A piece
// gets the share of its task's steps it runs.
static dqc_status batch_costs(dqc_handle* h, const TaskDesc* b, size_t n, unsigned long long first, int32_t sched,
                              double* cost) {
    if (sched == DQC_SCHED_LEARNED) {
        if (!h->d_model) {
            h->err = "no cost model set (dqc_set_cost_model)";
            return DQC_E_ARG;
        }
        const int blocks = (int)std::min<size_t>(1024, (n + 255) / 256);
        desc_cost_kernel<<<blocks, 256, 0, h->stream>>>(h->d_model, h->d_sdesc + first, n, h->d_cost, nullptr);
        h->hpred.resize(n);
        if (!hip_ok(h, hipGetLastError(), "cost launch") ||
            !hip_ok(h, hipMemcpyAsync(h->hpred.data(), h->d_cost, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream), "cost read") ||
            !hip_ok(h, hipStreamSynchronize(h->stream), "cost kernel"))
            return DQC_E_HIP;
        for (size_t i = 0; i < n; ++i) {  // predicted branches x task size
            const double steps = (double)std::max(1, b[i].tot - b[i].nm);
            const double share = (b[i].flags & TF_PIECE) ? std::min(1.0, (double)b[i].iters / steps) : 1.0;
            cost[i] = std::max(0.0, std::expm1((double)std::min(h->hpred[i], 60.f)) * share * (double)b[i].tot);
        }
        return DQC_OK;
    }
    for (size_t i = 0; i < n; ++i) {
        const double steps = (double)(b[i].tot - b[i].nm);
        cost[i] = ((b[i].flags & TF_PIECE) ? (double)b[i].iters : steps) * (double)b[i].tot;
    }
    return DQC_OK;
}

static dqc_status run_launches(dqc_handle* h, int32_t sched, int64_t max_launches, int32_t* finished);

dqc_status dqc_run(dqc_handle* h, int32_t sched) {
    int32_t fin = 0;
    return run_launches(h, sched, -1, &fin);
}

dqc_status dqc_step(dqc_handle* h, int32_t sched, int32_t launches, int32_t* finished) {
    if (!finished || launches < 1) return DQC_E_ARG;
    return run_launches(h, sched, launches, finished);
}

static KParams make_params(dqc_handle* h, unsigned long long ntask) {
    KParams P;
    P.n = h->n;
    P.oo = h->d_oo; P.io = h->d_io; P.to = h->d_to;
    P.on = h->d_on; P.in = h->d_in; P.tn = h->d_tn;
    P.mdo = h->d_mdo; P.mdi = h->d_mdi; P.ubo = h->d_ubo; P.ubi = h->d_ubi;
    P.min_size = h->min_size;
    P.host_addfail = h->cfg.host_addfail_check;
    P.profile = h->profile;
    P.sum_shift = h->sum_shift;
    P.in_desc = h->d_bdesc; P.in_vtx = h->d_bvtx; P.in_ntask = ntask;
    P.sp_desc = h->d_sdesc; P.sp_vtx = h->d_svtx;
    P.sp_cap_t = h->cfg.spill_tasks; P.sp_cap_v = h->cfg.spill_vertices;
    P.clq_ids = h->d_clq_ids; P.clq_meta = h->d_clq_meta;
    P.clq_cap_s = h->cfg.clique_sets; P.clq_cap_i = h->cfg.clique_ids;
    P.q = h->d_q;
    P.g = h->d_g;
    P.split_s = nullptr;
    P.lto = h->d_lto; P.ltn = h->d_ltn;
    P.scratch = h->d_scratch;
    P.stride = h->stride;
    P.log = h->d_log;
    P.log_cap = h->log_cap;
    P.kloc = h->kloc;
    P.gbase = h->d_gbase;
    P.max_depth = (uint32_t)std::max(0, h->cfg.max_depth);
    P.root_steps = std::max(0, h->cfg.root_branches);
    P.root_first = std::max(0, h->cfg.root_first);
    P.slice_ticks = h->cfg.launch_ms > 0 ? (unsigned long long)h->cfg.launch_ms * 100000ull : 0ull;  // 100 MHz
    P.maxtot = std::max<unsigned long long>(h->batch_maxtot, 1);
    return P;
}

static void fold_stats(dqc_handle* h, const GCtr& g) {
    h->st.launches = h->host_launches + (int64_t)g.steps;
    h->st.tasks_expanded = (int64_t)g.tasks + h->extra_tasks;
    h->st.branches = (int64_t)g.branches;
    h->st.raw_cliques = (int64_t)g.cliques + (int64_t)h->host_cliques.size();
    h->st.bytes_scanned = (int64_t)g.bytes;
    h->st.slow_sorts = (int64_t)g.slow_sorts;
    h->st.truncated_children = (int64_t)g.truncated;
    h->st.stack_pushbacks = (int64_t)g.pushbacks;
    h->st.time_slices = (int64_t)g.slices;
    h->st.wave_launches = h->wave_launches + (wave_resident(h) ? (int64_t)g.steps : 0);
    h->st.peak_stack_records = (int64_t)h->peak_v;
    h->st.peak_stack_tasks = (int64_t)h->peak_t;
    for (int i = 0; i < 8; ++i) h->phase[i] = g.phase[i];
}

// Launches one expansion over `ntask` work items and folds its counters into the stats.
// split_pass: the work items are tasks to split (the split-pass kernel runs them), and
// pend = the batch's tasks not yet finished.
static dqc_status launch(dqc_handle* h, const KParams& P, unsigned long long ntask, bool split_pass = false,
                         unsigned long long pend = ~0ull) {
    QCtr q{};
    q.next_task = 0;
    q.sp_alloc = (h->st_t << CNT_SHIFT) | h->st_v;
    q.root_first = ~0ull;
    q.pend = P.split_s ? 0 : (pend == ~0ull ? h->batch_pend : pend);
    q.def_alloc = (h->def_t << CNT_SHIFT) | h->def_v;
    q.used = ((h->st_t + h->def_t) << CNT_SHIFT) | (h->st_v + h->def_v);
    if (!hip_ok(h, hipMemcpyAsync(h->d_q, &q, sizeof q, hipMemcpyHostToDevice, h->stream), "ctr write")) return DQC_E_HIP;
    const bool wave = !P.split_s && !split_pass && h->batch_wave;
    unsigned long long gmax = (unsigned long long)(wave ? h->wave_wg : split_pass ? std::min(h->split_wg, h->launch_wg)
                                                                                  : h->launch_wg);
    if (!P.split_s) gmax = std::min(gmax, h->batch_grid);
    const int grid = (int)std::min<unsigned long long>(gmax, ntask);
    hipEventRecord(h->ev0, h->stream);
    if (split_pass && h->big) expand_kernel<true, SPLIT_T><<<grid, SPLIT_T, 0, h->stream>>>(P);
    else if (split_pass) expand_kernel<false, SPLIT_T><<<grid, SPLIT_T, 0, h->stream>>>(P);
    else if (h->big) expand_kernel<true><<<grid, NT, 0, h->stream>>>(P);
    else if (wave) expand_kernel<false, WAVE_T, WAVE_KW><<<grid, WAVE_T, 0, h->stream>>>(P);
    else expand_kernel<false><<<grid, NT, 0, h->stream>>>(P);
    h->wave_launches += wave ? 1 : 0;
    hipEventRecord(h->ev1, h->stream);
    if (!hip_ok(h, hipGetLastError(), "expand launch")) return DQC_E_HIP;
    GCtr g{};
    if (!hip_ok(h, hipMemcpyAsync(&q, h->d_q, sizeof q, hipMemcpyDeviceToHost, h->stream), "ctr read") ||
        !hip_ok(h, hipMemcpyAsync(&g, h->d_g, sizeof g, hipMemcpyDeviceToHost, h->stream), "ctr read") ||
        !hip_ok(h, hipStreamSynchronize(h->stream), "expand kernel"))
        return DQC_E_HIP;
    float ms = 0.f;
    hipEventElapsedTime(&ms, h->ev0, h->ev1);
    h->st.expand_kernel_ms += ms;
    h->host_launches++;
    fold_stats(h, g);
    if (g.error) {
        char st[200];
        std::snprintf(st, sizeof st, " [used %llu/%llu, request %llu, state %llx, pending %llu/%llu]",
                      g.spill_state[0] >> CNT_SHIFT, g.spill_state[0] & SIZE_MASK, g.spill_state[1], g.spill_state[2],
                      g.spill_state[3] >> CNT_SHIFT, g.spill_state[3] & SIZE_MASK);
        h->err = (g.error & ERR_SPILL) ? std::string("task stack full (raise spill_tasks / spill_vertices)") + st
               : std::string((g.error & ERR_CLIQUES) ? "clique store full (raise clique_sets / clique_ids)"
                                                     : "a task has more members than the member histograms hold");
        return (g.error & ERR_MEMBERS) ? DQC_E_UNSUPPORTED : DQC_E_CAPACITY;
    }
    h->st_t = q.sp_alloc >> CNT_SHIFT;
    h->st_v = q.sp_alloc & SIZE_MASK;
    h->def_t = q.def_alloc >> CNT_SHIFT;
    h->def_v = q.def_alloc & SIZE_MASK;
    h->peak_t = std::max(h->peak_t, h->st_t + h->def_t);
    h->peak_v = std::max(h->peak_v, h->st_v + h->def_v);
    h->last_root_first = q.root_first;
    h->max_depth = std::max(h->max_depth, h->st_t);
    return DQC_OK;
}

// The level-0 task expanded branch-parallel across all workgroups (see branch_root):
// a flags pass finds where the sequential branch loop would stop and which steps skip
// their branch, then every remaining branch runs as its own work item. The child of the
// smallest branch index keeps the host level-1 semantics (reference
// src/host_functions.cu:1110-1133: the first child goes to the CPU level-1 pass).
static dqc_status split_root(dqc_handle* h) {
    unsigned long long m = 0;
    dqc_status rc = pop_batch(h, DQC_SCHED_FIFO, &m);
    if (rc != DQC_OK) return rc;
    // the pop runs on the main stream, which does not order against null-stream copies
    if (!hip_ok(h, hipStreamSynchronize(h->stream), "pop")) return DQC_E_HIP;
    TaskDesc d;
    if (!hip_ok(h, hipMemcpy(&d, h->d_bdesc, sizeof d, hipMemcpyDeviceToHost), "root read")) return DQC_E_HIP;
    const int k = d.tot, ncov = d.ncov;
    // a bounded unit of work: only the first root_branches loop steps (dqc_config)
    const int S = h->cfg.root_branches > 0 ? std::min(k - ncov, std::max(0, h->cfg.root_first) + h->cfg.root_branches)
                                           : k - ncov;
    KParams P = make_params(h, 0);
    if (S > 0) {
        root_flags_kernel<<<(S + 255) / 256, 256, 0, h->stream>>>(P, h->d_bvtx + d.off, k, S, h->d_flags);
        if (!hip_ok(h, hipGetLastError(), "flags launch") || !hip_ok(h, hipStreamSynchronize(h->stream), "flags kernel"))
            return DQC_E_HIP;
    }
    std::vector<uint8_t> flags((size_t)std::max(S, 1));
    if (S > 0 && !hip_ok(h, hipMemcpy(flags.data(), h->d_flags, (size_t)S, hipMemcpyDeviceToHost), "flags read"))
        return DQC_E_HIP;
    // branches before the window (root_first) and in it
    std::vector<int32_t> head, work;
    const int R0 = std::max(0, h->cfg.root_first);
    for (int s = 0; s < S; ++s) {
        const uint8_t f = flags[s];
        if (f & 4) break;          // fewer than min_size vertices remain
        if (f & 1) continue;       // this step's removal skips the branch
        if (f & 2) {               // lookahead: the whole remaining set is emitted
            std::vector<int32_t> c((size_t)(k - s));
            for (int i = 0; i < k - s; ++i) c[i] = i;
            h->host_cliques.push_back(std::move(c));
            break;
        }
        (s < R0 ? head : work).push_back(s);
    }
    h->extra_tasks += 1;  // the level-0 task itself
    // Before the window only the first child matters: if one exists there, it takes the
    // host level-1 role and no child of the window does. Those branches run in chunks of
    // one per workgroup, from the first, until one writes a child (discarded).
    bool first_before = false;
    for (size_t c0 = 0; c0 < head.size() && !first_before; c0 += (size_t)h->launch_wg) {
        const size_t cn = std::min(head.size() - c0, (size_t)h->launch_wg);
        if (!hip_ok(h, hipMemcpy(h->d_split, head.data() + c0, sizeof(int32_t) * cn, hipMemcpyHostToDevice), "split upload"))
            return DQC_E_HIP;
        P.in_ntask = cn;
        P.split_s = h->d_split;
        rc = launch(h, P, cn);
        if (rc != DQC_OK) return rc;
        first_before = h->last_root_first != ~0ull;
    }
    if (work.empty()) {
        if (head.empty()) {  // no launch ran: the counters are the host's
            h->st.tasks_expanded += 1;
            h->st.raw_cliques = (int64_t)h->host_cliques.size();
        }
        return DQC_OK;
    }
    if (!hip_ok(h, hipMemcpy(h->d_split, work.data(), sizeof(int32_t) * work.size(), hipMemcpyHostToDevice), "split upload"))
        return DQC_E_HIP;
    P.in_ntask = work.size();
    P.split_s = h->d_split;
    rc = launch(h, P, work.size());
    if (rc != DQC_OK) return rc;
    // the first child keeps the host level-1 semantics (unless it lay before the window)
    if (!first_before && h->last_root_first != ~0ull && (h->last_root_first & 0xFFFFFFFFull) != 0xFFFFFFFFull) {
        const unsigned long long t = h->last_root_first & 0xFFFFFFFFull;
        const int32_t mode = MODE_HOST_L1;
        if (!hip_ok(h, hipMemcpy((char*)(h->d_sdesc + t) + offsetof(TaskDesc, mode), &mode, sizeof mode,
                                 hipMemcpyHostToDevice), "mode patch"))
            return DQC_E_HIP;
    }
    return DQC_OK;
}

// Enqueues one device-resident step on the handle's stream: pop, (costs, order), gather,
// expand. Nothing in it waits for the host.
static void enqueue_step(dqc_handle* h, int32_t sched) {
    const bool wave = wave_resident(h);
    const int wg = wave ? h->wave_wg : h->launch_wg;
    PopParams pp{h->d_sdesc, h->d_q, h->d_g, h->cfg.pool_tasks, h->cfg.pool_vertices, h->cfg.spill_tasks,
                 h->cfg.spill_vertices, (unsigned long long)wg};
    pop_kernel<<<1, POP_T, 0, h->stream>>>(pp);
    const uint32_t* perm = nullptr;
    const uint32_t* mark = nullptr;
    if (sched != DQC_SCHED_FIFO) {
        if (sched == DQC_SCHED_LEARNED) {
            const int blocks = (int)std::min<unsigned long long>(1024, (h->cfg.pool_tasks + 255) / 256);
            desc_cost_kernel<<<blocks, 256, 0, h->stream>>>(h->d_model, h->d_sdesc, 0, h->d_cost, h->d_q);
        }
        order_kernel<<<1, POP_T, 0, h->stream>>>(h->d_sdesc, h->d_q, h->d_cost, sched == DQC_SCHED_LEARNED,
                                                 h->cfg.split_branches, h->cfg.split_chunk, max_pieces(h), wg,
                                                 h->d_perm, h->d_mark);
        perm = h->d_perm;
        mark = h->d_mark;
    }
    gather_kernel<<<1024, 256, 0, h->stream>>>(h->d_sdesc, h->d_svtx, h->d_q, perm, mark, h->d_bdesc, h->d_bvtx);
    KParams P = make_params(h, ~0ull);
    P.maxtot = ~0ull;  // read from the device-resident pop
    if (h->big) expand_kernel<true><<<wg, NT, 0, h->stream>>>(P);
    else if (wave) expand_kernel<false, WAVE_T, WAVE_KW><<<wg, WAVE_T, 0, h->stream>>>(P);
    else expand_kernel<false><<<wg, NT, 0, h->stream>>>(P);
}

// Device-resident expansion: replays a captured HIP graph of resident_steps steps until
// the stack is empty (or max_steps steps ran); the host only reads the counters between
// replays.
static dqc_status resident_run(dqc_handle* h, int32_t sched, int64_t max_steps, int64_t* done) {
    if (sched == DQC_SCHED_LEARNED && !h->d_model) { h->err = "no cost model set (dqc_set_cost_model)"; return DQC_E_ARG; }
    if (!h->d_perm && (!dev_alloc(h, &h->d_perm, h->cfg.pool_tasks, "order alloc") ||
                       !dev_alloc(h, &h->d_mark, h->cfg.pool_tasks, "order alloc")))
        return DQC_E_HIP;
    QCtr q{};
    q.sp_alloc = (h->st_t << CNT_SHIFT) | h->st_v;
    q.def_alloc = (h->def_t << CNT_SHIFT) | h->def_v;
    q.used = ((h->st_t + h->def_t) << CNT_SHIFT) | (h->st_v + h->def_v);
    q.root_first = ~0ull;
    if (!hip_ok(h, hipMemcpyAsync(h->d_q, &q, sizeof q, hipMemcpyHostToDevice, h->stream), "ctr write")) return DQC_E_HIP;
    hipGraphExec_t& ex = h->gexec[sched];
    const int S = h->cfg.resident_steps;
    if (!ex) {
        hipGraph_t g = nullptr;
        if (!hip_ok(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed), "capture")) return DQC_E_HIP;
        for (int i = 0; i < S; ++i) enqueue_step(h, sched);
        if (!hip_ok(h, hipStreamEndCapture(h->stream, &g), "capture") ||
            !hip_ok(h, hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "graph instantiate")) {
            if (g) hipGraphDestroy(g);
            ex = nullptr;
            return DQC_E_HIP;
        }
        hipGraphDestroy(g);
    }
    while (max_steps < 0 || *done < max_steps) {
        hipEventRecord(h->ev0, h->stream);
        if (!hip_ok(h, hipGraphLaunch(ex, h->stream), "graph launch")) return DQC_E_HIP;
        hipEventRecord(h->ev1, h->stream);
        GCtr g{};
        if (!hip_ok(h, hipMemcpyAsync(&q, h->d_q, sizeof q, hipMemcpyDeviceToHost, h->stream), "ctr read") ||
            !hip_ok(h, hipMemcpyAsync(&g, h->d_g, sizeof g, hipMemcpyDeviceToHost, h->stream), "ctr read") ||
            !hip_ok(h, hipStreamSynchronize(h->stream), "expand graph"))
            return DQC_E_HIP;
        float ms = 0.f;
        hipEventElapsedTime(&ms, h->ev0, h->ev1);
        h->st.expand_kernel_ms += ms;
        *done += S;
        fold_stats(h, g);
        h->st_t = q.sp_alloc >> CNT_SHIFT;
        h->st_v = q.sp_alloc & SIZE_MASK;
        h->def_t = q.def_alloc >> CNT_SHIFT;
        h->def_v = q.def_alloc & SIZE_MASK;
        h->peak_t = std::max(h->peak_t, h->st_t + h->def_t);
        h->peak_v = std::max(h->peak_v, h->st_v + h->def_v);
        if (g.error) {
            h->err = (g.error & ERR_SPILL) ? "task stack full (raise spill_tasks / spill_vertices)"
                   : (g.error & ERR_CLIQUES) ? "clique store full (raise clique_sets / clique_ids)"
                   : (g.error & ERR_BATCH) ? "a task does not fit the batch buffer (raise pool_vertices)"
                   : "a task has more members than the member histograms hold";
            return (g.error & ERR_MEMBERS) ? DQC_E_UNSUPPORTED : DQC_E_CAPACITY;
        }
        if (h->st_t == 0 && h->def_t == 0) break;
    }
    return DQC_OK;
}

static dqc_status run_launches(dqc_handle* h, int32_t sched, int64_t max_launches, int32_t* finished) {
    if (!h || sched < DQC_SCHED_DEFAULT || sched > DQC_SCHED_LEARNED) return DQC_E_ARG;
    if (sched == DQC_SCHED_DEFAULT) sched = h->cfg.schedule;
    if (!h->d_mdo || !h->d_oo) { h->err = "graph and tables must be set first"; return DQC_E_ARG; }
    hipSetDevice(h->device);
    auto t0 = std::chrono::steady_clock::now();
    const bool progress = std::getenv("DQC_PROGRESS") != nullptr;
    auto last_report = t0;
    int64_t done = 0;
    while ((h->st_t > 0 || h->def_t > 0) && (max_launches < 0 || done < max_launches)) {
        ++done;
        if (h->st_t == 1 && h->root_pending) {
            h->root_pending = false;
            dqc_status rc = split_root(h);
            if (rc != DQC_OK) return rc;
        } else if (h->cfg.resident_steps > 0) {
            --done;
            dqc_status rc = resident_run(h, sched, max_launches, &done);
            if (rc != DQC_OK) return rc;
        } else {
            unsigned long long m = 0;
            dqc_status rc = pop_batch(h, sched, &m);
            if (rc != DQC_OK) return rc;
            const unsigned long long ns = h->batch_nsplit;
            if (ns > 0) {  // the split passes first, then the rest of the batch
                rc = launch(h, make_params(h, ns), ns, true);
                if (rc != DQC_OK) return rc;
            }
            if (m > ns) {
                KParams P = make_params(h, m - ns);
                P.in_desc = h->d_bdesc + ns;
                rc = launch(h, P, m - ns, false, h->batch_pend_rest);
                if (rc != DQC_OK) return rc;
            }
        }
        if (progress) {
            auto now = std::chrono::steady_clock::now();
            if (std::chrono::duration<double>(now - last_report).count() > 5.0) {
                last_report = now;
                std::fprintf(stderr, "dqc: %.1fs launches=%lld tasks=%lld stack=%llu cliques=%lld\n",
                             std::chrono::duration<double>(now - t0).count(), (long long)h->st.launches,
                             (long long)h->st.tasks_expanded, h->st_t, (long long)h->st.raw_cliques);
            }
        }
    }
    if (h->st_t == 0 && h->def_t == 0) {  // everything done: final counters
        GCtr g{};
        if (!hip_ok(h, hipMemcpy(&g, h->d_g, sizeof g, hipMemcpyDeviceToHost), "ctr read")) return DQC_E_HIP;
        fold_stats(h, g);
    }
    auto t1 = std::chrono::steady_clock::now();
    h->st.expand_wall_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    *finished = h->st_t == 0 && h->def_t == 0;
    return DQC_OK;
}

dqc_status dqc_set_cost_model(dqc_handle* h, const float* params, int64_t n_params) {
    if (!h || !params || n_params != dqc_cost::NPARAM) return DQC_E_ARG;
    hipSetDevice(h->device);
    if (!h->d_model) {
        if (!dev_alloc(h, &h->d_model, dqc_cost::NPARAM, "model alloc") ||
            !dev_alloc(h, &h->d_cost, h->cfg.pool_tasks, "model alloc"))
            return DQC_E_HIP;
    }
    if (!hip_ok(h, hipMemcpy(h->d_model, params, sizeof(float) * dqc_cost::NPARAM, hipMemcpyHostToDevice), "model upload"))
        return DQC_E_HIP;
    return DQC_OK;
}

dqc_status dqc_set_task_log(dqc_handle* h, int64_t capacity) {
    if (!h || capacity < 0) return DQC_E_ARG;
    hipSetDevice(h->device);
    drop_graphs(h);
    if (h->d_log) {
        hipFree(h->d_log);
        h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), (void*)h->d_log), h->allocs.end());
        h->d_log = nullptr;
    }
    h->log_cap = (unsigned long long)capacity;
    if (capacity > 0 && !dev_alloc(h, &h->d_log, (size_t)capacity, "log alloc")) return DQC_E_HIP;
    const unsigned long long zero = 0;
    if (!hip_ok(h, hipMemcpy((char*)h->d_g + offsetof(GCtr, log_n), &zero, sizeof zero, hipMemcpyHostToDevice), "log reset"))
        return DQC_E_HIP;
    return DQC_OK;
}

dqc_status dqc_fetch_task_log(dqc_handle* h, int64_t* count, int64_t* dropped, dqc_task_log* records) {
    if (!h || !count) return DQC_E_ARG;
    hipSetDevice(h->device);
    GCtr g{};
    if (!hip_ok(h, hipMemcpy(&g, h->d_g, sizeof g, hipMemcpyDeviceToHost), "ctr read")) return DQC_E_HIP;
    const unsigned long long n = std::min(g.log_n, h->log_cap);
    *count = (int64_t)n;
    if (dropped) *dropped = (int64_t)(g.log_n - n);
    if (records && n &&
        !hip_ok(h, hipMemcpy(records, h->d_log, sizeof(TaskLog) * n, hipMemcpyDeviceToHost), "log read"))
        return DQC_E_HIP;
    return DQC_OK;
}

dqc_status dqc_fetch_batch_cliques(dqc_handle* h, int64_t* count, int64_t* total_ids, int32_t* graph, int32_t* sizes,
                                   int32_t* ids) {
    if (!h || !count || !total_ids) return DQC_E_ARG;
    dqc_status rc = dqc_fetch_cliques(h, count, total_ids, sizes, ids);
    if (rc != DQC_OK || !graph || !sizes || !ids) return rc;
    size_t p = 0;
    for (int64_t c = 0; c < *count; ++c) {
        const int32_t len = sizes[c];
        const int32_t first = len > 0 ? ids[p] : 0;
        const int32_t g = (int32_t)(std::upper_bound(h->gbase.begin(), h->gbase.end(), first) - h->gbase.begin()) - 1;
        graph[c] = g;
        for (int32_t i = 0; i < len; ++i) ids[p + i] -= h->gbase[g];
        p += (size_t)len;
    }
    return DQC_OK;
}

dqc_status dqc_set_unit(dqc_handle* h, int32_t root_first, int32_t root_branches, int32_t max_depth) {
    if (!h || root_first < 0 || root_branches < 0 || max_depth < 0) return DQC_E_ARG;
    drop_graphs(h);  // captured steps hold the old kernel parameters
    h->cfg.root_first = root_first;
    h->cfg.root_branches = root_branches;
    h->cfg.max_depth = max_depth;
    return DQC_OK;
}

dqc_status dqc_batch_size(const dqc_handle* h, int32_t* ngraphs) {
    if (!h || !ngraphs) return DQC_E_ARG;
    *ngraphs = h->ngraphs;
    return DQC_OK;
}

dqc_status dqc_get_phases(const dqc_handle* h, uint64_t* ticks8) {
    if (!h || !ticks8) return DQC_E_ARG;
    for (int i = 0; i < 8; ++i) ticks8[i] = h->phase[i];
    return DQC_OK;
}

dqc_status dqc_fetch_cliques(dqc_handle* h, int64_t* count, int64_t* total_ids, int32_t* sizes, int32_t* ids) {
    if (!h || !count || !total_ids) return DQC_E_ARG;
    hipSetDevice(h->device);
    GCtr g{};
    if (!hip_ok(h, hipMemcpy(&g, h->d_g, sizeof g, hipMemcpyDeviceToHost), "ctr read")) return DQC_E_HIP;
    const unsigned long long ns = g.clq_alloc >> CNT_SHIFT, ni = g.clq_alloc & SIZE_MASK;
    unsigned long long hs = h->host_cliques.size(), hi = 0;
    for (auto& c : h->host_cliques) hi += c.size();
    *count = (int64_t)(ns + hs);
    *total_ids = (int64_t)(ni + hi);
    if (!sizes || !ids) return DQC_OK;
    size_t p = 0, k = 0;
    for (auto& c : h->host_cliques) {
        sizes[k++] = (int32_t)c.size();
        std::memcpy(ids + p, c.data(), sizeof(int32_t) * c.size());
        p += c.size();
    }
    if (ns == 0) return DQC_OK;
    std::vector<unsigned long long> meta(ns);
    std::vector<int32_t> raw(ni);
    if (!hip_ok(h, hipMemcpy(meta.data(), h->d_clq_meta, sizeof(unsigned long long) * ns, hipMemcpyDeviceToHost), "clique read") ||
        !hip_ok(h, hipMemcpy(raw.data(), h->d_clq_ids, sizeof(int32_t) * ni, hipMemcpyDeviceToHost), "clique read"))
        return DQC_E_HIP;
    for (unsigned long long s = 0; s < ns; ++s) {
        const unsigned long long off = meta[s] >> 24, len = meta[s] & 0xFFFFFFull;
        sizes[k++] = (int32_t)len;
        std::memcpy(ids + p, raw.data() + off, sizeof(int32_t) * len);
        p += len;
    }
    return DQC_OK;
}

dqc_status dqc_get_stats(const dqc_handle* h, dqc_stats* st) {
    if (!h || !st) return DQC_E_ARG;
    *st = h->st;
    return DQC_OK;
}

}  // extern "C"

// the allocation-free entry point launches the kernels above
#include "noalloc.hip"
