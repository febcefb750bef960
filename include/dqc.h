/* C-ABI of the MI355X maximal directed quasi-clique enumerator (libcudqc_amd.so).
 *
 * The reference (Mike12041204/cuDQC) has no plugin or FFI layer: its enumeration path
 * sits behind the CLI in program.cu and an in-process C++ host API. These entry points
 * are what that host API would bind to; each cites the reference interface it replaces.
 * Plain pointers and sizes only; every function returns a dqc_status (0 = DQC_OK) and
 * never calls exit(). A handle is single-threaded and owns all device memory it uses;
 * no caller pointer is retained after a call returns.
 */
#ifndef CUDQC_AMD_DQC_H
#define CUDQC_AMD_DQC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t dqc_status;
#define DQC_OK 0
#define DQC_E_HIP 1      /* HIP runtime error (message via dqc_last_error)        */
#define DQC_E_CAPACITY 2 /* a device pool or clique buffer is too small           */
#define DQC_E_ARG 3      /* invalid argument (reference program.cu:27-61 checks)   */
#define DQC_E_IO 4       /* file could not be read or written                      */
#define DQC_E_UNSUPPORTED 5 /* input outside what this build handles (e.g. size)  */

typedef struct dqc_handle dqc_handle;

/* Capacities are advisory in the way DS_Sizes.csv is for the reference
 * (reference src/common.cpp:521-557); 0 selects a default. The batch buffer and task
 * stack defaults follow the largest graph uploaded (k vertices): 1024 k batch records
 * (2^20 .. 2^27), 8192 k stack records (2^22 .. 2^30), stack tasks 1/16 of that. */
typedef struct {
    int32_t device;            /* HIP device ordinal                                  */
    int32_t workgroups;        /* persistent workgroups; 0 = as many as are resident  */
    uint64_t pool_tasks;       /* tasks expanded per launch (batch); 0 = 32 per WG    */
    uint64_t pool_vertices;    /* 24-byte vertex records the batch buffer holds       */
    uint64_t spill_tasks;      /* task descriptors the LIFO task stack holds          */
    uint64_t spill_vertices;   /* vertex records the LIFO task stack holds            */
    uint64_t clique_sets;      /* raw cliques held on the device                      */
    uint64_t clique_ids;       /* ids of raw cliques held on the device               */
    int32_t host_addfail_check;/* 1 (default 4-rank union) / 0 (single rank, see
                                  reference src/host_functions.cu:1081)                */
    int32_t split_branches;    /* SIZE / LEARNED schedules: a task with at least this
                                  many branches may be cut into pieces (of at least
                                  split_chunk branches; see split_pieces) that run in
                                  parallel; 0 = 1024, negative = never (output is
                                  identical)                                          */
    int32_t split_chunk;       /* least branches per piece; 0 = 32                    */
    int32_t schedule;          /* DQC_SCHED_* used by dqc_search_file and by
                                  DQC_SCHED_DEFAULT; 0 = DQC_SCHED_SIZE               */
    int32_t max_depth;         /* > 0: depth-bounded enumeration: children deeper than
                                  this (the level-0 task has depth 0) are counted in
                                  truncated_children and not expanded; 0 = no bound   */
    int32_t resident_steps;    /* > 0: batches are popped, ordered and expanded on the
                                  device, resident_steps per replay of one captured HIP
                                  graph, with no host round trip between them;
                                  0 = host-driven steps                               */
    int32_t root_branches;     /* > 0: only the level-0 loop steps s < root_first +
                                  root_branches of a single graph run (with max_depth, a
                                  fixed bounded unit of work: a window of level-0
                                  branches and their subtrees down to that depth);
                                  0 = all                                             */
    int32_t split_pieces;      /* most pieces one split task is cut into; 0 = as many
                                  as a launch has workgroups. A batch with several
                                  tasks to split gives each about 4 x workgroups /
                                  (their number) pieces, so the batch fills a launch  */
    int32_t root_first;        /* > 0: the unit starts at level-0 step root_first. The
                                  steps before it run only until the first one that
                                  writes a child (that child keeps the host level-1
                                  role, as in the reference); those branches emit no
                                  clique and count in no statistic; 0 = from step 0   */
    int32_t launch_ms;         /* > 0: a task still running after launch_ms in its
                                  launch pushes the rest of its steps back onto the
                                  stack as a piece and the launch ends soon after
                                  (output unchanged); 0 = launches run to completion  */
    int32_t wave_tasks;        /* > 0: a batch whose tasks all have at most wave_tasks
                                  vertices runs one task per 64-lane wavefront instead
                                  of one per 256-thread workgroup (condensed graphs of
                                  at most 2048 vertices; output identical); 0 = the
                                  default (DQC_WAVE_TASKS_DEFAULT), negative = never  */
} dqc_config;

typedef struct {
    int64_t tasks_expanded;    /* tasks whose branch loop ran                         */
    int64_t branches;          /* add-one-vertex steps                                */
    int64_t raw_cliques;       /* cliques emitted before the maximality filter        */
    int64_t maximal_cliques;   /* lines of the maximal result (search / finish calls) */
    int64_t bytes_scanned;     /* CSR + task-record bytes the kernels touched         */
    int64_t launches;          /* expansion kernel launches                           */
    int64_t slow_sorts;        /* children ordered by the fallback sort               */
    double expand_kernel_ms;   /* sum of HIP-event times of expansion launches        */
    double expand_wall_ms;     /* host wall time of dqc_run                           */
    int32_t n_condensed;       /* vertices after level-0 condensing                   */
    int32_t workgroups;
    int64_t truncated_children;/* children not expanded (depth-bounded runs)          */
    int64_t stack_pushbacks;   /* tasks whose remaining steps were deferred because
                                  the task stack was nearly full                      */
    int64_t peak_stack_tasks;  /* largest task-stack use seen between launches        */
    int64_t peak_stack_records;
    double preprocess_ms;      /* search entry points: graph load (or adoption), two-hop
                                  lists, level 0 + condense and upload                 */
    double rmnonmax_ms;        /* search entry points: the maximal-set filter         */
    int64_t time_slices;       /* tasks whose rest went back onto the stack at the end
                                  of a launch_ms time slice                           */
    int64_t wave_launches;     /* expansion launches that ran one task per wavefront  */
} dqc_stats;

/* Receives each maximal set (ascending ids, sets in the DQC-R file's order). */
typedef void (*dqc_set_sink)(void* ctx, const int32_t* ids, int32_t len);

/* One record per expanded task (or piece of a split task): the task's id, its parent's
 * id (UINT32_MAX for tasks pushed by the host), depth, members, size, the min / max out-
 * and in-degree inside the task, and the branches this record's run made. Summing
 * branches over a task's descendants gives the subtree cost the scheduling model learns.
 * No counterpart in the reference (its tasks carry no identity). */
typedef struct {
    uint32_t uid, parent;
    int32_t depth, nm, tot;
    uint16_t dmin_out, dmax_out, dmin_in, dmax_in;
    uint32_t branches, flags;  /* flags & 1: a piece of a split task */
} dqc_task_log;

/* Creates a handle on cfg->device (NULL cfg = defaults). */
dqc_status dqc_open(const dqc_config* cfg, dqc_handle** out);
dqc_status dqc_close(dqc_handle* h);
const char* dqc_last_error(const dqc_handle* h);

/* Uploads a condensed graph: out / in / two-hop CSR with 64-bit offsets.
 * Replaces the CPU_Graph -> GPU copy of reference src/host_functions.cu:346-367. */
dqc_status dqc_upload_graph(dqc_handle* h, int32_t n,
                            const uint64_t* out_off, const int32_t* out_nbr,
                            const uint64_t* in_off, const int32_t* in_nbr,
                            const uint64_t* two_off, const int32_t* two_nbr);

/* Batched form: ngraphs condensed graphs share one launch sequence. Graph g's lists are
 * given as its own CSR (local ids); inside the library its ids are shifted by a 32-aligned
 * base, and every task keeps its graph, so each graph's enumeration is the one it has
 * alone. Tables are shared (one gamma pair and minimum size for the batch). */
dqc_status dqc_upload_graph_batch(dqc_handle* h, int32_t ngraphs, const int32_t* nverts,
                                  const uint64_t* const* out_off, const int32_t* const* out_nbr,
                                  const uint64_t* const* in_off, const int32_t* const* in_nbr,
                                  const uint64_t* const* two_off, const int32_t* const* two_nbr);
/* Queues graph gid's level-0 task (records with local ids, as dqc_push_root). */
dqc_status dqc_push_graph_root(dqc_handle* h, int32_t gid, const uint32_t* records, int32_t nvert, int32_t ncov);
/* As dqc_fetch_cliques, with each clique's graph in graph[] and local ids. */
dqc_status dqc_fetch_batch_cliques(dqc_handle* h, int64_t* count, int64_t* total_ids, int32_t* graph, int32_t* sizes,
                                   int32_t* ids);

/* Minimum-degree tables md[k] = ceil(gamma (k-1)) and bound tables ub[d] = (int)(d/gamma),
 * len entries each (reference src/host_functions.cu:9-15, :2165-2166). */
dqc_status dqc_set_tables(dqc_handle* h, const int32_t* md_out, const int32_t* md_in,
                          const int32_t* ub_out, const int32_t* ub_in, int32_t len,
                          int32_t min_size);

/* Queues the level-0 task: nvert records of 6 uint32 {id, out_mem, out_can, in_mem,
 * in_can, lvl2adj} in enumeration order, the first ncov being the covered vertices
 * (reference src/host_functions.cu:668-682 builds this task). */
dqc_status dqc_push_root(dqc_handle* h, const uint32_t* records, int32_t nvert, int32_t ncov);

/* Expands every queued task to completion on the GPU (replaces the level loop of
 * reference src/host_functions.cu:92-285, the d_expand_level / d_transfer_buffers /
 * d_fill_from_buffer kernels and the cuTS MPI stealing). sched is one of DQC_SCHED_*;
 * the enumeration output and the tree counts do not depend on it. */
/* dqc_config.wave_tasks when left 0: batches of tasks of at most this many vertices run
 * one task per wavefront */
#define DQC_WAVE_TASKS_DEFAULT 256
#define DQC_SCHED_DEFAULT 0 /* the handle's dqc_config.schedule                      */
#define DQC_SCHED_FIFO 1    /* plain: batches popped from the top of the LIFO task stack,
                               expanded in stack order, never split                   */
#define DQC_SCHED_SIZE 2    /* each batch ordered by a size estimate, largest first; the
                               costliest tasks (>= split_branches branches) are cut
                               into pieces; a batch with no task that long keeps stack
                               order                                                  */
#define DQC_SCHED_LEARNED 3 /* as SIZE, with the cost model of dqc_set_cost_model
                               (predicted branches x task size) as the estimate       */
dqc_status dqc_run(dqc_handle* h, int32_t sched);

/* Queues any task (same record layout as dqc_push_root): nm members first, then ncov
 * covered vertices, then candidates; mode 0 = host level 0, 1 = host level 1, 2 = device
 * level (reference src/host_functions.cu:987-1134 vs src/device_kernels.cu:25-173). */
dqc_status dqc_push_task(dqc_handle* h, const uint32_t* records, int32_t nvert, int32_t nm, int32_t ncov,
                         int32_t mode);

/* Reads the task stack: meta gets 9 int32 per task {record offset, size, members, mode,
 * id, parent id, depth, flags, covered} and records the 6-uint32 records. NULL buffers
 * query the sizes. */
dqc_status dqc_fetch_tasks(dqc_handle* h, int64_t* ntasks, int64_t* nrecords, int32_t* meta, uint32_t* records);

/* Empties the task stack and the clique store (the graph and tables stay). */
dqc_status dqc_reset(dqc_handle* h);

/* Task log for the scheduling model (no counterpart in the reference): keeps up to
 * `capacity` dqc_task_log records of the tasks expanded from now on (0 = off). */
dqc_status dqc_set_task_log(dqc_handle* h, int64_t capacity);
/* *count records are copied to `records` (NULL queries); *dropped = records beyond the
 * capacity. */
dqc_status dqc_fetch_task_log(dqc_handle* h, int64_t* count, int64_t* dropped, dqc_task_log* records);

/* Cost model of the learned schedule (no counterpart in the reference): an MLP
 * 7 -> 64 -> 64 -> 1 (ReLU, fp32) over the features of dqc_task_features, predicting
 * log(1 + branches) of a task: its own loop's (the shipped model; the schedule weighs it
 * by the task size) or its whole subtree's. Parameters, DQC_MLP_PARAMS floats: W1[64][7],
 * b1[64], W2[64][64], b2[64], W3[64], b3. Forward, gradient and training run as HIP
 * kernels on the current device; all buffers are host memory. */
#define DQC_MLP_IN 7
#define DQC_MLP_HIDDEN 64
#define DQC_MLP_PARAMS (64 * 7 + 64 + 64 * 64 + 64 + 64 + 1)
/* x[7 * i ..] = {log1p(members), log1p(candidates), log1p(depth), log1p(min out-degree),
 * log1p(max out-degree), log1p(min in-degree), log1p(max in-degree)} of record i. */
dqc_status dqc_task_features(const dqc_task_log* records, int64_t n, float* x);
dqc_status dqc_mlp_forward(const float* params, const float* x, int64_t n, float* y);
/* Gradient of the mean squared error mean_i (y_i - t_i)^2 over all n samples. */
dqc_status dqc_mlp_gradient(const float* params, const float* x, const float* t, int64_t n, float* grad,
                            double* loss);
/* Adam on minibatches of `batch` samples (shuffled by seed) for `epochs` passes; params
 * is updated in place; *loss = final full-data mean squared error. */
dqc_status dqc_mlp_train(float* params, const float* x, const float* t, int64_t n, int32_t epochs, int32_t batch,
                         float lr, uint64_t seed, double* loss);
/* Installs the model DQC_SCHED_LEARNED orders and splits batches by. */
dqc_status dqc_set_cost_model(dqc_handle* h, const float* params, int64_t n_params);

/* Runs at most `launches` expansion batches (one batch = the tasks popped from the top
 * of the task stack and expanded by one kernel launch); *finished = 1 once the stack is
 * empty. dqc_run is dqc_step until finished. */
dqc_status dqc_step(dqc_handle* h, int32_t sched, int32_t launches, int32_t* finished);

/* Per-phase realtime-counter ticks summed over workgroups (100 MHz counter), recorded
 * only when DQC_PHASES is set in the environment at dqc_open: 0 other, 1 remove,
 * 2 lookahead, 3 child setup + add + diameter, 4 degree pruning, 5 critical, 6 child
 * sort + write, 7 clique checks. No reference counterpart (profiling aid). */
dqc_status dqc_get_phases(const dqc_handle* h, uint64_t* ticks8);

/* Host preprocessing + upload for one text graph into an open handle: load, two-hop
 * lists, tables, level 0 + condense, then queue the level-0 task (reference
 * program.cu:93-115 and src/host_functions.cu:55-76). Ready for dqc_step / dqc_run.
 * Optional outputs: condensed vertex count. */
dqc_status dqc_prepare_file(dqc_handle* h, const char* graph_path, double out_gamma, double in_gamma,
                            int32_t min_size, int32_t* n_condensed);

/* Raw cliques found so far: sizes (count entries) and their ids concatenated.
 * Pass NULL buffers to query *count / *total_ids. */
dqc_status dqc_fetch_cliques(dqc_handle* h, int64_t* count, int64_t* total_ids,
                             int32_t* sizes, int32_t* ids);

dqc_status dqc_get_stats(const dqc_handle* h, dqc_stats* st);

/* Whole drop-in path of reference program.cu:93-170 for one process: load the text
 * graph, build two-hop lists, tables, level 0 + condense, GPU expansion, then write
 * raw_path (DQC-T_<tag> format) and max_path (DQC-R_<tag>, RemoveNonMax output; not
 * created when there are no cliques). cfg may be NULL. Returns the maximal count in
 * *max_cliques. */
dqc_status dqc_search_file(const char* graph_path, double out_gamma, double in_gamma,
                           int32_t min_size, const char* raw_path, const char* max_path,
                           const dqc_config* cfg, int64_t* max_cliques, dqc_stats* st);

/* h_calculate_minimum_degrees (reference inc/host_functions.hpp:7,
 * src/host_functions.cu:9-15): md[0] = 0, md[k] = ceil(gamma (k - 1)) in IEEE double for
 * k = 1..n (n + 1 entries). No GPU needed. */
dqc_status dqc_min_degrees(double gamma, int32_t n, int32_t* md);

/* In-memory form of the reference's h_search (inc/host_functions.hpp:8-10) over an
 * uncondensed CPU_Graph (inc/common.hpp:76-94): out / in CSR with 64-bit offsets, each
 * list strictly ascending without self loops, in-lists the transpose of the out-lists
 * (checked; DQC_E_ARG otherwise); two-hop lists as GenLevel2NBs builds them, or NULL
 * two_off / two_nbr to have them built here. md_out / md_in are the caller's minimum-
 * degree tables (h_calculate_minimum_degrees output, md_len entries, used verbatim;
 * indices beyond md_len continue the closed form with the same gamma; md_len 0 with
 * NULL tables: the closed form throughout). Level 0 + condense, the GPU expansion,
 * RemoveNonMax. Outputs: raw_path (DQC-T_ format, the temp_results stream of
 * h_search) and max_path (DQC-R_ format, not created when empty), each optional; sink,
 * when given, receives every maximal set; orig_of (optional, n entries) gets the input
 * id of each condensed id, which the result files use (the reference discards this
 * map, src/host_functions.cu:919). Unlike h_search, the caller's arrays are not
 * modified. */
dqc_status dqc_search_graph(int32_t n, const uint64_t* out_off, const int32_t* out_nbr,
                            const uint64_t* in_off, const int32_t* in_nbr,
                            const uint64_t* two_off, const int32_t* two_nbr,
                            const int32_t* md_out, const int32_t* md_in, int32_t md_len,
                            double out_gamma, double in_gamma, int32_t min_size,
                            const char* raw_path, const char* max_path,
                            dqc_set_sink sink, void* sink_ctx, const dqc_config* cfg,
                            int64_t* max_cliques, dqc_stats* st, int32_t* orig_of);
/* The preprocessing half of dqc_search_graph into an open handle (as dqc_prepare_file):
 * checks, level 0 + condense, upload, level-0 task queued; ready for dqc_step / dqc_run. */
dqc_status dqc_prepare_graph(dqc_handle* h, int32_t n, const uint64_t* out_off, const int32_t* out_nbr,
                             const uint64_t* in_off, const int32_t* in_nbr,
                             const uint64_t* two_off, const int32_t* two_nbr,
                             const int32_t* md_out, const int32_t* md_in, int32_t md_len,
                             double out_gamma, double in_gamma, int32_t min_size,
                             int32_t* n_condensed, int32_t* orig_of);

/* Binary graph cache (replaces re-parsing the text file, reference src/common.cpp:28-57):
 * loads graph_path (text or binary) and writes bin_path in the binary format, with the
 * two-hop lists when with_twohop (built if the input lacks them). Every entry point
 * that takes a graph path accepts either format (detected by its magic bytes); a binary
 * file with two-hop lists skips building them. No GPU needed. */
dqc_status dqc_save_graph(const char* graph_path, const char* bin_path, int32_t with_twohop);

/* Host preprocessing, no GPU needed. dqc_load_graph: text or binary loader (+ two-hop
 * lists when with_twohop) with the semantics of CPU_Graph::CPU_Graph / GenLevel2NBs
 * (reference src/common.cpp:15-509); arrays are malloc'd, release them with dqc_free. */
dqc_status dqc_load_graph(const char* graph_path, int32_t with_twohop, int32_t* n, int64_t* m,
                          uint64_t** out_off, int32_t** out_nbr, uint64_t** in_off, int32_t** in_nbr,
                          uint64_t** two_off, int32_t** two_nbr);
/* Level-0 peel / cover / condense (reference src/host_functions.cu:461-934): the level-0
 * task (6 uint32 per vertex, enumeration order), its covered count and the input id of
 * every condensed id. Arrays are malloc'd. */
dqc_status dqc_level0(const char* graph_path, double out_gamma, double in_gamma, int32_t min_size,
                      int32_t* n_condensed, int32_t* ncov, uint32_t** records, int32_t** orig_of);
void dqc_free(void* p);

/* Host preprocessing of a batch (each graph as dqc_prepare_file, one thread per graph)
 * + upload of all graphs and their level-0 tasks into one handle; n_condensed[g] gets
 * graph g's condensed size. */
dqc_status dqc_prepare_batch(dqc_handle* h, const char* const* graph_paths, int32_t ngraphs, double out_gamma,
                             double in_gamma, int32_t min_size, int32_t* n_condensed);
/* Re-aims an open handle's bounded unit of work (the dqc_config fields root_first,
 * root_branches and max_depth) for the tasks queued from now on, e.g. to run several
 * windows of one prepared graph (dqc_reset, then dqc_push_root again). */
dqc_status dqc_set_unit(dqc_handle* h, int32_t root_first, int32_t root_branches, int32_t max_depth);
/* Graphs in the handle's id space (1 for a single graph). */
dqc_status dqc_batch_size(const dqc_handle* h, int32_t* ngraphs);
/* Result dump of a batched run (the per-graph h_dump_cliques + RemoveNonMax of the
 * reference's path, src/host_functions.cu:1208-1237 and program.cu:145-170): each
 * graph's raw cliques through the maximal-set filter; max_paths[g] (NULL entries skip)
 * gets graph g's maximal file (not created when empty), max_cliques[g] its count;
 * st (optional) the handle's stats with maximal_cliques and rmnonmax_ms set. */
dqc_status dqc_finish_batch(dqc_handle* h, const char* const* max_paths, int64_t* max_cliques, dqc_stats* st);
/* The whole path for a batch of graphs in one handle: max_paths[g] (NULL entries skip)
 * gets graph g's maximal file in RemoveNonMax format, max_cliques[g] its count. */
dqc_status dqc_search_batch(const char* const* graph_paths, int32_t ngraphs, double out_gamma, double in_gamma,
                            int32_t min_size, const char* const* max_paths, const dqc_config* cfg,
                            int64_t* max_cliques, dqc_stats* st);

/* RemoveNonMax (reference inc/Quick_rmnonmax.h:6): returns the number of maximal sets
 * written, 0 when the input holds none (no output file then), <0 on I/O error. */
int64_t dqc_remove_non_max(const char* in_path, const char* out_path);

/* Seeded synthetic workloads 1..5 in the reference text format; graph_index selects
 * one of the 256 graphs of workload 3. Returns the vertex count or <0. */
int32_t dqc_generate(int32_t config, int32_t graph_index, const char* path);

/* ---- Allocation-free batch expansion (no counterpart in the reference) ------------
 * For callers that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc. One expansion launch of a batch of
 * tasks, as dqc_step runs it, with every input and output in the caller's device memory
 * and the kernel scratch (launch counters, per-workgroup degree records and lists) in a
 * static device array of the library (about 92 MB per GPU, held until the process
 * ends). Nothing is allocated and nothing waits for the host: the launches are queued on
 * `stream` (a hipStream_t; NULL = the null stream) and the outcome lands in `counters`.
 * Limits: condensed graphs of at most DQC_NOALLOC_MAX_VERTICES vertices (larger ones:
 * DQC_E_UNSUPPORTED), DQC_NOALLOC_WORKGROUPS workgroups. Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU race, so order them (one stream, or
 * synchronise between them). */
#define DQC_NOALLOC_MAX_VERTICES 4096
#define DQC_NOALLOC_WORKGROUPS 256
#define DQC_NOALLOC_COUNTERS 9

/* Task descriptor in device memory (56 bytes): records [offset, offset + size) of the
 * record array, members first, then `covered` covered vertices (level-0 tasks), then
 * candidates; mode 0 / 1 / 2 as dqc_push_task; steps = size - members - covered for a
 * whole task (flags 0, first_step 0); flags & 1 marks a piece (a deferred rest of a
 * task, as this call writes them: it runs `steps` loop steps from its state, flags & 2 =
 * its first step removes, flags & 8 = a level-0 piece whose task already wrote its first
 * child; first_step = the task's loop steps before the piece). id, parent, depth, graph
 * and the degree extremes are carried along. */
typedef struct {
    uint64_t offset;
    int32_t size, members, covered, mode, steps, flags;
    uint32_t id, parent;
    uint16_t depth, graph, dmin_out, dmax_out, dmin_in, dmax_in;
    int32_t first_step;
} dqc_task_desc;

/* A condensed graph and its tables in device memory (as dqc_upload_graph and
 * dqc_set_tables take them on the host); table_len >= 2 (max(n, min_size) + 2) + 2. */
typedef struct {
    int32_t n;
    const uint64_t* out_off;
    const int32_t* out_nbr;
    const uint64_t* in_off;
    const int32_t* in_nbr;
    const uint64_t* two_off;
    const int32_t* two_nbr;
    const int32_t *md_out, *md_in, *ub_out, *ub_in;
    int32_t table_len, min_size;
    int32_t host_addfail_check; /* as dqc_config.host_addfail_check                   */
} dqc_device_graph;

/* Expands tasks[0 .. ntasks) (records: 6 uint32 per vertex, nrecords in all; the
 * largest task has max_task_size vertices) one level. Children go to out_tasks /
 * out_records from index 0 up; when those buffers come close to full, a task's
 * remaining steps are written as a piece from the top down (descriptors
 * [out_task_cap - n, out_task_cap), records likewise) instead of running now: feed both
 * back as tasks. Cliques: clique_ids, clique_meta ((first id << 24) | length each).
 * counters (device, DQC_NOALLOC_COUNTERS uint64) receives {children, child records,
 * pieces, piece records, tasks expanded, branches, cliques, clique ids, error bits
 * (1 tasks did not fit, 2 cliques did not fit, 4 too many members)}. Returns DQC_OK once
 * the launches are queued. A variant of dqc_step (one batch, one launch). */
dqc_status dqc_expand_noalloc(const dqc_device_graph* graph, const dqc_task_desc* tasks,
                              const uint32_t* records, int64_t ntasks, int64_t nrecords,
                              int32_t max_task_size, dqc_task_desc* out_tasks, uint32_t* out_records,
                              int64_t out_task_cap, int64_t out_record_cap, int32_t* clique_ids,
                              uint64_t* clique_meta, int64_t clique_cap, int64_t clique_id_cap,
                              uint64_t* counters, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CUDQC_AMD_DQC_H */
