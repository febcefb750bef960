// Allocation-free one-batch expansion (dqc_expand_noalloc). Compiled as part of
// engine.hip's translation unit (included at its end), so that it launches the same
// expansion kernel as dqc_step.
//
// Callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops) pass every input and output in their own device memory; the scratch
// the kernel needs (launch counters and the per-workgroup value arrays and lists) lives
// in the file-scope device array `scratch` below, NOALLOC_BYTES in all. Each GPU the
// library runs on holds its own copy until the process ends. Calls that overlap on one
// GPU share that copy and race; calls must be ordered (one stream, or synchronised).

namespace {

constexpr int NOALLOC_KMAX = 4096;  // largest condensed graph (vertices)
constexpr int NOALLOC_WG = 256;     // workgroups per launch (one per CU)
constexpr unsigned long long NOALLOC_STRIDE = 22ull * NOALLOC_KMAX;  // uint32 words per workgroup (even)
constexpr unsigned long long NOALLOC_CTR_WORDS = 256;                // QCtr + GCtr, padded
constexpr unsigned long long NOALLOC_WORDS = NOALLOC_CTR_WORDS + NOALLOC_STRIDE * NOALLOC_WG;
static_assert(sizeof(QCtr) + sizeof(GCtr) <= NOALLOC_CTR_WORDS * 4, "counter area");
static_assert(NOALLOC_STRIDE % 2 == 0, "8-byte aligned workgroup areas");

// The scratch array. In host code `scratch` names a host-side stand-in, never device
// memory: host code uses only the per-device address obtained from
// hipGetSymbolAddress (noalloc_scratch()).
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
uint32_t* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per device.
hipError_t noalloc_scratch(uint32_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        noalloc_ptr[dev] = (uint32_t*)p;
    }
    *out = noalloc_ptr[dev];
    return hipSuccess;
}

// Sets the launch counters for a batch of `ntask` tasks and `nrec` records (every field
// the expansion reads; nothing is left from an earlier call). One thread.
__global__ void noalloc_init_kernel(QCtr* q, GCtr* g, unsigned long long ntask, unsigned long long nrec) {
    QCtr qc{};
    qc.root_first = ~0ull;
    qc.ntask = ntask;
    qc.pend = (ntask << CNT_SHIFT) | nrec;
    *q = qc;
    *g = GCtr{};
}

// Copies the outcome into the caller's counters (DQC_NOALLOC_COUNTERS entries).
__global__ void noalloc_finish_kernel(const QCtr* q, const GCtr* g, unsigned long long* out) {
    out[0] = q->sp_alloc >> CNT_SHIFT;
    out[1] = q->sp_alloc & SIZE_MASK;
    out[2] = q->def_alloc >> CNT_SHIFT;
    out[3] = q->def_alloc & SIZE_MASK;
    out[4] = g->tasks;
    out[5] = g->branches;
    out[6] = g->clq_alloc >> CNT_SHIFT;
    out[7] = g->clq_alloc & SIZE_MASK;
    out[8] = g->error;
}

}  // namespace

static_assert(sizeof(dqc_task_desc) == sizeof(TaskDesc) && offsetof(dqc_task_desc, offset) == offsetof(TaskDesc, off) &&
                  offsetof(dqc_task_desc, size) == offsetof(TaskDesc, tot) &&
                  offsetof(dqc_task_desc, members) == offsetof(TaskDesc, nm) &&
                  offsetof(dqc_task_desc, covered) == offsetof(TaskDesc, ncov) &&
                  offsetof(dqc_task_desc, mode) == offsetof(TaskDesc, mode) &&
                  offsetof(dqc_task_desc, steps) == offsetof(TaskDesc, iters) &&
                  offsetof(dqc_task_desc, flags) == offsetof(TaskDesc, flags) &&
                  offsetof(dqc_task_desc, id) == offsetof(TaskDesc, uid) &&
                  offsetof(dqc_task_desc, parent) == offsetof(TaskDesc, parent) &&
                  offsetof(dqc_task_desc, depth) == offsetof(TaskDesc, depth) &&
                  offsetof(dqc_task_desc, graph) == offsetof(TaskDesc, gid) &&
                  offsetof(dqc_task_desc, dmin_out) == offsetof(TaskDesc, dmin_o) &&
                  offsetof(dqc_task_desc, dmax_in) == offsetof(TaskDesc, dmax_i) &&
                  offsetof(dqc_task_desc, first_step) == offsetof(TaskDesc, step0),
              "dqc_task_desc mirrors the kernels' task descriptor");

extern "C" dqc_status dqc_expand_noalloc(const dqc_device_graph* graph, const dqc_task_desc* tasks,
                                         const uint32_t* records, int64_t ntasks, int64_t nrecords,
                                         int32_t max_task_size, dqc_task_desc* out_tasks, uint32_t* out_records,
                                         int64_t out_task_cap, int64_t out_record_cap, int32_t* clique_ids,
                                         uint64_t* clique_meta, int64_t clique_cap, int64_t clique_id_cap,
                                         uint64_t* counters, void* stream) {
    if (!graph || !counters || ntasks < 0 || nrecords < 0 || (ntasks > 0 && (!tasks || !records)) ||
        max_task_size < 0 || out_task_cap < 0 || out_record_cap < 0 || clique_cap < 0 || clique_id_cap < 0 ||
        !graph->out_off || !graph->in_off || !graph->two_off || !graph->md_out || !graph->md_in || !graph->ub_out ||
        !graph->ub_in || graph->min_size < 2 || graph->n < 0 ||
        graph->table_len < 2 * (std::max(graph->n, graph->min_size) + 2) + 2)
        return DQC_E_ARG;
    // the limits the scratch array is sized for
    if (graph->n > NOALLOC_KMAX || max_task_size > graph->n || ntasks > (int64_t)MAX_ENTRIES ||
        out_task_cap > (int64_t)MAX_ENTRIES || clique_cap > (int64_t)MAX_ENTRIES || out_record_cap > (int64_t)MAX_RECORDS ||
        clique_id_cap > (int64_t)MAX_RECORDS)
        return DQC_E_UNSUPPORTED;
    uint32_t* base = nullptr;
    if (noalloc_scratch(&base) != hipSuccess) return DQC_E_HIP;
    hipStream_t s = (hipStream_t)stream;
    QCtr* q = (QCtr*)base;
    GCtr* g = (GCtr*)((char*)base + ((sizeof(QCtr) + 255) / 256) * 256);
    noalloc_init_kernel<<<1, 1, 0, s>>>(q, g, (unsigned long long)ntasks, (unsigned long long)nrecords);
    if (ntasks > 0) {
        KParams P{};
        P.n = graph->n;
        P.oo = (const unsigned long long*)graph->out_off;
        P.io = (const unsigned long long*)graph->in_off;
        P.to = (const unsigned long long*)graph->two_off;
        P.on = graph->out_nbr; P.in = graph->in_nbr; P.tn = graph->two_nbr;
        P.mdo = graph->md_out; P.mdi = graph->md_in; P.ubo = graph->ub_out; P.ubi = graph->ub_in;
        P.min_size = graph->min_size;
        P.host_addfail = graph->host_addfail_check;
        P.profile = 0;
        P.in_desc = (const TaskDesc*)tasks;
        P.in_vtx = (const dqc::TaskVertex*)records;
        P.in_ntask = (unsigned long long)ntasks;
        P.sp_desc = (TaskDesc*)out_tasks;
        P.sp_vtx = (dqc::TaskVertex*)out_records;
        P.sp_cap_t = (unsigned long long)out_task_cap;
        P.sp_cap_v = (unsigned long long)out_record_cap;
        P.clq_ids = clique_ids;
        P.clq_meta = (unsigned long long*)clique_meta;
        P.clq_cap_s = (unsigned long long)clique_cap;
        P.clq_cap_i = (unsigned long long)clique_id_cap;
        P.q = q;
        P.g = g;
        P.split_s = nullptr;
        P.lto = nullptr;
        P.ltn = nullptr;
        P.scratch = base + NOALLOC_CTR_WORDS;
        P.stride = NOALLOC_STRIDE;
        P.log = nullptr;
        P.log_cap = 0;
        P.kloc = std::max(graph->n, 1);
        P.sum_shift = 0;
        P.max_depth = 0;
        P.root_steps = 0;
        P.maxtot = (unsigned long long)std::max(max_task_size, 1);
        P.gbase = nullptr;
        const int grid = (int)std::min<int64_t>(NOALLOC_WG, ntasks);
        expand_kernel<false><<<grid, NT, 0, s>>>(P);
    }
    noalloc_finish_kernel<<<1, 1, 0, s>>>(q, g, (unsigned long long*)counters);
    return hipGetLastError() == hipSuccess ? DQC_OK : DQC_E_HIP;
}
