"""Python binding of the MI355X maximal directed quasi-clique enumerator.

The product is the native library ``cudqc_amd/lib/libcudqc_amd.so`` (C-ABI declared in
``include/dqc.h``); this module is a thin ctypes layer over it for tests and benchmarks.
Nothing here computes results on the CPU: if the library is missing, loading fails.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcudqc_amd.so")
CLI_PATH = os.path.join(_HERE, "bin", "cudqc")

DQC_OK = 0
STATUS = {0: "DQC_OK", 1: "DQC_E_HIP", 2: "DQC_E_CAPACITY", 3: "DQC_E_ARG", 4: "DQC_E_IO", 5: "DQC_E_UNSUPPORTED"}

# every symbol include/dqc.h declares
EXPORTS = [
    "dqc_open", "dqc_close", "dqc_last_error", "dqc_upload_graph", "dqc_set_tables", "dqc_push_root",
    "dqc_run", "dqc_fetch_cliques", "dqc_get_stats", "dqc_search_file", "dqc_remove_non_max", "dqc_generate",
    "dqc_load_graph", "dqc_level0", "dqc_free", "dqc_step", "dqc_get_phases", "dqc_prepare_file",
    "dqc_push_task", "dqc_fetch_tasks", "dqc_reset",
]


class Config(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32), ("workgroups", ctypes.c_int32),
        ("pool_tasks", ctypes.c_uint64), ("pool_vertices", ctypes.c_uint64),
        ("spill_tasks", ctypes.c_uint64), ("spill_vertices", ctypes.c_uint64),
        ("clique_sets", ctypes.c_uint64), ("clique_ids", ctypes.c_uint64),
        ("host_addfail_check", ctypes.c_int32),
        ("split_branches", ctypes.c_int32), ("split_chunk", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
    ]


SCHED_DEFAULT, SCHED_FIFO, SCHED_SIZE, SCHED_LEARNED = 0, 1, 2, 3


class Stats(ctypes.Structure):
    _fields_ = [
        ("tasks_expanded", ctypes.c_int64), ("branches", ctypes.c_int64), ("raw_cliques", ctypes.c_int64),
        ("maximal_cliques", ctypes.c_int64), ("bytes_scanned", ctypes.c_int64), ("launches", ctypes.c_int64),
        ("slow_sorts", ctypes.c_int64), ("expand_kernel_ms", ctypes.c_double), ("expand_wall_ms", ctypes.c_double),
        ("n_condensed", ctypes.c_int32), ("workgroups", ctypes.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class DQCError(RuntimeError):
    pass


_lib = None


def load():
    """Loads the native library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DQCError(f"native library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    lib.dqc_search_file.restype = ctypes.c_int32
    lib.dqc_search_file.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                    ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Config),
                                    ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(Stats)]
    lib.dqc_remove_non_max.restype = ctypes.c_int64
    lib.dqc_remove_non_max.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    lib.dqc_generate.restype = ctypes.c_int32
    lib.dqc_generate.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p]
    lib.dqc_last_error.restype = ctypes.c_char_p
    P64, P32, PU32 = ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint32)
    lib.dqc_load_graph.restype = ctypes.c_int32
    lib.dqc_load_graph.argtypes = [ctypes.c_char_p, ctypes.c_int32, P32, ctypes.POINTER(ctypes.c_int64)] + \
        [ctypes.POINTER(P64), ctypes.POINTER(P32)] * 3
    lib.dqc_level0.restype = ctypes.c_int32
    lib.dqc_level0.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_int32, P32, P32,
                               ctypes.POINTER(PU32), ctypes.POINTER(P32)]
    lib.dqc_free.argtypes = [ctypes.c_void_p]
    lib.dqc_open.restype = ctypes.c_int32
    lib.dqc_open.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)]
    lib.dqc_close.restype = ctypes.c_int32
    lib.dqc_close.argtypes = [ctypes.c_void_p]
    lib.dqc_last_error.argtypes = [ctypes.c_void_p]
    lib.dqc_prepare_file.restype = ctypes.c_int32
    lib.dqc_prepare_file.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int32, P32]
    lib.dqc_step.restype = ctypes.c_int32
    lib.dqc_step.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, P32]
    lib.dqc_run.restype = ctypes.c_int32
    lib.dqc_run.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    lib.dqc_get_stats.restype = ctypes.c_int32
    lib.dqc_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
    lib.dqc_get_phases.restype = ctypes.c_int32
    lib.dqc_get_phases.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    lib.dqc_push_task.restype = ctypes.c_int32
    lib.dqc_push_task.argtypes = [ctypes.c_void_p, PU32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    lib.dqc_fetch_tasks.restype = ctypes.c_int32
    lib.dqc_fetch_tasks.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), P32, PU32]
    lib.dqc_reset.restype = ctypes.c_int32
    lib.dqc_reset.argtypes = [ctypes.c_void_p]
    lib.dqc_fetch_cliques.restype = ctypes.c_int32
    lib.dqc_fetch_cliques.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                      P32, P32]
    _lib = lib
    return lib


def _b(p):
    return None if p is None else os.fsencode(p)


def search_file(graph_path, out_gamma, in_gamma, min_size, raw_path=None, max_path=None, config=None):
    """GPU enumeration of one graph file (reference program.cu path). Returns
    (maximal clique count, Stats)."""
    lib = load()
    cnt = ctypes.c_int64(0)
    st = Stats()
    cfg = ctypes.byref(config) if config is not None else None
    rc = lib.dqc_search_file(_b(graph_path), float(out_gamma), float(in_gamma), int(min_size), _b(raw_path),
                             _b(max_path), cfg, ctypes.byref(cnt), ctypes.byref(st))
    if rc != DQC_OK:
        raise DQCError(f"dqc_search_file failed: {STATUS.get(rc, rc)}")
    return cnt.value, st


class Engine:
    """A device handle (dqc_open) driven step by step: prepare one graph, then expand the
    task stack in batches. Used by the benchmark to time a bounded number of batches."""

    def __init__(self, config=None):
        self.lib = load()
        self.h = ctypes.c_void_p()
        rc = self.lib.dqc_open(ctypes.byref(config) if config is not None else None, ctypes.byref(self.h))
        self._check(rc, "dqc_open")

    def _check(self, rc, what):
        if rc != DQC_OK:
            msg = self.lib.dqc_last_error(self.h) if self.h else b""
            raise DQCError(f"{what} failed: {STATUS.get(rc, rc)}: {msg.decode(errors='replace') if msg else ''}")

    def prepare(self, graph_path, out_gamma, in_gamma, min_size):
        k = ctypes.c_int32()
        self._check(self.lib.dqc_prepare_file(self.h, _b(graph_path), out_gamma, in_gamma, int(min_size),
                                              ctypes.byref(k)), "dqc_prepare_file")
        return k.value

    def step(self, launches=1, sched=SCHED_DEFAULT):
        fin = ctypes.c_int32()
        self._check(self.lib.dqc_step(self.h, int(sched), int(launches), ctypes.byref(fin)), "dqc_step")
        return bool(fin.value)

    def run(self, sched=SCHED_DEFAULT):
        self._check(self.lib.dqc_run(self.h, int(sched)), "dqc_run")

    def stats(self):
        st = Stats()
        self._check(self.lib.dqc_get_stats(self.h, ctypes.byref(st)), "dqc_get_stats")
        return st

    def phases(self):
        arr = (ctypes.c_uint64 * 8)()
        self._check(self.lib.dqc_get_phases(self.h, arr), "dqc_get_phases")
        return list(arr)

    def cliques(self):
        cnt, tot = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.dqc_fetch_cliques(self.h, ctypes.byref(cnt), ctypes.byref(tot), None, None), "fetch")
        sizes = (ctypes.c_int32 * max(cnt.value, 1))()
        ids = (ctypes.c_int32 * max(tot.value, 1))()
        self._check(self.lib.dqc_fetch_cliques(self.h, ctypes.byref(cnt), ctypes.byref(tot), sizes, ids), "fetch")
        out, p = [], 0
        for i in range(cnt.value):
            out.append(list(ids[p:p + sizes[i]]))
            p += sizes[i]
        return out

    def reset(self):
        self._check(self.lib.dqc_reset(self.h), "dqc_reset")

    def push_task(self, records, nm, ncov=0, mode=2):
        """records: list of 6-tuples (id, out_mem, out_can, in_mem, in_can, lvl2adj)."""
        flat = (ctypes.c_uint32 * (6 * len(records)))(*[x for r in records for x in r])
        self._check(self.lib.dqc_push_task(self.h, flat, len(records), nm, ncov, mode), "dqc_push_task")

    def tasks(self):
        """The task stack as a list of (records, nm, mode)."""
        nt, nr = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), None, None), "fetch")
        meta = (ctypes.c_int32 * max(4 * nt.value, 1))()
        rec = (ctypes.c_uint32 * max(6 * nr.value, 1))()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), meta, rec), "fetch")
        out = []
        for i in range(nt.value):
            off, tot, nm, mode = meta[4 * i:4 * i + 4]
            recs = [tuple(rec[6 * (off + j):6 * (off + j) + 6]) for j in range(tot)]
            out.append((recs, nm, mode))
        return out

    def close(self):
        if self.h:
            self.lib.dqc_close(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def remove_non_max(in_path, out_path):
    """RemoveNonMax (reference inc/Quick_rmnonmax.h:6) on a raw clique file."""
    return load().dqc_remove_non_max(_b(in_path), _b(out_path))


def generate(config, path, graph_index=0):
    """Writes synthetic workload `config` (1..5) to `path`; returns its vertex count."""
    n = load().dqc_generate(int(config), int(graph_index), _b(path))
    if n < 0:
        raise DQCError(f"dqc_generate({config}, {graph_index}) failed: {n}")
    return n


def load_graph(path, with_twohop=True):
    """Host loader (+ two-hop lists): dict(n, m, out, in, two) of python adjacency lists."""
    lib = load()
    n = ctypes.c_int32()
    m = ctypes.c_int64()
    P64, P32 = ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32)
    ptrs = [P64() if i % 2 == 0 else P32() for i in range(6)]
    rc = lib.dqc_load_graph(_b(path), int(with_twohop), ctypes.byref(n), ctypes.byref(m), *[ctypes.byref(x) for x in ptrs])
    if rc != DQC_OK:
        raise DQCError(f"dqc_load_graph failed: {STATUS.get(rc, rc)}")
    N = n.value
    res = {"n": N, "m": m.value}
    for k, name in enumerate(["out", "in", "two"]):
        off = [ptrs[2 * k][i] for i in range(N + 1)]
        nb = [ptrs[2 * k + 1][i] for i in range(off[-1])] if N else []
        res[name] = [nb[off[i]:off[i + 1]] for i in range(N)]
    for x in ptrs:
        lib.dqc_free(x)
    return res


def level0(path, out_gamma, in_gamma, min_size):
    """Host level-0 task: (records as tuples (id, om, oc, im, ic, l2), ncov, orig_of)."""
    lib = load()
    k = ctypes.c_int32()
    ncov = ctypes.c_int32()
    rec = ctypes.POINTER(ctypes.c_uint32)()
    oo = ctypes.POINTER(ctypes.c_int32)()
    rc = lib.dqc_level0(_b(path), out_gamma, in_gamma, int(min_size), ctypes.byref(k), ctypes.byref(ncov),
                        ctypes.byref(rec), ctypes.byref(oo))
    if rc != DQC_OK:
        raise DQCError(f"dqc_level0 failed: {STATUS.get(rc, rc)}")
    K = k.value
    records = [tuple(rec[6 * i + j] for j in range(6)) for i in range(K)]
    orig = [oo[i] for i in range(K)]
    lib.dqc_free(rec)
    lib.dqc_free(oo)
    return records, ncov.value, orig


# Workload parameters (CLI order: out gamma, in gamma, min size)
WORKLOADS = {
    1: (0.8, 0.8, 10),
    2: (0.9, 0.9, 12),
    3: (0.85, 0.85, 10),
    4: (0.9, 0.9, 18),
    5: (0.9, 0.95, 20),
}
