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
    "dqc_push_task", "dqc_fetch_tasks", "dqc_reset", "dqc_set_task_log", "dqc_fetch_task_log",
    "dqc_task_features", "dqc_mlp_forward", "dqc_mlp_gradient", "dqc_mlp_train", "dqc_set_cost_model",
    "dqc_upload_graph_batch", "dqc_push_graph_root", "dqc_fetch_batch_cliques", "dqc_prepare_batch",
    "dqc_search_batch", "dqc_min_degrees", "dqc_search_graph", "dqc_prepare_graph", "dqc_save_graph",
    "dqc_expand_noalloc", "dqc_batch_size", "dqc_finish_batch", "dqc_set_unit",
]

MLP_IN, MLP_HIDDEN = 7, 64
MLP_PARAMS = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN + 1


class TaskLog(ctypes.Structure):
    """dqc_task_log: one expanded task (or piece of a split task)."""
    _fields_ = [("uid", ctypes.c_uint32), ("parent", ctypes.c_uint32), ("depth", ctypes.c_int32),
                ("nm", ctypes.c_int32), ("tot", ctypes.c_int32), ("dmin_out", ctypes.c_uint16),
                ("dmax_out", ctypes.c_uint16), ("dmin_in", ctypes.c_uint16), ("dmax_in", ctypes.c_uint16),
                ("branches", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class Config(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32), ("workgroups", ctypes.c_int32),
        ("pool_tasks", ctypes.c_uint64), ("pool_vertices", ctypes.c_uint64),
        ("spill_tasks", ctypes.c_uint64), ("spill_vertices", ctypes.c_uint64),
        ("clique_sets", ctypes.c_uint64), ("clique_ids", ctypes.c_uint64),
        ("host_addfail_check", ctypes.c_int32),
        ("split_branches", ctypes.c_int32), ("split_chunk", ctypes.c_int32),
        ("schedule", ctypes.c_int32), ("max_depth", ctypes.c_int32), ("resident_steps", ctypes.c_int32),
        ("root_branches", ctypes.c_int32), ("split_pieces", ctypes.c_int32), ("root_first", ctypes.c_int32),
        ("launch_ms", ctypes.c_int32),
        ("wave_tasks", ctypes.c_int32),
    ]


SCHED_DEFAULT, SCHED_FIFO, SCHED_SIZE, SCHED_LEARNED = 0, 1, 2, 3


class Stats(ctypes.Structure):
    _fields_ = [
        ("tasks_expanded", ctypes.c_int64), ("branches", ctypes.c_int64), ("raw_cliques", ctypes.c_int64),
        ("maximal_cliques", ctypes.c_int64), ("bytes_scanned", ctypes.c_int64), ("launches", ctypes.c_int64),
        ("slow_sorts", ctypes.c_int64), ("expand_kernel_ms", ctypes.c_double), ("expand_wall_ms", ctypes.c_double),
        ("n_condensed", ctypes.c_int32), ("workgroups", ctypes.c_int32),
        ("truncated_children", ctypes.c_int64), ("stack_pushbacks", ctypes.c_int64),
        ("peak_stack_tasks", ctypes.c_int64), ("peak_stack_records", ctypes.c_int64),
        ("preprocess_ms", ctypes.c_double), ("rmnonmax_ms", ctypes.c_double), ("time_slices", ctypes.c_int64),
        ("wave_launches", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class DQCError(RuntimeError):
    pass


# dqc_set_sink: receives each maximal set
SET_SINK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32)


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
    lib.dqc_push_root.restype = ctypes.c_int32
    lib.dqc_push_root.argtypes = [ctypes.c_void_p, PU32, ctypes.c_int32, ctypes.c_int32]
    lib.dqc_push_task.restype = ctypes.c_int32
    lib.dqc_push_task.argtypes = [ctypes.c_void_p, PU32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    lib.dqc_fetch_tasks.restype = ctypes.c_int32
    lib.dqc_fetch_tasks.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), P32, PU32]
    lib.dqc_reset.restype = ctypes.c_int32
    lib.dqc_reset.argtypes = [ctypes.c_void_p]
    lib.dqc_fetch_cliques.restype = ctypes.c_int32
    lib.dqc_fetch_cliques.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                      P32, P32]
    PF = ctypes.POINTER(ctypes.c_float)
    PD = ctypes.POINTER(ctypes.c_double)
    PI64 = ctypes.POINTER(ctypes.c_int64)
    lib.dqc_set_task_log.restype = ctypes.c_int32
    lib.dqc_set_task_log.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    lib.dqc_fetch_task_log.restype = ctypes.c_int32
    lib.dqc_fetch_task_log.argtypes = [ctypes.c_void_p, PI64, PI64, ctypes.POINTER(TaskLog)]
    lib.dqc_task_features.restype = ctypes.c_int32
    lib.dqc_task_features.argtypes = [ctypes.POINTER(TaskLog), ctypes.c_int64, PF]
    lib.dqc_mlp_forward.restype = ctypes.c_int32
    lib.dqc_mlp_forward.argtypes = [PF, PF, ctypes.c_int64, PF]
    lib.dqc_mlp_gradient.restype = ctypes.c_int32
    lib.dqc_mlp_gradient.argtypes = [PF, PF, PF, ctypes.c_int64, PF, PD]
    lib.dqc_mlp_train.restype = ctypes.c_int32
    lib.dqc_mlp_train.argtypes = [PF, PF, PF, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                  ctypes.c_uint64, PD]
    lib.dqc_set_cost_model.restype = ctypes.c_int32
    lib.dqc_set_cost_model.argtypes = [ctypes.c_void_p, PF, ctypes.c_int64]
    PPC = ctypes.POINTER(ctypes.c_char_p)
    lib.dqc_prepare_batch.restype = ctypes.c_int32
    lib.dqc_prepare_batch.argtypes = [ctypes.c_void_p, PPC, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int32, P32]
    lib.dqc_search_batch.restype = ctypes.c_int32
    lib.dqc_search_batch.argtypes = [PPC, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_int32, PPC,
                                     ctypes.POINTER(Config), PI64, ctypes.POINTER(Stats)]
    lib.dqc_fetch_batch_cliques.restype = ctypes.c_int32
    lib.dqc_fetch_batch_cliques.argtypes = [ctypes.c_void_p, PI64, PI64, P32, P32, P32]
    lib.dqc_min_degrees.restype = ctypes.c_int32
    lib.dqc_min_degrees.argtypes = [ctypes.c_double, ctypes.c_int32, P32]
    csr = [P64, P32, P64, P32, P64, P32]
    lib.dqc_search_graph.restype = ctypes.c_int32
    lib.dqc_search_graph.argtypes = [ctypes.c_int32] + csr + [P32, P32, ctypes.c_int32, ctypes.c_double,
                                                              ctypes.c_double, ctypes.c_int32, ctypes.c_char_p,
                                                              ctypes.c_char_p, SET_SINK, ctypes.c_void_p,
                                                              ctypes.POINTER(Config), PI64, ctypes.POINTER(Stats), P32]
    lib.dqc_prepare_graph.restype = ctypes.c_int32
    lib.dqc_prepare_graph.argtypes = [ctypes.c_void_p, ctypes.c_int32] + csr + [P32, P32, ctypes.c_int32,
                                                                                ctypes.c_double, ctypes.c_double,
                                                                                ctypes.c_int32, P32, P32]
    lib.dqc_finish_batch.restype = ctypes.c_int32
    lib.dqc_finish_batch.argtypes = [ctypes.c_void_p, PPC, PI64, ctypes.POINTER(Stats)]
    lib.dqc_set_unit.restype = ctypes.c_int32
    lib.dqc_set_unit.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    lib.dqc_batch_size.restype = ctypes.c_int32
    lib.dqc_batch_size.argtypes = [ctypes.c_void_p, P32]
    lib.dqc_save_graph.restype = ctypes.c_int32
    lib.dqc_save_graph.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32]
    _lib = lib
    return lib


def _b(p):
    return None if p is None else os.fsencode(p)


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def task_features(records):
    """7 features per TaskLog record (numpy float32 array n x 7)."""
    import numpy as np
    n = len(records)
    arr = (TaskLog * max(n, 1))(*records)
    x = np.zeros((max(n, 1), MLP_IN), dtype=np.float32)
    rc = load().dqc_task_features(arr, n, _fp(x))
    if rc != DQC_OK:
        raise DQCError(f"dqc_task_features failed: {STATUS.get(rc, rc)}")
    return x[:n]


def mlp_forward(params, x):
    """Cost-model forward pass on the GPU: predictions for the rows of x (n x 7)."""
    import numpy as np
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros(max(len(xx), 1), dtype=np.float32)
    rc = load().dqc_mlp_forward(_fp(p), _fp(xx), len(xx), _fp(y))
    if rc != DQC_OK:
        raise DQCError(f"dqc_mlp_forward failed: {STATUS.get(rc, rc)}")
    return y[:len(xx)]


def mlp_gradient(params, x, t):
    """Gradient (and value) of the mean squared error on the GPU."""
    import numpy as np
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    tt = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(MLP_PARAMS, dtype=np.float32)
    loss = ctypes.c_double()
    rc = load().dqc_mlp_gradient(_fp(p), _fp(xx), _fp(tt), len(xx), _fp(g), ctypes.byref(loss))
    if rc != DQC_OK:
        raise DQCError(f"dqc_mlp_gradient failed: {STATUS.get(rc, rc)}")
    return g, loss.value


def mlp_train(params, x, t, epochs=20, batch=1024, lr=1e-3, seed=1):
    """Adam training on the GPU; returns (new params, final mean squared error)."""
    import numpy as np
    p = np.array(params, dtype=np.float32, copy=True)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    tt = np.ascontiguousarray(t, dtype=np.float32)
    loss = ctypes.c_double()
    rc = load().dqc_mlp_train(_fp(p), _fp(xx), _fp(tt), len(xx), int(epochs), int(batch), float(lr), int(seed),
                              ctypes.byref(loss))
    if rc != DQC_OK:
        raise DQCError(f"dqc_mlp_train failed: {STATUS.get(rc, rc)}")
    return p, loss.value


def mlp_init(seed=0):
    """He-initialised parameters (W1, b1, W2, b2, W3, b3 flattened)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    h, i = MLP_HIDDEN, MLP_IN
    parts = [rng.normal(0, np.sqrt(2.0 / i), h * i), np.zeros(h), rng.normal(0, np.sqrt(2.0 / h), h * h),
             np.zeros(h), rng.normal(0, np.sqrt(1.0 / h), h), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


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


def _paths(paths):
    return (ctypes.c_char_p * max(len(paths), 1))(*[_b(p) for p in paths])


def search_batch(graph_paths, out_gamma, in_gamma, min_size, max_paths=None, config=None):
    """GPU enumeration of several graphs in one handle (batched form). Returns (list of
    maximal counts per graph, Stats)."""
    lib = load()
    n = len(graph_paths)
    cnt = (ctypes.c_int64 * max(n, 1))()
    st = Stats()
    mp = _paths(max_paths) if max_paths is not None else None
    rc = lib.dqc_search_batch(_paths(graph_paths), n, float(out_gamma), float(in_gamma), int(min_size), mp,
                              ctypes.byref(config) if config is not None else None, cnt, ctypes.byref(st))
    if rc != DQC_OK:
        raise DQCError(f"dqc_search_batch failed: {STATUS.get(rc, rc)}")
    return list(cnt[:n]), st


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

    def prepare_graph(self, g, out_gamma, in_gamma, min_size, md_out=None, md_in=None):
        """dqc_prepare_graph on a graph given as a dict of adjacency lists (out, in, and
        optionally two); returns the condensed vertex count."""
        n = len(g["out"])
        oo, on = _csr_arrays(g["out"])
        io, inn = _csr_arrays(g["in"])
        to, tn = _csr_arrays(g["two"]) if g.get("two") is not None else (None, None)
        ml = 0
        mo = mi = None
        if md_out is not None:
            ml = min(len(md_out), len(md_in))
            mo = (ctypes.c_int32 * ml)(*md_out[:ml])
            mi = (ctypes.c_int32 * ml)(*md_in[:ml])
        k = ctypes.c_int32()
        self._check(self.lib.dqc_prepare_graph(self.h, n, oo, on, io, inn, to, tn, mo, mi, ml, float(out_gamma),
                                               float(in_gamma), int(min_size), ctypes.byref(k), None),
                    "dqc_prepare_graph")
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

    def push_root(self, records, ncov):
        """Queues a level-0 task (expanded branch-parallel on a single graph)."""
        flat = (ctypes.c_uint32 * (6 * len(records)))(*[x for r in records for x in r])
        self._check(self.lib.dqc_push_root(self.h, flat, len(records), ncov), "dqc_push_root")

    def push_task(self, records, nm, ncov=0, mode=2):
        """records: list of 6-tuples (id, out_mem, out_can, in_mem, in_can, lvl2adj)."""
        flat = (ctypes.c_uint32 * (6 * len(records)))(*[x for r in records for x in r])
        self._check(self.lib.dqc_push_task(self.h, flat, len(records), nm, ncov, mode), "dqc_push_task")

    def tasks(self):
        """The task stack as a list of (records, nm, mode)."""
        nt, nr = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), None, None), "fetch")
        meta = (ctypes.c_int32 * max(9 * nt.value, 1))()
        rec = (ctypes.c_uint32 * max(6 * nr.value, 1))()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), meta, rec), "fetch")
        out = []
        for i in range(nt.value):
            off, tot, nm, mode = meta[9 * i:9 * i + 4]
            recs = [tuple(rec[6 * (off + j):6 * (off + j) + 6]) for j in range(tot)]
            out.append((recs, nm, mode))
        return out

    def prepare_batch(self, graph_paths, out_gamma, in_gamma, min_size):
        n = len(graph_paths)
        k = (ctypes.c_int32 * max(n, 1))()
        self._check(self.lib.dqc_prepare_batch(self.h, _paths(graph_paths), n, out_gamma, in_gamma, int(min_size), k),
                    "dqc_prepare_batch")
        return list(k[:n])

    def batch_cliques(self):
        """Raw cliques as (graph, local ids) pairs."""
        cnt, tot = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.dqc_fetch_batch_cliques(self.h, ctypes.byref(cnt), ctypes.byref(tot), None, None, None),
                    "fetch")
        g = (ctypes.c_int32 * max(cnt.value, 1))()
        sizes = (ctypes.c_int32 * max(cnt.value, 1))()
        ids = (ctypes.c_int32 * max(tot.value, 1))()
        self._check(self.lib.dqc_fetch_batch_cliques(self.h, ctypes.byref(cnt), ctypes.byref(tot), g, sizes, ids), "fetch")
        out, p = [], 0
        for i in range(cnt.value):
            out.append((g[i], list(ids[p:p + sizes[i]])))
            p += sizes[i]
        return out

    def set_unit(self, root_first, root_branches, max_depth):
        """dqc_set_unit: the bounded unit of work for tasks queued from now on."""
        self._check(self.lib.dqc_set_unit(self.h, int(root_first), int(root_branches), int(max_depth)),
                    "dqc_set_unit")

    def finish_batch(self, max_paths):
        """dqc_finish_batch: each graph's maximal file (None entries skip); returns
        (per-graph maximal counts, stats)."""
        n = len(max_paths)
        arr = (ctypes.c_char_p * n)(*[_b(p) for p in max_paths])
        counts = (ctypes.c_int64 * n)()
        st = Stats()
        self._check(self.lib.dqc_finish_batch(self.h, arr, counts, ctypes.byref(st)), "dqc_finish_batch")
        return list(counts), st

    def set_task_log(self, capacity):
        self._check(self.lib.dqc_set_task_log(self.h, int(capacity)), "dqc_set_task_log")

    def task_log(self):
        """The task log as a ctypes array of TaskLog (and the count dropped beyond capacity)."""
        n, dropped = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.dqc_fetch_task_log(self.h, ctypes.byref(n), ctypes.byref(dropped), None), "fetch log")
        arr = (TaskLog * max(n.value, 1))()
        self._check(self.lib.dqc_fetch_task_log(self.h, ctypes.byref(n), ctypes.byref(dropped), arr), "fetch log")
        return arr[:n.value], dropped.value

    def task_ids(self):
        """(uid, parent) of every task on the stack."""
        nt, nr = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), None, None), "fetch")
        meta = (ctypes.c_int32 * max(9 * nt.value, 1))()
        rec = (ctypes.c_uint32 * max(6 * nr.value, 1))()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), meta, rec), "fetch")
        return [(meta[9 * i + 4] & 0xFFFFFFFF, meta[9 * i + 5] & 0xFFFFFFFF) for i in range(nt.value)]

    def task_digests(self):
        """(SHA-1 of the records' bytes, nm, mode) of every task on the stack: a compact
        form of tasks() for large tasks (records are 6 uint32 each)."""
        import hashlib
        nt, nr = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), None, None), "fetch")
        meta = (ctypes.c_int32 * max(9 * nt.value, 1))()
        rec = (ctypes.c_uint32 * max(6 * nr.value, 1))()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), meta, rec), "fetch")
        mv = memoryview(rec).cast("B")
        out = []
        for i in range(nt.value):
            off, tot, nm, mode = meta[9 * i:9 * i + 4]
            out.append((hashlib.sha1(mv[24 * off:24 * (off + tot)]).hexdigest(), nm, mode))
        return out

    def task_meta(self):
        """(records, nm, mode, ncov) of every task on the stack."""
        nt, nr = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), None, None), "fetch")
        meta = (ctypes.c_int32 * max(9 * nt.value, 1))()
        rec = (ctypes.c_uint32 * max(6 * nr.value, 1))()
        self._check(self.lib.dqc_fetch_tasks(self.h, ctypes.byref(nt), ctypes.byref(nr), meta, rec), "fetch")
        out = []
        for i in range(nt.value):
            off, tot, nm, mode = meta[9 * i:9 * i + 4]
            out.append(([tuple(rec[6 * (off + j):6 * (off + j) + 6]) for j in range(tot)], nm, mode, meta[9 * i + 8]))
        return out

    def set_cost_model(self, params):
        import numpy as np
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.lib.dqc_set_cost_model(self.h, _fp(p), p.size), "dqc_set_cost_model")

    def close(self):
        if self.h:
            self.lib.dqc_close(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def min_degrees(gamma, n):
    """h_calculate_minimum_degrees: md[0..n] (reference src/host_functions.cu:9-15)."""
    arr = (ctypes.c_int32 * (n + 1))()
    rc = load().dqc_min_degrees(float(gamma), int(n), arr)
    if rc != DQC_OK:
        raise DQCError(f"dqc_min_degrees failed: {STATUS.get(rc, rc)}")
    return list(arr)


def _csr_arrays(lists):
    """Python adjacency lists -> (ctypes uint64 offsets, ctypes int32 ids)."""
    off = [0]
    for l in lists:
        off.append(off[-1] + len(l))
    o = (ctypes.c_uint64 * len(off))(*off)
    flat = [x for l in lists for x in l]
    ids = (ctypes.c_int32 * max(len(flat), 1))(*flat)
    return o, ids


def search_graph(out_lists, in_lists, two_lists, md_out, md_in, out_gamma, in_gamma, min_size, raw_path=None,
                 max_path=None, config=None, want_orig=False):
    """In-memory h_search (dqc_search_graph): adjacency lists of an uncondensed graph
    (two_lists None: built by the library), the caller's minimum-degree tables (None:
    the closed form). Returns (maximal sets in DQC-R order, stats[, orig_of])."""
    lib = load()
    n = len(out_lists)
    oo, on = _csr_arrays(out_lists)
    io, inn = _csr_arrays(in_lists)
    to, tn = _csr_arrays(two_lists) if two_lists is not None else (None, None)
    ml = 0
    mo = mi = None
    if md_out is not None:
        ml = min(len(md_out), len(md_in))
        mo = (ctypes.c_int32 * ml)(*md_out[:ml])
        mi = (ctypes.c_int32 * ml)(*md_in[:ml])
    sets = []

    def sink(_ctx, ids, length):
        sets.append(tuple(ids[i] for i in range(length)))

    cb = SET_SINK(sink)
    cnt = ctypes.c_int64()
    st = Stats()
    orig = (ctypes.c_int32 * max(n, 1))()
    rc = lib.dqc_search_graph(n, oo, on, io, inn, to, tn, mo, mi, ml, float(out_gamma), float(in_gamma), int(min_size),
                              _b(raw_path), _b(max_path), cb, None, ctypes.byref(config) if config else None,
                              ctypes.byref(cnt), ctypes.byref(st), orig)
    if rc != DQC_OK:
        raise DQCError(f"dqc_search_graph failed: {STATUS.get(rc, rc)}")
    assert cnt.value == len(sets)
    if want_orig:
        return sets, st, [orig[i] for i in range(st.n_condensed)]
    return sets, st


def save_graph(graph_path, bin_path, with_twohop=True):
    """Binary graph cache of a text (or binary) graph file (dqc_save_graph)."""
    rc = load().dqc_save_graph(_b(graph_path), _b(bin_path), int(bool(with_twohop)))
    if rc != DQC_OK:
        raise DQCError(f"dqc_save_graph failed: {STATUS.get(rc, rc)}")


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
