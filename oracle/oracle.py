"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

ORACLE — test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline, never by the
product package.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


class OrcStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in
                "n_input m_input n_condensed m_condensed twohop_len tasks branches raw_cliques max_cliques children bytes".split()] + \
               [(n, ctypes.c_double) for n in "t_load t_twohop t_level0 t_expand t_rmnonmax".split()]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_run.restype = ctypes.c_int64
    lib.orc_run.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.POINTER(OrcStats)]
    lib.orc_graph.restype = ctypes.c_int
    P64 = ctypes.POINTER(ctypes.c_int64)
    P32 = ctypes.POINTER(ctypes.c_int)
    lib.orc_graph.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P32, P64,
                              ctypes.POINTER(P64), ctypes.POINTER(P32), ctypes.POINTER(P64), ctypes.POINTER(P32),
                              ctypes.POINTER(P64), ctypes.POINTER(P32)]
    lib.orc_condense_map.restype = ctypes.c_int
    lib.orc_condense_map.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(P32)]
    lib.orc_free.argtypes = [ctypes.c_void_p]
    lib.orc_set_twohop_restricted.argtypes = [ctypes.c_int]
    lib.orc_twohop_restricted.restype = ctypes.c_int
    lib.orc_twohop_restricted.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, P32,
                                          ctypes.POINTER(P64), ctypes.POINTER(P32),
                                          ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32))]
    lib.orc_mindeg.argtypes = [ctypes.c_double, ctypes.c_int, P32]
    lib.orc_rmnonmax.restype = ctypes.c_int
    lib.orc_rmnonmax.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    lib.orc_open.restype = ctypes.c_void_p
    lib.orc_open.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_close.argtypes = [ctypes.c_void_p]
    lib.orc_root.restype = ctypes.c_int
    lib.orc_root.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int)]
    PU32 = ctypes.POINTER(ctypes.c_uint32)
    lib.orc_expand.restype = ctypes.c_int
    lib.orc_expand.argtypes = [ctypes.c_void_p, PU32, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int,
                               ctypes.POINTER(PU32), ctypes.POINTER(P64), ctypes.POINTER(P32), ctypes.POINTER(P32),
                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    _lib = lib
    return lib


def _b(p):
    return None if p is None else os.fsencode(p)


def set_twohop_restricted(on):
    """Later run / condense_map / Tasks build only the two-hop lists level 0 can read (those
    of the vertices passing its degree test, holding only such vertices); same results."""
    load().orc_set_twohop_restricted(int(bool(on)))


def twohop_restricted(path, mo, mi, threads=1):
    """Restricted two-hop build with degree flags out >= mo, in >= mi:
    (n, offsets, entries, whole list length per vertex)."""
    lib = load()
    n = ctypes.c_int()
    to = ctypes.POINTER(ctypes.c_int64)()
    tn = ctypes.POINTER(ctypes.c_int)()
    l2 = ctypes.POINTER(ctypes.c_uint32)()
    rc = lib.orc_twohop_restricted(_b(path), int(mo), int(mi), threads, ctypes.byref(n), ctypes.byref(to),
                                   ctypes.byref(tn), ctypes.byref(l2))
    if rc != 0:
        raise RuntimeError(f"orc_twohop_restricted failed: {rc}")
    N = n.value
    offs = [to[i] for i in range(N + 1)]
    ent = [tn[i] for i in range(offs[N])]
    full = [l2[i] for i in range(N)]
    for p in (to, tn, l2):
        lib.orc_free(p)
    return N, offs, ent, full


def run(graph_path, out_gamma, in_gamma, min_size, raw_path=None, max_path=None, threads=0,
        host_addfail_check=True, time_cap_s=0.0):
    """Whole CPU path; returns (maximal count or -3 when capped, OrcStats)."""
    st = OrcStats()
    r = load().orc_run(_b(graph_path), out_gamma, in_gamma, min_size, threads, int(bool(host_addfail_check)),
                       _b(raw_path), _b(max_path), float(time_cap_s), ctypes.byref(st))
    return r, st


def graph(path, with_twohop=True, threads=1):
    """Loader (+ two-hop) as python lists of CSR arrays: dict(n, m, out, in, two)."""
    lib = load()
    n = ctypes.c_int()
    m = ctypes.c_int64()
    ptrs = [ctypes.POINTER(ctypes.c_int64)() if i % 2 == 0 else ctypes.POINTER(ctypes.c_int)() for i in range(6)]
    rc = lib.orc_graph(_b(path), int(with_twohop), threads, ctypes.byref(n), ctypes.byref(m),
                       *[ctypes.byref(p) for p in ptrs])
    if rc != 0:
        raise RuntimeError(f"orc_graph failed: {rc}")
    N = n.value
    res = {"n": N, "m": m.value}
    names = ["out", "in", "two"]
    for k, name in enumerate(names):
        off = [ptrs[2 * k][i] for i in range(N + 1)]
        nb = [ptrs[2 * k + 1][i] for i in range(off[-1])] if N else []
        res[name] = [nb[off[i]:off[i + 1]] for i in range(N)]
    for p in ptrs:
        lib.orc_free(p)
    return res


class Tasks:
    """One graph prepared once; expands given tasks one level (device or host semantics)."""

    HOST_L0, HOST_L1, DEVICE = 0, 1, 2

    def __init__(self, path, out_gamma, in_gamma, min_size, threads=0, host_addfail_check=True):
        self.lib = load()
        self.h = self.lib.orc_open(_b(path), out_gamma, in_gamma, min_size, threads or (os.cpu_count() or 1),
                                   int(bool(host_addfail_check)))
        if not self.h:
            raise RuntimeError("orc_open failed")

    def root(self):
        ncov = ctypes.c_int()
        k = self.lib.orc_root(self.h, None, ctypes.byref(ncov))
        arr = (ctypes.c_uint32 * max(6 * k, 1))()
        self.lib.orc_root(self.h, arr, ctypes.byref(ncov))
        return [tuple(arr[6 * i:6 * i + 6]) for i in range(k)], ncov.value

    def expand(self, records, nm, ncov=0, mode=2, max_steps=0, first_step=0):
        """Returns (children as (records, nm), cliques as sorted id tuples, branches);
        max_steps > 0 runs only the loop steps [first_step, first_step + max_steps) after the
        covered prefix (the steps before first_step remove and test lookahead; their branches
        run only until one yields a child, and give no output)."""
        flat = (ctypes.c_uint32 * max(6 * len(records), 1))(*[x for r in records for x in r])
        kids = ctypes.POINTER(ctypes.c_uint32)()
        meta = ctypes.POINTER(ctypes.c_int64)()
        knm = ctypes.POINTER(ctypes.c_int)()
        clq = ctypes.POINTER(ctypes.c_int)()
        clen = ctypes.c_int64()
        br = ctypes.c_int64()
        n = self.lib.orc_expand(self.h, flat, len(records), nm, ncov, mode, int(max_steps), int(first_step),
                                ctypes.byref(kids), ctypes.byref(meta),
                                ctypes.byref(knm), ctypes.byref(clq), ctypes.byref(clen), ctypes.byref(br))
        out = []
        for c in range(n):
            off, size = meta[2 * c], meta[2 * c + 1]
            out.append(([tuple(kids[6 * (off + j):6 * (off + j) + 6]) for j in range(size)], knm[c]))
        cl, p = [], 0
        while p < clen.value:
            L = clq[p]
            cl.append(tuple(sorted(clq[p + 1:p + 1 + L])))
            p += 1 + L
        for x in (kids, meta, knm, clq):
            self.lib.orc_free(x)
        return out, sorted(cl), br.value

    def expand_digest(self, records, nm, ncov=0, mode=2, max_steps=0, first_step=0):
        """As expand, with each child as (SHA-1 of its records' bytes, nm): records are 6
        uint32 per vertex, the layout the library's task stack uses."""
        import hashlib
        flat = (ctypes.c_uint32 * max(6 * len(records), 1))(*[x for r in records for x in r])
        kids = ctypes.POINTER(ctypes.c_uint32)()
        meta = ctypes.POINTER(ctypes.c_int64)()
        knm = ctypes.POINTER(ctypes.c_int)()
        clq = ctypes.POINTER(ctypes.c_int)()
        clen = ctypes.c_int64()
        br = ctypes.c_int64()
        n = self.lib.orc_expand(self.h, flat, len(records), nm, ncov, mode, int(max_steps), int(first_step),
                                ctypes.byref(kids), ctypes.byref(meta), ctypes.byref(knm), ctypes.byref(clq),
                                ctypes.byref(clen), ctypes.byref(br))
        total = sum(meta[2 * c + 1] for c in range(n))
        buf = ctypes.string_at(kids, 24 * total) if total else b""
        out = []
        for c in range(n):
            off, size = meta[2 * c], meta[2 * c + 1]
            out.append((hashlib.sha1(buf[24 * off:24 * (off + size)]).hexdigest(), knm[c]))
        cl, p = [], 0
        while p < clen.value:
            L = clq[p]
            cl.append(tuple(sorted(clq[p + 1:p + 1 + L])))
            p += 1 + L
        for x in (kids, meta, knm, clq):
            self.lib.orc_free(x)
        return out, sorted(cl), br.value

    def close(self):
        if self.h:
            self.lib.orc_close(self.h)
            self.h = None


def condense_map(path, out_gamma, in_gamma, min_size):
    lib = load()
    p = ctypes.POINTER(ctypes.c_int)()
    k = lib.orc_condense_map(_b(path), out_gamma, in_gamma, min_size, 1, ctypes.byref(p))
    if k < 0:
        raise RuntimeError(f"orc_condense_map failed: {k}")
    res = [p[i] for i in range(k)]
    lib.orc_free(p)
    return res


def mindeg(gamma, length):
    arr = (ctypes.c_int * length)()
    load().orc_mindeg(gamma, length, arr)
    return list(arr)


def rmnonmax(in_path, out_path):
    return load().orc_rmnonmax(_b(in_path), _b(out_path))
