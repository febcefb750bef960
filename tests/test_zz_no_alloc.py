"""The allocation-free entry point (dqc_expand_noalloc) against the existing one-batch
expansion (dqc_push_task + dqc_step): on two calls in a row (the level-1 tasks of a
golden case, then a smaller batch built from the first call's own children), the
children (records, members, mode), the cliques, the tasks expanded and the branches must
be identical. The caller's buffers are allocated by the test through the HIP runtime
the library uses; the entry point itself allocates nothing."""
import ctypes
import os

import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2


class Hip:
    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), max(int(nbytes), 64)) == 0
        self.bufs.append(p.value)
        return p.value

    def put(self, ctype, values):
        arr = (ctype * max(len(values), 1))(*values)
        d = self.alloc(ctypes.sizeof(arr))
        assert self.lib.hipMemcpy(d, arr, ctypes.sizeof(arr), HIP_H2D) == 0
        return d

    def get(self, ctype, d, n):
        arr = (ctype * max(n, 1))()
        if n:
            assert self.lib.hipMemcpy(arr, d, ctypes.sizeof(ctype) * n, HIP_D2H) == 0
        return list(arr)[:n]

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


class TaskDesc(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("size", ctypes.c_int32), ("members", ctypes.c_int32),
                ("covered", ctypes.c_int32), ("mode", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("id", ctypes.c_uint32), ("parent", ctypes.c_uint32),
                ("depth", ctypes.c_uint16), ("graph", ctypes.c_uint16), ("dmin_out", ctypes.c_uint16),
                ("dmax_out", ctypes.c_uint16), ("dmin_in", ctypes.c_uint16), ("dmax_in", ctypes.c_uint16),
                ("first_step", ctypes.c_int32)]


class DeviceGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("out_off", ctypes.c_void_p), ("out_nbr", ctypes.c_void_p),
                ("in_off", ctypes.c_void_p), ("in_nbr", ctypes.c_void_p), ("two_off", ctypes.c_void_p),
                ("two_nbr", ctypes.c_void_p), ("md_out", ctypes.c_void_p), ("md_in", ctypes.c_void_p),
                ("ub_out", ctypes.c_void_p), ("ub_in", ctypes.c_void_p), ("table_len", ctypes.c_int32),
                ("min_size", ctypes.c_int32), ("host_addfail_check", ctypes.c_int32)]


def condensed_graph(dqc, path, go, gi, mn):
    """The condensed graph the library builds (lists restricted to the level-0 survivors,
    relabelled to condensed ids), from the loader's lists and the level-0 map."""
    g = dqc.load_graph(path)
    _, _, orig = dqc.level0(path, go, gi, mn)
    rank = {v: i for i, v in enumerate(orig)}
    lists = {}
    for name in ("out", "in", "two"):
        lists[name] = [sorted(rank[x] for x in g[name][v] if x in rank) for v in orig]
    return lists


def flat(lists):
    off = [0]
    for l in lists:
        off.append(off[-1] + len(l))
    return off, [x for l in lists for x in l]


def engine_expand(dqc, path, go, gi, mn, tasks):
    """The existing path: the tasks pushed, one dqc_step launch, plain FIFO."""
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.schedule = dqc.SCHED_FIFO
    with dqc.Engine(cfg) as eng:
        eng.prepare(path, go, gi, mn)
        eng.reset()
        for recs, nm, mode in tasks:
            eng.push_task(recs, nm, 0, mode)
        fin = eng.step(1)
        st = eng.stats()
        kids = eng.tasks()
        cl = sorted(tuple(sorted(c)) for c in eng.cliques())
        return sorted((tuple(r), nm, mode) for r, nm, mode in kids), cl, st.tasks_expanded, st.branches, fin


def noalloc_expand(dqc, hip, dg, tasks, out_cap_t=1 << 18, out_cap_v=1 << 24):
    lib = dqc.load()
    recs = []
    descs = (TaskDesc * max(len(tasks), 1))()
    maxtot = 0
    for i, (r, nm, mode) in enumerate(tasks):
        d = descs[i]
        d.offset = len(recs)
        d.size, d.members, d.covered, d.mode = len(r), nm, 0, mode
        d.steps, d.flags = len(r) - nm, 0
        d.id, d.parent, d.depth = 0, 0xFFFFFFFF, 1
        recs.extend(r)
        maxtot = max(maxtot, len(r))
    d_desc = hip.alloc(ctypes.sizeof(descs))
    assert hip.lib.hipMemcpy(d_desc, descs, ctypes.sizeof(descs), HIP_H2D) == 0
    d_rec = hip.put(ctypes.c_uint32, [x for r in recs for x in r])
    o_desc = hip.alloc(56 * out_cap_t)
    o_rec = hip.alloc(24 * out_cap_v)
    c_ids = hip.alloc(4 * (1 << 22))
    c_meta = hip.alloc(8 * (1 << 18))
    ctr = hip.alloc(8 * 9)
    lib.dqc_expand_noalloc.restype = ctypes.c_int32
    lib.dqc_expand_noalloc.argtypes = [ctypes.POINTER(DeviceGraph), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                       ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    rc = lib.dqc_expand_noalloc(ctypes.byref(dg), d_desc, d_rec, len(tasks), len(recs), maxtot, o_desc, o_rec,
                                out_cap_t, out_cap_v, c_ids, c_meta, 1 << 18, 1 << 22, ctr, None)
    assert rc == 0, rc
    hip.sync()
    c = hip.get(ctypes.c_uint64, ctr, 9)
    nkid, nkrec, npiece, _, ntask, nbr, ncl, ncli, err = c
    assert err == 0 and npiece == 0, c
    kd = (TaskDesc * max(nkid, 1))()
    if nkid:
        assert hip.lib.hipMemcpy(kd, o_desc, 56 * nkid, HIP_D2H) == 0
    kr = hip.get(ctypes.c_uint32, o_rec, 6 * nkrec)
    kids = []
    for i in range(nkid):
        d = kd[i]
        r = [tuple(kr[6 * (d.offset + j):6 * (d.offset + j) + 6]) for j in range(d.size)]
        kids.append((tuple(r), d.members, d.mode))
    meta = hip.get(ctypes.c_uint64, c_meta, ncl)
    ids = hip.get(ctypes.c_int32, c_ids, ncli)
    cl = sorted(tuple(sorted(ids[(m >> 24):(m >> 24) + (m & 0xFFFFFF)])) for m in meta)
    return sorted(kids), cl, ntask, nbr


@pytest.mark.parametrize("name", ["workload3_g0", "workload3_g2", "rand069"])
def test_noalloc_matches_step(dqc, manifest, name):
    c = next(x for x in manifest["cases"] if x["name"] == name)
    path = os.path.join(GOLDEN, c["graph"])
    go, gi, mn = c["out_gamma"], c["in_gamma"], c["min_size"]
    hip = Hip()
    try:
        lists = condensed_graph(dqc, path, go, gi, mn)
        n = len(lists["out"])
        tl = 2 * (max(n, mn) + 2) + 2
        md_o, md_i = dqc.min_degrees(go, tl - 1), dqc.min_degrees(gi, tl - 1)
        ub_o = [int(d / go) for d in range(tl)]
        ub_i = [int(d / gi) for d in range(tl)]
        dg = DeviceGraph()
        dg.n = n
        for name_, key in (("out", "out"), ("in", "in"), ("two", "two")):
            off, nb = flat(lists[key])
            setattr(dg, name_ + "_off", hip.put(ctypes.c_uint64, off))
            setattr(dg, name_ + "_nbr", hip.put(ctypes.c_int32, nb))
        dg.md_out, dg.md_in = hip.put(ctypes.c_int32, md_o), hip.put(ctypes.c_int32, md_i)
        dg.ub_out, dg.ub_in = hip.put(ctypes.c_int32, ub_o), hip.put(ctypes.c_int32, ub_i)
        dg.table_len, dg.min_size, dg.host_addfail_check = tl, mn, 1
        # the level-1 tasks: the level-0 task expanded by the library (branch-parallel)
        cfg = dqc.Config()
        cfg.host_addfail_check = 1
        with dqc.Engine(cfg) as eng:
            eng.prepare(path, go, gi, mn)
            eng.step(1)
            level1 = eng.tasks()
        assert level1
        first = noalloc_expand(dqc, hip, dg, level1)
        ref = engine_expand(dqc, path, go, gi, mn, level1)
        print(f"{name}: call 1: {len(level1)} tasks -> {len(first[0])} children, {first[3]} branches", flush=True)
        assert first[0] == ref[0] and first[1] == ref[1] and first[2] == ref[2] == len(level1) and first[3] == ref[3]
        # second call: a smaller batch, from the first call's own children
        take = max(1, min(len(level1) - 1, len(first[0]) // 4, 64))
        second_in = [(list(r), nm, mode) for r, nm, mode in first[0][:take]]
        assert 0 < len(second_in) < len(level1) or len(level1) == 1
        second = noalloc_expand(dqc, hip, dg, second_in)
        ref2 = engine_expand(dqc, path, go, gi, mn, second_in)
        print(f"{name}: call 2: {len(second_in)} tasks -> {len(second[0])} children, {second[3]} branches", flush=True)
        assert second[0] == ref2[0] and second[1] == ref2[1] and second[2] == ref2[2] and second[3] == ref2[3]
    finally:
        hip.free()


def test_noalloc_rejects_large_graph(dqc):
    lib = dqc.load()
    dg = DeviceGraph()
    dg.n = 4097
    dg.out_off = dg.in_off = dg.two_off = 1
    dg.md_out = dg.md_in = dg.ub_out = dg.ub_in = 1
    dg.table_len, dg.min_size = 2 * (4097 + 2) + 2, 3
    lib.dqc_expand_noalloc.restype = ctypes.c_int32
    ctr = ctypes.c_void_p(1)
    rc = lib.dqc_expand_noalloc(ctypes.byref(dg), None, None, ctypes.c_int64(0), ctypes.c_int64(0), 0, None, None,
                                ctypes.c_int64(0), ctypes.c_int64(0), None, None, ctypes.c_int64(0), ctypes.c_int64(0),
                                ctr, None)
    assert rc == 5  # DQC_E_UNSUPPORTED: beyond the scratch array's limit
