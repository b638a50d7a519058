"""Host-side driver of the engine (libhjoin.so): the drop-in execute() and resident tables.

    ctx = Context()                        # Contest::build_context (src/execute.cpp:610-612)
    result = ctx.execute(plan)             # Contest::execute (src/execute.cpp:598-608)
    ctx.close()                            # Contest::destroy_context (src/execute.cpp:614)

Errors raise EngineError with the engine's message (e.g. "Join key isn't of type int32_t",
the reference's message at src/execute.cpp:81). There is no CPU fallback: without the built
library every call raises abi.EngineMissing.
"""
import ctypes as C

import numpy as np

from . import abi
from .plan import ColumnarTable, result_from_c, table_to_c


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


class DeviceTable:
    """A ColumnarTable resident in HBM (hj_table_upload)."""

    def __init__(self, ctx, handle, num_rows):
        self.ctx = ctx
        self.handle = handle
        self.num_rows = num_rows

    def free(self):
        if self.handle:
            self.ctx.lib.hj_table_free(self.ctx.handle, self.handle)
            self.handle = None


class DeviceResult:
    """A join result in HBM (hj_execute_device); valid until free()."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle

    @property
    def num_rows(self):
        return int(self.ctx.lib.hj_dresult_num_rows(self.handle))

    def to_host(self):
        res = abi.hj_result()
        self.ctx._check(self.ctx.lib.hj_dresult_to_host(self.ctx.handle, self.handle, C.byref(res)))
        return result_from_c(res, self.ctx.lib.hj_result_free)

    def free(self):
        if self.handle:
            self.ctx.lib.hj_dresult_free(self.ctx.handle, self.handle)
            self.handle = None


class Context:
    def __init__(self, arena_bytes=0):
        self.lib = abi.load()
        h = C.c_void_p()
        rc = self.lib.hj_context_create(arena_bytes, C.byref(h))
        if rc:
            raise EngineError(rc, self.lib.hj_global_error().decode())
        self.handle = h

    def _check(self, rc):
        if rc:
            raise EngineError(rc, self.lib.hj_last_error(self.handle).decode())

    @property
    def stream(self):
        return self.lib.hj_context_stream(self.handle)

    def execute(self, plan):
        p, keep = plan.to_c()
        res = abi.hj_result()
        self._check(self.lib.hj_execute(self.handle, C.byref(p), C.byref(res)))
        return result_from_c(res, self.lib.hj_result_free)

    def upload(self, table: ColumnarTable):
        keep = []
        ht = table_to_c(table, keep)
        h = C.c_void_p()
        self._check(self.lib.hj_table_upload(self.handle, C.byref(ht), C.byref(h)))
        return DeviceTable(self, h, table.num_rows)

    def upload_cache(self, path):
        """Table::from_cache (src/build_table.cpp:147-203) straight into HBM
        (hj_table_upload_cache)"""
        h = C.c_void_p()
        self._check(self.lib.hj_table_upload_cache(self.handle, str(path).encode(), C.byref(h)))
        return DeviceTable(self, h, int(self.lib.hj_table_num_rows(h)))

    def execute_device(self, plan, dtables):
        # the C plan is rebuilt only when the plan changed (node count, root)
        sig = (len(plan.nodes), plan.root)
        cached = getattr(plan, "_c_device_plan", None)
        if cached is None or cached[0] != sig:
            cached = (sig,) + plan.to_c(with_inputs=False)
            plan._c_device_plan = cached
        _, p, keep = cached
        arr = (C.c_void_p * max(1, len(dtables)))(*[t.handle for t in dtables])
        h = C.c_void_p()
        self._check(self.lib.hj_execute_device(self.handle, C.byref(p), arr, C.byref(h)))
        return DeviceResult(self, h)

    def set_option(self, name, value):
        """engine options: "pipeline" (1 = fused probe pipeline for left-deep chains, 0 = off),
        "partitioned_build" (1 = pipeline tables built segment by segment in LDS, 0 = by
        compare-and-swap inserts into cleared tables), "gpu_pages" (1 = result pages written on
        the device, 0 = flat columns copied out and paged on the host; same bytes)"""
        self._check(self.lib.hj_set_option(self.handle, name.encode(), int(value)))

    def cache_stats(self):
        """directory cache: {"hits", "misses", "entries", "bytes"}"""
        v = [C.c_uint64() for _ in range(4)]
        self._check(self.lib.hj_dir_cache_stats(self.handle, *[C.byref(x) for x in v]))
        return dict(zip(("hits", "misses", "entries", "bytes"), (x.value for x in v)))

    def profile(self, prefix):
        """time engine kernels whose name starts with prefix ("" = all, None = off)"""
        self._check(self.lib.hj_profile_enable(
            self.handle, None if prefix is None else prefix.encode()))

    def profile_report(self):
        """{kernel name: (launches, total ms)} since the last report"""
        buf = C.create_string_buffer(1 << 16)
        self._check(self.lib.hj_profile_report(self.handle, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out

    def synchronize(self):
        self._check(self.lib.hj_synchronize(self.handle))

    def close(self):
        if self.handle:
            self.lib.hj_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Directory:
    """The GPU unchained directory of one key array (the design of include/unchained.h)."""

    def __init__(self, ctx, keys):
        self.ctx = ctx
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        h = C.c_void_p()
        ctx._check(ctx.lib.hj_directory_build(ctx.handle, k.ctypes.data if len(k) else None,
                                              len(k), C.byref(h)))
        self.handle = h

    def probe(self, keys):
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        cnt = np.zeros(max(1, len(k)), dtype=np.uint64)
        first = np.zeros(max(1, len(k)), dtype=np.uint64)
        self.ctx._check(self.ctx.lib.hj_directory_probe(
            self.ctx.handle, self.handle, k.ctypes.data if len(k) else None, len(k),
            cnt.ctypes.data, first.ctypes.data))
        return cnt[:len(k)], first[:len(k)]

    def probe_noalloc(self, keys):
        """probe() through hj_directory_probe_noalloc: no device allocation, at most
        abi.NOALLOC_MAX_PROBE_KEYS keys per call, calls on one device must not overlap"""
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        cnt = np.zeros(max(1, len(k)), dtype=np.uint64)
        first = np.zeros(max(1, len(k)), dtype=np.uint64)
        self.ctx._check(self.ctx.lib.hj_directory_probe_noalloc(
            self.ctx.handle, self.handle, k.ctypes.data if len(k) else None, len(k),
            cnt.ctypes.data, first.ctypes.data))
        return cnt[:len(k)], first[:len(k)]

    def probe_pairs(self, keys):
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        ptr = C.c_void_p()
        n = C.c_uint64()
        self.ctx._check(self.ctx.lib.hj_directory_probe_pairs(
            self.ctx.handle, self.handle, k.ctypes.data if len(k) else None, len(k),
            C.byref(ptr), C.byref(n)))
        out = np.zeros((n.value, 2), dtype=np.uint64)
        if n.value:
            C.memmove(out.ctypes.data, ptr, n.value * 16)
        self.ctx.lib.hj_free(ptr)
        return out

    def info(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx.lib.hj_directory_info(self.handle, C.byref(a), C.byref(b), C.byref(c))
        return {"num_slots": a.value, "num_tuples": b.value, "max_slot_len": c.value}

    def free(self):
        if self.handle:
            self.ctx.lib.hj_directory_free(self.ctx.handle, self.handle)
            self.handle = None


class Server:
    """A pool of contexts serving queries over resident tables (hj_server_*): submit() queues a
    query and returns a ticket, wait(ticket) returns (ColumnarTable, queued_ms, run_ms). Tables
    are uploaded with any Context and shared by every query."""

    def __init__(self, slots=4, arena_bytes=0, budget_bytes=None, dir_cache_bytes=None):
        self.lib = abi.load()
        h = C.c_void_p()
        rc = self.lib.hj_server_create(int(slots), int(arena_bytes), C.byref(h))
        if rc:
            raise EngineError(rc, self.lib.hj_global_error().decode())
        self.handle = h
        self._pending = {}
        if budget_bytes is not None:
            self.set_option("budget_bytes", budget_bytes)
        if dir_cache_bytes is not None:
            self.set_option("dir_cache_bytes", dir_cache_bytes)

    @property
    def slots(self):
        return int(self.lib.hj_server_slots(self.handle))

    def set_option(self, name, value):
        """"budget_bytes" (reservations of the running queries together), "dir_cache_bytes"
        (the slots' shared directory cache); set before the first submit"""
        rc = self.lib.hj_server_set_option(self.handle, name.encode(), int(value))
        if rc:
            raise EngineError(rc, f"server option {name}")

    def cache_stats(self):
        v = [C.c_uint64() for _ in range(4)]
        rc = self.lib.hj_server_cache_stats(self.handle, *[C.byref(x) for x in v])
        if rc:
            raise EngineError(rc, "server cache stats")
        return dict(zip(("hits", "misses", "entries", "bytes"), (x.value for x in v)))

    def submit(self, plan, dtables):
        sig = (len(plan.nodes), plan.root)
        cached = getattr(plan, "_c_device_plan", None)
        if cached is None or cached[0] != sig:
            cached = (sig,) + plan.to_c(with_inputs=False)
            plan._c_device_plan = cached
        _, p, keep = cached
        arr = (C.c_void_p * max(1, len(dtables)))(*[t.handle for t in dtables])
        t = C.c_uint64()
        rc = self.lib.hj_server_submit(self.handle, C.byref(p), arr, C.byref(t))
        if rc:
            raise EngineError(rc, "query refused" if rc == -6 else "invalid query")
        self._pending[t.value] = (p, keep, arr)  # alive until the query is done
        return t.value

    def wait(self, ticket, keep=True):
        """(ColumnarTable, queued_ms, run_ms); with keep=False the result is released at once
        and its page bytes come back in place of the table"""
        res = abi.hj_result()
        q, r = C.c_uint64(), C.c_uint64()
        rc = self.lib.hj_server_wait(self.handle, ticket, C.byref(res), C.byref(q), C.byref(r))
        self._pending.pop(ticket, None)
        if rc:
            raise EngineError(rc, self.lib.hj_server_last_error(self.handle).decode())
        if not keep:
            nbytes = sum(res.columns[j].num_pages for j in range(res.num_columns)) * 8192
            self.lib.hj_result_free(C.byref(res))
            return nbytes, q.value / 1e6, r.value / 1e6
        return result_from_c(res, self.lib.hj_result_free), q.value / 1e6, r.value / 1e6

    def close(self):
        if self.handle:
            self.lib.hj_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Many independent queries over resident tables in four shared launches (hj_batch_*).

    queries: list of (plan, dtables). params: None (the engine's defaults), a dict of
    hj_batch_params fields, or "tuned" (the committed autotuner table, tuning/batch_mi355x.json).
    run() executes every query; result(i) returns query i's rows as a ColumnarTable;
    info(i) -> (ran in the batch's launches, rows)."""

    def __init__(self, ctx, queries, params=None):
        self.ctx = ctx
        self.lib = ctx.lib
        if params == "tuned":
            from .autotune import tuned_params
            params = tuned_params(queries)
        self.params = dict(params or {})
        self._keep = []
        n = len(queries)
        plans = (C.POINTER(abi.hj_plan) * max(1, n))()
        tabs = (C.POINTER(C.c_void_p) * max(1, n))()
        for i, (plan, dts) in enumerate(queries):
            p, keep = plan.to_c(with_inputs=False)
            arr = (C.c_void_p * max(1, len(dts)))(*[t.handle for t in dts])
            self._keep.append((p, keep, arr))
            plans[i] = C.pointer(p)
            tabs[i] = C.cast(arr, C.POINTER(C.c_void_p))
        self._keep.append((plans, tabs))
        prm = None
        if params is not None:
            prm = abi.hj_batch_params()
            prm.use_graph = 1
            for k, v in self.params.items():
                setattr(prm, k, int(v))
        h = C.c_void_p()
        ctx._check(self.lib.hj_batch_create(ctx.handle, n, plans, tabs,
                                            C.byref(prm) if prm is not None else None,
                                            C.byref(h)))
        self.handle = h
        self.n = n

    def run(self):
        self.ctx._check(self.lib.hj_batch_run(self.ctx.handle, self.handle))

    def info(self, i):
        b, r = C.c_int32(), C.c_uint64()
        self.ctx._check(self.lib.hj_batch_info(self.handle, i, C.byref(b), C.byref(r)))
        return bool(b.value), int(r.value)

    def device_bytes(self):
        return int(self.lib.hj_batch_device_bytes(self.handle))

    def result(self, i):
        res = abi.hj_result()
        self.ctx._check(self.lib.hj_batch_result_to_host(self.ctx.handle, self.handle, i,
                                                         C.byref(res)))
        return result_from_c(res, self.lib.hj_result_free)

    def free(self):
        if self.handle:
            self.lib.hj_batch_free(self.ctx.handle, self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()
