"""Python mirror of the engine's host surfaces, over the C-ABI.

Names follow the reference's C++ API: ``build_hip_plan`` / ``generate_from_plan``
(buildMetalPlan / generateFromPlan, codegen/planning/metal_plan_builder.h:73-78),
``RuntimeCompiler`` (codegen/execution/runtime_compiler.h:9-41),
``HipGenericExecutor`` (MetalGenericExecutor, codegen/execution/metal_generic_executor.h:46-164)
and ``GenericResult`` (codegen/execution/metal_result_collector.h:19-60).
Errors surface as ``GdbError`` (``GdbOverflow`` for hash-table overflow), the
C-ABI's status codes, in place of the reference's C++ exceptions.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

GdbError = N.GdbError
GdbOverflow = N.GdbOverflow


def _hip():
    return N.load()[0]


def _eng():
    return N.load()[1]


def device_count() -> int:
    n = C.c_int(0)
    N.hip_check(_hip().gdb_device_count(C.byref(n)))
    return n.value


class Context:
    """One GPU: device context, default stream, allocations."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        N.hip_check(_hip().gdb_init(device, C.byref(self._h)))
        info = N.DeviceInfo()
        N.hip_check(_hip().gdb_device_info(self._h, C.byref(info)))
        self.info = info
        self.stream = _hip().gdb_default_stream(self._h)
        self._allocs: set[int] = set()

    @property
    def handle(self):
        return self._h

    @property
    def cu_count(self) -> int:
        return self.info.cu_count

    @property
    def arch(self) -> str:
        return self.info.arch.decode()

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        N.hip_check(_hip().gdb_alloc(self._h, max(int(nbytes), 16), C.byref(p)))
        self._allocs.add(p.value)
        return p.value

    def free(self, ptr: int) -> None:
        if ptr in self._allocs:
            N.hip_check(_hip().gdb_free(self._h, C.c_void_p(ptr)))
            self._allocs.discard(ptr)

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        ptr = self.alloc(arr.nbytes)
        N.hip_check(_hip().gdb_memcpy_h2d_async(C.c_void_p(ptr), arr.ctypes.data_as(C.c_void_p), arr.nbytes,
                                                 self.stream))
        self.sync()
        return ptr

    def download(self, ptr: int, nbytes: int, dtype=np.uint8) -> np.ndarray:
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        N.hip_check(_hip().gdb_memcpy_d2h_async(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), nbytes, self.stream))
        self.sync()
        return out

    def sync(self) -> None:
        N.hip_check(_hip().gdb_stream_sync(self.stream))

    def mem_info(self) -> tuple[int, int]:
        f, t = C.c_size_t(), C.c_size_t()
        N.hip_check(_hip().gdb_mem_info(self._h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def cache_stats(self) -> tuple[int, int, float]:
        h, m, ms = C.c_uint64(), C.c_uint64(), C.c_double()
        N.hip_check(_hip().gdb_cache_stats(self._h, C.byref(h), C.byref(m), C.byref(ms)))
        return h.value, m.value, ms.value

    def set_cache_dir(self, path: str) -> None:
        N.hip_check(_hip().gdb_cache_set_dir(self._h, path.encode()))

    def close(self) -> None:
        for p in list(self._allocs):
            self.free(p)
        if self._h:
            _hip().gdb_shutdown(self._h)
            self._h = C.c_void_p()


@dataclass
class DeviceTable:
    """A table resident in HBM: one device buffer per column
    (QueryColumns / loadQueryColumns, src/infra.h:930-1062)."""

    name: str
    rows: int
    columns: dict = field(default_factory=dict)  # column -> device pointer
    nbytes: int = 0
    key_max: dict = field(default_factory=dict)  # column -> largest value (device max-key scan, cached)

    @classmethod
    def from_host(cls, ctx: Context, name: str, cols: dict) -> "DeviceTable":
        rows = None
        t = cls(name, 0)
        for col, arr in cols.items():
            arr = np.ascontiguousarray(arr)
            r = arr.shape[0]
            if rows is None:
                rows = r
            elif rows != r:
                raise ValueError(f"table {name}: column {col} has {r} rows, expected {rows}")
            t.columns[col] = ctx.upload(arr)
            t.nbytes += arr.nbytes
        t.rows = rows or 0
        return t

    def free(self, ctx: Context) -> None:
        for p in self.columns.values():
            ctx.free(p)
        self.columns.clear()


class HipQueryPlan:
    """A built plan plus its generated HIP source (buildMetalPlan + generateFromPlan)."""

    def __init__(self, name: str, options: str = "", sql: str | None = None):
        self.name = name
        self.options = options
        self._h = C.c_void_p()
        if sql is None:
            N.eng_check(_eng().gdb_plan_build(name.encode(), options.encode(), C.byref(self._h)))
        else:
            N.eng_check(_eng().gdb_plan_build_sql(sql.encode(), options.encode(), C.byref(self._h)))
        self.compiled_for = None
        self.compile_ms = 0.0
        self.cache_hit = False
        self.plan_hash = 0

    @property
    def handle(self):
        return self._h

    def describe(self) -> str:
        return N.two_call_string(_eng().gdb_plan_describe, self._h)

    def source(self) -> str:
        return N.two_call_string(_eng().gdb_plan_source, self._h)

    def sql_layout(self) -> dict:
        """How to finish a SQL plan's rows (gdb_plan_sql_layout)."""
        import json

        return json.loads(N.two_call_string(_eng().gdb_plan_sql_layout, self._h))

    def phases(self) -> list[str]:
        n = C.c_int()
        N.eng_check(_eng().gdb_plan_phase_count(self._h, C.byref(n)))
        out = []
        for i in range(n.value):
            buf = C.create_string_buffer(256)
            N.eng_check(_eng().gdb_plan_phase_name(self._h, i, buf, 256))
            out.append(buf.value.decode())
        return out

    def set_launch(self, phase: str, block: int, items: int) -> None:
        N.eng_check(_eng().gdb_plan_set_launch(self._h, phase.encode(), block, items))

    def get_launch(self, phase: str) -> tuple[int, int]:
        b, i = C.c_int(), C.c_int()
        N.eng_check(_eng().gdb_plan_get_launch(self._h, phase.encode(), C.byref(b), C.byref(i)))
        return b.value, i.value

    def compile(self, ctx: Context) -> "HipQueryPlan":
        ms, hit, h = C.c_double(), C.c_int(), C.c_uint64()
        N.eng_check(_eng().gdb_plan_compile(ctx.handle, self._h, C.byref(ms), C.byref(hit), C.byref(h)))
        self.compiled_for = ctx
        self.compile_ms, self.cache_hit, self.plan_hash = ms.value, bool(hit.value), h.value
        return self

    @staticmethod
    def batch_kernel_source(plans: list, items: list[int]) -> str:
        """Source of the persistent batched kernel for pipelines plans[i] at
        items[i]; needs no device."""
        n = len(plans)
        hs = (C.c_void_p * n)(*[p._h.value for p in plans])
        its = (C.c_int * n)(*items)
        return N.two_call_string(_eng().gdb_batch_kernel_source, hs, its, n)

    @property
    def chunkable(self) -> bool:
        v = C.c_int()
        N.eng_check(_eng().gdb_plan_chunkable(self._h, C.byref(v)))
        return bool(v.value)

    def precompile(self, cache_dir: str, arch: str = "gfx950") -> bool:
        wc, h = C.c_int(), C.c_uint64()
        N.eng_check(_eng().gdb_plan_precompile(self._h, cache_dir.encode(), arch.encode(), C.byref(wc), C.byref(h)))
        return bool(wc.value)

    def __del__(self):
        try:
            if self._h:
                _eng().gdb_plan_free(self._h)
        except Exception:
            pass


def build_hip_plan(name: str, options: str = "") -> HipQueryPlan:
    return HipQueryPlan(name, options)


def build_sql_plan(sql: str, options: str = "") -> HipQueryPlan:
    """Plan a single-table SELECT from its text (gdb_plan_build_sql)."""
    return HipQueryPlan("SQL", options, sql=sql)


@dataclass
class Timing:
    total_ms: float
    alloc_ms: float
    phases: list  # (name, ms, grid, block, items)
    status: list  # per phase [overflow, dup, inserted, spills]


@dataclass
class GenericResult:
    columns: dict  # name -> numpy array (int64; int128 as Python-int object arrays; strings)
    canonical: str
    formats: dict = field(default_factory=dict)  # name -> (scale digits, is_date) of integer columns

    def rows(self) -> int:
        return 0 if not self.columns else len(next(iter(self.columns.values())))


class HostTable:
    """A table kept in host memory for streamed (out-of-core) execution.
    pinned=True copies the columns into page-locked memory so the chunk
    copies run at link speed and overlap the kernels."""

    def __init__(self, name: str, cols: dict, pinned: bool = True):
        self.name = name
        self.rows = int(next(iter(cols.values())).shape[0]) if cols else 0
        self._pinned = []
        self.columns = {}
        hip, _ = N.load()
        for k, a in cols.items():
            a = np.ascontiguousarray(a)
            if pinned and a.nbytes:
                p = C.c_void_p()
                N.hip_check(hip.gdb_host_alloc_pinned(a.nbytes, C.byref(p)))
                self._pinned.append(p)
                buf = (C.c_uint8 * a.nbytes).from_address(p.value)
                view = np.frombuffer(buf, dtype=a.dtype).reshape(a.shape)
                view[...] = a
                self.columns[k] = view
            else:
                self.columns[k] = a

    def nbytes(self) -> int:
        return sum(a.nbytes for a in self.columns.values())

    def close(self) -> None:
        hip, _ = N.load()
        self.columns = {}
        for p in self._pinned:
            hip.gdb_host_free_pinned(p)
        self._pinned = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipGenericExecutor:
    def __init__(self, ctx: Context, stream=None):
        self.ctx = ctx
        self._h = C.c_void_p()
        N.eng_check(_eng().gdb_executor_create(ctx.handle, stream, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def register_table(self, t: DeviceTable) -> None:
        N.eng_check(_eng().gdb_executor_register_table(self._h, t.name.encode(), t.rows))
        for col, ptr in t.columns.items():
            N.eng_check(_eng().gdb_executor_register_column(self._h, t.name.encode(), col.encode(), C.c_void_p(ptr)))

    def register_buffer(self, name: str, ptr: int, nbytes: int) -> None:
        N.eng_check(_eng().gdb_executor_register_buffer(self._h, name.encode(), C.c_void_p(ptr), nbytes))

    def register_symbol(self, name: str, value: int) -> None:
        N.eng_check(_eng().gdb_executor_register_symbol(self._h, name.encode(), int(value)))

    def get_symbol(self, name: str):
        v, f = C.c_uint64(), C.c_int()
        N.eng_check(_eng().gdb_executor_get_symbol(self._h, name.encode(), C.byref(v), C.byref(f)))
        return v.value if f.value else None

    def register_scalar(self, name: str, value: int) -> None:
        N.eng_check(_eng().gdb_executor_register_scalar(self._h, name.encode(), int(value)))

    def execute_streamed(self, plan: "HipQueryPlan", table: "HostTable", chunk_rows: int) -> N.ChunkStats:
        """Out-of-core run: `table` stays in host memory and is streamed into
        HBM in chunks of chunk_rows rows (phases scanning it run per chunk)."""
        names = list(table.columns)
        n = len(names)
        cn = (C.c_char_p * n)(*[x.encode() for x in names])
        ptrs = (C.c_void_p * n)(*[table.columns[x].ctypes.data for x in names])
        eb = (C.c_int * n)(*[table.columns[x].itemsize * (table.columns[x].shape[1] if table.columns[x].ndim == 2 else 1)
                             for x in names])
        st = N.ChunkStats()
        N.eng_check(_eng().gdb_executor_execute_streamed(self._h, plan.handle, table.name.encode(), table.rows, n, cn,
                                                         ptrs, eb, int(chunk_rows), C.byref(st)))
        return st

    def execute_streamed_colbin(self, plan: "HipQueryPlan", table: str, colbin_path, chunk_rows: int) -> N.ChunkStats:
        """Out-of-core run with `table` read from its .colbin file chunk by
        chunk (converted to the engine's formats in pinned staging); the
        table is never held whole in host memory."""
        st = N.ChunkStats()
        N.eng_check(_eng().gdb_executor_execute_streamed_colbin(self._h, plan.handle, table.encode(),
                                                                str(colbin_path).encode(), int(chunk_rows),
                                                                C.byref(st)))
        return st

    def set_skip_init(self, skip: bool) -> None:
        N.eng_check(_eng().gdb_executor_set_skip_init(self._h, 1 if skip else 0))

    def set_phase_timing(self, on: bool) -> None:
        """Per-phase device times (default) or only the whole run's, without
        the per-phase event records on the stream."""
        N.eng_check(_eng().gdb_executor_set_phase_timing(self._h, 1 if on else 0))

    def _timing(self, plan: HipQueryPlan, t: N.ExecTiming) -> Timing:
        names = plan.phases()
        ph = [(names[i] if i < len(names) else f"phase{i}", t.phase_ms[i], t.grid[i], t.block[i], t.items[i])
              for i in range(t.nphases)]
        st = [list(t.status[i]) for i in range(t.nphases)]
        return Timing(t.total_ms, t.alloc_ms, ph, st)

    def execute(self, plan: HipQueryPlan, warmup: int = 0, measured: int = 1, first: int = 0,
                last: int = -1) -> Timing:
        t = N.ExecTiming()
        N.eng_check(_eng().gdb_executor_execute_range(self._h, plan.handle, first, last, warmup, measured,
                                                       C.byref(t)))
        return self._timing(plan, t)

    def prepare(self, plan: HipQueryPlan) -> None:
        N.eng_check(_eng().gdb_executor_prepare(self._h, plan.handle))

    def enqueue(self, plan: HipQueryPlan) -> None:
        N.eng_check(_eng().gdb_executor_enqueue(self._h, plan.handle))

    def check_status(self, plan: HipQueryPlan) -> None:
        N.eng_check(_eng().gdb_executor_check_status(self._h, plan.handle))

    def read_buffer(self, name: str, nbytes: int, dtype=np.uint64, offset: int = 0) -> np.ndarray:
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        N.eng_check(_eng().gdb_executor_read_buffer(self._h, name.encode(), offset, out.ctypes.data_as(C.c_void_p),
                                                     nbytes))
        return out

    def buffer_ptr(self, name: str) -> tuple[int, int]:
        p, n = C.c_void_p(), C.c_size_t()
        N.eng_check(_eng().gdb_executor_buffer_ptr(self._h, name.encode(), C.byref(p), C.byref(n)))
        return p.value, n.value

    def owned_bytes(self) -> int:
        n = C.c_size_t()
        N.eng_check(_eng().gdb_executor_owned_bytes(self._h, C.byref(n)))
        return n.value

    def collect(self, plan: HipQueryPlan) -> GenericResult:
        r = C.c_void_p()
        N.eng_check(_eng().gdb_executor_collect(self._h, plan.handle, C.byref(r)))
        try:
            rows, cols = C.c_uint64(), C.c_int()
            N.eng_check(_eng().gdb_result_shape(r, C.byref(rows), C.byref(cols)))
            out, fmts = {}, {}
            for i in range(cols.value):
                name = C.create_string_buffer(128)
                typ = C.c_int()
                N.eng_check(_eng().gdb_result_column(r, i, name, 128, C.byref(typ)))
                nm = name.value.decode()
                if typ.value == 0:
                    a = np.empty(rows.value, dtype=np.int64)
                    N.eng_check(_eng().gdb_result_i64(r, i, a.ctypes.data_as(C.c_void_p)))
                elif typ.value == 1:
                    w = np.empty(2 * rows.value, dtype=np.int64)
                    N.eng_check(_eng().gdb_result_i128(r, i, w.ctypes.data_as(C.c_void_p)))
                    a = np.array([int(np.uint64(w[2 * k])) + (int(w[2 * k + 1]) << 64) for k in range(rows.value)],
                                 dtype=object)
                else:
                    width = 64
                    raw = C.create_string_buffer(width * max(1, rows.value))
                    N.eng_check(_eng().gdb_result_str(r, i, raw, width))
                    a = np.array([raw.raw[k * width:(k + 1) * width].rstrip(b"\0").decode() for k in range(rows.value)],
                                 dtype=object)
                out[nm] = a
                sc, dt = C.c_int(), C.c_int()
                N.eng_check(_eng().gdb_result_column_format(r, i, C.byref(sc), C.byref(dt)))
                fmts[nm] = (sc.value, bool(dt.value))
            canon = N.two_call_string(_eng().gdb_result_canonical, r) if rows.value <= 100000 else ""
            return GenericResult(out, canon, fmts)
        finally:
            _eng().gdb_result_free(r)

    def release(self) -> None:
        N.eng_check(_eng().gdb_executor_release(self._h))

    def close(self) -> None:
        if self._h:
            _eng().gdb_executor_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
