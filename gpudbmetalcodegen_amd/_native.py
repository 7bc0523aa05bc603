"""ctypes bindings of the two native libraries (include/gdb/gdb_hip.h,
include/gdb/gdb_engine.h).

The libraries are loaded from this package directory only. A missing library
is an ImportError-like failure with the build hint; nothing falls back to a
Python or CPU implementation.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
HIP_LIB = PKG / "libgdbhip.so"
ENGINE_LIB = PKG / "libgdbengine.so"

GDB_OK = 0
GDB_ERR_HIP = 1
GDB_ERR_COMPILE = 2
GDB_ERR_INVALID = 3
GDB_ERR_OOM = 4
GDB_ERR_OVERFLOW = 5
GDB_ERR_NOT_FOUND = 6
GDB_ERR_ADMISSION = 7
GDB_ERR_NO_DEVICE = 8

MAX_PHASES = 16


class GdbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[gdb status {code}] {msg}")
        self.code = code


class GdbOverflow(GdbError):
    pass


class DeviceInfo(C.Structure):
    _fields_ = [
        ("device", C.c_int), ("cu_count", C.c_int), ("lds_per_block", C.c_int), ("warp_size", C.c_int),
        ("clock_khz", C.c_int), ("mem_clock_khz", C.c_int), ("mem_bus_width", C.c_int), ("l2_bytes", C.c_int),
        ("hbm_bytes", C.c_longlong), ("arch", C.c_char * 64), ("name", C.c_char * 128),
    ]


class ChunkStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("stream_ms", C.c_double), ("chunks", C.c_uint64),
                ("bytes_copied", C.c_uint64), ("first_stream_phase", C.c_int), ("last_stream_phase", C.c_int)]


class ExecTiming(C.Structure):
    _fields_ = [
        ("total_ms", C.c_float), ("alloc_ms", C.c_float), ("nphases", C.c_int),
        ("phase_ms", C.c_float * MAX_PHASES), ("grid", C.c_uint * MAX_PHASES),
        ("block", C.c_uint * MAX_PHASES), ("items", C.c_uint * MAX_PHASES),
        ("status", (C.c_uint64 * 4) * MAX_PHASES),
    ]


_hip = None
_eng = None

vp = C.c_void_p
u64 = C.c_uint64
i64 = C.c_int64
sz = C.c_size_t
cstr = C.c_char_p
i32 = C.c_int
P = C.POINTER


def _sig(lib, name, args, res=C.c_int):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res


class FillDesc(C.Structure):
    """gdb_fill_desc (gdb_hip.h): one buffer of a device-resident fill table."""
    _fields_ = [("dst", C.c_void_p), ("words", C.c_uint64), ("period", C.c_uint64), ("pattern", C.c_uint64 * 4)]


def load() -> tuple[C.CDLL, C.CDLL]:
    """Load libgdbhip.so and libgdbengine.so (raises if they are not built)."""
    global _hip, _eng
    if _hip is not None:
        return _hip, _eng
    for p in (HIP_LIB, ENGINE_LIB):
        if not p.exists():
            raise ImportError(f"{p.name} is not built; run `python -m gpudbmetalcodegen_amd.build`")
    hip = C.CDLL(str(HIP_LIB), mode=C.RTLD_GLOBAL)
    eng = C.CDLL(str(ENGINE_LIB), mode=C.RTLD_GLOBAL)
    # gdb_hip.h
    _sig(hip, "gdb_last_error", [], C.c_char_p)
    _sig(hip, "gdb_device_count", [P(i32)])
    _sig(hip, "gdb_init", [i32, P(vp)])
    _sig(hip, "gdb_shutdown", [vp])
    _sig(hip, "gdb_device_info", [vp, P(DeviceInfo)])
    _sig(hip, "gdb_default_stream", [vp], vp)
    _sig(hip, "gdb_alloc", [vp, sz, P(vp)])
    _sig(hip, "gdb_free", [vp, vp])
    _sig(hip, "gdb_mem_info", [vp, P(sz), P(sz)])
    _sig(hip, "gdb_host_alloc_pinned", [sz, P(vp)])
    _sig(hip, "gdb_host_free_pinned", [vp])
    _sig(hip, "gdb_memcpy_h2d_async", [vp, vp, sz, vp])
    _sig(hip, "gdb_memcpy_d2h_async", [vp, vp, sz, vp])
    _sig(hip, "gdb_memcpy_d2d_async", [vp, vp, sz, vp])
    _sig(hip, "gdb_fill_async", [vp, sz, vp, i32, vp])
    _sig(hip, "gdb_fill_multi_async", [i32, P(vp), P(sz), P(vp), P(i32), vp])
    _sig(hip, "gdb_fill_table_async", [vp, i32, u64, vp])
    _sig(hip, "gdb_memset_async", [vp, i32, sz, vp])
    _sig(hip, "gdb_module_load", [vp, cstr, sz, P(cstr), i32, P(u64), P(vp)])
    _sig(hip, "gdb_module_precompile", [cstr, cstr, cstr, sz, P(cstr), i32, P(u64), P(i32)])
    _sig(hip, "gdb_module_unload", [vp])
    _sig(hip, "gdb_kernel_get", [vp, cstr, P(vp)])
    _sig(hip, "gdb_kernel_occupancy", [vp, i32, sz, P(i32)])
    _sig(hip, "gdb_kernel_attributes", [vp, P(i32), P(i32), P(i32), P(i32)])
    _sig(hip, "gdb_launch", [vp, C.c_uint, C.c_uint, C.c_uint, vp, sz, vp])
    _sig(hip, "gdb_cache_set_dir", [vp, cstr])
    _sig(hip, "gdb_cache_stats", [vp, P(u64), P(u64), P(C.c_double)])
    _sig(hip, "gdb_stream_create", [vp, P(vp)])
    _sig(hip, "gdb_stream_sync", [vp])
    _sig(hip, "gdb_stream_query", [vp, P(i32)])
    _sig(hip, "gdb_stream_destroy", [vp])
    _sig(hip, "gdb_stream_wait_event", [vp, vp])
    _sig(hip, "gdb_event_create", [P(vp)])
    _sig(hip, "gdb_event_record", [vp, vp])
    _sig(hip, "gdb_event_sync", [vp])
    _sig(hip, "gdb_event_query", [vp, P(i32)])
    _sig(hip, "gdb_event_elapsed_ms", [vp, vp, P(C.c_float)])
    _sig(hip, "gdb_event_destroy", [vp])
    _sig(hip, "gdb_device_sync", [vp])
    _sig(hip, "gdb_graph_begin_capture", [vp])
    _sig(hip, "gdb_graph_end_capture", [vp, P(vp)])
    _sig(hip, "gdb_graph_launch", [vp, vp])
    _sig(hip, "gdb_graph_destroy", [vp])
    # gdb_engine.h
    _sig(eng, "gdb_engine_last_error", [], C.c_char_p)
    _sig(eng, "gdb_gen_uniform_i32", [u64, u64, u64, i64, i64, vp])
    _sig(eng, "gdb_gen_uniform_i64", [u64, u64, u64, i64, i64, vp])
    _sig(eng, "gdb_gen_categorical_u8", [u64, u64, u64, vp, vp, i32, vp])
    _sig(eng, "gdb_gen_permutation_i32", [u64, u64, u64, i64, vp])
    _sig(eng, "gdb_gen_fanout_total", [u64, u64, u64, i32, i32, P(u64)])
    _sig(eng, "gdb_gen_fanout_fkey_i32", [u64, u64, u64, i32, i32, u64, vp])
    _sig(eng, "gdb_gen_zipf_i32", [u64, u64, u64, C.c_double, i64, vp])
    _sig(eng, "gdb_gen_vocabulary", [u64, i32, i32, i32, vp])
    _sig(eng, "gdb_gen_text", [u64, u64, u64, i32, vp, i32, vp])
    _sig(eng, "gdb_plan_build", [cstr, cstr, P(vp)])
    _sig(eng, "gdb_plan_build_sql", [cstr, cstr, P(vp)])
    _sig(eng, "gdb_plan_sql_layout", [vp, vp, sz, P(sz)])
    _sig(eng, "gdb_plan_describe", [vp, vp, sz, P(sz)])
    _sig(eng, "gdb_plan_source", [vp, vp, sz, P(sz)])
    _sig(eng, "gdb_plan_phase_count", [vp, P(i32)])
    _sig(eng, "gdb_plan_phase_name", [vp, i32, vp, sz])
    _sig(eng, "gdb_plan_set_launch", [vp, cstr, i32, i32])
    _sig(eng, "gdb_plan_get_launch", [vp, cstr, P(i32), P(i32)])
    _sig(eng, "gdb_plan_compile", [vp, vp, P(C.c_double), P(i32), P(u64)])
    _sig(eng, "gdb_plan_precompile", [vp, cstr, cstr, P(i32), P(u64)])
    _sig(eng, "gdb_plan_free", [vp])
    _sig(eng, "gdb_executor_create", [vp, vp, P(vp)])
    _sig(eng, "gdb_executor_destroy", [vp])
    _sig(eng, "gdb_executor_register_table", [vp, cstr, u64])
    _sig(eng, "gdb_executor_register_column", [vp, cstr, cstr, vp])
    _sig(eng, "gdb_executor_register_buffer", [vp, cstr, vp, sz])
    _sig(eng, "gdb_executor_register_symbol", [vp, cstr, u64])
    _sig(eng, "gdb_executor_get_symbol", [vp, cstr, P(u64), P(i32)])
    _sig(eng, "gdb_executor_register_scalar", [vp, cstr, i64])
    _sig(eng, "gdb_executor_set_skip_init", [vp, i32])
    _sig(eng, "gdb_executor_set_phase_timing", [vp, i32])
    _sig(eng, "gdb_executor_execute", [vp, vp, i32, i32, P(ExecTiming)])
    _sig(eng, "gdb_executor_execute_range", [vp, vp, i32, i32, i32, i32, P(ExecTiming)])
    _sig(eng, "gdb_executor_prepare", [vp, vp])
    _sig(eng, "gdb_executor_execute_streamed", [vp, vp, cstr, u64, i32, P(cstr), P(vp), P(i32), u64, P(ChunkStats)])
    _sig(eng, "gdb_executor_execute_streamed_colbin", [vp, vp, cstr, cstr, u64, P(ChunkStats)])
    _sig(eng, "gdb_plan_chunkable", [vp, P(i32)])
    _sig(eng, "gdb_executor_enqueue", [vp, vp])
    _sig(eng, "gdb_executor_check_status", [vp, vp])
    _sig(eng, "gdb_executor_read_buffer", [vp, cstr, sz, vp, sz])
    _sig(eng, "gdb_executor_buffer_ptr", [vp, cstr, P(vp), P(sz)])
    _sig(eng, "gdb_executor_owned_bytes", [vp, P(sz)])
    _sig(eng, "gdb_executor_release", [vp])
    _sig(eng, "gdb_executor_collect", [vp, vp, P(vp)])
    _sig(eng, "gdb_result_shape", [vp, P(u64), P(i32)])
    _sig(eng, "gdb_result_column", [vp, i32, vp, sz, P(i32)])
    _sig(eng, "gdb_result_i64", [vp, i32, vp])
    _sig(eng, "gdb_result_i128", [vp, i32, vp])
    _sig(eng, "gdb_result_str", [vp, i32, vp, sz])
    _sig(eng, "gdb_result_column_format", [vp, i32, P(i32), P(i32)])
    _sig(eng, "gdb_result_canonical", [vp, vp, sz, P(sz)])
    # TPC-H ingest and the golden check
    _sig(eng, "gdb_tpch_column", [cstr, cstr, P(i32), P(i32), P(i32)])
    _sig(eng, "gdb_tpch_table_columns", [cstr, vp, sz, P(sz)])
    _sig(eng, "gdb_tbl_to_colbin", [cstr, cstr, i32, i32, P(i32)])
    _sig(eng, "gdb_colbin_open", [cstr, cstr, P(vp)])
    _sig(eng, "gdb_colbin_shape", [vp, P(u64), P(i32)])
    _sig(eng, "gdb_colbin_column", [vp, i32, P(vp), P(u64), P(i32), P(i32)])
    _sig(eng, "gdb_colbin_close", [vp], None)
    _sig(eng, "gdb_compare_canonical", [cstr, cstr, C.c_double, C.c_double, vp, sz, P(sz)])
    _sig(eng, "gdb_result_free", [vp])
    # gdb_kernels.h (in libgdbhip.so)
    _sig(hip, "gdb_kernels_last_error", [], C.c_char_p)
    _sig(hip, "gdb_mlp_forward", [vp, vp, vp, i32, vp])
    _sig(hip, "gdb_mlp_backward", [vp, vp, vp, vp, vp, i32, C.c_float, vp])
    _sig(hip, "gdb_mlp_reduce_adam", [vp, vp, vp, vp, vp, i32, i32, C.c_float, C.c_float, C.c_float, C.c_float, vp])
    f32 = C.c_float
    _sig(hip, "gdb_mlp_train_step_noalloc", [vp, vp, vp, vp, vp, vp, vp, i32, f32, i32, f32, f32, f32, f32, vp])
    _sig(hip, "gdb_gen_device_zipf_i32", [vp, u64, C.c_double, i64, u64, vp])
    _sig(hip, "gdb_gen_device_uniform_i64", [vp, u64, i64, i64, u64, vp])
    _sig(hip, "gdb_copy_async", [vp, vp, sz, vp])
    # autotuner + batch (gdb_engine.h)
    _sig(eng, "gdb_autotune_phase", [vp, vp, cstr, vp, i32, vp, i32, cstr, vp, i32, i32, P(i32), P(i32), P(i64),
                                     P(C.c_float), vp, P(i32)])
    _sig(eng, "gdb_batch_create", [vp, P(vp)])
    _sig(eng, "gdb_batch_add", [vp, vp, vp, i32, P(i32)])
    _sig(eng, "gdb_batch_build", [vp, i32, u64, i32, P(C.c_uint), P(u64)])
    _sig(eng, "gdb_batch_set_level_chunk", [vp, i32, u64])
    _sig(eng, "gdb_batch_source", [vp, vp, sz, P(sz)])
    _sig(eng, "gdb_batch_kernel_source", [P(vp), P(i32), i32, vp, sz, P(sz)])
    _sig(eng, "gdb_batch_capture", [vp])
    _sig(eng, "gdb_batch_run", [vp, i32, P(C.c_float)])
    _sig(eng, "gdb_batch_destroy", [vp])
    # estimator (gdb_engine.h)
    _sig(eng, "gdb_est_features", [u64, i64, i64, i64, vp, vp, u64, vp, P(C.c_double)])
    _sig(eng, "gdb_est_sample", [vp, vp, vp, u64, u64, vp, vp])
    _sig(eng, "gdb_est_create", [vp, P(vp)])
    _sig(eng, "gdb_est_destroy", [vp])
    _sig(eng, "gdb_est_init", [vp, u64])
    _sig(eng, "gdb_est_train", [vp, vp, vp, i32, i32, i32, C.c_float, u64, vp])
    _sig(eng, "gdb_est_predict", [vp, vp, i32, vp])
    _sig(eng, "gdb_est_save", [vp, cstr])
    _sig(eng, "gdb_est_load", [vp, cstr])
    _sig(eng, "gdb_est_get_state", [vp, vp, vp, vp])
    _sig(eng, "gdb_est_set_state", [vp, vp, vp, vp])
    _sig(eng, "gdb_est_collect_runs", [vp, i32, u64, u64, u64, i64, i64, C.c_double, C.c_double, vp, vp, vp, vp])
    _hip, _eng = hip, eng
    return hip, eng


def hip_check(st: int) -> None:
    if st != GDB_OK:
        msg = _hip.gdb_last_error().decode(errors="replace")
        raise (GdbOverflow if st == GDB_ERR_OVERFLOW else GdbError)(st, msg)


def eng_check(st: int) -> None:
    if st != GDB_OK:
        msg = _eng.gdb_engine_last_error().decode(errors="replace")
        raise (GdbOverflow if st == GDB_ERR_OVERFLOW else GdbError)(st, msg)


def two_call_string(fn, *args) -> str:
    need = sz(0)
    eng_check(fn(*args, None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value)
    eng_check(fn(*args, buf, need.value, C.byref(need)))
    return buf.value.decode()
