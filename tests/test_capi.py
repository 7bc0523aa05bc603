"""C-ABI boundary checks that need no GPU: the libraries load, export every
function the headers declare, build and emit every plan, compile generated
sources with hipRTC into the code-object cache, and report errors through
status codes."""

import ctypes as C
import re
from pathlib import Path

import pytest

from gpudbmetalcodegen_amd import _native as N
from gpudbmetalcodegen_amd import engine as E

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include" / "gdb").glob("*.h"))

PLANS = [("Q1", ""), ("Q6", ""), ("Q3", ""), ("Q3", "probe_bitmap=0"), ("Q3_SIZE", ""), ("GROUPBY", ""),
         ("GROUPBY", "variant=global,filter=1"), ("GB_COUNT", ""), ("MAXKEY", "table=G,column=g_key"),
         ("Q13L", ""), ("Q13L", "negate=1"), ("Q22A", ""), ("PIPE_AGG", "lds_cap=4096"), ("PIPE_SEMI", ""),
         ("PIPE_ANTI", "")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(gdb_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert len(names) > 60
    assert "gdb_launch" in names and "gdb_plan_build" in names


def test_every_declared_symbol_is_exported():
    hip, eng = N.load()
    missing = [n for n in declared_functions() if not (hasattr(hip, n) or hasattr(eng, n))]
    assert missing == []


def test_no_gpu_paths_report_status():
    n = E.device_count()
    if n > 0:
        pytest.skip("a GPU is visible")
    hip, _ = N.load()
    h = C.c_void_p()
    assert hip.gdb_init(0, C.byref(h)) == N.GDB_ERR_NO_DEVICE
    assert b"no GPU" in hip.gdb_last_error()


def test_batch_command_fails_loudly_without_a_gpu():
    """The batched workload command refuses bad settings before touching the
    device and, with no GPU visible, stops with the runtime's error instead
    of falling back to anything."""
    from gpudbmetalcodegen_amd import batch as B
    with pytest.raises(SystemExit):
        B.main(["--replays", "x"])
    if E.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(N.GdbError, match="no GPU"):
        B.main(["--replays", "1", "--set", "4,12,2"])


@pytest.mark.parametrize("name,opts", PLANS)
def test_plan_build_describe_source(name, opts):
    p = E.build_hip_plan(name, opts)
    d = p.describe()
    src = p.source()
    assert f"Plan {name}" in d
    for ph in p.phases():
        assert f'extern "C" __global__' in src and f"{ph}_i" in src
    assert "gdb device runtime" in src


def test_unknown_plan_and_bad_options():
    with pytest.raises(E.GdbError) as ei:
        E.build_hip_plan("Q99")
    assert ei.value.code == N.GDB_ERR_NOT_FOUND
    with pytest.raises(E.GdbError):
        E.build_hip_plan("Q1", "items")  # not key=value


def test_set_launch_validates_variants():
    p = E.build_hip_plan("Q1")
    p.set_launch("Q1_reduce", 128, 4)
    assert p.get_launch("Q1_reduce") == (128, 4)
    with pytest.raises(E.GdbError):
        p.set_launch("Q1_reduce", 256, 3)   # no ITEMS=3 variant compiled
    with pytest.raises(E.GdbError):
        p.set_launch("Q1_reduce", 512, 4)   # above the launch bound
    with pytest.raises(E.GdbError):
        p.set_launch("nope", 256, 4)


def test_hiprtc_precompile_cache_hit_and_corruption(tmp_path):
    p = E.build_hip_plan("MAXKEY", "table=G,column=g_key")
    assert p.precompile(str(tmp_path)) is False        # compiled
    assert p.precompile(str(tmp_path)) is True         # cache hit
    files = list(tmp_path.glob("*.co"))
    assert len(files) == 1
    raw = bytearray(files[0].read_bytes())
    raw[-1] ^= 0xFF                                    # corrupt the payload
    files[0].write_bytes(bytes(raw))
    assert p.precompile(str(tmp_path)) is False        # checksum mismatch -> recompiled
    assert p.precompile(str(tmp_path)) is True


def test_hiprtc_compile_error_surfaces_log(tmp_path):
    hip, _ = N.load()
    src = b'extern "C" __global__ void broken(int* p) { p[0] = undefined_name; }'
    h, wc = C.c_uint64(), C.c_int()
    st = hip.gdb_module_precompile(str(tmp_path).encode(), b"gfx950", src, len(src), None, 0, C.byref(h),
                                   C.byref(wc))
    assert st == N.GDB_ERR_COMPILE
    msg = hip.gdb_last_error().decode()
    assert "undefined_name" in msg and "1: extern" in msg   # numbered source dump


def test_plan_hash_depends_on_source_and_options(tmp_path):
    hip, _ = N.load()
    src = b'extern "C" __global__ void k(int* p) { p[0] = 1; }'
    keys = set()
    for opts in ([], [b"-DX=1"]):
        arr = (C.c_char_p * max(1, len(opts)))(*opts) if opts else None
        h, wc = C.c_uint64(), C.c_int()
        assert hip.gdb_module_precompile(str(tmp_path).encode(), b"gfx950", src, len(src), arr, len(opts),
                                         C.byref(h), C.byref(wc)) == 0
        keys.add(h.value)
    assert len(keys) == 2


def test_batched_kernel_compiles_with_shared_lds(tmp_path):
    """Every ITEMS variant of a 128 KiB LDS group-by body in one persistent
    kernel: the bodies share one LDS buffer, so the kernel fits the CU's
    160 KiB instead of needing the sum."""
    agg = E.build_hip_plan("PIPE_AGG", "lds_cap=4096")
    semi, anti = E.build_hip_plan("PIPE_SEMI"), E.build_hip_plan("PIPE_ANTI")
    plans = [agg] * 5 + [semi, anti]
    items = [1, 2, 4, 8, 16, 4, 8]
    src = E.HipQueryPlan.batch_kernel_source(plans, items).encode()
    assert src.count(b"unsigned char _lds[") == 1 and b"__shared__ u64 _gl_" not in src
    hip, _ = N.load()
    h, wc = C.c_uint64(), C.c_int()
    st = hip.gdb_module_precompile(str(tmp_path).encode(), b"gfx950", src, len(src), None, 0, C.byref(h), C.byref(wc))
    assert st == 0, hip.gdb_last_error().decode()[-2000:]


def test_partial_table_plan_and_two_level_batch_kernel(tmp_path):
    """PIPE_AGG with partials=N: the scan phase stores LDS tables as partial
    tables and a PIPE_merge phase folds them (not chunkable: the partial
    count restarts once per run); a batch of it emits one persistent kernel
    per phase level, and the source compiles for gfx950."""
    agg = E.build_hip_plan("PIPE_AGG", "lds_cap=4096,partials=256")
    assert agg.phases() == ["PIPE_agg", "PIPE_merge"] and not agg.chunkable
    src = agg.source()
    assert "d_pipe_ht_pt" in src and "gdb_lds_gb_merge_at" in src
    assert E.build_hip_plan("PIPE_AGG", "lds_cap=4096").phases() == ["PIPE_agg"]
    semi = E.build_hip_plan("PIPE_SEMI")
    bsrc = E.HipQueryPlan.batch_kernel_source([agg, agg, semi], [4, 16, 8])
    assert "void __launch_bounds__(1024) gdb_batch_kernel(" in bsrc
    assert "void __launch_bounds__(1024) gdb_batch_kernel_l1(" in bsrc
    b = bsrc.encode()
    hip, _ = N.load()
    h, wc = C.c_uint64(), C.c_int()
    st = hip.gdb_module_precompile(str(tmp_path).encode(), b"gfx950", b, len(b), None, 0, C.byref(h), C.byref(wc))
    assert st == 0, hip.gdb_last_error().decode()[-2000:]


def test_q3_a_bitmap_prefix_in_plain_launches_only_when_asked():
    """Q3's a_lds_words=N copies the A bitmap's first N words into LDS once
    per block of the plain count and build launches (and raises their
    default block to 1024); without it neither phase declares the copy."""
    on = E.build_hip_plan("Q3", "a_lds_words=4096")
    src = on.source()
    assert src.count("__attribute__((aligned(16))) u32 _bmc_d_q3_a_bitmap[4096]") >= 2
    assert on.get_launch("Q3_count_b")[0] == 1024 and on.get_launch("Q3_build_b_table")[0] == 1024
    off = E.build_hip_plan("Q3")
    assert "u32 _bmc_d_q3_a_bitmap[" not in off.source()    # (the A-bitmap build's own u64 cache stays)
    assert off.get_launch("Q3_count_b")[0] == 256 and off.get_launch("Q3_build_b_table")[0] == 512


def test_join_bitmap_lds_cache_only_in_batched_bodies():
    """lds_bitmap_words=N keeps a bitmap prefix in LDS only in the batched
    (range-mode) body, whose LDS a group-by body sizes anyway; the plain
    launch is unchanged."""
    semi = E.build_hip_plan("PIPE_SEMI", "lds_bitmap_words=32768")
    assert "_bmc_" not in semi.source()
    agg = E.build_hip_plan("PIPE_AGG", "lds_cap=4096")
    src = E.HipQueryPlan.batch_kernel_source([agg, semi], [4, 8])
    assert "_bmc_d_pipe_bitmap[_w] = d_pipe_bitmap[_w]" in src


def test_chunkable_plans():
    """Out-of-core eligibility follows the associative-output rule: plans
    whose stream-table phases only accumulate are chunkable; a plan whose
    sizing pre-pass needs the whole table first is not."""
    for name, opts in [("Q1", ""), ("Q6", ""), ("Q3", ""), ("Q13L", ""), ("Q22A", ""), ("GROUPBY", ""),
                       ("PIPE_AGG", ""), ("PIPE_SEMI", "")]:
        assert E.build_hip_plan(name, opts).chunkable, name
    for name, opts in [("GROUPBY", "sizing=prepass"), ("GB_COUNT", ""), ("MAXKEY", "table=G,column=g_key")]:
        assert not E.build_hip_plan(name, opts).chunkable, name


def test_scalar_atomic_ablation_changes_the_reduction(monkeypatch):
    """GDB_SCALAR_ATOMIC=1 (the reference's GPUDB_SCALAR_ATOMIC ablation):
    small-bucket aggregates update the output with one global atomic per
    row instead of the register + wave + LDS block reduction."""
    from gpudbmetalcodegen_amd import engine as E

    base = E.build_hip_plan("Q1", "").source()
    assert "block reduction: wave DPP sums" in base and "reduction ablation" not in base
    monkeypatch.setenv("GDB_SCALAR_ATOMIC", "1")
    abl = E.build_hip_plan("Q1", "").source()
    assert "reduction ablation" in abl and "block reduction: wave DPP sums" not in abl
    monkeypatch.setenv("GDB_SCALAR_ATOMIC", "0")
    assert E.build_hip_plan("Q1", "").source() == base
