"""Buffer initialisation fills: gdb_fill_multi_async (descriptors passed as
kernel arguments) and gdb_fill_table_async (a descriptor table in device
memory, used by batched pipelines on every replay). Every pattern width the
two forms take (1, 2, 4 bytes; 1..4 words, including the 3-word patterns of
three-word aggregate buffers), over buffers from one pattern repeat up to 1M
words, equals numpy tiling of the pattern; forms they cannot take are
refused before any launch."""

import ctypes as C

import numpy as np
import pytest

from gpudbmetalcodegen_amd import _native as N

pytestmark = pytest.mark.gpu

# (pattern bytes, buffer bytes)
CASES = [(1, 8), (2, 64), (4, 4096), (8, 8), (8, 8 * (1 << 20) + 24), (16, 16 * 48), (24, 24), (24, 24 * 1000),
         (24, 24 * 43691), (24, 24 * 400000), (32, 32), (32, 32 * 4096), (32, 32 * 300000)]


def _pattern(pl: int, seed: int) -> bytes:
    return bytes(np.random.default_rng(seed).integers(1, 256, pl, dtype=np.uint8))


def _expect(pat: bytes, nbytes: int) -> np.ndarray:
    return np.frombuffer(pat * (nbytes // len(pat)), dtype=np.uint8)


def _garbage(ctx, ptrs_sizes):
    for p, n in ptrs_sizes:
        junk = np.full(n, 0xA5, np.uint8)
        N.hip_check(N.load()[0].gdb_memcpy_h2d_async(C.c_void_p(p), junk.ctypes.data_as(C.c_void_p), n, ctx.stream))
    ctx.sync()


def test_multi_buffer_fill_every_pattern_width(ctx):
    hip = N.load()[0]
    pats = [_pattern(pl, i) for i, (pl, _) in enumerate(CASES)]
    ptrs = [ctx.alloc(n) for _, n in CASES]
    try:
        _garbage(ctx, [(p, n) for p, (_, n) in zip(ptrs, CASES)])
        n = len(CASES)
        dst = (C.c_void_p * n)(*ptrs)
        nb = (C.c_size_t * n)(*[b for _, b in CASES])
        bufs = [C.create_string_buffer(p, len(p)) for p in pats]
        pp = (C.c_void_p * n)(*[C.cast(b, C.c_void_p) for b in bufs])
        pl = (C.c_int * n)(*[len(p) for p in pats])
        N.hip_check(hip.gdb_fill_multi_async(n, dst, nb, pp, pl, ctx.stream))
        ctx.sync()
        for p, (plen, b), pat in zip(ptrs, CASES, pats):
            assert np.array_equal(ctx.download(p, b), _expect(pat, b)), (plen, b)
    finally:
        for p in ptrs:
            ctx.free(p)


def test_fill_table_every_pattern_width_and_replays(ctx):
    hip = N.load()[0]
    pats = [_pattern(pl, 100 + i) for i, (pl, _) in enumerate(CASES)]
    ptrs = [ctx.alloc(n) for _, n in CASES]
    table = (N.FillDesc * len(CASES))()
    for d, p, (plen, b), pat in zip(table, ptrs, CASES, pats):
        word = pat * (8 // plen) if plen < 8 else pat
        period = len(word) // 8
        d.dst, d.words, d.period = p, b // 8, period
        for k in range(period):
            d.pattern[k] = int.from_bytes(word[8 * k:8 * k + 8], "little")
    dtab = ctx.upload(np.frombuffer(bytes(table), dtype=np.uint8))
    try:
        for _ in range(2):      # the same table on every replay, buffers dirtied in between
            _garbage(ctx, [(p, n) for p, (_, n) in zip(ptrs, CASES)])
            N.hip_check(hip.gdb_fill_table_async(C.c_void_p(dtab), len(CASES), max(b // 8 for _, b in CASES),
                                                 ctx.stream))
            ctx.sync()
            for p, (plen, b), pat in zip(ptrs, CASES, pats):
                assert np.array_equal(ctx.download(p, b), _expect(pat, b)), (plen, b)
    finally:
        ctx.free(dtab)
        for p in ptrs:
            ctx.free(p)


@pytest.mark.parametrize("plen,nbytes", [(3, 24), (24, 32), (40, 40), (8, 12)])
def test_multi_buffer_fill_refuses_forms_it_cannot_take(ctx, plen, nbytes):
    hip = N.load()[0]
    p = ctx.alloc(max(nbytes, 16))
    try:
        pat = C.create_string_buffer(b"\x11" * plen, plen)
        st = hip.gdb_fill_multi_async(1, (C.c_void_p * 1)(p), (C.c_size_t * 1)(nbytes),
                                      (C.c_void_p * 1)(C.cast(pat, C.c_void_p)), (C.c_int * 1)(plen), ctx.stream)
        assert st != 0
        assert b"gdb_fill_multi_async" in hip.gdb_last_error()
    finally:
        ctx.free(p)
