"""Streaming-copy probe: measures the achievable HBM copy bandwidth and gives
rocprofv3 FETCH_SIZE / WRITE_SIZE a launch of known size (1 GiB read + 1 GiB
written) to calibrate the counters' byte scale.

    python -m gpudbmetalcodegen_amd.copy_probe [--mib 1024] [--reps 10]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import sys

from . import _native as N
from . import engine as E


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args(argv)
    hip, _ = N.load()
    ctx = E.Context(0)
    nbytes = args.mib << 20
    src = ctx.alloc(nbytes)
    dst = ctx.alloc(nbytes)
    one = (C.c_uint8 * 1)(7)
    N.hip_check(hip.gdb_fill_async(C.c_void_p(src), nbytes, one, 1, ctx.stream))
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    N.hip_check(hip.gdb_event_create(C.byref(ev0)))
    N.hip_check(hip.gdb_event_create(C.byref(ev1)))
    N.hip_check(hip.gdb_copy_async(C.c_void_p(dst), C.c_void_p(src), nbytes, ctx.stream))  # warm
    N.hip_check(hip.gdb_event_record(ev0, ctx.stream))
    for _ in range(args.reps):
        N.hip_check(hip.gdb_copy_async(C.c_void_p(dst), C.c_void_p(src), nbytes, ctx.stream))
    N.hip_check(hip.gdb_event_record(ev1, ctx.stream))
    N.hip_check(hip.gdb_event_sync(ev1))
    ms = C.c_float()
    N.hip_check(hip.gdb_event_elapsed_ms(ev0, ev1, C.byref(ms)))
    per = ms.value / args.reps
    gbs = 2.0 * nbytes / (per / 1e3) / 1e9
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "ms_per_copy": per,
                      "copy_GBps_read_plus_write": gbs, "frac_of_8TBps": gbs / 8000.0}))
    ctx.free(src)
    ctx.free(dst)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
