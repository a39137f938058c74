"""Random-bucket read rate vs working-set size (measurement only).

    python3 -u gpu-ibs_amd/bench/gups_sweep.py

Drives hipibs_bench_gups (gups.hip) over device buffers from 2 MB to 27 GB,
with 64 B and 128 B buckets and 4 or 8 loads in flight per wave."""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import hipibs  # noqa: E402

b = C.CDLL(os.path.join(os.path.dirname(HERE), "lib", "libhipibs_bench.so"))
b.hipibs_bench_gups.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int, C.c_int,
                                C.POINTER(C.c_double), C.c_void_p]
b.hipibs_bench_gups_reads.argtypes = [C.c_uint32, C.c_uint64, C.c_int, C.c_int]
b.hipibs_bench_gups_reads.restype = C.c_uint64

for mb in [2, 16, 64, 128, 192, 256, 384, 512, 1024, 4096, 27000]:
    buf = hipibs.DeviceBuffer.zeros(mb << 20)
    for S in (8, 16):
        nb = (mb << 20) // (S * 8)
        for rounds in (4, 8):
            t = C.c_double(0)
            n = 500_000_000
            b.hipibs_bench_gups(buf.ptr, nb, S, n, rounds, 32, C.byref(t), None)  # warm-up
            b.hipibs_bench_gups(buf.ptr, nb, S, n, rounds, 32, C.byref(t), None)
            reads = b.hipibs_bench_gups_reads(S, n, rounds, 32)
            print(f"size {mb:6d} MB bucket {S * 8:3d} B rounds {rounds}: "
                  f"{reads / (t.value / 1e3) / 1e9:7.1f} G buckets/s  "
                  f"{reads * S * 8 / (t.value / 1e3) / 1e9:7.0f} GB/s", flush=True)
    buf.free()
