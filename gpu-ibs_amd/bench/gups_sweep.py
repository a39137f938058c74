"""Random-bucket read rate vs working-set size and load policy (measurement only).

    python3 -u gpu-ibs_amd/bench/gups_sweep.py            # sizes x bucket widths, plain loads
    python3 -u gpu-ibs_amd/bench/gups_sweep.py policies   # load policies at 27 GB and 256 MB
    python3 -u gpu-ibs_amd/bench/gups_sweep.py alloc      # allocation kinds at 24 GiB

Drives hipibs_bench_gups_policy (gups.hip) over device buffers from 2 MB to
27 GB, with 64 B and 128 B buckets and 4 or 8 loads in flight per wave.
Policies: 0 plain, 1 non-temporal (as the table probe loads), 2 / 3 relaxed
8 B atomic loads at system / device scope, 4 sc0, 5 sc0 nt, 6 sc1 nt, 7 nt
written in asm (a control for 1). Allocation kinds: hipMalloc's default,
fine-grained and uncached device memory (hipExtMallocWithFlags; an
uncached buffer is never freed)."""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import hipibs  # noqa: E402

b = C.CDLL(os.path.join(os.path.dirname(HERE), "lib", "libhipibs_bench.so"))
b.hipibs_bench_gups_policy.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int,
                                       C.c_int, C.c_int, C.POINTER(C.c_double), C.c_void_p]
b.hipibs_bench_gups_reads.argtypes = [C.c_uint32, C.c_uint64, C.c_int, C.c_int]
b.hipibs_bench_gups_reads.restype = C.c_uint64
b.hipibs_bench_alloc.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]
b.hipibs_bench_free.argtypes = [C.c_void_p, C.c_int]
N = 500_000_000


def rate(buf, mb, S, rounds, policy, wpc=32):
    nb = (mb << 20) // (S * 8)
    t = C.c_double(0)
    ptr = buf if isinstance(buf, int) else buf.ptr
    for _ in range(2):  # warm-up, then the measurement
        rc = b.hipibs_bench_gups_policy(ptr, nb, S, N, rounds, wpc, policy, C.byref(t), None)
    if rc != 0:
        return None
    reads = b.hipibs_bench_gups_reads(S, N, rounds, wpc)
    return reads / (t.value / 1e3) / 1e9, reads * S * 8 / (t.value / 1e3) / 1e9


if len(sys.argv) > 1 and sys.argv[1] == "alloc":
    mb = 24 << 10
    for kind, name in enumerate(("default", "fine-grained", "uncached")):
        p = C.c_void_p()
        if b.hipibs_bench_alloc(mb << 20, kind, C.byref(p)) != 0:
            print(f"{name}: allocation failed", flush=True)
            continue
        for S in (8, 16):
            for policy in (0, 1):
                r = rate(p.value, mb, S, 8, policy)
                print(f"alloc {name:12s} size {mb:6d} MB bucket {S * 8:3d} B policy {policy}: "
                      f"{r[0]:7.1f} G buckets/s", flush=True)
        if kind != 2:
            b.hipibs_bench_free(p, kind)
elif len(sys.argv) > 1 and sys.argv[1] == "policies":
    for mb in (256, 27000):
        buf = hipibs.DeviceBuffer.zeros(mb << 20)
        for policy in range(8):
            for S in (8, 16):
                r = rate(buf, mb, S, 4, policy)
                if r:
                    print(f"size {mb:6d} MB bucket {S * 8:3d} B policy {policy}: "
                          f"{r[0]:7.1f} G buckets/s", flush=True)
        buf.free()
else:
    for mb in [2, 16, 64, 128, 192, 256, 384, 512, 1024, 4096, 27000]:
        buf = hipibs.DeviceBuffer.zeros(mb << 20)
        for S in (8, 16):
            for rounds in (4, 8):
                r = rate(buf, mb, S, rounds, 0)
                print(f"size {mb:6d} MB bucket {S * 8:3d} B rounds {rounds}: "
                      f"{r[0]:7.1f} G buckets/s  {r[1]:7.0f} GB/s", flush=True)
        buf.free()
