// Random-sector read ceiling of the GPU (measurement only, not part of the
// product library): every wave reads random 64 B (or 128 B) buckets with
// exactly the access pattern of the table probe -- 4 (or 8) adjacent lanes
// load one bucket as 16 B each, ROUNDS loads in flight per wave -- over a
// table-sized buffer. The analogue of the reference's GUPS kernel
// (src/gpu/ProbeBench.cu:58-74), which reads 8 B words.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace
{

typedef unsigned long long U64x2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) U64x2 *GlobalU64x2;

__device__ inline uint64_t mix(uint64_t x)
{
    x ^= x >> 31;
    x *= 0x9E3779B97F4A7C15ull;
    x ^= x >> 29;
    x *= 0xBF58476D1CE4E5B9ull;
    return x ^ (x >> 32);
}

// Load policy of the bucket reads: 0 = plain 16 B loads; 1 = non-temporal
// 16 B loads (nt); 2 = two relaxed 8 B atomic loads at system scope
// (sc0 sc1); 3 = the same at device scope (sc1).
template <int POLICY>
__device__ inline U64x2 loadBucketPart(const uint64_t *p)
{
    if constexpr (POLICY == 0)
        return *(GlobalU64x2)(reinterpret_cast<const U64x2 *>(p));
    else if constexpr (POLICY == 1)
        return __builtin_nontemporal_load((GlobalU64x2)(reinterpret_cast<const U64x2 *>(p)));
    else if constexpr (POLICY == 2)
    {
        U64x2 v;
        v.x = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        v.y = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return v;
    }
    else
    {
        U64x2 v;
        v.x = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v.y = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return v;
    }
}

// Policies 4..7: four 16 B loads issued together in one asm block with
// cache-policy bits the builtins cannot express (4: sc0, 5: sc0 nt,
// 6: sc1 nt, 7: nt, the last as a control for the builtin's nt); the block
// waits for its loads before it ends.
#define GUPS_LOAD4(CPOL)                                                                         \
    asm volatile("global_load_dwordx4 %0, %4, off " CPOL "\n"                                     \
                 "global_load_dwordx4 %1, %5, off " CPOL "\n"                                     \
                 "global_load_dwordx4 %2, %6, off " CPOL "\n"                                     \
                 "global_load_dwordx4 %3, %7, off " CPOL "\n"                                     \
                 "s_waitcnt vmcnt(0)"                                                            \
                 : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3])                            \
                 : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3])                                    \
                 : "memory")

template <int POLICY>
__device__ inline void loadFour(const uint64_t *const a[4], U64x2 v[4])
{
    if constexpr (POLICY == 4)
        GUPS_LOAD4("sc0");
    else if constexpr (POLICY == 5)
        GUPS_LOAD4("sc0 nt");
    else if constexpr (POLICY == 6)
        GUPS_LOAD4("sc1 nt");
    else
        GUPS_LOAD4("nt");
}

template <int S, int POLICY>
__global__ __launch_bounds__(256) void gupsAsmKernel(const uint64_t *buf, uint64_t nBuckets,
                                                     uint64_t iters, uint64_t seed, uint64_t *out)
{
    constexpr int LPQ = S / 2;
    const int lane = threadIdx.x & 63;
    const int g = lane / LPQ, sub = lane % LPQ;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint64_t acc = 0;
    for (uint64_t it = 0; it < iters; ++it)
    {
        const uint64_t *a[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
            const uint64_t h = mix(seed ^ (((wave * iters + it) * 4 + r) * 64 + g));
            const uint64_t b = ((h >> 32) * nBuckets) >> 32;
            a[r] = buf + b * S + sub * 2;
        }
        U64x2 v[4];
        loadFour<POLICY>(a, v);
#pragma unroll
        for (int r = 0; r < 4; ++r)
            acc ^= v[r].x ^ v[r].y;
    }
    if (acc == 0x0123456789ABCDEFull)
        out[0] = acc;
}

template <int S, int ROUNDS, int POLICY = 0>
__global__ __launch_bounds__(256) void gupsKernel(const uint64_t *buf, uint64_t nBuckets,
                                                  uint64_t iters, uint64_t seed, uint64_t *out)
{
    constexpr int LPQ = S / 2;
    const int lane = threadIdx.x & 63;
    const int g = lane / LPQ, sub = lane % LPQ;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint64_t acc = 0;
    for (uint64_t it = 0; it < iters; ++it)
    {
        U64x2 v[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r)
        {
            const uint64_t h = mix(seed ^ (((wave * iters + it) * ROUNDS + r) * 64 + g));
            const uint64_t b = ((h >> 32) * nBuckets) >> 32;
            v[r] = loadBucketPart<POLICY>(buf + b * S + sub * 2);
        }
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r)
            acc ^= v[r].x ^ v[r].y;
    }
    if (acc == 0x0123456789ABCDEFull)
        out[0] = acc;
}

} // namespace

// Reads about n_reads random buckets of bucket_slots*8 bytes from d_buf
// (n_buckets buckets) with `rounds` loads in flight per wave; returns the
// kernel time in ms through *ms (HIP events on `stream`).
extern "C" int hipibs_bench_gups_policy(const void *d_buf, uint64_t n_buckets,
                                        uint32_t bucket_slots, uint64_t n_reads, int rounds,
                                        int waves_per_cu, int policy, double *ms, void *stream);

extern "C" int hipibs_bench_gups(const void *d_buf, uint64_t n_buckets, uint32_t bucket_slots,
                                 uint64_t n_reads, int rounds, int waves_per_cu, double *ms,
                                 void *stream)
{
    return hipibs_bench_gups_policy(d_buf, n_buckets, bucket_slots, n_reads, rounds, waves_per_cu,
                                    0, ms, stream);
}

// As hipibs_bench_gups with the load policy of gupsKernel (0..3, with 4 or
// 8 rounds; policies 2..3 with 64 B buckets only) or gupsAsmKernel (4..7,
// 64 B buckets, 4 rounds).
extern "C" int hipibs_bench_gups_policy(const void *d_buf, uint64_t n_buckets,
                                        uint32_t bucket_slots, uint64_t n_reads, int rounds,
                                        int waves_per_cu, int policy, double *ms, void *stream)
{
    if (policy < 0 || policy > 7 || (policy && (rounds != 4 && rounds != 8)) ||
        (policy >= 2 && bucket_slots != 8) || (policy >= 4 && rounds != 4))
        return -1;
    if (!d_buf || !ms || n_buckets == 0 || n_buckets >= (1ull << 32))
        return -1;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int blocks = cus * (waves_per_cu > 0 ? waves_per_cu : 32) / 4;
    const uint64_t waves = (uint64_t)blocks * 4;
    const int qpw = (bucket_slots == 16 ? 8 : 16) * rounds; // buckets per wave iteration
    const uint64_t iters = (n_reads + waves * qpw - 1) / (waves * qpw);
    uint64_t *out = nullptr;
    if (hipMalloc(&out, 8) != hipSuccess)
        return -2;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    const uint64_t *buf = static_cast<const uint64_t *>(d_buf);
#define GUPS_CASE(SS, RR, PP)                                                                    \
    if (bucket_slots == SS && rounds == RR && policy == PP)                                      \
        hipLaunchKernelGGL((gupsKernel<SS, RR, PP>), dim3(blocks), dim3(256), 0, s, buf,         \
                           n_buckets, iters, 0x5EEDull, out);
    GUPS_CASE(8, 1, 0)
    GUPS_CASE(8, 2, 0)
    GUPS_CASE(8, 4, 0)
    GUPS_CASE(8, 8, 0)
    GUPS_CASE(16, 4, 0)
    GUPS_CASE(16, 8, 0)
    GUPS_CASE(8, 4, 1)
    GUPS_CASE(8, 8, 1)
    GUPS_CASE(16, 4, 1)
    GUPS_CASE(16, 8, 1)
    GUPS_CASE(8, 4, 2)
    GUPS_CASE(8, 8, 2)
    GUPS_CASE(8, 4, 3)
    GUPS_CASE(8, 8, 3)
#undef GUPS_CASE
#define GUPS_ASM_CASE(PP)                                                                        \
    if (bucket_slots == 8 && rounds == 4 && policy == PP)                                        \
        hipLaunchKernelGGL((gupsAsmKernel<8, PP>), dim3(blocks), dim3(256), 0, s, buf,           \
                           n_buckets, iters, 0x5EEDull, out);
    GUPS_ASM_CASE(4)
    GUPS_ASM_CASE(5)
    GUPS_ASM_CASE(6)
    GUPS_ASM_CASE(7)
#undef GUPS_ASM_CASE
    (void)hipEventRecord(e1, s);
    hipError_t err = hipEventSynchronize(e1);
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (err != hipSuccess)
        return -3;
    *ms = t;
    return 0;
}

// Device memory of one allocation kind for the ceiling measurement: 0 =
// hipMalloc's default (coarse-grained), 1 = fine-grained, 2 = uncached
// (hipExtMallocWithFlags), filled with a byte pattern. Freeing uncached
// memory that a kernel has read can corrupt later allocations, so
// hipibs_bench_free refuses kind 2: such a buffer lives until the process
// ends.
extern "C" int hipibs_bench_alloc(uint64_t bytes, int kind, void **out)
{
    static const unsigned flags[3] = {hipDeviceMallocDefault, hipDeviceMallocFinegrained,
                                      hipDeviceMallocUncached};
    if (!out || kind < 0 || kind > 2 || bytes == 0)
        return -1;
    *out = nullptr;
    if (hipExtMallocWithFlags(out, bytes, flags[kind]) != hipSuccess)
        return -2;
    if (hipMemset(*out, 0x5A, bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return -3;
    return 0;
}

extern "C" int hipibs_bench_free(void *p, int kind)
{
    if (kind == 2)
        return -1;
    return hipFree(p) == hipSuccess ? 0 : -2;
}

// Buckets actually read by hipibs_bench_gups for these arguments.
extern "C" uint64_t hipibs_bench_gups_reads(uint32_t bucket_slots, uint64_t n_reads, int rounds,
                                            int waves_per_cu)
{
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint64_t waves = (uint64_t)(cus * (waves_per_cu > 0 ? waves_per_cu : 32) / 4) * 4;
    const int qpw = (bucket_slots == 16 ? 8 : 16) * rounds;
    const uint64_t iters = (n_reads + waves * qpw - 1) / (waves * qpw);
    return iters * waves * qpw;
}
