// Device tables: construction (insert kernel + host API) and the probe /
// probe-count kernels behind hipibs_table_probe{,_count{,_device}}.
//
// The table replaces the reference's two device structures -- the KMC
// LUT + binary-search image (src/gpu/DeviceKmerIndex.cu:28-142, ~log2(bucket)+3
// dependent loads per probe) and the double-hashed 8 B/key hash set
// (src/gpu/DeviceHashIndex.cu:31-195) -- with one bucketised open-addressing
// table whose probe costs one 64 B (or 128 B) HBM access in the common case.
#include "wave_probe.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace hipibs
{

// Stores `key` from bucket `home` on: walks the buckets, claiming the first
// EMPTY slot with a 64-bit CAS; a slot already holding the key ends the
// insert (duplicates are stored once). Slots of a bucket are claimed in
// order, so a concurrent insert of the same key always meets the first copy.
// Returns false when every bucket is full; *isNew / *disp report a new slot
// and its displacement (buckets past home).
template <int S>
__device__ inline bool insertOne(uint64_t *slots, uint64_t nBuckets, uint64_t key, uint64_t home,
                                 bool *isNew, uint32_t *disp)
{
    uint64_t b = home;
    *isNew = false;
    for (uint64_t d = 0; d < nBuckets; ++d)
    {
        unsigned long long *bucket = reinterpret_cast<unsigned long long *>(slots + b * S);
        for (int s = 0; s < S; ++s)
        {
            unsigned long long cur = __hip_atomic_load(bucket + s, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
            if (cur == key)
                return true;
            if (cur == EMPTY)
            {
                const unsigned long long prev = atomicCAS(bucket + s, EMPTY, key);
                if (prev == EMPTY)
                {
                    *isNew = true;
                    *disp = (uint32_t)min<uint64_t>(d, 0xffffffffu);
                    return true;
                }
                if (prev == key)
                    return true;
            }
        }
        b += 1;
        if (b == nBuckets)
            b = 0;
    }
    return false;
}

// One key per thread into a whole-k-mer table. counters[0] += newly stored
// keys, counters[1] = 1 when the table ran out of buckets, counters[2] =
// max displacement.
template <int S>
__global__ __launch_bounds__(256) void insertKernel(uint64_t *slots, const TableView T,
                                                    const uint64_t *keys, uint64_t n,
                                                    unsigned long long *counters)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t localNew = 0, localDisp = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const uint64_t key = keys[i];
        if (key == EMPTY)
            continue;
        bool isNew = false;
        uint32_t d = 0;
        if (!insertOne<S>(slots, T.nBuckets, key, homeOf(key, T), &isNew, &d))
            atomicExch(counters + 1, 1ull);
        else if (isNew)
        {
            localNew += 1;
            localDisp = max(localDisp, d);
        }
    }
    if (localNew)
        atomicAdd(counters + 0, (unsigned long long)localNew);
    if (localDisp)
        atomicMax(counters + 2, (unsigned long long)localDisp);
}

// One 33..64-base key per thread (keys[2i] high word, keys[2i+1] low word)
// into a table of 8 keys per 128 B bucket; a key already stored is not
// stored again. A thread claims a slot by moving its high word from EMPTY
// to its own (CAS) and then publishes the low word with a device-scope
// store. A thread that meets its own high word in a claimed slot reads the
// low word -- waiting, at most kWideSpin reads, while it still reads EMPTY
// (the claimer's store follows its CAS directly; within a wave the claim
// step precedes the compare step in program order) -- and stops if it is
// its own. A low word that really is all ones outlasts the wait and the key
// is stored in a further slot: a harmless extra copy. A walk ends after
// kWideMaxWalk buckets (a few at load <= 0.98), then the table reports
// overflow. A row whose high word is all ones holds no key and is skipped,
// except k = 64's T^32 A^32 (low word 0), which sets counters[3] instead
// of taking a slot. Counters as insertKernel.
constexpr uint64_t kWideMaxWalk = 4096;
constexpr uint32_t kWideSpin = 1u << 14;
__global__ __launch_bounds__(256) void insertWideKernel(uint64_t *slots, const TableView T,
                                                        const uint64_t *keys, uint64_t n,
                                                        unsigned long long *counters)
{
    uint64_t localNew = 0;
    uint32_t localDisp = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t walk = T.nBuckets < kWideMaxWalk ? T.nBuckets : kWideMaxWalk;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const uint64_t hi = keys[2 * i], lo = keys[2 * i + 1];
        if (hi == EMPTY)
        {
            // T^32 A^32 is stored as a flag; any other row with an all-ones
            // high word is no key (the encoder's mark of an invalid position)
            if (T.k == 64 && isWideTop(hi, lo) && atomicCAS(counters + 3, 0ull, 1ull) == 0ull)
                localNew += 1;
            continue;
        }
        uint64_t b = bucketOfWide(hi, lo, T.nBuckets);
        bool done = false;
        for (uint64_t d = 0; d < walk && !done; ++d)
        {
            unsigned long long *bucket = reinterpret_cast<unsigned long long *>(slots + b * 16);
            for (int q = 0; q < 8 && !done; ++q)
            {
                unsigned long long *hw = bucket + 2 * q, *lw = bucket + 2 * q + 1;
                unsigned long long cur =
                    __hip_atomic_load(hw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cur == EMPTY)
                {
                    const unsigned long long prev = atomicCAS(hw, EMPTY, hi);
                    if (prev == EMPTY)
                    {
                        __hip_atomic_store(lw, (unsigned long long)lo, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                        done = true;
                        localNew += 1;
                        localDisp = max(localDisp, (uint32_t)min<uint64_t>(d, 0xffffffffu));
                    }
                    else
                        cur = prev;
                }
                if (!done && cur == hi)
                {
                    unsigned long long l =
                        __hip_atomic_load(lw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    for (uint32_t w = 0; l == EMPTY && w < kWideSpin; ++w)
                        l = __hip_atomic_load(lw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    done = l == lo;
                }
            }
            b = b + 1 == T.nBuckets ? 0 : b + 1;
        }
        if (!done)
            atomicExch(counters + 1, 1ull);
    }
    if (localNew)
        atomicAdd(counters + 0, (unsigned long long)localNew);
    if (localDisp)
        atomicMax(counters + 2, (unsigned long long)localDisp);
}

// ---- minimizer tables ------------------------------------------------------
// A group = the keys sharing a 32-bit minimizer hash. The group map holds
// one 64-bit entry per group: hash in the low 32 bits, key count in bits
// 32..62, bit 63 set once the group's marker is stored. EMPTY = all ones.

__device__ inline uint64_t groupSlot(uint32_t minHash, uint64_t mask)
{
    uint64_t x = (uint64_t)minHash * 0xD6E8FEB86659FD93ull;
    x ^= x >> 32;
    return x & mask;
}

// Counts the keys of every group; *overflow = 1 if the map filled up (the
// walk of a key is bounded by the map size, so the kernel always ends).
__global__ __launch_bounds__(256) void groupCountKernel(const uint64_t *keys, uint64_t n,
                                                        uint32_t k, uint32_t mlen,
                                                        unsigned long long *map, uint64_t mask,
                                                        unsigned long long *overflow)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const uint64_t key = keys[i];
        if (key == EMPTY)
            continue;
        const uint32_t mh = minimizerOf(key, k, mlen);
        bool counted = false;
        uint64_t h = groupSlot(mh, mask);
        for (uint64_t step = 0; step <= mask && !counted; ++step, h = (h + 1) & mask)
        {
            unsigned long long cur = __hip_atomic_load(map + h, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
            if (cur == EMPTY)
            {
                cur = atomicCAS(map + h, EMPTY, (unsigned long long)mh | (1ull << 32));
                if (cur == EMPTY)
                {
                    counted = true;
                    break;
                }
            }
            // cur is a claimed entry here
            if ((uint32_t)cur == mh)
            {
                atomicAdd(map + h, 1ull << 32);
                counted = true;
            }
        }
        if (!counted)
            atomicExch(overflow, 1ull);
    }
}

// tally[0] += keys of groups of at most cap keys, tally[1] += larger
// groups (one marker each), tally[2] += their keys; tally[4] = 1 if a
// group's count reached 2^31 (bit 63 of its entry).
__global__ __launch_bounds__(256) void groupTallyKernel(const unsigned long long *map,
                                                        uint64_t size, uint32_t cap,
                                                        unsigned long long *tally)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long a = 0, g = 0, b = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride)
    {
        const unsigned long long e = map[i];
        if (e == EMPTY)
            continue;
        if (e >> 63)
            atomicExch(tally + 4, 1ull);
        const uint64_t c = (e >> 32) & 0x7fffffffull;
        if (c <= cap)
            a += c;
        else
        {
            g += 1;
            b += c;
        }
    }
    if (a)
        atomicAdd(tally + 0, a);
    if (g)
        atomicAdd(tally + 1, g);
    if (b)
        atomicAdd(tally + 2, b);
}

// Places every key: groups of at most cap keys into the minimizer-ordered
// part at their group's home bucket; larger groups' keys into the
// whole-k-mer part, and the first of them stores the group's marker at the
// group's home. counters: [0] keys in the minimizer part, [1] full flag,
// [2] its max displacement, [3] keys in the whole-k-mer part, [4] its max
// displacement, [5] markers.
template <int S>
__global__ __launch_bounds__(256) void placeKernel(const uint64_t *keys, uint64_t n, uint32_t k,
                                                   uint32_t mlen, unsigned long long *map,
                                                   uint64_t mask, uint32_t cap, uint64_t *slots,
                                                   uint64_t nBuckets, uint64_t *slots2,
                                                   uint64_t nBuckets2,
                                                   unsigned long long *counters)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t nA = 0, dA = 0, nB = 0, dB = 0, nM = 0;
    bool full = false;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const uint64_t key = keys[i];
        if (key == EMPTY)
            continue;
        const uint32_t mh = minimizerOf(key, k, mlen);
        uint64_t h = groupSlot(mh, mask);
        unsigned long long e = EMPTY;
        // every key's group was counted, and entries are never removed: the
        // walk meets the group before any EMPTY entry
        for (uint64_t step = 0; step <= mask; ++step, h = (h + 1) & mask)
        {
            e = map[h];
            if (e == EMPTY || (uint32_t)e == mh)
                break;
        }
        if (e == EMPTY || (uint32_t)e != mh)
        {
            full = true;
            continue;
        }
        const uint64_t c = (e >> 32) & 0x7fffffffull;
        bool isNew = false;
        uint32_t d = 0;
        if (c <= cap)
        {
            full |= !insertOne<S>(slots, nBuckets, key, bucketOfMinimizer(mh, nBuckets), &isNew,
                                  &d);
            if (isNew)
            {
                nA += 1;
                dA = max(dA, d);
            }
            continue;
        }
        full |= !insertOne<8>(slots2, nBuckets2, key, bucketOf(key, nBuckets2), &isNew, &d);
        if (isNew)
        {
            nB += 1;
            dB = max(dB, d);
        }
        const unsigned long long old = atomicOr(map + h, 1ull << 63);
        if (!(old >> 63))
        {
            full |= !insertOne<S>(slots, nBuckets, markerOf(mh), bucketOfMinimizer(mh, nBuckets),
                                  &isNew, &d);
            if (isNew)
            {
                nM += 1;
                dA = max(dA, d);
            }
        }
    }
    if (full)
        atomicExch(counters + 1, 1ull);
    if (nA)
        atomicAdd(counters + 0, (unsigned long long)nA);
    if (dA)
        atomicMax(counters + 2, (unsigned long long)dA);
    if (nB)
        atomicAdd(counters + 3, (unsigned long long)nB);
    if (dB)
        atomicMax(counters + 4, (unsigned long long)dB);
    if (nM)
        atomicAdd(counters + 5, (unsigned long long)nM);
}

__global__ void fillKernel(uint64_t *p, uint64_t n, uint64_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

// Last bucket j in [lo, hi] with lut[j] <= r (lut is non-decreasing and
// lut[lo] <= r).
__device__ inline uint64_t lutBucketOf(const uint64_t *lut, uint64_t lo, uint64_t hi, uint64_t r)
{
    while (lo < hi)
    {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if (lut[mid] <= r)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Keys of the KMC records [r0, r0+m) from their stored bytes (raw[0] is
// record r0): record r of bucket j (lut[j] <= r < lut[j+1]) holds the key
// ((j & prefixMask) << shift) | (its first sufBytes bytes, big-endian);
// the counter bytes that follow are skipped. j0..j1 are the buckets the
// records span. A thread decodes 8 consecutive records.
constexpr uint32_t kDecodePerThread = 8;
__global__ __launch_bounds__(256) void kmcDecodeKernel(const uint8_t *__restrict__ raw,
                                                       uint64_t r0, uint64_t m, uint32_t recBytes,
                                                       uint32_t sufBytes,
                                                       const uint64_t *__restrict__ lut,
                                                       uint64_t j0, uint64_t j1,
                                                       uint64_t prefixMask, uint32_t shift,
                                                       uint64_t *__restrict__ keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * kDecodePerThread;
    for (uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kDecodePerThread;
         i0 < m; i0 += stride)
    {
        uint64_t j = lutBucketOf(lut, j0, j1, r0 + i0);
        const uint64_t iEnd = i0 + kDecodePerThread < m ? i0 + kDecodePerThread : m;
        for (uint64_t i = i0; i < iEnd; ++i)
        {
            const uint64_t r = r0 + i;
            if (j < j1 && lut[j + 1] <= r)
                j = lutBucketOf(lut, j + 1, j1, r);
            const uint8_t *q = raw + i * recBytes;
            uint64_t v = 0;
            for (uint32_t b = 0; b < sufBytes; ++b)
                v = (v << 8) | q[b];
            keys[i] = (shift < 64 ? (j & prefixMask) << shift : 0ull) | v;
        }
    }
}

// kmcDecodeKernel for k = 33..64: record i gives keys[2i] (high word) and
// keys[2i+1] (low word); the suffix holds up to 16 bytes.
__global__ __launch_bounds__(256) void kmcDecodeWideKernel(
    const uint8_t *__restrict__ raw, uint64_t r0, uint64_t m, uint32_t recBytes,
    uint32_t sufBytes, const uint64_t *__restrict__ lut, uint64_t j0, uint64_t j1,
    uint64_t prefixMask, uint32_t shift, uint64_t *__restrict__ keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * kDecodePerThread;
    for (uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kDecodePerThread;
         i0 < m; i0 += stride)
    {
        uint64_t j = lutBucketOf(lut, j0, j1, r0 + i0);
        const uint64_t iEnd = i0 + kDecodePerThread < m ? i0 + kDecodePerThread : m;
        for (uint64_t i = i0; i < iEnd; ++i)
        {
            const uint64_t r = r0 + i;
            if (j < j1 && lut[j + 1] <= r)
                j = lutBucketOf(lut, j + 1, j1, r);
            const uint8_t *q = raw + i * recBytes;
            uint64_t hi = 0, lo = 0;
            for (uint32_t b = 0; b < sufBytes; ++b)
            {
                hi = (hi << 8) | (lo >> 56);
                lo = (lo << 8) | q[b];
            }
            const uint64_t pre = j & prefixMask;
            if (shift < 64)
            {
                lo |= pre << shift;
                hi |= shift ? pre >> (64 - shift) : 0ull;
            }
            else if (shift < 128)
                hi |= pre << (shift - 64);
            keys[2 * i] = hi;
            keys[2 * i + 1] = lo;
        }
    }
}

// Wide-key membership: codes[2i], codes[2i+1] = (high, low) of query i;
// otherwise as probeKernel.
__global__ __launch_bounds__(256) void probeWideKernel(const TableView T, const uint64_t *codes,
                                                       const uint8_t *valid, uint64_t n,
                                                       uint8_t *hits, unsigned long long *count)
{
    const int lane = laneId();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nWaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t nChunks = (n + 63) >> 6;
    uint64_t local = 0;
    for (uint64_t c = wave; c < nChunks; c += nWaves)
    {
        const uint64_t i = (c << 6) + (uint64_t)lane;
        const bool in = i < n;
        const bool v = in && (valid == nullptr || valid[i] != 0);
        const uint64_t hi = v ? codes[2 * i] : 0ull, lo = v ? codes[2 * i + 1] : 0ull;
        const uint64_t H =
            waveProbeWide(T, hi, lo, v ? (uint32_t)bucketOfWide(hi, lo, T.nBuckets) : 0u, v);
        if (hits && in)
            hits[i] = (uint8_t)((H >> lane) & 1ull);
        local += (uint64_t)__popcll(H);
    }
    if (count && lane == 0 && local)
        atomicAdd(count, (unsigned long long)local);
}

// 64 queries per wave per step; hits (nullable) gets one byte per query,
// count (nullable) the number of hits. Grid-stride over 64-query chunks.
template <int S>
__global__ __launch_bounds__(256) void probeKernel(const TableView T, const uint64_t *codes,
                                                   const uint8_t *valid, uint64_t n,
                                                   uint8_t *hits, unsigned long long *count)
{
    const int lane = laneId();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nWaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t nChunks = (n + 63) >> 6;
    uint64_t local = 0;
    for (uint64_t c = wave; c < nChunks; c += nWaves)
    {
        const uint64_t i = (c << 6) + (uint64_t)lane;
        const bool in = i < n;
        bool v = in && (valid == nullptr || valid[i] != 0);
        const uint64_t code = v ? codes[i] : 0ull;
        const uint64_t H = waveLookup<S>(T, code, v);
        if (hits && in)
            hits[i] = (uint8_t)((H >> lane) & 1ull);
        local += (uint64_t)__popcll(H);
    }
    if (count && lane == 0 && local)
        atomicAdd(count, (unsigned long long)local);
}

static int gridFor(uint64_t work, int block, int device)
{
    int cus = 256;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    const uint64_t want = (work + block - 1) / block;
    const uint64_t cap = (uint64_t)cus * 16;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(want, cap));
}

int launchTableInsert(hipibs_table *t, const uint64_t *dKeys, uint64_t n, hipStream_t s)
{
    if (n == 0)
        return 0;
    const int grid = gridFor(n, 256, t->device);
    auto *ctr = reinterpret_cast<unsigned long long *>(t->dCounters);
    if (t->k > 32)
        hipLaunchKernelGGL(insertWideKernel, dim3(grid), dim3(256), 0, s, t->dSlots, t->view(),
                           dKeys, n, ctr);
    else if (t->bucketSlots == 8)
        hipLaunchKernelGGL(insertKernel<8>, dim3(grid), dim3(256), 0, s, t->dSlots, t->view(),
                           dKeys, n, ctr);
    else
        hipLaunchKernelGGL(insertKernel<16>, dim3(grid), dim3(256), 0, s, t->dSlots, t->view(),
                           dKeys, n, ctr);
    HIPIBS_CHECK(hipGetLastError());
    return 0;
}

int launchProbe(const hipibs_table *t, const uint64_t *dCodes, const uint8_t *dValid, uint64_t n,
                uint8_t *dHits, uint64_t *dCount, hipStream_t s)
{
    if (n == 0)
        return 0;
    const uint64_t waves = (n + 63) / 64;
    const int grid = gridFor(waves * 64, 256, t->device);
    auto *cnt = reinterpret_cast<unsigned long long *>(dCount);
    if (t->k > 32)
        hipLaunchKernelGGL(probeWideKernel, dim3(grid), dim3(256), 0, s, t->view(), dCodes,
                           dValid, n, dHits, cnt);
    else if (t->bucketSlots == 8)
        hipLaunchKernelGGL(probeKernel<8>, dim3(grid), dim3(256), 0, s, t->view(), dCodes, dValid,
                           n, dHits, cnt);
    else
        hipLaunchKernelGGL(probeKernel<16>, dim3(grid), dim3(256), 0, s, t->view(), dCodes, dValid,
                           n, dHits, cnt);
    HIPIBS_CHECK(hipGetLastError());
    return 0;
}

// Allocates and clears a table sized for nKeys keys.
int allocTable(uint64_t nKeys, uint32_t k, const hipibs_table_opts *opts, hipibs_table **out,
               hipStream_t s, bool streamOrdered, bool clearSlots)
{
    hipibs_table_opts o{};
    if (opts)
        o = *opts;
    if (o.load_factor <= 0)
        o.load_factor = 0.6;
    if (k == 0 || k > 64)
        return setError(HIPIBS_E_ARG, "k must be in 1..64");
    const bool wide = k > 32; // 8 keys of 16 B per 128 B bucket
    if (wide && o.minimizer_len)
        return setError(HIPIBS_E_ARG, "minimizer tables need k <= 32");
    if (wide && o.bucket_slots == 8)
        return setError(HIPIBS_E_ARG, "tables of k > 32 have 128 B buckets (bucket_slots 16)");
    const uint32_t mlen = o.minimizer_len;
    if (mlen && (mlen < 5 || mlen > 31 || mlen > k))
        return setError(HIPIBS_E_ARG, "minimizer_len must be 0 (off) or 5..min(31, k)");
    if (o.bucket_slots == 0)
        o.bucket_slots = 16; // one 128 B bucket = one memory request on gfx950
    if (o.bucket_slots != 8 && o.bucket_slots != 16)
        return setError(HIPIBS_E_ARG, "bucket_slots must be 8 or 16");
    if (!(o.load_factor > 0 && o.load_factor < 0.98))
        return setError(HIPIBS_E_ARG, "load_factor must be in (0, 0.98)");
    const double slotsWanted = std::ceil((double)std::max<uint64_t>(nKeys, 1) / o.load_factor);
    uint64_t nBuckets = (uint64_t)std::ceil(slotsWanted / (wide ? 8 : o.bucket_slots));
    nBuckets = std::max<uint64_t>(nBuckets, 2);
    if (nBuckets >= (1ull << 32))
        return setError(HIPIBS_E_ARG, "table too large: more than 2^32 buckets");
    const uint64_t bytes = nBuckets * o.bucket_slots * 8ull;
    if (o.hbm_budget && bytes > o.hbm_budget)
        return setError(HIPIBS_E_NOMEM, "table of " + std::to_string(bytes) +
                                            " bytes exceeds the HBM budget of " +
                                            std::to_string(o.hbm_budget));
    HIPIBS_CHECK(hipSetDevice(o.device));
    auto *t = new hipibs_table();
    t->device = o.device;
    t->k = k;
    t->bucketSlots = o.bucket_slots;
    t->mlen = mlen;
    t->nBuckets = nBuckets;
    t->bytes = bytes;
    t->streamOrdered = streamOrdered;
    auto alloc = [&](void **p, uint64_t b) {
        return streamOrdered ? hipMallocAsync(p, b, s) : hipMalloc(p, b);
    };
    hipError_t e = alloc(reinterpret_cast<void **>(&t->dSlots), bytes);
    if (e != hipSuccess && !streamOrdered)
    {
        // the buffers kept for the next counter or build may be in the way
        (void)hipGetLastError();
        releaseSortCaches();
        e = alloc(reinterpret_cast<void **>(&t->dSlots), bytes);
    }
    if (e != hipSuccess)
    {
        delete t;
        return setError(HIPIBS_E_NOMEM, "allocation of " + std::to_string(bytes) +
                                             " table bytes failed: " + hipGetErrorString(e));
    }
    e = alloc(reinterpret_cast<void **>(&t->dCounters), 4 * sizeof(uint64_t));
    if (e != hipSuccess)
    {
        freeTableOn(t, s);
        return setError(HIPIBS_E_NOMEM, "allocation of table counters failed");
    }
    const uint64_t nSlots = nBuckets * o.bucket_slots;
    if (clearSlots)
        hipLaunchKernelGGL(fillKernel, dim3(gridFor(nSlots, 256, o.device)), dim3(256), 0, s,
                           t->dSlots, nSlots, EMPTY);
    hipLaunchKernelGGL(fillKernel, dim3(1), dim3(64), 0, s, t->dCounters, (uint64_t)4, 0ull);
    e = hipGetLastError();
    if (e != hipSuccess)
    {
        freeTableOn(t, s);
        return hipFail(e, "table clear");
    }
    *out = t;
    return 0;
}

int clearTableSlots(hipibs_table *t, hipStream_t s)
{
    const uint64_t nSlots = t->nBuckets * t->bucketSlots;
    hipLaunchKernelGGL(fillKernel, dim3(gridFor(nSlots, 256, t->device)), dim3(256), 0, s,
                       t->dSlots, nSlots, EMPTY);
    HIPIBS_CHECK(hipGetLastError());
    return 0;
}

void freeTableOn(hipibs_table *t, hipStream_t s)
{
    if (!t)
        return;
    (void)hipSetDevice(t->device);
    for (uint64_t *p : {t->dSlots, t->dSlots2, t->dCounters})
        if (p)
            (void)(t->streamOrdered ? hipFreeAsync(p, s) : hipFree(p));
    delete t;
}

// Reads back the build counters; fails when the table overflowed.
int finishTable(hipibs_table *t, hipStream_t s)
{
    uint64_t c[4] = {0, 0, 0, 0};
    HIPIBS_CHECK(hipMemcpyAsync(c, t->dCounters, sizeof(c), hipMemcpyDeviceToHost, s));
    HIPIBS_CHECK(hipStreamSynchronize(s));
    if (c[1])
        return setError(HIPIBS_E_FULL, t->k > 32 ? "table overflow: a key found no free slot "
                                                   "within 4096 buckets"
                                                 : "table overflow: no free slot for a key");
    t->nKeys = c[0];
    t->maxDisp = (uint32_t)c[2];
    if (t->k > 32)
        t->wideTop = c[3] ? 1u : 0u;
    return 0;
}

// Default keys per minimizer group kept in the minimizer-ordered part.
constexpr uint32_t kDefaultMinimizerCap = 32;

static uint64_t pow2AtLeast(uint64_t x)
{
    uint64_t p = 1;
    while (p < x)
        p <<= 1;
    return p;
}

static int allocSlots(uint64_t items, uint32_t slotsPerBucket, double loadFactor, int device,
                      uint64_t **out, uint64_t *nBuckets, hipStream_t s)
{
    const double slotsWanted = std::ceil((double)std::max<uint64_t>(items, 1) / loadFactor);
    uint64_t nb = std::max<uint64_t>(2, (uint64_t)std::ceil(slotsWanted / slotsPerBucket));
    if (nb >= (1ull << 32))
        return setError(HIPIBS_E_ARG, "table too large: more than 2^32 buckets");
    const uint64_t bytes = nb * slotsPerBucket * 8ull;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(out), bytes);
    if (e != hipSuccess)
        return setError(HIPIBS_E_NOMEM, "hipMalloc of " + std::to_string(bytes) +
                                            " table bytes failed: " + hipGetErrorString(e));
    hipLaunchKernelGGL(fillKernel, dim3(gridFor(nb * slotsPerBucket, 256, device)), dim3(256), 0,
                       s, *out, nb * slotsPerBucket, EMPTY);
    *nBuckets = nb;
    return 0;
}

// A minimizer table's build in three steps: begin() sizes and clears the
// group map for n keys; count() adds keys to their minimizer groups' counts
// (any number of calls, e.g. one per chunk while the rest of the keys are
// still being read); finish() sizes both parts from the counts and places
// all n keys, which must then be resident together.
struct MinimizerBuild
{
    hipibs_table_opts o{};
    uint32_t k = 0, mlen = 0, cap = 0;
    uint64_t mapSize = 0;
    unsigned long long *map = nullptr, *tally = nullptr;

    int begin(uint64_t n, uint32_t kk, const hipibs_table_opts *opts, hipStream_t s)
    {
        if (opts)
            o = *opts;
        if (o.load_factor <= 0)
            o.load_factor = 0.6;
        k = kk;
        mlen = o.minimizer_len;
        if (k == 0 || k > 32)
            return setError(HIPIBS_E_ARG, "k must be in 1..32");
        if (mlen < 5 || mlen > 31 || mlen > k)
            return setError(HIPIBS_E_ARG, "minimizer_len must be 0 (off) or 5..min(31, k)");
        if (o.bucket_slots == 0)
            o.bucket_slots = 16;
        if (o.bucket_slots != 8 && o.bucket_slots != 16)
            return setError(HIPIBS_E_ARG, "bucket_slots must be 8 or 16");
        if (!(o.load_factor > 0 && o.load_factor < 0.98))
            return setError(HIPIBS_E_ARG, "load_factor must be in (0, 0.98)");
        cap = o.minimizer_cap ? o.minimizer_cap : kDefaultMinimizerCap;
        if (cap > 4096)
            return setError(HIPIBS_E_ARG, "minimizer_cap must be at most 4096");
        // a group map entry holds its group's key count in bits 32..62 (bit
        // 63 is the placement's marker flag): with fewer than 2^32 keys a
        // count never wraps, and one that reaches 2^31 shows as bit 63 in
        // the tally
        if (n >= (1ull << 32))
            return setError(HIPIBS_E_ARG, "minimizer tables hold fewer than 2^32 keys");
        HIPIBS_CHECK(hipSetDevice(o.device));
        // group map: at most one group per key, so a map of >= n entries
        // always has room
        mapSize = pow2AtLeast(std::max<uint64_t>(1024, n));
        if (hipMalloc(reinterpret_cast<void **>(&map), mapSize * 8) != hipSuccess)
            return setError(HIPIBS_E_NOMEM, "hipMalloc of the minimizer group map failed");
        if (hipMalloc(reinterpret_cast<void **>(&tally), 5 * 8) != hipSuccess)
        {
            release();
            return setError(HIPIBS_E_NOMEM, "hipMalloc of the group tally failed");
        }
        hipLaunchKernelGGL(fillKernel, dim3(gridFor(mapSize, 256, o.device)), dim3(256), 0, s,
                           reinterpret_cast<uint64_t *>(map), mapSize, EMPTY);
        hipLaunchKernelGGL(fillKernel, dim3(1), dim3(64), 0, s,
                           reinterpret_cast<uint64_t *>(tally), (uint64_t)5, 0ull);
        if (hipGetLastError() != hipSuccess)
        {
            release();
            return setError(HIPIBS_E_HIP, "minimizer group map clear failed");
        }
        return 0;
    }

    int count(const uint64_t *dKeys, uint64_t m, hipStream_t s)
    {
        if (m)
            hipLaunchKernelGGL(groupCountKernel, dim3(gridFor(m, 256, o.device)), dim3(256), 0, s,
                               dKeys, m, k, mlen, map, mapSize - 1, tally + 3);
        return hipGetLastError() == hipSuccess
                   ? 0
                   : setError(HIPIBS_E_HIP, "minimizer group count launch failed");
    }

    // Frees the map and the tally once the stream is done with them.
    void release()
    {
        if (map)
            hipFree(map);
        if (tally)
            hipFree(tally);
        map = tally = nullptr;
    }

    int finish(const uint64_t *dKeys, uint64_t n, hipStream_t s, hipibs_table **out);
};

int MinimizerBuild::finish(const uint64_t *dKeys, uint64_t n, hipStream_t s, hipibs_table **out)
{
    *out = nullptr;
    const int dev = o.device;
    hipLaunchKernelGGL(groupTallyKernel, dim3(gridFor(mapSize, 256, dev)), dim3(256), 0, s, map,
                       mapSize, cap, tally);
    uint64_t tl[5] = {0, 0, 0, 0, 0};
    int rc = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(tl, tally, sizeof(tl), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        rc = setError(HIPIBS_E_HIP, "minimizer group count failed");
    if (!rc && tl[3])
        rc = setError(HIPIBS_E_HIP, "minimizer group map overflow");
    if (!rc && tl[4])
        rc = setError(HIPIBS_E_ARG, "a minimizer group holds 2^31 keys or more");
    hipibs_table *t = nullptr;
    if (!rc)
    {
        t = new hipibs_table();
        t->device = o.device;
        t->k = k;
        t->bucketSlots = o.bucket_slots;
        t->mlen = mlen;
        t->mcap = cap;
        rc = allocSlots(tl[0] + tl[1], o.bucket_slots, o.load_factor, dev, &t->dSlots,
                        &t->nBuckets, s);
        if (!rc)
            rc = allocSlots(tl[2], 8, o.load_factor, dev, &t->dSlots2, &t->nBuckets2, s);
        t->bytes = t->nBuckets * o.bucket_slots * 8ull + t->nBuckets2 * 64ull;
        if (!rc && o.hbm_budget && t->bytes > o.hbm_budget)
            rc = setError(HIPIBS_E_NOMEM, "table of " + std::to_string(t->bytes) +
                                              " bytes exceeds the HBM budget of " +
                                              std::to_string(o.hbm_budget));
        if (!rc && hipMalloc(reinterpret_cast<void **>(&t->dCounters), 8 * sizeof(uint64_t)) !=
                       hipSuccess)
            rc = setError(HIPIBS_E_NOMEM, "hipMalloc of table counters failed");
        if (!rc)
            hipLaunchKernelGGL(fillKernel, dim3(1), dim3(64), 0, s, t->dCounters, (uint64_t)8, 0ull);
    }
    if (!rc && n)
    {
        auto *ctr = reinterpret_cast<unsigned long long *>(t->dCounters);
        if (t->bucketSlots == 8)
            hipLaunchKernelGGL(placeKernel<8>, dim3(gridFor(n, 256, dev)), dim3(256), 0, s, dKeys,
                               n, k, mlen, map, mapSize - 1, cap, t->dSlots, t->nBuckets,
                               t->dSlots2, t->nBuckets2, ctr);
        else
            hipLaunchKernelGGL(placeKernel<16>, dim3(gridFor(n, 256, dev)), dim3(256), 0, s, dKeys,
                               n, k, mlen, map, mapSize - 1, cap, t->dSlots, t->nBuckets,
                               t->dSlots2, t->nBuckets2, ctr);
        if (hipGetLastError() != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "minimizer table placement launch failed");
    }
    uint64_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!rc && (hipMemcpyAsync(c, t->dCounters, sizeof(c), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        rc = setError(HIPIBS_E_HIP, "minimizer table placement failed");
    if (!rc && c[1])
        rc = setError(HIPIBS_E_FULL, "table overflow: no free slot for a key");
    (void)hipStreamSynchronize(s);
    release();
    if (rc)
    {
        hipibs_table_free(t);
        return rc;
    }
    t->nKeys = c[0] + c[3];
    t->maxDisp = (uint32_t)c[2];
    t->nKeys2 = c[3];
    t->maxDisp2 = (uint32_t)c[4];
    t->nMarkers = c[5];
    *out = t;
    return 0;
}

// A minimizer table from n device-resident keys: count the keys of every
// minimizer group, size both parts from the counts, then place the keys.
static int buildMinimizerTable(const uint64_t *dKeys, uint64_t n, uint32_t k,
                               const hipibs_table_opts *opts, hipStream_t s, hipibs_table **out)
{
    *out = nullptr;
    MinimizerBuild mb;
    int rc = mb.begin(n, k, opts, s);
    if (!rc)
        rc = mb.count(dKeys, n, s);
    if (rc)
    {
        (void)hipStreamSynchronize(s);
        mb.release();
        return rc;
    }
    return mb.finish(dKeys, n, s, out);
}

} // namespace hipibs

using namespace hipibs;

namespace hipibs
{
int buildKeysDevice(const uint64_t *d_keys, uint64_t n, uint32_t k, const hipibs_table_opts *opts,
                    hipibs_table **out, hipStream_t s, bool allowSorted)
{
    if (!out || (n && !d_keys))
        return setError(HIPIBS_E_ARG, "null argument");
    *out = nullptr;
    if (opts && opts->minimizer_len && k > 32)
        return setError(HIPIBS_E_ARG, "minimizer tables need k <= 32");
    if (opts && opts->minimizer_len)
        return buildMinimizerTable(d_keys, n, k, opts, s, out);
    hipibs_table *t = nullptr;
    // a sorted build writes every bucket: the slots are cleared only when
    // the keys go in one by one
    int rc = allocTable(n, k, opts, &t, s, false, !allowSorted);
    if (rc)
        return rc;
    bool sorted = false;
    if (allowSorted)
        rc = buildTableSorted(t, d_keys, n, s, &sorted);
    if (!rc && !sorted && allowSorted)
        rc = clearTableSlots(t, s);
    if (!rc && !sorted)
        rc = launchTableInsert(t, d_keys, n, s);
    if (!rc)
        rc = finishTable(t, s);
    if (rc)
    {
        hipibs_table_free(t);
        return rc;
    }
    *out = t;
    return 0;
}
} // namespace hipibs

extern "C" int hipibs_table_build_keys_device(const uint64_t *d_keys, uint64_t n, uint32_t k,
                                              const hipibs_table_opts *opts, hipibs_table **out,
                                              void *stream)
{
    return buildKeysDevice(d_keys, n, k, opts, out, reinterpret_cast<hipStream_t>(stream), true);
}

// memcpy on up to 4 host threads (staging copies into pinned memory are
// memory-bandwidth bound; one thread reaches a fraction of it, and builds
// may run on several threads at once).
static void parallelCopy(void *dst, const void *src, uint64_t bytes)
{
    const uint64_t piece = 16ull << 20;
    const unsigned n = (unsigned)std::min<uint64_t>(4, (bytes + piece - 1) / piece);
    if (n <= 1)
    {
        std::memcpy(dst, src, bytes);
        return;
    }
    const uint64_t per = (bytes + n - 1) / n;
    std::vector<std::thread> pool;
    for (unsigned i = 0; i < n; ++i)
        pool.emplace_back([=] {
            const uint64_t a = std::min(bytes, (uint64_t)i * per), b = std::min(bytes, a + per);
            std::memcpy(static_cast<char *>(dst) + a, static_cast<const char *>(src) + a, b - a);
        });
    for (auto &th : pool)
        th.join();
}

// Two pinned host buffers, two device buffers and a stream for staging host
// keys into table builds; callers that build often keep one.
struct HostStaging
{
    int device = 0;
    uint64_t chunk = 0; // keys per buffer
    uint64_t *pinned[2] = {nullptr, nullptr};
    uint64_t *dev[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    hipStream_t stream = nullptr;
};

namespace hipibs
{

void hostStagingFree(HostStaging *st)
{
    if (!st)
        return;
    (void)hipSetDevice(st->device);
    if (st->stream)
        (void)hipStreamSynchronize(st->stream);
    for (int i = 0; i < 2; ++i)
    {
        if (st->pinned[i])
            (void)hipHostFree(st->pinned[i]);
        if (st->dev[i])
            (void)hipFree(st->dev[i]);
        if (st->done[i])
            (void)hipEventDestroy(st->done[i]);
    }
    if (st->stream)
        (void)hipStreamDestroy(st->stream);
    delete st;
}

int hostStagingCreate(int device, uint64_t chunkKeys, HostStaging **out)
{
    *out = nullptr;
    HIPIBS_CHECK(hipSetDevice(device));
    auto *st = new HostStaging();
    st->device = device;
    st->chunk = std::max<uint64_t>(chunkKeys, 1);
    int rc = 0;
    if (hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking) != hipSuccess)
        rc = setError(HIPIBS_E_HIP, "stream creation failed");
    for (int i = 0; i < 2 && !rc; ++i)
        if (hipHostMalloc(reinterpret_cast<void **>(&st->pinned[i]), st->chunk * 8) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&st->dev[i]), st->chunk * 8) != hipSuccess ||
            hipEventCreateWithFlags(&st->done[i], hipEventDisableTiming) != hipSuccess)
            rc = setError(HIPIBS_E_NOMEM, "staging allocation failed");
    if (rc)
    {
        hostStagingFree(st);
        return rc;
    }
    *out = st;
    return 0;
}

// Streams host keys through the staging buffers into t (the copy of chunk
// i+1 overlaps the upload and insert of chunk i) and reads back the build
// counters.
static int buildFromHostChunks(hipibs_table *t, const uint64_t *hKeys, uint64_t n,
                               HostStaging *st)
{
    const uint64_t chunk = st->chunk;
    hipStream_t s = st->stream;
    int rc = 0;
    bool used[2] = {false, false};
    for (uint64_t off = 0, it = 0; off < n && !rc; off += chunk, ++it)
    {
        const int b = (int)(it & 1);
        const uint64_t m = std::min(chunk, n - off);
        if (used[b] && hipEventSynchronize(st->done[b]) != hipSuccess)
        {
            rc = setError(HIPIBS_E_HIP, "event wait failed");
            break;
        }
        parallelCopy(st->pinned[b], hKeys + off, m * 8);
        if (hipMemcpyAsync(st->dev[b], st->pinned[b], m * 8, hipMemcpyHostToDevice, s) !=
            hipSuccess)
        {
            rc = setError(HIPIBS_E_HIP, "key upload failed");
            break;
        }
        rc = launchTableInsert(t, st->dev[b], m, s);
        if (!rc && hipEventRecord(st->done[b], s) != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "event record failed");
        used[b] = true;
    }
    if (!rc)
        rc = finishTable(t, s);
    (void)hipStreamSynchronize(s);
    return rc;
}

// hipibs_table_build_keys with the caller's staging (its stream also clears
// the new table, so no other stream is involved).
int buildTableFromHost(const uint64_t *hKeys, uint64_t n, uint32_t k, const hipibs_table_opts *opts,
                       HostStaging *st, hipibs_table **out)
{
    *out = nullptr;
    hipibs_table *t = nullptr;
    int rc = allocTable(n, k, opts, &t, st->stream);
    if (rc)
        return rc;
    rc = buildFromHostChunks(t, hKeys, n, st);
    if (rc)
    {
        hipibs_table_free(t);
        return rc;
    }
    *out = t;
    return 0;
}

} // namespace hipibs

// Streams the records of a KMC database through two pinned staging buffers:
// the host only reads raw records (parallel preads, 4-9 B per key instead of
// 8 B keys over PCIe); the device decodes them with the LUT and hands each
// chunk of keys to `consume` (enqueued on s). The read of chunk i+1 overlaps
// the upload, decode and consumption of chunk i. Does not synchronise s at
// the end.
template <class Consume>
static int streamKmcKeys(const KmcRecords &src, int device, unsigned threads, hipStream_t s,
                         Consume consume)
{
    const uint64_t total = src.total;
    if (total == 0)
        return 0;
    const uint64_t nBuckets = src.lut.size() - 1;
    const uint64_t chunk = std::min<uint64_t>(total, 32ull << 20);
    const uint64_t rawBytes = chunk * src.recBytes;
    const uint32_t p = src.info.lut_prefix_length, k = src.info.kmer_length;
    const uint64_t prefixMask = (1ull << (2 * p)) - 1ull;
    const uint32_t shift = 2 * (k - p);
    uint8_t *pinned[2] = {nullptr, nullptr}, *raw[2] = {nullptr, nullptr};
    uint64_t *dKeys = nullptr, *dLut = nullptr;
    hipEvent_t copied[2] = {nullptr, nullptr};
    int rc = 0;
    for (int i = 0; i < 2 && !rc; ++i)
        if (hipHostMalloc(reinterpret_cast<void **>(&pinned[i]), rawBytes) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&raw[i]), rawBytes) != hipSuccess ||
            hipEventCreateWithFlags(&copied[i], hipEventDisableTiming) != hipSuccess)
            rc = setError(HIPIBS_E_NOMEM, "staging allocation failed");
    const bool wide = k > 32; // two words per key
    if (!rc && (hipMalloc(reinterpret_cast<void **>(&dKeys), chunk * (wide ? 16 : 8)) != hipSuccess ||
                hipMalloc(reinterpret_cast<void **>(&dLut), src.lut.size() * 8) != hipSuccess))
        rc = setError(HIPIBS_E_NOMEM, "device allocation for the KMC decode failed");
    if (!rc && hipMemcpyAsync(dLut, src.lut.data(), src.lut.size() * 8, hipMemcpyHostToDevice,
                              s) != hipSuccess)
        rc = setError(HIPIBS_E_HIP, "LUT upload failed");
    auto bucketOfRecord = [&](uint64_t r)
    {
        const uint64_t u = (uint64_t)(std::upper_bound(src.lut.begin(), src.lut.begin() + nBuckets,
                                                       r) -
                                      src.lut.begin());
        return u ? u - 1 : 0;
    };
    bool used[2] = {false, false};
    for (uint64_t r0 = 0, it = 0; r0 < total && !rc; r0 += chunk, ++it)
    {
        const int b = (int)(it & 1);
        const uint64_t m = std::min(chunk, total - r0);
        if (used[b] && hipEventSynchronize(copied[b]) != hipSuccess)
        {
            rc = setError(HIPIBS_E_HIP, "event wait failed");
            break;
        }
        rc = kmcReadRecords(src, r0, m, pinned[b], threads);
        if (rc)
            break;
        if (hipMemcpyAsync(raw[b], pinned[b], m * src.recBytes, hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipEventRecord(copied[b], s) != hipSuccess)
        {
            rc = setError(HIPIBS_E_HIP, "record upload failed");
            break;
        }
        used[b] = true;
        const uint64_t j0 = bucketOfRecord(r0), j1 = bucketOfRecord(r0 + m - 1);
        const int grid = gridFor((m + kDecodePerThread - 1) / kDecodePerThread, 256, device);
        if (wide)
            hipLaunchKernelGGL(kmcDecodeWideKernel, dim3(grid), dim3(256), 0, s, raw[b], r0, m,
                               src.recBytes, src.sufBytes, dLut, j0, j1, prefixMask, shift, dKeys);
        else
            hipLaunchKernelGGL(kmcDecodeKernel, dim3(grid), dim3(256), 0, s, raw[b], r0, m,
                               src.recBytes, src.sufBytes, dLut, j0, j1, prefixMask, shift, dKeys);
        if (hipGetLastError() != hipSuccess)
        {
            rc = setError(HIPIBS_E_HIP, "KMC decode launch failed");
            break;
        }
        rc = consume(dKeys, m);
    }
    // the staging buffers may only go once the stream has used them
    if (hipStreamSynchronize(s) != hipSuccess && !rc)
        rc = setError(HIPIBS_E_HIP, "KMC streaming failed on the device");
    for (int i = 0; i < 2; ++i)
    {
        if (pinned[i])
            hipHostFree(pinned[i]);
        if (raw[i])
            hipFree(raw[i]);
        if (copied[i])
            hipEventDestroy(copied[i]);
    }
    if (dKeys)
        hipFree(dKeys);
    if (dLut)
        hipFree(dLut);
    return rc;
}

static int hipibsBuildFromKmc(hipibs_table *t, const KmcRecords &src, unsigned threads)
{
    hipStream_t s;
    HIPIBS_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc = streamKmcKeys(src, t->device, threads, s, [&](const uint64_t *dKeys, uint64_t m)
                           { return launchTableInsert(t, dKeys, m, s); });
    if (!rc)
        rc = finishTable(t, s);
    hipStreamDestroy(s);
    return rc;
}

extern "C" int hipibs_table_build_keys(const uint64_t *h_keys, uint64_t n, uint32_t k,
                                       const hipibs_table_opts *opts, hipibs_table **out)
{
    if (!out || (n && !h_keys))
        return setError(HIPIBS_E_ARG, "null argument");
    *out = nullptr;
    if (k > 32 && k <= 64)
    {
        // two words per key: one upload, then the wide insert
        HIPIBS_CHECK(hipSetDevice(opts ? opts->device : 0));
        uint64_t *dAll = nullptr;
        if (hipMalloc(reinterpret_cast<void **>(&dAll), std::max<uint64_t>(2 * n, 1) * 8) !=
            hipSuccess)
            return setError(HIPIBS_E_NOMEM, "device key buffer allocation failed");
        int rc = hipMemcpy(dAll, h_keys, 2 * n * 8, hipMemcpyHostToDevice) == hipSuccess
                     ? 0
                     : setError(HIPIBS_E_HIP, "key upload failed");
        if (!rc)
            rc = hipibs_table_build_keys_device(dAll, n, k, opts, out, nullptr);
        (void)hipDeviceSynchronize();
        hipFree(dAll);
        return rc;
    }
    if (opts && opts->minimizer_len)
    {
        // the group counts need every key: upload them all first
        HIPIBS_CHECK(hipSetDevice(opts->device));
        uint64_t *dAll = nullptr;
        if (hipMalloc(reinterpret_cast<void **>(&dAll), std::max<uint64_t>(n, 1) * 8) != hipSuccess)
            return setError(HIPIBS_E_NOMEM, "device key buffer allocation failed");
        int rc = hipMemcpy(dAll, h_keys, n * 8, hipMemcpyHostToDevice) == hipSuccess
                     ? 0
                     : setError(HIPIBS_E_HIP, "key upload failed");
        if (!rc)
            rc = buildMinimizerTable(dAll, n, k, opts, nullptr, out);
        hipFree(dAll);
        return rc;
    }
    HostStaging *st = nullptr;
    int rc = hostStagingCreate(opts ? opts->device : 0, std::min<uint64_t>(std::max<uint64_t>(n, 1),
                                                                         32ull << 20), &st);
    if (rc)
        return rc;
    rc = buildTableFromHost(h_keys, n, k, opts, st, out);
    hostStagingFree(st);
    return rc;
}

extern "C" int hipibs_table_free(hipibs_table *t)
{
    freeTableOn(t, nullptr);
    return 0;
}

extern "C" int hipibs_table_info(const hipibs_table *t, uint64_t *n_keys, uint64_t *bytes,
                                 uint64_t *n_buckets, uint32_t *k, uint32_t *bucket_slots,
                                 int *device)
{
    if (!t)
        return setError(HIPIBS_E_ARG, "null table");
    if (n_keys)
        *n_keys = t->nKeys;
    if (bytes)
        *bytes = t->bytes;
    if (n_buckets)
        *n_buckets = t->nBuckets;
    if (k)
        *k = t->k;
    if (bucket_slots)
        *bucket_slots = t->bucketSlots;
    if (device)
        *device = t->device;
    return 0;
}

extern "C" int hipibs_table_probe(const hipibs_table *t, const uint64_t *d_codes,
                                  const uint8_t *d_valid, uint64_t n, uint8_t *d_hits,
                                  void *stream)
{
    if (!t || (n && (!d_codes || !d_hits)))
        return setError(HIPIBS_E_ARG, "null argument");
    return launchProbe(t, d_codes, d_valid, n, d_hits, nullptr,
                       reinterpret_cast<hipStream_t>(stream));
}

extern "C" int hipibs_table_probe_count_device(const hipibs_table *t, const uint64_t *d_keys,
                                               uint64_t n, uint64_t *d_count, void *stream)
{
    if (!t || !d_count || (n && !d_keys))
        return setError(HIPIBS_E_ARG, "null argument");
    return launchProbe(t, d_keys, nullptr, n, nullptr, d_count,
                       reinterpret_cast<hipStream_t>(stream));
}

// Host keys: pinned double-buffered chunks, one private stream per call, a
// mutex per table so concurrent callers (the intersect's consumer threads)
// serialise on the device as DeviceKmerIndex::probeCountHost does.
extern "C" int hipibs_table_probe_count(const hipibs_table *t, const uint64_t *h_keys, uint64_t n,
                                        uint64_t *out)
{
    if (!t || !out || (n && !h_keys))
        return setError(HIPIBS_E_ARG, "null argument");
    *out = 0;
    if (n == 0)
        return 0;
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    HIPIBS_CHECK(hipSetDevice(t->device));
    const uint64_t chunk = std::min<uint64_t>(n, 16ull << 20);
    const uint64_t kw = t->k > 32 ? 2 : 1; // words per key
    hipStream_t s;
    HIPIBS_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    uint64_t *pinned[2] = {nullptr, nullptr};
    uint64_t *dev[2] = {nullptr, nullptr};
    uint64_t *dCount = nullptr;
    hipEvent_t done[2] = {nullptr, nullptr};
    int rc = 0;
    for (int i = 0; i < 2 && !rc; ++i)
        if (hipHostMalloc(reinterpret_cast<void **>(&pinned[i]), chunk * 8 * kw) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&dev[i]), chunk * 8 * kw) != hipSuccess ||
            hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess)
            rc = setError(HIPIBS_E_NOMEM, "staging allocation failed");
    if (!rc && hipMalloc(reinterpret_cast<void **>(&dCount), 8) != hipSuccess)
        rc = setError(HIPIBS_E_NOMEM, "count allocation failed");
    if (!rc)
    {
        hipLaunchKernelGGL(fillKernel, dim3(1), dim3(64), 0, s, dCount, (uint64_t)1, 0ull);
        if (hipGetLastError() != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "count clear failed");
    }
    bool used[2] = {false, false};
    for (uint64_t off = 0, it = 0; off < n && !rc; off += chunk, ++it)
    {
        const int b = (int)(it & 1);
        const uint64_t m = std::min(chunk, n - off);
        if (used[b] && hipEventSynchronize(done[b]) != hipSuccess)
        {
            rc = setError(HIPIBS_E_HIP, "event wait failed");
            break;
        }
        parallelCopy(pinned[b], h_keys + off * kw, m * 8 * kw);
        if (hipMemcpyAsync(dev[b], pinned[b], m * 8 * kw, hipMemcpyHostToDevice, s) != hipSuccess)
        {
            rc = setError(HIPIBS_E_HIP, "key upload failed");
            break;
        }
        rc = launchProbe(t, dev[b], nullptr, m, nullptr, dCount, s);
        if (!rc && hipEventRecord(done[b], s) != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "event record failed");
        used[b] = true;
    }
    uint64_t c = 0;
    if (!rc && hipMemcpyAsync(&c, dCount, 8, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = setError(HIPIBS_E_HIP, "count download failed");
    if (hipStreamSynchronize(s) != hipSuccess && !rc)
        rc = setError(HIPIBS_E_HIP, "stream synchronize failed");
    for (int i = 0; i < 2; ++i)
    {
        if (pinned[i])
            hipHostFree(pinned[i]);
        if (dev[i])
            hipFree(dev[i]);
        if (done[i])
            hipEventDestroy(done[i]);
    }
    if (dCount)
        hipFree(dCount);
    hipStreamDestroy(s);
    if (!rc)
        *out = c;
    return rc;
}

extern "C" int hipibs_table_build_kmc(const char *prefix, const hipibs_table_opts *opts,
                                      hipibs_table **out)
{
    if (!prefix || !out)
        return setError(HIPIBS_E_ARG, "null argument");
    *out = nullptr;
    KmcRecords src;
    int rc = kmcOpenRecords(prefix, src);
    if (rc)
        return rc;
    const uint64_t n = src.total;
    hipibs_table *t = nullptr;
    if (opts && opts->minimizer_len && src.info.kmer_length > 32)
        return setError(HIPIBS_E_ARG, "minimizer tables need k <= 32");
    if (opts && opts->minimizer_len)
    {
        // the group counts need every key: decode them all into one buffer
        HIPIBS_CHECK(hipSetDevice(opts->device));
        uint64_t *dAll = nullptr;
        if (hipMalloc(reinterpret_cast<void **>(&dAll), std::max<uint64_t>(n, 1) * 8) != hipSuccess)
            return setError(HIPIBS_E_NOMEM, "device key buffer allocation failed");
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        {
            hipFree(dAll);
            return setError(HIPIBS_E_HIP, "stream creation failed");
        }
        // the groups are counted chunk by chunk while the records are read,
        // on a stream of their own so that the next chunk's upload does not
        // wait for them; the placement needs them all
        MinimizerBuild mb;
        hipStream_t cs = nullptr;
        hipEvent_t copied = nullptr;
        rc = mb.begin(n, src.info.kmer_length, opts, s);
        if (!rc && (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess ||
                    hipEventCreateWithFlags(&copied, hipEventDisableTiming) != hipSuccess))
            rc = setError(HIPIBS_E_HIP, "stream creation failed");
        uint64_t off = 0;
        if (!rc)
            rc = streamKmcKeys(src, opts->device, opts->load_threads, s,
                               [&](const uint64_t *dKeys, uint64_t m)
                               {
                                   if (hipMemcpyAsync(dAll + off, dKeys, m * 8,
                                                      hipMemcpyDeviceToDevice, s) != hipSuccess ||
                                       hipEventRecord(copied, s) != hipSuccess ||
                                       hipStreamWaitEvent(cs, copied, 0) != hipSuccess)
                                       return setError(HIPIBS_E_HIP, "key copy failed");
                                   off += m;
                                   return mb.count(dAll + off - m, m, cs);
                               });
        // the placement (on s) after every count
        if (!rc && (hipEventRecord(copied, cs) != hipSuccess ||
                    hipStreamWaitEvent(s, copied, 0) != hipSuccess))
            rc = setError(HIPIBS_E_HIP, "stream ordering failed");
        if (!rc)
            rc = mb.finish(dAll, n, s, &t);
        if (cs)
        {
            (void)hipStreamSynchronize(cs);
            (void)hipStreamDestroy(cs);
        }
        if (copied)
            (void)hipEventDestroy(copied);
        (void)hipStreamSynchronize(s);
        mb.release();
        hipStreamDestroy(s);
        hipFree(dAll);
    }
    else
    {
        rc = allocTable(n, src.info.kmer_length, opts, &t, nullptr);
        if (rc)
            return rc;
        rc = hipStreamSynchronize(nullptr) == hipSuccess ? 0
                                                         : setError(HIPIBS_E_HIP, "table clear failed");
        if (!rc)
            rc = hipibsBuildFromKmc(t, src, opts ? opts->load_threads : 0);
    }
    if (!rc && t->nKeys != n)
        rc = setError(HIPIBS_E_IO, "KMC database holds duplicate k-mers (" +
                                       std::to_string(n) + " records, " +
                                       std::to_string(t->nKeys) + " distinct)");
    if (rc)
    {
        hipibs_table_free(t);
        return rc;
    }
    *out = t;
    return 0;
}

// Keys of a KMC database found in the table, the records streamed and
// decoded on the device (no host key array): the counting half of
// KDBIntersect.
extern "C" int hipibs_table_probe_count_kmc(const hipibs_table *t, const char *prefix,
                                            unsigned load_threads, uint64_t *out)
{
    if (!t || !prefix || !out)
        return setError(HIPIBS_E_ARG, "null argument");
    *out = 0;
    KmcRecords src;
    int rc = kmcOpenRecords(prefix, src);
    if (rc)
        return rc;
    if (src.info.kmer_length != t->k)
        return setError(HIPIBS_E_ARG, "the KMC database and the table have different k");
    HIPIBS_CHECK(hipSetDevice(t->device));
    hipStream_t s;
    HIPIBS_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    uint64_t *dCount = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&dCount), 8) != hipSuccess)
        rc = setError(HIPIBS_E_NOMEM, "count allocation failed");
    if (!rc)
    {
        hipLaunchKernelGGL(fillKernel, dim3(1), dim3(64), 0, s, dCount, (uint64_t)1, 0ull);
        if (hipGetLastError() != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "count clear failed");
    }
    if (!rc)
        rc = streamKmcKeys(src, t->device, load_threads, s, [&](const uint64_t *dKeys, uint64_t m)
                           { return launchProbe(t, dKeys, nullptr, m, nullptr, dCount, s); });
    uint64_t c = 0;
    if (!rc && (hipMemcpyAsync(&c, dCount, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        rc = setError(HIPIBS_E_HIP, "count readback failed");
    if (dCount)
        hipFree(dCount);
    hipStreamDestroy(s);
    if (!rc)
        *out = c;
    return rc;
}

extern "C" int hipibs_table_minimizer_len(const hipibs_table *t, uint32_t *m)
{
    if (!t || !m)
        return setError(HIPIBS_E_ARG, "null argument");
    *m = t->mlen;
    return 0;
}
