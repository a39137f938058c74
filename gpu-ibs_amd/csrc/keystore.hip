// Compact host copies of sample keys for the serving engine, and table
// builds from them.
//
// A served sample's table is rebuilt whenever the sample is admitted after
// an eviction, so the engine keeps the sample's keys in host memory. Eight
// (or sixteen) bytes per key over PCIe, staged through a host memcpy, made
// admissions slower than the requests that wait for them. The copy kept
// here holds each 2k-bit code as a prefix of at most 22 bits, implied by
// the key's position, and a suffix of sufBytes whole bytes: keys are stored
// grouped by prefix (order inside a group is arbitrary), offs[p] = index of
// the first key of prefix p, offs[2^P] = n. At k = 25 that is 4 B per key
// plus a 1 MiB offset array. The copy lives in pinned host memory where the
// system grants it, so an admission uploads it by DMA directly.
//
// Registration (keyStoreFromDevice) is a counting sort by prefix on the
// device: a histogram, an exclusive scan on the host, an atomic scatter of
// the suffixes. An admission (buildTableFromStore) uploads the suffixes in
// chunks on one stream while another decodes each chunk (a binary search
// of the offsets gives a key's prefix) and inserts it into the table.
#include "hipibs_internal.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace hipibs
{

void keyStoreGeometry(uint32_t k, uint32_t &prefixBits, uint32_t &sufBytes)
{
    const uint32_t bits = 2 * k;
    sufBytes = bits > 22 ? (bits - 22 + 7) / 8 : 0;
    prefixBits = bits - 8 * sufBytes;
}

namespace
{

constexpr uint64_t kStoreChunk = 16ull << 20; // keys per uploaded chunk

int cusOf(int device)
{
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    return cus;
}

uint32_t gridOf(uint64_t items, int device)
{
    const uint64_t want = (items + 255) / 256;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)cusOf(device) * 16));
}

// The code of key i: lo = low 64 bits, hi = the rest (k > 32).
__device__ inline void loadCode(const uint64_t *keys, uint64_t i, bool wide, uint64_t &hi,
                                uint64_t &lo)
{
    if (wide)
    {
        hi = keys[2 * i];
        lo = keys[2 * i + 1];
    }
    else
    {
        hi = 0;
        lo = keys[i];
    }
}

// code >> S, for S = 8 * sufBytes < 128 (the prefix: at most 22 bits when
// the code has 2k bits).
__device__ inline uint64_t prefixOf(uint64_t hi, uint64_t lo, uint32_t S)
{
    if (S >= 64)
        return hi >> (S - 64);
    return S ? ((hi << (64 - S)) | (lo >> S)) : lo;
}

__global__ __launch_bounds__(256) void storeZeroKernel(uint32_t *p, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = 0;
}

// Keys per prefix. A code wider than its 2k bits (a caller's mistake)
// sets *bad and is not counted: its prefix is >= nPrefixes, or (two-word
// keys) its high word has bits above 2k - 64.
__global__ __launch_bounds__(256) void storeHistKernel(const uint64_t *__restrict__ keys, uint64_t n,
                                                      bool wide, uint32_t kbits, uint32_t S,
                                                      uint64_t nPrefixes,
                                                      uint32_t *__restrict__ cnt, uint32_t *bad)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        uint64_t hi, lo;
        loadCode(keys, i, wide, hi, lo);
        const uint64_t p = prefixOf(hi, lo, S);
        const bool tooWide = wide && kbits < 128 && (hi >> (kbits - 64)) != 0;
        if (p >= nPrefixes || tooWide)
            atomicOr(bad, 1u);
        else
            atomicAdd(cnt + p, 1u);
    }
}

// Each key's suffix (its low sufBytes bytes, little-endian) at the next
// free index of its prefix.
__global__ __launch_bounds__(256) void storeScatterKernel(const uint64_t *__restrict__ keys,
                                                         uint64_t n, bool wide, uint32_t S,
                                                         uint32_t sufBytes, uint64_t nPrefixes,
                                                         uint32_t *__restrict__ cursor,
                                                         uint8_t *__restrict__ suf)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        uint64_t hi, lo;
        loadCode(keys, i, wide, hi, lo);
        const uint64_t p = prefixOf(hi, lo, S);
        if (p >= nPrefixes)
            continue;
        const uint64_t at = (uint64_t)atomicAdd(cursor + p, 1u) * sufBytes;
        if (sufBytes == 4)
            *reinterpret_cast<uint32_t *>(suf + at) = (uint32_t)lo;
        else
            for (uint32_t b = 0; b < sufBytes; ++b)
                suf[at + b] = (uint8_t)(b < 8 ? lo >> (8 * b) : hi >> (8 * (b - 8)));
    }
}

// Keys i0 .. i0+m-1 from their suffixes (chunk-relative) and the offsets.
__global__ __launch_bounds__(256) void storeDecodeKernel(const uint32_t *__restrict__ offs,
                                                        uint64_t nPrefixes,
                                                        const uint8_t *__restrict__ suf,
                                                        uint32_t sufBytes, uint64_t i0,
                                                        uint64_t m, bool wide,
                                                        uint64_t *__restrict__ keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t S = 8 * sufBytes;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride)
    {
        const uint64_t i = i0 + j;
        // the prefix: the last p with offs[p] <= i (offs[0] = 0 <= i < offs[nPrefixes])
        uint64_t lo = 0, hi = nPrefixes;
        while (hi - lo > 1)
        {
            const uint64_t mid = (lo + hi) >> 1;
            if ((uint64_t)offs[mid] <= i)
                lo = mid;
            else
                hi = mid;
        }
        const uint64_t p = lo;
        uint64_t sLo = 0, sHi = 0;
        const uint8_t *q = suf + j * sufBytes;
        if (sufBytes == 4)
            sLo = *reinterpret_cast<const uint32_t *>(q);
        else
            for (uint32_t b = 0; b < sufBytes; ++b)
            {
                if (b < 8)
                    sLo |= (uint64_t)q[b] << (8 * b);
                else
                    sHi |= (uint64_t)q[b] << (8 * (b - 8));
            }
        uint64_t cHi, cLo;
        if (S >= 64)
        {
            cHi = (S > 64 ? (p << (S - 64)) : p) | sHi;
            cLo = sLo;
        }
        else
        {
            cHi = S ? (p >> (64 - S)) : 0;
            cLo = (S ? (p << S) : p) | sLo;
        }
        if (wide)
        {
            keys[2 * j] = cHi;
            keys[2 * j + 1] = cLo;
        }
        else
            keys[j] = cLo;
    }
}

} // namespace

void keyStoreFree(KeyStore &ks)
{
    if (ks.pinned)
    {
        (void)hipHostFree(ks.suf);
        (void)hipHostFree(ks.offs);
    }
    else
    {
        std::free(ks.suf);
        std::free(ks.offs);
    }
    ks.suf = nullptr;
    ks.offs = nullptr;
    ks.n = 0;
}

static int storeHostAlloc(KeyStore &ks, uint64_t sufBytesTotal, uint64_t offBytes)
{
    // pinned where granted (uploads by DMA), pageable otherwise
    if (hipHostMalloc(reinterpret_cast<void **>(&ks.suf), std::max<uint64_t>(sufBytesTotal, 1)) ==
            hipSuccess &&
        hipHostMalloc(reinterpret_cast<void **>(&ks.offs), offBytes) == hipSuccess)
    {
        ks.pinned = true;
        return 0;
    }
    if (ks.suf)
        (void)hipHostFree(ks.suf);
    ks.suf = nullptr;
    (void)hipGetLastError();
    ks.pinned = false;
    ks.suf = static_cast<uint8_t *>(std::malloc(std::max<uint64_t>(sufBytesTotal, 1)));
    ks.offs = static_cast<uint32_t *>(std::malloc(offBytes));
    if (!ks.suf || !ks.offs)
    {
        std::free(ks.suf);
        std::free(ks.offs);
        ks.suf = nullptr;
        ks.offs = nullptr;
        return setError(HIPIBS_E_NOMEM, "host memory for a sample's key copy");
    }
    return 0;
}

int keyStoreFromDevice(const uint64_t *dKeys, uint64_t n, uint32_t k, int device, hipStream_t s,
                       KeyStore &out)
{
    if (k == 0 || k > 64)
        return setError(HIPIBS_E_ARG, "k must be in 1..64");
    if (n >= (1ull << 32))
        return setError(HIPIBS_E_ARG, "a served sample holds fewer than 2^32 keys");
    if (n && !dKeys)
        return setError(HIPIBS_E_ARG, "null keys");
    KeyStore ks;
    ks.k = k;
    ks.n = n;
    keyStoreGeometry(k, ks.prefixBits, ks.sufBytes);
    const uint64_t nPrefixes = 1ull << ks.prefixBits;
    const bool wide = k > 32;
    const uint32_t S = 8 * ks.sufBytes;
    int rc = storeHostAlloc(ks, n * ks.sufBytes, (nPrefixes + 1) * 4);
    if (rc)
        return rc;
    HIPIBS_CHECK(hipSetDevice(device));
    uint32_t *dCnt = nullptr;
    uint8_t *dSuf = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&dCnt), (nPrefixes + 1) * 4) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&dSuf), std::max<uint64_t>(n * ks.sufBytes, 1)) !=
            hipSuccess)
        rc = setError(HIPIBS_E_NOMEM, "device memory for a sample's key copy");
    uint32_t bad = 0;
    std::vector<uint32_t> cnt;
    if (!rc)
    {
        hipLaunchKernelGGL(storeZeroKernel, dim3(gridOf(nPrefixes + 1, device)), dim3(256), 0, s,
                           dCnt, nPrefixes + 1);
        if (n)
            hipLaunchKernelGGL(storeHistKernel, dim3(gridOf(n, device)), dim3(256), 0, s, dKeys, n,
                               wide, 2 * k, S, nPrefixes, dCnt, dCnt + nPrefixes);
        cnt.resize(nPrefixes + 1);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(cnt.data(), dCnt, (nPrefixes + 1) * 4, hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "key histogram failed");
    }
    if (!rc)
    {
        bad = cnt[nPrefixes];
        if (bad)
            rc = setError(HIPIBS_E_ARG, "a key has bits above its 2k-bit code");
    }
    if (!rc)
    {
        uint64_t acc = 0;
        for (uint64_t p = 0; p < nPrefixes; ++p)
        {
            ks.offs[p] = (uint32_t)acc;
            acc += cnt[p];
        }
        ks.offs[nPrefixes] = (uint32_t)acc;
        if (acc != n)
            rc = setError(HIPIBS_E_HIP, "key histogram lost keys");
    }
    if (!rc && n)
    {
        // the offsets become the scatter's cursors
        if (hipMemcpyAsync(dCnt, ks.offs, (nPrefixes + 1) * 4, hipMemcpyHostToDevice, s) !=
            hipSuccess)
            rc = setError(HIPIBS_E_HIP, "offset upload failed");
        if (!rc)
        {
            hipLaunchKernelGGL(storeScatterKernel, dim3(gridOf(n, device)), dim3(256), 0, s, dKeys,
                               n, wide, S, ks.sufBytes, nPrefixes, dCnt, dSuf);
            if (hipGetLastError() != hipSuccess ||
                (ks.sufBytes && hipMemcpyAsync(ks.suf, dSuf, n * ks.sufBytes,
                                               hipMemcpyDeviceToHost, s) != hipSuccess) ||
                hipStreamSynchronize(s) != hipSuccess)
                rc = setError(HIPIBS_E_HIP, "suffix scatter failed");
        }
    }
    (void)hipFree(dCnt);
    (void)hipFree(dSuf);
    if (rc)
    {
        keyStoreFree(ks);
        return rc;
    }
    out = ks;
    return 0;
}

struct StoreStaging
{
    int device = 0;
    hipStream_t copy = nullptr, compute = nullptr;
    hipEvent_t copied[2] = {nullptr, nullptr}, consumed[2] = {nullptr, nullptr};
    uint8_t *dSuf[2] = {nullptr, nullptr};
    uint64_t *dKeys[2] = {nullptr, nullptr};
    uint32_t *dOffs = nullptr;
    uint64_t sufCap = 0, keyCap = 0, offCap = 0; // bytes
};

void storeStagingFree(StoreStaging *st)
{
    if (!st)
        return;
    (void)hipSetDevice(st->device);
    if (st->copy)
        (void)hipStreamSynchronize(st->copy);
    if (st->compute)
        (void)hipStreamSynchronize(st->compute);
    for (int i = 0; i < 2; ++i)
    {
        (void)hipFree(st->dSuf[i]);
        (void)hipFree(st->dKeys[i]);
        if (st->copied[i])
            (void)hipEventDestroy(st->copied[i]);
        if (st->consumed[i])
            (void)hipEventDestroy(st->consumed[i]);
    }
    (void)hipFree(st->dOffs);
    if (st->copy)
        (void)hipStreamDestroy(st->copy);
    if (st->compute)
        (void)hipStreamDestroy(st->compute);
    delete st;
}

int storeStagingCreate(int device, StoreStaging **out)
{
    *out = nullptr;
    HIPIBS_CHECK(hipSetDevice(device));
    auto *st = new StoreStaging();
    st->device = device;
    bool ok = hipStreamCreateWithFlags(&st->copy, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&st->compute, hipStreamNonBlocking) == hipSuccess;
    for (int i = 0; i < 2 && ok; ++i)
        ok = hipEventCreateWithFlags(&st->copied[i], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&st->consumed[i], hipEventDisableTiming) == hipSuccess;
    if (!ok)
    {
        storeStagingFree(st);
        return setError(HIPIBS_E_HIP, "staging stream/event creation failed");
    }
    *out = st;
    return 0;
}

// Makes both buffers of a pair at least `want` bytes (contents not kept).
template <class T>
static int growPair(T *(&p)[2], uint64_t &cap, uint64_t want)
{
    if (cap >= want)
        return 0;
    for (int i = 0; i < 2; ++i)
    {
        (void)hipFree(p[i]);
        p[i] = nullptr;
    }
    cap = 0;
    for (int i = 0; i < 2; ++i)
        if (hipMalloc(reinterpret_cast<void **>(&p[i]), want) != hipSuccess)
            return setError(HIPIBS_E_NOMEM, "staging allocation of " + std::to_string(want) +
                                                " bytes failed");
    cap = want;
    return 0;
}

int buildTableFromStore(const KeyStore &ks, const hipibs_table_opts *opts, StoreStaging *st,
                        hipibs_table **out, bool streamOrdered)
{
    *out = nullptr;
    HIPIBS_CHECK(hipSetDevice(st->device));
    const uint64_t n = ks.n;
    const uint64_t nPrefixes = 1ull << ks.prefixBits;
    const bool wide = ks.k > 32;
    const uint64_t keyBytes = wide ? 16 : 8;
    const bool whole = !(opts && opts->minimizer_len); // minimizer tables need every key at once
    const uint64_t chunk = std::min<uint64_t>(std::max<uint64_t>(n, 1), kStoreChunk);
    int rc = 0;
    if (st->offCap < (nPrefixes + 1) * 4)
    {
        (void)hipFree(st->dOffs);
        st->dOffs = nullptr;
        st->offCap = 0;
        if (hipMalloc(reinterpret_cast<void **>(&st->dOffs), (nPrefixes + 1) * 4) != hipSuccess)
            rc = setError(HIPIBS_E_NOMEM, "staging allocation failed");
        else
            st->offCap = (nPrefixes + 1) * 4;
    }
    if (!rc)
        rc = growPair(st->dSuf, st->sufCap, std::max<uint64_t>(chunk * ks.sufBytes, 4));
    if (!rc && whole)
        rc = growPair(st->dKeys, st->keyCap, chunk * keyBytes);
    uint64_t *dAll = nullptr; // minimizer tables: every key, held for this build only
    if (!rc && !whole &&
        hipMalloc(reinterpret_cast<void **>(&dAll), std::max<uint64_t>(n, 1) * keyBytes) !=
            hipSuccess)
        rc = setError(HIPIBS_E_NOMEM, "device memory for the sample's keys");
    hipibs_table *t = nullptr;
    if (!rc && whole)
        rc = allocTable(n, ks.k, opts, &t, st->compute, streamOrdered);
    if (!rc && hipMemcpyAsync(st->dOffs, ks.offs, (nPrefixes + 1) * 4, hipMemcpyHostToDevice,
                              st->copy) != hipSuccess)
        rc = setError(HIPIBS_E_HIP, "offset upload failed");
    // copy stream: suffixes of chunk c into dSuf[c & 1] once the decode of
    // chunk c-2 has consumed it; compute stream: decode + insert chunk c
    // once it has arrived
    bool used[2] = {false, false};
    for (uint64_t i0 = 0, c = 0; i0 < n && !rc; i0 += chunk, ++c)
    {
        const int b = (int)(c & 1);
        const uint64_t m = std::min(chunk, n - i0);
        if (used[b] && hipStreamWaitEvent(st->copy, st->consumed[b], 0) != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "stream wait failed");
        if (!rc && ks.sufBytes &&
            hipMemcpyAsync(st->dSuf[b], ks.suf + i0 * ks.sufBytes, m * ks.sufBytes,
                           hipMemcpyHostToDevice, st->copy) != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "suffix upload failed");
        if (!rc && (hipEventRecord(st->copied[b], st->copy) != hipSuccess ||
                    hipStreamWaitEvent(st->compute, st->copied[b], 0) != hipSuccess))
            rc = setError(HIPIBS_E_HIP, "event record failed");
        if (rc)
            break;
        uint64_t *dst = whole ? st->dKeys[b] : dAll + (wide ? 2 * i0 : i0);
        hipLaunchKernelGGL(storeDecodeKernel, dim3(gridOf(m, st->device)), dim3(256), 0,
                           st->compute, st->dOffs, nPrefixes, st->dSuf[b], ks.sufBytes, i0, m,
                           wide, dst);
        if (hipGetLastError() != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "key decode launch failed");
        if (!rc && whole)
            rc = launchTableInsert(t, dst, m, st->compute);
        if (!rc && hipEventRecord(st->consumed[b], st->compute) != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "event record failed");
        used[b] = true;
    }
    if (!rc && whole)
        rc = finishTable(t, st->compute); // synchronises the compute stream
    else if (!rc)
        rc = buildKeysDevice(dAll, n, ks.k, opts, &t, st->compute, false);
    (void)hipStreamSynchronize(st->copy);
    (void)hipStreamSynchronize(st->compute);
    if (dAll)
        (void)hipFree(dAll);
    if (rc)
    {
        freeTableOn(t, st->compute);
        return rc;
    }
    *out = t;
    return 0;
}

} // namespace hipibs
