// Internal definitions shared by the HIP translation units of libhipibs.
//
// Table layout in HBM: n_buckets buckets of `S` 64-bit slots (S = 8: one
// 64 B bucket; S = 16: 128 B), 256 B aligned, EMPTY = all ones. A key
// lives in bucket home(key) or, when that bucket was full at insert time,
// in one of the next `max_disp` buckets (linear probing at bucket
// granularity, wrapping). A probe therefore stops at the first of: key
// found, an EMPTY slot in the bucket, or max_disp+1 buckets visited.
// Canonical codes of k <= 32 are never all ones (k < 32: < 2^62; k = 32:
// min(x, ~x) < 2^63), so EMPTY can never collide with a key. Keys of
// 33..64 bases are (high, low) word pairs whose free mark is an all-ones
// high word; the one canonical code with such a high word is k = 64's
// T^32 A^32 (high all ones, low 0; its reverse complement is itself), which
// a table records in a flag of its own (TableView::wideTop).
#pragma once

#include <hip/hip_runtime.h>

#include <sys/mman.h>

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "hipibs.h"
#include "reads.hpp"

namespace hipibs
{

constexpr uint64_t EMPTY = ~0ull;

// A k-mer of 33..64 bases as one host integer (device code keeps the
// (high, low) word pair).
typedef unsigned __int128 Wide;

// The mix behind bucketOf: multiply-xorshift steps, each invertible, so
// keyUnmix recovers the key.
constexpr uint64_t kMixA = 0x9E3779B97F4A7C15ull, kMixB = 0xBF58476D1CE4E5B9ull;
__host__ __device__ inline uint64_t keyMix(uint64_t key)
{
    uint64_t h = (key ^ (key >> 31)) * kMixA;
    h ^= h >> 29;
    return h * kMixB;
}

// Multiplicative inverse of an odd 64-bit constant (Newton's iteration,
// each step doubles the correct low bits: 3 -> 96).
constexpr uint64_t inverseOdd(uint64_t a)
{
    uint64_t x = a;
    for (int i = 0; i < 5; ++i)
        x *= 2 - a * x;
    return x;
}

__host__ __device__ inline uint64_t keyUnmix(uint64_t h)
{
    constexpr uint64_t invA = inverseOdd(kMixA), invB = inverseOdd(kMixB);
    uint64_t x = h * invB;
    x ^= (x >> 29) ^ (x >> 58); // inverse of x ^= x >> 29
    x *= invA;
    return x ^ (x >> 31) ^ (x >> 62); // inverse of x ^= x >> 31
}

// Bucket of a key: the high 32 bits of keyMix scaled onto [0, n_buckets)
// (n_buckets < 2^32).
__host__ __device__ inline uint64_t bucketOf(uint64_t key, uint64_t nBuckets)
{
    return ((keyMix(key) >> 32) * nBuckets) >> 32;
}

// Bucket of a 33..64-base key (high word, low word).
__host__ __device__ inline uint64_t bucketOfWide(uint64_t hi, uint64_t lo, uint64_t nBuckets)
{
    return bucketOf(lo ^ (hi * 0xD6E8FEB86659FD93ull) ^ (hi >> 17), nBuckets);
}

// Reverses the order of the 32 2-bit groups of x.
__host__ __device__ inline uint64_t reverse2(uint64_t x)
{
    const uint64_t v = __builtin_bitreverse64(x);
    return ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
}

// Minimizer bucketing. The home bucket of a k-mer may come from its
// minimizer: the smallest 32-bit hash over its k-m+1 canonical m-mers
// (canonical = min of the m-mer and its reverse complement, so a k-mer and
// its reverse complement share it). Consecutive reference positions share
// a minimizer for ~(k-m)/2 positions on average, so they probe the same
// 128 B bucket and the wave's requests for them merge.
__host__ __device__ inline uint32_t mmerHash(uint64_t canonMmer)
{
    return (uint32_t)((canonMmer * 0x9E3779B97F4A7C15ull) >> 32);
}

__host__ __device__ inline uint32_t minimizerOf(uint64_t key, uint32_t k, uint32_t m)
{
    const uint64_t maskm = (1ull << (2 * m)) - 1ull;
    uint32_t best = 0xffffffffu;
    for (uint32_t i = 0; i + m <= k; ++i)
    {
        const uint64_t mm = (key >> (2 * (k - m - i))) & maskm;
        const uint64_t rc = reverse2(~mm & maskm) >> (64 - 2 * m);
        const uint32_t h = mmerHash(mm < rc ? mm : rc);
        best = h < best ? h : best;
    }
    return best;
}

// The minimum of hashes is skewed toward small values: re-mix it (the
// 32-bit finalizer of MurmurHash3, a bijection) before scaling onto the
// buckets (n_buckets < 2^32).
__host__ __device__ inline uint64_t bucketOfMinimizer(uint32_t minHash, uint64_t nBuckets)
{
    uint32_t x = minHash;
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return ((uint64_t)x * (uint32_t)nBuckets) >> 32;
}

// Device view of one table, passed by value to kernels. A minimizer table
// (mlen > 0) has two parts: minimizer-ordered buckets holding the keys of
// every minimizer group of at most `cap` keys plus one marker per larger
// group, and a whole-k-mer table (8-slot buckets) holding the larger
// groups' keys.
struct TableView
{
    const uint64_t *slots;
    uint64_t nBuckets;
    uint32_t maxDisp; // largest bucket displacement of any stored key
    uint32_t mlen;    // minimizer length of the home bucket, 0 = whole k-mer
    uint32_t k;
    uint32_t maxDisp2;
    const uint64_t *slots2; // the whole-k-mer part of a minimizer table
    uint64_t nBuckets2;
    uint32_t wideTop; // k = 64: 1 if the key T^32 A^32 (high word EMPTY, low 0) is present
    uint32_t pad;
};

// k = 64's T^32 A^32 as a (high, low) pair: the only canonical code whose
// high word is EMPTY.
__host__ __device__ inline bool isWideTop(uint64_t hi, uint64_t lo)
{
    return hi == EMPTY && lo == 0;
}

// The marker of a minimizer group (by its 32-bit minimizer hash) that keeps
// its keys in the whole-k-mer part: first and last base T, so it is never a
// canonical code for k = 32 (its reverse complement starts with A) and above
// every code for k < 32; never EMPTY.
__host__ __device__ inline uint64_t markerOf(uint32_t minHash)
{
    return 0xC000000000000003ull | ((uint64_t)minHash << 2);
}

__host__ __device__ inline uint64_t homeOf(uint64_t key, const TableView &T)
{
    return T.mlen ? bucketOfMinimizer(minimizerOf(key, T.k, T.mlen), T.nBuckets)
                  : bucketOf(key, T.nBuckets);
}

// Run-length summary of a contiguous range of positions of one window
// against one table (the monoid of the scan's reduction; see combine()).
// With nHit == 0 the range is one run of misses: lead == trail == nValid.
struct SegSummary
{
    uint32_t nValid;
    uint32_t nHit;
    uint32_t lead;  // valid misses before the first hit
    uint32_t trail; // valid misses after the last hit
    uint32_t innerVar;
    uint32_t innerDist;
};

// Window::addVariationOriginal's distance term for one gap (gap >= 1):
// d = gap - (k-1); if d <= 0 then d = |d + 1| (src/Window.hpp:29-36).
__host__ __device__ inline int32_t gapDistance(int32_t gap, int32_t k)
{
    int32_t d = gap - (k - 1);
    if (d <= 0)
    {
        d = d + 1;
        d = d < 0 ? -d : d;
    }
    return d;
}

__host__ __device__ inline SegSummary emptySummary()
{
    return SegSummary{0, 0, 0, 0, 0, 0};
}

// Associative combine of the summaries of two adjacent ranges A then B.
__host__ __device__ inline SegSummary combine(const SegSummary &A, const SegSummary &B, int32_t k)
{
    SegSummary r;
    r.nValid = A.nValid + B.nValid;
    r.nHit = A.nHit + B.nHit;
    if (A.nHit == 0 && B.nHit == 0)
    {
        r.lead = r.trail = r.nValid;
        r.innerVar = r.innerDist = 0;
    }
    else if (A.nHit == 0)
    {
        r.lead = A.nValid + B.lead;
        r.trail = B.trail;
        r.innerVar = B.innerVar;
        r.innerDist = B.innerDist;
    }
    else if (B.nHit == 0)
    {
        r.lead = A.lead;
        r.trail = A.trail + B.nValid;
        r.innerVar = A.innerVar;
        r.innerDist = A.innerDist;
    }
    else
    {
        const uint32_t g = A.trail + B.lead;
        r.lead = A.lead;
        r.trail = B.trail;
        r.innerVar = A.innerVar + B.innerVar + (g > 0 ? 1u : 0u);
        r.innerDist = A.innerDist + B.innerDist + (g > 0 ? (uint32_t)gapDistance((int32_t)g, k) : 0u);
    }
    return r;
}

// Window statistics of a whole window's summary: every maximal miss run,
// whether closed by a hit or by the window end, adds one variation and its
// distance term (src/KmerDatabase.hpp:596-672).
__host__ __device__ inline hipibs_window_stats finalizeSummary(const SegSummary &S, int32_t k)
{
    hipibs_window_stats w;
    w.total_kmers = (int32_t)S.nValid;
    w.observed_kmers = (int32_t)S.nHit;
    if (S.nHit == 0)
    {
        w.variations = S.nValid > 0 ? 1 : 0;
        w.kmer_distance = S.nValid > 0 ? gapDistance((int32_t)S.nValid, k) : 0;
        return w;
    }
    int32_t var = (int32_t)S.innerVar, dist = (int32_t)S.innerDist;
    if (S.lead > 0)
    {
        var += 1;
        dist += gapDistance((int32_t)S.lead, k);
    }
    if (S.trail > 0)
    {
        var += 1;
        dist += gapDistance((int32_t)S.trail, k);
    }
    w.variations = var;
    w.kmer_distance = dist;
    return w;
}

// ---- error reporting (thread-local last error) ----------------------------
int setError(int code, const std::string &msg);
int hipFail(hipError_t e, const char *what);

#define HIPIBS_CHECK(call)                                   \
    do                                                       \
    {                                                        \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess)                                \
            return ::hipibs::hipFail(e_, #call);             \
    } while (0)

} // namespace hipibs

// Host handle of a device table.
struct hipibs_table
{
    int device = 0;
    uint32_t k = 0;
    uint32_t bucketSlots = 8;
    uint64_t nBuckets = 0;
    uint64_t nKeys = 0;
    uint64_t bytes = 0;
    uint32_t maxDisp = 0;
    uint32_t mlen = 0;
    uint64_t *dSlots = nullptr;
    uint64_t *dCounters = nullptr; // [0] inserted, [1] full flag, [2] max displacement,
                                   // minimizer tables: [3] keys in the whole-k-mer part,
                                   // [4] its max displacement, [5] markers;
                                   // k > 32: [3] 1 if T^32 A^32 was inserted
    // minimizer tables only: the whole-k-mer part and the group cap
    uint64_t *dSlots2 = nullptr;
    uint64_t nBuckets2 = 0;
    uint32_t maxDisp2 = 0;
    uint32_t mcap = 0;
    uint64_t nKeys2 = 0, nMarkers = 0;
    uint32_t wideTop = 0; // k = 64: T^32 A^32 is a key (held outside the buckets)
    // slots and counters from hipMallocAsync (the device's default memory
    // pool): freed with hipFreeAsync, so neither allocation nor free waits
    // for the device
    bool streamOrdered = false;
    hipibs::TableView view() const
    {
        return hipibs::TableView{dSlots,   nBuckets, maxDisp,   mlen,    k,
                                 maxDisp2, dSlots2,  nBuckets2, wideTop, 0u};
    }
};

struct HostStaging;

namespace hipibs
{
// Launch helpers implemented in the .hip files (all asynchronous).
int launchTableInsert(hipibs_table *t, const uint64_t *dKeys, uint64_t n, hipStream_t s);
// A whole-k-mer table (k <= 32) filled from device keys by sorting them by
// table slice (count.hip); *used = false when it did nothing (small inputs,
// or no room for its buffers): insert with launchTableInsert then.
int buildTableSorted(hipibs_table *t, const uint64_t *dKeys, uint64_t n, hipStream_t s,
                     bool *used);
// hipibs_table_build_keys_device; allowSorted = false keeps whole-k-mer
// builds on the key-by-key insert (its buffers are allocated and freed with
// calls that wait for the whole device, which a server's builds must not).
int buildKeysDevice(const uint64_t *d_keys, uint64_t n, uint32_t k, const hipibs_table_opts *opts,
                    hipibs_table **out, hipStream_t s, bool allowSorted);
// Frees the buffers the process keeps for its next counter or table build.
void releaseSortCaches();
int launchProbe(const hipibs_table *t, const uint64_t *dCodes, const uint8_t *dValid,
                uint64_t n, uint8_t *dHits, uint64_t *dCount, hipStream_t s);
int launchScan(const hipibs_table *const *tables, int nTables, const uint64_t *dSeq,
               const uint64_t *dExc, uint64_t nBases, const hipibs_window *dWindows,
               uint32_t nWindows, uint32_t maxWindowBases, hipibs_window_stats *dStats,
               uint32_t segsPerWindow, void *dWork, uint64_t workBytes, hipStream_t s,
               TableView *dViews, bool profile = true);
int uploadViews(const hipibs_table *const *tables, int nTables, hipStream_t s, TableView **out);
// grouped: every table the windows name is a minimizer table (probed with
// waveProbeGrouped). segs > 1 splits every window into that many blocks
// (their summaries in dWork, nWindows * segs of them, folded in order).
int launchScanMixed(int bucketSlots, uint32_t k, const TableView *dViews, const uint64_t *dSeq,
                    const uint64_t *dExc, uint64_t nBases, const hipibs_window *dWindows,
                    uint32_t nWindows, hipibs_window_stats *dStats, hipStream_t s,
                    bool grouped = false, uint32_t segs = 1, SegSummary *dWork = nullptr);
int checkTables(const hipibs_table *const *tables, int nTables);
uint32_t autoSegsPerWindow(uint32_t nWindows, int nTables, uint32_t maxPositions);
uint64_t scanWorkBytes(uint32_t nWindows, int nTables, uint32_t segsPerWindow);
int kmcReadKeys(const std::string &prefix, uint64_t *keys, uint64_t cap, uint64_t *nOut,
                unsigned threads, hipibs_kmc_info *infoOut);

// Table construction steps shared by the build paths: slots sized for
// nKeys and cleared on s (clearSlots = false leaves them for a builder that
// writes every bucket; clearTableSlots clears them later); after the
// inserts, the build counters read back (synchronises s) and checked.
int allocTable(uint64_t nKeys, uint32_t k, const hipibs_table_opts *opts, hipibs_table **out,
               hipStream_t s, bool streamOrdered = false, bool clearSlots = true);
int clearTableSlots(hipibs_table *t, hipStream_t s);
int finishTable(hipibs_table *t, hipStream_t s);
// Frees t; a stream-ordered table's memory goes back to the pool after the
// work already on s.
void freeTableOn(hipibs_table *t, hipStream_t s);

// A sample's keys kept compactly in host memory (the serving engine's copy
// for re-admissions): each 2k-bit code is split into a prefix of
// prefixBits (<= 22) bits and a suffix of sufBytes bytes; keys are stored
// grouped by prefix (any order inside a group), offs[p] = index of the
// first key of prefix p, offs[2^prefixBits] = n, suffixes little-endian,
// sufBytes each. Pinned host memory where granted.
struct KeyStore
{
    uint32_t k = 0, prefixBits = 0, sufBytes = 0;
    uint64_t n = 0;
    uint32_t *offs = nullptr;
    uint8_t *suf = nullptr;
    bool pinned = false;
    uint64_t hostBytes() const { return n * sufBytes + ((1ull << prefixBits) + 1) * 4; }
};
void keyStoreGeometry(uint32_t k, uint32_t &prefixBits, uint32_t &sufBytes);
// From n device keys (one word each, two for k > 32): a counting sort by
// prefix on s (device memory for n * sufBytes bytes and the offsets while
// it runs). Rejects codes wider than 2k bits and n >= 2^32.
int keyStoreFromDevice(const uint64_t *dKeys, uint64_t n, uint32_t k, int device, hipStream_t s,
                       KeyStore &out);
void keyStoreFree(KeyStore &ks);
// A builder's reusable device staging (two streams, double buffers of
// 16 Mi keys' suffixes and decoded keys, the offsets).
struct StoreStaging;
int storeStagingCreate(int device, StoreStaging **out);
void storeStagingFree(StoreStaging *st);
// A table from a key copy: suffix chunks uploaded on one stream while the
// other decodes and inserts the previous one (a minimizer table decodes
// every key first, into device memory held for the build).
int buildTableFromStore(const KeyStore &ks, const hipibs_table_opts *opts, StoreStaging *st,
                        hipibs_table **out, bool streamOrdered = false);

// Reusable staging for builds from host keys (pinned + device buffers of
// chunkKeys keys each, and a stream).
int hostStagingCreate(int device, uint64_t chunkKeys, HostStaging **out);
void hostStagingFree(HostStaging *st);
int buildTableFromHost(const uint64_t *hKeys, uint64_t n, uint32_t k, const hipibs_table_opts *opts,
                       HostStaging *st, hipibs_table **out);

// An allocator whose default construction leaves the element as it is:
// vectors sized for bulk reads (a KMC LUT is 1 GiB at p = 9) skip the
// zero fill and are first touched by the threads that read into them.
template <class T>
struct NoInitAlloc : std::allocator<T>
{
    template <class U>
    struct rebind
    {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) noexcept
    {
    }
    // Arrays of 64 MiB or more are 2 MiB aligned and asked for huge pages:
    // the threads that fill them from the page cache then take one fault
    // per 2 MiB instead of one per 4 KiB (a 1 GB KMC LUT reads about twice
    // as fast).
    static constexpr size_t kHuge = size_t(64) << 20;
    T *allocate(size_t n)
    {
        const size_t bytes = n * sizeof(T);
        if (bytes < kHuge)
            return std::allocator<T>::allocate(n);
        void *p = nullptr;
        if (posix_memalign(&p, size_t(2) << 20, bytes) != 0)
            throw std::bad_alloc();
        (void)madvise(p, bytes, MADV_HUGEPAGE);
        return static_cast<T *>(p);
    }
    void deallocate(T *p, size_t n)
    {
        if (n * sizeof(T) < kHuge)
            std::allocator<T>::deallocate(p, n);
        else
            std::free(p);
    }
    template <class U>
    void construct(U *p) noexcept
    {
        ::new (static_cast<void *>(p)) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a)
    {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
using LutVector = std::vector<uint64_t, NoInitAlloc<uint64_t>>;

// The raw records of a KMC suffix file, for readers that decode on the
// device: record r of bucket j (lut[j] <= r < lut[j+1]) holds the key
// ((j mod 4^p) << 2(k-p)) | (its first sufBytes bytes, big-endian).
struct KmcRecords
{
    hipibs_kmc_info info{};
    LutVector lut; // one entry per bucket plus a guard > total
    std::string path;
    int fd = -1;
    uint32_t sufBytes = 0, recBytes = 0;
    uint64_t total = 0;
    KmcRecords() = default;
    KmcRecords(const KmcRecords &) = delete;
    KmcRecords &operator=(const KmcRecords &) = delete;
    ~KmcRecords();
};
int kmcOpenRecords(const std::string &prefix, KmcRecords &out);
int kmcWrite(const std::string &prefix, const uint64_t *keys, const uint32_t *counts, uint64_t n,
             uint32_t k, uint32_t p, uint32_t sigLen, uint32_t counterSize, uint32_t version,
             uint32_t minCount, uint32_t maxCount);

// Records [r0, r0+m) as stored (m * recBytes bytes) into dst, read by up to
// `threads` parallel preads (0 = every usable CPU).
int kmcReadRecords(const KmcRecords &src, uint64_t r0, uint64_t m, void *dst, unsigned threads);
} // namespace hipibs
