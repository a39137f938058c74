// k-mer counting from reads on the device: the step the reference's users
// run with KMC 3.2.4 before the scan (README.md:95: `kmc -k25 -ci2 -cs255`;
// scripts/make_test_corpus.sh:123: `kmc -k<K> -fm -ci1`). KMC is not part
// of the reference's sources; its published counting rules, restated here:
//   - every k-mer of every read (a k-mer never spans two records),
//   - a k-mer holding a byte other than ACGTacgt is skipped,
//   - counted in canonical form, min(code, reverse complement),
//   - kept when min_count <= count <= max_count (-ci, -cx), the stored count
//     clamped to counter_max (-cs).
//
// Counting table in HBM: n_buckets buckets of 128 B, 256 B aligned: the
// bucket's 8 keys (EMPTY = all ones) in its first 64 B, their 8 counts in
// the second. Home bucket bucketOf(key), linear probing at bucket
// granularity, wrapping.
//
// k <= 32 (sorted-slice counting): the k-mers of up to 16 chunks are hashed
// (keyMix, invertible) into a device buffer with a distinct-count sketch;
// the host sizes the table from the sketch (a power of two of buckets), the
// hashes are grouped by slice (512 buckets = 64 KB, the top bits of the
// hash) in two bucketing passes, and one workgroup per slice counts its
// k-mers in LDS and writes the slice back; the few k-mers whose probe runs
// past their slice are counted in the global table with atomics afterwards.
// k = 33..64: every k-mer is counted in the global table with memory-side
// atomics as it is read, and the host grows the table between chunks so
// that it stays at or below 0.7 load even if every position of the next
// chunk were a new key.
//
// Reads arrive as text: the bases of each record followed by one '\n' (any
// byte outside ACGTacgt ends a run of valid bases), in pinned chunks of
// `chunk_bytes`; a record cut by a chunk end continues in the next chunk
// after its last k-1 bases, so every k-mer is counted exactly once.
#include "hipibs_internal.hpp"
#include "wave_probe.hpp"

#include <algorithm>
#include <mutex>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace hipibs
{
namespace
{

constexpr int kCountWords = 16;    // 64-bit words per 128 B bucket
constexpr int kLanePositions = 32; // k-mer positions one lane rolls over
constexpr int kCountBlock = 256;
constexpr int kTilePositions = kCountBlock * kLanePositions;
constexpr int kTileBytes = kTilePositions + 64; // + k - 1 <= 63 bases
constexpr double kMaxLoad = 0.7;                // grow above this
constexpr double kGrowLoad = 0.5;               // load right after growing
constexpr uint64_t kDefaultChunk = 128ull << 20;
constexpr uint32_t kWideSpin = 1u << 14;        // reads of a low word being published

// Bucket layout. k <= 32: 8 keys in words 0..7, their counts in 8..15.
// k = 33..64: 5 keys as (high, low) in words 0..9, counts in 10..14, word
// 15 unused.
template <bool WIDE> struct Layout
{
    static constexpr int kSlots = 8;
    static constexpr int kKeyWords = 8;
    __host__ __device__ static int countWord(int s) { return 8 + s; }
};
template <> struct Layout<true>
{
    static constexpr int kSlots = 5;
    static constexpr int kKeyWords = 10;
    __host__ __device__ static int countWord(int s) { return 10 + s; }
};

// ACGT / acgt -> 0..3, any other byte 4 (the codec's base table,
// src/KmerCodec.hpp:105-118). The low two bits of (b >> 1) ^ (b >> 2) are
// 0, 1, 2, 3 for A, C, G, T in either case; among the bytes 0x40..0x7f,
// kValid marks their positions b & 31 (1, 3, 7, 20), shared by both cases.
// No branch: every lane of a wave takes the same path.
__device__ inline uint32_t baseCode(uint32_t b)
{
    constexpr uint32_t kValid = (1u << 1) | (1u << 3) | (1u << 7) | (1u << 20);
    const uint32_t ok = (uint32_t)((b & 0xC0u) == 0x40u) & ((kValid >> (b & 31u)) & 1u);
    return ok ? (((b >> 1) ^ (b >> 2)) & 3u) : 4u;
}

// Adds `add` to the count of key, storing the key first if it is absent.
// Keys are read with plain loads: a stored key never changes, and an EMPTY
// read from a stale line is settled by the compare-and-swap. Returns false
// only when every bucket was visited (excluded by the host's sizing).
__device__ inline bool countAdd(uint64_t *tab, uint64_t nBuckets, uint64_t key, uint64_t add,
                                bool *isNew)
{
    uint64_t b = bucketOf(key, nBuckets);
    *isNew = false;
    for (uint64_t d = 0; d < nBuckets; ++d)
    {
        unsigned long long *bucket = reinterpret_cast<unsigned long long *>(tab + b * kCountWords);
        const ulonglong2 *kv = reinterpret_cast<const ulonglong2 *>(bucket);
        uint64_t keys[8];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const ulonglong2 v = kv[q];
            keys[2 * q] = v.x;
            keys[2 * q + 1] = v.y;
        }
#pragma unroll
        for (int s = 0; s < 8; ++s)
        {
            uint64_t cur = keys[s];
            if (cur == EMPTY)
            {
                const unsigned long long prev = atomicCAS(bucket + s, EMPTY, key);
                *isNew = prev == EMPTY;
                cur = *isNew ? key : prev;
            }
            if (cur == key)
            {
                atomicAdd(bucket + Layout<false>::countWord(s), (unsigned long long)add);
                return true;
            }
        }
        b = b + 1 == nBuckets ? 0 : b + 1;
    }
    return false;
}

// countAdd for k = 33..64. A slot is claimed by moving its high word from
// EMPTY to the key's (CAS), and its low word is then published with a
// device-scope store; a thread meeting its own high word reads the low
// word, waiting (at most kWideSpin reads) while it still reads all ones,
// and counts there if it is its own. The one canonical code whose low word
// is all ones (k = 64: A^32 T^32) and the one whose high word is all ones
// (k = 64: T^32 A^32, which would read as a free slot) are counted by the
// caller outside the table, so a published low word is never all ones and
// a claimed high word never EMPTY.
__device__ inline bool countAddWide(uint64_t *tab, uint64_t nBuckets, uint64_t hi, uint64_t lo,
                                    uint64_t add, bool *isNew)
{
    uint64_t b = bucketOfWide(hi, lo, nBuckets);
    *isNew = false;
    for (uint64_t d = 0; d < nBuckets; ++d)
    {
        unsigned long long *bucket = reinterpret_cast<unsigned long long *>(tab + b * kCountWords);
        for (int s = 0; s < Layout<true>::kSlots; ++s)
        {
            unsigned long long *hw = bucket + 2 * s, *lw = bucket + 2 * s + 1;
            unsigned long long cur =
                __hip_atomic_load(hw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == EMPTY)
            {
                const unsigned long long prev = atomicCAS(hw, EMPTY, hi);
                if (prev == EMPTY)
                {
                    __hip_atomic_store(lw, (unsigned long long)lo, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    atomicAdd(bucket + Layout<true>::countWord(s), (unsigned long long)add);
                    *isNew = true;
                    return true;
                }
                cur = prev;
            }
            if (cur == hi)
            {
                unsigned long long l =
                    __hip_atomic_load(lw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                for (uint32_t w = 0; l == EMPTY && w < kWideSpin; ++w)
                    l = __hip_atomic_load(lw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (l == lo)
                {
                    atomicAdd(bucket + Layout<true>::countWord(s), (unsigned long long)add);
                    return true;
                }
            }
        }
        b = b + 1 == nBuckets ? 0 : b + 1;
    }
    return false;
}

__device__ inline uint32_t waveSum(uint32_t v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Counts the k-mers at positions skip .. n-k of text (16 B aligned).
// counters: [0] distinct keys stored, [1] k-mers counted, [2] 1 if an
// insert found no slot, [6] count of the k = 64 code A^32 T^32, [9] count
// of the k = 64 code T^32 A^32.
template <bool WIDE>
__global__ __launch_bounds__(kCountBlock) void countKernel(const uint8_t *__restrict__ text,
                                                           uint64_t n, uint64_t skip, uint32_t k,
                                                           uint64_t *tab, uint64_t nBuckets,
                                                           unsigned long long *counters)
{
    __shared__ __align__(16) uint8_t tile[kTileBytes];
    const uint64_t nPos = n >= k ? n - k + 1 : 0;
    const uint64_t nTiles = (nPos + kTilePositions - 1) / kTilePositions;
    // k <= 32: one word; k > 32: the high word holds the first k - 32 bases
    const uint64_t mask = WIDE ? (k == 64 ? ~0ull : (1ull << (2 * (k - 32))) - 1)
                               : (k == 32 ? ~0ull : (1ull << (2 * k)) - 1);
    const uint32_t shift = WIDE ? 2 * (k - 33) : 2 * (k - 1);
    uint32_t counted = 0, fresh = 0, full = 0;
    for (uint64_t t = blockIdx.x; t < nTiles; t += gridDim.x)
    {
        const uint64_t base = t * kTilePositions;
        const uint32_t bytes = (uint32_t)min<uint64_t>(kTilePositions + k - 1, n - base);
        __syncthreads(); // the previous tile is no longer read
        for (uint32_t i = threadIdx.x * 16; i < bytes; i += kCountBlock * 16)
        {
            if (i + 16 <= bytes)
                *reinterpret_cast<uint4 *>(tile + i) =
                    *reinterpret_cast<const uint4 *>(text + base + i);
            else
                for (uint32_t j = i; j < bytes; ++j)
                    tile[j] = text[base + j];
        }
        __syncthreads();
        // Positions [start, start + 32) of the tile: the k-mer at position p
        // ends at byte p + k - 1; `run` counts the valid bytes ending at i.
        const uint32_t start = threadIdx.x * kLanePositions;
        const uint32_t end = min<uint32_t>(bytes, start + kLanePositions + k - 1);
        uint64_t fwd = 0, rc = 0, fhi = 0, rhi = 0;
        uint32_t run = 0;
        for (uint32_t i = start; i < end; ++i)
        {
            const uint32_t c = baseCode(tile[i]);
            if (c > 3)
            {
                run = 0;
                continue;
            }
            ++run;
            if (WIDE)
            {
                fhi = ((fhi << 2) | (fwd >> 62)) & mask;
                fwd = (fwd << 2) | c;
                rc = (rc >> 2) | (rhi << 62);
                rhi = (rhi >> 2) | ((uint64_t)(3 - c) << shift);
            }
            else
            {
                fwd = ((fwd << 2) | c) & mask;
                rc = (rc >> 2) | ((uint64_t)(3 - c) << shift);
            }
            if (run >= k && base + i + 1 >= skip + k)
            {
                bool isNew;
                if (WIDE)
                {
                    const bool f = fhi < rhi || (fhi == rhi && fwd < rc);
                    const uint64_t h = f ? fhi : rhi, l = f ? fwd : rc;
                    if (l == EMPTY) // only A^32 T^32 (k = 64)
                        isNew = atomicAdd(counters + 6, 1ull) == 0;
                    else if (h == EMPTY) // only T^32 A^32 (k = 64)
                        isNew = atomicAdd(counters + 9, 1ull) == 0;
                    else
                        full |= !countAddWide(tab, nBuckets, h, l, 1, &isNew);
                }
                else
                    full |= !countAdd(tab, nBuckets, min(fwd, rc), 1, &isNew);
                counted += 1;
                fresh += isNew;
            }
        }
    }
    counted = waveSum(counted);
    fresh = waveSum(fresh);
    full = waveSum(full);
    if ((threadIdx.x & 63) == 0)
    {
        if (fresh)
            atomicAdd(counters + 0, (unsigned long long)fresh);
        if (counted)
            atomicAdd(counters + 1, (unsigned long long)counted);
        if (full)
            atomicExch(counters + 2, 1ull);
    }
}

// ---- sorted-slice counting (k <= 32) ---------------------------------------
// The k-mers of a flush are first written as h = keyMix(code), the table is
// sized, h is sorted by its top `sbits` bits (n_buckets = 2^(sbits+9), so
// those bits are the slice: the 512-bucket range that holds h's home
// bucket), and one workgroup per slice counts the slice's k-mers in LDS
// with LDS atomics before writing the slice back. A k-mer whose probe runs
// past the slice's last bucket is spilled and counted in the global table
// afterwards. No atomic reaches memory except the spills'.
//
// The sort is least-significant-digit first over digits of `digitBits` top
// bits (digit c = bits [64 - (c+1) digitBits, 64 - c digitBits) of h, at most
// kMaxDigits of them): the hashing kernel counts every digit's histogram,
// then each pass reads every hash once and writes it once, its place the
// digit's start + the hashes of that digit in earlier tiles (decoupled
// look-back across tiles) + its rank in its own tile. Passes after the first
// rank stably, so the result is ordered by all the digits sorted so far.
constexpr int kSliceLog = 9;
constexpr int kSliceBuckets = 1 << kSliceLog; // 64 KB of keys and counts
constexpr int kSliceBlock = 512;
constexpr int kHllLog = 10; // 1024 registers of a HyperLogLog sketch
constexpr int kHllSample = 3; // of the k-mers with hash bits 20..22 zero
constexpr int kMaxDigits = 3;
constexpr int kMaxDigitBits = 9;
constexpr int kBins = 1 << kMaxDigitBits;
constexpr int kTileDigits = 2; // per-tile counts of digits 1 and 2 (a first pass's digits)
constexpr int kRanges = 8;     // the last pass's input ranges, one per XCD

// The value the sort orders a hash by, its digits read from the top bits:
// the k-mer counter sorts its hashes by themselves (its table's slices are
// hash prefixes); a table build sorts its keys' hashes by the slice of
// their home bucket (monotone in the hash's top 32 bits for any bucket
// count) put at the top of the word.
struct SortKey
{
    uint64_t nBuckets; // 0: the hash itself
    uint32_t sliceLog; // buckets per slice = 2^sliceLog
    uint32_t sbits;    // bits of the slice index
    __host__ __device__ uint64_t operator()(uint64_t h) const
    {
        if (!nBuckets)
            return h;
        if (!sbits)
            return 0;
        return ((((h >> 32) * nBuckets) >> 32) >> sliceLog) << (64 - sbits);
    }
};
constexpr int kHashWaves = kCountBlock / 64;
constexpr uint32_t kRegion = 64 * kLanePositions; // hash slots of one wave per tile

// The counts of sort keys a hashing kernel takes for the sort, in LDS:
// digit 0 by range (the top rangeBits bits of digit 1: the last pass's
// ranges), digits 1 and 2 of the current tile (the first pass's places) and
// of the block.
struct DigitCounts
{
    uint32_t rng[kRanges][kBins];
    uint32_t tile[kTileDigits][kBins];
    uint32_t block[kTileDigits][kBins];

    __device__ void clear()
    {
        for (int i = threadIdx.x; i < kRanges * kBins; i += blockDim.x)
            (&rng[0][0])[i] = 0;
        for (int i = threadIdx.x; i < kTileDigits * kBins; i += blockDim.x)
        {
            (&tile[0][0])[i] = 0;
            (&block[0][0])[i] = 0;
        }
    }
    __device__ void add(uint64_t key, uint32_t digitBits, uint32_t rangeBits)
    {
        const uint32_t dmask = (1u << digitBits) - 1;
        atomicAdd(&rng[(key >> (64 - digitBits - rangeBits)) & ((1u << rangeBits) - 1)]
                      [(key >> (64 - digitBits)) & dmask],
                  1u);
#pragma unroll
        for (int dg = 1; dg < kMaxDigits; ++dg)
            atomicAdd(&tile[dg - 1][(key >> (64 - (dg + 1) * digitBits)) & dmask], 1u);
    }
    // The tile's counts out as row `tileIndex` of tileHist and into the
    // block's; call after a barrier that follows the tile's adds.
    __device__ void endTile(uint16_t *tileHist, uint64_t tileIndex)
    {
        for (int i = threadIdx.x; i < kTileDigits * kBins; i += blockDim.x)
        {
            const uint32_t c = (&tile[0][0])[i];
            tileHist[tileIndex * (kTileDigits * kBins) + i] = (uint16_t)c;
            (&block[0][0])[i] += c;
            (&tile[0][0])[i] = 0;
        }
    }
    // The block's counts into the global histograms (digitHist rows 1 and 2);
    // call after a barrier.
    __device__ void flush(unsigned long long *digitHist, unsigned long long *rangeHist)
    {
        for (int i = threadIdx.x; i < kTileDigits * kBins; i += blockDim.x)
            if (const uint32_t c = (&block[0][0])[i])
                atomicAdd(digitHist + kBins + i, (unsigned long long)c);
        for (int i = threadIdx.x; i < kRanges * kBins; i += blockDim.x)
            if (const uint32_t c = (&rng[0][0])[i])
                atomicAdd(rangeHist + i, (unsigned long long)c);
    }
};

// Writes h = keyMix(code) of the k-mers at positions skip .. n-k of text.
// Tile t (kTilePositions positions, the launch's tiles following tileBase)
// owns slots [(tileBase + t) kTilePositions, +kTilePositions) of out: wave w
// the region of kRegion slots from w kRegion, filled from its start, its
// fill in regionCount[(tileBase + t) kHashWaves + w]. Counts the k-mers in
// counters[1] and counters[8]; the histograms of digits 1 and 2 in
// digitHist (kMaxDigits x kBins, row 0 unused), digit 0's by range in
// rangeHist (kRanges x kBins: the range is the top rangeBits bits of digit
// 1, for the last pass of the sort); the tile's own counts of digits 1 and 2
// in tileHist (kTileDigits x kBins u16 per tile: the first pass of a sort
// over two or three digits reads them); and feeds the sampled ones to the
// sketch `hll` (register = max rank).
__global__ __launch_bounds__(kCountBlock) void countHashKernel(
    const uint8_t *__restrict__ text, uint64_t n, uint64_t skip, uint32_t k, uint64_t tileBase,
    uint32_t digitBits, uint64_t *__restrict__ out, uint32_t *__restrict__ regionCount,
    uint16_t *__restrict__ tileHist, uint32_t rangeBits, uint32_t *hll,
    unsigned long long *digitHist, unsigned long long *rangeHist, unsigned long long *counters)
{
    static_assert(kHashWaves * kRegion == kTilePositions, "one region per wave");
    static_assert(kTileDigits == kMaxDigits - 1, "per-tile counts of the digits below the top");
    __shared__ __align__(16) uint8_t tile[kTileBytes];
    __shared__ uint32_t reg[1 << kHllLog];
    __shared__ DigitCounts dc;
    for (int i = threadIdx.x; i < (1 << kHllLog); i += kCountBlock)
        reg[i] = 0;
    dc.clear();
    const uint64_t nPos = n >= k ? n - k + 1 : 0;
    const uint64_t nTiles = (nPos + kTilePositions - 1) / kTilePositions;
    const uint64_t mask = k == 32 ? ~0ull : (1ull << (2 * k)) - 1;
    const uint32_t shift = 2 * (k - 1);
    const int lane = laneId();
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    uint32_t counted = 0;
    for (uint64_t t = blockIdx.x; t < nTiles; t += gridDim.x)
    {
        const uint64_t base = t * kTilePositions;
        const uint32_t bytes = (uint32_t)min<uint64_t>(kTilePositions + k - 1, n - base);
        const uint64_t region = (tileBase + t) * kHashWaves + wv;
        uint64_t *wout = out + region * kRegion;
        __syncthreads(); // the previous tile is no longer read
        for (uint32_t i = threadIdx.x * 16; i < bytes; i += kCountBlock * 16)
        {
            if (i + 16 <= bytes)
                *reinterpret_cast<uint4 *>(tile + i) =
                    *reinterpret_cast<const uint4 *>(text + base + i);
            else
                for (uint32_t j = i; j < bytes; ++j)
                    tile[j] = text[base + j];
        }
        __syncthreads();
        // positions [start, start + 32) of the tile, every lane the same
        // number of steps so that the wave writes its k-mers together
        // (bytes start + d, read one aligned 4-byte word per 4 steps)
        const uint32_t start = threadIdx.x * kLanePositions;
        const uint32_t steps = kLanePositions + k - 1;
        uint64_t fwd = 0, rc = 0;
        uint32_t run = 0, wcount = 0;
        for (uint32_t d0 = 0; d0 < steps; d0 += 4)
        {
            const uint32_t word = *reinterpret_cast<const uint32_t *>(tile + start + d0);
#pragma unroll
            for (uint32_t b = 0; b < 4; ++b)
            {
                const uint32_t d = d0 + b;
                if (d >= steps)
                    break;
                const uint32_t i = start + d;
                const uint32_t c = i < bytes ? baseCode((word >> (8 * b)) & 0xffu) : 4u;
                run = c > 3 ? 0 : run + 1;
                fwd = ((fwd << 2) | (c & 3)) & mask;
                rc = (rc >> 2) | ((uint64_t)(3 - (c & 3)) << shift);
                const bool emit = run >= k && base + i + 1 >= skip + k;
                const uint64_t h = keyMix(min(fwd, rc));
                const uint64_t m = __ballot(emit);
                if (emit)
                {
                    wout[wcount + maskedRank(m)] = h;
                    dc.add(h, digitBits, rangeBits);
                    if (((h >> 20) & ((1u << kHllSample) - 1)) == 0)
                    {
                        const uint32_t rank =
                            (uint32_t)__builtin_ctzll((h >> 23) | (1ull << 40)) + 1u;
                        atomicMax(&reg[h & ((1u << kHllLog) - 1)], rank);
                    }
                }
                wcount += (uint32_t)__popcll(m);
                counted += emit;
            }
        }
        if (lane == 0)
            regionCount[region] = wcount;
        __syncthreads();
        dc.endTile(tileHist, tileBase + t);
    }
    counted = waveSum(counted);
    if (lane == 0 && counted)
    {
        atomicAdd(counters + 1, (unsigned long long)counted);
        atomicAdd(counters + 8, (unsigned long long)counted);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (1 << kHllLog); i += kCountBlock)
        if (reg[i])
            atomicMax(hll + i, reg[i]);
    dc.flush(digitHist, rangeHist);
}

// A table build's hashing: h = keyMix(key) of every row of keys[0..n) but
// EMPTY ones, in countHashKernel's regions (tile t: rows [t kTilePositions,
// +kTilePositions), wave w of it the kRegion rows from w kRegion, packed to
// the region's start, the fill in regionCount) with its counts of the sort
// key sk(h); counters[0] += hashes written.
__global__ __launch_bounds__(kCountBlock) void keyHashKernel(
    const uint64_t *__restrict__ keys, uint64_t n, SortKey sk, uint32_t digitBits,
    uint32_t rangeBits, uint64_t *__restrict__ out, uint32_t *__restrict__ regionCount,
    uint16_t *__restrict__ tileHist, unsigned long long *digitHist,
    unsigned long long *rangeHist, unsigned long long *counters)
{
    __shared__ DigitCounts dc;
    dc.clear();
    const uint64_t nTiles = (n + kTilePositions - 1) / kTilePositions;
    const int lane = laneId();
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    uint32_t written = 0;
    for (uint64_t t = blockIdx.x; t < nTiles; t += gridDim.x)
    {
        const uint64_t region = t * kHashWaves + wv;
        const uint64_t *src = keys + region * kRegion;
        uint64_t *wout = out + region * kRegion;
        const uint64_t left = n > region * kRegion ? n - region * kRegion : 0;
        const uint32_t rows = (uint32_t)min<uint64_t>(left, kRegion);
        __syncthreads(); // the previous tile's counts are out
        uint32_t wcount = 0;
        for (uint32_t r0 = 0; r0 < kRegion; r0 += 64 * 8)
        {
            uint64_t kv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
            {
                const uint32_t i = r0 + (uint32_t)j * 64 + lane;
                kv[j] = i < rows ? __builtin_nontemporal_load(src + i) : EMPTY;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
            {
                const bool ok = kv[j] != EMPTY;
                const uint64_t m = __ballot(ok);
                if (ok)
                {
                    const uint64_t h = keyMix(kv[j]);
                    wout[wcount + maskedRank(m)] = h;
                    dc.add(sk(h), digitBits, rangeBits);
                }
                wcount += (uint32_t)__popcll(m);
            }
        }
        if (lane == 0)
            regionCount[region] = wcount;
        written += wcount;
        __syncthreads();
        dc.endTile(tileHist, t);
    }
    __syncthreads();
    if (lane == 0 && written)
        atomicAdd(counters, (unsigned long long)written);
    dc.flush(digitHist, rangeHist);
}

// A tile's published count of one digit for the look-back: the flag in the
// top two bits (0 not yet, 1 the tile's own count, 2 the count of this and
// every earlier tile), the count below. 32-bit words serve while every
// digit holds fewer than 2^30 hashes.
template <class D> struct LookBack
{
    static constexpr int kShift = (int)sizeof(D) * 8 - 2;
    static constexpr D kOwn = D(1) << kShift;
    static constexpr D kAll = D(2) << kShift;
    static constexpr D kMask = kOwn - 1;
};
constexpr int kGroupBlock = 512;
constexpr int kGroupWaves = kGroupBlock / 64;
constexpr int kGroupPer = 16; // hashes per thread
constexpr int kGroupTile = kGroupBlock * kGroupPer;
constexpr int kWaveHashes = 64 * kGroupPer;
static_assert(kGroupTile == kTilePositions && kWaveHashes * 2 == (int)kRegion,
              "a first-pass tile is one hashing tile, a wave half a region");
constexpr uint32_t kLookBackSpins = 1u << 20; // polls of one descriptor before giving up
constexpr int kLookAhead = 8;                 // earlier tiles' words read per look-back step

// One pass of the sort by the digit at `shift` (bits wide): tile t places
// its hashes at the digit's start + the hashes of that digit in earlier
// tiles + their rank among its own.
//   kPassOffsets: the first pass, reading the hashing kernel's regions; each
//     tile's places come from tileOffset (the per-tile counts the hashing
//     kernel took, summed over tiles); ranks from LDS atomics.
//   kPassLookBack: the same with the earlier tiles' counts found by a
//     decoupled look-back (tiles taken in launch order from tileCounter) --
//     the first pass when there is one digit only.
//   kPassStable: a later pass over n dense hashes, already ordered by the
//     digits below this one (prevShift .. shift), with look-back.
//   kPassRanged: kPassStable for the last pass, its input cut into kRanges
//     ranges by the top bits of the digit below (rangeMeta: kRanges + 1
//     input starts, kRanges + 1 first tiles, kRanges x kBins output bases
//     from the hashing kernel's counts by range): blocks of class
//     blockIdx % kRanges take the tiles of their range in order (one
//     counter per range), and a look-back stops at the range's first tile,
//     so the ranges proceed side by side.
//   Later passes rank stably. The rank
//     keeps that order: a tile whose hashes hold at most two values of the
//     lower digits ranks them by (digit, which value) with LDS atomics;
//     otherwise a wave ranks 64 hashes at a time from ballots of the digit's
//     bits, the waves of a block covering consecutive stretches.
// A look-back that waits longer than kLookBackSpins polls sets *failed and
// takes the missing counts as zero, so that the launch always ends.
enum PassMode
{
    kPassOffsets,
    kPassLookBack,
    kPassStable,
    kPassRanged
};
template <int MODE, class D>
__global__ __launch_bounds__(kGroupBlock) void groupPassKernel(
    const uint64_t *__restrict__ in, const uint32_t *__restrict__ regionCount, uint64_t n,
    uint32_t shift, uint32_t bits, uint32_t prevShift, SortKey sk,
    const uint64_t *__restrict__ digitStart, const uint32_t *__restrict__ tileOffset,
    const uint64_t *__restrict__ rangeMeta, D *desc, unsigned int *tileCounter,
    unsigned long long *failed, uint64_t *__restrict__ out)
{
    using LB = LookBack<D>;
    constexpr bool kRegions = MODE == kPassOffsets || MODE == kPassLookBack;
    constexpr bool kStable = MODE == kPassStable || MODE == kPassRanged;
    __shared__ uint64_t stage[kGroupTile];
    __shared__ uint32_t cntRaw[kGroupWaves * kBins / 2]; // u16 [wave][bin], or u32 [bin][2]
    __shared__ uint16_t localStart[kBins];
    __shared__ uint64_t gbase[kBins];
    __shared__ uint32_t waveTot[kGroupWaves];
    __shared__ uint32_t tileSh, tileCount;
    __shared__ uint64_t keyEnds[2]; // lower-digit values of the tile's first and last hash
    uint16_t(*waveCnt)[kBins] = reinterpret_cast<uint16_t(*)[kBins]>(cntRaw);
    uint32_t *hist = cntRaw; // [digit][value of the lower digits - the first one]
    const int tid = threadIdx.x;
    const int lane = laneId();
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t bins = 1u << bits, dmask = bins - 1;
    const uint64_t pmask = shift > prevShift ? (1ull << (shift - prevShift)) - 1 : 0;
    const uint32_t xr = blockIdx.x % kRanges; // kPassRanged: the block's range
    if (tid == 0)
    {
        if (MODE == kPassOffsets)
        {
            // consecutive tiles on one XCD (blocks go to XCDs in turn), so
            // that one L2 joins the runs that neighbouring tiles write
            const uint32_t full = gridDim.x / kRanges, rem = gridDim.x % kRanges;
            tileSh = xr * full + min(xr, rem) + blockIdx.x / kRanges;
        }
        else
            tileSh = atomicAdd(tileCounter + (MODE == kPassRanged ? xr : 0u), 1u);
    }
    for (int i = tid; i < kGroupWaves * kBins / 2; i += kGroupBlock)
        cntRaw[i] = 0;
    __syncthreads();
    // the tile's descriptor row, the first row of its look-back, its input
    uint64_t t = tileSh, tFirst = 0, inBegin = 0, inEnd = 0;
    if (MODE == kPassRanged)
    {
        const uint64_t tiles = rangeMeta[kRanges + 1 + xr + 1] - rangeMeta[kRanges + 1 + xr];
        if (t >= tiles)
            return; // more blocks of this class than tiles in its range
        inBegin = rangeMeta[xr] + t * kGroupTile;
        inEnd = min<uint64_t>(rangeMeta[xr + 1], inBegin + kGroupTile);
        tFirst = rangeMeta[kRanges + 1 + xr];
        t += tFirst;
    }
    else if (!kRegions)
    {
        inBegin = t * kGroupTile;
        inEnd = min<uint64_t>(n, inBegin + kGroupTile);
    }
    // this wave's stretch of the tile
    const uint64_t *src;
    uint32_t have;
    if (kRegions)
    {
        const uint64_t region = t * (kGroupTile / kRegion) + (w >> 1);
        const uint32_t fill = regionCount[region];
        const uint32_t e0 = (uint32_t)(w & 1) * kWaveHashes;
        src = in + region * kRegion + e0;
        have = fill > e0 ? min<uint32_t>(fill - e0, kWaveHashes) : 0u;
    }
    else
    {
        const uint64_t e0 = inBegin + (uint64_t)w * kWaveHashes;
        src = in + e0;
        have = inEnd > e0 ? (uint32_t)min<uint64_t>(inEnd - e0, kWaveHashes) : 0u;
    }
    uint64_t v[kGroupPer];
    uint32_t rk[kGroupPer];
#pragma unroll
    for (int r = 0; r < kGroupPer; ++r)
    {
        const uint32_t idx = (uint32_t)r * 64 + lane;
        v[r] = idx < have ? __builtin_nontemporal_load(src + idx) : 0ull;
    }
    bool twoKeys = false; // kPassStable: at most two lower-digit values
    uint64_t key0 = 0;
    if (kStable)
    {
        const uint64_t cnt = inEnd - inBegin;
        const uint32_t lastW = (uint32_t)((cnt - 1) / kWaveHashes);
        const uint32_t lastI = (uint32_t)((cnt - 1) % kWaveHashes);
        if (w == 0 && lane == 0)
            keyEnds[0] = (sk(v[0]) >> prevShift) & pmask;
#pragma unroll
        for (int r = 0; r < kGroupPer; ++r)
            if ((uint32_t)w == lastW && (uint32_t)r * 64 + lane == lastI)
                keyEnds[1] = (sk(v[r]) >> prevShift) & pmask;
        __syncthreads();
        key0 = keyEnds[0];
        twoKeys = keyEnds[1] - key0 <= 1;
    }
    if (!kStable || twoKeys)
    {
#pragma unroll
        for (int r = 0; r < kGroupPer; ++r)
        {
            const bool ok = (uint32_t)r * 64 + lane < have;
            const uint64_t kv = sk(v[r]);
            const uint32_t d = (uint32_t)(kv >> shift) & dmask;
            const uint32_t sub = kStable ? (uint32_t)(((kv >> prevShift) & pmask) - key0) : 0u;
            rk[r] = ok ? atomicAdd(&hist[2 * d + sub], 1u) : 0u;
        }
    }
    else
    {
#pragma unroll
        for (int r = 0; r < kGroupPer; ++r)
        {
            const bool ok = (uint32_t)r * 64 + lane < have;
            const uint32_t d = (uint32_t)(sk(v[r]) >> shift) & dmask;
            // lanes holding the same digit, from one ballot per digit bit
            uint64_t peers = __ballot(ok);
            for (uint32_t b = 0; b < bits; ++b)
            {
                const bool set = (d >> b) & 1u;
                const uint64_t m = __ballot(set);
                peers &= set ? m : ~m;
            }
            const uint32_t below = maskedRank(peers);
            const uint32_t before = ok ? waveCnt[w][d] : 0u;
            // every lane has read its digit's count before a leader updates it
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (ok && below == 0)
                waveCnt[w][d] = (uint16_t)(before + __popcll(peers));
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            rk[r] = before + below;
        }
    }
    __syncthreads();
    // the tile's count of digit `tid` (and of its first lower-digit value);
    // the ballot path turns each wave's count into the wave's start among
    // the tile's hashes of that digit
    const bool atomics = !kStable || twoKeys;
    uint32_t tot = 0, firstPart = 0;
    if ((uint32_t)tid < bins)
    {
        if (atomics)
        {
            firstPart = hist[2 * tid];
            tot = firstPart + hist[2 * tid + 1];
        }
        else
            for (int ww = 0; ww < kGroupWaves; ++ww)
            {
                const uint32_t c = waveCnt[ww][tid];
                waveCnt[ww][tid] = (uint16_t)tot;
                tot += c;
            }
        if (MODE != kPassOffsets)
            __hip_atomic_store(desc + t * kBins + tid,
                               (t == tFirst ? LB::kAll : LB::kOwn) | (D)tot, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    // exclusive scan of the counts over the digits: each digit's start in stage
    uint32_t incl = tot;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint32_t y = shuffle32(incl, lane >= off ? lane - off : lane);
        if (lane >= off)
            incl += y;
    }
    if (lane == 63)
        waveTot[w] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int ww = 0; ww < w; ++ww)
        before += waveTot[ww];
    if ((uint32_t)tid < bins)
    {
        localStart[tid] = (uint16_t)(before + incl - tot);
        if (atomics)
            hist[2 * tid + 1] = firstPart; // the second value's start within the digit
    }
    if ((uint32_t)tid == bins - 1)
        tileCount = before + incl;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kGroupPer; ++r)
        if ((uint32_t)r * 64 + lane < have)
        {
            const uint64_t kv = sk(v[r]);
            const uint32_t d = (uint32_t)(kv >> shift) & dmask;
            uint32_t at = localStart[d] + rk[r];
            if (!atomics)
                at += waveCnt[w][d];
            else if (kStable && ((kv >> prevShift) & pmask) != key0)
                at += hist[2 * d + 1];
            stage[at] = v[r];
        }
    if ((uint32_t)tid < bins)
    {
        uint64_t excl = 0;
        if (MODE == kPassOffsets)
            excl = tileOffset[t * kBins + tid];
        else
        {
            // look-back: the hashes of digit `tid` in the tiles before this
            // one, kLookAhead tiles' words loaded at once
            D sum = 0;
            if (t > tFirst)
            {
                int64_t p = (int64_t)t - 1; // the nearest tile not yet added
                uint32_t spins = 0;
                for (bool done = false; !done;)
                {
                    D x[kLookAhead];
#pragma unroll
                    for (int j = 0; j < kLookAhead; ++j)
                        x[j] = p - j >= (int64_t)tFirst
                                   ? __hip_atomic_load(desc + (p - j) * kBins + tid, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT)
                                   : LB::kAll; // before tFirst (never reached: tFirst is whole)
                    // add words up to the first whole one; stop before one
                    // not published yet (polled again)
                    int used = 0;
                    bool stop = false;
#pragma unroll
                    for (int j = 0; j < kLookAhead; ++j)
                    {
                        const D flag = x[j] >> LB::kShift;
                        const bool take = !stop && flag != 0;
                        sum += take ? (x[j] & LB::kMask) : D(0);
                        used = take ? j + 1 : used;
                        done = done || (take && flag == 2);
                        stop = stop || flag != 1;
                    }
                    p -= used;
                    if (done)
                        break;
                    if (used == 0)
                    {
                        if (++spins >= kLookBackSpins)
                        {
                            atomicExch(failed, 1ull);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    else
                        spins = 0;
                }
                __hip_atomic_store(desc + t * kBins + tid, LB::kAll | ((sum + (D)tot) & LB::kMask),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            excl = (MODE == kPassRanged ? rangeMeta[2 * (kRanges + 1) + xr * kBins + tid]
                                        : digitStart[tid]) +
                   (uint64_t)sum;
        }
        gbase[tid] = excl;
    }
    __syncthreads();
    // the staged hashes of one digit are adjacent: so are their stores
    const uint32_t cnt = tileCount;
    for (uint32_t e = tid; e < cnt; e += kGroupBlock)
    {
        const uint64_t x = stage[e];
        const uint32_t d = (uint32_t)(sk(x) >> shift) & dmask;
        out[gbase[d] + (e - localStart[d])] = x;
    }
}

// The first pass's per-tile places from the hashing kernel's per-tile
// counts (tileHist: kTileDigits x kBins u16 per tile, column `col`), in
// groups of kScanTiles tiles: tileSumKernel sums each group, groupScanKernel
// turns the sums into each group's start (the digit's start + the earlier
// groups), tileOffsetKernel into each tile's.
constexpr int kScanTiles = 512;
__global__ __launch_bounds__(kBins) void tileSumKernel(const uint16_t *__restrict__ tileHist,
                                                       uint64_t nTiles, uint32_t col,
                                                       uint32_t bins, uint32_t *groupSum)
{
    const uint32_t d = threadIdx.x;
    if (d >= bins)
        return;
    const uint64_t t0 = (uint64_t)blockIdx.x * kScanTiles;
    const uint64_t t1 = min<uint64_t>(nTiles, t0 + kScanTiles);
    uint32_t sum = 0;
#pragma unroll 8
    for (uint64_t t = t0; t < t1; ++t)
        sum += tileHist[(t * kTileDigits + col) * kBins + d];
    groupSum[(uint64_t)blockIdx.x * kBins + d] = sum;
}

__global__ __launch_bounds__(kBins) void groupScanKernel(uint32_t *groupSum, uint64_t nGroups,
                                                         const uint64_t *__restrict__ digitStart,
                                                         uint32_t bins)
{
    const uint32_t d = threadIdx.x;
    if (d >= bins)
        return;
    uint32_t run = (uint32_t)digitStart[d];
#pragma unroll 8
    for (uint64_t g = 0; g < nGroups; ++g)
    {
        const uint32_t x = groupSum[g * kBins + d];
        groupSum[g * kBins + d] = run;
        run += x;
    }
}

__global__ __launch_bounds__(kBins) void tileOffsetKernel(const uint16_t *__restrict__ tileHist,
                                                          uint64_t nTiles, uint32_t col,
                                                          uint32_t bins,
                                                          const uint32_t *__restrict__ groupSum,
                                                          uint32_t *tileOffset)
{
    const uint32_t d = threadIdx.x;
    if (d >= bins)
        return;
    const uint64_t t0 = (uint64_t)blockIdx.x * kScanTiles;
    const uint64_t t1 = min<uint64_t>(nTiles, t0 + kScanTiles);
    uint32_t run = groupSum[(uint64_t)blockIdx.x * kBins + d];
#pragma unroll 8
    for (uint64_t t = t0; t < t1; ++t)
    {
        tileOffset[t * kBins + d] = run;
        run += tileHist[(t * kTileDigits + col) * kBins + d];
    }
}

// start[s] = the first index of the sorted hashes whose sort key's top
// sbits bits are >= s (s = 0 .. nSlices), by binary search.
__global__ __launch_bounds__(256) void sliceStartKernel(const uint64_t *__restrict__ hs, uint64_t n,
                                                        uint32_t sbits, uint64_t nSlices,
                                                        SortKey sk, uint64_t *start)
{
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nSlices)
        return;
    if (sbits == 0 || s == nSlices)
    {
        start[s] = s == 0 ? 0 : n;
        return;
    }
    uint64_t lo = 0, hi = n;
    while (lo < hi)
    {
        const uint64_t mid = lo + (hi - lo) / 2;
        if ((sk(hs[mid]) >> (64 - sbits)) < s)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[s] = lo;
}
// One workgroup per slice: loads the slice's keys into LDS (or EMPTY keys
// when the table is fresh), counts the slice's k-mers there -- a bucket's 8
// keys read as four 16 B words (stored swizzled against bank conflicts); the key's slot if it is there, else a
// compare-and-swap into the bucket's first EMPTY slot (read again if another
// key took it), then an LDS add to the slot's 32-bit count of this flush --
// and writes the slice back, each count added to the table's (with writeAll,
// an unfilled table, a slice without k-mers is written empty). A k-mer whose
// probe reaches the slice's end without its key or an EMPTY slot is
// appended to spill (counters[7]). counters[0] += keys stored. A flush
// holds fewer than 2^32 k-mers, so no count of one flush overflows.
__global__ __launch_bounds__(kSliceBlock) void countSliceKernel(
    const uint64_t *__restrict__ hs, const uint64_t *__restrict__ start, uint64_t nSlices,
    uint64_t *tab, uint64_t nBuckets, int fresh, int writeAll, uint64_t *spill,
    unsigned long long *counters)
{
    __shared__ __align__(16) unsigned long long keys[kSliceBuckets * 8];
    __shared__ uint32_t cnt[kSliceBuckets * 8];
    __shared__ uint32_t stored;
    if (threadIdx.x == 0)
        stored = 0;
    uint32_t mine = 0;
    // kSliceAhead hashes per thread are loaded before any is counted, so
    // their loads are in flight together
    constexpr int kSliceAhead = 16;
    do
    {
        const uint64_t s = blockIdx.x;
        if (s >= nSlices)
            break;
        const uint64_t lo = start[s], hi = start[s + 1];
        const uint64_t b0 = s << kSliceLog;
        const uint32_t nb = (uint32_t)min<uint64_t>(kSliceBuckets, nBuckets - b0);
        if (lo >= hi)
        {
            // no k-mers: an unfilled table still gets the slice's empty buckets
            if (writeAll)
                for (uint32_t q = threadIdx.x; q < nb * kCountWords; q += kSliceBlock)
                    tab[(b0 + q / kCountWords) * kCountWords + q % kCountWords] =
                        q % kCountWords < 8 ? EMPTY : 0ull;
            break;
        }
        uint64_t hv[kSliceAhead];
        // the first hashes are loaded before the LDS fill, so that the
        // fill overlaps their latency
#pragma unroll
        for (int a = 0; a < kSliceAhead; ++a)
        {
            const uint64_t i = lo + threadIdx.x + (uint64_t)a * kSliceBlock;
            hv[a] = i < hi ? hs[i] : EMPTY;
        }
        // a bucket's 16 B words are stored swizzled (word w of bucket b at
        // w ^ (b mod 4)), so that lanes reading the same word of different
        // buckets meet different LDS banks
        for (uint32_t q = threadIdx.x; q < nb * 8; q += kSliceBlock)
        {
            const uint32_t b = q / 8, slot = q % 8;
            keys[(b * 4 + ((slot >> 1) ^ (b & 3))) * 2 + (slot & 1)] =
                fresh ? EMPTY : tab[(b0 + b) * kCountWords + slot];
            cnt[q] = 0;
        }
        __syncthreads();
        for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (uint64_t)kSliceBlock * kSliceAhead)
        {
            if (i0 != lo + threadIdx.x)
#pragma unroll
                for (int a = 0; a < kSliceAhead; ++a)
                {
                    const uint64_t i = i0 + (uint64_t)a * kSliceBlock;
                    hv[a] = i < hi ? hs[i] : EMPTY;
                }
#pragma unroll
            for (int a = 0; a < kSliceAhead; ++a)
            {
                if (i0 + (uint64_t)a * kSliceBlock >= hi)
                    break;
                const uint64_t h = hv[a];
                const uint64_t key = keyUnmix(h);
                const uint64_t home = ((h >> 32) * nBuckets) >> 32;
                bool done = false;
                for (uint32_t b = (uint32_t)(home - b0); b < nb && !done;)
                {
                    const ulonglong2 *kb = reinterpret_cast<const ulonglong2 *>(keys) + b * 4;
                    uint64_t kk[8];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                    {
                        const ulonglong2 x = kb[q ^ (b & 3)];
                        kk[2 * q] = x.x;
                        kk[2 * q + 1] = x.y;
                    }
                    // keys take the first EMPTY slot and are never removed:
                    // a stored key sits before the first EMPTY one
                    int hit = 8, emp = 8;
#pragma unroll
                    for (int q = 7; q >= 0; --q)
                    {
                        hit = kk[q] == key ? q : hit;
                        emp = kk[q] == EMPTY ? q : emp;
                    }
                    if (hit < emp)
                    {
                        atomicAdd(&cnt[b * 8 + hit], 1u);
                        done = true;
                    }
                    else if (emp < 8)
                    {
                        const unsigned long long prev = atomicCAS(
                            &keys[(b * 4 + (((uint32_t)emp >> 1) ^ (b & 3))) * 2 + (emp & 1)],
                            EMPTY, (unsigned long long)key);
                        if (prev == EMPTY || prev == key)
                        {
                            mine += prev == EMPTY;
                            atomicAdd(&cnt[b * 8 + emp], 1u);
                            done = true;
                        }
                        // else another key took the slot: this bucket is read again
                    }
                    else
                        ++b;
                }
                if (!done)
                    spill[atomicAdd(counters + 7, 1ull)] = h;
            }
        }
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < nb * kCountWords; q += kSliceBlock)
        {
            const uint32_t b = q / kCountWords, w = q % kCountWords;
            uint64_t *dst = tab + (b0 + b) * kCountWords + w;
            if (w < 8)
                *dst = keys[(b * 4 + ((w >> 1) ^ (b & 3))) * 2 + (w & 1)];
            else if (fresh)
                *dst = cnt[b * 8 + w - 8];
            else if (const uint32_t c = cnt[b * 8 + w - 8])
                *dst += c;
        }
    } while (false);
    if (mine)
        atomicAdd(&stored, mine);
    __syncthreads();
    if (threadIdx.x == 0 && stored)
        atomicAdd(counters + 0, (unsigned long long)stored);
}

// A table build's slice: buckets [s << SLICE_LOG, +2^SLICE_LOG) of a
// whole-k-mer table of S key slots per bucket (32 KB), filled in LDS from
// the slice's sorted hashes and written out whole, empty buckets too.
//
// Counted placement (a slice of at most 256 x kSliceKeys keys, every one
// distinct): each key takes a rank in its home bucket (an LDS counter),
// a block scan turns the counts into runs -- bucket b's keys start at
// S x b plus the keys the buckets before it pushed past their end, i.e.
// o(b) = max(0, o(b-1) + count(b) - S), o(-1) = 0 -- and a key goes to
// slot start(home) + rank. Every slot from a key's home bucket up to its
// own is taken, as in a key-by-key insert (where a key takes the first
// EMPTY slot from its home on). Repeated keys share a home run: a key
// that meets its value earlier in its run sends the whole slice to the
// compare-and-swap placement below.
//
// Compare-and-swap placement (larger slices, or keys repeated): a key is
// placed as insertOne places it: from its home bucket on, in the first
// EMPTY slot (claimed by compare-and-swap; a bucket is read again when
// another key took it), stored once.
//
// Either way a key whose slot would lie past the slice's end is appended
// to spill (at most spillCap of them; spillCount counts them all) for the
// global insert. counters: [0] += keys stored, [2] = max displacement (as
// insertKernel).
constexpr int kSliceKeys = 16; // keys per thread of a counted slice

// Blocks loop over slices s = block, block + grid, ...: a slice's hashes
// are loaded into registers while the slice before it is placed, so that
// their latency overlaps that work.
// (128 VGPRs at most: four blocks of 33-35 KB of LDS share a CU)
template <int S, int SLICE_LOG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void
tableSliceKernel(const uint64_t *__restrict__ hs, const uint64_t *__restrict__ start,
                 uint64_t nSlices, uint64_t *slots, uint64_t nBuckets, uint64_t *spill,
                 uint64_t spillCap, unsigned long long *spillCount, unsigned long long *counters)
{
    constexpr uint32_t kB = 1u << SLICE_LOG;
    constexpr uint32_t W = S / 2;       // 16 B words per bucket
    constexpr uint32_t kPer = kB / 256; // buckets per thread in the run scan
    constexpr int kSBits = S == 16 ? 4 : 3;
    constexpr uint64_t kCap = 256ull * kSliceKeys; // most keys of a counted slice
    __shared__ __align__(16) unsigned long long keys[kB * S];
    __shared__ uint32_t runStart[kB];
    __shared__ int32_t waveSum[4], waveMin[4];
    __shared__ uint32_t stored, widest, repeated;
    if (threadIdx.x == 0)
        stored = widest = 0;
    uint32_t mine = 0, disp = 0;
    // slot p of the slice in LDS: a bucket's 16 B words are stored swizzled
    // (word w at w ^ (b mod W)), so that lanes reading the same word of
    // different buckets meet different LDS banks
    auto at = [](uint32_t p) {
        const uint32_t b = p >> kSBits, j = p & (S - 1);
        return (b * W + ((j >> 1) ^ (b & (W - 1)))) * 2 + (j & 1);
    };
    constexpr int kAhead = 8;
    uint64_t kv[kSliceKeys];
    uint64_t lo = 0, hi = 0;
    if (blockIdx.x < nSlices)
    {
        lo = start[blockIdx.x];
        hi = start[blockIdx.x + 1];
        if (hi - lo <= kCap)
#pragma unroll
            for (int r = 0; r < kSliceKeys; ++r)
            {
                const uint64_t i = lo + threadIdx.x + (uint64_t)r * 256;
                kv[r] = i < hi ? hs[i] : 0;
            }
    }
    for (uint64_t s = blockIdx.x; s < nSlices; s += gridDim.x)
    {
        const uint64_t sn = s + gridDim.x;
        const uint64_t loN = sn < nSlices ? start[sn] : 0, hiN = sn < nSlices ? start[sn + 1] : 0;
        const uint64_t b0 = s << SLICE_LOG;
        const uint32_t nb = (uint32_t)min<uint64_t>(kB, nBuckets - b0);
        const uint32_t limit = nb * S;
        bool counted = hi - lo <= kCap;
        for (uint32_t q = threadIdx.x; q < limit; q += 256)
            keys[q] = EMPTY;
        for (uint32_t q = threadIdx.x; q < kB; q += 256)
            runStart[q] = 0;
        if (threadIdx.x == 0)
            repeated = 0;
        __syncthreads();
        uint32_t hr[kSliceKeys]; // home << 16 | rank in the home bucket
        if (counted)
#pragma unroll
            for (int r = 0; r < kSliceKeys; ++r)
                if (lo + threadIdx.x + (uint64_t)r * 256 < hi)
                {
                    const uint64_t h = kv[r];
                    const uint32_t home = (uint32_t)((((h >> 32) * nBuckets) >> 32) - b0);
                    hr[r] = home << 16 | atomicAdd(&runStart[home], 1u);
                    kv[r] = keyUnmix(h);
                }
        // the next slice's hashes
        uint64_t kn[kSliceKeys];
        if (hiN - loN <= kCap)
#pragma unroll
            for (int r = 0; r < kSliceKeys; ++r)
            {
                const uint64_t i = loN + threadIdx.x + (uint64_t)r * 256;
                kn[r] = i < hiN ? hs[i] : 0;
            }
        __syncthreads();
        if (counted)
        {
            // runs: thread t scans buckets [t kPer, +kPer) of (count - S):
            // the segment's sum and its smallest prefix sum, combined over
            // the block (A then B: sum A + sum B, min(min A, sum A + min B))
            int32_t cnt[kPer];
            int32_t sum = 0, mn = INT32_MAX;
#pragma unroll
            for (uint32_t j = 0; j < kPer; ++j)
            {
                cnt[j] = (int32_t)runStart[threadIdx.x * kPer + j];
                sum += cnt[j] - S;
                mn = min(mn, sum);
            }
            const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1)
            {
                const int32_t ps = __shfl_up(sum, d, 64), pm = __shfl_up(mn, d, 64);
                if (lane >= (uint32_t)d)
                {
                    mn = min(pm, ps + mn);
                    sum = ps + sum;
                }
            }
            if (lane == 63)
            {
                waveSum[wave] = sum;
                waveMin[wave] = mn;
            }
            // the segment before this thread's: exclusive of the wave's
            // inclusive scan
            int32_t esum = __shfl_up(sum, 1, 64), emin = __shfl_up(mn, 1, 64);
            if (lane == 0)
            {
                esum = 0;
                emin = INT32_MAX;
            }
            __syncthreads();
            int32_t wsum = 0, wmin = INT32_MAX; // the waves before this one
            for (uint32_t w = 0; w < wave; ++w)
            {
                wmin = min(wmin, wsum + waveMin[w]);
                wsum += waveSum[w];
            }
            if (emin != INT32_MAX)
                wmin = min(wmin, wsum + emin);
            wsum += esum;
            // o(b-1) = P(b-1) - min(0, smallest prefix sum through b-1)
#pragma unroll
            for (uint32_t j = 0; j < kPer; ++j)
            {
                const uint32_t b = threadIdx.x * kPer + j;
                runStart[b] = S * b + (uint32_t)(wsum - min(0, wmin));
                wsum += cnt[j] - S;
                wmin = min(wmin, wsum);
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kSliceKeys; ++r)
                if (lo + threadIdx.x + (uint64_t)r * 256 < hi)
                {
                    const uint32_t home = hr[r] >> 16;
                    const uint32_t pos = runStart[home] + (hr[r] & 0xffffu);
                    if (pos < limit)
                        keys[at(pos)] = kv[r];
                }
            __syncthreads();
            // a repeated key meets its value earlier in its home run
            bool rep = false;
#pragma unroll
            for (int r = 0; r < kSliceKeys; ++r)
                if (lo + threadIdx.x + (uint64_t)r * 256 < hi)
                {
                    const uint32_t first = runStart[hr[r] >> 16];
                    const uint32_t pos = min(first + (hr[r] & 0xffffu), limit);
                    for (uint32_t p = first; p < pos; ++p)
                        rep |= keys[at(p)] == kv[r];
                }
            if (rep)
                repeated = 1;
            __syncthreads();
            counted = repeated == 0;
            if (counted)
            {
#pragma unroll
                for (int r = 0; r < kSliceKeys; ++r)
                    if (lo + threadIdx.x + (uint64_t)r * 256 < hi)
                    {
                        const uint32_t home = hr[r] >> 16;
                        const uint32_t pos = runStart[home] + (hr[r] & 0xffffu);
                        if (pos < limit)
                        {
                            mine += 1;
                            disp = max(disp, (pos >> kSBits) - home);
                        }
                        else
                        {
                            const unsigned long long k = atomicAdd(spillCount, 1ull);
                            if (k < spillCap)
                                spill[k] = kv[r];
                        }
                    }
            }
            else
            {
                for (uint32_t q = threadIdx.x; q < limit; q += 256)
                    keys[q] = EMPTY;
                __syncthreads();
            }
        }
        if (!counted)
        {
            uint64_t hv[kAhead];
            for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (uint64_t)blockDim.x * kAhead)
            {
#pragma unroll
                for (int a = 0; a < kAhead; ++a)
                {
                    const uint64_t i = i0 + (uint64_t)a * blockDim.x;
                    hv[a] = i < hi ? hs[i] : 0;
                }
#pragma unroll
                for (int a = 0; a < kAhead; ++a)
                {
                    if (i0 + (uint64_t)a * blockDim.x >= hi)
                        break;
                    const uint64_t h = hv[a];
                    const uint64_t key = keyUnmix(h);
                    const uint32_t home = (uint32_t)((((h >> 32) * nBuckets) >> 32) - b0);
                    bool done = false;
                    for (uint32_t b = home; b < nb && !done;)
                    {
                        const ulonglong2 *kb = reinterpret_cast<const ulonglong2 *>(keys) + b * W;
                        uint64_t kk[S];
#pragma unroll
                        for (int q = 0; q < (int)W; ++q)
                        {
                            const ulonglong2 x = kb[q ^ (b & (W - 1))];
                            kk[2 * q] = x.x;
                            kk[2 * q + 1] = x.y;
                        }
                        // keys take the first EMPTY slot and are never
                        // removed: a stored key sits before the first EMPTY one
                        int hit = S, emp = S;
#pragma unroll
                        for (int q = S - 1; q >= 0; --q)
                        {
                            hit = kk[q] == key ? q : hit;
                            emp = kk[q] == EMPTY ? q : emp;
                        }
                        if (hit < emp)
                            done = true;
                        else if (emp < S)
                        {
                            const unsigned long long prev = atomicCAS(
                                &keys[at(b * S + (uint32_t)emp)], EMPTY, (unsigned long long)key);
                            if (prev == EMPTY)
                            {
                                mine += 1;
                                disp = max(disp, b - home);
                                done = true;
                            }
                            else if (prev == key)
                                done = true;
                            // else another key took the slot: this bucket is read again
                        }
                        else
                            ++b;
                    }
                    if (!done)
                    {
                        const unsigned long long k = atomicAdd(spillCount, 1ull);
                        if (k < spillCap)
                            spill[k] = key;
                    }
                }
            }
        }
        __syncthreads();
        // 16 B stores of the slice's keys, unswizzled
        const ulonglong2 *src = reinterpret_cast<const ulonglong2 *>(keys);
        ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(slots + b0 * S);
        for (uint32_t q = threadIdx.x; q < nb * W; q += blockDim.x)
        {
            const uint32_t b = q / W;
            dst[q] = src[b * W + ((q % W) ^ (b & (W - 1)))];
        }
        // the next slice fills LDS once every thread has read it
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSliceKeys; ++r)
            kv[r] = kn[r];
        lo = loN;
        hi = hiN;
    }
    if (mine)
        atomicAdd(&stored, mine);
    if (disp)
        atomicMax(&widest, disp);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        if (stored)
            atomicAdd(counters + 0, (unsigned long long)stored);
        if (widest)
            atomicMax(counters + 2, (unsigned long long)widest);
    }
}

// The spilled k-mers (h form), counted in the global table.
__global__ __launch_bounds__(256) void countSpillKernel(const uint64_t *__restrict__ spill,
                                                        uint64_t n, uint64_t *tab,
                                                        uint64_t nBuckets,
                                                        unsigned long long *counters)
{
    uint32_t fresh = 0, full = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
    {
        bool isNew;
        full |= !countAdd(tab, nBuckets, keyUnmix(spill[i]), 1, &isNew);
        fresh += isNew;
    }
    fresh = waveSum(fresh);
    full = waveSum(full);
    if ((threadIdx.x & 63) == 0)
    {
        if (fresh)
            atomicAdd(counters + 0, (unsigned long long)fresh);
        if (full)
            atomicExch(counters + 2, 1ull);
    }
}

__global__ __launch_bounds__(256) void countFillKernel(uint64_t *tab, uint64_t words, int keyWords)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (uint64_t)gridDim.x * blockDim.x)
        tab[i] = (int)(i % kCountWords) < keyWords ? EMPTY : 0ull;
}

// Moves every (key, count) of the old table into the new one. Keys are
// distinct, so the thread that stores a key is the only writer of its
// count: a plain store (the low word of a two-word key: a device-scope
// store, as the counting insert publishes it).
template <bool WIDE>
__global__ __launch_bounds__(256) void countRehashKernel(const uint64_t *__restrict__ old,
                                                         uint64_t oldBuckets, uint64_t *tab,
                                                         uint64_t nBuckets,
                                                         unsigned long long *counters)
{
    using L = Layout<WIDE>;
    const uint64_t slots = oldBuckets * L::kSlots;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots;
         i += (uint64_t)gridDim.x * blockDim.x)
    {
        const uint64_t b = i / L::kSlots, s = i % L::kSlots;
        const uint64_t *ob = old + b * kCountWords;
        const uint64_t key = WIDE ? ob[2 * s] : ob[s];
        if (key == EMPTY)
            continue;
        const uint64_t lo = WIDE ? ob[2 * s + 1] : 0;
        const uint64_t c = ob[L::countWord((int)s)];
        uint64_t h = WIDE ? bucketOfWide(key, lo, nBuckets) : bucketOf(key, nBuckets);
        bool stored = false;
        for (uint64_t d = 0; d < nBuckets && !stored; ++d)
        {
            unsigned long long *bucket =
                reinterpret_cast<unsigned long long *>(tab + h * kCountWords);
            for (int q = 0; q < L::kSlots && !stored; ++q)
            {
                unsigned long long *kw = bucket + (WIDE ? 2 * q : q);
                if (*kw == EMPTY && atomicCAS(kw, EMPTY, key) == EMPTY)
                {
                    if (WIDE)
                        __hip_atomic_store(kw + 1, (unsigned long long)lo, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    bucket[L::countWord(q)] = c;
                    stored = true;
                }
            }
            h = h + 1 == nBuckets ? 0 : h + 1;
        }
        if (!stored)
            atomicExch(counters + 2, 1ull);
    }
}

// Keys with lo <= count <= hi, in no particular order (two words each when
// WIDE), with their counts clamped to cmax (counts may be null). filt: [0]
// kept, [1] below lo, [2] above hi. A wave claims its output range with
// one atomic add.
template <bool WIDE>
__global__ __launch_bounds__(256) void countExtractKernel(const uint64_t *__restrict__ tab,
                                                          uint64_t nBuckets, uint64_t lo,
                                                          uint64_t hi, uint64_t cmax,
                                                          uint64_t *keys, uint32_t *counts,
                                                          uint64_t cap,
                                                          unsigned long long *filt)
{
    using L = Layout<WIDE>;
    const uint64_t slots = nBuckets * L::kSlots;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t below = 0, above = 0;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x; w < slots; w += stride)
    {
        const uint64_t i = w + threadIdx.x;
        uint64_t key = EMPTY, low = 0, c = 0;
        if (i < slots)
        {
            const uint64_t b = i / L::kSlots, s = i % L::kSlots;
            const uint64_t *tb = tab + b * kCountWords;
            key = WIDE ? tb[2 * s] : tb[s];
            low = WIDE ? tb[2 * s + 1] : 0;
            c = tb[L::countWord((int)s)];
        }
        const bool present = key != EMPTY;
        const bool keep = present && c >= lo && c <= hi;
        below += present && c < lo;
        above += present && c > hi;
        const uint64_t m = __ballot(keep);
        if (!m)
            continue;
        uint64_t at = 0;
        if (laneId() == __builtin_ctzll(m))
            at = atomicAdd(filt + 0, (unsigned long long)__popcll(m));
        at = __shfl(at, __builtin_ctzll(m), 64);
        if (keep)
        {
            const uint64_t o = at + maskedRank(m);
            if (o < cap)
            {
                if (WIDE)
                {
                    keys[2 * o] = key;
                    keys[2 * o + 1] = low;
                }
                else
                    keys[o] = key;
                if (counts)
                    counts[o] = (uint32_t)min(c, cmax);
            }
        }
    }
    below = waveSum(below);
    above = waveSum(above);
    if ((threadIdx.x & 63) == 0)
    {
        if (below)
            atomicAdd(filt + 1, (unsigned long long)below);
        if (above)
            atomicAdd(filt + 2, (unsigned long long)above);
    }
}

int cuCount(int device)
{
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    return cus;
}

uint32_t gridFor(uint64_t items, uint64_t perBlock, int device)
{
    const uint64_t want = (items + perBlock - 1) / perBlock;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)cuCount(device) * 8));
}

// The buffers and steps of one sort by slice: a hashing kernel
// (countHashKernel, keyHashKernel) fills the regions and the counts, then
// sort() runs the passes of groupPassKernel and finds the slices' starts.
// A k-mer counter and a table build each keep one.
struct SliceSorter
{
    int device = 0;
    uint32_t digitBits = kMaxDigitBits;
    bool wideLookBack = false;            // 64-bit look-back words always
    unsigned long long *dDigit = nullptr; // digit histograms (rows 1, 2; row 0 from dRange)
    unsigned long long *dRange = nullptr; // digit 0 by range of digit 1
    uint32_t *dRegion = nullptr;          // the fill of every hashing region
    uint16_t *dTileHist = nullptr;        // every hashing tile's counts of digits 1 and 2
    uint64_t tileCap = 0;                 // hashing tiles both hold
    uint64_t *dDigitStart = nullptr;      // the digits' starts (kMaxDigits x (kBins + 1))
    uint64_t *dRangeMeta = nullptr;       // the last pass's ranges (see groupPassKernel)
    uint32_t *dGroupSum = nullptr;        // the first pass's per-group sums, then starts
    uint32_t *dTileOffset = nullptr;      // the first pass's per-tile places
    uint64_t tileOffsetCap = 0;
    void *dDesc = nullptr; // look-back words
    uint64_t descBytes = 0;
    unsigned int *dTileCounter = nullptr;
    unsigned long long *dBuild = nullptr; // a table build's counters (see buildTableSorted)
    uint64_t *dStart = nullptr; // the slices' starts
    uint64_t startCap = 0;
    std::vector<unsigned long long> hDigit, hRange;
    std::vector<uint64_t> hDigitStart, hRangeMeta;

    // Bits of digit 1 that cut the last pass's input into ranges.
    uint32_t rangeBits() const { return std::min<uint32_t>(3, digitBits); }

    // The histograms and small buffers, cleared (on the caller's device).
    int init(hipStream_t s)
    {
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dDigit),
                               kMaxDigits * kBins * sizeof(unsigned long long)));
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dRange),
                               kRanges * kBins * sizeof(unsigned long long)));
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dDigitStart),
                               kMaxDigits * (kBins + 1) * sizeof(uint64_t)));
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dRangeMeta),
                               (2 * (kRanges + 1) + kRanges * kBins) * sizeof(uint64_t)));
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dTileCounter),
                               kRanges * sizeof(unsigned int)));
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dBuild), 4 * sizeof(unsigned long long)));
        return clearCounts(s);
    }

    int clearCounts(hipStream_t s)
    {
        HIPIBS_CHECK(hipMemsetAsync(dDigit, 0, kMaxDigits * kBins * sizeof(unsigned long long), s));
        HIPIBS_CHECK(hipMemsetAsync(dRange, 0, kRanges * kBins * sizeof(unsigned long long), s));
        HIPIBS_CHECK(hipMemsetAsync(dBuild, 0, 4 * sizeof(unsigned long long), s));
        return 0;
    }

    // Regions and per-tile counts for `tiles` hashing tiles.
    int reserveTiles(uint64_t tiles)
    {
        if (tiles <= tileCap)
            return 0;
        if (dRegion)
            HIPIBS_CHECK(hipFree(dRegion));
        if (dTileHist)
            HIPIBS_CHECK(hipFree(dTileHist));
        dRegion = nullptr;
        dTileHist = nullptr;
        tileCap = 0;
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dRegion),
                               tiles * kHashWaves * sizeof(uint32_t)));
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dTileHist),
                               tiles * kTileDigits * kBins * sizeof(uint16_t)));
        tileCap = tiles;
        return 0;
    }

    void release()
    {
        for (void *p : {(void *)dDigit, (void *)dRange, (void *)dRegion, (void *)dTileHist,
                        (void *)dDigitStart, (void *)dRangeMeta, (void *)dGroupSum,
                        (void *)dTileOffset, dDesc, (void *)dTileCounter, (void *)dBuild,
                        (void *)dStart})
            if (p)
                (void)hipFree(p);
        *this = SliceSorter();
    }

    // Sorts the nHash hashes that the hashing kernel wrote in tilesFirst
    // tiles of src by the top sbits bits of sk (passes over digitBits-bit
    // digits, ping-ponging between src and dst); *sorted receives the
    // buffer holding the result, dStart the nSlices + 1 slice starts. Every
    // launch goes through timed(stage 1, launch); a look-back that timed out
    // sets *failed (device). `who` names the caller in errors.
    template <class Timed>
    int sort(uint64_t *src, uint64_t *dst, uint64_t nHash, uint64_t tilesFirst, uint32_t sbits,
             SortKey sk, unsigned long long *failed, hipStream_t stream, Timed &&timed,
             uint64_t **sorted, const std::string &who)
    {
        int rc = 0;
        const uint64_t nSlices = sbits ? (1ull << sbits) : 1;
        const uint32_t passes = std::max<uint32_t>(1, (sbits + digitBits - 1) / digitBits);
        if (passes > (uint32_t)kMaxDigits)
            return setError(HIPIBS_E_NOMEM, who + ": the table needs more than " +
                                                std::to_string(kMaxDigits * digitBits) +
                                                " bits of slice index");
        if (startCap < nSlices + 1)
        {
            if (dStart)
                HIPIBS_CHECK(hipFree(dStart));
            dStart = nullptr;
            startCap = 0;
            HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dStart), (nSlices + 1) * 8));
            startCap = nSlices + 1;
        }
        // every digit's start from its histogram
        const uint32_t bins = 1u << digitBits;
        hDigit.resize(kMaxDigits * kBins);
        hDigitStart.assign(kMaxDigits * (kBins + 1), 0);
        HIPIBS_CHECK(hipMemcpy(hDigit.data(), dDigit, hDigit.size() * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost));
        hRange.resize(kRanges * kBins);
        HIPIBS_CHECK(hipMemcpy(hRange.data(), dRange, hRange.size() * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost));
        for (uint32_t d = 0; d < bins; ++d)
        {
            hDigit[d] = 0;
            for (int x = 0; x < kRanges; ++x)
                hDigit[d] += hRange[x * kBins + d];
        }
        uint64_t widest = 0;
        for (int c = 0; c < kMaxDigits; ++c)
        {
            uint64_t acc = 0;
            for (uint32_t d = 0; d < bins; ++d)
            {
                hDigitStart[c * (kBins + 1) + d] = acc;
                acc += hDigit[c * kBins + d];
                widest = std::max<uint64_t>(widest, hDigit[c * kBins + d]);
            }
            hDigitStart[c * (kBins + 1) + bins] = acc;
            if (acc != nHash)
                return setError(HIPIBS_E_HIP, who + ": the slice sort lost hashes");
        }
        HIPIBS_CHECK(hipMemcpy(dDigitStart, hDigitStart.data(),
                               hDigitStart.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        // look-back words: 32 bits while no digit holds 2^30 hashes
        const bool narrow = widest < (1ull << 30) && !wideLookBack;
        const uint64_t tilesDense = (nHash + kGroupTile - 1) / kGroupTile;
        const uint64_t tilesMax = std::max(tilesFirst, tilesDense + kRanges);
        const uint64_t wordBytes = narrow ? 4 : 8;
        if (descBytes < tilesMax * kBins * wordBytes)
        {
            if (dDesc)
                HIPIBS_CHECK(hipFree(dDesc));
            dDesc = nullptr;
            descBytes = 0;
            HIPIBS_CHECK(hipMalloc(&dDesc, tilesMax * kBins * 8));
            descBytes = tilesMax * kBins * 8;
        }
        // the passes, least significant digit first; then the slices'
        // starts. The first pass over two or three digits takes its tiles'
        // places from the hashing kernel's per-tile counts; every other
        // pass finds them by look-back.
        const uint32_t prevShift = 64 - passes * digitBits;
        for (uint32_t q = 0; q < passes; ++q)
        {
            const uint32_t c = passes - 1 - q;
            const uint32_t shift = 64 - (c + 1) * digitBits;
            const uint64_t *starts = dDigitStart + c * (kBins + 1);
            if (q == 0 && c > 0)
            {
                const uint64_t groups = (tilesFirst + kScanTiles - 1) / kScanTiles;
                if (tileOffsetCap < tilesFirst)
                {
                    if (dGroupSum)
                        HIPIBS_CHECK(hipFree(dGroupSum));
                    if (dTileOffset)
                        HIPIBS_CHECK(hipFree(dTileOffset));
                    dGroupSum = nullptr;
                    dTileOffset = nullptr;
                    tileOffsetCap = 0;
                    HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dGroupSum),
                                           (groups + 1) * kBins * sizeof(uint32_t)));
                    HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dTileOffset),
                                           tilesFirst * kBins * sizeof(uint32_t)));
                    tileOffsetCap = tilesFirst;
                }
                if ((rc = timed(1, [&] {
                         tileSumKernel<<<(uint32_t)groups, kBins, 0, stream>>>(
                             dTileHist, tilesFirst, c - 1, bins, dGroupSum);
                         groupScanKernel<<<1, kBins, 0, stream>>>(dGroupSum, groups, starts, bins);
                         tileOffsetKernel<<<(uint32_t)groups, kBins, 0, stream>>>(
                             dTileHist, tilesFirst, c - 1, bins, dGroupSum, dTileOffset);
                         groupPassKernel<kPassOffsets, uint32_t>
                             <<<(uint32_t)tilesFirst, kGroupBlock, 0, stream>>>(
                                 src, dRegion, nHash, shift, digitBits, prevShift, sk, starts,
                                 dTileOffset, nullptr, static_cast<uint32_t *>(dDesc), dTileCounter,
                                 failed, dst);
                     })))
                    return rc;
            }
            else if (q > 0 && c == 0)
            {
                // the last pass in kRanges ranges: their input starts (the
                // counts by range are ordered as the input is, by digit 1),
                // first tiles, and output bases (digit 0's start + the
                // earlier ranges' counts of it)
                const uint32_t nr = 1u << rangeBits();
                hRangeMeta.assign(2 * (kRanges + 1) + kRanges * kBins, 0);
                uint64_t at = 0, tile = 0, longest = 0;
                for (int x = 0; x <= kRanges; ++x)
                {
                    hRangeMeta[x] = at;
                    hRangeMeta[kRanges + 1 + x] = tile;
                    if (x == kRanges)
                        break;
                    uint64_t cnt = 0;
                    for (uint32_t d = 0; d < bins; ++d)
                    {
                        hRangeMeta[2 * (kRanges + 1) + x * kBins + d] =
                            x ? hRangeMeta[2 * (kRanges + 1) + (x - 1) * kBins + d] +
                                    hRange[(x - 1) * kBins + d]
                              : hDigitStart[d];
                        cnt += hRange[x * kBins + d];
                    }
                    if ((uint32_t)x >= nr && cnt)
                        return setError(HIPIBS_E_HIP, who + ": a hash outside the ranges");
                    const uint64_t tiles = (cnt + kGroupTile - 1) / kGroupTile;
                    at += cnt;
                    tile += tiles;
                    longest = std::max(longest, tiles);
                }
                if (at != nHash)
                    return setError(HIPIBS_E_HIP, who + ": the slice sort lost hashes");
                HIPIBS_CHECK(hipMemcpy(dRangeMeta, hRangeMeta.data(),
                                       hRangeMeta.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
                HIPIBS_CHECK(hipMemsetAsync(dDesc, 0, tile * kBins * wordBytes, stream));
                HIPIBS_CHECK(hipMemsetAsync(dTileCounter, 0, kRanges * sizeof(unsigned int), stream));
                if ((rc = timed(1, [&] {
                         const dim3 grid((uint32_t)(kRanges * longest));
                         if (narrow)
                             groupPassKernel<kPassRanged, uint32_t><<<grid, kGroupBlock, 0, stream>>>(
                                 src, dRegion, nHash, shift, digitBits, prevShift, sk, starts,
                                 nullptr, dRangeMeta, static_cast<uint32_t *>(dDesc), dTileCounter,
                                 failed, dst);
                         else
                             groupPassKernel<kPassRanged, uint64_t><<<grid, kGroupBlock, 0, stream>>>(
                                 src, dRegion, nHash, shift, digitBits, prevShift, sk, starts,
                                 nullptr, dRangeMeta, static_cast<uint64_t *>(dDesc), dTileCounter,
                                 failed, dst);
                     })))
                    return rc;
            }
            else
            {
                const uint64_t tiles = q == 0 ? tilesFirst : tilesDense;
                HIPIBS_CHECK(hipMemsetAsync(dDesc, 0, tiles * kBins * wordBytes, stream));
                HIPIBS_CHECK(hipMemsetAsync(dTileCounter, 0, 4, stream));
                if ((rc = timed(1, [&] {
                         const dim3 grid((uint32_t)tiles);
                         if (q == 0 && narrow)
                             groupPassKernel<kPassLookBack, uint32_t><<<grid, kGroupBlock, 0, stream>>>(
                                 src, dRegion, nHash, shift, digitBits, prevShift, sk, starts, nullptr,
                                 nullptr, static_cast<uint32_t *>(dDesc), dTileCounter, failed, dst);
                         else if (q == 0)
                             groupPassKernel<kPassLookBack, uint64_t><<<grid, kGroupBlock, 0, stream>>>(
                                 src, dRegion, nHash, shift, digitBits, prevShift, sk, starts, nullptr,
                                 nullptr, static_cast<uint64_t *>(dDesc), dTileCounter, failed, dst);
                         else if (narrow)
                             groupPassKernel<kPassStable, uint32_t><<<grid, kGroupBlock, 0, stream>>>(
                                 src, dRegion, nHash, shift, digitBits, prevShift, sk, starts, nullptr,
                                 nullptr, static_cast<uint32_t *>(dDesc), dTileCounter, failed, dst);
                         else
                             groupPassKernel<kPassStable, uint64_t><<<grid, kGroupBlock, 0, stream>>>(
                                 src, dRegion, nHash, shift, digitBits, prevShift, sk, starts, nullptr,
                                 nullptr, static_cast<uint64_t *>(dDesc), dTileCounter, failed, dst);
                     })))
                    return rc;
            }
            std::swap(src, dst);
        }
        if ((rc = timed(1, [&] {
                 sliceStartKernel<<<(uint32_t)((nSlices + 1 + 255) / 256), 256, 0, stream>>>(
                     src, nHash, sbits, nSlices, sk, dStart);
             })))
            return rc;
        *sorted = src;
        return 0;
    }
};

} // namespace
} // namespace hipibs

using namespace hipibs;

namespace
{
// The hash buffers of sorted-slice counting are reused by the process's
// next counter: the largest pair a freed counter returns is kept (one pair
// per process; hipibs_counter_release_cache frees it). Allocating and
// freeing tens of GB per counter can cost the runtime seconds.
struct HashCache
{
    std::mutex m;
    int device = -1;
    uint64_t *p[2] = {nullptr, nullptr};
    uint64_t cap = 0;
} gHashCache;

// Takes the cached pair if it is on `device` and holds `need` entries.
bool takeCachedHash(int device, uint64_t need, uint64_t *out[2], uint64_t *cap)
{
    std::lock_guard<std::mutex> g(gHashCache.m);
    if (!gHashCache.p[0] || gHashCache.device != device || gHashCache.cap < need)
        return false;
    out[0] = gHashCache.p[0];
    out[1] = gHashCache.p[1];
    *cap = gHashCache.cap;
    gHashCache.p[0] = gHashCache.p[1] = nullptr;
    gHashCache.cap = 0;
    return true;
}

// Keeps a freed counter's pair if it is larger than the cached one (the
// smaller pair is freed); called with the pair's device current.
void returnHash(int device, uint64_t *p[2], uint64_t cap)
{
    std::lock_guard<std::mutex> g(gHashCache.m);
    if (gHashCache.p[0] && gHashCache.cap >= cap)
    {
        for (int i = 0; i < 2; ++i)
            (void)hipFree(p[i]);
        return;
    }
    if (gHashCache.p[0])
    {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(gHashCache.device);
        for (int i = 0; i < 2; ++i)
            (void)hipFree(gHashCache.p[i]);
        (void)hipSetDevice(cur);
    }
    gHashCache.device = device;
    gHashCache.p[0] = p[0];
    gHashCache.p[1] = p[1];
    gHashCache.cap = cap;
}

// A freed counter's table is kept for the process's next counter: the
// largest one so far (hipibs_counter_release_cache frees it). A counter
// allocates its table once it knows its size, and a new allocation of
// tens of GB that reuses memory the process freed can cost the runtime a
// second or more.
struct CountTableCache
{
    std::mutex m;
    int device = -1;
    uint64_t *p = nullptr;
    uint64_t bytes = 0;
} gCountTable;

bool takeCountTable(int device, uint64_t need, uint64_t **out, uint64_t *bytes)
{
    std::lock_guard<std::mutex> g(gCountTable.m);
    if (!gCountTable.p || gCountTable.device != device || gCountTable.bytes < need)
        return false;
    *out = gCountTable.p;
    *bytes = gCountTable.bytes;
    gCountTable.p = nullptr;
    gCountTable.bytes = 0;
    return true;
}

// Keeps p if it is larger than the kept table (the smaller is freed);
// called with p's device current and its stream idle.
void returnCountTable(int device, uint64_t *p, uint64_t bytes)
{
    std::lock_guard<std::mutex> g(gCountTable.m);
    if (gCountTable.p && gCountTable.bytes >= bytes)
    {
        (void)hipFree(p);
        return;
    }
    if (gCountTable.p)
    {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(gCountTable.device);
        (void)hipFree(gCountTable.p);
        (void)hipSetDevice(cur);
    }
    gCountTable.device = device;
    gCountTable.p = p;
    gCountTable.bytes = bytes;
}

// The sorter of table builds is kept for the process's next build (its
// per-tile buffers grow with the largest build), one per process.
struct BuildSorterCache
{
    std::mutex m;
    int device = -1;
    SliceSorter *so = nullptr;
} gBuildSorter;

SliceSorter *takeBuildSorter(int device)
{
    std::lock_guard<std::mutex> g(gBuildSorter.m);
    if (!gBuildSorter.so || gBuildSorter.device != device)
        return nullptr;
    SliceSorter *so = gBuildSorter.so;
    gBuildSorter.so = nullptr;
    return so;
}

void returnBuildSorter(int device, SliceSorter *so)
{
    std::lock_guard<std::mutex> g(gBuildSorter.m);
    if (gBuildSorter.so)
    {
        so->release();
        delete so;
        return;
    }
    gBuildSorter.device = device;
    gBuildSorter.so = so;
}

// The sort buffer of table builds (8 B per key plus room for spills) is
// kept for the process's next build: the largest one so far.
struct ScratchCache
{
    std::mutex m;
    int device = -1;
    uint64_t *p = nullptr;
    uint64_t cap = 0; // entries
} gScratch;

bool takeScratch(int device, uint64_t need, uint64_t **out, uint64_t *cap)
{
    std::lock_guard<std::mutex> g(gScratch.m);
    if (!gScratch.p || gScratch.device != device || gScratch.cap < need)
        return false;
    *out = gScratch.p;
    *cap = gScratch.cap;
    gScratch.p = nullptr;
    gScratch.cap = 0;
    return true;
}

void returnScratch(int device, uint64_t *p, uint64_t cap)
{
    std::lock_guard<std::mutex> g(gScratch.m);
    if (gScratch.p && gScratch.cap >= cap)
    {
        (void)hipFree(p);
        return;
    }
    if (gScratch.p)
    {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(gScratch.device);
        (void)hipFree(gScratch.p);
        (void)hipSetDevice(cur);
    }
    gScratch.device = device;
    gScratch.p = p;
    gScratch.cap = cap;
}

void releaseBuildSorter()
{
    {
        std::lock_guard<std::mutex> g(gCountTable.m);
        if (gCountTable.p)
        {
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(gCountTable.device);
            (void)hipFree(gCountTable.p);
            (void)hipSetDevice(cur);
            gCountTable.p = nullptr;
            gCountTable.bytes = 0;
        }
    }
    {
        std::lock_guard<std::mutex> g(gScratch.m);
        if (gScratch.p)
        {
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(gScratch.device);
            (void)hipFree(gScratch.p);
            (void)hipSetDevice(cur);
            gScratch.p = nullptr;
            gScratch.cap = 0;
        }
    }
    std::lock_guard<std::mutex> g(gBuildSorter.m);
    if (!gBuildSorter.so)
        return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(gBuildSorter.device);
    gBuildSorter.so->release();
    delete gBuildSorter.so;
    gBuildSorter.so = nullptr;
    (void)hipSetDevice(cur);
}
} // namespace

struct hipibs_counter : ReadSink
{
    int device = 0;
    uint32_t k = 0;
    bool wide = false; // k > 32: two-word keys, 5 per bucket
    uint64_t chunkBytes = 0;
    hipStream_t stream = nullptr;
    uint64_t *dTab = nullptr;
    uint64_t tabBytes = 0; // dTab's allocation (a kept table may be larger)
    uint64_t nBuckets = 0;
    unsigned long long *dCounters = nullptr; // [0] distinct, [1] counted, [2] full, [3..5] filter,
                                             // [6] count of A^32 T^32 (k = 64), [7] spilled
                                             // and [8] hashed k-mers of the open flush,
                                             // [9] count of T^32 A^32 (k = 64),
                                             // [10] a sort pass's look-back timed out
    unsigned long long *hCounters = nullptr; // pinned copy
    // sorted-slice counting (k <= 32): the hashed k-mers of the open flush
    // (dHash[0]; grouped by slice through both buffers, the other one then
    // takes the spills)
    uint64_t *dHash[2] = {nullptr, nullptr};
    uint64_t hashCap = 0;    // entries of each buffer
    uint64_t hashBudget = 0; // most entries one flush holds
    uint64_t nSlots = 0;     // slots the open flush's hashing tiles own in dHash[0]
    uint64_t nHash = 0;      // hashes of the open flush (after a sync)
    uint32_t *dHll = nullptr; // the flush's distinct-count sketch
    // the sort by slice (its digit width from the environment
    // HIPIBS_SORT_DIGIT_BITS, 1..9: narrower digits take more passes, which
    // the tests use; HIPIBS_SORT_LOOKBACK_WIDE=1 forces 64-bit look-back
    // words)
    SliceSorter sorter;
    bool tableUnfilled = false; // dTab allocated but not yet written
    char *hText[2] = {nullptr, nullptr};     // pinned chunks (host text only)
    uint8_t *dText = nullptr;
    // device time per stage (0 hashing or, k > 32, counting; 1 grouping by
    // slice; 2 slice counts; 3 spills), from HIP events around every launch,
    // read after the stream's next synchronisation
    std::vector<hipEvent_t> evPool;
    size_t evUsed = 0;
    std::vector<std::pair<size_t, int>> evPending;
    double stageMs[4] = {0, 0, 0, 0};
    double kernelMs = 0;                   // their sum
    uint64_t launches = 0;
    uint64_t spilled = 0, flushes = 0;     // sorted-slice counting
    // new keys a flush sizes the table for, per distinct k-mer the sketch
    // estimates (environment HIPIBS_COUNT_SKETCH_SCALE; a small value
    // undersizes the table so that most k-mers take the spill path, which
    // the tests use)
    double sketchScale = 1.15;
    int cur = 0;          // chunk being filled
    uint64_t len = 0;     // bytes in it
    uint64_t recLen = 0;  // bases of the open record so far
    uint64_t reads = 0;   // records
    uint64_t distinct = 0, counted = 0;
    int err = 0;

    int failed() const override { return err; }

    // Launches with HIP events around it, its time going to `stage`.
    template <class F> int timed(int stage, F &&launch)
    {
        while (evUsed + 2 > evPool.size())
        {
            hipEvent_t e = nullptr;
            HIPIBS_CHECK(hipEventCreate(&e));
            evPool.push_back(e);
        }
        HIPIBS_CHECK(hipEventRecord(evPool[evUsed], stream));
        launch();
        HIPIBS_CHECK(hipGetLastError());
        HIPIBS_CHECK(hipEventRecord(evPool[evUsed + 1], stream));
        evPending.push_back({evUsed, stage});
        evUsed += 2;
        launches += 1;
        return 0;
    }

    // Reads the times of the launches since the last call (stream idle).
    int collectTimes()
    {
        for (const auto &pe : evPending)
        {
            float ms = 0;
            HIPIBS_CHECK(hipEventElapsedTime(&ms, evPool[pe.first], evPool[pe.first + 1]));
            stageMs[pe.second] += ms;
            kernelMs += ms;
        }
        evPending.clear();
        evUsed = 0;
        return 0;
    }
    void bases(const char *p, size_t n) override
    {
        while (n && !err)
        {
            if (len == chunkBytes)
            {
                flush(true);
                continue;
            }
            const size_t m = std::min<size_t>(n, chunkBytes - len);
            std::memcpy(hText[cur] + len, p, m);
            len += m;
            recLen += m;
            p += m;
            n -= m;
        }
    }
    void recordEnd() override
    {
        if (err)
            return;
        if (len == chunkBytes)
            flush(false);
        hText[cur][len++] = '\n';
        recLen = 0;
        reads += 1;
    }

    // Submits the filled chunk; with midRecord the next chunk starts with the
    // open record's last k-1 bases (fewer if it is shorter).
    void flush(bool midRecord)
    {
        const uint64_t carry = midRecord ? std::min<uint64_t>(recLen, k - 1) : 0;
        const int done = cur;
        const uint64_t doneLen = len;
        err = submit(done, doneLen);
        cur ^= 1;
        len = 0;
        if (!err && carry)
        {
            std::memcpy(hText[cur], hText[done] + doneLen - carry, carry);
            len = carry;
        }
    }

    // Waits for the previous chunk, then copies chunk `slot` to the device
    // and counts it. The other pinned chunk is free once this returns.
    int submit(int slot, uint64_t n)
    {
        int rc = sync();
        if (rc)
            return rc;
        const uint64_t positions = n >= k ? n - k + 1 : 0;
        if (!positions)
            return 0;
        if (!wide)
        {
            HIPIBS_CHECK(hipMemcpyAsync(dText, hText[slot], n, hipMemcpyHostToDevice, stream));
            return appendHashes(dText, n, 0);
        }
        rc = reserve(distinct + positions);
        if (rc)
            return rc;
        HIPIBS_CHECK(hipMemcpyAsync(dText, hText[slot], n, hipMemcpyHostToDevice, stream));
        return launchCount(dText, n, 0);
    }

    // k <= 32: hashes the k-mers at positions skip .. n-k of device text
    // into the open flush (called with the stream synchronised), making
    // room first: the buffers grow up to hashBudget entries, beyond that
    // the open flush is counted.
    int appendHashes(const uint8_t *text, uint64_t n, uint64_t skip)
    {
        const uint64_t positions = n >= k + skip ? n - k + 1 - skip : 0;
        if (!positions)
            return 0;
        int rc = 0;
        if ((rc = budget()))
            return rc;
        // the launch's tiles own whole slots ranges, the skipped positions' too
        const uint64_t tiles = (n - k + 1 + kTilePositions - 1) / kTilePositions;
        const uint64_t slots = tiles * kTilePositions;
        if (slots > hashBudget)
            return setError(HIPIBS_E_NOMEM, "k-mer counter: a text piece exceeds the flush budget");
        if (nSlots + slots > hashBudget && (rc = flushSorted()))
            return rc;
        if ((rc = hashRoom(nSlots + slots)))
            return rc;
        if (!dHll)
        {
            HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dHll), (4u << kHllLog)));
            HIPIBS_CHECK(hipMemsetAsync(dHll, 0, 4u << kHllLog, stream));
            if ((rc = sorter.init(stream)) ||
                (rc = sorter.reserveTiles(hashBudget / kTilePositions + 1)))
                return rc;
        }
        const uint64_t tileBase = nSlots / kTilePositions;
        if ((rc = timed(0, [&] {
                 countHashKernel<<<gridFor(n - k + 1, kTilePositions, device), kCountBlock, 0,
                                   stream>>>(text, n, skip, k, tileBase, sorter.digitBits,
                                             dHash[0], sorter.dRegion, sorter.dTileHist,
                                             sorter.rangeBits(), dHll, sorter.dDigit,
                                             sorter.dRange, dCounters);
             })))
            return rc;
        HIPIBS_CHECK(hipMemcpyAsync(hCounters, dCounters, 9 * sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost, stream));
        nSlots += slots;
        return 0;
    }

    // The most hash slots one flush holds: 16 read chunks' hashing tiles
    // (2 Gi k-mers at the default chunk), both buffers within a quarter of
    // the free memory, at least one chunk's tiles.
    int budget()
    {
        if (hashBudget)
            return 0;
        size_t freeB = 0, totalB = 0;
        HIPIBS_CHECK(hipMemGetInfo(&freeB, &totalB));
        const uint64_t chunkSlots =
            (chunkBytes + kTilePositions - 1) / kTilePositions * kTilePositions;
        hashBudget = std::min<uint64_t>(16 * chunkSlots, freeB / 4 / 16 / kTilePositions *
                                                             kTilePositions);
        // places within a flush are 32-bit (the first pass's per-tile ones)
        hashBudget = std::min<uint64_t>(hashBudget, 1ull << 31);
        hashBudget = std::max<uint64_t>(hashBudget, chunkSlots);
        return 0;
    }

    // Grows the hash buffers to hold `need` entries (at most the budget),
    // keeping the open flush's; called with the stream synchronised.
    int hashRoom(uint64_t need)
    {
        if (need <= hashCap)
            return 0;
        if (!hashCap && takeCachedHash(device, need, dHash, &hashCap))
            return 0;
        {
            const uint64_t want = std::min<uint64_t>(
                hashBudget, std::max<uint64_t>({need, 2 * hashCap, 1ull << 22}));
            uint64_t *nb[2] = {nullptr, nullptr};
            for (int i = 0; i < 2; ++i)
                if (hipMalloc(reinterpret_cast<void **>(&nb[i]), want * 8) != hipSuccess)
                {
                    (void)hipGetLastError();
                    if (nb[0])
                        (void)hipFree(nb[0]);
                    return setError(HIPIBS_E_NOMEM, "k-mer counter: no device memory for " +
                                                        std::to_string(want) + " hashed k-mers");
                }
            if (nSlots)
                HIPIBS_CHECK(hipMemcpyAsync(nb[0], dHash[0], nSlots * 8, hipMemcpyDeviceToDevice,
                                            stream));
            HIPIBS_CHECK(hipStreamSynchronize(stream));
            for (int i = 0; i < 2; ++i)
            {
                if (dHash[i])
                    HIPIBS_CHECK(hipFree(dHash[i]));
                dHash[i] = nb[i];
            }
            hashCap = want;
        }
        return 0;
    }

    // Distinct values among the hashed k-mers of the open flush, from the
    // sketch (HyperLogLog over the 1/8 of them with hash bits 20..22 zero;
    // linear counting while registers are empty).
    double sketchEstimate()
    {
        uint32_t reg[1 << kHllLog];
        if (hipMemcpy(reg, dHll, sizeof(reg), hipMemcpyDeviceToHost) != hipSuccess)
            return -1;
        const double m = (double)(1 << kHllLog);
        double sum = 0;
        int zeros = 0;
        for (uint32_t r : reg)
        {
            sum += std::ldexp(1.0, -(int)r);
            zeros += r == 0;
        }
        double e = 0.7213 / (1.0 + 1.079 / m) * m * m / sum;
        if (e <= 2.5 * m && zeros)
            e = m * std::log(m / zeros);
        return e * (1 << kHllSample);
    }

    // Counts the open flush: sizes the table from the sketch, sorts the
    // hashes by slice, counts every slice in LDS, then the spills in the
    // global table (after growing it for them).
    int flushSorted()
    {
        int rc = sync();
        if (rc || wide)
            return rc;
        nHash = hCounters[8];
        if (!nHash)
        {
            nSlots = 0; // tiles without k-mers
            return 0;
        }
        const double est = sketchEstimate();
        if (est < 0)
            return setError(HIPIBS_E_HIP, "k-mer counter: reading the sketch failed");
        const uint64_t newKeys = std::min<uint64_t>(nHash, (uint64_t)(est * sketchScale) + 4096);
        if ((rc = reserve(distinct + newKeys, distinct == 0)))
            return rc;
        const int lg = 63 - __builtin_clzll(nBuckets); // a power of two
        const uint32_t sbits = lg > kSliceLog ? (uint32_t)(lg - kSliceLog) : 0u;
        const uint64_t nSlices = sbits ? (1ull << sbits) : 1;
        uint64_t *src = nullptr;
        auto timedStage = [&](int stage, auto &&launch) { return timed(stage, launch); };
        if ((rc = sorter.sort(dHash[0], dHash[1], nHash, nSlots / kTilePositions, sbits,
                              SortKey{0, 0, 0}, dCounters + 10, stream, timedStage, &src,
                              "k-mer counter")))
            return rc;
        uint64_t *dst = src == dHash[0] ? dHash[1] : dHash[0];
        // the grouped hashes are in src; dst takes the spills
        const uint64_t *sorted = src;
        uint64_t *other = dst;
        if ((rc = timed(2, [&] {
                 countSliceKernel<<<(uint32_t)nSlices, kSliceBlock, 0, stream>>>(
                     sorted, sorter.dStart, nSlices, dTab, nBuckets, distinct == 0, tableUnfilled,
                     other, dCounters);
             })))
            return rc;
        HIPIBS_CHECK(hipMemcpyAsync(hCounters, dCounters, 11 * sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost, stream));
        HIPIBS_CHECK(hipStreamSynchronize(stream));
        if ((rc = collectTimes()))
            return rc;
        if (hCounters[10])
            return setError(HIPIBS_E_HIP, "k-mer counter: a sort pass's look-back timed out");
        tableUnfilled = false; // every slice written
        distinct = hCounters[0];
        const uint64_t spills = hCounters[7];
        if (spills)
        {
            // the spills sit in `other`, which a growth leaves in place
            if ((rc = reserve(distinct + spills)))
                return rc;
            if ((rc = timed(3, [&] {
                     countSpillKernel<<<gridFor(spills, 256 * 4, device), 256, 0, stream>>>(
                         other, spills, dTab, nBuckets, dCounters);
                 })))
                return rc;
            HIPIBS_CHECK(hipMemcpyAsync(hCounters, dCounters, 9 * sizeof(unsigned long long),
                                        hipMemcpyDeviceToHost, stream));
            HIPIBS_CHECK(hipStreamSynchronize(stream));
            if ((rc = collectTimes()))
                return rc;
            if (hCounters[2])
                return setError(HIPIBS_E_FULL, "k-mer counter: an insert found no free slot");
            distinct = hCounters[0];
        }
        spilled += spills;
        HIPIBS_CHECK(hipMemsetAsync(dCounters + 7, 0, 2 * sizeof(unsigned long long), stream));
        HIPIBS_CHECK(hipMemsetAsync(dHll, 0, 4u << kHllLog, stream));
        if ((rc = sorter.clearCounts(stream)))
            return rc;
        HIPIBS_CHECK(hipStreamSynchronize(stream));
        hCounters[7] = hCounters[8] = 0;
        nHash = 0;
        nSlots = 0;
        flushes += 1;
        return 0;
    }

    // Counts positions skip .. n-k of device text (room reserved).
    int launchCount(const uint8_t *text, uint64_t n, uint64_t skip)
    {
        const int rc = timed(0, [&] {
            if (wide)
                countKernel<true><<<gridFor(n - k + 1, kTilePositions, device), kCountBlock, 0,
                                    stream>>>(text, n, skip, k, dTab, nBuckets, dCounters);
            else
                countKernel<false><<<gridFor(n - k + 1, kTilePositions, device), kCountBlock, 0,
                                     stream>>>(text, n, skip, k, dTab, nBuckets, dCounters);
        });
        if (rc)
            return rc;
        HIPIBS_CHECK(hipMemcpyAsync(hCounters, dCounters, 3 * sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost, stream));
        return 0;
    }

    // Device-resident text in pieces of at most chunkBytes: a piece starts
    // 16 B aligned at or before the previous piece's last k-1 bytes and
    // skips the positions that piece counted.
    int countDevice(const uint8_t *text, uint64_t n)
    {
        if (!wide && n >= k)
        {
            // room for the whole text's k-mers at once (within the budget)
            int rc = sync();
            if (!rc)
                rc = budget();
            if (!rc)
                rc = hashRoom(std::min<uint64_t>(
                    hashBudget, nSlots + (n - k + 1) +
                                    (n / std::max<uint64_t>(chunkBytes - 64, 1) + 2) *
                                        kTilePositions));
            if (rc)
                return rc;
        }
        uint64_t off = 0, from = 0; // piece start, first position to count
        while (n >= k && from <= n - k)
        {
            const uint64_t len = std::min<uint64_t>(chunkBytes, n - off);
            if (len < k || off + len - k + 1 <= from)
                break;
            int rc = sync();
            if (!rc && !wide)
                rc = appendHashes(text + off, len, from - off);
            else
            {
                if (!rc)
                    rc = reserve(distinct + (off + len - k + 1 - from));
                if (!rc)
                    rc = launchCount(text + off, len, from - off);
            }
            if (rc)
                return rc;
            if (off + len == n)
                break;
            from = off + len - k + 1;
            off = from & ~15ull;
        }
        return sync();
    }

    int sync()
    {
        HIPIBS_CHECK(hipStreamSynchronize(stream));
        if (int rc = collectTimes())
            return rc;
        if (hCounters[2])
            return setError(HIPIBS_E_FULL, "k-mer counter: an insert found no free slot");
        distinct = hCounters[0];
        counted = hCounters[1];
        return 0;
    }

    // Makes room for `need` keys at <= kMaxLoad, growing to kGrowLoad and
    // at least doubling (called with the stream idle).
    int reserve(uint64_t need, bool lazyFill = false)
    {
        const int slots = wide ? Layout<true>::kSlots : Layout<false>::kSlots;
        if ((double)need <= kMaxLoad * (double)(nBuckets * slots))
            return 0;
        uint64_t want = (uint64_t)((double)need / kGrowLoad) / slots + 1;
        want = std::max<uint64_t>(want, 2 * nBuckets);
        want = std::max<uint64_t>(want, 1024);
        if (!wide) // a power of two: slices of the sorted-slice count
            want = 1ull << (64 - __builtin_clzll(want - 1));
        if (want >= (1ull << 32))
            return setError(HIPIBS_E_NOMEM, "k-mer counter: more than 2^32 buckets needed");
        uint64_t *t = nullptr;
        uint64_t tBytes = want * kCountWords * 8;
        if (!takeCountTable(device, tBytes, &t, &tBytes))
        {
            hipError_t e = hipMalloc(reinterpret_cast<void **>(&t), tBytes);
            if (e != hipSuccess)
            {
                // the buffers kept for the next counter or build may be in
                // the way (a kept table too small for this one among them)
                (void)hipGetLastError();
                releaseBuildSorter();
                e = hipMalloc(reinterpret_cast<void **>(&t), tBytes);
            }
            if (e != hipSuccess)
            {
                (void)hipGetLastError();
                return setError(HIPIBS_E_NOMEM, "k-mer counter: no device memory for " +
                                                    std::to_string(want) + " buckets (" +
                                                    std::to_string(distinct) +
                                                    " distinct k-mers so far)");
            }
        }
        const uint64_t words = want * kCountWords;
        // a new table with no keys to move in may stay unfilled when the
        // caller's slice counts write every slice (lazyFill)
        tableUnfilled = lazyFill && !dTab;
        if (!tableUnfilled)
        {
            countFillKernel<<<gridFor(words, 256 * 4, device), 256, 0, stream>>>(
                t, words, wide ? Layout<true>::kKeyWords : Layout<false>::kKeyWords);
            HIPIBS_CHECK(hipGetLastError());
        }
        if (dTab)
        {
            if (wide)
                countRehashKernel<true><<<gridFor(nBuckets * slots, 256 * 4, device), 256, 0,
                                          stream>>>(dTab, nBuckets, t, want, dCounters);
            else
                countRehashKernel<false><<<gridFor(nBuckets * slots, 256 * 4, device), 256, 0,
                                           stream>>>(dTab, nBuckets, t, want, dCounters);
            HIPIBS_CHECK(hipGetLastError());
            HIPIBS_CHECK(hipStreamSynchronize(stream));
            // freed at once: a growing counter may need the memory next
            HIPIBS_CHECK(hipFree(dTab));
        }
        dTab = t;
        tabBytes = tBytes;
        nBuckets = want;
        return 0;
    }

    // The pinned chunks and the device chunk, on first use of host text.
    int hostBuffers()
    {
        if (dText)
            return 0;
        for (int i = 0; i < 2; ++i)
            HIPIBS_CHECK(hipHostMalloc(reinterpret_cast<void **>(&hText[i]), chunkBytes,
                                       hipHostMallocDefault));
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dText), chunkBytes));
        return 0;
    }

    // Submits what is buffered and waits for every chunk.
    int drain()
    {
        if (err)
            return err;
        if (len)
        {
            err = submit(cur, len);
            cur ^= 1;
            len = 0;
            recLen = 0;
            if (err)
                return err;
        }
        return err = sync();
    }

    // drain, then counts the open flush.
    int finish()
    {
        if (drain())
            return err;
        return err = flushSorted();
    }

    // Keys with lo <= count <= hi into dKeys (two words each for k > 32;
    // and clamped counts into dCounts if non-null), at most cap; filt
    // receives kept / below / above.
    int extract(uint64_t lo, uint64_t hi, uint64_t cmax, uint64_t *dKeys, uint32_t *dCounts,
                uint64_t cap, hipibs_count_filter *filt)
    {
        HIPIBS_CHECK(hipMemsetAsync(dCounters + 3, 0, 3 * sizeof(unsigned long long), stream));
        if (nBuckets)
        {
            const uint64_t slots =
                nBuckets * (wide ? Layout<true>::kSlots : Layout<false>::kSlots);
            if (wide)
                countExtractKernel<true><<<gridFor(slots, 256 * 8, device), 256, 0, stream>>>(
                    dTab, nBuckets, lo, hi, cmax, dKeys, dCounts, cap, dCounters + 3);
            else
                countExtractKernel<false><<<gridFor(slots, 256 * 8, device), 256, 0, stream>>>(
                    dTab, nBuckets, lo, hi, cmax, dKeys, dCounts, cap, dCounters + 3);
            HIPIBS_CHECK(hipGetLastError());
        }
        HIPIBS_CHECK(hipMemcpyAsync(hCounters + 3, dCounters + 3, 7 * sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost, stream));
        HIPIBS_CHECK(hipStreamSynchronize(stream));
        filt->kept = hCounters[3];
        filt->below = hCounters[4];
        filt->above = hCounters[5];
        // the k = 64 codes A^32 T^32 and T^32 A^32, counted outside the table
        const struct
        {
            uint64_t count;
            uint64_t key[2];
        } special[2] = {{hCounters[6], {0, EMPTY}}, {hCounters[9], {EMPTY, 0}}};
        for (const auto &sp : special)
        {
            const uint64_t sc = sp.count;
            if (!sc)
                continue;
            if (sc < lo)
                filt->below += 1;
            else if (sc > hi)
                filt->above += 1;
            else
            {
                const uint64_t o = filt->kept++;
                if (o < cap)
                {
                    const uint32_t c = (uint32_t)std::min<uint64_t>(sc, cmax);
                    HIPIBS_CHECK(hipMemcpy(dKeys + 2 * o, sp.key, 16, hipMemcpyHostToDevice));
                    if (dCounts)
                        HIPIBS_CHECK(hipMemcpy(dCounts + o, &c, 4, hipMemcpyHostToDevice));
                }
            }
        }
        return 0;
    }
};

static int checkFilter(uint32_t lo, uint32_t hi, uint32_t cmax)
{
    if (lo > hi)
        return setError(HIPIBS_E_ARG, "k-mer counter: min_count > max_count");
    if (cmax == 0)
        return setError(HIPIBS_E_ARG, "k-mer counter: counter_max must be >= 1");
    return 0;
}

namespace hipibs
{
// Whole-k-mer tables of k <= 32 from device keys by sorting, in place of
// one compare-and-swap walk per key: the keys' hashes are sorted by the
// table slice (32 KB of buckets) of their home bucket, each slice is filled
// in LDS and written out whole, and the few keys that overflow their slice
// are inserted by insertKernel. The sort needs one buffer of about 9 B per
// key (kept for the next build); the table's own slots are its other. The table (allocated; its slots need not be
// cleared: every bucket is written) gets the same keys, each stored once
// from its home bucket on. Returns 0 with
// *used = false (nothing done) below HIPIBS_SORTED_BUILD_MIN keys (default
// 2^20), when the sort's buffer does not fit, or
// when the slice index needs more than three digits (the environment's
// HIPIBS_SORT_DIGIT_BITS / HIPIBS_SORT_LOOKBACK_WIDE apply as to counters).
int buildTableSorted(hipibs_table *t, const uint64_t *dKeys, uint64_t n, hipStream_t s,
                     bool *used)
{
    *used = false;
    uint64_t minKeys = 1ull << 20;
    if (const char *e = std::getenv("HIPIBS_SORTED_BUILD_MIN"))
        minKeys = std::strtoull(e, nullptr, 10);
    if (t->k > 32 || t->mlen || n == 0 || n < minKeys)
        return 0;
    uint32_t digitBits = kMaxDigitBits;
    if (const char *db = std::getenv("HIPIBS_SORT_DIGIT_BITS"))
        digitBits = (uint32_t)std::min(kMaxDigitBits, std::max(1, std::atoi(db)));
    const char *lw = std::getenv("HIPIBS_SORT_LOOKBACK_WIDE");
    const bool wideLookBack = lw && std::atoi(lw) != 0;
    const int S = (int)t->bucketSlots;
    const uint32_t sliceLog = S == 16 ? 8 : 9; // 32 KB of keys per slice
    const uint64_t nB = t->nBuckets;
    const uint64_t nSlices = (nB + (1ull << sliceLog) - 1) >> sliceLog;
    const uint32_t sbits = nSlices > 1 ? (uint32_t)(64 - __builtin_clzll(nSlices - 1)) : 0u;
    if (std::max<uint32_t>(1, (sbits + digitBits - 1) / digitBits) > (uint32_t)kMaxDigits)
        return 0;
    const uint64_t tiles = (n + kTilePositions - 1) / kTilePositions;
    const uint64_t slots = tiles * kTilePositions;
    // one sort buffer of its own: the sorted hashes, then room for the keys
    // that overflow their slice; the table's slots (written whole at the end)
    // serve as the other -- the hashing writes there when the passes are
    // odd, so that the last pass ends in the buffer
    const uint32_t passes = std::max<uint32_t>(1, (sbits + digitBits - 1) / digitBits);
    if (nB * (uint64_t)S < slots)
        return 0;
    const uint64_t spillRoom = slots / 8 + 65536;
    uint64_t *buf = nullptr;
    uint64_t cap = 0;
    const bool cached = takeScratch(t->device, slots + spillRoom, &buf, &cap);
    size_t freeB = 0, totalB = 0;
    HIPIBS_CHECK(hipMemGetInfo(&freeB, &totalB));
    const uint64_t need = (cached ? 0 : (slots + spillRoom) * 8) +
                          tiles * (kTileDigits * kBins * 2 + kHashWaves * 4 + kBins * 12) +
                          (8ull << sbits) + (64ull << 20);
    if ((double)need > 0.9 * (double)freeB)
    {
        if (cached)
            returnScratch(t->device, buf, cap);
        return 0;
    }
    if (!cached)
    {
        // room for larger builds to come (the buffer is kept for them): half
        // again as many entries, when that fits
        cap = slots + spillRoom;
        const uint64_t roomy = cap / 2 * 3;
        if ((double)(need + (roomy - cap) * 8) < 0.8 * (double)freeB)
            cap = roomy;
        if (hipMalloc(reinterpret_cast<void **>(&buf), cap * 8) != hipSuccess)
        {
            (void)hipGetLastError();
            return 0; // no room: the caller inserts key by key
        }
    }
    uint64_t *first = passes % 2 == 0 ? buf : t->dSlots; // the hashing's output
    uint64_t *second = first == buf ? t->dSlots : buf;
    SliceSorter *so = takeBuildSorter(t->device);
    int rc = 0;
    if (so)
        rc = so->clearCounts(s);
    else
    {
        so = new SliceSorter();
        so->device = t->device;
        rc = so->init(s);
    }
    so->digitBits = digitBits;
    so->wideLookBack = wideLookBack;
    // [0] hashes, [1] spills, [2] a look-back timed out (cleared with the counts)
    unsigned long long *dCnt = so->dBuild;
    auto cleanup = [&] {
        (void)hipStreamSynchronize(s);
        returnScratch(t->device, buf, cap);
        returnBuildSorter(t->device, so);
    };
    if (!rc)
        rc = so->reserveTiles(tiles);
    const SortKey sk{nB, sliceLog, sbits};
    if (!rc)
    {
        keyHashKernel<<<gridFor(n, kTilePositions, t->device), kCountBlock, 0, s>>>(
            dKeys, n, sk, so->digitBits, so->rangeBits(), first, so->dRegion, so->dTileHist,
            so->dDigit, so->dRange, dCnt);
        if (hipGetLastError() != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "table build: hashing launch failed");
    }
    unsigned long long hc[4] = {0, 0, 0, 0};
    if (!rc && (hipMemcpyAsync(hc, dCnt, sizeof(hc), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        rc = setError(HIPIBS_E_HIP, "table build: hashing failed");
    uint64_t *sorted = nullptr;
    if (!rc && !hc[0])
        rc = clearTableSlots(t, s); // no keys: every bucket EMPTY
    if (!rc && hc[0])
    {
        auto plain = [&](int, auto &&launch) {
            launch();
            return hipGetLastError() == hipSuccess
                       ? 0
                       : setError(HIPIBS_E_HIP, "table build: a sort launch failed");
        };
        rc = so->sort(first, second, hc[0], tiles, sbits, sk, dCnt + 2, s, plain, &sorted,
                      "table build");
        if (!rc && sorted != buf)
            rc = setError(HIPIBS_E_HIP, "table build: the sort ended in the table's slots");
        uint64_t *spill = buf + slots;
        const uint64_t spillCap = cap - slots;
        if (!rc)
        {
            auto *ctr = reinterpret_cast<unsigned long long *>(t->dCounters);
            // as many blocks as are resident at once (each loops over slices)
            int perCu = 0;
            if (S == 16)
                (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &perCu, tableSliceKernel<16, 8>, 256, 0);
            else
                (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &perCu, tableSliceKernel<8, 9>, 256, 0);
            const uint32_t grid = (uint32_t)std::min<uint64_t>(
                nSlices, (uint64_t)cuCount(t->device) * (uint64_t)std::max(1, perCu));
            if (S == 16)
                tableSliceKernel<16, 8><<<grid, 256, 0, s>>>(
                    sorted, so->dStart, nSlices, t->dSlots, nB, spill, spillCap, dCnt + 1, ctr);
            else
                tableSliceKernel<8, 9><<<grid, 256, 0, s>>>(
                    sorted, so->dStart, nSlices, t->dSlots, nB, spill, spillCap, dCnt + 1, ctr);
            if (hipGetLastError() != hipSuccess)
                rc = setError(HIPIBS_E_HIP, "table build: slice launch failed");
        }
        if (!rc && (hipMemcpyAsync(hc, dCnt, sizeof(hc), hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess))
            rc = setError(HIPIBS_E_HIP, "table build: slice fill failed");
        if (!rc && hc[2])
            rc = setError(HIPIBS_E_HIP, "table build: a sort pass's look-back timed out");
        // keys past their slice's end: inserted from their home bucket on
        // (more than the room holds: every key, the ones stored are found)
        if (!rc && hc[1])
            rc = hc[1] <= spillCap ? launchTableInsert(t, spill, hc[1], s)
                                   : launchTableInsert(t, dKeys, n, s);
    }
    cleanup();
    if (!rc)
        *used = true;
    return rc;
}

// The buffers kept for the next counter or table build, freed.
void releaseSortCaches()
{
    hipibs_counter_release_cache();
}
} // namespace hipibs

extern "C" int hipibs_counter_create(uint32_t k, int device, uint64_t expected_keys,
                                     uint64_t chunk_bytes, hipibs_counter **out)
{
    if (!out)
        return setError(HIPIBS_E_ARG, "null argument");
    *out = nullptr;
    if (k < 1 || k > 64)
        return setError(HIPIBS_E_ARG, "k-mer counter: k must be 1..64");
    if (chunk_bytes == 0)
        chunk_bytes = kDefaultChunk;
    if (chunk_bytes < 64 || chunk_bytes > (1ull << 30))
        return setError(HIPIBS_E_ARG, "k-mer counter: chunk_bytes must be 64 B .. 1 GiB");
    int nDev = 0;
    if (hipGetDeviceCount(&nDev) != hipSuccess || nDev == 0)
    {
        (void)hipGetLastError();
        return setError(HIPIBS_E_NODEV, "no usable GPU");
    }
    if (device < 0 || device >= nDev)
        return setError(HIPIBS_E_ARG, "no such device");
    HIPIBS_CHECK(hipSetDevice(device));
    auto *c = new hipibs_counter();
    c->device = device;
    c->k = k;
    c->wide = k > 32;
    c->chunkBytes = chunk_bytes;
    if (const char *sc = std::getenv("HIPIBS_COUNT_SKETCH_SCALE"))
        c->sketchScale = std::min(100.0, std::max(0.0, std::atof(sc)));
    if (const char *db = std::getenv("HIPIBS_SORT_DIGIT_BITS"))
        c->sorter.digitBits = (uint32_t)std::min(kMaxDigitBits, std::max(1, std::atoi(db)));
    if (const char *lw = std::getenv("HIPIBS_SORT_LOOKBACK_WIDE"))
        c->sorter.wideLookBack = std::atoi(lw) != 0;
    auto fail = [c](int rc)
    {
        hipibs_counter_free(c);
        return rc;
    };
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&c->dCounters), 16 * sizeof(unsigned long long));
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void **>(&c->hCounters), 16 * sizeof(unsigned long long),
                          hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipMemset(c->dCounters, 0, 16 * sizeof(unsigned long long));
    if (e != hipSuccess)
        return fail(hipFail(e, "k-mer counter buffers"));
    std::memset(c->hCounters, 0, 16 * sizeof(unsigned long long));
    if (expected_keys)
    {
        const int rc = c->reserve(expected_keys);
        if (rc)
            return fail(rc);
    }
    *out = c;
    return 0;
}

extern "C" int hipibs_counter_add_text(hipibs_counter *c, const char *h_text, uint64_t n)
{
    if (!c || (n && !h_text))
        return setError(HIPIBS_E_ARG, "null argument");
    if (c->err)
        return setError(c->err, "k-mer counter: an earlier call failed");
    HIPIBS_CHECK(hipSetDevice(c->device));
    if ((c->err = c->hostBuffers()))
        return c->err;
    c->bases(h_text, n);
    c->recordEnd();
    return c->err;
}

extern "C" int hipibs_counter_add_text_device(hipibs_counter *c, const char *d_text, uint64_t n)
{
    if (!c || (n && !d_text))
        return setError(HIPIBS_E_ARG, "null argument");
    if (reinterpret_cast<uintptr_t>(d_text) & 15)
        return setError(HIPIBS_E_ARG, "k-mer counter: device text must be 16 B aligned");
    if (c->err)
        return setError(c->err, "k-mer counter: an earlier call failed");
    HIPIBS_CHECK(hipSetDevice(c->device));
    int rc = c->drain(); // text buffered by earlier calls goes first
    if (!rc)
        rc = c->countDevice(reinterpret_cast<const uint8_t *>(d_text), n);
    if (rc)
        return c->err = rc;
    c->reads += 1;
    return 0;
}

extern "C" int hipibs_counter_add_file(hipibs_counter *c, const char *path)
{
    if (!c || !path)
        return setError(HIPIBS_E_ARG, "null argument");
    if (c->err)
        return setError(c->err, "k-mer counter: an earlier call failed");
    HIPIBS_CHECK(hipSetDevice(c->device));
    if ((c->err = c->hostBuffers()))
        return c->err;
    const int rc = readRecords(path, *c, nullptr);
    if (rc && !c->err)
        c->err = rc; // the file's records before the error may be counted: the counter is spent
    return rc;
}

extern "C" int hipibs_counter_stats(hipibs_counter *c, hipibs_count_stats *out)
{
    if (!c || !out)
        return setError(HIPIBS_E_ARG, "null argument");
    HIPIBS_CHECK(hipSetDevice(c->device));
    const int rc = c->finish();
    if (rc)
        return rc;
    out->reads = c->reads;
    out->kmers = c->counted;
    out->distinct = c->distinct;
    out->kernel_ms = c->kernelMs;
    out->launches = c->launches;
    out->flushes = c->flushes;
    out->spilled = c->spilled;
    for (int i = 0; i < 4; ++i)
        out->stage_ms[i] = c->stageMs[i];
    out->table_bytes = c->nBuckets * kCountWords * 8;
    return 0;
}

extern "C" int hipibs_counter_keys(hipibs_counter *c, uint32_t min_count, uint32_t max_count,
                                   uint32_t counter_max, uint64_t *h_keys, uint32_t *h_counts,
                                   uint64_t cap, hipibs_count_filter *filt)
{
    if (!c || !filt || (cap && !h_keys))
        return setError(HIPIBS_E_ARG, "null argument");
    int rc = checkFilter(min_count, max_count, counter_max);
    if (rc)
        return rc;
    HIPIBS_CHECK(hipSetDevice(c->device));
    if ((rc = c->finish()))
        return rc;
    const uint64_t m = std::min<uint64_t>(cap, c->distinct);
    uint64_t *dKeys = nullptr;
    uint32_t *dCounts = nullptr;
    hipError_t e = hipSuccess;
    if (m)
    {
        e = hipMalloc(reinterpret_cast<void **>(&dKeys), m * 8 * (c->wide ? 2 : 1));
        if (e == hipSuccess && h_counts)
            e = hipMalloc(reinterpret_cast<void **>(&dCounts), m * 4);
    }
    if (e == hipSuccess)
        rc = c->extract(min_count, max_count, counter_max, dKeys, dCounts, m, filt);
    const uint64_t got = std::min<uint64_t>(m, filt->kept);
    if (e == hipSuccess && !rc && got)
    {
        e = hipMemcpy(h_keys, dKeys, got * 8 * (c->wide ? 2 : 1), hipMemcpyDeviceToHost);
        if (e == hipSuccess && h_counts)
            e = hipMemcpy(h_counts, dCounts, got * 4, hipMemcpyDeviceToHost);
    }
    (void)hipFree(dKeys);
    (void)hipFree(dCounts);
    if (rc)
        return rc;
    if (e != hipSuccess)
        return hipFail(e, "copying counted k-mers");
    return 0;
}

extern "C" int hipibs_counter_build_table(hipibs_counter *c, uint32_t min_count,
                                          uint32_t max_count, const hipibs_table_opts *opts,
                                          hipibs_table **out, hipibs_count_filter *filt)
{
    if (!c || !out)
        return setError(HIPIBS_E_ARG, "null argument");
    *out = nullptr;
    int rc = checkFilter(min_count, max_count, 1);
    if (rc)
        return rc;
    hipibs_table_opts o{};
    if (opts)
        o = *opts;
    else
        o.device = c->device;
    if (o.device != c->device)
        return setError(HIPIBS_E_ARG, "k-mer counter: the table must be on the counter's device");
    HIPIBS_CHECK(hipSetDevice(c->device));
    if ((rc = c->finish()))
        return rc;
    hipibs_count_filter f{};
    uint64_t *dKeys = nullptr;
    if (c->distinct)
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&dKeys),
                               c->distinct * 8 * (c->wide ? 2 : 1)));
    rc = c->extract(min_count, max_count, 1, dKeys, nullptr, c->distinct, &f);
    if (!rc)
        rc = hipibs_table_build_keys_device(dKeys, f.kept, c->k, &o, out, c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dKeys);
    if (filt)
        *filt = f;
    return rc;
}

// KMC's counter width for -cs: the bytes that hold counter_max.
static uint32_t counterBytes(uint32_t cmax)
{
    uint32_t b = 1;
    while (b < 4 && (uint64_t)cmax >= (1ull << (8 * b)))
        ++b;
    return b;
}

// Prefix length of the written database: (k - p) % 4 == 0 and p <= 12, the
// one with the smallest files (records of (k - p) / 4 + counter bytes each,
// plus a LUT of 8 B per prefix per signature bin).
static uint32_t lutPrefixLen(uint64_t n, uint32_t k, uint32_t cbytes, uint32_t bins)
{
    uint32_t best = k % 4;
    double bestBytes = -1;
    for (uint32_t p = k % 4; p <= std::min<uint32_t>(12, k); p += 4)
    {
        const double bytes = (double)n * ((k - p) / 4 + cbytes) +
                             (double)(1ull << (2 * p)) * bins * 8.0;
        if (bestBytes < 0 || bytes < bestBytes)
        {
            bestBytes = bytes;
            best = p;
        }
    }
    return best;
}

extern "C" int hipibs_counter_write_kmc(hipibs_counter *c, const char *prefix, uint32_t min_count,
                                        uint32_t max_count, uint32_t counter_max,
                                        hipibs_count_filter *filt)
{
    if (!c || !prefix)
        return setError(HIPIBS_E_ARG, "null argument");
    int rc = checkFilter(min_count, max_count, counter_max);
    if (rc)
        return rc;
    HIPIBS_CHECK(hipSetDevice(c->device));
    if ((rc = c->finish()))
        return rc;
    std::vector<uint64_t> keys(c->distinct * (c->wide ? 2 : 1));
    std::vector<uint32_t> counts(c->distinct);
    hipibs_count_filter f{};
    rc = hipibs_counter_keys(c, min_count, max_count, counter_max, keys.data(), counts.data(),
                             c->distinct, &f);
    if (rc)
        return rc;
    if (filt)
        *filt = f;
    // KMC2 layout with 9-base signatures (KMC's default); k < 9 shortens
    // them, and k < 5 writes the KMC1 layout, which has none.
    const uint32_t sig = std::min<uint32_t>(9, c->k);
    const uint32_t version = sig >= 5 ? 0x200 : 0;
    const uint32_t cbytes = counterBytes(counter_max);
    const uint32_t p = lutPrefixLen(f.kept, c->k, cbytes, version ? 512 : 1);
    return kmcWrite(prefix, keys.data(), counts.data(), f.kept, c->k, p, version ? sig : 0, cbytes,
                    version, min_count, max_count);
}

extern "C" void hipibs_counter_free(hipibs_counter *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    if (c->dTab)
        returnCountTable(c->device, c->dTab, c->tabBytes);
    if (c->dCounters)
        (void)hipFree(c->dCounters);
    if (c->dText)
        (void)hipFree(c->dText);
    if (c->dHash[0])
        returnHash(c->device, c->dHash, c->hashCap);
    if (c->dHll)
        (void)hipFree(c->dHll);
    c->sorter.release();
    for (char *h : c->hText)
        if (h)
            (void)hipHostFree(h);
    if (c->hCounters)
        (void)hipHostFree(c->hCounters);
    for (hipEvent_t e : c->evPool)
        (void)hipEventDestroy(e);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" void hipibs_counter_release_cache(void)
{
    releaseBuildSorter();
    std::lock_guard<std::mutex> g(gHashCache.m);
    if (!gHashCache.p[0])
        return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(gHashCache.device);
    for (int i = 0; i < 2; ++i)
        (void)hipFree(gHashCache.p[i]);
    (void)hipSetDevice(cur);
    gHashCache.p[0] = gHashCache.p[1] = nullptr;
    gHashCache.cap = 0;
}
