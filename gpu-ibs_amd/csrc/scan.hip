// The fused scan: encode + canonicalise + probe + per-window run-length
// reduction, plus the mapper's hit-mask and coverage kernels.
//
// Work decomposition. A block (4 waves) owns one segment of one window
// against one table (grid.y = table). The segment's positions are walked
// in 64-position words aligned to the packed stream (word m = positions
// 64m .. 64m+63); each wave owns a contiguous quarter of the words. For a
// word, the wave reads 3 sequence words and 2 exception words with scalar
// loads (every lane needs bases from the same <= 96-base span), each lane
// cuts its k-mer out of them with a funnel shift, forms the reverse
// complement as the complement of the low-first packing and the forward
// code by reversing the 2-bit groups, and the 64 canonical codes are probed
// cooperatively (wave_probe.hpp). The hit and valid masks are ballots, so
// the run-length statistics of the word follow from bit arithmetic: each
// hit lane measures the miss run that it closes (misses between the
// previous hit of the word and itself, counted in valid positions only, so
// a run continues across N bases exactly as the reference's compacted code
// stream does: src/KmerCodec.hpp:38-61, src/KmerDatabase.hpp:596-672); the
// run that crosses the word seam is carried in a wave-uniform counter. The
// four wave summaries are folded in LDS with the associative combine of
// hipibs_internal.hpp, and windows split over several blocks are folded by
// a second, tiny kernel.
#include "wave_probe.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace hipibs
{

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_WAVES = SCAN_BLOCK / 64;
// 64-position words per step of a minimizer-table scan (two words per step
// keep twice the lines in flight per wave but cost a wave per SIMD: slower).
constexpr int kGroupWords = 1;

// MIXED: every window names its own table (hipibs_window.reserved) and its
// result goes to row w; otherwise grid.y is the table and results are
// table-major. WIDE: k = 33..64, two-word codes (wideCanonical) probed in a
// table of 16 B keys (waveProbeWide); a fourth sequence word is read.
// GROUPED: every table is a minimizer table, probed with waveProbeGrouped
// (each run of positions sharing a minimizer loads its home line once).
template <int S, bool MIXED, bool WIDE = false, bool GROUPED = false>
__global__ __launch_bounds__(SCAN_BLOCK) void scanKernel(
    const TableView *__restrict__ tables, const uint64_t *__restrict__ seq,
    const uint64_t *__restrict__ exc, uint64_t nBases, const hipibs_window *__restrict__ windows,
    uint32_t nWindows, uint32_t segs, uint32_t k, hipibs_window_stats *__restrict__ stats,
    SegSummary *__restrict__ work)
{
    __shared__ SegSummary waveSum[SCAN_WAVES];
    __shared__ __align__(16) uint64_t groupStage[GROUPED ? SCAN_WAVES : 1]
                                                [GROUPED ? GroupStage<S, kGroupWords>::kWords : 2];
    const uint32_t unit = blockIdx.x;
    const uint32_t w = unit / segs, s = unit - w * segs;
    const hipibs_window win = windows[w];
    const uint32_t t = MIXED ? win.reserved : blockIdx.y;
    const TableView T = tables[t];
    const uint64_t row = MIXED ? (uint64_t)w : (uint64_t)t * nWindows + w;
    uint64_t g0, g1;
    segmentRange(win, nBases, k, s, segs, g0, g1);

    const int lane = laneId();
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint64_t kmask1 = k >= 64 ? ~0ull : (1ull << k) - 1ull;
    const uint64_t kmask2 = k >= 32 ? ~0ull : (1ull << (2 * k)) - 1ull;
    const uint32_t fwdShift = k >= 32 ? 0u : 64 - 2 * k;

    uint32_t nV = 0, nH = 0, lead = 0, T0 = 0, uVar = 0, uDist = 0;
    uint32_t lVar = 0, lDist = 0;
    bool seen = false;
    if (g0 < g1)
    {
        const uint64_t mFirst = g0 >> 6, mLast = (g1 - 1) >> 6;
        const uint64_t nWords = mLast - mFirst + 1;
        const uint64_t per = (nWords + SCAN_WAVES - 1) / SCAN_WAVES;
        const uint64_t mBeg = mFirst + std::min<uint64_t>(nWords, (uint64_t)wv * per);
        const uint64_t mEnd = mFirst + std::min<uint64_t>(nWords, (uint64_t)(wv + 1) * per);
        const int j = lane & 31;
        const uint32_t mlen = T.mlen;
        const uint64_t maskm2 = mlen ? (1ull << (2 * mlen)) - 1ull : 0ull;
        const uint32_t mShift = 64 - 2 * mlen;
        const int win = (int)(k - mlen + 1); // m-mers per k-mer
        // Canonical code, validity and (minimizer tables) minimizer hash of
        // position 64m + lane.
        auto encode = [&](uint64_t m, uint64_t &canon, bool &valid, uint32_t &minh) {
            const uint64_t s0 = seq[2 * m], s1 = seq[2 * m + 1], s2 = seq[2 * m + 2];
            const uint64_t e0 = exc[m], e1 = exc[m + 1];
            const uint64_t a = lane < 32 ? s0 : s1;
            const uint64_t b = lane < 32 ? s1 : s2;
            const uint64_t x = (j ? ((a >> (2 * j)) | (b << (64 - 2 * j))) : a) & kmask2;
            const uint64_t ev = lane ? ((e0 >> lane) | (e1 << (64 - lane))) : e0;
            const uint64_t p = (m << 6) + (uint64_t)lane;
            valid = ((ev & kmask1) == 0) && p >= g0 && p < g1;
            const uint64_t rc = ~x & kmask2;
            const uint64_t fwd = reverse2(x) >> fwdShift;
            canon = fwd < rc ? fwd : rc;
            minh = 0;
            if (GROUPED || mlen)
            {
                // canonical m-mer hashes at positions p and p+64, then the
                // minimum over each k-mer's window of `win` m-mers by a
                // doubling sliding minimum across the wave
                const uint64_t s3 = seq[2 * m + 3], s4 = seq[2 * m + 4];
                const uint64_t a2 = lane < 32 ? s2 : s3;
                const uint64_t b2 = lane < 32 ? s3 : s4;
                uint64_t y = j ? ((a2 >> (2 * j)) | (b2 << (64 - 2 * j))) : a2;
                y &= maskm2;
                // the m-mer at p is the k-mer's first m bases: its forward
                // code is the top of fwd, its reverse complement the bottom
                // of rc
                const uint64_t flo = fwd >> (2 * (k - mlen)), rlo = rc & maskm2;
                const uint64_t rhi = ~y & maskm2, fhi = reverse2(y) >> mShift;
                minh = waveWindowMin(mmerHash(flo < rlo ? flo : rlo),
                                     mmerHash(fhi < rhi ? fhi : rhi), win);
            }
        };
        // Canonical code, validity and the hash of the m-mer at position
        // 64m + lane (the k-mer's first m bases: the top of its forward code,
        // the bottom of its reverse complement).
        auto encodeFirst = [&](uint64_t m, uint64_t &canon, bool &valid, uint32_t &L) {
            const uint64_t s0 = seq[2 * m], s1 = seq[2 * m + 1], s2 = seq[2 * m + 2];
            const uint64_t e0 = exc[m], e1 = exc[m + 1];
            const uint64_t a = lane < 32 ? s0 : s1;
            const uint64_t b = lane < 32 ? s1 : s2;
            const uint64_t x = (j ? ((a >> (2 * j)) | (b << (64 - 2 * j))) : a) & kmask2;
            const uint64_t ev = lane ? ((e0 >> lane) | (e1 << (64 - lane))) : e0;
            const uint64_t p = (m << 6) + (uint64_t)lane;
            valid = ((ev & kmask1) == 0) && p >= g0 && p < g1;
            const uint64_t rc = ~x & kmask2;
            const uint64_t fwd = reverse2(x) >> fwdShift;
            canon = fwd < rc ? fwd : rc;
            const uint64_t flo = fwd >> (2 * (k - mlen)), rlo = rc & maskm2;
            L = mmerHash(flo < rlo ? flo : rlo);
        };
        // Run-length statistics of one word from its hit and valid masks.
        auto fold = [&](uint64_t H, uint64_t V) {
            const uint64_t Miss = V & ~H;
            if (H)
            {
                const int f = __builtin_ctzll(H);
                const int l = 63 - __builtin_clzll(H);
                const uint32_t firstGap = T0 + (uint32_t)__popcll(Miss & ((1ull << f) - 1ull));
                if (seen)
                {
                    if (firstGap)
                    {
                        uVar += 1;
                        uDist += (uint32_t)gapDistance((int32_t)firstGap, (int32_t)k);
                    }
                }
                else
                {
                    lead = firstGap;
                    seen = true;
                }
                T0 = (uint32_t)__popcll(Miss >> l);
                // runs closed inside the word, one per hit lane with an
                // earlier hit in the word
                if (((H >> lane) & 1ull) && maskedRank(H))
                {
                    // the previous hit: H's highest bit below this lane
                    // (lane >= 1 here), moved to the top of the word
                    const int prev = lane - 1 - __builtin_clzll(H << (64 - lane));
                    // misses between the previous hit and this lane
                    const uint32_t gap = maskedRank(Miss & ~((2ull << prev) - 1ull));
                    if (gap)
                    {
                        lVar += 1;
                        lDist += (uint32_t)gapDistance((int32_t)gap, (int32_t)k);
                    }
                }
            }
            else
                T0 += (uint32_t)__popcll(Miss);
            nV += (uint32_t)__popcll(V);
            nH += (uint32_t)__popcll(H);
        };
        if constexpr (GROUPED && kGroupWords == 1)
        {
            // one word behind: the next word's first m-mers (encoded for its
            // own step) end this word's windows. Word mEnd is read within
            // the packed arrays' pad words.
            uint64_t canon[1] = {0};
            uint32_t L = 0, minh[1] = {0};
            bool valid[1] = {false};
            if (mBeg < mEnd)
                encodeFirst(mBeg, canon[0], valid[0], L);
            for (uint64_t m = mBeg; m < mEnd; ++m)
            {
                uint64_t canonN;
                bool validN;
                uint32_t LN;
                encodeFirst(m + 1, canonN, validN, LN);
                minh[0] = waveWindowMin(L, LN, win);
                uint64_t H[1];
                waveProbeGrouped<S, 1>(T, canon, minh, valid, groupStage[wv], H);
                fold(H[0], ballot(valid[0]));
                canon[0] = canonN;
                valid[0] = validN;
                L = LN;
            }
        }
        else if constexpr (GROUPED)
        {
            // kGroupWords words per step: their home lines are in flight
            // together
            for (uint64_t m = mBeg; m < mEnd; m += kGroupWords)
            {
                uint64_t canon[kGroupWords], H[kGroupWords];
                uint32_t minh[kGroupWords];
                bool valid[kGroupWords];
#pragma unroll
                for (int i = 0; i < kGroupWords; ++i)
                {
                    canon[i] = 0;
                    minh[i] = 0;
                    valid[i] = false;
                    if (m + i < mEnd)
                        encode(m + i, canon[i], valid[i], minh[i]);
                }
                waveProbeGrouped<S, kGroupWords>(T, canon, minh, valid, groupStage[wv], H);
#pragma unroll
                for (int i = 0; i < kGroupWords; ++i)
                    if (m + i < mEnd)
                        fold(H[i], ballot(valid[i]));
            }
        }
        else
            for (uint64_t m = mBeg; m < mEnd; ++m)
            {
                uint64_t H;
                bool valid;
                if constexpr (WIDE)
                {
                    const uint64_t s0 = seq[2 * m], s1 = seq[2 * m + 1], s2 = seq[2 * m + 2];
                    const uint64_t s3 = seq[2 * m + 3];
                    const uint64_t e0 = exc[m], e1 = exc[m + 1];
                    const uint64_t a = lane < 32 ? s0 : s1;
                    const uint64_t b = lane < 32 ? s1 : s2;
                    const uint64_t c = lane < 32 ? s2 : s3;
                    const uint64_t x0 = j ? ((a >> (2 * j)) | (b << (64 - 2 * j))) : a;
                    const uint64_t x1 = j ? ((b >> (2 * j)) | (c << (64 - 2 * j))) : b;
                    const uint64_t ev = lane ? ((e0 >> lane) | (e1 << (64 - lane))) : e0;
                    const uint64_t p = (m << 6) + (uint64_t)lane;
                    valid = ((ev & kmask1) == 0) && p >= g0 && p < g1;
                    uint64_t hi, lo;
                    wideCanonical(x0, x1, k, hi, lo);
                    H = waveProbeWide(T, hi, lo,
                                      valid ? (uint32_t)bucketOfWide(hi, lo, T.nBuckets) : 0u,
                                      valid);
                }
                else
                {
                    uint64_t canon;
                    uint32_t minh;
                    encode(m, canon, valid, minh);
                    if (mlen)
                        H = waveProbeMinimizer<S>(T, canon, minh, valid);
                    else
                        H = waveProbe<S>(T, canon, (uint32_t)bucketOf(canon, T.nBuckets), valid);
                }
                fold(H, ballot(valid));
            }
    }
    lVar = waveSumU32(lVar);
    lDist = waveSumU32(lDist);
    if (lane == 0)
    {
        SegSummary r;
        r.nValid = nV;
        r.nHit = nH;
        r.lead = seen ? lead : nV;
        r.trail = T0;
        r.innerVar = uVar + lVar;
        r.innerDist = uDist + lDist;
        waveSum[wv] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        SegSummary acc = waveSum[0];
#pragma unroll
        for (int i = 1; i < SCAN_WAVES; ++i)
            acc = combine(acc, waveSum[i], (int32_t)k);
        if (segs == 1)
            stats[row] = finalizeSummary(acc, (int32_t)k);
        else
            work[row * segs + s] = acc;
    }
}

// Folds the segs summaries of each (table, window) in order.
__global__ void foldKernel(const SegSummary *__restrict__ work, uint32_t nItems, uint32_t segs,
                           uint32_t k, hipibs_window_stats *__restrict__ stats)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nItems)
        return;
    const SegSummary *p = work + (uint64_t)i * segs;
    SegSummary acc = p[0];
    for (uint32_t s = 1; s < segs; ++s)
        acc = combine(acc, p[s], (int32_t)k);
    stats[i] = finalizeSummary(acc, (int32_t)k);
}

// Mapper: hit mask over positions pos0 .. pos0+nPos-1 (bit i = position
// pos0+i), one 64-position word per wave step, each lane gathering its own
// k-mer (positions need not be aligned to the packed words). GROUPED: a
// minimizer table, probed with waveProbeGrouped (the wave's lanes hold
// consecutive positions, so runs of them share a minimizer).
template <int S, bool WIDE = false, bool GROUPED = false>
__global__ __launch_bounds__(256) void hitmaskKernel(const TableView T, const uint64_t *seq,
                                                     const uint64_t *exc, uint64_t pos0,
                                                     uint64_t nPos, uint64_t contigEnd, uint32_t k,
                                                     uint32_t *hitWords)
{
    __shared__ __align__(16) uint64_t
        groupStage[GROUPED ? 4 : 1][GROUPED ? GroupStage<S, 1>::kWords : 2];
    const int lane = laneId();
    const int wv = GROUPED ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nWaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t nChunks = (nPos + 63) >> 6;
    const uint64_t kmask1 = k >= 64 ? ~0ull : (1ull << k) - 1ull;
    const uint64_t kmask2 = k >= 32 ? ~0ull : (1ull << (2 * k)) - 1ull;
    const uint32_t fwdShift = k >= 32 ? 0u : 64 - 2 * k;
    const uint64_t nOutWords = (nPos + 31) >> 5;
    for (uint64_t c = wave; c < nChunks; c += nWaves)
    {
        const uint64_t i = (c << 6) + (uint64_t)lane;
        const uint64_t p = pos0 + i;
        bool valid = i < nPos && p + k <= contigEnd;
        uint64_t canon = 0, chi = 0;
        if (valid)
        {
            const uint64_t wi = p >> 5;
            const uint32_t sh = (uint32_t)(p & 31) * 2;
            const uint64_t a = seq[wi], b = seq[wi + 1];
            uint64_t x = sh ? ((a >> sh) | (b << (64 - sh))) : a;
            const uint64_t ej = p >> 6;
            const uint32_t es = (uint32_t)(p & 63);
            const uint64_t ea = exc[ej], eb = exc[ej + 1];
            const uint64_t ev = es ? ((ea >> es) | (eb << (64 - es))) : ea;
            valid = (ev & kmask1) == 0;
            if constexpr (WIDE)
            {
                const uint64_t c2 = seq[wi + 2];
                const uint64_t x1 = sh ? ((b >> sh) | (c2 << (64 - sh))) : b;
                wideCanonical(x, x1, k, chi, canon);
            }
            else
            {
                x &= kmask2;
                const uint64_t rc = ~x & kmask2;
                const uint64_t fwd = reverse2(x) >> fwdShift;
                canon = fwd < rc ? fwd : rc;
            }
        }
        uint64_t H;
        if constexpr (WIDE)
            H = waveProbeWide(T, chi, canon,
                              valid ? (uint32_t)bucketOfWide(chi, canon, T.nBuckets) : 0u, valid);
        else if constexpr (GROUPED)
        {
            const uint64_t code[1] = {canon};
            const uint32_t minh[1] = {valid ? minimizerOf(canon, k, T.mlen) : 0u};
            const bool v[1] = {valid};
            uint64_t Hw[1];
            waveProbeGrouped<S, 1>(T, code, minh, v, groupStage[wv], Hw);
            H = Hw[0];
        }
        else
            H = waveLookup<S>(T, canon, valid);
        if (lane < 2)
        {
            const uint64_t o = 2 * c + (uint64_t)lane;
            if (o < nOutWords)
                hitWords[o] = (uint32_t)(H >> (32 * lane));
        }
    }
}

// Bit j of the low 8 bits of b as byte j (0 or 1) of a 64-bit word.
__device__ inline uint64_t bitsToBytes(uint32_t b)
{
    const uint64_t x = ((uint64_t)(b & 0xFFu) * 0x0101010101010101ull) & 0x8040201008040201ull;
    const uint64_t y = ((x & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | x;
    return (y >> 7) & 0x0101010101010101ull;
}

// Running sums of the bytes of v (byte j = sum of bytes 0..j; sums < 256).
__device__ inline uint64_t bytePrefix(uint64_t v)
{
    v += v << 8;
    v += v << 16;
    return v + (v << 32);
}

// coverage[i] = popcount of hit bits p in [i-k+1, i] (clamped to the
// positions), as a byte (k <= 64). A wave takes 1024 consecutive bases,
// 16 per lane (bases i0..i0+15 of lane l, i0 = wave base + 16 l): lane l
// loads hit word (wave base)/32 - 2 + l, and each lane takes the four
// words of positions i0-64 .. i0+63 from its neighbours (ds_bpermute).
// The lane's first base counts its window; the other 15 come in bytes at
// once -- coverage[i0+t] = coverage[i0] + (hits entering at i0+1..i0+t) -
// (hits leaving at i0+1-k..i0+t-k), both running sums of bits spread to
// bytes -- and the 16 bytes are one 16 B store (consecutive lanes,
// consecutive stores).
constexpr uint32_t kCovPerWave = 1024;
__global__ __launch_bounds__(256) void coverageKernel(const uint32_t *__restrict__ hitWords,
                                                      uint64_t nPos, uint64_t nBases, uint32_t k,
                                                      uint8_t *__restrict__ cov)
{
    const int lane = laneId();
    const uint64_t nWaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t nWords = (nPos + 31) >> 5;
    const uint64_t nChunks = (nBases + kCovPerWave - 1) / kCovPerWave;
    const bool vec = (reinterpret_cast<uintptr_t>(cov) & 15) == 0;
    for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < nChunks;
         c += nWaves)
    {
        const uint64_t W0 = c * kCovPerWave;
        // word (W0/32) - 2 + lane (positions < 0 or >= nPos read as 0)
        const int64_t w = (int64_t)(W0 >> 5) - 2 + lane;
        uint32_t v = 0;
        if (w >= 0 && (uint64_t)w < nWords)
        {
            v = hitWords[w];
            const int64_t last = (int64_t)nPos - 32 * w; // positions of word w below nPos
            if (last < 32)
                v &= (1u << last) - 1u;
        }
        const uint64_t i0 = W0 + 16 * (uint64_t)lane;
        // this lane's words (W0/32) + lane/2 - 2 .. +1 are lanes lane/2 .. lane/2+3
        uint64_t wv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            wv[q] = shuffle32(v, (lane >> 1) + q);
        // L: positions i0-64 .. i0-1, R: positions i0 .. i0+63 (bit 0 = the first)
        uint64_t L, R;
        if (lane & 1)
        {
            L = (wv[0] >> 16) | (wv[1] << 16) | (wv[2] << 48);
            R = (wv[2] >> 16) | (wv[3] << 16);
        }
        else
        {
            L = wv[0] | (wv[1] << 32);
            R = wv[2] | (wv[3] << 32);
        }
        const uint64_t c0 = (R & 1ull) + (k > 1 ? (uint64_t)__popcll(L >> (65 - k)) : 0ull);
        const uint32_t in = (uint32_t)R & 0xFFFEu;
        const uint32_t out = (uint32_t)(k == 64 ? L : (L >> (64 - k)) | (R << k)) & 0xFFFEu;
        const uint64_t in0 = bytePrefix(bitsToBytes(in)), out0 = bytePrefix(bitsToBytes(out));
        const uint64_t in1 = bytePrefix(bitsToBytes(in >> 8)) + (in0 >> 56) * 0x0101010101010101ull;
        const uint64_t out1 =
            bytePrefix(bitsToBytes(out >> 8)) + (out0 >> 56) * 0x0101010101010101ull;
        const uint64_t base = c0 * 0x0101010101010101ull;
        // (base + entered) >= left in every byte: no borrow crosses a byte
        const uint64_t lo = base + in0 - out0, hi = base + in1 - out1;
        if (vec && i0 + 16 <= nBases)
            *reinterpret_cast<uint4 *>(cov + i0) =
                make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
        else
            for (uint64_t j = 0; j < 16 && i0 + j < nBases; ++j)
                cov[i0 + j] = (uint8_t)((j < 8 ? lo >> (8 * j) : hi >> (8 * (j - 8))) & 0xFF);
    }
}

// Canonical code per position (all ones when invalid).
template <bool WIDE>
__global__ void encodeKernel(const uint64_t *__restrict__ seq, const uint64_t *__restrict__ exc,
                             uint64_t pos0, uint64_t nPos, uint64_t contigEnd, uint32_t k,
                             uint64_t *__restrict__ codes, uint8_t *__restrict__ valid)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t kmask1 = k >= 64 ? ~0ull : (1ull << k) - 1ull;
    const uint64_t kmask2 = k >= 32 ? ~0ull : (1ull << (2 * k)) - 1ull;
    const uint32_t fwdShift = k >= 32 ? 0u : 64 - 2 * k;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nPos; i += stride)
    {
        const uint64_t p = pos0 + i;
        bool ok = p + k <= contigEnd;
        uint64_t canon = EMPTY, hi = EMPTY;
        if (ok)
        {
            const uint64_t wi = p >> 5;
            const uint32_t sh = (uint32_t)(p & 31) * 2;
            const uint64_t a = seq[wi], b = seq[wi + 1];
            uint64_t x = sh ? ((a >> sh) | (b << (64 - sh))) : a;
            x &= kmask2;
            const uint64_t ej = p >> 6;
            const uint32_t es = (uint32_t)(p & 63);
            const uint64_t ea = exc[ej], eb = exc[ej + 1];
            const uint64_t ev = es ? ((ea >> es) | (eb << (64 - es))) : ea;
            ok = (ev & kmask1) == 0;
            if (ok && WIDE)
            {
                const uint64_t x0 = sh ? ((a >> sh) | (b << (64 - sh))) : a;
                const uint64_t c = seq[wi + 2];
                const uint64_t x1 = sh ? ((b >> sh) | (c << (64 - sh))) : b;
                wideCanonical(x0, x1, k, hi, canon);
            }
            else if (ok)
            {
                const uint64_t rc = ~x & kmask2;
                const uint64_t fwd = reverse2(x) >> fwdShift;
                canon = fwd < rc ? fwd : rc;
            }
        }
        if (WIDE)
        {
            codes[2 * i] = hi;
            codes[2 * i + 1] = canon;
        }
        else
            codes[i] = canon;
        if (valid)
            valid[i] = ok ? 1 : 0;
    }
}

static int deviceCUs(int device)
{
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    return cus;
}

uint32_t autoSegsPerWindow(uint32_t nWindows, int nTables, uint32_t maxPositions)
{
    // Aim for >= 8192 blocks (32 per CU on a 256-CU part) without cutting a
    // segment below 4096 positions.
    const uint64_t units = (uint64_t)std::max<uint32_t>(nWindows, 1) * std::max(nTables, 1);
    uint64_t segs = (8192 + units - 1) / units;
    const uint64_t maxSegs = std::max<uint64_t>(1, (uint64_t)maxPositions / 4096);
    segs = std::min(segs, maxSegs);
    return (uint32_t)std::max<uint64_t>(1, segs);
}

uint64_t scanWorkBytes(uint32_t nWindows, int nTables, uint32_t segs)
{
    if (segs <= 1)
        return 0;
    return (uint64_t)nWindows * nTables * segs * sizeof(SegSummary) + 64;
}

// Per-thread profiling of the scan kernel (HIP events on the launch stream).
static thread_local bool tProfile = false;
static thread_local hipEvent_t tEv[2] = {nullptr, nullptr};
static thread_local bool tEvRecorded = false;

int launchScan(const hipibs_table *const *tables, int nTables, const uint64_t *dSeq,
               const uint64_t *dExc, uint64_t nBases, const hipibs_window *dWindows,
               uint32_t nWindows, uint32_t maxWindowBases, hipibs_window_stats *dStats,
               uint32_t segsPerWindow, void *dWork, uint64_t workBytes, hipStream_t s,
               TableView *dViews, bool profile)
{
    const uint32_t k = tables[0]->k;
    const uint32_t maxPos = maxWindowBases >= k ? maxWindowBases - k + 1 : 0;
    uint32_t segs = segsPerWindow ? segsPerWindow : autoSegsPerWindow(nWindows, nTables, maxPos);
    if (segs > 1 && (!dWork || workBytes < scanWorkBytes(nWindows, nTables, segs)))
        return setError(HIPIBS_E_ARG, "scan scratch too small for the segment split");
    if ((uint64_t)nWindows * segs >= (1ull << 31) || nTables > 65535)
        return setError(HIPIBS_E_ARG, "too many windows or tables for one launch");
    if (nWindows == 0)
        return 0;
    const dim3 grid(nWindows * segs, (uint32_t)nTables);
    auto *work = reinterpret_cast<SegSummary *>(dWork);
    profile = profile && tProfile;
    if (profile)
    {
        if (!tEv[0])
        {
            HIPIBS_CHECK(hipEventCreate(&tEv[0]));
            HIPIBS_CHECK(hipEventCreate(&tEv[1]));
        }
        HIPIBS_CHECK(hipEventRecord(tEv[0], s));
    }
    if (k > 32)
        hipLaunchKernelGGL((scanKernel<16, false, true>), grid, dim3(SCAN_BLOCK), 0, s, dViews,
                           dSeq, dExc, nBases, dWindows, nWindows, segs, k, dStats, work);
    else if (tables[0]->mlen && tables[0]->bucketSlots == 8)
        hipLaunchKernelGGL((scanKernel<8, false, false, true>), grid, dim3(SCAN_BLOCK), 0, s,
                           dViews, dSeq, dExc, nBases, dWindows, nWindows, segs, k, dStats, work);
    else if (tables[0]->mlen)
        hipLaunchKernelGGL((scanKernel<16, false, false, true>), grid, dim3(SCAN_BLOCK), 0, s,
                           dViews, dSeq, dExc, nBases, dWindows, nWindows, segs, k, dStats, work);
    else if (tables[0]->bucketSlots == 8)
        hipLaunchKernelGGL((scanKernel<8, false>), grid, dim3(SCAN_BLOCK), 0, s, dViews, dSeq, dExc,
                           nBases, dWindows, nWindows, segs, k, dStats, work);
    else
        hipLaunchKernelGGL((scanKernel<16, false>), grid, dim3(SCAN_BLOCK), 0, s, dViews, dSeq, dExc,
                           nBases, dWindows, nWindows, segs, k, dStats, work);
    HIPIBS_CHECK(hipGetLastError());
    if (profile)
    {
        HIPIBS_CHECK(hipEventRecord(tEv[1], s));
        tEvRecorded = true;
    }
    if (segs > 1)
    {
        const uint32_t items = nWindows * (uint32_t)nTables;
        hipLaunchKernelGGL(foldKernel, dim3((items + 255) / 256), dim3(256), 0, s, work, items,
                           segs, k, dStats);
        HIPIBS_CHECK(hipGetLastError());
    }
    return 0;
}

// One launch over windows that each name their table (windows[i].reserved
// indexes dViews); results row i per window; one segment per window.
int launchScanMixed(int bucketSlots, uint32_t k, const TableView *dViews, const uint64_t *dSeq,
                    const uint64_t *dExc, uint64_t nBases, const hipibs_window *dWindows,
                    uint32_t nWindows, hipibs_window_stats *dStats, hipStream_t s, bool grouped,
                    uint32_t segs, SegSummary *dWork)
{
    if (nWindows == 0)
        return 0;
    if (segs == 0 || (segs > 1 && !dWork))
        return setError(HIPIBS_E_ARG, "segments need a summary buffer");
    const dim3 grid(nWindows * segs, 1);
    SegSummary *work = segs > 1 ? dWork : nullptr;
    if (k > 32)
        hipLaunchKernelGGL((scanKernel<16, true, true>), grid, dim3(SCAN_BLOCK), 0, s, dViews,
                           dSeq, dExc, nBases, dWindows, nWindows, segs, k, dStats, work);
    else if (grouped && bucketSlots == 8)
        hipLaunchKernelGGL((scanKernel<8, true, false, true>), grid, dim3(SCAN_BLOCK), 0, s,
                           dViews, dSeq, dExc, nBases, dWindows, nWindows, segs, k, dStats, work);
    else if (grouped)
        hipLaunchKernelGGL((scanKernel<16, true, false, true>), grid, dim3(SCAN_BLOCK), 0, s,
                           dViews, dSeq, dExc, nBases, dWindows, nWindows, segs, k, dStats, work);
    else if (bucketSlots == 8)
        hipLaunchKernelGGL((scanKernel<8, true>), grid, dim3(SCAN_BLOCK), 0, s, dViews, dSeq, dExc,
                           nBases, dWindows, nWindows, segs, k, dStats, work);
    else
        hipLaunchKernelGGL((scanKernel<16, true>), grid, dim3(SCAN_BLOCK), 0, s, dViews, dSeq, dExc,
                           nBases, dWindows, nWindows, segs, k, dStats, work);
    HIPIBS_CHECK(hipGetLastError());
    if (segs > 1)
    {
        hipLaunchKernelGGL(foldKernel, dim3((nWindows + 255) / 256), dim3(256), 0, s, dWork,
                           nWindows, segs, k, dStats);
        HIPIBS_CHECK(hipGetLastError());
    }
    return 0;
}

// Table views live in a small device array; one per thread, grown on demand
// (the launch reads it asynchronously, so it is only replaced after a
// device synchronisation).
struct ViewCache
{
    TableView *d = nullptr;
    int cap = 0;
    int device = -1;
    std::vector<TableView> last; // contents of d as last uploaded
};
static thread_local ViewCache tViews;

// Every field counts: a freed table's address may come back for a table of
// another shape (views hold no padding bytes of undefined value: view()
// sets every member).
static bool sameViews(const std::vector<TableView> &a, const std::vector<TableView> &b)
{
    return a.size() == b.size() &&
           (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(TableView)) == 0);
}

int uploadViews(const hipibs_table *const *tables, int nTables, hipStream_t s, TableView **out)
{
    std::vector<TableView> v(nTables);
    for (int i = 0; i < nTables; ++i)
        v[i] = tables[i]->view();
    if (tViews.d && tViews.device == tables[0]->device && sameViews(v, tViews.last))
    {
        *out = tViews.d;
        return 0;
    }
    // The previous contents may still be read by queued kernels.
    HIPIBS_CHECK(hipDeviceSynchronize());
    if (tViews.cap < nTables || tViews.device != tables[0]->device)
    {
        if (tViews.d)
            (void)hipFree(tViews.d);
        tViews.d = nullptr;
        tViews.last.clear();
        const int cap = std::max(nTables, 64);
        HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&tViews.d), cap * sizeof(TableView)));
        tViews.cap = cap;
        tViews.device = tables[0]->device;
    }
    HIPIBS_CHECK(hipMemcpy(tViews.d, v.data(), nTables * sizeof(TableView), hipMemcpyHostToDevice));
    tViews.last = v;
    (void)s;
    *out = tViews.d;
    return 0;
}

int checkTables(const hipibs_table *const *tables, int nTables)
{
    if (!tables || nTables <= 0)
        return setError(HIPIBS_E_ARG, "no tables");
    for (int i = 0; i < nTables; ++i)
    {
        if (!tables[i])
            return setError(HIPIBS_E_ARG, "null table");
        if (tables[i]->k != tables[0]->k || tables[i]->device != tables[0]->device ||
            tables[i]->bucketSlots != tables[0]->bucketSlots || tables[i]->mlen != tables[0]->mlen)
            return setError(HIPIBS_E_ARG,
                            "tables differ in k, device, bucket width or minimizer length");
    }
    return 0;
}

// ---- coverage text: comma-separated decimal values, as the mapper writes ----

constexpr uint32_t kTextPerThread = 16;
constexpr uint32_t kTextBlock = 256 * kTextPerThread; // bases per block

__device__ inline uint32_t decimalDigits(uint32_t v)
{
    return v >= 100 ? 3u : v >= 10 ? 2u : 1u;
}

// Text bytes of each block's bases, one separator counted per value.
__global__ __launch_bounds__(256) void coverageTextLenKernel(const uint8_t *__restrict__ cov,
                                                             uint64_t n, uint64_t *blockLen)
{
    __shared__ uint32_t part[4];
    const uint64_t base = (uint64_t)blockIdx.x * kTextBlock + threadIdx.x * kTextPerThread;
    uint32_t len = 0;
    for (uint32_t i = 0; i < kTextPerThread; ++i)
        if (base + i < n)
            len += decimalDigits(cov[base + i]) + 1u;
    for (int off = 32; off > 0; off >>= 1)
        len += __shfl_down(len, off, 64);
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = len;
    __syncthreads();
    if (threadIdx.x == 0)
        blockLen[blockIdx.x] = (uint64_t)part[0] + part[1] + part[2] + part[3];
}

// Writes every block's values at blockOff[block]: decimal digits, ',' after
// every value but the contig's last.
__global__ __launch_bounds__(256) void coverageTextKernel(const uint8_t *__restrict__ cov,
                                                          uint64_t n,
                                                          const uint64_t *__restrict__ blockOff,
                                                          char *__restrict__ text)
{
    __shared__ uint32_t scan[256];
    const uint32_t tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * kTextBlock + tid * kTextPerThread;
    uint8_t v[kTextPerThread];
    uint32_t len = 0;
    for (uint32_t i = 0; i < kTextPerThread; ++i)
    {
        v[i] = base + i < n ? cov[base + i] : 0;
        if (base + i < n)
            len += decimalDigits(v[i]) + 1u;
    }
    scan[tid] = len;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1)
    {
        const uint32_t x = tid >= off ? scan[tid - off] : 0u;
        __syncthreads();
        scan[tid] += x;
        __syncthreads();
    }
    uint64_t o = blockOff[blockIdx.x] + scan[tid] - len;
    for (uint32_t i = 0; i < kTextPerThread && base + i < n; ++i)
    {
        const uint32_t x = v[i];
        if (x >= 100)
            text[o++] = (char)('0' + x / 100);
        if (x >= 10)
            text[o++] = (char)('0' + (x / 10) % 10);
        text[o++] = (char)('0' + x % 10);
        if (base + i + 1 < n)
            text[o++] = ',';
    }
}

// Lays out the text of d_cov[0..n): exclusive block offsets (device) and
// the total length. Synchronises s.
static int coverageTextLayout(const uint8_t *dCov, uint64_t n, hipStream_t s, uint64_t **dOff,
                              uint64_t *total)
{
    const uint64_t nBlocks = (n + kTextBlock - 1) / kTextBlock;
    std::vector<uint64_t> h(nBlocks);
    HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(dOff), nBlocks * 8));
    hipLaunchKernelGGL(coverageTextLenKernel, dim3((uint32_t)nBlocks), dim3(256), 0, s, dCov, n,
                       *dOff);
    int rc = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h.data(), *dOff, nBlocks * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        rc = setError(HIPIBS_E_HIP, "coverage text layout failed");
    uint64_t acc = 0;
    for (uint64_t b = 0; b < nBlocks; ++b)
    {
        const uint64_t l = h[b];
        h[b] = acc;
        acc += l;
    }
    if (!rc && (hipMemcpyAsync(*dOff, h.data(), nBlocks * 8, hipMemcpyHostToDevice, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        rc = setError(HIPIBS_E_HIP, "coverage text layout failed");
    if (rc)
    {
        hipFree(*dOff);
        *dOff = nullptr;
        return rc;
    }
    *total = acc - 1; // no separator after the last value
    return 0;
}

static int launchCoverageText(const uint8_t *dCov, uint64_t n, const uint64_t *dOff, char *dText,
                              hipStream_t s)
{
    const uint64_t nBlocks = (n + kTextBlock - 1) / kTextBlock;
    hipLaunchKernelGGL(coverageTextKernel, dim3((uint32_t)nBlocks), dim3(256), 0, s, dCov, n, dOff,
                       dText);
    HIPIBS_CHECK(hipGetLastError());
    return 0;
}

} // namespace hipibs

using namespace hipibs;

extern "C" int hipibs_coverage_text_device(const uint8_t *d_cov, uint64_t n_bases, char *d_text,
                                           uint64_t text_cap, uint64_t *text_len, void *stream)
{
    if (!text_len || (n_bases && !d_cov))
        return setError(HIPIBS_E_ARG, "null argument");
    *text_len = 0;
    if (n_bases == 0)
        return 0;
    if (n_bases > (uint64_t)UINT32_MAX * kTextBlock)
        return setError(HIPIBS_E_ARG, "contig too long for the coverage text kernels");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    uint64_t *dOff = nullptr, total = 0;
    int rc = coverageTextLayout(d_cov, n_bases, s, &dOff, &total);
    if (rc)
        return rc;
    *text_len = total;
    if (d_text)
    {
        if (total > text_cap)
            rc = setError(HIPIBS_E_ARG, "coverage text buffer too small (" + std::to_string(total) +
                                            " bytes needed)");
        else
            rc = launchCoverageText(d_cov, n_bases, dOff, d_text, s);
        if (!rc && hipStreamSynchronize(s) != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "coverage text failed");
    }
    hipFree(dOff);
    return rc;
}

extern "C" int hipibs_coverage_text_stream(const uint8_t *d_cov, uint64_t n_bases,
                                           hipibs_sink_fn sink, void *user, void *stream)
{
    if (!sink || (n_bases && !d_cov))
        return setError(HIPIBS_E_ARG, "null argument");
    if (n_bases == 0)
        return 0;
    if (n_bases > (uint64_t)UINT32_MAX * kTextBlock)
        return setError(HIPIBS_E_ARG, "contig too long for the coverage text kernels");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    uint64_t *dOff = nullptr, total = 0;
    int rc = coverageTextLayout(d_cov, n_bases, s, &dOff, &total);
    if (rc)
        return rc;
    char *dText = nullptr, *pinned[2] = {nullptr, nullptr};
    hipEvent_t ready[2] = {nullptr, nullptr};
    const uint64_t chunk = std::min<uint64_t>(total, 64ull << 20);
    if (hipMalloc(reinterpret_cast<void **>(&dText), total) != hipSuccess)
        rc = setError(HIPIBS_E_NOMEM, "coverage text allocation failed");
    for (int i = 0; i < 2 && !rc; ++i)
        if (hipHostMalloc(reinterpret_cast<void **>(&pinned[i]), chunk) != hipSuccess ||
            hipEventCreateWithFlags(&ready[i], hipEventDisableTiming) != hipSuccess)
            rc = setError(HIPIBS_E_NOMEM, "staging allocation failed");
    if (!rc)
        rc = launchCoverageText(d_cov, n_bases, dOff, dText, s);
    // chunk i is copied while the sink consumes chunk i-1
    const uint64_t nChunks = rc ? 0 : (total + chunk - 1) / chunk;
    for (uint64_t c = 0; c <= nChunks && !rc; ++c)
    {
        if (c < nChunks)
        {
            const uint64_t o = c * chunk, m = std::min(chunk, total - o);
            if (hipMemcpyAsync(pinned[c & 1], dText + o, m, hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipEventRecord(ready[c & 1], s) != hipSuccess)
            {
                rc = setError(HIPIBS_E_HIP, "coverage text download failed");
                break;
            }
        }
        if (c > 0)
        {
            const uint64_t p = c - 1, o = p * chunk, m = std::min(chunk, total - o);
            if (hipEventSynchronize(ready[p & 1]) != hipSuccess)
            {
                rc = setError(HIPIBS_E_HIP, "coverage text download failed");
                break;
            }
            if (sink(user, pinned[p & 1], m) != 0)
                rc = setError(HIPIBS_E_IO, "coverage text sink failed");
        }
    }
    hipStreamSynchronize(s);
    for (int i = 0; i < 2; ++i)
    {
        if (pinned[i])
            hipHostFree(pinned[i]);
        if (ready[i])
            hipEventDestroy(ready[i]);
    }
    if (dText)
        hipFree(dText);
    hipFree(dOff);
    return rc;
}

extern "C" uint64_t hipibs_scan_work_bytes(uint32_t n_windows, int n_tables,
                                           uint32_t max_window_bases, uint32_t k,
                                           uint32_t segs_per_window, uint32_t *segs_resolved)
{
    const uint32_t maxPos = max_window_bases >= k ? max_window_bases - k + 1 : 0;
    const uint32_t segs =
        segs_per_window ? segs_per_window : autoSegsPerWindow(n_windows, n_tables, maxPos);
    if (segs_resolved)
        *segs_resolved = segs;
    return scanWorkBytes(n_windows, n_tables, segs);
}

extern "C" int hipibs_scan_windows_device(const hipibs_table *const *tables, int n_tables,
                                          const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                                          uint64_t n_bases, const hipibs_window *d_windows,
                                          uint32_t n_windows, uint32_t max_window_bases,
                                          hipibs_window_stats *d_stats, uint32_t segs_per_window,
                                          void *d_work, uint64_t work_bytes, void *stream)
{
    int rc = checkTables(tables, n_tables);
    if (rc)
        return rc;
    if (n_windows && (!d_seq_words || !d_exc_words || !d_windows || !d_stats))
        return setError(HIPIBS_E_ARG, "null buffer");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    HIPIBS_CHECK(hipSetDevice(tables[0]->device));
    TableView *views = nullptr;
    rc = uploadViews(tables, n_tables, s, &views);
    if (rc)
        return rc;
    return launchScan(tables, n_tables, d_seq_words, d_exc_words, n_bases, d_windows, n_windows,
                      max_window_bases, d_stats, segs_per_window, d_work, work_bytes, s, views);
}

// ---- the scan without allocation -----------------------------------------
// Scratch of hipibs_scan_windows_device_noalloc: the launch's table views
// (at most kNoAllocTables, in the first kNoAllocViewBytes), then the
// segment summaries of a window split over several blocks. One copy per
// GPU, held until the process ends; every call on a GPU shares it (calls
// that overlap on one GPU race). Nothing in it needs clearing: every block
// writes its summary before the fold reads it.
constexpr int kNoAllocTables = 64;
constexpr uint64_t kNoAllocViewBytes = 4096;
constexpr uint64_t kNoAllocScratchBytes = 16ull << 20;
static_assert(kNoAllocTables * sizeof(TableView) <= kNoAllocViewBytes, "views exceed their part");
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Device address of `scratch` per device (the host-side name is not one):
// taken once per device, with that device current.
constexpr int kMaxDevices = 64;
static std::mutex gScratchMu;
static void *gScratch[kMaxDevices] = {};

static int scratchOfDevice(int device, uint8_t **out)
{
    if (device < 0 || device >= kMaxDevices)
        return setError(HIPIBS_E_ARG, "device ordinal out of range for the scan scratch");
    std::lock_guard<std::mutex> lk(gScratchMu);
    if (!gScratch[device])
    {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return hipFail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                           "hipGetSymbolAddress of the scan scratch");
        gScratch[device] = p;
    }
    *out = static_cast<uint8_t *>(gScratch[device]);
    return 0;
}

// Table views travel as kernel arguments (no host buffer outlives the call).
constexpr int kViewsPerLaunch = 16;
struct ViewBatch
{
    TableView v[kViewsPerLaunch];
};

__global__ void setViewsKernel(const ViewBatch b, uint32_t n, TableView *dst)
{
    if (threadIdx.x < n)
        dst[threadIdx.x] = b.v[threadIdx.x];
}

extern "C" uint64_t hipibs_scan_noalloc_max_work_bytes(void)
{
    return kNoAllocScratchBytes - kNoAllocViewBytes;
}

extern "C" int hipibs_scan_windows_device_noalloc(const hipibs_table *const *tables, int n_tables,
                                                  const uint64_t *d_seq_words,
                                                  const uint64_t *d_exc_words, uint64_t n_bases,
                                                  const hipibs_window *d_windows,
                                                  uint32_t n_windows, uint32_t max_window_bases,
                                                  hipibs_window_stats *d_stats,
                                                  uint32_t segs_per_window, void *stream)
{
    int rc = checkTables(tables, n_tables);
    if (rc)
        return rc;
    if (n_tables > kNoAllocTables)
        return setError(HIPIBS_E_ARG, "the scan without allocation takes at most " +
                                          std::to_string(kNoAllocTables) + " tables");
    if (n_windows && (!d_seq_words || !d_exc_words || !d_windows || !d_stats))
        return setError(HIPIBS_E_ARG, "null buffer");
    const uint32_t k = tables[0]->k;
    const uint32_t maxPos = max_window_bases >= k ? max_window_bases - k + 1 : 0;
    const uint32_t segs =
        segs_per_window ? segs_per_window : autoSegsPerWindow(n_windows, n_tables, maxPos);
    const uint64_t work = scanWorkBytes(n_windows, n_tables, segs);
    if (work > kNoAllocScratchBytes - kNoAllocViewBytes)
        return setError(HIPIBS_E_ARG,
                        "windows x tables x segments need " + std::to_string(work) +
                            " bytes of scratch, more than the " +
                            std::to_string(kNoAllocScratchBytes - kNoAllocViewBytes) +
                            " the scan without allocation holds: fewer windows per call or "
                            "segs_per_window = 1");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    HIPIBS_CHECK(hipSetDevice(tables[0]->device));
    uint8_t *base = nullptr;
    rc = scratchOfDevice(tables[0]->device, &base);
    if (rc)
        return rc;
    auto *views = reinterpret_cast<TableView *>(base);
    for (int c = 0; c < n_tables; c += kViewsPerLaunch)
    {
        ViewBatch b{};
        const int m = std::min(kViewsPerLaunch, n_tables - c);
        for (int i = 0; i < m; ++i)
            b.v[i] = tables[c + i]->view();
        hipLaunchKernelGGL(setViewsKernel, dim3(1), dim3(64), 0, s, b, (uint32_t)m, views + c);
        HIPIBS_CHECK(hipGetLastError());
    }
    return launchScan(tables, n_tables, d_seq_words, d_exc_words, n_bases, d_windows, n_windows,
                      max_window_bases, d_stats, segs, base + kNoAllocViewBytes,
                      kNoAllocScratchBytes - kNoAllocViewBytes, s, views, false);
}

extern "C" int hipibs_hitmask_device(const hipibs_table *t, const uint64_t *d_seq_words,
                                     const uint64_t *d_exc_words, uint64_t pos0, uint64_t n_pos,
                                     uint64_t contig_end, uint32_t *d_hit_words, void *stream)
{
    if (!t || (n_pos && (!d_seq_words || !d_exc_words || !d_hit_words)))
        return setError(HIPIBS_E_ARG, "null argument");
    if (n_pos == 0)
        return 0;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint64_t chunks = (n_pos + 63) / 64;
    const uint64_t grid = std::min<uint64_t>((chunks + 3) / 4, (uint64_t)deviceCUs(t->device) * 16);
    if (t->k > 32)
        hipLaunchKernelGGL((hitmaskKernel<16, true>), dim3((uint32_t)grid), dim3(256), 0, s,
                           t->view(), d_seq_words, d_exc_words, pos0, n_pos, contig_end, t->k,
                           d_hit_words);
    else if (t->mlen && t->bucketSlots == 8)
        hipLaunchKernelGGL((hitmaskKernel<8, false, true>), dim3((uint32_t)grid), dim3(256), 0, s,
                           t->view(), d_seq_words, d_exc_words, pos0, n_pos, contig_end, t->k,
                           d_hit_words);
    else if (t->mlen)
        hipLaunchKernelGGL((hitmaskKernel<16, false, true>), dim3((uint32_t)grid), dim3(256), 0,
                           s, t->view(), d_seq_words, d_exc_words, pos0, n_pos, contig_end, t->k,
                           d_hit_words);
    else if (t->bucketSlots == 8)
        hipLaunchKernelGGL(hitmaskKernel<8>, dim3((uint32_t)grid), dim3(256), 0, s, t->view(),
                           d_seq_words, d_exc_words, pos0, n_pos, contig_end, t->k, d_hit_words);
    else
        hipLaunchKernelGGL(hitmaskKernel<16>, dim3((uint32_t)grid), dim3(256), 0, s, t->view(),
                           d_seq_words, d_exc_words, pos0, n_pos, contig_end, t->k, d_hit_words);
    HIPIBS_CHECK(hipGetLastError());
    return 0;
}

extern "C" int hipibs_coverage_device(const uint32_t *d_hit_words, uint64_t n_pos, uint64_t n_bases,
                                      uint32_t k, uint8_t *d_cov, void *stream)
{
    if (k == 0 || k > 64 || (n_bases && !d_cov) || (n_pos && !d_hit_words))
        return setError(HIPIBS_E_ARG, "bad argument");
    if (n_pos > n_bases)
        return setError(HIPIBS_E_ARG, "more positions than bases");
    if (n_bases == 0)
        return 0;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int dev = 0;
    (void)hipGetDevice(&dev);
    const uint64_t waves = (n_bases + kCovPerWave - 1) / kCovPerWave;
    const uint64_t grid = std::min<uint64_t>((waves + 3) / 4, (uint64_t)deviceCUs(dev) * 16);
    hipLaunchKernelGGL(coverageKernel, dim3((uint32_t)grid), dim3(256), 0, s, d_hit_words, n_pos,
                       n_bases, k, d_cov);
    HIPIBS_CHECK(hipGetLastError());
    return 0;
}

extern "C" int hipibs_set_profiling(int on)
{
    tProfile = on != 0;
    tEvRecorded = false;
    return 0;
}

extern "C" double hipibs_last_kernel_ms(void)
{
    if (!tProfile || !tEvRecorded)
        return -1.0;
    float ms = -1.0f;
    if (hipEventSynchronize(tEv[1]) != hipSuccess)
        return -1.0;
    if (hipEventElapsedTime(&ms, tEv[0], tEv[1]) != hipSuccess)
        return -1.0;
    return (double)ms;
}

extern "C" int hipibs_encode_device(const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                                    uint64_t pos0, uint64_t n_pos, uint64_t contig_end, uint32_t k,
                                    uint64_t *d_codes, uint8_t *d_valid, void *stream)
{
    if (k == 0 || k > 64 || (n_pos && (!d_seq_words || !d_exc_words || !d_codes)))
        return setError(HIPIBS_E_ARG, "bad argument");
    if (n_pos == 0)
        return 0;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int dev = 0;
    (void)hipGetDevice(&dev);
    const uint64_t grid = std::min<uint64_t>((n_pos + 255) / 256, (uint64_t)deviceCUs(dev) * 16);
    if (k > 32)
        hipLaunchKernelGGL(encodeKernel<true>, dim3((uint32_t)grid), dim3(256), 0, s, d_seq_words,
                           d_exc_words, pos0, n_pos, contig_end, k, d_codes, d_valid);
    else
        hipLaunchKernelGGL(encodeKernel<false>, dim3((uint32_t)grid), dim3(256), 0, s,
                           d_seq_words, d_exc_words, pos0, n_pos, contig_end, k, d_codes, d_valid);
    HIPIBS_CHECK(hipGetLastError());
    return 0;
}
