// Wavefront-cooperative membership probe of 64 queries (one per lane) in a
// bucketised table. Device-only.
//
// A bucket of S 8-byte slots (S = 8: 64 B, S = 16: 128 B) is read by
// LPQ = S/2 adjacent lanes, 16 B each, so one load instruction of the wave
// fetches QPR = 64/LPQ whole buckets as contiguous, aligned requests (one
// 64 B or 128 B HBM access per query). The 64 queries take LPQ such rounds;
// all of their loads are issued before any compare so they are in flight
// together. The match / empty-slot flags come back as wave ballots, from
// which each lane reads its own query's result. Queries whose bucket was
// full without a match (rare) continue to the following buckets in compact
// rounds until found, an empty slot is seen, or the table's maximum
// displacement is exhausted.
#pragma once

#include "hipibs_internal.hpp"

namespace hipibs
{

struct alignas(16) SlotPair
{
    uint64_t a, b;
};

// 16-byte load of bucket slots through the global address space (a generic
// pointer would compile to a flat load, which also counts against the LDS
// wait counter), non-temporal: random bucket reads have no reuse, and the
// nt hint raised the random 64 B-bucket read rate over a 27 GB buffer from
// 48.5 to 54 G/s (every read is still one 128 B memory request).
typedef unsigned long long U64x2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) U64x2 *GlobalU64x2;
__device__ inline SlotPair loadPair(const uint64_t *p)
{
    const U64x2 v = __builtin_nontemporal_load((GlobalU64x2)(reinterpret_cast<const U64x2 *>(p)));
    return SlotPair{v.x, v.y};
}

__device__ inline int laneId()
{
    return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of m below this lane.
__device__ inline uint32_t maskedRank(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline uint64_t ballot(bool p)
{
    return __ballot(p);
}

__device__ inline uint32_t shuffle32(uint32_t v, int src)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}

__device__ inline uint64_t shuffle64(uint64_t v, int src)
{
    const uint32_t lo = shuffle32((uint32_t)v, src);
    const uint32_t hi = shuffle32((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// Position of the n-th (0-based) set bit of m; requires n < popcount(m).
__device__ inline int selectBit(uint64_t m, uint32_t n)
{
    int pos = 0;
    uint32_t c = __popc((uint32_t)m);
    if (n >= c)
    {
        n -= c;
        pos += 32;
        m >>= 32;
    }
    c = __popc((uint32_t)m & 0xFFFFu);
    if (n >= c)
    {
        n -= c;
        pos += 16;
        m >>= 16;
    }
    c = __popc((uint32_t)m & 0xFFu);
    if (n >= c)
    {
        n -= c;
        pos += 8;
        m >>= 8;
    }
    c = __popc((uint32_t)m & 0xFu);
    if (n >= c)
    {
        n -= c;
        pos += 4;
        m >>= 4;
    }
    c = __popc((uint32_t)m & 0x3u);
    if (n >= c)
    {
        n -= c;
        pos += 2;
        m >>= 2;
    }
    c = (uint32_t)m & 1u;
    if (n >= c)
        pos += 1;
    return pos;
}

// Continues the queries of mask NC (home bucket full without a match) in
// the buckets after home, in compact rounds of QPR queries: each stops at
// its code (H), at `alt` (ALT only: A), at an EMPTY slot or after the
// table's maximum displacement. Must be called by all 64 lanes.
template <int S, bool ALT>
__device__ inline void probeOverflow(const TableView &T, uint64_t code, uint32_t home, uint64_t alt,
                                     uint64_t NC, uint64_t &H, uint64_t &A)
{
    constexpr int LPQ = S / 2;
    constexpr int QPR = 64 / LPQ;
    constexpr uint32_t GMASK = (1u << LPQ) - 1u;
    const int lane = laneId();
    const int g = lane / LPQ, sub = lane % LPQ;
    // Overflow rounds: pending lane i continues at home_i + disp_i.
    uint32_t disp = 1;
    while (NC)
    {
        const bool pending = (NC >> lane) & 1ull;
        const uint32_t nPend = (uint32_t)__popcll(NC);
        // group g serves the pending lane of rank g
        const bool serving = (uint32_t)g < nPend;
        const int src = serving ? selectBit(NC, (uint32_t)g) : 0;
        const uint64_t qcode = shuffle64(code, src);
        const uint64_t qhome = shuffle32(home, src);
        const uint32_t qdisp = shuffle32(disp, src);
        uint64_t qalt = 0;
        if constexpr (ALT)
            qalt = shuffle64(alt, src);
        SlotPair kv2{EMPTY, EMPTY};
        if (serving)
        {
            uint64_t b = qhome + qdisp;
            if (b >= T.nBuckets)
                b -= T.nBuckets;
            kv2 = loadPair(T.slots + b * S + (uint32_t)sub * 2u);
        }
        const uint64_t M2 = ballot(serving && (kv2.a == qcode || kv2.b == qcode));
        const uint64_t E2 = ballot(serving && (kv2.a == EMPTY || kv2.b == EMPTY));
        uint64_t A2 = 0;
        if constexpr (ALT)
            A2 = ballot(serving && (kv2.a == qalt || kv2.b == qalt));
        const uint32_t rank = maskedRank(NC);
        const bool served = pending && rank < (uint32_t)QPR;
        const uint32_t sh = (served ? rank : 0u) * (uint32_t)LPQ;
        const bool f2 = served && ((M2 >> sh) & GMASK) != 0;
        const bool e2 = served && ((E2 >> sh) & GMASK) != 0;
        const bool a2 = ALT && served && !f2 && ((A2 >> sh) & GMASK) != 0;
        const bool done = served && (f2 || a2 || e2 || disp >= T.maxDisp);
        H |= ballot(f2);
        if constexpr (ALT)
            A |= ballot(a2);
        if (served && !done)
            disp += 1;
        NC &= ~ballot(done);
    }
}

// Returns the 64-bit hit mask of the wave: bit i set when lane i's query is
// valid and its code is stored in the table; `home` is the query's home
// bucket (homeOf). Must be called by all 64 lanes. With ALT, a query also
// stops at the slot value `alt` (its minimizer group's marker) and *altHit
// gets the mask of the queries that stopped there.
template <int S, bool ALT = false>
__device__ inline uint64_t waveProbe(const TableView T, uint64_t code, uint32_t home, bool valid,
                                     uint64_t alt = 0, uint64_t *altHit = nullptr)
{
    constexpr int LPQ = S / 2;      // lanes per query
    constexpr int QPR = 64 / LPQ;   // queries per round
    constexpr int ROUNDS = 64 / QPR;
    constexpr uint32_t GMASK = (1u << LPQ) - 1u;
    const int lane = laneId();
    const int g = lane / LPQ, sub = lane % LPQ;
    const uint64_t V = ballot(valid);

    // All shuffles first, then all bucket loads: the loads are global
    // (vmcnt-only) so the LDS permutes never wait on them, and all ROUNDS
    // loads of the wave are in flight together.
    uint64_t qc[ROUNDS], qa[ROUNDS];
    uint32_t qh[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r)
    {
        qc[r] = shuffle64(code, r * QPR + g);
        qh[r] = shuffle32(home, r * QPR + g);
        if constexpr (ALT)
            qa[r] = shuffle64(alt, r * QPR + g);
    }
    SlotPair kv[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r)
    {
        kv[r] = SlotPair{EMPTY, EMPTY};
        if ((V >> (r * QPR + g)) & 1ull)
        {
            kv[r] = loadPair(T.slots + (uint64_t)qh[r] * S + (uint32_t)sub * 2u);
        }
    }
    uint64_t M[ROUNDS], E[ROUNDS], AM[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r)
    {
        const bool vq = (V >> (r * QPR + g)) & 1ull;
        M[r] = ballot(vq && (kv[r].a == qc[r] || kv[r].b == qc[r]));
        E[r] = ballot(vq && (kv[r].a == EMPTY || kv[r].b == EMPTY));
        if constexpr (ALT)
            AM[r] = ballot(vq && (kv[r].a == qa[r] || kv[r].b == qa[r]));
    }
    const int rOwn = lane / QPR, gOwn = lane % QPR;
    uint64_t Mo = M[0], Eo = E[0], Ao = 0;
    if constexpr (ALT)
        Ao = AM[0];
#pragma unroll
    for (int r = 1; r < ROUNDS; ++r)
        if (rOwn == r)
        {
            Mo = M[r];
            Eo = E[r];
            if constexpr (ALT)
                Ao = AM[r];
        }
    const bool found = ((Mo >> (gOwn * LPQ)) & GMASK) != 0;
    const bool hasEmpty = ((Eo >> (gOwn * LPQ)) & GMASK) != 0;
    const bool marked = ALT && ((Ao >> (gOwn * LPQ)) & GMASK) != 0;
    uint64_t H = ballot(valid && found);
    uint64_t A = 0;
    if constexpr (ALT)
        A = ballot(valid && !found && marked);
    uint64_t NC = ballot(valid && !found && !marked && !hasEmpty);
    if (NC && T.maxDisp)
        probeOverflow<S, ALT>(T, code, home, alt, NC, H, A);
    if constexpr (ALT)
        *altHit = A;
    return H;
}

// waveProbe for tables of 33..64-base keys: a bucket is 8 keys of 16 B
// (high word, low word), so 8 lanes read one query's bucket, one key each;
// a slot whose high word is EMPTY is free. The one canonical code with an
// all-ones high word (k = 64: T^32 A^32) never enters the buckets: its
// presence is T.wideTop, and a lane querying it answers from that flag.
// Must be called by all 64 lanes.
__device__ inline uint64_t waveProbeWide(const TableView T, uint64_t hi, uint64_t lo,
                                         uint32_t home, bool valid)
{
    constexpr int LPQ = 8, QPR = 8, ROUNDS = 8, S = 16;
    constexpr uint32_t GMASK = 0xFFu;
    const int lane = laneId();
    const int g = lane / LPQ, sub = lane % LPQ;
    const bool top = valid && isWideTop(hi, lo);
    const uint64_t topHit = T.wideTop ? ballot(top) : 0ull;
    valid = valid && !top;
    const uint64_t V = ballot(valid);
    uint64_t qh[ROUNDS], ql[ROUNDS];
    uint32_t qb[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r)
    {
        qh[r] = shuffle64(hi, r * QPR + g);
        ql[r] = shuffle64(lo, r * QPR + g);
        qb[r] = shuffle32(home, r * QPR + g);
    }
    SlotPair kv[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r)
    {
        kv[r] = SlotPair{EMPTY, EMPTY};
        if ((V >> (r * QPR + g)) & 1ull)
            kv[r] = loadPair(T.slots + (uint64_t)qb[r] * S + (uint32_t)sub * 2u);
    }
    uint64_t M[ROUNDS], E[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r)
    {
        const bool vq = (V >> (r * QPR + g)) & 1ull;
        M[r] = ballot(vq && kv[r].a == qh[r] && kv[r].b == ql[r]);
        E[r] = ballot(vq && kv[r].a == EMPTY);
    }
    const int rOwn = lane / QPR, gOwn = lane % QPR;
    uint64_t Mo = M[0], Eo = E[0];
#pragma unroll
    for (int r = 1; r < ROUNDS; ++r)
        if (rOwn == r)
        {
            Mo = M[r];
            Eo = E[r];
        }
    const bool found = ((Mo >> (gOwn * LPQ)) & GMASK) != 0;
    const bool hasEmpty = ((Eo >> (gOwn * LPQ)) & GMASK) != 0;
    uint64_t H = ballot(valid && found) | topHit;
    uint64_t NC = ballot(valid && !found && !hasEmpty);
    if (NC == 0 || T.maxDisp == 0)
        return H;
    // Overflow rounds: pending lane i continues at home_i + disp_i.
    uint32_t disp = 1;
    while (NC)
    {
        const bool pending = (NC >> lane) & 1ull;
        const uint32_t nPend = (uint32_t)__popcll(NC);
        const bool serving = (uint32_t)g < nPend;
        const int src = serving ? selectBit(NC, (uint32_t)g) : 0;
        const uint64_t qhi = shuffle64(hi, src), qlo = shuffle64(lo, src);
        const uint64_t qhome = shuffle32(home, src);
        const uint32_t qdisp = shuffle32(disp, src);
        SlotPair kv2{EMPTY, EMPTY};
        if (serving)
        {
            uint64_t b = qhome + qdisp;
            if (b >= T.nBuckets)
                b -= T.nBuckets;
            kv2 = loadPair(T.slots + b * S + (uint32_t)sub * 2u);
        }
        const uint64_t M2 = ballot(serving && kv2.a == qhi && kv2.b == qlo);
        const uint64_t E2 = ballot(serving && kv2.a == EMPTY);
        const uint32_t rank = maskedRank(NC);
        const bool served = pending && rank < (uint32_t)QPR;
        const uint32_t sh = (served ? rank : 0u) * (uint32_t)LPQ;
        const bool f2 = served && ((M2 >> sh) & GMASK) != 0;
        const bool e2 = served && ((E2 >> sh) & GMASK) != 0;
        const bool done = served && (f2 || e2 || disp >= T.maxDisp);
        H |= ballot(f2);
        if (served && !done)
            disp += 1;
        NC &= ~ballot(done);
    }
    return H;
}

// Canonical code of the k = 33..64 bases whose low-first 2-bit packing is
// x0 (bases 0..31) and x1 (bases 32..63, only the first k-32 used): the
// reverse complement is the complement of the packing; the forward code
// puts the first base in the most significant bits, through the 2-bit
// group reversal of each word.
__device__ inline void wideCanonical(uint64_t x0, uint64_t x1, uint32_t k, uint64_t &hi,
                                     uint64_t &lo)
{
    const uint32_t k2 = k - 32; // bases in the high word
    const uint64_t m2 = k2 == 32 ? ~0ull : (1ull << (2 * k2)) - 1ull;
    x1 &= m2;
    const uint64_t rhi = ~x1 & m2, rlo = ~x0;
    const uint64_t r0 = reverse2(x0);
    const uint64_t r1 = reverse2(x1) >> (64 - 2 * k2);
    const uint64_t fhi = k2 == 32 ? r0 : r0 >> (64 - 2 * k2);
    const uint64_t flo = k2 == 32 ? r1 : (r0 << (2 * k2)) | r1;
    const bool f = fhi < rhi || (fhi == rhi && flo < rlo);
    hi = f ? fhi : rhi;
    lo = f ? flo : rlo;
}

// Membership in a minimizer table of the codes whose minimizer hash is
// minHash: the minimizer-ordered part first; queries that meet their
// group's marker there continue in the whole-k-mer part.
template <int S>
__device__ inline uint64_t waveProbeMinimizer(const TableView &T, uint64_t code, uint32_t minHash,
                                              bool valid)
{
    const uint32_t home = valid ? (uint32_t)bucketOfMinimizer(minHash, T.nBuckets) : 0u;
    uint64_t A = 0;
    uint64_t H = waveProbe<S, true>(T, code, home, valid, markerOf(minHash), &A);
    if (A)
    {
        const TableView T2{T.slots2, T.nBuckets2, T.maxDisp2, 0u, T.k, 0u, nullptr, 0ull};
        const bool v2 = (A >> laneId()) & 1ull;
        H |= waveProbe<8>(T2, code, v2 ? (uint32_t)bucketOf(code, T.nBuckets2) : 0u, v2);
    }
    return H;
}

// Sliding minimum of m-mer hashes across the wave: L holds the hash at
// position base + lane, Hh the one at base + 64 + lane; lane i gets the
// minimum over positions base + i .. base + i + win - 1 (win <= 65), the
// minimizer hash of the k-mer at base + i when win = k - m + 1. Doubling
// spans, then one overlapping step for the remainder. Must be called by
// all 64 lanes.
__device__ inline uint32_t waveWindowMin(uint32_t L, uint32_t Hh, int win)
{
    const int lane = laneId();
    int span = 1;
    while (2 * span <= win)
    {
        const int src = (lane + span) & 63;
        const uint32_t l2 = shuffle32(L, src), h2 = shuffle32(Hh, src);
        const bool inLo = lane + span < 64;
        L = min(L, inLo ? l2 : h2);
        Hh = min(Hh, inLo ? h2 : 0xffffffffu);
        span *= 2;
    }
    const int d = win - span;
    const int src = (lane + d) & 63;
    const uint32_t l2 = shuffle32(L, src), h2 = shuffle32(Hh, src);
    return min(L, lane + d < 64 ? l2 : h2);
}

// Orders this wave's LDS accesses around a hand-over between lanes: the
// writes before it are visible to every lane of the wave after it, and no
// access moves across it. Wavefront-scope fences emit no instructions (a
// wave's LDS operations complete in order); they keep the compiler from
// reordering the accesses of different lanes, which it cannot see.
__device__ inline void waveLdsSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wave LDS staging of waveProbeGrouped over NW words: up to kLines home
// lines, each S slots plus 16 B of padding (lanes that read different lines
// at the same offset then fall in different banks), and the lines' home
// buckets.
template <int S, int NW>
struct GroupStage
{
    static constexpr int kLines = 32 * NW;
    static constexpr int kStride = S + 2; // 64-bit words per staged line
    static constexpr int kWords = kLines * kStride + kLines / 2;
};

// waveProbeMinimizer for the scan's consecutive positions, NW words of 64
// positions at once (lane i holds position i of every word): lanes in runs
// of equal minimizer hash share one home bucket, so the wave loads each
// run's home line once (LPQ lanes x 16 B; the lines of all NW words are in
// flight together, up to kLines per pass), stages the lines in LDS and
// every lane searches its runs' lines there for its code, its group's
// marker (the group lives in the whole-k-mer part) and an EMPTY last slot
// (slots fill from the first, so then the key is not displaced). Queries
// whose home line is full without their key or marker continue in
// probeOverflow; marked ones probe the whole-k-mer part. `stage` is this
// wave's GroupStage<S, NW>::kWords words of LDS, 16 B aligned. H[w] gets
// word w's hit mask. Must be called by all 64 lanes.
template <int S, int NW>
__device__ inline void waveProbeGrouped(const TableView &T, const uint64_t (&code)[NW],
                                        const uint32_t (&minHash)[NW], const bool (&valid)[NW],
                                        uint64_t *stage, uint64_t (&H)[NW])
{
    using G = GroupStage<S, NW>;
    constexpr int LPQ = S / 2;           // lanes per line
    constexpr int LPR = 64 / LPQ;        // lines per load round
    constexpr int ROUNDS = G::kLines / LPR;
    uint64_t *lines = stage;
    uint32_t *homes = reinterpret_cast<uint32_t *>(stage + G::kLines * G::kStride);
    const int lane = laneId();
    const int g = lane / LPQ, sub = lane % LPQ;
    uint32_t run[NW], home[NW], below[NW];
    bool leader[NW];
    uint32_t nL = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w)
    {
        const uint64_t V = ballot(valid[w]);
        const uint32_t prevHash = shuffle32(minHash[w], lane > 0 ? lane - 1 : 0);
        const bool prevValid = lane > 0 && ((V >> ((lane - 1) & 63)) & 1ull);
        leader[w] = valid[w] && !(prevValid && prevHash == minHash[w]);
        const uint64_t LM = ballot(leader[w]);
        // runs are numbered across the words: word w's after word w-1's
        below[w] = nL + maskedRank(LM);
        // a valid lane that is not a leader has a valid lane of its run below
        run[w] = leader[w] ? below[w] : below[w] - 1u;
        home[w] = valid[w] ? (uint32_t)bucketOfMinimizer(minHash[w], T.nBuckets) : 0u;
        nL += (uint32_t)__popcll(LM);
        H[w] = 0;
    }
    uint64_t A[NW], NC[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w)
        A[w] = NC[w] = 0;
    for (uint32_t base = 0; base < nL; base += G::kLines)
    {
#pragma unroll
        for (int w = 0; w < NW; ++w)
            if (leader[w] && below[w] - base < (uint32_t)G::kLines)
                homes[below[w] - base] = home[w];
        waveLdsSync(); // homes[] are read by the lanes that load the lines
        const uint32_t nLines = min(nL - base, (uint32_t)G::kLines);
        U64x2 kv[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r)
        {
            const uint32_t li = (uint32_t)(r * LPR + g);
            if (li < nLines)
                kv[r] = __builtin_nontemporal_load((GlobalU64x2)(reinterpret_cast<const U64x2 *>(
                    T.slots + (uint64_t)homes[li] * S + (uint32_t)sub * 2u)));
        }
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r)
        {
            const uint32_t li = (uint32_t)(r * LPR + g);
            if (li < nLines)
                *reinterpret_cast<U64x2 *>(lines + li * G::kStride + sub * 2) = kv[r];
        }
        waveLdsSync(); // lines[] are searched by the lanes of their runs
#pragma unroll
        for (int w = 0; w < NW; ++w)
        {
            const uint32_t li = run[w] - base;
            const bool mine = valid[w] && li < (uint32_t)G::kLines;
            const uint64_t marker = markerOf(minHash[w]);
            bool found = false, marked = false, hasEmpty = false;
            if (mine)
            {
                const uint64_t *ln = lines + li * G::kStride;
#pragma unroll
                for (int q = 0; q < S; q += 2)
                {
                    const U64x2 v = *reinterpret_cast<const U64x2 *>(ln + q);
                    found = found || v.x == code[w] || v.y == code[w];
                    marked = marked || v.x == marker || v.y == marker;
                    if (q == S - 2)
                        hasEmpty = v.y == EMPTY;
                }
            }
            H[w] |= ballot(mine && found);
            A[w] |= ballot(mine && !found && marked);
            NC[w] |= ballot(mine && !found && !marked && !hasEmpty);
        }
        waveLdsSync(); // the next pass overwrites homes[] and lines[]
    }
#pragma unroll
    for (int w = 0; w < NW; ++w)
    {
        if (NC[w] && T.maxDisp)
            probeOverflow<S, true>(T, code[w], home[w], markerOf(minHash[w]), NC[w], H[w], A[w]);
        if (A[w])
        {
            const TableView T2{T.slots2, T.nBuckets2, T.maxDisp2, 0u, T.k, 0u, nullptr, 0ull};
            const bool v2 = (A[w] >> lane) & 1ull;
            H[w] |= waveProbe<8>(T2, code[w], v2 ? (uint32_t)bucketOf(code[w], T.nBuckets2) : 0u,
                                 v2);
        }
    }
}

// Membership of each lane's code in T, whole-k-mer or minimizer table.
template <int S>
__device__ inline uint64_t waveLookup(const TableView &T, uint64_t code, bool valid)
{
    if (T.mlen == 0)
        return waveProbe<S>(T, code, valid ? (uint32_t)bucketOf(code, T.nBuckets) : 0u, valid);
    return waveProbeMinimizer<S>(T, code, valid ? minimizerOf(code, T.k, T.mlen) : 0u, valid);
}

__device__ inline uint32_t waveSumU32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v += (uint32_t)__shfl_xor((int)v, o, 64);
    return v;
}

// Positions [g0, g1) of window w for segment s of segs: the window's
// positions split into segs 64-aligned pieces.
__device__ inline void segmentRange(const hipibs_window &win, uint64_t nBases, uint32_t k,
                                    uint32_t s, uint32_t segs, uint64_t &g0, uint64_t &g1)
{
    uint64_t npos = win.n_bases >= k ? (uint64_t)win.n_bases - k + 1 : 0;
    // defensive clamp to the stream (the host checks windows it can see)
    if (win.base + (uint64_t)win.n_bases > nBases)
        npos = (nBases >= win.base + k) ? nBases - win.base - k + 1 : 0;
    const uint64_t per = ((npos + segs - 1) / segs + 63) & ~63ull;
    g0 = win.base + std::min<uint64_t>(npos, (uint64_t)s * per);
    g1 = win.base + std::min<uint64_t>(npos, (uint64_t)(s + 1) * per);
}

} // namespace hipibs
