/*
 * oracle.c -- CPU restatement of the reference's k-mer IBS scan path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the
 * checker. The product (libhipibs.so) never calls it.
 *
 * Plain C, straight-line loops, no attempt at speed. Every function names
 * the reference lines it restates. How each function is pinned (golden
 * vectors, the compiled reference under oracle/_ref, closed forms, brute
 * force) is listed in tests/test_oracle.py; a function without such a
 * check would say "parity unpinned" here -- none does.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Base table of KmerCodec (src/KmerCodec.hpp:105-118). */
static int base_code(unsigned char c)
{
    switch (c)
    {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

/* KmerCodec::encodePositions (src/KmerCodec.hpp:67-100): for i in
 * 0..n-k, codes[i] = canonical code of seq[i..i+k), valid[i] = all k bases
 * ACGT (codes[i] = 0 when invalid). Returns the number of positions. */
int64_t orc_encode_positions(const char *seq, int64_t n, uint32_t k, uint64_t *codes,
                             uint8_t *valid)
{
    if (k == 0 || k > 32 || n < (int64_t)k)
        return 0;
    const uint64_t mask = k == 32 ? ~0ull : (1ull << (2 * k)) - 1;
    const uint32_t rcShift = 2 * (k - 1);
    uint64_t fwd = 0, rc = 0;
    uint32_t run = 0;
    for (int64_t j = 0; j < n; ++j)
    {
        int c = base_code((unsigned char)seq[j]);
        if (c > 3)
        {
            run = 0;
            c = 0;
        }
        else
            ++run;
        fwd = ((fwd << 2) | (uint64_t)c) & mask;
        rc = (rc >> 2) | ((uint64_t)(3 - c) << rcShift);
        if (j >= (int64_t)k - 1)
        {
            const int64_t i = j - (k - 1);
            const int ok = run >= k;
            valid[i] = (uint8_t)ok;
            codes[i] = ok ? (fwd < rc ? fwd : rc) : 0;
        }
    }
    return n - k + 1;
}

/* Window geometry of chunkSequences (src/Utils.hpp:296-327) for one contig
 * of length len: windows [j, min(j+W, len)) for j = 0, W-k, 2(W-k), ...
 * while j < len. Writes up to cap (start, end) pairs; returns the count,
 * or -1 when W <= k (the reference throws). */
int64_t orc_windows(int64_t len, int64_t W, int64_t k, int64_t *starts, int64_t *ends, int64_t cap)
{
    if (W <= k)
        return -1;
    int64_t n = 0;
    for (int64_t j = 0; j < len; j += W - k)
    {
        if (n < cap)
        {
            starts[n] = j;
            ends[n] = j + W > len ? len : j + W;
        }
        ++n;
    }
    return n;
}

static int cmp_u64(const void *a, const void *b)
{
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* Membership in a sorted key array (the set semantics of
 * CKMCFile::CheckKmer / KmcRaIndex::probe, src/gpu/ProbeCore.hpp:113-129). */
static int member(const uint64_t *keys, int64_t nkeys, uint64_t code)
{
    int64_t lo = 0, hi = nkeys - 1;
    while (lo <= hi)
    {
        const int64_t mid = lo + (hi - lo) / 2;
        if (keys[mid] == code)
            return 1;
        if (keys[mid] < code)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return 0;
}

void orc_sort_keys(uint64_t *keys, int64_t n)
{
    qsort(keys, (size_t)n, sizeof(uint64_t), cmp_u64);
}

/* Window::addVariationOriginal (src/Window.hpp:29-36). */
static void add_variation(int32_t *var, int32_t *dist, int32_t gap, int32_t k)
{
    *var += 1;
    int32_t d = gap - (k - 1);
    if (d <= 0)
        d = abs(d + 1);
    *dist += d;
}

/* KmerDatabase::getStatsFromSequence, CPU path
 * (src/KmerDatabase.hpp:546-673): canonical codes of the window's valid
 * positions in order (KmerCodec::getCanonicalCodes, src/KmerCodec.hpp:38-61),
 * membership per code, and the gap run-length loop. out = {total,
 * observed, variations, kmer_distance}. keys must be sorted. */
void orc_window_stats(const char *seq, int64_t n, uint32_t k, const uint64_t *keys, int64_t nkeys,
                      int32_t out[4])
{
    out[0] = out[1] = out[2] = out[3] = 0;
    if (n < (int64_t)k)
        return;
    const int64_t np = n - k + 1;
    uint64_t *codes = (uint64_t *)malloc((size_t)np * 8);
    uint8_t *valid = (uint8_t *)malloc((size_t)np);
    orc_encode_positions(seq, n, k, codes, valid);
    int32_t gap = 0;
    for (int64_t i = 0; i < np; ++i)
    {
        if (!valid[i])
            continue;
        out[0] += 1;
        if (!member(keys, nkeys, codes[i]))
            gap += 1;
        else
        {
            out[1] += 1;
            if (gap > 0)
            {
                add_variation(&out[2], &out[3], gap, (int32_t)k);
                gap = 0;
            }
        }
    }
    if (gap > 0)
        add_variation(&out[2], &out[3], gap, (int32_t)k);
    free(codes);
    free(valid);
}

/* The same statistics from a per-position hit flag array (hit[i] set only
 * for valid positions): the consumer of the GPU path, BitStats::run +
 * addVariationOriginal (src/gpu/EncodeProbeCore.hpp:338-370). */
void orc_stats_from_flags(const uint8_t *valid, const uint8_t *hit, int64_t np, uint32_t k,
                          int32_t out[4])
{
    out[0] = out[1] = out[2] = out[3] = 0;
    int32_t gap = 0;
    for (int64_t i = 0; i < np; ++i)
    {
        if (!valid[i])
            continue;
        out[0] += 1;
        if (hit[i])
        {
            out[1] += 1;
            if (gap > 0)
            {
                add_variation(&out[2], &out[3], gap, (int32_t)k);
                gap = 0;
            }
        }
        else
            gap += 1;
    }
    if (gap > 0)
        add_variation(&out[2], &out[3], gap, (int32_t)k);
}

/* Per-base coverage of one contig, KmerDatabase::getMappingFromSequence +
 * serializeCoverage (src/KmerDatabase.hpp:756-802): +1 at each valid hit
 * position p, -1 at p+k, prefix sum over the contig's bases. */
void orc_coverage(const char *seq, int64_t n, uint32_t k, const uint64_t *keys, int64_t nkeys,
                  int32_t *cov)
{
    int32_t *diff = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    if (n >= (int64_t)k)
    {
        const int64_t np = n - k + 1;
        uint64_t *codes = (uint64_t *)malloc((size_t)np * 8);
        uint8_t *valid = (uint8_t *)malloc((size_t)np);
        orc_encode_positions(seq, n, k, codes, valid);
        for (int64_t i = 0; i < np; ++i)
            if (valid[i] && member(keys, nkeys, codes[i]))
            {
                diff[i] += 1;
                diff[i + k] -= 1;
            }
        free(codes);
        free(valid);
    }
    int32_t c = 0;
    for (int64_t i = 0; i < n; ++i)
    {
        c += diff[i];
        cov[i] = c;
    }
    free(diff);
}

/* KDBIntersect's count (src/KDBIntersect.cpp:240-275): keys of a present
 * in b. Both arrays sorted, distinct. */
int64_t orc_intersect(const uint64_t *a, int64_t na, const uint64_t *b, int64_t nb)
{
    int64_t i = 0, j = 0, c = 0;
    while (i < na && j < nb)
    {
        if (a[i] == b[j])
        {
            ++c;
            ++i;
            ++j;
        }
        else if (a[i] < b[j])
            ++i;
        else
            ++j;
    }
    return c;
}

/* ---- KMC database parsing, KmcRaIndex::readHeaderOf / loadPre / loadSuf /
 * walkBuckets (src/gpu/KmcRaIndex.hpp:497-708, :437-451). Returns the
 * number of k-mers (keys written up to cap) or a negative error. ---- */
static int read_at(FILE *f, long off, void *dst, size_t n)
{
    if (fseek(f, off, SEEK_SET) != 0)
        return 0;
    return fread(dst, 1, n, f) == n;
}

int64_t orc_kmc_read(const char *prefix, uint64_t *keys, int64_t cap, uint32_t *k_out)
{
    char path[4096];
    snprintf(path, sizeof(path), "%s.kmc_pre", prefix);
    FILE *f = fopen(path, "rb");
    if (!f)
        return -1;
    fseek(f, 0, SEEK_END);
    const long fsize = ftell(f);
    uint32_t version = 0;
    unsigned char off8 = 0;
    if (!read_at(f, fsize - 12, &version, 4) || !read_at(f, fsize - 8, &off8, 1))
    {
        fclose(f);
        return -2;
    }
    const long hpos = fsize - ((long)off8 + 8);
    uint32_t h[7] = {0};
    uint64_t total = 0;
    uint32_t k, p;
    uint64_t nLut;
    if (version == 0x200)
    {
        read_at(f, hpos, h, 28);
        read_at(f, hpos + 28, &total, 8);
        k = h[0];
        p = h[3];
        const uint64_t sigEntries = (1ull << (2 * h[4])) + 1;
        const uint64_t lutArea = (uint64_t)fsize - 12 - (sigEntries * 4 + off8 + 8);
        nLut = lutArea / 8 + 1;
    }
    else if (version == 0)
    {
        read_at(f, hpos, h, 24);
        read_at(f, hpos + 24, &total, 8);
        k = h[0];
        p = h[3];
        nLut = (1ull << (2 * p)) + 1;
    }
    else
    {
        fclose(f);
        return -3;
    }
    const uint32_t counter = h[2];
    uint64_t *lut = (uint64_t *)malloc(nLut * 8);
    read_at(f, 4, lut, nLut * 8);
    fclose(f);
    lut[nLut - 1] = total + 1;
    if (k_out)
        *k_out = k;
    snprintf(path, sizeof(path), "%s.kmc_suf", prefix);
    f = fopen(path, "rb");
    if (!f)
    {
        free(lut);
        return -4;
    }
    const uint32_t sufBytes = (k - p) / 4, rec = sufBytes + counter;
    unsigned char *buf = (unsigned char *)malloc(rec ? rec : 1);
    const uint32_t shift = 2 * (k - p);
    const uint64_t single = 1ull << (2 * p);
    int64_t w = 0;
    fseek(f, 4, SEEK_SET);
    for (uint64_t j = 0; j + 1 < nLut; ++j)
    {
        const uint64_t hi = lut[j + 1] < total ? lut[j + 1] : total;
        for (uint64_t r = lut[j]; r < hi; ++r)
        {
            if (fread(buf, 1, rec, f) != rec)
            {
                free(buf);
                free(lut);
                fclose(f);
                return -5;
            }
            uint64_t v = 0;
            for (uint32_t b = 0; b < sufBytes; ++b)
                v = (v << 8) | buf[b];
            if (w < cap)
                keys[w] = ((j % single) << shift) | v;
            ++w;
        }
    }
    free(buf);
    free(lut);
    fclose(f);
    return w;
}

/* All windows of one contig (orc_windows geometry), statistics per window
 * as orc_window_stats. out gets 6 int64 per window: start, end, total,
 * observed, variations, distance. Returns the window count (-1 if W <= k). */
int64_t orc_scan_contig(const char *seq, int64_t n, int64_t W, uint32_t k, const uint64_t *keys,
                        int64_t nkeys, int64_t *out, int64_t cap)
{
    if (W <= (int64_t)k)
        return -1;
    int64_t nw = 0;
    for (int64_t j = 0; j < n; j += W - k)
    {
        const int64_t e = j + W > n ? n : j + W;
        if (nw < cap)
        {
            int32_t st[4];
            orc_window_stats(seq + j, e - j, k, keys, nkeys, st);
            int64_t *o = out + 6 * nw;
            o[0] = j;
            o[1] = e;
            o[2] = st[0];
            o[3] = st[1];
            o[4] = st[2];
            o[5] = st[3];
        }
        ++nw;
    }
    return nw;
}

/* k-mer counting with KMC 3.2.4's rules, the step that makes the
 * reference's sample databases (README.md:95 `kmc -k25 -ci2 -cs255`; KMC
 * is an external tool, not in the reference's sources): every position of
 * text whose k bytes are ACGTacgt (a record separator or N ends a run)
 * gives its canonical code (orc_encode_positions); the codes are sorted and
 * each distinct code is written with its number of occurrences, ascending.
 * Writes at most cap pairs; returns the number of distinct codes and the
 * number of k-mers in *n_kmers. */
static void radix_sort_u64(uint64_t *a, uint64_t *tmp, int64_t n)
{
    for (int shift = 0; shift < 64; shift += 8)
    {
        int64_t count[257] = {0};
        for (int64_t i = 0; i < n; ++i)
            count[((a[i] >> shift) & 255) + 1] += 1;
        for (int b = 0; b < 256; ++b)
            count[b + 1] += count[b];
        for (int64_t i = 0; i < n; ++i)
            tmp[count[(a[i] >> shift) & 255]++] = a[i];
        memcpy(a, tmp, (size_t)n * sizeof(uint64_t));
    }
}

int64_t orc_count_kmers(const char *text, int64_t n, uint32_t k, uint64_t *keys,
                        uint32_t *counts, int64_t cap, int64_t *n_kmers)
{
    *n_kmers = 0;
    if (k == 0 || k > 32 || n < (int64_t)k)
        return 0;
    const int64_t np = n - k + 1;
    uint64_t *codes = malloc((size_t)np * sizeof(uint64_t));
    uint64_t *tmp = malloc((size_t)np * sizeof(uint64_t));
    uint8_t *valid = malloc((size_t)np);
    if (!codes || !tmp || !valid)
    {
        free(codes);
        free(tmp);
        free(valid);
        return -1;
    }
    orc_encode_positions(text, n, k, codes, valid);
    int64_t m = 0;
    for (int64_t i = 0; i < np; ++i)
        if (valid[i])
            codes[m++] = codes[i];
    free(valid);
    radix_sort_u64(codes, tmp, m);
    free(tmp);
    int64_t d = 0;
    for (int64_t i = 0; i < m;)
    {
        int64_t j = i;
        while (j < m && codes[j] == codes[i])
            ++j;
        if (d < cap)
        {
            keys[d] = codes[i];
            counts[d] = (uint32_t)(j - i);
        }
        ++d;
        i = j;
    }
    free(codes);
    *n_kmers = m;
    return d;
}

/* ---- k = 33..64: two-word codes ---------------------------------------
 * The reference takes its string path for k > 32 (KmerDatabase.hpp:104-123):
 * getCanonicalKmersLegacy (src/Utils.hpp:102-125) keeps every k-length
 * substring made only of ACGT after upper-casing, canonical = the smaller
 * of it and its reverse complement in string order; the statistics loop is
 * getStatsFromSequenceLegacy (src/KmerDatabase.hpp:675-709), coverage
 * getMappingFromSequenceLegacy (:804-845). Restated here with the code of
 * a k-mer as a 128-bit integer (first base most significant, A=0 C=1 G=2
 * T=3), whose numeric order on ACGT strings is the string order. Keys are
 * (high word, low word) pairs, sorted as 128-bit integers. */
typedef unsigned __int128 u128;

static u128 wide_of(const uint64_t *p)
{
    return ((u128)p[0] << 64) | p[1];
}

int64_t orc_encode_positions_wide(const char *seq, int64_t n, uint32_t k, uint64_t *codes,
                                  uint8_t *valid)
{
    if (k < 33 || k > 64 || n < (int64_t)k)
        return 0;
    const u128 mask = k == 64 ? ~(u128)0 : (((u128)1 << (2 * k)) - 1);
    const uint32_t rcShift = 2 * (k - 1);
    u128 fwd = 0, rc = 0;
    uint32_t run = 0;
    for (int64_t j = 0; j < n; ++j)
    {
        int c = base_code((unsigned char)seq[j]);
        if (c > 3)
        {
            run = 0;
            c = 0;
        }
        else
            ++run;
        fwd = ((fwd << 2) | (u128)c) & mask;
        rc = (rc >> 2) | ((u128)(3 - c) << rcShift);
        if (j >= (int64_t)k - 1)
        {
            const int64_t i = j - (k - 1);
            const int ok = run >= k;
            const u128 v = ok ? (fwd < rc ? fwd : rc) : 0;
            valid[i] = (uint8_t)ok;
            codes[2 * i] = (uint64_t)(v >> 64);
            codes[2 * i + 1] = (uint64_t)v;
        }
    }
    return n - k + 1;
}

static int cmp_u128(const void *a, const void *b)
{
    const u128 x = wide_of((const uint64_t *)a), y = wide_of((const uint64_t *)b);
    return x < y ? -1 : (x > y ? 1 : 0);
}

void orc_sort_keys_wide(uint64_t *keys, int64_t n)
{
    qsort(keys, (size_t)n, 2 * sizeof(uint64_t), cmp_u128);
}

static int member_wide(const uint64_t *keys, int64_t nkeys, u128 code)
{
    int64_t lo = 0, hi = nkeys - 1;
    while (lo <= hi)
    {
        const int64_t mid = lo + (hi - lo) / 2;
        const u128 v = wide_of(keys + 2 * mid);
        if (v == code)
            return 1;
        if (v < code)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return 0;
}

/* Hit flag per position (valid positions only) against sorted pair keys. */
static void hits_wide(const char *seq, int64_t n, uint32_t k, const uint64_t *keys,
                      int64_t nkeys, uint8_t *valid, uint8_t *hit)
{
    const int64_t np = n - k + 1;
    uint64_t *codes = (uint64_t *)malloc((size_t)np * 16);
    orc_encode_positions_wide(seq, n, k, codes, valid);
    for (int64_t i = 0; i < np; ++i)
        hit[i] = valid[i] && member_wide(keys, nkeys, wide_of(codes + 2 * i));
    free(codes);
}

/* Window rows as orc_scan_contig, for k = 33..64. */
int64_t orc_scan_contig_wide(const char *seq, int64_t n, int64_t W, uint32_t k,
                             const uint64_t *keys, int64_t nkeys, int64_t *out, int64_t cap)
{
    if (W <= (int64_t)k || k < 33 || k > 64)
        return -1;
    int64_t nw = 0;
    for (int64_t j = 0; j < n; j += W - k)
    {
        const int64_t e = j + W > n ? n : j + W;
        if (nw < cap)
        {
            int32_t st[4] = {0, 0, 0, 0};
            if (e - j >= (int64_t)k)
            {
                const int64_t np = e - j - k + 1;
                uint8_t *valid = (uint8_t *)malloc((size_t)np);
                uint8_t *hit = (uint8_t *)malloc((size_t)np);
                hits_wide(seq + j, e - j, k, keys, nkeys, valid, hit);
                orc_stats_from_flags(valid, hit, np, k, st);
                free(valid);
                free(hit);
            }
            int64_t *o = out + 6 * nw;
            o[0] = j;
            o[1] = e;
            o[2] = st[0];
            o[3] = st[1];
            o[4] = st[2];
            o[5] = st[3];
        }
        ++nw;
    }
    return nw;
}

/* Per-base coverage of one contig as orc_coverage, for k = 33..64. */
void orc_coverage_wide(const char *seq, int64_t n, uint32_t k, const uint64_t *keys,
                       int64_t nkeys, int32_t *cov)
{
    int32_t *diff = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    if (n >= (int64_t)k && k >= 33 && k <= 64)
    {
        const int64_t np = n - k + 1;
        uint8_t *valid = (uint8_t *)malloc((size_t)np);
        uint8_t *hit = (uint8_t *)malloc((size_t)np);
        hits_wide(seq, n, k, keys, nkeys, valid, hit);
        for (int64_t i = 0; i < np; ++i)
            if (hit[i])
            {
                diff[i] += 1;
                diff[i + k] -= 1;
            }
        free(valid);
        free(hit);
    }
    int32_t c = 0;
    for (int64_t i = 0; i < n; ++i)
    {
        c += diff[i];
        cov[i] = c;
    }
    free(diff);
}
