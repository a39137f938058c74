// ref_driver.cpp -- a thin extern "C" shell over the reference's own host
// code, compiled from the sources where they lie (see ref_build.sh) into
// oracle/_ref/libref.so. TEST INFRASTRUCTURE ONLY: it is the checker the
// oracle and the fixtures are pinned against, and the CPU baseline of the
// benchmark; the product never loads it.
//
// Reference code exercised (everything that computes something):
//   readSequenceFile, chunkSequences            src/Utils.hpp:170-327
//   KmerCodec::encodePositions / getCanonicalCodes  src/KmerCodec.hpp:38-100
//   Window::addVariationOriginal                src/Window.hpp:29-36
//   BitStats::run                               src/gpu/EncodeProbeCore.hpp:338-370
//   CKMCFile::OpenForRA / IsKmer / CheckKmer    KMC/kmc_api/kmc_file.cpp
//   CKmerAPI::from_int                          KMC/kmc_api/kmer_api.h:521-544
//   KmcRaIndex (load, probe, forEachKmer)       src/gpu/KmcRaIndex.hpp
//   fastibs::probeBatched (CPU-BATCHED)         src/CpuProbeBatched.hpp:214-268
//   getCanonicalKmersLegacy, reverseComplement  src/Utils.hpp:102-125 (k > 32)
//   CKmerAPI::from_string                       KMC/kmc_api/kmer_api.h
// What this file adds is glue: the TSV / coverage text formats of
// KmerDatabase::writeStats / serializeCoverage (src/KmerDatabase.hpp:756-772,
// :935-948; KmerDatabase.hpp itself needs Boost, absent here) and the
// thread fan-out of the baseline timer.
#include "Utils.hpp"
#include "KmerCodec.hpp"
#include "Window.hpp"
#include "gpu/KmcRaIndex.hpp"
#include "gpu/EncodeProbeCore.hpp"
#include "CpuProbeBatched.hpp"
#include "kmc_file.h"

#include <charconv>
#include <chrono>
#include <fstream>
#include <thread>

namespace
{

bool openRa(CKMCFile &db, const std::string &prefix)
{
    return db.OpenForRA(prefix);
}

// Per-window statistics from codec output and hit flags, through the
// reference's BitStats::run and Window::addVariationOriginal.
Window statsOf(const SeqWindow &w, uint32_t k, const std::vector<uint8_t> &valid,
               const std::vector<uint8_t> &hit)
{
    const uint32_t nPos = (uint32_t)valid.size();
    std::vector<uint32_t> vw((nPos + 31) / 32 + 1, 0), hw((nPos + 31) / 32 + 1, 0);
    for (uint32_t p = 0; p < nPos; ++p)
    {
        if (valid[p])
            vw[p / 32] |= 1u << (p % 32);
        if (hit[p])
            hw[p / 32] |= 1u << (p % 32);
    }
    Window stats;
    stats.id = *w.id;
    stats.start = (int)w.start;
    stats.end = (int)w.end;
    fastibs::BitStats bs;
    fastibs::BitStats::run(hw.data(), vw.data(), 0, nPos, bs,
                           [&](uint64_t gap) { stats.addVariationOriginal((int)gap, (int)k); });
    stats.totalKmers = (int)bs.totalKmers;
    stats.observedKmers = (int)bs.observedKmers;
    return stats;
}

void writeTsv(const std::string &out, const std::vector<Window> &rows)
{
    std::ofstream f(out);
    f << "seqname\tstart\tend\ttotal_kmers\tobserved_kmers\tvariations\tkmer_distance\n";
    for (const auto &s : rows)
        f << s.id << '\t' << s.start << '\t' << s.end << '\t' << s.totalKmers << '\t'
          << s.observedKmers << '\t' << s.variations << '\t' << s.kmerDistance << std::endl;
}

} // namespace

extern "C" {

// Scan TSV of a FASTA against a KMC database. mode 0: CKMCFile::IsKmer per
// code (CPU-OPT); mode 1: KmcRaIndex + probeBatched, K=16 (CPU-BATCHED).
int ref_scan_tsv(const char *dbPrefix, const char *fasta, int W, const char *outTsv, int mode)
{
    try
    {
        const bool gz = std::string(fasta).find(".gz") != std::string::npos;
        ParseResult parsed = readSequenceFile(fasta, gz);
        CKMCFile db;
        std::unique_ptr<KmcRaIndex> idx;
        uint32_t k;
        if (mode == 0)
        {
            if (!openRa(db, dbPrefix))
                return -1;
            k = db.KmerLength();
        }
        else
        {
            idx = std::make_unique<KmcRaIndex>(dbPrefix);
            k = idx->getParams().kmerLength;
        }
        std::vector<SeqWindow> windows;
        chunkSequences(parsed.ids, parsed.sequences, (size_t)W, k, windows);
        std::vector<Window> rows;
        rows.reserve(windows.size());
        KmerCodec codec(k);
        CKmerAPI kmer(k);
        std::vector<uint64_t> codes;
        std::vector<uint8_t> valid, hit;
        for (const auto &w : windows)
        {
            codec.encodePositions(w.seq, codes, valid);
            hit.assign(codes.size(), 0);
            if (mode == 0)
            {
                for (size_t i = 0; i < codes.size(); ++i)
                    if (valid[i])
                    {
                        kmer.from_int(codes[i], k);
                        hit[i] = db.IsKmer(kmer) ? 1 : 0;
                    }
            }
            else
                fastibs::probeBatched(*idx, codes.data(), valid.data(), codes.size(), hit.data(),
                                      16);
            rows.push_back(statsOf(w, k, valid, hit));
        }
        writeTsv(outTsv, rows);
        return (int)rows.size();
    }
    catch (const std::exception &e)
    {
        fprintf(stderr, "ref_scan_tsv: %s\n", e.what());
        return -2;
    }
}

// Coverage file (fastibsmapper format: id line, comma-separated coverage
// line, per contig) through CKMCFile::IsKmer.
int ref_coverage(const char *dbPrefix, const char *fasta, const char *outPath)
{
    try
    {
        const bool gz = std::string(fasta).find(".gz") != std::string::npos;
        ParseResult parsed = readSequenceFile(fasta, gz);
        CKMCFile db;
        if (!openRa(db, dbPrefix))
            return -1;
        const uint32_t k = db.KmerLength();
        KmerCodec codec(k);
        CKmerAPI kmer(k);
        std::ofstream f(outPath);
        std::vector<uint64_t> codes;
        std::vector<uint8_t> valid;
        for (size_t c = 0; c < parsed.sequences.size(); ++c)
        {
            const std::string &seq = parsed.sequences[c];
            codec.encodePositions(seq, codes, valid);
            std::vector<int> diff(seq.size() + 1, 0);
            for (size_t i = 0; i < codes.size(); ++i)
            {
                if (!valid[i])
                    continue;
                kmer.from_int(codes[i], k);
                if (db.IsKmer(kmer))
                {
                    diff[i] += 1;
                    diff[i + k] -= 1;
                }
            }
            std::string line;
            char buf[12];
            int cov = 0;
            for (size_t i = 0; i < seq.size(); ++i)
            {
                cov += diff[i];
                auto r = std::to_chars(buf, buf + sizeof(buf), cov);
                line.append(buf, r.ptr);
                line.push_back(',');
            }
            if (!line.empty())
                line.pop_back();
            f << parsed.ids[c] << '\n' << line << std::endl;
        }
        return (int)parsed.sequences.size();
    }
    catch (const std::exception &e)
    {
        fprintf(stderr, "ref_coverage: %s\n", e.what());
        return -2;
    }
}

// The reference's string path (CPU-BASE, the one it takes for k > 32):
// getCanonicalKmersLegacy + CKmerAPI::from_string + CKMCFile::IsKmer per
// window, with the gap loop of KmerDatabase::getStatsFromSequenceLegacy
// (src/KmerDatabase.hpp:675-709; glue, that header needs Boost).
int ref_scan_tsv_legacy(const char *dbPrefix, const char *fasta, int W, const char *outTsv)
{
    try
    {
        const bool gz = std::string(fasta).find(".gz") != std::string::npos;
        ParseResult parsed = readSequenceFile(fasta, gz);
        CKMCFile db;
        if (!openRa(db, dbPrefix))
            return -1;
        const uint32_t k = db.KmerLength();
        std::vector<SeqWindow> windows;
        chunkSequences(parsed.ids, parsed.sequences, (size_t)W, k, windows);
        std::vector<Window> rows;
        rows.reserve(windows.size());
        for (const auto &w : windows)
        {
            Window stats;
            stats.id = *w.id;
            stats.start = (int)w.start;
            stats.end = (int)w.end;
            std::string seq(w.seq);
            const auto kmers = getCanonicalKmersLegacy(seq, (int)k);
            stats.totalKmers = (int)kmers.size();
            int gap = 0;
            CKmerAPI km;
            for (const auto &s : kmers)
            {
                km.from_string(s);
                if (!db.IsKmer(km))
                    gap += 1;
                else
                {
                    stats.observedKmers += 1;
                    if (gap > 0)
                    {
                        stats.addVariationOriginal(gap, (int)k);
                        gap = 0;
                    }
                }
            }
            if (gap > 0)
                stats.addVariationOriginal(gap, (int)k);
            rows.push_back(stats);
        }
        writeTsv(outTsv, rows);
        return (int)rows.size();
    }
    catch (const std::exception &e)
    {
        fprintf(stderr, "ref_scan_tsv_legacy: %s\n", e.what());
        return -2;
    }
}

// The string path's coverage (KmerDatabase::getMappingFromSequenceLegacy,
// src/KmerDatabase.hpp:804-845; glue): per position, the upper-cased
// k-mer, skipped if not ACGT, canonical by reverseComplement and string
// order, CKMCFile::IsKmer; a hit adds 1 to its k bases.
int ref_coverage_legacy(const char *dbPrefix, const char *fasta, const char *outPath)
{
    try
    {
        const bool gz = std::string(fasta).find(".gz") != std::string::npos;
        ParseResult parsed = readSequenceFile(fasta, gz);
        CKMCFile db;
        if (!openRa(db, dbPrefix))
            return -1;
        const size_t k = db.KmerLength();
        std::ofstream f(outPath);
        for (size_t c = 0; c < parsed.sequences.size(); ++c)
        {
            const std::string &seq = parsed.sequences[c];
            std::vector<int> cov(seq.size(), 0);
            for (size_t i = 0; i + k <= seq.size(); ++i)
            {
                std::string km = seq.substr(i, k);
                for (auto &ch : km)
                    ch = (char)toupper((unsigned char)ch);
                if (km.find_first_not_of("ACGT") != std::string::npos)
                    continue;
                const std::string rc = reverseComplement(km);
                CKmerAPI api;
                api.from_string(km > rc ? rc : km);
                if (db.IsKmer(api))
                    for (size_t j = 0; j < k; ++j)
                        cov[i + j] += 1;
            }
            std::string line;
            char buf[12];
            for (size_t i = 0; i < seq.size(); ++i)
            {
                auto r = std::to_chars(buf, buf + sizeof(buf), cov[i]);
                line.append(buf, r.ptr);
                line.push_back(',');
            }
            if (!line.empty())
                line.pop_back();
            f << parsed.ids[c] << '\n' << line << std::endl;
        }
        return (int)parsed.sequences.size();
    }
    catch (const std::exception &e)
    {
        fprintf(stderr, "ref_coverage_legacy: %s\n", e.what());
        return -2;
    }
}

// Membership of n k-mer strings of length k (concatenated) through
// CKmerAPI::from_string + CKMCFile::IsKmer.
long long ref_kmc_is_kmer_strings(const char *prefix, const char *kmers, long long n, uint8_t *hit)
{
    CKMCFile db;
    if (!openRa(db, prefix))
        return -1;
    const size_t k = db.KmerLength();
    CKmerAPI api;
    long long c = 0;
    for (long long i = 0; i < n; ++i)
    {
        api.from_string(std::string(kmers + i * k, k));
        hit[i] = db.IsKmer(api) ? 1 : 0;
        c += hit[i];
    }
    return c;
}

// Every k-mer of a database via KmcRaIndex::forEachKmer; returns the count.
long long ref_list_keys(const char *prefix, uint64_t *out, long long cap)
{
    try
    {
        KmcRaIndex idx(prefix);
        long long n = 0;
        idx.forEachKmer([&](uint64_t key)
                        {
                            if (n < cap)
                                out[n] = key;
                            ++n;
                        });
        return n;
    }
    catch (const std::exception &e)
    {
        fprintf(stderr, "ref_list_keys: %s\n", e.what());
        return -1;
    }
}

// Membership of n codes through CKMCFile::IsKmer (1/0 per code).
long long ref_kmc_is_kmer(const char *prefix, const uint64_t *codes, long long n, uint8_t *hit)
{
    CKMCFile db;
    if (!openRa(db, prefix))
        return -1;
    const uint32_t k = db.KmerLength();
    CKmerAPI kmer(k);
    long long c = 0;
    for (long long i = 0; i < n; ++i)
    {
        kmer.from_int(codes[i], k);
        hit[i] = db.IsKmer(kmer) ? 1 : 0;
        c += hit[i];
    }
    return c;
}

// Membership of n codes through KmcRaIndex::probe.
long long ref_ra_probe(const char *prefix, const uint64_t *codes, long long n, uint8_t *hit)
{
    try
    {
        KmcRaIndex idx(prefix);
        long long c = 0;
        for (long long i = 0; i < n; ++i)
        {
            hit[i] = idx.probe(codes[i]) ? 1 : 0;
            c += hit[i];
        }
        return c;
    }
    catch (const std::exception &e)
    {
        fprintf(stderr, "ref_ra_probe: %s\n", e.what());
        return -1;
    }
}

// KDBIntersect's count: the database with fewer k-mers enumerated, probed
// against the other one's host image.
long long ref_intersect(const char *a, const char *b)
{
    try
    {
        KmcRaIndex ia(a), ib(b);
        const bool aSmall = ia.getParams().totalKmers <= ib.getParams().totalKmers;
        const KmcRaIndex &small = aSmall ? ia : ib;
        const KmcRaIndex &big = aSmall ? ib : ia;
        long long c = 0;
        small.forEachKmer([&](uint64_t key) { c += big.probe(key) ? 1 : 0; });
        return c;
    }
    catch (const std::exception &e)
    {
        fprintf(stderr, "ref_intersect: %s\n", e.what());
        return -1;
    }
}

// ---- CPU baseline: the reference's CPU-BATCHED scan over an in-memory
// sequence, windows spread over `threads` threads. Per window: codec ->
// probeBatched(K) -> gap run-lengths (the loop of
// src/KmerDatabase.hpp:625-642). Returns elapsed seconds; stats gets
// {total, observed, variations, distance} per window. ----
void *ref_index_open(const char *prefix)
{
    try
    {
        return new KmcRaIndex(prefix);
    }
    catch (const std::exception &e)
    {
        fprintf(stderr, "ref_index_open: %s\n", e.what());
        return nullptr;
    }
}

void ref_index_close(void *h)
{
    delete static_cast<KmcRaIndex *>(h);
}

double ref_batched_scan(void *h, const char *seq, const long long *wStart, const long long *wLen,
                        long long nWin, int K, int threads, int *stats)
{
    const KmcRaIndex &idx = *static_cast<KmcRaIndex *>(h);
    const uint32_t k = idx.getParams().kmerLength;
    const auto t0 = std::chrono::steady_clock::now();
    std::atomic<long long> next{0};
    auto worker = [&]()
    {
        KmerCodec codec(k);
        std::vector<uint64_t> codes;
        std::vector<uint8_t> hits;
        for (;;)
        {
            const long long i = next.fetch_add(1);
            if (i >= nWin)
                break;
            codec.getCanonicalCodes(std::string_view(seq + wStart[i], (size_t)wLen[i]), codes);
            hits.assign(codes.size(), 0);
            fastibs::probeBatched(idx, codes.data(), nullptr, codes.size(), hits.data(),
                                  (unsigned)K);
            Window w;
            w.totalKmers = (int)codes.size();
            int gap = 0;
            for (const auto hit : hits)
            {
                if (!hit)
                    gap += 1;
                else
                {
                    w.observedKmers += 1;
                    if (gap > 0)
                    {
                        w.addVariationOriginal(gap, (int)k);
                        gap = 0;
                    }
                }
            }
            if (gap > 0)
                w.addVariationOriginal(gap, (int)k);
            stats[4 * i + 0] = w.totalKmers;
            stats[4 * i + 1] = w.observedKmers;
            stats[4 * i + 2] = w.variations;
            stats[4 * i + 3] = w.kmerDistance;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(1, threads); ++t)
        pool.emplace_back(worker);
    for (auto &t : pool)
        t.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The reference's CPU-BATCHED whole scan of one FASTA against one KMC
// database, phase by phase as its processReference runs it
// (src/KmerDatabase.hpp:950-1065 with FASTIBS_CPU_BATCHED): open the
// KmcRaIndex, read the FASTA (readSequenceFile), cut windows
// (chunkSequences), per window codec -> probeBatched(K) -> statistics on
// `threads` threads, write the TSV. phases[0..4] = index load, FASTA read,
// windows, probe + statistics, write (seconds); returns the number of
// windows, or < 0 on failure. probes receives the valid positions probed.
long long ref_batched_whole_scan(const char *dbPrefix, const char *fasta, int W,
                                 const char *outTsv, int K, int threads, double *phases,
                                 long long *probes)
{
    try
    {
        using clk = std::chrono::steady_clock;
        auto t = clk::now();
        auto lap = [&t]() {
            const auto n = clk::now();
            const double s = std::chrono::duration<double>(n - t).count();
            t = n;
            return s;
        };
        KmcRaIndex idx(dbPrefix);
        const uint32_t k = idx.getParams().kmerLength;
        phases[0] = lap();
        const bool gz = std::string(fasta).find(".gz") != std::string::npos;
        ParseResult parsed = readSequenceFile(fasta, gz);
        phases[1] = lap();
        std::vector<SeqWindow> windows;
        chunkSequences(parsed.ids, parsed.sequences, (size_t)W, k, windows);
        phases[2] = lap();
        std::vector<Window> rows(windows.size());
        std::atomic<long long> next{0}, probed{0};
        auto worker = [&]()
        {
            KmerCodec codec(k);
            std::vector<uint64_t> codes;
            std::vector<uint8_t> valid, hit;
            long long mine = 0;
            for (;;)
            {
                const long long i = next.fetch_add(1);
                if (i >= (long long)windows.size())
                    break;
                codec.encodePositions(windows[i].seq, codes, valid);
                hit.assign(codes.size(), 0);
                fastibs::probeBatched(idx, codes.data(), valid.data(), codes.size(), hit.data(),
                                      (unsigned)K);
                rows[i] = statsOf(windows[i], k, valid, hit);
                mine += rows[i].totalKmers;
            }
            probed += mine;
        };
        std::vector<std::thread> pool;
        for (int i = 0; i < std::max(1, threads); ++i)
            pool.emplace_back(worker);
        for (auto &th : pool)
            th.join();
        phases[3] = lap();
        writeTsv(outTsv, rows);
        phases[4] = lap();
        *probes = probed.load();
        return (long long)rows.size();
    }
    catch (const std::exception &e)
    {
        fprintf(stderr, "ref_batched_whole_scan: %s\n", e.what());
        return -2;
    }
}

} // extern "C"
