// KmerDatabase: one KMC sample against references -- the scan (per-window
// IBS table) and the mapper (per-base coverage) -- with the reference's
// public interface (src/KmerDatabase.hpp:44-1598) over the MI355X library.
// There is no CPU path: the device table is the only probe structure, and
// a machine without a usable GPU gets an error, not a silent fallback.
#pragma once

#include "Hipibs.hpp"
#include "KmerCodec.hpp"
#include "ScanPipeline.hpp"
#include "SeqIO.hpp"
#include "Window.hpp"

#include <cerrno>
#include <charconv>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <future>
#include <iostream>
#include <string>
#include <thread>
#include <utility>
#include <vector>

class KmerDatabase
{
public:
    // Device and batch knobs (the reference's GpuOptions, set from the CLI).
    struct GpuOptions
    {
        int device = 0;
        uint64_t batchBases = 128ull << 20;
        int depth = 3;
        double loadFactor = 0; // 0: chosen with the table kind (0.3 minimizer, 0.6 whole k-mer)
        uint32_t bucketSlots = 16;
        bool bucketAuto = true; // a panel may time 64 B vs 128 B buckets
        uint64_t hbmBudget = 0;
        unsigned loadThreads = 0;
        int minimizerLen = -1;     // -1: chosen by use; 0: whole-k-mer table; 5..31: minimizer
        uint32_t minimizerCap = 0; // 0: 16 keys per minimizer group
    };
    static GpuOptions &gpuOptions()
    {
        static GpuOptions o;
        return o;
    }

    // What a table is probed with: the consecutive positions of a reference
    // (scan, mapper), or keys in no order (intersect) / many tables per
    // launch (panel).
    enum class TableUse
    {
        Scan,
        Lookup,
    };

    // The table options for a database of n k-mers of length k. Unless
    // given, scans of k = 13..32 get a minimizer table (m = k - 4, groups of
    // up to 16 keys, load 0.2: 40 B per key; runs of positions sharing a
    // minimizer read one bucket), at load 0.3 (26.7 B per key) when that
    // does not fit 80% of the device's free memory with its build, else a
    // whole-k-mer table; every other use a whole-k-mer table at load 0.6
    // (13.3 B per key).
    static hipibs_table_opts tableOptions(uint32_t k, TableUse use, uint64_t n = 0)
    {
        const GpuOptions &g = gpuOptions();
        hipibs_table_opts o = hipibs::defaultOpts();
        o.device = g.device;
        o.bucket_slots = g.bucketSlots;
        o.hbm_budget = g.hbmBudget;
        o.load_threads = g.loadThreads;
        uint32_t m = 0;
        if (g.minimizerLen > 0)
            m = (uint32_t)g.minimizerLen;
        else if (g.minimizerLen < 0 && use == TableUse::Scan && k >= 13 && k <= 32)
            m = k - 4;
        double lf = g.loadFactor;
        if (m && g.minimizerLen < 0 && lf <= 0 && n)
        {
            // the build holds every key (8 B) and a group map (<= 16 B per key) beside
            // the table
            uint64_t freeB = 0, totalB = 0;
            if (hipibs_device_mem(g.device, &freeB, &totalB) == 0 && freeB)
            {
                auto fits = [&](double load) {
                    return (double)n * (8.0 * 1.07 / load + 24.0) <= 0.8 * (double)freeB;
                };
                if (fits(0.2))
                    lf = 0.2;
                else if (fits(0.3))
                    lf = 0.3;
                else
                    m = 0;
            }
        }
        o.minimizer_len = m;
        o.minimizer_cap = m ? (g.minimizerCap ? g.minimizerCap : 16) : 0;
        o.load_factor = lf > 0 ? lf : (m ? 0.2 : 0.6);
        return o;
    }
    // Reads the header only; the device table is built on first use.
    uint64_t kmerCount() const { return info_.total_kmers; }

    explicit KmerDatabase(std::string sourcePath, bool listing = false)
        : sourcePath_(std::move(sourcePath))
    {
        (void)listing;
        hipibs::check(hipibs_kmc_header(sourcePath_.c_str(), &info_), "reading the KMC header");
        kmerSize = info_.kmer_length;
        if (kmerSize == 0 || needsStringPath(kmerSize))
            throw std::runtime_error("k = " + std::to_string(kmerSize) +
                                     ": this build supports k <= 64 only");
    }

    uint32_t getKmerSize() const { return kmerSize; }
    uint64_t totalKmers() const { return info_.total_kmers; }
    const std::string &sourcePath() const { return sourcePath_; }

    void printKMCInfo() const
    {
        std::cout << "KMC version: " << (info_.kmc_version == 0x200 ? "KMC2" : "KMC1") << "\n"
                  << "k: " << info_.kmer_length << "\n"
                  << "Signature length: " << info_.signature_len << "\n"
                  << "LUT prefix length: " << info_.lut_prefix_length << "\n"
                  << "Counter size: " << info_.counter_size << "\n"
                  << "Min count: " << info_.min_count << "\n"
                  << "Max count: " << info_.max_count << "\n"
                  << "Total k-mers: " << info_.total_kmers << std::endl;
    }

    bool gpuActive() const { return table_ != nullptr; }
    const hipibs_table *table() const { return table_.get(); }

    // Builds the device table (kind and geometry by tableOptions): the
    // database's raw records are read on host threads into pinned buffers,
    // decoded into keys and inserted on the device. A table built for one
    // use serves the others too (only more slowly).
    void initGpuIndex(TableUse use = TableUse::Scan)
    {
        if (table_)
            return;
        std::string why;
        if (hipibs::deviceCount(why) < 1)
            throw std::runtime_error("no usable GPU: " + why);
        const hipibs_table_opts o = tableOptions(kmerSize, use, info_.total_kmers);
        const auto t0 = std::chrono::steady_clock::now();
        table_ = hipibs::tableFromKmc(sourcePath_, o);
        const auto ti = hipibs::info(table_.get());
        std::cout << "Device table: " << ti.keys << " keys, " << ti.bytes / 1e9 << " GB, "
                  << ti.slots * 8 << " B buckets, "
                  << (o.minimizer_len ? "minimizer m = " + std::to_string(o.minimizer_len) +
                                            ", cap " + std::to_string(o.minimizer_cap)
                                      : std::string("whole k-mer"))
                  << ", load " << o.load_factor << ", built in "
                  << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count()
                  << " s" << std::endl;
    }

    void releaseGpuIndex() { table_.reset(); }

    // One window's statistics (the per-window call of the reference's
    // CPU and GPU-sync paths).
    Window getStatsFromSequence(const SeqWindow &window)
    {
        initGpuIndex();
        std::vector<SeqWindow> one{window};
        hipibs::ScanOptions so;
        so.batchBases = std::max<uint64_t>(1 << 16, window.end - window.start + 128);
        so.depth = 1;
        return hipibs::scanWindowsStreamed({table_.get()}, one, so)[0][0];
    }

    // ---- the reference's process-wide switches, kept for its callers ----

    // The device path holds a k-mer in one 64-bit word (k <= 32) or two
    // (k <= 64); a larger k would need the reference's string path, which
    // this build does not have: the constructor rejects such databases.
    static bool needsStringPath(uint32_t k) { return k > 64; }

    // Host threads for parsing, packing and KMC reads: FASTIBS_THREADS=N,
    // else every CPU this process may use.
    static unsigned poolThreads()
    {
        if (const char *v = std::getenv("FASTIBS_THREADS"))
        {
            const long n = std::atol(v);
            if (n > 0)
                return (unsigned)n;
        }
        return seqio_detail::hostThreads();
    }

    // One device per process (the --device ordinal).
    static int devicesToUse(int available) { return available > 0 ? 1 : 0; }

    // The device path is the only path: there is no CPU probe and no
    // separate synchronous device path.
    static bool gpuRequested() { return true; }
    static bool gpuSyncRequested() { return false; }

    // One window against several databases (panel mode's per-window call):
    // out[i] = dbs[i]->getStatsFromSequence(window), from one device scan
    // over all the databases' tables (all of one k).
    static void getStatsFromSequenceMulti(const SeqWindow &window,
                                          const std::vector<KmerDatabase *> &dbs, Window *out)
    {
        if (dbs.empty())
            return;
        for (KmerDatabase *d : dbs)
            d->initGpuIndex(dbs.size() > 1 ? TableUse::Lookup : TableUse::Scan);
        std::vector<SeqWindow> one{window};
        hipibs::ScanOptions so;
        so.batchBases = std::max<uint64_t>(1 << 16, window.end - window.start + 128);
        so.depth = 1;
        // one launch per table kind (a table built earlier for a single-sample
        // scan may be a minimizer table; one launch takes one kind)
        std::vector<bool> done(dbs.size(), false);
        for (size_t i = 0; i < dbs.size(); ++i)
        {
            if (done[i])
                continue;
            uint32_t mi = 0;
            hipibs_table_minimizer_len(dbs[i]->table(), &mi);
            std::vector<size_t> idx;
            std::vector<const hipibs_table *> tabs;
            for (size_t j = i; j < dbs.size(); ++j)
            {
                uint32_t mj = 0;
                hipibs_table_minimizer_len(dbs[j]->table(), &mj);
                if (!done[j] && mj == mi)
                {
                    idx.push_back(j);
                    tabs.push_back(dbs[j]->table());
                    done[j] = true;
                }
            }
            const auto res = hipibs::scanWindowsStreamed(tabs, one, so);
            for (size_t t = 0; t < idx.size(); ++t)
                out[idx[t]] = res[t][0];
        }
    }

    // A mapper chunk's hits into its contig's difference array: +1 at every
    // hit position p (contig coordinates, chunk.start + offset) and -1 at
    // p + k, added atomically (chunks of one contig may run concurrently).
    void accumulateMappingChunk(const SeqWindow &chunk, std::vector<int> &diff)
    {
        const std::vector<uint32_t> words = hitWords(chunk.seq);
        const size_t npos = chunk.seq.size() >= kmerSize ? chunk.seq.size() - kmerSize + 1 : 0;
        for (size_t i = 0; i < npos; ++i)
            if ((words[i >> 5] >> (i & 31)) & 1u)
            {
                const size_t p = chunk.start + i;
                __atomic_fetch_add(&diff[p], 1, __ATOMIC_RELAXED);
                __atomic_fetch_sub(&diff[p + kmerSize], 1, __ATOMIC_RELAXED);
            }
    }

    // The reference's coverage line from a difference array: prefix sums
    // joined by ',' (to_chars), no trailing separator.
    static std::string serializeCoverage(const std::vector<int> &diff, size_t seqSize)
    {
        std::string out;
        out.reserve(2 * seqSize);
        char buf[12];
        int coverage = 0;
        for (size_t i = 0; i < seqSize; ++i)
        {
            coverage += diff[i];
            const auto r = std::to_chars(buf, buf + sizeof(buf), coverage);
            out.append(buf, r.ptr);
            out.push_back(',');
        }
        if (!out.empty())
            out.pop_back();
        return out;
    }

    // One contig's (id, coverage line), hit mask, coverage and text made on
    // the device.
    std::pair<std::string, std::string> getMappingFromSequence(const std::string &id,
                                                               const std::string &sequence)
    {
        initGpuIndex();
        const int dev = gpuOptions().device;
        const uint64_t n = sequence.size();
        if (n == 0)
            return {id, std::string()};
        const uint64_t npos = n >= kmerSize ? n - kmerSize + 1 : 0;
        const std::vector<uint32_t> words = hitWords(sequence);
        hipibs::DeviceBuffer dhit(dev, std::max<uint64_t>(1, words.size()) * 4), dcov(dev, n),
            dtxt(dev, 4 * n);
        if (!words.empty())
            hipibs::check(hipibs_memcpy(dhit.get(), words.data(), words.size() * 4, 1), "upload");
        hipibs::check(hipibs_coverage_device(dhit.as<uint32_t>(), npos, n, kmerSize,
                                             dcov.as<uint8_t>(), nullptr),
                      "coverage");
        uint64_t len = 0;
        hipibs::check(hipibs_coverage_text_device(dcov.as<uint8_t>(), n, dtxt.as<char>(), 4 * n,
                                                  &len, nullptr),
                      "coverage text");
        std::string text(len, '\0');
        hipibs::check(hipibs_memcpy(text.data(), dtxt.get(), len, 2), "download");
        return {id, std::move(text)};
    }

    struct ReferenceWindows
    {
        ParseResult parsed;
        std::vector<SeqWindow> chunks;
        double readSec = 0, windowSec = 0;
    };

    void readReferenceWindows(const std::string &refPath, int windowSize, ReferenceWindows &out)
    {
        const auto t0 = std::chrono::steady_clock::now();
        std::cout << "Reading sequences" << std::endl;
        out.parsed = readSequenceFile(refPath, refPath.find(".gz") != std::string::npos);
        const auto t1 = std::chrono::steady_clock::now();
        out.readSec = std::chrono::duration<double>(t1 - t0).count();
        chunkSequences(out.parsed.ids, out.parsed.sequences, (size_t)windowSize, kmerSize,
                       out.chunks);
        out.windowSec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        std::cout << "Number of chunks: " << out.chunks.size() << std::endl;
    }

    static hipibs::ScanOptions scanOptions()
    {
        hipibs::ScanOptions so;
        so.batchBases = gpuOptions().batchBases;
        so.depth = gpuOptions().depth;
        return so;
    }

    // "<out>.partial", then rename: a failed write leaves no file under
    // the final name (src/KmerDatabase.hpp:899-931).
    template <typename WriteFn>
    static void writeOutputFile(const std::string &outPath, WriteFn write)
    {
        const std::string tmp = outPath + ".partial";
        {
            errno = 0;
            std::ofstream out(tmp, std::ios::binary);
            if (!out.is_open())
                throw std::runtime_error("Unable to open " + tmp + " for writing" +
                                         (errno ? std::string(": ") + std::strerror(errno) : ""));
            write(out);
            out.close();
            if (!out)
            {
                std::error_code ec;
                std::filesystem::remove(tmp, ec);
                throw std::runtime_error("Writing " + outPath + " failed (disk full or I/O error)");
            }
        }
        std::error_code ec;
        std::filesystem::rename(tmp, outPath, ec);
        if (ec)
        {
            std::error_code ig;
            std::filesystem::remove(tmp, ig);
            throw std::runtime_error("Unable to move " + tmp + " to " + outPath + ": " +
                                     ec.message());
        }
    }

    // The TSV (src/KmerDatabase.hpp:935-948 format).
    static void writeStats(const std::string &outPath, const std::vector<Window> &rows)
    {
        writeOutputFile(outPath, [&rows](std::ofstream &f) {
            std::string buf = "seqname\tstart\tend\ttotal_kmers\tobserved_kmers\tvariations\t"
                              "kmer_distance\n";
            char num[16];
            auto put = [&](long long v, char sep) {
                auto r = std::to_chars(num, num + sizeof(num), v);
                buf.append(num, r.ptr);
                buf.push_back(sep);
            };
            for (const Window &w : rows)
            {
                buf += w.id;
                buf.push_back('\t');
                put(w.start, '\t');
                put(w.end, '\t');
                put(w.totalKmers, '\t');
                put(w.observedKmers, '\t');
                put(w.variations, '\t');
                put(w.kmerDistance, '\n');
                if (buf.size() > (1u << 22))
                {
                    f.write(buf.data(), (std::streamsize)buf.size());
                    buf.clear();
                }
            }
            f.write(buf.data(), (std::streamsize)buf.size());
        });
    }

    void processReference(std::string refPath, std::string outPath, int windowSize = 50000)
    {
        using clk = std::chrono::steady_clock;
        auto t = clk::now();
        auto lap = [&t]() {
            const auto n = clk::now();
            const double s = std::chrono::duration<double>(n - t).count();
            t = n;
            return s;
        };
        // the reference is read and cut into windows on its own thread while
        // the table is built
        ReferenceWindows rw;
        auto refRead = std::async(std::launch::async,
                                  [&] { readReferenceWindows(refPath, windowSize, rw); });
        try
        {
            initGpuIndex();
        }
        catch (...)
        {
            refRead.wait();
            throw;
        }
        const double indexSec = lap();
        refRead.get();
        lap();
        std::cout << "Kmer size: " << kmerSize << "\nK-mer path: fused device encode+probe+"
                  << "window reduction (MI355X)" << std::endl;
        hipibs::ScanTimes st;
        auto res = hipibs::scanWindowsStreamed({table_.get()}, rw.chunks, scanOptions(), &st);
        const double pipeSec = lap();
        writeStats(outPath, res[0]);
        const double writeSec = lap();
        std::cout << "Scan phases: index " << indexSec << " s, read " << rw.readSec
                  << " s, windows " << rw.windowSec << " s, pipeline " << pipeSec << " s ("
                  << st.batches << " batches, pack " << st.packSec << " s, device wait "
                  << st.waitSec << " s), write " << writeSec << " s" << std::endl;
    }

    // Per-base coverage of every contig (produceMappingGpu,
    // src/KmerDatabase.hpp:1544-1581): the contig is packed on host threads,
    // its hit mask, coverage and coverage text are made on the device, and
    // the text streams from pinned buffers into the output file.
    void produceMapping(std::string refPath, std::string outPath)
    {
        using clk = std::chrono::steady_clock;
        std::cout << "Reading sequences" << std::endl;
        auto refRead = std::async(std::launch::async, [&] {
            return readSequenceFile(refPath, refPath.find(".gz") != std::string::npos);
        });
        try
        {
            initGpuIndex();
        }
        catch (...)
        {
            refRead.wait();
            throw;
        }
        std::cout << "Kmer size: " << kmerSize << "\nK-mer path: device hit mask + coverage"
                  << std::endl;
        ParseResult r = refRead.get();
        const auto t0 = clk::now();
        const int dev = gpuOptions().device;
        const unsigned threads = std::min(64u, seqio_detail::hostThreads());
        std::vector<uint64_t> sw, ew;
        double packSec = 0, devSec = 0, textSec = 0;
        std::cout << "Writing mapping to file" << std::endl;
        writeOutputFile(outPath, [&](std::ofstream &f) {
            for (size_t c = 0; c < r.sequences.size(); ++c)
            {
                const std::string &s = r.sequences[c];
                const uint64_t n = s.size();
                const auto tp = clk::now();
                const uint64_t nsw = (n + 31) / 32 + HIPIBS_PACK_PAD_WORDS;
                const uint64_t new_ = (n + 63) / 64 + HIPIBS_PACK_PAD_WORDS;
                sw.assign(nsw, 0);
                ew.assign(new_, 0);
                // 64-base-aligned pieces write disjoint words
                const uint64_t piece = 1ull << 22;
                seqio_detail::parallelFor((n + piece - 1) / piece, threads, [&](size_t i) {
                    const uint64_t lo = i * piece, hi = std::min<uint64_t>(n, lo + piece);
                    hipibs::check(hipibs_pack_bases(s.data() + lo, hi - lo, lo, sw.data(),
                                                    ew.data()),
                                  "pack");
                });
                const auto td = clk::now();
                const uint64_t npos = n >= kmerSize ? n - kmerSize + 1 : 0;
                hipibs::DeviceBuffer dcov(dev, n);
                {
                    hipibs::DeviceBuffer dsw(dev, nsw * 8), dew(dev, new_ * 8),
                        dhit(dev, ((npos + 31) / 32 + 1) * 4);
                    hipibs::check(hipibs_memcpy(dsw.get(), sw.data(), nsw * 8, 1), "upload");
                    hipibs::check(hipibs_memcpy(dew.get(), ew.data(), new_ * 8, 1), "upload");
                    if (npos)
                        hipibs::check(hipibs_hitmask_device(table_.get(), dsw.as<uint64_t>(),
                                                            dew.as<uint64_t>(), 0, npos, n,
                                                            dhit.as<uint32_t>(), nullptr),
                                      "hit mask");
                    hipibs::check(hipibs_coverage_device(dhit.as<uint32_t>(), npos, n, kmerSize,
                                                         dcov.as<uint8_t>(), nullptr),
                                  "coverage");
                    hipibs::check(hipibs_dev_synchronize(dev), "coverage");
                }
                const auto tt = clk::now();
                f << r.ids[c] << '\n';
                hipibs::check(hipibs_coverage_text_stream(
                                  dcov.as<uint8_t>(), n,
                                  [](void *u, const char *d, uint64_t m) -> int {
                                      auto &o = *static_cast<std::ofstream *>(u);
                                      o.write(d, (std::streamsize)m);
                                      return o ? 0 : 1;
                                  },
                                  &f, nullptr),
                              "coverage text");
                f << '\n';
                packSec += std::chrono::duration<double>(td - tp).count();
                devSec += std::chrono::duration<double>(tt - td).count();
                textSec += std::chrono::duration<double>(clk::now() - tt).count();
            }
        });
        std::cout << "Mapping phases: pack " << packSec << " s, device " << devSec
                  << " s, text + write " << textSec << " s (total "
                  << std::chrono::duration<double>(clk::now() - t0).count() << " s)" << std::endl;
    }

    uint32_t kmerSize = 0;

private:
    // Hit mask of every position of seq (bit i of word i/32: the k-mer at
    // i is valid and in the table), from the device.
    std::vector<uint32_t> hitWords(std::string_view seq)
    {
        initGpuIndex();
        const int dev = gpuOptions().device;
        const uint64_t n = seq.size();
        const uint64_t npos = n >= kmerSize ? n - kmerSize + 1 : 0;
        std::vector<uint32_t> words((npos + 31) / 32, 0);
        if (npos == 0)
            return words;
        const uint64_t nsw = (n + 31) / 32 + HIPIBS_PACK_PAD_WORDS;
        const uint64_t new_ = (n + 63) / 64 + HIPIBS_PACK_PAD_WORDS;
        std::vector<uint64_t> sw(nsw, 0), ew(new_, 0);
        hipibs::check(hipibs_pack_bases(seq.data(), n, 0, sw.data(), ew.data()), "pack");
        hipibs::DeviceBuffer dsw(dev, nsw * 8), dew(dev, new_ * 8), dhit(dev, (words.size() + 1) * 4);
        hipibs::check(hipibs_memcpy(dsw.get(), sw.data(), nsw * 8, 1), "upload");
        hipibs::check(hipibs_memcpy(dew.get(), ew.data(), new_ * 8, 1), "upload");
        hipibs::check(hipibs_hitmask_device(table_.get(), dsw.as<uint64_t>(), dew.as<uint64_t>(), 0,
                                            npos, n, dhit.as<uint32_t>(), nullptr),
                      "hit mask");
        hipibs::check(hipibs_dev_synchronize(dev), "hit mask");
        hipibs::check(hipibs_memcpy(words.data(), dhit.get(), words.size() * 4, 2), "download");
        return words;
    }

    std::string sourcePath_;
    hipibs_kmc_info info_{};
    hipibs::TablePtr table_;
};
