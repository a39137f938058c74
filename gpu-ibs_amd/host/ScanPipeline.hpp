// The streamed scan of a window list against one or more device tables:
// windows are grouped into batches that never split a window (as
// buildBatches, src/gpu/EncodeProbeCore.hpp:103-134); the windows of a
// batch that come from the same contig share one packed region (they
// overlap by k bases), packed on host threads into the slot's pinned
// buffers while earlier slots run on the device (the ring of
// KmerDatabase::runEngineLoop, src/KmerDatabase.hpp:1260-1494, with the
// per-window statistics computed on the device instead of the host).
#pragma once

#include "Hipibs.hpp"
#include "SeqIO.hpp"
#include "Window.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

namespace hipibs
{

struct ScanOptions
{
    uint64_t batchBases = 128ull << 20; // bases per slot (0.375 B each on the device)
    int depth = 3;                      // slots in flight
    unsigned packThreads = 0;           // 0 = hardware concurrency
};

struct ScanTimes
{
    double packSec = 0, waitSec = 0, totalSec = 0;
    size_t batches = 0;
};

namespace detail
{
struct Region
{
    const char *text; // first base of the region
    uint64_t n;       // bases
    uint64_t at;      // slot base offset (64-aligned)
};
struct Batch
{
    size_t lo = 0, hi = 0; // window index range
    std::vector<Region> regions;
    std::vector<uint64_t> winBase; // slot base of each window
    uint64_t bases = 0;            // slot bases used (64-aligned)
};

inline std::vector<Batch> planBatches(const std::vector<SeqWindow> &w, uint64_t cap, uint32_t maxWin)
{
    std::vector<Batch> out;
    Batch cur;
    const char *regBase = nullptr; // contig base pointer of the open region
    uint64_t regStart = 0, regEnd = 0;
    auto closeRegion = [&]() {
        if (!regBase)
            return;
        const uint64_t n = regEnd - regStart;
        cur.regions.push_back(Region{regBase + regStart, n, cur.bases});
        cur.bases += (n + 63) / 64 * 64;
        regBase = nullptr;
    };
    auto closeBatch = [&](size_t i) {
        closeRegion();
        cur.hi = i;
        out.push_back(std::move(cur));
        cur = Batch{};
        cur.lo = i;
    };
    auto round64 = [](uint64_t x) { return (x + 63) / 64 * 64; };
    for (size_t i = 0; i < w.size(); ++i)
    {
        const char *contig = w[i].seq.data() - w[i].start;
        bool extend = regBase == contig && w[i].start >= regStart;
        const uint64_t projected =
            extend ? cur.bases + round64(std::max<uint64_t>(regEnd, w[i].end) - regStart)
                   : cur.bases + (regBase ? round64(regEnd - regStart) : 0) +
                         round64(w[i].end - w[i].start);
        if ((projected + 64 > cap || i - cur.lo + 1 > maxWin) && i > cur.lo)
        {
            closeBatch(i);
            extend = false;
        }
        if (!extend)
        {
            closeRegion();
            regBase = contig;
            regStart = w[i].start;
            regEnd = w[i].end;
        }
        else
            regEnd = std::max<uint64_t>(regEnd, w[i].end);
        // the open region will sit at slot offset cur.bases
        cur.winBase.push_back(cur.bases + (w[i].start - regStart));
    }
    if (w.size() > cur.lo)
        closeBatch(w.size());
    return out;
}
} // namespace detail

// Statistics of every window against every table: result[t][i].
inline std::vector<std::vector<Window>> scanWindowsStreamed(
    const std::vector<const hipibs_table *> &tables, const std::vector<SeqWindow> &windows,
    const ScanOptions &opt, ScanTimes *times = nullptr)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    std::vector<std::vector<Window>> res(tables.size(), std::vector<Window>(windows.size()));
    for (auto &r : res)
        for (size_t i = 0; i < windows.size(); ++i)
        {
            r[i].id = *windows[i].id;
            r[i].start = (int)windows[i].start;
            r[i].end = (int)windows[i].end;
        }
    if (windows.empty())
        return res;
    uint64_t maxWin = 0;
    for (const auto &w : windows)
        maxWin = std::max<uint64_t>(maxWin, w.end - w.start);
    const uint64_t cap = std::max<uint64_t>(opt.batchBases, maxWin + 128);
    const uint32_t maxWinPerBatch =
        (uint32_t)std::min<uint64_t>(windows.size(), std::max<uint64_t>(4096, cap / 64));
    std::vector<detail::Batch> batches = detail::planBatches(windows, cap, maxWinPerBatch);
    ScanEngine eng(tables, cap, maxWinPerBatch, std::max(1, opt.depth));
    const unsigned T = opt.packThreads ? opt.packThreads
                                       : std::max(1u, std::thread::hardware_concurrency());
    const int depth = eng.depth();
    std::vector<long> inSlot(depth, -1);
    double packSec = 0, waitSec = 0;

    auto collect = [&](int s) {
        const auto tw = clk::now();
        eng.wait(s);
        waitSec += std::chrono::duration<double>(clk::now() - tw).count();
        const detail::Batch &b = batches[(size_t)inSlot[s]];
        const hipibs_window_stats *st = eng.slot(s).stats;
        const size_t n = b.hi - b.lo;
        for (size_t t = 0; t < tables.size(); ++t)
            for (size_t j = 0; j < n; ++j)
            {
                const hipibs_window_stats &x = st[t * n + j];
                Window &w = res[t][b.lo + j];
                w.totalKmers = x.total_kmers;
                w.observedKmers = x.observed_kmers;
                w.variations = x.variations;
                w.kmerDistance = x.kmer_distance;
            }
        inSlot[s] = -1;
    };

    for (size_t bi = 0; bi < batches.size(); ++bi)
    {
        const int s = (int)(bi % depth);
        if (inSlot[s] >= 0)
            collect(s);
        const auto tp = clk::now();
        const detail::Batch &b = batches[bi];
        const hipibs_host_slot &sl = eng.slot(s);
        // pack: regions split into 64-aligned pieces of up to 4 Mbases
        struct Piece
        {
            const char *text;
            uint64_t n, at;
        };
        std::vector<Piece> pieces;
        for (const auto &r : b.regions)
            for (uint64_t o = 0; o < r.n; o += (4ull << 20))
                pieces.push_back(Piece{r.text + o, std::min<uint64_t>(4ull << 20, r.n - o), r.at + o});
        std::atomic<size_t> next{0};
        auto packer = [&]() {
            for (;;)
            {
                const size_t p = next.fetch_add(1);
                if (p >= pieces.size())
                    break;
                const Piece &q = pieces[p];
                // the piece's words start zeroed (the pinned buffer is reused)
                std::memset(sl.seq_words + q.at / 32, 0, ((q.n + 63) / 64) * 16);
                std::memset(sl.exc_words + q.at / 64, 0, ((q.n + 63) / 64) * 8);
                hipibs_pack_bases(q.text, q.n, q.at, sl.seq_words, sl.exc_words);
            }
        };
        const unsigned nt = (unsigned)std::min<size_t>(T, pieces.size());
        if (nt <= 1)
            packer();
        else
        {
            std::vector<std::thread> pool;
            for (unsigned t = 0; t < nt; ++t)
                pool.emplace_back(packer);
            for (auto &th : pool)
                th.join();
        }
        const size_t n = b.hi - b.lo;
        for (size_t j = 0; j < n; ++j)
        {
            const SeqWindow &w = windows[b.lo + j];
            sl.windows[j] = hipibs_window{b.winBase[j], (uint32_t)(w.end - w.start), 0};
        }
        // pad words after the last region: zeroed so they are defined
        std::memset(sl.seq_words + b.bases / 32, 0, HIPIBS_PACK_PAD_WORDS * 8);
        std::memset(sl.exc_words + b.bases / 64, 0, HIPIBS_PACK_PAD_WORDS * 8);
        packSec += std::chrono::duration<double>(clk::now() - tp).count();
        eng.submit(s, (uint32_t)n, b.bases / 32, b.bases / 64);
        inSlot[s] = (long)bi;
    }
    for (int s = 0; s < depth; ++s)
        if (inSlot[s] >= 0)
            collect(s);
    if (times)
    {
        times->packSec = packSec;
        times->waitSec = waitSec;
        times->batches = batches.size();
        times->totalSec = std::chrono::duration<double>(clk::now() - t0).count();
    }
    return res;
}

} // namespace hipibs
