// KmerPanel: N samples against a reference in one reference pass (the
// reference's panel mode, src/KmerPanel.hpp:24-465). Samples are grouped so
// that a group's tables fit the device memory budget (largest first within
// the caller's order); each group is one multi-table scan: one encode of
// every position, one probe per table, statistics per table -- the outputs
// of a group are byte-identical to single-sample runs.
#pragma once

#include "KmerDatabase.hpp"

#include <algorithm>
#include <atomic>
#include <future>
#include <cmath>
#include <memory>

class KmerPanel
{
public:
    KmerPanel(std::vector<std::string> sourcePaths, int resident = 0)
        : resident_(resident)
    {
        for (auto &p : sourcePaths)
            dbs_.push_back(std::make_unique<KmerDatabase>(p));
        for (auto &d : dbs_)
            if (d->getKmerSize() != dbs_[0]->getKmerSize())
                throw std::runtime_error("panel databases must share k");
    }

    uint32_t getKmerSize() const { return dbs_[0]->getKmerSize(); }
    size_t size() const { return dbs_.size(); }

    // Bytes the device table of database i will take (the sizing rules of
    // the table builds: whole-k-mer buckets of 16 keys, or 8 two-word keys
    // for k > 32; a minimizer table's two parts hold at most n keys plus a
    // marker per group of more than cap keys, each part rounded up to whole
    // buckets).
    uint64_t tableBytes(size_t i) const
    {
        const uint32_t k = dbs_[i]->getKmerSize();
        const uint64_t n = std::max<uint64_t>(1, dbs_[i]->totalKmers());
        const hipibs_table_opts o =
            KmerDatabase::tableOptions(k, KmerDatabase::TableUse::Lookup);
        const uint32_t slots = o.bucket_slots ? o.bucket_slots : 16;
        if (o.minimizer_len)
        {
            const double items = (double)n + std::ceil((double)n / (o.minimizer_cap + 1.0));
            return (uint64_t)std::ceil(std::ceil(items / o.load_factor) * 8.0) +
                   3ull * slots * 8 + 3ull * 64;
        }
        const uint32_t perBucket = k > 32 ? 8 : slots;
        const uint64_t buckets = std::max<uint64_t>(
            2, (uint64_t)std::ceil(std::ceil((double)n / o.load_factor) / perBucket));
        return buckets * slots * 8;
    }

    // Device bytes a build holds beside the tables already resident: a
    // minimizer table's build holds every key (8 B) and its group map (a
    // power of two >= n entries of 8 B); a whole-k-mer build from KMC
    // records streams them through small fixed staging (in the reserve).
    uint64_t buildTransientBytes() const
    {
        const hipibs_table_opts o =
            KmerDatabase::tableOptions(getKmerSize(), KmerDatabase::TableUse::Lookup);
        if (!o.minimizer_len)
            return 0;
        uint64_t most = 0;
        for (auto &d : dbs_)
        {
            const uint64_t n = d->totalKmers();
            uint64_t map = 1024;
            while (map < n)
                map <<= 1;
            most = std::max<uint64_t>(most, n * 8 + map * 8);
        }
        return most;
    }

    // Device bytes the panel's tables may take.
    uint64_t tableBudget() const
    {
        const auto &g = KmerDatabase::gpuOptions();
        uint64_t freeB = 0, totalB = 0;
        hipibs::check(hipibs_device_mem(g.device, &freeB, &totalB), "device memory query");
        const uint64_t reserve = reserveBytes_ + (4ull << 30) + buildTransientBytes();
        uint64_t budget = freeB > reserve ? freeB - reserve : 0;
        if (g.hbmBudget)
            budget = std::min(budget, g.hbmBudget);
        return budget;
    }

    // Groups of database indices, in order, each within `budget` bytes
    // (0 = the whole table budget).
    std::vector<std::vector<size_t>> planRounds(uint64_t budget = 0) const
    {
        if (budget == 0)
            budget = tableBudget();
        std::vector<std::vector<size_t>> rounds;
        std::vector<size_t> cur;
        uint64_t used = 0;
        for (size_t i = 0; i < dbs_.size(); ++i)
        {
            const uint64_t b = tableBytes(i);
            if (b > budget)
                throw std::runtime_error("database " + dbs_[i]->sourcePath() +
                                         " alone exceeds the device memory budget");
            if (!cur.empty() && (used + b > budget || (resident_ > 0 && (int)cur.size() >= resident_)))
            {
                rounds.push_back(cur);
                cur.clear();
                used = 0;
            }
            cur.push_back(i);
            used += b;
        }
        if (!cur.empty())
            rounds.push_back(cur);
        return rounds;
    }

    // The whole reference packed once into device memory: contigs
    // concatenated, every window a (base, length) range of it.
    struct DeviceReference
    {
        hipibs::DeviceBuffer seq, exc, windows;
        uint64_t nBases = 0, bytes = 0;
        uint32_t nWindows = 0;
    };

    static void packReference(const KmerDatabase::ReferenceWindows &rw, int device,
                              DeviceReference &out)
    {
        const auto &seqs = rw.parsed.sequences;
        std::vector<uint64_t> off(seqs.size() + 1, 0);
        for (size_t c = 0; c < seqs.size(); ++c)
            off[c + 1] = off[c] + (seqs[c].size() + 63) / 64 * 64; // contigs 64-aligned
        const uint64_t n = off.back();
        const uint64_t nsw = n / 32 + HIPIBS_PACK_PAD_WORDS, new_ = n / 64 + HIPIBS_PACK_PAD_WORDS;
        std::vector<uint64_t> sw(nsw, 0), ew(new_, 0);
        // pieces of <= 4 Mbases, 64-aligned in the stream, packed on threads
        struct Piece
        {
            const char *text;
            uint64_t n, at;
        };
        std::vector<Piece> pieces;
        for (size_t c = 0; c < seqs.size(); ++c)
            for (uint64_t o = 0; o < seqs[c].size(); o += (4ull << 20))
                pieces.push_back(Piece{seqs[c].data() + o,
                                       std::min<uint64_t>(4ull << 20, seqs[c].size() - o), off[c] + o});
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t i; (i = next.fetch_add(1)) < pieces.size();)
                hipibs_pack_bases(pieces[i].text, pieces[i].n, pieces[i].at, sw.data(), ew.data());
        };
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < std::max(1u, std::thread::hardware_concurrency()); ++t)
            pool.emplace_back(work);
        for (auto &th : pool)
            th.join();
        std::vector<hipibs_window> win(rw.chunks.size());
        std::vector<const std::string *> contigId(seqs.size());
        for (size_t i = 0; i < rw.chunks.size(); ++i)
        {
            const SeqWindow &w = rw.chunks[i];
            const size_t c = (size_t)(w.id - rw.parsed.ids.data());
            win[i] = hipibs_window{off[c] + w.start, (uint32_t)(w.end - w.start), 0};
        }
        out.seq = hipibs::DeviceBuffer(device, nsw * 8);
        out.exc = hipibs::DeviceBuffer(device, new_ * 8);
        out.windows = hipibs::DeviceBuffer(device, std::max<size_t>(1, win.size()) * sizeof(hipibs_window));
        hipibs::check(hipibs_memcpy(out.seq.get(), sw.data(), nsw * 8, 1), "reference upload");
        hipibs::check(hipibs_memcpy(out.exc.get(), ew.data(), new_ * 8, 1), "reference upload");
        hipibs::check(hipibs_memcpy(out.windows.get(), win.data(), win.size() * sizeof(hipibs_window), 1),
                      "window upload");
        out.nBases = n;
        out.nWindows = (uint32_t)win.size();
        out.bytes = nsw * 8 + new_ * 8 + win.size() * sizeof(hipibs_window);
    }

    // Bucket width: the smallest database's table is built with 64 B and
    // with 128 B buckets and scanned over the first 1024 windows; the faster
    // width builds every table of the panel.
    void tuneBucketWidth(const DeviceReference &dref)
    {
        using clk = std::chrono::steady_clock;
        auto &g = KmerDatabase::gpuOptions();
        size_t pick = 0;
        for (size_t i = 1; i < dbs_.size(); ++i)
            if (dbs_[i]->kmerCount() < dbs_[pick]->kmerCount())
                pick = i;
        const uint32_t nw = std::min<uint32_t>(dref.nWindows, 1024);
        const uint64_t wb = hipibs_panel_work_bytes(nw, 1, 1);
        hipibs::DeviceBuffer stats(g.device, (uint64_t)nw * sizeof(hipibs_window_stats));
        hipibs::DeviceBuffer work(g.device, std::max<uint64_t>(wb, 8));
        const hipibs_scan_config c{1, 16, {0, 0}};
        double best = 1e30;
        uint32_t bestSlots = g.bucketSlots;
        for (uint32_t slots : {8u, 16u})
        {
            g.bucketSlots = slots;
            dbs_[pick]->initGpuIndex(KmerDatabase::TableUse::Lookup);
            const hipibs_table *t = dbs_[pick]->table();
            double sec = 0;
            for (int rep = 0; rep < 3; ++rep)
            {
                const auto t0 = clk::now();
                hipibs::check(hipibs_panel_scan_device(&t, 1, dref.seq.as<uint64_t>(),
                                                       dref.exc.as<uint64_t>(), dref.nBases,
                                                       dref.windows.as<hipibs_window>(), nw,
                                                       stats.as<hipibs_window_stats>(), &c,
                                                       work.get(), wb, nullptr),
                              "bucket-width trial");
                hipibs::check(hipibs_dev_synchronize(g.device), "bucket-width trial");
                if (rep > 0)
                    sec += std::chrono::duration<double>(clk::now() - t0).count();
            }
            dbs_[pick]->releaseGpuIndex();
            if (sec < best)
            {
                best = sec;
                bestSlots = slots;
            }
        }
        g.bucketSlots = bestSlots;
        std::cout << "Autotuned bucket width: " << bestSlots * 8 << " B" << std::endl;
    }

    // outPaths[i] is database i's TSV; an empty path skips the database.
    // Device-resident path: the reference is packed and uploaded once;
    // each memory-budgeted round builds its tables and replays one captured
    // graph of chunked panel-kernel steps (launch knobs autotuned once).
    void processReference(const std::string &refPath, const std::vector<std::string> &outPaths,
                          int windowSize)
    {
        using clk = std::chrono::steady_clock;
        std::string why;
        if (hipibs::deviceCount(why) < 1)
            throw std::runtime_error("no usable GPU: " + why);
        const auto t0 = clk::now();
        KmerDatabase::ReferenceWindows rw;
        dbs_[0]->readReferenceWindows(refPath, windowSize, rw);
        const auto &g = KmerDatabase::gpuOptions();
        DeviceReference dref;
        packReference(rw, g.device, dref);
        const double ingestSec = std::chrono::duration<double>(clk::now() - t0).count();
        reserveBytes_ = dref.bytes;
        // More than one memory group: plan them at half the budget, so the
        // next group's tables can be built while the current group scans.
        auto rounds = planRounds();
        const bool overlap = rounds.size() > 1;
        if (overlap)
            rounds = planRounds(tableBudget() / 2);
        std::cout << "Panel: " << dbs_.size() << " databases in " << rounds.size()
                  << " device round(s); reference " << dref.nBases << " bases, "
                  << dref.nWindows << " windows, ingest " << ingestSec << " s" << std::endl;
        // k > 32: 128 B buckets only, and the tables of a group are scanned
        // by the multi-table scan launch (the panel kernel holds 64-bit codes)
        const bool wide = dbs_[0]->getKmerSize() > 32;
        if (g.bucketAuto && dref.nWindows && !wide)
            tuneBucketWidth(dref);
        hipibs_scan_config cfg{0, 0, {0, 0}};
        bool tuned = false;
        double buildSec = 0, waitSec = 0, scanSec = 0, writeSec = 0;
        std::vector<std::vector<size_t>> todos;
        for (const auto &round : rounds)
        {
            std::vector<size_t> todo;
            for (size_t i : round)
                if (!outPaths[i].empty())
                    todo.push_back(i);
            if (!todo.empty())
                todos.push_back(todo);
        }
        // builds of group r+1 run on their own thread during the scan of r
        auto buildGroup = [this](const std::vector<size_t> &todo) {
            const auto tb = clk::now();
            for (size_t i : todo)
                dbs_[i]->initGpuIndex(KmerDatabase::TableUse::Lookup);
            return std::chrono::duration<double>(clk::now() - tb).count();
        };
        std::future<double> nextBuild;
        if (!todos.empty())
            nextBuild = std::async(std::launch::async, buildGroup, todos[0]);
        for (size_t r = 0; r < todos.size(); ++r)
        {
            const std::vector<size_t> &todo = todos[r];
            auto tb = clk::now();
            buildSec += nextBuild.get(); // builder time (overlapped when grouped)
            if (r + 1 < todos.size())
                nextBuild = std::async(std::launch::async, buildGroup, todos[r + 1]);
            std::vector<const hipibs_table *> tabs;
            for (size_t i : todo)
                tabs.push_back(dbs_[i]->table());
            auto ts = clk::now();
            waitSec += std::chrono::duration<double>(ts - tb).count();
            const int nt = (int)tabs.size();
            hipibs::DeviceBuffer stats(g.device, std::max<uint64_t>(1, (uint64_t)dref.nWindows * nt) *
                                                     sizeof(hipibs_window_stats));
            uint32_t segs = 0;
            const uint64_t wb =
                wide ? std::max<uint64_t>(8, hipibs_scan_work_bytes(dref.nWindows, nt,
                                                                     (uint32_t)windowSize,
                                                                     dbs_[0]->getKmerSize(), 0,
                                                                     &segs))
                     : hipibs_panel_work_bytes(dref.nWindows, nt, 4);
            hipibs::DeviceBuffer work(g.device, wb);
            if (!tuned && dref.nWindows && !wide)
            {
                double ms = 0;
                hipibs::check(hipibs_panel_autotune(tabs.data(), nt, dref.seq.as<uint64_t>(),
                                                    dref.exc.as<uint64_t>(), dref.nBases,
                                                    dref.windows.as<hipibs_window>(), dref.nWindows,
                                                    1024, stats.as<hipibs_window_stats>(),
                                                    work.get(), wb, &cfg, &ms),
                              "panel autotune");
                std::cout << "Autotuned panel launch: " << cfg.waves_per_cu << " waves/CU, "
                          << cfg.segs_per_window << " segment(s)/window" << std::endl;
                tuned = true;
            }
            std::vector<hipibs_window_stats> host((uint64_t)dref.nWindows * nt);
            if (dref.nWindows && wide)
            {
                hipibs::check(hipibs_scan_windows_device(
                                  tabs.data(), nt, dref.seq.as<uint64_t>(), dref.exc.as<uint64_t>(),
                                  dref.nBases, dref.windows.as<hipibs_window>(), dref.nWindows,
                                  (uint32_t)windowSize, stats.as<hipibs_window_stats>(), segs,
                                  work.get(), wb, nullptr),
                              "scan of the group");
                hipibs::check(hipibs_dev_synchronize(g.device), "scan of the group");
                hipibs::check(hipibs_memcpy(host.data(), stats.get(),
                                            host.size() * sizeof(hipibs_window_stats), 2),
                              "statistics download");
            }
            else if (dref.nWindows)
            {
                hipibs_graph *graph = nullptr;
                hipibs::check(hipibs_panel_graph_create(tabs.data(), nt, dref.seq.as<uint64_t>(),
                                                        dref.exc.as<uint64_t>(), dref.nBases,
                                                        dref.windows.as<hipibs_window>(),
                                                        dref.nWindows,
                                                        stats.as<hipibs_window_stats>(), &cfg, 8,
                                                        work.get(), wb, &graph),
                              "panel graph");
                const int rc = hipibs_graph_launch(graph, nullptr);
                const int rc2 = rc ? rc : hipibs_dev_synchronize(g.device);
                hipibs_graph_free(graph);
                hipibs::check(rc2, "panel graph launch");
                hipibs::check(hipibs_memcpy(host.data(), stats.get(),
                                            host.size() * sizeof(hipibs_window_stats), 2),
                              "statistics download");
            }
            auto tw = clk::now();
            scanSec += std::chrono::duration<double>(tw - ts).count();
            for (size_t j = 0; j < todo.size(); ++j)
            {
                std::vector<Window> rows(rw.chunks.size());
                for (size_t i = 0; i < rw.chunks.size(); ++i)
                {
                    const SeqWindow &w = rw.chunks[i];
                    const hipibs_window_stats &x = host[(uint64_t)j * dref.nWindows + i];
                    rows[i].id = *w.id;
                    rows[i].start = (int)w.start;
                    rows[i].end = (int)w.end;
                    rows[i].totalKmers = x.total_kmers;
                    rows[i].observedKmers = x.observed_kmers;
                    rows[i].variations = x.variations;
                    rows[i].kmerDistance = x.kmer_distance;
                }
                KmerDatabase::writeStats(outPaths[todo[j]], rows);
                dbs_[todo[j]]->releaseGpuIndex();
            }
            writeSec += std::chrono::duration<double>(clk::now() - tw).count();
        }
        std::cout << "Panel phases: ingest " << ingestSec << " s, table builds " << buildSec
                  << " s (waited " << waitSec << " s" << (overlap ? ", overlapped with scans" : "")
                  << "), scans " << scanSec << " s, write " << writeSec << " s" << std::endl;
    }

private:
    std::vector<std::unique_ptr<KmerDatabase>> dbs_;
    int resident_;
    uint64_t reserveBytes_ = 0; // device bytes held outside the tables
};
