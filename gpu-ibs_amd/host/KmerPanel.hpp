// KmerPanel: N samples against a reference in one reference pass (the
// reference's panel mode, src/KmerPanel.hpp:24-465). Samples are grouped so
// that a group's tables fit the device memory budget (largest first within
// the caller's order); each group is one multi-table scan: one encode of
// every position, one probe per table, statistics per table -- the outputs
// of a group are byte-identical to single-sample runs.
#pragma once

#include "KmerDatabase.hpp"

#include <algorithm>
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

    // Bytes the device table of database i will take.
    uint64_t tableBytes(size_t i) const
    {
        const auto &g = KmerDatabase::gpuOptions();
        const double slots = std::ceil((double)std::max<uint64_t>(1, dbs_[i]->totalKmers()) /
                                       g.loadFactor);
        const uint64_t buckets = std::max<uint64_t>(2, (uint64_t)std::ceil(slots / g.bucketSlots));
        return buckets * g.bucketSlots * 8;
    }

    // Groups of database indices, in order, each within the memory budget.
    std::vector<std::vector<size_t>> planRounds() const
    {
        const auto &g = KmerDatabase::gpuOptions();
        uint64_t freeB = 0, totalB = 0;
        hipibs::check(hipibs_device_mem(g.device, &freeB, &totalB), "device memory query");
        const uint64_t reserve = g.batchBases / 2 * (uint64_t)g.depth + (4ull << 30);
        uint64_t budget = freeB > reserve ? freeB - reserve : 0;
        if (g.hbmBudget)
            budget = std::min(budget, g.hbmBudget);
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

    // outPaths[i] is database i's TSV; an empty path skips the database.
    void processReference(const std::string &refPath, const std::vector<std::string> &outPaths,
                          int windowSize)
    {
        std::string why;
        if (hipibs::deviceCount(why) < 1)
            throw std::runtime_error("no usable GPU: " + why);
        KmerDatabase::ReferenceWindows rw;
        dbs_[0]->readReferenceWindows(refPath, windowSize, rw);
        const auto rounds = planRounds();
        std::cout << "Panel: " << dbs_.size() << " databases in " << rounds.size()
                  << " device round(s)" << std::endl;
        for (const auto &round : rounds)
        {
            std::vector<size_t> todo;
            for (size_t i : round)
                if (!outPaths[i].empty())
                    todo.push_back(i);
            if (todo.empty())
                continue;
            std::vector<const hipibs_table *> tabs;
            for (size_t i : todo)
            {
                dbs_[i]->initGpuIndex();
                tabs.push_back(dbs_[i]->table());
            }
            auto res = hipibs::scanWindowsStreamed(tabs, rw.chunks, KmerDatabase::scanOptions());
            for (size_t j = 0; j < todo.size(); ++j)
            {
                KmerDatabase::writeStats(outPaths[todo[j]], res[j]);
                dbs_[todo[j]]->releaseGpuIndex();
            }
        }
    }

private:
    std::vector<std::unique_ptr<KmerDatabase>> dbs_;
    int resident_;
};
