// Device-backed KmerDatabase members that the tools do not call, for
// tests/test_tools.py to check against the golden fixtures:
//   test_api <fasta> <out dir> <W> <db prefix>...
// writes
//   <out>/mapping.txt         getMappingFromSequence per contig (the mapper's file format)
//   <out>/mapping_chunks.txt  accumulateMappingChunk over chunkForMapping chunks on 4
//                             threads + serializeCoverage(diff, n)
//   <out>/multi_<i>.tsv       getStatsFromSequenceMulti per window, database i's rows
//   <out>/single_<i>.tsv      getStatsFromSequence per window
// and prints the process-wide switches and, for DeviceKmerIndex built from
// the second database, probeCountHost / probeHost of the first database's
// keys.
#include "DeviceKmerIndex.hpp"
#include "KmerDatabase.hpp"

#include <atomic>
#include <fstream>
#include <memory>
#include <thread>

int main(int argc, char **argv)
{
    if (argc < 5)
    {
        std::cerr << "usage: " << argv[0] << " <fasta> <out dir> <W> <db prefix>..." << std::endl;
        return 1;
    }
    try
    {
        const std::string fasta = argv[1], out = argv[2];
        const int W = std::atoi(argv[3]);
        std::vector<std::unique_ptr<KmerDatabase>> dbs;
        std::vector<KmerDatabase *> raw;
        for (int i = 4; i < argc; ++i)
        {
            dbs.push_back(std::make_unique<KmerDatabase>(argv[i]));
            raw.push_back(dbs.back().get());
        }
        const uint32_t k = dbs[0]->getKmerSize();
        ParseResult r = readSequenceFile(fasta);
        {
            std::ofstream m(out + "/mapping.txt", std::ios::binary);
            for (size_t c = 0; c < r.sequences.size(); ++c)
            {
                const auto idText = dbs[0]->getMappingFromSequence(r.ids[c], r.sequences[c]);
                m << idText.first << '\n' << idText.second << '\n';
            }
        }
        {
            std::vector<SeqWindow> chunks;
            std::vector<size_t> contigOf;
            chunkForMapping(r.ids, r.sequences, 100000, k, chunks, contigOf);
            std::vector<std::vector<int>> diffs;
            for (const auto &s : r.sequences)
                diffs.emplace_back(s.size() + 1, 0);
            std::atomic<size_t> next{0};
            std::vector<std::thread> pool;
            for (int t = 0; t < 4; ++t)
                pool.emplace_back([&] {
                    for (size_t j; (j = next++) < chunks.size();)
                        dbs[0]->accumulateMappingChunk(chunks[j], diffs[contigOf[j]]);
                });
            for (auto &th : pool)
                th.join();
            std::ofstream m(out + "/mapping_chunks.txt", std::ios::binary);
            for (size_t c = 0; c < r.sequences.size(); ++c)
                m << r.ids[c] << '\n'
                  << KmerDatabase::serializeCoverage(diffs[c], r.sequences[c].size()) << '\n';
        }
        std::vector<SeqWindow> wins;
        chunkSequences(r.ids, r.sequences, (size_t)W, k, wins);
        std::vector<std::vector<Window>> multi(dbs.size(), std::vector<Window>(wins.size())),
            single = multi;
        std::vector<Window> row(dbs.size());
        for (size_t w = 0; w < wins.size(); ++w)
        {
            KmerDatabase::getStatsFromSequenceMulti(wins[w], raw, row.data());
            for (size_t i = 0; i < dbs.size(); ++i)
            {
                multi[i][w] = row[i];
                single[i][w] = dbs[i]->getStatsFromSequence(wins[w]);
            }
        }
        for (size_t i = 0; i < dbs.size(); ++i)
        {
            KmerDatabase::writeStats(out + "/multi_" + std::to_string(i) + ".tsv", multi[i]);
            KmerDatabase::writeStats(out + "/single_" + std::to_string(i) + ".tsv", single[i]);
        }
        if (dbs.size() > 1)
        {
            auto shards = DeviceKmerIndex::buildShards(argv[5], 1);
            DeviceKmerIndex &idx = *shards[0];
            std::vector<uint64_t> q(dbs[0]->totalKmers());
            uint64_t nq = 0;
            hipibs::check(hipibs_kmc_read_keys(argv[4], q.data(), q.size(), &nq, 0), "read keys");
            q.resize(nq);
            std::vector<uint8_t> hits;
            idx.probeHost(q, hits);
            uint64_t sum = 0;
            for (uint8_t h : hits)
                sum += h;
            std::cout << "intersect 0-1: probeCountHost " << idx.probeCountHost(q)
                      << " probeHost " << sum << " devices " << DeviceKmerIndex::deviceCount()
                      << " footprint>0 " << (idx.footprintBytes() > 0) << " freeMemory>0 "
                      << (DeviceKmerIndex::freeMemory(idx.device()) > 0) << std::endl;
        }
        std::cout << "needsStringPath(" << k << ") " << KmerDatabase::needsStringPath(k)
                  << ", needsStringPath(64) " << KmerDatabase::needsStringPath(64)
                  << ", needsStringPath(65) " << KmerDatabase::needsStringPath(65)
                  << ", devicesToUse(4) " << KmerDatabase::devicesToUse(4) << ", gpuRequested "
                  << KmerDatabase::gpuRequested() << ", gpuSyncRequested "
                  << KmerDatabase::gpuSyncRequested() << ", poolThreads>0 "
                  << (KmerDatabase::poolThreads() > 0) << std::endl;
        return 0;
    }
    catch (const std::exception &e)
    {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
}
