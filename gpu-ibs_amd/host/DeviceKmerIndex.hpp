// DeviceKmerIndex with the reference's member names
// (src/gpu/DeviceKmerIndex.cuh:26-124) over the C ABI: a sample's device
// table built from its KMC database, the asynchronous batched probe, the
// synchronous host-staged probe and probe count, and the device queries.
// The reference builds it from a host KmcRaIndex image (LUT + suffix
// records) and probes by binary search; here the records stream from the
// KMC files into a hash table. One GPU per process, so buildShards makes
// one shard; the reference's L2-persisting-LUT ablation has no
// counterpart (there is no LUT on the device).
#pragma once

#include "Hipibs.hpp"

#include <cstring>
#include <mutex>

class DeviceKmerIndex
{
public:
    // Builds the table of the KMC database at `prefix` on `device`.
    explicit DeviceKmerIndex(const std::string &prefix, int device = 0) : dev_(device)
    {
        hipibs_table_opts o = hipibs::defaultOpts();
        o.device = device;
        table_ = hipibs::tableFromKmc(prefix, o);
        footprint_ = hipibs::info(table_.get()).bytes;
    }
    DeviceKmerIndex(const DeviceKmerIndex &) = delete;
    DeviceKmerIndex &operator=(const DeviceKmerIndex &) = delete;

    // One shard per device in use; this build uses one device.
    static std::vector<std::unique_ptr<DeviceKmerIndex>> buildShards(const std::string &prefix,
                                                                     int nDevices)
    {
        if (nDevices != 1)
            throw std::runtime_error("this build runs on one GPU: buildShards needs nDevices == 1");
        std::vector<std::unique_ptr<DeviceKmerIndex>> v;
        v.push_back(std::make_unique<DeviceKmerIndex>(prefix, 0));
        return v;
    }

    int device() const { return dev_; }
    const hipibs_table *table() const { return table_.get(); }
    size_t footprintBytes() const { return footprint_; }

    // Asynchronous batched probe of device-resident queries: d_hits[i] = 1
    // if d_valid (nullable) marks i valid and d_queries[i] is in the table.
    void probe(const uint64_t *d_queries, const uint8_t *d_valid, size_t n, uint8_t *d_hits,
               void *stream) const
    {
        hipibs::check(hipibs_table_probe(table_.get(), d_queries, d_valid, n, d_hits, stream),
                      "probe");
    }

    // Synchronous probe of host queries through device staging buffers
    // (grown on demand). Thread-safe: calls are serialised.
    void probeHost(const std::vector<uint64_t> &queries, std::vector<uint8_t> &hits)
    {
        std::lock_guard<std::mutex> lk(stageMutex_);
        hits.assign(queries.size(), 0);
        if (queries.empty())
            return;
        ensureStage(queries.size());
        hipibs::check(hipibs_memcpy(stageQ_.get(), queries.data(), queries.size() * 8, 1),
                      "query upload");
        probe(stageQ_.as<uint64_t>(), nullptr, queries.size(), stageH_.as<uint8_t>(), nullptr);
        hipibs::check(hipibs_dev_synchronize(dev_), "probe");
        hipibs::check(hipibs_memcpy(hits.data(), stageH_.get(), queries.size(), 2), "hit download");
    }

    // Number of the queries present in the table (only the count comes
    // back). Thread-safe.
    uint64_t probeCountHost(const std::vector<uint64_t> &queries)
    {
        uint64_t count = 0;
        hipibs::check(hipibs_table_probe_count(table_.get(), queries.data(), queries.size(),
                                               &count),
                      "probe count");
        return count;
    }

    // Number of usable devices (0 without a device or driver) and why the
    // last call found none.
    static int deviceCount()
    {
        std::string why;
        const int n = hipibs::deviceCount(why);
        lastError() = why;
        return n;
    }
    static const char *deviceCountError() { return lastError().c_str(); }

    // Free memory of `device` in bytes; 0 when it cannot be queried.
    static size_t freeMemory(int device)
    {
        uint64_t freeB = 0, totalB = 0;
        return hipibs_device_mem(device, &freeB, &totalB) == HIPIBS_OK ? (size_t)freeB : 0;
    }

private:
    static std::string &lastError()
    {
        static std::string s;
        return s;
    }
    void ensureStage(size_t n)
    {
        if (n <= stageCap_)
            return;
        stageQ_ = hipibs::DeviceBuffer(dev_, n * 8);
        stageH_ = hipibs::DeviceBuffer(dev_, n);
        stageCap_ = n;
    }

    int dev_ = 0;
    hipibs::TablePtr table_;
    size_t footprint_ = 0;
    std::mutex stageMutex_;
    hipibs::DeviceBuffer stageQ_, stageH_;
    size_t stageCap_ = 0;
};
