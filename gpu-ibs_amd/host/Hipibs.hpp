// C++ face of the C ABI (include/hipibs.h): RAII handles that throw
// std::runtime_error with the library's message on failure, which is the
// reference's error convention (cudaCheck, src/gpu/EncodeProbe.cu:15-20).
#pragma once

#include "hipibs.h"

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hipibs
{

inline void check(int rc, const char *what)
{
    if (rc != HIPIBS_OK)
        throw std::runtime_error(std::string(what) + ": " + hipibs_last_error());
}

struct TableDeleter
{
    void operator()(hipibs_table *t) const { hipibs_table_free(t); }
};
using TablePtr = std::unique_ptr<hipibs_table, TableDeleter>;

inline hipibs_table_opts defaultOpts()
{
    hipibs_table_opts o{};
    o.load_factor = 0.6;
    o.bucket_slots = 16;
    return o;
}

inline TablePtr tableFromKmc(const std::string &prefix, const hipibs_table_opts &o)
{
    hipibs_table *t = nullptr;
    check(hipibs_table_build_kmc(prefix.c_str(), &o, &t), "building the device table");
    return TablePtr(t);
}

struct TableInfo
{
    uint64_t keys = 0, bytes = 0, buckets = 0;
    uint32_t k = 0, slots = 0;
    int device = 0;
};

inline TableInfo info(const hipibs_table *t)
{
    TableInfo i;
    check(hipibs_table_info(t, &i.keys, &i.bytes, &i.buckets, &i.k, &i.slots, &i.device),
          "table info");
    return i;
}

inline int deviceCount(std::string &why)
{
    int n = 0;
    char buf[512] = {0};
    hipibs_device_count(&n, buf, sizeof(buf));
    why = buf;
    return n;
}

// Device buffer owned by the library.
class DeviceBuffer
{
public:
    DeviceBuffer() = default;
    DeviceBuffer(int device, uint64_t bytes) { check(hipibs_dev_alloc(device, bytes, &p_), "alloc"); }
    ~DeviceBuffer() { hipibs_dev_free(p_); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : p_(o.p_) { o.p_ = nullptr; }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept
    {
        std::swap(p_, o.p_);
        return *this;
    }
    void *get() const { return p_; }
    template <typename T> T *as() const { return static_cast<T *>(p_); }

private:
    void *p_ = nullptr;
};

// The streaming scan engine (pinned slots, one stream each).
class ScanEngine
{
public:
    ScanEngine(const std::vector<const hipibs_table *> &tables, uint64_t batchBases,
               uint32_t maxWindows, int depth)
    {
        check(hipibs_scan_create(tables.data(), (int)tables.size(), batchBases, maxWindows, depth,
                                 &s_),
              "creating the scan engine");
        slots_.resize(depth);
        for (int i = 0; i < depth; ++i)
            check(hipibs_scan_slot(s_, i, &slots_[i]), "scan slot");
    }
    ~ScanEngine() { hipibs_scan_free(s_); }
    ScanEngine(const ScanEngine &) = delete;
    ScanEngine &operator=(const ScanEngine &) = delete;
    int depth() const { return (int)slots_.size(); }
    const hipibs_host_slot &slot(int i) const { return slots_[i]; }
    void submit(int i, uint32_t nWindows, uint64_t seqWords, uint64_t excWords)
    {
        check(hipibs_scan_submit(s_, i, nWindows, seqWords, excWords), "scan submit");
    }
    void wait(int i) { check(hipibs_scan_wait(s_, i), "scan wait"); }

private:
    hipibs_scan *s_ = nullptr;
    std::vector<hipibs_host_slot> slots_;
};

} // namespace hipibs
