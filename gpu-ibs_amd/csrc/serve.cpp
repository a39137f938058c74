// Serving engine: scan queries (sample, reference range, window size)
// against sample tables kept resident in HBM, with continuous batching.
//
// One worker thread owns the device work. Each turn it (1) retires
// finished launches, handing their statistics to requests and unpinning
// their tables; (2) if a launch slot is free, fills it from the pending
// requests in arrival order -- a request whose sample is not resident is
// admitted by building its table from the engine's compact host copy of
// the sample's keys on a builder thread (at most one build per builder in
// flight), after evicting
// tables that no launch holds, the least requested first, as far as the
// HBM budget requires; a table is only evicted for a sample requested at
// least as often, unless the waiting request is older than kForceAdmitMs
// -- and launches one mixed scan in which every
// window names its own table, so requests for different samples, ranges
// and window sizes share a launch; (3) sleeps until a submission or a
// completion if nothing moved. Up to max_inflight launches are in flight
// on their own streams, so requests that arrive while others run are
// admitted into the next launch and overlap on the device. A request's
// windows may span launches; its statistics do not depend on how it was
// batched (the per-window scan is independent of the launch).
#include "hipibs_internal.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace hipibs;

namespace
{

using Clock = std::chrono::steady_clock;

struct Request
{
    uint64_t id = 0;
    uint32_t sample = 0;
    std::vector<hipibs_window> wins;
    std::vector<hipibs_window_stats> stats;
    uint32_t nextWin = 0, doneWins = 0;
    Clock::time_point tSubmit, tFirstLaunch, tDone;
    bool launched = false;
};

struct Sample
{
    KeyStore store;        // the keys (samples registered by key arrays)
    std::string kmcPrefix; // set: the table is built from this KMC database
    uint64_t n = 0;
    uint64_t kmcStage = 0; // KMC samples: device bytes of the record decode's staging
    hipibs_table *table = nullptr;
    uint64_t bytes = 0;    // bytes counted against the budget: while building the table's
                           // estimate plus the build's transient memory, then the table's own
    bool building = false; // queued on or running in the builder thread
    int pins = 0;
    uint64_t lastUse = 0;
    uint64_t hits = 0; // requests received, halved every kAgeEvery requests
};

// Admission policy constants.
constexpr uint64_t kAgeEvery = 4096; // requests between halvings of the hit counts
constexpr double kForceAdmitMs = 50.0; // a request waiting longer may evict anything

struct Built
{
    uint32_t id;
    hipibs_table *table;
    std::string error;
};

// Bases of window positions one launch may take (keeps launches short so a
// small request never waits behind a whole chromosome of others).
constexpr uint64_t LAUNCH_BASES = 256ull << 20;
// A launch of few windows splits each into segments (one block each) until
// it has about kLaunchBlocksPerCu blocks per CU, keeping segments of at least
// kMinSegBases bases; at most kSlotSegs summaries per launch.
constexpr uint32_t kLaunchBlocksPerCu = 8;
constexpr uint64_t kMinSegBases = 4096;
constexpr uint32_t kSlotSegs = 1u << 15;

struct Part
{
    std::shared_ptr<Request> req;
    uint32_t first, count, offset;
};

struct Slot
{
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipibs_window *hWin = nullptr, *dWin = nullptr;
    hipibs_window_stats *hStats = nullptr, *dStats = nullptr;
    TableView *hViews = nullptr, *dViews = nullptr;
    SegSummary *dWork = nullptr; // segment summaries of a split launch
    bool busy = false;
    std::vector<Part> parts;
    std::vector<uint32_t> pinned;
};

} // namespace

struct hipibs_serve
{
    int device = 0;
    uint32_t k = 0;
    uint64_t budget = 0;
    hipibs_table_opts topts{};
    uint32_t maxWindows = 0;
    int cus = 256; // compute units of the device
    uint64_t *dSeq = nullptr, *dExc = nullptr;
    uint64_t nBases = 0;
    std::unordered_map<uint32_t, Sample> samples;
    uint64_t resident = 0, residentTables = 0, useClock = 0; // worker-owned
    uint64_t buildsInFlight = 0, requestsSeen = 0;           // worker-owned
    std::vector<Slot> slots;
    std::deque<std::shared_ptr<Request>> incoming, pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> completed;
    std::deque<uint64_t> completedOrder;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::thread worker;
    std::atomic<bool> stop{false};
    std::string workerError;
    hipibs_serve_counters counters{};
    // asynchronous admission: the builder thread turns queued sample ids
    // into tables while the worker keeps launching resident requests
    std::vector<std::thread> builders;
    std::deque<uint32_t> buildQueue;
    std::deque<hipibs_table *> freeQueue; // evicted tables, freed by the builders
    std::vector<Built> built;
    std::mutex bmu;
    std::condition_variable bcv;
};

namespace
{

double ms(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Evicts unpinned tables until `need` more bytes fit the budget: the
// least requested first (ties: least recently used). Without `force`, a
// table requested more often than the candidate (candHits) is never
// evicted for it. False if the bytes cannot be freed.
bool makeRoom(hipibs_serve *s, uint64_t need, uint64_t candHits, bool force)
{
    while (s->resident + need > s->budget)
    {
        Sample *victim = nullptr;
        for (auto &kv : s->samples)
            if (kv.second.table && kv.second.pins == 0 &&
                (!victim || kv.second.hits < victim->hits ||
                 (kv.second.hits == victim->hits && kv.second.lastUse < victim->lastUse)))
                victim = &kv.second;
        if (!victim || (!force && victim->hits > candHits))
            return false;
        // freeing can wait for the device; the builders do it off this thread
        {
            std::lock_guard<std::mutex> lk(s->bmu);
            s->freeQueue.push_back(victim->table);
        }
        s->bcv.notify_one();
        victim->table = nullptr;
        s->resident -= victim->bytes;
        victim->bytes = 0;
        s->residentTables -= 1;
        std::lock_guard<std::mutex> lk(s->mu);
        s->counters.evictions += 1;
        s->counters.resident_bytes = s->resident;
        s->counters.resident_tables = s->residentTables;
    }
    return true;
}

// Upper bound of the bytes of a table of n keys built with the engine's
// options (the sizing rules of the table builds).
uint64_t tableBytesFor(const hipibs_serve *s, uint64_t n)
{
    const double lf = s->topts.load_factor > 0 ? s->topts.load_factor : 0.6;
    const uint32_t slots = s->topts.bucket_slots ? s->topts.bucket_slots : 16;
    const uint32_t keysPerBucket = s->k > 32 ? 8 : slots; // k > 32: 8 keys of 16 B
    if (s->topts.minimizer_len && s->k <= 32)
    {
        // Two parts: minimizer-ordered buckets (keys of groups of at most
        // cap keys + one marker per larger group) and a whole-k-mer part
        // of 64 B buckets; together at most n keys + n / (cap + 1) markers
        // at the load factor, each part rounded up to whole buckets (>= 2).
        const uint32_t cap = s->topts.minimizer_cap ? s->topts.minimizer_cap : 32;
        const double items = (double)n + std::ceil((double)n / (cap + 1.0));
        const uint64_t slotBytes = (uint64_t)std::ceil(std::ceil(items / lf) * 8.0);
        return slotBytes + 3ull * slots * 8 + 3ull * 64;
    }
    const uint64_t buckets = std::max<uint64_t>(
        2, (uint64_t)std::ceil(std::ceil((double)std::max<uint64_t>(n, 1) / lf) / keysPerBucket));
    return buckets * slots * 8;
}

// Device bytes a build holds beside its table while it runs: a KMC
// sample's record staging and decoded keys, and for a minimizer table every
// key on the device plus the group map (a power of two >= n entries of
// 8 B). (Each builder thread's staging for builds from the key copies --
// two chunks of 16 Mi keys' suffixes and decoded keys, and the prefix
// offsets -- is allocated once and held outside the budget.)
uint64_t buildTransientFor(const hipibs_serve *s, const Sample &sm)
{
    const uint64_t keyBytes = s->k > 32 ? 16 : 8;
    uint64_t b = 0;
    if (!sm.kmcPrefix.empty())
        b += sm.kmcStage + std::min<uint64_t>(sm.n, 32ull << 20) * keyBytes;
    if (s->topts.minimizer_len && s->k <= 32)
    {
        uint64_t map = 1024;
        while (map < sm.n)
            map <<= 1;
        b += sm.n * keyBytes + map * 8;
    }
    return b;
}

// After a build: the sample's reservation becomes its table's real size.
void settleBytes(hipibs_serve *s, Sample &sm)
{
    uint64_t bytes = sm.bytes;
    if (sm.table)
        (void)hipibs_table_info(sm.table, nullptr, &bytes, nullptr, nullptr, nullptr, nullptr);
    s->resident = s->resident - sm.bytes + bytes;
    sm.bytes = bytes;
}

// Builds sample id's table now (preload: from its keys on the device,
// dKeys, or its KMC database) -- or, async, queues it on a builder thread
// -- after making room. 0 = resident, 1 = not (yet), <0 = error. The bytes
// of a queued build are reserved against the budget. `force`: the waiting
// request is old enough to evict any idle table.
int ensureResident(hipibs_serve *s, uint32_t id, bool async, bool force,
                   const uint64_t *dKeys = nullptr)
{
    Sample &sm = s->samples.at(id);
    sm.lastUse = ++s->useClock;
    if (sm.table)
        return 0;
    if (sm.building)
        return 1;
    if (async && s->buildsInFlight >= s->builders.size())
        return 1;
    const uint64_t tableBytes = tableBytesFor(s, sm.n);
    if (tableBytes > s->budget)
        return setError(HIPIBS_E_NOMEM, "a sample table alone exceeds the serving HBM budget");
    // the build's transient memory is reserved as far as the budget allows
    const uint64_t need = std::min(s->budget, tableBytes + buildTransientFor(s, sm));
    if (!makeRoom(s, need, sm.hits, force || !async))
        return 1;
    sm.bytes = need;
    s->resident += need;
    if (async)
    {
        sm.building = true;
        s->buildsInFlight += 1;
        {
            std::lock_guard<std::mutex> lk(s->bmu);
            s->buildQueue.push_back(id);
        }
        s->bcv.notify_one();
        return 1;
    }
    hipibs_table_opts o = s->topts;
    o.device = s->device;
    int rc = !sm.kmcPrefix.empty()
                 ? hipibs_table_build_kmc(sm.kmcPrefix.c_str(), &o, &sm.table)
                 : dKeys ? buildKeysDevice(dKeys, sm.n, s->k, &o, &sm.table, nullptr, false)
                         : setError(HIPIBS_E_ARG, "no device keys for a preload");
    if (rc)
    {
        s->resident -= need;
        sm.bytes = 0;
        return rc;
    }
    settleBytes(s, sm);
    s->residentTables += 1;
    std::lock_guard<std::mutex> lk(s->mu);
    s->counters.admissions += 1;
    s->counters.resident_bytes = s->resident;
    s->counters.resident_tables = s->residentTables;
    return 0;
}

void builderLoop(hipibs_serve *s)
{
    (void)hipSetDevice(s->device);
    StoreStaging *staging = nullptr; // kept across builds
    for (;;)
    {
        uint32_t id = 0;
        bool haveBuild = false;
        std::vector<hipibs_table *> toFree;
        {
            std::unique_lock<std::mutex> lk(s->bmu);
            s->bcv.wait(lk, [s] {
                return s->stop || !s->buildQueue.empty() || !s->freeQueue.empty();
            });
            toFree.assign(s->freeQueue.begin(), s->freeQueue.end());
            s->freeQueue.clear();
            if (!s->stop && !s->buildQueue.empty())
            {
                id = s->buildQueue.front();
                s->buildQueue.pop_front();
                haveBuild = true;
            }
        }
        if (!toFree.empty())
        {
            const auto t0 = Clock::now();
            for (hipibs_table *t : toFree)
                freeTableOn(t, nullptr); // stream-ordered tables: no wait for the device
            const uint64_t us = (uint64_t)(ms(t0, Clock::now()) * 1000.0);
            std::lock_guard<std::mutex> lk(s->mu);
            s->counters.evict_us += us;
        }
        if (!haveBuild)
        {
            if (s->stop)
                break;
            continue;
        }
        const Sample &sm = s->samples.at(id); // keys and n never change
        hipibs_table_opts o = s->topts;
        o.device = s->device;
        hipibs_table *t = nullptr;
        const auto t0 = Clock::now();
        int rc = 0;
        if (!sm.kmcPrefix.empty())
            rc = hipibs_table_build_kmc(sm.kmcPrefix.c_str(), &o, &t);
        else
        {
            if (!staging)
                rc = storeStagingCreate(s->device, &staging);
            if (!rc)
                rc = buildTableFromStore(sm.store, &o, staging, &t, true);
        }
        const uint64_t us = (uint64_t)(ms(t0, Clock::now()) * 1000.0);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->counters.build_us += us;
        }
        {
            std::lock_guard<std::mutex> lk(s->bmu);
            s->built.push_back(Built{id, rc ? nullptr : t, rc ? hipibs_last_error() : ""});
        }
        s->cvWork.notify_one();
    }
    storeStagingFree(staging);
}

// Worker side of a finished build: the table becomes resident.
int collectBuilds(hipibs_serve *s, bool &progress)
{
    std::vector<Built> done;
    {
        std::lock_guard<std::mutex> lk(s->bmu);
        done.swap(s->built);
    }
    for (auto &b : done)
    {
        Sample &sm = s->samples.at(b.id);
        sm.building = false;
        s->buildsInFlight -= 1;
        if (!b.table)
        {
            s->resident -= sm.bytes;
            sm.bytes = 0;
            return setError(HIPIBS_E_HIP, "table build failed: " + b.error);
        }
        sm.table = b.table;
        settleBytes(s, sm);
        s->residentTables += 1;
        progress = true;
        std::lock_guard<std::mutex> lk(s->mu);
        s->counters.admissions += 1;
        s->counters.resident_bytes = s->resident;
        s->counters.resident_tables = s->residentTables;
    }
    return 0;
}

void retire(hipibs_serve *s, Slot &sl)
{
    const auto now = Clock::now();
    std::vector<std::shared_ptr<Request>> finished;
    for (auto &p : sl.parts)
    {
        std::copy(sl.hStats + p.offset, sl.hStats + p.offset + p.count,
                  p.req->stats.begin() + p.first);
        p.req->doneWins += p.count;
        if (p.req->doneWins == p.req->wins.size())
        {
            p.req->tDone = now;
            finished.push_back(p.req);
        }
    }
    for (uint32_t id : sl.pinned)
        s->samples.at(id).pins -= 1;
    sl.parts.clear();
    sl.pinned.clear();
    sl.busy = false;
    if (!finished.empty())
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto &r : finished)
        {
            s->completed[r->id] = r;
            s->completedOrder.push_back(r->id);
        }
        s->counters.completed += finished.size();
        s->cvDone.notify_all();
    }
}

// Fills a free slot from the pending queue and launches it. Returns true
// if a launch was made.
int fillAndLaunch(hipibs_serve *s, Slot &sl, bool &launched)
{
    launched = false;
    uint32_t used = 0;
    uint64_t bases = 0;
    std::vector<uint32_t> viewOf; // sample id per view index
    std::unordered_map<uint32_t, uint32_t> viewIdx;
    for (auto it = s->pending.begin();
         it != s->pending.end() && used < s->maxWindows && bases < LAUNCH_BASES;)
    {
        const std::shared_ptr<Request> &r = *it;
        const int rc = ensureResident(s, r->sample, true,
                                      ms(r->tSubmit, Clock::now()) > kForceAdmitMs);
        if (rc < 0)
            return rc;
        if (rc == 1)
        {
            ++it; // being built, or memory held by in-flight launches: later
            continue;
        }
        uint32_t vi;
        auto f = viewIdx.find(r->sample);
        if (f == viewIdx.end())
        {
            vi = (uint32_t)viewOf.size();
            viewIdx[r->sample] = vi;
            viewOf.push_back(r->sample);
            s->samples.at(r->sample).pins += 1;
            sl.pinned.push_back(r->sample);
        }
        else
            vi = f->second;
        uint32_t cnt = 0;
        while (r->nextWin + cnt < r->wins.size() && used + cnt < s->maxWindows &&
               (bases < LAUNCH_BASES || cnt == 0))
            bases += r->wins[r->nextWin + cnt++].n_bases;
        for (uint32_t j = 0; j < cnt; ++j)
        {
            hipibs_window w = r->wins[r->nextWin + j];
            w.reserved = vi;
            sl.hWin[used + j] = w;
        }
        if (!r->launched)
        {
            r->launched = true;
            r->tFirstLaunch = Clock::now();
        }
        sl.parts.push_back(Part{r, r->nextWin, cnt, used});
        r->nextWin += cnt;
        used += cnt;
        if (r->nextWin == r->wins.size())
            it = s->pending.erase(it);
        else
            ++it;
    }
    if (used == 0)
    {
        for (uint32_t id : sl.pinned)
            s->samples.at(id).pins -= 1;
        sl.pinned.clear();
        return 0;
    }
    for (size_t v = 0; v < viewOf.size(); ++v)
        sl.hViews[v] = s->samples.at(viewOf[v]).table->view();
    const int slots = s->samples.at(viewOf[0]).table->bucketSlots;
    HIPIBS_CHECK(hipMemcpyAsync(sl.dViews, sl.hViews, viewOf.size() * sizeof(TableView),
                                hipMemcpyHostToDevice, sl.stream));
    HIPIBS_CHECK(hipMemcpyAsync(sl.dWin, sl.hWin, used * sizeof(hipibs_window),
                                hipMemcpyHostToDevice, sl.stream));
    // few windows (a lone small request): segments keep the CUs busy, so
    // its latency is not one block's walk over a whole window
    uint32_t segs = 1;
    const uint64_t want = (uint64_t)s->cus * kLaunchBlocksPerCu;
    if (used < want)
        segs = (uint32_t)std::min<uint64_t>(
            {(want + used - 1) / used, std::max<uint64_t>(1, bases / used / kMinSegBases),
             kSlotSegs / used});
    segs = std::max<uint32_t>(segs, 1);
    // every table of the engine is built with the engine's options
    int rc = launchScanMixed(slots, s->k, sl.dViews, s->dSeq, s->dExc, s->nBases, sl.dWin, used,
                             sl.dStats, sl.stream, s->topts.minimizer_len != 0 && s->k <= 32,
                             segs, sl.dWork);
    if (rc)
        return rc;
    HIPIBS_CHECK(hipMemcpyAsync(sl.hStats, sl.dStats, used * sizeof(hipibs_window_stats),
                                hipMemcpyDeviceToHost, sl.stream));
    HIPIBS_CHECK(hipEventRecord(sl.done, sl.stream));
    sl.busy = true;
    launched = true;
    std::lock_guard<std::mutex> lk(s->mu);
    s->counters.launches += 1;
    s->counters.windows += used;
    s->counters.resident_bytes = s->resident;
    s->counters.resident_tables = s->residentTables;
    return 0;
}

void workerLoop(hipibs_serve *s)
{
    if (hipSetDevice(s->device) != hipSuccess)
    {
        s->workerError = "hipSetDevice failed in the serving worker";
        return;
    }
    while (!s->stop)
    {
        bool progress = false;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            while (!s->incoming.empty())
            {
                s->samples.at(s->incoming.front()->sample).hits += 1;
                if (++s->requestsSeen % kAgeEvery == 0)
                    for (auto &kv : s->samples)
                        kv.second.hits /= 2;
                s->pending.push_back(std::move(s->incoming.front()));
                s->incoming.pop_front();
                progress = true;
            }
        }
        if (collectBuilds(s, progress))
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->workerError = hipibs_last_error();
            s->stop = true;
            s->cvDone.notify_all();
            return;
        }
        for (auto &sl : s->slots)
            if (sl.busy && hipEventQuery(sl.done) == hipSuccess)
            {
                retire(s, sl);
                progress = true;
            }
        for (auto &sl : s->slots)
        {
            if (sl.busy || s->pending.empty())
                continue;
            bool launched = false;
            const int rc = fillAndLaunch(s, sl, launched);
            if (rc)
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->workerError = hipibs_last_error();
                s->stop = true;
                s->cvDone.notify_all();
                return;
            }
            progress |= launched;
        }
        if (!progress)
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cvWork.wait_for(lk, std::chrono::microseconds(200));
        }
    }
}

void destroy(hipibs_serve *s)
{
    s->stop = true;
    s->cvWork.notify_all();
    {
        std::lock_guard<std::mutex> lk(s->bmu);
    }
    s->bcv.notify_all();
    if (s->worker.joinable())
        s->worker.join();
    for (auto &b : s->builders)
        if (b.joinable())
            b.join();
    for (auto &b : s->built)
        hipibs_table_free(b.table);
    for (hipibs_table *t : s->freeQueue)
        hipibs_table_free(t);
    s->freeQueue.clear();
    (void)hipSetDevice(s->device);
    for (auto &sl : s->slots)
    {
        if (sl.stream)
            (void)hipStreamSynchronize(sl.stream);
        (void)hipHostFree(sl.hWin);
        (void)hipHostFree(sl.hStats);
        (void)hipHostFree(sl.hViews);
        (void)hipFree(sl.dWin);
        (void)hipFree(sl.dStats);
        (void)hipFree(sl.dViews);
        (void)hipFree(sl.dWork);
        if (sl.done)
            (void)hipEventDestroy(sl.done);
        if (sl.stream)
            (void)hipStreamDestroy(sl.stream);
    }
    for (auto &kv : s->samples)
    {
        hipibs_table_free(kv.second.table);
        keyStoreFree(kv.second.store);
    }
    (void)hipFree(s->dSeq);
    (void)hipFree(s->dExc);
    // the pool's retained memory goes back once the frees above are done
    hipMemPool_t pool = nullptr;
    if (hipDeviceSynchronize() == hipSuccess &&
        hipDeviceGetDefaultMemPool(&pool, s->device) == hipSuccess && pool)
        (void)hipMemPoolTrimTo(pool, 0);
    delete s;
}

} // namespace

extern "C" int hipibs_serve_create(int device, uint32_t k, uint64_t hbm_budget,
                                   const hipibs_table_opts *topts, int max_inflight,
                                   uint32_t max_windows_per_launch, hipibs_serve **out)
{
    if (!out || k == 0 || k > 64 || max_inflight < 1 || max_inflight > 16 ||
        max_windows_per_launch == 0 || hbm_budget == 0)
        return setError(HIPIBS_E_ARG, "bad serving engine arguments");
    *out = nullptr;
    HIPIBS_CHECK(hipSetDevice(device));
    auto *s = new hipibs_serve();
    s->device = device;
    s->k = k;
    s->budget = hbm_budget;
    if (topts)
        s->topts = *topts;
    s->maxWindows = max_windows_per_launch;
    // admissions allocate tables from the device's default memory pool and
    // evictions return them there (hipMallocAsync / hipFreeAsync): freed
    // memory stays in the pool for the next admission instead of going back
    // to the driver, and neither call waits for the device
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool)
    {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    s->slots.resize(max_inflight);
    bool ok = true;
    for (auto &sl : s->slots)
    {
        const uint32_t W = max_windows_per_launch;
        ok = ok && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&sl.hWin), W * sizeof(hipibs_window)) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&sl.hStats), W * sizeof(hipibs_window_stats)) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&sl.hViews), W * sizeof(TableView)) == hipSuccess &&
             hipMalloc(reinterpret_cast<void **>(&sl.dWin), W * sizeof(hipibs_window)) == hipSuccess &&
             hipMalloc(reinterpret_cast<void **>(&sl.dStats), W * sizeof(hipibs_window_stats)) == hipSuccess &&
             hipMalloc(reinterpret_cast<void **>(&sl.dViews), W * sizeof(TableView)) == hipSuccess &&
             hipMalloc(reinterpret_cast<void **>(&sl.dWork), kSlotSegs * sizeof(SegSummary)) == hipSuccess;
    }
    (void)hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, device);
    if (!ok)
    {
        destroy(s);
        return setError(HIPIBS_E_NOMEM, "serving engine allocation failed");
    }
    *out = s;
    return 0;
}

extern "C" int hipibs_serve_set_reference(hipibs_serve *s, const uint64_t *h_seq_words,
                                          const uint64_t *h_exc_words, uint64_t n_bases)
{
    if (!s || !h_seq_words || !h_exc_words)
        return setError(HIPIBS_E_ARG, "null argument");
    if (s->worker.joinable())
        return setError(HIPIBS_E_ARG, "the reference is set before the first sample");
    HIPIBS_CHECK(hipSetDevice(s->device));
    const uint64_t nsw = (n_bases + 31) / 32 + HIPIBS_PACK_PAD_WORDS;
    const uint64_t new_ = (n_bases + 63) / 64 + HIPIBS_PACK_PAD_WORDS;
    HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&s->dSeq), nsw * 8));
    HIPIBS_CHECK(hipMalloc(reinterpret_cast<void **>(&s->dExc), new_ * 8));
    HIPIBS_CHECK(hipMemcpy(s->dSeq, h_seq_words, nsw * 8, hipMemcpyHostToDevice));
    HIPIBS_CHECK(hipMemcpy(s->dExc, h_exc_words, new_ * 8, hipMemcpyHostToDevice));
    s->nBases = n_bases;
    return 0;
}

static int checkRegistration(hipibs_serve *s, uint32_t sample_id)
{
    if (!s->dSeq)
        return setError(HIPIBS_E_ARG, "set the reference first");
    if (s->worker.joinable())
        return setError(HIPIBS_E_ARG, "samples are registered before the first request");
    if (s->samples.count(sample_id))
        return setError(HIPIBS_E_ARG, "sample id already registered");
    return 0;
}

static int addSample(hipibs_serve *s, uint32_t sample_id, Sample sm, int preload,
                     const uint64_t *dKeys)
{
    int rc = checkRegistration(s, sample_id);
    if (rc)
    {
        keyStoreFree(sm.store);
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->counters.host_bytes += sm.store.suf ? sm.store.hostBytes() : 0;
    }
    s->samples[sample_id] = std::move(sm);
    if (preload)
    {
        rc = ensureResident(s, sample_id, false, true, dKeys);
        if (rc < 0)
            return rc;
        if (rc == 1)
            return setError(HIPIBS_E_NOMEM, "preloaded samples exceed the HBM budget");
    }
    return 0;
}

// A sample from keys on the device: the engine's host copy, then (preload)
// the table from the same device keys.
static int addSampleDevice(hipibs_serve *s, uint32_t sample_id, const uint64_t *dKeys,
                           uint64_t n, int preload)
{
    Sample sm;
    sm.n = n;
    int rc = keyStoreFromDevice(dKeys, n, s->k, s->device, nullptr, sm.store);
    if (rc)
        return rc;
    return addSample(s, sample_id, std::move(sm), preload, dKeys);
}

extern "C" int hipibs_serve_add_sample(hipibs_serve *s, uint32_t sample_id, const uint64_t *h_keys,
                                       uint64_t n_keys, int preload)
{
    if (!s || (n_keys && !h_keys))
        return setError(HIPIBS_E_ARG, "null argument");
    int rc = checkRegistration(s, sample_id);
    if (rc)
        return rc;
    HIPIBS_CHECK(hipSetDevice(s->device));
    const uint64_t bytes = std::max<uint64_t>(n_keys, 1) * (s->k > 32 ? 16 : 8);
    uint64_t *dKeys = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&dKeys), bytes) != hipSuccess)
        return setError(HIPIBS_E_NOMEM, "device memory to register a sample's keys");
    rc = n_keys && hipMemcpy(dKeys, h_keys, n_keys * (s->k > 32 ? 16 : 8),
                             hipMemcpyHostToDevice) != hipSuccess
             ? setError(HIPIBS_E_HIP, "key upload failed")
             : 0;
    if (!rc)
        rc = addSampleDevice(s, sample_id, dKeys, n_keys, preload);
    (void)hipFree(dKeys);
    return rc;
}

extern "C" int hipibs_serve_add_sample_device(hipibs_serve *s, uint32_t sample_id,
                                              const uint64_t *d_keys, uint64_t n_keys, int preload)
{
    if (!s || (n_keys && !d_keys))
        return setError(HIPIBS_E_ARG, "null argument");
    int rc = checkRegistration(s, sample_id);
    if (rc)
        return rc;
    HIPIBS_CHECK(hipSetDevice(s->device));
    return addSampleDevice(s, sample_id, d_keys, n_keys, preload);
}

extern "C" int hipibs_serve_add_sample_kmc(hipibs_serve *s, uint32_t sample_id,
                                           const char *kmc_prefix, int preload)
{
    if (!s || !kmc_prefix)
        return setError(HIPIBS_E_ARG, "null argument");
    hipibs_kmc_info info{};
    int rc = hipibs_kmc_header(kmc_prefix, &info);
    if (rc)
        return rc;
    if (info.kmer_length != s->k)
        return setError(HIPIBS_E_ARG, "the KMC database's k differs from the engine's");
    Sample sm;
    sm.kmcPrefix = kmc_prefix;
    sm.n = info.total_kmers;
    // the decode's staging: two raw record chunks (of at most 32 Mi records),
    // and the LUT, which the .kmc_pre file holds (its size bounds the LUT's)
    const uint64_t recBytes =
        (info.kmer_length - info.lut_prefix_length) / 4 + info.counter_size;
    struct stat st{};
    const uint64_t preBytes =
        stat((std::string(kmc_prefix) + ".kmc_pre").c_str(), &st) == 0 ? (uint64_t)st.st_size : 0;
    sm.kmcStage = 2 * std::min<uint64_t>(info.total_kmers, 32ull << 20) * recBytes + preBytes;
    return addSample(s, sample_id, std::move(sm), preload, nullptr);
}

extern "C" int hipibs_serve_submit(hipibs_serve *s, uint64_t request_id, uint32_t sample_id,
                                   uint64_t base, uint64_t n_bases, uint32_t W)
{
    if (!s)
        return setError(HIPIBS_E_ARG, "null engine");
    if (!s->samples.count(sample_id))
        return setError(HIPIBS_E_ARG, "unknown sample id");
    if (W <= s->k)
        return setError(HIPIBS_E_ARG, "window size must be larger than k");
    if (base + n_bases > s->nBases)
        return setError(HIPIBS_E_ARG, "range outside the reference");
    auto r = std::make_shared<Request>();
    r->id = request_id;
    r->sample = sample_id;
    // chunkSequences geometry over the range (src/Utils.hpp:296-327)
    for (uint64_t j = 0; j < n_bases; j += W - s->k)
        r->wins.push_back(hipibs_window{base + j, (uint32_t)std::min<uint64_t>(W, n_bases - j), 0});
    r->stats.resize(r->wins.size());
    r->tSubmit = Clock::now();
    if (r->wins.empty())
    {
        r->tFirstLaunch = r->tDone = r->tSubmit;
        std::lock_guard<std::mutex> lk(s->mu);
        s->completed[request_id] = r;
        s->completedOrder.push_back(request_id);
        s->counters.completed += 1;
        s->cvDone.notify_all();
        return 0;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->workerError.empty())
            return setError(HIPIBS_E_HIP, "serving worker stopped: " + s->workerError);
        s->incoming.push_back(std::move(r));
        s->counters.submitted += 1;
        if (!s->worker.joinable())
        {
            // concurrent builds overlap one sample's host-side staging with
            // another's upload and insert
            for (int i = 0; i < 3; ++i)
                s->builders.emplace_back(builderLoop, s);
            s->worker = std::thread(workerLoop, s);
        }
    }
    s->cvWork.notify_one();
    return 0;
}

extern "C" int hipibs_serve_poll(hipibs_serve *s, uint64_t *ids, uint32_t max, int timeout_ms,
                                 uint32_t *n_out)
{
    if (!s || !n_out || (max && !ids))
        return setError(HIPIBS_E_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->completedOrder.empty() && timeout_ms > 0)
        s->cvDone.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                           [s] { return !s->completedOrder.empty() || !s->workerError.empty(); });
    if (!s->workerError.empty() && s->completedOrder.empty())
        return setError(HIPIBS_E_HIP, "serving worker stopped: " + s->workerError);
    uint32_t n = 0;
    while (n < max && !s->completedOrder.empty())
    {
        ids[n++] = s->completedOrder.front();
        s->completedOrder.pop_front();
    }
    *n_out = n;
    return 0;
}

extern "C" int hipibs_serve_fetch(hipibs_serve *s, uint64_t request_id, hipibs_window_stats *out,
                                  uint32_t cap, uint32_t *n_windows, double *latency_ms,
                                  double *queue_ms)
{
    if (!s)
        return setError(HIPIBS_E_ARG, "null engine");
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->completed.find(request_id);
        if (it == s->completed.end())
            return setError(HIPIBS_E_ARG, "request not completed (or already fetched)");
        r = it->second;
        if (out) // a call without a buffer only reports the sizes
            s->completed.erase(it);
    }
    if (n_windows)
        *n_windows = (uint32_t)r->wins.size();
    if (out)
        std::copy(r->stats.begin(), r->stats.begin() + std::min<size_t>(cap, r->stats.size()), out);
    if (latency_ms)
        *latency_ms = ms(r->tSubmit, r->tDone);
    if (queue_ms)
        *queue_ms = ms(r->tSubmit, r->tFirstLaunch);
    return 0;
}

extern "C" int hipibs_serve_get_counters(hipibs_serve *s, hipibs_serve_counters *out)
{
    if (!s || !out)
        return setError(HIPIBS_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->counters;
    return 0;
}

extern "C" int hipibs_serve_free(hipibs_serve *s)
{
    if (s)
        destroy(s);
    return 0;
}
