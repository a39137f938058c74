// Streaming scan engine: a ring of `depth` slots, each with pinned host
// buffers, device buffers, its own stream and a completion event. A submit
// enqueues H2D of the packed bases, exception mask and windows, the fused
// scan, and D2H of the per-window statistics on the slot's stream, so the
// host packs slot i+1 while slot i is on the device (the role of
// EncodeProbeEngine, src/gpu/EncodeProbe.cu:469-1022; here the statistics,
// not per-position bit masks, come back: 16 B per window instead of
// 2 x B/8 bytes per batch).
#include "hipibs_internal.hpp"

#include <algorithm>
#include <vector>

using namespace hipibs;

namespace
{

struct Slot
{
    hipibs_host_slot host{};
    uint64_t *dSeq = nullptr, *dExc = nullptr;
    hipibs_window *dWin = nullptr;
    hipibs_window_stats *dStats = nullptr;
    void *dWork = nullptr;
    uint64_t workBytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool inFlight = false;
};

} // namespace

struct hipibs_scan
{
    std::vector<const hipibs_table *> tables;
    int device = 0;
    uint32_t k = 0;
    TableView *dViews = nullptr;
    std::vector<Slot> slots;
};

static void destroyScan(hipibs_scan *s)
{
    if (!s)
        return;
    hipSetDevice(s->device);
    for (auto &sl : s->slots)
    {
        if (sl.stream)
            hipStreamSynchronize(sl.stream);
        if (sl.host.seq_words)
            hipHostFree(sl.host.seq_words);
        if (sl.host.exc_words)
            hipHostFree(sl.host.exc_words);
        if (sl.host.windows)
            hipHostFree(sl.host.windows);
        if (sl.host.stats)
            hipHostFree(sl.host.stats);
        if (sl.dSeq)
            hipFree(sl.dSeq);
        if (sl.dExc)
            hipFree(sl.dExc);
        if (sl.dWin)
            hipFree(sl.dWin);
        if (sl.dStats)
            hipFree(sl.dStats);
        if (sl.dWork)
            hipFree(sl.dWork);
        if (sl.done)
            hipEventDestroy(sl.done);
        if (sl.stream)
            hipStreamDestroy(sl.stream);
    }
    if (s->dViews)
        hipFree(s->dViews);
    delete s;
}

extern "C" int hipibs_scan_create(const hipibs_table *const *tables, int n_tables,
                                  uint64_t batch_bases, uint32_t max_windows, int depth,
                                  hipibs_scan **out)
{
    if (!out)
        return setError(HIPIBS_E_ARG, "null argument");
    *out = nullptr;
    int rc = checkTables(tables, n_tables);
    if (rc)
        return rc;
    if (depth < 1 || depth > 16 || batch_bases == 0 || max_windows == 0)
        return setError(HIPIBS_E_ARG, "bad engine geometry");
    auto *s = new hipibs_scan();
    s->tables.assign(tables, tables + n_tables);
    s->device = tables[0]->device;
    s->k = tables[0]->k;
    if (hipSetDevice(s->device) != hipSuccess)
    {
        delete s;
        return setError(HIPIBS_E_HIP, "hipSetDevice failed");
    }
    std::vector<TableView> v(n_tables);
    for (int i = 0; i < n_tables; ++i)
        v[i] = tables[i]->view();
    const uint64_t seqCap = batch_bases / 32 + HIPIBS_PACK_PAD_WORDS + 2;
    const uint64_t excCap = batch_bases / 64 + HIPIBS_PACK_PAD_WORDS + 2;
    // work bound: units * segs <= 8192 + units (see autoSegsPerWindow)
    const uint64_t units = (uint64_t)max_windows * n_tables;
    const uint64_t workBytes = (units + 8192) * sizeof(SegSummary) + 64;
    bool ok = hipMalloc(reinterpret_cast<void **>(&s->dViews), n_tables * sizeof(TableView)) ==
                  hipSuccess &&
              hipMemcpy(s->dViews, v.data(), n_tables * sizeof(TableView),
                        hipMemcpyHostToDevice) == hipSuccess;
    s->slots.resize(depth);
    for (int i = 0; i < depth && ok; ++i)
    {
        Slot &sl = s->slots[i];
        sl.host.seq_cap = seqCap;
        sl.host.exc_cap = excCap;
        sl.host.window_cap = max_windows;
        sl.host.n_tables = (uint32_t)n_tables;
        sl.workBytes = workBytes;
        ok = hipHostMalloc(reinterpret_cast<void **>(&sl.host.seq_words), seqCap * 8) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&sl.host.exc_words), excCap * 8) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&sl.host.windows),
                           max_windows * sizeof(hipibs_window)) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&sl.host.stats),
                           units * sizeof(hipibs_window_stats)) == hipSuccess &&
             hipMalloc(reinterpret_cast<void **>(&sl.dSeq), seqCap * 8) == hipSuccess &&
             hipMalloc(reinterpret_cast<void **>(&sl.dExc), excCap * 8) == hipSuccess &&
             hipMalloc(reinterpret_cast<void **>(&sl.dWin), max_windows * sizeof(hipibs_window)) ==
                 hipSuccess &&
             hipMalloc(reinterpret_cast<void **>(&sl.dStats), units * sizeof(hipibs_window_stats)) ==
                 hipSuccess &&
             hipMalloc(&sl.dWork, workBytes) == hipSuccess &&
             hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess;
    }
    if (!ok)
    {
        destroyScan(s);
        return setError(HIPIBS_E_NOMEM, "scan engine allocation failed");
    }
    *out = s;
    return 0;
}

extern "C" int hipibs_scan_slot(hipibs_scan *s, int slot, hipibs_host_slot *out)
{
    if (!s || !out || slot < 0 || slot >= (int)s->slots.size())
        return setError(HIPIBS_E_ARG, "bad slot");
    *out = s->slots[slot].host;
    return 0;
}

extern "C" int hipibs_scan_submit(hipibs_scan *s, int slot, uint32_t n_windows, uint64_t n_seq_words,
                                  uint64_t n_exc_words)
{
    if (!s || slot < 0 || slot >= (int)s->slots.size())
        return setError(HIPIBS_E_ARG, "bad slot");
    Slot &sl = s->slots[slot];
    if (sl.inFlight)
        return setError(HIPIBS_E_ARG, "slot is in flight: wait for it first");
    if (n_windows > sl.host.window_cap ||
        n_seq_words + HIPIBS_PACK_PAD_WORDS > sl.host.seq_cap ||
        n_exc_words + HIPIBS_PACK_PAD_WORDS > sl.host.exc_cap)
        return setError(HIPIBS_E_ARG, "batch exceeds the slot capacity");
    const uint64_t nBases = std::min(n_seq_words * 32, n_exc_words * 64);
    uint32_t maxBases = 0;
    for (uint32_t i = 0; i < n_windows; ++i)
    {
        const hipibs_window &w = sl.host.windows[i];
        if (w.base + (uint64_t)w.n_bases > nBases)
            return setError(HIPIBS_E_ARG, "window outside the packed batch");
        maxBases = std::max(maxBases, w.n_bases);
    }
    HIPIBS_CHECK(hipSetDevice(s->device));
    const uint64_t seqCopy = n_seq_words + HIPIBS_PACK_PAD_WORDS;
    const uint64_t excCopy = n_exc_words + HIPIBS_PACK_PAD_WORDS;
    // Once a copy is queued, the slot's host buffers are in use: a failure
    // from then on waits for the stream before returning, so the caller may
    // refill the slot at once.
    auto failQueued = [&](int rc)
    {
        (void)hipStreamSynchronize(sl.stream);
        return rc;
    };
    hipError_t e = hipMemcpyAsync(sl.dSeq, sl.host.seq_words, seqCopy * 8, hipMemcpyHostToDevice,
                                  sl.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sl.dExc, sl.host.exc_words, excCopy * 8, hipMemcpyHostToDevice,
                           sl.stream);
    if (e == hipSuccess && n_windows)
        e = hipMemcpyAsync(sl.dWin, sl.host.windows, n_windows * sizeof(hipibs_window),
                           hipMemcpyHostToDevice, sl.stream);
    if (e != hipSuccess)
        return failQueued(hipFail(e, "scan batch upload"));
    const int nT = (int)s->tables.size();
    int rc = launchScan(s->tables.data(), nT, sl.dSeq, sl.dExc, nBases, sl.dWin, n_windows, maxBases,
                        sl.dStats, 0, sl.dWork, sl.workBytes, sl.stream, s->dViews);
    if (rc)
        return failQueued(rc);
    // results are table-major with stride n_windows
    if (n_windows)
        e = hipMemcpyAsync(sl.host.stats, sl.dStats,
                           (uint64_t)n_windows * nT * sizeof(hipibs_window_stats),
                           hipMemcpyDeviceToHost, sl.stream);
    if (e == hipSuccess)
        e = hipEventRecord(sl.done, sl.stream);
    if (e != hipSuccess)
        return failQueued(hipFail(e, "scan statistics download"));
    sl.inFlight = true;
    return 0;
}

extern "C" int hipibs_scan_wait(hipibs_scan *s, int slot)
{
    if (!s || slot < 0 || slot >= (int)s->slots.size())
        return setError(HIPIBS_E_ARG, "bad slot");
    Slot &sl = s->slots[slot];
    if (!sl.inFlight)
        return 0;
    sl.inFlight = false;
    HIPIBS_CHECK(hipEventSynchronize(sl.done));
    return 0;
}

extern "C" int hipibs_scan_free(hipibs_scan *s)
{
    destroyScan(s);
    return 0;
}
