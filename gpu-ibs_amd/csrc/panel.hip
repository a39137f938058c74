// Panel scan: one persistent launch scans a group of sample tables over a
// shared packed reference. Every 64-position word is encoded once (codec,
// validity, the table-independent part of the bucket hash -- or, for
// minimizer tables, the positions' minimizers) and probed against each
// table of the group in turn; each table's run-length state
// lives in LDS (wave-uniform part + per-lane accumulators), so the result
// per (table, window) is the same as a single-table scan. Blocks pull
// (window, segment) items from a counter until the work runs out, so the
// grid size (waves per CU) and the segment length are free launch knobs,
// which hipibs_panel_autotune times on a sample of the windows. A group's
// steps (counter reset, panel kernel per <= 16 tables, fold) can be
// captured once into a hipGraph and replayed per chunk.
#include "wave_probe.hpp"

#include <algorithm>
#include <chrono>
#include <vector>

namespace hipibs
{

constexpr int PANEL_BLOCK = 256;
constexpr int PANEL_WAVES = PANEL_BLOCK / 64;

struct alignas(16) WaveState
{
    uint32_t nH, lead, T0, uVar, uDist, seen, pad0, pad1;
};

// The table-independent half of bucketOf(): the mixed key's high word.
__device__ inline uint64_t keyMixHi(uint64_t key)
{
    uint64_t h = (key ^ (key >> 31)) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    return h >> 32;
}

// GROUPED: every table is a minimizer table of the same minimizer length;
// each word's minimizers are computed once and every table is probed with
// waveProbeGrouped (a run of positions sharing a minimizer loads its home
// line once).
template <int S, int MT, bool GROUPED = false>
__global__ __launch_bounds__(PANEL_BLOCK) void panelKernel(
    const TableView *__restrict__ tables, int nT, const uint64_t *__restrict__ seq,
    const uint64_t *__restrict__ exc, uint64_t nBases, const hipibs_window *__restrict__ windows,
    uint32_t nWindows, uint32_t segs, uint32_t k, hipibs_window_stats *__restrict__ stats,
    SegSummary *__restrict__ work, uint32_t *__restrict__ counter, uint32_t tableBase,
    uint64_t statsStride)
{
    __shared__ TableView tv[MT];
    __shared__ WaveState st[PANEL_WAVES][MT];
    __shared__ uint32_t lacc[PANEL_WAVES][MT][64][2];
    __shared__ SegSummary sum[PANEL_WAVES][MT];
    __shared__ __align__(16) uint64_t groupStage[GROUPED ? PANEL_WAVES : 1]
                                                [GROUPED ? GroupStage<S, 1>::kWords : 2];
    __shared__ uint32_t item;
    const int lane = laneId();
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (threadIdx.x < (unsigned)nT)
        tv[threadIdx.x] = tables[threadIdx.x];
    const uint64_t kmask1 = (1ull << k) - 1ull;
    const uint64_t kmask2 = k == 32 ? ~0ull : (1ull << (2 * k)) - 1ull;
    const uint32_t fwdShift = 64 - 2 * k;
    const uint64_t lanesBelow = (1ull << lane) - 1ull;
    const int j = lane & 31;
    const uint32_t nItems = nWindows * segs;
    for (;;)
    {
        if (threadIdx.x == 0)
            item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = item;
        __syncthreads();
        if (it >= nItems)
            break;
        const uint32_t w = it / segs, s = it - w * segs;
        uint64_t g0, g1;
        segmentRange(windows[w], nBases, k, s, segs, g0, g1);
        for (int t = 0; t < nT; ++t)
        {
            if (lane == 0)
                st[wv][t] = WaveState{0, 0, 0, 0, 0, 0, 0, 0};
            lacc[wv][t][lane][0] = 0;
            lacc[wv][t][lane][1] = 0;
        }
        uint32_t nV = 0;
        if (g0 < g1)
        {
            const uint64_t mFirst = g0 >> 6, mLast = (g1 - 1) >> 6;
            const uint64_t nWords = mLast - mFirst + 1;
            const uint64_t per = (nWords + PANEL_WAVES - 1) / PANEL_WAVES;
            const uint64_t mBeg = mFirst + std::min<uint64_t>(nWords, (uint64_t)wv * per);
            const uint64_t mEnd = mFirst + std::min<uint64_t>(nWords, (uint64_t)(wv + 1) * per);
            const uint32_t mlen = GROUPED ? tv[0].mlen : 0u;
            const uint64_t maskm2 = GROUPED ? (1ull << (2 * mlen)) - 1ull : 0ull;
            const int win = (int)(k - mlen + 1); // m-mers per k-mer
            // canonical code and validity of position 64m + lane; GROUPED:
            // the hash of its first m-mer (the top of the forward code, the
            // bottom of the reverse complement)
            auto encodeWord = [&](uint64_t m, uint64_t &canon, bool &valid, uint32_t &L) {
                const uint64_t s0 = seq[2 * m], s1 = seq[2 * m + 1], s2 = seq[2 * m + 2];
                const uint64_t e0 = exc[m], e1 = exc[m + 1];
                const uint64_t a = lane < 32 ? s0 : s1;
                const uint64_t b = lane < 32 ? s1 : s2;
                uint64_t x = j ? ((a >> (2 * j)) | (b << (64 - 2 * j))) : a;
                x &= kmask2;
                const uint64_t ev = lane ? ((e0 >> lane) | (e1 << (64 - lane))) : e0;
                const uint64_t p = (m << 6) + (uint64_t)lane;
                valid = ((ev & kmask1) == 0) && p >= g0 && p < g1;
                const uint64_t rc = ~x & kmask2;
                const uint64_t fwd = reverse2(x) >> fwdShift;
                canon = fwd < rc ? fwd : rc;
                L = 0;
                if constexpr (GROUPED)
                {
                    const uint64_t flo = fwd >> (2 * (k - mlen)), rlo = rc & maskm2;
                    L = mmerHash(flo < rlo ? flo : rlo);
                }
            };
            uint64_t canon = 0;
            bool valid = false;
            uint32_t L = 0;
            if (GROUPED && mBeg < mEnd)
                encodeWord(mBeg, canon, valid, L);
            for (uint64_t m = mBeg; m < mEnd; ++m)
            {
                uint32_t minh = 0;
                uint64_t hi = 0, canonN = 0;
                bool validN = false;
                uint32_t LN = 0;
                if constexpr (GROUPED)
                {
                    // one word behind: the next word's first m-mers end this
                    // word's minimizer windows (word mEnd is read within the
                    // packed arrays' pad words)
                    encodeWord(m + 1, canonN, validN, LN);
                    minh = waveWindowMin(L, LN, win);
                }
                else
                {
                    encodeWord(m, canon, valid, L);
                    hi = keyMixHi(canon);
                }
                const uint64_t V = ballot(valid);
                nV += (uint32_t)__popcll(V);
                for (int t = 0; t < nT; ++t)
                {
                    const TableView T = tv[t];
                    uint64_t H;
                    if constexpr (GROUPED)
                    {
                        const uint64_t code[1] = {canon};
                        const uint32_t mh[1] = {minh};
                        const bool v[1] = {valid};
                        uint64_t Hw[1];
                        waveProbeGrouped<S, 1>(T, code, mh, v, groupStage[wv], Hw);
                        H = Hw[0];
                    }
                    else
                    {
                        const uint32_t home = (uint32_t)((hi * T.nBuckets) >> 32);
                        H = waveProbe<S>(T, canon, home, valid);
                    }
                    const uint64_t Miss = V & ~H;
                    WaveState ws = st[wv][t];
                    if (H)
                    {
                        const int f = __builtin_ctzll(H);
                        const int l = 63 - __builtin_clzll(H);
                        const uint32_t firstGap =
                            ws.T0 + (uint32_t)__popcll(Miss & ((1ull << f) - 1ull));
                        if (ws.seen)
                        {
                            if (firstGap)
                            {
                                ws.uVar += 1;
                                ws.uDist += (uint32_t)gapDistance((int32_t)firstGap, (int32_t)k);
                            }
                        }
                        else
                        {
                            ws.lead = firstGap;
                            ws.seen = 1;
                        }
                        ws.T0 = (uint32_t)__popcll(Miss >> l);
                        if ((H >> lane) & 1ull)
                        {
                            const uint64_t below = H & lanesBelow;
                            if (below)
                            {
                                const int prev = 63 - __builtin_clzll(below);
                                const uint32_t gap = (uint32_t)__popcll(
                                    Miss & lanesBelow & ~((2ull << prev) - 1ull));
                                if (gap)
                                {
                                    lacc[wv][t][lane][0] += 1;
                                    lacc[wv][t][lane][1] +=
                                        (uint32_t)gapDistance((int32_t)gap, (int32_t)k);
                                }
                            }
                        }
                    }
                    else
                        ws.T0 += (uint32_t)__popcll(Miss);
                    ws.nH += (uint32_t)__popcll(H);
                    if (lane == 0)
                        st[wv][t] = ws;
                }
                if constexpr (GROUPED)
                {
                    canon = canonN;
                    valid = validN;
                    L = LN;
                }
            }
        }
        for (int t = 0; t < nT; ++t)
        {
            const uint32_t lv = waveSumU32(lacc[wv][t][lane][0]);
            const uint32_t ld = waveSumU32(lacc[wv][t][lane][1]);
            if (lane == 0)
            {
                const WaveState ws = st[wv][t];
                sum[wv][t] = SegSummary{nV, ws.nH, ws.seen ? ws.lead : nV, ws.T0, ws.uVar + lv,
                                        ws.uDist + ld};
            }
        }
        __syncthreads();
        if (threadIdx.x < (unsigned)nT)
        {
            const int t = (int)threadIdx.x;
            SegSummary acc = sum[0][t];
#pragma unroll
            for (int i = 1; i < PANEL_WAVES; ++i)
                acc = combine(acc, sum[i][t], (int32_t)k);
            const uint64_t tt = tableBase + (uint64_t)t;
            if (segs == 1)
                stats[tt * statsStride + w] = finalizeSummary(acc, (int32_t)k);
            else
                work[(tt * nWindows + w) * segs + s] = acc;
        }
        __syncthreads();
    }
}

__global__ void panelResetKernel(uint32_t *counter, int n)
{
    if (threadIdx.x < (unsigned)n)
        counter[threadIdx.x] = 0;
}

// Item i = t * nWindows + w -> stats[t * statsStride + w].
__global__ void panelFoldKernel(const SegSummary *__restrict__ work, uint32_t nWindows,
                                uint32_t nItems, uint32_t segs, uint32_t k,
                                hipibs_window_stats *__restrict__ stats, uint64_t statsStride)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nItems)
        return;
    const SegSummary *p = work + (uint64_t)i * segs;
    SegSummary acc = p[0];
    for (uint32_t s = 1; s < segs; ++s)
        acc = combine(acc, p[s], (int32_t)k);
    const uint32_t t = i / nWindows, w = i - t * nWindows;
    stats[(uint64_t)t * statsStride + w] = finalizeSummary(acc, (int32_t)k);
}

constexpr int PANEL_MAX_TABLES = 16;

// Scratch layout: [counters: 64 x u32][views: n_tables x TableView][summaries]
static uint64_t panelWorkBytes(uint32_t nWindows, int nTables, uint32_t segs)
{
    uint64_t b = 256 + 256 * (((uint64_t)nTables * sizeof(TableView) + 255) / 256);
    if (segs > 1)
        b += (uint64_t)nWindows * nTables * segs * sizeof(SegSummary);
    return b;
}

static int cuCount(int device)
{
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    return cus;
}

struct PanelLaunch
{
    const hipibs_table *const *tables;
    int nTables;
    const uint64_t *seq, *exc;
    uint64_t nBases;
    const hipibs_window *windows;
    uint32_t nWindows;
    hipibs_window_stats *stats;
    uint32_t segs, wpc;
    char *work;
    uint64_t statsStride; // windows per table row of the statistics
};

// Enqueues the steps of one panel scan on s (graph-capturable: kernels only).
static int enqueuePanel(const PanelLaunch &L, hipStream_t s)
{
    const hipibs_table *t0 = L.tables[0];
    const uint32_t k = t0->k;
    uint32_t *counters = reinterpret_cast<uint32_t *>(L.work);
    auto *views = reinterpret_cast<TableView *>(L.work + 256);
    auto *sums = reinterpret_cast<SegSummary *>(
        L.work + 256 + 256 * (((uint64_t)L.nTables * sizeof(TableView) + 255) / 256));
    const int groups = (L.nTables + PANEL_MAX_TABLES - 1) / PANEL_MAX_TABLES;
    if (groups > 64)
        return setError(HIPIBS_E_ARG, "too many tables for one panel launch (max 1024)");
    hipLaunchKernelGGL(panelResetKernel, dim3(1), dim3(64), 0, s, counters, groups);
    HIPIBS_CHECK(hipGetLastError());
    const int blocks = std::max(1, cuCount(t0->device) * (int)L.wpc / PANEL_WAVES);
    for (int g = 0; g < groups; ++g)
    {
        const int base = g * PANEL_MAX_TABLES;
        const int n = std::min(PANEL_MAX_TABLES, L.nTables - base);
#define PANEL_LAUNCH(SS, MT, GR)                                                                 \
    hipLaunchKernelGGL((panelKernel<SS, MT, GR>), dim3(blocks), dim3(PANEL_BLOCK), 0, s,          \
                       views + base, n, L.seq, L.exc, L.nBases, L.windows, L.nWindows, L.segs, k, \
                       L.stats, sums, counters + g, (uint32_t)base, L.statsStride)
#define PANEL_LAUNCH_MT(SS, GR)                                                                  \
    do                                                                                           \
    {                                                                                            \
        if (n <= 4)                                                                              \
            PANEL_LAUNCH(SS, 4, GR);                                                             \
        else if (n <= 8)                                                                         \
            PANEL_LAUNCH(SS, 8, GR);                                                             \
        else                                                                                     \
            PANEL_LAUNCH(SS, 16, GR);                                                            \
    } while (0)
        const bool grouped = t0->mlen != 0; // checkTables: every table alike
        if (t0->bucketSlots == 8)
        {
            if (grouped)
                PANEL_LAUNCH_MT(8, true);
            else
                PANEL_LAUNCH_MT(8, false);
        }
        else
        {
            if (grouped)
                PANEL_LAUNCH_MT(16, true);
            else
                PANEL_LAUNCH_MT(16, false);
        }
#undef PANEL_LAUNCH_MT
#undef PANEL_LAUNCH
        HIPIBS_CHECK(hipGetLastError());
    }
    if (L.segs > 1)
    {
        const uint32_t items = L.nWindows * (uint32_t)L.nTables;
        hipLaunchKernelGGL(panelFoldKernel, dim3((items + 255) / 256), dim3(256), 0, s, sums,
                           L.nWindows, items, L.segs, k, L.stats, L.statsStride);
        HIPIBS_CHECK(hipGetLastError());
    }
    return 0;
}

static int preparePanel(const hipibs_table *const *tables, int n_tables, const uint64_t *d_seq,
                        const uint64_t *d_exc, uint64_t n_bases, const hipibs_window *d_windows,
                        uint32_t n_windows, hipibs_window_stats *d_stats,
                        const hipibs_scan_config *cfg, void *d_work, uint64_t work_bytes,
                        PanelLaunch &L)
{
    int rc = checkTables(tables, n_tables);
    if (rc)
        return rc;
    if (tables[0]->k > 32)
        return setError(HIPIBS_E_ARG, "panel scans need k <= 32 (scan tables of k > 32 one "
                                      "launch per table with hipibs_scan_windows_device)");
    if (n_windows && (!d_seq || !d_exc || !d_windows || !d_stats))
        return setError(HIPIBS_E_ARG, "null buffer");
    const uint32_t segs = cfg && cfg->segs_per_window ? cfg->segs_per_window : 1;
    const uint32_t wpc = cfg && cfg->waves_per_cu ? cfg->waves_per_cu : 16;
    if (segs > 4096 || wpc > 64)
        return setError(HIPIBS_E_ARG, "bad panel configuration");
    if ((uint64_t)n_windows * segs >= (1ull << 31))
        return setError(HIPIBS_E_ARG, "too many windows for one launch");
    if (!d_work || work_bytes < panelWorkBytes(n_windows, n_tables, segs))
        return setError(HIPIBS_E_ARG, "panel scratch too small (hipibs_panel_work_bytes)");
    std::vector<TableView> v(n_tables);
    for (int i = 0; i < n_tables; ++i)
        v[i] = tables[i]->view();
    HIPIBS_CHECK(hipSetDevice(tables[0]->device));
    // the views are copied before anything is enqueued (and so before a capture)
    HIPIBS_CHECK(hipMemcpy(static_cast<char *>(d_work) + 256, v.data(), v.size() * sizeof(TableView),
                           hipMemcpyHostToDevice));
    L = PanelLaunch{tables,    n_tables,  d_seq, d_exc, n_bases, d_windows,
                    n_windows, d_stats,   segs,  wpc,   static_cast<char *>(d_work), n_windows};
    return 0;
}

} // namespace hipibs

using namespace hipibs;

struct hipibs_graph
{
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
};

extern "C" uint64_t hipibs_panel_work_bytes(uint32_t n_windows, int n_tables,
                                            uint32_t segs_per_window)
{
    return panelWorkBytes(n_windows, n_tables, segs_per_window ? segs_per_window : 1);
}

extern "C" int hipibs_panel_scan_device(const hipibs_table *const *tables, int n_tables,
                                        const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                                        uint64_t n_bases, const hipibs_window *d_windows,
                                        uint32_t n_windows, hipibs_window_stats *d_stats,
                                        const hipibs_scan_config *cfg, void *d_work,
                                        uint64_t work_bytes, void *stream)
{
    PanelLaunch L;
    int rc = preparePanel(tables, n_tables, d_seq_words, d_exc_words, n_bases, d_windows,
                          n_windows, d_stats, cfg, d_work, work_bytes, L);
    if (rc || n_windows == 0)
        return rc;
    return enqueuePanel(L, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int hipibs_panel_graph_create(const hipibs_table *const *tables, int n_tables,
                                         const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                                         uint64_t n_bases, const hipibs_window *d_windows,
                                         uint32_t n_windows, hipibs_window_stats *d_stats,
                                         const hipibs_scan_config *cfg, uint32_t n_chunks,
                                         void *d_work, uint64_t work_bytes, hipibs_graph **out)
{
    if (!out)
        return setError(HIPIBS_E_ARG, "null argument");
    *out = nullptr;
    n_chunks = std::max<uint32_t>(1, std::min<uint32_t>(n_chunks, std::max<uint32_t>(1, n_windows)));
    PanelLaunch L;
    int rc = preparePanel(tables, n_tables, d_seq_words, d_exc_words, n_bases, d_windows,
                          n_windows, d_stats, cfg, d_work, work_bytes, L);
    if (rc)
        return rc;
    hipStream_t s;
    HIPIBS_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    auto *g = new hipibs_graph();
    g->device = tables[0]->device;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess)
    {
        // chunk c: windows [c*per, (c+1)*per), written into the rows of the
        // whole table-major result (stride n_windows)
        const uint32_t per = (n_windows + n_chunks - 1) / n_chunks;
        for (uint32_t c = 0; c < n_chunks && !rc; ++c)
        {
            const uint32_t lo = c * per, hi = std::min(n_windows, lo + per);
            if (lo >= hi)
                break;
            PanelLaunch C = L;
            C.windows = L.windows + lo;
            C.nWindows = hi - lo;
            C.stats = L.stats + lo;
            rc = enqueuePanel(C, s);
        }
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s, &graph);
        if (!rc && e2 != hipSuccess)
            rc = hipFail(e2, "hipStreamEndCapture");
        g->graph = graph;
        if (!rc && hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess)
            rc = setError(HIPIBS_E_HIP, "hipGraphInstantiate failed");
    }
    else
        rc = hipFail(e, "hipStreamBeginCapture");
    (void)hipStreamDestroy(s);
    if (rc)
    {
        if (g->exec)
            (void)hipGraphExecDestroy(g->exec);
        if (g->graph)
            (void)hipGraphDestroy(g->graph);
        delete g;
        return rc;
    }
    *out = g;
    return 0;
}

extern "C" int hipibs_graph_launch(hipibs_graph *g, void *stream)
{
    if (!g || !g->exec)
        return setError(HIPIBS_E_ARG, "null graph");
    HIPIBS_CHECK(hipSetDevice(g->device));
    HIPIBS_CHECK(hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream)));
    return 0;
}

extern "C" int hipibs_graph_free(hipibs_graph *g)
{
    if (!g)
        return 0;
    if (g->exec)
        (void)hipGraphExecDestroy(g->exec);
    if (g->graph)
        (void)hipGraphDestroy(g->graph);
    delete g;
    return 0;
}

// Times candidate (waves per CU, segments per window) pairs on the first
// `sample_windows` windows and returns the fastest.
extern "C" int hipibs_panel_autotune(const hipibs_table *const *tables, int n_tables,
                                     const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                                     uint64_t n_bases, const hipibs_window *d_windows,
                                     uint32_t n_windows, uint32_t sample_windows,
                                     hipibs_window_stats *d_stats, void *d_work,
                                     uint64_t work_bytes, hipibs_scan_config *best,
                                     double *best_ms)
{
    if (!best)
        return setError(HIPIBS_E_ARG, "null argument");
    const uint32_t n = std::max<uint32_t>(1, std::min(n_windows, sample_windows ? sample_windows : 4096));
    // occupancy from 8 to 32 waves per CU; one segment per window or four
    const uint32_t wpcs[] = {8, 16, 24, 32};
    const uint32_t segss[] = {1, 4};
    hipEvent_t e0, e1;
    HIPIBS_CHECK(hipSetDevice(tables && n_tables > 0 && tables[0] ? tables[0]->device : 0));
    HIPIBS_CHECK(hipEventCreate(&e0));
    HIPIBS_CHECK(hipEventCreate(&e1));
    hipStream_t s;
    HIPIBS_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    double bestT = 1e30;
    hipibs_scan_config bestC{1, 16, {0, 0}};
    int rc = 0;
    for (uint32_t wpc : wpcs)
        for (uint32_t segs : segss)
        {
            if (panelWorkBytes(n, n_tables, segs) > work_bytes)
                continue;
            hipibs_scan_config c{segs, wpc, {0, 0}};
            double t = 0;
            for (int rep = 0; rep < 2 && !rc; ++rep)
            {
                (void)hipEventRecord(e0, s);
                rc = hipibs_panel_scan_device(tables, n_tables, d_seq_words, d_exc_words, n_bases,
                                              d_windows, n, d_stats, &c, d_work, work_bytes, s);
                (void)hipEventRecord(e1, s);
                (void)hipEventSynchronize(e1);
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (rep > 0)
                    t += ms;
            }
            if (rc)
                break;
            if (t < bestT)
            {
                bestT = t;
                bestC = c;
            }
        }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc)
        return rc;
    *best = bestC;
    if (best_ms)
        *best_ms = bestT;
    return 0;
}
