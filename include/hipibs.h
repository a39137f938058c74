/*
 * hipibs.h -- C ABI of the MI355X k-mer IBS scan library (libhipibs.so).
 *
 * Plain C types only: pointers, sizes, status codes. Every function returns
 * 0 on success and a negative HIPIBS_E_* code on failure; the message of the
 * last failure on the calling thread is hipibs_last_error(). Pointers named
 * d_* are device pointers on the table's device; h_* are host pointers.
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *
 * Data formats shared by every entry point:
 *  - k-mer key: canonical 2-bit packed code, A=0 C=1 G=2 T=3, first base in
 *    the most significant used bits (src/KmerCodec.hpp:18-60). k <= 32: one
 *    uint64_t per key. 33 <= k <= 64 (the reference's string path,
 *    src/KmerDatabase.hpp:104-123, :675-709): two uint64_t per key, the high
 *    word first, so every key array, code array and count argument below
 *    holds 2n words for n keys when the table's (or database's) k is above
 *    32. Tables of k > 32 have 128 B buckets of 8 keys, and their builds
 *    skip keys whose high word is all ones (hipibs_encode_device's mark of
 *    an invalid position); panel-kernel launches (hipibs_panel_*) and
 *    minimizer tables take k <= 32.
 *  - packed bases: base b of a stream lives at bits 2*(b%32) of 64-bit word
 *    b/32 (code of a non-ACGT base is 0); its exception bit (set when the
 *    byte is not one of ACGTacgt) is bit b%64 of word b/64 of a separate
 *    array (src/gpu/EncodeProbeCore.hpp:40-50 uses the same two arrays).
 *    Both arrays must carry HIPIBS_PACK_PAD_WORDS readable words past the
 *    last base.
 *  - window: a contiguous base range [base, base+n_bases) of the packed
 *    stream; its k-mer positions are base .. base+n_bases-k (none when
 *    n_bases < k), as cut by chunkSequences (src/Utils.hpp:296-327).
 *  - window statistics: int32 total_kmers, observed_kmers, variations,
 *    kmer_distance, with the run-length rule of
 *    Window::addVariationOriginal (src/Window.hpp:29-36) applied to the
 *    valid-position stream (src/KmerDatabase.hpp:546-673).
 */
#ifndef HIPIBS_H
#define HIPIBS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIPIBS_OK 0
#define HIPIBS_E_ARG (-1)      /* invalid argument */
#define HIPIBS_E_HIP (-2)      /* HIP runtime error */
#define HIPIBS_E_IO (-3)       /* file missing, unreadable or malformed */
#define HIPIBS_E_NOMEM (-4)    /* device or host allocation failed / budget */
#define HIPIBS_E_FULL (-5)     /* table capacity exhausted during a build */
#define HIPIBS_E_NODEV (-6)    /* no usable GPU */

#define HIPIBS_PACK_PAD_WORDS 8

typedef struct hipibs_table hipibs_table;
typedef struct hipibs_scan hipibs_scan;

/* Table construction options (zero-initialised fields take the defaults).
 * Tables scanned in one launch must share k, bucket width and minimizer
 * length. */
typedef struct hipibs_table_opts {
    double load_factor;     /* keys / slots, default 0.6 */
    uint32_t bucket_slots;  /* 16 (128 B bucket, one memory request; default) or 8 (64 B) */
    int device;             /* HIP device ordinal, default 0 */
    uint64_t hbm_budget;    /* max table bytes, 0 = no limit */
    uint32_t load_threads;  /* host threads of a KMC read, 0 = all CPUs */
    uint32_t minimizer_len; /* 0 (default): home bucket from a hash of the whole k-mer;
                               5..min(31, k): a minimizer table -- keys sharing a minimizer of
                               that length share buckets, so consecutive positions of a
                               reference share memory requests */
    uint32_t minimizer_cap; /* minimizer tables: a minimizer with more keys than this
                               (0 = 32) leaves one marker in the minimizer-ordered
                               buckets and its keys go to a whole-k-mer part (repeat
                               families would otherwise form long probe chains) */
} hipibs_table_opts;

typedef struct hipibs_window {
    uint64_t base;     /* first base of the window in the packed stream */
    uint32_t n_bases;  /* window length in bases */
    uint32_t reserved; /* must be 0 */
} hipibs_window;

typedef struct hipibs_window_stats {
    int32_t total_kmers;
    int32_t observed_kmers;
    int32_t variations;
    int32_t kmer_distance;
} hipibs_window_stats;

/* Header of a KMC database (<prefix>.kmc_pre), as KmcRaIndex::Header. */
typedef struct hipibs_kmc_info {
    uint32_t kmc_version;   /* 0 (KMC1) or 0x200 (KMC2) */
    uint32_t kmer_length;
    uint32_t mode;
    uint32_t counter_size;
    uint32_t lut_prefix_length;
    uint32_t signature_len; /* 0 for KMC1 */
    uint32_t min_count;
    uint32_t max_count;
    uint64_t total_kmers;
} hipibs_kmc_info;

/* Host slot of a streaming scan: pinned buffers the caller fills before
 * hipibs_scan_submit and reads after hipibs_scan_wait. */
typedef struct hipibs_host_slot {
    uint64_t *seq_words;           /* packed bases (capacity seq_cap words) */
    uint64_t *exc_words;           /* exception mask (capacity exc_cap words) */
    hipibs_window *windows;        /* windows, bases relative to this slot */
    hipibs_window_stats *stats;    /* n_tables x n_windows results, table-major */
    uint64_t seq_cap, exc_cap;     /* capacities in words (pad included) */
    uint32_t window_cap;
    uint32_t n_tables;
} hipibs_host_slot;

/* ---- errors and devices ------------------------------------------------ */

/* Message of the last failure on this thread ("" if none). */
const char *hipibs_last_error(void);

/* Replaces DeviceKmerIndex::deviceCount / deviceCountError
 * (src/gpu/DeviceKmerIndex.cuh:30-40): number of usable GPUs in *n; on
 * failure *n = 0 and the reason is copied to why (if non-NULL). */
int hipibs_device_count(int *n, char *why, size_t why_len);

/* Replaces DeviceKmerIndex::freeMemory (src/gpu/DeviceKmerIndex.cu:354-367). */
int hipibs_device_mem(int device, uint64_t *free_bytes, uint64_t *total_bytes);

/* ---- KMC databases (host) ----------------------------------------------- */

/* Replaces KmcRaIndex::readHeader (src/gpu/KmcRaIndex.hpp:104-110). */
int hipibs_kmc_header(const char *prefix, hipibs_kmc_info *out);

/* Replaces KmcRaIndex(prefix) + forEachKmer (src/gpu/KmcRaIndex.hpp:124-164):
 * every k-mer of the database as its canonical code, in on-disk order.
 * Writes at most cap keys; *n receives total_kmers. */
int hipibs_kmc_read_keys(const char *prefix, uint64_t *h_keys, uint64_t cap,
                         uint64_t *n, uint32_t threads);

/* Writes a KMC database readable by CKMCFile (KMC/kmc_api/kmc_file.cpp:178-369)
 * and KmcRaIndex: keys must be distinct canonical codes (any order),
 * counts (nullable = all 1) are clamped to the counter width. version is 0
 * (KMC1) or 0x200 (KMC2, sig_len 5..11); (k - p) % 4 == 0 is required.
 * No counterpart in the reference (it reads databases written by KMC 3.2.4). */
int hipibs_kmc_write(const char *prefix, const uint64_t *h_keys, const uint32_t *h_counts,
                     uint64_t n, uint32_t k, uint32_t p, uint32_t sig_len,
                     uint32_t counter_size, uint32_t version, uint32_t min_count,
                     uint32_t max_count);

/* ---- packing (host) ------------------------------------------------------ */

/* Packs n bases of text into the 2-bit + exception layout, starting at base
 * offset `at` of the destination arrays. Words that also hold bases before
 * `at` are OR-ed into, so their bits for bases at..  must be zero (true for
 * consecutive calls into a zeroed buffer); all other words are overwritten.
 * The (tail) words past the last base keep unspecified bits for later bases. Same base table as KmerCodec::baseCode
 * (src/KmerCodec.hpp:105-118) and packWindow (src/gpu/EncodeProbeCore.hpp:157-186). */
int hipibs_pack_bases(const char *h_text, uint64_t n, uint64_t at, uint64_t *h_seq_words,
                      uint64_t *h_exc_words);

/* ---- device tables ------------------------------------------------------- */

/* Replaces KmcRaIndex(prefix) + DeviceKmerIndex::buildShards
 * (src/gpu/KmcRaIndex.hpp:124-135, src/gpu/DeviceKmerIndex.cu:144-183):
 * reads the database and builds a device-resident bucketised open-addressing
 * table of its canonical keys (the raw suffix records are read on host
 * threads into pinned buffers and decoded into keys on the device). */
int hipibs_table_build_kmc(const char *prefix, const hipibs_table_opts *opts,
                           hipibs_table **out);

/* Builds a table from host keys (duplicates are stored once). The analogue
 * of DeviceHashIndex's build from an enumeration (src/gpu/DeviceHashIndex.cu:121-153). */
int hipibs_table_build_keys(const uint64_t *h_keys, uint64_t n, uint32_t k,
                            const hipibs_table_opts *opts, hipibs_table **out);

/* Same from device keys on `stream` (synchronises before returning). A
 * whole-k-mer table of k <= 32 from 2^20 keys or more is built by sorting
 * the keys' hashes by table slice and filling each 32 KB slice in LDS (the
 * same keys in the same probe order as a key-by-key insert); the sort uses
 * the table's own slots and one buffer of about 9 B per key, which the
 * process keeps for its next build (see hipibs_counter_release_cache). */
int hipibs_table_build_keys_device(const uint64_t *d_keys, uint64_t n, uint32_t k,
                                   const hipibs_table_opts *opts, hipibs_table **out,
                                   void *stream);

/* Replaces DeviceKmerIndex's destructor. NULL is accepted. */
int hipibs_table_free(hipibs_table *t);

/* Replaces DeviceKmerIndex::footprintBytes / view (src/gpu/DeviceKmerIndex.cuh:86-110). */
int hipibs_table_info(const hipibs_table *t, uint64_t *n_keys, uint64_t *bytes,
                      uint64_t *n_buckets, uint32_t *k, uint32_t *bucket_slots, int *device);

/* Minimizer length the table's home buckets use (0 = whole-k-mer hash).
 * No counterpart in the reference (its index has no such choice). */
int hipibs_table_minimizer_len(const hipibs_table *t, uint32_t *m);

/* Replaces DeviceKmerIndex::probe(d_q, d_valid, n, d_hits, stream)
 * (src/gpu/DeviceKmerIndex.cuh:60-66): d_hits[i] = 1 if d_valid[i] (NULL =
 * all valid) and d_codes[i] is in the table, else 0. Asynchronous. */
int hipibs_table_probe(const hipibs_table *t, const uint64_t *d_codes, const uint8_t *d_valid,
                       uint64_t n, uint8_t *d_hits, void *stream);

/* Replaces DeviceKmerIndex::probeCountHost (src/gpu/DeviceKmerIndex.cu:264-306):
 * number of host keys present in the table. Thread-safe; synchronous. */
int hipibs_table_probe_count(const hipibs_table *t, const uint64_t *h_keys, uint64_t n,
                             uint64_t *out);

/* Device-key variant of hipibs_table_probe_count; the count is written to
 * d_count (a device uint64, accumulated: the caller zeroes it). Asynchronous. */
int hipibs_table_probe_count_device(const hipibs_table *t, const uint64_t *d_keys, uint64_t n,
                                    uint64_t *d_count, void *stream);

/* Replaces the counting half of KDBIntersect::gpuIntersect, enumerateAndCount
 * + probeCountHost (src/KDBIntersect.cpp:133-238, src/gpu/ParallelEnum.hpp:181-362):
 * number of the KMC database's k-mers present in the table. The records are
 * read raw on up to load_threads host threads (0 = all usable CPUs) and
 * decoded on the device; no host key array. The database's k must equal the
 * table's. Synchronous. */
int hipibs_table_probe_count_kmc(const hipibs_table *t, const char *kmc_prefix,
                                 unsigned load_threads, uint64_t *out);

/* ---- k-mer counting from reads ------------------------------------------- */

/* Replaces the `kmc -k<K> -ci<A> -cx<B> -cs<C>` step that produces the
 * reference's sample databases (README.md:95, scripts/make_test_corpus.sh:123;
 * KMC 3.2.4 is an external tool, not part of the reference's sources). The
 * canonical k-mers of the reads are counted in a device hash table: a k-mer
 * never spans two records and is skipped if it holds a byte other than
 * ACGTacgt. The result becomes a device table (no files), host arrays, or a
 * KMC database. A counter is used by one thread at a time. */
typedef struct hipibs_counter hipibs_counter;

typedef struct hipibs_count_stats {
    uint64_t reads;       /* records (add_file) and add_text calls */
    uint64_t kmers;       /* k-mers counted: positions with k valid bases */
    uint64_t distinct;    /* distinct canonical k-mers */
    uint64_t table_bytes; /* device bytes of the counting table */
    uint64_t launches;    /* counting kernel launches */
    double kernel_ms;     /* their device time (HIP events around each launch on the
                             counter's stream) */
    uint64_t flushes;     /* k <= 32: sorted-slice counting rounds */
    uint64_t spilled;     /* k <= 32: k-mers whose probe left their 512-bucket slice and
                             were counted in the global table */
    double stage_ms[4];   /* kernel_ms by stage: hashing (k > 32: counting), grouping by
                             slice, slice counts, spills */
} hipibs_count_stats;

typedef struct hipibs_count_filter {
    uint64_t kept;  /* distinct k-mers with min_count <= count <= max_count */
    uint64_t below; /* count < min_count */
    uint64_t above; /* count > max_count */
} hipibs_count_filter;

/* k in 1..64 (for k > 32, key arrays hold two words per key, high first).
 * expected_keys (0 = unknown) sizes the first table; the table
 * grows as needed. chunk_bytes (0 = 128 MiB, at most 1 GiB) is the text
 * moved to the device at a time: two pinned host buffers and one device
 * buffer of that size, allocated when host text first arrives. */
int hipibs_counter_create(uint32_t k, int device, uint64_t expected_keys, uint64_t chunk_bytes,
                          hipibs_counter **out);
/* Counts a FASTA or FASTQ file, plain or gzip (decided by its first byte);
 * asynchronous with respect to the device. After a failure the counter only
 * accepts hipibs_counter_free. */
int hipibs_counter_add_file(hipibs_counter *c, const char *path);
/* Counts n bytes of text as one record: any byte other than ACGTacgt (a
 * '\n', say) ends a run of k-mers, so several reads may be passed at once. */
int hipibs_counter_add_text(hipibs_counter *c, const char *h_text, uint64_t n);
/* Counts n bytes of device-resident text (16 B aligned, on the counter's
 * device) as one record, by the rules of hipibs_counter_add_text; returns
 * when the device has counted it. For callers whose reads are already in
 * HBM. */
int hipibs_counter_add_text_device(hipibs_counter *c, const char *d_text, uint64_t n);
/* Waits for the device and reports totals. */
int hipibs_counter_stats(hipibs_counter *c, hipibs_count_stats *out);
/* The k-mers with min_count <= count <= max_count, in no particular order,
 * into h_keys (at most cap) and, if non-NULL, their counts clamped to
 * counter_max into h_counts. */
int hipibs_counter_keys(hipibs_counter *c, uint32_t min_count, uint32_t max_count,
                        uint32_t counter_max, uint64_t *h_keys, uint32_t *h_counts,
                        uint64_t cap, hipibs_count_filter *filt);
/* A device table of the kept k-mers, built on the device from the counting
 * table (opts as hipibs_table_build_kmc; NULL = defaults on the counter's
 * device). filt may be NULL. */
int hipibs_counter_build_table(hipibs_counter *c, uint32_t min_count, uint32_t max_count,
                               const hipibs_table_opts *opts, hipibs_table **out,
                               hipibs_count_filter *filt);
/* Writes the kept k-mers as a KMC database (<prefix>.kmc_pre/.kmc_suf; the
 * KMC2 layout with 9-base signatures, counters of the width counter_max
 * needs; header min/max count = min_count / max_count), readable by CKMCFile,
 * KmcRaIndex and hipibs_table_build_kmc. filt may be NULL. */
int hipibs_counter_write_kmc(hipibs_counter *c, const char *prefix, uint32_t min_count,
                             uint32_t max_count, uint32_t counter_max, hipibs_count_filter *filt);
/* Frees the counter. Its table, and for k <= 32 its two buffers of hashed
 * k-mers (8 B per k-mer of one counting round, up to 16 chunks' worth), are
 * kept for the process's next counter on the same device: the largest table
 * and the largest pair freed so far, one of each per process. */
void hipibs_counter_free(hipibs_counter *c);
/* Frees the buffers kept for the process's next counter or table build:
 * a counter's table, the pair of hashed k-mer buffers, a table build's
 * sort buffer and sort state (no counter needs to exist). A table
 * allocation that fails frees them and tries once more. */
void hipibs_counter_release_cache(void);

/* ---- the fused scan (device-resident inputs) ------------------------------ */

/* Fused encode + canonicalise + probe + per-window run-length reduction over
 * device-resident packed bases: for every table i and window j,
 * d_stats[i*n_windows + j] receives the window's statistics against table i
 * (the device half of EncodeProbeEngine::submit + BitStats::run +
 * Window::addVariationOriginal: src/gpu/EncodeProbe.cu:152-193,
 * src/gpu/EncodeProbeCore.hpp:338-370, src/Window.hpp:29-36). All tables
 * must share k, bucket width and device. Every window must lie inside
 * [0, n_bases) and be at most max_window_bases long. segs_per_window splits
 * each window's positions into that many block work units (0 = chosen by
 * the library); when it resolves to more than 1, d_work must hold
 * hipibs_scan_work_bytes() bytes of scratch. Asynchronous on `stream`. */
int hipibs_scan_windows_device(const hipibs_table *const *tables, int n_tables,
                               const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                               uint64_t n_bases, const hipibs_window *d_windows,
                               uint32_t n_windows, uint32_t max_window_bases,
                               hipibs_window_stats *d_stats, uint32_t segs_per_window,
                               void *d_work, uint64_t work_bytes, void *stream);

/* hipibs_scan_windows_device for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. It calls no allocating
 * routine: its scratch (the launch's table views and segment summaries) is
 * a library-held device array of 16 MiB per GPU, kept until the process
 * ends. Limits: at most 64 tables, and n_windows x n_tables x (resolved
 * segments per window, when above 1) x 24 B at most
 * hipibs_scan_noalloc_max_work_bytes(); larger calls fail with
 * HIPIBS_E_ARG. Every call on one GPU shares that array, so calls that
 * overlap on one GPU (on different streams, or from several threads) race:
 * serialise them. Asynchronous on `stream`. */
int hipibs_scan_windows_device_noalloc(const hipibs_table *const *tables, int n_tables,
                                       const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                                       uint64_t n_bases, const hipibs_window *d_windows,
                                       uint32_t n_windows, uint32_t max_window_bases,
                                       hipibs_window_stats *d_stats, uint32_t segs_per_window,
                                       void *stream);

/* Segment-summary bytes hipibs_scan_windows_device_noalloc can hold. */
uint64_t hipibs_scan_noalloc_max_work_bytes(void);

/* Scratch bytes of hipibs_scan_windows_device for these arguments (0 when
 * the resolved split is one segment per window); *segs_resolved (nullable)
 * receives the split the library will use. */
uint64_t hipibs_scan_work_bytes(uint32_t n_windows, int n_tables, uint32_t max_window_bases,
                                uint32_t k, uint32_t segs_per_window, uint32_t *segs_resolved);

/* ---- panel scans (many samples, one reference pass) ----------------------- */

/* Launch knobs of the persistent panel scan. Zero fields take defaults. */
typedef struct hipibs_scan_config {
    uint32_t segs_per_window; /* block work units per window (default 1) */
    uint32_t waves_per_cu;    /* persistent grid size (default 16) */
    uint32_t reserved[2];
} hipibs_scan_config;

/* Scratch bytes of a panel scan of n_windows windows against n_tables. */
uint64_t hipibs_panel_work_bytes(uint32_t n_windows, int n_tables, uint32_t segs_per_window);

/* The panel scan (KmerPanel's device pass, src/KmerPanel.hpp:400-439, with
 * encodeProbeWordMulti's one-encode-many-probes shape,
 * src/gpu/EncodeProbeCore.hpp:291-328): one persistent launch per 16 tables,
 * every position encoded once and probed against each table; results as
 * hipibs_scan_windows_device (table-major). Tables need whole-k-mer
 * buckets (minimizer_len 0). Asynchronous on `stream`. */
int hipibs_panel_scan_device(const hipibs_table *const *tables, int n_tables,
                             const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                             uint64_t n_bases, const hipibs_window *d_windows, uint32_t n_windows,
                             hipibs_window_stats *d_stats, const hipibs_scan_config *cfg,
                             void *d_work, uint64_t work_bytes, void *stream);

/* Times (waves per CU, segments per window) candidates on the first
 * sample_windows windows (0 = 4096) and returns the fastest in *best. No
 * counterpart in the reference (its launch shape is fixed). Synchronous. */
int hipibs_panel_autotune(const hipibs_table *const *tables, int n_tables,
                          const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                          uint64_t n_bases, const hipibs_window *d_windows, uint32_t n_windows,
                          uint32_t sample_windows, hipibs_window_stats *d_stats, void *d_work,
                          uint64_t work_bytes, hipibs_scan_config *best, double *best_ms);

typedef struct hipibs_graph hipibs_graph;

/* Captures a panel scan split into n_chunks window ranges (per chunk: counter
 * reset, panel kernel(s), fold) into a hipGraph; hipibs_graph_launch replays
 * it. The captured pointers must stay valid. No counterpart in the reference. */
int hipibs_panel_graph_create(const hipibs_table *const *tables, int n_tables,
                              const uint64_t *d_seq_words, const uint64_t *d_exc_words,
                              uint64_t n_bases, const hipibs_window *d_windows, uint32_t n_windows,
                              hipibs_window_stats *d_stats, const hipibs_scan_config *cfg,
                              uint32_t n_chunks, void *d_work, uint64_t work_bytes,
                              hipibs_graph **out);
int hipibs_graph_launch(hipibs_graph *g, void *stream);
int hipibs_graph_free(hipibs_graph *g);

/* Per-position hit mask over positions [pos0, pos0+n_pos) of a packed
 * stream: bit (p-pos0)%32 of word (p-pos0)/32 of d_hit_words is set when
 * position p is a valid k-mer of the stream's contig (all k bases ACGT and
 * p+k <= contig_end) found in the table. The device half of the mapper
 * (src/KmerDatabase.hpp:1544-1581). Asynchronous. */
int hipibs_hitmask_device(const hipibs_table *t, const uint64_t *d_seq_words,
                          const uint64_t *d_exc_words, uint64_t pos0, uint64_t n_pos,
                          uint64_t contig_end, uint32_t *d_hit_words, void *stream);

/* Canonical code of every position p in [pos0, pos0+n_pos) of a packed
 * stream: d_codes[p-pos0] = canonical code when the k bases at p are all
 * ACGT and p+k <= contig_end, else all ones; d_valid (nullable) gets 1/0.
 * k <= 64; for k > 32 d_codes holds two words per position (high first).
 * The encode half of the fused kernel on its own, as
 * KmerCodec::encodePositions (src/KmerCodec.hpp:67-100). Asynchronous. */
int hipibs_encode_device(const uint64_t *d_seq_words, const uint64_t *d_exc_words, uint64_t pos0,
                         uint64_t n_pos, uint64_t contig_end, uint32_t k, uint64_t *d_codes,
                         uint8_t *d_valid, void *stream);

/* Per-base coverage of one contig from its hit mask (positions 0..n_pos-1,
 * n_pos = n_bases-k+1, or 0): d_cov[i] = number of set positions p with
 * p <= i <= p+k-1, i < n_bases (src/KmerDatabase.hpp:756-802). Asynchronous. */
int hipibs_coverage_device(const uint32_t *d_hit_words, uint64_t n_pos, uint64_t n_bases,
                           uint32_t k, uint8_t *d_cov, void *stream);

/* The mapper's coverage line of one contig, written on the device: the
 * decimal values of d_cov[0..n_bases) joined by ',' (no newline), as
 * KmerDatabase::serializeCoverage formats them with to_chars
 * (src/KmerDatabase.hpp:756-772). *text_len receives the length; with
 * d_text NULL only the length is computed, else the text is written when it
 * fits text_cap (at most 4 * n_bases bytes). Synchronous on stream. */
int hipibs_coverage_text_device(const uint8_t *d_cov, uint64_t n_bases, char *d_text,
                                uint64_t text_cap, uint64_t *text_len, void *stream);

/* Receives consecutive pieces of a text; returns 0 to continue, non-zero to
 * stop with an error. */
typedef int (*hipibs_sink_fn)(void *user, const char *data, uint64_t n);

/* The same text handed to `sink` in order, in pieces of at most 64 MiB
 * copied through pinned double buffers (the copy of piece i+1 overlaps the
 * sink's work on piece i): the device half of the mapper's serialise +
 * write (src/KmerDatabase.hpp:756-772, src/FastIBSMapper.cpp:20-171).
 * Synchronous on stream. */
int hipibs_coverage_text_stream(const uint8_t *d_cov, uint64_t n_bases, hipibs_sink_fn sink,
                                void *user, void *stream);

/* Timing of the last scan kernel launched by this thread, in ms, measured
 * with HIP events on the launch stream (valid after the stream completes;
 * negative when profiling is off). hipibs_set_profiling(1) turns it on. */
int hipibs_set_profiling(int on);
double hipibs_last_kernel_ms(void);

/* ---- serving: resident tables, continuous batching ------------------------- */

typedef struct hipibs_serve hipibs_serve;

typedef struct hipibs_serve_counters {
    uint64_t submitted, completed, launches, windows;
    uint64_t admissions, evictions;   /* tables built / freed for memory */
    uint64_t resident_bytes, resident_tables;
    uint64_t build_us;                /* summed wall time of table builds */
    uint64_t evict_us;                /* summed wall time of eviction frees */
    uint64_t host_bytes;              /* the engine's host copies of sample keys */
} hipibs_serve_counters;

/* A serving engine on `device` for k-mers of length k: sample tables are
 * built (with topts) on demand from the engine's host copy of each
 * registered sample's keys (or its KMC database) and kept resident
 * within hbm_budget bytes; idle tables are evicted least requested first
 * (request counts halved every 4096 requests), and only for a sample
 * requested at least as often unless its request has waited 50 ms; up to
 * max_inflight launches of at most max_windows_per_launch windows run at
 * once, each batching windows of any pending requests.
 * No counterpart in the reference (it has no serving mode). */
int hipibs_serve_create(int device, uint32_t k, uint64_t hbm_budget,
                        const hipibs_table_opts *topts, int max_inflight,
                        uint32_t max_windows_per_launch, hipibs_serve **out);

/* The packed reference (words as hipibs_pack_bases, pad included), copied
 * to the device; set once, before samples are added. */
int hipibs_serve_set_reference(hipibs_serve *s, const uint64_t *h_seq_words,
                               const uint64_t *h_exc_words, uint64_t n_bases);

/* Registers a sample's canonical keys (n_keys codes of 2k bits, two words
 * each for k > 32); preload != 0 builds its table now. The engine keeps
 * its own compact copy in host memory (pinned where the system grants
 * it): each code's top <= 22 bits are implied by its position among keys
 * grouped by them, the rest kept as whole bytes -- 4 B per key at k = 25,
 * 5 B at k = 31 -- plus 4 B per prefix value; admissions upload that copy
 * and decode it on the device. The caller's array may be freed on return.
 * Registration uses device memory for the keys and their suffixes while
 * it runs. All samples are registered before the first request; fewer
 * than 2^32 keys per sample. */
int hipibs_serve_add_sample(hipibs_serve *s, uint32_t sample_id, const uint64_t *h_keys,
                            uint64_t n_keys, int preload);

/* hipibs_serve_add_sample from keys in device memory (complete before the
 * call); a preloaded table is built from them directly. */
int hipibs_serve_add_sample_device(hipibs_serve *s, uint32_t sample_id, const uint64_t *d_keys,
                                   uint64_t n_keys, int preload);

/* Registers a sample by its KMC database (the prefix of .kmc_pre/.kmc_suf,
 * which must stay readable): admissions stream its raw records through the
 * device decode, so the host keeps no key array. Same rules as
 * hipibs_serve_add_sample; the database's k must be the engine's. */
int hipibs_serve_add_sample_kmc(hipibs_serve *s, uint32_t sample_id, const char *kmc_prefix,
                                int preload);

/* Queues a scan of bases [base, base+n_bases) of the reference (one
 * contig's range) in windows of W (chunkSequences geometry within the
 * range) against one sample. Returns at once. */
int hipibs_serve_submit(hipibs_serve *s, uint64_t request_id, uint32_t sample_id, uint64_t base,
                        uint64_t n_bases, uint32_t W);

/* Ids of completed, not yet fetched requests (waits up to timeout_ms for
 * the first one). */
int hipibs_serve_poll(hipibs_serve *s, uint64_t *ids, uint32_t max, int timeout_ms,
                      uint32_t *n_out);

/* A completed request's per-window statistics (up to cap), window count,
 * submit-to-completion latency and submit-to-first-launch queueing time in
 * ms; the request is forgotten afterwards. With out == NULL only the sizes
 * and times are reported and the request is kept. */
int hipibs_serve_fetch(hipibs_serve *s, uint64_t request_id, hipibs_window_stats *out,
                       uint32_t cap, uint32_t *n_windows, double *latency_ms, double *queue_ms);

int hipibs_serve_get_counters(hipibs_serve *s, hipibs_serve_counters *out);

/* Stops the worker (in-flight launches finish), frees every table. */
int hipibs_serve_free(hipibs_serve *s);

/* ---- device memory for FFI callers without a HIP binding ------------------
 * No counterpart in the reference (its callers are CUDA C++). Synchronous. */
int hipibs_dev_alloc(int device, uint64_t bytes, void **out);
int hipibs_dev_free(void *p);
/* kind: 1 host->device, 2 device->host, 3 device->device. */
int hipibs_memcpy(void *dst, const void *src, uint64_t bytes, int kind);
int hipibs_dev_memset(void *dst, int value, uint64_t bytes);
int hipibs_dev_synchronize(int device);

/* ---- streaming scan engine (host buffers, pinned, double-buffered) -------- */

/* Replaces the EncodeProbeEngine constructors + caps + host(slot)
 * (src/gpu/EncodeProbe.cuh:46-120): `depth` slots of pinned host and device
 * buffers sized for batch_bases bases and max_windows windows against
 * n_tables tables (all with the same k and device). */
int hipibs_scan_create(const hipibs_table *const *tables, int n_tables, uint64_t batch_bases,
                       uint32_t max_windows, int depth, hipibs_scan **out);

/* The pinned host buffers of slot `slot` (valid until hipibs_scan_free). */
int hipibs_scan_slot(hipibs_scan *s, int slot, hipibs_host_slot *out);

/* Replaces EncodeProbeEngine::submit (src/gpu/EncodeProbe.cu:941-1022):
 * uploads n_seq_words / n_exc_words words and n_windows windows of the
 * slot, runs the fused scan, and downloads the statistics, all
 * asynchronously on the slot's stream. The slot must not be in flight. */
int hipibs_scan_submit(hipibs_scan *s, int slot, uint32_t n_windows, uint64_t n_seq_words,
                       uint64_t n_exc_words);

/* Replaces EncodeProbeEngine::wait (src/gpu/EncodeProbe.cu:1144-1154). */
int hipibs_scan_wait(hipibs_scan *s, int slot);

int hipibs_scan_free(hipibs_scan *s);

#ifdef __cplusplus
}
#endif

#endif /* HIPIBS_H */
