// Sequence input and window geometry of the scan (the reference's
// src/Utils.hpp:127-327, re-implemented): FASTA/FASTQ reading (optionally
// gzip), the scan's overlapping windows and the mapper's position-
// partitioning chunks.
#pragma once

#include <fcntl.h>
#include <sched.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

struct ParseResult
{
    std::vector<std::string> ids;
    std::vector<std::string> sequences;
};

// A window or mapper chunk: id points into ParseResult::ids, seq views
// into ParseResult::sequences (which must outlive it); start/end are
// contig coordinates.
struct SeqWindow
{
    const std::string *id;
    size_t start, end;
    std::string_view seq;
};

// Whole gzip file into memory; a truncated or corrupt stream is an error.
inline std::string decompressGzip(const std::string &path)
{
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f)
        throw std::runtime_error("Failed to open file " + path);
    gzbuffer(f, 1 << 20);
    std::string out;
    std::vector<char> buf(1 << 22);
    for (;;)
    {
        const int n = gzread(f, buf.data(), (unsigned)buf.size());
        if (n < 0)
        {
            int err = 0;
            const std::string msg = gzerror(f, &err);
            gzclose(f);
            throw std::runtime_error("gzip read error in " + path + ": " + msg);
        }
        if (n == 0)
            break;
        out.append(buf.data(), (size_t)n);
    }
    int err = 0;
    const char *msg = gzerror(f, &err);
    if (err != Z_OK && err != Z_STREAM_END)
    {
        const std::string m = msg ? msg : "";
        gzclose(f);
        throw std::runtime_error("gzip stream of " + path + " is truncated or corrupt: " + m);
    }
    if (gzclose(f) != Z_OK)
        throw std::runtime_error("gzip stream of " + path + " is truncated or corrupt");
    return out;
}

namespace seqio_detail
{

// CPUs this process may run on.
inline unsigned hostThreads()
{
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0)
        return std::max(1, CPU_COUNT(&set));
    return std::max(1u, std::thread::hardware_concurrency());
}

// fn(i) for every i in [0, n), spread over up to `threads` threads; the
// first exception thrown by fn is rethrown.
template <class F>
void parallelFor(size_t n, unsigned threads, F fn)
{
    threads = (unsigned)std::min<size_t>(std::max(1u, threads), n);
    if (threads <= 1)
    {
        for (size_t i = 0; i < n; ++i)
            fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::atomic<bool> failed{false};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&]
                          {
                              for (size_t i; !failed && (i = next++) < n;)
                              {
                                  try
                                  {
                                      fn(i);
                                  }
                                  catch (...)
                                  {
                                      if (!failed.exchange(true))
                                          err = std::current_exception();
                                  }
                              }
                          });
    for (auto &th : pool)
        th.join();
    if (err)
        std::rethrow_exception(err);
}

// A whole file in memory, read by parallel pread calls (no zero fill).
struct FileBytes
{
    std::unique_ptr<char[]> data;
    size_t size = 0;
};

inline FileBytes readWholeFile(const std::string &path, unsigned threads)
{
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0)
        throw std::runtime_error("Failed to open file " + path);
    struct stat st;
    if (fstat(fd, &st) != 0)
    {
        ::close(fd);
        throw std::runtime_error("Failed to open file " + path);
    }
    FileBytes fb;
    fb.size = (size_t)st.st_size;
    fb.data.reset(new char[fb.size ? fb.size : 1]);
    const size_t piece = 64ull << 20;
    std::atomic<bool> bad{false};
    parallelFor((fb.size + piece - 1) / piece, threads, [&](size_t i)
                {
                    size_t off = i * piece;
                    const size_t end = std::min(fb.size, off + piece);
                    while (off < end)
                    {
                        const ssize_t r = ::pread(fd, fb.data.get() + off, end - off, (off_t)off);
                        if (r <= 0)
                        {
                            bad = true;
                            return;
                        }
                        off += (size_t)r;
                    }
                });
    ::close(fd);
    if (bad)
        throw std::runtime_error("Failed to read file " + path);
    return fb;
}

// FASTQ: '@' id lines, sequence lines up to a '+' line, whose next line is
// quality and skipped.
inline ParseResult parseSerial(const char *data, size_t n, bool fastq)
{
    const char head = fastq ? '@' : '>';
    ParseResult r;
    std::string id, seq;
    bool skipQuality = false;
    size_t pos = 0;
    while (pos < n)
    {
        const char *p = data + pos;
        const void *nl = std::memchr(p, '\n', n - pos);
        const size_t len = nl ? (size_t)((const char *)nl - p) : n - pos;
        pos += len + 1;
        if (len == 0)
            continue;
        if (fastq && p[0] == '+')
        {
            skipQuality = true;
            continue;
        }
        if (skipQuality)
        {
            skipQuality = false;
            continue;
        }
        if (p[0] == head)
        {
            if (!seq.empty())
            {
                r.ids.push_back(std::move(id));
                r.sequences.push_back(std::move(seq));
                seq.clear();
            }
            id.assign(p + 1, len - 1);
        }
        else
            seq.append(p, len);
    }
    if (!seq.empty())
    {
        r.ids.push_back(std::move(id));
        r.sequences.push_back(std::move(seq));
    }
    return r;
}

// FASTA with the serial parser's line rules, in parallel: a record starts
// at every '>' that begins a line; its id is the rest of that line, its
// sequence the concatenation of the following lines (up to the next
// record) without their '\n'; records with no sequence are dropped.
// `piece` is the largest span of input one task handles.
inline ParseResult parseFastaParallel(const char *d, size_t n, unsigned threads,
                                      size_t piece = 32ull << 20)
{
    piece = std::max<size_t>(piece, 1);
    const size_t nPieces = (n + piece - 1) / piece;
    std::vector<std::vector<size_t>> found(nPieces);
    parallelFor(nPieces, threads, [&](size_t i)
                {
                    const size_t lo = i * piece, hi = std::min(n, lo + piece);
                    for (size_t q = lo; q < hi;)
                    {
                        const void *g = std::memchr(d + q, '>', hi - q);
                        if (!g)
                            break;
                        const size_t at = (size_t)((const char *)g - d);
                        if (at == 0 || d[at - 1] == '\n')
                            found[i].push_back(at);
                        q = at + 1;
                    }
                });
    std::vector<size_t> heads;
    for (auto &f : found)
        heads.insert(heads.end(), f.begin(), f.end());
    found.clear();
    struct Rec
    {
        size_t idBegin, idEnd, bodyBegin, bodyEnd;
    };
    std::vector<Rec> recs(heads.size());
    for (size_t j = 0; j < heads.size(); ++j)
    {
        const size_t h = heads[j];
        const size_t next = j + 1 < heads.size() ? heads[j + 1] : n;
        const void *nl = std::memchr(d + h, '\n', next - h);
        const size_t e = nl ? (size_t)((const char *)nl - d) : next;
        recs[j] = Rec{h + 1, e, std::min(e + 1, next), next};
    }
    // tasks: spans of record bodies of at most `piece` bytes
    struct Task
    {
        size_t rec, begin, end, out;
    };
    std::vector<Task> tasks;
    for (size_t j = 0; j < recs.size(); ++j)
        for (size_t b = recs[j].bodyBegin; b < recs[j].bodyEnd; b += piece)
            tasks.push_back(Task{j, b, std::min(recs[j].bodyEnd, b + piece), 0});
    std::vector<size_t> bases(tasks.size());
    parallelFor(tasks.size(), threads, [&](size_t i)
                {
                    const Task &t = tasks[i];
                    bases[i] = (t.end - t.begin) - (size_t)std::count(d + t.begin, d + t.end, '\n');
                });
    std::vector<size_t> recLen(recs.size(), 0);
    for (size_t i = 0; i < tasks.size(); ++i)
    {
        tasks[i].out = recLen[tasks[i].rec];
        recLen[tasks[i].rec] += bases[i];
    }
    std::vector<size_t> keep;
    for (size_t j = 0; j < recs.size(); ++j)
        if (recLen[j])
            keep.push_back(j);
    std::vector<size_t> slot(recs.size(), 0);
    ParseResult r;
    r.ids.resize(keep.size());
    r.sequences.resize(keep.size());
    for (size_t q = 0; q < keep.size(); ++q)
    {
        slot[keep[q]] = q;
        r.ids[q].assign(d + recs[keep[q]].idBegin, recs[keep[q]].idEnd - recs[keep[q]].idBegin);
    }
    parallelFor(keep.size(), threads, [&](size_t q) { r.sequences[q].resize(recLen[keep[q]]); });
    parallelFor(tasks.size(), threads, [&](size_t i)
                {
                    const Task &t = tasks[i];
                    if (!recLen[t.rec])
                        return;
                    char *out = r.sequences[slot[t.rec]].data() + t.out;
                    for (size_t q = t.begin; q < t.end;)
                    {
                        const void *nl = std::memchr(d + q, '\n', t.end - q);
                        const size_t stop = nl ? (size_t)((const char *)nl - d) : t.end;
                        std::memcpy(out, d + q, stop - q);
                        out += stop - q;
                        q = stop + 1;
                    }
                });
    return r;
}

} // namespace seqio_detail

// FASTA or FASTQ, decided by the first byte ('>' or '@'). Lines are split
// on '\n' only (a '\r' stays in the sequence and becomes an invalid base),
// empty lines are ignored, a record with no sequence is dropped. FASTA is
// read and parsed on all usable CPUs (`threads` = 0) or on `threads`.
inline ParseResult readSequenceFile(const std::string &path, bool compressed = false,
                                    unsigned threads = 0)
{
    if (threads == 0)
        threads = std::min(64u, seqio_detail::hostThreads());
    std::string gz;
    seqio_detail::FileBytes fb;
    const char *data;
    size_t n;
    if (compressed)
    {
        gz = decompressGzip(path);
        data = gz.data();
        n = gz.size();
    }
    else
    {
        fb = seqio_detail::readWholeFile(path, threads);
        data = fb.data.get();
        n = fb.size;
    }
    if (n == 0)
        throw std::runtime_error("Empty file");
    const bool fastq = data[0] == '@';
    if (!fastq && data[0] != '>')
        throw std::runtime_error("Unknown file type");
    std::cout << "Parsing " << (fastq ? "fastq" : "fasta") << " data\n";
    if (fastq)
        return seqio_detail::parseSerial(data, n, true);
    return seqio_detail::parseFastaParallel(data, n, threads);
}

// Scan windows: [j, min(j+W, len)) for j = 0, W-k, 2(W-k), ... while
// j < len; consecutive windows share k bases (one k-mer position).
inline void chunkSequences(const std::vector<std::string> &ids,
                           const std::vector<std::string> &seqs, size_t W, size_t k,
                           std::vector<SeqWindow> &out)
{
    if (W <= k)
        throw std::invalid_argument("window size " + std::to_string(W) +
                                    " must be larger than k (" + std::to_string(k) + ")");
    out.clear();
    for (size_t c = 0; c < seqs.size(); ++c)
    {
        const std::string_view s = seqs[c];
        for (size_t j = 0; j < s.size(); j += W - k)
        {
            const size_t e = std::min(j + W, s.size());
            out.push_back(SeqWindow{&ids[c], j, e, s.substr(j, e - j)});
        }
    }
}

// Mapper chunks: every k-mer position of a contig in exactly one chunk
// (chunks overlap by k-1 bases); contigs shorter than k give none.
inline void chunkForMapping(const std::vector<std::string> &ids,
                            const std::vector<std::string> &seqs, size_t chunkSize, size_t k,
                            std::vector<SeqWindow> &chunks, std::vector<size_t> &chunkContig)
{
    chunkSize = std::max(chunkSize, k);
    chunks.clear();
    chunkContig.clear();
    for (size_t c = 0; c < seqs.size(); ++c)
    {
        const std::string_view s = seqs[c];
        for (size_t o = 0; o + k <= s.size();)
        {
            const size_t len = std::min(chunkSize, s.size() - o);
            chunks.push_back(SeqWindow{&ids[c], o, o + len, s.substr(o, len)});
            chunkContig.push_back(c);
            o += len - k + 1;
        }
    }
}
