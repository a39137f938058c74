// Sequence input and window geometry of the scan (the reference's
// src/Utils.hpp:127-327, re-implemented): FASTA/FASTQ reading (optionally
// gzip), the scan's overlapping windows and the mapper's position-
// partitioning chunks.
#pragma once

#include <zlib.h>

#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <string_view>
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

// FASTA or FASTQ, decided by the first byte ('>' or '@'). Lines are split
// on '\n' only (a '\r' stays in the sequence and becomes an invalid base),
// empty lines are ignored, a record with no sequence is dropped.
inline ParseResult readSequenceFile(const std::string &path, bool compressed = false)
{
    std::string data;
    if (compressed)
        data = decompressGzip(path);
    else
    {
        std::ifstream f(path, std::ios::binary | std::ios::ate);
        if (!f)
            throw std::runtime_error("Failed to open file " + path);
        data.resize((size_t)f.tellg());
        f.seekg(0);
        f.read(data.data(), (std::streamsize)data.size());
    }
    if (data.empty())
        throw std::runtime_error("Empty file");
    const bool fastq = data[0] == '@';
    if (!fastq && data[0] != '>')
        throw std::runtime_error("Unknown file type");
    std::cout << "Parsing " << (fastq ? "fastq" : "fasta") << " data\n";
    const char head = fastq ? '@' : '>';
    ParseResult r;
    std::string id, seq;
    bool skipQuality = false;
    size_t pos = 0;
    const size_t n = data.size();
    while (pos < n)
    {
        const char *p = data.data() + pos;
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
