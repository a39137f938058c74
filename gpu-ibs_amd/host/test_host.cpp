// Host-only checks of the C++ layer (run by tests/test_tools.py):
// KmerCodec against a string brute force, window and mapper geometry, the
// batch planner's invariants, the Window distance rule.
// `test_host reads <file>` prints the records the k-mer counter's reader
// takes from a read file, one line each.
#include "../csrc/reads.hpp"
#include "KmerCodec.hpp"
#include "ScanPipeline.hpp"
#include "SeqIO.hpp"
#include "Window.hpp"

#include <cstdio>
#include <random>

static int fails = 0;
#define CHECK(c)                                                            \
    do                                                                      \
    {                                                                       \
        if (!(c))                                                           \
        {                                                                   \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++fails;                                                        \
        }                                                                   \
    } while (0)

static bool stringCanonical(const std::string &w, uint64_t &code)
{
    std::string rc(w.rbegin(), w.rend());
    for (char &c : rc)
    {
        switch (c)
        {
        case 'A': c = 'T'; break;
        case 'C': c = 'G'; break;
        case 'G': c = 'C'; break;
        case 'T': c = 'A'; break;
        default: return false;
        }
    }
    const std::string &m = std::min(w, rc);
    code = 0;
    for (char c : m)
        code = (code << 2) | (uint64_t)(c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : 3);
    return true;
}

// Prints each record's bases followed by '\n'.
struct PrintSink : hipibs::ReadSink
{
    void bases(const char *p, size_t n) override { std::fwrite(p, 1, n, stdout); }
    void recordEnd() override { std::fputc('\n', stdout); }
    int failed() const override { return 0; }
};

int main(int argc, char **argv)
{
    if (argc == 3 && std::string(argv[1]) == "reads")
    {
        PrintSink sink;
        uint64_t n = 0;
        const int rc = hipibs::readRecords(argv[2], sink, &n);
        std::fflush(stdout);
        if (rc)
        {
            std::fprintf(stderr, "%s\n", hipibs_last_error());
            return 2;
        }
        std::fprintf(stderr, "records %llu\n", (unsigned long long)n);
        return 0;
    }
    std::mt19937_64 rng(20260830);
    const char alpha[] = "ACGTACGTacgtNnXY-";
    for (uint32_t k : {1u, 4u, 15u, 25u, 31u, 32u})
        for (size_t len : {0ul, 3ul, 31ul, 32ul, 33ul, 100ul, 399ul})
        {
            std::string s(len, 'A');
            for (auto &c : s)
                c = alpha[rng() % (sizeof(alpha) - 1)];
            KmerCodec codec(k);
            std::vector<uint64_t> codes, compact;
            std::vector<uint8_t> valid;
            codec.encodePositions(s, codes, valid);
            codec.getCanonicalCodes(s, compact);
            CHECK(codes.size() == (len >= k ? len - k + 1 : 0));
            size_t ci = 0;
            for (size_t i = 0; i < codes.size(); ++i)
            {
                std::string w = s.substr(i, k);
                for (auto &c : w)
                    c = (char)std::toupper((unsigned char)c);
                uint64_t want = 0;
                const bool ok = stringCanonical(w, want);
                CHECK(valid[i] == ok);
                if (ok)
                {
                    CHECK(codes[i] == want);
                    CHECK(ci < compact.size() && compact[ci] == want);
                    ++ci;
                }
            }
            CHECK(ci == compact.size());
        }
    // windows: [j, min(j+W, n)) for j = 0, W-k, ...
    std::vector<std::string> ids{"a", "b"}, seqs{std::string(1000, 'A'), std::string(30, 'C')};
    std::vector<SeqWindow> w;
    chunkSequences(ids, seqs, 100, 25, w);
    size_t idx = 0;
    for (size_t c = 0; c < 2; ++c)
        for (size_t j = 0; j < seqs[c].size(); j += 75, ++idx)
        {
            CHECK(w[idx].id == &ids[c] && w[idx].start == j &&
                  w[idx].end == std::min(j + 100, seqs[c].size()));
        }
    CHECK(idx == w.size());
    bool threw = false;
    try
    {
        chunkSequences(ids, seqs, 25, 25, w);
    }
    catch (const std::invalid_argument &)
    {
        threw = true;
    }
    CHECK(threw);
    // mapper chunks partition the positions
    std::vector<SeqWindow> mc;
    std::vector<size_t> cc;
    chunkForMapping(ids, seqs, 64, 25, mc, cc);
    size_t pos = 0;
    for (size_t i = 0; i < mc.size(); ++i)
        if (cc[i] == 0)
        {
            CHECK(mc[i].start == pos);
            pos += (mc[i].end - mc[i].start) - 25 + 1;
        }
    CHECK(pos == 1000 - 25 + 1);
    // batch planner: every window inside its batch's slot range, batches
    // cover the windows in order, windows of one region keep their offsets
    chunkSequences(ids, seqs, 100, 25, w);
    for (uint64_t cap : {256ull, 512ull, 4096ull})
    {
        auto batches = hipibs::detail::planBatches(w, cap, 7);
        size_t next = 0;
        for (const auto &b : batches)
        {
            CHECK(b.lo == next && b.hi > b.lo && b.hi - b.lo <= 7);
            CHECK(b.bases + 64 <= cap || b.hi - b.lo == 1);
            for (size_t j = b.lo; j < b.hi; ++j)
            {
                const uint64_t base = b.winBase[j - b.lo];
                CHECK(base + (w[j].end - w[j].start) <= b.bases);
                bool inRegion = false;
                for (const auto &r : b.regions)
                    if (base >= r.at && base + (w[j].end - w[j].start) <= r.at + r.n &&
                        r.text + (base - r.at) == w[j].seq.data())
                        inRegion = true;
                CHECK(inRegion);
            }
            next = b.hi;
        }
        CHECK(next == w.size());
    }
    Window win;
    win.addVariationOriginal(1, 25);
    win.addVariationOriginal(23, 25);
    win.addVariationOriginal(24, 25);
    win.addVariationOriginal(30, 25);
    CHECK(win.variations == 4 && win.kmerDistance == 22 + 0 + 1 + 6);

    // the parallel FASTA parser against the serial line rules, on inputs
    // with '>' inside lines, empty lines, '\r', records without sequence
    // and no final newline, cut into tasks of 1..13 bytes
    const char fa[] = "ACGTN\r\n>x y\n>";
    for (int trial = 0; trial < 400; ++trial)
    {
        std::string in = ">";
        const size_t len = rng() % 200;
        for (size_t i = 0; i < len; ++i)
            in.push_back(fa[rng() % (sizeof(fa) - 1)]);
        if (rng() % 2)
            in.push_back('\n');
        const ParseResult a = seqio_detail::parseSerial(in.data(), in.size(), false);
        const size_t piece = 1 + rng() % 13;
        const ParseResult b = seqio_detail::parseFastaParallel(in.data(), in.size(),
                                                                1 + (unsigned)(rng() % 4), piece);
        CHECK(a.ids == b.ids && a.sequences == b.sequences);
    }
    std::printf("%s\n", fails ? "host checks FAILED" : "host checks ok");
    return fails ? 1 : 0;
}
