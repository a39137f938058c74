// KmerCodec: canonical 2-bit k-mer codes of a sequence (k <= 32), the host
// half of the codec the device kernels implement. Interface of the
// reference's KmerCodec (src/KmerCodec.hpp:18-123): A=0 C=1 G=2 T=3 with
// the first base most significant, canonical = min(forward, reverse
// complement), any byte other than ACGTacgt invalidates the k windows that
// contain it.
#pragma once

#include <array>
#include <cstdint>
#include <string_view>
#include <vector>

class KmerCodec
{
public:
    static constexpr uint8_t INVALID_BASE = 4;

    explicit KmerCodec(uint32_t kmerSize) : k_(kmerSize) {}

    static uint8_t baseCode(char base) { return codes()[static_cast<uint8_t>(base)]; }

    uint32_t kmerSize() const { return k_; }

    // Canonical codes of the valid positions, in order (compacted).
    void getCanonicalCodes(std::string_view seq, std::vector<uint64_t> &out) const
    {
        out.clear();
        if (seq.size() < k_)
            return;
        out.reserve(seq.size() - k_ + 1);
        roll(seq, [&](size_t, bool ok, uint64_t c) {
            if (ok)
                out.push_back(c);
        });
    }

    // One entry per position i in [0, n-k]: codes[i] (unspecified when
    // !valid[i]; 0 here) and valid[i].
    void encodePositions(std::string_view seq, std::vector<uint64_t> &codes,
                         std::vector<uint8_t> &valid) const
    {
        codes.clear();
        valid.clear();
        if (seq.size() < k_)
            return;
        codes.resize(seq.size() - k_ + 1);
        valid.resize(seq.size() - k_ + 1);
        roll(seq, [&](size_t i, bool ok, uint64_t c) {
            valid[i] = ok;
            codes[i] = ok ? c : 0;
        });
    }

private:
    // Rolling pass: forward code shifted in at the low end, reverse
    // complement shifted in at the high end, both over 2k bits; a run
    // counter of consecutive valid bases decides validity.
    template <typename Emit>
    void roll(std::string_view seq, Emit &&emit) const
    {
        const uint64_t mask = k_ >= 32 ? ~0ull : (1ull << (2 * k_)) - 1;
        const uint32_t top = 2 * (k_ - 1);
        uint64_t fwd = 0, rev = 0;
        uint32_t run = 0;
        const auto &t = codes();
        for (size_t j = 0; j < seq.size(); ++j)
        {
            uint64_t c = t[static_cast<uint8_t>(seq[j])];
            if (c == INVALID_BASE)
            {
                run = 0;
                c = 0;
            }
            else
                ++run;
            fwd = ((fwd << 2) | c) & mask;
            rev = (rev >> 2) | ((3 - c) << top);
            if (j + 1 >= k_)
                emit(j + 1 - k_, run >= k_, fwd < rev ? fwd : rev);
        }
    }

    static const std::array<uint8_t, 256> &codes()
    {
        static const std::array<uint8_t, 256> t = [] {
            std::array<uint8_t, 256> a{};
            a.fill(INVALID_BASE);
            const char *u = "ACGT", *l = "acgt";
            for (uint8_t i = 0; i < 4; ++i)
                a[(uint8_t)u[i]] = a[(uint8_t)l[i]] = i;
            return a;
        }();
        return t;
    }

    uint32_t k_;
};
