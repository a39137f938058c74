// KMC database reader and writer, base packing, error state and device
// queries (host code of libhipibs).
//
// On-disk format (KMC 3.2.4, versions KMC1 = 0 and KMC2 = 0x200), as read
// by CKMCFile (KMC/kmc_api/kmc_file.cpp:178-325) and KmcRaIndex
// (src/gpu/KmcRaIndex.hpp:497-708):
//   .kmc_pre: "KMCP" | LUT: u64 record offsets, one per (bin, prefix), bins
//             consecutive, 4^p per bin (KMC1: one bin) | guard u64 |
//             [KMC2: signature map, u32 x (4^sigLen + 1)] | header |
//             header_offset (u32; only its low byte is read) | "KMCP".
//             The header ends 8 bytes before the end of file with the
//             format version; header_offset is its length.
//   .kmc_suf: "KMCS" | records | "KMCS"; record = (k-p)/4 suffix bytes,
//             most significant first, then counter_size counter bytes,
//             least significant first. Records of one (bin, prefix) are
//             sorted by suffix.
// A k-mer's LUT entry is bin*4^p + (code >> 2(k-p)); its bin is
// sigMap[signature] (KMC2) where signature is the minimum, over the
// k-sigLen+1 m-mers, of the m-mer's normalised value (KMC/kmc_api/mmer.h:40-90).
#include "hipibs_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fcntl.h>
#include <sched.h>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

namespace hipibs
{

static thread_local std::string tLastError;

int setError(int code, const std::string &msg)
{
    tLastError = msg;
    return code;
}

int hipFail(hipError_t e, const char *what)
{
    tLastError = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? HIPIBS_E_NOMEM : HIPIBS_E_HIP;
}

namespace
{

struct Fd
{
    int fd = -1;
    explicit Fd(const std::string &path, int flags = O_RDONLY, mode_t mode = 0644)
    {
        fd = ::open(path.c_str(), flags, mode);
    }
    ~Fd()
    {
        if (fd >= 0)
            ::close(fd);
    }
};

bool preadAll(int fd, void *dst, size_t bytes, uint64_t off)
{
    auto *p = static_cast<char *>(dst);
    while (bytes)
    {
        const ssize_t r = ::pread(fd, p, bytes, (off_t)off);
        if (r <= 0)
        {
            if (r < 0 && errno == EINTR)
                continue;
            return false;
        }
        p += r;
        bytes -= (size_t)r;
        off += (uint64_t)r;
    }
    return true;
}

bool writeAll(int fd, const void *src, size_t bytes)
{
    auto *p = static_cast<const char *>(src);
    while (bytes)
    {
        const ssize_t r = ::write(fd, p, bytes);
        if (r <= 0)
        {
            if (r < 0 && errno == EINTR)
                continue;
            return false;
        }
        p += r;
        bytes -= (size_t)r;
    }
    return true;
}

uint64_t fileSize(int fd)
{
    const off_t end = ::lseek(fd, 0, SEEK_END);
    return end < 0 ? 0 : (uint64_t)end;
}

unsigned usableCpus()
{
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0)
    {
        const int n = CPU_COUNT(&set);
        if (n > 0)
            return (unsigned)n;
    }
    const unsigned hc = std::thread::hardware_concurrency();
    return hc ? hc : 1u;
}

struct PreFile
{
    hipibs_kmc_info info{};
    uint64_t headerOffset = 0;
    uint64_t fileBytes = 0;
    LutVector lut; // with the guard entry
    std::vector<uint32_t> sigMap;
};

// preadAll of [off, off + bytes) into dst by up to `threads` threads in
// pieces of at least 4 MiB (each thread also first-touches its part).
bool preadParallel(int fd, void *dst, uint64_t bytes, uint64_t off, unsigned threads)
{
    const uint64_t minPiece = 4ull << 20;
    threads = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>({threads, 64, bytes / minPiece}));
    if (threads == 1)
        return preadAll(fd, dst, bytes, off);
    const uint64_t per = (bytes + threads - 1) / threads;
    std::atomic<bool> bad{false};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            const uint64_t a = std::min(bytes, (uint64_t)t * per), b = std::min(bytes, a + per);
            if (a < b && !preadAll(fd, static_cast<char *>(dst) + a, b - a, off + a))
                bad = true;
        });
    for (auto &th : pool)
        th.join();
    return !bad;
}

// lut[0 .. n-1] non-decreasing, checked on up to `threads` threads.
bool monotonic(const uint64_t *lut, uint64_t n, unsigned threads)
{
    threads = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>({threads, 64, n / (1u << 20)}));
    const uint64_t per = (n + threads - 1) / threads;
    std::atomic<bool> bad{false};
    auto check = [&](unsigned t) {
        const uint64_t a = std::min(n, (uint64_t)t * per), b = std::min(n, a + per);
        for (uint64_t j = std::max<uint64_t>(a, 1); j < b; ++j)
            if (lut[j] < lut[j - 1])
            {
                bad = true;
                return;
            }
    };
    if (threads == 1)
        check(0);
    else
    {
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < threads; ++t)
            pool.emplace_back(check, t);
        for (auto &th : pool)
            th.join();
    }
    return !bad;
}

int readHeaderOf(int fd, const std::string &path, PreFile &pf)
{
    const uint64_t fsize = fileSize(fd);
    char head[4], tail[4];
    if (fsize < 24 || !preadAll(fd, head, 4, 0) || !preadAll(fd, tail, 4, fsize - 4) ||
        std::memcmp(head, "KMCP", 4) || std::memcmp(tail, "KMCP", 4))
        return setError(HIPIBS_E_IO, "bad KMC prefix file markers: " + path);
    pf.fileBytes = fsize;
    uint32_t version = 0;
    unsigned char off8 = 0;
    if (!preadAll(fd, &version, 4, fsize - 12) || !preadAll(fd, &off8, 1, fsize - 8))
        return setError(HIPIBS_E_IO, "short read of KMC header: " + path);
    if (version != 0 && version != 0x200)
        return setError(HIPIBS_E_IO, "unsupported KMC format version in " + path);
    pf.headerOffset = off8;
    if (pf.headerOffset + 8 > fsize)
        return setError(HIPIBS_E_IO, "bad KMC header offset in " + path);
    const uint64_t hpos = fsize - (pf.headerOffset + 8);
    hipibs_kmc_info &I = pf.info;
    I.kmc_version = version;
    if (version == 0x200)
    {
        uint32_t h[7];
        if (!preadAll(fd, h, sizeof(h), hpos) || !preadAll(fd, &I.total_kmers, 8, hpos + 28))
            return setError(HIPIBS_E_IO, "short read of KMC header: " + path);
        I.kmer_length = h[0];
        I.mode = h[1];
        I.counter_size = h[2];
        I.lut_prefix_length = h[3];
        I.signature_len = h[4];
        I.min_count = h[5];
        I.max_count = h[6];
    }
    else
    {
        uint32_t h[6];
        if (!preadAll(fd, h, sizeof(h), hpos) || !preadAll(fd, &I.total_kmers, 8, hpos + 24))
            return setError(HIPIBS_E_IO, "short read of KMC header: " + path);
        I.kmer_length = h[0];
        I.mode = h[1];
        I.counter_size = h[2];
        I.lut_prefix_length = h[3];
        I.signature_len = 0;
        I.min_count = h[4];
        I.max_count = h[5];
    }
    return 0;
}

int loadPre(const std::string &prefix, PreFile &pf, bool withLut)
{
    const std::string path = prefix + ".kmc_pre";
    Fd f(path);
    if (f.fd < 0)
        return setError(HIPIBS_E_IO, "cannot open " + path);
    int rc = readHeaderOf(f.fd, path, pf);
    if (rc || !withLut)
        return rc;
    const hipibs_kmc_info &I = pf.info;
    if (I.mode != 0)
        return setError(HIPIBS_E_IO, "KMC: Quake-compatible counters are not supported");
    if (I.kmer_length == 0 || I.kmer_length > 64)
        return setError(HIPIBS_E_IO, "KMC: k must be in 1..64 for this loader");
    if (I.lut_prefix_length > I.kmer_length || (I.kmer_length - I.lut_prefix_length) % 4 != 0)
        return setError(HIPIBS_E_IO, "KMC: unexpected suffix geometry");
    if (I.lut_prefix_length > 15)
        return setError(HIPIBS_E_IO, "KMC: LUT prefix length too large");
    const uint64_t single = 1ull << (2 * I.lut_prefix_length);
    if (I.kmc_version == 0x200)
    {
        if (I.signature_len == 0 || I.signature_len > 12)
            return setError(HIPIBS_E_IO, "KMC: bad signature length");
        const uint64_t sigEntries = (1ull << (2 * I.signature_len)) + 1;
        const uint64_t fixed = 12 + sigEntries * 4 + pf.headerOffset + 8;
        if (pf.fileBytes < fixed + 4)
            return setError(HIPIBS_E_IO, "KMC: prefix file too small");
        const uint64_t lutArea = pf.fileBytes - fixed;
        if (lutArea % 8 || (lutArea / 8) % single)
            return setError(HIPIBS_E_IO, "KMC: LUT area is not a whole number of bins");
        const uint64_t entries = lutArea / 8 + 1;
        pf.lut.resize(entries);
        pf.sigMap.resize(sigEntries);
        if (!preadParallel(f.fd, pf.lut.data(), entries * 8, 4, usableCpus()) ||
            !preadAll(f.fd, pf.sigMap.data(), sigEntries * 4, 4 + lutArea + 8))
            return setError(HIPIBS_E_IO, "short read of KMC LUT: " + path);
        const uint64_t nBins = (entries - 1) / single;
        for (uint32_t v : pf.sigMap)
            if (v >= nBins)
                return setError(HIPIBS_E_IO, "KMC: signature map names a missing bin");
    }
    else
    {
        pf.lut.resize(single + 1);
        if (!preadParallel(f.fd, pf.lut.data(), (single + 1) * 8, 4, usableCpus()))
            return setError(HIPIBS_E_IO, "short read of KMC LUT: " + path);
    }
    // The guard entry, as both readers set it.
    pf.lut.back() = I.total_kmers + 1;
    if (!monotonic(pf.lut.data(), pf.lut.size() - 1, usableCpus()))
        return setError(HIPIBS_E_IO, "KMC: LUT is not monotonic");
    return 0;
}

} // namespace

// Reads every record's suffix and rebuilds the key, multithreaded over
// contiguous record slices (each thread locates its first bucket by binary
// search over the monotonic LUT and walks forward).
KmcRecords::~KmcRecords()
{
    if (fd >= 0)
        ::close(fd);
}

int kmcOpenRecords(const std::string &prefix, KmcRecords &out)
{
    PreFile pf;
    int rc = loadPre(prefix, pf, true);
    if (rc)
        return rc;
    const hipibs_kmc_info &I = pf.info;
    out.path = prefix + ".kmc_suf";
    out.fd = ::open(out.path.c_str(), O_RDONLY);
    if (out.fd < 0)
        return setError(HIPIBS_E_IO, "cannot open " + out.path);
    const uint64_t fsize = fileSize(out.fd);
    char head[4], tail[4];
    if (fsize < 8 || !preadAll(out.fd, head, 4, 0) || !preadAll(out.fd, tail, 4, fsize - 4) ||
        std::memcmp(head, "KMCS", 4) || std::memcmp(tail, "KMCS", 4))
        return setError(HIPIBS_E_IO, "bad KMC suffix file markers: " + out.path);
    out.info = I;
    out.total = I.total_kmers;
    out.sufBytes = (I.kmer_length - I.lut_prefix_length) / 4;
    out.recBytes = out.sufBytes + I.counter_size;
    if (out.recBytes == 0 && out.total > 0)
        return setError(HIPIBS_E_IO, "KMC: zero-length records");
    if (fsize - 8 < out.total * (uint64_t)out.recBytes)
        return setError(HIPIBS_E_IO, "KMC: suffix file smaller than its k-mer count");
    out.lut = std::move(pf.lut);
    return 0;
}

int kmcReadRecords(const KmcRecords &src, uint64_t r0, uint64_t m, void *dst, unsigned threads)
{
    const uint64_t bytes = m * src.recBytes;
    const uint64_t off0 = 4 + r0 * src.recBytes;
    if (threads == 0)
        threads = usableCpus();
    const uint64_t minPiece = 4ull << 20;
    threads = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>({threads, 64, bytes / minPiece}));
    const uint64_t per = (bytes + threads - 1) / threads;
    std::atomic<bool> bad{false};
    auto work = [&](unsigned t)
    {
        const uint64_t a = std::min(bytes, (uint64_t)t * per), b = std::min(bytes, a + per);
        if (a < b && !preadAll(src.fd, static_cast<char *>(dst) + a, b - a, off0 + a))
            bad = true;
    };
    if (threads == 1)
        work(0);
    else
    {
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < threads; ++t)
            pool.emplace_back(work, t);
        for (auto &th : pool)
            th.join();
    }
    if (bad)
        return setError(HIPIBS_E_IO, "short read of KMC suffix records: " + src.path);
    return 0;
}

int kmcReadKeys(const std::string &prefix, uint64_t *keys, uint64_t cap, uint64_t *nOut,
                unsigned threads, hipibs_kmc_info *infoOut)
{
    PreFile pf;
    int rc = loadPre(prefix, pf, true);
    if (rc)
        return rc;
    const hipibs_kmc_info I = pf.info;
    if (infoOut)
        *infoOut = I;
    if (nOut)
        *nOut = I.total_kmers;
    const uint64_t total = I.total_kmers;
    const std::string sufPath = prefix + ".kmc_suf";
    Fd f(sufPath);
    if (f.fd < 0)
        return setError(HIPIBS_E_IO, "cannot open " + sufPath);
    const uint64_t fsize = fileSize(f.fd);
    char head[4], tail[4];
    if (fsize < 8 || !preadAll(f.fd, head, 4, 0) || !preadAll(f.fd, tail, 4, fsize - 4) ||
        std::memcmp(head, "KMCS", 4) || std::memcmp(tail, "KMCS", 4))
        return setError(HIPIBS_E_IO, "bad KMC suffix file markers: " + sufPath);
    const uint32_t sufBytes = (I.kmer_length - I.lut_prefix_length) / 4;
    const uint32_t recBytes = sufBytes + I.counter_size;
    if (recBytes == 0 && total > 0)
        return setError(HIPIBS_E_IO, "KMC: zero-length records");
    if (fsize - 8 < total * (uint64_t)recBytes)
        return setError(HIPIBS_E_IO, "KMC: suffix file smaller than its k-mer count");
    const uint64_t want = std::min(cap, total);
    if (want == 0)
        return 0;
    if (!keys)
        return setError(HIPIBS_E_ARG, "null key buffer");
    const uint32_t shift = 2 * (I.kmer_length - I.lut_prefix_length);
    const uint64_t prefixMask = (1ull << (2 * I.lut_prefix_length)) - 1;
    const bool wide = I.kmer_length > 32; // two words per key: high, low
    const LutVector &lut = pf.lut;
    const uint64_t nBuckets = lut.size() - 1;

    if (threads == 0)
        threads = usableCpus();
    threads = std::max(1u, std::min(threads, 64u));
    if (want < (1ull << 20))
        threads = 1;
    std::vector<std::thread> pool;
    std::vector<std::string> errs(threads);
    const uint64_t per = (want + threads - 1) / threads;
    auto work = [&](unsigned t)
    {
        const uint64_t first = std::min<uint64_t>(want, (uint64_t)t * per);
        const uint64_t last = std::min<uint64_t>(want, first + per);
        if (first >= last)
            return;
        // bucket j holds records [lut[j], min(lut[j+1], total))
        uint64_t j = (uint64_t)(std::upper_bound(lut.begin(), lut.begin() + nBuckets, first) -
                                lut.begin());
        j = j ? j - 1 : 0;
        std::vector<unsigned char> buf((size_t)1 << 20);
        const uint64_t recsPerChunk = std::max<uint64_t>(1, buf.size() / recBytes);
        uint64_t r = first;
        while (r < last)
        {
            const uint64_t m = std::min(recsPerChunk, last - r);
            if (!preadAll(f.fd, buf.data(), m * recBytes, 4 + r * recBytes))
            {
                errs[t] = "short read of KMC suffix records";
                return;
            }
            const unsigned char *q = buf.data();
            for (uint64_t i = 0; i < m; ++i, ++r, q += recBytes)
            {
                while (j + 1 < nBuckets && std::min<uint64_t>(lut[j + 1], total) <= r)
                    ++j;
                if (wide)
                {
                    Wide v = 0;
                    for (uint32_t b = 0; b < sufBytes; ++b)
                        v = (v << 8) | q[b];
                    if (shift < 128)
                        v |= (Wide)(j & prefixMask) << shift;
                    keys[2 * r] = (uint64_t)(v >> 64);
                    keys[2 * r + 1] = (uint64_t)v;
                    continue;
                }
                uint64_t v = 0;
                for (uint32_t b = 0; b < sufBytes; ++b)
                    v = (v << 8) | q[b];
                keys[r] = ((j & prefixMask) << shift) | v;
            }
        }
    };
    if (threads == 1)
        work(0);
    else
    {
        for (unsigned t = 0; t < threads; ++t)
            pool.emplace_back(work, t);
        for (auto &th : pool)
            th.join();
    }
    for (auto &e : errs)
        if (!e.empty())
            return setError(HIPIBS_E_IO, e + ": " + sufPath);
    return 0;
}

// ---- writer ------------------------------------------------------------

namespace
{

bool mmerAllowed(uint32_t mmer, uint32_t len)
{
    if ((mmer & 0x3f) == 0x3f || (mmer & 0x3f) == 0x3b || (mmer & 0x3c) == 0x3c)
        return false;
    for (uint32_t j = 0; j + 3 < len; ++j)
    {
        if ((mmer & 0xf) == 0)
            return false;
        mmer >>= 2;
    }
    if (mmer == 0 || mmer == 0x04 || (mmer & 0xf) == 0)
        return false;
    return true;
}

std::vector<uint32_t> normTable(uint32_t len)
{
    const uint32_t special = 1u << (2 * len);
    std::vector<uint32_t> norm(special);
    for (uint32_t i = 0; i < special; ++i)
    {
        uint32_t rev = 0, m = i;
        for (uint32_t j = 0; j < len; ++j)
        {
            rev = (rev << 2) | (3 - (m & 3));
            m >>= 2;
        }
        const uint32_t a = mmerAllowed(i, len) ? i : special;
        const uint32_t b = mmerAllowed(rev, len) ? rev : special;
        norm[i] = std::min(a, b);
    }
    return norm;
}

template <typename Key>
uint32_t signatureOf(Key code, uint32_t k, uint32_t len, const std::vector<uint32_t> &norm)
{
    const uint32_t mask = (1u << (2 * len)) - 1;
    uint32_t cur = 0, best = 0xffffffffu;
    for (uint32_t i = 0; i < k; ++i)
    {
        cur = ((cur << 2) | (uint32_t)((code >> (2 * (k - 1 - i))) & 3)) & mask;
        if (i + 1 >= len)
            best = std::min(best, norm[cur]);
    }
    return best;
}

// Key i of a caller's array: one word (k <= 32) or a (high, low) pair.
template <typename Key> Key keyAt(const uint64_t *keys, uint64_t i);
template <> uint64_t keyAt<uint64_t>(const uint64_t *keys, uint64_t i)
{
    return keys[i];
}
template <> Wide keyAt<Wide>(const uint64_t *keys, uint64_t i)
{
    return ((Wide)keys[2 * i] << 64) | keys[2 * i + 1];
}

// All ones in the low `bits` bits of Key.
template <typename Key> Key lowMask(uint32_t bits)
{
    return bits >= 8 * sizeof(Key) ? ~(Key)0 : (((Key)1 << bits) - 1);
}

template <typename Key>
int kmcWriteT(const std::string &prefix, const uint64_t *keys, const uint32_t *counts, uint64_t n,
              uint32_t k, uint32_t p, uint32_t sigLen, uint32_t counterSize, uint32_t version,
              uint32_t minCount, uint32_t maxCount)
{
    if (version != 0 && version != 0x200)
        return setError(HIPIBS_E_ARG, "kmc_write: version must be 0 or 0x200");
    if (version == 0x200 && (sigLen < 5 || sigLen > 11 || sigLen > k))
        return setError(HIPIBS_E_ARG, "kmc_write: signature length must be 5..11 and <= k");
    if (counterSize > 4)
        return setError(HIPIBS_E_ARG, "kmc_write: counter size must be 0..4");
    if (n && !keys)
        return setError(HIPIBS_E_ARG, "kmc_write: null keys");
    const uint64_t single = 1ull << (2 * p);
    const uint32_t nBins = version == 0x200 ? 512 : 1;
    std::vector<uint32_t> norm;
    std::vector<uint32_t> sigMap;
    if (version == 0x200)
    {
        norm = normTable(sigLen);
        sigMap.resize((1ull << (2 * sigLen)) + 1);
        for (size_t s = 0; s < sigMap.size(); ++s)
            sigMap[s] = (uint32_t)(s % nBins);
    }
    // Records sorted by (bin, key) == (bin, prefix, suffix): bins computed
    // in parallel, a counting sort by bin, then each bin sorted on its own.
    struct Rec
    {
        Key key;
        uint64_t count;
    };
    const Key kmask = lowMask<Key>(2 * k);
    const uint32_t shift = 2 * (k - p);
    auto prefixOf = [shift](Key key)
    { return shift >= 8 * sizeof(Key) ? 0ull : (uint64_t)(key >> shift); };
    const Key sufMask = lowMask<Key>(shift);
    unsigned T = std::max(1u, std::min(usableCpus(), 64u));
    if (n < (1ull << 20))
        T = 1;
    std::vector<uint16_t> binOf(n);
    std::vector<std::vector<uint64_t>> hist(T, std::vector<uint64_t>(nBins, 0));
    std::atomic<bool> wide{false};
    auto parallel = [T](auto &&fn)
    {
        if (T == 1)
        {
            fn(0u);
            return;
        }
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < T; ++t)
            pool.emplace_back(fn, t);
        for (auto &th : pool)
            th.join();
    };
    const uint64_t per = (n + T - 1) / T;
    parallel([&](unsigned t)
             {
                 const uint64_t lo = std::min<uint64_t>(n, (uint64_t)t * per);
                 const uint64_t hi = std::min<uint64_t>(n, lo + per);
                 for (uint64_t i = lo; i < hi; ++i)
                 {
                     const Key key = keyAt<Key>(keys, i);
                     if (key & ~kmask)
                         wide = true;
                     const uint32_t b =
                         version == 0x200 ? sigMap[signatureOf(key & kmask, k, sigLen, norm)] : 0;
                     binOf[i] = (uint16_t)b;
                     hist[t][b] += 1;
                 }
             });
    if (wide)
        return setError(HIPIBS_E_ARG, "kmc_write: key wider than 2k bits");
    std::vector<uint64_t> binStart(nBins + 1, 0);
    for (uint32_t b = 0; b < nBins; ++b)
    {
        uint64_t c = 0;
        for (unsigned t = 0; t < T; ++t)
            c += hist[t][b];
        binStart[b + 1] = binStart[b] + c;
    }
    std::vector<Rec> recs(n);
    parallel([&](unsigned t)
             {
                 std::vector<uint64_t> pos(nBins);
                 for (uint32_t b = 0; b < nBins; ++b)
                 {
                     uint64_t o = binStart[b];
                     for (unsigned u = 0; u < t; ++u)
                         o += hist[u][b];
                     pos[b] = o;
                 }
                 const uint64_t lo = std::min<uint64_t>(n, (uint64_t)t * per);
                 const uint64_t hi = std::min<uint64_t>(n, lo + per);
                 for (uint64_t i = lo; i < hi; ++i)
                     recs[pos[binOf[i]]++] = Rec{keyAt<Key>(keys, i), counts ? counts[i] : 1ull};
             });
    std::vector<uint16_t>().swap(binOf);
    std::vector<uint64_t> lut((uint64_t)nBins * single + 1, 0);
    std::atomic<uint32_t> nextBin{0};
    std::atomic<bool> dup{false};
    parallel([&](unsigned)
             {
                 for (;;)
                 {
                     const uint32_t b = nextBin.fetch_add(1);
                     if (b >= nBins)
                         break;
                     Rec *r0 = recs.data() + binStart[b], *r1 = recs.data() + binStart[b + 1];
                     std::sort(r0, r1, [](const Rec &x, const Rec &y) { return x.key < y.key; });
                     uint64_t *cnt = lut.data() + (uint64_t)b * single;
                     for (Rec *r = r0; r < r1; ++r)
                     {
                         if (r > r0 && r->key == r[-1].key)
                             dup = true;
                         cnt[prefixOf(r->key)] += 1;
                     }
                 }
             });
    if (dup)
        return setError(HIPIBS_E_ARG, "kmc_write: duplicate key");
    {
        uint64_t acc = 0;
        for (size_t j = 0; j + 1 < lut.size(); ++j)
        {
            const uint64_t c = lut[j];
            lut[j] = acc;
            acc += c;
        }
        lut.back() = acc;
    }
    // suffix file
    const uint32_t sufBytes = (k - p) / 4;
    const uint32_t recBytes = sufBytes + counterSize;
    const uint64_t cmax = counterSize == 0 ? 1 : (counterSize >= 4 ? 0xffffffffull
                                                                      : (1ull << (8 * counterSize)) - 1);
    {
        const std::string path = prefix + ".kmc_suf";
        Fd f(path, O_WRONLY | O_CREAT | O_TRUNC);
        if (f.fd < 0)
            return setError(HIPIBS_E_IO, "cannot create " + path);
        std::vector<unsigned char> buf;
        buf.reserve(1 << 20);
        buf.insert(buf.end(), {'K', 'M', 'C', 'S'});
        for (const Rec &o : recs)
        {
            const Key suf = o.key & sufMask;
            for (int b = (int)sufBytes - 1; b >= 0; --b)
                buf.push_back((unsigned char)(suf >> (8 * b)));
            const uint64_t c = std::min<uint64_t>(o.count, cmax);
            for (uint32_t b = 0; b < counterSize; ++b)
                buf.push_back((unsigned char)(c >> (8 * b)));
            if (buf.size() >= (1u << 20) - recBytes)
            {
                if (!writeAll(f.fd, buf.data(), buf.size()))
                    return setError(HIPIBS_E_IO, "write failed: " + path);
                buf.clear();
            }
        }
        buf.insert(buf.end(), {'K', 'M', 'C', 'S'});
        if (!writeAll(f.fd, buf.data(), buf.size()))
            return setError(HIPIBS_E_IO, "write failed: " + path);
    }
    // prefix file
    {
        const std::string path = prefix + ".kmc_pre";
        Fd f(path, O_WRONLY | O_CREAT | O_TRUNC);
        if (f.fd < 0)
            return setError(HIPIBS_E_IO, "cannot create " + path);
        std::vector<unsigned char> out;
        auto put = [&out](const void *src, size_t bytes)
        {
            const auto *c = static_cast<const unsigned char *>(src);
            out.insert(out.end(), c, c + bytes);
        };
        put("KMCP", 4);
        put(lut.data(), lut.size() * 8); // LUT + guard
        if (version == 0x200)
            put(sigMap.data(), sigMap.size() * 4);
        const uint32_t headerBytes = 64;
        unsigned char hdr[64];
        std::memset(hdr, 0, sizeof(hdr));
        const uint64_t total = n;
        if (version == 0x200)
        {
            const uint32_t h[7] = {k, 0, counterSize, p, sigLen, minCount, maxCount};
            std::memcpy(hdr, h, sizeof(h));
            std::memcpy(hdr + 28, &total, 8);
            hdr[36] = 0; // stored !both_strands: canonical k-mers
        }
        else
        {
            const uint32_t h[6] = {k, 0, counterSize, p, minCount, maxCount};
            std::memcpy(hdr, h, sizeof(h));
            std::memcpy(hdr + 24, &total, 8);
            hdr[32] = 0;
            const uint32_t maxHi = 0;
            std::memcpy(hdr + 33, &maxHi, 4);
        }
        std::memcpy(hdr + headerBytes - 4, &version, 4);
        put(hdr, headerBytes);
        put(&headerBytes, 4);
        put("KMCP", 4);
        if (!writeAll(f.fd, out.data(), out.size()))
            return setError(HIPIBS_E_IO, "write failed: " + path);
    }
    return 0;
}

} // namespace

int kmcWrite(const std::string &prefix, const uint64_t *keys, const uint32_t *counts, uint64_t n,
             uint32_t k, uint32_t p, uint32_t sigLen, uint32_t counterSize, uint32_t version,
             uint32_t minCount, uint32_t maxCount)
{
    if (k == 0 || k > 64 || p > k || (k - p) % 4 != 0 || p > 12)
        return setError(HIPIBS_E_ARG, "kmc_write: need k <= 64, p <= 12, (k - p) % 4 == 0");
    return k <= 32 ? kmcWriteT<uint64_t>(prefix, keys, counts, n, k, p, sigLen, counterSize,
                                         version, minCount, maxCount)
                   : kmcWriteT<Wide>(prefix, keys, counts, n, k, p, sigLen, counterSize, version,
                                     minCount, maxCount);
}

} // namespace hipibs

using namespace hipibs;

extern "C" const char *hipibs_last_error(void)
{
    return tLastError.c_str();
}

extern "C" int hipibs_device_count(int *n, char *why, size_t why_len)
{
    if (!n)
        return setError(HIPIBS_E_ARG, "null argument");
    *n = 0;
    int c = 0;
    const hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess || c <= 0)
    {
        const std::string msg = e != hipSuccess ? std::string("hipGetDeviceCount: ") +
                                                      hipGetErrorString(e)
                                                : std::string("no HIP devices");
        if (why && why_len)
        {
            std::strncpy(why, msg.c_str(), why_len - 1);
            why[why_len - 1] = 0;
        }
        return setError(HIPIBS_E_NODEV, msg);
    }
    *n = c;
    return 0;
}

extern "C" int hipibs_device_mem(int device, uint64_t *free_bytes, uint64_t *total_bytes)
{
    HIPIBS_CHECK(hipSetDevice(device));
    size_t f = 0, t = 0;
    HIPIBS_CHECK(hipMemGetInfo(&f, &t));
    if (free_bytes)
        *free_bytes = f;
    if (total_bytes)
        *total_bytes = t;
    return 0;
}

extern "C" int hipibs_kmc_header(const char *prefix, hipibs_kmc_info *out)
{
    if (!prefix || !out)
        return setError(HIPIBS_E_ARG, "null argument");
    PreFile pf;
    int rc = loadPre(prefix, pf, false);
    if (!rc)
        *out = pf.info;
    return rc;
}

extern "C" int hipibs_kmc_read_keys(const char *prefix, uint64_t *h_keys, uint64_t cap, uint64_t *n,
                                    uint32_t threads)
{
    if (!prefix)
        return setError(HIPIBS_E_ARG, "null prefix");
    return kmcReadKeys(prefix, h_keys, cap, n, threads, nullptr);
}

extern "C" int hipibs_kmc_write(const char *prefix, const uint64_t *h_keys, const uint32_t *h_counts,
                                uint64_t n, uint32_t k, uint32_t p, uint32_t sig_len,
                                uint32_t counter_size, uint32_t version, uint32_t min_count,
                                uint32_t max_count)
{
    if (!prefix)
        return setError(HIPIBS_E_ARG, "null prefix");
    return kmcWrite(prefix, h_keys, h_counts, n, k, p, sig_len, counter_size, version, min_count,
                    max_count);
}

// Same table as KmerCodec::baseCode: ACGT/acgt -> 0..3, anything else 4.
static const unsigned char *baseTable()
{
    static unsigned char t[256];
    static bool init = []
    {
        std::memset(t, 4, sizeof(t));
        t[(unsigned char)'A'] = t[(unsigned char)'a'] = 0;
        t[(unsigned char)'C'] = t[(unsigned char)'c'] = 1;
        t[(unsigned char)'G'] = t[(unsigned char)'g'] = 2;
        t[(unsigned char)'T'] = t[(unsigned char)'t'] = 3;
        return true;
    }();
    (void)init;
    return t;
}

extern "C" int hipibs_pack_bases(const char *h_text, uint64_t n, uint64_t at, uint64_t *h_seq_words,
                                 uint64_t *h_exc_words)
{
    if (n && (!h_text || !h_seq_words || !h_exc_words))
        return setError(HIPIBS_E_ARG, "null argument");
    const unsigned char *T = baseTable();
    uint64_t b = at;
    uint64_t i = 0;
    // unaligned head, base by base
    while (i < n && (b & 63))
    {
        const unsigned char c = T[(unsigned char)h_text[i]];
        if (c > 3)
            h_exc_words[b >> 6] |= 1ull << (b & 63);
        else
            h_seq_words[b >> 5] |= (uint64_t)c << (2 * (b & 31));
        ++i;
        ++b;
    }
    // aligned body: 64 bases -> 2 sequence words + 1 exception word
    while (i + 64 <= n)
    {
        uint64_t s0 = 0, s1 = 0, e = 0;
        for (int j = 0; j < 32; ++j)
        {
            const unsigned char c = T[(unsigned char)h_text[i + j]];
            e |= (uint64_t)(c > 3) << j;
            s0 |= (uint64_t)(c & 3) << (2 * j);
        }
        for (int j = 0; j < 32; ++j)
        {
            const unsigned char c = T[(unsigned char)h_text[i + 32 + j]];
            e |= (uint64_t)(c > 3) << (32 + j);
            s1 |= (uint64_t)(c & 3) << (2 * j);
        }
        // non-ACGT bytes carry code 0 (the table's 4 has its low bits clear)
        h_seq_words[b >> 5] = s0;
        h_seq_words[(b >> 5) + 1] = s1;
        h_exc_words[b >> 6] = e;
        i += 64;
        b += 64;
    }
    if (i < n && (b & 63) == 0)
    {
        // the tail's words hold no earlier bases: clear them first
        h_seq_words[b >> 5] = 0;
        h_seq_words[(b >> 5) + 1] = 0;
        h_exc_words[b >> 6] = 0;
    }
    while (i < n)
    {
        const unsigned char c = T[(unsigned char)h_text[i]];
        if (c > 3)
            h_exc_words[b >> 6] |= 1ull << (b & 63);
        else
            h_seq_words[b >> 5] |= (uint64_t)c << (2 * (b & 31));
        ++i;
        ++b;
    }
    return 0;
}
