// Command-line input rules of the tools (the reference's src/CliInput.hpp
// and src/GpuCli.hpp, re-implemented): dataset folder -> KMC prefix,
// reference-file filter, results folder, strict integer parsing, output
// naming, and the device options this build understands.
#pragma once

#include "KmerDatabase.hpp"

#include <cerrno>
#include <climits>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace cli
{
namespace fs = std::filesystem;

inline bool parseInt(const std::string &text, long long lo, long long hi, long long &out)
{
    const char *s = text.c_str();
    char *end = nullptr;
    errno = 0;
    const long long v = std::strtoll(s, &end, 10);
    if (end == s || *end != '\0' || errno == ERANGE || v < lo || v > hi)
        return false;
    out = v;
    return true;
}

inline bool windowFitsK(long long W, unsigned k) { return W > (long long)k; }

// First "<prefix>.kmc_*" entry of a dataset folder.
inline std::string findPrefix(const std::string &dir)
{
    std::error_code ec;
    for (fs::directory_iterator it(dir, ec), end; !ec && it != end; it.increment(ec))
    {
        const std::string f = it->path().filename().string();
        const size_t p = f.find(".kmc_");
        if (p != std::string::npos)
            return f.substr(0, p);
    }
    return "";
}

inline bool resolveDatabase(const std::string &dir, std::string &prefixPath)
{
    std::error_code ec;
    if (!fs::is_directory(dir, ec))
    {
        std::cerr << "Error: KMC database folder not found: " << dir;
        if (fs::exists(dir + ".kmc_pre", ec))
            std::cerr << " (this is a KMC prefix; pass the folder that holds it)";
        std::cerr << std::endl;
        return false;
    }
    const std::string prefix = findPrefix(dir);
    const std::string path = dir + "/" + prefix;
    if (prefix.empty() || !fs::exists(path + ".kmc_pre", ec) || !fs::exists(path + ".kmc_suf", ec))
    {
        std::cerr << "Error: no KMC database (<prefix>.kmc_pre and <prefix>.kmc_suf) in " << dir
                  << std::endl;
        return false;
    }
    prefixPath = path;
    return true;
}

inline std::string stripGz(const std::string &f)
{
    return f.size() > 3 && f.compare(f.size() - 3, 3, ".gz") == 0 ? f.substr(0, f.size() - 3) : f;
}

inline bool isReferenceFile(const std::string &f)
{
    const std::string ext = fs::path(stripGz(f)).extension().string();
    return ext == ".fasta" || ext == ".fa" || ext == ".fna";
}

inline bool listReferences(const std::string &dir, std::vector<fs::path> &out)
{
    out.clear();
    std::error_code ec;
    if (!fs::is_directory(dir, ec))
    {
        std::cerr << "Error: reference folder not found: " << dir << std::endl;
        return false;
    }
    for (fs::directory_iterator it(dir, ec), end; !ec && it != end; it.increment(ec))
    {
        const std::string name = it->path().filename().string();
        if (isReferenceFile(name))
            out.push_back(it->path());
        else
            std::cout << "Skipping " << name << " (not a reference file: .fasta, .fa or .fna, "
                      << "optionally .gz)" << std::endl;
    }
    if (ec || out.empty())
    {
        std::cerr << "Error: no reference files (.fasta, .fa or .fna, optionally .gz) in " << dir
                  << std::endl;
        return false;
    }
    return true;
}

inline bool prepareResultsFolder(const std::string &dir)
{
    std::error_code ec;
    if (fs::is_directory(dir, ec))
        return true;
    if (fs::exists(dir, ec))
    {
        std::cerr << "Error: results path exists and is not a folder: " << dir << std::endl;
        return false;
    }
    fs::create_directories(dir, ec);
    if (ec || !fs::is_directory(dir))
    {
        std::cerr << "Error: cannot create the results folder " << dir << std::endl;
        return false;
    }
    return true;
}

// "<name>_genome..." -> "<name>" (src/FastIBS.cpp:14-25).
inline std::string referenceName(const std::string &filename)
{
    const size_t p = filename.find("_genome");
    return p == std::string::npos ? filename : filename.substr(0, p);
}

inline void removeTrailingSlash(std::string &p)
{
    if (!p.empty() && p.back() == '/')
        p.pop_back();
}

inline void logReferenceError(const std::string &ref, const std::exception &e)
{
    std::cerr << "Error processing reference " << ref << ": " << e.what() << " (also in log.txt)"
              << std::endl;
    std::ofstream log("log.txt", std::ios::app);
    if (log)
        log << "Error processing reference: " << ref << "\n" << e.what() << std::endl;
}

struct Options
{
    std::vector<std::string> dbs; // --db=
    std::string dbList;           // --db-list=
    int resident = 0;             // --resident=
};

inline const char *optionsUsage()
{
    return "Options:\n"
           "  --gpu                 use the GPU (the only path of this build; accepted for\n"
           "                        compatibility)\n"
           "  --device=N            HIP device ordinal (default 0)\n"
           "  --batch-mb=N          bases per streamed batch, in MiB (default 128)\n"
           "  --streams=N           batches in flight (default 3)\n"
           "  --load-factor=F       device table load factor (default 0.2 for minimizer\n"
           "                        tables, 0.3 if memory is short, 0.6 for whole-k-mer)\n"
           "  --bucket-bytes=64|128 device table bucket width (default 128; a panel\n"
           "                        times both unless this is given)\n"
           "  --hbm-budget-gb=N     cap on device table memory\n"
           "  --minimizer=M[,CAP]   minimizer tables (k <= 32, M in 5..min(31, k), groups of\n"
           "                        up to CAP keys, default 16): runs of positions that share\n"
           "                        a minimizer read their bucket once. Default for one\n"
           "                        sample at k = 13..32: M = k - 4; panels and intersects\n"
           "                        default to whole-k-mer tables (M given: minimizer tables\n"
           "                        for every sample); --minimizer=0: whole-k-mer tables\n"
           "  --db=<folder>         (fastibs) another sample in the same reference pass\n"
           "  --db-list=<file>      (fastibs) sample folders, one per line\n"
           "  --resident=N          (fastibs) at most N sample tables on the device at once\n";
}

// Parses argv[first..]; false (after an error line) on anything unknown.
inline bool parseOptions(int argc, char **argv, int first, Options &o, bool panel)
{
    auto &g = KmerDatabase::gpuOptions();
    for (int i = first; i < argc; ++i)
    {
        const std::string a = argv[i];
        const size_t eq = a.find('=');
        const std::string key = a.substr(0, eq);
        const std::string val = eq == std::string::npos ? "" : a.substr(eq + 1);
        long long v = 0;
        if (a == "--gpu")
            continue;
        if (a == "--cpu")
        {
            std::cerr << "Error: --cpu: this build has no CPU path (the scan runs on the GPU)"
                      << std::endl;
            return false;
        }
        if (key == "--device" && parseInt(val, 0, 1024, v))
            g.device = (int)v;
        else if (key == "--batch-mb" && parseInt(val, 1, 1 << 20, v))
            g.batchBases = (uint64_t)v << 20;
        else if (key == "--streams" && parseInt(val, 1, 16, v))
            g.depth = (int)v;
        else if (key == "--load-factor" && !val.empty())
        {
            char *end = nullptr;
            const double f = std::strtod(val.c_str(), &end);
            if (*end || !(f > 0 && f < 0.98))
            {
                std::cerr << "Error: --load-factor must be in (0, 0.98)" << std::endl;
                return false;
            }
            g.loadFactor = f;
        }
        else if (key == "--bucket-bytes" && (val == "64" || val == "128"))
        {
            g.bucketSlots = val == "64" ? 8 : 16;
            g.bucketAuto = false;
        }
        else if (key == "--hbm-budget-gb" && parseInt(val, 1, 1 << 20, v))
            g.hbmBudget = (uint64_t)v << 30;
        else if (key == "--minimizer" && !val.empty())
        {
            const size_t comma = val.find(',');
            long long cap = 16;
            if (val == "0")
                v = 0;
            else if (!parseInt(val.substr(0, comma), 5, 31, v) ||
                     (comma != std::string::npos &&
                      !parseInt(val.substr(comma + 1), 1, 4096, cap)))
            {
                std::cerr << "Error: --minimizer=M[,CAP] needs M in 5..31 and CAP in 1..4096 "
                             "(or 0: whole-k-mer table)"
                          << std::endl;
                return false;
            }
            g.minimizerLen = (int)v;
            g.minimizerCap = (uint32_t)cap;
        }
        else if (panel && key == "--db" && !val.empty())
            o.dbs.push_back(val);
        else if (panel && key == "--db-list" && !val.empty())
            o.dbList = val;
        else if (panel && key == "--resident" && parseInt(val, 1, 1 << 20, v))
            o.resident = (int)v;
        else
        {
            std::cerr << "Error: unknown or malformed option: " << a << std::endl;
            return false;
        }
    }
    return true;
}

inline bool wantsHelp(int argc, char **argv)
{
    for (int i = 1; i < argc; ++i)
        if (std::string(argv[i]) == "-h" || std::string(argv[i]) == "--help")
            return true;
    return false;
}

inline bool wantsVersion(int argc, char **argv)
{
    return argc == 2 && std::string(argv[1]) == "--version";
}

} // namespace cli
