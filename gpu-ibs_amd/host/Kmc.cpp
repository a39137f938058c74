// kmc: the k-mer counting command the reference's users run before the
// scan (README.md:95 `kmc -k25 -ci2 -cs255 -t16 -m32 @S1.reads out tmp`,
// scripts/make_test_corpus.sh:123 `kmc -k25 -fm -ci1 -t4 in.fasta out tmp`),
// counted on the GPU and written as the same KMC database (<out>.kmc_pre,
// <out>.kmc_suf). KMC 3.2.4 itself is an external tool; this command takes
// the subset of its command line that counting canonical k-mers of FASTA or
// FASTQ reads needs:
//
//   kmc [options] <input file | @list of input files> <output prefix> <work dir>
//     -k<len>   k-mer length, 1..64 (default 25)
//     -ci<n>    drop k-mers seen fewer than n times (default 2)
//     -cx<n>    drop k-mers seen more than n times (default 1e9)
//     -cs<n>    largest stored count (default 255)
//     -fq -fa -fm  input format; the first byte of each file decides anyway
//     -p9       signature length (only 9, KMC's default)
//     -j<file>  also write the statistics as JSON
//     -t -m -sm -r -hp -v -n -sf -sp -sr  accepted; the GPU does the work
//
// Not available: -b (non-canonical counting), -fbam, -fkmc, -o kff, -w.
#include "Hipibs.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace
{

int usage()
{
    std::cerr << "usage: kmc [options] <input file | @file with input names> <output prefix> "
                 "<working directory>\n"
                 "  -k<len> (1..64, default 25)  -ci<n> (default 2)  -cx<n> (default 1e9)\n"
                 "  -cs<n> (default 255)  -fq/-fa/-fm  -j<json file>\n";
    return 1;
}

bool number(const std::string &s, double &out)
{
    if (s.empty())
        return false;
    char *end = nullptr;
    out = std::strtod(s.c_str(), &end);
    return end && *end == '\0';
}

} // namespace

int main(int argc, char **argv)
{
    uint32_t k = 25;
    double ci = 2, cx = 1e9, cs = 255;
    std::string json;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i)
    {
        const std::string a = argv[i];
        if (a.size() < 2 || a[0] != '-')
        {
            pos.push_back(a);
            continue;
        }
        double v = 0;
        auto opt = [&](const char *name) { return a.rfind(name, 0) == 0; };
        if (opt("-ci") || opt("-cx") || opt("-cs"))
        {
            if (!number(a.substr(3), v) || v < 0)
                return usage();
            (a[2] == 'i' ? ci : a[2] == 'x' ? cx : cs) = v;
        }
        else if (opt("-k"))
        {
            if (!number(a.substr(2), v) || v < 1)
                return usage();
            if (v > 64)
            {
                std::cerr << "kmc: k = " << v << " is not supported (k <= 64)\n";
                return 1;
            }
            k = (uint32_t)v;
        }
        else if (a == "-fq" || a == "-fa" || a == "-fm")
            continue;
        else if (opt("-f"))
        {
            std::cerr << "kmc: input format " << a << " is not supported (FASTA or FASTQ only)\n";
            return 1;
        }
        else if (opt("-p"))
        {
            if (a != "-p9")
            {
                std::cerr << "kmc: only the default signature length (-p9) is supported\n";
                return 1;
            }
        }
        else if (opt("-j"))
            json = a.substr(2);
        else if (a == "-b" || opt("-o") || a == "-w")
        {
            std::cerr << "kmc: option " << a << " is not supported\n";
            return 1;
        }
        else if (opt("-t") || opt("-m") || a == "-sm" || a == "-r" || a == "-hp" || a == "-v" ||
                 opt("-n") || opt("-sf") || opt("-sp") || opt("-sr"))
            continue;
        else
        {
            std::cerr << "kmc: unknown option " << a << "\n";
            return usage();
        }
    }
    if (pos.size() != 3)
        return usage();
    if (ci > 4294967295.0 || cx > 4294967295.0 || cs > 4294967295.0 || cs < 1 || ci > cx)
    {
        std::cerr << "kmc: counts out of range (need 1 <= cs, ci <= cx < 2^32)\n";
        return 1;
    }
    std::vector<std::string> inputs;
    if (pos[0][0] == '@')
    {
        std::ifstream list(pos[0].substr(1));
        if (!list)
        {
            std::cerr << "kmc: cannot open " << pos[0].substr(1) << "\n";
            return 1;
        }
        for (std::string line; std::getline(list, line);)
        {
            while (!line.empty() && (line.back() == '\r' || line.back() == ' '))
                line.pop_back();
            if (!line.empty())
                inputs.push_back(line);
        }
    }
    else
        inputs.push_back(pos[0]);
    if (inputs.empty())
    {
        std::cerr << "kmc: no input files\n";
        return 1;
    }

    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    hipibs_counter *c = nullptr;
    try
    {
        hipibs::check(hipibs_counter_create(k, 0, 0, 0, &c), "creating the k-mer counter");
        for (const auto &in : inputs)
            hipibs::check(hipibs_counter_add_file(c, in.c_str()), ("counting " + in).c_str());
        hipibs_count_stats st{};
        hipibs::check(hipibs_counter_stats(c, &st), "counting");
        const auto t1 = Clock::now();
        hipibs_count_filter f{};
        hipibs::check(hipibs_counter_write_kmc(c, pos[1].c_str(), (uint32_t)ci, (uint32_t)cx,
                                               (uint32_t)cs, &f),
                      "writing the KMC database");
        hipibs_counter_free(c);
        c = nullptr;
        const auto t2 = Clock::now();
        const double s1 = std::chrono::duration<double>(t1 - t0).count();
        const double s2 = std::chrono::duration<double>(t2 - t1).count();
        std::printf("1st stage: %.3fs\n2nd stage: %.3fs\nTotal    : %.3fs\nTmp size : 0MB\n\n",
                    s1, s2, s1 + s2);
        std::printf("Stats:\n");
        std::printf("   No. of k-mers below min. threshold : %12llu\n",
                    (unsigned long long)f.below);
        std::printf("   No. of k-mers above max. threshold : %12llu\n",
                    (unsigned long long)f.above);
        std::printf("   No. of unique k-mers               : %12llu\n",
                    (unsigned long long)st.distinct);
        std::printf("   No. of unique counted k-mers       : %12llu\n",
                    (unsigned long long)f.kept);
        std::printf("   Total no. of k-mers                : %12llu\n",
                    (unsigned long long)st.kmers);
        std::printf("   Total no. of reads                 : %12llu\n",
                    (unsigned long long)st.reads);
        if (!json.empty())
        {
            std::ofstream j(json);
            j << "{\n  \"1st_stage\": \"" << s1 << "s\",\n  \"2nd_stage\": \"" << s2
              << "s\",\n  \"Total\": \"" << s1 + s2 << "s\",\n  \"Stats\": {\n"
              << "    \"#k-mers_below_min_threshold\": " << f.below << ",\n"
              << "    \"#k-mers_above_max_threshold\": " << f.above << ",\n"
              << "    \"#Unique_k-mers\": " << st.distinct << ",\n"
              << "    \"#Unique_counted_k-mers\": " << f.kept << ",\n"
              << "    \"#Total no. of k-mers\": " << st.kmers << ",\n"
              << "    \"#Total_reads\": " << st.reads << "\n  }\n}\n";
            if (!j)
                throw std::runtime_error("cannot write " + json);
        }
    }
    catch (const std::exception &e)
    {
        if (c)
            hipibs_counter_free(c);
        std::cerr << "kmc: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
