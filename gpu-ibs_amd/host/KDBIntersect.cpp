// KDBIntersect: number of k-mers shared by two KMC databases (the
// reference's src/KDBIntersect.cpp). The database with more k-mers (ties:
// the larger .kmc_pre, then the first argument) becomes the device table;
// the other's records are read raw on host threads and decoded + probed on
// the device in pinned double-buffered chunks. stdout carries the count only.
#include "CliInput.hpp"

#include <filesystem>

using namespace std;

static int run(int argc, char **argv)
{
    if (argc < 3 || cli::wantsHelp(argc, argv))
    {
        cerr << "Usage: " << argv[0] << " <db1 folder or prefix> <db2 folder or prefix> [options]"
             << endl;
        return argc < 3 ? 1 : 0;
    }
    cli::Options opt;
    if (!cli::parseOptions(argc, argv, 3, opt, false))
        return 1;
    string p[2];
    for (int i = 0; i < 2; ++i)
    {
        string a = argv[1 + i];
        cli::removeTrailingSlash(a);
        if (filesystem::is_directory(a))
        {
            if (!cli::resolveDatabase(a, p[i]))
                return 1;
        }
        else
            p[i] = a;
    }
    hipibs_kmc_info h[2];
    for (int i = 0; i < 2; ++i)
        hipibs::check(hipibs_kmc_header(p[i].c_str(), &h[i]), "reading a KMC header");
    if (h[0].kmer_length != h[1].kmer_length)
        throw runtime_error("the databases have different k");
    auto preBytes = [](const string &pre) {
        error_code ec;
        const auto s = filesystem::file_size(pre + ".kmc_pre", ec);
        return ec ? 0ull : (unsigned long long)s;
    };
    int big = 0;
    if (h[1].total_kmers > h[0].total_kmers ||
        (h[1].total_kmers == h[0].total_kmers && preBytes(p[1]) > preBytes(p[0])))
        big = 1;
    const int small = 1 - big;
    KmerDatabase db(p[big]);
    const auto t0 = chrono::steady_clock::now();
    db.initGpuIndex(KmerDatabase::TableUse::Lookup);
    const auto t1 = chrono::steady_clock::now();
    uint64_t count = 0;
    hipibs::check(hipibs_table_probe_count_kmc(db.table(), p[small].c_str(),
                                               KmerDatabase::gpuOptions().loadThreads, &count),
                  "counting the smaller database's k-mers");
    const auto t2 = chrono::steady_clock::now();
    cerr << "[gpu] intersect: table " << chrono::duration<double>(t1 - t0).count() << " s, "
         << h[small].total_kmers << " queries streamed in "
         << chrono::duration<double>(t2 - t1).count() << " s" << endl;
    fprintf(stdout, "%llu\n", (unsigned long long)count);
    fflush(stdout);
    return 0;
}

int main(int argc, char **argv)
{
    // all chatter (including the library wrapper's) goes to stderr
    cout.rdbuf(cerr.rdbuf());
    try
    {
        return run(argc, argv);
    }
    catch (const exception &e)
    {
        cerr << "Error: " << e.what() << endl;
        return 1;
    }
}
