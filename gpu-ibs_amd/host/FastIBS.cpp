// fastibs: the per-window IBS table of one sample (or a panel) against
// every reference in a folder (the reference's src/FastIBS.cpp).
#include "CliInput.hpp"
#include "KmerPanel.hpp"

using namespace std;

static void usage(const char *argv0)
{
    cout << "\nFastIBS - IBS Distance Calculator (MI355X)\n"
         << "Usage:\n  " << argv0
         << " <sourcePath> <referencePath> <resultsFolder> <windowSize> [options]\n\n"
         << "  <sourcePath>     folder with a KMC database\n"
         << "  <referencePath>  folder with reference genomes (.fasta/.fa/.fna[.gz])\n"
         << "  <resultsFolder>  output folder\n"
         << "  <windowSize>     window length (integer larger than k)\n\n"
         << cli::optionsUsage()
         << "\nOutput: <db>_v_<reference>_<W>.tsv with columns seqname, start, end,\n"
         << "total_kmers, observed_kmers, variations, kmer_distance. Existing outputs are\n"
         << "skipped; failures are logged to log.txt and the exit status is 1.\n";
}

static int run(int argc, char **argv)
{
    if (cli::wantsVersion(argc, argv))
    {
        cout << "fastibs (hipibs, MI355X)" << endl;
        return 0;
    }
    if (cli::wantsHelp(argc, argv))
    {
        usage(argv[0]);
        return 0;
    }
    if (argc < 5)
    {
        usage(argv[0]);
        return 1;
    }
    string sourcePath = argv[1], referencePath = argv[2], resultsFolder = argv[3];
    long long W = 0;
    if (!cli::parseInt(argv[4], 1, INT_MAX, W))
    {
        cerr << "Invalid <windowSize>: " << argv[4] << " (expected a positive integer)" << endl;
        usage(argv[0]);
        return 1;
    }
    cli::Options opt;
    if (!cli::parseOptions(argc, argv, 5, opt, true))
    {
        usage(argv[0]);
        return 1;
    }
    cli::removeTrailingSlash(sourcePath);
    cli::removeTrailingSlash(referencePath);
    cli::removeTrailingSlash(resultsFolder);
    vector<string> dbDirs = opt.dbs;
    if (!opt.dbList.empty())
    {
        ifstream list(opt.dbList);
        if (!list)
        {
            cerr << "Error: cannot open --db-list file: " << opt.dbList << endl;
            return 1;
        }
        string line;
        while (getline(list, line))
        {
            const size_t a = line.find_first_not_of(" \t\r");
            if (a == string::npos || line[a] == '#')
                continue;
            const size_t b = line.find_last_not_of(" \t\r");
            dbDirs.push_back(line.substr(a, b - a + 1));
        }
    }
    dbDirs.insert(dbDirs.begin(), sourcePath);
    vector<string> prefixes, names;
    for (auto &d : dbDirs)
    {
        string d2 = d;
        cli::removeTrailingSlash(d2);
        string prefix;
        if (!cli::resolveDatabase(d2, prefix))
            return 1;
        prefixes.push_back(prefix);
        names.push_back(filesystem::path(d2).filename().string());
    }
    vector<filesystem::path> refs;
    if (!cli::listReferences(referencePath, refs))
        return 1;
    if (!cli::prepareResultsFolder(resultsFolder))
        return 1;
    // One sample: the streamed drop-in path (pinned, double-buffered batches
    // of the reference). Several: the panel path (reference resident on the
    // device, memory-budgeted rounds of tables, one graph replay per round).
    unique_ptr<KmerDatabase> single;
    unique_ptr<KmerPanel> panel;
    uint32_t k = 0;
    if (prefixes.size() == 1)
    {
        single = make_unique<KmerDatabase>(prefixes[0]);
        single->printKMCInfo();
        k = single->getKmerSize();
    }
    else
    {
        panel = make_unique<KmerPanel>(prefixes, opt.resident);
        k = panel->getKmerSize();
    }
    if (!cli::windowFitsK(W, k))
    {
        cerr << "Error: <windowSize> " << W << " must be larger than the database's k (" << k
             << ")" << endl;
        return 1;
    }
    int failed = 0;
    const auto t0 = chrono::steady_clock::now();
    for (const auto &ref : refs)
    {
        const string refName = cli::referenceName(ref.filename().string());
        cout << "Processing reference: " << refName << endl;
        vector<string> outs(prefixes.size());
        bool any = false;
        for (size_t i = 0; i < prefixes.size(); ++i)
        {
            const string out = resultsFolder + "/" + names[i] + "_v_" + refName + "_" +
                               to_string(W) + ".tsv";
            if (filesystem::exists(out))
                cout << "File already exists, skipping... " << out << endl;
            else
            {
                outs[i] = out;
                any = true;
            }
        }
        if (!any)
            continue;
        try
        {
            if (single)
                single->processReference(ref.string(), outs[0], (int)W);
            else
                panel->processReference(ref.string(), outs, (int)W);
        }
        catch (const exception &e)
        {
            ++failed;
            cli::logReferenceError(refName, e);
        }
    }
    cout << "Reference processed in "
         << chrono::duration<double>(chrono::steady_clock::now() - t0).count() << " seconds"
         << endl;
    if (failed)
        cerr << failed << " reference(s) failed; see log.txt" << endl;
    return failed ? 1 : 0;
}

int main(int argc, char **argv)
{
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
