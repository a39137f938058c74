// fastibsmapper: per-base k-mer coverage of every reference in a folder
// (the reference's src/FastIBSMapper.cpp): <db>_v_<stem>.txt, one id line
// and one comma-separated coverage line per contig.
#include "CliInput.hpp"

using namespace std;

static void usage(const char *argv0)
{
    cout << "\nFastIBS - Reference Mapping Tool (MI355X)\nUsage:\n  " << argv0
         << " <sourcePath> <referencePath> <resultsFolder> [options]\n\n"
         << cli::optionsUsage();
}

static int run(int argc, char **argv)
{
    if (cli::wantsVersion(argc, argv))
    {
        cout << "fastibsmapper (hipibs, MI355X)" << endl;
        return 0;
    }
    if (cli::wantsHelp(argc, argv))
    {
        usage(argv[0]);
        return 0;
    }
    if (argc < 4)
    {
        usage(argv[0]);
        return 1;
    }
    string sourcePath = argv[1], referencePath = argv[2], resultsFolder = argv[3];
    cli::Options opt;
    if (!cli::parseOptions(argc, argv, 4, opt, false))
    {
        usage(argv[0]);
        return 1;
    }
    cli::removeTrailingSlash(sourcePath);
    cli::removeTrailingSlash(referencePath);
    cli::removeTrailingSlash(resultsFolder);
    string prefix;
    if (!cli::resolveDatabase(sourcePath, prefix))
        return 1;
    vector<filesystem::path> refs;
    if (!cli::listReferences(referencePath, refs))
        return 1;
    if (!cli::prepareResultsFolder(resultsFolder))
        return 1;
    const string database = filesystem::path(sourcePath).filename().string();
    KmerDatabase db(prefix);
    db.printKMCInfo();
    int failed = 0;
    for (const auto &ref : refs)
    {
        const string stem = filesystem::path(cli::stripGz(ref.filename().string())).stem().string();
        const string out = resultsFolder + "/" + database + "_v_" + stem + ".txt";
        cout << "Processing reference: " << stem << endl;
        if (filesystem::exists(out))
        {
            cout << "File already exists, skipping..." << endl;
            continue;
        }
        try
        {
            db.produceMapping(ref.string(), out);
        }
        catch (const exception &e)
        {
            ++failed;
            cli::logReferenceError(stem, e);
        }
    }
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
