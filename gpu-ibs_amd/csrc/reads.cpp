// Streaming reader of read files for the k-mer counter: FASTA (one or
// many lines per record) and FASTQ, plain or gzip (zlib reads both), handed
// to a sink as the bases of each record followed by the record's end. Lines
// of any length are streamed without being held whole.
//
// Record rules (what KMC 3.2.4, which the reference runs before the scan,
// counts: README.md:95, scripts/make_test_corpus.sh:123):
//   FASTA  a '>' line starts a record; its sequence is the concatenation of
//          the following lines up to the next '>' line.
//   FASTQ  an '@' line starts a record; sequence lines follow up to a line
//          starting with '+'; quality lines are then skipped until they hold
//          as many characters as the sequence (so a quality line may start
//          with '@' or '+').
//   Lines are split on '\n' only and empty lines are skipped; the format is
//   decided by the first byte of the file.
#include "hipibs_internal.hpp"

#include <zlib.h>

#include <cstring>
#include <vector>

namespace hipibs
{

int readRecords(const std::string &path, ReadSink &sink, uint64_t *nRecords)
{
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f)
        return setError(HIPIBS_E_IO, "cannot open " + path);
    gzbuffer(f, 1u << 20);
    std::vector<char> buf(8u << 20);

    enum Line
    {
        kHeader,  // skip to the end of the line
        kSeq,     // sequence bytes
        kPlus,    // the '+' line
        kQual     // quality bytes, counted
    };
    int fmt = 0; // '>' or '@' once known
    bool atStart = true, inRecord = false;
    Line line = kHeader;
    uint64_t seqLen = 0, qualLen = 0, records = 0;
    int rc = 0;
    for (;;)
    {
        const int got = gzread(f, buf.data(), (unsigned)buf.size());
        if (got < 0)
        {
            int err = 0;
            const std::string msg = gzerror(f, &err);
            rc = setError(HIPIBS_E_IO, "read error in " + path + ": " + msg);
            break;
        }
        if (got == 0)
            break;
        const char *p = buf.data(), *end = p + got;
        while (p < end)
        {
            if (atStart)
            {
                if (*p == '\n')
                {
                    ++p; // empty line
                    continue;
                }
                if (!fmt)
                {
                    if (*p != '>' && *p != '@')
                    {
                        rc = setError(HIPIBS_E_IO, path + ": unknown file type (not FASTA or FASTQ)");
                        break;
                    }
                    fmt = *p;
                }
                if (fmt == '>')
                    line = *p == '>' ? kHeader : kSeq;
                else if ((line == kPlus || line == kQual) && qualLen < seqLen)
                    line = kQual; // a quality line, whatever its first byte
                else if (*p == '@')
                    line = kHeader;
                else if (*p == '+' && inRecord && (line == kHeader || line == kSeq))
                    line = kPlus;
                else if (inRecord && (line == kHeader || line == kSeq))
                    line = kSeq;
                else
                {
                    rc = setError(HIPIBS_E_IO, path + ": malformed FASTQ record " +
                                                   std::to_string(records + 1));
                    break;
                }
                if (line == kHeader)
                {
                    if (inRecord)
                    {
                        sink.recordEnd();
                        ++records;
                    }
                    inRecord = true;
                    seqLen = qualLen = 0;
                }
                atStart = false;
            }
            const char *nl = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
            const char *stop = nl ? nl : end;
            const size_t n = (size_t)(stop - p);
            if (line == kSeq && n)
            {
                sink.bases(p, n);
                seqLen += n;
            }
            else if (line == kQual)
                qualLen += n;
            if (sink.failed())
            {
                rc = sink.failed();
                break;
            }
            p = stop;
            if (nl)
            {
                ++p;
                atStart = true;
            }
        }
        if (rc)
            break;
    }
    gzclose(f);
    if (rc)
        return rc;
    if (!fmt)
        return setError(HIPIBS_E_IO, path + ": empty file");
    if (inRecord)
    {
        sink.recordEnd();
        ++records;
    }
    if (sink.failed())
        return sink.failed();
    if (nRecords)
        *nRecords = records;
    return 0;
}

} // namespace hipibs
