// Reader of read files for the k-mer counter (reads.cpp). Free of HIP so
// that host-only programs can use it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

namespace hipibs
{

// Receiver of the records of a read file: the bases of a record arrive in
// one or more bases() calls, then recordEnd(). failed() != 0 stops the
// reader, which returns that status.
struct ReadSink
{
    virtual ~ReadSink() = default;
    virtual void bases(const char *p, size_t n) = 0;
    virtual void recordEnd() = 0;
    virtual int failed() const = 0;
};

// Streams the records of a FASTA or FASTQ file, plain or gzip, into sink;
// 0 or a HIPIBS_E_* status (message in hipibs_last_error()).
int readRecords(const std::string &path, ReadSink &sink, uint64_t *nRecords);

} // namespace hipibs
