"""The k-mer counter (hipibs_counter_*, bin/kmc) against the oracle's
counting (corpus.count_kmers: KMC's rules, pinned in test_oracle.py) and,
through tables it builds, against the reference's golden TSVs. Bit-exact:
keys and counts are integers."""
import gzip
import os
import subprocess
import sys

import numpy as np
import pytest

import corpus
from util import ROOT, sha256

import hipibs

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "gpu-ibs_amd", "bin")
ALL = 0xFFFFFFFF


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fastq_bytes(reads, lower=False):
    """FASTQ of code rows; quality strings cycle through '!'..'I', so some
    quality lines start with '+' or '@'."""
    out = []
    for i, r in enumerate(reads):
        s = corpus.to_text(r, lower=lower)
        q = bytes(33 + (i + j) % 41 for j in range(len(s)))
        out.append(b"@r%d\n%s\n+\n%s\n" % (i, s, q))
    return b"".join(out)


def fasta_bytes(records, width=60):
    return corpus.fasta([(b"c%d" % i, corpus.to_text(r)) for i, r in enumerate(records)], width)


def ragged_counts(records, k, ci=1, cs=ALL):
    """count_kmers over records of any lengths."""
    codes = []
    for r in records:
        if r.size >= k:
            c, v = corpus.canonical_codes(r[None, :], k)
            codes.append(c[v])
    allc = np.concatenate(codes) if codes else np.zeros(0, np.uint64)
    keys, counts = np.unique(allc, return_counts=True)
    keep = counts >= ci
    return keys[keep], np.minimum(counts[keep], cs).astype(np.uint32), allc.size


def sorted_pairs(keys, counts):
    o = np.argsort(keys, kind="stable")
    return keys[o], counts[o]


@pytest.fixture(scope="module")
def small_reads():
    return corpus.small_reads()


def test_counts_match_oracle_fastq(tmp_path, small_reads):
    _, reads = small_reads
    r = reads[0]
    path = tmp_path / "s0.fq"
    path.write_bytes(fastq_bytes(r))
    note("count s0 fastq")
    c = hipibs.KmerCounter(25)
    c.add_file(str(path))
    st = c.stats()
    want_k, want_c = corpus.count_kmers(r, 25, ci=1, cs=ALL)
    _, valid = corpus.canonical_codes(r, 25)
    assert st["reads"] == r.shape[0]
    assert st["kmers"] == int(valid.sum())
    assert st["distinct"] == want_k.size
    keys, counts, filt = c.keys(1, ALL, ALL)
    keys, counts = sorted_pairs(keys, counts)
    np.testing.assert_array_equal(keys, want_k)
    np.testing.assert_array_equal(counts, want_c)
    assert filt == {"kept": want_k.size, "below": 0, "above": 0}
    c.free()


def test_reads_to_table_scan_matches_golden(tmp_path, small_reads, expected):
    """reads -> GPU counts (-ci2) -> device table -> scan == the reference's
    TSVs for the three samples; the KMC database the counter writes gives
    the same table."""
    ref, reads = small_reads
    seq = corpus.to_text(ref)
    for i, r in enumerate(reads):
        note(f"reads -> table s{i}")
        path = tmp_path / f"s{i}.fq.gz"
        path.write_bytes(gzip.compress(fastq_bytes(r), compresslevel=1))
        c = hipibs.KmerCounter(25)
        c.add_file(str(path))
        t = c.build_table(2, ALL)
        assert t.info()["n_keys"] == expected["small"][f"s{i}"]["n_keys"]
        prefix = str(tmp_path / f"db{i}")
        f = c.write_kmc(prefix, 2, 1_000_000_000, 255)
        assert f["kept"] == expected["small"][f"s{i}"]["n_keys"]
        c.free()
        info = hipibs.kmc_header(prefix)
        assert (info.kmer_length, info.counter_size, info.min_count, info.total_kmers) == (
            25, 1, 2, f["kept"])
        t2 = hipibs.Table.from_kmc(prefix)
        with open(os.path.join(ROOT, "tests", "golden", f"small_s{i}_50000.tsv"), "rb") as fh:
            golden = fh.read()
        for tab in (t, t2):
            rows = hipibs.scan_sequences([tab], [b"chr1"], [seq], 50000)[0]
            assert hipibs.format_tsv(rows) == golden
            tab.free()


@pytest.mark.parametrize("chunk", [64, 4096, 1 << 20])
def test_chunk_seams_and_growth(tmp_path, chunk):
    """Long multi-line FASTA records cut by small chunks, a table grown
    from its smallest size many times: the counts do not change."""
    rng = np.random.default_rng(chunk)
    recs = []
    for n in (200_000, 31, 32, 0, 90_001, 5):
        r = rng.integers(0, 4, n).astype(np.uint8)
        if n > 1000:
            r[rng.integers(0, n, n // 500)] = 4
            r[1000:1100] = 4
        recs.append(r)
    path = tmp_path / "long.fa"
    path.write_bytes(fasta_bytes(recs))
    k = 31
    want_k, want_c, total = ragged_counts(recs, k)
    note(f"chunk {chunk}")
    c = hipibs.KmerCounter(k, chunk_bytes=chunk)
    c.add_file(str(path))
    st = c.stats()
    assert st["kmers"] == total and st["distinct"] == want_k.size
    # a table sized from the sketch: probes rarely leave their slice
    assert st["spilled"] <= total // 50, st
    assert st["reads"] == 6  # the empty record counts as a read with no k-mers
    keys, counts, _ = c.keys(1, ALL, ALL)
    keys, counts = sorted_pairs(keys, counts)
    np.testing.assert_array_equal(keys, want_k)
    np.testing.assert_array_equal(counts, want_c)
    c.free()


@pytest.mark.parametrize("k", [1, 9, 15, 32])
def test_edge_reads(tmp_path, k):
    """Reads shorter than k, all N, lowercase, CRLF line ends (the '\\r' is
    a non-base byte), text passed directly, every k in range."""
    rng = np.random.default_rng(100 + k)
    reads = [rng.integers(0, 4, int(n)).astype(np.uint8) for n in rng.integers(0, 120, 300)]
    reads[3] = np.full(50, 4, np.uint8)
    reads[4] = np.zeros(k - 1 if k > 1 else 0, np.uint8)
    reads[5] = np.zeros(k, np.uint8)
    lines = []
    for i, r in enumerate(reads):
        s = corpus.to_text(r, lower=bool(i % 2))
        lines.append(b"@q%d\r\n%s\r\n+\r\n%s\r\n" % (i, s, b"I" * len(s)))
    path = tmp_path / "crlf.fq"
    path.write_bytes(b"".join(lines))
    want_k, want_c, total = ragged_counts(reads, k)
    c = hipibs.KmerCounter(k, chunk_bytes=1000)
    c.add_file(str(path))
    # the same reads once more, as one text blob of '\n'-separated reads
    c.add_text(b"\n".join(corpus.to_text(r) for r in reads))
    st = c.stats()
    assert st["kmers"] == 2 * total
    keys, counts, _ = c.keys(1, ALL, ALL)
    keys, counts = sorted_pairs(keys, counts)
    np.testing.assert_array_equal(keys, want_k)
    np.testing.assert_array_equal(counts, 2 * want_c)
    c.free()


@pytest.mark.parametrize("chunk", [64, 1000, 0])
def test_device_text_equals_oracle(small_reads, chunk):
    """Text already in HBM, counted in aligned overlapping pieces; a
    misaligned pointer is refused."""
    _, reads = small_reads
    r = reads[0][:3000]
    text = b"\n".join(corpus.to_text(x) for x in r)
    want_k, want_c = corpus.count_kmers(r, 25, ci=1, cs=ALL)
    _, valid = corpus.canonical_codes(r, 25)
    buf = hipibs.DeviceBuffer.from_array(np.frombuffer(text, np.uint8))
    c = hipibs.KmerCounter(25, chunk_bytes=chunk)
    c.add_text_device(buf.ptr, len(text))
    st = c.stats()
    assert st["kmers"] == int(valid.sum())
    # a flush holds 16 chunks of k-mers: small chunks count in many rounds
    # into a table that already holds keys
    assert (st["flushes"] > 100) if chunk == 64 else (st["flushes"] >= 1)
    keys, counts, _ = c.keys(1, ALL, ALL)
    keys, counts = sorted_pairs(keys, counts)
    np.testing.assert_array_equal(keys, want_k)
    np.testing.assert_array_equal(counts, want_c)
    with pytest.raises(hipibs.HipibsError, match="aligned"):
        c.add_text_device(buf.ptr + 3, 100)
    c.free()
    buf.free()


@pytest.mark.parametrize("chunk", [1000, 0])
def test_spilled_kmers_counted_in_global_table(small_reads, monkeypatch, chunk):
    """A table sized far below the sketch's estimate: most k-mers' probes
    leave their slice and are counted by the spill path in the global
    table, which grows for them; the counts stay exact. With 1000 B chunks
    the spills of later rounds meet keys stored by earlier ones."""
    _, reads = small_reads
    r = reads[1][:4000]
    text = b"\n".join(corpus.to_text(x) for x in r)
    want_k, want_c = corpus.count_kmers(r, 25, ci=1, cs=ALL)
    monkeypatch.setenv("HIPIBS_COUNT_SKETCH_SCALE", "0")
    c = hipibs.KmerCounter(25, chunk_bytes=chunk)
    monkeypatch.delenv("HIPIBS_COUNT_SKETCH_SCALE")
    half = len(text) // 2
    half = text.index(b"\n", half)
    c.add_text(text[:half])
    c.add_text(text[half + 1:])
    st = c.stats()
    note(f"spill test chunk {chunk}: {st}")
    assert st["spilled"] > (st["kmers"] // 4 if chunk == 0 else 0)
    keys, counts, _ = c.keys(1, ALL, ALL)
    keys, counts = sorted_pairs(keys, counts)
    np.testing.assert_array_equal(keys, want_k)
    np.testing.assert_array_equal(counts, want_c)
    assert st["distinct"] == want_k.size
    c.free()


def slice_passes(table_bytes, digit_bits):
    """Passes of the counter's slice sort for a table of table_bytes (128 B
    buckets, 512-bucket slices, digits of digit_bits top bits)."""
    sbits = max(0, (table_bytes // 128).bit_length() - 1 - 9)
    return max(1, -(-sbits // digit_bits))


@pytest.mark.parametrize("bits,wide", [(9, "0"), (9, "1"), (4, "0"), (5, "1")])
def test_slice_sort_passes_and_lookback_words(small_reads, monkeypatch, bits, wide):
    """The slice sort over two or three digits (every pass after the first
    ranks stably) with 32- or 64-bit look-back words. The table is sized for
    every hashed k-mer (1.5 M, 2^19 buckets: 10 slice bits); with 512 values
    per digit a tile of the second pass spans several values of the first
    pass's digit, so a rank that is not stable would misplace k-mers. Keys
    and counts equal the oracle's."""
    _, reads = small_reads
    r = reads[2][:12000]
    text = b"\n".join(corpus.to_text(x) for x in r)
    want_k, want_c = corpus.count_kmers(r, 25, ci=1, cs=ALL)
    monkeypatch.setenv("HIPIBS_SORT_DIGIT_BITS", str(bits))
    monkeypatch.setenv("HIPIBS_SORT_LOOKBACK_WIDE", wide)
    monkeypatch.setenv("HIPIBS_COUNT_SKETCH_SCALE", "40")
    c = hipibs.KmerCounter(25)
    for v in ("HIPIBS_SORT_DIGIT_BITS", "HIPIBS_SORT_LOOKBACK_WIDE", "HIPIBS_COUNT_SKETCH_SCALE"):
        monkeypatch.delenv(v)
    c.add_text(text)
    st = c.stats()
    passes = slice_passes(st["table_bytes"], bits)
    note(f"slice sort: digit bits {bits}, wide look-back {wide}, passes {passes}: {st}")
    assert passes >= 2
    keys, counts, _ = c.keys(1, ALL, ALL)
    keys, counts = sorted_pairs(keys, counts)
    np.testing.assert_array_equal(keys, want_k)
    np.testing.assert_array_equal(counts, want_c)
    c.free()


def test_repetitive_reads_fill_few_digits(monkeypatch):
    """Reads drawn from 40 short sequences: most digit values are empty,
    a few hold every hash, and tiles of the later passes hold runs of one
    value; with 3-bit digits the sort takes three passes."""
    seqs = [corpus.random_bases(0x5EED, 70 + i, 150) for i in range(40)]
    pick = corpus.splitmix(0x5EED, 200, 30000) % np.uint64(40)
    r = np.stack([seqs[int(i)] for i in pick])
    text = b"\n".join(corpus.to_text(x) for x in r)
    want_k, want_c = corpus.count_kmers(r, 21, ci=1, cs=ALL)
    monkeypatch.setenv("HIPIBS_SORT_DIGIT_BITS", "3")
    monkeypatch.setenv("HIPIBS_COUNT_SKETCH_SCALE", "100")
    c = hipibs.KmerCounter(21)
    monkeypatch.delenv("HIPIBS_SORT_DIGIT_BITS")
    monkeypatch.delenv("HIPIBS_COUNT_SKETCH_SCALE")
    c.add_text(text)
    st = c.stats()
    note(f"repetitive reads: passes {slice_passes(st['table_bytes'], 3)}: {st}")
    assert slice_passes(st["table_bytes"], 3) == 3
    keys, counts, _ = c.keys(1, ALL, ALL)
    keys, counts = sorted_pairs(keys, counts)
    np.testing.assert_array_equal(keys, want_k)
    np.testing.assert_array_equal(counts, want_c)
    c.free()


def test_hash_buffers_reused_across_counters(small_reads):
    """Counters alive together each keep their own buffers; a freed
    counter's hash buffers and table serve the next counter, larger or
    smaller (a kept table still holds the last counter's keys and counts),
    and hipibs_counter_release_cache frees them between counters: every
    count equals the oracle's."""
    _, reads = small_reads
    sets = [reads[0][:6000], reads[1][:1500], reads[2][:9000], reads[0][:500]]
    texts = [b"\n".join(corpus.to_text(x) for x in r) for r in sets]
    wants = [corpus.count_kmers(r, 25, ci=1, cs=ALL) for r in sets]

    def check(c, i):
        keys, counts, _ = c.keys(1, ALL, ALL)
        keys, counts = sorted_pairs(keys, counts)
        np.testing.assert_array_equal(keys, wants[i][0])
        np.testing.assert_array_equal(counts, wants[i][1])

    a, b = hipibs.KmerCounter(25), hipibs.KmerCounter(25)
    a.add_text(texts[0])
    b.add_text(texts[1])
    check(a, 0)
    check(b, 1)
    a.free()
    c = hipibs.KmerCounter(25)  # takes a's buffers
    c.add_text(texts[2])
    check(c, 2)
    b.free()
    c.free()
    e = hipibs.KmerCounter(25)  # the largest kept table, full of c's counts
    e.add_text(texts[1])
    check(e, 1)
    e.free()
    hipibs._lib.hipibs_counter_release_cache()
    d = hipibs.KmerCounter(25)
    d.add_text(texts[3])
    check(d, 3)
    d.free()
    hipibs._lib.hipibs_counter_release_cache()
    hipibs._lib.hipibs_counter_release_cache()  # nothing kept: no-op


def test_filters(small_reads):
    _, reads = small_reads
    r = reads[2][:20000]
    c = hipibs.KmerCounter(25, expected_keys=10)
    c.add_text(b"\n".join(corpus.to_text(x) for x in r))
    all_k, all_c = corpus.count_kmers(r, 25, ci=1, cs=ALL)
    for ci, cx, cs in ((2, ALL, 255), (3, 7, 5), (1, 1, 1), (50, ALL, 255)):
        keys, counts, f = c.keys(ci, cx, cs)
        keys, counts = sorted_pairs(keys, counts)
        keep = (all_c >= ci) & (all_c <= cx)
        np.testing.assert_array_equal(keys, all_k[keep])
        np.testing.assert_array_equal(counts, np.minimum(all_c[keep], cs))
        assert f == {"kept": int(keep.sum()), "below": int((all_c < ci).sum()),
                     "above": int((all_c > cx).sum())}
    c.free()


def test_kmc_command(tmp_path, small_reads):
    """bin/kmc with the reference README's options over an @list of gzip
    FASTQ files: one database of both files' k-mers."""
    _, reads = small_reads
    halves = (reads[1][:30000], reads[1][30000:])
    names = []
    for i, h in enumerate(halves):
        p = tmp_path / f"part{i}.fastq.gz"
        p.write_bytes(gzip.compress(fastq_bytes(h), compresslevel=1))
        names.append(str(p))
    lst = tmp_path / "S1.reads"
    lst.write_text("\n".join(names) + "\n")
    (tmp_path / "tmp").mkdir()
    out = str(tmp_path / "S1")
    js = str(tmp_path / "S1.json")
    r = subprocess.run([os.path.join(BIN, "kmc"), "-k25", "-ci2", "-cs255", "-t16", "-m32",
                        f"-j{js}", f"@{lst}", out, str(tmp_path / "tmp")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    want_k, _ = corpus.count_kmers(reads[1], 25, ci=2, cs=255)
    assert "No. of unique counted k-mers       : %12d" % want_k.size in r.stdout
    np.testing.assert_array_equal(np.sort(hipibs.kmc_read_keys(out)), want_k)
    info = hipibs.kmc_header(out)
    assert (info.kmc_version, info.signature_len, info.counter_size) == (0x200, 9, 1)
    assert os.path.getsize(js) > 0


def test_counter_errors(tmp_path):
    for k in (0, 65):
        with pytest.raises(hipibs.HipibsError):
            hipibs.KmerCounter(k)
    c = hipibs.KmerCounter(25)
    with pytest.raises(hipibs.HipibsError, match="cannot open"):
        c.add_file(str(tmp_path / "missing.fq"))
    c.free()
    bad = tmp_path / "bad.txt"
    bad.write_bytes(b"ACGT\n")
    c = hipibs.KmerCounter(25)
    with pytest.raises(hipibs.HipibsError, match="unknown file type"):
        c.add_file(str(bad))
    with pytest.raises(hipibs.HipibsError, match="earlier call failed"):
        c.add_text(b"ACGT")
    c.free()
    c = hipibs.KmerCounter(25)
    with pytest.raises(hipibs.HipibsError):
        c.keys(5, 2, 255)
    c.free()


def wide_counts(records, k):
    """(keys (n, 2), counts, k-mers) of records of any lengths, k = 33..64."""
    rows = []
    for r in records:
        if r.size >= k:
            h, l, v = corpus.canonical_codes_wide(r[None, :], k)
            rows.append(np.stack([h[v], l[v]], axis=1))
    allr = np.concatenate(rows) if rows else np.zeros((0, 2), np.uint64)
    keys, counts = np.unique(allr, axis=0, return_counts=True)
    return keys, counts.astype(np.uint32), allr.shape[0]


def sorted_rows(keys, counts):
    o = np.lexsort((keys[:, 1], keys[:, 0]))
    return keys[o], counts[o]


@pytest.mark.parametrize("k", [33, 47, 64])
def test_wide_counts_match_oracle(tmp_path, k):
    """k = 33..64: FASTQ reads of a small genome (so k-mers repeat) cut into
    64 B pieces with many table growths; at k = 64 the one code whose low
    word is all ones (A^32 T^32) and the one whose high word is all ones
    (T^32 A^32) are among them."""
    rng = np.random.default_rng(k)
    genome = rng.integers(0, 4, 20000).astype(np.uint8)
    reads = []
    for i in range(1500):
        ln = int(rng.integers(40, 150))
        st = int(rng.integers(0, genome.size - ln))
        r = genome[st:st + ln].copy()
        if i % 2:
            r = (3 - r)[::-1].copy()
        if i % 9 == 0:
            r[int(rng.integers(0, ln))] = 4
        reads.append(r)
    pal = np.array([0] * 32 + [3] * 32, dtype=np.uint8)  # A^32 T^32, its own reverse complement
    top = np.array([3] * 32 + [0] * 32, dtype=np.uint8)  # T^32 A^32, likewise
    for j in range(3):
        reads.append(np.concatenate([genome[:10], pal, genome[j:j + 5]]))
    for j in range(4):
        reads.append(np.concatenate([genome[j:j + 7], top, genome[:3]]))
    path = tmp_path / "r.fq"
    path.write_bytes(b"".join(b"@r%d\n%s\n+\n%s\n" % (i, corpus.to_text(r), b"I" * r.size)
                              for i, r in enumerate(reads)))
    want_k, want_c, total = wide_counts(reads, k)
    note(f"wide count k={k}")
    c = hipibs.KmerCounter(k, chunk_bytes=64)
    c.add_file(str(path))
    st = c.stats()
    assert (st["kmers"], st["distinct"]) == (total, want_k.shape[0])
    keys, counts, _ = c.keys(1, ALL, ALL)
    keys, counts = sorted_rows(keys, counts)
    np.testing.assert_array_equal(keys, want_k)
    np.testing.assert_array_equal(counts, want_c)
    if k == 64:
        ones = np.uint64(2 ** 64 - 1)
        i = np.flatnonzero((want_k[:, 0] == 0) & (want_k[:, 1] == ones))
        assert i.size == 1 and want_c[i[0]] >= 3
        j = np.flatnonzero((want_k[:, 0] == ones) & (want_k[:, 1] == 0))
        assert j.size == 1 and want_c[j[0]] >= 4
        # both travel into a table built from the counts, and probe as present
        t = c.build_table(1, ALL)
        assert t.info()["n_keys"] == want_k.shape[0]
        assert t.probe_count(want_k[np.concatenate([i, j])]) == 2
        t.free()
    f = c.keys(2, 5, 3)[2]
    assert f["kept"] == int(((want_c >= 2) & (want_c <= 5)).sum())
    c.free()


def test_wide_counter_to_table_matches_reference_string_path(tmp_path, expected):
    """k = 47: the edge corpus's sample sequences counted (-ci1) give the
    golden key set; their table, built directly and through the KMC database
    the counter writes, scans to the reference string path's TSV."""
    k = 47
    recs, sample = corpus.edge_sequences(k)
    _, keys = corpus.edge_config(k)
    path = tmp_path / "sample.fa"
    path.write_bytes(corpus.fasta([(b"s%d" % i, corpus.to_text(s)) for i, s in enumerate(sample)]))
    c = hipibs.KmerCounter(k)
    c.add_file(str(path))
    got, _, _ = c.keys(1, ALL, ALL)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    np.testing.assert_array_equal(got, keys)
    t = c.build_table(1, ALL)
    prefix = str(tmp_path / "db")
    c.write_kmc(prefix, 1, 1_000_000_000, 255)
    c.free()
    t2 = hipibs.Table.from_kmc(prefix)
    want = expected["edge_wide"][str(k)]["tsv"]["1000"]
    ids = [r[0] for r in recs]
    seqs = [r[1] for r in recs]
    for tab in (t, t2):
        tsv = hipibs.format_tsv(hipibs.scan_sequences([tab], ids, seqs, 1000)[0])
        assert (sha256(tsv), len(tsv)) == (want["sha256"], want["bytes"])
        tab.free()


def test_kmc_command_k35(tmp_path):
    """`kmc -k35 -ci1 -fm` as the reference's test script runs it for its
    k > 32 corpus (scripts/run_tests.sh:70)."""
    recs, sample = corpus.edge_sequences(35)
    fa = tmp_path / "sample_reads.fasta"
    fa.write_bytes(corpus.fasta([(b"s%d" % i, corpus.to_text(s)) for i, s in enumerate(sample)]))
    (tmp_path / "tmp").mkdir()
    out = str(tmp_path / "sample")
    r = subprocess.run([os.path.join(BIN, "kmc"), "-k35", "-ci1", "-fm", "-t4", str(fa), out,
                        str(tmp_path / "tmp")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    _, keys = corpus.edge_config(35)
    back = hipibs.kmc_read_keys(out)
    np.testing.assert_array_equal(np.unique(back, axis=0), keys)
    assert hipibs.kmc_header(out).kmer_length == 35
