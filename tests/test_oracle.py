"""Pins the oracle (oracle/oracle.c) before anything is checked against it.

Each oracle function is checked against something independent of it:
  orc_encode_positions  brute force: substring -> reverse complement ->
                        lexicographic min, the string oracle of
                        tests/phase0_test.cpp:39-56 (k in 4..32, lowercase,
                        non-ACGT bytes, short inputs)
  orc_windows           the closed-form loop of src/Utils.hpp:310 and the
                        reference's own windows in the golden TSVs
  orc_window_stats /    the golden TSVs written by the reference's code
  orc_scan_contig       (CPU-OPT and CPU-BATCHED agree), the closed form of
                        the distance term for every gap 1..3k, invariants
  orc_stats_from_flags  equality with orc_window_stats on the same input
  orc_coverage          the golden coverage SHA-256s (reference code) and a
                        brute-force k-increment count (tests/phase0_test.cpp:140-155)
  orc_intersect         the golden KDBIntersect counts and a Python set
  orc_kmc_read          keys written by the KMC writer, read back and listed
                        by the reference's KmcRaIndex when the databases were made
  orc_*_wide            (k = 33..64) the string oracle for k up to 64; the
                        golden TSVs and coverage of the reference's string
                        path (getCanonicalKmersLegacy + CKMCFile::IsKmer on
                        strings) at k = 33, 47, 64
  orc_count_kmers /     (KMC's counting rules, the k-mer counter's oracle) a
  corpus.count_kmers    dictionary count over substrings of each read's text
                        with the string canonical form; each other on the
                        small samples' reads; the golden key counts of the
                        three samples, whose databases the reference's code read
"""
import os

import numpy as np
import pytest

import corpus
from util import (REF_PATH, ROOT, coverage_line, oracle, oracle_count, oracle_coverage_text,
                  oracle_coverage_text_wide, oracle_scan_rows, oracle_scan_rows_wide, sha256,
                  sorted_keys, sorted_keys_wide)

import hipibs

COMP = bytes.maketrans(b"ACGT", b"TGCA")


def string_canonical(s: bytes, k: int):
    """Per-position canonical code by string manipulation (no rolling)."""
    out = []
    code = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}
    for i in range(len(s) - k + 1):
        w = s[i:i + k].upper()
        if any(c not in b"ACGT" for c in w):
            out.append(None)
            continue
        rc = w.translate(COMP)[::-1]
        m = min(w, rc)
        v = 0
        for c in m:
            v = (v << 2) | code[c]
        out.append(v)
    return out


@pytest.mark.parametrize("k", [1, 4, 15, 25, 31, 32])
def test_codec_against_string_oracle(k):
    lib = oracle()
    rng = np.random.default_rng(20260830 + k)
    alphabet = np.frombuffer(b"ACGTACGTACGTacgtNnXY-", dtype=np.uint8)
    for length in [0, k - 1, k, k + 1, 57, 399]:
        if length < 0:
            continue
        s = alphabet[rng.integers(0, alphabet.size, length)].tobytes()
        np_ = max(0, length - k + 1)
        codes = np.zeros(max(np_, 1), dtype=np.uint64)
        valid = np.zeros(max(np_, 1), dtype=np.uint8)
        n = lib.orc_encode_positions(s, length, k, codes.ctypes.data, valid.ctypes.data)
        assert n == np_
        want = string_canonical(s, k)
        for i in range(np_):
            if want[i] is None:
                assert valid[i] == 0
            else:
                assert valid[i] == 1 and int(codes[i]) == want[i], (k, length, i)


def test_window_geometry_closed_form():
    lib = oracle()
    for L, W, k in [(1_000_000, 50000, 25), (100, 33, 25), (26, 27, 25), (0, 100, 25),
                    (49976, 50000, 25), (123457, 1000, 31)]:
        cap = L // max(1, W - k) + 3
        st = np.zeros(cap, dtype=np.int64)
        en = np.zeros(cap, dtype=np.int64)
        n = lib.orc_windows(L, W, k, st.ctypes.data, en.ctypes.data, cap)
        j = list(range(0, L, W - k))
        assert n == len(j)
        assert st[:n].tolist() == j
        assert en[:n].tolist() == [min(x + W, L) for x in j]
    assert lib.orc_windows(100, 25, 25, None, None, 0) == -1


def gap_distance(gap, k):
    d = gap - (k - 1)
    return abs(d + 1) if d <= 0 else d


@pytest.mark.parametrize("k", [15, 25, 31, 32])
def test_distance_term_closed_form(k):
    """A window of all-miss valid positions of length L has one run: the
    oracle must give (1, f(L)) with f from src/Window.hpp:29-36, and
    f(1) = k-3, f(k-2) = 0, f(k-1) = 1, f(g >= k) = g-k+1."""
    lib = oracle()
    assert gap_distance(1, k) == k - 3
    assert gap_distance(k - 2, k) == 0
    assert gap_distance(k - 1, k) == 1
    assert gap_distance(3 * k, k) == 2 * k + 1
    keys = np.zeros(1, dtype=np.uint64) + np.uint64(2 ** 62)  # matches nothing
    rng = np.random.default_rng(k)
    for L in range(1, 3 * k + 1):
        s = corpus.to_text(rng.integers(0, 4, L + k - 1).astype(np.uint8))
        out = np.zeros(4, dtype=np.int32)
        lib.orc_window_stats(s, len(s), k, keys.ctypes.data, 1, out.ctypes.data)
        assert out.tolist() == [L, 0, 1, gap_distance(L, k)]


def test_stats_from_flags_equals_codes_path():
    lib = oracle()
    rng = np.random.default_rng(5)
    k = 25
    codes_all = rng.integers(0, 4, 3000).astype(np.uint8)
    codes_all[rng.random(3000) < 0.01] = 4
    s = corpus.to_text(codes_all)
    np_ = len(s) - k + 1
    codes = np.zeros(np_, dtype=np.uint64)
    valid = np.zeros(np_, dtype=np.uint8)
    lib.orc_encode_positions(s, len(s), k, codes.ctypes.data, valid.ctypes.data)
    keys = sorted_keys(np.unique(codes[valid.astype(bool)][rng.random(int(valid.sum())) < 0.6]))
    a = np.zeros(4, dtype=np.int32)
    lib.orc_window_stats(s, len(s), k, keys.ctypes.data, keys.size, a.ctypes.data)
    hit = (np.isin(codes, keys) & valid.astype(bool)).astype(np.uint8)
    b = np.zeros(4, dtype=np.int32)
    lib.orc_stats_from_flags(valid.ctypes.data, hit.ctypes.data, np_, k, b.ctypes.data)
    assert a.tolist() == b.tolist()
    # invariants
    tot, obs, var, _ = a.tolist()
    assert obs <= tot and var <= min(tot - obs, obs + 1)


def test_small_config_tsv_matches_reference(small_corpus, expected):
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    for i, (keys, _) in enumerate(samples):
        rows = oracle_scan_rows([b"chr1"], [seq], 50000, 25, sorted_keys(keys))
        tsv = hipibs.format_tsv(rows)
        with open(os.path.join(ROOT, "tests", "golden", f"small_s{i}_50000.tsv"), "rb") as f:
            assert tsv == f.read()
        assert sha256(tsv) == expected["small"][f"s{i}"]["tsv_sha256"]


@pytest.mark.parametrize("k", [25, 31, 32, 15])
def test_edge_tsv_and_coverage_match_reference(k, expected):
    recs, keys = corpus.edge_config(k)
    ids = [r[0] for r in recs]
    seqs = [r[1] for r in recs]
    ks = sorted_keys(keys)
    e = expected["edge"][str(k)]
    assert e["n_keys"] == keys.size
    for W, want in e["tsv"].items():
        tsv = hipibs.format_tsv(oracle_scan_rows(ids, seqs, int(W), k, ks))
        assert (sha256(tsv), len(tsv)) == (want["sha256"], want["bytes"]), (k, W)
    cov = oracle_coverage_text(ids, seqs, k, ks)
    assert (sha256(cov), len(cov)) == (e["coverage_sha256"], e["coverage_bytes"])


def test_coverage_brute_force():
    lib = oracle()
    rng = np.random.default_rng(9)
    k = 7
    c = rng.integers(0, 4, 500).astype(np.uint8)
    c[rng.random(500) < 0.03] = 4
    s = corpus.to_text(c)
    np_ = len(s) - k + 1
    codes = np.zeros(np_, dtype=np.uint64)
    valid = np.zeros(np_, dtype=np.uint8)
    lib.orc_encode_positions(s, len(s), k, codes.ctypes.data, valid.ctypes.data)
    keys = sorted_keys(np.unique(codes[valid.astype(bool)])[::2])
    cov = np.zeros(len(s), dtype=np.int32)
    lib.orc_coverage(s, len(s), k, keys.ctypes.data, keys.size, cov.ctypes.data)
    want = np.zeros(len(s), dtype=np.int32)
    for p in range(np_):
        if valid[p] and int(codes[p]) in set(keys.tolist()):
            want[p:p + k] += 1
    assert np.array_equal(cov, want)
    assert coverage_line(np.array([0, 12, 3])) == b"0,12,3"


def test_small_coverage_matches_reference(small_corpus, expected):
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    keys, _ = samples[0]
    cov = oracle_coverage_text([b"chr1"], [seq], 25, sorted_keys(keys))
    assert sha256(cov) == expected["small"]["s0"]["coverage_sha256"]


def test_intersect_matches_reference(small_corpus, expected):
    lib = oracle()
    _, samples = small_corpus
    ks = [sorted_keys(s[0]) for s in samples]
    for pair, want in expected["small"]["intersect"].items():
        a, b = map(int, pair.split("-"))
        got = lib.orc_intersect(ks[a].ctypes.data, ks[a].size, ks[b].ctypes.data, ks[b].size)
        assert got == want
        assert got == len(set(ks[a].tolist()) & set(ks[b].tolist()))


@pytest.mark.parametrize("variant", [("kmc2_p9_s9_c1", 9, 9, 1, 0x200),
                                     ("kmc2_p5_s5_c2", 5, 5, 2, 0x200),
                                     ("kmc2_p9_s11_c4", 9, 11, 4, 0x200),
                                     ("kmc1_p9_c1", 9, 0, 1, 0), ("kmc1_p5_c3", 5, 0, 3, 0),
                                     ("kmc2_p1_s7_c0", 1, 7, 0, 0x200)])
def test_oracle_kmc_reader(tmp_path, variant, expected):
    name, p, s, c, v = variant
    _, keys = corpus.edge_config(25)
    pre = str(tmp_path / name)
    hipibs.kmc_write(pre, keys, 25, p=p, sig_len=max(s, 5), counter_size=c, version=v,
                     min_count=1, max_count=255 if c else 1)
    out = np.zeros(keys.size, dtype=np.uint64)
    import ctypes as C
    kk = C.c_uint32(0)
    n = oracle().orc_kmc_read(pre.encode(), out.ctypes.data, out.size, C.byref(kk))
    assert n == keys.size == expected["kmc_formats"][name]["n_keys"] and kk.value == 25
    assert np.array_equal(np.sort(out), np.sort(keys))


def brute_count(texts, k):
    """Counts of canonical k-mers over the reads' text, by substrings."""
    counts = {}
    for t in texts:
        for v in string_canonical(t, k):
            if v is not None:
                counts[v] = counts.get(v, 0) + 1
    return counts


@pytest.mark.parametrize("k", [1, 5, 25, 32])
def test_count_kmers_against_string_count(k):
    """count_kmers (numpy, vectorised over a read matrix) == a dictionary
    count over every read's substrings; k-mers never span reads, N breaks
    them, ci / cs filter and clamp."""
    rng = np.random.default_rng(k)
    base = rng.integers(0, 4, 400).astype(np.uint8)
    reads = []
    for i in range(60):
        s = int(rng.integers(0, 400 - 40))
        r = base[s:s + 40].copy()
        if i % 3 == 0:
            r = np.where(r < 4, 3 - r, 4)[::-1].copy()  # the other strand
        if i % 7 == 0:
            r[int(rng.integers(0, 40))] = 4
        reads.append(r)
    reads = np.stack(reads)
    want = brute_count([corpus.to_text(r) for r in reads], k)
    for ci, cs in ((1, 1 << 30), (2, 255), (3, 4)):
        keys, counts = corpus.count_kmers(reads, k, ci=ci, cs=cs)
        exp = {v: min(c, cs) for v, c in want.items() if c >= ci}
        assert dict(zip(keys.tolist(), counts.tolist())) == exp
    ok, oc, nk = oracle_count(b"\n".join(corpus.to_text(r, lower=True) for r in reads), k)
    assert dict(zip(ok.tolist(), oc.tolist())) == want
    assert nk == sum(want.values())
    # the reads share k-mers, so counts above 1 and the filters are exercised
    assert max(want.values()) > 4


def test_count_kmers_reproduces_golden_key_counts(expected):
    _, reads = corpus.small_reads()
    for i, r in enumerate(reads):
        keys, counts = corpus.count_kmers(r, 25, ci=1, cs=1 << 31)
        assert int((counts >= 2).sum()) == expected["small"][f"s{i}"]["n_keys"]
        if i == 0:
            ok, oc, _ = oracle_count(b"\n".join(corpus.to_text(x) for x in r), 25)
            np.testing.assert_array_equal(ok, keys)
            np.testing.assert_array_equal(oc, counts)


@pytest.mark.parametrize("k", [33, 40, 47, 63, 64])
def test_wide_codec_against_string_oracle(k):
    rng = np.random.default_rng(k)
    seq = bytes(rng.choice(list(b"ACGTacgt" * 25 + b"N"), 3000).tolist())
    np_ = len(seq) - k + 1
    codes = np.zeros(2 * np_, dtype=np.uint64)
    valid = np.zeros(np_, dtype=np.uint8)
    assert oracle().orc_encode_positions_wide(seq, len(seq), k, codes.ctypes.data,
                                              valid.ctypes.data) == np_
    want = string_canonical(seq, k)
    assert sum(v is not None for v in want) > 100
    for i, w in enumerate(want):
        assert bool(valid[i]) == (w is not None), i
        if w is not None:
            assert (int(codes[2 * i]) << 64) | int(codes[2 * i + 1]) == w, i


@pytest.mark.parametrize("k", [33, 47, 64])
def test_wide_edge_tsv_and_coverage_match_reference(k, expected):
    """The oracle's two-word path against the reference's string path."""
    recs, keys = corpus.edge_config(k)
    ids = [r[0] for r in recs]
    seqs = [r[1] for r in recs]
    ks = sorted_keys_wide(keys)
    e = expected["edge_wide"][str(k)]
    assert e["n_keys"] == ks.shape[0]
    for W, want in e["tsv"].items():
        tsv = hipibs.format_tsv(oracle_scan_rows_wide(ids, seqs, int(W), k, ks))
        assert (sha256(tsv), len(tsv)) == (want["sha256"], want["bytes"]), (k, W)
    cov = oracle_coverage_text_wide(ids, seqs, k, ks)
    assert (sha256(cov), len(cov)) == (e["coverage_sha256"], e["coverage_bytes"])


@pytest.mark.parametrize("k", [33, 48, 64])
def test_wide_kmc_writer_round_trip(tmp_path, k):
    """KMC databases of k > 32 written by the library read back key for key,
    and the reference's CKMCFile finds each key given as a string."""
    import ctypes as C
    recs, keys = corpus.edge_config(k)
    keys = keys[::7]
    p = str(tmp_path / "db")
    counts = (np.arange(keys.shape[0]) % 300 + 1).astype(np.uint32)
    hipibs.kmc_write(p, keys, k, counts=counts, sig_len=9, counter_size=2, version=0x200,
                     min_count=1, max_count=65535)
    back = hipibs.kmc_read_keys(p)
    np.testing.assert_array_equal(np.unique(back, axis=0), keys)
    info = hipibs.kmc_header(p)
    assert (info.kmer_length, info.counter_size, info.total_kmers) == (k, 2, keys.shape[0])
    if k < 64:  # a key wider than 2k bits (every 128-bit value is a 64-mer code)
        bad = keys[:2].copy()
        bad[0, 0] = np.uint64(1) << np.uint64(63)
        with pytest.raises(hipibs.HipibsError):
            hipibs.kmc_write(str(tmp_path / "bad"), bad, k)
    if not os.path.exists(REF_PATH):
        return
    ref = C.CDLL(REF_PATH)
    ref.ref_kmc_is_kmer_strings.argtypes = [C.c_char_p, C.c_char_p, C.c_longlong, C.c_void_p]
    ref.ref_kmc_is_kmer_strings.restype = C.c_longlong
    sub = keys[::10]
    text = "".join("".join("ACGT"[(((int(h) << 64) | int(l)) >> (2 * (k - 1 - i))) & 3]
                           for i in range(k)) for h, l in sub.tolist()).encode()
    hit = np.zeros(sub.shape[0], dtype=np.uint8)
    assert ref.ref_kmc_is_kmer_strings(p.encode(), text, sub.shape[0],
                                       hit.ctypes.data) == sub.shape[0]


def test_reference_batched_whole_scan_writes_the_golden_tsv(tmp_path, small_corpus):
    """The CPU baseline's whole-scan driver (the reference's KmcRaIndex,
    readSequenceFile, chunkSequences, probeBatched and statistics, timed
    phase by phase) writes the golden TSV of each small sample."""
    import ctypes as C
    if not os.path.exists(REF_PATH):
        pytest.skip("oracle/_ref/libref.so not built (no reference sources)")
    ref_lib = C.CDLL(REF_PATH)
    f = ref_lib.ref_batched_whole_scan
    f.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_void_p,
                  C.POINTER(C.c_longlong)]
    f.restype = C.c_longlong
    ref, samples = small_corpus
    fa = tmp_path / "ref.fasta"
    fa.write_bytes(corpus.fasta([(b"chr1", corpus.to_text(ref))]))
    for i, (keys, counts) in enumerate(samples):
        p = str(tmp_path / f"s{i}")
        hipibs.kmc_write(p, keys, 25, counts=counts, p=9, sig_len=9, counter_size=1,
                         version=0x200, min_count=2, max_count=255)
        out = tmp_path / f"s{i}.tsv"
        phases = np.zeros(5)
        probes = C.c_longlong(0)
        n = f(p.encode(), str(fa).encode(), 50000, str(out).encode(), 16, 3,
              phases.ctypes.data, C.byref(probes))
        assert n == 21 and (phases >= 0).all()
        with open(os.path.join(ROOT, "tests", "golden", f"small_s{i}_50000.tsv"), "rb") as g:
            want = g.read()
        assert out.read_bytes() == want
        assert probes.value == sum(int(l.split(b"\t")[3]) for l in want.splitlines()[1:])
