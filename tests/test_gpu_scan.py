"""Parity of the HIP path with the reference (golden fixtures) and the
oracle, through the C ABI. Bit-exact everywhere: the path is integer."""
import os
import sys

import numpy as np
import pytest

import corpus
from util import (ROOT, coverage_line, oracle, oracle_scan_rows, oracle_scan_rows_wide, sha256,
                  sorted_keys, sorted_keys_wide)

import hipibs

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_dbs(tmp_path_factory, small_corpus):
    d = tmp_path_factory.mktemp("small")
    _, samples = small_corpus
    out = []
    for i, (keys, counts) in enumerate(samples):
        p = str(d / f"s{i}")
        hipibs.kmc_write(p, keys, 25, counts=counts, p=9, sig_len=9, counter_size=1,
                         version=0x200, min_count=2, max_count=255)
        out.append(p)
    return out


def test_device_present():
    n, why = hipibs.device_count()
    assert n >= 1, why


def test_small_scan_matches_golden(small_corpus, small_dbs, expected):
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    for i, p in enumerate(small_dbs):
        note(f"small scan s{i}")
        t = hipibs.Table.from_kmc(p)
        info = t.info()
        assert info["n_keys"] == samples[i][0].size
        rows = hipibs.scan_sequences([t], [b"chr1"], [seq], 50000)[0]
        tsv = hipibs.format_tsv(rows)
        with open(os.path.join(ROOT, "tests", "golden", f"small_s{i}_50000.tsv"), "rb") as f:
            assert tsv == f.read()
        t.free()


def test_panel_scan_equals_single(small_corpus, small_dbs):
    """One launch over three tables == three single-table scans."""
    ref, _ = small_corpus
    seq = corpus.to_text(ref)
    tabs = [hipibs.Table.from_kmc(p) for p in small_dbs]
    multi = hipibs.scan_sequences(tabs, [b"chr1"], [seq], 10000)
    for i, t in enumerate(tabs):
        single = hipibs.scan_sequences([t], [b"chr1"], [seq], 10000)[0]
        assert multi[i] == single
    for t in tabs:
        t.free()


@pytest.mark.parametrize("k", [25, 31, 32, 15])
def test_edge_scans_match_golden(k, expected):
    recs, keys = corpus.edge_config(k)
    ids = [r[0] for r in recs]
    seqs = [r[1] for r in recs]
    t = hipibs.Table.from_keys(keys, k)
    assert t.info()["n_keys"] == keys.size
    for W, want in expected["edge"][str(k)]["tsv"].items():
        note(f"edge k={k} W={W}")
        rows = hipibs.scan_sequences([t], ids, seqs, int(W), batch_bases=1 << 22)[0]
        tsv = hipibs.format_tsv(rows)
        assert (sha256(tsv), len(tsv)) == (want["sha256"], want["bytes"]), (k, W)
    t.free()


@pytest.mark.parametrize("slots,lf,mlen,cap", [(8, 0.6, 0, 0), (16, 0.6, 0, 0), (8, 0.9, 0, 0),
                                               (16, 0.95, 0, 0), (8, 0.3, 0, 0), (16, 0.6, 13, 0),
                                               (8, 0.6, 17, 0), (16, 0.9, 5, 0), (16, 0.6, 17, 0),
                                               (16, 0.6, 13, 1), (16, 0.9, 15, 2), (8, 0.6, 11, 4),
                                               (16, 0.6, 9, 0), (16, 0.35, 21, 16),
                                               (8, 0.35, 25, 2)])
def test_table_geometry_does_not_change_results(small_corpus, slots, lf, mlen, cap):
    """Bucket width, load factor and home-bucket rule (whole-k-mer hash,
    minimizer tables of several lengths and group caps, so that groups
    overflow to the whole-k-mer part behind markers) never change the
    statistics."""
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    keys = samples[1][0]
    note(f"geometry slots={slots} lf={lf} mlen={mlen} cap={cap}")
    t = hipibs.Table.from_keys(keys, 25, bucket_slots=slots, load_factor=lf, minimizer_len=mlen,
                               minimizer_cap=cap)
    assert t.info()["minimizer_len"] == mlen
    assert t.info()["n_keys"] == keys.size
    rows = hipibs.scan_sequences([t], [b"chr1"], [seq], 5000)[0]
    want = oracle_scan_rows([b"chr1"], [seq], 5000, 25, sorted_keys(keys))
    assert rows == want
    t.free()


@pytest.mark.parametrize("k,mlen,cap", [(25, 13, 1), (25, 15, 3), (31, 17, 2), (32, 17, 1),
                                        (15, 7, 2), (31, 13, 0), (25, 21, 16), (31, 31, 1)])
def test_minimizer_table_every_build_path(tmp_path, k, mlen, cap):
    """Minimizer tables built from host keys, device keys and a KMC database
    hold exactly the key set (markers never answer as keys), and their scans
    and hit masks equal the oracle's on the edge corpus."""
    import ctypes as C
    recs, keys = corpus.edge_config(k)
    ids = [r[0] for r in recs]
    seqs = [r[1] for r in recs]
    note(f"minimizer table k={k} m={mlen} cap={cap}")
    other = np.unique(corpus.splitmix(k, 91, 30000) & np.uint64((1 << (2 * k)) - 1 if k < 32
                                                              else 0xFFFFFFFFFFFFFFFF))
    other = np.unique(np.concatenate([np.minimum(other, corpus.revcomp(other, k)), keys[::4]]))
    want_other = int(np.isin(other, keys).sum())
    pre = str(tmp_path / "db")
    hipibs.kmc_write(pre, keys, k, sig_len=9 if k > 9 else 5, counter_size=1, version=0x200,
                     min_count=1, max_count=255)
    d_keys = D.from_array(keys)
    opts = dict(minimizer_len=mlen, minimizer_cap=cap)
    tables = [hipibs.Table.from_keys(keys, k, **opts),
              hipibs.Table.from_device_keys(d_keys.ptr, keys.size, k, **opts),
              hipibs.Table.from_kmc(pre, **opts)]
    hipibs.synchronize()
    want_rows = oracle_scan_rows(ids, seqs, 5000, k, sorted_keys(keys))
    for t in tables:
        assert t.info()["n_keys"] == keys.size
        assert t.probe_count(keys) == keys.size
        assert t.probe_count(other) == want_other
        rows = hipibs.scan_sequences([t], ids, seqs, 5000)[0]
        assert rows == want_rows
    # hit mask of the longest contig against the oracle's coverage rule
    c = max(range(len(seqs)), key=lambda i: len(seqs[i]))
    seq = seqs[c]
    sw, ew, _ = hipibs.pack_sequences([seq])
    npos = len(seq) - k + 1
    d_sw, d_ew = D.from_array(sw), D.from_array(ew)
    d_hit = D.zeros(((npos + 31) // 32 + 1) * 4)
    hipibs._check(hipibs._lib.hipibs_hitmask_device(tables[2].handle, d_sw.ptr, d_ew.ptr, 0, npos,
                                                    len(seq), d_hit.ptr, None), "hitmask")
    hipibs.synchronize()
    words = d_hit.to_array(np.uint32, (npos + 31) // 32)
    bits = ((words[:, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(-1)[:npos]
    lib = oracle()
    enc = np.zeros(npos, dtype=np.uint64)
    val = np.zeros(npos, dtype=np.uint8)
    assert lib.orc_encode_positions(seq, len(seq), k, enc.ctypes.data, val.ctypes.data) == npos
    want_bits = (val.astype(bool) & np.isin(enc, keys)).astype(np.int64)
    assert np.array_equal(bits.astype(np.int64), want_bits)
    for t in tables:
        t.free()


def test_minimizer_len_limits():
    """Minimizer lengths outside 5..min(31, k) are refused."""
    keys = corpus.sequence_kmers([corpus.random_bases(5, 5, 4000)], 32)
    for k, m in ((32, 32), (25, 26), (25, 4)):
        with pytest.raises(hipibs.HipibsError, match="minimizer_len"):
            hipibs.Table.from_keys(keys[:100] >> np.uint64(2 * (32 - k)), k, minimizer_len=m)
    t = hipibs.Table.from_keys(keys, 32, minimizer_len=31)
    assert t.probe_count(keys) == keys.size
    t.free()


def test_probe_count_intersect(small_corpus, expected):
    _, samples = small_corpus
    tabs = [hipibs.Table.from_keys(s[0], 25) for s in samples]
    for pair, want in expected["small"]["intersect"].items():
        a, b = map(int, pair.split("-"))
        small, big = (a, b) if samples[a][0].size <= samples[b][0].size else (b, a)
        assert tabs[big].probe_count(samples[small][0]) == want
        assert tabs[small].probe_count(samples[big][0]) == want
    # self-intersection is the whole set; random codes mostly miss
    assert tabs[0].probe_count(samples[0][0]) == samples[0][0].size
    rnd = corpus.splitmix(3, 3, 200000) & np.uint64((1 << 50) - 1)
    rnd = np.minimum(rnd, corpus.revcomp(rnd, 25))
    want = int(np.isin(rnd, samples[0][0]).sum())
    assert tabs[0].probe_count(rnd) == want
    for t in tabs:
        t.free()


def test_duplicate_keys_stored_once():
    keys = corpus.splitmix(5, 5, 100000) & np.uint64((1 << 50) - 1)
    dup = np.concatenate([keys, keys[::3], keys[::7]])
    t = hipibs.Table.from_keys(dup, 25, load_factor=0.9)
    assert t.info()["n_keys"] == np.unique(keys).size
    assert t.probe_count(keys) == keys.size
    t.free()


D = hipibs.DeviceBuffer


@pytest.mark.parametrize("slots,lf,n,bits", [(16, 0.6, 3_000_000, 9), (16, 0.95, 2_000_000, 9),
                                             (8, 0.9, 1_500_000, 9), (16, 0.6, 200_000, 9),
                                             (8, 0.3, 700_000, 5), (16, 0.6, 3_000_000, 4)])
def test_sorted_table_build_equals_insert(small_corpus, monkeypatch, slots, lf, n, bits):
    """Whole-k-mer tables from device keys are built by sorting the keys by
    table slice (above HIPIBS_SORTED_BUILD_MIN keys): the same key set as
    the key-by-key insert -- repeated keys stored once, EMPTY rows skipped,
    keys past their slice's end (load 0.95) inserted afterwards -- over one,
    two and three sort passes (narrow digits forced) and both bucket widths;
    every key answers, non-keys answer as set arithmetic says, and a scan of
    the 1 Mbp reference with the sample's keys added gives the same rows as
    the oracle for both builds."""
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    sample = samples[1][0][:n]
    rnd = corpus.splitmix(n, 17, n) & np.uint64((1 << 50) - 1)
    keys = np.concatenate([sample, rnd, rnd[::5], sample[::3]])
    keys[::100003] = np.uint64(0xFFFFFFFFFFFFFFFF)  # EMPTY rows: no key
    distinct = np.unique(keys[keys != np.uint64(0xFFFFFFFFFFFFFFFF)])
    other = corpus.splitmix(n, 29, 200_000) & np.uint64((1 << 50) - 1)
    want_other = int(np.isin(other, distinct).sum())
    d = D.from_array(keys)
    note(f"sorted build slots={slots} lf={lf} n={keys.size} digit bits={bits}")
    monkeypatch.setenv("HIPIBS_SORT_DIGIT_BITS", str(bits))
    monkeypatch.setenv("HIPIBS_SORTED_BUILD_MIN", "1")
    t_sort = hipibs.Table.from_device_keys(d.ptr, keys.size, 25, bucket_slots=slots,
                                           load_factor=lf)
    monkeypatch.setenv("HIPIBS_SORTED_BUILD_MIN", str(1 << 62))
    t_ins = hipibs.Table.from_device_keys(d.ptr, keys.size, 25, bucket_slots=slots,
                                          load_factor=lf)
    monkeypatch.delenv("HIPIBS_SORTED_BUILD_MIN")
    monkeypatch.delenv("HIPIBS_SORT_DIGIT_BITS")
    nb = t_sort.info()["n_buckets"]
    note(f"  {nb} buckets, {-(-nb // (256 if slots == 16 else 512))} slices")
    want_rows = oracle_scan_rows([b"chr1"], [seq], 50000, 25, sorted_keys(distinct))
    for t in (t_sort, t_ins):
        assert t.info()["n_keys"] == distinct.size
        assert t.probe_count(distinct) == distinct.size
        assert t.probe_count(other) == want_other
        assert hipibs.scan_sequences([t], [b"chr1"], [seq], 50000)[0] == want_rows
    t_sort.free()
    t_ins.free()
    d.free()


def home_buckets(keys, n_buckets):
    """Home bucket of every key: the top 32 bits of the library's
    multiply-xorshift mix scaled onto the bucket count (used here only to
    choose keys that crowd one slice)."""
    x = keys.astype(np.uint64)
    with np.errstate(over="ignore"):
        h = (x ^ (x >> np.uint64(31))) * np.uint64(0x9E3779B97F4A7C15)
        h ^= h >> np.uint64(29)
        h = h * np.uint64(0xBF58476D1CE4E5B9)
        return ((h >> np.uint64(32)) * np.uint64(n_buckets)) >> np.uint64(32)


@pytest.mark.parametrize("slots", [16, 8])
def test_sorted_table_build_distinct_and_crowded_slices(small_corpus, monkeypatch, slots):
    """Distinct keys: every slice but one is filled by rank and run scan (no
    repeated key sends it to the compare-and-swap walk); one slice is
    crowded with 6,000 keys for its 4,096 slots, so it takes the walk and
    its overflow is inserted afterwards. The sorted build holds the same
    keys as the key-by-key insert, and a scan with either table gives the
    oracle's rows."""
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    mask = np.uint64((1 << 50) - 1)
    base = np.unique(np.concatenate([samples[1][0], corpus.splitmix(slots, 41, 1_500_000) & mask]))
    lf = 0.6
    nb = max(2, -(-int(np.ceil(base.size / lf)) // slots))  # as the table sizes itself
    per_slice = 256 if slots == 16 else 512
    cand = np.setdiff1d(corpus.splitmix(slots, 43, 8_000_000) & mask, base)
    crowd = cand[home_buckets(cand, nb) // per_slice == 3][:6000]
    assert crowd.size == 6000
    keys = np.concatenate([base[crowd.size:], crowd])  # same row count, all distinct
    assert np.unique(keys).size == keys.size == base.size
    other = corpus.splitmix(slots, 47, 200_000) & mask
    want_other = int(np.isin(other, keys).sum())
    d = D.from_array(keys)
    note(f"sorted build, distinct keys, slots={slots}: {keys.size} keys, {nb} buckets, "
         f"slice 3 holds {crowd.size}")
    monkeypatch.setenv("HIPIBS_SORTED_BUILD_MIN", "1")
    t_sort = hipibs.Table.from_device_keys(d.ptr, keys.size, 25, bucket_slots=slots,
                                           load_factor=lf)
    monkeypatch.setenv("HIPIBS_SORTED_BUILD_MIN", str(1 << 62))
    t_ins = hipibs.Table.from_device_keys(d.ptr, keys.size, 25, bucket_slots=slots,
                                          load_factor=lf)
    monkeypatch.delenv("HIPIBS_SORTED_BUILD_MIN")
    assert t_sort.info()["n_buckets"] == nb
    want_rows = oracle_scan_rows([b"chr1"], [seq], 50000, 25, sorted_keys(keys))
    for t in (t_sort, t_ins):
        assert t.info()["n_keys"] == keys.size
        assert t.probe_count(keys) == keys.size
        assert t.probe_count(crowd) == crowd.size
        assert t.probe_count(other) == want_other
        assert hipibs.scan_sequences([t], [b"chr1"], [seq], 50000)[0] == want_rows
    t_sort.free()
    t_ins.free()
    d.free()


def test_coverage_matches_golden(small_corpus, expected):
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    keys = samples[0][0]
    t = hipibs.Table.from_keys(keys, 25)
    sw, ew, off = hipibs.pack_sequences([seq])
    d_sw, d_ew = D.from_array(sw), D.from_array(ew)
    n = len(seq)
    npos = n - 25 + 1
    d_hit = D.zeros(((npos + 31) // 32 + 1) * 4)
    d_cov = D.zeros(n)
    hipibs._check(hipibs._lib.hipibs_hitmask_device(t.handle, d_sw.ptr, d_ew.ptr, 0, npos, n,
                                                    d_hit.ptr, None), "hitmask")
    hipibs._check(hipibs._lib.hipibs_coverage_device(d_hit.ptr, npos, n, 25, d_cov.ptr, None),
                  "coverage")
    hipibs.synchronize()
    cov = d_cov.to_array(np.uint8, n)
    text = b"chr1\n" + coverage_line(cov) + b"\n"
    assert sha256(text) == expected["small"]["s0"]["coverage_sha256"]
    t.free()


def test_coverage_kernel_against_brute_force():
    """hipibs_coverage_device on random hit masks: every k in 1..64 class,
    lengths around the 16-base chunks and 32-bit words, contigs shorter
    than k, an output pointer off 16 B alignment, bits past n_pos ignored."""
    rng = np.random.default_rng(77)
    cases = [(k, n) for k in (1, 2, 15, 25, 31, 32, 33, 47, 63, 64)
             for n in (0, 1, 15, 16, 17, 31, 33, 63, 64, 65, 100, 1000, 4099)]
    for k, n in cases:
        npos = n - k + 1 if n >= k else 0
        bits = (rng.random(max(npos, 1)) < 0.6)[:npos]
        words = np.zeros((npos + 31) // 32 + 2, dtype=np.uint32)
        for p in np.flatnonzero(bits):
            words[p // 32] |= np.uint32(1) << np.uint32(p % 32)
        if npos % 32:
            words[npos // 32] |= ~((np.uint32(1) << np.uint32(npos % 32)) - np.uint32(1))
        want = np.zeros(n, dtype=np.int64)
        for p in np.flatnonzero(bits):
            want[p:p + k] += 1
        d_hit = D.from_array(words)
        for off in (0, 1):
            d_cov = D.zeros(n + 32)
            ptr = d_cov.ptr + off
            hipibs._check(hipibs._lib.hipibs_coverage_device(d_hit.ptr, npos, n, k, ptr, None),
                          "coverage")
            hipibs.synchronize()
            got = d_cov.to_array(np.uint8, n + 32)[off:off + n].astype(np.int64)
            assert np.array_equal(got, want), (k, n, off)
            d_cov.free()
        d_hit.free()


@pytest.mark.parametrize("n", [1, 2, 15, 16, 17, 4095, 4096, 4097, 12288, 100003])
def test_coverage_text_matches_the_serialisation(n):
    """Device text of a coverage line == the values joined by ',' (1-3
    digits, no separator after the last), through both text entry points."""
    import ctypes as C
    rng = np.random.default_rng(n)
    cov = rng.integers(0, 33, n).astype(np.uint8)
    big = rng.random(n) < 0.05
    cov[big] = rng.integers(100, 256, int(big.sum()))
    want = coverage_line(cov)
    d_cov = D.from_array(cov)
    ln = C.c_uint64(0)
    hipibs._check(hipibs._lib.hipibs_coverage_text_device(d_cov.ptr, n, None, 0, C.byref(ln), None),
                  "text length")
    assert ln.value == len(want)
    d_txt = D.zeros(4 * n)
    hipibs._check(hipibs._lib.hipibs_coverage_text_device(d_cov.ptr, n, d_txt.ptr, 4 * n,
                                                          C.byref(ln), None), "text")
    assert d_txt.to_array(np.uint8, ln.value).tobytes() == want
    small = D.zeros(max(1, len(want) - 1))
    assert hipibs._lib.hipibs_coverage_text_device(d_cov.ptr, n, small.ptr, len(want) - 1,
                                                   C.byref(ln), None) != 0
    parts = []
    SINK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_char), C.c_uint64)
    sink = SINK(lambda u, d, m: parts.append(C.string_at(d, m)) or 0)
    hipibs._check(hipibs._lib.hipibs_coverage_text_stream(d_cov.ptr, n, sink, None, None),
                  "text stream")
    assert b"".join(parts) == want


def test_encode_device_matches_oracle():
    rng = np.random.default_rng(11)
    for k in (15, 25, 31, 32):
        c = rng.integers(0, 4, 20000).astype(np.uint8)
        c[rng.random(20000) < 0.01] = 4
        s = corpus.to_text(c, lower=bool(k % 2))
        sw, ew, _ = hipibs.pack_sequences([s])
        d_sw, d_ew = D.from_array(sw), D.from_array(ew)
        np_ = len(s) - k + 1
        d_codes = D.zeros(np_ * 8)
        d_valid = D.zeros(np_)
        hipibs._check(hipibs._lib.hipibs_encode_device(d_sw.ptr, d_ew.ptr, 0, np_, len(s), k,
                                                       d_codes.ptr, d_valid.ptr, None), "encode")
        hipibs.synchronize()
        codes = d_codes.to_array(np.uint64, np_)
        valid = d_valid.to_array(np.uint8, np_)
        oc = np.zeros(np_, dtype=np.uint64)
        ov = np.zeros(np_, dtype=np.uint8)
        oracle().orc_encode_positions(s, len(s), k, oc.ctypes.data, ov.ctypes.data)
        assert np.array_equal(valid, ov)
        assert np.array_equal(codes[ov == 1], oc[ov == 1])
        assert np.all(codes[ov == 0] == np.uint64(2 ** 64 - 1))


def test_encode_device_wide_matches_oracle():
    """Two-word codes (k = 33..64): (high, low) per position."""
    rng = np.random.default_rng(12)
    for k in (33, 47, 63, 64):
        c = rng.integers(0, 4, 20000).astype(np.uint8)
        c[rng.random(20000) < 0.003] = 4
        s = corpus.to_text(c, lower=bool(k % 2))
        sw, ew, _ = hipibs.pack_sequences([s])
        d_sw, d_ew = D.from_array(sw), D.from_array(ew)
        np_ = len(s) - k + 1
        d_codes = D.zeros(np_ * 16)
        d_valid = D.zeros(np_)
        hipibs._check(hipibs._lib.hipibs_encode_device(d_sw.ptr, d_ew.ptr, 0, np_, len(s), k,
                                                       d_codes.ptr, d_valid.ptr, None), "encode")
        hipibs.synchronize()
        codes = d_codes.to_array(np.uint64, 2 * np_).reshape(np_, 2)
        valid = d_valid.to_array(np.uint8, np_)
        oc = np.zeros(2 * np_, dtype=np.uint64)
        ov = np.zeros(np_, dtype=np.uint8)
        oracle().orc_encode_positions_wide(s, len(s), k, oc.ctypes.data, ov.ctypes.data)
        oc = oc.reshape(np_, 2)
        assert np.array_equal(valid, ov) and ov.sum() > 1000
        assert np.array_equal(codes[ov == 1], oc[ov == 1])
        assert np.all(codes[ov == 0] == np.uint64(2 ** 64 - 1))


def test_probe_bytes_matches_membership(small_corpus):
    _, samples = small_corpus
    keys = samples[0][0]
    t = hipibs.Table.from_keys(keys, 25)
    q = np.concatenate([keys[::5], np.minimum(corpus.splitmix(8, 8, 100000) & np.uint64((1 << 50) - 1),
                                              corpus.revcomp(corpus.splitmix(8, 8, 100000) &
                                                             np.uint64((1 << 50) - 1), 25))])
    valid = (np.arange(q.size) % 7 != 3).astype(np.uint8)
    d_q, d_v, d_h = D.from_array(q), D.from_array(valid), D.zeros(q.size)
    hipibs._check(hipibs._lib.hipibs_table_probe(t.handle, d_q.ptr, d_v.ptr, q.size, d_h.ptr, None),
                  "probe")
    hipibs.synchronize()
    hits = d_h.to_array(np.uint8, q.size)
    want = (np.isin(q, keys) & (valid == 1)).astype(np.uint8)
    assert np.array_equal(hits, want)
    t.free()


@pytest.mark.parametrize("segs", [1, 2, 3, 7, 16])
def test_segment_split_invariance(small_corpus, segs):
    """Windows split over several blocks fold to the same statistics
    (the run-length carry across block seams)."""
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    keys = samples[2][0]
    t = hipibs.Table.from_keys(keys, 25)
    sw, ew, _ = hipibs.pack_sequences([seq])
    wins = hipibs.chunk_sequences([len(seq)], 50000, 25)
    win = np.zeros(len(wins), dtype=hipibs.WINDOW_DTYPE)
    for i, (_, s, e) in enumerate(wins):
        win[i] = (s, e - s, 0)
    d_sw, d_ew, d_win = D.from_array(sw), D.from_array(ew), D.from_array(win)
    d_st = D.zeros(len(wins) * 16)
    import ctypes as C
    res = C.c_uint32(0)
    wb = hipibs._lib.hipibs_scan_work_bytes(len(wins), 1, 50000, 25, segs, C.byref(res))
    assert res.value == segs
    d_work = D.zeros(max(wb, 8))
    arr = hipibs._table_array([t])
    hipibs._check(hipibs._lib.hipibs_scan_windows_device(
        arr, 1, d_sw.ptr, d_ew.ptr, len(seq), d_win.ptr, len(wins), 50000, d_st.ptr, segs,
        d_work.ptr, wb, None), "scan")
    hipibs.synchronize()
    st = d_st.to_array(np.int32).reshape(-1, 4)
    want = oracle_scan_rows([b"chr1"], [seq], 50000, 25, sorted_keys(keys))
    got = [(b"chr1", s, e) + tuple(int(x) for x in st[i]) for i, (_, s, e) in enumerate(wins)]
    assert got == want
    t.free()


def test_build_from_kmc_equals_from_keys(small_dbs, small_corpus):
    _, samples = small_corpus
    a = hipibs.Table.from_kmc(small_dbs[1], load_threads=3)
    keys = hipibs.kmc_read_keys(small_dbs[1])
    assert np.array_equal(np.sort(keys), samples[1][0])
    assert a.info()["n_keys"] == keys.size
    assert a.probe_count(samples[0][0]) == int(np.isin(samples[0][0], samples[1][0]).sum())
    a.free()


@pytest.mark.parametrize("variant", [(25, 9, 9, 1, 0x200), (25, 5, 5, 2, 0x200),
                                     (25, 9, 11, 4, 0x200), (25, 9, 0, 1, 0), (25, 1, 7, 0, 0x200),
                                     (32, 8, 9, 1, 0x200), (32, 0, 9, 1, 0x200), (31, 7, 9, 3, 0),
                                     (15, 3, 5, 1, 0x200)])
def test_build_from_kmc_decodes_every_layout(tmp_path, variant):
    """Table builds decode KMC records on the device (LUT bucket search +
    big-endian suffix): every layout gives exactly the written key set."""
    k, p, s, c, v = variant
    note(f"kmc layout k={k} p={p} sig={s} counter={c} version={v:#x}")
    _, keys = corpus.edge_config(k)
    counts = (np.arange(keys.size) % 300 + 1).astype(np.uint32)
    pre = str(tmp_path / "db")
    hipibs.kmc_write(pre, keys, k, counts=counts, p=p, sig_len=max(s, 5), counter_size=c,
                     version=v, min_count=1, max_count=2 ** 31)
    t = hipibs.Table.from_kmc(pre, load_threads=3)
    assert t.info()["n_keys"] == keys.size
    assert t.probe_count(keys) == keys.size
    other = np.unique(corpus.splitmix(k, 77, 20000) & np.uint64((1 << (2 * k)) - 1 if k < 32
                                                              else 0xFFFFFFFFFFFFFFFF))
    other = np.unique(np.minimum(other, corpus.revcomp(other, k)))
    other = np.unique(np.concatenate([other, keys[::5]]))
    want = int(np.isin(other, keys).sum())
    assert t.probe_count(other) == want
    # the intersect's streamed count: records decoded on the device
    assert t.probe_count_kmc(pre, threads=2) == keys.size
    pre2 = str(tmp_path / "other")
    hipibs.kmc_write(pre2, other, k, p=p, sig_len=max(s, 5), counter_size=c, version=v,
                     min_count=1, max_count=2 ** 31)
    assert t.probe_count_kmc(pre2) == want
    t.free()


@pytest.mark.parametrize("kind", ["whole", "minimizer"])
def test_build_from_kmc_across_chunks(tmp_path, kind):
    """More records than one staging chunk (32M): the decode of every chunk
    starts at its own record and bucket range; a minimizer table counts its
    groups chunk by chunk while the records stream, and must come out the
    same size as the one built from the same keys on the host (the sizes
    follow from the group counts)."""
    k = 31
    keys = np.unique(corpus.splitmix(31, 5, 40_000_000) & np.uint64((1 << 62) - 1))
    keys = np.unique(np.minimum(keys, corpus.revcomp(keys, k)))
    note(f"kmc across chunks ({kind}): {keys.size} keys")
    pre = str(tmp_path / "big")
    hipibs.kmc_write(pre, keys, k, p=7, sig_len=9, counter_size=1, version=0x200,
                     min_count=1, max_count=255)
    opts = dict(minimizer_len=27, minimizer_cap=16, load_factor=0.2) if kind == "minimizer" else {}
    t = hipibs.Table.from_kmc(pre, **opts)
    assert t.info()["n_keys"] == keys.size
    assert t.probe_count(keys) == keys.size
    probe = keys[::97] ^ np.uint64(1)
    assert t.probe_count(probe) == int(np.isin(probe, keys).sum())
    assert t.probe_count_kmc(pre) == keys.size
    if kind == "minimizer":
        assert t.info()["minimizer_len"] == 27
        t_host = hipibs.Table.from_keys(keys, k, **opts)
        assert t_host.info()["bytes"] == t.info()["bytes"]
        assert t_host.info()["n_keys"] == keys.size
        t_host.free()
    t.free()


def _packed_windows(seq, W, k):
    sw, ew, _ = hipibs.pack_sequences([seq])
    wins = hipibs.chunk_sequences([len(seq)], W, k)
    win = np.zeros(len(wins), dtype=hipibs.WINDOW_DTYPE)
    for i, (_, s, e) in enumerate(wins):
        win[i] = (s, e - s, 0)
    return D.from_array(sw), D.from_array(ew), D.from_array(win), wins


PANEL_MINIMIZER = {"whole": dict(bucket_slots=8),
                   "m21": dict(minimizer_len=21, minimizer_cap=4, load_factor=0.35,
                               bucket_slots=16),
                   "m15s8": dict(minimizer_len=15, minimizer_cap=16, load_factor=0.3,
                                 bucket_slots=8)}


@pytest.mark.parametrize("n_tables,kind", [(1, "whole"), (3, "whole"), (17, "whole"),
                                           (1, "m21"), (3, "m21"), (17, "m21"), (5, "m15s8")])
def test_panel_kernel_equals_single_scans(small_corpus, n_tables, kind):
    """The persistent one-encode-many-probes panel kernel (16 tables per
    launch, so 17 spans two) gives each table's single-scan statistics, for
    several launch configurations; over whole-k-mer tables and over
    minimizer tables (each word's minimizers computed once, every table
    probed with the grouped probe; a cap of 4 sends repeat groups to the
    whole-k-mer part)."""
    import ctypes as C
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    keysets = [samples[i % 3][0][: (samples[i % 3][0].size * (i + 5)) // (n_tables + 5)]
               for i in range(n_tables)]
    opts = PANEL_MINIMIZER[kind]
    tabs = [hipibs.Table.from_keys(ks, 25, **opts) for ks in keysets]
    d_sw, d_ew, d_win, wins = _packed_windows(seq, 20000, 25)
    nw = len(wins)
    arr = hipibs._table_array(tabs)
    want = [oracle_scan_rows([b"chr1"], [seq], 20000, 25, sorted_keys(ks)) for ks in keysets]
    for segs, wpc in ((1, 16), (3, 8), (2, 32)):
        note(f"panel n={n_tables} {kind} segs={segs} wpc={wpc}")
        wb = hipibs._lib.hipibs_panel_work_bytes(nw, n_tables, segs)
        d_work = D.zeros(wb)
        d_st = D.zeros(nw * n_tables * 16)
        cfg = hipibs.ScanConfig(segs, wpc)
        hipibs._check(hipibs._lib.hipibs_panel_scan_device(
            arr, n_tables, d_sw.ptr, d_ew.ptr, len(seq), d_win.ptr, nw, d_st.ptr, C.byref(cfg),
            d_work.ptr, wb, None), "panel")
        hipibs.synchronize()
        st = d_st.to_array(np.int32).reshape(n_tables, nw, 4)
        for t in range(n_tables):
            got = [(b"chr1", s, e) + tuple(int(x) for x in st[t][i])
                   for i, (_, s, e) in enumerate(wins)]
            assert got == want[t], (t, segs, wpc)
    for t in tabs:
        t.free()


def test_panel_graph_replays_and_autotune(small_corpus):
    """A captured panel graph (4 chunks) replayed three times, with eager
    work between the replays, matches the eager result every time; the
    autotuner returns a configuration that runs."""
    import ctypes as C
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    tabs = [hipibs.Table.from_keys(samples[i][0], 25) for i in range(3)]
    d_sw, d_ew, d_win, wins = _packed_windows(seq, 10000, 25)
    nw = len(wins)
    arr = hipibs._table_array(tabs)
    best = hipibs.ScanConfig()
    ms = C.c_double(0)
    wb = hipibs._lib.hipibs_panel_work_bytes(nw, 3, 4)
    d_work = D.zeros(wb)
    d_eager = D.zeros(nw * 3 * 16)
    hipibs._check(hipibs._lib.hipibs_panel_autotune(
        arr, 3, d_sw.ptr, d_ew.ptr, len(seq), d_win.ptr, nw, 64, d_eager.ptr, d_work.ptr, wb,
        C.byref(best), C.byref(ms)), "autotune")
    assert best.segs_per_window in (1, 2, 4) and best.waves_per_cu in (8, 12, 16, 24, 32)
    assert ms.value > 0
    hipibs._check(hipibs._lib.hipibs_panel_scan_device(
        arr, 3, d_sw.ptr, d_ew.ptr, len(seq), d_win.ptr, nw, d_eager.ptr, C.byref(best),
        d_work.ptr, wb, None), "panel")
    hipibs.synchronize()
    eager = d_eager.to_array(np.int32)
    d_g = D.zeros(nw * 3 * 16)
    g = C.c_void_p()
    hipibs._check(hipibs._lib.hipibs_panel_graph_create(
        arr, 3, d_sw.ptr, d_ew.ptr, len(seq), d_win.ptr, nw, d_g.ptr, C.byref(best), 4,
        d_work.ptr, wb, C.byref(g)), "graph")
    for rep in range(3):
        hipibs._check(hipibs._lib.hipibs_dev_memset(C.c_void_p(d_g.ptr), 0xFF, nw * 3 * 16),
                      "memset")
        hipibs._check(hipibs._lib.hipibs_graph_launch(g, None), "graph launch")
        hipibs.synchronize()
        assert np.array_equal(d_g.to_array(np.int32), eager), rep
        # eager work between replays
        assert tabs[0].probe_count(samples[1][0][:1000]) >= 0
    hipibs._lib.hipibs_graph_free(g)
    for t in tabs:
        t.free()


def test_serving_results_same_alone_and_in_the_mix(small_corpus, tmp_path):
    """Continuous batching: every request's statistics are identical when
    served alone and when served mixed with other samples, ranges and
    window sizes, with an HBM budget that forces evictions and
    re-admissions (one sample registered by its KMC database in the mix);
    and equal to the oracle."""
    import time
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    sw, ew, _ = hipibs.pack_sequences([seq])
    n = len(seq)
    keys = [s[0] for s in samples]
    one_table = hipibs.Table.from_keys(keys[0], 25).info()["bytes"]
    rng = np.random.default_rng(42)
    reqs = []
    for i in range(40):
        sid = int(rng.integers(0, 3))
        W = int(rng.choice([1000, 5000, 10000]))
        ln = int(rng.integers(30, n // 3))
        base = int(rng.integers(0, n - ln))
        reqs.append((i, sid, base, ln, W))

    kmc2 = str(tmp_path / "s2")
    hipibs.kmc_write(kmc2, keys[2], 25, p=9, sig_len=9, counter_size=1, version=0x200,
                     min_count=1, max_count=255)

    def serve(phases, budget, burst, via_kmc=False):
        eng = hipibs.ServeEngine(25, budget, max_inflight=3, max_windows_per_launch=512)
        eng.set_reference(sw, ew, n)
        for sid in range(3):
            if via_kmc and sid == 2:
                eng.add_sample_kmc(sid, kmc2)
            else:
                eng.add_sample(sid, keys[sid])
        out = {}
        for batch in phases:  # each phase completes before the next arrives
            pending = set()
            for j, (rid, sid, base, ln, W) in enumerate(batch):
                eng.submit(rid, sid, base, ln, W)
                pending.add(rid)
                if burst and j % 7 == 6:
                    time.sleep(0.002)  # later arrivals while earlier ones run
            while pending:
                for rid in eng.poll(timeout_ms=2000):
                    st, lat, q = eng.fetch(rid)
                    assert lat >= q >= 0
                    out[rid] = st
                    pending.discard(rid)
        c = eng.counters()
        eng.free()
        return out, c

    alone = {}
    for r in reqs:
        o, _ = serve([[r]], 3 * one_table, False)
        alone.update(o)
    # one table fits: the first phase admits every sample it asks for; the
    # second asks for all three again, so at least two are admitted again
    first, second = reqs[:20], reqs[20:]
    assert {r[1] for r in second} == {0, 1, 2}
    mixed, c = serve([first, second], int(1.5 * one_table), True, via_kmc=True)
    assert c["evictions"] >= 2 and c["admissions"] >= len({r[1] for r in first}) + 2
    assert c["launches"] < len(reqs)  # requests shared launches
    for rid, sid, base, ln, W in reqs:
        assert np.array_equal(alone[rid], mixed[rid]), rid
    for rid, sid, base, ln, W in reqs[:4]:
        want = oracle_scan_rows([b"c"], [seq[base:base + ln]], W, 25, sorted_keys(keys[sid]))
        got = [tuple(int(x) for x in row) for row in mixed[rid]]
        assert got == [w[3:] for w in want], rid


@pytest.mark.parametrize("k", [11, 12, 21, 25, 31, 32, 33, 47, 64])
def test_serving_key_copy_round_trip(k):
    """Admissions rebuild tables from the engine's compact host copy of a
    sample's keys (the top <= 22 bits of each code implied by its position
    among keys grouped by them, the rest kept as whole bytes: 0 bytes at
    k = 11, 1 at k = 12, ..., 14 at k = 64). A budget of 1.2 times the
    largest table makes nearly every request re-admit its sample from that
    copy. The
    reference holds T^i A^(k-i) runs, so codes whose prefix is all ones
    are hit; samples come from host keys and from device keys (with A^k,
    code 0). Every request equals the oracle."""
    wide = k > 32
    rnd = corpus.random_bases(0x5E7E, k, 40_000)
    runs = np.concatenate([np.full(k, 3, np.uint8), np.full(k, 0, np.uint8),
                           np.full(k + 3, 0, np.uint8)])
    codes = np.concatenate([rnd, runs, corpus.random_bases(0x5E7E, 100 + k, 2_000)])
    seq = corpus.to_text(codes)
    sw, ew, _ = hipibs.pack_sequences([seq])
    n = len(seq)
    every = (corpus.sequence_kmers_wide if wide else corpus.sequence_kmers)([codes], k)
    rng = np.random.default_rng(k)
    samples = [every, every[rng.random(every.shape[0]) < 0.5], every[::3]]
    samples = [np.ascontiguousarray(x) for x in samples]
    one = hipibs.Table.from_keys(samples[0], k)
    budget = int(1.2 * one.info()["bytes"])
    one.free()
    eng = hipibs.ServeEngine(k, budget, max_inflight=2, max_windows_per_launch=256)
    eng.set_reference(sw, ew, n)
    eng.add_sample(0, samples[0])
    eng.add_sample(1, samples[1])
    d = hipibs.DeviceBuffer.from_array(samples[2])
    eng.add_sample_device(2, d.ptr, samples[2].shape[0])
    d.free()
    reqs = []
    for i in range(9):
        ln = int(rng.integers(2 * k, n))
        reqs.append((i, i % 3, int(rng.integers(0, n - ln + 1)), ln, int(rng.choice([300, 5000]))))
    reqs.append((9, 0, rnd.size - 5, 3 * k + 10, 2 * k))  # the T/A runs
    out = {}
    for r in reqs:  # one at a time: each admits its sample again
        eng.submit(*r)
        got = []
        while not got:
            got = eng.poll(timeout_ms=2000)
        assert got == [r[0]]
        out[r[0]] = eng.fetch(r[0])[0]
    c = eng.counters()
    eng.free()
    assert c["evictions"] >= 6 and c["admissions"] >= 8
    bpk = {11: 0, 12: 1, 21: 3, 25: 4, 31: 5, 32: 6, 33: 6, 47: 9, 64: 14}[k]
    assert c["host_bytes"] == sum(x.shape[0] for x in samples) * bpk + 3 * 4 * (
        (1 << (2 * k - 8 * bpk)) + 1)
    for rid, sid, base, ln, W in reqs:
        if wide:
            want = oracle_scan_rows_wide([b"c"], [seq[base:base + ln]], W, k,
                                         sorted_keys_wide(samples[sid]))
        else:
            want = oracle_scan_rows([b"c"], [seq[base:base + ln]], W, k, sorted_keys(samples[sid]))
        assert [tuple(int(x) for x in row) for row in out[rid]] == [w[3:] for w in want], rid


def test_serving_rejects_codes_wider_than_2k_bits():
    seq = corpus.to_text(corpus.random_bases(3, 3, 5_000))
    sw, ew, _ = hipibs.pack_sequences([seq])
    for k, bad in ((25, np.array([1 << 50], dtype=np.uint64)),
                   (40, np.array([[1 << 16, 0]], dtype=np.uint64))):
        eng = hipibs.ServeEngine(k, 1 << 30, max_inflight=1, max_windows_per_launch=64)
        eng.set_reference(sw, ew, len(seq))
        with pytest.raises(hipibs.HipibsError, match="2k-bit"):
            eng.add_sample(0, bad)
        eng.free()


def test_minimizer_serving_equals_oracle(small_corpus):
    """Serving with minimizer tables (the grouped probe in the mixed-table
    launch; a cap of 4 keys per group sends repeat groups to the whole-k-mer
    part), a budget that forces evictions: every request equals the
    oracle."""
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    sw, ew, _ = hipibs.pack_sequences([seq])
    n = len(seq)
    keys = [s[0] for s in samples]
    opts = dict(minimizer_len=21, minimizer_cap=4, load_factor=0.35)
    one_table = hipibs.Table.from_keys(keys[0], 25, **opts).info()["bytes"]
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(18):
        ln = int(rng.integers(30, n // 3))
        reqs.append((i, i % 3, int(rng.integers(0, n - ln)), ln,
                     int(rng.choice([1000, 5000, 10000]))))
    eng = hipibs.ServeEngine(25, int(1.5 * one_table), max_inflight=3,
                             max_windows_per_launch=512, **opts)
    eng.set_reference(sw, ew, n)
    for sid in range(3):
        eng.add_sample(sid, keys[sid])
    got = {}
    for phase in (reqs[:9], reqs[9:]):
        pending = set()
        for r in phase:
            eng.submit(*r)
            pending.add(r[0])
        while pending:
            for rid in eng.poll(timeout_ms=2000):
                got[rid] = eng.fetch(rid)[0]
                pending.discard(rid)
    c = eng.counters()
    eng.free()
    assert c["evictions"] >= 1
    for rid, sid, base, ln, W in reqs:
        want = oracle_scan_rows([b"c"], [seq[base:base + ln]], W, 25, sorted_keys(keys[sid]))
        assert [tuple(int(x) for x in row) for row in got[rid]] == [w[3:] for w in want], rid


def test_minimizer_serving_budget_counts_real_table_bytes(small_corpus):
    """The serving budget holds minimizer tables at their real size: three
    samples of the same keys (tables of equal bytes) under a budget of two
    and a half tables; after every phase the engine's resident bytes are
    exactly its resident tables' bytes, never above the budget, and every
    request equals the oracle."""
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    sw, ew, _ = hipibs.pack_sequences([seq])
    n = len(seq)
    keys = samples[0][0]
    opts = dict(minimizer_len=21, minimizer_cap=8, load_factor=0.3)
    one_table = hipibs.Table.from_keys(keys, 25, **opts).info()["bytes"]
    budget = int(2.5 * one_table)
    eng = hipibs.ServeEngine(25, budget, max_inflight=2, max_windows_per_launch=256, **opts)
    eng.set_reference(sw, ew, n)
    for sid in range(3):
        eng.add_sample(sid, keys)
    want = {}
    for phase in range(4):
        pending = set()
        for j in range(3):
            rid = phase * 3 + j
            sid = (phase + j) % 3
            eng.submit(rid, sid, 1000 * rid, 200_000, 10_000)
            pending.add(rid)
            want[rid] = sid
        while pending:
            for rid in eng.poll(timeout_ms=2000):
                rows = eng.fetch(rid)[0]
                base = 1000 * rid
                ref_rows = oracle_scan_rows([b"c"], [seq[base:base + 200_000]], 10_000, 25,
                                            sorted_keys(keys))
                assert [tuple(int(x) for x in r) for r in rows] == [w[3:] for w in ref_rows]
                pending.discard(rid)
        c = eng.counters()
        assert c["resident_bytes"] == c["resident_tables"] * one_table, c
        assert c["resident_bytes"] <= budget
    assert eng.counters()["evictions"] >= 1
    eng.free()


def coverage_text_gpu(t, ids, seqs, k):
    """fastibsmapper file content from the device hit mask + coverage."""
    parts = []
    for cid, s in zip(ids, seqs):
        n = len(s)
        npos = n - k + 1 if n >= k else 0
        sw, ew, _ = hipibs.pack_sequences([s])
        d_sw, d_ew = D.from_array(sw), D.from_array(ew)
        d_hit = D.zeros(((npos + 31) // 32 + 1) * 4)
        d_cov = D.zeros(max(n, 1))
        if npos:
            hipibs._check(hipibs._lib.hipibs_hitmask_device(t.handle, d_sw.ptr, d_ew.ptr, 0, npos,
                                                            n, d_hit.ptr, None), "hitmask")
        hipibs._check(hipibs._lib.hipibs_coverage_device(d_hit.ptr, npos, n, k, d_cov.ptr, None),
                      "coverage")
        hipibs.synchronize()
        cov = d_cov.to_array(np.uint8, n) if n else np.zeros(0, np.uint8)
        parts.append(cid + b"\n" + coverage_line(cov) + b"\n")
    return b"".join(parts)


@pytest.mark.parametrize("k", [33, 47, 64])
def test_wide_tables_match_reference_string_path(tmp_path, k, expected):
    """k > 32 (two-word keys): tables built from a KMC database (16 B
    suffixes decoded on the device) and from host keys with duplicates scan
    to the reference's string-path TSVs and coverage; membership and the
    streamed intersect against set arithmetic."""
    recs, keys = corpus.edge_config(k)
    ids = [r[0] for r in recs]
    seqs = [r[1] for r in recs]
    p = str(tmp_path / "db")
    hipibs.kmc_write(p, keys, k, counts=None, sig_len=9, counter_size=1, version=0x200,
                     min_count=1, max_count=255)
    note(f"wide k={k}: builds")
    t1 = hipibs.Table.from_kmc(p)
    t2 = hipibs.Table.from_keys(np.concatenate([keys, keys[:5000]]), k)
    for t in (t1, t2):
        info = t.info()
        assert (info["n_keys"], info["k"]) == (keys.shape[0], k)
    e = expected["edge_wide"][str(k)]
    for W, want in e["tsv"].items():
        note(f"wide k={k} W={W}")
        both = hipibs.scan_sequences([t1, t2], ids, seqs, int(W), batch_bases=1 << 22)
        for rows in both:
            tsv = hipibs.format_tsv(rows)
            assert (sha256(tsv), len(tsv)) == (want["sha256"], want["bytes"]), (k, W)
    note(f"wide k={k}: coverage")
    cov = coverage_text_gpu(t1, ids, seqs, k)
    assert (sha256(cov), len(cov)) == (e["coverage_sha256"], e["coverage_bytes"])
    assert t1.probe_count(keys) == keys.shape[0]
    rng = np.random.default_rng(k)
    q = rng.integers(0, 1 << 63, size=(20000, 2), dtype=np.uint64)
    q[:, 0] &= np.uint64((1 << (2 * (k - 32))) - 1) if k < 64 else np.uint64(~0 & (2**64 - 1))
    q[:10000] = keys[rng.integers(0, keys.shape[0], 10000)]
    pair = np.dtype([("h", "<u8"), ("l", "<u8")])
    want_hits = int(np.isin(np.ascontiguousarray(q).view(pair),
                            np.ascontiguousarray(keys).view(pair)).sum())
    assert t2.probe_count(q) == want_hits
    sub = np.unique(np.concatenate([keys[::3], sorted_keys_wide(q[10000:])]), axis=0)
    p2 = str(tmp_path / "db2")
    hipibs.kmc_write(p2, sub, k, counts=None, sig_len=9, counter_size=0, version=0x200,
                     min_count=1, max_count=1)
    inter = int(np.isin(sub.view(pair), np.ascontiguousarray(keys).view(pair)).sum())
    assert t1.probe_count_kmc(p2) == inter
    t1.free()
    t2.free()


def test_k64_palindromes_scan_like_the_oracle(tmp_path):
    """k = 64: T^32 A^32 is the one canonical code whose high word is all ones
    (the tables' free mark) and A^32 T^32 the one whose low word is. Contigs
    holding both, with N runs beside them, scan and map like the oracle
    whether the table holds them (from host keys, from device keys, from a
    KMC database) or not."""
    k = 64
    rng = np.random.default_rng(64)
    pal = np.array([0] * 32 + [3] * 32, dtype=np.uint8)
    top = np.array([3] * 32 + [0] * 32, dtype=np.uint8)
    contigs = []
    for i in range(6):
        parts = []
        for j in range(8):
            parts.append(rng.integers(0, 4, int(rng.integers(1, 300))).astype(np.uint8))
            parts.append(top if (i + j) % 2 else pal)
            if j % 3 == 2:
                parts.append(np.full(int(rng.integers(1, 70)), 4, np.uint8))
        contigs.append(np.concatenate(parts))
    ids = [b"c%d" % i for i in range(len(contigs))]
    seqs = [corpus.to_text(c) for c in contigs]
    every = sorted_keys_wide(np.concatenate([corpus.sequence_kmers_wide([c], k) for c in contigs]))
    ones = np.uint64(2 ** 64 - 1)
    is_top = (every[:, 0] == ones) & (every[:, 1] == 0)
    is_pal = (every[:, 0] == 0) & (every[:, 1] == ones)
    assert is_top.sum() == 1 and is_pal.sum() == 1
    half = every[rng.random(every.shape[0]) < 0.5]
    with_both = sorted_keys_wide(np.concatenate([half, every[is_top | is_pal]]))
    without_top = np.ascontiguousarray(with_both[~((with_both[:, 0] == ones) &
                                                   (with_both[:, 1] == 0))])
    for keys, name in ((with_both, "with"), (without_top, "without")):
        p = str(tmp_path / f"db_{name}")
        hipibs.kmc_write(p, keys, k, counts=None, sig_len=9, counter_size=1, version=0x200,
                         min_count=1, max_count=255)
        d = D.from_array(np.ascontiguousarray(keys[::-1]))
        tabs = [hipibs.Table.from_keys(keys, k), hipibs.Table.from_kmc(p),
                hipibs.Table.from_device_keys(d.ptr, keys.shape[0], k)]
        for W in (100, 1000):
            note(f"k=64 palindromes ({name} T^32A^32) W={W}")
            want = oracle_scan_rows_wide(ids, seqs, W, k, keys)
            for t, rows in zip(tabs, hipibs.scan_sequences(tabs, ids, seqs, W)):
                assert t.info()["n_keys"] == keys.shape[0]
                assert rows == want
        assert tabs[0].probe_count(every) == keys.shape[0]
        cov = coverage_text_gpu(tabs[1], ids, seqs, k)
        from util import oracle_coverage_text_wide
        assert cov == oracle_coverage_text_wide(ids, seqs, k, keys)
        for t in tabs:
            t.free()
        d.free()


def test_wide_device_build_stores_repeated_keys_once():
    """Device keys of k > 32 may repeat, as for k <= 32: 100,000 copies of one
    key among distinct ones are stored once, and quickly; rows whose high
    word is all ones (the encoder's invalid mark) are skipped."""
    import time
    recs, keys = corpus.edge_config(47)
    none = np.full((1000, 2), 2 ** 64 - 1, dtype=np.uint64)  # rows with no key: skipped
    rep = np.ascontiguousarray(np.concatenate([np.repeat(keys[:1], 100_000, axis=0), keys,
                                               keys[::5], none]))
    rng = np.random.default_rng(5)
    rep = np.ascontiguousarray(rep[rng.permutation(rep.shape[0])])
    d = D.from_array(rep)
    t0 = time.time()
    t = hipibs.Table.from_device_keys(d.ptr, rep.shape[0], 47)
    assert time.time() - t0 < 30
    assert t.info()["n_keys"] == keys.shape[0]
    assert t.probe_count(keys) == keys.shape[0]
    t.free()
    d.free()


def test_wide_serving_equals_single_scans(tmp_path, expected):
    """Serving at k = 47: samples from host keys and from a KMC database,
    a budget of one table (evictions, re-admissions), mixed windows; every
    request equals the reference string path's rows for its range."""
    k = 47
    recs, keys = corpus.edge_config(k)
    seq = recs[0][1]  # chr1, 240 kbp
    sw, ew, _ = hipibs.pack_sequences([seq])
    n = len(seq)
    samples = [keys, keys[::2], keys[1::3]]
    kmc = str(tmp_path / "s2")
    hipibs.kmc_write(kmc, samples[2], k, sig_len=9, counter_size=1, version=0x200,
                     min_count=1, max_count=255)
    one = hipibs.Table.from_keys(samples[0], k)
    budget = int(1.5 * one.info()["bytes"])
    one.free()
    eng = hipibs.ServeEngine(k, budget, max_inflight=3, max_windows_per_launch=512)
    eng.set_reference(sw, ew, n)
    eng.add_sample(0, samples[0])
    eng.add_sample(1, samples[1])
    eng.add_sample_kmc(2, kmc)
    rng = np.random.default_rng(9)
    reqs = []
    for i in range(24):
        ln = int(rng.integers(100, n // 2))
        reqs.append((i, i % 3, int(rng.integers(0, n - ln)), ln, int(rng.choice([1000, 5000]))))
    out = {}
    for phase in (reqs[:12], reqs[12:]):
        for r in phase:
            eng.submit(*r)
        pending = {r[0] for r in phase}
        while pending:
            for rid in eng.poll(timeout_ms=2000):
                out[rid] = eng.fetch(rid)[0]
                pending.discard(rid)
    c = eng.counters()
    eng.free()
    # at most two of the three tables fit together, and the second phase asks
    # for all three again
    assert c["evictions"] >= 1 and c["admissions"] >= 4
    for rid, sid, base, ln, W in reqs:
        want = oracle_scan_rows_wide([b"c"], [seq[base:base + ln]], W, k,
                                     sorted_keys_wide(samples[sid]))
        got = [tuple(int(x) for x in row) for row in out[rid]]
        assert got == [w[3:] for w in want], rid


def test_wide_tables_refused_where_unsupported(tmp_path):
    recs, keys = corpus.edge_config(47)
    with pytest.raises(hipibs.HipibsError, match="minimizer"):
        hipibs.Table.from_keys(keys[:100], 47, minimizer_len=9)
    with pytest.raises(hipibs.HipibsError, match="128 B"):
        hipibs.Table.from_keys(keys[:100], 47, bucket_slots=8)
