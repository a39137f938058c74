"""Host-side pieces of libhipibs and its Python layer, on the CPU: the C
ABI exports, the KMC reader/writer against the oracle's reader and the
reference-written expectations, packing, FASTA parsing and window
geometry, and argument errors. No GPU compute is called here."""
import os
import subprocess

import numpy as np
import pytest

import corpus
from util import ROOT, oracle

import hipibs


def test_library_exports_every_declared_symbol():
    import re
    with open(os.path.join(ROOT, "include", "hipibs.h")) as f:
        decl = set(re.findall(r"\b(hipibs_[a-z_]+)\s*\(", f.read()))
    out = subprocess.run(["nm", "-D", "--defined-only", hipibs.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert decl <= syms, decl - syms
    assert decl == set(hipibs.EXPORTED)


def test_no_device_is_reported_cleanly():
    n, why = hipibs.device_count()
    if n == 0:
        assert why


@pytest.mark.parametrize("variant", [(25, 9, 9, 1, 0x200), (25, 5, 5, 2, 0x200),
                                     (25, 9, 11, 4, 0x200), (25, 9, 0, 1, 0), (25, 1, 7, 0, 0x200),
                                     (32, 8, 9, 1, 0x200), (32, 0, 9, 1, 0x200), (31, 7, 9, 3, 0),
                                     (15, 3, 5, 1, 0x200)])
def test_kmc_round_trip(tmp_path, variant):
    k, p, s, c, v = variant
    _, keys = corpus.edge_config(k)
    keys = keys[::3]
    counts = (np.arange(keys.size) % 300 + 1).astype(np.uint32)
    pre = str(tmp_path / "db")
    hipibs.kmc_write(pre, keys, k, counts=counts, p=p, sig_len=max(s, 5), counter_size=c,
                     version=v, min_count=1, max_count=2 ** 31)
    info = hipibs.kmc_header(pre)
    assert (info.kmer_length, info.lut_prefix_length, info.counter_size, info.total_kmers,
            info.kmc_version) == (k, p, c, keys.size, v)
    for threads in (1, 3):
        got = hipibs.kmc_read_keys(pre, threads=threads)
        assert np.array_equal(np.sort(got), np.sort(keys))
    import ctypes as C
    out = np.zeros(keys.size, dtype=np.uint64)
    kk = C.c_uint32(0)
    assert oracle().orc_kmc_read(pre.encode(), out.ctypes.data, out.size, C.byref(kk)) == keys.size
    assert np.array_equal(np.sort(out), np.sort(keys))


def test_kmc_reader_on_large_multithreaded(tmp_path):
    keys = np.unique(corpus.splitmix(1, 2, 3_000_000) & np.uint64((1 << 50) - 1))
    keys = np.minimum(keys, corpus.revcomp(keys, 25))
    keys = np.unique(keys)
    pre = str(tmp_path / "big")
    hipibs.kmc_write(pre, keys, 25)
    got = hipibs.kmc_read_keys(pre, threads=7)
    assert np.array_equal(np.sort(got), keys)


def test_kmc_errors(tmp_path):
    with pytest.raises(hipibs.HipibsError, match="cannot open"):
        hipibs.kmc_header(str(tmp_path / "missing"))
    bad = tmp_path / "bad.kmc_pre"
    bad.write_bytes(b"XXXX" + b"\0" * 100 + b"XXXX")
    with pytest.raises(hipibs.HipibsError, match="markers"):
        hipibs.kmc_header(str(tmp_path / "bad"))
    keys = np.array([1, 2, 2], dtype=np.uint64)
    with pytest.raises(hipibs.HipibsError, match="duplicate"):
        hipibs.kmc_write(str(tmp_path / "d"), keys, 25)
    with pytest.raises(hipibs.HipibsError, match="wider"):
        hipibs.kmc_write(str(tmp_path / "w"), np.array([1 << 60], dtype=np.uint64), 25)
    with pytest.raises(hipibs.HipibsError):
        hipibs.kmc_write(str(tmp_path / "g"), keys[:1], 25, p=6)  # (k - p) % 4 != 0
    # truncated suffix file
    good = str(tmp_path / "t")
    hipibs.kmc_write(good, np.arange(1000, dtype=np.uint64) * 7919, 25)
    suf = tmp_path / "t.kmc_suf"
    data = suf.read_bytes()
    suf.write_bytes(data[:len(data) // 2] + b"KMCS")
    with pytest.raises(hipibs.HipibsError, match="smaller"):
        hipibs.kmc_read_keys(good)


def unpack(sw, ew, n):
    bases = np.zeros(n, dtype=np.uint8)
    for b in range(n):
        bases[b] = (int(sw[b >> 5]) >> (2 * (b & 31))) & 3
        if (int(ew[b >> 6]) >> (b & 63)) & 1:
            bases[b] = 4
    return bases


def test_pack_layout():
    rng = np.random.default_rng(3)
    seqs = []
    for L in (0, 1, 31, 64, 65, 200, 1000):
        c = rng.integers(0, 5, L).astype(np.uint8)
        t = corpus.to_text(c)
        if L > 10:
            t = t[:5] + b"x\rac" + t[9:]
        seqs.append(t)
    sw, ew, off = hipibs.pack_sequences(seqs)
    total = int(off[-1])
    got = unpack(sw, ew, total)
    table = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3, ord("a"): 0, ord("c"): 1,
             ord("g"): 2, ord("t"): 3}
    want = np.array([table.get(ch, 4) for ch in b"".join(seqs)], dtype=np.uint8)
    assert np.array_equal(got, want)
    assert sw.size == (total + 31) // 32 + hipibs.PACK_PAD_WORDS


def test_read_sequence_file(tmp_path):
    f = tmp_path / "r.fasta"
    f.write_bytes(b">c1 desc\nACGT\nac\r\n\n>empty\n>c2\nNNNN\nA")
    r = hipibs.read_sequence_file(str(f))
    assert r.ids == [b"c1 desc", b"c2"]
    assert r.sequences == [b"ACGTac\r", b"NNNNA"]
    q = tmp_path / "r.fastq"
    q.write_bytes(b"@r1\nACGT\n+\nIIII\n@r2\nGG\n+\nII\n")
    r = hipibs.read_sequence_file(str(q))
    assert r.ids == [b"r1", b"r2"] and r.sequences == [b"ACGT", b"GG"]
    import gzip
    g = tmp_path / "r.fa.gz"
    g.write_bytes(gzip.compress(b">x\nACGT\n"))
    assert hipibs.read_sequence_file(str(g)).sequences == [b"ACGT"]
    e = tmp_path / "e.fa"
    e.write_bytes(b"")
    with pytest.raises(ValueError):
        hipibs.read_sequence_file(str(e))


def test_chunk_sequences_matches_oracle():
    lib = oracle()
    for L, W, k in [(1_000_000, 50000, 25), (100, 33, 25), (0, 100, 25), (49976, 50000, 25)]:
        got = hipibs.chunk_sequences([L], W, k)
        cap = L // (W - k) + 3
        st = np.zeros(cap, dtype=np.int64)
        en = np.zeros(cap, dtype=np.int64)
        n = lib.orc_windows(L, W, k, st.ctypes.data, en.ctypes.data, cap)
        assert [(s, e) for _, s, e in got] == list(zip(st[:n].tolist(), en[:n].tolist()))
    with pytest.raises(ValueError):
        hipibs.chunk_sequences([10], 25, 25)


def test_output_names():
    assert hipibs.reference_name("MorexV3_genome.fasta") == "MorexV3"
    assert hipibs.reference_name("ref.fa") == "ref.fa"
    assert hipibs.output_name("/data/BW_01002/", "test_genome.fasta", 50000) == \
        "BW_01002_v_test_50000.tsv"
