"""Writes the golden fixtures under tests/golden/ from the reference's own
code, compiled from its sources into oracle/_ref/libref.so (make -C oracle).

    python tests/golden/make_golden.py

Inputs are regenerated deterministically by corpus.py; KMC databases are
written by libhipibs's KMC writer and first checked to read back, key for
key, through the reference's KmcRaIndex and CKMCFile. Expected outputs:
  - scan TSVs through the reference's CPU-OPT probe (CKMCFile::IsKmer) and
    CPU-BATCHED probe (KmcRaIndex + probeBatched), which must agree;
  - fastibsmapper coverage files (SHA-256 + size);
  - KDBIntersect counts;
  - for k > 32 (the edge corpus at k = 33, 47, 64), scan TSVs and coverage
    through the reference's string path (getCanonicalKmersLegacy +
    CKmerAPI::from_string + CKMCFile::IsKmer), the databases first checked
    key for key through CKMCFile with the k-mers as strings.
Only data is committed (expected.json and the small TSVs); the reference
never travels.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "gpu-ibs_amd"))

import corpus  # noqa: E402
import hipibs  # noqa: E402

EDGE_KS = (25, 31, 32, 15)
WIDE_KS = (33, 47, 64)
WIDE_WS = (50000, 1000, 100)
EDGE_WS = (50000, 5000, 1000, 100, 33)
SMALL_W = 50000


def load_ref():
    path = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    lib = C.CDLL(path)
    lib.ref_scan_tsv.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    lib.ref_coverage.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
    lib.ref_list_keys.argtypes = [C.c_char_p, C.c_void_p, C.c_longlong]
    lib.ref_list_keys.restype = C.c_longlong
    lib.ref_kmc_is_kmer.argtypes = [C.c_char_p, C.c_void_p, C.c_longlong, C.c_void_p]
    lib.ref_kmc_is_kmer.restype = C.c_longlong
    lib.ref_ra_probe.argtypes = [C.c_char_p, C.c_void_p, C.c_longlong, C.c_void_p]
    lib.ref_ra_probe.restype = C.c_longlong
    lib.ref_intersect.argtypes = [C.c_char_p, C.c_char_p]
    lib.ref_intersect.restype = C.c_longlong
    lib.ref_scan_tsv_legacy.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_char_p]
    lib.ref_coverage_legacy.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
    lib.ref_kmc_is_kmer_strings.argtypes = [C.c_char_p, C.c_char_p, C.c_longlong, C.c_void_p]
    lib.ref_kmc_is_kmer_strings.restype = C.c_longlong
    return lib


def wide_strings(keys, k):
    """(n, 2) high/low codes as concatenated k-mer strings."""
    out = []
    for hi, lo in keys.tolist():
        v = (hi << 64) | lo
        out.append("".join("ACGT"[(v >> (2 * (k - 1 - i))) & 3] for i in range(k)))
    return "".join(out).encode()


def check_db_wide(ref, prefix, keys, k):
    """Every key of a k > 32 database is found by CKMCFile, given as a string,
    and the library's reader lists exactly the keys."""
    hit = np.zeros(keys.shape[0], dtype=np.uint8)
    assert ref.ref_kmc_is_kmer_strings(prefix.encode(), wide_strings(keys, k), keys.shape[0],
                                       hit.ctypes.data) == keys.shape[0], "CKMCFile misses a key"
    back = hipibs.kmc_read_keys(prefix)
    assert np.array_equal(np.unique(back, axis=0), keys), "reader lists other keys"


def sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        h.update(f.read())
    return h.hexdigest(), os.path.getsize(path)


def check_db(ref, prefix, keys):
    """The written database lists exactly `keys` through the reference."""
    n = keys.size
    out = np.zeros(max(n, 1), dtype=np.uint64)
    got = ref.ref_list_keys(prefix.encode(), out.ctypes.data, n)
    assert got == n, (got, n)
    assert np.array_equal(np.sort(out[:n]), np.sort(keys)), "KmcRaIndex listing differs"
    sample = keys[:: max(1, n // 20000)]
    hit = np.zeros(sample.size, dtype=np.uint8)
    assert ref.ref_kmc_is_kmer(prefix.encode(), sample.ctypes.data, sample.size,
                               hit.ctypes.data) == sample.size, "CKMCFile misses a key"


def scan_both(ref, prefix, fa, W, out):
    a, b = out + ".opt", out + ".bat"
    na = ref.ref_scan_tsv(prefix.encode(), fa.encode(), W, a.encode(), 0)
    nb = ref.ref_scan_tsv(prefix.encode(), fa.encode(), W, b.encode(), 1)
    assert na > 0 and na == nb, (na, nb)
    ta, tb = open(a, "rb").read(), open(b, "rb").read()
    assert ta == tb, "CPU-OPT and CPU-BATCHED disagree"
    return ta


def main():
    ref = load_ref()
    exp = {"small": {}, "edge": {}, "kmc_formats": {}}
    with tempfile.TemporaryDirectory() as td:
        # ---- small configuration ----
        refc, samples = corpus.small_config()
        fa = os.path.join(td, "small_genome.fasta")
        with open(fa, "wb") as f:
            f.write(corpus.fasta([(b"chr1", corpus.to_text(refc))]))
        exp["small"]["fasta_sha256"] = sha(fa)[0]
        prefixes = []
        for i, (keys, counts) in enumerate(samples):
            p = os.path.join(td, f"s{i}")
            hipibs.kmc_write(p, keys, 25, counts=counts, p=9, sig_len=9, counter_size=1,
                             version=0x200, min_count=2, max_count=255)
            check_db(ref, p, keys)
            prefixes.append(p)
            tsv = scan_both(ref, p, fa, SMALL_W, os.path.join(td, f"s{i}.tsv"))
            name = f"small_s{i}_{SMALL_W}.tsv"
            with open(os.path.join(HERE, name), "wb") as f:
                f.write(tsv)
            cov = os.path.join(td, f"s{i}.cov")
            assert ref.ref_coverage(p.encode(), fa.encode(), cov.encode()) == 1
            exp["small"][f"s{i}"] = {"n_keys": int(keys.size), "tsv": name,
                                     "tsv_sha256": hashlib.sha256(tsv).hexdigest(),
                                     "coverage_sha256": sha(cov)[0],
                                     "coverage_bytes": sha(cov)[1]}
        inter = {}
        for a in range(3):
            for b in range(a + 1, 3):
                inter[f"{a}-{b}"] = int(ref.ref_intersect(prefixes[a].encode(),
                                                           prefixes[b].encode()))
        exp["small"]["intersect"] = inter
        print("small:", exp["small"]["intersect"])

        # ---- edge corpus at several k and window sizes ----
        for k in EDGE_KS:
            recs, keys = corpus.edge_config(k)
            fa = os.path.join(td, f"edge{k}_genome.fasta")
            with open(fa, "wb") as f:
                f.write(corpus.fasta(recs))
            p = os.path.join(td, f"edge{k}")
            hipibs.kmc_write(p, keys, k, counts=None, sig_len=min(9, k), counter_size=1,
                             version=0x200, min_count=1, max_count=255)
            check_db(ref, p, keys)
            e = {"n_keys": int(keys.size), "fasta_sha256": sha(fa)[0], "tsv": {}}
            for W in EDGE_WS + (k + 1,):
                if W <= k:
                    continue
                tsv = scan_both(ref, p, fa, W, os.path.join(td, f"e{k}_{W}.tsv"))
                e["tsv"][str(W)] = {"sha256": hashlib.sha256(tsv).hexdigest(),
                                    "bytes": len(tsv)}
            cov = os.path.join(td, f"e{k}.cov")
            assert ref.ref_coverage(p.encode(), fa.encode(), cov.encode()) == len(recs)
            e["coverage_sha256"], e["coverage_bytes"] = sha(cov)
            exp["edge"][str(k)] = e
            print("edge k", k, "keys", keys.size)

        # ---- k > 32: the reference's string path ----
        exp["edge_wide"] = {}
        for k in WIDE_KS:
            recs, keys = corpus.edge_config(k)
            fa = os.path.join(td, f"wide{k}_genome.fasta")
            with open(fa, "wb") as f:
                f.write(corpus.fasta(recs))
            p = os.path.join(td, f"wide{k}")
            hipibs.kmc_write(p, keys, k, counts=None, sig_len=9, counter_size=1,
                             version=0x200, min_count=1, max_count=255)
            check_db_wide(ref, p, keys, k)
            e = {"n_keys": int(keys.shape[0]), "fasta_sha256": sha(fa)[0], "tsv": {}}
            for W in WIDE_WS + (k + 1,):
                out = os.path.join(td, f"w{k}_{W}.tsv")
                assert ref.ref_scan_tsv_legacy(p.encode(), fa.encode(), W, out.encode()) > 0
                e["tsv"][str(W)] = dict(zip(("sha256", "bytes"), sha(out)))
            cov = os.path.join(td, f"w{k}.cov")
            assert ref.ref_coverage_legacy(p.encode(), fa.encode(), cov.encode()) == len(recs)
            e["coverage_sha256"], e["coverage_bytes"] = sha(cov)
            exp["edge_wide"][str(k)] = e
            print("edge k", k, "keys", keys.shape[0])

        # ---- KMC format variants: one key set, several layouts ----
        recs, keys = corpus.edge_config(25)
        rng_codes = np.concatenate([corpus.splitmix(77, 1, 25000) & np.uint64((1 << 50) - 1),
                                    keys[::max(1, keys.size // 25000)]])
        rng_codes = np.minimum(rng_codes, corpus.revcomp(rng_codes, 25))
        variants = [("kmc2_p9_s9_c1", 9, 9, 1, 0x200), ("kmc2_p5_s5_c2", 5, 5, 2, 0x200),
                    ("kmc2_p9_s11_c4", 9, 11, 4, 0x200), ("kmc1_p9_c1", 9, 0, 1, 0),
                    ("kmc1_p5_c3", 5, 0, 3, 0), ("kmc2_p1_s7_c0", 1, 7, 0, 0x200)]
        for name, p, s, c, v in variants:
            pre = os.path.join(td, name)
            hipibs.kmc_write(pre, keys, 25, counts=None, p=p, sig_len=max(s, 5), counter_size=c,
                             version=v, min_count=1, max_count=255 if c else 1)
            check_db(ref, pre, keys)
            hit = np.zeros(rng_codes.size, dtype=np.uint8)
            nhit_ra = ref.ref_ra_probe(pre.encode(), rng_codes.ctypes.data, rng_codes.size,
                                       hit.ctypes.data)
            exp["kmc_formats"][name] = {"n_keys": int(keys.size),
                                        "random_probe_hits": int(nhit_ra)}
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump(exp, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "expected.json"))


if __name__ == "__main__":
    main()
