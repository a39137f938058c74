"""Deterministic synthetic corpora for the parity tests (numpy only).

Every random draw comes from splitmix64 over a (seed, stream, index)
counter, so the corpora are identical on every machine and numpy version.
Bases are A=0 C=1 G=2 T=3; N is 4 (written as 'N').

The small configuration (1 Mbp reference, k=25, three samples at SNP rates
0.1% / 1% / 5% with 2% of 50 kbp windows replaced by random sequence,
10x 150 bp reads at 0.5% errors, k-mers counted with min count 2 and
counters capped at 255) is the reference's CPU-runnable case; the edge
corpus reproduces the contig shapes of scripts/make_test_corpus.sh:57-117
(lowercase, N runs of 1 / k-1 / k / 3k, a contig shorter than k, one of
exactly k) with our own generator.
"""
from __future__ import annotations

import numpy as np

GOLD = np.uint64(0x9E3779B97F4A7C15)
ACGT = np.frombuffer(b"ACGTN", dtype=np.uint8)


def splitmix(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """n pseudo-random uint64 of (seed, stream), indices start..start+n-1."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0xD6E8FEB86659FD93) & (2**64 - 1))
        x = base + (np.arange(start, start + n, dtype=np.uint64) + np.uint64(1)) * GOLD
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    return (splitmix(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def random_bases(seed: int, stream: int, n: int) -> np.ndarray:
    return (splitmix(seed, stream, n) >> np.uint64(62)).astype(np.uint8)


def make_reference(n: int, seed: int, n_frac: float = 0.02, run_lo: int = 100,
                   run_hi: int = 1000) -> np.ndarray:
    """Uniform ACGT codes with N runs (code 4) of length U[run_lo, run_hi]
    covering about n_frac of the positions."""
    ref = random_bases(seed, 1, n)
    want = int(n * n_frac)
    r = splitmix(seed, 2, 4096)
    placed, i = 0, 0
    while placed < want and i + 1 < r.size:
        ln = run_lo + int(r[i] % np.uint64(run_hi - run_lo + 1))
        st = int(r[i + 1] % np.uint64(max(1, n - ln)))
        ref[st:st + ln] = 4
        placed += ln
        i += 2
    return ref


def mutate(ref: np.ndarray, snp_rate: float, seed: int, window: int = 50000,
           replace_frac: float = 0.02) -> np.ndarray:
    """i.i.d. substitutions at snp_rate (never to the same base, N kept),
    plus max(1, round(replace_frac * windows)) windows of `window` bases
    replaced by random sequence."""
    hap = ref.copy()
    n = hap.size
    u = uniform(seed, 3, n)
    delta = (1 + (splitmix(seed, 4, n) % np.uint64(3))).astype(np.uint8)
    snp = (u < snp_rate) & (hap < 4)
    hap[snp] = (hap[snp] + delta[snp]) % 4
    n_win = max(1, n // window)
    n_rep = max(1, int(round(replace_frac * n_win)))
    picks = splitmix(seed, 5, n_rep) % np.uint64(n_win)
    for j, w in enumerate(picks):
        s = int(w) * window
        e = min(n, s + window)
        rnd = random_bases(seed, 100 + j, e - s)
        keep_n = hap[s:e] == 4
        seg = rnd.copy()
        seg[keep_n] = 4
        hap[s:e] = seg
    return hap


def simulate_reads(hap: np.ndarray, coverage: float, read_len: int, err: float,
                   seed: int) -> np.ndarray:
    """(R, read_len) codes: uniform starts, random strand, i.i.d.
    substitution errors at rate err (N stays N)."""
    n = hap.size
    R = int(coverage * n / read_len)
    starts = (splitmix(seed, 6, R) % np.uint64(n - read_len + 1)).astype(np.int64)
    idx = starts[:, None] + np.arange(read_len)[None, :]
    reads = hap[idx]
    u = uniform(seed, 7, R * read_len).reshape(R, read_len)
    delta = (1 + (splitmix(seed, 8, R * read_len) % np.uint64(3))).astype(np.uint8).reshape(
        R, read_len)
    e = (u < err) & (reads < 4)
    reads[e] = (reads[e] + delta[e]) % 4
    strand = (splitmix(seed, 9, R) & np.uint64(1)).astype(bool)
    rc = np.where(reads[strand] < 4, 3 - reads[strand], 4)[:, ::-1]
    reads[strand] = rc
    return reads


def canonical_codes(codes: np.ndarray, k: int):
    """Canonical k-mer codes of every k-window of each row of a 2-D code
    array (N = 4); returns (codes (R, L-k+1) uint64, valid bool)."""
    R, L = codes.shape
    P = L - k + 1
    fwd = np.zeros((R, P), dtype=np.uint64)
    rc = np.zeros((R, P), dtype=np.uint64)
    bad = np.zeros((R, P), dtype=bool)
    c64 = codes.astype(np.uint64) & np.uint64(3)
    for j in range(k):
        col = c64[:, j:j + P]
        fwd = (fwd << np.uint64(2)) | col
        rc = rc | ((np.uint64(3) - col) << np.uint64(2 * j))
        bad |= codes[:, j:j + P] > 3
    return np.minimum(fwd, rc), ~bad


def count_kmers(reads: np.ndarray, k: int, ci: int = 2, cs: int = 255):
    """KMC semantics: canonical k-mers of reads, k-mers with a non-ACGT base
    skipped, keep count >= ci, counters capped at cs."""
    codes, valid = canonical_codes(reads, k)
    keys, counts = np.unique(codes[valid], return_counts=True)
    keep = counts >= ci
    return keys[keep], np.minimum(counts[keep], cs).astype(np.uint32)


def sequence_kmers(seqs, k: int) -> np.ndarray:
    """Distinct canonical k-mers of whole sequences (ci = 1)."""
    out = []
    for s in seqs:
        if s.size >= k:
            c, v = canonical_codes(s[None, :], k)
            out.append(c[v])
    if not out:
        return np.zeros(0, dtype=np.uint64)
    return np.unique(np.concatenate(out))


def canonical_codes_wide(codes: np.ndarray, k: int):
    """canonical_codes for k = 33..64: codes as (high, low) uint64 words of
    the 2k-bit code (first base most significant); returns (hi, lo, valid)."""
    R, L = codes.shape
    P = L - k + 1
    fh = np.zeros((R, P), dtype=np.uint64)
    fl = np.zeros((R, P), dtype=np.uint64)
    rh = np.zeros((R, P), dtype=np.uint64)
    rl = np.zeros((R, P), dtype=np.uint64)
    bad = np.zeros((R, P), dtype=bool)
    c64 = codes.astype(np.uint64) & np.uint64(3)
    for j in range(k):
        col = c64[:, j:j + P]
        fh = (fh << np.uint64(2)) | (fl >> np.uint64(62))
        fl = (fl << np.uint64(2)) | col
        comp = np.uint64(3) - col
        if j < 32:
            rl |= comp << np.uint64(2 * j)
        else:
            rh |= comp << np.uint64(2 * (j - 32))
        bad |= codes[:, j:j + P] > 3
    if k < 64:
        fh &= np.uint64((1 << (2 * (k - 32))) - 1)
    f_lt = (fh < rh) | ((fh == rh) & (fl < rl))
    return np.where(f_lt, fh, rh), np.where(f_lt, fl, rl), ~bad


def sequence_kmers_wide(seqs, k: int) -> np.ndarray:
    """Distinct canonical k-mers (k = 33..64) of whole sequences as an
    (n, 2) array of (high, low) words, sorted."""
    out = []
    for s in seqs:
        if s.size >= k:
            h, l, v = canonical_codes_wide(s[None, :], k)
            out.append(np.stack([h[v], l[v]], axis=1))
    if not out:
        return np.zeros((0, 2), dtype=np.uint64)
    return np.unique(np.concatenate(out), axis=0)


def revcomp(codes: np.ndarray, k: int) -> np.ndarray:
    """Reverse complement of 2-bit packed k-mer codes."""
    c = np.asarray(codes, dtype=np.uint64)
    out = np.zeros_like(c)
    for j in range(k):
        out = (out << np.uint64(2)) | (np.uint64(3) - ((c >> np.uint64(2 * j)) & np.uint64(3)))
    return out


def to_text(codes: np.ndarray, lower: bool = False) -> bytes:
    t = ACGT[codes].tobytes()
    return t.lower().replace(b"n", b"N") if lower else t


def fasta(records, width: int = 80) -> bytes:
    out = []
    for name, text in records:
        out.append(b">" + name + b"\n")
        for i in range(0, len(text), width):
            out.append(text[i:i + width] + b"\n")
    return b"".join(out)


# ---- configurations --------------------------------------------------------

SMALL_SEED = 0x1B50000 + 100 * 1


def small_reads(seed: int = SMALL_SEED, n: int = 1_000_000):
    """The 1 Mbp reference and its three samples' reads: 150 bp at 10x of
    haplotypes with 0.1% / 1% / 5% SNPs, 0.5% substitution errors."""
    ref = make_reference(n, seed)
    reads = []
    for i, rate in enumerate((0.001, 0.01, 0.05)):
        hap = mutate(ref, rate, seed + i + 1)
        reads.append(simulate_reads(hap, 10.0, 150, 0.005, seed + 10 + i))
    return ref, reads


def small_config(seed: int = SMALL_SEED, n: int = 1_000_000, k: int = 25):
    """The 1 Mbp reference and its three samples' (keys, counts)."""
    ref, reads = small_reads(seed, n)
    return ref, [count_kmers(r, k, ci=2, cs=255) for r in reads]


def edge_config(k: int, seed: int = 0x1B5E000):
    """Edge-case reference contigs and a sample built from a mutated copy
    (1% SNPs plus three dropout blocks), every k-mer of the copy kept."""
    records, sample_seqs = edge_sequences(k, seed)
    keys = sequence_kmers(sample_seqs, k) if k <= 32 else sequence_kmers_wide(sample_seqs, k)
    return records, keys


def edge_sequences(k: int, seed: int = 0x1B5E000):
    """edge_config's reference records and the sample's sequences (codes)."""
    def with_n_runs(a):
        a = a.copy()
        for st, ln in ((1000, 1), (5000, k - 1), (9000, k), (13000, 3 * k)):
            if st + ln <= a.size:
                a[st:st + ln] = 4
        return a

    contigs = [
        (b"chr1", random_bases(seed, 11, 240_000), False),
        (b"chr2_lowercase", random_bases(seed, 12, 100_000), True),
        (b"chr3_withN", with_n_runs(random_bases(seed, 13, 60_000)), False),
        (b"chr4_short", random_bases(seed, 14, k - 3), False),
        (b"chr5_exactk", random_bases(seed, 15, k), False),
        (b"chr6_mixed", with_n_runs(random_bases(seed, 16, 50_000)), True),
    ]
    sample_seqs = []
    for ci, (name, codes, lower) in enumerate(contigs):
        s = codes.copy()
        u = uniform(seed, 20 + ci, s.size)
        d = (1 + (splitmix(seed, 30 + ci, s.size) % np.uint64(3))).astype(np.uint8)
        m = (u < 0.01) & (s < 4)
        s[m] = (s[m] + d[m]) % 4
        keep = np.ones(s.size, dtype=bool)
        if ci == 0:
            for a, b in ((20_000, 25_000), (100_000, 102_000), (180_000, 180_100)):
                keep[a:b] = False
        sample_seqs.append(s[keep])
    records = [(name, to_text(codes, lower)) for name, codes, lower in contigs]
    return records, sample_seqs
