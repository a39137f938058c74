"""The benchmark configurations at their stated sizes, on the GPU: the
2 Gbp x 2e9-key k=25 scan, the 4.2 Gbp k=31 scan with per-base coverage,
the 48-sample panel and the 96-sample serving mix.

Everything is generated on the GPU from fixed seeds (the generators of
bench.py). A sample's keys are the k-mers of a mutated haplotype of the
reference (made distinct) kept by a hash of the key, plus random
canonical codes; a random code may repeat a kept k-mer, and every build
stores a key once.

What pins the results at these sizes:
- two table layouts with different probes (whole-k-mer buckets at load
  0.6, minimizer buckets with their grouped probe) give identical
  statistics on every window, and a scan repeated gives the same bytes;
- the oracle on slices: the oracle encodes the slice's text itself; its
  key set is the sample's keys that occur in the slice, found by a sorted
  search (numpy sort of the slice's codes, torch.searchsorted of the
  sample's key array) that shares no code with the library's tables;
- for coverage, sum = k x hit positions and each window's observed count =
  the hit-mask bits of its positions.
"""
import ctypes as C
import os
import sys
import time

import numpy as np
import pytest

from util import ROOT, oracle, oracle_scan_rows

sys.path.insert(0, ROOT)
import bench  # noqa: E402

import hipibs  # noqa: E402

pytestmark = pytest.mark.gpu

SEED = 0x1B50200


def note(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    yield t
    t.cuda.empty_cache()


# ---- sample construction ---------------------------------------------------------

def mutate(torch, codes, rate, g, lo=0, hi=None):
    """i.i.d. substitutions at `rate` over codes[lo:hi] (N stays N), in place."""
    hi = codes.numel() if hi is None else hi
    step = 1 << 28
    for b0 in range(lo, hi, step):
        b1 = min(hi, b0 + step)
        seg = codes[b0:b1]
        m = (torch.rand(b1 - b0, device="cuda", generator=g) < rate) & (seg < 4)
        d = torch.randint(1, 4, (b1 - b0,), dtype=torch.uint8, device="cuda", generator=g)
        seg[m] = (seg[m] + d[m]) % 4


def hap_codes(torch, hap, offs, k):
    """Canonical code per position of every chromosome of hap (-1 = invalid)."""
    hsw, hew = bench.pack(torch, hap)
    out = [bench.encode_range(torch, hsw, hew, int(offs[c]), int(offs[c + 1] - offs[c]) - k + 1,
                              int(offs[c + 1]), k) for c in range(len(offs) - 1)]
    del hsw, hew
    return out


def keep_mask(torch, x, frac, salt):
    """A k-mer is kept when a hash of it (and the sample's salt) falls below
    frac: the decision depends on the key only."""
    h = ((x ^ salt) * -7046029254386353131) >> 32
    return (h & 0xFFFFFFFF) < int(frac * 2 ** 32)


def distinct_codes(torch, codes):
    """The distinct valid codes of a list of code tensors, as disjoint
    classes (each made unique on its own, so no sort spans all of them)."""
    n_parts = 4
    vs = [c[c >= 0] for c in codes]
    cls = [bench.code_class(torch, v, n_parts).to(torch.uint8) for v in vs]

    def parts_of(p):
        for v, c in zip(vs, cls):
            yield v[c == p]
    out = bench.unique_partitioned(torch, parts_of, n_parts)
    del vs, cls
    return out


def sample_keys(torch, hap_distinct, target, k, seed):
    """~target keys: 80% of them distinct k-mers of the haplotype (kept by a
    hash of the key), the rest random canonical codes (which may repeat a
    kept one: the builds store a key once). int64 on the device."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n_hap = sum(x.numel() for x in hap_distinct)
    frac = min(1.0, 0.8 * target / max(1, n_hap))
    parts = [x[keep_mask(torch, x, frac, seed * 0x9E3779B1)] for x in hap_distinct]
    kept = sum(p.numel() for p in parts)
    parts.append(bench.random_canonical(torch, max(0, int(target) - kept), k, g))
    keys = torch.cat(parts)
    del parts
    return keys


def slice_keys(torch, keys, text, k):
    """The keys of the sample that occur in text, as the oracle's sorted
    uint64 key array: the oracle encodes text, numpy sorts its distinct
    codes, torch.searchsorted finds the sample keys among them."""
    lib = oracle()
    npos = len(text) - k + 1
    codes = np.zeros(npos, dtype=np.uint64)
    valid = np.zeros(npos, dtype=np.uint8)
    lib.orc_encode_positions(text, len(text), k, codes.ctypes.data, valid.ctypes.data)
    u = np.unique(codes[valid == 1])
    if u.size == 0:
        return u
    su = torch.from_numpy(u.view(np.int64)).cuda()
    found = []
    step = 1 << 28
    for b0 in range(0, keys.numel(), step):
        x = keys[b0:b0 + step]
        i = torch.searchsorted(su, x).clamp_(max=su.numel() - 1)
        found.append(x[su[i] == x])
    hit = torch.cat(found).cpu().numpy().view(np.uint64)
    return np.unique(hit)


def text_of(torch, codes):
    return np.frombuffer(b"ACGTN", dtype=np.uint8)[codes.cpu().numpy()].tobytes()


# ---- device scans ------------------------------------------------------------------

def windows_of(lens, W, k):
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    wins = hipibs.chunk_sequences(lens, W, k)
    win = np.zeros(len(wins), dtype=hipibs.WINDOW_DTYPE)
    win["base"] = np.array([offs[w[0]] + w[1] for w in wins], dtype=np.uint64)
    win["n_bases"] = np.array([w[2] - w[1] for w in wins], dtype=np.uint32)
    return wins, win


class Scanner:
    """hipibs_scan_windows_device over a resident packed reference."""

    def __init__(self, torch, sw, ew, n, win, W, k):
        self.torch, self.sw, self.ew, self.n, self.W, self.k = torch, sw, ew, n, W, k
        self.nw = win.size
        self.d_win = torch.from_numpy(win.view(np.uint8).copy()).cuda()

    def scan(self, tables):
        torch = self.torch
        nt = len(tables)
        d_st = torch.full((nt * self.nw * 4,), -1, dtype=torch.int32, device="cuda")
        res = C.c_uint32(0)
        wb = hipibs._lib.hipibs_scan_work_bytes(self.nw, nt, self.W, self.k, 0, C.byref(res))
        d_work = torch.zeros(max(wb, 8), dtype=torch.uint8, device="cuda")
        hipibs._check(hipibs._lib.hipibs_scan_windows_device(
            hipibs._table_array(tables), nt, self.sw.data_ptr(), self.ew.data_ptr(), self.n,
            self.d_win.data_ptr(), self.nw, self.W, d_st.data_ptr(), 0, d_work.data_ptr(), wb,
            None), "scan")
        torch.cuda.synchronize()
        return d_st.cpu().numpy().reshape(nt, self.nw, 4)


def check_slice(torch, ref, keys, stats, wins, offs, lens, c, a, L, W, k, label):
    """Oracle statistics of bases [a, a+L) of chromosome c (a on the window
    grid) against the windows of the full scan that lie inside it."""
    c0 = int(offs[c]) + int(a)
    text = text_of(torch, ref[c0:c0 + int(L)])
    ks = slice_keys(torch, keys, text, k)
    want = oracle_scan_rows([b"s"], [text], W, k, ks)
    first = next(i for i, w in enumerate(wins) if w[0] == c and w[1] == a)
    n_cmp = 0
    for j, row in enumerate(want):
        s, e = row[1], row[2]
        if e - s < W and a + e != lens[c]:
            continue  # cut by the slice's end, not by the chromosome's
        w = wins[first + j]
        assert (w[0], w[1], w[2]) == (c, a + s, a + e)
        assert tuple(int(x) for x in stats[first + j]) == row[3:], (label, j)
        n_cmp += 1
    nN = text.count(b"N")
    note(f"{label}: {n_cmp} windows of {L/1e6:.0f} Mbp equal the oracle "
         f"({ks.size} slice keys, {nN} N bases)")
    assert n_cmp >= L // W - 2 and nN > 0
    return text, ks


# ---- C2: 2 Gbp x 2e9 keys, k = 25 -----------------------------------------------

@pytest.mark.timeout(600)
def test_c2_scan_2gbp_2e9_keys_full_size(torch):
    k, W = 25, 50000
    n = 2_000_000_000
    t0 = time.time()
    ref = bench.gen_reference(torch, n, SEED)
    sw, ew = bench.pack(torch, ref)
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED + 1)
    hap = ref.clone()
    mutate(torch, hap, 0.01, g)
    offs = np.array([0, n], dtype=np.int64)
    hd = distinct_codes(torch, hap_codes(torch, hap, offs, k))
    del hap
    keys = sample_keys(torch, hd, 2.0e9, k, SEED + 2)
    del hd
    torch.cuda.empty_cache()
    note(f"C2: reference and {keys.numel()/1e9:.2f} G sample keys ({time.time()-t0:.0f} s)")
    t_whole = hipibs.Table.from_device_keys(keys.data_ptr(), keys.numel(), k)
    t_min = hipibs.Table.from_device_keys(keys.data_ptr(), keys.numel(), k, minimizer_len=21,
                                          minimizer_cap=16, load_factor=0.2)
    iw, im = t_whole.info(), t_min.info()
    note(f"C2: tables {iw['n_keys']/1e9:.3f} G distinct keys, {iw['bytes']/1e9:.1f} GB "
         f"(whole k-mer) / {im['bytes']/1e9:.1f} GB (minimizer)")
    assert iw["n_keys"] == im["n_keys"] and 1.95e9 < iw["n_keys"] <= 2.0e9
    wins, win = windows_of([n], W, k)
    sc = Scanner(torch, sw, ew, n, win, W, k)
    a = np.stack([sc.scan([t_whole])[0], sc.scan([t_min])[0]])
    b = sc.scan([t_whole])
    assert (a >= 0).all()
    assert np.array_equal(a[0], a[1]), "whole-k-mer and minimizer tables differ"
    assert np.array_equal(a[0], b[0]), "repeated scan differs"
    probes, hits = int(a[0, :, 0].sum()), int(a[0, :, 1].sum())
    note(f"C2: {len(wins)} windows identical across layouts and repeats; {probes/1e9:.3f} G "
         f"valid positions, hit rate {hits/probes:.3f}")
    assert 1.8e9 < probes < 2.0e9
    t_whole.free()
    t_min.free()
    # oracle on the first and the last 16 Mbp (the chromosome's ends, N runs)
    L = 16_000_000
    check_slice(torch, ref, keys, a[0], wins, offs, [n], 0, 0, L, W, k, "C2 first 16 Mbp")
    m = (n - L) // (W - k)
    check_slice(torch, ref, keys, a[0], wins, offs, [n], 0, m * (W - k), n - m * (W - k), W, k,
                "C2 last 16 Mbp")
    del keys, ref, sw, ew, sc
    torch.cuda.empty_cache()


# ---- C3: 4.2 Gbp in 7 chromosomes, k = 31, coverage -------------------------------

def hitmask(torch, t, sw, ew, c0, L, k):
    npos = L - k + 1
    d_hit = torch.zeros((npos + 31) // 32 + 1, dtype=torch.int32, device="cuda")
    hipibs._check(hipibs._lib.hipibs_hitmask_device(t.handle, sw.data_ptr(), ew.data_ptr(), c0,
                                                    npos, c0 + L, d_hit.data_ptr(), None),
                  "hitmask")
    return d_hit, npos


def coverage(torch, d_hit, npos, L, k):
    d_cov = torch.zeros(L, dtype=torch.uint8, device="cuda")
    hipibs._check(hipibs._lib.hipibs_coverage_device(d_hit.data_ptr(), npos, L, k,
                                                     d_cov.data_ptr(), None), "coverage")
    return d_cov


def bit_prefix(torch, d_hit, npos):
    """cumsum over positions of the hit bits (int64, length npos + 1)."""
    w = d_hit[:(npos + 31) // 32].to(torch.int64) & 0xFFFFFFFF
    bits = ((w.unsqueeze(1) >> torch.arange(32, device="cuda")) & 1).reshape(-1)[:npos]
    del w
    return torch.cat([torch.zeros(1, dtype=torch.int64, device="cuda"), bits.cumsum(0)])


@pytest.mark.timeout(900)
def test_c3_large_genome_k31_with_coverage_full_size(torch):
    k, W = bench.LARGE_K, 50000
    lens = [int(m * 1e6) // 64 * 64 for m in bench.LARGE_CHROM_MBP]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(offs[-1])
    t0 = time.time()
    ref = bench.gen_reference(torch, n, SEED + 3, rep_frac=0.60, ident_lo=0.80, ident_hi=0.99,
                              n_frac=0.1625)
    sw, ew = bench.pack(torch, ref)
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED + 4)
    hap = ref.clone()
    mutate(torch, hap, 0.005, g)
    rng = np.random.default_rng(SEED + 5)
    for _ in range(30):  # introgressed blocks at ~3% divergence
        c = int(rng.integers(0, len(lens)))
        L = int(rng.uniform(5e6, 20e6))
        s0 = int(offs[c] + rng.integers(0, lens[c] - L))
        mutate(torch, hap, 0.0251, g, s0, s0 + L)
    hd = distinct_codes(torch, hap_codes(torch, hap, offs, k))
    del hap
    keys = sample_keys(torch, hd, 2.1e9, k, SEED + 6)
    del hd
    torch.cuda.empty_cache()
    nN = int((ref == 4).sum())
    note(f"C3: {n/1e9:.2f} Gbp in {len(lens)} chromosomes ({nN/n*100:.1f}% N), "
         f"{keys.numel()/1e9:.2f} G sample keys ({time.time()-t0:.0f} s)")
    t_whole = hipibs.Table.from_device_keys(keys.data_ptr(), keys.numel(), k)
    t_min = hipibs.Table.from_device_keys(keys.data_ptr(), keys.numel(), k, minimizer_len=21,
                                          minimizer_cap=16, load_factor=0.25)
    assert t_whole.info()["n_keys"] == t_min.info()["n_keys"]
    wins, win = windows_of(lens, W, k)
    sc = Scanner(torch, sw, ew, n, win, W, k)
    a = np.stack([sc.scan([t_whole])[0], sc.scan([t_min])[0]])
    b = sc.scan([t_min])
    assert (a >= 0).all()
    assert np.array_equal(a[0], a[1]) and np.array_equal(a[1], b[0])
    note(f"C3: {len(wins)} windows identical across layouts and repeats")
    # per chromosome: hit masks of both tables equal; coverage sum = k x hits;
    # each window's observed count = the hit bits of its positions
    wstart = np.array([w[1] for w in wins])
    wend = np.array([w[2] for w in wins])
    wchrom = np.array([w[0] for w in wins])
    cov_first = cov_last = None
    total_hits = 0
    for c in range(len(lens)):
        h1, npos = hitmask(torch, t_whole, sw, ew, int(offs[c]), lens[c], k)
        h2, _ = hitmask(torch, t_min, sw, ew, int(offs[c]), lens[c], k)
        nwords = (npos + 31) // 32
        assert torch.equal(h1[:nwords], h2[:nwords]), c
        del h2
        cov = coverage(torch, h1, npos, lens[c], k)
        cs = bit_prefix(torch, h1, npos)
        hits = int(cs[-1])
        assert int(cov.sum(dtype=torch.int64)) == k * hits, c
        sel = wchrom == c
        p0 = torch.from_numpy(wstart[sel]).cuda()
        p1 = torch.from_numpy(np.maximum(wstart[sel], wend[sel] - k + 1)).cuda()
        obs = (cs[p1] - cs[p0]).cpu().numpy()
        assert np.array_equal(obs, a[0][sel][:, 1]), c
        total_hits += hits
        if c == 0:
            cov_first = cov[:16_000_000].cpu().numpy()
        if c == len(lens) - 1:
            cov_last = cov[-16_000_000:].cpu().numpy()
        del h1, cov, cs
    note(f"C3: hit masks equal across layouts; coverage sums and window counts agree "
         f"({total_hits/1e9:.3f} G hit positions)")
    t_whole.free()
    t_min.free()
    # oracle on chr1's first and chr7's last 16 Mbp: statistics and coverage
    L = 16_000_000
    text, ks = check_slice(torch, ref, keys, a[0], wins, offs, lens, 0, 0, L, W, k,
                           "C3 chr1 first 16 Mbp")
    want = np.zeros(L, dtype=np.int32)
    oracle().orc_coverage(text, L, k, ks.ctypes.data, ks.size, want.ctypes.data)
    assert np.array_equal(cov_first[:L - k + 1], want[:L - k + 1])
    c = len(lens) - 1
    m = (lens[c] - L) // (W - k)
    a0 = m * (W - k)
    text, ks = check_slice(torch, ref, keys, a[0], wins, offs, lens, c, a0, lens[c] - a0, W, k,
                           "C3 chr7 last 16 Mbp")
    want = np.zeros(len(text), dtype=np.int32)
    oracle().orc_coverage(text, len(text), k, ks.ctypes.data, ks.size, want.ctypes.data)
    tail = min(len(text) - (k - 1), cov_last.size)
    assert np.array_equal(cov_last[-tail:], want[-tail:])
    note("C3: coverage of both slices equals the oracle")
    del keys, ref, sw, ew, sc
    torch.cuda.empty_cache()


# ---- C4: the 48-sample panel, full key counts, HBM-budgeted groups ----------------

C4_N, C4_GROUP_BYTES = 48, 120e9


def c4_plan():
    """48 samples of log-uniform 0.2-2.5e9 keys and SNP rates U[0.1%, 5%],
    in memory groups of at most C4_GROUP_BYTES of whole-k-mer tables at
    load 0.6 (13.3 B/key)."""
    rng = np.random.default_rng(SEED + 40)
    sizes = np.exp(rng.uniform(np.log(0.2e9), np.log(2.5e9), C4_N))
    rates = rng.uniform(0.001, 0.05, C4_N)
    groups, cur, cur_bytes = [], [], 0.0
    for i in range(C4_N):
        tb = np.ceil(np.ceil(sizes[i] / 0.6) / 16) * 128
        if cur and cur_bytes + tb > C4_GROUP_BYTES:
            groups.append(cur)
            cur, cur_bytes = [], 0.0
        cur.append(i)
        cur_bytes += tb
    groups.append(cur)
    return sizes, rates, groups


C4_SIZES, C4_RATES, C4_GROUPS = c4_plan()


@pytest.fixture(scope="module")
def c4_reference(torch):
    k, W, n = 25, 50000, 2_000_000_000
    ref = bench.gen_reference(torch, n, SEED)
    sw, ew = bench.pack(torch, ref)
    wins, win = windows_of([n], W, k)
    env = dict(k=k, W=W, n=n, ref=ref, sw=sw, ew=ew, nw=len(wins),
               sc=Scanner(torch, sw, ew, n, win, W, k), text=text_of(torch, ref[:1_000_000]))
    yield env
    env.clear()
    torch.cuda.empty_cache()


def test_c4_plan_spans_at_least_three_groups():
    assert len(C4_GROUPS) >= 3 and sum(len(g) for g in C4_GROUPS) == C4_N
    assert 40e9 < C4_SIZES.sum() < 50e9


@pytest.mark.timeout(600)
@pytest.mark.parametrize("gi", range(len(C4_GROUPS)))
def test_c4_panel_48_samples_full_key_counts(torch, c4_reference, gi):
    """One memory group of the 48-sample panel at full key counts: the
    panel launch over the group's tables equals one scan per table, and
    each sample's first 1 Mbp equals the oracle."""
    E = c4_reference
    k, W, n, ref, nw = E["k"], E["W"], E["n"], E["ref"], E["nw"]
    offs = np.array([0, n], dtype=np.int64)
    members = C4_GROUPS[gi]
    tabs = []
    for i in members:
        g = torch.Generator(device="cuda")
        g.manual_seed(SEED + 1000 + i)
        hap = ref.clone()
        mutate(torch, hap, float(C4_RATES[i]), g)
        hd = distinct_codes(torch, hap_codes(torch, hap, offs, k))
        del hap
        keys = sample_keys(torch, hd, float(C4_SIZES[i]), k, SEED + 2000 + i)
        del hd
        torch.cuda.empty_cache()
        tabs.append(hipibs.Table.from_device_keys(keys.data_ptr(), keys.numel(), k))
        ks = slice_keys(torch, keys, E["text"], k)
        del keys
        torch.cuda.empty_cache()
        want = oracle_scan_rows([b"s"], [E["text"]], W, k, ks)
        tabs[-1].oracle_rows = [r[3:] for r in want if r[2] - r[1] == W]
    single = E["sc"].scan(tabs)
    nt = len(tabs)
    d_pst = torch.full((nt * nw * 4,), -1, dtype=torch.int32, device="cuda")
    pwb = hipibs._lib.hipibs_panel_work_bytes(nw, nt, 0)
    d_pwork = torch.zeros(max(pwb, 8), dtype=torch.uint8, device="cuda")
    hipibs._check(hipibs._lib.hipibs_panel_scan_device(
        hipibs._table_array(tabs), nt, E["sw"].data_ptr(), E["ew"].data_ptr(), n,
        E["sc"].d_win.data_ptr(), nw, d_pst.data_ptr(), None, d_pwork.data_ptr(), pwb, None),
        "panel")
    torch.cuda.synchronize()
    pst = d_pst.cpu().numpy().reshape(nt, nw, 4)
    assert np.array_equal(pst, single)
    del d_pst, d_pwork
    keys_total = 0
    for j, t in enumerate(tabs):
        got = [tuple(int(x) for x in single[j][r]) for r in range(len(t.oracle_rows))]
        assert got == t.oracle_rows, members[j]
        keys_total += t.info()["n_keys"]
        t.free()
    torch.cuda.empty_cache()
    note(f"C4 group {gi}: {len(members)} samples, {keys_total/1e9:.1f} G distinct keys; "
         f"panel == single scans == oracle")


# ---- C5: served queries, 96 samples, 32 resident ---------------------------------

@pytest.mark.timeout(900)
def test_c5_serving_96_samples_32_resident(torch):
    """96 samples of U[0.3, 1.5]e9 / 10 keys (a tenth of the stated counts:
    the host holds every sample's keys) registered with a serving engine
    whose HBM budget holds 32 mean-sized tables; ~240 requests of (sample,
    chromosome, range 1 Mbp-chromosome log-uniform, W in 10/50/100 kbp)
    arriving at ~100 per second. Every request equals the same range
    scanned alone against its sample's table, and the small ones equal the
    oracle."""
    k = 25
    n = 2_000_000_000
    nchrom = 8
    lens = [n // nchrom] * nchrom
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ref = bench.gen_reference(torch, n, SEED)
    sw, ew = bench.pack(torch, ref)
    h_sw = sw.cpu().numpy().view(np.uint64)
    h_ew = ew.cpu().numpy().view(np.uint64)
    rng = np.random.default_rng(SEED + 50)
    S = 96
    scale = 0.1
    sizes = rng.uniform(0.3e9, 1.5e9, S) * scale
    # four haplotypes; a sample keeps its own hash-chosen share of one
    haps = []
    for h, rate in enumerate((0.001, 0.01, 0.03, 0.05)):
        g = torch.Generator(device="cuda")
        g.manual_seed(SEED + 60 + h)
        hap = ref.clone()
        mutate(torch, hap, rate, g)
        hd = distinct_codes(torch, hap_codes(torch, hap, np.array([0, n]), k))
        del hap
        haps.append([x[keep_mask(torch, x, 0.25, h + 1)] for x in hd])  # a quarter each
        del hd
    host_keys = []
    for i in range(S):
        kk = sample_keys(torch, haps[i % 4], float(sizes[i]), k, SEED + 3000 + i)
        host_keys.append(kk.cpu().numpy().view(np.uint64))
        del kk
    del haps
    torch.cuda.empty_cache()
    mean_tbl = float(np.mean([np.ceil(np.ceil(x.size / 0.6) / 16) * 128 for x in host_keys]))
    budget = int(32 * mean_tbl)
    note(f"C5: {S} samples, {sum(x.size for x in host_keys)/1e9:.1f} G host keys, budget "
         f"{budget/1e9:.1f} GB (32 mean tables)")
    reqs = []
    for r in range(240):
        sid = int(rng.integers(0, S))
        c = int(rng.integers(0, nchrom))
        ln = int(min(lens[c], np.exp(rng.uniform(np.log(1e6), np.log(lens[c])))))
        if r % 10 == 0:
            ln = int(rng.integers(1_000_000, 1_200_000))  # small: checked by the oracle
        base = int(offs[c] + rng.integers(0, lens[c] - ln + 1))
        Wr = int(rng.choice([10_000, 50_000, 100_000]))
        reqs.append((r, sid, base, ln, Wr))
    eng = hipibs.ServeEngine(k, budget, max_inflight=3, max_windows_per_launch=65536)
    eng.set_reference(h_sw, h_ew, n)
    for sid in range(S):
        eng.add_sample(sid, host_keys[sid])
    served = {}
    t0 = time.time()
    arrivals = np.cumsum(rng.exponential(1 / 100.0, len(reqs)))
    for (rid, sid, base, ln, Wr), at in zip(reqs, arrivals):
        while time.time() - t0 < at:
            for done in eng.poll(timeout_ms=1):
                served[done] = eng.fetch(done)[0]
        eng.submit(rid, sid, base, ln, Wr)
    while len(served) < len(reqs):
        for done in eng.poll(timeout_ms=2000):
            served[done] = eng.fetch(done)[0]
        assert time.time() - t0 < 600, "serving did not finish"
    cnt = eng.counters()
    eng.free()
    note(f"C5: {len(reqs)} requests served in {time.time()-t0:.1f} s; {cnt['launches']} "
         f"launches, {cnt['admissions']} admissions, {cnt['evictions']} evictions")
    assert cnt["evictions"] > 0 and cnt["launches"] < len(reqs)
    # alone: each request's range scanned against its sample's own table
    by_sample = {}
    for r in reqs:
        by_sample.setdefault(r[1], []).append(r)
    n_oracle = 0
    for sid, rs in sorted(by_sample.items()):
        t = hipibs.Table.from_keys(host_keys[sid], k)
        for rid, _, base, ln, Wr in rs:
            wins = hipibs.chunk_sequences([ln], Wr, k)
            win = np.zeros(len(wins), dtype=hipibs.WINDOW_DTYPE)
            win["base"] = np.array([base + w[1] for w in wins], dtype=np.uint64)
            win["n_bases"] = np.array([w[2] - w[1] for w in wins], dtype=np.uint32)
            alone = Scanner(torch, sw, ew, n, win, Wr, k).scan([t])[0]
            assert np.array_equal(served[rid], alone), rid
            if ln <= 1_200_000:
                text = text_of(torch, ref[base:base + ln])
                d_keys = torch.from_numpy(host_keys[sid].view(np.int64)).cuda()
                ks = slice_keys(torch, d_keys, text, k)
                del d_keys
                want = oracle_scan_rows([b"s"], [text], Wr, k, ks)
                assert [tuple(int(x) for x in row) for row in alone] == [w[3:] for w in want]
                n_oracle += 1
        t.free()
    note(f"C5: every request equals its scan alone; {n_oracle} equal the oracle")
    assert n_oracle >= 20
    del ref, sw, ew
    torch.cuda.empty_cache()
