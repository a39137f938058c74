"""hipibs_scan_windows_device_noalloc (the scan whose scratch is a
library-held device array) against hipibs_scan_windows_device (scratch
from the caller), through the C ABI: two calls in a row, the second with a
smaller input, several tables and segment splits; sizes beyond the
array's limits are refused."""
import ctypes as C

import numpy as np
import pytest

import corpus
from util import oracle_scan_rows, sorted_keys

import hipibs

pytestmark = pytest.mark.gpu
D = hipibs.DeviceBuffer


def _windows(lengths, W, k):
    wins = hipibs.chunk_sequences(lengths, W, k)
    offs = np.concatenate([[0], np.cumsum(lengths)])
    win = np.zeros(len(wins), dtype=hipibs.WINDOW_DTYPE)
    for i, (c, s, e) in enumerate(wins):
        win[i] = (offs[c] + s, e - s, 0)
    return wins, win


def _scan_both(tabs, seqs, W, k, segs):
    """Statistics of both entry points (n_tables x n_windows x 4) on the same input."""
    sw, ew, _ = hipibs.pack_sequences(seqs)
    wins, win = _windows([len(s) for s in seqs], W, k)
    n_bases = sum(len(s) for s in seqs)
    d_sw, d_ew, d_win = D.from_array(sw), D.from_array(ew), D.from_array(win)
    nt = len(tabs)
    arr = hipibs._table_array(tabs)
    out = []
    # the entry point with caller scratch
    res = C.c_uint32(0)
    wb = hipibs._lib.hipibs_scan_work_bytes(len(wins), nt, W, k, segs, C.byref(res))
    d_work = D.zeros(max(wb, 8))
    d_a = D.zeros(len(wins) * nt * 16)
    hipibs._check(hipibs._lib.hipibs_scan_windows_device(
        arr, nt, d_sw.ptr, d_ew.ptr, n_bases, d_win.ptr, len(wins), W, d_a.ptr, segs,
        d_work.ptr, wb, None), "scan")
    # the one without allocation; its output buffer starts with garbage
    d_b = D.from_array(np.full(len(wins) * nt * 4, -7, dtype=np.int32))
    hipibs._check(hipibs._lib.hipibs_scan_windows_device_noalloc(
        arr, nt, d_sw.ptr, d_ew.ptr, n_bases, d_win.ptr, len(wins), W, d_b.ptr, segs, None),
        "scan without allocation")
    hipibs.synchronize()
    a = d_a.to_array(np.int32).reshape(nt, len(wins), 4)
    b = d_b.to_array(np.int32).reshape(nt, len(wins), 4)
    for d in (d_sw, d_ew, d_win, d_work, d_a, d_b):
        d.free()
    return wins, res.value, a, b


@pytest.mark.parametrize("segs", [0, 1, 5])
def test_noalloc_scan_equals_scan_with_caller_scratch(small_corpus, segs):
    ref, samples = small_corpus
    seq = corpus.to_text(ref)
    tabs = [hipibs.Table.from_keys(s[0], 25) for s in samples]
    tabs.append(hipibs.Table.from_keys(samples[1][0], 25, minimizer_len=21, minimizer_cap=8,
                                       load_factor=0.3))
    # first call: the whole 1 Mbp reference in 5 kbp windows against the
    # three whole-k-mer tables; second call: a smaller input (two pieces,
    # 1 kbp windows) against two tables -- so the second reuses scratch the
    # first filled with more summaries and views
    calls = [(tabs[:3], [seq], 5000), (tabs[1:3], [seq[:70_000], seq[500_000:533_333]], 1000)]
    for tset, seqs, W in calls:
        wins, resolved, a, b = _scan_both(tset, seqs, W, 25, segs)
        if segs:
            assert resolved == segs
        np.testing.assert_array_equal(a, b)
    # a minimizer table through the same scratch, checked against the oracle
    seqs = [seq[:200_000]]
    wins, _, a, b = _scan_both([tabs[3]], seqs, 10_000, 25, segs)
    np.testing.assert_array_equal(a, b)
    want = oracle_scan_rows([b"c"], seqs, 10_000, 25, sorted_keys(samples[1][0]))
    assert [tuple(int(x) for x in r) for r in b[0]] == [w[3:] for w in want]
    for t in tabs:
        t.free()


def test_noalloc_scan_refuses_what_exceeds_its_scratch(small_corpus):
    ref, samples = small_corpus
    t = hipibs.Table.from_keys(samples[0][0], 25)
    cap = hipibs._lib.hipibs_scan_noalloc_max_work_bytes()
    assert cap >= 1 << 20
    arr = hipibs._table_array([t] * 65)
    d = D.zeros(64)
    rc = hipibs._lib.hipibs_scan_windows_device_noalloc(arr, 65, d.ptr, d.ptr, 64, d.ptr, 1, 64,
                                                       d.ptr, 1, None)
    assert rc == -1 and b"at most 64 tables" in hipibs._lib.hipibs_last_error()
    # windows x segments whose summaries need more than the scratch holds
    n_win = cap // (24 * 64) + 1
    rc = hipibs._lib.hipibs_scan_windows_device_noalloc(hipibs._table_array([t]), 1, d.ptr, d.ptr,
                                                       64, d.ptr, n_win, 50000, d.ptr, 64, None)
    assert rc == -1 and b"scratch" in hipibs._lib.hipibs_last_error()
    d.free()
    t.free()
