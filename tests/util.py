"""Test helpers: the oracle (oracle/liboracle.so) through ctypes, and
small conversions shared by the CPU and GPU tests."""
import ctypes as C
import hashlib
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libref.so")

_orc = None


def oracle():
    global _orc
    if _orc is None:
        lib = C.CDLL(ORACLE_PATH)
        i64, u32, vp = C.c_int64, C.c_uint32, C.c_void_p
        lib.orc_encode_positions.argtypes = [C.c_char_p, i64, u32, vp, vp]
        lib.orc_encode_positions.restype = i64
        lib.orc_windows.argtypes = [i64, i64, i64, vp, vp, i64]
        lib.orc_windows.restype = i64
        lib.orc_sort_keys.argtypes = [vp, i64]
        lib.orc_window_stats.argtypes = [C.c_char_p, i64, u32, vp, i64, vp]
        lib.orc_stats_from_flags.argtypes = [vp, vp, i64, u32, vp]
        lib.orc_coverage.argtypes = [C.c_char_p, i64, u32, vp, i64, vp]
        lib.orc_intersect.argtypes = [vp, i64, vp, i64]
        lib.orc_intersect.restype = i64
        lib.orc_kmc_read.argtypes = [C.c_char_p, vp, i64, C.POINTER(u32)]
        lib.orc_kmc_read.restype = i64
        lib.orc_scan_contig.argtypes = [C.c_char_p, i64, i64, u32, vp, i64, vp, i64]
        lib.orc_scan_contig.restype = i64
        lib.orc_count_kmers.argtypes = [C.c_char_p, i64, u32, vp, vp, i64, C.POINTER(i64)]
        lib.orc_count_kmers.restype = i64
        lib.orc_encode_positions_wide.argtypes = [C.c_char_p, i64, u32, vp, vp]
        lib.orc_encode_positions_wide.restype = i64
        lib.orc_scan_contig_wide.argtypes = [C.c_char_p, i64, i64, u32, vp, i64, vp, i64]
        lib.orc_scan_contig_wide.restype = i64
        lib.orc_coverage_wide.argtypes = [C.c_char_p, i64, u32, vp, i64, vp]
        _orc = lib
    return _orc


def sorted_keys(keys):
    return np.sort(np.ascontiguousarray(keys, dtype=np.uint64))


def oracle_scan_rows(ids, seqs, W, k, keys_sorted):
    """TSV rows of every contig through the oracle."""
    lib = oracle()
    rows = []
    for cid, s in zip(ids, seqs):
        cap = len(s) // max(1, W - k) + 2
        out = np.zeros((cap, 6), dtype=np.int64)
        n = lib.orc_scan_contig(s, len(s), W, k, keys_sorted.ctypes.data, keys_sorted.size,
                                out.ctypes.data, cap)
        assert 0 <= n <= cap
        for r in out[:n]:
            rows.append((cid,) + tuple(int(x) for x in r))
    return rows


def oracle_coverage_text(ids, seqs, k, keys_sorted):
    """fastibsmapper file content through the oracle."""
    lib = oracle()
    parts = []
    for cid, s in zip(ids, seqs):
        cov = np.zeros(len(s), dtype=np.int32)
        lib.orc_coverage(s, len(s), k, keys_sorted.ctypes.data, keys_sorted.size, cov.ctypes.data)
        parts.append(cid + b"\n" + coverage_line(cov) + b"\n")
    return b"".join(parts)


def sorted_keys_wide(keys):
    """(n, 2) high/low codes sorted as 128-bit integers, contiguous."""
    return np.ascontiguousarray(np.unique(np.asarray(keys, dtype=np.uint64), axis=0))


def oracle_scan_rows_wide(ids, seqs, W, k, keys2):
    """TSV rows for k = 33..64 through the oracle (keys2 from sorted_keys_wide)."""
    lib = oracle()
    rows = []
    for cid, s in zip(ids, seqs):
        cap = len(s) // max(1, W - k) + 2
        out = np.zeros((cap, 6), dtype=np.int64)
        n = lib.orc_scan_contig_wide(s, len(s), W, k, keys2.ctypes.data, keys2.shape[0],
                                     out.ctypes.data, cap)
        assert 0 <= n <= cap
        for r in out[:n]:
            rows.append((cid,) + tuple(int(x) for x in r))
    return rows


def oracle_coverage_text_wide(ids, seqs, k, keys2):
    lib = oracle()
    parts = []
    for cid, s in zip(ids, seqs):
        cov = np.zeros(len(s), dtype=np.int32)
        lib.orc_coverage_wide(s, len(s), k, keys2.ctypes.data, keys2.shape[0], cov.ctypes.data)
        parts.append(cid + b"\n" + coverage_line(cov) + b"\n")
    return b"".join(parts)


def coverage_line(cov):
    """Comma-joined decimal coverage (serializeCoverage,
    src/KmerDatabase.hpp:756-772)."""
    if len(cov) == 0:
        return b""
    return ",".join(map(str, np.asarray(cov).tolist())).encode()


def sha256(b):
    return hashlib.sha256(b).hexdigest()


def oracle_count(text, k):
    """(keys, counts, k-mers) of text by the oracle's counter (ascending)."""
    lib = oracle()
    cap = max(1, len(text) - k + 1)
    keys = np.zeros(cap, dtype=np.uint64)
    counts = np.zeros(cap, dtype=np.uint32)
    nk = C.c_int64(0)
    d = lib.orc_count_kmers(text, len(text), k, keys.ctypes.data, counts.ctypes.data, cap,
                            C.byref(nk))
    assert 0 <= d <= cap
    return keys[:d], counts[:d], nk.value
