"""Python bindings of libhipibs (the MI355X k-mer IBS scan) over its C ABI.

The library is the product; this module is a thin ctypes layer used by the
tests, the benchmark and scripting users. It never falls back to a CPU
implementation: if libhipibs.so is missing, import fails loudly.

Host-side helpers mirror the reference's host code so results line up with
its tools: read_sequence_file (src/Utils.hpp:170-255), chunk_sequences
(src/Utils.hpp:296-327), output naming and the TSV writer
(src/FastIBS.cpp:14-25, src/KmerDatabase.hpp:935-948).
"""
from __future__ import annotations

import ctypes as C
import gzip
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libhipibs.so")
PACK_PAD_WORDS = 8

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libhipibs.so not found at {LIB_PATH}: build it first "
        "(python -c 'import __graft_entry__ as g; g.build()' or make -C gpu-ibs_amd)")

_lib = C.CDLL(LIB_PATH)

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
vp = C.c_void_p


class TableOpts(C.Structure):
    _fields_ = [("load_factor", C.c_double), ("bucket_slots", C.c_uint32), ("device", C.c_int),
                ("hbm_budget", C.c_uint64), ("load_threads", C.c_uint32),
                ("minimizer_len", C.c_uint32), ("minimizer_cap", C.c_uint32)]


class Window(C.Structure):
    _fields_ = [("base", C.c_uint64), ("n_bases", C.c_uint32), ("reserved", C.c_uint32)]


class WindowStats(C.Structure):
    _fields_ = [("total_kmers", C.c_int32), ("observed_kmers", C.c_int32),
                ("variations", C.c_int32), ("kmer_distance", C.c_int32)]


class KmcInfo(C.Structure):
    _fields_ = [("kmc_version", C.c_uint32), ("kmer_length", C.c_uint32), ("mode", C.c_uint32),
                ("counter_size", C.c_uint32), ("lut_prefix_length", C.c_uint32),
                ("signature_len", C.c_uint32), ("min_count", C.c_uint32),
                ("max_count", C.c_uint32), ("total_kmers", C.c_uint64)]


class ScanConfig(C.Structure):
    _fields_ = [("segs_per_window", C.c_uint32), ("waves_per_cu", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class ServeCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "launches", "windows",
                                         "admissions", "evictions", "resident_bytes",
                                         "resident_tables", "build_us", "evict_us",
                                         "host_bytes")]


class CountStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("reads", "kmers", "distinct", "table_bytes",
                                         "launches")] + [("kernel_ms", C.c_double)] + \
        [(n, C.c_uint64) for n in ("flushes", "spilled")] + [("stage_ms", C.c_double * 4)]


class CountFilter(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("kept", "below", "above")]


class HostSlot(C.Structure):
    _fields_ = [("seq_words", u64p), ("exc_words", u64p), ("windows", C.POINTER(Window)),
                ("stats", C.POINTER(WindowStats)), ("seq_cap", C.c_uint64),
                ("exc_cap", C.c_uint64), ("window_cap", C.c_uint32), ("n_tables", C.c_uint32)]


WINDOW_DTYPE = np.dtype([("base", "<u8"), ("n_bases", "<u4"), ("reserved", "<u4")])
STATS_DTYPE = np.dtype([("total_kmers", "<i4"), ("observed_kmers", "<i4"),
                        ("variations", "<i4"), ("kmer_distance", "<i4")])

# (name, restype, argtypes): every symbol include/hipibs.h declares.
_SIGS = [
    ("hipibs_last_error", C.c_char_p, []),
    ("hipibs_device_count", C.c_int, [C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    ("hipibs_device_mem", C.c_int, [C.c_int, u64p, u64p]),
    ("hipibs_kmc_header", C.c_int, [C.c_char_p, C.POINTER(KmcInfo)]),
    ("hipibs_kmc_read_keys", C.c_int, [C.c_char_p, u64p, C.c_uint64, u64p, C.c_uint32]),
    ("hipibs_kmc_write", C.c_int, [C.c_char_p, u64p, u32p, C.c_uint64, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("hipibs_pack_bases", C.c_int, [C.c_char_p, C.c_uint64, C.c_uint64, u64p, u64p]),
    ("hipibs_table_build_kmc", C.c_int, [C.c_char_p, C.POINTER(TableOpts), C.POINTER(vp)]),
    ("hipibs_table_build_keys", C.c_int, [u64p, C.c_uint64, C.c_uint32, C.POINTER(TableOpts),
                                          C.POINTER(vp)]),
    ("hipibs_table_build_keys_device", C.c_int, [vp, C.c_uint64, C.c_uint32,
                                                 C.POINTER(TableOpts), C.POINTER(vp), vp]),
    ("hipibs_table_free", C.c_int, [vp]),
    ("hipibs_table_minimizer_len", C.c_int, [vp, u32p]),
    ("hipibs_table_info", C.c_int, [vp, u64p, u64p, u64p, u32p, u32p, C.POINTER(C.c_int)]),
    ("hipibs_table_probe", C.c_int, [vp, vp, vp, C.c_uint64, vp, vp]),
    ("hipibs_table_probe_count", C.c_int, [vp, u64p, C.c_uint64, u64p]),
    ("hipibs_table_probe_count_device", C.c_int, [vp, vp, C.c_uint64, vp, vp]),
    ("hipibs_table_probe_count_kmc", C.c_int, [vp, C.c_char_p, C.c_uint, u64p]),
    ("hipibs_counter_create", C.c_int, [C.c_uint32, C.c_int, C.c_uint64, C.c_uint64,
                                        C.POINTER(vp)]),
    ("hipibs_counter_add_file", C.c_int, [vp, C.c_char_p]),
    ("hipibs_counter_add_text", C.c_int, [vp, C.c_char_p, C.c_uint64]),
    ("hipibs_counter_add_text_device", C.c_int, [vp, vp, C.c_uint64]),
    ("hipibs_counter_stats", C.c_int, [vp, C.POINTER(CountStats)]),
    ("hipibs_counter_keys", C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, u64p, u32p,
                                      C.c_uint64, C.POINTER(CountFilter)]),
    ("hipibs_counter_build_table", C.c_int, [vp, C.c_uint32, C.c_uint32, C.POINTER(TableOpts),
                                             C.POINTER(vp), C.POINTER(CountFilter)]),
    ("hipibs_counter_write_kmc", C.c_int, [vp, C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.POINTER(CountFilter)]),
    ("hipibs_counter_free", None, [vp]),
    ("hipibs_counter_release_cache", None, []),
    ("hipibs_coverage_text_device", C.c_int, [vp, C.c_uint64, vp, C.c_uint64, u64p, vp]),
    ("hipibs_coverage_text_stream", C.c_int, [vp, C.c_uint64, vp, vp, vp]),
    ("hipibs_scan_windows_device", C.c_int, [C.POINTER(vp), C.c_int, vp, vp, C.c_uint64, vp,
                                             C.c_uint32, C.c_uint32, vp, C.c_uint32, vp,
                                             C.c_uint64, vp]),
    ("hipibs_scan_work_bytes", C.c_uint64, [C.c_uint32, C.c_int, C.c_uint32, C.c_uint32,
                                            C.c_uint32, u32p]),
    ("hipibs_scan_windows_device_noalloc", C.c_int, [C.POINTER(vp), C.c_int, vp, vp,
                                                     C.c_uint64, vp, C.c_uint32, C.c_uint32, vp,
                                                     C.c_uint32, vp]),
    ("hipibs_scan_noalloc_max_work_bytes", C.c_uint64, []),
    ("hipibs_hitmask_device", C.c_int, [vp, vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, vp, vp]),
    ("hipibs_coverage_device", C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint32, vp, vp]),
    ("hipibs_encode_device", C.c_int, [vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32,
                                       vp, vp, vp]),
    ("hipibs_set_profiling", C.c_int, [C.c_int]),
    ("hipibs_last_kernel_ms", C.c_double, []),
    ("hipibs_scan_create", C.c_int, [C.POINTER(vp), C.c_int, C.c_uint64, C.c_uint32, C.c_int,
                                     C.POINTER(vp)]),
    ("hipibs_scan_slot", C.c_int, [vp, C.c_int, C.POINTER(HostSlot)]),
    ("hipibs_scan_submit", C.c_int, [vp, C.c_int, C.c_uint32, C.c_uint64, C.c_uint64]),
    ("hipibs_scan_wait", C.c_int, [vp, C.c_int]),
    ("hipibs_scan_free", C.c_int, [vp]),
    ("hipibs_panel_work_bytes", C.c_uint64, [C.c_uint32, C.c_int, C.c_uint32]),
    ("hipibs_panel_scan_device", C.c_int, [C.POINTER(vp), C.c_int, vp, vp, C.c_uint64, vp,
                                           C.c_uint32, vp, C.POINTER(ScanConfig), vp, C.c_uint64,
                                           vp]),
    ("hipibs_panel_autotune", C.c_int, [C.POINTER(vp), C.c_int, vp, vp, C.c_uint64, vp,
                                        C.c_uint32, C.c_uint32, vp, vp, C.c_uint64,
                                        C.POINTER(ScanConfig), C.POINTER(C.c_double)]),
    ("hipibs_panel_graph_create", C.c_int, [C.POINTER(vp), C.c_int, vp, vp, C.c_uint64, vp,
                                            C.c_uint32, vp, C.POINTER(ScanConfig), C.c_uint32,
                                            vp, C.c_uint64, C.POINTER(vp)]),
    ("hipibs_graph_launch", C.c_int, [vp, vp]),
    ("hipibs_graph_free", C.c_int, [vp]),
    ("hipibs_serve_create", C.c_int, [C.c_int, C.c_uint32, C.c_uint64, C.POINTER(TableOpts),
                                      C.c_int, C.c_uint32, C.POINTER(vp)]),
    ("hipibs_serve_set_reference", C.c_int, [vp, u64p, u64p, C.c_uint64]),
    ("hipibs_serve_add_sample", C.c_int, [vp, C.c_uint32, u64p, C.c_uint64, C.c_int]),
    ("hipibs_serve_add_sample_device", C.c_int, [vp, C.c_uint32, vp, C.c_uint64, C.c_int]),
    ("hipibs_serve_add_sample_kmc", C.c_int, [vp, C.c_uint32, C.c_char_p, C.c_int]),
    ("hipibs_serve_submit", C.c_int, [vp, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64,
                                      C.c_uint32]),
    ("hipibs_serve_poll", C.c_int, [vp, u64p, C.c_uint32, C.c_int, u32p]),
    ("hipibs_serve_fetch", C.c_int, [vp, C.c_uint64, vp, C.c_uint32, u32p,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("hipibs_serve_get_counters", C.c_int, [vp, C.POINTER(ServeCounters)]),
    ("hipibs_serve_free", C.c_int, [vp]),
    ("hipibs_dev_alloc", C.c_int, [C.c_int, C.c_uint64, C.POINTER(vp)]),
    ("hipibs_dev_free", C.c_int, [vp]),
    ("hipibs_memcpy", C.c_int, [vp, vp, C.c_uint64, C.c_int]),
    ("hipibs_dev_memset", C.c_int, [vp, C.c_int, C.c_uint64]),
    ("hipibs_dev_synchronize", C.c_int, [C.c_int]),
]
for _name, _res, _args in _SIGS:
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = [s[0] for s in _SIGS]


class HipibsError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = _lib.hipibs_last_error().decode(errors="replace")
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def _check(rc: int, where: str) -> None:
    if rc != 0:
        raise HipibsError(rc, where)


def _ptr(a: np.ndarray, t=u64p):
    return a.ctypes.data_as(t)


# ---- devices ----------------------------------------------------------------

def device_count() -> Tuple[int, str]:
    n = C.c_int(0)
    why = C.create_string_buffer(512)
    _lib.hipibs_device_count(C.byref(n), why, 512)
    return n.value, why.value.decode(errors="replace")


def device_mem(device: int = 0) -> Tuple[int, int]:
    f, t = C.c_uint64(0), C.c_uint64(0)
    _check(_lib.hipibs_device_mem(device, C.byref(f), C.byref(t)), "hipibs_device_mem")
    return f.value, t.value


# ---- KMC databases ----------------------------------------------------------

def kmc_header(prefix: str) -> KmcInfo:
    info = KmcInfo()
    _check(_lib.hipibs_kmc_header(prefix.encode(), C.byref(info)), "hipibs_kmc_header")
    return info


def kmc_read_keys(prefix: str, threads: int = 0) -> np.ndarray:
    """Keys in on-disk order: uint64 (k <= 32) or (n, 2) high/low words."""
    info = kmc_header(prefix)
    wide = info.kmer_length > 32
    keys = np.empty((info.total_kmers, 2) if wide else info.total_kmers, dtype=np.uint64)
    n = C.c_uint64(0)
    _check(_lib.hipibs_kmc_read_keys(prefix.encode(), _ptr(keys), info.total_kmers, C.byref(n),
                                     threads), "hipibs_kmc_read_keys")
    return keys


def _key_count(keys: np.ndarray, k: int) -> int:
    """Number of keys in a uint64 array (k <= 32) or an (n, 2) one."""
    if k > 32:
        if keys.ndim != 2 or keys.shape[1] != 2:
            raise ValueError("keys of k > 32 are an (n, 2) array of (high, low) words")
        return keys.shape[0]
    return keys.size


def kmc_write(prefix: str, keys: np.ndarray, k: int, counts: Optional[np.ndarray] = None,
              p: Optional[int] = None, sig_len: int = 9, counter_size: int = 1,
              version: int = 0x200, min_count: int = 1, max_count: int = 255) -> None:
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    n = _key_count(keys, k)
    if p is None:
        p = _default_lut_prefix(k)
    cp = None
    if counts is not None:
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        cp = _ptr(counts, u32p)
    _check(_lib.hipibs_kmc_write(prefix.encode(), _ptr(keys), cp, n, k, p, sig_len,
                                 counter_size, version, min_count, max_count), "hipibs_kmc_write")


def _default_lut_prefix(k: int) -> int:
    """Largest p <= min(k, 12) with (k - p) % 4 == 0 and at least one suffix
    byte when k allows it (KMC's own choice depends on the database size)."""
    for p in range(min(k, 12), -1, -1):
        if (k - p) % 4 == 0 and (k - p) >= 4:
            if p <= 9:
                return p
    for p in range(min(k, 12), -1, -1):
        if (k - p) % 4 == 0:
            return p
    raise ValueError(f"no LUT prefix length for k={k}")


# ---- packing ------------------------------------------------------------------

def pack_words(n_bases: int) -> Tuple[int, int]:
    """Word counts (sequence, exception) of a packed stream of n_bases,
    padding included."""
    return (n_bases + 31) // 32 + PACK_PAD_WORDS, (n_bases + 63) // 64 + PACK_PAD_WORDS


def pack_sequences(seqs: Sequence[bytes]) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Concatenates the contigs into one packed stream. Returns (seq_words,
    exc_words, contig_offsets) with offsets[i] = first base of contig i."""
    offsets = np.zeros(len(seqs) + 1, dtype=np.uint64)
    for i, s in enumerate(seqs):
        offsets[i + 1] = offsets[i] + len(s)
    total = int(offsets[-1])
    ns, ne = pack_words(total)
    sw = np.zeros(ns, dtype=np.uint64)
    ew = np.zeros(ne, dtype=np.uint64)
    for i, s in enumerate(seqs):
        if len(s):
            _check(_lib.hipibs_pack_bases(bytes(s), len(s), int(offsets[i]), _ptr(sw), _ptr(ew)),
                   "hipibs_pack_bases")
    return sw, ew, offsets


# ---- host input handling mirroring the reference ------------------------------

@dataclass
class ParseResult:
    ids: List[bytes]
    sequences: List[bytes]


def read_sequence_file(path: str) -> ParseResult:
    """readSequenceFile (src/Utils.hpp:170-255): '>' or '@' decides FASTA or
    FASTQ; lines split on '\\n' only (a trailing '\\r' stays a base); empty
    lines skipped; records with an empty sequence dropped."""
    if ".gz" in path:
        with gzip.open(path, "rb") as f:
            data = f.read()
    else:
        with open(path, "rb") as f:
            data = f.read()
    if not data:
        raise ValueError("Empty file")
    if data[:1] == b">":
        fastq = False
    elif data[:1] == b"@":
        fastq = True
    else:
        raise ValueError("Unknown file type")
    id_char = b"@" if fastq else b">"
    ids: List[bytes] = []
    seqs: List[bytes] = []
    cur_id = b""
    cur: List[bytes] = []
    quality = False
    for line in data.split(b"\n"):
        if not line:
            continue
        if fastq and line[:1] == b"+":
            quality = True
            continue
        if quality:
            quality = False
            continue
        if line[:1] == id_char:
            if cur:
                s = b"".join(cur)
                if s:
                    ids.append(cur_id)
                    seqs.append(s)
                cur = []
            cur_id = line[1:]
        else:
            cur.append(line)
    if cur:
        s = b"".join(cur)
        if s:
            ids.append(cur_id)
            seqs.append(s)
    return ParseResult(ids, seqs)


def chunk_sequences(lengths: Sequence[int], W: int, k: int) -> List[Tuple[int, int, int]]:
    """chunkSequences (src/Utils.hpp:296-327): (contig, start, end) per
    window, start = 0, W-k, 2(W-k), ... while start < len."""
    if W <= k:
        raise ValueError(f"window size {W} must be larger than k ({k})")
    out = []
    for c, n in enumerate(lengths):
        j = 0
        while j < n:
            out.append((c, j, min(j + W, n)))
            j += W - k
    return out


def format_tsv(rows: Sequence[Tuple[bytes, int, int, int, int, int, int]]) -> bytes:
    """KmerDatabase::writeStats (src/KmerDatabase.hpp:935-948)."""
    parts = [b"seqname\tstart\tend\ttotal_kmers\tobserved_kmers\tvariations\tkmer_distance\n"]
    for r in rows:
        parts.append(b"%s\t%d\t%d\t%d\t%d\t%d\t%d\n" % tuple(r))
    return b"".join(parts)


# ---- device memory ---------------------------------------------------------------

class DeviceBuffer:
    """A device allocation owned by libhipibs (no other HIP runtime needed)."""

    def __init__(self, nbytes: int, device: int = 0):
        p = vp()
        _check(_lib.hipibs_dev_alloc(device, int(nbytes), C.byref(p)), "hipibs_dev_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray, device: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 1), device)
        _check(_lib.hipibs_memcpy(vp(b.ptr), a.ctypes.data_as(vp), a.nbytes, 1), "hipibs_memcpy")
        return b

    @classmethod
    def zeros(cls, nbytes: int, device: int = 0) -> "DeviceBuffer":
        b = cls(max(int(nbytes), 1), device)
        _check(_lib.hipibs_dev_memset(vp(b.ptr), 0, b.nbytes), "hipibs_dev_memset")
        return b

    def to_array(self, dtype, count: Optional[int] = None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(n, dtype=dt)
        _check(_lib.hipibs_memcpy(out.ctypes.data_as(vp), vp(self.ptr), out.nbytes, 2),
               "hipibs_memcpy")
        return out

    def free(self) -> None:
        if self.ptr:
            _lib.hipibs_dev_free(vp(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(device: int = 0) -> None:
    _check(_lib.hipibs_dev_synchronize(device), "hipibs_dev_synchronize")


# ---- tables -------------------------------------------------------------------

def _opts(load_factor=0.6, bucket_slots=0, device=0, hbm_budget=0, load_threads=0,
          minimizer_len=0, minimizer_cap=0):
    return TableOpts(load_factor, bucket_slots, device, hbm_budget, load_threads, minimizer_len,
                     minimizer_cap)


class Table:
    """A device-resident k-mer table (hipibs_table)."""

    def __init__(self, handle: int):
        self.handle = C.c_void_p(handle)

    @classmethod
    def from_keys(cls, keys: np.ndarray, k: int, **kw) -> "Table":
        """keys: uint64 codes (k <= 32) or an (n, 2) array of (high, low)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        h = vp()
        o = _opts(**kw)
        _check(_lib.hipibs_table_build_keys(_ptr(keys), _key_count(keys, k), k, C.byref(o),
                                            C.byref(h)), "hipibs_table_build_keys")
        return cls(h.value)

    @classmethod
    def from_kmc(cls, prefix: str, **kw) -> "Table":
        h = vp()
        o = _opts(**kw)
        _check(_lib.hipibs_table_build_kmc(prefix.encode(), C.byref(o), C.byref(h)),
               "hipibs_table_build_kmc")
        return cls(h.value)

    @classmethod
    def from_device_keys(cls, d_keys: int, n: int, k: int, stream: int = 0, **kw) -> "Table":
        h = vp()
        o = _opts(**kw)
        _check(_lib.hipibs_table_build_keys_device(vp(d_keys), n, k, C.byref(o), C.byref(h),
                                                   vp(stream)), "hipibs_table_build_keys_device")
        return cls(h.value)

    def info(self) -> dict:
        nk, nb, nbk = C.c_uint64(), C.c_uint64(), C.c_uint64()
        k, bs, dev = C.c_uint32(), C.c_uint32(), C.c_int()
        _check(_lib.hipibs_table_info(self.handle, C.byref(nk), C.byref(nb), C.byref(nbk),
                                      C.byref(k), C.byref(bs), C.byref(dev)), "hipibs_table_info")
        d = dict(n_keys=nk.value, bytes=nb.value, n_buckets=nbk.value, k=k.value,
                 bucket_slots=bs.value, device=dev.value)
        ml = C.c_uint32(0)
        _check(_lib.hipibs_table_minimizer_len(self.handle, C.byref(ml)),
               "hipibs_table_minimizer_len")
        d["minimizer_len"] = ml.value
        return d

    def probe_count(self, keys: np.ndarray) -> int:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        n = _key_count(keys, self.info()["k"])
        out = C.c_uint64(0)
        _check(_lib.hipibs_table_probe_count(self.handle, _ptr(keys), n, C.byref(out)),
               "hipibs_table_probe_count")
        return out.value

    def probe_count_kmc(self, prefix: str, threads: int = 0) -> int:
        """Keys of a KMC database present in the table (records decoded on
        the device)."""
        out = C.c_uint64(0)
        _check(_lib.hipibs_table_probe_count_kmc(self.handle, prefix.encode(), threads,
                                                 C.byref(out)), "hipibs_table_probe_count_kmc")
        return out.value

    def free(self) -> None:
        if self.handle:
            _lib.hipibs_table_free(self.handle)
            self.handle = C.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _table_array(tables: Sequence[Table]):
    arr = (vp * len(tables))(*[t.handle.value for t in tables])
    return arr


# ---- the streaming scan engine --------------------------------------------------

class ScanEngine:
    """hipibs_scan: pinned double-buffered slots (the EncodeProbeEngine role)."""

    def __init__(self, tables: Sequence[Table], batch_bases: int, max_windows: int, depth: int = 2):
        self.tables = list(tables)
        h = vp()
        _check(_lib.hipibs_scan_create(_table_array(self.tables), len(self.tables), batch_bases,
                                       max_windows, depth, C.byref(h)), "hipibs_scan_create")
        self.handle = h
        self.depth = depth
        self.slots = []
        for i in range(depth):
            s = HostSlot()
            _check(_lib.hipibs_scan_slot(self.handle, i, C.byref(s)), "hipibs_scan_slot")
            self.slots.append(s)

    def slot_arrays(self, i: int):
        s = self.slots[i]
        seq = np.ctypeslib.as_array(s.seq_words, shape=(s.seq_cap,))
        exc = np.ctypeslib.as_array(s.exc_words, shape=(s.exc_cap,))
        win = np.ctypeslib.as_array(C.cast(s.windows, C.POINTER(C.c_uint8)),
                                    shape=(s.window_cap * WINDOW_DTYPE.itemsize,)).view(WINDOW_DTYPE)
        st = np.ctypeslib.as_array(C.cast(s.stats, C.POINTER(C.c_uint8)),
                                   shape=(s.window_cap * s.n_tables * STATS_DTYPE.itemsize,)
                                   ).view(STATS_DTYPE)
        return seq, exc, win, st

    def submit(self, i: int, n_windows: int, n_seq_words: int, n_exc_words: int) -> None:
        _check(_lib.hipibs_scan_submit(self.handle, i, n_windows, n_seq_words, n_exc_words),
               "hipibs_scan_submit")

    def wait(self, i: int) -> None:
        _check(_lib.hipibs_scan_wait(self.handle, i), "hipibs_scan_wait")

    def free(self) -> None:
        if self.handle:
            _lib.hipibs_scan_free(self.handle)
            self.handle = vp(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def scan_sequences(tables: Sequence[Table], ids: Sequence[bytes], seqs: Sequence[bytes], W: int,
                   batch_bases: int = 1 << 26, depth: int = 2) -> List[List[tuple]]:
    """Per-window statistics of every contig against every table through the
    streaming engine; returns one list of TSV rows per table.

    A batch holds whole windows (never split, as buildBatches,
    src/gpu/EncodeProbeCore.hpp:103-134); consecutive windows of a contig
    share one packed region of the batch (they overlap by k bases), so a
    batch costs 0.375 B per distinct base, not per window base."""
    k = tables[0].info()["k"]
    wins = chunk_sequences([len(s) for s in seqs], W, k)
    if not wins:
        return [[] for _ in tables]
    wc = np.array([w[0] for w in wins], dtype=np.int64)
    ws = np.array([w[1] for w in wins], dtype=np.int64)
    we = np.array([w[2] for w in wins], dtype=np.int64)
    max_win = int((we - ws).max())
    batch_bases = max(batch_bases, max_win + 128)
    max_windows = int(min(len(wins), 1 << 18))
    # group windows into batches of regions
    batches = []  # list of (win_lo, win_hi, regions=[(contig, rs, re, first_win)])
    lo = 0
    regions = []
    used = 0
    for i in range(len(wins)):
        c, s_, e_ = int(wc[i]), int(ws[i]), int(we[i])
        if regions and regions[-1][0] == c and s_ >= regions[-1][1]:
            grow = max(0, e_ - regions[-1][2])
            extra = grow
            new_region = False
        else:
            extra = (e_ - s_) + 64
            new_region = True
        if (used + extra + 64 > batch_bases or i - lo >= max_windows) and i > lo:
            batches.append((lo, i, regions))
            lo, regions, used = i, [], 0
            new_region, extra = True, (e_ - s_) + 64
        if new_region:
            regions.append([c, s_, e_, i])
        else:
            regions[-1][2] = max(regions[-1][2], e_)
        used += extra
    batches.append((lo, len(wins), regions))

    eng = ScanEngine(tables, batch_bases, max_windows=max_windows, depth=depth)
    nt = len(tables)
    stats = np.zeros((nt, len(wins), 4), dtype=np.int32)
    try:
        pending = {}
        for bi, (blo, bhi, regs) in enumerate(batches):
            slot = bi % depth
            if slot in pending:
                eng.wait(slot)
                _collect(eng, slot, pending.pop(slot), stats)
            seq, exc, win, _ = eng.slot_arrays(slot)
            at = 0
            base_of = np.zeros(bhi - blo, dtype=np.int64)
            for r, (c, rs, re, first) in enumerate(regs):
                last = regs[r + 1][3] if r + 1 < len(regs) else bhi
                n = re - rs
                seq[at // 32:(at + n + 63) // 32 + 2] = 0
                exc[at // 64:(at + n + 63) // 64 + 2] = 0
                text = seqs[c][rs:re]
                if n:
                    _check(_lib.hipibs_pack_bases(bytes(text), n, at, seq.ctypes.data_as(u64p),
                                                  exc.ctypes.data_as(u64p)), "hipibs_pack_bases")
                base_of[first - blo:last - blo] = at - rs
                at += (n + 63) // 64 * 64
            nb = bhi - blo
            wv = win[:nb]
            wv["base"] = (ws[blo:bhi] + base_of).astype(np.uint64)
            wv["n_bases"] = (we[blo:bhi] - ws[blo:bhi]).astype(np.uint32)
            wv["reserved"] = 0
            eng.submit(slot, nb, (at + 31) // 32, (at + 63) // 64)
            pending[slot] = (blo, bhi)
        for slot in sorted(pending):
            eng.wait(slot)
            _collect(eng, slot, pending[slot], stats)
    finally:
        eng.free()
    rows = []
    for t in range(nt):
        st = stats[t].tolist()
        rows.append([(ids[int(wc[i])], int(ws[i]), int(we[i])) + tuple(st[i])
                     for i in range(len(wins))])
    return rows


def _collect(eng: ScanEngine, slot: int, rng: Tuple[int, int], stats: np.ndarray) -> None:
    _, _, _, st = eng.slot_arrays(slot)
    lo, hi = rng
    n = hi - lo
    arr = st[:n * stats.shape[0]].view(np.int32).reshape(stats.shape[0], n, 4)
    stats[:, lo:hi, :] = arr


class KmerCounter:
    """Canonical k-mer counts of reads on the device (hipibs_counter): the
    counting step KMC performs for the reference (`kmc -k -ci -cx -cs`)."""

    def __init__(self, k: int, device: int = 0, expected_keys: int = 0, chunk_bytes: int = 0):
        h = vp()
        _check(_lib.hipibs_counter_create(k, device, expected_keys, chunk_bytes, C.byref(h)),
               "hipibs_counter_create")
        self.handle = h
        self.k = k
        self.device = device

    def add_file(self, path: str) -> None:
        _check(_lib.hipibs_counter_add_file(self.handle, str(path).encode()),
               "hipibs_counter_add_file")

    def add_text(self, text: bytes) -> None:
        _check(_lib.hipibs_counter_add_text(self.handle, text, len(text)),
               "hipibs_counter_add_text")

    def add_text_device(self, d_text: int, n: int) -> None:
        """Counts n bytes of text at device address d_text (16 B aligned)."""
        _check(_lib.hipibs_counter_add_text_device(self.handle, vp(d_text), n),
               "hipibs_counter_add_text_device")

    def stats(self) -> dict:
        st = CountStats()
        _check(_lib.hipibs_counter_stats(self.handle, C.byref(st)), "hipibs_counter_stats")
        out = {n: getattr(st, n) for n, _ in CountStats._fields_}
        out["stage_ms"] = list(st.stage_ms)
        return out

    def keys(self, min_count: int = 1, max_count: int = 0xFFFFFFFF, counter_max: int = 255):
        """(keys, counts uint32, filter dict), in no particular order; keys
        uint64 (k <= 32) or (n, 2) high/low words."""
        n = self.stats()["distinct"]
        keys = np.empty((n, 2) if self.k > 32 else n, dtype=np.uint64)
        counts = np.empty(n, dtype=np.uint32)
        f = CountFilter()
        _check(_lib.hipibs_counter_keys(self.handle, min_count, max_count, counter_max,
                                        _ptr(keys), _ptr(counts, u32p), n, C.byref(f)),
               "hipibs_counter_keys")
        kept = min(f.kept, n)
        return keys[:kept], counts[:kept], {nm: getattr(f, nm) for nm, _ in CountFilter._fields_}

    def build_table(self, min_count: int = 1, max_count: int = 0xFFFFFFFF, **kw) -> "Table":
        kw.setdefault("device", self.device)
        o = _opts(**kw)
        h = vp()
        f = CountFilter()
        _check(_lib.hipibs_counter_build_table(self.handle, min_count, max_count, C.byref(o),
                                               C.byref(h), C.byref(f)),
               "hipibs_counter_build_table")
        return Table(h.value)

    def write_kmc(self, prefix: str, min_count: int = 2, max_count: int = 1_000_000_000,
                  counter_max: int = 255) -> dict:
        f = CountFilter()
        _check(_lib.hipibs_counter_write_kmc(self.handle, str(prefix).encode(), min_count,
                                             max_count, counter_max, C.byref(f)),
               "hipibs_counter_write_kmc")
        return {nm: getattr(f, nm) for nm, _ in CountFilter._fields_}

    def free(self) -> None:
        if self.handle:
            _lib.hipibs_counter_free(self.handle)
            self.handle = vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ServeEngine:
    """hipibs_serve: resident sample tables, continuous batching of scan
    queries, per-request latency."""

    def __init__(self, k: int, hbm_budget: int, max_inflight: int = 3,
                 max_windows_per_launch: int = 65536, device: int = 0, **table_opts):
        h = vp()
        o = _opts(device=device, **table_opts)
        _check(_lib.hipibs_serve_create(device, k, int(hbm_budget), C.byref(o), max_inflight,
                                        max_windows_per_launch, C.byref(h)), "hipibs_serve_create")
        self.handle = h
        self.k = k
        self._keep = []

    def set_reference(self, seq_words: np.ndarray, exc_words: np.ndarray, n_bases: int) -> None:
        self._keep += [seq_words, exc_words]
        _check(_lib.hipibs_serve_set_reference(self.handle, _ptr(seq_words), _ptr(exc_words),
                                               n_bases), "hipibs_serve_set_reference")

    def add_sample(self, sample_id: int, keys: np.ndarray, preload: bool = False) -> None:
        """keys: uint64 codes (k <= 32) or an (n, 2) array of (high, low)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        _check(_lib.hipibs_serve_add_sample(self.handle, sample_id, _ptr(keys),
                                            _key_count(keys, self.k), int(preload)),
               "hipibs_serve_add_sample")

    def add_sample_device(self, sample_id: int, d_keys: int, n_keys: int,
                          preload: bool = False) -> None:
        """d_keys: device address of n_keys codes (two words each for k > 32),
        complete before the call."""
        _check(_lib.hipibs_serve_add_sample_device(self.handle, sample_id, vp(d_keys), n_keys,
                                                   int(preload)),
               "hipibs_serve_add_sample_device")

    def add_sample_kmc(self, sample_id: int, kmc_prefix: str, preload: bool = False) -> None:
        _check(_lib.hipibs_serve_add_sample_kmc(self.handle, sample_id, kmc_prefix.encode(),
                                                int(preload)), "hipibs_serve_add_sample_kmc")

    def submit(self, request_id: int, sample_id: int, base: int, n_bases: int, W: int) -> None:
        _check(_lib.hipibs_serve_submit(self.handle, request_id, sample_id, base, n_bases, W),
               "hipibs_serve_submit")

    def poll(self, timeout_ms: int = 0, max_n: int = 4096) -> List[int]:
        ids = np.zeros(max_n, dtype=np.uint64)
        n = C.c_uint32(0)
        _check(_lib.hipibs_serve_poll(self.handle, _ptr(ids), max_n, timeout_ms, C.byref(n)),
               "hipibs_serve_poll")
        return [int(x) for x in ids[:n.value]]

    def fetch(self, request_id: int):
        """(stats array (n, 4) int32, latency ms, queue ms)."""
        n = C.c_uint32(0)
        lat, q = C.c_double(0), C.c_double(0)
        _check(_lib.hipibs_serve_fetch(self.handle, request_id, None, 0, C.byref(n), None, None),
               "hipibs_serve_fetch")
        out = np.zeros((max(1, n.value), 4), dtype=np.int32)
        _check(_lib.hipibs_serve_fetch(self.handle, request_id, out.ctypes.data_as(vp),
                                       out.shape[0], C.byref(n), C.byref(lat), C.byref(q)),
               "hipibs_serve_fetch")
        if n.value > out.shape[0]:
            raise RuntimeError("request has more windows than the fetch buffer")
        return out[:n.value].copy(), lat.value, q.value

    def counters(self) -> dict:
        c = ServeCounters()
        _check(_lib.hipibs_serve_get_counters(self.handle, C.byref(c)), "hipibs_serve_get_counters")
        return {f: getattr(c, f) for f, _ in ServeCounters._fields_}

    def free(self) -> None:
        if self.handle:
            _lib.hipibs_serve_free(self.handle)
            self.handle = vp(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def reference_name(filename: str) -> str:
    """getReferenceName (src/FastIBS.cpp:14-25): cut at the first '_genome'."""
    pos = filename.find("_genome")
    return filename[:pos] if pos >= 0 else filename


def output_name(db_folder: str, ref_file: str, W: int) -> str:
    """<db folder name>_v_<reference name>_<W>.tsv (src/FastIBS.cpp:99)."""
    return f"{os.path.basename(db_folder.rstrip('/'))}_v_{reference_name(ref_file)}_{W}.tsv"
