"""ctypes binding of the CPU oracle (oracle/dakc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product (dakc_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libdakc_oracle.so"
OWNER_SEED = 0x9E3779B97F4A7C15

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    lib.dko_char2base.restype = C.c_uint8
    lib.dko_char2base.argtypes = [C.c_ubyte]
    lib.dko_murmur64a_u64.restype = C.c_uint64
    lib.dko_murmur64a_u64.argtypes = [C.c_uint64, C.c_uint64]
    lib.dko_murmur64a_u128.restype = C.c_uint64
    lib.dko_murmur64a_u128.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    lib.dko_extract.restype = C.c_int64
    lib.dko_extract.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_uint32,
                                C.c_int32, C.c_uint64, C.c_void_p, C.c_uint64]
    lib.dko_count.restype = C.c_int
    lib.dko_count.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_uint32,
                              C.c_int32, C.c_uint64, P(P(C.c_uint64)), P(P(C.c_uint32)), P(C.c_uint64),
                              P(C.c_uint64), C.c_void_p]
    lib.dko_count_mt.restype = C.c_int
    lib.dko_count_mt.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int32,
                                 C.c_uint64, C.c_int, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64),
                                 P(C.c_double)]
    lib.dko_count_keys_mt.restype = C.c_int
    lib.dko_count_keys_mt.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, P(C.c_uint64),
                                      P(C.c_uint64), P(C.c_double)]
    lib.dko_extract_mt.restype = C.c_int64
    lib.dko_extract_mt.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_uint32,
                                   C.c_int32, C.c_uint64, C.c_int, C.c_void_p, C.c_uint64]
    lib.dko_count_hitter.restype = C.c_int
    lib.dko_count_hitter.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int32,
                                     C.c_uint64, C.c_uint32, P(P(C.c_uint64)), P(P(C.c_uint32)), P(C.c_uint64),
                                     C.c_void_p, P(P(C.c_uint64)), P(P(C.c_uint32)), P(C.c_uint64), C.c_void_p,
                                     C.c_void_p]
    lib.dko_table_checksum.restype = C.c_uint64
    lib.dko_table_checksum.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]
    lib.dko_free.restype = None
    lib.dko_free.argtypes = [C.c_void_p]
    lib.dko_convert.restype = C.c_int
    lib.dko_convert.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_void_p, C.c_size_t,
                                P(C.c_uint64), P(C.c_uint32)]
    lib.dko_histogram.restype = C.c_uint32
    lib.dko_histogram.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]
    lib.dko_filter.restype = C.c_uint64
    lib.dko_filter.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_uint32, C.c_void_p,
                               C.c_uint32, P(C.c_uint64)]
    lib.dko_query.restype = None
    lib.dko_query.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint32,
                              C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p]
    _lib = lib
    return lib


def _buf(reads) -> np.ndarray:
    if isinstance(reads, np.ndarray):
        return np.ascontiguousarray(reads).view(np.uint8).reshape(-1)
    return np.frombuffer(bytes(reads), dtype=np.uint8)


class OracleError(RuntimeError):
    def __init__(self, status):
        self.status = status
        super().__init__(f"oracle status {status}")


def extract(reads, read_len, k, canonical=False, num_pes=1, owner_bin=-1, seed=0) -> np.ndarray:
    """Every k-mer of the records in record order, (n, key_words) u64."""
    lib = load()
    b = _buf(reads)
    ptr = b.ctypes.data if b.size else None
    n = lib.dko_extract(ptr, b.size, read_len, k, int(canonical), num_pes, owner_bin, seed, None, 0)
    if n < 0:
        raise OracleError(-n)
    kw = 1 if k <= 32 else 2
    out = np.zeros((max(n, 1), kw), np.uint64)
    lib.dko_extract(ptr, b.size, read_len, k, int(canonical), num_pes, owner_bin, seed,
                    out.ctypes.data, n)
    return out[:n]


def extract_mt(reads, read_len, k, threads, canonical=False, num_pes=1, owner_bin=-1, seed=0,
               cap=None) -> np.ndarray:
    """extract() on `threads` CPU threads (same keys, same order)."""
    lib = load()
    b = _buf(reads)
    ptr = b.ctypes.data if b.size else None
    kw = 1 if k <= 32 else 2
    if cap is None:
        n = lib.dko_extract_mt(ptr, b.size, read_len, k, int(canonical), num_pes, owner_bin, seed, threads,
                               None, 0)
        if n < 0:
            raise OracleError(-n)
        cap = n
    out = np.zeros((max(cap, 1), kw), np.uint64)
    n = lib.dko_extract_mt(ptr, b.size, read_len, k, int(canonical), num_pes, owner_bin, seed, threads,
                           out.ctypes.data, cap)
    if n < 0:
        raise OracleError(-n)
    if n > cap:
        raise OracleError(3)
    return out[:n]


def count(reads, read_len, k, canonical=False, num_pes=1, owner_bin=-1, seed=0):
    """Returns (keys (d, kw) u64, counts (d,) u32, total n, bin_offsets or None)."""
    lib = load()
    b = _buf(reads)
    ptr = b.ctypes.data if b.size else None
    keys = C.POINTER(C.c_uint64)()
    counts = C.POINTER(C.c_uint32)()
    d = C.c_uint64()
    n = C.c_uint64()
    all_bins = num_pes > 1 and owner_bin < 0
    offs = np.zeros(num_pes + 1, np.uint64) if all_bins else None
    st = lib.dko_count(ptr, b.size, read_len, k, int(canonical), num_pes, owner_bin, seed, C.byref(keys),
                       C.byref(counts), C.byref(d), C.byref(n),
                       offs.ctypes.data if offs is not None else None)
    if st:
        raise OracleError(st)
    kw = 1 if k <= 32 else 2
    dn = int(d.value)
    try:
        if dn:
            kk = np.ctypeslib.as_array(keys, shape=(dn * kw,)).copy().reshape(dn, kw)
            cc = np.ctypeslib.as_array(counts, shape=(dn,)).copy()
        else:
            kk = np.zeros((0, kw), np.uint64)
            cc = np.zeros(0, np.uint32)
    finally:
        lib.dko_free(C.cast(keys, C.c_void_p))
        lib.dko_free(C.cast(counts, C.c_void_p))
    return kk, cc, int(n.value), offs


def count_mt(reads, read_len, k, threads, canonical=False, num_pes=1, owner_bin=-1, seed=0):
    """DAKC's P-PE shape on `threads` CPU threads; returns (n, d, checksum, seconds)."""
    lib = load()
    b = _buf(reads)
    n, d, cs, el = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_double()
    st = lib.dko_count_mt(b.ctypes.data if b.size else None, b.size, read_len, k,
                          1 if canonical else 0, num_pes, owner_bin, seed, threads,
                          C.byref(n), C.byref(d), C.byref(cs), C.byref(el))
    if st:
        raise OracleError(st)
    return int(n.value), int(d.value), int(cs.value), float(el.value)


def count_keys_mt(keys: np.ndarray, k: int, threads: int):
    """Sort + run-length reduce of a key array on `threads` CPU threads;
    returns (d, checksum, seconds)."""
    lib = load()
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    kw = 1 if k <= 32 else 2
    n = keys.size // kw
    d, cs, el = C.c_uint64(), C.c_uint64(), C.c_double()
    st = lib.dko_count_keys_mt(keys.ctypes.data if n else None, n, k, threads, C.byref(d), C.byref(cs),
                               C.byref(el))
    if st:
        raise OracleError(st)
    return int(d.value), int(cs.value), float(el.value)


def checksum(keys: np.ndarray, counts: np.ndarray) -> int:
    lib = load()
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    return int(lib.dko_table_checksum(keys.ctypes.data, counts.ctypes.data, counts.size,
                                      keys.shape[1] if keys.ndim == 2 else 1))


def char2base(c: int) -> int:
    return int(load().dko_char2base(c))


def murmur64a_u64(key: int, seed: int = OWNER_SEED) -> int:
    return int(load().dko_murmur64a_u64(key, seed))


def murmur64a_u128(lo: int, hi: int, seed: int = OWNER_SEED) -> int:
    return int(load().dko_murmur64a_u128(lo, hi, seed))


def histogram(counts: np.ndarray, num_buckets: int):
    """(hist (num_buckets,) u64, min_bucket) of a table's counts."""
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    hist = np.zeros(max(num_buckets, 1), dtype=np.uint64)
    m = load().dko_histogram(counts.ctypes.data, counts.size, num_buckets, hist.ctypes.data)
    return hist[:num_buckets], int(m)


def filter_table(keys: np.ndarray, counts: np.ndarray, min_count: int, bin_offsets=None):
    """(keys, counts, total, bin_offsets) of the entries with count >= min_count."""
    keys = np.array(keys, dtype=np.uint64, order="C", copy=True)
    kw = keys.shape[1] if keys.ndim == 2 else 1
    counts = np.array(counts, dtype=np.uint32, copy=True)
    offs = None if bin_offsets is None else np.array(bin_offsets, dtype=np.uint64, copy=True)
    total = C.c_uint64()
    n = load().dko_filter(keys.ctypes.data, counts.ctypes.data, counts.size, kw, min_count,
                          offs.ctypes.data if offs is not None else None,
                          (offs.size - 1) if offs is not None else 0, C.byref(total))
    keys = keys.reshape(-1, kw)[:n]
    return keys, counts[:n], int(total.value), offs


def query(keys: np.ndarray, counts: np.ndarray, queries: np.ndarray, bin_offsets=None, seed=0) -> np.ndarray:
    """Count of each query (0 when absent) in a sorted table."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    kw = keys.shape[1] if keys.ndim == 2 else 1
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    queries = np.ascontiguousarray(queries, dtype=np.uint64).reshape(-1, kw)
    offs = None if bin_offsets is None else np.ascontiguousarray(bin_offsets, dtype=np.uint64)
    out = np.zeros(queries.shape[0], dtype=np.uint32)
    load().dko_query(keys.ctypes.data, counts.ctypes.data, counts.size, kw,
                     offs.ctypes.data if offs is not None else None,
                     (offs.size - 1) if offs is not None else 0, seed, queries.ctypes.data, queries.shape[0],
                     out.ctypes.data)
    return out


def convert(data, fastq: bool, read_len: int = 0):
    """(fixed-width record bytes as uint8 array, width) of FASTQ or line input."""
    lib = load()
    b = _buf(data)
    n, L = C.c_uint64(), C.c_uint32()
    p = b.ctypes.data if b.size else None
    st = lib.dko_convert(p, b.size, 1 if fastq else 0, read_len, None, 0, C.byref(n), C.byref(L))
    if st:
        raise OracleError(st)
    out = np.zeros(max(int(n.value), 1), dtype=np.uint8)
    st = lib.dko_convert(p, b.size, 1 if fastq else 0, read_len, out.ctypes.data, out.size, C.byref(n),
                         C.byref(L))
    if st:
        raise OracleError(st)
    return out[:int(n.value)], int(L.value)


class HitterStats(C.Structure):
    _fields_ = [("buckets", C.c_uint64), ("normal_keys", C.c_uint64), ("heavy_pairs", C.c_uint64)]


def count_hitter(reads, read_len, k, canonical=False, num_pes=1, owner_bin=-1, seed=0, bucket_size=10000):
    """HITTER=1: (light (keys, counts, offsets), heavy (keys, counts, offsets), stats dict)."""
    lib = load()
    b = _buf(reads)
    ptr = b.ctypes.data if b.size else None
    lk, hk = C.POINTER(C.c_uint64)(), C.POINTER(C.c_uint64)()
    lc, hc = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint32)()
    dl, dh = C.c_uint64(), C.c_uint64()
    nb = num_pes if num_pes > 1 and owner_bin < 0 else 1
    lo, ho = np.zeros(nb + 1, np.uint64), np.zeros(nb + 1, np.uint64)
    st = HitterStats()
    r = lib.dko_count_hitter(ptr, b.size, read_len, k, int(canonical), num_pes, owner_bin, seed, bucket_size,
                             C.byref(lk), C.byref(lc), C.byref(dl), lo.ctypes.data, C.byref(hk), C.byref(hc),
                             C.byref(dh), ho.ctypes.data, C.byref(st))
    if r:
        raise OracleError(r)
    kw = 1 if k <= 32 else 2

    def take(kp, cp, d):
        d = int(d.value)
        try:
            if d:
                return (np.ctypeslib.as_array(kp, shape=(d * kw,)).copy().reshape(d, kw),
                        np.ctypeslib.as_array(cp, shape=(d,)).copy())
            return np.zeros((0, kw), np.uint64), np.zeros(0, np.uint32)
        finally:
            lib.dko_free(C.cast(kp, C.c_void_p))
            lib.dko_free(C.cast(cp, C.c_void_p))

    light = take(lk, lc, dl) + (lo,)
    heavy = take(hk, hc, dh) + (ho,)
    return light, heavy, {n: int(getattr(st, n)) for n, _ in HitterStats._fields_}
