"""ctypes binding of libdakc_amd.so (include/dakc.h).

Loads the in-tree library only; there is no fallback: if the HIP library is
missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libdakc_amd.so"

DAKC_OK = 0
DAKC_NOMEM = 3
STATUS_NAMES = {
    0: "DAKC_OK", 1: "DAKC_E_INVAL", 2: "DAKC_E_FORMAT", 3: "DAKC_E_NOMEM", 4: "DAKC_E_TOOBIG",
    5: "DAKC_E_DEVICE", 6: "DAKC_E_NODEVICE", 7: "DAKC_E_REJECTED", 8: "DAKC_E_NOTFOUND",
    9: "DAKC_E_PENDING", 10: "DAKC_E_CANCELLED",
}
FMT_FIXED, FMT_LINES, FMT_FASTQ = 0, 1, 2
NOALLOC_SCRATCH = 256 << 20  # DAKC_NOALLOC_SCRATCH
ALGO_AUTO, ALGO_LSD, ALGO_MSD = 0, 1, 2

# every symbol include/dakc.h declares
EXPORTS = [
    "dakc_status_str", "dakc_abi_version", "dakc_ctx_create", "dakc_ctx_destroy", "dakc_count",
    "dakc_count_device", "dakc_last_stats", "dakc_count_batch", "dakc_submit", "dakc_poll",
    "dakc_wait", "dakc_table_free", "dakc_table_serialize", "dakc_gen_nbytes", "dakc_generate_reads",
    "dakc_generate_reads_device", "dakc_device_alloc", "dakc_device_free", "dakc_memcpy_h2d",
    "dakc_memcpy_d2h", "dakc_memcpy_d2d", "dakc_batch_info", "dakc_serve_get_stats", "dakc_batch_create", "dakc_batch_run",
    "dakc_batch_tables", "dakc_batch_destroy", "dakc_autotune", "dakc_tuning_save",
    "dakc_tuning_load", "dakc_table_histogram", "dakc_table_filter", "dakc_table_query",
    "dakc_convert_reads", "dakc_plan_job", "dakc_count_noalloc", "dakc_partition", "dakc_count_keys",
    "dakc_count_hitter",
]


class DakcError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        msg = status_str(status) if _lib is not None else name
        super().__init__(f"{where}: {name} ({msg})")


class Job(C.Structure):
    _fields_ = [
        ("reads", C.c_void_p), ("nbytes", C.c_uint64), ("format", C.c_int32),
        ("read_len", C.c_int32), ("k", C.c_int32), ("canonical", C.c_int32),
        ("num_pes", C.c_uint32), ("owner_bin", C.c_int32), ("reads_on_device", C.c_int32),
        ("table_on_device", C.c_int32), ("hash_seed", C.c_uint64),
    ]


class Table(C.Structure):
    _fields_ = [
        ("keys", C.POINTER(C.c_uint64)), ("counts", C.POINTER(C.c_uint32)), ("n", C.c_uint64),
        ("total", C.c_uint64), ("bin_offsets", C.POINTER(C.c_uint64)), ("num_bins", C.c_uint32),
        ("key_words", C.c_int32), ("k", C.c_int32), ("on_device", C.c_int32), ("priv", C.c_void_p),
        ("hash_seed", C.c_uint64),
    ]


class Opts(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_inflight", C.c_int32), ("algo", C.c_int32),
                ("max_large", C.c_int32), ("batch_solo_items", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [
        ("hist_ms", C.c_double), ("pass0_ms", C.c_double), ("passes_ms", C.c_double),
        ("rle_ms", C.c_double), ("total_ms", C.c_double), ("npasses", C.c_uint32),
        ("tiles0", C.c_uint32), ("n", C.c_uint64), ("d", C.c_uint64), ("sweep_ms", C.c_double * 18),
        ("path", C.c_uint32), ("msd_reruns", C.c_uint32), ("part2_ms", C.c_double),
        ("final_ms", C.c_double),
    ]


class PlanInfo(C.Structure):
    _fields_ = [("path", C.c_uint32), ("key_words", C.c_uint32), ("kmer_slots", C.c_uint64),
                ("device_bytes", C.c_uint64), ("b1", C.c_uint32), ("b2", C.c_uint32), ("b3", C.c_uint32),
                ("first_level_buckets", C.c_uint32), ("l1_tiles", C.c_uint32), ("l2_groups", C.c_uint32),
                ("final_items", C.c_uint32), ("final_threads", C.c_uint32), ("radix_passes", C.c_uint32)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class HitterStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("buckets", "normal_keys", "heavy_pairs", "kmers", "bytes_sent",
                                          "bytes_without")]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class ServeStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "admitted", "completed", "failed", "lsd_reruns",
                                          "budget", "peak_device_bytes", "resident_bytes",
                                          "pool_used", "queued")] + [("admit_host_ms", C.c_double),
                                                                     ("init_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class GenParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("num_reads", C.c_uint64), ("first_read", C.c_uint64),
        ("genome_len", C.c_uint64), ("sub_rate", C.c_double), ("n_rate", C.c_double),
        ("read_len", C.c_int32), ("min_len", C.c_int32), ("n_run_max", C.c_int32),
        ("both_strands", C.c_int32),
    ]


_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m dakc_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    P, V = C.POINTER, C.c_void_p
    sig = {
        "dakc_status_str": (C.c_char_p, [C.c_int]),
        "dakc_abi_version": (C.c_int, []),
        "dakc_ctx_create": (C.c_int, [C.c_int, P(Opts), P(V)]),
        "dakc_ctx_destroy": (C.c_int, [V]),
        "dakc_count": (C.c_int, [V, P(Job), P(Table)]),
        "dakc_count_device": (C.c_int, [V, P(Job), P(Table)]),
        "dakc_last_stats": (C.c_int, [V, P(Stats)]),
        "dakc_count_batch": (C.c_int, [V, P(Job), C.c_size_t, P(Table)]),
        "dakc_submit": (C.c_int, [V, P(Job), P(C.c_uint64)]),
        "dakc_poll": (C.c_int, [V, C.c_uint64, P(C.c_int)]),
        "dakc_wait": (C.c_int, [V, C.c_uint64, P(Table)]),
        "dakc_table_free": (None, [P(Table)]),
        "dakc_table_serialize": (C.c_uint64, [P(Table), V, C.c_uint64]),
        "dakc_gen_nbytes": (C.c_uint64, [P(GenParams)]),
        "dakc_generate_reads": (C.c_int, [P(GenParams), V, C.c_uint64, C.c_int]),
        "dakc_generate_reads_device": (C.c_int, [V, P(GenParams), V, C.c_uint64]),
        "dakc_device_alloc": (C.c_int, [V, C.c_uint64, P(V)]),
        "dakc_device_free": (C.c_int, [V, V]),
        "dakc_memcpy_h2d": (C.c_int, [V, V, V, C.c_uint64]),
        "dakc_memcpy_d2h": (C.c_int, [V, V, V, C.c_uint64]),
        "dakc_memcpy_d2d": (C.c_int, [V, V, V, C.c_uint64]),
        "dakc_batch_create": (C.c_int, [V, P(Job), C.c_size_t, P(V)]),
        "dakc_batch_info": (C.c_int, [V, P(C.c_size_t), P(C.c_uint64)]),
        "dakc_serve_get_stats": (C.c_int, [V, P(ServeStats)]),
        "dakc_batch_run": (C.c_int, [V, P(C.c_double)]),
        "dakc_batch_tables": (C.c_int, [V, P(Table)]),
        "dakc_batch_destroy": (C.c_int, [V]),
        "dakc_autotune": (C.c_int, [V, P(Job), C.c_int, P(C.c_double)]),
        "dakc_tuning_save": (C.c_int, [V, C.c_char_p]),
        "dakc_tuning_load": (C.c_int, [V, C.c_char_p]),
        "dakc_table_histogram": (C.c_int, [V, P(Table), C.c_uint32, V, P(C.c_uint32)]),
        "dakc_table_filter": (C.c_int, [V, P(Table), C.c_uint32]),
        "dakc_table_query": (C.c_int, [V, P(Table), V, C.c_uint64, C.c_int, V]),
        "dakc_convert_reads": (C.c_int, [V, V, C.c_uint64, C.c_int, C.c_uint32, V, C.c_uint64,
                                         P(C.c_uint64), P(C.c_uint32)]),
        "dakc_plan_job": (C.c_int, [P(Job), C.c_int, C.c_uint32, P(PlanInfo)]),
        "dakc_count_noalloc": (C.c_int, [V, P(Job), V, V, C.c_uint64, V, P(C.c_uint64), P(C.c_uint64)]),
        "dakc_partition": (C.c_int, [V, P(Job), V, C.c_uint64, V, P(C.c_uint64)]),
        "dakc_count_keys": (C.c_int, [V, V, C.c_uint64, C.c_int, C.c_uint32, C.c_uint64, C.c_int, C.c_int,
                                      P(Table)]),
        "dakc_count_hitter": (C.c_int, [V, P(Job), C.c_uint32, P(Table), P(Table), P(HitterStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def status_str(s: int) -> str:
    return load().dakc_status_str(int(s)).decode()


def check(status: int, where: str) -> None:
    if status != DAKC_OK:
        raise DakcError(status, where)


def gen_params(seed: int, num_reads: int, read_len: int = 150, *, min_len: int = 0,
               genome_len: int = 0, sub_rate: float = 0.0, n_rate: float = 0.0,
               n_run_max: int = 10, both_strands: bool = True, first_read: int = 0) -> GenParams:
    return GenParams(seed=seed, num_reads=num_reads, first_read=first_read, genome_len=genome_len,
                     sub_rate=sub_rate, n_rate=n_rate, read_len=read_len, min_len=min_len,
                     n_run_max=n_run_max, both_strands=1 if both_strands else 0)


def generate_reads(p: GenParams, threads: int = 0) -> np.ndarray:
    """Host generator: the record buffer as a uint8 array."""
    lib = load()
    nb = lib.dakc_gen_nbytes(C.byref(p))
    buf = np.empty(nb, dtype=np.uint8)
    check(lib.dakc_generate_reads(C.byref(p), buf.ctypes.data_as(C.c_void_p), nb,
                                  threads or min(16, os.cpu_count() or 1)), "dakc_generate_reads")
    return buf


def table_to_numpy(t: Table) -> tuple[np.ndarray, np.ndarray, np.ndarray | None]:
    """Copy a host table into numpy: keys (n, key_words) u64, counts (n,) u32, bin offsets."""
    n, kw = int(t.n), int(t.key_words)
    if n:
        keys = np.ctypeslib.as_array(t.keys, shape=(n * kw,)).copy().reshape(n, kw)
        counts = np.ctypeslib.as_array(t.counts, shape=(n,)).copy()
    else:
        keys = np.zeros((0, kw), np.uint64)
        counts = np.zeros(0, np.uint32)
    offs = None
    if t.num_bins:
        offs = np.ctypeslib.as_array(t.bin_offsets, shape=(t.num_bins + 1,)).copy()
    return keys, counts, offs


def plan_job(job: Job, algo: int = ALGO_AUTO, num_cus: int = 0) -> dict:
    """The geometry a job would be counted with (host only, no device call)."""
    info = PlanInfo()
    check(load().dakc_plan_job(C.byref(job), algo, num_cus, C.byref(info)), "dakc_plan_job")
    return info.as_dict()
