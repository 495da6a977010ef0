"""Python face of the counting path, mirroring the reference's driver surface.

The reference drives one PE as (dakc/src/main/main.cpp:40-49)

    fqreader fq(file, READLEN, rank, size);  char* chunk = fq.read_file();
    kmercounter km(chunk, vectordbg);         // runs perform_kcount()

with KMERLEN / READLEN fixed at compile time (dakc/Makefile:11).  Here the same
steps are `FqReader(path, read_len).read_file()` and
`KmerCounter(k=..., read_len=...).count(chunk)`, with k and read_len chosen at
run time; the result is the sorted (kmer, count) table the reference keeps in
`lightdbg` (kcounter.hpp:76-79, kcounter.cpp:385-411).  Every call goes through
the C ABI of libdakc_amd.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import Job, Opts, Stats, Table, check


@dataclass
class KmerTable:
    """Sorted (kmer, count) table.  keys: (n, key_words) uint64 (lo, hi for
    k > 32); counts: (n,) uint32; bin_offsets: owner-bin boundaries when all P
    bins were counted."""
    keys: np.ndarray
    counts: np.ndarray
    k: int
    total: int
    bin_offsets: np.ndarray | None = None

    @property
    def key_words(self) -> int:
        return self.keys.shape[1]

    def __len__(self) -> int:
        return int(self.counts.shape[0])

    def serialize(self) -> bytes:
        """Little-endian records: u64 kmer (or u64 lo, u64 hi) then u32 count."""
        kw = self.key_words
        rec = np.zeros(len(self), dtype=[("k", "<u8", (kw,)), ("c", "<u4")])
        rec["k"] = self.keys
        rec["c"] = self.counts
        return rec.tobytes()


class FqReader:
    """fqreader (dakc/src/fqreader/fqreader.cpp:14-90) for the single-PE case:
    the whole file of read_len-byte records, each followed by '\\n'."""

    def __init__(self, filename: str, read_len: int):
        self.filename = filename
        self.read_len = read_len

    def read_file(self) -> np.ndarray:
        try:
            return np.fromfile(self.filename, dtype=np.uint8)
        except OSError as e:  # the reference prints and exits(2) (fqreader.cpp:31-36)
            raise FileNotFoundError(f"could not open the input file {self.filename}") from e


def _as_bytes(reads) -> np.ndarray:
    if isinstance(reads, np.ndarray):
        return np.ascontiguousarray(reads).view(np.uint8).reshape(-1)
    if isinstance(reads, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(reads), dtype=np.uint8)
    raise TypeError("reads must be bytes or a uint8 numpy array")


class KmerCounter:
    """One GPU context; count() replaces kmercounter(read_chunk, vectordbg)."""

    def __init__(self, k: int = 31, read_len: int = 150, *, canonical: bool = False,
                 num_pes: int = 1, owner_bin: int = -1, hash_seed: int = 0, device: int = 0,
                 hbm_budget: int = 0, max_inflight: int = 0, algo: int = 0, max_large: int = 0,
                 batch_solo_items: int = 0):
        self.k, self.read_len, self.canonical = k, read_len, canonical
        self.num_pes, self.owner_bin, self.hash_seed = num_pes, owner_bin, hash_seed
        self._libh = _lib.load()
        self._ctx = C.c_void_p()
        opts = Opts(hbm_budget=hbm_budget, max_inflight=max_inflight, algo=algo, max_large=max_large,
                    batch_solo_items=batch_solo_items)
        check(self._libh.dakc_ctx_create(device, C.byref(opts), C.byref(self._ctx)),
              "dakc_ctx_create")

    def close(self) -> None:
        if self._ctx:
            self._libh.dakc_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._ctx

    def job(self, reads_ptr: int, nbytes: int, *, on_device: bool = False, **over) -> Job:
        p = dict(k=self.k, read_len=self.read_len, canonical=self.canonical,
                 num_pes=self.num_pes, owner_bin=self.owner_bin, table_on_device=False,
                 format=_lib.FMT_FIXED, hash_seed=self.hash_seed)
        p.update(over)
        return Job(reads=reads_ptr, nbytes=nbytes, format=p["format"], read_len=p["read_len"],
                   k=p["k"], canonical=1 if p["canonical"] else 0, num_pes=p["num_pes"],
                   owner_bin=p["owner_bin"], reads_on_device=1 if on_device else 0,
                   table_on_device=1 if p["table_on_device"] else 0, hash_seed=p["hash_seed"])

    def count(self, reads, *, dev_ptr: int = 0, nbytes: int = 0, **over) -> KmerTable:
        """Count one read buffer into a host table.  With dev_ptr the reads are
        already in device memory (nbytes long)."""
        if dev_ptr:
            job = self.job(dev_ptr, nbytes, on_device=True, **over)
        else:
            buf = _as_bytes(reads)
            job = self.job(buf.ctypes.data if buf.size else 0, int(buf.size), **over)
        t = Table()
        check(self._libh.dakc_count(self._ctx, C.byref(job), C.byref(t)), "dakc_count")
        try:
            keys, counts, offs = _lib.table_to_numpy(t)
            return KmerTable(keys, counts, int(t.k), int(t.total), offs)
        finally:
            self._libh.dakc_table_free(C.byref(t))

    def count_device(self, dev_ptr: int, nbytes: int, **over) -> Table:
        """Count reads already resident in device memory; the table stays on
        the device (valid until the next call).  Returns the raw Table."""
        job = self.job(dev_ptr, nbytes, on_device=True, **over)
        t = Table()
        check(self._libh.dakc_count_device(self._ctx, C.byref(job), C.byref(t)), "dakc_count_device")
        return t

    def count_noalloc(self, dev_ptr: int, nbytes: int, keys_out: int, counts_out: int, cap: int,
                      **over) -> tuple[int, int, int, np.ndarray | None]:
        """dakc_count_noalloc: count device-resident reads into the caller's
        device buffers (cap entries) with no allocation.  Returns (status,
        distinct, total, bin offsets or None); the status is returned, not
        raised, so a caller can grow its buffers on DAKC_E_NOMEM."""
        job = self.job(dev_ptr, nbytes, on_device=True, **over)
        all_bins = job.num_pes > 1 and job.owner_bin < 0
        offs = np.zeros(job.num_pes + 1, np.uint64) if all_bins else None
        d, n = C.c_uint64(), C.c_uint64()
        st = self._libh.dakc_count_noalloc(self._ctx, C.byref(job), keys_out or None, counts_out or None, cap,
                                           offs.ctypes.data if all_bins else None, C.byref(d), C.byref(n))
        return int(st), int(d.value), int(n.value), offs

    def partition(self, reads=None, *, dev_ptr: int = 0, nbytes: int = 0, keys_out: int = 0, cap: int = 0,
                  **over) -> tuple[int, np.ndarray]:
        """Phase 1 (dakc_partition): the k-mers routed to their owner bins in
        device memory keys_out (cap keys).  Returns (k-mers, bin offsets);
        with keys_out = 0 only the count is returned (DAKC_E_NOMEM taken as
        'tell me the size')."""
        if dev_ptr:
            job = self.job(dev_ptr, nbytes, on_device=True, **over)
        else:
            buf = _as_bytes(reads)
            job = self.job(buf.ctypes.data if buf.size else 0, int(buf.size), **over)
        nb = job.num_pes if job.num_pes > 1 and job.owner_bin < 0 else 1
        offs = np.zeros(nb + 1, np.uint64)
        n = C.c_uint64()
        st = self._libh.dakc_partition(self._ctx, C.byref(job), keys_out or None, cap, offs.ctypes.data,
                                       C.byref(n))
        if not (st == _lib.DAKC_NOMEM and not keys_out):
            check(st, "dakc_partition")
        return int(n.value), offs

    def count_keys(self, keys=None, *, dev_ptr: int = 0, n: int = 0, k: int | None = None, num_pes: int = 1,
                   hash_seed: int = 0, on_device_table: bool = False):
        """Phase 2 (dakc_count_keys): the table of a key array, host (n, key_words)
        u64 array or device pointer with n keys.  Returns a KmerTable, or the
        raw device Table with on_device_table."""
        k = self.k if k is None else k
        if dev_ptr:
            ptr, on_dev = dev_ptr, 1
        else:
            arr = np.ascontiguousarray(keys, dtype=np.uint64)
            n = arr.shape[0] if arr.ndim == 2 else arr.size // (1 if k <= 32 else 2)
            ptr, on_dev = (arr.ctypes.data if n else 0), 0
        t = Table()
        check(self._libh.dakc_count_keys(self._ctx, ptr or None, n, k, num_pes, hash_seed, on_dev,
                                         1 if on_device_table else 0, C.byref(t)), "dakc_count_keys")
        if on_device_table:
            return t
        try:
            keys_, counts, offs = _lib.table_to_numpy(t)
            return KmerTable(keys_, counts, int(t.k), int(t.total), offs)
        finally:
            self._libh.dakc_table_free(C.byref(t))

    def count_hitter(self, reads=None, *, dev_ptr: int = 0, nbytes: int = 0, bucket_size: int = 0,
                     table_on_device: bool = False, **over):
        """HITTER=1 (dakc_count_hitter): (light, heavy, stats).  Host tables are
        KmerTables; with table_on_device the raw device Tables (valid until the
        next HITTER count on this context)."""
        if dev_ptr:
            job = self.job(dev_ptr, nbytes, on_device=True, table_on_device=table_on_device, **over)
        else:
            buf = _as_bytes(reads)
            job = self.job(buf.ctypes.data if buf.size else 0, int(buf.size), table_on_device=table_on_device,
                           **over)
        lt, ht, st = Table(), Table(), _lib.HitterStats()
        check(self._libh.dakc_count_hitter(self._ctx, C.byref(job), bucket_size, C.byref(lt), C.byref(ht),
                                           C.byref(st)), "dakc_count_hitter")
        if table_on_device:
            return lt, ht, st.as_dict()
        out = []
        for t in (lt, ht):
            try:
                keys, counts, offs = _lib.table_to_numpy(t)
                out.append(KmerTable(keys, counts, int(t.k), int(t.total), offs))
            finally:
                self._libh.dakc_table_free(C.byref(t))
        return out[0], out[1], st.as_dict()

    def last_stats(self) -> Stats:
        s = Stats()
        check(self._libh.dakc_last_stats(self._ctx, C.byref(s)), "dakc_last_stats")
        return s

    def count_batch(self, read_sets, **over) -> list[KmerTable]:
        bufs = [_as_bytes(r) for r in read_sets]
        jobs = (Job * len(bufs))(*[self.job(b.ctypes.data if b.size else 0, int(b.size), **over)
                                   for b in bufs])
        outs = (Table * len(bufs))()
        st = self._libh.dakc_count_batch(self._ctx, jobs, len(bufs), outs)
        try:
            check(st, "dakc_count_batch")
            res = []
            for t in outs:
                keys, counts, offs = _lib.table_to_numpy(t)
                res.append(KmerTable(keys, counts, int(t.k), int(t.total), offs))
            return res
        finally:
            for t in outs:
                self._libh.dakc_table_free(C.byref(t))

    def count_batch_device(self, read_sets=None, *, dev_ptrs=None, sizes=None, **over) -> list["DeviceTable"]:
        """dakc_count_batch with device tables of their own (table_on_device):
        host read buffers, or device pointers with their sizes.  Each table
        keeps its device memory until DeviceTable.free()."""
        if dev_ptrs is not None:
            keep = []
            jobs = [self.job(int(p), int(n), on_device=True, table_on_device=True, **over)
                    for p, n in zip(dev_ptrs, sizes)]
        else:
            keep = [_as_bytes(r) for r in read_sets]
            jobs = [self.job(b.ctypes.data if b.size else 0, int(b.size), table_on_device=True, **over)
                    for b in keep]
        arr = (Job * len(jobs))(*jobs)
        outs = (Table * len(jobs))()
        st = self._libh.dakc_count_batch(self._ctx, arr, len(jobs), outs)
        tabs = [DeviceTable(self, outs[i]) for i in range(len(jobs))]
        if st != _lib.DAKC_OK:
            for t in tabs:
                t.free()
            check(st, "dakc_count_batch")
        return tabs

    # captured batches and the autotuner
    def batch_create(self, read_sets, *, on_device: bool = False, sizes=None, **over) -> "Batch":
        """Stage many read sets for a captured multi-job launch.  read_sets are
        host buffers, or device pointers (with `sizes`) when on_device."""
        if on_device:
            jobs = [self.job(int(ptr), int(n), on_device=True, **over) for ptr, n in zip(read_sets, sizes)]
            keep = []
        else:
            keep = [_as_bytes(r) for r in read_sets]
            jobs = [self.job(b.ctypes.data if b.size else 0, int(b.size), **over) for b in keep]
        arr = (Job * len(jobs))(*jobs)
        h = C.c_void_p()
        check(self._libh.dakc_batch_create(self._ctx, arr, len(jobs), C.byref(h)), "dakc_batch_create")
        return Batch(self, h, len(jobs), keep)

    def autotune(self, reads=None, *, dev_ptr: int = 0, nbytes: int = 0, trials: int = 2, **over) -> float:
        if reads is not None:
            buf = _as_bytes(reads)
            job = self.job(buf.ctypes.data if buf.size else 0, int(buf.size), **over)
        else:
            job = self.job(dev_ptr, nbytes, on_device=True, **over)
        ms = C.c_double()
        check(self._libh.dakc_autotune(self._ctx, C.byref(job), trials, C.byref(ms)), "dakc_autotune")
        return ms.value

    def tuning_save(self, path: str) -> None:
        check(self._libh.dakc_tuning_save(self._ctx, str(path).encode()), "dakc_tuning_save")

    def tuning_load(self, path: str) -> None:
        check(self._libh.dakc_tuning_load(self._ctx, str(path).encode()), "dakc_tuning_load")

    # served mode
    def submit(self, reads, *, dev_ptr: int = 0, nbytes: int = 0, **over) -> tuple[int, np.ndarray]:
        """Queue a job; returns (ticket, buffer) -- keep the buffer alive until wait().
        With dev_ptr the reads are already in device memory (nbytes long)."""
        if dev_ptr:
            buf = np.zeros(0, np.uint8)
            job = self.job(dev_ptr, nbytes, on_device=True, **over)
        else:
            buf = _as_bytes(reads)
            job = self.job(buf.ctypes.data if buf.size else 0, int(buf.size), **over)
        tk = C.c_uint64()
        check(self._libh.dakc_submit(self._ctx, C.byref(job), C.byref(tk)), "dakc_submit")
        return int(tk.value), buf

    def serve_stats(self) -> dict:
        st = _lib.ServeStats()
        check(self._libh.dakc_serve_get_stats(self._ctx, C.byref(st)), "dakc_serve_get_stats")
        return st.as_dict()

    def poll(self, ticket: int) -> bool:
        d = C.c_int()
        check(self._libh.dakc_poll(self._ctx, ticket, C.byref(d)), "dakc_poll")
        return bool(d.value)

    def wait_device(self, ticket: int) -> "DeviceTable":
        """Result of a job submitted with table_on_device=True: the table stays
        in device memory until the returned DeviceTable is freed."""
        t = Table()
        check(self._libh.dakc_wait(self._ctx, ticket, C.byref(t)), "dakc_wait")
        return DeviceTable(self, t)

    def wait(self, ticket: int) -> KmerTable:
        t = Table()
        check(self._libh.dakc_wait(self._ctx, ticket, C.byref(t)), "dakc_wait")
        if t.on_device:
            return DeviceTable(self, t).to_host()
        try:
            keys, counts, offs = _lib.table_to_numpy(t)
            return KmerTable(keys, counts, int(t.k), int(t.total), offs)
        finally:
            self._libh.dakc_table_free(C.byref(t))

    # device memory helpers
    # before the count: FASTQ / line input to fixed-width records on the device
    def convert_reads(self, dev_ptr: int, nbytes: int, fmt: int, read_len: int = 0) -> tuple[int, int, int]:
        """(device pointer, bytes, width) of the fixed-width records made from
        FASTQ (fmt=FMT_FASTQ) or line (FMT_LINES) input in device memory, as
        syngen/fq2txtmaker.sh makes them; free the pointer with device_free."""
        n, L = C.c_uint64(), C.c_uint32()
        check(self._libh.dakc_convert_reads(self._ctx, dev_ptr or None, nbytes, fmt, read_len, None, 0,
                                            C.byref(n), C.byref(L)), "dakc_convert_reads")
        out = self.device_alloc(max(int(n.value), 4))
        st = self._libh.dakc_convert_reads(self._ctx, dev_ptr or None, nbytes, fmt, read_len, out,
                                           max(int(n.value), 4), C.byref(n), C.byref(L))
        if st:
            self.device_free(out)
            check(st, "dakc_convert_reads")
        return out, int(n.value), int(L.value)

    # after the count, on a device table (from count_device, or a served
    # DeviceTable)
    def table_histogram(self, t, num_buckets: int) -> tuple[np.ndarray, int]:
        """(hist[c] for c < num_buckets, min_bucket) as construct_kmer_histogram
        computes them (pakman_star/src/kmer_histogram.cpp:81-116)."""
        raw = t.t if isinstance(t, DeviceTable) else t
        hist = np.zeros(max(num_buckets, 1), dtype=np.uint64)
        m = C.c_uint32()
        check(self._libh.dakc_table_histogram(self._ctx, C.byref(raw), num_buckets, hist.ctypes.data,
                                              C.byref(m)), "dakc_table_histogram")
        return hist[:num_buckets], int(m.value)

    def table_filter(self, t, min_count: int) -> None:
        """Keep the entries with count >= min_count, in place."""
        raw = t.t if isinstance(t, DeviceTable) else t
        check(self._libh.dakc_table_filter(self._ctx, C.byref(raw), min_count), "dakc_table_filter")
        if isinstance(t, DeviceTable):
            t.n, t.total = int(raw.n), int(raw.total)

    def table_query(self, t, queries: np.ndarray | None = None, *, dev_queries: int = 0, nq: int = 0,
                    dev_out: int = 0) -> np.ndarray | None:
        """Counts of host queries ((nq,) or (nq, key_words) u64; returns an
        array), or of nq device queries into dev_out (returns None)."""
        raw = t.t if isinstance(t, DeviceTable) else t
        if dev_queries:
            check(self._libh.dakc_table_query(self._ctx, C.byref(raw), dev_queries, nq, 1, dev_out),
                  "dakc_table_query")
            return None
        q = np.ascontiguousarray(queries, dtype=np.uint64)
        n = q.size // int(raw.key_words)
        out = np.zeros(n, dtype=np.uint32)
        check(self._libh.dakc_table_query(self._ctx, C.byref(raw), q.ctypes.data if n else None, n, 0,
                                          out.ctypes.data if n else None), "dakc_table_query")
        return out

    def device_alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        check(self._libh.dakc_device_alloc(self._ctx, nbytes, C.byref(p)), "dakc_device_alloc")
        return int(p.value)

    def device_free(self, ptr: int) -> None:
        check(self._libh.dakc_device_free(self._ctx, C.c_void_p(ptr)), "dakc_device_free")

    def generate_device(self, p: _lib.GenParams, dev_ptr: int, cap: int) -> None:
        check(self._libh.dakc_generate_reads_device(self._ctx, C.byref(p), C.c_void_p(dev_ptr), cap),
              "dakc_generate_reads_device")

    def d2h(self, dev_ptr: int, nbytes: int, dtype=np.uint8) -> np.ndarray:
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        if nbytes:
            check(self._libh.dakc_memcpy_d2h(self._ctx, out.ctypes.data_as(C.c_void_p),
                                             C.c_void_p(dev_ptr), nbytes), "dakc_memcpy_d2h")
        return out

    def d2d(self, dst: int, src: int, nbytes: int) -> None:
        check(self._libh.dakc_memcpy_d2d(self._ctx, C.c_void_p(dst), C.c_void_p(src), nbytes),
              "dakc_memcpy_d2d")

    def h2d(self, dev_ptr: int, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes:
            check(self._libh.dakc_memcpy_h2d(self._ctx, C.c_void_p(dev_ptr),
                                             arr.ctypes.data_as(C.c_void_p), arr.nbytes),
                  "dakc_memcpy_h2d")


class DeviceTable:
    """A served result kept in device memory (keys/counts are device pointers)."""

    def __init__(self, kc: KmerCounter, t: Table):
        self._kc, self.t = kc, t
        self.n, self.total, self.key_words = int(t.n), int(t.total), int(t.key_words)

    def to_host(self) -> KmerTable:
        t = self.t
        kw, n = self.key_words, self.n
        keys = self._kc.d2h(_ptr(t.keys), n * kw * 8, np.uint64).reshape(n, kw)
        counts = self._kc.d2h(_ptr(t.counts), n * 4, np.uint32)
        offs = None
        if t.bin_offsets:
            offs = np.ctypeslib.as_array(t.bin_offsets, shape=(int(t.num_bins) + 1,)).copy()
        res = KmerTable(keys, counts, int(t.k), self.total, offs)
        self.free()
        return res

    def free(self) -> None:
        if self.t is not None:
            self._kc._libh.dakc_table_free(C.byref(self.t))
            self.t = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(p) -> int:
    return C.cast(p, C.c_void_p).value or 0


class Batch:
    """A captured multi-job count: run() replays the hipGraph, tables() returns
    one KmerTable per read set."""

    def __init__(self, kc: KmerCounter, handle, njobs: int, keep):
        self._kc, self._h, self.njobs, self._keep = kc, handle, njobs, keep

    def info(self) -> tuple[int, int]:
        """(jobs counted by the captured launch, device bytes held)."""
        n, b = C.c_size_t(), C.c_uint64()
        check(self._kc._libh.dakc_batch_info(self._h, C.byref(n), C.byref(b)), "dakc_batch_info")
        return int(n.value), int(b.value)

    def run(self) -> float:
        ms = C.c_double()
        check(self._kc._libh.dakc_batch_run(self._h, C.byref(ms)), "dakc_batch_run")
        return ms.value

    def tables(self) -> list[KmerTable]:
        outs = (Table * self.njobs)()
        st = self._kc._libh.dakc_batch_tables(self._h, outs)
        try:
            check(st, "dakc_batch_tables")
            res = []
            for t in outs:
                keys, counts, offs = _lib.table_to_numpy(t)
                res.append(KmerTable(keys, counts, int(t.k), int(t.total), offs))
            return res
        finally:
            for t in outs:
                self._kc._libh.dakc_table_free(C.byref(t))

    def close(self) -> None:
        if self._h:
            self._kc._libh.dakc_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def gpu_available() -> bool:
    """True when a GPU context can be created (does not initialise torch)."""
    try:
        lib = _lib.load()
    except (FileNotFoundError, OSError):
        return False
    ctx = C.c_void_p()
    st = lib.dakc_ctx_create(int(os.environ.get("DAKC_DEVICE", "0")), None, C.byref(ctx))
    if st == 0:
        lib.dakc_ctx_destroy(ctx)
        return True
    return False
