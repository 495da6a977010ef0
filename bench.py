"""Benchmark of the k-mer counting hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload genome|genome_k55|iid]

A step is one full count (decode, extract, radix sort, run-length reduce) of
the genome-like workload: a 100 Mbp uniform-ACGT genome, 20,000,000 reads of
150 bp sampled at 30x from either strand with 0.5% substitutions, k = 31,
forward k-mers in one u64, counts u32.  The reads are generated on the device
and are resident in HBM before timing starts; the table stays in HBM.

Prints one JSON line: k-mers counted per second (whole job), the roofline of
the dominant kernel measured with HIP events on the stream it runs on, and
beside them (none of it in the timed region):
  parity       the whole table against the oracle's threaded count of the same
               reads regenerated on the host (k-mers, distinct, checksum);
  cpu_baseline the oracle's DAKC P-PE shape on every host core this process
               may use, and on one, over a bounded sample of the workload;
  batched      config 4's read sets (a bounded seeded subset, and the 128
               smallest): one captured multi-job launch vs plain per-set counts;
  served       Poisson-arrival requests (k in {21,31,55}): p50/p99 latency and
               k-mers/s of continuous batching vs one job at a time.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # genome-like single job: 1e8 bp genome, 30x of 150 bp reads, 0.5% errors, k=31
    "genome": dict(seed=2025, num_reads=20_000_000, read_len=150, genome_len=100_000_000,
                   sub_rate=0.005, k=31),
    # the same reads at k=55: two-word keys (1.92e9 k-mers)
    "genome_k55": dict(seed=2025, num_reads=20_000_000, read_len=150, genome_len=100_000_000,
                       sub_rate=0.005, k=55),
    # plain single job: 1M i.i.d. reads x 150, canonical k=31
    "iid": dict(seed=2024, num_reads=1_000_000, read_len=150, genome_len=0, sub_rate=0.0, k=31,
                canonical=True),
}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/traffic.json), or None."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    tag = {"genome": "genome_k31", "genome_k55": "genome_k55", "iid": "iid_k31"}.get(workload)
    data = json.loads(f.read_text()).get(tag, {}).get("kernels", {})
    for name, v in data.items():
        if f"dk::{kernel}<" in name:
            return v["traffic_bytes"]
    return None


def host_cpus() -> dict:
    """What the host offers this process: logical CPUs, the ones it may run on,
    the thread share the environment sets (OMP_NUM_THREADS), the CPU model."""
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(share)) if share and share.isdigit() and int(share) > 0 else affinity
    return {"nproc": os.cpu_count(), "affinity": affinity, "omp_num_threads": share, "model": model,
            "threads": threads}


def _timed_runs(fn, runs: int = 5):
    """1 untimed warm-up, then `runs` timed; (median s, min s, last result)."""
    fn()
    times, res = [], None
    for _ in range(runs):
        res = fn()
        times.append(res[-1])
    return float(np.median(times)), float(min(times)), res


def cpu_baseline(w: dict, sample_reads: int, threads: int, t1_reads: int) -> dict:
    """The oracle's DAKC P-PE shape (T workers = T PEs over shared-memory owner
    bins) on the host cores, over the first reads of the same workload: T =
    every core this process may use, and T = 1.  The timer covers the count
    phases only (file read excluded, as kcounter.cpp:262)."""
    import oracle
    from dakc_amd import _lib

    can = w.get("canonical", False)
    p = _lib.gen_params(w["seed"], sample_reads, w["read_len"], genome_len=w["genome_len"],
                        sub_rate=w["sub_rate"], min_len=w.get("min_len", 0))
    reads = _lib.generate_reads(p, threads=threads)
    run = lambda r, T: oracle.count_mt(r, w["read_len"], w["k"], T, canonical=can)  # noqa: E731
    med, mn, (n, d, cs, _) = _timed_runs(lambda: run(reads, threads))
    one = reads[:t1_reads * (w["read_len"] + 1)]
    med1, mn1, (n1, _, _, _) = _timed_runs(lambda: run(one, 1))
    cpus = host_cpus()
    return {"value": n / med, "unit": "k-mers/s", "cores": threads, "kind": "port",
            "sample": f"first {sample_reads} reads of the same workload ({n} k-mers), oracle P-PE shape "
                      f"(P = T = {threads} threads); median of 5 timed runs after 1 warm-up",
            "median_s": med, "min_s": mn, "min_value": n / mn,
            "t1": {"value": n1 / med1, "min_value": n1 / mn1, "median_s": med1, "min_s": mn1,
                   "sample": f"first {t1_reads} reads ({n1} k-mers), one thread"},
            "host": cpus}


def full_parity(kc, t, w: dict, dreads: int, nbytes: int, threads: int) -> dict:
    """Outside the timed region: the GPU's device table `t` of the whole
    workload against the oracle's threaded count of the same reads, made again
    on the host by the same counter-based generator (the bytes are compared
    with the device's first).  Compared by (k-mers, distinct, order-independent
    checksum of the (kmer, count) multiset)."""
    import oracle
    from dakc_amd import _lib
    from dakc_amd.kcounter import _ptr
    t0 = time.perf_counter()
    kb = 8 * int(t.key_words)
    d = int(t.n)
    keys = kc.d2h(_ptr(t.keys), d * kb, np.uint64).reshape(d, kb // 8)
    counts = kc.d2h(_ptr(t.counts), d * 4, np.uint32)
    if kb == 8:
        ascending = bool(np.all(keys[1:, 0] > keys[:-1, 0]))
    else:
        hi, lo = keys[:, 1], keys[:, 0]
        ascending = bool(np.all((hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1]))))
    gpu = {"n": int(t.total), "d": d, "checksum": oracle.checksum(keys, counts),
           "strictly_ascending": d < 2 or ascending, "counts_ge_1": bool(d == 0 or counts.min() >= 1)}
    del keys, counts
    p = _lib.gen_params(w["seed"], w["num_reads"], w["read_len"], genome_len=w["genome_len"],
                        sub_rate=w["sub_rate"], min_len=w.get("min_len", 0))
    host = _lib.generate_reads(p, threads=threads)
    same_input = host.size == nbytes and np.array_equal(kc.d2h(dreads, nbytes, np.uint8), host)
    n, dd, cs, el = oracle.count_mt(host, w["read_len"], w["k"], threads, canonical=w.get("canonical", False))
    del host
    cpu = {"n": n, "d": dd, "checksum": cs, "count_s": el}
    ok = (gpu["n"], gpu["d"], gpu["checksum"]) == (n, dd, cs) and gpu["strictly_ascending"] and gpu["counts_ge_1"]
    return {"equal": bool(same_input and ok),
            "input_bytes_equal": bool(same_input), "gpu": gpu, "cpu_oracle": cpu, "cpu_threads": threads,
            "what": "whole workload: GPU table (n, d, checksum, ascending keys) vs the oracle's threaded "
                    "count of the same reads regenerated on the host",
            "seconds": time.perf_counter() - t0}


def pre_count(kc, dreads: int, w: dict, nreads: int = 4_000_000) -> dict:
    """FASTQ of the first `nreads` reads of the workload (header "@r", the read,
    "+", a quality line), converted on the device to fixed-width records (one
    untimed call, then one timed); the records must equal the generator's."""
    from dakc_amd import _lib
    L = w["read_len"]
    nreads = min(nreads, w["num_reads"])
    fixed = kc.d2h(dreads, nreads * (L + 1), np.uint8).reshape(nreads, L + 1)
    rec = np.empty((nreads, 3 + (L + 1) + 2 + (L + 1)), dtype=np.uint8)
    rec[:, 0:3] = np.frombuffer(b"@r\n", np.uint8)
    rec[:, 3:4 + L] = fixed
    rec[:, 4 + L:6 + L] = np.frombuffer(b"+\n", np.uint8)
    rec[:, 6 + L:6 + 2 * L] = ord("F")
    rec[:, 6 + 2 * L] = ord("\n")
    fq = rec.reshape(-1)
    dfq = kc.device_alloc(fq.size)
    kc.h2d(dfq, fq)
    out, n, width = kc.convert_reads(dfq, fq.size, _lib.FMT_FASTQ)
    kc.device_free(out)
    t0 = time.perf_counter()
    out, n, width = kc.convert_reads(dfq, fq.size, _lib.FMT_FASTQ)
    s = time.perf_counter() - t0
    same = width == L and n == fixed.size and np.array_equal(kc.d2h(out, n, np.uint8), fixed.reshape(-1))
    kc.device_free(out)
    kc.device_free(dfq)
    return {"format": "fastq", "reads": nreads, "in_bytes": int(fq.size), "out_bytes": int(n), "ms": s * 1e3,
            "GBps_in_plus_out": (fq.size + n) / s / 1e9, "records_equal_generator": bool(same),
            "timing": "wall clock around the synchronous call (newline count, offsets, longest read, "
                      "records), one run, not in the timed counts"}


def post_count(kc, t, kb: int) -> dict:
    """After the count, on its device table `t` (once each, after one untimed
    call): the 64-bucket count histogram, point queries (5M keys taken from
    the table, 5M random keys) and the solid-k-mer filter at the histogram's
    minimum bucket, which changes the table and so runs last."""
    from dakc_amd.kcounter import _ptr
    d = int(t.n)
    kc.table_histogram(t, 64)
    t0 = time.perf_counter()
    hist, m = kc.table_histogram(t, 64)
    h_s = time.perf_counter() - t0
    rng = np.random.default_rng(11)
    blocks, per = 100, 50_000
    present = [kc.d2h(_ptr(t.keys) + int(s) * kb, per * kb, np.uint64)
               for s in np.linspace(0, max(d - per, 0), blocks).astype(np.int64)] if d >= per else []
    kw = kb // 8
    q = np.concatenate([np.concatenate(present).reshape(-1, kw) if present else np.zeros((0, kw), np.uint64),
                        rng.integers(0, 2**62, size=(blocks * per, kw), dtype=np.uint64)])
    dq, do = kc.device_alloc(q.nbytes), kc.device_alloc(q.shape[0] * 4)
    kc.h2d(dq, q)
    kc.table_query(t, dev_queries=dq, nq=q.shape[0], dev_out=do)
    t0 = time.perf_counter()
    kc.table_query(t, dev_queries=dq, nq=q.shape[0], dev_out=do)
    q_s = time.perf_counter() - t0
    hits = int((kc.d2h(do, q.shape[0] * 4, np.uint32) > 0).sum())
    kc.device_free(dq)
    kc.device_free(do)
    t0 = time.perf_counter()
    kc.table_filter(t, m)
    f_s = time.perf_counter() - t0
    kept = int(t.n)
    return {"histogram": {"buckets": 64, "ms": h_s * 1e3, "alg_bytes": d * 4,
                          "GBps": d * 4 / h_s / 1e9, "min_bucket": m},
            "query": {"nq": int(q.shape[0]), "hits": hits, "ms": q_s * 1e3,
                      "queries_per_s": q.shape[0] / q_s},
            "filter": {"min_count": m, "kept": kept, "ms": f_s * 1e3,
                       "alg_bytes": (d + kept) * (kb + 4), "GBps": (d + kept) * (kb + 4) / f_s / 1e9},
            "timing": "wall clock around each synchronous call (one run, not in the timed counts)"}


def _pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


def _table_checksum(kc, t) -> tuple[int, int, int]:
    """(k-mers, distinct, order-independent checksum) of a device or host table."""
    import oracle
    from dakc_amd.kcounter import DeviceTable, KmerTable, _ptr
    if isinstance(t, KmerTable):
        return int(t.total), len(t), oracle.checksum(t.keys, t.counts)
    raw = t.t if isinstance(t, DeviceTable) else t
    d, kw = int(raw.n), int(raw.key_words)
    keys = kc.d2h(_ptr(raw.keys), d * kw * 8, np.uint64).reshape(d, kw)
    counts = kc.d2h(_ptr(raw.counts), d * 4, np.uint32)
    return int(raw.total), d, oracle.checksum(keys, counts)


def batch_sets(cap_reads: float, nsets: int = 256):
    """Config 4's 256 read sets (sizes log-uniform 1e4-5e6 reads, seeded), taken
    in seeded order while they fit `cap_reads` reads, and the 128 smallest."""
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(5e6), nsets)).astype(np.int64)
    chosen, tot = [], 0
    for i, nr in enumerate(sizes):
        if tot + nr > cap_reads and chosen:
            continue
        chosen.append((i, int(nr)))
        tot += int(nr)
    small = sorted(((i, int(nr)) for i, nr in enumerate(sizes)), key=lambda x: x[1])[:nsets // 2]
    return chosen, small


def _batch_legs(kc, sets, steps, warmup, tune_trials=3):
    """Device-resident read sets: autotune (one set per factor of 2 in size),
    then plain per-set counts vs one captured multi-job launch; every batched
    table checked against its plain count by checksum."""
    from dakc_amd import _lib
    ptrs, nbs = [], []
    for i, nr in sets:
        p = _lib.gen_params(7000 + i, nr, 250, min_len=100, sub_rate=0.005)
        nb = nr * 251
        d = kc.device_alloc(nb)
        kc.generate_device(p, d, nb)
        ptrs.append(d)
        nbs.append(nb)
    classes = {}
    for d, nb in zip(ptrs, nbs):  # the median set of each class represents it
        classes.setdefault(int(np.log2(nb / 251)), []).append((nb, d))
    for members in classes.values():
        nb, d = sorted(members)[len(members) // 2]
        kc.autotune(dev_ptr=d, nbytes=nb, trials=tune_trials)
    # plain: one count per set, in turn (the largest set first, untimed, so
    # the workspace is at its full size); best of two passes
    big = int(np.argmax(nbs))
    kc.count_device(ptrs[big], nbs[big])
    plain_s = 1e30
    for _ in range(2):
        t0 = time.perf_counter()
        for d, nb in zip(ptrs, nbs):
            kc.count_device(d, nb)
        plain_s = min(plain_s, time.perf_counter() - t0)
    plain = [_table_checksum(kc, kc.count_device(d, nb)) for d, nb in zip(ptrs, nbs)]  # untimed
    kmers = sum(x[0] for x in plain)
    b = kc.batch_create(ptrs, on_device=True, sizes=nbs)
    in_launch, held = b.info()
    for _ in range(warmup):
        b.run()
    t0 = time.perf_counter()
    dev_ms = [b.run() for _ in range(steps)]
    batch_s = (time.perf_counter() - t0) / steps
    tabs = b.tables()
    same = all(_table_checksum(kc, t) == x for t, x in zip(tabs, plain))
    b.close()
    for d in ptrs:
        kc.device_free(d)
    return dict(sets=len(sets), reads=int(sum(nr for _, nr in sets)), kmers=kmers, plain_s=plain_s,
                batch_s=batch_s, device_ms=float(np.mean(dev_ms)), in_launch=in_launch, held=held,
                same=same)


def run_batched(steps: int, warmup: int, cap_reads: float, with_small: bool = True, tune_trials: int = 3) -> dict:
    from dakc_amd import KmerCounter
    chosen, small = batch_sets(cap_reads)
    kc = KmerCounter(k=31, read_len=250)
    print(f"[bench] batch: {len(chosen)} read sets, {sum(n for _, n in chosen)} reads", file=sys.stderr,
          flush=True)
    r = _batch_legs(kc, chosen, steps, warmup, tune_trials)
    rs = None
    if with_small:
        print(f"[bench] batch: small half, {len(small)} sets", file=sys.stderr, flush=True)
        rs = _batch_legs(kc, small, steps, warmup, tune_trials)
    import tempfile
    with tempfile.NamedTemporaryFile("r", suffix=".json") as tf:
        kc.tuning_save(tf.name)
        tuning = json.load(open(tf.name))["entries"]
    kc.close()
    out = {"workload": f"batched: {r['sets']} of the 256 read sets of config 4 (sizes log-uniform 1e4-5e6 "
                       f"reads, taken in seeded order up to {cap_reads:g} reads), lengths 100-250 M-padded "
                       f"to 250, 0.5% subs, k=31; device-resident reads, autotuned geometry",
           "sets": r["sets"], "reads": r["reads"], "kmers": r["kmers"],
           "batched_kmers_per_s": r["kmers"] / r["batch_s"], "batched_ms": r["batch_s"] * 1e3,
           "plain_kmers_per_s": r["kmers"] / r["plain_s"], "plain_ms": r["plain_s"] * 1e3,
           "batched_over_plain": r["plain_s"] / r["batch_s"], "device_ms": r["device_ms"],
           "sets_in_captured_launch": r["in_launch"], "batch_device_bytes": r["held"],
           "tables_match_plain": r["same"],
           "timing": f"plain: best of 2 passes of one count per set; batched: mean of {steps} graph "
                     f"replays after {warmup}; tables compared by checksum of every (kmer, count)",
           "tuning": tuning}
    if rs is not None:
        out["small_sets"] = {"workload": f"the {rs['sets']} smallest of the 256 sets ({rs['reads']} reads)",
                             "batched_kmers_per_s": rs["kmers"] / rs["batch_s"],
                             "batched_ms": rs["batch_s"] * 1e3,
                             "plain_kmers_per_s": rs["kmers"] / rs["plain_s"], "plain_ms": rs["plain_s"] * 1e3,
                             "batched_over_plain": rs["plain_s"] / rs["batch_s"],
                             "sets_in_captured_launch": rs["in_launch"], "tables_match_plain": rs["same"]}
    return out


def run_batched_full(window_kmers: float = 3e9, tune_trials: int = 1, check_sets: int = 16) -> dict:
    """Config 4 at its full size: all 256 read sets (about 2.1e8 reads, 3e10
    k-mers), generated into HBM once.  The tables (about 12 bytes per k-mer:
    these reads have almost no repeated k-mer) outgrow host and device memory
    together, so both legs stream: plain counts one set at a time, batched
    counts windows of sets with dakc_count_batch (device tables of their own,
    chunked inside by the HBM budget) and frees each window's tables.  Every
    table is checked against its plain count by (k-mers, distinct);
    `check_sets` of the sets below 1e6 reads also by checksum of every
    (kmer, count).  Timed: the counting calls only."""
    from dakc_amd import KmerCounter, _lib
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(5e6), 256)).astype(np.int64)
    kc = KmerCounter(k=31, read_len=250)
    ptrs, nbs = [], []
    for i, nr in enumerate(sizes):
        p = _lib.gen_params(7000 + i, int(nr), 250, min_len=100, sub_rate=0.005)
        nb = int(nr) * 251
        d = kc.device_alloc(nb)
        kc.generate_device(p, d, nb)
        ptrs.append(d)
        nbs.append(nb)
    print(f"[bench] batch-full: 256 read sets, {int(sizes.sum())} reads, {sum(nbs) / 1e9:.1f} GB on the device",
          file=sys.stderr, flush=True)
    classes = {}
    for d, nb in zip(ptrs, nbs):
        classes.setdefault(int(np.log2(nb / 251)), []).append((nb, d))
    for members in classes.values():
        nb, d = sorted(members)[len(members) // 2]
        kc.autotune(dev_ptr=d, nbytes=nb, trials=tune_trials)
    small = [i for i in range(256) if sizes[i] < 1_000_000]
    checked = set(small[:: max(1, len(small) // check_sets)][:check_sets])
    # plain: one count per set (the largest first, untimed, sizes the workspace)
    big = int(np.argmax(nbs))
    kc.count_device(ptrs[big], nbs[big])
    plain_s, plain_fp, plain_cs = 0.0, [], {}
    for i, (d, nb) in enumerate(zip(ptrs, nbs)):
        t0 = time.perf_counter()
        t = kc.count_device(d, nb)
        plain_s += time.perf_counter() - t0
        plain_fp.append((int(t.total), int(t.n)))
        if i in checked:
            plain_cs[i] = _table_checksum(kc, t)
    kmers = sum(x[0] for x in plain_fp)
    print(f"[bench] batch-full: plain {plain_s:.2f} s, {kmers} k-mers", file=sys.stderr, flush=True)
    # batched: windows of consecutive sets of about window_kmers k-mer slots
    windows, cur, slots = [], [], 0
    for i, nr in enumerate(sizes):
        cur.append(i)
        slots += int(nr) * 220
        if slots >= window_kmers:
            windows.append(cur)
            cur, slots = [], 0
    if cur:
        windows.append(cur)
    batch_s, same_fp, same_cs = 0.0, True, True
    for w in windows:
        t0 = time.perf_counter()
        tabs = kc.count_batch_device(dev_ptrs=[ptrs[i] for i in w], sizes=[nbs[i] for i in w])
        batch_s += time.perf_counter() - t0
        for i, t in zip(w, tabs):
            same_fp &= (t.total, t.n) == plain_fp[i]
            if i in checked:
                same_cs &= _table_checksum(kc, t) == plain_cs[i]
            t.free()
    print(f"[bench] batch-full: batched {batch_s:.2f} s in {len(windows)} windows", file=sys.stderr, flush=True)
    for d in ptrs:
        kc.device_free(d)
    kc.close()
    return {"workload": "config 4 at full size: all 256 read sets (sizes log-uniform 1e4-5e6 reads, seeded), "
                        "lengths 100-250 M-padded to 250, 0.5% subs, k=31; device-resident reads, autotuned geometry",
            "sets": 256, "reads": int(sizes.sum()), "kmers": kmers, "windows": len(windows),
            "window_kmer_slots": window_kmers,
            "batched_kmers_per_s": kmers / batch_s, "batched_s": batch_s,
            "plain_kmers_per_s": kmers / plain_s, "plain_s": plain_s,
            "batched_over_plain": plain_s / batch_s,
            "tables_match_plain": bool(same_fp and same_cs),
            "checked": f"(k-mers, distinct) of all 256 tables; checksum of every (kmer, count) of "
                       f"{len(checked)} sets below 1e6 reads",
            "timing": "wall clock around the counting calls only: plain, one dakc_count_device per set; "
                      "batched, one dakc_count_batch per window with device tables (freed after the checks)"}


def bench_batch(a) -> int:
    """Many read sets of uneven size: plain per-set launches vs one captured
    multi-job launch (autotuned), same device-resident inputs."""
    if a.batch_full:
        r = run_batched_full()
        out = {"metric": "k-mers counted/s (batched, config 4 at full size)", "value": r["batched_kmers_per_s"],
               "unit": "k-mers/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": r["batched_s"] * 1e3,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
               "data": "synthetic (seeded, generated on the device)",
               "config": {"workload": r["workload"], "reads": r["reads"], "kmers_per_step": r["kmers"],
                          "parallelism": "single GPU"},
               "batched_full": r}
        print(json.dumps(out), flush=True)
        return 0
    r = run_batched(a.steps, a.warmup, a.batch_reads_cap)
    out = {"metric": "k-mers counted/s (batched)", "value": r["batched_kmers_per_s"], "unit": "k-mers/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": r["batched_ms"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
           "data": "synthetic (seeded, generated on the device)",
           "config": {"workload": r["workload"], "reads": r["reads"], "kmers_per_step": r["kmers"],
                      "parallelism": "single GPU"},
           "batched": r}
    if not a.no_cpu:
        threads = a.cpu_threads or host_cpus()["threads"]
        w = dict(seed=7000, read_len=250, min_len=100, genome_len=0, sub_rate=0.005, k=31)
        out["cpu_baseline"] = cpu_baseline(w, 400_000, threads, 100_000)
    print(json.dumps(out), flush=True)
    return 0


def run_served(nreq: int, slots: int, max_large: int, load: float = 0.75, check: int = 8) -> dict:
    """Poisson-arrival counting requests: continuous batching (dakc_submit /
    dakc_poll / dakc_wait, jobs overlapping on the device) vs one job at a time
    in arrival order; latency = arrival -> table ready.  Every served table's
    (k-mers, distinct) equals its count alone; `check` of them are compared by
    checksum of every entry."""
    import threading
    from dakc_amd import KmerCounter, _lib
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(1e7), nreq)).astype(np.int64)
    ks = rng.choice([21, 31, 55], nreq)
    kc = KmerCounter(read_len=150, max_inflight=slots, max_large=max_large)
    reqs = []
    for i in range(nreq):
        p = _lib.gen_params(9000 + i, int(sizes[i]), 150, genome_len=50_000_000, sub_rate=0.005)
        nb = int(sizes[i]) * 151
        d = kc.device_alloc(nb)
        kc.generate_device(p, d, nb)
        reqs.append((d, nb, int(ks[i])))
    # service times alone (also the per-request reference totals); both forms
    # take device-resident reads and leave the table in device memory.  The
    # largest request of each k first, so the plain workspace is at its full
    # size before anything is timed.
    for kk in sorted(set(int(x) for x in ks)):
        d, nb, k = max((r for r in reqs if r[2] == kk), key=lambda r: r[1])
        kc.count_device(d, nb, k=k)
    service, totals = [], []
    checked = set(int(i) for i in np.random.default_rng(6).choice(nreq, size=min(check, nreq), replace=False))
    sums = {}
    for i, (d, nb, k) in enumerate(reqs):
        t0 = time.perf_counter()
        t = kc.count_device(d, nb, k=k)
        service.append(time.perf_counter() - t0)
        totals.append((int(t.total), int(t.n)))
        if i in checked:
            sums[i] = _table_checksum(kc, t)
    # served warm-up: the first submit allocates the pool
    for d, nb, k in reqs[:4]:
        kc.wait_device(kc.submit(None, dev_ptr=d, nbytes=nb, k=k, table_on_device=True)[0]).free()
    kmers = sum(x[0] for x in totals)
    print(f"[bench] serve: {nreq} requests, {kmers} k-mers, alone {sum(service):.3f} s",
          file=sys.stderr, flush=True)
    lam = load / float(np.mean(service))  # `load` of one-at-a-time capacity
    arrivals = np.cumsum(rng.exponential(1.0 / lam, nreq))
    # plain: one at a time, in arrival order
    start = time.perf_counter()
    lat_plain = []
    for (d, nb, k), at in zip(reqs, arrivals):
        now = time.perf_counter() - start
        if now < at:
            time.sleep(at - now)
        kc.count_device(d, nb, k=k)
        lat_plain.append(time.perf_counter() - start - at)
    plain_span = time.perf_counter() - start
    print(f"[bench] serve: plain leg {plain_span:.3f} s", file=sys.stderr, flush=True)
    # served: submit on arrival, poll for completion
    done_t = [None] * nreq
    tickets = [None] * nreq
    start = time.perf_counter()
    errors = []

    def submitter():
        try:
            for i, ((d, nb, k), at) in enumerate(zip(reqs, arrivals)):
                now = time.perf_counter() - start
                if now < at:
                    time.sleep(at - now)
                tickets[i] = kc.submit(None, dev_ptr=d, nbytes=nb, k=k, table_on_device=True)[0]
        except Exception as e:  # surfaced by the polling loop below
            errors.append(e)

    th = threading.Thread(target=submitter)
    th.start()
    pending = set(range(nreq))
    same = True
    mix = {}
    while pending:
        if errors:
            raise errors[0]
        for i in list(pending):
            if tickets[i] is not None and kc.poll(tickets[i]):
                done_t[i] = time.perf_counter() - start
                pending.discard(i)
                t = kc.wait_device(tickets[i])  # done: returns at once
                same = same and (t.total, t.n) == totals[i]
                if i in checked:
                    mix[i] = t  # compared after the timed leg
                else:
                    t.free()
        time.sleep(0.0002)
    th.join()
    served_span = max(done_t)
    same = same and all(_table_checksum(kc, mix[i]) == sums[i] for i in checked)
    for t in mix.values():
        t.free()
    print(f"[bench] serve: served leg {served_span:.3f} s", file=sys.stderr, flush=True)
    lat_served = [done_t[i] - arrivals[i] for i in range(nreq)]
    # saturated: every request queued at once (closed loop), makespan
    t0 = time.perf_counter()
    tks = [kc.submit(None, dev_ptr=d, nbytes=nb, k=k, table_on_device=True)[0] for d, nb, k in reqs]
    for i, tk in enumerate(tks):
        t = kc.wait_device(tk)
        same = same and (t.total, t.n) == totals[i]
        t.free()
    sat_served = time.perf_counter() - t0
    sat_plain = float(np.sum(service))
    sstats = kc.serve_stats()
    for d, nb, k in reqs:
        kc.device_free(d)
    kc.close()
    return {"workload": f"served: {nreq} Poisson-arrival requests at {load:.0%} of one-at-a-time capacity, "
                        "sizes log-uniform 1e4-1e7 reads x 150 bp (genome-like, 0.5% subs), k uniform in "
                        "{21,31,55}; device-resident reads, device tables", "requests": nreq, "kmers": kmers,
            "arrival_rate_per_s": lam, "slots": slots, "max_large": max_large,
            "served": {"p50_ms": _pct(lat_served, 50) * 1e3, "p99_ms": _pct(lat_served, 99) * 1e3,
                       "kmers_per_s": kmers / served_span, "span_s": served_span},
            "plain": {"p50_ms": _pct(lat_plain, 50) * 1e3, "p99_ms": _pct(lat_plain, 99) * 1e3,
                      "kmers_per_s": kmers / plain_span, "span_s": plain_span},
            "saturated": {"served_kmers_per_s": kmers / sat_served, "plain_kmers_per_s": kmers / sat_plain},
            "tables_match_alone": same,
            "tables_checked_by_checksum": len(checked), "serve_stats": sstats}


def bench_serve(a) -> int:
    r = run_served(a.serve_requests, a.serve_slots, a.serve_max_large)
    out = {"metric": "served p50/p99 latency and k-mers/s", "value": r["served"]["kmers_per_s"],
           "unit": "k-mers/s", "n_gpus": 1, "steps": r["requests"], "warmup": 4,
           "ms_per_step": r["served"]["span_s"] * 1e3 / r["requests"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u64",
           "data": "synthetic (seeded, generated on the device)",
           "config": {"workload": r["workload"], "kmers": r["kmers"],
                      "parallelism": f"single GPU, {a.serve_slots} job slots, "
                                     f"{a.serve_max_large} GPU-filling job(s) at once"},
           "served_vs_plain": r}
    print(json.dumps(out), flush=True)
    return 0


# config 3: one PE's share of a 64-PE run -- a 2 Gbp genome, 250 bp reads at
# 20x (1.6e8 reads) with 1% substitutions and 0.1% of bases in N runs (runs of
# 1-10 bases: start probability 0.001 / 5.5), k = 63, P = 64
PE64 = dict(seed=3033, num_reads=160_000_000, read_len=250, genome_len=2_000_000_000, sub_rate=0.01,
            n_rate=0.001 / 5.5, n_run_max=10, k=63, num_pes=64)


def _pe64_params(first: int, count: int):
    from dakc_amd import _lib
    w = PE64
    return _lib.gen_params(w["seed"], count, w["read_len"], genome_len=w["genome_len"], sub_rate=w["sub_rate"],
                           n_rate=w["n_rate"], n_run_max=w["n_run_max"], first_read=first)


def hitter_leg(kc, dreads: int, nbytes: int, table_checksum=None, **job) -> dict:
    """HITTER=1 on device reads (once, after one untimed call): what leaves the
    send side with and without the per-bucket pre-aggregation; light + heavy
    checked against the plain table (sizes, and the additive checksum of every
    entry) when its checksum is given."""
    kc.count_hitter(dev_ptr=dreads, nbytes=nbytes, table_on_device=True, **job)
    t0 = time.perf_counter()
    lt, ht, st = kc.count_hitter(dev_ptr=dreads, nbytes=nbytes, table_on_device=True, **job)
    el = time.perf_counter() - t0
    out = {"bucket_size": 10000, "stats": st, "ms": el * 1e3, "light": int(lt.n), "heavy": int(ht.n),
           "bytes_ratio": st["bytes_sent"] / max(st["bytes_without"], 1),
           "timing": "wall clock around one synchronous call (bucket chain, record-order k-mers, per-bucket "
                     "counts, both owner-side counts and the merge), not in the timed counts"}
    if table_checksum is not None:
        n, d, cs = table_checksum
        _, dl, cl = _table_checksum(kc, lt)
        _, dh, ch = _table_checksum(kc, ht)
        out["union_is_plain_table"] = bool(dl + dh == d and (cl + ch) % 2**64 == cs
                                           and int(lt.total) + int(ht.total) == n)
    from dakc_amd import _lib
    for t in (lt, ht):
        kc._libh.dakc_table_free(_lib.C.byref(t))
    return out


def bench_pe64(a) -> int:
    """Config 3 at its real shape on one GPU: (i) PE 0's Phase 1, its 1/64 of
    the reads partitioned into the 64 owner bins; (ii) PE 0's Phase 2, owner
    bin 0 of all 1.6e8 reads (every sender's k-mers for PE 0, selected from
    device-generated chunks) counted; (iii) what HITTER saves PE 0's send side.
    Parity: two bins of (i) against the oracle's extraction, element for
    element; (ii)'s table against the oracle's streamed bin-0 count by checksum."""
    import oracle
    from dakc_amd import KmerCounter, _lib
    w = PE64
    L, k, P = w["read_len"], w["k"], w["num_pes"]
    threads = a.cpu_threads or host_cpus()["threads"]
    kc = KmerCounter(k=k, read_len=L, num_pes=P, owner_bin=-1)
    kb = 16
    # (i) PE 0's block of records (fqreader.cpp:43-52 gives PE p the p-th 1/64)
    nr0 = w["num_reads"] // P
    nb0 = nr0 * (L + 1)
    d0 = kc.device_alloc(nb0)
    kc.generate_device(_pe64_params(0, nr0), d0, nb0)
    n0, _ = kc.partition(dev_ptr=d0, nbytes=nb0)
    dk = kc.device_alloc(n0 * kb)
    kc.partition(dev_ptr=d0, nbytes=nb0, keys_out=dk, cap=n0)  # warm-up
    times = []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        n0b, offs = kc.partition(dev_ptr=d0, nbytes=nb0, keys_out=dk, cap=n0)
        times.append(time.perf_counter() - t0)
    part_s = float(np.median(times))
    host0 = _lib.generate_reads(_pe64_params(0, nr0), threads=threads)
    same_in = np.array_equal(kc.d2h(d0, nb0, np.uint8), host0)
    ok = same_in and n0b == n0
    for b in (0, 37):
        got = kc.d2h(dk + int(offs[b]) * kb, int(offs[b + 1] - offs[b]) * kb, np.uint64).reshape(-1, 2)
        try:
            want = oracle.extract_mt(host0, L, k, threads, False, P, b, cap=got.shape[0] + 1)
        except oracle.OracleError:  # the oracle's bin is larger than the GPU's
            want = None
        ok = ok and want is not None and np.array_equal(got, want)
    del host0
    part = {"workload": f"PE 0's {nr0} reads (1/64 of config 3) partitioned into 64 owner bins, k=63",
            "kmers": n0, "ms": part_s * 1e3, "kmers_per_s": n0 / part_s,
            "alg_bytes": 2 * nb0 + n0 * kb, "GBps": (2 * nb0 + n0 * kb) / part_s / 1e9,
            "bins_0_and_37_equal_oracle": bool(ok),
            "timing": f"median of {a.steps} synchronous calls (histogram pass, host size check, sweep)"}
    print(f"[bench] pe64: partition {n0} k-mers in {part_s * 1e3:.1f} ms", file=sys.stderr, flush=True)
    hit = hitter_leg(kc, d0, nb0, None, k=k, read_len=L, num_pes=P, owner_bin=-1)
    kc.device_free(dk)
    kc.device_free(d0)
    # (ii) bin 0 of all reads: chunks generated on the device, bin 0 kept
    chunk = 10_000_000
    nchunks = (w["num_reads"] + chunk - 1) // chunk
    cap0 = int(w["num_reads"] * (L - k + 1) / P * 1.1) + (1 << 20)
    dbin = kc.device_alloc(cap0 * kb)
    dch = kc.device_alloc(chunk * (L + 1))
    n_bin, filter_s, gen_s = 0, 0.0, 0.0
    for c in range(nchunks):
        nr = min(chunk, w["num_reads"] - c * chunk)
        nb = nr * (L + 1)
        t0 = time.perf_counter()
        kc.generate_device(_pe64_params(c * chunk, nr), dch, nb)
        t1 = time.perf_counter()
        m, _ = kc.partition(dev_ptr=dch, nbytes=nb, keys_out=dbin + n_bin * kb, cap=cap0 - n_bin, owner_bin=0)
        filter_s += time.perf_counter() - t1
        gen_s += t1 - t0
        n_bin += m
    kc.device_free(dch)
    print(f"[bench] pe64: bin 0 holds {n_bin} k-mers", file=sys.stderr, flush=True)
    kc.count_keys(dev_ptr=dbin, n=n_bin, k=k, on_device_table=True)  # warm-up
    times = []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        t = kc.count_keys(dev_ptr=dbin, n=n_bin, k=k, on_device_table=True)
        times.append(time.perf_counter() - t0)
    cnt_s = float(np.median(times))
    st = kc.last_stats()
    gpu = _table_checksum(kc, t)
    # the oracle: the same reads regenerated on the host chunk by chunk, bin 0 kept
    t0 = time.perf_counter()
    parts = []
    for c in range(nchunks):
        nr = min(chunk, w["num_reads"] - c * chunk)
        h = _lib.generate_reads(_pe64_params(c * chunk, nr), threads=threads)
        cap = int(nr * (L - k + 1) / P * 1.3) + 65536  # one pass (about 1/64 of the windows)
        try:
            parts.append(oracle.extract_mt(h, L, k, threads, False, P, 0, cap=cap))
        except oracle.OracleError:
            parts.append(oracle.extract_mt(h, L, k, threads, False, P, 0))
        del h
    keys0 = np.concatenate(parts)
    del parts
    dcpu, cscpu, cnt_cpu_s = oracle.count_keys_mt(keys0, k, threads)
    cpu_s = time.perf_counter() - t0
    ncpu = keys0.shape[0]
    del keys0
    kc.device_free(dbin)
    kc.close()
    bin0 = {"workload": "owner bin 0 of all 1.6e8 reads (every PE's k-mers for PE 0), k=63 as 2 x u64",
            "kmers": n_bin, "distinct": gpu[1], "count_ms": cnt_s * 1e3, "count_kmers_per_s": n_bin / cnt_s,
            "count_alg_GBps": (4 * n_bin * kb + gpu[1] * (kb + 4)) / cnt_s / 1e9,
            "stages_ms": {"l1": st.pass0_ms, "l2": st.part2_ms, "final": st.final_ms, "path": int(st.path)},
            "select_ms": filter_s * 1e3, "generate_ms": gen_s * 1e3,
            "parity": {"equal": bool(gpu == (ncpu, dcpu, cscpu)), "gpu": list(gpu), "cpu_oracle": [ncpu, dcpu, cscpu],
                       "cpu_threads": threads, "cpu_count_s": cnt_cpu_s, "cpu_leg_s": cpu_s},
            "timing": f"count: median of {a.steps} dakc_count_keys calls on the bin-0 keys in HBM; select: "
                      "the bin-0 filter over 16 chunks of reads (the k-mers other PEs would send)"}
    out = {"metric": "k-mers counted/s (one PE's share of a 64-PE run)", "value": n_bin / cnt_s,
           "unit": "k-mers/s", "n_gpus": 1, "steps": a.steps, "warmup": 1, "ms_per_step": cnt_s * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64x2",
           "data": "synthetic (seeded, generated on the device, regenerated on the host for the oracle)",
           "config": {"workload": "config 3: 2 Gbp genome, 250 bp at 20x, 1% subs, 0.1% N runs, k=63, P=64",
                      "parallelism": "single GPU (one PE's share)"},
           "partition": part, "bin0": bin0, "hitter": hit}
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="genome", choices=sorted(WORKLOADS) + ["pe64"],
                    help="pe64: config 3, one PE's share of a 64-PE run (its own line)")
    ap.add_argument("--mode", default="plain", choices=["plain", "batch", "serve"],
                    help="plain: one job per count (the headline); batch: many read sets per "
                         "captured launch vs one launch per set; serve: Poisson-arrival requests, "
                         "continuous batching vs one job at a time")
    ap.add_argument("--batch-full", action="store_true",
                    help="batch mode: config 4 at full size (all 256 read sets, streamed in windows)")
    ap.add_argument("--batch-reads-cap", type=float, default=2e7,
                    help="batch mode: read sets are taken in seeded order until this many reads")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-slots", type=int, default=4, help="serve mode: jobs in flight")
    ap.add_argument("--serve-max-large", type=int, default=1,
                    help="serve mode: GPU-filling jobs in flight at once")
    ap.add_argument("--cpu-sample-reads", type=int, default=1_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-post", action="store_true",
                    help="skip the FASTQ conversion and the after-the-count histogram / query / "
                         "filter measurements")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the whole-workload check of the GPU table against the oracle's threaded count")
    ap.add_argument("--no-batched", action="store_true", help="skip the bounded batched sub-object")
    ap.add_argument("--no-hitter", action="store_true", help="skip the HITTER=1 sub-object")
    ap.add_argument("--no-served", action="store_true", help="skip the bounded served sub-object")
    ap.add_argument("--line-batch-reads", type=float, default=4e6,
                    help="plain mode's batched sub-object: read sets taken up to this many reads")
    ap.add_argument("--line-serve-requests", type=int, default=150,
                    help="plain mode's served sub-object: Poisson requests")
    ap.add_argument("--algo", default="auto", choices=["auto", "lsd"])
    ap.add_argument("--copy-probe", action="store_true",
                    help="also copy the read buffer device-to-device once (a plain copy kernel "
                         "of known size, the scale for rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU's count; --gpus {a.gpus} is not supported in this build "
              "(the path shards by owner hash, one exchange step; see DESIGN.md)", file=sys.stderr)
        return 2

    from dakc_amd import KmerCounter, _lib

    if a.workload == "pe64":
        return bench_pe64(a)
    if a.mode == "batch":
        return bench_batch(a)
    if a.mode == "serve":
        return bench_serve(a)

    w = WORKLOADS[a.workload]
    kc = KmerCounter(k=w["k"], read_len=w["read_len"], canonical=w.get("canonical", False),
                     algo=_lib.ALGO_LSD if a.algo == "lsd" else _lib.ALGO_AUTO)
    p = _lib.gen_params(w["seed"], w["num_reads"], w["read_len"], genome_len=w["genome_len"],
                        sub_rate=w["sub_rate"])
    nbytes = w["num_reads"] * (w["read_len"] + 1)
    dreads = kc.device_alloc(nbytes)
    kc.generate_device(p, dreads, nbytes)
    print(f"[bench] generated {nbytes / 1e9:.2f} GB of reads on the device", file=sys.stderr,
          flush=True)

    if a.copy_probe:
        dcopy = kc.device_alloc(nbytes)
        kc.d2d(dcopy, dreads, nbytes)
        kc.device_free(dcopy)
        print(f"[bench] copy probe: {nbytes} bytes device-to-device", file=sys.stderr, flush=True)

    for i in range(a.warmup):
        t = kc.count_device(dreads, nbytes)
        print(f"[bench] warmup {i}: {t.total} k-mers, {t.n} distinct", file=sys.stderr, flush=True)
    n_kmers = d = 0
    acc = {"hist": 0.0, "pass0": 0.0, "key_passes": 0.0, "rle": 0.0, "part2": 0.0, "final": 0.0}
    sweep_sum = np.zeros(18)
    npasses = path = reruns = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        t = kc.count_device(dreads, nbytes)
        s = kc.last_stats()
        n_kmers, d = int(t.total), int(t.n)
        sweep_sum += np.array(s.sweep_ms[:])
        acc["hist"] += s.hist_ms
        acc["pass0"] += s.pass0_ms
        acc["key_passes"] += s.passes_ms
        acc["rle"] += s.rle_ms
        acc["part2"] += s.part2_ms
        acc["final"] += s.final_ms
        npasses, path, reruns = int(s.npasses), int(s.path), int(s.msd_reruns)
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / a.steps
    value = n_kmers / (ms_per_step / 1e3)
    ms = {k: v / a.steps for k, v in acc.items()}

    # algorithmic bytes per launch of each kernel (compulsory reads + writes):
    # in = record bytes, n = k-mers, d = distinct, w = key bytes, c = 4 (count)
    kb, inb = 8 * (1 if w["k"] <= 32 else 2), nbytes
    kernels = {}
    if path == 2:
        kernels["k_l1"] = (ms["pass0"], inb + n_kmers * kb)
        kernels["k_l2"] = (ms["part2"], 2 * n_kmers * kb)
        kernels["k_fin"] = (ms["final"], n_kmers * kb + d * (kb + 4))
        b_alg = inb + 4 * n_kmers * kb + d * (kb + 4)
    else:
        kernels["k_hist_reads"] = (ms["hist"], inb)
        kernels["k_sweep<from_reads>"] = (ms["pass0"], inb + n_kmers * kb)
        key_pass_ms = sweep_sum[1:npasses] / a.steps
        kernels["k_sweep<keys>"] = (float(key_pass_ms.mean()) if npasses > 1 else 0.0, 2 * n_kmers * kb)
        kernels["k_rle"] = (ms["rle"], n_kmers * kb + d * (kb + 4))
        b_alg = 2 * inb + n_kmers * kb + 2 * (npasses - 1) * n_kmers * kb + n_kmers * kb + d * (kb + 4)
    total_kernel_ms = {name: (t_ms * (npasses - 1) if name == "k_sweep<keys>" else t_ms)
                       for name, (t_ms, _) in kernels.items()}
    dom = max(total_kernel_ms, key=total_kernel_ms.get)
    dom_ms, dom_bytes = kernels[dom]
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9 if dom_ms > 0 else 0.0
    traffic = measured_traffic(a.workload, dom)
    threads = a.cpu_threads or host_cpus()["threads"]
    parity = None
    if not a.no_parity:
        print("[bench] parity: whole table against the oracle's threaded count", file=sys.stderr, flush=True)
        parity = full_parity(kc, t, w, dreads, nbytes, threads)
    hit = None
    if not a.no_hitter:
        print("[bench] hitter: the same reads with per-bucket pre-aggregation", file=sys.stderr, flush=True)
        csum = (parity["gpu"]["n"], parity["gpu"]["d"], parity["gpu"]["checksum"]) if parity else None
        hit = hitter_leg(kc, dreads, nbytes, csum, k=w["k"], read_len=w["read_len"],
                         canonical=w.get("canonical", False))
        t = kc.count_device(dreads, nbytes)  # the HITTER count used the workspace: count again
    post = post_count(kc, t, kb) if not a.no_post else None  # changes the table: after the parity check
    pre = pre_count(kc, dreads, w) if not a.no_post else None
    stages = {
        "path": "msd" if path == 2 else "lsd", "msd_reruns": reruns, "distinct": d,
        "kernels": {name: {"avg_ms": t_ms, "alg_bytes": nb_,
                           "GBps": nb_ / (t_ms / 1e3) / 1e9 if t_ms > 0 else None}
                    for name, (t_ms, nb_) in kernels.items()},
        "whole_job_GBps_alg": b_alg / (ms_per_step / 1e3) / 1e9,
        "B_alg_per_kmer": b_alg / max(n_kmers, 1),
        "B_ref_per_kmer": (inb + n_kmers * kb + 8 * n_kmers * kb) / max(n_kmers, 1),
    }
    out = {
        "metric": "k-mers counted/s",
        "value": value,
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded counter-based generator, generated on the device)",
        "config": {"workload": f"{a.workload}: {w['num_reads']} reads x {w['read_len']} bp, "
                               f"genome {w['genome_len']} bp, {w['sub_rate'] * 100:.1f}% subs, "
                               f"k={w['k']}" + (" canonical" if w.get("canonical") else ""),
                   "k": w["k"], "kmers_per_step": n_kmers,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                     "avg_ms": dom_ms, "alg_bytes_per_launch": dom_bytes},
        "stages": stages,
        "parity": parity,
        "hitter": hit,
        "pre_count": pre,
        "post_count": post,
    }
    kc.device_free(dreads)
    kc.close()
    if not a.no_cpu:
        print(f"[bench] cpu baseline: {threads} threads and 1", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, min(a.cpu_sample_reads, w["num_reads"]), threads,
                                           min(a.cpu_sample_reads // 2, w["num_reads"]))
    # the metric's second half, bounded so the whole line takes a few minutes
    if not a.no_batched:
        out["batched"] = run_batched(max(3, min(a.steps, 5)), 1, a.line_batch_reads, with_small=True,
                                     tune_trials=1)
    if not a.no_served:
        out["served"] = run_served(a.line_serve_requests, a.serve_slots, a.serve_max_large)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
