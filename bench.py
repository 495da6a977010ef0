"""Benchmark of the k-mer counting hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload genome|iid]

A step is one full count (decode, extract, radix sort, run-length reduce) of
the genome-like workload: a 100 Mbp uniform-ACGT genome, 20,000,000 reads of
150 bp sampled at 30x from either strand with 0.5% substitutions, k = 31,
forward k-mers in one u64, counts u32.  The reads are generated on the device
and are resident in HBM before timing starts; the table stays in HBM.

Prints one JSON line: k-mers counted per second (whole job), the roofline of
the dominant kernel measured with HIP events on the stream it runs on, and a
CPU baseline (the oracle's DAKC P-PE shape on the host cores) on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
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
    # plain single job: 1M i.i.d. reads x 150, k=31
    "iid": dict(seed=2024, num_reads=1_000_000, read_len=150, genome_len=0, sub_rate=0.0, k=31),
}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/traffic.json), or None."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    tag = {"genome": "genome_k31", "iid": "iid_k31"}.get(workload)
    data = json.loads(f.read_text()).get(tag, {}).get("kernels", {})
    for name, v in data.items():
        if name.startswith(f"dk::{kernel}<"):
            return v["traffic_bytes"]
    return None


def cpu_baseline(w: dict, sample_reads: int, threads: int) -> dict:
    import oracle
    from dakc_amd import _lib

    p = _lib.gen_params(w["seed"], sample_reads, w["read_len"], genome_len=w["genome_len"],
                        sub_rate=w["sub_rate"])
    reads = _lib.generate_reads(p, threads=threads)
    oracle.count_mt(reads, w["read_len"], w["k"], threads)  # warm-up
    times = []
    n = 0
    for _ in range(3):
        n, d, cs, el = oracle.count_mt(reads, w["read_len"], w["k"], threads)
        times.append(el)
    t = float(np.median(times))
    return {"value": n / t, "unit": "k-mers/s", "cores": threads, "kind": "port",
            "sample": f"first {sample_reads} reads of the same workload ({n} k-mers), "
                      f"oracle P-PE shape (P = threads), median of 3 after 1 warm-up, "
                      f"{t:.3f} s each"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="genome", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample-reads", type=int, default=1_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
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

    w = WORKLOADS[a.workload]
    kc = KmerCounter(k=w["k"], read_len=w["read_len"],
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
                               f"k={w['k']}", "k": w["k"], "kmers_per_step": n_kmers,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                     "avg_ms": dom_ms, "alg_bytes_per_launch": dom_bytes},
        "stages": stages,
    }
    kc.device_free(dreads)
    kc.close()
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(w, min(a.cpu_sample_reads, w["num_reads"]), threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
