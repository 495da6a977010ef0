"""HITTER=1 on the GPU (dakc_count_hitter) against the oracle's restatement
(dko_count_hitter): the light and heavy tables entry for entry and the send
statistics, with the reference's bucket rule; their union is the plain
count's table."""
import random

import numpy as np
import pytest

import oracle
from dakc_amd import KmerCounter, _lib

pytestmark = pytest.mark.gpu


def repetitive_reads(seed, n, L, period=400, n_rate=0.01):
    rng = random.Random(seed)
    g = "".join(rng.choice("ACGT") for _ in range(period))
    out = []
    for _ in range(n):
        st = rng.randint(0, len(g) - L)
        s = list(g[st:st + L])
        for i in range(L):
            if rng.random() < n_rate:
                s[i] = "N"
        ln = rng.randint(L // 2, L)
        out.append("".join(s[:ln]) + "M" * (L - ln) + "\n")
    return np.frombuffer("".join(out).encode(), np.uint8)


@pytest.fixture(scope="module")
def kc():
    c = KmerCounter()
    yield c
    c.close()


def check(kc, data, L, k, bucket, canonical=False, P=1, b=-1, seed=0, device=False):
    if device:
        d = kc.device_alloc(max(data.size, 4))
        kc.h2d(d, data)
        lt, ht, st = kc.count_hitter(dev_ptr=d, nbytes=data.size, bucket_size=bucket, k=k, read_len=L,
                                     canonical=canonical, num_pes=P, owner_bin=b, hash_seed=seed)
        kc.device_free(d)
    else:
        lt, ht, st = kc.count_hitter(data, bucket_size=bucket, k=k, read_len=L, canonical=canonical, num_pes=P,
                                     owner_bin=b, hash_seed=seed)
    (lk, lc, lo), (hk, hc, ho), ost = oracle.count_hitter(data, L, k, canonical, P, b, seed, bucket)
    assert np.array_equal(lt.keys, lk) and np.array_equal(lt.counts, lc)
    assert np.array_equal(ht.keys, hk) and np.array_equal(ht.counts, hc)
    if P > 1 and b < 0:
        assert np.array_equal(lt.bin_offsets, lo) and np.array_equal(ht.bin_offsets, ho)
    for key in ("buckets", "normal_keys", "heavy_pairs"):
        assert st[key] == ost[key], key
    assert st["kmers"] == lt.total + ht.total
    return lt, ht, st


@pytest.mark.parametrize("k,bucket", [(5, 300), (11, 500), (21, 700), (31, 1000), (40, 1200), (63, 2000)])
@pytest.mark.parametrize("canonical", [False, True])
def test_hitter_matches_oracle(kc, k, bucket, canonical):
    data = repetitive_reads(k * 13 + bucket, 600, 80)
    lt, ht, st = check(kc, data, 80, k, bucket, canonical)
    assert len(ht) > 0 and st["heavy_pairs"] > 0


@pytest.mark.parametrize("P,b", [(5, -1), (5, 2), (64, 0), (64, -1)])
def test_hitter_owner_bins(kc, P, b):
    data = repetitive_reads(P + b, 800, 100)
    check(kc, data, 100, 27, 2500, P=P, b=b, seed=0xFEED)
    check(kc, data, 100, 45, 2500, P=P, b=b)


def test_hitter_default_buckets_genome_reads(kc):
    # the reference's KCOUNT_BUCKET_SIZE, reads at 30x of a 1 Mbp genome in
    # generator order (random positions: heavy hitters are rare) and sorted
    # by position (coverage-local: most k-mers are heavy in their bucket)
    p = _lib.gen_params(41, 200_000, 150, genome_len=1_000_000, sub_rate=0.005, both_strands=False)
    data = _lib.generate_reads(p)
    lt, ht, st = check(kc, data, 150, 31, 0, device=True)
    full = kc.count(data)
    merged = np.concatenate([lt.keys[:, 0], ht.keys[:, 0]])
    order = np.argsort(merged)
    assert np.array_equal(merged[order], full.keys[:, 0])
    assert np.array_equal(np.concatenate([lt.counts, ht.counts])[order], full.counts)
    # coverage-local order: reads sorted by their first 20 bases stand in for position order
    recs = data.reshape(-1, 151)
    srt = recs[np.lexsort(recs[:, :20].T[::-1])].reshape(-1)
    lt2, ht2, st2 = check(kc, srt, 150, 31, 0)
    assert st2["bytes_sent"] < st["bytes_sent"] and st["bytes_without"] == st2["bytes_without"]


def test_hitter_device_tables_empty_and_errors(kc):
    data = repetitive_reads(5, 300, 60)
    lt, ht, st = kc.count_hitter(data, bucket_size=700, k=15, read_len=60, table_on_device=True)
    assert lt.on_device == 1 and ht.on_device == 1
    (lk, lc, _), (hk, hc, _), _ = oracle.count_hitter(data, 60, 15, bucket_size=700)
    from dakc_amd.kcounter import _ptr
    assert np.array_equal(kc.d2h(_ptr(lt.keys), int(lt.n) * 8, np.uint64), lk[:, 0])
    assert np.array_equal(kc.d2h(_ptr(ht.counts), int(ht.n) * 4, np.uint32), hc)
    kc._libh.dakc_table_free(_lib.C.byref(lt))
    kc._libh.dakc_table_free(_lib.C.byref(ht))
    lt, ht, st = kc.count_hitter(b"", k=31, read_len=150)
    assert len(lt) == len(ht) == 0 and st["buckets"] == 0
    with pytest.raises(_lib.DakcError):
        kc.count_hitter(data, bucket_size=60, k=15, read_len=60)  # a bucket no larger than a record
    with pytest.raises(_lib.DakcError):
        kc.count_hitter(data, bucket_size=20000, k=15, read_len=60)  # over the 10240 limit
