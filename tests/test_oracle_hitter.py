"""CPU pins of the oracle's HITTER=1 restatement (oracle/dakc_oracle.c,
dko_count_hitter, after kcounter.cpp:56-104, 176-245, 325-384): the light and
heavy tables and the send statistics equal an independent set-based
statement in tests/brute.py; light and heavy are disjoint and their union is
the HITTER=0 table; the bucket rule is the reference's record-granular one."""
import random

import numpy as np
import pytest

import brute
import oracle


def as_dict(keys, counts):
    if keys.shape[1] == 1:
        return {int(k[0]): int(c) for k, c in zip(keys, counts)}
    return {int(k[0]) | (int(k[1]) << 64): int(c) for k, c in zip(keys, counts)}


def repetitive_reads(seed, n, L, period=None, n_rate=0.01):
    """reads over a short genome (k-mers repeat inside a bucket), with N runs"""
    rng = random.Random(seed)
    g = "".join(rng.choice("ACGT") for _ in range(period or 400))
    out = []
    for _ in range(n):
        st = rng.randint(0, len(g) - L)
        s = list(g[st:st + L])
        for i in range(L):
            if rng.random() < n_rate:
                s[i] = "N"
        ln = rng.randint(L // 2, L)
        out.append("".join(s[:ln]) + "M" * (L - ln) + "\n")
    return "".join(out).encode()


@pytest.mark.parametrize("k,bucket", [(5, 300), (11, 500), (21, 700), (31, 1000), (40, 1200)])
@pytest.mark.parametrize("canonical", [False, True])
def test_hitter_matches_bruteforce(k, bucket, canonical):
    L = 60
    data = repetitive_reads(k * 7 + bucket, 90, L)
    (lk, lc, _), (hk, hc, _), st = oracle.count_hitter(data, L, k, canonical, bucket_size=bucket)
    wl, wh, ws = brute.hitter(data, L, k, bucket, canonical)
    assert as_dict(lk, lc) == wl and as_dict(hk, hc) == wh and st == ws
    assert len(wh) > 0  # the repeats make heavy hitters
    for keys in (lk, hk):  # ascending, as the reference's tables
        v = [int(k0[0]) | (int(k0[-1]) << 64 if keys.shape[1] == 2 else 0) for k0 in keys]
        assert v == sorted(v) and len(set(v)) == len(v)


@pytest.mark.parametrize("P,b", [(5, -1), (5, 2), (64, 0)])
def test_hitter_owner_bins(P, b):
    L, k, bucket = 80, 15, 200
    seed = 0xABCDEF
    data = repetitive_reads(3, 120, L)
    (lk, lc, lo), (hk, hc, ho), st = oracle.count_hitter(data, L, k, False, P, b, seed, bucket)
    wl, wh, ws = brute.hitter(data, L, k, bucket, False, P, b, seed)
    assert as_dict(lk, lc) == wl and as_dict(hk, hc) == wh and st == ws
    if b < 0:
        for o in range(P):  # each bin holds only its owner's k-mers
            for keys, offs in ((lk, lo), (hk, ho)):
                seg = keys[int(offs[o]):int(offs[o + 1])]
                assert all(brute.owner(int(x[0]), k, P, seed) == o for x in seg)


@pytest.mark.parametrize("k", [21, 31, 55])
def test_light_and_heavy_are_the_hitter0_table(k):
    L = 150
    data = repetitive_reads(k, 400, L, period=3000)
    (lk, lc, _), (hk, hc, _), _ = oracle.count_hitter(data, L, k, bucket_size=10000)
    fk, fc, n, _ = oracle.count(data, L, k)
    light, heavy = as_dict(lk, lc), as_dict(hk, hc)
    assert not (light.keys() & heavy.keys())
    assert {**light, **heavy} == as_dict(fk, fc)
    assert sum(light.values()) + sum(heavy.values()) == n


def test_bucket_rule_and_send_statistics():
    # 150-base records of 120 31-mers each: a bucket of 10000 takes records
    # while it holds <= 9850 k-mers, i.e. 83 records (9960 k-mers)
    rng = random.Random(5)
    recs = ["".join(rng.choice("ACGT") for _ in range(150)) for _ in range(200)]
    data = ("\n".join(recs) + "\n").encode()
    _, _, st = oracle.count_hitter(data, 150, 31, bucket_size=10000)
    assert st == {"buckets": 3, "normal_keys": 200 * 120, "heavy_pairs": 0}
    # the same record three times in one bucket: all its k-mers heavy
    data3 = ((recs[0] + "\n") * 3).encode()
    (lk, lc, _), (hk, hc, _), st = oracle.count_hitter(data3, 150, 31, bucket_size=10000)
    assert len(lc) == 0 and len(hc) == 120 and set(hc.tolist()) == {3}
    assert st == {"buckets": 1, "normal_keys": 0, "heavy_pairs": 120}
    # ... but in three buckets (a bucket per record) nothing is heavy
    (lk, lc, _), (hk, hc, _), st = oracle.count_hitter(data3, 150, 31, bucket_size=151)
    assert len(hc) == 0 and len(lc) == 120 and st["buckets"] == 3
    # empty input, and a bucket no larger than a record
    (lk, lc, _), (hk, hc, _), st = oracle.count_hitter(b"", 150, 31)
    assert len(lc) == len(hc) == 0 and st["buckets"] == 0
    with pytest.raises(oracle.OracleError):
        oracle.count_hitter(data, 150, 31, bucket_size=150)
