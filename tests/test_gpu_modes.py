"""Batched and served modes: every job's table equals its plain count and the
oracle, alone and in the mix."""
import random

import numpy as np
import pytest

import oracle
from dakc_amd import KmerCounter, _lib

pytestmark = pytest.mark.gpu


def read_sets(seed, n, width=250):
    rng = random.Random(seed)
    sets = []
    for i in range(n):
        nreads = int(np.exp(rng.uniform(np.log(500), np.log(30_000))))
        p = _lib.gen_params(seed * 1000 + i, nreads, width, min_len=100, sub_rate=0.005,
                            genome_len=rng.choice([0, 200_000]))
        sets.append(_lib.generate_reads(p))
    return sets


def test_batch_matches_plain_and_oracle():
    sets = read_sets(1, 12)
    with KmerCounter(k=31, read_len=250) as kc:
        outs = kc.count_batch(sets)
        for s, t in zip(sets, outs):
            keys, counts, n, _ = oracle.count(s, 250, 31)
            assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)


def test_served_mix_equals_alone():
    sets = read_sets(2, 10)
    ks = [21, 31, 55] * 4
    with KmerCounter(read_len=250, max_inflight=3) as kc:
        alone = [kc.count(s, k=k) for s, k in zip(sets, ks)]
        tickets = [kc.submit(s, k=k) for s, k in zip(sets, ks)]
        mixed = [kc.wait(t) for t, _buf in tickets]
        for a, m in zip(alone, mixed):
            assert a.total == m.total
            assert np.array_equal(a.keys, m.keys) and np.array_equal(a.counts, m.counts)


def test_served_rejects_over_budget_and_survives_bad_job():
    with KmerCounter(k=31, read_len=150, hbm_budget=64 << 20, max_inflight=2) as kc:
        big = np.zeros(151 * 400_000, np.uint8)   # needs far more than 64 MiB
        with pytest.raises(Exception) as e:
            kc.submit(big)
        assert getattr(e.value, "status", None) == 7
        with pytest.raises(Exception) as e:        # not whole records: refused at submit
            kc.submit(b"ACGT" * 40)
        assert getattr(e.value, "status", None) == 2
        small = _lib.generate_reads(_lib.gen_params(3, 1000, 150))
        t, buf = kc.submit(small)
        assert kc.wait(t).total == 1000 * 120


def test_served_bad_record_then_good():
    good = _lib.generate_reads(_lib.gen_params(3, 2000, 150))
    bad = good.copy()
    bad[150] = ord("X")
    with KmerCounter(k=31, read_len=150, max_inflight=2) as kc:
        t1, b1 = kc.submit(bad)
        t2, b2 = kc.submit(good)
        with pytest.raises(Exception) as e:
            kc.wait(t1)
        assert getattr(e.value, "status", None) == 2
        t = kc.wait(t2)
        assert t.total == 2000 * 120
