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


def test_captured_batch_replays_match_oracle():
    sets = read_sets(3, 9)
    want = [oracle.count(s, 250, 31) for s in sets]
    with KmerCounter(k=31, read_len=250) as kc:
        with kc.batch_create(sets) as b:
            for rep in range(3):
                ms = b.run()
                assert ms > 0
                # eager work between replays must not disturb the captured one
                kc.count(sets[rep], k=21)
                got = b.tables()
                for (keys, counts, n, _), t in zip(want, got):
                    assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)


def test_captured_batch_device_inputs_and_mixed_k():
    rng = np.random.default_rng(4)
    with KmerCounter(k=55, read_len=150) as kc:
        params = [_lib.gen_params(50 + i, int(rng.integers(2000, 20000)), 150, sub_rate=0.005,
                                  genome_len=100_000) for i in range(5)]
        host = [_lib.generate_reads(p) for p in params]
        ptrs = []
        try:
            for p, h in zip(params, host):
                d = kc.device_alloc(h.size)
                kc.generate_device(p, d, h.size)
                ptrs.append(d)
            with kc.batch_create(ptrs, on_device=True, sizes=[h.size for h in host]) as b:
                b.run()
                got = b.tables()
            for h, t in zip(host, got):
                keys, counts, n, _ = oracle.count(h, 150, 55)
                assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
        finally:
            for d in ptrs:
                kc.device_free(d)


def test_autotune_keeps_counts_exact(tmp_path):
    p = _lib.gen_params(8, 40_000, 150, genome_len=300_000, sub_rate=0.005)
    data = _lib.generate_reads(p)
    keys, counts, n, _ = oracle.count(data, 150, 31)
    with KmerCounter(k=31, read_len=150) as kc:
        best = kc.autotune(data, trials=1)
        assert best > 0
        t = kc.count(data)
        assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
        path = tmp_path / "tune.json"
        kc.tuning_save(path)
        assert "size_class" in path.read_text()
    with KmerCounter(k=31, read_len=150) as kc2:
        kc2.tuning_load(path)
        t = kc2.count(data)
        assert np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
