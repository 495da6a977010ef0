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


def test_batch_device_tables_match_oracle():
    """dakc_count_batch with table_on_device: every job's table lands in device
    memory of its own (also the job that overflows the MSD path and is
    recounted on the LSD path, and an owner-bin job), equals the oracle's, and
    outlives the next batch on the same context."""
    sets = read_sets(7, 6)
    sets.append(np.frombuffer((b"A" * 250 + b"\n") * 30_000, np.uint8))  # one k-mer 6.6e6 times
    with KmerCounter(k=31, read_len=250) as kc:
        first = kc.count_batch_device(sets)
        kc.count_batch_device(sets[:2])[0].free()  # reuses the workspace
        for s, t in zip(sets, first):
            assert t.t.on_device == 1
            keys, counts, n, _ = oracle.count(s, 250, 31)
            h = t.to_host()
            assert h.total == n and np.array_equal(h.keys, keys) and np.array_equal(h.counts, counts)
        binned = kc.count_batch_device(sets[:3], num_pes=8, owner_bin=-1)
        for s, t in zip(sets[:3], binned):
            keys, counts, n, offs = oracle.count(s, 250, 31, False, 8, -1)
            h = t.to_host()
            assert np.array_equal(h.keys, keys) and np.array_equal(h.counts, counts)
            assert np.array_equal(h.bin_offsets, offs)


def test_batch_and_serve_canonical_match_oracle():
    """Canonical jobs, whose final items are sized per quarter of the key
    range, in one batch with a low solo threshold (some jobs get final
    launches of their own, the rest share the multi-job launch), then served
    side by side: every table equals the oracle's."""
    sets = read_sets(5, 8)
    want = [oracle.count(s, 250, 31, True) for s in sets]
    with KmerCounter(k=31, read_len=250, canonical=True, batch_solo_items=64, max_inflight=3) as kc:
        for t, (keys, counts, n, _) in zip(kc.count_batch(sets), want):
            assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
        tickets = [kc.submit(s) for s in sets]
        for (tk, _buf), (keys, counts, n, _) in zip(tickets, want):
            t = kc.wait(tk)
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


def mixed_tuning(path):
    """A tuning file whose size classes ask for every final-level size (and
    one for the LSD path); read_sets(11, 8) falls in classes 17, 19, 20, 22."""
    rows = []
    for cls in range(12, 28):  # classes: log2 of the k-mer slots
        fin, algo = {17: 256, 19: 512, 22: 1024}.get(cls, (256, 512, 1024)[cls % 3]), "msd"
        if cls == 20:
            algo = "lsd"
        rows.append(f'  {{"key_words": 1, "size_class": {cls}, "fin_nt": {fin}, "fill_pct": 60, '
                    f'"algo": "{algo}", "ms": 1.0}}')
    path.write_text('{\n "entries": [\n' + ",\n".join(rows) + "\n ]\n}\n")


def test_captured_batch_takes_every_job_across_tuned_classes(tmp_path):
    """Size classes tuned to different final-level sizes (and one to the LSD
    path) still go into the one captured launch, each job keeping its own
    final-level size (one final launch per size), with exact counts; the
    uncaptured batch groups them the same way."""
    path = tmp_path / "tune.json"
    mixed_tuning(path)
    sets = read_sets(11, 8)
    want = [oracle.count(s, 250, 31) for s in sets]
    with KmerCounter(k=31, read_len=250) as kc:
        kc.tuning_load(path)
        with kc.batch_create(sets) as b:
            in_launch, held = b.info()
            assert in_launch == len(sets) and held > 0
            b.run()
            got = b.tables()
        for (keys, counts, n, _), t in zip(want, got):
            assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
        for (keys, counts, n, _), t in zip(want, kc.count_batch(sets)):
            assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)


@pytest.mark.parametrize("solo", [1, 512])
def test_batch_solo_final_launches_match_oracle(tmp_path, solo):
    """Jobs with at least `solo` final-level items get a final launch of their
    own (solo=1: every job; 512: the larger sets, the rest share the
    work-queue launch), beside jobs of other final-level sizes: exact counts,
    captured and uncaptured."""
    path = tmp_path / "tune.json"
    mixed_tuning(path)
    sets = read_sets(11, 8)
    want = [oracle.count(s, 250, 31) for s in sets]
    with KmerCounter(k=31, read_len=250, batch_solo_items=solo) as kc:
        kc.tuning_load(path)
        with kc.batch_create(sets) as b:
            assert b.info()[0] == len(sets)
            for _ in range(2):
                b.run()
                got = b.tables()
                for (keys, counts, n, _), t in zip(want, got):
                    assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
        for (keys, counts, n, _), t in zip(want, kc.count_batch(sets)):
            assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)

def test_served_device_tables_match_oracle():
    """Served jobs that keep their table in device memory: each table equals
    the oracle's, whether read back at once or after later jobs ran."""
    rng = np.random.default_rng(12)
    params = [_lib.gen_params(300 + i, int(rng.integers(3000, 40000)), 150, sub_rate=0.005,
                              genome_len=150_000) for i in range(6)]
    host = [_lib.generate_reads(p) for p in params]
    ks = [21, 31, 55, 31, 21, 55]
    with KmerCounter(read_len=150, max_inflight=3) as kc:
        tickets = [kc.submit(h, k=k, table_on_device=True)[0] for h, k in zip(host, ks)]
        held = [kc.wait_device(tk) for tk in tickets]
        # a few more jobs while the first tables stay resident
        more = [kc.submit(h, k=k)[0] for h, k in zip(host[:3], ks[:3])]
        for h, k, tk in zip(host[:3], ks[:3], more):
            t = kc.wait(tk)
            keys, counts, n, _ = oracle.count(h, 150, k)
            assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
        for h, k, dt in zip(host, ks, held):
            keys, counts, n, _ = oracle.count(h, 150, k)
            t = dt.to_host()
            assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)


def test_served_jobs_queue_for_pool_memory():
    """A pool that holds only a few jobs at a time: later jobs wait for memory
    (not fail) and every table is exact."""
    params = [_lib.gen_params(400 + i, 20_000, 150, genome_len=500_000, sub_rate=0.01) for i in range(10)]
    host = [_lib.generate_reads(p) for p in params]
    with KmerCounter(k=31, read_len=150, hbm_budget=160 << 20, max_inflight=4) as kc:
        tickets = [kc.submit(h)[0] for h in host]
        for h, tk in zip(host, tickets):
            t = kc.wait(tk)
            keys, counts, n, _ = oracle.count(h, 150, 31)
            assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
        st = kc.serve_stats()
        assert st["completed"] == len(host) and st["failed"] == 0
        assert st["peak_device_bytes"] <= st["budget"] and st["pool_used"] == 0


def test_served_held_device_tables_fill_pool_then_fail_cleanly():
    """Device tables the client keeps count against the pool: once only they
    are in the way a new job fails with out-of-memory, and freeing them lets
    the next job run."""
    host = _lib.generate_reads(_lib.gen_params(77, 20_000, 150))  # iid: table ~ all k-mers
    keys, counts, n, _ = oracle.count(host, 150, 31)
    with KmerCounter(k=31, read_len=150, hbm_budget=128 << 20, max_inflight=2) as kc:
        held, status = [], None
        for _ in range(12):
            tk = kc.submit(host, table_on_device=True)[0]
            try:
                held.append(kc.wait_device(tk))
            except Exception as e:
                status = getattr(e, "status", None)
                break
        assert status == 3 and len(held) >= 1
        for dt in held:
            assert dt.total == n and dt.n == len(counts)
        for dt in held:
            dt.free()
        assert kc.serve_stats()["resident_bytes"] == 0
        t = kc.wait(kc.submit(host)[0])
        assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)


def test_served_close_cancels_jobs_that_cannot_fit():
    """Finished tables nobody has claimed fill a small pool, so the jobs behind
    them can only wait; closing the context cancels those jobs and returns."""
    import threading
    import time
    host = _lib.generate_reads(_lib.gen_params(78, 20_000, 150))  # iid: table ~ all k-mers
    kc = KmerCounter(k=31, read_len=150, hbm_budget=128 << 20, max_inflight=2)
    njobs = 8
    for _ in range(njobs):
        kc.submit(host, table_on_device=True)
    deadline = time.time() + 60
    st = kc.serve_stats()
    while time.time() < deadline:
        st = kc.serve_stats()
        # settled: nothing running, some tables done and unclaimed, the rest queued
        if st["completed"] >= 1 and st["queued"] >= 1 and st["admitted"] == st["completed"] + st["failed"]:
            time.sleep(0.2)
            if kc.serve_stats() == st:
                break
        time.sleep(0.05)
    assert st["completed"] >= 1 and st["queued"] >= 1, st
    closer = threading.Thread(target=kc.close, daemon=True)
    closer.start()
    closer.join(timeout=60)
    assert not closer.is_alive(), "dakc_ctx_destroy did not return"
