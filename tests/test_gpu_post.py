"""After the count, on device tables: the count-frequency histogram and its
minimum bucket, the solid-k-mer filter (in place, owner-bin offsets moved) and
point queries, each against the oracle applied to the same table.  The
counts themselves are checked against the oracle first."""
import numpy as np
import pytest

import oracle
from dakc_amd import KmerCounter, _lib
from dakc_amd.kcounter import _ptr

pytestmark = pytest.mark.gpu

CASES = [  # (name, gen kwargs, count kwargs)
    ("k31_genome", dict(num_reads=60_000, read_len=150, genome_len=300_000, sub_rate=0.01), dict(k=31)),
    ("k55_two_words", dict(num_reads=30_000, read_len=150, genome_len=200_000, sub_rate=0.01), dict(k=55)),
    ("k31_p64_bins", dict(num_reads=40_000, read_len=150, genome_len=200_000, sub_rate=0.01),
     dict(k=31, num_pes=64)),
    ("k21_canonical_p7", dict(num_reads=30_000, read_len=100, genome_len=100_000, sub_rate=0.02, n_rate=0.001),
     dict(k=21, canonical=True, num_pes=7)),
]


def host_copy(kc, t):
    n, kw = int(t.n), int(t.key_words)
    keys = kc.d2h(_ptr(t.keys), n * kw * 8, np.uint64).reshape(n, kw) if n else np.zeros((0, kw), np.uint64)
    counts = kc.d2h(_ptr(t.counts), n * 4, np.uint32) if n else np.zeros(0, np.uint32)
    offs = np.ctypeslib.as_array(t.bin_offsets, shape=(int(t.num_bins) + 1,)).copy() if t.bin_offsets else None
    return keys, counts, offs


def device_count(kc, seed, gen, cnt):
    p = _lib.gen_params(seed, gen["num_reads"], gen["read_len"], genome_len=gen["genome_len"],
                        sub_rate=gen["sub_rate"], n_rate=gen.get("n_rate", 0.0))
    reads = _lib.generate_reads(p)
    d = kc.device_alloc(reads.size)
    kc.h2d(d, reads)
    t = kc.count_device(d, reads.size, **cnt)
    return reads, d, t


@pytest.mark.parametrize("name,gen,cnt", CASES, ids=[c[0] for c in CASES])
def test_histogram_filter_query_match_oracle(name, gen, cnt):
    with KmerCounter(read_len=gen["read_len"]) as kc:
        reads, d, t = device_count(kc, 77, gen, cnt)
        try:
            keys, counts, offs = host_copy(kc, t)
            wk, wc, wn, woffs = oracle.count(reads, gen["read_len"], cnt["k"], cnt.get("canonical", False),
                                             cnt.get("num_pes", 1))
            assert np.array_equal(keys, wk) and np.array_equal(counts, wc) and int(t.total) == wn
            # histogram and its minimum bucket; 20000 buckets take the HBM-atomic path
            for nb in (0, 1, 2, 3, 64, 20000):
                hist, m = kc.table_histogram(t, nb)
                whist, wm = oracle.histogram(counts, nb)
                assert np.array_equal(hist, whist) and m == wm, nb
            # queries: present keys, random keys (nearly all absent), host and device buffers
            rng = np.random.default_rng(5)
            kw = keys.shape[1]
            q = np.concatenate([keys[rng.integers(0, len(keys), 3000)],
                                rng.integers(0, 2**62, size=(3000, kw), dtype=np.uint64)])
            if kw == 2 and cnt["k"] < 64:  # keep random keys inside 2k bits
                q[3000:, 1] &= np.uint64((1 << (2 * cnt["k"] - 64)) - 1)
            want = oracle.query(keys, counts, q, offs)
            assert np.array_equal(kc.table_query(t, q), want)
            dq = kc.device_alloc(q.nbytes)
            do = kc.device_alloc(q.shape[0] * 4)
            kc.h2d(dq, q)
            kc.table_query(t, dev_queries=dq, nq=q.shape[0], dev_out=do)
            assert np.array_equal(kc.d2h(do, q.shape[0] * 4, np.uint32), want)
            kc.device_free(dq)
            kc.device_free(do)
            # filter step by step: a filter of a filtered table is the filter
            # by the larger cutoff; the last one empties the table
            _, m64 = oracle.histogram(counts, 64)
            cut = 0
            for mc in (1, 2, 7, m64, 3, 10**9):
                kc.table_filter(t, mc)
                cut = max(cut, mc)
                fk, fc, fo = host_copy(kc, t)
                ek, ec, etot, eoffs = oracle.filter_table(keys, counts, cut, offs)
                assert np.array_equal(fk, ek) and np.array_equal(fc, ec) and int(t.total) == etot, mc
                if offs is not None:
                    assert np.array_equal(fo, eoffs), mc
            assert int(t.n) == 0
        finally:
            kc.device_free(d)


def test_filter_then_query_on_served_device_table():
    """A served table kept in device memory: filter in place, then query it."""
    gen = dict(num_reads=50_000, read_len=150, genome_len=250_000, sub_rate=0.01)
    p = _lib.gen_params(91, gen["num_reads"], 150, genome_len=gen["genome_len"], sub_rate=gen["sub_rate"])
    reads = _lib.generate_reads(p)
    with KmerCounter(k=31, read_len=150, max_inflight=2) as kc:
        tk, _buf = kc.submit(reads, table_on_device=True)
        dt = kc.wait_device(tk)
        keys, counts, _, _ = oracle.count(reads, 150, 31)
        hist, m = kc.table_histogram(dt, 40)
        assert (hist, m)[1] == oracle.histogram(counts, 40)[1]
        kc.table_filter(dt, m)
        ek, ec, etot, _ = oracle.filter_table(keys, counts, m)
        assert dt.n == len(ec) and dt.total == etot
        q = np.concatenate([keys[::97], keys[::89] + np.uint64(1)])
        assert np.array_equal(kc.table_query(dt, q), oracle.query(ek, ec, q))
        t = dt.to_host()
        assert np.array_equal(t.keys, ek) and np.array_equal(t.counts, ec)


def test_post_ops_refuse_host_tables():
    with KmerCounter(k=31, read_len=150) as kc:
        p = _lib.gen_params(3, 100, 150)
        reads = _lib.generate_reads(p)
        t = kc.count(reads)  # a host table
        from dakc_amd._lib import Table, DakcError
        raw = Table()
        raw.on_device = 0
        with pytest.raises(DakcError):
            kc.table_filter(raw, 2)
        assert len(t) > 0
