"""GPU parity at the configs' real sizes, where the geometry reaches limits the
small tests do not: config 1 exactly (1M i.i.d. reads x 150, canonical k=31)
against the oracle's table entry for entry, and 5M genome-like reads (6e8
k-mers) against the oracle's threaded count by checksum of the whole
(kmer, count) multiset, forward, canonical and with all 64 owner bins."""
import numpy as np
import pytest

import oracle
from dakc_amd import KmerCounter, _lib

pytestmark = pytest.mark.gpu


def threads():
    import os
    share = os.environ.get("OMP_NUM_THREADS")
    n = len(os.sched_getaffinity(0))
    return min(n, int(share)) if share and share.isdigit() else n


@pytest.fixture(scope="module")
def config1():
    reads = _lib.generate_reads(_lib.gen_params(2024, 1_000_000, 150), threads=threads())
    return reads, oracle.count(reads, 150, 31, True)


@pytest.mark.parametrize("algo", [_lib.ALGO_AUTO, _lib.ALGO_LSD], ids=["msd", "lsd"])
def test_config1_iid_canonical_k31_exact(config1, algo):
    reads, (keys, counts, n, _) = config1
    with KmerCounter(k=31, read_len=150, canonical=True, algo=algo) as kc:
        t = kc.count(reads)
        st = kc.last_stats()
        assert st.path == (2 if algo == _lib.ALGO_AUTO else 1) and st.msd_reruns == 0
    assert t.total == n == 120_000_000
    assert np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)


@pytest.fixture(scope="module")
def genome5m():
    p = _lib.gen_params(2026, 5_000_000, 150, genome_len=25_000_000, sub_rate=0.005)
    return _lib.generate_reads(p, threads=threads())


def table_sum(t):
    return int(t.total), len(t), oracle.checksum(t.keys, t.counts)


@pytest.mark.parametrize("canonical", [False, True])
def test_genome_5m_reads_checksum(genome5m, canonical):
    n, d, cs, _ = oracle.count_mt(genome5m, 150, 31, threads(), canonical=canonical)
    with KmerCounter(k=31, read_len=150, canonical=canonical) as kc:
        t = kc.count(genome5m)
        assert kc.last_stats().path == 2
    assert table_sum(t) == (n, d, cs)
    k = t.keys[:, 0]
    assert np.all(k[1:] > k[:-1]) and t.counts.min() >= 1


@pytest.mark.parametrize("P", [64, 8])
def test_genome_5m_reads_all_owner_bins(genome5m, P):
    # every bin at once stays on the MSD path; each bin is ascending and holds
    # only its owner's k-mers; the union is the single-PE table
    n, d, cs, _ = oracle.count_mt(genome5m, 150, 31, threads())
    with KmerCounter(k=31, read_len=150, num_pes=P, owner_bin=-1) as kc:
        t = kc.count(genome5m)
        st = kc.last_stats()
    assert st.path == 2 and st.msd_reruns == 0
    assert table_sum(t) == (n, d, cs)
    offs = t.bin_offsets
    assert offs[0] == 0 and offs[-1] == d and np.all(np.diff(offs.astype(np.int64)) > 0)
    rng = np.random.default_rng(1)
    for b in rng.choice(P, 4, replace=False):
        seg = t.keys[int(offs[b]):int(offs[b + 1]), 0]
        assert np.all(seg[1:] > seg[:-1])
        for x in seg[rng.integers(0, seg.size, 50)]:
            assert oracle.murmur64a_u64(int(x)) % P == b
