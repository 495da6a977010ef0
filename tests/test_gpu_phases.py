"""The two phases of a P-PE count run apart, against the oracle: Phase 1
(dakc_partition) puts each owner's k-mers in its bin in record order, so bin
o equals the oracle's extraction filtered to owner o element for element;
Phase 2 (dakc_count_keys) turns any array of keys into the table the plain
count makes of the reads they came from; partition then count_keys per bin is
the P-PE count."""
import numpy as np
import pytest

import oracle
from dakc_amd import DakcError, KmerCounter, _lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[_lib.ALGO_AUTO, _lib.ALGO_LSD], ids=["msd", "lsd"])
def kc(request):
    c = KmerCounter(algo=request.param)
    yield c
    c.close()


@pytest.fixture(scope="module")
def reads():
    p = _lib.gen_params(31, 30_000, 150, min_len=120, genome_len=400_000, sub_rate=0.01, n_rate=0.001)
    return _lib.generate_reads(p)


def partition(kc, data, **over):
    kw = 1 if over["k"] <= 32 else 2
    n, _ = kc.partition(data, **over)
    d = kc.device_alloc(max(n, 1) * kw * 8)
    try:
        n2, offs = kc.partition(data, keys_out=d, cap=n, **over)
        assert n2 == n
        return kc.d2h(d, n * kw * 8, np.uint64).reshape(n, kw), offs
    finally:
        kc.device_free(d)


CASES = [(31, False, 64, 0), (31, True, 7, 0x5EED), (55, False, 64, 0), (21, False, 1, 0),
         (63, True, 256, 0x1234567), (32, False, 3, 0)]


@pytest.mark.parametrize("k,canonical,P,seed", CASES)
def test_partition_bins_are_owner_subsequences(kc, reads, k, canonical, P, seed):
    keys, offs = partition(kc, reads, k=k, read_len=150, canonical=canonical, num_pes=P, owner_bin=-1,
                           hash_seed=seed)
    if P == 1:
        assert offs.tolist() == [0, keys.shape[0]]
        assert np.array_equal(keys, oracle.extract(reads, 150, k, canonical))
        return
    assert offs.size == P + 1 and offs[0] == 0 and offs[-1] == keys.shape[0]
    for o in range(P):
        want = oracle.extract(reads, 150, k, canonical, P, o, seed)
        assert np.array_equal(keys[int(offs[o]):int(offs[o + 1])], want), o
    # one bin kept: that bin alone
    b = P // 3
    kb, ob = partition(kc, reads, k=k, read_len=150, canonical=canonical, num_pes=P, owner_bin=b, hash_seed=seed)
    assert ob.tolist() == [0, kb.shape[0]]
    assert np.array_equal(kb, oracle.extract(reads, 150, k, canonical, P, b, seed))


def same_table(t, o):
    keys, counts, n, offs = o
    assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
    if offs is not None:
        assert np.array_equal(t.bin_offsets, offs)


@pytest.mark.parametrize("k,canonical", [(31, False), (31, True), (55, False), (63, True), (16, False)])
def test_count_keys_equals_count_of_the_reads(kc, reads, k, canonical):
    keys = oracle.extract(reads, 150, k, canonical)
    want = oracle.count(reads, 150, k, canonical)
    same_table(kc.count_keys(keys, k=k), want)
    # any order, from device memory
    shuffled = keys[np.random.default_rng(k).permutation(keys.shape[0])]
    d = kc.device_alloc(max(shuffled.nbytes, 8))
    try:
        kc.h2d(d, shuffled)
        same_table(kc.count_keys(dev_ptr=d, n=shuffled.shape[0], k=k), want)
        # all owner bins
        same_table(kc.count_keys(dev_ptr=d, n=shuffled.shape[0], k=k, num_pes=7, hash_seed=77),
                   oracle.count(reads, 150, k, canonical, 7, -1, 77))
    finally:
        kc.device_free(d)


@pytest.mark.parametrize("k", [31, 55])
def test_partition_then_count_keys_is_the_p_pe_count(kc, reads, k):
    P = 16
    kw = 1 if k <= 32 else 2
    n, _ = kc.partition(reads, k=k, read_len=150, num_pes=P, owner_bin=-1)
    d = kc.device_alloc(n * kw * 8)
    try:
        n2, offs = kc.partition(reads, keys_out=d, cap=n, k=k, read_len=150, num_pes=P, owner_bin=-1)
        allk, allc, _, aoffs = oracle.count(reads, 150, k, False, P, -1)
        for o in range(P):
            nb = int(offs[o + 1] - offs[o])
            t = kc.count_keys(dev_ptr=d + int(offs[o]) * kw * 8, n=nb, k=k)
            seg = slice(int(aoffs[o]), int(aoffs[o + 1]))
            assert t.total == nb
            assert np.array_equal(t.keys, allk[seg]) and np.array_equal(t.counts, allc[seg]), o
    finally:
        kc.device_free(d)


def test_phases_empty_input_and_limits(kc):
    n, offs = kc.partition(b"", k=31, read_len=150, num_pes=4, owner_bin=-1)
    assert n == 0 and offs.tolist() == [0] * 5
    t = kc.count_keys(np.zeros((0, 1), np.uint64), k=31)
    assert t.total == 0 and len(t) == 0
    data = _lib.generate_reads(_lib.gen_params(32, 100, 150))
    d = kc.device_alloc(64)
    try:
        with pytest.raises(DakcError) as e:  # 12000 k-mers do not fit 8 keys
            kc.partition(data, keys_out=d, cap=8, k=31, read_len=150)
        assert e.value.status == 3
    finally:
        kc.device_free(d)
    with pytest.raises(DakcError):
        kc.count_keys(np.zeros((4, 1), np.uint64), k=65)
