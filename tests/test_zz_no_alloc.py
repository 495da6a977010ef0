"""The allocation-free count (dakc_count_noalloc) through the C ABI: the same
table as dakc_count_device on the same device reads, on calls in a row (the
second smaller, so it finds the first call's contents in the shared array),
for one- and two-word keys, all owner bins and an LSD recount; and its limits."""
import numpy as np
import pytest

from dakc_amd import KmerCounter, _lib
from dakc_amd.kcounter import _ptr

pytestmark = pytest.mark.gpu


def device_table(kc, t):
    kw, n = int(t.key_words), int(t.n)
    keys = kc.d2h(_ptr(t.keys), n * kw * 8, np.uint64).reshape(n, kw)
    counts = kc.d2h(_ptr(t.counts), n * 4, np.uint32)
    offs = np.ctypeslib.as_array(t.bin_offsets, shape=(t.num_bins + 1,)).copy() if t.num_bins else None
    return keys, counts, int(t.total), offs


def both(kc, data, **over):
    """(count_device table, noalloc table) of host reads `data`."""
    d = kc.device_alloc(max(data.size, 4))
    kc.h2d(d, data)
    ref = device_table(kc, kc.count_device(d, data.size, **over))
    kw = 1 if over.get("k", kc.k) <= 32 else 2
    cap = max(len(ref[1]), 1)
    dk, dc = kc.device_alloc(cap * kw * 8), kc.device_alloc(cap * 4)
    try:
        st, nd, nt, offs = kc.count_noalloc(d, data.size, dk, dc, cap, **over)
        assert st == 0, _lib.STATUS_NAMES.get(st)
        got = (kc.d2h(dk, nd * kw * 8, np.uint64).reshape(nd, kw), kc.d2h(dc, nd * 4, np.uint32), nt, offs)
    finally:
        kc.device_free(dk)
        kc.device_free(dc)
        kc.device_free(d)
    return ref, got


def same(ref, got):
    assert got[2] == ref[2]
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    if ref[3] is not None:
        assert np.array_equal(got[3], ref[3])


@pytest.mark.parametrize("k", [31, 55])
def test_noalloc_equals_count_device_on_calls_in_a_row(k):
    big = _lib.generate_reads(_lib.gen_params(801, 40_000, 150, genome_len=400_000, sub_rate=0.005))
    small = _lib.generate_reads(_lib.gen_params(802, 7_000, 150, genome_len=90_000, sub_rate=0.01,
                                                n_rate=0.001))
    with KmerCounter(k=k, read_len=150) as kc:
        for data in (big, small):
            same(*both(kc, data))
        for data in (big, small):
            same(*both(kc, data, canonical=True))
        for P, b in [(7, -1), (64, 3)]:
            for data in (big, small):
                same(*both(kc, data, num_pes=P, owner_bin=b))


def test_noalloc_lsd_recount_and_iid_reads():
    # poly-A reads overflow the MSD path's LDS and are recounted on the LSD
    # path inside the same array; i.i.d. reads make every k-mer distinct
    polya = np.frombuffer((b"A" * 150 + b"\n") * 3000, np.uint8)
    iid = _lib.generate_reads(_lib.gen_params(803, 30_000, 150))
    with KmerCounter(k=31, read_len=150) as kc:
        for data in (polya, iid, polya):
            same(*both(kc, data))


def test_noalloc_limits():
    data = _lib.generate_reads(_lib.gen_params(804, 5_000, 150, genome_len=50_000, sub_rate=0.01))
    with KmerCounter(k=31, read_len=150) as kc:
        d = kc.device_alloc(data.size)
        kc.h2d(d, data)
        ref = device_table(kc, kc.count_device(d, data.size))
        dk, dc = kc.device_alloc(8 * 16), kc.device_alloc(4 * 16)
        try:
            # a table larger than the caller's buffers: nothing written, the size told
            st, nd, nt, _ = kc.count_noalloc(d, data.size, dk, dc, 16)
            assert st == 3 and nd == len(ref[1]) and nt == ref[2]
            # a job whose workspace exceeds the array is refused before any launch
            big_nb = 2_000_000 * 151
            dbig = kc.device_alloc(big_nb)
            try:
                kc.generate_device(_lib.gen_params(805, 2_000_000, 150), dbig, big_nb)
                st, _, _, _ = kc.count_noalloc(dbig, big_nb, dk, dc, 16)
                assert st == 4
            finally:
                kc.device_free(dbig)
            # host reads are refused
            job = kc.job(int(data.ctypes.data), data.size)
            n1, n2 = _lib.C.c_uint64(), _lib.C.c_uint64()
            assert kc._libh.dakc_count_noalloc(kc.handle, _lib.C.byref(job), dk, dc, 16, None,
                                               _lib.C.byref(n1), _lib.C.byref(n2)) == 1
            # empty input
            st, nd, nt, _ = kc.count_noalloc(d, 0, dk, dc, 16)
            assert (st, nd, nt) == (0, 0, 0)
        finally:
            kc.device_free(dk)
            kc.device_free(dc)
            kc.device_free(d)
