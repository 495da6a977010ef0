"""GPU parity: the HIP path, called through the C ABI, against the CPU oracle
and the golden fixtures.  Integer work, so every comparison is bit-exact."""
import ctypes
import hashlib
import json
import random

import numpy as np
import pytest

import oracle
from dakc_amd import DakcError, KmerCounter, _lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[_lib.ALGO_AUTO, _lib.ALGO_LSD], ids=["msd", "lsd"])
def kc(request):
    c = KmerCounter(algo=request.param)
    c.algo = request.param
    yield c
    c.close()


def assert_same(gpu, ora):
    keys, counts, n, offs = ora
    assert gpu.total == n
    assert gpu.keys.shape == keys.shape, (gpu.keys.shape, keys.shape)
    assert np.array_equal(gpu.keys, keys)
    assert np.array_equal(gpu.counts, counts)
    if offs is not None:
        assert np.array_equal(gpu.bin_offsets, offs)


def run_both(kc, data, L, k, canonical=False, P=1, b=-1, seed=0):
    g = kc.count(data, k=k, read_len=L, canonical=canonical, num_pes=P, owner_bin=b, hash_seed=seed)
    o = oracle.count(data, L, k, canonical, P, b, seed)
    assert_same(g, o)
    return g


def test_golden_fixtures(kc, golden_dir):
    exp = json.loads((golden_dir / "expected.json").read_text())
    for name, e in exp.items():
        data = (golden_dir / e["reads"]).read_bytes()
        t = kc.count(data, k=e["k"], read_len=e["read_len"], canonical=e["canonical"],
                     num_pes=e["num_pes"], owner_bin=e["owner_bin"])
        blob = t.serialize()
        assert hashlib.sha256(blob).hexdigest() == e["sha256"], name
        assert t.total == e["total"] and len(t) == e["distinct"], name
        if "bin_offsets" in e:
            assert t.bin_offsets.tolist() == e["bin_offsets"], name


def messy_reads(seed, nreads, L):
    rng = random.Random(seed)
    g = "".join(rng.choice("ACGT") for _ in range(5000))
    out = []
    for _ in range(nreads):
        ln = rng.randint(0, L)
        st = rng.randint(0, len(g) - L)
        s = list(g[st:st + ln])
        for i in range(ln):
            u = rng.random()
            if u < 0.02:
                s[i] = rng.choice("acgtNRm \r\x01\xff")
        out.append("".join(s) + "M" * (L - ln) + "\n")
    return "".join(out).encode("latin-1")


@pytest.mark.parametrize("k", [1, 2, 5, 16, 21, 31, 32, 33, 40, 55, 63, 64])
@pytest.mark.parametrize("canonical", [False, True])
def test_messy_reads_all_k(kc, k, canonical):
    L = 150 if k <= 32 else 250
    run_both(kc, messy_reads(k * 7 + canonical, 700, L), L, k, canonical)


@pytest.mark.parametrize("L", [150, 61])
def test_every_byte_value(kc, L):
    """Every byte value 1-255 inside reads of ACGT/acgt, at every offset within
    a 4-byte word (records of odd width start at every alignment), singly and
    in clumps: the decode treats each byte exactly as the oracle does."""
    rng = np.random.default_rng(L)
    n = 255 * 8
    recs = np.frombuffer(b"ACGTacgt", np.uint8)[rng.integers(0, 8, size=(n, L))]
    for i in range(n):
        v = 1 + i % 255
        p = (i // 255) + 4 * int(rng.integers(0, (L - 8) // 4))  # offsets 0..7 of a word pair
        recs[i, p] = v
        if i % 3 == 0:  # a clump of arbitrary bytes after it
            q = int(rng.integers(1, 6))
            recs[i, p + 1:min(L, p + 1 + q)] = rng.integers(1, 256, size=min(L, p + 1 + q) - p - 1)
    data = np.concatenate([recs, np.full((n, 1), 10, np.uint8)], axis=1).tobytes()
    for k in (1, 5, 21, 31):
        run_both(kc, data, L, k)


@pytest.mark.parametrize("L", [1, 7, 31, 32, 33, 63, 64, 65, 100, 257, 1000, 4096])
def test_read_widths(kc, L):
    data = messy_reads(L, 200 if L < 1000 else 20, L)
    for k in sorted({1, min(L, 21), min(L, 31)}):
        run_both(kc, data, L, k)


def test_iid_reads_100k(kc):
    p = _lib.gen_params(42, 100_000, 150)
    data = _lib.generate_reads(p)
    t = run_both(kc, data, 150, 31)
    assert t.total == 100_000 * 120
    st = kc.last_stats()
    assert st.path == (2 if kc.algo == _lib.ALGO_AUTO else 1) and st.msd_reruns == 0


def test_genome_reads_repeats(kc):
    p = _lib.gen_params(7, 200_000, 150, genome_len=1_000_000, sub_rate=0.005)
    data = _lib.generate_reads(p)
    t = run_both(kc, data, 150, 31)
    assert t.counts.max() > 10
    st = kc.last_stats()  # repeats still fit the default final items
    assert st.path == (2 if kc.algo == _lib.ALGO_AUTO else 1) and st.msd_reruns == 0
    run_both(kc, data, 150, 31, canonical=True)
    run_both(kc, data, 150, 21)


def test_k63_owner_bins(kc):
    p = _lib.gen_params(3, 20_000, 250, genome_len=400_000, sub_rate=0.01, n_rate=0.001)
    data = _lib.generate_reads(p)
    run_both(kc, data, 250, 63, P=64, b=0)
    run_both(kc, data, 250, 63, P=64, b=-1)
    run_both(kc, data, 250, 63)


def test_owner_bins_k31(kc):
    p = _lib.gen_params(4, 50_000, 150, genome_len=500_000, sub_rate=0.005)
    data = _lib.generate_reads(p)
    for P, b in [(64, 0), (64, 63), (64, -1), (7, -1), (256, -1), (1000, 3)]:
        run_both(kc, data, 150, 31, P=P, b=b)


def test_owner_hash_seed(kc):
    # a named owner seed moves k-mers between bins the same way on the GPU as
    # in the oracle; 0 is the reference's seed; queries find keys in the
    # seeded bins
    p = _lib.gen_params(14, 30_000, 150, genome_len=300_000, sub_rate=0.005)
    data = _lib.generate_reads(p)
    seed = 0xD1B54A32D192ED03
    for P, b in [(64, 5), (64, -1), (7, -1)]:
        g = run_both(kc, data, 150, 31, P=P, b=b, seed=seed)
        run_both(kc, data, 150, 55, P=P, b=b, seed=seed)
    assert not np.array_equal(g.bin_offsets, run_both(kc, data, 150, 31, P=7, b=-1).bin_offsets)
    d = kc.device_alloc(data.size)
    try:
        kc.h2d(d, data)
        t = kc.count_device(d, data.size, num_pes=7, owner_bin=-1, hash_seed=seed)
        assert t.hash_seed == seed
        keys, counts, _, offs = oracle.count(data, 150, 31, False, 7, -1, seed)
        q = np.concatenate([keys[::97, 0], np.arange(1000, dtype=np.uint64)])
        assert np.array_equal(kc.table_query(t, q), oracle.query(keys, counts, q, offs, seed))
    finally:
        kc.device_free(d)


def test_runs_longer_than_a_tile(kc):
    # identical reads: every k-mer's run spans many tiles (gallop path)
    read = ("ACGTTGCA" * 19)[:150]
    data = ((read + "\n") * 3000).encode()
    t = run_both(kc, data, 150, 31)
    assert t.counts.min() >= 3000
    polya = (("A" * 150 + "\n") * 2000).encode()
    t = run_both(kc, polya, 150, 31)
    assert len(t) == 1 and t.counts[0] == 2000 * 120
    # one k-mer fills a whole sub-bucket: the MSD path must fall back to LSD
    st = kc.last_stats()
    assert st.path == 1
    if kc.algo == _lib.ALGO_AUTO:
        assert st.msd_reruns == 1


def test_empty_and_short(kc):
    t = kc.count(b"", k=31, read_len=150)
    assert len(t) == 0 and t.total == 0
    t = kc.count(b"ACGT\n" * 10, k=31, read_len=4)   # k > read_len
    assert len(t) == 0 and t.total == 0
    t = kc.count((b"N" * 150 + b"\n") * 10, k=31, read_len=150)
    assert len(t) == 0


def test_format_errors(kc):
    good = (b"ACGT" * 37 + b"AC\n") * 20
    assert kc.count(good, k=31, read_len=150).total == 20 * 120
    assert kc.count(good[:-1], k=31, read_len=150).total == 20 * 120
    for bad in (good[:-2], good[:150] + b"X" + good[151:], b"\x00" + good[1:],
                good[:151] + b"\x00" + good[152:]):
        with pytest.raises(DakcError) as e:
            kc.count(bad, k=31, read_len=150)
        assert e.value.status == 2
    for kw in ({"k": 0}, {"k": 65}, {"read_len": 0}, {"num_pes": 0}, {"num_pes": 4, "owner_bin": 4}):
        with pytest.raises(DakcError) as e:
            kc.count(good, **{"k": 31, "read_len": 150, **kw})
        assert e.value.status == 1
    # a later call still works
    assert kc.count(good, k=31, read_len=150).total == 20 * 120


def test_device_generator_matches_host(kc):
    for p in (_lib.gen_params(11, 20_000, 150, genome_len=100_000, sub_rate=0.005),
              _lib.gen_params(12, 5_000, 250, min_len=100, sub_rate=0.005, n_rate=0.001),
              _lib.gen_params(13, 3_001, 151)):
        host = _lib.generate_reads(p)
        dptr = kc.device_alloc(host.size)
        try:
            kc.generate_device(p, dptr, host.size)
            dev = kc.d2h(dptr, host.size)
        finally:
            kc.device_free(dptr)
        assert np.array_equal(host, dev)


def test_device_resident_input(kc):
    p = _lib.gen_params(21, 30_000, 150, genome_len=300_000, sub_rate=0.005)
    host = _lib.generate_reads(p)
    want = kc.count(host)
    dptr = kc.device_alloc(host.size)
    try:
        kc.generate_device(p, dptr, host.size)
        t = kc.count_device(dptr, host.size)
        assert t.on_device == 1 and t.n == len(want)
        kptr = ctypes.cast(t.keys, ctypes.c_void_p).value
        cptr = ctypes.cast(t.counts, ctypes.c_void_p).value
        keys = kc.d2h(kptr, t.n * 8, np.uint64)
        counts = kc.d2h(cptr, t.n * 4, np.uint32)
        assert np.array_equal(keys, want.keys[:, 0]) and np.array_equal(counts, want.counts)
        kc._libh.dakc_table_free(ctypes.byref(t))
    finally:
        kc.device_free(dptr)


def test_low_complexity_prefixes(kc):
    # many distinct k-mers sharing long prefixes land in the same in-LDS bin
    rng = random.Random(5)
    recs = []
    for _ in range(3000):
        tail = "".join(rng.choice("ACGT") for _ in range(40))
        recs.append("A" * 110 + tail + "\n")
    data = "".join(recs).encode()
    run_both(kc, data, 150, 31)
    run_both(kc, data, 150, 21, canonical=True)


def test_repeat_runs_identical(kc):
    p = _lib.gen_params(5, 60_000, 150, genome_len=200_000, sub_rate=0.005)
    data = _lib.generate_reads(p)
    a = kc.count(data)
    for _ in range(3):
        b = kc.count(data)
        assert np.array_equal(a.keys, b.keys) and np.array_equal(a.counts, b.counts)


def test_final_tables_near_full():
    """Reads with 2% substitutions at 10x coverage: 58% of the k-mers are
    distinct, so a final item of about 13,700 keys fills its 10,240-slot hash
    table to about 78% (long probe chains, some wrapping past the last slot,
    and items whose chains run out take the general path).  The count stays
    on the MSD path and matches the oracle."""
    p = _lib.gen_params(11, 200_000, 150, genome_len=3_000_000, sub_rate=0.02)
    data = _lib.generate_reads(p)
    with KmerCounter(k=31, read_len=150) as kc2:
        for canonical in (False, True):
            g = kc2.count(data, canonical=canonical)
            st = kc2.last_stats()
            assert st.path == 2 and st.msd_reruns == 0
            ora = oracle.count(data, 150, 31, canonical)
            assert_same(g, ora)
            assert 0.5 < len(ora[1]) / ora[2] < 0.7  # the load the test is for


def test_items_larger_than_lds_go_digit_by_digit(tmp_path):
    """A tuned fill above 100% packs more second-level digits into a final
    item than LDS holds: every item is then counted digit by digit in its
    workgroup, still on the MSD path, with exact counts."""
    p = _lib.gen_params(7, 200_000, 150, genome_len=1_000_000, sub_rate=0.005)
    data = _lib.generate_reads(p)
    path = tmp_path / "tune.json"
    path.write_text('{\n "entries": [\n  {"key_words": 1, "size_class": 24, "fin_nt": 1024, '
                    '"fill_pct": 150, "algo": "msd", "ms": 1.0}\n ]\n}\n')
    with KmerCounter(k=31, read_len=150) as kc2:
        kc2.tuning_load(path)
        for canonical in (False, True):
            g = kc2.count(data, canonical=canonical)
            st = kc2.last_stats()
            assert st.path == 2 and st.msd_reruns == 0
            assert_same(g, oracle.count(data, 150, 31, canonical))
