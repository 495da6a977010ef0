"""CPU tests that pin the oracle (oracle/dakc_oracle.c) to things other than
its own formulas: the reference's char2base table as a golden vector, the
SMHasher verification value of MurmurHash64A, brute-force substring counting
(tests/brute.py), committed golden fixtures, and invariants (bin union, sizes,
canonical identity)."""
import hashlib
import json
import random

import numpy as np
import pytest

import brute
import oracle


def test_char2base_matches_reference_table(golden_dir):
    lut = (golden_dir / "char2base_lut.bin").read_bytes()
    assert len(lut) == 256
    got = bytes(oracle.char2base(c) for c in range(256))
    assert got == lut


def test_char2base_roundtrip_base2char():
    # base2char {'C','A','T','G'} (common.hpp:94-97) inverts char2base on ACGT
    for code, ch in enumerate("CATG"):
        assert oracle.char2base(ord(ch)) == code
        assert oracle.char2base(ord(ch.lower())) == code


def test_murmur_general_smhasher_verification():
    # published SMHasher verification value of MurmurHash64A
    assert brute.smhasher_verification(brute.murmur64a) == 0x1F0D3804


def test_murmur_oracle_matches_general_form():
    rng = random.Random(1)
    for _ in range(2000):
        key = rng.getrandbits(64)
        seed = rng.choice([oracle.OWNER_SEED, 0, rng.getrandbits(64)])
        assert oracle.murmur64a_u64(key, seed) == brute.murmur64a(key.to_bytes(8, "little"), seed)
        lo, hi = rng.getrandbits(64), rng.getrandbits(64)
        assert oracle.murmur64a_u128(lo, hi, seed) == brute.murmur64a(
            lo.to_bytes(8, "little") + hi.to_bytes(8, "little"), seed)


def random_reads(seed, nreads, L, alphabet="ACGTACGTACGTacgtNMRX \r", mpad=True):
    rng = random.Random(seed)
    recs = []
    for _ in range(nreads):
        ln = rng.randint(0, L) if mpad else L
        s = "".join(rng.choice(alphabet) for _ in range(ln)) + "M" * (L - ln)
        recs.append(s + "\n")
    return "".join(recs).encode("latin-1")


@pytest.mark.parametrize("k", [1, 2, 5, 21, 31, 32, 33, 47, 55, 63, 64])
@pytest.mark.parametrize("canonical", [False, True])
def test_extract_matches_bruteforce(k, canonical):
    L = 80 if k <= 32 else 130
    data = random_reads(100 + k, 40, L, alphabet="ACGTACGTACGTACGTACGTacgN")
    want = sorted(brute.kmers(data, L, k, canonical))
    got = oracle.extract(data, L, k, canonical)
    if k <= 32:
        gotv = sorted(int(x) for x in got[:, 0])
    else:
        gotv = sorted(int(lo) | (int(hi) << 64) for lo, hi in got)
    assert gotv == want


@pytest.mark.parametrize("k,L", [(3, 20), (21, 150), (31, 150), (55, 250), (63, 250)])
def test_count_matches_bruteforce(k, L):
    # a small genome so k-mers repeat
    rng = random.Random(k)
    g = "".join(rng.choice("ACGT") for _ in range(400))
    recs = []
    for _ in range(60):
        ln = rng.randint(max(1, L // 2), L)
        st = rng.randint(0, len(g) - min(ln, len(g)))
        s = (g * 2)[st:st + ln]
        s = "".join(c if rng.random() > 0.01 else "N" for c in s)
        recs.append(s + "M" * (L - len(s)) + "\n")
    data = "".join(recs).encode()
    keys, counts, n, _ = oracle.count(data, L, k)
    want = brute.count(data, L, k)
    got = [(int(r[0]) if k <= 32 else int(r[0]) | (int(r[1]) << 64), int(c)) for r, c in zip(keys, counts)]
    assert got == want
    assert n == sum(c for _, c in want)


def _ser(keys, counts):
    kw = keys.shape[1]
    rec = np.zeros(len(counts), dtype=[("k", "<u8", (kw,)), ("c", "<u4")])
    rec["k"] = keys
    rec["c"] = counts
    return rec.tobytes()


def test_oracle_matches_golden_fixtures(golden_dir):
    exp = json.loads((golden_dir / "expected.json").read_text())
    for name, e in exp.items():
        data = (golden_dir / e["reads"]).read_bytes()
        keys, counts, n, offs = oracle.count(data, e["read_len"], e["k"], e["canonical"],
                                             e["num_pes"], e["owner_bin"])
        blob = _ser(keys, counts)
        assert hashlib.sha256(blob).hexdigest() == e["sha256"], name
        assert n == e["total"] and len(counts) == e["distinct"], name
        if "bin_offsets" in e:
            assert [int(x) for x in offs] == e["bin_offsets"], name
        if "table" in e:
            assert blob == (golden_dir / e["table"]).read_bytes(), name


@pytest.mark.parametrize("k", [21, 31, 63])
def test_owner_bins_partition_the_table(k):
    L = 150 if k <= 32 else 250
    data = random_reads(7, 50, L, alphabet="ACGT", mpad=False)
    P = 64
    full_k, full_c, n, _ = oracle.count(data, L, k)
    allk, allc, n2, offs = oracle.count(data, L, k, num_pes=P, owner_bin=-1)
    assert n2 == n and offs[0] == 0 and offs[-1] == len(allc)
    parts = []
    total = 0
    for b in range(P):
        bk, bc, nb, _ = oracle.count(data, L, k, num_pes=P, owner_bin=b)
        seg = slice(int(offs[b]), int(offs[b + 1]))
        assert np.array_equal(bk, allk[seg]) and np.array_equal(bc, allc[seg])
        parts.append((bk, bc))
        total += nb
    assert total == n
    uk = np.concatenate([p[0] for p in parts])
    uc = np.concatenate([p[1] for p in parts])
    order = np.lexsort(tuple(uk[:, i] for i in range(uk.shape[1])))
    assert np.array_equal(uk[order], full_k) and np.array_equal(uc[order], full_c)


@pytest.mark.parametrize("k", [21, 31, 55, 63])
def test_canonical_identity(k):
    # odd k: no k-mer is its own reverse complement, so
    # canon_count(c) = fwd_count(c) + fwd_count(revcomp(c))
    L = 150 if k <= 32 else 250
    data = random_reads(11, 30, L, alphabet="ACGTACGTACGTN")
    fk, fc, _, _ = oracle.count(data, L, k)
    ck, cc, _, _ = oracle.count(data, L, k, canonical=True)
    val = (lambda r: int(r[0])) if k <= 32 else (lambda r: int(r[0]) | (int(r[1]) << 64))
    fwd = {val(r): int(c) for r, c in zip(fk, fc)}
    canon = {val(r): int(c) for r, c in zip(ck, cc)}
    want = {}
    for v, c in fwd.items():
        m = min(v, brute.revcomp_value(v, k))
        want[m] = want.get(m, 0) + c
    assert canon == want


def test_k1_last_base_quirk():
    # kcounter_funcs.cpp:97 drops a run that starts at READLEN-1: with k = 1 the
    # base in the last position is not counted when the one before it breaks
    data = b"ACNA\nACGA\nAAAm\n"
    want = sorted(brute.kmers(data, 4, 1))
    assert want == sorted([1, 0, 1, 0, 3, 1, 1, 1, 1])
    assert sorted(int(x) for x in oracle.extract(data, 4, 1)[:, 0]) == want
    assert oracle.count(b"A\nC\n", 1, 1)[2] == 0   # READLEN 1: nothing is ever emitted


def test_format_errors_and_empty():
    ok = b"ACGT\n" * 3
    assert oracle.count(ok, 4, 2)[2] == 9
    assert oracle.count(ok[:-1], 4, 2)[2] == 9           # missing final newline accepted
    for bad in (ok[:-2], b"ACGTX" + ok[5:], b"\x00CGT\n" + ok[5:]):
        with pytest.raises(oracle.OracleError) as e:
            oracle.count(bad, 4, 2)
        assert e.value.status == 2
    k, c, n, _ = oracle.count(b"", 150, 31)
    assert n == 0 and len(c) == 0
    with pytest.raises(oracle.OracleError):
        oracle.count(ok, 4, 65)


def test_threaded_shape_matches_single_pe():
    data = random_reads(3, 3000, 150, alphabet="ACGTACGTACGTN", mpad=False)
    for canonical in (False, True):
        keys, counts, n, _ = oracle.count(data, 150, 31, canonical=canonical)
        for T in (1, 3, 8):
            n2, d2, cs, el = oracle.count_mt(data, 150, 31, T, canonical=canonical)
            assert (n2, d2) == (n, len(counts))
            assert cs == oracle.checksum(keys, counts)


@pytest.mark.parametrize("k", [31, 63])
def test_threaded_shape_two_word_keys_and_owner_bins(k):
    # the threaded count of one owner bin (and of all bins) equals the single-PE
    # table of that bin; the all-bins checksum is the P=1 table's (each k-mer
    # has one owner)
    L = 250
    data = random_reads(13, 600, L, alphabet="ACGTACGTACGTN", mpad=False)
    seed = 0x1234567
    fk, fc, fn, _ = oracle.count(data, L, k)
    for T in (1, 5):
        n2, d2, cs, _ = oracle.count_mt(data, L, k, T, num_pes=64, owner_bin=-1)
        assert (n2, d2, cs) == (fn, len(fc), oracle.checksum(fk, fc))
        for b in (0, 17):
            bk, bc, bn, _ = oracle.count(data, L, k, num_pes=64, owner_bin=b, seed=seed)
            n3, d3, cs3, _ = oracle.count_mt(data, L, k, T, num_pes=64, owner_bin=b, seed=seed)
            assert (n3, d3, cs3) == (bn, len(bc), oracle.checksum(bk, bc))


@pytest.mark.parametrize("k", [21, 55])
def test_key_array_count_and_threaded_extract(k):
    L = 150
    data = random_reads(17, 900, L, alphabet="ACGTACGTACGTN", mpad=True)
    keys = oracle.extract(data, L, k, canonical=True)
    for T in (1, 4, 7):
        assert np.array_equal(oracle.extract_mt(data, L, k, T, canonical=True), keys)
    # counting the extracted keys is counting the reads
    fk, fc, fn, _ = oracle.count(data, L, k, canonical=True)
    assert fn == keys.shape[0]
    for T in (1, 6):
        d, cs, _ = oracle.count_keys_mt(keys, k, T)
        assert (d, cs) == (len(fc), oracle.checksum(fk, fc))
    d, cs, _ = oracle.count_keys_mt(np.zeros((0, keys.shape[1]), np.uint64), k, 3)
    assert (d, cs) == (0, 0)


def test_owner_seed_matches_bruteforce():
    # a named seed changes which bin a k-mer belongs to; 0 stands for the
    # reference's seed (kcounter.cpp:52)
    data = random_reads(19, 40, 60, alphabet="ACGTACGTACGTACGTN", mpad=False)
    for k in (15, 40):
        for seed in (0x5EED, 2**64 - 1):
            for b in (0, 3):
                keys, counts, n, _ = oracle.count(data, 60, k, num_pes=5, owner_bin=b, seed=seed)
                want = brute.count(data, 60, k, num_pes=5, owner_bin=b, seed=seed)
                got = [(int(r[0]) | (int(r[1]) << 64) if keys.shape[1] == 2 else int(r[0]), int(c))
                       for r, c in zip(keys, counts)]
                assert got == want
        a = oracle.count(data, 60, k, num_pes=5, owner_bin=-1)
        z = oracle.count(data, 60, k, num_pes=5, owner_bin=-1, seed=oracle.OWNER_SEED)
        o = oracle.count(data, 60, k, num_pes=5, owner_bin=-1, seed=0x5EED)
        assert np.array_equal(a[0], z[0]) and np.array_equal(a[3], z[3])
        assert a[3][-1] > 50 and not np.array_equal(a[3], o[3])


@pytest.mark.parametrize("target", ["oracle_asan", "oracle_tsan"])
def test_sanitized(target, tmp_path):
    """Every oracle entry point over messy records (tests/oracle_sanitize.c),
    built with AddressSanitizer + UndefinedBehaviorSanitizer and, for the
    threaded count and extraction, ThreadSanitizer: no report, and the
    cross-checks between the entry points hold."""
    import os
    import subprocess
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    out = tmp_path / "_build"
    r = subprocess.run(["make", "-s", "-f", str(root / "oracle" / "Makefile"), f"OUT={out}", str(out / target)],
                       capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr and "cannot find" in r.stderr:
        pytest.skip(f"{target}: sanitizer runtime not installed")
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")
    run = subprocess.run([str(out / target)], capture_output=True, text=True, env=env, timeout=600)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "all checks passed" in run.stdout
