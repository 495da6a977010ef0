"""The oracle's after-the-count functions (count-frequency histogram and its
minimum bucket, solid-k-mer filter, point queries), each checked against an
independent statement: numpy's bincount and boolean masks, Python's min (the
first minimum, as std::min_element), and dict lookups."""
import numpy as np
import pytest

import oracle


def sorted_table(seed, d, kw=1, num_bins=0):
    """A sorted (per owner bin) table of d distinct keys with a genome-like
    count mix: many 1s, a peak near 30, a long tail."""
    rng = np.random.default_rng(seed)
    if kw == 1:
        keys = np.unique(rng.integers(0, 2**62, size=d + d // 8, dtype=np.uint64))[:d].reshape(-1, 1)
    else:
        raw = rng.integers(0, 2**62, size=(d + d // 8, 2), dtype=np.uint64)
        raw = np.unique(raw, axis=0)[:d]
        keys = raw[np.lexsort((raw[:, 0], raw[:, 1]))]  # (lo, hi) ordered by hi, then lo
    kind = rng.random(d)
    counts = np.where(kind < 0.6, 1, np.where(kind < 0.95, rng.poisson(30, d) + 1,
                                              rng.integers(60, 5000, d))).astype(np.uint32)
    offs = None
    if num_bins:
        own = np.array([(oracle.murmur64a_u64(int(k[0])) if kw == 1 else
                         oracle.murmur64a_u128(int(k[0]), int(k[1]))) % num_bins for k in keys])
        order = np.argsort(own, kind="stable")  # keys stay sorted inside each bin
        keys, counts, own = keys[order], counts[order], own[order]
        offs = np.searchsorted(own, np.arange(num_bins + 1)).astype(np.uint64)
    return keys, counts, offs


@pytest.mark.parametrize("nb", [1, 2, 3, 8, 40, 100, 7000])
def test_histogram_matches_bincount_and_first_minimum(nb):
    _, counts, _ = sorted_table(1, 20000)
    hist, m = oracle.histogram(counts, nb)
    want = np.bincount(counts[counts < nb], minlength=nb)[:nb]
    assert np.array_equal(hist, want)
    # std::min_element over [1, nb): Python's min also keeps the first minimum;
    # an empty range (nb == 1) gives its end, index 1
    assert m == (1 if nb == 1 else min(range(1, nb), key=lambda c: int(want[c])))
    # every entry is binned or at/above nb; the binned counts add up
    assert int(hist.sum()) + int((counts >= nb).sum()) == counts.size
    assert int((hist * np.arange(nb, dtype=np.uint64)).sum()) + int(counts[counts >= nb].sum()) == \
        int(counts.sum())


def test_histogram_ties_zero_bucket_and_no_buckets():
    # hist[1..5] = 5, 2, 7, 2, 9: two minima, the first (2) wins; hist[0] = 0
    # is never a candidate; counts >= nb are not binned
    counts = np.array([1] * 5 + [2] * 2 + [3] * 7 + [4] * 2 + [5] * 9 + [6, 99], dtype=np.uint32)
    hist, m = oracle.histogram(counts, 6)
    assert list(hist) == [0, 5, 2, 7, 2, 9] and m == 2
    _, m0 = oracle.histogram(counts, 0)
    assert m0 == 0
    _, m1 = oracle.histogram(counts, 1)
    assert m1 == 1


@pytest.mark.parametrize("kw", [1, 2])
@pytest.mark.parametrize("min_count", [0, 1, 2, 17, 31, 10**9])
def test_filter_matches_mask(kw, min_count):
    keys, counts, _ = sorted_table(2 + kw, 5000, kw)
    k2, c2, total, _ = oracle.filter_table(keys, counts, min_count)
    mask = counts >= min_count
    assert np.array_equal(k2, keys[mask]) and np.array_equal(c2, counts[mask])
    assert total == int(counts[mask].sum())


def test_filter_moves_bin_offsets():
    keys, counts, offs = sorted_table(5, 3000, 1, num_bins=16)
    offs[3] = offs[4]  # an empty bin too (its entries counted with bin 2)
    for m in (2, 30, 10**9):
        _, _, _, new = oracle.filter_table(keys, counts, m, offs)
        keep = counts >= m
        assert list(new) == [int(keep[:int(o)].sum()) for o in offs]


@pytest.mark.parametrize("kw", [1, 2])
@pytest.mark.parametrize("num_bins", [0, 7])
def test_query_matches_dict(kw, num_bins):
    keys, counts, offs = sorted_table(7 + kw, 4000, kw, num_bins)
    lut = {tuple(int(x) for x in k): int(c) for k, c in zip(keys, counts)}
    rng = np.random.default_rng(9)
    present = keys[rng.integers(0, len(keys), 500)]
    absent = rng.integers(0, 2**62, size=(500, kw), dtype=np.uint64)
    q = np.concatenate([present, absent])
    got = oracle.query(keys, counts, q, offs)
    assert list(got) == [lut.get(tuple(int(x) for x in k), 0) for k in q]
