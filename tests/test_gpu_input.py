"""FASTQ and newline-delimited input converted on the GPU (dakc_convert_reads,
and dakc_count with those formats) against the oracle's restatement of
syngen/fq2txtmaker.sh; at a larger size, FASTQ written from generated reads
converts back to the generator's own fixed-width buffer byte for byte."""
import random

import numpy as np
import pytest

import oracle
from dakc_amd import KmerCounter, _lib
from dakc_amd._lib import DakcError
from test_oracle_input import KNOWN

pytestmark = pytest.mark.gpu


def convert_on_device(kc, data: bytes, fmt, read_len=0):
    d = kc.device_alloc(max(len(data), 4))
    try:
        if data:
            kc.h2d(d, np.frombuffer(data, dtype=np.uint8))
        out, n, L = kc.convert_reads(d, len(data), fmt, read_len)
        try:
            got = kc.d2h(out, n, np.uint8) if n else np.zeros(0, np.uint8)
        finally:
            kc.device_free(out)
        return bytes(got), L
    finally:
        kc.device_free(d)


def random_input(seed, fastq, nrec=2000):
    rng = random.Random(seed)
    recs = []
    for i in range(nrec):
        n = rng.choice([0, 31, 100, 150, 250, rng.randint(0, 400)])
        seq = bytes(rng.choice(b"ACGTN M\r") if rng.random() < 0.02 else rng.choice(b"ACGT") for _ in range(n))
        recs.append(b"@r%d\n%s\n+\n%s\n" % (i, seq, b"#" * n) if fastq else seq + b"\n")
    data = b"".join(recs)
    return data[:-1] if seed % 2 else data  # odd seeds: no final newline


def test_convert_known_answers():
    with KmerCounter() as kc:
        for data, fastq, want, width in KNOWN:
            got, L = convert_on_device(kc, data, _lib.FMT_FASTQ if fastq else _lib.FMT_LINES)
            assert got == want and L == width, data


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("fastq", [True, False])
def test_convert_matches_oracle(seed, fastq):
    data = random_input(seed, fastq)
    want, width = oracle.convert(data, fastq)
    with KmerCounter() as kc:
        got, L = convert_on_device(kc, data, _lib.FMT_FASTQ if fastq else _lib.FMT_LINES)
        assert got == bytes(want) and L == width
        got, L = convert_on_device(kc, data, _lib.FMT_FASTQ if fastq else _lib.FMT_LINES, read_len=500)
        assert got == bytes(oracle.convert(data, fastq, 500)[0]) and L == 500


@pytest.mark.parametrize("fastq", [True, False])
@pytest.mark.parametrize("k", [21, 31, 55])
def test_count_fastq_and_lines(fastq, k):
    data = random_input(10 + k, fastq, nrec=3000)
    fixed, L = oracle.convert(data, fastq)
    keys, counts, n, _ = oracle.count(fixed, L, k)
    fmt = _lib.FMT_FASTQ if fastq else _lib.FMT_LINES
    with KmerCounter(k=k) as kc:
        t = kc.count(np.frombuffer(data, np.uint8), format=fmt, read_len=0)  # host bytes
        assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)
        d = kc.device_alloc(len(data))
        kc.h2d(d, np.frombuffer(data, np.uint8))
        t = kc.count(None, dev_ptr=d, nbytes=len(data), format=fmt, read_len=0)  # device bytes
        kc.device_free(d)
        assert t.total == n and np.array_equal(t.keys, keys) and np.array_equal(t.counts, counts)


def test_fastq_round_trip_at_size():
    """FASTQ written from 200k generated reads (100-250 bp, M-padded to 250)
    converts back to the generator's buffer byte for byte, and counts alike."""
    p = _lib.gen_params(4242, 200_000, 250, min_len=100, genome_len=1_000_000, sub_rate=0.01, n_rate=0.001)
    fixed = _lib.generate_reads(p)
    recs = fixed.reshape(-1, 251)[:, :250]
    lens = 250 - np.argmax(recs[:, ::-1] != ord("M"), axis=1)
    assert lens.max() == 250
    parts = []
    for i, (r, n) in enumerate(zip(recs, lens)):
        s = r[:n].tobytes()
        parts.append(b"@gen%d\n%s\n+\n%s\n" % (i, s, b"F" * n))
    fq = b"".join(parts)
    with KmerCounter(k=31, read_len=250) as kc:
        got, L = convert_on_device(kc, fq, _lib.FMT_FASTQ)
        assert L == 250 and got == fixed.tobytes()
        a = kc.count(np.frombuffer(fq, np.uint8), format=_lib.FMT_FASTQ, read_len=0)
        b = kc.count(fixed)
        assert a.total == b.total and np.array_equal(a.keys, b.keys) and np.array_equal(a.counts, b.counts)


def test_convert_errors():
    with KmerCounter() as kc:
        data = b"@a\nACGTACGT\n+\nIIIIIIII\n"
        with pytest.raises(DakcError) as e:
            convert_on_device(kc, data, _lib.FMT_FASTQ, read_len=4)  # a read longer than the width
        assert e.value.status == 2  # DAKC_E_FORMAT
        d = kc.device_alloc(64)
        try:
            with pytest.raises(DakcError):
                kc.convert_reads(d + 1, 8, _lib.FMT_FASTQ)  # not 4-byte aligned
            with pytest.raises(DakcError):
                kc.convert_reads(d, 8, _lib.FMT_FIXED)  # nothing to convert
        finally:
            kc.device_free(d)
