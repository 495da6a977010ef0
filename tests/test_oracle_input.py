"""The oracle's input conversion (FASTQ or newline-delimited reads to the
fixed-width records of syngen/fq2txtmaker.sh), checked against hand-written
known answers and an independent statement of the awk pipeline in Python
bytes operations."""
import random

import numpy as np
import pytest

import oracle


def awk_fq2txt(data: bytes, fastq: bool, read_len: int = 0):
    """awk 'NR%4==2' | printf "%-Ns" | sed 's/ /M/g', in bytes operations:
    lines split at '\\n' only, a final '\\n' ends the last line."""
    lines = data.split(b"\n")
    if lines[-1] == b"":
        lines.pop()
    sel = lines[1::4] if fastq else lines
    width = read_len or max([len(x) for x in sel] + [1])
    return b"".join(x.ljust(width, b" ").replace(b" ", b"M") + b"\n" for x in sel), width


KNOWN = [
    # (input, fastq, records, width)
    (b"@r1\nACGT\n+\nIIII\n@r2\nAC\n+\nII\n", True, b"ACGT\nACMM\n", 4),
    (b"@r1\nACGT\n+\nIIII", True, b"ACGT\n", 4),                 # no final newline
    (b"@r1\nAC GT\n+\nIIIII\n", True, b"ACMGT\n", 5),            # a space becomes M
    (b"@r1\nACGT\r\n+\nIIII\r\n", True, b"ACGT\r\n", 5),         # CR stays part of the line
    (b"@r1\nACGT\n+\n", True, b"ACGT\n", 4),                     # a cut record still gives its read
    (b"@r1\n", True, b"", 1),                                    # no read line
    (b"", True, b"", 1),
    (b"ACGT\nAC\n", False, b"ACGT\nACMM\n", 4),
    (b"ACGT\n\nA", False, b"ACGT\nMMMM\nAMMM\n", 4),             # empty line, no final newline
    (b"\n\n", False, b"M\nM\n", 1),                              # only empty lines: width 1
]


@pytest.mark.parametrize("data,fastq,want,width", KNOWN)
def test_known_answers(data, fastq, want, width):
    out, L = oracle.convert(data, fastq)
    assert bytes(out) == want and L == width


def test_read_len_pads_wider_and_refuses_longer():
    out, L = oracle.convert(b"@a\nACG\n+\nIII\n", True, read_len=6)
    assert bytes(out) == b"ACGMMM\n" and L == 6
    with pytest.raises(oracle.OracleError) as e:
        oracle.convert(b"@a\nACGTACGT\n+\nIIIIIIII\n", True, read_len=4)
    assert e.value.status == 2  # format


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("fastq", [True, False])
def test_matches_awk_pipeline(seed, fastq):
    rng = random.Random(seed)
    recs = []
    for i in range(rng.randint(0, 300)):
        n = rng.choice([0, 1, 5, 99, 150, 250, rng.randint(0, 300)])
        seq = bytes(rng.choice(b"ACGTNacgt M\r") if rng.random() < 0.05 else rng.choice(b"ACGT") for _ in range(n))
        recs.append(b"@read%d len=%d\n%s\n+\n%s\n" % (i, n, seq, b"I" * n) if fastq else seq + b"\n")
    data = b"".join(recs)
    if data and rng.random() < 0.5:
        data = data[:-1]  # drop the final newline
    want, width = awk_fq2txt(data, fastq)
    out, L = oracle.convert(data, fastq)
    assert bytes(out) == want and L == width
