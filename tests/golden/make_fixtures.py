"""Write the golden fixtures under tests/golden/.

    python tests/golden/make_fixtures.py [--reference /root/reference]

1. char2base_lut.bin: the 256 output bytes of the reference's char2base table
   (dakc/src/common/common.hpp:72-92), read as data from the reference header
   when it is present; this is the golden vector of that function.
2. reads_*.txt: small seeded read sets in the reference's record format with
   repeats (counts > 1), lower case, N runs, stray bytes and M padding.
3. expected.json (+ table_*.bin for a few cases): tables computed by the
   brute-force restatement in tests/brute.py, which shares no code with the C
   oracle or the HIP kernels.  Serialized as little-endian u64 kmer (or u64 lo,
   u64 hi) + u32 count per entry; digests are SHA-256 of that byte string.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import random
import re
import struct
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import brute  # noqa: E402


def lut_from_reference(ref: Path) -> bytes:
    text = (ref / "dakc/src/common/common.hpp").read_text()
    m = re.search(r"base_map\[256\]\s*=\s*\{(.*?)\};", text, re.S)
    body = re.sub(r"/\*.*?\*/", "", m.group(1))
    vals = [int(v, 16) for v in re.findall(r"0x[0-9A-Fa-f]+", body)]
    assert len(vals) == 256, len(vals)
    return bytes(vals)


def make_reads(seed: int, nreads: int, width: int, min_len: int, genome_len: int) -> bytes:
    rng = random.Random(seed)
    genome = "".join(rng.choice("ACGT") for _ in range(genome_len))
    out = []
    for _ in range(nreads):
        ln = rng.randint(min_len, width)
        st = rng.randint(0, genome_len - ln)
        s = list(genome[st:st + ln])
        if rng.random() < 0.5:
            s = list("".join(s)[::-1].translate(str.maketrans("ACGT", "TGCA")))
        for i in range(ln):
            u = rng.random()
            if u < 0.01:
                s[i] = rng.choice("ACGT")          # substitution
            elif u < 0.015:
                s[i] = s[i].lower()                # soft-masked base
            elif u < 0.017:
                s[i] = rng.choice("NRY \r-")      # breaks the run
        if rng.random() < 0.1:                      # an N run
            a = rng.randint(0, ln - 1)
            for i in range(a, min(ln, a + rng.randint(1, 10))):
                s[i] = "N"
        if rng.random() < 0.05 and ln > 10:         # an early 'm' ends the read
            s[rng.randint(1, ln - 1)] = "m"
        rec = "".join(s) + "M" * (width - ln)
        out.append(rec + "\n")
    return "".join(out).encode("latin-1")


def serialize(table: list[tuple[int, int]], k: int) -> bytes:
    b = bytearray()
    for v, c in table:
        if k <= 32:
            b += struct.pack("<QI", v, c)
        else:
            b += struct.pack("<QQI", v & ((1 << 64) - 1), v >> 64, c)
    return bytes(b)


CASES = [
    # (reads file, read_len, k, canonical, P, owner_bin, keep full table)
    ("reads_a.txt", 150, 21, False, 1, -1, False),
    ("reads_a.txt", 150, 21, True, 1, -1, False),
    ("reads_a.txt", 150, 31, False, 1, -1, True),
    ("reads_a.txt", 150, 31, True, 1, -1, False),
    ("reads_a.txt", 150, 31, False, 64, 0, False),
    ("reads_a.txt", 150, 31, False, 64, -1, False),
    ("reads_a.txt", 150, 32, False, 1, -1, False),
    ("reads_b.txt", 250, 55, False, 1, -1, False),
    ("reads_b.txt", 250, 55, True, 1, -1, True),
    ("reads_b.txt", 250, 63, False, 1, -1, False),
    ("reads_b.txt", 250, 63, True, 1, -1, False),
    ("reads_b.txt", 250, 63, False, 64, 0, False),
    ("reads_b.txt", 250, 63, False, 64, -1, False),
    ("reads_b.txt", 250, 64, False, 1, -1, False),
]


def case_name(f, L, k, canon, P, b):
    return f"{f.split('.')[0]}_k{k}_{'canon' if canon else 'fwd'}_P{P}_b{b}"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    ref = Path(a.reference)
    if (ref / "dakc/src/common/common.hpp").exists():
        (HERE / "char2base_lut.bin").write_bytes(lut_from_reference(ref))
    files = {
        "reads_a.txt": make_reads(20240601, 64, 150, 100, 3000),
        "reads_b.txt": make_reads(20240602, 40, 250, 180, 4000),
    }
    for name, data in files.items():
        (HERE / name).write_bytes(data)
    expected = {}
    for f, L, k, canon, P, b, keep in CASES:
        data = files[f]
        table = brute.count(data, L, k, canon, P, b)
        blob = serialize(table, k)
        nm = case_name(f, L, k, canon, P, b)
        e = {"reads": f, "read_len": L, "k": k, "canonical": canon, "num_pes": P, "owner_bin": b,
             "distinct": len(table), "total": sum(c for _, c in table),
             "sha256": hashlib.sha256(blob).hexdigest()}
        if P > 1 and b < 0:
            owners = [brute.owner(v, k, P) for v, _ in table]
            e["bin_offsets"] = [sum(1 for o in owners if o < p) for p in range(P + 1)]
        if keep:
            (HERE / f"table_{nm}.bin").write_bytes(blob)
            e["table"] = f"table_{nm}.bin"
        expected[nm] = e
    (HERE / "expected.json").write_text(json.dumps(expected, indent=1, sort_keys=True) + "\n")
    print(f"wrote {len(expected)} cases")


if __name__ == "__main__":
    main()
