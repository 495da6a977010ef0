"""Independent brute-force restatement of DAKC's single-PE count, for tiny inputs.

Works on Python strings and big integers only, so it shares no code or bit
tricks with the C oracle or the HIP kernels:
  - records: read_len bytes + '\\n' (fqreader.cpp:43, kcounter_funcs.cpp:102);
  - a read ends at its first 'M'/'m' (common.hpp:80,96 map M to 0xF0, which
    ends the record loop at kcounter_funcs.cpp:92-94);
  - maximal runs of ACGT/acgt (every other byte maps to 0xFF and splits,
    common.hpp:72-92, kcounter_funcs.cpp:88-91);
  - the run still open at the record's end only if it starts before
    read_len - 1 (kcounter_funcs.cpp:97; matters for k = 1 only);
  - every length-k substring of a run, read as a base-4 number with the
    digits C0 A1 T2 G3 (common.hpp:78-81), first base most significant
    (kcounter_funcs.cpp:25-38);
  - owner = MurmurHash64A (the general-length public algorithm) of the key's
    little-endian bytes, mod P (kcounter.cpp:29-54).
"""
from __future__ import annotations

import re
from collections import Counter

M64 = (1 << 64) - 1
DIGIT = {"C": "0", "A": "1", "T": "2", "G": "3"}
COMP = str.maketrans("ACGT", "TGCA")
OWNER_SEED = 0x9E3779B97F4A7C15


class FormatError(ValueError):
    pass


def murmur64a(data: bytes, seed: int) -> int:
    """Appleby's MurmurHash64A for any byte string (blocks, tail, finaliser)."""
    m, r = 0xC6A4A7935BD1E995, 47
    h = (seed ^ (len(data) * m)) & M64
    nblocks = len(data) // 8
    for b in range(nblocks):
        k = int.from_bytes(data[8 * b:8 * b + 8], "little")
        k = (k * m) & M64
        k ^= k >> r
        k = (k * m) & M64
        h ^= k
        h = (h * m) & M64
    tail = data[8 * nblocks:]
    if tail:
        for i in range(len(tail) - 1, -1, -1):
            h ^= tail[i] << (8 * i)
        h = (h * m) & M64
    h ^= h >> r
    h = (h * m) & M64
    h ^= h >> r
    return h


def smhasher_verification(hashfn) -> int:
    """SMHasher's VerificationTest for a 64-bit hash(key_bytes, seed)."""
    key = bytearray(256)
    hashes = bytearray()
    for i in range(256):
        key[i] = i
        hashes += hashfn(bytes(key[:i]), 256 - i).to_bytes(8, "little")
    return hashfn(bytes(hashes), 0) & 0xFFFFFFFF


def records(buf: bytes, read_len: int) -> list[bytes]:
    stride = read_len + 1
    if len(buf) == 0:
        return []
    n, rem = divmod(len(buf), stride)
    if rem == read_len:
        n += 1
    elif rem:
        raise FormatError("not a whole number of records")
    recs = []
    for i in range(n):
        rec = buf[i * stride:i * stride + read_len]
        if rec[:1] == b"\x00":
            raise FormatError("record starts with NUL")
        sep = buf[i * stride + read_len:i * stride + read_len + 1]
        if sep and sep != b"\n":
            raise FormatError("separator is not a newline")
        recs.append(rec)
    return recs


def value(s: str) -> int:
    return int("".join(DIGIT[c] for c in s), 4)


def kmers(buf: bytes, read_len: int, k: int, canonical: bool = False,
          num_pes: int = 1, owner_bin: int = -1, seed: int = OWNER_SEED) -> list[int]:
    out = []
    for rec in records(buf, read_len):
        s = rec.decode("latin-1")
        cut = min([i for i, ch in enumerate(s) if ch in "Mm"], default=len(s))
        for mt in re.finditer(r"[ACGTacgt]+", s[:cut]):
            # kcounter_funcs.cpp:97 -- the run still open at the end of the
            # record is emitted only if it starts before read_len - 1
            if mt.end() == read_len and mt.start() >= read_len - 1:
                continue
            run = mt.group().upper()
            for i in range(len(run) - k + 1):
                sub = run[i:i + k]
                v = value(sub)
                if canonical:
                    v = min(v, value(sub[::-1].translate(COMP)))
                if num_pes > 1 and owner_bin >= 0 and owner(v, k, num_pes, seed) != owner_bin:
                    continue
                out.append(v)
    return out


def owner(v: int, k: int, num_pes: int, seed: int = OWNER_SEED) -> int:
    nbytes = 8 if k <= 32 else 16
    return murmur64a(v.to_bytes(nbytes, "little"), seed) % num_pes


def count(buf: bytes, read_len: int, k: int, canonical: bool = False,
          num_pes: int = 1, owner_bin: int = -1, seed: int = OWNER_SEED) -> list[tuple[int, int]]:
    """Sorted (kmer, count); with all P bins, ordered by (owner, kmer)."""
    c = Counter(kmers(buf, read_len, k, canonical, num_pes, owner_bin, seed))
    if num_pes > 1 and owner_bin < 0:
        return sorted(c.items(), key=lambda kv: (owner(kv[0], k, num_pes, seed), kv[0]))
    return sorted(c.items())


def revcomp_value(v: int, k: int) -> int:
    s = "".join("CATG"[(v >> (2 * (k - 1 - j))) & 3] for j in range(k))
    return value(s[::-1].translate(COMP))


def hitter(buf: bytes, read_len: int, k: int, bucket_size: int, canonical: bool = False, num_pes: int = 1,
           owner_bin: int = -1, seed: int = OWNER_SEED):
    """HITTER=1 as sets: buckets of whole records (a record joins the bucket
    while the bucket holds at most bucket_size - read_len k-mers); a k-mer is
    heavy when it is seen three or more times in some bucket; light and heavy
    entries carry the k-mer's total count.  Returns ({key: count} light,
    {key: count} heavy, stats) for the owners kept (all, or owner_bin)."""
    recs = list(records(buf, read_len))
    per = [kmers(r + b"\n", read_len, k, canonical) for r in recs]
    buckets, cur, n = [], [], 0
    for km in per:
        if n > bucket_size - read_len:
            buckets.append(cur)
            cur, n = [], 0
        cur.extend(km)
        n += len(km)
    if recs:
        buckets.append(cur)
    keep = (lambda v: True) if num_pes <= 1 or owner_bin < 0 else (
        lambda v: owner(v, k, num_pes, seed) == owner_bin)
    heavy_set, total = set(), Counter()
    stats = {"buckets": 0, "normal_keys": 0, "heavy_pairs": 0}
    for bk in buckets:
        if not bk:
            continue
        stats["buckets"] += 1
        for v, c in Counter(bk).items():
            if not keep(v):
                continue
            total[v] += c
            if c >= 3:
                heavy_set.add(v)
                stats["heavy_pairs"] += 1
            else:
                stats["normal_keys"] += c
    light = {v: c for v, c in total.items() if v not in heavy_set}
    heavy = {v: c for v, c in total.items() if v in heavy_set}
    return light, heavy, stats
