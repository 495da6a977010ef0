// dk_gen.hpp -- counter-based synthetic read generator (host and device).
//
// Every output byte is a pure function of (parameters, byte index), so the
// host and the GPU produce identical buffers and any slice can be generated
// on its own.  Streams of a SplitMix64-style hash (mix64 of a per-stream key
// plus a Weyl sequence) supply the randomness:
//   - i.i.d. reads: base j of read r = 2 bits of hash(IID, r*W + j/32);
//   - genome reads: genome base i = 2 bits of hash(GENOME, i/32); read r starts
//     at mulhi(hash(START, r), G - len + 1), strand from hash(STRAND, r);
//   - substitution: hash(ERR, r*L + j) high word < p*2^32 -> base + 1 + (low % 3);
//   - N runs: a run of 1 + (low16 % n_run_max) N's starts at j when
//     hash(NRUN, r*L + j) high word < q*2^32;
//   - variable length: len = min_len + hash(LEN, r) % (L - min_len + 1), then M padding.
// Letters are "ACGT"[code] (complement = code ^ 3).
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define DK_HD __host__ __device__ __forceinline__
#else
#define DK_HD inline
#endif

namespace dk {

struct GenCore {
  uint64_t key_iid, key_genome, key_start, key_strand, key_err, key_nrun, key_len;
  uint64_t num_reads, first_read, genome_len;
  uint32_t sub_thr, n_thr;
  uint32_t L, min_len, n_run_max, both_strands;
};

DK_HD uint64_t gen_mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}
DK_HD uint64_t gen_stream_key(uint64_t seed, uint64_t stream) {
  return gen_mix64(seed + (stream + 1) * 0xD1B54A32D192ED03ULL);
}
DK_HD uint64_t gen_h(uint64_t key, uint64_t i) { return gen_mix64(key + (i + 1) * 0x9E3779B97F4A7C15ULL); }
DK_HD uint64_t gen_mulhi(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

DK_HD uint32_t gen_read_len(const GenCore &g, uint64_t r) {
  if (g.min_len >= g.L) return g.L;
  return g.min_len + (uint32_t)(gen_h(g.key_len, r) % (uint64_t)(g.L - g.min_len + 1));
}

struct GenRead {
  uint32_t len;
  uint32_t rev;
  uint64_t start;
};

DK_HD GenRead gen_read_info(const GenCore &g, uint64_t r) {
  GenRead ri;
  ri.len = gen_read_len(g, r);
  ri.start = 0;
  ri.rev = 0;
  if (g.genome_len) {
    ri.start = gen_mulhi(gen_h(g.key_start, r), g.genome_len - ri.len + 1);
    ri.rev = g.both_strands ? (uint32_t)(gen_h(g.key_strand, r) & 1) : 0u;
  }
  return ri;
}

// base code (0..3, "ACGT") of position j of read r, before N runs
DK_HD uint32_t gen_base(const GenCore &g, uint64_t r, uint32_t j, const GenRead &ri) {
  uint32_t code;
  if (g.genome_len == 0) {
    const uint64_t wpr = (g.L + 31) / 32;
    code = (uint32_t)(gen_h(g.key_iid, r * wpr + (j >> 5)) >> (2 * (j & 31))) & 3u;
  } else {
    const uint64_t pos = ri.rev ? ri.start + ri.len - 1 - j : ri.start + j;
    code = (uint32_t)(gen_h(g.key_genome, pos >> 5) >> (2 * (pos & 31))) & 3u;
    if (ri.rev) code ^= 3u;
  }
  if (g.sub_thr) {
    const uint64_t e = gen_h(g.key_err, r * g.L + j);
    if ((uint32_t)(e >> 32) < g.sub_thr) code = (code + 1 + (uint32_t)(e & 0xFFFFFFFFu) % 3u) & 3u;
  }
  return code;
}

DK_HD bool gen_is_n(const GenCore &g, uint64_t r, uint32_t j) {
  if (!g.n_thr) return false;
  const uint32_t lo = j + 1 > g.n_run_max ? j + 1 - g.n_run_max : 0;
  for (uint32_t s = lo; s <= j; ++s) {
    const uint64_t u = gen_h(g.key_nrun, r * g.L + s);
    if ((uint32_t)(u >> 32) < g.n_thr) {
      const uint32_t run = 1 + (uint32_t)(u & 0xFFFFu) % g.n_run_max;
      if (s + run > j) return true;
    }
  }
  return false;
}

// byte at offset `off` of the generated buffer
DK_HD uint8_t gen_byte(const GenCore &g, uint64_t off) {
  const uint64_t stride = (uint64_t)g.L + 1;
  const uint64_t r = g.first_read + off / stride;
  const uint32_t j = (uint32_t)(off % stride);
  if (j == g.L) return '\n';
  const GenRead ri = gen_read_info(g, r);
  if (j >= ri.len) return 'M';
  if (gen_is_n(g, r, j)) return 'N';
  return (uint8_t)(0x54474341u >> (8 * gen_base(g, r, j, ri)));  // "ACGT"[code]
}

}  // namespace dk
