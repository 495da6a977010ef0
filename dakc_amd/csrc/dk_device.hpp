// dk_device.hpp -- device-side building blocks shared by the counting kernels.
//
// Encoding (common.hpp:72-92 of the reference): C/c=0, A/a=1, T/t=2, G/g=3,
// M/m ends a read, every other byte breaks a run.  k-mers are packed MSB-first
// (first base in the highest bits, kcounter_funcs.cpp:25-38) over 2k bits:
// one u64 for k <= 32, a (lo, hi) pair for 33 <= k <= 64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dk_status.hpp"

namespace dk {

struct K2 {
  uint64_t lo, hi;
};

#define DK_DEV __device__ __forceinline__

// ------------------------------------------------------------------- keys --
DK_DEV bool key_eq(uint64_t a, uint64_t b) { return a == b; }
DK_DEV bool key_eq(const K2 &a, const K2 &b) { return a.lo == b.lo && a.hi == b.hi; }
DK_DEV bool key_lt(uint64_t a, uint64_t b) { return a < b; }
DK_DEV bool key_lt(const K2 &a, const K2 &b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }

// 8-bit digit at bit offset `shift` of the 2k-bit value
DK_DEV uint32_t key_digit(uint64_t x, uint32_t shift) { return (uint32_t)(x >> shift) & 0xFFu; }
DK_DEV uint32_t key_digit(const K2 &x, uint32_t shift) {
  return shift >= 64 ? (uint32_t)(x.hi >> (shift - 64)) & 0xFFu : (uint32_t)(x.lo >> shift) & 0xFFu;
}

// kcounter.cpp:29-48 -- MurmurHash64A of one 8-byte key
DK_DEV uint64_t murmur64a(uint64_t key, uint64_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ULL;
  uint64_t h = seed ^ (8 * m);
  uint64_t k = key * m;
  k ^= k >> 47;
  k *= m;
  h ^= k;
  h *= m;
  h ^= h >> 47;
  h *= m;
  h ^= h >> 47;
  return h;
}
// MurmurHash64A over the 16 little-endian bytes lo||hi (2x64-bit keys)
DK_DEV uint64_t murmur64a(const K2 &key, uint64_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ULL;
  uint64_t h = seed ^ (16 * m);
  uint64_t k = key.lo * m;
  k ^= k >> 47;
  k *= m;
  h ^= k;
  h *= m;
  k = key.hi * m;
  k ^= k >> 47;
  k *= m;
  h ^= k;
  h *= m;
  h ^= h >> 47;
  h *= m;
  h ^= h >> 47;
  return h;
}
constexpr uint64_t OWNER_SEED = 0x9E3779B97F4A7C15ULL;  // kcounter.cpp:52, the default seed

// owner PE of a k-mer (kcounter.cpp:50-54 with the job's seed); a power-of-two
// P (as the reference's 64-PE runs) takes the low bits, which is the same
// remainder without the 64-bit division
template <typename KeyT>
DK_DEV uint32_t owner_of(const KeyT &x, uint32_t num_pes, uint64_t seed) {
  const uint64_t h = murmur64a(x, seed);
  if ((num_pes & (num_pes - 1)) == 0) return (uint32_t)h & (num_pes - 1);
  return (uint32_t)(h % (uint64_t)num_pes);
}

// bits [shift, shift + nbits) of the key value, nbits <= 32
DK_DEV uint32_t key_bits(uint64_t x, uint32_t shift, uint32_t nbits) {
  if (!nbits) return 0;
  return (uint32_t)(x >> shift) & (0xFFFFFFFFu >> (32 - nbits));
}
DK_DEV uint32_t key_bits(const K2 &x, uint32_t shift, uint32_t nbits) {
  if (!nbits) return 0;
  uint64_t v;
  if (shift >= 64) v = x.hi >> (shift - 64);
  else v = shift ? (x.lo >> shift) | (x.hi << (64 - shift)) : x.lo;
  return (uint32_t)v & (0xFFFFFFFFu >> (32 - nbits));
}

// --------------------------------------------------------- record bytes --
// 0x80 in each byte of x that is zero, 0 elsewhere (no carries between bytes)
__host__ __device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
  return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}

// Four record bytes decoded at once: w holds them little-endian (the first
// base in the low byte), the first nv (0..4) lie inside the read.  code8 is
// their 2-bit codes, first base in the top bits, 0 for a base that is not
// ACGT or lies past the read; inv4 has bit i set for base i not ACGT (or
// past the read); mmask has 0x80 in the bytes that are 'M' or 'm'.  The same
// as the byte-at-a-time decode: u = b & 0xDF (upper case), tt = ASCII bits
// 1-2 (A0 C1 T2 G3), valid iff u is the letter of tt, code C0 A1 T2 G3.
// The letter and the code of each byte are 4-entry byte lookups by tt
// (v_perm_b32 picks byte tt of its second operand): full-rate instructions
// where the arithmetic form took two quarter-rate multiplies.
struct Dec4 {
  uint32_t code8, inv4, mmask;
};
DK_DEV Dec4 decode4(uint32_t w, uint32_t nv) {
  const uint32_t pm = nv >= 4 ? 0x80808080u : ((1u << (8 * nv)) - 1u) & 0x80808080u;
  const uint32_t u = w & 0xDFDFDFDFu;
  const uint32_t tt = (w >> 1) & 0x03030303u;
  const uint32_t letter = __builtin_amdgcn_perm(0u, 0x47544341u, tt);  // G T C A for tt 3 2 1 0
  const uint32_t okm = zero_bytes(u ^ letter) & pm;
  const uint32_t c = __builtin_amdgcn_perm(0u, 0x03020001u, tt) & ((okm >> 7) * 3u);  // A->1, C->0
  Dec4 d;
  d.code8 = (c * 0x40100401u) >> 24;                    // fields at bits 0, 8, 16, 24
  d.inv4 = ~(((okm >> 7) * 0x00204081u) >> 21) & 0xFu;  // bits 7, 15, 23, 31 gathered
  d.mmask = zero_bytes(u ^ 0x4D4D4D4Du) & pm;
  return d;
}

// reverse the order of the 32 2-bit groups of a word
DK_DEV uint64_t rev2(uint64_t x) {
  x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
  x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
  return __builtin_bswap64(x);
}
// reverse complement over 2k bits (complement of code c is c ^ 3)
DK_DEV uint64_t revcomp(uint64_t x, int k) { return rev2(~x) >> (64 - 2 * k); }
DK_DEV K2 revcomp(const K2 &x, int k) {
  uint64_t nh = rev2(~x.lo), nl = rev2(~x.hi);  // 128-bit reversal
  int s = 128 - 2 * k;                           // 0 <= s <= 62 for 33 <= k <= 64
  K2 r;
  r.lo = s ? (nl >> s) | (nh << (64 - s)) : nl;
  r.hi = nh >> s;
  return r;
}

// key of the window that starts `sh` bits into code word c0 (c1, c2 follow)
template <typename KeyT>
DK_DEV KeyT window_key(uint64_t c0, uint64_t c1, uint64_t c2, uint32_t sh, int k);
template <>
DK_DEV uint64_t window_key<uint64_t>(uint64_t c0, uint64_t c1, uint64_t, uint32_t sh, int k) {
  uint64_t t = sh ? (c0 << sh) | (c1 >> (64 - sh)) : c0;
  return t >> (64 - 2 * k);
}
template <>
DK_DEV K2 window_key<K2>(uint64_t c0, uint64_t c1, uint64_t c2, uint32_t sh, int k) {
  uint64_t th = sh ? (c0 << sh) | (c1 >> (64 - sh)) : c0;
  uint64_t tl = sh ? (c1 << sh) | (c2 >> (64 - sh)) : c1;
  int s = 128 - 2 * k;
  K2 r;
  r.lo = s ? (tl >> s) | (th << (64 - s)) : tl;
  r.hi = th >> s;
  return r;
}

template <typename KeyT>
DK_DEV KeyT canon(const KeyT &x, int k) {
  KeyT r = revcomp(x, k);
  return key_lt(r, x) ? r : x;
}

// ------------------------------------------------------ decoupled look-back --
// Flag word: [63:50] epoch, [49:48] status (1 aggregate, 2 inclusive), [47:0] value.
// A word whose epoch differs from the current pass's is "not yet published",
// so the flag arrays never need clearing between passes.
constexpr uint64_t LB_AGG = 1, LB_INC = 2;
DK_DEV uint64_t lb_pack(uint32_t epoch, uint64_t status, uint64_t v) {
  return ((uint64_t)epoch << 50) | (status << 48) | (v & 0xFFFFFFFFFFFFULL);
}
DK_DEV void lb_store(uint64_t *p, uint64_t w) {
  __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DK_DEV uint64_t lb_load(uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// A look-back wait gives up after this many polls (several seconds, far
// beyond any predecessor's time to publish): it sets ERR_STALL in the job's
// error word and carries on as if the predecessor had published an inclusive
// prefix of 0.  Prefixes can then only come out too small, so every write
// stays inside the buffers, the successors find a published value and end,
// and the host reports DAKC_E_DEVICE instead of the GPU hanging.
constexpr uint32_t LB_SPIN_MAX = 1u << 22;

// Exclusive prefix of `slot` over tiles [0, tile) with stride `stride` words
// between consecutive tiles.  Tiles are taken in ticket order, so every
// predecessor is already resident and will publish: the wait ends, and is
// bounded by LB_SPIN_MAX polls besides.
DK_DEV uint64_t lb_lookback(uint64_t *flags, uint32_t tile, uint32_t stride, uint32_t slot,
                            uint32_t epoch, uint32_t *err) {
  uint64_t excl = 0;
  int64_t t = (int64_t)tile - 1;
  uint32_t polls = 0;
  while (t >= 0) {
    uint64_t w = lb_load(flags + (uint64_t)t * stride + slot);
    uint32_t e = (uint32_t)(w >> 50);
    uint64_t st = (w >> 48) & 3;
    if (e != epoch || st == 0) {
      if (++polls > LB_SPIN_MAX) {
        atomicOr(err, ERR_STALL);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += w & 0xFFFFFFFFFFFFULL;
    if (st == LB_INC) break;
    --t;
  }
  return excl;
}

// lanes below this one in the wavefront
DK_DEV uint64_t lanemask_lt() {
  uint32_t lane = __lane_id();
  return lane ? (~0ULL >> (64 - lane)) : 0ULL;
}

}  // namespace dk
