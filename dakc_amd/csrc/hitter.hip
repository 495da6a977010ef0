// hitter.hip -- HITTER=1 on the GPU: per-bucket pre-aggregation on the send
// side of a P-PE run, and the owner side's merge into light and heavy tables
// (kcounter.cpp:56-104, 176-245, 325-384; off in the shipped build,
// dakc/Makefile:11).
//
// Send side.  A bucket is what read_till_buf_max (kcounter_funcs.cpp:69-111)
// fills: whole records, taken while the bucket holds at most bucket_size -
// READLEN k-mers.  The chain of bucket starts is sequential in the reference;
// here every record's k-mer count is scanned into a prefix P, each start s
// gets its bucket's end J(s) by a search in P, and the chain from record 0 is
// found by lifting (J^(2^t) for every t), so each bucket finds its own start in
// log steps.  The k-mers themselves come in record order from the partition
// kernels (one bin), so bucket j is the key range [P(start_j), P(start_j+1)).
// One workgroup per bucket counts it in an LDS hash table; a k-mer seen once
// or twice goes out that many times as a plain key, one seen three or more
// times as one (key, count) pair (send2sendbuf, kcounter.cpp:176-210).
//
// Owner side.  The plain keys are counted (the dakc_count_keys path), the
// pair keys too, and the pairs' counts are added into the latter by binary
// search (sort_and_merge_duplicate_kmer_packets, kcounter.cpp:56-104); every
// light run whose k-mer is heavy moves its count to the heavy table
// (binary_search, kcounter.cpp:342-371), and the light table drops it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "dk_api_internal.hpp"

namespace dk {

constexpr uint32_t HIT_NT = 1024;
constexpr uint32_t HIT_MAX_BUCKET = 10240;            // bucket_size limit: a bucket holds fewer keys
constexpr int HIT_ITEMS = HIT_MAX_BUCKET / HIT_NT;    // keys per thread
constexpr uint32_t HIT_HT1 = 12288;                   // one-word table slots (load <= 0.84)
constexpr uint32_t HIT_HT2 = 6144;                    // two-word table slots, per hash class
constexpr int HIT_SPT1 = HIT_HT1 / HIT_NT, HIT_SPT2 = HIT_HT2 / HIT_NT;
constexpr uint32_t SCAN_TILE = 1024;                  // records per workgroup of the count / scan kernels

// meta words of a HITTER run (u64): plain keys, pairs, buckets, error word,
// bucket count, k-mers of the kept owners (plain keys + the pairs' counts)
enum { HM_NORMAL = 0, HM_PAIRS = 1, HM_BUCKETS = 2, HM_ERR = 3, HM_NB = 4, HM_KEPT = 5, HM_WORDS = 8 };

// k-mers of one record (kcounter_funcs.cpp:69-111 with get_kmers :40-67):
// ACGT runs split by any other byte, the record ended by M/m; the run still
// open at the end counts only when it starts before READLEN-1 (:97).
DK_DEV uint32_t record_kmers(const uint8_t *rec, uint32_t L, uint32_t k) {
  uint32_t n = 0, left = 0, i = 0;
  for (; i < L; ++i) {
    const uint32_t u = rec[i] & 0xDFu;  // upper case
    if (u == 'A' || u == 'C' || u == 'G' || u == 'T') continue;
    if (u == 'M') break;
    if (i - left >= k) n += i - left - k + 1;
    left = i + 1;
  }
  if (left + 1 < L && i - left >= k) n += i - left - k + 1;
  return n;
}

DK_DEV uint32_t block_sum(uint32_t v, uint32_t *s_wave) {
  const uint32_t tid = threadIdx.x, lane = tid & 63;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) s_wave[tid >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
  for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += s_wave[w];
  __syncthreads();
  return t;
}

// Exclusive scan over the workgroup (blockDim.x threads, one u32 each).
DK_DEV uint32_t block_excl(uint32_t v, uint32_t *s_wave, uint32_t *total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  uint32_t pre = 0, all = 0;
  for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
    if (w < wave) pre += s_wave[w];
    all += s_wave[w];
  }
  __syncthreads();
  if (total) *total = all;
  return pre + x - v;
}

// per-record k-mer counts and their sums per tile of SCAN_TILE records; the
// record-format checks of the count
__global__ __launch_bounds__(SCAN_TILE) void k_hit_reccount(ReadsDesc rd, uint32_t *c, uint32_t *tile_sums,
                                                            unsigned long long *meta) {
  __shared__ uint32_t s_wave[SCAN_TILE / 64];
  const uint64_t r = (uint64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  uint32_t v = 0;
  if (r < rd.nrec) {
    const uint8_t *rec = rd.reads + r * rd.stride;
    v = record_kmers(rec, rd.L, rd.k);
    uint32_t e = 0;
    if (rec[0] == 0) e |= ERR_NUL_RECORD;
    if ((r + 1) * rd.stride <= rd.nbytes && rec[rd.L] != '\n') e |= ERR_SEPARATOR;
    if (e) atomicOr(reinterpret_cast<unsigned int *>(&meta[HM_ERR]), e);
    c[r] = v;
  }
  const uint32_t t = block_sum(v, s_wave);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = t;
}

// exclusive prefix of the tile sums, one workgroup (each thread a contiguous
// range of tiles)
__global__ __launch_bounds__(1024) void k_hit_scan_top(const uint32_t *tile_sums, uint32_t ntiles,
                                                       uint32_t *tile_off) {
  __shared__ uint32_t s_wave[16];
  const uint32_t per = (ntiles + 1023) / 1024, t0 = threadIdx.x * per;
  uint32_t sum = 0;
  for (uint32_t i = t0; i < t0 + per && i < ntiles; ++i) sum += tile_sums[i];
  uint32_t run = block_excl(sum, s_wave, nullptr);
  for (uint32_t i = t0; i < t0 + per && i < ntiles; ++i) {
    tile_off[i] = run;
    run += tile_sums[i];
  }
}

// P[r] = k-mers of the records before r, r <= nrec
__global__ __launch_bounds__(SCAN_TILE) void k_hit_scan_apply(const uint32_t *c, const uint32_t *tile_off,
                                                              uint64_t nrec, uint32_t *P) {
  __shared__ uint32_t s_wave[SCAN_TILE / 64];
  const uint64_t r = (uint64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  const uint32_t v = r < nrec ? c[r] : 0u;
  const uint32_t ex = block_excl(v, s_wave, nullptr) + tile_off[blockIdx.x];
  if (r < nrec) P[r] = ex;
  if (r + 1 == nrec) P[nrec] = ex + v;
}

// J0[s] = the end of the bucket that starts at record s: the first e > s with
// P[e] > P[s] + bp (the bucket takes records while it holds <= bp k-mers),
// else nrec; J0[nrec] = nrec.  Galloping search from s + 1.
__global__ __launch_bounds__(256) void k_hit_next(const uint32_t *P, uint64_t nrec, uint32_t bp, uint32_t *J0) {
  const uint64_t s = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (s > nrec) return;
  if (s + 1 >= nrec) {
    J0[s] = (uint32_t)nrec;
    return;
  }
  const uint64_t target = (uint64_t)P[s] + bp;
  uint64_t e;
  if (P[s + 1] > target) {
    e = s + 1;
  } else {
    uint64_t lo = s + 1, hi, step = 1;  // P[lo] <= target
    for (;;) {
      const uint64_t nx = lo + step;
      if (nx >= nrec) {
        hi = nrec;
        break;
      }
      if (P[nx] > target) {
        hi = nx;
        break;
      }
      lo = nx;
      step <<= 1;
    }
    while (hi - lo > 1) {
      const uint64_t mid = lo + ((hi - lo) >> 1);
      if (P[mid] <= target) lo = mid;
      else hi = mid;
    }
    e = hi;
  }
  J0[s] = (uint32_t)e;
}

// one lifting level: Jn = Jt o Jt over [0, nrec]
__global__ __launch_bounds__(256) void k_hit_lift(const uint32_t *Jt, uint32_t *Jn, uint64_t n1) {
  const uint64_t s = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (s < n1) Jn[s] = Jt[Jt[s]];
}

// the number of buckets: the most jumps from record 0 that stay below nrec, + 1
__global__ void k_hit_count_buckets(const uint32_t *J, uint32_t T, uint64_t nrec, unsigned long long *meta) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t s = 0, cnt = 0;
  for (int t = (int)T - 1; t >= 0; --t) {
    const uint32_t x = J[(uint64_t)t * (nrec + 1) + s];
    if (x < nrec) {
      s = x;
      cnt += 1ull << t;
    }
  }
  meta[HM_NB] = nrec ? cnt + 1 : 0;
}

// start record of bucket j: J0 applied j times to record 0, by the binary digits of j
__global__ __launch_bounds__(256) void k_hit_starts(const uint32_t *J, uint32_t T, uint64_t nrec, uint64_t nb,
                                                    uint32_t *starts) {
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < nb) {
    uint64_t s = 0;
    for (uint32_t t = 0; t < T; ++t)
      if ((j >> t) & 1) s = J[(uint64_t)t * (nrec + 1) + s];
    starts[j] = (uint32_t)s;
  }
  if (j == 0) starts[nb] = (uint32_t)nrec;
}

struct HitLocalArgs {
  const void *keys;       // every k-mer in record order
  const uint32_t *P;      // P[r]: k-mers before record r
  const uint32_t *starts; // bucket starts, nb + 1
  void *norm;             // plain keys out
  void *pkeys;            // pair keys out
  uint32_t *pcounts;      // pair counts out
  unsigned long long *meta;
  uint32_t num_pes;       // owner filter (owner_bin >= 0, num_pes > 1)
  int32_t owner_bin;
  uint64_t seed;
};

DK_DEV uint32_t hit_home(uint64_t h, uint32_t ht) { return (uint32_t)((((h * 0x9E3779B97F4A7C15ull) >> 32) * ht) >> 32); }
DK_DEV uint64_t hit_hash(const K2 &x) {
  uint64_t h = x.lo ^ ((x.hi << 31) | (x.hi >> 33)) ^ 0xD1B54A32D192ED03ull;
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 32;
  return h;
}

template <typename KeyT>
DK_DEV bool hit_keep(const HitLocalArgs &a, const KeyT &x) {
  return a.num_pes <= 1 || a.owner_bin < 0 || owner_of(x, a.num_pes, a.seed) == (uint32_t)a.owner_bin;
}

// Emit one workgroup's (key, count) entries: counts 1-2 as that many plain
// keys, 3+ as a pair.  mine[] holds up to SPT entries of this thread
// (count 0 = none); places come from a workgroup scan and one bump per stream.
template <typename KeyT, int SPT>
DK_DEV void hit_emit(const HitLocalArgs &a, const KeyT (&x)[SPT], const uint32_t (&cnt)[SPT], uint32_t *s_wave,
                     unsigned long long *s_base) {
  uint32_t nn = 0, np = 0, kept = 0;
#pragma unroll
  for (int q = 0; q < SPT; ++q) {
    if (!cnt[q]) continue;
    kept += cnt[q];
    if (cnt[q] <= 2) nn += cnt[q];
    else ++np;
  }
  uint32_t tn, tp, tk;
  const uint32_t on = block_excl(nn, s_wave, &tn);
  const uint32_t op = block_excl(np, s_wave, &tp);
  block_excl(kept, s_wave, &tk);
  if (threadIdx.x == 0) {
    s_base[0] = tn ? atomicAdd(&a.meta[HM_NORMAL], (unsigned long long)tn) : 0;
    s_base[1] = tp ? atomicAdd(&a.meta[HM_PAIRS], (unsigned long long)tp) : 0;
    if (tk) atomicAdd(&a.meta[HM_KEPT], (unsigned long long)tk);
  }
  __syncthreads();
  uint64_t bn = s_base[0] + on, bp = s_base[1] + op;
  KeyT *norm = reinterpret_cast<KeyT *>(a.norm);
  KeyT *pk = reinterpret_cast<KeyT *>(a.pkeys);
#pragma unroll
  for (int q = 0; q < SPT; ++q) {
    if (!cnt[q]) continue;
    if (cnt[q] <= 2) {
      norm[bn++] = x[q];
      if (cnt[q] == 2) norm[bn++] = x[q];
    } else {
      pk[bp] = x[q];
      a.pcounts[bp++] = cnt[q];
    }
  }
  __syncthreads();  // s_base and s_wave are reused by the next emission
}

// one-word keys: the bucket's keys into a table of HIT_HT1 slots (CAS on the
// key; the one value no slot can hold, all ones (k = 32), counted apart)
__global__ __launch_bounds__(HIT_NT) void k_hit_local1(HitLocalArgs a) {
  constexpr uint64_t EMPTY = ~0ull;
  __shared__ uint64_t tk[HIT_HT1];
  __shared__ uint32_t tc[HIT_HT1];
  __shared__ uint32_t s_wave[HIT_NT / 64], s_ones;
  __shared__ unsigned long long s_base[2];
  const uint32_t tid = threadIdx.x, j = blockIdx.x;
  const uint64_t b0 = a.P[a.starts[j]], b1 = a.P[a.starts[j + 1]];
  if (b1 == b0) return;  // a bucket with no k-mer sends nothing
  for (uint32_t i = tid; i < HIT_HT1; i += HIT_NT) {
    tk[i] = EMPTY;
    tc[i] = 0;
  }
  if (tid == 0) s_ones = 0;
  __syncthreads();
  const uint64_t *keys = reinterpret_cast<const uint64_t *>(a.keys);
#pragma unroll
  for (int u = 0; u < HIT_ITEMS; ++u) {
    const uint64_t i = b0 + (uint64_t)u * HIT_NT + tid;
    if (i >= b1) continue;
    const uint64_t x = keys[i];
    if (x == EMPTY) {
      atomicAdd(&s_ones, 1u);
      continue;
    }
    // fewer distinct keys than slots: the probe ends within one lap
    uint32_t h = hit_home(x, HIT_HT1);
    for (uint32_t pr = 0; pr < HIT_HT1; ++pr) {
      const uint64_t o = atomicCAS((unsigned long long *)&tk[h], EMPTY, x);
      if (o == EMPTY || o == x) {
        atomicAdd(&tc[h], 1u);
        break;
      }
      h = h + 1 == HIT_HT1 ? 0u : h + 1;
    }
  }
  __syncthreads();
  if (tid == 0) atomicAdd(&a.meta[HM_BUCKETS], 1ull);
  uint64_t x[HIT_SPT1 + 1];
  uint32_t c[HIT_SPT1 + 1];
#pragma unroll
  for (int q = 0; q < HIT_SPT1; ++q) {
    const uint32_t i = tid + q * HIT_NT;
    x[q] = tk[i];
    c[q] = x[q] != EMPTY && hit_keep(a, x[q]) ? tc[i] : 0u;
  }
  x[HIT_SPT1] = EMPTY;
  c[HIT_SPT1] = tid == 0 && s_ones && hit_keep(a, EMPTY) ? s_ones : 0u;
  hit_emit<uint64_t, HIT_SPT1 + 1>(a, x, c, s_wave, s_base);
}

// two-word keys: a table of HIT_HT2 slots per class of the keys' hash (two
// classes by the lowest hash bit to begin with); a slot is claimed by a CAS
// on its state word (0 empty, 1 being written, 2 ready), so the 16-byte key
// is written once and then only compared.  A class with more distinct keys
// than slots is split into two by the next hash bit and redone.
__global__ __launch_bounds__(HIT_NT) void k_hit_local2(HitLocalArgs a) {
  __shared__ uint64_t klo[HIT_HT2], khi[HIT_HT2];
  __shared__ uint32_t tc[HIT_HT2], tst[HIT_HT2];
  __shared__ uint32_t s_wave[HIT_NT / 64], s_over, s_sp;
  __shared__ uint64_t s_stack[66];  // classes still to count: mask << 32 | value of the hash's low bits
  __shared__ unsigned long long s_base[2];
  const uint32_t tid = threadIdx.x, j = blockIdx.x;
  const uint64_t b0 = a.P[a.starts[j]], b1 = a.P[a.starts[j + 1]];
  if (b1 == b0) return;
  const K2 *keys = reinterpret_cast<const K2 *>(a.keys);
  K2 v[HIT_ITEMS];
  uint32_t hv[HIT_ITEMS];
#pragma unroll
  for (int u = 0; u < HIT_ITEMS; ++u) {
    const uint64_t i = b0 + (uint64_t)u * HIT_NT + tid;
    v[u] = K2{0, 0};
    hv[u] = 0;
    if (i < b1) {
      v[u] = keys[i];
      hv[u] = (uint32_t)hit_hash(v[u]);
    }
  }
  if (tid == 0) {
    s_stack[0] = (1ull << 32) | 1u;
    s_stack[1] = (1ull << 32) | 0u;
    s_sp = 2;
    atomicAdd(&a.meta[HM_BUCKETS], 1ull);
  }
  __syncthreads();
  while (s_sp) {  // uniform: written by thread 0 between barriers
    const uint64_t cls = s_stack[s_sp - 1];
    const uint32_t mask = (uint32_t)(cls >> 32), val = (uint32_t)cls;
    __syncthreads();
    if (tid == 0) {
      --s_sp;
      s_over = 0;
    }
    for (uint32_t i = tid; i < HIT_HT2; i += HIT_NT) {
      tc[i] = 0;
      tst[i] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < HIT_ITEMS; ++u) {
      if (b0 + (uint64_t)u * HIT_NT + tid >= b1 || (hv[u] & mask) != val) continue;
      uint32_t h = hit_home(hv[u] ^ (hv[u] >> 16) * 0x45D9F3Bu, HIT_HT2);
      bool done = false;
      // probes over distinct slots (a lap at most), plus rereads of a slot
      // another lane is writing (a bounded spin)
      for (uint32_t pr = 0, spin = 0; pr < HIT_HT2 && spin < (1u << 20); ++spin) {
        const uint32_t st = __hip_atomic_load(&tst[h], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (st == 0) {
          if (atomicCAS(&tst[h], 0u, 1u) == 0u) {
            klo[h] = v[u].lo;
            khi[h] = v[u].hi;
            atomicAdd(&tc[h], 1u);
            __hip_atomic_store(&tst[h], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            done = true;
            break;
          }
          continue;  // another key took the slot: read it again
        }
        if (st == 1) continue;  // being written
        if (klo[h] == v[u].lo && khi[h] == v[u].hi) {
          atomicAdd(&tc[h], 1u);
          done = true;
          break;
        }
        h = h + 1 == HIT_HT2 ? 0u : h + 1;
        ++pr;
      }
      if (!done) s_over = 1;
    }
    __syncthreads();
    if (s_over) {  // split the class by the next hash bit (32 bits at most)
      if (tid == 0) {
        const uint32_t m2 = mask * 2 + 1;
        if (mask == 0xFFFFFFFFu || s_sp + 2 > 66) {
          atomicOr(reinterpret_cast<unsigned int *>(&a.meta[HM_ERR]), ERR_OVERFLOW);
        } else {
          s_stack[s_sp++] = ((uint64_t)m2 << 32) | val;
          s_stack[s_sp++] = ((uint64_t)m2 << 32) | (val | (mask + 1));
        }
      }
      __syncthreads();
      continue;
    }
    K2 x[HIT_SPT2];
    uint32_t c[HIT_SPT2];
#pragma unroll
    for (int q = 0; q < HIT_SPT2; ++q) {
      const uint32_t i = tid + q * HIT_NT;
      x[q] = K2{klo[i], khi[i]};
      c[q] = tst[i] == 2 && hit_keep(a, x[q]) ? tc[i] : 0u;
    }
    hit_emit<K2, HIT_SPT2>(a, x, c, s_wave, s_base);
  }
}

template <typename KeyT>
DK_DEV uint64_t hit_lower_bound(const KeyT *keys, uint64_t lo, uint64_t hi, const KeyT &x) {
  while (lo < hi) {
    const uint64_t mid = lo + ((hi - lo) >> 1);
    if (key_lt(keys[mid], x)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

struct HitMergeArgs {
  const void *src;     // keys to look up
  uint32_t *src_counts;
  uint64_t n;
  const void *heavy;   // heavy table keys (ascending per owner bin)
  uint32_t *heavy_counts;
  uint64_t dh;
  const uint64_t *hoff;  // heavy bin offsets (num_pes + 1) when num_pes > 1
  uint32_t num_pes;
  uint64_t seed;
  int zero_src;          // merge: a found entry's count moves to the heavy table
  unsigned long long *meta;
};

// For each source entry: its key's place in the heavy table (in its owner
// bin); add its count there.  Pairs: every key is there.  Light runs: the
// ones found are zeroed (the light table then drops them).
template <typename KeyT>
__global__ __launch_bounds__(256) void k_hit_merge(HitMergeArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const KeyT x = reinterpret_cast<const KeyT *>(a.src)[i];
  const KeyT *h = reinterpret_cast<const KeyT *>(a.heavy);
  uint64_t lo = 0, hi = a.dh;
  if (a.num_pes > 1) {
    const uint32_t o = owner_of(x, a.num_pes, a.seed);
    lo = a.hoff[o];
    hi = a.hoff[o + 1];
  }
  const uint64_t p = hit_lower_bound(h, lo, hi, x);
  if (p < hi && key_eq(h[p], x)) {
    atomicAdd(&a.heavy_counts[p], a.src_counts[i]);
    if (a.zero_src) a.src_counts[i] = 0;
  } else if (!a.zero_src) {
    atomicOr(reinterpret_cast<unsigned int *>(&a.meta[HM_ERR]), ERR_OVERFLOW);  // a pair key not found: a bug
  }
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace dk

using namespace dk;

namespace {

// a device table of the context's HITTER buffers, or its host copy
int hit_out(dakc_ctx *c, const dakc_table &t, DevBuf &bk, DevBuf &bc, bool to_host, dakc_table *out) {
  *out = t;
  out->priv = nullptr;
  out->bin_offsets = nullptr;
  const size_t kb = (size_t)t.n * t.key_words * 8, cb = (size_t)t.n * 4;
  if (t.bin_offsets) {
    out->bin_offsets = (uint64_t *)std::malloc(((size_t)t.num_bins + 1) * 8);
    if (!out->bin_offsets) return DAKC_E_NOMEM;
    std::memcpy(out->bin_offsets, t.bin_offsets, ((size_t)t.num_bins + 1) * 8);
  }
  if (!to_host) {
    out->keys = (uint64_t *)bk.p;
    out->counts = (uint32_t *)bc.p;
    out->on_device = 1;
    return DAKC_OK;
  }
  out->on_device = 0;
  out->keys = (uint64_t *)std::malloc(kb ? kb : 8);
  out->counts = (uint32_t *)std::malloc(cb ? cb : 4);
  if (!out->keys || !out->counts) return DAKC_E_NOMEM;
  if (t.n && (hipMemcpy(out->keys, bk.p, kb, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(out->counts, bc.p, cb, hipMemcpyDeviceToHost) != hipSuccess))
    return DAKC_E_DEVICE;
  return DAKC_OK;
}

// count n keys (device) with dakc_count_keys and move the device table into
// (bk, bc); the table's bin offsets are kept in `offs` (host)
int hit_count_into(dakc_ctx *c, const void *keys, uint64_t n, int k, uint32_t P, uint64_t seed, DevBuf &bk,
                   DevBuf &bc, dakc_table *t, std::vector<uint64_t> &offs) {
  dakc_table w;
  int s = dakc_count_keys(c, (const uint64_t *)keys, n, k, P, seed, 1, 1, &w);
  if (s) return s;
  const size_t kb = (size_t)w.n * w.key_words * 8, cb = (size_t)w.n * 4;
  if ((s = bk.ensure(std::max<size_t>(kb, 8), c->stream, false)) || (s = bc.ensure(std::max<size_t>(cb, 4), c->stream, false))) {
    dakc_table_free(&w);
    return s;
  }
  if ((kb && hipMemcpyAsync(bk.p, w.keys, kb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) ||
      (cb && hipMemcpyAsync(bc.p, w.counts, cb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)) {
    dakc_table_free(&w);
    return DAKC_E_DEVICE;
  }
  *t = w;
  offs.clear();
  if (w.bin_offsets) offs.assign(w.bin_offsets, w.bin_offsets + w.num_bins + 1);
  dakc_table_free(&w);  // a workspace table: only forgets it
  t->keys = (uint64_t *)bk.p;
  t->counts = (uint32_t *)bc.p;
  t->bin_offsets = offs.empty() ? nullptr : offs.data();
  t->priv = nullptr;
  return DAKC_OK;
}

}  // namespace

extern "C" int dakc_count_hitter(dakc_ctx *c, const dakc_job *job, uint32_t bucket_size, dakc_table *light,
                                 dakc_table *heavy, dakc_hitter_stats *stats) {
  if (!c || !job || !light || !heavy) return DAKC_E_INVAL;
  std::memset(light, 0, sizeof(*light));
  std::memset(heavy, 0, sizeof(*heavy));
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (job->format != DAKC_FMT_FIXED) return DAKC_E_INVAL;
  if (!bucket_size) bucket_size = 10000;  // KCOUNT_BUCKET_SIZE (dakc/Makefile:11)
  if (bucket_size <= (uint32_t)job->read_len || bucket_size > HIT_MAX_BUCKET) return DAKC_E_INVAL;
  hipSetDevice(c->device);
  Plan pl;
  int s = make_plan(job, &pl);
  if (s) return s;
  hipStream_t st = c->stream;
  HitterBufs &hb = c->hit;
  // the reads on the device
  const uint8_t *dreads = (const uint8_t *)job->reads;
  if (!job->reads_on_device && pl.nbytes) {
    if ((s = hb.reads.ensure(pl.nbytes + 256, st, false))) return s;
    if (hipMemcpyAsync(hb.reads.p, job->reads, pl.nbytes, hipMemcpyHostToDevice, st) != hipSuccess)
      return DAKC_E_DEVICE;
    dreads = (const uint8_t *)hb.reads.p;
  }
  const uint64_t nrec = pl.nrec;
  const uint32_t bp = bucket_size - pl.L;
  // 1. the bucket starts
  const uint64_t ntiles = (nrec + SCAN_TILE - 1) / SCAN_TILE;
  // at most total / (bp + 1) + 1 buckets: every bucket but the last holds more than bp k-mers
  const uint64_t nb_max = pl.nmax / ((uint64_t)bp + 1) + 1;
  uint32_t T = 1;
  while (T < 40 && (1ull << T) <= nb_max) ++T;
  if ((s = hb.meta.ensure(256, st, false)) || (s = hb.c.ensure(std::max<uint64_t>(nrec, 1) * 4, st, false)) ||
      (s = hb.tiles.ensure(std::max<uint64_t>(ntiles, 1) * 8, st, false)) ||
      (s = hb.P.ensure((nrec + 1) * 4, st, false)) || (s = hb.J.ensure((uint64_t)T * (nrec + 1) * 4, st, false)) ||
      (s = hb.starts.ensure((nb_max + 2) * 4, st, false)))
    return s;
  auto *meta = (unsigned long long *)hb.meta.p;
  launch_fill_u64((uint64_t *)meta, 0, HM_WORDS, st);
  const ReadsDesc rd = plan_reads(pl, dreads);
  uint32_t *dc = (uint32_t *)hb.c.p, *tsum = (uint32_t *)hb.tiles.p, *toff = tsum + std::max<uint64_t>(ntiles, 1);
  uint32_t *P = (uint32_t *)hb.P.p, *J = (uint32_t *)hb.J.p, *starts = (uint32_t *)hb.starts.p;
  if (nrec) {
    k_hit_reccount<<<(uint32_t)ntiles, SCAN_TILE, 0, st>>>(rd, dc, tsum, meta);
    k_hit_scan_top<<<1, 1024, 0, st>>>(tsum, (uint32_t)ntiles, toff);
    k_hit_scan_apply<<<(uint32_t)ntiles, SCAN_TILE, 0, st>>>(dc, toff, nrec, P);
    const uint32_t g = (uint32_t)((nrec + 1 + 255) / 256);
    k_hit_next<<<g, 256, 0, st>>>(P, nrec, bp, J);
    for (uint32_t t = 1; t < T; ++t)
      k_hit_lift<<<g, 256, 0, st>>>(J + (uint64_t)(t - 1) * (nrec + 1), J + (uint64_t)t * (nrec + 1), nrec + 1);
    k_hit_count_buckets<<<1, 64, 0, st>>>(J, T, nrec, meta);
  }
  unsigned long long hm[HM_WORDS];
  uint32_t ptotal = 0;  // P[nrec]: the k-mers of all records
  if (hipMemcpyAsync(hm, meta, sizeof(hm), hipMemcpyDeviceToHost, st) != hipSuccess ||
      (nrec && hipMemcpyAsync(&ptotal, P + nrec, 4, hipMemcpyDeviceToHost, st) != hipSuccess) ||
      hipStreamSynchronize(st) != hipSuccess)
    return DAKC_E_DEVICE;
  if ((s = status_of_err((uint32_t)hm[HM_ERR]))) return s == STATUS_MSD_OVERFLOW ? DAKC_E_DEVICE : s;
  const uint64_t nb = hm[HM_NB];
  if (nb > nb_max) return DAKC_E_DEVICE;
  if (nb) k_hit_starts<<<(uint32_t)((nb + 1 + 255) / 256), 256, 0, st>>>(J, T, nrec, nb, starts);
  // 2. every k-mer in record order (the partition kernels with one bin)
  const size_t kb = (size_t)pl.kw * 8;
  const uint64_t n = pl.nmax;  // slots bound the k-mers
  if ((s = hb.keys.ensure(std::max<uint64_t>(n, 1) * kb, st, false))) return s;
  dakc_job one = *job;
  one.reads = (const char *)dreads;
  one.reads_on_device = 1;
  one.num_pes = 1;
  one.owner_bin = -1;
  uint64_t offs1[2] = {0, 0}, nk = 0;
  if ((s = dakc_partition(c, &one, (uint64_t *)hb.keys.p, n, offs1, &nk))) return s;
  if (nk != ptotal) return DAKC_E_DEVICE;  // the record counts and the extraction disagree
  // 3. the buckets' local counts: plain keys and pairs
  const uint64_t npair_max = nk / 3 + 1;
  if ((s = hb.norm.ensure(std::max<uint64_t>(nk, 1) * kb, st, false)) ||
      (s = hb.pkeys.ensure(npair_max * kb, st, false)) || (s = hb.pcounts.ensure(npair_max * 4, st, false)))
    return s;
  HitLocalArgs ha;
  ha.keys = hb.keys.p;
  ha.P = P;
  ha.starts = starts;
  ha.norm = hb.norm.p;
  ha.pkeys = hb.pkeys.p;
  ha.pcounts = (uint32_t *)hb.pcounts.p;
  ha.meta = meta;
  ha.num_pes = pl.num_pes;
  ha.owner_bin = pl.owner_bin;
  ha.seed = pl.seed;
  if (nb) {
    if (pl.kw == 1) k_hit_local1<<<(uint32_t)nb, HIT_NT, 0, st>>>(ha);
    else k_hit_local2<<<(uint32_t)nb, HIT_NT, 0, st>>>(ha);
  }
  if (hipMemcpyAsync(hm, meta, sizeof(hm), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return DAKC_E_DEVICE;
  if (hm[HM_ERR]) return DAKC_E_DEVICE;  // a class that could not be split further
  const uint64_t nnorm = hm[HM_NORMAL], npair = hm[HM_PAIRS], kept = hm[HM_KEPT];
  // 4. the owner side: light candidates and heavy keys counted, pairs added
  const uint32_t Pc = pl.owner_digit ? pl.num_pes : 1;
  dakc_table tl, th;
  std::vector<uint64_t> ol, oh;
  if ((s = hit_count_into(c, hb.norm.p, nnorm, pl.k, Pc, pl.seed, hb.lk, hb.lc, &tl, ol))) return s;
  if ((s = hit_count_into(c, hb.pkeys.p, npair, pl.k, Pc, pl.seed, hb.hk, hb.hc, &th, oh))) return s;
  if ((s = hb.hoff.ensure((oh.size() + 1) * 8, st, false))) return s;
  if (!oh.empty() && hipMemcpyAsync(hb.hoff.p, oh.data(), oh.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess)
    return DAKC_E_DEVICE;
  launch_fill_u32((uint32_t *)hb.hc.p, 0, th.n, st);
  HitMergeArgs ma;
  ma.heavy = hb.hk.p;
  ma.heavy_counts = (uint32_t *)hb.hc.p;
  ma.dh = th.n;
  ma.hoff = (const uint64_t *)hb.hoff.p;
  ma.num_pes = Pc;
  ma.seed = pl.seed;
  ma.meta = meta;
  auto merge = [&](const void *src, uint32_t *cnt, uint64_t m, int zero) {
    ma.src = src;
    ma.src_counts = cnt;
    ma.n = m;
    ma.zero_src = zero;
    if (!m) return;
    const uint32_t g = (uint32_t)((m + 255) / 256);
    if (pl.kw == 1) k_hit_merge<uint64_t><<<g, 256, 0, st>>>(ma);
    else k_hit_merge<K2><<<g, 256, 0, st>>>(ma);
  };
  merge(hb.pkeys.p, (uint32_t *)hb.pcounts.p, npair, 0);  // the pairs' counts into the heavy table
  if (th.n) merge(tl.keys, tl.counts, tl.n, 1);           // heavy runs of the plain keys move over
  if (hipMemcpyAsync(hm, meta, sizeof(hm), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return DAKC_E_DEVICE;
  if (hm[HM_ERR]) return DAKC_E_DEVICE;
  tl.on_device = th.on_device = 1;
  if (th.n && (s = table_filter(c, &tl, 1))) return s;  // the light table drops the moved runs
  th.total = kept - tl.total;
  const bool to_host = !job->table_on_device;
  if ((s = hit_out(c, tl, hb.lk, hb.lc, to_host, light)) || (s = hit_out(c, th, hb.hk, hb.hc, to_host, heavy))) {
    dakc_table_free(light);
    dakc_table_free(heavy);
    return s;
  }
  if (stats) {
    stats->buckets = hm[HM_BUCKETS];
    stats->normal_keys = nnorm;
    stats->heavy_pairs = npair;
    stats->kmers = kept;
    stats->bytes_sent = nnorm * kb + npair * (kb + 8);
    stats->bytes_without = kept * kb;
  }
  return DAKC_OK;
}
