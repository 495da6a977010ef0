// post.hip -- after the count, on a device table: the count-frequency
// histogram with its minimum bucket, the solid-k-mer filter, and point
// queries.
//
// The histogram and the cutoff follow pakman_star's construct_kmer_histogram
// (pakman_star/src/kmer_histogram.cpp:81-116: hist[c] for c < num_buckets,
// min_bucket = the first minimum over c >= 1) and its filter keeps the
// entries with count >= min_bucket in order (kmer_histogram.cpp:138-140).
// Queries look a k-mer up in the sorted table (the lookup kcounter.cpp:381-383
// describes), inside its owner bin when the table holds all P bins.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "dk_api_internal.hpp"
#include "dk_device.hpp"

namespace dk {

// ------------------------------------------------------------- histogram --

constexpr int PH_NT = 256;
constexpr uint32_t PH_REG = 16;           // counts below this are tallied in registers
constexpr uint32_t PH_LDS_MAX = 16384;    // buckets held in LDS; beyond, global atomics

// hist[c] += entries with count c, c < nb.  Most entries of a genome table
// have count 1 or a count near the coverage: the small counts go to per-thread
// registers and one wave sum each, so no LDS address takes a whole wavefront's
// atomics; the rest add into an LDS histogram (or straight into HBM when nb is
// larger than LDS holds).
template <bool LDS>
__global__ __launch_bounds__(PH_NT) void k_count_hist(const uint32_t *counts, uint64_t d, uint32_t nb,
                                                      unsigned long long *hist) {
  __shared__ uint32_t s_h[LDS ? PH_LDS_MAX : 1];
  if (LDS)
    for (uint32_t i = threadIdx.x; i < nb; i += PH_NT) s_h[i] = 0;
  __syncthreads();
  uint32_t reg[PH_REG];
#pragma unroll
  for (uint32_t j = 0; j < PH_REG; ++j) reg[j] = 0;
  const uint64_t stride = (uint64_t)gridDim.x * PH_NT;
  for (uint64_t i = (uint64_t)blockIdx.x * PH_NT + threadIdx.x; i < d; i += stride) {
    const uint32_t c = counts[i];
    if (c < PH_REG) {
#pragma unroll
      for (uint32_t j = 0; j < PH_REG; ++j) reg[j] += c == j ? 1u : 0u;
    } else if (c < nb) {
      if (LDS) atomicAdd(&s_h[c], 1u);
      else atomicAdd(&hist[c], 1ull);
    }
  }
  // wave sums of the register tallies, one atomic per bucket per wave
#pragma unroll
  for (uint32_t j = 0; j < PH_REG; ++j) {
    uint32_t v = reg[j];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v && j < nb) {
      if (LDS) atomicAdd(&s_h[j], v);
      else atomicAdd(&hist[j], (unsigned long long)v);
    }
  }
  if (LDS) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += PH_NT)
      if (s_h[i]) atomicAdd(&hist[i], (unsigned long long)s_h[i]);
  }
}

// ---------------------------------------------------------------- filter --
//
// Stable in-place compaction of (keys, counts) to the entries with count >=
// min_count, by decoupled look-back over tiles taken in ticket order.  In
// place is safe: a tile loads all its entries before it publishes its
// aggregate, it writes only after every earlier tile has published, and it
// writes below its own end, so no write lands on an entry not yet loaded.

constexpr int PF_NT = 512;
constexpr int PF_ITEMS = 16;
constexpr uint32_t PF_TILE = PF_NT * PF_ITEMS;
constexpr int PF_NW = PF_NT / 64;

struct FilterArgs {
  void *keys;
  uint32_t *counts;
  uint64_t d;
  uint32_t min_count;
  uint32_t nbins;           // owner-bin offsets to move (0: none)
  const uint64_t *off_in;   // nbins + 1 old offsets
  uint64_t *off_out;        // nbins + 1 new offsets
  uint64_t *flags;          // look-back words, one per tile, zeroed
  uint32_t *ticket;
  unsigned long long *out;  // [0] kept entries, [1] sum of kept counts
  uint32_t *err;            // ERR_STALL when a look-back wait gave up
};

// W: 64-bit words per key (held as plain words: an array of two-word structs
// would go to scratch)
template <int W>
__global__ __launch_bounds__(PF_NT) void k_filter(FilterArgs a) {
  __shared__ uint32_t s_row[PF_ITEMS][PF_NW];  // kept per (item row, wave), then its exclusive prefix
  __shared__ uint32_t s_tile, s_total;
  __shared__ uint64_t s_excl;
  __shared__ uint32_t s_rank[PF_TILE];  // exclusive kept rank of each position (bin offsets only)
  const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid / 64;
  if (tid == 0) s_tile = atomicAdd(a.ticket, 1u);
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
  const uint64_t t0 = (uint64_t)tile * PF_TILE;
  uint64_t *keys = reinterpret_cast<uint64_t *>(a.keys);
  // striped: item u of thread tid is entry t0 + u * NT + tid (coalesced)
  uint64_t k[PF_ITEMS][W];
  uint32_t c[PF_ITEMS];
  uint32_t keep = 0;  // bit u: item u is kept
#pragma unroll
  for (int u = 0; u < PF_ITEMS; ++u) {
    const uint64_t i = t0 + (uint64_t)u * PF_NT + tid;
    c[u] = 0;
#pragma unroll
    for (int x = 0; x < W; ++x) k[u][x] = 0;
    if (i < a.d) {
#pragma unroll
      for (int x = 0; x < W; ++x) k[u][x] = keys[i * W + x];
      c[u] = a.counts[i];
    }
    if (i < a.d && c[u] >= a.min_count) keep |= 1u << u;
  }
  uint64_t ball[PF_ITEMS];
  unsigned long long csum = 0;
#pragma unroll
  for (int u = 0; u < PF_ITEMS; ++u) {
    ball[u] = __ballot((keep >> u) & 1u);
    if ((keep >> u) & 1u) csum += c[u];
    if (lane == 0) s_row[u][w] = (uint32_t)__popcll(ball[u]);
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int u = 0; u < PF_ITEMS; ++u)
      for (int v = 0; v < PF_NW; ++v) {
        const uint32_t x = s_row[u][v];
        s_row[u][v] = run;
        run += x;
      }
    s_total = run;
    const uint32_t epoch = 1;
    if (tile == 0) {
      lb_store(a.flags, lb_pack(epoch, LB_INC, run));
      s_excl = 0;
    } else {
      lb_store(a.flags + tile, lb_pack(epoch, LB_AGG, run));
      const uint64_t ex = lb_lookback(a.flags, tile, 1, 0, epoch, a.err);
      lb_store(a.flags + tile, lb_pack(epoch, LB_INC, ex + run));
      s_excl = ex;
    }
  }
  // the sum of kept counts does not depend on order
  for (int off = 32; off > 0; off >>= 1) csum += __shfl_xor(csum, off);
  if (lane == 0 && csum) atomicAdd(&a.out[1], csum);
  __syncthreads();
  const uint64_t excl = s_excl;
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int u = 0; u < PF_ITEMS; ++u) {
    const uint32_t r = s_row[u][w] + (uint32_t)__popcll(ball[u] & lt);
    if ((keep >> u) & 1u) {
#pragma unroll
      for (int x = 0; x < W; ++x) keys[(excl + r) * W + x] = k[u][x];
      a.counts[excl + r] = c[u];
    }
    if (a.nbins) s_rank[u * PF_NT + tid] = r;
  }
  const uint64_t t1 = std::min<uint64_t>(t0 + PF_TILE, a.d);
  if (a.nbins) {
    __syncthreads();
    for (uint32_t b = tid; b <= a.nbins; b += PF_NT) {
      const uint64_t p = a.off_in[b];
      if (p >= t0 && p < t1) a.off_out[b] = excl + s_rank[p - t0];
      else if (p == a.d && t1 == a.d) a.off_out[b] = excl + s_total;
    }
  }
  if (tid == 0 && t1 == a.d) a.out[0] = excl + s_total;
}

// ----------------------------------------------------------------- query --

struct QueryArgs {
  const void *keys;
  const uint32_t *counts;
  uint64_t d;
  uint32_t nbins;           // 0: one sorted run
  const uint64_t *offs;     // nbins + 1
  uint64_t seed;            // owner-hash seed of the bins
  const void *queries;
  uint64_t nq;
  uint32_t *out;
};

// One query per thread: lower bound in its owner bin (or the whole table).
template <typename KeyT>
__global__ __launch_bounds__(256) void k_query(QueryArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.nq) return;
  const KeyT *keys = reinterpret_cast<const KeyT *>(a.keys);
  const KeyT q = reinterpret_cast<const KeyT *>(a.queries)[i];
  uint64_t lo = 0, end = a.d;
  if (a.nbins) {
    const uint32_t o = owner_of(q, a.nbins, a.seed);
    lo = a.offs[o];
    end = a.offs[o + 1];
  }
  uint64_t hi = end;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (key_lt(keys[mid], q)) lo = mid + 1;
    else hi = mid;
  }
  a.out[i] = lo < end && key_eq(keys[lo], q) ? a.counts[lo] : 0u;
}

}  // namespace dk

using namespace dk;

// ------------------------------------------------------------- C ABI --

static int post_check(dakc_ctx *ctx, const dakc_table *t) {
  if (!ctx || !t) return DAKC_E_INVAL;
  if (!t->on_device) return DAKC_E_INVAL;  // device tables only: the work runs on the GPU
  if (t->n && (!t->keys || !t->counts)) return DAKC_E_INVAL;
  if (t->key_words != 1 && t->key_words != 2) return DAKC_E_INVAL;
  hipSetDevice(ctx->device);
  return DAKC_OK;
}

static int sync_status(hipStream_t st) {
  if (hipGetLastError() != hipSuccess) return DAKC_E_DEVICE;
  return hipStreamSynchronize(st) == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
}

extern "C" int dakc_table_histogram(dakc_ctx *ctx, const dakc_table *t, uint32_t num_buckets,
                                    uint64_t *hist, uint32_t *min_bucket) {
  int s = post_check(ctx, t);
  if (s) return s;
  if (num_buckets && !hist) return DAKC_E_INVAL;
  if (num_buckets == 0) {  // no histogram, no cutoff (kmer_histogram.cpp:79 skips it)
    if (min_bucket) *min_bucket = 0;
    return DAKC_OK;
  }
  hipStream_t st = ctx->stream;
  if ((s = ctx->post_a.ensure((size_t)num_buckets * 8, st, false))) return s;
  auto *dh = (unsigned long long *)ctx->post_a.p;
  launch_fill_u64((uint64_t *)dh, 0, num_buckets, st);
  if (t->n) {
    const uint64_t per = (uint64_t)PH_NT * 16;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((t->n + per - 1) / per, (uint64_t)ctx->num_cus * 8);
    if (num_buckets <= PH_LDS_MAX)
      k_count_hist<true><<<grid, PH_NT, 0, st>>>(t->counts, t->n, num_buckets, dh);
    else
      k_count_hist<false><<<grid, PH_NT, 0, st>>>(t->counts, t->n, num_buckets, dh);
  }
  if (hipMemcpyAsync(hist, dh, (size_t)num_buckets * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
    return DAKC_E_DEVICE;
  if ((s = sync_status(st))) return s;
  if (min_bucket) {
    // std::min_element over [1, num_buckets): the first minimum; an empty
    // range (num_buckets == 1) gives its end, 1
    uint32_t m = 1;
    for (uint32_t c = 2; c < num_buckets; ++c)
      if (hist[c] < hist[m]) m = c;
    *min_bucket = m;
  }
  return DAKC_OK;
}

extern "C" int dakc_table_filter(dakc_ctx *ctx, dakc_table *t, uint32_t min_count) {
  int s = post_check(ctx, t);
  if (s) return s;
  if (t->n == 0 || min_count <= 1) return DAKC_OK;  // every entry has count >= 1
  return dk::table_filter(ctx, t, min_count);
}

// Stable in-place compaction of a device table to its entries with count >=
// min_count (with min_count = 1: drop the entries whose count was zeroed).
int dk::table_filter(dakc_ctx *ctx, dakc_table *t, uint32_t min_count) {
  int s;
  if (t->n == 0) return DAKC_OK;
  hipStream_t st = ctx->stream;
  const uint64_t ntiles = (t->n + PF_TILE - 1) / PF_TILE;
  if (ntiles > 0xFFFFFFFFull) return DAKC_E_TOOBIG;
  const uint32_t nbins = t->bin_offsets ? t->num_bins : 0;
  // scratch: flags | ticket + sums | old offsets | new offsets
  const size_t fl = ntiles * 8, meta = 64, ob = ((size_t)nbins + 1) * 8;
  if ((s = ctx->post_a.ensure(fl + meta + 2 * ob, st, false))) return s;
  uint8_t *p = (uint8_t *)ctx->post_a.p;
  FilterArgs a;
  a.keys = t->keys;
  a.counts = t->counts;
  a.d = t->n;
  a.min_count = min_count;
  a.nbins = nbins;
  a.flags = (uint64_t *)p;
  a.ticket = (uint32_t *)(p + fl);
  a.out = (unsigned long long *)(p + fl + 8);
  a.err = (uint32_t *)(p + fl + 24);
  a.off_in = (const uint64_t *)(p + fl + meta);
  a.off_out = (uint64_t *)(p + fl + meta + ob);
  launch_fill_u64((uint64_t *)p, 0, (fl + meta) / 8, st);
  if (nbins && hipMemcpyAsync((void *)a.off_in, t->bin_offsets, ob, hipMemcpyHostToDevice, st) != hipSuccess)
    return DAKC_E_DEVICE;
  if (t->key_words == 1) k_filter<1><<<(uint32_t)ntiles, PF_NT, 0, st>>>(a);
  else k_filter<2><<<(uint32_t)ntiles, PF_NT, 0, st>>>(a);
  unsigned long long out[3];  // kept, sum of kept counts, error word
  std::vector<uint64_t> offs(nbins + 1);
  if (hipMemcpyAsync(out, a.out, 24, hipMemcpyDeviceToHost, st) != hipSuccess ||
      (nbins && hipMemcpyAsync(offs.data(), a.off_out, ob, hipMemcpyDeviceToHost, st) != hipSuccess))
    return DAKC_E_DEVICE;
  if ((s = sync_status(st))) return s;
  if ((s = status_of_err((uint32_t)out[2]))) return s;  // the table is unchanged on the host side
  t->n = out[0];
  t->total = out[1];
  if (nbins) std::memcpy(t->bin_offsets, offs.data(), ob);
  return DAKC_OK;
}

extern "C" int dakc_table_query(dakc_ctx *ctx, const dakc_table *t, const uint64_t *queries, uint64_t nq,
                                int on_device, uint32_t *counts_out) {
  int s = post_check(ctx, t);
  if (s) return s;
  if (nq && (!queries || !counts_out)) return DAKC_E_INVAL;
  if (nq == 0) return DAKC_OK;
  hipStream_t st = ctx->stream;
  const size_t qb = (size_t)nq * t->key_words * 8, cb = (size_t)nq * 4;
  const uint32_t nbins = t->bin_offsets ? t->num_bins : 0;
  const size_t ob = nbins ? ((size_t)nbins + 1) * 8 : 0;
  const size_t qoff = (ob + 255) & ~(size_t)255;
  const size_t coff = qoff + (on_device ? 0 : (qb + 255) & ~(size_t)255);
  if ((s = ctx->post_b.ensure(coff + (on_device ? 0 : cb) + 256, st, false))) return s;
  uint8_t *p = (uint8_t *)ctx->post_b.p;
  QueryArgs a;
  a.keys = t->keys;
  a.counts = t->counts;
  a.d = t->n;
  a.nbins = nbins;
  a.offs = (const uint64_t *)p;
  a.seed = t->hash_seed ? t->hash_seed : OWNER_SEED;
  a.nq = nq;
  a.queries = on_device ? (const void *)queries : (const void *)(p + qoff);
  a.out = on_device ? counts_out : (uint32_t *)(p + coff);
  if (nbins && hipMemcpyAsync(p, t->bin_offsets, ob, hipMemcpyHostToDevice, st) != hipSuccess)
    return DAKC_E_DEVICE;
  if (!on_device && hipMemcpyAsync(p + qoff, queries, qb, hipMemcpyHostToDevice, st) != hipSuccess)
    return DAKC_E_DEVICE;
  const uint64_t grid = (nq + 255) / 256;
  if (grid > 0xFFFFFFFFull) return DAKC_E_TOOBIG;
  if (t->key_words == 1) k_query<uint64_t><<<(uint32_t)grid, 256, 0, st>>>(a);
  else k_query<K2><<<(uint32_t)grid, 256, 0, st>>>(a);
  if (!on_device && hipMemcpyAsync(counts_out, a.out, cb, hipMemcpyDeviceToHost, st) != hipSuccess)
    return DAKC_E_DEVICE;
  return sync_status(st);
}
