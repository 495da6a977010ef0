// kcount.hip -- gfx950 kernels of the k-mer counting hot path.
//
// Pipeline for one job (replaces kcounter.cpp:247-414, HITTER=0):
//   k_hist_reads  read records once: every radix digit histogram of every
//                 k-mer (+ owner digit), record-format checks
//   k_scan_digits exclusive digit starts, n
//   k_sweep<from_reads=true>   LSD pass 0: decode + 2-bit pack + sliding
//                 window straight from the records into digit-0 order
//   k_sweep<from_reads=false>  remaining LSD passes (onesweep: stable wave
//                 ballot ranking, decoupled look-back, LDS-staged scatter)
//   k_rle         run-length reduce of the sorted keys -> (kmer, count)
//   k_bin_offsets owner-bin boundaries when all P owner bins are counted
#include "dk_device.hpp"
#include "dk_kernels.hpp"
#include <algorithm>
#include <cstring>

namespace dk {

// ------------------------------------------------------------- utilities --

// s_waitcnt operand that waits for every vector memory operation (vmcnt 0)
// and for nothing else (expcnt 7, lgkmcnt 15)
constexpr int VMCNT0 = 0x0F70;

// The thread index, opaque to the compiler: per-thread address arithmetic is
// recomputed in each tile of a persistent loop instead of being hoisted out
// of it and spilled.
DK_DEV uint32_t opaque_tid() {
  uint32_t t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}


__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
// n 32-bit words, 16 bytes per thread and step (both pointers 16-byte aligned)
__global__ void k_copy_u32(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src, uint64_t n) {
  const uint64_t nv = n / 4, stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (uint64_t i = t; i < nv; i += stride)
    reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(src)[i];
  for (uint64_t i = nv * 4 + t; i < n; i += stride) dst[i] = src[i];
}

// Sum of the first `below` of NW per-wave values s[0, NW) (NW a power of two
// <= 64), in every lane: each lane reads one value and a butterfly of NW
// lanes adds them; one LDS read and log2(NW) shuffles instead of a chain of
// dependent reads.  Every lane must be active.
template <int NW, typename T>
DK_DEV T sum_waves_below(const T *s, uint32_t below, uint32_t lane) {
  static_assert(NW >= 1 && NW <= 64 && (NW & (NW - 1)) == 0, "a power of two of at most 64");
  const uint32_t l = lane & (NW - 1);
  T t = l < below ? s[l] : T(0);
#pragma unroll
  for (int o = NW / 2; o >= 1; o >>= 1) t += __shfl_xor(t, o, 64);
  return t;
}

// Exclusive scan of 256 u32 values held by threads [0, 256); all threads call.
// Returns the exclusive prefix for threads < 256; total in *s_total.
template <int NT>
DK_DEV uint32_t scan256(uint32_t v, uint32_t *s_wave, uint32_t *s_total) {
  static_assert(NT >= 256, "needs 256 threads");
  const uint32_t tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
  uint32_t x = tid < 256 ? v : 0;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (tid < 256 && lane == 63) s_wave[wave] = x;
  __syncthreads();
  const uint32_t pre = sum_waves_below<4>(s_wave, wave, lane);
  if (tid == 255) *s_total = pre + x;
  __syncthreads();
  return pre + x - (tid < 256 ? v : 0);
}

// Exclusive scan over all NT threads of one u32 each.  SHFL: the earlier
// waves' totals are added by shuffles (one LDS read); otherwise by a chain of
// LDS reads, which needs fewer registers in a kernel at its register limit.
template <int NT, bool SHFL = true>
DK_DEV uint32_t scan_block(uint32_t v, uint32_t *s_wave, uint32_t *s_total) {
  const uint32_t tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  uint32_t pre = 0;
  if constexpr (SHFL) {
    pre = sum_waves_below<NT / 64>(s_wave, wave, lane);
  } else {
    for (uint32_t w = 0; w < wave; ++w) pre += s_wave[w];
  }
  if (tid == NT - 1) *s_total = pre + x;
  __syncthreads();
  return pre + x - v;
}

// ------------------------------------------------------------ extraction --

// LDS scratch of one tile of records: raw bytes, packed 2-bit codes, masks of
// positions that cannot be in a k-mer, and each record's first 'M'.
template <int RAWB, int CODEW, int MAXRT>
struct ExtractSmemT {
  static constexpr int RAW_BYTES = RAWB, CODE_WORDS = CODEW, MAX_RECS = MAXRT;
  uint32_t raw[RAWB / 4 + 4];
  uint64_t code[CODEW];  // MSB-first 2-bit codes, 32 bases per word
  uint32_t inv[CODEW];   // LSB-first: bit q set = base 32w+q cannot be in a k-mer
  uint32_t mpos[MAXRT];  // first 'M' (end of read) per record, L if none
};
using ExtractSmem = ExtractSmemT<RAW_CAP, 896, 128>;        // 512-thread tiles
using ExtractSmemBig = ExtractSmemT<L1_RAW, L1_CODEW, L1_MAXRT>;  // 1024-thread tiles

// Decode one tile of records and emit the k-mers of ITEMS slots per thread.
// Slot s of the tile is (record s / S, start position s % S); thread t holds
// slots wave*ITEMS*64 + j*64 + lane.  Restates kcounter_funcs.cpp:40-111: a
// window is a k-mer iff its k bases are all ACGT (0xFF splits runs) and it
// starts before the record's first 'M' (0xF0 ends the read).
template <typename KeyT, int NT, int ITEMS, bool CHECK, typename SM>
DK_DEV void extract_tile(const ReadsDesc &rd, uint32_t tile, SM &sm, KeyT (&keys)[ITEMS],
                         uint32_t &vmask, uint32_t *err) {
  const uint32_t tid = threadIdx.x;
  const uint64_t r0 = (uint64_t)tile * rd.RT;
  const uint64_t left = rd.nrec - r0;
  const uint32_t nr = (uint32_t)(left < rd.RT ? left : rd.RT);
  const uint64_t b0 = r0 * rd.stride;
  uint64_t b1 = b0 + (uint64_t)nr * rd.stride;
  if (b1 > rd.nbytes) b1 = rd.nbytes;
  const uint64_t a0 = b0 & ~3ULL;
  const uint32_t nwords = (uint32_t)((b1 - a0 + 3) >> 2);
  const uint32_t *src32 = reinterpret_cast<const uint32_t *>(rd.reads + a0);
  // every word load in flight at once, then the LDS stores
  constexpr int RW = (SM::RAW_BYTES / 4 + 4 + NT - 1) / NT;
  uint32_t wv[RW];
#pragma unroll
  for (int j = 0; j < RW; ++j) {
    const uint32_t i = tid + j * NT;
    wv[j] = 0;
    if (i < nwords && a0 + 4ULL * i + 4 <= rd.nbytes) wv[j] = src32[i];
  }
#pragma unroll
  for (int j = 0; j < RW; ++j) {
    const uint32_t i = tid + j * NT;
    if (i < nwords) {
      const uint64_t addr = a0 + 4ULL * i;
      if (addr + 4 > rd.nbytes)
        for (uint32_t b = 0; b < 4; ++b)
          if (addr + b < rd.nbytes) wv[j] |= (uint32_t)rd.reads[addr + b] << (8 * b);
      sm.raw[i] = wv[j];
    }
  }
  for (uint32_t r = tid; r < nr; r += NT) sm.mpos[r] = rd.L;
  __syncthreads();

  const uint8_t *raw = reinterpret_cast<const uint8_t *>(sm.raw) + (b0 - a0);
  const uint32_t wpl = (rd.L + 31) >> 5;
  for (uint32_t it = tid; it < nr * rd.wpr; it += NT) {
    const uint32_t r = it / rd.wpr, w = it - r * rd.wpr;
    uint64_t code = 0;
    uint32_t inv = 0xFFFFFFFFu;
    if (w < wpl) {
      const uint8_t *rec = raw + (uint64_t)r * rd.stride;
      const uint32_t j0 = w * 32;
      uint32_t mloc = 0xFFFFFFFFu;
      inv = 0;
#pragma unroll 8
      for (uint32_t q = 0; q < 32; ++q) {
        const uint32_t j = j0 + q;
        const uint32_t b = j < rd.L ? rec[j] : 0u;
        const uint32_t u = b & 0xDFu;  // upper-case letters
        const bool ok = (u == 'A') | (u == 'C') | (u == 'G') | (u == 'T');
        const uint32_t t = (b >> 1) & 3u;             // A0 C1 T2 G3 (ASCII bits 1-2)
        const uint32_t c = t ^ ((t >> 1) ^ 1u);       // -> C0 A1 T2 G3
        code = (code << 2) | (ok ? c : 0u);
        inv |= (ok ? 0u : 1u) << q;
        if (u == 'M' && mloc == 0xFFFFFFFFu) mloc = j;
      }
      if (mloc != 0xFFFFFFFFu) atomicMin(&sm.mpos[r], mloc);
      if (CHECK) {
        if (w == 0 && rec[0] == 0) atomicOr(err, ERR_NUL_RECORD);
        if (w == wpl - 1 && b0 + (uint64_t)(r + 1) * rd.stride <= rd.nbytes && rec[rd.L] != '\n')
          atomicOr(err, ERR_SEPARATOR);
      }
    }
    sm.code[it] = code;
    sm.inv[it] = inv;
  }
  __syncthreads();

  const uint32_t lane = __lane_id(), wave = tid >> 6;
  vmask = 0;
  if (rd.S == 0) return;  // reads shorter than k: no k-mer slots
  const int k = (int)rd.k;
  const uint64_t kmask = k >= 64 ? ~0ULL : ((1ULL << k) - 1);
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t slot = wave * ITEMS * 64 + j * 64 + lane;
    uint32_t r = (uint32_t)((float)slot * rd.invS);
    if (r * rd.S > slot) --r;
    else if ((r + 1) * rd.S <= slot) ++r;
    const uint32_t p = slot - r * rd.S;
    bool v = r < nr;
    KeyT key{};
    if (v) {
      const uint32_t base = r * rd.wpr + (p >> 5), sh = p & 31;
      const uint64_t m01 = (uint64_t)sm.inv[base] | ((uint64_t)sm.inv[base + 1] << 32);
      const uint64_t m2 = sm.inv[base + 2];
      const uint64_t win = (m01 >> sh) | (sh ? (m2 << (64 - sh)) : 0ULL);
      v = ((win & kmask) == 0) && (p + rd.k <= sm.mpos[r]);
      // kcounter_funcs.cpp:97: the run left open at the end of a record is
      // emitted only if it starts before READLEN-1, so a lone base in the last
      // position never forms a 1-mer
      if (k == 1 && p == rd.L - 1 && (rd.L < 2 || ((sm.inv[r * rd.wpr + ((p - 1) >> 5)] >> ((p - 1) & 31)) & 1)))
        v = false;
      if (v) {
        key = window_key<KeyT>(sm.code[base], sm.code[base + 1], sm.code[base + 2], 2 * sh, k);
        if (rd.canonical) key = canon(key, k);
        if (rd.num_pes > 1 && rd.owner_bin >= 0 && owner_of(key, rd.num_pes, rd.seed) != (uint32_t)rd.owner_bin)
          v = false;
      }
    }
    keys[j] = key;
    vmask |= (v ? 1u : 0u) << j;
  }
}

// ------------------------------------------------------------- histogram --

template <typename KeyT>
DK_DEV uint32_t digit_eval(const DigitSpec &ds, const KeyT &key, uint64_t seed) {
  if (ds.mode == DIG_OWNER) return owner_of(key, ds.num_pes, seed);
  uint32_t v = key_bits(key, ds.shift, ds.nbits);
  if (ds.mode == DIG_MSD && ds.num_pes > 1) v += owner_of(key, ds.num_pes, seed) << ds.nbits;
  return v;
}

// Digit histograms (<= 256 bins each) of every k-mer of the records; also the
// record-format checks.  Persistent grid: LDS counts, one flush per group.
template <typename KeyT, int NT, int ITEMS>
__global__ __launch_bounds__(NT) void k_hist_reads(ReadsDesc rd, HistSpecs hs, uint32_t ntiles,
                                                   uint64_t *hist, uint32_t *err) {
  __shared__ ExtractSmem ex;
  __shared__ uint32_t s_h[MAX_PASSES * 256];
  const int nh = hs.n * 256;
  for (int i = threadIdx.x; i < nh; i += NT) s_h[i] = 0;
  // (the first barrier inside extract_tile orders the clear before use)
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    KeyT keys[ITEMS];
    uint32_t vmask;
    extract_tile<KeyT, NT, ITEMS, true>(rd, tile, ex, keys, vmask, err);
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      if (!((vmask >> j) & 1)) continue;
      for (int p = 0; p < hs.n; ++p) atomicAdd(&s_h[p * 256 + digit_eval(hs.d[p], keys[j], rd.seed)], 1u);
    }
    __syncthreads();
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nh; i += NT)
    if (s_h[i]) atomicAdd(reinterpret_cast<unsigned long long *>(&hist[i]), (unsigned long long)s_h[i]);
}

__global__ __launch_bounds__(256) void k_scan_digits(int npasses, const uint64_t *hist,
                                                     uint64_t *digit_start, uint64_t *n_dev) {
  __shared__ uint64_t s[256];
  const int d = threadIdx.x;
  for (int p = 0; p < npasses; ++p) {
    s[d] = hist[p * 256 + d];
    __syncthreads();
    if (d == 0) {
      uint64_t run = 0;
      for (int i = 0; i < 256; ++i) {
        uint64_t c = s[i];
        s[i] = run;
        run += c;
      }
      if (p == 0) n_dev[0] = run;
    }
    __syncthreads();
    digit_start[p * 256 + d] = s[d];
    __syncthreads();
  }
}

// -------------------------------------------------------------- LSD pass --

struct SweepArgs {
  ReadsDesc rd;
  const void *in;
  void *out;
  const uint64_t *n_dev;
  const uint64_t *digit_start;
  uint64_t *flags;
  uint32_t *ticket;
  DigitSpec dig;
  uint32_t epoch;
  uint64_t key_cap;  // keys the in / out buffers hold
  uint32_t *err;     // the job's error word
};

template <typename KeyT, int NT, int ITEMS, bool FROM_READS>
__global__ __launch_bounds__(NT) void k_sweep(SweepArgs a) {
  constexpr int WAVES = NT / 64;
  constexpr int TILE = NT * ITEMS;
  struct SortSmem {
    KeyT keys[TILE];
    uint8_t dig[TILE];
  };
  union Smem {
    ExtractSmem ex;
    SortSmem so;
  };
  static_assert(!FROM_READS || sizeof(ExtractSmem) <= sizeof(SortSmem), "LDS union");
  __shared__ Smem sm;
  __shared__ uint16_t s_wh[WAVES * 256];
  __shared__ uint32_t s_cnt[256], s_bbase[256], s_wave[WAVES];
  __shared__ uint64_t s_gbase[256];
  __shared__ uint32_t s_tile, s_total;

  const uint32_t tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(a.ticket, 1u);
  for (uint32_t i = tid; i < WAVES * 256; i += NT) s_wh[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  uint64_t ntiles, n = 0;
  if (FROM_READS) {
    ntiles = (a.rd.nrec + a.rd.RT - 1) / a.rd.RT;
  } else {
    n = a.n_dev[0];
    if (n > a.key_cap) n = a.key_cap;  // never index past the key buffers
    ntiles = (n + TILE - 1) / TILE;
  }
  if (tile >= ntiles) return;

  KeyT keys[ITEMS];
  uint32_t vmask = 0;
  if (FROM_READS) {
    extract_tile<KeyT, NT, ITEMS, false>(a.rd, tile, sm.ex, keys, vmask, nullptr);
  } else {
    const KeyT *in = reinterpret_cast<const KeyT *>(a.in);
    const uint64_t t0 = (uint64_t)tile * TILE + wave * ITEMS * 64 + lane;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint64_t idx = t0 + j * 64;
      if (idx < n) {
        keys[j] = in[idx];
        vmask |= 1u << j;
      } else {
        keys[j] = KeyT{};
      }
    }
  }

  // stable rank within the tile: order (wave, item, lane); wave-level digit
  // match by 8 ballots, running per-wave digit counts in LDS
  uint16_t rank[ITEMS];
  uint8_t digs[ITEMS];
  uint16_t *wh = s_wh + wave * 256;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const bool v = (vmask >> j) & 1;
    uint32_t d = 0;
    if (v) d = digit_eval(a.dig, keys[j], a.rd.seed);
    uint64_t peers = __ballot(v);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t below = __popcll(peers & lanemask_lt());
    uint32_t c = 0;
    if (v) c = wh[d];
    if (v && below == 0) wh[d] = (uint16_t)(c + __popcll(peers));
    rank[j] = (uint16_t)(c + below);
    digs[j] = (uint8_t)d;
  }
  __syncthreads();

  if (tid < 256) {
    uint32_t sum = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
      const uint32_t c = s_wh[w * 256 + tid];
      s_wh[w * 256 + tid] = (uint16_t)sum;
      sum += c;
    }
    s_cnt[tid] = sum;
  }
  __syncthreads();
  const uint32_t cnt = tid < 256 ? s_cnt[tid] : 0;
  const uint32_t bb = scan256<NT>(cnt, s_wave, &s_total);
  if (tid < 256) {
    s_bbase[tid] = bb;
    uint64_t *fl = a.flags + (uint64_t)tile * 256;
    uint64_t excl = 0;
    if (tile == 0) {
      lb_store(fl + tid, lb_pack(a.epoch, LB_INC, cnt));
    } else {
      lb_store(fl + tid, lb_pack(a.epoch, LB_AGG, cnt));
      excl = lb_lookback(a.flags, tile, 256, tid, a.epoch, a.err);
      lb_store(fl + tid, lb_pack(a.epoch, LB_INC, excl + cnt));
    }
    s_gbase[tid] = a.digit_start[tid] + excl;
  }
  __syncthreads();

#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    if ((vmask >> j) & 1) {
      const uint32_t d = digs[j];
      const uint32_t pos = s_bbase[d] + s_wh[wave * 256 + d] + rank[j];
      sm.so.keys[pos] = keys[j];
      sm.so.dig[pos] = (uint8_t)d;
    }
  }
  __syncthreads();
  KeyT *out = reinterpret_cast<KeyT *>(a.out);
  const uint32_t total = s_total;
  for (uint32_t i = tid; i < total; i += NT) {
    const uint32_t d = sm.so.dig[i];
    out[s_gbase[d] + (i - s_bbase[d])] = sm.so.keys[i];
  }
}

// ------------------------------------------------------ run-length reduce --

// kcounter.cpp:389-411: collapse equal neighbours of the sorted keys.
// Each tile emits one (kmer, count) per run that starts inside it; a run
// that continues past the tile end is measured by a galloping search.
template <typename KeyT, int NT, int ITEMS>
__global__ __launch_bounds__(NT) void k_rle(const KeyT *in, KeyT *out_keys, uint32_t *out_counts,
                                            uint64_t *n_dev, uint64_t *flags, uint32_t *ticket,
                                            uint32_t epoch, uint64_t key_cap, uint32_t *err) {
  constexpr int TILE = NT * ITEMS;
  __shared__ KeyT s_k[TILE];
  __shared__ uint16_t s_hp[TILE + 1];
  __shared__ uint32_t s_wave[NT / 64], s_total, s_tile;
  __shared__ uint64_t s_excl;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t n = n_dev[0] < key_cap ? n_dev[0] : key_cap;
  const uint64_t ntiles = (n + TILE - 1) / TILE;
  if (tile >= ntiles) return;
  const uint64_t t0 = (uint64_t)tile * TILE;
  const uint32_t cnt = (uint32_t)((n - t0) < (uint64_t)TILE ? (n - t0) : TILE);
  for (uint32_t i = tid; i < cnt; i += NT) s_k[i] = in[t0 + i];
  __syncthreads();

  uint32_t heads = 0;
  bool hf[ITEMS];
#pragma unroll
  for (int q = 0; q < ITEMS; ++q) {
    const uint32_t i = tid * ITEMS + q;
    bool h = false;
    if (i < cnt) {
      if (i > 0) h = !key_eq(s_k[i], s_k[i - 1]);
      else h = (t0 == 0) || !key_eq(s_k[0], in[t0 - 1]);
    }
    hf[q] = h;
    heads += h;
  }
  const uint32_t off = scan_block<NT>(heads, s_wave, &s_total);
  const uint32_t H = s_total;
  {
    uint32_t o = off;
#pragma unroll
    for (int q = 0; q < ITEMS; ++q)
      if (hf[q]) s_hp[o++] = (uint16_t)(tid * ITEMS + q);
  }
  if (tid == 0) {
    uint64_t excl = 0;
    if (tile == 0) {
      lb_store(flags, lb_pack(epoch, LB_INC, H));
    } else {
      lb_store(flags + tile, lb_pack(epoch, LB_AGG, H));
      excl = lb_lookback(flags, tile, 1, 0, epoch, err);
      lb_store(flags + tile, lb_pack(epoch, LB_INC, excl + H));
    }
    s_excl = excl;
    if (tile == ntiles - 1) n_dev[1] = excl + H;
  }
  __syncthreads();
  const uint64_t excl = s_excl;
  for (uint32_t h = tid; h < H; h += NT) {
    const uint32_t i = s_hp[h];
    const KeyT key = s_k[i];
    uint64_t end;
    if (h + 1 < H) {
      end = t0 + s_hp[h + 1];
    } else {
      uint64_t lo = t0 + cnt;  // first index past the tile
      if (lo >= n || !key_eq(in[lo], key)) {
        end = lo;
      } else {  // run continues: gallop then bisect for its end
        uint64_t step = 1, hi;
        for (;;) {
          const uint64_t pr = lo + step;
          if (pr >= n) { hi = n; break; }
          if (!key_eq(in[pr], key)) { hi = pr; break; }
          lo = pr;
          step <<= 1;
        }
        while (hi - lo > 1) {
          const uint64_t mid = lo + ((hi - lo) >> 1);
          if (key_eq(in[mid], key)) lo = mid;
          else hi = mid;
        }
        end = hi;
      }
    }
    const uint64_t c = end - (t0 + i);
    out_keys[excl + h] = key;
    out_counts[excl + h] = c > 0xFFFFFFFFULL ? 0xFFFFFFFFu : (uint32_t)c;
  }
}

// owner-bin boundaries of a table sorted by (owner, kmer): lower bound search
template <typename KeyT>
__global__ __launch_bounds__(256) void k_bin_offsets(const KeyT *keys, const uint64_t *n_dev,
                                                     uint32_t num_pes, uint64_t seed, uint64_t *offsets) {
  const uint64_t d = n_dev[1];
  for (uint32_t o = threadIdx.x; o <= num_pes; o += blockDim.x) {
    uint64_t lo = 0, hi = d;
    if (o == num_pes) lo = d;
    while (lo < hi) {
      const uint64_t mid = lo + ((hi - lo) >> 1);
      if (owner_of(keys[mid], num_pes, seed) < o) lo = mid + 1;
      else hi = mid;
    }
    offsets[o] = lo;
  }
}

// ------------------------------------------------------------- MSD path --
//
// One GPU's counting without the LSD passes: an MSD sweep from the records
// puts every k-mer in one of nb1 first-level buckets (top bits, and the owner
// when all P bins are counted); k_part2 splits each bucket tile by the next b2
// bits inside LDS and writes the tile back in digit order with its segment
// starts; k_final gathers one (bucket, digit) sub-bucket from all tiles of its
// bucket into LDS, sorts it there, run-length reduces it and writes its
// (kmer, count) entries at the position the decoupled look-back gives.

// exclusive scan in place of n <= 4*NT u32 in LDS; s[n] = *s_total = total.
// SHFL: the earlier waves' totals are added by shuffles (one LDS read);
// otherwise by a loop of LDS reads, which needs fewer registers in a kernel
// at its register limit.  Every thread adds up the wave totals itself, so
// one barrier publishes them and a second ends the scan.
template <int NT, bool SHFL = true>
DK_DEV void scan_lds(uint32_t *s, uint32_t n, uint32_t *s_wave, uint32_t *s_total) {
  constexpr uint32_t NW = NT / 64;
  const uint32_t tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
  uint32_t v[4], sum = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t i = tid * 4 + q;
    v[q] = i < n ? s[i] : 0;
    sum += v[q];
  }
  uint32_t x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  if constexpr (SHFL) {
    pre = sum_waves_below<NW>(s_wave, wave, lane);
    tot = sum_waves_below<NW>(s_wave, NW, lane);
  } else {
    // only the waves holding values have a non-zero total (4 * 64 per wave)
    const uint32_t nwv = min(NW, (n + 255) / 256);
#pragma nounroll
    for (uint32_t w = 0; w < nwv; ++w) {  // one read in flight: no registers held for all
      const uint32_t sw = s_wave[w];
      if (w < wave) pre += sw;
      tot += sw;
    }
  }
  pre += x - sum;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t i = tid * 4 + q;
    if (i < n) s[i] = pre;
    pre += v[q];
  }
  if (tid == 0) {
    s[n] = tot;
    *s_total = tot;
  }
  __syncthreads();
}


// Workgroup i is dispatched to XCD i % 8; give each XCD a contiguous range of
// tiles so the offset words neighbouring tiles write meet in one L2.
DK_DEV uint32_t xcd_remap(uint32_t bid, uint32_t n) {
  const uint32_t x = bid & 7, k = bid >> 3, per = n >> 3, rem = n & 7;
  return x * per + (x < rem ? x : rem) + k;
}

// LDS scratch of gather_segments for up to NSEG segments and CAP keys.
template <int NSEG, int CAP>
struct GatherSmem {
  static constexpr uint32_t NW = (CAP / 32 + 2) & ~1u;
  alignas(8) uint32_t wpre[NW];  // segments that start before position 32 * w
  uint32_t bits[NW];             // bit p: a non-empty segment starts at position p
  uint32_t delta[NSEG];          // k-th non-empty segment: source offset - first position
};

// Gather nseg segments into the workgroup's registers.  Segment q holds
// row1[q] - row0[q] keys at src[region0 + q * stride + row0[q]] (the bounds
// come in registers, from load_rows); q starts at
// position pref[q] of the concatenation.  Thread t receives positions
// u * NT + t (u < ITEMS), so each load instruction of a wavefront reads 64
// consecutive positions, mostly from one segment.  A position's segment is
// the number of segment starts at or before it, from a bitmask of the starts
// and per-word prefix counts (no search).  Returns the total; when it exceeds
// CAP mine = 0.  Keys past this thread's mine are left as they were (any
// value, not used).  Offsets must fit 32 bits (q * stride + row0[q] <
// 2^32).
// Segment bounds of thread t: row0[q], row1[q] for q = t * TPT + u (0 past nseg).
template <int TPT>
DK_DEV void load_rows(const uint16_t *row0, const uint16_t *row1, uint32_t nseg, uint32_t (&r0)[TPT],
                      uint32_t (&r1)[TPT]) {
  const uint32_t tid = opaque_tid();
#pragma unroll
  for (int u = 0; u < TPT; ++u) {
    const uint32_t q = tid * TPT + u;
    r0[u] = r1[u] = 0;
    if (q < nseg) {
      r0[u] = row0[q];
      r1[u] = row1[q];
    }
  }
}

// SHFL as in scan_block.
template <typename KeyT, int NT, int ITEMS, int TPT, int NSEG, bool SHFL>
DK_DEV uint32_t gather_segments(const KeyT *src, const uint32_t (&r0)[TPT], const uint32_t (&r1)[TPT],
                                uint32_t nseg, uint64_t region0, uint32_t stride,
                                GatherSmem<NSEG, NT * ITEMS> &g, uint32_t *s_wave,
                                uint32_t *s_total, KeyT (&v)[ITEMS], uint32_t &mine) {
  constexpr uint32_t CAP = NT * ITEMS, NW = GatherSmem<NSEG, NT * ITEMS>::NW;
  static_assert(NT * TPT >= NSEG, "segments per thread");
  static_assert(NW >= NT / 32, "scan scratch");
  const uint32_t tid = opaque_tid(), lane = tid & 63;
  uint32_t o0v[TPT], lenv[TPT], sum = 0, cnt = 0;
#pragma unroll
  for (int u = 0; u < TPT; ++u) {
    const uint32_t q = tid * TPT + u;
    o0v[u] = lenv[u] = 0;
    if (q < nseg) {
      o0v[u] = r0[u];
      lenv[u] = r1[u] - o0v[u];
      if (lenv[u] > CAP + 1) lenv[u] = CAP + 1;  // keeps the sum in 32 bits
    }
    sum += lenv[u];
    cnt += lenv[u] != 0;
  }
  for (uint32_t w = tid; w < NW; w += NT) g.bits[w] = 0;
  // exclusive prefix of (keys, non-empty segments) over the threads
  uint64_t x = ((uint64_t)cnt << 32) | sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  uint64_t *s_w64 = reinterpret_cast<uint64_t *>(g.wpre);  // NT / 64 words, reused below
  if (lane == 63) s_w64[tid >> 6] = x;
  __syncthreads();
  uint64_t pre = 0;
  if constexpr (SHFL) {
    pre = sum_waves_below<NT / 64>(s_w64, tid >> 6, lane);
  } else {
    // waves past the segments (TPT per thread) hold none: their totals are 0
    const uint32_t nwv = (nseg + 64 * TPT - 1) / (64 * TPT);
    const uint32_t wb = (tid >> 6) < nwv ? (tid >> 6) : nwv;
    for (uint32_t w = 0; w < wb; ++w) pre += s_w64[w];
  }
  if (tid == NT - 1) *s_total = (uint32_t)(pre + x);
  pre += x - (((uint64_t)cnt << 32) | sum);
  __syncthreads();
  const uint32_t tot = *s_total;
  const bool fits = tot <= CAP;  // uniform
  uint32_t pos = (uint32_t)pre, k = (uint32_t)(pre >> 32);
  if (tid == 0) g.wpre[0] = 0;
#pragma unroll
  for (int u = 0; u < TPT; ++u) {
    const uint32_t q = tid * TPT + u;
    if (fits && lenv[u]) {
      g.delta[k] = q * stride + o0v[u] - pos;
      atomicOr(&g.bits[pos >> 5], 1u << (pos & 31));
      // words whose first position lies in (pos, pos + len]: k + 1 starts precede them
      for (uint32_t w = (pos + 32) >> 5; w <= (pos + lenv[u]) >> 5; ++w) g.wpre[w] = k + 1;
      pos += lenv[u];
      ++k;
    }
  }
  __syncthreads();
  mine = fits && tot > tid ? (tot - tid + NT - 1) / NT : 0;
  if (mine > (uint32_t)ITEMS) mine = ITEMS;
  // in chunks of CH keys: the word reads, then the segment-offset reads, then
  // the key loads, so each group of LDS reads waits once and the loads of a
  // chunk are in flight while the next chunk's offsets are found
  constexpr int CH = ITEMS % 5 == 0 ? 5 : 4;
#pragma unroll
  for (int c0 = 0; c0 < ITEMS; c0 += CH) {
    // LDS reads unconditional (positions past the total read in-range words
    // whose values are not used), so each group of reads waits once
    uint32_t wp[CH], wb[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const uint32_t w = ((c0 + i < ITEMS ? c0 + i : 0) * NT + tid) >> 5;
      wp[i] = g.wpre[w];
      wb[i] = g.bits[w];
    }
    __builtin_amdgcn_sched_barrier(0);  // keep each group of reads together
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const uint32_t p = (c0 + i < ITEMS ? c0 + i : 0) * NT + tid;
      uint32_t seg = wp[i] + __popc(wb[i] & (0xFFFFFFFFu >> (31 - (p & 31)))) - 1;
      seg = seg < (uint32_t)NSEG ? seg : (uint32_t)NSEG - 1;
      wp[i] = g.delta[seg] + p;
    }
    __builtin_amdgcn_sched_barrier(0);
    // a slot past this thread's keys loads nothing and is left as it was
    // (zeroing them first made the compiler copy every later key register
    // at each load's branch)
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int u = c0 + i;
      if (u < ITEMS && (uint32_t)u < mine) v[u] = src[region0 + wp[i]];
    }
  }
  return tot;
}

// Job owning work item `item` (items numbered across jobs by the prefix
// `pre` of njobs + 1 entries), found by the whole workgroup at once: every
// thread tests its own jobs, one coalesced read instead of a chain of
// dependent ones.  `item` is uniform; ends with a barrier unless njobs == 1.
DK_DEV uint32_t find_job_wg(const uint32_t *pre, uint32_t njobs, uint32_t item, uint32_t *s_job) {
  if (njobs == 1) return 0;
  for (uint32_t j = opaque_tid(); j < njobs; j += blockDim.x)
    if (pre[j] <= item && item < pre[j + 1]) *s_job = j;
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*s_job);
}

// The descriptors and item prefixes of a multi-job launch are written before
// it and only read by kernels: read them through the scalar data path (a
// uniform index gives scalar loads into SGPRs).
#define DK_CONSTANT __attribute__((address_space(4)))
DK_DEV JobDesc job_desc(const BatchArgs &a, uint32_t j) {
  static_assert(sizeof(JobDesc) % 4 == 0, "JobDesc is copied in 32-bit words");
  const DK_CONSTANT uint32_t *src = (const DK_CONSTANT uint32_t *)(a.jobs + j);
  JobDesc d;
  uint32_t *dst = reinterpret_cast<uint32_t *>(&d);
#pragma unroll
  for (uint32_t i = 0; i < sizeof(JobDesc) / 4; ++i) dst[i] = src[i];
  return d;
}
DK_DEV uint32_t pre_at(const uint32_t *pre, uint32_t j) {
  return ((const DK_CONSTANT uint32_t *)pre)[j];
}

// The job owning the current item, with its item range [lo, hi): the items a
// persistent workgroup takes mostly stay in one job, so it is looked up again
// only when an item leaves the range.  Uniform; may end with a barrier.
struct JobCursor {
  uint32_t job = 0, lo = 0, hi = 0;
};
DK_DEV void cursor_seek(JobCursor &c, const uint32_t *pre, uint32_t njobs, uint32_t it, uint32_t *s_job) {
  if (it >= c.lo && it < c.hi) return;
  c.job = find_job_wg(pre, njobs, it, s_job);
  c.lo = pre_at(pre, c.job);
  c.hi = pre_at(pre, c.job + 1);
}

// Staged keys [0, tot) of a workgroup's LDS to o[0, tot), thread t taking
// positions t + u * NT (coalesced): a chunk of CH LDS reads, then its
// stores.  Reads stay in range (positions < NT * ITEMS) and are not stored
// past the total.  The chunk is held as 64-bit words, which stay in
// registers for two-word keys too (an array of key structs went to scratch).
template <typename KeyT, int NT, int ITEMS>
DK_DEV void write_staged(const KeyT *stage, KeyT *o, uint32_t tot, uint32_t tid) {
  constexpr int CH = ITEMS % 5 == 0 ? 5 : 4;
  constexpr int W = sizeof(KeyT) / 8;
  static_assert(W == 1 || W == 2, "one- or two-word keys");
#pragma unroll
  for (int c0 = 0; c0 < ITEMS; c0 += CH) {
    uint64_t w[CH * W];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      if (c0 + i < ITEMS) {
        const KeyT x = stage[tid + (c0 + i) * NT];
        if constexpr (W == 1) {
          w[i] = x;
        } else {
          w[2 * i] = x.lo;
          w[2 * i + 1] = x.hi;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      if (c0 + i < ITEMS && tid + (c0 + i) * NT < tot) {
        if constexpr (W == 1) o[tid + (c0 + i) * NT] = w[i];
        else o[tid + (c0 + i) * NT] = KeyT{w[2 * i], w[2 * i + 1]};
      }
    }
  }
}

// ---------------------------------------------------------- L1 extraction --
//
// k_l1 tiles are record-aligned: a record's S window slots are split over
// TPR = ceil(S / ITEMS) consecutive threads of ITEMS consecutive slots each,
// so a thread's windows slide along one record: after the first, each k-mer
// is the previous one shifted by one base (kcounter_funcs.cpp:25-38 does the
// same per read).

constexpr int L1_RAWW = L1_RAW / 4 + 4;  // LDS words of an L1 tile's raw records (+ alignment)

// Geometry of a tile's raw bytes: records r0 .. r0 + nr - 1 of the job start
// at byte b0; word i of the tile is the 4 bytes at a0 + 4i, a0 = b0 rounded
// down to 4, nwords of them.
struct L1Span {
  uint64_t r0, b0, a0;
  uint32_t nr, nwords;
};
DK_DEV L1Span l1_span(const ReadsDesc &rd, uint32_t tile) {
  L1Span t;
  t.r0 = (uint64_t)tile * rd.RT;
  const uint64_t left = rd.nrec - t.r0;
  t.nr = (uint32_t)(left < rd.RT ? left : rd.RT);
  t.b0 = t.r0 * rd.stride;
  uint64_t b1 = t.b0 + (uint64_t)t.nr * rd.stride;
  if (b1 > rd.nbytes) b1 = rd.nbytes;
  t.a0 = t.b0 & ~3ULL;
  t.nwords = (uint32_t)((b1 - t.a0 + 3) >> 2);
  return t;
}

// Start loading a tile's raw words straight into LDS (raw[i] = word i, no
// registers involved; the loads complete under vmcnt).  A word that would
// read past the job's end is left out: l1_patch_tail assembles it.
template <int NT, int RAWW>
DK_DEV void l1_prefetch_raw(const ReadsDesc &rd, uint32_t tile, uint32_t *raw) {
  const uint32_t tid = opaque_tid(), wbase = tid & ~63u;
  const L1Span t = l1_span(rd, tile);
  const uint32_t *src32 = reinterpret_cast<const uint32_t *>(rd.reads + t.a0);
#pragma unroll
  for (int j = 0; j < (RAWW + NT - 1) / NT; ++j) {
    const uint32_t i = tid + j * NT;
    if (i < t.nwords && t.a0 + 4ULL * i + 4 <= rd.nbytes)
      __builtin_amdgcn_global_load_lds(src32 + i, raw + j * NT + wbase, 4, 0, 0);
  }
}

// The job's last word, when the job does not end on a 4-byte boundary.
DK_DEV void l1_patch_tail(const ReadsDesc &rd, uint32_t tile, uint32_t *raw) {
  const L1Span t = l1_span(rd, tile);
  const uint32_t i = t.nwords - 1;
  const uint64_t addr = t.a0 + 4ULL * i;
  if (opaque_tid() == 0 && t.nwords && addr + 4 > rd.nbytes) {
    uint32_t w = 0;
    for (uint32_t b = 0; b < 4; ++b)
      if (addr + b < rd.nbytes) w |= (uint32_t)rd.reads[addr + b] << (8 * b);
    raw[i] = w;
  }
}

// Sliding-window helpers.  key_push appends base b to a forward k-mer;
// rc_push prepends the complement of b to its reverse complement.
DK_DEV void key_push(uint64_t &x, uint32_t b, uint64_t kmask) { x = ((x << 2) | b) & kmask; }
DK_DEV void key_push(K2 &x, uint32_t b, uint64_t kmask) {
  x.hi = ((x.hi << 2) | (x.lo >> 62)) & kmask;
  x.lo = (x.lo << 2) | b;
}
DK_DEV void rc_push(uint64_t &x, uint32_t b, int k) { x = (x >> 2) | ((uint64_t)(b ^ 3u) << (2 * k - 2)); }
DK_DEV void rc_push(K2 &x, uint32_t b, int k) {  // 33 <= k <= 64: the new base is in the high word
  x.lo = (x.lo >> 2) | (x.hi << 62);
  x.hi = (x.hi >> 2) | ((uint64_t)(b ^ 3u) << (2 * k - 66));
}

// LDS of an L1 tile's decoded records: packed 2-bit codes and masks of
// positions that cannot be in a k-mer.
struct L1Code {
  uint64_t code[L1_CODEW];  // MSB-first 2-bit codes, 32 bases per word
  uint32_t inv[L1_CODEW];   // LSB-first: bit q set = base 32w+q cannot be in a k-mer
};

// Decode the tile's raw records (in `raw`, records r0 .. r0 + nr - 1 of the
// job) into 2-bit codes and invalid-base masks in LDS and each record's first
// 'M' into mpos, then build this thread's ITEMS windows.  Restates
// kcounter_funcs.cpp:40-111: a window is a k-mer iff its k bases are all ACGT
// (0xFF splits runs) and it starts before the record's first 'M' (0xF0 ends
// the read).  The caller has set mpos[r] to L and made raw and mpos visible
// (barrier); after the decode's closing barrier `after_decode` runs, and raw
// is no longer read.
template <typename KeyT, int NT, int ITEMS, typename AfterDecode>
DK_DEV void l1_windows(const ReadsDesc &rd, uint32_t tile, const uint32_t *raw32, L1Code &sm,
                       uint32_t *mpos, KeyT (&keys)[ITEMS], uint32_t &vmask, uint32_t *err,
                       AfterDecode after_decode) {
  const uint32_t tid = opaque_tid();
  const L1Span t = l1_span(rd, tile);
  const uint32_t nr = t.nr;
  const uint64_t b0 = t.b0;
  const uint8_t *raw = reinterpret_cast<const uint8_t *>(raw32) + (b0 - t.a0);
  const uint32_t wpl = (rd.L + 31) >> 5;
  static_assert(L1_CODEW <= 2 * NT, "two code words per thread at most");
#pragma nounroll
  for (int pass = 0; pass < 2; ++pass) {
    const uint32_t it = tid + pass * NT;
    if (it >= nr * rd.wpr) break;
    const uint32_t r = it / rd.wpr, w = it - r * rd.wpr;
    uint64_t code = 0;
    uint32_t inv = 0xFFFFFFFFu;
    if (w < wpl) {
      const uint8_t *rec = raw + (uint64_t)r * rd.stride;
      const uint32_t j0 = w * 32;
      // bytes j0 .. j0 + 31 of the record: nine aligned LDS words (a word
      // past the buffer only holds bytes past the read, which are masked),
      // realigned and decoded four at a time
      const uint32_t off = (uint32_t)(b0 - t.a0) + r * rd.stride + j0;
      const uint32_t i0 = off >> 2, sh = 8 * (off & 3);
      uint32_t dw[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) dw[q] = raw32[i0 + q < (uint32_t)L1_RAWW ? i0 + q : L1_RAWW - 1];
      const uint32_t nvalid = rd.L - j0;  // bases of the read from j0 on (>= 1)
      uint32_t mloc = 0xFFFFFFFFu;
      inv = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint32_t x = (uint32_t)((((uint64_t)dw[q + 1] << 32) | dw[q]) >> sh);
        const Dec4 d = decode4(x, nvalid > 4u * q ? nvalid - 4u * q : 0u);
        code |= (uint64_t)d.code8 << (56 - 8 * q);
        inv |= d.inv4 << (4 * q);
        if (d.mmask && mloc == 0xFFFFFFFFu) mloc = j0 + 4 * q + (__builtin_ctz(d.mmask) >> 3);
      }
      if (mloc != 0xFFFFFFFFu) atomicMin(&mpos[r], mloc);
      if (w == 0 && rec[0] == 0) atomicOr(err, ERR_NUL_RECORD);
      if (w == wpl - 1 && b0 + (uint64_t)(r + 1) * rd.stride <= rd.nbytes && rec[rd.L] != '\n')
        atomicOr(err, ERR_SEPARATOR);
    }
    sm.code[it] = code;
    sm.inv[it] = inv;
  }
  __syncthreads();
  after_decode();  // raw is free from here on

  vmask = 0;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) keys[j] = KeyT{};
  if (rd.S == 0) return;  // reads shorter than k: no k-mer slots
  const int k = (int)rd.k;
  const uint32_t tpr = (rd.S + ITEMS - 1) / ITEMS;
  const uint32_t r = tid / tpr, p0 = (tid - r * tpr) * ITEMS;
  if (r >= nr) return;
  const uint32_t nslots = rd.S - p0 < (uint32_t)ITEMS ? rd.S - p0 : (uint32_t)ITEMS;
  const uint32_t rb = r * rd.wpr, w0 = rb + (p0 >> 5), sh0 = p0 & 31;
  const uint32_t mp = mpos[r];
  // invalid-base bits from p0 on: vlo = bits [p0, p0 + 64), vhi = the next 32
  const uint64_t i01 = (uint64_t)sm.inv[w0] | ((uint64_t)sm.inv[w0 + 1] << 32);
  uint64_t i23 = sm.inv[w0 + 2];
  if (sizeof(KeyT) == 16) i23 |= (uint64_t)sm.inv[w0 + 3] << 32;
  const uint64_t vlo = sh0 ? (i01 >> sh0) | (i23 << (64 - sh0)) : i01;
  const uint64_t vhi = i23 >> sh0;
  // the bases entering the window at slots 1..ITEMS-1: positions p0 + k on
  const uint32_t q0 = p0 + (uint32_t)k, wq = rb + (q0 >> 5), sq = 2 * (q0 & 31);
  const uint64_t c0 = sm.code[wq], c1 = sm.code[wq + 1];
  const uint64_t nb = sq ? (c0 << sq) | (c1 >> (64 - sq)) : c0;
  const uint64_t kbits = k >= 64 ? ~0ULL : ((1ULL << k) - 1);  // one bit per base
  const uint64_t kmask = sizeof(KeyT) == 8 ? (k >= 32 ? ~0ULL : ((1ULL << (2 * k)) - 1))
                                           : (k >= 64 ? ~0ULL : ((1ULL << (2 * k - 64)) - 1));
  // Which windows are k-mers, for all of them at once.  Window j (position
  // p0 + j) must be one of this thread's slots and end at or before the
  // record's first 'M' (p0 + j + k <= mp): j < nl.  None of its k bases may
  // be invalid: the ITEMS windows span ITEMS - 1 + k positions from p0 (the
  // read's end bounds the windows that count) and reads seldom hold an
  // invalid base, so one test clears every window and only a thread whose
  // span holds one tests them one by one.
  const int lim = (int)mp - k - (int)p0 + 1;
  const uint32_t nl = lim <= 0 ? 0u : ((uint32_t)lim < nslots ? (uint32_t)lim : nslots);
  uint32_t ok = (1u << nl) - 1;  // nl <= ITEMS < 32
  const uint32_t span = min((uint32_t)(ITEMS - 1 + k), rd.L - p0);  // <= 46 one-word, <= 71 two-word
  const bool anybad = span < 64 ? (vlo & ((1ULL << span) - 1)) != 0
                                : (vlo != 0 || (vhi & ((1ULL << (span - 64)) - 1)) != 0);
  if (anybad) {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint64_t win = j ? (vlo >> j) | (vhi << (64 - j)) : vlo;
      if ((win & kbits) != 0) ok &= ~(1u << j);
    }
  }
  // kcounter_funcs.cpp:97: the run left open at the end of a record is
  // emitted only if it starts before READLEN-1, so a lone base in the last
  // position never forms a 1-mer
  if (k == 1) {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t p = p0 + j;
      if (((ok >> j) & 1) && p == rd.L - 1 &&
          (rd.L < 2 || ((sm.inv[rb + ((p - 1) >> 5)] >> ((p - 1) & 31)) & 1)))
        ok &= ~(1u << j);
    }
  }
  KeyT fw = window_key<KeyT>(sm.code[w0], sm.code[w0 + 1], sm.code[w0 + 2], 2 * sh0, k);
  KeyT rc = rd.canonical ? revcomp(fw, k) : fw;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    if (j > 0) {
      const uint32_t b = (uint32_t)(nb >> (64 - 2 * j)) & 3u;
      key_push(fw, b, kmask);
      if (rd.canonical) rc_push(rc, b, k);
    }
    keys[j] = rd.canonical ? (key_lt(rc, fw) ? rc : fw) : fw;
  }
  if (rd.num_pes > 1 && rd.owner_bin >= 0) {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
      if (((ok >> j) & 1) && owner_of(keys[j], rd.num_pes, rd.seed) != (uint32_t)rd.owner_bin) ok &= ~(1u << j);
  }
  vmask = ok;
}

// L1: tiles of records -> their k-mers in first-level bucket order, each
// written contiguously at tile * tcap1 with the tile's bucket starts in off1.
// Also the record-format checks and the k-mer total.  Persistent: workgroup w
// runs on XCD w % 8, each XCD owns a contiguous range of tiles (so the offset
// words of neighbouring tiles meet in one L2) and its workgroups take them in
// turn; the next tile's records stream into LDS while the current one is
// processed.
template <typename KeyT, int NT, int ITEMS, bool ONE>
__global__ __launch_bounds__(NT) void k_l1(BatchArgs a, JobDesc one) {
  constexpr int TILE = NT * ITEMS;
  constexpr int RAWW = L1_RAWW;
  union Smem {
    L1Code ex;
    KeyT stage[TILE];
  };
  __shared__ Smem sm;
  __shared__ uint32_t s_raw[RAWW];
  __shared__ uint32_t s_h[1024 + 1];
  __shared__ uint32_t s_mpos[L1_MAXRT];  // first 'M' per record (outside the stage, set a tile ahead)
  __shared__ uint32_t s_wave[NT / 64], s_total;
  __shared__ uint32_t s_job;
  static_assert(NT >= 1024 && L1_MAXRT <= NT, "one histogram slot and record per thread");
  const uint32_t nitems = ONE ? one.m.ntiles1 : pre_at(a.pre1, a.njobs);
  const uint32_t G = gridDim.x, x = blockIdx.x & 7, nx = (G - x + 7) >> 3;
  const uint32_t per = nitems >> 3, rem = nitems & 7;
  const uint32_t lo = x * per + (x < rem ? x : rem);
  const uint32_t hi = lo + per + (x < rem ? 1u : 0u);
  uint32_t item = lo + (blockIdx.x >> 3);
  JobCursor cur;  // the job of the tile being loaded next
  // A tile's LDS set-up is done during the tile before it: its records
  // stream in (and its job's last word is assembled) once the previous
  // records are decoded, its digit counts are cleared and its record ends
  // set to the read length after the previous tile's staging.  One barrier
  // at the top of the tile then orders all of it before the decode.
  const uint32_t tid0 = opaque_tid();
  if (item < hi) {
    if (!ONE) cursor_seek(cur, a.pre1, a.njobs, item, &s_job);
    const ReadsDesc rdf = ONE ? one.rd : job_desc(a, cur.job).rd;
    l1_prefetch_raw<NT, RAWW>(rdf, item - cur.lo, s_raw);
    l1_patch_tail(rdf, item - cur.lo, s_raw);
    s_h[tid0] = 0;
    if (tid0 == 0) s_h[1024] = 0;
    if (tid0 < L1_MAXRT) s_mpos[tid0] = rdf.L;
  }
  __builtin_amdgcn_s_waitcnt(VMCNT0);
  for (; item < hi; item += nx) {
    const JobDesc J = ONE ? one : job_desc(a, cur.job);
    const ReadsDesc rd = J.rd;
    const MsdArgs m = J.m;
    KeyT *out1 = reinterpret_cast<KeyT *>(J.out1);
    uint64_t *n_dev = J.nd;
    uint32_t *err = J.err;
    const uint32_t tile = item - cur.lo;
    const uint32_t tid = opaque_tid();
    __syncthreads();  // the previous tile is written out; this tile's records and set-up are in LDS
    KeyT keys[ITEMS];
    uint32_t vmask, dr[ITEMS];  // digit << 16 | rank within the digit (< 2^14 keys per tile)
    uint32_t next_L = 0;        // read length of the next tile's job
    const uint32_t nitem = item + nx;
    l1_windows<KeyT, NT, ITEMS>(rd, tile, s_raw, sm.ex, s_mpos, keys, vmask, err, [&]() {
      // the records are decoded: the next tile's stream into the raw buffer
      if (nitem < hi) {
        if (!ONE) cursor_seek(cur, a.pre1, a.njobs, nitem, &s_job);
        const ReadsDesc rdn = ONE ? one.rd : job_desc(a, cur.job).rd;
        l1_prefetch_raw<NT, RAWW>(rdn, nitem - cur.lo, s_raw);
        l1_patch_tail(rdn, nitem - cur.lo, s_raw);
        next_L = rdn.L;
      }
    });
    // digit << 16 | rank of each slot; the digit is the top bits alone, or
    // owner * 2^b1 + top bits.  The rank is joined to the digit after the
    // branch, so the atomics of a thread stay in flight together (joined
    // inside, every atomic waited for its result before the next went out).
    uint32_t dg[ITEMS], rk[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) dg[j] = rk[j] = 0;
    if (m.d1.num_pes <= 1) {
#pragma unroll
      for (int j = 0; j < ITEMS; ++j)
        if ((vmask >> j) & 1) {
          dg[j] = key_bits(keys[j], m.d1.shift, m.d1.nbits);
          rk[j] = atomicAdd(&s_h[dg[j]], 1u);
        }
    } else {
#pragma unroll
      for (int j = 0; j < ITEMS; ++j)
        if ((vmask >> j) & 1) {
          dg[j] = digit_eval(m.d1, keys[j], rd.seed);
          rk[j] = atomicAdd(&s_h[dg[j]], 1u);
        }
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) dr[j] = (dg[j] << 16) | rk[j];
    __syncthreads();
    scan_lds<NT>(s_h, m.nb1, s_wave, &s_total);
    const uint32_t total = s_h[m.nb1];
    // the next tile's records arrived while this one was processed: settle
    // them before this tile's stores, so the next iteration waits for nothing
    __builtin_amdgcn_s_waitcnt(VMCNT0);
    uint16_t *off = m.off1 + tile;
    if (tid <= m.nb1) off[(uint64_t)tid * m.ntiles1] = (uint16_t)s_h[tid];
    if (tid == 0 && m.nb1 >= NT) off[(uint64_t)NT * m.ntiles1] = (uint16_t)s_h[NT];
    if (tid == 0 && total)
      atomicAdd(reinterpret_cast<unsigned long long *>(n_dev), (unsigned long long)total);
    // places first (every slot reads its digit's start, an empty one s_h[0]),
    // then the writes: a read inside the write's branch waited for each
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) dr[j] = s_h[dr[j] >> 16] + (dr[j] & 0xFFFFu);
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
      if ((vmask >> j) & 1) sm.stage[dr[j]] = keys[j];
    __syncthreads();
    if (nitem < hi) {  // the next tile's digit counts and record ends (this tile's are read)
      s_h[tid] = 0;
      if (tid == 0) s_h[1024] = 0;
      if (tid < L1_MAXRT) s_mpos[tid] = next_L;
    }
    write_staged<KeyT, NT, ITEMS>(sm.stage, out1 + (uint64_t)tile * m.tcap1, total, tid);
  }
}

// L1 from a key array (the owner side of a P-PE run counts the k-mers it
// received, kcounter.cpp:385-411 over vectordbg): tile t is keys [t * TILE,
// (t + 1) * TILE) of rd.reads (rd.nrec keys), ranked by the first-level digit
// and written contiguously at tile * tcap1 with its bucket starts in off1,
// the layout k_l1 leaves for the second level.  One tile per workgroup.
template <typename KeyT, int NT, int ITEMS>
__global__ __launch_bounds__(NT) void k_l1_keys(JobDesc one) {
  constexpr int TILE = NT * ITEMS;
  __shared__ KeyT stage[TILE];
  __shared__ uint32_t s_h[1024 + 1];
  __shared__ uint32_t s_wave[NT / 64], s_total;
  static_assert(NT >= 1024, "one histogram slot per thread");
  const ReadsDesc &rd = one.rd;
  const MsdArgs &m = one.m;
  const uint32_t tile = blockIdx.x, tid = threadIdx.x;
  const KeyT *in = reinterpret_cast<const KeyT *>(rd.reads);
  const uint64_t t0 = (uint64_t)tile * TILE;
  KeyT keys[ITEMS];
  uint32_t vmask = 0;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint64_t i = t0 + (uint64_t)j * NT + tid;
    keys[j] = KeyT{};
    if (i < rd.nrec) {
      keys[j] = in[i];
      vmask |= 1u << j;
    }
  }
  s_h[tid] = 0;
  if (tid == 0) s_h[1024] = 0;
  __syncthreads();
  // digit << 16 | rank within the digit, joined after the branch so the
  // atomics stay in flight together
  uint32_t dr[ITEMS], rk[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    dr[j] = rk[j] = 0;
    if ((vmask >> j) & 1) {
      dr[j] = digit_eval(m.d1, keys[j], rd.seed);
      rk[j] = atomicAdd(&s_h[dr[j]], 1u);
    }
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) dr[j] = (dr[j] << 16) | rk[j];
  __syncthreads();
  scan_lds<NT>(s_h, m.nb1, s_wave, &s_total);
  const uint32_t total = s_h[m.nb1];
  uint16_t *off = m.off1 + tile;
  if (tid <= m.nb1) off[(uint64_t)tid * m.ntiles1] = (uint16_t)s_h[tid];
  if (tid == 0 && m.nb1 >= NT) off[(uint64_t)NT * m.ntiles1] = (uint16_t)s_h[NT];
  if (tid == 0 && total) atomicAdd(reinterpret_cast<unsigned long long *>(one.nd), (unsigned long long)total);
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) dr[j] = s_h[dr[j] >> 16] + (dr[j] & 0xFFFFu);  // places, then the writes
#pragma unroll
  for (int j = 0; j < ITEMS; ++j)
    if ((vmask >> j) & 1) stage[dr[j]] = keys[j];
  __syncthreads();
  write_staged<KeyT, NT, ITEMS>(stage, reinterpret_cast<KeyT *>(one.out1) + (uint64_t)tile * m.tcap1, total, tid);
}

// Digit histograms (<= 256 bins each) of a key array, for the LSD path over
// received keys.  Grid-stride; LDS counts flushed once per workgroup.
template <typename KeyT, int NT>
__global__ __launch_bounds__(NT) void k_hist_keys(const KeyT *keys, uint64_t n, HistSpecs hs, uint64_t seed,
                                                  uint64_t *hist) {
  __shared__ uint32_t s_h[MAX_PASSES * 256];
  const int nh = hs.n * 256;
  for (int i = threadIdx.x; i < nh; i += NT) s_h[i] = 0;
  __syncthreads();
  for (uint64_t i = blockIdx.x * (uint64_t)NT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * NT) {
    const KeyT x = keys[i];
    for (int p = 0; p < hs.n; ++p) atomicAdd(&s_h[p * 256 + digit_eval(hs.d[p], x, seed)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nh; i += NT)
    if (s_h[i]) atomicAdd(reinterpret_cast<unsigned long long *>(&hist[i]), (unsigned long long)s_h[i]);
}

// L2: bucket b, group g of G consecutive L1 tiles -> its keys in second-level
// digit order, written at (b * ngroups + g) * cap2, with the digit starts in
// off2[b][g][digit].  Items are numbered g * nb1 + b.  Persistent like k_l1:
// each XCD owns a contiguous range of items, so the workgroups of an XCD take
// consecutive buckets of one group, whose segments are neighbours in every L1
// tile (the cache lines they share stay in that XCD's L2).  The next item's
// keys are gathered into a second register set while the current item is
// split and written.
template <typename KeyT, int NT, int ITEMS, bool ONE>
__global__ __launch_bounds__(NT) void k_l2(BatchArgs a, JobDesc one) {
  constexpr int CAP = NT * ITEMS;
  static_assert(NT >= 1024, "one second-level digit (nd2 <= 1024) per thread");
  __shared__ KeyT stage[CAP];
  __shared__ GatherSmem<NT, CAP> s_g;
  __shared__ uint32_t s_h[1024 + 1];
  __shared__ uint32_t s_wave[NT / 64], s_total, s_gwave[NT / 64], s_gtotal;
  __shared__ uint32_t s_job;
  const uint32_t nitems = ONE ? one.m.nb1 * one.m.ngroups : pre_at(a.pre2, a.njobs);
  const uint32_t G = gridDim.x, x = blockIdx.x & 7, nx = (G - x + 7) >> 3;
  const uint32_t per = nitems >> 3, rem = nitems & 7;
  const uint32_t lo = x * per + (x < rem ? x : rem);
  const uint32_t hi = lo + per + (x < rem ? 1u : 0u);
  uint32_t item = lo + (blockIdx.x >> 3);

  // gather item `it` into dst (loads left in flight); returns the key count
  auto gather_item = [&](uint32_t it, JobCursor &jc, KeyT (&dst)[ITEMS], uint32_t &mn) -> uint32_t {
    if (!ONE) cursor_seek(jc, a.pre2, a.njobs, it, &s_job);
    const JobDesc Jg = ONE ? one : job_desc(a, jc.job);
    const MsdArgs mg = Jg.m;
    const uint32_t wgi = it - jc.lo;
    const uint32_t gg = wgi / mg.nb1, bb = wgi - gg * mg.nb1;
    const uint32_t t0 = gg * mg.G;
    const uint32_t nt = mg.ntiles1 - t0 < mg.G ? mg.ntiles1 - t0 : mg.G;
    const uint16_t *row0 = mg.off1 + (uint64_t)bb * mg.ntiles1 + t0;
    uint32_t r0[1], r1[1];
    load_rows<1>(row0, row0 + mg.ntiles1, nt, r0, r1);
    uint32_t t = gather_segments<KeyT, NT, ITEMS, 1, NT, true>(
        reinterpret_cast<const KeyT *>(Jg.out1), r0, r1, nt,
        (uint64_t)t0 * mg.tcap1, mg.tcap1, s_g, s_gwave, &s_gtotal, dst, mn);
    if (t > (uint32_t)CAP) {
      if (threadIdx.x == 0) atomicOr(Jg.err, ERR_OVERFLOW);
      t = 0;
    }
    return t;
  };

  KeyT v[ITEMS];
  JobCursor cur;  // the job of the item being processed
  uint32_t mine = 0, tot = 0;
  if (item < hi) tot = gather_item(item, cur, v, mine);
  __builtin_amdgcn_s_waitcnt(VMCNT0);
  for (; item < hi; item += nx) {
    const JobDesc J = ONE ? one : job_desc(a, cur.job);
    const MsdArgs m = J.m;
    KeyT *out2 = reinterpret_cast<KeyT *>(J.out2);
    const uint32_t wg = item - cur.lo;
    const uint32_t g = wg / m.nb1, b = wg - g * m.nb1;
    const uint32_t tid = opaque_tid();
    s_h[tid] = 0;  // the previous item's split is done (its stage write-out follows a barrier)
    if (tid == 0) s_h[1024] = 0;
    __syncthreads();
    // (the join inside the branch makes each atomic wait for its result
    // before the next goes out; the deferred join of k_l1 measured 1-2%
    // slower here, where the ranking overlaps the next item's gather)
    uint32_t dr[ITEMS];  // digit << 16 | rank within the digit (< 2^14 keys per item)
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      dr[u] = 0;
      if ((uint32_t)u < mine) {
        const uint32_t d = key_bits(v[u], m.d2.shift, m.d2.nbits);  // levels 2 and 3 split by key bits
        dr[u] = (d << 16) | atomicAdd(&s_h[d], 1u);
      }
    }
    // the next item's keys stream in while this one is split; the gather's
    // barriers also order the ranking before the scan
    KeyT vn[ITEMS];
    JobCursor cn = cur;
    uint32_t mn = 0, tn = 0;
    const uint32_t nitem = item + nx;
    if (nitem < hi) tn = gather_item(nitem, cn, vn, mn);
    else __syncthreads();
    scan_lds<NT>(s_h, m.nd2, s_wave, &s_total);
    // this item's digit starts, one contiguous row of off2 [bucket][group][digit]
    uint16_t *off = m.off2 + ((uint64_t)b * m.ngroups + g) * (m.nd2 + 1);
    if (tid <= m.nd2) off[tid] = (uint16_t)(tot ? s_h[tid] : 0);
    if (tid == 0 && m.nd2 >= NT) off[NT] = (uint16_t)(tot ? s_h[NT] : 0);
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) dr[u] = s_h[dr[u] >> 16] + (dr[u] & 0xFFFFu);  // places, then the writes
#pragma unroll
    for (int u = 0; u < ITEMS; ++u)
      if ((uint32_t)u < mine) stage[dr[u]] = v[u];
    __syncthreads();
    // the next item's keys have arrived: settle them before this item's
    // stores, so the next iteration waits for nothing
    __builtin_amdgcn_s_waitcnt(VMCNT0);
    write_staged<KeyT, NT, ITEMS>(stage, out2 + ((uint64_t)b * m.ngroups + g) * m.cap2, tot, tid);
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) v[u] = vn[u];
    cur = cn;
    mine = mn;
    tot = tn;
  }
}

// Exclusive prefix of sub-bucket `idx` over all earlier ones, one wavefront
// reading 64 predecessors' flags at a time.  Called by a whole wavefront.
// Bounded like lb_lookback: the wave polls its 64 predecessors together (one
// poll count for the wave, in a scalar register); when it gives up, the lanes
// still waiting take their predecessor as published with an inclusive prefix
// of 0 and ERR_STALL is set.
DK_DEV uint64_t lb_lookback_wave(uint64_t *flags, uint32_t idx, uint32_t epoch, uint32_t *err) {
  const uint32_t lane = __lane_id();
  uint64_t excl = 0;
  int64_t base = (int64_t)idx - 1;
  while (base >= 0) {
    const int64_t t = base - (int64_t)lane;
    uint64_t val = 0;
    uint32_t st = LB_INC;
    bool ready = t < 0;
    for (uint32_t polls = 0;; ++polls) {
      if (!ready) {
        const uint64_t w = lb_load(flags + t);
        st = (uint32_t)((w >> 48) & 3);
        if ((uint32_t)(w >> 50) == epoch && st != 0) {
          val = w & 0xFFFFFFFFFFFFULL;
          ready = true;
        }
      }
      if (!__ballot(!ready)) break;
      if (polls >= LB_SPIN_MAX) {
        if (!ready) {
          atomicOr(err, ERR_STALL);
          st = LB_INC;
          val = 0;
        }
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    const uint64_t inc = __ballot(st == LB_INC);
    const uint32_t first = inc ? (uint32_t)__ffsll((unsigned long long)inc) - 1 : 64u;
    uint64_t v = lane <= first ? val : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    excl += v;
    if (inc) break;
    base -= 64;
  }
  return excl;
}

// Publish sub-bucket sb's distinct count H and return its exclusive prefix
// over the earlier sub-buckets (decoupled look-back by wavefront 0).  Called
// by the whole workgroup; ends with a barrier.
template <typename KeyT>
DK_DEV uint32_t fin_bin(const MsdArgs &m, const FinItem &f, const KeyT &x) {
  return ((key_bits(x, m.d2.shift, m.d2.nbits) - f.s0) << m.d3.nbits) | key_bits(x, m.d3.shift, m.d3.nbits);
}

// Bins of the final level's general path: as many bits below the digit as
// keep the item's cnt << b3 bins within nb3 (the plan's b3 is sized for its
// widest items; an item of few digits, as canonical counts make where keys
// are dense, gets finer bins, and the general path's ranking costs grow with
// the square of a bin's keys).
DK_DEV uint32_t fin_gen_b3(const MsdArgs &m, uint32_t cnt, uint32_t nb3) {
  uint32_t b3 = 0;
  while (b3 < m.d2.shift && b3 < 11 && (cnt << (b3 + 1)) <= nb3) ++b3;
  return b3;
}
template <typename KeyT>
DK_DEV uint32_t fin_gen_bin(const MsdArgs &m, uint32_t s0, uint32_t b3, const KeyT &x) {
  return ((key_bits(x, m.d2.shift, m.d2.nbits) - s0) << b3) | key_bits(x, m.d2.shift - b3, b3);
}

DK_DEV uint64_t fin_prefix(uint64_t *flags, uint32_t sb, uint32_t nsb, uint32_t H, uint64_t *n_dev,
                           uint32_t *err, uint64_t *s_excl) {
  constexpr uint32_t epoch = 1;  // flags are zeroed before every launch
  const uint32_t tid = opaque_tid();
  if ((tid >> 6) == 0) {
    if (tid == 0) lb_store(flags + sb, lb_pack(epoch, sb == 0 ? LB_INC : LB_AGG, H));
    uint64_t excl = 0;
    if (sb > 0) {
      excl = lb_lookback_wave(flags, sb, epoch, err);
      if (tid == 0) lb_store(flags + sb, lb_pack(epoch, LB_INC, excl + H));
    }
    if (tid == 0) {
      *s_excl = excl;
      if (sb == nsb - 1) n_dev[1] = excl + H;
    }
  }
  __syncthreads();
  return *s_excl;
}

constexpr uint64_t HASH_EMPTY = ~0ull;  // no k-mer of k < 32 bases has all 64 bits set

// Empty hash table of fin_distinct: HT keys of HASH_EMPTY, HT zero counts.
template <int NT, int HT>
DK_DEV void fin_clear(uint64_t *lds) {
  uint32_t *tc = reinterpret_cast<uint32_t *>(lds + HT);
  const uint32_t tid = opaque_tid();
#pragma unroll
  for (int j = 0; j < HT / NT; ++j) {
    lds[tid + j * NT] = HASH_EMPTY;
    tc[tid + j * NT] = 0;
  }
}

// Distinct-key path of the final level (keys of k < 32 bases): the gathered
// keys go into an LDS hash table (equal keys add to one count), the occupied
// slots are split by the next b3 bits, the few distinct keys of each bin are
// ranked by counting, and each (key, count) pair goes from registers to its
// place in the output.  The table is HT slots of key + count; the binned
// pairs later take its place.
// Returns false, having changed nothing the other path needs, when a key
// finds no slot within 32 probes.  Uniform over the workgroup.
// Slot after probe pr (0, 1, ...) of a probe sequence that steps 1, 2, 3, ...
// from the home slot (offsets 0, 1, 3, 6, ...), wrapping at HT.  Growing steps
// break up the runs of occupied slots that unit steps pile keys into, so
// chains stay shorter at the same load; every copy of a key still walks the
// same sequence.
template <int HT>
DK_DEV uint32_t probe_next(uint32_t h, int pr) {
  static_assert(HT > 32, "a step is below HT");
  const uint32_t n = h + (uint32_t)pr + 1;  // pr + 1 <= 32 < HT
  return n >= (uint32_t)HT ? n - (uint32_t)HT : n;
}

template <int NT, int ITEMS, int HT>
DK_DEV bool fin_distinct(const JobDesc &J, const MsdArgs &m, const uint64_t (&keys)[ITEMS],
                         uint32_t mine, uint64_t *lds, uint32_t *s_h, uint32_t *s_wave,
                         uint32_t *s_total, uint32_t *s_fail, uint64_t *s_excl, uint32_t sb,
                         uint32_t nsb, const FinItem &fi, bool synced) {
  constexpr int SPT = HT / NT;  // table slots per thread
  const uint32_t tid = opaque_tid();
  uint64_t *tk = lds;                                     // HT keys, then the ordered keys
  uint32_t *tc = reinterpret_cast<uint32_t *>(lds + HT);  // HT counts, then the ordered counts
  uint32_t *hc = reinterpret_cast<uint32_t *>(reinterpret_cast<uint8_t *>(J.nd) + 24);  // fail, ok
  // *s_fail: the caller's skip decision (set before the keys were loaded);
  // the caller has cleared the table (fin_clear) while the keys loaded, and
  // passed a barrier since when `synced`
  if (!synced) __syncthreads();
  if (*s_fail) return false;
  // first probes of all keys back to back; collisions walk on from there
  uint32_t hh[ITEMS];
  uint64_t old[ITEMS];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    if ((uint32_t)u < mine) {
      hh[u] = (uint32_t)((((keys[u] * 0x9E3779B97F4A7C15ull) >> 32) * HT) >> 32);  // [0, HT)
      old[u] = atomicCAS((unsigned long long *)&tk[hh[u]], HASH_EMPTY, keys[u]);
    }
  }
  bool lost = false;
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    if ((uint32_t)u < mine) {
      const uint64_t x = keys[u];
      uint32_t h = hh[u];
      uint64_t o = old[u];
      for (int pr = 0; pr < 32 && o != HASH_EMPTY && o != x; ++pr) {
        h = probe_next<HT>(h, pr);
        o = atomicCAS((unsigned long long *)&tk[h], HASH_EMPTY, x);
      }
      if (o == HASH_EMPTY || o == x) atomicAdd(&tc[h], 1u);
      else lost = true;
    }
  }
  if (lost) *s_fail = 1;
  __syncthreads();
  const bool fits = !*s_fail;
  if (tid == 0) atomicAdd(hc + (fits ? 1 : 0), 1u);
  if (!fits) {
    __syncthreads();  // everyone has read s_fail before the other path reuses LDS
    return false;
  }
  // split the occupied slots by the next b3 bits; counts move with their keys
  const uint32_t nd3 = fi.nbins;
  {
    uint64_t ek[SPT];
    uint32_t ec[SPT], ed[SPT], er[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      ek[j] = tk[tid + j * NT];
      ec[j] = tc[tid + j * NT];
      ed[j] = er[j] = 0;
      if (ek[j] != HASH_EMPTY) {
        ed[j] = fin_bin(m, fi, ek[j]);
        er[j] = atomicAdd(&s_h[ed[j]], 1u);
      }
    }
    __syncthreads();
    scan_lds<NT, false>(s_h, nd3, s_wave, s_total);  // at the register limit
    // places first (an empty slot reads s_h[0]), then the writes: a read
    // inside the write's branch waited for each
#pragma unroll
    for (int j = 0; j < SPT; ++j) er[j] += s_h[ed[j]];
#pragma unroll
    for (int j = 0; j < SPT; ++j)
      if (ek[j] != HASH_EMPTY) {
        tk[er[j]] = ek[j];
        tc[er[j]] = ec[j];
      }
  }
  const uint32_t H = s_h[nd3];
  constexpr uint32_t epoch = 1;
  if (tid == 0) lb_store(J.flags + sb, lb_pack(epoch, sb == 0 ? LB_INC : LB_AGG, H));  // early
  __syncthreads();
  // order the distinct keys of each bin, positions tid, tid + NT, ... < H:
  // place = bin start + number of smaller keys in the bin
  // (the pairs, then their bins' bounds, each group of reads waiting once;
  // positions at or past H read table slots and bin 0, and are not used)
  uint64_t x[SPT];
  uint32_t xc[SPT], place[SPT], bnd[SPT];
  bool big = false;
#pragma unroll
  for (int i = 0; i < SPT; ++i) {
    x[i] = tk[tid + i * NT];
    xc[i] = tc[tid + i * NT];
  }
#pragma unroll
  for (int i = 0; i < SPT; ++i) {
    const uint32_t d = tid + i * NT < H ? fin_bin(m, fi, x[i]) : 0u;
    bnd[i] = s_h[d] | (s_h[d + 1] << 16);  // bin starts are below HT <= 2^16
  }
#pragma unroll
  for (int i = 0; i < SPT; ++i) {
    const uint32_t p = tid + i * NT;
    if (p < H) {
      const uint32_t lo = bnd[i] & 0xFFFFu, hi = bnd[i] >> 16;
      uint32_t r = 0;
      if (hi - lo > 1) {
        if (hi - lo > 256) {
          big = true;
        } else {
          // all 8 reads go out before one wait (a read under `t0 + q < hi`
          // waited for each); past hi they see the next bins or, past H <=
          // HT, the counts after the keys, and are masked out
          for (uint32_t t0 = lo; t0 < hi; t0 += 8) {
            uint64_t y[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) y[q] = tk[t0 + q];
#pragma unroll
            for (int q = 0; q < 8; ++q) r += (t0 + q < hi) & (y[q] < x[i]);
          }
        }
      }
      place[i] = lo + r;
    }
  }
  if (big) atomicOr(J.err, ERR_OVERFLOW);  // recounted on the LSD path
  // the bins are contiguous, so a wave's 64 pairs land in a few lines
  const uint64_t excl = fin_prefix(J.flags, sb, nsb, H, J.nd, J.err, s_excl);
  uint64_t *out_keys = reinterpret_cast<uint64_t *>(J.tkeys);
#pragma unroll
  for (int i = 0; i < SPT; ++i)
    if (tid + i * NT < H) {
      out_keys[excl + place[i]] = x[i];
      J.tcounts[excl + place[i]] = xc[i];
    }
  return true;
}

// Distinct-key path of the final level for two-word keys (k = 33..64): the
// gathered keys go into an LDS table of HT slots of (claim word, hi, count).
// A key's claim word is lo ^ hi * C (C odd, so lo = word ^ hi * C): it mixes
// both words, so keys that share a low word (k-mers that differ only in their
// first 32 - k bases, as substitution errors often make) still differ in it.
// First every key walks to the slot holding its claim word or claims an
// empty one by a CAS, and a claimer writes its high word; after a barrier
// every key compares its slot's high word with its own and adds 1 to the
// count (a mismatch, two keys with one claim word, sends the item to the
// other path).  As in fin_distinct the first probes of all keys go out back
// to back; the occupied slots are split by the next b3 bits through a 16-bit
// slot index per key in bin order, the distinct keys of each bin are ranked
// by counting, and each pair is written to its place.
// Returns false, having changed nothing the other path needs, when a key
// finds no slot within 32 probes or its claim word is the empty word.
constexpr uint64_t K2_MIX = 0x9E3779B97F4A7C15ull;

DK_DEV uint64_t k2_claim(const K2 &x) { return x.lo ^ (x.hi * K2_MIX); }

template <int NT, int ITEMS, int HT>
DK_DEV bool fin_distinct2(const JobDesc &J, const MsdArgs &m, const K2 (&keys)[ITEMS], uint32_t mine,
                          uint8_t *lds, uint32_t *s_h, uint32_t *s_wave, uint32_t *s_total, uint32_t *s_fail,
                          uint64_t *s_excl, uint32_t sb, uint32_t nsb, const FinItem &fi) {
  constexpr int SPT = HT / NT;  // table slots per thread
  const uint32_t tid = opaque_tid();
  uint64_t *kc = reinterpret_cast<uint64_t *>(lds), *khi = kc + HT;
  uint32_t *tc = reinterpret_cast<uint32_t *>(khi + HT);
  uint16_t *ix = reinterpret_cast<uint16_t *>(tc + HT);
  uint32_t *hc = reinterpret_cast<uint32_t *>(reinterpret_cast<uint8_t *>(J.nd) + 24);  // fail, ok
#pragma unroll
  for (int j = 0; j < SPT; ++j) {
    kc[tid + j * NT] = HASH_EMPTY;
    tc[tid + j * NT] = 0;
  }
  __syncthreads();
  if (*s_fail) return false;
  bool lost = false;
  uint32_t hh[ITEMS];
  uint64_t old[ITEMS];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    if ((uint32_t)u < mine) {
      const uint64_t c = k2_claim(keys[u]);
      lost |= c == HASH_EMPTY;
      hh[u] = (uint32_t)((((c * 0xC2B2AE3D27D4EB4Full) >> 32) * HT) >> 32);  // [0, HT)
      old[u] = atomicCAS((unsigned long long *)&kc[hh[u]], HASH_EMPTY, c);
    }
  }
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    if ((uint32_t)u < mine) {
      const uint64_t c = k2_claim(keys[u]);
      uint32_t h = hh[u];
      uint64_t o = old[u];
      for (int pr = 0; pr < 32 && o != HASH_EMPTY && o != c; ++pr) {
        h = probe_next<HT>(h, pr);
        o = atomicCAS((unsigned long long *)&kc[h], HASH_EMPTY, c);
      }
      if (o == HASH_EMPTY) khi[h] = keys[u].hi;  // claimed here
      else lost |= o != c;
      hh[u] = h;
    }
  }
  __syncthreads();  // every claimer's high word is written
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    if ((uint32_t)u < mine) {
      if (khi[hh[u]] == keys[u].hi) atomicAdd(&tc[hh[u]], 1u);
      else lost = true;
    }
  }
  if (lost) *s_fail = 1;
  __syncthreads();
  const bool fits = !*s_fail;
  if (tid == 0) atomicAdd(hc + (fits ? 1 : 0), 1u);
  if (!fits) {
    __syncthreads();  // everyone has read s_fail before the other path reuses LDS
    return false;
  }
  // the keys of the occupied slots
  auto key_at = [&](uint32_t sl) {
    const uint64_t hi = khi[sl];
    return K2{kc[sl] ^ (hi * K2_MIX), hi};
  };
  // split the occupied slots by the next b3 bits: ix[p] is the slot of the
  // p-th key in bin order, and the keys stay where they are
  const uint32_t nd3 = fi.nbins;
  {
    uint32_t ed[SPT], er[SPT];
    uint32_t occ = 0;
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      const uint32_t i = tid + j * NT;
      if (kc[i] != HASH_EMPTY) {
        occ |= 1u << j;
        ed[j] = fin_bin(m, fi, key_at(i));
        er[j] = atomicAdd(&s_h[ed[j]], 1u);
      }
    }
    __syncthreads();
    scan_lds<NT, false>(s_h, nd3, s_wave, s_total);
#pragma unroll
    for (int j = 0; j < SPT; ++j)
      if ((occ >> j) & 1) ix[s_h[ed[j]] + er[j]] = (uint16_t)(tid + j * NT);
  }
  const uint32_t H = s_h[nd3];
  constexpr uint32_t epoch = 1;
  if (tid == 0) lb_store(J.flags + sb, lb_pack(epoch, sb == 0 ? LB_INC : LB_AGG, H));  // early
  __syncthreads();
  // place = bin start + number of smaller keys in the bin
  uint32_t place[SPT];
  bool big = false;
#pragma unroll
  for (int i = 0; i < SPT; ++i) {
    const uint32_t p = tid + i * NT;
    place[i] = 0;
    if (p < H) {
      const uint32_t sl = ix[p];
      const K2 x = key_at(sl);
      const uint32_t d = fin_bin(m, fi, x);
      const uint32_t lo = s_h[d], hi = s_h[d + 1];
      uint32_t r = 0;
      if (hi - lo > 1) {
        if (hi - lo > 256) {
          big = true;
        } else {
          for (uint32_t t0 = lo; t0 < hi; t0 += 4) {
            K2 y[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              y[q] = key_at(t0 + q < hi ? ix[t0 + q] : sl);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) r += key_lt(y[q], x) ? 1u : 0u;
          }
        }
      }
      place[i] = lo + r;
    }
  }
  if (big) atomicOr(J.err, ERR_OVERFLOW);  // recounted on the LSD path
  const uint64_t excl = fin_prefix(J.flags, sb, nsb, H, J.nd, J.err, s_excl);
  K2 *out_keys = reinterpret_cast<K2 *>(J.tkeys);
#pragma unroll
  for (int i = 0; i < SPT; ++i) {
    const uint32_t p = tid + i * NT;
    if (p < H) {
      const uint32_t sl = ix[p];
      out_keys[excl + place[i]] = key_at(sl);
      J.tcounts[excl + place[i]] = tc[sl];
    }
  }
  return true;
}

// Final: one sub-bucket (bucket b, digit s) gathered from the L2 groups of b
// into LDS, split by the next b3 bits, each bin ordered by counting, run-length
// reduced, and written at the place the decoupled look-back gives.
// Four waves per SIMD (128 VGPRs at most): one 1024-, two 512- or four
// 256-thread workgroups per CU.
template <typename KeyT, int NT, int ITEMS, int NB3, int HT, bool ONE>
__global__ __launch_bounds__(NT, 4) void k_fin(BatchArgs a, JobDesc one) {
  constexpr int CAP = NT * ITEMS;
  constexpr int TPT = FIN_MAX_GROUPS / NT;
  __shared__ KeyT B[CAP + 8];  // the rank reads up to 7 keys past a bin's end
  __shared__ union {
    GatherSmem<FIN_MAX_GROUPS, CAP> seg;
    uint16_t hp[CAP];
  } s_u;
  __shared__ uint32_t s_h[NB3 + 1];
  __shared__ uint32_t s_mixed[NB3 / 32];
  __shared__ uint32_t s_wave[NT / 64], s_wmin[NT / 64], s_total, s_sb, s_job, s_fail, s_next, s_dup;
  __shared__ uint64_t s_excl;
  const uint32_t tid = threadIdx.x;

  // the overflow counters of the distinct-key path (fail, ok) of a job
  auto fail_word = [](const JobDesc &J) {
    return __hip_atomic_load(reinterpret_cast<const uint64_t *>(reinterpret_cast<const uint8_t *>(J.nd) + 24),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  // group rows of sub-bucket sb (loads left in flight)
  auto rows_of = [&](const MsdArgs &m, const FinItem &f, uint32_t (&r0)[TPT], uint32_t (&r1)[TPT]) {
    // digit starts s0 and s0 + cnt of every group of the bucket (rows of the
    // bucket's groups lie together, so the items of a bucket share them in L2)
    const uint16_t *base = m.off2 + (uint64_t)f.b * m.ngroups * (m.nd2 + 1);
    const uint32_t tid = opaque_tid();
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
      const uint32_t q = tid * TPT + u;
      r0[u] = r1[u] = 0;
      if (q < m.ngroups) {
        const uint16_t *row = base + (uint64_t)q * (m.nd2 + 1);
        r0[u] = row[f.s0];
        r1[u] = row[f.s0 + f.cnt];
      }
    }
  };

  // Set by run when its item ended on the one-word distinct-key path, which
  // reads no LDS after its last barrier (it writes its pairs from registers).
  bool regs_out = false;

  // One final item fi of job J, its group rows in r0/r1 and the job's
  // overflow counters in fo (tid 0).  `between` runs right after the gather,
  // when the rows are free again, and returns whether it passed a barrier
  // (uniform); every thread returns together.  Returns
  // true, before `between`, when the item holds more keys than LDS and more
  // than one digit: the caller then runs its digits one by one.
  auto run = [&](const JobDesc &J, const FinItem &fi, uint32_t (&r0)[TPT], uint32_t (&r1)[TPT], uint64_t fo,
                 auto &&between) -> bool {
    const uint32_t tid = opaque_tid();  // per item: no address is kept across items
    // distinct-key path: skip it once most earlier sub-buckets of the job
    // overflowed the table
    if (tid == 0) s_fail = (uint32_t)fo > (uint32_t)(fo >> 32) + 32 ? 1u : 0u;
    const MsdArgs m = J.m;
    const KeyT *in2 = reinterpret_cast<const KeyT *>(J.out2);
    KeyT *out_keys = reinterpret_cast<KeyT *>(J.tkeys);
    uint32_t *out_counts = J.tcounts;
    uint64_t *n_dev = J.nd;
    uint32_t *err = J.err;
    const uint32_t nd3 = fi.nbins;
    // cleared bins and mixed flags: the gather's barriers order these
    // stores before their first use (the split, after the inserts)
    for (uint32_t i = tid; i <= nd3; i += NT) s_h[i] = 0;
    for (uint32_t w = tid; w < NB3 / 32; w += NT) s_mixed[w] = 0;
    // look-back slots are digits: this item's count goes to its last digit
    const uint32_t nsb = m.nb1 * m.nd2;
    const uint32_t sb = fi.b * m.nd2 + fi.s0 + fi.cnt - 1;
    const uint32_t b = fi.b;

    KeyT keys[ITEMS];
    uint32_t mine;
    const uint32_t tot = gather_segments<KeyT, NT, ITEMS, TPT, FIN_MAX_GROUPS, false>(
        in2, r0, r1, m.ngroups, (uint64_t)b * m.ngroups * m.cap2, m.cap2, s_u.seg, s_wave, &s_total,
        keys, mine);
    if (tot > (uint32_t)CAP && fi.cnt > 1) return true;
    // the distinct-key path's table is cleared while the keys load (it
    // shares no LDS with the gather)
    if constexpr (sizeof(KeyT) == 8)
      if (J.rd.k < 32) fin_clear<NT, HT>(reinterpret_cast<uint64_t *>(B));
    const bool synced = between();
    if (tot > (uint32_t)CAP && tid == 0) atomicOr(err, ERR_OVERFLOW);
    const uint32_t mcnt = tot > (uint32_t)CAP ? 0 : tot;
    // the item's inner digits add nothing to the prefix
    if (tid + 1 < fi.cnt) lb_store(J.flags + sb + 1 - fi.cnt + tid, lb_pack(1, LB_AGG, 0));

    if constexpr (sizeof(KeyT) == 8) {
      static_assert((size_t)HT * 12 <= sizeof(B) && HT % NT == 0, "hash table does not fit the key buffer");
      if (J.rd.k < 32 && fin_distinct<NT, ITEMS, HT>(J, m, keys, mine, reinterpret_cast<uint64_t *>(B),
                                                         s_h, s_wave, &s_total, &s_fail, &s_excl, sb,
                                                         nsb, fi, synced)) {
        regs_out = true;
        return false;
      }
    } else {
      static_assert((size_t)HT * 22 <= sizeof(B) && HT % NT == 0 && HT <= 65536,
                    "hash table does not fit the key buffer");
      if (fin_distinct2<NT, ITEMS, HT>(J, m, keys, mine, reinterpret_cast<uint8_t *>(B), s_h, s_wave, &s_total,
                                       &s_fail, &s_excl, sb, nsb, fi))
        return false;
    }

    // split by the item's own bin bits into B (unstable, LDS atomics); its
    // bins past the distinct-key path's are cleared first
    if (tid == 0) s_dup = 0;  // set below when two keys of the item are equal
    const uint32_t gb3 = fin_gen_b3(m, fi.cnt, NB3), gbins = fi.cnt << gb3;
    if (gbins > nd3) {
      for (uint32_t i = nd3 + 1 + tid; i <= gbins; i += NT) s_h[i] = 0;
      __syncthreads();
    }
    uint32_t dig[ITEMS], rank[ITEMS];
  #pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      dig[u] = rank[u] = 0;
      if ((uint32_t)u < mine) {
        dig[u] = fin_gen_bin(m, fi.s0, gb3, keys[u]);
        rank[u] = atomicAdd(&s_h[dig[u]], 1u);
      }
    }
    __syncthreads();
    scan_lds<NT, false>(s_h, gbins, s_wave, &s_total);  // at the register limit
  #pragma unroll
    for (int u = 0; u < ITEMS; ++u) rank[u] += s_h[dig[u]];  // places, then the writes
  #pragma unroll
    for (int u = 0; u < ITEMS; ++u)
      if ((uint32_t)u < mine) B[rank[u]] = keys[u];
    __syncthreads();

    // Bins whose keys are not all equal (mixed) are ranked below; a mixed bin
    // longer than RANK_MAX flags the job, which is then recounted on the LSD
    // path.  Keys are reloaded in B's order (position tid + q * NT).
    constexpr uint32_t RANK_MAX = 128;
  #pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
      const uint32_t i = tid + q * NT;
      if (i < mcnt) {
        const KeyT x = B[i];
        keys[q] = x;
        dig[q] = fin_gen_bin(m, fi.s0, gb3, x);
        if (!key_eq(B[s_h[dig[q]]], x)) atomicOr(&s_mixed[dig[q] >> 5], 1u << (dig[q] & 31));
      }
    }
    __syncthreads();
    // The place of a key in sorted order is its bin start + the keys of its
    // bin that precede it (smaller, or equal and stored before it); a bin of
    // copies of one k-mer keeps its order.  s_map[place] = position in B.
    uint16_t *s_map = s_u.hp;
    bool bad = false, dup = false;
  #pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
      const uint32_t i = tid + q * NT;
      if (i < mcnt) {
        const KeyT x = keys[q];
        const uint32_t lo = s_h[dig[q]], hi = s_h[dig[q] + 1];
        uint32_t place = i;
        if ((s_mixed[dig[q] >> 5] >> (dig[q] & 31)) & 1) {
          if (hi - lo <= RANK_MAX) {
            // four independent LDS reads per step; the keys read past the
            // bin's end (inside B's padding at most) are masked out
            uint32_t r = 0, e = 0;
            for (uint32_t t0 = lo; t0 < hi; t0 += 4) {
              KeyT y[4];
  #pragma unroll
              for (int j = 0; j < 4; ++j) y[j] = B[t0 + j];
  #pragma unroll
              for (int j = 0; j < 4; ++j) {
                const bool in = t0 + j < hi, eq = in && key_eq(y[j], x);
                r += (in && (key_lt(y[j], x) || (t0 + j < i && eq))) ? 1u : 0u;
                e += eq ? 1u : 0u;
              }
            }
            place = lo + r;
            dup |= e > 1;  // x itself is one of them
          } else {
            bad = true;
          }
        } else {
          dup |= hi - lo > 1;  // a bin of copies of one k-mer
        }
        s_map[place] = (uint16_t)i;
      }
    }
    if (bad) atomicOr(err, ERR_OVERFLOW);  // recounted on the LSD path
    if (dup) s_dup = 1;
    __syncthreads();

    if (!s_dup) {
      // No two keys of the item are equal: each place is a run of one and
      // its own output index, so thread t writes places t, t + NT, ...
      // (consecutive lanes, consecutive entries).
      const uint64_t excl = fin_prefix(J.flags, sb, nsb, mcnt, n_dev, err, &s_excl);
  #pragma unroll
      for (int q = 0; q < ITEMS; ++q) {
        const uint32_t pl = tid + q * NT;
        if (pl < mcnt) {
          const uint32_t sp = s_map[pl];
          out_keys[excl + pl] = B[sp < mcnt ? sp : 0];
          out_counts[excl + pl] = 1;
        }
      }
      return false;
    }

    // Runs in place order, places tid * ITEMS .. + ITEMS - 1 per thread: a
    // place starts a run (a head) when its key differs from the previous
    // place's (keys of different bins differ).  A run ends at the next head:
    // in this thread's places, else at the first head of a later thread
    // (a suffix minimum over the threads), else at mcnt.
    uint32_t heads = 0, hmask = 0, first = 0xFFFFFFFFu;
    {
      const uint32_t p0 = tid * ITEMS;
      KeyT prev{};
      // a place's word is < mcnt unless a bin went unranked (then the job
      // is recounted); the clamp keeps such reads inside B
      auto at = [&](uint32_t pl) -> KeyT {
        const uint32_t sp = s_map[pl];
        return B[sp < mcnt ? sp : 0];
      };
      if (p0 > 0 && p0 < mcnt) prev = at(p0 - 1);
  #pragma unroll
      for (int q = 0; q < ITEMS; ++q) {
        const uint32_t pl = p0 + q;
        if (pl < mcnt) {
          const KeyT x = at(pl);
          const bool h = pl == 0 || !key_eq(x, prev);
          hmask |= (h ? 1u : 0u) << q;
          heads += h;
          if (h && first == 0xFFFFFFFFu) first = pl;
          prev = x;
        }
      }
    }
    // the prefix of the head counts and the suffix minimum of the first heads
    uint32_t o = 0, after = 0;
    {
      const uint32_t lane = tid & 63, wave = tid >> 6;
      uint32_t x = heads, mn = first;
  #pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64), z = __shfl_down(mn, d, 64);
        if (lane >= (uint32_t)d) x += y;
        if (lane + d < 64) mn = min(mn, z);
      }
      if (lane == 63) s_wave[wave] = x;
      if (lane == 0) s_wmin[wave] = mn;
      const uint32_t nx = __shfl_down(mn, 1, 64);
      __syncthreads();
      uint32_t pre = 0, later = lane == 63 ? 0xFFFFFFFFu : nx;
      for (uint32_t w = 0; w < NT / 64; ++w) {
        const uint32_t sw = s_wave[w], mw = s_wmin[w];
        if (w < wave) pre += sw;
        if (w > wave) later = min(later, mw);
      }
      if (tid == NT - 1) s_total = pre + x;
      o = pre + x - heads;
      after = later == 0xFFFFFFFFu ? mcnt : later;
      __syncthreads();
    }
    const uint32_t H = s_total;
    const uint64_t excl = fin_prefix(J.flags, sb, nsb, H, n_dev, err, &s_excl);
    {
      uint32_t end = after;  // the next head after the place at hand
      o += heads;
  #pragma unroll
      for (int q = ITEMS - 1; q >= 0; --q) {
        if ((hmask >> q) & 1) {
          const uint32_t pl = tid * ITEMS + q;
          --o;
          const uint32_t sp = s_map[pl];
          out_keys[excl + o] = B[sp < mcnt ? sp : 0];
          out_counts[excl + o] = end - pl;
          end = pl;
        }
      }
    }
    return false;
  };

  if constexpr (ONE) {
    // One job: persistent over the tickets.  The next ticket is taken before
    // this sub-bucket's gather and read after it, and the next sub-bucket's
    // group rows and overflow counters load while this one is sorted and
    // written.  Tickets keep the look-back's predecessors running.
    const uint32_t n = a.nfin;
    if (tid == 0) s_sb = atomicAdd(a.ticket, 1u);
    __syncthreads();
    uint32_t t = __builtin_amdgcn_readfirstlane(s_sb);
    uint32_t r0[TPT], r1[TPT];
    uint64_t fo = 0;
    if (t < n) {
      rows_of(one.m, fin_item_of(one.m, t), r0, r1);
      if (tid == 0) fo = fail_word(one);
    }
    while (t < n) {
      uint32_t nt = n;  // not persistent: this workgroup's one item only
      if (tid == 0 && a.persist) nt = atomicAdd(a.ticket, 1u);  // its value is needed only after the gather
      uint32_t tn = n;
      uint64_t fon = 0;
      auto fetch_next = [&] {
        if (tid == 0) s_next = nt;
        __syncthreads();
        tn = __builtin_amdgcn_readfirstlane(s_next);
        if (tn < n) {
          rows_of(one.m, fin_item_of(one.m, tn), r0, r1);
          if (tid == 0) fon = fail_word(one);
        }
      };
      // passes: the item, then, when it holds more keys than LDS, its
      // digits one at a time (one call site, so one copy of the code)
      const FinItem f = fin_item_of(one.m, t);
      FinItem cur = f;
      bool digits = false;
      uint32_t d = 0;
      regs_out = false;
      for (;;) {
        const bool split = run(one, cur, r0, r1, fo, [&] {
          if (digits) return false;
          fetch_next();  // its barrier follows the table clear
          return true;
        });
        if (split) digits = true;
        else if (digits) ++d;
        if (!digits) break;
        __syncthreads();  // the previous pass is done with LDS
        if (d == f.cnt) {
          fetch_next();
          break;
        }
        cur = FinItem{f.b, f.s0 + d, 1, 1u << one.m.d3.nbits};
        rows_of(one.m, cur, r0, r1);
      }
      // this item's LDS is free (s_next was read after fetch_next's barrier):
      // an item that wrote its pairs from registers read none since its
      // look-back barrier, so the next item's first stores need no barrier
      if (digits || !regs_out) __syncthreads();
      t = tn;
      fo = fon;
    }
  } else {
    // several jobs: one sub-bucket per workgroup, its job looked up
    if (tid == 0) s_sb = atomicAdd(a.ticket, 1u);  // items in ticket order, job by job
    __syncthreads();
    const uint32_t item = a.item0 + __builtin_amdgcn_readfirstlane(s_sb);
    const uint32_t job = find_job_wg(a.pre3, a.njobs, item, &s_job);
    const JobDesc J = job_desc(a, job);
    uint32_t r0[TPT], r1[TPT];
    const FinItem f = fin_item_of(J.m, item - pre_at(a.pre3, job));
    const uint64_t fo = tid == 0 ? fail_word(J) : 0;
    rows_of(J.m, f, r0, r1);
    FinItem cur = f;
    bool digits = false;  // more keys than LDS holds: then digit by digit
    uint32_t d = 0;
    for (;;) {
      const bool split = run(J, cur, r0, r1, fo, [] { return false; });
      if (split) digits = true;
      else if (digits) ++d;
      if (!digits || d == f.cnt) break;
      __syncthreads();
      cur = FinItem{f.b, f.s0 + d, 1, 1u << J.m.d3.nbits};
      rows_of(J.m, cur, r0, r1);
    }
  }
}

// ------------------------------------------------------------- launchers --

void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st) {
  if (!n) return;
  uint64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  k_fill_u32<<<(uint32_t)g, 256, 0, st>>>(p, v, n);
}
void launch_copy_u32(uint32_t *dst, const uint32_t *src, uint64_t n, hipStream_t st) {
  if (!n) return;
  uint64_t g = (n / 4 + 255) / 256;
  g = g < 1 ? 1 : g > 8192 ? 8192 : g;
  k_copy_u32<<<(uint32_t)g, 256, 0, st>>>(dst, src, n);
}
void launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t st) {
  if (!n) return;
  uint64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  k_fill_u64<<<(uint32_t)g, 256, 0, st>>>(p, v, n);
}

void launch_hist_reads(int kw, const ReadsDesc &rd, const HistSpecs &hs, const Workspace &ws,
                       uint32_t ntiles, int grid, hipStream_t st) {
  if (kw == 1)
    k_hist_reads<uint64_t, SWEEP_NT, SWEEP_ITEMS_1>
        <<<grid, SWEEP_NT, 0, st>>>(rd, hs, ntiles, ws.hist, ws.err);
  else
    k_hist_reads<K2, SWEEP_NT, SWEEP_ITEMS_2><<<grid, SWEEP_NT, 0, st>>>(rd, hs, ntiles, ws.hist, ws.err);
}

void launch_scan_digits(int npasses, const Workspace &ws, hipStream_t st) {
  k_scan_digits<<<1, 256, 0, st>>>(npasses, ws.hist, ws.digit_start, ws.n_dev);
}

void launch_sweep(int kw, bool from_reads, const ReadsDesc &rd, const void *in, void *out,
                  const DigitSpec &dig, const uint64_t *digit_start, const Workspace &ws,
                  uint32_t ticket_slot, uint32_t epoch, uint32_t grid, hipStream_t st) {
  SweepArgs a;
  a.rd = rd;
  a.in = in;
  a.out = out;
  a.n_dev = ws.n_dev;
  a.digit_start = digit_start;
  a.flags = ws.flags;
  a.ticket = ws.tickets + ticket_slot;
  a.dig = dig;
  a.epoch = epoch;
  a.key_cap = ws.key_cap;
  a.err = ws.err;
  if (!grid) return;
  if (kw == 1) {
    if (from_reads) k_sweep<uint64_t, SWEEP_NT, SWEEP_ITEMS_1, true><<<grid, SWEEP_NT, 0, st>>>(a);
    else k_sweep<uint64_t, SWEEP_NT, SWEEP_ITEMS_1, false><<<grid, SWEEP_NT, 0, st>>>(a);
  } else {
    if (from_reads) k_sweep<K2, SWEEP_NT, SWEEP_ITEMS_2, true><<<grid, SWEEP_NT, 0, st>>>(a);
    else k_sweep<K2, SWEEP_NT, SWEEP_ITEMS_2, false><<<grid, SWEEP_NT, 0, st>>>(a);
  }
}

void launch_rle(int kw, const void *sorted, void *out_keys, const Workspace &ws, uint32_t epoch,
                uint32_t grid, hipStream_t st) {
  if (!grid) return;
  uint32_t *ticket = ws.tickets + MAX_PASSES;
  if (kw == 1)
    k_rle<uint64_t, SWEEP_NT, SWEEP_ITEMS_1><<<grid, SWEEP_NT, 0, st>>>(
        (const uint64_t *)sorted, (uint64_t *)out_keys, ws.counts, ws.n_dev, ws.rle_flags, ticket,
        epoch, ws.key_cap, ws.err);
  else
    k_rle<K2, SWEEP_NT, SWEEP_ITEMS_2><<<grid, SWEEP_NT, 0, st>>>(
        (const K2 *)sorted, (K2 *)out_keys, ws.counts, ws.n_dev, ws.rle_flags, ticket, epoch,
        ws.key_cap, ws.err);
}

void launch_bin_offsets_at(int kw, const void *table_keys, const uint64_t *nd, uint32_t num_pes,
                           uint64_t seed, uint64_t *offsets, hipStream_t st) {
  if (kw == 1)
    k_bin_offsets<uint64_t><<<1, 256, 0, st>>>((const uint64_t *)table_keys, nd, num_pes, seed, offsets);
  else
    k_bin_offsets<K2><<<1, 256, 0, st>>>((const K2 *)table_keys, nd, num_pes, seed, offsets);
}

void launch_bin_offsets(int kw, const void *table_keys, uint32_t num_pes, uint64_t seed, const Workspace &ws,
                        hipStream_t st) {
  launch_bin_offsets_at(kw, table_keys, ws.n_dev, num_pes, seed, ws.bin_offsets, st);
}

void launch_l1_keys(int kw, const JobDesc &one, uint32_t ntiles, hipStream_t st) {
  if (!ntiles) return;
  if (kw == 1) k_l1_keys<uint64_t, MSD_NT, MSD_ITEMS_1><<<ntiles, MSD_NT, 0, st>>>(one);
  else k_l1_keys<K2, MSD_NT, MSD_ITEMS_2><<<ntiles, MSD_NT, 0, st>>>(one);
}

void launch_hist_keys(int kw, const void *keys, uint64_t n, const HistSpecs &hs, uint64_t seed, uint64_t *hist,
                      int grid, hipStream_t st) {
  if (!n || grid <= 0) return;
  if (kw == 1) k_hist_keys<uint64_t, 512><<<grid, 512, 0, st>>>((const uint64_t *)keys, n, hs, seed, hist);
  else k_hist_keys<K2, 512><<<grid, 512, 0, st>>>((const K2 *)keys, n, hs, seed, hist);
}

// compute units of the stream's device (one persistent L1 workgroup each)
static uint32_t device_cus(hipStream_t st) {
  static uint32_t cache[64] = {};
  int dev = 0;
  if (hipStreamGetDevice(st, &dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cache[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    cache[dev] = (uint32_t)n;
  }
  return cache[dev];
}

template <typename KeyT, int ITEMS, int NT3, int NB3, int HT, bool ONE>
static void msd_levels(const BatchArgs &a, const JobDesc &one, uint32_t n1, uint32_t n2, uint32_t n3,
                       hipStream_t st) {
  if (n1) k_l1<KeyT, MSD_NT, ITEMS, ONE><<<std::min(n1, device_cus(st)), MSD_NT, 0, st>>>(a, one);
  if (n2) k_l2<KeyT, L2_NT, ITEMS, ONE><<<std::min(n2, device_cus(st)), L2_NT, 0, st>>>(a, one);
  constexpr int FI = sizeof(KeyT) == 8 ? FIN_ITEMS_1 : FIN_ITEMS_2;
  if (n3) {
    // one job: persistent workgroups (as many as fit beside each other, by
    // thread count) over the tickets; several jobs: one workgroup per item
    // (two-word keys run one item per workgroup: persistent, the k=55 genome
    // count's final level measured 113 ms against 51 ms)
    BatchArgs af = a;
    af.nfin = n3;
    af.persist = a.persist && sizeof(KeyT) == 8 ? 1u : 0u;
    const uint32_t grid = ONE && af.persist ? std::min(n3, device_cus(st) * (uint32_t)(1024 / NT3)) : n3;
    k_fin<KeyT, NT3, FI, NB3, HT, ONE><<<grid, NT3, 0, st>>>(af, one);
  }
}

template <typename KeyT, int ITEMS, bool ONE>
static void msd_levels_nt(uint32_t fin_nt, const BatchArgs &a, const JobDesc &one, uint32_t n1,
                          uint32_t n2, uint32_t n3, hipStream_t st) {
  constexpr bool W1 = sizeof(KeyT) == 8;
  // hash slots (12 bytes each for one-word keys, 22 for two-word keys) in the
  // final level's key buffer; LDS bins
  // of a final workgroup as plan_choose_path assumes (fin_bins)
  if (fin_nt == 1024) msd_levels<KeyT, ITEMS, 1024, 2048, W1 ? 10240 : 5120, ONE>(a, one, n1, n2, n3, st);
  else if (fin_nt == 512) msd_levels<KeyT, ITEMS, 512, 1024, W1 ? 5120 : 2560, ONE>(a, one, n1, n2, n3, st);
  else msd_levels<KeyT, ITEMS, 256, 512, W1 ? 2560 : 1280, ONE>(a, one, n1, n2, n3, st);
}

void launch_msd_batch(int kw, uint32_t fin_nt, const BatchArgs &a, const JobDesc *one, uint32_t n1,
                      uint32_t n2, uint32_t n3, hipStream_t st) {
  JobDesc none;
  std::memset(&none, 0, sizeof(none));
  if (kw == 1) {
    if (one) msd_levels_nt<uint64_t, MSD_ITEMS_1, true>(fin_nt, a, *one, n1, n2, n3, st);
    else msd_levels_nt<uint64_t, MSD_ITEMS_1, false>(fin_nt, a, none, n1, n2, n3, st);
  } else {
    if (one) msd_levels_nt<K2, MSD_ITEMS_2, true>(fin_nt, a, *one, n1, n2, n3, st);
    else msd_levels_nt<K2, MSD_ITEMS_2, false>(fin_nt, a, none, n1, n2, n3, st);
  }
}

}  // namespace dk
