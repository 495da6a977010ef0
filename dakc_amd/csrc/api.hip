// api.hip -- the C ABI of libdakc_amd (include/dakc.h): contexts, job
// planning, device workspaces and the launch sequence of one count.
//
// Replaces the reference's kmercounter constructor + perform_kcount
// (kcounter.hpp:84-99, kcounter.cpp:247-414): the record buffer goes in,
// every stage runs on the GPU, the sorted (kmer, count) table comes out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/dakc.h"
#include "dk_api_internal.hpp"
#include "dk_gen.hpp"
#include "dk_kernels.hpp"

using namespace dk;

// ------------------------------------------------------------------ status --

extern "C" const char *dakc_status_str(int s) {
  switch (s) {
  case DAKC_OK: return "ok";
  case DAKC_E_INVAL: return "invalid argument";
  case DAKC_E_FORMAT: return "malformed read buffer (records must be read_len bytes + '\\n')";
  case DAKC_E_NOMEM: return "out of memory";
  case DAKC_E_TOOBIG: return "job has 2^32 or more k-mer slots";
  case DAKC_E_DEVICE: return "HIP runtime error";
  case DAKC_E_NODEVICE: return "no usable GPU";
  case DAKC_E_REJECTED: return "job exceeds the served-mode HBM budget";
  case DAKC_E_NOTFOUND: return "unknown ticket";
  case DAKC_E_PENDING: return "job still running";
  case DAKC_E_CANCELLED: return "job cancelled";
  default: return "unknown status";
  }
}

extern "C" int dakc_abi_version(void) { return DAKC_ABI_VERSION; }

namespace dk {

// ------------------------------------------------------------ workspaces --

DevArena::~DevArena() {
  if (base) {
    hipSetDevice(device);
    hipFree(base);
  }
}

int DevArena::init(int dev, size_t bytes) {
  device = dev;
  size = bytes & ~(size_t)255;
  if (hipMalloc(&base, size) != hipSuccess) {
    (void)hipGetLastError();
    base = nullptr;
    size = 0;
    return DAKC_E_NOMEM;
  }
  free_ranges[0] = size;
  return DAKC_OK;
}

void *DevArena::alloc(size_t bytes) {
  bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> g(mu);
  for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it) {
    if (it->second < bytes) continue;
    const size_t off = it->first, len = it->second;
    free_ranges.erase(it);
    if (len > bytes) free_ranges[off + bytes] = len - bytes;
    used += bytes;
    peak = std::max(peak, used);
    return (char *)base + off;
  }
  return nullptr;
}

void *DevArena::alloc_top(size_t bytes) {
  bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> g(mu);
  for (auto it = free_ranges.rbegin(); it != free_ranges.rend(); ++it) {
    if (it->second < bytes) continue;
    const size_t off = it->first, len = it->second;
    if (len > bytes)
      it->second = len - bytes;  // the range keeps its bottom part
    else
      free_ranges.erase(off);
    used += bytes;
    peak = std::max(peak, used);
    return (char *)base + off + (len - bytes);
  }
  return nullptr;
}

size_t DevArena::largest_free() {
  std::lock_guard<std::mutex> g(mu);
  size_t m = 0;
  for (const auto &r : free_ranges) m = std::max(m, r.second);
  return m;
}

void DevArena::release(void *p, size_t bytes) {
  if (!p) return;
  bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> g(mu);
  size_t off = (size_t)((char *)p - (char *)base), len = bytes;
  used -= bytes;
  auto next = free_ranges.lower_bound(off);
  if (next != free_ranges.end() && off + len == next->first) {
    len += next->second;
    next = free_ranges.erase(next);
  }
  if (next != free_ranges.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      prev->second += len;
      return;
    }
  }
  free_ranges[off] = len;
}

int DevBuf::ensure(size_t bytes, hipStream_t st, bool zero) {
  if (bytes <= cap) return DAKC_OK;
  if (p) {
    hipStreamSynchronize(st);
    release();
  }
  size_t want = (bytes + 4095) & ~(size_t)4095;
  if (arena) {
    p = arena->alloc(want);
    if (!p) return DAKC_E_NOMEM;
  } else if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    return DAKC_E_NOMEM;
  }
  cap = want;
  if (zero && hipMemsetAsync(p, 0, want, st) != hipSuccess) return DAKC_E_DEVICE;
  return DAKC_OK;
}

void DevBuf::release() {
  if (p) {
    if (arena)
      arena->release(p, cap);
    else
      hipFree(p);
  }
  p = nullptr;
  cap = 0;
}

JobWs::~JobWs() {
  for (auto *b : {&hist, &dstart, &ndev, &err, &tickets, &flags, &rle_flags, &keys_a, &keys_b,
                  &counts, &binoff, &reads, &msd_off1, &msd_off2, &arena, &meta, &desc})
    b->release();
  for (auto &e : ev)
    if (e) hipEventDestroy(e);
  if (stage_ev) {
    hipEventSynchronize(stage_ev);
    hipEventDestroy(stage_ev);
  }
  if (hmeta) hipHostFree(hmeta);
  if (hstage) hipHostFree(hstage);
}

int JobWs::init(hipStream_t st, DevArena *pool) {
  stream = st;
  // the per-job buffers come from the pool when there is one; the small
  // per-slot buffers stay on hipMalloc
  for (auto *b : {&flags, &rle_flags, &keys_a, &keys_b, &counts, &reads, &msd_off2, &arena}) b->arena = pool;
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) return DAKC_E_DEVICE;
  if (hipEventCreateWithFlags(&stage_ev, hipEventDisableTiming) != hipSuccess) return DAKC_E_DEVICE;
  if (hipHostMalloc((void **)&hmeta, 256, hipHostMallocDefault) != hipSuccess) {
    hmeta = nullptr;
    return DAKC_E_NOMEM;
  }
  std::memset(hmeta, 0, 256);
  return DAKC_OK;
}

void JobWs::release_job_buffers() {
  for (auto *b : {&flags, &rle_flags, &keys_a, &keys_b, &counts, &reads, &msd_off2, &arena}) b->release();
}

void JobWs::detach_job_buffers(std::vector<std::pair<void *, size_t>> &out) {
  for (auto *b : {&flags, &rle_flags, &keys_a, &keys_b, &counts, &reads, &msd_off2, &arena}) {
    if (b->p) out.emplace_back(b->p, b->cap);
    b->p = nullptr;
    b->cap = 0;
  }
}

uint32_t JobWs::next_epoch() {
  if (epoch + 1 >= (1u << 14)) {
    launch_fill_u64((uint64_t *)flags.p, 0, flags.cap / 8, stream);
    launch_fill_u64((uint64_t *)rle_flags.p, 0, rle_flags.cap / 8, stream);
    epoch = 0;
  }
  return ++epoch;
}

Workspace JobWs::view() const {
  Workspace w;
  w.hist = (uint64_t *)hist.p;
  w.digit_start = (uint64_t *)dstart.p;
  w.n_dev = (uint64_t *)ndev.p;
  w.err = (uint32_t *)err.p;
  w.tickets = (uint32_t *)tickets.p;
  w.flags = (uint64_t *)flags.p;
  w.rle_flags = (uint64_t *)rle_flags.p;
  w.keys_a = keys_a.p;
  w.keys_b = keys_b.p;
  w.counts = (uint32_t *)counts.p;
  w.bin_offsets = (uint64_t *)binoff.p;
  w.key_cap = 0;
  return w;
}

size_t msd_job_bytes(const Plan &pl, bool copy_reads);

size_t job_device_bytes(const Plan &pl) {
  if (pl.msd) return msd_job_bytes(pl, true) + 8192;
  const size_t kb = (size_t)pl.kw * 8;
  size_t b = 0;
  b += 2 * (size_t)MAX_PASSES * 256 * 8 + 64 + 4096;
  b += (size_t)std::max<uint64_t>(pl.tiles0, pl.tiles_keys) * 256 * 8 + (size_t)pl.tiles_keys * 8;
  b += 2 * (size_t)pl.nmax * kb + (size_t)pl.nmax * 4;
  b += (size_t)pl.nbytes + 256;
  return b;
}

int JobWs::reserve(const Plan &pl, bool need_reads) {
  int s;
  const size_t kb = (size_t)pl.kw * 8;
  if (pl.msd) return DAKC_OK;  // msd_stage sizes the arena for the job
  if ((s = hist.ensure((size_t)MAX_PASSES * 256 * 8, stream, false))) return s;
  if ((s = dstart.ensure((size_t)MAX_PASSES * 256 * 8, stream, false))) return s;
  if ((s = ndev.ensure(64, stream, false))) return s;
  if ((s = err.ensure(64, stream, false))) return s;
  if ((s = tickets.ensure((MAX_PASSES + 2) * 4, stream, false))) return s;
  size_t ft = std::max<uint64_t>(pl.tiles0, pl.tiles_keys);
  const size_t fl_before = flags.cap, rf_before = rle_flags.cap;
  if ((s = flags.ensure(std::max<size_t>(ft, 1) * 256 * 8, stream, true))) return s;
  if ((s = rle_flags.ensure(std::max<size_t>(pl.tiles_keys, 1) * 8, stream, true))) return s;
  if (flags.cap != fl_before || rf_before != rle_flags.cap) {
    // freshly zeroed arrays: restart epochs for both so no stale word matches
    launch_fill_u64((uint64_t *)flags.p, 0, flags.cap / 8, stream);
    launch_fill_u64((uint64_t *)rle_flags.p, 0, rle_flags.cap / 8, stream);
    epoch = 0;
  }
  if ((s = keys_a.ensure(std::max<size_t>(pl.nmax, 1) * kb + 256, stream, false))) return s;
  if ((s = keys_b.ensure(std::max<size_t>(pl.nmax, 1) * kb + 256, stream, false))) return s;
  if ((s = counts.ensure(std::max<size_t>(pl.nmax, 1) * 4 + 256, stream, false))) return s;
  if ((s = binoff.ensure(((size_t)pl.num_pes + 1) * 8 + 64, stream, false))) return s;
  if (need_reads && (s = reads.ensure(pl.nbytes + 256, stream, false))) return s;
  return DAKC_OK;
}

// ------------------------------------------------------------- planning --

int make_plan(const dakc_job *job, Plan *pl) {
  if (!job || !pl) return DAKC_E_INVAL;
  if (job->format != DAKC_FMT_FIXED) return DAKC_E_INVAL;
  if (job->k < 1 || job->k > 64) return DAKC_E_INVAL;
  if (job->read_len < 1 || job->read_len > MAX_READ_LEN) return DAKC_E_INVAL;
  if (job->num_pes < 1) return DAKC_E_INVAL;
  if (job->owner_bin >= 0 && (uint32_t)job->owner_bin >= job->num_pes) return DAKC_E_INVAL;
  if (job->owner_bin < -1) return DAKC_E_INVAL;
  if (job->num_pes > 256 && job->owner_bin == -1) return DAKC_E_INVAL;  // one owner digit
  if (job->nbytes && !job->reads) return DAKC_E_INVAL;
  if (job->reads_on_device && ((uintptr_t)job->reads & 3)) return DAKC_E_INVAL;
  Plan &p = *pl;
  p = Plan();
  p.kw = job->k <= 32 ? 1 : 2;
  p.k = job->k;
  p.L = (uint32_t)job->read_len;
  p.stride = p.L + 1;
  p.nbytes = job->nbytes;
  p.num_pes = job->num_pes;
  p.owner_bin = job->num_pes > 1 ? job->owner_bin : -1;
  p.seed = job->hash_seed ? job->hash_seed : OWNER_SEED;
  const uint64_t rem = p.nbytes % p.stride;
  p.nrec = p.nbytes / p.stride;
  p.last_has_nl = 1;
  if (rem == p.L) {
    p.nrec += 1;
    p.last_has_nl = 0;
  } else if (rem != 0) {
    return DAKC_E_FORMAT;
  }
  p.S = p.L >= (uint32_t)p.k ? p.L - p.k + 1 : 0;
  const unsigned __int128 nm = (unsigned __int128)p.nrec * p.S;
  if (nm >= ((unsigned __int128)1 << 32)) return DAKC_E_TOOBIG;
  p.nmax = (uint64_t)nm;
  const int tile = tile_keys(p.kw);
  p.tile = tile;
  const uint32_t wpr = (p.L + 31) / 32 + 2;
  uint32_t rt = p.S ? tile / p.S : tile;
  rt = std::min<uint32_t>(rt, 128);
  rt = std::min<uint32_t>(rt, RAW_CAP / p.stride);
  rt = std::min<uint32_t>(rt, 896 / wpr);
  rt = std::max<uint32_t>(rt, 1);
  p.RT = rt;
  p.wpr = wpr;
  p.tiles0 = (uint32_t)((p.nrec + rt - 1) / rt);
  p.tiles_keys = (uint32_t)((p.nmax + tile - 1) / tile);
  p.ndig = (2 * p.k + 7) / 8;
  p.owner_digit = (p.num_pes > 1 && p.owner_bin < 0) ? 1 : 0;
  p.npasses = p.ndig + p.owner_digit;
  p.canonical = job->canonical ? 1 : 0;
  return DAKC_OK;
}

int make_plan_keys(uint64_t n, int k, uint32_t num_pes, uint64_t seed, Plan *pl) {
  if (!pl || k < 1 || k > 64 || num_pes < 1 || num_pes > 256) return DAKC_E_INVAL;
  if (n >= ((uint64_t)1 << 32)) return DAKC_E_TOOBIG;
  Plan &p = *pl;
  p = Plan();
  p.from_keys = 1;
  p.kw = k <= 32 ? 1 : 2;
  p.k = k;
  p.nmax = n;
  p.nrec = n;  // the key count, where a record plan has its records
  p.nbytes = n * (uint64_t)p.kw * 8;
  p.num_pes = num_pes;
  p.owner_bin = -1;
  p.seed = seed ? seed : OWNER_SEED;
  p.tile = tile_keys(p.kw);
  p.tiles0 = p.tiles_keys = (uint32_t)((n + p.tile - 1) / p.tile);
  p.ndig = (2 * p.k + 7) / 8;
  p.owner_digit = num_pes > 1 ? 1 : 0;
  p.npasses = p.ndig + p.owner_digit;
  return DAKC_OK;
}

// MSD geometry.  A final item should hold at most tune_fill_pct percent of
// what one final workgroup sorts in LDS on average (by default 90 for
// one-word keys and 75 for two-word keys: the k = 55 genome count's final
// level took 40.4 ms at 90, where sub-buckets over LDS capacity are redone
// digit by digit and hold up the look-back of every later item, and 27.2 ms
// at 75; the k = 31 count's took 17.3 ms at 90 and 20.2 ms at 75).  The
// coarsest split that meets it fixes b1 (<= 10: 1024 first-level buckets,
// fewer key bits when owner bins share them); the second level then splits up
// to two bits finer (b2 <= 10) and a final item takes `pack` consecutive
// second-level digits, as many as the target holds, so items come close to
// the target instead of to the nearest power of two below it.  L1 tiles take
// as many records as fill 15360 (8192) slots; an L2 group gathers G
// consecutive L1 tiles of a bucket, G sized so a group fills 75% of an L2
// workgroup on average.  b3 splits each digit of an item into LDS bins (pack
// << b3 bins in all).
int plan_choose_path(Plan *pl, int algo) {
  Plan &p = *pl;
  p.msd = 0;
  if (algo == DAKC_ALGO_LSD || p.nmax == 0) return 0;
  const uint32_t bits = 2 * (uint32_t)p.k;
  const uint32_t fin_nt = p.tune_fin_nt == 256 ? 256u : p.tune_fin_nt == 1024 ? 1024u : 512u;
  const uint64_t cap3 = (uint64_t)fin_cap(p.kw, (int)fin_nt), cap1 = (uint64_t)msd_tile(p.kw),
                 cap2 = (uint64_t)l2_cap(p.kw);
  // at most the tuned fill, and small enough for a final item per CU (a
  // small job in few large items would leave most of the GPU idle)
  // keys expected to reach the sort: with one owner bin of P counted, the
  // owner filter keeps about 1/P of the slots (MurmurHash64A spreads them)
  const uint64_t nexp = p.owner_bin >= 0 && p.num_pes > 1 ? (p.nmax + p.num_pes - 1) / p.num_pes : p.nmax;
  const uint64_t target = std::min<uint64_t>(
      cap3 * (uint64_t)(p.tune_fill_pct ? p.tune_fill_pct : p.kw == 2 ? 75 : 90) / 100,
      std::max<uint64_t>(cap3 / 8, nexp / std::max(p.num_cus, 1u)));
  // With all P owner bins counted, the owner is the top of the first-level
  // digit: MurmurHash64A spreads the k-mers evenly over the owners, so each
  // owner block needs key bits only for its own 1/P of them.
  const uint32_t owners = p.owner_digit ? p.num_pes : 1;
  if (owners > 1024) return 0;
  const uint64_t nown = (nexp + owners - 1) / owners;
  uint32_t b12 = 0;
  while ((nown >> b12) > target) ++b12;
  if (b12 > bits) return 0;  // more splitting than the key has bits: few distinct keys, LSD
  uint32_t b1max = 0;
  while ((owners << (b1max + 1)) <= 1024 && b1max < 10) ++b1max;
  uint32_t b1 = std::min(std::min(b1max, (b12 + 1) / 2), bits);
  if (b12 > b1 + 10) b1 = std::min(std::min(b1max, b12 - 10), bits);
  // the tuned first-level width: one bit more makes the final items gather
  // from half as many second-level groups (longer segments), one bit less
  // the reverse; the second level keeps the rest
  if (p.tune_b1_shift) {
    const int64_t t = (int64_t)b1 + p.tune_b1_shift;
    b1 = (uint32_t)std::max<int64_t>(0, std::min<int64_t>(t, std::min(b1max, std::min(bits, b12))));
  }
  const uint32_t b2 = std::min(std::min<uint32_t>(10, b12 > b1 ? b12 - b1 : 0), bits - b1);
  if (b12 > b1 + b2) return 0;  // sub-buckets would be too large on average
  // L1 tile geometry: a record's slots take ceil(S / items) threads (a key
  // array: tiles of cap1 keys)
  const uint32_t wpr = p.from_keys ? 0 : (p.L + 31) / 32 + 2;
  const uint32_t l1_items = p.kw == 1 ? MSD_ITEMS_1 : MSD_ITEMS_2;
  uint32_t rt = 1;
  if (!p.from_keys) {
    rt = p.S ? (uint32_t)MSD_NT / ((p.S + l1_items - 1) / l1_items) : (uint32_t)cap1;
    rt = std::min<uint32_t>(rt, L1_MAXRT);
    rt = std::min<uint32_t>(rt, L1_RAW / p.stride);
    rt = std::min<uint32_t>(rt, L1_CODEW / wpr);
    rt = std::max<uint32_t>(rt, 1);
    if ((uint64_t)rt * p.S > cap1) return 0;
  }
  const uint32_t b2f = std::min(std::min(10u, bits - b1), b2 + 2);
  // Final items are sized to fit where keys are densest: an item over LDS
  // capacity is redone digit by digit, and its later digits' look-back slots
  // stay unpublished meanwhile, so every later item waits on it.  Canonical
  // k-mers (the smaller of a k-mer and its reverse complement) fall linearly
  // from twice the mean at the bottom of the key range to none at the top, so
  // each quarter of an owner's top-bit range gets items sized for its own
  // densest point (2, 1.5, 1, 0.5 x the mean).
  const double per_digit = (double)nown / (double)(1ull << (b1 + b2f));
  const uint32_t lc = std::min<uint32_t>(2, b1), nd2 = 1u << b2f;
  uint32_t packmax = 1;
  for (uint32_t c = 0; c < 4; ++c) {
    p.packc[c] = p.ipbc[c] = 0;
    if (c >= (1u << lc)) continue;
    const double dens = p.canonical ? 2.0 * (1.0 - (double)c / (double)(1u << lc)) : 1.0;
    uint32_t pk = (uint32_t)std::max(1.0, std::floor((double)target / std::max(per_digit * dens, 1e-9)));
    pk = std::min<uint32_t>(std::min<uint32_t>(pk, 16), nd2);
    p.packc[c] = pk;
    p.ipbc[c] = (nd2 + pk - 1) / pk;
    packmax = std::max(packmax, pk);
  }
  const uint32_t nb3 = fin_nt >= 1024 ? 2048 : fin_nt >= 512 ? 1024 : 512;  // LDS bins of a final workgroup
  uint32_t b3 = 0;
  while (b3 < 11 && b3 < bits - b1 - b2f && (packmax << (b3 + 1)) <= nb3) ++b3;
  p.b1 = b1;
  p.b2 = b2f;
  p.b3 = b3;
  p.pack = packmax;
  p.cls_log = lc;
  p.fin_nt = fin_nt;
  p.nb1 = owners << b1;
  p.nd2 = nd2;
  uint32_t per_block = 0;  // items of an owner block (2^b1 buckets)
  for (uint32_t c = 0; c < 4; ++c) per_block += (1u << (b1 - lc)) * p.ipbc[c];
  p.nsb = owners * per_block;
  p.RT1 = rt;
  p.wpr1 = wpr;
  if (p.from_keys) {
    p.ntiles1 = (uint32_t)((p.nmax + cap1 - 1) / cap1);
    p.tcap1 = (uint32_t)cap1;
  } else {
    p.ntiles1 = (uint32_t)((p.nrec + rt - 1) / rt);
    p.tcap1 = rt * p.S;
  }
  // canonical k-mers (min of a k-mer and its reverse complement) are twice as
  // dense as the mean at the low end of the key range: size groups for that
  const double per_tile_bucket = (double)nexp / p.ntiles1 / p.nb1 * (p.canonical ? 2.0 : 1.0);
  uint32_t G = (uint32_t)std::max(1.0, std::floor(0.75 * (double)cap2 / std::max(per_tile_bucket, 1e-9)));
  G = std::min<uint32_t>(G, L2_NT);
  G = std::min<uint32_t>(G, p.ntiles1);
  p.G = std::max<uint32_t>(G, 1);
  p.ngroups = (p.ntiles1 + p.G - 1) / p.G;
  if (p.ngroups > (uint32_t)FIN_MAX_GROUPS) return 0;
  p.cap2 = (uint32_t)cap2;
  p.msd = 1;
  return 1;
}

ReadsDesc plan_reads(const Plan &p, const uint8_t *dreads) {
  ReadsDesc rd;
  rd.reads = dreads;
  rd.nbytes = p.nbytes;
  rd.nrec = p.nrec;
  rd.L = p.L;
  rd.stride = p.stride;
  rd.k = (uint32_t)p.k;
  rd.S = p.S;
  rd.RT = p.RT;
  rd.wpr = p.wpr;
  rd.invS = p.S ? 1.0f / (float)p.S : 0.0f;
  rd.canonical = p.canonical;
  rd.num_pes = p.num_pes;
  rd.owner_bin = p.owner_bin;
  rd.last_has_nl = p.last_has_nl;
  rd.seed = p.seed;
  return rd;
}

// Enqueue the whole count of one job on ws.stream; ws.loc says where the
// table and the status words will be.
// ------------------------------------------------------ multi-job MSD --

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
// A job of a multi-job launch with this many final-level items (default)
// fills the GPU a few hundred workgroups deep on its own: it gets a
// final-level launch of its own, which saves every workgroup the lookup of
// its job.
constexpr uint32_t FIN_SOLO_ITEMS = 32768;
uint32_t solo_items(const dakc_ctx *c) {
  return c->opts.batch_solo_items > 0 ? (uint32_t)c->opts.batch_solo_items : FIN_SOLO_ITEMS;
}
constexpr size_t META_STRIDE = 32;  // per job: u64 n, u64 d, u32 err

// Arena layout: all jobs' look-back flags first (one fill clears them), then
// per job its reads copy (host input only), L1 tiles / table keys, L2 regions,
// counts, offsets and owner-bin offsets.
void msd_layout(MsdBatch &b, const std::vector<Plan> &plans, const std::vector<bool> &copy_reads,
                uint32_t solo) {
  const size_t nj = plans.size();
  b.plans = plans;
  b.kw = nj ? plans[0].kw : 1;
  for (auto *v : {&b.off_reads, &b.off_out1, &b.off_out2, &b.off_counts, &b.off_off1, &b.off_off2,
                  &b.off_flags, &b.off_binoff})
    v->assign(nj, 0);
  b.pre1.assign(nj + 1, 0);
  b.pre2.assign(nj + 1, 0);
  b.pre3.assign(nj + 1, 0);
  size_t o = 0;
  b.flags_off = 0;
  for (size_t j = 0; j < nj; ++j) {
    b.off_flags[j] = o;
    o += (size_t)plans[j].nb1 * plans[j].nd2 * 8;  // final look-back: one word per second-level digit
  }
  b.flags_words = o / 8;
  o = al256(o);
  for (size_t j = 0; j < nj; ++j) {
    const Plan &p = plans[j];
    const size_t kb = (size_t)p.kw * 8;
    if (copy_reads[j]) {
      b.off_reads[j] = o;
      o = al256(o + p.nbytes + 4);
    } else {
      b.off_reads[j] = (size_t)-1;
    }
    b.off_out1[j] = o;
    o = al256(o + std::max<size_t>((size_t)p.ntiles1 * p.tcap1, p.nmax) * kb);
    b.off_out2[j] = o;
    o = al256(o + (size_t)p.nb1 * p.ngroups * p.cap2 * kb);
    b.off_counts[j] = o;
    o = al256(o + std::max<size_t>(p.nmax, 1) * 4);
    b.off_off1[j] = o;
    o = al256(o + ((size_t)p.nb1 + 1) * p.ntiles1 * 2);
    b.off_off2[j] = o;
    o = al256(o + (size_t)p.nb1 * (p.nd2 + 1) * p.ngroups * 2);
    b.off_binoff[j] = o;
    o = al256(o + ((size_t)p.num_pes + 1) * 8);
    b.pre1[j + 1] = b.pre1[j] + p.ntiles1;
    b.pre2[j + 1] = b.pre2[j] + p.nb1 * p.ngroups;
    b.pre3[j + 1] = b.pre3[j] + p.nsb;
  }
  b.fin_runs.clear();
  for (size_t j = 0; j < nj; ++j) {
    const uint32_t fin = plans[j].fin_nt;
    if (nj > 1 && plans[j].nsb >= solo) {
      b.fin_runs.push_back({fin, b.pre3[j], b.pre3[j + 1], (int32_t)j});
      continue;
    }
    if (b.fin_runs.empty() || b.fin_runs.back().solo >= 0 || b.fin_runs.back().fin_nt != fin ||
        b.fin_runs.back().item1 != b.pre3[j])
      b.fin_runs.push_back({fin, b.pre3[j], b.pre3[j], -1});
    b.fin_runs.back().item1 = b.pre3[j + 1];
  }
  b.arena_bytes = o;
  b.meta_bytes = al256(nj * META_STRIDE);
  // descriptors, item prefixes, one ticket per final-level run
  b.desc_bytes = al256(nj * sizeof(JobDesc)) + al256(3 * (nj + 1) * 4) + al256(b.fin_runs.size() * 4 + 4);
}

size_t msd_job_bytes(const Plan &pl, bool copy_reads) {
  MsdBatch b;
  msd_layout(b, {pl}, {copy_reads});
  return b.arena_bytes + b.meta_bytes + b.desc_bytes;
}

void msd_describe(const MsdBatch &b, size_t j, const uint8_t *dreads, uint8_t *ar, uint8_t *mt, JobDesc &D,
                  TableLoc &L) {
  const Plan &p = b.plans[j];
  std::memset(&D, 0, sizeof(D));
  D.rd = plan_reads(p, dreads);
  D.rd.RT = p.RT1;
  D.rd.wpr = p.wpr1;
  const uint32_t bits = 2 * (uint32_t)p.k;
  MsdArgs &m = D.m;
  m.d1 = {DIG_MSD, bits - p.b1, p.b1, p.owner_digit ? p.num_pes : 1u};
  m.d2 = {DIG_BITS, bits - p.b1 - p.b2, p.b2, 1};
  m.d3 = {DIG_BITS, bits - p.b1 - p.b2 - p.b3, p.b3, 1};
  m.nb1 = p.nb1;
  m.nd2 = p.nd2;
  m.ntiles1 = p.ntiles1;
  m.tcap1 = p.tcap1;
  m.G = p.G;
  m.ngroups = p.ngroups;
  m.cap2 = p.cap2;
  m.fin_nt = p.fin_nt;
  m.cls = p.cls_log | (p.b1 - p.cls_log) << 4;
  m.packc = 0;
  m.ipbc[0] = m.ipbc[1] = 0;
  for (int c = 0; c < 4; ++c) {
    m.packc |= p.packc[c] << (8 * c);
    m.ipbc[c >> 1] |= p.ipbc[c] << (16 * (c & 1));
  }
  m.off1 = (uint16_t *)(ar + b.off_off1[j]);
  m.off2 = (uint16_t *)(ar + b.off_off2[j]);
  D.out1 = ar + b.off_out1[j];
  D.out2 = ar + b.off_out2[j];
  D.tcounts = (uint32_t *)(ar + b.off_counts[j]);
  D.tkeys = D.out1;
  if (j < b.dest_keys.size() && b.dest_keys[j] && b.dest_counts[j]) {  // a table of its own
    D.tkeys = b.dest_keys[j];
    D.tcounts = b.dest_counts[j];
  }
  D.nd = (uint64_t *)(mt + j * META_STRIDE);
  D.err = (uint32_t *)(mt + j * META_STRIDE + 16);
  D.flags = (uint64_t *)(ar + b.off_flags[j]);
  L = TableLoc();
  L.keys = D.tkeys;
  L.counts = D.tcounts;
  L.binoff = p.owner_digit ? (uint64_t *)(ar + b.off_binoff[j]) : nullptr;
  L.nd = D.nd;
  L.err = D.err;
}

int msd_stage(JobWs &ws, MsdBatch &b, const std::vector<const dakc_job *> &jobs) {
  const size_t nj = b.plans.size();
  int s;
  if ((s = ws.arena.ensure(std::max<size_t>(b.arena_bytes, 256), ws.stream, false))) return s;
  if ((s = ws.meta.ensure(std::max<size_t>(b.meta_bytes, 256), ws.stream, false))) return s;
  if ((s = ws.desc.ensure(b.desc_bytes, ws.stream, false))) return s;
  uint8_t *ar = (uint8_t *)ws.arena.p;
  uint8_t *mt = (uint8_t *)ws.meta.p;
  uint8_t *dd = (uint8_t *)ws.desc.p;
  std::vector<JobDesc> desc(nj);
  b.locs.assign(nj, TableLoc());
  for (size_t j = 0; j < nj; ++j) {
    const Plan &p = b.plans[j];
    const uint8_t *dreads = (const uint8_t *)jobs[j]->reads;
    if (b.off_reads[j] != (size_t)-1) {
      if (p.nbytes && hipMemcpyAsync(ar + b.off_reads[j], jobs[j]->reads, p.nbytes,
                                     hipMemcpyHostToDevice, ws.stream) != hipSuccess)
        return DAKC_E_DEVICE;
      dreads = ar + b.off_reads[j];
    }
    msd_describe(b, j, dreads, ar, mt, desc[j], b.locs[j]);
  }
  // argument block: JobDesc[nj] | pre1 | pre2 | pre3 | ticket, staged in
  // pinned memory once the previous upload from it has finished
  hipEventSynchronize(ws.stage_ev);
  if (ws.hstage_cap < b.desc_bytes) {
    if (ws.hstage) hipHostFree(ws.hstage);
    ws.hstage_cap = 0;
    if (hipHostMalloc((void **)&ws.hstage, b.desc_bytes, hipHostMallocDefault) != hipSuccess) {
      ws.hstage = nullptr;
      return DAKC_E_NOMEM;
    }
    ws.hstage_cap = b.desc_bytes;
  }
  b.descs = desc;
  uint8_t *host = ws.hstage;
  std::memset(host, 0, b.desc_bytes);
  std::memcpy(host, desc.data(), nj * sizeof(JobDesc));
  uint32_t *pre = (uint32_t *)(host + al256(nj * sizeof(JobDesc)));
  for (size_t j = 0; j <= nj; ++j) {
    pre[j] = b.pre1[j];
    pre[nj + 1 + j] = b.pre2[j];
    pre[2 * (nj + 1) + j] = b.pre3[j];
  }
  if (hipMemcpyAsync(dd, host, b.desc_bytes, hipMemcpyHostToDevice, ws.stream) != hipSuccess ||
      hipEventRecord(ws.stage_ev, ws.stream) != hipSuccess)
    return DAKC_E_DEVICE;
  return DAKC_OK;
}

static BatchArgs batch_args(uint8_t *dd, const MsdBatch &b, bool persist) {
  const size_t nj = b.plans.size();
  uint32_t *pre = (uint32_t *)(dd + al256(nj * sizeof(JobDesc)));
  BatchArgs a;
  a.jobs = (const JobDesc *)dd;
  a.pre1 = pre;
  a.pre2 = pre + (nj + 1);
  a.pre3 = pre + 2 * (nj + 1);
  a.njobs = (uint32_t)nj;
  a.ticket = pre + 3 * (nj + 1);
  a.item0 = 0;
  a.persist = persist ? 1u : 0u;
  return a;
}

// Fill kernels and the three levels; no host transfer, so a stream capture of
// this sequence replays it exactly.  arena / meta / desc: the job's device
// buffers (desc holds the argument block; a one-job launch reads only its
// tickets from it); ev: 3 + MAX_PASSES stage events, or null.
void msd_enqueue_at(hipStream_t st, const MsdBatch &b, uint8_t *arena, uint8_t *meta, uint8_t *desc,
                    bool persist, hipEvent_t *ev) {
  const size_t nj = b.plans.size();
  const BatchArgs a = batch_args(desc, b, persist);
  launch_fill_u64((uint64_t *)meta, 0, b.meta_bytes / 8, st);
  launch_fill_u64((uint64_t *)arena + b.flags_off / 8, 0, b.flags_words, st);
  launch_fill_u32(a.ticket, 0, std::max<size_t>(b.fin_runs.size(), 1), st);
  if (ev) hipEventRecord(ev[0], st);
  const JobDesc *one = nj == 1 ? &b.descs[0] : nullptr;
  const uint32_t fin0 = b.fin_runs.empty() ? 256u : b.fin_runs[0].fin_nt;
  if (one && b.plans[0].from_keys) launch_l1_keys(b.kw, *one, b.pre1[1], st);
  else launch_msd_batch(b.kw, fin0, a, one, b.pre1[nj], 0, 0, st);
  if (ev) hipEventRecord(ev[1], st);
  launch_msd_batch(b.kw, fin0, a, one, 0, b.pre2[nj], 0, st);
  if (ev) hipEventRecord(ev[2], st);
  for (size_t r = 0; r < b.fin_runs.size(); ++r) {
    const MsdBatch::FinRun &fr = b.fin_runs[r];
    BatchArgs ar = a;
    ar.ticket = a.ticket + r;
    ar.item0 = fr.solo >= 0 ? 0u : fr.item0;  // a solo launch numbers its job's items from 0
    launch_msd_batch(b.kw, fr.fin_nt, ar, fr.solo >= 0 ? &b.descs[fr.solo] : one, 0, 0,
                     fr.item1 - fr.item0, st);
  }
  for (size_t j = 0; j < nj; ++j)
    if (b.plans[j].owner_digit)
      launch_bin_offsets_at(b.kw, b.locs[j].keys, b.locs[j].nd, b.plans[j].num_pes, b.plans[j].seed,
                            b.locs[j].binoff, st);
  if (ev) hipEventRecord(ev[2 + MAX_PASSES], st);
}

void msd_enqueue(JobWs &ws, const MsdBatch &b, bool events) {
  msd_enqueue_at(ws.stream, b, (uint8_t *)ws.arena.p, (uint8_t *)ws.meta.p, (uint8_t *)ws.desc.p,
                 ws.fin_persistent, events ? ws.ev : nullptr);
}

int msd_finish(const MsdBatch &b, size_t j, const void *meta_host, uint64_t *n, uint64_t *d) {
  const uint8_t *m = (const uint8_t *)meta_host + j * META_STRIDE;
  uint64_t nd[2];
  uint32_t e;
  std::memcpy(nd, m, 16);
  std::memcpy(&e, m + 16, 4);
  if (int s = status_of_err(e)) return s;
  *n = nd[0];
  *d = b.plans[j].nmax ? nd[1] : 0;
  return DAKC_OK;
}

int enqueue_count(JobWs &ws, const Plan &pl, const dakc_job *job, const uint8_t *dreads,
                  int num_cus, bool events) {
  if (pl.msd) {
    msd_layout(ws.single, {pl}, {!job->reads_on_device});
    int s = msd_stage(ws, ws.single, {job});
    if (s) return s;
    msd_enqueue(ws, ws.single, events);
    ws.loc = ws.single.locs[0];
    if (hipMemcpyAsync(ws.hmeta, ws.loc.nd, META_STRIDE, hipMemcpyDeviceToHost, ws.stream) != hipSuccess)
      return DAKC_E_DEVICE;
    return hipGetLastError() == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
  }
  hipStream_t st = ws.stream;
  Workspace w = ws.view();
  w.key_cap = pl.nmax;  // reserve() sized keys_a, keys_b and counts for nmax keys
  if (!job->reads_on_device && pl.nbytes) {
    if (hipMemcpyAsync(ws.reads.p, job->reads, pl.nbytes, hipMemcpyHostToDevice, st) != hipSuccess)
      return DAKC_E_DEVICE;
    dreads = (const uint8_t *)ws.reads.p;
  }
  lsd_enqueue(w, pl, dreads, st, num_cus, [&ws] { return ws.next_epoch(); }, events ? ws.ev : nullptr,
              ws.loc);
  if (hipMemcpyAsync(ws.hmeta, w.n_dev, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(ws.hmeta + 16, w.err, 4, hipMemcpyDeviceToHost, st) != hipSuccess)
    return DAKC_E_DEVICE;
  return hipGetLastError() == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
}

// The LSD path of one job on workspace w (keys_a / keys_b / counts for
// w.key_cap keys, flags for max(tiles0, tiles_keys) tiles): histograms, one
// sweep per digit, run-length reduce, owner-bin offsets.  next_epoch gives
// each look-back pass a fresh flag epoch; ev: stage events or null.
void lsd_enqueue(const Workspace &w, const Plan &pl, const uint8_t *dreads, hipStream_t st, int num_cus,
                 const std::function<uint32_t()> &next_epoch, hipEvent_t *ev, TableLoc &loc) {
  const ReadsDesc rd = plan_reads(pl, dreads);
  launch_fill_u64(w.hist, 0, (uint64_t)pl.npasses * 256, st);
  launch_fill_u64(w.n_dev, 0, 2, st);
  launch_fill_u32(w.err, 0, 1, st);
  launch_fill_u32(w.tickets, 0, MAX_PASSES + 2, st);
  if (ev) hipEventRecord(ev[0], st);
  auto pass_digit = [&](int p) -> DigitSpec {
    if (pl.owner_digit && p == pl.npasses - 1) return DigitSpec{DIG_OWNER, 0, 8, pl.num_pes};
    return DigitSpec{DIG_BITS, 8u * (uint32_t)p, 8, 1};
  };
  if (pl.nrec > 0) {
    HistSpecs hs;
    hs.n = pl.npasses;
    for (int p = 0; p < pl.npasses; ++p) hs.d[p] = pass_digit(p);
    const int grid = (int)std::min<uint64_t>(pl.tiles0, (uint64_t)num_cus * 4);
    if (pl.from_keys) launch_hist_keys(pl.kw, dreads, pl.nmax, hs, pl.seed, w.hist, grid, st);
    else launch_hist_reads(pl.kw, rd, hs, w, pl.tiles0, grid, st);
    launch_scan_digits(pl.npasses, w, st);
  }
  if (ev) hipEventRecord(ev[1], st);
  void *cur = w.keys_a, *other = w.keys_b;
  if (pl.nmax > 0) {
    for (int p = 0; p < pl.npasses; ++p) {
      const uint32_t ep = next_epoch();
      if (p == 0 && pl.from_keys) {  // the first pass reads the key array (left unchanged)
        launch_sweep(pl.kw, false, rd, dreads, cur, pass_digit(0), w.digit_start, w, 0, ep, pl.tiles_keys, st);
      } else if (p == 0) {
        launch_sweep(pl.kw, true, rd, nullptr, cur, pass_digit(0), w.digit_start, w, 0, ep,
                     pl.tiles0, st);
      } else {
        launch_sweep(pl.kw, false, rd, cur, other, pass_digit(p), w.digit_start + 256 * p, w, p, ep,
                     pl.tiles_keys, st);
        std::swap(cur, other);
      }
      if (ev) hipEventRecord(ev[2 + p], st);
    }
    const uint32_t ep = next_epoch();
    launch_rle(pl.kw, cur, other, w, ep, pl.tiles_keys, st);
    if (pl.owner_digit) launch_bin_offsets(pl.kw, other, pl.num_pes, pl.seed, w, st);
  }
  if (ev) hipEventRecord(ev[2 + MAX_PASSES], st);
  loc.keys = other;
  loc.counts = w.counts;
  loc.binoff = w.bin_offsets;
  loc.nd = w.n_dev;
  loc.err = w.err;
}

// After the stream has drained: n, d and the error word, from the pinned
// copy the enqueue made of them.
int finish_count(JobWs &ws, const Plan &pl, uint64_t *n, uint64_t *d) {
  if (pl.msd) return msd_finish(ws.single, 0, ws.hmeta, n, d);
  uint64_t nd[2] = {0, 0};
  uint32_t e = 0;
  std::memcpy(nd, ws.hmeta, 16);
  std::memcpy(&e, ws.hmeta + 16, 4);
  if (int s = status_of_err(e)) return s;
  if (pl.nmax == 0) nd[1] = 0;
  *n = nd[0];
  *d = nd[1];
  return DAKC_OK;
}

}  // namespace dk

// --------------------------------------------------------------- context --

extern "C" int dakc_ctx_create(int device, const dakc_opts *opts, dakc_ctx **out) {
  if (!out) return DAKC_E_INVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return DAKC_E_NODEVICE;
  }
  if (device < 0 || device >= ndev) return DAKC_E_INVAL;
  if (hipSetDevice(device) != hipSuccess) return DAKC_E_DEVICE;
  // once per device: the allocation-free count's array (its entry point
  // reports a failure here)
  noalloc_setup(device);
  dakc_ctx *c = new (std::nothrow) dakc_ctx();
  if (!c) return DAKC_E_NOMEM;
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return DAKC_E_DEVICE;
  }
  c->num_cus = prop.multiProcessorCount;
  if (opts) c->opts = *opts;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return DAKC_E_DEVICE;
  }
  int s = c->ws.init(c->stream);
  if (s) {
    dakc_ctx_destroy(c);
    return s;
  }
  *out = c;
  return DAKC_OK;
}

extern "C" int dakc_ctx_destroy(dakc_ctx *c) {
  if (!c) return DAKC_E_INVAL;
  hipSetDevice(c->device);
  serve_shutdown(c);
  if (c->stream) hipStreamSynchronize(c->stream);
  c->ws.~JobWs();
  new (&c->ws) JobWs();
  for (DevBuf *b : {&c->post_a, &c->post_b, &c->conv_in, &c->conv_nl, &c->conv_out}) b->release();
  c->hit.release();
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return DAKC_OK;
}

// ----------------------------------------------------------- plain count --

static void fill_stats(dakc_stats &st, JobWs &ws, const Plan &pl, uint64_t n, uint64_t d) {
  std::memset(&st, 0, sizeof(st));
  float ms = 0;
  st.npasses = pl.nmax ? (uint32_t)(pl.msd ? 1 : pl.npasses) : 0;
  st.tiles0 = pl.tiles0;
  st.n = n;
  st.d = d;
  st.path = pl.msd ? 2 : 1;
  hipEventElapsedTime(&ms, ws.ev[0], ws.ev[1]);
  st.hist_ms = ms;
  if (pl.nmax && pl.msd) {
    st.hist_ms = 0;
    hipEventElapsedTime(&ms, ws.ev[0], ws.ev[1]);
    st.pass0_ms = st.sweep_ms[0] = ms;
    hipEventElapsedTime(&ms, ws.ev[1], ws.ev[2]);
    st.part2_ms = ms;
    hipEventElapsedTime(&ms, ws.ev[2], ws.ev[2 + MAX_PASSES]);
    st.final_ms = ms;
  } else if (pl.nmax) {
    hipEventElapsedTime(&ms, ws.ev[1], ws.ev[2]);
    st.pass0_ms = ms;
    st.sweep_ms[0] = ms;
    for (int p = 1; p < pl.npasses; ++p) {
      hipEventElapsedTime(&ms, ws.ev[1 + p], ws.ev[2 + p]);
      st.sweep_ms[p] = ms;
      st.passes_ms += ms;
    }
    hipEventElapsedTime(&ms, ws.ev[1 + pl.npasses], ws.ev[2 + MAX_PASSES]);
    st.rle_ms = ms;
  }
  hipEventElapsedTime(&ms, ws.ev[0], ws.ev[2 + MAX_PASSES]);
  st.total_ms = ms;
}

// FASTQ or line input: convert it on the device into ctx->conv_out and
// describe that as a fixed-width job (host bytes are staged in ctx->conv_in).
static int convert_job(dakc_ctx *c, const dakc_job *job, dakc_job *fixed) {
  if (job->read_len < 0 || (job->nbytes && !job->reads)) return DAKC_E_INVAL;
  const uint8_t *in = (const uint8_t *)job->reads;
  int s;
  if (!job->reads_on_device && job->nbytes) {
    if ((s = c->conv_in.ensure(job->nbytes, c->stream, false))) return s;
    if (hipMemcpyAsync(c->conv_in.p, job->reads, job->nbytes, hipMemcpyHostToDevice, c->stream) != hipSuccess)
      return DAKC_E_DEVICE;
    in = (const uint8_t *)c->conv_in.p;
  }
  uint64_t nb = 0;
  uint32_t L = 0;
  if ((s = convert_reads(c, in, job->nbytes, job->format, (uint32_t)job->read_len, nullptr, 0, &c->conv_out,
                         &nb, &L)))
    return s;
  *fixed = *job;
  fixed->format = DAKC_FMT_FIXED;
  fixed->reads = (const char *)c->conv_out.p;
  fixed->nbytes = nb;
  fixed->read_len = (int32_t)L;
  fixed->reads_on_device = 1;
  return DAKC_OK;
}

static int count_common(dakc_ctx *c, const dakc_job *job, dakc_table *out, bool to_host) {
  if (!c || !job || !out) return DAKC_E_INVAL;
  std::memset(out, 0, sizeof(*out));
  hipSetDevice(c->device);
  dakc_job fixed;
  if (job->format == DAKC_FMT_FASTQ || job->format == DAKC_FMT_LINES) {
    int s = convert_job(c, job, &fixed);
    if (s) return s;
    job = &fixed;
  }
  Plan pl;
  int s = make_plan(job, &pl);
  if (s) return s;
  return count_planned(c, pl, job, out, to_host);
}

namespace dk {
int count_planned(dakc_ctx *c, Plan &pl, const dakc_job *job, dakc_table *out, bool to_host) {
  int s;
  apply_tuning(c, &pl, c->opts.algo);
  JobWs &ws = c->ws;
  uint64_t n = 0, d = 0;
  uint32_t reruns = 0;
  bool freed = false;  // the workspace was emptied once for this path
  for (;;) {
    s = ws.reserve(pl, !job->reads_on_device);
    if (!s) s = enqueue_count(ws, pl, job, (const uint8_t *)job->reads, c->num_cus, true);
    if (s == DAKC_E_NOMEM) {
      // out of device memory: give back what the workspace holds from earlier
      // counts and try again; then fall back from the MSD path to the LSD
      // path, whose workspace is smaller
      if (hipStreamSynchronize(ws.stream) != hipSuccess) return DAKC_E_DEVICE;
      ws.release_job_buffers();
      if (!freed) {
        freed = true;
        continue;
      }
      if (pl.msd) {
        plan_choose_path(&pl, DAKC_ALGO_LSD);
        freed = true;
        continue;
      }
    }
    if (s) return s;
    if (hipStreamSynchronize(ws.stream) != hipSuccess) return DAKC_E_DEVICE;
    s = finish_count(ws, pl, &n, &d);
    if (s == STATUS_MSD_OVERFLOW && pl.msd) {  // skewed input: count it on the LSD path
      plan_choose_path(&pl, DAKC_ALGO_LSD);
      ++reruns;
      continue;
    }
    if (s) return s;
    break;
  }
  const TableLoc loc = ws.loc;
  fill_stats(c->last, ws, pl, n, d);
  c->last.msd_reruns = reruns;

  out->n = d;
  out->total = n;
  out->key_words = pl.kw;
  out->k = pl.k;
  out->num_bins = pl.owner_digit ? pl.num_pes : 0;
  out->hash_seed = pl.seed;
  if (pl.owner_digit) {
    out->bin_offsets = (uint64_t *)std::malloc(((size_t)pl.num_pes + 1) * 8);
    if (!out->bin_offsets) return DAKC_E_NOMEM;
    if (d) {
      if (hipMemcpy(out->bin_offsets, loc.binoff, ((size_t)pl.num_pes + 1) * 8,
                    hipMemcpyDeviceToHost) != hipSuccess)
        return DAKC_E_DEVICE;
    } else {
      std::memset(out->bin_offsets, 0, ((size_t)pl.num_pes + 1) * 8);
    }
  }
  if (!to_host) {
    out->keys = (uint64_t *)loc.keys;
    out->counts = loc.counts;
    out->on_device = 1;
    return DAKC_OK;
  }
  const size_t kbytes = (size_t)d * pl.kw * 8;
  out->keys = (uint64_t *)std::malloc(kbytes ? kbytes : 8);
  out->counts = (uint32_t *)std::malloc(d ? (size_t)d * 4 : 4);
  if (!out->keys || !out->counts) {
    dakc_table_free(out);
    return DAKC_E_NOMEM;
  }
  if (d) {
    if (hipMemcpy(out->keys, loc.keys, kbytes, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(out->counts, loc.counts, (size_t)d * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      dakc_table_free(out);
      return DAKC_E_DEVICE;
    }
  }
  return DAKC_OK;
}
}  // namespace dk

extern "C" int dakc_plan_job(const dakc_job *job, int algo, uint32_t num_cus, dakc_plan_info *out) {
  if (!job || !out) return DAKC_E_INVAL;
  std::memset(out, 0, sizeof(*out));
  if (job->format != DAKC_FMT_FIXED) return DAKC_E_INVAL;
  Plan pl;
  if (int s = make_plan(job, &pl)) return s;
  pl.num_cus = num_cus ? num_cus : 256;
  plan_choose_path(&pl, algo);
  out->path = pl.msd ? 2 : 1;
  out->key_words = (uint32_t)pl.kw;
  out->kmer_slots = pl.nmax;
  out->device_bytes = job_device_bytes(pl);
  if (pl.msd) {
    out->b1 = pl.b1;
    out->b2 = pl.b2;
    out->b3 = pl.b3;
    out->first_level_buckets = pl.nb1;
    out->l1_tiles = pl.ntiles1;
    out->l2_groups = pl.ngroups;
    out->final_items = pl.nsb;
    out->final_threads = pl.fin_nt;
  } else {
    out->radix_passes = (uint32_t)pl.npasses;
  }
  return DAKC_OK;
}

extern "C" int dakc_count(dakc_ctx *ctx, const dakc_job *job, dakc_table *out) {
  return count_common(ctx, job, out, true);
}

extern "C" int dakc_count_device(dakc_ctx *ctx, const dakc_job *job, dakc_table *out) {
  return count_common(ctx, job, out, false);
}

extern "C" int dakc_last_stats(dakc_ctx *ctx, dakc_stats *out) {
  if (!ctx || !out) return DAKC_E_INVAL;
  *out = ctx->last;
  return DAKC_OK;
}

extern "C" void dakc_table_free(dakc_table *t) {
  if (!t) return;
  if (t->on_device && t->priv) {
    dev_table_release(t);
  } else if (!t->on_device) {
    std::free(t->keys);
    std::free(t->counts);
  }
  std::free(t->bin_offsets);
  std::memset(t, 0, sizeof(*t));
}

extern "C" uint64_t dakc_table_serialize(const dakc_table *t, void *buf, uint64_t cap) {
  if (!t || t->on_device) return 0;
  const uint64_t rec = (uint64_t)t->key_words * 8 + 4;
  const uint64_t need = rec * t->n;
  if (!buf) return need;
  if (cap < need) return 0;
  uint8_t *o = (uint8_t *)buf;
  for (uint64_t i = 0; i < t->n; ++i) {
    std::memcpy(o, t->keys + i * t->key_words, (size_t)t->key_words * 8);
    std::memcpy(o + t->key_words * 8, t->counts + i, 4);
    o += rec;
  }
  return need;
}

// ----------------------------------------------------------- batched count --

extern "C" int dakc_count_batch(dakc_ctx *ctx, const dakc_job *jobs, size_t njobs,
                                dakc_table *outs) {
  if (!ctx || (njobs && (!jobs || !outs))) return DAKC_E_INVAL;
  return batch_count(ctx, jobs, njobs, outs);
}

// ------------------------------------------------------- device helpers --

extern "C" int dakc_device_alloc(dakc_ctx *ctx, uint64_t nbytes, void **dptr) {
  if (!ctx || !dptr) return DAKC_E_INVAL;
  hipSetDevice(ctx->device);
  if (hipMalloc(dptr, nbytes ? nbytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    return DAKC_E_NOMEM;
  }
  return DAKC_OK;
}
extern "C" int dakc_device_free(dakc_ctx *ctx, void *dptr) {
  if (!ctx) return DAKC_E_INVAL;
  hipSetDevice(ctx->device);
  return hipFree(dptr) == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
}
extern "C" int dakc_memcpy_h2d(dakc_ctx *ctx, void *dst, const void *src, uint64_t n) {
  if (!ctx) return DAKC_E_INVAL;
  hipSetDevice(ctx->device);
  return hipMemcpy(dst, src, n, hipMemcpyHostToDevice) == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
}
extern "C" int dakc_memcpy_d2h(dakc_ctx *ctx, void *dst, const void *src, uint64_t n) {
  if (!ctx) return DAKC_E_INVAL;
  hipSetDevice(ctx->device);
  return hipMemcpy(dst, src, n, hipMemcpyDeviceToHost) == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
}

extern "C" int dakc_memcpy_d2d(dakc_ctx *ctx, void *dst, const void *src, uint64_t n) {
  if (!ctx) return DAKC_E_INVAL;
  hipSetDevice(ctx->device);
  if (hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) return DAKC_E_DEVICE;
  return hipStreamSynchronize(ctx->stream) == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
}

// ---------------------------------------------------------- generator --

namespace dk {

int make_gen_core(const dakc_gen_params *p, GenCore *g) {
  if (!p || !g) return DAKC_E_INVAL;
  if (p->read_len < 1 || p->read_len > MAX_READ_LEN) return DAKC_E_INVAL;
  if (p->min_len < 0 || p->min_len > p->read_len) return DAKC_E_INVAL;
  if (p->genome_len && p->genome_len < (uint64_t)p->read_len) return DAKC_E_INVAL;
  if (p->sub_rate < 0 || p->sub_rate > 1 || p->n_rate < 0 || p->n_rate > 1) return DAKC_E_INVAL;
  if (p->n_rate > 0 && (p->n_run_max < 1 || p->n_run_max > 4096)) return DAKC_E_INVAL;
  GenCore &c = *g;
  c.key_iid = gen_stream_key(p->seed, 0);
  c.key_genome = gen_stream_key(p->seed, 1);
  c.key_start = gen_stream_key(p->seed, 2);
  c.key_strand = gen_stream_key(p->seed, 3);
  c.key_err = gen_stream_key(p->seed, 4);
  c.key_nrun = gen_stream_key(p->seed, 5);
  c.key_len = gen_stream_key(p->seed, 6);
  c.num_reads = p->num_reads;
  c.first_read = p->first_read;
  c.genome_len = p->genome_len;
  auto thr = [](double r) -> uint32_t {
    double t = r * 4294967296.0;
    if (t >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)t;
  };
  c.sub_thr = thr(p->sub_rate);
  c.n_thr = thr(p->n_rate);
  c.L = (uint32_t)p->read_len;
  c.min_len = (uint32_t)(p->min_len ? p->min_len : p->read_len);
  c.n_run_max = (uint32_t)std::max(p->n_run_max, 1);
  c.both_strands = p->both_strands ? 1u : 0u;
  return DAKC_OK;
}

}  // namespace dk

extern "C" uint64_t dakc_gen_nbytes(const dakc_gen_params *p) {
  if (!p || p->read_len < 1) return 0;
  return p->num_reads * ((uint64_t)p->read_len + 1);
}

extern "C" int dakc_generate_reads(const dakc_gen_params *p, char *out, uint64_t cap, int threads) {
  GenCore g;
  int s = make_gen_core(p, &g);
  if (s) return s;
  const uint64_t nb = dakc_gen_nbytes(p);
  if (!out || cap < nb) return DAKC_E_INVAL;
  if (threads < 1) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  threads = std::min<int>(threads, 64);
  const uint64_t R = p->num_reads, stride = (uint64_t)p->read_len + 1;
  auto work = [&](uint64_t r0, uint64_t r1) {
    for (uint64_t r = r0; r < r1; ++r) {
      const uint64_t gr = g.first_read + r;
      uint8_t *rec = (uint8_t *)out + r * stride;
      const GenRead ri = gen_read_info(g, gr);
      for (uint32_t j = 0; j < g.L; ++j) {
        if (j >= ri.len) rec[j] = 'M';
        else if (gen_is_n(g, gr, j)) rec[j] = 'N';
        else rec[j] = (uint8_t)(0x54474341u >> (8 * gen_base(g, gr, j, ri)));
      }
      rec[g.L] = '\n';
    }
  };
  if (threads == 1 || R < 1024) {
    work(0, R);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t) th.emplace_back(work, R * t / threads, R * (t + 1) / threads);
    for (auto &t : th) t.join();
  }
  return DAKC_OK;
}

__global__ void k_gen_reads(GenCore g, uint8_t *out, uint64_t nbytes) {
  const uint64_t nw = (nbytes + 3) / 4;
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < nw;
       q += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t o = 4 * q;
    if (o + 4 <= nbytes) {
      uint32_t w = (uint32_t)gen_byte(g, o) | ((uint32_t)gen_byte(g, o + 1) << 8) |
                   ((uint32_t)gen_byte(g, o + 2) << 16) | ((uint32_t)gen_byte(g, o + 3) << 24);
      reinterpret_cast<uint32_t *>(out)[q] = w;
    } else {
      for (uint64_t b = o; b < nbytes; ++b) out[b] = gen_byte(g, b);
    }
  }
}

extern "C" int dakc_generate_reads_device(dakc_ctx *ctx, const dakc_gen_params *p, void *dev_out,
                                          uint64_t cap) {
  if (!ctx) return DAKC_E_INVAL;
  GenCore g;
  int s = make_gen_core(p, &g);
  if (s) return s;
  const uint64_t nb = dakc_gen_nbytes(p);
  if (!dev_out || cap < nb || ((uintptr_t)dev_out & 3)) return DAKC_E_INVAL;
  hipSetDevice(ctx->device);
  if (nb) {
    const uint64_t nw = (nb + 3) / 4;
    const uint64_t grid = std::min<uint64_t>((nw + 255) / 256, (uint64_t)ctx->num_cus * 64);
    k_gen_reads<<<(uint32_t)grid, 256, 0, ctx->stream>>>(g, (uint8_t *)dev_out, nb);
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return DAKC_E_DEVICE;
  return DAKC_OK;
}
