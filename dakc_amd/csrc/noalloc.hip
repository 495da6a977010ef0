// noalloc.hip -- the allocation-free count (dakc_count_noalloc).
//
// For callers that cannot allocate device memory: the whole workspace of the
// count lives in one file-scope device array of DAKC_NOALLOC_SCRATCH bytes.
// Each GPU the library runs on holds its own copy of that array until the
// process ends.  Its device address is looked up once per device, when a
// context is created on it (with that device current), and only that address
// is handed to kernels and copies.  Every call on one GPU shares the array and
// finds the previous call's contents in it: the counters, flags and tickets a
// count needs zeroed are cleared by fill kernels at its start, and calls that
// overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "dk_api_internal.hpp"

namespace dk {

static __device__ __align__(256) uint8_t scratch[DAKC_NOALLOC_SCRATCH];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_scratch[MAX_DEVICES] = {};  // device address of `scratch` on each device
int g_status[MAX_DEVICES] = {};     // DAKC_OK once looked up, else why not

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

// Called by dakc_ctx_create with `device` current.
int noalloc_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return DAKC_E_INVAL;
  std::lock_guard<std::mutex> g(g_mu);
  if (g_scratch[device]) return DAKC_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    (void)hipGetLastError();
    g_status[device] = DAKC_E_DEVICE;
    return DAKC_E_DEVICE;
  }
  g_scratch[device] = p;
  g_status[device] = DAKC_OK;
  return DAKC_OK;
}

// LSD workspace of plan pl cut from `base` (bytes needed when base is null).
static size_t lsd_layout(const Plan &pl, uint8_t *base, Workspace *w) {
  const size_t kb = (size_t)pl.kw * 8;
  const size_t ft = std::max<uint64_t>(std::max<uint64_t>(pl.tiles0, pl.tiles_keys), 1);
  size_t o = 0;
  auto cut = [&](size_t bytes) {
    uint8_t *p = base ? base + o : nullptr;
    o = al256(o + bytes);
    return p;
  };
  Workspace v;
  v.hist = (uint64_t *)cut((size_t)MAX_PASSES * 256 * 8);
  v.digit_start = (uint64_t *)cut((size_t)MAX_PASSES * 256 * 8);
  v.n_dev = (uint64_t *)cut(64);
  v.err = (uint32_t *)cut(64);
  v.tickets = (uint32_t *)cut((MAX_PASSES + 2) * 4);
  v.flags = (uint64_t *)cut(ft * 256 * 8);
  v.rle_flags = (uint64_t *)cut((size_t)std::max<uint32_t>(pl.tiles_keys, 1) * 8);
  v.keys_a = cut(std::max<size_t>(pl.nmax, 1) * kb + 256);
  v.keys_b = cut(std::max<size_t>(pl.nmax, 1) * kb + 256);
  v.counts = (uint32_t *)cut(std::max<size_t>(pl.nmax, 1) * 4 + 256);
  v.bin_offsets = (uint64_t *)cut(((size_t)pl.num_pes + 1) * 8 + 64);
  v.key_cap = pl.nmax;
  if (w) *w = v;
  return o;
}

// MSD layout of plan pl: [arena | meta | argument block] (bytes needed).
static size_t msd_bytes(const MsdBatch &b) { return al256(b.arena_bytes) + al256(b.meta_bytes) + al256(b.desc_bytes); }

}  // namespace dk

using namespace dk;

extern "C" int dakc_count_noalloc(dakc_ctx *ctx, const dakc_job *job, uint64_t *keys_out, uint32_t *counts_out,
                                  uint64_t cap, uint64_t *bin_offsets_out, uint64_t *n_distinct,
                                  uint64_t *n_total) {
  if (!ctx || !job || !n_distinct || !n_total) return DAKC_E_INVAL;
  *n_distinct = *n_total = 0;
  if (job->format != DAKC_FMT_FIXED || (job->nbytes && !job->reads_on_device)) return DAKC_E_INVAL;
  if (cap && (!keys_out || !counts_out)) return DAKC_E_INVAL;
  if (ctx->device < 0 || ctx->device >= MAX_DEVICES) return DAKC_E_INVAL;
  void *sp;
  {
    std::lock_guard<std::mutex> g(g_mu);
    sp = g_scratch[ctx->device];
    if (!sp) return g_status[ctx->device] ? g_status[ctx->device] : DAKC_E_DEVICE;
  }
  uint8_t *base = (uint8_t *)sp;
  if (hipSetDevice(ctx->device) != hipSuccess) return DAKC_E_DEVICE;
  Plan pl;
  if (int s = make_plan(job, &pl)) return s;
  apply_tuning(ctx, &pl, ctx->opts.algo);
  const bool all_bins = pl.owner_digit != 0;
  if (all_bins && !bin_offsets_out) return DAKC_E_INVAL;
  hipStream_t st = ctx->stream;
  uint8_t *hmeta = ctx->ws.hmeta;  // pinned, made when the context was
  // the MSD path when its workspace fits, else the LSD path when that fits
  MsdBatch b;
  if (pl.msd) {
    msd_layout(b, {pl}, {false});
    if (msd_bytes(b) > DAKC_NOALLOC_SCRATCH) plan_choose_path(&pl, DAKC_ALGO_LSD);
  }
  if (!pl.msd && lsd_layout(pl, nullptr, nullptr) > DAKC_NOALLOC_SCRATCH) return DAKC_E_TOOBIG;
  uint64_t n = 0, d = 0;
  TableLoc loc;
  int s = DAKC_OK;
  if (pl.msd) {
    uint8_t *ar = base, *mt = base + al256(b.arena_bytes), *dd = mt + al256(b.meta_bytes);
    b.descs.assign(1, JobDesc());
    b.locs.assign(1, TableLoc());
    msd_describe(b, 0, (const uint8_t *)job->reads, ar, mt, b.descs[0], b.locs[0]);
    msd_enqueue_at(st, b, ar, mt, dd, ctx->ws.fin_persistent, nullptr);
    loc = b.locs[0];
    if (hipMemcpyAsync(hmeta, loc.nd, 32, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return DAKC_E_DEVICE;
    s = msd_finish(b, 0, hmeta, &n, &d);
    if (s == STATUS_MSD_OVERFLOW) {  // skewed input: the LSD path, if its workspace fits
      plan_choose_path(&pl, DAKC_ALGO_LSD);
      if (lsd_layout(pl, nullptr, nullptr) > DAKC_NOALLOC_SCRATCH) return DAKC_E_TOOBIG;
    } else if (s) {
      return s;
    }
  }
  if (!pl.msd) {
    Workspace w;
    lsd_layout(pl, base, &w);
    // the flags hold the last call's epochs: clear them, then count epochs from 1
    launch_fill_u64(w.flags, 0, (uint64_t)std::max<uint64_t>(std::max<uint64_t>(pl.tiles0, pl.tiles_keys), 1) * 256,
                    st);
    launch_fill_u64(w.rle_flags, 0, std::max<uint32_t>(pl.tiles_keys, 1), st);
    uint32_t epoch = 0;
    lsd_enqueue(w, pl, (const uint8_t *)job->reads, st, ctx->num_cus, [&epoch] { return ++epoch; }, nullptr,
                loc);
    if (hipMemcpyAsync(hmeta, w.n_dev, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(hmeta + 16, w.err, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return DAKC_E_DEVICE;
    uint64_t nd[2];
    uint32_t e;
    std::memcpy(nd, hmeta, 16);
    std::memcpy(&e, hmeta + 16, 4);
    if ((s = status_of_err(e))) return s == STATUS_MSD_OVERFLOW ? DAKC_E_DEVICE : s;
    n = nd[0];
    d = pl.nmax ? nd[1] : 0;
  }
  *n_total = n;
  *n_distinct = d;
  if (d > cap) return DAKC_E_NOMEM;  // *n_distinct tells the caller how large to make the table
  if (d && (hipMemcpyAsync(keys_out, loc.keys, d * pl.kw * 8, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(counts_out, loc.counts, d * 4, hipMemcpyDeviceToDevice, st) != hipSuccess))
    return DAKC_E_DEVICE;
  if (all_bins) {
    if (d) {
      if (hipMemcpyAsync(bin_offsets_out, loc.binoff, ((size_t)pl.num_pes + 1) * 8, hipMemcpyDeviceToHost, st) !=
          hipSuccess)
        return DAKC_E_DEVICE;
    } else {
      std::memset(bin_offsets_out, 0, ((size_t)pl.num_pes + 1) * 8);
    }
  }
  return hipStreamSynchronize(st) == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
}
