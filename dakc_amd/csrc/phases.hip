// phases.hip -- the two phases of a P-PE count as entry points of their own.
//
// The reference runs one PE's count as Phase 1, parse and send every k-mer to
// its owner PE (kcounter.cpp:146-174: add_in_normal_packet, empty_packets),
// then Phase 2, the owner's sort and run-length reduce of what it received
// (kcounter.cpp:123-144 recv_kmer into vectordbg, :385-411).  dakc_count does
// both at once; dakc_partition is Phase 1 alone (the per-owner bins a
// multi-GPU exchange would ship) and dakc_count_keys is Phase 2 alone (a
// table from any array of keys, such as a received bin).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "dk_api_internal.hpp"

using namespace dk;

// Phase 1 on the device: the onesweep histogram + sweep with the owner as
// the digit (a stable scatter into per-owner bins: within a bin the k-mers
// keep their record order), or with one digit when a single bin is kept.
extern "C" int dakc_partition(dakc_ctx *c, const dakc_job *job, uint64_t *keys_out, uint64_t cap,
                              uint64_t *bin_offsets_out, uint64_t *n_out) {
  if (!c || !job || !bin_offsets_out || !n_out) return DAKC_E_INVAL;
  *n_out = 0;
  if (job->format != DAKC_FMT_FIXED) return DAKC_E_INVAL;
  if (cap && !keys_out) return DAKC_E_INVAL;
  hipSetDevice(c->device);
  Plan pl;
  int s = make_plan(job, &pl);
  if (s) return s;
  JobWs &ws = c->ws;
  hipStream_t st = ws.stream;
  if ((s = ws.hist.ensure((size_t)MAX_PASSES * 256 * 8, st, false)) ||
      (s = ws.dstart.ensure((size_t)MAX_PASSES * 256 * 8, st, false)) || (s = ws.ndev.ensure(64, st, false)) ||
      (s = ws.err.ensure(64, st, false)) || (s = ws.tickets.ensure((MAX_PASSES + 2) * 4, st, false)))
    return s;
  const size_t fl_before = ws.flags.cap;
  if ((s = ws.flags.ensure(std::max<size_t>(pl.tiles0, 1) * 256 * 8, st, true))) return s;
  if (ws.flags.cap != fl_before) {  // fresh flag words: restart the epochs (rle_flags too)
    launch_fill_u64((uint64_t *)ws.flags.p, 0, ws.flags.cap / 8, st);
    if (ws.rle_flags.p) launch_fill_u64((uint64_t *)ws.rle_flags.p, 0, ws.rle_flags.cap / 8, st);
    ws.epoch = 0;
  }
  const uint8_t *dreads = (const uint8_t *)job->reads;
  if (!job->reads_on_device && pl.nbytes) {
    if ((s = ws.reads.ensure(pl.nbytes + 256, st, false))) return s;
    if (hipMemcpyAsync(ws.reads.p, job->reads, pl.nbytes, hipMemcpyHostToDevice, st) != hipSuccess)
      return DAKC_E_DEVICE;
    dreads = (const uint8_t *)ws.reads.p;
  }
  Workspace w = ws.view();
  w.key_cap = cap;
  const ReadsDesc rd = plan_reads(pl, dreads);
  const uint32_t nbins = pl.owner_digit ? pl.num_pes : 1;
  const DigitSpec dig = pl.owner_digit ? DigitSpec{DIG_OWNER, 0, 8, pl.num_pes} : DigitSpec{DIG_BITS, 0, 0, 1};
  launch_fill_u64(w.hist, 0, 256, st);
  launch_fill_u64(w.n_dev, 0, 2, st);
  launch_fill_u32(w.err, 0, 1, st);
  launch_fill_u32(w.tickets, 0, MAX_PASSES + 2, st);
  if (pl.nrec) {
    HistSpecs hs;
    hs.n = 1;
    hs.d[0] = dig;
    launch_hist_reads(pl.kw, rd, hs, w, pl.tiles0, (int)std::min<uint64_t>(pl.tiles0, (uint64_t)c->num_cus * 4),
                      st);
    launch_scan_digits(1, w, st);
  }
  uint8_t *hm = ws.hmeta;
  if (hipMemcpyAsync(hm, w.n_dev, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hm + 16, w.err, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return DAKC_E_DEVICE;
  uint64_t n;
  uint32_t e;
  std::memcpy(&n, hm, 8);
  std::memcpy(&e, hm + 16, 4);
  if ((s = status_of_err(e))) return s;
  *n_out = n;
  if (n > cap) return DAKC_E_NOMEM;  // *n_out tells the caller how large to make keys_out
  if (n) launch_sweep(pl.kw, true, rd, nullptr, keys_out, dig, w.digit_start, w, 0, ws.next_epoch(), pl.tiles0, st);
  std::vector<uint64_t> starts(nbins);
  if (hipMemcpyAsync(starts.data(), w.digit_start, (size_t)nbins * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hm + 16, w.err, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return DAKC_E_DEVICE;
  std::memcpy(&e, hm + 16, 4);
  if ((s = status_of_err(e))) return s;
  for (uint32_t o = 0; o < nbins; ++o) bin_offsets_out[o] = n ? starts[o] : 0;
  bin_offsets_out[nbins] = n;
  return DAKC_OK;
}

// Phase 2: the plain count's path (MSD levels with the first level reading
// key tiles, or the LSD passes over the keys) on a plan of n keys.
extern "C" int dakc_count_keys(dakc_ctx *c, const uint64_t *keys, uint64_t n, int k, uint32_t num_pes,
                               uint64_t hash_seed, int keys_on_device, int table_on_device, dakc_table *out) {
  if (!c || !out || (n && !keys)) return DAKC_E_INVAL;
  std::memset(out, 0, sizeof(*out));
  if (keys_on_device && ((uintptr_t)keys & 7)) return DAKC_E_INVAL;
  hipSetDevice(c->device);
  Plan pl;
  int s = make_plan_keys(n, k, num_pes, hash_seed, &pl);
  if (s) return s;
  dakc_job job{};
  job.reads = (const char *)keys;
  job.nbytes = pl.nbytes;
  job.format = DAKC_FMT_FIXED;
  job.k = k;
  job.num_pes = num_pes;
  job.owner_bin = -1;
  job.reads_on_device = keys_on_device ? 1 : 0;
  job.hash_seed = hash_seed;
  return count_planned(c, pl, &job, out, !table_on_device);
}
