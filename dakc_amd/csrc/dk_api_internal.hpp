// dk_api_internal.hpp -- host-side state behind the C ABI (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/dakc.h"
#include "dk_device.hpp"
#include "dk_kernels.hpp"

namespace dk {

// Grow-only device buffer.
// Host-side allocator over one device allocation (the served pool): best fit
// over free ranges, coalesced on release, 256-byte granularity.  Shared by the
// served scheduler and every device result table cut from it, so the memory
// lives until the last of them is gone.
struct DevArena {
  void *base = nullptr;
  size_t size = 0;
  int device = 0;
  std::mutex mu;
  std::map<size_t, size_t> free_ranges;  // offset -> length
  size_t used = 0, peak = 0;
  ~DevArena();
  int init(int dev, size_t bytes);
  // lowest-address first fit (job buffers) / highest-address fit (result
  // tables), so long-lived tables pack at the top and leave the bottom
  // contiguous; nullptr when no free range is large enough
  void *alloc(size_t bytes);
  void *alloc_top(size_t bytes);
  void release(void *p, size_t bytes);
  size_t largest_free();  // bytes of the largest free range
};

// A grow-only device buffer: from hipMalloc, or (served slots) from an arena.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  DevArena *arena = nullptr;
  int ensure(size_t bytes, hipStream_t st, bool zero);
  void release();
};

// Geometry and parameters of one counting job.
struct Plan {
  int kw = 1, k = 31, tile = 0;
  uint32_t L = 0, stride = 0, S = 0, RT = 1, wpr = 0;
  uint64_t nbytes = 0, nrec = 0, nmax = 0;
  uint32_t tiles0 = 0, tiles_keys = 0;
  uint32_t num_pes = 1;
  int32_t owner_bin = -1;
  uint64_t seed = OWNER_SEED;  // owner-hash seed
  int ndig = 0, owner_digit = 0, npasses = 0;
  uint32_t canonical = 0, last_has_nl = 1;
  // the input is an array of nmax keys (the owner side of a P-PE run), not records
  int from_keys = 0;
  // MSD path (msd = 1): first-level bits b1 (with the owner when all bins
  // are counted), second-level b2, in-LDS b3; L1 tiles of RT1 records.  A
  // final item is `pack` consecutive second-level digits of a bucket (ipb
  // items per bucket, nsb items in all).
  int msd = 0;
  uint32_t b1 = 0, b2 = 0, b3 = 0, nb1 = 1, nd2 = 1, nsb = 1, pack = 1;
  uint32_t cls_log = 0, packc[4] = {1, 0, 0, 0}, ipbc[4] = {1, 0, 0, 0};  // final-item classes (MsdArgs)
  uint32_t RT1 = 1, wpr1 = 0, ntiles1 = 0, tcap1 = 0, G = 1, ngroups = 0, cap2 = 0, fin_nt = 256;
  // tunables set before plan_choose_path (autotuner; defaults otherwise),
  // and the CUs of the device the job runs on
  uint32_t tune_fin_nt = 1024, tune_fill_pct = 0, num_cus = 256;  // fill 0: the key width's default
  int32_t tune_b1_shift = 0;  // first-level key bits over (or under) the planner's choice
};

// path choice (dakc_opts.algo)
int plan_choose_path(Plan *pl, int algo);

// Where a finished count left its table and status words (device pointers).
struct TableLoc {
  void *keys = nullptr;
  uint32_t *counts = nullptr;
  uint64_t *binoff = nullptr;
  uint64_t *nd = nullptr;  // [0] k-mers, [1] distinct
  uint32_t *err = nullptr;
};

// A laid-out multi-job MSD launch: arena offsets per job and the device
// argument block (JobDesc table, work-item prefixes, ticket).
struct MsdBatch {
  int kw = 1;
  // Final-level launches.  A job with at least FIN_SOLO_ITEMS sub-buckets has
  // a launch of its own (solo = its index: its descriptor is a kernel
  // argument, as in a one-job count); consecutive smaller jobs with one
  // workgroup size share a work-queue launch over their items [item0, item1)
  // (solo = -1), where each workgroup looks its job up.
  struct FinRun {
    uint32_t fin_nt, item0, item1;
    int32_t solo;
  };
  std::vector<FinRun> fin_runs;
  std::vector<Plan> plans;
  std::vector<size_t> off_reads, off_out1, off_out2, off_counts, off_off1, off_off2, off_flags,
      off_binoff;
  std::vector<uint32_t> pre1, pre2, pre3;
  size_t arena_bytes = 0, flags_off = 0, flags_words = 0;
  size_t meta_bytes = 0, desc_bytes = 0;
  std::vector<TableLoc> locs;
  std::vector<JobDesc> descs;  // host copies of the job descriptors
  // where job j's final level writes its table (empty or null: the arena)
  std::vector<void *> dest_keys;
  std::vector<uint32_t *> dest_counts;
};

// Device workspace of one in-flight job (one per served slot).
struct JobWs {
  DevBuf hist, dstart, ndev, err, tickets, flags, rle_flags, keys_a, keys_b, counts, binoff, reads,
      msd_off1, msd_off2, arena, meta, desc;
  TableLoc loc;     // table of the last enqueued count
  MsdBatch single;  // layout of the last single-job MSD count
  // one-job final levels: persistent workgroups over all the items (plain
  // counts), or one item per workgroup (served slots, whose jobs share the
  // GPU: short workgroups let the other slots' kernels in between)
  bool fin_persistent = true;
  hipStream_t stream = nullptr;
  hipEvent_t ev[MAX_PASSES + 4] = {};
  uint32_t epoch = 0;
  // pinned host memory: the job's status words land in hmeta (copied at the
  // end of every enqueue, so finishing a job makes no HIP call), the MSD
  // argument block is staged in hstage (async upload; stage_ev marks its end)
  uint8_t *hmeta = nullptr, *hstage = nullptr;
  size_t hstage_cap = 0;
  hipEvent_t stage_ev = nullptr;
  ~JobWs();
  int init(hipStream_t st, DevArena *arena = nullptr);
  void release_job_buffers();  // return the per-job buffers now (no device work may use them)
  // hand the per-job buffers over as (pointer, bytes) ranges, to be released
  // once the device work still reading them has finished
  void detach_job_buffers(std::vector<std::pair<void *, size_t>> &out);
  int reserve(const Plan &pl, bool need_reads);
  uint32_t next_epoch();
  Workspace view() const;
  size_t bytes_held() const {
    size_t t = 0;
    for (const DevBuf *b : {&hist, &dstart, &ndev, &err, &tickets, &flags, &rle_flags, &keys_a, &keys_b,
                            &counts, &binoff, &reads, &msd_off1, &msd_off2, &arena, &meta, &desc})
      t += b->cap;
    return t;
  }
};

int make_plan(const dakc_job *job, Plan *pl);
// the kernels' view of a plan's records (or keys) at device address dreads
ReadsDesc plan_reads(const Plan &p, const uint8_t *dreads);
// plan of counting an array of n keys of k bases (owner bins when num_pes > 1)
int make_plan_keys(uint64_t n, int k, uint32_t num_pes, uint64_t seed, Plan *pl);
// the count of a planned job on the context's workspace (dakc_count[_device]
// after the plan): table to host memory or left on the device
int count_planned(dakc_ctx *c, Plan &pl, const dakc_job *job, dakc_table *out, bool to_host);
// multi-job MSD: lay out, stage (allocate, upload arguments, copy host reads),
// enqueue (fills + kernels; capturable in a hipGraph), collect per-job status
// solo: final-level items from which a job of a multi-job layout gets a
// final-level launch of its own
void msd_layout(MsdBatch &b, const std::vector<Plan> &plans, const std::vector<bool> &copy_reads,
                uint32_t solo = 0xFFFFFFFFu);
uint32_t solo_items(const dakc_ctx *c);
int msd_stage(JobWs &ws, MsdBatch &b, const std::vector<const dakc_job *> &jobs);
// job j's descriptor and table location in a laid-out batch: its buffers at
// the layout's offsets in `arena`, its status words at slot j of `meta`
void msd_describe(const MsdBatch &b, size_t j, const uint8_t *dreads, uint8_t *arena, uint8_t *meta,
                  JobDesc &D, TableLoc &L);
void msd_enqueue(JobWs &ws, const MsdBatch &b, bool events);
void msd_enqueue_at(hipStream_t st, const MsdBatch &b, uint8_t *arena, uint8_t *meta, uint8_t *desc,
                    bool persist, hipEvent_t *ev);
void lsd_enqueue(const Workspace &w, const Plan &pl, const uint8_t *dreads, hipStream_t st, int num_cus,
                 const std::function<uint32_t()> &next_epoch, hipEvent_t *ev, TableLoc &loc);
int msd_finish(const MsdBatch &b, size_t j, const void *meta_host, uint64_t *n, uint64_t *d);
size_t job_device_bytes(const Plan &pl);
int enqueue_count(JobWs &ws, const Plan &pl, const dakc_job *job, const uint8_t *dreads,
                  int num_cus, bool events);
int finish_count(JobWs &ws, const Plan &pl, uint64_t *n, uint64_t *d);

struct ServeState;  // served mode (serve.hip)

}  // namespace dk

namespace dk {
// autotuned MSD parameters for one (key width, job-size class)
struct TuneEntry {
  uint32_t fin_nt = 1024, fill_pct = 0, algo = 0;  // fill 0: the key width's default; algo: 0 MSD, 1 LSD
  int32_t b1_shift = 0;                             // first-level digit width - the planner's
  double ms = 0;
};
int size_class(uint64_t nmax);
void apply_tuning(const dakc_ctx *c, Plan *pl, int algo);
}  // namespace dk

namespace dk {
// device buffers of the HITTER count (grow-only, held by the context): the
// bucket chain, the k-mers in record order, the send streams, and the light
// and heavy tables of the last count
struct HitterBufs {
  DevBuf reads, meta, c, tiles, P, J, starts, keys, norm, pkeys, pcounts, lk, lc, hk, hc, hoff;
  void release() {
    for (DevBuf *b : {&reads, &meta, &c, &tiles, &P, &J, &starts, &keys, &norm, &pkeys, &pcounts, &lk, &lc, &hk,
                      &hc, &hoff})
      b->release();
  }
};
}  // namespace dk

struct dakc_ctx {
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  dakc_opts opts = {};
  dk::JobWs ws;
  dakc_stats last = {};
  dk::ServeState *serve = nullptr;
  std::mutex serve_init_mu;
  std::mutex tune_mu;
  std::map<std::pair<int, int>, dk::TuneEntry> tuning;  // (kw, size class) -> parameters
  dk::DevBuf post_a, post_b;  // scratch of the table histogram / filter / query
  dk::DevBuf conv_in, conv_nl, conv_out;  // FASTQ / line input: staged bytes, newlines, records
  dk::HitterBufs hit;
  // arenas holding the device tables of dakc_count_batch jobs with
  // table_on_device; kept when their tables are freed, for the next ones
  std::vector<std::shared_ptr<dk::DevArena>> table_pool;
  std::mutex table_pool_mu;
};

namespace dk {
int batch_count(dakc_ctx *c, const dakc_job *jobs, size_t njobs, dakc_table *outs);
int convert_reads(dakc_ctx *c, const uint8_t *in, uint64_t nbytes, int format, uint32_t read_len,
                  uint8_t *out, uint64_t cap, DevBuf *grow, uint64_t *out_nbytes, uint32_t *out_read_len);
void serve_shutdown(dakc_ctx *c);
// look up the allocation-free count's device array on `device` (current)
int noalloc_setup(int device);
// stable in-place compaction of a device table to its entries with count >= min_count
int table_filter(dakc_ctx *ctx, dakc_table *t, uint32_t min_count);

// Owner of a device-resident result table (dakc_table.priv): the two ranges
// of the served pool that hold its keys and its counts, kept from the job's
// own buffers when the rest of them went back to the pool.  `resident` counts
// the bytes of the live tables of one served context.
struct DevTableHold {
  std::shared_ptr<DevArena> arena;
  std::pair<void *, size_t> parts[2] = {};
  size_t bytes = 0;
  std::shared_ptr<std::atomic<uint64_t>> resident;
};
void dev_table_release(dakc_table *t);
}  // namespace dk
