// batch.hip -- batched counting (one multi-job launch per stage, optionally
// replayed from a captured hipGraph) and the autotuner of the MSD geometry.
//
// Many read sets of uneven size are laid out in one arena and counted by the
// work-queue kernels of the MSD path: every stage numbers its work items across
// all jobs, so a few hundred small jobs fill the GPU like one large job and
// cost one launch per stage instead of a few per job.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dk_api_internal.hpp"

namespace dk {

int size_class(uint64_t nmax) {
  int c = 0;
  while (c < 63 && (nmax >> (c + 1)) != 0) ++c;
  return c;  // one class per factor of 2 in k-mer slots: the sub-bucket count,
             // which sets the final level's time, moves by whole bits
}

void apply_tuning(const dakc_ctx *c, Plan *pl, int algo) {
  TuneEntry e;
  {
    std::lock_guard<std::mutex> g(const_cast<dakc_ctx *>(c)->tune_mu);
    auto it = c->tuning.find({pl->kw, size_class(pl->nmax)});
    if (it != c->tuning.end()) e = it->second;
  }
  pl->tune_fin_nt = e.fin_nt;
  pl->tune_fill_pct = e.fill_pct;
  pl->tune_b1_shift = e.b1_shift;
  pl->num_cus = (uint32_t)std::max(c->num_cus, 1);
  plan_choose_path(pl, algo == DAKC_ALGO_AUTO && e.algo == 1 ? DAKC_ALGO_LSD : algo);
}

size_t msd_job_bytes(const Plan &pl, bool copy_reads);

static size_t hbm_budget(const dakc_ctx *c) {
  if (c->opts.hbm_budget) return c->opts.hbm_budget;
  size_t free_b = 0, total_b = 0;
  hipMemGetInfo(&free_b, &total_b);
  return (size_t)(free_b * 0.6);
}

// copy job j's table out of the arena into library-owned host memory
static int table_to_host(const MsdBatch &b, size_t j, uint64_t n, uint64_t d, dakc_table *out) {
  const Plan &p = b.plans[j];
  const TableLoc &L = b.locs[j];
  std::memset(out, 0, sizeof(*out));
  out->n = d;
  out->total = n;
  out->key_words = p.kw;
  out->k = p.k;
  out->num_bins = p.owner_digit ? p.num_pes : 0;
  out->hash_seed = p.seed;
  const size_t kb = (size_t)d * p.kw * 8;
  out->keys = (uint64_t *)std::malloc(kb ? kb : 8);
  out->counts = (uint32_t *)std::malloc(d ? d * 4 : 4);
  if (p.owner_digit) out->bin_offsets = (uint64_t *)std::calloc((size_t)p.num_pes + 1, 8);
  if (!out->keys || !out->counts || (p.owner_digit && !out->bin_offsets)) {
    dakc_table_free(out);
    return DAKC_E_NOMEM;
  }
  if (d && (hipMemcpy(out->keys, L.keys, kb, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(out->counts, L.counts, d * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            (p.owner_digit && hipMemcpy(out->bin_offsets, L.binoff, ((size_t)p.num_pes + 1) * 8,
                                        hipMemcpyDeviceToHost) != hipSuccess))) {
    dakc_table_free(out);
    return DAKC_E_DEVICE;
  }
  return DAKC_OK;
}

// Copy a device table (src: keys and counts in device memory, the rest of its
// fields set; owner-bin offsets in host memory, or binoff_dev on the device)
// into device memory of its own from the context's table pool, which
// dakc_table_free gives back (a DevTableHold, as for a served table).  For
// jobs with table_on_device, whose tables may together outgrow host memory.
static int copy_words(void *dst, const void *src, size_t bytes, hipStream_t st) {
  if (!bytes) return DAKC_OK;
  if ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0 && bytes % 4 == 0) {
    launch_copy_u32((uint32_t *)dst, (const uint32_t *)src, bytes / 4, st);
    return hipGetLastError() == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
  }
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st) == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
}

// Room for a table's keys (kb bytes) and counts (cb) in the context's table
// pool: the first arena with room for both, else a new arena of at least
// 4 GiB.  Freed tables give their ranges back to their arena, which the pool
// keeps, so the tables of a stream of batches reuse memory instead of
// allocating it again (large allocations after frees measured up to seconds
// each on the MI355X).  Arenas with no live table are let go only when a new
// one cannot be allocated.
// Give the pool's empty arenas back to the device (caller holds table_pool_mu).
static void pool_drop_empty_locked(dakc_ctx *c) {
  c->table_pool.erase(std::remove_if(c->table_pool.begin(), c->table_pool.end(),
                                     [](const std::shared_ptr<DevArena> &x) { return x->largest_free() == x->size; }),
                      c->table_pool.end());
}

static std::shared_ptr<DevArena> pool_alloc(dakc_ctx *c, size_t kb, size_t cb, void **keys, void **counts) {
  std::lock_guard<std::mutex> g(c->table_pool_mu);
  for (auto &a : c->table_pool) {
    void *k = a->alloc(kb);
    if (!k) continue;
    void *n = a->alloc(cb);
    if (!n) {
      a->release(k, kb);
      continue;
    }
    *keys = k;
    *counts = n;
    return a;
  }
  const size_t need = kb + cb;
  auto a = std::make_shared<DevArena>();
  if (a->init(c->device, std::max<size_t>(need, (size_t)4 << 30)) && a->init(c->device, need)) {
    pool_drop_empty_locked(c);
    if (a->init(c->device, need)) return nullptr;
  }
  *keys = a->alloc(kb);
  *counts = a->alloc(cb);
  c->table_pool.push_back(a);
  return a;
}

static int own_device_table(dakc_ctx *c, hipStream_t st, const dakc_table &src, const uint64_t *binoff_dev,
                            dakc_table *out) {
  std::memset(out, 0, sizeof(*out));
  const size_t kb = std::max<size_t>(((size_t)src.n * src.key_words * 8 + 255) & ~(size_t)255, 256);
  const size_t cb = std::max<size_t>(((size_t)src.n * 4 + 255) & ~(size_t)255, 256);
  void *keys = nullptr, *counts = nullptr;
  auto arena = pool_alloc(c, kb, cb, &keys, &counts);
  if (!arena) return DAKC_E_NOMEM;
  auto *h = new (std::nothrow) DevTableHold();
  uint64_t *binoff = src.num_bins ? (uint64_t *)std::calloc((size_t)src.num_bins + 1, 8) : nullptr;
  if (!h || (src.num_bins && !binoff)) {
    delete h;
    std::free(binoff);
    arena->release(keys, kb);
    arena->release(counts, cb);
    return DAKC_E_NOMEM;
  }
  h->arena = arena;
  h->parts[0] = {keys, kb};
  h->parts[1] = {counts, cb};
  h->bytes = kb + cb;
  // copied by a kernel on the context's stream (after the count that wrote src)
  const size_t nk = (size_t)src.n * src.key_words * 8, nc = (size_t)src.n * 4;
  bool ok = copy_words(keys, src.keys, nk, st) == DAKC_OK && copy_words(counts, src.counts, nc, st) == DAKC_OK &&
            hipStreamSynchronize(st) == hipSuccess;
  if (ok && binoff) {
    const size_t ob = ((size_t)src.num_bins + 1) * 8;
    ok = binoff_dev ? hipMemcpy(binoff, binoff_dev, ob, hipMemcpyDeviceToHost) == hipSuccess
                    : (src.bin_offsets ? (std::memcpy(binoff, src.bin_offsets, ob), true) : false);
  }
  *out = src;
  out->keys = (uint64_t *)keys;
  out->counts = (uint32_t *)counts;
  out->bin_offsets = binoff;
  out->on_device = 1;
  out->priv = h;
  if (!ok) {
    dakc_table_free(out);
    return DAKC_E_DEVICE;
  }
  return DAKC_OK;
}

// job j's table out of the arena: host memory, or device memory of its own
// when the job asks for a device table
static int table_out(dakc_ctx *c, hipStream_t st, const MsdBatch &b, size_t j, uint64_t n, uint64_t d,
                     bool on_device, dakc_table *out) {
  if (!on_device) return table_to_host(b, j, n, d, out);
  const Plan &p = b.plans[j];
  const TableLoc &L = b.locs[j];
  dakc_table src;
  std::memset(&src, 0, sizeof(src));
  src.keys = (uint64_t *)L.keys;
  src.counts = (uint32_t *)L.counts;
  src.n = d;
  src.total = n;
  src.key_words = p.kw;
  src.k = p.k;
  src.num_bins = p.owner_digit ? p.num_pes : 0;
  src.hash_seed = p.seed;
  return own_device_table(c, st, src, p.owner_digit ? (const uint64_t *)L.binoff : nullptr, out);
}

// Job j's table, written by the final level into pool memory reserved for it
// (keys kb, counts cb bytes): the ranges past the table go back to the pool
// and the table owns the rest.
static int table_in_place(const MsdBatch &b, size_t j, uint64_t n, uint64_t d, const std::shared_ptr<DevArena> &arena,
                          size_t kb, size_t cb, dakc_table *out) {
  const Plan &p = b.plans[j];
  const TableLoc &L = b.locs[j];
  void *keys = b.dest_keys[j];
  uint32_t *counts = b.dest_counts[j];
  const size_t ku = std::max<size_t>(((size_t)d * p.kw * 8 + 255) & ~(size_t)255, 256);
  const size_t cu = std::max<size_t>(((size_t)d * 4 + 255) & ~(size_t)255, 256);
  std::memset(out, 0, sizeof(*out));
  auto *h = new (std::nothrow) DevTableHold();
  uint64_t *binoff = p.owner_digit ? (uint64_t *)std::calloc((size_t)p.num_pes + 1, 8) : nullptr;
  bool ok = h && (!p.owner_digit || binoff);
  if (ok && p.owner_digit)
    ok = hipMemcpy(binoff, L.binoff, ((size_t)p.num_pes + 1) * 8, hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) {
    const int st = h && (!p.owner_digit || binoff) ? DAKC_E_DEVICE : DAKC_E_NOMEM;
    delete h;
    std::free(binoff);
    arena->release(keys, kb);
    arena->release(counts, cb);
    return st;
  }
  if (kb > ku) arena->release((uint8_t *)keys + ku, kb - ku);
  if (cb > cu) arena->release((uint8_t *)counts + cu, cb - cu);
  h->arena = arena;
  h->parts[0] = {keys, ku};
  h->parts[1] = {counts, cu};
  h->bytes = ku + cu;
  out->keys = (uint64_t *)keys;
  out->counts = counts;
  out->n = d;
  out->total = n;
  out->key_words = p.kw;
  out->k = p.k;
  out->num_bins = p.owner_digit ? p.num_pes : 0;
  out->hash_seed = p.seed;
  out->bin_offsets = binoff;
  out->on_device = 1;
  out->priv = h;
  return DAKC_OK;
}

// Count the jobs listed in `idx` (all MSD-planned with the same key width,
// ordered by final-level size) in one multi-job launch on ws; overflowed jobs
// go to `redo`.
static int run_msd_group(dakc_ctx *c, JobWs &ws, const dakc_job *jobs, const std::vector<Plan> &plans,
                         const std::vector<size_t> &idx, dakc_table *outs, std::vector<int> &status,
                         std::vector<size_t> &redo) {
  MsdBatch b;
  std::vector<Plan> ps;
  std::vector<bool> copy;
  std::vector<const dakc_job *> js;
  for (size_t i : idx) {
    ps.push_back(plans[i]);
    copy.push_back(!jobs[i].reads_on_device);
    js.push_back(&jobs[i]);
  }
  msd_layout(b, ps, copy, solo_items(c));
  // A job with table_on_device has its final level write straight into pool
  // memory of its own, reserved for the most entries it can have (one per
  // k-mer slot) and trimmed to the table afterwards; when the pool cannot
  // give it, the table is written in the arena and copied out.
  struct Reserved {
    std::shared_ptr<DevArena> arena;
    size_t kb = 0, cb = 0;
  };
  std::vector<Reserved> res(idx.size());
  b.dest_keys.assign(idx.size(), nullptr);
  b.dest_counts.assign(idx.size(), nullptr);
  for (size_t j = 0; j < idx.size(); ++j) {
    if (!jobs[idx[j]].table_on_device) continue;
    const Plan &p = ps[j];
    res[j].kb = std::max<size_t>(((size_t)p.nmax * p.kw * 8 + 255) & ~(size_t)255, 256);
    res[j].cb = std::max<size_t>(((size_t)p.nmax * 4 + 255) & ~(size_t)255, 256);
    void *k = nullptr, *n = nullptr;
    res[j].arena = pool_alloc(c, res[j].kb, res[j].cb, &k, &n);
    if (res[j].arena) {
      b.dest_keys[j] = k;
      b.dest_counts[j] = (uint32_t *)n;
    }
  }
  auto unreserve = [&](size_t j) {
    if (!res[j].arena) return;
    res[j].arena->release(b.dest_keys[j], res[j].kb);
    res[j].arena->release(b.dest_counts[j], res[j].cb);
    res[j].arena.reset();
  };
  int s = msd_stage(ws, b, js);
  if (s == DAKC_E_NOMEM) {
    // the reservations may be what left the workspace no room: drop them and
    // the pool's empty arenas (the tables are then written in the workspace
    // and copied out) and stage again
    bool any = false;
    for (size_t j = 0; j < idx.size(); ++j) {
      any |= res[j].arena != nullptr;
      unreserve(j);
      b.dest_keys[j] = nullptr;
      b.dest_counts[j] = nullptr;
    }
    if (any) {
      {
        std::lock_guard<std::mutex> g(c->table_pool_mu);
        pool_drop_empty_locked(c);
      }
      s = msd_stage(ws, b, js);
    }
  }
  if (!s) {
    msd_enqueue(ws, b, false);
    if (hipStreamSynchronize(ws.stream) != hipSuccess) s = DAKC_E_DEVICE;
  }
  std::vector<uint8_t> meta(b.meta_bytes);
  if (!s && hipMemcpy(meta.data(), ws.meta.p, b.meta_bytes, hipMemcpyDeviceToHost) != hipSuccess)
    s = DAKC_E_DEVICE;
  if (s) {
    for (size_t j = 0; j < idx.size(); ++j) unreserve(j);
    return s;
  }
  for (size_t j = 0; j < idx.size(); ++j) {
    uint64_t n = 0, d = 0;
    int st = msd_finish(b, j, meta.data(), &n, &d);
    if (st == STATUS_MSD_OVERFLOW) {
      unreserve(j);
      redo.push_back(idx[j]);
      continue;
    }
    if (!st && res[j].arena) {
      st = table_in_place(b, j, n, d, res[j].arena, res[j].kb, res[j].cb, &outs[idx[j]]);
      res[j].arena.reset();  // the table holds it now, or gave it back
    } else {
      unreserve(j);
      if (!st) st = table_out(c, ws.stream, b, j, n, d, jobs[idx[j]].table_on_device != 0, &outs[idx[j]]);
    }
    status[idx[j]] = st;
  }
  return DAKC_OK;
}

int batch_count(dakc_ctx *c, const dakc_job *jobs, size_t njobs, dakc_table *outs) {
  hipSetDevice(c->device);
  std::vector<Plan> plans(njobs);
  std::vector<int> status(njobs, DAKC_OK);
  std::vector<size_t> plain;  // counted one at a time (LSD path, or no MSD geometry)
  std::map<int, std::vector<size_t>> groups;  // by key width
  for (size_t i = 0; i < njobs; ++i) {
    std::memset(&outs[i], 0, sizeof(outs[i]));
    status[i] = make_plan(&jobs[i], &plans[i]);
    if (status[i]) continue;
    apply_tuning(c, &plans[i], c->opts.algo);
    if (plans[i].msd) groups[plans[i].kw].push_back(i);
    else plain.push_back(i);
  }
  const size_t budget = hbm_budget(c);
  std::vector<size_t> redo;
  for (auto &kv : groups) {
    // chunks whose arena fits the budget, largest jobs first
    std::vector<size_t> ids = kv.second;
    std::sort(ids.begin(), ids.end(), [&](size_t a, size_t b) { return plans[a].nmax > plans[b].nmax; });
    std::vector<size_t> chunk;
    size_t bytes = 0;
    for (size_t k = 0; k <= ids.size(); ++k) {
      const size_t jb = k < ids.size() ? msd_job_bytes(plans[ids[k]], !jobs[ids[k]].reads_on_device) : 0;
      if (k == ids.size() || (!chunk.empty() && bytes + jb > budget)) {
        if (!chunk.empty()) {
          // small jobs of one final-level size are consecutive (one shared
          // final launch), large ones follow
          std::stable_sort(chunk.begin(), chunk.end(), [&](size_t a, size_t b) {
            return plans[a].fin_nt != plans[b].fin_nt ? plans[a].fin_nt < plans[b].fin_nt
                                                      : plans[a].nsb < plans[b].nsb;
          });
          int s = run_msd_group(c, c->ws, jobs, plans, chunk, outs, status, redo);
          if (s)
            for (size_t i : chunk) status[i] = s;
        }
        chunk.clear();
        bytes = 0;
      }
      if (k < ids.size()) {
        chunk.push_back(ids[k]);
        bytes += jb;
      }
    }
  }
  // jobs without an MSD geometry, and MSD overflows, on the plain path
  dakc_job lsd_job;
  for (size_t i : redo) plain.push_back(i);
  const int saved_algo = c->opts.algo;
  for (size_t i : plain) {
    if (std::find(redo.begin(), redo.end(), i) != redo.end()) c->opts.algo = DAKC_ALGO_LSD;
    lsd_job = jobs[i];
    if (lsd_job.table_on_device) {
      // the workspace table of dakc_count_device, copied to memory of its own
      dakc_table ws_table;
      status[i] = dakc_count_device(c, &lsd_job, &ws_table);
      if (!status[i]) status[i] = own_device_table(c, c->ws.stream, ws_table, nullptr, &outs[i]);
      dakc_table_free(&ws_table);
    } else {
      status[i] = dakc_count(c, &lsd_job, &outs[i]);
    }
    c->opts.algo = saved_algo;
  }
  int first = DAKC_OK;
  for (size_t i = 0; i < njobs; ++i)
    if (status[i] && !first) first = status[i];
  return first;
}

}  // namespace dk

using namespace dk;

// ------------------------------------------------------- captured batches --

struct dakc_batch {
  dakc_ctx *ctx = nullptr;
  JobWs ws;
  hipStream_t stream = nullptr;
  MsdBatch layout;
  std::vector<dakc_job> jobs;
  std::vector<Plan> plans;
  std::vector<size_t> msd_idx, plain_idx;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool ran = false;
};

extern "C" int dakc_batch_create(dakc_ctx *ctx, const dakc_job *jobs, size_t njobs, dakc_batch **out) {
  if (!ctx || !out || (njobs && !jobs)) return DAKC_E_INVAL;
  *out = nullptr;
  hipSetDevice(ctx->device);
  auto *b = new (std::nothrow) dakc_batch();
  if (!b) return DAKC_E_NOMEM;
  b->ctx = ctx;
  b->jobs.assign(jobs, jobs + njobs);
  b->plans.resize(njobs);
  size_t lead = njobs;  // the largest MSD job: its key width rules the launch
  for (size_t i = 0; i < njobs; ++i) {
    int s = make_plan(&jobs[i], &b->plans[i]);
    if (s) {
      delete b;
      return s;
    }
    apply_tuning(ctx, &b->plans[i], ctx->opts.algo);
    Plan &p = b->plans[i];
    // a captured batch is one multi-job MSD launch: take every job it can
    if (!p.msd && ctx->opts.algo != DAKC_ALGO_LSD) plan_choose_path(&p, DAKC_ALGO_MSD);
    if (p.msd && (lead == njobs || p.nmax > b->plans[lead].nmax)) lead = i;
  }
  const int kw = lead < njobs ? b->plans[lead].kw : -1;
  for (size_t i = 0; i < njobs; ++i) {
    const Plan &p = b->plans[i];
    if (p.msd && p.kw == kw) b->msd_idx.push_back(i);
    else b->plain_idx.push_back(i);
  }
  // every job keeps its tuned final-level size; small jobs of one size are
  // consecutive (one shared final launch), large ones follow
  std::stable_sort(b->msd_idx.begin(), b->msd_idx.end(), [&](size_t x, size_t y) {
    const Plan &p = b->plans[x], &q = b->plans[y];
    return p.fin_nt != q.fin_nt ? p.fin_nt < q.fin_nt : p.nsb < q.nsb;
  });
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      b->ws.init(b->stream) || hipEventCreate(&b->e0) != hipSuccess ||
      hipEventCreate(&b->e1) != hipSuccess) {
    dakc_batch_destroy(b);
    return DAKC_E_DEVICE;
  }
  std::vector<Plan> ps;
  std::vector<bool> copy;
  std::vector<const dakc_job *> js;
  for (size_t i : b->msd_idx) {
    ps.push_back(b->plans[i]);
    copy.push_back(!jobs[i].reads_on_device);
    js.push_back(&b->jobs[i]);
  }
  msd_layout(b->layout, ps, copy, solo_items(ctx));
  int s = msd_stage(b->ws, b->layout, js);  // host reads are copied once, here
  if (s) {
    dakc_batch_destroy(b);
    return s;
  }
  // capture fills + three levels; replays need no host work
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    dakc_batch_destroy(b);
    return DAKC_E_DEVICE;
  }
  msd_enqueue(b->ws, b->layout, false);
  if (hipStreamEndCapture(b->stream, &b->graph) != hipSuccess ||
      hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    dakc_batch_destroy(b);
    return DAKC_E_DEVICE;
  }
  *out = b;
  return DAKC_OK;
}

extern "C" int dakc_batch_info(const dakc_batch *b, size_t *jobs_in_launch, uint64_t *device_bytes) {
  if (!b) return DAKC_E_INVAL;
  if (jobs_in_launch) *jobs_in_launch = b->msd_idx.size();
  if (device_bytes) *device_bytes = b->ws.bytes_held();
  return DAKC_OK;
}

extern "C" int dakc_batch_run(dakc_batch *b, double *device_ms) {
  if (!b) return DAKC_E_INVAL;
  hipSetDevice(b->ctx->device);
  if (b->msd_idx.empty()) {
    if (device_ms) *device_ms = 0;
    b->ran = true;
    return DAKC_OK;
  }
  if (hipEventRecord(b->e0, b->stream) != hipSuccess ||
      hipGraphLaunch(b->exec, b->stream) != hipSuccess ||
      hipEventRecord(b->e1, b->stream) != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess)
    return DAKC_E_DEVICE;
  float ms = 0;
  hipEventElapsedTime(&ms, b->e0, b->e1);
  if (device_ms) *device_ms = ms;
  b->ran = true;
  return DAKC_OK;
}

extern "C" int dakc_batch_tables(dakc_batch *b, dakc_table *outs) {
  if (!b || !outs) return DAKC_E_INVAL;
  hipSetDevice(b->ctx->device);
  const size_t nj = b->jobs.size();
  for (size_t i = 0; i < nj; ++i) std::memset(&outs[i], 0, sizeof(outs[i]));
  int first = DAKC_OK;
  std::vector<size_t> redo;
  if (!b->msd_idx.empty()) {
    if (!b->ran) {
      int s = dakc_batch_run(b, nullptr);
      if (s) return s;
    }
    std::vector<uint8_t> meta(b->layout.meta_bytes);
    if (hipMemcpy(meta.data(), b->ws.meta.p, meta.size(), hipMemcpyDeviceToHost) != hipSuccess)
      return DAKC_E_DEVICE;
    for (size_t j = 0; j < b->msd_idx.size(); ++j) {
      uint64_t n = 0, d = 0;
      int st = msd_finish(b->layout, j, meta.data(), &n, &d);
      if (st == STATUS_MSD_OVERFLOW) {
        redo.push_back(b->msd_idx[j]);
        continue;
      }
      if (!st) st = table_to_host(b->layout, j, n, d, &outs[b->msd_idx[j]]);
      if (st && !first) first = st;
    }
  }
  const int saved = b->ctx->opts.algo;
  for (size_t i : b->plain_idx) {
    int st = dakc_count(b->ctx, &b->jobs[i], &outs[i]);
    if (st && !first) first = st;
  }
  b->ctx->opts.algo = DAKC_ALGO_LSD;
  for (size_t i : redo) {
    int st = dakc_count(b->ctx, &b->jobs[i], &outs[i]);
    if (st && !first) first = st;
  }
  b->ctx->opts.algo = saved;
  return first;
}

extern "C" int dakc_batch_destroy(dakc_batch *b) {
  if (!b) return DAKC_E_INVAL;
  hipSetDevice(b->ctx->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->e0) hipEventDestroy(b->e0);
  if (b->e1) hipEventDestroy(b->e1);
  b->ws.~JobWs();
  new (&b->ws) JobWs();
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return DAKC_OK;
}

// -------------------------------------------------------------- autotuner --

// Time the candidate MSD geometries (final-level workgroup of 256, 512 or 1024
// threads, target final-item fill 60/75/90%) and the LSD path on `job`, keep
// the fastest for the job's (key width, size class).  The count is exact for
// every candidate; only the time differs.
extern "C" int dakc_autotune(dakc_ctx *ctx, const dakc_job *job, int trials, double *best_ms) {
  if (!ctx || !job) return DAKC_E_INVAL;
  if (trials < 1) trials = 1;
  Plan base;
  int s = make_plan(job, &base);
  if (s) return s;
  struct Cand {
    uint32_t fin, fill, algo;
    int32_t b1_shift;
  };
  // the default geometry first: another candidate replaces the best so far
  // only when it is more than 5% faster (the median of `trials` timed runs)
  // (fill 0: the key width's default, 90 or 75)
  const Cand cands[] = {{1024, 0, 0, 0},  {1024, 90, 0, 0}, {1024, 75, 0, 0}, {1024, 60, 0, 0},
                        {1024, 0, 0, 1},  {512, 90, 0, 0},  {512, 75, 0, 0},  {512, 60, 0, 0},
                        {256, 90, 0, 0},  {1024, 0, 1, 0}};
  TuneEntry best;
  best.ms = 1e300;
  dakc_table t;
  // the table entry before tuning, kept when no candidate counts the job
  TuneEntry before;
  bool had = false;
  {
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    auto it = ctx->tuning.find({base.kw, size_class(base.nmax)});
    if ((had = it != ctx->tuning.end())) before = it->second;
  }
  int last_err = DAKC_OK;
  bool any = false;
  for (const Cand &cd : cands) {
    {
      std::lock_guard<std::mutex> g(ctx->tune_mu);
      TuneEntry e;
      e.fin_nt = cd.fin;
      e.fill_pct = cd.fill;
      e.algo = cd.algo;
      e.b1_shift = cd.b1_shift;
      ctx->tuning[{base.kw, size_class(base.nmax)}] = e;
    }
    std::vector<double> ms;
    bool ok = true;
    for (int r = 0; r < trials + 1 && ok; ++r) {  // first run warms the workspace
      s = dakc_count_device(ctx, job, &t);
      dakc_table_free(&t);
      if (s) {
        ok = false;
        last_err = s;
        break;
      }
      dakc_stats st;
      dakc_last_stats(ctx, &st);
      if (r > 0) ms.push_back(st.total_ms + (st.msd_reruns ? 1e9 : 0));
    }
    if (!ok) continue;
    any = true;
    std::sort(ms.begin(), ms.end());
    const double med = ms[ms.size() / 2];
    if (med < best.ms * (best.ms < 1e300 ? 0.95 : 1.0)) {
      best.fin_nt = cd.fin;
      best.fill_pct = cd.fill;
      best.algo = cd.algo;
      best.b1_shift = cd.b1_shift;
      best.ms = med;
    }
  }
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  if (!any) {  // no candidate counted the job: the table stays as it was
    if (had) ctx->tuning[{base.kw, size_class(base.nmax)}] = before;
    else ctx->tuning.erase({base.kw, size_class(base.nmax)});
    return last_err;
  }
  ctx->tuning[{base.kw, size_class(base.nmax)}] = best;
  if (best_ms) *best_ms = best.ms;
  return DAKC_OK;
}

extern "C" int dakc_tuning_save(dakc_ctx *ctx, const char *path) {
  if (!ctx || !path) return DAKC_E_INVAL;
  std::FILE *f = std::fopen(path, "w");
  if (!f) return DAKC_E_INVAL;
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  std::fprintf(f, "{\n \"entries\": [\n");
  size_t i = 0;
  for (auto &kv : ctx->tuning) {
    std::fprintf(f, "  {\"key_words\": %d, \"size_class\": %d, \"fin_nt\": %u, \"fill_pct\": %u, "
                 "\"algo\": \"%s\", \"ms\": %.4f, \"b1_shift\": %d}%s\n",
                 kv.first.first, kv.first.second, kv.second.fin_nt, kv.second.fill_pct,
                 kv.second.algo ? "lsd" : "msd", kv.second.ms, kv.second.b1_shift,
                 ++i < ctx->tuning.size() ? "," : "");
  }
  std::fprintf(f, " ]\n}\n");
  std::fclose(f);
  return DAKC_OK;
}

extern "C" int dakc_tuning_load(dakc_ctx *ctx, const char *path) {
  if (!ctx || !path) return DAKC_E_INVAL;
  std::FILE *f = std::fopen(path, "r");
  if (!f) return DAKC_E_INVAL;
  char line[512];
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  while (std::fgets(line, sizeof line, f)) {
    int kw, cls;
    unsigned fin, fill;
    char algo[8] = {0};
    double ms;
    int shift = 0;
    const int got = std::sscanf(line, " {\"key_words\": %d, \"size_class\": %d, \"fin_nt\": %u, \"fill_pct\": %u, "
                                      "\"algo\": \"%3[a-z]\", \"ms\": %lf, \"b1_shift\": %d}",
                                &kw, &cls, &fin, &fill, algo, &ms, &shift);
    if (got >= 6) {  // files without b1_shift (6 fields) keep the planner's width
      TuneEntry e;
      e.fin_nt = fin == 256 ? 256 : fin == 1024 ? 1024 : 512;
      // above 100 the items outgrow LDS and are counted digit by digit; 0 is
      // the key width's default
      e.fill_pct = fill == 0 ? 0u : std::min(400u, std::max(30u, fill));
      e.algo = std::strcmp(algo, "lsd") == 0 ? 1 : 0;
      e.ms = ms;
      e.b1_shift = got == 7 ? std::max(-2, std::min(2, shift)) : 0;
      ctx->tuning[{kw, cls}] = e;
    }
  }
  std::fclose(f);
  return DAKC_OK;
}
