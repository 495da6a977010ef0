// serve.hip -- served mode (continuous batching) and the batched entry point.
//
// Served mode keeps a small pool of device-resident job slots, each with its
// own stream and workspace.  A scheduler thread admits queued jobs into free
// slots as long as the slots' device footprint stays inside the HBM budget,
// so requests that arrive while others run start as soon as a slot and memory
// free up, and their kernels overlap on the GPU.  Tables are exact, so a job's
// result does not depend on what else is in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "dk_api_internal.hpp"

namespace dk {

using Ranges = std::vector<std::pair<void *, size_t>>;

// Ticket life: QUEUED -> RUNNING (in a slot) -> COPYING (the table moves to
// the top of the pool; the slot already takes its next job) -> done.
struct Ticket {
  dakc_job job;
  Plan plan;
  int status = DAKC_E_PENDING;
  bool done = false;
  dakc_table table = {};     // device table (priv: DevTableHold) once done
  hipEvent_t copied = nullptr;
  Ranges pending_free;       // the job's buffers, released when the copy is done
};

struct Slot {
  JobWs ws;
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;
  uint64_t ticket = 0;  // 0 = free
  bool large = false;   // the job fills the GPU by itself
  bool msd = false;     // on the MSD path: ws.ev[2] marks its second level done
};

// A large job holds the GPU to the end, or, under a backlog (pipelined), only
// until its final level: the next large job's first levels then fill the GPU
// beside the final level instead of after it.  That raises throughput but
// stretches the earlier job's final level, so it is kept for backlogs.
static bool holds_large(Slot &sl, bool pipelined) {
  if (!sl.ticket || !sl.large) return false;
  return !pipelined || !sl.msd || hipEventQuery(sl.ws.ev[2]) == hipErrorNotReady;
}

// A job fills the GPU by itself when its first level alone is 16 waves of
// workgroups over all CUs; two such jobs side by side only share the GPU and
// stretch each other's latency, so they run one at a time while smaller jobs
// fill the other slots.
static bool is_large(const Plan &p, int num_cus) {
  return p.nmax >= (uint64_t)num_cus * 16 * (uint64_t)msd_tile(p.kw);
}

// Most pool bytes a job can hold once done: its workspace while it runs,
// then its table (every k-mer distinct at worst) until the client claims it.
static size_t job_pool_bound(const Plan &p) {
  return job_device_bytes(p) + (size_t)p.nmax * ((size_t)p.kw * 8 + 4) + 1024;
}

struct ServeState {
  dakc_ctx *ctx = nullptr;
  dakc_serve_stats stats = {};
  std::shared_ptr<DevArena> arena;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done;
  std::deque<uint64_t> queue;
  std::vector<uint64_t> copying;
  std::vector<hipEvent_t> spare_events;
  std::map<uint64_t, std::unique_ptr<Ticket>> tickets;
  std::vector<std::unique_ptr<Slot>> slots;
  std::shared_ptr<std::atomic<uint64_t>> resident = std::make_shared<std::atomic<uint64_t>>(0);
  uint64_t unclaimed = 0;  // bytes of finished tables nobody has waited for yet
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread sched;
};

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

void dev_table_release(dakc_table *t) {
  auto *h = (DevTableHold *)t->priv;
  if (!h) return;
  for (auto &r : h->parts)
    if (r.second) h->arena->release(r.first, r.second);
  if (h->resident) h->resident->fetch_sub(h->bytes);
  delete h;
  std::free(t->bin_offsets);
  std::memset(t, 0, sizeof(*t));
}

static void fail(ServeState &s, Ticket &t, int st) {
  t.status = st;
  t.done = true;
  ++s.stats.failed;
  s.cv_done.notify_all();
}

// return a slot's job buffers at once (after its stream has drained)
static void drop_job_buffers(Slot &sl) {
  hipStreamSynchronize(sl.stream);
  sl.ws.release_job_buffers();
}

// Give back every part of the job's buffers outside the kept ranges.
static void release_except(DevArena &a, const Ranges &job, const std::pair<void *, size_t> *keep, int nkeep) {
  for (const auto &R : job) {
    uintptr_t cur = (uintptr_t)R.first;
    const uintptr_t end = cur + R.second;
    std::vector<std::pair<uintptr_t, uintptr_t>> ks;
    for (int i = 0; i < nkeep; ++i) {
      const uintptr_t b = (uintptr_t)keep[i].first, e = b + keep[i].second;
      if (keep[i].second && b >= cur && e <= end) ks.emplace_back(b, e);
    }
    std::sort(ks.begin(), ks.end());
    for (auto &k : ks) {
      if (k.first > cur) a.release((void *)cur, k.first - cur);
      cur = std::max(cur, k.second);
    }
    if (end > cur) a.release((void *)cur, end - cur);
  }
}

// Tables up to this size move to the top of the pool (a copy of well under a
// millisecond), so the bottom stays contiguous for job buffers; larger ones
// stay where the count wrote them: moving them would cost a few percent of
// their count in HBM traffic.
constexpr size_t TABLE_MOVE_MAX = (size_t)32 << 20;

// The slot's count has finished: check it (rerun an LDS overflow on the LSD
// path), then move a small table to a range at the top of the pool with a
// device copy on the slot's stream; the job's buffers go back once the copy
// is done, and the slot takes its next job at once (its stream orders that
// job after the copy).  A large table, or one the top has no room for, stays
// where the count left it: the two ranges of the job's buffers that hold its
// keys and counts become the table's and the rest goes back to the pool.
static void retire(ServeState &s, Slot &sl) {
  const uint64_t id = sl.ticket;
  Ticket &t = *s.tickets[id];
  uint64_t n = 0, d = 0;
  int st = finish_count(sl.ws, t.plan, &n, &d);
  if (st == STATUS_MSD_OVERFLOW && t.plan.msd) {
    // a sub-bucket did not fit in LDS: count this job again on the LSD path
    plan_choose_path(&t.plan, DAKC_ALGO_LSD);
    ++s.stats.lsd_reruns;
    sl.msd = false;
    drop_job_buffers(sl);
    st = sl.ws.reserve(t.plan, !t.job.reads_on_device);
    if (!st)
      st = enqueue_count(sl.ws, t.plan, &t.job, (const uint8_t *)t.job.reads, s.ctx->num_cus, false);
    if (!st && hipEventRecord(sl.done_ev, sl.stream) != hipSuccess) st = DAKC_E_DEVICE;
    if (!st) return;  // still running
    drop_job_buffers(sl);
    sl.ticket = 0;
    if (st == DAKC_E_NOMEM) {  // wait for memory like a new job, at the head of the queue
      s.queue.push_front(id);
      return;
    }
    fail(s, t, st);
    return;
  }
  sl.ticket = 0;
  dakc_table &o = t.table;
  auto *h = st ? nullptr : new (std::nothrow) DevTableHold();
  if (!st && !h) st = DAKC_E_NOMEM;
  if (!st && t.plan.owner_digit) {
    o.bin_offsets = (uint64_t *)std::calloc(t.plan.num_pes + 1, 8);
    st = !o.bin_offsets ? DAKC_E_NOMEM
         : hipMemcpy(o.bin_offsets, sl.ws.loc.binoff, ((size_t)t.plan.num_pes + 1) * 8,
                     hipMemcpyDeviceToHost) != hipSuccess ? DAKC_E_DEVICE : DAKC_OK;
  }
  if (st) {
    delete h;
    std::free(o.bin_offsets);
    o.bin_offsets = nullptr;
    drop_job_buffers(sl);
    fail(s, t, st);
    return;
  }
  o.n = d;
  o.total = n;
  o.key_words = t.plan.kw;
  o.k = t.plan.k;
  o.num_bins = t.plan.owner_digit ? t.plan.num_pes : 0;
  o.hash_seed = t.plan.seed;
  o.on_device = 1;
  h->arena = s.arena;
  h->resident = s.resident;
  o.priv = h;
  const size_t kb = (size_t)d * t.plan.kw * 8, cb = (size_t)d * 4;
  const size_t tb = al256(kb) + al256(cb);
  void *dst = d && tb <= TABLE_MOVE_MAX ? s.arena->alloc_top(tb) : nullptr;
  hipEvent_t ev = nullptr;
  if (dst) {
    if (!s.spare_events.empty()) {
      ev = s.spare_events.back();
      s.spare_events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      ev = nullptr;
    }
  }
  if (dst && ev &&
      hipMemcpyAsync(dst, sl.ws.loc.keys, kb, hipMemcpyDeviceToDevice, sl.stream) == hipSuccess &&
      hipMemcpyAsync((char *)dst + al256(kb), sl.ws.loc.counts, cb, hipMemcpyDeviceToDevice,
                     sl.stream) == hipSuccess &&
      hipEventRecord(ev, sl.stream) == hipSuccess) {
    o.keys = (uint64_t *)dst;
    o.counts = (uint32_t *)((char *)dst + al256(kb));
    h->parts[0] = {dst, tb};
    h->bytes = tb;
    t.copied = ev;
    sl.ws.detach_job_buffers(t.pending_free);
    s.resident->fetch_add(h->bytes);
    s.copying.push_back(id);
    return;
  }
  if (dst) {  // could not start the copy: keep the table in place instead
    hipStreamSynchronize(sl.stream);
    s.arena->release(dst, tb);
  }
  if (ev) s.spare_events.push_back(ev);
  o.keys = (uint64_t *)sl.ws.loc.keys;
  o.counts = sl.ws.loc.counts;
  // the kept ranges, widened to the pool's 256-byte granularity
  auto span = [](const void *p, size_t bytes) {
    const uintptr_t b = (uintptr_t)p & ~(uintptr_t)255;
    const uintptr_t e = ((uintptr_t)p + bytes + 255) & ~(uintptr_t)255;
    return std::make_pair((void *)b, (size_t)(bytes ? e - b : 0));
  };
  h->parts[0] = span(o.keys, kb);
  h->parts[1] = span(o.counts, cb);
  if (h->parts[1].second && h->parts[0].second &&
      (uintptr_t)h->parts[1].first < (uintptr_t)h->parts[0].first + h->parts[0].second &&
      (uintptr_t)h->parts[0].first < (uintptr_t)h->parts[1].first + h->parts[1].second) {
    // keys and counts share a 256-byte line: keep one range over both
    const uintptr_t b = std::min((uintptr_t)h->parts[0].first, (uintptr_t)h->parts[1].first);
    const uintptr_t e = std::max((uintptr_t)h->parts[0].first + h->parts[0].second,
                                 (uintptr_t)h->parts[1].first + h->parts[1].second);
    h->parts[0] = {(void *)b, (size_t)(e - b)};
    h->parts[1] = {nullptr, 0};
  }
  h->bytes = h->parts[0].second + h->parts[1].second;
  Ranges job;
  sl.ws.detach_job_buffers(job);
  release_except(*s.arena, job, h->parts, 2);
  s.resident->fetch_add(h->bytes);
  s.unclaimed += h->bytes;
  t.status = DAKC_OK;
  t.done = true;
  ++s.stats.completed;
  s.cv_done.notify_all();
}

static void scheduler(ServeState *s) {
  hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    bool progressed = false, blocked = false;
    // retire finished slots first: slots and memory free up for admission
    bool any_running = false;
    for (auto &sl : s->slots) {
      if (!sl->ticket) continue;
      if (hipEventQuery(sl->done_ev) != hipErrorNotReady) {
        retire(*s, *sl);
        progressed = true;
      }
      if (sl->ticket) any_running = true;
    }
    // finished copies: release the job's buffers, the ticket is done
    for (size_t i = 0; i < s->copying.size();) {
      Ticket &t = *s->tickets[s->copying[i]];
      if (hipEventQuery(t.copied) != hipErrorNotReady) {
        s->spare_events.push_back(t.copied);
        t.copied = nullptr;
        for (auto &r : t.pending_free) s->arena->release(r.first, r.second);
        t.pending_free.clear();
        s->unclaimed += ((DevTableHold *)t.table.priv)->bytes;
        t.status = DAKC_OK;
        t.done = true;
        ++s->stats.completed;
        s->cv_done.notify_all();
        s->copying[i] = s->copying.back();
        s->copying.pop_back();
        progressed = true;
      } else {
        ++i;
      }
    }
    // admit queued jobs in FIFO order into free slots while their buffers fit
    // the pool; a large job waits while another large one runs, and smaller
    // jobs behind it may go first meanwhile (a job that does not fit the
    // pool keeps its place and everything behind it waits).  Large jobs that
    // shared the GPU would finish together, later than one at a time would
    // finish the first: the latency tail of a burst.
    // A job goes before queued jobs only if the pool keeps room for the
    // largest of their workspaces beside the overtaker's workspace and table:
    // a client that claims tables in submission order then never waits on a
    // job that unclaimed tables of later jobs keep out of the pool.
    const int max_large = s->ctx->opts.max_large > 0 ? s->ctx->opts.max_large : 1;
    size_t passed = 0;  // largest workspace among the queued jobs passed over
    int large_running = 0;
    const bool pipelined = s->queue.size() > 4 * s->slots.size();
    for (auto &sl : s->slots) large_running += holds_large(*sl, pipelined);
    for (auto qi = s->queue.begin(); !s->stop && qi != s->queue.end();) {
      Slot *slp = nullptr;
      for (auto &sl : s->slots)
        if (!sl->ticket) {
          slp = sl.get();
          break;
        }
      if (!slp) break;
      Slot &sl = *slp;
      const uint64_t id = *qi;
      Ticket &t = *s->tickets[id];
      const bool large = is_large(t.plan, s->ctx->num_cus);
      if ((large && large_running >= max_large) ||
          (passed && s->arena->largest_free() < passed + job_pool_bound(t.plan))) {
        passed = std::max(passed, job_device_bytes(t.plan));
        ++qi;
        continue;
      }
      const auto h0 = std::chrono::steady_clock::now();
      int st = sl.ws.reserve(t.plan, !t.job.reads_on_device);
      if (!st)  // with stage events: ws.ev[2] tells when its final level starts
        st = enqueue_count(sl.ws, t.plan, &t.job, (const uint8_t *)t.job.reads, s->ctx->num_cus,
                           true);
      if (!st && hipEventRecord(sl.done_ev, sl.stream) != hipSuccess) st = DAKC_E_DEVICE;
      s->stats.admit_host_ms +=
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
      if (st == DAKC_E_NOMEM) {
        drop_job_buffers(sl);
        // wait for running jobs and copies, or for the client to claim finished
        // tables; when only tables it already holds are in the way, the job fails
        if (any_running || !s->copying.empty() || s->unclaimed) {
          blocked = true;
          break;
        }
        qi = s->queue.erase(qi);
        fail(*s, t, DAKC_E_NOMEM);
        progressed = true;
        continue;
      }
      qi = s->queue.erase(qi);
      if (st) {
        drop_job_buffers(sl);
        fail(*s, t, st);
        progressed = true;
        continue;
      }
      ++s->stats.admitted;
      sl.ticket = id;
      sl.large = large;
      sl.msd = t.plan.msd;
      large_running += large;
      any_running = true;
      progressed = true;
    }
    if (s->stop) {
      // shutting down: jobs not admitted yet are cancelled (a job waiting
      // for pool room that only unclaimed tables hold would otherwise wait
      // for ever); the loop ends once the running jobs and copies are done
      for (uint64_t id : s->queue) fail(*s, *s->tickets[id], DAKC_E_CANCELLED);
      s->queue.clear();
      if (!any_running && s->copying.empty()) break;
    }
    if (!progressed) {
      if (any_running || !s->copying.empty()) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
      } else if (blocked && !s->stop) {
        // only unclaimed tables hold the pool: look again when a wait claims one
        s->cv_sched.wait_for(lk, std::chrono::milliseconds(1));
      } else {
        s->cv_sched.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      }
    }
  }
}

static ServeState *serve_get(dakc_ctx *c, int *status) {
  std::lock_guard<std::mutex> g(c->serve_init_mu);
  if (c->serve) return c->serve;
  auto *s = new (std::nothrow) ServeState();
  if (!s) {
    *status = DAKC_E_NOMEM;
    return nullptr;
  }
  s->ctx = c;
  hipSetDevice(c->device);
  size_t free_b = 0, total_b = 0;
  hipMemGetInfo(&free_b, &total_b);
  const size_t budget = c->opts.hbm_budget ? c->opts.hbm_budget : (size_t)(free_b * 0.6);
  // one device allocation for the whole pool: job buffers and result tables
  // are cut from it on the host, so admission never goes back to the driver
  const auto t0 = std::chrono::steady_clock::now();
  s->arena = std::make_shared<DevArena>();
  if (int st = s->arena->init(c->device, budget)) {
    *status = st;
    delete s;
    return nullptr;
  }
  s->stats.init_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const int nslots = c->opts.max_inflight > 0 ? c->opts.max_inflight : 4;
  for (int i = 0; i < nslots; ++i) {
    auto sl = std::make_unique<Slot>();
    if (hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl->done_ev, hipEventDisableTiming) != hipSuccess ||
        sl->ws.init(sl->stream, s->arena.get())) {
      *status = DAKC_E_DEVICE;
      delete s;
      return nullptr;
    }
    sl->ws.fin_persistent = false;  // slots share the GPU: short final workgroups
    s->slots.push_back(std::move(sl));
  }
  s->sched = std::thread(scheduler, s);
  c->serve = s;
  return s;
}

void serve_shutdown(dakc_ctx *c) {
  ServeState *s = c->serve;
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_sched.notify_all();
  if (s->sched.joinable()) s->sched.join();
  for (auto &sl : s->slots)
    if (sl->stream) hipStreamSynchronize(sl->stream);
  for (auto &kv : s->tickets) {
    Ticket &t = *kv.second;
    if (t.copied) hipEventDestroy(t.copied);
    for (auto &r : t.pending_free) s->arena->release(r.first, r.second);
    dakc_table_free(&t.table);
  }
  for (hipEvent_t e : s->spare_events) hipEventDestroy(e);
  for (auto &sl : s->slots) {
    sl->ws.~JobWs();
    new (&sl->ws) JobWs();
    if (sl->done_ev) hipEventDestroy(sl->done_ev);
    if (sl->stream) hipStreamDestroy(sl->stream);
  }
  delete s;  // the pool itself lives on while result tables hold it
  c->serve = nullptr;
}

}  // namespace dk

using namespace dk;

extern "C" int dakc_submit(dakc_ctx *ctx, const dakc_job *job, uint64_t *ticket) {
  if (!ctx || !job || !ticket) return DAKC_E_INVAL;
  Plan pl;
  int st = make_plan(job, &pl);
  if (st) return st;
  ServeState *s = serve_get(ctx, &st);
  if (!s) return st;
  // the footprint of the path the job will take: a job whose MSD workspace
  // can never fit the pool takes the LSD path when that one fits
  apply_tuning(ctx, &pl, ctx->opts.algo);
  if (job_device_bytes(pl) > s->arena->size) {
    if (!pl.msd) return DAKC_E_REJECTED;
    Plan lsd = pl;
    plan_choose_path(&lsd, DAKC_ALGO_LSD);
    if (job_device_bytes(lsd) > s->arena->size) return DAKC_E_REJECTED;
    pl = lsd;
  }
  auto t = std::make_unique<Ticket>();
  t->job = *job;
  t->plan = pl;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return DAKC_E_CANCELLED;
    *ticket = s->next_ticket++;
    ++s->stats.submitted;
    s->tickets[*ticket] = std::move(t);
    s->queue.push_back(*ticket);
  }
  s->cv_sched.notify_all();
  return DAKC_OK;
}

extern "C" int dakc_poll(dakc_ctx *ctx, uint64_t ticket, int *done) {
  if (!ctx || !done) return DAKC_E_INVAL;
  ServeState *s = ctx->serve;
  if (!s) return DAKC_E_NOTFOUND;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return DAKC_E_NOTFOUND;
  *done = it->second->done ? 1 : 0;
  return DAKC_OK;
}

extern "C" int dakc_wait(dakc_ctx *ctx, uint64_t ticket, dakc_table *out) {
  if (!ctx || !out) return DAKC_E_INVAL;
  std::memset(out, 0, sizeof(*out));
  ServeState *s = ctx->serve;
  if (!s) return DAKC_E_NOTFOUND;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return DAKC_E_NOTFOUND;
  Ticket *t = it->second.get();
  s->cv_done.wait(lk, [&] { return t->done; });
  int st = t->status;
  dakc_table dev = t->table;
  const bool want_device = t->job.table_on_device != 0;
  const uint64_t tbytes = !st && dev.priv ? ((DevTableHold *)dev.priv)->bytes : 0;
  s->tickets.erase(it);
  if (st) return st;
  if (want_device) {
    // the client holds the table from here on
    s->unclaimed -= tbytes;
    lk.unlock();
    s->cv_sched.notify_all();
    *out = dev;
    return DAKC_OK;
  }
  lk.unlock();
  // host table: copied here, in the caller's thread, then the device copy goes
  hipSetDevice(ctx->device);
  dakc_table &o = *out;
  o = dev;
  o.on_device = 0;
  o.priv = nullptr;
  o.bin_offsets = nullptr;
  const size_t kb = (size_t)dev.n * dev.key_words * 8, cb = (size_t)dev.n * 4;
  o.keys = (uint64_t *)std::malloc(kb ? kb : 8);
  o.counts = (uint32_t *)std::malloc(cb ? cb : 4);
  if (!o.keys || !o.counts) st = DAKC_E_NOMEM;
  else if ((kb && hipMemcpy(o.keys, dev.keys, kb, hipMemcpyDeviceToHost) != hipSuccess) ||
           (cb && hipMemcpy(o.counts, dev.counts, cb, hipMemcpyDeviceToHost) != hipSuccess))
    st = DAKC_E_DEVICE;
  if (st) {
    dakc_table_free(&o);
  } else {
    o.bin_offsets = dev.bin_offsets;  // host memory: moves to the returned table
    dev.bin_offsets = nullptr;
  }
  dakc_table_free(&dev);
  {
    // only now is the device copy gone; until here it counted as unclaimed,
    // so the scheduler waited for it instead of failing a job
    std::lock_guard<std::mutex> g(s->mu);
    s->unclaimed -= tbytes;
  }
  s->cv_sched.notify_all();
  return st;
}

extern "C" int dakc_serve_get_stats(dakc_ctx *ctx, dakc_serve_stats *out) {
  if (!ctx || !out) return DAKC_E_INVAL;
  std::memset(out, 0, sizeof(*out));
  ServeState *s = ctx->serve;
  if (!s) return DAKC_OK;
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->stats;
  out->budget = s->arena->size;
  out->resident_bytes = s->resident->load();
  {
    std::lock_guard<std::mutex> ga(s->arena->mu);
    out->pool_used = s->arena->used;
    out->peak_device_bytes = s->arena->peak;
  }
  out->queued = s->queue.size();
  return DAKC_OK;
}
