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

struct Ticket {
  dakc_job job;
  Plan plan;
  int status = DAKC_E_PENDING;
  bool done = false;
  dakc_table table = {};
  int slot = -1;
};

struct Slot {
  JobWs ws;
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;
  uint64_t ticket = 0;  // 0 = free
};

struct ServeState {
  dakc_ctx *ctx = nullptr;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done;
  std::deque<uint64_t> queue;
  std::map<uint64_t, std::unique_ptr<Ticket>> tickets;
  std::vector<std::unique_ptr<Slot>> slots;
  uint64_t next_ticket = 1;
  size_t budget = 0;
  bool stop = false;
  std::thread sched;
};

static size_t slots_footprint(const ServeState &s) {
  size_t b = 0;
  for (auto &sl : s.slots)
    for (const DevBuf *d : {&sl->ws.hist, &sl->ws.dstart, &sl->ws.ndev, &sl->ws.err, &sl->ws.tickets,
                            &sl->ws.flags, &sl->ws.rle_flags, &sl->ws.keys_a, &sl->ws.keys_b,
                            &sl->ws.counts, &sl->ws.binoff, &sl->ws.reads, &sl->ws.msd_off1,
                            &sl->ws.msd_off2, &sl->ws.arena, &sl->ws.meta, &sl->ws.desc})
      b += d->cap;
  return b;
}

static size_t slot_growth(const Slot &sl, const Plan &pl, bool need_reads) {
  // bytes the slot's grow-only buffers would add for this plan (upper bound)
  const size_t need = job_device_bytes(pl) - (need_reads ? 0 : pl.nbytes);
  size_t have = 0;
  for (const DevBuf *d : {&sl.ws.flags, &sl.ws.rle_flags, &sl.ws.keys_a, &sl.ws.keys_b,
                          &sl.ws.counts, &sl.ws.reads, &sl.ws.msd_off2, &sl.ws.arena})
    have += d->cap;
  return need > have ? need - have : 0;
}

// copy a finished slot's table to host memory and release the slot
static void complete(ServeState &s, Slot &sl) {
  Ticket &t = *s.tickets[sl.ticket];
  uint64_t n = 0, d = 0;
  int st = finish_count(sl.ws, t.plan, &n, &d);
  if (st == STATUS_MSD_OVERFLOW && t.plan.msd) {
    // a sub-bucket did not fit in LDS: count this job again on the LSD path
    plan_choose_path(&t.plan, DAKC_ALGO_LSD);
    st = sl.ws.reserve(t.plan, !t.job.reads_on_device);
    if (!st)
      st = enqueue_count(sl.ws, t.plan, &t.job, (const uint8_t *)t.job.reads, s.ctx->num_cus, false);
    if (!st && hipStreamSynchronize(sl.stream) != hipSuccess) st = DAKC_E_DEVICE;
    if (!st) st = finish_count(sl.ws, t.plan, &n, &d);
  }
  if (!st) {
    dakc_table &o = t.table;
    o.n = d;
    o.total = n;
    o.key_words = t.plan.kw;
    o.k = t.plan.k;
    o.num_bins = t.plan.owner_digit ? t.plan.num_pes : 0;
    const size_t kb = (size_t)d * t.plan.kw * 8;
    o.keys = (uint64_t *)std::malloc(kb ? kb : 8);
    o.counts = (uint32_t *)std::malloc(d ? d * 4 : 4);
    if (t.plan.owner_digit) o.bin_offsets = (uint64_t *)std::calloc(t.plan.num_pes + 1, 8);
    if (!o.keys || !o.counts || (t.plan.owner_digit && !o.bin_offsets)) {
      st = DAKC_E_NOMEM;
    } else if (d) {
      if (hipMemcpyAsync(o.keys, sl.ws.loc.keys, kb, hipMemcpyDeviceToHost, sl.stream) != hipSuccess ||
          hipMemcpyAsync(o.counts, sl.ws.loc.counts, d * 4, hipMemcpyDeviceToHost, sl.stream) !=
              hipSuccess ||
          (t.plan.owner_digit &&
           hipMemcpyAsync(o.bin_offsets, sl.ws.loc.binoff, (t.plan.num_pes + 1) * 8,
                          hipMemcpyDeviceToHost, sl.stream) != hipSuccess) ||
          hipStreamSynchronize(sl.stream) != hipSuccess)
        st = DAKC_E_DEVICE;
    }
    if (st) dakc_table_free(&o);
  }
  t.status = st;
  t.done = true;
  t.slot = -1;
  sl.ticket = 0;
}

static void scheduler(ServeState *s) {
  hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    // admit: first queued job into a free slot if the budget allows (FIFO)
    bool progressed = false;
    while (!s->queue.empty()) {
      Ticket &t = *s->tickets[s->queue.front()];
      Slot *free_slot = nullptr;
      size_t best = (size_t)-1;
      const bool need_reads = !t.job.reads_on_device;
      for (auto &sl : s->slots)
        if (!sl->ticket) {
          const size_t g = slot_growth(*sl, t.plan, need_reads);
          if (g < best) best = g, free_slot = sl.get();
        }
      if (!free_slot) break;
      if (slots_footprint(*s) + best > s->budget) {
        // shrink idle slots to make room before giving up
        for (auto &sl : s->slots)
          if (!sl->ticket && sl.get() != free_slot) {
            sl->ws.keys_a.release();
            sl->ws.keys_b.release();
            sl->ws.counts.release();
            sl->ws.reads.release();
            sl->ws.flags.release();
            sl->ws.rle_flags.release();
            sl->ws.msd_off2.release();
            sl->ws.arena.release();
          }
        if (slots_footprint(*s) + best > s->budget) break;  // wait for running jobs
      }
      const uint64_t id = s->queue.front();
      s->queue.pop_front();
      Slot &sl = *free_slot;
      int st = sl.ws.reserve(t.plan, need_reads);
      if (!st)
        st = enqueue_count(sl.ws, t.plan, &t.job, (const uint8_t *)t.job.reads, s->ctx->num_cus,
                           false);
      if (!st && hipEventRecord(sl.done_ev, sl.stream) != hipSuccess) st = DAKC_E_DEVICE;
      if (st) {
        t.status = st;
        t.done = true;
        s->cv_done.notify_all();
        continue;
      }
      sl.ticket = id;
      t.slot = 1;
      progressed = true;
    }
    // retire finished slots
    bool any_running = false;
    for (auto &sl : s->slots) {
      if (!sl->ticket) continue;
      hipError_t q = hipEventQuery(sl->done_ev);
      if (q != hipErrorNotReady) {
        complete(*s, *sl);
        progressed = true;
        s->cv_done.notify_all();
      } else {
        any_running = true;
      }
    }
    if (s->stop && s->queue.empty() && !any_running) break;
    if (!progressed) {
      if (any_running) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
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
  s->budget = c->opts.hbm_budget ? c->opts.hbm_budget : (size_t)(free_b * 0.6);
  const int nslots = c->opts.max_inflight > 0 ? c->opts.max_inflight : 4;
  for (int i = 0; i < nslots; ++i) {
    auto sl = std::make_unique<Slot>();
    if (hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl->done_ev, hipEventDisableTiming) != hipSuccess ||
        sl->ws.init(sl->stream)) {
      *status = DAKC_E_DEVICE;
      delete s;
      return nullptr;
    }
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
  for (auto &kv : s->tickets) dakc_table_free(&kv.second->table);
  for (auto &sl : s->slots) {
    if (sl->stream) hipStreamSynchronize(sl->stream);
    sl->ws.~JobWs();
    new (&sl->ws) JobWs();
    if (sl->done_ev) hipEventDestroy(sl->done_ev);
    if (sl->stream) hipStreamDestroy(sl->stream);
  }
  delete s;
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
  if (job_device_bytes(pl) > s->budget) return DAKC_E_REJECTED;
  apply_tuning(ctx, &pl, ctx->opts.algo);
  auto t = std::make_unique<Ticket>();
  t->job = *job;
  t->plan = pl;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return DAKC_E_CANCELLED;
    *ticket = s->next_ticket++;
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
  const int st = t->status;
  if (!st) *out = t->table;
  s->tickets.erase(it);
  return st;
}
