// dk_api_internal.hpp -- host-side state behind the C ABI (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/dakc.h"
#include "dk_kernels.hpp"

namespace dk {

// Grow-only device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
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
  int ndig = 0, owner_digit = 0, npasses = 0;
  uint32_t canonical = 0, last_has_nl = 1;
  // MSD path (msd = 1): first-level bits b1 (with the owner when all bins
  // are counted), second-level b2, in-LDS b3; L1 tiles of RT1 records
  int msd = 0;
  uint32_t b1 = 0, b2 = 0, b3 = 0, nb1 = 1, nd2 = 1, nsb = 1;
  uint32_t RT1 = 1, wpr1 = 0, ntiles1 = 0, tcap1 = 0, G = 1, ngroups = 0, cap2 = 0, fin_nt = 256;
};

// internal status: an MSD sub-bucket did not fit in LDS; rerun on the LSD path
constexpr int STATUS_MSD_OVERFLOW = -100;

// path choice (dakc_opts.algo)
int plan_choose_path(Plan *pl, int algo);

// Device workspace of one in-flight job (one per served slot).
struct JobWs {
  DevBuf hist, dstart, ndev, err, tickets, flags, rle_flags, keys_a, keys_b, counts, binoff, reads,
      msd_off1, msd_off2;
  hipStream_t stream = nullptr;
  hipEvent_t ev[MAX_PASSES + 4] = {};
  uint32_t epoch = 0;
  ~JobWs();
  int init(hipStream_t st);
  int reserve(const Plan &pl, bool need_reads);
  uint32_t next_epoch();
  Workspace view() const;
};

int make_plan(const dakc_job *job, Plan *pl);
size_t job_device_bytes(const Plan &pl);
int enqueue_count(JobWs &ws, const Plan &pl, const uint8_t *dreads, int num_cus, bool events,
                  void **table_keys);
int finish_count(JobWs &ws, const Plan &pl, uint64_t *n, uint64_t *d);

struct ServeState;  // served mode (serve.hip)

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
};

namespace dk {
int batch_count(dakc_ctx *c, const dakc_job *jobs, size_t njobs, dakc_table *outs);
void serve_shutdown(dakc_ctx *c);
}  // namespace dk
