// dk_status.hpp -- the error bits the kernels leave in a job's error word and
// the status the host reports for them.  Plain C++ (no HIP), so the mapping
// is also compiled and checked on its own by the CPU tests.
#pragma once
#include <stdint.h>

#include "../../include/dakc.h"

namespace dk {

enum : uint32_t {
  ERR_SEPARATOR = 1u,   // a record's separator byte is not '\n'
  ERR_NUL_RECORD = 2u,  // a record starts with NUL
  ERR_OVERFLOW = 4u,    // an MSD item, group or bin did not fit in LDS (recount on the LSD path)
  ERR_STALL = 8u,       // a decoupled look-back wait gave up (bounded polls): results are invalid
};

// internal status: an MSD sub-bucket did not fit in LDS; rerun on the LSD path
constexpr int STATUS_MSD_OVERFLOW = -100;

// A wait that gave up voids everything else the launch computed, so it is
// reported first; a malformed input is reported before an overflow, which
// would only lead to a recount of the same malformed input.
inline int status_of_err(uint32_t e) {
  if (e & ERR_STALL) return DAKC_E_DEVICE;
  if (e & (ERR_SEPARATOR | ERR_NUL_RECORD)) return DAKC_E_FORMAT;
  if (e & ERR_OVERFLOW) return STATUS_MSD_OVERFLOW;
  return DAKC_OK;
}

}  // namespace dk
