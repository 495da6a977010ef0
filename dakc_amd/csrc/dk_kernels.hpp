// dk_kernels.hpp -- host-visible launch interface of the counting kernels.
// Internal to libdakc_amd; the public surface is include/dakc.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dk {

// Fixed-width record input (fqreader.cpp:43, kcounter_funcs.cpp:102): each
// record is L bytes followed by '\n'.
struct ReadsDesc {
  const uint8_t *reads;  // device pointer, 4-byte aligned
  uint64_t nbytes;
  uint64_t nrec;       // number of records
  uint32_t L;          // record width (READLEN)
  uint32_t stride;     // L + 1
  uint32_t k;
  uint32_t S;          // k-mer start slots per record: L - k + 1 (0 if L < k)
  uint32_t RT;         // records per tile
  uint32_t wpr;        // packed 32-base words per record in LDS (incl. 2 pad words)
  float invS;          // 1/S for slot -> (record, position)
  uint32_t canonical;
  uint32_t num_pes;    // P
  int32_t owner_bin;   // -1: all bins; >= 0: keep only k-mers owned by this bin
  uint32_t last_has_nl;  // 0 when the final record has no trailing '\n'
};

// error bits reported through the job's err word
enum : uint32_t { ERR_SEPARATOR = 1u, ERR_NUL_RECORD = 2u };

struct Workspace {
  uint64_t *hist;         // (MAX_PASSES) x 256 digit counts
  uint64_t *digit_start;  // (MAX_PASSES) x 256 exclusive starts
  uint64_t *n_dev;        // [0] = n k-mers, [1] = d distinct
  uint32_t *err;
  uint32_t *tickets;      // MAX_PASSES + 1 tile tickets
  uint64_t *flags;        // look-back flags: max_tiles x 256
  uint64_t *rle_flags;    // look-back flags of the run-length reduce: max_tiles
  void *keys_a, *keys_b;  // nmax keys each (8 or 16 B per key)
  uint32_t *counts;       // nmax
  uint64_t *bin_offsets;  // P + 1
};

constexpr int MAX_PASSES = 18;  // 16 digits of a 128-bit key + owner digit + spare

// Tile geometry of the radix passes (keys per tile = NT * ITEMS)
constexpr int SWEEP_NT = 512;
constexpr int SWEEP_ITEMS_1 = 15;  // 1 x u64 keys: 7680 per tile
constexpr int SWEEP_ITEMS_2 = 8;   // 2 x u64 keys: 4096 per tile
constexpr int RAW_CAP = 16384;     // LDS bytes for one tile of raw records
constexpr int MAX_READ_LEN = 12000;

inline int tile_keys(int kw) { return SWEEP_NT * (kw == 1 ? SWEEP_ITEMS_1 : SWEEP_ITEMS_2); }

// Launchers (all asynchronous on `st`).
void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st);
void launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t st);
void launch_hist_reads(int kw, const ReadsDesc &rd, int ndig, int owner_digit, const Workspace &ws,
                       uint32_t ntiles, int grid, hipStream_t st);
void launch_scan_digits(int npasses, const Workspace &ws, hipStream_t st);
// one LSD pass; pass 0 reads the records, later passes read keys
void launch_sweep(int kw, bool from_reads, const ReadsDesc &rd, const void *in, void *out,
                  uint32_t shift, bool owner_pass, uint32_t num_pes, const uint64_t *digit_start,
                  const Workspace &ws, uint32_t ticket_slot, uint32_t epoch, uint32_t grid,
                  hipStream_t st);
void launch_rle(int kw, const void *sorted, void *out_keys, const Workspace &ws, uint32_t epoch,
                uint32_t grid, hipStream_t st);
void launch_bin_offsets(int kw, const void *table_keys, uint32_t num_pes, const Workspace &ws,
                        hipStream_t st);

}  // namespace dk
