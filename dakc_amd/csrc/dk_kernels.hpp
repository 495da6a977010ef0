// dk_kernels.hpp -- host-visible launch interface of the counting kernels.
// Internal to libdakc_amd; the public surface is include/dakc.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dk_status.hpp"

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
  uint64_t seed;         // owner-hash seed
};

// error bits reported through the job's err word: dk_status.hpp

// Which digit of a key a pass sorts or partitions by.
//   DIG_BITS : bits [shift, shift + nbits) of the 2k-bit k-mer value
//   DIG_OWNER: MurmurHash64A owner % num_pes
//   DIG_MSD  : owner * 2^nbits + bits [shift, shift + nbits) (owner part only
//              when num_pes > 1), so buckets come out in (owner, prefix) order
enum : uint32_t { DIG_BITS = 0, DIG_OWNER = 1, DIG_MSD = 2 };
struct DigitSpec {
  uint32_t mode;
  uint32_t shift;
  uint32_t nbits;
  uint32_t num_pes;
};

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
  uint64_t key_cap;       // keys each of keys_a / keys_b holds
};

constexpr int MAX_PASSES = 18;  // 16 digits of a 128-bit key + owner digit + spare
#define MAX_PASSES_DECL 18

// Tile geometry of the radix passes (keys per tile = NT * ITEMS)
constexpr int SWEEP_NT = 512;
constexpr int SWEEP_ITEMS_1 = 15;  // 1 x u64 keys: 7680 per tile
constexpr int SWEEP_ITEMS_2 = 8;   // 2 x u64 keys: 4096 per tile
constexpr int RAW_CAP = 16384;     // LDS bytes for one tile of raw records
constexpr int MAX_READ_LEN = 4096;  // keeps L - k + 1 <= keys per tile

inline int tile_keys(int kw) { return SWEEP_NT * (kw == 1 ? SWEEP_ITEMS_1 : SWEEP_ITEMS_2); }

// MSD levels 1 and 2: one 1024-thread workgroup per CU handles a tile of
// 15360 u64 (8192 2x u64) keys in LDS.  Final level: 256-thread workgroups,
// four per CU, each holding a sub-bucket of up to 3840 u64 (2048 2x u64)
// keys, 512-thread workgroups, two per CU, holding twice that, or one
// 1024-thread workgroup per CU holding four times that.
constexpr int MSD_NT = 1024;
constexpr int MSD_ITEMS_1 = 15;
constexpr int MSD_ITEMS_2 = 8;
inline int msd_tile(int kw) { return MSD_NT * (kw == 1 ? MSD_ITEMS_1 : MSD_ITEMS_2); }
// L2 workgroups (NT threads; 1024 measured faster than two 512s per CU) and
// the keys one holds
constexpr int L2_NT = 1024;
inline int l2_cap(int kw) { return L2_NT * (kw == 1 ? MSD_ITEMS_1 : MSD_ITEMS_2); }
constexpr int L1_RAW = 20480;    // LDS bytes of raw records per L1 tile
constexpr int L1_CODEW = 1792;   // packed code words per L1 tile
constexpr int L1_MAXRT = 256;    // records per L1 tile
constexpr int FIN_NT = 256;
constexpr int FIN_ITEMS_1 = 15;
constexpr int FIN_ITEMS_2 = 8;
constexpr int FIN_MAX_GROUPS = 1024;  // L2 groups per bucket the final level can gather
inline int fin_cap(int kw, int nt) { return nt * (kw == 1 ? FIN_ITEMS_1 : FIN_ITEMS_2); }

// Launchers (all asynchronous on `st`).
void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st);
void launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t st);
// copy of n 32-bit words; dst and src 16-byte aligned
void launch_copy_u32(uint32_t *dst, const uint32_t *src, uint64_t n, hipStream_t st);
struct HistSpecs {
  DigitSpec d[MAX_PASSES_DECL];
  int n;
};
void launch_hist_reads(int kw, const ReadsDesc &rd, const HistSpecs &hs, const Workspace &ws,
                       uint32_t ntiles, int grid, hipStream_t st);
void launch_scan_digits(int npasses, const Workspace &ws, hipStream_t st);
// one LSD pass; pass 0 reads the records, later passes read keys
void launch_sweep(int kw, bool from_reads, const ReadsDesc &rd, const void *in, void *out,
                  const DigitSpec &dig, const uint64_t *digit_start, const Workspace &ws,
                  uint32_t ticket_slot, uint32_t epoch, uint32_t grid, hipStream_t st);

// MSD path (three tile-local levels, no global histogram):
//   k_l1  records -> k-mers in first-level bucket order, per tile, with the
//         tile's bucket starts (off1[bucket][tile])
//   k_l2  bucket b, group of G consecutive L1 tiles -> keys in second-level
//         digit order, with the digit starts (off2[b][group][digit])
//   k_fin sub-bucket (b, digit) gathered from b's groups into LDS, sorted,
//         run-length reduced, written in place by the decoupled look-back
struct MsdArgs {
  DigitSpec d1, d2, d3;
  uint32_t nb1;      // first-level buckets (owner x 2^b1)
  uint32_t nd2;      // second-level digits (2^b2)
  uint32_t ntiles1;  // L1 tiles
  uint32_t tcap1;    // keys capacity of one L1 tile (records per tile * slots)
  uint32_t G;        // L1 tiles per L2 group
  uint32_t ngroups;  // L2 groups per bucket
  uint32_t cap2;     // keys capacity of one (bucket, group)
  uint32_t fin_nt;   // threads of a final-level workgroup (256, 512 or 1024)
  // Final items, in bucket order.  An owner block's 2^b1 buckets form 2^cl
  // classes of 2^bl consecutive buckets each (ranges of the top bits); class
  // c's items take pack_c second-level digits, ipb_c items per bucket.  Kept
  // in four words (few scalar registers in the final level's loop):
  uint32_t cls;    // cl | bl << 4
  uint32_t packc;  // pack_c in byte c
  uint32_t ipbc[2];  // ipb_c = ceil(nd2 / pack_c) in halfword c (0: class unused)
  uint16_t *off1;    // [nb1 + 1][ntiles1]
  uint16_t *off2;    // [nb1][ngroups][nd2 + 1]
};

// A final item: bucket b, second-level digits [s0, s0 + cnt) of it.  Its
// keys go to LDS bin (digit - s0) << b3 | the next b3 bits, which orders
// them as the keys order.
struct FinItem {
  uint32_t b, s0, cnt, nbins;
};
__host__ __device__ inline FinItem fin_item_of(const MsdArgs &m, uint32_t it) {
  const uint32_t cl = m.cls & 15, bl = (m.cls >> 4) & 15;
  const uint32_t i0 = m.ipbc[0] & 0xFFFFu, i1 = m.ipbc[0] >> 16, i2 = m.ipbc[1] & 0xFFFFu, i3 = m.ipbc[1] >> 16;
  // items of an owner block before classes 1, 2, 3 and in all of it
  const uint32_t e0 = i0 << bl, e1 = e0 + (i1 << bl), e2 = e1 + (i2 << bl), e3 = e2 + (i3 << bl);
  const uint32_t blk = it / e3;
  uint32_t r = it - blk * e3;
  // classes in order; an unused class has no items and is never chosen
  const uint32_t c = (r >= e0 ? 1u : 0u) + (r >= e1 ? 1u : 0u) + (r >= e2 ? 1u : 0u);
  r -= c == 0 ? 0u : c == 1 ? e0 : c == 2 ? e1 : e2;
  const uint32_t ipb = c == 0 ? i0 : c == 1 ? i1 : c == 2 ? i2 : i3;
  const uint32_t pack = (m.packc >> (8 * c)) & 0xFFu;
  const uint32_t bi = r / ipb;
  FinItem f;
  f.b = (((blk << cl) + c) << bl) + bi;
  f.s0 = (r - bi * ipb) * pack;
  f.cnt = m.nd2 - f.s0 < pack ? m.nd2 - f.s0 : pack;
  f.nbins = f.cnt << m.d3.nbits;
  return f;
}

// One job of a multi-job MSD launch.  Work items of every stage are numbered
// across the jobs of a launch (job j owns items [pre[j], pre[j+1])).
struct JobDesc {
  ReadsDesc rd;       // records; RT / wpr are the L1 tile geometry
  MsdArgs m;
  void *out1;         // L1 tiles (tile * tcap1 keys each); later the table keys
  void *out2;         // L2 regions ((b * ngroups + g) * cap2 keys each)
  uint32_t *tcounts;  // table counts
  void *tkeys;        // table keys: out1, or memory the table owns (batched device tables)
  uint64_t *nd;       // [0] k-mers, [1] distinct
  uint64_t *flags;    // nsb look-back words, zeroed before the launch
  uint32_t *err;      // this job's error bits
  uint32_t pad;
};
struct BatchArgs {
  const JobDesc *jobs;
  const uint32_t *pre1, *pre2, *pre3;  // njobs + 1 prefixes of L1 / L2 / final items
  uint32_t njobs;
  uint32_t *ticket;                    // final-level ticket of this launch
  uint32_t item0;                      // first final-level item of this launch
  uint32_t nfin;                       // final-level items of this launch
  uint32_t persist;                    // one-job final level: 1 persistent workgroups, 0 one item each
};
// launch the MSD levels for every job of `a` (n1 L1 items, n2 L2 items, n3
// final items from a.item0 on; a zero count skips the level).  kw is shared
// by all jobs, fin_nt by the jobs whose final items the launch covers.
// one: the descriptor of a single-job launch (passed as a kernel argument), or
// null when the launch holds several jobs (read from a.jobs)
void launch_msd_batch(int kw, uint32_t fin_nt, const BatchArgs &a, const JobDesc *one, uint32_t n1,
                      uint32_t n2, uint32_t n3, hipStream_t st);
void launch_rle(int kw, const void *sorted, void *out_keys, const Workspace &ws, uint32_t epoch,
                uint32_t grid, hipStream_t st);
// the first MSD level over a key array (one.rd.reads: one.rd.nrec keys), one
// workgroup per tile of msd_tile(kw) keys
void launch_l1_keys(int kw, const JobDesc &one, uint32_t ntiles, hipStream_t st);
// the LSD digit histograms of a key array
void launch_hist_keys(int kw, const void *keys, uint64_t n, const HistSpecs &hs, uint64_t seed, uint64_t *hist,
                      int grid, hipStream_t st);
void launch_bin_offsets(int kw, const void *table_keys, uint32_t num_pes, uint64_t seed, const Workspace &ws,
                        hipStream_t st);
void launch_bin_offsets_at(int kw, const void *table_keys, const uint64_t *nd, uint32_t num_pes, uint64_t seed,
                           uint64_t *offsets, hipStream_t st);

}  // namespace dk
