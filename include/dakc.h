/*
 * dakc.h -- C ABI of libdakc_amd: DAKC's k-mer counting hot path on one MI355X.
 *
 * The reference (Souvadra/dakc) has no plugin API or FFI; its seam is the C++
 * constructor
 *     kmercounter(char* read_chunk, std::vector<kmer_t>& vectordbg)
 *     (dakc/src/kcounter/kcounter.hpp:84-99, called from dakc/src/main/main.cpp:42-49)
 * which runs the whole count via perform_kcount() (dakc/src/kcounter/kcounter.cpp:247-507)
 * with k, READLEN and the PE count fixed at compile time (dakc/Makefile:11,
 * common.hpp:19-52).  These entry points replace that seam with plain
 * pointers and sizes: read buffer in, k at run time, sorted (kmer, count)
 * table out.  Errors are status codes; no C++ exception crosses this ABI.
 *
 * Semantics (bit-exact with a single-PE reference run):
 *   - input: records of exactly read_len bytes followed by '\n'
 *     (fqreader.cpp:43, kcounter_funcs.cpp:102); the last '\n' may be missing;
 *   - decode: C/c=0 A/a=1 T/t=2 G/g=3; M/m ends the read; any other byte
 *     splits it (common.hpp:72-92, kcounter_funcs.cpp:86-99);
 *   - k-mers: forward strand, first base in the highest bits, 2k bits
 *     (kcounter_funcs.cpp:25-67); canonical = min(kmer, reverse complement)
 *     on request (no reference counterpart);
 *   - owner: MurmurHash64A(kmer, seed) % num_pes with the reference's seed
 *     0x9E3779B97F4A7C15 unless the job names another (kcounter.cpp:29-54);
 *     for 2x64-bit keys MurmurHash64A over lo||hi;
 *   - output: ascending by k-mer (within each owner bin), no duplicates,
 *     counts >= 1 (kcounter.cpp:385-411).
 */
#ifndef DAKC_H
#define DAKC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DAKC_ABI_VERSION 2

typedef enum dakc_status {
  DAKC_OK = 0,
  DAKC_E_INVAL = 1,     /* bad argument (k out of 1..64, read_len, owner bin, NULL) */
  DAKC_E_FORMAT = 2,    /* buffer is not whole records / separator not '\n' / NUL record */
  DAKC_E_NOMEM = 3,     /* host or device allocation failed */
  DAKC_E_TOOBIG = 4,    /* job has 2^32 or more k-mer slots (counts are u32) */
  DAKC_E_DEVICE = 5,    /* HIP runtime error */
  DAKC_E_NODEVICE = 6,  /* no usable GPU */
  DAKC_E_REJECTED = 7,  /* served mode: job can never fit the HBM budget */
  DAKC_E_NOTFOUND = 8,  /* unknown ticket */
  DAKC_E_PENDING = 9,   /* served mode: job not finished yet (dakc_poll) */
  DAKC_E_CANCELLED = 10 /* served mode: job was cancelled */
} dakc_status;

/* read buffer formats */
#define DAKC_FMT_FIXED 0 /* read_len bytes + '\n' per record (the reference's .txt) */
#define DAKC_FMT_LINES 1 /* one read per '\n'-ended line, any lengths */
#define DAKC_FMT_FASTQ 2 /* 4-line FASTQ records; the second line of each is the read */

typedef struct dakc_job {
  const char *reads;   /* host pointer, or device pointer when reads_on_device */
  uint64_t nbytes;
  int32_t format;      /* DAKC_FMT_*; LINES and FASTQ are converted on the device
                        * (dakc_convert_reads) by dakc_count / dakc_count_device
                        * only; batched and served jobs take FIXED */
  int32_t read_len;    /* READLEN: record width in bases (1..4096); for LINES /
                        * FASTQ 0 = the longest read */
  int32_t k;           /* KMERLEN: 1..64 (k <= 32: one u64 per key, else two) */
  int32_t canonical;   /* 0: forward strand (the reference), 1: min(x, revcomp x) */
  uint32_t num_pes;    /* P; 1 = single PE */
  int32_t owner_bin;   /* -1: all P owner bins; b >= 0: only k-mers owned by PE b */
  int32_t reads_on_device; /* reads is a device pointer (4-byte aligned) */
  int32_t table_on_device; /* served mode (dakc_wait) and dakc_count_batch: the
                            * table is returned in device memory owned by the
                            * table (keys/counts are device pointers;
                            * dakc_table_free releases it) */
  uint64_t hash_seed;  /* seed of the owner hash; 0 = the reference's
                        * 0x9E3779B97F4A7C15 (kcounter.cpp:52) */
} dakc_job;

/* The (kmer, count) table.  keys holds n * key_words u64 (lo, hi per key for
 * key_words == 2); counts holds n u32.  With num_pes > 1 and owner_bin == -1
 * the table is the concatenation of the P owner tables and bin_offsets (P + 1
 * entries) marks their boundaries; otherwise bin_offsets is NULL.
 * Memory is owned by the library: release with dakc_table_free.  A device
 * table from dakc_count_device lives in the context's workspace (valid until
 * the next count on it; dakc_table_free only forgets it); one from a served
 * or batched job with table_on_device owns its device memory until
 * dakc_table_free. */
typedef struct dakc_table {
  uint64_t *keys;
  uint32_t *counts;
  uint64_t n;           /* distinct k-mers */
  uint64_t total;       /* k-mers counted (sum of counts) */
  uint64_t *bin_offsets;
  uint32_t num_bins;
  int32_t key_words;
  int32_t k;
  int32_t on_device;    /* 1: keys/counts are device pointers (dakc_count_device) */
  void *priv;
  uint64_t hash_seed;   /* owner-hash seed the bins were formed with (queries use it) */
} dakc_table;

typedef struct dakc_ctx dakc_ctx;

#define DAKC_ALGO_AUTO 0 /* MSD partition + in-LDS count, LSD radix sort when a bucket overflows */
#define DAKC_ALGO_LSD 1  /* LSD radix sort of all k-mers, then run-length reduce */
#define DAKC_ALGO_MSD 2  /* same as AUTO (the LSD rerun still guards skewed inputs) */

typedef struct dakc_opts {
  uint64_t hbm_budget;  /* served mode: device bytes the job pool may hold (0 = 60% of free) */
  int32_t max_inflight; /* served mode: concurrent jobs on the device (0 = 4) */
  int32_t algo;         /* DAKC_ALGO_* */
  int32_t max_large;    /* served mode: jobs that fill the GPU by themselves (first level
                         * >= 16 waves of workgroups) in flight at once (0 = 1) */
  int32_t batch_solo_items; /* batched: a job with at least this many final-level work items
                             * (sub-buckets) gets a final-level launch of its own instead of
                             * sharing the work-queue launch of the smaller jobs (0 = 32768) */
} dakc_opts;

/* Stage timings of the last dakc_count/dakc_count_device on a context,
 * measured with HIP events on the context's stream (milliseconds). */
typedef struct dakc_stats {
  double hist_ms, pass0_ms, passes_ms, rle_ms, total_ms;
  uint32_t npasses;     /* radix passes executed (including pass 0) */
  uint32_t tiles0;      /* tiles of pass 0 */
  uint64_t n;           /* k-mers */
  uint64_t d;           /* distinct */
  double sweep_ms[18];  /* per radix pass (LSD path) */
  uint32_t path;        /* 1: LSD radix sort, 2: MSD partition + in-LDS count */
  uint32_t msd_reruns;  /* 1 when an MSD bucket overflowed LDS and the LSD path reran */
  double part2_ms;      /* MSD: second-level tile partition */
  double final_ms;      /* MSD: gather + in-LDS sort + run-length reduce */
} dakc_stats;

const char *dakc_status_str(int status);
int dakc_abi_version(void);

int dakc_ctx_create(int device, const dakc_opts *opts, dakc_ctx **out);
int dakc_ctx_destroy(dakc_ctx *ctx);

/* Plain count: one job, table copied to host memory. */
int dakc_count(dakc_ctx *ctx, const dakc_job *job, dakc_table *out);
/* Same, table left in device memory owned by the context (valid until the
 * next call on ctx or dakc_table_free). */
int dakc_count_device(dakc_ctx *ctx, const dakc_job *job, dakc_table *out);
int dakc_last_stats(dakc_ctx *ctx, dakc_stats *out);

/* The geometry a job would be counted with on a GPU of num_cus compute units
 * (0 = 256), worked out on the host with no device call: for diagnostics and
 * tuning (no counterpart in the reference, whose geometry is fixed at
 * compile time).  algo is a DAKC_ALGO_*; the path the count takes can still
 * change at run time (an LSD recount after an LDS overflow). */
typedef struct dakc_plan_info {
  uint32_t path;                 /* 1: LSD radix sort, 2: MSD partition + in-LDS count */
  uint32_t key_words;            /* 1 (k <= 32) or 2 */
  uint64_t kmer_slots;           /* window slots of the records: a bound on k-mers and distinct */
  uint64_t device_bytes;         /* device workspace the count holds */
  uint32_t b1, b2, b3;           /* MSD: key bits split by the first, second and final level */
  uint32_t first_level_buckets;  /* MSD: owners x 2^b1 */
  uint32_t l1_tiles, l2_groups;  /* MSD: first-level tiles, second-level groups per bucket */
  uint32_t final_items;          /* MSD: final-level work items */
  uint32_t final_threads;        /* MSD: threads of a final-level workgroup */
  uint32_t radix_passes;         /* LSD: digit passes (with the owner digit) */
} dakc_plan_info;
int dakc_plan_job(const dakc_job *job, int algo, uint32_t num_cus, dakc_plan_info *out);

/* ------------------------------------------- the two phases of a P-PE run --
 * dakc_count does both phases of one PE's count at once.  These run them
 * apart, as a P-PE run on several GPUs would (each GPU partitions its reads,
 * an exchange ships bin o to PE o, each PE counts what it received). */

/* Phase 1, the send side: every k-mer of the reads routed to its owner PE
 * (kcounter.cpp:146-174, add_in_normal_packet / empty_packets, with
 * owner_pe kcounter.cpp:50-54).  keys_out (device, cap keys of key_words
 * u64) receives the owner bins one after another, each bin's k-mers in
 * record order; bin_offsets_out (host) receives num_pes + 1 boundaries
 * (all bins), or 2 when one bin is kept (owner_bin >= 0) or num_pes == 1.
 * *n_out = k-mers routed; more than cap: DAKC_E_NOMEM, nothing written.
 * The job's reads may be host or device memory (DAKC_FMT_FIXED). */
int dakc_partition(dakc_ctx *ctx, const dakc_job *job, uint64_t *keys_out, uint64_t cap,
                   uint64_t *bin_offsets_out, uint64_t *n_out);

/* Phase 2, the owner side: the sorted (kmer, count) table of an array of n
 * keys of k bases (key_words u64 each, in any order), as the owner sorts and
 * run-length reduces the k-mers it received (kcounter.cpp:123-144 recv_kmer
 * into vectordbg, :385-411).  With num_pes > 1 the table is split into the
 * P owner bins (hash_seed as in dakc_job).  keys are device memory when
 * keys_on_device (8-byte aligned), else host memory; they are not changed.
 * table_on_device: the table stays in the context's workspace as with
 * dakc_count_device, else it is copied to host memory as with dakc_count. */
int dakc_count_keys(dakc_ctx *ctx, const uint64_t *keys, uint64_t n, int k, uint32_t num_pes,
                    uint64_t hash_seed, int keys_on_device, int table_on_device, dakc_table *out);

/* HITTER=1 (kcounter.cpp:56-104, 176-245, 325-384; off in the shipped build,
 * dakc/Makefile:11): the send side pre-aggregates.  Records go into a bucket
 * while it holds at most bucket_size - read_len k-mers (read_till_buf_max,
 * kcounter_funcs.cpp:69-111; 0 = the reference's KCOUNT_BUCKET_SIZE, 10000;
 * read_len < bucket_size <= 10240).  A k-mer seen once or twice in its
 * bucket is sent as that many plain k-mers, one seen three or more times as
 * one (k-mer, count) pair.  The owner side gives two tables, both ascending
 * per owner bin and disjoint: heavy (the k-mers sent as a pair from some
 * bucket, with their total count) and light (all others); their union is
 * dakc_count's table.  The job's table_on_device chooses device tables in the
 * context's workspace (valid until the next HITTER count) or host copies.
 * stats (may be NULL) counts what the send side ships. */
typedef struct dakc_hitter_stats {
  uint64_t buckets;        /* buckets sent (with at least one k-mer) */
  uint64_t normal_keys;    /* plain k-mers sent */
  uint64_t heavy_pairs;    /* (k-mer, count) pairs sent */
  uint64_t kmers;          /* k-mers of the owners counted (plain + the pairs' counts) */
  uint64_t bytes_sent;     /* normal_keys * key bytes + heavy_pairs * (key bytes + 8: a u64 count,
                            * as the reference's heavy packet, kcounter.cpp:176-188) */
  uint64_t bytes_without;  /* kmers * key bytes: what the same owners receive with HITTER=0 */
} dakc_hitter_stats;
int dakc_count_hitter(dakc_ctx *ctx, const dakc_job *job, uint32_t bucket_size, dakc_table *light,
                      dakc_table *heavy, dakc_hitter_stats *stats);

/* Allocation-free count.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It allocates nothing, on
 * the device or in pinned host memory: its whole workspace is one fixed
 * device array of DAKC_NOALLOC_SCRATCH bytes that each GPU the library runs
 * on holds until the process ends.  Every call on one GPU shares that array,
 * so calls that overlap on one GPU (from any context) race.
 *   - reads: device memory (job->reads_on_device = 1), DAKC_FMT_FIXED;
 *   - the table goes to the caller's device buffers keys_out (cap * key_words
 *     u64) and counts_out (cap u32), ascending as dakc_count_device leaves
 *     it; bin_offsets_out (host, num_pes + 1) receives the owner-bin
 *     boundaries when all P bins are counted (required then, else unused);
 *   - *n_distinct / *n_total: table entries and k-mers.  When the table has
 *     more than cap entries nothing is written, the status is DAKC_E_NOMEM
 *     and *n_distinct says how many there are;
 *   - a job whose workspace does not fit the array is DAKC_E_TOOBIG (with
 *     k <= 32 about 10M k-mer window slots fit, with k <= 64 about 6M).
 * Same counts as dakc_count_device (MSD path, LSD recount on LDS overflow). */
#define DAKC_NOALLOC_SCRATCH (256ull << 20)
int dakc_count_noalloc(dakc_ctx *ctx, const dakc_job *job, uint64_t *keys_out, uint32_t *counts_out,
                       uint64_t cap, uint64_t *bin_offsets_out, uint64_t *n_distinct, uint64_t *n_total);

/* Batched: many independent jobs, one table each (outs[i] for jobs[i]).  Jobs
 * are counted together, one multi-job launch per stage, in chunks that fit the
 * HBM budget.  A job with table_on_device gets a device table of its own, so a
 * caller streaming more read sets than host memory holds tables for can count
 * them window by window and release each window's tables. */
int dakc_count_batch(dakc_ctx *ctx, const dakc_job *jobs, size_t njobs, dakc_table *outs);

/* Captured batch: plan and stage the jobs once (host reads are copied to the
 * device here), capture the whole pipeline in a hipGraph, then replay it with
 * dakc_batch_run as often as wanted; dakc_batch_tables returns one table per
 * job.  Device inputs must stay valid until dakc_batch_destroy. */
typedef struct dakc_batch dakc_batch;
int dakc_batch_create(dakc_ctx *ctx, const dakc_job *jobs, size_t njobs, dakc_batch **out);
/* How many of the batch's jobs the captured launch counts (the rest -- a
 * second key width, or jobs too large for the multi-job path -- are counted
 * by dakc_batch_tables one by one) and the device bytes the batch holds. */
int dakc_batch_info(const dakc_batch *b, size_t *jobs_in_launch, uint64_t *device_bytes);
int dakc_batch_run(dakc_batch *b, double *device_ms);
int dakc_batch_tables(dakc_batch *b, dakc_table *outs);
int dakc_batch_destroy(dakc_batch *b);

/* Autotuner: time the candidate geometries (final-level workgroup size,
 * sub-bucket fill, MSD or LSD) on a representative job and keep the fastest
 * for jobs of its key width and size class; save/load the table as JSON. */
int dakc_autotune(dakc_ctx *ctx, const dakc_job *job, int trials, double *best_ms);
int dakc_tuning_save(dakc_ctx *ctx, const char *path);
int dakc_tuning_load(dakc_ctx *ctx, const char *path);

/* Served: continuous batching.  submit copies the job descriptor (the reads
 * buffer must stay valid until completion); wait blocks for the table.  Jobs
 * are admitted in FIFO order when their buffers fit the served pool, one
 * device allocation of opts.hbm_budget bytes (default 60% of free memory at
 * the first submit) that also holds the device result tables; a job that can
 * only wait for tables the client already holds fails with DAKC_E_NOMEM. */
int dakc_submit(dakc_ctx *ctx, const dakc_job *job, uint64_t *ticket);
int dakc_poll(dakc_ctx *ctx, uint64_t ticket, int *done);
int dakc_wait(dakc_ctx *ctx, uint64_t ticket, dakc_table *out);

/* Served-mode counters of a context (all zero before the first submit). */
typedef struct dakc_serve_stats {
  uint64_t submitted, admitted, completed, failed;
  uint64_t lsd_reruns;         /* jobs whose MSD count overflowed LDS and reran */
  uint64_t budget;             /* bytes of the served pool (one device allocation) */
  uint64_t peak_device_bytes;  /* peak bytes of the pool in use */
  uint64_t resident_bytes;     /* live device result tables now */
  uint64_t pool_used;          /* pool bytes in use now (job buffers + tables) */
  uint64_t queued;             /* jobs waiting for a slot or memory now */
  double admit_host_ms;        /* host time spent admitting (buffers, enqueue) */
  double init_ms;              /* time to allocate the pool on the first submit */
} dakc_serve_stats;
int dakc_serve_get_stats(dakc_ctx *ctx, dakc_serve_stats *out);

void dakc_table_free(dakc_table *t);

/* Serialize a host table as little-endian records: u64 kmer (k <= 32) or
 * u64 lo, u64 hi (k <= 64), then u32 count.  Returns bytes written or
 * needed when buf is NULL. */
uint64_t dakc_table_serialize(const dakc_table *t, void *buf, uint64_t cap);

/* ------------------------------------------------ before the count --------
 * FASTQ (DAKC_FMT_FASTQ) or newline-delimited reads (DAKC_FMT_LINES) in device
 * memory (4-byte aligned) become fixed-width records on the device, as
 * syngen/fq2txtmaker.sh:12-18 makes them with awk: the read line of each
 * record (every line for LINES), padded with 'M' to read_len (0: the longest
 * read; a longer read is DAKC_E_FORMAT), every space made 'M', each followed
 * by '\n'.  out (device, cap bytes) receives *out_nbytes bytes of records
 * of width *out_read_len; with out == NULL only the sizes are returned. */
int dakc_convert_reads(dakc_ctx *ctx, const void *in, uint64_t nbytes, int format, uint32_t read_len,
                       void *out, uint64_t cap, uint64_t *out_nbytes, uint32_t *out_read_len);

/* ------------------------------------------------ after the count ---------
 * On a device table (on_device == 1; DAKC_E_INVAL otherwise).  The reference
 * stops at the table; these are the steps its consumers take next. */

/* Count-frequency histogram: hist[c] = entries with count c for c <
 * num_buckets (hist[0] = 0; counts >= num_buckets are not binned), and
 * *min_bucket = the first c in [1, num_buckets) with the fewest entries (1
 * when num_buckets == 1; 0, with no histogram, when num_buckets == 0).
 * Replaces construct_kmer_histogram's first pass
 * (pakman_star/src/kmer_histogram.cpp:79-116). */
int dakc_table_histogram(dakc_ctx *ctx, const dakc_table *t, uint32_t num_buckets, uint64_t *hist,
                         uint32_t *min_bucket);
/* Solid-k-mer cutoff: keep the entries with count >= min_count, in order, in
 * place; n, total (sum of the kept counts) and bin_offsets follow.  Replaces
 * the valid-entry selection of kmer_histogram.cpp:138-140. */
int dakc_table_filter(dakc_ctx *ctx, dakc_table *t, uint32_t min_count);
/* Point queries: counts_out[i] = count of queries[i] (key_words u64 each, lo
 * then hi) in the table, 0 when absent; looked up in the query's owner bin
 * when the table holds all P bins.  Queries are taken as given (a canonical
 * table wants canonical queries).  queries and counts_out are device
 * pointers when on_device, else host memory.  The lookup the reference
 * describes for its sorted table (dakc/src/kcounter/kcounter.cpp:381-383). */
int dakc_table_query(dakc_ctx *ctx, const dakc_table *t, const uint64_t *queries, uint64_t nq,
                     int on_device, uint32_t *counts_out);

/* ------------------------------------------------ synthetic reads ---------
 * Seeded, counter-based generator of reads in the fixed-width format, so the
 * same bytes can be made on the host and on the device. */
typedef struct dakc_gen_params {
  uint64_t seed;
  uint64_t num_reads;
  uint64_t first_read;  /* generate reads [first_read, first_read + num_reads) */
  uint64_t genome_len;  /* 0: i.i.d. uniform bases; else reads sampled from a genome */
  double sub_rate;      /* per-base substitution probability */
  double n_rate;        /* per-position probability that an N run starts */
  int32_t read_len;     /* record width */
  int32_t min_len;      /* read lengths uniform in [min_len, read_len]; rest M-padded */
  int32_t n_run_max;    /* N runs have length uniform in [1, n_run_max] */
  int32_t both_strands; /* genome reads from either strand */
} dakc_gen_params;

uint64_t dakc_gen_nbytes(const dakc_gen_params *p);
int dakc_generate_reads(const dakc_gen_params *p, char *out, uint64_t cap, int threads);
int dakc_generate_reads_device(dakc_ctx *ctx, const dakc_gen_params *p, void *dev_out, uint64_t cap);

/* device memory helpers for callers without their own allocator */
int dakc_device_alloc(dakc_ctx *ctx, uint64_t nbytes, void **dptr);
int dakc_device_free(dakc_ctx *ctx, void *dptr);
int dakc_memcpy_h2d(dakc_ctx *ctx, void *dst, const void *src, uint64_t nbytes);
int dakc_memcpy_d2h(dakc_ctx *ctx, void *dst, const void *src, uint64_t nbytes);
int dakc_memcpy_d2d(dakc_ctx *ctx, void *dst, const void *src, uint64_t nbytes);

#ifdef __cplusplus
}
#endif
#endif
