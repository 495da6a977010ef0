/*
 * dakc_oracle.h -- CPU restatement of DAKC's single-PE k-mer counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (dakc_amd/, include/)
 * links, loads or calls this code.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, and only as the checker / the timed CPU
 * baseline, never as the thing measured or shipped.
 *
 * Each function cites the reference file:line whose behaviour it restates
 * (paths relative to the Souvadra/dakc source tree).  Pins (what checks each
 * function against something other than its own formula) are listed in
 * tests/test_oracle.py:
 *   dko_char2base      exhaustive 256-byte golden vector extracted from the
 *                      reference's table (tests/golden/char2base_lut.bin)
 *   dko_extract        brute-force substring enumeration in Python
 *   dko_murmur64a_*    the general-length public MurmurHash64A written
 *                      independently in Python; len=8 must equal the
 *                      reference's single-block form
 *   dko_count          collections.Counter + sorted() on tiny inputs, golden
 *                      fixtures, bin-union == P=1 table, canonical identity
 */
#ifndef DAKC_ORACLE_H
#define DAKC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: same numbering as include/dakc.h */
#define DKO_OK 0
#define DKO_E_INVAL 1
#define DKO_E_FORMAT 2
#define DKO_E_NOMEM 3
#define DKO_E_TOOBIG 4

/* kcounter.cpp:52 -- the reference's owner-hash seed; a seed argument of 0
 * stands for it */
#define DKO_OWNER_SEED 0x9E3779B97F4A7C15ULL

/* common.hpp:72-92: ASCII -> 2-bit code (C0 A1 T2 G3), 0xF0 for M/m, 0xFF otherwise */
uint8_t dko_char2base(unsigned char c);

/* kcounter.cpp:29-48: MurmurHash64A of one 8-byte key */
uint64_t dko_murmur64a_u64(uint64_t key, uint64_t seed);

/* owner hash for a 2x64-bit key: MurmurHash64A over 16 little-endian bytes
 * (lo word then hi word).  Build decision; no reference counterpart. */
uint64_t dko_murmur64a_u128(uint64_t lo, uint64_t hi, uint64_t seed);

/* Extract every k-mer of a fixed-width record buffer (kcounter_funcs.cpp:25-111).
 * Records are read_len bytes + '\n'.  Keys are written as key_words() words each
 * (1 for k<=32: the value; 2 for k<=64: lo, hi).  Only k-mers whose owner ==
 * owner_bin are kept when num_pes > 1 and owner_bin >= 0.
 * Returns the k-mer count, or -(status) on error.  out may be NULL (count only). */
int64_t dko_extract(const char *reads, size_t nbytes, int read_len, int k, int canonical,
                    uint32_t num_pes, int32_t owner_bin, uint64_t seed, uint64_t *out, uint64_t cap);

/* Full single-PE count (kcounter.cpp:247-414, HITTER=0): extract, sort ascending
 * as unsigned, run-length reduce.  With num_pes > 1 and owner_bin == -1 the
 * result is the concatenation of the num_pes owner tables in owner order and
 * bin_offsets (num_pes+1 entries, may be NULL) gives their boundaries.
 * *keys_out (d * key_words u64) and *counts_out (d u32) are malloc'd; free with
 * dko_free.  Returns a status. */
int dko_count(const char *reads, size_t nbytes, int read_len, int k, int canonical,
              uint32_t num_pes, int32_t owner_bin, uint64_t seed, uint64_t **keys_out, uint32_t **counts_out,
              uint64_t *d_out, uint64_t *n_out, uint64_t *bin_offsets);

/* DAKC's P-PE shape on shared memory, for timing the CPU baseline and checking
 * large tables: `threads` workers each parse 1/threads of the records
 * (fqreader.cpp:43-52), route k-mers to threads worker bins by MurmurHash64A
 * (kcounter.cpp:50-54, 162-174), barrier, then each sorts and run-length
 * reduces its own bin (kcounter.cpp:385-411).  With num_pes > 1 and
 * owner_bin >= 0 only owner bin owner_bin's k-mers are counted.  Returns the
 * total k-mer count, distinct count and an order-independent checksum of the
 * (kmer, count) multiset (dko_table_checksum), so the result can be compared
 * with any table of the same entries (owner-bin order does not change it).
 * *elapsed_s is the wall time of the counting phases (file read excluded, as
 * kcounter.cpp:262).  k <= 64. */
int dko_count_mt(const char *reads, size_t nbytes, int read_len, int k, int canonical, uint32_t num_pes,
                 int32_t owner_bin, uint64_t seed, int threads, uint64_t *n_out, uint64_t *d_out,
                 uint64_t *checksum_out, double *elapsed_s);

/* The same Phase 2 over an array of n keys (kw words each, as dko_extract
 * writes them): the owner's sort and run-length reduce of its received
 * k-mers (kcounter.cpp:385-411).  Returns distinct count and checksum. */
int dko_count_keys_mt(const uint64_t *keys, uint64_t n, int k, int threads, uint64_t *d_out,
                      uint64_t *checksum_out, double *elapsed_s);

/* dko_extract on `threads` workers, keys in record order (out may be NULL:
 * count only; at most cap keys are written).  Returns the k-mer count or
 * -(status). */
int64_t dko_extract_mt(const char *reads, size_t nbytes, int read_len, int k, int canonical, uint32_t num_pes,
                       int32_t owner_bin, uint64_t seed, int threads, uint64_t *out, uint64_t cap);

/* order-independent checksum of a table: sum over entries of mix(key words, count) */
uint64_t dko_table_checksum(const uint64_t *keys, const uint32_t *counts, uint64_t d, int key_words);

void dko_free(void *p);

/* HITTER=1 (kcounter.cpp:56-104, 176-245, 325-384): records are taken into a
 * bucket while it holds at most bucket_size - read_len k-mers
 * (read_till_buf_max, kcounter_funcs.cpp:69-111; KCOUNT_BUCKET_SIZE 10000 in
 * dakc/Makefile:11); each bucket is sorted and run-length reduced, a k-mer
 * seen once or twice in it is sent to its owner as that many plain k-mers, one
 * seen three or more times as one (k-mer, count) pair.  Each owner merges its
 * heavy pairs into the heavy table, counts its plain k-mers, adds the runs of
 * heavy k-mers to the heavy table and keeps the others as the light table.
 * Both tables are ascending per owner bin and disjoint; their union is the
 * HITTER=0 table.  With num_pes > 1 and owner_bin == -1 both are the
 * concatenation of the P owners' tables (offsets: num_pes + 1 entries each,
 * may be NULL); with owner_bin >= 0 only that owner's.  Tables are malloc'd
 * (dko_free).  bucket_size (0: the reference's 10000) must exceed read_len. */
typedef struct {
    uint64_t buckets;      /* buckets flushed (with at least one k-mer) */
    uint64_t normal_keys;  /* plain k-mers sent */
    uint64_t heavy_pairs;  /* (k-mer, count) pairs sent */
} dko_hitter_stats;
int dko_count_hitter(const char *reads, size_t nbytes, int read_len, int k, int canonical, uint32_t num_pes,
                     int32_t owner_bin, uint64_t seed, uint32_t bucket_size, uint64_t **light_keys,
                     uint32_t **light_counts, uint64_t *d_light, uint64_t *light_offsets, uint64_t **heavy_keys,
                     uint32_t **heavy_counts, uint64_t *d_heavy, uint64_t *heavy_offsets, dko_hitter_stats *st);

/* Before the count: FASTQ / line input to fixed-width records (syngen/fq2txtmaker.sh:12-18). */
int dko_convert(const char *in, size_t n, int fastq, int read_len, char *out, size_t cap,
                uint64_t *out_n, uint32_t *width);

/* After the count (pakman_star/src/kmer_histogram.cpp:79-140; kcounter.cpp:381-383). */
uint32_t dko_histogram(const uint32_t *counts, uint64_t d, uint32_t num_buckets, uint64_t *hist);
uint64_t dko_filter(uint64_t *keys, uint32_t *counts, uint64_t d, int key_words, uint32_t min_count,
                    uint64_t *bin_offsets, uint32_t num_bins, uint64_t *total_out);
void dko_query(const uint64_t *keys, const uint32_t *counts, uint64_t d, int key_words,
               const uint64_t *bin_offsets, uint32_t num_bins, uint64_t seed, const uint64_t *queries,
               uint64_t nq, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
