/*
 * dakc_oracle.c -- CPU restatement of DAKC's single-PE k-mer counting path.
 *
 * TEST INFRASTRUCTURE ONLY (see dakc_oracle.h): the checker for the HIP path
 * and the timed CPU baseline.  Written from the behaviour of the reference
 * (Souvadra/dakc), not copied from it.  Every function cites what it restates.
 *
 * Deliberate, documented departures from the reference (its behaviour there is
 * undefined or reads stale memory):
 *   - empty input gives an empty table (kcounter.cpp:389-411 would emit one
 *     stale entry);
 *   - a buffer whose length is not a whole number of records, a separator byte
 *     that is not '\n', or a record that starts with NUL is a format error
 *     (the reference over-reads, mis-parses, or silently stops parsing);
 *   - a final record without its '\n' is accepted (kcounter_funcs.cpp:102-103
 *     would read past the buffer).
 */
#define _GNU_SOURCE
#include "dakc_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned __int128 u128;

/* common.hpp:72-92 (char2base): C/c->0, A/a->1, T/t->2, G/g->3, M/m->0xF0,
 * every other byte -> 0xFF.  Written as a switch, not as a copy of the table. */
uint8_t dko_char2base(unsigned char c) {
    switch (c) {
    case 'C': case 'c': return 0;
    case 'A': case 'a': return 1;
    case 'T': case 't': return 2;
    case 'G': case 'g': return 3;
    case 'M': case 'm': return 0xF0;
    default: return 0xFF;
    }
}

/* kcounter.cpp:29-48 (MurmurHash64A, one 8-byte block, len = 8) */
uint64_t dko_murmur64a_u64(uint64_t key, uint64_t seed) {
    const uint64_t m = 0xc6a4a7935bd1e995ULL;
    const int r = 47;
    uint64_t h = seed ^ (8 * m);
    uint64_t k = key * m;
    k ^= k >> r;
    k *= m;
    h ^= k;
    h *= m;
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

/* MurmurHash64A (Appleby's public algorithm) over the 16 bytes lo||hi. */
uint64_t dko_murmur64a_u128(uint64_t lo, uint64_t hi, uint64_t seed) {
    const uint64_t m = 0xc6a4a7935bd1e995ULL;
    const int r = 47;
    uint64_t h = seed ^ (16 * m);
    uint64_t blocks[2] = {lo, hi};
    for (int b = 0; b < 2; b++) {
        uint64_t k = blocks[b] * m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

static uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}

uint64_t dko_table_checksum(const uint64_t *keys, const uint32_t *counts, uint64_t d, int key_words) {
    uint64_t s = 0;
    for (uint64_t i = 0; i < d; i++) {
        uint64_t lo = keys[i * key_words];
        uint64_t hi = key_words == 2 ? keys[i * key_words + 1] : 0;
        s += mix64(mix64(lo + 0x9E3779B97F4A7C15ULL) ^ (hi + 0xD1B54A32D192ED03ULL) ^ ((uint64_t)counts[i] << 17));
    }
    return s;
}

void dko_free(void *p) { free(p); }

/* ---------------------------------------------------------------- format --- */

/* Record layout of fqreader.cpp:43 and kcounter_funcs.cpp:102: read_len bytes
 * then '\n'.  Returns the record count or -(status). */
static int64_t count_records(const char *reads, size_t nbytes, int read_len) {
    if (read_len < 1) return -DKO_E_INVAL;
    if (nbytes == 0) return 0;
    if (!reads) return -DKO_E_INVAL;
    size_t stride = (size_t)read_len + 1;
    size_t nrec = nbytes / stride;
    size_t rem = nbytes % stride;
    if (rem == (size_t)read_len) nrec += 1; /* last record without '\n' */
    else if (rem != 0) return -DKO_E_FORMAT;
    for (size_t r = 0; r < nrec; r++) {
        const char *rec = reads + r * stride;
        if (rec[0] == '\0') return -DKO_E_FORMAT;
        if ((r + 1) * stride <= nbytes && rec[read_len] != '\n') return -DKO_E_FORMAT;
    }
    return (int64_t)nrec;
}

static int key_words_for(int k) { return k <= 32 ? 1 : 2; }

static u128 mask_for(int k) {
    return k >= 64 ? ~(u128)0 : (((u128)1 << (2 * k)) - 1);
}

/* Emission context for one run of valid codes. */
typedef struct {
    int k, canonical, kw;
    uint32_t num_pes;
    int32_t owner_bin;
    uint64_t seed;
    uint64_t *out;
    uint64_t cap, n;
} emit_ctx;

/* the seed a job names, 0 standing for the reference's (kcounter.cpp:52) */
static inline uint64_t seed_or_default(uint64_t seed) { return seed ? seed : DKO_OWNER_SEED; }

/* MurmurHash64A of a key of kw words (kcounter.cpp:29-48 for kw = 1) */
static inline uint64_t key_hash(u128 key, int kw, uint64_t seed) {
    return kw == 1 ? dko_murmur64a_u64((uint64_t)key, seed)
                   : dko_murmur64a_u128((uint64_t)key, (uint64_t)(key >> 64), seed);
}

/* kcounter.cpp:50-54 (owner_pe) */
static inline uint32_t owner_of(u128 key, int kw, uint32_t num_pes, uint64_t seed) {
    return (uint32_t)(key_hash(key, kw, seed) % num_pes);
}

static inline void emit_key(emit_ctx *c, u128 key) {
    if (c->num_pes > 1 && c->owner_bin >= 0 &&
        owner_of(key, c->kw, c->num_pes, c->seed) != (uint32_t)c->owner_bin)
        return;
    if (c->out && c->n < c->cap) {
        c->out[c->n * c->kw] = (uint64_t)key;
        if (c->kw == 2) c->out[c->n * c->kw + 1] = (uint64_t)(key >> 64);
    }
    c->n++;
}

/* kcounter_funcs.cpp:40-67 (get_kmers) with set_kmer_fast / update_kmer_fast
 * (:25-38): the first k-mer of a run is built from k codes, MSB first; each
 * next k-mer shifts in one code and masks to 2k bits (KMER_MASK common.hpp:61).
 * Canonical mode (no reference counterpart) keeps min(x, revcomp(x)). */
static void emit_run(emit_ctx *c, const uint8_t *codes, int len) {
    int k = c->k;
    if (len < k) return;
    const u128 mask = mask_for(k);
    u128 x = 0, rc = 0;
    for (int j = 0; j < k; j++) {
        x = (x << 2) | codes[j];
        rc = (rc >> 2) | ((u128)(codes[j] ^ 3) << (2 * (k - 1)));
    }
    emit_key(c, c->canonical && rc < x ? rc : x);
    for (int i = k; i < len; i++) {
        x = ((x << 2) | codes[i]) & mask;
        rc = (rc >> 2) | ((u128)(codes[i] ^ 3) << (2 * (k - 1)));
        emit_key(c, c->canonical && rc < x ? rc : x);
    }
}

/* kcounter_funcs.cpp:69-111 (read_till_buf_max), one record: decode each byte
 * (common.hpp:72-92); a 0xFF closes the current run (:88-91); a 0xF0 ('M')
 * ends the record (:92-94); the final run is emitted at the end (:97-99). */
static void parse_record(emit_ctx *c, const char *rec, int read_len, uint8_t *codes) {
    int left = 0, i;
    for (i = 0; i < read_len; i++) {
        uint8_t b = dko_char2base((unsigned char)rec[i]);
        codes[i] = b;
        if (b == 0xFF) {
            emit_run(c, codes + left, i - left);
            left = i + 1;
        }
        if (b == 0xF0) break;
    }
    if (left < read_len - 1) emit_run(c, codes + left, i - left);
}

static int64_t extract_range(const char *reads, int read_len, uint64_t r0, uint64_t r1, emit_ctx *c) {
    uint8_t *codes = (uint8_t *)malloc((size_t)read_len + 1);
    if (!codes) return -DKO_E_NOMEM;
    size_t stride = (size_t)read_len + 1;
    for (uint64_t r = r0; r < r1; r++) parse_record(c, reads + r * stride, read_len, codes);
    free(codes);
    return (int64_t)c->n;
}

int64_t dko_extract(const char *reads, size_t nbytes, int read_len, int k, int canonical,
                    uint32_t num_pes, int32_t owner_bin, uint64_t seed, uint64_t *out, uint64_t cap) {
    if (k < 1 || k > 64 || num_pes < 1) return -DKO_E_INVAL;
    if (num_pes > 1 && owner_bin >= (int32_t)num_pes) return -DKO_E_INVAL;
    int64_t nrec = count_records(reads, nbytes, read_len);
    if (nrec < 0) return nrec;
    emit_ctx c = {k, canonical, key_words_for(k), num_pes, owner_bin, seed_or_default(seed), out, cap, 0};
    return extract_range(reads, read_len, 0, (uint64_t)nrec, &c);
}

/* ------------------------------------------------------------------ sort --- */

/* Ascending unsigned LSD radix sort, 8-bit digits, constant digits skipped.
 * Stands in for ska_sort (kcounter.cpp:388, ska_sort.hpp:1417-1421): the sorted
 * order of integer keys is unique, so the sort algorithm does not matter. */
static void sort_u64(uint64_t *a, uint64_t *tmp, uint64_t n, int bits) {
    uint64_t *src = a, *dst = tmp;
    for (int sh = 0; sh < bits; sh += 8) {
        uint64_t cnt[256] = {0};
        for (uint64_t i = 0; i < n; i++) cnt[(src[i] >> sh) & 255]++;
        int trivial = 0;
        for (int d = 0; d < 256; d++) if (cnt[d] == n) trivial = 1;
        if (trivial) continue;
        uint64_t s = 0;
        for (int d = 0; d < 256; d++) { uint64_t t = cnt[d]; cnt[d] = s; s += t; }
        for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & 255]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(uint64_t));
}

/* same for (lo, hi) pairs stored interleaved, ordered by the 128-bit value */
static void sort_u128(uint64_t *a, uint64_t *tmp, uint64_t n, int bits) {
    uint64_t *src = a, *dst = tmp;
    for (int sh = 0; sh < bits; sh += 8) {
        int w = sh >= 64 ? 1 : 0, s8 = sh & 63;
        uint64_t cnt[256] = {0};
        for (uint64_t i = 0; i < n; i++) cnt[(src[2 * i + w] >> s8) & 255]++;
        int trivial = 0;
        for (int d = 0; d < 256; d++) if (cnt[d] == n) trivial = 1;
        if (trivial) continue;
        uint64_t s = 0;
        for (int d = 0; d < 256; d++) { uint64_t t = cnt[d]; cnt[d] = s; s += t; }
        for (uint64_t i = 0; i < n; i++) {
            uint64_t p = cnt[(src[2 * i + w] >> s8) & 255]++;
            dst[2 * p] = src[2 * i];
            dst[2 * p + 1] = src[2 * i + 1];
        }
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, 2 * n * sizeof(uint64_t));
}

static void sort_keys(uint64_t *a, uint64_t *tmp, uint64_t n, int k) {
    if (key_words_for(k) == 1) sort_u64(a, tmp, n, 2 * k);
    else sort_u128(a, tmp, n, 2 * k);
}

/* kcounter.cpp:389-411: collapse equal neighbours of a sorted array into
 * (kmer, count).  Writes in place over `a` (d <= n); returns d.  An empty input
 * gives d = 0 (the reference would emit a stale entry here). */
static uint64_t rle(uint64_t *a, uint64_t n, int kw, uint32_t *counts) {
    if (n == 0) return 0;
    uint64_t d = 0, run = 1;
    for (uint64_t i = 1; i <= n; i++) {
        int same = 0;
        if (i < n) {
            same = a[i * kw] == a[(i - 1) * kw];
            if (kw == 2) same = same && a[i * kw + 1] == a[(i - 1) * kw + 1];
        }
        if (same) { run++; continue; }
        for (int w = 0; w < kw; w++) a[d * kw + w] = a[(i - 1) * kw + w];
        counts[d++] = run > 0xFFFFFFFFULL ? 0xFFFFFFFFu : (uint32_t)run;
        run = 1;
    }
    return d;
}

int dko_count(const char *reads, size_t nbytes, int read_len, int k, int canonical,
              uint32_t num_pes, int32_t owner_bin, uint64_t seed, uint64_t **keys_out, uint32_t **counts_out,
              uint64_t *d_out, uint64_t *n_out, uint64_t *bin_offsets) {
    if (!keys_out || !counts_out || !d_out) return DKO_E_INVAL;
    *keys_out = NULL; *counts_out = NULL; *d_out = 0;
    seed = seed_or_default(seed);
    int64_t n = dko_extract(reads, nbytes, read_len, k, canonical, num_pes, owner_bin, seed, NULL, 0);
    if (n < 0) return (int)-n;
    if ((uint64_t)n > 0xFFFFFFFFULL) return DKO_E_TOOBIG;
    int kw = key_words_for(k);
    uint64_t *keys = (uint64_t *)malloc(((size_t)n + 1) * kw * sizeof(uint64_t));
    uint64_t *tmp = (uint64_t *)malloc(((size_t)n + 1) * kw * sizeof(uint64_t));
    uint32_t *counts = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    if (!keys || !tmp || !counts) { free(keys); free(tmp); free(counts); return DKO_E_NOMEM; }
    dko_extract(reads, nbytes, read_len, k, canonical, num_pes, owner_bin, seed, keys, (uint64_t)n);
    if (n_out) *n_out = (uint64_t)n;
    uint64_t d = 0;
    if (num_pes > 1 && owner_bin < 0) {
        /* all owner bins: partition by owner (kcounter.cpp:162-174 routes each
         * k-mer to its owner), then each owner sorts and reduces its own bin. */
        uint64_t *cnt = (uint64_t *)calloc((size_t)num_pes + 1, sizeof(uint64_t));
        uint32_t *own = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
        if (!cnt || !own) { free(cnt); free(own); free(keys); free(tmp); free(counts); return DKO_E_NOMEM; }
        for (int64_t i = 0; i < n; i++) {
            u128 key = keys[i * kw];
            if (kw == 2) key |= (u128)keys[i * kw + 1] << 64;
            own[i] = owner_of(key, kw, num_pes, seed);
            cnt[own[i] + 1]++;
        }
        for (uint32_t p = 0; p < num_pes; p++) cnt[p + 1] += cnt[p];
        uint64_t *pos = (uint64_t *)malloc((size_t)num_pes * sizeof(uint64_t));
        if (!pos) { free(cnt); free(own); free(keys); free(tmp); free(counts); return DKO_E_NOMEM; }
        memcpy(pos, cnt, (size_t)num_pes * sizeof(uint64_t));
        for (int64_t i = 0; i < n; i++) {
            uint64_t p = pos[own[i]]++;
            for (int w = 0; w < kw; w++) tmp[p * kw + w] = keys[i * kw + w];
        }
        memcpy(keys, tmp, (size_t)n * kw * sizeof(uint64_t));
        for (uint32_t p = 0; p < num_pes; p++) {
            uint64_t b0 = cnt[p], b1 = cnt[p + 1];
            sort_keys(keys + b0 * kw, tmp, b1 - b0, k);
            uint64_t db = rle(keys + b0 * kw, b1 - b0, kw, counts + d);
            if (bin_offsets) bin_offsets[p] = d;
            memmove(keys + d * kw, keys + b0 * kw, db * kw * sizeof(uint64_t));
            d += db;
        }
        if (bin_offsets) bin_offsets[num_pes] = d;
        free(pos); free(own); free(cnt);
    } else {
        sort_keys(keys, tmp, (uint64_t)n, k);
        d = rle(keys, (uint64_t)n, kw, counts);
        if (bin_offsets) { bin_offsets[0] = 0; bin_offsets[1] = d; }
    }
    free(tmp);
    *keys_out = keys;
    *counts_out = counts;
    *d_out = d;
    return DKO_OK;
}

/* -------------------------------------------------- threaded P-PE shape --- */

/* Shared-memory stand-in for DAKC's P-PE run, used to time the CPU baseline
 * and to check large GPU tables by checksum.  T workers each act as one PE:
 * Phase 1 parses a block of records (fqreader.cpp:43-52) and routes every
 * k-mer to its worker bin (kcounter.cpp:162-174 routes it to its owner PE);
 * after a barrier, Phase 2 sorts and run-length reduces each bin
 * (kcounter.cpp:385-411).  A worker bin is MurmurHash64A % T, the reference's
 * owner_pe over T PEs, taken from (hash / P) when the job itself keeps only
 * owner bin b of P (so the kept k-mers still spread over the workers). */
typedef struct {
    int tid, T, kw;
    uint64_t seed, div;          /* worker bin = (hash / div) % T */
    /* Phase 1 input: records [r0, r1), or keys [k0, k1) of a key array */
    const char *reads;
    int read_len, k, canonical;
    uint32_t num_pes;
    int32_t owner_bin;
    uint64_t r0, r1;
    const uint64_t *keys_in;
    uint64_t k0, k1;
    uint64_t *local, nlocal;     /* this worker's k-mers (Phase 1), capacity then count */
    uint64_t *cnt;               /* T x T: cnt[t*T + o] */
    uint64_t *global, *tmp;      /* n keys each (kw words per key) */
    uint64_t *bin_start;         /* T + 1 */
    uint64_t d, checksum;
    pthread_barrier_t *bar;
} mt_arg;

static inline u128 key_at(const uint64_t *a, uint64_t i, int kw) {
    return kw == 1 ? (u128)a[i] : ((u128)a[2 * i] | ((u128)a[2 * i + 1] << 64));
}
static inline void key_put(uint64_t *a, uint64_t i, int kw, u128 x) {
    if (kw == 1) {
        a[i] = (uint64_t)x;
    } else {
        a[2 * i] = (uint64_t)x;
        a[2 * i + 1] = (uint64_t)(x >> 64);
    }
}
static inline uint32_t worker_bin(const mt_arg *a, u128 x) {
    return (uint32_t)((key_hash(x, a->kw, a->seed) / a->div) % (uint64_t)a->T);
}

static void *mt_worker(void *p) {
    mt_arg *a = (mt_arg *)p;
    const int T = a->T, kw = a->kw;
    /* Phase 1: this worker's records (or its slice of the key array) */
    if (a->reads) {
        emit_ctx c = {a->k, a->canonical, kw, a->num_pes, a->owner_bin, a->seed, a->local, a->nlocal, 0};
        extract_range(a->reads, a->read_len, a->r0, a->r1, &c);
        a->nlocal = c.n;
    } else {
        a->nlocal = a->k1 - a->k0;
        memcpy(a->local, a->keys_in + a->k0 * kw, (size_t)a->nlocal * kw * sizeof(uint64_t));
    }
    uint64_t *mycnt = a->cnt + (size_t)a->tid * T;
    for (int o = 0; o < T; o++) mycnt[o] = 0;
    for (uint64_t i = 0; i < a->nlocal; i++) mycnt[worker_bin(a, key_at(a->local, i, kw))]++;
    pthread_barrier_wait(a->bar);
    /* worker bins: bin o holds worker 0's k-mers for o, then worker 1's, ... */
    uint64_t *pos = (uint64_t *)malloc((size_t)T * sizeof(uint64_t));
    if (!pos) {
        a->d = 0;
        return NULL;
    }
    for (int o = 0; o < T; o++) {
        uint64_t s = 0;
        for (int o2 = 0; o2 < o; o2++) for (int t = 0; t < T; t++) s += a->cnt[(size_t)t * T + o2];
        for (int t = 0; t < a->tid; t++) s += a->cnt[(size_t)t * T + o];
        pos[o] = s;
        if (a->tid == 0) a->bin_start[o] = s;
    }
    if (a->tid == 0) {
        uint64_t tot = 0;
        for (size_t i = 0; i < (size_t)T * T; i++) tot += a->cnt[i];
        a->bin_start[T] = tot;
    }
    for (uint64_t i = 0; i < a->nlocal; i++) {
        const u128 x = key_at(a->local, i, kw);
        key_put(a->global, pos[worker_bin(a, x)]++, kw, x);
    }
    free(pos);
    pthread_barrier_wait(a->bar);
    /* Phase 2: each worker sorts and reduces its own bin (kcounter.cpp:385-411) */
    const uint64_t b0 = a->bin_start[a->tid], b1 = a->bin_start[a->tid + 1];
    uint64_t *g = a->global + b0 * kw, *t = a->tmp + b0 * kw;
    sort_keys(g, t, b1 - b0, a->k);
    uint32_t *counts = (uint32_t *)t; /* reuse scratch */
    const uint64_t d = rle(g, b1 - b0, kw, counts);
    a->d = d;
    a->checksum = dko_table_checksum(g, counts, d, kw);
    return NULL;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Run the T workers over records or keys; nmax bounds the k-mers. */
static int run_mt(mt_arg *proto, uint64_t nunits, uint64_t per_unit, uint64_t nmax, int threads,
                  uint64_t *n_out, uint64_t *d_out, uint64_t *checksum_out, double *elapsed_s) {
    const int T = threads, kw = proto->kw;
    const size_t kb = (size_t)kw * sizeof(uint64_t);
    uint64_t *local = (uint64_t *)malloc((size_t)(nmax + 1) * kb);
    uint64_t *global = (uint64_t *)malloc((size_t)(nmax + 1) * kb);
    uint64_t *tmp = (uint64_t *)malloc((size_t)(nmax + 1) * kb);
    uint64_t *cnt = (uint64_t *)calloc((size_t)T * T, sizeof(uint64_t));
    uint64_t *bin_start = (uint64_t *)calloc((size_t)T + 1, sizeof(uint64_t));
    mt_arg *args = (mt_arg *)calloc((size_t)T, sizeof(mt_arg));
    pthread_t *th = (pthread_t *)calloc((size_t)T, sizeof(pthread_t));
    if (!local || !global || !tmp || !cnt || !bin_start || !args || !th) {
        free(local); free(global); free(tmp); free(cnt); free(bin_start); free(args); free(th);
        return DKO_E_NOMEM;
    }
    /* touch the buffers once so page faults are not timed */
    memset(local, 0, (size_t)(nmax + 1) * kb);
    memset(global, 0, (size_t)(nmax + 1) * kb);
    memset(tmp, 0, (size_t)(nmax + 1) * kb);
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)T);
    const double t0 = now_s();
    for (int t = 0; t < T; t++) {
        mt_arg *a = &args[t];
        *a = *proto;
        a->tid = t;
        a->T = T;
        const uint64_t u0 = nunits * (uint64_t)t / T, u1 = nunits * (uint64_t)(t + 1) / T;
        if (a->reads) {
            a->r0 = u0;
            a->r1 = u1;
        } else {
            a->k0 = u0;
            a->k1 = u1;
        }
        a->local = local + u0 * per_unit * kw;
        a->nlocal = (u1 - u0) * per_unit;
        a->cnt = cnt; a->global = global; a->tmp = tmp; a->bin_start = bin_start; a->bar = &bar;
        pthread_create(&th[t], NULL, mt_worker, a);
    }
    uint64_t d = 0, cs = 0;
    for (int t = 0; t < T; t++) {
        pthread_join(th[t], NULL);
        d += args[t].d;
        cs += args[t].checksum;
    }
    const double t1 = now_s();
    pthread_barrier_destroy(&bar);
    if (n_out) *n_out = bin_start[T];
    if (d_out) *d_out = d;
    if (checksum_out) *checksum_out = cs;
    if (elapsed_s) *elapsed_s = t1 - t0;
    free(local); free(global); free(tmp); free(cnt); free(bin_start); free(args); free(th);
    return DKO_OK;
}

int dko_count_mt(const char *reads, size_t nbytes, int read_len, int k, int canonical, uint32_t num_pes,
                 int32_t owner_bin, uint64_t seed, int threads, uint64_t *n_out, uint64_t *d_out,
                 uint64_t *checksum_out, double *elapsed_s) {
    if (threads < 1 || threads > 1024 || k < 1 || k > 64 || num_pes < 1) return DKO_E_INVAL;
    if (num_pes > 1 && owner_bin >= (int32_t)num_pes) return DKO_E_INVAL;
    int64_t nrec = count_records(reads, nbytes, read_len);
    if (nrec < 0) return (int)-nrec;
    const uint64_t per_read = read_len >= k ? (uint64_t)(read_len - k + 1) : 0;
    mt_arg proto;
    memset(&proto, 0, sizeof(proto));
    proto.kw = key_words_for(k);
    proto.seed = seed_or_default(seed);
    proto.div = num_pes > 1 && owner_bin >= 0 ? num_pes : 1;
    proto.reads = nrec ? reads : "";
    proto.read_len = read_len;
    proto.k = k;
    proto.canonical = canonical != 0;
    proto.num_pes = num_pes;
    proto.owner_bin = num_pes > 1 ? owner_bin : -1;
    return run_mt(&proto, (uint64_t)nrec, per_read, (uint64_t)nrec * per_read, threads, n_out, d_out,
                  checksum_out, elapsed_s);
}

int dko_count_keys_mt(const uint64_t *keys, uint64_t n, int k, int threads, uint64_t *d_out,
                      uint64_t *checksum_out, double *elapsed_s) {
    if (threads < 1 || threads > 1024 || k < 1 || k > 64 || (n && !keys)) return DKO_E_INVAL;
    mt_arg proto;
    memset(&proto, 0, sizeof(proto));
    proto.kw = key_words_for(k);
    proto.seed = DKO_OWNER_SEED;
    proto.div = 1;
    proto.k = k;
    proto.keys_in = keys;
    uint64_t nn = 0;
    return run_mt(&proto, n, 1, n, threads, &nn, d_out, checksum_out, elapsed_s);
}

typedef struct {
    emit_ctx c;
    const char *reads;
    int read_len;
    uint64_t r0, r1;
} ex_arg;

static void *ex_worker(void *p) {
    ex_arg *a = (ex_arg *)p;
    extract_range(a->reads, a->read_len, a->r0, a->r1, &a->c);
    return NULL;
}

int64_t dko_extract_mt(const char *reads, size_t nbytes, int read_len, int k, int canonical, uint32_t num_pes,
                       int32_t owner_bin, uint64_t seed, int threads, uint64_t *out, uint64_t cap) {
    if (threads < 1 || threads > 1024 || k < 1 || k > 64 || num_pes < 1) return -DKO_E_INVAL;
    if (num_pes > 1 && owner_bin >= (int32_t)num_pes) return -DKO_E_INVAL;
    int64_t nrec = count_records(reads, nbytes, read_len);
    if (nrec < 0) return nrec;
    const int kw = key_words_for(k), T = threads;
    const uint64_t per_read = read_len >= k ? (uint64_t)(read_len - k + 1) : 0;
    ex_arg *args = (ex_arg *)calloc((size_t)T, sizeof(ex_arg));
    pthread_t *th = (pthread_t *)calloc((size_t)T, sizeof(pthread_t));
    uint64_t *scratch = (uint64_t *)malloc(((size_t)nrec * per_read + 1) * kw * sizeof(uint64_t));
    if (!args || !th || !scratch) {
        free(args); free(th); free(scratch);
        return -DKO_E_NOMEM;
    }
    for (int t = 0; t < T; t++) {
        ex_arg *a = &args[t];
        a->reads = reads;
        a->read_len = read_len;
        a->r0 = (uint64_t)nrec * t / T;
        a->r1 = (uint64_t)nrec * (t + 1) / T;
        emit_ctx c = {k, canonical, kw, num_pes, num_pes > 1 ? owner_bin : -1, seed_or_default(seed),
                      scratch + a->r0 * per_read * kw, (a->r1 - a->r0) * per_read, 0};
        a->c = c;
        pthread_create(&th[t], NULL, ex_worker, a);
    }
    uint64_t o = 0;
    for (int t = 0; t < T; t++) {
        pthread_join(th[t], NULL);
        const uint64_t n = args[t].c.n;
        if (out && o + n <= cap) memcpy(out + o * kw, args[t].c.out, (size_t)n * kw * sizeof(uint64_t));
        o += n;
    }
    free(args); free(th); free(scratch);
    return (int64_t)o;
}

/* ------------------------------------------------------- after the count --- */

/* Count-frequency histogram and its minimum bucket, as construct_kmer_histogram
 * computes them (pakman_star/src/kmer_histogram.cpp:81-116): hist[c] counts the
 * entries with count c < num_buckets (:87-92), and min_bucket is the index of
 * std::min_element over [1, num_buckets) (:109-110): the first minimum, or
 * num_buckets (= 1) when that range is empty.  num_buckets == 0 skips the
 * histogram and gives min_bucket 0 (:79-81). */
uint32_t dko_histogram(const uint32_t *counts, uint64_t d, uint32_t num_buckets, uint64_t *hist) {
    if (num_buckets == 0) return 0;
    for (uint32_t c = 0; c < num_buckets; c++) hist[c] = 0;
    for (uint64_t i = 0; i < d; i++)
        if (counts[i] < num_buckets) hist[counts[i]]++;
    uint32_t m = 1;
    for (uint32_t c = 2; c < num_buckets; c++)
        if (hist[c] < hist[m]) m = c;
    return m;
}

/* Keep the entries with count >= min_count, in table order, in place
 * (kmer_histogram.cpp:138-140 selects k_count >= min_bucket).  Returns the kept
 * count; *total_out = sum of the kept counts; bin_offsets (num_bins + 1
 * entries, or NULL) become the kept entries' bin boundaries. */
uint64_t dko_filter(uint64_t *keys, uint32_t *counts, uint64_t d, int key_words, uint32_t min_count,
                    uint64_t *bin_offsets, uint32_t num_bins, uint64_t *total_out) {
    uint64_t o = 0, total = 0, b = 0;
    for (uint64_t i = 0; i <= d; i++) {
        /* boundaries at position i move to the kept entries before i */
        while (bin_offsets && b <= num_bins && bin_offsets[b] == i) bin_offsets[b++] = o;
        if (i == d) break;
        if (counts[i] < min_count) continue;
        for (int x = 0; x < key_words; x++) keys[o * key_words + x] = keys[i * key_words + x];
        counts[o] = counts[i];
        total += counts[i];
        o++;
    }
    if (total_out) *total_out = total;
    return o;
}

static int key_cmp(const uint64_t *a, const uint64_t *b, int kw) {
    for (int x = kw - 1; x >= 0; x--) /* the high word decides first */
        if (a[x] != b[x]) return a[x] < b[x] ? -1 : 1;
    return 0;
}

/* Count of each query in a sorted table (0 when absent), by binary search in
 * the query's owner bin (MurmurHash64A % num_bins, kcounter.cpp:50-54) when
 * bin_offsets is given, else in the whole table: the lookup kcounter.cpp:381-383
 * describes for the sorted lightdbg. */
void dko_query(const uint64_t *keys, const uint32_t *counts, uint64_t d, int key_words,
               const uint64_t *bin_offsets, uint32_t num_bins, uint64_t seed, const uint64_t *queries,
               uint64_t nq, uint32_t *out) {
    seed = seed_or_default(seed);
    for (uint64_t i = 0; i < nq; i++) {
        const uint64_t *q = queries + i * key_words;
        uint64_t lo = 0, hi = d;
        if (bin_offsets && num_bins) {
            uint64_t h = key_words == 1 ? dko_murmur64a_u64(q[0], seed) : dko_murmur64a_u128(q[0], q[1], seed);
            uint32_t o = (uint32_t)(h % num_bins);
            lo = bin_offsets[o];
            hi = bin_offsets[o + 1];
        }
        out[i] = 0;
        while (lo < hi) {
            uint64_t mid = lo + (hi - lo) / 2;
            int c = key_cmp(keys + mid * key_words, q, key_words);
            if (c == 0) { out[i] = counts[mid]; break; }
            if (c < 0) lo = mid + 1; else hi = mid;
        }
    }
}

/* ------------------------------------------------------ before the count --- */

/* FASTQ (fastq != 0) or newline-delimited reads to the fixed-width format, as
 * syngen/fq2txtmaker.sh does with awk: the lines with NR % 4 == 2 (:12; every
 * line for plain reads), where a line ends at '\n' and a last line without
 * its '\n' still counts; the longest such line sets the width (:15); each line
 * is left-justified in that width with spaces and every space becomes 'M'
 * (:18, printf "%-Ns" | sed 's/ /M/g'), then '\n'.  read_len > 0 sets the
 * width instead (a longer line is DKO_E_FORMAT); the width is at least 1.
 * out == NULL: sizes only.  *out_n = bytes written, *width = record width. */
int dko_convert(const char *in, size_t n, int fastq, int read_len, char *out, size_t cap,
                uint64_t *out_n, uint32_t *width) {
    if (read_len < 0 || read_len > 4096) return DKO_E_INVAL;
    uint64_t longest = 0, nrec = 0, line = 0;
    for (size_t s = 0; s < n; line++) {
        size_t e = s;
        while (e < n && in[e] != '\n') e++;
        if (!fastq || line % 4 == 1) {
            if (e - s > longest) longest = e - s;
            nrec++;
        }
        s = e + 1;
    }
    if (read_len && longest > (uint64_t)read_len) return DKO_E_FORMAT;
    if (!read_len && longest > 4096) return DKO_E_TOOBIG;
    uint32_t L = read_len ? (uint32_t)read_len : (longest ? (uint32_t)longest : 1u);
    uint64_t need = nrec * ((uint64_t)L + 1);
    if (out_n) *out_n = need;
    if (width) *width = L;
    if (!out) return DKO_OK;
    if (cap < need) return DKO_E_INVAL;
    uint64_t o = 0;
    line = 0;
    for (size_t s = 0; s < n; line++) {
        size_t e = s;
        while (e < n && in[e] != '\n') e++;
        if (!fastq || line % 4 == 1) {
            for (uint32_t p = 0; p < L; p++) {
                char c = s + p < e ? in[s + p] : ' ';
                out[o++] = c == ' ' ? 'M' : c;
            }
            out[o++] = '\n';
        }
        s = e + 1;
    }
    return DKO_OK;
}

/* ------------------------------------------------------------ HITTER=1 --- */

/* A (key, count) pair of the heavy stream, kept with its owner. */
typedef struct {
    u128 key;
    uint64_t count;
    uint32_t owner;
} hit_pair;

static int hit_pair_cmp(const void *a, const void *b) {
    const hit_pair *x = (const hit_pair *)a, *y = (const hit_pair *)b;
    if (x->owner != y->owner) return x->owner < y->owner ? -1 : 1;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return 0;
}

/* grow-only arrays of the two send streams */
typedef struct {
    uint64_t *keys, n, cap;  /* normal stream: kw words per key */
    hit_pair *pairs;
    uint64_t np, pcap;
    int kw, oom;
} hit_streams;

static void hit_put_key(hit_streams *s, u128 x) {
    if (s->n == s->cap) {
        uint64_t nc = s->cap ? 2 * s->cap : 1024;
        uint64_t *p = (uint64_t *)realloc(s->keys, (size_t)nc * s->kw * sizeof(uint64_t));
        if (!p) { s->oom = 1; return; }
        s->keys = p;
        s->cap = nc;
    }
    key_put(s->keys, s->n++, s->kw, x);
}

static void hit_put_pair(hit_streams *s, u128 x, uint64_t c, uint32_t o) {
    if (s->np == s->pcap) {
        uint64_t nc = s->pcap ? 2 * s->pcap : 256;
        hit_pair *p = (hit_pair *)realloc(s->pairs, (size_t)nc * sizeof(hit_pair));
        if (!p) { s->oom = 1; return; }
        s->pairs = p;
        s->pcap = nc;
    }
    s->pairs[s->np].key = x;
    s->pairs[s->np].count = c;
    s->pairs[s->np].owner = o;
    s->np++;
}

/* kcounter.cpp:212-245 (flush_buffer with HITTER) and :176-210
 * (send2sendbuf): the bucket's k-mers sorted and run-length reduced; a k-mer
 * seen once or twice goes to its owner as that many plain k-mers, one seen
 * three or more times as one (k-mer, count) pair.  An empty bucket sends
 * nothing (the reference would send kcount_buffer[0] with count 1). */
static void hit_flush(uint64_t *buf, uint64_t *tmp, uint64_t nb, int k, int kw, uint32_t num_pes,
                      int32_t owner_bin, uint64_t seed, hit_streams *s, dko_hitter_stats *st) {
    if (nb == 0) return;
    st->buckets++;
    sort_keys(buf, tmp, nb, k);
    uint64_t run = 1;
    for (uint64_t i = 1; i <= nb; i++) {
        if (i < nb && key_at(buf, i, kw) == key_at(buf, i - 1, kw)) {
            run++;
            continue;
        }
        const u128 x = key_at(buf, i - 1, kw);
        const uint32_t o = num_pes > 1 ? owner_of(x, kw, num_pes, seed) : 0;
        if (num_pes <= 1 || owner_bin < 0 || o == (uint32_t)owner_bin) {
            if (run <= 2) {
                for (uint64_t c = 0; c < run; c++) hit_put_key(s, x);
                st->normal_keys += run;
            } else {
                hit_put_pair(s, x, run, o);
                st->heavy_pairs++;
            }
        }
        run = 1;
    }
}

int dko_count_hitter(const char *reads, size_t nbytes, int read_len, int k, int canonical, uint32_t num_pes,
                     int32_t owner_bin, uint64_t seed, uint32_t bucket_size, uint64_t **light_keys,
                     uint32_t **light_counts, uint64_t *d_light, uint64_t *light_offsets, uint64_t **heavy_keys,
                     uint32_t **heavy_counts, uint64_t *d_heavy, uint64_t *heavy_offsets, dko_hitter_stats *st) {
    if (!light_keys || !light_counts || !d_light || !heavy_keys || !heavy_counts || !d_heavy || !st)
        return DKO_E_INVAL;
    *light_keys = *heavy_keys = NULL;
    *light_counts = *heavy_counts = NULL;
    *d_light = *d_heavy = 0;
    memset(st, 0, sizeof(*st));
    if (k < 1 || k > 64 || num_pes < 1 || read_len < 1) return DKO_E_INVAL;
    if (num_pes > 1 && owner_bin >= (int32_t)num_pes) return DKO_E_INVAL;
    if (bucket_size == 0) bucket_size = 10000; /* KCOUNT_BUCKET_SIZE, dakc/Makefile:11 */
    /* the while condition of read_till_buf_max compares with bucket_size -
     * READLEN: below 0 it would never stop filling the buffer */
    if (bucket_size <= (uint32_t)read_len) return DKO_E_INVAL;
    seed = seed_or_default(seed);
    int64_t nrec = count_records(reads, nbytes, read_len);
    if (nrec < 0) return (int)-nrec;
    const int kw = key_words_for(k);
    const size_t stride = (size_t)read_len + 1;
    const uint64_t per_read = read_len >= k ? (uint64_t)(read_len - k + 1) : 0;
    const uint64_t bcap = bucket_size + per_read + 1;
    uint64_t *buf = (uint64_t *)malloc((size_t)bcap * kw * sizeof(uint64_t));
    uint64_t *tmp = (uint64_t *)malloc((size_t)bcap * kw * sizeof(uint64_t));
    uint8_t *codes = (uint8_t *)malloc((size_t)read_len + 1);
    hit_streams s;
    memset(&s, 0, sizeof(s));
    s.kw = kw;
    if (!buf || !tmp || !codes) {
        free(buf); free(tmp); free(codes);
        return DKO_E_NOMEM;
    }
    /* kcounter.cpp:247-305 with read_till_buf_max (kcounter_funcs.cpp:69-111):
     * records are parsed into the bucket while it holds at most
     * bucket_size - READLEN k-mers; then the bucket is flushed */
    uint64_t r = 0;
    while (r < (uint64_t)nrec) {
        emit_ctx c = {k, canonical, kw, 1, -1, seed, buf, bcap, 0};
        while (c.n <= (uint64_t)(bucket_size - (uint32_t)read_len) && r < (uint64_t)nrec) {
            parse_record(&c, reads + r * stride, read_len, codes);
            r++;
        }
        hit_flush(buf, tmp, c.n, k, kw, num_pes, owner_bin, seed, &s, st);
    }
    free(buf); free(tmp); free(codes);
    if (s.oom) {
        free(s.keys); free(s.pairs);
        return DKO_E_NOMEM;
    }
    /* the owner side, kcounter.cpp:325-384, per owner in owner order:
     * heavy pairs sorted and merged (sort_and_merge_duplicate_kmer_packets,
     * :56-104); the plain k-mers sorted and run-length reduced; a run whose
     * k-mer is heavy adds to the heavy count (binary_search, :60-73),
     * every other run is a light entry */
    const int all = num_pes > 1 && owner_bin < 0;
    const uint32_t P = all ? num_pes : 1;
    if (s.np) qsort(s.pairs, (size_t)s.np, sizeof(hit_pair), hit_pair_cmp);  /* no pairs: s.pairs may be NULL */
    uint64_t hd = 0;  /* merged heavy pairs, in place */
    for (uint64_t i = 0; i < s.np; i++) {
        if (hd && s.pairs[hd - 1].key == s.pairs[i].key && s.pairs[hd - 1].owner == s.pairs[i].owner)
            s.pairs[hd - 1].count += s.pairs[i].count;
        else
            s.pairs[hd++] = s.pairs[i];
    }
    /* plain k-mers grouped by owner, each group sorted */
    uint64_t *own = (uint64_t *)calloc((size_t)P + 1, sizeof(uint64_t));
    uint64_t *grp = (uint64_t *)malloc(((size_t)s.n + 1) * kw * sizeof(uint64_t));
    uint64_t *gtmp = (uint64_t *)malloc(((size_t)s.n + 1) * kw * sizeof(uint64_t));
    uint64_t *lk = (uint64_t *)malloc(((size_t)s.n + 1) * kw * sizeof(uint64_t));
    uint32_t *lc = (uint32_t *)malloc(((size_t)s.n + 1) * sizeof(uint32_t));
    uint64_t *hk = (uint64_t *)malloc(((size_t)hd + 1) * kw * sizeof(uint64_t));
    uint32_t *hc = (uint32_t *)malloc(((size_t)hd + 1) * sizeof(uint32_t));
    if (!own || !grp || !gtmp || !lk || !lc || !hk || !hc) {
        free(own); free(grp); free(gtmp); free(lk); free(lc); free(hk); free(hc); free(s.keys); free(s.pairs);
        return DKO_E_NOMEM;
    }
    for (uint64_t i = 0; i < s.n; i++)
        own[(all ? owner_of(key_at(s.keys, i, kw), kw, num_pes, seed) : 0) + 1]++;
    for (uint32_t o = 0; o < P; o++) own[o + 1] += own[o];
    {
        uint64_t *pos = (uint64_t *)malloc((size_t)P * sizeof(uint64_t));
        if (!pos) {
            free(own); free(grp); free(gtmp); free(lk); free(lc); free(hk); free(hc); free(s.keys); free(s.pairs);
            return DKO_E_NOMEM;
        }
        memcpy(pos, own, (size_t)P * sizeof(uint64_t));
        for (uint64_t i = 0; i < s.n; i++) {
            const u128 x = key_at(s.keys, i, kw);
            key_put(grp, pos[all ? owner_of(x, kw, num_pes, seed) : 0]++, kw, x);
        }
        free(pos);
    }
    free(s.keys);
    uint64_t dl = 0, h0 = 0;
    for (uint32_t o = 0; o < P; o++) {
        uint64_t h1 = h0;
        while (h1 < hd && (!all || s.pairs[h1].owner == o)) h1++;
        if (light_offsets) light_offsets[o] = dl;
        if (heavy_offsets) heavy_offsets[o] = h0;
        const uint64_t b0 = own[o], b1 = own[o + 1];
        sort_keys(grp + b0 * kw, gtmp, b1 - b0, k);
        uint64_t run = 1;
        for (uint64_t i = b0 + 1; i <= b1 && b1 > b0; i++) {
            if (i < b1 && key_at(grp, i, kw) == key_at(grp, i - 1, kw)) {
                run++;
                continue;
            }
            const u128 x = key_at(grp, i - 1, kw);
            uint64_t lo = h0, hi = h1; /* lower bound of x among this owner's heavy pairs */
            while (lo < hi) {
                const uint64_t mid = lo + (hi - lo) / 2;
                if (s.pairs[mid].key < x) lo = mid + 1;
                else hi = mid;
            }
            if (lo < h1 && s.pairs[lo].key == x) {
                s.pairs[lo].count += run;
            } else {
                key_put(lk, dl, kw, x);
                lc[dl++] = run > 0xFFFFFFFFULL ? 0xFFFFFFFFu : (uint32_t)run;
            }
            run = 1;
        }
        h0 = h1;
    }
    if (light_offsets) light_offsets[P] = dl;
    if (heavy_offsets) heavy_offsets[P] = hd;
    for (uint64_t i = 0; i < hd; i++) {
        key_put(hk, i, kw, s.pairs[i].key);
        hc[i] = s.pairs[i].count > 0xFFFFFFFFULL ? 0xFFFFFFFFu : (uint32_t)s.pairs[i].count;
    }
    free(own); free(grp); free(gtmp); free(s.pairs);
    *light_keys = lk;
    *light_counts = lc;
    *d_light = dl;
    *heavy_keys = hk;
    *heavy_counts = hc;
    *d_heavy = hd;
    return DKO_OK;
}
