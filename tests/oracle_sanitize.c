/* Host sanitizer run of the oracle (test infrastructure): built with
 * -fsanitize=address,undefined together with oracle/dakc_oracle.c and run by
 * tests/test_oracle.py::test_sanitized.  It drives every oracle entry point
 * over messy seeded records (N and M bytes, lower case, invalid bytes, short
 * and empty records) and checks the cross-relations the oracle's tests rely
 * on: the threaded count's (n, d, checksum) against the single-PE table, the
 * threaded extraction's length, HITTER's light + heavy checksum against the
 * plain table's, the histogram total, the filter's kept count and the query
 * hits.  A sanitizer report aborts the run; a mismatch exits with status 1. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../oracle/dakc_oracle.h"

static uint64_t rng = 0x243F6A8885A308D3ull;
static uint32_t next32(void) {
  rng = rng * 6364136223846793005ull + 1442695040888963407ull;
  return (uint32_t)(rng >> 33);
}

/* n records of read_len bytes + '\n' */
static char *messy_reads(int n, int read_len) {
  static const char alpha[] = "ACGTACGTACGTacgtNNMnX\x01\xff";
  char *r = malloc((size_t)n * (read_len + 1) + 1);
  for (int i = 0; i < n; ++i) {
    char *p = r + (size_t)i * (read_len + 1);
    const int kind = (int)(next32() % 8);
    for (int j = 0; j < read_len; ++j) {
      const uint32_t v = next32();
      p[j] = kind == 0 ? alpha[v % (sizeof(alpha) - 1)] : "ACGT"[v & 3];
      if (kind == 1 && v % 97 == 0) p[j] = 'N';
      if (kind == 2 && j == read_len / 2) p[j] = 'M';
    }
    p[read_len] = '\n';
  }
  r[(size_t)n * (read_len + 1)] = 0;
  return r;
}

static int fails = 0;
#define CHECK(c, ...)                     \
  do {                                    \
    if (!(c)) {                           \
      fprintf(stderr, "FAIL: " __VA_ARGS__); \
      fputc('\n', stderr);                \
      ++fails;                            \
    }                                     \
  } while (0)

int main(void) {
  const int read_len = 60, nrec = 400;
  char *reads = messy_reads(nrec, read_len);
  const size_t nbytes = (size_t)nrec * (read_len + 1);
  const int ks[] = {1, 5, 16, 31, 32, 33, 55, 63, 64};
  const uint32_t pes[] = {1, 7, 64};
  for (size_t ki = 0; ki < sizeof(ks) / sizeof(ks[0]); ++ki) {
    const int k = ks[ki], kw = k <= 32 ? 1 : 2;
    for (int canon = 0; canon < 2; ++canon) {
      for (size_t pi = 0; pi < sizeof(pes) / sizeof(pes[0]); ++pi) {
        const uint32_t P = pes[pi];
        const int32_t bins[] = {-1, 0, 3};
        for (size_t bi = 0; bi < 3; ++bi) {
          const int32_t b = bins[bi];
          if (P == 1 && b > 0) continue;
          const uint64_t seed = bi == 2 ? 12345 : 0;
          uint64_t *keys = NULL, d = 0, n = 0, offs[65];
          uint32_t *counts = NULL;
          int s = dko_count(reads, nbytes, read_len, k, canon, P, b, seed, &keys, &counts, &d, &n,
                            b < 0 ? offs : NULL);
          CHECK(s == 0, "count k=%d status %d", k, s);
          const uint64_t cs = dko_table_checksum(keys, counts, d, kw);
          uint64_t n2 = 0, d2 = 0, cs2 = 0;
          double el = 0;
          s = dko_count_mt(reads, nbytes, read_len, k, canon, P, b, seed, 3, &n2, &d2, &cs2, &el);
          CHECK(s == 0 && n2 == n && d2 == d && cs2 == cs, "count_mt k=%d P=%u b=%d", k, P, b);
          const int64_t ne = dko_extract_mt(reads, nbytes, read_len, k, canon, P, b, seed, 3, NULL, 0);
          CHECK(ne == (int64_t)n, "extract_mt k=%d: %lld vs %llu", k, (long long)ne, (unsigned long long)n);
          if (d) {
            /* histogram total, filter's kept count, query hits */
            uint64_t hist[16];
            dko_histogram(counts, d, 16, hist);
            uint64_t tot = 0, below = 0;
            for (int i = 0; i < 16; ++i) tot += hist[i];
            for (uint64_t i = 0; i < d; ++i) below += counts[i] < 16;
            CHECK(tot == below, "histogram k=%d", k);
            uint32_t *q = malloc(d * sizeof(uint32_t));
            dko_query(keys, counts, d, kw, b < 0 ? offs : NULL, b < 0 ? P : 0, seed, keys, kw == 1 ? d : 0, q);
            if (kw == 1)
              for (uint64_t i = 0; i < d; ++i) CHECK(q[i] == counts[i], "query k=%d", k);
            free(q);
            uint64_t keep = 0;
            for (uint64_t i = 0; i < d; ++i) keep += counts[i] >= 2;
            uint64_t total = 0;
            const uint64_t kept = dko_filter(keys, counts, d, kw, 2, b < 0 ? offs : NULL, b < 0 ? P : 0, &total);
            CHECK(kept == keep, "filter k=%d", k);
          }
          dko_free(keys);
          dko_free(counts);
          if (b < 0 && k >= 16) {
            uint64_t *lk = NULL, *hk = NULL, dl = 0, dh = 0, lo[65], ho[65];
            uint32_t *lc = NULL, *hc = NULL;
            dko_hitter_stats st;
            s = dko_count_hitter(reads, nbytes, read_len, k, canon, P, b, seed, 400, &lk, &lc, &dl, lo, &hk, &hc,
                                 &dh, ho, &st);
            CHECK(s == 0, "hitter k=%d status %d", k, s);
            CHECK(dko_table_checksum(lk, lc, dl, kw) + dko_table_checksum(hk, hc, dh, kw) == cs,
                  "hitter union k=%d", k);
            dko_free(lk);
            dko_free(lc);
            dko_free(hk);
            dko_free(hc);
          }
        }
      }
    }
  }
  /* FASTQ in: four-line records of 30..49 bases, and the same bytes as plain lines */
  {
    char fq[4096];
    size_t m = 0;
    for (int i = 0; i < 20; ++i) {
      m += (size_t)snprintf(fq + m, sizeof(fq) - m, "@r%d\n", i);
      for (int j = 0; j < 30 + i; ++j) fq[m++] = "ACGTN"[next32() % 5];
      m += (size_t)snprintf(fq + m, sizeof(fq) - m, "\n+\n");
      for (int j = 0; j < 30 + i; ++j) fq[m++] = 'I';
      fq[m++] = '\n';
    }
    char out[8192];
    uint64_t on = 0;
    uint32_t w = 0;
    int s = dko_convert(fq, m, 1, 0, out, sizeof(out), &on, &w);
    /* out_n: bytes of fixed-width records, width: the longest read (49) */
    CHECK(s == 0 && w == 49 && on == 20ull * (w + 1), "convert fastq: status %d bytes %llu", s,
          (unsigned long long)on);
    s = dko_convert(fq, m, 0, 0, out, sizeof(out), &on, &w);
    CHECK(s == 0 && w == 49 && on == 80ull * (w + 1), "convert lines: status %d bytes %llu", s,
          (unsigned long long)on);
  }
  free(reads);
  if (fails) fprintf(stderr, "%d checks failed\n", fails);
  else printf("oracle sanitizer run: all checks passed\n");
  return fails ? 1 : 0;
}
