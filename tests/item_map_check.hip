// Host check of fin_item_of (dk_kernels.hpp): over many plan geometries the
// final items, in index order, cover every (bucket, second-level digit) exactly
// once and in order, within bounds.  Built and run by tests/test_item_map.py.
#include "dk_kernels.hpp"
#include <cstdio>
#include <vector>
using namespace dk;
int main() {
  long bad = 0, geos = 0;
  for (uint32_t b1 = 0; b1 <= 10; ++b1)
    for (uint32_t b2 = 0; b2 <= 10; b2 += (b2 < 3 ? 1 : 3))
      for (uint32_t owners : {1u, 3u, 8u})
        for (int canon = 0; canon < 2; ++canon)
          for (uint32_t base : {1u, 2u, 3u, 5u, 7u, 16u}) {
            if ((owners << b1) > 1024) continue;
            MsdArgs m{};
            const uint32_t lc = b1 < 2 ? b1 : 2, nd2 = 1u << b2;
            m.nb1 = owners << b1; m.nd2 = nd2; m.cls = lc | (b1 - lc) << 4;
            m.d3.nbits = 3;
            uint32_t per_block = 0;
            for (uint32_t c = 0; c < 4; ++c) {
              uint32_t pk = 0, ip = 0;
              if (c < (1u << lc)) {
                pk = canon ? base * (c + 1) : base;  // packs grow toward the top, as canonical ones do
                if (pk > 16) pk = 16;
                if (pk > nd2) pk = nd2;
                ip = (nd2 + pk - 1) / pk;
              }
              m.packc |= pk << (8 * c);
              m.ipbc[c >> 1] |= ip << (16 * (c & 1));
              per_block += ip << (b1 - lc);
            }
            const uint32_t n = owners * per_block;
            uint64_t expect = 0;  // next (b * nd2 + digit) to cover
            for (uint32_t it = 0; it < n; ++it) {
              const FinItem f = fin_item_of(m, it);
              const uint64_t at = (uint64_t)f.b * nd2 + f.s0;
              if (f.b >= m.nb1 || f.cnt == 0 || f.s0 + f.cnt > nd2 || at != expect ||
                  f.nbins != (f.cnt << 3)) {
                if (bad < 5) printf("bad b1=%u b2=%u P=%u c=%d base=%u it=%u -> b=%u s0=%u cnt=%u (expect %llu)\n",
                                    b1, b2, owners, canon, base, it, f.b, f.s0, f.cnt, (unsigned long long)expect);
                ++bad;
                break;
              }
              expect = at + f.cnt;
            }
            if (expect != (uint64_t)m.nb1 * nd2) {
              if (bad < 5) printf("coverage b1=%u b2=%u P=%u: %llu of %u\n", b1, b2, owners, (unsigned long long)expect, m.nb1 * nd2);
              ++bad;
            }
            ++geos;
          }
  printf("%ld geometries, %ld bad\n", geos, bad);
  return bad != 0;
}
