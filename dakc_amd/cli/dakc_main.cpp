// dakc -- command-line driver: file in, k, sorted (kmer, count) table out.
//
// Mirrors the reference's main (dakc/src/main/main.cpp:21-55) and arg_parser
// (dakc/src/main/parser.hpp:35-55), whose only runtime flag is -f; the values
// the reference fixes at compile time (KMERLEN, READLEN, dakc/Makefile:11) are
// flags here.  Unlike the reference, the table is written out (-o).
//
//   dakc -f reads.txt [-k 31] [-r 150] [-c] [-p P] [-b bin] [-s seed] [-o table.bin]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>

#include "dakc.hpp"

static void usage() {
  std::printf(
      "usage: dakc -f <reads.txt> [-k K] [-r READLEN] [-c] [-p P] [-b BIN] [-s SEED] [-o out.bin] [-d DEV]\n"
      "  -f  record file: READLEN bytes + '\\n' per read ('M'-padded, as syngen/fq2txtmaker.sh)\n"
      "  -k  k-mer length 1..64 (default 31)\n"
      "  -r  record width READLEN (default 150)\n"
      "  -c  canonical k-mers (min of k-mer and reverse complement)\n"
      "  -p  number of PEs P for the owner hash (default 1)\n"
      "  -b  count only the k-mers owned by PE BIN (default: all bins)\n"
      "  -s  owner-hash seed (default 0: the reference's 0x9E3779B97F4A7C15)\n"
      "  -o  write the table: u64 kmer (or u64 lo,u64 hi) + u32 count per entry\n");
}

int main(int argc, char **argv) {
  std::string file, out;
  int k = 31, read_len = 150, canonical = 0, dev = 0;
  long p = 1, bin = -1;
  unsigned long long seed = 0;
  int opt;
  while ((opt = getopt(argc, argv, "hf:k:r:cp:b:s:o:d:")) != -1) {
    switch (opt) {
    case 'f': file = optarg; break;
    case 'k': k = std::atoi(optarg); break;
    case 'r': read_len = std::atoi(optarg); break;
    case 'c': canonical = 1; break;
    case 'p': p = std::atol(optarg); break;
    case 'b': bin = std::atol(optarg); break;
    case 's': seed = std::strtoull(optarg, nullptr, 0); break;
    case 'o': out = optarg; break;
    case 'd': dev = std::atoi(optarg); break;
    case 'h':
    default: usage(); return opt == 'h' ? 0 : 2;
    }
  }
  if (file.empty()) {
    usage();
    return 2;
  }
  std::printf("file: %s  READLEN: %d  k: %d  canonical: %d  P: %ld  bin: %ld\n", file.c_str(),
              read_len, k, canonical, p, bin);
  try {
    dakc::context ctx(dev);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<char> chunk = dakc::fqreader(file, read_len).read_file();
    auto t1 = std::chrono::steady_clock::now();
    std::printf("Reading time: %.6f seconds\n", std::chrono::duration<double>(t1 - t0).count());
    dakc::kmercounter km(ctx, chunk, k, read_len, canonical != 0, (uint32_t)p, (int32_t)bin, seed);
    auto t2 = std::chrono::steady_clock::now();
    std::printf("kmer counting time: %.6f seconds\n", std::chrono::duration<double>(t2 - t1).count());
    std::printf("global_kmers: %llu\nglobal_distinct_kmers: %llu\n", (unsigned long long)km.total(),
                (unsigned long long)km.size());
    if (!out.empty()) {
      auto bytes = km.serialize();
      std::FILE *f = std::fopen(out.c_str(), "wb");
      if (!f || std::fwrite(bytes.data(), 1, bytes.size(), f) != bytes.size()) {
        std::fprintf(stderr, "cannot write %s\n", out.c_str());
        if (f) std::fclose(f);
        return 2;
      }
      std::fclose(f);
    }
  } catch (const dakc::error &e) {
    std::fprintf(stderr, "dakc: %s\n", e.what());
    return e.status == DAKC_E_INVAL && std::strstr(e.what(), "could not open") ? 2 : 1;
  }
  return 0;
}
