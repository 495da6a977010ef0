// hpp_example -- the smallest driver of include/dakc.hpp, written as the
// reference's main (dakc/src/main/main.cpp:40-49) reads with the classes
// swapped: fqreader reads the record file, kmercounter counts it in its
// constructor, and the table is written out.
//
//   hpp_example <reads.txt> <k> <read_len> <canonical 0|1> <num_pes> <owner_bin> <out.bin>
#include <cstdio>
#include <cstdlib>

#include "dakc.hpp"

int main(int argc, char **argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s reads.txt k read_len canonical num_pes owner_bin out.bin\n", argv[0]);
    return 2;
  }
  const int k = std::atoi(argv[2]), read_len = std::atoi(argv[3]);
  try {
    dakc::context ctx(0);
    dakc::fqreader fq(argv[1], read_len);
    std::vector<char> chunk = fq.read_file();
    dakc::kmercounter km(ctx, chunk, k, read_len, std::atoi(argv[4]) != 0, (uint32_t)std::atoi(argv[5]),
                         (int32_t)std::atoi(argv[6]));
    const std::vector<unsigned char> bytes = km.serialize();
    std::FILE *f = std::fopen(argv[7], "wb");
    if (!f || std::fwrite(bytes.data(), 1, bytes.size(), f) != bytes.size()) {
      std::fprintf(stderr, "cannot write %s\n", argv[7]);
      if (f) std::fclose(f);
      return 2;
    }
    std::fclose(f);
    std::printf("total %llu distinct %llu\n", (unsigned long long)km.total(), (unsigned long long)km.size());
  } catch (const dakc::error &e) {
    std::fprintf(stderr, "hpp_example: %s\n", e.what());
    return 1;
  }
  return 0;
}
