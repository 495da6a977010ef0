// dakc.hpp -- header-only C++ face of the C ABI in dakc.h, shaped like the
// reference's driver classes so a DAKC driver switches with few edits:
//
//   reference (dakc/src/main/main.cpp:40-49)      here
//   fqreader fq(file, READLEN, rank, size);       dakc::fqreader fq(file, read_len);
//   char* chunk = fq.read_file();                 std::vector<char> chunk = fq.read_file();
//   kmercounter km(chunk, vectordbg);             dakc::kmercounter km(ctx, chunk, k, read_len);
//   (table in km.lightdbg, never exported)        km.lightdbg() / km.size() / km.serialize()
//
// fqreader: dakc/src/fqreader/fqreader.hpp:22-44, fqreader.cpp:14-90 (single PE).
// kmercounter: dakc/src/kcounter/kcounter.hpp:73-113 (runs the count in its
// constructor, like perform_kcount, kcounter.cpp:247).
#ifndef DAKC_HPP
#define DAKC_HPP

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "dakc.h"

namespace dakc {

struct error : std::runtime_error {
  int status;
  error(int s, const std::string &where)
      : std::runtime_error(where + ": " + dakc_status_str(s)), status(s) {}
};

inline void check(int s, const char *where) {
  if (s != DAKC_OK) throw error(s, where);
}

class context {
 public:
  explicit context(int device = 0, const dakc_opts *opts = nullptr) {
    check(dakc_ctx_create(device, opts, &ctx_), "dakc_ctx_create");
  }
  ~context() { if (ctx_) dakc_ctx_destroy(ctx_); }
  context(const context &) = delete;
  context &operator=(const context &) = delete;
  dakc_ctx *get() const { return ctx_; }

 private:
  dakc_ctx *ctx_ = nullptr;
};

// Reads the whole file of read_len-byte records (the single-PE block of
// fqreader.cpp:43-52).  A missing file throws (the reference exits with 2).
class fqreader {
 public:
  fqreader(std::string filename, int read_len) : filename_(std::move(filename)), read_len_(read_len) {}
  std::vector<char> read_file() const {
    std::FILE *f = std::fopen(filename_.c_str(), "rb");
    if (!f) throw error(DAKC_E_INVAL, "could not open the input file " + filename_);
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t got;
    while ((got = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
    std::fclose(f);
    return buf;
  }
  int read_len() const { return read_len_; }

 private:
  std::string filename_;
  int read_len_;
};

// Counts one read chunk on construction and owns the sorted table.
class kmercounter {
 public:
  kmercounter(const context &ctx, const std::vector<char> &read_chunk, int k, int read_len,
              bool canonical = false, uint32_t num_pes = 1, int32_t owner_bin = -1,
              uint64_t hash_seed = 0) {
    dakc_job job{};
    job.reads = read_chunk.data();
    job.nbytes = read_chunk.size();
    job.format = DAKC_FMT_FIXED;
    job.read_len = read_len;
    job.k = k;
    job.canonical = canonical ? 1 : 0;
    job.num_pes = num_pes;
    job.owner_bin = owner_bin;
    job.hash_seed = hash_seed;  // 0: the reference's owner seed (kcounter.cpp:52)
    check(dakc_count(ctx.get(), &job, &table_), "dakc_count");
  }
  ~kmercounter() { dakc_table_free(&table_); }
  kmercounter(const kmercounter &) = delete;
  kmercounter &operator=(const kmercounter &) = delete;

  const dakc_table &lightdbg() const { return table_; }
  uint64_t size() const { return table_.n; }
  uint64_t total() const { return table_.total; }
  std::vector<unsigned char> serialize() const {
    std::vector<unsigned char> out(dakc_table_serialize(&table_, nullptr, 0));
    if (!out.empty()) dakc_table_serialize(&table_, out.data(), out.size());
    return out;
  }

 private:
  dakc_table table_{};
};

}  // namespace dakc

#endif
