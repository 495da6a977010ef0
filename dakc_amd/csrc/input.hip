// input.hip -- the step before the count, on the device: FASTQ (4-line
// records) or newline-delimited reads of any lengths become the fixed-width
// format the count takes, as syngen/fq2txtmaker.sh:12-18 does with awk: the
// second line of each FASTQ record (NR % 4 == 2; every line for plain lines),
// padded to the longest such line, every space made 'M', each followed by
// '\n'.  Lines are split at '\n' only, and a last line without its '\n' is
// still a line (awk's record rules).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "dk_api_internal.hpp"
#include "dk_device.hpp"

namespace dk {

constexpr int LN_NT = 256;
constexpr int LN_ROWS = 16;                       // words per thread per tile
constexpr uint64_t LN_TILE = (uint64_t)LN_NT * LN_ROWS * 4;  // bytes per tile
constexpr int LN_NW = LN_NT / 64;

// newline mask of the 4 bytes at byte offset b (bytes past the end are not
// newlines); `in` is 4-byte aligned
DK_DEV uint32_t nl_mask(const uint8_t *in, uint64_t nbytes, uint64_t b) {
  if (b >= nbytes) return 0;
  uint32_t w;
  if (b + 4 <= nbytes) {
    w = *reinterpret_cast<const uint32_t *>(in + b);
  } else {
    w = 0;  // bytes past the end read as 0, never '\n'
    for (uint64_t i = b; i < nbytes; ++i) w |= (uint32_t)in[i] << (8 * (i - b));
  }
  uint32_t m = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) m |= (((w >> (8 * i)) & 0xFFu) == '\n' ? 1u : 0u) << i;
  return m;
}

// total newlines (grid-stride; one atomic per wave)
__global__ __launch_bounds__(LN_NT) void k_count_nl(const uint8_t *in, uint64_t nbytes,
                                                    unsigned long long *total) {
  unsigned long long c = 0;
  const uint64_t stride = (uint64_t)gridDim.x * LN_NT * 4;
  for (uint64_t b = ((uint64_t)blockIdx.x * LN_NT + threadIdx.x) * 4; b < nbytes; b += stride)
    c += __popc(nl_mask(in, nbytes, b));
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (__lane_id() == 0 && c) atomicAdd(total, c);
}

DK_DEV uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

struct LinesArgs {
  const uint8_t *in;
  uint64_t nbytes;
  uint64_t *nl;     // newline byte offsets, in order
  uint64_t *flags;  // look-back words, one per tile, zeroed
  uint32_t *ticket;
  uint32_t *err;    // ERR_STALL when a look-back wait gave up
};

// Newline offsets in byte order: tiles in ticket order, each thread reading
// words u * NT + tid of its tile (row u is 1 KB of consecutive bytes), ranks
// by row and wave prefix, the tile's offset by decoupled look-back.
__global__ __launch_bounds__(LN_NT) void k_find_nl(LinesArgs a) {
  __shared__ uint32_t s_row[LN_ROWS][LN_NW];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_excl;
  const uint32_t tid = threadIdx.x, w = tid / 64, lane = __lane_id();
  if (tid == 0) s_tile = atomicAdd(a.ticket, 1u);
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
  const uint64_t b0 = (uint64_t)tile * LN_TILE;
  uint32_t mask[LN_ROWS], excl[LN_ROWS];
#pragma unroll
  for (int u = 0; u < LN_ROWS; ++u) {
    mask[u] = nl_mask(a.in, a.nbytes, b0 + ((uint64_t)u * LN_NT + tid) * 4);
    const uint32_t c = (uint32_t)__popc(mask[u]);
    const uint32_t inc = wave_incl_scan(c);
    excl[u] = inc - c;
    if (lane == 63) s_row[u][w] = inc;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int u = 0; u < LN_ROWS; ++u)
      for (int v = 0; v < LN_NW; ++v) {
        const uint32_t x = s_row[u][v];
        s_row[u][v] = run;
        run += x;
      }
    const uint32_t epoch = 1;
    if (tile == 0) {
      lb_store(a.flags, lb_pack(epoch, LB_INC, run));
      s_excl = 0;
    } else {
      lb_store(a.flags + tile, lb_pack(epoch, LB_AGG, run));
      const uint64_t ex = lb_lookback(a.flags, tile, 1, 0, epoch, a.err);
      lb_store(a.flags + tile, lb_pack(epoch, LB_INC, ex + run));
      s_excl = ex;
    }
  }
  __syncthreads();
  const uint64_t base = s_excl;
#pragma unroll
  for (int u = 0; u < LN_ROWS; ++u) {
    uint64_t r = base + s_row[u][w] + excl[u];
    const uint64_t b = b0 + ((uint64_t)u * LN_NT + tid) * 4;
    for (uint32_t m = mask[u]; m; m &= m - 1) a.nl[r++] = b + (uint32_t)__ffs(m) - 1;
  }
}

struct ExpandArgs {
  const uint8_t *in;
  uint64_t nbytes;
  const uint64_t *nl;
  uint64_t nnl;        // newlines
  uint64_t nrec;       // records out
  uint32_t step, first;  // record r is line first + step * r
  uint32_t L;          // record width
  uint8_t *out;        // nrec * (L + 1) bytes
  unsigned int *maxlen;
};

DK_DEV void line_span(const ExpandArgs &a, uint64_t i, uint64_t &s, uint64_t &e) {
  s = i ? a.nl[i - 1] + 1 : 0;
  e = i < a.nnl ? a.nl[i] : a.nbytes;
}

// longest selected line (one thread per record)
__global__ __launch_bounds__(256) void k_max_line(ExpandArgs a) {
  unsigned int m = 0;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < a.nrec; r += stride) {
    uint64_t s, e;
    line_span(a, a.first + (uint64_t)a.step * r, s, e);
    const uint64_t len = e - s;
    m = max(m, len > 0xFFFFFFFFull ? 0xFFFFFFFFu : (unsigned int)len);
  }
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned int)__shfl_xor(m, off));
  if (__lane_id() == 0 && m) atomicMax(a.maxlen, m);
}

// one wavefront per record: copy the line, spaces to 'M', pad with 'M', '\n'
__global__ __launch_bounds__(256) void k_expand(ExpandArgs a) {
  const uint64_t wave = ((uint64_t)blockIdx.x * 256 + threadIdx.x) / 64;
  const uint64_t nwaves = (uint64_t)gridDim.x * 4;
  const uint32_t lane = __lane_id();
  for (uint64_t r = wave; r < a.nrec; r += nwaves) {
    uint64_t s, e;
    line_span(a, a.first + (uint64_t)a.step * r, s, e);
    const uint64_t len = e - s;
    uint8_t *o = a.out + r * ((uint64_t)a.L + 1);
    for (uint32_t p = lane; p <= a.L; p += 64) {
      uint8_t c = 'M';
      if (p == a.L) c = '\n';
      else if (p < len) {
        c = a.in[s + p];
        if (c == ' ') c = 'M';
      }
      o[p] = c;
    }
  }
}

}  // namespace dk

using namespace dk;

// Convert on ctx's stream; in and out are device memory.  The records go to
// `grow` (grown to fit) when it is given, else to out (cap bytes; nullptr:
// sizes only).  Scratch comes from ctx->post_a and ctx->conv_nl.
int dk::convert_reads(dakc_ctx *c, const uint8_t *in, uint64_t nbytes, int format, uint32_t read_len,
                      uint8_t *out, uint64_t cap, DevBuf *grow, uint64_t *out_nbytes, uint32_t *out_read_len) {
  if (format != DAKC_FMT_FASTQ && format != DAKC_FMT_LINES) return DAKC_E_INVAL;
  if (nbytes && ((uintptr_t)in & 3)) return DAKC_E_INVAL;
  if (read_len > (uint32_t)MAX_READ_LEN) return DAKC_E_INVAL;
  hipStream_t st = c->stream;
  int s;
  // newline count, then offsets
  const uint64_t ntiles = (nbytes + LN_TILE - 1) / LN_TILE;
  if (ntiles > 0xFFFFFFFFull) return DAKC_E_TOOBIG;
  if ((s = c->post_a.ensure(ntiles * 8 + 64, st, false))) return s;
  uint64_t *flags = (uint64_t *)c->post_a.p;
  // [0] total nl, [1] max len, [2] ticket, [3] error word
  auto *meta = (unsigned long long *)((uint8_t *)c->post_a.p + ntiles * 8);
  uint32_t *ticket = (uint32_t *)(meta + 2);
  uint32_t *err = (uint32_t *)(meta + 3);
  launch_fill_u64((uint64_t *)c->post_a.p, 0, ntiles + 8, st);
  if (nbytes) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((nbytes + LN_NT * 64 - 1) / (LN_NT * 64),
                                                       (uint64_t)c->num_cus * 8);
    k_count_nl<<<grid, LN_NT, 0, st>>>(in, nbytes, meta);
  }
  unsigned long long nnl = 0;
  uint8_t last = '\n';
  if (hipMemcpyAsync(&nnl, meta, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      (nbytes && hipMemcpyAsync(&last, in + nbytes - 1, 1, hipMemcpyDeviceToHost, st) != hipSuccess) ||
      hipStreamSynchronize(st) != hipSuccess)
    return DAKC_E_DEVICE;
  if ((s = c->conv_nl.ensure(std::max<uint64_t>(nnl, 1) * 8, st, false))) return s;
  LinesArgs la{in, nbytes, (uint64_t *)c->conv_nl.p, flags, ticket, err};
  if (nbytes) k_find_nl<<<(uint32_t)ntiles, LN_NT, 0, st>>>(la);
  // lines: one per newline, plus a last one without its '\n'
  const uint64_t nlines = nnl + (nbytes && last != '\n' ? 1 : 0);
  ExpandArgs ea;
  ea.in = in;
  ea.nbytes = nbytes;
  ea.nl = (const uint64_t *)c->conv_nl.p;
  ea.nnl = nnl;
  ea.step = format == DAKC_FMT_FASTQ ? 4 : 1;
  ea.first = format == DAKC_FMT_FASTQ ? 1 : 0;
  ea.nrec = format == DAKC_FMT_FASTQ ? (nlines + 2) / 4 : nlines;  // lines 1, 5, 9, ... (NR % 4 == 2)
  ea.maxlen = (unsigned int *)(meta + 1);
  ea.out = nullptr;  // set once the output buffer is known
  if (ea.nrec) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((ea.nrec + 255) / 256, (uint64_t)c->num_cus * 8);
    k_max_line<<<grid, 256, 0, st>>>(ea);
  }
  unsigned long long m3[3];  // max len, ticket, error word
  if (hipMemcpyAsync(m3, meta + 1, 24, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return DAKC_E_DEVICE;
  if ((s = status_of_err((uint32_t)m3[2]))) return s;
  const unsigned int maxlen = ea.nrec ? (unsigned int)m3[0] : 0u;
  if (read_len && maxlen > read_len) return DAKC_E_FORMAT;  // a read longer than the record width
  if (!read_len && maxlen > (unsigned int)MAX_READ_LEN) return DAKC_E_TOOBIG;
  // the longest read (at least 1: empty reads become "M\n", which hold no k-mer)
  const uint32_t L = read_len ? read_len : std::max(1u, maxlen);
  const uint64_t need = ea.nrec * ((uint64_t)L + 1);
  if (out_nbytes) *out_nbytes = need;
  if (out_read_len) *out_read_len = L;
  if (grow) {
    if ((s = grow->ensure(std::max<uint64_t>(need, 4), st, false))) return s;
    out = (uint8_t *)grow->p;
    cap = grow->cap;
  }
  if (!out) return DAKC_OK;
  if (cap < need) return DAKC_E_INVAL;
  ea.out = out;
  ea.L = L;
  if (ea.nrec) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((ea.nrec + 3) / 4, (uint64_t)c->num_cus * 16);
    k_expand<<<grid, 256, 0, st>>>(ea);
  }
  if (hipGetLastError() != hipSuccess) return DAKC_E_DEVICE;
  return hipStreamSynchronize(st) == hipSuccess ? DAKC_OK : DAKC_E_DEVICE;
}

extern "C" int dakc_convert_reads(dakc_ctx *ctx, const void *in, uint64_t nbytes, int format,
                                  uint32_t read_len, void *out, uint64_t cap, uint64_t *out_nbytes,
                                  uint32_t *out_read_len) {
  if (!ctx || (nbytes && !in)) return DAKC_E_INVAL;
  hipSetDevice(ctx->device);
  return convert_reads(ctx, (const uint8_t *)in, nbytes, format, read_len, (uint8_t *)out, cap, nullptr,
                       out_nbytes, out_read_len);
}
