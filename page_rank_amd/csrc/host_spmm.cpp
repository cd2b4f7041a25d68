// Host (CPU) SpMM of a PBR matrix: Y = A X for X, Y row-major in host memory. It is the
// product's counterpart of the reference's pbr_batched_matmul_cpu
// (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:17-68), which pbr_matmul calls for CPU
// tensors (Backend/graph_link/graph_link/__init__.py:130-134).
//
// Same arithmetic as the reference, element for element: every output row first takes the
// products of its stripe's tiles, tile by tile in the matrix's tile order and inside a tile
// in bit order (row-major), then the products of its remainder entries in CSR order; each
// product is rounded and added on its own (no fused multiply-add: the file is compiled with
// -ffp-contract=off for the baseline x86-64 target). The reference walks all
// tiles and then all remainder rows on one thread; here threads own disjoint ranges of
// stripes, which changes nothing in any row's order of additions, so Y is bit-identical to
// the reference's for every matrix whose set bits lie inside the matrix (all builders'
// output; values are located through block_offsets rather than one running counter).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/graph_link_hip.h"

namespace gl {
void set_error(const char* fmt, ...);
}

namespace {

template <typename I, typename S>
struct HostPbr {
  int64_t rows, cols, br, bc, T;
  const uint64_t* codes;
  const I* coords;
  const I* offsets;
  const S* data;
  const I* rptr;
  const I* rcol;
  const S* rval;
};

// Rows [s0 * br, s1 * br) of Y: zero them, add the tiles of stripes s0..s1-1 (tile lists
// `order[sp[s] .. sp[s+1])`, each in the matrix's tile order), then the remainder rows.
template <typename I, typename S>
void stripes(const HostPbr<I, S>& m, const std::vector<int64_t>& sp,
             const std::vector<int64_t>& order, const S* X, int64_t ldx, int64_t K, S* Y,
             int64_t ldy, int64_t s0, int64_t s1) {
  const int64_t r_lo = s0 * m.br;
  const int64_t r_hi = std::min<int64_t>(s1 * m.br, m.rows);
  for (int64_t r = r_lo; r < r_hi; ++r) std::memset(Y + r * ldy, 0, sizeof(S) * K);
  for (int64_t s = s0; s < s1; ++s) {
    for (int64_t q = sp[s]; q < sp[s + 1]; ++q) {
      const int64_t t = order[q];
      const uint64_t code = m.codes[t];
      const int64_t row0 = (int64_t)m.coords[2 * t];
      const int64_t col0 = (int64_t)m.coords[2 * t + 1];
      int64_t vi = (int64_t)m.offsets[t];
      for (int64_t r = 0; r < m.br; ++r) {
        const int64_t gr = row0 + r;
        if (gr >= m.rows) break;  // (rows past the edge hold no entries)
        S* y = Y + gr * ldy;
        for (int64_t c = 0; c < m.bc; ++c) {
          if (!((code >> (r * m.bc + c)) & 1ull)) continue;
          const int64_t gc = col0 + c;
          const S v = m.data[vi++];
          if (gc >= m.cols) continue;
          const S* x = X + gc * ldx;
          for (int64_t k = 0; k < K; ++k) y[k] += v * x[k];
        }
      }
    }
  }
  for (int64_t r = r_lo; r < r_hi; ++r) {
    S* y = Y + r * ldy;
    for (int64_t j = (int64_t)m.rptr[r]; j < (int64_t)m.rptr[r + 1]; ++j) {
      const S v = m.rval[j];
      const S* x = X + (int64_t)m.rcol[j] * ldx;
      for (int64_t k = 0; k < K; ++k) y[k] += v * x[k];
    }
  }
}

template <typename I, typename S>
int host_spmm(int64_t rows, int64_t cols, int32_t br, int32_t bc, int64_t T,
              const uint64_t* codes, const I* coords, const I* offsets, const S* data,
              const I* rptr, const I* rcol, const S* rval, const S* X, int64_t ldx, int64_t K,
              S* Y, int64_t ldy, int32_t threads) {
  if (rows < 0 || cols < 0 || T < 0 || K < 1) {
    gl::set_error("bad sizes (rows=%lld cols=%lld tiles=%lld K=%lld)", (long long)rows,
                  (long long)cols, (long long)T, (long long)K);
    return GL_ERR_INVALID;
  }
  if (br < 1 || bc < 1 || br * bc > 64) {
    gl::set_error("block size %dx%d does not fit a 64-bit code", br, bc);
    return GL_ERR_INVALID;
  }
  if (ldx < K || ldy < K) {
    gl::set_error("leading dimension smaller than K");
    return GL_ERR_INVALID;
  }
  if ((T > 0 && !(codes && coords && offsets && data)) || (rows > 0 && !rptr) ||
      (rows > 0 && !Y) || (cols > 0 && !X)) {
    gl::set_error("null array");
    return GL_ERR_INVALID;
  }
  if (rows == 0) return GL_OK;
  const HostPbr<I, S> m{rows, cols, br, bc, T, codes, coords, offsets, data, rptr, rcol, rval};
  const int64_t ns = (rows + br - 1) / br;
  // tiles grouped by stripe, keeping the matrix's tile order inside each stripe (a stable
  // counting sort: any tile order works, the builders' stripe-ascending one included)
  std::vector<int64_t> sp(ns + 1, 0), order(T);
  for (int64_t t = 0; t < T; ++t) {
    const int64_t s = (int64_t)coords[2 * t] / br;
    if (s < 0 || s >= ns || (int64_t)coords[2 * t] % br != 0) {
      gl::set_error("tile %lld: row0 %lld is not a stripe start inside the matrix",
                    (long long)t, (long long)coords[2 * t]);
      return GL_ERR_INVALID;
    }
    ++sp[s + 1];
  }
  for (int64_t s = 0; s < ns; ++s) sp[s + 1] += sp[s];
  {
    std::vector<int64_t> fill(sp.begin(), sp.end() - 1);
    for (int64_t t = 0; t < T; ++t) order[fill[(int64_t)coords[2 * t] / br]++] = t;
  }
  int64_t nt = threads > 0 ? threads : (int64_t)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  // about 64k multiply-adds per thread at least, or the thread costs more than it saves
  const int64_t work = (rows + T * br * bc) * K;
  nt = std::min<int64_t>(nt, std::max<int64_t>(1, work / 65536));
  nt = std::min<int64_t>(nt, ns);
  if (nt <= 1) {
    stripes(m, sp, order, X, ldx, K, Y, ldy, 0, ns);
    return GL_OK;
  }
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int64_t i = 0; i < nt; ++i) {
    const int64_t a = ns * i / nt, b = ns * (i + 1) / nt;
    pool.emplace_back([&, a, b] { stripes(m, sp, order, X, ldx, K, Y, ldy, a, b); });
  }
  for (auto& th : pool) th.join();
  return GL_OK;
}

}  // namespace

extern "C" {

#define GL_HOST_SPMM(NAME, I, S)                                                            \
  int NAME(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,              \
           int64_t num_tiles, const uint64_t* codes, const I* coords, const I* offsets,     \
           const S* data, const I* rem_indptr, const I* rem_col_ind, const S* rem_vals,     \
           const S* X, int64_t ldx, int64_t K, S* Y, int64_t ldy, int32_t threads) {        \
    return host_spmm<I, S>(rows, cols, block_rows, block_cols, num_tiles, codes, coords,    \
                           offsets, data, rem_indptr, rem_col_ind, rem_vals, X, ldx, K, Y,  \
                           ldy, threads);                                                   \
  }
GL_HOST_SPMM(gl_pbr_spmm_host_i32_f32, int32_t, float)
GL_HOST_SPMM(gl_pbr_spmm_host_i64_f32, int64_t, float)
GL_HOST_SPMM(gl_pbr_spmm_host_i32_f64, int32_t, double)
GL_HOST_SPMM(gl_pbr_spmm_host_i64_f64, int64_t, double)
#undef GL_HOST_SPMM

}  // extern "C"
