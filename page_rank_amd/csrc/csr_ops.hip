// CSR utilities around the PPR path, on the device:
//   - column normalization of a transition matrix (A[target, source] -> column-stochastic),
//     bit-identical to the float32 host arithmetic of the reference's
//     normalize_transition_matrix (Backend/graph_link/graph_link/__init__.py:184-207), with
//     dangling columns reported as a mask instead of densified;
//   - PBR -> CSR (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:289-338).
#include <hipcub/hipcub.hpp>

#include "gl_common.hpp"

namespace gl {

// ---------------------------------------------------------------------------------------
// Column sums in the host's order. The host computes a float32 column sum with
// numpy.add.reduceat over the column's entries (rows ascending): the first entry plus the
// pairwise sum of the rest, where pairwise(n) is a plain left-to-right sum below 8 entries,
// 8 interleaved accumulators combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) plus a
// left-to-right tail up to 128 entries, and pairwise(h) + pairwise(n-h) above, with
// h = n/2 rounded down to a multiple of 8. Every float add below is one IEEE add in that
// order, so the sums match bit for bit.
// ---------------------------------------------------------------------------------------
__device__ float pw_leaf(const float* __restrict__ v, const int32_t* __restrict__ perm,
                         int64_t s, int64_t n) {
  if (n < 8) {
    float r = -0.0f;
    for (int64_t i = 0; i < n; ++i) r = __fadd_rn(r, v[perm[s + i]]);
    return r;
  }
  float r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = v[perm[s + j]];
  int64_t i = 8;
  for (; i < n - (n % 8); i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = __fadd_rn(r[j], v[perm[s + i + j]]);
  }
  float res = __fadd_rn(__fadd_rn(__fadd_rn(r[0], r[1]), __fadd_rn(r[2], r[3])),
                        __fadd_rn(__fadd_rn(r[4], r[5]), __fadd_rn(r[6], r[7])));
  for (; i < n; ++i) res = __fadd_rn(res, v[perm[s + i]]);
  return res;
}

__device__ float pw_sum(const float* __restrict__ v, const int32_t* __restrict__ perm,
                        int64_t s0, int64_t n0) {
  struct Frame {
    int64_t s, n;
    int state;
    float left;
  };
  Frame st[48];  // depth <= log2(2^31 / 128) + 1
  int sp = 0;
  st[0] = {s0, n0, 0, 0.f};
  float ret = 0.f;
  while (sp >= 0) {
    Frame& f = st[sp];
    if (f.n <= 128) {
      ret = pw_leaf(v, perm, f.s, f.n);
      --sp;
      continue;
    }
    int64_t h = f.n / 2;
    h -= h % 8;
    if (f.state == 0) {
      f.state = 1;
      st[sp + 1] = {f.s, h, 0, 0.f};
      ++sp;
    } else if (f.state == 1) {
      f.left = ret;
      f.state = 2;
      st[sp + 1] = {f.s + h, f.n - h, 0, 0.f};
      ++sp;
    } else {
      ret = __fadd_rn(f.left, ret);
      --sp;
    }
  }
  return ret;
}

__global__ void k_col_keys(int64_t nnz, const int32_t* __restrict__ idx, int32_t* __restrict__ keys,
                           int32_t* __restrict__ ids, int32_t* __restrict__ counts) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = idx[e];
    keys[e] = c;
    ids[e] = (int32_t)e;
    atomicAdd(&counts[c], 1);
  }
}

__global__ void k_col_sums(int64_t cols, const int32_t* __restrict__ colptr,
                           const float* __restrict__ vals, const int32_t* __restrict__ perm,
                           float* __restrict__ sums, float* __restrict__ inv,
                           uint8_t* __restrict__ dangling) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < cols;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p0 = colptr[c], n = colptr[c + 1] - p0;
    float s = 0.f;
    if (n == 1) s = vals[perm[p0]];
    else if (n > 1) s = __fadd_rn(vals[perm[p0]], pw_sum(vals, perm, p0 + 1, n - 1));
    if (sums) sums[c] = s;
    inv[c] = s > 0.f ? __fdiv_rn(1.0f, s) : 0.f;
    if (dangling) dangling[c] = s == 0.f ? 1 : 0;
  }
}

__global__ void k_col_scale(int64_t nnz, const int32_t* __restrict__ idx,
                            const float* __restrict__ vals, const float* __restrict__ inv,
                            float* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = __fmul_rn(vals[e], inv[idx[e]]);
}

static int nblk(int64_t n, int cap = 65536) {
  const int64_t b = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}
static int64_t a256(int64_t x) { return (x + 255) & ~(int64_t)255; }
static int key_bits(int64_t n) {
  int b = 1;
  while (b < 31 && ((int64_t)1 << b) < n) ++b;
  return b;
}

struct NormWs {
  int32_t* counts;  // cols + 1, then the exclusive scan (colptr)
  int32_t* keys[2];
  int32_t* ids[2];
  float* inv;
  void* tmp;
  size_t tmp_bytes;
  int64_t bytes;
};

static int norm_carve(char* base, int64_t cols, int64_t nnz, NormWs& w) {
  size_t sort_b = 0, scan_b = 0;
  GL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_b, (const int32_t*)nullptr,
                                            (int32_t*)nullptr, (const int32_t*)nullptr,
                                            (int32_t*)nullptr, (int)std::max<int64_t>(nnz, 1),
                                            0, key_bits(cols)));
  GL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_b, (int32_t*)nullptr,
                                          (int32_t*)nullptr, (int)(cols + 1)));
  int64_t off = 0;
  auto take = [&](int64_t n) {
    char* p = base ? base + off : nullptr;
    off += a256(n);
    return p;
  };
  w.counts = (int32_t*)take((cols + 1) * 4);
  w.keys[0] = (int32_t*)take(nnz * 4);
  w.keys[1] = (int32_t*)take(nnz * 4);
  w.ids[0] = (int32_t*)take(nnz * 4);
  w.ids[1] = (int32_t*)take(nnz * 4);
  w.inv = (float*)take(cols * 4);
  w.tmp_bytes = std::max(sort_b, scan_b);
  w.tmp = take((int64_t)w.tmp_bytes);
  w.bytes = off;
  return GL_OK;
}

// ---------------------------------------------------------------------------------------
// PBR -> CSR: per-row counts (tile bits + remainder), exclusive scan, unsorted fill with a
// cursor per row, then a per-row sort by column (no duplicate columns exist: tiles and
// remainder are disjoint), so the output is canonical CSR whatever the tile order.
// ---------------------------------------------------------------------------------------
// bad: bit 0 = a set tile bit outside the matrix, bit 1 = tile values outside data[V],
// bit 2 = a remainder entry outside the matrix or rem_col[R].
__global__ void k_p2c_count(int64_t rows, int64_t cols, int32_t br, int32_t bc, int64_t T,
                            int64_t V, int64_t R, const uint64_t* __restrict__ codes,
                            const int32_t* __restrict__ coords, const int32_t* __restrict__ offsets,
                            const int32_t* __restrict__ rem_ptr, const int32_t* __restrict__ rem_col,
                            int32_t* __restrict__ cnt, int32_t* __restrict__ bad) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < T;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t c = codes[t];
    const int64_t r0 = coords[2 * t], c0 = coords[2 * t + 1];
    const int64_t o = offsets[t];
    if (o < 0 || o + __popcll(c) > V) atomicOr(bad, 2);
    const uint64_t rowmask = bc >= 64 ? ~0ull : ((1ull << bc) - 1ull);
    for (int r = 0; r < br; ++r) {
      const uint64_t bits = (c >> (r * bc)) & rowmask;
      if (!bits) continue;
      const int hi = 63 - __builtin_clzll(bits);  // highest set column of this row
      if (r0 < 0 || c0 < 0 || r0 + r >= rows || c0 + hi >= cols) {
        atomicOr(bad, 1);
        continue;
      }
      atomicAdd(&cnt[r0 + r], __popcll(bits));
    }
  }
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = rem_ptr[r], b = rem_ptr[r + 1];
    if (a < 0 || b < a || b > R) {
      atomicOr(bad, 4);
      continue;
    }
    for (int32_t j = a; j < b; ++j)
      if (rem_col[j] < 0 || rem_col[j] >= cols) atomicOr(bad, 4);
    if (b > a) atomicAdd(&cnt[r], b - a);
  }
}

// Values are payloads (copied, never computed): V is float or double.
template <typename V>
__global__ void k_p2c_fill(int64_t rows, int32_t br, int32_t bc, int64_t T,
                           const uint64_t* __restrict__ codes, const int32_t* __restrict__ coords,
                           const int32_t* __restrict__ offsets, const V* __restrict__ data,
                           const int32_t* __restrict__ rem_ptr, const int32_t* __restrict__ rem_col,
                           const V* __restrict__ rem_val, int32_t* __restrict__ cursor,
                           int32_t* __restrict__ col_out, V* __restrict__ val_out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < T;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t c = codes[t];
    const int64_t r0 = coords[2 * t], c0 = coords[2 * t + 1];
    int32_t v = offsets[t];
    const uint64_t rowmask = bc >= 64 ? ~0ull : ((1ull << bc) - 1ull);
    for (int r = 0; r < br; ++r) {
      uint64_t bits = (c >> (r * bc)) & rowmask;
      if (!bits) continue;
      const int32_t n = __popcll(bits);
      int32_t pos = atomicAdd(&cursor[r0 + r], n);
      while (bits) {
        const int j = __builtin_ctzll(bits);
        bits &= bits - 1;
        col_out[pos] = (int32_t)(c0 + j);
        val_out[pos] = data[v];
        ++pos;
        ++v;
      }
    }
  }
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = rem_ptr[r], b = rem_ptr[r + 1];
    if (b == a) continue;
    int32_t pos = atomicAdd(&cursor[r], b - a);
    for (int32_t j = a; j < b; ++j, ++pos) {
      col_out[pos] = rem_col[j];
      val_out[pos] = rem_val[j];
    }
  }
}

template <typename V>
struct P2cWs {
  int32_t* cnt;     // rows + 1 counts, scanned in place into indptr order
  int32_t* cursor;  // rows
  int32_t* flag;
  int32_t* keys;    // nnz (unsorted columns)
  V* vals;          // nnz (unsorted values)
  void* tmp;
  size_t tmp_bytes;
  int64_t bytes;
};

template <typename V>
static int p2c_carve(char* base, int64_t rows, int64_t nnz, P2cWs<V>& w) {
  size_t sort_b = 0, scan_b = 0;
  GL_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(
      nullptr, sort_b, (const int32_t*)nullptr, (int32_t*)nullptr, (const V*)nullptr,
      (V*)nullptr, (int)std::max<int64_t>(nnz, 1), (int)std::max<int64_t>(rows, 1),
      (const int32_t*)nullptr, (const int32_t*)nullptr, 0, 31));
  GL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_b, (int32_t*)nullptr,
                                          (int32_t*)nullptr, (int)(rows + 1)));
  int64_t off = 0;
  auto take = [&](int64_t n) {
    char* p = base ? base + off : nullptr;
    off += a256(n);
    return p;
  };
  w.cnt = (int32_t*)take((rows + 1) * 4);
  w.cursor = (int32_t*)take(rows * 4);
  w.flag = (int32_t*)take(4);
  w.keys = (int32_t*)take(nnz * 4);
  w.vals = (V*)take(nnz * (int64_t)sizeof(V));
  w.tmp_bytes = std::max(sort_b, scan_b);
  w.tmp = take((int64_t)w.tmp_bytes);
  w.bytes = off;
  return GL_OK;
}

template <typename V>
int64_t p2c_workspace_bytes(int64_t rows, int64_t nnz) {
  if (rows < 0 || nnz < 0 || rows >= INT32_MAX || nnz >= INT32_MAX) return -1;
  P2cWs<V> w;
  if (p2c_carve<V>(nullptr, rows, nnz, w) != GL_OK) return -1;
  return w.bytes;
}

template <typename V>
int pbr_to_csr(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
               int64_t num_tiles, int64_t num_vals, int64_t num_rem, const uint64_t* codes,
               const int32_t* coords, const int32_t* offsets, const V* data,
               const int32_t* rem_indptr, const int32_t* rem_col_ind, const V* rem_vals,
               int32_t* indptr, int32_t* indices, V* vals, int64_t nnz, void* ws,
               int64_t ws_bytes, void* stream) {
  GL_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0 && rows < INT32_MAX && nnz < INT32_MAX,
             "bad sizes");
  GL_REQUIRE(block_rows >= 1 && block_cols >= 1 && block_rows * block_cols <= 64,
             "block %dx%d does not fit a 64-bit code", block_rows, block_cols);
  GL_REQUIRE(num_tiles >= 0 && num_vals >= 0 && num_rem >= 0, "bad sizes");
  GL_REQUIRE(num_tiles == 0 || (codes && coords && offsets && data), "null tile arrays");
  GL_REQUIRE(num_rem == 0 || (rem_col_ind && rem_vals), "null remainder arrays");
  GL_REQUIRE(rem_indptr && indptr && (nnz == 0 || (indices && vals)), "null CSR arrays");
  P2cWs<V> w;
  int rc = p2c_carve<V>((char*)ws, rows, nnz, w);
  if (rc) return rc;
  GL_REQUIRE(ws != nullptr && ws_bytes >= w.bytes, "workspace too small: %lld < %lld bytes",
             (long long)ws_bytes, (long long)w.bytes);
  hipStream_t st = (hipStream_t)stream;
  GL_HIP(hipMemsetAsync(w.cnt, 0, (rows + 1) * 4, st));
  GL_HIP(hipMemsetAsync(w.flag, 0, 4, st));
  hipLaunchKernelGGL(k_p2c_count, dim3(nblk(std::max(num_tiles, rows))), dim3(256), 0, st, rows,
                     cols, block_rows, block_cols, num_tiles, num_vals, num_rem, codes, coords,
                     offsets, rem_indptr, rem_col_ind, w.cnt, w.flag);
  GL_LAUNCH_CHECK("k_p2c_count");
  size_t tb = w.tmp_bytes;
  GL_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.cnt, indptr, (int)(rows + 1), st));
  int32_t h[2] = {0, 0};
  GL_HIP(hipMemcpyAsync(&h[0], indptr + rows, 4, hipMemcpyDeviceToHost, st));
  GL_HIP(hipMemcpyAsync(&h[1], w.flag, 4, hipMemcpyDeviceToHost, st));
  GL_HIP(hipStreamSynchronize(st));
  GL_REQUIRE(!(h[1] & 1), "a tile has a set bit outside the %lldx%lld matrix",
             (long long)rows, (long long)cols);
  GL_REQUIRE(!(h[1] & 2), "tile offsets point outside the %lld tile values", (long long)num_vals);
  GL_REQUIRE(!(h[1] & 4), "the remainder CSR has entries outside the matrix");
  GL_REQUIRE(h[0] == nnz, "the matrix holds %d entries but the output has room for %lld",
             h[0], (long long)nnz);
  if (nnz == 0) return GL_OK;
  GL_HIP(hipMemcpyAsync(w.cursor, indptr, rows * 4, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_p2c_fill<V>, dim3(nblk(std::max(num_tiles, rows))), dim3(256), 0, st, rows,
                     block_rows, block_cols, num_tiles, codes, coords, offsets, data, rem_indptr,
                     rem_col_ind, rem_vals, w.cursor, w.keys, w.vals);
  GL_LAUNCH_CHECK("k_p2c_fill");
  tb = w.tmp_bytes;
  GL_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(w.tmp, tb, w.keys, indices, w.vals, vals,
                                                     (int)nnz, (int)rows, indptr, indptr + 1, 0,
                                                     key_bits(cols), st));
  return GL_OK;
}

}  // namespace gl

using namespace gl;

extern "C" {

int64_t gl_csr_normalize_workspace_bytes(int64_t cols, int64_t nnz) {
  if (cols < 0 || nnz < 0 || cols >= INT32_MAX || nnz >= INT32_MAX) return -1;
  NormWs w;
  if (norm_carve(nullptr, cols, nnz, w) != GL_OK) return -1;
  return w.bytes;
}

int gl_csr_normalize_columns_f32(int64_t cols, int64_t nnz, const int32_t* indices,
                                 const float* vals, float* out_vals, float* col_sums,
                                 uint8_t* dangling, void* ws, int64_t ws_bytes, void* stream) {
  GL_REQUIRE(cols >= 0 && nnz >= 0 && cols < INT32_MAX && nnz < INT32_MAX, "bad sizes");
  GL_REQUIRE(nnz == 0 || (indices && vals && out_vals), "null CSR arrays");
  NormWs w;
  int rc = norm_carve((char*)ws, cols, nnz, w);
  if (rc) return rc;
  GL_REQUIRE(ws != nullptr && ws_bytes >= w.bytes, "workspace too small: %lld < %lld bytes",
             (long long)ws_bytes, (long long)w.bytes);
  hipStream_t st = (hipStream_t)stream;
  if (cols == 0) return GL_OK;
  GL_HIP(hipMemsetAsync(w.counts, 0, (cols + 1) * 4, st));
  if (nnz > 0) {
    hipLaunchKernelGGL(k_col_keys, dim3(nblk(nnz)), dim3(256), 0, st, nnz, indices, w.keys[0],
                       w.ids[0], w.counts);
    GL_LAUNCH_CHECK("k_col_keys");
    size_t tb = w.tmp_bytes;
    // stable LSD radix sort: entries of one column keep their CSR (row-ascending) order
    GL_HIP(hipcub::DeviceRadixSort::SortPairs(w.tmp, tb, w.keys[0], w.keys[1], w.ids[0],
                                              w.ids[1], (int)nnz, 0, key_bits(cols), st));
  }
  size_t tb = w.tmp_bytes;
  GL_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.counts, w.counts, (int)(cols + 1), st));
  hipLaunchKernelGGL(k_col_sums, dim3(nblk(cols)), dim3(256), 0, st, cols, w.counts, vals,
                     w.ids[1], col_sums, w.inv, dangling);
  GL_LAUNCH_CHECK("k_col_sums");
  if (nnz > 0) {
    hipLaunchKernelGGL(k_col_scale, dim3(nblk(nnz)), dim3(256), 0, st, nnz, indices, vals,
                       w.inv, out_vals);
    GL_LAUNCH_CHECK("k_col_scale");
  }
  return GL_OK;
}

int64_t gl_pbr_to_csr_workspace_bytes(int64_t rows, int64_t nnz) {
  return p2c_workspace_bytes<float>(rows, nnz);
}

int64_t gl_pbr_to_csr_f64_workspace_bytes(int64_t rows, int64_t nnz) {
  return p2c_workspace_bytes<double>(rows, nnz);
}

int gl_pbr_to_csr_i32_f32(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                          int64_t num_tiles, int64_t num_vals, int64_t num_rem,
                          const uint64_t* codes, const int32_t* coords,
                          const int32_t* offsets, const float* data, const int32_t* rem_indptr,
                          const int32_t* rem_col_ind, const float* rem_vals, int32_t* indptr,
                          int32_t* indices, float* vals, int64_t nnz, void* ws, int64_t ws_bytes,
                          void* stream) {
  return pbr_to_csr<float>(rows, cols, block_rows, block_cols, num_tiles, num_vals, num_rem,
                           codes, coords, offsets, data, rem_indptr, rem_col_ind, rem_vals,
                           indptr, indices, vals, nnz, ws, ws_bytes, stream);
}

int gl_pbr_to_csr_i32_f64(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                          int64_t num_tiles, int64_t num_vals, int64_t num_rem,
                          const uint64_t* codes, const int32_t* coords,
                          const int32_t* offsets, const double* data,
                          const int32_t* rem_indptr, const int32_t* rem_col_ind,
                          const double* rem_vals, int32_t* indptr, int32_t* indices,
                          double* vals, int64_t nnz, void* ws, int64_t ws_bytes, void* stream) {
  return pbr_to_csr<double>(rows, cols, block_rows, block_cols, num_tiles, num_vals, num_rem,
                            codes, coords, offsets, data, rem_indptr, rem_col_ind, rem_vals,
                            indptr, indices, vals, nnz, ws, ws_bytes, stream);
}

}  // extern "C"
