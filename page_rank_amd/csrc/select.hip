// Top-k per column of a dense row-major [N][K] float matrix (the ranks a PPR run returns),
// by radix selection on 64-bit keys:
//   key = order-preserving bits of the value << 32 | (0xFFFFFFFF - row)
// Keys are unique, so "the k largest keys" is exact and deterministic: larger value first,
// and among equal values the smaller row first. Eight passes of 8-bit digits (high to low)
// narrow the k-th largest key of every column; one pass then collects the k keys at or
// above it, and a segmented sort orders them.
#include <hipcub/hipcub.hpp>

#include "gl_common.hpp"

namespace gl {

__device__ __forceinline__ uint64_t sel_key(float v, int64_t row) {
  uint32_t u = __float_as_uint(v);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // total order of IEEE floats
  return ((uint64_t)u << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)row);
}

struct SelState {
  uint64_t prefix;  // key bits fixed so far (above `shift + 8`)
  int64_t need;     // keys still to take among those matching the prefix
};

// Histogram of digit `shift` over the keys of each column that match its prefix.
template <bool LDS>
__global__ void k_sel_hist(const float* __restrict__ x, int64_t N, int64_t K, int64_t ld,
                           int shift, const SelState* __restrict__ st,
                           uint32_t* __restrict__ hist) {
  extern __shared__ uint32_t sh[];
  if (LDS) {
    for (int64_t i = threadIdx.x; i < K * 256; i += blockDim.x) sh[i] = 0;
    __syncthreads();
  }
  const uint64_t hi_mask = shift >= 56 ? 0ull : (~0ull << (shift + 8));
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * K;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e / K, col = e % K;
    const uint64_t key = sel_key(x[row * ld + col], row);
    if ((key & hi_mask) != (st[col].prefix & hi_mask)) continue;
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    if (LDS) atomicAdd(&sh[col * 256 + d], 1u);
    else atomicAdd(&hist[col * 256 + d], 1u);
  }
  if (LDS) {
    __syncthreads();
    for (int64_t i = threadIdx.x; i < K * 256; i += blockDim.x)
      if (sh[i]) atomicAdd(&hist[i], sh[i]);
  }
}

// Fix digit `shift` of every column's threshold and clear the histogram for the next pass.
__global__ void k_sel_pick(int64_t K, int shift, SelState* __restrict__ st,
                           uint32_t* __restrict__ hist) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < K;
       c += (int64_t)gridDim.x * blockDim.x) {
    SelState s = st[c];
    int64_t above = 0;
    int d = 255;
    for (; d > 0; --d) {
      const int64_t h = hist[c * 256 + d];
      if (above + h >= s.need) break;
      above += h;
    }
    s.prefix |= (uint64_t)d << shift;
    s.need -= above;
    st[c] = s;
    for (int i = 0; i < 256; ++i) hist[c * 256 + i] = 0;
  }
}

// Keys at or above the threshold (exactly k per column) with their rows.
__global__ void k_sel_collect(const float* __restrict__ x, int64_t N, int64_t K, int64_t ld,
                              int64_t k, const SelState* __restrict__ st,
                              uint32_t* __restrict__ cnt, uint64_t* __restrict__ keys,
                              int64_t* __restrict__ rows) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * K;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e / K, col = e % K;
    const uint64_t key = sel_key(x[row * ld + col], row);
    if (key < st[col].prefix) continue;
    const uint32_t slot = atomicAdd(&cnt[col], 1u);
    if (slot < (uint32_t)k) {
      keys[col * k + slot] = key;
      rows[col * k + slot] = row;
    }
  }
}

__global__ void k_sel_init(int64_t K, int64_t k, SelState* __restrict__ st,
                           uint32_t* __restrict__ cnt, int32_t* __restrict__ seg) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c <= K;
       c += (int64_t)gridDim.x * blockDim.x) {
    seg[c] = (int32_t)(c * k);
    if (c < K) {
      st[c] = SelState{0ull, k};
      cnt[c] = 0;
    }
  }
}

// Sorted keys -> values [k][K] and rows [k][K] (the layout of torch.topk over dim 0).
__global__ void k_sel_out(int64_t K, int64_t k, const uint64_t* __restrict__ keys,
                          const int64_t* __restrict__ rows, float* __restrict__ vals_out,
                          int64_t* __restrict__ rows_out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < K * k;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = i / k, j = i % k;
    uint32_t u = (uint32_t)(keys[i] >> 32);
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    vals_out[j * K + c] = __uint_as_float(u);
    rows_out[j * K + c] = 0xFFFFFFFFll - (int64_t)(uint32_t)(keys[i] & 0xFFFFFFFFull);
  }
}

struct SelWs {
  SelState* st;
  uint32_t* hist;
  uint32_t* cnt;
  int32_t* seg;
  uint64_t* keys[2];
  int64_t* rows[2];
  void* tmp;
  size_t tmp_bytes;
  int64_t bytes;
};

static int sel_carve(char* base, int64_t K, int64_t k, SelWs& w) {
  size_t sort_b = 0;
  const int64_t n = std::max<int64_t>(K * k, 1);
  GL_HIP(hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
      nullptr, sort_b, (const uint64_t*)nullptr, (uint64_t*)nullptr, (const int64_t*)nullptr,
      (int64_t*)nullptr, (int)n, (int)std::max<int64_t>(K, 1), (const int32_t*)nullptr,
      (const int32_t*)nullptr, 0, 64));
  int64_t off = 0;
  auto take = [&](int64_t b) {
    char* p = base ? base + off : nullptr;
    off += (b + 255) & ~(int64_t)255;
    return p;
  };
  w.st = (SelState*)take(K * sizeof(SelState));
  w.hist = (uint32_t*)take(K * 256 * 4);
  w.cnt = (uint32_t*)take(K * 4);
  w.seg = (int32_t*)take((K + 1) * 4);
  w.keys[0] = (uint64_t*)take(n * 8);
  w.keys[1] = (uint64_t*)take(n * 8);
  w.rows[0] = (int64_t*)take(n * 8);
  w.rows[1] = (int64_t*)take(n * 8);
  w.tmp_bytes = sort_b;
  w.tmp = take((int64_t)sort_b);
  w.bytes = off;
  return GL_OK;
}

}  // namespace gl

using namespace gl;

extern "C" {

int64_t gl_topk_workspace_bytes(int64_t K, int64_t k) {
  if (K < 1 || k < 1 || K * k >= INT32_MAX) return -1;
  SelWs w;
  if (sel_carve(nullptr, K, k, w) != GL_OK) return -1;
  return w.bytes;
}

int gl_topk_columns_f32(const float* x, int64_t N, int64_t K, int64_t ld, int64_t k,
                        float* values, int64_t* rows, void* ws, int64_t ws_bytes, void* stream) {
  GL_REQUIRE(x && values && rows, "null pointer");
  GL_REQUIRE(N >= 1 && N < INT32_MAX && K >= 1 && ld >= K && k >= 1 && k <= N,
             "bad sizes (N=%lld K=%lld ld=%lld k=%lld)", (long long)N, (long long)K,
             (long long)ld, (long long)k);
  GL_REQUIRE(K * k < INT32_MAX, "K * k too large");
  SelWs w;
  int rc = sel_carve((char*)ws, K, k, w);
  if (rc) return rc;
  GL_REQUIRE(ws != nullptr && ws_bytes >= w.bytes, "workspace too small: %lld < %lld bytes",
             (long long)ws_bytes, (long long)w.bytes);
  hipStream_t st = (hipStream_t)stream;
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((N * K + 255) / 256, 8192));
  GL_HIP(hipMemsetAsync(w.hist, 0, K * 256 * 4, st));
  hipLaunchKernelGGL(k_sel_init, dim3((unsigned)((K + 256) / 256)), dim3(256), 0, st, K, k,
                     w.st, w.cnt, w.seg);
  GL_LAUNCH_CHECK("k_sel_init");
  const bool lds = K <= 64;
  for (int shift = 56; shift >= 0; shift -= 8) {
    if (lds)
      hipLaunchKernelGGL(k_sel_hist<true>, dim3(nb), dim3(256), (size_t)(K * 256 * 4), st, x, N,
                         K, ld, shift, w.st, w.hist);
    else
      hipLaunchKernelGGL(k_sel_hist<false>, dim3(nb), dim3(256), 0, st, x, N, K, ld, shift,
                         w.st, w.hist);
    GL_LAUNCH_CHECK("k_sel_hist");
    hipLaunchKernelGGL(k_sel_pick, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, st, K, shift,
                       w.st, w.hist);
    GL_LAUNCH_CHECK("k_sel_pick");
  }
  hipLaunchKernelGGL(k_sel_collect, dim3(nb), dim3(256), 0, st, x, N, K, ld, k, w.st, w.cnt,
                     w.keys[0], w.rows[0]);
  GL_LAUNCH_CHECK("k_sel_collect");
  size_t tb = w.tmp_bytes;
  GL_HIP(hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
      w.tmp, tb, w.keys[0], w.keys[1], w.rows[0], w.rows[1], (int)(K * k), (int)K, w.seg,
      w.seg + 1, 0, 64, st));
  hipLaunchKernelGGL(k_sel_out, dim3((unsigned)std::min<int64_t>((K * k + 255) / 256, 65535)),
                     dim3(256), 0, st, K, k, w.keys[1], w.rows[1], values, rows);
  GL_LAUNCH_CHECK("k_sel_out");
  return GL_OK;
}

}  // extern "C"
