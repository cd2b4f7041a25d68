// FP64 SpMM of a float64 PBR matrix. It keeps float64 matrices working end to end, as the
// reference's templated kernels do (pbr_spmm_cuda_i32_f64); the FP32 kernels remain the fast
// path. The PPR init / update steps in both precisions are in ppr_step.hip.
//
// SpMM: one wave per stripe (XCD-aware persistent grid). Lane (r, q) owns row r of the
// stripe and RHS columns q, q + L, ... (L = 64 / B lanes per row); it walks the stripe's
// tiles, the set bits of its row in bit order (values are packed in that order), then its
// row's remainder entries, accumulating in double. Every output element is written once.
#include <algorithm>

#include "gl_common.hpp"

namespace gl {

int grid_for(int64_t work_items, int wg_per_cu);

struct PbrView64 {
  int64_t rows, cols, num_stripes;
  const uint64_t* __restrict__ codes;
  const int32_t* __restrict__ coords;
  const int32_t* __restrict__ offsets;
  const double* __restrict__ data;
  const int32_t* __restrict__ stripe_ptr;
  const int32_t* __restrict__ rem_ptr;
  const int32_t* __restrict__ rem_col;
  const double* __restrict__ rem_val;
};

template <int B>
__global__ void __launch_bounds__(256) k_spmm_f64(PbrView64 m, const double* __restrict__ X,
                                                  int64_t ldx, int64_t K,
                                                  double* __restrict__ Y, int64_t ldy,
                                                  int accumulate) {
  constexpr int L = kWave / B;
  constexpr uint32_t kRowMask = (1u << B) - 1u;
  const int lane = threadIdx.x & (kWave - 1);
  const int r = lane / L;
  const int q = lane % L;
  const Sched sc = xcd_sched(m.num_stripes);
  for (int64_t s = sc.first; s < sc.end; s += sc.stride) {
    const int64_t row = s * B + r;
    const bool rok = row < m.rows;
    const int t0 = m.stripe_ptr[s], t1 = m.stripe_ptr[s + 1];
    const int j0 = rok ? m.rem_ptr[row] : 0;
    const int j1 = rok ? m.rem_ptr[row + 1] : 0;
    for (int64_t f0 = 0; f0 < K; f0 += L) {
      const int64_t f = f0 + q;
      const bool fok = f < K;
      double acc = 0.0;
      for (int t = t0; t < t1; ++t) {
        const uint64_t code = m.codes[t];
        const int64_t c0 = m.coords[2 * (int64_t)t + 1];
        const int rb = r * B;
        const uint32_t bits = (uint32_t)(code >> rb) & kRowMask;
        int idx = m.offsets[t] + (rb == 0 ? 0 : __popcll(code & ((1ull << rb) - 1ull)));
#pragma unroll
        for (int c = 0; c < B; ++c) {
          if ((bits >> c) & 1u) {
            const double v = m.data[idx++];
            const int64_t col = c0 + c;
            if (fok && col < m.cols) acc = __builtin_fma(v, X[col * ldx + f], acc);
          }
        }
      }
      for (int j = j0; j < j1; ++j)
        if (fok) acc = __builtin_fma(m.rem_val[j], X[(int64_t)m.rem_col[j] * ldx + f], acc);
      if (rok && fok) {
        double* p = Y + row * ldy + f;
        *p = accumulate ? *p + acc : acc;
      }
    }
  }
}

}  // namespace gl

using namespace gl;

extern "C" {

int gl_pbr_spmm_i32_f64(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                        int64_t num_tiles, const uint64_t* codes, const int32_t* coords,
                        const int32_t* offsets, const double* data, const int32_t* stripe_ptr,
                        const int32_t* rem_indptr, const int32_t* rem_col_ind,
                        const double* rem_vals, const double* X, int64_t ldx, int64_t K,
                        double* Y, int64_t ldy, int32_t accumulate, void* stream) {
  GL_REQUIRE(rows >= 0 && cols >= 0 && num_tiles >= 0, "negative size");
  GL_REQUIRE(block_rows == block_cols &&
                 (block_rows == 2 || block_rows == 4 || block_rows == 8),
             "Unsupported block size %dx%d: only 2x2, 4x4 and 8x8 are supported", block_rows,
             block_cols);
  GL_REQUIRE(rows <= INT32_MAX && cols <= INT32_MAX, "rows/cols exceed int32 indices");
  GL_REQUIRE(num_tiles == 0 || (codes && coords && offsets && data), "null tile array");
  GL_REQUIRE((stripe_ptr && rem_indptr) || rows == 0, "null stripe_ptr or remainder indptr");
  GL_REQUIRE(K >= 1, "K must be >= 1");
  GL_REQUIRE(ldx >= K && ldy >= K, "leading dimension smaller than K");
  GL_REQUIRE((X != nullptr || cols == 0) && (Y != nullptr || rows == 0), "null X or Y");
  if (rows == 0) return GL_OK;
  PbrView64 m{rows, cols, (rows + block_rows - 1) / block_rows, codes, coords, offsets, data,
              stripe_ptr, rem_indptr, rem_col_ind, rem_vals};
  const int nwg = grid_for(m.num_stripes, 0);
  hipStream_t st = (hipStream_t)stream;
  switch (block_rows) {
    case 2:
      hipLaunchKernelGGL(k_spmm_f64<2>, dim3(nwg), dim3(256), 0, st, m, X, ldx, K, Y, ldy,
                         accumulate);
      break;
    case 4:
      hipLaunchKernelGGL(k_spmm_f64<4>, dim3(nwg), dim3(256), 0, st, m, X, ldx, K, Y, ldy,
                         accumulate);
      break;
    default:
      hipLaunchKernelGGL(k_spmm_f64<8>, dim3(nwg), dim3(256), 0, st, m, X, ldx, K, Y, ldy,
                         accumulate);
      break;
  }
  GL_LAUNCH_CHECK("k_spmm_f64");
  return GL_OK;
}

}  // extern "C"
