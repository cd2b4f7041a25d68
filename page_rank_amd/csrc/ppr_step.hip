// The reference's per-iteration PPR kernels as separate calls, in FP32 and FP64: one-hot
// init, the in-place update for a column-stochastic matrix, and the older dangling-to-seed
// pair (column sums of Y, then the update that sends the missing mass to the seed)
// (Backend/graph_link/kernels/ppr_kernel.cu:13-216, bound at
// Backend/graph_link/bindings/bindings_kernels.cu:148-161). A host loop built from them
// runs one SpMM plus one update per iteration; the fused, graph-captured loop of
// gl_ppr_run_i32_f32 is the fast path.
//
// Where the reference adds per-block float atomics into `errors` / `col_sums`, these kernels
// write per-block partial sums (a fixed row range per block, threads combined in thread
// order) and a second kernel adds the partials in block order, so results do not depend on
// scheduling.
#include <algorithm>
#include <mutex>

#include "gl_common.hpp"

namespace gl {

// Scratch of gl_ppr_update_normalized_f32_noalloc: the per-block partial sums that the
// workspace holds for gl_ppr_update_normalized_f32. A file-scope device array, so the call
// allocates nothing; every GPU the library runs on holds its own copy (kScratchBytes) until
// the process ends. In host code the array's name is not a device address: its address on
// each device comes from hipGetSymbolAddress, taken once per device.
constexpr int64_t kScratchBytes = 4 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (nullptr and the last error set on
// failure).
static void* scratch_ptr() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    hip_fail(e, "hipGetDevice", __FILE__, __LINE__);
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d is beyond the %d devices the scratch table covers", dev, kMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)",
               __FILE__, __LINE__);
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// X = one-hot columns: X[i][k] = (i == src[k]). The reference's init only sets the ones
// into a zeroed X (ppr_kernel.cu:199-216); this writes every element, so X need not be
// zeroed first.
template <typename T>
__global__ void k_ppr_init(T* __restrict__ X, const int32_t* __restrict__ src, int64_t N,
                           int64_t K) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * K;
       e += (int64_t)gridDim.x * blockDim.x)
    X[e] = (e / K == (int64_t)src[e % K]) ? T(1) : T(0);
}

// Block b owns rows [b * rpb, (b + 1) * rpb). A wave covers W = min(K, 64) columns and
// RP = 64 / W rows per step; the four waves of the block interleave rows. `part` gets the
// block's sum per column (threads combined in a fixed order).
//   MODE 0 (normalized, in place, Xn == Xc): Xn = alpha Y + (1 - alpha) [i == src]; part = |Xn - Xc|
//   MODE 1 (dangling to seed): Xn = alpha Y + [i == src] (alpha (1 - colsum) + (1 - alpha));
//          part = |Xn - Xc|
//   MODE 2 (column sums of Y): part = Y
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_ppr_cols(const T* __restrict__ Y, const T* Xc, T* Xn,
                                                  const int32_t* __restrict__ src,
                                                  const T* __restrict__ colsum, T alpha,
                                                  int64_t N, int64_t K, int64_t rpb,
                                                  T* __restrict__ part) {
  __shared__ T red[256];
  const int W = (int)(K < kWave ? K : kWave);
  const int RP = kWave / W;
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int lc = lane % W, lr = lane / W;
  const bool lane_ok = lr < RP;
  const int64_t r0 = blockIdx.x * rpb;
  const int64_t r1 = r0 + rpb < N ? r0 + rpb : N;
  const T comp = T(1) - alpha;
  for (int64_t k0 = 0; k0 < K; k0 += W) {
    const int64_t k = k0 + lc;
    const bool ok = lane_ok && k < K;
    T e = T(0);
    if (ok) {
      const int64_t sk = MODE == 2 ? -1 : (int64_t)src[k];
      const T tm = MODE == 1 ? alpha * (T(1) - colsum[k]) + comp : comp;
      for (int64_t i = r0 + wid * RP + lr; i < r1; i += 4 * RP) {
        const int64_t ix = i * K + k;
        if constexpr (MODE == 2) {
          e += Y[ix];
        } else {
          T nv = alpha * Y[ix];
          if (i == sk) nv += tm;
          const T d = nv - Xc[ix];
          e += d < T(0) ? -d : d;
          Xn[ix] = nv;
        }
      }
    }
    red[threadIdx.x] = e;
    __syncthreads();
    if (threadIdx.x < W && k0 + threadIdx.x < K) {
      T sum = T(0);
      for (int w = 0; w < 4; ++w)
        for (int rr = 0; rr < RP; ++rr) sum += red[w * kWave + rr * W + threadIdx.x];
      part[blockIdx.x * K + k0 + threadIdx.x] = sum;
    }
    __syncthreads();
  }
}

// out[k] (+)= sum over blocks of part[b][k], in block order.
template <typename T>
__global__ void k_ppr_colsum_final(const T* __restrict__ part, int64_t nb, int64_t K,
                                   T* __restrict__ out, int accumulate) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < K;
       k += (int64_t)gridDim.x * blockDim.x) {
    T s = T(0);
    for (int64_t b = 0; b < nb; ++b) s += part[b * K + k];
    out[k] = accumulate ? out[k] + s : s;
  }
}

constexpr int64_t kUpdBlocks = 1024;

// Blocks and rows per block of k_ppr_cols for N rows and K columns.
void update_shape(int64_t N, int64_t K, int64_t& nb, int64_t& rpb) {
  const int64_t W = K < kWave ? K : kWave;
  const int64_t step = 4 * (kWave / W);  // rows one block advances per loop step
  rpb = (N + kUpdBlocks - 1) / kUpdBlocks;
  rpb = ((rpb + step - 1) / step) * step;
  if (rpb < step) rpb = step;
  nb = (N + rpb - 1) / rpb;
  if (nb < 1) nb = 1;
}

template <typename T>
int64_t cols_workspace_bytes(int64_t N, int64_t K) {
  if (N < 1 || K < 1) return -1;
  int64_t nb, rpb;
  update_shape(N, K, nb, rpb);
  return nb * K * (int64_t)sizeof(T);
}

template <typename T>
int ppr_init(T* X, const int32_t* src, int64_t N, int64_t K, void* stream) {
  GL_REQUIRE(N >= 1 && K >= 1, "bad sizes (N=%lld K=%lld)", (long long)N, (long long)K);
  GL_REQUIRE(X && src, "null pointer");
  const int64_t nb = std::min<int64_t>((N * K + 255) / 256, 65536);
  hipLaunchKernelGGL(k_ppr_init<T>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, X,
                     src, N, K);
  GL_LAUNCH_CHECK("k_ppr_init");
  return GL_OK;
}

// One k_ppr_cols pass of MODE over [N][K] plus the block-order sum into out[K]. `part` holds
// nb * K partials (nb from update_shape).
template <typename T, int MODE>
int ppr_cols(const T* Y, const T* Xc, T* Xn, const int32_t* src, const T* colsum, T alpha,
             int64_t N, int64_t K, T* out, int accumulate, T* part, hipStream_t st) {
  int64_t nb, rpb;
  update_shape(N, K, nb, rpb);
  hipLaunchKernelGGL((k_ppr_cols<T, MODE>), dim3((unsigned)nb), dim3(256), 0, st, Y, Xc, Xn,
                     src, colsum, alpha, N, K, rpb, part);
  GL_LAUNCH_CHECK("k_ppr_cols");
  hipLaunchKernelGGL(k_ppr_colsum_final<T>, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, st,
                     (const T*)part, nb, K, out, accumulate);
  GL_LAUNCH_CHECK("k_ppr_colsum_final");
  return GL_OK;
}

template <typename T>
int check_ws(int64_t N, int64_t K, const void* ws, int64_t ws_bytes) {
  const int64_t need = cols_workspace_bytes<T>(N, K);
  if (ws == nullptr || ws_bytes < need) {
    set_error("workspace too small: %lld < %lld bytes", (long long)ws_bytes, (long long)need);
    return GL_ERR_WORKSPACE;
  }
  return GL_OK;
}

template <typename T>
int ppr_update_normalized(const T* Y, T* X, const int32_t* src, double alpha, int64_t N,
                          int64_t K, T* errors, void* ws, int64_t ws_bytes, void* stream) {
  GL_REQUIRE(N >= 1 && K >= 1, "bad sizes (N=%lld K=%lld)", (long long)N, (long long)K);
  GL_REQUIRE(Y && X && src && errors, "null pointer");
  int rc = check_ws<T>(N, K, ws, ws_bytes);
  if (rc) return rc;
  return ppr_cols<T, 0>(Y, X, X, src, nullptr, (T)alpha, N, K, errors, 0, (T*)ws,
                        (hipStream_t)stream);
}

template <typename T>
int ppr_missing_mass(const T* Y, T* col_sums, int64_t N, int64_t K, void* ws,
                     int64_t ws_bytes, void* stream) {
  GL_REQUIRE(N >= 1 && K >= 1, "bad sizes (N=%lld K=%lld)", (long long)N, (long long)K);
  GL_REQUIRE(Y && col_sums, "null pointer");
  int rc = check_ws<T>(N, K, ws, ws_bytes);
  if (rc) return rc;
  return ppr_cols<T, 2>(Y, nullptr, nullptr, nullptr, nullptr, T(0), N, K, col_sums, 0, (T*)ws,
                        (hipStream_t)stream);
}

template <typename T>
int ppr_update(const T* Y, const T* X_cur, T* X_next, const int32_t* src, const T* col_sums,
               double alpha, int64_t N, int64_t K, T* errors, void* ws, int64_t ws_bytes,
               void* stream) {
  GL_REQUIRE(N >= 1 && K >= 1, "bad sizes (N=%lld K=%lld)", (long long)N, (long long)K);
  GL_REQUIRE(Y && X_cur && X_next && src && col_sums && errors, "null pointer");
  GL_REQUIRE(X_cur != X_next, "X_cur and X_next must be different buffers");
  int rc = check_ws<T>(N, K, ws, ws_bytes);
  if (rc) return rc;
  return ppr_cols<T, 1>(Y, X_cur, X_next, src, col_sums, (T)alpha, N, K, errors, 1, (T*)ws,
                        (hipStream_t)stream);
}

}  // namespace gl

using namespace gl;

extern "C" {

int gl_ppr_init_f32(float* X, const int32_t* source_nodes, int64_t N, int64_t K, void* stream) {
  return ppr_init<float>(X, source_nodes, N, K, stream);
}

int gl_ppr_init_f64(double* X, const int32_t* source_nodes, int64_t N, int64_t K,
                    void* stream) {
  return ppr_init<double>(X, source_nodes, N, K, stream);
}

int64_t gl_ppr_update_f32_workspace_bytes(int64_t N, int64_t K) {
  return cols_workspace_bytes<float>(N, K);
}

int64_t gl_ppr_update_f64_workspace_bytes(int64_t N, int64_t K) {
  return cols_workspace_bytes<double>(N, K);
}

int gl_ppr_update_normalized_f32(const float* Y, float* X, const int32_t* source_nodes,
                                 float alpha, int64_t N, int64_t K, float* errors, void* ws,
                                 int64_t ws_bytes, void* stream) {
  return ppr_update_normalized<float>(Y, X, source_nodes, alpha, N, K, errors, ws, ws_bytes,
                                      stream);
}

int gl_ppr_update_normalized_f32_noalloc(const float* Y, float* X, const int32_t* source_nodes,
                                         float alpha, int64_t N, int64_t K, float* errors,
                                         void* stream) {
  GL_REQUIRE(N >= 1 && K >= 1, "bad sizes (N=%lld K=%lld)", (long long)N, (long long)K);
  GL_REQUIRE(Y && X && source_nodes && errors, "null pointer");
  const int64_t need = cols_workspace_bytes<float>(N, K);
  GL_REQUIRE(need <= kScratchBytes,
             "N=%lld, K=%lld needs %lld bytes of partial sums, above the %lld-byte static "
             "scratch of this entry point (any N works for K <= 1024); use "
             "gl_ppr_update_normalized_f32 with a workspace",
             (long long)N, (long long)K, (long long)need, (long long)kScratchBytes);
  void* part = scratch_ptr();
  if (!part) return GL_ERR_HIP;
  // every partial that the sum reads is written by this call's first kernel: nothing of an
  // earlier call is read, so the scratch needs no clearing
  return ppr_cols<float, 0>(Y, X, X, source_nodes, nullptr, alpha, N, K, errors, 0,
                            (float*)part, (hipStream_t)stream);
}

int gl_ppr_update_normalized_f64(const double* Y, double* X, const int32_t* source_nodes,
                                 double alpha, int64_t N, int64_t K, double* errors, void* ws,
                                 int64_t ws_bytes, void* stream) {
  return ppr_update_normalized<double>(Y, X, source_nodes, alpha, N, K, errors, ws, ws_bytes,
                                       stream);
}

int gl_ppr_missing_mass_f32(const float* Y, float* col_sums, int64_t N, int64_t K, void* ws,
                            int64_t ws_bytes, void* stream) {
  return ppr_missing_mass<float>(Y, col_sums, N, K, ws, ws_bytes, stream);
}

int gl_ppr_missing_mass_f64(const double* Y, double* col_sums, int64_t N, int64_t K, void* ws,
                            int64_t ws_bytes, void* stream) {
  return ppr_missing_mass<double>(Y, col_sums, N, K, ws, ws_bytes, stream);
}

int gl_ppr_update_f32(const float* Y, const float* X_cur, float* X_next,
                      const int32_t* source_nodes, const float* col_sums, float alpha,
                      int64_t N, int64_t K, float* errors, void* ws, int64_t ws_bytes,
                      void* stream) {
  return ppr_update<float>(Y, X_cur, X_next, source_nodes, col_sums, alpha, N, K, errors, ws,
                           ws_bytes, stream);
}

int gl_ppr_update_f64(const double* Y, const double* X_cur, double* X_next,
                      const int32_t* source_nodes, const double* col_sums, double alpha,
                      int64_t N, int64_t K, double* errors, void* ws, int64_t ws_bytes,
                      void* stream) {
  return ppr_update<double>(Y, X_cur, X_next, source_nodes, col_sums, alpha, N, K, errors, ws,
                            ws_bytes, stream);
}

}  // extern "C"
