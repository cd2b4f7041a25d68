// C ABI: error reporting, device queries, the stripe index and the SpMM entry points.
#include <cstdarg>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "gl_common.hpp"
#include "launch.hpp"
#include "stripe_kernels.hpp"
#include "stripe_mfma.hpp"

namespace gl {

static thread_local std::string t_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_last_error = buf;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  return GL_ERR_HIP;
}

namespace {
struct PerDevice {
  bool init = false;
  DeviceInfo info;
  std::vector<StreamSet*> free_sets;  // stream sets not leased by any call
};
std::mutex g_mu;
std::deque<PerDevice> g_dev;  // a deque: growing it keeps references to its elements

PerDevice& per_device_locked(int dev) {
  if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
  PerDevice& p = g_dev[dev];
  if (!p.init) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
      p.info.num_cus = prop.multiProcessorCount;
      p.info.lds_bytes = (int)prop.maxSharedMemoryPerMultiProcessor;
      p.info.clock_khz = prop.clockRate;
      p.info.hbm_bytes = (int64_t)prop.totalGlobalMem;
    }
    p.init = true;
  }
  return p;
}
}  // namespace

const DeviceInfo& device_info() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_mu);
  return per_device_locked(dev).info;
}

// Stream sets are leased for the length of one call and returned to a per-device free
// list, so two calls that run at once (threads of a server) never share a stream or an
// event. A returned set may still hold queued work of the call that used it; the next
// call's work then queues behind it, and an event wait issued earlier is not affected by a
// later record of the same event, so reuse is safe. Sets live until the process ends.
StreamSet* acquire_stream_set() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    PerDevice& p = per_device_locked(dev);
    if (!p.free_sets.empty()) {
      StreamSet* s = p.free_sets.back();
      p.free_sets.pop_back();
      return s;
    }
  }
  StreamSet* s = new StreamSet;
  s->device = dev;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming);
  if (e != hipSuccess) {
    set_error("HIP error %d (%s) creating a stream set", (int)e, hipGetErrorString(e));
    delete s;  // its stream or events, if created, stay (this path only runs on failure)
    return nullptr;
  }
  return s;
}

void release_stream_set(StreamSet* s) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(g_mu);
  per_device_locked(s->device).free_sets.push_back(s);
}

// SpMM grids: up to 32 workgroups per CU, more than the device holds at once, so the
// hardware hands out the later ones as earlier ones finish and every CU stays busy to the
// end. Measured on MI355X (1M x 1M, 8x8) against 8 per CU: 1 column 114 -> 106 us,
// 2 columns 161 -> 155 us, 64 columns 479 -> 467 us, 8 and 32 columns unchanged.
constexpr int kSpmmWgPerCu = 32;

int grid_for(int64_t work_items, int wg_per_cu) {
  const int cus = device_info().num_cus;
  int64_t want = (work_items + 3) / 4;  // 4 waves per workgroup, one item per wave
  int64_t cap = (int64_t)cus * (wg_per_cu > 0 ? wg_per_cu : 8);
  int64_t n = want < cap ? want : cap;
  n = ((n + kNumXcd - 1) / kNumXcd) * kNumXcd;
  if (n < kNumXcd) n = kNumXcd;
  return (int)n;
}

// Columns per launch for K RHS columns. Two columns run in the 4-column kernel, whose
// float4 X-row loads beat the 2-column one (measured 124 vs 154 us on 1M x 1M, 8x8).
int choose_kw(int64_t K) {
  if (K <= 1) return 1;
  if (K <= 4) return 4;
  if (K <= 8) return 8;
  if (K <= 16) return 16;
  if (K <= 32) return 32;
  return 64;
}

bool valid_kw(int kw) {
  return kw == 1 || kw == 2 || kw == 4 || kw == 8 || kw == 16 || kw == 32 || kw == 64;
}

// stripe_ptr[s] = first tile t with coords[2t] / B >= s (tiles are stripe-ascending).
__global__ void k_stripe_index(int64_t num_stripes, int32_t B, int64_t T,
                               const int32_t* __restrict__ coords,
                               int32_t* __restrict__ stripe_ptr) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s <= num_stripes;
       s += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = T;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)(coords[2 * mid] / B) < s) lo = mid + 1;
      else hi = mid;
    }
    stripe_ptr[s] = (int32_t)lo;
  }
}

// Execution layout of an 8x8 PBR matrix: one wave per tile, lane l = column-major
// position 8c + r. The transposed code is the wave's ballot of "bit (8r + c) is set"; each
// value moves from its row-major rank to its column-major rank inside the tile's range.
__global__ void k_exec_layout(int64_t T, const uint64_t* __restrict__ codes,
                              const int32_t* __restrict__ coords,
                              const int32_t* __restrict__ offsets, const float* __restrict__ data,
                              TileRec* __restrict__ trec, float* __restrict__ tdata) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  for (int64_t t = w0; t < T + kTileRecPad; t += waves) {
    if (t >= T) {
      if (lane == 0) trec[t] = TileRec{0u, 0u, 0, T > 0 ? offsets[T] : 0};
      continue;
    }
    const uint64_t code = codes[t];
    const int r = lane & 7, c = lane >> 3;
    const int pos = 8 * r + c;
    const bool has = (code >> pos) & 1ull;
    const uint64_t tcode = __ballot(has);
    const int off = offsets[t];
    if (has) {
      const int rm = __popcll(code & ((1ull << pos) - 1ull));
      const int cm = __popcll(tcode & ((1ull << lane) - 1ull));
      tdata[off + cm] = data[off + rm];
    }
    if (lane == 0) trec[t] = TileRec{(uint32_t)tcode, (uint32_t)(tcode >> 32), coords[2 * t + 1], off};
  }
}

}  // namespace gl

using namespace gl;

extern "C" {

const char* gl_last_error(void) { return t_last_error.c_str(); }

int32_t gl_abi_version(void) { return 10; }

int gl_device_info(int32_t* num_cus, int32_t* lds_bytes_per_cu, int32_t* clock_khz,
                   int64_t* hbm_bytes) {
  const DeviceInfo& d = device_info();
  if (num_cus) *num_cus = d.num_cus;
  if (lds_bytes_per_cu) *lds_bytes_per_cu = d.lds_bytes;
  if (clock_khz) *clock_khz = d.clock_khz;
  if (hbm_bytes) *hbm_bytes = d.hbm_bytes;
  return GL_OK;
}

int gl_pbr_stripe_index(int64_t rows, int32_t block_rows, int64_t num_tiles,
                        const int32_t* coords, int32_t* stripe_ptr, void* stream) {
  GL_REQUIRE(rows >= 0 && num_tiles >= 0, "negative size");
  GL_REQUIRE(block_rows > 0, "block_rows must be positive");
  GL_REQUIRE(stripe_ptr != nullptr, "stripe_ptr is null");
  GL_REQUIRE(num_tiles == 0 || coords != nullptr, "coords is null");
  const int64_t ns = (rows + block_rows - 1) / block_rows;
  const int64_t n = ns + 1;
  const int threads = 256;
  const int64_t blocks = (n + threads - 1) / threads < 65536 ? (n + threads - 1) / threads : 65536;
  hipLaunchKernelGGL(k_stripe_index, dim3((unsigned)blocks), dim3(threads), 0,
                     (hipStream_t)stream, ns, block_rows, num_tiles, coords, stripe_ptr);
  GL_LAUNCH_CHECK("k_stripe_index");
  return GL_OK;
}

static int check_pbr(int64_t rows, int64_t cols, int32_t br, int32_t bc, int64_t T,
                     const uint64_t* codes, const int32_t* coords, const int32_t* offsets,
                     const float* data, const int32_t* stripe_ptr) {
  GL_REQUIRE(rows >= 0 && cols >= 0 && T >= 0, "negative size");
  GL_REQUIRE(br == bc && (br == 2 || br == 4 || br == 8),
             "Unsupported block size %dx%d: only 2x2, 4x4 and 8x8 are supported", br, bc);
  GL_REQUIRE(rows <= INT32_MAX && cols <= INT32_MAX, "rows/cols exceed int32 indices");
  GL_REQUIRE(T == 0 || (codes && coords && offsets && data), "null tile array");
  GL_REQUIRE(stripe_ptr != nullptr || rows == 0, "null stripe_ptr");
  return GL_OK;
}

static int check_dense(int64_t rows, int64_t cols, const float* X, int64_t ldx, int64_t K,
                       float* Y, int64_t ldy) {
  GL_REQUIRE(K >= 1, "K must be >= 1");
  GL_REQUIRE(ldx >= K && ldy >= K, "leading dimension smaller than K");
  GL_REQUIRE((X != nullptr || cols == 0) && (Y != nullptr || rows == 0), "null X or Y");
  return GL_OK;
}

int gl_pbr_block_spmm_i32_f32(int64_t rows, int64_t cols, int32_t block_rows,
                              int32_t block_cols, int64_t num_tiles, int64_t num_vals,
                              const uint64_t* codes,
                              const int32_t* coords, const int32_t* offsets, const float* data,
                              const int32_t* stripe_ptr, const float* X, int64_t ldx,
                              int64_t K, float* Y, int64_t ldy, int32_t accumulate, int32_t kw,
                              void* stream) {
  int rc = check_pbr(rows, cols, block_rows, block_cols, num_tiles, codes, coords, offsets,
                     data, stripe_ptr);
  if (rc) return rc;
  if ((rc = check_dense(rows, cols, X, ldx, K, Y, ldy))) return rc;
  if (kw == 0) kw = choose_kw(K);
  GL_REQUIRE(valid_kw(kw), "kw must be one of 1,2,4,8,16,32,64");
  if (rows == 0) return GL_OK;
  PbrView m{rows, cols, (rows + block_rows - 1) / block_rows, num_tiles, num_vals, 0,
            codes, coords, offsets, data, stripe_ptr, nullptr, nullptr, nullptr};
  const int nwg = grid_for(m.num_stripes, kSpmmWgPerCu);
  for (int64_t k0 = 0, w = kw; k0 < K; k0 += w) {
    w = chunk_kw(kw, K - k0);
    RhsView d{X, ldx, K, k0, cols * ldx};
    EpiStore e{Y, ldy, accumulate, fits_u24(rows, ldy)};
    if ((rc = launch_stripe<EpiStore, true, false>(m, d, block_rows, (int)w, e, nwg,
                                                   (hipStream_t)stream)))
      return rc;
  }
  return GL_OK;
}

int gl_csr_spmm_i32_f32(int64_t rows, int64_t cols, int64_t nnz, const int32_t* indptr,
                        const int32_t* col_ind, const float* vals, const float* X, int64_t ldx,
                        int64_t K, float* Y, int64_t ldy, int32_t accumulate, int32_t kw,
                        void* stream) {
  GL_REQUIRE(rows >= 0 && cols >= 0, "negative size");
  GL_REQUIRE(indptr != nullptr || rows == 0, "null indptr");
  int rc = check_dense(rows, cols, X, ldx, K, Y, ldy);
  if (rc) return rc;
  if (kw == 0) kw = choose_kw(K);
  GL_REQUIRE(valid_kw(kw), "kw must be one of 1,2,4,8,16,32,64");
  if (rows == 0) return GL_OK;
  // Row groups of 8 rows; one wave owns a group (the kernel's "stripe").
  PbrView m{rows, cols, (rows + 7) / 8, 0, 0, nnz, nullptr, nullptr, nullptr, nullptr, nullptr,
            indptr, col_ind, vals};
  const int nwg = grid_for(m.num_stripes, kSpmmWgPerCu);
  for (int64_t k0 = 0, w = kw; k0 < K; k0 += w) {
    w = chunk_kw(kw, K - k0);
    RhsView d{X, ldx, K, k0, cols * ldx};
    EpiStore e{Y, ldy, accumulate, fits_u24(rows, ldy)};
    if ((rc = launch_stripe<EpiStore, false, true>(m, d, 8, (int)w, e, nwg,
                                                   (hipStream_t)stream)))
      return rc;
  }
  return GL_OK;
}

int64_t gl_pbr_exec_layout_bytes(int64_t num_tiles) {
  return num_tiles < 0 ? 0 : (num_tiles + kTileRecPad) * (int64_t)sizeof(TileRec);
}

int gl_pbr_exec_layout(int64_t num_tiles, int64_t num_vals, const uint64_t* codes,
                       const int32_t* coords, const int32_t* offsets, const float* data,
                       void* tile_records, float* values_cm, void* stream) {
  GL_REQUIRE(num_tiles >= 0 && num_vals >= 0, "negative size");
  GL_REQUIRE(tile_records != nullptr && offsets != nullptr, "null tile records or offsets");
  GL_REQUIRE(num_tiles == 0 || (codes && coords && data && values_cm), "null tile array");
  const int64_t waves = num_tiles + kTileRecPad;
  const int64_t blocks = std::min<int64_t>((waves + 3) / 4, 65536);
  hipLaunchKernelGGL(k_exec_layout, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     num_tiles, codes, coords, offsets, data, (TileRec*)tile_records, values_cm);
  GL_LAUNCH_CHECK("k_exec_layout");
  return GL_OK;
}

int gl_pbr_spmm_i32_f32(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                        int64_t num_tiles, int64_t num_vals, int64_t num_rem,
                        const uint64_t* codes, const int32_t* coords,
                        const int32_t* offsets, const float* data, const int32_t* stripe_ptr,
                        const int32_t* rem_indptr, const int32_t* rem_col_ind,
                        const float* rem_vals, const float* X, int64_t ldx, int64_t K,
                        float* Y, int64_t ldy, int32_t accumulate, int32_t mode, int32_t kw,
                        void* stream) {
  return gl_pbr_spmm_exec_i32_f32(rows, cols, block_rows, block_cols, num_tiles, num_vals,
                                  num_rem, codes, coords, offsets, data, stripe_ptr, rem_indptr,
                                  rem_col_ind, rem_vals, nullptr, nullptr, X, ldx, K, Y, ldy,
                                  accumulate, mode, kw, stream);
}

int gl_pbr_spmm_exec_i32_f32(int64_t rows, int64_t cols, int32_t block_rows,
                             int32_t block_cols, int64_t num_tiles, int64_t num_vals,
                             int64_t num_rem, const uint64_t* codes, const int32_t* coords,
                             const int32_t* offsets, const float* data,
                             const int32_t* stripe_ptr, const int32_t* rem_indptr,
                             const int32_t* rem_col_ind, const float* rem_vals,
                             const void* tile_records, const float* values_cm, const float* X,
                             int64_t ldx, int64_t K, float* Y, int64_t ldy, int32_t accumulate,
                             int32_t mode, int32_t kw, void* stream) {
  int rc = check_pbr(rows, cols, block_rows, block_cols, num_tiles, codes, coords, offsets,
                     data, stripe_ptr);
  if (rc) return rc;
  if ((rc = check_dense(rows, cols, X, ldx, K, Y, ldy))) return rc;
  GL_REQUIRE(rem_indptr != nullptr || rows == 0, "null remainder indptr");
  GL_REQUIRE(mode == GL_SPMM_FUSED || mode == GL_SPMM_SPLIT, "unknown mode %d", mode);
  if (kw == 0) kw = choose_kw(K);
  GL_REQUIRE(valid_kw(kw), "kw must be one of 1,2,4,8,16,32,64");
  if (rows == 0) return GL_OK;
  hipStream_t st = (hipStream_t)stream;
  if (mode == GL_SPMM_SPLIT) {
    rc = gl_pbr_block_spmm_i32_f32(rows, cols, block_rows, block_cols, num_tiles, num_vals,
                                   codes, coords, offsets, data, stripe_ptr, X, ldx, K, Y, ldy,
                                   accumulate, kw, stream);
    if (rc) return rc;
    // the remainder kernel follows the block kernel on the same stream: both own whole
    // rows of Y, so the second one adds without atomics (the fused mode is the one that
    // runs tiles and remainder side by side, inside one launch)
    rc = gl_csr_spmm_i32_f32(rows, cols, num_rem, rem_indptr, rem_col_ind, rem_vals, X, ldx, K,
                             Y, ldy, 1, kw, stream);
    if (rc) return rc;
    return GL_OK;
  }
  PbrView m{rows, cols, (rows + block_rows - 1) / block_rows, num_tiles, num_vals, num_rem,
            codes, coords, offsets, data, stripe_ptr, rem_indptr, rem_col_ind, rem_vals};
  if (block_rows == 8) {
    m.trec = tile_records;
    m.tdata = values_cm;
  }
  const int nwg = grid_for(m.num_stripes, kSpmmWgPerCu);
  for (int64_t k0 = 0, w = kw; k0 < K; k0 += w) {
    w = chunk_kw(kw, K - k0);
    RhsView d{X, ldx, K, k0, cols * ldx};
    EpiStore e{Y, ldy, accumulate, fits_u24(rows, ldy)};
    if ((rc = launch_stripe<EpiStore, true, true>(m, d, block_rows, (int)w, e, nwg, st)))
      return rc;
  }
  return GL_OK;
}

}  // extern "C"
