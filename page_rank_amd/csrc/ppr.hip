// Personalized PageRank power iteration: setup kernels, the per-iteration retire kernel and
// the host loop (eager or hipGraph-captured, with device-side stopping).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "gl_common.hpp"
#include "launch.hpp"
#include "stripe_kernels.hpp"

namespace gl {

constexpr int kMaxWg = 4096;      // cap of persistent workgroups (partials rows)
constexpr int kRetireThreads = 1024;

__global__ void k_fill(float* __restrict__ x, int64_t n, float v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

// One-hot (seed-set) start: X[row, col] += w for each seed entry (entries are unique).
__global__ void k_seed_scatter(float* __restrict__ x, int64_t ld, const int32_t* __restrict__ srow,
                               const int32_t* __restrict__ scol, const float* __restrict__ sw,
                               int64_t n) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n) x[(int64_t)srow[e] * ld + scol[e]] = sw[e];
}

// sptr[s] = first seed entry (sorted by row) whose row >= s * B.
__global__ void k_seed_stripes(int32_t* __restrict__ sptr, int64_t num_stripes, int32_t B,
                               const int32_t* __restrict__ srow, int64_t n) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s <= num_stripes;
       s += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = n;
    const int64_t r = s * B;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)srow[mid] < r) lo = mid + 1;
      else hi = mid;
    }
    sptr[s] = (int32_t)lo;
  }
}

// Dangling mass of the start vector, dsum[k] = sum_{i: dmask[i]} X[i, k], in a fixed order:
// block b sums rows [b * rpb, (b + 1) * rpb) (thread t: column k0 + t % W, every RP-th row
// from t / W), combines its threads in thread order, and k_dangling_final adds the block
// partials in block order. W = min(K, 256) columns per pass, RP = 256 / W.
constexpr int kDangBlocks = 1024;
static_assert(kDangBlocks <= kMaxWg, "dangling partials reuse the step's partial rows");

__global__ void __launch_bounds__(256) k_dangling_partial(const float* __restrict__ x,
                                                          int64_t N, int64_t K,
                                                          const uint8_t* __restrict__ dmask,
                                                          int64_t rpb, float* __restrict__ part) {
  __shared__ float red[256];
  const int W = (int)(K < 256 ? K : 256);
  const int RP = 256 / W;
  const int t = threadIdx.x;
  const int lc = t % W, lr = t / W;
  const int64_t r0 = blockIdx.x * rpb;
  const int64_t r1 = r0 + rpb < N ? r0 + rpb : N;
  for (int64_t k0 = 0; k0 < K; k0 += W) {
    const int64_t k = k0 + lc;
    float acc = 0.f;
    if (lr < RP && k < K)
      for (int64_t i = r0 + lr; i < r1; i += RP)
        if (dmask[i]) acc += x[i * K + k];
    red[t] = acc;
    __syncthreads();
    if (t < W && k0 + t < K) {
      float sum = 0.f;
      for (int r = 0; r < RP; ++r) sum += red[r * W + t];
      part[blockIdx.x * K + k0 + t] = sum;
    }
    __syncthreads();
  }
}

__global__ void k_dangling_final(const float* __restrict__ part, int64_t nb, int64_t K,
                                 float* __restrict__ dsum) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < K;
       k += (int64_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int64_t b = 0; b < nb; ++b) s += part[b * K + k];
    dsum[k] = s;
  }
}

struct ColState {
  int32_t* frozen;
  int32_t* iters;
  int32_t* conv;
  float* dsum;
  float* cu;
  float* tele;
  float* res_prev;  // residual of the previous pass (extrapolation)
  float* beta;      // extrapolation factor of the pass just retired (0 = none)
};

__device__ __forceinline__ void set_coeffs(const ColState& c, int k, float alpha, int64_t N,
                                           int dmode) {
  const float d = c.dsum[k];
  c.cu[k] = dmode == GL_DANGLING_UNIFORM ? alpha * d / (float)N : 0.f;
  c.tele[k] = (1.f - alpha) + (dmode == GL_DANGLING_PERSONALIZATION ? alpha * d : 0.f);
}

__global__ void k_ctl_init(PprCtl* ctl, ColState c, int64_t K, float alpha, int64_t N,
                           int dmode) {
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    c.frozen[k] = 0;
    c.iters[k] = 0;
    c.conv[k] = 0;
    c.res_prev[k] = 0.f;
    c.beta[k] = 0.f;
    set_coeffs(c, k, alpha, N, dmode);
  }
  if (threadIdx.x == 0) {
    ctl->iter = 0;
    ctl->active = (int32_t)K;
    ctl->kick = 0;
  }
}

// After each pass: sum the per-workgroup partials of every column (double), record the
// residual, freeze converged columns (or all when retiring together), stop everything at
// max_iter, and refresh the dangling coefficients for the next pass. One workgroup:
// consecutive threads take consecutive columns (coalesced rows of the partial arrays) and
// kRetireThreads / KC slices of threads split the workgroup rows; an LDS tree in fixed order
// combines the slices, so the sums are deterministic. The column residuals go through
// `res` (K floats of the workspace; the barriers order those global writes and reads within
// the workgroup), so K has no LDS-imposed limit.
constexpr int kUnroll = 8;
__global__ void __launch_bounds__(kRetireThreads)
k_retire(PprCtl* ctl, ColState c, int64_t K, int nwg, const float* __restrict__ part_res,
         const float* __restrict__ part_dang, float* __restrict__ res, float* __restrict__ hist,
         float tol, float alpha, int64_t N, int dmode, int together, int max_iter,
         int extrap_every) {
  __shared__ double s_r[kRetireThreads];
  __shared__ double s_d[kRetireThreads];
  __shared__ int s_max_bits;
  __shared__ int s_active;
  __shared__ int s_kick;
  if (ctl->active <= 0) return;  // the pass was skipped: nothing to retire
  const int it = ctl->iter;
  if (threadIdx.x == 0) {
    s_max_bits = 0;
    s_active = 0;
    s_kick = 0;
  }
  // an extrapolation follows passes extrap_every, 2 * extrap_every, ...
  const bool kick_pass = extrap_every > 0 && (it + 1) % extrap_every == 0;
  int KC = 64;  // columns per sweep: smallest power of two >= K, at most 64
  while (KC > 1 && KC / 2 >= K) KC /= 2;
  const int S = kRetireThreads / KC;  // slices
  const int kk = threadIdx.x % KC;
  const int sl = threadIdx.x / KC;
  __syncthreads();
  for (int64_t k0 = 0; k0 < K; k0 += KC) {
    const int64_t k = k0 + kk;
    double r = 0.0, dg = 0.0;
    if (k < K && !c.frozen[k]) {
      // kUnroll independent accumulators, so that many partial loads are in flight (each
      // comes from another XCD's writes), combined in a fixed order
      double ra[kUnroll], da[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) ra[u] = da[u] = 0.0;
      int w = sl;
      for (; w + (kUnroll - 1) * S < nwg; w += kUnroll * S) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          ra[u] += part_res[(int64_t)(w + u * S) * K + k];
          da[u] += part_dang[(int64_t)(w + u * S) * K + k];
        }
      }
      for (; w < nwg; w += S) {
        ra[0] += part_res[(int64_t)w * K + k];
        da[0] += part_dang[(int64_t)w * K + k];
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        r += ra[u];
        dg += da[u];
      }
    }
    s_r[threadIdx.x] = r;
    s_d[threadIdx.x] = dg;
    __syncthreads();
    for (int h = S / 2; h >= 1; h /= 2) {
      if (sl < h) {
        s_r[threadIdx.x] += s_r[threadIdx.x + h * KC];
        s_d[threadIdx.x] += s_d[threadIdx.x + h * KC];
      }
      __syncthreads();
    }
    if (sl == 0 && k < K) {
      const float rf = (float)s_r[kk];
      res[k] = rf;  // read back below by other threads of this workgroup
      if (!c.frozen[k]) {
        // beta[k] keeps the previous iterate's dangling mass until the sweep below sets
        // the column's extrapolation factor
        c.beta[k] = c.dsum[k];
        c.dsum[k] = (float)s_d[kk];
        atomicMax(&s_max_bits, __float_as_int(rf));  // rf >= 0: int order == float order
      }
    }
    __syncthreads();
  }
  const float maxres = __int_as_float(s_max_bits);
  const bool last = it + 1 >= max_iter;
  for (int64_t k = threadIdx.x; k < K; k += blockDim.x) {
    if (c.frozen[k]) {
      if (hist) hist[(int64_t)it * K + k] = 0.f;
      continue;
    }
    const float r = res[k];
    if (hist) hist[(int64_t)it * K + k] = r;
    const bool conv = together ? (maxres < tol) : (r < tol);
    const float dsum_prev = c.beta[k];
    float beta = 0.f;
    if (conv || last) {
      c.frozen[k] = 1;
      c.iters[k] = it + 1;
      c.conv[k] = conv ? 1 : 0;
    } else {
      atomicAdd(&s_active, 1);
      // Extrapolation of the dominant error mode (Aitken's delta-squared on the residual
      // ratio, as Kamvar et al. apply it to PageRank): once the residual shrinks by a
      // steady ratio rho per pass, the remaining error is about rho / (1 - rho) times the
      // last update, so X_{t+1} += beta (X_{t+1} - X_t) with beta = rho / (1 - rho). Only
      // for slowly mixing columns (rho in [0.5, 0.97]), capped at 4.
      const float rp = c.res_prev[k];
      if (kick_pass && rp > 0.f) {
        const float rho = r / rp;
        if (rho >= 0.5f && rho <= 0.97f) beta = fminf(rho / (1.f - rho), 4.f);
      }
      if (beta > 0.f) {
        // the dangling mass is linear in X: that of the extrapolated iterate
        c.dsum[k] = c.dsum[k] + beta * (c.dsum[k] - dsum_prev);
        s_kick = 1;  // benign race: every writer stores 1
      }
      set_coeffs(c, (int)k, alpha, N, dmode);
    }
    c.beta[k] = beta;
    c.res_prev[k] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl->iter = it + 1;
    ctl->active = s_active;
    ctl->kick = s_kick;
  }
}

// The extrapolation step that k_retire scheduled: xn[i][k] += beta[k] (xn[i][k] - xc[i][k])
// for every column with beta > 0 (xn = X_{t+1}, the next pass's input; xc = X_t). A grid
// stride loop over rows, consecutive threads on consecutive columns; exits at once when no
// column extrapolates.
__global__ void __launch_bounds__(256) k_extrapolate(const PprCtl* __restrict__ ctl,
                                                     const float* __restrict__ beta,
                                                     const float* __restrict__ xc,
                                                     float* __restrict__ xn, int64_t N,
                                                     int64_t K) {
  if (!ctl->kick) return;
  const int64_t total = N * K;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float b = beta[e % K];
    if (b > 0.f) xn[e] = __builtin_fmaf(b, xn[e] - xc[e], xn[e]);
  }
}

// The whole run in one workgroup, for graphs so small that a pass of the multi-workgroup
// loop (a step kernel, the retire kernel and a graph replay every few passes: about 20 us
// whatever the size) costs more than the pass's own work on one CU. Every pass, thread t
// computes column k = t % KC of rows sl, sl + S, ... (KC = smallest power of two >= K,
// S = kSingleThreads / KC, sl = t / KC) exactly as the step epilogue does:
//   X_{t+1}[i,k] = alpha (A X_t)[i,k] + cu[k] + w[i,k] tele[k]   (frozen columns copied)
// with the L1 residual and the dangling mass of its elements. Per column, the threads' sums
// are combined in double in a fixed order (within each wave, then across the 16 waves); then
// the retire rule of k_retire (tolerance,
// together or per column, max_iter, extrapolation factor, dangling coefficients) runs on the
// column state, which lives in LDS for the whole launch, and an extrapolation pass follows
// when a column was given a factor. X_t and X_{t+1} are global buffers that only this
// workgroup touches; the barriers order its writes and reads. The loop ends when no column
// is active, at the latest after pass max_iter (which freezes every column), and leaves the
// final iterate in x0buf. Tiles are decoded per row from the code: the row's bits
// (code >> r B) and its first value offset + popcount of the bits before them.
constexpr int kSingleThreads = 1024;
constexpr int kSingleMaxK = 64;

__global__ void __launch_bounds__(kSingleThreads)
k_ppr_single(PbrView m, int32_t B, int64_t K, float* x0buf, float* x1buf, ColState c,
             PprCtl* ctl, const int32_t* __restrict__ sptr, const int32_t* __restrict__ srow,
             const int32_t* __restrict__ scol, const float* __restrict__ sw,
             const uint8_t* __restrict__ dmask, float* __restrict__ hist, float tol, float alpha,
             int dmode, int together, int max_iter, int extrap_every) {
  __shared__ double s_r[kSingleThreads / 64 * kSingleMaxK];
  __shared__ double s_d[kSingleThreads / 64 * kSingleMaxK];
  __shared__ float s_cu[kSingleMaxK], s_tele[kSingleMaxK], s_dsum[kSingleMaxK];
  __shared__ float s_res[kSingleMaxK], s_rprev[kSingleMaxK], s_beta[kSingleMaxK];
  __shared__ int s_frozen[kSingleMaxK];
  __shared__ int s_max_bits, s_active, s_kick;
  const int t = threadIdx.x;
  const int64_t N = m.rows;
  int KC = kSingleMaxK;
  while (KC > 1 && KC / 2 >= K) KC /= 2;
  const int S = kSingleThreads / KC;
  const int k = t % KC;
  const int sl = t / KC;
  const bool on = k < K;
  if (t < K) {
    s_cu[t] = c.cu[t];
    s_tele[t] = c.tele[t];
    s_dsum[t] = c.dsum[t];
    s_rprev[t] = 0.f;
    s_beta[t] = 0.f;
    s_frozen[t] = 0;
  }
  if (t == 0) {
    s_max_bits = 0;
    s_active = 0;
    s_kick = 0;
  }
  __syncthreads();
  const uint64_t rmask = B == 8 ? 0xffull : ((1ull << B) - 1);
  float* xc = x0buf;
  float* xn = x1buf;
  int it = 0;
  for (;;) {
    float res = 0.f, dang = 0.f;
    if (on) {
      const bool frz = s_frozen[k] != 0;
      const float cuv = s_cu[k], tev = s_tele[k];
      for (int64_t i = sl; i < N; i += S) {
        const float xold = xc[i * K + k];
        float xv = xold;
        if (!frz) {
          const int64_t s = i / B;
          const int r = (int)(i - s * B);
          float acc = 0.f;
          for (int32_t tt = m.stripe_ptr[s]; tt < m.stripe_ptr[s + 1]; ++tt) {
            const uint64_t code = m.codes[tt];
            uint64_t bits = (code >> (r * B)) & rmask;
            int32_t v = m.offsets[tt] + __popcll(code & ((1ull << (r * B)) - 1));
            const int64_t col0 = m.coords[2 * tt + 1];
            while (bits) {
              const int cc = __ffsll((long long)bits) - 1;
              bits &= bits - 1;
              acc = __builtin_fmaf(m.data[v++], xc[(col0 + cc) * K + k], acc);
            }
          }
          for (int32_t j = m.rem_ptr[i]; j < m.rem_ptr[i + 1]; ++j)
            acc = __builtin_fmaf(m.rem_val[j], xc[(int64_t)m.rem_col[j] * K + k], acc);
          float w = 0.f;
          for (int32_t e = sptr[s]; e < sptr[s + 1]; ++e)
            if (srow[e] == i && scol[e] == k) w += sw[e];
          xv = __builtin_fmaf(alpha, acc, cuv);
          xv = __builtin_fmaf(w, tev, xv);
          res += fabsf(xv - xold);
          if (dmask != nullptr && dmask[i]) dang += xv;
        }
        xn[i * K + k] = xv;
      }
    }
    // per column: a butterfly over the lanes of a wave that hold the column (lane distances
    // 32, 16, ..., KC; both partners add the same two values, so every lane ends with the
    // same bits), then the 16 wave sums in wave order
    {
      double r = res, d = dang;
      for (int off = 32; off >= KC; off >>= 1) {
        r += __shfl_xor(r, off, 64);
        d += __shfl_xor(d, off, 64);
      }
      const int lane = t & 63;
      if (lane < KC) {
        s_r[(t >> 6) * KC + lane] = r;
        s_d[(t >> 6) * KC + lane] = d;
      }
    }
    __syncthreads();
    // per column (the same rule as k_retire)
    float dsum_prev = 0.f;
    if (t < K && !s_frozen[t]) {
      double r = 0.0, d = 0.0;
      for (int wv = 0; wv < kSingleThreads / 64; ++wv) {
        r += s_r[wv * KC + t];
        d += s_d[wv * KC + t];
      }
      const float rf = (float)r;
      s_res[t] = rf;
      dsum_prev = s_dsum[t];
      s_dsum[t] = (float)d;
      atomicMax(&s_max_bits, __float_as_int(rf));  // rf >= 0: int order == float order
    }
    __syncthreads();
    if (t < K) {
      if (s_frozen[t]) {
        if (hist) hist[(int64_t)it * K + t] = 0.f;
      } else {
        const float r = s_res[t];
        const float maxres = __int_as_float(s_max_bits);
        if (hist) hist[(int64_t)it * K + t] = r;
        const bool conv = together ? (maxres < tol) : (r < tol);
        float beta = 0.f;
        if (conv || it + 1 >= max_iter) {
          s_frozen[t] = 1;
          c.iters[t] = it + 1;
          c.conv[t] = conv ? 1 : 0;
        } else {
          atomicAdd(&s_active, 1);
          const float rp = s_rprev[t];
          if (extrap_every > 0 && (it + 1) % extrap_every == 0 && rp > 0.f) {
            const float rho = r / rp;
            if (rho >= 0.5f && rho <= 0.97f) beta = fminf(rho / (1.f - rho), 4.f);
          }
          if (beta > 0.f) {
            s_dsum[t] = s_dsum[t] + beta * (s_dsum[t] - dsum_prev);
            s_kick = 1;
          }
          const float d = s_dsum[t];
          s_cu[t] = dmode == GL_DANGLING_UNIFORM ? alpha * d / (float)N : 0.f;
          s_tele[t] = (1.f - alpha) + (dmode == GL_DANGLING_PERSONALIZATION ? alpha * d : 0.f);
        }
        s_beta[t] = beta;
        s_rprev[t] = r;
      }
    }
    __syncthreads();
    const int active = s_active;
    if (s_kick && on) {
      // each thread moves its own elements: X_{t+1} += beta (X_{t+1} - X_t)
      const float b = s_beta[k];
      if (b > 0.f)
        for (int64_t i = sl; i < N; i += S) {
          const float y = xn[i * K + k];
          xn[i * K + k] = __builtin_fmaf(b, y - xc[i * K + k], y);
        }
    }
    ++it;
    float* tmp = xc;
    xc = xn;
    xn = tmp;
    __syncthreads();  // every thread has read s_active and s_kick; X_{t+1} is written
    if (active == 0) break;
    if (t == 0) {
      s_max_bits = 0;
      s_active = 0;
      s_kick = 0;
    }
  }
  // the last pass wrote xc; its own elements go to x0buf by the thread that wrote them
  if (xc != x0buf && on)
    for (int64_t i = sl; i < N; i += S) x0buf[i * K + k] = xc[i * K + k];
  if (t < K) c.frozen[t] = 1;
  if (t == 0) {
    ctl->iter = it;
    ctl->active = 0;
    ctl->kick = 0;
  }
}

// Whether the one-workgroup run is chosen for a graph of nnz entries and K columns. Measured
// on MI355X, its pass costs about 1 us per 1,000 entry gathers (about 100 us for the
// 10,000-node, 99k-entry SBM, against about 22 us for a pass of the multi-workgroup loop), so
// it is the default only for graphs of a few thousand entries.
static bool single_wg_fits(int64_t nnz, int64_t K) {
  return K <= kSingleMaxK && nnz * K <= ((int64_t)1 << 13);
}

static int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }

struct Workspace {
  float* xb;
  PprCtl* ctl;
  int32_t* frozen;
  float* dsum;
  float* cu;
  float* tele;
  float* res;
  float* res_prev;
  float* beta;
  float* part_res;
  float* part_dang;
  int32_t* sptr;
  int32_t* srow;
  int32_t* scol;
  float* sw;
  int64_t bytes;
};

static Workspace carve(char* base, int64_t rows, int32_t B, int64_t K, int64_t E) {
  Workspace w;
  int64_t off = 0;
  auto take = [&](int64_t n) {
    char* p = base ? base + off : nullptr;
    off += align256(n);
    return p;
  };
  const int64_t ns = (rows + B - 1) / B;
  w.xb = (float*)take(rows * K * 4);
  w.ctl = (PprCtl*)take(sizeof(PprCtl));
  w.frozen = (int32_t*)take(K * 4);
  w.dsum = (float*)take(K * 4);
  w.cu = (float*)take(K * 4);
  w.tele = (float*)take(K * 4);
  w.res = (float*)take(K * 4);
  w.res_prev = (float*)take(K * 4);
  w.beta = (float*)take(K * 4);
  w.part_res = (float*)take((int64_t)kMaxWg * K * 4);
  w.part_dang = (float*)take((int64_t)kMaxWg * K * 4);
  w.sptr = (int32_t*)take((ns + 1) * 4);
  w.srow = (int32_t*)take(std::max<int64_t>(E, 1) * 4);
  w.scol = (int32_t*)take(std::max<int64_t>(E, 1) * 4);
  w.sw = (float*)take(std::max<int64_t>(E, 1) * 4);
  w.bytes = off;
  return w;
}

static int blocks_for(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (int)b;
}

}  // namespace gl

using namespace gl;

extern "C" {

int64_t gl_ppr_workspace_bytes(int64_t rows, int32_t block_rows, int64_t K,
                               int64_t num_seed_entries) {
  if (rows < 0 || K < 1 || block_rows < 1 || num_seed_entries < 0) return -1;
  return carve(nullptr, rows, block_rows, K, num_seed_entries).bytes;
}

int gl_ppr_run_i32_f32(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                       int64_t num_tiles, int64_t num_vals, int64_t num_rem,
                       const uint64_t* codes, const int32_t* coords,
                       const int32_t* offsets, const float* data, const int32_t* stripe_ptr,
                       const int32_t* rem_indptr, const int32_t* rem_col_ind,
                       const float* rem_vals, int64_t K, const int32_t* seed_ptr,
                       const int32_t* seed_idx, const float* seed_w,
                       const uint8_t* dangling_mask, const float* X0, const gl_ppr_opts* opts,
                       float* X_out, int32_t* iterations, int32_t* converged, float* history,
                       void* workspace, int64_t workspace_bytes, int32_t* total_iterations,
                       void* stream) {
  GL_REQUIRE(opts != nullptr, "opts is null");
  GL_REQUIRE(rows > 0 && rows == cols, "PPR needs a square matrix with rows > 0");
  GL_REQUIRE(block_rows == block_cols && (block_rows == 2 || block_rows == 4 || block_rows == 8),
             "Unsupported block size %dx%d: only 2x2, 4x4 and 8x8 are supported", block_rows,
             block_cols);
  GL_REQUIRE(rows <= INT32_MAX, "rows exceed int32 indices");
  GL_REQUIRE(K >= 1 && K <= INT32_MAX, "K must be in [1, 2^31)");
  GL_REQUIRE(seed_ptr && seed_idx, "seed arrays are null");
  GL_REQUIRE(X_out && iterations && converged, "null output");
  GL_REQUIRE(stripe_ptr && rem_indptr, "null stripe_ptr / remainder indptr");
  GL_REQUIRE(num_tiles == 0 || (codes && coords && offsets && data), "null tile arrays");
  GL_REQUIRE(opts->max_iter >= 0, "max_iter must be >= 0");
  GL_REQUIRE(opts->extrapolate_every >= 0 && opts->extrapolate_every != 1,
             "extrapolate_every must be 0 (off) or >= 2");
  GL_REQUIRE(opts->dangling_mode >= 0 && opts->dangling_mode <= 2, "bad dangling_mode");
  GL_REQUIRE(opts->dangling_mode == GL_DANGLING_NONE || dangling_mask != nullptr,
             "dangling_mask required for dangling_mode %d", opts->dangling_mode);
  GL_REQUIRE(opts->single_wg <= 0 || K <= kSingleMaxK,
             "the one-workgroup run takes at most %d columns (K = %lld)", kSingleMaxK,
             (long long)K);
  int kw = opts->kw ? opts->kw : choose_kw(K);
  GL_REQUIRE(valid_kw(kw), "kw must be one of 1,2,4,8,16,32,64");

  // ---- seed entries: merge duplicates, sort by (row, col) ----
  GL_REQUIRE(seed_ptr[0] == 0, "seed_ptr[0] must be 0");
  struct Ent {
    int32_t row, col;
    float w;
  };
  std::vector<Ent> ents;
  for (int64_t k = 0; k < K; ++k) {
    GL_REQUIRE(seed_ptr[k + 1] >= seed_ptr[k], "seed_ptr not monotone");
    GL_REQUIRE(seed_ptr[k + 1] > seed_ptr[k], "seed set %lld is empty", (long long)k);
    for (int32_t e = seed_ptr[k]; e < seed_ptr[k + 1]; ++e) {
      const int32_t i = seed_idx[e];
      GL_REQUIRE(i >= 0 && i < rows, "seed node %d out of range", i);
      const float w = seed_w ? seed_w[e] : 1.f / (float)(seed_ptr[k + 1] - seed_ptr[k]);
      ents.push_back({i, (int32_t)k, w});
    }
  }
  std::sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
    return a.row != b.row ? a.row < b.row : a.col < b.col;
  });
  std::vector<int32_t> hrow, hcol;
  std::vector<float> hw;
  for (const Ent& e : ents) {
    if (!hrow.empty() && hrow.back() == e.row && hcol.back() == e.col) {
      hw.back() += e.w;
    } else {
      hrow.push_back(e.row);
      hcol.push_back(e.col);
      hw.push_back(e.w);
    }
  }
  const int64_t E = (int64_t)hrow.size();

  const int64_t need = carve(nullptr, rows, block_rows, K, E).bytes;
  GL_REQUIRE(workspace != nullptr, "workspace is null");
  if (workspace_bytes < need) {
    set_error("workspace too small: %lld < %lld bytes", (long long)workspace_bytes,
              (long long)need);
    return GL_ERR_WORKSPACE;
  }
  Workspace w = carve((char*)workspace, rows, block_rows, K, E);
  const int64_t ns = (rows + block_rows - 1) / block_rows;

  // the run's own stream (graph capture needs one; the caller's may be the null stream),
  // leased for this call only, so concurrent runs never share it
  StreamLease lease;
  if (!lease.s) return GL_ERR_HIP;
  hipStream_t user = (hipStream_t)stream;
  hipStream_t st = lease.s->stream;
  GL_HIP(hipEventRecord(lease.s->ev_in, user));
  GL_HIP(hipStreamWaitEvent(st, lease.s->ev_in, 0));

  GL_HIP(hipMemcpyAsync(w.srow, hrow.data(), E * 4, hipMemcpyHostToDevice, st));
  GL_HIP(hipMemcpyAsync(w.scol, hcol.data(), E * 4, hipMemcpyHostToDevice, st));
  GL_HIP(hipMemcpyAsync(w.sw, hw.data(), E * 4, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_seed_stripes, dim3(blocks_for(ns + 1, 256)), dim3(256), 0, st, w.sptr,
                     ns, block_rows, w.srow, E);
  GL_LAUNCH_CHECK("k_seed_stripes");

  // ---- start vector in X_out (buffer 0); X0 == X_out is a warm start in place ----
  const int64_t NK = rows * K;
  if (X0) {
    if (X0 != X_out) GL_HIP(hipMemcpyAsync(X_out, X0, NK * 4, hipMemcpyDeviceToDevice, st));
  } else {
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(NK, 256)), dim3(256), 0, st, X_out, NK, 0.f);
    GL_LAUNCH_CHECK("k_fill");
    hipLaunchKernelGGL(k_seed_scatter, dim3(blocks_for(E, 256)), dim3(256), 0, st, X_out, K,
                       w.srow, w.scol, w.sw, E);
    GL_LAUNCH_CHECK("k_seed_scatter");
  }
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(K, 256)), dim3(256), 0, st, w.dsum, K, 0.f);
  if (opts->dangling_mode != GL_DANGLING_NONE) {
    // block partials in w.part_res (kMaxWg >= kDangBlocks rows of K), unused until the loop
    const int64_t rpb = (rows + kDangBlocks - 1) / kDangBlocks;
    const int64_t nb = (rows + rpb - 1) / rpb;
    hipLaunchKernelGGL(k_dangling_partial, dim3((unsigned)nb), dim3(256), 0, st, X_out, rows, K,
                       dangling_mask, rpb, w.part_res);
    GL_LAUNCH_CHECK("k_dangling_partial");
    hipLaunchKernelGGL(k_dangling_final, dim3(blocks_for(K, 256)), dim3(256), 0, st,
                       (const float*)w.part_res, nb, K, w.dsum);
    GL_LAUNCH_CHECK("k_dangling_final");
  }
  ColState cs{w.frozen, iterations, converged, w.dsum, w.cu, w.tele, w.res_prev, w.beta};
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1024), 0, st, w.ctl, cs, K, opts->alpha, rows,
                     opts->dangling_mode);
  GL_LAUNCH_CHECK("k_ctl_init");
  if (history && opts->max_iter > 0) {
    hipLaunchKernelGGL(k_fill, dim3(blocks_for((int64_t)opts->max_iter * K, 256)), dim3(256),
                       0, st, history, (int64_t)opts->max_iter * K, 0.f);
    GL_LAUNCH_CHECK("k_fill");
  }
  if (opts->max_iter == 0) {
    // no pass: X_out holds the start vector, every column 0 iterations, not converged
    // (the reference's loop body never runs and returns (X, 0, False))
    if (total_iterations) *total_iterations = 0;
    GL_HIP(hipEventRecord(lease.s->ev_out, st));
    GL_HIP(hipStreamWaitEvent(user, lease.s->ev_out, 0));
    return GL_OK;
  }

  PbrView m{rows, cols, ns, num_tiles, num_vals, num_rem, codes, coords, offsets, data,
            stripe_ptr, rem_indptr, rem_col_ind, rem_vals};
  const int together = opts->retire_mode == GL_RETIRE_TOGETHER ? 1 : 0;
  static thread_local PprCtl* h_ctl = nullptr;
  if (!h_ctl) GL_HIP(hipHostMalloc((void**)&h_ctl, sizeof(PprCtl), hipHostMallocDefault));

  if (opts->single_wg > 0 || (opts->single_wg == 0 && single_wg_fits(num_vals + num_rem, K))) {
    hipLaunchKernelGGL(k_ppr_single, dim3(1), dim3(kSingleThreads), 0, st, m, block_rows, K,
                       X_out, w.xb, cs, w.ctl, (const int32_t*)w.sptr, (const int32_t*)w.srow,
                       (const int32_t*)w.scol, (const float*)w.sw,
                       opts->dangling_mode != GL_DANGLING_NONE ? dangling_mask : nullptr,
                       history, opts->tol, opts->alpha, opts->dangling_mode, together,
                       opts->max_iter, opts->extrapolate_every);
    GL_LAUNCH_CHECK("k_ppr_single");
    GL_HIP(hipMemcpyAsync(h_ctl, w.ctl, sizeof(PprCtl), hipMemcpyDeviceToHost, st));
    GL_HIP(hipStreamSynchronize(st));
    if (total_iterations) *total_iterations = h_ctl->iter;
    GL_HIP(hipEventRecord(lease.s->ev_out, st));
    GL_HIP(hipStreamWaitEvent(user, lease.s->ev_out, 0));
    return GL_OK;
  }

  int nwg = grid_for(ns, opts->grid_mult);
  if (opts->grid_mult == 0) {
    // default grid: what the device holds of the kernel at once
    const EpiPpr e0{X_out, w.xb, K, opts->alpha, w.cu, w.tele, w.frozen,
                    opts->dangling_mode != GL_DANGLING_NONE ? dangling_mask : nullptr,
                    w.sptr, w.srow, w.scol, w.sw, w.ctl, w.part_res, w.part_dang, K, 0,
                    fits_u24(rows, K), rows * K, 0.f, 0.f, 0, 0};
    nwg = stripe_grid<EpiPpr, true, true>(m, RhsView{X_out, K, K, 0, rows * K}, block_rows, kw,
                                          e0, nwg);
  }
  if (nwg > kMaxWg) nwg = kMaxWg - (kMaxWg % kNumXcd);

  float* bufs[2] = {X_out, w.xb};
  const int extrap_blocks = blocks_for(NK, 256) < 8 * device_info().num_cus
                                ? blocks_for(NK, 256)
                                : 8 * device_info().num_cus;
  auto one_iteration = [&](int parity) -> int {
    const float* xc = bufs[parity];
    float* xn = bufs[parity ^ 1];
    for (int64_t k0 = 0, cw = kw; k0 < K; k0 += cw) {
      cw = chunk_kw(kw, K - k0);
      RhsView d{xc, K, K, k0, rows * K};
      EpiPpr e{xc, xn, K, opts->alpha, w.cu, w.tele, w.frozen,
               opts->dangling_mode != GL_DANGLING_NONE ? dangling_mask : nullptr,
               w.sptr, w.srow, w.scol, w.sw, w.ctl, w.part_res, w.part_dang, K, k0,
               fits_u24(rows, K), rows * K, 0.f, 0.f, 0, 0};
      int rc = launch_stripe<EpiPpr, true, true>(m, d, block_rows, (int)cw, e, nwg, st);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_retire, dim3(1), dim3(kRetireThreads), 0, st, w.ctl, cs, K, nwg,
                       w.part_res, w.part_dang, w.res, history, opts->tol, opts->alpha, rows,
                       opts->dangling_mode, together, opts->max_iter, opts->extrapolate_every);
    GL_LAUNCH_CHECK("k_retire");
    if (opts->extrapolate_every > 0) {
      hipLaunchKernelGGL(k_extrapolate, dim3(extrap_blocks), dim3(256), 0, st, w.ctl, w.beta,
                         xc, xn, rows, K);
      GL_LAUNCH_CHECK("k_extrapolate");
    }
    return GL_OK;
  };

  const int M = opts->graph_iters > 0 ? opts->graph_iters : 0;
  int done_iters = 0;
  if (M > 0) {
    // M iterations per graph; M even keeps the buffer parity of every replay identical.
    const int Mg = (M % 2) ? M + 1 : M;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = GL_OK;
    for (int i = 0; i < Mg && rc == GL_OK; ++i) rc = one_iteration(i & 1);
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    GL_HIP(ce);
    GL_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    for (;;) {
      GL_HIP(hipGraphLaunch(exec, st));
      GL_HIP(hipMemcpyAsync(h_ctl, w.ctl, sizeof(PprCtl), hipMemcpyDeviceToHost, st));
      GL_HIP(hipStreamSynchronize(st));
      if (h_ctl->active <= 0) break;
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    done_iters = h_ctl->iter;
  } else {
    for (int it = 0;; ++it) {
      int rc = one_iteration(it & 1);
      if (rc) return rc;
      GL_HIP(hipMemcpyAsync(h_ctl, w.ctl, sizeof(PprCtl), hipMemcpyDeviceToHost, st));
      GL_HIP(hipStreamSynchronize(st));
      if (h_ctl->active <= 0) break;
    }
    done_iters = h_ctl->iter;
  }
  // the last executed pass wrote buffer (done_iters % 2); every column's final value is there
  if (done_iters % 2 == 1)
    GL_HIP(hipMemcpyAsync(X_out, w.xb, NK * 4, hipMemcpyDeviceToDevice, st));
  if (total_iterations) *total_iterations = done_iters;
  GL_HIP(hipEventRecord(lease.s->ev_out, st));
  GL_HIP(hipStreamWaitEvent(user, lease.s->ev_out, 0));
  return GL_OK;
}

}  // extern "C"
