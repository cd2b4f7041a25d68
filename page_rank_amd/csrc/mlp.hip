// Learned warm start for PPR: per-node features, a 12 -> 64 -> 64 -> 1 MLP (ReLU, ReLU,
// softplus) with hand-written forward / backward / Adam kernels, and the assembly of a
// normalized warm-start matrix X0[N][K].
//
// Features of (node i, seed set k), all FP32:
//   0  in-degree bucket  b(d) = round(8 * log2(1 + d)) / 8, d clamped to 2^24 (exact)
//   1  out-degree bucket (same)
//   2..7  one-hot hop distance from the seed set: 0, 1, 2, 3, 4, more
//   8..10 forward-push values log1p(N * q_t) / 8 for t = 1, 2, 3, where
//         q_t = (1 - alpha) * sum_{j<t} alpha^j (A^j p_k)[i]  (the first t power-iteration
//         terms; v_j = A^j p_k come from the PBR SpMM)
//   11 constant 1
// Beyond hop 2 the push values are zero, so the network output depends only on the two
// degree buckets and the hop class (3, 4 or more): those pairs read a table evaluated once
// per (in-bucket, out-bucket, class); pairs within 2 hops are listed and evaluated one by
// one. The column is then normalized to sum 1.
#include <algorithm>
#include <cmath>

#include "gl_common.hpp"

namespace gl {

constexpr int kIn = 12;
constexpr int kHid = 64;
// parameter layout: W1[64][12] b1[64] W2[64][64] b2[64] w3[64] b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kIn;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kNumParams = kOffB3 + 1;  // 5057
constexpr int kBuckets = 193;           // round(8 * log2(1 + 2^24)) + 1

// Evaluated in double: 8 * log2(1 + d) never comes closer than 5.9e-8 to a rounding tie
// for d <= 2^24, far above the double error, so the bucket is the exact one (float32
// misplaces 126 degrees below 2^24).
__device__ __forceinline__ int deg_bucket(int d) {
  const double x = log2(1.0 + (double)min(d, 1 << 24));
  return min(kBuckets - 1, (int)rint(8.0 * x));
}

__device__ __forceinline__ float softplus(float o) {
  return o > 20.f ? o : log1pf(expf(o));
}
__device__ __forceinline__ float sigmoid(float o) { return 1.f / (1.f + expf(-o)); }

// Forward pieces of one sample. Weight indices are uniform across the wave (scalar loads).
__device__ __forceinline__ void mlp_h1(const float* __restrict__ w, const float (&f)[kIn],
                                       float (&h1)[kHid]) {
#pragma unroll
  for (int j = 0; j < kHid; ++j) {
    float s = w[kOffB1 + j];
#pragma unroll
    for (int i = 0; i < kIn; ++i) s = __builtin_fmaf(w[kOffW1 + j * kIn + i], f[i], s);
    h1[j] = fmaxf(s, 0.f);
  }
}
__device__ __forceinline__ float mlp_h2(const float* __restrict__ w, const float (&h1)[kHid],
                                        int j) {
  float s = w[kOffB2 + j];
#pragma unroll
  for (int i = 0; i < kHid; ++i) s = __builtin_fmaf(w[kOffW2 + j * kHid + i], h1[i], s);
  return fmaxf(s, 0.f);
}
// pre-activation output; h2 is folded into the sum and never held as an array
__device__ __forceinline__ float mlp_out(const float* __restrict__ w, const float (&f)[kIn]) {
  float h1[kHid];
  mlp_h1(w, f, h1);
  float o = w[kOffB3];
#pragma unroll
  for (int j = 0; j < kHid; ++j) o = __builtin_fmaf(w[kOffW3 + j], mlp_h2(w, h1, j), o);
  return o;
}

// s[m] = softplus(MLP(feat[m])), one sample per thread. Weight indices are wave-uniform,
// so the weights arrive through the scalar cache as SGPR operands of the FMAs.
__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ w,
                                                     const float* __restrict__ feat, int64_t M,
                                                     float* __restrict__ out) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < M;
       m += (int64_t)gridDim.x * blockDim.x) {
    float f[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) f[i] = feat[m * kIn + i];
    out[m] = softplus(mlp_out(w, f));
  }
}

// One pass of training over M samples: weighted L1 loss sum_m w_m |s_m - t_m| (s the
// softplus output). A wave takes 64 samples at a time (one per lane) and keeps their h1,
// dL/dh2 and, in turn, h2 and dL/dh1 in shared memory; lane j then owns row j of W2 and W1
// and entry j of b1, b2, w3 and sums the per-sample outer products over the 64 samples.
// Each block (one wave) writes its gradient sum as one partial.
__global__ void __launch_bounds__(64)
k_mlp_train(const float* __restrict__ w, const float* __restrict__ feat,
            const float* __restrict__ target, const float* __restrict__ weight, int64_t M,
            float* __restrict__ partial, float* __restrict__ loss) {
  __shared__ float sh1[64][kHid + 1];
  __shared__ float sdh2[64][kHid + 1];
  __shared__ float sx[64][kHid + 1];  // h2, then dL/dh1
  __shared__ float sf[64][kIn];
  __shared__ float sdo[64];
  const int lane = threadIdx.x;
  float gW2[kHid], gW1[kIn];
#pragma unroll
  for (int i = 0; i < kHid; ++i) gW2[i] = 0.f;
#pragma unroll
  for (int i = 0; i < kIn; ++i) gW1[i] = 0.f;
  float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, lsum = 0.f;
  for (int64_t base = blockIdx.x * (int64_t)64; base < M; base += (int64_t)gridDim.x * 64) {
    const int64_t m = base + lane;
    const bool ok = m < M;
    float f[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) f[i] = ok ? feat[m * kIn + i] : 0.f;
    float o = w[kOffB3];
    {
      float h1[kHid];
      mlp_h1(w, f, h1);
#pragma unroll
      for (int j = 0; j < kHid; ++j) {
        sh1[lane][j] = h1[j];
        const float h2 = mlp_h2(w, h1, j);
        sx[lane][j] = h2;
        o = __builtin_fmaf(w[kOffW3 + j], h2, o);
      }
    }
    const float s = softplus(o);
    const float t = ok ? target[m] : 0.f;
    const float wt = ok ? weight[m] : 0.f;
    lsum += wt * fabsf(s - t);
    const float g = (s > t ? wt : (s < t ? -wt : 0.f)) * sigmoid(o);  // dL/do
#pragma unroll
    for (int j = 0; j < kHid; ++j) sdh2[lane][j] = sx[lane][j] > 0.f ? g * w[kOffW3 + j] : 0.f;
#pragma unroll
    for (int i = 0; i < kIn; ++i) sf[lane][i] = f[i];
    sdo[lane] = g;
    __syncthreads();
    // phase 1: lane j sums dh2_q[j] * h1_q and g_q * h2_q[j] over the 64 samples
    for (int q = 0; q < 64; ++q) {
      const float d2 = sdh2[q][lane];
#pragma unroll
      for (int i = 0; i < kHid; ++i) gW2[i] = __builtin_fmaf(d2, sh1[q][i], gW2[i]);
      gb2 += d2;
      gw3 = __builtin_fmaf(sdo[q], sx[q][lane], gw3);
    }
    if (lane == 0)
      for (int q = 0; q < 64; ++q) gb3 += sdo[q];
    __syncthreads();
    // dh1 = W2^T dh2 (masked by ReLU) for this lane's sample, into sx
#pragma unroll
    for (int i = 0; i < kHid; ++i) {
      float acc = 0.f;
#pragma unroll
      for (int j = 0; j < kHid; ++j) acc = __builtin_fmaf(w[kOffW2 + j * kHid + i], sdh2[lane][j], acc);
      sx[lane][i] = sh1[lane][i] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    // phase 2: lane j sums dh1_q[j] * f_q and dh1_q[j]
    for (int q = 0; q < 64; ++q) {
      const float d1 = sx[q][lane];
#pragma unroll
      for (int i = 0; i < kIn; ++i) gW1[i] = __builtin_fmaf(d1, sf[q][i], gW1[i]);
      gb1 += d1;
    }
    __syncthreads();
  }
  float* out = partial + (int64_t)blockIdx.x * kNumParams;
#pragma unroll
  for (int i = 0; i < kHid; ++i) out[kOffW2 + lane * kHid + i] = gW2[i];
#pragma unroll
  for (int i = 0; i < kIn; ++i) out[kOffW1 + lane * kIn + i] = gW1[i];
  out[kOffB1 + lane] = gb1;
  out[kOffB2 + lane] = gb2;
  out[kOffW3 + lane] = gw3;
  if (lane == 0) out[kOffB3] = gb3;
  float l = lsum;
#pragma unroll
  for (int msk = 1; msk < 64; msk <<= 1) l += __shfl_xor(l, msk, 64);
  if (lane == 0) atomicAdd(loss, l);
}

__global__ void k_grad_sum(const float* __restrict__ partial, int nblk, float* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNumParams) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * kNumParams + i];
  g[i] = s;
}

// Sum the per-block gradients, scale, and take one Adam step.
__global__ void k_adam(float* __restrict__ params, const float* __restrict__ partial, int nblk,
                       float scale, float* __restrict__ m1, float* __restrict__ m2, float lr,
                       float b1, float b2, float eps, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNumParams) return;
  float g = 0.f;
  for (int b = 0; b < nblk; ++b) g += partial[(int64_t)b * kNumParams + i];
  g *= scale;
  const float a = b1 * m1[i] + (1.f - b1) * g;
  const float v = b2 * m2[i] + (1.f - b2) * g * g;
  m1[i] = a;
  m2[i] = v;
  const float ah = a / (1.f - powf(b1, (float)step));
  const float vh = v / (1.f - powf(b2, (float)step));
  params[i] -= lr * ah / (sqrtf(vh) + eps);
}

// ---- degrees from the PBR matrix (row i = in-edges of i, column j = out-edges of j) ----
__global__ void k_tile_degrees(int64_t T, int32_t B, const uint64_t* __restrict__ codes,
                               const int32_t* __restrict__ coords, int64_t n,
                               int32_t* __restrict__ indeg, int32_t* __restrict__ outdeg) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < T;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t c = codes[t];
    const int64_t r0 = coords[2 * t], c0 = coords[2 * t + 1];
    for (int r = 0; r < B; ++r) {
      const int cnt = __popcll((c >> (r * B)) & ((1ull << B) - 1ull));
      if (cnt && r0 + r < n) atomicAdd(&indeg[r0 + r], cnt);
    }
    for (int col = 0; col < B; ++col) {
      int cnt = 0;
      for (int r = 0; r < B; ++r) cnt += (c >> (r * B + col)) & 1ull;
      if (cnt && c0 + col < n) atomicAdd(&outdeg[c0 + col], cnt);
    }
  }
}
__global__ void k_rem_degrees(int64_t n, const int32_t* __restrict__ rptr,
                              const int32_t* __restrict__ rcol, int32_t* __restrict__ indeg,
                              int32_t* __restrict__ outdeg) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int a = rptr[r], b = rptr[r + 1];
    if (b > a) atomicAdd(&indeg[r], b - a);
    for (int j = a; j < b; ++j) atomicAdd(&outdeg[rcol[j]], 1);
  }
}

// ---- features ----
struct WsView {
  int64_t N, K;
  float alpha;
  const int32_t* indeg;
  const int32_t* outdeg;
  const float* v[5];  // v_j = A^j p, [N][K]
};

__device__ __forceinline__ void pair_features(const WsView& s, int64_t i, int64_t k,
                                              float (&f)[kIn], int& hop) {
  const int64_t ix = i * s.K + k;
  float vv[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) vv[j] = s.v[j][ix];
  hop = 5;
#pragma unroll
  for (int j = 4; j >= 0; --j)
    if (vv[j] > 0.f) hop = j;
  f[0] = deg_bucket(s.indeg[i]) * 0.125f;
  f[1] = deg_bucket(s.outdeg[i]) * 0.125f;
#pragma unroll
  for (int h = 0; h < 6; ++h) f[2 + h] = hop == h ? 1.f : 0.f;
  const float a = s.alpha, c = 1.f - s.alpha, n = (float)s.N;
  const float q1 = c * vv[0];
  const float q2 = q1 + c * a * vv[1];
  const float q3 = q2 + c * a * a * vv[2];
  f[8] = log1pf(n * q1) * 0.125f;
  f[9] = log1pf(n * q2) * 0.125f;
  f[10] = log1pf(n * q3) * 0.125f;
  f[11] = 1.f;
}

// features of listed pairs (i = idx / K, k = idx % K)
__global__ void k_pair_features(WsView s, const int64_t* __restrict__ pairs, int64_t M,
                                float* __restrict__ feat) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < M;
       m += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = pairs[m];
    float f[kIn];
    int hop;
    pair_features(s, p / s.K, p % s.K, f, hop);
#pragma unroll
    for (int i = 0; i < kIn; ++i) feat[m * kIn + i] = f[i];
  }
}

// table of (in-bucket, out-bucket, class) features, class 0/1/2 = hop 3 / 4 / more
__global__ void k_table_features(float* __restrict__ feat) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= kBuckets * kBuckets * 3) return;
  const int cls = m % 3, bo = (m / 3) % kBuckets, bi = m / 3 / kBuckets;
  float* f = feat + (int64_t)m * kIn;
  f[0] = bi * 0.125f;
  f[1] = bo * 0.125f;
  for (int h = 0; h < 6; ++h) f[2 + h] = (h == 3 + cls) ? 1.f : 0.f;
  f[8] = f[9] = f[10] = 0.f;
  f[11] = 1.f;
}

// Table row of every node: in-bucket * kBuckets + out-bucket (once per graph).
__global__ void k_node_buckets(int64_t n, const int32_t* __restrict__ indeg,
                               const int32_t* __restrict__ outdeg, int32_t* __restrict__ nb) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    nb[i] = deg_bucket(indeg[i]) * kBuckets + deg_bucket(outdeg[i]);
}

// X0 for pairs beyond hop 2 from the table; pairs within hop 2 appended to `near`. One wave
// per row (lanes over the K columns, coalesced in the [N][K] layout). Each wave collects its
// near pairs in an LDS buffer and reserves their slots with one counter update per flush:
// one update per row instead would serialize millions of atomics on the one counter.
constexpr int kNearBuf = 512;  // pairs per wave buffer (>= 64: a row chunk always fits)
// The 4th power-iteration iterate from the seed start, from the feature products
// v_j = A^j P (j = 0..4): x4 = (1 - alpha) (v0 + alpha v1 + alpha^2 v2 + alpha^3 v3)
// + alpha^4 v4.
__device__ __forceinline__ float iterate4(float v0, float v1, float v2, float v3, float v4,
                                          float alpha) {
  const float a2 = alpha * alpha, a3 = a2 * alpha, a4 = a3 * alpha;
  return (1.f - alpha) * (v0 + alpha * v1 + a2 * v2 + a3 * v3) + a4 * v4;
}
// x4 (may be null): also write the 4th iterate, from the products this kernel reads anyway.
__global__ void __launch_bounds__(256) k_assemble(WsView s, const int32_t* __restrict__ nb,
                                                  const float* __restrict__ table,
                                                  float* __restrict__ x0,
                                                  int64_t* __restrict__ near, int64_t cap,
                                                  unsigned long long* __restrict__ count,
                                                  float* __restrict__ x4) {
  __shared__ int64_t buf_all[4][kNearBuf];
  const int lane = threadIdx.x & 63;
  int64_t* __restrict__ buf = buf_all[threadIdx.x >> 6];
  const int64_t w0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int nbuf = 0;  // wave-uniform
  auto flush = [&]() {
    wave_lds_sync();
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(count, (unsigned long long)nbuf);
    base = __shfl(base, 0, 64);
    for (int j = lane; j < nbuf; j += 64)
      if ((int64_t)(base + j) < cap) near[base + j] = buf[j];
    wave_lds_sync();
    nbuf = 0;
  };
  for (int64_t i = w0; i < s.N; i += nw) {
    const int64_t tb = (int64_t)nb[i] * 3;
    for (int64_t k0 = 0; k0 < s.K; k0 += 64) {
      const int64_t k = k0 + lane;
      const bool ok = k < s.K;
      const int64_t ix = i * s.K + k;
      int hop = 5;
      float vj[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
      if (ok) {
#pragma unroll
        for (int j = 4; j >= 0; --j) {
          vj[j] = s.v[j][ix];
          if (vj[j] > 0.f) hop = j;
        }
        if (x4 != nullptr) x4[ix] = iterate4(vj[0], vj[1], vj[2], vj[3], vj[4], s.alpha);
      }
      const bool is_near = ok && hop <= 2;
      const uint64_t m = __ballot(is_near);
      const int c = __popcll(m);
      if (nbuf + c > kNearBuf) flush();
      if (is_near) buf[nbuf + __popcll(m & below)] = ix;
      nbuf += c;
      if (ok) x0[ix] = is_near ? 0.f : table[tb + (hop - 3)];
    }
  }
  if (nbuf > 0) flush();
}

__global__ void k_scatter_pairs(const int64_t* __restrict__ pairs, const float* __restrict__ val,
                                int64_t M, float* __restrict__ x0) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m < M) x0[pairs[m]] = val[m];
}

// Column sums of x [N][K] in a fixed order, so that repeated runs agree bit for bit. A wave
// walks rows (wave w of the grid takes rows w, w + nw, ...), lane k keeps column k (+64j) in
// a register; the 4 waves of a workgroup are added in wave order into part[block][K], and
// k_colsum_final adds the workgroups' partials in a fixed order.
constexpr int kColsumBlocks = 1024;
__global__ void __launch_bounds__(256) k_colsum(const float* __restrict__ x, int64_t N, int64_t K,
                                                float* __restrict__ part) {
  __shared__ float acc[4][1024];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int64_t w0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t k0 = 0; k0 < K; k0 += 64) {
    const int64_t k = k0 + lane;
    float a = 0.f;
    if (k < K)
      for (int64_t i = w0; i < N; i += nw) a += x[i * K + k];
    if (k < K) acc[wid][k] = a;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += blockDim.x)
    part[blockIdx.x * K + k] = ((acc[0][k] + acc[1][k]) + acc[2][k]) + acc[3][k];
}
// sums[k] = the nb partials of column k: wave w of the 16 takes partials w, w + 16, ... into
// 8 accumulators (partial b into accumulator (b / 16) % 8), then the waves are added in order.
__global__ void __launch_bounds__(1024) k_colsum_final(const float* __restrict__ part, int nb,
                                                       int64_t K, float* __restrict__ sums) {
  __shared__ float red[16][64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t k = blockIdx.x * 64LL + lane;
  float a[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) a[j] = 0.f;
  if (k < K) {
    int b = w;
    for (; b + 16 * 7 < nb; b += 16 * 8) {
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] += part[(int64_t)(b + 16 * j) * K + k];
    }
#pragma unroll
    for (int j = 0; j < 7; ++j)
      if (b + 16 * j < nb) a[j] += part[(int64_t)(b + 16 * j) * K + k];
  }
  red[w][lane] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  __syncthreads();
  if (w == 0 && k < K) {
    float t = 0.f;
    for (int i = 0; i < 16; ++i) t += red[i][lane];
    sums[k] = t;
  }
}
// x0 <- max(0, x4 + lam (x0 / msum - x4)) for the network's raw start x0 with column sums
// msum and the 4th iterate x4, and the workgroups' partial column sums of the result (the
// order of k_colsum, so k_colsum_final can add them): the network's start normalized, the
// blend and the next column sums in one pass.
__global__ void __launch_bounds__(256) k_blend_sums(float* __restrict__ x0,
                                                    const float* __restrict__ x4, int64_t N,
                                                    int64_t K, const float* __restrict__ msum,
                                                    float lam, float* __restrict__ part) {
  __shared__ float acc[4][1024];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int64_t w0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t k0 = 0; k0 < K; k0 += 64) {
    const int64_t k = k0 + lane;
    float a = 0.f;
    if (k < K) {
      const float sk = msum[k];
      const float inv = sk > 0.f ? 1.f / sk : 0.f;
      for (int64_t i = w0; i < N; i += nw) {
        const int64_t e = i * K + k;
        const float xi = x4[e];
        const float y = fmaxf(0.f, xi + lam * (x0[e] * inv - xi));
        x0[e] = y;
        a += y;
      }
      acc[wid][k] = a;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += blockDim.x)
    part[blockIdx.x * K + k] = ((acc[0][k] + acc[1][k]) + acc[2][k]) + acc[3][k];
}

__global__ void k_colscale(float* __restrict__ x, int64_t N, int64_t K, const float* __restrict__ sums) {
  const int lane = threadIdx.x & 63;
  const int64_t w0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t k0 = 0; k0 < K; k0 += 64) {
    const int64_t k = k0 + lane;
    if (k >= K) continue;
    const float sk = sums[k];
    const float inv = sk > 0.f ? 1.f / sk : 0.f;
    for (int64_t i = w0; i < N; i += nw) x[i * K + k] *= inv;
  }
}

static int nblocks(int64_t n, int t = 256, int cap = 65536) {
  int64_t b = (n + t - 1) / t;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

}  // namespace gl

using namespace gl;

extern "C" {

int32_t gl_mlp_num_params(void) { return kNumParams; }
int32_t gl_mlp_num_features(void) { return kIn; }
int32_t gl_ws_num_buckets(void) { return kBuckets; }

int gl_mlp_forward(const float* params, const float* feat, int64_t M, float* out, void* stream) {
  GL_REQUIRE(params && (M == 0 || (feat && out)), "null pointer");
  if (M == 0) return GL_OK;
  hipLaunchKernelGGL(k_mlp_forward, dim3(nblocks(M, 256, 4096)), dim3(256), 0,
                     (hipStream_t)stream, params, feat, M, out);
  GL_LAUNCH_CHECK("k_mlp_forward");
  return GL_OK;
}

int64_t gl_mlp_train_partials(void) { return (int64_t)1024 * kNumParams; }

int gl_mlp_train_step(float* params, const float* feat, const float* target,
                      const float* weight, int64_t M, float grad_scale, float lr, float* adam_m,
                      float* adam_v, int32_t step, float* partial, float* loss, void* stream) {
  GL_REQUIRE(params && feat && target && weight && adam_m && adam_v && partial && loss,
             "null pointer");
  GL_REQUIRE(M > 0 && step >= 1, "bad sizes");
  const int nblk = (int)std::min<int64_t>(1024, (M + 63) / 64);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_train, dim3(nblk), dim3(64), 0, st, params, feat, target, weight, M,
                     partial, loss);
  GL_LAUNCH_CHECK("k_mlp_train");
  hipLaunchKernelGGL(k_adam, dim3((kNumParams + 255) / 256), dim3(256), 0, st, params, partial,
                     nblk, grad_scale, adam_m, adam_v, lr, 0.9f, 0.999f, 1e-8f, step);
  GL_LAUNCH_CHECK("k_adam");
  return GL_OK;
}

// Gradient only (no update): sums into grad[kNumParams] (for checks).
int gl_mlp_grad(const float* params, const float* feat, const float* target, const float* weight,
                int64_t M, float* partial, float* grad, float* loss, void* stream) {
  GL_REQUIRE(params && feat && target && weight && partial && grad && loss, "null pointer");
  GL_REQUIRE(M > 0, "bad sizes");
  const int nblk = (int)std::min<int64_t>(1024, (M + 63) / 64);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_train, dim3(nblk), dim3(64), 0, st, params, feat, target, weight, M,
                     partial, loss);
  GL_LAUNCH_CHECK("k_mlp_train");
  hipLaunchKernelGGL(k_grad_sum, dim3((kNumParams + 255) / 256), dim3(256), 0, st, partial, nblk,
                     grad);
  GL_LAUNCH_CHECK("k_grad_sum");
  return GL_OK;
}

int gl_ws_degrees(int64_t n, int32_t block_rows, int64_t num_tiles, const uint64_t* codes,
                  const int32_t* coords, const int32_t* rem_indptr, const int32_t* rem_col_ind,
                  int32_t* indeg, int32_t* outdeg, void* stream) {
  GL_REQUIRE(n >= 0 && num_tiles >= 0 && indeg && outdeg && rem_indptr, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  GL_HIP(hipMemsetAsync(indeg, 0, n * 4, st));
  GL_HIP(hipMemsetAsync(outdeg, 0, n * 4, st));
  if (num_tiles > 0) {
    hipLaunchKernelGGL(k_tile_degrees, dim3(nblocks(num_tiles)), dim3(256), 0, st, num_tiles,
                       block_rows, codes, coords, n, indeg, outdeg);
    GL_LAUNCH_CHECK("k_tile_degrees");
  }
  hipLaunchKernelGGL(k_rem_degrees, dim3(nblocks(n)), dim3(256), 0, st, n, rem_indptr,
                     rem_col_ind, indeg, outdeg);
  GL_LAUNCH_CHECK("k_rem_degrees");
  return GL_OK;
}

static WsView make_view(int64_t N, int64_t K, float alpha, const int32_t* indeg,
                        const int32_t* outdeg, const float* const* v) {
  WsView s;
  s.N = N;
  s.K = K;
  s.alpha = alpha;
  s.indeg = indeg;
  s.outdeg = outdeg;
  for (int j = 0; j < 5; ++j) s.v[j] = v[j];
  return s;
}

int gl_ws_pair_features(int64_t N, int64_t K, float alpha, const int32_t* indeg,
                        const int32_t* outdeg, const float* const* v5, const int64_t* pairs,
                        int64_t M, float* feat, void* stream) {
  GL_REQUIRE(v5 && indeg && outdeg && (M == 0 || (pairs && feat)), "null pointer");
  if (M == 0) return GL_OK;
  WsView s = make_view(N, K, alpha, indeg, outdeg, v5);
  hipLaunchKernelGGL(k_pair_features, dim3(nblocks(M)), dim3(256), 0, (hipStream_t)stream, s,
                     pairs, M, feat);
  GL_LAUNCH_CHECK("k_pair_features");
  return GL_OK;
}

int64_t gl_ws_table_size(void) { return (int64_t)kBuckets * kBuckets * 3; }

int gl_ws_table_features(float* feat, void* stream) {
  GL_REQUIRE(feat != nullptr, "null pointer");
  const int n = kBuckets * kBuckets * 3;
  hipLaunchKernelGGL(k_table_features, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     feat);
  GL_LAUNCH_CHECK("k_table_features");
  return GL_OK;
}

int gl_ws_node_buckets(int64_t n, const int32_t* indeg, const int32_t* outdeg, int32_t* nb,
                       void* stream) {
  GL_REQUIRE(n >= 0 && (n == 0 || (indeg && outdeg && nb)), "bad arguments");
  if (n == 0) return GL_OK;
  hipLaunchKernelGGL(k_node_buckets, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, n,
                     indeg, outdeg, nb);
  GL_LAUNCH_CHECK("k_node_buckets");
  return GL_OK;
}

int gl_ws_assemble(int64_t N, int64_t K, float alpha, const int32_t* indeg,
                   const int32_t* outdeg, const float* const* v5, const int32_t* node_buckets,
                   const float* table, float* x0, int64_t* near, int64_t near_cap,
                   uint64_t* near_count, float* x4, void* stream) {
  GL_REQUIRE(v5 && node_buckets && table && x0 && near && near_count, "null pointer");
  WsView s = make_view(N, K, alpha, indeg, outdeg, v5);
  hipStream_t st = (hipStream_t)stream;
  GL_HIP(hipMemsetAsync(near_count, 0, 8, st));
  hipLaunchKernelGGL(k_assemble, dim3(nblocks(N * 64, 256, 16384)), dim3(256), 0, st, s,
                     node_buckets, table, x0, near, near_cap, (unsigned long long*)near_count,
                     x4);
  GL_LAUNCH_CHECK("k_assemble");
  return GL_OK;
}

int gl_ws_scatter(const int64_t* pairs, const float* val, int64_t M, float* x0, void* stream) {
  if (M == 0) return GL_OK;
  GL_REQUIRE(pairs && val && x0, "null pointer");
  hipLaunchKernelGGL(k_scatter_pairs, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, pairs, val, M, x0);
  GL_LAUNCH_CHECK("k_scatter_pairs");
  return GL_OK;
}

int gl_ws_blend_start(float* x0, const float* x4, int64_t N, int64_t K, float lam,
                      float* sums, void* stream) {
  GL_REQUIRE(x0 && x4 && sums && N >= 1 && K >= 1 && K <= 1024, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  // sums[0, K): the network start's column sums; sums[K, 2K): the result's;
  // sums[2K, 2K + nb * K): the workgroups' partials of either
  const int nb = nblocks(N * 64, 256, kColsumBlocks);
  float* part = sums + 2 * K;
  hipLaunchKernelGGL(k_colsum, dim3(nb), dim3(256), 0, st, x0, N, K, part);
  GL_LAUNCH_CHECK("k_colsum");
  hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)((K + 63) / 64)), dim3(1024), 0, st, part,
                     nb, K, sums);
  GL_LAUNCH_CHECK("k_colsum_final");
  hipLaunchKernelGGL(k_blend_sums, dim3(nb), dim3(256), 0, st, x0, x4, N, K, sums, lam, part);
  GL_LAUNCH_CHECK("k_blend_sums");
  hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)((K + 63) / 64)), dim3(1024), 0, st, part,
                     nb, K, sums + K);
  GL_LAUNCH_CHECK("k_colsum_final");
  hipLaunchKernelGGL(k_colscale, dim3(nblocks(N * 64, 256, 16384)), dim3(256), 0, st, x0, N, K,
                     sums + K);
  GL_LAUNCH_CHECK("k_colscale");
  return GL_OK;
}

int gl_ws_normalize(float* x, int64_t N, int64_t K, float* sums, void* stream) {
  GL_REQUIRE(x && sums && K >= 1 && K <= 1024, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  // sums[0, K): the column sums; sums[K, K + nb * K): the workgroups' partials
  const int nb = nblocks(N * 64, 256, kColsumBlocks);
  float* part = sums + K;
  hipLaunchKernelGGL(k_colsum, dim3(nb), dim3(256), 0, st, x, N, K, part);
  GL_LAUNCH_CHECK("k_colsum");
  hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)((K + 63) / 64)), dim3(1024), 0, st, part,
                     nb, K, sums);
  GL_LAUNCH_CHECK("k_colsum_final");
  hipLaunchKernelGGL(k_colscale, dim3(nblocks(N * 64, 256, 16384)), dim3(256), 0, st, x, N, K,
                     sums);
  GL_LAUNCH_CHECK("k_colscale");
  return GL_OK;
}

}  // extern "C"
