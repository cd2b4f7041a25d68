// Stripe-owner SpMM kernels for PBR matrices on gfx950 (wave64).
//
// Work unit: one stripe = one block row (BR consecutive matrix rows). One wave owns a stripe
// end to end: it multiplies every tile of the stripe, then the stripe's CSR-remainder rows,
// then runs an epilogue on the finished rows (plain store for SpMM, the fused power-iteration
// step for PPR). Owner-computes means no atomics on Y and deterministic results; the
// reference instead launches one block per tile and atomically adds into a zeroed Y
// (Backend/graph_link/kernels/pbr_kernel.cu:144-153, :326-328, :456).
//
// Two lane layouts, chosen by KW (RHS columns handled by one launch):
//   bit-lane (KW <= 8): lane l <-> bit p = l % (B*B) of a tile (one 8x8 tile per wave, four
//       4x4 or sixteen 2x2 tiles). Each lane fetches its value with a popcount prefix of the
//       bits below it, multiplies it with the KW-wide X row of its tile column and keeps the
//       product per lane; lanes of one matrix row are summed once per stripe.
//   feature-lane (KW in 16,32,64): lane f <-> RHS column k0+f; 64/KW tiles per wave. The
//       tile's B*B values are spread one per lane over each 16-lane DPP row and broadcast to
//       the FMA with row_newbcast, so each X element is loaded exactly once per tile and the
//       packed values never leave registers.
// Templates: B (block edge 2/4/8), KW, TILES (multiply tiles), REM (multiply the CSR rows),
// Epi (epilogue functor).
#pragma once

#include "gl_common.hpp"

#ifndef GL_FUSED_DPP_FMA
#define GL_FUSED_DPP_FMA 1
#endif

namespace gl {

struct PbrView {
  int64_t rows, cols;
  int64_t num_stripes;
  const uint64_t* __restrict__ codes;
  const int32_t* __restrict__ coords;
  const int32_t* __restrict__ offsets;
  const float* __restrict__ data;
  const int32_t* __restrict__ stripe_ptr;
  const int32_t* __restrict__ rem_ptr;
  const int32_t* __restrict__ rem_col;
  const float* __restrict__ rem_val;
};

struct RhsView {
  const float* __restrict__ x;
  int64_t ldx;
  int64_t K;   // total RHS columns
  int64_t k0;  // first column of this launch
};

// --------------------------------------------------------------------------------------
// Tile FMA block of the feature-lane layout: acc[r] += sum_c V[r][c] * x[c], where V[r][c]
// is held by lane (r*B+c) % 16 of vector v[(r*B+c)/16] in every 16-lane DPP row.
// --------------------------------------------------------------------------------------
#define GL_F(a, v, x, n) \
  "v_fmac_f32_dpp %" #a ", %" #v ", %" #x " row_newbcast:" #n " row_mask:0xf bank_mask:0xf\n"
#define GL_ROW8_LO(r, v) \
  GL_F(r, v, 12, 0) GL_F(r, v, 13, 1) GL_F(r, v, 14, 2) GL_F(r, v, 15, 3) \
  GL_F(r, v, 16, 4) GL_F(r, v, 17, 5) GL_F(r, v, 18, 6) GL_F(r, v, 19, 7)
#define GL_ROW8_HI(r, v) \
  GL_F(r, v, 12, 8) GL_F(r, v, 13, 9) GL_F(r, v, 14, 10) GL_F(r, v, 15, 11) \
  GL_F(r, v, 16, 12) GL_F(r, v, 17, 13) GL_F(r, v, 18, 14) GL_F(r, v, 19, 15)

template <int B>
struct TileFma;

template <>
struct TileFma<8> {
  static __device__ __forceinline__ void run(float (&a)[8], const float (&v)[4],
                                             const float (&x)[8]) {
#if GL_FUSED_DPP_FMA
    // s_nop 1: two wait states between the VALU write of v[] and its first DPP read.
    asm("s_nop 1\n" GL_ROW8_LO(0, 8) GL_ROW8_HI(1, 8) GL_ROW8_LO(2, 9) GL_ROW8_HI(3, 9)
            GL_ROW8_LO(4, 10) GL_ROW8_HI(5, 10) GL_ROW8_LO(6, 11) GL_ROW8_HI(7, 11)
        : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]),
          "+v"(a[6]), "+v"(a[7])
        : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(x[0]), "v"(x[1]), "v"(x[2]),
          "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]));
#else
#define GL_BC(r, c) a[r] = __builtin_fmaf(row_bcast<((r) * 8 + (c)) % 16>(v[((r) * 8 + (c)) / 16]), x[c], a[r])
#define GL_BROW(r) GL_BC(r, 0); GL_BC(r, 1); GL_BC(r, 2); GL_BC(r, 3); GL_BC(r, 4); GL_BC(r, 5); GL_BC(r, 6); GL_BC(r, 7)
    GL_BROW(0); GL_BROW(1); GL_BROW(2); GL_BROW(3); GL_BROW(4); GL_BROW(5); GL_BROW(6); GL_BROW(7);
#undef GL_BROW
#undef GL_BC
#endif
  }
};

template <>
struct TileFma<4> {
  static __device__ __forceinline__ void run(float (&a)[4], const float (&v)[1],
                                             const float (&x)[4]) {
#if GL_FUSED_DPP_FMA
#define GL_F4(a_, n, xx) GL_F(a_, 4, xx, n)
    asm("s_nop 1\n"
        GL_F4(0, 0, 5) GL_F4(0, 1, 6) GL_F4(0, 2, 7) GL_F4(0, 3, 8)
        GL_F4(1, 4, 5) GL_F4(1, 5, 6) GL_F4(1, 6, 7) GL_F4(1, 7, 8)
        GL_F4(2, 8, 5) GL_F4(2, 9, 6) GL_F4(2, 10, 7) GL_F4(2, 11, 8)
        GL_F4(3, 12, 5) GL_F4(3, 13, 6) GL_F4(3, 14, 7) GL_F4(3, 15, 8)
        : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3])
        : "v"(v[0]), "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]));
#undef GL_F4
#else
#define GL_BC(r, c) a[r] = __builtin_fmaf(row_bcast<(r) * 4 + (c)>(v[0]), x[c], a[r])
    GL_BC(0, 0); GL_BC(0, 1); GL_BC(0, 2); GL_BC(0, 3);
    GL_BC(1, 0); GL_BC(1, 1); GL_BC(1, 2); GL_BC(1, 3);
    GL_BC(2, 0); GL_BC(2, 1); GL_BC(2, 2); GL_BC(2, 3);
    GL_BC(3, 0); GL_BC(3, 1); GL_BC(3, 2); GL_BC(3, 3);
#undef GL_BC
#endif
  }
};

template <>
struct TileFma<2> {
  static __device__ __forceinline__ void run(float (&a)[2], const float (&v)[1],
                                             const float (&x)[2]) {
#if GL_FUSED_DPP_FMA
    asm("s_nop 1\n" GL_F(0, 2, 3, 0) GL_F(0, 2, 4, 1) GL_F(1, 2, 3, 2) GL_F(1, 2, 4, 3)
        : "+v"(a[0]), "+v"(a[1])
        : "v"(v[0]), "v"(x[0]), "v"(x[1]));
#else
    a[0] = __builtin_fmaf(row_bcast<0>(v[0]), x[0], a[0]);
    a[0] = __builtin_fmaf(row_bcast<1>(v[0]), x[1], a[0]);
    a[1] = __builtin_fmaf(row_bcast<2>(v[0]), x[0], a[1]);
    a[1] = __builtin_fmaf(row_bcast<3>(v[0]), x[1], a[1]);
#endif
  }
};
#undef GL_ROW8_LO
#undef GL_ROW8_HI
#undef GL_F

// --------------------------------------------------------------------------------------
// Epilogues. Each lane of the epilogue owns one (row, column) at a time; put() is called
// once per output element. begin_stripe() runs (wave-uniform) before the stripe's puts and
// finish() once per lane at the end with the lane's column slot (or -1).
// --------------------------------------------------------------------------------------
struct EpiStore {
  float* __restrict__ y;
  int64_t ldy;
  int32_t accumulate;
  __device__ __forceinline__ bool active() const { return true; }
  __device__ __forceinline__ void begin_stripe(int64_t) {}
  __device__ __forceinline__ void put(int64_t row, int64_t col, float acc) {
    float* p = y + row * ldy + col;
    *p = accumulate ? (*p + acc) : acc;
  }
  __device__ __forceinline__ void finish(int, float*) {}
};

// Device control block of a PPR run.
struct PprCtl {
  int32_t iter;    // completed iterations
  int32_t active;  // columns still iterating
  int32_t pad[2];
};

// Fused power-iteration step (north-star subsystem d):
//   X_{t+1}[i,k] = alpha * (A X_t)[i,k] + cu[k] + tele[k] * p_k[i]
// with cu[k] = alpha * dangling_k / N (uniform dangling) and tele[k] = (1 - alpha)
// (+ alpha * dangling_k in personalization mode), frozen columns copied unchanged, the L1
// residual |X_{t+1} - X_t| and the next dangling mass accumulated per lane and written as
// per-workgroup partials (summed by the retire kernel: deterministic, no float atomics to
// HBM). The reference's equivalent is ppr_update_normalized_kernel
// (Backend/graph_link/kernels/ppr_kernel.cu:130-168), run as a separate pass over Y.
struct EpiPpr {
  const float* __restrict__ xcur;
  float* __restrict__ xnext;
  int64_t ld;
  float alpha;
  const float* __restrict__ cu;
  const float* __restrict__ tele;
  const int32_t* __restrict__ frozen;
  const uint8_t* __restrict__ dmask;
  const int32_t* __restrict__ sptr;
  const int32_t* __restrict__ srow;
  const int32_t* __restrict__ scol;
  const float* __restrict__ sw;
  const PprCtl* __restrict__ ctl;
  float* __restrict__ part_res;   // [gridDim.x][K]
  float* __restrict__ part_dang;  // [gridDim.x][K]
  int64_t K;
  int64_t k0;
  // per-lane state
  float res;
  float dang;
  int32_t e0, e1;

  __device__ __forceinline__ bool active() const { return ctl->active > 0; }
  __device__ __forceinline__ void begin_stripe(int64_t s) {
    e0 = sptr[s];
    e1 = sptr[s + 1];
  }
  __device__ __forceinline__ void put(int64_t row, int64_t col, float acc) {
    const int64_t ix = row * ld + col;
    const float xc = xcur[ix];
    float xn = xc;
    if (!frozen[col]) {
      float w = 0.f;
      for (int e = e0; e < e1; ++e)
        if (srow[e] == row && scol[e] == col) w += sw[e];
      xn = __builtin_fmaf(alpha, acc, cu[col]);
      xn = __builtin_fmaf(w, tele[col], xn);
      res += fabsf(xn - xc);
      if (dmask != nullptr && dmask[row]) dang += xn;
    }
    xnext[ix] = xn;
  }
  // slot: the lane's column minus k0 (or -1); red: 2*64 floats of shared memory.
  __device__ __forceinline__ void finish(int slot, float* red) {
    if (threadIdx.x < 128) red[threadIdx.x] = 0.f;
    __syncthreads();
    if (slot >= 0) {
      atomicAdd(&red[slot], res);
      atomicAdd(&red[64 + slot], dang);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const int64_t col = k0 + threadIdx.x;
      if (col < K) {
        part_res[blockIdx.x * K + col] = red[threadIdx.x];
        part_dang[blockIdx.x * K + col] = red[64 + threadIdx.x];
      }
    }
  }
};

// --------------------------------------------------------------------------------------
// Feature-lane kernel, KW in {16, 32, 64}.
// --------------------------------------------------------------------------------------
template <int B, int KW, bool TILES, bool REM, class Epi>
__global__ void __launch_bounds__(256) k_stripe_feat(PbrView m, RhsView d, Epi epi) {
  constexpr int G = kWave / KW;           // tiles in flight per wave
  constexpr int P = B * B;                 // bits per tile
  constexpr int NV = (P + 15) / 16;        // value registers per lane
  constexpr int L = kWave / B;             // lanes per row in the remainder phase
  constexpr int FPL = (KW / L) > 0 ? (KW / L) : 1;  // RHS columns per lane there
  __shared__ float red[128];
  __shared__ float trans[4][B][KW];        // remainder row-layout -> feature-layout

  if (!epi.active()) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int g = lane / KW;
  const int f = lane % KW;
  const int o = lane & 15;
  const int64_t colk = d.k0 + f;
  const bool fok = colk < d.K;
  const float* __restrict__ X = d.x + d.k0;
  // remainder-phase layout: lane = (row rr, slot q), columns q*FPL .. q*FPL+FPL-1
  const int rr = lane / L;
  const int q = lane % L;
  const bool qok = q * FPL < KW;

  const Sched sc = xcd_sched(m.num_stripes);
  for (int64_t s = sc.first; s < sc.end; s += sc.stride) {
    const int64_t r0 = s * B;
    float acc[B];
#pragma unroll
    for (int r = 0; r < B; ++r) acc[r] = 0.f;

    if constexpr (TILES) {
      const int t0 = m.stripe_ptr[s];
      const int t1 = m.stripe_ptr[s + 1];
      for (int tc = t0; tc < t1; tc += kWave) {
        // stage up to 64 tiles' metadata, one tile per lane
        const int tl = tc + lane;
        uint64_t mcode = 0;
        int mc0 = 0, moff = 0;
        if (tl < t1) {
          mcode = m.codes[tl];
          mc0 = m.coords[2 * (int64_t)tl + 1];
          moff = m.offsets[tl];
        }
        const int tce = min(t1, tc + kWave);

        for (int tb = tc; tb < tce; tb += G) {
          const int src = tb - tc + g;  // lane holding this group's tile metadata
          const bool tv = tb + g < tce;
          const uint32_t clo = __shfl((uint32_t)mcode, src, kWave);
          const uint32_t chi = __shfl((uint32_t)(mcode >> 32), src, kWave);
          const int c0 = __shfl(mc0, src, kWave);
          const int off = __shfl(moff, src, kWave);
          const uint64_t code = tv ? (((uint64_t)chi << 32) | clo) : 0ull;
          float v[NV];
#pragma unroll
          for (int j = 0; j < NV; ++j) {
            const int p = 16 * j + o;
            v[j] = 0.f;
            if (p < P && ((code >> p) & 1ull))
              v[j] = m.data[off + popc64(code & ((1ull << p) - 1ull))];
          }
          float x[B];
#pragma unroll
          for (int c = 0; c < B; ++c) {
            const int64_t col = (int64_t)c0 + c;
            x[c] = (tv && fok && col < m.cols) ? X[col * d.ldx + f] : 0.f;
          }
          TileFma<B>::run(acc, v, x);
        }
      }
    }

    if constexpr (REM) {
      // Row layout: lanes of row rr walk that row's remainder entries in order; all B rows
      // advance together so their loads overlap.
      float racc[FPL];
#pragma unroll
      for (int u = 0; u < FPL; ++u) racc[u] = 0.f;
      const int64_t row = r0 + rr;
      int j = 0, j1 = 0;
      if (row < m.rows) {
        j = m.rem_ptr[row];
        j1 = m.rem_ptr[row + 1];
      }
      const int kq = q * FPL;
      for (; __any(j < j1); j += 1) {
        if (j < j1 && qok) {
          const float a = m.rem_val[j];
          const float* xr = X + (int64_t)m.rem_col[j] * d.ldx + kq;
#pragma unroll
          for (int u = 0; u < FPL; ++u)
            if (d.k0 + kq + u < d.K) racc[u] = __builtin_fmaf(a, xr[u], racc[u]);
        }
      }
      // transpose through LDS: (row rr, columns kq..) -> (column f, all rows)
      if (qok) {
#pragma unroll
        for (int u = 0; u < FPL; ++u)
          if (kq + u < KW) trans[wid][rr][kq + u] = racc[u];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (g == 0) {
#pragma unroll
        for (int r = 0; r < B; ++r) acc[r] += trans[wid][r][f];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // combine the tile groups: every group ends with the full row sums
    if constexpr (G >= 2) {
#pragma unroll
      for (int r = 0; r < B; ++r) acc[r] += shfl_xor(acc[r], 32);
    }
    if constexpr (G >= 4) {
#pragma unroll
      for (int r = 0; r < B; ++r) acc[r] += shfl_xor(acc[r], 16);
    }

    epi.begin_stripe(s);
    if (fok) {
#pragma unroll
      for (int r = 0; r < B; ++r) {
        const int64_t row = r0 + r;
        if ((r % G) == g && row < m.rows) epi.put(row, colk, acc[r]);
      }
    }
  }
  epi.finish(fok ? f : -1, red);
}

// --------------------------------------------------------------------------------------
// Bit-lane kernel, KW in {1, 2, 4, 8}.
// --------------------------------------------------------------------------------------
template <int B, int KW, bool TILES, bool REM, class Epi>
__global__ void __launch_bounds__(256) k_stripe_bit(PbrView m, RhsView d, Epi epi) {
  constexpr int P = B * B;
  constexpr int G = kWave / P;   // tiles per wave
  constexpr int L = kWave / B;   // lanes per matrix row
  __shared__ float red[128];

  if (!epi.active()) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int g = lane / P;
  const int p = lane % P;
  const int rl = p / B;          // row of this lane inside the stripe
  const int cl = p % B;          // column inside the tile
  const int q = g * B + cl;      // slot among the L lanes of row rl
  const uint64_t below = (1ull << p) - 1ull;
  const float* __restrict__ X = d.x + d.k0;
  const int kwv = (int)min<int64_t>(KW, d.K - d.k0);  // valid columns of this launch

  const Sched sc = xcd_sched(m.num_stripes);
  for (int64_t s = sc.first; s < sc.end; s += sc.stride) {
    const int64_t r0 = s * B;
    float acc[KW];
#pragma unroll
    for (int k = 0; k < KW; ++k) acc[k] = 0.f;

    if constexpr (TILES) {
      const int t0 = m.stripe_ptr[s];
      const int t1 = m.stripe_ptr[s + 1];
      for (int tc = t0; tc < t1; tc += kWave) {
        const int tl = tc + lane;
        uint64_t mcode = 0;
        int mc0 = 0, moff = 0;
        if (tl < t1) {
          mcode = m.codes[tl];
          mc0 = m.coords[2 * (int64_t)tl + 1];
          moff = m.offsets[tl];
        }
        const int tce = min(t1, tc + kWave);

        for (int tb = tc; tb < tce; tb += G) {
          const int src = tb - tc + g;
          const bool tv = tb + g < tce;
          const uint32_t clo = __shfl((uint32_t)mcode, src, kWave);
          const uint32_t chi = __shfl((uint32_t)(mcode >> 32), src, kWave);
          const int c0 = __shfl(mc0, src, kWave);
          const int off = __shfl(moff, src, kWave);
          const uint64_t code = tv ? (((uint64_t)chi << 32) | clo) : 0ull;
          if ((code >> p) & 1ull) {
            const int64_t col = (int64_t)c0 + cl;
            const float v = m.data[off + popc64(code & below)];
            if (col < m.cols) {
              const float* xr = X + col * d.ldx;
#pragma unroll
              for (int k = 0; k < KW; ++k)
                if (k < kwv) acc[k] = __builtin_fmaf(v, xr[k], acc[k]);
            }
          }
        }
      }
    }

    if constexpr (REM) {
      const int64_t row = r0 + rl;
      int j = 0, j1 = 0;
      if (row < m.rows) {
        j = m.rem_ptr[row] + q;
        j1 = m.rem_ptr[row + 1];
      }
      for (; __any(j < j1); j += L) {
        if (j < j1) {
          const float a = m.rem_val[j];
          const float* xr = X + (int64_t)m.rem_col[j] * d.ldx;
#pragma unroll
          for (int k = 0; k < KW; ++k)
            if (k < kwv) acc[k] = __builtin_fmaf(a, xr[k], acc[k]);
        }
      }
    }

    // sum the L lanes of each row: bits of cl (1 .. B/2) and of g (P .. 32)
#pragma unroll
    for (int msk = 1; msk < B; msk <<= 1) {
#pragma unroll
      for (int k = 0; k < KW; ++k) acc[k] += shfl_xor(acc[k], msk);
    }
#pragma unroll
    for (int msk = P; msk < kWave; msk <<= 1) {
#pragma unroll
      for (int k = 0; k < KW; ++k) acc[k] += shfl_xor(acc[k], msk);
    }

    epi.begin_stripe(s);
    const int64_t row = r0 + rl;
    if (q < kwv && row < m.rows) {
      float mine = acc[0];
#pragma unroll
      for (int k = 1; k < KW; ++k)
        if (q == k) mine = acc[k];
      epi.put(row, d.k0 + q, mine);
    }
  }
  epi.finish(q < kwv ? q : -1, red);
}

}  // namespace gl
