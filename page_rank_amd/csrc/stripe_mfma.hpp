// Stripe-owner SpMM kernel for 8x8 PBR matrices at 32 or 8 RHS columns per launch whose tile
// products run on the matrix cores (gfx950), reading the matrix's execution layout (TileRec
// below). The description is of the 32-column form; at 8 columns each lane holds one column
// instead of four (one dword of the X block, two instructions per tile).
//
// It does the tile work of k_stripe_feat<8, 32> with fewer instructions: one dwordx4 X load
// per tile (instead of four dword loads), tile metadata through scalar loads, a three-
// instruction value decode per tile (two mbcnt and a v_cndmask whose lane mask is the
// transposed tile code), and 8 v_mfma_f32_4x4x1_16b_f32 per tile instead of 32 DPP FMAs.
// Measured on MI355X at 1M rows and 32 columns: tiles alone 149 -> 130 us, the fused product
// with the 5M-entry remainder 210 -> 197 us, where the launch's random 128-byte line reads
// (remainder gathers, tile X rows that miss L2, the matrix) run at about the rate a
// random-line read benchmark reaches on this chip (56 G lines/s).
//
// Layout (l = lane, the wave owns stripe rows r0 .. r0 + 7, columns k0 .. k0 + 31):
//   X block of a tile: lane l loads X[c0 + (l >> 3)][k0 + 4 (l & 7) .. + 3] (one float4), i.e.
//       tile column c = l >> 3, columns of quad q = l & 7: 8 rows x 128 bytes, contiguous.
//   Product: v_mfma_f32_4x4x1_16b_f32 (16 blocks of 4x4 outer products, FP32). Block
//       b = l >> 2 holds tile column c = b >> 1; with component j of the lanes' float4 as the
//       B operand, block b multiplies columns 4 (l & 7) + j of X row c. Its A operand (rows
//       4h .. 4h + 3 of tile column c) is broadcast from block ABID = h of its pair of blocks
//       (CBSZ 1): lane 8c + 4h + i holds the tile value at row 4h + i, column c, which is bit
//       l of the transposed code. One tile is 8 instructions (h = 0, 1; j = 0..3) into 8
//       accumulators of 4 rows each.
//   Each accumulator block sums one tile column; at the end of the stripe the partial sums of
//       the 8 tile columns (lane bits 3-5) are added with permlane32 / permlane16 swaps and a
//       DPP rotate, in a fixed order, leaving lane l with row l >> 3, columns 4 (l & 7) .. + 3:
//       the layout of the remainder rows (8 lanes per row, a float4 each) and of the
//       output store (one dwordx4 per lane).
// The tile values of a batch of up to 8 tiles are contiguous (at most 512 floats): they are
// loaded with float4 loads, staged in LDS, and read back per lane.
#pragma once

#include <type_traits>

#include "stripe_kernels.hpp"

namespace gl {

typedef float v4f __attribute__((ext_vector_type(4)));

// v_permlane32_swap(x, y) exchanges the high half of x with the low half of y: the sum of
// its results holds x's two halves added in lanes 0-31 and y's in lanes 32-63.
__device__ __forceinline__ float swap32_sum(float x, float y) {
  const auto p = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, x),
                                                  __builtin_bit_cast(uint32_t, y), false, false);
  return __builtin_bit_cast(float, (uint32_t)p[0]) + __builtin_bit_cast(float, (uint32_t)p[1]);
}
// v_permlane16_swap(x, y) exchanges the odd 16-lane rows of x with the even rows of y: the
// sum of its results holds x's rows 2k and 2k+1 added in row 2k and y's in row 2k+1.
__device__ __forceinline__ float swap16_sum(float x, float y) {
  const auto p = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(uint32_t, x),
                                                  __builtin_bit_cast(uint32_t, y), false, false);
  return __builtin_bit_cast(float, (uint32_t)p[0]) + __builtin_bit_cast(float, (uint32_t)p[1]);
}
// Lanes with bit 3 clear get x[l] + x[l ^ 8], lanes with bit 3 set y[l] + y[l ^ 8]
// (row_ror:8 inside a 16-lane row is lane ^ 8).
__device__ __forceinline__ float swap8_sum(float x, float y, bool b3) {
  const float keep = b3 ? y : x;
  const float send = b3 ? x : y;
  return keep + dpp<0x128>(send);
}

// Execution layout of an 8x8 PBR matrix, derived once at upload (k_exec_layout): one
// 16-byte record per tile, read with scalar loads, and the tile values in column-major bit
// order. The record holds the tile's code with its 8x8 bit matrix transposed (bit 8c + r set
// <=> row r, column c holds a value), so that bit l of it belongs to the lane that multiplies
// row l & 7 of tile column l >> 3, the tile's first X row (block column x 8) and its value
// offset (the same as in `offsets`: the count of values per tile does not change). The
// record array carries kTileRecPad zero records past the last tile, so a batch of 8 records
// can always be loaded whole.
struct TileRec {
  uint32_t tlo, thi;
  int32_t c0;
  int32_t off;
};
constexpr int kTileRecPad = 8;

// Stage capacity of the tile values of one batch (8 tiles of at most 64 values), plus one
// zero slot read by positions without a value.
constexpr int kMxVals = 512;
constexpr int kMxBatch = 8;  // tiles per batch

typedef __attribute__((address_space(4))) const TileRec c_TileRec;



// Largest value of v over the wave, with DPP and permlane swaps (no LDS round trips).
__device__ __forceinline__ int wave_max_fast(int v) {
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));   // lane ^ 1
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));   // lane ^ 2
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true));  // 8-lane half-row mirror
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true));  // 16-lane row mirror
  const auto a = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
  v = max((int)a[0], (int)a[1]);
  const auto b = __builtin_amdgcn_permlane32_swap((uint32_t)v, (uint32_t)v, false, false);
  return __builtin_amdgcn_readfirstlane(max((int)b[0], (int)b[1]));
}

// v_cndmask with a wave-uniform 64-bit lane mask: lane l gets b if bit l of mask is set,
// else a.
__device__ __forceinline__ uint32_t sel_by_mask(uint32_t a, uint32_t b, uint64_t mask) {
  uint32_t r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(mask));
  return r;
}

// KW: RHS columns of the launch, 32 (a float4 of columns per lane) or 8 (one column per
// lane: the X block of a tile is one dword per lane, and each tile takes 2 instructions).
// ALIGNED (32 columns): X rows and the launch's first column are 16-byte aligned, so each X
// block row segment loads with one 128-bit buffer load whose row offset is a scalar.
template <int KW, bool TILES, bool REM, bool ALIGNED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_stripe_mx(PbrView m, RhsView d, EpiStore epi) {
  static_assert(KW == 8 || KW == 32, "8 or 32 columns per launch");
  constexpr int B = 8;
  constexpr int L = 8;          // remainder: lanes per row
  constexpr int CPL = KW / 8;   // RHS columns per lane (tiles and remainder)
  constexpr int FPL = CPL;
  constexpr int EB = GL_REM_EB;
  __shared__ float vstage[4][kMxVals + 4];

  const PbrLd<true> ld(m, d);
  const Ld<true> tval(m.tdata, m.num_vals * 4);
  c_TileRec* __restrict__ trec = (c_TileRec*)m.trec;
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int rl = lane >> 3;  // the lane's output row inside the stripe (and remainder row)
  const int q = lane & 7;    // its group of columns: CPL q .. CPL q + CPL - 1
  const int kvalid = (int)min<int64_t>(KW, d.K - d.k0);
  const int nv = kvalid - CPL * q;  // columns of the group inside the matrix
  const uint32_t ldb = (uint32_t)d.ldx * 4u;
  const bool b3 = (lane >> 3) & 1;
  // byte offset of this lane's columns of an X block (row c0 + (lane >> 3), group q)
  const uint32_t xo = ((uint32_t)(lane >> 3) * (uint32_t)d.ldx + (uint32_t)(CPL * q)) * 4u;
  float* __restrict__ vs = vstage[wid];
  const uint32_t zslot = 4u * kMxVals;  // byte address of the zero slot
  if (lane == 0) vs[kMxVals] = 0.f;

  const Sched sc = xcd_sched(m.num_stripes);
  int64_t s = sc.first;
  int ta0 = 0, ta1 = 0, tb0 = 0, tb1 = 0;
  if (TILES && s < sc.end) {
    ta0 = sload(m.stripe_ptr, s);
    ta1 = sload(m.stripe_ptr, s + 1);
  }
  if (TILES && s + sc.stride < sc.end) {
    tb0 = sload(m.stripe_ptr, s + sc.stride);
    tb1 = sload(m.stripe_ptr, s + sc.stride + 1);
  }
  RowSpan rs{0, 0};
  if constexpr (REM) rs = load_row_span(ld, m, s * B + rl, s < sc.end);

  for (; s < sc.end; s += sc.stride) {
    const int64_t r0 = s * B;
    const int64_t sn = s + sc.stride;
    const int64_t snn = sn + sc.stride;
    // ---- prefetch: stripe pointers two stripes ahead, row spans one ahead
    int tc0 = 0, tc1 = 0;
    RowSpan rsn{0, 0};
    if (TILES && snn < sc.end) {
      tc0 = sload(m.stripe_ptr, snn);
      tc1 = sload(m.stripe_ptr, snn + 1);
    }
    if constexpr (REM) rsn = load_row_span(ld, m, sn * B + rl, sn < sc.end);
    // ---- first remainder entries of this stripe
    int scol = 0;
    float sval = 0.f;
    if constexpr (REM) {
      const bool ok = q < rs.n();
      scol = (int)ld.rcol.u32(rs.j0 + q, ok);
      sval = ld.rval.f32(rs.j0 + q, ok);
    }

    float out[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) out[j] = 0.f;
    if constexpr (TILES) {
      v4f acc[2][CPL];
      // One batch of up to 8 tiles from tile tb. FIRST: the stripe's first batch, whose first
      // tile starts the accumulators (a zero C operand instead of zeroing them).
      auto batch = [&](int tb, auto first_tag) {
        constexpr bool FIRST = decltype(first_tag)::value;
        const int nb = min(kMxBatch, ta1 - tb);
        // the batch's records (scalar loads; the padding makes 8 always readable)
        TileRec r[kMxBatch];
#pragma unroll
        for (int u = 0; u < kMxBatch; ++u) {
          r[u].tlo = trec[tb + u].tlo;
          r[u].thi = trec[tb + u].thi;
          r[u].c0 = trec[tb + u].c0;
          r[u].off = trec[tb + u].off;
        }
        // values of the batch: [offset of tile tb, end of its last tile)
        int bend = r[0].off + __popc(r[0].tlo) + __popc(r[0].thi);
#pragma unroll
        for (int u = 1; u < kMxBatch; ++u)
          if (u < nb) bend = r[u].off + __popc(r[u].tlo) + __popc(r[u].thi);
        const int boff = r[0].off;
        const int nval = bend - boff;
        const float4 v0 = tval.f32x4(boff + 4 * lane, 4 * lane < nval);
        float4 v1 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (nval > 256) v1 = tval.f32x4(boff + 256 + 4 * lane, 256 + 4 * lane < nval);
        float xv[kMxBatch][CPL];
#pragma unroll
        for (int u = 0; u < kMxBatch; ++u) {
          // a slot past the batch's last tile reads that tile's X block (its mask is empty)
          const int c0 = u < nb ? r[u].c0 : r[0].c0;
          const uint32_t so = (uint32_t)c0 * ldb;
          if constexpr (CPL == 1) {
            xv[u][0] = __builtin_bit_cast(
                float, __builtin_amdgcn_raw_buffer_load_b32(ld.x.r, xo, so, 0));
          } else if constexpr (ALIGNED) {
            const v4f t = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(ld.x.r, xo, so, 0));
            xv[u][0] = t.x;
            xv[u][1] = t.y;
            xv[u][2] = t.z;
            xv[u][3] = t.w;
          } else {
            const uint32_t vo = xo + so;
#pragma unroll
            for (int j = 0; j < 4; ++j)
              xv[u][j] = __builtin_bit_cast(
                  float, __builtin_amdgcn_raw_buffer_load_b32(ld.x.r, vo + 4u * j, 0, 0));
          }
        }
        *reinterpret_cast<float4*>(vs + 4 * lane) = v0;
        if (nval > 256) *reinterpret_cast<float4*>(vs + 256 + 4 * lane) = v1;
        wave_lds_sync();
        float av[kMxBatch];
#pragma unroll
        for (int u = 0; u < kMxBatch; ++u) {
          // stage byte address of this lane's value: 4 x (offset + set bits below the lane)
          const uint32_t mlo = u < nb ? r[u].tlo : 0u;
          const uint32_t mhi = u < nb ? r[u].thi : 0u;
          const uint32_t idx = __builtin_amdgcn_mbcnt_hi(
              mhi, __builtin_amdgcn_mbcnt_lo(mlo, (uint32_t)(r[u].off - boff)));
          const uint32_t a =
              sel_by_mask(zslot, idx * 4u, ((uint64_t)mhi << 32) | (uint64_t)mlo);
          av[u] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(vs) + a);
        }
        wave_lds_sync();
#pragma unroll
        for (int u = 0; u < kMxBatch; ++u) {
#pragma unroll
          for (int j = 0; j < CPL; ++j) {
            const v4f c0v = (FIRST && u == 0) ? v4f{0.f, 0.f, 0.f, 0.f} : acc[0][j];
            const v4f c1v = (FIRST && u == 0) ? v4f{0.f, 0.f, 0.f, 0.f} : acc[1][j];
            acc[0][j] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[u], xv[u][j], c0v, 1, 0, 0);
            acc[1][j] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[u], xv[u][j], c1v, 1, 1, 0);
          }
        }
            };
      if (ta0 < ta1) {
        batch(ta0, std::true_type{});
        for (int tb = ta0 + kMxBatch; tb < ta1; tb += kMxBatch) batch(tb, std::false_type{});
        // add the 8 tile columns' partial sums (lane bits 5, 4, 3 <-> row bits h, v1, v0)
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          float r1[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) r1[v] = swap32_sum(acc[0][j][v], acc[1][j][v]);
          const float r2a = swap16_sum(r1[0], r1[2]);
          const float r2b = swap16_sum(r1[1], r1[3]);
          out[j] = swap8_sum(r2a, r2b, b3);
        }
      }
    }

    if constexpr (REM) {
      // row lanes: the 8 lanes of row rl hold a float4 each and walk that row's entries
      float racc[FPL];
#pragma unroll
      for (int i = 0; i < FPL; ++i) racc[i] = 0.f;
      const int nmax = wave_max_fast(rs.n());
      const int gbase = lane & ~(L - 1);
      for (int jb = 0; jb < nmax; jb += L) {
        if (jb > 0) {
          const bool ok = jb + q < rs.n();
          scol = (int)ld.rcol.u32(rs.j0 + jb + q, ok);
          sval = ld.rval.f32(rs.j0 + jb + q, ok);
        }
        for (int ib = 0; ib < L && jb + ib < nmax; ib += EB) {
          float xr[EB][FPL];
          float a[EB];
#pragma unroll
          for (int i = 0; i < EB; ++i) {
            const int c = __shfl(scol, gbase + ib + i, kWave);
            a[i] = __shfl(sval, gbase + ib + i, kWave);
            const bool ok = jb + ib + i < rs.n();
            load_row<FPL>(ld.xr, xrow<true>(c, d.ldx) + CPL * q, true, ok ? nv : 0, xr[i]);
          }
#pragma unroll
          for (int i = 0; i < EB; ++i)
#pragma unroll
            for (int u = 0; u < FPL; ++u) racc[u] = __builtin_fmaf(a[i], xr[i][u], racc[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < FPL; ++u) out[u] += racc[u];
    }

    const int64_t row = r0 + rl;
    if (row < m.rows && nv > 0) {
      float* __restrict__ yp = epi.y + elem(row, epi.ldy, d.k0 + CPL * q, epi.u24);
      if constexpr (CPL == 1) {
        if (epi.accumulate) *yp += out[0];
        else __builtin_nontemporal_store(out[0], yp);
      } else if (nv >= 4 && (reinterpret_cast<uintptr_t>(yp) & 15u) == 0) {
        float4 o = make_float4(out[0], out[1], out[2], out[3]);
        if (epi.accumulate) {
          const float4 y0 = *reinterpret_cast<const float4*>(yp);
          o.x += y0.x;
          o.y += y0.y;
          o.z += y0.z;
          o.w += y0.w;
        }
        // non-temporal: Y is not read again by this launch, and its lines would otherwise
        // push tile X rows out of L2 (measured 200 -> 189 us at 1M rows, 32 columns)
        __builtin_nontemporal_store(v4f{o.x, o.y, o.z, o.w}, reinterpret_cast<v4f*>(yp));
      } else {
#pragma unroll
        for (int j = 0; j < CPL; ++j)
          if (j < nv) yp[j] = epi.accumulate ? yp[j] + out[j] : out[j];
      }
    }
    // ---- rotate the pipeline
    ta0 = tb0;
    ta1 = tb1;
    tb0 = tc0;
    tb1 = tc1;
    rs = rsn;
  }
}

}  // namespace gl
