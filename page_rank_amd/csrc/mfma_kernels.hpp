// Matrix-core variant of the 8x8 tile product for densely tiled matrices (SpMM epilogue).
//
// One wave owns a unit of two adjacent stripes (16 output rows). Tile i of the first stripe
// and tile i of the second form a pair, and the pair is one block-diagonal product on
// v_mfma_f32_16x16x4_f32, Y^T[16 features][16 rows] += X^T[16 features][16] * B[16][16 rows],
// taken as four K=4 steps:
//   step 0 / 1: columns 0-3 / 4-7 of the first stripe's tile (B zero on rows 8-15),
//   step 2 / 3: columns 0-3 / 4-7 of the second stripe's tile (B zero on rows 0-7),
// once per 16-feature block. Operand layouts (checked on gfx950): lane l = j + 16 k holds
// A[j][k] and B[k][j]; C[i][j] is register i % 4 of lane j + 16 (i / 4). So lane l feeds
// X[c0 + k][f0 + j] and the tile value at (row j, column k), and ends with features
// 4 (l / 16) .. +3 of row l % 16. The matrix pipe does the multiply-adds while the vector
// unit decodes tile bits, and the remainder rows (4 lanes per row, the wave's 16 rows) are
// folded in through LDS before the stores.
#pragma once

#include "stripe_kernels.hpp"

namespace gl {

typedef float gl_v4f __attribute__((ext_vector_type(4)));

template <int KW, bool REM, bool BUF>
__global__ void __launch_bounds__(256) k_stripe_mfma(PbrView m, RhsView d, EpiStore epi) {
  constexpr int NFB = KW / 16;        // 16-feature blocks
  constexpr int L = 4;                // remainder lanes per row (16 rows per unit)
  constexpr int FPL = KW / L;         // remainder columns per lane (8 or 16)
  constexpr int EB = 2;               // remainder entries per batch
  constexpr int PB = KW == 32 ? 4 : 2; // tile pairs whose loads are issued together
  __shared__ float racc[4][16 * KW];

  const PbrLd<BUF> ld(m, d);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int kk = lane >> 4;           // MFMA k index (0..3)
  const int jj = lane & 15;           // MFMA feature (A) / row (B, C) index
  const int kvalid = (int)min<int64_t>(KW, d.K - d.k0);
  const bool vec = (d.ldx & 3) == 0 && kvalid == KW;
  // float4 stores of finished rows: plain stores, 16-byte aligned rows, a full column chunk
  const bool st4 = !epi.accumulate && kvalid == KW && (epi.ldy & 3) == 0 && (d.k0 & 3) == 0 &&
                   ((uintptr_t)epi.y & 15) == 0;
  float* __restrict__ ra = racc[wid];
  // remainder lanes
  const int rr = lane / L;            // row 0..15 of the unit
  const int kq = (lane % L) * FPL;
  const int nvq = kvalid - kq;

  const int64_t units = (m.num_stripes + 1) / 2;
  const Sched sc = xcd_sched(units);
  for (int64_t u = sc.first; u < sc.end; u += sc.stride) {
    const int64_t sa = 2 * u, sb = 2 * u + 1;
    const int64_t r0 = sa * 8;
    const int ta0 = m.stripe_ptr[sa], ta1 = m.stripe_ptr[sa + 1];
    const int tb0 = ta1;
    const int tb1 = sb < m.num_stripes ? m.stripe_ptr[sb + 1] : ta1;
    // remainder row span of this lane's row
    RowSpan rs{0, 0};
    if constexpr (REM) rs = load_row_span(ld, m, r0 + rr);

    gl_v4f c[NFB];
#pragma unroll
    for (int f = 0; f < NFB; ++f) c[f] = gl_v4f{0.f, 0.f, 0.f, 0.f};

    const int na = ta1 - ta0, nb = tb1 - tb0;
    const int np = max(na, nb);
    for (int i0 = 0; i0 < np; i0 += 32) {
      // lanes 0..31: tiles i0.. of the first stripe; lanes 32..63: of the second
      const int ti = lane < 32 ? ta0 + i0 + lane : tb0 + i0 + (lane - 32);
      const bool tok = lane < 32 ? (i0 + lane < na) : (i0 + lane - 32 < nb);
      const uint64_t code = ld.codes.u64(ti, tok);
      const uint32_t mclo = (uint32_t)code, mchi = (uint32_t)(code >> 32);
      const int mc0 = (int)ld.coords.u32(2 * (int64_t)ti + 1, tok);
      const int moff = (int)ld.offsets.u32(ti, tok);
      const int pe = min(32, np - i0);
      for (int p0 = 0; p0 < pe; p0 += PB) {
        // loads of PB tile pairs first, then their products
        float bv[PB][2];
        float av[PB][NFB][4];
#pragma unroll
        for (int q = 0; q < PB; ++q) {
          const int p = p0 + q;  // < 32: PB divides 32
          // B operand: row jj of the unit lives in the first (jj < 8) or second tile
          const int src = jj < 8 ? p : 32 + p;
          const bool bt = jj < 8 ? (i0 + p < na) : (i0 + p < nb);
          const uint32_t clo = __shfl(mclo, src, kWave);
          const uint32_t chi = __shfl(mchi, src, kWave);
          const int off = __shfl(moff, src, kWave);
          const int q0 = (jj & 7) * 8 + kk, q1 = q0 + 4;
          bv[q][0] = ld.data.f32(off + code_prefix(clo, chi, q0), bt && code_bit(clo, chi, q0));
          bv[q][1] = ld.data.f32(off + code_prefix(clo, chi, q1), bt && code_bit(clo, chi, q1));
          // A operand: X rows of both tiles' columns
          const bool okA = i0 + p < na, okB = i0 + p < nb;
          const int c0a = __shfl(mc0, p, kWave), c0b = __shfl(mc0, 32 + p, kWave);
          const int ca[4] = {c0a + kk, c0a + 4 + kk, c0b + kk, c0b + 4 + kk};
          const bool cok[4] = {okA, okA, okB, okB};
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) {
            const bool okc = cok[s4] && (BUF || ca[s4] < m.cols);
            const int64_t xb = xrow<BUF>(ca[s4], d.ldx) + jj;
#pragma unroll
            for (int f = 0; f < NFB; ++f)
              av[q][f][s4] = ld.x.f32(xb + f * 16, okc && f * 16 + jj < kvalid);
          }
        }
#pragma unroll
        for (int q = 0; q < PB; ++q) {
          const float b0 = jj < 8 ? bv[q][0] : 0.f, b1 = jj < 8 ? bv[q][1] : 0.f;
          const float b2 = jj < 8 ? 0.f : bv[q][0], b3 = jj < 8 ? 0.f : bv[q][1];
#pragma unroll
          for (int f = 0; f < NFB; ++f) {
            c[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][f][0], b0, c[f], 0, 0, 0);
            c[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][f][1], b1, c[f], 0, 0, 0);
            c[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][f][2], b2, c[f], 0, 0, 0);
            c[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][f][3], b3, c[f], 0, 0, 0);
          }
        }
      }
    }

    if constexpr (REM) {
      // row lanes: the 4 lanes of row rr hold FPL columns each and walk that row's entries
      float racc2[FPL];
#pragma unroll
      for (int i = 0; i < FPL; ++i) racc2[i] = 0.f;
      const int nmax = wave_max(rs.n());
      for (int jb = 0; jb < nmax; jb += EB) {
        float xv[EB][FPL];
        float a[EB];
#pragma unroll
        for (int i = 0; i < EB; ++i) {
          const bool ok = jb + i < rs.n();
          const int cidx = (int)ld.rcol.u32(rs.j0 + jb + i, ok);
          a[i] = ld.rval.f32(rs.j0 + jb + i, ok);
          load_row<FPL>(ld.x, xrow<BUF>(cidx, d.ldx) + kq, vec, ok ? nvq : 0, xv[i]);
        }
#pragma unroll
        for (int i = 0; i < EB; ++i)
#pragma unroll
          for (int v = 0; v < FPL; ++v) racc2[v] = __builtin_fmaf(a[i], xv[i][v], racc2[v]);
      }
#pragma unroll
      for (int v = 0; v < FPL; ++v) ra[rr * KW + kq + v] = racc2[v];
      wave_lds_sync();
#pragma unroll
      for (int f = 0; f < NFB; ++f)
#pragma unroll
        for (int r = 0; r < 4; ++r) c[f][r] += ra[jj * KW + f * 16 + 4 * kk + r];
      wave_lds_sync();
    }

    // lane holds features f*16 + 4 kk .. +3 of row jj
    const int64_t row = r0 + jj;
    if (row < m.rows) {
      if (st4) {
        float* yr = epi.y + elem(row, epi.ldy, d.k0 + 4 * kk, epi.u24);
#pragma unroll
        for (int f = 0; f < NFB; ++f)
          *(float4*)(yr + f * 16) = make_float4(c[f][0], c[f][1], c[f][2], c[f][3]);
      } else {
#pragma unroll
        for (int f = 0; f < NFB; ++f)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int fe = f * 16 + 4 * kk + r;
            if (fe < kvalid) epi.template put<BUF>(0, row, d.k0 + fe, c[f][r]);
          }
      }
    }
  }
}

}  // namespace gl
