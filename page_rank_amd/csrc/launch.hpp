// Launch helpers shared by the SpMM and PPR entry points.
#pragma once

#include "gl_common.hpp"
#include "stripe_kernels.hpp"
#include "stripe_mfma.hpp"

#include <algorithm>
#include <atomic>
#include <type_traits>

namespace gl {

// Persistent grid: enough 4-wave workgroups for `work_items` waves, capped at wg_per_cu
// per CU (0 = 8), rounded to a multiple of the XCD count.
int grid_for(int64_t work_items, int wg_per_cu);
int choose_kw(int64_t K);
bool valid_kw(int kw);
// Width of the launch for columns [k0, K) taken kw at a time: the last, partial chunk runs
// in the narrowest kernel that covers it (a 64-column launch with 32 live columns costs
// about 1.5x a 32-column one).
inline int chunk_kw(int kw, int64_t left) {
  if (left >= kw) return kw;
  const int c = choose_kw(left);
  return c < kw ? c : kw;
}

// One launch of the bit-lane (FEAT false) or feature-lane kernel. The VEC form exists only
// where a lane's slice of an X row is 4 or more floats: bit-lane kernels of 4 and 8 columns,
// feature-lane remainders with KW / (64 / B) >= 4.
// Workgroups of `kernel` (256 threads) that one CU holds at once.
template <class F>
int resident_per_cu(F kernel) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, 256, 0) != hipSuccess) n = 0;
  return n;
}

// One launch of the bit-lane (FEAT false) or feature-lane kernel, or, with `occ` set, no
// launch: *occ = workgroups per CU of the kernel that would run. The VEC form exists only
// where a lane's slice of an X row is 4 or more floats: bit-lane kernels of 4 and 8 columns,
// feature-lane remainders with KW / (64 / B) >= 4.
#define GL_GO(KERN)                                                        \
  do {                                                                     \
    static std::atomic<int> occ_cached{-1};                                \
    if (occ) {                                                             \
      if (occ_cached.load() < 0) occ_cached.store(resident_per_cu(KERN));  \
      *occ = occ_cached.load();                                            \
    } else {                                                               \
      hipLaunchKernelGGL(KERN, grid, block, 0, st, m, d, e);               \
    }                                                                      \
  } while (0)
template <int B, int KW, bool FEAT, bool TILES, bool REM, bool BUF, class Epi, int U>
void launch_vec(const dim3& grid, hipStream_t st, const PbrView& m, const RhsView& d,
                const Epi& e, bool vec, int* occ) {
  constexpr bool useful = FEAT ? KW / (kWave / B) >= 4 : KW >= 4;
  const dim3 block(256);
  if constexpr (useful) {
    if (vec) {
      if constexpr (FEAT) GL_GO((k_stripe_feat<B, KW, TILES, REM, BUF, Epi, U, true>));
      else GL_GO((k_stripe_bit<B, KW, TILES, REM, BUF, Epi, U, true>));
      return;
    }
  }
  if constexpr (FEAT) GL_GO((k_stripe_feat<B, KW, TILES, REM, BUF, Epi, U, false>));
  else GL_GO((k_stripe_bit<B, KW, TILES, REM, BUF, Epi, U, false>));
}

template <int B, class Epi, bool TILES, bool REM, bool BUF>
int launch_stripe_bb(const PbrView& m, const RhsView& d, int kw, const Epi& e, int nwg,
                     hipStream_t st, int* occ) {
  const dim3 grid(nwg), block(256);
  // tile batch: a whole stripe at once when 8x8 stripes hold about 8 tiles (measured on
  // MI355X: -16 % at 1 column, -3 % at 32; narrower batches win on sparser tilings). For the
  // feature-lane kernels that is 8 / (64 / KW) groups, 2 at 16 columns (8 columns computed
  // with kw 16: 234 -> 196 us against 4 groups) and 4 at 32; 64 columns stay at 4 (registers).
  const bool wide = B == 8 && m.num_tiles >= 6 * m.num_stripes;
  // float4 X-row loads (the VEC kernels, no runtime branch per X load). The plain-pointer
  // form needs 16-byte aligned rows and a full column chunk. Buffer loads need only dword
  // alignment and are range-checked per dword, so there a partial last chunk reads the
  // columns past K as well: the next row's values (0 past the array), which only reach
  // accumulators of columns that are not stored.
  const bool full = std::min<int64_t>(kw, d.K - d.k0) == kw;
  const bool vec = BUF || ((d.ldx & 3) == 0 && full);
  if constexpr (B == 8 && BUF && TILES && std::is_same_v<Epi, EpiStore>) {
    if (m.trec != nullptr && m.tdata != nullptr && (kw == 32 || kw == 8)) {
      if (kw == 8) {
        GL_GO((k_stripe_mx<8, TILES, REM, false>));
      } else if ((d.ldx & 3) == 0 && (d.k0 & 3) == 0 &&
                 (reinterpret_cast<uintptr_t>(d.x) & 15u) == 0) {
        GL_GO((k_stripe_mx<32, TILES, REM, true>));
      } else {
        GL_GO((k_stripe_mx<32, TILES, REM, false>));
      }
      if (!occ) GL_LAUNCH_CHECK("k_stripe_mx");
      return GL_OK;
    }
  }
  switch (kw) {
#define GL_KB(KW)                                                                        \
    case KW:                                                                             \
      if (wide)                                                                          \
        launch_vec<B, KW, false, TILES, REM, BUF, Epi, (B == 8 ? 8 : 2)>(grid, st, m, d, e, \
                                                                         vec, occ);      \
      else                                                                               \
        launch_vec<B, KW, false, TILES, REM, BUF, Epi, (B == 8 ? 4 : 2)>(grid, st, m, d, e, \
                                                                         vec, occ);      \
      break;
#define GL_KF(KW)                                                                        \
    case KW:                                                                             \
      if (wide)                                                                          \
        launch_vec<B, KW, true, TILES, REM, BUF, Epi, (KW == 16 ? 2 : 4)>(grid, st, m, d, e, \
                                                                          vec, occ);     \
      else                                                                               \
        launch_vec<B, KW, true, TILES, REM, BUF, Epi, 1>(grid, st, m, d, e, vec, occ);   \
      break;
    GL_KB(1) GL_KB(2) GL_KB(4) GL_KB(8)
    GL_KF(16) GL_KF(32) GL_KF(64)
#undef GL_KB
#undef GL_KF
    default:
      set_error("unsupported kw %d", kw);
      return GL_ERR_INVALID;
  }
  if (!occ) GL_LAUNCH_CHECK("k_stripe");
  return GL_OK;
}

// Buffer-descriptor loads when every array of the product is below 2 GB and X rows
// are addressable with 24-bit products.
inline bool buffer_ok(const PbrView& m, const RhsView& d) {
  return m.num_tiles * 8 < kBufLimit && m.num_vals * 4 < kBufLimit &&
         m.num_rem * 4 < kBufLimit && (m.rows + 1) * 4 < kBufLimit && d.len * 4 < kBufLimit &&
         m.cols < (1 << 24) && d.ldx < (1 << 24);
}

template <int B, class Epi, bool TILES, bool REM>
int launch_stripe_b(const PbrView& m, const RhsView& d, int kw, const Epi& e, int nwg,
                    hipStream_t st, int* occ) {
  if (buffer_ok(m, d))
    return launch_stripe_bb<B, Epi, TILES, REM, true>(m, d, kw, e, nwg, st, occ);
  return launch_stripe_bb<B, Epi, TILES, REM, false>(m, d, kw, e, nwg, st, occ);
}

template <class Epi, bool TILES, bool REM>
int launch_stripe(const PbrView& m, const RhsView& d, int B, int kw, const Epi& e, int nwg,
                  hipStream_t st, int* occ = nullptr) {
  switch (B) {
    case 2: return launch_stripe_b<2, Epi, TILES, REM>(m, d, kw, e, nwg, st, occ);
    case 4: return launch_stripe_b<4, Epi, TILES, REM>(m, d, kw, e, nwg, st, occ);
    case 8: return launch_stripe_b<8, Epi, TILES, REM>(m, d, kw, e, nwg, st, occ);
    default:
      set_error("Unsupported block size %d: only 2x2, 4x4 and 8x8 are supported", B);
      return GL_ERR_INVALID;
  }
}

#undef GL_GO

// Persistent grid for the kernel that launch_stripe would run: `nwg` workgroups, cut to
// what the device holds at once, so that no workgroup waits for another to end before it
// starts its stripes.
template <class Epi, bool TILES, bool REM>
int stripe_grid(const PbrView& m, const RhsView& d, int B, int kw, const Epi& e, int nwg) {
  int occ = 0;
  if (launch_stripe<Epi, TILES, REM>(m, d, B, kw, e, nwg, nullptr, &occ) || occ < 1) return nwg;
  int64_t cap = (int64_t)occ * device_info().num_cus;
  cap -= cap % kNumXcd;
  return cap >= kNumXcd && cap < nwg ? (int)cap : nwg;
}

}  // namespace gl
