// Launch helpers shared by the SpMM and PPR entry points.
#pragma once

#include "gl_common.hpp"
#include "stripe_kernels.hpp"

namespace gl {

hipEvent_t lib_event(int i);
// Persistent grid: enough 4-wave workgroups for `work_items` waves, capped at wg_per_cu
// per CU (0 = 8), rounded to a multiple of the XCD count.
int grid_for(int64_t work_items, int wg_per_cu);
int choose_kw(int64_t K);
bool valid_kw(int kw);

template <int B, class Epi, bool TILES, bool REM>
int launch_stripe_b(const PbrView& m, const RhsView& d, int kw, const Epi& e, int nwg,
                    hipStream_t st) {
  const dim3 grid(nwg), block(256);
  switch (kw) {
    case 1: hipLaunchKernelGGL((k_stripe_bit<B, 1, TILES, REM, Epi>), grid, block, 0, st, m, d, e); break;
    case 2: hipLaunchKernelGGL((k_stripe_bit<B, 2, TILES, REM, Epi>), grid, block, 0, st, m, d, e); break;
    case 4: hipLaunchKernelGGL((k_stripe_bit<B, 4, TILES, REM, Epi>), grid, block, 0, st, m, d, e); break;
    case 8: hipLaunchKernelGGL((k_stripe_bit<B, 8, TILES, REM, Epi>), grid, block, 0, st, m, d, e); break;
    case 16: hipLaunchKernelGGL((k_stripe_feat<B, 16, TILES, REM, Epi>), grid, block, 0, st, m, d, e); break;
    case 32: hipLaunchKernelGGL((k_stripe_feat<B, 32, TILES, REM, Epi>), grid, block, 0, st, m, d, e); break;
    case 64: hipLaunchKernelGGL((k_stripe_feat<B, 64, TILES, REM, Epi>), grid, block, 0, st, m, d, e); break;
    default:
      set_error("unsupported kw %d", kw);
      return GL_ERR_INVALID;
  }
  GL_LAUNCH_CHECK("k_stripe");
  return GL_OK;
}

template <class Epi, bool TILES, bool REM>
int launch_stripe(const PbrView& m, const RhsView& d, int B, int kw, const Epi& e, int nwg,
                  hipStream_t st) {
  switch (B) {
    case 2: return launch_stripe_b<2, Epi, TILES, REM>(m, d, kw, e, nwg, st);
    case 4: return launch_stripe_b<4, Epi, TILES, REM>(m, d, kw, e, nwg, st);
    case 8: return launch_stripe_b<8, Epi, TILES, REM>(m, d, kw, e, nwg, st);
    default:
      set_error("Unsupported block size %d: only 2x2, 4x4 and 8x8 are supported", B);
      return GL_ERR_INVALID;
  }
}

}  // namespace gl
