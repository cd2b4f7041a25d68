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

// Remainder entries whose X-row gathers are issued together (feature-lane kernel).
#ifndef GL_REM_EB
#define GL_REM_EB 4
#endif
// The same for the 64-column feature-lane kernel. Measured on the 4M-node PPR step, with the
// loop bound a scalar: 2 entries 883 PPR/s, 1 entry 857, 4 entries 844 (fewer waves).
#ifndef GL_REM_EB64
#define GL_REM_EB64 2
#endif
// The same for the plain-pointer form (arrays of 2 GB or more): 1 entry measured 15 % slower
// than 2 for the 16M-node PPR step.
#ifndef GL_REM_EB64_PTR
#define GL_REM_EB64_PTR 2
#endif
// Minimum waves per SIMD of the 64-column feature-lane kernel (caps its VGPRs).
#ifndef GL_FEAT64_WAVES
#define GL_FEAT64_WAVES 4
#endif

namespace gl {

struct PbrView {
  int64_t rows, cols;
  int64_t num_stripes;
  int64_t num_tiles, num_vals, num_rem;  // array lengths (buffer-descriptor bounds)
  const uint64_t* __restrict__ codes;
  const int32_t* __restrict__ coords;
  const int32_t* __restrict__ offsets;
  const float* __restrict__ data;
  const int32_t* __restrict__ stripe_ptr;
  const int32_t* __restrict__ rem_ptr;
  const int32_t* __restrict__ rem_col;
  const float* __restrict__ rem_val;
  // execution layout of 8x8 matrices (tile records and column-major values), or null
  const void* __restrict__ trec = nullptr;
  const float* __restrict__ tdata = nullptr;
};

struct RhsView {
  const float* __restrict__ x;
  int64_t ldx;
  int64_t K;   // total RHS columns
  int64_t k0;  // first column of this launch
  int64_t len; // elements of x (rows of X times ldx)
};

// --------------------------------------------------------------------------------------
// Predicated loads. Ld<true> goes through a raw buffer descriptor: a lane whose predicate
// is false reads an offset beyond the descriptor's bound and gets 0 from the hardware range
// check, so no branch guards the load and all 32-bit offsets stay in VGPRs (X column steps
// ride in the scalar offset). Ld<false> is the plain-pointer form for arrays of 2 GB or
// more. Both return 0 for a false predicate.
// --------------------------------------------------------------------------------------
constexpr uint32_t kOob = 0x80000000u;        // never inside a descriptor (< 2 GB)
constexpr int64_t kBufLimit = 0x7fff0000ll;    // largest array (bytes) for Ld<true>

template <bool BUF>
struct Ld;

template <>
struct Ld<true> {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Ld(const void* p, int64_t bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                            0x00020000)) {}
  __device__ __forceinline__ uint32_t u32(int64_t i, bool ok, uint32_t sb = 0) const {
    return __builtin_amdgcn_raw_buffer_load_b32(r, ok ? (uint32_t)i * 4u : kOob, sb, 0);
  }
  __device__ __forceinline__ float f32(int64_t i, bool ok, uint32_t sb = 0) const {
    return __builtin_bit_cast(float, u32(i, ok, sb));
  }
  // Wide loads are built from dword loads at consecutive offsets (the backend merges them
  // into buffer_load_dwordx2/x4); the vector-result buffer builtins are not used.
  __device__ __forceinline__ uint64_t u64(int64_t i, bool ok) const {
    const uint32_t o = ok ? (uint32_t)i * 8u : kOob;
    const uint32_t lo = __builtin_amdgcn_raw_buffer_load_b32(r, o, 0, 0);
    const uint32_t hi = __builtin_amdgcn_raw_buffer_load_b32(r, o + 4u, 0, 0);
    return ((uint64_t)hi << 32) | lo;
  }
  __device__ __forceinline__ float4 f32x4(int64_t i, bool ok) const {
    const uint32_t o = ok ? (uint32_t)i * 4u : kOob;
    return make_float4(
        __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, o, 0, 0)),
        __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, o + 4u, 0, 0)),
        __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, o + 8u, 0, 0)),
        __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, o + 12u, 0, 0)));
  }
};

template <>
struct Ld<false> {
  const char* p;
  __device__ __forceinline__ Ld(const void* q, int64_t) : p((const char*)q) {}
  __device__ __forceinline__ uint32_t u32(int64_t i, bool ok, uint32_t sb = 0) const {
    return ok ? *(const uint32_t*)(p + i * 4 + sb) : 0u;
  }
  __device__ __forceinline__ float f32(int64_t i, bool ok, uint32_t sb = 0) const {
    return ok ? *(const float*)(p + i * 4 + sb) : 0.f;
  }
  __device__ __forceinline__ uint64_t u64(int64_t i, bool ok) const {
    return ok ? *(const uint64_t*)(p + i * 8) : 0ull;
  }
  __device__ __forceinline__ float4 f32x4(int64_t i, bool ok) const {
    return ok ? *(const float4*)(p + i * 4) : make_float4(0.f, 0.f, 0.f, 0.f);
  }
};

// The arrays of one SpMM as predicated loaders.
template <bool BUF>
struct PbrLd {
  Ld<BUF> codes, coords, offsets, data, rptr, rcol, rval;
  Ld<BUF> x;
  Ld<BUF> xr;  // X rows gathered by remainder entries
  __device__ __forceinline__ PbrLd(const struct PbrView& m, const RhsView& d);
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
    // s_nop 1: two wait states between the VALU write of v[] and its first DPP read.
    asm("s_nop 1\n" GL_ROW8_LO(0, 8) GL_ROW8_HI(1, 8) GL_ROW8_LO(2, 9) GL_ROW8_HI(3, 9)
            GL_ROW8_LO(4, 10) GL_ROW8_HI(5, 10) GL_ROW8_LO(6, 11) GL_ROW8_HI(7, 11)
        : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]),
          "+v"(a[6]), "+v"(a[7])
        : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(x[0]), "v"(x[1]), "v"(x[2]),
          "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]));
  }
};

template <>
struct TileFma<4> {
  static __device__ __forceinline__ void run(float (&a)[4], const float (&v)[1],
                                             const float (&x)[4]) {
#define GL_F4(a_, n, xx) GL_F(a_, 4, xx, n)
    asm("s_nop 1\n"
        GL_F4(0, 0, 5) GL_F4(0, 1, 6) GL_F4(0, 2, 7) GL_F4(0, 3, 8)
        GL_F4(1, 4, 5) GL_F4(1, 5, 6) GL_F4(1, 6, 7) GL_F4(1, 7, 8)
        GL_F4(2, 8, 5) GL_F4(2, 9, 6) GL_F4(2, 10, 7) GL_F4(2, 11, 8)
        GL_F4(3, 12, 5) GL_F4(3, 13, 6) GL_F4(3, 14, 7) GL_F4(3, 15, 8)
        : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3])
        : "v"(v[0]), "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]));
#undef GL_F4
  }
};

template <>
struct TileFma<2> {
  static __device__ __forceinline__ void run(float (&a)[2], const float (&v)[1],
                                             const float (&x)[2]) {
    asm("s_nop 1\n" GL_F(0, 2, 3, 0) GL_F(0, 2, 4, 1) GL_F(1, 2, 3, 2) GL_F(1, 2, 4, 3)
        : "+v"(a[0]), "+v"(a[1])
        : "v"(v[0]), "v"(x[0]), "v"(x[1]));
  }
};
#undef GL_ROW8_LO
#undef GL_ROW8_HI
#undef GL_F

// --------------------------------------------------------------------------------------
// Epilogues. A lane's output column is fixed for the launch (lane_init, once). At the
// start of each stripe, begin_stripe() (wave-uniform) and prefetch(i, row, col) for the
// lane's output elements run; at its end put(i, row, col, acc) once per output element;
// finish() once per lane at the end with the lane's column slot (or -1).
// --------------------------------------------------------------------------------------
// Element offset row * ld + col; `u24` (set by the host when rows and ld are below 2^24 and
// their product below 2^32) selects a full-rate 24-bit multiply.
__device__ __forceinline__ int64_t elem(int64_t row, int64_t ld, int64_t col, bool u24) {
  return u24 ? (int64_t)__umul24((uint32_t)row, (uint32_t)ld) + col : row * ld + col;
}
inline bool fits_u24(int64_t rows, int64_t ld) {
  return rows < (1 << 24) && ld < (1 << 24) && rows * ld < (int64_t)1 << 32;
}

struct EpiStore {
  float* __restrict__ y;
  int64_t ldy;
  int32_t accumulate;
  bool u24;
  __device__ __forceinline__ bool active() const { return true; }
  __device__ __forceinline__ void lane_init(int64_t, bool) {}
  __device__ __forceinline__ void begin_stripe(int64_t) {}
  template <bool BUF>
  __device__ __forceinline__ void prefetch(int, int64_t, int64_t, bool) {}
  template <bool BUF>
  __device__ __forceinline__ void put(int, int64_t row, int64_t col, float acc) {
    float* p = y + elem(row, ldy, col, u24);
    if (accumulate) *p += acc;
    else *p = acc;
  }
  __device__ __forceinline__ void finish(int) {}
};

// Device control block of a PPR run.
struct PprCtl {
  int32_t iter;    // completed iterations
  int32_t active;  // columns still iterating
  int32_t kick;    // the pass just retired is followed by an extrapolation
  int32_t pad;
};

// Fused power-iteration step:
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
  bool u24;
  int64_t xlen;    // elements of xcur (and of xnext)
  // per-lane state
  float res;
  float dang;
  int32_t e0, e1;
  int32_t frz;     // this lane's column is frozen
  float cuv, tev;  // cu[col], tele[col] of this lane's column
  float xc[8];     // X_t at this lane's output elements of the current stripe

  __device__ __forceinline__ bool active() const { return ctl->active > 0; }
  // The lane's column is fixed for the whole launch: read its state once.
  __device__ __forceinline__ void lane_init(int64_t col, bool ok) {
    frz = ok ? frozen[col] : 1;
    cuv = ok ? cu[col] : 0.f;
    tev = ok ? tele[col] : 0.f;
  }
  __device__ __forceinline__ void begin_stripe(int64_t s) {
    e0 = sload(sptr, s);
    e1 = sload(sptr, s + 1);
  }
  // Element offset of (row, col) in X_t / X_{t+1}. With BUF (the buffer-descriptor
  // kernels, whose launch guarantees arrays below 2 GB and rows, ld below 2^24) it is a
  // 24-bit multiply with no runtime test.
  template <bool BUF>
  __device__ __forceinline__ int64_t at(int64_t row, int64_t col) const {
    if constexpr (BUF) return (int64_t)__umul24((uint32_t)row, (uint32_t)ld) + col;
    else return elem(row, ld, col, u24);
  }
  // Issued at the start of a stripe, so X_t arrives while the stripe is multiplied. With BUF
  // a predicated buffer load (no branch around it).
  template <bool BUF>
  __device__ __forceinline__ void prefetch(int i, int64_t row, int64_t col, bool ok) {
    if constexpr (BUF) xc[i] = Ld<true>(xcur, xlen * 4).f32(at<true>(row, col), ok);
    else xc[i] = ok ? xcur[at<false>(row, col)] : 0.f;
  }
  template <bool BUF>
  __device__ __forceinline__ void put(int i, int64_t row, int64_t col, float acc) {
    const int64_t ix = at<BUF>(row, col);
    const float x0 = xc[i];
    float xn = x0;
    if (!frz) {
      float w = 0.f;
      for (int e = e0; e < e1; ++e)
        if (srow[e] == row && scol[e] == col) w += sw[e];
      xn = __builtin_fmaf(alpha, acc, cuv);
      xn = __builtin_fmaf(w, tev, xn);
      res += fabsf(xn - x0);
      if (dmask != nullptr && dmask[row]) dang += xn;
    }
    // non-temporal: the next iterate is read only by the next pass (measured +1 % PPR/s)
    __builtin_nontemporal_store(xn, xnext + ix);
  }
  // slot: the lane's column minus k0 (or -1). The workgroup's partial of a column is the
  // sum of its lanes' values taken in thread order, so partials (and the residuals built
  // from them) do not depend on which wave finishes first. Once per workgroup and launch.
  __device__ __forceinline__ void finish(int slot) {
    __shared__ float s_res[256], s_dang[256];
    __shared__ int s_slot[256];
    __shared__ float s_q[2][4][64];
    const int t = threadIdx.x;
    s_res[t] = res;
    s_dang[t] = dang;
    s_slot[t] = slot;
    __syncthreads();
    {  // thread (quarter q, slot c) sums the lanes of wave q whose slot is c
      const int c = t & 63, q = t >> 6;
      float r = 0.f, dg = 0.f;
      for (int i = 64 * q; i < 64 * q + 64; ++i) {
        if (s_slot[i] == c) {
          r += s_res[i];
          dg += s_dang[i];
        }
      }
      s_q[0][q][c] = r;
      s_q[1][q][c] = dg;
    }
    __syncthreads();
    if (t < 64) {
      const int64_t col = k0 + t;
      if (col < K) {
        part_res[blockIdx.x * K + col] = ((s_q[0][0][t] + s_q[0][1][t]) + s_q[0][2][t]) + s_q[0][3][t];
        part_dang[blockIdx.x * K + col] = ((s_q[1][0][t] + s_q[1][1][t]) + s_q[1][2][t]) + s_q[1][3][t];
      }
    }
  }
};

// --------------------------------------------------------------------------------------
// Cross-lane sums. DPP for neighbours inside a 16-lane row, bpermute for wider strides.
// --------------------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
// Sum over the lanes that differ only in bit `MASK` of the lane id.
template <int MASK>
__device__ __forceinline__ float xor_sum(float v) {
  if constexpr (MASK == 1) return v + dpp<0xB1>(v);  // quad_perm [1,0,3,2]
  else if constexpr (MASK == 2) return v + dpp<0x4E>(v);  // quad_perm [2,3,0,1]
  else return v + __shfl_xor(v, MASK, kWave);
}
// Sum over aligned groups of 8 lanes (row_half_mirror completes the 8-lane sum once the
// two quads are uniform).
__device__ __forceinline__ float sum8(float v) {
  v = xor_sum<1>(v);
  v = xor_sum<2>(v);
  return v + dpp<0x141>(v);
}

template <bool BUF>
__device__ __forceinline__ PbrLd<BUF>::PbrLd(const PbrView& m, const RhsView& d)
    : codes(m.codes, m.num_tiles * 8),
      coords(m.coords, m.num_tiles * 8),
      offsets(m.offsets, (m.num_tiles + 1) * 4),
      data(m.data, m.num_vals * 4),
      rptr(m.rem_ptr, (m.rows + 1) * 4),
      rcol(m.rem_col, m.num_rem * 4),
      rval(m.rem_val, m.num_rem * 4),
      x(d.x + d.k0, (d.len - d.k0) * 4),
      xr(d.x + d.k0, (d.len - d.k0) * 4) {}

// N (1..8) consecutive floats of X starting at element i, zeros for lanes with valid <= 0.
// `wide` (wave-uniform: 16-byte aligned rows and a full column chunk) selects float4 loads.
template <int N, bool BUF>
__device__ __forceinline__ void load_row(const Ld<BUF>& xl, int64_t i, bool wide, int valid,
                                         float (&out)[N]) {
  if constexpr (N >= 4) {
    if (wide) {
#pragma unroll
      for (int k = 0; k < N; k += 4) {
        const float4 t = xl.f32x4(i + k, valid > 0);
        out[k] = t.x;
        out[k + 1] = t.y;
        out[k + 2] = t.z;
        out[k + 3] = t.w;
      }
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < N; ++k) out[k] = xl.f32(i + k, k < valid);
}

// Bit p of a tile code given as two 32-bit words, and the number of set bits below it
// (the index of p's value inside the tile's packed values); 32-bit ALU only.
__device__ __forceinline__ bool code_bit(uint32_t clo, uint32_t chi, int p) {
  const uint32_t w = p < 32 ? clo : chi;
  return (w >> (p & 31)) & 1u;
}
__device__ __forceinline__ int code_prefix(uint32_t clo, uint32_t chi, int p) {
  const uint32_t w = p < 32 ? clo : chi;
  const int base = p < 32 ? 0 : __popc(clo);
  return base + __popc(w & ((1u << (p & 31)) - 1u));
}
// Element offset of row r of X: 24-bit multiply on the buffer path (rows and ldx below
// 2^24 there), 64-bit otherwise.
template <bool BUF>
__device__ __forceinline__ int64_t xrow(int64_t r, int64_t ldx) {
  if constexpr (BUF) return (int64_t)__umul24((uint32_t)r, (uint32_t)ldx);
  else return r * ldx;
}
// The result is returned as a scalar: loops bounded by it are then uniform, which lets the
// compiler keep their loads in flight together instead of masking every iteration.
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v = max(v, __shfl_xor(v, m, kWave));
  return __builtin_amdgcn_readfirstlane(v);
}

// Interleaved form of load_row for a lane slice of N = 4 H floats (H >= 2) among L lanes of
// a row: chunk h of lane q holds columns h * 4L + 4q .. +3, so the L lanes' h-th float4 loads
// read one contiguous 16 L-byte span of the row (a single 128-byte line at 64 columns)
// instead of every line of it. `valid`: columns of the row that may be read (0: none).
template <int N, int L, bool BUF>
__device__ __forceinline__ void load_row_il(const Ld<BUF>& xl, int64_t row0, int q, bool wide,
                                            int valid, float (&out)[N]) {
  static_assert(N % 4 == 0 && N >= 8, "interleaved rows take whole float4 chunks");
#pragma unroll
  for (int h = 0; h < N / 4; ++h) {
    const int c = h * 4 * L + 4 * q;
    if (wide) {
      const float4 t = xl.f32x4(row0 + c, c < valid);
      out[4 * h] = t.x;
      out[4 * h + 1] = t.y;
      out[4 * h + 2] = t.z;
      out[4 * h + 3] = t.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) out[4 * h + j] = xl.f32(row0 + c + j, c + j < valid);
    }
  }
}

// Metadata of up to 64 tiles [t, t1), one tile per lane.
struct TileMeta {
  uint32_t clo, chi;
  int c0, off;
};
template <bool BUF>
__device__ __forceinline__ TileMeta load_tile_meta(const PbrLd<BUF>& ld, int t, int t1,
                                                   int lane) {
  const int tl = t + lane;
  const bool ok = tl < t1;
  const uint64_t c = ld.codes.u64(tl, ok);
  TileMeta r;
  r.clo = (uint32_t)c;
  r.chi = (uint32_t)(c >> 32);
  r.c0 = (int)ld.coords.u32(2 * (int64_t)tl + 1, ok);
  r.off = (int)ld.offsets.u32(tl, ok);
  return r;
}

// Remainder entries of one matrix row: [j0, j1). The length is taken where it is used: a
// subtraction right after the two loads makes the wave wait for every load in flight
// (measured: 2-4.5 % slower SpMM at 2-64 columns and PPR steps). EAGER keeps that early
// wait, which the one-column bit-lane kernel measured 6 % faster with.
struct RowSpan {
  int j0, j1;
  __device__ __forceinline__ int n() const { return j1 - j0; }
};
template <bool EAGER = false, bool BUF>
__device__ __forceinline__ RowSpan load_row_span(const PbrLd<BUF>& ld, const PbrView& m,
                                                 int64_t row, bool pred = true) {
  const bool ok = pred && row < m.rows;
  RowSpan r;
  r.j0 = (int)ld.rptr.u32(row, ok);
  r.j1 = (int)ld.rptr.u32(row + 1, ok);
  if constexpr (EAGER) asm volatile("" : "+v"(r.j0), "+v"(r.j1));
  return r;
}

// --------------------------------------------------------------------------------------
// Both kernels walk their stripes with a software pipeline: stripe pointers are loaded two
// stripes ahead, tile metadata and remainder row spans one stripe ahead, and the first
// remainder entries of a stripe are requested before its tile loads, so each stripe costs
// about two dependent memory round trips.
// --------------------------------------------------------------------------------------

// Feature-lane kernel, KW in {16, 32, 64}. UF: tile groups whose loads are issued together
// (the launcher picks a whole stripe, 2 groups at 16 columns and 4 at 32 and 64, for 8x8
// matrices with about 8 tiles per stripe, else 1). VEC: X rows load as float4 (buffer
// launches always; plain-pointer launches for aligned rows and a full column chunk).
template <int B, int KW, bool TILES, bool REM, bool BUF, class Epi, int UF, bool VEC>
__global__ void __launch_bounds__(256)
// the 64-column kernel is held to 128 VGPRs (4 waves per SIMD)
__attribute__((amdgpu_waves_per_eu(KW == 64 ? GL_FEAT64_WAVES : 1)))
k_stripe_feat(PbrView m, RhsView d, Epi epi) {
  constexpr int G = kWave / KW;           // tiles in flight per wave
  constexpr int P = B * B;                // bits per tile
  constexpr int NV = (P + 15) / 16;       // value registers per lane
  constexpr int U = UF;
  constexpr int L = kWave / B;            // remainder: lanes per row
  constexpr int FPL = (KW / L) > 0 ? (KW / L) : 1;  // remainder: RHS columns per lane
  // remainder: X gathers issued together (fewer at 64 columns, where each holds 8 floats
  // per lane and registers decide the occupancy)
  constexpr int EBW = KW >= 64 ? (BUF ? GL_REM_EB64 : GL_REM_EB64_PTR) : GL_REM_EB;
  constexpr int EB = L < EBW ? L : EBW;
  static_assert(L % EB == 0, "remainder batches must not straddle a block of L entries");
  __shared__ float racc[4][B * KW];

  if (!epi.active()) return;
  const PbrLd<BUF> ld(m, d);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int g = lane / KW;
  const int f = lane % KW;
  const int o = lane & 15;
  // Split decode: the R = 4 / G DPP rows of a tile each decode NL = NV / R of its 16-position
  // chunks (row ri takes chunks ri * NL ..), and permlane swaps give every row all of them.
  // (buffer-descriptor kernels only: with plain-pointer loads it measured 3 % slower at 16M rows)
  constexpr bool kSplit = B == 8 && G <= 2 && BUF;
  constexpr int R = kSplit ? 4 / G : 1;
  constexpr int NL = NV / R;
  const int ri = (lane >> 4) & (R - 1);
  const bool whi = R == 2 ? ri == 1 : ri >= 2;  // this row's chunks lie in the code's high word
  uint32_t sbit[NL], smask[NL];                 // bit of each decoded position, bits below it
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int sh = (R == 2 ? 16 * j : 16 * (ri & 1)) + o;
    sbit[j] = 1u << sh;
    smask[j] = sbit[j] - 1u;
  }
  const int64_t colk = d.k0 + f;
  const bool fok = colk < d.K;
  epi.lane_init(colk, fok);
  const int kvalid = (int)min<int64_t>(KW, d.K - d.k0);
  // float4 rows and a full column chunk: the launcher's VEC, a compile-time constant, so the
  // X-row loads carry no runtime branch between the float4 and the per-element form
  const bool vec = VEC;
  const uint32_t ldb = (uint32_t)d.ldx * 4u;  // bytes per X row
  const int rr = lane / L;                // remainder: row of this lane
  const int qq = lane % L;                // remainder: slot in the row
  const int kq = qq * FPL;
  const int nvq = kvalid - kq;
  // slices of 8+ floats per lane are interleaved float4 chunks (load_row_il)
  constexpr bool IL = FPL >= 8;
  const int gbase = lane & ~(L - 1);
  float* __restrict__ ra = racc[wid];

  // Values and X rows of one batch of G*U tiles (U groups of G tiles, one tile per group).
  struct FeatBatch {
    float v[U][NV];
    float x[U][B];
    __device__ __forceinline__ void clear() {
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int j = 0; j < NV; ++j) v[u][j] = 0.f;
#pragma unroll
        for (int c = 0; c < B; ++c) x[u][c] = 0.f;
      }
    }
  };
  // Batch starting at tile tb of the 64-tile chunk [tc, tce) whose metadata is in `meta`.
  auto load_batch = [&](FeatBatch& bt, const TileMeta& meta, int tc, int tb, int tce) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool tv = tb + u * G + g < tce;
      // A slot past the chunk's last tile takes that tile's metadata, with its values kept
      // at 0. Its X loads then need no slot predicate, which the compiler would otherwise
      // turn into a divergent branch around every load.
      const int src = min(tb - tc + u * G + g, tce - tc - 1) & (kWave - 1);
      const uint32_t clo = __shfl(meta.clo, src, kWave);
      const uint32_t chi = __shfl(meta.chi, src, kWave);
      const int c0 = __shfl(meta.c0, src, kWave);
      const int off = __shfl(meta.off, src, kWave);
      if constexpr (kSplit) {
        const uint32_t w = whi ? chi : clo;
        const int base = off + (whi ? (int)__popc(clo) : 0);
        uint32_t dv[NL];
#pragma unroll
        for (int j = 0; j < NL; ++j) {
          const bool has = tv && (w & sbit[j]) != 0u;
          dv[j] = ld.data.u32(base + (int)__popc(w & smask[j]), has);
        }
        // permlane16_swap(a, a) = (row 2k of a in rows 2k, 2k+1; row 2k+1 of a in both);
        // permlane32_swap(a, a) = (rows 0-1 of a in both halves; rows 2-3 in both)
        if constexpr (R == 2) {
#pragma unroll
          for (int j = 0; j < NL; ++j) {
            const auto pr = __builtin_amdgcn_permlane16_swap(dv[j], dv[j], false, false);
            bt.v[u][j] = __builtin_bit_cast(float, (uint32_t)pr[0]);
            bt.v[u][NL + j] = __builtin_bit_cast(float, (uint32_t)pr[1]);
          }
        } else {
          const auto p16 = __builtin_amdgcn_permlane16_swap(dv[0], dv[0], false, false);
          const auto qa = __builtin_amdgcn_permlane32_swap((uint32_t)p16[0], (uint32_t)p16[0],
                                                           false, false);
          const auto qb = __builtin_amdgcn_permlane32_swap((uint32_t)p16[1], (uint32_t)p16[1],
                                                           false, false);
          bt.v[u][0] = __builtin_bit_cast(float, (uint32_t)qa[0]);
          bt.v[u][1] = __builtin_bit_cast(float, (uint32_t)qb[0]);
          bt.v[u][2] = __builtin_bit_cast(float, (uint32_t)qa[1]);
          bt.v[u][3] = __builtin_bit_cast(float, (uint32_t)qb[1]);
        }
      } else {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          const int p = 16 * j + o;
          const bool has = tv && p < P && code_bit(clo, chi, p);
          bt.v[u][j] = ld.data.f32(off + code_prefix(clo, chi, p), has);
        }
      }
      const int64_t xb = xrow<BUF>(c0, d.ldx) + f;
#pragma unroll
      for (int c = 0; c < B; ++c) {
        const bool okc = fok && (BUF || (int64_t)c0 + c < m.cols);
        bt.x[u][c] = ld.x.f32(xb, okc, (uint32_t)c * ldb);
      }
    }
  };

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
  TileMeta mt{0u, 0u, 0, 0};
  if (TILES) mt = load_tile_meta(ld, ta0, ta1, lane);
  RowSpan rs{0, 0};
  if constexpr (REM) rs = load_row_span(ld, m, s * B + rr, s < sc.end);

  for (; s < sc.end; s += sc.stride) {
    const int64_t r0 = s * B;
    const int64_t sn = s + sc.stride;
    const int64_t snn = sn + sc.stride;
    epi.begin_stripe(s);
#pragma unroll
    for (int i = 0; i < B / G; ++i)
      epi.template prefetch<BUF>(i, r0 + i * G + g, colk, fok && r0 + i * G + g < m.rows);
    // ---- prefetch for the next stripe
    TileMeta mtn{0u, 0u, 0, 0};
    int tc0 = 0, tc1 = 0;
    RowSpan rsn{0, 0};
    if (TILES) mtn = load_tile_meta(ld, tb0, tb1, lane);
    if (TILES && snn < sc.end) {
      tc0 = sload(m.stripe_ptr, snn);
      tc1 = sload(m.stripe_ptr, snn + 1);
    }
    if constexpr (REM) rsn = load_row_span(ld, m, sn * B + rr, sn < sc.end);
    // ---- first remainder entries of this stripe
    int scol = 0;
    float sval = 0.f;
    if constexpr (REM) {
      const bool ok = qq < rs.n();
      scol = (int)ld.rcol.u32(rs.j0 + qq, ok);
      sval = ld.rval.f32(rs.j0 + qq, ok);
    }

    float acc[B];
#pragma unroll
    for (int r = 0; r < B; ++r) acc[r] = 0.f;

    if constexpr (TILES) {
      // Chunks of 64 tiles. The first uses the metadata prefetched a stripe ago; later ones
      // (stripes of more than 64 tiles) load their own, in a loop of their own, so the
      // first chunk never waits on a metadata load of this stripe.
      auto chunk = [&](const TileMeta& meta, int tc) {
        const int tce = min(ta1, tc + kWave);
        for (int tb = tc; tb < tce; tb += G * U) {
          FeatBatch P1;
          load_batch(P1, meta, tc, tb, tce);
#pragma unroll
          for (int u = 0; u < U; ++u) TileFma<B>::run(acc, P1.v[u], P1.x[u]);
        }
      };
      if (ta0 < ta1) chunk(mt, ta0);
      for (int tc = ta0 + kWave; tc < ta1; tc += kWave)
        chunk(load_tile_meta(ld, tc, ta1, lane), tc);
    }

    if constexpr (REM) {
      // row lanes: the L lanes of row rr hold FPL columns each and walk that row's entries
      float racc2[FPL];
#pragma unroll
      for (int i = 0; i < FPL; ++i) racc2[i] = 0.f;
      const int nmax = wave_max(rs.n());
      for (int jb = 0; jb < nmax; jb += L) {
        if (jb > 0) {
          const bool ok = jb + qq < rs.n();
          scol = (int)ld.rcol.u32(rs.j0 + jb + qq, ok);
          sval = ld.rval.f32(rs.j0 + jb + qq, ok);
        }
        for (int ib = 0; ib < L && jb + ib < nmax; ib += EB) {
          float xv[EB][FPL];
          float a[EB];
#pragma unroll
          for (int i = 0; i < EB; ++i) {
            const int c = __shfl(scol, gbase + ib + i, kWave);
            a[i] = __shfl(sval, gbase + ib + i, kWave);
            const bool ok = jb + ib + i < rs.n();
            if constexpr (IL)
              load_row_il<FPL, L>(ld.xr, xrow<BUF>(c, d.ldx), qq, vec, ok ? kvalid : 0, xv[i]);
            else
              load_row<FPL>(ld.xr, xrow<BUF>(c, d.ldx) + kq, vec, ok ? nvq : 0, xv[i]);
          }
#pragma unroll
          for (int i = 0; i < EB; ++i)
#pragma unroll
            for (int u = 0; u < FPL; ++u) racc2[u] = __builtin_fmaf(a[i], xv[i][u], racc2[u]);
        }
      }
      // transpose through LDS: (row rr, this lane's columns) -> (column f, all rows)
      if constexpr (IL) {
#pragma unroll
        for (int u = 0; u < FPL; ++u)
          ra[rr * KW + (u >> 2) * (4 * L) + 4 * qq + (u & 3)] = racc2[u];
      } else if (kq < KW) {
#pragma unroll
        for (int u = 0; u < FPL; ++u)
          if (kq + u < KW) ra[rr * KW + kq + u] = racc2[u];
      }
      wave_lds_sync();
      if (g == 0) {
#pragma unroll
        for (int r = 0; r < B; ++r) acc[r] += ra[r * KW + f];
      }
      wave_lds_sync();
    }

    // combine the tile groups: every group ends with the full row sums
    if constexpr (G >= 2) {
#pragma unroll
      for (int r = 0; r < B; ++r) acc[r] = xor_sum<32>(acc[r]);
    }
    if constexpr (G >= 4) {
#pragma unroll
      for (int r = 0; r < B; ++r) acc[r] = xor_sum<16>(acc[r]);
    }

    if (fok) {
#pragma unroll
      for (int r = 0; r < B; ++r) {
        const int64_t row = r0 + r;
        if ((r % G) == g && row < m.rows) epi.template put<BUF>(r / G, row, colk, acc[r]);
      }
    }
    // ---- rotate the pipeline
    ta0 = tb0;
    ta1 = tb1;
    tb0 = tc0;
    tb1 = tc1;
    mt = mtn;
    rs = rsn;
  }
  epi.finish(fok ? f : -1);
}

// --------------------------------------------------------------------------------------
// Bit-lane kernel, KW in {1, 2, 4, 8}.
// --------------------------------------------------------------------------------------
// UB: tile groups whose loads are issued together (the launcher picks 8 for 8x8 matrices
// with about 8 tiles per stripe, where one batch then covers a stripe, else 4; 2 for
// smaller blocks). VEC as for the feature-lane kernel.
template <int B, int KW, bool TILES, bool REM, bool BUF, class Epi, int UB, bool VEC>
__global__ void __launch_bounds__(256) k_stripe_bit(PbrView m, RhsView d, Epi epi) {
  constexpr int P = B * B;
  constexpr int G = kWave / P;      // tiles per wave
  constexpr int U = UB;
  constexpr int L = kWave / B;      // lanes per matrix row
  // Densely tiled 8x8 matrices (UB == 8, whole-stripe batches) at one RHS column: lane
  // (row r, slot q) walks the 8 bits of row r of tile q of each 8-tile pass, instead of one
  // lane per tile bit. Measured on MI355X (1M x 1M, 8 tiles per stripe): tiles 76 vs 93 us,
  // the fused SpMM 113.5 vs 128.5 us. On sparse tilings most of the 8 slots would stay empty.
  constexpr bool ROWLANE = B == 8 && UB == 8 && KW == 1;

  if (!epi.active()) return;
  const PbrLd<BUF> ld(m, d);
  const int lane = threadIdx.x & (kWave - 1);
  const int g = lane / P;
  const int p = lane % P;
  const int rl = p / B;             // row of this lane inside the stripe
  const int cl = p % B;             // column inside the tile
  const int q = g * B + cl;         // slot among the L lanes of row rl
  // row-lane tiles: code bits of the rows above row rl (value offset of its first bit)
  const uint32_t rmask_lo = rl >= 4 ? 0xFFFFFFFFu : ((1u << (8 * rl)) - 1u);
  const uint32_t rmask_hi = rl >= 4 ? ((1u << (8 * (rl - 4))) - 1u) : 0u;
  const int kwv = (int)min<int64_t>(KW, d.K - d.k0);  // valid columns of this launch
  const bool vec = VEC;  // float4 rows, full column chunk (compile-time, as above)
  epi.lane_init(d.k0 + q, q < kwv);

  // Value and X row of this lane's bit in each of U tiles.
  struct BitBatch {
    float v[U];
    float x[U][KW];
    __device__ __forceinline__ void clear() {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        v[u] = 0.f;
#pragma unroll
        for (int k = 0; k < KW; ++k) x[u][k] = 0.f;
      }
    }
  };
  // Batch starting at tile tb of the 64-tile chunk [tc, tce) whose metadata is in `meta`.
  auto load_batch = [&](BitBatch& bt, const TileMeta& meta, int tc, int tb, int tce) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int src = (tb - tc + u * G + g) & (kWave - 1);
      const bool tv = tb + u * G + g < tce;
      const uint32_t clo = __shfl(meta.clo, src, kWave);
      const uint32_t chi = __shfl(meta.chi, src, kWave);
      const int c0 = __shfl(meta.c0, src, kWave);
      const int off = __shfl(meta.off, src, kWave);
      const int64_t col = (int64_t)c0 + cl;
      const bool has = tv && code_bit(clo, chi, p) && (BUF || col < m.cols);
      bt.v[u] = ld.data.f32(off + code_prefix(clo, chi, p), has);
      load_row<KW>(ld.x, xrow<BUF>(col, d.ldx), vec, has ? kwv : 0, bt.x[u]);
    }
  };

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
  TileMeta mt{0u, 0u, 0, 0};
  if (TILES) mt = load_tile_meta(ld, ta0, ta1, lane);
  RowSpan rs{0, 0};
  if constexpr (REM) rs = load_row_span<KW == 1>(ld, m, s * B + rl, s < sc.end);

  for (; s < sc.end; s += sc.stride) {
    const int64_t r0 = s * B;
    const int64_t sn = s + sc.stride;
    const int64_t snn = sn + sc.stride;
    epi.begin_stripe(s);
    epi.template prefetch<BUF>(0, r0 + rl, d.k0 + q, q < kwv && r0 + rl < m.rows);
    TileMeta mtn{0u, 0u, 0, 0};
    int tc0 = 0, tc1 = 0;
    RowSpan rsn{0, 0};
    if (TILES) mtn = load_tile_meta(ld, tb0, tb1, lane);
    if (TILES && snn < sc.end) {
      tc0 = sload(m.stripe_ptr, snn);
      tc1 = sload(m.stripe_ptr, snn + 1);
    }
    if constexpr (REM) rsn = load_row_span<KW == 1>(ld, m, sn * B + rl, sn < sc.end);
    // first remainder entry of this lane's row slot
    int scol = 0;
    float sval = 0.f;
    if constexpr (REM) {
      const bool ok = q < rs.n();
      scol = (int)ld.rcol.u32(rs.j0 + q, ok);
      sval = ld.rval.f32(rs.j0 + q, ok);
    }

    float acc[KW];
#pragma unroll
    for (int k = 0; k < KW; ++k) acc[k] = 0.f;

    if constexpr (TILES && ROWLANE) {
      // lane (rl, q): row rl of tile q of each pass of 8 tiles. The row's 8 bits are byte rl
      // of the code, its values start after the set bits of the rows above it.
      const bool xwide = BUF && d.ldx == 1;  // X column 0 holds the 8 inputs contiguously
      // The chunk loop keeps its in-loop metadata reload and the per-lane float4 test: for
      // this kernel, peeling the first chunk and hoisting the test (as the other paths do)
      // measured 7 % slower.
      for (int tc = ta0; tc < ta1; tc += kWave) {
        if (tc != ta0) mt = load_tile_meta(ld, tc, ta1, lane);
        const int tce = min(ta1, tc + kWave);
        for (int tb = tc; tb < tce; tb += B) {
          const int src = (tb - tc + q) & (kWave - 1);
          const bool tv = tb + q < tce;
          const uint32_t clo = __shfl(mt.clo, src, kWave);
          const uint32_t chi = __shfl(mt.chi, src, kWave);
          const int c0 = __shfl(mt.c0, src, kWave);
          const int off = __shfl(mt.off, src, kWave);
          const uint32_t bits = tv ? ((rl < 4 ? clo >> (8 * rl) : chi >> (8 * (rl - 4))) & 0xFFu)
                                   : 0u;
          const int base = off + __popc(clo & rmask_lo) + __popc(chi & rmask_hi);
          float xv[B];
          if (xwide && c0 + B <= m.cols) {
            const float4 a = ld.x.f32x4(c0, tv);
            const float4 b = ld.x.f32x4(c0 + 4, tv);
            xv[0] = a.x; xv[1] = a.y; xv[2] = a.z; xv[3] = a.w;
            xv[4] = b.x; xv[5] = b.y; xv[6] = b.z; xv[7] = b.w;
          } else {
#pragma unroll
            for (int c = 0; c < B; ++c)
              xv[c] = ld.x.f32(xrow<BUF>((int64_t)c0 + c, d.ldx),
                               tv && (BUF || (int64_t)c0 + c < m.cols));
          }
          float vv[B];
#pragma unroll
          for (int c = 0; c < B; ++c)
            vv[c] = ld.data.f32(base + __popc(bits & ((1u << c) - 1u)), (bits >> c) & 1u);
#pragma unroll
          for (int c = 0; c < B; ++c) acc[0] = __builtin_fmaf(vv[c], xv[c], acc[0]);
        }
      }
    } else if constexpr (TILES) {
      // chunks of 64 tiles, the first with the prefetched metadata (as in the feature kernel)
      auto chunk = [&](const TileMeta& meta, int tc) {
        const int tce = min(ta1, tc + kWave);
        for (int tb = tc; tb < tce; tb += G * U) {
          BitBatch P1;
          load_batch(P1, meta, tc, tb, tce);
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < KW; ++k) acc[k] = __builtin_fmaf(P1.v[u], P1.x[u][k], acc[k]);
        }
      };
      if (ta0 < ta1) chunk(mt, ta0);
      for (int tc = ta0 + kWave; tc < ta1; tc += kWave)
        chunk(load_tile_meta(ld, tc, ta1, lane), tc);
    }

    if constexpr (REM) {
      // lane (rl, q) takes entries q, q+L, ... of its own row: products go straight into
      // the lane's accumulators and are summed with the tile products below
      for (int jb = 0; __any(jb < rs.n()); jb += L) {
        const bool ok = jb + q < rs.n();
        float xv[KW];
        if (jb > 0) {
          scol = (int)ld.rcol.u32(rs.j0 + jb + q, ok);
          sval = ld.rval.f32(rs.j0 + jb + q, ok);
        }
        load_row<KW>(ld.xr, xrow<BUF>(scol, d.ldx), vec, ok ? kwv : 0, xv);
#pragma unroll
        for (int k = 0; k < KW; ++k) acc[k] = __builtin_fmaf(sval, xv[k], acc[k]);
      }
    }

    // sum the lanes of each row: bits of cl (1 .. B/2) and of g (P .. 32)
#pragma unroll
    for (int k = 0; k < KW; ++k) {
      float t = acc[k];
      if constexpr (B == 8) t = sum8(t);
      if constexpr (B == 4) t = xor_sum<2>(xor_sum<1>(t));
      if constexpr (B == 2) t = xor_sum<1>(t);
      if constexpr (P <= 4) t = xor_sum<4>(t);
      if constexpr (P <= 8) t = xor_sum<8>(t);
      if constexpr (P <= 16) t = xor_sum<16>(t);
      if constexpr (P <= 32) t = xor_sum<32>(t);
      acc[k] = t;
    }

    const int64_t row = r0 + rl;
    if (q < kwv && row < m.rows) {
      float mine = acc[0];
#pragma unroll
      for (int k = 1; k < KW; ++k)
        if (q == k) mine = acc[k];
      epi.template put<BUF>(0, row, d.k0 + q, mine);
    }
    ta0 = tb0;
    ta1 = tb1;
    tb0 = tc0;
    tb1 = tc1;
    mt = mtn;
    rs = rsn;
  }
  epi.finish(q < kwv ? q : -1);
}

}  // namespace gl
