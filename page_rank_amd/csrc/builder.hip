// GPU PBR builder: canonical CSR -> PBR tiles + CSR remainder,
// bit-exact with the reference builder (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:340-477).
//
// No sort: a group of GL lanes owns one stripe (lane r <-> matrix row r0 + r). Each row's
// entries are sorted by column, hence by block column; the group repeatedly takes the
// minimum block column b over the rows' cursors (a k-way merge), every row whose cursor sits
// on b consumes its run of entries in block b, the group ORs the runs' bits into the tile
// code, and the tile is kept when popcount(code) >= min_nnz_per_block. Kept tiles write
// their code, coords and values (row r's run lands at popcount(code & bits-below-row-r));
// rejected tiles append their entries to the remainder rows, which therefore come out
// sorted by column. Pass 1 counts (tiles and values per stripe, remainder entries per row,
// distinct blocks per stripe), exclusive scans turn counts into offsets, pass 2 fills.
//
// Tile order inside a stripe is canonical (block columns ascending). The reference emits
// them in std::unordered_map iteration order instead (pbr_matrix.cpp:402); pass 2 can also
// list every distinct block (block column, first row, kept tile or -1) so that the host can
// replay the reference's container operations and gl_pbr_permute_tiles can reorder.
#include <algorithm>
#include <atomic>
#include <type_traits>
#include <cmath>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gl_common.hpp"

namespace gl {

struct BuildArgs {
  int64_t rows, cols;
  int32_t br, bc, min_nnz;
  int64_t num_stripes;
  const int32_t* __restrict__ indptr;
  const int32_t* __restrict__ indices;
};

template <int GL>
__device__ __forceinline__ int32_t group_min(int32_t v) {
#pragma unroll
  for (int m = 1; m < GL; m <<= 1) v = min(v, __shfl_xor(v, m, kWave));
  return v;
}
template <int GL>
__device__ __forceinline__ uint64_t group_or(uint64_t v) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
  for (int m = 1; m < GL; m <<= 1) {
    lo |= __shfl_xor(lo, m, kWave);
    hi |= __shfl_xor(hi, m, kWave);
  }
  return ((uint64_t)hi << 32) | lo;
}
template <int GL>
__device__ __forceinline__ int32_t group_sum(int32_t v) {
#pragma unroll
  for (int m = 1; m < GL; m <<= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

struct FillOut {
  // counts (pass 1) / offsets (pass 2)
  int32_t* nt;        // [ns] tiles per stripe            | tile_base [ns+1]
  int32_t* nv;        // [ns] tile values per stripe      | val_base  [ns+1]
  int32_t* nr;        // [rows] remainder entries per row | rem_ptr   [rows+1]
  int32_t* nd;        // [ns] distinct blocks per stripe  | blk_base  [ns+1]
  // pass-2 outputs
  uint64_t* codes;
  int32_t* coords;
  int32_t* offsets;
  void* data;
  int32_t* rem_col;
  void* rem_val;
  const void* values;
  int32_t* blk_bc;    // optional block list for the reference-order replay
  int8_t* blk_rmin;
  int32_t* blk_tile;
  int32_t* big;       // stripes left to the merge kernel (more distinct blocks than the
                      // LDS table holds): flags [ns] in the count pass, their exclusive
                      // prefix sums [ns+1] in the fill pass; NULL = the merge kernel takes
                      // every stripe
  int32_t* stripe_ptr;  // optional: first tile of each stripe [ns+1]
  // count pass -> fill pass, indexed by CSR entry
  uint64_t* wcode;    // codes of a stripe's distinct blocks, from its first entry on
  int32_t* wkey;      // their block columns
  uint16_t* winfo;    // per entry: position of its block in the list | bit << 8
};

template <int GL, bool FILL, typename V>
__global__ void __launch_bounds__(256) k_build(BuildArgs a, FillOut o) {
  constexpr int SPW = kWave / GL;  // stripes per wave
  constexpr int32_t kNone = INT32_MAX;
  const int lane = threadIdx.x & (kWave - 1);
  const int r = lane % GL;
  const int gidx = lane / GL;
  // uniform across the wave: scalar registers
  const int64_t wave = __builtin_amdgcn_readfirstlane(
      (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kWave));
  const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / kWave;
  const V* __restrict__ vals = (const V*)o.values;

  for (int64_t sb = wave * SPW; sb < a.num_stripes; sb += nwaves * SPW) {
    const int64_t s = sb + gidx;
    const int64_t row = s * a.br + r;
    // stripes the hash kernels left: flags in the count pass, their exclusive prefix sums
    // (scanned with the counts) in the fill pass
    const bool mine =
        s < a.num_stripes &&
        (o.big == nullptr || (FILL ? o.big[s + 1] != o.big[s] : o.big[s] != 0));
    const bool lane_ok = mine && r < a.br && row < a.rows;
    int32_t p = 0, pe = 0;
    if (lane_ok) {
      p = a.indptr[row];
      pe = a.indptr[row + 1];
    }
    // pass 2 running offsets (group-uniform values kept in every lane of the group)
    int32_t t_next = 0, v_next = 0, d_next = 0, rem_next = 0;
    if (FILL && mine) {
      t_next = o.nt[s];
      v_next = o.nv[s];
      d_next = o.nd[s];
      if (lane_ok) rem_next = o.nr[row];
    }
    int32_t cnt_t = 0, cnt_v = 0, cnt_d = 0, cnt_r = 0;
    const int32_t rshift = r * a.bc;
    while (__any(p < pe)) {
      const int32_t mybc = p < pe ? a.indices[p] / a.bc : kNone;
      const int32_t b = group_min<GL>(mybc);
      const bool part = (b != kNone) && (mybc == b);
      int32_t q = p;
      uint64_t bits = 0;
      if (part) {
        const int32_t lim = (b + 1) * a.bc;  // first column past block b
        while (q < pe) {
          const int32_t c = a.indices[q];
          if (c >= lim) break;
          bits |= 1ull << (uint64_t)(rshift + (c - b * a.bc));
          ++q;
        }
      }
      const uint64_t code = group_or<GL>(bits);
      const int cnt = __popcll(code);
      const bool kept = cnt >= a.min_nnz;
      if (b != kNone) {
        if (!FILL) {
          cnt_d += 1;
          if (kept) {
            cnt_t += 1;
            cnt_v += cnt;
          } else if (part) {
            cnt_r += q - p;
          }
        } else {
          if (kept) {
            const int32_t t = t_next++;
            const int32_t off = v_next;
            v_next += cnt;
            if (r == 0) {
              o.codes[t] = code;
              o.coords[2 * (int64_t)t] = (int32_t)(s * a.br);
              o.coords[2 * (int64_t)t + 1] = b * a.bc;
              o.offsets[t] = off;
            }
            if (part) {
              const uint64_t below = rshift >= 64 ? ~0ull : ((1ull << (uint64_t)rshift) - 1ull);
              int32_t dst = off + __popcll(code & below);
              V* out = (V*)o.data;
              for (int32_t j = p; j < q; ++j) out[dst++] = vals[j];
            }
          } else if (part) {
            V* out = (V*)o.rem_val;
            for (int32_t j = p; j < q; ++j) {
              o.rem_col[rem_next] = a.indices[j];
              out[rem_next] = vals[j];
              ++rem_next;
            }
          }
          if (o.blk_bc != nullptr) {
            const int32_t rmin = group_min<GL>(part ? r : kNone);
            const int32_t d = d_next++;
            if (r == 0) {
              o.blk_bc[d] = b;
              o.blk_rmin[d] = (int8_t)rmin;
              o.blk_tile[d] = kept ? (t_next - 1) : -1;
            }
          }
        }
      }
      if (part) p = q;
    }
    if (!FILL && mine) {
      if (r == 0) {
        o.nt[s] = cnt_t;
        o.nv[s] = cnt_v;
        o.nd[s] = cnt_d;
      }
      if (lane_ok) o.nr[row] = cnt_r;
    }
  }
}

// ---------------------------------------------------------------------------------------
// Wave-per-stripe builder, two passes that hash each entry once.
//
// Count pass: a wave reads its stripe's CSR segment 64 entries at a time (coalesced) and
// inserts every entry's block column into an LDS hash table (compare-and-swap on the key,
// 64-bit OR of the entry's bit into the code). It then lists the occupied blocks (in table
// order, j = 0..D-1) into workspace arrays indexed by the stripe's first entry E0 (a stripe
// never has more distinct blocks than entries, so the list fits in [E0, E0 + D)), writes
// every entry's block index and bit as one 16-bit word, and counts tiles and tile values per
// stripe and remainder entries per row (a row's entries are a lane range of each chunk).
//
// Fill pass (after the scans): the wave reads the stripe's block list back, ranks the kept
// blocks by block column (canonical tile order), gives each its tile index and value offset
// by prefix sums, and places every entry from its 16-bit word alone: a kept entry at
// value_offset + popcount(code & bits below its bit), a rejected one at the stripe's
// remainder offset + its rank among the stripe's rejected entries. Remainder rows are stored
// one after the other and each keeps CSR (column) order, so that rank is the entry's
// position in its remainder row plus the remainder entries of the rows above it; its column
// is block * block_cols + bit % block_cols. The fill pass reads no column index and does no
// hashing. Stripes with more than kHashCap distinct blocks are flagged by the count pass and
// left to the merge kernel in both passes. Results equal the merge kernel's entry for entry.
// ---------------------------------------------------------------------------------------
constexpr int kHashCap = 256;
// Grid of the count pass, in workgroups per CU (its LDS allows 9).
constexpr int kHashCountWgPerCu = 8;
// Chunks of 64 entries per stripe whose loads leave together at the top of the stripe and
// stay in registers for the stripe's sweeps. Longer stripes load the rest chunk by chunk.
constexpr int kHashPre = 4;
// stripes of at most kHashCap entries are loaded whole by the prefetched chunks
static_assert(kHashPre * kWave == kHashCap, "prefetched chunks must cover kHashCap entries");

struct CountLds {
  unsigned long long code[kHashCap];
  int32_t key[kHashCap];
  int32_t jof[kHashCap];  // slot -> list index j, | 256 when the block is rejected
  int32_t rows[65];       // CSR start of each row of the stripe (block_rows + 1 values)
  int32_t flag;           // table overflow
};

// Indexed by the block's position j in the stripe's block list.
struct FillLds {
  unsigned long long code[kHashCap];  // block code
  int32_t key[kHashCap];              // block column
  union {
    int32_t lkey[kHashCap];           // listed blocks: block columns (until ranked)
    int32_t voff[kHashCap];           // then a block's first value offset when kept, else -1
  };
  uint8_t lidx[kHashCap];             // listed blocks: j, then j by block column
};

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int hash_slot(int32_t b) {
  return (int)(((uint32_t)b * 2654435761u) >> 24);  // kHashCap = 256
}
// bits [0, n) of a 64-bit mask, n in [0, 64]
__device__ __forceinline__ uint64_t low_bits(int n) {
  return n >= 64 ? ~0ull : ((1ull << (uint64_t)n) - 1ull);
}
// Row (inside the stripe) of entry e: the last k < br with rows[k] <= e.
__device__ __forceinline__ int row_of(const int32_t* rows, int br, int32_t e) {
  int lo = 0, hi = br - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rows[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// P2: block_cols is a power of two, so an entry's block column and bit come from a shift and
// a mask instead of an integer division (both passes are bound by VALU issue).
template <bool P2>
__global__ void __launch_bounds__(256, kHashCountWgPerCu) k_hash_count(BuildArgs a, FillOut o) {
  __shared__ CountLds sh_all[4];
  const int lane = threadIdx.x & (kWave - 1);
  CountLds& h = sh_all[threadIdx.x / kWave];
  // uniform across the wave: scalar registers
  // stripe indices in 32 bits (fewer scalar registers): stripes < 2^31, and unsigned sums of
  // a stripe and two grid strides stay below 2^32
  const uint32_t wave = __builtin_amdgcn_readfirstlane(
      (blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const uint32_t num_stripes = (uint32_t)a.num_stripes;
  const int br = a.br, bc = a.bc;
  const int bc_shift = P2 ? __builtin_ctz((unsigned)bc) : 0;
  const uint64_t below_lane = low_bits(lane);
  auto block_of = [&](int32_t col) { return P2 ? (col >> bc_shift) : col / bc; };
  auto bit_of = [&](int r, int32_t col, int32_t b) {
    return P2 ? ((r << bc_shift) | (col & (bc - 1))) : r * bc + (col - b * bc);
  };
  // CSR row starts of a stripe: row k in lane k, row 64 in lane 0's second register
  auto load_rows = [&](uint32_t st, int32_t& rp0, int32_t& rp1) {
    rp0 = rp1 = 0;
    if (st >= num_stripes) return;
    const int64_t rb = (int64_t)st * br;
    if (lane <= br) rp0 = a.indptr[min<int64_t>(rb + lane, a.rows)];
    if (lane == 0 && br >= kWave) rp1 = a.indptr[min<int64_t>(rb + kWave, a.rows)];
  };
  auto end_of = [&](int32_t rp0, int32_t rp1) {
    return br >= kWave ? __builtin_amdgcn_readfirstlane(rp1) : __builtin_amdgcn_readlane(rp0, br);
  };
  auto load_cols = [&](int32_t e0, int32_t e1, int32_t* col) {
#pragma unroll
    for (int k = 0; k < kHashPre; ++k) {
      const int32_t e = e0 + k * kWave + lane;
      col[k] = e < e1 ? a.indices[e] : 0;
    }
  };
  // Software pipeline: a stripe's row starts are loaded two stripes ahead and its first
  // column chunks one stripe ahead, both before the previous stripe's stores, so the wait
  // for a stripe's columns overlaps the previous stripe's stores instead of following them.
  int32_t rp0, rp1, nrp0, nrp1;
  load_rows(wave, rp0, rp1);
  int32_t E0 = __builtin_amdgcn_readfirstlane(rp0), E1 = end_of(rp0, rp1);
  int32_t pcol[kHashPre];
  load_cols(E0, E1, pcol);
  load_rows(wave + nwaves, nrp0, nrp1);

  for (uint32_t s = wave; s < num_stripes; s += nwaves) {
    const int64_t r0 = (int64_t)s * br;
    if (lane <= br) h.rows[lane] = rp0;
    if (lane == 0 && br >= kWave) h.rows[kWave] = rp1;
    for (int i = lane; i < kHashCap; i += kWave) {
      h.key[i] = -1;
      h.code[i] = 0ull;
    }
    if (lane == 0) h.flag = 0;
    // row starts 1..7 as scalars for the common block_rows <= 8 (a row is then found by
    // compares instead of a search in LDS)
    int32_t rs8[7];
#pragma unroll
    for (int k = 0; k < 7; ++k)
      rs8[k] = k + 1 < br ? __builtin_amdgcn_readlane(rp0, k + 1) : INT32_MAX;
    // lane k < br: row k covers entries [rk, rk1)
    // (the 65th row start sits in lane 0's second register; read it with the whole wave
    // active)
    const int32_t rk = rp0;
    const int32_t row64 = __builtin_amdgcn_readfirstlane(rp1);
    const int32_t next_start = __shfl_down(rp0, 1, kWave);
    const int32_t rk1 = (br >= kWave && lane == kWave - 1) ? row64 : next_start;
    wsync();
    // Row of each lane's entry in the chunk of 64 entries from c: compares against the
    // scalar row starts for block_rows <= 8 (a scalar mask of row-start lanes counted with
    // mbcnt took more scalar instructions than it saved), a search in LDS for taller blocks.
    auto chunk_row = [&](int32_t c) {
      const int32_t e = c + lane;
      if (br <= 8) {
        int r = 0;
#pragma unroll
        for (int k = 0; k < 7; ++k) r += rs8[k] <= e ? 1 : 0;
        return r;
      }
      return row_of(h.rows, br, e);
    };
    // Returns the entry's slot | bit << 8. CHECK: the table may fill (more than kHashCap
    // entries in the stripe); -1 and the overflow flag then. With at most kHashCap entries a
    // full table holds every key of the stripe, so the probe always ends.
    auto insert = [&](auto check_tag, int r, int32_t col) -> int32_t {
      constexpr bool CHECK = decltype(check_tag)::value;
      const int32_t b = block_of(col);
      const int bit = bit_of(r, col, b);
      int slot = hash_slot(b);
      int probes = 0;
      for (;;) {
        const int32_t old = atomicCAS(&h.key[slot], -1, b);
        if (old == -1 || old == b) break;
        slot = (slot + 1) & (kHashCap - 1);
        if (CHECK && ++probes == kHashCap) {
          slot = -1;
          break;
        }
      }
      if (CHECK && slot < 0) {
        h.flag = 1;
        return -1;
      }
      atomicOr(&h.code[slot], 1ull << (uint64_t)bit);
      return slot | (bit << 8);
    };
    int32_t pinf[kHashPre];
    if (E1 - E0 <= kHashCap) {
#pragma unroll
      for (int k = 0; k < kHashPre; ++k) {
        const int32_t c = E0 + k * kWave;
        const int r = chunk_row(c);
        pinf[k] = c + lane < E1 ? insert(std::false_type{}, r, pcol[k]) : -1;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kHashPre; ++k) {
        const int32_t c = E0 + k * kWave;
        const int r = chunk_row(c);
        pinf[k] = c + lane < E1 ? insert(std::true_type{}, r, pcol[k]) : -1;
      }
      for (int32_t c = E0 + kHashPre * kWave; c < E1; c += kWave) {
        const int32_t e = c + lane;
        const int r = chunk_row(c);
        if (e < E1) insert(std::true_type{}, r, a.indices[e]);
      }
    }
    // ---- next stripe: its columns and the row starts of the one after, issued before this
    // stripe's stores
    const int32_t cE0 = E0, cE1 = E1;
    rp0 = nrp0;
    rp1 = nrp1;
    if (s + nwaves < num_stripes) {
      E0 = __builtin_amdgcn_readfirstlane(rp0);
      E1 = end_of(rp0, rp1);
      load_cols(E0, E1, pcol);
      load_rows(s + 2 * nwaves, nrp0, nrp1);
    }
    wsync();
    if (h.flag) {
      if (lane == 0) o.big[s] = 1;  // the merge kernel takes this stripe
      wsync();
      continue;
    }
    if (lane == 0) o.big[s] = 0;

    // ---- list the occupied blocks (table order) and count kept tiles and their values
    int32_t D = 0, nt = 0, nv = 0;
#pragma unroll
    for (int t = 0; t < kHashCap / kWave; ++t) {
      const int i = t * kWave + lane;
      const int32_t k = h.key[i];
      const bool occ = k >= 0;
      const uint64_t code = h.code[i];
      const int32_t cnt = __popcll(code);
      const bool kept = occ && cnt >= a.min_nnz;
      const uint64_t m = __ballot(occ);
      const int32_t j = D + __popcll(m & below_lane);
      if (occ) {
        h.jof[i] = j | (kept ? 0 : 256);
        o.wkey[cE0 + j] = k;
        o.wcode[cE0 + j] = code;
      }
      D += __popcll(m);
      nt += __popcll(__ballot(kept));
      nv += kept ? cnt : 0;
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) nv += __shfl_xor(nv, m, kWave);
    if (lane == 0) {
      o.nt[s] = nt;
      o.nv[s] = nv;
      o.nd[s] = D;
    }
    wsync();

    // ---- entries: block index and bit of each, rejected entries per row
    int32_t nrow = 0;
    auto sweep = [&](int32_t c, int32_t col, int32_t inf) {
      const int32_t e = c + lane;
      const bool valid = e < cE1;
      int slot = 0, bit = 0;
      if (valid) {
        if (inf >= 0) {
          slot = inf & 255;
          bit = inf >> 8;
        } else {
          const int r = chunk_row(c);
          const int32_t b = block_of(col);
          bit = bit_of(r, col, b);
          slot = hash_slot(b);
          while (h.key[slot] != b) slot = (slot + 1) & (kHashCap - 1);
        }
      }
      const int32_t x = valid ? h.jof[slot] : 0;
      if (valid) o.winfo[e] = (uint16_t)((x & 255) | (bit << 8));
      const uint64_t rm = __ballot(valid && (x & 256) != 0);
      if (lane < br) {
        const int lo = min(max(rk - c, 0), kWave), hi = min(max(rk1 - c, 0), kWave);
        nrow += __popcll(rm & low_bits(hi) & ~low_bits(lo));
      }
    };
#pragma unroll
    for (int k = 0; k < kHashPre; ++k) {
      const int32_t c = cE0 + k * kWave;
      if (c < cE1) sweep(c, 0, pinf[k]);
    }
    for (int32_t c = cE0 + kHashPre * kWave; c < cE1; c += kWave) {
      const int32_t e = c + lane;
      sweep(c, e < cE1 ? a.indices[e] : 0, -1);
    }
    if (lane < br && r0 + lane < a.rows) o.nr[r0 + lane] = nrow;
    wsync();
  }
}

// 8 waves per SIMD (64 VGPRs), which is also what its LDS allows (8 workgroups per CU).
constexpr int kHashFillWavesPerSimd = 8;
// ALL: the reference-order block list (every distinct block) is written too.
template <typename V, bool P2, bool ALL>
__global__ void __launch_bounds__(256, kHashFillWavesPerSimd) k_hash_fill(BuildArgs a, FillOut o) {
  __shared__ FillLds sh_all[4];
  const int lane = threadIdx.x & (kWave - 1);
  FillLds& h = sh_all[threadIdx.x / kWave];
  // stripe indices in 32 bits (fewer scalar registers): stripes < 2^31, and unsigned sums of
  // a stripe and two grid strides stay below 2^32
  const uint32_t wave = __builtin_amdgcn_readfirstlane(
      (blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const uint32_t num_stripes = (uint32_t)a.num_stripes;
  const V* __restrict__ vals = (const V*)o.values;
  V* __restrict__ data = (V*)o.data;
  V* __restrict__ rem_val = (V*)o.rem_val;
  const int br = a.br, bc = a.bc;
  const int bc_shift = P2 ? __builtin_ctz((unsigned)bc) : 0;
  const uint64_t below_lane = low_bits(lane);
  constexpr bool all_blocks = ALL;
  // a stripe's scalars, scalar loads issued a stripe ahead: first and end entry, tile /
  // value / distinct-block offsets, next distinct-block offset, remainder offset, overflow
  struct Head {
    int32_t E0, E1, tile, val, d0, d1, rem, big;
  };
  auto load_head = [&](uint32_t st) {
    Head hd{0, 0, 0, 0, 0, 0, 0, 0};
    if (st < num_stripes) {
      const int64_t rb = (int64_t)st * br;
      hd.E0 = sload(a.indptr, rb);
      hd.E1 = sload(a.indptr, min<int64_t>(rb + br, a.rows));
      hd.tile = sload(o.nt, st);
      hd.val = sload(o.nv, st);
      hd.d0 = sload(o.nd, st);
      hd.d1 = sload(o.nd, st + 1);
      hd.rem = sload(o.nr, rb);
      hd.big = sload(o.big, st + 1) - sload(o.big, st);
    }
    return hd;
  };
  Head head = load_head(wave);
  // the last wave closes the offset arrays: offsets[T] = V, stripe_ptr[ns] = T
  if (wave == nwaves - 1 && lane == 0) {
    const int32_t T = o.nt[a.num_stripes];
    o.offsets[T] = o.nv[a.num_stripes];
    if (o.stripe_ptr) o.stripe_ptr[a.num_stripes] = T;
  }

  for (uint32_t s = wave; s < num_stripes; s += nwaves) {
    const int64_t r0 = (int64_t)s * br;
    const Head cur = head;
    const int32_t E0 = cur.E0, E1 = cur.E1;
    const int32_t tile_base = cur.tile, val_base = cur.val, d_base = cur.d0;
    const int32_t D = cur.d1 - cur.d0;
    const int32_t rem_base = cur.rem;
    if (lane == 0 && o.stripe_ptr) o.stripe_ptr[s] = tile_base;
    if (cur.big) {  // the merge kernel fills this stripe
      head = load_head(s + nwaves);
      continue;
    }
    // block list and the first chunks of entries, then the next stripe's scalars
    int32_t bkey[kHashCap / kWave];
    uint64_t bcode[kHashCap / kWave];
#pragma unroll
    for (int t = 0; t < kHashCap / kWave; ++t) {
      const int j = t * kWave + lane;
      bkey[t] = 0;
      bcode[t] = 0ull;
      if (t * kWave < D && j < D) {
        bkey[t] = o.wkey[E0 + j];
        bcode[t] = o.wcode[E0 + j];
      }
    }
    int32_t pinf[kHashPre];
    V pval[kHashPre];
#pragma unroll
    for (int k = 0; k < kHashPre; ++k) {
      const int32_t e = E0 + k * kWave + lane;
      pinf[k] = e < E1 ? (int32_t)o.winfo[e] : 0;
      pval[k] = e < E1 ? vals[e] : V(0);
    }
    head = load_head(s + nwaves);

    // ---- listed blocks: the kept ones, or every block when the reference-order block list
    // is written
    int32_t L = 0;
#pragma unroll
    for (int t = 0; t < kHashCap / kWave; ++t) {
      if (t * kWave >= D) break;
      const int j = t * kWave + lane;
      const bool valid = j < D;
      const bool listed = valid && (all_blocks || __popcll(bcode[t]) >= a.min_nnz);
      if (valid) {
        h.key[j] = bkey[t];
        h.code[j] = bcode[t];
      }
      const uint64_t m = __ballot(listed);
      if (listed) {
        const int pos = L + __popcll(m & below_lane);
        h.lkey[pos] = bkey[t];
        h.lidx[pos] = (uint8_t)j;
      }
      L += __popcll(m);
    }
    wsync();
    // rank of each listed block = number of smaller block columns (unique)
    if (L <= kWave) {  // one listed block per lane; the keys come by lane reads
      const int32_t mykey = lane < L ? h.lkey[lane] : INT32_MAX;
      const int32_t myj = lane < L ? h.lidx[lane] : 0;
      int32_t rank = 0;
      for (int i = 0; i < L; ++i) rank += __builtin_amdgcn_readlane(mykey, i) < mykey ? 1 : 0;
      wsync();
      if (lane < L) h.lidx[rank] = (uint8_t)myj;
    } else {  // broadcast LDS reads
      int32_t mykey[kHashCap / kWave], myj[kHashCap / kWave], rank[kHashCap / kWave];
#pragma unroll
      for (int t = 0; t < kHashCap / kWave; ++t) {
        const int q = t * kWave + lane;
        mykey[t] = q < L ? h.lkey[q] : INT32_MAX;
        myj[t] = q < L ? h.lidx[q] : 0;
        rank[t] = 0;
      }
      for (int i = 0; i < L; ++i) {
        const int32_t ki = h.lkey[i];
#pragma unroll
        for (int t = 0; t < kHashCap / kWave; ++t) rank[t] += ki < mykey[t] ? 1 : 0;
      }
      wsync();
#pragma unroll
      for (int t = 0; t < kHashCap / kWave; ++t)
        if (t * kWave + lane < L) h.lidx[rank[t]] = (uint8_t)myj[t];
    }
    wsync();
    // the ranked keys are done with: their space holds the value offsets, -1 = rejected
    for (int j = lane; j < D; j += kWave) h.voff[j] = -1;
    wsync();

    // ---- tile indices and value offsets by prefix sums in block-column order
    int32_t kcar = 0, vcar = 0;
    for (int q0 = 0; q0 < L; q0 += kWave) {
      const int q = q0 + lane;
      const bool valid = q < L;
      const int j = valid ? h.lidx[q] : 0;
      const uint64_t code = valid ? h.code[j] : 0ull;
      const int32_t key = h.key[j];
      const int32_t cnt = __popcll(code);
      const bool kept = valid && cnt >= a.min_nnz;
      const uint64_t km = __ballot(kept);
      const int32_t krank = kcar + __popcll(km & below_lane);
      int32_t vin = kept ? cnt : 0;  // inclusive scan of kept counts
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int32_t y = __shfl_up(vin, d, kWave);
        if (lane >= d) vin += y;
      }
      const int32_t vex = vcar + vin - (kept ? cnt : 0);
      if (kept) {
        const int32_t t = tile_base + krank;
        h.voff[j] = val_base + vex;
        o.codes[t] = code;
        o.coords[2 * (int64_t)t] = (int32_t)r0;
        o.coords[2 * (int64_t)t + 1] = key * bc;
        o.offsets[t] = val_base + vex;
      }
      if (all_blocks && valid) {
        o.blk_bc[d_base + q] = key;
        o.blk_rmin[d_base + q] = (int8_t)(__builtin_ctzll(code) / bc);
        o.blk_tile[d_base + q] = kept ? tile_base + krank : -1;
      }
      kcar += __popcll(km);
      vcar += __shfl(vin, kWave - 1, kWave);
    }
    wsync();

    // ---- entries: kept values into their tile, rejected ones to the remainder
    int32_t carry = 0;  // rejected entries of the stripe so far
    auto sweep = [&](int32_t c, int32_t inf, V val) {
      const int32_t e = c + lane;
      const bool valid = e < E1;
      const int j = inf & 255;
      const int bit = inf >> 8;
      const int32_t vo = valid ? h.voff[j] : 0;
      const bool rej = valid && vo < 0;
      const uint64_t rm = __ballot(rej);
      if (rej) {
        const int32_t pos = rem_base + carry + __popcll(rm & below_lane);
        const int32_t in_blk = P2 ? (bit & (bc - 1)) : bit % bc;
        o.rem_col[pos] = h.key[j] * bc + in_blk;
        rem_val[pos] = val;
      } else if (valid) {
        data[vo + __popcll(h.code[j] & low_bits(bit))] = val;
      }
      carry += __popcll(rm);
    };
#pragma unroll
    for (int k = 0; k < kHashPre; ++k) {
      const int32_t c = E0 + k * kWave;
      if (c < E1) sweep(c, pinf[k], pval[k]);
    }
    for (int32_t c = E0 + kHashPre * kWave; c < E1; c += kWave) {
      const int32_t e = c + lane;
      const bool ok = e < E1;
      sweep(c, ok ? (int32_t)o.winfo[e] : 0, ok ? vals[e] : V(0));
    }
    wsync();
  }
}

// ---- exclusive scan of int32 counts: out[0..n] (out[n] = total) ----
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ int32_t block_excl_scan(int32_t v, int32_t* sh, int32_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  if (wid == 0) {
    int32_t w = lane < (int)(blockDim.x >> 6) ? sh[lane] : 0;
    int32_t z = w;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(z, o, kWave);
      if (lane >= o) z += y;
    }
    if (lane < (int)(blockDim.x >> 6)) sh[lane] = z - w;
    if (lane == (int)(blockDim.x >> 6) - 1) sh[64] = z;
  }
  __syncthreads();
  const int32_t r = sh[wid] + x - v;
  *total = sh[64];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_tiles(const int32_t* __restrict__ in, int32_t* __restrict__ out, int64_t n,
             int32_t* __restrict__ bsum) {
  __shared__ int32_t sh[65];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < n ? in[base + i] : 0;
    s += v[i];
  }
  int32_t total;
  int32_t run = block_excl_scan(s, sh, &total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// single block: exclusive scan of the block sums in place, grand total to out[n]
__global__ void __launch_bounds__(kScanThreads)
k_scan_sums(int32_t* __restrict__ bsum, int64_t nb, int32_t* __restrict__ out, int64_t n) {
  __shared__ int32_t sh[65];
  int32_t carry = 0;
  for (int64_t c0 = 0; c0 < nb; c0 += kScanThreads) {
    const int64_t i = c0 + threadIdx.x;
    const int32_t v = i < nb ? bsum[i] : 0;
    int32_t total;
    const int32_t e = block_excl_scan(v, sh, &total);
    if (i < nb) bsum[i] = carry + e;
    carry += total;
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ void k_scan_add(int32_t* __restrict__ out, int64_t n, const int32_t* __restrict__ bsum) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] += bsum[i / kScanTile];
}

int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

// out[0..n]: exclusive scan of in[0..n-1]; in and out may alias.
int excl_scan(const int32_t* in, int32_t* out, int64_t n, int32_t* tmp, hipStream_t st) {
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb > 0) {
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, out, n,
                       tmp);
    GL_LAUNCH_CHECK("k_scan_tiles");
  }
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, tmp, nb, out, n);
  GL_LAUNCH_CHECK("k_scan_sums");
  if (nb > 1) {
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, n,
                       tmp);
    GL_LAUNCH_CHECK("k_scan_add");
  }
  return GL_OK;
}

// ---- the builder's scans (tiles, tile values and distinct blocks per stripe, remainder
// entries per row, overflowing stripes) in two launches: tile-local scans of all arrays,
// then one that adds each tile's carry (the sum of the tile sums before it, read by the tile
// itself) and writes out[n] and the totals, to device memory and to pinned host memory, so
// the host reads them after a stream synchronization without a copy.
// Arrays of more than kScanMaxTiles tiles take the three-launch scan above instead.
constexpr int kScanArrays = 5;
constexpr int kScanMaxTiles = 2048;
struct ScanSet {
  int32_t* a[kScanArrays];
  int64_t n[kScanArrays];
  int32_t* bsum[kScanArrays];
  int32_t tiles[kScanArrays];
  int32_t first[kScanArrays + 1];   // first workgroup of array i in the tile launch
  int32_t afirst[kScanArrays + 1];  // ... in the add launch (at least one per array)
  int32_t* totals;                  // device
  int32_t* host_totals;             // pinned host memory (device-visible), or NULL
};

__device__ __forceinline__ int scan_array(const int32_t* first, int32_t blk) {
  int i = 0;
  while (i < kScanArrays - 1 && blk >= first[i + 1]) ++i;
  return i;
}

__global__ void __launch_bounds__(kScanThreads) k_scanset_tiles(ScanSet p) {
  __shared__ int32_t sh[65];
  const int i = scan_array(p.first, (int32_t)blockIdx.x);
  const int64_t j = blockIdx.x - p.first[i];
  int32_t* __restrict__ a = p.a[i];
  const int64_t n = p.n[i];
  const int64_t base = j * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? a[base + k] : 0;
    s += v[k];
  }
  int32_t total;
  int32_t run = block_excl_scan(s, sh, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) a[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) p.bsum[i][j] = total;
}

__device__ __forceinline__ void put_total(const ScanSet& p, int i, int32_t total) {
  p.totals[i] = total;
  if (p.host_totals) {
    p.host_totals[i] = total;
    __threadfence_system();
  }
}

__global__ void __launch_bounds__(kScanThreads) k_scanset_add(ScanSet p) {
  __shared__ int32_t sh[65];
  const int i = scan_array(p.afirst, (int32_t)blockIdx.x);
  const int64_t j = blockIdx.x - p.afirst[i];
  int32_t* __restrict__ a = p.a[i];
  const int64_t n = p.n[i];
  const int32_t* __restrict__ bs = p.bsum[i];
  int32_t s = 0;
  for (int64_t k = threadIdx.x; k < j; k += kScanThreads) s += bs[k];
  int32_t carry;
  (void)block_excl_scan(s, sh, &carry);
  if (carry != 0) {
    const int64_t base = j * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t idx = base + (int64_t)k * kScanThreads + threadIdx.x;
      if (idx < n) a[idx] += carry;
    }
  }
  if (threadIdx.x == 0 && j + 1 >= p.tiles[i]) {
    const int32_t total = carry + (p.tiles[i] > 0 ? bs[j] : 0);
    a[n] = total;
    put_total(p, i, total);
  }
}

// totals of arrays scanned by the three-launch scan
__global__ void k_scanset_totals(ScanSet p) {
  if (threadIdx.x < kScanArrays) put_total(p, threadIdx.x, p.a[threadIdx.x][p.n[threadIdx.x]]);
}

// In-place exclusive scans of a[i][0..n[i]) with a[i][n[i]] = total, totals[i] = total (and
// host_totals[i] when not NULL). bsum: kScanMaxTiles ints per array, or scan_tmp_elems()
// for the three-launch scan.
int scan_set(int32_t* const a[kScanArrays], const int64_t n[kScanArrays], int32_t* bsum,
             int32_t* totals, int32_t* host_totals, hipStream_t st) {
  ScanSet p{};
  int64_t max_tiles = 0;
  for (int i = 0; i < kScanArrays; ++i) {
    p.a[i] = a[i];
    p.n[i] = n[i];
    max_tiles = std::max(max_tiles, (n[i] + kScanTile - 1) / kScanTile);
  }
  p.totals = totals;
  p.host_totals = host_totals;
  if (max_tiles > kScanMaxTiles) {
    for (int i = 0; i < kScanArrays; ++i) {
      int rc = excl_scan(a[i], a[i], n[i], bsum, st);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_scanset_totals, dim3(1), dim3(64), 0, st, p);
    GL_LAUNCH_CHECK("k_scanset_totals");
    return GL_OK;
  }
  p.first[0] = p.afirst[0] = 0;
  for (int i = 0; i < kScanArrays; ++i) {
    p.bsum[i] = bsum + (int64_t)i * kScanMaxTiles;
    p.tiles[i] = (int32_t)((n[i] + kScanTile - 1) / kScanTile);
    p.first[i + 1] = p.first[i] + p.tiles[i];
    p.afirst[i + 1] = p.afirst[i] + std::max(p.tiles[i], 1);
  }
  if (p.first[kScanArrays] > 0) {
    hipLaunchKernelGGL(k_scanset_tiles, dim3((unsigned)p.first[kScanArrays]),
                       dim3(kScanThreads), 0, st, p);
    GL_LAUNCH_CHECK("k_scanset_tiles");
  }
  hipLaunchKernelGGL(k_scanset_add, dim3((unsigned)p.afirst[kScanArrays]), dim3(kScanThreads),
                     0, st, p);
  GL_LAUNCH_CHECK("k_scanset_add");
  return GL_OK;
}

// ---- workspace layout ----
// Per stripe: counts / offsets (nt, nv, nd), overflow flags; scan scratch; the four totals;
// per CSR entry (count pass -> fill pass): block lists (code, block column) at the stripe's
// first entry and each entry's block index and bit.
struct BuildWs {
  int32_t *nt, *nv, *nd, *tmp, *tot;
  int32_t* big;
  uint64_t* wcode;
  int32_t* wkey;
  uint16_t* winfo;
  int64_t bytes;
};
static int64_t al(int64_t x) { return (x + 255) & ~(int64_t)255; }
static BuildWs carve_build(char* base, int64_t rows, int32_t br, int64_t nnz) {
  const int64_t ns = (rows + br - 1) / br;
  const int64_t tmpn = std::max<int64_t>(scan_tmp_elems(std::max(rows, ns) + 1),
                                         kScanArrays * (int64_t)kScanMaxTiles);
  BuildWs w;
  int64_t off = 0;
  auto take = [&](int64_t n) {
    char* p = base ? base + off : nullptr;
    off += al(n);
    return (void*)p;
  };
  w.nt = (int32_t*)take((ns + 1) * 4);
  w.nv = (int32_t*)take((ns + 1) * 4);
  w.nd = (int32_t*)take((ns + 1) * 4);
  w.tmp = (int32_t*)take(tmpn * 4);
  w.tot = (int32_t*)take(kScanArrays * 4);
  w.big = (int32_t*)take((ns + 1) * 4);
  w.wcode = (uint64_t*)take(nnz * 8);
  w.wkey = (int32_t*)take(nnz * 4);
  w.winfo = (uint16_t*)take(nnz * 2);
  w.bytes = off;
  return w;
}

static int group_lanes(int br) {
  int g = 1;
  while (g < br) g <<= 1;
  return g;
}

// Waves: the count pass uses kHashCountWgPerCu workgroups per CU, the fill pass fills the
// CUs once (its occupancy, from its LDS).
template <bool FILL, typename V, bool P2>
static void launch_build_hash(const BuildArgs& a, const FillOut& o, hipStream_t st) {
  static std::atomic<int> per_cu{-1};
  if (per_cu.load() < 0) {
    int n = kHashCountWgPerCu;
    if (FILL &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_hash_fill<V, P2, false>, 256, 0) !=
             hipSuccess ||
         n < 1))
      n = 4;
    per_cu.store(n);
  }
  const int64_t cap = (int64_t)device_info().num_cus * per_cu.load() * 4;
  const int64_t nw = std::max<int64_t>(4, std::min(a.num_stripes, cap));
  const unsigned groups = (unsigned)((nw + 3) / 4);
  if constexpr (FILL) {
    if (o.blk_bc != nullptr)
      hipLaunchKernelGGL((k_hash_fill<V, P2, true>), dim3(groups), dim3(256), 0, st, a, o);
    else
      hipLaunchKernelGGL((k_hash_fill<V, P2, false>), dim3(groups), dim3(256), 0, st, a, o);
  }
  else
    hipLaunchKernelGGL((k_hash_count<P2>), dim3(groups), dim3(256), 0, st, a, o);
}

template <bool FILL, typename V>
static int launch_build(const BuildArgs& a, const FillOut& o, hipStream_t st,
                        bool merge = true) {
  if (o.big != nullptr) {
    if ((a.bc & (a.bc - 1)) == 0) launch_build_hash<FILL, V, true>(a, o, st);
    else launch_build_hash<FILL, V, false>(a, o, st);
    GL_LAUNCH_CHECK("k_hash_count / k_hash_fill");
  }
  if (!merge) return GL_OK;
  const int GLn = group_lanes(a.br);
  const int spw = kWave / GLn;
  const int64_t waves = (a.num_stripes + spw - 1) / spw;
  const int64_t cap = (int64_t)device_info().num_cus * 32;
  const int64_t nw = std::max<int64_t>(1, std::min(waves, cap));
  const dim3 grid((unsigned)((nw + 3) / 4)), block(256);
  switch (GLn) {
#define GL_B(G) case G: hipLaunchKernelGGL((k_build<G, FILL, V>), grid, block, 0, st, a, o); break;
    GL_B(1) GL_B(2) GL_B(4) GL_B(8) GL_B(16) GL_B(32) GL_B(64)
#undef GL_B
    default:
      set_error("block_rows %d too large", a.br);
      return GL_ERR_INVALID;
  }
  GL_LAUNCH_CHECK("k_build");
  return GL_OK;
}

__global__ void k_tile_popc(const uint64_t* __restrict__ codes, const int32_t* __restrict__ perm,
                            int64_t T, int32_t* __restrict__ cnt) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < T) cnt[t] = __popcll(codes[perm[t]]);
}

template <typename V>
__global__ void k_tile_gather(const int32_t* __restrict__ perm, int64_t T,
                              const uint64_t* __restrict__ codes, const int32_t* __restrict__ coords,
                              const int32_t* __restrict__ offsets, const V* __restrict__ data,
                              const int32_t* __restrict__ new_off, uint64_t* __restrict__ codes_o,
                              int32_t* __restrict__ coords_o, V* __restrict__ data_o) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int32_t src = perm[t];
  codes_o[t] = codes[src];
  coords_o[2 * t] = coords[2 * (int64_t)src];
  coords_o[2 * t + 1] = coords[2 * (int64_t)src + 1];
  const int32_t a = offsets[src], n = offsets[src + 1] - a, d = new_off[t];
  for (int32_t i = 0; i < n; ++i) data_o[d + i] = data[a + i];
}

// Reference emission order without the container: the reference builder keeps one
// std::unordered_map<index_t, ...>(sqrt(num_col_blocks)), emplaces each stripe's blocks in
// CSR first-occurrence order, iterates the map and clear()s it (pbr_matrix.cpp:374-428).
// libstdc++'s _Hashtable (identity hash on the key converted to size_t, unique keys) gives
// that iteration a fixed order: a node entering an empty bucket goes to the front of the
// one list and becomes its bucket's "before" node for the bucket that was first; a node
// entering an occupied bucket goes right after that bucket's "before" node; a rehash relinks
// the list front to back the same way (hashtable.h, _M_insert_bucket_begin and
// _M_rehash_aux). That order is replayed here with arrays, per stripe on worker threads.
// Only the bucket count needs the library: it depends on the number of elements alone, and
// it can change only in a stripe holding more blocks than every stripe before it (clear()
// keeps the buckets and the rehash threshold), so one real map of the same type, fed
// placeholder keys for those stripes only, gives the bucket count before every insert.
// a mod d for a fixed d without a division (Lemire, Kaser and Kurz, "Faster Remainder by
// Direct Computation", 2019): M = floor((2^128 - 1) / d) + 1, a mod d = high 64 bits of
// ((M * a) mod 2^128) * d.
struct FastMod64 {
  uint64_t d = 1;
  unsigned __int128 m = 0;
  void set(uint64_t div) {
    d = div;
    m = ~(unsigned __int128)0 / div + 1;
  }
  uint64_t operator()(uint64_t a) const {
    const unsigned __int128 low = m * a;
    const unsigned __int128 bottom = ((unsigned __int128)(uint64_t)low * d) >> 64;
    const unsigned __int128 top = (unsigned __int128)(uint64_t)(low >> 64) * d;
    return (uint64_t)((bottom + top) >> 64);
  }
};

// Returns false when a stripe's kept blocks do not carry consecutive tile indices.
template <typename I>
static bool replay(int64_t ns, int64_t cols, int32_t br, int32_t bc, const int32_t* blk_ptr,
                   const int32_t* blk_bc, const int8_t* blk_rmin, const int32_t* blk_tile,
                   int32_t* perm) {
  const uint64_t ncb = ((uint64_t)cols + bc - 1) / bc;
  if (ns <= 0) return true;
  // bucket counts: bstart[s] before stripe s; a stripe with more blocks than any before it
  // (rec_off[s] >= 0) takes rec_bc[rec_off[s] + i] at its i-th insert
  std::vector<uint64_t> bstart(ns);
  std::vector<int64_t> rec_off(ns, -1);
  std::vector<uint64_t> rec_bc;
  int32_t maxd = 0;
  {
    std::unordered_map<I, int32_t> probe((size_t)std::sqrt((double)ncb));
    uint64_t cur = probe.bucket_count();
    for (int64_t st = 0; st < ns; ++st) {
      bstart[st] = cur;
      const int32_t D = blk_ptr[st + 1] - blk_ptr[st];
      if (D > maxd) {
        rec_off[st] = (int64_t)rec_bc.size();
        probe.clear();
        for (int32_t i = 0; i < D; ++i) {
          probe.emplace((I)i, 0);
          rec_bc.push_back(probe.bucket_count());
        }
        maxd = D;
        cur = probe.bucket_count();
      }
    }
  }
  const uint64_t nbkt_max = rec_bc.empty() ? bstart[0] : std::max(bstart[0], rec_bc.back());

  std::atomic<bool> ok{true};
  auto run = [&](int64_t s_begin, int64_t s_end) {
    constexpr int32_t kNil = -1, kFront = -2;  // end of list; the list's front ("before begin")
    std::vector<int32_t> nxt(maxd), ord(maxd), before(nbkt_max), stamp(nbkt_max, 0);
    std::vector<uint64_t> hv(maxd);
    std::vector<int32_t> cnt(br + 1);
    int32_t gen = 0;
    FastMod64 mod;
    for (int64_t st = s_begin; st < s_end; ++st) {
      const int32_t d0 = blk_ptr[st], D = blk_ptr[st + 1] - d0;
      if (D == 0) continue;
      // insertion order: CSR order of first occurrence = (first row, block column)
      std::fill(cnt.begin(), cnt.end(), 0);
      for (int32_t j = 0; j < D; ++j) cnt[blk_rmin[d0 + j] + 1]++;
      for (int r = 0; r < br; ++r) cnt[r + 1] += cnt[r];
      for (int32_t j = 0; j < D; ++j) ord[cnt[blk_rmin[d0 + j]]++] = j;
      uint64_t nb = bstart[st];
      if (mod.d != nb) mod.set(nb);
      int32_t head = kNil;
      ++gen;
      auto after = [&](int32_t prev) { return prev == kFront ? head : nxt[prev]; };
      auto link_after = [&](int32_t prev, int32_t node) {
        if (prev == kFront) head = node;
        else nxt[prev] = node;
      };
      for (int32_t i = 0; i < D; ++i) {
        const int32_t node = ord[i];
        const uint64_t key = (uint64_t)st * ncb + (uint64_t)blk_bc[d0 + node];
        hv[node] = (uint64_t)(int64_t)(I)key;  // the key as index_t, then as size_t
        const uint64_t want = rec_off[st] >= 0 ? rec_bc[rec_off[st] + i] : nb;
        if (want != nb) {  // rehash: relink the list front to back into new buckets
          nb = want;
          mod.set(nb);
          ++gen;
          int32_t p = head, bfirst = 0;
          head = kNil;
          while (p != kNil) {
            const int32_t nx = nxt[p];
            const int32_t b = (int32_t)mod(hv[p]);
            if (stamp[b] != gen) {
              nxt[p] = head;
              head = p;
              stamp[b] = gen;
              before[b] = kFront;
              if (nxt[p] != kNil) before[bfirst] = p;
              bfirst = b;
            } else {
              const int32_t prev = before[b];
              nxt[p] = after(prev);
              link_after(prev, p);
            }
            p = nx;
          }
        }
        const int32_t b = (int32_t)mod(hv[node]);
        if (stamp[b] == gen) {
          const int32_t prev = before[b];
          nxt[node] = after(prev);
          link_after(prev, node);
        } else {
          nxt[node] = head;
          head = node;
          if (nxt[node] != kNil) before[(int32_t)mod(hv[nxt[node]])] = node;
          stamp[b] = gen;
          before[b] = kFront;
        }
      }
      // emission: kept blocks in list order, at this stripe's first tile index (the kept
      // blocks, in block-column order, carry consecutive tile indices from there)
      int32_t out = -1, kept = 0;
      for (int32_t j = 0; j < D; ++j) {
        const int32_t t = blk_tile[d0 + j];
        if (t < 0) continue;
        if (out < 0) out = t;
        if (t != out + kept) ok = false;
        ++kept;
      }
      for (int32_t p = head; p != kNil && out >= 0; p = nxt[p])
        if (blk_tile[d0 + p] >= 0) perm[out++] = blk_tile[d0 + p];
    }
  };
  const int64_t hw = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
  const int64_t nt = std::min<int64_t>({hw, 16, (ns + 4095) / 4096});
  if (nt <= 1) {
    run(0, ns);
  } else {
    std::vector<std::thread> pool;
    for (int64_t t = 0; t < nt; ++t)
      pool.emplace_back(run, ns * t / nt, ns * (t + 1) / nt);
    for (auto& th : pool) th.join();
  }
  return ok.load();
}

}  // namespace gl

using namespace gl;

extern "C" {

int64_t gl_pbr_build_workspace_bytes(int64_t rows, int32_t block_rows, int64_t nnz) {
  if (rows < 0 || block_rows < 1 || nnz < 0) return -1;
  return carve_build(nullptr, rows, block_rows, nnz).bytes;
}

static int check_build(int64_t rows, int64_t cols, int64_t nnz, int32_t br, int32_t bc,
                       int32_t min_nnz, const int32_t* indptr, const int32_t* indices) {
  GL_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0, "negative size");
  GL_REQUIRE(rows < INT32_MAX && cols < INT32_MAX && nnz < INT32_MAX,
             "rows/cols/nnz exceed int32 indices");
  GL_REQUIRE(br >= 1 && bc >= 1 && (int64_t)br * bc <= 64,
             "block %dx%d does not fit a 64-bit code", br, bc);
  GL_REQUIRE(min_nnz >= 0 && min_nnz <= 255, "min_nnz_per_block must be in [0, 255]");
  // indices / values may be null for an empty matrix (never dereferenced then)
  GL_REQUIRE(indptr != nullptr, "null CSR indptr");
  return GL_OK;
}

int gl_pbr_build_count(int64_t rows, int64_t cols, int64_t nnz, int32_t block_rows,
                       int32_t block_cols, int32_t min_nnz_per_block, const int32_t* indptr,
                       const int32_t* indices, int32_t* rem_indptr, void* workspace,
                       int64_t workspace_bytes, int64_t* counts, void* stream) {
  int rc = check_build(rows, cols, nnz, block_rows, block_cols, min_nnz_per_block, indptr,
                       indices);
  if (rc) return rc;
  GL_REQUIRE(rem_indptr && counts, "null output");
  const int64_t need = carve_build(nullptr, rows, block_rows, nnz).bytes;
  if (!workspace || workspace_bytes < need) {
    set_error("build workspace too small (%lld < %lld)", (long long)workspace_bytes,
              (long long)need);
    return GL_ERR_WORKSPACE;
  }
  BuildWs w = carve_build((char*)workspace, rows, block_rows, nnz);
  hipStream_t st = (hipStream_t)stream;
  const int64_t ns = (rows + block_rows - 1) / block_rows;
  BuildArgs a{rows, cols, block_rows, block_cols, min_nnz_per_block, ns, indptr, indices};
  FillOut o{};
  o.nt = w.nt;
  o.nv = w.nv;
  o.nr = rem_indptr;
  o.nd = w.nd;
  o.big = w.big;  // written for every stripe by the count pass
  o.wcode = w.wcode;
  o.wkey = w.wkey;
  o.winfo = w.winfo;
  if (ns > 0 && (rc = launch_build<false, uint32_t>(a, o, st))) return rc;
  // the last scan kernel also writes the totals to pinned host memory of this thread
  // (coherent and portable: any device of the process may be the current one)
  static thread_local int32_t* h_tot = nullptr;
  if (!h_tot)
    GL_HIP(hipHostMalloc((void**)&h_tot, kScanArrays * sizeof(int32_t),
                         hipHostMallocCoherent | hipHostMallocPortable));
  int32_t* d_tot = nullptr;
  GL_HIP(hipHostGetDevicePointer((void**)&d_tot, h_tot, 0));
  int32_t* arrs[kScanArrays] = {w.nt, w.nv, rem_indptr, w.nd, w.big};
  const int64_t lens[kScanArrays] = {ns, ns, rows, ns, ns};
  if ((rc = scan_set(arrs, lens, w.tmp, w.tot, d_tot, st))) return rc;
  GL_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < kScanArrays; ++i) counts[i] = h_tot[i];
  return GL_OK;
}

int gl_pbr_build_fill(int64_t rows, int64_t cols, int64_t nnz, int32_t block_rows,
                      int32_t block_cols, int32_t min_nnz_per_block, int64_t num_tiles,
                      int64_t num_big_stripes, const int32_t* indptr, const int32_t* indices, const void* values,
                      int32_t value_bytes, void* workspace, int64_t workspace_bytes,
                      uint64_t* codes, int32_t* coords, int32_t* offsets, void* data,
                      const int32_t* rem_indptr, int32_t* rem_col_ind, void* rem_vals,
                      int32_t* stripe_ptr, int32_t* blk_ptr, int32_t* blk_bc, int8_t* blk_rmin,
                      int32_t* blk_tile, void* stream) {
  int rc = check_build(rows, cols, nnz, block_rows, block_cols, min_nnz_per_block, indptr,
                       indices);
  if (rc) return rc;
  GL_REQUIRE(value_bytes == 4 || value_bytes == 8, "value_bytes must be 4 or 8");
  GL_REQUIRE(offsets && rem_indptr, "null output");
  const int64_t need = carve_build(nullptr, rows, block_rows, nnz).bytes;
  if (!workspace || workspace_bytes < need) {
    set_error("build workspace too small");
    return GL_ERR_WORKSPACE;
  }
  BuildWs w = carve_build((char*)workspace, rows, block_rows, nnz);
  hipStream_t st = (hipStream_t)stream;
  const int64_t ns = (rows + block_rows - 1) / block_rows;
  BuildArgs a{rows, cols, block_rows, block_cols, min_nnz_per_block, ns, indptr, indices};
  FillOut o{};
  o.nt = w.nt;
  o.nv = w.nv;
  o.nr = const_cast<int32_t*>(rem_indptr);
  o.nd = w.nd;
  o.codes = codes;
  o.coords = coords;
  o.offsets = offsets;
  o.data = data;
  o.rem_col = rem_col_ind;
  o.rem_val = rem_vals;
  o.values = values;
  o.blk_bc = blk_bc;
  o.blk_rmin = blk_rmin;
  o.blk_tile = blk_tile;
  o.big = w.big;  // as the count pass left it on this workspace
  o.stripe_ptr = stripe_ptr;
  o.wcode = w.wcode;
  o.wkey = w.wkey;
  o.winfo = w.winfo;
  if (ns > 0) {
    // the fill kernel also writes offsets[T] = V and stripe_ptr[ns] = T; the merge kernel
    // runs when the count pass left it stripes (or their number is not given)
    const bool merge = num_big_stripes != 0;
    rc = value_bytes == 4 ? launch_build<true, uint32_t>(a, o, st, merge)
                          : launch_build<true, uint64_t>(a, o, st, merge);
    if (rc) return rc;
  } else {
    GL_HIP(hipMemcpyAsync(offsets + num_tiles, w.nv, 4, hipMemcpyDeviceToDevice, st));
    if (stripe_ptr) GL_HIP(hipMemcpyAsync(stripe_ptr, w.nt, 4, hipMemcpyDeviceToDevice, st));
  }
  // the distinct-block pointer per stripe when asked for
  if (blk_ptr) GL_HIP(hipMemcpyAsync(blk_ptr, w.nd, (ns + 1) * 4, hipMemcpyDeviceToDevice, st));
  return GL_OK;
}

// Host replay of the reference's std::unordered_map emission order (host arrays).
// index_bytes selects the key type of the reference builder instantiation (int32 or int64).
int gl_pbr_reference_perm(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                          int32_t index_bytes, const int32_t* blk_ptr, const int32_t* blk_bc,
                          const int8_t* blk_rmin, const int32_t* blk_tile, int32_t* perm) {
  GL_REQUIRE(index_bytes == 4 || index_bytes == 8, "index_bytes must be 4 or 8");
  GL_REQUIRE(block_rows >= 1 && block_cols >= 1, "bad block size");
  const int64_t ns = (rows + block_rows - 1) / block_rows;
  const bool ok =
      index_bytes == 4
          ? replay<int32_t>(ns, cols, block_rows, block_cols, blk_ptr, blk_bc, blk_rmin, blk_tile,
                            perm)
          : replay<int64_t>(ns, cols, block_rows, block_cols, blk_ptr, blk_bc, blk_rmin, blk_tile,
                            perm);
  GL_REQUIRE(ok, "block list: a stripe's kept blocks do not carry consecutive tile indices");
  return GL_OK;
}

// Reorder tiles: out tile t = in tile perm[t]; values move with their tile.
int gl_pbr_permute_tiles(int64_t T, const int32_t* perm, const uint64_t* codes,
                         const int32_t* coords, const int32_t* offsets, const void* data,
                         int32_t value_bytes, uint64_t* codes_out, int32_t* coords_out,
                         int32_t* offsets_out, void* data_out, void* workspace,
                         int64_t workspace_bytes, void* stream) {
  GL_REQUIRE(T >= 0, "negative size");
  GL_REQUIRE(value_bytes == 4 || value_bytes == 8, "value_bytes must be 4 or 8");
  if (T == 0) return GL_OK;
  const int64_t need = scan_tmp_elems(T + 1) * 4;
  GL_REQUIRE(workspace && workspace_bytes >= need, "permute workspace too small");
  int32_t* tmp = (int32_t*)workspace;
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = (unsigned)((T + 255) / 256);
  hipLaunchKernelGGL(k_tile_popc, dim3(nb), dim3(256), 0, st, codes, perm, T, offsets_out);
  GL_LAUNCH_CHECK("k_tile_popc");
  int rc = excl_scan(offsets_out, offsets_out, T, tmp, st);
  if (rc) return rc;
  if (value_bytes == 4)
    hipLaunchKernelGGL(k_tile_gather<uint32_t>, dim3(nb), dim3(256), 0, st, perm, T, codes,
                       coords, offsets, (const uint32_t*)data, offsets_out, codes_out,
                       coords_out, (uint32_t*)data_out);
  else
    hipLaunchKernelGGL(k_tile_gather<uint64_t>, dim3(nb), dim3(256), 0, st, perm, T, codes,
                       coords, offsets, (const uint64_t*)data, offsets_out, codes_out,
                       coords_out, (uint64_t*)data_out);
  GL_LAUNCH_CHECK("k_tile_gather");
  return GL_OK;
}

int64_t gl_pbr_permute_workspace_bytes(int64_t T) { return scan_tmp_elems(T + 1) * 4; }

}  // extern "C"
