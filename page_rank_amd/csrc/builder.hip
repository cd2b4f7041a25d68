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
#include <cmath>
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
  uint8_t* big;       // stripes left to the merge kernel (more distinct blocks than the
                      // LDS table holds); NULL = the merge kernel takes every stripe
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
    const bool mine = s < a.num_stripes && (o.big == nullptr || o.big[s] != 0);
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
// Wave-per-stripe builder. A stripe's entries are one contiguous CSR segment; the wave reads
// it 64 entries at a time (coalesced) and inserts every entry's block column into an LDS
// hash table (atomic CAS on the key, atomic OR of the entry's bit into the code, atomic min
// of its row). The occupied keys are then put in ascending order (rank = number of smaller
// keys, counted with broadcast LDS reads), which gives
// the canonical tile order, tile indices and value offsets by prefix sums; a second sweep
// over the entries writes kept values at popcount(code & bits below) and rejected entries
// at their rank inside their row (ballot prefix counts), so remainder rows keep column
// order. Results are the merge kernel's, entry for entry. Stripes with more than kHashCap
// distinct blocks are flagged and left to the merge kernel.
// ---------------------------------------------------------------------------------------
constexpr int kHashCap = 256;
// Register budget of the fill pass: 7 waves per SIMD (72 VGPRs), which is also what its LDS
// allows (7 workgroups of 4 waves per CU).
constexpr int kHashWavesPerSimd = 7;
// Grid of the count pass, in workgroups per CU: 7 of the 8 its LDS allows (0.71 ms for the
// whole 1M x 1M build, against 0.77 ms with 8 and with 4).
constexpr int kHashCountWgPerCu = 7;
// Chunks of 64 entries per stripe whose column (and, in the fill pass, value) loads leave
// together at the top of the stripe and stay in registers for both sweeps: one memory round
// trip instead of one per chunk per sweep. Longer stripes load the rest chunk by chunk.
constexpr int kHashPre = 4;

// FILL: the fill pass's arrays (value offsets, slots in key order) exist only there, so the
// count pass needs less LDS per wave and holds more waves. A block's first row is not kept:
// it is the lowest set bit of its code divided by block_cols.
template <bool FILL>
struct HashLds {
  static constexpr int kF = FILL ? kHashCap : 1;
  int32_t key[kHashCap];
  uint32_t lo[kHashCap], hi[kHashCap];
  int32_t voff[kHashCap];  // per slot when kept: offset of its first value (fill) or 1 (count);
                           // -1 if not
  int32_t sslot[kF];       // slots in ascending key order
  int32_t rows[65];        // CSR start of each row of the stripe (block_rows + 1 values)
  int32_t rcnt[64];        // rejected entries per row so far
  int32_t flag;            // table overflow
};

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int hash_slot(int32_t b) {
  return (int)(((uint32_t)b * 2654435761u) >> 24);  // kHashCap = 256
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
// a mask instead of an integer division (the kernel is bound by VALU issue).
template <bool FILL, typename V, bool P2>
__global__ void __launch_bounds__(256, FILL ? kHashWavesPerSimd : 1)
    k_build_hash(BuildArgs a, FillOut o) {
  __shared__ HashLds<FILL> sh_all[4];
  const int lane = threadIdx.x & (kWave - 1);
  HashLds<FILL>& h = sh_all[threadIdx.x / kWave];
  // uniform across the wave: scalar registers
  const int64_t wave = __builtin_amdgcn_readfirstlane(
      (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kWave));
  const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / kWave;
  const V* __restrict__ vals = (const V*)o.values;
  const int br = a.br, bc = a.bc;
  const int bc_shift = P2 ? __builtin_ctz((unsigned)bc) : 0;
  const uint64_t below_lane = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  // block column of a column, and the entry's bit in its block's code (row r of the stripe)
  auto block_of = [&](int32_t col) { return P2 ? (col >> bc_shift) : col / bc; };
  auto bit_of = [&](int r, int32_t col, int32_t b) {
    return P2 ? ((r << bc_shift) | (col & (bc - 1))) : r * bc + (col - b * bc);
  };

  // Per-stripe loads that do not depend on the stripe's entries: CSR row starts (row k of the
  // stripe in lane k, row 64 in lane 0's second register) and, in the fill pass, the stripe's
  // output offsets and its rows' remainder offsets. They are issued a stripe ahead.
  struct StripeHead {
    int32_t rp0, rp1, nr;
    int32_t nt, nv, nd;
  };
  auto load_head = [&](int64_t st) {
    StripeHead hd{0, 0, 0, 0, 0, 0};
    if (st >= a.num_stripes) return hd;
    const int64_t rb = st * br;
    if (lane <= br) hd.rp0 = a.indptr[min<int64_t>(rb + lane, a.rows)];
    if (lane == 0 && br >= kWave) hd.rp1 = a.indptr[min<int64_t>(rb + kWave, a.rows)];
    if constexpr (FILL) {
      if (lane < br && rb + lane < a.rows) hd.nr = o.nr[rb + lane];
      hd.nt = o.nt[st];
      hd.nv = o.nv[st];
      hd.nd = o.nd[st];
    }
    return hd;
  };
  StripeHead head = load_head(wave);

  for (int64_t s = wave; s < a.num_stripes; s += nwaves) {
    const int64_t r0 = s * br;
    const StripeHead cur = head;
    if (lane <= br) h.rows[lane] = cur.rp0;
    if (lane == 0 && br >= kWave) h.rows[kWave] = cur.rp1;
    // fill: remainder entries of each row so far start at the row's remainder offset
    if (lane < br) h.rcnt[lane] = FILL ? cur.nr : 0;
    for (int i = lane; i < kHashCap; i += kWave) {
      h.key[i] = -1;
      h.lo[i] = 0u;
      h.hi[i] = 0u;
    }
    if (lane == 0) h.flag = 0;
    wsync();
    const int32_t E0 = __builtin_amdgcn_readfirstlane(h.rows[0]);
    const int32_t E1 = __builtin_amdgcn_readfirstlane(h.rows[br]);
    // this stripe's first kHashPre chunks, then the next stripe's head
    int32_t pcol[kHashPre];
    V pval[kHashPre];
#pragma unroll
    for (int k = 0; k < kHashPre; ++k) {
      const int32_t e = E0 + k * kWave + lane;
      pcol[k] = e < E1 ? a.indices[e] : 0;
      if constexpr (FILL) pval[k] = e < E1 ? vals[e] : V(0);
    }
    head = load_head(s + nwaves);
    // row starts 1..7 in registers for the common block_rows <= 8 (a row is then found by
    // compares instead of a search in LDS)
    int32_t rs8[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) rs8[k] = k + 1 < br ? h.rows[k + 1] : INT32_MAX;
    auto row_at = [&](int32_t e) {
      if (br <= 8) {
        int r = 0;
#pragma unroll
        for (int k = 0; k < 7; ++k) r += rs8[k] <= e ? 1 : 0;
        return r;
      }
      return row_of(h.rows, br, e);
    };

    // ---- insert every entry's block
    // returns the entry's slot | bit << 8 | row << 14 (-1 if the table is full), which the
    // entry sweep takes instead of finding the slot again
    auto insert = [&](int32_t e, int32_t col) -> int32_t {
      const int r = row_at(e);
      const int32_t b = block_of(col);
      const int bit = bit_of(r, col, b);
      int slot = hash_slot(b);
      int probes = 0;
      for (;;) {
        const int32_t old = atomicCAS(&h.key[slot], -1, b);
        if (old == -1 || old == b) break;
        slot = (slot + 1) & (kHashCap - 1);
        if (++probes == kHashCap) {
          slot = -1;
          break;
        }
      }
      if (slot < 0) {
        h.flag = 1;
        return -1;
      }
      if (bit < 32) atomicOr(&h.lo[slot], 1u << bit);
      else atomicOr(&h.hi[slot], 1u << (bit - 32));
      return slot | (bit << 8) | (r << 14);
    };
    int32_t pinf[kHashPre];
#pragma unroll
    for (int k = 0; k < kHashPre; ++k) {
      const int32_t e = E0 + k * kWave + lane;
      pinf[k] = e < E1 ? insert(e, pcol[k]) : -1;
    }
    for (int32_t c = E0 + kHashPre * kWave; c < E1; c += kWave) {
      const int32_t e = c + lane;
      if (e < E1) insert(e, a.indices[e]);
    }
    wsync();
    if (h.flag) {
      if (!FILL && lane == 0) o.big[s] = 1;  // the merge kernel takes this stripe
      wsync();
      continue;
    }

    if constexpr (!FILL) {
      // counts do not depend on the tile order: sweep the table as it is (per-row remainder
      // counts from the kept codes, with LDS atomics, measured slower than the entry sweep)
      int32_t nd = 0, nt = 0, nv = 0;
#pragma unroll
      for (int t = 0; t < kHashCap / kWave; ++t) {
        const int i = t * kWave + lane;
        const bool occ = h.key[i] >= 0;
        const int32_t cnt = __popc(h.lo[i]) + __popc(h.hi[i]);
        const bool kept = occ && cnt >= a.min_nnz;
        h.voff[i] = kept ? 1 : -1;
        nd += occ ? 1 : 0;
        nt += kept ? 1 : 0;
        nv += kept ? cnt : 0;
      }
#pragma unroll
      for (int m = 1; m < kWave; m <<= 1) {
        nd += __shfl_xor(nd, m, kWave);
        nt += __shfl_xor(nt, m, kWave);
        nv += __shfl_xor(nv, m, kWave);
      }
      if (lane == 0) {
        o.nt[s] = nt;
        o.nv[s] = nv;
        o.nd[s] = nd;
      }
    } else {
    // ---- listed keys, compacted (keys into voff[], free until the kept blocks are numbered,
    // slots into sslot[]), then ranked: sslot[] lists the slots in ascending key order. The
    // list holds every occupied block when the reference-order block list is written, else
    // only the kept ones (on remainder-heavy graphs a stripe has a few kept blocks among
    // a hundred), whose voff[] entries are the only ones set >= 0 below.
    const bool all_blocks = o.blk_bc != nullptr;
    int32_t D = 0;
#pragma unroll
    for (int t = 0; t < kHashCap / kWave; ++t) {
      const int i = t * kWave + lane;
      const int32_t k = h.key[i];
      const bool listed =
          k >= 0 && (all_blocks || __popc(h.lo[i]) + __popc(h.hi[i]) >= a.min_nnz);
      const uint64_t m = __ballot(listed);
      if (listed) {
        const int pos = D + __popcll(m & below_lane);
        h.voff[pos] = k;
        h.sslot[pos] = i;
      }
      D += __popcll(m);
    }
    // rank of each key = number of smaller keys (keys are unique); broadcast reads, no
    // stage barriers
    wsync();
    {
      int32_t mykey[kHashCap / kWave], myslot[kHashCap / kWave], rank[kHashCap / kWave];
#pragma unroll
      for (int t = 0; t < kHashCap / kWave; ++t) {
        const int j = t * kWave + lane;
        mykey[t] = j < D ? h.voff[j] : INT32_MAX;
        myslot[t] = j < D ? h.sslot[j] : -1;
        rank[t] = 0;
      }
      for (int i = 0; i < D; ++i) {
        const int32_t ki = h.voff[i];
#pragma unroll
        for (int t = 0; t < kHashCap / kWave; ++t) rank[t] += ki < mykey[t] ? 1 : 0;
      }
      wsync();
#pragma unroll
      for (int t = 0; t < kHashCap / kWave; ++t) {
        if (t * kWave + lane < D) h.sslot[rank[t]] = myslot[t];
        // kept-only list: every unlisted block is rejected (the listed ones are set below)
        if (!all_blocks) h.voff[t * kWave + lane] = -1;
      }
    }
    wsync();

    // ---- kept blocks: tile indices and value offsets by prefix sums over the sorted list
    int32_t tile_base = 0, val_base = 0, d_base = 0;
    if (FILL) {
      tile_base = cur.nt;
      val_base = cur.nv;
      d_base = cur.nd;
    }
    int32_t kcar = 0, vcar = 0;
    for (int j0 = 0; j0 < D; j0 += kWave) {
      const int j = j0 + lane;
      const bool valid = j < D;
      const int slot = valid ? h.sslot[j] : 0;
      const uint64_t code = valid ? (((uint64_t)h.hi[slot] << 32) | h.lo[slot]) : 0ull;
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
      if (valid) {
        const int32_t t = kept ? (FILL ? tile_base + krank : 1) : -1;
        h.voff[slot] = kept ? val_base + vex : -1;
        if (FILL) {
          const int32_t key = h.key[slot];
          if (kept) {
            o.codes[t] = code;
            o.coords[2 * (int64_t)t] = (int32_t)r0;
            o.coords[2 * (int64_t)t + 1] = key * bc;
            o.offsets[t] = val_base + vex;
          }
          if (o.blk_bc != nullptr) {
            o.blk_bc[d_base + j] = key;
            o.blk_rmin[d_base + j] = (int8_t)(__builtin_ctzll(code) / bc);
            o.blk_tile[d_base + j] = kept ? t : -1;
          }
        }
      }
      kcar += __popcll(km);
      vcar += __shfl(vin, kWave - 1, kWave);
    }
    }  // FILL
    wsync();

    // ---- entries: kept values into their tile, rejected ones to the remainder row
    // inf: what insert returned for the entry, or -1 to find its slot again
    auto sweep = [&](int32_t c, int32_t col, V val, int32_t inf) {
      const int32_t e = c + lane;
      const bool valid = e < E1;
      int r = 0, slot = 0, bit = 0;
      if (valid) {
        if (inf >= 0) {
          slot = inf & 255;
          bit = (inf >> 8) & 63;
          r = inf >> 14;
        } else {
          r = row_at(e);
          const int32_t b = block_of(col);
          bit = bit_of(r, col, b);
          slot = hash_slot(b);
          while (h.key[slot] != b) slot = (slot + 1) & (kHashCap - 1);
        }
      }
      const bool rej = valid && h.voff[slot] < 0;
      const uint64_t rm = __ballot(rej);
      // lanes of this row in this chunk: [ls, ...]
      const int ls = valid ? max(0, h.rows[r] - c) : 0;
      const uint64_t from_ls = ls == 0 ? ~0ull : ~(~0ull >> (64 - ls));
      const int32_t before = __popcll(rm & below_lane & from_ls);
      if constexpr (FILL) {
        if (valid) {
          if (rej) {
            const int32_t pos = h.rcnt[r] + before;
            o.rem_col[pos] = col;
            ((V*)o.rem_val)[pos] = val;
          } else {
            const uint64_t code = ((uint64_t)h.hi[slot] << 32) | h.lo[slot];
            const uint64_t below = bit == 0 ? 0ull : (~0ull >> (64 - bit));
            ((V*)o.data)[h.voff[slot] + __popcll(code & below)] = val;
          }
        }
      }
      // the last lane of each row in this chunk advances that row's count
      const bool last = valid && (lane == kWave - 1 || e + 1 == h.rows[r + 1]);
      const int32_t in_row = __popcll(rm & (below_lane | (1ull << lane)) & from_ls);
      wsync();
      if (last) h.rcnt[r] += in_row;
      wsync();
    };
#pragma unroll
    for (int k = 0; k < kHashPre; ++k) {
      const int32_t c = E0 + k * kWave;
      if (c < E1) sweep(c, pcol[k], FILL ? pval[k] : V(0), pinf[k]);
    }
    for (int32_t c = E0 + kHashPre * kWave; c < E1; c += kWave) {
      const int32_t e = c + lane;
      const bool ok = e < E1;
      sweep(c, ok ? a.indices[e] : 0, (FILL && ok) ? vals[e] : V(0), -1);
    }
    if (!FILL && lane < br && r0 + lane < a.rows) o.nr[r0 + lane] = h.rcnt[lane];
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

// ---- workspace layout ----
struct BuildWs {
  int32_t *nt, *nv, *nd, *tmp;
  uint8_t* big;
  int64_t bytes;
};
static int64_t al(int64_t x) { return (x + 255) & ~(int64_t)255; }
static BuildWs carve_build(char* base, int64_t rows, int32_t br) {
  const int64_t ns = (rows + br - 1) / br;
  const int64_t tmpn = scan_tmp_elems(std::max(rows, ns) + 1);
  BuildWs w;
  int64_t off = 0;
  auto take = [&](int64_t n) {
    char* p = base ? base + off : nullptr;
    off += al(n);
    return (int32_t*)p;
  };
  w.nt = take((ns + 1) * 4);
  w.nv = take((ns + 1) * 4);
  w.nd = take((ns + 1) * 4);
  w.tmp = take(tmpn * 4);
  w.big = (uint8_t*)take(ns);
  w.bytes = off;
  return w;
}

static int group_lanes(int br) {
  int g = 1;
  while (g < br) g <<= 1;
  return g;
}

template <bool FILL, typename V, bool P2>
static void launch_build_hash(const BuildArgs& a, const FillOut& o, hipStream_t st) {
  // waves: the fill pass fills the CUs once (fewer or more workgroups were slower), the
  // count pass uses kHashCountWgPerCu
  static std::atomic<int> per_cu{-1};
  if (per_cu.load() < 0) {
    int n = kHashCountWgPerCu;
    if (FILL &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_build_hash<FILL, V, P2>, 256, 0) !=
             hipSuccess ||
         n < 1))
      n = 4;
    per_cu.store(n);
  }
  const int64_t cap = (int64_t)device_info().num_cus * per_cu.load() * 4;
  const int64_t nw = std::max<int64_t>(4, std::min(a.num_stripes, cap));
  hipLaunchKernelGGL((k_build_hash<FILL, V, P2>), dim3((unsigned)((nw + 3) / 4)), dim3(256), 0,
                     st, a, o);
}

template <bool FILL, typename V>
static int launch_build(const BuildArgs& a, const FillOut& o, hipStream_t st) {
  if (o.big != nullptr) {
    if ((a.bc & (a.bc - 1)) == 0) launch_build_hash<FILL, V, true>(a, o, st);
    else launch_build_hash<FILL, V, false>(a, o, st);
    GL_LAUNCH_CHECK("k_build_hash");
  }
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

template <typename I>
static void replay(int64_t ns, int64_t cols, int32_t br, int32_t bc, const int32_t* blk_ptr,
                   const int32_t* blk_bc, const int8_t* blk_rmin, const int32_t* blk_tile,
                   int32_t* perm) {
  const uint64_t ncb = ((uint64_t)cols + bc - 1) / bc;
  std::unordered_map<I, int32_t> m((size_t)std::sqrt((double)ncb));
  std::vector<int32_t> order;
  std::vector<int32_t> bucket_start(br + 1);
  int64_t out = 0;
  for (int64_t s = 0; s < ns; ++s) {
    const int32_t d0 = blk_ptr[s], d1 = blk_ptr[s + 1];
    // insertion order of the reference: CSR order of first occurrence = (first row, block col)
    std::fill(bucket_start.begin(), bucket_start.end(), 0);
    for (int32_t d = d0; d < d1; ++d) bucket_start[blk_rmin[d] + 1]++;
    for (int i = 0; i < br; ++i) bucket_start[i + 1] += bucket_start[i];
    order.assign(d1 - d0, 0);
    for (int32_t d = d0; d < d1; ++d) order[bucket_start[blk_rmin[d]]++] = d;
    for (int32_t d : order) {
      const uint64_t block_index = (uint64_t)s * ncb + (uint64_t)blk_bc[d];
      m.emplace((I)block_index, d);
    }
    for (const auto& kv : m)
      if (blk_tile[kv.second] >= 0) perm[out++] = blk_tile[kv.second];
    m.clear();
  }
}

}  // namespace gl

using namespace gl;

extern "C" {

int64_t gl_pbr_build_workspace_bytes(int64_t rows, int32_t block_rows) {
  if (rows < 0 || block_rows < 1) return -1;
  return carve_build(nullptr, rows, block_rows).bytes;
}

static int check_build(int64_t rows, int64_t cols, int32_t br, int32_t bc, int32_t min_nnz,
                       const int32_t* indptr, const int32_t* indices) {
  GL_REQUIRE(rows >= 0 && cols >= 0, "negative size");
  GL_REQUIRE(rows < INT32_MAX && cols < INT32_MAX, "rows/cols exceed int32 indices");
  GL_REQUIRE(br >= 1 && bc >= 1 && (int64_t)br * bc <= 64,
             "block %dx%d does not fit a 64-bit code", br, bc);
  GL_REQUIRE(min_nnz >= 0 && min_nnz <= 255, "min_nnz_per_block must be in [0, 255]");
  // indices / values may be null for an empty matrix (never dereferenced then)
  GL_REQUIRE(indptr != nullptr, "null CSR indptr");
  return GL_OK;
}

int gl_pbr_build_count(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                       int32_t min_nnz_per_block, const int32_t* indptr,
                       const int32_t* indices, int32_t* rem_indptr, void* workspace,
                       int64_t workspace_bytes, int64_t* counts, void* stream) {
  int rc = check_build(rows, cols, block_rows, block_cols, min_nnz_per_block, indptr, indices);
  if (rc) return rc;
  GL_REQUIRE(rem_indptr && counts, "null output");
  const int64_t need = carve_build(nullptr, rows, block_rows).bytes;
  if (!workspace || workspace_bytes < need) {
    set_error("build workspace too small (%lld < %lld)", (long long)workspace_bytes,
              (long long)need);
    return GL_ERR_WORKSPACE;
  }
  BuildWs w = carve_build((char*)workspace, rows, block_rows);
  hipStream_t st = (hipStream_t)stream;
  const int64_t ns = (rows + block_rows - 1) / block_rows;
  BuildArgs a{rows, cols, block_rows, block_cols, min_nnz_per_block, ns, indptr, indices};
  FillOut o{};
  o.nt = w.nt;
  o.nv = w.nv;
  o.nr = rem_indptr;
  o.nd = w.nd;
  o.big = w.big;
  if (ns > 0) GL_HIP(hipMemsetAsync(w.big, 0, ns, st));
  if (ns > 0 && (rc = launch_build<false, uint32_t>(a, o, st))) return rc;
  if ((rc = excl_scan(w.nt, w.nt, ns, w.tmp, st))) return rc;
  if ((rc = excl_scan(w.nv, w.nv, ns, w.tmp, st))) return rc;
  if ((rc = excl_scan(w.nd, w.nd, ns, w.tmp, st))) return rc;
  if ((rc = excl_scan(rem_indptr, rem_indptr, rows, w.tmp, st))) return rc;
  int32_t h[4];
  GL_HIP(hipMemcpyAsync(&h[0], w.nt + ns, 4, hipMemcpyDeviceToHost, st));
  GL_HIP(hipMemcpyAsync(&h[1], w.nv + ns, 4, hipMemcpyDeviceToHost, st));
  GL_HIP(hipMemcpyAsync(&h[2], rem_indptr + rows, 4, hipMemcpyDeviceToHost, st));
  GL_HIP(hipMemcpyAsync(&h[3], w.nd + ns, 4, hipMemcpyDeviceToHost, st));
  GL_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < 4; ++i) counts[i] = h[i];
  return GL_OK;
}

int gl_pbr_build_fill(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                      int32_t min_nnz_per_block, int64_t num_tiles, const int32_t* indptr,
                      const int32_t* indices,
                      const void* values, int32_t value_bytes, void* workspace,
                      int64_t workspace_bytes, uint64_t* codes, int32_t* coords,
                      int32_t* offsets, void* data, const int32_t* rem_indptr,
                      int32_t* rem_col_ind, void* rem_vals, int32_t* blk_ptr, int32_t* blk_bc,
                      int8_t* blk_rmin, int32_t* blk_tile, void* stream) {
  int rc = check_build(rows, cols, block_rows, block_cols, min_nnz_per_block, indptr, indices);
  if (rc) return rc;
  GL_REQUIRE(value_bytes == 4 || value_bytes == 8, "value_bytes must be 4 or 8");
  GL_REQUIRE(offsets && rem_indptr, "null output");
  const int64_t need = carve_build(nullptr, rows, block_rows).bytes;
  if (!workspace || workspace_bytes < need) {
    set_error("build workspace too small");
    return GL_ERR_WORKSPACE;
  }
  BuildWs w = carve_build((char*)workspace, rows, block_rows);
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
  if (ns > 0) {
    rc = value_bytes == 4 ? launch_build<true, uint32_t>(a, o, st)
                          : launch_build<true, uint64_t>(a, o, st);
    if (rc) return rc;
  }
  // offsets[T] = V, and the distinct-block pointer per stripe when asked for
  GL_HIP(hipMemcpyAsync(offsets + num_tiles, w.nv + ns, 4, hipMemcpyDeviceToDevice, st));
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
  if (index_bytes == 4)
    replay<int32_t>(ns, cols, block_rows, block_cols, blk_ptr, blk_bc, blk_rmin, blk_tile, perm);
  else
    replay<int64_t>(ns, cols, block_rows, block_cols, blk_ptr, blk_bc, blk_rmin, blk_tile, perm);
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
