// ORACLE — test infrastructure only. Nothing in the product imports, links or calls this.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
//
// CPU restatement of the reference's PBR algorithms (Dayan135/Page_Rank):
//   orc_csr_to_pbr_*   builder      Backend/graph_link/pbr_matrix/pbr_matrix.cpp:340-477
//   orc_pbr_spmm_*     CPU SpMM     Backend/graph_link/pbr_matrix/pbr_matrix.cpp:17-68
//   orc_pbr_to_csr_*   round trip   Backend/graph_link/pbr_matrix/pbr_matrix.cpp:289-338
// Field layout follows Backend/graph_link/pbr_matrix/pbr_matrix.hpp:98-167.
//
// Emission order of tiles inside a stripe: the reference iterates a
// std::unordered_map<index_t, ...> that was built with the bucket hint sqrt(num_col_blocks)
// (pbr_matrix.cpp:374), filled in CSR order with the key block_row*num_col_blocks+block_col
// truncated to index_t (pbr_matrix.cpp:386-389), and clear()ed after every stripe
// (pbr_matrix.cpp:426-428), so bucket counts persist across stripes. This restatement drives
// the same library container through the same sequence of operations, which reproduces that
// order with the C++ standard library of this image (libstdc++ of GCC 11). order=1 emits the
// canonical order instead (block columns ascending), which is what the product builder's
// canonical mode produces.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

template <typename I, typename S>
struct Pbr {
  I rows = 0, cols = 0, br = 0, bc = 0, nnz = 0;
  std::vector<uint64_t> codes;
  std::vector<I> coords;   // flattened (row0, col0) pairs
  std::vector<I> offsets;  // T + 1
  std::vector<S> data;
  std::vector<I> rptr, rcol;
  std::vector<S> rval;
};

template <typename I, typename S>
struct StripeBlock {
  uint64_t code = 0;
  I row0 = 0, col0 = 0;
  std::vector<S> vals;
};

template <typename I, typename S>
Pbr<I, S>* build(I rows, I cols, const I* indptr, const I* indices, const S* data, I br, I bc,
                 uint8_t min_nnz, int order) {
  auto* out = new Pbr<I, S>();
  out->rows = rows;
  out->cols = cols;
  out->br = br;
  out->bc = bc;
  out->nnz = rows > 0 ? indptr[rows] : 0;
  const uint64_t ncb = ((uint64_t)cols + bc - 1) / bc;
  std::vector<std::vector<std::pair<I, S>>> bins(rows);
  out->offsets.push_back(0);
  I cur_off = 0;

  std::unordered_map<I, StripeBlock<I, S>> stripe((size_t)std::sqrt((double)ncb));
  std::vector<I> first_seen;  // canonical order: keys sorted by block column

  auto flush = [&]() {
    auto emit = [&](StripeBlock<I, S>& b) {
      const int cnt = __builtin_popcountll(b.code);
      if (cnt >= (int)min_nnz) {
        out->codes.push_back(b.code);
        out->coords.push_back(b.row0);
        out->coords.push_back(b.col0);
        out->data.insert(out->data.end(), b.vals.begin(), b.vals.end());
        cur_off += (I)cnt;
        out->offsets.push_back(cur_off);
      } else {
        size_t k = 0;
        for (I r = 0; r < br; ++r)
          for (I c = 0; c < bc; ++c)
            if ((b.code >> (uint64_t)(r * bc + c)) & 1ull)
              bins[b.row0 + r].emplace_back((I)(b.col0 + c), b.vals[k++]);
      }
    };
    if (order == 0) {
      for (auto& kv : stripe) emit(kv.second);
    } else {
      std::vector<StripeBlock<I, S>*> bl;
      for (auto& kv : stripe) bl.push_back(&kv.second);
      std::sort(bl.begin(), bl.end(),
                [](const StripeBlock<I, S>* a, const StripeBlock<I, S>* b) {
                  return a->col0 < b->col0;
                });
      for (auto* b : bl) emit(*b);
    }
    stripe.clear();
  };

  for (uint64_t row = 0; row < (uint64_t)rows; ++row) {
    const I block_row = (I)(row / (uint64_t)br);
    const uint64_t roff = row % (uint64_t)br;
    for (uint64_t j = (uint64_t)indptr[row]; j < (uint64_t)indptr[row + 1]; ++j) {
      const uint64_t col = (uint64_t)indices[j];
      const uint64_t bcol = col / (uint64_t)bc;
      const uint64_t bit = roff * (uint64_t)bc + col % (uint64_t)bc;
      const uint64_t block_index = (uint64_t)block_row * ncb + bcol;
      auto ins = stripe.emplace((I)block_index, StripeBlock<I, S>{});
      StripeBlock<I, S>& b = ins.first->second;
      if (ins.second) {
        b.row0 = (I)(block_row * br);
        b.col0 = (I)(bcol * (uint64_t)bc);
      }
      b.code |= 1ull << bit;
      b.vals.push_back(data[j]);
    }
    if ((row + 1) / (uint64_t)br > (uint64_t)block_row) flush();
  }
  flush();  // trailing partial stripe

  out->rptr.assign((size_t)rows + 1, 0);
  for (I r = 0; r < rows; ++r) {
    std::sort(bins[r].begin(), bins[r].end());
    out->rptr[r + 1] = out->rptr[r] + (I)bins[r].size();
  }
  for (auto& b : bins)
    for (auto& e : b) {
      out->rcol.push_back(e.first);
      out->rval.push_back(e.second);
    }
  return out;
}

// Y[B][rows][K] (+)= A X, tiles in stored order then the remainder, accumulating in A.
template <typename I, typename S, typename A>
void spmm(I rows, I cols, I br, I bc, int64_t T, const uint64_t* codes, const I* coords,
          const S* data, const I* rptr, const I* rcol, const S* rval, const S* X, int64_t K,
          S* Y) {
  std::vector<A> acc((size_t)rows * K, A(0));
  int64_t di = 0;
  for (int64_t t = 0; t < T; ++t) {
    const uint64_t code = codes[t];
    const I row0 = coords[2 * t], col0 = coords[2 * t + 1];
    for (I r = 0; r < br; ++r) {
      const int64_t gr = (int64_t)row0 + r;
      if (gr >= rows) continue;
      for (I c = 0; c < bc; ++c) {
        const int64_t gc = (int64_t)col0 + c;
        if (gc >= cols) continue;
        if ((code >> (uint64_t)(r * bc + c)) & 1ull) {
          const A v = (A)data[di++];
          for (int64_t w = 0; w < K; ++w) acc[gr * K + w] += v * (A)X[gc * K + w];
        }
      }
    }
  }
  for (int64_t r = 0; r < rows; ++r)
    for (I j = rptr[r]; j < rptr[r + 1]; ++j) {
      const A v = (A)rval[j];
      const int64_t c = rcol[j];
      for (int64_t w = 0; w < K; ++w) acc[r * K + w] += v * (A)X[c * K + w];
    }
  for (size_t i = 0; i < acc.size(); ++i) Y[i] = (S)acc[i];
}

// PBR -> sorted CSR. strict=1 keeps the reference's refusal of tiles in a trailing partial
// stripe (pbr_matrix.cpp:308-310) and returns -1 for them.
template <typename I, typename S>
int to_csr(I rows, I cols, I br, I bc, int64_t T, const uint64_t* codes, const I* coords,
           const S* data, const I* rptr, const I* rcol, const S* rval, I* indptr, I* indices,
           S* out, int strict) {
  std::vector<std::vector<std::pair<I, S>>> bins(rows);
  int64_t di = 0;
  for (int64_t t = 0; t < T; ++t) {
    const I row0 = coords[2 * t], col0 = coords[2 * t + 1];
    for (I r = 0; r < br; ++r) {
      if ((int64_t)row0 + r >= rows) {
        if (strict) return -1;
        continue;
      }
      for (I c = 0; c < bc; ++c) {
        if ((int64_t)col0 + c >= cols) continue;
        if ((codes[t] >> (uint64_t)(r * bc + c)) & 1ull)
          bins[row0 + r].emplace_back((I)(col0 + c), data[di++]);
      }
    }
  }
  for (I r = 0; r < rows; ++r)
    for (I j = rptr[r]; j < rptr[r + 1]; ++j) bins[r].emplace_back(rcol[j], rval[j]);
  I n = 0;
  indptr[0] = 0;
  for (I r = 0; r < rows; ++r) {
    std::sort(bins[r].begin(), bins[r].end());
    for (auto& e : bins[r]) {
      indices[n] = e.first;
      out[n] = e.second;
      ++n;
    }
    indptr[r + 1] = n;
  }
  return 0;
}

}  // namespace

#define ORC_VARIANT(SUF, I, S)                                                                 \
  extern "C" void* orc_csr_to_pbr_##SUF(I rows, I cols, const I* indptr, const I* indices,    \
                                        const S* data, I br, I bc, int32_t min_nnz,            \
                                        int32_t order, int64_t* sizes) {                       \
    auto* p = build<I, S>(rows, cols, indptr, indices, data, br, bc, (uint8_t)min_nnz, order); \
    sizes[0] = (int64_t)p->codes.size();                                                       \
    sizes[1] = (int64_t)p->data.size();                                                        \
    sizes[2] = (int64_t)p->rcol.size();                                                        \
    return p;                                                                                  \
  }                                                                                            \
  extern "C" void orc_pbr_get_##SUF(void* h, uint64_t* codes, I* coords, I* offsets, S* data, \
                                    I* rptr, I* rcol, S* rval) {                              \
    auto* p = (Pbr<I, S>*)h;                                                                   \
    std::copy(p->codes.begin(), p->codes.end(), codes);                                       \
    std::copy(p->coords.begin(), p->coords.end(), coords);                                     \
    std::copy(p->offsets.begin(), p->offsets.end(), offsets);                                  \
    std::copy(p->data.begin(), p->data.end(), data);                                           \
    std::copy(p->rptr.begin(), p->rptr.end(), rptr);                                           \
    std::copy(p->rcol.begin(), p->rcol.end(), rcol);                                           \
    std::copy(p->rval.begin(), p->rval.end(), rval);                                           \
  }                                                                                            \
  extern "C" void orc_pbr_free_##SUF(void* h) { delete (Pbr<I, S>*)h; }                       \
  extern "C" void orc_pbr_spmm_##SUF(I rows, I cols, I br, I bc, int64_t T,                   \
                                     const uint64_t* codes, const I* coords, const S* data,    \
                                     const I* rptr, const I* rcol, const S* rval, const S* X,  \
                                     int64_t K, S* Y, int32_t wide) {                          \
    if (wide)                                                                                  \
      spmm<I, S, double>(rows, cols, br, bc, T, codes, coords, data, rptr, rcol, rval, X, K,   \
                         Y);                                                                   \
    else                                                                                       \
      spmm<I, S, S>(rows, cols, br, bc, T, codes, coords, data, rptr, rcol, rval, X, K, Y);    \
  }                                                                                            \
  extern "C" int orc_pbr_to_csr_##SUF(I rows, I cols, I br, I bc, int64_t T,                  \
                                      const uint64_t* codes, const I* coords, const S* data,   \
                                      const I* rptr, const I* rcol, const S* rval, I* indptr,  \
                                      I* indices, S* out, int32_t strict) {                    \
    return to_csr<I, S>(rows, cols, br, bc, T, codes, coords, data, rptr, rcol, rval, indptr,  \
                        indices, out, strict);                                                 \
  }

ORC_VARIANT(i32_f32, int32_t, float)
ORC_VARIANT(i64_f32, int64_t, float)
ORC_VARIANT(i32_f64, int32_t, double)
ORC_VARIANT(i64_f64, int64_t, double)

// Bucket count that std::unordered_map<int32_t, int> reports after being constructed with
// hint n — lets tests pin the hashtable model of the emission order against the library.
extern "C" int64_t orc_bucket_count_after(int64_t hint, const int32_t* keys, int64_t nkeys,
                                          int64_t* counts_after_each) {
  std::unordered_map<int32_t, int> m((size_t)hint);
  for (int64_t i = 0; i < nkeys; ++i) {
    m.emplace(keys[i], 0);
    if (counts_after_each) counts_after_each[i] = (int64_t)m.bucket_count();
  }
  return (int64_t)m.bucket_count();
}

// Bucket count after every emplace when the keys of consecutive stripes are inserted into one
// std::unordered_map<I, int>(hint) that is clear()ed between stripes (the reference builder's
// sequence of container operations, pbr_matrix.cpp:374-428).
template <typename I>
static void bucket_trace(int64_t hint, const I* keys, const int64_t* sptr, int64_t ns,
                         int64_t* counts) {
  std::unordered_map<I, int> m((size_t)hint);
  for (int64_t s = 0; s < ns; ++s) {
    for (int64_t i = sptr[s]; i < sptr[s + 1]; ++i) {
      m.emplace(keys[i], 0);
      counts[i] = (int64_t)m.bucket_count();
    }
    m.clear();
  }
}
extern "C" void orc_bucket_trace_i32(int64_t hint, const int32_t* keys, const int64_t* sptr,
                                     int64_t ns, int64_t* counts) {
  bucket_trace<int32_t>(hint, keys, sptr, ns, counts);
}
extern "C" void orc_bucket_trace_i64(int64_t hint, const int64_t* keys, const int64_t* sptr,
                                     int64_t ns, int64_t* counts) {
  bucket_trace<int64_t>(hint, keys, sptr, ns, counts);
}
