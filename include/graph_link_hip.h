/*
 * graph_link_hip.h — C ABI of the MI355X (gfx950) Personalized-PageRank hot path.
 *
 * Every entry point:
 *   - takes plain device pointers, int64 sizes and an explicit hipStream_t (passed as void*)
 *     (the host SpMM, gl_pbr_spmm_host_*, takes host pointers and no stream);
 *   - returns GL_OK (0) or a GL_ERR_* code, with a message available from gl_last_error();
 *   - never throws across the ABI and never allocates device memory: scratch space is a
 *     caller-provided workspace whose size comes from the matching *_workspace_bytes query;
 *   - may be called from several host threads at once: each call that needs a stream of
 *     its own (the PPR run) leases one that no other call uses until it returns.
 *
 * Matrix layout on the device is the reference's PBR ("packed block representation"):
 *   codes    u64[T]     occupancy bitmask of each BRxBC tile, bit (r*BC + c)
 *   coords   i32[2T]    (row0, col0) of each tile, multiples of BR / BC
 *   offsets  i32[T+1]   prefix of popcounts: tile t owns data[offsets[t] .. offsets[t+1])
 *   data     f32[V]     packed tile values in bit order
 *   rem_*    CSR        indptr i32[rows+1], col i32[R], val f32[R] for the tiles below
 *                       min_nnz_per_block
 * plus one derived index, stripe_ptr i32[num_stripes+1] (tiles of block row s are
 * [stripe_ptr[s], stripe_ptr[s+1]) — tiles are stored stripe-ascending by every builder).
 * Dense right-hand sides X and results Y are row-major [rows][K] with leading dimension ld.
 */
#ifndef GRAPH_LINK_HIP_H
#define GRAPH_LINK_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GL_OK 0
#define GL_ERR_INVALID 1     /* bad argument (shape, pointer, unsupported block size)   */
#define GL_ERR_HIP 2         /* a HIP runtime call or kernel launch failed              */
#define GL_ERR_WORKSPACE 3   /* caller workspace too small                              */
#define GL_ERR_NOTCONVERGED 4 /* not an error code of any call; reserved for callers    */

/* Last error message of the calling thread ("" when none). */
const char* gl_last_error(void);
/* ABI version; bumped on any signature change. */
int32_t gl_abi_version(void);
/* Compute units and LDS bytes of the current device (for launch sizing and reports). */
int gl_device_info(int32_t* num_cus, int32_t* lds_bytes_per_cu, int32_t* clock_khz, int64_t* hbm_bytes);

/* ------------------------------------------------------------------------------------
 * Derived index of a PBR matrix.
 * stripe_ptr[s] = first tile whose coords row0 / block_rows >= s, for s in [0, num_stripes].
 * Replaces nothing in the reference (its kernels launch one block per tile,
 * Backend/graph_link/kernels/pbr_kernel.cu:357-364); computed once per upload.
 * ---------------------------------------------------------------------------------- */
int gl_pbr_stripe_index(int64_t rows, int32_t block_rows, int64_t num_tiles,
                        const int32_t* coords, int32_t* stripe_ptr, void* stream);

/* ------------------------------------------------------------------------------------
 * SpMM  Y = A X  (or Y += A X when accumulate != 0), A in PBR, X/Y row-major [rows][K].
 * Replaces pbr_spmm_cuda_{i32}_{f32} (Backend/graph_link/bindings/bindings_kernels.cu:13-83,
 * bound at :142-145) and the launchers launch_pbr_spmm / launch_csr_spmm
 * (Backend/graph_link/kernels/pbr_kernel.cu:336-396, :460-478).
 *   mode GL_SPMM_FUSED : one stripe-owner kernel does tiles + remainder, plain stores.
 *   mode GL_SPMM_SPLIT : block kernel (b), then the CSR-remainder kernel (c) adding into
 *                        the same rows, both on `stream` (each owns whole rows of Y, so
 *                        there are no atomics; the fused mode runs the two parts side by
 *                        side inside one launch).
 * num_vals = offsets[num_tiles] (tile values), num_rem = rem_indptr[rows] (remainder nnz).
 * kw: RHS columns per launch (1,2,4,8 bit-lane kernel; 16,32,64 feature-lane kernel);
 *     0 lets the library choose. Block sizes 2x2, 4x4, 8x8 (as the reference).
 * ---------------------------------------------------------------------------------- */
#define GL_SPMM_FUSED 0
#define GL_SPMM_SPLIT 1
int gl_pbr_spmm_i32_f32(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                        int64_t num_tiles, int64_t num_vals, int64_t num_rem,
                        const uint64_t* codes, const int32_t* coords,
                        const int32_t* offsets, const float* data, const int32_t* stripe_ptr,
                        const int32_t* rem_indptr, const int32_t* rem_col_ind,
                        const float* rem_vals, const float* X, int64_t ldx, int64_t K,
                        float* Y, int64_t ldy, int32_t accumulate, int32_t mode, int32_t kw,
                        void* stream);

/* Execution layout of an 8x8 PBR matrix, derived once (at upload) from the reference's
 * arrays and passed beside them to gl_pbr_spmm_exec_i32_f32. No counterpart in the
 * reference: the same matrix, re-encoded for the matrix-core tile kernel.
 *   tile_records: gl_pbr_exec_layout_bytes(num_tiles) bytes (16 per tile plus 8 padding
 *                 records): the tile's code with its 8x8 bit matrix transposed, its first X
 *                 row and its value offset;
 *   values_cm:    num_vals floats, each tile's values in column-major bit order. */
int64_t gl_pbr_exec_layout_bytes(int64_t num_tiles);
int gl_pbr_exec_layout(int64_t num_tiles, int64_t num_vals, const uint64_t* codes,
                       const int32_t* coords, const int32_t* offsets, const float* data,
                       void* tile_records, float* values_cm, void* stream);

/* gl_pbr_spmm_i32_f32 with the execution layout (both pointers null: the same as
 * gl_pbr_spmm_i32_f32). With 8x8 tiles, fused mode and 8 or 32 RHS columns per launch, the
 * tiles are multiplied on the matrix cores from the layout; every other case ignores it. */
int gl_pbr_spmm_exec_i32_f32(int64_t rows, int64_t cols, int32_t block_rows,
                             int32_t block_cols, int64_t num_tiles, int64_t num_vals,
                             int64_t num_rem, const uint64_t* codes, const int32_t* coords,
                             const int32_t* offsets, const float* data,
                             const int32_t* stripe_ptr, const int32_t* rem_indptr,
                             const int32_t* rem_col_ind, const float* rem_vals,
                             const void* tile_records, const float* values_cm, const float* X,
                             int64_t ldx, int64_t K, float* Y, int64_t ldy, int32_t accumulate,
                             int32_t mode, int32_t kw, void* stream);

/* Block-tile part alone. Rows with no tile get 0 (or keep Y). Replaces launch_pbr_spmm
 * (Backend/graph_link/kernels/pbr_kernel.cu:336-396). */
int gl_pbr_block_spmm_i32_f32(int64_t rows, int64_t cols, int32_t block_rows,
                              int32_t block_cols, int64_t num_tiles, int64_t num_vals,
                              const uint64_t* codes,
                              const int32_t* coords, const int32_t* offsets, const float* data,
                              const int32_t* stripe_ptr, const float* X, int64_t ldx,
                              int64_t K, float* Y, int64_t ldy, int32_t accumulate, int32_t kw,
                              void* stream);

/* CSR SpMM alone (the PBR remainder, or any CSR matrix).
 * Replaces launch_csr_spmm (Backend/graph_link/kernels/pbr_kernel.cu:404-478). */
int gl_csr_spmm_i32_f32(int64_t rows, int64_t cols, int64_t nnz, const int32_t* indptr,
                        const int32_t* col_ind, const float* vals, const float* X, int64_t ldx,
                        int64_t K, float* Y, int64_t ldy, int32_t accumulate, int32_t kw,
                        void* stream);

/* ------------------------------------------------------------------------------------
 * Personalized PageRank power iteration (fused step + the persistent loop).
 * Replaces run_personalized_pagerank (Backend/graph_link/graph_link/__init__.py:210-271),
 * init_ppr (Backend/graph_link/kernels/ppr_kernel.cu:199-216), ppr_update_normalized
 * (:130-187) and the dangling->seed step missing_mass + ppr_update (:13-117).
 *
 * Column k of X is the PPR vector of seed set k: personalization p_k[i] = seed_w[e] for the
 * entries e in [seed_ptr[k], seed_ptr[k+1]) with seed_idx[e] == i (weights of one set sum
 * to 1; duplicate nodes are allowed and add up). Seed arrays are HOST arrays.
 *   X_{t+1} = alpha * A X_t + alpha * dangling_k + (1 - alpha) p_k
 * dangling_mode: GL_DANGLING_NONE (A is column-stochastic already, as the reference's
 *   normalize_transition_matrix makes it), GL_DANGLING_UNIFORM (mass of rows flagged in
 *   dangling_mask spread as 1/N), GL_DANGLING_PERSONALIZATION (that mass returns to p_k).
 * retire_mode: GL_RETIRE_TOGETHER stops every column when max_k residual_k < tol (the
 *   reference); GL_RETIRE_PER_COLUMN freezes each column at its own residual_k < tol.
 * residual_k = sum_i |X_{t+1}[i,k] - X_t[i,k]| (L1), as the reference.
 * X0: warm start [rows][K] (NULL = seed one-hot start, as the reference; X0 == X_out is a
 *   warm start in place).
 * K (seed sets) may be any positive count; the workspace grows with it.
 * max_iter 0 runs no pass: X_out is the start vector, every column has 0 iterations and is
 *   not converged (the reference's loop with max_iterations=0 returns (X, 0, False)).
 * Outputs: X_out [rows][K] (ld = K); iterations[K], converged[K] (device int32);
 *   history [max_iter][K] residuals (device, may be NULL); *total_iterations (host).
 * graph_iters > 0 captures that many iterations into one hipGraph and polls the device
 *   stop flag once per replay; 0 launches eagerly and polls every iteration.
 * ---------------------------------------------------------------------------------- */
#define GL_DANGLING_NONE 0
#define GL_DANGLING_UNIFORM 1
#define GL_DANGLING_PERSONALIZATION 2
#define GL_RETIRE_TOGETHER 0
#define GL_RETIRE_PER_COLUMN 1

typedef struct gl_ppr_opts {
  float alpha;
  float tol;
  int32_t max_iter;
  int32_t dangling_mode;
  int32_t retire_mode;
  int32_t graph_iters;
  int32_t kw;          /* RHS columns per launch, 0 = library choice */
  int32_t grid_mult;   /* workgroups per CU of the persistent launch, 0 = as many as the
                          device holds of the step kernel at once */
  int32_t extrapolate_every; /* 0 = plain power iteration (the reference); P >= 2 = after
                          passes P, 2P, ... each active column whose residual ratio rho =
                          r_t / r_{t-1} lies in [0.5, 0.97] is extrapolated,
                          X += rho / (1 - rho) (X_t - X_{t-1}) (at most 4x), on the device */
  int32_t single_wg;   /* 0 = library choice: the whole run in one launch of one workgroup
                          when K <= 64 and (num_vals + num_rem) * K <= 2^13 (small graphs,
                          where a pass of the multi-workgroup loop is bound by its launches);
                          1 = always (K <= 64); -1 = never. Same step, residual, retire and
                          extrapolation rules either way; graph_iters, kw and grid_mult apply
                          to the multi-workgroup loop only */
} gl_ppr_opts;

int64_t gl_ppr_workspace_bytes(int64_t rows, int32_t block_rows, int64_t K,
                               int64_t num_seed_entries);

int gl_ppr_run_i32_f32(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                       int64_t num_tiles, int64_t num_vals, int64_t num_rem,
                       const uint64_t* codes, const int32_t* coords,
                       const int32_t* offsets, const float* data, const int32_t* stripe_ptr,
                       const int32_t* rem_indptr, const int32_t* rem_col_ind,
                       const float* rem_vals, int64_t K, const int32_t* seed_ptr,
                       const int32_t* seed_idx, const float* seed_w,
                       const uint8_t* dangling_mask, const float* X0, const gl_ppr_opts* opts,
                       float* X_out, int32_t* iterations, int32_t* converged, float* history,
                       void* workspace, int64_t workspace_bytes, int32_t* total_iterations,
                       void* stream);


/* ------------------------------------------------------------------------------------
 * GPU PBR builder. Replaces csr_to_pbr
 * (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:340-477, wrapper
 * Backend/graph_link/graph_link/__init__.py:137-156). Input: canonical CSR on the device
 * (int32 indptr/indices, rows sorted by column, no duplicates); values are copied as opaque
 * 4- or 8-byte payloads, so float32 and float64 matrices build bit-exactly.
 * Two calls on the same workspace (gl_pbr_build_workspace_bytes(rows, block_rows, nnz):
 * per-stripe counts plus 14 bytes per CSR entry, which carry each entry's block and bit from
 * the count pass to the fill pass so that the fill pass neither reads column indices nor
 * hashes):
 *   gl_pbr_build_count: fills rem_indptr[rows+1] and returns counts[5] = {T, V, R, D, B}
 *                       (tiles, tile values, remainder entries, distinct blocks, stripes
 *                       with more distinct blocks than the hash table holds) on the host;
 *   gl_pbr_build_fill : given T and B (B < 0: unknown), writes codes[T], coords[2T],
 *                       offsets[T+1], data[V], rem_col_ind[R],
 *                       rem_vals[R] in canonical order (block columns ascending inside a
 *                       stripe), stripe_ptr[ns+1] (first tile of each stripe) when not NULL,
 *                       and, when blk_bc != NULL, the distinct-block list
 *                       blk_ptr[ns+1], blk_bc[D], blk_rmin[D], blk_tile[D] used to restore
 *                       the reference's emission order.
 * Block sizes: any block_rows x block_cols <= 64 (the reference's bitset<64> code).
 * ---------------------------------------------------------------------------------- */
int64_t gl_pbr_build_workspace_bytes(int64_t rows, int32_t block_rows, int64_t nnz);
int gl_pbr_build_count(int64_t rows, int64_t cols, int64_t nnz, int32_t block_rows,
                       int32_t block_cols, int32_t min_nnz_per_block, const int32_t* indptr,
                       const int32_t* indices, int32_t* rem_indptr, void* workspace,
                       int64_t workspace_bytes, int64_t* counts, void* stream);
int gl_pbr_build_fill(int64_t rows, int64_t cols, int64_t nnz, int32_t block_rows,
                      int32_t block_cols, int32_t min_nnz_per_block, int64_t num_tiles,
                      int64_t num_big_stripes, const int32_t* indptr,
                      const int32_t* indices, const void* values,
                      int32_t value_bytes, void* workspace, int64_t workspace_bytes,
                      uint64_t* codes, int32_t* coords, int32_t* offsets, void* data,
                      const int32_t* rem_indptr, int32_t* rem_col_ind, void* rem_vals,
                      int32_t* stripe_ptr, int32_t* blk_ptr, int32_t* blk_bc, int8_t* blk_rmin,
                      int32_t* blk_tile, void* stream);
/* Reference emission order: replays the reference builder's std::unordered_map<index_t,...>
 * (bucket hint sqrt(num_col_blocks), inserts in CSR first-occurrence order, clear() per
 * stripe; pbr_matrix.cpp:374-428) on HOST copies of the block list and returns perm[T]
 * (new tile -> canonical tile). index_bytes = 4 or 8 selects the key type. */
int gl_pbr_reference_perm(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                          int32_t index_bytes, const int32_t* blk_ptr, const int32_t* blk_bc,
                          const int8_t* blk_rmin, const int32_t* blk_tile, int32_t* perm);
/* Device gather of tiles into perm order (values move with their tile). */
int64_t gl_pbr_permute_workspace_bytes(int64_t num_tiles);
int gl_pbr_permute_tiles(int64_t num_tiles, const int32_t* perm, const uint64_t* codes,
                         const int32_t* coords, const int32_t* offsets, const void* data,
                         int32_t value_bytes, uint64_t* codes_out, int32_t* coords_out,
                         int32_t* offsets_out, void* data_out, void* workspace,
                         int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------------------
 * Learned warm start (no counterpart in the reference).
 * MLP 12 -> 64 -> 64 -> 1, ReLU, ReLU, softplus; parameters packed as
 * W1[64][12] b1[64] W2[64][64] b2[64] w3[64] b3 (gl_mlp_num_params() floats).
 * Features of (node i, seed set k): degree buckets round(8 log2(1+deg))/8 (in, out), one-hot
 * hop 0..4/more, log1p(N q_t)/8 for the first three power-iteration terms q_1..q_3, and 1.
 * ---------------------------------------------------------------------------------- */
int32_t gl_mlp_num_params(void);
int32_t gl_mlp_num_features(void);
int32_t gl_ws_num_buckets(void);
/* out[m] = softplus(MLP(feat[m])) for feat [M][12]. */
int gl_mlp_forward(const float* params, const float* feat, int64_t M, float* out, void* stream);
/* Floats of the per-block gradient buffer the training calls need. */
int64_t gl_mlp_train_partials(void);
/* One Adam step on the weighted L1 loss sum_m weight[m] |s_m - target[m]| over M samples,
 * gradient multiplied by grad_scale; *loss += the loss. */
int gl_mlp_train_step(float* params, const float* feat, const float* target,
                      const float* weight, int64_t M, float grad_scale, float lr, float* adam_m,
                      float* adam_v, int32_t step, float* partial, float* loss, void* stream);
/* Gradient of the same loss into grad[gl_mlp_num_params()] (no update). */
int gl_mlp_grad(const float* params, const float* feat, const float* target, const float* weight,
                int64_t M, float* partial, float* grad, float* loss, void* stream);
/* In- and out-degree of every node from the PBR matrix (row i = in-edges of i). */
int gl_ws_degrees(int64_t n, int32_t block_rows, int64_t num_tiles, const uint64_t* codes,
                  const int32_t* coords, const int32_t* rem_indptr, const int32_t* rem_col_ind,
                  int32_t* indeg, int32_t* outdeg, void* stream);
/* Features [M][12] of the listed pairs (pair = i*K + k); v5 = 5 device pointers to
 * A^j P, j = 0..4, each [N][K]. */
int gl_ws_pair_features(int64_t N, int64_t K, float alpha, const int32_t* indeg,
                        const int32_t* outdeg, const float* const* v5, const int64_t* pairs,
                        int64_t M, float* feat, void* stream);
/* Features of the (in-bucket, out-bucket, hop class 3/4/more) table, gl_ws_table_size() rows. */
int64_t gl_ws_table_size(void);
int gl_ws_table_features(float* feat, void* stream);
/* Table row of every node: nb[i] = in-bucket * gl_ws_num_buckets() + out-bucket (once per
 * graph, from gl_ws_degrees). */
int gl_ws_node_buckets(int64_t n, const int32_t* indeg, const int32_t* outdeg, int32_t* nb,
                       void* stream);
/* X0 from the table for pairs beyond hop 2; pairs within hop 2 appended to near (count in
 * *near_count on the device, list truncated at near_cap). node_buckets from
 * gl_ws_node_buckets. x4 (may be NULL): also receives the 4th power-iteration iterate from
 * the seed start, rebuilt from the feature products (v5 = A^j P, j = 0..4):
 * x4 = (1 - alpha) (v0 + alpha v1 + alpha^2 v2 + alpha^3 v3) + alpha^4 v4. */
int gl_ws_assemble(int64_t N, int64_t K, float alpha, const int32_t* indeg,
                   const int32_t* outdeg, const float* const* v5, const int32_t* node_buckets,
                   const float* table, float* x0, int64_t* near, int64_t near_cap,
                   uint64_t* near_count, float* x4, void* stream);
/* x0[pairs[m]] = val[m]. */
int gl_ws_scatter(const int64_t* pairs, const float* val, int64_t M, float* x0, void* stream);
/* The blended start: x0 (the network's start, not normalized) <- max(0, x4 + lam (x0 / s - x4))
 * with s its column sums and x4 from gl_ws_assemble, then every column scaled to sum 1: the
 * four products the features needed count as the first four passes, and the network's start
 * moves that iterate a fraction lam of the way. K <= 1024; sums: K * 1026 floats of scratch
 * (the network start's column sums in [0, K), the result's in [K, 2K), fixed order). */
int gl_ws_blend_start(float* x0, const float* x4, int64_t N, int64_t K, float lam, float* sums,
                      void* stream);
/* Scale every column of x [N][K] (K <= 1024) to sum 1. sums: K * 1025 floats of scratch; the
 * first K receive the column sums, taken in a fixed order (repeated calls agree bit for bit). */
int gl_ws_normalize(float* x, int64_t N, int64_t K, float* sums, void* stream);

/* ------------------------------------------------------------------------------------
 * Column normalization of a transition matrix A[target, source] given as CSR on the device
 * (entries in CSR order: rows ascending, columns ascending inside a row). Replaces
 * normalize_transition_matrix (Backend/graph_link/graph_link/__init__.py:184-207) without its
 * dense 1/N block: out_vals[e] = vals[e] * (1 / colsum[indices[e]]), with the float32 column
 * sums taken in the same order as the host's numpy reduction (bit-identical output when no
 * column is dangling); dangling[c] = 1 for a zero column sum (col_sums and dangling may be
 * NULL). The matrix keeps its structure; dangling columns are then handled inside the PPR
 * run (GL_DANGLING_UNIFORM with this mask) instead of densified.
 * ---------------------------------------------------------------------------------- */
int64_t gl_csr_normalize_workspace_bytes(int64_t cols, int64_t nnz);
int gl_csr_normalize_columns_f32(int64_t cols, int64_t nnz, const int32_t* indices,
                                 const float* vals, float* out_vals, float* col_sums,
                                 uint8_t* dangling, void* workspace, int64_t workspace_bytes,
                                 void* stream);

/* PBR -> canonical CSR on the device (tiles in any order, remainder rows in any order).
 * Replaces pbr_to_csr (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:289-338, bound at
 * Backend/graph_link/bindings/bindings.cpp:75-78); unlike it, a last block row that only
 * partly lies inside the matrix is accepted (its out-of-matrix bits must be clear).
 * nnz = entries the output arrays hold (indptr[rows+1], indices[nnz], vals[nnz]); the call
 * fails with GL_ERR_INVALID if the matrix holds a different count, a tile bit lies outside
 * the matrix, or offsets / remainder indices are out of range. Synchronizes `stream` once. */
int64_t gl_pbr_to_csr_workspace_bytes(int64_t rows, int64_t nnz);
int gl_pbr_to_csr_i32_f32(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                          int64_t num_tiles, int64_t num_vals, int64_t num_rem,
                          const uint64_t* codes, const int32_t* coords, const int32_t* offsets,
                          const float* data, const int32_t* rem_indptr,
                          const int32_t* rem_col_ind, const float* rem_vals, int32_t* indptr,
                          int32_t* indices, float* vals, int64_t nnz, void* workspace,
                          int64_t workspace_bytes, void* stream);
/* The same for a float64 matrix (pbr_matrix.cpp:289-338 instantiated for double, bound at
 * bindings.cpp:77-78); values are copied, never computed, so the output is exact. */
int64_t gl_pbr_to_csr_f64_workspace_bytes(int64_t rows, int64_t nnz);
int gl_pbr_to_csr_i32_f64(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                          int64_t num_tiles, int64_t num_vals, int64_t num_rem,
                          const uint64_t* codes, const int32_t* coords, const int32_t* offsets,
                          const double* data, const int32_t* rem_indptr,
                          const int32_t* rem_col_ind, const double* rem_vals, int32_t* indptr,
                          int32_t* indices, double* vals, int64_t nnz, void* workspace,
                          int64_t workspace_bytes, void* stream);

/* Top-k of every column of a dense row-major x [N][ld] (K columns): values [k][K] and
 * rows [k][K] (int64), largest first, equal values by ascending row (deterministic).
 * Replaces the torch.topk over the rank vector of Tests/wiki_ppr.py:187. */
int64_t gl_topk_workspace_bytes(int64_t K, int64_t k);
int gl_topk_columns_f32(const float* x, int64_t N, int64_t K, int64_t ld, int64_t k,
                        float* values, int64_t* rows, void* workspace, int64_t workspace_bytes,
                        void* stream);

/* ------------------------------------------------------------------------------------
 * FP64 variants (float64 PBR matrices), a compatibility path beside the FP32 kernels.
 * ---------------------------------------------------------------------------------- */
/* Y = A X (or Y += A X) for a float64 PBR and float64 X/Y, row-major [rows][K]; the
 * block tiles and the remainder in one launch. Replaces pbr_spmm_cuda_i32_f64
 * (Backend/graph_link/bindings/bindings_kernels.cu:13-83, bound at :144) with its
 * launchers (Backend/graph_link/kernels/pbr_kernel.cu:336-396, :460-478). */
int gl_pbr_spmm_i32_f64(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                        int64_t num_tiles, const uint64_t* codes, const int32_t* coords,
                        const int32_t* offsets, const double* data, const int32_t* stripe_ptr,
                        const int32_t* rem_indptr, const int32_t* rem_col_ind,
                        const double* rem_vals, const double* X, int64_t ldx, int64_t K,
                        double* Y, int64_t ldy, int32_t accumulate, void* stream);

/* ------------------------------------------------------------------------------------
 * The reference's per-iteration PPR kernels as separate calls (FP32 and FP64), for a host
 * loop of SpMM + update per iteration (Backend/graph_link/graph_link/__init__.py:240-260).
 * X, Y are device [N][K] row-major (ld = K). Sums over rows (residuals, column sums) are
 * per-block partials in the workspace, added in block order (the reference adds float
 * atomics), so repeated calls agree bit for bit.
 * ---------------------------------------------------------------------------------- */
/* X [N][K] = one-hot columns, X[i][k] = (i == source_nodes[k]); the whole of X is written.
 * Replaces init_ppr_cuda_float / init_ppr_cuda_double
 * (Backend/graph_link/kernels/ppr_kernel.cu:199-216, bound at bindings_kernels.cu:148-149),
 * which also need X zeroed beforehand. */
int gl_ppr_init_f32(float* X, const int32_t* source_nodes, int64_t N, int64_t K, void* stream);
int gl_ppr_init_f64(double* X, const int32_t* source_nodes, int64_t N, int64_t K,
                    void* stream);

/* Workspace of the update and column-sum calls below (per-block partial sums). */
int64_t gl_ppr_update_f32_workspace_bytes(int64_t N, int64_t K);
int64_t gl_ppr_update_f64_workspace_bytes(int64_t N, int64_t K);

/* X <- alpha Y + (1 - alpha) e_src in place; errors[k] = L1 norm of column k's update.
 * Replaces ppr_update_normalized_cuda_float / _double (ppr_kernel.cu:130-187, bound at
 * bindings_kernels.cu:160-161). */
int gl_ppr_update_normalized_f32(const float* Y, float* X, const int32_t* source_nodes,
                                 float alpha, int64_t N, int64_t K, float* errors,
                                 void* workspace, int64_t workspace_bytes, void* stream);
int gl_ppr_update_normalized_f64(const double* Y, double* X, const int32_t* source_nodes,
                                 double alpha, int64_t N, int64_t K, double* errors,
                                 void* workspace, int64_t workspace_bytes, void* stream);

/* gl_ppr_update_normalized_f32 without a workspace: the per-block partial sums live in a
 * static device array of the library (4 MiB on every GPU it runs on, held until the process
 * ends), so the call allocates nothing. No counterpart in the reference (whose update needs
 * no workspace, but whose errors are float atomics): it is for callers that cannot allocate,
 * such as code under HIP graph capture and tight serving loops. Limit: N * K small enough
 * that the partials fit 4 MiB (any N for K <= 1024); larger sizes return GL_ERR_INVALID.
 * Every call on one GPU shares that GPU's array, so calls that overlap on one GPU (several
 * streams or threads at once) race; calls in sequence on one stream are safe. */
int gl_ppr_update_normalized_f32_noalloc(const float* Y, float* X, const int32_t* source_nodes,
                                         float alpha, int64_t N, int64_t K, float* errors,
                                         void* stream);

/* col_sums[k] = sum_i Y[i][k] (overwritten). Replaces missing_mass_cuda_float / _double
 * (ppr_kernel.cu:13-48, bound at bindings_kernels.cu:152-153). */
int gl_ppr_missing_mass_f32(const float* Y, float* col_sums, int64_t N, int64_t K,
                            void* workspace, int64_t workspace_bytes, void* stream);
int gl_ppr_missing_mass_f64(const double* Y, double* col_sums, int64_t N, int64_t K,
                            void* workspace, int64_t workspace_bytes, void* stream);

/* X_next = alpha Y + [i == src_k] (alpha (1 - col_sums[k]) + (1 - alpha)): the step for a
 * matrix whose dangling columns are empty, their mass returned to the seed (networkx's
 * default). errors[k] += L1 norm of X_next - X_cur (added, as the reference, which leaves
 * zeroing errors to the caller). Replaces ppr_update_cuda_float / _double
 * (ppr_kernel.cu:50-117, bound at bindings_kernels.cu:156-157). */
int gl_ppr_update_f32(const float* Y, const float* X_cur, float* X_next,
                      const int32_t* source_nodes, const float* col_sums, float alpha,
                      int64_t N, int64_t K, float* errors, void* workspace,
                      int64_t workspace_bytes, void* stream);
int gl_ppr_update_f64(const double* Y, const double* X_cur, double* X_next,
                      const int32_t* source_nodes, const double* col_sums, double alpha,
                      int64_t N, int64_t K, double* errors, void* workspace,
                      int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------------------
 * Host (CPU) SpMM: Y = A X with the PBR arrays, X and Y in HOST memory (row-major, leading
 * dimensions ldx / ldy), for the four index / value types of the reference's classes.
 * Replaces pbr_batched_matmul_cpu (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:17-68,
 * bound at Backend/graph_link/bindings/bindings.cpp:65-68), which pbr_matmul uses for CPU
 * tensors (Backend/graph_link/graph_link/__init__.py:130-134). Any block size with
 * block_rows * block_cols <= 64 and any tile order. Every row adds its tiles' products in
 * tile and bit order, then its remainder entries, each product rounded on its own, so Y is
 * bit-identical to the reference's. `threads` host threads (0 = all cores) own disjoint
 * stripe ranges. Needs no GPU.
 * ---------------------------------------------------------------------------------- */
int gl_pbr_spmm_host_i32_f32(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                             int64_t num_tiles, const uint64_t* codes, const int32_t* coords,
                             const int32_t* offsets, const float* data,
                             const int32_t* rem_indptr, const int32_t* rem_col_ind,
                             const float* rem_vals, const float* X, int64_t ldx, int64_t K,
                             float* Y, int64_t ldy, int32_t threads);
int gl_pbr_spmm_host_i64_f32(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                             int64_t num_tiles, const uint64_t* codes, const int64_t* coords,
                             const int64_t* offsets, const float* data,
                             const int64_t* rem_indptr, const int64_t* rem_col_ind,
                             const float* rem_vals, const float* X, int64_t ldx, int64_t K,
                             float* Y, int64_t ldy, int32_t threads);
int gl_pbr_spmm_host_i32_f64(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                             int64_t num_tiles, const uint64_t* codes, const int32_t* coords,
                             const int32_t* offsets, const double* data,
                             const int32_t* rem_indptr, const int32_t* rem_col_ind,
                             const double* rem_vals, const double* X, int64_t ldx, int64_t K,
                             double* Y, int64_t ldy, int32_t threads);
int gl_pbr_spmm_host_i64_f64(int64_t rows, int64_t cols, int32_t block_rows, int32_t block_cols,
                             int64_t num_tiles, const uint64_t* codes, const int64_t* coords,
                             const int64_t* offsets, const double* data,
                             const int64_t* rem_indptr, const int64_t* rem_col_ind,
                             const double* rem_vals, const double* X, int64_t ldx, int64_t K,
                             double* Y, int64_t ldy, int32_t threads);

#ifdef __cplusplus
}
#endif

#endif /* GRAPH_LINK_HIP_H */
