/*
 * kpm_amd.h -- C ABI of the MI355X (gfx950) Chebyshev moment engine.
 *
 * Every buffer argument named d_* or documented as "device" is a device
 * pointer that the caller owns (in the Python host these come from
 * torch.Tensor.data_ptr()); the library only borrows it for the duration of
 * the call.  Plans own their workspaces (block partials, counters, the step
 * counter) and nothing else.  All entry points return KPM_OK (0) on success
 * and a nonzero status otherwise; kpm_last_error() then describes the failure
 * (thread-local).  Streams are hipStream_t passed as void*; NULL is the
 * default stream.  No entry point synchronises the stream unless stated.
 *
 * Complex values are interleaved (re, im): KPM_C64 = 2 x float,
 * KPM_C128 = 2 x double.  Vector blocks are row-major N x W ("W" = block
 * width = number of random vectors handled together), so one matrix row of a
 * block is W contiguous complex numbers.
 *
 * The reference this replaces is kwant's pure-Python kpm module; its only
 * "FFI" is numpy/scipy, so each entry point below names the kwant/kpm.py
 * lines whose numerical work it performs.
 */
#ifndef KPM_AMD_H
#define KPM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KPM_OK 0
#define KPM_ERR_ARG 1         /* bad argument: maps to ValueError      */
#define KPM_ERR_DEVICE 2      /* HIP runtime failure: RuntimeError      */
#define KPM_ERR_UNSUPPORTED 3 /* no kernel variant for this shape       */
#define KPM_ERR_ALLOC 4       /* device allocation failed               */

#define KPM_C64 1
#define KPM_C128 2

/* dots_mode of kpm_cheb_steps / kpm_block_dots */
#define KPM_DOTS_NONE 0     /* recursion only (Correlator expansions)         */
#define KPM_DOTS_DOUBLING 1 /* <a_n|a_n>, Re <a_{n+1}|a_n> per column         */
#define KPM_DOTS_PROJECT 2  /* <w|a_n>, <w|a_{n+1}> per column (operator path) */
#define KPM_DOTS_LOCAL 3    /* per row: sum_c conj(w[r,c]) a_{n+1}[r,c]       */
#define KPM_DOTS_LOCAL_COLS 4 /* per row and column: conj(w[r,c]) a_{n+1}[r,c] */

typedef void* kpm_stream_t;

/* A CSR matrix in device memory (borrowed).  indptr has n_rows+1 int32
 * entries, indices nnz int32 entries, values nnz complex entries of
 * value_type.  This is the csr_matrix that kwant/kpm.py:181 builds from the
 * int32-indexed COO of kwant/_system.pyx:118-119.
 * row_nnz_hint: a row length that nearly all rows stay within (for example
 * the 99th percentile), 0 if unknown.  It only selects how many entries per
 * row the recursion kernel requests ahead; results never depend on it. */
typedef struct {
  int32_t n_rows;
  int32_t n_cols;
  int64_t nnz;
  const int32_t* indptr;
  const int32_t* indices;
  const void* values;
  int32_t value_type;
  int32_t row_nnz_hint;
} kpm_csr;

/* The same matrix in block-CSR with 4 x 4 blocks (the norbs x norbs site
 * blocks kwant/_system.pyx:94-153 assembles for 4-orbital sites): indptr has
 * n_block_rows+1 int32 entries, indices nnzb int32 block columns, values
 * nnzb x 4 x 4 complex entries of value_type, each block row-major.  Zeros
 * inside a block are stored.  block_size must be 4. */
typedef struct {
  int32_t n_block_rows;
  int32_t n_block_cols;
  int64_t nnzb;
  const int32_t* indptr;
  const int32_t* indices;
  const void* values;
  int32_t value_type;
  int32_t block_size;
} kpm_bsr;

typedef struct kpm_plan kpm_plan;

/* Last error message of the calling thread ("" if none). */
const char* kpm_last_error(void);

/* ABI version (major*100 + minor). */
int kpm_abi_version(void);

/* Compute units and gfx architecture number (e.g. 950) of a device. */
int kpm_device_info(int32_t device, int32_t* num_cu, int32_t* gfx_arch);

/* Number of recursion kernel variants for a (width, value_type) pair, and
 * the launch shape of each: lanes per matrix row, complex values per lane,
 * waves per workgroup, the residency rounds the grid is sized for (rows per
 * workgroup = rows / (CUs x resident workgroups per CU x rounds)), the row
 * slots each wavefront processes per iteration (2: the pipelined loop for
 * 4 to 16 lanes per row), the interleave chunk (0: each workgroup owns
 * contiguous rows; c: the workgroups of an XCD deal a range out in chunks of
 * c wavefront iterations) and the resident workgroups per CU the grid leaves
 * spare.  Used by the host autotuner. */
int kpm_variant_count(int32_t width, int32_t value_type, int32_t* count);
int kpm_variant_info(int32_t width, int32_t value_type, int32_t variant,
                     int32_t* lanes_per_row, int32_t* vec,
                     int32_t* waves, int32_t* rounds, int32_t* unroll,
                     int32_t* chunk, int32_t* spare);

/* Create a recursion plan for an N x width vector block whose rows are split
 * into n_segments independent systems (segment s owns rows
 * [seg_row_begin[s], seg_row_begin[s+1]) -- host array).  One segment is the
 * plain SpectralDensity case; several are a batched disorder ensemble whose
 * block-diagonal matrix is passed as one CSR.  Each segment gets its own
 * rescaling (a, b) and its own moments.  variant selects the kernel launch
 * shape (see kpm_variant_info); -1 picks the default.
 * Replaces the per-system setup of kwant/kpm.py:163-252. */
int kpm_plan_create(int32_t n_rows, int32_t width, int32_t value_type,
                    int32_t n_segments, const int32_t* seg_row_begin,
                    int32_t variant, kpm_plan** out);
int kpm_plan_destroy(kpm_plan* plan);

/* Rescaling of each segment: host array of n_segments (a, b) pairs so the
 * recursion applies (H v - b v) / a  (kwant/kpm.py:1174-1175, 1182-1183).
 * Synchronous upload. */
int kpm_plan_set_scale(kpm_plan* plan, const double* h_ab);

/* Grid of the exact sums of Im <a_{n+1}|a_n> (KPM_DOTS_DOUBLING), fixed from
 * the block's start vectors a0 (n_rows x width device block, the plan's value
 * type): per segment and column, with 2^E > <a0|a0>, every term is cut into
 * integers on the grid 2^(E-73) and summed exactly, so the imaginary part of
 * mu_{2n+1} (kwant/kpm.py:470-473), rounding noise of an exactly-zero sum,
 * comes out the same for every launch shape, block width and resumed run --
 * and kwant's real_if_close (kwant/kpm.py:379) sees a nonzero imaginary part
 * as it does with numpy's vdot.  Required before doubling-mode steps; call it
 * once per block, before the first step (the grid must not change while a
 * block is extended).  Fails with KPM_ERR_ARG when a0 holds NaN or Inf.
 * Synchronises the stream (reads <a0|a0> back). */
int kpm_plan_set_dots_grid(kpm_plan* plan, const void* a0, kpm_stream_t stream);

/* Step counter n of the plan (host side): the next kpm_cheb_steps launch
 * computes a_{n+1}.  The stream argument is unused (kept for symmetry). */
int kpm_plan_set_step(kpm_plan* plan, int32_t n, kpm_stream_t stream);
int kpm_plan_get_step(kpm_plan* plan, int32_t* n, kpm_stream_t stream);

/* Run n_steps Chebyshev steps of the three-term recursion over the vector
 * block (kwant/kpm.py:436-484 with the rescaled matvec of :1182-1183):
 *   n == 0 : a_1     = (H a_0 - b a_0) / a
 *   n >= 1 : a_{n+1} = 2 (H a_n - b a_n) / a - a_{n-1}
 * a_k lives in ring slot (k mod slots): device buffer `ring`, slot k at
 * ring + k*slot_stride complex elements, each slot an n_rows x width block.
 * With slots == 2 the update is in place (a_{n+1} overwrites a_{n-1}).
 * The step index n comes from the plan's counter, which each step advances.
 * Each launch writes one (sum, error) partial per workgroup and column; the
 * next step's launch first reduces the previous step's partials (in
 * workgroup order) into that step's record, and one reduction kernel after
 * the last step of the call finishes the final record.
 *
 * dots_mode KPM_DOTS_DOUBLING writes, per step, segment and column c,
 *   d_dots[((n - out_base) * n_segments + s) * width + c] = 4 doubles
 *   { <a_n|a_n>, 0, Re <a_{n+1}|a_n>, Im <a_{n+1}|a_n> }
 * (the two vdot()s of kwant/kpm.py:444-445, 470-473; at n == 0 the second
 * is <a_1|a_0>, the conjugate of mu_1).  <a_n|a_n> is real (numpy's vdot of
 * a vector with itself has a zero imaginary part); Im <a_{n+1}|a_n> is the
 * exact sum on the grid of kpm_plan_set_dots_grid (zero for Hermitian H in
 * exact arithmetic, rounding noise in practice).  KPM_DOTS_PROJECT writes
 * { <w|a_n>, <w|a_{n+1}> } with w the device block `w` (the matrix-operator
 * moments vdot(a_0, O a_n) of kwant/kpm.py:193-195, 484 with w = O^H a_0).
 * Sums are FP64 and deterministic (fixed order), so repeated runs and
 * resumed runs are bitwise identical.
 *
 * The site-resolved operator path (kwant.operator.Density with sum=False,
 * kwant/operator.pyx:771-801, evaluated as op(a_0, a_{n+1}) by
 * kwant/kpm.py:484) writes per matrix row instead of per column, with
 * w = M^H a_0 for the block-diagonal onsite matrix M, as complex128
 * (re, im) pairs and without workgroup partials:
 *   KPM_DOTS_LOCAL:      d_dots[((n - out_base) * n_rows + r) * 2 + {0, 1}]
 *                        = sum_c conj(w[r,c]) a_{n+1}[r,c]  (the block's
 *                        vectors summed, for mean=True; a TwoSum sum
 *                        rounded once)
 *   KPM_DOTS_LOCAL_COLS: d_dots[(((n - out_base) * n_rows + r) * width + c) * 2
 *                        + {0, 1}] = conj(w[r,c]) a_{n+1}[r,c]  (mean=False)
 * The caller sums the rows of each site.
 *
 * The call fails with KPM_ERR_ARG (nothing launched) when a step's slot
 * (n - out_base) would fall outside [0, out_cap). */
int kpm_cheb_steps(kpm_plan* plan, const kpm_csr* h, void* ring,
                   int64_t slot_stride, int32_t slots, const void* w,
                   double* d_dots, int32_t out_base, int32_t out_cap,
                   int32_t n_steps, int32_t dots_mode, kpm_stream_t stream);

/* kpm_cheb_steps with the matrix in 4 x 4 block-CSR: the same recursion
 * (kwant/kpm.py:436-473 with the rescaled matvec of :1182-1183), the four
 * rows of a site gathered as one block.  Modes KPM_DOTS_DOUBLING and
 * KPM_DOTS_NONE; the plan's rows and segment boundaries must be multiples
 * of 4 and its variant a 4-wave shape with at most 16 lanes per row.  With
 * the blocks' columns in order the moments are bitwise those of the
 * column-sorted CSR (the padding zeros add exact zeros). */
int kpm_bsr_steps(kpm_plan* plan, const kpm_bsr* h, void* ring, int64_t slot_stride,
                  int32_t slots, double* d_dots, int32_t out_base, int32_t out_cap,
                  int32_t n_steps, int32_t dots_mode, kpm_stream_t stream);

/* Two Chebyshev steps per launch over tiles of consecutive rows (one
 * segment).  The same recursion as kpm_cheb_steps (kwant/kpm.py:464-473 with
 * the rescaled matvec of kwant/kpm.py:1182-1183), a_{n+1} and a_{n+2} from
 * one launch: a workgroup computes a_{n+1} of its tile and of the tile's halo
 * (the rows outside the tile that its entries reach) into LDS, then a_{n+2}
 * of the tile from LDS, so the vectors cross HBM four times per two steps
 * instead of six.  Every row of a_{n+1}, a_{n+2} and every dot is bitwise
 * what two kpm_cheb_steps launches give.  Tiles pay off when the row order
 * makes them compact patches of the lattice (few halo rows per tile row).
 *
 * kpm_tiled_create builds the tiling of the device CSR matrix h (the same
 * matrix later passed to kpm_cheb_steps2): tile t owns rows
 * [tile_row_begin[t], tile_row_begin[t+1]) (host array of n_tiles + 1), at
 * most the variant's max_tile_rows each.  It reads the pattern back, computes
 * each tile's halo and the LDS slot of every entry's column, and packs the
 * first entries of every row (columns, LDS slots, values) so that a row's
 * entries need no row-extent read first; it fails with KPM_ERR_UNSUPPORTED
 * when a tile and its halo do not fit in LDS.  The packed values are a copy:
 * build the tiling again after changing the matrix.  variant selects the
 * launch shape (kpm_tiled_variant_info: columns per workgroup, waves per
 * workgroup, the largest tile); -1 picks the default.  The tiling holds its
 * own partials workspace; use one tiling per plan at a time.  Synchronous.
 * Replaces the per-vector loop of kwant/kpm.py:436-477 for the plain DOS
 * path, like kpm_cheb_steps. */
typedef struct kpm_tiled kpm_tiled;
int kpm_tiled_variant_count(int32_t width, int32_t value_type, int32_t* count);
int kpm_tiled_variant_info(int32_t width, int32_t value_type, int32_t variant,
                           int32_t* group_width, int32_t* waves, int32_t* max_tile_rows);
int kpm_tiled_create(const kpm_csr* h, int32_t width, int32_t n_tiles,
                     const int32_t* h_tile_row_begin, int32_t variant, kpm_tiled** out);
int kpm_tiled_destroy(kpm_tiled* tiles);
/* Tiles, the largest tile, the largest halo, all halo rows and their matrix
 * entries, the rows of a_n the tiles gather (each tile's rows, its halo and
 * the halo's own neighbours, summed over tiles), the persistent workgroups of
 * a launch and the LDS bytes each uses. */
int kpm_tiled_info(const kpm_tiled* tiles, int32_t* n_tiles, int32_t* max_tile_rows,
                   int32_t* max_halo_rows, int64_t* halo_rows, int64_t* halo_entries,
                   int64_t* gather_rows, int32_t* workgroups, int32_t* lds_bytes);

/* n_pairs launches of two steps each, from the plan's step counter n >= 1
 * (run step 0, and an odd last step, with kpm_cheb_steps), advancing it by
 * 2 n_pairs.  The ring needs at least four slots (a_{n-1} .. a_{n+2} are read
 * and written by one launch; the halo rows of one tile are other tiles' rows,
 * so nothing is updated in place).  dots_mode KPM_DOTS_DOUBLING (records as
 * kpm_cheb_steps, two per launch) or KPM_DOTS_NONE.  Plan: one segment, the
 * tiling's row count, width and value type, scale (and, for the doubling
 * mode, dots grid) set. */
int kpm_cheb_steps2(kpm_plan* plan, kpm_tiled* tiles, const kpm_csr* h, void* ring,
                    int64_t slot_stride, int32_t slots, double* d_dots, int32_t out_base,
                    int32_t out_cap, int32_t n_pairs, int32_t dots_mode, kpm_stream_t stream);

/* The steps of kpm_cheb_steps for one system (one segment) without a plan
 * and without allocating: no counterpart in the reference; for callers that
 * cannot call hipMalloc (code under HIP graph capture, tight serving loops).
 * The workspace (workgroup partials, workgroup row ranges, the rescaling) is
 * a fixed 16 MiB device array per GPU inside the library, filled by kernels
 * on `stream`: a call is a fixed sequence of kernel launches with no host
 * synchronisation and no host-to-device copy.  The call fails with
 * KPM_ERR_UNSUPPORTED when its workspace exceeds that array (about
 * 2 x blocks x 4 x width x 16 bytes, blocks ~ CUs x resident workgroups x
 * residency rounds of the variant).  Every call on a GPU shares the array:
 * calls that overlap on one GPU race (serialise them on one stream).
 * Arguments as kpm_cheb_steps, with the plan's state passed explicitly:
 * (a, b) the rescaling, width and variant the plan's, first_step the step
 * counter n of the first step.  grid_exp: device array of width int32, the
 * grid of the exact Im <a_{n+1}|a_n> sums (kpm_plan_set_dots_grid): written
 * from ring slot 0 (a_0) when first_step == 0, read otherwise -- keep it
 * with the ring between calls; required in the doubling mode.  Unlike
 * kpm_plan_set_dots_grid it does not report NaN/Inf in a_0 (that would need
 * a synchronisation); such moments come out NaN.  Asynchronous. */
int kpm_cheb_steps_static(const kpm_csr* h, double a, double b, int32_t width,
                          int32_t variant, void* ring, int64_t slot_stride, int32_t slots,
                          const void* w, int32_t* grid_exp, int32_t first_step,
                          int32_t n_steps, double* d_dots, int32_t out_base, int32_t out_cap,
                          int32_t dots_mode, kpm_stream_t stream);

/* Dots of one block without a matrix product: per segment and column,
 * DOUBLING: { <x|x>, 0, 0, 0 }; PROJECT: { <w|x>, 0, 0 }.  Written to
 * d_out[s * width + c] (4 doubles).  This is the odd-moment tail
 * vdot(alpha_next, alpha_next) of kwant/kpm.py:474-477.  LOCAL and
 * LOCAL_COLS write the per-row record of kpm_cheb_steps for x in place of
 * a_{n+1} (the first site-resolved moment op(a_0, a_0), kwant/kpm.py:447). */
int kpm_block_dots(kpm_plan* plan, const void* x, const void* w,
                   double* d_out, int32_t dots_mode, kpm_stream_t stream);

/* Bond records of a current-like operator (kwant.operator.Current,
 * kwant/operator.pyx:940-1001, evaluated as op(a_0, a_n) by
 * kwant/kpm.py:484): for every entry k of the square CSR `bonds` (row r,
 * column s, value h -- the Hamiltonian entries of the hoppings asked for)
 * and every column c of the n_rows x width blocks,
 *   t_k(c) = i (conj(h) conj(bra[s,c]) v[r,c] - conj(w[r,c]) h ket[s,c])
 * with v = M ket and w = M^H bra for the block-diagonal onsite matrix M
 * (v = ket, w = bra when M = 1).  per_column == 0: d_out[k] = sum_c t_k(c)
 * (complex128, the vectors of the block summed for mean=True); otherwise
 * d_out[k * width + c] = t_k(c).  The caller sums the entries of each
 * hopping.  All blocks have the bond matrix's value type. */
int kpm_bond_records(const kpm_csr* bonds, const void* bra, const void* ket,
                     const void* v, const void* w, int32_t width,
                     int32_t per_column, double* d_out, kpm_stream_t stream);

/* y = scale * op x for an n_rows x width block (op: n_rows x n_cols CSR of the
 * same value type; x: n_cols x width; y: n_rows x width), for n_slots
 * consecutive slots (x + k*x_stride, y + k*y_stride).  This is the operator
 * application op.dot(v) of kwant/kpm.py:1098-1099, used for
 * Psi = op2 T_n(H)|r> (kwant/kpm.py:708-718) and Omega_0 = op1|r>
 * (kwant/kpm.py:697-706). */
int kpm_spmm(const kpm_csr* op, const void* x, int64_t x_stride, void* y,
             int64_t y_stride, int32_t width, int32_t n_slots,
             kpm_stream_t stream);

/* Correlator moment matrix (kwant/kpm.py:685-695):
 *   mu[m, n] += sum_k conj(A[m, k]) * B[n, k]
 * A: m_rows x k complex, row stride lda (complex elements); B: n_rows x k,
 * row stride ldb; mu: FP64 complex m_rows x n_rows, row stride ldc,
 * accumulated in place.  precision:
 *   KPM_GEMM_FP64  : complex128 inputs, FP64 MFMA
 *   KPM_GEMM_FP32  : complex64 inputs, FP32 MFMA, FP32 partial sums over
 *                    k-chunks of k_chunk, then FP64 accumulation
 *   KPM_GEMM_FP32X3: complex128 inputs split hi+lo into FP32, three FP32 MFMA
 *                    products (hi*hi + hi*lo + lo*hi), FP64 accumulation
 *                    across k-chunks.  The FP32 MFMA rounds products and its
 *                    accumulator to 24 bits, so this is no more accurate than
 *                    KPM_GEMM_FP32 (measured: 2.6e-7 relative either way);
 *                    retired from the Python API, kept here for ABI
 *                    stability.  kpm_corr_gemm_sliced is the split-precision
 *                    mode that works.
 * The workgroups take their output tiles in XCD-local groups (4 x 8 tiles of
 * 64 per XCD at a time, the same k range), so each operand chunk is fetched
 * from HBM about once per XCD.
 * k is split into slices to fill the GPU when a workspace is given: work is a
 * device FP64 buffer of work_elems doubles (2 * m_rows * n_rows per slice;
 * NULL = no split).  Slices are added in order (deterministic).
 * Asynchronous on stream. */
#define KPM_GEMM_FP64 0
#define KPM_GEMM_FP32 1
#define KPM_GEMM_FP32X3 2
/* OR-ed into precision: use the vector-ALU (FMA) kernel instead of the
 * matrix cores -- a cross-check of the MFMA path. */
#define KPM_GEMM_VECTOR_ALU 16
int kpm_corr_gemm(const void* A, int64_t lda, const void* B, int64_t ldb,
                  double* mu, int64_t ldc, int32_t m_rows, int32_t n_rows,
                  int64_t k, int32_t precision, int32_t k_chunk,
                  double* work, int64_t work_elems, kpm_stream_t stream);

/* The same product, mu[m, n] += sum_k conj(A[m, k]) B[n, k], from
 * complex128 A and B on the BF16 matrix cores, exact in every term kept:
 * per panel of k and per operand row, each real component is cut into
 * `slices` (1..4) integer slices of <= 8 bits against the row's largest
 * exponent; slice products and their level sums over 256 k' are exact in the
 * FP32 accumulator and are folded into FP64.  The error is the slice tail and
 * the dropped cross levels: about 2^(-8 slices) of max|A_m| max|B_n| per term
 * (slices = 4: FP64-class accuracy for well-scaled rows).  work: device
 * buffer of at least kpm_corr_gemm_sliced_work() bytes.  Asynchronous. */
int kpm_corr_gemm_sliced_work(int32_t m_rows, int32_t n_rows, int64_t k, int32_t slices,
                              int64_t* bytes);
int kpm_corr_gemm_sliced(const void* A, int64_t lda, const void* B, int64_t ldb,
                         double* mu, int64_t ldc, int32_t m_rows, int32_t n_rows,
                         int64_t k, int32_t slices, void* work, int64_t work_bytes,
                         kpm_stream_t stream);

/* The same product (kwant/kpm.py:685-695) on the INT8 matrix cores
 * (v_mfma_i32_32x32x32_i8), after Ootomo, Ozaki and Yokota (2024): each real
 * component is cut into `slices` (1..6) integer slices of <= 7 bits
 * (|q| <= 64) against its row's largest exponent; slice products are summed
 * exactly in int32 over at most 65536 k' per workgroup and folded into FP64
 * once.  The error is about 2^(-7 slices) of max|A_m| max|B_n| per term
 * (slices = 5: below the 4-slice BF16 form's).  work: device buffer of at
 * least kpm_corr_gemm_islice_work() bytes.  Asynchronous. */
int kpm_corr_gemm_islice_work(int32_t m_rows, int32_t n_rows, int64_t k, int32_t slices,
                              int64_t* bytes);
int kpm_corr_gemm_islice(const void* A, int64_t lda, const void* B, int64_t ldb,
                         double* mu, int64_t ldc, int32_t m_rows, int32_t n_rows,
                         int64_t k, int32_t slices, void* work, int64_t work_bytes,
                         kpm_stream_t stream);

/* The same product with B cut into slices once and reused: the Correlator
 * multiplies Psi (kwant/kpm.py:708-718) by every chunk of Omega
 * (kwant/kpm.py:685-695).  kpm_islice_prep_rows slices rows
 * [row0, row0 + rows) of an n_rows x k operand (B points at row row0) into
 * `prep` (kpm_islice_prep_bytes bytes); rows may be filled in any order and
 * call, each row independently (its own exponent per k-panel).  Then
 * kpm_corr_gemm_islice_prep multiplies an m_rows x k A by all n_rows prepared
 * rows: mu (m_rows x n_rows) += conj(A) B^T, with a work buffer of
 * kpm_corr_gemm_islice_prep_work bytes.  Results are bitwise those of
 * kpm_corr_gemm_islice on the same operands.  Asynchronous. */
int kpm_islice_prep_bytes(int32_t n_rows, int64_t k, int32_t slices, int64_t* bytes);
int kpm_islice_prep_rows(const void* B, int64_t ldb, int32_t row0, int32_t rows,
                         int32_t n_rows, int64_t k, int32_t slices, void* prep,
                         int64_t prep_bytes, kpm_stream_t stream);
int kpm_corr_gemm_islice_prep_work(int32_t m_rows, int32_t n_rows, int64_t k,
                                   int32_t slices, int64_t* bytes);
int kpm_corr_gemm_islice_prep(const void* A, int64_t lda, int32_t m_rows, const void* prep,
                              int64_t prep_bytes, int32_t n_rows, int64_t k, int32_t slices,
                              double* mu, int64_t ldc, void* work, int64_t work_bytes,
                              kpm_stream_t stream);

/* Exact moments after a uniform on-site shift (no counterpart in kwant/kpm.py;
 * it acts on the moments kwant/kpm.py:436-477 produces).  nu holds moments
 * <T_k(t)> in a window t = (E - b') / a'; the moments in the window
 * y = alpha t + beta (for H + mu: alpha = a'/a, beta = (b' + mu - b)/a) are
 *   mu_n = out_scale[n] * sum_{k<=n} C_nk in_scale[k] nu_k,
 *   T_n(alpha t + beta) = sum_k C_nk T_k(t)   (three-term recurrence in FP64).
 * Element (z, p, k) of nu is nu[z*nu_bs + p*nu_rs + k], likewise out; beta is a
 * device array of rows doubles; in_scale / out_scale are device arrays of
 * num_moments doubles or NULL (1).  Requires |alpha| + |beta[p]| <= 1 (the
 * caller checks beta): then |C_nk| <= 2 and the recurrence is stable.
 * num_moments <= 1024.  Asynchronous on stream. */
int kpm_moment_shift(const float* nu, int64_t nu_bs, int64_t nu_rs, int32_t batch,
                     int64_t rows, int32_t num_moments, const double* in_scale,
                     const double* out_scale, double alpha, const double* beta,
                     float* out, int64_t out_bs, int64_t out_rs, kpm_stream_t stream);

/* ---- cubic 4-orbital disorder ensemble generator --------------------------
 * Replaces the host assembly kwant would do through hamiltonian_submatrix
 * (kwant/kpm.py:177-184 -> kwant/_system.pyx:233-338) for the batched
 * ensemble: a system of side L (4 L^3 orbitals, sites s = (x L + y) L + z,
 * rows 4 s + o) with random Hermitian on-site blocks A_s = (X + X^H)/2,
 * X_ij = sigma (g + i g'), g standard normals from the counter-based stream of
 * `seed`, and hopping t * I_4 to the 6 nearest neighbours (open boundaries).
 * kpm_cubic4_nnz gives its entry count, 40 L^3 - 24 L^2.  kpm_gen_cubic4
 * writes the system's rows [row0, row0 + 4L^3) of a block-diagonal CSR whose
 * entries for it start at nz0 (indices are global: row0 + local column), and
 * raises *radius_bits (a non-negative float stored as its bits, zeroed by the
 * caller) to at least the Gershgorin radius max_r sum_j |H_rj| (spectrum in
 * [-radius, radius]).  Asynchronous on stream. */
int kpm_cubic4_nnz(int32_t L, int64_t* nnz);
int kpm_gen_cubic4(int32_t L, uint64_t seed, double sigma, double t, int64_t row0,
                   int64_t nz0, int32_t* indptr, int32_t* indices, void* values,
                   int32_t value_type, unsigned int* radius_bits, kpm_stream_t stream);

/* The same system's on-site blocks alone, for the lattice-stencil recursion:
 * onsite[(row0 + r) * 4 + j] = A_s[o, j] (complex of value_type) for the
 * system's rows r = 4 s + o -- bitwise the values kpm_gen_cubic4 writes --
 * and the same Gershgorin radius.  Asynchronous on stream. */
int kpm_gen_cubic4_onsite(int32_t L, uint64_t seed, double sigma, double t, int64_t row0,
                          void* onsite, int32_t value_type, unsigned int* radius_bits,
                          kpm_stream_t stream);

/* The Chebyshev steps of an ensemble of such systems without a stored matrix
 * (the lattice stencil): kpm_cheb_steps (kwant/kpm.py:464-473 with the
 * rescaled matvec of kwant/kpm.py:1182-1183) for the block-diagonal matrix
 * of the systems of sides h_sides[0 .. n_systems), stacked in that order,
 * with on-site blocks `onsite` (n_rows x 4 complex, kpm_gen_cubic4_onsite)
 * and hopping t I_4.  A workgroup marches columns (y, z, orbital) of one
 * system along x, keeping a_n of the x neighbours in registers; the row sums
 * take the CSR's entries in its column order, so vectors and dots are bitwise
 * those of kpm_cheb_steps on kpm_gen_cubic4's CSR.  kpm_march_create
 * builds the work list (waves: 4 or 8 per workgroup; depth: planes, 1 or 2,
 * that the streams are requested ahead, 3 or 4 the same with each plane's
 * a_n exchanged through LDS inside the workgroup; tile_y, 1 to 1024: each
 * workgroup takes waves * 16 / (lanes per row) consecutive sites of a plane
 * in band order -- bands of tile_y rows along y, z-major inside a band --
 * with lanes per row width / 2 for KPM_C64, width for KPM_C128);
 * kpm_cheb_steps_march
 * runs n_steps steps from the plan's counter (the plan: one segment per
 * system, this row count, width and value type; scale, and for
 * KPM_DOTS_DOUBLING the dots grid, set).  Dots as kpm_cheb_steps; modes
 * KPM_DOTS_DOUBLING and KPM_DOTS_NONE. */
typedef struct kpm_march kpm_march;
int kpm_march_create(int32_t n_systems, const int32_t* h_sides, int32_t width,
                     int32_t value_type, int32_t waves, int32_t depth, int32_t tile_y,
                     kpm_march** out);
int kpm_march_destroy(kpm_march* m);
int kpm_cheb_steps_march(kpm_plan* plan, kpm_march* m, const void* onsite, double t,
                         void* ring, int64_t slot_stride, int32_t slots, double* d_dots,
                         int32_t out_base, int32_t out_cap, int32_t n_steps, int32_t dots_mode,
                         kpm_stream_t stream);

/* ---- moment surrogate (no counterpart in kwant/kpm.py) ------------------
 * A small MLP that maps Hamiltonian parameters to Jackson-damped moments,
 * trained on moments from kpm_cheb_steps and evaluated in place of the
 * recursion inside parameter sweeps.  All buffers are FP32 device memory. */

/* Batched strided GEMM, z = 0 .. batch-1:
 *   C[z][m, n] = sum_k A[z*a_bs + m*a_rs + k*a_cs] * B[z*b_bs + n*b_rs + k*b_cs]
 *                (+ bias[z*bias_bs + n])
 * operands rounded to BF16, FP32 accumulation on the matrix cores.  flags =
 * one epilogue | optional modifiers:
 *   KPM_MLP_PLAIN     : C = acc + bias
 *   KPM_MLP_BIAS_GELU : C = Z = acc + bias, C2 = gelu(Z) (exact erf form),
 *                       C2 with the strides of C
 *   KPM_MLP_GELU_GRAD : C = acc * gelu'(aux[z*aux_bs + m*aux_rs + n])
 *   KPM_MLP_BF16X3    : split each operand x = hi + lo (both BF16) and form
 *                       hi*hi + hi*lo + lo*hi (about 16 significant bits)
 *   KPM_MLP_ACCUMULATE: C += acc + bias (plain epilogue only)
 * C element (z, m, n) is at c + z*c_bs + m*c_rs + n.  bias may be NULL.
 * Asynchronous on stream. */
#define KPM_MLP_PLAIN 0
#define KPM_MLP_BIAS_GELU 1
#define KPM_MLP_GELU_GRAD 2
#define KPM_MLP_BF16X3 16
#define KPM_MLP_ACCUMULATE 32
typedef struct kpm_mlp_gemm_desc {
  const float* a;
  int64_t a_rs, a_cs, a_bs;
  const float* b;
  int64_t b_rs, b_cs, b_bs;
  float* c;
  int64_t c_rs, c_bs;
  float* c2;
  const float* bias;
  int64_t bias_bs;
  const float* aux;
  int64_t aux_rs, aux_bs;
  int32_t m, n, k, batch;
  int32_t flags;
} kpm_mlp_gemm_desc;
int kpm_mlp_gemm(const kpm_mlp_gemm_desc* desc, kpm_stream_t stream);
/* out[z*out_bs + n] = sum_m A[z*a_bs + m*lda + n] (bias gradients), rows
 * summed in order. */
int kpm_mlp_colsum(const float* A, int64_t lda, int64_t a_bs, int32_t M, int32_t N,
                   int32_t batch, float* out, int64_t out_bs, kpm_stream_t stream);
/* dy = scale * (y - t) over n elements; partial[b] = sum over workgroup b of
 * (y - t)^2 in FP64 (n_partial workgroups; the loss is their sum). */
int kpm_mlp_mse(const float* y, const float* t, float* dy, int64_t n, float scale,
                double* partial, int32_t n_partial, kpm_stream_t stream);
/* Ensemble statistics per row p of y[z*z_stride + p*row_stride + n]
 * (members z < members, n < width):
 *   mean[p*mean_stride + n] = average over members           (mean may be NULL)
 *   dev[p] = (1/width) sum_n (sample standard deviation over members)
 *                                                             (dev may be NULL)
 * With y = densities at Chebyshev nodes, dev is the L1 norm of the ensemble's
 * pointwise standard deviation: the surrogate's spread (fallback gate).  dev
 * needs members >= 2 (KPM_ERR_ARG otherwise). */
int kpm_mlp_ensemble(const float* y, int64_t z_stride, int64_t row_stride,
                     int32_t members, int64_t rows, int32_t width, float* mean,
                     int64_t mean_stride, float* dev, kpm_stream_t stream);
/* One Adam update of n parameters at step t >= 1 (bias-corrected).  With
 * sched != NULL (device memory: lr, 1/(1-b1^t), 1/(1-b2^t), ema decay) lr, t
 * and ema_decay are ignored and the kernel reads them when it runs, so one
 * captured graph can replay many steps of a schedule.  With ema != NULL the
 * updated parameters are folded into ema = d ema + (1 - d) p. */
int kpm_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                 float b1, float b2, float eps, int32_t t, const float* sched,
                 float* ema, float ema_decay, kpm_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* KPM_AMD_H */
