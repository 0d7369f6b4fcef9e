// Internal declarations shared by the HIP translation units of libkpm_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/kpm_amd.h"

namespace kpm {

// Number of XCDs of an MI355X; workgroups are dispatched round-robin over
// them, so logical block ids are remapped to give each XCD a contiguous range
// of rows (its own L2 then sees the neighbour rows of its gathers).
constexpr int kNumXcd = 8;

// Workgroup -> (k split, output tile) of the moment-matrix GEMMs.
// Workgroups are dealt round-robin to the XCDs, so hardware block hb is
// remapped to logical block (hb mod 8) * (blocks / 8) + hb / 8: each XCD owns
// a contiguous run of logical blocks, split-major (blocks must be a multiple
// of 8).  Within a split the tiles are ordered in groups of 4 row tiles x 8
// column tiles when the tile grid allows, so the 32 workgroups an XCD holds
// at once read 4 tiles' rows of A and 8 tiles' rows of B over the same k
// range: each operand chunk comes from HBM once per XCD and is reused from
// its L2 by the group.  Logical blocks past mt x nt x splits give split >=
// the split count (grid padding).
constexpr int kGemmGroupM = 4, kGemmGroupN = 8;

__device__ inline void gemm_tile_of(int hb, int nblocks, int mt, int nt, int* m_tile,
                                    int* n_tile, int* split) {
  const int per = nblocks / kNumXcd;
  const int lb = (hb % kNumXcd) * per + hb / kNumXcd;
  const int tiles = mt * nt;
  *split = lb / tiles;
  const int t = lb - *split * tiles;
  if (mt % kGemmGroupM == 0 && nt % kGemmGroupN == 0) {
    const int g = t / (kGemmGroupM * kGemmGroupN), w = t - g * (kGemmGroupM * kGemmGroupN);
    const int groups_n = nt / kGemmGroupN;
    *m_tile = (g / groups_n) * kGemmGroupM + w / kGemmGroupN;
    *n_tile = (g % groups_n) * kGemmGroupN + w % kGemmGroupN;
  } else {
    *m_tile = t / nt;
    *n_tile = t % nt;
  }
}

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

struct KernelShape {
  int lpr;    // lanes per matrix row
  int vec;    // complex values per lane
  int waves;  // wavefronts per workgroup
  int rounds; // residency rounds the grid is sized for
  int unroll; // row slots per wave iteration (scalar-metadata shapes)
  int chunk;  // 0: contiguous rows per workgroup; c > 0: the workgroups of an
              // XCD interleave chunks of c wavefront iterations (rounds 1)
  int spare;  // resident workgroups per CU left unused by the grid
  int iters;  // row iterations per workgroup (set by the plan)
};

// Launch-side parameters of the recursion kernel (passed by value).
struct StepParams {
  const int32_t* indptr;
  const int32_t* indices;
  const void* values;
  char* ring;
  int64_t slot_bytes;
  int32_t slots;
  const void* w;
  double* dots;            // this step's record: nseg x W x 4 doubles
  const int2* blk_rows;    // logical block -> [first row, row_end)
  const int64_t* blk_stride; // logical block -> rows from one chunk to the next
  int32_t chunk_iters;     // wavefront iterations per chunk (0: one chunk)
  const int32_t* blk_seg;  // logical block -> segment
  const int32_t* seg_blk;  // segment -> first logical block (nseg + 1)
  const double* scale;     // segment -> (a, b)
  double* part;            // this launch's block partials, (sum, error) pairs
  const double* prev_part; // previous step's partials to reduce first (or null)
  double* prev_dots;       // ... into this record (nseg x W x 4 doubles)
  const int32_t* im_exp;   // segment x column -> k = 51 - E, the grid 2^-k of the
                           // exact Im <a_{n+1}|a_n> sums (doubling mode; null: 0)
  double* local;           // site-resolved record of this step (KPM_DOTS_LOCAL*)
  int32_t local_cols;      // 1: one value per row and column, 0: per row
  int32_t nblocks;
  int32_t nblocks_launch;
  int32_t nseg;
  int32_t xcd;
};

// Launchers implemented in kpm_cheb.hip.  mode: dots mode; n: the step (the
// launch computes a_{n+1}); only_dots: no matrix product, dots of ring slot n
// (kpm_block_dots).  A dots-mode launch writes workgroup partials to p.part
// and, when p.prev_part is set, first reduces those of the previous step into
// p.prev_dots; launch_reduce reduces p.part into p.dots on its own.
// row_nnz_hint: the matrix's typical row length (kpm_csr), 0 if unknown.
int launch_cheb(int value_type, const KernelShape& ks, int mode, bool only_dots, int n,
                int row_nnz_hint, const StepParams& p, hipStream_t stream);
int launch_reduce(int mode, int width, const StepParams& p, hipStream_t stream);
// One block-CSR step (4 x 4 blocks): the kernel shape's lanes per row and
// complex values per lane, its wavefront iterations (ks.iters), mode
// KPM_DOTS_DOUBLING or KPM_DOTS_NONE.
int launch_bsr(const KernelShape& ks, int mode, int n, const kpm_bsr* h, const StepParams& p,
               hipStream_t stream);
int launch_bond(int width, const kpm_csr* b, const void* bra, const void* ket, const void* v,
                const void* w, int per_column, double* out, hipStream_t stream);
int launch_spmm(int value_type, int width, const kpm_csr* op, const void* x,
                int64_t x_stride, void* y, int64_t y_stride, int n_slots,
                hipStream_t stream);
// At most kMaxVariants launch shapes per (width, value type).
constexpr int kMaxVariants = 96;
int variants_for(int width, int value_type, KernelShape* out, int max_out);
int blocks_per_cu(int value_type, const KernelShape& ks, int* out);

// The state of a plan that the tiled two-step launches (kpm_tiled.hip)
// share with kpm_cheb_steps: shape, rescaling, dots grid and step counter.
struct PlanState {
  int32_t n_rows, width, value_type, n_segments, step;
  bool scale_set, grid_set;
  const double* scale;     // device (a, b) per segment
  const int32_t* im_exp;   // device grid exponents per segment and column
};
int plan_state(const kpm_plan* plan, PlanState* out);
// Advance the plan's step counter by `by` steps.
int plan_advance(kpm_plan* plan, int by);

int launch_corr_gemm(const void* A, int64_t lda, const void* B, int64_t ldb,
                     double* mu, int64_t ldc, int m_rows, int n_rows, int64_t k,
                     int precision, int k_chunk, double* work, int64_t work_elems,
                     hipStream_t stream);

}  // namespace kpm
