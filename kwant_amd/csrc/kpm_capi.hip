// C ABI of libkpm_amd: plans, argument validation, error reporting.
//
// The Python host (kwant_amd/_engine.py) binds these through ctypes; every
// device buffer is a caller-owned torch tensor, passed as a raw pointer.

#include <stdio.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

#include "kpm_internal.h"

namespace kpm {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return KPM_OK;
  return fail(KPM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace kpm

using namespace kpm;

struct kpm_plan {
  int32_t n_rows = 0, width = 0, value_type = 0, nseg = 0;
  KernelShape shape{};
  int32_t nblocks = 0, nblocks_launch = 0;
  int32_t step = 0;  // the next kpm_cheb_steps launch computes a_{step+1}
  int2* blk_rows = nullptr;
  int64_t* blk_stride = nullptr;
  int32_t* blk_seg = nullptr;
  int32_t* seg_blk = nullptr;
  double* scale = nullptr;
  double* part = nullptr;   // two buffers of part_half doubles (step parity)
  int64_t part_half = 0;
  int32_t* im_exp = nullptr;    // nseg x width grid exponents of the exact Im sums
  double* grid_dots = nullptr;  // nseg x width x 4: <a_0|a_0> for the grid
  bool scale_set = false;
  bool grid_set = false;
  std::vector<int32_t> seg_rows;  // segment row boundaries (host copy)
};

namespace {

template <typename T>
int dev_alloc(T** p, size_t count, const char* what) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(KPM_ERR_ALLOC, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
  }
  return KPM_OK;
}

template <typename T>
int upload(T* d, const std::vector<T>& h, const char* what) {
  return check_hip(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), what);
}

void free_plan(kpm_plan* p) {
  if (!p) return;
  (void)hipFree(p->blk_rows);
  (void)hipFree(p->blk_stride);
  (void)hipFree(p->blk_seg);
  (void)hipFree(p->seg_blk);
  (void)hipFree(p->scale);
  (void)hipFree(p->part);
  (void)hipFree(p->im_exp);
  (void)hipFree(p->grid_dots);
  delete p;
}

int fill_params(const kpm_plan* plan, StepParams& p) {
  p.blk_rows = plan->blk_rows;
  p.blk_stride = plan->blk_stride;
  p.chunk_iters = plan->shape.chunk;
  p.blk_seg = plan->blk_seg;
  p.seg_blk = plan->seg_blk;
  p.scale = plan->scale;
  p.part = plan->part;
  p.im_exp = plan->grid_set ? plan->im_exp : nullptr;
  p.prev_part = nullptr;
  p.prev_dots = nullptr;
  p.nblocks = plan->nblocks;
  p.nblocks_launch = plan->nblocks_launch;
  p.nseg = plan->nseg;
  p.xcd = kNumXcd;
  return KPM_OK;
}

size_t elem_bytes(int value_type) { return value_type == KPM_C128 ? 16 : 8; }

}  // namespace

namespace kpm {

int plan_state(const kpm_plan* plan, PlanState* out) {
  if (!plan || !out) return fail(KPM_ERR_ARG, "plan is NULL");
  out->n_rows = plan->n_rows;
  out->width = plan->width;
  out->value_type = plan->value_type;
  out->n_segments = plan->nseg;
  out->step = plan->step;
  out->scale_set = plan->scale_set;
  out->grid_set = plan->grid_set;
  out->scale = plan->scale;
  out->im_exp = plan->im_exp;
  return KPM_OK;
}

int plan_advance(kpm_plan* plan, int by) {
  if (!plan) return fail(KPM_ERR_ARG, "plan is NULL");
  plan->step += by;
  return KPM_OK;
}

}  // namespace kpm

extern "C" {

const char* kpm_last_error(void) { return g_last_error.c_str(); }

int kpm_abi_version(void) { return 107; }

int kpm_device_info(int32_t device, int32_t* num_cu, int32_t* gfx_arch) {
  hipDeviceProp_t prop;
  int rc = check_hip(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  if (rc) return rc;
  if (num_cu) *num_cu = prop.multiProcessorCount;
  if (gfx_arch) {
    int arch = 0;
    if (sscanf(prop.gcnArchName, "gfx%x", &arch) == 1) {
      // "gfx950" parses as hex 0x950; report the decimal-looking digits.
      char buf[16];
      snprintf(buf, sizeof(buf), "%x", arch);
      arch = atoi(buf);
    }
    *gfx_arch = arch;
  }
  return KPM_OK;
}

int kpm_variant_count(int32_t width, int32_t value_type, int32_t* count) {
  if (!count) return fail(KPM_ERR_ARG, "count is NULL");
  KernelShape ks[kMaxVariants];
  *count = variants_for(width, value_type, ks, kMaxVariants);
  return KPM_OK;
}

int kpm_variant_info(int32_t width, int32_t value_type, int32_t variant, int32_t* lanes_per_row,
                     int32_t* vec, int32_t* waves, int32_t* rounds, int32_t* unroll,
                     int32_t* chunk, int32_t* spare) {
  KernelShape ks[kMaxVariants];
  const int n = variants_for(width, value_type, ks, kMaxVariants);
  if (variant < 0 || variant >= n) return fail(KPM_ERR_ARG, "variant out of range");
  if (lanes_per_row) *lanes_per_row = ks[variant].lpr;
  if (vec) *vec = ks[variant].vec;
  if (waves) *waves = ks[variant].waves;
  if (rounds) *rounds = ks[variant].rounds;
  if (unroll) *unroll = ks[variant].unroll;
  if (chunk) *chunk = ks[variant].chunk;
  if (spare) *spare = ks[variant].spare;
  return KPM_OK;
}

int kpm_plan_create(int32_t n_rows, int32_t width, int32_t value_type, int32_t n_segments,
                    const int32_t* seg_row_begin, int32_t variant, kpm_plan** out) {
  if (!out) return fail(KPM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (value_type != KPM_C64 && value_type != KPM_C128)
    return fail(KPM_ERR_ARG, "value_type must be KPM_C64 or KPM_C128");
  if (n_rows <= 0) return fail(KPM_ERR_ARG, "n_rows must be positive");
  if (n_segments <= 0 || !seg_row_begin) return fail(KPM_ERR_ARG, "need at least one segment");
  if (seg_row_begin[0] != 0 || seg_row_begin[n_segments] != n_rows)
    return fail(KPM_ERR_ARG, "segments must cover rows [0, n_rows)");
  for (int s = 0; s < n_segments; ++s)
    if (seg_row_begin[s + 1] <= seg_row_begin[s])
      return fail(KPM_ERR_ARG, "every segment needs at least one row");
  KernelShape ks[kMaxVariants];
  const int nv = variants_for(width, value_type, ks, kMaxVariants);
  if (nv == 0) return fail(KPM_ERR_UNSUPPORTED, "block width must be a power of two <= 128 "
                                                "(<= 64 for complex128)");
  if (variant < 0) variant = 0;
  if (variant >= nv) return fail(KPM_ERR_ARG, "variant out of range");

  kpm_plan* p = new kpm_plan();
  p->n_rows = n_rows;
  p->width = width;
  p->value_type = value_type;
  p->nseg = n_segments;
  p->shape = ks[variant];
  p->seg_rows.assign(seg_row_begin, seg_row_begin + n_segments + 1);
  int occ = 0, cus = 0, dev = 0;
  int rc0 = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (!rc0)
    rc0 = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                    "hipDeviceGetAttribute");
  if (!rc0) rc0 = blocks_per_cu(value_type, p->shape, &occ);
  if (rc0) {
    delete p;
    return rc0;
  }
  occ -= p->shape.spare;
  if (occ < 1) occ = 1;
  const int64_t rows_per_iter = int64_t(p->shape.waves) * 64 / p->shape.lpr;
  const int64_t target = int64_t(cus) * occ * p->shape.rounds;
  int64_t rpb = (int64_t(n_rows) + target - 1) / target;
  // blocks of a multiple of 4 x rows per wavefront iteration: measured 4 %
  // faster on the graphene block than the finest split
  const int64_t align = rows_per_iter * 4;
  rpb = (rpb + align - 1) / align * align;
  if (rpb < align) rpb = align;
  p->shape.iters = int(rpb / rows_per_iter);

  std::vector<int2> blk_rows;
  std::vector<int32_t> blk_seg, seg_blk;
  for (int s = 0; s < n_segments; ++s) {
    seg_blk.push_back(int32_t(blk_rows.size()));
    const int r0 = seg_row_begin[s], r1 = seg_row_begin[s + 1];
    for (int64_t r = r0; r < r1; r += rpb) {
      blk_rows.push_back(make_int2(int(r), int(r + rpb < r1 ? r + rpb : r1)));
      blk_seg.push_back(s);
    }
  }
  seg_blk.push_back(int32_t(blk_rows.size()));
  p->nblocks = int32_t(blk_rows.size());
  p->nblocks_launch = (p->nblocks + kNumXcd - 1) / kNumXcd * kNumXcd;

  // contiguous workgroups: one chunk each
  std::vector<int64_t> blk_stride(blk_rows.size(), int64_t(n_rows));
  if (p->shape.chunk > 0) {
    // Interleaved: the workgroups of one segment on one XCD (a run of
    // consecutive logical blocks, see the kernel's XCD remap) share the run's
    // rows in chunks of C rows, dealt round-robin, so the run sweeps its
    // range as one front G*C rows wide.
    const int64_t C = rows_per_iter * p->shape.chunk;
    const int per = p->nblocks_launch / kNumXcd;
    int64_t max_chunks = 1;
    for (int b0 = 0; b0 < p->nblocks;) {
      int b1 = b0 + 1;
      while (b1 < p->nblocks && blk_seg[b1] == blk_seg[b0] && b1 / per == b0 / per) ++b1;
      const int64_t r_begin = blk_rows[b0].x, r_end = blk_rows[b1 - 1].y;
      const int64_t g = b1 - b0;
      for (int64_t k = 0; k < g; ++k) {
        const int64_t first = r_begin + k * C < r_end ? r_begin + k * C : r_end;
        blk_rows[b0 + k] = make_int2(int(first), int(r_end));
        blk_stride[b0 + k] = g * C;
        const int64_t chunks = (r_end - first + g * C - 1) / (g * C);
        if (chunks > max_chunks) max_chunks = chunks;
      }
      b0 = b1;
    }
    p->shape.iters = int(max_chunks * p->shape.chunk);
  }

  const int maxnc = 4;  // doubling: 2 TwoSum accumulators + 2 grid sums
  int rc = KPM_OK;
  if (!rc) rc = dev_alloc(&p->blk_rows, blk_rows.size(), "blk_rows");
  if (!rc) rc = dev_alloc(&p->blk_stride, blk_stride.size(), "blk_stride");
  if (!rc) rc = dev_alloc(&p->blk_seg, blk_seg.size(), "blk_seg");
  if (!rc) rc = dev_alloc(&p->seg_blk, seg_blk.size(), "seg_blk");
  if (!rc) rc = dev_alloc(&p->scale, size_t(2) * n_segments, "scale");
  p->part_half = int64_t(p->nblocks) * maxnc * width * 2;
  if (!rc) rc = dev_alloc(&p->part, size_t(2 * p->part_half), "partials");
  if (!rc) rc = dev_alloc(&p->im_exp, size_t(n_segments) * width, "grid exponents");
  if (!rc) rc = dev_alloc(&p->grid_dots, size_t(n_segments) * width * 4, "grid dots");
  if (!rc) rc = upload(p->blk_rows, blk_rows, "upload blk_rows");
  if (!rc) rc = upload(p->blk_stride, blk_stride, "upload blk_stride");
  if (!rc) rc = upload(p->blk_seg, blk_seg, "upload blk_seg");
  if (!rc) rc = upload(p->seg_blk, seg_blk, "upload seg_blk");
  if (!rc) rc = check_hip(hipDeviceSynchronize(), "plan setup");
  if (rc) {
    free_plan(p);
    return rc;
  }
  *out = p;
  return KPM_OK;
}

int kpm_plan_destroy(kpm_plan* plan) {
  if (!plan) return KPM_OK;
  hipError_t e = hipDeviceSynchronize();
  free_plan(plan);
  return check_hip(e, "plan destroy");
}

int kpm_plan_set_scale(kpm_plan* plan, const double* h_ab) {
  if (!plan || !h_ab) return fail(KPM_ERR_ARG, "NULL argument");
  for (int s = 0; s < plan->nseg; ++s) {
    const double a = h_ab[2 * s];
    if (!(a > 0.0)) return fail(KPM_ERR_ARG, "scale a must be positive and finite");
  }
  int rc = check_hip(hipMemcpy(plan->scale, h_ab, sizeof(double) * 2 * plan->nseg,
                               hipMemcpyHostToDevice),
                     "upload scale");
  if (!rc) plan->scale_set = true;
  return rc;
}

int kpm_plan_set_dots_grid(kpm_plan* plan, const void* a0, kpm_stream_t stream) {
  if (!plan || !a0) return fail(KPM_ERR_ARG, "NULL argument");
  if (!plan->scale_set) return fail(KPM_ERR_ARG, "scale not set (kpm_plan_set_scale)");
  hipStream_t s = (hipStream_t)stream;
  // <a_0|a_0> per segment and column, as Re <w|x> of the projection mode with
  // w = x = a_0 (record slot 0): a TwoSum of non-negative terms, so its
  // correctly rounded value whatever the launch shape
  StepParams p{};
  p.ring = reinterpret_cast<char*>(const_cast<void*>(a0));
  p.slot_bytes = 0;
  p.slots = 1;
  p.w = a0;
  p.dots = plan->grid_dots;
  fill_params(plan, p);
  p.im_exp = nullptr;
  int rc = launch_cheb(plan->value_type, plan->shape, KPM_DOTS_PROJECT, true, 0, 0, p, s);
  if (!rc) rc = launch_reduce(KPM_DOTS_PROJECT, plan->width, p, s);
  const int64_t cols = int64_t(plan->nseg) * plan->width;
  std::vector<double> dots(static_cast<size_t>(cols) * 4);
  std::vector<int32_t> kexp(static_cast<size_t>(cols));
  if (!rc)
    rc = check_hip(hipMemcpyAsync(dots.data(), plan->grid_dots, sizeof(double) * dots.size(),
                                  hipMemcpyDeviceToHost, s),
                   "copy grid dots");
  if (!rc) rc = check_hip(hipStreamSynchronize(s), "grid dots");
  if (rc) return rc;
  for (int64_t c = 0; c < cols; ++c) {
    const double mu0 = dots[size_t(c) * 4];
    if (!std::isfinite(mu0))
      return fail(KPM_ERR_ARG, "start vectors contain NaN or Inf (column " +
                                   std::to_string(c % plan->width) + ")");
    int e = 0;
    if (mu0 > 0.0) (void)frexp(mu0, &e);  // mu0 < 2^e
    kexp[size_t(c)] = 51 - e;
  }
  rc = check_hip(hipMemcpy(plan->im_exp, kexp.data(), sizeof(int32_t) * kexp.size(),
                           hipMemcpyHostToDevice),
                 "upload grid exponents");
  if (!rc) plan->grid_set = true;
  return rc;
}

int kpm_plan_set_step(kpm_plan* plan, int32_t n, kpm_stream_t stream) {
  (void)stream;
  if (!plan) return fail(KPM_ERR_ARG, "plan is NULL");
  if (n < 0) return fail(KPM_ERR_ARG, "step must be non-negative");
  plan->step = n;
  return KPM_OK;
}

int kpm_plan_get_step(kpm_plan* plan, int32_t* n, kpm_stream_t stream) {
  (void)stream;
  if (!plan || !n) return fail(KPM_ERR_ARG, "NULL argument");
  *n = plan->step;
  return KPM_OK;
}

static int check_csr(const kpm_plan* plan, const kpm_csr* h) {
  if (!h) return fail(KPM_ERR_ARG, "matrix is NULL");
  if (h->value_type != plan->value_type)
    return fail(KPM_ERR_ARG, "matrix value type differs from the plan's");
  if (h->n_rows != plan->n_rows || h->n_cols != plan->n_rows)
    return fail(KPM_ERR_ARG, "matrix must be square with the plan's row count");
  if (!h->indptr || (h->nnz > 0 && (!h->indices || !h->values)))
    return fail(KPM_ERR_ARG, "matrix arrays are NULL");
  if (h->nnz < 0 || h->nnz > INT32_MAX) return fail(KPM_ERR_ARG, "nnz out of int32 range");
  return KPM_OK;
}

int kpm_cheb_steps(kpm_plan* plan, const kpm_csr* h, void* ring, int64_t slot_stride,
                   int32_t slots, const void* w, double* d_dots, int32_t out_base,
                   int32_t out_cap, int32_t n_steps, int32_t dots_mode,
                   kpm_stream_t stream) {
  if (!plan) return fail(KPM_ERR_ARG, "plan is NULL");
  int rc = check_csr(plan, h);
  if (rc) return rc;
  if (!plan->scale_set) return fail(KPM_ERR_ARG, "scale not set (kpm_plan_set_scale)");
  if (!ring) return fail(KPM_ERR_ARG, "ring is NULL");
  if (slots < 2) return fail(KPM_ERR_ARG, "ring needs at least two slots");
  if (slot_stride < int64_t(plan->n_rows) * plan->width)
    return fail(KPM_ERR_ARG, "slot_stride smaller than n_rows * width");
  if (dots_mode < KPM_DOTS_NONE || dots_mode > KPM_DOTS_LOCAL_COLS)
    return fail(KPM_ERR_ARG, "unknown dots_mode");
  if (dots_mode != KPM_DOTS_NONE && (!d_dots || out_cap <= 0))
    return fail(KPM_ERR_ARG, "dots output buffer missing");
  if (dots_mode == KPM_DOTS_DOUBLING && !plan->grid_set)
    return fail(KPM_ERR_ARG, "dots grid not set (kpm_plan_set_dots_grid)");
  const bool local = dots_mode == KPM_DOTS_LOCAL || dots_mode == KPM_DOTS_LOCAL_COLS;
  if ((dots_mode == KPM_DOTS_PROJECT || local) && !w)
    return fail(KPM_ERR_ARG, "projection vector missing");
  if (n_steps < 0) return fail(KPM_ERR_ARG, "n_steps must be non-negative");
  if (dots_mode != KPM_DOTS_NONE &&
      (plan->step - out_base < 0 || plan->step + n_steps - out_base > out_cap))
    return fail(KPM_ERR_ARG, "dots slots [step - out_base, + n_steps) outside [0, out_cap)");
  StepParams p{};
  p.indptr = h->indptr;
  p.indices = h->indices;
  p.values = h->values;
  p.ring = reinterpret_cast<char*>(ring);
  p.slot_bytes = slot_stride * int64_t(elem_bytes(plan->value_type));
  p.slots = slots;
  p.w = w;
  fill_params(plan, p);
  // Partials alternate between two buffers by step parity: step n writes
  // buffer n % 2 and reduces step n-1's buffer into its record first, so a
  // step costs one launch; the last step of the call is reduced by a
  // standalone launch.
  const int64_t rec = int64_t(plan->nseg) * plan->width * 4;
  const int64_t half = plan->part_half;
  // site-resolved records: one complex128 per row (or per row and column)
  const int64_t local_rec =
      int64_t(plan->n_rows) * 2 * (dots_mode == KPM_DOTS_LOCAL_COLS ? plan->width : 1);
  p.local_cols = dots_mode == KPM_DOTS_LOCAL_COLS;
  for (int i = 0; i < n_steps; ++i) {
    const int n = plan->step;
    p.part = plan->part + (n & 1) * half;
    p.prev_part = nullptr;
    p.prev_dots = nullptr;
    p.local = local ? d_dots + int64_t(n - out_base) * local_rec : nullptr;
    if (dots_mode != KPM_DOTS_NONE && !local && i > 0) {
      p.prev_part = plan->part + ((n - 1) & 1) * half;
      p.prev_dots = d_dots + int64_t(n - 1 - out_base) * rec;
    }
    rc = launch_cheb(plan->value_type, plan->shape, dots_mode, false, n, h->row_nnz_hint, p,
                     (hipStream_t)stream);
    if (rc) return rc;
    plan->step = n + 1;
  }
  if (dots_mode != KPM_DOTS_NONE && !local && n_steps > 0) {
    const int last = plan->step - 1;
    p.part = plan->part + (last & 1) * half;
    p.dots = d_dots + int64_t(last - out_base) * rec;
    rc = launch_reduce(dots_mode, plan->width, p, (hipStream_t)stream);
  }
  return rc;
}

int kpm_bsr_steps(kpm_plan* plan, const kpm_bsr* h, void* ring, int64_t slot_stride,
                  int32_t slots, double* d_dots, int32_t out_base, int32_t out_cap,
                  int32_t n_steps, int32_t dots_mode, kpm_stream_t stream) {
  if (!plan || !h) return fail(KPM_ERR_ARG, "NULL argument");
  if (h->block_size != 4) return fail(KPM_ERR_UNSUPPORTED, "block size must be 4");
  if (h->value_type != plan->value_type)
    return fail(KPM_ERR_ARG, "matrix value type differs from the plan's");
  if (int64_t(h->n_block_rows) * 4 != plan->n_rows || h->n_block_cols != h->n_block_rows)
    return fail(KPM_ERR_ARG, "block matrix must be square with the plan's rows / 4 blocks");
  if (!h->indptr || (h->nnzb > 0 && (!h->indices || !h->values)))
    return fail(KPM_ERR_ARG, "block matrix arrays are NULL");
  if (h->nnzb < 0 || h->nnzb > INT32_MAX / 16)
    return fail(KPM_ERR_ARG, "nnzb out of range");
  for (int32_t r : plan->seg_rows)
    if (r % 4) return fail(KPM_ERR_ARG, "segment boundaries must be multiples of 4");
  if (!plan->scale_set) return fail(KPM_ERR_ARG, "scale not set (kpm_plan_set_scale)");
  if (!ring) return fail(KPM_ERR_ARG, "ring is NULL");
  if (slots < 2) return fail(KPM_ERR_ARG, "ring needs at least two slots");
  if (slot_stride < int64_t(plan->n_rows) * plan->width)
    return fail(KPM_ERR_ARG, "slot_stride smaller than n_rows * width");
  if (dots_mode != KPM_DOTS_DOUBLING && dots_mode != KPM_DOTS_NONE)
    return fail(KPM_ERR_ARG, "dots_mode must be KPM_DOTS_DOUBLING or KPM_DOTS_NONE");
  if (dots_mode == KPM_DOTS_DOUBLING && (!d_dots || out_cap <= 0))
    return fail(KPM_ERR_ARG, "dots output buffer missing");
  if (dots_mode == KPM_DOTS_DOUBLING && !plan->grid_set)
    return fail(KPM_ERR_ARG, "dots grid not set (kpm_plan_set_dots_grid)");
  if (n_steps < 0) return fail(KPM_ERR_ARG, "n_steps must be non-negative");
  if (dots_mode != KPM_DOTS_NONE &&
      (plan->step - out_base < 0 || plan->step + n_steps - out_base > out_cap))
    return fail(KPM_ERR_ARG, "dots slots [step - out_base, + n_steps) outside [0, out_cap)");
  if (plan->shape.lpr > 16 || plan->shape.waves != 4)
    return fail(KPM_ERR_UNSUPPORTED,
                "block-CSR steps need a 4-wave variant with at most 16 lanes per row");
  StepParams p{};
  p.ring = reinterpret_cast<char*>(ring);
  p.slot_bytes = slot_stride * int64_t(elem_bytes(plan->value_type));
  p.slots = slots;
  fill_params(plan, p);
  const int64_t rec = int64_t(plan->nseg) * plan->width * 4;
  const int64_t half = plan->part_half;
  int rc = KPM_OK;
  for (int i = 0; i < n_steps; ++i) {
    const int n = plan->step;
    p.part = plan->part + (n & 1) * half;
    p.prev_part = nullptr;
    p.prev_dots = nullptr;
    if (dots_mode != KPM_DOTS_NONE && i > 0) {
      p.prev_part = plan->part + ((n - 1) & 1) * half;
      p.prev_dots = d_dots + int64_t(n - 1 - out_base) * rec;
    }
    rc = launch_bsr(plan->shape, dots_mode, n, h, p, (hipStream_t)stream);
    if (rc) return rc;
    plan->step = n + 1;
  }
  if (dots_mode != KPM_DOTS_NONE && n_steps > 0) {
    const int last = plan->step - 1;
    p.part = plan->part + (last & 1) * half;
    p.dots = d_dots + int64_t(last - out_base) * rec;
    rc = launch_reduce(dots_mode, plan->width, p, (hipStream_t)stream);
  }
  return rc;
}

int kpm_block_dots(kpm_plan* plan, const void* x, const void* w, double* d_out,
                   int32_t dots_mode, kpm_stream_t stream) {
  if (!plan) return fail(KPM_ERR_ARG, "plan is NULL");
  if (!x || !d_out) return fail(KPM_ERR_ARG, "NULL buffer");
  if (dots_mode < KPM_DOTS_DOUBLING || dots_mode > KPM_DOTS_LOCAL_COLS)
    return fail(KPM_ERR_ARG, "dots_mode must be DOUBLING, PROJECT, LOCAL or LOCAL_COLS");
  const bool local = dots_mode == KPM_DOTS_LOCAL || dots_mode == KPM_DOTS_LOCAL_COLS;
  if ((dots_mode == KPM_DOTS_PROJECT || local) && !w)
    return fail(KPM_ERR_ARG, "projection vector missing");
  if (!plan->scale_set) return fail(KPM_ERR_ARG, "scale not set (kpm_plan_set_scale)");
  if (dots_mode == KPM_DOTS_DOUBLING && !plan->grid_set)
    return fail(KPM_ERR_ARG, "dots grid not set (kpm_plan_set_dots_grid)");
  StepParams p{};
  p.ring = reinterpret_cast<char*>(const_cast<void*>(x));
  p.slot_bytes = 0;
  p.slots = 1;
  p.w = w;
  p.dots = d_out;
  p.local = local ? d_out : nullptr;
  p.local_cols = dots_mode == KPM_DOTS_LOCAL_COLS;
  fill_params(plan, p);
  int rc = launch_cheb(plan->value_type, plan->shape, dots_mode, true, 0, 0, p,
                       (hipStream_t)stream);
  if (!rc && !local) rc = launch_reduce(dots_mode, plan->width, p, (hipStream_t)stream);
  return rc;
}

int kpm_spmm(const kpm_csr* op, const void* x, int64_t x_stride, void* y, int64_t y_stride,
             int32_t width, int32_t n_slots, kpm_stream_t stream) {
  if (!op || !x || !y) return fail(KPM_ERR_ARG, "NULL argument");
  if (op->n_rows <= 0 || op->n_cols <= 0) return fail(KPM_ERR_ARG, "empty operator");
  if (!op->indptr || (op->nnz > 0 && (!op->indices || !op->values)))
    return fail(KPM_ERR_ARG, "operator arrays are NULL");
  if (n_slots <= 0) return fail(KPM_ERR_ARG, "n_slots must be positive");
  if (n_slots > 1 && (x_stride < int64_t(op->n_cols) * width ||
                      y_stride < int64_t(op->n_rows) * width))
    return fail(KPM_ERR_ARG, "slot strides too small");
  return launch_spmm(op->value_type, width, op, x, x_stride, y, y_stride, n_slots,
                     (hipStream_t)stream);
}

int kpm_bond_records(const kpm_csr* bonds, const void* bra, const void* ket, const void* v,
                     const void* w, int32_t width, int32_t per_column, double* d_out,
                     kpm_stream_t stream) {
  if (!bonds || !bra || !ket || !v || !w || !d_out) return fail(KPM_ERR_ARG, "NULL argument");
  if (bonds->n_rows <= 0 || bonds->n_cols != bonds->n_rows)
    return fail(KPM_ERR_ARG, "bond matrix must be square and non-empty");
  if (!bonds->indptr || (bonds->nnz > 0 && (!bonds->indices || !bonds->values)))
    return fail(KPM_ERR_ARG, "bond matrix arrays are NULL");
  if (bonds->nnz == 0) return KPM_OK;
  return launch_bond(width, bonds, bra, ket, v, w, per_column != 0, d_out, (hipStream_t)stream);
}

int kpm_corr_gemm(const void* A, int64_t lda, const void* B, int64_t ldb, double* mu,
                  int64_t ldc, int32_t m_rows, int32_t n_rows, int64_t k, int32_t precision,
                  int32_t k_chunk, double* work, int64_t work_elems, kpm_stream_t stream) {
  if (!A || !B || !mu) return fail(KPM_ERR_ARG, "NULL argument");
  if (m_rows <= 0 || n_rows <= 0 || k <= 0) return fail(KPM_ERR_ARG, "empty GEMM");
  if (lda < k || ldb < k || ldc < n_rows) return fail(KPM_ERR_ARG, "leading dimension too small");
  const int base = precision & ~KPM_GEMM_VECTOR_ALU;
  if (base < KPM_GEMM_FP64 || base > KPM_GEMM_FP32X3 ||
      (precision & ~(KPM_GEMM_VECTOR_ALU | 3)) != 0)
    return fail(KPM_ERR_ARG, "unknown precision");
  if (work_elems < 0) return fail(KPM_ERR_ARG, "negative workspace size");
  return launch_corr_gemm(A, lda, B, ldb, mu, ldc, m_rows, n_rows, k, precision, k_chunk,
                          work, work_elems, (hipStream_t)stream);
}

}  // extern "C"
