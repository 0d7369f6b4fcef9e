// kpm_cheb_steps_static: the Chebyshev steps of kpm_cheb_steps for one
// system without a plan and without any device allocation.
//
// kpm_cheb_steps keeps its workspace -- the workgroup partials of the dot
// products, the row range of every workgroup, the rescaling -- in a plan that
// kpm_plan_create allocates.  Callers that cannot allocate (code under HIP
// graph capture, tight serving loops) use this entry point instead: the same
// kernels run with that workspace in `scratch`, a file-scope device array of
// kStaticBytes.  The workgroup layout is written into it by a small kernel from
// scalar arguments (no host-to-device copy), so a call is a fixed sequence of
// kernel launches on the caller's stream.  The grid of the exact
// Im <a_{n+1}|a_n> sums lives in a caller buffer (grid_exp) that travels with
// the ring between calls.  Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU race, and the array keeps the last call's
// contents.

#include <mutex>

#include "kpm_internal.h"

namespace kpm {

constexpr int64_t kStaticBytes = int64_t(16) << 20;  // 16 MiB per GPU

static __device__ __align__(256) char scratch[kStaticBytes];

namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` on each GPU (the host symbol is a stand-in, not
// a device address), looked up once per device with that device current.
char* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

int scratch_of_current_device(char** out) {
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices) return fail(KPM_ERR_UNSUPPORTED, "device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    if (rc) return rc;
    g_scratch[dev] = static_cast<char*>(ptr);
  }
  *out = g_scratch[dev];
  return KPM_OK;
}

// Layout of one segment of n_rows rows in nblocks workgroups of rpb rows
// (the same split kpm_plan_create makes): contiguous ranges, or, with an
// interleave of C rows, the workgroups of each XCD run (per consecutive
// blocks) dealing their run's rows out in chunks of C.
struct StaticLayout {
  int64_t n_rows, rpb, C;
  int32_t nblocks, per;
};

__host__ __device__ inline void static_block(const StaticLayout& L, int b, int64_t* first,
                                             int64_t* end, int64_t* stride) {
  if (L.C == 0) {
    *first = int64_t(b) * L.rpb;
    *end = *first + L.rpb < L.n_rows ? *first + L.rpb : L.n_rows;
    *stride = L.n_rows;
    return;
  }
  const int b0 = b / L.per * L.per;
  const int b1 = b0 + L.per < L.nblocks ? b0 + L.per : L.nblocks;
  const int64_t r_begin = int64_t(b0) * L.rpb;
  const int64_t r_end = int64_t(b1) * L.rpb < L.n_rows ? int64_t(b1) * L.rpb : L.n_rows;
  const int64_t g = b1 - b0;
  const int64_t f = r_begin + int64_t(b - b0) * L.C;
  *first = f < r_end ? f : r_end;
  *end = r_end;
  *stride = g * L.C;
}

__global__ void static_layout_kernel(StaticLayout L, int2* blk_rows, int64_t* blk_stride,
                                     int32_t* blk_seg, int32_t* seg_blk, double* scale,
                                     double a, double b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < L.nblocks) {
    int64_t first, end, stride;
    static_block(L, i, &first, &end, &stride);
    blk_rows[i] = make_int2(int(first), int(end));
    blk_stride[i] = stride;
    blk_seg[i] = 0;
  }
  if (i == 0) {
    seg_blk[0] = 0;
    seg_blk[1] = L.nblocks;
    scale[0] = a;
    scale[1] = b;
  }
}

// k = 51 - e with <a_0|a_0> < 2^e per column (0 for a zero column), as
// kpm_plan_set_dots_grid computes it on the host.
__global__ void static_grid_kernel(const double* dots, int width, int32_t* grid_exp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= width) return;
  const double mu0 = dots[int64_t(c) * 4];
  int e = 0;
  if (mu0 > 0.0 && isfinite(mu0)) (void)frexp(mu0, &e);
  grid_exp[c] = 51 - e;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace
}  // namespace kpm

using namespace kpm;

extern "C" int kpm_cheb_steps_static(const kpm_csr* h, double a, double b, int32_t width,
                                     int32_t variant, void* ring, int64_t slot_stride,
                                     int32_t slots, const void* w, int32_t* grid_exp,
                                     int32_t first_step, int32_t n_steps, double* d_dots,
                                     int32_t out_base, int32_t out_cap, int32_t dots_mode,
                                     kpm_stream_t stream) {
  if (!h) return fail(KPM_ERR_ARG, "matrix is NULL");
  if (h->value_type != KPM_C64 && h->value_type != KPM_C128)
    return fail(KPM_ERR_ARG, "value_type must be KPM_C64 or KPM_C128");
  if (h->n_rows <= 0 || h->n_cols != h->n_rows)
    return fail(KPM_ERR_ARG, "matrix must be square and non-empty");
  if (!h->indptr || (h->nnz > 0 && (!h->indices || !h->values)))
    return fail(KPM_ERR_ARG, "matrix arrays are NULL");
  if (h->nnz < 0 || h->nnz > INT32_MAX) return fail(KPM_ERR_ARG, "nnz out of int32 range");
  if (!(a > 0.0)) return fail(KPM_ERR_ARG, "scale a must be positive and finite");
  if (!ring) return fail(KPM_ERR_ARG, "ring is NULL");
  if (slots < 2) return fail(KPM_ERR_ARG, "ring needs at least two slots");
  if (slot_stride < int64_t(h->n_rows) * width)
    return fail(KPM_ERR_ARG, "slot_stride smaller than n_rows * width");
  if (dots_mode < KPM_DOTS_NONE || dots_mode > KPM_DOTS_LOCAL_COLS)
    return fail(KPM_ERR_ARG, "unknown dots_mode");
  if (dots_mode != KPM_DOTS_NONE && (!d_dots || out_cap <= 0))
    return fail(KPM_ERR_ARG, "dots output buffer missing");
  if (dots_mode == KPM_DOTS_DOUBLING && !grid_exp)
    return fail(KPM_ERR_ARG, "grid_exp buffer missing (doubling mode)");
  const bool local = dots_mode == KPM_DOTS_LOCAL || dots_mode == KPM_DOTS_LOCAL_COLS;
  if ((dots_mode == KPM_DOTS_PROJECT || local) && !w)
    return fail(KPM_ERR_ARG, "projection vector missing");
  if (first_step < 0 || n_steps < 0) return fail(KPM_ERR_ARG, "steps must be non-negative");
  if (dots_mode != KPM_DOTS_NONE &&
      (first_step - out_base < 0 || first_step + n_steps - out_base > out_cap))
    return fail(KPM_ERR_ARG, "dots slots [first_step - out_base, + n_steps) outside "
                             "[0, out_cap)");
  KernelShape ks[kMaxVariants];
  const int nv = variants_for(width, h->value_type, ks, kMaxVariants);
  if (nv == 0) return fail(KPM_ERR_UNSUPPORTED, "block width must be a power of two <= 128 "
                                                "(<= 64 for complex128)");
  if (variant < 0) variant = 0;
  if (variant >= nv) return fail(KPM_ERR_ARG, "variant out of range");
  KernelShape shape = ks[variant];

  // the workgroup split of kpm_plan_create for one segment
  int dev = 0, cus = 0, occ = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (!rc)
    rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                   "hipDeviceGetAttribute");
  if (!rc) rc = blocks_per_cu(h->value_type, shape, &occ);
  if (rc) return rc;
  occ -= shape.spare;
  if (occ < 1) occ = 1;
  const int64_t n_rows = h->n_rows;
  const int64_t rows_per_iter = int64_t(shape.waves) * 64 / shape.lpr;
  const int64_t target = int64_t(cus) * occ * shape.rounds;
  int64_t rpb = (n_rows + target - 1) / target;
  const int64_t align = rows_per_iter * 4;
  rpb = (rpb + align - 1) / align * align;
  if (rpb < align) rpb = align;
  shape.iters = int(rpb / rows_per_iter);
  const int32_t nblocks = int32_t((n_rows + rpb - 1) / rpb);
  const int32_t nblocks_launch = (nblocks + kNumXcd - 1) / kNumXcd * kNumXcd;
  StaticLayout L{n_rows, rpb, shape.chunk > 0 ? rows_per_iter * shape.chunk : 0, nblocks,
                 nblocks_launch / kNumXcd};
  if (L.C > 0) {
    int64_t max_chunks = 1;
    for (int i = 0; i < nblocks; ++i) {
      int64_t first, end, stride;
      static_block(L, i, &first, &end, &stride);
      const int64_t chunks = (end - first + stride - 1) / stride;
      if (chunks > max_chunks) max_chunks = chunks;
    }
    shape.iters = int(max_chunks * shape.chunk);
  }

  // the workspace inside the static array
  const int64_t part_half = int64_t(nblocks) * 4 * width * 2;
  size_t off = 0;
  const size_t o_part = off;
  off = align256(off + sizeof(double) * size_t(2 * part_half));
  const size_t o_rows = off;
  off = align256(off + sizeof(int2) * size_t(nblocks));
  const size_t o_stride = off;
  off = align256(off + sizeof(int64_t) * size_t(nblocks));
  const size_t o_seg = off;
  off = align256(off + sizeof(int32_t) * size_t(nblocks));
  const size_t o_segblk = off;
  off = align256(off + sizeof(int32_t) * 2);
  const size_t o_scale = off;
  off = align256(off + sizeof(double) * 2);
  const size_t o_grid = off;
  off = align256(off + sizeof(double) * 4 * size_t(width));
  if (int64_t(off) > kStaticBytes)
    return fail(KPM_ERR_UNSUPPORTED,
                "workspace of " + std::to_string(off) + " bytes exceeds the " +
                    std::to_string(kStaticBytes) +
                    " bytes of the static array (fewer rows, a narrower block or a shape "
                    "with fewer residency rounds)");
  char* base = nullptr;
  rc = scratch_of_current_device(&base);
  if (rc) return rc;

  hipStream_t s = (hipStream_t)stream;
  StepParams p{};
  p.indptr = h->indptr;
  p.indices = h->indices;
  p.values = h->values;
  p.ring = reinterpret_cast<char*>(ring);
  p.slot_bytes = slot_stride * int64_t(h->value_type == KPM_C128 ? 16 : 8);
  p.slots = slots;
  p.w = w;
  p.blk_rows = reinterpret_cast<int2*>(base + o_rows);
  p.blk_stride = reinterpret_cast<int64_t*>(base + o_stride);
  p.chunk_iters = shape.chunk;
  p.blk_seg = reinterpret_cast<int32_t*>(base + o_seg);
  p.seg_blk = reinterpret_cast<int32_t*>(base + o_segblk);
  p.scale = reinterpret_cast<double*>(base + o_scale);
  p.nblocks = nblocks;
  p.nblocks_launch = nblocks_launch;
  p.nseg = 1;
  p.xcd = kNumXcd;
  double* part = reinterpret_cast<double*>(base + o_part);
  hipLaunchKernelGGL(static_layout_kernel, dim3((nblocks + 255) / 256), dim3(256), 0, s, L,
                     const_cast<int2*>(p.blk_rows), const_cast<int64_t*>(p.blk_stride),
                     const_cast<int32_t*>(p.blk_seg), const_cast<int32_t*>(p.seg_blk),
                     const_cast<double*>(p.scale), a, b);
  rc = check_hip(hipGetLastError(), "static_layout_kernel launch");
  if (rc) return rc;

  if (dots_mode == KPM_DOTS_DOUBLING && first_step == 0) {
    // the grid of the exact sums from <a_0|a_0> (ring slot 0), on the device:
    // Re <w|x> of the projection mode with w = x = a_0 (record slot 0)
    StepParams q = p;
    q.ring = reinterpret_cast<char*>(ring);
    q.slot_bytes = 0;
    q.slots = 1;
    q.w = ring;
    q.part = part;
    q.im_exp = nullptr;
    q.dots = reinterpret_cast<double*>(base + o_grid);
    rc = launch_cheb(h->value_type, shape, KPM_DOTS_PROJECT, true, 0, 0, q, s);
    if (!rc) rc = launch_reduce(KPM_DOTS_PROJECT, width, q, s);
    if (rc) return rc;
    hipLaunchKernelGGL(static_grid_kernel, dim3((width + 63) / 64), dim3(64), 0, s, q.dots,
                       int(width), grid_exp);
    rc = check_hip(hipGetLastError(), "static_grid_kernel launch");
    if (rc) return rc;
  }
  p.im_exp = dots_mode == KPM_DOTS_DOUBLING ? grid_exp : nullptr;

  // the step loop of kpm_cheb_steps, with the step counter from the caller
  const int64_t rec = int64_t(width) * 4;
  const int64_t local_rec =
      n_rows * 2 * (dots_mode == KPM_DOTS_LOCAL_COLS ? int64_t(width) : 1);
  p.local_cols = dots_mode == KPM_DOTS_LOCAL_COLS;
  for (int i = 0; i < n_steps; ++i) {
    const int n = first_step + i;
    p.part = part + (n & 1) * part_half;
    p.prev_part = nullptr;
    p.prev_dots = nullptr;
    p.local = local ? d_dots + int64_t(n - out_base) * local_rec : nullptr;
    if (dots_mode != KPM_DOTS_NONE && !local && i > 0) {
      p.prev_part = part + ((n - 1) & 1) * part_half;
      p.prev_dots = d_dots + int64_t(n - 1 - out_base) * rec;
    }
    rc = launch_cheb(h->value_type, shape, dots_mode, false, n, h->row_nnz_hint, p, s);
    if (rc) return rc;
  }
  if (dots_mode != KPM_DOTS_NONE && !local && n_steps > 0) {
    const int last = first_step + n_steps - 1;
    p.part = part + (last & 1) * part_half;
    p.dots = d_dots + int64_t(last - out_base) * rec;
    rc = launch_reduce(dots_mode, width, p, s);
  }
  return rc;
}
