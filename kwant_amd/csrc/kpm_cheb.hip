// Chebyshev recursion kernels of the KPM moment engine (gfx950).
//
// One launch = one Chebyshev step over an N x W block of random vectors
// (row-major: a matrix row of the block is W contiguous complex numbers):
//
//   a_{n+1}[i, :] = 2 * ( sum_k H[i, j_k] a_n[j_k, :]  -  b a_n[i, :] ) / a
//                   - a_{n-1}[i, :]                                  (n >= 1)
//   a_1[i, :]     = ( sum_k H[i, j_k] a_0[j_k, :]  -  b a_0[i, :] ) / a
//
// which is kwant/kpm.py:466 (and :442, :481) with the rescaled matvec of
// kwant/kpm.py:1182-1183 fused into the SpMM epilogue: no temporaries, one
// read of a_n (gathers hit L2), one streaming read of a_{n-1}, one streaming
// write of a_{n+1} into the same slot (in place).  The two dot products of
// the doubling identity (kwant/kpm.py:470-473) are accumulated in FP64 in
// the same pass: one lane owns VEC columns of the block, so per-column sums
// need no shuffles until the end of the workgroup; workgroup partials are
// then reduced by the last workgroup of a 32-block group and the group
// partials by the last group of the segment (atomic tickets, fixed summation
// order => bitwise reproducible).  Every addition is a TwoSum and the
// (sum, error) pair is carried through every level and rounded once at the
// end, so a column's result is the correctly rounded sum of its terms (up to
// the error of the error term) whatever the launch shape or block width.
//
// Wave layout: LPR lanes per matrix row, VEC complex values per lane
// (W = LPR * VEC), 64 / LPR rows per wavefront.  Loads are 8 or 16 bytes per
// lane, so a row of the block is read as one coalesced 512-1024 byte access
// when W*sizeof(complex) allows.

#include "kpm_internal.h"

namespace kpm {

typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

template <typename T, int VEC>
struct CLane {
  T re[VEC];
  T im[VEC];
};

template <typename T, int VEC>
__device__ inline void zero(CLane<T, VEC>& c) {
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    c.re[v] = T(0);
    c.im[v] = T(0);
  }
}

template <typename T, int VEC>
struct LaneIO;

template <>
struct LaneIO<float, 1> {
  __device__ static CLane<float, 1> load(const float* p) {
    f2v v = *reinterpret_cast<const f2v*>(p);
    return {{v.x}, {v.y}};
  }
  __device__ static CLane<float, 1> load_nt(const float* p) {
    f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
    return {{v.x}, {v.y}};
  }
  __device__ static void store_nt(float* p, const CLane<float, 1>& c) {
    f2v v = {c.re[0], c.im[0]};
    __builtin_nontemporal_store(v, reinterpret_cast<f2v*>(p));
  }
  __device__ static void store(float* p, const CLane<float, 1>& c) {
    f2v v = {c.re[0], c.im[0]};
    *reinterpret_cast<f2v*>(p) = v;
  }
};

template <>
struct LaneIO<float, 2> {
  __device__ static CLane<float, 2> load(const float* p) {
    f4v v = *reinterpret_cast<const f4v*>(p);
    return {{v.x, v.z}, {v.y, v.w}};
  }
  __device__ static CLane<float, 2> load_nt(const float* p) {
    f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
    return {{v.x, v.z}, {v.y, v.w}};
  }
  __device__ static void store_nt(float* p, const CLane<float, 2>& c) {
    f4v v = {c.re[0], c.im[0], c.re[1], c.im[1]};
    __builtin_nontemporal_store(v, reinterpret_cast<f4v*>(p));
  }
  __device__ static void store(float* p, const CLane<float, 2>& c) {
    f4v v = {c.re[0], c.im[0], c.re[1], c.im[1]};
    *reinterpret_cast<f4v*>(p) = v;
  }
};

template <>
struct LaneIO<double, 1> {
  __device__ static CLane<double, 1> load(const double* p) {
    d2v v = *reinterpret_cast<const d2v*>(p);
    return {{v.x}, {v.y}};
  }
  __device__ static CLane<double, 1> load_nt(const double* p) {
    d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
    return {{v.x}, {v.y}};
  }
  __device__ static void store_nt(double* p, const CLane<double, 1>& c) {
    d2v v = {c.re[0], c.im[0]};
    __builtin_nontemporal_store(v, reinterpret_cast<d2v*>(p));
  }
  __device__ static void store(double* p, const CLane<double, 1>& c) {
    d2v v = {c.re[0], c.im[0]};
    *reinterpret_cast<d2v*>(p) = v;
  }
};

// Complex matrix value k.
template <typename T>
__device__ inline void load_value(const void* values, int64_t k, T& re, T& im);
template <>
__device__ inline void load_value<float>(const void* values, int64_t k, float& re,
                                         float& im) {
  f2v v = reinterpret_cast<const f2v*>(values)[k];
  re = v.x;
  im = v.y;
}
template <>
__device__ inline void load_value<double>(const void* values, int64_t k, double& re,
                                          double& im) {
  d2v v = reinterpret_cast<const d2v*>(values)[k];
  re = v.x;
  im = v.y;
}

// TwoSum accumulation: (s, c) += t, with c collecting the exact rounding
// error of every addition (Knuth).
__device__ inline void acc2(double& s, double& c, double t) {
  const double s1 = s + t;
  const double bp = s1 - s;
  const double e = (s - (s1 - bp)) + (t - bp);
  s = s1;
  c += e;
}

__device__ inline unsigned ticket(unsigned* ctr) {
  return __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_rel(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_coh(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double load_coh(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Number of FP64 accumulators per column for a dots mode.
template <int MODE>
struct NComp {
  static constexpr int value = MODE == KPM_DOTS_PROJECT ? 4 : (MODE == KPM_DOTS_DOUBLING ? 3 : 1);
};

// Position of accumulator c in the 4-double output record of a column.
template <int MODE>
__device__ inline int out_pos(int c) {
  if (MODE == KPM_DOTS_DOUBLING) return c == 0 ? 0 : c + 1;
  return c;
}

template <typename T, int VEC, int LPR, int WAVES, int MODE>
__global__ __launch_bounds__(WAVES * 64) void cheb_kernel(StepParams p, int iters,
                                                          int only_dots) {
  constexpr int W = LPR * VEC;
  constexpr int RPW = 64 / LPR;
  constexpr int NC = NComp<MODE>::value;
  using IO = LaneIO<T, VEC>;

  const int hb = blockIdx.x;
  const int per = p.nblocks_launch / p.xcd;
  const int lb = (hb % p.xcd) * per + hb / p.xcd;
  if (lb >= p.nblocks) return;

  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int grp = lane / LPR;
  const int col = (lane % LPR) * VEC;

  const int n = __builtin_amdgcn_readfirstlane(
      __hip_atomic_load(p.ctrl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  const int seg = p.blk_seg[lb];
  const int2 rr = p.blk_rows[lb];
  const double sa = p.scale[2 * seg];
  const double sb = p.scale[2 * seg + 1];
  const T mul = T((n == 0 ? 1.0 : 2.0) / sa);
  const T bsh = T(sb);

  const T* cur = reinterpret_cast<const T*>(p.ring + int64_t(n % p.slots) * p.slot_bytes);
  T* nxt = reinterpret_cast<T*>(p.ring + int64_t((n + 1) % p.slots) * p.slot_bytes);
  const T* prv = reinterpret_cast<const T*>(
      p.ring + int64_t((n + p.slots - 1) % p.slots) * p.slot_bytes);
  const T* wv = reinterpret_cast<const T*>(p.w);

  double as[NC][VEC], ac[NC][VEC];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      as[c][v] = 0.0;
      ac[c][v] = 0.0;
    }

  for (int it = 0; it < iters; ++it) {
    const int row = rr.x + (it * WAVES + wave) * RPW + grp;
    if (row >= rr.y) break;
    const int64_t roff = 2 * (int64_t(row) * W + col);
    CLane<T, VEC> xs = IO::load(cur + roff);
    CLane<T, VEC> y;
    zero(y);
    if (!only_dots) {
      CLane<T, VEC> xp;
      if (n > 0)
        xp = IO::load_nt(prv + roff);
      else
        zero(xp);
      const int k0 = p.indptr[row];
      const int k1 = p.indptr[row + 1];
      CLane<T, VEC> acc;
      zero(acc);
      for (int k = k0; k < k1; k += 4) {
        T hr[4], hi[4];
        CLane<T, VEC> x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (k + u < k1) {
            const int j = p.indices[k + u];
            load_value<T>(p.values, k + u, hr[u], hi[u]);
            x[u] = IO::load(cur + 2 * (int64_t(j) * W + col));
          } else {
            hr[u] = T(0);
            hi[u] = T(0);
            zero(x[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            acc.re[v] = fma(hr[u], x[u].re[v], acc.re[v]);
            acc.re[v] = fma(-hi[u], x[u].im[v], acc.re[v]);
            acc.im[v] = fma(hr[u], x[u].im[v], acc.im[v]);
            acc.im[v] = fma(hi[u], x[u].re[v], acc.im[v]);
          }
      }
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        y.re[v] = fma(acc.re[v] - bsh * xs.re[v], mul, -xp.re[v]);
        y.im[v] = fma(acc.im[v] - bsh * xs.im[v], mul, -xp.im[v]);
      }
      IO::store_nt(nxt + roff, y);
    }
    if (MODE == KPM_DOTS_DOUBLING) {
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const double xr = xs.re[v], xi = xs.im[v];
        const double yr = y.re[v], yi = y.im[v];
        acc2(as[0][v], ac[0][v], fma(xr, xr, xi * xi));
        acc2(as[1][v], ac[1][v], fma(yr, xr, yi * xi));
        acc2(as[2][v], ac[2][v], fma(yr, xi, -(yi * xr)));
      }
    } else if (MODE == KPM_DOTS_PROJECT) {
      CLane<T, VEC> ww = IO::load(wv + roff);
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const double wr = ww.re[v], wi = ww.im[v];
        const double xr = xs.re[v], xi = xs.im[v];
        const double yr = y.re[v], yi = y.im[v];
        acc2(as[0][v], ac[0][v], fma(wr, xr, wi * xi));
        acc2(as[1][v], ac[1][v], fma(wr, xi, -(wi * xr)));
        acc2(as[2][v], ac[2][v], fma(wr, yr, wi * yi));
        acc2(as[3][v], ac[3][v], fma(wr, yi, -(wi * yr)));
      }
    }
  }

  if (MODE == KPM_DOTS_NONE) {
    // Only the step counter: the last workgroup to finish advances it.
    __syncthreads();
    if (tid == 0) {
      const unsigned t = ticket(p.done_cnt);
      if (t == unsigned(p.nblocks - 1)) {
        store_rel(p.done_cnt, 0u);
        __hip_atomic_store(p.ctrl, n + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    return;
  }

  // ---- workgroup reduction: row groups of a wave, then waves via LDS ----
#pragma unroll
  for (int off = LPR; off < 64; off <<= 1) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const double os = __shfl_xor(as[c][v], off);
        const double oc = __shfl_xor(ac[c][v], off);
        acc2(as[c][v], ac[c][v], os);
        ac[c][v] += oc;
      }
  }
  __shared__ double red_s[WAVES][NC][W];
  __shared__ double red_c[WAVES][NC][W];
  __shared__ int s_flag;
  if (grp == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        red_s[wave][c][col + v] = as[c][v];
        red_c[wave][c][col + v] = ac[c][v];
      }
  }
  __syncthreads();
  // Partials are (sum, error) pairs: part[((blk * NC + c) * W + col) * 2 + {0,1}].
  for (int t = tid; t < NC * W; t += WAVES * 64) {
    const int c = t / W, cc = t % W;
    double s = 0.0, e = 0.0;
#pragma unroll
    for (int wv2 = 0; wv2 < WAVES; ++wv2) {
      acc2(s, e, red_s[wv2][c][cc]);
      e += red_c[wv2][c][cc];
    }
    double* dst = p.part + ((int64_t(lb) * NC + c) * W + cc) * 2;
    store_coh(dst, s);
    store_coh(dst + 1, e);
  }
  __threadfence();
  __syncthreads();

  // ---- group level: the last workgroup of the group sums its partials ----
  const int gi = p.blk_grp[lb];
  if (tid == 0) {
    const unsigned t = ticket(p.grp_cnt + gi);
    const int gsize = p.grp_blk[gi + 1] - p.grp_blk[gi];
    s_flag = (t == unsigned(gsize - 1));
    if (s_flag) store_rel(p.grp_cnt + gi, 0u);
  }
  __syncthreads();
  if (!s_flag) return;
  __threadfence();
  {
    const int b0 = p.grp_blk[gi], b1 = p.grp_blk[gi + 1];
    for (int t = tid; t < NC * W; t += WAVES * 64) {
      const int c = t / W, cc = t % W;
      double s = 0.0, e = 0.0;
      for (int b = b0; b < b1; ++b) {
        const double* src = p.part + ((int64_t(b) * NC + c) * W + cc) * 2;
        acc2(s, e, load_coh(src));
        e += load_coh(src + 1);
      }
      double* dst = p.gpart + ((int64_t(gi) * NC + c) * W + cc) * 2;
      store_coh(dst, s);
      store_coh(dst + 1, e);
    }
  }
  __threadfence();
  __syncthreads();

  // ---- segment level: the last group of the segment writes the moments ----
  if (tid == 0) {
    const unsigned t = ticket(p.seg_cnt + seg);
    const int ngr = p.seg_grp[seg + 1] - p.seg_grp[seg];
    s_flag = (t == unsigned(ngr - 1));
    if (s_flag) store_rel(p.seg_cnt + seg, 0u);
  }
  __syncthreads();
  if (!s_flag) return;
  __threadfence();
  const int slot = only_dots ? 0 : n - p.out_base;
  const bool ok = slot >= 0 && slot < p.out_cap;
  {
    const int g0 = p.seg_grp[seg], g1 = p.seg_grp[seg + 1];
    for (int t = tid; t < NC * W; t += WAVES * 64) {
      const int c = t / W, cc = t % W;
      double s = 0.0, e = 0.0;
      for (int g = g0; g < g1; ++g) {
        const double* src = p.gpart + ((int64_t(g) * NC + c) * W + cc) * 2;
        acc2(s, e, load_coh(src));
        e += load_coh(src + 1);
      }
      if (ok) {
        double* rec = p.dots + ((int64_t(slot) * p.nseg + seg) * W + cc) * 4;
        rec[out_pos<MODE>(c)] = s + e;
        if (MODE == KPM_DOTS_DOUBLING && c == 0) rec[1] = 0.0;
      }
    }
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) {
    if (!ok) __hip_atomic_fetch_or(p.ctrl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned t = ticket(p.done_cnt);
    if (t == unsigned(p.nseg - 1)) {
      store_rel(p.done_cnt, 0u);
      if (!only_dots)
        __hip_atomic_store(p.ctrl, n + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// y[slot] = op x[slot] for an n_rows x W block (no recursion, no dots).
template <typename T, int VEC, int LPR, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void spmm_kernel(const int32_t* indptr,
                                                          const int32_t* indices,
                                                          const void* values, const T* x,
                                                          int64_t x_stride, T* y,
                                                          int64_t y_stride, int n_rows) {
  constexpr int W = LPR * VEC;
  constexpr int RPW = 64 / LPR;
  using IO = LaneIO<T, VEC>;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int row = (blockIdx.x * WAVES + wave) * RPW + lane / LPR;
  const int col = (lane % LPR) * VEC;
  if (row >= n_rows) return;
  const T* xs = x + 2 * int64_t(blockIdx.y) * x_stride;
  T* ys = y + 2 * int64_t(blockIdx.y) * y_stride;
  const int k0 = indptr[row], k1 = indptr[row + 1];
  CLane<T, VEC> acc;
  zero(acc);
  for (int k = k0; k < k1; k += 4) {
    T hr[4], hi[4];
    CLane<T, VEC> xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k + u < k1) {
        const int j = indices[k + u];
        load_value<T>(values, k + u, hr[u], hi[u]);
        xv[u] = IO::load(xs + 2 * (int64_t(j) * W + col));
      } else {
        hr[u] = T(0);
        hi[u] = T(0);
        zero(xv[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        acc.re[v] = fma(hr[u], xv[u].re[v], acc.re[v]);
        acc.re[v] = fma(-hi[u], xv[u].im[v], acc.re[v]);
        acc.im[v] = fma(hr[u], xv[u].im[v], acc.im[v]);
        acc.im[v] = fma(hi[u], xv[u].re[v], acc.im[v]);
      }
  }
  IO::store(ys + 2 * (int64_t(row) * W + col), acc);
}

// ----------------------------------------------------------------------------
// Variant table and dispatch.

namespace {

constexpr int kWaves = 4;

template <typename T, int VEC, int LPR>
int launch_one(int mode, bool only_dots, int iters, const StepParams& p,
               hipStream_t stream) {
  dim3 grid(p.nblocks_launch), block(kWaves * 64);
  switch (mode) {
    case KPM_DOTS_NONE:
      hipLaunchKernelGGL((cheb_kernel<T, VEC, LPR, kWaves, KPM_DOTS_NONE>), grid, block, 0,
                         stream, p, iters, int(only_dots));
      break;
    case KPM_DOTS_DOUBLING:
      hipLaunchKernelGGL((cheb_kernel<T, VEC, LPR, kWaves, KPM_DOTS_DOUBLING>), grid, block,
                         0, stream, p, iters, int(only_dots));
      break;
    case KPM_DOTS_PROJECT:
      hipLaunchKernelGGL((cheb_kernel<T, VEC, LPR, kWaves, KPM_DOTS_PROJECT>), grid, block,
                         0, stream, p, iters, int(only_dots));
      break;
    default:
      return fail(KPM_ERR_ARG, "unknown dots mode");
  }
  return check_hip(hipGetLastError(), "cheb_kernel launch");
}

template <typename T, int VEC, int LPR>
int launch_spmm_one(const kpm_csr* op, const void* x, int64_t xs, void* y, int64_t ys,
                    int n_slots, hipStream_t stream) {
  constexpr int RPB = kWaves * 64 / LPR;
  dim3 grid((op->n_rows + RPB - 1) / RPB, n_slots), block(kWaves * 64);
  hipLaunchKernelGGL((spmm_kernel<T, VEC, LPR, kWaves>), grid, block, 0, stream, op->indptr,
                     op->indices, op->values, reinterpret_cast<const T*>(x), xs,
                     reinterpret_cast<T*>(y), ys, op->n_rows);
  return check_hip(hipGetLastError(), "spmm_kernel launch");
}

// (vec, lpr) dispatch for a value type; returns KPM_ERR_UNSUPPORTED if none.
template <typename T, typename F>
int dispatch_shape(int vec, int lpr, F&& f);

#define KPM_SHAPE_CASE(V, L) \
  if (vec == V && lpr == L) return f.template operator()<T, V, L>();

template <typename T, typename F>
int dispatch_shape(int vec, int lpr, F&& f) {
  KPM_SHAPE_CASE(1, 1)
  KPM_SHAPE_CASE(1, 2)
  KPM_SHAPE_CASE(1, 4)
  KPM_SHAPE_CASE(1, 8)
  KPM_SHAPE_CASE(1, 16)
  KPM_SHAPE_CASE(1, 32)
  KPM_SHAPE_CASE(1, 64)
  if constexpr (sizeof(T) == 4) {
    KPM_SHAPE_CASE(2, 1)
    KPM_SHAPE_CASE(2, 2)
    KPM_SHAPE_CASE(2, 4)
    KPM_SHAPE_CASE(2, 8)
    KPM_SHAPE_CASE(2, 16)
    KPM_SHAPE_CASE(2, 32)
    KPM_SHAPE_CASE(2, 64)
  }
  return fail(KPM_ERR_UNSUPPORTED, "no kernel for this (vec, lanes-per-row) shape");
}
#undef KPM_SHAPE_CASE

}  // namespace

int variants_for(int width, int value_type, KernelShape* out, int max_out) {
  if (width < 1 || width > 128 || (width & (width - 1)) != 0) return 0;
  int n = 0;
  auto add = [&](int lpr, int vec, int iters) {
    if (n < max_out) out[n] = KernelShape{lpr, vec, kWaves, iters};
    ++n;
  };
  // Rows per workgroup = iters * waves * 64 / lpr; two sizes per shape.
  if (value_type == KPM_C64) {
    if (width >= 2) {
      const int lpr = width / 2;
      const int rpw = 64 / lpr;
      add(lpr, 2, 1024 / (kWaves * rpw) > 0 ? 1024 / (kWaves * rpw) : 1);
      add(lpr, 2, 256 / (kWaves * rpw) > 0 ? 256 / (kWaves * rpw) : 1);
    }
    if (width <= 64) {
      const int lpr = width;
      const int rpw = 64 / lpr;
      add(lpr, 1, 1024 / (kWaves * rpw) > 0 ? 1024 / (kWaves * rpw) : 1);
      add(lpr, 1, 256 / (kWaves * rpw) > 0 ? 256 / (kWaves * rpw) : 1);
    }
  } else if (value_type == KPM_C128) {
    if (width <= 64) {
      const int lpr = width;
      const int rpw = 64 / lpr;
      add(lpr, 1, 1024 / (kWaves * rpw) > 0 ? 1024 / (kWaves * rpw) : 1);
      add(lpr, 1, 256 / (kWaves * rpw) > 0 ? 256 / (kWaves * rpw) : 1);
    }
  }
  return n;
}

namespace {

struct ChebLauncher {
  int mode;
  bool only_dots;
  int iters;
  const StepParams& p;
  hipStream_t s;
  template <typename T, int V, int LP>
  int operator()() {
    return launch_one<T, V, LP>(mode, only_dots, iters, p, s);
  }
};

struct SpmmLauncher {
  const kpm_csr* op;
  const void* x;
  int64_t xs;
  void* y;
  int64_t ys;
  int n_slots;
  hipStream_t s;
  template <typename T, int V, int LP>
  int operator()() {
    return launch_spmm_one<T, V, LP>(op, x, xs, y, ys, n_slots, s);
  }
};

}  // namespace

int launch_cheb(int value_type, const KernelShape& ks, int mode, bool only_dots,
                const StepParams& p, hipStream_t stream) {
  ChebLauncher l{mode, only_dots, ks.iters, p, stream};
  if (value_type == KPM_C64) return dispatch_shape<float>(ks.vec, ks.lpr, l);
  if (value_type == KPM_C128) return dispatch_shape<double>(ks.vec, ks.lpr, l);
  return fail(KPM_ERR_ARG, "unknown value type");
}

int launch_spmm(int value_type, int width, const kpm_csr* op, const void* x, int64_t x_stride,
                void* y, int64_t y_stride, int n_slots, hipStream_t stream) {
  if (width < 1 || width > 128 || (width & (width - 1)) != 0)
    return fail(KPM_ERR_UNSUPPORTED, "spmm width must be a power of two <= 128");
  SpmmLauncher l{op, x, x_stride, y, y_stride, n_slots, stream};
  if (value_type == KPM_C64) {
    if (width >= 2) return dispatch_shape<float>(2, width / 2, l);
    return dispatch_shape<float>(1, 1, l);
  }
  if (value_type == KPM_C128) {
    if (width > 64) return fail(KPM_ERR_UNSUPPORTED, "complex128 spmm width must be <= 64");
    return dispatch_shape<double>(1, width, l);
  }
  return fail(KPM_ERR_ARG, "unknown value type");
}

}  // namespace kpm
