// Device-side building blocks shared by the recursion kernels of libkpm_amd
// (the one-step kernels of kpm_cheb.hip and the two-step tiled kernel of
// kpm_tiled.hip): complex lane vectors and their loads and stores, the
// FP64 dot accumulators (TwoSum pairs and exact grid sums), the workgroup
// partials and their reduction.
#pragma once

#include <type_traits>

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

// Number of TwoSum (sum, error) accumulators per column for a dots mode.
// The doubling mode sums <a_n|a_n> and Re <a_{n+1}|a_n> this way.
template <int MODE>
struct NComp {
  static constexpr int value = MODE == KPM_DOTS_PROJECT ? 4 : (MODE == KPM_DOTS_DOUBLING ? 2 : 1);
};

// Im <a_{n+1}|a_n> of the doubling mode.  For a Hermitian H it is zero in
// exact arithmetic (it equals -Im <a_{n-1}|a_n>, down to Im <a_1|a_0> = 0),
// so what the vectors give is rounding noise: a sum that cancels.  A TwoSum
// chain rounds such a sum differently for every launch shape, so it is
// summed exactly instead.  Each term t (an FP64 value, computed the same way
// in every shape) is cut on a fixed grid into two integers,
//   t * 2^k = h + l 2^-22 (+ a tail below 2^-23, dropped),   |l| <= 2^21,
// with k = 51 - E and 2^E > <a_0|a_0> of the column (the start vectors'
// squared norm; sum |t| <= ||a_{n+1}|| ||a_n|| <= <a_0|a_0> while the
// spectrum of the rescaled H lies in [-1, 1]).  Both integer sums stay below
// 2^52 for up to 2^31 rows, so every FP64 addition of them is exact and the
// result does not depend on the order: the same bits for every launch shape,
// block width and resumed run.  The grid 2^(E-73) is 2^-73 of <a_0|a_0>.
template <int MODE>
struct NQuant {
  static constexpr int value = MODE == KPM_DOTS_DOUBLING ? 2 : 0;
};
constexpr double kQuantStep = 4194304.0;          // 2^22
constexpr double kQuantInv = 1.0 / 4194304.0;     // 2^-22

__device__ inline void quant_add(double& h, double& l, double t, int k) {
  const double scaled = ldexp(t, k);
  const double hi = rint(scaled);
  h += hi;
  l += rint((scaled - hi) * kQuantStep);          // scaled - hi is exact
}

// Position of accumulator c in the 4-double output record of a column.
template <int MODE>
__device__ inline int out_pos(int c) {
  if (MODE == KPM_DOTS_DOUBLING) return c == 0 ? 0 : 2;
  return c;
}

// sum_k H[row, j_k] x[j_k, cols] for U slots of RPW (1 or 2) consecutive rows
// of a wavefront (rows row0 .. row0 + U*RPW - 1, row0 wave-uniform; slot u,
// lane group g owns row row0 + u*RPW + g).  indptr and every entry are read
// with scalar loads (uniform addresses); each lane group selects its own row's
// entry, so a chunk of 4 entries per row costs one scalar wait and the
// gathers of all U slots are in flight together.  Surplus slots of a chunk
// reuse a valid entry with a zero weight.
template <typename T, int VEC, int RPW, int U>
__device__ inline void spmm_rows_scalar(const int32_t* __restrict__ indptr,
                                        const int32_t* __restrict__ indices,
                                        const void* __restrict__ values, const T* cur, int row0,
                                        int rows_left, int grp, int col, int W,
                                        CLane<T, VEC> (&acc)[U]) {
  using IO = LaneIO<T, VEC>;
  constexpr int Q = U * RPW;
  const int nrows = rows_left < Q ? rows_left : Q;
  int kb[Q + 1];
  kb[0] = indptr[row0];
#pragma unroll
  for (int q = 1; q <= Q; ++q) kb[q] = q <= nrows ? indptr[row0 + q] : kb[q - 1];
  int len[Q];
  int nmax = 0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    len[q] = kb[q + 1] - kb[q];
    nmax = len[q] > nmax ? len[q] : nmax;
  }
  const int safe = kb[Q] - 1;  // a valid entry whenever nmax > 0
  int my_len[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    my_len[u] = len[u * RPW];
#pragma unroll
    for (int r = 1; r < RPW; ++r)
      if (grp == r) my_len[u] = len[u * RPW + r];
    zero(acc[u]);
  }
  for (int c = 0; c < nmax; c += 4) {
    int j[U][4];
    T hr[U][4], hi[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int v4 = 0; v4 < 4; ++v4) {
        int jr[RPW];
        T hrr[RPW], hir[RPW];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
          const int q = u * RPW + r;
          const int e = c + v4 < len[q] ? kb[q] + c + v4 : safe;
          jr[r] = indices[e];
          load_value<T>(values, e, hrr[r], hir[r]);
          if (c + v4 >= len[q]) hrr[r] = hir[r] = T(0);
        }
        j[u][v4] = jr[0];
        hr[u][v4] = hrr[0];
        hi[u][v4] = hir[0];
#pragma unroll
        for (int r = 1; r < RPW; ++r)
          if (grp == r) {
            j[u][v4] = jr[r];
            hr[u][v4] = hrr[r];
            hi[u][v4] = hir[r];
          }
      }
    CLane<T, VEC> x[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int v4 = 0; v4 < 4; ++v4) {
        x[u][v4] = IO::load(cur + 2 * (int64_t(j[u][v4]) * W + col));
        if (c + v4 >= my_len[u]) zero(x[u][v4]);
      }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int v4 = 0; v4 < 4; ++v4)
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          acc[u].re[v] = fma(hr[u][v4], x[u][v4].re[v], acc[u].re[v]);
          acc[u].re[v] = fma(-hi[u][v4], x[u][v4].im[v], acc[u].re[v]);
          acc[u].im[v] = fma(hr[u][v4], x[u][v4].im[v], acc[u].im[v]);
          acc[u].im[v] = fma(hi[u][v4], x[u][v4].re[v], acc[u].im[v]);
        }
  }
}

// FP64 dot contributions of one row (lane columns) to the accumulators:
// TwoSum pairs (as, ac) and, in the doubling mode, the exact grid sums aq of
// Im <a_{n+1}|a_n> with the columns' grid exponents qk.
//
// fg (complex64 vectors, doubling mode, grid set): all three sums are exact
// sums on the single grid 2^-k = 2^(E-51): each term is rounded to the grid
// (3 FP64 operations instead of a 6-operation TwoSum or the two-level grid)
// and the integer sums are order-independent.  The rounding of each term,
// at most 2^-52 of <a_0|a_0>, is far below what the FP32 vectors carry
// (their recursion is good to ~1e-7), and the sums stay the same bits for
// every launch shape.  as holds grid units until finish_grid_dots.
template <typename T, int VEC, int MODE, int NC, int NQ>
__device__ inline void add_dots(double (&as)[NC][VEC], double (&ac)[NC][VEC],
                                double (&aq)[NQ][VEC], const int (&qk)[VEC], bool fg,
                                const CLane<T, VEC>& xs, const CLane<T, VEC>& y,
                                const CLane<T, VEC>& ww) {
  if constexpr (MODE == KPM_DOTS_DOUBLING && std::is_same<T, float>::value) {
    if (fg) {
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const double xr = xs.re[v], xi = xs.im[v];
        const double yr = y.re[v], yi = y.im[v];
        as[0][v] += rint(ldexp(fma(xr, xr, xi * xi), qk[v]));
        as[1][v] += rint(ldexp(fma(yr, xr, yi * xi), qk[v]));
        aq[0][v] += rint(ldexp(fma(yr, xi, -(yi * xr)), qk[v]));
      }
      return;
    }
  }
  if constexpr (MODE == KPM_DOTS_DOUBLING) {
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      const double xr = xs.re[v], xi = xs.im[v];
      const double yr = y.re[v], yi = y.im[v];
      acc2(as[0][v], ac[0][v], fma(xr, xr, xi * xi));
      acc2(as[1][v], ac[1][v], fma(yr, xr, yi * xi));
      const double t = fma(yr, xi, -(yi * xr));
      quant_add(aq[0][v], aq[1][v], t, qk[v]);
    }
  } else if (MODE == KPM_DOTS_PROJECT) {
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

// The single-grid sums (fg) back from grid units to values (exact scaling),
// so that the workgroup reduction adds them as (sum, 0) pairs: sums of grid
// multiples below 2^52 units are exact in every order.
template <typename T, int MODE, int NC, int VEC>
__device__ inline void finish_grid_dots(bool fg, double (&as)[NC][VEC], const int (&qk)[VEC]) {
  if constexpr (MODE == KPM_DOTS_DOUBLING && std::is_same<T, float>::value) {
    if (fg) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int v = 0; v < VEC; ++v) as[c][v] = ldexp(as[c][v], -qk[v]);
    }
  }
}

// Site-resolved operator record (KPM_DOTS_LOCAL) of one matrix row: the
// FP64 products conj(w[row, c]) x[row, c] of the lane's columns, either
// stored per column (p.local_cols) or summed over the block's columns -- a
// TwoSum over the lane's columns, then a butterfly over the row's LPR lanes
// carrying the (sum, error) pairs, rounded once -- and stored by the row's
// first lane.  Every lane of the wavefront calls this (the butterfly
// shuffles); rows past the end pass live = false and store nothing.
template <typename T, int VEC, int LPR>
__device__ inline void write_local(const StepParams& p, int row, bool live, int col,
                                   const CLane<T, VEC>& x, const CLane<T, VEC>& ww) {
  constexpr int W = LPR * VEC;
  double re[VEC], im[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    const double wr = ww.re[v], wi = ww.im[v];
    const double xr = x.re[v], xi = x.im[v];
    re[v] = fma(wr, xr, wi * xi);
    im[v] = fma(wr, xi, -(wi * xr));
  }
  d2v* out = reinterpret_cast<d2v*>(p.local);
  if (p.local_cols) {
    if (live) {
#pragma unroll
      for (int v = 0; v < VEC; ++v) out[int64_t(row) * W + col + v] = d2v{re[v], im[v]};
    }
    return;
  }
  double sr = 0.0, cr = 0.0, si = 0.0, ci = 0.0;
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    acc2(sr, cr, re[v]);
    acc2(si, ci, im[v]);
  }
#pragma unroll
  for (int off = 1; off < LPR; off <<= 1) {
    const double osr = __shfl_xor(sr, off), ocr = __shfl_xor(cr, off);
    const double osi = __shfl_xor(si, off), oci = __shfl_xor(ci, off);
    acc2(sr, cr, osr);
    cr += ocr;
    acc2(si, ci, osi);
    ci += oci;
  }
  if (live && col == 0) out[row] = d2v{sr + cr, si + ci};
}

// Entries per chunk of the per-lane path: all index and value loads of a
// chunk are issued before its gathers, so a row of up to CHV entries costs one
// index-to-gather round trip.
#ifndef KPM_VEC_CHUNK
#define KPM_VEC_CHUNK 4
#endif
constexpr int CHV = KPM_VEC_CHUNK;


// The same product for one row per lane group with per-lane (vector) loads of
// the matrix entries (several rows per wavefront).  Also loads the row's own
// a_n (xs) and a_{n-1} (xp), after the first chunk of matrix entries.
template <typename T, int VEC>
__device__ inline CLane<T, VEC> spmm_row_vector(const int32_t* __restrict__ indptr,
                                                const int32_t* __restrict__ indices,
                                                const void* __restrict__ values, const T* cur,
                                                int row, bool live, int col, int W,
                                                CLane<T, VEC>& xs, CLane<T, VEC>& xp,
                                                const T* prv, int64_t roff, bool has_prev) {
  using IO = LaneIO<T, VEC>;
  const int k0 = live ? indptr[row] : 0;
  const int k1 = live ? indptr[row + 1] : 0;
  int j0[CHV];
  T hr0[CHV], hi0[CHV];
  if (k0 < k1) {
#pragma unroll
    for (int u = 0; u < CHV; ++u) {
      const int kk = k0 + u < k1 ? k0 + u : k1 - 1;
      j0[u] = indices[kk];
      load_value<T>(values, kk, hr0[u], hi0[u]);
      if (k0 + u >= k1) hr0[u] = hi0[u] = T(0);
    }
  }
  xs = IO::load(cur + roff);
  if (has_prev)
    xp = IO::load_nt(prv + roff);
  else
    zero(xp);
  CLane<T, VEC> acc;
  zero(acc);
  for (int k = k0; k < k1; k += CHV) {
    T hr[CHV], hi[CHV];
    int j[CHV];
    CLane<T, VEC> x[CHV];
    if (k == k0) {
#pragma unroll
      for (int u = 0; u < CHV; ++u) {
        j[u] = j0[u];
        hr[u] = hr0[u];
        hi[u] = hi0[u];
      }
    } else {
#pragma unroll
      for (int u = 0; u < CHV; ++u) {
        const int kk = k + u < k1 ? k + u : k1 - 1;
        j[u] = indices[kk];
        load_value<T>(values, kk, hr[u], hi[u]);
        if (k + u >= k1) hr[u] = hi[u] = T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < CHV; ++u) {
      x[u] = IO::load(cur + 2 * (int64_t(j[u]) * W + col));
      if (k + u >= k1) zero(x[u]);
    }
#pragma unroll
    for (int u = 0; u < CHV; ++u)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        acc.re[v] = fma(hr[u], x[u].re[v], acc.re[v]);
        acc.re[v] = fma(-hi[u], x[u].im[v], acc.re[v]);
        acc.im[v] = fma(hr[u], x[u].im[v], acc.im[v]);
        acc.im[v] = fma(hi[u], x[u].re[v], acc.im[v]);
      }
  }
  return acc;
}

// First row of wavefront iteration it of a workgroup whose chunks of ipc
// iterations (rows_per_iter rows each) start at begin, begin + stride, ...
__device__ inline int64_t iter_row0(int it, int64_t begin, int64_t stride, int ipc,
                                    int rows_per_iter) {
  const int m = it / ipc;
  return begin + m * stride + int64_t(it - m * ipc) * rows_per_iter;
}


// One wavefront sums the workgroup partials of one (segment, accumulator,
// column) and writes it to the dots record: lane l takes the segment's
// blocks l, l + 64, ... in order, then a fixed butterfly combines the lanes.
// The doubling mode's exact grid sums of Im <a_{n+1}|a_n> are one more
// group: one wavefront adds up the three integer levels (exact) and scales
// them back.  Groups q = first, first + stride, ... of the
// nseg * (NC + grid) * W are done.
template <int MODE>
__device__ inline void reduce_sums(const double* __restrict__ part, int nblocks,
                                   const int32_t* __restrict__ seg_blk, int nseg, int W,
                                   const int32_t* __restrict__ qexp,
                                   double* __restrict__ rec_out, int first, int stride) {
  constexpr int NC = NComp<MODE>::value;
  constexpr int NQ = NQuant<MODE>::value;
  constexpr int NG = NC + (NQ > 0 ? 1 : 0);
  const int lane = threadIdx.x & 63;
  const int total = nseg * NG * W;
  for (int q = first; q < total; q += stride) {
    const int seg = q / (NG * W);
    const int rem = q - seg * NG * W;
    const int c = rem / W, cc = rem - (rem / W) * W;
    const int b0 = seg_blk[seg], b1 = seg_blk[seg + 1];
    double* rec = rec_out + (int64_t(seg) * W + cc) * 4;
    if (NQ > 0 && c == NC) {
      double lev[NQ > 0 ? NQ : 1];
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        const double* src = part + (int64_t(NC + j) * W + cc) * nblocks * 2;
        double s = 0.0;
        for (int b = b0 + lane; b < b1; b += 64) s += src[2 * b];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
        lev[j] = s;
      }
      if (lane == 0) {
        // t 2^k = h + l 2^-22: both terms are scaled back exactly, then
        // added with one rounding
        const int k = qexp != nullptr ? qexp[int64_t(seg) * W + cc] : 0;
        rec[3] = ldexp(lev[0], -k) + ldexp(lev[1] * kQuantInv, -k);
      }
      continue;
    }
    const double* src = part + (int64_t(c) * W + cc) * nblocks * 2;
    double s = 0.0, e = 0.0;
#pragma unroll 4
    for (int b = b0 + lane; b < b1; b += 64) {
      acc2(s, e, src[2 * b]);
      e += src[2 * b + 1];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double os = __shfl_xor(s, off);
      const double oe = __shfl_xor(e, off);
      acc2(s, e, os);
      e += oe;
    }
    if (lane == 0) {
      rec[out_pos<MODE>(c)] = s + e;
      if (MODE == KPM_DOTS_DOUBLING && c == 0) rec[1] = 0.0;  // <a_n|a_n> is real
    }
  }
}


// Workgroup reduction of the dot accumulators (row groups of a wave, then
// waves via LDS) and the store of the workgroup's partials.
//
// part, nblocks: the destination and its workgroup count; the lanes hold
// columns col_base + col of a block w_full columns wide.
// red_s, red_c: WAVES x NC x W doubles of LDS each, red_q: WAVES x NQA x W.
template <int VEC, int LPR, int WAVES, int NC, int NQ, int NQA>
__device__ inline void store_partials_to(double* part, int nblocks, int w_full, int col_base,
                                         int lb, int wave, int grp, int col, int tid,
                                         double (&as)[NC][VEC], double (&ac)[NC][VEC],
                                         double (&aq)[NQA][VEC], double* red_s, double* red_c,
                                         double* red_q) {
  constexpr int W = LPR * VEC;
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
#pragma unroll
    for (int c = 0; c < NQ; ++c)
#pragma unroll
      for (int v = 0; v < VEC; ++v) aq[c][v] += __shfl_xor(aq[c][v], off);  // exact
  }
  if (grp == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        red_s[(wave * NC + c) * W + col + v] = as[c][v];
        red_c[(wave * NC + c) * W + col + v] = ac[c][v];
      }
#pragma unroll
    for (int c = 0; c < NQ; ++c)
#pragma unroll
      for (int v = 0; v < VEC; ++v) red_q[(wave * NQA + c) * W + col + v] = aq[c][v];
  }
  __syncthreads();
  // Workgroup partials are (sum, error) pairs, stored column-major over
  // workgroups -- part[((c * W + col) * nblocks + blk) * 2 + {0, 1}] -- so that
  // the reduction of one sum reads contiguous memory.  The grid sums follow
  // the TwoSum accumulators as c = NC .. NC + NQ - 1, with a zero error.
  for (int t = tid; t < (NC + NQ) * W; t += WAVES * 64) {
    const int c = t / W, cc = t % W;
    double s = 0.0, e = 0.0;
    if (c < NC) {
#pragma unroll
      for (int wv2 = 0; wv2 < WAVES; ++wv2) {
        acc2(s, e, red_s[(wv2 * NC + c) * W + cc]);
        e += red_c[(wv2 * NC + c) * W + cc];
      }
    } else {
#pragma unroll
      for (int wv2 = 0; wv2 < WAVES; ++wv2) s += red_q[(wv2 * NQA + c - NC) * W + cc];
    }
    double* dst = part + ((int64_t(c) * w_full + col_base + cc) * nblocks + lb) * 2;
    dst[0] = s;
    dst[1] = e;
  }
}


template <int VEC, int LPR, int WAVES, int NC, int NQ, int NQA>
__device__ inline void store_partials(const StepParams& p, int lb, int wave, int grp, int col,
                                      int tid, double (&as)[NC][VEC], double (&ac)[NC][VEC],
                                      double (&aq)[NQA][VEC]) {
  constexpr int W = LPR * VEC;
  __shared__ double red_s[WAVES * NC * W];
  __shared__ double red_c[WAVES * NC * W];
  __shared__ double red_q[WAVES * NQA * W];
  store_partials_to<VEC, LPR, WAVES, NC, NQ, NQA>(p.part, p.nblocks, W, 0, lb, wave, grp, col, tid,
                                                   as, ac, aq, red_s, red_c, red_q);
}

}  // namespace kpm
