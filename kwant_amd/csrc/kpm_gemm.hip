// Correlator moment-matrix GEMM: mu[m, n] += sum_k conj(A[m, k]) * B[n, k]
// (the Omega^H Psi product of kwant/kpm.py:693, with the vector index folded
// into k).  Both operands are k-contiguous rows of complex numbers.
//
// The k range is split over gridDim.z slices so that a 1024 x 1024 output with
// k ~ 1e7 still fills the chip; each slice writes its FP64 partial tile to a
// workspace and a second kernel adds the slices to mu in slice order
// (deterministic, no atomics).
//
// corr_gemm_fma is the portable LDS-tiled FMA kernel (64 x 64 output tile,
// 4 x 4 complex outputs per thread).  precision:
//   FP64   : complex128 inputs and FP64 arithmetic;
//   FP32   : complex64 inputs, FP32 products and partial sums over k-chunks of
//            k_chunk, each chunk folded into FP64 accumulators;
//   FP32X3 : complex128 inputs split into hi = fp32(x), lo = fp32(x - hi);
//            hi*hi + hi*lo + lo*hi in FP32 per k-chunk, folded into FP64.

#include <type_traits>

#include "kpm_internal.h"

namespace kpm {

namespace {

constexpr int TM = 64;
constexpr int TN = 64;
constexpr int KT = 16;

template <int PREC>
struct GemmIn;
template <>
struct GemmIn<KPM_GEMM_FP64> {
  using T = double;
};
template <>
struct GemmIn<KPM_GEMM_FP32> {
  using T = float;
};
template <>
struct GemmIn<KPM_GEMM_FP32X3> {
  using T = double;
};

// conj(a) * b accumulated into (re, im)
template <typename F>
__device__ inline void cmac_conj(F& re, F& im, F ar, F ai, F br, F bi) {
  re = fma(ar, br, re);
  re = fma(ai, bi, re);
  im = fma(ar, bi, im);
  im = fma(-ai, br, im);
}

template <int PREC>
__global__ __launch_bounds__(256) void corr_gemm_fma(const void* Av, int64_t lda,
                                                     const void* Bv, int64_t ldb,
                                                     double* out, int64_t ldo, int m_rows,
                                                     int n_rows, int64_t k, int64_t kps,
                                                     int k_chunk, int accumulate) {
  using T = typename GemmIn<PREC>::T;
  constexpr bool SPLIT = PREC == KPM_GEMM_FP32X3;
  constexpr bool F32 = PREC != KPM_GEMM_FP64;
  using F = typename std::conditional<F32, float, double>::type;

  const T* A = reinterpret_cast<const T*>(Av);
  const T* B = reinterpret_cast<const T*>(Bv);
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int64_t kb = int64_t(blockIdx.z) * kps;
  const int64_t ke = kb + kps < k ? kb + kps : k;

  __shared__ F a_re[KT][TM], a_im[KT][TM], b_re[KT][TN], b_im[KT][TN];
  __shared__ F al_re[SPLIT ? KT : 1][SPLIT ? TM : 1], al_im[SPLIT ? KT : 1][SPLIT ? TM : 1];
  __shared__ F bl_re[SPLIT ? KT : 1][SPLIT ? TN : 1], bl_im[SPLIT ? KT : 1][SPLIT ? TN : 1];

  double acc_re[4][4], acc_im[4][4];
  F part_re[4][4], part_im[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc_re[i][j] = acc_im[i][j] = 0.0;
      part_re[i][j] = part_im[i][j] = F(0);
    }

  int64_t since_fold = 0;
  for (int64_t k0 = kb; k0 < ke; k0 += KT) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      const int r = e / KT, c = e % KT;
      const int64_t kk = k0 + c;
      T ar = T(0), ai = T(0), br = T(0), bi = T(0);
      if (kk < ke) {
        if (m0 + r < m_rows) {
          ar = A[2 * (int64_t(m0 + r) * lda + kk)];
          ai = A[2 * (int64_t(m0 + r) * lda + kk) + 1];
        }
        if (n0 + r < n_rows) {
          br = B[2 * (int64_t(n0 + r) * ldb + kk)];
          bi = B[2 * (int64_t(n0 + r) * ldb + kk) + 1];
        }
      }
      a_re[c][r] = F(ar);
      a_im[c][r] = F(ai);
      b_re[c][r] = F(br);
      b_im[c][r] = F(bi);
      if constexpr (SPLIT) {
        al_re[c][r] = F(ar - double(F(ar)));
        al_im[c][r] = F(ai - double(F(ai)));
        bl_re[c][r] = F(br - double(F(br)));
        bl_im[c][r] = F(bi - double(F(bi)));
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < KT; ++kk) {
      F xr[4], xi[4], yr[4], yi[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        xr[i] = a_re[kk][ty + 16 * i];
        xi[i] = a_im[kk][ty + 16 * i];
        yr[i] = b_re[kk][tx + 16 * i];
        yi[i] = b_im[kk][tx + 16 * i];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if constexpr (F32)
            cmac_conj(part_re[i][j], part_im[i][j], xr[i], xi[i], yr[j], yi[j]);
          else
            cmac_conj(acc_re[i][j], acc_im[i][j], xr[i], xi[i], yr[j], yi[j]);
        }
      if constexpr (SPLIT) {
        F lr[4], li[4], mr[4], mi[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          lr[i] = al_re[kk][ty + 16 * i];
          li[i] = al_im[kk][ty + 16 * i];
          mr[i] = bl_re[kk][tx + 16 * i];
          mi[i] = bl_im[kk][tx + 16 * i];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            cmac_conj(part_re[i][j], part_im[i][j], xr[i], xi[i], mr[j], mi[j]);
            cmac_conj(part_re[i][j], part_im[i][j], lr[i], li[i], yr[j], yi[j]);
          }
      }
    }
    __syncthreads();
    if constexpr (F32) {
      since_fold += KT;
      if (since_fold >= k_chunk || k0 + KT >= ke) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            acc_re[i][j] += double(part_re[i][j]);
            acc_im[i][j] += double(part_im[i][j]);
            part_re[i][j] = part_im[i][j] = F(0);
          }
        since_fold = 0;
      }
    }
  }

  double* o = out + int64_t(blockIdx.z) * (accumulate ? 0 : int64_t(m_rows) * ldo * 2);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + ty + 16 * i, n = n0 + tx + 16 * j;
      if (m < m_rows && n < n_rows) {
        double* p = o + 2 * (int64_t(m) * ldo + n);
        if (accumulate) {
          p[0] += acc_re[i][j];
          p[1] += acc_im[i][j];
        } else {
          p[0] = acc_re[i][j];
          p[1] = acc_im[i][j];
        }
      }
    }
}

// mu[m, n] += sum_s work[s][m][n], slices in order.
__global__ void add_slices(const double* work, int slices, double* mu, int64_t ldc, int m_rows,
                           int n_rows) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t total = int64_t(m_rows) * n_rows;
  if (idx >= total) return;
  const int m = int(idx / n_rows), n = int(idx % n_rows);
  double re = 0.0, im = 0.0;
  for (int s = 0; s < slices; ++s) {
    const double* p = work + 2 * (int64_t(s) * m_rows * n_rows + int64_t(m) * n_rows + n);
    re += p[0];
    im += p[1];
  }
  mu[2 * (int64_t(m) * ldc + n)] += re;
  mu[2 * (int64_t(m) * ldc + n) + 1] += im;
}

template <int PREC>
int launch_fma(const void* A, int64_t lda, const void* B, int64_t ldb, double* mu, int64_t ldc,
               int m_rows, int n_rows, int64_t k, int k_chunk, double* work,
               int64_t work_elems, hipStream_t stream) {
  const int tiles = ((m_rows + TM - 1) / TM) * ((n_rows + TN - 1) / TN);
  int64_t slices = 1;
  const int64_t per_slice = 2 * int64_t(m_rows) * n_rows;
  if (work && work_elems >= per_slice) {
    const int64_t want = (2048 + tiles - 1) / tiles;
    const int64_t by_k = (k + 4095) / 4096;
    slices = want < by_k ? want : by_k;
    if (slices > work_elems / per_slice) slices = work_elems / per_slice;
    if (slices > 65535) slices = 65535;
    if (slices < 1) slices = 1;
  }
  int64_t kps = (k + slices - 1) / slices;
  kps = (kps + KT - 1) / KT * KT;
  slices = (k + kps - 1) / kps;
  if (k_chunk <= 0) k_chunk = 256;
  dim3 grid((n_rows + TN - 1) / TN, (m_rows + TM - 1) / TM, unsigned(slices));
  if (slices == 1) {
    hipLaunchKernelGGL(corr_gemm_fma<PREC>, grid, dim3(256), 0, stream, A, lda, B, ldb, mu, ldc,
                       m_rows, n_rows, k, kps, k_chunk, 1);
    return check_hip(hipGetLastError(), "corr_gemm launch");
  }
  hipLaunchKernelGGL(corr_gemm_fma<PREC>, grid, dim3(256), 0, stream, A, lda, B, ldb, work,
                     int64_t(n_rows), m_rows, n_rows, k, kps, k_chunk, 0);
  int rc = check_hip(hipGetLastError(), "corr_gemm launch");
  if (rc) return rc;
  const int64_t total = int64_t(m_rows) * n_rows;
  hipLaunchKernelGGL(add_slices, dim3(unsigned((total + 255) / 256)), dim3(256), 0, stream, work,
                     int(slices), mu, ldc, m_rows, n_rows);
  return check_hip(hipGetLastError(), "add_slices launch");
}

}  // namespace

int launch_corr_gemm(const void* A, int64_t lda, const void* B, int64_t ldb, double* mu,
                     int64_t ldc, int m_rows, int n_rows, int64_t k, int precision, int k_chunk,
                     double* work, int64_t work_elems, hipStream_t stream) {
  switch (precision) {
    case KPM_GEMM_FP64:
      return launch_fma<KPM_GEMM_FP64>(A, lda, B, ldb, mu, ldc, m_rows, n_rows, k, k_chunk,
                                       work, work_elems, stream);
    case KPM_GEMM_FP32:
      return launch_fma<KPM_GEMM_FP32>(A, lda, B, ldb, mu, ldc, m_rows, n_rows, k, k_chunk,
                                       work, work_elems, stream);
    case KPM_GEMM_FP32X3:
      return launch_fma<KPM_GEMM_FP32X3>(A, lda, B, ldb, mu, ldc, m_rows, n_rows, k, k_chunk,
                                         work, work_elems, stream);
  }
  return fail(KPM_ERR_ARG, "unknown precision");
}

}  // namespace kpm
