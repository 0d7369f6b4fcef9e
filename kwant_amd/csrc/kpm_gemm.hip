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
                                                     int k_chunk, int accumulate, int nblocks,
                                                     int splits) {
  using T = typename GemmIn<PREC>::T;
  constexpr bool SPLIT = PREC == KPM_GEMM_FP32X3;
  constexpr bool F32 = PREC != KPM_GEMM_FP64;
  using F = typename std::conditional<F32, float, double>::type;

  int m_tile, n_tile, split;
  gemm_tile_of(blockIdx.x, nblocks, (m_rows + TM - 1) / TM, (n_rows + TN - 1) / TN, &m_tile,
               &n_tile, &split);
  if (split >= splits) return;  // grid padding
  const T* A = reinterpret_cast<const T*>(Av);
  const T* B = reinterpret_cast<const T*>(Bv);
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int m0 = m_tile * TM, n0 = n_tile * TN;
  const int64_t kb = int64_t(split) * kps;
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

  double* o = out + int64_t(split) * (accumulate ? 0 : int64_t(m_rows) * ldo * 2);
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

// ---------------------------------------------------------------------------
// Matrix-core kernel.  The complex product is two real GEMMs over the
// interleaved k axis k' = 2k + {0: re, 1: im}:
//   Re mu = sum_k' A'[m, k'] B'[n, k']     with B' = (Br, Bi)
//   Im mu = sum_k' A'[m, k'] B''[n, k']    with B'' = (Bi, -Br)
// (conj(A) B = (Ar Br + Ai Bi) + i (Ar Bi - Ai Br)).  B'' is formed while
// reading B' from LDS.  One v_mfma_*_16x16x4 computes a 16 x 16 tile over 4
// k'; a workgroup of 4 waves owns a 64 x 64 complex tile, each wave 2 x 2
// MFMA tiles, for both the real and the imaginary accumulators.
// Operand layout of the 16x16x4 MFMAs: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; result register r of lane l is D[4 (l / 16) + r][l % 16]
// for FP32 and D[l / 16 + 4 r][l % 16] for FP64.

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4m __attribute__((ext_vector_type(4)));

template <typename F>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = d4v;
  __device__ static acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  using acc_t = f4m;
  __device__ static acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int MB = 64;            // complex rows / cols per workgroup tile
constexpr int MKC = 16;           // complex k per LDS stage
constexpr int MKR = 2 * MKC;      // real k' per stage
constexpr int MLD = MKR + 1;      // padded LDS row stride

template <int PREC>
__global__ __launch_bounds__(256) void corr_gemm_mfma(const void* Av, int64_t lda,
                                                      const void* Bv, int64_t ldb,
                                                      double* out, int64_t ldo, int m_rows,
                                                      int n_rows, int64_t k, int64_t kps,
                                                      int k_chunk, int accumulate, int nblocks,
                                                      int splits) {
  using T = typename GemmIn<PREC>::T;
  constexpr bool SPLIT = PREC == KPM_GEMM_FP32X3;
  constexpr bool F32 = PREC != KPM_GEMM_FP64;
  using F = typename std::conditional<F32, float, double>::type;
  using MF = Mfma<F>;
  using acc_t = typename MF::acc_t;

  int m_tile, n_tile, split;
  gemm_tile_of(blockIdx.x, nblocks, (m_rows + MB - 1) / MB, (n_rows + MB - 1) / MB, &m_tile,
               &n_tile, &split);
  if (split >= splits) return;  // grid padding
  const T* A = reinterpret_cast<const T*>(Av);
  const T* B = reinterpret_cast<const T*>(Bv);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = m_tile * MB, n0 = n_tile * MB;
  const int64_t kb = int64_t(split) * kps;
  const int64_t ke = kb + kps < k ? kb + kps : k;

  __shared__ F as_[MB * MLD], bs_[MB * MLD];
  __shared__ F al_[SPLIT ? MB * MLD : 1], bl_[SPLIT ? MB * MLD : 1];

  acc_t cr[2][2], ci[2][2];
  double dr[2][2][4], di[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cr[i][j][r] = F(0);
        ci[i][j][r] = F(0);
        dr[i][j][r] = 0.0;
        di[i][j][r] = 0.0;
      }
    }

  const int lr = tid >> 2;        // staged row
  const int lq = (tid & 3) * 4;   // first staged complex k of this thread
  // Register prefetch: the next stage's operands are requested from global
  // memory before the current stage is multiplied out, so their latency
  // overlaps the matrix-core work.
  T pa[4][2], pb[4][2];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t kc = k0 + lq + e;
      pa[e][0] = pa[e][1] = pb[e][0] = pb[e][1] = T(0);
      if (kc < ke) {
        if (m0 + lr < m_rows) {
          const T* p = A + 2 * (int64_t(m0 + lr) * lda + kc);
          pa[e][0] = p[0];
          pa[e][1] = p[1];
        }
        if (n0 + lr < n_rows) {
          const T* p = B + 2 * (int64_t(n0 + lr) * ldb + kc);
          pb[e][0] = p[0];
          pb[e][1] = p[1];
        }
      }
    }
  };
  int64_t since_fold = 0;
  if (kb < ke) fetch(kb);
  for (int64_t k0 = kb; k0 < ke; k0 += MKC) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const T ar = pa[e][0], ai = pa[e][1], br = pb[e][0], bi = pb[e][1];
      const int o = lr * MLD + 2 * (lq + e);
      as_[o] = F(ar);
      as_[o + 1] = F(ai);
      bs_[o] = F(br);
      bs_[o + 1] = F(bi);
      if constexpr (SPLIT) {
        al_[o] = F(ar - double(F(ar)));
        al_[o + 1] = F(ai - double(F(ai)));
        bl_[o] = F(br - double(F(br)));
        bl_[o + 1] = F(bi - double(F(bi)));
      }
    }
    __syncthreads();
    if (k0 + MKC < ke) fetch(k0 + MKC);
#pragma unroll
    for (int kk = 0; kk < MKR / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      const F sgn = (kr & 1) ? F(-1) : F(1);
      F a[2], b[2], bsw[2];
      F la[2], lb[2], lbsw[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int ra = (wm * 32 + i * 16 + (lane & 15)) * MLD;
        const int rb = (wn * 32 + i * 16 + (lane & 15)) * MLD;
        a[i] = as_[ra + kr];
        b[i] = bs_[rb + kr];
        bsw[i] = sgn * bs_[rb + (kr ^ 1)];
        if constexpr (SPLIT) {
          la[i] = al_[ra + kr];
          lb[i] = bl_[rb + kr];
          lbsw[i] = sgn * bl_[rb + (kr ^ 1)];
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          cr[i][j] = MF::run(a[i], b[j], cr[i][j]);
          ci[i][j] = MF::run(a[i], bsw[j], ci[i][j]);
          if constexpr (SPLIT) {
            cr[i][j] = MF::run(a[i], lb[j], cr[i][j]);
            cr[i][j] = MF::run(la[i], b[j], cr[i][j]);
            ci[i][j] = MF::run(a[i], lbsw[j], ci[i][j]);
            ci[i][j] = MF::run(la[i], bsw[j], ci[i][j]);
          }
        }
    }
    __syncthreads();
    if constexpr (F32) {
      since_fold += MKC;
      if (since_fold >= k_chunk || k0 + MKC >= ke) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              dr[i][j][r] += double(cr[i][j][r]);
              di[i][j][r] += double(ci[i][j][r]);
              cr[i][j][r] = F(0);
              ci[i][j][r] = F(0);
            }
        since_fold = 0;
      }
    }
  }

  double* o = out + int64_t(split) * (accumulate ? 0 : int64_t(m_rows) * ldo * 2);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // result row of register r: FP32 16x16x4 holds rows 4 (l/16) + r,
        // FP64 16x16x4 holds rows (l/16) + 4 r (measured on gfx950)
        const int m = m0 + wm * 32 + i * 16 + (F32 ? 4 * (lane >> 4) + r : (lane >> 4) + 4 * r);
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m < m_rows && n < n_rows) {
          const double vr = F32 ? dr[i][j][r] : double(cr[i][j][r]);
          const double vi = F32 ? di[i][j][r] : double(ci[i][j][r]);
          double* p = o + 2 * (int64_t(m) * ldo + n);
          if (accumulate) {
            p[0] += vr;
            p[1] += vi;
          } else {
            p[0] = vr;
            p[1] = vi;
          }
        }
      }
}

template <int PREC>
int launch_fma(const void* A, int64_t lda, const void* B, int64_t ldb, double* mu, int64_t ldc,
               int m_rows, int n_rows, int64_t k, int k_chunk, double* work,
               int64_t work_elems, hipStream_t stream, bool use_mfma) {
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
  // a 1-D grid, a multiple of the XCD count: the kernels map workgroups to
  // (split, tile) in XCD-local groups (gemm_tile_of)
  const int64_t blocks = (int64_t(tiles) * slices + kNumXcd - 1) / kNumXcd * kNumXcd;
  const dim3 grid(static_cast<unsigned>(blocks));
  static_assert(TM == MB && TN == MB && KT == MKC, "tile shapes of the two kernels differ");
  auto kern = use_mfma ? corr_gemm_mfma<PREC> : corr_gemm_fma<PREC>;
  if (slices == 1) {
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, stream, A, lda, B, ldb, mu, ldc, m_rows, n_rows,
                       k, kps, k_chunk, 1, int(blocks), int(slices));
    return check_hip(hipGetLastError(), "corr_gemm launch");
  }
  hipLaunchKernelGGL(kern, grid, dim3(256), 0, stream, A, lda, B, ldb, work, int64_t(n_rows),
                     m_rows, n_rows, k, kps, k_chunk, 0, int(blocks), int(slices));
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
  const bool use_mfma = !(precision & KPM_GEMM_VECTOR_ALU);
  switch (precision & ~KPM_GEMM_VECTOR_ALU) {
    case KPM_GEMM_FP64:
      return launch_fma<KPM_GEMM_FP64>(A, lda, B, ldb, mu, ldc, m_rows, n_rows, k, k_chunk,
                                       work, work_elems, stream, use_mfma);
    case KPM_GEMM_FP32:
      return launch_fma<KPM_GEMM_FP32>(A, lda, B, ldb, mu, ldc, m_rows, n_rows, k, k_chunk,
                                       work, work_elems, stream, use_mfma);
    case KPM_GEMM_FP32X3:
      return launch_fma<KPM_GEMM_FP32X3>(A, lda, B, ldb, mu, ldc, m_rows, n_rows, k, k_chunk,
                                         work, work_elems, stream, use_mfma);
  }
  return fail(KPM_ERR_ARG, "unknown precision");
}

}  // namespace kpm
