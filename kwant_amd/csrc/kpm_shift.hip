// Exact Chebyshev moments of H + mu from the moments of H.
//
// If nu_k = <T_k(t)> are the moments of a spectral measure in the rescaled
// variable t = (E - b') / a', the moments of the same measure in a second
// window, y = (E + mu - b) / a = alpha t + beta, are
//     mu_n = sum_{k <= n} C_nk nu_k,   T_n(alpha t + beta) = sum_k C_nk T_k(t).
// The coefficient vectors obey the three-term recurrence of T_n,
//     C_{n+1} = 2 alpha (t . C_n) + 2 beta C_n - C_{n-1},
// with multiplication by t in the Chebyshev basis
//     (t . c)_j = (c_{j-1} + c_{j+1}) / 2  (c_{-1} = 0),  plus c_0 / 2 at j = 1,
// which is t T_k = (T_{k+1} + T_{|k-1|}) / 2.  When |alpha| + |beta| <= 1 the
// second window lies inside the first one's domain, every C_n is the
// Chebyshev series of a polynomial bounded by 1 on [-1, 1] (|C_nk| <= 2), and
// the recurrence is stable.  With the same random vectors the result is the
// recursion's own stochastic estimate, not only its expectation: a uniform
// on-site shift leaves the eigenvectors, hence the vector weights, unchanged.
//
// One wavefront per (member, row): lane l keeps coefficients k = l R .. l R + R - 1
// of C_{n-1}, C_n in FP64 registers, exchanges the two boundary values with its
// neighbours by cross-lane moves, and reduces sum_k C_nk nu_k with a butterfly.

#include "kpm_internal.h"

namespace kpm {

namespace {

template <int R>
__global__ __launch_bounds__(256) void shift_kernel(const float* __restrict__ nu, int64_t nu_bs,
                                                    int64_t nu_rs, int64_t rows, int m,
                                                    const double* __restrict__ in_scale,
                                                    const double* __restrict__ out_scale,
                                                    double alpha, const double* __restrict__ beta,
                                                    float* __restrict__ out, int64_t out_bs,
                                                    int64_t out_rs) {
  const int lane = threadIdx.x & 63;
  const int64_t p = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (p >= rows) return;
  const int64_t z = blockIdx.y;
  const float* src = nu + z * nu_bs + p * nu_rs;
  float* dst = out + z * out_bs + p * out_rs;
  const double bp = beta[p];

  double v[R], cp[R], cc[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int k = lane * R + j;
    v[j] = 0.0;
    if (k < m) v[j] = double(src[k]) * (in_scale ? in_scale[k] : 1.0);
    cp[j] = k == 0 ? 1.0 : 0.0;                      // T_0 = 1
    cc[j] = k == 0 ? bp : (k == 1 ? alpha : 0.0);    // T_1(alpha t + beta)
  }
  auto reduce = [](double s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
  };
  auto emit = [&](int n, double s) {
    if (lane == 0 && n < m) dst[n] = float(s * (out_scale ? out_scale[n] : 1.0));
  };
  {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      s0 += cp[j] * v[j];
      s1 += cc[j] * v[j];
    }
    emit(0, reduce(s0));
    emit(1, reduce(s1));
  }
  for (int n = 1; n + 1 < m; ++n) {
    // neighbours across lane boundaries: c_{lR-1} from lane l-1, c_{lR+R} from lane l+1
    double left = __shfl_up(cc[R - 1], 1, 64);
    double right = __shfl_down(cc[0], 1, 64);
    if (lane == 0) left = 0.0;
    if (lane == 63) right = 0.0;
    double s = 0.0;
    double nx[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const double lo = j == 0 ? left : cc[j - 1];
      const double hi = j == R - 1 ? right : cc[j + 1];
      double tc = 0.5 * (lo + hi);
      if (lane * R + j == 1) tc += 0.5 * (R == 1 ? left : cc[0]);
      nx[j] = 2.0 * (alpha * tc + bp * cc[j]) - cp[j];
      s += nx[j] * v[j];
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      cp[j] = cc[j];
      cc[j] = nx[j];
    }
    emit(n + 1, reduce(s));
  }
}

template <int R>
void launch_shift(const float* nu, int64_t nu_bs, int64_t nu_rs, int batch, int64_t rows, int m,
                  const double* in_scale, const double* out_scale, double alpha,
                  const double* beta, float* out, int64_t out_bs, int64_t out_rs,
                  hipStream_t s) {
  dim3 grid(unsigned((rows + 3) / 4), batch);
  hipLaunchKernelGGL(shift_kernel<R>, grid, dim3(256), 0, s, nu, nu_bs, nu_rs, rows, m, in_scale,
                     out_scale, alpha, beta, out, out_bs, out_rs);
}

}  // namespace

}  // namespace kpm

using namespace kpm;

extern "C" int kpm_moment_shift(const float* nu, int64_t nu_bs, int64_t nu_rs, int32_t batch,
                                int64_t rows, int32_t num_moments, const double* in_scale,
                                const double* out_scale, double alpha, const double* beta,
                                float* out, int64_t out_bs, int64_t out_rs,
                                kpm_stream_t stream) {
  if (!nu || !beta || !out) return fail(KPM_ERR_ARG, "NULL buffer");
  if (batch <= 0 || batch > 65535 || rows <= 0 || num_moments < 2)
    return fail(KPM_ERR_ARG, "empty shift");
  if (num_moments > 1024) return fail(KPM_ERR_UNSUPPORTED, "at most 1024 moments");
  if (nu_rs < num_moments || out_rs < num_moments) return fail(KPM_ERR_ARG, "row stride < moments");
  if (!(std::fabs(alpha) <= 1.0)) return fail(KPM_ERR_ARG, "|alpha| must be <= 1");
  if ((rows + 3) / 4 > (int64_t(1) << 31) - 1) return fail(KPM_ERR_ARG, "too many rows");
  hipStream_t s = (hipStream_t)stream;
  const int r = (num_moments + 63) / 64;
  if (r <= 1)
    launch_shift<1>(nu, nu_bs, nu_rs, batch, rows, num_moments, in_scale, out_scale, alpha, beta,
                    out, out_bs, out_rs, s);
  else if (r <= 2)
    launch_shift<2>(nu, nu_bs, nu_rs, batch, rows, num_moments, in_scale, out_scale, alpha, beta,
                    out, out_bs, out_rs, s);
  else if (r <= 4)
    launch_shift<4>(nu, nu_bs, nu_rs, batch, rows, num_moments, in_scale, out_scale, alpha, beta,
                    out, out_bs, out_rs, s);
  else if (r <= 8)
    launch_shift<8>(nu, nu_bs, nu_rs, batch, rows, num_moments, in_scale, out_scale, alpha, beta,
                    out, out_bs, out_rs, s);
  else
    launch_shift<16>(nu, nu_bs, nu_rs, batch, rows, num_moments, in_scale, out_scale, alpha, beta,
                     out, out_bs, out_rs, s);
  return check_hip(hipGetLastError(), "moment shift launch");
}
