// Kernels of the moment surrogate: a small MLP trained and evaluated on the
// GPU with BF16 matrix cores and FP32 accumulation (no reference counterpart:
// the surrogate maps Hamiltonian parameters to Jackson-damped KPM moments).
//
// All tensors live in FP32; GEMM operands are rounded to BF16 (round to
// nearest even) while they are staged in LDS, products accumulate in FP32 on
// v_mfma_f32_16x16x32_bf16.  The GEMM is strided and batched,
//     C[z][m, n] = sum_k A[z*a_bs + m*a_rs + k*a_cs] * B[z*b_bs + n*b_rs + k*b_cs],
// so the forward product (X W^T), the weight gradient (dZ^T X) and the input
// gradient (dZ W) of every member of an ensemble are one launch each.
// With KPM_MLP_BF16X3 each operand is split into hi = bf16(x) and
// lo = bf16(x - hi) and the product is hi*hi + hi*lo + lo*hi (about 16
// significant bits instead of 8; used where inputs need more resolution).
//
// Operand layout of v_mfma_f32_16x16x32_bf16: lane l supplies
// A[l % 16][8 (l / 16) .. 8 (l / 16) + 7] and B[8 (l / 16) .. + 7][l % 16];
// result register r of lane l is D[4 (l / 16) + r][l % 16].

#include "kpm_internal.h"

namespace kpm {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4a __attribute__((ext_vector_type(4)));

constexpr int GB = 64;       // rows / cols per workgroup tile
#ifndef KPM_MLP_GK
#define KPM_MLP_GK 32
#endif
constexpr int GK = KPM_MLP_GK;  // k per stage (a multiple of 32, one MFMA k-step each);
                                // deeper stages mean fewer load-to-use waits
constexpr int GLD = GK + 8;  // LDS row stride (bf16), keeps 16-byte alignment

__device__ inline float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ inline float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) +
         x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// Stage a GB x GK tile of a strided operand into LDS as bf16 (hi and, for
// SPLIT, lo parts).  Consecutive threads walk k when k has unit stride and
// rows otherwise, so global loads coalesce either way.  All loads of the tile
// are issued before the first conversion, so their latencies overlap.
constexpr int kStageLoads = GB * GK / 256;

__device__ inline void stage_load(const float* __restrict__ X, int64_t rs, int64_t cs, int r0,
                                  int rows, int k0, int K, int tid, float (&v)[kStageLoads]) {
  const bool kfast = cs == 1;
#pragma unroll
  for (int q = 0; q < kStageLoads; ++q) {
    const int e = tid + 256 * q;
    const int r = kfast ? e / GK : e % GB;
    const int c = kfast ? e % GK : e / GB;
    v[q] = (r0 + r < rows && k0 + c < K) ? X[int64_t(r0 + r) * rs + int64_t(k0 + c) * cs]
                                         : 0.0f;
  }
}

template <bool SPLIT>
__device__ inline void stage_store(const float (&v)[kStageLoads], bool kfast, __bf16* hi,
                                   __bf16* lo, int tid) {
#pragma unroll
  for (int q = 0; q < kStageLoads; ++q) {
    const int e = tid + 256 * q;
    const int r = kfast ? e / GK : e % GB;
    const int c = kfast ? e % GK : e / GB;
    const __bf16 h = (__bf16)v[q];
    hi[r * GLD + c] = h;
    if (SPLIT) lo[r * GLD + c] = (__bf16)(v[q] - (float)h);
  }
}

template <int EPI, bool SPLIT>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(kpm_mlp_gemm_desc d, int accumulate) {
  __shared__ __bf16 ah[GB * GLD], bh[GB * GLD];
  __shared__ __bf16 al[SPLIT ? GB * GLD : 1], bl[SPLIT ? GB * GLD : 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * GB, n0 = blockIdx.x * GB;
  const int64_t z = blockIdx.z;
  const float* A = d.a + z * d.a_bs;
  const float* B = d.b + z * d.b_bs;

  f4a acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0f;

  // software pipeline: the global loads of stage k0 + GK are in flight while
  // stage k0 is multiplied out of LDS
  float va[kStageLoads], vb[kStageLoads];
  stage_load(A, d.a_rs, d.a_cs, m0, d.m, 0, d.k, tid, va);
  stage_load(B, d.b_rs, d.b_cs, n0, d.n, 0, d.k, tid, vb);
  for (int k0 = 0; k0 < d.k; k0 += GK) {
    stage_store<SPLIT>(va, d.a_cs == 1, ah, al, tid);
    stage_store<SPLIT>(vb, d.b_cs == 1, bh, bl, tid);
    __syncthreads();
    if (k0 + GK < d.k) {
      stage_load(A, d.a_rs, d.a_cs, m0, d.m, k0 + GK, d.k, tid, va);
      stage_load(B, d.b_rs, d.b_cs, n0, d.n, k0 + GK, d.k, tid, vb);
    }
    const int steps = (d.k - k0 + 31) / 32 < GK / 32 ? (d.k - k0 + 31) / 32 : GK / 32;
    for (int kk = 0; kk < steps; ++kk) {
      const int col = kk * 32 + 8 * (lane >> 4);
      bf8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = *reinterpret_cast<const bf8*>(&ah[(wm * 32 + i * 16 + (lane & 15)) * GLD + col]);
        b[i] = *reinterpret_cast<const bf8*>(&bh[(wn * 32 + i * 16 + (lane & 15)) * GLD + col]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
      if (SPLIT) {
        bf8 x[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
          x[i] = *reinterpret_cast<const bf8*>(
              &bl[(wn * 32 + i * 16 + (lane & 15)) * GLD + col]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], x[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
          x[i] = *reinterpret_cast<const bf8*>(
              &al[(wm * 32 + i * 16 + (lane & 15)) * GLD + col]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x[i], b[j], acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }

  float* C = d.c + z * d.c_bs;
  const float* bias = d.bias ? d.bias + z * d.bias_bs : nullptr;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= d.m || n >= d.n) continue;
        float v = acc[i][j][r];
        if (bias) v += bias[n];
        float* c = C + int64_t(m) * d.c_rs + n;
        if (EPI == KPM_MLP_PLAIN) {
          *c = accumulate ? *c + v : v;
        } else if (EPI == KPM_MLP_BIAS_GELU) {
          *c = v;  // pre-activation Z
          d.c2[z * d.c_bs + int64_t(m) * d.c_rs + n] = gelu(v);
        } else {
          *c = v * gelu_grad(d.aux[z * d.aux_bs + int64_t(m) * d.aux_rs + n]);
        }
      }
}

// out[z][n] = sum_m A[z*bs + m*lda + n] (bias gradients): 64 columns per
// workgroup, 4 row groups each summing rows g, g+4, ... in order, then the 4
// partials added in group order (deterministic).
__global__ __launch_bounds__(256) void colsum_kernel(const float* __restrict__ A, int64_t lda,
                                                     int64_t bs, int M, int N,
                                                     float* __restrict__ out, int64_t out_bs) {
  __shared__ float part[4][64];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + c;
  const float* a = A + blockIdx.y * bs;
  float s = 0.0f;
  if (n < N)
    for (int m = g; m < M; m += 4) s += a[int64_t(m) * lda + n];
  part[g][c] = s;
  __syncthreads();
  if (g == 0 && n < N) out[blockIdx.y * out_bs + n] = ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
}

// dy = scale * (y - t); per-workgroup partial sums of (y - t)^2 in FP64.
__global__ void mse_kernel(const float* __restrict__ y, const float* __restrict__ t,
                           float* __restrict__ dy, int64_t n, float scale,
                           double* __restrict__ partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float e = y[i] - t[i];
    dy[i] = scale * e;
    s += double(e) * e;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) with bias corrections bc1 = 1/(1-b1^t), bc2 = 1/(1-b2^t).
// With sched != NULL the step reads (lr, bc1, bc2, ema decay) from device
// memory, so a captured graph of many steps follows a schedule the host
// writes between replays.  With ema != NULL the updated parameters are also
// folded into an exponential moving average, ema = d ema + (1 - d) p
// (d = 0 copies p).
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float b1, float b2, float eps, float bc1, float bc2,
                            const float* __restrict__ sched, float* __restrict__ ema,
                            float decay) {
  if (sched) {
    lr = sched[0];
    bc1 = sched[1];
    bc2 = sched[2];
    decay = sched[3];
  }
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi * bc1) / (sqrtf(vi * bc2) + eps);
    p[i] = pi;
    if (ema) ema[i] = decay * ema[i] + (1.0f - decay) * pi;
  }
}

// Per row p of an ensemble y[z][p][:] (members z = 0 .. Z-1, `width` values):
//   mean[p][n] = (1/Z) sum_z y[z][p][n]                             (if mean)
//   dev[p]     = (1/width) sum_n sqrt(sum_z (y[z][p][n] - mean_n)^2 / (Z-1))
//                                                                   (if dev)
// i.e. the L1 norm of the pointwise sample standard deviation (0 for Z = 1).
// One wavefront per row; members summed in order, lanes reduced by a fixed
// butterfly, so the result is deterministic.
__global__ __launch_bounds__(256) void ensemble_kernel(const float* __restrict__ y, int64_t z_stride,
                                                       int64_t row_stride, int members,
                                                       int64_t rows, int width,
                                                       float* __restrict__ mean,
                                                       int64_t mean_stride,
                                                       float* __restrict__ dev) {
  const int lane = threadIdx.x & 63;
  const int64_t p = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= rows) return;
  const float* yp = y + p * row_stride;
  float s = 0.0f;
  for (int n = lane; n < width; n += 64) {
    float mu = 0.0f;
    for (int q = 0; q < members; ++q) mu += yp[q * z_stride + n];
    mu *= 1.0f / float(members);
    if (mean) mean[p * mean_stride + n] = mu;
    if (dev && members > 1) {
      float var = 0.0f;
      for (int q = 0; q < members; ++q) {
        const float d = yp[q * z_stride + n] - mu;
        var += d * d;
      }
      s += sqrtf(var / float(members - 1));
    }
  }
  if (!dev) return;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) dev[p] = s / float(width);
}

template <int EPI>
void launch_gemm(const kpm_mlp_gemm_desc& d, bool split, int accumulate, hipStream_t s) {
  dim3 grid((d.n + GB - 1) / GB, (d.m + GB - 1) / GB, d.batch), block(256);
  if (split)
    hipLaunchKernelGGL((mlp_gemm_kernel<EPI, true>), grid, block, 0, s, d, accumulate);
  else
    hipLaunchKernelGGL((mlp_gemm_kernel<EPI, false>), grid, block, 0, s, d, accumulate);
}

}  // namespace

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_mlp_gemm(const kpm_mlp_gemm_desc* d, kpm_stream_t stream) {
  if (!d) return fail(KPM_ERR_ARG, "NULL descriptor");
  if (!d->a || !d->b || !d->c) return fail(KPM_ERR_ARG, "NULL operand");
  if (d->m <= 0 || d->n <= 0 || d->k <= 0 || d->batch <= 0)
    return fail(KPM_ERR_ARG, "empty GEMM");
  if (d->batch > 65535 || d->m > 65535 * GB) return fail(KPM_ERR_ARG, "GEMM grid too large");
  if (d->c_rs < d->n) return fail(KPM_ERR_ARG, "row stride of C smaller than N");
  const int epi = d->flags & 3;
  const bool split = (d->flags & KPM_MLP_BF16X3) != 0;
  const int accumulate = (d->flags & KPM_MLP_ACCUMULATE) != 0;
  if (d->flags & ~(3 | KPM_MLP_BF16X3 | KPM_MLP_ACCUMULATE))
    return fail(KPM_ERR_ARG, "unknown GEMM flags");
  if (accumulate && epi != KPM_MLP_PLAIN)
    return fail(KPM_ERR_ARG, "accumulate only with the plain epilogue");
  hipStream_t s = (hipStream_t)stream;
  switch (epi) {
    case KPM_MLP_PLAIN:
      launch_gemm<KPM_MLP_PLAIN>(*d, split, accumulate, s);
      break;
    case KPM_MLP_BIAS_GELU:
      if (!d->c2) return fail(KPM_ERR_ARG, "activation output missing");
      launch_gemm<KPM_MLP_BIAS_GELU>(*d, split, 0, s);
      break;
    case KPM_MLP_GELU_GRAD:
      if (!d->aux || d->aux_rs < d->n) return fail(KPM_ERR_ARG, "pre-activation input missing");
      launch_gemm<KPM_MLP_GELU_GRAD>(*d, split, 0, s);
      break;
    default:
      return fail(KPM_ERR_ARG, "unknown epilogue");
  }
  return check_hip(hipGetLastError(), "mlp_gemm launch");
}

int kpm_mlp_colsum(const float* A, int64_t lda, int64_t a_bs, int32_t M, int32_t N,
                   int32_t batch, float* out, int64_t out_bs, kpm_stream_t stream) {
  if (!A || !out || M <= 0 || N <= 0 || batch <= 0 || batch > 65535 || lda < N)
    return fail(KPM_ERR_ARG, "bad colsum arguments");
  hipLaunchKernelGGL(colsum_kernel, dim3((N + 63) / 64, batch), dim3(256), 0,
                     (hipStream_t)stream, A, lda, a_bs, M, N, out, out_bs);
  return check_hip(hipGetLastError(), "colsum launch");
}

int kpm_mlp_mse(const float* y, const float* t, float* dy, int64_t n, float scale,
                double* partial, int32_t n_partial, kpm_stream_t stream) {
  if (!y || !t || !dy || !partial || n <= 0 || n_partial <= 0 || n_partial > 65535)
    return fail(KPM_ERR_ARG, "bad mse arguments");
  hipLaunchKernelGGL(mse_kernel, dim3(n_partial), dim3(256), 0, (hipStream_t)stream, y, t, dy, n,
                     scale, partial);
  return check_hip(hipGetLastError(), "mse launch");
}

int kpm_mlp_ensemble(const float* y, int64_t z_stride, int64_t row_stride, int32_t members,
                     int64_t rows, int32_t width, float* mean, int64_t mean_stride,
                     float* dev, kpm_stream_t stream) {
  if (!y || members <= 0 || rows <= 0 || width <= 0 || row_stride < width ||
      (mean && mean_stride < width) || (!mean && !dev))
    return fail(KPM_ERR_ARG, "bad ensemble arguments");
  if (dev && members < 2)
    return fail(KPM_ERR_ARG, "the spread needs at least two members");
  if ((rows + 3) / 4 > (int64_t(1) << 31) - 1) return fail(KPM_ERR_ARG, "too many rows");
  hipLaunchKernelGGL(ensemble_kernel, dim3(unsigned((rows + 3) / 4)), dim3(256), 0,
                     (hipStream_t)stream, y, z_stride, row_stride, members, rows, width, mean,
                     mean_stride, dev);
  return check_hip(hipGetLastError(), "ensemble launch");
}

int kpm_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                 float b2, float eps, int32_t t, const float* sched, float* ema,
                 float ema_decay, kpm_stream_t stream) {
  if (!p || !g || !m || !v || n <= 0 || (t <= 0 && !sched))
    return fail(KPM_ERR_ARG, "bad adam arguments");
  if (sched) t = 1;
  if (!(b1 >= 0.0f && b1 < 1.0f && b2 >= 0.0f && b2 < 1.0f && lr >= 0.0f && eps > 0.0f))
    return fail(KPM_ERR_ARG, "bad adam hyper-parameters");
  const float bc1 = float(1.0 / (1.0 - std::pow(double(b1), double(t))));
  const float bc2 = float(1.0 / (1.0 - std::pow(double(b2), double(t))));
  const int64_t blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
  hipLaunchKernelGGL(adam_kernel, dim3(unsigned(blocks)), dim3(256), 0, (hipStream_t)stream, p,
                     g, m, v, n, lr, b1, b2, eps, bc1, bc2, sched, ema, ema_decay);
  return check_hip(hipGetLastError(), "adam launch");
}

}  // extern "C"
