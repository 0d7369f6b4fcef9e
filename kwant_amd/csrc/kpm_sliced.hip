// Correlator moment matrix mu[m, n] += sum_k conj(A[m, k]) B[n, k] from
// complex128 operands on the BF16 matrix cores, exact in every term it keeps
// (a split-integer scheme after Ozaki et al., "Error-free transformations of
// matrix multiplication", Numer. Algorithms 59 (2012)).
//
// The k range is processed in panels.  For each panel and each operand row r
// the largest component magnitude gives an exponent e_r (max < 2^e_r), and
// every real component x (re and im interleaved, k' = 2k + {0, 1}) is cut into
// S integer slices of at most 8 bits:
//     v = x 2^(7 - e_r);  q_1 = rint(v);  v = (v - q_1) 256;  q_2 = rint(v); ...
//     x = 2^(e_r - 7) (q_1 + q_2 2^-8 + ... + q_S 2^(8 - 8S)) + O(2^(e_r - 8S - 8))
// |q_i| <= 128, so each q is exact in BF16 and each product q_i q'_j is an
// integer of at most 2^14 -- exact in the FP32 accumulator of
// v_mfma_f32_16x16x32_bf16.  Products are grouped by level l = i + j (0-based
// slices): a level has at most S pairs, so over 256 real k' its partial sum is
// an integer of magnitude <= S 256 2^14 <= 2^24 for S <= 4 and stays exact.
// Every 256 k' the level sums are folded into FP64 as sum_l acc_l 2^(-8l)
// (exact scaling), and the tile is finally scaled by 2^(e_m + e_n - 14).
// Levels l >= S are dropped: the result equals the exact product of the
// S-slice truncations up to those cross terms, about 2^(-8S) of
// max|A_m| max|B_n| per term.
//
// Re mu = sum_k' A'[m, k'] B'[n, k'] with the interleaved real rows, and
// Im mu = sum_k' A'[m, k'] B''[n, k'] with B'' = (b_im, -b_re), formed from
// B's slices in registers (swap and negate within a pair: exact).

#include "kpm_internal.h"

namespace kpm {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4a __attribute__((ext_vector_type(4)));
typedef short s8v __attribute__((ext_vector_type(8)));
typedef double d2v __attribute__((ext_vector_type(2)));

constexpr int SB = 64;           // output tile (rows and cols)
constexpr int SK = 64;           // real k' per LDS stage
constexpr int SLD = SK + 8;      // LDS row stride in bf16 (16-byte aligned rows)
constexpr int FOLD = 256;        // real k' per exact FP32 interval
constexpr int kMaxSlices = 4;

// e[r] such that max |x| over the panel's 2 kn components of row r is < 2^e
// (0 for an all-zero row), in three launches: clear, a grid of row spans
// that each fold their maximum's exponent in with an integer atomicMax
// (order-independent), and the decode of the biased exponent.  A row holds
// up to 2^21 components, so one workgroup per row would leave most CUs idle.
constexpr int kExpBias = 1100;     // > -(smallest double exponent): biased values > 0
constexpr int kExpSpan = 16384;    // components per workgroup

__global__ __launch_bounds__(256) void exponent_clear(int* __restrict__ e, int rows) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) e[r] = 0;
}

__global__ __launch_bounds__(256) void row_exponent_kernel(const double* __restrict__ X,
                                                           int64_t ld, int rows, int64_t k0,
                                                           int64_t kn, int* __restrict__ e) {
  const int r = blockIdx.y;
  if (r >= rows) return;
  const double* x = X + 2 * (int64_t(r) * ld + k0);
  const int64_t n2 = 2 * kn;
  const int64_t begin = int64_t(blockIdx.x) * kExpSpan;
  const int64_t end = begin + kExpSpan < n2 ? begin + kExpSpan : n2;
  double m = 0.0;
  // (re, im) pairs: 16-byte aligned, n2 and begin are even
  for (int64_t i = begin + 2 * threadIdx.x; i < end; i += 2 * blockDim.x) {
    const d2v v = *reinterpret_cast<const d2v*>(x + i);
    m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
    if (m > 0.0) {
      int ex = 0;
      frexp(m, &ex);
      atomicMax(e + r, ex + kExpBias);
    }
  }
}

__global__ __launch_bounds__(256) void exponent_decode(int* __restrict__ e, int rows) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows && e[r] != 0) e[r] -= kExpBias;
}

// S slices of every component of the panel, rows padded to kp2 with zeros:
// out[(s * rows + r) * kp2 + k'].
__global__ __launch_bounds__(256) void slice_kernel(const double* __restrict__ X, int64_t ld,
                                                    int rows, int64_t k0, int64_t kn,
                                                    const int* __restrict__ e, int S,
                                                    __bf16* __restrict__ out, int64_t kp2) {
  const int r = blockIdx.y;
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= kp2 || r >= rows) return;
  double v = c < 2 * kn ? X[2 * (int64_t(r) * ld + k0) + c] : 0.0;
  v = ldexp(v, 7 - e[r]);
  for (int s = 0; s < S; ++s) {
    const double q = rint(v);
    out[(int64_t(s) * rows + r) * kp2 + c] = (__bf16)float(q);
    v = (v - q) * 256.0;
  }
}

__device__ inline bf8 swap_negate(bf8 b) {
  // (b1, -b0, b3, -b2, ...): the B'' = (im, -re) pairing, exact in bf16
  s8v u = __builtin_bit_cast(s8v, b);
  s8v w;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    w[2 * p] = u[2 * p + 1];
    w[2 * p + 1] = short(u[2 * p] ^ short(0x8000));
  }
  return __builtin_bit_cast(bf8, w);
}

template <int S>
__global__ __launch_bounds__(256) void sliced_gemm_kernel(
    const __bf16* __restrict__ As, const __bf16* __restrict__ Bs, int m_rows, int n_rows,
    int64_t kp2, int64_t kps, const int* __restrict__ ea, const int* __restrict__ eb,
    double* __restrict__ out, int64_t ldo, int accumulate, int nblocks, int zs) {
  __shared__ __bf16 la[S][SB * SLD];
  __shared__ __bf16 lb[S][SB * SLD];
  const int mt = (m_rows + SB - 1) / SB, nt = (n_rows + SB - 1) / SB;
  int m_tile, n_tile, split;
  gemm_tile_of(blockIdx.x, nblocks, mt, nt, &m_tile, &n_tile, &split);
  if (split >= zs) return;  // padding of the grid to a multiple of 8
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = m_tile * SB, n0 = n_tile * SB;
  const int64_t kb = int64_t(split) * kps;
  const int64_t ke = kb + kps < kp2 ? kb + kps : kp2;

  f4a acc_r[S][2][2], acc_i[S][2][2];
  double dr[2][2][4], di[2][2][4];
#pragma unroll
  for (int l = 0; l < S; ++l)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc_r[l][i][j][r] = acc_i[l][i][j][r] = 0.0f;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) dr[i][j][r] = di[i][j][r] = 0.0;

  // stage S slices of both tiles (64 rows x 64 k' bf16 = 8 x 16 B per row);
  // the next stage's global loads are in flight during this stage's products
  bf8 va[S][2], vb[S][2];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int e = tid + 256 * q;  // 512 chunks of 8 bf16 per tile
        const int row = e >> 3, col = (e & 7) * 8;
        va[s][q] = bf8{};
        vb[s][q] = bf8{};
        if (m0 + row < m_rows)
          va[s][q] = *reinterpret_cast<const bf8*>(As + (int64_t(s) * m_rows + m0 + row) * kp2 + k0 + col);
        if (n0 + row < n_rows)
          vb[s][q] = *reinterpret_cast<const bf8*>(Bs + (int64_t(s) * n_rows + n0 + row) * kp2 + k0 + col);
      }
  };
  if (kb < ke) fetch(kb);
  int since = 0;
  for (int64_t k0 = kb; k0 < ke; k0 += SK) {
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int e = tid + 256 * q;
        const int row = e >> 3, col = (e & 7) * 8;
        *reinterpret_cast<bf8*>(&la[s][row * SLD + col]) = va[s][q];
        *reinterpret_cast<bf8*>(&lb[s][row * SLD + col]) = vb[s][q];
      }
    __syncthreads();
    if (k0 + SK < ke) fetch(k0 + SK);
#pragma unroll
    for (int kk = 0; kk < SK; kk += 32) {
      const int col = kk + 8 * (lane >> 4);
      // every slice's fragments once per k-step, then all the products
      bf8 a[S][2], b[S][2], bs[S][2];
#pragma unroll
      for (int q = 0; q < S; ++q)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          a[q][t] = *reinterpret_cast<const bf8*>(&la[q][(wm * 32 + t * 16 + (lane & 15)) * SLD + col]);
          b[q][t] = *reinterpret_cast<const bf8*>(&lb[q][(wn * 32 + t * 16 + (lane & 15)) * SLD + col]);
        }
#pragma unroll
      for (int q = 0; q < S; ++q)
#pragma unroll
        for (int t = 0; t < 2; ++t) bs[q][t] = swap_negate(b[q][t]);
#pragma unroll
      for (int l = 0; l < S; ++l)
#pragma unroll
        for (int sa = 0; sa <= l; ++sa)
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              acc_r[l][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[sa][i], b[l - sa][j], acc_r[l][i][j], 0, 0, 0);
              acc_i[l][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[sa][i], bs[l - sa][j], acc_i[l][i][j], 0, 0, 0);
            }
    }
    __syncthreads();
    since += SK;
    if (since >= FOLD || k0 + SK >= ke) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int l = S - 1; l >= 0; --l) {
              sr += ldexp(double(acc_r[l][i][j][r]), -8 * l);
              si += ldexp(double(acc_i[l][i][j][r]), -8 * l);
              acc_r[l][i][j][r] = 0.0f;
              acc_i[l][i][j][r] = 0.0f;
            }
            dr[i][j][r] += sr;
            di[i][j][r] += si;
          }
      since = 0;
    }
  }

  double* o = out + int64_t(split) * int64_t(m_rows) * ldo * 2;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= m_rows || n >= n_rows) continue;
        const int sc = ea[m] + eb[n] - 14;
        double* p = o + 2 * (int64_t(m) * ldo + n);
        const double vr = ldexp(dr[i][j][r], sc), vi = ldexp(di[i][j][r], sc);
        if (accumulate) {
          p[0] += vr;
          p[1] += vi;
        } else {
          p[0] = vr;
          p[1] = vi;
        }
      }
}

__global__ void add_partials(const double* __restrict__ work, int slices, double* mu,
                             int64_t ldc, int m_rows, int n_rows) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(m_rows) * n_rows) return;
  const int m = int(t / n_rows), n = int(t % n_rows);
  double sr = 0.0, si = 0.0;
  for (int z = 0; z < slices; ++z) {
    const double* p = work + 2 * (int64_t(z) * m_rows * n_rows + t);
    sr += p[0];
    si += p[1];
  }
  mu[2 * (int64_t(m) * ldc + n)] += sr;
  mu[2 * (int64_t(m) * ldc + n) + 1] += si;
}

struct SlicedPlan {
  int64_t panel;      // complex k per panel
  int64_t kp2;        // padded real k' per panel row
  int64_t zs;         // k-splits per panel
  int64_t kps;        // real k' per split (multiple of SK)
  int64_t ldk;        // row pitch of the int8 slice arrays in bytes
  int64_t slice_bytes, exp_bytes, part_bytes;
};

SlicedPlan sliced_plan(int m_rows, int n_rows, int64_t k, int slices) {
  SlicedPlan p;
  p.panel = k < (int64_t(1) << 20) ? k : (int64_t(1) << 20);
  p.kp2 = (2 * p.panel + SK - 1) / SK * SK;
  p.ldk = p.kp2;
  const int64_t tiles = int64_t((m_rows + SB - 1) / SB) * ((n_rows + SB - 1) / SB);
  int64_t zs = (2048 + tiles - 1) / tiles;
  const int64_t by_k = (p.kp2 + 4 * FOLD - 1) / (4 * FOLD);
  if (zs > by_k) zs = by_k;
  if (zs < 1) zs = 1;
  p.kps = ((p.kp2 + zs - 1) / zs + SK - 1) / SK * SK;
  p.zs = (p.kp2 + p.kps - 1) / p.kps;
  p.slice_bytes = int64_t(slices) * (m_rows + n_rows) * p.kp2 * 2;
  p.exp_bytes = int64_t(m_rows + n_rows) * 4;
  p.part_bytes = p.zs * int64_t(m_rows) * n_rows * 2 * 8;
  return p;
}

template <int S>
void launch_sliced_gemm(const __bf16* As, const __bf16* Bs, int m_rows, int n_rows,
                        const SlicedPlan& p, const int* ea, const int* eb, double* part,
                        int accumulate, hipStream_t s) {
  const int64_t tiles = int64_t((n_rows + SB - 1) / SB) * ((m_rows + SB - 1) / SB);
  // a multiple of the XCD count (the kernel's remap needs it; the padding
  // workgroups return at once)
  const int64_t blocks = (tiles * p.zs + kNumXcd - 1) / kNumXcd * kNumXcd;
  hipLaunchKernelGGL(sliced_gemm_kernel<S>, dim3(unsigned(blocks)), dim3(256), 0, s, As, Bs,
                     m_rows, n_rows, p.kp2, p.kps, ea, eb, part, int64_t(n_rows), accumulate,
                     int(blocks), int(p.zs));
}

// ---------------------------------------------------------------------------
// The same product on the INT8 matrix cores (v_mfma_i32_32x32x32_i8), after
// Ootomo, Ozaki and Yokota, "DGEMM on integer matrix multiplication unit"
// (Int. J. High Perform. Comput. Appl., 2024): slices of at most 7 bits and
// exact int32 accumulation.
//     v = x 2^(6 - e_r);  q_1 = rint(v);  v = (v - q_1) 128;  q_2 = rint(v); ...
//     x = 2^(e_r - 6) (q_1 + q_2 2^-7 + ... + q_S 2^(7 - 7S)) + t,  |t| <= 2^(e_r - 7S)
// |v| < 64 before each rint, so |q_i| <= 64 fits int8 and every product is at
// most 2^12 in magnitude.  A level l = i + j (0-based slices) has at most S
// pairs, so its sum over 65536 k' is below S 2^28 < 2^31 for S <= 6: the
// int32 accumulator is exact over that interval.  Every 65536 k' (and at the
// end of the workgroup's k range) each level is folded into the FP64 partial
// output as sum_l acc_l 2^(-7l), scaled by 2^(e_m + e_n - 12).  Against the
// BF16 form an int8 slice carries one bit less, but takes half the bytes,
// multiplies at twice the rate and needs no FP32 fold every 256 k'.

typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));

constexpr int IB = 64;               // output tile (rows and cols)
constexpr int IKS = 64;              // real k' per LDS stage (two MFMA k-steps)
constexpr int ILD = IKS + 16;        // LDS row stride in bytes (16-byte aligned rows)
constexpr int64_t IFOLD = 65536;     // real k' per exact int32 interval
constexpr int kMaxISlices = 6;
// word 3 of a raw (stride 0) buffer descriptor on this GPU family: 32-bit
// data format, no swizzle
constexpr int kBufferWord3 = 0x00020000;
// Slice rows are kp2 bytes (a power of two at full panels) plus this much
// padding apart, so that the rows a workgroup reads at one k' do not all
// fall on the same L2 set and channel.
constexpr int kIPad = 256;

// S int8 slices of every component of the panel, 4 components per thread,
// rows padded to kp2 with zeros: out[(s * slice_rows + r) * ldk + k'] (the
// rows may be a part of a larger sliced operand of slice_rows rows).
__global__ __launch_bounds__(256) void islice_kernel(const double* __restrict__ X, int64_t ld,
                                                     int rows, int64_t k0, int64_t kn,
                                                     const int* __restrict__ e, int S,
                                                     int8_t* __restrict__ out, int64_t kp2,
                                                     int64_t ldk, int64_t slice_rows) {
  const int r = blockIdx.y;
  const int64_t c = 4 * (int64_t(blockIdx.x) * blockDim.x + threadIdx.x);
  if (c >= kp2 || r >= rows) return;
  const double* x = X + 2 * (int64_t(r) * ld + k0);
  double v[4];
  if (c + 4 <= 2 * kn) {
    // (re, im) pairs are 16-byte aligned and c is a multiple of 4
    const d2v p0 = *reinterpret_cast<const d2v*>(x + c);
    const d2v p1 = *reinterpret_cast<const d2v*>(x + c + 2);
    v[0] = p0.x;
    v[1] = p0.y;
    v[2] = p1.x;
    v[3] = p1.y;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = c + j < 2 * kn ? x[c + j] : 0.0;
  }
  const int sh = 6 - e[r];
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = ldexp(v[j], sh);
  for (int s = 0; s < S; ++s) {
    uint32_t packed = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double q = rint(v[j]);
      packed |= (uint32_t(int(q)) & 0xFFu) << (8 * j);
      v[j] = (v[j] - q) * 128.0;
    }
    *reinterpret_cast<uint32_t*>(out + (int64_t(s) * slice_rows + r) * ldk + c) = packed;
  }
}

// (b1, -b0, b3, -b2, ...) on packed int8: the B'' = (im, -re) pairing.  The
// even bytes are negated as 0x100 - b in 16-bit lanes (no borrow crosses a
// lane; the low byte of the difference is -b mod 256) and moved up one byte.
__device__ inline i4v swap_negate_i8(i4v b) {
  i4v w;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const uint32_t u = uint32_t(b[p]);
    const uint32_t neg = 0x01000100u - (u & 0x00FF00FFu);
    w[p] = int(((u >> 8) & 0x00FF00FFu) | ((neg & 0x00FF00FFu) << 8));
  }
  return w;
}

static_assert(IKS == 64, "the staging loads 4 x 16 B per row and stage");

template <int S>
__device__ __forceinline__ void islice_gemm_body(
    const int8_t* __restrict__ As, const int8_t* __restrict__ Bs, int m_rows, int n_rows,
    int64_t kp2, int64_t ldk, int64_t kps, const int* __restrict__ ea,
    const int* __restrict__ eb, double* __restrict__ out, int64_t ldo, int slot0, int nblocks,
    int zs) {
  __shared__ __attribute__((aligned(16))) int8_t la[S][IB * ILD];
  __shared__ __attribute__((aligned(16))) int8_t lb[S][IB * ILD];
  const int mt = (m_rows + IB - 1) / IB, nt = (n_rows + IB - 1) / IB;
  int m_tile, n_tile, split;
  gemm_tile_of(blockIdx.x, nblocks, mt, nt, &m_tile, &n_tile, &split);
  if (split >= zs) return;  // padding of the grid to a multiple of 8
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = m_tile * IB, n0 = n_tile * IB;
  // kps <= IFOLD (islice_plan), so the int32 level sums are exact over the
  // whole range and are folded once, at the end
  const int64_t kb = int64_t(split) * kps;
  const int64_t ke = kb + kps < kp2 ? kb + kps : kp2;

  // one 32 x 32 tile per wave: 16 int32 per lane and level, Re and Im
  i16v acc_r[S], acc_i[S];
#pragma unroll
  for (int l = 0; l < S; ++l)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc_r[l][r] = acc_i[l][r] = 0;

  // staging: per slice and operand 64 rows x 64 B = 256 threads x 16 B; the
  // next stage's global loads are in flight during this stage's products.
  // Buffer loads: one descriptor per slice and operand covers the tile's
  // valid rows (num_records = rows x ldk < 2^28), so a lane needs a single
  // 32-bit offset, the stage's k' goes in the scalar offset, and rows past
  // the matrix read as zeros (out of the descriptor's range).
  const int srow = tid >> 2, scol = (tid & 3) * 16;
  const int rows_a = m_rows - m0 < IB ? m_rows - m0 : IB;
  const int rows_b = n_rows - n0 < IB ? n_rows - n0 : IB;
  const int voff = srow * int(ldk) + scol;
  __amdgpu_buffer_rsrc_t ra[S], rb[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    ra[s] = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int8_t*>(As + (int64_t(s) * m_rows + m0) * ldk), short(0),
        rows_a * int(ldk), kBufferWord3);
    rb[s] = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int8_t*>(Bs + (int64_t(s) * n_rows + n0) * ldk), short(0),
        rows_b * int(ldk), kBufferWord3);
  }
  i4v va[S], vb[S];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      va[s] = __builtin_bit_cast(i4v, __builtin_amdgcn_raw_buffer_load_b128(ra[s], voff, int(k0), 0));
      vb[s] = __builtin_bit_cast(i4v, __builtin_amdgcn_raw_buffer_load_b128(rb[s], voff, int(k0), 0));
    }
  };

  if (kb < ke) fetch(kb);
  const int fr = lane & 31, fc = 16 * (lane >> 5);
  for (int64_t k0 = kb; k0 < ke; k0 += IKS) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      *reinterpret_cast<i4v*>(&la[s][srow * ILD + scol]) = va[s];
      *reinterpret_cast<i4v*>(&lb[s][srow * ILD + scol]) = vb[s];
    }
    __syncthreads();
    if (k0 + IKS < ke) fetch(k0 + IKS);
#pragma unroll
    for (int kk = 0; kk < IKS; kk += 32) {
      // A and B fragments hold the same k' bytes lane for lane, so the
      // products pair a_re with b_re and a_im with b_im, and after the swap
      // a_re with b_im and a_im with -b_re.
      i4v a[S];
#pragma unroll
      for (int q = 0; q < S; ++q)
        a[q] = *reinterpret_cast<const i4v*>(&la[q][(wm * 32 + fr) * ILD + kk + fc]);
#pragma unroll
      for (int sb = 0; sb < S; ++sb) {
        const i4v b = *reinterpret_cast<const i4v*>(&lb[sb][(wn * 32 + fr) * ILD + kk + fc]);
        const i4v bs = swap_negate_i8(b);
#pragma unroll
        for (int sa = 0; sa + sb < S; ++sa) {
          acc_r[sa + sb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[sa], b, acc_r[sa + sb], 0, 0, 0);
          acc_i[sa + sb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[sa], bs, acc_i[sa + sb], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }

  // Fold the exact level sums into this (panel, split)'s FP64 partial output
  // slot; each element belongs to one lane of one workgroup.  32 x 32 MFMA
  // layout: lane l holds column l mod 32 and rows 8 (r / 4) + 4 (l / 32) + r mod 4.
  double* o = out + int64_t(slot0 + split) * int64_t(m_rows) * ldo * 2;
  const int n = n0 + wn * 32 + (lane & 31);
  if (n >= n_rows) return;
  const int en = eb[n] - 12;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (m >= m_rows) continue;
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int l = S - 1; l >= 0; --l) {
      sr += ldexp(double(acc_r[l][r]), -7 * l);
      si += ldexp(double(acc_i[l][r]), -7 * l);
    }
    const int sc = ea[m] + en;
    double* p = o + 2 * (int64_t(m) * ldo + n);
    p[0] = ldexp(sr, sc);
    p[1] = ldexp(si, sc);
  }
}

// Two waves per SIMD: up to 5 slices, the accumulators (32 registers per
// slice) and the staging fit in 256 registers (measured with 4 slices: 0.143
// against 0.196 s per 1024 x 1024 x 2^21 product at one wave); 6 slices do
// not, and run at one wave per SIMD.
template <int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
islice_gemm_kernel(const int8_t* __restrict__ As, const int8_t* __restrict__ Bs, int m_rows,
                   int n_rows, int64_t kp2, int64_t ldk, int64_t kps, const int* __restrict__ ea,
                   const int* __restrict__ eb, double* __restrict__ out, int64_t ldo, int slot0,
                   int nblocks, int zs) {
  islice_gemm_body<S>(As, Bs, m_rows, n_rows, kp2, ldk, kps, ea, eb, out, ldo, slot0, nblocks,
                      zs);
}

__global__ __launch_bounds__(256) void islice_gemm_kernel_wide(
    const int8_t* __restrict__ As, const int8_t* __restrict__ Bs, int m_rows, int n_rows,
    int64_t kp2, int64_t ldk, int64_t kps, const int* __restrict__ ea, const int* __restrict__ eb,
    double* __restrict__ out, int64_t ldo, int slot0, int nblocks, int zs) {
  islice_gemm_body<kMaxISlices>(As, Bs, m_rows, n_rows, kp2, ldk, kps, ea, eb, out, ldo, slot0,
                                nblocks, zs);
}

SlicedPlan islice_plan(int m_rows, int n_rows, int64_t k, int slices) {
  SlicedPlan p;
  p.panel = k < (int64_t(1) << 20) ? k : (int64_t(1) << 20);
  p.kp2 = (2 * p.panel + IKS - 1) / IKS * IKS;
  const int64_t tiles = int64_t((m_rows + IB - 1) / IB) * ((n_rows + IB - 1) / IB);
  int64_t zs = (2048 + tiles - 1) / tiles;
  const int64_t by_k = (p.kp2 + 16383) / 16384;
  if (zs > by_k) zs = by_k;
  const int64_t exact = (p.kp2 + IFOLD - 1) / IFOLD;  // at most IFOLD k' per split
  if (zs < exact) zs = exact;
  if (zs < 1) zs = 1;
  p.kps = ((p.kp2 + zs - 1) / zs + IKS - 1) / IKS * IKS;
  if (p.kps > IFOLD) p.kps = IFOLD;
  p.zs = (p.kp2 + p.kps - 1) / p.kps;
  const int64_t panels = (k + p.panel - 1) / p.panel;
  p.ldk = p.kp2 + kIPad;
  p.slice_bytes = int64_t(slices) * (m_rows + n_rows) * p.ldk;
  p.exp_bytes = int64_t(m_rows + n_rows) * 4;
  // one FP64 partial slot per (panel, split)
  p.part_bytes = panels * p.zs * int64_t(m_rows) * n_rows * 2 * 8;
  return p;
}

template <int S>
void launch_islice_gemm(const int8_t* As, const int8_t* Bs, int m_rows, int n_rows,
                        const SlicedPlan& p, const int* ea, const int* eb, double* part,
                        int slot0, hipStream_t s) {
  const int64_t tiles = int64_t((n_rows + IB - 1) / IB) * ((m_rows + IB - 1) / IB);
  const int64_t blocks = (tiles * p.zs + kNumXcd - 1) / kNumXcd * kNumXcd;
  if constexpr (S == kMaxISlices)
    hipLaunchKernelGGL(islice_gemm_kernel_wide, dim3(unsigned(blocks)), dim3(256), 0, s, As, Bs,
                       m_rows, n_rows, p.kp2, p.ldk, p.kps, ea, eb, part, int64_t(n_rows), slot0,
                       int(blocks), int(p.zs));
  else
    hipLaunchKernelGGL(islice_gemm_kernel<S>, dim3(unsigned(blocks)), dim3(256), 0, s, As, Bs,
                       m_rows, n_rows, p.kp2, p.ldk, p.kps, ea, eb, part, int64_t(n_rows), slot0,
                       int(blocks), int(p.zs));
}

}  // namespace

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_corr_gemm_sliced_work(int32_t m_rows, int32_t n_rows, int64_t k, int32_t slices,
                              int64_t* bytes) {
  if (!bytes || m_rows <= 0 || n_rows <= 0 || k <= 0 || slices < 1 || slices > kMaxSlices)
    return fail(KPM_ERR_ARG, "bad sliced GEMM shape");
  if (m_rows > 65535 || n_rows > 65535)
    return fail(KPM_ERR_ARG, "at most 65535 rows per operand");
  const SlicedPlan p = sliced_plan(m_rows, n_rows, k, slices);
  *bytes = p.slice_bytes + p.exp_bytes + p.part_bytes + 256;
  return KPM_OK;
}

int kpm_corr_gemm_sliced(const void* A, int64_t lda, const void* B, int64_t ldb, double* mu,
                         int64_t ldc, int32_t m_rows, int32_t n_rows, int64_t k,
                         int32_t slices, void* work, int64_t work_bytes, kpm_stream_t stream) {
  if (!A || !B || !mu || !work) return fail(KPM_ERR_ARG, "NULL buffer");
  if (m_rows <= 0 || n_rows <= 0 || k <= 0) return fail(KPM_ERR_ARG, "empty GEMM");
  if (m_rows > 65535 || n_rows > 65535)
    return fail(KPM_ERR_ARG, "at most 65535 rows per operand");
  if (slices < 1 || slices > kMaxSlices) return fail(KPM_ERR_ARG, "slices must be 1..4");
  if (lda < k || ldb < k || ldc < n_rows) return fail(KPM_ERR_ARG, "leading dimension too small");
  const SlicedPlan p = sliced_plan(m_rows, n_rows, k, slices);
  if (work_bytes < p.slice_bytes + p.exp_bytes + p.part_bytes + 256)
    return fail(KPM_ERR_ARG, "work buffer too small (kpm_corr_gemm_sliced_work)");
  hipStream_t s = (hipStream_t)stream;
  char* w = static_cast<char*>(work);
  double* part = reinterpret_cast<double*>(w);
  int* ea = reinterpret_cast<int*>(w + p.part_bytes);
  int* eb = ea + m_rows;
  __bf16* As = reinterpret_cast<__bf16*>(w + ((p.part_bytes + p.exp_bytes + 255) / 256) * 256);
  __bf16* Bs = As + int64_t(slices) * m_rows * p.kp2;
  const double* a = static_cast<const double*>(A);
  const double* b = static_cast<const double*>(B);
  int panel_no = 0;
  for (int64_t k0 = 0; k0 < k; k0 += p.panel, ++panel_no) {
    const int64_t kn = k - k0 < p.panel ? k - k0 : p.panel;
    const unsigned spans = unsigned((2 * kn + kExpSpan - 1) / kExpSpan);
    hipLaunchKernelGGL(exponent_clear, dim3((m_rows + 255) / 256), dim3(256), 0, s, ea, m_rows);
    hipLaunchKernelGGL(exponent_clear, dim3((n_rows + 255) / 256), dim3(256), 0, s, eb, n_rows);
    hipLaunchKernelGGL(row_exponent_kernel, dim3(spans, m_rows), dim3(256), 0, s, a, lda,
                       m_rows, k0, kn, ea);
    hipLaunchKernelGGL(row_exponent_kernel, dim3(spans, n_rows), dim3(256), 0, s, b, ldb,
                       n_rows, k0, kn, eb);
    hipLaunchKernelGGL(exponent_decode, dim3((m_rows + 255) / 256), dim3(256), 0, s, ea, m_rows);
    hipLaunchKernelGGL(exponent_decode, dim3((n_rows + 255) / 256), dim3(256), 0, s, eb, n_rows);
    const dim3 sg(unsigned((p.kp2 + 255) / 256));
    hipLaunchKernelGGL(slice_kernel, dim3(sg.x, m_rows), dim3(256), 0, s, a, lda, m_rows, k0,
                       kn, ea, int(slices), As, p.kp2);
    hipLaunchKernelGGL(slice_kernel, dim3(sg.x, n_rows), dim3(256), 0, s, b, ldb, n_rows, k0,
                       kn, eb, int(slices), Bs, p.kp2);
    const int acc = panel_no > 0;
    switch (slices) {
      case 1: launch_sliced_gemm<1>(As, Bs, m_rows, n_rows, p, ea, eb, part, acc, s); break;
      case 2: launch_sliced_gemm<2>(As, Bs, m_rows, n_rows, p, ea, eb, part, acc, s); break;
      case 3: launch_sliced_gemm<3>(As, Bs, m_rows, n_rows, p, ea, eb, part, acc, s); break;
      default: launch_sliced_gemm<4>(As, Bs, m_rows, n_rows, p, ea, eb, part, acc, s); break;
    }
    int rc = check_hip(hipGetLastError(), "sliced GEMM launch");
    if (rc) return rc;
  }
  const int64_t total = int64_t(m_rows) * n_rows;
  hipLaunchKernelGGL(add_partials, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, part,
                     int(p.zs), mu, ldc, m_rows, n_rows);
  return check_hip(hipGetLastError(), "sliced GEMM partial sum");
}

int kpm_corr_gemm_islice_work(int32_t m_rows, int32_t n_rows, int64_t k, int32_t slices,
                              int64_t* bytes) {
  if (!bytes || m_rows <= 0 || n_rows <= 0 || k <= 0 || slices < 1 || slices > kMaxISlices)
    return fail(KPM_ERR_ARG, "bad int8-sliced GEMM shape");
  if (m_rows > 65535 || n_rows > 65535)
    return fail(KPM_ERR_ARG, "at most 65535 rows per operand");
  const SlicedPlan p = islice_plan(m_rows, n_rows, k, slices);
  *bytes = p.slice_bytes + p.exp_bytes + p.part_bytes + 256;
  return KPM_OK;
}

int kpm_corr_gemm_islice(const void* A, int64_t lda, const void* B, int64_t ldb, double* mu,
                         int64_t ldc, int32_t m_rows, int32_t n_rows, int64_t k,
                         int32_t slices, void* work, int64_t work_bytes, kpm_stream_t stream) {
  if (!A || !B || !mu || !work) return fail(KPM_ERR_ARG, "NULL buffer");
  if (m_rows <= 0 || n_rows <= 0 || k <= 0) return fail(KPM_ERR_ARG, "empty GEMM");
  if (m_rows > 65535 || n_rows > 65535)
    return fail(KPM_ERR_ARG, "at most 65535 rows per operand");
  if (slices < 1 || slices > kMaxISlices) return fail(KPM_ERR_ARG, "slices must be 1..6");
  if (lda < k || ldb < k || ldc < n_rows) return fail(KPM_ERR_ARG, "leading dimension too small");
  const SlicedPlan p = islice_plan(m_rows, n_rows, k, slices);
  if (work_bytes < p.slice_bytes + p.exp_bytes + p.part_bytes + 256)
    return fail(KPM_ERR_ARG, "work buffer too small (kpm_corr_gemm_islice_work)");
  hipStream_t s = (hipStream_t)stream;
  char* w = static_cast<char*>(work);
  double* part = reinterpret_cast<double*>(w);
  int* ea = reinterpret_cast<int*>(w + p.part_bytes);
  int* eb = ea + m_rows;
  // the row pitch ldk is a multiple of 64, so both slice arrays start
  // 64-byte aligned
  int8_t* As = reinterpret_cast<int8_t*>(w + ((p.part_bytes + p.exp_bytes + 255) / 256) * 256);
  int8_t* Bs = As + int64_t(slices) * m_rows * p.ldk;
  const double* a = static_cast<const double*>(A);
  const double* b = static_cast<const double*>(B);
  int panel_no = 0;
  for (int64_t k0 = 0; k0 < k; k0 += p.panel, ++panel_no) {
    const int64_t kn = k - k0 < p.panel ? k - k0 : p.panel;
    const unsigned spans = unsigned((2 * kn + kExpSpan - 1) / kExpSpan);
    hipLaunchKernelGGL(exponent_clear, dim3((m_rows + 255) / 256), dim3(256), 0, s, ea, m_rows);
    hipLaunchKernelGGL(exponent_clear, dim3((n_rows + 255) / 256), dim3(256), 0, s, eb, n_rows);
    hipLaunchKernelGGL(row_exponent_kernel, dim3(spans, m_rows), dim3(256), 0, s, a, lda,
                       m_rows, k0, kn, ea);
    hipLaunchKernelGGL(row_exponent_kernel, dim3(spans, n_rows), dim3(256), 0, s, b, ldb,
                       n_rows, k0, kn, eb);
    hipLaunchKernelGGL(exponent_decode, dim3((m_rows + 255) / 256), dim3(256), 0, s, ea, m_rows);
    hipLaunchKernelGGL(exponent_decode, dim3((n_rows + 255) / 256), dim3(256), 0, s, eb, n_rows);
    const unsigned sg = unsigned((p.kp2 / 4 + 255) / 256);
    hipLaunchKernelGGL(islice_kernel, dim3(sg, m_rows), dim3(256), 0, s, a, lda, m_rows, k0, kn,
                       ea, int(slices), As, p.kp2, p.ldk, int64_t(m_rows));
    hipLaunchKernelGGL(islice_kernel, dim3(sg, n_rows), dim3(256), 0, s, b, ldb, n_rows, k0, kn,
                       eb, int(slices), Bs, p.kp2, p.ldk, int64_t(n_rows));
    const int slot0 = panel_no * int(p.zs);
    switch (slices) {
      case 1: launch_islice_gemm<1>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      case 2: launch_islice_gemm<2>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      case 3: launch_islice_gemm<3>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      case 4: launch_islice_gemm<4>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      case 5: launch_islice_gemm<5>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      default: launch_islice_gemm<6>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
    }
    int rc = check_hip(hipGetLastError(), "int8-sliced GEMM launch");
    if (rc) return rc;
  }
  // the (panel, split) slots in order: deterministic
  const int64_t total = int64_t(m_rows) * n_rows;
  hipLaunchKernelGGL(add_partials, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, part,
                     panel_no * int(p.zs), mu, ldc, m_rows, n_rows);
  return check_hip(hipGetLastError(), "int8-sliced GEMM partial sum");
}

// A prepared (pre-sliced) B operand: per panel, the rows' exponents (a
// 256-byte aligned block of n_rows int32) and then S x n_rows slice rows of
// ldk bytes.  The layout depends on (n_rows, k, slices) only.
static int64_t prep_panel_bytes(int n_rows, const SlicedPlan& p, int slices) {
  const int64_t exp_block = (int64_t(n_rows) * 4 + 255) / 256 * 256;
  return exp_block + int64_t(slices) * n_rows * p.ldk;
}

int kpm_islice_prep_bytes(int32_t n_rows, int64_t k, int32_t slices, int64_t* bytes) {
  if (!bytes || n_rows <= 0 || k <= 0 || slices < 1 || slices > kMaxISlices)
    return fail(KPM_ERR_ARG, "bad prepared-operand shape");
  if (n_rows > 65535) return fail(KPM_ERR_ARG, "at most 65535 rows per operand");
  const SlicedPlan p = islice_plan(1, n_rows, k, slices);
  const int64_t panels = (k + p.panel - 1) / p.panel;
  *bytes = panels * prep_panel_bytes(n_rows, p, slices);
  return KPM_OK;
}

int kpm_islice_prep_rows(const void* B, int64_t ldb, int32_t row0, int32_t rows, int32_t n_rows,
                         int64_t k, int32_t slices, void* prep, int64_t prep_bytes,
                         kpm_stream_t stream) {
  if (!B || !prep) return fail(KPM_ERR_ARG, "NULL buffer");
  if (n_rows <= 0 || n_rows > 65535 || k <= 0 || slices < 1 || slices > kMaxISlices)
    return fail(KPM_ERR_ARG, "bad prepared-operand shape");
  if (rows <= 0 || row0 < 0 || row0 + rows > n_rows)
    return fail(KPM_ERR_ARG, "rows outside the prepared operand");
  if (ldb < k) return fail(KPM_ERR_ARG, "leading dimension too small");
  const SlicedPlan p = islice_plan(1, n_rows, k, slices);
  const int64_t panels = (k + p.panel - 1) / p.panel;
  const int64_t pb = prep_panel_bytes(n_rows, p, slices);
  if (prep_bytes < panels * pb)
    return fail(KPM_ERR_ARG, "prepared operand too small (kpm_islice_prep_bytes)");
  hipStream_t s = (hipStream_t)stream;
  const double* b = static_cast<const double*>(B);
  const int64_t exp_block = (int64_t(n_rows) * 4 + 255) / 256 * 256;
  int panel_no = 0;
  for (int64_t k0 = 0; k0 < k; k0 += p.panel, ++panel_no) {
    const int64_t kn = k - k0 < p.panel ? k - k0 : p.panel;
    char* base = static_cast<char*>(prep) + panel_no * pb;
    int* e = reinterpret_cast<int*>(base) + row0;
    int8_t* out = reinterpret_cast<int8_t*>(base + exp_block) + int64_t(row0) * p.ldk;
    const unsigned spans = unsigned((2 * kn + kExpSpan - 1) / kExpSpan);
    hipLaunchKernelGGL(exponent_clear, dim3((rows + 255) / 256), dim3(256), 0, s, e, rows);
    hipLaunchKernelGGL(row_exponent_kernel, dim3(spans, rows), dim3(256), 0, s, b, ldb, rows, k0,
                       kn, e);
    hipLaunchKernelGGL(exponent_decode, dim3((rows + 255) / 256), dim3(256), 0, s, e, rows);
    const unsigned sg = unsigned((p.kp2 / 4 + 255) / 256);
    hipLaunchKernelGGL(islice_kernel, dim3(sg, rows), dim3(256), 0, s, b, ldb, rows, k0, kn, e,
                       int(slices), out, p.kp2, p.ldk, int64_t(n_rows));
    int rc = check_hip(hipGetLastError(), "operand slicing launch");
    if (rc) return rc;
  }
  return KPM_OK;
}

int kpm_corr_gemm_islice_prep_work(int32_t m_rows, int32_t n_rows, int64_t k, int32_t slices,
                                   int64_t* bytes) {
  if (!bytes || m_rows <= 0 || n_rows <= 0 || k <= 0 || slices < 1 || slices > kMaxISlices)
    return fail(KPM_ERR_ARG, "bad int8-sliced GEMM shape");
  if (m_rows > 65535 || n_rows > 65535)
    return fail(KPM_ERR_ARG, "at most 65535 rows per operand");
  const SlicedPlan p = islice_plan(m_rows, n_rows, k, slices);
  *bytes = p.part_bytes + int64_t(m_rows) * 4 + int64_t(slices) * m_rows * p.ldk + 256;
  return KPM_OK;
}

int kpm_corr_gemm_islice_prep(const void* A, int64_t lda, int32_t m_rows, const void* prep,
                              int64_t prep_bytes, int32_t n_rows, int64_t k, int32_t slices,
                              double* mu, int64_t ldc, void* work, int64_t work_bytes,
                              kpm_stream_t stream) {
  if (!A || !prep || !mu || !work) return fail(KPM_ERR_ARG, "NULL buffer");
  if (m_rows <= 0 || n_rows <= 0 || k <= 0) return fail(KPM_ERR_ARG, "empty GEMM");
  if (m_rows > 65535 || n_rows > 65535)
    return fail(KPM_ERR_ARG, "at most 65535 rows per operand");
  if (slices < 1 || slices > kMaxISlices) return fail(KPM_ERR_ARG, "slices must be 1..6");
  if (lda < k || ldc < n_rows) return fail(KPM_ERR_ARG, "leading dimension too small");
  const SlicedPlan p = islice_plan(m_rows, n_rows, k, slices);
  const int64_t panels = (k + p.panel - 1) / p.panel;
  const int64_t pb = prep_panel_bytes(n_rows, p, slices);
  if (prep_bytes < panels * pb)
    return fail(KPM_ERR_ARG, "prepared operand too small (kpm_islice_prep_bytes)");
  if (work_bytes < p.part_bytes + int64_t(m_rows) * 4 + int64_t(slices) * m_rows * p.ldk + 256)
    return fail(KPM_ERR_ARG, "work buffer too small (kpm_corr_gemm_islice_prep_work)");
  hipStream_t s = (hipStream_t)stream;
  char* w = static_cast<char*>(work);
  double* part = reinterpret_cast<double*>(w);
  int* ea = reinterpret_cast<int*>(w + p.part_bytes);
  int8_t* As = reinterpret_cast<int8_t*>(
      w + ((p.part_bytes + int64_t(m_rows) * 4 + 255) / 256) * 256);
  const double* a = static_cast<const double*>(A);
  const int64_t exp_block = (int64_t(n_rows) * 4 + 255) / 256 * 256;
  int panel_no = 0;
  for (int64_t k0 = 0; k0 < k; k0 += p.panel, ++panel_no) {
    const int64_t kn = k - k0 < p.panel ? k - k0 : p.panel;
    const char* base = static_cast<const char*>(prep) + panel_no * pb;
    const int* eb = reinterpret_cast<const int*>(base);
    const int8_t* Bs = reinterpret_cast<const int8_t*>(base + exp_block);
    const unsigned spans = unsigned((2 * kn + kExpSpan - 1) / kExpSpan);
    hipLaunchKernelGGL(exponent_clear, dim3((m_rows + 255) / 256), dim3(256), 0, s, ea, m_rows);
    hipLaunchKernelGGL(row_exponent_kernel, dim3(spans, m_rows), dim3(256), 0, s, a, lda,
                       m_rows, k0, kn, ea);
    hipLaunchKernelGGL(exponent_decode, dim3((m_rows + 255) / 256), dim3(256), 0, s, ea, m_rows);
    const unsigned sg = unsigned((p.kp2 / 4 + 255) / 256);
    hipLaunchKernelGGL(islice_kernel, dim3(sg, m_rows), dim3(256), 0, s, a, lda, m_rows, k0, kn,
                       ea, int(slices), As, p.kp2, p.ldk, int64_t(m_rows));
    const int slot0 = panel_no * int(p.zs);
    switch (slices) {
      case 1: launch_islice_gemm<1>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      case 2: launch_islice_gemm<2>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      case 3: launch_islice_gemm<3>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      case 4: launch_islice_gemm<4>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      case 5: launch_islice_gemm<5>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
      default: launch_islice_gemm<6>(As, Bs, m_rows, n_rows, p, ea, eb, part, slot0, s); break;
    }
    int rc = check_hip(hipGetLastError(), "int8-sliced GEMM launch");
    if (rc) return rc;
  }
  const int64_t total = int64_t(m_rows) * n_rows;
  hipLaunchKernelGGL(add_partials, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, part,
                     panel_no * int(p.zs), mu, ldc, m_rows, n_rows);
  return check_hip(hipGetLastError(), "int8-sliced GEMM partial sum");
}

}  // extern "C"
