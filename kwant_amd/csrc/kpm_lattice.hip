// Device-side generator of the cubic 4-orbital disorder ensemble: writes the
// CSR of each system straight into its diagonal block of one block-diagonal
// matrix (no host assembly, no upload), plus a Gershgorin radius per system.
//
// System of side L: sites s = (x L + y) L + z, orbital rows r = 4 s + o.
// H = A_s (on-site, 4 x 4 Hermitian) + t * I_4 to the 6 nearest neighbours
// (open boundaries).  A_s = (X + X^H) / 2 with X_ij = sigma (g_re + i g_im),
// g standard normals from a counter-based generator:
//   u(seed, c) = (SplitMix64(seed ^ SplitMix64(c)) >> 11) * 2^-53    in [0, 1)
//   g(seed, k) = sqrt(-2 ln(1 - u(seed, 2k))) cos(2 pi u(seed, 2k + 1))  (Box-Muller)
//   k = (16 s + 4 i + j) * 2 + {0: re, 1: im}.
// SplitMix64 is the mixing function of Steele, Lea and Flood (OOPSLA 2014).
// Row entries are written in increasing column order:
//   x-1, y-1, z-1 neighbours, the 4 on-site columns, z+1, y+1, x+1,
// so each row has 4 + deg(s) entries and the row offsets have a closed form
// (cubic4_entries_before).

#include "kpm_internal.h"

namespace kpm {

namespace {

__host__ __device__ inline uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline double unit_uniform(uint64_t seed, uint64_t c) {
  return double(splitmix(seed ^ splitmix(c)) >> 11) * (1.0 / 9007199254740992.0);
}

__device__ inline double std_normal(uint64_t seed, uint64_t k) {
  const double u1 = 1.0 - unit_uniform(seed, 2 * k);
  const double u2 = unit_uniform(seed, 2 * k + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Entries of all rows of sites before site s = (x L + y) L + z: every site
// holds 4 rows of (4 + deg) entries, and the neighbours that exist among the
// earlier sites are counted per direction in closed form.
__host__ __device__ inline int64_t cubic4_entries_before(int64_t L, int64_t x, int64_t y,
                                                         int64_t z) {
  const int64_t s = (x * L + y) * L + z;
  const int64_t cols = x * L + y;  // complete z-columns before s
  int64_t nb = 0;
  nb += cols * (L - 1) + (z > 0 ? z - 1 : 0);                        // z-1
  nb += cols * (L - 1) + z;                                          // z+1
  nb += x * (L - 1) * L + (y > 0 ? (y - 1) * L : 0) + (y > 0 ? z : 0);  // y-1
  nb += x * (L - 1) * L + y * L + (y < L - 1 ? z : 0);               // y+1
  nb += (x > 0 ? (x - 1) * L * L + y * L + z : 0);                   // x-1
  nb += x * L * L + (x < L - 1 ? y * L + z : 0);                     // x+1
  return 4 * (4 * s + nb);
}

// On-site entry (o, j) of site s: (X + X^H)_oj / 2 with X_ij = sigma (g_re + i g_im).
__device__ inline void cubic4_onsite(uint64_t seed, double sigma, int64_t s, int o, int j,
                                     double& ar, double& ai) {
  // no FMA contraction: every product is rounded on its own, so (o, j) and
  // (j, o) add and subtract the same two numbers and the block is Hermitian
  // bit for bit (A[j, o] = conj A[o, j], Im A[o, o] = +0), as the oracle's
#pragma clang fp contract(off)
  const uint64_t kij = uint64_t(16 * s + 4 * o + j) * 2;
  const uint64_t kji = uint64_t(16 * s + 4 * j + o) * 2;
  const double xr = sigma * std_normal(seed, kij), xi = sigma * std_normal(seed, kij + 1);
  const double yr = sigma * std_normal(seed, kji), yi = sigma * std_normal(seed, kji + 1);
  ar = 0.5 * (xr + yr);
  ai = 0.5 * (xi - yi);
}

// wavefront maximum of a non-negative row sum, then one atomic per wavefront;
// non-negative floats order like their bit patterns, rounded up so the
// radius is an upper bound
__device__ inline void radius_max(double rowsum, unsigned int* radius_bits) {
  float v = float(rowsum) * (1.0f + 1e-6f);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(radius_bits, __float_as_uint(v));
}

template <typename T>
__global__ __launch_bounds__(256) void cubic4_kernel(int L, uint64_t seed, double sigma,
                                                     double t, int64_t row0, int64_t nz0,
                                                     int32_t* __restrict__ indptr,
                                                     int32_t* __restrict__ indices,
                                                     T* __restrict__ values,
                                                     unsigned int* __restrict__ radius_bits) {
  const int64_t n = 4 * int64_t(L) * L * L;
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double rowsum = 0.0;
  if (r < n) {
    const int64_t s = r >> 2;
    const int o = int(r & 3);
    const int64_t z = s % L, y = (s / L) % L, x = s / (int64_t(L) * L);
    const int deg = (x > 0) + (x < L - 1) + (y > 0) + (y < L - 1) + (z > 0) + (z < L - 1);
    int64_t k = nz0 + cubic4_entries_before(L, x, y, z) + int64_t(o) * (4 + deg);
    indptr[row0 + r] = int32_t(k);
    if (r == n - 1) indptr[row0 + n] = int32_t(k + 4 + deg);
    auto emit = [&](int64_t col, double re, double im) {
      indices[k] = int32_t(row0 + col);
      values[2 * k] = T(re);
      values[2 * k + 1] = T(im);
      ++k;
    };
    const int64_t LL = int64_t(L) * L;
    if (x > 0) emit(4 * (s - LL) + o, t, 0.0);
    if (y > 0) emit(4 * (s - L) + o, t, 0.0);
    if (z > 0) emit(4 * (s - 1) + o, t, 0.0);
    for (int j = 0; j < 4; ++j) {
      double ar, ai;
      cubic4_onsite(seed, sigma, s, o, j, ar, ai);
      emit(4 * s + j, ar, ai);
      rowsum += sqrt(ar * ar + ai * ai);
    }
    if (z < L - 1) emit(4 * (s + 1) + o, t, 0.0);
    if (y < L - 1) emit(4 * (s + L) + o, t, 0.0);
    if (x < L - 1) emit(4 * (s + LL) + o, t, 0.0);
    rowsum += deg * fabs(t);
  }
  radius_max(rowsum, radius_bits);
}

// The same system as cubic4_kernel for the lattice-stencil recursion: only
// the on-site blocks, onsite[4 (row0 + r) + j] = A_s[o, j] for row r = 4 s + o
// (the hoppings t I_4 are implicit), and the same Gershgorin radius.
template <typename T>
__global__ __launch_bounds__(256) void cubic4_onsite_kernel(int L, uint64_t seed, double sigma,
                                                            double t, int64_t row0,
                                                            T* __restrict__ onsite,
                                                            unsigned int* __restrict__ radius_bits) {
  const int64_t n = 4 * int64_t(L) * L * L;
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double rowsum = 0.0;
  if (r < n) {
    const int64_t s = r >> 2;
    const int o = int(r & 3);
    const int64_t z = s % L, y = (s / L) % L, x = s / (int64_t(L) * L);
    const int deg = (x > 0) + (x < L - 1) + (y > 0) + (y < L - 1) + (z > 0) + (z < L - 1);
    for (int j = 0; j < 4; ++j) {
      double ar, ai;
      cubic4_onsite(seed, sigma, s, o, j, ar, ai);
      onsite[2 * ((row0 + r) * 4 + j)] = T(ar);
      onsite[2 * ((row0 + r) * 4 + j) + 1] = T(ai);
      rowsum += sqrt(ar * ar + ai * ai);
    }
    rowsum += deg * fabs(t);
  }
  radius_max(rowsum, radius_bits);
}

}  // namespace

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_cubic4_nnz(int32_t L, int64_t* nnz) {
  if (!nnz || L < 1) return fail(KPM_ERR_ARG, "bad cubic4 size");
  const int64_t l = L;
  *nnz = 40 * l * l * l - 24 * l * l;
  return KPM_OK;
}

int kpm_gen_cubic4(int32_t L, uint64_t seed, double sigma, double t, int64_t row0, int64_t nz0,
                   int32_t* indptr, int32_t* indices, void* values, int32_t value_type,
                   unsigned int* radius_bits, kpm_stream_t stream) {
  if (!indptr || !indices || !values || !radius_bits) return fail(KPM_ERR_ARG, "NULL buffer");
  if (L < 1 || row0 < 0 || nz0 < 0) return fail(KPM_ERR_ARG, "bad cubic4 block position");
  const int64_t n = 4 * int64_t(L) * L * L;
  const int64_t nnz = 40 * int64_t(L) * L * L - 24 * int64_t(L) * L;
  if (row0 + n >= (int64_t(1) << 31) || nz0 + nnz >= (int64_t(1) << 31))
    return fail(KPM_ERR_ARG, "ensemble too large for int32 CSR indices");
  const dim3 grid(unsigned((n + 255) / 256)), block(256);
  hipStream_t s = (hipStream_t)stream;
  if (value_type == KPM_C64)
    hipLaunchKernelGGL(cubic4_kernel<float>, grid, block, 0, s, L, seed, sigma, t, row0, nz0,
                       indptr, indices, static_cast<float*>(values), radius_bits);
  else if (value_type == KPM_C128)
    hipLaunchKernelGGL(cubic4_kernel<double>, grid, block, 0, s, L, seed, sigma, t, row0, nz0,
                       indptr, indices, static_cast<double*>(values), radius_bits);
  else
    return fail(KPM_ERR_ARG, "unknown value type");
  return check_hip(hipGetLastError(), "cubic4 generator launch");
}

int kpm_gen_cubic4_onsite(int32_t L, uint64_t seed, double sigma, double t, int64_t row0,
                          void* onsite, int32_t value_type, unsigned int* radius_bits,
                          kpm_stream_t stream) {
  if (!onsite || !radius_bits) return fail(KPM_ERR_ARG, "NULL buffer");
  if (L < 1 || row0 < 0) return fail(KPM_ERR_ARG, "bad cubic4 block position");
  const int64_t n = 4 * int64_t(L) * L * L;
  if (row0 + n >= (int64_t(1) << 31)) return fail(KPM_ERR_ARG, "ensemble too large");
  const dim3 grid(unsigned((n + 255) / 256)), block(256);
  hipStream_t s = (hipStream_t)stream;
  if (value_type == KPM_C64)
    hipLaunchKernelGGL(cubic4_onsite_kernel<float>, grid, block, 0, s, L, seed, sigma, t, row0,
                       static_cast<float*>(onsite), radius_bits);
  else if (value_type == KPM_C128)
    hipLaunchKernelGGL(cubic4_onsite_kernel<double>, grid, block, 0, s, L, seed, sigma, t, row0,
                       static_cast<double*>(onsite), radius_bits);
  else
    return fail(KPM_ERR_ARG, "unknown value type");
  return check_hip(hipGetLastError(), "cubic4 on-site generator launch");
}

}  // extern "C"
