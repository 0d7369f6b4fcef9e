// The Chebyshev step of cubic multi-orbital lattices without a stored matrix
// (gfx950): the batched ensemble of kpm_lattice.hip's cubic systems, side L,
// 4 orbitals per site, on-site 4 x 4 blocks A_s and hoppings t I_4 to the 6
// nearest neighbours (open boundaries).  Row r = 4 s + o of a system, site
// s = (x L + y) L + z.  The step is kwant/kpm.py:466 with the rescaled matvec
// of kwant/kpm.py:1182-1183 and the doubling dots of kwant/kpm.py:470-473,
// as cheb_kernel computes it from the CSR that kpm_gen_cubic4 writes -- with
// the same operations in the same order, entry by entry:
//   x-1, y-1, z-1 neighbours (t), the 4 on-site columns (A_s[o, j]),
//   z+1, y+1, x+1 neighbours (t),
// so the vectors and the dots are bitwise the CSR path's.
//
// A lane group owns one column (y, z, o) of a system and marches it along x:
// its rows are r(x) = row0 + x 4L^2 + c with c = 4 (y L + z) + o.  a_n of
// rows x-1, x and x+1 stays in the group's registers, so the x neighbours --
// 4L^2 rows away, the gathers that miss L2 in the CSR kernel -- are never
// gathered; the y and z neighbours and the site's other orbitals are nearby
// rows that other lane groups of the same or neighbouring workgroups read at
// the same x.  The matrix bytes are the on-site blocks (4 complex per row)
// instead of the CSR's 10 entries and row pointer.  The streams of a_n beyond
// the next plane, a_{n-1} and the on-site block are requested one or two
// planes ahead.

#include <vector>

#include "kpm_device.h"

namespace kpm {

struct MarchParams {
  char* ring;
  int64_t slot_bytes;
  int32_t slots;
  const int4* jobs;          // logical block -> (row0, L, first site, sites) in band order
  const int32_t* blk_seg;    // logical block -> system (segment)
  const int32_t* seg_blk;    // system -> first logical block (nseg + 1)
  const double* scale;       // system -> (a, b)
  const int32_t* im_exp;     // system x column grid exponents (doubling mode)
  const void* onsite;        // n_rows x 4 complex: A_s[o, j] of row 4 s + o
  double t;
  double* part;              // this step's partials
  const double* prev_part;   // the previous step's partials to reduce first (or null)
  double* prev_dots;
  int32_t nblocks;
  int32_t nblocks_launch;
  int32_t nseg;
  int32_t tile_y;            // rows along y of a band
};

// PD: planes the streams of a_n (beyond the x + 1 plane), a_{n-1} and the
// on-site values are requested ahead (1 or 2).  LX: the workgroup's columns
// put this plane's a_n in LDS (double-buffered, one barrier per plane), and
// the site's other orbitals and the y and z neighbours inside the workgroup's
// tile of sites are read from there instead of through the vector memory path.
// A workgroup's columns are the 4 orbitals of consecutive sites of one
// system's plane in band order (bands of tile_y rows, z-major inside a band),
// so a tile holds up to tile_y rows of y neighbours.
template <typename T, int VEC, int LPR, int WAVES, int MODE, int PD, bool LX>
__global__ __launch_bounds__(WAVES * 64) void march_kernel(MarchParams p, int n) {
  constexpr int W = LPR * VEC;
  constexpr int RPW = 64 / LPR;
  constexpr int NC = NComp<MODE>::value;
  constexpr int NQ = NQuant<MODE>::value;
  constexpr int NQA = NQ > 0 ? NQ : 1;
  constexpr bool dots = MODE == KPM_DOTS_DOUBLING;
  constexpr bool fg = dots && std::is_same<T, float>::value;
  static_assert(LPR >= 4, "four on-site values are shared within a lane group");
  using IO = LaneIO<T, VEC>;
  using CL = CLane<T, VEC>;

  const int hb = blockIdx.x;
  const int per = p.nblocks_launch / kNumXcd;
  const int lb = (hb % kNumXcd) * per + hb / kNumXcd;
  if (lb >= p.nblocks) return;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int grp = lane / LPR;
  const int sub = lane % LPR;
  const int col = sub * VEC;

  if constexpr (dots) {
    if (p.prev_part != nullptr)
      reduce_sums<MODE>(p.prev_part, p.nblocks, p.seg_blk, p.nseg, W, p.im_exp, p.prev_dots,
                        lb * WAVES + wave, p.nblocks * WAVES);
  }
  // the workgroup's tile: `cnt` consecutive sites from site K0 of a plane in
  // band order -- bands of tile_y rows along y (the last one may have fewer),
  // each band's sites z-major: K = band tile_y L + z rows + ty; slot = 4 (K -
  // K0) + o.  Lanes past the last site of a system take site K0 (every address
  // they form is valid) and store nothing.
  const int4 job = p.jobs[lb];
  const int seg = p.blk_seg[lb];
  const int64_t row0 = job.x;
  const int L = job.y;
  const int K0 = job.z, cnt = job.w;
  const int slot = wave * RPW + grp;
  const int o = slot & 3, kk = slot >> 2;
  const bool live = kk < cnt;
  const int TY = p.tile_y;
  const int K = K0 + (live ? kk : 0);
  const int band = K / (TY * L);
  const int kb = K - band * TY * L;
  const int rb = min(TY, L - band * TY);  // rows of this band
  const int z = kb / rb, ty = kb - z * rb;
  const int y = band * TY + ty;
  const int c = 4 * (y * L + z) + o;
  const int64_t P = 4 * int64_t(L) * L;  // rows from one x plane to the next
  const double sa = p.scale[2 * seg], sb = p.scale[2 * seg + 1];
  const T mul = T((n == 0 ? 1.0 : 2.0) / sa);
  const T bsh = T(sb);
  const T ht = T(p.t), hz = T(0);
  const T* cur = reinterpret_cast<const T*>(p.ring + int64_t(n % p.slots) * p.slot_bytes);
  T* nxt = reinterpret_cast<T*>(p.ring + int64_t((n + 1) % p.slots) * p.slot_bytes);
  const T* prv = reinterpret_cast<const T*>(
      p.ring + int64_t((n + p.slots - 1) % p.slots) * p.slot_bytes);
  const T* hv = reinterpret_cast<const T*>(p.onsite);

  double as[NC][VEC], ac[NC][VEC], aq[NQA][VEC];
  int qk[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
#pragma unroll
    for (int k = 0; k < NC; ++k) as[k][v] = ac[k][v] = 0.0;
#pragma unroll
    for (int k = 0; k < NQA; ++k) aq[k][v] = 0.0;
    qk[v] = NQ > 0 && p.im_exp != nullptr ? p.im_exp[seg * W + col + v] : 0;
  }
  CL wz;
  zero(wz);

  auto off = [&](int64_t row) { return 2 * (row * W + col); };
  // one FMA group of the row sum, as cheb_kernel applies a matrix entry (hr, hi)
  auto entry = [&](CL& acc, T hr, T hi, const CL& xe) {
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      acc.re[v] = fma(hr, xe.re[v], acc.re[v]);
      acc.re[v] = fma(-hi, xe.im[v], acc.re[v]);
      acc.im[v] = fma(hr, xe.im[v], acc.im[v]);
      acc.im[v] = fma(hi, xe.re[v], acc.im[v]);
    }
  };
  // lane `sub` holds on-site value sub mod 4 of the row; the group shares them
  auto onsite_part = [&](int64_t row, T& re, T& im) {
    load_value<T>(hv, row * 4 + (sub & 3), re, im);
  };

  __shared__ __attribute__((aligned(16))) T lds_x[LX ? 2 * WAVES * RPW * W * 2 : 2];
  const int64_t r0 = row0 + c;
  // with the LDS exchange, which in-plane neighbours are in the tile (the
  // same at every x): d sites away in band order
  auto in_tile = [&](int d) { return LX && unsigned(kk + d) < unsigned(cnt); };
  const bool zm_lds = z > 0 && in_tile(-rb), zp_lds = z < L - 1 && in_tile(rb);
  const bool ym_lds = ty > 0 && in_tile(-1), yp_lds = ty + 1 < rb && in_tile(1);
  // registers of the march: a_n of planes x-1 .. x+1+PD, a_{n-1} and the
  // on-site values of planes x .. x+PD-1 (slot k holds plane x + k)
  CL am, an[PD + 2], xq[PD];
  T hre[PD], him[PD];
  zero(am);
#pragma unroll
  for (int k = 0; k < PD + 1; ++k) {
    if (k < L) an[k] = IO::load(cur + off(r0 + k * P)); else zero(an[k]);
  }
  zero(an[PD + 1]);
#pragma unroll
  for (int k = 0; k < PD; ++k) {
    if (n > 0 && k < L) xq[k] = IO::load_nt(prv + off(r0 + k * P)); else zero(xq[k]);
    hre[k] = him[k] = T(0);
    if (k < L) onsite_part(r0 + k * P, hre[k], him[k]);
  }

  for (int x = 0; x < L; ++x) {
    const int64_t r = r0 + int64_t(x) * P;
    // ahead: a_n at plane x + 1 + PD, a_{n-1} and the on-site values at x + PD
    CL anew, xnew;
    T hrn = T(0), hin = T(0);
    if (x + 1 + PD < L) anew = IO::load(cur + off(r + (1 + PD) * P)); else zero(anew);
    if (x + PD < L) {
      if (n > 0) xnew = IO::load_nt(prv + off(r + PD * P)); else zero(xnew);
      onsite_part(r + PD * P, hrn, hin);
    } else {
      zero(xnew);
    }
    const CL& acur = an[0];
    const CL& ap = an[1];
    // this plane's neighbours: those read from memory are requested first,
    // so that their latency overlaps the LDS exchange
    CL gym, gzm, gzp, gyp, site[4];
    if (y > 0 && !ym_lds) gym = IO::load(cur + off(r - 4 * L)); else zero(gym);
    if (y < L - 1 && !yp_lds) gyp = IO::load(cur + off(r + 4 * L)); else zero(gyp);
    if (z > 0 && !zm_lds) gzm = IO::load(cur + off(r - 4)); else zero(gzm);
    if (z < L - 1 && !zp_lds) gzp = IO::load(cur + off(r + 4)); else zero(gzp);
    if constexpr (LX) {
      // every slot puts its a_n in LDS; the site's other orbitals and the
      // neighbours inside the tile come from there
      T* buf = lds_x + (x & 1) * (WAVES * RPW * W * 2);
      IO::store(buf + 2 * (slot * W + col), acur);
      __syncthreads();
      auto from = [&](int s) { return IO::load(buf + 2 * (s * W + col)); };
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j != o) site[j] = from(slot - o + j); else site[j] = acur;
      }
      if (zm_lds) gzm = from(slot - 4 * rb);
      if (zp_lds) gzp = from(slot + 4 * rb);
      if (ym_lds) gym = from(slot - 4);
      if (yp_lds) gyp = from(slot + 4);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j != o) site[j] = IO::load(cur + off(r - o + j)); else site[j] = acur;
      }
    }

    T er[4], ei[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      er[j] = __shfl(hre[0], j, LPR);
      ei[j] = __shfl(him[0], j, LPR);
    }
    CL acc;
    zero(acc);
    if (x > 0) entry(acc, ht, hz, am);
    if (y > 0) entry(acc, ht, hz, gym);
    if (z > 0) entry(acc, ht, hz, gzm);
#pragma unroll
    for (int j = 0; j < 4; ++j) entry(acc, er[j], ei[j], site[j]);
    if (z < L - 1) entry(acc, ht, hz, gzp);
    if (y < L - 1) entry(acc, ht, hz, gyp);
    if (x < L - 1) entry(acc, ht, hz, ap);
    CL yv;
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      yv.re[v] = fma(acc.re[v] - bsh * acur.re[v], mul, -xq[0].re[v]);
      yv.im[v] = fma(acc.im[v] - bsh * acur.im[v], mul, -xq[0].im[v]);
    }
    if (live) {
      IO::store_nt(nxt + off(r), yv);
      if constexpr (dots) add_dots<T, VEC, MODE, NC, NQA>(as, ac, aq, qk, fg, acur, yv, wz);
    }
    am = an[0];
#pragma unroll
    for (int k = 0; k < PD + 1; ++k) an[k] = an[k + 1];
    an[PD] = anew;
#pragma unroll
    for (int k = 0; k + 1 < PD; ++k) {
      xq[k] = xq[k + 1];
      hre[k] = hre[k + 1];
      him[k] = him[k + 1];
    }
    xq[PD - 1] = xnew;
    hre[PD - 1] = hrn;
    him[PD - 1] = hin;
  }
  if constexpr (!dots) return;
  finish_grid_dots<T, MODE>(fg, as, qk);
  __shared__ double red_s[WAVES * NC * W];
  __shared__ double red_c[WAVES * NC * W];
  __shared__ double red_q[WAVES * NQA * W];
  store_partials_to<VEC, LPR, WAVES, NC, NQ, NQA>(p.part, p.nblocks, W, 0, lb, wave, grp, col,
                                                   tid, as, ac, aq, red_s, red_c, red_q);
}

namespace {


template <typename T, int VEC, int LPR, int WAVES, int PD, bool LX>
int launch_march_w(int mode, int n, const MarchParams& p, hipStream_t s) {
  const dim3 grid(p.nblocks_launch), block(WAVES * 64);
  if (mode == KPM_DOTS_DOUBLING)
    hipLaunchKernelGGL((march_kernel<T, VEC, LPR, WAVES, KPM_DOTS_DOUBLING, PD, LX>), grid, block,
                       0, s, p, n);
  else
    hipLaunchKernelGGL((march_kernel<T, VEC, LPR, WAVES, KPM_DOTS_NONE, PD, LX>), grid, block, 0,
                       s, p, n);
  return check_hip(hipGetLastError(), "march_kernel launch");
}

// depth: 1 or 2 planes ahead; depth + 2 (3 or 4): the same with the LDS
// exchange of the plane
template <typename T, int VEC, int LPR>
int launch_march_v(int waves, int depth, int mode, int n, const MarchParams& p, hipStream_t s) {
  if (waves == 8) {
    switch (depth) {
      case 2: return launch_march_w<T, VEC, LPR, 8, 2, false>(mode, n, p, s);
      case 3: return launch_march_w<T, VEC, LPR, 8, 1, true>(mode, n, p, s);
      case 4: return launch_march_w<T, VEC, LPR, 8, 2, true>(mode, n, p, s);
      default: return launch_march_w<T, VEC, LPR, 8, 1, false>(mode, n, p, s);
    }
  }
  switch (depth) {
    case 2: return launch_march_w<T, VEC, LPR, 4, 2, false>(mode, n, p, s);
    case 3: return launch_march_w<T, VEC, LPR, 4, 1, true>(mode, n, p, s);
    case 4: return launch_march_w<T, VEC, LPR, 4, 2, true>(mode, n, p, s);
    default: return launch_march_w<T, VEC, LPR, 4, 1, false>(mode, n, p, s);
  }
}

int launch_march(int value_type, int width, int waves, int depth, int mode, int n,
                 const MarchParams& p, hipStream_t s) {
  if (value_type == KPM_C64) {
    switch (width) {
      case 8: return launch_march_v<float, 2, 4>(waves, depth, mode, n, p, s);
      case 16: return launch_march_v<float, 2, 8>(waves, depth, mode, n, p, s);
      case 32: return launch_march_v<float, 2, 16>(waves, depth, mode, n, p, s);
      case 64: return launch_march_v<float, 2, 32>(waves, depth, mode, n, p, s);
      default: break;
    }
  } else if (value_type == KPM_C128) {
    switch (width) {
      case 4: return launch_march_v<double, 1, 4>(waves, depth, mode, n, p, s);
      case 8: return launch_march_v<double, 1, 8>(waves, depth, mode, n, p, s);
      case 16: return launch_march_v<double, 1, 16>(waves, depth, mode, n, p, s);
      case 32: return launch_march_v<double, 1, 32>(waves, depth, mode, n, p, s);
      case 64: return launch_march_v<double, 1, 64>(waves, depth, mode, n, p, s);
      default: break;
    }
  }
  return fail(KPM_ERR_UNSUPPORTED, "no lattice-stencil kernel for this width and value type");
}

int lanes_per_row(int value_type, int width) {
  return value_type == KPM_C64 ? width / 2 : width;
}

}  // namespace

}  // namespace kpm

using namespace kpm;

struct kpm_march {
  int32_t n_rows = 0, width = 0, value_type = 0, nseg = 0, waves = 0, depth = 1, tile_y = 1;
  int32_t nblocks = 0, nblocks_launch = 0;
  int4* jobs = nullptr;
  int32_t* blk_seg = nullptr;
  int32_t* seg_blk = nullptr;
  double* part = nullptr;  // two buffers of part_half doubles (step parity)
  int64_t part_half = 0;
};

namespace {

void free_march(kpm_march* m) {
  if (!m) return;
  (void)hipFree(m->jobs);
  (void)hipFree(m->blk_seg);
  (void)hipFree(m->seg_blk);
  (void)hipFree(m->part);
  delete m;
}

template <typename T>
int march_alloc(T** p, size_t count, const char* what) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(KPM_ERR_ALLOC, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
  }
  return KPM_OK;
}

template <typename T>
int march_upload(T* d, const std::vector<T>& h, const char* what) {
  return check_hip(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), what);
}

}  // namespace

extern "C" {

int kpm_march_create(int32_t n_systems, const int32_t* h_sides, int32_t width,
                     int32_t value_type, int32_t waves, int32_t depth, int32_t tile_y,
                     kpm_march** out) {
  if (!out) return fail(KPM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (n_systems <= 0 || !h_sides) return fail(KPM_ERR_ARG, "need at least one system");
  if (value_type != KPM_C64 && value_type != KPM_C128)
    return fail(KPM_ERR_ARG, "value_type must be KPM_C64 or KPM_C128");
  if (waves != 4 && waves != 8) return fail(KPM_ERR_ARG, "waves must be 4 or 8");
  if (depth < 1 || depth > 4) return fail(KPM_ERR_ARG, "depth must be 1 to 4");
  const int lpr = lanes_per_row(value_type, width);
  if (lpr < 4 || lpr > 64 || (width & (width - 1)) != 0 ||
      (value_type == KPM_C64 && width > 64) || (value_type == KPM_C128 && width > 64))
    return fail(KPM_ERR_UNSUPPORTED, "no lattice-stencil kernel for this width and value type");
  const int S = waves * 64 / lpr / 4;  // sites per workgroup
  if (tile_y < 1 || tile_y > 1024) return fail(KPM_ERR_ARG, "tile_y must be 1 to 1024");
  std::vector<int4> jobs;
  std::vector<int32_t> blk_seg, seg_blk;
  int64_t row0 = 0;
  for (int s = 0; s < n_systems; ++s) {
    const int64_t L = h_sides[s];
    if (L < 1 || L > 1290) return fail(KPM_ERR_ARG, "system side out of range");
    if (row0 + 4 * L * L * L >= (int64_t(1) << 31))
      return fail(KPM_ERR_ARG, "ensemble too large for int32 rows");
    seg_blk.push_back(int32_t(jobs.size()));
    // S consecutive sites of the plane in band order per workgroup
    for (int64_t k0 = 0; k0 < L * L; k0 += S) {
      const int64_t cnt = L * L - k0 < S ? L * L - k0 : S;
      jobs.push_back(make_int4(int(row0), int(L), int(k0), int(cnt)));
      blk_seg.push_back(s);
    }
    row0 += 4 * L * L * L;
  }
  seg_blk.push_back(int32_t(jobs.size()));
  kpm_march* m = new kpm_march();
  m->n_rows = int32_t(row0);
  m->width = width;
  m->value_type = value_type;
  m->nseg = n_systems;
  m->waves = waves;
  m->depth = depth;
  m->tile_y = tile_y;
  m->nblocks = int32_t(jobs.size());
  m->nblocks_launch = (m->nblocks + kNumXcd - 1) / kNumXcd * kNumXcd;
  const int maxnc = 4;  // doubling: 2 TwoSum accumulators + 2 grid sums
  m->part_half = int64_t(m->nblocks) * maxnc * width * 2;
  int rc = march_alloc(&m->jobs, jobs.size(), "march jobs");
  if (!rc) rc = march_alloc(&m->blk_seg, blk_seg.size(), "march blk_seg");
  if (!rc) rc = march_alloc(&m->seg_blk, seg_blk.size(), "march seg_blk");
  if (!rc) rc = march_alloc(&m->part, size_t(2 * m->part_half), "march partials");
  if (!rc) rc = march_upload(m->jobs, jobs, "upload march jobs");
  if (!rc) rc = march_upload(m->blk_seg, blk_seg, "upload march blk_seg");
  if (!rc) rc = march_upload(m->seg_blk, seg_blk, "upload march seg_blk");
  if (!rc) rc = check_hip(hipDeviceSynchronize(), "march setup");
  if (rc) {
    free_march(m);
    return rc;
  }
  *out = m;
  return KPM_OK;
}

int kpm_march_destroy(kpm_march* m) {
  if (!m) return KPM_OK;
  hipError_t e = hipDeviceSynchronize();
  free_march(m);
  return check_hip(e, "march destroy");
}

int kpm_cheb_steps_march(kpm_plan* plan, kpm_march* m, const void* onsite, double t,
                         void* ring, int64_t slot_stride, int32_t slots, double* d_dots,
                         int32_t out_base, int32_t out_cap, int32_t n_steps, int32_t dots_mode,
                         kpm_stream_t stream) {
  if (!plan || !m) return fail(KPM_ERR_ARG, "plan or stencil is NULL");
  PlanState ps{};
  int rc = plan_state(plan, &ps);
  if (rc) return rc;
  if (ps.n_segments != m->nseg) return fail(KPM_ERR_ARG, "plan segments differ from systems");
  if (ps.n_rows != m->n_rows || ps.width != m->width || ps.value_type != m->value_type)
    return fail(KPM_ERR_ARG, "plan and stencil shapes differ");
  if (!ps.scale_set) return fail(KPM_ERR_ARG, "scale not set (kpm_plan_set_scale)");
  if (!ring || !onsite) return fail(KPM_ERR_ARG, "ring or on-site blocks NULL");
  if (slots < 2) return fail(KPM_ERR_ARG, "ring needs at least two slots");
  if (slot_stride < int64_t(ps.n_rows) * ps.width)
    return fail(KPM_ERR_ARG, "slot_stride smaller than n_rows * width");
  if (dots_mode != KPM_DOTS_DOUBLING && dots_mode != KPM_DOTS_NONE)
    return fail(KPM_ERR_ARG, "stencil steps take KPM_DOTS_DOUBLING or KPM_DOTS_NONE");
  const bool dots = dots_mode == KPM_DOTS_DOUBLING;
  if (dots && (!d_dots || out_cap <= 0)) return fail(KPM_ERR_ARG, "dots output buffer missing");
  if (dots && !ps.grid_set) return fail(KPM_ERR_ARG, "dots grid not set (kpm_plan_set_dots_grid)");
  if (n_steps < 0) return fail(KPM_ERR_ARG, "n_steps must be non-negative");
  if (dots && (ps.step - out_base < 0 || ps.step + n_steps - out_base > out_cap))
    return fail(KPM_ERR_ARG, "dots slots [step - out_base, + n_steps) outside [0, out_cap)");
  MarchParams p{};
  p.ring = reinterpret_cast<char*>(ring);
  p.slot_bytes = slot_stride * int64_t(m->value_type == KPM_C128 ? 16 : 8);
  p.slots = slots;
  p.jobs = m->jobs;
  p.blk_seg = m->blk_seg;
  p.seg_blk = m->seg_blk;
  p.scale = ps.scale;
  p.im_exp = ps.grid_set ? ps.im_exp : nullptr;
  p.onsite = onsite;
  p.t = t;
  p.nblocks = m->nblocks;
  p.nblocks_launch = m->nblocks_launch;
  p.nseg = m->nseg;
  p.tile_y = m->tile_y;
  const int64_t rec = int64_t(m->nseg) * m->width * 4;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n_steps; ++i) {
    const int n = ps.step + i;
    p.part = m->part + (n & 1) * m->part_half;
    p.prev_part = nullptr;
    p.prev_dots = nullptr;
    if (dots && i > 0) {
      p.prev_part = m->part + ((n - 1) & 1) * m->part_half;
      p.prev_dots = d_dots + int64_t(n - 1 - out_base) * rec;
    }
    rc = launch_march(m->value_type, m->width, m->waves, m->depth, dots_mode, n, p, s);
    if (rc) return rc;
    rc = plan_advance(plan, 1);
    if (rc) return rc;
  }
  if (dots && n_steps > 0) {
    const int last = ps.step + n_steps - 1;
    StepParams q{};
    q.part = m->part + (last & 1) * m->part_half;
    q.nblocks = m->nblocks;
    q.seg_blk = m->seg_blk;
    q.nseg = m->nseg;
    q.im_exp = p.im_exp;
    q.dots = d_dots + int64_t(last - out_base) * rec;
    rc = launch_reduce(dots_mode, m->width, q, s);
  }
  return rc;
}

}  // extern "C"
