// Two Chebyshev steps per launch over compact row tiles (gfx950).
//
// One launch computes a_{n+1} and a_{n+2} from a_{n-1} and a_n (n >= 1):
//   a_{n+1}[i] = 2 (sum_k H[i, j_k] a_n[j_k]     - b a_n[i])     / a - a_{n-1}[i]
//   a_{n+2}[i] = 2 (sum_k H[i, j_k] a_{n+1}[j_k] - b a_{n+1}[i]) / a - a_n[i]
// which is kwant/kpm.py:466 twice, with the rescaled matvec of
// kwant/kpm.py:1182-1183 fused into the epilogue, and the doubling dots of
// kwant/kpm.py:470-473 for both steps.
//
// The rows are cut into tiles of consecutive rows -- a compact patch of the
// lattice when the row order is a tiled site order.  A workgroup takes one
// tile and one group of WC columns of the vector block at a time:
//   phase A: a_{n+1} of the tile's rows and of its halo (the rows outside the
//            tile that the tile's entries reach) into LDS, gathering a_n from
//            HBM / L2; the tile's own rows are also written to their ring
//            slot, and their a_n is kept in LDS too;
//   phase B: a_{n+2} of the tile's rows, gathering a_{n+1} from LDS through
//            local entry indices (column j -> LDS slot of j).
// Per two steps the vectors cross HBM four times (read a_{n-1}, read a_n,
// write a_{n+1}, write a_{n+2}) plus the halo rows, against six times for two
// one-step launches.  A halo row's a_{n+1} is recomputed by every tile that
// needs it with its own tile's operations (entries in order, the same FMA
// sequence), so every row of a_{n+1} and a_{n+2} is bitwise the one-step
// kernel's, and the dot sums (order-independent: exact grid sums for
// complex64, TwoSum pairs rounded once for complex128) are bitwise too.
//
// Workgroups are persistent: a multiple of 8 x (column groups) of them, as
// many as fit on the device at once.  XCD x (hardware block b runs on XCD
// b mod 8) owns tiles [x T / 8, (x + 1) T / 8); within an XCD, workgroup k
// takes column group k mod CG and every (XCD workgroups / CG)-th tile from
// k / CG on, so the XCD sweeps its tiles as one front and the halo rows that
// neighbouring tiles share meet in its L2.  The dots of both steps stay in
// registers across a workgroup's tiles and are stored once, as workgroup
// partials that the next launch (or a reduction launch) sums in fixed order.

#include <algorithm>
#include <vector>

#include "kpm_device.h"

namespace kpm {

struct TiledParams {
  char* ring;
  int64_t slot_bytes;
  int32_t slots;
  int32_t width;               // columns of the vector block
  const int32_t* tile_begin;   // n_tiles + 1 row boundaries
  const int32_t* halo_begin;   // n_tiles + 1 offsets into halo_rows
  const int32_t* halo_rows;    // each tile's halo rows, ascending
  const double* scale;         // (a, b)
  const int32_t* im_exp;       // width grid exponents (doubling mode) or null
  double* part_a;              // this launch's partials of step n
  double* part_b;              // ... and of step n + 1
  const double* prev_a;        // the previous launch's partials (or null) ...
  const double* prev_b;
  double* prev_dots_a;         // ... and the records they are reduced into
  double* prev_dots_b;
  const int32_t* seg_blk;      // {0, nb}: one segment of nb partial blocks
  int32_t n_tiles;
  int32_t groups;              // column groups (width / WC)
  int32_t nb;                  // partial blocks per column group
  int32_t slot_cap;            // LDS slots of a_{n+1} (tile rows + halo)
  int32_t tile_cap;            // LDS rows of a_n (tile rows)
};

// Workgroup iterations over a tile's rows in phase B (a tile has at most
// kTileIters x rows-per-iteration rows), unrolled so that every entry of the
// phase is requested before the barrier that opens it.
constexpr int kTileIters = 4;

// The first PF entries of every row, packed so that a row's entries need no
// row-extent load first: col_pf[r * PF + e] is the column of entry e (-1: the
// row has fewer entries), and the last slot holds -2 - j when the row has
// more than PF entries (the rest follow from the CSR); lidx_pf the same with
// LDS slots (tile rows only); val_pf the values (0 where there is none).
template <typename T, int VEC, int LPR, int WAVES, int MODE, int PF>
__global__ __launch_bounds__(WAVES * 64) void tiled_kernel(TiledParams p,
                                                           const int32_t* __restrict__ indptr,
                                                           const int32_t* __restrict__ indices,
                                                           const void* __restrict__ values,
                                                           const int32_t* __restrict__ lidx,
                                                           const int32_t* __restrict__ col_pf,
                                                           const int32_t* __restrict__ lidx_pf,
                                                           const void* __restrict__ val_pf,
                                                           int n) {
  constexpr int WC = LPR * VEC;     // columns per work item
  constexpr int RPW = 64 / LPR;     // rows per wavefront
  constexpr int G = WAVES * RPW;    // rows per workgroup iteration
  constexpr int NB = kTileIters;
  constexpr int NC = NComp<MODE>::value;
  constexpr int NQ = NQuant<MODE>::value;
  constexpr int NQA = NQ > 0 ? NQ : 1;
  constexpr int NCH = (PF + LPR - 1) / LPR;
  constexpr bool dots = MODE == KPM_DOTS_DOUBLING;
  constexpr bool fg = dots && std::is_same<T, float>::value;
  using IO = LaneIO<T, VEC>;
  using CL = CLane<T, VEC>;
  extern __shared__ __attribute__((aligned(16))) char lds_raw[];
  // LDS: a_{n+1} slots (tile rows, then halo rows), a_n of the tile rows,
  // the halo rows' ids
  T* lds = reinterpret_cast<T*>(lds_raw);
  T* lds_x = lds + size_t(p.slot_cap) * WC * 2;
  int32_t* lds_h = reinterpret_cast<int32_t*>(lds_x + size_t(p.tile_cap) * WC * 2);

  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int grp = lane / LPR;
  const int sub = lane % LPR;
  const int gid = wave * RPW + grp;  // row group of the workgroup
  const int lcol = sub * VEC;        // lane's columns within the work item
  const int W = p.width;

  const int per = gridDim.x / kNumXcd;
  const int xcd = blockIdx.x % kNumXcd, kx = blockIdx.x / kNumXcd;
  const int cg = kx % p.groups, kk = kx / p.groups;
  const int fronts = per / p.groups;
  const int lb = xcd * fronts + kk;  // partial block within the column group
  const int col = cg * WC + lcol;

  if constexpr (dots) {
    if (p.prev_a != nullptr) {
      // the previous launch's two steps: their partials are complete
      const int first = blockIdx.x * WAVES + wave, stride = gridDim.x * WAVES;
      reduce_sums<MODE>(p.prev_a, p.nb, p.seg_blk, 1, W, p.im_exp, p.prev_dots_a, first, stride);
      reduce_sums<MODE>(p.prev_b, p.nb, p.seg_blk, 1, W, p.im_exp, p.prev_dots_b, first, stride);
    }
  }

  const double sa = p.scale[0], sb = p.scale[1];
  const T mul = T(2.0 / sa);
  const T bsh = T(sb);
  const T* am = reinterpret_cast<const T*>(p.ring + int64_t((n - 1) % p.slots) * p.slot_bytes);
  const T* a0 = reinterpret_cast<const T*>(p.ring + int64_t(n % p.slots) * p.slot_bytes);
  T* a1 = reinterpret_cast<T*>(p.ring + int64_t((n + 1) % p.slots) * p.slot_bytes);
  T* a2 = reinterpret_cast<T*>(p.ring + int64_t((n + 2) % p.slots) * p.slot_bytes);

  double asA[NC][VEC], acA[NC][VEC], aqA[NQA][VEC];
  double asB[NC][VEC], acB[NC][VEC], aqB[NQA][VEC];
  int qk[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
#pragma unroll
    for (int c = 0; c < NC; ++c) asA[c][v] = acA[c][v] = asB[c][v] = acB[c][v] = 0.0;
#pragma unroll
    for (int c = 0; c < NQA; ++c) aqA[c][v] = aqB[c][v] = 0.0;
    qk[v] = NQ > 0 && p.im_exp != nullptr ? p.im_exp[col + v] : 0;
  }
  CL wz;
  zero(wz);

  // The packed entries of a row, cooperatively: lane `sub` of the row's
  // group takes entries sub, sub + LPR, ... of the first PF, and the group
  // shares them by shuffles.  ix_pf: packed columns (phase A) or LDS slots
  // (phase B).
  auto head = [&](const int32_t* __restrict__ ix_pf, int row, bool live, int (&j)[NCH],
                  T (&hr)[NCH], T (&hi)[NCH]) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int e = c * LPR + sub;
      j[c] = -1;
      hr[c] = T(0);
      hi[c] = T(0);
      if (live && e < PF) {
        const int64_t k = int64_t(row) * PF + e;
        j[c] = ix_pf[k];
        load_value<T>(val_pf, k, hr[c], hi[c]);
      }
    }
  };
  // sum_k H[row, j_k] x[j_k, cols] with x rows at src + j * ld (complex
  // elements): the first PF entries from the packed head, the rest (rows
  // flagged longer) from the CSR in chunks of LPR, in order -- the FMA
  // sequence of the one-step kernels.  ix: the CSR's column index array
  // (global columns in phase A, LDS slots in phase B).
  auto row_sum = [&](const T* src, int ld, const int32_t* __restrict__ ix, int row,
                     const int (&j)[NCH], const T (&hr)[NCH], const T (&hi)[NCH]) {
    CL acc;
    zero(acc);
    CL x[PF];
    T er[PF], ei[PF];
    bool has[PF];
    int last = -1;
#pragma unroll
    for (int e = 0; e < PF; ++e) {
      const int jj = __shfl(j[e / LPR], e % LPR, LPR);
      er[e] = __shfl(hr[e / LPR], e % LPR, LPR);
      ei[e] = __shfl(hi[e / LPR], e % LPR, LPR);
      has[e] = jj != -1;
      const int jr = jj >= 0 ? jj : -2 - jj;
      if (has[e])
        x[e] = IO::load(src + 2 * (int64_t(jr) * ld + lcol));
      else
        zero(x[e]);
      if (e == PF - 1) last = jj;
    }
#pragma unroll
    for (int e = 0; e < PF; ++e) {
      if (has[e]) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          acc.re[v] = fma(er[e], x[e].re[v], acc.re[v]);
          acc.re[v] = fma(-ei[e], x[e].im[v], acc.re[v]);
          acc.im[v] = fma(er[e], x[e].im[v], acc.im[v]);
          acc.im[v] = fma(ei[e], x[e].re[v], acc.im[v]);
        }
      }
    }
    if (last <= -2) {
      // a row longer than PF: the whole group takes this loop together
      const int k1 = indptr[row + 1];
      for (int kt = indptr[row] + PF; kt < k1; kt += LPR) {
        int jt = 0;
        T htr = T(0), hti = T(0);
        if (kt + sub < k1) {
          jt = ix[kt + sub];
          load_value<T>(values, kt + sub, htr, hti);
        }
#pragma unroll
        for (int s = 0; s < LPR; ++s) {
          const int jj = __shfl(jt, s, LPR);
          const T tr = __shfl(htr, s, LPR);
          const T ti = __shfl(hti, s, LPR);
          if (kt + s < k1) {
            const CL xe = IO::load(src + 2 * (int64_t(jj) * ld + lcol));
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
              acc.re[v] = fma(tr, xe.re[v], acc.re[v]);
              acc.re[v] = fma(-ti, xe.im[v], acc.re[v]);
              acc.im[v] = fma(tr, xe.im[v], acc.im[v]);
              acc.im[v] = fma(ti, xe.re[v], acc.im[v]);
            }
          }
        }
      }
    }
    return acc;
  };

  const int t_lo = int((int64_t(xcd) * p.n_tiles) / kNumXcd);
  const int t_hi = int((int64_t(xcd + 1) * p.n_tiles) / kNumXcd);
  int t = t_lo + kk;
  int r0 = 0, R = 0, h0 = 0, H = 0;
  // phase A's pipeline registers: iteration it's row and packed entries
  int rA = 0;
  bool lA = false;
  int jA[NCH];
  T hrA[NCH], hiA[NCH];
  auto open_tile = [&](int tt) {
    // a tile's extent, and the entries of its first phase-A iteration
    r0 = p.tile_begin[tt];
    R = p.tile_begin[tt + 1] - r0;
    h0 = p.halo_begin[tt];
    H = p.halo_begin[tt + 1] - h0;
    lA = gid < R;
    rA = r0 + (lA ? gid : 0);
    head(col_pf, rA, lA, jA, hrA, hiA);
  };
  if (t < t_hi) open_tile(t);
  while (t < t_hi) {
    for (int s = tid; s < H; s += WAVES * 64) lds_h[s] = p.halo_rows[h0 + s];
    __syncthreads();  // halo rows visible; the previous tile's phase B is done

    // ---- phase A: a_{n+1} of the tile's rows (iterations 0 .. NI-1) and of
    // its halo (iterations NI ..): while iteration it gathers, the entries of
    // it + 1 are in flight.
    const int NI = (R + G - 1) / G;
    const int NA = NI + (H + G - 1) / G;
#pragma unroll 1
    for (int it = 0; it < NA; ++it) {
      const int nx = it + 1;
      int rB = r0;
      bool lB = false;
      if (nx < NI) {
        const int s = gid + G * nx;
        lB = s < R;
        rB = r0 + (lB ? s : 0);
      } else if (nx < NA) {
        const int s = gid + G * (nx - NI);
        lB = s < H;
        rB = lB ? lds_h[s] : r0;
      }
      int jB[NCH];
      T hrB[NCH], hiB[NCH];
      head(col_pf, rB, lB, jB, hrB, hiB);
      const bool inner = it < NI;
      const int64_t off = 2 * (int64_t(rA) * W + col);
      CL xs = IO::load(a0 + off);
      CL xp = inner ? IO::load_nt(am + off) : IO::load(am + off);
      const CL acc = row_sum(a0 + 2 * cg * WC, W, indices, rA, jA, hrA, hiA);
      CL y;
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        y.re[v] = fma(acc.re[v] - bsh * xs.re[v], mul, -xp.re[v]);
        y.im[v] = fma(acc.im[v] - bsh * xs.im[v], mul, -xp.im[v]);
      }
      const int s = inner ? gid + G * it : R + gid + G * (it - NI);
      if (lA) IO::store(lds + 2 * (int64_t(s) * WC + lcol), y);
      if (inner) {
        if (lA) {
          IO::store_nt(a1 + off, y);
          IO::store(lds_x + 2 * (int64_t(s) * WC + lcol), xs);
        } else {
          zero(xs);
          zero(y);
        }
        if constexpr (dots) add_dots<T, VEC, MODE, NC, NQA>(asA, acA, aqA, qk, fg, xs, y, wz);
      }
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        jA[c] = jB[c];
        hrA[c] = hrB[c];
        hiA[c] = hiB[c];
      }
      rA = rB;
      lA = lB;
    }
    // every entry of phase B, requested before the barrier
    int jb[NB][NCH];
    T hrb[NB][NCH], hib[NB][NCH];
#pragma unroll
    for (int it = 0; it < NB; ++it) {
      const int s = gid + G * it;
      head(lidx_pf, r0 + (s < R ? s : 0), s < R, jb[it], hrb[it], hib[it]);
    }
    __syncthreads();  // a_{n+1} of the tile and its halo, a_n of the tile in LDS

    // ---- phase B: a_{n+2} of the tile's rows from LDS (local slots); the
    // next tile's first entries are requested meanwhile
    const int rb0 = r0, RB = R;
    t += fronts;
    if (t < t_hi) open_tile(t);
#pragma unroll
    for (int it = 0; it < NB; ++it) {
      if (G * it >= RB) break;
      const int s0 = gid + G * it;
      const bool live = s0 < RB;
      const int s = live ? s0 : 0;
      const CL acc = row_sum(lds, WC, lidx, rb0 + s, jb[it], hrb[it], hib[it]);
      CL y1 = IO::load(lds + 2 * (int64_t(s) * WC + lcol));
      const CL xn = IO::load(lds_x + 2 * (int64_t(s) * WC + lcol));
      CL y2;
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        y2.re[v] = fma(acc.re[v] - bsh * y1.re[v], mul, -xn.re[v]);
        y2.im[v] = fma(acc.im[v] - bsh * y1.im[v], mul, -xn.im[v]);
      }
      if (live) {
        IO::store_nt(a2 + 2 * (int64_t(rb0 + s) * W + col), y2);
      } else {
        zero(y1);
        zero(y2);
      }
      if constexpr (dots) add_dots<T, VEC, MODE, NC, NQA>(asB, acB, aqB, qk, fg, y1, y2, wz);
    }
  }
  if constexpr (!dots) return;

  finish_grid_dots<T, MODE>(fg, asA, qk);
  finish_grid_dots<T, MODE>(fg, asB, qk);
  // the workgroup reduction reuses the front of the dynamic LDS (the host
  // sizes it for both): every wave is past its last phase B first
  double* red_s = reinterpret_cast<double*>(lds_raw);
  double* red_c = red_s + WAVES * NC * WC;
  double* red_q = red_c + WAVES * NC * WC;
  __syncthreads();
  store_partials_to<VEC, LPR, WAVES, NC, NQ, NQA>(p.part_a, p.nb, W, cg * WC, lb, wave, grp, lcol,
                                                  tid, asA, acA, aqA, red_s, red_c, red_q);
  __syncthreads();  // the reduction arrays are reused
  store_partials_to<VEC, LPR, WAVES, NC, NQ, NQA>(p.part_b, p.nb, W, cg * WC, lb, wave, grp, lcol,
                                                  tid, asB, acB, aqB, red_s, red_c, red_q);
}

// The packed values: val_pf[r * PF + e] = values[indptr[r] + e] for e below
// the row's length, 0 past it.
template <typename T>
__global__ __launch_bounds__(256) void pack_values_kernel(const int32_t* __restrict__ indptr,
                                                          const void* __restrict__ values,
                                                          void* __restrict__ val_pf, int n_rows,
                                                          int pf) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= int64_t(n_rows) * pf) return;
  const int r = int(i / pf), e = int(i - int64_t(r) * pf);
  const int k0 = indptr[r], k1 = indptr[r + 1];
  T re = T(0), im = T(0);
  if (k0 + e < k1) load_value<T>(values, k0 + e, re, im);
  T* out = reinterpret_cast<T*>(val_pf);
  out[2 * i] = re;
  out[2 * i + 1] = im;
}

// ----------------------------------------------------------------------------
// Launch shapes and dispatch.

namespace {

struct TiledShape {
  int lpr;        // lanes per row
  int vec;        // complex values per lane (group width = lpr * vec)
  int waves;      // wavefronts per workgroup
  int tile_rows;  // the largest tile: kTileIters x rows per workgroup
                  // iteration; with a quarter as many halo rows it fits the
                  // LDS share of a workgroup at full occupancy (one 16-wave
                  // or two 8-wave workgroups per CU)
};

// complex64: 8, 16 or 32 lanes x 2 columns (groups of 16, 32 or 64 columns:
// 128 to 512 bytes per row); complex128: 8, 16 or 32 lanes x 1.
constexpr TiledShape kC64Shapes[] = {{8, 2, 16, 512},  {8, 2, 8, 256},  {16, 2, 16, 256},
                                     {16, 2, 8, 128},  {32, 2, 16, 128}, {32, 2, 8, 64}};
constexpr TiledShape kC128Shapes[] = {{8, 1, 16, 512}, {16, 1, 16, 256}, {32, 1, 16, 128}};
constexpr int kTiledPrefetch[] = {4, 8};

int tiled_shapes(int width, int value_type, TiledShape* out, int max_out) {
  if (value_type != KPM_C64 && value_type != KPM_C128) return 0;
  const TiledShape* table = value_type == KPM_C64 ? kC64Shapes : kC128Shapes;
  const int count = value_type == KPM_C64 ? int(sizeof(kC64Shapes) / sizeof(TiledShape))
                                          : int(sizeof(kC128Shapes) / sizeof(TiledShape));
  int n = 0;
  for (int i = 0; i < count; ++i) {
    const int wc = table[i].lpr * table[i].vec;
    if (width % wc != 0) continue;
    if (n < max_out) out[n] = table[i];
    ++n;
  }
  return n < max_out ? n : max_out;
}

template <typename T, int VEC, int LPR, int WAVES, int MODE, int PF>
const void* kernel_ptr() {
  return reinterpret_cast<const void*>(&tiled_kernel<T, VEC, LPR, WAVES, MODE, PF>);
}

// The kernel of (shape, mode, prefetch depth) for a value type, or null.
template <typename T, int VEC, int LPR, int WAVES>
const void* pick_mode(int mode, int pf) {
  if (mode == KPM_DOTS_DOUBLING)
    return pf == 4 ? kernel_ptr<T, VEC, LPR, WAVES, KPM_DOTS_DOUBLING, 4>()
                   : kernel_ptr<T, VEC, LPR, WAVES, KPM_DOTS_DOUBLING, 8>();
  if (mode == KPM_DOTS_NONE)
    return pf == 4 ? kernel_ptr<T, VEC, LPR, WAVES, KPM_DOTS_NONE, 4>()
                   : kernel_ptr<T, VEC, LPR, WAVES, KPM_DOTS_NONE, 8>();
  return nullptr;
}

const void* tiled_kernel_for(int value_type, const TiledShape& s, int mode, int pf) {
  if (value_type == KPM_C64) {
    if (s.lpr == 8 && s.waves == 16) return pick_mode<float, 2, 8, 16>(mode, pf);
    if (s.lpr == 8 && s.waves == 8) return pick_mode<float, 2, 8, 8>(mode, pf);
    if (s.lpr == 16 && s.waves == 16) return pick_mode<float, 2, 16, 16>(mode, pf);
    if (s.lpr == 16 && s.waves == 8) return pick_mode<float, 2, 16, 8>(mode, pf);
    if (s.lpr == 32 && s.waves == 16) return pick_mode<float, 2, 32, 16>(mode, pf);
    if (s.lpr == 32 && s.waves == 8) return pick_mode<float, 2, 32, 8>(mode, pf);
  } else if (value_type == KPM_C128) {
    if (s.lpr == 8 && s.waves == 16) return pick_mode<double, 1, 8, 16>(mode, pf);
    if (s.lpr == 16 && s.waves == 16) return pick_mode<double, 1, 16, 16>(mode, pf);
    if (s.lpr == 32 && s.waves == 16) return pick_mode<double, 1, 32, 16>(mode, pf);
  }
  return nullptr;
}

int tiled_prefetch(int32_t row_nnz_hint) {
  for (int pf : kTiledPrefetch)
    if (row_nnz_hint > 0 && row_nnz_hint <= pf) return pf;
  return kTiledPrefetch[sizeof(kTiledPrefetch) / sizeof(int) - 1];
}

constexpr int kTiledMaxShapes = 8;

}  // namespace

}  // namespace kpm

using namespace kpm;

struct kpm_tiled {
  int32_t n_rows = 0, width = 0, value_type = 0, n_tiles = 0;
  int64_t nnz = 0;
  const void* values = nullptr;  // the matrix whose entries were packed
  TiledShape shape{};
  int32_t pf = 0;          // packed entries per row
  int32_t groups = 0;      // column groups
  int32_t slot_cap = 0;    // max tile rows + halo rows
  int32_t halo_cap = 0;    // max halo rows
  int32_t max_tile = 0;    // max tile rows
  int64_t halo_total = 0;  // halo rows over all tiles
  int64_t halo_entries = 0;  // matrix entries of those rows
  int64_t gather_rows = 0;   // rows of a_n the tiles gather (tile, halo, halo's reach)
  size_t lds_bytes = 0;
  int32_t nwg = 0, nb = 0; // persistent workgroups; partial blocks per column group
  int32_t* tile_begin = nullptr;
  int32_t* halo_begin = nullptr;
  int32_t* halo_rows = nullptr;
  int32_t* lidx = nullptr;     // nnz: LDS slot of every entry of its row's tile
  int32_t* col_pf = nullptr;   // n_rows x pf packed columns
  int32_t* lidx_pf = nullptr;  // n_rows x pf packed LDS slots
  void* val_pf = nullptr;      // n_rows x pf packed values
  int32_t* seg_blk = nullptr;
  double* part = nullptr;  // four buffers of part_q doubles: (step parity in the launch) x (launch parity)
  int64_t part_q = 0;
};

namespace {

template <typename T>
int tdev_alloc(T** p, size_t count, const char* what) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(KPM_ERR_ALLOC, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
  }
  return KPM_OK;
}

template <typename T>
int tupload(T* d, const std::vector<T>& h, const char* what) {
  if (h.empty()) return KPM_OK;
  return check_hip(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), what);
}

void free_tiled(kpm_tiled* t) {
  if (!t) return;
  (void)hipFree(t->tile_begin);
  (void)hipFree(t->halo_begin);
  (void)hipFree(t->halo_rows);
  (void)hipFree(t->lidx);
  (void)hipFree(t->col_pf);
  (void)hipFree(t->lidx_pf);
  (void)hipFree(t->val_pf);
  (void)hipFree(t->seg_blk);
  (void)hipFree(t->part);
  delete t;
}

// The first pf entries of each row of (indptr, ix) with the continuation
// flag: -1 past the row's end, -2 - j in the last slot when more follow.
std::vector<int32_t> pack_entries(const std::vector<int32_t>& indptr,
                                  const std::vector<int32_t>& ix, int32_t n_rows, int pf) {
  std::vector<int32_t> out(size_t(n_rows) * pf, -1);
  for (int32_t r = 0; r < n_rows; ++r) {
    const int64_t k0 = indptr[r], len = int64_t(indptr[r + 1]) - k0;
    for (int e = 0; e < pf && e < len; ++e) out[size_t(r) * pf + e] = ix[size_t(k0 + e)];
    if (len > pf) out[size_t(r) * pf + pf - 1] = -2 - out[size_t(r) * pf + pf - 1];
  }
  return out;
}

}  // namespace

extern "C" {

int kpm_tiled_variant_count(int32_t width, int32_t value_type, int32_t* count) {
  if (!count) return fail(KPM_ERR_ARG, "count is NULL");
  TiledShape s[kTiledMaxShapes];
  *count = tiled_shapes(width, value_type, s, kTiledMaxShapes);
  return KPM_OK;
}

int kpm_tiled_variant_info(int32_t width, int32_t value_type, int32_t variant,
                           int32_t* group_width, int32_t* waves, int32_t* max_tile_rows) {
  TiledShape s[kTiledMaxShapes];
  const int nv = tiled_shapes(width, value_type, s, kTiledMaxShapes);
  if (variant < 0 || variant >= nv) return fail(KPM_ERR_ARG, "tiled variant out of range");
  if (group_width) *group_width = s[variant].lpr * s[variant].vec;
  if (waves) *waves = s[variant].waves;
  if (max_tile_rows) *max_tile_rows = s[variant].tile_rows;
  return KPM_OK;
}

int kpm_tiled_create(const kpm_csr* h, int32_t width, int32_t n_tiles,
                     const int32_t* h_tile_row_begin, int32_t variant, kpm_tiled** out) {
  if (!out) return fail(KPM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!h || !h_tile_row_begin) return fail(KPM_ERR_ARG, "NULL argument");
  const int32_t n_rows = h->n_rows, value_type = h->value_type;
  const int64_t nnz = h->nnz;
  if (n_rows <= 0 || h->n_cols != n_rows) return fail(KPM_ERR_ARG, "need a square matrix");
  if (n_tiles <= 0) return fail(KPM_ERR_ARG, "need at least one tile");
  if (nnz < 0 || nnz > INT32_MAX) return fail(KPM_ERR_ARG, "nnz out of int32 range");
  if (!h->indptr || (nnz > 0 && (!h->indices || !h->values)))
    return fail(KPM_ERR_ARG, "matrix arrays are NULL");
  TiledShape shapes[kTiledMaxShapes];
  const int nv = tiled_shapes(width, value_type, shapes, kTiledMaxShapes);
  if (nv == 0)
    return fail(KPM_ERR_UNSUPPORTED, "no tiled kernel for this width and value type");
  if (variant < 0) variant = 0;
  if (variant >= nv) return fail(KPM_ERR_ARG, "tiled variant out of range");
  const TiledShape sh = shapes[variant];
  const int wc = sh.lpr * sh.vec;
  if (sh.tile_rows != kTileIters * sh.waves * 64 / sh.lpr)
    return fail(KPM_ERR_UNSUPPORTED, "tiled shape table inconsistent with the kernel");
  if (h_tile_row_begin[0] != 0 || h_tile_row_begin[n_tiles] != n_rows)
    return fail(KPM_ERR_ARG, "tiles must cover rows [0, n_rows)");
  int32_t max_tile = 0;
  for (int32_t t = 0; t < n_tiles; ++t) {
    const int32_t r = h_tile_row_begin[t + 1] - h_tile_row_begin[t];
    if (r <= 0) return fail(KPM_ERR_ARG, "every tile needs at least one row");
    if (r > sh.tile_rows)
      return fail(KPM_ERR_ARG, "a tile has " + std::to_string(r) + " rows; this variant takes " +
                                   std::to_string(sh.tile_rows));
    if (r > max_tile) max_tile = r;
  }
  // the pattern, read back from the device
  std::vector<int32_t> indptr(static_cast<size_t>(n_rows) + 1);
  std::vector<int32_t> indices(static_cast<size_t>(nnz));
  int rc = check_hip(hipMemcpy(indptr.data(), h->indptr, indptr.size() * 4, hipMemcpyDeviceToHost),
                     "copy indptr");
  if (!rc && nnz > 0)
    rc = check_hip(hipMemcpy(indices.data(), h->indices, indices.size() * 4,
                             hipMemcpyDeviceToHost),
                   "copy indices");
  if (rc) return rc;
  if (indptr[0] != 0 || int64_t(indptr[n_rows]) != nnz)
    return fail(KPM_ERR_ARG, "indptr must run from 0 to nnz");
  for (int32_t r = 0; r < n_rows; ++r)
    if (indptr[r + 1] < indptr[r]) return fail(KPM_ERR_ARG, "indptr must not decrease");
  // Halo of each tile (the columns its entries reach outside it, ascending)
  // and the local index of every entry: its column's LDS slot.
  std::vector<int32_t> halo_begin(size_t(n_tiles) + 1, 0), halo_rows;
  std::vector<int32_t> lidx(static_cast<size_t>(nnz));
  std::vector<int32_t> tile_begin(h_tile_row_begin, h_tile_row_begin + n_tiles + 1);
  std::vector<int32_t> outside, reach;
  int32_t halo_cap = 0, slot_cap = 0;
  int64_t halo_entries = 0, gather_rows = 0;
  for (int32_t t = 0; t < n_tiles; ++t) {
    const int32_t r0 = tile_begin[t], r1 = tile_begin[t + 1];
    outside.clear();
    for (int64_t k = indptr[r0]; k < indptr[r1]; ++k) {
      const int32_t j = indices[k];
      if (j < 0 || j >= n_rows) return fail(KPM_ERR_ARG, "column index out of range");
      if (j < r0 || j >= r1) outside.push_back(j);
    }
    std::sort(outside.begin(), outside.end());
    outside.erase(std::unique(outside.begin(), outside.end()), outside.end());
    const int32_t hcount = int32_t(outside.size());
    halo_rows.insert(halo_rows.end(), outside.begin(), outside.end());
    halo_begin[size_t(t) + 1] = int32_t(halo_rows.size());
    for (int64_t k = indptr[r0]; k < indptr[r1]; ++k) {
      const int32_t j = indices[k];
      if (j >= r0 && j < r1) {
        lidx[size_t(k)] = j - r0;
      } else {
        const auto it = std::lower_bound(outside.begin(), outside.end(), j);
        lidx[size_t(k)] = (r1 - r0) + int32_t(it - outside.begin());
      }
    }
    if (hcount > halo_cap) halo_cap = hcount;
    if (r1 - r0 + hcount > slot_cap) slot_cap = r1 - r0 + hcount;
    // what phase A gathers beyond the tile and its halo: the halo rows' reach
    reach.clear();
    for (const int32_t hr : outside) {
      halo_entries += indptr[hr + 1] - indptr[hr];
      for (int64_t k = indptr[hr]; k < indptr[hr + 1]; ++k) {
        const int32_t j = indices[k];
        if (j < 0 || j >= n_rows) return fail(KPM_ERR_ARG, "column index out of range");
        if ((j < r0 || j >= r1) && !std::binary_search(outside.begin(), outside.end(), j))
          reach.push_back(j);
      }
    }
    std::sort(reach.begin(), reach.end());
    reach.erase(std::unique(reach.begin(), reach.end()), reach.end());
    gather_rows += int64_t(r1 - r0) + hcount + int64_t(reach.size());
  }
  const int pf = tiled_prefetch(h->row_nnz_hint);
  const std::vector<int32_t> col_pf = pack_entries(indptr, indices, n_rows, pf);
  const std::vector<int32_t> lidx_pf = pack_entries(indptr, lidx, n_rows, pf);
  const size_t elem = value_type == KPM_C128 ? 16 : 8;
  // a_{n+1} slots, a_n of the tile rows and halo row ids; at the end the
  // workgroup reduction of the dots (2 x 2 TwoSum and 2 grid doubles per
  // wave and column) reuses it
  size_t lds_bytes = size_t(slot_cap + max_tile) * wc * elem + size_t(halo_cap) * 4;
  const size_t red_bytes = size_t(sh.waves) * 6 * wc * sizeof(double);
  if (lds_bytes < red_bytes) lds_bytes = red_bytes;

  kpm_tiled* p = new kpm_tiled();
  p->n_rows = n_rows;
  p->width = width;
  p->value_type = value_type;
  p->n_tiles = n_tiles;
  p->nnz = nnz;
  p->values = h->values;
  p->shape = sh;
  p->pf = pf;
  p->groups = width / wc;
  p->slot_cap = slot_cap;
  p->halo_cap = halo_cap;
  p->max_tile = max_tile;
  p->halo_total = int64_t(halo_rows.size());
  p->halo_entries = halo_entries;
  p->gather_rows = gather_rows;
  p->lds_bytes = lds_bytes;

  int dev = 0, cus = 0, max_lds = 0;
  rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (!rc)
    rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                   "hipDeviceGetAttribute");
  if (!rc)
    rc = check_hip(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
                   "hipDeviceGetAttribute");
  if (!rc && lds_bytes > size_t(max_lds))
    rc = fail(KPM_ERR_UNSUPPORTED, "tiles need " + std::to_string(lds_bytes) +
                                       " bytes of LDS (tile rows + halo too large)");
  // occupancy with the LDS request, for the modes of this prefetch depth
  int occ = 1 << 20;
  for (int mode : {KPM_DOTS_DOUBLING, KPM_DOTS_NONE}) {
    if (rc) break;
    const void* k = tiled_kernel_for(value_type, sh, mode, pf);
    if (!k) {
      rc = fail(KPM_ERR_UNSUPPORTED, "no tiled kernel for this shape");
      break;
    }
    rc = check_hip(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(lds_bytes)),
                   "hipFuncSetAttribute");
    int o = 0;
    if (!rc)
      rc = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k, sh.waves * 64,
                                                                  lds_bytes),
                     "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (!rc && o < occ) occ = o;
  }
  if (!rc && occ < 1) rc = fail(KPM_ERR_UNSUPPORTED, "tiled kernel does not fit on a CU");
  if (rc) {
    delete p;
    return rc;
  }
  // persistent grid: every workgroup resident, a multiple of 8 x groups
  const int unit = kNumXcd * p->groups;
  int nwg = cus * occ / unit * unit;
  if (nwg < unit) nwg = unit;
  p->nwg = nwg;
  p->nb = nwg / p->groups;
  const int maxnc = 4;  // doubling: 2 TwoSum accumulators + 2 grid sums
  p->part_q = int64_t(p->nb) * maxnc * width * 2;
  const std::vector<int32_t> seg_blk = {0, p->nb};
  const size_t packed = size_t(n_rows) * pf;
  if (!rc) rc = tdev_alloc(&p->tile_begin, tile_begin.size(), "tile_begin");
  if (!rc) rc = tdev_alloc(&p->halo_begin, halo_begin.size(), "halo_begin");
  if (!rc) rc = tdev_alloc(&p->halo_rows, halo_rows.size(), "halo_rows");
  if (!rc) rc = tdev_alloc(&p->lidx, lidx.size(), "local indices");
  if (!rc) rc = tdev_alloc(&p->col_pf, packed, "packed columns");
  if (!rc) rc = tdev_alloc(&p->lidx_pf, packed, "packed local indices");
  if (!rc) rc = tdev_alloc(reinterpret_cast<char**>(&p->val_pf), packed * elem, "packed values");
  if (!rc) rc = tdev_alloc(&p->seg_blk, seg_blk.size(), "seg_blk");
  if (!rc) rc = tdev_alloc(&p->part, size_t(4 * p->part_q), "tiled partials");
  if (!rc) rc = tupload(p->tile_begin, tile_begin, "upload tile_begin");
  if (!rc) rc = tupload(p->halo_begin, halo_begin, "upload halo_begin");
  if (!rc) rc = tupload(p->halo_rows, halo_rows, "upload halo_rows");
  if (!rc) rc = tupload(p->lidx, lidx, "upload local indices");
  if (!rc) rc = tupload(p->col_pf, col_pf, "upload packed columns");
  if (!rc) rc = tupload(p->lidx_pf, lidx_pf, "upload packed local indices");
  if (!rc) rc = tupload(p->seg_blk, seg_blk, "upload seg_blk");
  if (!rc) {
    const int64_t total = int64_t(packed);
    const dim3 grid(unsigned((total + 255) / 256));
    if (value_type == KPM_C64)
      hipLaunchKernelGGL(pack_values_kernel<float>, grid, dim3(256), 0, 0, h->indptr, h->values,
                         p->val_pf, n_rows, pf);
    else
      hipLaunchKernelGGL(pack_values_kernel<double>, grid, dim3(256), 0, 0, h->indptr,
                         h->values, p->val_pf, n_rows, pf);
    rc = check_hip(hipGetLastError(), "pack_values_kernel launch");
  }
  if (!rc) rc = check_hip(hipDeviceSynchronize(), "tiled setup");
  if (rc) {
    free_tiled(p);
    return rc;
  }
  *out = p;
  return KPM_OK;
}

int kpm_tiled_destroy(kpm_tiled* t) {
  if (!t) return KPM_OK;
  hipError_t e = hipDeviceSynchronize();
  free_tiled(t);
  return check_hip(e, "tiled destroy");
}

int kpm_tiled_info(const kpm_tiled* t, int32_t* n_tiles, int32_t* max_tile_rows,
                   int32_t* max_halo_rows, int64_t* halo_rows, int64_t* halo_entries,
                   int64_t* gather_rows, int32_t* workgroups, int32_t* lds_bytes) {
  if (!t) return fail(KPM_ERR_ARG, "tiling is NULL");
  if (n_tiles) *n_tiles = t->n_tiles;
  if (max_tile_rows) *max_tile_rows = t->max_tile;
  if (max_halo_rows) *max_halo_rows = t->halo_cap;
  if (halo_rows) *halo_rows = t->halo_total;
  if (halo_entries) *halo_entries = t->halo_entries;
  if (gather_rows) *gather_rows = t->gather_rows;
  if (workgroups) *workgroups = t->nwg;
  if (lds_bytes) *lds_bytes = int32_t(t->lds_bytes);
  return KPM_OK;
}

int kpm_cheb_steps2(kpm_plan* plan, kpm_tiled* tiles, const kpm_csr* h, void* ring,
                    int64_t slot_stride, int32_t slots, double* d_dots, int32_t out_base,
                    int32_t out_cap, int32_t n_pairs, int32_t dots_mode, kpm_stream_t stream) {
  if (!plan || !tiles) return fail(KPM_ERR_ARG, "plan or tiling is NULL");
  if (!h) return fail(KPM_ERR_ARG, "matrix is NULL");
  PlanState ps{};
  int rc = plan_state(plan, &ps);
  if (rc) return rc;
  if (ps.n_segments != 1) return fail(KPM_ERR_UNSUPPORTED, "tiled steps take one segment");
  if (ps.value_type != tiles->value_type || h->value_type != tiles->value_type)
    return fail(KPM_ERR_ARG, "value types of plan, tiling and matrix differ");
  if (ps.width != tiles->width) return fail(KPM_ERR_ARG, "plan and tiling widths differ");
  if (ps.n_rows != tiles->n_rows || h->n_rows != tiles->n_rows || h->n_cols != tiles->n_rows)
    return fail(KPM_ERR_ARG, "matrix, plan and tiling row counts differ");
  if (h->nnz != tiles->nnz || h->values != tiles->values)
    return fail(KPM_ERR_ARG, "the matrix is not the one the tiling was built from");
  if (!h->indptr || (h->nnz > 0 && (!h->indices || !h->values)))
    return fail(KPM_ERR_ARG, "matrix arrays are NULL");
  if (!ps.scale_set) return fail(KPM_ERR_ARG, "scale not set (kpm_plan_set_scale)");
  if (!ring) return fail(KPM_ERR_ARG, "ring is NULL");
  if (slots < 4) return fail(KPM_ERR_ARG, "two steps per launch need a ring of >= 4 slots");
  if (slot_stride < int64_t(ps.n_rows) * ps.width)
    return fail(KPM_ERR_ARG, "slot_stride smaller than n_rows * width");
  if (dots_mode != KPM_DOTS_DOUBLING && dots_mode != KPM_DOTS_NONE)
    return fail(KPM_ERR_ARG, "tiled steps take KPM_DOTS_DOUBLING or KPM_DOTS_NONE");
  if (dots_mode == KPM_DOTS_DOUBLING && (!d_dots || out_cap <= 0))
    return fail(KPM_ERR_ARG, "dots output buffer missing");
  if (dots_mode == KPM_DOTS_DOUBLING && !ps.grid_set)
    return fail(KPM_ERR_ARG, "dots grid not set (kpm_plan_set_dots_grid)");
  if (n_pairs < 0) return fail(KPM_ERR_ARG, "n_pairs must be non-negative");
  if (n_pairs > 0 && ps.step < 1)
    return fail(KPM_ERR_ARG, "tiled steps start from step 1 (run step 0 with kpm_cheb_steps)");
  if (dots_mode == KPM_DOTS_DOUBLING &&
      (ps.step - out_base < 0 || ps.step + 2 * n_pairs - out_base > out_cap))
    return fail(KPM_ERR_ARG, "dots slots [step - out_base, + 2 n_pairs) outside [0, out_cap)");
  const void* kern = tiled_kernel_for(tiles->value_type, tiles->shape, dots_mode, tiles->pf);
  if (!kern) return fail(KPM_ERR_UNSUPPORTED, "no tiled kernel for this shape");
  // the kernel's dynamic LDS limit follows this tiling (another tiling of the
  // same shape may have set a smaller one)
  rc = check_hip(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(tiles->lds_bytes)),
                 "hipFuncSetAttribute");
  if (rc) return rc;
  TiledParams tp{};
  tp.ring = reinterpret_cast<char*>(ring);
  tp.slot_bytes = slot_stride * int64_t(tiles->value_type == KPM_C128 ? 16 : 8);
  tp.slots = slots;
  tp.width = ps.width;
  tp.tile_begin = tiles->tile_begin;
  tp.halo_begin = tiles->halo_begin;
  tp.halo_rows = tiles->halo_rows;
  tp.scale = ps.scale;
  tp.im_exp = ps.grid_set ? ps.im_exp : nullptr;
  tp.seg_blk = tiles->seg_blk;
  tp.n_tiles = tiles->n_tiles;
  tp.groups = tiles->groups;
  tp.nb = tiles->nb;
  tp.slot_cap = tiles->slot_cap;
  tp.tile_cap = tiles->max_tile;
  const bool dots = dots_mode == KPM_DOTS_DOUBLING;
  const int64_t rec = int64_t(ps.width) * 4;
  hipStream_t s = (hipStream_t)stream;
  int n = ps.step;
  for (int i = 0; i < n_pairs; ++i) {
    const int q = i & 1;
    tp.part_a = tiles->part + (2 * q) * tiles->part_q;
    tp.part_b = tiles->part + (2 * q + 1) * tiles->part_q;
    tp.prev_a = tp.prev_b = nullptr;
    tp.prev_dots_a = tp.prev_dots_b = nullptr;
    if (dots && i > 0) {
      tp.prev_a = tiles->part + (2 * (1 - q)) * tiles->part_q;
      tp.prev_b = tiles->part + (2 * (1 - q) + 1) * tiles->part_q;
      tp.prev_dots_a = d_dots + int64_t(n - 2 - out_base) * rec;
      tp.prev_dots_b = d_dots + int64_t(n - 1 - out_base) * rec;
    }
    void* args[] = {&tp,
                    const_cast<int32_t**>(&h->indptr),
                    const_cast<int32_t**>(&h->indices),
                    const_cast<void**>(&h->values),
                    &tiles->lidx,
                    &tiles->col_pf,
                    &tiles->lidx_pf,
                    &tiles->val_pf,
                    &n};
    rc = check_hip(hipLaunchKernel(kern, dim3(tiles->nwg), dim3(tiles->shape.waves * 64), args,
                                   tiles->lds_bytes, s),
                   "tiled_kernel launch");
    if (rc) return rc;
    n += 2;
    rc = plan_advance(plan, 2);
    if (rc) return rc;
  }
  if (dots && n_pairs > 0) {
    const int q = (n_pairs - 1) & 1;
    for (int half = 0; half < 2; ++half) {
      StepParams p{};
      p.part = tiles->part + (2 * q + half) * tiles->part_q;
      p.nblocks = tiles->nb;
      p.seg_blk = tiles->seg_blk;
      p.nseg = 1;
      p.im_exp = tp.im_exp;
      p.dots = d_dots + int64_t(n - 2 + half - out_base) * rec;
      rc = launch_reduce(dots_mode, ps.width, p, s);
      if (rc) return rc;
    }
  }
  return KPM_OK;
}

}  // extern "C"
