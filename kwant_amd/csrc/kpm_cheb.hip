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
// need no shuffles until the end of the workgroup.  Workgroup partials are
// stored with plain stores (no cross-workgroup synchronisation inside the
// streaming kernel: on a multi-XCD part an agent-scope fence per workgroup
// costs an L2 write-back); the next step's launch reduces them, one
// wavefront per sum in fixed order, before its own rows (the kernel boundary
// makes them visible), and cheb_reduce does so for the last step of a call.
// Every addition is a TwoSum and the
// (sum, error) pair is carried through every level and rounded once at the
// end, so a column's result is the correctly rounded sum of its terms (up to
// the error of the error term) whatever the launch shape or block width, and
// repeated or resumed runs are bitwise identical.
//
// Wave layout: LPR lanes per matrix row, VEC complex values per lane
// (W = LPR * VEC), 64 / LPR rows per wavefront.  Loads are 8 or 16 bytes per
// lane, so a row of the block is read as one coalesced 512-1024 byte access
// when W*sizeof(complex) allows.

#include "kpm_device.h"

namespace kpm {

// The CSR arrays are separate __restrict__ arguments: nothing the kernel
// stores can alias them, so for one or two rows per wavefront (LPR >= 32) the
// compiler reads them with scalar loads, which complete on their own counter
// instead of queueing behind the vector gathers.
//
// PF (pipelined lane-group shapes only): entries per row requested one
// iteration ahead, chosen at launch from the matrix's row-length hint.
template <typename T, int VEC, int LPR, int WAVES, int MODE, int U, int PF>
__global__ __launch_bounds__(WAVES * 64) void cheb_kernel(StepParams p,
                                                          const int32_t* __restrict__ indptr,
                                                          const int32_t* __restrict__ indices,
                                                          const void* __restrict__ values,
                                                          int n, int iters, int only_dots) {
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

  if (MODE != KPM_DOTS_NONE && MODE != KPM_DOTS_LOCAL && p.prev_part != nullptr) {
    // the previous step's dots: its partials are complete (kernel boundary)
    reduce_sums<MODE>(p.prev_part, p.nblocks, p.seg_blk, p.nseg, W, p.im_exp, p.prev_dots,
                      lb * WAVES + wave, p.nblocks * WAVES);
  }
  const int seg = p.blk_seg[lb];
  const int2 rr = p.blk_rows[lb];
  // The workgroup's rows: chunks of ipc wavefront iterations, the first at
  // rr.x, the next ones stride rows further on, up to rr.y.  A contiguous
  // workgroup has one chunk of all its iterations; interleaved workgroups of
  // one XCD take turns over a range so that together they sweep it in step.
  const int64_t stride = p.blk_stride[lb];
  const int ipc = p.chunk_iters > 0 ? p.chunk_iters : iters;
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
  // exact grid sums (doubling mode) and the columns' grid scales
  constexpr int NQ = NQuant<MODE>::value;
  constexpr int NQA = NQ > 0 ? NQ : 1;
  double aq[NQA][VEC];
  int qk[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
#pragma unroll
    for (int c = 0; c < NQA; ++c) aq[c][v] = 0.0;
    qk[v] = NQ > 0 && p.im_exp != nullptr ? p.im_exp[seg * W + col + v] : 0;
  }
  // complex64 doubling steps always run with the grid set (the C entry points
  // require it; <a_0|a_0> itself is taken in the projection mode)
  constexpr bool fg = MODE == KPM_DOTS_DOUBLING && std::is_same<T, float>::value;

  if constexpr (RPW <= 2) {
    // In each chunk a wave owns a contiguous range of ipc * RPW rows and walks
    // it U row slots at a time: all streaming loads and gathers of the U
    // slots are issued before the first is consumed.
    for (int64_t cb = rr.x; cb < rr.y; cb += stride) {
    const int64_t wb64 = cb + int64_t(wave) * ipc * RPW;
    if (wb64 >= rr.y) break;
    const int wbeg = int(wb64);
    const int wend = wb64 + int64_t(ipc) * RPW < rr.y ? int(wb64 + int64_t(ipc) * RPW) : rr.y;
    for (int r0 = __builtin_amdgcn_readfirstlane(wbeg); r0 < wend; r0 += U * RPW) {
      CLane<T, VEC> xs[U], xp[U], y[U], ww[U];
      int64_t roff[U];
      bool live[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int row = r0 + u * RPW + grp;
        live[u] = row < wend;
        roff[u] = 2 * (int64_t(live[u] ? row : r0) * W + col);
        // a separate, early load of the row's own a_n (it also brings the row
        // into L2 ahead of the gathers of its neighbours)
        xs[u] = IO::load(cur + roff[u]);
        if (!only_dots && n > 0)
          xp[u] = IO::load_nt(prv + roff[u]);
        else
          zero(xp[u]);
        if (MODE == KPM_DOTS_PROJECT || MODE == KPM_DOTS_LOCAL)
          ww[u] = IO::load(wv + roff[u]);
        else
          zero(ww[u]);
        zero(y[u]);
      }
      if (!only_dots) {
        CLane<T, VEC> acc[U];
        spmm_rows_scalar<T, VEC, RPW, U>(indptr, indices, values, cur, r0, wend - r0, grp, col,
                                         W, acc);
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            y[u].re[v] = fma(acc[u].re[v] - bsh * xs[u].re[v], mul, -xp[u].re[v]);
            y[u].im[v] = fma(acc[u].im[v] - bsh * xs[u].im[v], mul, -xp[u].im[v]);
          }
          if (live[u]) IO::store_nt(nxt + roff[u], y[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!live[u]) {
          zero(xs[u]);
          zero(y[u]);
        }
        add_dots<T, VEC, MODE, NC, NQA>(as, ac, aq, qk, fg, xs[u], y[u], ww[u]);
        if constexpr (MODE == KPM_DOTS_LOCAL)
          write_local<T, VEC, LPR>(p, r0 + u * RPW + grp, live[u], col, only_dots ? xs[u] : y[u],
                                   ww[u]);
      }
    }
    }
  } else if constexpr (U == 2) {
    // Software-pipelined lane groups.  While iteration it gathers its
    // neighbours, the first PF entries of iteration it + 1's row and the
    // extent of iteration it + 2's row are already requested, so an iteration
    // waits on one round of memory latency (the gathers) instead of three
    // (extent, entries, gathers).  A group loads its row's entries
    // cooperatively (lane s of the group takes entries s, s + LPR, ...) and
    // shuffles them to the group; entries past PF (rows of higher degree) are
    // handled chunk by chunk after the prefetched ones, in the same order, so
    // the sum is bitwise the other shapes' sum.
    static_assert(PF > 0, "pipelined lane groups need a prefetch depth");
    constexpr int NCH = (PF + LPR - 1) / LPR;
    const int sub = lane % LPR;
    auto extent = [&](int it_, int& k0, int& k1) {
      const int64_t r = iter_row0(it_, rr.x, stride, ipc, WAVES * RPW) + wave * RPW + grp;
      k0 = 0;
      k1 = 0;
      if (!only_dots && it_ < iters && r < rr.y) {
        k0 = indptr[r];
        k1 = indptr[r + 1];
      }
    };
    auto head = [&](int k0, int k1, int* j, T* hr, T* hi) {
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const int e = c * LPR + sub;
        j[c] = 0;
        hr[c] = T(0);
        hi[c] = T(0);
        if (e < PF && k0 + e < k1) {
          j[c] = indices[k0 + e];
          load_value<T>(values, k0 + e, hr[c], hi[c]);
        }
      }
    };
    int kA0, kA1, kB0, kB1;
    int jA[NCH];
    T hrA[NCH], hiA[NCH];
    extent(0, kA0, kA1);
    head(kA0, kA1, jA, hrA, hiA);
    extent(1, kB0, kB1);
    for (int it = 0; it < iters; ++it) {
      const int64_t r64 = iter_row0(it, rr.x, stride, ipc, WAVES * RPW) + wave * RPW;
      if (r64 >= rr.y) break;
      const int row0 = __builtin_amdgcn_readfirstlane(int(r64));
      const int row = row0 + grp;
      const bool live = row < rr.y;
      const int64_t roff = 2 * (int64_t(live ? row : row0) * W + col);
      int jB[NCH];
      T hrB[NCH], hiB[NCH];
      int kC0, kC1;
      head(kB0, kB1, jB, hrB, hiB);
      extent(it + 2, kC0, kC1);
      CLane<T, VEC> y, xs, ww;
      zero(y);
      xs = IO::load(cur + roff);
      if (!only_dots) {
        CLane<T, VEC> xp, acc;
        if (n > 0)
          xp = IO::load_nt(prv + roff);
        else
          zero(xp);
        zero(acc);
        const int deg = kA1 - kA0;
        CLane<T, VEC> x[PF];
        T hr[PF], hi[PF];
#pragma unroll
        for (int e = 0; e < PF; ++e) {
          const int j = __shfl(jA[e / LPR], e % LPR, LPR);
          hr[e] = __shfl(hrA[e / LPR], e % LPR, LPR);
          hi[e] = __shfl(hiA[e / LPR], e % LPR, LPR);
          if (e < deg)
            x[e] = IO::load(cur + 2 * (int64_t(j) * W + col));
          else
            zero(x[e]);
        }
#pragma unroll
        for (int e = 0; e < PF; ++e) {
          if (e < deg) {
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
              acc.re[v] = fma(hr[e], x[e].re[v], acc.re[v]);
              acc.re[v] = fma(-hi[e], x[e].im[v], acc.re[v]);
              acc.im[v] = fma(hr[e], x[e].im[v], acc.im[v]);
              acc.im[v] = fma(hi[e], x[e].re[v], acc.im[v]);
            }
          }
        }
        for (int kt = kA0 + PF; kt < kA1; kt += LPR) {
          // entries past the prefetched ones: the whole group takes this loop
          // together (one row per group), so the shuffles stay in the group
          int jt = 0;
          T htr = T(0), hti = T(0);
          if (kt + sub < kA1) {
            jt = indices[kt + sub];
            load_value<T>(values, kt + sub, htr, hti);
          }
#pragma unroll
          for (int s = 0; s < LPR; ++s) {
            const int j = __shfl(jt, s, LPR);
            const T er = __shfl(htr, s, LPR);
            const T ei = __shfl(hti, s, LPR);
            if (kt + s < kA1) {
              const CLane<T, VEC> xe = IO::load(cur + 2 * (int64_t(j) * W + col));
#pragma unroll
              for (int v = 0; v < VEC; ++v) {
                acc.re[v] = fma(er, xe.re[v], acc.re[v]);
                acc.re[v] = fma(-ei, xe.im[v], acc.re[v]);
                acc.im[v] = fma(er, xe.im[v], acc.im[v]);
                acc.im[v] = fma(ei, xe.re[v], acc.im[v]);
              }
            }
          }
        }
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          y.re[v] = fma(acc.re[v] - bsh * xs.re[v], mul, -xp.re[v]);
          y.im[v] = fma(acc.im[v] - bsh * xs.im[v], mul, -xp.im[v]);
        }
        if (live) IO::store_nt(nxt + roff, y);
      }
      if (MODE == KPM_DOTS_PROJECT || MODE == KPM_DOTS_LOCAL)
        ww = IO::load(wv + roff);
      else
        zero(ww);
      if (!live) {
        zero(xs);
        zero(y);
      }
      add_dots<T, VEC, MODE, NC, NQA>(as, ac, aq, qk, fg, xs, y, ww);
      if constexpr (MODE == KPM_DOTS_LOCAL)
        write_local<T, VEC, LPR>(p, row, live, col, only_dots ? xs : y, ww);
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        jA[c] = jB[c];
        hrA[c] = hrB[c];
        hiA[c] = hiB[c];
      }
      kA0 = kB0;
      kA1 = kB1;
      kB0 = kC0;
      kB1 = kC1;
    }
  } else {
    for (int it = 0; it < iters; ++it) {
      // rows row0 .. row0 + RPW - 1 belong to this wave; lane group grp owns one
      const int64_t r64 = iter_row0(it, rr.x, stride, ipc, WAVES * RPW) + wave * RPW;
      if (r64 >= rr.y) break;
      const int row0 = __builtin_amdgcn_readfirstlane(int(r64));
      const int row = row0 + grp;
      const bool live = row < rr.y;
      const int64_t roff = 2 * (int64_t(live ? row : row0) * W + col);
      CLane<T, VEC> y, xs, ww;
      zero(y);
      if (!only_dots) {
        // per-lane matrix entries are vector loads, which complete in order:
        // request them before the streaming row loads
        CLane<T, VEC> xp;
        CLane<T, VEC> acc = spmm_row_vector<T, VEC>(indptr, indices, values, cur,
                                                    live ? row : row0, live, col, W, xs, xp,
                                                    prv, roff, n > 0);
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          y.re[v] = fma(acc.re[v] - bsh * xs.re[v], mul, -xp.re[v]);
          y.im[v] = fma(acc.im[v] - bsh * xs.im[v], mul, -xp.im[v]);
        }
        if (live) IO::store_nt(nxt + roff, y);
      } else {
        xs = IO::load(cur + roff);
      }
      if (MODE == KPM_DOTS_PROJECT || MODE == KPM_DOTS_LOCAL)
        ww = IO::load(wv + roff);
      else
        zero(ww);
      if (!live) {
        zero(xs);
        zero(y);
      }
      add_dots<T, VEC, MODE, NC, NQA>(as, ac, aq, qk, fg, xs, y, ww);
      if constexpr (MODE == KPM_DOTS_LOCAL)
        write_local<T, VEC, LPR>(p, row, live, col, only_dots ? xs : y, ww);
    }
  }

  if (MODE == KPM_DOTS_NONE || MODE == KPM_DOTS_LOCAL) return;

  finish_grid_dots<T, MODE>(fg, as, qk);
  store_partials<VEC, LPR, WAVES, NC, NQ, NQA>(p, lb, wave, grp, col, tid, as, ac, aq);
}

// Standalone reduction (after the last step of a call and for kpm_block_dots):
// one wavefront per sum.
template <int MODE>
__global__ __launch_bounds__(256) void cheb_reduce(const double* __restrict__ part, int nblocks,
                                                   const int32_t* __restrict__ seg_blk,
                                                   int nseg, int W,
                                                   const int32_t* __restrict__ qexp,
                                                   double* __restrict__ out) {
  reduce_sums<MODE>(part, nblocks, seg_blk, nseg, W, qexp, out,
                    blockIdx.x * 4 + (threadIdx.x >> 6), gridDim.x * 4);
}

// The Chebyshev step with the matrix in 4 x 4 block-CSR (the norbs x norbs
// site blocks kwant assembles, kwant/_system.pyx:94-153).  The lane layout is
// cheb_kernel's per-lane one (LPR lanes x VEC columns per row) with the four
// rows of a block-row in consecutive lane groups, so the 4 * LPR lanes of a
// block-row load a neighbour block's four rows of a_n as one contiguous
// 4 * W complex read and exchange them with shuffles; each lane then applies
// its row of the 4 x 4 block.  Blocks follow in column order and a block's
// columns in order, so every row sum is the same FMA sequence as the CSR
// kernel's on the same (column-sorted) entries -- the padding zeros of a
// block add exact zeros.  Dots, partials and the epilogue are cheb_kernel's.
template <typename T, int VEC, int LPR, int WAVES, int MODE>
__global__ __launch_bounds__(WAVES * 64) void bsr_kernel(StepParams p,
                                                         const int32_t* __restrict__ bptr,
                                                         const int32_t* __restrict__ bidx,
                                                         const void* __restrict__ bval, int n,
                                                         int iters) {
  static_assert(LPR <= 16 && 64 % (4 * LPR) == 0, "a block-row needs 4 lane groups");
  constexpr int W = LPR * VEC;
  constexpr int RPW = 64 / LPR;
  constexpr int NC = NComp<MODE>::value;
  constexpr int NQ = NQuant<MODE>::value;
  constexpr int NQA = NQ > 0 ? NQ : 1;
  using IO = LaneIO<T, VEC>;

  const int hb = blockIdx.x;
  const int per = p.nblocks_launch / p.xcd;
  const int lb = (hb % p.xcd) * per + hb / p.xcd;
  if (lb >= p.nblocks) return;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int grp = lane / LPR;
  const int sub = lane % LPR;
  const int col = sub * VEC;
  const int brow = grp & 3;                  // row within the block
  const int gbase = (grp & ~3) * LPR + sub;  // lane of block row 0, same columns

  if (MODE == KPM_DOTS_DOUBLING && p.prev_part != nullptr)
    reduce_sums<MODE>(p.prev_part, p.nblocks, p.seg_blk, p.nseg, W, p.im_exp, p.prev_dots,
                      lb * WAVES + wave, p.nblocks * WAVES);
  const int seg = p.blk_seg[lb];
  const int2 rr = p.blk_rows[lb];
  const int64_t stride = p.blk_stride[lb];
  const int ipc = p.chunk_iters > 0 ? p.chunk_iters : iters;
  const double sa = p.scale[2 * seg];
  const double sb = p.scale[2 * seg + 1];
  const T mul = T((n == 0 ? 1.0 : 2.0) / sa);
  const T bsh = T(sb);
  const T* cur = reinterpret_cast<const T*>(p.ring + int64_t(n % p.slots) * p.slot_bytes);
  T* nxt = reinterpret_cast<T*>(p.ring + int64_t((n + 1) % p.slots) * p.slot_bytes);
  const T* prv = reinterpret_cast<const T*>(
      p.ring + int64_t((n + p.slots - 1) % p.slots) * p.slot_bytes);

  double as[NC][VEC], ac[NC][VEC], aq[NQA][VEC];
  int qk[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
#pragma unroll
    for (int c = 0; c < NC; ++c) as[c][v] = ac[c][v] = 0.0;
#pragma unroll
    for (int c = 0; c < NQA; ++c) aq[c][v] = 0.0;
    qk[v] = NQ > 0 && p.im_exp != nullptr ? p.im_exp[seg * W + col + v] : 0;
  }
  // complex64 doubling steps always run with the grid set (the C entry points
  // require it; <a_0|a_0> itself is taken in the projection mode)
  constexpr bool fg = MODE == KPM_DOTS_DOUBLING && std::is_same<T, float>::value;

  for (int it = 0; it < iters; ++it) {
    const int64_t r64 = iter_row0(it, rr.x, stride, ipc, WAVES * RPW) + wave * RPW;
    if (r64 >= rr.y) break;
    const int row0 = __builtin_amdgcn_readfirstlane(int(r64));
    const int row = row0 + grp;
    const bool live = row < rr.y;  // whole block-rows: range ends are multiples of 4
    const int64_t roff = 2 * (int64_t(live ? row : row0) * W + col);
    const int br = (live ? row : row0) >> 2;
    const int k0 = live ? bptr[br] : 0;
    const int len = live ? bptr[br + 1] - k0 : 0;
    int lmax = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_xor(lmax, off);
      lmax = o > lmax ? o : lmax;
    }
    CLane<T, VEC> xs = IO::load(cur + roff), xp, y, acc;
    if (n > 0)
      xp = IO::load_nt(prv + roff);
    else
      zero(xp);
    zero(acc);
    for (int t = 0; t < lmax; ++t) {
      const bool act = t < len;
      const int k = act ? k0 + t : 0;
      CLane<T, VEC> xm;
      T hr[4], hi[4];
      if (act) {
        const int j = bidx[k];
        xm = IO::load(cur + 2 * (int64_t(4 * j + brow) * W + col));
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) load_value<T>(bval, int64_t(k) * 16 + brow * 4 + kk, hr[kk], hi[kk]);
      } else {
        zero(xm);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) hr[kk] = hi[kk] = T(0);
      }
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        CLane<T, VEC> xk;
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          xk.re[v] = __shfl(xm.re[v], gbase + kk * LPR);
          xk.im[v] = __shfl(xm.im[v], gbase + kk * LPR);
        }
        if (act) {
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            acc.re[v] = fma(hr[kk], xk.re[v], acc.re[v]);
            acc.re[v] = fma(-hi[kk], xk.im[v], acc.re[v]);
            acc.im[v] = fma(hr[kk], xk.im[v], acc.im[v]);
            acc.im[v] = fma(hi[kk], xk.re[v], acc.im[v]);
          }
        }
      }
    }
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      y.re[v] = fma(acc.re[v] - bsh * xs.re[v], mul, -xp.re[v]);
      y.im[v] = fma(acc.im[v] - bsh * xs.im[v], mul, -xp.im[v]);
    }
    if (live) IO::store_nt(nxt + roff, y);
    if (!live) {
      zero(xs);
      zero(y);
    }
    CLane<T, VEC> ww;
    zero(ww);
    add_dots<T, VEC, MODE, NC, NQA>(as, ac, aq, qk, fg, xs, y, ww);
  }
  if (MODE == KPM_DOTS_NONE) return;
  finish_grid_dots<T, MODE>(fg, as, qk);
  store_partials<VEC, LPR, WAVES, NC, NQ, NQA>(p, lb, wave, grp, col, tid, as, ac, aq);
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

// The same product for SG consecutive slots per workgroup: a row's entries are
// loaded once and applied to every slot, so the matrix crosses HBM once per SG
// slots instead of once per slot.  Each (row, slot) sum is spmm_kernel's
// sequence of FMAs (padding entries included), so the results are bitwise
// equal.
template <typename T, int VEC, int LPR, int WAVES, int SG>
__global__ __launch_bounds__(WAVES * 64) void spmm_slots_kernel(
    const int32_t* indptr, const int32_t* indices, const void* values, const T* x,
    int64_t x_stride, T* y, int64_t y_stride, int n_rows, int n_slots) {
  constexpr int W = LPR * VEC;
  constexpr int RPW = 64 / LPR;
  using IO = LaneIO<T, VEC>;
  using CL = CLane<T, VEC>;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int row = (blockIdx.x * WAVES + wave) * RPW + lane / LPR;
  const int col = (lane % LPR) * VEC;
  if (row >= n_rows) return;
  const int s0 = blockIdx.y * SG;
  const int ns = n_slots - s0 < SG ? n_slots - s0 : SG;
  const T* xs = x + 2 * int64_t(s0) * x_stride;
  T* ys = y + 2 * int64_t(s0) * y_stride;
  const int k0 = indptr[row], k1 = indptr[row + 1];
  CL acc[SG];
#pragma unroll
  for (int s = 0; s < SG; ++s) zero(acc[s]);
  for (int k = k0; k < k1; k += 4) {
    T hr[4], hi[4];
    int j[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      hr[u] = hi[u] = T(0);
      j[u] = -1;
      if (k + u < k1) {
        j[u] = indices[k + u];
        load_value<T>(values, k + u, hr[u], hi[u]);
      }
    }
#pragma unroll
    for (int s = 0; s < SG; ++s) {
      if (s < ns) {
        CL xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (j[u] >= 0)
            xv[u] = IO::load(xs + 2 * (int64_t(s) * x_stride + int64_t(j[u]) * W + col));
          else
            zero(xv[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            acc[s].re[v] = fma(hr[u], xv[u].re[v], acc[s].re[v]);
            acc[s].re[v] = fma(-hi[u], xv[u].im[v], acc[s].re[v]);
            acc[s].im[v] = fma(hr[u], xv[u].im[v], acc[s].im[v]);
            acc[s].im[v] = fma(hi[u], xv[u].re[v], acc[s].im[v]);
          }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < SG; ++s)
    if (s < ns) IO::store(ys + 2 * (int64_t(s) * y_stride + int64_t(row) * W + col), acc[s]);
}

// Bond records of a current-like operator (kwant.operator.Current,
// kwant/operator.pyx:940-1001): for every entry k of the bond matrix B (row
// r, column s = indices[k], value h = values[k]) and every block column c,
//   t_k(c) = i (conj(h) conj(bra[s,c]) v[r,c] - conj(w[r,c]) h ket[s,c])
// with v = M ket and w = M^H bra for the block-diagonal onsite matrix M.
// Summed over the block's columns the entry's h factors out:
//   sum_c t_k(c) = i (conj(h) A - h B),  A = sum_c conj(bra[s,c]) v[r,c],
//                                        B = sum_c conj(w[r,c]) ket[s,c],
// so a column costs two FP64 complex products.  A lane group of LPR lanes
// owns a row; lane `sub` holds G column groups of VEC, group g at columns
// (g * LPR + sub) * VEC (each load instruction reads LPR * VEC contiguous
// columns).  Column sums are plain FP64 additions in a fixed order (lane
// sums, then a butterfly over the group), so a (width, value type) always
// gives the same bits.  out[k] = complex128 record, or per column
// out[k * W + c] = t_k(c) (per_column).
template <typename T, int VEC, int LPR, int G, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void bond_kernel(
    const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const void* __restrict__ values, const T* __restrict__ bra, const T* __restrict__ ket,
    const T* __restrict__ v, const T* __restrict__ w, double* __restrict__ out, int per_column,
    int n_rows) {
  constexpr int W = LPR * VEC * G;
  constexpr int RPW = 64 / LPR;
  using IO = LaneIO<T, VEC>;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  // XCD-aware order: hardware block hb runs on XCD hb % kNumXcd; each XCD
  // takes one contiguous range of row blocks, so the rows that a row's
  // entries gather from stay in that XCD's L2
  const int per = gridDim.x / kNumXcd;
  const int lb = (blockIdx.x % kNumXcd) * per + blockIdx.x / kNumXcd;
  const int row = (lb * WAVES + wave) * RPW + lane / LPR;
  const int sub = lane % LPR;
  if (row >= n_rows) return;  // whole lane groups leave together
  auto colof = [&](int g) { return (g * LPR + sub) * VEC; };
  // the row's v and w stay in their storage type (registers decide the
  // occupancy, and occupancy the number of gathers in flight)
  CLane<T, VEC> vv[G], ww[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    vv[g] = IO::load(v + 2 * (int64_t(row) * W + colof(g)));
    ww[g] = IO::load(w + 2 * (int64_t(row) * W + colof(g)));
  }
  d2v* rec = reinterpret_cast<d2v*>(out);
  const int k1 = indptr[row + 1];
  for (int k = indptr[row]; k < k1; ++k) {
    const int s = indices[k];
    T hr_, hi_;
    load_value<T>(values, k, hr_, hi_);
    const double hr = hr_, hi = hi_;
    CLane<T, VEC> bs[G], ks[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      bs[g] = IO::load(bra + 2 * (int64_t(s) * W + colof(g)));
      ks[g] = IO::load(ket + 2 * (int64_t(s) * W + colof(g)));
    }
    if (per_column) {
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int u = 0; u < VEC; ++u) {
          const double br = bs[g].re[u], bi = bs[g].im[u];
          const double kr = ks[g].re[u], ki = ks[g].im[u];
          const double vr = vv[g].re[u], vi = vv[g].im[u];
          const double wr = ww[g].re[u], wi = ww[g].im[u];
          // a = conj(bra) v, b = conj(w) ket, t = i (conj(h) a - h b)
          const double ar = fma(br, vr, bi * vi);
          const double ai = fma(br, vi, -(bi * vr));
          const double bre = fma(wr, kr, wi * ki);
          const double bim = fma(wr, ki, -(wi * kr));
          const double dr = fma(hr, ar, hi * ai) - fma(hr, bre, -(hi * bim));
          const double di = fma(hr, ai, -(hi * ar)) - fma(hr, bim, hi * bre);
          rec[int64_t(k) * W + colof(g) + u] = d2v{-di, dr};
        }
      continue;
    }
    double ar = 0.0, ai = 0.0, bre = 0.0, bim = 0.0;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int u = 0; u < VEC; ++u) {
        const double br = bs[g].re[u], bi = bs[g].im[u];
        const double kr = ks[g].re[u], ki = ks[g].im[u];
        const double vr = vv[g].re[u], vi = vv[g].im[u];
        const double wr = ww[g].re[u], wi = ww[g].im[u];
        ar = fma(br, vr, fma(bi, vi, ar));
        ai = fma(br, vi, fma(-bi, vr, ai));
        bre = fma(wr, kr, fma(wi, ki, bre));
        bim = fma(wr, ki, fma(-wi, kr, bim));
      }
#pragma unroll
    for (int off = 1; off < LPR; off <<= 1) {
      ar += __shfl_xor(ar, off);
      ai += __shfl_xor(ai, off);
      bre += __shfl_xor(bre, off);
      bim += __shfl_xor(bim, off);
    }
    if (sub == 0) {
      // i (conj(h) A - h B)
      const double dr = fma(hr, ar, hi * ai) - fma(hr, bre, -(hi * bim));
      const double di = fma(hr, ai, -(hi * ar)) - fma(hr, bim, hi * bre);
      rec[k] = d2v{-di, dr};
    }
  }
}

// ----------------------------------------------------------------------------
// Variant table and dispatch.

namespace {

#ifndef KPM_WAVES
#define KPM_WAVES 4
#endif
constexpr int kWaves = KPM_WAVES;
// Wide workgroups for the pipelined lane-group loop: one workgroup fills a
// CU, so an XCD runs a quarter as many row streams and a row read as the
// +-plane neighbour of a 3D lattice is more often still in its L2
// (256-system cubic ensemble: 10.4 vs 10.9 ms per step).
constexpr int kWideWaves = 16;

// Prefetch depths of the pipelined lane-group loop: the smallest one that
// covers the matrix's typical row length is launched (padding a row's
// prefetch past its length costs shuffles and registers, and the registers
// cost occupancy).
constexpr int kPrefetchDepths[] = {4, 6, 8, 10, 12, 16};
constexpr int kDefaultPrefetch = 12;

template <int MODE>
int launch_reduce_mode(const StepParams& p, int width, hipStream_t stream) {
  constexpr int NG = NComp<MODE>::value + (NQuant<MODE>::value > 0 ? 1 : 0);
  const int sums = p.nseg * NG * width;
  hipLaunchKernelGGL((cheb_reduce<MODE>), dim3((sums + 3) / 4), dim3(256), 0, stream, p.part,
                     p.nblocks, p.seg_blk, p.nseg, width, p.im_exp, p.dots);
  return check_hip(hipGetLastError(), "cheb_reduce launch");
}

template <typename T, int VEC, int LPR, int U, int PF, int WV = kWaves>
int launch_u(int mode, bool only_dots, int iters, int n, const StepParams& p,
             hipStream_t stream) {
  dim3 grid(p.nblocks_launch), block(WV * 64);
  switch (mode) {
    case KPM_DOTS_NONE:
      hipLaunchKernelGGL((cheb_kernel<T, VEC, LPR, WV, KPM_DOTS_NONE, U, PF>), grid, block, 0,
                         stream, p, p.indptr, p.indices, p.values, n, iters, int(only_dots));
      break;
    case KPM_DOTS_DOUBLING:
      hipLaunchKernelGGL((cheb_kernel<T, VEC, LPR, WV, KPM_DOTS_DOUBLING, U, PF>), grid, block,
                         0, stream, p, p.indptr, p.indices, p.values, n, iters,
                         int(only_dots));
      break;
    case KPM_DOTS_PROJECT:
      hipLaunchKernelGGL((cheb_kernel<T, VEC, LPR, WV, KPM_DOTS_PROJECT, U, PF>), grid, block,
                         0, stream, p, p.indptr, p.indices, p.values, n, iters,
                         int(only_dots));
      break;
    case KPM_DOTS_LOCAL:
    case KPM_DOTS_LOCAL_COLS:  // the same kernel; p.local_cols selects the record
      hipLaunchKernelGGL((cheb_kernel<T, VEC, LPR, WV, KPM_DOTS_LOCAL, U, PF>), grid, block, 0,
                         stream, p, p.indptr, p.indices, p.values, n, iters, int(only_dots));
      break;
    default:
      return fail(KPM_ERR_ARG, "unknown dots mode");
  }
  return check_hip(hipGetLastError(), "cheb_kernel launch");
}

// unroll 2: row-slot unrolling for the scalar-metadata shapes (LPR >= 32),
// the pipelined loop for lane groups of 4 to 16 lanes with prefetch depth pf
// (one of kPrefetchDepths).
template <typename T, int VEC, int LPR, int WV>
int launch_pipelined(int mode, bool only_dots, int iters, int n, int pf, const StepParams& p,
                     hipStream_t stream) {
  switch (pf) {
    case 4: return launch_u<T, VEC, LPR, 2, 4, WV>(mode, only_dots, iters, n, p, stream);
    case 6: return launch_u<T, VEC, LPR, 2, 6, WV>(mode, only_dots, iters, n, p, stream);
    case 8: return launch_u<T, VEC, LPR, 2, 8, WV>(mode, only_dots, iters, n, p, stream);
    case 10: return launch_u<T, VEC, LPR, 2, 10, WV>(mode, only_dots, iters, n, p, stream);
    case 12: return launch_u<T, VEC, LPR, 2, 12, WV>(mode, only_dots, iters, n, p, stream);
    case 16: return launch_u<T, VEC, LPR, 2, 16, WV>(mode, only_dots, iters, n, p, stream);
    default: return fail(KPM_ERR_ARG, "unsupported prefetch depth");
  }
}

template <typename T, int VEC, int LPR>
int launch_one(int mode, bool only_dots, int iters, int n, int unroll, int pf, int waves,
               const StepParams& p, hipStream_t stream) {
  if constexpr (LPR >= 32) {
    if (unroll == 2) return launch_u<T, VEC, LPR, 2, 0>(mode, only_dots, iters, n, p, stream);
  } else if constexpr (LPR >= 4) {
    if (unroll == 2) {
      if (waves == kWideWaves)
        return launch_pipelined<T, VEC, LPR, kWideWaves>(mode, only_dots, iters, n, pf, p,
                                                         stream);
      return launch_pipelined<T, VEC, LPR, kWaves>(mode, only_dots, iters, n, pf, p, stream);
    }
  }
  return launch_u<T, VEC, LPR, 1, 0>(mode, only_dots, iters, n, p, stream);
}

template <typename T, int VEC, int LPR>
int launch_spmm_one(const kpm_csr* op, const void* x, int64_t xs, void* y, int64_t ys,
                    int n_slots, hipStream_t stream) {
  constexpr int RPB = kWaves * 64 / LPR;
  // slots per workgroup of the multi-slot kernel
  constexpr int SG = 8;
  if (n_slots >= SG) {
    dim3 grid((op->n_rows + RPB - 1) / RPB, (n_slots + SG - 1) / SG), block(kWaves * 64);
    hipLaunchKernelGGL((spmm_slots_kernel<T, VEC, LPR, kWaves, SG>), grid, block, 0, stream,
                       op->indptr, op->indices, op->values, reinterpret_cast<const T*>(x), xs,
                       reinterpret_cast<T*>(y), ys, op->n_rows, n_slots);
    return check_hip(hipGetLastError(), "spmm_slots_kernel launch");
  }
  dim3 grid((op->n_rows + RPB - 1) / RPB, n_slots), block(kWaves * 64);
  hipLaunchKernelGGL((spmm_kernel<T, VEC, LPR, kWaves>), grid, block, 0, stream, op->indptr,
                     op->indices, op->values, reinterpret_cast<const T*>(x), xs,
                     reinterpret_cast<T*>(y), ys, op->n_rows);
  return check_hip(hipGetLastError(), "spmm_kernel launch");
}

template <typename T, int VEC, int LPR, int G>
int launch_bond_one(const kpm_csr* b, const void* bra, const void* ket, const void* v,
                    const void* w, int per_column, double* out, hipStream_t stream) {
  constexpr int RPB = kWaves * 64 / LPR;
  // a multiple of kNumXcd blocks (the kernel's XCD-aware order needs it)
  const int blocks = (b->n_rows + RPB - 1) / RPB;
  dim3 grid((blocks + kNumXcd - 1) / kNumXcd * kNumXcd), block(kWaves * 64);
  hipLaunchKernelGGL((bond_kernel<T, VEC, LPR, G, kWaves>), grid, block, 0, stream, b->indptr,
                     b->indices, b->values, reinterpret_cast<const T*>(bra),
                     reinterpret_cast<const T*>(ket), reinterpret_cast<const T*>(v),
                     reinterpret_cast<const T*>(w), out, per_column, b->n_rows);
  return check_hip(hipGetLastError(), "bond_kernel launch");
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
  // A variant is a lane layout plus how many residency rounds the grid is
  // sized for (1: every workgroup resident at once, one contiguous row range
  // each; R: R times as many, R times shorter).  Rows per workgroup are set by
  // the plan from the kernel's occupancy and the device's CU count.  Many
  // rounds keep the resident workgroups of an XCD on one narrow window of
  // rows, so neighbours far apart in the row order (the next plane of a 3D
  // lattice) are read while still in L2; few rounds suit matrices whose
  // neighbours are close (2D lattices).
  auto add = [&](int lpr, int vec) {
    // unroll 2: two row slots per wave (LPR >= 32) or the pipelined
    // lane-group loop (LPR 4 .. 16), which comes first for lane groups
    const int n_unroll = lpr >= 4 ? 2 : 1;
    const int first = lpr >= 4 && lpr < 32 ? 2 : 1;
    for (int ui = 0; ui < n_unroll; ++ui) {
      const int unroll = ui == 0 ? first : 3 - first;
      for (int rounds : {2, 1, 8, 32, 128}) {
        if (n < max_out) out[n] = KernelShape{lpr, vec, kWaves, rounds, unroll, 0, 0, 0};
        ++n;
      }
      // The interleaved sweep needs every workgroup resident at once; the
      // occupancy query can promise one workgroup per CU more than the
      // dispatcher places (measured on the 2M-site graphene block with 64
      // vectors: 0.87 ms per step at the reported 7 workgroups per CU, 0.68 ms
      // at 6), so each chunk also comes with one spare slot per CU.
      for (int chunk : {8, 32})
        for (int spare = 0; spare <= 1; ++spare) {
          if (n < max_out) out[n] = KernelShape{lpr, vec, kWaves, 1, unroll, chunk, spare, 0};
          ++n;
        }
      if (unroll == 2 && lpr < 32)
        for (int rounds : {2, 1, 8, 32, 128}) {
          if (n < max_out) out[n] = KernelShape{lpr, vec, kWideWaves, rounds, 2, 0, 0, 0};
          ++n;
        }
    }
  };
  // Further interleave chunks, listed after every shape above so that those
  // keep their variant numbers.
  auto add_chunks = [&](int lpr, int vec) {
    const int n_unroll = lpr >= 4 ? 2 : 1;
    const int first = lpr >= 4 && lpr < 32 ? 2 : 1;
    for (int ui = 0; ui < n_unroll; ++ui) {
      const int unroll = ui == 0 ? first : 3 - first;
      for (int chunk : {4, 16})
        for (int spare = 0; spare <= 1; ++spare) {
          if (n < max_out) out[n] = KernelShape{lpr, vec, kWaves, 1, unroll, chunk, spare, 0};
          ++n;
        }
    }
  };
  if (value_type == KPM_C64) {
    if (width >= 2) add(width / 2, 2);
    if (width <= 64) add(width, 1);
    if (width >= 2) add_chunks(width / 2, 2);
    if (width <= 64) add_chunks(width, 1);
  } else if (value_type == KPM_C128) {
    if (width <= 64) {
      add(width, 1);
      add_chunks(width, 1);
    }
  }
  return n < max_out ? n : max_out;  // never more than the caller's array holds
}

namespace {

struct OccQuery {
  int* out;
  int unroll;
  int waves;
  template <typename T, int V, int LP>
  int operator()() {
    // the pipelined lane-group shapes are sized with the default depth
    if constexpr (LP >= 4 && LP < 32) {
      if (unroll == 2 && waves == kWideWaves)
        return check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             out,
                             cheb_kernel<T, V, LP, kWideWaves, KPM_DOTS_DOUBLING, 2,
                                         kDefaultPrefetch>,
                             kWideWaves * 64, 0),
                         "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    }
    if constexpr (LP >= 4) {
      if (unroll == 2)
        return check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             out,
                             cheb_kernel<T, V, LP, kWaves, KPM_DOTS_DOUBLING, 2,
                                         LP >= 32 ? 0 : kDefaultPrefetch>,
                             kWaves * 64, 0),
                         "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    }
    return check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         out, cheb_kernel<T, V, LP, kWaves, KPM_DOTS_DOUBLING, 1, 0>, kWaves * 64,
                         0),
                     "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  }
};

}  // namespace

int blocks_per_cu(int value_type, const KernelShape& ks, int* out) {
  OccQuery q{out, ks.unroll, ks.waves};
  if (value_type == KPM_C64) return dispatch_shape<float>(ks.vec, ks.lpr, q);
  if (value_type == KPM_C128) return dispatch_shape<double>(ks.vec, ks.lpr, q);
  return fail(KPM_ERR_ARG, "unknown value type");
}

namespace {

struct ChebLauncher {
  int mode;
  bool only_dots;
  int iters;
  int n;
  int unroll;
  int pf;
  int waves;
  const StepParams& p;
  hipStream_t s;
  template <typename T, int V, int LP>
  int operator()() {
    return launch_one<T, V, LP>(mode, only_dots, iters, n, unroll, pf, waves, p, s);
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

int prefetch_depth(int32_t row_nnz_hint) {
  if (row_nnz_hint <= 0) return kDefaultPrefetch;
  for (int pf : kPrefetchDepths)
    if (row_nnz_hint <= pf) return pf;
  return kPrefetchDepths[sizeof(kPrefetchDepths) / sizeof(int) - 1];
}

int launch_cheb(int value_type, const KernelShape& ks, int mode, bool only_dots, int n,
                int row_nnz_hint, const StepParams& p, hipStream_t stream) {
  ChebLauncher l{mode,   only_dots, ks.iters, n, ks.unroll, prefetch_depth(row_nnz_hint),
                 ks.waves, p,         stream};
  if (value_type == KPM_C64) return dispatch_shape<float>(ks.vec, ks.lpr, l);
  if (value_type == KPM_C128) return dispatch_shape<double>(ks.vec, ks.lpr, l);
  return fail(KPM_ERR_ARG, "unknown value type");
}

int launch_reduce(int mode, int width, const StepParams& p, hipStream_t stream) {
  if (mode == KPM_DOTS_DOUBLING) return launch_reduce_mode<KPM_DOTS_DOUBLING>(p, width, stream);
  if (mode == KPM_DOTS_PROJECT) return launch_reduce_mode<KPM_DOTS_PROJECT>(p, width, stream);
  return KPM_OK;
}

namespace {

struct BsrLauncher {
  int mode;
  int iters;
  int n;
  const kpm_bsr* h;
  const StepParams& p;
  hipStream_t s;
  template <typename T, int V, int LP>
  int operator()() {
    if constexpr (LP > 16) {
      return fail(KPM_ERR_UNSUPPORTED, "block-CSR steps need at most 16 lanes per row");
    } else {
      dim3 grid(p.nblocks_launch), block(kWaves * 64);
      if (mode == KPM_DOTS_DOUBLING)
        hipLaunchKernelGGL((bsr_kernel<T, V, LP, kWaves, KPM_DOTS_DOUBLING>), grid, block, 0, s,
                           p, h->indptr, h->indices, h->values, n, iters);
      else
        hipLaunchKernelGGL((bsr_kernel<T, V, LP, kWaves, KPM_DOTS_NONE>), grid, block, 0, s, p,
                           h->indptr, h->indices, h->values, n, iters);
      return check_hip(hipGetLastError(), "bsr_kernel launch");
    }
  }
};

}  // namespace

int launch_bsr(const KernelShape& ks, int mode, int n, const kpm_bsr* h, const StepParams& p,
               hipStream_t stream) {
  if (ks.waves != kWaves) return fail(KPM_ERR_UNSUPPORTED, "block-CSR steps run 4-wave shapes");
  if (mode != KPM_DOTS_DOUBLING && mode != KPM_DOTS_NONE)
    return fail(KPM_ERR_ARG, "block-CSR steps take KPM_DOTS_DOUBLING or KPM_DOTS_NONE");
  BsrLauncher l{mode, ks.iters, n, h, p, stream};
  if (h->value_type == KPM_C64) return dispatch_shape<float>(ks.vec, ks.lpr, l);
  if (h->value_type == KPM_C128) return dispatch_shape<double>(ks.vec, ks.lpr, l);
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

int launch_bond(int width, const kpm_csr* b, const void* bra, const void* ket, const void* v,
                const void* w, int per_column, double* out, hipStream_t stream) {
  // lanes per row at most 8 (three butterfly levels), the rest of the row in
  // G column groups per lane
#define KPM_BOND(T, VEC, LPR, G) \
  return launch_bond_one<T, VEC, LPR, G>(b, bra, ket, v, w, per_column, out, stream)
  if (b->value_type == KPM_C64) {
    switch (width) {
      case 1: KPM_BOND(float, 1, 1, 1);
      case 2: KPM_BOND(float, 2, 1, 1);
      case 4: KPM_BOND(float, 2, 2, 1);
      case 8: KPM_BOND(float, 2, 4, 1);
      case 16: KPM_BOND(float, 2, 8, 1);
      case 32: KPM_BOND(float, 2, 8, 2);
      case 64: KPM_BOND(float, 2, 8, 4);
      default: break;
    }
  } else if (b->value_type == KPM_C128) {
    switch (width) {
      case 1: KPM_BOND(double, 1, 1, 1);
      case 2: KPM_BOND(double, 1, 2, 1);
      case 4: KPM_BOND(double, 1, 4, 1);
      case 8: KPM_BOND(double, 1, 8, 1);
      case 16: KPM_BOND(double, 1, 8, 2);
      case 32: KPM_BOND(double, 1, 8, 4);
      case 64: KPM_BOND(double, 1, 8, 8);
      default: break;
    }
  } else {
    return fail(KPM_ERR_ARG, "unknown value type");
  }
#undef KPM_BOND
  return fail(KPM_ERR_UNSUPPORTED, "bond width must be a power of two <= 64");
}

}  // namespace kpm
