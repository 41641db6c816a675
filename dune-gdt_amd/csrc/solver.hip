// solver.hip -- operator application and Jacobi-preconditioned CG on gfx950 (FP64).
//
// Replaces ConstMatrixOperator::apply (matrix_.mv, dune/gdt/operators/matrix-based.hh:121-129)
// and apply_inverse (XT::LA::Solver, matrix-based.hh:148-159).
//
// Storage: the assembled CSR matrix is converted once per assembly into
//   * DIA  -- one value array per diagonal, no column indices, when every problem instance
//             has <= 32 distinct (col - row) offsets (structured grids).  When the matrix is
//             exactly symmetric (checked bit for bit on the device after every assembly) only
//             the diagonals with offset >= 0 are read: A(i, i-o) = A(i-o, i) is taken from the
//             upper diagonal at row i-o, which halves the bytes read per application;
//   * SELL-64 -- sliced ELLPACK with slices of 64 rows (one wavefront), column-major inside a
//             slice so every load of a wave is one contiguous 512-byte segment.
// One lane owns a row and sums it in ascending column order (the CSR order), so DIA, symmetric
// DIA and SELL give the same bits for the same matrix.
//
// Jacobi-PCG runs as CG on the symmetrically scaled matrix D^{-1/2} A D^{-1/2} (same iterates in
// exact arithmetic; the scaled matrix has a unit diagonal that is never stored or read, and the
// preconditioner costs no vector), three launches per iteration, each one pass over its vectors:
//   k_cg_direction : stopping test from the previous update's partials; p = r + beta p into the other
//                    of two p buffers; every second step x += the two updates deferred since
//   k_cg_spmv      : q = Ah p and the block partials of p.q (k_cg_spmv_win: the same with the p entries
//                    of a row block staged in LDS, for operators larger than the last-level cache;
//                    k_cg_spmv_blk: 4 x 4 element blocks)
//   k_cg_update    : alpha from the p.q partials; r -= alpha q; block partials of r.r and |D^{1/2} r|^2
// Data cross blocks only at kernel boundaries.  Every block of an instance reduces that
// instance's block partials itself, in block order, with a fixed butterfly / wave order, so all
// blocks obtain bitwise identical alpha and beta and the result does not depend on scheduling;
// no fence or atomic is used.  The per-instance scalars live in two records selected by the
// parity of the iteration: iteration j's direction kernel reads record j%2 and writes record
// (j+1)%2, which its update kernel then completes.  Every block belongs to exactly one problem instance, so a batch of independent
// problems (a union grid) is solved in the same launches with per-instance scalars and stopping
// tests; blocks of finished instances exit at once.
//
// Launch geometry: a block owns R = block * items consecutive rows, each lane `items` rows
// strided by the block size (coalesced, all loads of a lane issued before use).  Hardware
// workgroups are dealt round-robin to the 8 XCDs; logical block L is given to hardware block
// (L % q) * 8 + L / q (q = ceil(nblocks / 8)) so every XCD streams one contiguous eighth of the
// rows and the stencil neighbours of its rows stay in its own L2.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>

#include <hip/hip_ext.h>

#include "hgdt_internal.hpp"
#include "block_reduce.hpp"

namespace hgdt {
namespace {

constexpr int MAXDIAG = 32;
enum { FMT_DIA = HGDT_FORMAT_DIA, FMT_SELL = HGDT_FORMAT_SELL, FMT_DIAS = 3, FMT_BLK = HGDT_FORMAT_BLOCK };

// column windows of an instance's symmetric DIA stencil: the offsets 0 and +-o of its strict upper
// diagonals, grouped into at most WIN_MAX ranges [lo, hi] (relative to the row) no wider than the row
// block; offsets in no range never meet a column of the instance
constexpr int WIN_MAX = 4, WIN_MAXU = 8;
constexpr int64_t WIN_LDS = 64 * 1024;
// automatic choice: staged when the application moves more than the 256 MiB last-level cache (MALL) of the
// GPU holds; below that, the direct kernel, whose cache-line requests the MALL then serves, is faster
constexpr int64_t WIN_AUTO_BYTES = int64_t(256) << 20;
struct DiaWindows {
  int32_t nwin;
  int32_t span[WIN_MAX];  // hi_w - lo_w
  int32_t cspan[WIN_MAX]; // sum of the spans of the windows before w
  int32_t zero_w, zero_rel; // the window holding offset 0, and its position in it (cspan_w + 0 - lo_w)
  // the window of p(i - o_u) and p(i + o_u) of strict upper slot u (-1: no column of the instance) and its
  // position cspan_w + (+-o_u - lo_w): LDS index = lane + w R + position for blocks of R rows
  int32_t dn_w[WIN_MAXU], dn_rel[WIN_MAXU], up_w[WIN_MAXU], up_rel[WIN_MAXU];
  int64_t lo[WIN_MAX];
};

struct PcgParams {
  int64_t n;
  // A: row partition of the operator application
  int a_nb;
  const int64_t* a_row0;
  const int32_t* a_inst;
  const int32_t* a_ib0; // first A-block of every instance
  // V: row partition of the vector kernels
  int v_nb;
  const int64_t* v_row0;
  const int32_t* v_inst;
  const int32_t* v_ib0;
  // blocks of unfinished instances (NULL: every block); a_nb / v_nb are then the list lengths
  const int32_t* a_list;
  const int32_t* v_list;
  // the blocks of each XCD are walked top down (set per launch: successive kernels of an iteration walk
  // in opposite orders, so each starts on the rows the previous one touched last)
  int rev;
  const int64_t* inst_row0;
  // DIA
  int ndiag;
  const int64_t* dia_off; // ninst * ndiag, ascending, sentinel n for unused slots
  const int32_t* dia_k0;  // ninst: index of offset 0
  const double* dia;      // ndiag * n
  // the solver's scaled copy of a symmetric DIA matrix: the strict upper diagonals with a nonzero entry,
  // s_nu slots (values scaled[u * n + row]), offsets s_off[inst * s_nu + u] ascending (sentinel n: padding)
  int s_nu;
  const int64_t* s_off;
  // SELL
  const int64_t* block_slice0; // per A-block
  const int64_t* slice_ptr;
  const int32_t* sell_cols;
  const double* sell_vals;
  // 4 x 4 element blocks (value planes of blk_ne elements, blk_mn neighbour slots)
  int64_t blk_ne;
  int blk_mn;
  const int32_t* blk_nb;
  const int32_t* blk_mir;
  const double* blk;
  // symmetric scaling D^{-1/2} A D^{-1/2}: values in the layout of the format, h = D^{-1/2}, s = D^{1/2}
  const double* scaled;
  const double* dh;
  const double* dsq;
  // vectors of the scaled system (x and r scaled)
  const double* b;
  const double* ax;
  double* x;
  double* r;
  double* p;
  double* p2; // second p buffer: the p formed at step j is in p2 for even j, in p for odd j
  double* q;
  double* part1; // 1 per A-block: p.q
  double* part2; // 2 per V-block: r.r of the scaled residual, r.r of the residual
  double* part3; // 1 per V-block: b.b (initialisation)
  PcgScalars* S; // 2 * ninst records
  int ninst;
  double rtol2;
  int64_t max_iter;
  int zero_guess;
  const DiaWindows* win; // ninst: column windows of the DIA stencils (p staged in LDS)
};

// y_i = sum_j A_ij v(j) for the ITEMS rows r0 + threadIdx.x + j * blockDim.x (< r1) of block b, whose rows
// start at rb (a block may process its rows in several chunks starting at r0).  `vals` is the
// value storage of the format (DIA diagonals or SELL values); v(c) returns the vector entry at column c.
// UNIT: the diagonal entry is 1 and not read (the symmetrically scaled matrix of the CG solver).
template <int FMT, int ITEMS, bool UNIT, class VSRC>
__device__ __forceinline__ void apply_rows(const PcgParams& P, const double* __restrict__ vals, int b, int inst,
                                           int64_t rb, int64_t r0, int64_t r1, VSRC v, double (&acc)[ITEMS])
{
  const int64_t B = blockDim.x;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j)
    acc[j] = 0.;
  if constexpr (FMT == FMT_DIA || FMT == FMT_DIAS) {
    const int64_t lo = P.inst_row0[inst], hi = P.inst_row0[inst + 1];
    const int64_t* __restrict__ off = P.dia_off + int64_t(inst) * P.ndiag;
    const double* __restrict__ dia = vals;
    const int64_t n = P.n;
    const int k0 = P.dia_k0[inst];
    if constexpr (FMT == FMT_DIA) {
#pragma unroll 3
      for (int k = 0; k < P.ndiag; ++k) {
        const int64_t o = off[k];
        const double* __restrict__ dk = dia + int64_t(k) * n;
        const bool unit = UNIT && (k == k0);
        double a[ITEMS], x[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
          const int64_t i = r0 + threadIdx.x + j * B;
          const int64_t c = i + o;
          const bool ok = (i < r1) && (c >= lo) && (c < hi);
          a[j] = (ok && !unit) ? dk[i] : 1.;
          x[j] = ok ? v(c) : 0.;
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
          acc[j] = unit ? acc[j] + x[j] : fma(a[j], x[j], acc[j]);
      }
    } else {
      // lower triangle from the upper diagonals (descending offset = ascending column), then upper.
      // Strict upper slot u has offset uo[u] and its values at ub[u * n + row].  The assembled matrix's
      // storage (!UNIT) holds every diagonal of the pattern; the solver's scaled copy (UNIT) holds only the
      // strict upper diagonals with a nonzero entry in the instance (s_nu slots, offsets s_off), which
      // leaves out FMAs that would add 0 * x: the same sums.
      const int ku = UNIT ? P.s_nu : P.ndiag - 1 - k0; // (trailing slots may be padding: offset n)
      const int64_t* __restrict__ uo = UNIT ? P.s_off + int64_t(inst) * P.s_nu : off + k0 + 1;
      const double* __restrict__ ub = UNIT ? dia : dia + int64_t(k0 + 1) * n;
      if constexpr (ITEMS == 1) {
        // the loads of up to 8 diagonals are issued before their FMAs (few memory round trips per
        // row); the FMAs run in the order of the loops below, so the bits are the same
        constexpr int MU = 8;
        const int64_t i = r0 + threadIdx.x;
        const bool in = i < r1;
        double s = 0.;
        for (int top = ku - 1; top >= 0; top -= MU) { // lower part, descending offset
          double a[MU], x[MU];
#pragma unroll
          for (int t = 0; t < MU; ++t) {
            const int u = top - t;
            a[t] = x[t] = 0.;
            if (u >= 0) {
              const int64_t cl = i - uo[u];
              if (in && cl >= lo && cl < hi) {
                a[t] = ub[int64_t(u) * n + cl];
                x[t] = v(cl);
              }
            }
          }
#pragma unroll
          for (int t = 0; t < MU; ++t)
            if (top - t >= 0)
              s = fma(a[t], x[t], s);
        }
        if constexpr (UNIT) {
          if (in)
            s += v(i);
        } else {
          const double ad = in ? dia[int64_t(k0) * n + i] : 0.;
          const double xd = in ? v(i) : 0.;
          s = fma(ad, xd, s);
        }
        for (int bot = 0; bot < ku; bot += MU) { // upper part, ascending offset
          double a[MU], x[MU];
#pragma unroll
          for (int t = 0; t < MU; ++t) {
            const int u = bot + t;
            a[t] = x[t] = 0.;
            if (u < ku) {
              const int64_t cu = i + uo[u];
              if (in && cu >= lo && cu < hi) {
                a[t] = ub[int64_t(u) * n + i];
                x[t] = v(cu);
              }
            }
          }
#pragma unroll
          for (int t = 0; t < MU; ++t)
            if (bot + t < ku)
              s = fma(a[t], x[t], s);
        }
        acc[0] = s;
        return;
      }
#pragma unroll 2
      for (int u = ku - 1; u >= 0; --u) {
        const int64_t o = uo[u];
        const double* __restrict__ dk = ub + int64_t(u) * n;
        double a[ITEMS], x[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
          const int64_t i = r0 + threadIdx.x + j * B;
          const int64_t c = i - o;
          const bool ok = (i < r1) && (c >= lo) && (c < hi);
          a[j] = ok ? dk[c] : 0.;
          x[j] = ok ? v(c) : 0.;
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
          acc[j] = fma(a[j], x[j], acc[j]);
      }
      // the diagonal, between the lower and the upper part (unit in the solver's scaled copy)
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const int64_t i = r0 + threadIdx.x + j * B;
        if (i < r1) {
          if constexpr (UNIT)
            acc[j] += v(i);
          else
            acc[j] = fma(dia[int64_t(k0) * n + i], v(i), acc[j]);
        }
      }
#pragma unroll 2
      for (int u = 0; u < ku; ++u) {
        const int64_t o = uo[u];
        const double* __restrict__ dk = ub + int64_t(u) * n;
        double a[ITEMS], x[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
          const int64_t i = r0 + threadIdx.x + j * B;
          const int64_t c = i + o;
          const bool ok = (i < r1) && (c >= lo) && (c < hi);
          a[j] = ok ? dk[i] : 0.;
          x[j] = ok ? v(c) : 0.;
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
          acc[j] = fma(a[j], x[j], acc[j]);
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const int64_t i = r0 + threadIdx.x + j * B;
      if (i < r1) {
        const int64_t local = i - rb;
        const int64_t sl = P.block_slice0[b] + (local >> 6);
        const int lane = int(local & 63);
        const int64_t s0 = P.slice_ptr[sl];
        const int w = int((P.slice_ptr[sl + 1] - s0) >> 6);
        const int32_t* __restrict__ cc = P.sell_cols + s0 + lane;
        const double* __restrict__ vv = vals + s0 + lane;
        double s = 0.;
        for (int k = 0; k < w; ++k)
          s = fma(vv[int64_t(k) * 64], v(cc[int64_t(k) * 64]), s);
        acc[j] = s;
      }
    }
  }
}

template <int FMT, int ITEMS>
__global__ void __launch_bounds__(1024) k_apply(PcgParams P, const double* __restrict__ x, double* __restrict__ y)
{
  const int b = logical_block(P.a_nb);
  if (b < 0)
    return;
  const int inst = P.a_inst[b];
  const int64_t rb = P.a_row0[b], r1 = P.a_row0[b + 1];
  const double* vals = (FMT == FMT_SELL) ? P.sell_vals : P.dia;
  for (int64_t r0 = rb; r0 < r1; r0 += int64_t(blockDim.x) * ITEMS) {
    double acc[ITEMS];
    apply_rows<FMT, ITEMS, false>(P, vals, b, inst, rb, r0, r1, [&](int64_t c) { return x[c]; }, acc);
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
      if (i < r1)
        y[i] = acc[j];
    }
  }
}

// 4 x 4 element blocks: y = A x over the 4 rows of element E (lane per element).  Each row sums its
// terms in ascending column order (neighbours below E, the own block, neighbours above E), the order of
// DIA and SELL, so the formats give the same bits.  SYM (exactly symmetric matrix): only the own
// block's upper triangle and the blocks of the neighbours above E are read from E's record; the block
// with a neighbour F below E is F's block with E (read from F's record, transposed).  UNIT: the unit
// diagonal of the scaled matrix, not read.
template <bool SYM, bool UNIT, class XS>
__device__ __forceinline__ void blk_rows_seq(const PcgParams& P, const double* __restrict__ v, int64_t E, XS xs,
                                         double (&y)[4])
{
  const int64_t ne = P.blk_ne;
  const int mn = P.blk_mn;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    y[i] = 0.;
  int s = 0;
  for (; s < mn; ++s) { // neighbours below E
    const int32_t F = P.blk_nb[int64_t(s) * ne + E];
    if (F < 0 || F > E)
      break;
    const double* __restrict__ B =
        SYM ? v + int64_t(16 + 16 * P.blk_mir[int64_t(s) * ne + E]) * ne + F : v + int64_t(16 + 16 * s) * ne + E;
    double b[16], x[4];
#pragma unroll
    for (int k = 0; k < 16; ++k)
      b[k] = B[int64_t(k) * ne];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      x[j] = xs(4 * int64_t(F) + j);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        y[i] = fma(SYM ? b[4 * j + i] : b[4 * i + j], x[j], y[i]);
  }
  { // own block
    double o[16], x[4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c)
        o[4 * a + c] = ((SYM && c < a) || (UNIT && c == a)) ? 0. : v[int64_t(4 * a + c) * ne + E];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      x[j] = xs(4 * E + j);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (UNIT && j == i)
          y[i] += x[j];
        else
          y[i] = fma((SYM && j < i) ? o[4 * j + i] : o[4 * i + j], x[j], y[i]);
      }
  }
  for (; s < mn; ++s) { // neighbours above E
    const int32_t F = P.blk_nb[int64_t(s) * ne + E];
    if (F < 0)
      break;
    const double* __restrict__ B = v + int64_t(16 + 16 * s) * ne + E;
    double b[16], x[4];
#pragma unroll
    for (int k = 0; k < 16; ++k)
      b[k] = B[int64_t(k) * ne];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      x[j] = xs(4 * int64_t(F) + j);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        y[i] = fma(b[4 * i + j], x[j], y[i]);
  }
}

// the same sums for elements with at most MAXS neighbours: the slots' neighbours are read first, and the
// blocks are loaded two at a time (loads of two blocks in flight per lane; the FMAs keep the order)
template <int MAXS, bool SYM, bool UNIT, class XS>
__device__ __forceinline__ void blk_rows_pairs(const PcgParams& P, const double* __restrict__ v, int64_t E, XS xs,
                                               double (&y)[4])
{
  const int64_t ne = P.blk_ne;
  const int mn = P.blk_mn;
  int32_t Fs[MAXS], ts[MAXS];
#pragma unroll
  for (int s = 0; s < MAXS; ++s) {
    Fs[s] = s < mn ? P.blk_nb[int64_t(s) * ne + E] : -1;
    ts[s] = (SYM && s < mn) ? P.blk_mir[int64_t(s) * ne + E] : 0;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
    y[i] = 0.;
  // the own block's loads are issued first, so they are in flight with the first neighbour blocks
  double o[16], xo[4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      o[4 * a + c] = ((SYM && c < a) || (UNIT && c == a)) ? 0. : v[int64_t(4 * a + c) * ne + E];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    xo[j] = xs(4 * E + j);
  // neighbours below E (a prefix of the slots), two blocks at a time
#pragma unroll
  for (int s0 = 0; s0 < MAXS; s0 += 2) {
    double b[2][16], x[2][4];
    bool on[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int s = s0 + g;
      on[g] = s < MAXS && Fs[s] >= 0 && Fs[s] < E;
      if (on[g]) {
        const double* __restrict__ B =
            SYM ? v + int64_t(16 + 16 * ts[s]) * ne + Fs[s] : v + int64_t(16 + 16 * s) * ne + E;
#pragma unroll
        for (int k = 0; k < 16; ++k)
          b[g][k] = B[int64_t(k) * ne];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          x[g][j] = xs(4 * int64_t(Fs[s]) + j);
      }
    }
#pragma unroll
    for (int g = 0; g < 2; ++g)
      if (on[g])
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            y[i] = fma(SYM ? b[g][4 * j + i] : b[g][4 * i + j], x[g][j], y[i]);
  }
  { // own block (loaded before the neighbours below)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (UNIT && j == i)
          y[i] += xo[j];
        else
          y[i] = fma((SYM && j < i) ? o[4 * j + i] : o[4 * i + j], xo[j], y[i]);
      }
  }
  // neighbours above E, two blocks at a time
#pragma unroll
  for (int s0 = 0; s0 < MAXS; s0 += 2) {
    double b[2][16], x[2][4];
    bool on[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int s = s0 + g;
      on[g] = s < MAXS && Fs[s] > E;
      if (on[g]) {
        const double* __restrict__ B = v + int64_t(16 + 16 * s) * ne + E;
#pragma unroll
        for (int k = 0; k < 16; ++k)
          b[g][k] = B[int64_t(k) * ne];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          x[g][j] = xs(4 * int64_t(Fs[s]) + j);
      }
    }
#pragma unroll
    for (int g = 0; g < 2; ++g)
      if (on[g])
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            y[i] = fma(b[g][4 * i + j], x[g][j], y[i]);
  }
}

template <bool SYM, bool UNIT, class XS>
__device__ __forceinline__ void blk_rows(const PcgParams& P, const double* __restrict__ v, int64_t E, XS xs,
                                         double (&y)[4])
{
  if (P.blk_mn <= 4)
    blk_rows_pairs<4, SYM, UNIT>(P, v, E, xs, y);
  else
    blk_rows_seq<SYM, UNIT>(P, v, E, xs, y);
}

// y = A x with element blocks (unscaled values), one lane per element (XCD-aware 1-d launch of 256 lanes)
template <bool SYM>
__global__ void __launch_bounds__(256) k_blk_apply(PcgParams P, const double* __restrict__ x, double* __restrict__ y)
{
  const int64_t E = xcd_index(P.blk_ne);
  if (E >= P.blk_ne)
    return;
  const int64_t i0 = 4 * E;
  double acc[4];
  blk_rows<SYM, false>(P, P.blk, E, [&](int64_t c) { return x[c]; }, acc);
#pragma unroll
  for (int i = 0; i < 4; ++i)
    y[i0 + i] = acc[i];
}

// diagonal scaling and exact-symmetry test of the element blocks: h = D^{-1/2}, s = D^{1/2} from the own
// blocks' diagonals (flags[0]: a non-positive diagonal); flags[1]: an entry differs from its transposed
// one (own block, or the block with a neighbour below against that neighbour's block)
__global__ void k_blk_prepare(PcgParams P, double* __restrict__ dh, double* __restrict__ dsq, int* flags)
{
  const int64_t E = xcd_index(P.blk_ne);
  if (E >= P.blk_ne)
    return;
  const int64_t ne = P.blk_ne;
  const double* __restrict__ v = P.blk;
  bool ok = true;
  for (int a = 0; a < 4; ++a) {
    double d = v[int64_t(5 * a) * ne + E];
    if (!(d > 0.)) {
      flags[0] = 1;
      d = 1.;
    }
    const double sq = sqrt(d);
    dh[4 * E + a] = 1. / sq;
    dsq[4 * E + a] = sq;
    for (int c = a + 1; c < 4; ++c)
      ok = ok && v[int64_t(4 * a + c) * ne + E] == v[int64_t(4 * c + a) * ne + E];
  }
  for (int s = 0; s < P.blk_mn; ++s) {
    const int32_t F = P.blk_nb[int64_t(s) * ne + E];
    if (F < 0 || F > E)
      break;
    const int t = P.blk_mir[int64_t(s) * ne + E];
    for (int k = 0; k < 16; ++k)
      ok = ok && v[int64_t(16 + 16 * s + k) * ne + E] == v[int64_t(16 + 16 * t + 4 * (k % 4) + k / 4) * ne + F];
  }
  if (!ok)
    flags[1] = 1;
}

// scaled element blocks: Ah(i, j) = a_ij (h_i h_j); slots past the last neighbour hold 0
__global__ void k_blk_scale(PcgParams P, const double* __restrict__ dh, double* __restrict__ dst)
{
  const int64_t E = xcd_index(P.blk_ne);
  if (E >= P.blk_ne)
    return;
  const int64_t ne = P.blk_ne;
  const double* __restrict__ v = P.blk;
  for (int k = 0; k < 16; ++k)
    dst[int64_t(k) * ne + E] = v[int64_t(k) * ne + E] * (dh[4 * E + k / 4] * dh[4 * E + k % 4]);
  for (int s = 0; s < P.blk_mn; ++s) {
    const int32_t F = P.blk_nb[int64_t(s) * ne + E];
    for (int k = 0; k < 16; ++k) {
      const int64_t m = int64_t(16 + 16 * s + k) * ne + E;
      dst[m] = F < 0 ? 0. : v[m] * (dh[4 * E + k / 4] * dh[4 * int64_t(F) + k % 4]);
    }
  }
}

// Jacobi-preconditioned CG run as CG on the symmetrically scaled system
//   Ah = D^{-1/2} A D^{-1/2},  xh = D^{1/2} x,  rh = D^{-1/2} r,
// which in exact arithmetic has the iterates of Jacobi-PCG (z = D^{-1} r, rh.rh = r.z,
// p = D^{-1/2} ph, the same alpha and beta).  The unpreconditioned residual for the stopping test is
// r = D^{1/2} rh.
//
// initialisation: r = b - A x0, rh = D^{-1/2} r, xh = D^{1/2} x0 (0 for a zero guess);
// block partials of rh.rh and r.r (the form step 0 reduces) and of b.b
template <int ITEMS>
__global__ void __launch_bounds__(1024) k_cg_init(PcgParams P)
{
  const int b = logical_block(P.v_nb);
  if (b < 0)
    return;
  const int64_t r0 = P.v_row0[b], r1 = P.v_row0[b + 1];
  double v[3] = {0., 0., 0.};
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    if (i < r1) {
      const double bi = P.b[i];
      const double ri = P.ax ? bi - P.ax[i] : bi;
      const double rh = P.dh[i] * ri;
      P.r[i] = rh;
      P.x[i] = P.zero_guess ? 0. : P.dsq[i] * P.x[i];
      v[0] += rh * rh, v[1] += ri * ri, v[2] += bi * bi;
    }
  }
  block_sum<3>(v);
  if (threadIdx.x == 0) {
    P.part2[int64_t(b) * 2 + 0] = v[0];
    P.part2[int64_t(b) * 2 + 1] = v[1];
    P.part3[b] = v[2];
  }
}

// one block per instance: record 0 (tolerance from b.b, no iteration done yet)
__global__ void __launch_bounds__(256) k_cg_init_scalars(PcgParams P)
{
  const int inst = blockIdx.x;
  double t[1];
  reduce_partials<1>(P.part3, P.v_ib0[inst], P.v_ib0[inst + 1], t);
  if (threadIdx.x == 0) {
    PcgScalars S{};
    S.bb = t[0];
    S.tol2 = P.rtol2 * t[0];
    S.iters = 0;
    S.done = 0;
    P.S[inst] = S;
  }
}

// CG direction update of step j (par = j & 1), on the vector row partition:
//   rho_j = rh.rh and |r|^2 from the partials of the previous update (or of the initialisation),
//   stopping test; p_j = rh + beta p_{j-1} with beta = rho_j / rho_{j-1} (p_0 = rh).
// p_j is written to p2 for even j and to p for odd j, so p_{j-1} and p_{j-2} are both at hand, and the
// iterate is updated every second step: at even j >= 2, xh += alpha_{j-2} p_{j-2}, then += alpha_{j-1}
// p_{j-1} (the operations of one update per step, in the same order), and at the step that stops the
// instance, whatever is still pending (the update kernel does it after a breakdown).
// Reads record par, writes record par ^ 1.
template <int ITEMS>
__global__ void __launch_bounds__(1024) k_cg_direction(PcgParams P, int par)
{
  const int b = active_block(P.v_nb, P.v_list, P.rev);
  if (b < 0)
    return;
  const int64_t r0 = P.v_row0[b], r1 = P.v_row0[b + 1];
  const double* __restrict__ pin = par ? P.p2 : P.p; // p_{j-1}
  double* __restrict__ pout = par ? P.p : P.p2;      // p_j, over p_{j-2}
  // every load that needs only the block's rows or record is issued before any is waited for: the rows
  // of r and p_{j-1} (p_{j-1} also at the first step, where it is not used), the scalars, the partials of
  // the previous update and, on pair steps, the iterate's rows and p_{j-2}
  double pv[ITEMS], rv[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    const bool ok = i < r1;
    pv[j] = ok ? pin[i] : 0.;
    rv[j] = ok ? P.r[i] : 0.;
  }
  const int inst = P.v_inst[b];
  const PcgScalars cur = P.S[par * P.ninst + inst];
  PcgScalars* nxt = P.S + (par ^ 1) * P.ninst + inst;
  const bool writer = (b == P.v_ib0[inst] && threadIdx.x == 0);
  const int done = pcg_state(cur);
  Partials<2> pt;
  pt.issue(P.part2, P.v_ib0[inst], done ? P.v_ib0[inst] : P.v_ib0[inst + 1]);
  const int64_t it = cur.iters;
  const bool first = (it == 0);
  const bool pair = !first && (it & 1) == 0; // both pending updates of the iterate are due
  double ov[ITEMS], xv[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    const bool ok = i < r1 && pair && !done;
    ov[j] = ok ? pout[i] : 0.;
    xv[j] = ok ? P.x[i] : 0.;
  }
  if (done) {
    if (writer) {
      *nxt = cur;
      nxt->done = done;
    }
    return;
  }
  const double rho_prev = cur.rz, alpha = cur.alpha, alpha_prev = cur.alpha_prev, tol2 = cur.tol2;
  double t[2];
  pt.sum(P.part2, t);
  const int status = (t[1] <= tol2) ? 1 : (it >= P.max_iter) ? 2 : 0;
  const double beta = (status == 0 && !first) ? t[0] / rho_prev : 0.;
  if (!first) {
    if (!pair && status != 0) // the instance stops at an odd step: its one pending update
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
        xv[j] = i < r1 ? P.x[i] : 0.;
      }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
      if (i < r1) {
        if (pair)
          P.x[i] = fma(alpha, pv[j], fma(alpha_prev, ov[j], xv[j]));
        else if (status != 0)
          P.x[i] = fma(alpha, pv[j], xv[j]);
        if (status == 0)
          pout[i] = fma(beta, pv[j], rv[j]);
      }
    }
  } else if (status == 0) {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
      if (i < r1)
        pout[i] = rv[j];
    }
  }
  if (writer) {
    nxt->bb = cur.bb;
    nxt->tol2 = tol2;
    nxt->alpha = alpha;
    nxt->alpha_prev = alpha;
    nxt->pq = cur.pq;
    nxt->rr = t[1];
    nxt->rz = t[0];
    nxt->beta = beta;
    nxt->iters = status ? it : it + 1;
    nxt->done = status;
    nxt->brk = 0;
  }
}

// q = Ah p and the block partials of p.q, on the operator's row partition (record par ^ 1, read only)
template <int FMT, int ITEMS>
__global__ void __launch_bounds__(1024) k_cg_spmv(PcgParams P, int par)
{
  const int b = active_block(P.a_nb, P.a_list, P.rev);
  if (b < 0)
    return;
  // a single instance is instance 0: its offsets and bounds are then read without waiting for a_inst
  const int inst = P.ninst == 1 ? 0 : P.a_inst[b];
  const int64_t rb = P.a_row0[b], r1 = P.a_row0[b + 1];
  const double* __restrict__ p = par ? P.p : P.p2; // p_j (written by the direction update of step j)
  // the stopping flag is read with the block's first rows and tested before the first store, so its load
  // does not delay theirs (a block of a finished instance reads its rows and discards them)
  const int done = P.S[(par ^ 1) * P.ninst + inst].done;
  double v[1] = {0.};
  for (int64_t r0 = rb; r0 < r1; r0 += int64_t(blockDim.x) * ITEMS) {
    double acc[ITEMS];
    apply_rows<FMT, ITEMS, FMT != FMT_SELL>(P, P.scaled, b, inst, rb, r0, r1, [&](int64_t c) { return p[c]; },
                                            acc);
    if (done)
      return;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
      if (i < r1) {
        P.q[i] = acc[j];
        v[0] += p[i] * acc[j];
      }
    }
  }
  block_sum<1>(v);
  if (threadIdx.x == 0)
    P.part1[b] = v[0];
}

// q = Ah p and the block partials of p.q with symmetric DIA storage, one row per lane (a block's rows at
// most blockDim.x), at most WIN_MAXU strict upper diagonals, and the p entries of the block's rows staged
// in LDS: every column window row + [lo_w, hi_w] of the instance is loaded once per block (coalesced),
// and each row reads its 2 ku + 1 entries of p from LDS instead of from the cache hierarchy.  The FMAs
// run in the order of the direct kernel (lower part by descending offset, the unit diagonal, the upper
// part by ascending offset), so q and the partials have the same bits.
__global__ void __launch_bounds__(1024) k_cg_spmv_win(PcgParams P, int par)
{
  extern __shared__ double win[];
  const int b = active_block(P.a_nb, P.a_list, P.rev);
  if (b < 0)
    return;
  const int inst = P.a_inst[b];
  if (P.S[(par ^ 1) * P.ninst + inst].done)
    return;
  const int64_t rb = P.a_row0[b], r1 = P.a_row0[b + 1];
  const double* __restrict__ p = par ? P.p : P.p2; // p_j (written by the direction update of step j)
  const int64_t lo = P.inst_row0[inst], hi = P.inst_row0[inst + 1];
  const int64_t* __restrict__ off = P.s_off + int64_t(inst) * P.s_nu; // the scaled copy's strict upper slots
  const double* __restrict__ dia = P.scaled;
  const int64_t n = P.n;
  const int ku = P.s_nu; // <= WIN_MAXU (checked by the host)
  const DiaWindows& W = P.win[inst];
  const int t = threadIdx.x, R = blockDim.x;
  const int64_t i = rb + t;
  const bool in = i < r1;
  const int rows = int(r1 - rb);
  // the windows of p first, into registers (a window spans at most R + rows <= 2 R entries: two per
  // lane): window w holds p(rb + lo_w + s), s < rows + span_w, at w R + cspan_w + s (0 outside the
  // instance)
  double pw[WIN_MAX][2];
#pragma unroll
  for (int w = 0; w < WIN_MAX; ++w)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      pw[w][j] = 0.;
      if (w < W.nwin) {
        const int s = t + j * R;
        const int64_t g = rb + W.lo[w] + s;
        if (s < rows + W.span[w] && g >= lo && g < hi)
          pw[w][j] = p[g];
      }
    }
  // then the matrix values (the lower part from the mirrored rows), whose loads stay in flight across the
  // staging of p
  double al[WIN_MAXU], au[WIN_MAXU];
#pragma unroll
  for (int u = 0; u < WIN_MAXU; ++u) {
    al[u] = au[u] = 0.;
    if (u < ku) {
      const int64_t o = off[u];
      const double* __restrict__ du = dia + int64_t(u) * n;
      if (in && i - o >= lo)
        al[u] = du[i - o];
      if (in && i + o < hi)
        au[u] = du[i];
    }
  }
#pragma unroll
  for (int w = 0; w < WIN_MAX; ++w)
    if (w < W.nwin)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        if (t + j * R < rows + W.span[w])
          win[w * R + W.cspan[w] + t + j * R] = pw[w][j];
  __syncthreads();
  double xl[WIN_MAXU], xu[WIN_MAXU];
#pragma unroll
  for (int u = 0; u < WIN_MAXU; ++u) {
    xl[u] = (u < ku && W.dn_w[u] >= 0) ? win[t + W.dn_w[u] * R + W.dn_rel[u]] : 0.;
    xu[u] = (u < ku && W.up_w[u] >= 0) ? win[t + W.up_w[u] * R + W.up_rel[u]] : 0.;
  }
  const double pi = win[t + W.zero_w * R + W.zero_rel];
  double s = 0.;
#pragma unroll
  for (int u = WIN_MAXU - 1; u >= 0; --u) // lower part, descending offset
    if (u < ku)
      s = fma(al[u], xl[u], s);
  if (in)
    s += pi;
#pragma unroll
  for (int u = 0; u < WIN_MAXU; ++u) // upper part, ascending offset
    if (u < ku)
      s = fma(au[u], xu[u], s);
  double v[1] = {0.};
  if (in) {
    P.q[i] = s;
    v[0] += pi * s;
  }
  block_sum<1>(v);
  if (t == 0)
    P.part1[b] = v[0];
}

// q = Ah p and the block partials of p.q with element blocks (lane per element of the operator's row blocks)
template <bool SYM>
__global__ void __launch_bounds__(256) k_cg_spmv_blk(PcgParams P, int par)
{
  const int b = active_block(P.a_nb, P.a_list, P.rev);
  if (b < 0)
    return;
  const int inst = P.a_inst[b];
  if (P.S[(par ^ 1) * P.ninst + inst].done)
    return;
  const double* __restrict__ p = par ? P.p : P.p2; // p_j (written by the direction update of step j)
  const int64_t i0 = P.a_row0[b] + 4 * int64_t(threadIdx.x);
  double v[1] = {0.};
  if (i0 < P.a_row0[b + 1]) {
    double acc[4];
    blk_rows<SYM, true>(P, P.scaled, i0 / 4, [&](int64_t c) { return p[c]; }, acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      P.q[i0 + i] = acc[i];
      v[0] += p[i0 + i] * acc[i];
    }
  }
  block_sum<1>(v);
  if (threadIdx.x == 0)
    P.part1[b] = v[0];
}

// CG update of step j on the vector row partition: alpha = rho_j / p.q, rh -= alpha q, block partials
// of rh.rh and |D^{1/2} rh|^2.  Works on record par ^ 1 (written by the step kernel): stores alpha
// there for the iterate's deferred updates, or the breakdown flag `brk`.  Its blocks test only `done`,
// which this kernel never writes, so a block that starts after the writer block has flagged a breakdown
// still applies the pending half update of the iterate (the next direction update turns `brk` into
// done = 3).
template <int ITEMS>
__global__ void __launch_bounds__(1024) k_cg_update(PcgParams P, int par)
{
  const int b = active_block(P.v_nb, P.v_list, P.rev);
  if (b < 0)
    return;
  const int64_t r0 = P.v_row0[b], r1 = P.v_row0[b + 1];
  // the rows, the record (read whole, before any store: the writer block's stores below touch only alpha,
  // pq and brk, of which only alpha is read, and only on a breakdown, which no block writes alpha at) and
  // the partials of p.q are all issued before any is waited for
  double qv[ITEMS], rv[ITEMS], sv[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    const bool ok = i < r1;
    qv[j] = ok ? P.q[i] : 0.;
    rv[j] = ok ? P.r[i] : 0.;
    sv[j] = ok ? P.dsq[i] : 0.;
  }
  const int inst = P.v_inst[b];
  PcgScalars* S = P.S + (par ^ 1) * P.ninst + inst;
  const PcgScalars Sv = *S;
  Partials<1> pp;
  pp.issue(P.part1, P.a_ib0[inst], Sv.done ? P.a_ib0[inst] : P.a_ib0[inst + 1]);
  if (Sv.done)
    return;
  const bool writer = (b == P.v_ib0[inst] && threadIdx.x == 0);
  double pq[1];
  pp.sum(P.part1, pq);
  if (!(pq[0] > 0.)) {
    // breakdown (A not positive definite on the Krylov space): every block of the instance sees it.  At
    // an odd step j the direction update deferred alpha_{j-1} p_{j-1} (p_{j-1} in p2): applied here, so
    // the stopped iterate is the one of an update at every step
    if ((Sv.iters & 1) == 0)
      for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x)
        P.x[i] = fma(Sv.alpha, P.p2[i], P.x[i]);
    if (writer)
      S->brk = 1;
    return;
  }
  const double alpha = Sv.rz / pq[0];
  if (writer)
    S->alpha = alpha, S->pq = pq[0];
  double v[2] = {0., 0.};
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    if (i < r1) {
      const double ri = fma(-alpha, qv[j], rv[j]);
      P.r[i] = ri;
      const double ru = sv[j] * ri;
      v[0] += ri * ri, v[1] += ru * ru;
    }
  }
  block_sum<2>(v);
  if (threadIdx.x == 0)
    P.part2[int64_t(b) * 2 + 0] = v[0], P.part2[int64_t(b) * 2 + 1] = v[1];
}

// x = D^{-1/2} xh after the last step
__global__ void k_cg_finish(int64_t n, const double* __restrict__ dh, double* __restrict__ x)
{
  const int64_t i = xcd_index(n);
  if (i < n)
    x[i] = dh[i] * x[i];
}

// true residual at exit, on the vector row partition: block partials of |b - A x|^2 with ax = A x computed
// by the (unscaled) operator application
template <int ITEMS>
__global__ void __launch_bounds__(1024) k_true_residual(PcgParams P)
{
  const int b = logical_block(P.v_nb);
  if (b < 0)
    return;
  const int64_t r0 = P.v_row0[b], r1 = P.v_row0[b + 1];
  double v[1] = {0.};
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    if (i < r1) {
      const double ri = P.b[i] - P.ax[i];
      v[0] += ri * ri;
    }
  }
  block_sum<1>(v);
  if (threadIdx.x == 0)
    P.part3[b] = v[0];
}

// one block per instance: |b - A x| / |b| (|b - A x| when b = 0), with b.b from the instance's record
__global__ void __launch_bounds__(256) k_true_reduce(PcgParams P, double* __restrict__ out)
{
  const int inst = blockIdx.x;
  double t[1];
  reduce_partials<1>(P.part3, P.v_ib0[inst], P.v_ib0[inst + 1], t);
  if (threadIdx.x == 0) {
    const double bb = P.S[inst].bb;
    out[inst] = bb > 0. ? sqrt(t[0] / bb) : sqrt(t[0]);
  }
}

// scaled DIA storage, one diagonal after the other (coalesced): Ah(i, i+o) = a(i, i+o) (h_i h_{i+o});
// entries outside the instance are 0.  Diagonals below `kfirst` (the lower ones, never read by the
// symmetric format) and the main diagonal (unit, never read) are skipped.
__global__ void k_scale_dia(int64_t n, int ndiag, const int64_t* __restrict__ off, const int32_t* __restrict__ k0v,
                            int lower_too, const int32_t* __restrict__ row_inst, const int64_t* __restrict__ inst_row0,
                            const double* __restrict__ dia, const double* __restrict__ dh, double* __restrict__ dst)
{
  const int64_t i = xcd_index(n);
  const int k = blockIdx.y;
  if (i >= n)
    return;
  const int inst = row_inst[i];
  const int k0 = k0v[inst];
  if (k == k0 || (!lower_too && k < k0))
    return;
  const int64_t lo = inst_row0[inst], hi = inst_row0[inst + 1];
  const int64_t c = i + off[int64_t(inst) * ndiag + k];
  const int64_t at = int64_t(k) * n + i;
  dst[at] = (c >= lo && c < hi) ? dia[at] * (dh[i] * dh[c]) : 0.;
}

// values of the scaled matrix in the storage layout: Ah_ij = a_ij (h_i h_j); the product h_i h_j is
// symmetric in i and j, so an exactly symmetric A gives an exactly symmetric Ah
__global__ void k_scale_values(int64_t n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ cols,
                               const double* __restrict__ vals, const int64_t* __restrict__ pos,
                               const double* __restrict__ dh, double* __restrict__ dst)
{
  const int64_t i = xcd_index(n);
  if (i >= n)
    return;
  const double hi = dh[i];
  for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
    dst[pos[k]] = vals[k] * (hi * dh[cols[k]]);
}

__global__ void k_scatter_values(int64_t nnz, const double* __restrict__ vals, const int64_t* __restrict__ pos,
                                 double* __restrict__ dst)
{
  const int64_t s = xcd_index(nnz);
  if (s < nnz)
    dst[pos[s]] = vals[s];
}

__global__ void k_fill_int(int n, int* __restrict__ dst, int v)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n)
    dst[s] = v;
}

__global__ void k_fill(int64_t n, double* __restrict__ dst, double v)
{
  const int64_t s = xcd_index(n);
  if (s < n)
    dst[s] = v;
}

// Jacobi scaling from the CSR diagonal: h_i = 1 / sqrt(a_ii), s_i = sqrt(a_ii); flags a non-positive diagonal
__global__ void k_jacobi(int64_t n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ cols,
                         const double* __restrict__ vals, double* __restrict__ dh, double* __restrict__ dsq, int* bad)
{
  const int64_t i = xcd_index(n);
  if (i >= n)
    return;
  double d = 0.;
  for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
    if (cols[k] == i)
      d = vals[k];
  if (!(d > 0.)) {
    bad[0] = 1;
    d = 1.;
  }
  const double sq = sqrt(d);
  dh[i] = 1. / sq;
  dsq[i] = sq;
}

// DIA preparation in one pass over the rows: Jacobi scaling from the offset-0 diagonal (flagging a
// non-positive entry) and exact symmetry: for every stored negative offset o, A(i, i+o) must equal
// A(i+o, i), stored on the mirrored diagonal mirror[k] (offset -o) at row i+o
__global__ void k_dia_prepare(int64_t n, int ndiag, const int64_t* __restrict__ off, const int32_t* __restrict__ k0v,
                              const int32_t* __restrict__ mirror, const int32_t* __restrict__ row_inst,
                              const int64_t* __restrict__ inst_row0, const double* __restrict__ dia,
                              double* __restrict__ dh, double* __restrict__ dsq, int* flags, int* nz)
{
  const int64_t i = xcd_index(n);
  if (i >= n)
    return;
  const int inst = row_inst[i];
  const int64_t lo = inst_row0[inst], hi = inst_row0[inst + 1];
  const int64_t* o = off + int64_t(inst) * ndiag;
  const int32_t* mir = mirror + int64_t(inst) * ndiag;
  const int k0 = k0v[inst];
  double d = dia[int64_t(k0) * n + i];
  if (!(d > 0.)) {
    flags[0] = 1;
    d = 1.;
  }
  const double sq = sqrt(d);
  dh[i] = 1. / sq;
  dsq[i] = sq;
  bool ok = true;
  for (int k = 0; k < k0; ++k) {
    const int64_t c = i + o[k];
    if (c < lo || c >= hi)
      continue; // outside the instance: padding, never read
    const int m = mir[k];
    if (m < 0) {
      ok = false;
      continue;
    }
    // A(c, c - o) on the upper diagonal m: every upper entry inside the instance is met once here, so the
    // upper diagonals that hold a nonzero entry are flagged on the way (same value stored by every writer)
    const double up = dia[int64_t(m) * n + c];
    ok = ok && (up == dia[int64_t(k) * n + i]);
    wave_flag(up != 0., nz + int64_t(inst) * ndiag + m);
  }
  if (!ok)
    flags[1] = 1;
}

// ---------------------------------------------------------------------------------------------------------
// structured assembly: P1 Laplace forms and constant / per-element sources on the Kuhn tetrahedra of
// axis-parallel lattices (the grid walk of MatrixOperator::assemble, matrix-based.hh:496-500, with the local
// matrices of LocalElementIntegralBilinearForm(LocalLaplaceIntegrand), local/bilinear-forms/integrals.hh:97-134,
// local/integrands/laplace.hh:74-102, and the scatter of bilinear-form-assemblers.hh:110-128), written straight
// into the solver's symmetric DIA storage.  One lane per DoF (vertex) visits the 24 tetrahedra around it in the
// walk order (cells ascending, the six tetrahedra of a cell in permutation order) and adds each one's entries
// of its row: the diagonal and the entry with the next vertex on the tetrahedron's path, the only nonzero
// strict upper entries of a path tetrahedron's Laplace matrix (so the face and body diagonals of the stencil
// stay exactly 0, as the reference's assembly also leaves them).  The DIA slots k0 + u, u = dx + 2 dy + 4 dz,
// hold the entries with the vertex (x + dx, y + dy, z + dz) (ascending offsets 1, nx+1, nx+2, (nx+1)(ny+1), ...).
// Dirichlet rows get a unit diagonal, Dirichlet columns 0 (dirichlet-constraints.hh:153-184); the row's
// D^{-1/2} and D^{1/2} and the nonzero flags of the diagonals are written on the way.
// ---------------------------------------------------------------------------------------------------------
constexpr int DIRECT_MAX_FORMS = 8, DIRECT_MAX_SOURCES = 4;
struct DirectParams {
  int64_t n;
  int ninst;
  const DirectInst* inst;
  int nforms;
  const double* field[DIRECT_MAX_FORMS];
  double value[DIRECT_MAX_FORMS];
  double scaling[DIRECT_MAX_FORMS];
  int nsrc;
  const double* src_data[DIRECT_MAX_SOURCES]; // per element (NULL: the constant)
  double src_value[DIRECT_MAX_SOURCES];
  const uint8_t* mask; // Dirichlet DoFs (NULL: no constraints)
  double* dia;         // row i of slot u at dia[u * n + i]: u = 0 the diagonal, u = 1 .. 7 the upper diagonals
  double* rhs;
  double* dh;
  double* dsq;
  int* flags; // [0] non-positive diagonal; nonzero flag of slot u of instance t at nz[t * ndiag + k0 + u]
  int* nz;
  int ndiag, k0;
};

__global__ void __launch_bounds__(256) k_direct_p1_tets(DirectParams P)
{
  constexpr int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  const int64_t i = xcd_index(P.n);
  if (i >= P.n)
    return;
  int t0 = 0, t1 = P.ninst - 1; // the instance: last first-DoF <= i
  while (t0 < t1) {
    const int mid = (t0 + t1 + 1) >> 1;
    if (P.inst[mid].dof0 <= i)
      t0 = mid;
    else
      t1 = mid - 1;
  }
  const DirectInst I = P.inst[t0];
  const int64_t nvx = I.nx + 1, nvy = I.ny + 1, local = i - I.dof0;
  const int64_t x = local % nvx, y = (local / nvx) % nvy, z = local / (nvx * nvy);
  double ad = 0., ax = 0., ay = 0., az = 0., b = 0.; // diagonal, entries with +e_x, +e_y, +e_z
  for (int64_t cz = z - 1; cz <= z; ++cz) {
    if (cz < 0 || cz >= I.nz)
      continue;
    for (int64_t cy = y - 1; cy <= y; ++cy) {
      if (cy < 0 || cy >= I.ny)
        continue;
      for (int64_t cx = x - 1; cx <= x; ++cx) {
        if (cx < 0 || cx >= I.nx)
          continue;
        const int corner = int(x - cx) | (int(y - cy) << 1) | (int(z - cz) << 2);
        const int64_t e0 = I.elem0 + 6 * (cx + I.nx * (cy + I.ny * cz));
#pragma unroll
        for (int p = 0; p < 6; ++p) {
          const int c1 = 1 << PERM[p][0], c2 = c1 | (1 << PERM[p][1]);
          const int li = corner == 0 ? 0 : corner == c1 ? 1 : corner == c2 ? 2 : corner == 7 ? 3 : -1;
          if (li < 0)
            continue;
          const int64_t e = e0 + p;
          // unit-coefficient row li: diagonal wT[prev axis] + wT[next axis], entry -wT[next axis] with the next
          // path vertex (one step along the next axis)
          const int prev = li > 0 ? PERM[p][li - 1] : -1, next = li < 3 ? PERM[p][li] : -1;
          const double kd = (prev >= 0 ? I.wT[prev] : 0.) + (next >= 0 ? I.wT[next] : 0.);
          const double ku = next >= 0 ? -I.wT[next] : 0.;
          double cd = 0., cu = 0.;
          for (int f = 0; f < P.nforms; ++f) {
            const double c = P.field[f] ? P.field[f][e] : P.value[f];
            cd += P.scaling[f] * (c * kd);
            cu += P.scaling[f] * (c * ku);
          }
          ad += cd;
          if (next == 0)
            ax += cu;
          else if (next == 1)
            ay += cu;
          else if (next == 2)
            az += cu;
          for (int sidx = 0; sidx < P.nsrc; ++sidx)
            b += (P.src_data[sidx] ? P.src_data[sidx][e] : P.src_value[sidx]) * I.quarter_vol;
        }
      }
    }
  }
  // Dirichlet constraints: unit row, zero columns, b_i = 0
  const int64_t S = nvx * nvy;
  if (P.mask) {
    if (P.mask[i]) {
      ad = 1., ax = ay = az = 0., b = 0.;
    } else {
      if (x + 1 < nvx && P.mask[i + 1])
        ax = 0.;
      if (y + 1 < nvy && P.mask[i + nvx])
        ay = 0.;
      if (z < I.nz && P.mask[i + S])
        az = 0.;
    }
  }
  const int64_t n = P.n;
  P.dia[i] = ad;
  P.dia[1 * n + i] = ax;
  P.dia[2 * n + i] = ay;
  P.dia[3 * n + i] = 0.;
  P.dia[4 * n + i] = az;
  P.dia[5 * n + i] = 0.;
  P.dia[6 * n + i] = 0.;
  P.dia[7 * n + i] = 0.;
  P.rhs[i] = b;
  double d = ad;
  if (!(d > 0.)) {
    P.flags[0] = 1;
    d = 1.;
  }
  const double sq = sqrt(d);
  P.dh[i] = 1. / sq;
  P.dsq[i] = sq;
  int* nzi = P.nz + int64_t(t0) * P.ndiag + P.k0;
  wave_flag(ax != 0., nzi + 1);
  wave_flag(ay != 0., nzi + 2);
  wave_flag(az != 0., nzi + 4);
}

// CSR values from symmetric DIA storage of which only the diagonal and the upper diagonals are written: slot s
// sits at pos[s] = k n + i; a lower entry (k < k0) is the upper entry of its mirrored diagonal at row i + o
__global__ void k_values_from_upper(int64_t nnz, int64_t n, int ndiag, const int64_t* __restrict__ pos,
                                    const int64_t* __restrict__ off, const int32_t* __restrict__ k0v,
                                    const int32_t* __restrict__ mirror, const int32_t* __restrict__ row_inst,
                                    const double* __restrict__ dia, double* __restrict__ values)
{
  const int64_t s = xcd_index(nnz);
  if (s >= nnz)
    return;
  const int64_t p = pos[s];
  const int k = int(p / n);
  const int64_t i = p - int64_t(k) * n;
  const int inst = row_inst[i];
  if (k >= k0v[inst]) {
    values[s] = dia[p];
  } else {
    const int64_t t = int64_t(inst) * ndiag + k;
    values[s] = dia[int64_t(mirror[t]) * n + i + off[t]];
  }
}

// the solver's scaled copy of a symmetric DIA matrix: Ah(i, i+o) = a(i, i+o) (h_i h_{i+o}) for the kept strict
// upper diagonals (kidx: index of slot u's diagonal in the full layout, -1 for padding), compact
// (dst[u * n + i]); entries outside the instance are 0
__global__ void k_scale_upper(int64_t n, int s_nu, const int64_t* __restrict__ s_off, const int32_t* __restrict__ kidx,
                              const int32_t* __restrict__ row_inst, const int64_t* __restrict__ inst_row0,
                              const double* __restrict__ dia, const double* __restrict__ dh, double* __restrict__ dst)
{
  const int64_t i = xcd_index(n);
  if (i >= n)
    return;
  const int inst = row_inst[i];
  const int64_t lo = inst_row0[inst], hi = inst_row0[inst + 1];
  const double hi_i = dh[i];
  for (int u = 0; u < s_nu; ++u) {
    const int k = kidx[int64_t(inst) * s_nu + u];
    const int64_t c = i + s_off[int64_t(inst) * s_nu + u];
    dst[int64_t(u) * n + i] = (k >= 0 && c >= lo && c < hi) ? dia[int64_t(k) * n + i] * (hi_i * dh[c]) : 0.;
  }
}

template <class F>
void dispatch_items(int items, F&& f)
{
  switch (items) {
    case 1:
      f(std::integral_constant<int, 1>{});
      break;
    case 2:
      f(std::integral_constant<int, 2>{});
      break;
    case 4:
      f(std::integral_constant<int, 4>{});
      break;
    case 8:
      f(std::integral_constant<int, 8>{});
      break;
    default:
      HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "items_per_thread must be 1, 2, 4 or 8");
  }
}

// rows a lane of the operator application loads together: 2 for items = 2, else 1 (a block with
// items = 4, 8, 16 rows per lane processes them in chunks of block rows)
inline int apply_lane_rows(int items)
{
  return items == 2 ? 2 : 1;
}

// 1-d launches use the XCD-aware index (xcd_index)
inline unsigned blocks_for(int64_t n, int block)
{
  return xcd_grid(n, block);
}

inline unsigned hw_blocks(int64_t nb)
{
  return unsigned(((nb + 7) / 8) * 8);
}

} // namespace

// per-operator solver workspace that the header does not need to know about
struct SolverCache {
  // lists of the blocks of unfinished instances (two buffers: one may still be read by a queued chunk)
  DeviceArray<int32_t> active[2];
  std::vector<int32_t> active_host[2];
  DeviceArray<int64_t> block_slice0;
  DeviceArray<int32_t> dia_k0;
  DeviceArray<int32_t> dia_mirror;
  // column windows of the DIA stencils (k_cg_spmv_win): usable when every instance has at most WIN_MAX
  // windows and WIN_MAXU strict upper diagonals; the LDS of a block of R rows is at most
  // (win_nmax R + win_span) doubles, and blocks of R rows load a window with two entries per lane when its
  // span is at most R (win_wmax: the largest span)
  std::vector<std::vector<int64_t>> dia_offs; // per instance: its DIA offsets, ascending
  std::vector<int32_t> dia_k0_host;
  int win_block = 0; // row block size the windows were planned for (0: not planned)
  std::vector<DiaWindows> win_host;
  DeviceArray<DiaWindows> win; // uploaded when the staged kernel is first used
  bool win_ok = false;
  int win_nmax = 0;
  int64_t win_span = 0, win_wmax = 0;
  DeviceArray<int32_t> row_inst;
  // [0] non-positive diagonal, [1] asymmetric, then (DIA) per instance and diagonal: holds a nonzero entry
  DeviceArray<int> flags;
  int* host_flags = nullptr;
  int64_t host_flags_capacity = 0;
  // the solver's scaled copy of a symmetric DIA matrix keeps the strict upper diagonals with a nonzero entry:
  // per instance their offsets (host), and as tables [ninst][s_nu] their offsets (sentinel n) and their index
  // in the full layout (-1 for padding)
  int s_nu = -1;
  std::vector<std::vector<int64_t>> s_offs;
  std::vector<int32_t> s_kidx_host;
  DeviceArray<int64_t> s_off;
  DeviceArray<int32_t> s_kidx;
  PcgScalars* host_scalars[2] = {nullptr, nullptr};
  int host_capacity = 0;
  // true relative residual per instance at the solver's exit (device and pinned host copy)
  DeviceArray<double> true_relres;
  double* host_true = nullptr;
  int host_true_capacity = 0;
  hipEvent_t copied[2] = {nullptr, nullptr};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  std::vector<hipEvent_t> timing_events;
  hipStream_t own_stream = nullptr;
  hipEvent_t join = nullptr;
  hipGraphExec_t graph = nullptr;
  const double* graph_x = nullptr;
  const double* graph_b = nullptr;
  int64_t graph_maxit = 0;
  int graph_chunk = 0;
  int64_t graph_version = -1;
  int64_t version = 0;
  ~SolverCache()
  {
    for (auto& h : host_scalars)
      if (h)
        (void)hipHostFree(h);
    if (host_true)
      (void)hipHostFree(host_true);
    if (host_flags)
      (void)hipHostFree(host_flags);
    for (auto& e : copied)
      if (e)
        (void)hipEventDestroy(e);
    for (auto e : timing_events)
      (void)hipEventDestroy(e);
    if (t0)
      (void)hipEventDestroy(t0);
    if (t1)
      (void)hipEventDestroy(t1);
    if (join)
      (void)hipEventDestroy(join);
    if (graph)
      (void)hipGraphExecDestroy(graph);
    if (own_stream)
      (void)hipStreamDestroy(own_stream);
  }
};

static std::mutex g_cache_mutex;

SolverCache& cache_of(Operator& op)
{
  if (!op.solver_cache) {
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    op.solver_cache = std::make_shared<SolverCache>();
  }
  return *op.solver_cache;
}

void RowPartition::build(const std::vector<int64_t>& inst_dof, int block_size, int items_per_thread)
{
  block = block_size, items = items_per_thread;
  const int64_t R = int64_t(block) * items;
  row0.clear(), inst.clear(), inst_block0.assign(1, 0);
  const int ninst = int(inst_dof.size()) - 1;
  for (int in = 0; in < ninst; ++in) {
    for (int64_t r = inst_dof[in]; r < inst_dof[in + 1]; r += R) {
      row0.push_back(r);
      inst.push_back(in);
    }
    inst_block0.push_back(int32_t(row0.size()));
  }
  row0.push_back(inst_dof.back());
  nblocks = int64_t(inst.size());
  if (nblocks == 0)
    HGDT_THROW(HGDT_ERR_OPERATOR, "operator has no rows");
  if (nblocks >= (int64_t(1) << 31) / 8)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "too many row blocks");
  d_row0.upload(row0);
  d_inst.upload(inst);
  d_inst_block0.upload(inst_block0);
}

void RowPartition::build_counts(const std::vector<int64_t>& inst_dof, const std::vector<int64_t>& blocks_per_inst,
                                int block_size, int items_per_thread)
{
  block = block_size, items = items_per_thread;
  row0.clear(), inst.clear(), inst_block0.assign(1, 0);
  const int ninst = int(inst_dof.size()) - 1;
  for (int in = 0; in < ninst; ++in) {
    for (int64_t t = 0; t < blocks_per_inst[in]; ++t) {
      row0.push_back(inst_dof[in]);
      inst.push_back(in);
    }
    inst_block0.push_back(int32_t(row0.size()));
  }
  row0.push_back(inst_dof.back());
  nblocks = int64_t(inst.size());
  if (nblocks == 0)
    HGDT_THROW(HGDT_ERR_OPERATOR, "operator has no rows");
  if (nblocks >= (int64_t(1) << 31) / 8)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "too many row blocks");
  d_row0.upload(row0);
  d_inst.upload(inst);
  d_inst_block0.upload(inst_block0);
}

static bool blk_wanted(const Operator& op);

// row-block partitions of the operator (host set-up, O(number of blocks))
static void build_partitions(Operator& op)
{
  SolverFormat& F = op.fmt;
  SolverCache& C = cache_of(op);
  const Space& s = *op.space;
  // default launch shapes: one row per lane for the operator application (many blocks keep enough
  // loads in flight); the vector kernels take ~1024 blocks so that every block's reduction of an
  // instance's partials stays small
  if (!op.launch[3].user)
    op.launch[3].block = 256, op.launch[3].items = 1;
  // the matrix-free kernels give every lane one row (continuous spaces, at most 256 lanes per block) or one
  // element of a tile of elements (discontinuous spaces: 256 lanes, 4 rows each)
  const bool mf = op.requested_format == HGDT_FORMAT_MATRIX_FREE;
  const bool mf_tiles = mf && s.kind == HGDT_DISCONTINUOUS_LAGRANGE;
  // element blocks: one element (4 rows) per lane, at most 256 lanes per block (registers for two blocks
  // of loads in flight)
  const int block3 = mf_tiles ? 256 : (mf || blk_wanted(op)) ? std::min(op.launch[3].block, 256) : op.launch[3].block;
  const int items3 = (mf_tiles || blk_wanted(op)) ? 4 : mf ? 1 : op.launch[3].items;
  if (!op.launch[4].user) {
    op.launch[4].block = 256;
    const int64_t want = (op.n + 256 * 1024 - 1) / (256 * 1024);
    op.launch[4].items = want <= 1 ? 1 : want <= 2 ? 2 : want <= 4 ? 4 : 8;
  }
  const std::vector<int64_t> tiles = mf_tiles ? mf_tile_counts(op) : std::vector<int64_t>{};
  if (!tiles.empty())
    F.A.build_counts(s.inst_dof, tiles, block3, items3);
  else
    F.A.build(s.inst_dof, block3, items3);
  F.V.build(s.inst_dof, op.launch[4].block, op.launch[4].items);
  F.d_inst_row0.upload(s.inst_dof);
  if (mf)
    mf_block_table(op);
  F.partitions_ready = true;
  ++C.version; // captured graphs hold the launch shapes
}

// storage structure of the operator (host set-up): DIA offsets and CSR -> DIA positions, or the
// SELL-64 slices of the A partition
// element-block storage is the default for discontinuous spaces with 4 DoFs per element
static bool blk_wanted(const Operator& op)
{
  const Space& s = *op.space;
  const bool fits = s.kind == HGDT_DISCONTINUOUS_LAGRANGE && s.local_size == 4 && !s.grid->mixed && op.n % 4 == 0 &&
                    op.n > 0 && op.square();
  return fits && (op.requested_format == HGDT_FORMAT_BLOCK || op.requested_format == HGDT_FORMAT_AUTO);
}

// 4 x 4 element blocks from the CSR pattern: false when the rows of an element do not share their columns
// in whole aligned blocks of 4 including the element's own block (then another storage is used)
static bool build_blocks(Operator& op)
{
  SolverFormat& F = op.fmt;
  const int64_t ne = op.n / 4;
  std::vector<int32_t> cnt(ne, 0);
  std::atomic<bool> ok{true};
  std::mutex m;
  int mn = 0;
  parallel_for(ne, [&](int64_t e0, int64_t e1) {
    int lmax = 0;
    for (int64_t E = e0; E < e1 && ok; ++E) {
      const int64_t t0 = op.row_ptr[4 * E], len = op.row_ptr[4 * E + 1] - t0;
      bool self = false;
      bool good = len > 0 && len % 4 == 0;
      for (int a = 0; a < 4 && good; ++a) {
        const int64_t ta = op.row_ptr[4 * E + a];
        good = op.row_ptr[4 * E + a + 1] - ta == len;
        for (int64_t k = 0; k < len && good; ++k) {
          const int64_t c0 = op.cols[t0 + (k / 4) * 4];
          good = c0 % 4 == 0 && op.cols[ta + k] == c0 + k % 4;
          self = self || c0 == 4 * E;
        }
      }
      if (!good || !self) {
        ok = false;
        break;
      }
      cnt[E] = int32_t(len / 4 - 1);
      lmax = std::max(lmax, cnt[E]);
    }
    std::lock_guard<std::mutex> lock(m);
    mn = std::max(mn, lmax);
  });
  if (!ok)
    return false;
  // neighbour slots in ascending order; slot of the element in each neighbour's list
  std::vector<int32_t> nb(std::max<int64_t>(int64_t(mn) * ne, 1), -1), mir(std::max<int64_t>(int64_t(mn) * ne, 1), -1);
  parallel_for(ne, [&](int64_t e0, int64_t e1) {
    for (int64_t E = e0; E < e1; ++E) {
      const int64_t t0 = op.row_ptr[4 * E], nq = (op.row_ptr[4 * E + 1] - t0) / 4;
      int s = 0;
      for (int64_t q = 0; q < nq; ++q) {
        const int32_t Fb = op.cols[t0 + 4 * q] / 4;
        if (Fb != E)
          nb[int64_t(s++) * ne + E] = Fb;
      }
    }
  });
  std::atomic<bool> sym_pattern{true};
  int64_t up = 0, all = 0;
  parallel_for(ne, [&](int64_t e0, int64_t e1) {
    int64_t lup = 0, lall = 0;
    for (int64_t E = e0; E < e1; ++E)
      for (int s = 0; s < cnt[E]; ++s) {
        const int32_t Fb = nb[int64_t(s) * ne + E];
        ++lall;
        lup += Fb > E;
        int t = -1;
        for (int u = 0; u < cnt[Fb]; ++u)
          if (nb[int64_t(u) * ne + Fb] == E)
            t = u;
        if (t < 0)
          sym_pattern = false;
        mir[int64_t(s) * ne + E] = t;
      }
    std::lock_guard<std::mutex> lock(m);
    up += lup, all += lall;
  });
  if (!sym_pattern)
    return false;
  std::vector<int64_t> pos(std::max<int64_t>(op.nnz, 1));
  parallel_for(ne, [&](int64_t e0, int64_t e1) {
    for (int64_t E = e0; E < e1; ++E)
      for (int a = 0; a < 4; ++a) {
        const int64_t ta = op.row_ptr[4 * E + a], len = op.row_ptr[4 * E + a + 1] - ta;
        int s = 0;
        for (int64_t q = 0; q < len / 4; ++q) {
          const int32_t Fb = op.cols[ta + 4 * q] / 4;
          const int64_t base = (Fb == E) ? 0 : 16 + 16 * int64_t(s++);
          for (int c = 0; c < 4; ++c)
            pos[ta + 4 * q + c] = (base + 4 * a + c) * ne + E;
        }
      }
  });
  F.blk_ne = ne, F.blk_mn = mn, F.blk_up = up, F.blk_all = all;
  F.d_blk_nb.upload(nb);
  F.d_blk_mir.upload(mir);
  F.d_blk_pos.upload(pos);
  F.d_blk.allocate(std::max<int64_t>(int64_t(16 + 16 * mn) * ne, 1));
  return true;
}

// column windows of every instance's stencil for the symmetric DIA kernel with p staged in LDS, for row
// blocks of R rows: the sorted offsets {0, +-o : o a strict upper offset inside the instance} split into at
// most WIN_MAX contiguous groups of span at most R that load the fewest entries, sum of (R + span) over the
// groups (dynamic programming over at most 2 WIN_MAXU + 1 offsets).  Not usable (win_ok false) when an
// instance has more than WIN_MAXU strict upper slots or no such split exists.
static void plan_windows(Operator& op, int R)
{
  SolverCache& C = cache_of(op);
  const Space& s = *op.space;
  const auto& offs = C.s_offs; // the strict upper offsets of the solver's scaled copy, per instance
  const int ninst = int(offs.size());
  std::vector<DiaWindows> w(ninst);
  bool ok = ninst > 0 && C.s_nu <= WIN_MAXU;
  int nmax = 0;
  int64_t span = 0, wmax = 0;
  for (int in = 0; in < ninst && ok; ++in) {
    const int64_t size = s.inst_dof[in + 1] - s.inst_dof[in];
    std::vector<int64_t> o{0};
    for (size_t u = 0; u < offs[in].size(); ++u)
      if (offs[in][u] < size)
        o.push_back(offs[in][u]), o.push_back(-offs[in][u]);
    std::sort(o.begin(), o.end());
    const int m = int(o.size());
    constexpr int64_t INF = INT64_MAX / 4;
    // best[j][k]: fewest entries for the first j offsets in k groups; from[j][k]: start of the last group
    std::vector<std::array<int64_t, WIN_MAX + 1>> best(m + 1), from(m + 1);
    for (auto& b : best)
      b.fill(INF);
    best[0][0] = 0;
    for (int j = 1; j <= m; ++j)
      for (int k = 1; k <= WIN_MAX; ++k)
        for (int a0 = j - 1; a0 >= 0 && o[j - 1] - o[a0] <= R; --a0)
          if (best[a0][k - 1] < INF && best[a0][k - 1] + R + (o[j - 1] - o[a0]) < best[j][k])
            best[j][k] = best[a0][k - 1] + R + (o[j - 1] - o[a0]), from[j][k] = a0;
    int kb = 0;
    for (int k = 1; k <= WIN_MAX; ++k)
      if (best[m][k] < INF && (kb == 0 || best[m][k] < best[m][kb]))
        kb = k;
    if (kb == 0) {
      ok = false;
      break;
    }
    std::vector<int64_t> lo(kb), hi(kb);
    for (int j = m, k = kb; k > 0; --k) {
      const int a0 = int(from[j][k]);
      lo[k - 1] = o[a0], hi[k - 1] = o[j - 1];
      j = a0;
    }
    DiaWindows& d = w[in];
    d = DiaWindows{};
    d.nwin = kb;
    int64_t cs = 0;
    for (int v = 0; v < d.nwin; ++v) {
      d.lo[v] = lo[v];
      d.span[v] = int32_t(hi[v] - lo[v]); // <= R
      d.cspan[v] = int32_t(cs);
      cs += hi[v] - lo[v];
      wmax = std::max(wmax, hi[v] - lo[v]);
    }
    // window and position of an offset (-1: in no window, i.e. never a column of the instance)
    auto place = [&](int64_t x, int32_t& win, int32_t& rel) {
      win = -1, rel = 0;
      for (int v = 0; v < d.nwin; ++v)
        if (x >= d.lo[v] && x <= d.lo[v] + d.span[v])
          win = v, rel = int32_t(d.cspan[v] + (x - d.lo[v]));
    };
    place(0, d.zero_w, d.zero_rel);
    for (int u = 0; u < WIN_MAXU; ++u) {
      d.dn_w[u] = d.up_w[u] = -1;
      if (size_t(u) < offs[in].size() && offs[in][u] < size) {
        place(-offs[in][u], d.dn_w[u], d.dn_rel[u]);
        place(offs[in][u], d.up_w[u], d.up_rel[u]);
      }
    }
    nmax = std::max(nmax, int(d.nwin));
    span = std::max(span, cs);
  }
  C.win_ok = ok;
  C.win_nmax = nmax, C.win_span = span, C.win_wmax = wmax;
  C.win_block = R;
  C.win.release();
  C.win_host = ok ? std::move(w) : std::vector<DiaWindows>{};
}

// bytes of LDS that k_cg_spmv_win needs for row blocks of `block` rows (0: the kernel is not used)
static int64_t win_lds_bytes(const Operator& op)
{
  const SolverFormat& F = op.fmt;
  if (!op.solver_cache)
    return 0;
  const SolverCache& C = *op.solver_cache;
  if (!C.win_ok || C.win_block != F.A.block || !F.symmetric || F.format != HGDT_FORMAT_DIA || F.A.items != 1 ||
      C.win_wmax > F.A.block)
    return 0;
  if (op.column_staging == 0 || (op.column_staging < 0 && F.step_bytes <= WIN_AUTO_BYTES))
    return 0;
  const int64_t bytes = (int64_t(C.win_nmax) * F.A.block + C.win_span) * int64_t(sizeof(double));
  return bytes <= WIN_LDS ? bytes : 0;
}

// device flag array and its pinned host copy, at least `count` ints
static void flag_buffers(SolverCache& C, int64_t count)
{
  if (int64_t(C.flags.count) < count)
    C.flags.allocate(size_t(count));
  if (C.host_flags_capacity < count) {
    if (C.host_flags)
      HGDT_HIP(hipHostFree(C.host_flags));
    C.host_flags = nullptr;
    HGDT_HIP(hipHostMalloc((void**)&C.host_flags, size_t(count) * sizeof(int)));
    C.host_flags_capacity = count;
  }
}

// tables of the solver's scaled copy of a symmetric DIA matrix from the per-instance nonzero flags of the
// diagonals (nz[inst * ndiag + k]): the strict upper diagonals that hold a nonzero entry, ascending.  A
// diagonal whose entries are all exactly zero adds only 0 * p to every row, so leaving it out changes no
// sum.  The tables are uploaded (and captured graphs and column windows invalidated) only when they change.
static void compact_upper(Operator& op, const int* nz)
{
  SolverCache& C = cache_of(op);
  const SolverFormat& F = op.fmt;
  const int ninst = int(C.dia_offs.size()), nd = F.ndiag;
  std::vector<std::vector<int32_t>> kept(ninst);
  int snu = 0;
  for (int in = 0; in < ninst; ++in) {
    for (int k = C.dia_k0_host[in] + 1; k < int(C.dia_offs[in].size()); ++k)
      if (nz[int64_t(in) * nd + k])
        kept[in].push_back(k);
    snu = std::max(snu, int(kept[in].size()));
  }
  std::vector<int32_t> kidx(std::max<int64_t>(int64_t(ninst) * snu, 1), -1);
  std::vector<int64_t> off(std::max<int64_t>(int64_t(ninst) * snu, 1), op.n);
  std::vector<std::vector<int64_t>> offs(ninst);
  for (int in = 0; in < ninst; ++in)
    for (size_t u = 0; u < kept[in].size(); ++u) {
      kidx[int64_t(in) * snu + u] = kept[in][u];
      off[int64_t(in) * snu + u] = C.dia_offs[in][kept[in][u]];
      offs[in].push_back(C.dia_offs[in][kept[in][u]]);
    }
  if (snu == C.s_nu && kidx == C.s_kidx_host)
    return;
  C.s_kidx.upload(kidx);
  C.s_off.upload(off);
  C.s_kidx_host = std::move(kidx);
  C.s_offs = std::move(offs);
  C.s_nu = snu;
  C.win_ok = false, C.win_block = 0; // the column windows follow the kept offsets
  C.win.release();
  ++C.version; // captured graphs hold the tables and the slot count
}

static void build_structure(Operator& op)
{
  SolverFormat& F = op.fmt;
  SolverCache& C = cache_of(op);
  const Space& s = *op.space;
  const int ninst = int(s.inst_dof.size()) - 1;
  C.win_ok = false, C.win_block = 0;
  C.win.release();
  F.d_dia.release(), F.d_dia_pos.release(), F.d_sell_vals.release(), F.d_sell_cols.release();
  F.d_sell_pos.release(), F.d_scaled.release();
  F.d_mf_ref.release();
  F.mf_kind = -1;
  if (!op.square() && op.requested_format != HGDT_FORMAT_AUTO && op.requested_format != HGDT_FORMAT_SELL)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "an operator between two spaces is stored in SELL format only");
  if (op.requested_format == HGDT_FORMAT_MATRIX_FREE) {
    mf_setup(op);
    F.format = HGDT_FORMAT_MATRIX_FREE;
    F.ndiag = 0, F.nup = 0;
    F.structure_ready = true;
    F.valid = false;
    ++C.version;
    return;
  }

  F.d_blk_nb.release(), F.d_blk_mir.release(), F.d_blk.release(), F.d_blk_pos.release();
  if (op.requested_format == HGDT_FORMAT_BLOCK && !blk_wanted(op))
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "block storage needs a discontinuous space with 4 DoFs per element");
  if (blk_wanted(op)) {
    if (build_blocks(op)) {
      F.format = HGDT_FORMAT_BLOCK;
      F.ndiag = F.blk_mn, F.nup = 0;
      F.structure_ready = true;
      F.valid = false;
      ++C.version;
      return;
    }
    if (op.requested_format == HGDT_FORMAT_BLOCK)
      HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "block storage: the pattern does not consist of whole 4 x 4 element blocks");
  }

  // DIA detection: distinct (col - row) offsets per instance
  std::vector<std::vector<int64_t>> offs(ninst);
  bool dia_ok = op.requested_format != HGDT_FORMAT_SELL && op.square();
  for (int in = 0; in < ninst && dia_ok; ++in) {
    std::vector<int64_t>& o = offs[in];
    std::mutex m;
    std::atomic<bool> ok{true};
    parallel_for(s.inst_dof[in + 1] - s.inst_dof[in], [&](int64_t b0, int64_t b1) {
      std::vector<int64_t> loc;
      for (int64_t i = s.inst_dof[in] + b0; i < s.inst_dof[in] + b1 && ok; ++i)
        for (int64_t t = op.row_ptr[i]; t < op.row_ptr[i + 1]; ++t) {
          const int64_t d = int64_t(op.cols[t]) - i;
          auto it = std::lower_bound(loc.begin(), loc.end(), d);
          if (it == loc.end() || *it != d) {
            loc.insert(it, d);
            if (loc.size() > size_t(MAXDIAG)) {
              ok = false;
              break;
            }
          }
        }
      std::lock_guard<std::mutex> lock(m);
      for (int64_t d : loc) {
        auto it = std::lower_bound(o.begin(), o.end(), d);
        if (it == o.end() || *it != d)
          o.insert(it, d);
      }
    });
    if (!ok || o.size() > size_t(MAXDIAG) || !std::binary_search(o.begin(), o.end(), int64_t(0)))
      dia_ok = false;
  }
  if (op.requested_format == HGDT_FORMAT_DIA && !dia_ok)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "DIA storage requested but an instance has more than 32 diagonals");
  if (dia_ok) {
    F.format = HGDT_FORMAT_DIA;
    int nd = 0;
    for (auto& o : offs)
      nd = std::max(nd, int(o.size()));
    F.ndiag = nd;
    std::vector<int64_t> table(int64_t(ninst) * nd, op.n); // sentinel offset: never inside an instance
    std::vector<int32_t> k0(ninst, 0);
    int nup = 0;
    for (int in = 0; in < ninst; ++in) {
      for (size_t k = 0; k < offs[in].size(); ++k)
        table[int64_t(in) * nd + k] = offs[in][k];
      k0[in] = int32_t(std::lower_bound(offs[in].begin(), offs[in].end(), int64_t(0)) - offs[in].begin());
      nup = std::max(nup, int(offs[in].size()) - k0[in]);
    }
    F.nup = nup;
    std::vector<int64_t> pos(std::max<int64_t>(op.nnz, 1));
    std::vector<int32_t> row_inst(std::max<int64_t>(op.n, 1));
    for (int in = 0; in < ninst; ++in) {
      const auto& o = offs[in];
      parallel_for(s.inst_dof[in + 1] - s.inst_dof[in], [&](int64_t b0, int64_t b1) {
        for (int64_t i = s.inst_dof[in] + b0; i < s.inst_dof[in] + b1; ++i) {
          row_inst[i] = in;
          for (int64_t t = op.row_ptr[i]; t < op.row_ptr[i + 1]; ++t) {
            const int64_t k = std::lower_bound(o.begin(), o.end(), int64_t(op.cols[t]) - i) - o.begin();
            pos[t] = k * op.n + i;
          }
        }
      });
    }
    // mirror[k]: index of the diagonal with offset -offset[k] (-1 when the pattern is not symmetric)
    std::vector<int32_t> mirror(int64_t(ninst) * nd, -1);
    for (int in = 0; in < ninst; ++in)
      for (size_t k = 0; k < offs[in].size(); ++k) {
        auto it = std::lower_bound(offs[in].begin(), offs[in].end(), -offs[in][k]);
        if (it != offs[in].end() && *it == -offs[in][k])
          mirror[int64_t(in) * nd + k] = int32_t(it - offs[in].begin());
      }
    C.dia_mirror.upload(mirror);
    C.dia_offs = offs, C.dia_k0_host = k0;
    C.s_nu = -1, C.s_kidx_host.clear(), C.s_offs.clear(); // compact tables rebuilt at the next preparation
    C.win_block = 0; // windows planned for the row blocks of the solver
    F.d_dia_offsets.upload(table);
    C.dia_k0.upload(k0);
    C.row_inst.upload(row_inst);
    F.d_dia_pos.upload(pos);
    F.d_dia.allocate(std::max<int64_t>(int64_t(nd) * op.n, 1));
  } else {
    F.format = HGDT_FORMAT_SELL;
    std::vector<int64_t> slice_ptr{0}, bslice0;
    for (int64_t b = 0; b < F.A.nblocks; ++b) {
      bslice0.push_back(int64_t(slice_ptr.size()) - 1);
      for (int64_t r = F.A.row0[b]; r < F.A.row0[b + 1]; r += 64) {
        int64_t w = 0;
        for (int64_t i = r; i < std::min(r + 64, F.A.row0[b + 1]); ++i)
          w = std::max(w, op.row_ptr[i + 1] - op.row_ptr[i]);
        slice_ptr.push_back(slice_ptr.back() + 64 * w);
      }
    }
    bslice0.push_back(int64_t(slice_ptr.size()) - 1);
    F.nslices = int64_t(slice_ptr.size()) - 1;
    const int64_t total = slice_ptr.back();
    std::vector<int32_t> scols(std::max<int64_t>(total, 1));
    std::vector<int64_t> pos(std::max<int64_t>(op.nnz, 1));
    parallel_for(F.A.nblocks, [&](int64_t b0, int64_t b1) {
      for (int64_t b = b0; b < b1; ++b) {
        int64_t sl = bslice0[b];
        for (int64_t r = F.A.row0[b]; r < F.A.row0[b + 1]; r += 64, ++sl) {
          const int64_t w = (slice_ptr[sl + 1] - slice_ptr[sl]) / 64;
          for (int64_t i = r; i < r + 64; ++i) {
            const int lane = int(i - r);
            const int64_t len = (i < F.A.row0[b + 1]) ? op.row_ptr[i + 1] - op.row_ptr[i] : 0;
            for (int64_t k = 0; k < w; ++k) {
              const int64_t at = slice_ptr[sl] + k * 64 + lane;
              if (k < len) {
                scols[at] = op.cols[op.row_ptr[i] + k];
                pos[op.row_ptr[i] + k] = at;
              } else {
                scols[at] = int32_t(std::min<int64_t>(i, op.ncols - 1)); // padding: zero value, valid column
              }
            }
          }
        }
      }
    });
    F.d_slice_ptr.upload(slice_ptr);
    C.block_slice0.upload(bslice0);
    F.d_sell_cols.upload(scols);
    F.d_sell_pos.upload(pos);
    F.d_sell_vals.allocate(std::max<int64_t>(total, 1));
  }
  F.structure_ready = true;
  F.valid = false;
  ++C.version;
}

// algorithmic bytes of one operator application in the format in use: matrix values (+ int32 columns for
// SELL), x read, y written
static int64_t apply_bytes(const Operator& op)
{
  const SolverFormat& F = op.fmt;
  if (F.format == HGDT_FORMAT_MATRIX_FREE)
    return mf_bytes(op, MF_APPLY);
  if (F.format == HGDT_FORMAT_BLOCK) // own block (upper triangle if symmetric), neighbour blocks read
    return 8 * ((F.symmetric ? 10 : 16) * F.blk_ne + 16 * (F.symmetric ? F.blk_up : F.blk_all)) + 16 * op.n;
  if (F.format == HGDT_FORMAT_DIA)
    return int64_t(F.symmetric ? F.nup : F.ndiag) * 8 * op.n + 16 * op.n;
  return 12 * op.nnz + 8 * op.n + 8 * op.ncols;
}

// algorithmic bytes of the CG solver's operator application: the scaled matrix without its unit diagonal
// (DIA) or with it (SELL, + int32 columns), p read, q written
static int64_t step_bytes(const Operator& op)
{
  const SolverFormat& F = op.fmt;
  if (F.format == HGDT_FORMAT_MATRIX_FREE) // coefficients and vectors (with the direction update when fused)
    return mf_bytes(op, mf_fused_direction(op) ? MF_PCG_DIR : MF_PCG);
  if (F.format == HGDT_FORMAT_BLOCK) // the unit diagonal is not read
    return 8 * ((F.symmetric ? 6 : 12) * F.blk_ne + 16 * (F.symmetric ? F.blk_up : F.blk_all)) + 16 * op.n;
  if (F.format == HGDT_FORMAT_DIA && F.symmetric) // the strict upper diagonals with a nonzero entry
    return int64_t(std::max(op.solver_cache ? op.solver_cache->s_nu : 0, 0)) * 8 * op.n + 16 * op.n;
  if (F.format == HGDT_FORMAT_DIA)
    return int64_t(F.ndiag - 1) * 8 * op.n + 16 * op.n;
  return 12 * op.nnz + 16 * op.n;
}

static PcgParams make_params(Operator& op)
{
  SolverFormat& F = op.fmt;
  SolverCache& C = cache_of(op);
  PcgParams P{};
  P.n = op.n;
  P.a_nb = int(F.A.nblocks);
  P.a_row0 = F.A.d_row0.ptr;
  P.a_inst = F.A.d_inst.ptr;
  P.a_ib0 = F.A.d_inst_block0.ptr;
  P.v_nb = int(F.V.nblocks);
  P.v_row0 = F.V.d_row0.ptr;
  P.v_inst = F.V.d_inst.ptr;
  P.v_ib0 = F.V.d_inst_block0.ptr;
  P.inst_row0 = F.d_inst_row0.ptr;
  P.ndiag = F.ndiag;
  P.dia_off = F.d_dia_offsets.ptr;
  P.dia_k0 = C.dia_k0.ptr;
  P.s_nu = std::max(C.s_nu, 0);
  P.s_off = C.s_off.ptr;
  P.dia = F.d_dia.ptr;
  P.win = C.win.ptr;
  P.block_slice0 = C.block_slice0.ptr;
  P.slice_ptr = F.d_slice_ptr.ptr;
  P.sell_cols = F.d_sell_cols.ptr;
  P.sell_vals = F.d_sell_vals.ptr;
  P.blk_ne = F.blk_ne;
  P.blk_mn = F.blk_mn;
  P.blk_nb = F.d_blk_nb.ptr;
  P.blk_mir = F.d_blk_mir.ptr;
  P.blk = F.d_blk.ptr;
  P.scaled = F.d_scaled.ptr;
  P.dh = F.d_dh.ptr;
  P.dsq = F.d_dsq.ptr;
  P.ninst = int(op.space->inst_dof.size()) - 1;
  return P;
}

static int kernel_format(const Operator& op)
{
  if (op.fmt.format == HGDT_FORMAT_SELL)
    return FMT_SELL;
  if (op.fmt.format == HGDT_FORMAT_BLOCK)
    return FMT_BLK;
  return op.fmt.symmetric ? FMT_DIAS : FMT_DIA;
}

// the structured assembly applies: P1 on the Kuhn tetrahedra of axis-parallel lattices (every instance), Laplace
// element forms only, constant or per-element sources, symmetric DIA storage with the Kuhn stencil's 15 diagonals
bool Operator::direct_applicable()
{
  if (direct_state >= 0)
    return direct_state == 1;
  direct_state = 0;
  if (borrowed || !square()) // a caller's matrix or an operator between two spaces: the general path
    return false;
  if (dirichlet && (dirichlet_only_clear || !dirichlet_ensure_symmetry)) // writes unit rows and columns
    return false;
  const Space& s = *space;
  const Grid& g = *s.grid;
  // every instance a lattice (a structured grid, or a union of structured grids, which keeps their lattices)
  if (g.ref.geom != HGDT_TET || int(g.lattice.size()) != g.ninst() || g.ninst() < 1)
    return false;
  if (s.kind != HGDT_CONTINUOUS_LAGRANGE || s.order != 1 || stencil != HGDT_STENCIL_ELEMENT)
    return false;
  if (eforms.empty() || eforms.size() > size_t(DIRECT_MAX_FORMS) || !cforms.empty() || !bforms.empty() ||
      !bsources.empty() || sources.size() > size_t(DIRECT_MAX_SOURCES))
    return false;
  for (const auto& f : eforms)
    if (f.kind != EF_LAPLACE || f.points || f.tensor_kind)
      return false;
  for (const auto& src : sources)
    if (src.kind != HGDT_SOURCE_CONSTANT && src.kind != HGDT_SOURCE_PER_ELEMENT)
      return false;
  if (!(requested_format == HGDT_FORMAT_AUTO || requested_format == HGDT_FORMAT_DIA) || !allow_symmetric)
    return false;
  if (!fmt.partitions_ready)
    build_partitions(*this);
  if (!fmt.structure_ready)
    build_structure(*this);
  SolverCache& C = cache_of(*this);
  if (fmt.format != HGDT_FORMAT_DIA || fmt.ndiag != 15)
    return false;
  std::vector<DirectInst> rec(g.ninst());
  for (int in = 0; in < g.ninst(); ++in) {
    const Lattice& L = g.lattice[in];
    const int64_t nvx = L.n[0] + 1, nvy = L.n[1] + 1, S = nvx * nvy;
    const int64_t up[7] = {1, nvx, nvx + 1, S, S + 1, S + nvx, S + nvx + 1};
    std::vector<int64_t> expect;
    for (int u = 6; u >= 0; --u)
      expect.push_back(-up[u]);
    expect.push_back(0);
    for (int u = 0; u < 7; ++u)
      expect.push_back(up[u]);
    if (C.dia_offs[in] != expect || C.dia_k0_host[in] != 7 ||
        s.inst_dof[in + 1] - s.inst_dof[in] != S * (L.n[2] + 1) || g.inst_elem[in + 1] - g.inst_elem[in] != 6 * L.n[0] * L.n[1] * L.n[2])
      return false;
    DirectInst& r = rec[in];
    r = DirectInst{};
    r.dof0 = s.inst_dof[in], r.elem0 = g.inst_elem[in];
    r.nx = int32_t(L.n[0]), r.ny = int32_t(L.n[1]), r.nz = int32_t(L.n[2]);
    const double vol = L.h[0] * L.h[1] * L.h[2] / 6.;
    for (int a = 0; a < 3; ++a)
      r.wT[a] = vol / (L.h[a] * L.h[a]);
    r.quarter_vol = vol / 4.;
  }
  d_direct_inst.upload(rec);
  direct_state = 1;
  return true;
}

void Operator::assemble_direct(hipStream_t stream)
{
  SolverFormat& F = fmt;
  SolverCache& C = cache_of(*this);
  const int ninst = int(space->inst_dof.size()) - 1;
  F.d_dh.allocate(std::max<int64_t>(n, 1));
  F.d_dsq.allocate(std::max<int64_t>(n, 1));
  const int64_t nflags = 2 + int64_t(ninst) * F.ndiag;
  flag_buffers(C, nflags);
  const int B = 256;
  hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(nflags, B)), dim3(B), 0, stream, int(nflags), C.flags.ptr, 0);
  DirectParams P{};
  P.n = n, P.ninst = ninst, P.inst = d_direct_inst.ptr;
  P.nforms = int(eforms.size());
  for (int f = 0; f < P.nforms; ++f)
    P.field[f] = eforms[f].field, P.value[f] = eforms[f].value, P.scaling[f] = eforms[f].scaling;
  P.nsrc = int(sources.size());
  for (int k = 0; k < P.nsrc; ++k) {
    P.src_data[k] = sources[k].kind == HGDT_SOURCE_PER_ELEMENT ? sources[k].data : nullptr;
    P.src_value[k] = sources[k].value;
  }
  P.mask = dirichlet ? d_dof_mask.ptr : nullptr;
  P.ndiag = F.ndiag, P.k0 = 7;
  P.dia = F.d_dia.ptr + int64_t(7) * n;
  P.rhs = d_rhs.ptr, P.dh = F.d_dh.ptr, P.dsq = F.d_dsq.ptr;
  P.flags = C.flags.ptr, P.nz = C.flags.ptr + 2;
  hipLaunchKernelGGL(k_direct_p1_tets, dim3(blocks_for(n, B)), dim3(B), 0, stream, P);
  HGDT_HIP(hipGetLastError());
  F.valid = false;
  direct_storage = true;
  values_stale = true;
}

int Operator::solver_upper_diagonals() const
{
  if (!fmt.valid || fmt.format != HGDT_FORMAT_DIA || !fmt.symmetric || !solver_cache)
    return -1;
  return solver_cache->s_nu;
}

void Operator::materialize_values(hipStream_t stream)
{
  if (!values_stale)
    return;
  if (!direct_storage || fmt.format != HGDT_FORMAT_DIA)
    HGDT_THROW(HGDT_ERR_OPERATOR, "CSR values requested from a storage that does not hold the assembly");
  SolverCache& C = cache_of(*this);
  hipLaunchKernelGGL(k_values_from_upper, dim3(blocks_for(nnz, 256)), dim3(256), 0, stream, nnz, n, fmt.ndiag,
                     fmt.d_dia_pos.ptr, fmt.d_dia_offsets.ptr, C.dia_k0.ptr, C.dia_mirror.ptr, C.row_inst.ptr,
                     fmt.d_dia.ptr, d_values.ptr);
  HGDT_HIP(hipGetLastError());
  HGDT_HIP(hipStreamSynchronize(stream)); // the values may be read on another stream or by a copy next
  values_stale = false;
}

void Operator::prepare_solver(hipStream_t stream)
{
  prepare_storage(stream);
  // column windows for the current row blocks, on the device once the staged application is chosen
  // (set-up, never captured)
  SolverCache& C = cache_of(*this);
  if (fmt.format == HGDT_FORMAT_DIA && fmt.symmetric && fmt.A.items == 1 && column_staging != 0 &&
      C.win_block != fmt.A.block)
    plan_windows(*this, fmt.A.block);
  if (win_lds_bytes(*this) > 0 && !C.win.ptr)
    C.win.upload(C.win_host);
}

void Operator::prepare_storage(hipStream_t stream)
{
  SolverFormat& F = fmt;
  SolverCache& C = cache_of(*this);
  if (!F.partitions_ready)
    build_partitions(*this);
  if (!F.structure_ready)
    build_structure(*this);
  if (F.valid)
    return;
  F.d_dh.allocate(std::max<int64_t>(n, 1));
  F.d_dsq.allocate(std::max<int64_t>(n, 1));
  const int ninst = int(space->inst_dof.size()) - 1;
  const int64_t nflags = 2 + (F.format == HGDT_FORMAT_DIA ? int64_t(ninst) * F.ndiag : 0);
  flag_buffers(C, nflags);
  if (F.format == HGDT_FORMAT_MATRIX_FREE) {
    // Jacobi scaling from the diagonal of the operator, recomputed from the coefficients
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, 2, C.flags.ptr, 0);
    MfVectors v;
    v.out_dh = F.d_dh.ptr, v.out_dsq = F.d_dsq.ptr, v.flags = C.flags.ptr;
    mf_launch(*this, MF_DIAG, v, stream);
    HGDT_HIP(hipMemcpyAsync(C.host_flags, C.flags.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
    HGDT_HIP(hipStreamSynchronize(stream));
    F.diag_positive = (C.host_flags[0] == 0);
    F.symmetric = false;
    F.bytes = apply_bytes(*this);
    F.step_bytes = step_bytes(*this);
    F.valid = true;
    return;
  }
  if (matrix_free_assembly)
    HGDT_THROW(HGDT_ERR_OPERATOR,
               "the operator was assembled in matrix-free mode and holds no matrix: assemble again in this format");
  const int B = 256;
  if (F.format == HGDT_FORMAT_BLOCK) {
    const int64_t total = int64_t(16 + 16 * F.blk_mn) * F.blk_ne;
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(total, B)), dim3(B), 0, stream, total, F.d_blk.ptr, 0.);
    hipLaunchKernelGGL(k_scatter_values, dim3(blocks_for(nnz, B)), dim3(B), 0, stream, nnz, d_values.ptr,
                       F.d_blk_pos.ptr, F.d_blk.ptr);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, 2, C.flags.ptr, 0);
    const PcgParams P = make_params(*this);
    hipLaunchKernelGGL(k_blk_prepare, dim3(blocks_for(F.blk_ne, B)), dim3(B), 0, stream, P, F.d_dh.ptr, F.d_dsq.ptr,
                       C.flags.ptr);
    HGDT_HIP(hipGetLastError());
    HGDT_HIP(hipMemcpyAsync(C.host_flags, C.flags.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
    HGDT_HIP(hipStreamSynchronize(stream));
    F.diag_positive = (C.host_flags[0] == 0);
    const bool sym = C.host_flags[1] == 0 && allow_symmetric;
    F.d_scaled.allocate(std::max<int64_t>(total, 1));
    hipLaunchKernelGGL(k_blk_scale, dim3(blocks_for(F.blk_ne, B)), dim3(B), 0, stream, P, F.d_dh.ptr,
                       F.d_scaled.ptr);
    HGDT_HIP(hipGetLastError());
    if (sym != F.symmetric)
      ++C.version; // captured graphs hold the kernel variant
    F.symmetric = sym;
    F.bytes = apply_bytes(*this);
    F.step_bytes = step_bytes(*this);
    F.valid = true;
    return;
  }
  const bool dia = F.format == HGDT_FORMAT_DIA;
  double* dst = dia ? F.d_dia.ptr : F.d_sell_vals.ptr;
  const int64_t total = dia ? int64_t(F.ndiag) * n : int64_t(F.d_sell_vals.count);
  if (!square()) {
    // an operator between two spaces: its SELL values for apply(); no Jacobi scaling, no solver copy
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(total, B)), dim3(B), 0, stream, total, dst, 0.);
    hipLaunchKernelGGL(k_scatter_values, dim3(blocks_for(nnz, B)), dim3(B), 0, stream, nnz, d_values.ptr,
                       F.d_sell_pos.ptr, dst);
    HGDT_HIP(hipGetLastError());
    F.diag_positive = false;
    F.symmetric = false;
    F.bytes = apply_bytes(*this);
    F.step_bytes = 0;
    F.valid = true;
    return;
  }
  if (dia && direct_storage) {
    // the structured assembly wrote the diagonal and the upper diagonals, D^{-1/2}, D^{1/2} and the flags
    // (symmetric by construction); nothing to convert
  } else {
    if (values_stale)
      materialize_values(stream);
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(total, B)), dim3(B), 0, stream, total, dst, 0.);
    hipLaunchKernelGGL(k_scatter_values, dim3(blocks_for(nnz, B)), dim3(B), 0, stream, nnz, d_values.ptr,
                       dia ? F.d_dia_pos.ptr : F.d_sell_pos.ptr, dst);
    hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(nflags, B)), dim3(B), 0, stream, int(nflags), C.flags.ptr, 0);
    if (dia)
      hipLaunchKernelGGL(k_dia_prepare, dim3(blocks_for(n, B)), dim3(B), 0, stream, n, F.ndiag, F.d_dia_offsets.ptr,
                         C.dia_k0.ptr, C.dia_mirror.ptr, C.row_inst.ptr, F.d_inst_row0.ptr, F.d_dia.ptr, F.d_dh.ptr,
                         F.d_dsq.ptr, C.flags.ptr, C.flags.ptr + 2);
    else
      hipLaunchKernelGGL(k_jacobi, dim3(blocks_for(n, B)), dim3(B), 0, stream, n, d_row_ptr.ptr, d_cols.ptr,
                         d_values.ptr, F.d_dh.ptr, F.d_dsq.ptr, C.flags.ptr);
    HGDT_HIP(hipGetLastError());
  }
  HGDT_HIP(hipMemcpyAsync(C.host_flags, C.flags.ptr, nflags * sizeof(int), hipMemcpyDeviceToHost, stream));
  HGDT_HIP(hipStreamSynchronize(stream));
  F.diag_positive = (C.host_flags[0] == 0);
  const bool sym = dia && C.host_flags[1] == 0 && allow_symmetric;
  if (dia && !sym && direct_storage)
    HGDT_THROW(HGDT_ERR_OPERATOR, "structured assembly into symmetric storage met a non-symmetric operator");
  // the scaled matrix for the CG solver: symmetric DIA as its strict upper diagonals with a nonzero entry,
  // otherwise in the layout of the storage
  if (dia && sym) {
    compact_upper(*this, C.host_flags + 2);
    F.d_scaled.allocate(std::max<int64_t>(int64_t(C.s_nu) * n, 1));
    hipLaunchKernelGGL(k_scale_upper, dim3(blocks_for(n, B)), dim3(B), 0, stream, n, C.s_nu, C.s_off.ptr,
                       C.s_kidx.ptr, C.row_inst.ptr, F.d_inst_row0.ptr, F.d_dia.ptr, F.d_dh.ptr, F.d_scaled.ptr);
  } else if (dia) {
    F.d_scaled.allocate(std::max<int64_t>(total, 1));
    hipLaunchKernelGGL(k_scale_dia, dim3(blocks_for(n, B), unsigned(F.ndiag)), dim3(B), 0, stream, n, F.ndiag,
                       F.d_dia_offsets.ptr, C.dia_k0.ptr, 1, C.row_inst.ptr, F.d_inst_row0.ptr, F.d_dia.ptr,
                       F.d_dh.ptr, F.d_scaled.ptr);
  } else {
    F.d_scaled.allocate(std::max<int64_t>(total, 1));
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(total, B)), dim3(B), 0, stream, total, F.d_scaled.ptr, 0.);
    hipLaunchKernelGGL(k_scale_values, dim3(blocks_for(n, B)), dim3(B), 0, stream, n, d_row_ptr.ptr, d_cols.ptr,
                       d_values.ptr, F.d_sell_pos.ptr, F.d_dh.ptr, F.d_scaled.ptr);
  }
  HGDT_HIP(hipGetLastError());
  if (sym != F.symmetric)
    ++C.version; // captured graphs hold the kernel variant
  F.symmetric = sym;
  F.bytes = apply_bytes(*this);
  F.step_bytes = step_bytes(*this);
  F.valid = true;
}

void Operator::apply(const double* x, double* y, hipStream_t stream)
{
  prepare_solver(stream);
  if (fmt.format == HGDT_FORMAT_MATRIX_FREE) {
    MfVectors v;
    v.x = x, v.y = y;
    mf_launch(*this, MF_APPLY, v, stream);
    return;
  }
  PcgParams P = make_params(*this);
  const int block = fmt.A.block;
  const unsigned grid = hw_blocks(fmt.A.nblocks);
  const int kf = kernel_format(*this);
  if (kf == FMT_BLK) {
    const unsigned g1 = xcd_grid(fmt.blk_ne, 256);
    if (fmt.symmetric)
      hipLaunchKernelGGL(k_blk_apply<true>, dim3(g1), dim3(256), 0, stream, P, x, y);
    else
      hipLaunchKernelGGL(k_blk_apply<false>, dim3(g1), dim3(256), 0, stream, P, x, y);
    HGDT_HIP(hipGetLastError());
    return;
  }
  dispatch_items(apply_lane_rows(fmt.A.items), [&](auto IT) {
    constexpr int I = decltype(IT)::value;
    if (kf == FMT_DIA)
      hipLaunchKernelGGL((k_apply<FMT_DIA, I>), dim3(grid), dim3(block), 0, stream, P, x, y);
    else if (kf == FMT_DIAS)
      hipLaunchKernelGGL((k_apply<FMT_DIAS, I>), dim3(grid), dim3(block), 0, stream, P, x, y);
    else
      hipLaunchKernelGGL((k_apply<FMT_SELL, I>), dim3(grid), dim3(block), 0, stream, P, x, y);
  });
  HGDT_HIP(hipGetLastError());
}

// launch of k(P, par); with events, they are bound to the dispatch itself (ev0 to its start, ev1 to its end;
// either may be NULL), not recorded as separate markers around it
template <class K>
static void launch_pcg(K kernel, unsigned grid, int block, hipStream_t stream, const PcgParams& P, int par,
                       hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, size_t lds = 0)
{
  if (!ev0 && !ev1) {
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, stream, P, par);
    return;
  }
  PcgParams p = P;
  int q = par;
  void* args[] = {&p, &q};
  HGDT_HIP(hipExtLaunchKernel(reinterpret_cast<const void*>(kernel), dim3(grid), dim3(block), args, lds, stream,
                              ev0, ev1, 0));
}

// ev0 / ev1 time the iteration's operator application (ev0 NULL: only ev1 is bound, to its end), ev_upd is
// bound to the end of the iteration's update
static void launch_iteration(const Operator& op, const PcgParams& P, int par, hipStream_t stream, hipEvent_t ev0,
                             hipEvent_t ev1, hipEvent_t ev_upd = nullptr)
{
  const RowPartition& A = op.fmt.A;
  const RowPartition& V = op.fmt.V;
  const int kf = kernel_format(op);
  const bool fused = mf_fused_direction(op);
  // block order: the operator application walks the blocks of every XCD upward, the vector kernels
  // downward, so the direction update ends and the update starts on the rows the application touches
  // first and last
  PcgParams Pd = P, Pa = P, Pu = P;
  Pd.rev = 1, Pa.rev = 0, Pu.rev = 1;
  // timed application: from the end of the direction update to the end of the application (the kernels
  // run back to back, so this is the application's own interval; an event bound to the start of a
  // dispatch also counts the dispatch's set-up); the caller times an application fused with the direction
  // update from the end of the previous update
  if (!fused) {
    dispatch_items(V.items, [&](auto IT) {
      constexpr int I = decltype(IT)::value;
      launch_pcg(k_cg_direction<I>, hw_blocks(P.v_nb), V.block, stream, Pd, par, nullptr, ev0);
    });
    ev0 = nullptr;
  }
  if (op.fmt.format == HGDT_FORMAT_MATRIX_FREE) {
    MfVectors v;
    v.x = par ? P.p : P.p2, v.y = P.q, v.dh = P.dh, v.part = P.part1, v.S = P.S, v.par = par;
    v.list = P.a_list, v.nb = P.a_nb, v.rev = Pa.rev;
    if (fused) {
      // the direction update inside the application: p of the previous step in one buffer, the new p in
      // the other (the parity of the step picks them)
      v.x = par ? P.p2 : P.p, v.p_out = par ? P.p : P.p2;
      v.r = P.r, v.xh = P.x, v.part2 = P.part2, v.v_ib0 = P.v_ib0, v.a_ib0 = P.a_ib0, v.max_iter = P.max_iter;
    }
    mf_launch(op, fused ? MF_PCG_DIR : MF_PCG, v, stream, ev0, ev1);
  } else if (kf == FMT_BLK) {
    const unsigned grid = hw_blocks(P.a_nb);
    if (op.fmt.symmetric)
      launch_pcg(k_cg_spmv_blk<true>, grid, A.block, stream, Pa, par, ev0, ev1);
    else
      launch_pcg(k_cg_spmv_blk<false>, grid, A.block, stream, Pa, par, ev0, ev1);
  } else dispatch_items(apply_lane_rows(A.items), [&](auto IT) {
    constexpr int I = decltype(IT)::value;
    const unsigned grid = hw_blocks(P.a_nb);
    const int64_t lds = win_lds_bytes(op); // > 0: the p entries of a row block staged in LDS
    if (kf == FMT_DIA)
      launch_pcg(k_cg_spmv<FMT_DIA, I>, grid, A.block, stream, Pa, par, ev0, ev1);
    else if (kf == FMT_DIAS && lds > 0 && P.win)
      launch_pcg(k_cg_spmv_win, grid, A.block, stream, Pa, par, ev0, ev1, size_t(lds));
    else if (kf == FMT_DIAS)
      launch_pcg(k_cg_spmv<FMT_DIAS, I>, grid, A.block, stream, Pa, par, ev0, ev1);
    else
      launch_pcg(k_cg_spmv<FMT_SELL, I>, grid, A.block, stream, Pa, par, ev0, ev1);
  });
  dispatch_items(V.items, [&](auto IT) {
    constexpr int I = decltype(IT)::value;
    launch_pcg(k_cg_update<I>, hw_blocks(P.v_nb), V.block, stream, Pu, par, nullptr, ev_upd);
  });
}

void Operator::apply_inverse(const double* b, double* x, const hgdt_solver_options& opt, hgdt_solver_stats& st,
                             hipStream_t user_stream)
{
  SolverCache& C = cache_of(*this);
  if (opt.check_every < 1 || opt.max_iter < 1 || !(opt.rtol >= 0.))
    HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "invalid solver options");
  if (!square())
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "apply_inverse needs an operator whose source and range spaces are the same");
  if (dirichlet && !dirichlet_only_clear && !dirichlet_ensure_symmetry)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED,
               "Jacobi-PCG needs a symmetric operator: unit rows without their columns (ensure_symmetry false) make "
               "it non-symmetric");
  prepare_solver(user_stream);
  if (!fmt.diag_positive)
    HGDT_THROW(HGDT_ERR_SOLVER_FAILED, "Jacobi-PCG needs a positive diagonal: the matrix is not SPD");
  const int ninst = int(space->inst_dof.size()) - 1;
  d_r.allocate(std::max<int64_t>(n, 1));
  d_p.allocate(std::max<int64_t>(n, 1));
  d_p2.allocate(std::max<int64_t>(n, 1));
  d_q.allocate(std::max<int64_t>(n, 1));
  d_partials.allocate(std::max<int64_t>(fmt.A.nblocks + 3 * fmt.V.nblocks, 1));
  d_scalars.allocate(2 * ninst);
  if (C.host_capacity < ninst) {
    for (auto& h : C.host_scalars) {
      if (h)
        HGDT_HIP(hipHostFree(h));
      HGDT_HIP(hipHostMalloc((void**)&h, sizeof(PcgScalars) * ninst));
    }
    C.host_capacity = ninst;
  }
  for (auto& e : C.copied)
    if (!e)
      HGDT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!C.t0)
    HGDT_HIP(hipEventCreate(&C.t0));
  if (!C.t1)
    HGDT_HIP(hipEventCreate(&C.t1));
  // graphs cannot be captured on the legacy default stream: run on an owned stream joined to the caller's
  hipStream_t stream = user_stream;
  bool graph = opt.use_graph && !opt.time_spmv;
  if (graph) {
    if (!C.own_stream)
      HGDT_HIP(hipStreamCreateWithFlags(&C.own_stream, hipStreamNonBlocking));
    if (!C.join)
      HGDT_HIP(hipEventCreateWithFlags(&C.join, hipEventDisableTiming));
    HGDT_HIP(hipEventRecord(C.join, user_stream));
    HGDT_HIP(hipStreamWaitEvent(C.own_stream, C.join, 0));
    stream = C.own_stream;
  }
  PcgParams P = make_params(*this);
  P.b = b, P.x = x, P.r = d_r.ptr, P.p = d_p.ptr, P.p2 = d_p2.ptr, P.q = d_q.ptr;
  P.part1 = d_partials.ptr;
  P.part2 = d_partials.ptr + fmt.A.nblocks;
  P.part3 = d_partials.ptr + fmt.A.nblocks + 2 * fmt.V.nblocks;
  P.S = d_scalars.ptr;
  P.rtol2 = opt.rtol * opt.rtol, P.max_iter = opt.max_iter, P.zero_guess = opt.zero_initial_guess ? 1 : 0;
  HGDT_HIP(hipEventRecord(C.t0, stream));
  if (!opt.zero_initial_guess) {
    apply(x, d_q.ptr, stream);
    P.ax = d_q.ptr;
  }
  dispatch_items(fmt.V.items, [&](auto IT) {
    constexpr int I = decltype(IT)::value;
    hipLaunchKernelGGL((k_cg_init<I>), dim3(hw_blocks(fmt.V.nblocks)), dim3(fmt.V.block), 0, stream, P);
  });
  hipLaunchKernelGGL(k_cg_init_scalars, dim3(unsigned(ninst)), dim3(256), 0, stream, P);
  HGDT_HIP(hipGetLastError());
  P.ax = nullptr;

  // an even number of steps per chunk keeps the parity of every captured launch fixed and leaves the
  // current scalars in record 0 after each chunk; the last step of a solve runs in the chunk after the
  // one whose update met the tolerance (the stopping test is the first thing a step does)
  const int chunk = opt.check_every + (opt.check_every & 1);
  if (graph && (C.graph == nullptr || C.graph_x != x || C.graph_chunk != chunk || C.graph_version != C.version
                || C.graph_b != b || C.graph_maxit != opt.max_iter)) {
    if (C.graph)
      HGDT_HIP(hipGraphExecDestroy(C.graph));
    C.graph = nullptr;
    hipGraph_t g;
    HGDT_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < chunk; ++it)
      launch_iteration(*this, P, it & 1, stream, nullptr, nullptr);
    HGDT_HIP(hipStreamEndCapture(stream, &g));
    HGDT_HIP(hipGraphInstantiate(&C.graph, g, nullptr, nullptr, 0));
    HGDT_HIP(hipGraphDestroy(g));
    C.graph_x = x, C.graph_b = b, C.graph_chunk = chunk, C.graph_version = C.version;
    C.graph_maxit = opt.max_iter;
  }
  size_t nev = 0;
  int64_t issued = 0; // iterations launched so far
  std::vector<int64_t> timed_iter;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> timed_pair; // (start, end) of each timed application
  const int64_t stride = opt.time_spmv > 0 ? opt.time_spmv : 0;
  // timed iterations: issued % stride == phase, and the next one when stride > 2 (both parities: the
  // iterate is updated every second step), phase 1 when stride > 1 (the first iteration of a chunk follows
  // the convergence check's copy, which an interval starting at the previous iteration's end would include)
  const int64_t phase = stride > 1 ? 1 : 0;
  auto timed = [&](int64_t i) {
    if (!stride)
      return false;
    const int64_t m = i % stride;
    return m == phase || (stride > 2 && m == phase + 1);
  };
  const bool fused_dir = mf_fused_direction(*this);
  auto timing_event = [&]() {
    if (C.timing_events.size() < nev + 1) {
      hipEvent_t e;
      HGDT_HIP(hipEventCreate(&e));
      C.timing_events.push_back(e);
    }
    return C.timing_events[nev++];
  };
  hipEvent_t prev_upd = nullptr; // bound to the end of the previous iteration's update (fused application)
  auto launch_chunk = [&]() {
    if (graph) {
      HGDT_HIP(hipGraphLaunch(C.graph, stream));
      issued += chunk;
    } else {
      for (int it = 0; it < chunk; ++it, ++issued) {
        hipEvent_t e0 = nullptr, e1 = nullptr, eu = nullptr;
        if (timed(issued)) {
          timed_iter.push_back(issued);
          e1 = timing_event();
          // unfused: e0 is bound to the end of the direction update; fused: the previous update's end, or
          // the start of the first iteration's application
          e0 = (fused_dir && prev_upd) ? nullptr : timing_event();
          timed_pair.emplace_back(e0 ? e0 : prev_upd, e1);
        }
        if (fused_dir && timed(issued + 1))
          eu = timing_event();
        launch_iteration(*this, P, it & 1, stream, e0, e1, eu);
        prev_upd = eu;
      }
    }
    HGDT_HIP(hipGetLastError());
  };
  auto all_done = [&](const PcgScalars* h) {
    for (int i = 0; i < ninst; ++i)
      if (pcg_state(h[i]) == 0)
        return false;
    return true;
  };
  // Batches: once at least 15 % of the launched blocks belong to finished instances, later chunks are
  // launched over the blocks of the unfinished instances only (lists read through active_block), eagerly.
  // An instance found finished after chunk k did its last step in chunk k and the copy of its record to
  // the other parity in chunk k + 1 (queued before this check), so its blocks have nothing left to do from
  // chunk k + 2 on, the first chunk launched with the new lists.
  int slot = 0;
  int64_t compactions = 0;
  auto compact = [&](const PcgScalars* h) {
    const RowPartition& A = fmt.A;
    const RowPartition& V = fmt.V;
    int64_t na = 0, nv = 0;
    for (int i = 0; i < ninst; ++i)
      if (pcg_state(h[i]) == 0) {
        na += A.inst_block0[i + 1] - A.inst_block0[i];
        nv += V.inst_block0[i + 1] - V.inst_block0[i];
      }
    if (!(na < 0.85 * P.a_nb || nv < 0.85 * P.v_nb) || na == 0 || nv == 0)
      return;
    slot ^= 1;
    std::vector<int32_t>& L = C.active_host[slot];
    L.clear();
    for (int i = 0; i < ninst; ++i)
      if (pcg_state(h[i]) == 0)
        for (int32_t b = A.inst_block0[i]; b < A.inst_block0[i + 1]; ++b)
          L.push_back(b);
    for (int i = 0; i < ninst; ++i)
      if (pcg_state(h[i]) == 0)
        for (int32_t b = V.inst_block0[i]; b < V.inst_block0[i + 1]; ++b)
          L.push_back(b);
    DeviceArray<int32_t>& D = C.active[slot];
    D.allocate(size_t(A.nblocks + V.nblocks));
    HGDT_HIP(hipMemcpyAsync(D.ptr, L.data(), L.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    P.a_list = D.ptr, P.a_nb = int(na);
    P.v_list = D.ptr + na, P.v_nb = int(nv);
    graph = false;
    ++compactions;
  };
  // double-buffered convergence checks: chunk k+1 is queued before the flags of chunk k are read
  const int64_t max_chunks = (opt.max_iter + 1 + chunk - 1) / chunk + 1;
  int cur = 0;
  int64_t launched = 0;
  launch_chunk();
  ++launched;
  HGDT_HIP(hipMemcpyAsync(C.host_scalars[cur], d_scalars.ptr, sizeof(PcgScalars) * ninst, hipMemcpyDeviceToHost,
                          stream));
  HGDT_HIP(hipEventRecord(C.copied[cur], stream));
  while (true) {
    const int nxt = cur ^ 1;
    bool queued = false;
    if (launched < max_chunks) {
      launch_chunk();
      ++launched;
      HGDT_HIP(hipMemcpyAsync(C.host_scalars[nxt], d_scalars.ptr, sizeof(PcgScalars) * ninst,
                              hipMemcpyDeviceToHost, stream));
      HGDT_HIP(hipEventRecord(C.copied[nxt], stream));
      queued = true;
    }
    HGDT_HIP(hipEventSynchronize(C.copied[cur]));
    if (all_done(C.host_scalars[cur]) || !queued) {
      if (!queued)
        break;
      cur = nxt;
      HGDT_HIP(hipEventSynchronize(C.copied[cur]));
      break;
    }
    if (ninst > 1)
      compact(C.host_scalars[cur]);
    cur = nxt;
  }
  hipLaunchKernelGGL(k_cg_finish, dim3(blocks_for(n, 256)), dim3(256), 0, stream, n, fmt.d_dh.ptr, x);
  HGDT_HIP(hipGetLastError());
  // true residual |b - A x| / |b| of every instance at exit, with the unscaled operator (one application and
  // one pass over b and A x, over all blocks whatever the narrowed launches held)
  if (opt.true_residual) {
    if (C.host_true_capacity < ninst) {
      if (C.host_true)
        HGDT_HIP(hipHostFree(C.host_true));
      C.host_true = nullptr;
      HGDT_HIP(hipHostMalloc((void**)&C.host_true, sizeof(double) * ninst));
      C.host_true_capacity = ninst;
    }
    C.true_relres.allocate(size_t(ninst));
    apply(x, d_q.ptr, stream);
    PcgParams T = make_params(*this);
    T.b = b, T.ax = d_q.ptr, T.part3 = P.part3, T.S = P.S;
    dispatch_items(fmt.V.items, [&](auto IT) {
      constexpr int I = decltype(IT)::value;
      hipLaunchKernelGGL((k_true_residual<I>), dim3(hw_blocks(fmt.V.nblocks)), dim3(fmt.V.block), 0, stream, T);
    });
    hipLaunchKernelGGL(k_true_reduce, dim3(unsigned(ninst)), dim3(256), 0, stream, T, C.true_relres.ptr);
    HGDT_HIP(hipGetLastError());
    HGDT_HIP(hipMemcpyAsync(C.host_true, C.true_relres.ptr, sizeof(double) * ninst, hipMemcpyDeviceToHost, stream));
  }
  HGDT_HIP(hipEventRecord(C.t1, stream));
  HGDT_HIP(hipEventSynchronize(C.t1));
  if (graph) {
    HGDT_HIP(hipEventRecord(C.join, stream));
    HGDT_HIP(hipStreamWaitEvent(user_stream, C.join, 0));
  }
  const PcgScalars* h = C.host_scalars[cur];
  last_iters.assign(ninst, 0), last_relres.assign(ninst, 0.);
  last_true_relres.assign(ninst, opt.true_residual ? 0. : -1.);
  st = hgdt_solver_stats{};
  st.converged = 1;
  st.true_relres = opt.true_residual ? 0. : -1.;
  for (int i = 0; i < ninst; ++i) {
    last_iters[i] = h[i].iters;
    last_relres[i] = h[i].bb > 0. ? std::sqrt(h[i].rr / h[i].bb) : 0.;
    if (opt.true_residual) {
      last_true_relres[i] = C.host_true[i];
      if (!(C.host_true[i] <= st.true_relres)) // a NaN propagates
        st.true_relres = C.host_true[i];
    }
    st.iterations = std::max<int64_t>(st.iterations, h[i].iters);
    st.relres = std::max(st.relres, last_relres[i]);
    if (pcg_state(h[i]) != 1)
      st.converged = 0;
    if (pcg_state(h[i]) == 3)
      st.breakdown = 1;
  }
  float ms = 0.f;
  HGDT_HIP(hipEventElapsedTime(&ms, C.t0, C.t1));
  st.solve_ms = ms;
  if (opt.time_spmv && nev > 0) {
    // only iterations that did work: the early-exit launches after convergence are skipped
    double total = 0.;
    int64_t cnt = 0;
    for (size_t k = 0; k < timed_pair.size(); ++k) {
      if (timed_iter[k] < st.iterations) {
        float t = 0.f;
        HGDT_HIP(hipEventElapsedTime(&t, timed_pair[k].first, timed_pair[k].second));
        total += t, ++cnt;
      }
    }
    st.spmv_avg_ms = cnt ? total / cnt : 0.;
    st.spmv_count = cnt;
  }
  st.step_bytes = fmt.step_bytes;
  // one iteration: the application, the direction update (36 B per row on average over the two parities;
  // inside the application for the matrix-free SWIPDG operator) and the update (32 B per row)
  st.iteration_bytes = fmt.step_bytes + (fused_dir ? 0 : 36 * n) + 32 * n;
  st.compactions = compactions;
  st.column_staged = (kernel_format(*this) == FMT_DIAS && win_lds_bytes(*this) > 0 && C.win.ptr) ? 1 : 0;
  if (!st.converged) {
    if (st.breakdown)
      HGDT_THROW(HGDT_ERR_SOLVER_FAILED, "Jacobi-PCG breakdown: p.Ap <= 0 (matrix not SPD)");
    HGDT_THROW(HGDT_ERR_SOLVER_FAILED, "Jacobi-PCG did not reach the requested tolerance within max_iter");
  }
}

} // namespace hgdt
