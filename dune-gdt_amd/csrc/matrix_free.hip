// matrix_free.hip -- matrix-free operator application on structured grids (gfx950, FP64).
//
// Replaces the product with the assembled matrix (ConstMatrixOperator::apply,
// dune/gdt/operators/matrix-based.hh:121-129) by the grid walk of LocalizableOperator::apply
// (operators/localizable-operator.hh:352-379, local/assembler/operator-applicators.hh:82-87,219-227):
// range = sum over the elements and intersections of the local operators applied to the source,
// the local operators being those of the appended bilinear forms.  Nothing of the matrix is stored.
//
// On a structured grid every element of an instance has the same shape, so each local matrix is a
// coefficient-dependent scalar times a reference matrix:
//   * element forms: kappa_e times the unit-coefficient Laplace or product matrix
//     (local/integrands/laplace.hh:74-102, product.hh:104-130);
//   * SWIPDG blocks of an inner intersection (laplace-ipdg.hh:107-186, ipdg.hh:113-171): with
//     G_ab[i][j] = int (grad phi^b_j . n) phi^a_i and M_ab[i][j] = int phi^b_j phi^a_i over the face
//     (a, b in {inside, outside}),
//       in_in   = -s_c wm k_in (G_ii + s G_ii^T) + s_c p M_ii
//       in_out  =  s_c (-wp k_out G_io + s wm k_in G_oi^T) - s_c p M_io
//       out_in  =  s_c ( wm k_in G_oi - s wp k_out G_io^T) - s_c p M_io^T
//       out_out =  s_c wp k_out (G_oo + s G_oo^T) + s_c p M_oo
//     (s_c the scaling, s the symmetry prefactor, wm = d+ / (d+ + d-), wp = d- / (d+ + d-) from the
//     weights d = n . (omega n), p = sigma (d+ d- / (d+ + d-)) / h_F);
//   * boundary blocks (laplace-ipdg.hh:340-368, ipdg.hh:254-282): -s_c k (G + s G^T) and
//     s_c sigma omega / h_F M.
// The reference matrices are computed once per instance on the host, with the quadrature rules the
// assembly kernels use, so the application equals the assembled operator's up to the association of
// the sums.  Per application a kernel reads one coefficient per element and field, the source vector,
// and writes the range vector:
//   continuous Q1: one lane per DoF (vertex), summing over the 4 (8) cells around it; constrained DoFs
//     have unit rows and zero columns (tools/dirichlet-constraints.hh:153-184);
//   discontinuous Q1 on quadrilaterals: one lane per element and its 4 DoFs, the element matrix and the
//     blocks of its 4 intersections: an inner intersection is visited from the element with the smaller
//     index (InnerIntersectionsOnce), which is its inside.
// Modes: y = A x (apply); q = D^{-1/2} A D^{-1/2} p with the block partials of p.q (the CG solver's
// operator application, its row blocks being the solver's); and the diagonal, giving D^{-1/2} and D^{1/2}.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>

#include <hip/hip_ext.h>

#include "block_reduce.hpp"
#include "hgdt_internal.hpp"

namespace hgdt {
namespace {

constexpr int MF_MAXF = 8;
enum { MFK_CG_QUAD = 0, MFK_CG_HEX = 1, MFK_DG_QUAD = 2 };

struct MfParams {
  int ninst;
  // per row block: first DoF and first element of the instance, cells n0 n1 n2, instance, first and last
  // row of the block relative to the instance's first DoF
  const int64_t* blk;
  int64_t ref_stride; // doubles per instance of the reference data: h0 h1 1/h1 1/h0, element, face, boundary matrices
  int64_t eoff, coff, boff; // offsets of the three groups inside an instance's block
  int dirichlet;
  int nef;
  const double* ecoef[MF_MAXF];
  double evalue[MF_MAXF], escale[MF_MAXF];
  int ncf, nbf;
  int fkind[2][MF_MAXF]; // [0] inner, [1] boundary forms
  double fparam[2][MF_MAXF], fscale[2][MF_MAXF];
  const double* fkappa[2][MF_MAXF];
  double fkv[2][MF_MAXF];
  const double* fomega[2][MF_MAXF];
  double fov[2][MF_MAXF];
  int nb; // row blocks (the solver's operator partition): one row or element per lane
  MfVectors v;
};

// coefficient of a form at element e: the field's value, or the constant
__device__ __forceinline__ double cf(const double* __restrict__ f, double value, int64_t e)
{
  return f ? f[e] : value;
}

// y += s (M v) for the N x N row-major matrix M (T: s (M^T v)); DIAG: y_i += s M_ii
template <int N, bool T, bool DIAG>
__device__ __forceinline__ void mv(const double* __restrict__ M, double s, const double* v, double* y)
{
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if constexpr (DIAG) {
      y[i] = fma(s, M[i * N + i], y[i]);
    } else {
      double t = 0.;
#pragma unroll
      for (int j = 0; j < N; ++j)
        t = fma(T ? M[j * N + i] : M[i * N + j], v[j], t);
      y[i] = fma(s, t, y[i]);
    }
  }
}

// source entry: x (apply), D^{-1/2} p (solver)
template <int MODE>
__device__ __forceinline__ double src(const MfParams& P, int64_t i)
{
  if constexpr (MODE == MF_PCG)
    return P.v.dh[i] * P.v.x[i];
  else
    return P.v.x[i];
}

// result of row i: y (apply), q and the p.q partial (solver), D^{-1/2}, D^{1/2} (diagonal)
template <int MODE>
__device__ __forceinline__ void store_row(const MfParams& P, int64_t i, double y, double& part)
{
  if constexpr (MODE == MF_APPLY) {
    P.v.y[i] = y;
  } else if constexpr (MODE == MF_PCG) {
    const double q = P.v.dh[i] * y;
    P.v.y[i] = q;
    part += P.v.x[i] * q;
  } else {
    double d = y;
    if (!(d > 0.)) {
      P.v.flags[0] = 1;
      d = 1.;
    }
    const double sq = sqrt(d);
    P.v.out_dh[i] = 1. / sq;
    P.v.out_dsq[i] = sq;
  }
}

template <int MODE>
__device__ __forceinline__ void finish_block(const MfParams& P, int b, double part)
{
  if constexpr (MODE == MF_PCG) {
    double v[1] = {part};
    block_sum<1>(v);
    if (threadIdx.x == 0)
      P.v.part[b] = v[0];
  }
}

// the 4 rows i .. i + 3 of an element (discontinuous spaces, i a multiple of 4): raw source values xo and,
// for the solver, D^{-1/2} (ho); VEC: the vectors are 16-byte aligned, two 16-byte accesses per array
template <int MODE, bool VEC>
__device__ __forceinline__ void load4(const MfParams& P, int64_t i, double* xo, double* ho)
{
  if constexpr (VEC) {
    const double2* __restrict__ x = reinterpret_cast<const double2*>(P.v.x + i);
    const double2 x0 = x[0], x1 = x[1];
    xo[0] = x0.x, xo[1] = x0.y, xo[2] = x1.x, xo[3] = x1.y;
    if constexpr (MODE == MF_PCG) {
      const double2* __restrict__ h = reinterpret_cast<const double2*>(P.v.dh + i);
      const double2 h0 = h[0], h1 = h[1];
      ho[0] = h0.x, ho[1] = h0.y, ho[2] = h1.x, ho[3] = h1.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      xo[k] = P.v.x[i + k];
      if constexpr (MODE == MF_PCG)
        ho[k] = P.v.dh[i + k];
    }
  }
}

template <bool VEC>
__device__ __forceinline__ void store4(double* __restrict__ out, int64_t i, const double* v)
{
  if constexpr (VEC) {
    double2* o = reinterpret_cast<double2*>(out + i);
    o[0] = make_double2(v[0], v[1]);
    o[1] = make_double2(v[2], v[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      out[i + k] = v[k];
  }
}

// continuous Q1 on quadrilaterals (D = 2) or hexahedra (D = 3): one lane per DoF.  Every load of a lane
// (the source values of the 3^D neighbourhood, the coefficients of the 2^D cells around the vertex) is
// issued before the solver's convergence flag is examined.  NEF >= 0: the number of element forms is known
// at compile time; -1: taken from the parameters.
template <int D, int MODE, int NEF>
__global__ void __launch_bounds__(256) k_mf_cg(MfParams P, const double* __restrict__ ref)
{
  constexpr int NV = 1 << D, NS = (D == 2) ? 9 : 27;
  constexpr bool DIAG = (MODE == MF_DIAG);
  constexpr int EC = NEF > 0 ? NEF : 1;
  const int b = logical_block(P.nb);
  if (b < 0)
    return;
  const int64_t* __restrict__ B = P.blk + 8 * int64_t(b);
  const int64_t d0 = B[0], e0 = B[1];
  const uint32_t nx = uint32_t(B[2]), ny = uint32_t(B[3]), nz = uint32_t(B[4]);
  const int inst = int(B[5]);
  const uint32_t r1 = uint32_t(B[7]);
  const uint32_t v = uint32_t(B[6]) + threadIdx.x; // row of the instance
  const bool row = v < r1;
  const uint32_t vx = nx + 1, vy = ny + 1;
  const int64_t sy = vx, sz = int64_t(vx) * vy;
  const uint32_t ix = v % vx, iy = (v / vx) % vy, iz = (D == 3) ? v / (vx * vy) : 0u;
  auto boundary = [&](uint32_t x, uint32_t y, uint32_t z) {
    return x == 0u || x == nx || y == 0u || y == ny || (D == 3 && (z == 0u || z == nz));
  };
  const bool fixed = P.dirichlet && boundary(ix, iy, iz); // unit row of a constrained DoF
  const int64_t i = d0 + v;
  // source on the 3^D neighbourhood; zero outside the instance and on constrained DoFs (zero columns)
  double s[NS];
#pragma unroll
  for (int t = 0; t < NS; ++t) {
    s[t] = 0.;
    if constexpr (!DIAG) {
      const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = (D == 3) ? t / 9 - 1 : 0;
      const uint32_t jx = ix + dx, jy = iy + dy, jz = iz + dz; // wraps to > n below zero
      const bool in = jx <= nx && jy <= ny && (D == 2 || jz <= nz);
      const bool self = dx == 0 && dy == 0 && dz == 0;
      if (row && in && (self || !fixed) && !(P.dirichlet && !self && boundary(jx, jy, jz)))
        s[t] = src<MODE>(P, i + dx + dy * sy + dz * sz);
    }
  }
  // cell c = (cx, cy, cz) in {0, 1}^D is the cell (ix - 1 + cx, ...), holding the vertex as local vertex
  // l = (1 - cx) + 2 (1 - cy) + 4 (1 - cz)
  bool valid[NV];
  int64_t cell[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    const uint32_t ex = ix - 1 + cx, ey = iy - 1 + cy, ez = (D == 3) ? iz - 1 + cz : 0u;
    valid[c] = row && !fixed && ex < nx && ey < ny && (D == 2 || ez < nz);
    cell[c] = e0 + (int64_t(ez) * ny + ey) * nx + ex;
  }
  double kc[EC][NV];
  if constexpr (NEF >= 0) {
#pragma unroll
    for (int f = 0; f < NEF; ++f)
#pragma unroll
      for (int c = 0; c < NV; ++c)
        kc[f][c] = valid[c] ? cf(P.ecoef[f], P.evalue[f], cell[c]) : 0.;
  }
  if (MODE == MF_PCG && P.v.S[(P.v.par ^ 1) * P.ninst + inst].done)
    return;
  const double* __restrict__ R = ref + int64_t(inst) * P.ref_stride + P.eoff;
  double y = 0.;
  if (fixed) {
    y = DIAG ? 1. : s[NS / 2];
  } else {
#pragma unroll
    for (int f = 0; f < (NEF >= 0 ? NEF : P.nef); ++f) {
      const double sc = P.escale[f];
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        if (!valid[c])
          continue;
        const double k = (NEF >= 0) ? kc[NEF >= 0 ? f : 0][c] : cf(P.ecoef[f], P.evalue[f], cell[c]);
        const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const int l = (1 - cx) | ((1 - cy) << 1) | ((D == 3 ? (1 - cz) : 0) << 2);
        const double* __restrict__ rw = R + f * NV * NV + l * NV;
        if constexpr (DIAG) {
          y = fma(sc * k, rw[l], y);
        } else {
          double t = 0.;
#pragma unroll
          for (int j = 0; j < NV; ++j) {
            const int tx = cx + (j & 1), ty = cy + ((j >> 1) & 1), tz = (D == 3) ? cz + (j >> 2) : 0;
            t = fma(rw[j], s[tx + 3 * ty + 9 * tz], t);
          }
          y = fma(sc * k, t, y);
        }
      }
    }
  }
  double part = 0.;
  if (row)
    store_row<MODE>(P, i, y, part);
  finish_block<MODE>(P, b, part);
}

// Structural zeros of the Q1 face matrices.  On a face normal to x (o = 0) the inside element's shape
// functions 1, 3 (reference x = 1) and the outside element's 0, 2 do not vanish; normal to y (o = 1) the
// inside's 2, 3 and the outside's 0, 1.  A matrix entry (i, j) is bit 4 i + j of a mask.
constexpr uint32_t face_dofs(int o, bool inside)
{
  return o == 0 ? (inside ? 0xAu : 0x5u) : (inside ? 0xCu : 0x3u);
}
constexpr uint32_t mask_of(uint32_t rows, uint32_t cols, bool either)
{
  uint32_t m = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const bool r = (rows >> i) & 1, c = (cols >> j) & 1;
      if (either ? (r || c) : (r && c))
        m |= 1u << (4 * i + j);
    }
  return m;
}
// masks of the 6 coupling slots (A_in, G_io, G_oi^T, A_out, G_oi, G_io^T) and the 4 penalty slots (M_ii,
// M_io, M_io^T, M_oo) of orientation o, and of the boundary matrices of side sd (coupling / penalty)
constexpr uint32_t coupling_mask(int o, int slot)
{
  const uint32_t fi = face_dofs(o, true), fo = face_dofs(o, false);
  switch (slot) {
    case 0: return mask_of(fi, fi, true);
    case 1: return mask_of(fi, 0xFu, false);
    case 2: return mask_of(0xFu, fo, false);
    case 3: return mask_of(fo, fo, true);
    case 4: return mask_of(fo, 0xFu, false);
    default: return mask_of(0xFu, fi, false);
  }
}
constexpr uint32_t penalty_mask(int o, int slot)
{
  const uint32_t fi = face_dofs(o, true), fo = face_dofs(o, false);
  switch (slot) {
    case 0: return mask_of(fi, fi, false);
    case 1: return mask_of(fi, fo, false);
    case 2: return mask_of(fo, fi, false);
    default: return mask_of(fo, fo, false);
  }
}
constexpr uint32_t boundary_dofs(int sd)
{
  return face_dofs(sd >> 1, sd & 1);
}

// y += s (M v) over the entries of MASK (DIAG: y_i += s M_ii)
template <uint32_t MASK, bool DIAG>
__device__ __forceinline__ void mvm(const double* __restrict__ M, double s, const double* v, double* y)
{
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if constexpr (DIAG) {
      if ((MASK >> (5 * i)) & 1u)
        y[i] = fma(s, M[5 * i], y[i]);
    } else {
      if (((MASK >> (4 * i)) & 0xFu) == 0u)
        continue;
      double t = 0.;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if ((MASK >> (4 * i + j)) & 1u)
          t = fma(M[4 * i + j], v[j], t);
      y[i] = fma(s, t, y[i]);
    }
  }
}

// 1 / x for x > 0: hardware reciprocal and two Newton steps (correct to round-off)
__device__ __forceinline__ double recip(double x)
{
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.), r, r);
  return fma(fma(-x, r, 1.), r, r);
}

// the inner forms' contributions of side SD (0 x-, 1 x+, 2 y-, 3 y+) to the rows of element e
template <int SD, int MODE, int NCF>
__device__ __forceinline__ void dg_side(const MfParams& P, const double* __restrict__ R, int ncf, int64_t e, int64_t en,
                                        const double* a, const double* bv, double* y)
{
  constexpr bool DIAG = (MODE == MF_DIAG);
  constexpr int o = SD >> 1;
  constexpr bool inside = SD & 1; // the neighbour across x+ / y+ has the larger index: this element is inside
  const double ilen = R[2 + o];
#pragma unroll
  for (int f = 0; f < (NCF >= 0 ? NCF : ncf); ++f) {
    const double sc = P.fscale[0][f], prm = P.fparam[0][f];
    const double* __restrict__ of = P.fomega[0][f];
    const double* __restrict__ G = R + P.coff + (2 * f + o) * 96;
    const double os = cf(of, P.fov[0][f], e), on = cf(of, P.fov[0][f], en);
    // weights of the intersection from its inside (-) and outside (+) elements
    const double dm = inside ? os : on, dp = inside ? on : os;
    const double inv = recip(dp + dm);
    if (P.fkind[0][f] == FF_INNER_COUPLING) {
      const double* __restrict__ kf = P.fkappa[0][f];
      const double ks = cf(kf, P.fkv[0][f], e), kn = cf(kf, P.fkv[0][f], en);
      const double kin = inside ? ks : kn, kout = inside ? kn : ks;
      const double wm = dp * inv, wp = dm * inv;
      if constexpr (inside) {
        mvm<coupling_mask(o, 0), DIAG>(G, -sc * wm * kin, a, y);
        if constexpr (!DIAG) {
          mvm<coupling_mask(o, 1), false>(G + 16, -sc * wp * kout, bv, y);
          mvm<coupling_mask(o, 2), false>(G + 32, sc * prm * wm * kin, bv, y);
        }
      } else {
        mvm<coupling_mask(o, 3), DIAG>(G + 48, sc * wp * kout, a, y);
        if constexpr (!DIAG) {
          mvm<coupling_mask(o, 4), false>(G + 64, sc * wm * kin, bv, y);
          mvm<coupling_mask(o, 5), false>(G + 80, -sc * prm * wp * kout, bv, y);
        }
      }
    } else {
      const double pen = sc * (prm * (dp * dm * inv)) * ilen;
      if constexpr (inside) {
        mvm<penalty_mask(o, 0), DIAG>(G, pen, a, y);
        if constexpr (!DIAG)
          mvm<penalty_mask(o, 1), false>(G + 16, -pen, bv, y);
      } else {
        mvm<penalty_mask(o, 3), DIAG>(G + 48, pen, a, y);
        if constexpr (!DIAG)
          mvm<penalty_mask(o, 2), false>(G + 32, -pen, bv, y);
      }
    }
  }
}

// the boundary forms' contributions of side SD
template <int SD, int MODE, int NBF>
__device__ __forceinline__ void dg_boundary(const MfParams& P, const double* __restrict__ R, int nbf, int64_t e,
                                            const double* a, double* y)
{
  constexpr bool DIAG = (MODE == MF_DIAG);
  const double ilen = R[2 + (SD >> 1)];
#pragma unroll
  for (int f = 0; f < (NBF >= 0 ? NBF : nbf); ++f) {
    const double sc = P.fscale[1][f], prm = P.fparam[1][f];
    const double* __restrict__ B = R + P.boff + (4 * f + SD) * 16;
    if (P.fkind[1][f] == FF_DIRICHLET_COUPLING) // -s_c k (G + s G^T), stored combined
      mvm<mask_of(boundary_dofs(SD), boundary_dofs(SD), true), DIAG>(B, -sc * cf(P.fkappa[1][f], P.fkv[1][f], e), a, y);
    else // s_c sigma omega / h_F M
      mvm<mask_of(boundary_dofs(SD), boundary_dofs(SD), false), DIAG>(
          B, sc * (prm * cf(P.fomega[1][f], P.fov[1][f], e)) * ilen, a, y);
  }
}

// discontinuous Q1 on quadrilaterals: a block works on a tile of TX x TY elements, one lane per element (its
// 4 consecutive DoFs).  The source values of the tile and of a one-element halo (D^{-1/2} p for the solver)
// are first staged in LDS with coalesced loads (the 4-DoF vectors of a row of elements are contiguous), so
// each lane reads its own and its 4 neighbours' values from LDS.  Sides: 0 x-, 1 x+, 2 y-, 3 y+; across
// x+ / y+ the neighbour has the larger index (this element is the inside of that intersection), across
// x- / y- the smaller (the outside).  Face matrices per form and orientation (host set-up): coupling A_in =
// G_ii + s G_ii^T, G_io, G_oi^T, A_out = G_oo + s G_oo^T, G_oi, G_io^T; penalty M_ii, M_io, M_io^T, M_oo;
// only their structurally nonzero entries are read.  NEF, NCF, NBF >= 0: form counts known at compile time
// (loops unrolled); -1: runtime loops.  VEC: the vectors are 16-byte aligned (16-byte accesses).
constexpr int TX = 64, TY = 4;

template <int MODE, int NEF, int NCF, int NBF, bool VEC>
__global__ void __launch_bounds__(TX * TY) k_mf_dg(MfParams P, const double* __restrict__ ref)
{
  constexpr bool DIAG = (MODE == MF_DIAG);
  __shared__ __align__(16) double sp[TY + 2][TX + 2][4]; // source of the tile and its halo
  const int nef = NEF >= 0 ? NEF : P.nef, ncf = NCF >= 0 ? NCF : P.ncf, nbf = NBF >= 0 ? NBF : P.nbf;
  const int b = logical_block(P.nb);
  if (b < 0)
    return;
  const int64_t* __restrict__ Bk = P.blk + 8 * int64_t(b);
  const int64_t d0 = Bk[0], e0 = Bk[1];
  const uint32_t nx = uint32_t(Bk[2]), ny = uint32_t(Bk[3]);
  const int inst = int(Bk[5]);
  const uint32_t x0 = uint32_t(Bk[6]), y0 = uint32_t(Bk[7]);
  if (MODE == MF_PCG && P.v.S[(P.v.par ^ 1) * P.ninst + inst].done)
    return;
  const uint32_t tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const uint32_t ex = x0 + tx, ey = y0 + ty;
  const bool row = ex < nx && ey < ny;
  if constexpr (!DIAG) {
    // staged columns x0 + xs .. x0 + xe - 1 of the rows y0 - 1 .. y0 + TY (clipped to the instance)
    const int xs = x0 > 0u ? -1 : 0;
    const int xe = int(min(uint32_t(TX + 1), nx - x0));
    constexpr int U = VEC ? 2 : 4; // accesses per element and array
    const int per_row = (xe - xs) * U;
    for (int u = threadIdx.x; u < (TY + 2) * per_row; u += TX * TY) {
      const int r = u / per_row, w = u % per_row;
      const int64_t yy = int64_t(y0) - 1 + r;
      if (yy < 0 || yy >= int64_t(ny))
        continue;
      const int col = xs + w / U; // relative to x0
      const int64_t dof = d0 + 4 * (yy * nx + x0 + col) + (4 / U) * (w % U);
      double* dst = &sp[r][col + 1][(4 / U) * (w % U)];
      if constexpr (VEC) {
        double2 xv = *reinterpret_cast<const double2*>(P.v.x + dof);
        if constexpr (MODE == MF_PCG) {
          const double2 hv = *reinterpret_cast<const double2*>(P.v.dh + dof);
          xv.x *= hv.x, xv.y *= hv.y;
        }
        *reinterpret_cast<double2*>(dst) = xv;
      } else {
        dst[0] = src<MODE>(P, dof);
      }
    }
    __syncthreads();
  }
  const double* __restrict__ R = ref + int64_t(inst) * P.ref_stride;
  const uint32_t c = ey * nx + ex;
  const int64_t e = e0 + c;
  const bool has[4] = {row && ex > 0u, row && ex + 1u < nx, row && ey > 0u, row && ey + 1u < ny};
  double a[4], bv[4][4], y[4] = {0., 0., 0., 0.};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = (!DIAG && row) ? sp[ty + 1][tx + 1][k] : 0.;
    bv[0][k] = (!DIAG && has[0]) ? sp[ty + 1][tx][k] : 0.;
    bv[1][k] = (!DIAG && has[1]) ? sp[ty + 1][tx + 2][k] : 0.;
    bv[2][k] = (!DIAG && has[2]) ? sp[ty][tx + 1][k] : 0.;
    bv[3][k] = (!DIAG && has[3]) ? sp[ty + 2][tx + 1][k] : 0.;
  }
  if (row) {
#pragma unroll
    for (int f = 0; f < nef; ++f)
      mv<4, false, DIAG>(R + P.eoff + 16 * f, P.escale[f] * cf(P.ecoef[f], P.evalue[f], e), a, y);
    if (has[0])
      dg_side<0, MODE, NCF>(P, R, ncf, e, e - 1, a, bv[0], y);
    else
      dg_boundary<0, MODE, NBF>(P, R, nbf, e, a, y);
    if (has[1])
      dg_side<1, MODE, NCF>(P, R, ncf, e, e + 1, a, bv[1], y);
    else
      dg_boundary<1, MODE, NBF>(P, R, nbf, e, a, y);
    if (has[2])
      dg_side<2, MODE, NCF>(P, R, ncf, e, e - nx, a, bv[2], y);
    else
      dg_boundary<2, MODE, NBF>(P, R, nbf, e, a, y);
    if (has[3])
      dg_side<3, MODE, NCF>(P, R, ncf, e, e + nx, a, bv[3], y);
    else
      dg_boundary<3, MODE, NBF>(P, R, nbf, e, a, y);
  }
  double part = 0.;
  if (row) {
    const int64_t i0 = d0 + 4 * int64_t(c);
    if constexpr (MODE == MF_APPLY) {
      store4<VEC>(P.v.y, i0, y);
    } else if constexpr (MODE == MF_PCG) {
      // q = D^{-1/2} y; p.q = (D^{-1/2} p).y
      double xo[4], ho[4];
      load4<MODE, VEC>(P, i0, xo, ho);
      double q[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        q[k] = ho[k] * y[k];
        part += xo[k] * q[k];
      }
      store4<VEC>(P.v.y, i0, q);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        store_row<MODE>(P, i0 + k, y[k], part);
    }
  }
  finish_block<MODE>(P, b, part);
}

// ---------------------------------------------------------------------------------------------------
// host set-up
// ---------------------------------------------------------------------------------------------------

// Q1 shape functions on the reference cube: phi_b = prod_a (bit a of b ? x_a : 1 - x_a)
void q1_shape(int D, const double* x, double* phi, double (*dphi)[3])
{
  for (int bb = 0; bb < (1 << D); ++bb) {
    phi[bb] = 1.;
    for (int a = 0; a < D; ++a) {
      const bool hi = (bb >> a) & 1;
      phi[bb] *= hi ? x[a] : 1. - x[a];
      double g = hi ? 1. : -1.;
      for (int c = 0; c < D; ++c)
        if (c != a)
          g *= ((bb >> c) & 1) ? x[c] : 1. - x[c];
      dphi[bb][a] = g;
    }
  }
}

[[noreturn]] void unsupported(const std::string& why)
{
  HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "matrix-free application: " + why);
}

unsigned hw_grid(int64_t nb)
{
  return unsigned(((nb + 7) / 8) * 8);
}

} // namespace

void mf_setup(Operator& op)
{
  const Space& s = *op.space;
  const Grid& g = *s.grid;
  SolverFormat& F = op.fmt;
  if (g.lattice.empty() || int(g.lattice.size()) != g.ninst())
    unsupported("needs a structured grid or a union of structured grids");
  if (g.ref.geom != HGDT_QUAD && g.ref.geom != HGDT_HEX)
    unsupported("provided on grids of quadrilaterals or hexahedra");
  if (s.order != 1)
    unsupported("provided for order-1 spaces");
  const bool dg = s.kind == HGDT_DISCONTINUOUS_LAGRANGE;
  if (dg && g.ref.geom != HGDT_QUAD)
    unsupported("discontinuous spaces are provided on quadrilateral grids");
  if (!dg && (!op.cforms.empty() || !op.bforms.empty()))
    unsupported("intersection forms are provided on discontinuous spaces");
  if (dg && op.dirichlet)
    unsupported("Dirichlet constraints are provided on continuous spaces");
  if (op.eforms.size() > size_t(MF_MAXF) || op.cforms.size() > size_t(MF_MAXF) || op.bforms.size() > size_t(MF_MAXF))
    unsupported("too many forms appended");
  const int D = g.ref.dim, NV = 1 << D;
  const int ninst = g.ninst();

  // the DoF numbering (and for discontinuous spaces the neighbours) must be those of the lattice
  std::atomic<bool> ok{true};
  for (int in = 0; in < ninst; ++in) {
    const Lattice& L = g.lattice[in];
    const int64_t e0 = g.inst_elem[in], d0 = s.inst_dof[in];
    const int64_t n0 = L.n[0], n1 = L.n[1], n2 = (D == 3) ? L.n[2] : 1;
    const int64_t ncell = n0 * n1 * n2;
    const int64_t ndof = dg ? 4 * ncell : (n0 + 1) * (n1 + 1) * (D == 3 ? n2 + 1 : 1);
    if (g.inst_elem[in + 1] - e0 != ncell || s.inst_dof[in + 1] - d0 != ndof)
      unsupported("the grid does not match its lattice");
    parallel_for(ncell, [&](int64_t c0, int64_t c1) {
      for (int64_t c = c0; c < c1 && ok; ++c) {
        const int64_t i = c % n0, j = (c / n0) % n1, k = c / (n0 * n1);
        for (int bb = 0; bb < NV; ++bb) {
          const int64_t want = dg ? d0 + 4 * c + bb
                                  : d0 + ((k + ((bb >> 2) & 1)) * (n1 + 1) + j + ((bb >> 1) & 1)) * (n0 + 1) + i + (bb & 1);
          if (s.dofs[(e0 + c) * NV + bb] != want)
            ok = false;
        }
        if (dg) {
          const int64_t e = e0 + c;
          const int64_t want[4] = {i > 0 ? e - 1 : -1, i < n0 - 1 ? e + 1 : -1, j > 0 ? e - n0 : -1,
                                   j < n1 - 1 ? e + n0 : -1};
          for (int f = 0; f < 4; ++f)
            if (g.nb[e * 4 + f] != want[f])
              ok = false;
        }
      }
    });
  }
  if (!ok)
    unsupported("the DoF numbering or the neighbours differ from the lattice's");
  if (op.dirichlet) {
    // the constrained DoFs must be the boundary vertices of every lattice
    op.build_dirichlet();
    int64_t nbnd = 0;
    for (int in = 0; in < ninst; ++in) {
      const Lattice& L = g.lattice[in];
      int64_t inner = 1, all = 1;
      for (int a = 0; a < D; ++a)
        inner *= std::max<int64_t>(L.n[a] - 1, 0), all *= L.n[a] + 1;
      nbnd += all - inner;
    }
    bool same = int64_t(op.dirichlet_dofs.size()) == nbnd;
    for (size_t t = 0; t < op.dirichlet_dofs.size() && same; ++t) {
      const int64_t dof = op.dirichlet_dofs[t];
      const int in = int(std::upper_bound(s.inst_dof.begin(), s.inst_dof.end(), dof) - s.inst_dof.begin()) - 1;
      const Lattice& L = g.lattice[in];
      int64_t v = dof - s.inst_dof[in];
      bool bnd = false;
      for (int a = 0; a < D; ++a) {
        const int64_t idx = v % (L.n[a] + 1);
        v /= (L.n[a] + 1);
        bnd = bnd || idx == 0 || idx == L.n[a];
      }
      same = bnd;
    }
    if (!same)
      unsupported("the constrained DoFs are not the boundary vertices");
  }

  // reference matrices per instance, with the assembly kernels' quadrature rules (element rules of
  // order 0 + 1 + 1 + over_integrate, face rules likewise: integrals.hh:115, laplace.hh:74-79,
  // product.hh:88-100, laplace-ipdg.hh:95-105, ipdg.hh:100-109,245-250)
  const int nef = int(op.eforms.size()), ncf = int(op.cforms.size()), nbf = int(op.bforms.size());
  const int64_t eoff = 4, coff = eoff + int64_t(nef) * NV * NV, boff = coff + int64_t(ncf) * 2 * 96;
  const int64_t stride = boff + int64_t(nbf) * 4 * 16;
  std::vector<double> ref(size_t(stride) * ninst, 0.);
  for (int in = 0; in < ninst; ++in) {
    const Lattice& L = g.lattice[in];
    double* r = &ref[size_t(in) * stride];
    const double h[3] = {L.h[0], L.h[1], D == 3 ? L.h[2] : 1.};
    // header: h0, h1 and the inverse face lengths 1 / h1 (faces normal to x), 1 / h0 (normal to y)
    r[0] = h[0], r[1] = h[1], r[2] = 1. / h[1], r[3] = 1. / h[0];
    const double vol = h[0] * h[1] * (D == 3 ? h[2] : 1.);
    for (int f = 0; f < nef; ++f) {
      const auto& E = op.eforms[f];
      const Quadrature q = element_rule(g.ref, 0 + 1 + 1 + E.over_integrate);
      double* M = r + eoff + int64_t(f) * NV * NV;
      for (int p = 0; p < q.size(); ++p) {
        double phi[8], dphi[8][3];
        q1_shape(D, &q.x[size_t(p) * D], phi, dphi);
        const double w = q.w[p] * vol;
        for (int i = 0; i < NV; ++i)
          for (int j = 0; j < NV; ++j) {
            double v = 0.;
            if (E.kind == EF_LAPLACE)
              for (int a = 0; a < D; ++a)
                v += (dphi[j][a] / h[a]) * (dphi[i][a] / h[a]);
            else
              v = phi[i] * phi[j];
            M[i * NV + j] += w * v;
          }
      }
    }
    // faces: orientation o (0: normal +x, 1: normal +y); the inside element holds the face at reference
    // coordinate 1 in direction o, the outside element at 0; t runs along the other direction
    for (int f = 0; f < ncf; ++f) {
      const auto& C = op.cforms[f];
      const Quadrature q = line_rule(0 + 1 + 1 + C.over_integrate);
      for (int o = 0; o < 2; ++o) {
        // face integrals G_ii, G_io, G_oi, G_oo (coupling) or M_ii, M_io, M_oo (penalty)
        double I[4][16] = {};
        const double flen = h[1 - o];
        for (int p = 0; p < q.size(); ++p) {
          double xin[2], xout[2];
          xin[o] = 1., xout[o] = 0., xin[1 - o] = xout[1 - o] = q.x[p];
          double pin[8], pout[8], din[8][3], dout[8][3];
          q1_shape(2, xin, pin, din);
          q1_shape(2, xout, pout, dout);
          const double w = flen * q.w[p];
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
              if (C.kind == FF_INNER_COUPLING) {
                const double gin_j = din[j][o] / h[o], gout_j = dout[j][o] / h[o];
                I[0][i * 4 + j] += w * (gin_j * pin[i]);
                I[1][i * 4 + j] += w * (gout_j * pin[i]);
                I[2][i * 4 + j] += w * (gin_j * pout[i]);
                I[3][i * 4 + j] += w * (gout_j * pout[i]);
              } else {
                I[0][i * 4 + j] += w * (pin[j] * pin[i]);
                I[1][i * 4 + j] += w * (pout[j] * pin[i]);
                I[3][i * 4 + j] += w * (pout[j] * pout[i]);
              }
            }
        }
        double* G = r + coff + int64_t(2 * f + o) * 96;
        const double s_ = C.param;
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) {
            const int ij = i * 4 + j, ji = j * 4 + i;
            if (C.kind == FF_INNER_COUPLING) {
              G[0 + ij] = I[0][ij] + s_ * I[0][ji]; // A_in
              G[16 + ij] = I[1][ij];                // G_io
              G[32 + ij] = I[2][ji];                // G_oi^T
              G[48 + ij] = I[3][ij] + s_ * I[3][ji]; // A_out
              G[64 + ij] = I[2][ij];                // G_oi
              G[80 + ij] = I[1][ji];                // G_io^T
            } else {
              G[0 + ij] = I[0][ij];  // M_ii
              G[16 + ij] = I[1][ij]; // M_io
              G[32 + ij] = I[1][ji]; // M_io^T
              G[48 + ij] = I[3][ij]; // M_oo
            }
          }
      }
    }
    // boundary sides 0 x-, 1 x+, 2 y-, 3 y+ with the outward normal
    for (int f = 0; f < nbf; ++f) {
      const auto& Bd = op.bforms[f];
      const Quadrature q = line_rule(0 + 1 + 1 + Bd.over_integrate);
      for (int side = 0; side < 4; ++side) {
        const int o = side >> 1;
        const double sign = (side & 1) ? 1. : -1.;
        double I[16] = {};
        const double flen = h[1 - o];
        for (int p = 0; p < q.size(); ++p) {
          double x[2];
          x[o] = (side & 1) ? 1. : 0., x[1 - o] = q.x[p];
          double phi[8], dphi[8][3];
          q1_shape(2, x, phi, dphi);
          const double w = flen * q.w[p];
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
              I[i * 4 + j] += (Bd.kind == FF_DIRICHLET_COUPLING) ? w * ((sign * dphi[j][o] / h[o]) * phi[i])
                                                                : w * (phi[j] * phi[i]);
        }
        // coupling: G + s G^T (G_ij = int (grad phi_j . n) phi_i); penalty: M
        double* B = r + boff + int64_t(4 * f + side) * 16;
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j)
            B[i * 4 + j] = (Bd.kind == FF_DIRICHLET_COUPLING) ? I[i * 4 + j] + Bd.param * I[j * 4 + i] : I[i * 4 + j];
      }
    }
  }
  // the kernels read only the structurally nonzero entries of the face matrices: the others must vanish
  for (int in = 0; in < ninst && dg; ++in) {
    const double* r = &ref[size_t(in) * stride];
    for (int f = 0; f < ncf; ++f)
      for (int o = 0; o < 2; ++o)
        for (int slot = 0; slot < 6; ++slot) {
          const bool coupling = op.cforms[f].kind == FF_INNER_COUPLING;
          if (!coupling && slot >= 4)
            continue;
          const uint32_t m = coupling ? coupling_mask(o, slot) : penalty_mask(o, slot);
          for (int t = 0; t < 16; ++t)
            if (!((m >> t) & 1u) && r[coff + (2 * f + o) * 96 + slot * 16 + t] != 0.)
              HGDT_THROW(HGDT_ERR_OPERATOR, "matrix-free set-up: face matrix entry outside the Q1 face pattern");
        }
    for (int f = 0; f < nbf; ++f)
      for (int sd = 0; sd < 4; ++sd) {
        const uint32_t m = mask_of(boundary_dofs(sd), boundary_dofs(sd), op.bforms[f].kind == FF_DIRICHLET_COUPLING);
        for (int t = 0; t < 16; ++t)
          if (!((m >> t) & 1u) && r[boff + (4 * f + sd) * 16 + t] != 0.)
            HGDT_THROW(HGDT_ERR_OPERATOR, "matrix-free set-up: boundary matrix entry outside the Q1 face pattern");
      }
  }
  F.mf_kind = dg ? MFK_DG_QUAD : (D == 2 ? MFK_CG_QUAD : MFK_CG_HEX);
  F.mf_ref_stride = stride;
  F.d_mf_ref.upload(ref);
  // distinct coefficient arrays read per application
  std::vector<const double*> fields;
  for (const auto& E : op.eforms)
    fields.push_back(E.field);
  for (const auto* forms : {&op.cforms, &op.bforms})
    for (const auto& Fm : *forms)
      fields.push_back(Fm.kind == FF_INNER_COUPLING || Fm.kind == FF_DIRICHLET_COUPLING ? Fm.kappa : nullptr),
          fields.push_back(Fm.kind == FF_DIRICHLET_COUPLING ? nullptr : Fm.omega);
  fields.erase(std::remove(fields.begin(), fields.end(), nullptr), fields.end());
  std::sort(fields.begin(), fields.end());
  F.mf_fields = int(std::unique(fields.begin(), fields.end()) - fields.begin());
}

void mf_block_table(Operator& op)
{
  const Space& s = *op.space;
  const Grid& g = *s.grid;
  const RowPartition& A = op.fmt.A;
  if (int(g.lattice.size()) != g.ninst())
    return; // not a structured grid: mf_setup refuses the operator
  const int D = g.ref.dim;
  const bool tiles = s.kind == HGDT_DISCONTINUOUS_LAGRANGE;
  std::vector<int64_t> blk(8 * size_t(A.nblocks));
  for (int64_t b = 0; b < A.nblocks; ++b) {
    const int in = A.inst[b];
    const Lattice& L = g.lattice[in];
    int64_t* r = &blk[8 * size_t(b)];
    r[0] = s.inst_dof[in], r[1] = g.inst_elem[in];
    r[2] = L.n[0], r[3] = L.n[1], r[4] = (D == 3) ? L.n[2] : 1;
    r[5] = in;
    if (tiles) { // tile origin (element column and row)
      const int64_t t = b - A.inst_block0[in], ntx = (L.n[0] + TX - 1) / TX;
      r[6] = (t % ntx) * TX, r[7] = (t / ntx) * TY;
    } else { // rows of the block relative to the instance
      r[6] = A.row0[b] - s.inst_dof[in], r[7] = A.row0[b + 1] - s.inst_dof[in];
    }
  }
  op.fmt.d_mf_blk.upload(blk);
}

std::vector<int64_t> mf_tile_counts(const Operator& op)
{
  const Grid& g = *op.space->grid;
  if (int(g.lattice.size()) != g.ninst() || g.ref.dim != 2)
    return {};
  std::vector<int64_t> t(g.ninst());
  for (int in = 0; in < g.ninst(); ++in)
    t[in] = ((g.lattice[in].n[0] + TX - 1) / TX) * ((g.lattice[in].n[1] + TY - 1) / TY);
  return t;
}

int64_t mf_bytes(const Operator& op, bool scaled)
{
  return (scaled ? 24 : 16) * op.n + int64_t(8) * op.fmt.mf_fields * op.space->grid->ne;
}

void mf_launch(const Operator& op, int mode, const MfVectors& v, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1)
{
  const SolverFormat& F = op.fmt;
  MfParams P{};
  P.ninst = int(op.space->inst_dof.size()) - 1;
  P.blk = F.d_mf_blk.ptr;
  P.ref_stride = F.mf_ref_stride;
  const int D = op.space->grid->ref.dim, NV = 1 << D;
  P.eoff = 4;
  P.coff = P.eoff + int64_t(op.eforms.size()) * NV * NV;
  P.boff = P.coff + int64_t(op.cforms.size()) * 2 * 96;
  P.dirichlet = op.dirichlet ? 1 : 0;
  P.nef = int(op.eforms.size());
  for (int f = 0; f < P.nef; ++f) {
    const auto& E = op.eforms[f];
    P.ecoef[f] = E.field, P.evalue[f] = E.value, P.escale[f] = E.scaling;
  }
  P.ncf = int(op.cforms.size()), P.nbf = int(op.bforms.size());
  for (int grp = 0; grp < 2; ++grp) {
    const auto& forms = grp ? op.bforms : op.cforms;
    for (size_t f = 0; f < forms.size(); ++f) {
      const auto& Fm = forms[f];
      P.fkind[grp][f] = Fm.kind, P.fparam[grp][f] = Fm.param, P.fscale[grp][f] = Fm.scaling;
      P.fkappa[grp][f] = Fm.kappa, P.fkv[grp][f] = Fm.kappa_value;
      P.fomega[grp][f] = Fm.omega, P.fov[grp][f] = Fm.omega_value;
    }
  }
  P.nb = int(F.A.nblocks);
  P.v = v;
  const unsigned grid = hw_grid(F.A.nblocks);
  const int block = F.A.block; // one row / element per lane (the partition is built so)
  const bool ok = (F.mf_kind == MFK_DG_QUAD)
                      ? (block == TX * TY && F.A.items == 4 && int64_t(F.A.nblocks) * TX * TY * 4 >= op.n)
                      : (block <= 256 && F.A.items == 1 && int64_t(block) * F.A.nblocks >= op.n);
  if (!ok)
    HGDT_THROW(HGDT_ERR_OPERATOR, "matrix-free application: row partition does not match the kernels");
  const void* k = nullptr;
  auto pick = [&](auto k0, auto k1, auto k2) {
    k = mode == MF_APPLY ? reinterpret_cast<const void*>(k0)
                         : mode == MF_PCG ? reinterpret_cast<const void*>(k1) : reinterpret_cast<const void*>(k2);
  };
  // compile-time form counts for the usual operators: one element form (Laplace), and for SWIPDG one
  // element form, coupling + penalty on the inner and on the boundary intersections
  // 16-byte vector accesses when every vector of the launch is 16-byte aligned
  auto aligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  const bool vec = aligned(v.x) && aligned(v.y) && aligned(v.dh);
  const bool one = op.eforms.size() == 1;
  const bool swipdg = one && op.cforms.size() == 2 && op.bforms.size() == 2;
  if (F.mf_kind == MFK_CG_QUAD && one)
    pick(k_mf_cg<2, MF_APPLY, 1>, k_mf_cg<2, MF_PCG, 1>, k_mf_cg<2, MF_DIAG, 1>);
  else if (F.mf_kind == MFK_CG_QUAD)
    pick(k_mf_cg<2, MF_APPLY, -1>, k_mf_cg<2, MF_PCG, -1>, k_mf_cg<2, MF_DIAG, -1>);
  else if (F.mf_kind == MFK_CG_HEX && one)
    pick(k_mf_cg<3, MF_APPLY, 1>, k_mf_cg<3, MF_PCG, 1>, k_mf_cg<3, MF_DIAG, 1>);
  else if (F.mf_kind == MFK_CG_HEX)
    pick(k_mf_cg<3, MF_APPLY, -1>, k_mf_cg<3, MF_PCG, -1>, k_mf_cg<3, MF_DIAG, -1>);
  else if (F.mf_kind == MFK_DG_QUAD && swipdg && vec)
    pick(k_mf_dg<MF_APPLY, 1, 2, 2, true>, k_mf_dg<MF_PCG, 1, 2, 2, true>, k_mf_dg<MF_DIAG, 1, 2, 2, true>);
  else if (F.mf_kind == MFK_DG_QUAD && swipdg)
    pick(k_mf_dg<MF_APPLY, 1, 2, 2, false>, k_mf_dg<MF_PCG, 1, 2, 2, false>, k_mf_dg<MF_DIAG, 1, 2, 2, false>);
  else if (F.mf_kind == MFK_DG_QUAD && vec)
    pick(k_mf_dg<MF_APPLY, -1, -1, -1, true>, k_mf_dg<MF_PCG, -1, -1, -1, true>,
         k_mf_dg<MF_DIAG, -1, -1, -1, true>);
  else if (F.mf_kind == MFK_DG_QUAD)
    pick(k_mf_dg<MF_APPLY, -1, -1, -1, false>, k_mf_dg<MF_PCG, -1, -1, -1, false>,
         k_mf_dg<MF_DIAG, -1, -1, -1, false>);
  else
    HGDT_THROW(HGDT_ERR_OPERATOR, "matrix-free data not prepared");
  const double* ref = F.d_mf_ref.ptr;
  void* args[] = {&P, &ref};
  if (ev0)
    HGDT_HIP(hipExtLaunchKernel(k, dim3(grid), dim3(block), args, 0, stream, ev0, ev1, 0));
  else
    HGDT_HIP(hipLaunchKernel(k, dim3(grid), dim3(block), args, 0, stream));
}

} // namespace hgdt
