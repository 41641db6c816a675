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
constexpr int TX = 32, TY = 8; // tile of the discontinuous kernel (elements): halo rows 1.3x the tile
enum { MFK_CG_QUAD = 0, MFK_CG_HEX = 1, MFK_DG_QUAD = 2 };

// the parameters of an appended form, in one record (a kernel's few forms then span few lines of its argument
// block, which the kernel reads before anything else)
struct MfElementForm {
  const double* coef;
  double value, scale;
  int kind;
};
struct MfFaceForm {
  const double* kappa;
  const double* omega;
  double kv, ov, param, scale;
  int kind;
};

struct MfParams {
  int ninst;
  // per row block: first DoF and first element of the instance, cells n0 n1 n2, instance, first and last
  // row of the block relative to the instance's first DoF
  const int64_t* blk;
  int64_t ref_stride; // doubles per instance of the reference data: h0 h1 1/h1 1/h0, element, face, boundary matrices
  int64_t eoff; // offset of the element matrices inside an instance's block (continuous spaces)
  int dirichlet;
  int nef, ncf, nbf;
  int nb; // row blocks (the solver's operator partition): one row or element per lane
  const double* single; // the one coefficient field of every form (kernels that prefetch it), or NULL
  MfVectors v;
  // the forms last: a kernel reads the first nef / ncf / nbf records of each array
  MfElementForm ef[MF_MAXF];
  MfFaceForm ff[2][MF_MAXF]; // [0] inner, [1] boundary forms
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
  if constexpr (MODE == MF_PCG || MODE == MF_PCG_DIR) {
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

// the 4 entries i .. i + 3 of a (16-byte accesses when VEC)
template <bool VEC>
__device__ __forceinline__ void ld4(const double* __restrict__ a, int64_t i, double* o)
{
  if constexpr (VEC) {
    const double2 a0 = reinterpret_cast<const double2*>(a + i)[0], a1 = reinterpret_cast<const double2*>(a + i)[1];
    o[0] = a0.x, o[1] = a0.y, o[2] = a1.x, o[3] = a1.y;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      o[k] = a[i + k];
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
  const int b = active_block(P.nb, P.v.list, P.v.rev);
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
        kc[f][c] = valid[c] ? cf(P.ef[f].coef, P.ef[f].value, cell[c]) : 0.;
  }
  if (MODE == MF_PCG && P.v.S[(P.v.par ^ 1) * P.ninst + inst].done)
    return;
  const double* __restrict__ R = ref + int64_t(inst) * P.ref_stride + P.eoff;
  double y = 0.;
  if (fixed) {
    y = DIAG ? 1. : s[NS / 2];
  } else {
    for (int f = 0; f < (NEF >= 0 ? NEF : P.nef); ++f) {
      const double sc = P.ef[f].scale;
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        if (!valid[c])
          continue;
        const double k = (NEF >= 0) ? kc[NEF >= 0 ? f : 0][c] : cf(P.ef[f].coef, P.ef[f].value, cell[c]);
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

// 1 / x for x > 0: hardware reciprocal and two Newton steps (correct to round-off)
__device__ __forceinline__ double recip(double x)
{
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.), r, r);
  return fma(fma(-x, r, 1.), r, r);
}

// ---- discontinuous Q1 on rectangles, in closed form ----
// Every integral of the element and intersection forms is exact under the reference's rules (order
// 0 + 1 + 1 + over_integrate >= 2: Gauss rules of >= 2 points per direction; the integrands are of degree
// <= 2 per direction on a rectangle), so the local operators are applied in closed form.  Local DoF i has
// bits (i & 1, i >> 1) = reference (x, y) vertex coordinates.  On a face normal to direction o the trace of
// a Q1 function is linear along the face in the basis (1 - t, t); its two coefficients are the DoFs
// fidx(o, side, m), m = 0, 1 (side 1: reference coordinate 1 in direction o).  With W = [2 1; 1 2]:
//   face mass        int theta_m theta_n = len / 6 W[m][n]
//   element Laplace  (h_y / 6 h_x) s_i s_j W[y_i][y_j] + (h_x / 6 h_y) s'_i s'_j W[x_i][x_j]
//                    (s = +-1 by the x bit, s' by the y bit)
//   element product  h_x h_y / 36 W[x_i][x_j] W[y_i][y_j]
// and for an inner intersection with inside L (its side 1) and outside R (its side 0), jump J = u_L - u_R,
// F = wm k_L dn u_L + wp k_R dn u_R (laplace-ipdg.hh:107-186, ipdg.hh:113-171):
//   rows of L:  - int F psi_i - s wm k_L int J dn psi_i + p int J psi_i
//   rows of R:  + int F psi_i - s wp k_R int J dn psi_i - p int J psi_i
// and on a boundary intersection with outward normal (laplace-ipdg.hh:340-368, ipdg.hh:254-282):
//   - int k dn u psi_i - s int k u dn psi_i + sigma omega / h_F int u psi_i
constexpr int fidx(int o, int side, int m)
{
  return o == 0 ? side + 2 * m : 2 * side + m;
}

struct DgGeom {
  double h[2], ih[2]; // cell widths and their inverses
};

// coefficients of the element and of its neighbours across sides 0 .. 3 (x-, x+, y-, y+): a field's value or
// the constant.  PRE: every form's coefficient is one field (MfParams::single) or a constant, and that
// field's values at the element and at its neighbours are loaded once, at the start of the kernel
template <bool PRE>
struct Coefs {
  int64_t e;     // the element
  int64_t en[4]; // its neighbours
  double own_v = 0., nb_v[4] = {0., 0., 0., 0.};
  __device__ __forceinline__ double own(const double* f, double value) const
  {
    if constexpr (PRE)
      return f ? own_v : value;
    else
      return cf(f, value, e);
  }
  __device__ __forceinline__ double nb(const double* f, double value, int side) const
  {
    if constexpr (PRE)
      return f ? nb_v[side] : value;
    else
      return cf(f, value, en[side]);
  }
};

// y += c (W v) on the face DoFs of `side` (v: 2 face coefficients already multiplied by len / 6)
template <int O, int SIDE>
__device__ __forceinline__ void add_face(double c, const double (&wv)[2], double* y)
{
  y[fidx(O, SIDE, 0)] = fma(c, wv[0], y[fidx(O, SIDE, 0)]);
  y[fidx(O, SIDE, 1)] = fma(c, wv[1], y[fidx(O, SIDE, 1)]);
}
// y_i += c dn(psi_i) v[m_i] for all i: dn psi_i = +-1 / h_o by side (the derivative along +e_o)
template <int O>
__device__ __forceinline__ void add_dn(double c, const double (&wv)[2], double* y)
{
  y[fidx(O, 1, 0)] = fma(c, wv[0], y[fidx(O, 1, 0)]);
  y[fidx(O, 1, 1)] = fma(c, wv[1], y[fidx(O, 1, 1)]);
  y[fidx(O, 0, 0)] = fma(-c, wv[0], y[fidx(O, 0, 0)]);
  y[fidx(O, 0, 1)] = fma(-c, wv[1], y[fidx(O, 0, 1)]);
}
__device__ __forceinline__ void wmul(double c6, const double (&v)[2], double (&out)[2])
{
  out[0] = c6 * (2. * v[0] + v[1]);
  out[1] = c6 * (v[0] + 2. * v[1]);
}

// element forms
template <class KC>
__device__ __forceinline__ void dg_element(const MfParams& P, int nef, const KC& K, const DgGeom& g, const double* a,
                                           double* y)
{
  for (int f = 0; f < nef; ++f) {
    const double c = P.ef[f].scale * K.own(P.ef[f].coef, P.ef[f].value);
    if (P.ef[f].kind == EF_LAPLACE) {
      const double cx = c * (g.h[1] * g.ih[0]) * (1. / 6.), cy = c * (g.h[0] * g.ih[1]) * (1. / 6.);
      const double dx[2] = {a[1] - a[0], a[3] - a[2]}, dy[2] = {a[2] - a[0], a[3] - a[1]};
      const double wx0 = 2. * dx[0] + dx[1], wx1 = dx[0] + 2. * dx[1];
      const double wy0 = 2. * dy[0] + dy[1], wy1 = dy[0] + 2. * dy[1];
      y[0] = fma(-cx, wx0, fma(-cy, wy0, y[0]));
      y[1] = fma(cx, wx0, fma(-cy, wy1, y[1]));
      y[2] = fma(-cx, wx1, fma(cy, wy0, y[2]));
      y[3] = fma(cx, wx1, fma(cy, wy1, y[3]));
    } else {
      const double cm = c * (g.h[0] * g.h[1]) * (1. / 36.);
      double t[4]; // W in x, then W in y
      t[0] = 2. * a[0] + a[1], t[1] = a[0] + 2. * a[1], t[2] = 2. * a[2] + a[3], t[3] = a[2] + 2. * a[3];
      y[0] = fma(cm, 2. * t[0] + t[2], y[0]);
      y[1] = fma(cm, 2. * t[1] + t[3], y[1]);
      y[2] = fma(cm, t[0] + 2. * t[2], y[2]);
      y[3] = fma(cm, t[1] + 2. * t[3], y[3]);
    }
  }
}

// the inner forms of side SD (0 x-, 1 x+, 2 y-, 3 y+; across x+ / y+ this element is the inside)
template <int SD, int NCF, class KC>
__device__ __forceinline__ void dg_inner(const MfParams& P, int ncf, const KC& K, const DgGeom& g, const double* a,
                                         const double* bv, double* y)
{
  constexpr int O = SD >> 1;
  constexpr bool INSIDE = SD & 1;
  const double* uL = INSIDE ? a : bv;
  const double* uR = INSIDE ? bv : a;
  const double ihn = g.ih[O], c6 = g.h[1 - O] * (1. / 6.), ilen = g.ih[1 - O];
  double J[2], DL[2], DR[2], YJ[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    J[m] = uL[fidx(O, 1, m)] - uR[fidx(O, 0, m)];
    DL[m] = (uL[fidx(O, 1, m)] - uL[fidx(O, 0, m)]) * ihn;
    DR[m] = (uR[fidx(O, 1, m)] - uR[fidx(O, 0, m)]) * ihn;
  }
  wmul(c6, J, YJ);
  for (int f = 0; f < (NCF >= 0 ? NCF : ncf); ++f) {
    const double sc = P.ff[0][f].scale, prm = P.ff[0][f].param;
    const double* __restrict__ of = P.ff[0][f].omega;
    const double os = K.own(of, P.ff[0][f].ov), on = K.nb(of, P.ff[0][f].ov, SD);
    // weights of the intersection from its inside (-) and outside (+) elements
    const double dm = INSIDE ? os : on, dp = INSIDE ? on : os;
    const double inv = recip(dp + dm);
    if (P.ff[0][f].kind == FF_INNER_COUPLING) {
      const double* __restrict__ kf = P.ff[0][f].kappa;
      const double ks = K.own(kf, P.ff[0][f].kv), kn = K.nb(kf, P.ff[0][f].kv, SD);
      const double kin = INSIDE ? ks : kn, kout = INSIDE ? kn : ks;
      const double wm = dp * inv, wp = dm * inv;
      double F[2], YF[2];
      F[0] = (wm * kin) * DL[0] + (wp * kout) * DR[0];
      F[1] = (wm * kin) * DL[1] + (wp * kout) * DR[1];
      wmul(c6, F, YF);
      if constexpr (INSIDE) {
        add_face<O, 1>(-sc, YF, y);
        add_dn<O>(-sc * prm * wm * kin * ihn, YJ, y);
      } else {
        add_face<O, 0>(sc, YF, y);
        add_dn<O>(-sc * prm * wp * kout * ihn, YJ, y);
      }
    } else {
      const double pen = sc * (prm * (dp * dm * inv)) * ilen;
      if constexpr (INSIDE)
        add_face<O, 1>(pen, YJ, y);
      else
        add_face<O, 0>(-pen, YJ, y);
    }
  }
}

// the boundary forms on side SD (outward normal +-e_o)
template <int SD, int NBF, class KC>
__device__ __forceinline__ void dg_boundary(const MfParams& P, int nbf, const KC& K, const DgGeom& g, const double* a,
                                            double* y)
{
  constexpr int O = SD >> 1, SIDE = SD & 1;
  const double sign = SIDE ? 1. : -1.;
  const double ihn = g.ih[O], c6 = g.h[1 - O] * (1. / 6.), ilen = g.ih[1 - O];
  double T[2], Dn[2], YT[2], YD[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    T[m] = a[fidx(O, SIDE, m)];
    Dn[m] = sign * ((a[fidx(O, 1, m)] - a[fidx(O, 0, m)]) * ihn);
  }
  wmul(c6, T, YT);
  wmul(c6, Dn, YD);
  for (int f = 0; f < (NBF >= 0 ? NBF : nbf); ++f) {
    const double sc = P.ff[1][f].scale, prm = P.ff[1][f].param;
    if (P.ff[1][f].kind == FF_DIRICHLET_COUPLING) {
      const double k = K.own(P.ff[1][f].kappa, P.ff[1][f].kv);
      add_face<O, SIDE>(-sc * k, YD, y);
      add_dn<O>(-sc * prm * k * sign * ihn, YT, y);
    } else {
      add_face<O, SIDE>(sc * (prm * K.own(P.ff[1][f].omega, P.ff[1][f].ov)) * ilen, YT, y);
    }
  }
}

// y = (local operator of element e) applied to its own values a and its neighbours' bv
template <int NEF, int NCF, int NBF, class KC>
__device__ __forceinline__ void dg_local(const MfParams& P, int nef, int ncf, int nbf, const KC& K,
                                         const bool (&has)[4], const DgGeom& g, const double* a,
                                         const double (&bv)[4][4], double* y)
{
  dg_element(P, NEF >= 0 ? NEF : nef, K, g, a, y);
  if (has[0])
    dg_inner<0, NCF>(P, ncf, K, g, a, bv[0], y);
  else
    dg_boundary<0, NBF>(P, nbf, K, g, a, y);
  if (has[1])
    dg_inner<1, NCF>(P, ncf, K, g, a, bv[1], y);
  else
    dg_boundary<1, NBF>(P, nbf, K, g, a, y);
  if (has[2])
    dg_inner<2, NCF>(P, ncf, K, g, a, bv[2], y);
  else
    dg_boundary<2, NBF>(P, nbf, K, g, a, y);
  if (has[3])
    dg_inner<3, NCF>(P, ncf, K, g, a, bv[3], y);
  else
    dg_boundary<3, NBF>(P, nbf, K, g, a, y);
}

// discontinuous Q1 on quadrilaterals: a block works on a tile of TX x TY elements, one lane per element (its
// 4 consecutive DoFs).  The source values of the tile and of a one-element halo (D^{-1/2} p for the solver)
// are staged in LDS with coalesced loads (the 4-DoF vectors of a row of elements are contiguous), so each
// lane reads its own and its 4 neighbours' values from LDS.  The diagonal is the local operator applied to
// the 4 unit vectors.  NEF, NCF, NBF >= 0: form counts known at compile time (loops unrolled); -1: runtime
// loops.  VEC: the vectors are 16-byte aligned (16-byte accesses).  PRE: the coefficients come from one field
// (Coefs).
// Latency: every load that depends only on the block's record is issued before any result is waited for
// (the staged vectors, the coefficients, the solver's scalars and partials), so the kernel pays about one
// memory round trip before its arithmetic instead of one per stage.
template <int MODE, int NEF, int NCF, int NBF, bool VEC, bool PRE>
__global__ void __launch_bounds__(TX * TY) k_mf_dg(MfParams P, const double* __restrict__ ref)
{
  constexpr bool DIAG = (MODE == MF_DIAG);
  constexpr bool DIR = (MODE == MF_PCG_DIR);
  constexpr bool SCALED = (MODE == MF_PCG || DIR);
  __shared__ __align__(16) double sp[TY + 2][TX + 2][4]; // source of the tile and its halo
  // DIR: D^{-1/2} of the tile's own elements (staged with the source; the tile's own rows are touched
  // only by the staging, which also writes their new p and updates their iterate)
  __shared__ __align__(16) double sh[DIR ? TY : 1][DIR ? TX : 1][4];
  const int nef = NEF >= 0 ? NEF : P.nef, ncf = NCF >= 0 ? NCF : P.ncf, nbf = NBF >= 0 ? NBF : P.nbf;
  const int b = active_block(P.nb, P.v.list, P.v.rev);
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
  const uint32_t c = ey * nx + ex;
  const int64_t i0 = d0 + 4 * int64_t(c); // first DoF of the lane's element
  const int64_t e = e0 + c;
  const bool has[4] = {row && ex > 0u, row && ex + 1u < nx, row && ey > 0u, row && ey + 1u < ny};
  Coefs<PRE> K;
  K.e = e;
  K.en[0] = e - 1, K.en[1] = e + 1, K.en[2] = e - int64_t(nx), K.en[3] = e + int64_t(nx);
  if constexpr (PRE) {
    if (P.single) {
      K.own_v = row ? P.single[e] : 0.;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        K.nb_v[s] = has[s] ? P.single[K.en[s]] : 0.;
    }
  }
  // staged entries: columns x0 + xs .. x0 + xe - 1 of the rows y0 - 1 .. y0 + TY (clipped to the instance),
  // U accesses of W doubles per element and array; entry u of the block is (row u / per_row, access
  // u % per_row), lane t takes the entries t + k TX TY
  constexpr int U = VEC ? 2 : 4, W = 4 / U;
  constexpr int NR = ((TY + 2) * (TX + 2) * U + TX * TY - 1) / (TX * TY);
  const int xs = x0 > 0u ? -1 : 0;
  const int xe = int(min(uint32_t(TX + 1), nx - x0));
  const int per_row = (xe - xs) * U;
  auto entry = [&](int k, int& r, int& col, int& k0, int64_t& dof) {
    const int u = int(threadIdx.x) + k * TX * TY;
    r = u / per_row;
    const int w = u % per_row;
    col = xs + w / U; // relative to x0
    k0 = W * (w % U);
    const int64_t yy = int64_t(y0) - 1 + r;
    dof = d0 + 4 * (yy * nx + x0 + col) + k0;
    return u < (TY + 2) * per_row && yy >= 0 && yy < int64_t(ny);
  };
  auto ld = [](const double* __restrict__ a, int64_t i, double (&o)[W]) {
    if constexpr (VEC) {
      const double2 v = *reinterpret_cast<const double2*>(a + i);
      o[0] = v.x, o[1] = v.y;
    } else {
      o[0] = a[i];
    }
  };
  auto st = [](double* __restrict__ a, int64_t i, const double (&v)[W]) {
    if constexpr (VEC)
      *reinterpret_cast<double2*>(a + i) = make_double2(v[0], v[1]);
    else
      a[i] = v[0];
  };
  // the staged vectors' loads: the source (DIR: r, and the previous p, read even at the first step, where it
  // is not used) and D^{-1/2}
  double lx[NR][W], lp[NR][W], lh[NR][W];
  if constexpr (!DIAG) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      int r, col, k0;
      int64_t dof;
#pragma unroll
      for (int j = 0; j < W; ++j)
        lx[k][j] = lp[k][j] = lh[k][j] = 0.;
      if (entry(k, r, col, k0, dof)) {
        ld(DIR ? P.v.r : P.v.x, dof, lx[k]);
        if constexpr (DIR)
          ld(P.v.x, dof, lp[k]);
        if constexpr (SCALED)
          ld(P.v.dh, dof, lh[k]);
      }
    }
  }
  // DIR: the step's scalars and the instance's partial range, read after the staged vectors' loads are
  // issued (which need only the block's record) and before any store of the kernel
  PcgScalars S{};
  int pb0 = 0, pb1 = 0;
  if constexpr (DIR) {
    S = P.v.S[P.v.par * P.ninst + inst];
    pb0 = P.v.v_ib0[inst], pb1 = P.v.v_ib0[inst + 1];
  }
  // DIR: the direction update of the step (as k_cg_direction: stopping test on the previous update's
  // partials, p = r + beta p, and every second step the iterate's two deferred updates) on the block's
  // elements.  The staged source is formed from r and the previous p, which live in another buffer than
  // the new p, so a neighbouring block's writes never reach this block's halo.
  double alpha = 0., alpha_prev = 0., beta = 0.;
  bool first = false, pair = false; // pair: the iterate's two pending updates are due (even step >= 2)
  double lo[DIR ? NR : 1][W], lxh[DIR ? NR : 1][W]; // pair steps: p_{j-2} and the iterate of the own entries
  if constexpr (DIR) {
    PcgScalars* nxt = P.v.S + (P.v.par ^ 1) * P.ninst + inst;
    const bool writer = (b == P.v.a_ib0[inst] && threadIdx.x == 0);
    const int64_t it = S.iters;
    first = (it == 0);
    pair = !first && (it & 1) == 0;
    const int done = pcg_state(S);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
#pragma unroll
      for (int j = 0; j < W; ++j)
        lo[k][j] = lxh[k][j] = 0.;
      int r, col, k0;
      int64_t dof;
      if (!done && pair && entry(k, r, col, k0, dof) && r >= 1 && r <= TY && col >= 0 && col < TX) {
        ld(P.v.p_out, dof, lo[k]);
        ld(P.v.xh, dof, lxh[k]);
      }
    }
    Partials<2> pt;
    pt.issue(P.v.part2, pb0, done ? pb0 : pb1);
    if (done) {
      if (writer) {
        *nxt = S;
        nxt->done = done;
      }
      return;
    }
    alpha = S.alpha, alpha_prev = S.alpha_prev;
    double t[2];
    pt.sum(P.v.part2, t);
    const int status = (t[1] <= S.tol2) ? 1 : (it >= P.v.max_iter) ? 2 : 0;
    beta = (status == 0 && !first) ? t[0] / S.rz : 0.;
    if (writer) {
      nxt->bb = S.bb;
      nxt->tol2 = S.tol2;
      nxt->alpha = alpha;
      nxt->alpha_prev = alpha;
      nxt->pq = S.pq;
      nxt->rr = t[1];
      nxt->rz = t[0];
      nxt->beta = beta;
      nxt->iters = status ? it : it + 1;
      nxt->done = status;
      nxt->brk = 0;
    }
    if (status != 0) {
      // the instance stops here: only the iterate's pending updates remain (uniform over the block)
      if (row && !first) {
        double pv[4], xv[4];
        ld4<VEC>(P.v.x, i0, pv);
        ld4<VEC>(P.v.xh, i0, xv);
        if (pair) {
          double ov[4];
          ld4<VEC>(P.v.p_out, i0, ov);
#pragma unroll
          for (int k = 0; k < 4; ++k)
            xv[k] = fma(alpha_prev, ov[k], xv[k]);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
          xv[k] = fma(alpha, pv[k], xv[k]);
        store4<VEC>(P.v.xh, i0, xv);
      }
      return;
    }
  }
  if constexpr (!DIAG) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      int r, col, k0;
      int64_t dof;
      if (!entry(k, r, col, k0, dof))
        continue;
      double xv[W];
#pragma unroll
      for (int j = 0; j < W; ++j)
        xv[j] = (DIR && !first) ? fma(beta, lp[k][j], lx[k][j]) : lx[k][j];
      if constexpr (DIR) {
        // an entry of the tile itself (not of the halo): its new p, and the iterate on pair steps
        if (r >= 1 && r <= TY && col >= 0 && col < TX) {
          if (pair) {
            double xh[W];
#pragma unroll
            for (int j = 0; j < W; ++j)
              xh[j] = fma(alpha, lp[k][j], fma(alpha_prev, lo[k][j], lxh[k][j]));
            st(P.v.xh, dof, xh);
          }
          st(P.v.p_out, dof, xv);
#pragma unroll
          for (int j = 0; j < W; ++j)
            sh[r - 1][col][k0 + j] = lh[k][j];
        }
      }
#pragma unroll
      for (int j = 0; j < W; ++j)
        sp[r][col + 1][k0 + j] = SCALED ? xv[j] * lh[k][j] : xv[j];
    }
    __syncthreads();
  }
  const double* __restrict__ R = ref + int64_t(inst) * P.ref_stride;
  DgGeom g;
  g.h[0] = R[0], g.h[1] = R[1], g.ih[0] = R[3], g.ih[1] = R[2];
  double y[4] = {0., 0., 0., 0.};
  if constexpr (DIAG) {
    // diagonal: column k of the element's own block, entry k
    const double z[4][4] = {};
    if (row)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double a[4] = {0., 0., 0., 0.}, yk[4] = {0., 0., 0., 0.};
        a[k] = 1.;
        dg_local<NEF, NCF, NBF>(P, nef, ncf, nbf, K, has, g, a, z, yk);
        y[k] = yk[k];
      }
  } else {
    double a[4], bv[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = row ? sp[ty + 1][tx + 1][k] : 0.;
      bv[0][k] = has[0] ? sp[ty + 1][tx][k] : 0.;
      bv[1][k] = has[1] ? sp[ty + 1][tx + 2][k] : 0.;
      bv[2][k] = has[2] ? sp[ty][tx + 1][k] : 0.;
      bv[3][k] = has[3] ? sp[ty + 2][tx + 1][k] : 0.;
    }
    if (row)
      dg_local<NEF, NCF, NBF>(P, nef, ncf, nbf, K, has, g, a, bv, y);
  }
  double part = 0.;
  if (row) {
    if constexpr (MODE == MF_APPLY) {
      store4<VEC>(P.v.y, i0, y);
    } else if constexpr (SCALED) {
      // q = D^{-1/2} y; p.q
      double q[4];
      if constexpr (DIR) {
        // D^{-1/2} and D^{-1/2} p of the element from LDS: p.q = (D^{-1/2} p).y
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          q[k] = sh[ty][tx][k] * y[k];
          part += sp[ty + 1][tx + 1][k] * y[k];
        }
      } else {
        double xo[4], ho[4];
        load4<MODE, VEC>(P, i0, xo, ho);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          q[k] = ho[k] * y[k];
          part += xo[k] * q[k];
        }
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
  if (op.borrowed || !op.square())
    unsupported("an operator on a caller's matrix or between two spaces keeps its matrix");
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
  if (op.dirichlet && (op.dirichlet_only_clear || !op.dirichlet_ensure_symmetry))
    unsupported("Dirichlet constraints as unit rows and columns (the default of DirichletConstraints::apply)");
  if (op.eforms.size() > size_t(MF_MAXF) || op.cforms.size() > size_t(MF_MAXF) || op.bforms.size() > size_t(MF_MAXF))
    unsupported("too many forms appended");
  for (const auto* forms : {&op.cforms, &op.bforms})
    for (const auto& f : *forms)
      if (f.kappa_t || f.omega_t)
        unsupported("intersection coefficients are scalar (not matrix-valued)");
  for (const auto& f : op.eforms)
    if (f.points || f.tensor_kind)
      unsupported("element coefficients are scalar, per element or constant (not matrix-valued or given at "
                  "quadrature points)");
  for (const auto& b : op.bforms)
    if (b.mask)
      unsupported("boundary forms are applied on every boundary intersection (no boundary-info filter)");
  if (!op.dirichlet_faces.empty())
    unsupported("Dirichlet constraints are applied on every boundary intersection (no boundary info)");
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

  // per instance: header h0, h1, 1 / h1, 1 / h0; continuous spaces also the reference element matrices,
  // with the assembly kernels' quadrature rules (order 0 + 1 + 1 + over_integrate: integrals.hh:115,
  // laplace.hh:74-79, product.hh:88-100).  Discontinuous spaces apply their forms in closed form.
  const int nef = int(op.eforms.size());
  const int64_t eoff = 4;
  const int64_t stride = dg ? 4 : eoff + int64_t(nef) * NV * NV;
  std::vector<double> ref(size_t(stride) * ninst, 0.);
  for (int in = 0; in < ninst; ++in) {
    const Lattice& L = g.lattice[in];
    double* r = &ref[size_t(in) * stride];
    const double h[3] = {L.h[0], L.h[1], D == 3 ? L.h[2] : 1.};
    r[0] = h[0], r[1] = h[1], r[2] = 1. / h[1], r[3] = 1. / h[0];
    if (dg)
      continue;
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

int64_t mf_bytes(const Operator& op, int mode)
{
  // apply: x read, y written; solver: p, D^{-1/2} read, q written; with the direction update: r, p of the
  // previous step, D^{-1/2} read, p, q written (40 B), and every second step also the iterate and the p
  // before read and the iterate written (64 B): 52 B per launch on average.  Plus one coefficient per
  // element and distinct field.
  const int64_t per_dof = mode == MF_APPLY ? 16 : mode == MF_PCG ? 24 : 52;
  return per_dof * op.n + int64_t(8) * op.fmt.mf_fields * op.space->grid->ne;
}

bool mf_fused_direction(const Operator& op)
{
  return op.fmt.format == HGDT_FORMAT_MATRIX_FREE && op.fmt.mf_kind == MFK_DG_QUAD;
}

void mf_launch(const Operator& op, int mode, const MfVectors& v, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1)
{
  const SolverFormat& F = op.fmt;
  MfParams P{};
  P.ninst = int(op.space->inst_dof.size()) - 1;
  P.blk = F.d_mf_blk.ptr;
  P.ref_stride = F.mf_ref_stride;
  P.eoff = 4;
  P.dirichlet = op.dirichlet ? 1 : 0;
  P.nef = int(op.eforms.size());
  for (int f = 0; f < P.nef; ++f) {
    const auto& E = op.eforms[f];
    P.ef[f].kind = E.kind, P.ef[f].coef = E.field, P.ef[f].value = E.value, P.ef[f].scale = E.scaling;
  }
  P.ncf = int(op.cforms.size()), P.nbf = int(op.bforms.size());
  for (int grp = 0; grp < 2; ++grp) {
    const auto& forms = grp ? op.bforms : op.cforms;
    for (size_t f = 0; f < forms.size(); ++f) {
      const auto& Fm = forms[f];
      P.ff[grp][f].kind = Fm.kind, P.ff[grp][f].param = Fm.param, P.ff[grp][f].scale = Fm.scaling;
      P.ff[grp][f].kappa = Fm.kappa, P.ff[grp][f].kv = Fm.kappa_value;
      P.ff[grp][f].omega = Fm.omega, P.ff[grp][f].ov = Fm.omega_value;
    }
  }
  P.nb = v.nb >= 0 ? v.nb : int(F.A.nblocks);
  P.v = v;
  // the coefficient field shared by every form, if there is one (every other coefficient a constant)
  bool one_field = true;
  auto see = [&](const double* f) {
    if (f && P.single && f != P.single)
      one_field = false;
    else if (f)
      P.single = f;
  };
  for (const auto& E : op.eforms)
    see(E.field);
  for (const auto* forms : {&op.cforms, &op.bforms})
    for (const auto& Fm : *forms)
      see(Fm.kappa), see(Fm.omega);
  if (!one_field)
    P.single = nullptr;
  const unsigned grid = hw_grid(P.nb);
  const int block = F.A.block; // one row / element per lane (the partition is built so)
  const bool ok = (F.mf_kind == MFK_DG_QUAD)
                      ? (block == TX * TY && F.A.items == 4 && int64_t(F.A.nblocks) * TX * TY * 4 >= op.n)
                      : (block <= 256 && F.A.items == 1 && int64_t(block) * F.A.nblocks >= op.n);
  if (!ok)
    HGDT_THROW(HGDT_ERR_OPERATOR, "matrix-free application: row partition does not match the kernels");
  const void* k = nullptr;
  auto pick = [&](auto k0, auto k1, auto k2, const void* k3) {
    k = mode == MF_APPLY ? reinterpret_cast<const void*>(k0)
        : mode == MF_PCG ? reinterpret_cast<const void*>(k1)
        : mode == MF_DIAG ? reinterpret_cast<const void*>(k2)
                          : k3;
  };
  auto kv = [](auto kernel) { return reinterpret_cast<const void*>(kernel); };
  // compile-time form counts for the usual operators: one element form (Laplace), and for SWIPDG one
  // element form, coupling + penalty on the inner and on the boundary intersections
  // 16-byte vector accesses when every vector of the launch is 16-byte aligned
  auto aligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  const bool vec = aligned(v.x) && aligned(v.y) && aligned(v.dh) && aligned(v.r) && aligned(v.xh) &&
                   aligned(v.p_out);
  const bool one = op.eforms.size() == 1;
  const bool swipdg = one && op.cforms.size() == 2 && op.bforms.size() == 2;
  if (F.mf_kind == MFK_CG_QUAD && one)
    pick(k_mf_cg<2, MF_APPLY, 1>, k_mf_cg<2, MF_PCG, 1>, k_mf_cg<2, MF_DIAG, 1>, nullptr);
  else if (F.mf_kind == MFK_CG_QUAD)
    pick(k_mf_cg<2, MF_APPLY, -1>, k_mf_cg<2, MF_PCG, -1>, k_mf_cg<2, MF_DIAG, -1>, nullptr);
  else if (F.mf_kind == MFK_CG_HEX && one)
    pick(k_mf_cg<3, MF_APPLY, 1>, k_mf_cg<3, MF_PCG, 1>, k_mf_cg<3, MF_DIAG, 1>, nullptr);
  else if (F.mf_kind == MFK_CG_HEX)
    pick(k_mf_cg<3, MF_APPLY, -1>, k_mf_cg<3, MF_PCG, -1>, k_mf_cg<3, MF_DIAG, -1>, nullptr);
  else if (F.mf_kind == MFK_DG_QUAD && swipdg && vec && one_field)
    pick(k_mf_dg<MF_APPLY, 1, 2, 2, true, true>, k_mf_dg<MF_PCG, 1, 2, 2, true, true>,
         k_mf_dg<MF_DIAG, 1, 2, 2, true, true>, kv(k_mf_dg<MF_PCG_DIR, 1, 2, 2, true, true>));
  else if (F.mf_kind == MFK_DG_QUAD && swipdg && vec)
    pick(k_mf_dg<MF_APPLY, 1, 2, 2, true, false>, k_mf_dg<MF_PCG, 1, 2, 2, true, false>,
         k_mf_dg<MF_DIAG, 1, 2, 2, true, false>, kv(k_mf_dg<MF_PCG_DIR, 1, 2, 2, true, false>));
  else if (F.mf_kind == MFK_DG_QUAD && swipdg)
    pick(k_mf_dg<MF_APPLY, 1, 2, 2, false, false>, k_mf_dg<MF_PCG, 1, 2, 2, false, false>,
         k_mf_dg<MF_DIAG, 1, 2, 2, false, false>, kv(k_mf_dg<MF_PCG_DIR, 1, 2, 2, false, false>));
  else if (F.mf_kind == MFK_DG_QUAD && vec)
    pick(k_mf_dg<MF_APPLY, -1, -1, -1, true, false>, k_mf_dg<MF_PCG, -1, -1, -1, true, false>,
         k_mf_dg<MF_DIAG, -1, -1, -1, true, false>, kv(k_mf_dg<MF_PCG_DIR, -1, -1, -1, true, false>));
  else if (F.mf_kind == MFK_DG_QUAD)
    pick(k_mf_dg<MF_APPLY, -1, -1, -1, false, false>, k_mf_dg<MF_PCG, -1, -1, -1, false, false>,
         k_mf_dg<MF_DIAG, -1, -1, -1, false, false>, kv(k_mf_dg<MF_PCG_DIR, -1, -1, -1, false, false>));
  else
    HGDT_THROW(HGDT_ERR_OPERATOR, "matrix-free data not prepared");
  if (!k)
    HGDT_THROW(HGDT_ERR_OPERATOR, "matrix-free application: mode not provided for this operator");
  const double* ref = F.d_mf_ref.ptr;
  void* args[] = {&P, &ref};
  if (ev0 || ev1)
    HGDT_HIP(hipExtLaunchKernel(k, dim3(grid), dim3(block), args, 0, stream, ev0, ev1, 0));
  else
    HGDT_HIP(hipLaunchKernel(k, dim3(grid), dim3(block), args, 0, stream));
}

} // namespace hgdt
