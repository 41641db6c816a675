// assembly.hip -- device assembly for libhgdt.so (gfx950, FP64).
//
// The reference's grid walk (XT::Grid::Walker driven by MatrixOperator::assemble,
// dune/gdt/operators/matrix-based.hh:496-500) becomes three launches:
//   1. k_element_local: one element per lane.  Geometry (J, J^{-1}, |det J|) and
//      basis data live in registers; the quadrature rules are read from a small
//      uniform table.  Computes the sum of the appended element forms
//      (LocalElementIntegralBilinearForm::apply2, local/bilinear-forms/integrals.hh:97-134
//      with LocalLaplaceIntegrand, local/integrands/laplace.hh:81-102, and
//      LocalElementProductIntegrand, local/integrands/product.hh:104-130) and the
//      right-hand side functional (local/functionals/integrals.hh:72-97), and stores
//      them entry-major (coalesced across lanes) in the staging buffer.
//   2. k_inner_local / k_boundary_local: one intersection per lane, SWIPDG coupling
//      and penalty terms (local/integrands/laplace-ipdg.hh:107-186,340-368,
//      local/integrands/ipdg.hh:113-171,254-282) through the face quadrature loops of
//      integrals.hh:197-267,338-369.
//   3. k_gather_values / k_gather_rhs: one matrix slot (resp. DoF) per lane sums its
//      contributions in the fixed walk order given by the precomputed scatter map and
//      applies the Dirichlet flag (tools/dirichlet-constraints.hh:153-184).  No atomics:
//      results are bitwise reproducible.

#include <algorithm>
#include <mutex>

#include "hgdt_internal.hpp"

namespace hgdt {
namespace {

constexpr int MAX_EFORMS = 8;
constexpr int MAX_FFORMS = 8;
constexpr int MAX_SOURCES = 4;
constexpr int MAX_BSOURCES = 4;

struct QRule {
  int offset; // into the point / weight table
  int npts;
};

struct ElementParams {
  int64_t ne;
  const double* coords;
  const int32_t* conn;
  // one part of a mixed grid: the grid index of every element (coefficient fields and per-element sources
  // are indexed by it) and, per source given at quadrature points, where each element's points start
  // (NULL: element e, point q at e * npts + q)
  const int32_t* eid;
  const int64_t* src_qoff[MAX_SOURCES];
  // element forms whose coefficient is given at the points of their rule (NULL: field or value), and on mixed
  // grids where each element's points start
  const double* points[MAX_EFORMS];
  const int64_t* ef_qoff[MAX_EFORMS];
  // Laplace forms with a matrix-valued diffusion: 1 per element (tensor, d x d row-major), 2 constant (tvalue)
  int tensor_kind[MAX_EFORMS];
  const double* tensor[MAX_EFORMS];
  double tvalue[MAX_EFORMS][9];
  const double* qtab; // points (dim each) then weights, per rule
  int nforms;
  int kind[MAX_EFORMS];
  const double* field[MAX_EFORMS];
  double value[MAX_EFORMS];
  double scaling[MAX_EFORMS];
  QRule rule[MAX_EFORMS];
  int nsrc;
  int src_kind[MAX_SOURCES];
  const double* src_data[MAX_SOURCES];
  double src_value[MAX_SOURCES];
  QRule src_rule[MAX_SOURCES];
  double* staging;
  int64_t elem_base, rhs_base;
  int write_matrix;
};

struct FaceParams {
  int64_t nfaces;
  const int32_t* faces; // inner: 4 ints per face, boundary: 2 ints per face
  const double* coords;
  const int32_t* conn;
  const double* qtab; // line rule points then weights
  int nforms;
  int kind[MAX_FFORMS];
  double param[MAX_FFORMS];
  const double* kappa[MAX_FFORMS];
  double kappa_value[MAX_FFORMS];
  const double* omega[MAX_FFORMS];
  double omega_value[MAX_FFORMS];
  // matrix-valued diffusion / weight: d x d per element (NULL: the scalars)
  const double* kappa_t[MAX_FFORMS];
  const double* omega_t[MAX_FFORMS];
  double scaling[MAX_FFORMS];
  QRule rule[MAX_FFORMS];
  const uint8_t* mask[MAX_FFORMS]; // boundary forms: applied on the intersections with a nonzero byte (NULL: all)
  double* staging;
  int64_t base;
  QRule brule[MAX_BSOURCES]; // line rules of the boundary right-hand side terms
  // every appended form is symmetric (symmetry prefactor 1 or a penalty): the blocks are stored
  // exactly symmetric (in_in, out_out and boundary blocks from their upper triangle, out_in as the
  // transpose of in_out), so the assembled matrix is symmetric bit for bit
  int symmetric;
  // mixed grids: the part of every element (a launch takes the faces whose inside / outside elements are of
  // parts part_in / part_out); conn holds cstride entries per element; the staged blocks have rows of ls
  // entries (the largest local size).  One element type: etype NULL, cstride the type's vertices, ls its ND
  const int8_t* etype;
  int part_in, part_out;
  int ls;
  int64_t cstride;
};

// right-hand side terms on the boundary intersections
struct BoundaryRhsParams {
  int64_t nfaces;
  const int32_t* faces; // 2 ints per boundary face: element, local face
  const double* coords;
  const int32_t* conn;
  const double* qtab; // line rules
  int nterms;
  int kind[MAX_BSOURCES];
  double param[MAX_BSOURCES];
  const double* coeff[MAX_BSOURCES];
  double coeff_value[MAX_BSOURCES];
  const double* data[MAX_BSOURCES];
  double value[MAX_BSOURCES];
  QRule rule[MAX_BSOURCES];
  const uint8_t* mask[MAX_BSOURCES]; // applied on the intersections with a nonzero byte (NULL: all)
  const double* coeff_t[MAX_BSOURCES]; // matrix-valued coefficient, d x d per element (NULL: coeff)
  double* staging;
  int64_t base;
  const int8_t* etype; // as FaceParams
  int part_in;
  int ls;
  int64_t cstride;
};

template <int GEOM>
struct Traits;
template <>
struct Traits<HGDT_QUAD> {
  static constexpr int D = 2, NV = 4, ND = 4;
  static constexpr bool SIMPLEX = false;
  static constexpr double VOL = 1.;
};
template <>
struct Traits<HGDT_TRI> {
  static constexpr int D = 2, NV = 3, ND = 3;
  static constexpr bool SIMPLEX = true;
  static constexpr double VOL = 0.5;
};
template <>
struct Traits<HGDT_HEX> {
  static constexpr int D = 3, NV = 8, ND = 8;
  static constexpr bool SIMPLEX = false;
  static constexpr double VOL = 1.;
};
template <>
struct Traits<HGDT_TET> {
  static constexpr int D = 3, NV = 4, ND = 4;
  static constexpr bool SIMPLEX = true;
  static constexpr double VOL = 1. / 6.;
};

// affine element geometry in registers
template <int D>
struct Geo {
  double x0[D];
  double J[D][D];
  double Ji[D][D];
  double det; // |det J|
};

// (stride: entries of conn per element, the type's vertex count, or 4 in a mixed grid's connectivity)
template <int GEOM>
__device__ __forceinline__ void load_geometry(const double* __restrict__ coords, const int32_t* __restrict__ conn,
                                              int64_t e, Geo<Traits<GEOM>::D>& g, int64_t stride = Traits<GEOM>::NV)
{
  using T = Traits<GEOM>;
  constexpr int D = T::D;
  int32_t v[D + 1];
  v[0] = conn[e * stride];
#pragma unroll
  for (int k = 0; k < D; ++k)
    v[k + 1] = conn[e * stride + (T::SIMPLEX ? (k + 1) : (1 << k))];
#pragma unroll
  for (int a = 0; a < D; ++a)
    g.x0[a] = coords[int64_t(v[0]) * D + a];
#pragma unroll
  for (int k = 0; k < D; ++k)
#pragma unroll
    for (int a = 0; a < D; ++a)
      g.J[a][k] = coords[int64_t(v[k + 1]) * D + a] - g.x0[a];
  bool diag = true;
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (a != k && g.J[a][k] != 0.)
        diag = false;
  double det;
  if (diag) {
    det = 1.;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      det *= g.J[a][a];
#pragma unroll
      for (int k = 0; k < D; ++k)
        g.Ji[a][k] = (a == k) ? 1. / g.J[a][a] : 0.;
    }
  } else if constexpr (D == 2) {
    det = g.J[0][0] * g.J[1][1] - g.J[0][1] * g.J[1][0];
    const double r = 1. / det;
    g.Ji[0][0] = g.J[1][1] * r, g.Ji[0][1] = -g.J[0][1] * r;
    g.Ji[1][0] = -g.J[1][0] * r, g.Ji[1][1] = g.J[0][0] * r;
  } else {
    const double(&A)[D][D] = g.J;
    const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    const double r = 1. / det;
    g.Ji[0][0] = c00 * r, g.Ji[1][0] = c01 * r, g.Ji[2][0] = c02 * r;
    g.Ji[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * r;
    g.Ji[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * r;
    g.Ji[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * r;
    g.Ji[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * r;
    g.Ji[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * r;
    g.Ji[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * r;
  }
  g.det = fabs(det);
}

// P1 / Q1 shape functions on the reference element: values and reference gradients
template <int GEOM>
__device__ __forceinline__ void shape(const double* xh, double* phi, double (*dphi)[Traits<GEOM>::D])
{
  using T = Traits<GEOM>;
  constexpr int D = T::D;
  if constexpr (T::SIMPLEX) {
    double l0 = 1.;
#pragma unroll
    for (int a = 0; a < D; ++a)
      l0 -= xh[a];
    phi[0] = l0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      phi[a + 1] = xh[a];
      dphi[0][a] = -1.;
#pragma unroll
      for (int c = 0; c < D; ++c)
        dphi[c + 1][a] = (a == c) ? 1. : 0.;
    }
  } else {
#pragma unroll
    for (int b = 0; b < T::NV; ++b) {
      double v = 1.;
#pragma unroll
      for (int a = 0; a < D; ++a)
        v *= ((b >> a) & 1) ? xh[a] : (1. - xh[a]);
      phi[b] = v;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        double g = ((b >> a) & 1) ? 1. : -1.;
#pragma unroll
        for (int c = 0; c < D; ++c)
          if (c != a)
            g *= ((b >> c) & 1) ? xh[c] : (1. - xh[c]);
        dphi[b][a] = g;
      }
    }
  }
}

// Lagrange basis of order 1 or 2 in the local node order of lagrange_node_vertices() (setup.cpp),
// LagrangeLocalFiniteElement<EquidistantPointSet> as created by local/finite-elements/lagrange.hh:141
template <int GEOM, int ORDER>
struct Basis {
  static constexpr int D = Traits<GEOM>::D;
  static constexpr int ND = (ORDER == 1) ? Traits<GEOM>::NV
                            : Traits<GEOM>::SIMPLEX ? (D + 1) * (D + 2) / 2
                                                    : (D == 2 ? 9 : 27);
};

// 1d quadratic Lagrange polynomials on the nodes 0, 1/2, 1 and their derivatives
__device__ __forceinline__ double lag2(int m, double t)
{
  return m == 0 ? 2. * (t - 0.5) * (t - 1.) : m == 1 ? -4. * t * (t - 1.) : 2. * t * (t - 0.5);
}
__device__ __forceinline__ double dlag2(int m, double t)
{
  return m == 0 ? 4. * t - 3. : m == 1 ? 4. - 8. * t : 4. * t - 1.;
}

template <int GEOM, int ORDER>
__device__ __forceinline__ void basis(const double* xh, double* phi, double (*dphi)[Traits<GEOM>::D])
{
  using T = Traits<GEOM>;
  constexpr int D = T::D, ND = Basis<GEOM, ORDER>::ND;
  if constexpr (ORDER == 1) {
    shape<GEOM>(xh, phi, dphi);
  } else if constexpr (T::SIMPLEX) {
    // barycentric coordinates: l_0 = 1 - sum x, l_{a+1} = x_a; vertices l (2l - 1), edges 4 l_a l_b
    double lam[D + 1], gl[D + 1][D];
    lam[0] = 1.;
#pragma unroll
    for (int a = 0; a < D; ++a)
      lam[0] -= xh[a], lam[a + 1] = xh[a];
#pragma unroll
    for (int v = 0; v <= D; ++v)
#pragma unroll
      for (int a = 0; a < D; ++a)
        gl[v][a] = (v == 0) ? -1. : (v == a + 1 ? 1. : 0.);
#pragma unroll
    for (int v = 0; v <= D; ++v) {
      phi[v] = lam[v] * (2. * lam[v] - 1.);
#pragma unroll
      for (int a = 0; a < D; ++a)
        dphi[v][a] = (4. * lam[v] - 1.) * gl[v][a];
    }
    int t = D + 1;
#pragma unroll
    for (int va = 0; va <= D; ++va)
#pragma unroll
      for (int vb = va + 1; vb <= D; ++vb, ++t) {
        phi[t] = 4. * lam[va] * lam[vb];
#pragma unroll
        for (int a = 0; a < D; ++a)
          dphi[t][a] = 4. * (lam[va] * gl[vb][a] + lam[vb] * gl[va][a]);
      }
  } else {
#pragma unroll
    for (int t = 0; t < ND; ++t) {
      const int idx[3] = {t % 3, (t / 3) % 3, t / 9};
      double v = 1.;
#pragma unroll
      for (int a = 0; a < D; ++a)
        v *= lag2(idx[a], xh[a]);
      phi[t] = v;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        double g = dlag2(idx[a], xh[a]);
#pragma unroll
        for (int c = 0; c < D; ++c)
          if (c != a)
            g *= lag2(idx[c], xh[c]);
        dphi[t][a] = g;
      }
    }
  }
}

// (K g) . h for a matrix-valued diffusion K (row-major D x D): (weight * ansatz_grad) * test_grad,
// local/integrands/laplace.hh:99-101
template <int D>
__device__ __forceinline__ double kmat_dot(const double (&K)[D * D], const double* g, const double* h)
{
  double s = 0.;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double kg = 0.;
#pragma unroll
    for (int b = 0; b < D; ++b)
      kg += K[a * D + b] * g[b];
    s += kg * h[a];
  }
  return s;
}

// the form's matrix-valued diffusion at grid element ge, if it has one
template <int D, class PP>
__device__ __forceinline__ bool load_tensor(const PP& P, int f, int64_t ge, double (&K)[D * D])
{
  if (P.tensor_kind[f] == 0)
    return false;
#pragma unroll
  for (int t = 0; t < D * D; ++t)
    K[t] = P.tensor_kind[f] == 1 ? P.tensor[f][ge * (D * D) + t] : P.tvalue[f][t];
  return true;
}

// physical gradient g = J^{-T} gh (spaces/basis/default.hh:152-175)
template <int D>
__device__ __forceinline__ void to_physical(const Geo<D>& g, const double* gh, double* out)
{
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double s = 0.;
#pragma unroll
    for (int c = 0; c < D; ++c)
      s += g.Ji[c][a] * gh[c];
    out[a] = s;
  }
}

template <int GEOM, int ORDER>
__global__ void __launch_bounds__(256) k_element_local(ElementParams P)
{
  using T = Traits<GEOM>;
  constexpr int D = T::D, ND = Basis<GEOM, ORDER>::ND;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= P.ne)
    return;
  const int64_t ge = P.eid ? P.eid[e] : e;
  Geo<D> g;
  load_geometry<GEOM>(P.coords, P.conn, e, g);
  const double* __restrict__ qt = P.qtab;
  if (P.write_matrix) {
    double A[ND][ND];
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int j = 0; j < ND; ++j)
        A[i][j] = 0.;
    for (int f = 0; f < P.nforms; ++f) {
      const double c = P.field[f] ? P.field[f][ge] : P.value[f];
      const double s = P.scaling[f];
      const double* __restrict__ cp = P.points[f]; // the coefficient at the rule's points of the element
      if (cp)
        cp += P.ef_qoff[f] ? P.ef_qoff[f][e] : e * P.rule[f].npts;
      double K[D * D];
      const bool tk = P.kind[f] == EF_LAPLACE && load_tensor<D>(P, f, ge, K);
      if (P.kind[f] == EF_LAPLACE && T::SIMPLEX && ORDER == 1 && !cp) {
        // P1 gradients are constant: the integrand is constant on the element and any
        // rule of the order chosen by integrals.hh:115 reduces to one evaluation
        double phi[ND], dphi[ND][D], G[ND][D];
        double xc[D];
#pragma unroll
        for (int a = 0; a < D; ++a)
          xc[a] = 1. / (D + 1);
        shape<GEOM>(xc, phi, dphi);
#pragma unroll
        for (int i = 0; i < ND; ++i)
          to_physical<D>(g, dphi[i], G[i]);
        const double factor = g.det * T::VOL;
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
          for (int j = i; j < ND; ++j) {
            double dotv = 0.;
            if (tk) {
              dotv = kmat_dot<D>(K, G[j], G[i]);
            } else {
#pragma unroll
              for (int a = 0; a < D; ++a)
                dotv += (c * G[j][a]) * G[i][a];
            }
            A[i][j] += s * (dotv * factor);
          }
        continue;
      }
      const QRule R = P.rule[f];
      for (int q = 0; q < R.npts; ++q) {
        double xh[D];
#pragma unroll
        for (int a = 0; a < D; ++a)
          xh[a] = qt[R.offset * D + q * D + a];
        const double w = qt[R.offset * D + R.npts * D + q];
        double phi[ND], dphi[ND][D];
        basis<GEOM, ORDER>(xh, phi, dphi);
        const double factor = g.det * w;
        const double cq = cp ? cp[q] : c;
        if (P.kind[f] == EF_LAPLACE) {
          double G[ND][D];
#pragma unroll
          for (int i = 0; i < ND; ++i)
            to_physical<D>(g, dphi[i], G[i]);
#pragma unroll
          for (int i = 0; i < ND; ++i)
#pragma unroll
            for (int j = i; j < ND; ++j) {
              double dotv = 0.;
              if (tk) {
                dotv = kmat_dot<D>(K, G[j], G[i]);
              } else {
#pragma unroll
                for (int a = 0; a < D; ++a)
                  dotv += (cq * G[j][a]) * G[i][a];
              }
              A[i][j] += s * (dotv * factor);
            }
        } else {
#pragma unroll
          for (int i = 0; i < ND; ++i)
#pragma unroll
            for (int j = i; j < ND; ++j)
              A[i][j] += s * (((cq * phi[i]) * phi[j]) * factor);
        }
      }
    }
    // the Laplace and product forms are symmetric: only the upper triangle (i <= j) is computed and
    // stored, entry-major (entry t of every element contiguous, coalesced across lanes); the scatter
    // reads entry (min, max) for both (i, j) and (j, i), so assembled matrices are symmetric bit for bit
    double* __restrict__ out = P.staging + P.elem_base;
    int t = 0;
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int j = i; j < ND; ++j, ++t)
        out[int64_t(t) * P.ne + e] = A[i][j];
  }
  if (P.nsrc > 0) {
    double b[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i)
      b[i] = 0.;
    for (int sidx = 0; sidx < P.nsrc; ++sidx) {
      const QRule R = P.src_rule[sidx];
      for (int q = 0; q < R.npts; ++q) {
        double xh[D];
#pragma unroll
        for (int a = 0; a < D; ++a)
          xh[a] = qt[R.offset * D + q * D + a];
        const double w = qt[R.offset * D + R.npts * D + q];
        double phi[ND], dphi[ND][D];
        basis<GEOM, ORDER>(xh, phi, dphi);
        double fv;
        if (P.src_kind[sidx] == HGDT_SOURCE_CONSTANT)
          fv = P.src_value[sidx];
        else if (P.src_kind[sidx] == HGDT_SOURCE_PER_ELEMENT)
          fv = P.src_data[sidx][ge];
        else
          fv = P.src_data[sidx][(P.src_qoff[sidx] ? P.src_qoff[sidx][e] : e * R.npts) + q];
#pragma unroll
        for (int i = 0; i < ND; ++i)
          b[i] += (phi[i] * fv) * g.det * w;
      }
    }
    double* __restrict__ out = P.staging + P.rhs_base;
#pragma unroll
    for (int i = 0; i < ND; ++i)
      out[int64_t(i) * P.ne + e] = b[i];
  }
}

// element matrices of an element with many local DoFs (Q2 on hexahedra: 27, whose 378 upper-triangle
// entries do not fit a lane's registers): one lane per (element, row i) accumulates the row's entries over
// the quadrature points and writes those with j >= i (and the right-hand side's entry i).  The basis
// functions are the products of their 1-d factors, formed per point in the order of basis<>() (the same
// values and gradients), and every entry is accumulated as in k_element_local.
__device__ __forceinline__ double pick3(int m, const double* v)
{
  return m == 0 ? v[0] : (m == 1 ? v[1] : v[2]);
}

template <int GEOM, int ORDER>
__global__ void __launch_bounds__(256) k_element_rows(ElementParams P)
{
  using T = Traits<GEOM>;
  constexpr int D = T::D, ND = Basis<GEOM, ORDER>::ND;
  static_assert(!T::SIMPLEX && ORDER == 2 && D == 3, "row lanes are provided for Q2 hexahedra");
  const int64_t lane = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (lane >= P.ne * ND)
    return;
  const int64_t e = lane / ND;
  const int64_t ge = P.eid ? P.eid[e] : e;
  const int i = int(lane % ND);
  const int ix = i % 3, iy = (i / 3) % 3, iz = i / 9;
  Geo<D> g;
  load_geometry<GEOM>(P.coords, P.conn, e, g);
  const double* __restrict__ qt = P.qtab;
  // 1-d factors at a point: l[a][m] = lag2(m, x_a), dl[a][m] = lag2'(m, x_a)
  auto factors = [&](const double* xh, double (&l)[3][3], double (&dl)[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m)
        l[a][m] = lag2(m, xh[a]), dl[a][m] = dlag2(m, xh[a]);
  };
  if (P.write_matrix) {
    double row[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j)
      row[j] = 0.;
    for (int f = 0; f < P.nforms; ++f) {
      const double c = P.field[f] ? P.field[f][ge] : P.value[f];
      const double s = P.scaling[f];
      const QRule R = P.rule[f];
      const double* __restrict__ cp = P.points[f]; // the coefficient at the rule's points of the element
      if (cp)
        cp += P.ef_qoff[f] ? P.ef_qoff[f][e] : e * R.npts;
      double K[9];
      const bool tk = P.kind[f] == EF_LAPLACE && load_tensor<3>(P, f, ge, K);
      for (int q = 0; q < R.npts; ++q) {
        double xh[D];
#pragma unroll
        for (int a = 0; a < D; ++a)
          xh[a] = qt[R.offset * D + q * D + a];
        const double w = qt[R.offset * D + R.npts * D + q];
        double l[3][3], dl[3][3];
        factors(xh, l, dl);
        const double lx = pick3(ix, l[0]), ly = pick3(iy, l[1]), lz = pick3(iz, l[2]);
        const double factor = g.det * w;
        const double cq = cp ? cp[q] : c;
        if (P.kind[f] == EF_LAPLACE) {
          const double gh[3] = {(pick3(ix, dl[0]) * ly) * lz, (pick3(iy, dl[1]) * lx) * lz,
                                (pick3(iz, dl[2]) * lx) * ly};
          double Gi[3];
          to_physical<3>(g, gh, Gi);
#pragma unroll
          for (int j = 0; j < ND; ++j) {
            const int jx = j % 3, jy = (j / 3) % 3, jz = j / 9;
            const double ghj[3] = {(dl[0][jx] * l[1][jy]) * l[2][jz], (dl[1][jy] * l[0][jx]) * l[2][jz],
                                   (dl[2][jz] * l[0][jx]) * l[1][jy]};
            double Gj[3];
            to_physical<3>(g, ghj, Gj);
            double dotv = 0.;
            if (tk) {
              dotv = kmat_dot<3>(K, Gj, Gi);
            } else {
#pragma unroll
              for (int a = 0; a < 3; ++a)
                dotv += (cq * Gj[a]) * Gi[a];
            }
            row[j] += s * (dotv * factor);
          }
        } else {
          const double phii = (lx * ly) * lz;
#pragma unroll
          for (int j = 0; j < ND; ++j) {
            const int jx = j % 3, jy = (j / 3) % 3, jz = j / 9;
            const double phij = (l[0][jx] * l[1][jy]) * l[2][jz];
            row[j] += s * (((cq * phii) * phij) * factor);
          }
        }
      }
    }
    // packed upper triangle, entry-major: entry (i, j >= i) at t = i ND - i (i - 1) / 2 + (j - i)
    double* __restrict__ out = P.staging + P.elem_base;
    const int t0 = i * ND - i * (i - 1) / 2 - i;
#pragma unroll
    for (int j = 0; j < ND; ++j)
      if (j >= i)
        out[int64_t(t0 + j) * P.ne + e] = row[j];
  }
  if (P.nsrc > 0) {
    double b = 0.;
    for (int sidx = 0; sidx < P.nsrc; ++sidx) {
      const QRule R = P.src_rule[sidx];
      for (int q = 0; q < R.npts; ++q) {
        double xh[D];
#pragma unroll
        for (int a = 0; a < D; ++a)
          xh[a] = qt[R.offset * D + q * D + a];
        const double w = qt[R.offset * D + R.npts * D + q];
        const double phii = (lag2(ix, xh[0]) * lag2(iy, xh[1])) * lag2(iz, xh[2]);
        double fv;
        if (P.src_kind[sidx] == HGDT_SOURCE_CONSTANT)
          fv = P.src_value[sidx];
        else if (P.src_kind[sidx] == HGDT_SOURCE_PER_ELEMENT)
          fv = P.src_data[sidx][ge];
        else
          fv = P.src_data[sidx][(P.src_qoff[sidx] ? P.src_qoff[sidx][e] : e * R.npts) + q];
        b += (phii * fv) * g.det * w;
      }
    }
    P.staging[P.rhs_base + int64_t(i) * P.ne + e] = b;
  }
}

// ---------------------------------------------------------------------------
// intersections: segments (2d), triangles and parallelograms (3d)
// ---------------------------------------------------------------------------
// local vertex k of reference face f: simplices number face f opposite vertex D - f, cubes hold face
// 2a + s on the vertices whose bit a equals s (the reference's face numbering; vertices ascending)
template <int GEOM>
__device__ __forceinline__ int face_vertex(int f, int k)
{
  using T = Traits<GEOM>;
  if constexpr (T::SIMPLEX)
    return k + (k >= T::D - f ? 1 : 0);
  const int a = f >> 1, s = f & 1;
  return (k & ((1 << a) - 1)) | (s << a) | ((k >> a) << (a + 1));
}

// affine face X(s, t) = P0 + s E1 + t E2 over the corners P0, P1, P2 of the face (2d: X(s) = P0 + s E1);
// integration element |E1| (2d) or |E1 x E2| (3d); unit normal pointing away from the inside element's
// centroid; diameter h = largest distance between two corners (the length in 2d)
template <int D>
struct FaceFrame {
  double P0[D], E1[D], E2[D], n[D];
  double ie, h;
  __device__ __forceinline__ void point(double s, double t, double* X) const
  {
#pragma unroll
    for (int a = 0; a < D; ++a)
      X[a] = (D == 2) ? P0[a] + s * E1[a] : P0[a] + s * E1[a] + t * E2[a];
  }
};

template <int GEOM>
__device__ __forceinline__ void face_frame(const double* __restrict__ coords, const int32_t* __restrict__ conn,
                                           int64_t e, int f, FaceFrame<Traits<GEOM>::D>& F,
                                           int64_t stride = Traits<GEOM>::NV)
{
  using T = Traits<GEOM>;
  constexpr int D = T::D, NC = (D == 2) ? 2 : (T::SIMPLEX ? 3 : 4);
  double P[NC][D], cen[D];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int32_t v = conn[e * stride + face_vertex<GEOM>(f, k)];
#pragma unroll
    for (int a = 0; a < D; ++a)
      P[k][a] = coords[int64_t(v) * D + a];
  }
#pragma unroll
  for (int a = 0; a < D; ++a)
    cen[a] = 0.;
  for (int v = 0; v < T::NV; ++v) {
    const int32_t vv = conn[e * stride + v];
#pragma unroll
    for (int a = 0; a < D; ++a)
      cen[a] += coords[int64_t(vv) * D + a] / T::NV;
  }
#pragma unroll
  for (int a = 0; a < D; ++a)
    F.P0[a] = P[0][a], F.E1[a] = P[1][a] - P[0][a], F.E2[a] = 0.;
  if constexpr (D == 2) {
    F.ie = F.h = sqrt(F.E1[0] * F.E1[0] + F.E1[1] * F.E1[1]);
    double nx = F.E1[1] / F.ie, ny = -F.E1[0] / F.ie;
    if (nx * (P[0][0] - cen[0]) + ny * (P[0][1] - cen[1]) < 0.)
      nx = -nx, ny = -ny;
    F.n[0] = nx + 0., F.n[1] = ny + 0.;
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a)
      F.E2[a] = P[2][a] - P[0][a];
    double N[3];
    N[0] = F.E1[1] * F.E2[2] - F.E1[2] * F.E2[1];
    N[1] = F.E1[2] * F.E2[0] - F.E1[0] * F.E2[2];
    N[2] = F.E1[0] * F.E2[1] - F.E1[1] * F.E2[0];
    F.ie = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    double side = 0.;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      side += N[a] * (P[0][a] - cen[a]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
      F.n[a] = (side < 0. ? -N[a] : N[a]) / F.ie + 0.;
    double h2 = 0.;
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
      for (int l = k + 1; l < NC; ++l) {
        double d2 = 0.;
#pragma unroll
        for (int a = 0; a < 3; ++a)
          d2 += (P[k][a] - P[l][a]) * (P[k][a] - P[l][a]);
        h2 = fmax(h2, d2);
      }
    F.h = sqrt(h2);
  }
}

// point q of a face rule (FD = D - 1 coordinates per point, weights after the points)
template <int FD>
__device__ __forceinline__ void face_qp(const double* __restrict__ qtab, const QRule& R, int q, double& s, double& t,
                                        double& w)
{
  const double* __restrict__ b = qtab + int64_t(R.offset) * FD;
  s = b[q * FD];
  t = (FD == 2) ? b[q * FD + 1] : 0.;
  w = b[R.npts * FD + q];
}

template <int GEOM, int ORDER>
__device__ __forceinline__ void eval_at(const Geo<Traits<GEOM>::D>& g, const double* X, double* phi,
                                        double (*G)[Traits<GEOM>::D])
{
  constexpr int D = Traits<GEOM>::D, ND = Basis<GEOM, ORDER>::ND;
  double xh[D], dx[D];
#pragma unroll
  for (int a = 0; a < D; ++a)
    dx[a] = X[a] - g.x0[a];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double v = g.Ji[a][0] * dx[0];
#pragma unroll
    for (int b = 1; b < D; ++b)
      v += g.Ji[a][b] * dx[b];
    xh[a] = v;
  }
  double dphi[ND][D];
  basis<GEOM, ORDER>(xh, phi, dphi);
#pragma unroll
  for (int i = 0; i < ND; ++i)
    to_physical<D>(g, dphi[i], G[i]);
}

// (k G) . n, or (K G) . n with a matrix-valued diffusion K (its d x d block of the element; laplace-ipdg.hh:
// 158-186: (diffusion * grad) * normal)
template <int D>
__device__ __forceinline__ double flux(double k, const double* G, const double* n, const double* K = nullptr)
{
  if (K) {
    double v = 0.;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      double kg = 0.;
#pragma unroll
      for (int b = 0; b < D; ++b)
        kg += K[a * D + b] * G[b];
      v += kg * n[a];
    }
    return v;
  }
  double v = (k * G[0]) * n[0];
#pragma unroll
  for (int a = 1; a < D; ++a)
    v += (k * G[a]) * n[a];
  return v;
}

// n . (w n), or n . (W n) with a matrix-valued weight (laplace-ipdg.hh:149-150, ipdg.hh:149-153)
template <int D>
__device__ __forceinline__ double n_w_n(const double* n, double w, const double* W = nullptr)
{
  if (W) {
    double v = 0.;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      double wn = 0.;
#pragma unroll
      for (int b = 0; b < D; ++b)
        wn += W[a * D + b] * n[b];
      v += n[a] * wn;
    }
    return v;
  }
  double v = n[0] * (w * n[0]);
#pragma unroll
  for (int a = 1; a < D; ++a)
    v += n[a] * (w * n[a]);
  return v;
}

// the d x d block of element e of a tensor field (NULL: none)
template <int D>
__device__ __forceinline__ const double* tblock(const double* t, int64_t e)
{
  return t ? t + e * (D * D) : nullptr;
}

// one inner intersection per lane: the four blocks (in_in, in_out, out_in, out_out) of every appended
// coupling form, RB rows of each block at a time (all rows at once up to 4 basis functions).  With
// symmetric forms the diagonal blocks are accumulated from their upper triangle and out_in as the transpose
// of in_out: the stored blocks are symmetric bit for bit
template <int GEOM, int ORDER>
__global__ void __launch_bounds__(256) k_inner_local(FaceParams P)
{
  using T = Traits<GEOM>;
  constexpr int D = T::D, ND = Basis<GEOM, ORDER>::ND, K2 = ND * ND;
  // rows of the four blocks held at once: all of them up to 4 basis functions, else at most 16 / ND
  constexpr int RB = ND <= 4 ? ND : (16 / ND > 0 ? 16 / ND : 1);
  const int64_t fi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (fi >= P.nfaces)
    return;
  const int64_t e = P.faces[4 * fi + 0], o = P.faces[4 * fi + 2];
  const int f = P.faces[4 * fi + 1];
  Geo<D> gin, gout;
  load_geometry<GEOM>(P.coords, P.conn, e, gin);
  load_geometry<GEOM>(P.coords, P.conn, o, gout);
  FaceFrame<D> F;
  face_frame<GEOM>(P.coords, P.conn, e, f, F);
  const double* n = F.n;
  const bool sym = P.symmetric != 0;
  double* __restrict__ out = P.staging + P.base + fi * (4 * K2);
#pragma unroll
  for (int i0 = 0; i0 < ND; i0 += RB) {
    double B[4][RB][ND];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int j = 0; j < ND; ++j)
          B[k][r][j] = 0.;
    for (int fm = 0; fm < P.nforms; ++fm) {
      const double kin = P.kappa[fm] ? P.kappa[fm][e] : P.kappa_value[fm];
      const double kout = P.kappa[fm] ? P.kappa[fm][o] : P.kappa_value[fm];
      const double win = P.omega[fm] ? P.omega[fm][e] : P.omega_value[fm];
      const double wout = P.omega[fm] ? P.omega[fm][o] : P.omega_value[fm];
      const double *Kin = tblock<D>(P.kappa_t[fm], e), *Kout = tblock<D>(P.kappa_t[fm], o);
      const double *Win = tblock<D>(P.omega_t[fm], e), *Wout = tblock<D>(P.omega_t[fm], o);
      const double sc = P.scaling[fm], prm = P.param[fm];
      const QRule R = P.rule[fm];
      for (int q = 0; q < R.npts; ++q) {
        double s, t, wq, X[D];
        face_qp<D - 1>(P.qtab, R, q, s, t, wq);
        F.point(s, t, X);
        double pin[ND], pout[ND], Gin[ND][D], Gout[ND][D];
        eval_at<GEOM, ORDER>(gin, X, pin, Gin);
        eval_at<GEOM, ORDER>(gout, X, pout, Gout);
        const double delta_plus = n_w_n<D>(n, wout, Wout);
        const double delta_minus = n_w_n<D>(n, win, Win);
        const double factor = F.ie * wq;
        if (P.kind[fm] == FF_INNER_COUPLING) {
          const double wm = delta_plus / (delta_plus + delta_minus);
          const double wp = delta_minus / (delta_plus + delta_minus);
          double fin[ND], fout[ND]; // (kappa grad phi) . n
#pragma unroll
          for (int i = 0; i < ND; ++i) {
            fin[i] = flux<D>(kin, Gin[i], n, Kin);
            fout[i] = flux<D>(kout, Gout[i], n, Kout);
          }
#pragma unroll
          for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int j = 0; j < ND; ++j) {
              const int i = i0 + r;
              const int a = (sym && i > j) ? j : i, b = (sym && i > j) ? i : j;
              const double ii = -1.0 * wm * fin[b] * pin[a] + -1.0 * prm * wm * pin[b] * fin[a];
              const double io = -1.0 * wp * fout[j] * pin[i] + prm * wm * pout[j] * fin[i];
              const double oi = sym ? -1.0 * wp * fout[i] * pin[j] + prm * wm * pout[i] * fin[j]
                                    : wm * fin[j] * pout[i] + -1.0 * prm * wp * pin[j] * fout[i];
              const double oo = wp * fout[b] * pout[a] + prm * wp * pout[b] * fout[a];
              B[0][r][j] += sc * (ii * factor);
              B[1][r][j] += sc * (io * factor);
              B[2][r][j] += sc * (oi * factor);
              B[3][r][j] += sc * (oo * factor);
            }
        } else {
          const double weight = (delta_plus * delta_minus) / (delta_plus + delta_minus);
          const double pen = (prm * weight) / F.h;
#pragma unroll
          for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int j = 0; j < ND; ++j) {
              const int i = i0 + r;
              const int a = (sym && i > j) ? j : i, b = (sym && i > j) ? i : j;
              B[0][r][j] += sc * ((pen * pin[b] * pin[a]) * factor);
              B[1][r][j] += sc * ((-1.0 * pen * pout[j] * pin[i]) * factor);
              B[2][r][j] += sc * ((sym ? -1.0 * pen * pout[i] * pin[j] : -1.0 * pen * pin[j] * pout[i]) * factor);
              B[3][r][j] += sc * ((pen * pout[b] * pout[a]) * factor);
            }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int j = 0; j < ND; ++j)
          out[k * K2 + (i0 + r) * ND + j] = B[k][r][j];
  }
}

// block K of the four blocks of an inner intersection between elements of types GIN (inside) and GOUT
// (outside) on a mixed grid: 0 in_in [NI][NI], 1 in_out [NI][NO], 2 out_in [NO][NI], 3 out_out [NO][NO]; every
// entry is the expression of k_inner_local for the same rows and columns, accumulated in the same order
// (forms, then points), so faces between elements of one type give k_inner_local's bits.  Row chunks of at
// most 16 entries per lane.
template <int K, int GIN, int GOUT, int ORDER>
__device__ __forceinline__ void mixed_face_block(const FaceParams& P, int64_t e, int64_t o, const Geo<Traits<GIN>::D>& gin,
                                                 const Geo<Traits<GIN>::D>& gout, const FaceFrame<Traits<GIN>::D>& F,
                                                 bool sym, double* __restrict__ out)
{
  constexpr int D = Traits<GIN>::D, NI = Basis<GIN, ORDER>::ND, NO = Basis<GOUT, ORDER>::ND;
  constexpr int NR = (K < 2) ? NI : NO, NC = (K == 0 || K == 2) ? NI : NO;
  constexpr int RB = NC <= 4 ? NR : (16 / NC > 0 ? 16 / NC : 1);
  const double* n = F.n;
  const int LS = P.ls;
#pragma unroll
  for (int i0 = 0; i0 < NR; i0 += RB) {
    double B[RB][NC];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int j = 0; j < NC; ++j)
        B[r][j] = 0.;
    for (int fm = 0; fm < P.nforms; ++fm) {
      const double kin = P.kappa[fm] ? P.kappa[fm][e] : P.kappa_value[fm];
      const double kout = P.kappa[fm] ? P.kappa[fm][o] : P.kappa_value[fm];
      const double win = P.omega[fm] ? P.omega[fm][e] : P.omega_value[fm];
      const double wout = P.omega[fm] ? P.omega[fm][o] : P.omega_value[fm];
      const double *Kin = tblock<D>(P.kappa_t[fm], e), *Kout = tblock<D>(P.kappa_t[fm], o);
      const double *Win = tblock<D>(P.omega_t[fm], e), *Wout = tblock<D>(P.omega_t[fm], o);
      const double sc = P.scaling[fm], prm = P.param[fm];
      const QRule R = P.rule[fm];
      for (int q = 0; q < R.npts; ++q) {
        double s, t, wq, X[D];
        face_qp<D - 1>(P.qtab, R, q, s, t, wq);
        F.point(s, t, X);
        double pin[NI], pout[NO], Gin[NI][D], Gout[NO][D];
        eval_at<GIN, ORDER>(gin, X, pin, Gin);
        eval_at<GOUT, ORDER>(gout, X, pout, Gout);
        const double delta_plus = n_w_n<D>(n, wout, Wout);
        const double delta_minus = n_w_n<D>(n, win, Win);
        const double factor = F.ie * wq;
        if (P.kind[fm] == FF_INNER_COUPLING) {
          const double wm = delta_plus / (delta_plus + delta_minus);
          const double wp = delta_minus / (delta_plus + delta_minus);
          double fin[NI], fout[NO];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            fin[i] = flux<D>(kin, Gin[i], n, Kin);
#pragma unroll
          for (int i = 0; i < NO; ++i)
            fout[i] = flux<D>(kout, Gout[i], n, Kout);
#pragma unroll
          for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int j = 0; j < NC; ++j) {
              const int i = i0 + r;
              if (i >= NR)
                continue;
              const int a = (sym && i > j) ? j : i, b = (sym && i > j) ? i : j;
              double v;
              if constexpr (K == 0)
                v = -1.0 * wm * fin[b] * pin[a] + -1.0 * prm * wm * pin[b] * fin[a];
              else if constexpr (K == 1)
                v = -1.0 * wp * fout[j] * pin[i] + prm * wm * pout[j] * fin[i];
              else if constexpr (K == 2)
                v = sym ? -1.0 * wp * fout[i] * pin[j] + prm * wm * pout[i] * fin[j]
                        : wm * fin[j] * pout[i] + -1.0 * prm * wp * pin[j] * fout[i];
              else
                v = wp * fout[b] * pout[a] + prm * wp * pout[b] * fout[a];
              B[r][j] += sc * (v * factor);
            }
        } else {
          const double weight = (delta_plus * delta_minus) / (delta_plus + delta_minus);
          const double pen = (prm * weight) / F.h;
#pragma unroll
          for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int j = 0; j < NC; ++j) {
              const int i = i0 + r;
              if (i >= NR)
                continue;
              const int a = (sym && i > j) ? j : i, b = (sym && i > j) ? i : j;
              double v;
              if constexpr (K == 0)
                v = pen * pin[b] * pin[a];
              else if constexpr (K == 1)
                v = -1.0 * pen * pout[j] * pin[i];
              else if constexpr (K == 2)
                v = sym ? -1.0 * pen * pout[i] * pin[j] : -1.0 * pen * pin[j] * pout[i];
              else
                v = pen * pout[b] * pout[a];
              B[r][j] += sc * (v * factor);
            }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int j = 0; j < NC; ++j)
        if (i0 + r < NR)
          out[int64_t(K) * LS * LS + (i0 + r) * LS + j] = B[r][j];
  }
}

// inner intersections of a mixed grid between an inside element of part P.part_in (type GIN) and an
// outside element of part P.part_out (type GOUT): one lane per intersection of the walk, the lanes of
// other type pairs idle; blocks staged with rows of P.ls entries
template <int GIN, int GOUT, int ORDER>
__global__ void __launch_bounds__(256) k_inner_mixed(FaceParams P)
{
  constexpr int D = Traits<GIN>::D;
  static_assert(Traits<GOUT>::D == D, "both sides of an intersection have the grid's dimension");
  const int64_t fi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (fi >= P.nfaces)
    return;
  const int64_t e = P.faces[4 * fi + 0], o = P.faces[4 * fi + 2];
  const int f = P.faces[4 * fi + 1];
  if (P.etype[e] != P.part_in || P.etype[o] != P.part_out)
    return;
  Geo<D> gin, gout;
  load_geometry<GIN>(P.coords, P.conn, e, gin, P.cstride);
  load_geometry<GOUT>(P.coords, P.conn, o, gout, P.cstride);
  FaceFrame<D> F;
  face_frame<GIN>(P.coords, P.conn, e, f, F, P.cstride);
  const bool sym = P.symmetric != 0;
  double* __restrict__ out = P.staging + P.base + fi * (4 * int64_t(P.ls) * P.ls);
  mixed_face_block<0, GIN, GOUT, ORDER>(P, e, o, gin, gout, F, sym, out);
  mixed_face_block<1, GIN, GOUT, ORDER>(P, e, o, gin, gout, F, sym, out);
  mixed_face_block<2, GIN, GOUT, ORDER>(P, e, o, gin, gout, F, sym, out);
  mixed_face_block<3, GIN, GOUT, ORDER>(P, e, o, gin, gout, F, sym, out);
}

template <int GEOM, int ORDER>
__global__ void __launch_bounds__(256) k_boundary_local(FaceParams P)
{
  using T = Traits<GEOM>;
  constexpr int D = T::D, ND = Basis<GEOM, ORDER>::ND;
  const int64_t fi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (fi >= P.nfaces)
    return;
  const int64_t e = P.faces[2 * fi + 0];
  const int f = P.faces[2 * fi + 1];
  if (P.etype && P.etype[e] != P.part_in)
    return;
  Geo<D> gin;
  load_geometry<GEOM>(P.coords, P.conn, e, gin, P.cstride);
  FaceFrame<D> F;
  face_frame<GEOM>(P.coords, P.conn, e, f, F, P.cstride);
  const double* n = F.n;
  double B[ND][ND];
#pragma unroll
  for (int i = 0; i < ND; ++i)
#pragma unroll
    for (int j = 0; j < ND; ++j)
      B[i][j] = 0.;
  for (int fm = 0; fm < P.nforms; ++fm) {
    if (P.mask[fm] && !P.mask[fm][fi]) // ApplyOn::CustomBoundaryIntersections: not one of this form's faces
      continue;
    const double kin = P.kappa[fm] ? P.kappa[fm][e] : P.kappa_value[fm];
    const double win = P.omega[fm] ? P.omega[fm][e] : P.omega_value[fm];
    const double *Kin = tblock<D>(P.kappa_t[fm], e), *Win = tblock<D>(P.omega_t[fm], e);
    const double sc = P.scaling[fm], prm = P.param[fm];
    const QRule R = P.rule[fm];
    for (int q = 0; q < R.npts; ++q) {
      double s, t, wq, X[D];
      face_qp<D - 1>(P.qtab, R, q, s, t, wq);
      F.point(s, t, X);
      double pin[ND], Gin[ND][D];
      eval_at<GEOM, ORDER>(gin, X, pin, Gin);
      const double factor = F.ie * wq;
      if (P.kind[fm] == FF_DIRICHLET_COUPLING) {
        double fin[ND];
#pragma unroll
        for (int i = 0; i < ND; ++i)
          fin[i] = flux<D>(kin, Gin[i], n, Kin);
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
          for (int j = 0; j < ND; ++j)
            B[i][j] += sc * ((-1.0 * fin[j] * pin[i] + -1.0 * prm * pin[j] * fin[i]) * factor);
      } else {
        const double pen = (prm * n_w_n<D>(n, win, Win)) / F.h;
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
          for (int j = 0; j < ND; ++j)
            B[i][j] += sc * ((pen * pin[j] * pin[i]) * factor);
      }
    }
  }
  const int LS = P.ls;
  double* __restrict__ out = P.staging + P.base + fi * (int64_t(LS) * LS);
#pragma unroll
  for (int i = 0; i < ND; ++i)
#pragma unroll
    for (int j = 0; j < ND; ++j)
      out[i * LS + j] = (P.symmetric && i > j) ? B[j][i] : B[i][j];
}

// boundary functionals (LocalIntersectionIntegralFunctional::apply, local/functionals/integrals.hh:157-185),
// one boundary intersection per lane, all appended terms summed into the face's local vector:
//   DirichletCoupling as a unary integrand: -s g (kappa grad psi_i) . n   (laplace-ipdg.hh:300-320)
//   BoundaryPenalty(sigma, w).with_ansatz(g): sigma (n . w n) / h_F g psi_i (ipdg.hh:254-282)
//   LocalIntersectionProductIntegrand(w).with_ansatz(g): (w g) psi_i     (product.hh:355-377)
template <int GEOM, int ORDER>
__global__ void __launch_bounds__(256) k_boundary_rhs(BoundaryRhsParams P)
{
  using T = Traits<GEOM>;
  constexpr int D = T::D, ND = Basis<GEOM, ORDER>::ND;
  const int64_t fi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (fi >= P.nfaces)
    return;
  const int64_t e = P.faces[2 * fi + 0];
  const int f = P.faces[2 * fi + 1];
  if (P.etype && P.etype[e] != P.part_in)
    return;
  Geo<D> gin;
  load_geometry<GEOM>(P.coords, P.conn, e, gin, P.cstride);
  FaceFrame<D> F;
  face_frame<GEOM>(P.coords, P.conn, e, f, F, P.cstride);
  const double* n = F.n;
  double b[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i)
    b[i] = 0.;
  for (int tm = 0; tm < P.nterms; ++tm) {
    if (P.mask[tm] && !P.mask[tm][fi])
      continue;
    const double c = P.coeff[tm] ? P.coeff[tm][e] : P.coeff_value[tm];
    const double* Ct = tblock<D>(P.coeff_t[tm], e);
    const double prm = P.param[tm];
    const QRule R = P.rule[tm];
    for (int q = 0; q < R.npts; ++q) {
      double s, t, wq, X[D];
      face_qp<D - 1>(P.qtab, R, q, s, t, wq);
      F.point(s, t, X);
      double pin[ND], Gin[ND][D];
      eval_at<GEOM, ORDER>(gin, X, pin, Gin);
      const double g = P.data[tm] ? P.data[tm][fi * R.npts + q] : P.value[tm];
      const double factor = F.ie * wq;
      if (P.kind[tm] == HGDT_BOUNDARY_DIRICHLET_COUPLING) {
#pragma unroll
        for (int i = 0; i < ND; ++i)
          b[i] += (-1.0 * prm * g * flux<D>(c, Gin[i], n, Ct)) * factor;
      } else if (P.kind[tm] == HGDT_BOUNDARY_PENALTY) {
        const double pen = (prm * n_w_n<D>(n, c, Ct)) / F.h;
#pragma unroll
        for (int i = 0; i < ND; ++i)
          b[i] += (pen * g * pin[i]) * factor;
      } else {
#pragma unroll
        for (int i = 0; i < ND; ++i)
          b[i] += ((c * g) * pin[i]) * factor;
      }
    }
  }
  double* __restrict__ out = P.staging + P.base + fi * P.ls;
#pragma unroll
  for (int i = 0; i < ND; ++i)
    out[i] = b[i];
}

// ---- intersections of Q2 hexahedra: 27 basis functions per side, one lane per (face, row) ----
// The 1-d factors of the basis at the physical point X of an element (its reference coordinates by the
// inverse map, as eval_at), and basis function j's value and physical gradient from them.
__device__ __forceinline__ void q2hex_factors(const Geo<3>& g, const double* X, double (&l)[3][3], double (&dl)[3][3])
{
  double xh[3], dx[3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
    dx[a] = X[a] - g.x0[a];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double v = g.Ji[a][0] * dx[0];
#pragma unroll
    for (int b = 1; b < 3; ++b)
      v += g.Ji[a][b] * dx[b];
    xh[a] = v;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int m = 0; m < 3; ++m)
      l[a][m] = lag2(m, xh[a]), dl[a][m] = dlag2(m, xh[a]);
}
__device__ __forceinline__ void q2hex_basis(const Geo<3>& g, const double (&l)[3][3], const double (&dl)[3][3], int j,
                                            double& phi, double (&G)[3])
{
  const int jx = j % 3, jy = (j / 3) % 3, jz = j / 9;
  const double lx = pick3(jx, l[0]), ly = pick3(jy, l[1]), lz = pick3(jz, l[2]);
  phi = (lx * ly) * lz;
  const double gh[3] = {(pick3(jx, dl[0]) * ly) * lz, (pick3(jy, dl[1]) * lx) * lz, (pick3(jz, dl[2]) * lx) * ly};
  to_physical<3>(g, gh, G);
}

// inner intersections: lane (face, r), r < 27: row r of in_in and in_out; r >= 27: row r - 27 of out_in and
// out_out.  Every entry is k_inner_local's expression (the diagonal blocks from their upper triangle and
// out_in as the transpose of in_out for symmetric forms, so the entries of two lanes that mirror each
// other are the same bits), accumulated over forms and points in the same order.
__global__ void __launch_bounds__(256) k_inner_rows_q2hex(FaceParams P)
{
  constexpr int ND = 27, D = 3;
  const int64_t lane = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (lane >= P.nfaces * 2 * ND)
    return;
  const int64_t fi = lane / (2 * ND);
  const int rr = int(lane % (2 * ND));
  const bool outside_rows = rr >= ND;
  const int i = rr % ND;
  const int64_t e = P.faces[4 * fi + 0], o = P.faces[4 * fi + 2];
  const int f = P.faces[4 * fi + 1];
  Geo<D> gin, gout;
  load_geometry<HGDT_HEX>(P.coords, P.conn, e, gin);
  load_geometry<HGDT_HEX>(P.coords, P.conn, o, gout);
  FaceFrame<D> F;
  face_frame<HGDT_HEX>(P.coords, P.conn, e, f, F);
  const double* n = F.n;
  const bool sym = P.symmetric != 0;
  double A[ND], B[ND]; // inside rows: in_in, in_out; outside rows: out_in, out_out
#pragma unroll
  for (int j = 0; j < ND; ++j)
    A[j] = B[j] = 0.;
  for (int fm = 0; fm < P.nforms; ++fm) {
    const double kin = P.kappa[fm] ? P.kappa[fm][e] : P.kappa_value[fm];
    const double kout = P.kappa[fm] ? P.kappa[fm][o] : P.kappa_value[fm];
    const double win = P.omega[fm] ? P.omega[fm][e] : P.omega_value[fm];
    const double wout = P.omega[fm] ? P.omega[fm][o] : P.omega_value[fm];
    const double *Kin = tblock<D>(P.kappa_t[fm], e), *Kout = tblock<D>(P.kappa_t[fm], o);
    const double *Win = tblock<D>(P.omega_t[fm], e), *Wout = tblock<D>(P.omega_t[fm], o);
    const double sc = P.scaling[fm], prm = P.param[fm];
    const QRule R = P.rule[fm];
    for (int q = 0; q < R.npts; ++q) {
      double s, t, wq, X[D];
      face_qp<D - 1>(P.qtab, R, q, s, t, wq);
      F.point(s, t, X);
      double li[3][3], dli[3][3], lo[3][3], dlo[3][3];
      q2hex_factors(gin, X, li, dli);
      q2hex_factors(gout, X, lo, dlo);
      const double delta_plus = n_w_n<D>(n, wout, Wout);
      const double delta_minus = n_w_n<D>(n, win, Win);
      const double factor = F.ie * wq;
      const bool coupling = P.kind[fm] == FF_INNER_COUPLING;
      const double wm = delta_plus / (delta_plus + delta_minus);
      const double wp = delta_minus / (delta_plus + delta_minus);
      const double pen = (prm * ((delta_plus * delta_minus) / (delta_plus + delta_minus))) / F.h;
      // the lane's own row function (inside or outside) and its flux
      double pr, Gr[3];
      if (outside_rows)
        q2hex_basis(gout, lo, dlo, i, pr, Gr);
      else
        q2hex_basis(gin, li, dli, i, pr, Gr);
      const double fr = flux<D>(outside_rows ? kout : kin, Gr, n, outside_rows ? Kout : Kin);
      for (int j = 0; j < ND; ++j) {
        double pin, Gin[3], pout, Gout[3];
        q2hex_basis(gin, li, dli, j, pin, Gin);
        q2hex_basis(gout, lo, dlo, j, pout, Gout);
        const double fin = flux<D>(kin, Gin, n, Kin), fout = flux<D>(kout, Gout, n, Kout);
        double va, vb;
        if (!outside_rows) {
          // in_in (i, j): from the upper triangle when symmetric; in_out (i, j)
          const bool sw = sym && i > j;
          const double pa = sw ? pin : pr, fa = sw ? fin : fr, pb = sw ? pr : pin, fb = sw ? fr : fin;
          if (coupling) {
            va = -1.0 * wm * fb * pa + -1.0 * prm * wm * pb * fa;
            vb = -1.0 * wp * fout * pr + prm * wm * pout * fr;
          } else {
            va = pen * pb * pa;
            vb = -1.0 * pen * pout * pr;
          }
        } else {
          // out_in (i, j): the transpose of in_out when symmetric; out_out (i, j) from the upper triangle
          const bool sw = sym && i > j;
          const double pa = sw ? pout : pr, fa = sw ? fout : fr, pb = sw ? pr : pout, fb = sw ? fr : fout;
          if (coupling) {
            va = sym ? -1.0 * wp * fr * pin + prm * wm * pr * fin : wm * fin * pr + -1.0 * prm * wp * pin * fr;
            vb = wp * fb * pa + prm * wp * pb * fa;
          } else {
            va = sym ? -1.0 * pen * pr * pin : -1.0 * pen * pin * pr;
            vb = pen * pb * pa;
          }
        }
        A[j] += sc * (va * factor);
        B[j] += sc * (vb * factor);
      }
    }
  }
  constexpr int K2 = ND * ND;
  double* __restrict__ out = P.staging + P.base + fi * (4 * K2);
  const int ka = outside_rows ? 2 : 0, kb = outside_rows ? 3 : 1;
#pragma unroll
  for (int j = 0; j < ND; ++j) {
    out[ka * K2 + i * ND + j] = A[j];
    out[kb * K2 + i * ND + j] = B[j];
  }
}

// boundary intersections of Q2 hexahedra: lane (face, row i); for symmetric forms an entry (i, j) with
// i > j takes the value k_boundary_local stores there, B[j][i]
__global__ void __launch_bounds__(256) k_boundary_rows_q2hex(FaceParams P)
{
  constexpr int ND = 27, D = 3;
  const int64_t lane = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (lane >= P.nfaces * ND)
    return;
  const int64_t fi = lane / ND;
  const int i = int(lane % ND);
  const int64_t e = P.faces[2 * fi + 0];
  const int f = P.faces[2 * fi + 1];
  Geo<D> gin;
  load_geometry<HGDT_HEX>(P.coords, P.conn, e, gin);
  FaceFrame<D> F;
  face_frame<HGDT_HEX>(P.coords, P.conn, e, f, F);
  const double* n = F.n;
  double row[ND];
#pragma unroll
  for (int j = 0; j < ND; ++j)
    row[j] = 0.;
  for (int fm = 0; fm < P.nforms; ++fm) {
    if (P.mask[fm] && !P.mask[fm][fi])
      continue;
    const double kin = P.kappa[fm] ? P.kappa[fm][e] : P.kappa_value[fm];
    const double win = P.omega[fm] ? P.omega[fm][e] : P.omega_value[fm];
    const double *Kin = tblock<D>(P.kappa_t[fm], e), *Win = tblock<D>(P.omega_t[fm], e);
    const double sc = P.scaling[fm], prm = P.param[fm];
    const QRule R = P.rule[fm];
    for (int q = 0; q < R.npts; ++q) {
      double s, t, wq, X[D];
      face_qp<D - 1>(P.qtab, R, q, s, t, wq);
      F.point(s, t, X);
      double l[3][3], dl[3][3];
      q2hex_factors(gin, X, l, dl);
      double pr, Gr[3];
      q2hex_basis(gin, l, dl, i, pr, Gr);
      const double fr = flux<D>(kin, Gr, n, Kin);
      const double factor = F.ie * wq;
      const double pen = (prm * n_w_n<D>(n, win, Win)) / F.h;
      for (int j = 0; j < ND; ++j) {
        double pj, Gj[3];
        q2hex_basis(gin, l, dl, j, pj, Gj);
        const double fj = flux<D>(kin, Gj, n, Kin);
        // entry (r, c) of k_boundary_local: -fin[c] pin[r] - prm pin[c] fin[r] or pen pin[c] pin[r]; the stored
        // (i, j) is (r, c) = (j, i) when symmetric and i > j
        const bool sw = P.symmetric && i > j;
        const double prr = sw ? pj : pr, frr = sw ? fj : fr, pc = sw ? pr : pj, fc = sw ? fr : fj;
        const double v = (P.kind[fm] == FF_DIRICHLET_COUPLING) ? -1.0 * fc * prr + -1.0 * prm * pc * frr
                                                               : pen * pc * prr;
        row[j] += sc * (v * factor);
      }
    }
  }
  double* __restrict__ out = P.staging + P.base + fi * (ND * ND);
#pragma unroll
  for (int j = 0; j < ND; ++j)
    out[i * ND + j] = row[j];
}

// boundary functionals on Q2 hexahedra: lane (face, row i), the terms of k_boundary_rhs for entry i
__global__ void __launch_bounds__(256) k_boundary_rhs_rows_q2hex(BoundaryRhsParams P)
{
  constexpr int ND = 27, D = 3;
  const int64_t lane = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (lane >= P.nfaces * ND)
    return;
  const int64_t fi = lane / ND;
  const int i = int(lane % ND);
  const int64_t e = P.faces[2 * fi + 0];
  const int f = P.faces[2 * fi + 1];
  Geo<D> gin;
  load_geometry<HGDT_HEX>(P.coords, P.conn, e, gin);
  FaceFrame<D> F;
  face_frame<HGDT_HEX>(P.coords, P.conn, e, f, F);
  const double* n = F.n;
  double b = 0.;
  for (int tm = 0; tm < P.nterms; ++tm) {
    if (P.mask[tm] && !P.mask[tm][fi])
      continue;
    const double c = P.coeff[tm] ? P.coeff[tm][e] : P.coeff_value[tm];
    const double* Ct = tblock<D>(P.coeff_t[tm], e);
    const double prm = P.param[tm];
    const QRule R = P.rule[tm];
    for (int q = 0; q < R.npts; ++q) {
      double s, t, wq, X[D];
      face_qp<D - 1>(P.qtab, R, q, s, t, wq);
      F.point(s, t, X);
      double l[3][3], dl[3][3];
      q2hex_factors(gin, X, l, dl);
      double pr, Gr[3];
      q2hex_basis(gin, l, dl, i, pr, Gr);
      const double g = P.data[tm] ? P.data[tm][fi * R.npts + q] : P.value[tm];
      const double factor = F.ie * wq;
      if (P.kind[tm] == HGDT_BOUNDARY_DIRICHLET_COUPLING)
        b += (-1.0 * prm * g * flux<D>(c, Gr, n, Ct)) * factor;
      else if (P.kind[tm] == HGDT_BOUNDARY_PENALTY)
        b += (((prm * n_w_n<D>(n, c, Ct)) / F.h) * g * pr) * factor;
      else
        b += ((c * g) * pr) * factor;
    }
  }
  P.staging[P.base + fi * ND + i] = b;
}

// fixed-order gather of the staged local contributions into the CSR values
// sum of staging[contrib[c]] for c = c0 .. c1 - 1, added in that order (the walk order); the loads of U
// contributions are issued together so that a lane keeps U index / value pairs in flight
template <int U>
__device__ __forceinline__ double ordered_sum(const uint32_t* __restrict__ contrib, const double* __restrict__ staging,
                                              uint32_t c0, uint32_t c1, double v = 0.)
{
  for (uint32_t c = c0; c < c1; c += U) {
    uint32_t k[U];
    double x[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      k[u] = (c + u < c1) ? contrib[c + u] : 0u;
#pragma unroll
    for (int u = 0; u < U; ++u)
      x[u] = (c + u < c1) ? staging[k[u]] : 0.;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (c + u < c1)
        v += x[u];
  }
  return v;
}

// blocks are mapped XCD-aware (xcd_index): each XCD sums a contiguous range of slots, whose contributions
// come from a compact set of elements, so the staged lines they share are found in that XCD's L2
// accumulate: the contributions are added to the value held, in walk order (add_to_entry on a caller's
// matrix, bilinear-form-assemblers.hh:126); otherwise the slot starts at 0
__global__ void __launch_bounds__(256) k_gather_values(int64_t nnz, const uint32_t* __restrict__ slot_ptr,
                                                       const uint32_t* __restrict__ contrib,
                                                       const double* __restrict__ staging,
                                                       const uint8_t* __restrict__ flag, double* __restrict__ values,
                                                       int accumulate)
{
  const int64_t s = xcd_index(nnz);
  if (s >= nnz)
    return;
  const uint8_t fl = flag ? flag[s] : 0;
  double v;
  if (fl == 1)
    v = 0.;
  else if (fl == 2)
    v = 1.;
  else
    v = ordered_sum<8>(contrib, staging, slot_ptr[s], slot_ptr[s + 1], accumulate ? values[s] : 0.);
  values[s] = v;
}

__global__ void __launch_bounds__(256) k_gather_rhs(int64_t n, const uint32_t* __restrict__ ptr,
                                                    const uint32_t* __restrict__ contrib,
                                                    const double* __restrict__ staging,
                                                    const uint8_t* __restrict__ mask, double* __restrict__ rhs)
{
  const int64_t i = xcd_index(n);
  if (i >= n)
    return;
  double v = 0.;
  if (!(mask && mask[i]))
    v = ordered_sum<8>(contrib, staging, ptr[i], ptr[i + 1]);
  rhs[i] = v;
}

// adjoint element kernel: g_e = -kappa_e lambda_e^T Khat_e u_e with Khat_e recomputed element-locally
template <int GEOM, int ORDER>
__global__ void __launch_bounds__(256) k_adjoint_gradient(int64_t ne, const double* __restrict__ coords,
                                                          const int32_t* __restrict__ conn,
                                                          const int32_t* __restrict__ eid,
                                                          const int32_t* __restrict__ dofs,
                                                          const double* __restrict__ qtab, QRule R,
                                                          const double* __restrict__ kappa, double kappa_value,
                                                          const double* __restrict__ u,
                                                          const double* __restrict__ lambda,
                                                          double* __restrict__ grad)
{
  using T = Traits<GEOM>;
  constexpr int D = T::D, ND = Basis<GEOM, ORDER>::ND;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne)
    return;
  Geo<D> g;
  load_geometry<GEOM>(coords, conn, e, g);
  double ue[ND], le[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    const int32_t d = dofs[e * ND + i];
    ue[i] = u[d], le[i] = lambda[d];
  }
  double acc = 0.;
  if constexpr (T::SIMPLEX && ORDER == 1) {
    double phi[ND], dphi[ND][D], G[ND][D], xc[D];
#pragma unroll
    for (int a = 0; a < D; ++a)
      xc[a] = 1. / (D + 1);
    shape<GEOM>(xc, phi, dphi);
    double gu[D], gl[D];
#pragma unroll
    for (int a = 0; a < D; ++a)
      gu[a] = 0., gl[a] = 0.;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      to_physical<D>(g, dphi[i], G[i]);
#pragma unroll
      for (int a = 0; a < D; ++a)
        gu[a] += ue[i] * G[i][a], gl[a] += le[i] * G[i][a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
      acc += gl[a] * gu[a];
    acc *= g.det * T::VOL;
  } else {
    for (int q = 0; q < R.npts; ++q) {
      double xh[D];
#pragma unroll
      for (int a = 0; a < D; ++a)
        xh[a] = qtab[R.offset * D + q * D + a];
      const double w = qtab[R.offset * D + R.npts * D + q];
      double phi[ND], dphi[ND][D];
      basis<GEOM, ORDER>(xh, phi, dphi);
      double gu[D], gl[D];
#pragma unroll
      for (int a = 0; a < D; ++a)
        gu[a] = 0., gl[a] = 0.;
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        double G[D];
        to_physical<D>(g, dphi[i], G);
#pragma unroll
        for (int a = 0; a < D; ++a)
          gu[a] += ue[i] * G[a], gl[a] += le[i] * G[a];
      }
      double dotv = 0.;
#pragma unroll
      for (int a = 0; a < D; ++a)
        dotv += gl[a] * gu[a];
      acc += dotv * (g.det * w);
    }
  }
  // one part of a mixed grid: its elements' grid indices (coefficient and gradient)
  const int64_t ge = eid ? eid[e] : e;
  const double k = kappa ? kappa[ge] : kappa_value;
  grad[ge] = -k * acc;
}

// BilinearForm(e, e).apply2() for e = u_h - g, element-local part: out[e] = int e^2 (product integrand,
// weight 1) and out[ne + e] = int (kappa grad e) . grad e (Laplace integrand), with g given by its values
// and physical gradients at the points of rule R (operators/bilinear-form.hh:344-452,
// local/bilinear-forms/integrals.hh:97-134)
template <int GEOM, int ORDER>
__global__ void __launch_bounds__(256) k_error_local(int64_t ne, const double* __restrict__ coords,
                                                     const int32_t* __restrict__ conn,
                                                     const int32_t* __restrict__ dofs,
                                                     const double* __restrict__ qtab, QRule R,
                                                     const double* __restrict__ u,
                                                     const double* __restrict__ gval,
                                                     const double* __restrict__ ggrad,
                                                     const double* __restrict__ kappa, double kappa_value,
                                                     double* __restrict__ out, const int32_t* __restrict__ eid,
                                                     const int64_t* __restrict__ qoff, int64_t ne_out)
{
  // eid, qoff (one part of a mixed grid): the grid index of every element, where its points start in
  // gval / ggrad; the element's values go to out[eid], out[ne_out + eid] (NULL: the element's own index)
  using T = Traits<GEOM>;
  constexpr int D = T::D, ND = Basis<GEOM, ORDER>::ND;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne)
    return;
  const int64_t ge = eid ? eid[e] : e;
  Geo<D> g;
  load_geometry<GEOM>(coords, conn, e, g);
  double ue[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i)
    ue[i] = u[dofs[e * ND + i]];
  const double k = kappa ? kappa[ge] : kappa_value;
  double l2 = 0., h1 = 0.;
  for (int q = 0; q < R.npts; ++q) {
    double xh[D];
#pragma unroll
    for (int a = 0; a < D; ++a)
      xh[a] = qtab[R.offset * D + q * D + a];
    const double w = qtab[R.offset * D + R.npts * D + q];
    double phi[ND], dphi[ND][D];
    basis<GEOM, ORDER>(xh, phi, dphi);
    double ev = 0., eg[D];
#pragma unroll
    for (int a = 0; a < D; ++a)
      eg[a] = 0.;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double G[D];
      to_physical<D>(g, dphi[i], G);
      ev += ue[i] * phi[i];
#pragma unroll
      for (int a = 0; a < D; ++a)
        eg[a] += ue[i] * G[a];
    }
    const int64_t qi = (qoff ? qoff[e] : e * R.npts) + q;
    if (gval)
      ev -= gval[qi];
    if (ggrad)
#pragma unroll
      for (int a = 0; a < D; ++a)
        eg[a] -= ggrad[qi * D + a];
    const double f = g.det * w;
    double kd = 0.;
#pragma unroll
    for (int a = 0; a < D; ++a)
      kd += (k * eg[a]) * eg[a];
    l2 += (ev * ev) * f;
    h1 += kd * f;
  }
  out[ge] = l2;
  out[ne_out + ge] = h1;
}

// fixed-order sums of element values over chunks (one block of 256 per chunk and component);
// a chunk never straddles two instances
__global__ void __launch_bounds__(256) k_sum_chunks(const double* __restrict__ vals, int64_t ne,
                                                    const int64_t* __restrict__ chunk0, double* __restrict__ part,
                                                    int64_t nchunks)
{
  __shared__ double sh[256];
  const int64_t c = blockIdx.x;
  const int comp = blockIdx.y;
  const int64_t e0 = chunk0[c], e1 = chunk0[c + 1];
  double s = 0.;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += 256)
    s += vals[comp * ne + e];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    part[comp * nchunks + c] = sh[0];
}

inline unsigned grid_for(int64_t n, int block)
{
  return unsigned((n + block - 1) / block);
}

// quadrature table kept per operator: appended rules (element rules use dim coords per point)
struct QTable {
  std::vector<double> data;
  QRule add(const Quadrature& q)
  {
    const int dim = q.dim;
    // offset in units such that points start at offset*dim and weights follow
    while (data.size() % dim)
      data.push_back(0.);
    QRule r;
    r.offset = int(data.size() / dim);
    r.npts = q.size();
    data.insert(data.end(), q.x.begin(), q.x.end());
    // weights start at offset*dim + npts*dim
    data.insert(data.end(), q.w.begin(), q.w.end());
    return r;
  }
};

struct TableCache {
  DeviceArray<double> elem, line;
};

// (geometry, order) pairs with device kernels: orders 1 and 2 on every geometry
template <class F>
void dispatch_fe(int geom, int order, F&& f)
{
  using I = std::integral_constant<int, 1>;
  using II = std::integral_constant<int, 2>;
  if (order == 1) {
    switch (geom) {
      case HGDT_QUAD:
        return f(std::integral_constant<int, HGDT_QUAD>{}, I{});
      case HGDT_TRI:
        return f(std::integral_constant<int, HGDT_TRI>{}, I{});
      case HGDT_HEX:
        return f(std::integral_constant<int, HGDT_HEX>{}, I{});
      case HGDT_TET:
        return f(std::integral_constant<int, HGDT_TET>{}, I{});
    }
  } else if (order == 2) {
    switch (geom) {
      case HGDT_QUAD:
        return f(std::integral_constant<int, HGDT_QUAD>{}, II{});
      case HGDT_TRI:
        return f(std::integral_constant<int, HGDT_TRI>{}, II{});
      case HGDT_HEX:
        return f(std::integral_constant<int, HGDT_HEX>{}, II{});
      case HGDT_TET:
        return f(std::integral_constant<int, HGDT_TET>{}, II{});
    }
  }
  HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "no device kernels for this geometry and polynomial order");
}

template <class F>
void dispatch_geom(int geom, F&& f)
{
  switch (geom) {
    case HGDT_QUAD:
      f(std::integral_constant<int, HGDT_QUAD>{});
      break;
    case HGDT_TRI:
      f(std::integral_constant<int, HGDT_TRI>{});
      break;
    case HGDT_HEX:
      f(std::integral_constant<int, HGDT_HEX>{});
      break;
    case HGDT_TET:
      f(std::integral_constant<int, HGDT_TET>{});
      break;
    default:
      HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "unknown geometry");
  }
}

// (geometry, order) pairs with intersection kernels: every geometry at order 1, simplices and quadrilaterals
// at order 2 (a Q2 hexahedron's 27 basis functions on both sides of a face do not fit a lane)
template <class F>
void dispatch_faces(int geom, int order, F&& f)
{
  using I = std::integral_constant<int, 1>;
  using II = std::integral_constant<int, 2>;
  if (order == 1) {
    switch (geom) {
      case HGDT_QUAD:
        return f(std::integral_constant<int, HGDT_QUAD>{}, I{});
      case HGDT_TRI:
        return f(std::integral_constant<int, HGDT_TRI>{}, I{});
      case HGDT_HEX:
        return f(std::integral_constant<int, HGDT_HEX>{}, I{});
      case HGDT_TET:
        return f(std::integral_constant<int, HGDT_TET>{}, I{});
    }
  } else if (order == 2) {
    switch (geom) {
      case HGDT_QUAD:
        return f(std::integral_constant<int, HGDT_QUAD>{}, II{});
      case HGDT_TRI:
        return f(std::integral_constant<int, HGDT_TRI>{}, II{});
      case HGDT_TET:
        return f(std::integral_constant<int, HGDT_TET>{}, II{});
    }
  }
  HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "no intersection kernels for this geometry and polynomial order");
}

} // namespace

// (inside type, outside type, order) of the faces of a mixed grid of triangles and quadrilaterals
template <class F>
void dispatch_mixed_faces(int gin, int gout, int order, F&& f)
{
  using Q = std::integral_constant<int, HGDT_QUAD>;
  using T = std::integral_constant<int, HGDT_TRI>;
  auto by_order = [&](auto GI, auto GO) {
    if (order == 1)
      f(GI, GO, std::integral_constant<int, 1>{});
    else if (order == 2)
      f(GI, GO, std::integral_constant<int, 2>{});
    else
      HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "intersection integrals on mixed grids need order 1 or 2");
  };
  if (gin == HGDT_QUAD && gout == HGDT_QUAD)
    by_order(Q{}, Q{});
  else if (gin == HGDT_QUAD && gout == HGDT_TRI)
    by_order(Q{}, T{});
  else if (gin == HGDT_TRI && gout == HGDT_QUAD)
    by_order(T{}, Q{});
  else if (gin == HGDT_TRI && gout == HGDT_TRI)
    by_order(T{}, T{});
  else
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "mixed grids hold triangles and quadrilaterals");
}

// quadrature tables of the appended forms: element rules and face rules (dim - 1 coordinates per point)
// (one ElementParams per part of a mixed grid, each with the rules of its element type)
static void build_tables(const Operator& op, QTable& etab, QTable& ltab, std::vector<ElementParams>& PE,
                         FaceParams& C, FaceParams& B)
{
  const Grid& g = *op.space->grid;
  const int k = op.space->order;
  PE.assign(std::max<size_t>(1, g.parts.size()), ElementParams{});
  for (size_t part = 0; part < PE.size(); ++part) {
    ElementParams& P = PE[part];
    const RefElement& r = g.mixed ? g.parts[part].ref : g.ref;
    P.nforms = int(op.eforms.size());
    for (int f = 0; f < P.nforms; ++f) {
      const auto& F = op.eforms[f];
      P.kind[f] = F.kind, P.field[f] = F.field, P.value[f] = F.value, P.scaling[f] = F.scaling;
      P.points[f] = F.points;
      P.tensor_kind[f] = F.tensor_kind, P.tensor[f] = F.tensor;
      for (int t = 0; t < 9; ++t)
        P.tvalue[f][t] = F.tensor_value[t];
      // integrand order: coefficient (0 when piecewise constant) + test + ansatz (laplace.hh:74-79,
      // product.hh:88-100)
      P.rule[f] = etab.add(element_rule(r, F.coeff_order + k + k + F.over_integrate));
    }
    P.nsrc = int(op.sources.size());
    for (int s = 0; s < P.nsrc; ++s) {
      const auto& S = op.sources[s];
      P.src_kind[s] = S.kind, P.src_data[s] = S.data, P.src_value[s] = S.value;
      // product integrand with the source as ansatz: 0 + test order + f order (conversion.hh:90-93)
      P.src_rule[s] = etab.add(element_rule(r, 0 + k + S.f_order + S.over_integrate));
    }
  }
  auto fill_face = [&](const std::vector<FaceFormDesc>& forms, FaceParams& F) {
    F.nforms = int(forms.size());
    for (int f = 0; f < F.nforms; ++f) {
      const auto& D = forms[f];
      F.kind[f] = D.kind, F.param[f] = D.param, F.kappa[f] = D.kappa, F.kappa_value[f] = D.kappa_value;
      F.omega[f] = D.omega, F.omega_value[f] = D.omega_value, F.scaling[f] = D.scaling, F.mask[f] = D.mask;
      F.kappa_t[f] = D.kappa_t, F.omega_t[f] = D.omega_t;
      // coupling / penalty orders: 0 + 0 + test + ansatz (laplace-ipdg.hh:95-105, ipdg.hh:100-109,245-250)
      F.rule[f] = ltab.add(face_rule(g.ref, 0 + k + k + D.over_integrate));
    }
  };
  fill_face(op.cforms, C);
  fill_face(op.bforms, B);
  for (size_t t = 0; t < op.bsources.size() && t < size_t(MAX_BSOURCES); ++t)
    B.brule[t] = ltab.add(face_rule(g.ref, op.boundary_source_order(op.bsources[t])));
  // the SWIPDG terms are symmetric for symmetry prefactor 1 (laplace-ipdg.hh:107-186,340-368); the
  // penalties always are (ipdg.hh:113-171,254-282)
  bool sym = true;
  for (const auto* forms : {&op.cforms, &op.bforms})
    for (const auto& D : *forms)
      if ((D.kind == FF_INNER_COUPLING || D.kind == FF_DIRICHLET_COUPLING) && D.param != 1.)
        sym = false;
  C.symmetric = B.symmetric = sym ? 1 : 0;
}

void Operator::prepare_assembly()
{
  if (eforms.size() > size_t(MAX_EFORMS) || sources.size() > size_t(MAX_SOURCES) || cforms.size() > size_t(MAX_FFORMS)
      || bforms.size() > size_t(MAX_FFORMS))
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "too many local forms appended");
  if (bsources.size() > size_t(MAX_BSOURCES))
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "too many boundary functionals appended");

  if (tables_ready)
    return;
  QTable etab, ltab;
  std::vector<ElementParams> PE;
  FaceParams C{}, B{};
  build_tables(*this, etab, ltab, PE, C, B);
  d_qtab_elem.upload(etab.data.empty() ? std::vector<double>{0.} : etab.data);
  d_qtab_line.upload(ltab.data.empty() ? std::vector<double>{0.} : ltab.data);
  // mixed grids: the start of every element's points in the element-ordered quadrature-point data
  const Grid& g = *space->grid;
  part_qoff.clear();
  if (g.mixed)
    for (size_t s = 0; s < sources.size(); ++s) {
      const std::vector<int64_t> off =
          sources[s].kind == HGDT_SOURCE_PER_QUADRATURE_POINT
              ? element_point_offsets(g, 0 + space->order + sources[s].f_order + sources[s].over_integrate)
              : std::vector<int64_t>{};
      for (const GridPart& P : g.parts) {
        part_qoff.push_back(std::make_unique<DeviceArray<int64_t>>());
        if (off.empty())
          continue;
        std::vector<int64_t> po(P.eid.size());
        for (size_t el = 0; el < po.size(); ++el)
          po[el] = off[P.eid[el]];
        part_qoff.back()->upload(po);
      }
    }
  part_eqoff.clear();
  if (g.mixed)
    for (const auto& F : eforms) {
      const std::vector<int64_t> off =
          F.points ? element_point_offsets(g, F.coeff_order + 2 * space->order + F.over_integrate)
                   : std::vector<int64_t>{};
      for (const GridPart& P : g.parts) {
        part_eqoff.push_back(std::make_unique<DeviceArray<int64_t>>());
        if (off.empty())
          continue;
        std::vector<int64_t> po(P.eid.size());
        for (size_t el = 0; el < po.size(); ++el)
          po[el] = off[P.eid[el]];
        part_eqoff.back()->upload(po);
      }
    }
  tables_ready = true;
}

void launch_assembly(Operator& op, hipStream_t stream, bool matrix)
{
  const Grid& g = *op.space->grid;
  const ScatterMaps& M = op.maps;
  QTable etab, ltab;
  std::vector<ElementParams> PE;
  FaceParams C{}, B{};
  build_tables(op, etab, ltab, PE, C, B);
  const int eb = op.launch[0].block, fb = op.launch[1].block, gb = op.launch[2].block;
  const bool emat = matrix && !op.eforms.empty();
  if (emat || !op.sources.empty()) {
    // one launch per element type (a mixed grid's parts, each over its own elements and staging)
    for (size_t part = 0; part < PE.size(); ++part) {
      ElementParams& P = PE[part];
      const RefElement& r = g.mixed ? g.parts[part].ref : g.ref;
      const int64_t ne = g.mixed ? g.parts[part].ne() : g.ne;
      if (ne == 0)
        continue;
      P.ne = ne;
      P.coords = g.d_coords.ptr;
      P.conn = g.mixed ? g.parts[part].d_conn.ptr : g.d_conn.ptr;
      P.eid = g.mixed ? g.parts[part].d_eid.ptr : nullptr;
      for (int s = 0; s < P.nsrc && g.mixed; ++s)
        P.src_qoff[s] = op.part_qoff[s * g.parts.size() + part]->ptr;
      for (int f = 0; f < P.nforms && g.mixed; ++f)
        P.ef_qoff[f] = op.part_eqoff[f * g.parts.size() + part]->ptr;
      P.qtab = op.d_qtab_elem.ptr;
      P.staging = op.d_staging.ptr;
      P.elem_base = M.part_elem_base[part];
      P.rhs_base = M.part_rhs_base[part];
      P.write_matrix = emat ? 1 : 0;
      if (r.geom == HGDT_HEX && op.space->order == 2) // 27 local DoFs: one lane per element row
        hipLaunchKernelGGL((k_element_rows<HGDT_HEX, 2>), dim3(grid_for(ne * 27, eb)), dim3(eb), 0, stream, P);
      else
        dispatch_fe(r.geom, op.space->order, [&](auto G, auto O) {
          if constexpr (decltype(G)::value == HGDT_HEX && decltype(O)::value == 2)
            HGDT_THROW(HGDT_ERR_OPERATOR, "Q2 hexahedra are assembled by row lanes");
          else
            hipLaunchKernelGGL((k_element_local<decltype(G)::value, decltype(O)::value>), dim3(grid_for(ne, eb)),
                               dim3(eb), 0, stream, P);
        });
      HGDT_HIP(hipGetLastError());
    }
  }
  const int64_t nin = int64_t(M.inner_in.size()), nbd = int64_t(M.bnd_e.size());
  // faces: one element type, or (mixed grids) one launch per pair of part types over every face of the
  // walk, each lane keeping the faces of its pair; staged blocks with rows of the space's local size
  const int nparts = g.mixed ? int(g.parts.size()) : 1;
  const bool q2hex = !g.mixed && g.ref.geom == HGDT_HEX && op.space->order == 2;
  auto part_geom = [&](int p) { return g.mixed ? g.parts[p].ref.geom : g.ref.geom; };
  for (FaceParams* F : {&C, &B})
    F->etype = g.mixed ? g.d_etype.ptr : nullptr, F->ls = op.space->local_size, F->cstride = g.ref.nverts;
  if (matrix && !op.cforms.empty() && nin > 0) {
    C.nfaces = nin, C.faces = M.d_inner.ptr, C.coords = g.d_coords.ptr, C.conn = g.d_conn.ptr;
    C.qtab = op.d_qtab_line.ptr, C.staging = op.d_staging.ptr, C.base = M.inner_base;
    if (q2hex) // 27 basis functions per side: one lane per (face, row)
      hipLaunchKernelGGL(k_inner_rows_q2hex, dim3(grid_for(nin * 54, fb)), dim3(fb), 0, stream, C);
    else if (!g.mixed)
      dispatch_faces(g.ref.geom, op.space->order, [&](auto G, auto O) {
        hipLaunchKernelGGL((k_inner_local<decltype(G)::value, decltype(O)::value>), dim3(grid_for(nin, fb)), dim3(fb), 0, stream, C);
      });
    else
      for (int pi = 0; pi < nparts; ++pi)
        for (int po = 0; po < nparts; ++po) {
          C.part_in = pi, C.part_out = po;
          dispatch_mixed_faces(part_geom(pi), part_geom(po), op.space->order, [&](auto GI, auto GO, auto O) {
            hipLaunchKernelGGL((k_inner_mixed<decltype(GI)::value, decltype(GO)::value, decltype(O)::value>),
                               dim3(grid_for(nin, fb)), dim3(fb), 0, stream, C);
          });
        }
    HGDT_HIP(hipGetLastError());
  }
  if (matrix && !op.bforms.empty() && nbd > 0) {
    B.nfaces = nbd, B.faces = M.d_bnd.ptr, B.coords = g.d_coords.ptr, B.conn = g.d_conn.ptr;
    B.qtab = op.d_qtab_line.ptr, B.staging = op.d_staging.ptr, B.base = M.bnd_base;
    if (q2hex)
      hipLaunchKernelGGL(k_boundary_rows_q2hex, dim3(grid_for(nbd * 27, fb)), dim3(fb), 0, stream, B);
    else
      for (int pi = 0; pi < nparts; ++pi) {
        B.part_in = pi;
        dispatch_faces(part_geom(pi), op.space->order, [&](auto G, auto O) {
          hipLaunchKernelGGL((k_boundary_local<decltype(G)::value, decltype(O)::value>), dim3(grid_for(nbd, fb)), dim3(fb), 0, stream, B);
        });
      }
    HGDT_HIP(hipGetLastError());
  }
  if (matrix && (!op.eforms.empty() || !op.cforms.empty() || !op.bforms.empty())) {
    hipLaunchKernelGGL(k_gather_values, dim3(xcd_grid(op.nnz, gb)), dim3(gb), 0, stream, op.nnz, M.d_slot_ptr.ptr,
                       M.d_contrib.ptr, op.d_staging.ptr, op.dirichlet ? op.d_slot_flag.ptr : nullptr,
                       op.d_values.ptr, op.borrowed ? 1 : 0);
    HGDT_HIP(hipGetLastError());
  }
  if (!op.bsources.empty() && nbd > 0) {
    BoundaryRhsParams BR{};
    BR.nfaces = nbd, BR.faces = M.d_bnd.ptr, BR.coords = g.d_coords.ptr, BR.conn = g.d_conn.ptr;
    BR.qtab = op.d_qtab_line.ptr, BR.staging = op.d_staging.ptr, BR.base = M.brhs_base;
    BR.nterms = int(op.bsources.size());
    for (int t = 0; t < BR.nterms; ++t) {
      const auto& S = op.bsources[t];
      BR.kind[t] = S.kind, BR.param[t] = S.param, BR.coeff[t] = S.coeff, BR.coeff_value[t] = S.coeff_value;
      BR.coeff_t[t] = S.coeff_t;
      BR.data[t] = S.data, BR.value[t] = S.value, BR.rule[t] = B.brule[t], BR.mask[t] = S.mask;
    }
    BR.etype = g.mixed ? g.d_etype.ptr : nullptr, BR.ls = op.space->local_size, BR.cstride = g.ref.nverts;
    if (q2hex)
      hipLaunchKernelGGL(k_boundary_rhs_rows_q2hex, dim3(grid_for(nbd * 27, fb)), dim3(fb), 0, stream, BR);
    else
      for (int pi = 0; pi < nparts; ++pi) {
        BR.part_in = pi;
        dispatch_faces(part_geom(pi), op.space->order, [&](auto G, auto O) {
          hipLaunchKernelGGL((k_boundary_rhs<decltype(G)::value, decltype(O)::value>), dim3(grid_for(nbd, fb)), dim3(fb), 0, stream, BR);
        });
      }
    HGDT_HIP(hipGetLastError());
  }
  if (!op.sources.empty() || !op.bsources.empty()) {
    hipLaunchKernelGGL(k_gather_rhs, dim3(xcd_grid(op.n, gb)), dim3(gb), 0, stream, op.n, M.d_rhs_ptr.ptr,
                       M.d_rhs_contrib.ptr, op.d_staging.ptr, op.dirichlet ? op.d_dof_mask.ptr : nullptr,
                       op.d_rhs.ptr);
    HGDT_HIP(hipGetLastError());
  }
}

void Operator::adjoint_gradient(int form_index, const double* u, const double* lambda, double* grad,
                                hipStream_t stream)
{
  const Grid& g = *space->grid;
  int seen = -1;
  const ElementFormDesc* F = nullptr;
  for (const auto& f : eforms)
    if (++seen == form_index)
      F = &f;
  if (!F)
    HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "form_index does not name an appended element form");
  if (F->kind != EF_LAPLACE)
    HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "the adjoint gradient is defined for a Laplace form");
  if (F->tensor_kind)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "the adjoint gradient is taken with respect to a scalar diffusion");
  if (F->points)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED,
               "the adjoint gradient is taken with respect to a per-element coefficient, not one given at points");
  prepare_assembly();
  QTable etab, ltab;
  std::vector<ElementParams> PE;
  FaceParams C{}, B{};
  build_tables(*this, etab, ltab, PE, C, B);
  const int block = launch[0].block;
  // one launch per element type (a mixed grid's parts, each over its own elements, DoFs and rule)
  for (size_t part = 0; part < PE.size(); ++part) {
    const int64_t ne = g.mixed ? g.parts[part].ne() : g.ne;
    if (ne == 0)
      continue;
    const int geom = g.mixed ? g.parts[part].ref.geom : g.ref.geom;
    const int32_t* conn = g.mixed ? g.parts[part].d_conn.ptr : g.d_conn.ptr;
    const int32_t* eid = g.mixed ? g.parts[part].d_eid.ptr : nullptr;
    const int32_t* dofs = g.mixed ? space->parts[part].d_dofs.ptr : space->d_dofs.ptr;
    const QRule R = PE[part].rule[form_index];
    dispatch_fe(geom, space->order, [&](auto G, auto O) {
      hipLaunchKernelGGL((k_adjoint_gradient<decltype(G)::value, decltype(O)::value>), dim3(grid_for(ne, block)),
                         dim3(block), 0, stream, ne, g.d_coords.ptr, conn, eid, dofs, d_qtab_elem.ptr, R, F->field,
                         F->value, u, lambda, grad);
    });
  }
  HGDT_HIP(hipGetLastError());
}

void error_norms(const Space& s, const double* dofs, int order, const double* gval, const double* ggrad,
                 const double* kappa, double kappa_value, double* host_out, hipStream_t stream)
{
  const Grid& g = *s.grid;
  if (s.kind != HGDT_CONTINUOUS_LAGRANGE && s.kind != HGDT_DISCONTINUOUS_LAGRANGE)
    HGDT_THROW(HGDT_ERR_SPACE, "unknown space kind");
  if (order < 0 || order > 20)
    HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "quadrature order must be in [0, 20]");
  QTable tab;
  const QRule R = tab.add(element_rule(g.mixed ? g.parts[0].ref : g.ref, order));
  std::vector<QRule> part_rule; // mixed grids: the rule of every part
  for (const GridPart& P : g.parts)
    part_rule.push_back(tab.add(element_rule(P.ref, order)));
  DeviceArray<double> d_tab, d_vals, d_part;
  DeviceArray<int64_t> d_chunk;
  d_tab.upload(tab.data);
  // chunks of at most 65536 elements inside each instance
  constexpr int64_t CH = 65536;
  std::vector<int64_t> chunk0; // start of every chunk, closed by ne
  for (int i = 0; i < g.ninst(); ++i)
    for (int64_t e = g.inst_elem[i]; e < g.inst_elem[i + 1]; e += CH)
      chunk0.push_back(e);
  const int64_t nchunks = int64_t(chunk0.size());
  chunk0.push_back(g.ne);
  d_chunk.upload(chunk0);
  d_vals.allocate(std::max<int64_t>(2 * g.ne, 1));
  d_part.allocate(std::max<int64_t>(2 * nchunks, 1));
  if (g.ne > 0 && g.mixed) {
    // one launch per element type, every element's values at its grid index
    const std::vector<int64_t> off = element_point_offsets(g, order);
    std::vector<DeviceArray<int64_t>> d_qoff(g.parts.size());
    for (size_t p = 0; p < g.parts.size(); ++p) {
      const GridPart& P = g.parts[p];
      std::vector<int64_t> po(P.eid.size());
      for (size_t el = 0; el < po.size(); ++el)
        po[el] = off[P.eid[el]];
      d_qoff[p].upload(po);
      if (P.ne() == 0)
        continue;
      dispatch_fe(P.ref.geom, s.order, [&](auto G, auto O) {
        hipLaunchKernelGGL((k_error_local<decltype(G)::value, decltype(O)::value>), dim3(grid_for(P.ne(), 256)),
                           dim3(256), 0, stream, P.ne(), g.d_coords.ptr, P.d_conn.ptr, s.parts[p].d_dofs.ptr,
                           d_tab.ptr, part_rule[p], dofs, gval, ggrad, kappa, kappa_value, d_vals.ptr, P.d_eid.ptr,
                           d_qoff[p].ptr, g.ne);
      });
      HGDT_HIP(hipGetLastError());
    }
    // the offsets are freed at return, after the stream synchronization below
    hipLaunchKernelGGL(k_sum_chunks, dim3(unsigned(nchunks), 2), dim3(256), 0, stream, d_vals.ptr, g.ne, d_chunk.ptr,
                       d_part.ptr, nchunks);
    HGDT_HIP(hipGetLastError());
    std::vector<double> part(2 * nchunks, 0.);
    HGDT_HIP(hipMemcpyAsync(part.data(), d_part.ptr, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    HGDT_HIP(hipStreamSynchronize(stream));
    for (int comp = 0; comp < 2; ++comp) {
      double acc = 0.;
      for (int64_t c = 0; c < nchunks; ++c)
        acc += part[comp * nchunks + c];
      host_out[comp] = acc;
    }
    return;
  }
  if (g.ne > 0) {
    dispatch_fe(g.ref.geom, s.order, [&](auto G, auto O) {
      hipLaunchKernelGGL((k_error_local<decltype(G)::value, decltype(O)::value>), dim3(grid_for(g.ne, 256)), dim3(256),
                         0, stream, g.ne,
                         g.d_coords.ptr, g.d_conn.ptr, s.d_dofs.ptr, d_tab.ptr, R, dofs, gval, ggrad, kappa,
                         kappa_value, d_vals.ptr, nullptr, nullptr, g.ne);
    });
    HGDT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_chunks, dim3(unsigned(nchunks), 2), dim3(256), 0, stream, d_vals.ptr, g.ne, d_chunk.ptr,
                       d_part.ptr, nchunks);
    HGDT_HIP(hipGetLastError());
  }
  std::vector<double> part(2 * nchunks, 0.);
  if (g.ne > 0) {
    HGDT_HIP(hipMemcpyAsync(part.data(), d_part.ptr, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    HGDT_HIP(hipStreamSynchronize(stream));
  }
  // per instance, chunks summed in order on the host (a handful of values)
  for (int i = 0; i < g.ninst(); ++i)
    for (int comp = 0; comp < 2; ++comp) {
      double acc = 0.;
      for (int64_t c = 0; c < nchunks; ++c)
        if (chunk0[c] >= g.inst_elem[i] && chunk0[c] < g.inst_elem[i + 1])
          acc += part[comp * nchunks + c];
      host_out[2 * i + comp] = acc;
    }
}

// ---------------------------------------------------------------------------------------------------------
// error_norms without allocations: the same kernels over a fixed device array `scratch` (one copy per GPU,
// held until the process ends).  Its regions: the quadrature table, the chunk table, the chunk partials and
// the element values of one window.  A window is a run of whole chunks (of NA_CHUNK elements, never
// straddling an instance) holding at most NA_WINDOW elements; k_error_local runs on the window's elements
// (pointer offsets) and k_sum_chunks sums each of its chunks exactly as error_norms does, so the results
// are the same bits.  Calls that overlap on one GPU (two host threads, or two streams) race on the array.
// ---------------------------------------------------------------------------------------------------------
namespace {
constexpr int64_t NA_CHUNK = 65536;             // elements per chunk (as error_norms)
constexpr int64_t NA_WINDOW = 16 * NA_CHUNK;     // elements per window: 2 values each
constexpr int64_t NA_TABLE = 65536;              // doubles of quadrature table
constexpr int64_t NA_CHUNKS = 8192;              // chunk-table entries: chunks + one end mark per window
constexpr int64_t NA_VALUES_OFF = 0;                                   // bytes
constexpr int64_t NA_TABLE_OFF = NA_VALUES_OFF + 8 * 2 * NA_WINDOW;
constexpr int64_t NA_CHUNKS_OFF = NA_TABLE_OFF + 8 * NA_TABLE;
constexpr int64_t NA_PART_OFF = NA_CHUNKS_OFF + 8 * NA_CHUNKS;
constexpr int64_t NA_BYTES = NA_PART_OFF + 8 * 2 * NA_CHUNKS;         // 16.75 MiB

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[NA_MAX_DEVICES] = {};

// device address of `scratch` on the current device, looked up once per device
unsigned char* scratch_on_current_device()
{
  int dev = -1;
  HGDT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEVICES)
    HGDT_THROW(HGDT_ERR_DEVICE, "device index beyond the scratch table of the no-allocation entry points");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HGDT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[dev];
}
} // namespace

int64_t error_norms_scratch_bytes()
{
  return NA_BYTES;
}

void error_norms_noalloc(const Space& s, const double* dofs, int order, const double* gval, const double* ggrad,
                         const double* kappa, double kappa_value, double* host_out, hipStream_t stream)
{
  const Grid& g = *s.grid;
  if (s.kind != HGDT_CONTINUOUS_LAGRANGE && s.kind != HGDT_DISCONTINUOUS_LAGRANGE)
    HGDT_THROW(HGDT_ERR_SPACE, "unknown space kind");
  if (order < 0 || order > 20)
    HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "quadrature order must be in [0, 20]");
  if (g.mixed)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "no-allocation error norms need a grid of one element type");
  QTable tab;
  const QRule R = tab.add(element_rule(g.ref, order));
  if (int64_t(tab.data.size()) > NA_TABLE)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "no-allocation error norms: the quadrature rule of this order needs " +
                                           std::to_string(tab.data.size()) + " table entries, more than the " +
                                           std::to_string(NA_TABLE) + " of the fixed scratch array");
  // chunks (as error_norms), then windows of whole chunks; the device chunk table holds every window's
  // starts relative to the window followed by the window's end
  std::vector<int64_t> chunk0;
  for (int i = 0; i < g.ninst(); ++i)
    for (int64_t e = g.inst_elem[i]; e < g.inst_elem[i + 1]; e += NA_CHUNK)
      chunk0.push_back(e);
  const int64_t nchunks = int64_t(chunk0.size());
  chunk0.push_back(g.ne);
  std::vector<int64_t> win0{0}; // first chunk of every window, closed by nchunks
  for (int64_t c = 0; c < nchunks; ++c)
    if (chunk0[c + 1] - chunk0[win0.back()] > NA_WINDOW)
      win0.push_back(c);
  win0.push_back(nchunks);
  const int64_t nwin = int64_t(win0.size()) - 1;
  if (nchunks + nwin > NA_CHUNKS)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "no-allocation error norms: the grid needs " + std::to_string(nchunks) +
                                           " chunks of at most 65536 elements in " + std::to_string(nwin) +
                                           " windows, more than the fixed scratch array's " +
                                           std::to_string(NA_CHUNKS) + " chunk-table entries");
  std::vector<int64_t> table;
  table.reserve(size_t(nchunks + nwin));
  for (int64_t w = 0; w < nwin; ++w)
    for (int64_t c = win0[w]; c <= win0[w + 1]; ++c)
      table.push_back(chunk0[c] - chunk0[win0[w]]);
  unsigned char* base = scratch_on_current_device();
  double* d_vals = reinterpret_cast<double*>(base + NA_VALUES_OFF);
  double* d_tab = reinterpret_cast<double*>(base + NA_TABLE_OFF);
  int64_t* d_chunk = reinterpret_cast<int64_t*>(base + NA_CHUNKS_OFF);
  double* d_part = reinterpret_cast<double*>(base + NA_PART_OFF);
  std::vector<double> part(2 * nchunks, 0.);
  if (g.ne > 0) {
    HGDT_HIP(hipMemcpyAsync(d_tab, tab.data.data(), tab.data.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    HGDT_HIP(hipMemcpyAsync(d_chunk, table.data(), table.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream));
    const int nv = g.ref.nverts, nd = s.local_size, D = g.ref.dim;
    for (int64_t w = 0; w < nwin; ++w) {
      const int64_t e0 = chunk0[win0[w]], e1 = chunk0[win0[w + 1]], ne = e1 - e0;
      const int64_t nc = win0[w + 1] - win0[w];
      dispatch_fe(g.ref.geom, s.order, [&](auto G, auto O) {
        hipLaunchKernelGGL((k_error_local<decltype(G)::value, decltype(O)::value>), dim3(grid_for(ne, 256)),
                           dim3(256), 0, stream, ne, g.d_coords.ptr, g.d_conn.ptr + e0 * nv,
                           s.d_dofs.ptr + e0 * nd, d_tab, R, dofs, gval ? gval + e0 * R.npts : nullptr,
                           ggrad ? ggrad + e0 * R.npts * D : nullptr, kappa ? kappa + e0 : nullptr, kappa_value,
                           d_vals, nullptr, nullptr, ne);
      });
      HGDT_HIP(hipGetLastError());
      // window w's chunk starts sit at table offset win0[w] + w; its partials at 2 win0[w] as [comp][chunk]
      hipLaunchKernelGGL(k_sum_chunks, dim3(unsigned(nc), 2), dim3(256), 0, stream, d_vals, ne,
                         d_chunk + win0[w] + w, d_part + 2 * win0[w], nc);
      HGDT_HIP(hipGetLastError());
    }
    HGDT_HIP(hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    HGDT_HIP(hipStreamSynchronize(stream));
  }
  for (int i = 0; i < g.ninst(); ++i)
    for (int comp = 0; comp < 2; ++comp) {
      double acc = 0.;
      for (int64_t w = 0; w < nwin; ++w)
        for (int64_t c = win0[w]; c < win0[w + 1]; ++c)
          if (chunk0[c] >= g.inst_elem[i] && chunk0[c] < g.inst_elem[i + 1])
            acc += part[2 * win0[w] + comp * (win0[w + 1] - win0[w]) + (c - win0[w])];
      host_out[2 * i + comp] = acc;
    }
}

namespace {
__global__ void k_tensor_symmetric(int64_t ne, int d, const double* __restrict__ t, int* __restrict__ flag)
{
  const int64_t e = xcd_index(ne);
  if (e >= ne)
    return;
  bool sym = true;
  for (int a = 0; a < d; ++a)
    for (int b = a + 1; b < d; ++b)
      sym = sym && t[e * d * d + a * d + b] == t[e * d * d + b * d + a];
  if (!sym)
    *flag = 1; // every writer stores the same value
}
} // namespace

bool tensor_field_symmetric(const double* tensor, int64_t ne, int d)
{
  DeviceArray<int> flag;
  flag.allocate(1);
  int h = 0;
  HGDT_HIP(hipMemcpy(flag.ptr, &h, sizeof(int), hipMemcpyHostToDevice));
  if (ne > 0)
    hipLaunchKernelGGL(k_tensor_symmetric, dim3(xcd_grid(ne, 256)), dim3(256), 0, nullptr, ne, d, tensor, flag.ptr);
  HGDT_HIP(hipGetLastError());
  HGDT_HIP(hipMemcpy(&h, flag.ptr, sizeof(int), hipMemcpyDeviceToHost));
  return h == 0;
}

} // namespace hgdt
