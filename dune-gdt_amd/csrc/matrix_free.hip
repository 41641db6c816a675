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
  const int64_t* lat;  // per instance: cells n0 n1 n2, first element, first DoF
  const double* ref;   // per instance (ref_stride doubles): h0 h1 h2 0, element, face, boundary matrices
  int64_t ref_stride;
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
  // row blocks (the solver's operator partition; discontinuous spaces: multiples of 4 rows)
  int nb;
  const int64_t* row0;
  const int32_t* binst;
  MfVectors v;
};

__device__ __forceinline__ double coef(const double* f, double value, int64_t e)
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

// continuous Q1 on quadrilaterals (D = 2) or hexahedra (D = 3): one lane per DoF
template <int D, int MODE>
__global__ void __launch_bounds__(256) k_mf_cg(MfParams P)
{
  constexpr int NV = 1 << D, NS = (D == 2) ? 9 : 27;
  constexpr bool DIAG = (MODE == MF_DIAG);
  const int b = logical_block(P.nb);
  if (b < 0)
    return;
  const int inst = P.binst[b];
  if (MODE == MF_PCG && P.v.S[(P.v.par ^ 1) * P.ninst + inst].done)
    return;
  const int64_t* __restrict__ L = P.lat + 5 * int64_t(inst);
  const int64_t e0 = L[3], d0 = L[4];
  const uint32_t nx = uint32_t(L[0]), ny = uint32_t(L[1]), nz = uint32_t(L[2]);
  const uint32_t vx = nx + 1, vy = ny + 1;
  const int64_t sy = vx, sz = int64_t(vx) * vy;
  const double* __restrict__ R = P.ref + int64_t(inst) * P.ref_stride + P.eoff;
  const int64_t r0 = P.row0[b], r1 = P.row0[b + 1];
  double part = 0.;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
    const uint32_t v = uint32_t(i - d0);
    const uint32_t ix = v % vx, iy = (v / vx) % vy, iz = (D == 3) ? v / (vx * vy) : 0u;
    auto boundary = [&](uint32_t x, uint32_t y, uint32_t z) {
      return x == 0u || x == nx || y == 0u || y == ny || (D == 3 && (z == 0u || z == nz));
    };
    double y = 0.;
    if (P.dirichlet && boundary(ix, iy, iz)) {
      // unit row of a constrained DoF
      y = DIAG ? 1. : src<MODE>(P, i);
    } else {
      // source on the 3^D neighbourhood; zero outside the instance and on constrained DoFs (zero columns)
      double s[NS];
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        s[t] = 0.;
        if constexpr (!DIAG) {
          const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = (D == 3) ? t / 9 - 1 : 0;
          const uint32_t jx = ix + dx, jy = iy + dy, jz = iz + dz; // wraps to > n below zero
          const bool in = jx <= nx && jy <= ny && (D == 2 || jz <= nz);
          if (in && !(P.dirichlet && boundary(jx, jy, jz)))
            s[t] = src<MODE>(P, i + dx + dy * sy + dz * sz);
        }
      }
      // the cells around the vertex: cell (ix - 1 + cx, ...) holds it as local vertex l
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const uint32_t ex = ix - 1 + cx, ey = iy - 1 + cy, ez = iz - 1 + cz;
        if (ex >= nx || ey >= ny || (D == 3 && ez >= nz))
          continue;
        const int l = (1 - cx) | ((1 - cy) << 1) | ((D == 3 ? (1 - cz) : 0) << 2);
        const int64_t e = e0 + (int64_t(D == 3 ? ez : 0u) * ny + ey) * nx + ex;
        for (int f = 0; f < P.nef; ++f) {
          const double sf = P.escale[f] * coef(P.ecoef[f], P.evalue[f], e);
          const double* __restrict__ row = R + f * NV * NV + l * NV;
          if constexpr (DIAG) {
            y = fma(sf, row[l], y);
          } else {
            double t = 0.;
#pragma unroll
            for (int j = 0; j < NV; ++j) {
              const int tx = cx + (j & 1), ty = cy + ((j >> 1) & 1), tz = (D == 3) ? cz + (j >> 2) : 1;
              t = fma(row[j], s[tx + 3 * ty + 9 * (D == 3 ? tz : 0)], t);
            }
            y = fma(sf, t, y);
          }
        }
      }
    }
    store_row<MODE>(P, i, y, part);
  }
  finish_block<MODE>(P, b, part);
}

// discontinuous Q1 on quadrilaterals: one lane per element (4 consecutive DoFs)
template <int MODE>
__global__ void __launch_bounds__(256) k_mf_dg(MfParams P)
{
  constexpr bool DIAG = (MODE == MF_DIAG);
  const int b = logical_block(P.nb);
  if (b < 0)
    return;
  const int inst = P.binst[b];
  if (MODE == MF_PCG && P.v.S[(P.v.par ^ 1) * P.ninst + inst].done)
    return;
  const int64_t* __restrict__ L = P.lat + 5 * int64_t(inst);
  const int64_t nx = L[0], ny = L[1], e0 = L[3], d0 = L[4];
  const double* __restrict__ R = P.ref + int64_t(inst) * P.ref_stride;
  const double len[2] = {R[1], R[0]}; // faces normal to x have length h1, faces normal to y h0
  const int64_t r0 = P.row0[b], r1 = P.row0[b + 1];
  double part = 0.;
  for (int64_t i0 = r0 + 4 * int64_t(threadIdx.x); i0 < r1; i0 += 4 * int64_t(blockDim.x)) {
    const int64_t c = (i0 - d0) >> 2, e = e0 + c;
    const int64_t ci = c % nx, cj = c / nx;
    double a[4], y[4] = {0., 0., 0., 0.};
#pragma unroll
    for (int k = 0; k < 4; ++k)
      a[k] = DIAG ? 0. : src<MODE>(P, i0 + k);
    for (int f = 0; f < P.nef; ++f)
      mv<4, false, DIAG>(R + P.eoff + 16 * f, P.escale[f] * coef(P.ecoef[f], P.evalue[f], e), a, y);
    // sides: 0 x-, 1 x+, 2 y-, 3 y+; the neighbour across x+ / y+ has the larger index (this element is
    // the inside of that intersection), across x- / y- the smaller (this element is the outside)
    for (int side = 0; side < 4; ++side) {
      const int o = side >> 1;
      const bool inside = side & 1;
      const bool inner = (side == 0) ? ci > 0 : (side == 1) ? ci < nx - 1 : (side == 2) ? cj > 0 : cj < ny - 1;
      if (inner) {
        const int64_t step = (o == 0) ? 1 : nx;
        const int64_t nb = inside ? e + step : e - step;
        const int64_t ein = inside ? e : nb, eout = inside ? nb : e;
        double bv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          bv[k] = DIAG ? 0. : src<MODE>(P, (inside ? i0 + 4 * step : i0 - 4 * step) + k);
        for (int f = 0; f < P.ncf; ++f) {
          const double* __restrict__ G = R + P.coff + (2 * f + o) * 64; // G_ii G_io G_oi G_oo or M_ii M_io - M_oo
          const double sc = P.fscale[0][f], prm = P.fparam[0][f];
          const double dm = coef(P.fomega[0][f], P.fov[0][f], ein), dp = coef(P.fomega[0][f], P.fov[0][f], eout);
          if (P.fkind[0][f] == FF_INNER_COUPLING) {
            const double kin = coef(P.fkappa[0][f], P.fkv[0][f], ein);
            const double kout = coef(P.fkappa[0][f], P.fkv[0][f], eout);
            const double wm = dp / (dp + dm), wp = dm / (dp + dm);
            if (inside) {
              const double s0 = -sc * wm * kin;
              mv<4, false, DIAG>(G, s0, a, y);
              mv<4, true, DIAG>(G, s0 * prm, a, y);
              if constexpr (!DIAG) {
                mv<4, false, false>(G + 16, -sc * wp * kout, bv, y);
                mv<4, true, false>(G + 32, sc * prm * wm * kin, bv, y);
              }
            } else {
              const double t0 = sc * wp * kout;
              mv<4, false, DIAG>(G + 48, t0, a, y);
              mv<4, true, DIAG>(G + 48, t0 * prm, a, y);
              if constexpr (!DIAG) {
                mv<4, false, false>(G + 32, sc * wm * kin, bv, y);
                mv<4, true, false>(G + 16, -sc * prm * wp * kout, bv, y);
              }
            }
          } else {
            const double pen = sc * ((prm * ((dp * dm) / (dp + dm))) / len[o]);
            mv<4, false, DIAG>(G + (inside ? 0 : 48), pen, a, y);
            if constexpr (!DIAG) {
              if (inside)
                mv<4, false, false>(G + 16, -pen, bv, y);
              else
                mv<4, true, false>(G + 16, -pen, bv, y);
            }
          }
        }
      } else {
        for (int f = 0; f < P.nbf; ++f) {
          const double* __restrict__ B = R + P.boff + (4 * f + side) * 16;
          const double sc = P.fscale[1][f], prm = P.fparam[1][f];
          if (P.fkind[1][f] == FF_DIRICHLET_COUPLING) {
            const double s0 = -sc * coef(P.fkappa[1][f], P.fkv[1][f], e);
            mv<4, false, DIAG>(B, s0, a, y);
            mv<4, true, DIAG>(B, s0 * prm, a, y);
          } else {
            mv<4, false, DIAG>(B, sc * ((prm * coef(P.fomega[1][f], P.fov[1][f], e)) / len[o]), a, y);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      store_row<MODE>(P, i0 + k, y[k], part);
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
  const int64_t eoff = 4, coff = eoff + int64_t(nef) * NV * NV, boff = coff + int64_t(ncf) * 2 * 64;
  const int64_t stride = boff + int64_t(nbf) * 4 * 16;
  std::vector<double> ref(size_t(stride) * ninst, 0.);
  std::vector<int64_t> lat(5 * size_t(ninst));
  for (int in = 0; in < ninst; ++in) {
    const Lattice& L = g.lattice[in];
    double* r = &ref[size_t(in) * stride];
    const double h[3] = {L.h[0], L.h[1], D == 3 ? L.h[2] : 1.};
    r[0] = h[0], r[1] = h[1], r[2] = h[2];
    lat[5 * in + 0] = L.n[0], lat[5 * in + 1] = L.n[1], lat[5 * in + 2] = (D == 3) ? L.n[2] : 1;
    lat[5 * in + 3] = g.inst_elem[in], lat[5 * in + 4] = s.inst_dof[in];
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
        double* G = r + coff + int64_t(2 * f + o) * 64;
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
                G[0 + i * 4 + j] += w * (gin_j * pin[i]);
                G[16 + i * 4 + j] += w * (gout_j * pin[i]);
                G[32 + i * 4 + j] += w * (gin_j * pout[i]);
                G[48 + i * 4 + j] += w * (gout_j * pout[i]);
              } else {
                G[0 + i * 4 + j] += w * (pin[j] * pin[i]);
                G[16 + i * 4 + j] += w * (pout[j] * pin[i]);
                G[48 + i * 4 + j] += w * (pout[j] * pout[i]);
              }
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
        double* B = r + boff + int64_t(4 * f + side) * 16;
        const double flen = h[1 - o];
        for (int p = 0; p < q.size(); ++p) {
          double x[2];
          x[o] = (side & 1) ? 1. : 0., x[1 - o] = q.x[p];
          double phi[8], dphi[8][3];
          q1_shape(2, x, phi, dphi);
          const double w = flen * q.w[p];
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
              B[i * 4 + j] += (Bd.kind == FF_DIRICHLET_COUPLING) ? w * ((sign * dphi[j][o] / h[o]) * phi[i])
                                                                : w * (phi[j] * phi[i]);
        }
      }
    }
  }
  F.mf_kind = dg ? MFK_DG_QUAD : (D == 2 ? MFK_CG_QUAD : MFK_CG_HEX);
  F.mf_ref_stride = stride;
  F.d_mf_ref.upload(ref);
  F.d_mf_lat.upload(lat);
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

int64_t mf_bytes(const Operator& op, bool scaled)
{
  return (scaled ? 24 : 16) * op.n + int64_t(8) * op.fmt.mf_fields * op.space->grid->ne;
}

void mf_launch(const Operator& op, int mode, const MfVectors& v, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1)
{
  const SolverFormat& F = op.fmt;
  MfParams P{};
  P.ninst = int(op.space->inst_dof.size()) - 1;
  P.lat = F.d_mf_lat.ptr;
  P.ref = F.d_mf_ref.ptr;
  P.ref_stride = F.mf_ref_stride;
  const int D = op.space->grid->ref.dim, NV = 1 << D;
  P.eoff = 4;
  P.coff = P.eoff + int64_t(op.eforms.size()) * NV * NV;
  P.boff = P.coff + int64_t(op.cforms.size()) * 2 * 64;
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
  P.row0 = F.A.d_row0.ptr;
  P.binst = F.A.d_inst.ptr;
  P.v = v;
  const unsigned grid = hw_grid(F.A.nblocks);
  const int block = std::min(F.A.block, 256);
  const void* k = nullptr;
  auto pick = [&](auto k0, auto k1, auto k2) {
    k = mode == MF_APPLY ? reinterpret_cast<const void*>(k0)
                         : mode == MF_PCG ? reinterpret_cast<const void*>(k1) : reinterpret_cast<const void*>(k2);
  };
  if (F.mf_kind == MFK_CG_QUAD)
    pick(k_mf_cg<2, MF_APPLY>, k_mf_cg<2, MF_PCG>, k_mf_cg<2, MF_DIAG>);
  else if (F.mf_kind == MFK_CG_HEX)
    pick(k_mf_cg<3, MF_APPLY>, k_mf_cg<3, MF_PCG>, k_mf_cg<3, MF_DIAG>);
  else if (F.mf_kind == MFK_DG_QUAD)
    pick(k_mf_dg<MF_APPLY>, k_mf_dg<MF_PCG>, k_mf_dg<MF_DIAG>);
  else
    HGDT_THROW(HGDT_ERR_OPERATOR, "matrix-free data not prepared");
  void* args[] = {&P};
  if (ev0)
    HGDT_HIP(hipExtLaunchKernel(k, dim3(grid), dim3(block), args, 0, stream, ev0, ev1, 0));
  else
    HGDT_HIP(hipLaunchKernel(k, dim3(grid), dim3(block), args, 0, stream));
}

} // namespace hgdt
