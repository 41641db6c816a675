// oracle.cpp -- CPU restatement of dune-gdt's global assembly and linear-solve path.
//
// TEST INFRASTRUCTURE ONLY: this is the checker.  It is loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
// product library.  It follows the reference's loop nests one to one:
//   * per element: bind basis, quadrature loop accumulating a dense local
//     matrix (dune/gdt/local/bilinear-forms/integrals.hh:97-134), then
//     add_to_entry of scaling * local into the sorted CSR row
//     (dune/gdt/local/assembler/bilinear-form-assemblers.hh:110-128);
//   * per inner intersection visited once / per boundary intersection: face
//     quadrature loops (integrals.hh:197-267, 338-369) and the four-block /
//     one-block scatters (bilinear-form-assemblers.hh:238-278, 380-396);
//   * right-hand side functional (dune/gdt/local/functionals/integrals.hh:72-97,
//     dune/gdt/local/assembler/functional-assemblers.hh:77-86);
//   * Dirichlet constraints (dune/gdt/tools/dirichlet-constraints.hh:85-184);
//   * sparsity patterns (dune/gdt/tools/sparsity-pattern.hh:33-144);
//   * Jacobi-preconditioned CG for apply_inverse
//     (dune/gdt/operators/matrix-based.hh:148-159; the ISTL solver behind it is
//     not vendored, the PCG recursion below is the textbook one).
// Reference paths are relative to the dune-gdt source tree.
//
// Build-defined conventions (DUNE's come from un-vendored dune-geometry/grid):
// reference cube vertex b <-> bits of b (x = bit 0, y = bit 1, z = bit 2);
// reference simplex vertex 0 = origin, vertex k = e_k; structured vertices and
// elements lexicographic with x fastest; Kuhn split of each cube into d!
// simplices ordered by the lexicographic order of the axis permutations.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_error;

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// reference elements
// ---------------------------------------------------------------------------
struct RefElement {
  int geom = 0, dim = 0, nverts = 0, nfaces = 0;
  bool simplex = false;
  std::vector<std::array<double, 3>> verts;
  std::vector<std::vector<int>> faces; // local vertex ids of each face, DUNE face numbering
};

RefElement make_ref(int geom)
{
  RefElement r;
  r.geom = geom;
  switch (geom) {
    case OR_QUAD:
      r.dim = 2, r.nverts = 4, r.nfaces = 4;
      r.verts = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}};
      r.faces = {{0, 2}, {1, 3}, {0, 1}, {2, 3}};
      break;
    case OR_TRI:
      r.dim = 2, r.nverts = 3, r.nfaces = 3, r.simplex = true;
      r.verts = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
      r.faces = {{0, 1}, {0, 2}, {1, 2}};
      break;
    case OR_HEX:
      r.dim = 3, r.nverts = 8, r.nfaces = 6;
      for (int b = 0; b < 8; ++b)
        r.verts.push_back({double(b & 1), double((b >> 1) & 1), double((b >> 2) & 1)});
      r.faces = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
      break;
    case OR_TET:
      r.dim = 3, r.nverts = 4, r.nfaces = 4, r.simplex = true;
      r.verts = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      r.faces = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
      break;
    default:
      throw Error("unknown geometry type");
  }
  return r;
}

// ---------------------------------------------------------------------------
// quadrature (dune-geometry QuadratureRules<D,d>::rule(type, order) is not
// vendored; these rules integrate polynomials of the requested order exactly)
// ---------------------------------------------------------------------------
struct Quadrature {
  int dim = 0;
  std::vector<std::array<double, 3>> x;
  std::vector<double> w;
};

// Gauss-Legendre with n points on [0,1], ascending
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w)
{
  x.assign(n, 0.), w.assign(n, 0.);
  for (int i = 0; i < n; ++i) {
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1., p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2. * k - 1.) * t * p1 - (k - 1.) * p0) / k;
        p0 = p1, p1 = p2;
      }
      const double pn = (n == 1) ? t : p1, pnm1 = (n == 1) ? 1. : p0;
      const double dp = n * (t * pn - pnm1) / (t * t - 1.);
      const double dt = pn / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-17)
        break;
    }
    double p0 = 1., p1 = t;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2. * k - 1.) * t * p1 - (k - 1.) * p0) / k;
      p0 = p1, p1 = p2;
    }
    const double pn = (n == 1) ? t : p1, pnm1 = (n == 1) ? 1. : p0;
    const double dp = n * (t * pn - pnm1) / (t * t - 1.);
    x[i] = 0.5 * (1. - t);
    w[i] = 1. / ((1. - t * t) * dp * dp);
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i)
    idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return x[a] < x[b]; });
  std::vector<double> xs(n), ws(n);
  for (int i = 0; i < n; ++i)
    xs[i] = x[idx[i]], ws[i] = w[idx[i]];
  x = xs, w = ws;
}

Quadrature line_rule(int order)
{
  Quadrature q;
  q.dim = 1;
  std::vector<double> x, w;
  gauss_legendre01(order / 2 + 1, x, w);
  for (size_t i = 0; i < x.size(); ++i)
    q.x.push_back({x[i], 0, 0}), q.w.push_back(w[i]);
  return q;
}

Quadrature cube_rule(int dim, int order)
{
  Quadrature q;
  q.dim = dim;
  std::vector<double> x, w;
  gauss_legendre01(order / 2 + 1, x, w);
  const int n = int(x.size());
  const int total = (dim == 2) ? n * n : n * n * n;
  for (int t = 0; t < total; ++t) {
    const int i = t % n, j = (t / n) % n, k = t / (n * n);
    if (dim == 2)
      q.x.push_back({x[i], x[j], 0}), q.w.push_back(w[i] * w[j]);
    else
      q.x.push_back({x[i], x[j], x[k]}), q.w.push_back(w[i] * w[j] * w[k]);
  }
  return q;
}

Quadrature simplex_rule(int dim, int order)
{
  Quadrature q;
  q.dim = dim;
  if (order <= 1) {
    if (dim == 2)
      q.x.push_back({1. / 3., 1. / 3., 0}), q.w.push_back(0.5);
    else
      q.x.push_back({0.25, 0.25, 0.25}), q.w.push_back(1. / 6.);
    return q;
  }
  if (order == 2) {
    if (dim == 2) {
      const double a = 1. / 6., b = 2. / 3.;
      q.x = {{a, a, 0}, {b, a, 0}, {a, b, 0}};
      q.w = {1. / 6., 1. / 6., 1. / 6.};
    } else {
      const double a = (5. + 3. * std::sqrt(5.)) / 20., b = (5. - std::sqrt(5.)) / 20.;
      q.x = {{b, b, b}, {a, b, b}, {b, a, b}, {b, b, a}};
      q.w = {1. / 24., 1. / 24., 1. / 24., 1. / 24.};
    }
    return q;
  }
  // conical product (collapsed Gauss-Legendre) rule, exact to the given order
  std::vector<double> xu, wu, xv, wv, xw, ww;
  if (dim == 2) {
    gauss_legendre01((order + 3) / 2, xu, wu);
    gauss_legendre01((order + 2) / 2, xv, wv);
    for (size_t i = 0; i < xu.size(); ++i)
      for (size_t j = 0; j < xv.size(); ++j) {
        q.x.push_back({xu[i], (1. - xu[i]) * xv[j], 0});
        q.w.push_back(wu[i] * wv[j] * (1. - xu[i]));
      }
  } else {
    gauss_legendre01((order + 4) / 2, xu, wu);
    gauss_legendre01((order + 3) / 2, xv, wv);
    gauss_legendre01((order + 2) / 2, xw, ww);
    for (size_t i = 0; i < xu.size(); ++i)
      for (size_t j = 0; j < xv.size(); ++j)
        for (size_t k = 0; k < xw.size(); ++k) {
          const double u = xu[i], v = xv[j];
          q.x.push_back({u, (1. - u) * v, (1. - u) * (1. - v) * xw[k]});
          q.w.push_back(wu[i] * wv[j] * ww[k] * (1. - u) * (1. - u) * (1. - v));
        }
  }
  return q;
}

Quadrature element_rule(const RefElement& ref, int order)
{
  return ref.simplex ? simplex_rule(ref.dim, order) : cube_rule(ref.dim, order);
}

// ---------------------------------------------------------------------------
// Lagrange local finite elements (dune/gdt/local/finite-elements/lagrange.hh:65-186
// wraps dune-localfunctions; order() as in local/finite-elements/wrapper.hh:125-127)
// ---------------------------------------------------------------------------
struct LocalFE {
  const RefElement* ref = nullptr;
  int order = 1, size = 0;
  std::vector<std::array<double, 3>> nodes;
  std::vector<std::vector<int>> node_subentity; // local vertices spanning the node's subentity
  // simplex order 2: node -> (a, b) vertex pair (a == b for vertices)
  std::vector<std::array<int, 2>> simplex_pair;

  static double lag1d(int k, int m, double t)
  {
    double v = 1.;
    for (int l = 0; l <= k; ++l)
      if (l != m)
        v *= (t - double(l) / k) / (double(m) / k - double(l) / k);
    return v;
  }
  static double dlag1d(int k, int m, double t)
  {
    double s = 0.;
    for (int q = 0; q <= k; ++q) {
      if (q == m)
        continue;
      double v = 1. / (double(m) / k - double(q) / k);
      for (int l = 0; l <= k; ++l)
        if (l != m && l != q)
          v *= (t - double(l) / k) / (double(m) / k - double(l) / k);
      s += v;
    }
    return s;
  }

  void init(const RefElement* r, int k)
  {
    ref = r, order = k;
    const int d = r->dim;
    if (!r->simplex) {
      const int n1 = k + 1;
      size = (d == 2) ? n1 * n1 : n1 * n1 * n1;
      for (int t = 0; t < size; ++t) {
        const int idx[3] = {t % n1, (t / n1) % n1, t / (n1 * n1)};
        std::array<double, 3> x = {0, 0, 0};
        for (int a = 0; a < d; ++a)
          x[a] = double(idx[a]) / k;
        nodes.push_back(x);
        std::vector<int> sub;
        for (int b = 0; b < r->nverts; ++b) {
          bool ok = true;
          for (int a = 0; a < d; ++a) {
            const int bit = (b >> a) & 1;
            if (idx[a] == 0 && bit != 0)
              ok = false;
            if (idx[a] == k && bit != 1)
              ok = false;
          }
          if (ok)
            sub.push_back(b);
        }
        node_subentity.push_back(sub);
      }
    } else {
      if (k < 1 || k > 2)
        throw Error("simplex Lagrange elements of order 1 and 2 only");
      for (int v = 0; v < r->nverts; ++v) {
        nodes.push_back(r->verts[v]);
        node_subentity.push_back({v});
        simplex_pair.push_back({v, v});
      }
      if (k == 2) {
        for (int a = 0; a < r->nverts; ++a)
          for (int b = a + 1; b < r->nverts; ++b) {
            std::array<double, 3> x;
            for (int c = 0; c < 3; ++c)
              x[c] = 0.5 * (r->verts[a][c] + r->verts[b][c]);
            nodes.push_back(x);
            node_subentity.push_back({a, b});
            simplex_pair.push_back({a, b});
          }
      }
      size = int(nodes.size());
    }
  }

  // barycentric coordinates and their (constant) reference gradients
  void bary(const double* x, double* lam, double (*glam)[3]) const
  {
    const int d = ref->dim;
    lam[0] = 1.;
    for (int a = 0; a < d; ++a)
      lam[0] -= x[a];
    for (int a = 0; a < d; ++a)
      lam[a + 1] = x[a];
    for (int v = 0; v <= d; ++v)
      for (int a = 0; a < 3; ++a)
        glam[v][a] = 0.;
    for (int a = 0; a < d; ++a)
      glam[0][a] = -1., glam[a + 1][a] = 1.;
  }

  void evaluate(const double* x, double* vals) const
  {
    const int d = ref->dim;
    if (!ref->simplex) {
      const int n1 = order + 1;
      for (int t = 0; t < size; ++t) {
        const int idx[3] = {t % n1, (t / n1) % n1, t / (n1 * n1)};
        double v = 1.;
        for (int a = 0; a < d; ++a)
          v *= lag1d(order, idx[a], x[a]);
        vals[t] = v;
      }
    } else {
      double lam[4], glam[4][3];
      bary(x, lam, glam);
      for (int t = 0; t < size; ++t) {
        const int a = simplex_pair[t][0], b = simplex_pair[t][1];
        if (order == 1)
          vals[t] = lam[a];
        else if (a == b)
          vals[t] = lam[a] * (2. * lam[a] - 1.);
        else
          vals[t] = 4. * lam[a] * lam[b];
      }
    }
  }

  // reference gradients, grads[t*3 + a]
  void jacobians(const double* x, double* grads) const
  {
    const int d = ref->dim;
    if (!ref->simplex) {
      const int n1 = order + 1;
      for (int t = 0; t < size; ++t) {
        const int idx[3] = {t % n1, (t / n1) % n1, t / (n1 * n1)};
        for (int a = 0; a < 3; ++a)
          grads[t * 3 + a] = 0.;
        for (int a = 0; a < d; ++a) {
          double v = dlag1d(order, idx[a], x[a]);
          for (int c = 0; c < d; ++c)
            if (c != a)
              v *= lag1d(order, idx[c], x[c]);
          grads[t * 3 + a] = v;
        }
      }
    } else {
      double lam[4], glam[4][3];
      bary(x, lam, glam);
      for (int t = 0; t < size; ++t) {
        const int a = simplex_pair[t][0], b = simplex_pair[t][1];
        for (int c = 0; c < 3; ++c) {
          if (order == 1)
            grads[t * 3 + c] = glam[a][c];
          else if (a == b)
            grads[t * 3 + c] = (4. * lam[a] - 1.) * glam[a][c];
          else
            grads[t * 3 + c] = 4. * (glam[a][c] * lam[b] + lam[a] * glam[b][c]);
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// affine element geometry (integrationElement, jacobianInverseTransposed,
// spaces/basis/default.hh:152-175 applies J^{-T} to reference gradients)
// ---------------------------------------------------------------------------
struct Geometry {
  int dim = 0;
  double x0[3] = {0, 0, 0};
  double J[3][3] = {{0}}, Jinv[3][3] = {{0}};
  double det = 0., integration_element = 0.;

  void global(const double* xh, double* x) const
  {
    for (int a = 0; a < dim; ++a) {
      double s = x0[a];
      for (int c = 0; c < dim; ++c)
        s += J[a][c] * xh[c];
      x[a] = s;
    }
  }
  void local(const double* x, double* xh) const
  {
    double dx[3];
    for (int a = 0; a < dim; ++a)
      dx[a] = x[a] - x0[a];
    for (int a = 0; a < dim; ++a) {
      double s = 0.;
      for (int c = 0; c < dim; ++c)
        s += Jinv[a][c] * dx[c];
      xh[a] = s;
    }
  }
  // physical gradient g = J^{-T} gh
  void transform_gradient(const double* gh, double* g) const
  {
    for (int a = 0; a < dim; ++a) {
      double s = 0.;
      for (int c = 0; c < dim; ++c)
        s += Jinv[c][a] * gh[c];
      g[a] = s;
    }
  }
};

// ---------------------------------------------------------------------------
// grid
// ---------------------------------------------------------------------------
struct Mesh {
  // mixed meshes (triangles and quadrilaterals): ref holds the strides of conn, nb and nbface (those of the
  // quadrilateral: 4 vertices, 4 faces; -1 vertices and -2 neighbours past an element's own), eref(e) the
  // element's own reference element
  RefElement ref;
  bool mixed = false;
  std::vector<int8_t> etype; // mixed: 0 quadrilateral, 1 triangle
  RefElement refs[2];
  const RefElement& eref(int64_t e) const
  {
    return mixed ? refs[etype[e]] : ref;
  }
  int64_t nv = 0, ne = 0;
  std::vector<double> coords; // nv * dim
  std::vector<int32_t> conn;  // ne * nverts
  std::vector<int32_t> nb;    // ne * nfaces, neighbour element or -1
  std::vector<int32_t> nbface;

  Geometry geometry(int64_t e) const
  {
    Geometry g;
    const int d = ref.dim;
    g.dim = d;
    const int32_t* c = &conn[e * ref.nverts];
    for (int a = 0; a < d; ++a)
      g.x0[a] = coords[int64_t(c[0]) * d + a];
    for (int k = 0; k < d; ++k) {
      const int lv = eref(e).simplex ? (k + 1) : (1 << k);
      for (int a = 0; a < d; ++a)
        g.J[a][k] = coords[int64_t(c[lv]) * d + a] - g.x0[a];
    }
    bool diagonal = true;
    for (int a = 0; a < d; ++a)
      for (int k = 0; k < d; ++k)
        if (a != k && g.J[a][k] != 0.)
          diagonal = false;
    if (diagonal) {
      g.det = 1.;
      for (int a = 0; a < d; ++a)
        g.det *= g.J[a][a], g.Jinv[a][a] = 1. / g.J[a][a];
    } else if (d == 2) {
      g.det = g.J[0][0] * g.J[1][1] - g.J[0][1] * g.J[1][0];
      g.Jinv[0][0] = g.J[1][1] / g.det, g.Jinv[0][1] = -g.J[0][1] / g.det;
      g.Jinv[1][0] = -g.J[1][0] / g.det, g.Jinv[1][1] = g.J[0][0] / g.det;
    } else {
      const double(*A)[3] = g.J;
      g.det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
              + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
      g.Jinv[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / g.det;
      g.Jinv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / g.det;
      g.Jinv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / g.det;
      g.Jinv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / g.det;
      g.Jinv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / g.det;
      g.Jinv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / g.det;
      g.Jinv[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / g.det;
      g.Jinv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / g.det;
      g.Jinv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / g.det;
    }
    g.integration_element = std::fabs(g.det);
    return g;
  }

  void check_affine() const
  {
    if (ref.simplex)
      return;
    const int d = ref.dim;
    for (int64_t e = 0; e < ne; ++e) {
      if (eref(e).simplex)
        continue;
      const int32_t* c = &conn[e * ref.nverts];
      for (int b = 0; b < ref.nverts; ++b)
        for (int a = 0; a < d; ++a) {
          double expect = coords[int64_t(c[0]) * d + a];
          for (int k = 0; k < d; ++k)
            if ((b >> k) & 1)
              expect += coords[int64_t(c[1 << k]) * d + a] - coords[int64_t(c[0]) * d + a];
          const double got = coords[int64_t(c[b]) * d + a];
          if (std::fabs(expect - got) > 1e-12 * (1. + std::fabs(got)))
            throw Error("oracle supports affine (parallelogram/parallelepiped) cube elements only");
        }
    }
  }

  void build_neighbors()
  {
    const int nf = ref.nfaces;
    nb.assign(ne * nf, -1);
    nbface.assign(ne * nf, -1);
    std::map<std::vector<int32_t>, std::pair<int64_t, int>> open;
    for (int64_t e = 0; e < ne; ++e)
      for (int f = 0; f < nf; ++f) {
        if (f >= eref(e).nfaces) {
          nb[e * nf + f] = -2;
          continue;
        }
        std::vector<int32_t> key;
        for (int lv : eref(e).faces[f])
          key.push_back(conn[e * ref.nverts + lv]);
        std::sort(key.begin(), key.end());
        auto it = open.find(key);
        if (it == open.end()) {
          open.emplace(key, std::make_pair(e, f));
        } else {
          const int64_t o = it->second.first;
          const int of = it->second.second;
          nb[e * nf + f] = int32_t(o), nbface[e * nf + f] = of;
          nb[o * nf + of] = int32_t(e), nbface[o * nf + of] = f;
          open.erase(it);
        }
      }
  }
};

Mesh* structured_mesh(int geom, const int32_t* cells, const double* lo, const double* hi)
{
  auto* m = new Mesh;
  m->ref = make_ref(geom);
  const int d = m->ref.dim;
  int64_t n[3] = {1, 1, 1}, nvx[3] = {1, 1, 1};
  double h[3] = {0, 0, 0};
  for (int a = 0; a < d; ++a) {
    if (cells[a] < 1)
      throw Error("cells must be positive");
    n[a] = cells[a], nvx[a] = cells[a] + 1, h[a] = (hi[a] - lo[a]) / cells[a];
  }
  m->nv = nvx[0] * nvx[1] * nvx[2];
  m->coords.resize(m->nv * d);
  for (int64_t v = 0; v < m->nv; ++v) {
    const int64_t idx[3] = {v % nvx[0], (v / nvx[0]) % nvx[1], v / (nvx[0] * nvx[1])};
    for (int a = 0; a < d; ++a)
      m->coords[v * d + a] = lo[a] + idx[a] * h[a];
  }
  const int64_t ncell = n[0] * n[1] * n[2];
  auto vid = [&](int64_t i, int64_t j, int64_t k) { return int32_t((k * nvx[1] + j) * nvx[0] + i); };
  for (int64_t c = 0; c < ncell; ++c) {
    const int64_t i = c % n[0], j = (c / n[0]) % n[1], k = c / (n[0] * n[1]);
    int32_t corner[8];
    for (int b = 0; b < (1 << d); ++b)
      corner[b] = vid(i + (b & 1), j + ((b >> 1) & 1), k + ((b >> 2) & 1));
    if (geom == OR_QUAD || geom == OR_HEX) {
      for (int b = 0; b < (1 << d); ++b)
        m->conn.push_back(corner[b]);
    } else if (geom == OR_TRI) {
      // permutations (x,y), (y,x): v0, v0+e_p0, v0+e_p0+e_p1
      m->conn.insert(m->conn.end(), {corner[0], corner[1], corner[3]});
      m->conn.insert(m->conn.end(), {corner[0], corner[2], corner[3]});
    } else {
      static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
      for (auto& p : perms) {
        const int b1 = 1 << p[0], b2 = b1 | (1 << p[1]);
        m->conn.insert(m->conn.end(), {corner[0], corner[b1], corner[b2], corner[7]});
      }
    }
  }
  m->ne = int64_t(m->conn.size()) / m->ref.nverts;
  m->build_neighbors();
  return m;
}

// ---------------------------------------------------------------------------
// spaces + mappers (spaces/mapper/continuous.hh:139-182,
// spaces/mapper/discontinuous.hh:102-125)
// ---------------------------------------------------------------------------
struct Space {
  const Mesh* mesh = nullptr;
  bool continuous = true;
  int order = 1;
  LocalFE fe;                   // mixed meshes: the quadrilateral's (the largest: fe.size is the stride)
  LocalFE fes[2];               // mixed meshes: the local finite element of every element type
  int64_t size = 0;
  std::vector<int32_t> indices; // ne * fe.size (mixed meshes: -1 past an element's own local size)
  const LocalFE& fe_of(int64_t e) const
  {
    return mesh->mixed ? fes[mesh->etype[e]] : fe;
  }

  void init(const Mesh* m, bool cont, int k)
  {
    mesh = m, continuous = cont, order = k;
    if (m->mixed) {
      init_mixed(m, cont, k);
      return;
    }
    fe.init(&m->ref, k);
    const int ls = fe.size;
    indices.resize(m->ne * ls);
    if (!cont) {
      // offset[element] + local index, offsets in element order
      for (int64_t e = 0; e < m->ne; ++e)
        for (int i = 0; i < ls; ++i)
          indices[e * ls + i] = int32_t(e * ls + i);
      size = m->ne * ls;
      return;
    }
    if (k == 1) {
      // Q1/P1 nodes coincide with the reference vertices: global DoF = vertex index
      for (int64_t e = 0; e < m->ne; ++e)
        for (int i = 0; i < ls; ++i)
          indices[e * ls + i] = m->conn[e * m->ref.nverts + fe.node_subentity[i][0]];
      size = m->nv;
      return;
    }
    // higher order: one DoF per (subentity, position); order <= 2 has one node per subentity
    std::map<std::vector<int32_t>, int32_t> ids;
    for (int64_t e = 0; e < m->ne; ++e)
      for (int i = 0; i < ls; ++i) {
        std::vector<int32_t> key;
        for (int lv : fe.node_subentity[i])
          key.push_back(m->conn[e * m->ref.nverts + lv]);
        std::sort(key.begin(), key.end());
        if (key.size() > 1 && order > 2)
          throw Error("continuous spaces of order > 2 not supported");
        auto it = ids.find(key);
        if (it == ids.end())
          it = ids.emplace(key, int32_t(ids.size())).first;
        indices[e * ls + i] = it->second;
      }
    size = int64_t(ids.size());
  }

  // the mappers over a mixed mesh (spaces/mapper/continuous.hh:117-182, discontinuous.hh:102-125): every
  // element with its own type's local finite element; continuous order 1: the vertices, order 2: one DoF per
  // subentity in order of first appearance in the walk; discontinuous: offset[element] + local index
  void init_mixed(const Mesh* m, bool cont, int k)
  {
    fe.init(&m->refs[0], k);
    fes[0].init(&m->refs[0], k), fes[1].init(&m->refs[1], k);
    const int ls = fe.size;
    indices.assign(m->ne * ls, -1);
    if (!cont) {
      int64_t off = 0;
      for (int64_t e = 0; e < m->ne; ++e)
        for (int i = 0; i < fe_of(e).size; ++i)
          indices[e * ls + i] = int32_t(off++);
      size = off;
      return;
    }
    std::map<std::vector<int32_t>, int32_t> ids;
    for (int64_t e = 0; e < m->ne; ++e)
      for (int i = 0; i < fe_of(e).size; ++i) {
        std::vector<int32_t> key;
        for (int lv : fe_of(e).node_subentity[i])
          key.push_back(m->conn[e * m->ref.nverts + lv]);
        if (k == 1) {
          indices[e * ls + i] = key[0];
          continue;
        }
        std::sort(key.begin(), key.end());
        auto it = ids.find(key);
        if (it == ids.end())
          it = ids.emplace(key, int32_t(ids.size())).first;
        indices[e * ls + i] = it->second;
      }
    size = k == 1 ? m->nv : int64_t(ids.size());
  }
};

// ---------------------------------------------------------------------------
// coefficient fields (XT GridFunction restricted to element-wise constants)
// ---------------------------------------------------------------------------
struct Coefficient {
  std::vector<double> values; // empty -> constant
  double constant = 1.;
  double at(int64_t e) const
  {
    return values.empty() ? constant : values[e];
  }
  void set(const double* v, double c, int64_t ne)
  {
    constant = c;
    if (v)
      values.assign(v, v + ne);
  }
};

enum ElementFormKind { EF_LAPLACE = 0, EF_PRODUCT = 1 };
enum FaceFormKind { FF_INNER_COUPLING = 0, FF_INNER_PENALTY = 1, FF_DIRICHLET_COUPLING = 2, FF_BOUNDARY_PENALTY = 3 };

struct ElementForm {
  int kind;
  Coefficient coeff;
  int over_integrate;
  double scaling;
  // a coefficient varying inside the elements (an XT GridFunction of order coeff_order): its values at the
  // points of each element's rule of order coeff_order + 2 k + over_integrate, element-ordered, element e's
  // from points_off[e]; empty: coeff
  std::vector<double> points;
  std::vector<int64_t> points_off;
  int coeff_order = 0;
  // a matrix-valued diffusion (GridFunction<E, d, d>): d x d per element, or d x d for all (Laplace forms)
  std::vector<double> tensor;
};
struct FaceForm {
  int kind;
  double param; // symmetry prefactor or penalty
  Coefficient kappa, omega;
  int over_integrate;
  double scaling;
  // boundary forms: the intersections the form is applied on, by boundary intersection in walk order
  // (ApplyOn::CustomBoundaryIntersections(boundary_info, type)); empty: every boundary intersection
  std::vector<uint8_t> mask;
  // matrix-valued diffusion and weight (GridFunction<E, d, d>): d x d per element; empty: the scalars above
  std::vector<double> kappa_t, omega_t;
};
struct Source {
  int kind;
  Coefficient f;
  int f_order;
  int over_integrate;
};
// right-hand side terms on boundary intersections (LocalIntersectionIntegralFunctional)
enum { BS_DIRICHLET_COUPLING = 0, BS_BOUNDARY_PENALTY = 1, BS_PRODUCT = 2 };
struct BoundarySource {
  int kind;
  double param;         // symmetry prefactor or penalty
  Coefficient coeff;    // diffusion (coupling) or weight (penalty, product)
  std::vector<double> g; // data at the face quadrature points [boundary face][point], empty -> g_const
  double g_const = 0.;
  int g_order = 0;
  int over_integrate = 0;
  std::vector<uint8_t> mask; // as FaceForm::mask
  std::vector<double> coeff_t; // matrix-valued diffusion / weight, d x d per element; empty: coeff
};

// bound local basis: values and physical gradients at a reference point
struct BoundBasis {
  const LocalFE* fe = nullptr;
  const Geometry* geo = nullptr;
  std::vector<double> vals, rgrads, grads; // grads[i*3+a]
  void bind(const LocalFE* f, const Geometry* g)
  {
    fe = f, geo = g;
    vals.resize(f->size), rgrads.resize(3 * f->size), grads.assign(3 * f->size, 0.);
  }
  void evaluate(const double* xh)
  {
    fe->evaluate(xh, vals.data());
  }
  void jacobians(const double* xh)
  {
    fe->jacobians(xh, rgrads.data());
    for (int i = 0; i < fe->size; ++i)
      geo->transform_gradient(&rgrads[3 * i], &grads[3 * i]);
  }
};

inline double dot(const double* a, const double* b, int d)
{
  double s = 0.;
  for (int i = 0; i < d; ++i)
    s += a[i] * b[i];
  return s;
}

// (kappa I g) . h with the reference's matrix-vector then dot evaluation order
inline double kgrad_dot(double kappa, const double* g, const double* h, int d)
{
  double s = 0.;
  for (int a = 0; a < d; ++a)
    s += (kappa * g[a]) * h[a];
  return s;
}
// (k g) . n, or (K g) . n with the element's d x d block K of a matrix-valued diffusion (laplace-ipdg.hh:158-186)
inline double flux_n(double k, const double* K, const double* g, const double* n, int d);
// n . (w n), or n . (W n) (laplace-ipdg.hh:149-150, ipdg.hh:149-153)
inline double n_w_n(double w, const double* W, const double* n, int d)
{
  double v = 0.;
  for (int a = 0; a < d; ++a) {
    double wn = 0.;
    if (W)
      for (int b = 0; b < d; ++b)
        wn += W[a * d + b] * n[b];
    else
      wn = w * n[a];
    v += n[a] * wn;
  }
  return v;
}
// (K g) . h for a matrix-valued weight K (d x d, row-major): (weight * ansatz_grad) * test_grad,
// local/integrands/laplace.hh:99-101
inline double kmat_grad_dot(const double* K, const double* g, const double* h, int d)
{
  double s = 0.;
  for (int a = 0; a < d; ++a) {
    double kg = 0.;
    for (int b = 0; b < d; ++b)
      kg += K[a * d + b] * g[b];
    s += kg * h[a];
  }
  return s;
}

inline double flux_n(double k, const double* K, const double* g, const double* n, int d)
{
  return K ? kmat_grad_dot(K, g, n, d) : kgrad_dot(k, g, n, d);
}

struct Operator {
  Space* space = nullptr;  // range space: rows (test functions), right-hand side, constraints
  Space* source = nullptr; // source space: columns (ansatz functions); the range space unless set
  int stencil = 0;
  int64_t n = 0;
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> cols;
  std::vector<double> values, rhs;
  std::vector<ElementForm> eforms;
  std::vector<FaceForm> cforms, bforms;
  std::vector<Source> sources;
  std::vector<BoundarySource> bsources;
  std::vector<int64_t> bface_index; // boundary intersections numbered in walk order (element, then face)
  void number_boundary_faces()
  {
    const Mesh& m = *space->mesh;
    const int nf = m.ref.nfaces;
    bface_index.assign(m.ne * nf, -1);
    int64_t c = 0;
    for (int64_t e = 0; e < m.ne; ++e)
      for (int f = 0; f < nf; ++f)
        if (m.nb[e * nf + f] == -1)
          bface_index[e * nf + f] = c++;
  }
  bool dirichlet = false;
  bool only_clear = false, ensure_symmetry = true; // the variant of DirichletConstraints::apply
  std::vector<int32_t> dirichlet_dofs;
  // boundary_info_.type(intersection) == DirichletBoundary() (dirichlet-constraints.hh:93), by boundary
  // intersection in walk order; empty: every boundary intersection is a Dirichlet intersection
  std::vector<uint8_t> dirichlet_faces;
  // filter of the boundary forms and functionals appended next (empty: every boundary intersection)
  std::vector<uint8_t> pending_mask;
  bool on_boundary_subset(const std::vector<uint8_t>& mask, int64_t e, int f) const
  {
    return mask.empty() || mask[bface_index[e * space->mesh->ref.nfaces + f]] != 0;
  }

  // sparsity pattern, tools/sparsity-pattern.hh:33-50 (element) and :117-144 (element+intersection)
  void build_pattern()
  {
    const Mesh& m = *space->mesh;
    const int ls = space->fe.size, nf = m.ref.nfaces;
    n = space->size;
    std::vector<std::set<int32_t>> pattern(n);
    for (int64_t e = 0; e < m.ne; ++e) {
      const int32_t* gi = &space->indices[e * ls];
      const int32_t* si = &source->indices[e * ls];
      const int le = space->fe_of(e).size;
      if (stencil != OR_STENCIL_INTERSECTION)
        for (int i = 0; i < le; ++i)
          for (int j = 0; j < le; ++j)
            pattern[gi[i]].insert(si[j]);
      if (stencil != OR_STENCIL_ELEMENT)
        for (int f = 0; f < nf; ++f) {
          const int32_t o = m.nb[e * nf + f];
          if (o < 0)
            continue;
          const int32_t* gj = &source->indices[int64_t(o) * ls];
          for (int i = 0; i < le; ++i)
            for (int j = 0; j < source->fe_of(o).size; ++j)
              pattern[gi[i]].insert(gj[j]);
        }
    }
    row_ptr.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i)
      row_ptr[i + 1] = row_ptr[i] + int64_t(pattern[i].size());
    cols.resize(row_ptr[n]);
    for (int64_t i = 0; i < n; ++i) {
      int64_t k = row_ptr[i];
      for (int32_t c : pattern[i])
        cols[k++] = c;
    }
    values.assign(row_ptr[n], 0.);
    rhs.assign(n, 0.);
  }

  // ISTL-style add_to_entry: locate the column inside the sorted row
  inline void add_to_entry(int32_t i, int32_t j, double v, bool atomic)
  {
    const int32_t* b = &cols[row_ptr[i]];
    const int32_t* e = &cols[row_ptr[i + 1]];
    const int32_t* it = std::lower_bound(b, e, j);
    if (it == e || *it != j)
      throw Error("add_to_entry: entry not in sparsity pattern");
    double& slot = values[row_ptr[i] + (it - b)];
    if (atomic) {
#pragma omp atomic
      slot += v;
    } else {
      slot += v;
    }
  }

  // LocalElementIntegralBilinearForm::apply2, local/bilinear-forms/integrals.hh:97-134
  void element_local(const ElementForm& F, int64_t e, const Geometry& geo, BoundBasis& basis,
                     std::vector<double>& result) const
  {
    const int ls = space->fe_of(e).size, d = geo.dim;
    result.assign(ls * ls, 0.);
    const int basis_order = space->fe.order;
    // LocalLaplaceIntegrand::order (local/integrands/laplace.hh:74-79) and
    // LocalElementProductIntegrand::order (local/integrands/product.hh:88-100): weight order (0 for a
    // piecewise constant coefficient) + test + ansatz
    const int order = F.coeff_order + basis_order + basis_order + F.over_integrate;
    const Quadrature q = element_rule(space->mesh->eref(e), order);
    std::vector<double> iv(ls * ls);
    for (size_t qp = 0; qp < q.w.size(); ++qp) {
      const double* xh = q.x[qp].data();
      const double factor = geo.integration_element * q.w[qp];
      // local_weight_->evaluate(x) (laplace.hh:92, product.hh:118)
      const double c = F.points.empty() ? F.coeff.at(e) : F.points[size_t(F.points_off[e]) + qp];
      if (F.kind == EF_LAPLACE && !F.tensor.empty()) {
        const double* K = &F.tensor[F.tensor.size() == size_t(d * d) ? 0 : size_t(e) * d * d];
        basis.jacobians(xh);
        for (int i = 0; i < ls; ++i)
          for (int j = 0; j < ls; ++j)
            iv[i * ls + j] = 0. + kmat_grad_dot(K, &basis.grads[3 * j], &basis.grads[3 * i], d);
      } else if (F.kind == EF_LAPLACE) {
        // local/integrands/laplace.hh:94-101
        basis.jacobians(xh);
        for (int i = 0; i < ls; ++i)
          for (int j = 0; j < ls; ++j)
            iv[i * ls + j] = 0. + kgrad_dot(c, &basis.grads[3 * j], &basis.grads[3 * i], d);
      } else {
        // local/integrands/product.hh:120-128
        basis.evaluate(xh);
        for (int i = 0; i < ls; ++i)
          for (int j = 0; j < ls; ++j)
            iv[i * ls + j] = (c * basis.vals[i]) * basis.vals[j];
      }
      for (int i = 0; i < ls; ++i)
        for (int j = 0; j < ls; ++j)
          result[i * ls + j] += iv[i * ls + j] * factor;
    }
  }

  // face quadrature support: inside/outside reference points, unit outer normal, face volume
  struct FaceGeo {
    std::vector<std::array<double, 3>> xin, xout;
    std::vector<double> w;
    double normal[3] = {0, 0, 0};
    double integration_element = 0., diameter = 0.;
  };

  FaceGeo face_geometry(int64_t e, int f, int64_t o, const Geometry& gin, const Geometry* gout, int order) const
  {
    const Mesh& m = *space->mesh;
    if (m.ref.dim == 3)
      return face_geometry_3d(f, gin, gout, order);
    if (m.ref.dim != 2)
      throw Error("oracle face integrals are implemented for 2d and 3d grids");
    FaceGeo fg;
    const RefElement& r = m.eref(e); // the inside element's (mixed meshes: its own type)
    const int a = r.faces[f][0], b = r.faces[f][1];
    double Pa[3], Pb[3];
    gin.global(r.verts[a].data(), Pa);
    gin.global(r.verts[b].data(), Pb);
    const double tx = Pb[0] - Pa[0], ty = Pb[1] - Pa[1];
    const double len = std::sqrt(tx * tx + ty * ty);
    fg.integration_element = len, fg.diameter = len;
    double nx = ty / len, ny = -tx / len;
    // orient away from the inside element's centroid
    double cen[2] = {0, 0};
    for (int v = 0; v < r.nverts; ++v) {
      double P[3];
      gin.global(r.verts[v].data(), P);
      cen[0] += P[0] / r.nverts, cen[1] += P[1] / r.nverts;
    }
    if (nx * (Pa[0] - cen[0]) + ny * (Pa[1] - cen[1]) < 0.)
      nx = -nx, ny = -ny;
    fg.normal[0] = nx + 0., fg.normal[1] = ny + 0.;
    const Quadrature q = line_rule(order);
    for (size_t qp = 0; qp < q.w.size(); ++qp) {
      const double t = q.x[qp][0];
      std::array<double, 3> xh = {0, 0, 0};
      for (int c = 0; c < 2; ++c)
        xh[c] = (1. - t) * r.verts[a][c] + t * r.verts[b][c];
      fg.xin.push_back(xh);
      if (gout) {
        double X[3], xo[3] = {0, 0, 0};
        gin.global(xh.data(), X);
        gout->local(X, xo);
        fg.xout.push_back({xo[0], xo[1], 0.});
      }
      fg.w.push_back(q.w[qp]);
    }
    (void)e, (void)o;
    return fg;
  }

  // 3d faces of affine elements (integrals.hh:226-233, 352-356): a triangle (tetrahedra) or a parallelogram
  // (hexahedra) X(s, t) = P0 + s (P1 - P0) + t (P2 - P0) over the corners P0, P1, P2 (P3) of reference face f,
  // integration element |(P1 - P0) x (P2 - P0)|, the simplex rule of the reference triangle or the tensor Gauss
  // rule of the square; the normal points away from the inside element, the diameter is the largest distance
  // between two corners (XT::Grid::diameter, the default of ipdg.hh:27-38 for d > 1)
  FaceGeo face_geometry_3d(int f, const Geometry& gin, const Geometry* gout, int order) const
  {
    const Mesh& m = *space->mesh;
    const auto& fv = m.ref.faces[f];
    const int nc = int(fv.size());
    FaceGeo fg;
    double P[4][3], cen[3] = {0, 0, 0};
    for (int k = 0; k < nc; ++k)
      gin.global(m.ref.verts[fv[k]].data(), P[k]);
    for (int v = 0; v < m.ref.nverts; ++v) {
      double X[3];
      gin.global(m.ref.verts[v].data(), X);
      for (int c = 0; c < 3; ++c)
        cen[c] += X[c] / m.ref.nverts;
    }
    double E1[3], E2[3], N[3];
    for (int c = 0; c < 3; ++c)
      E1[c] = P[1][c] - P[0][c], E2[c] = P[2][c] - P[0][c];
    N[0] = E1[1] * E2[2] - E1[2] * E2[1];
    N[1] = E1[2] * E2[0] - E1[0] * E2[2];
    N[2] = E1[0] * E2[1] - E1[1] * E2[0];
    const double area = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    fg.integration_element = area;
    double side = 0.;
    for (int c = 0; c < 3; ++c)
      side += N[c] * (P[0][c] - cen[c]);
    for (int c = 0; c < 3; ++c)
      fg.normal[c] = (side < 0. ? -N[c] : N[c]) / area + 0.;
    double diam = 0.;
    for (int a = 0; a < nc; ++a)
      for (int b = a + 1; b < nc; ++b) {
        double d2 = 0.;
        for (int c = 0; c < 3; ++c)
          d2 += (P[a][c] - P[b][c]) * (P[a][c] - P[b][c]);
        diam = std::max(diam, std::sqrt(d2));
      }
    fg.diameter = diam;
    const bool simplex = m.ref.simplex;
    const Quadrature q = simplex ? simplex_rule(2, order) : cube_rule(2, order);
    for (size_t qp = 0; qp < q.w.size(); ++qp) {
      const double s = q.x[qp][0], t = q.x[qp][1];
      std::array<double, 3> xh = {0, 0, 0};
      for (int c = 0; c < 3; ++c)
        xh[c] = simplex ? (1. - s - t) * m.ref.verts[fv[0]][c] + s * m.ref.verts[fv[1]][c] + t * m.ref.verts[fv[2]][c]
                        : (1. - s) * (1. - t) * m.ref.verts[fv[0]][c] + s * (1. - t) * m.ref.verts[fv[1]][c] +
                              (1. - s) * t * m.ref.verts[fv[2]][c] + s * t * m.ref.verts[fv[3]][c];
      fg.xin.push_back(xh);
      if (gout) {
        double X[3], xo[3] = {0, 0, 0};
        gin.global(xh.data(), X);
        gout->local(X, xo);
        fg.xout.push_back({xo[0], xo[1], xo[2]});
      }
      fg.w.push_back(q.w[qp]);
    }
    return fg;
  }

  // LocalCouplingIntersectionIntegralBilinearForm::apply2 (integrals.hh:197-267) with
  // InnerCoupling (local/integrands/laplace-ipdg.hh:107-186) or InnerPenalty (local/integrands/ipdg.hh:113-171)
  void coupling_local(const FaceForm& F, int64_t e, int64_t o, const FaceGeo& fg, BoundBasis& bin,
                      BoundBasis& bout, std::vector<double> blocks[4]) const
  {
    // blocks in_in [li][li], in_out [li][lo], out_in [lo][li], out_out [lo][lo] (li, lo: the local sizes of
    // the inside and outside elements, different between a triangle and a quadrilateral)
    const int li = bin.fe->size, lo = bout.fe->size, d = space->mesh->ref.dim;
    const int rows[4] = {li, li, lo, lo}, colsz[4] = {li, lo, li, lo};
    for (int k = 0; k < 4; ++k)
      blocks[k].assign(rows[k] * colsz[k], 0.);
    std::vector<double> iv[4];
    for (int k = 0; k < 4; ++k)
      iv[k].assign(rows[k] * colsz[k], 0.);
    const double* n = fg.normal;
    const double kin = F.kappa.at(e), kout = F.kappa.at(o);
    const double win = F.omega.at(e), wout = F.omega.at(o);
    const size_t dd = size_t(d) * d;
    const double* Kin = F.kappa_t.empty() ? nullptr : &F.kappa_t[size_t(e) * dd];
    const double* Kout = F.kappa_t.empty() ? nullptr : &F.kappa_t[size_t(o) * dd];
    const double* Win = F.omega_t.empty() ? nullptr : &F.omega_t[size_t(e) * dd];
    const double* Wout = F.omega_t.empty() ? nullptr : &F.omega_t[size_t(o) * dd];
    for (size_t qp = 0; qp < fg.w.size(); ++qp) {
      for (int k = 0; k < 4; ++k)
        std::fill(iv[k].begin(), iv[k].end(), 0.);
      bin.evaluate(fg.xin[qp].data()), bin.jacobians(fg.xin[qp].data());
      bout.evaluate(fg.xout[qp].data()), bout.jacobians(fg.xout[qp].data());
      // delta = n . (omega n)
      const double delta_plus = n_w_n(wout, Wout, n, d), delta_minus = n_w_n(win, Win, n, d);
      if (F.kind == FF_INNER_COUPLING) {
        const double s = F.param;
        const double weight_minus = delta_plus / (delta_plus + delta_minus);
        const double weight_plus = delta_minus / (delta_plus + delta_minus);
        for (int i = 0; i < li; ++i) {
          const double kgt_in = flux_n(kin, Kin, &bin.grads[3 * i], n, d);
          for (int j = 0; j < li; ++j) {
            iv[0][i * li + j] += -1.0 * weight_minus * flux_n(kin, Kin, &bin.grads[3 * j], n, d) * bin.vals[i];
            iv[0][i * li + j] += -1.0 * s * weight_minus * bin.vals[j] * kgt_in;
          }
          for (int j = 0; j < lo; ++j) {
            iv[1][i * lo + j] += -1.0 * weight_plus * flux_n(kout, Kout, &bout.grads[3 * j], n, d) * bin.vals[i];
            iv[1][i * lo + j] += s * weight_minus * bout.vals[j] * kgt_in;
          }
        }
        for (int i = 0; i < lo; ++i) {
          const double kgt_out = flux_n(kout, Kout, &bout.grads[3 * i], n, d);
          for (int j = 0; j < li; ++j) {
            iv[2][i * li + j] += weight_minus * flux_n(kin, Kin, &bin.grads[3 * j], n, d) * bout.vals[i];
            iv[2][i * li + j] += -1.0 * s * weight_plus * bin.vals[j] * kgt_out;
          }
          for (int j = 0; j < lo; ++j) {
            iv[3][i * lo + j] += weight_plus * flux_n(kout, Kout, &bout.grads[3 * j], n, d) * bout.vals[i];
            iv[3][i * lo + j] += s * weight_plus * bout.vals[j] * kgt_out;
          }
        }
      } else {
        const double weight = (delta_plus * delta_minus) / (delta_plus + delta_minus); // half harmonic average
        const double penalty = (F.param * weight) / fg.diameter;
        for (int i = 0; i < li; ++i) {
          for (int j = 0; j < li; ++j)
            iv[0][i * li + j] += penalty * bin.vals[j] * bin.vals[i];
          for (int j = 0; j < lo; ++j)
            iv[1][i * lo + j] += -1.0 * penalty * bout.vals[j] * bin.vals[i];
        }
        for (int i = 0; i < lo; ++i) {
          for (int j = 0; j < li; ++j)
            iv[2][i * li + j] += -1.0 * penalty * bin.vals[j] * bout.vals[i];
          for (int j = 0; j < lo; ++j)
            iv[3][i * lo + j] += penalty * bout.vals[j] * bout.vals[i];
        }
      }
      for (int k = 0; k < 4; ++k)
        for (size_t t = 0; t < iv[k].size(); ++t)
          blocks[k][t] += iv[k][t] * fg.integration_element * fg.w[qp];
    }
  }

  // LocalIntersectionIntegralBilinearForm::apply2 (integrals.hh:338-369) with
  // DirichletCoupling (laplace-ipdg.hh:340-368) or BoundaryPenalty (ipdg.hh:254-282)
  void boundary_local(const FaceForm& F, int64_t e, const FaceGeo& fg, BoundBasis& bin,
                      std::vector<double>& result) const
  {
    const int ls = bin.fe->size, d = space->mesh->ref.dim;
    result.assign(ls * ls, 0.);
    std::vector<double> iv(ls * ls);
    const double* n = fg.normal;
    const double kin = F.kappa.at(e), win = F.omega.at(e);
    const size_t dd = size_t(d) * d;
    const double* Kin = F.kappa_t.empty() ? nullptr : &F.kappa_t[size_t(e) * dd];
    const double* Win = F.omega_t.empty() ? nullptr : &F.omega_t[size_t(e) * dd];
    for (size_t qp = 0; qp < fg.w.size(); ++qp) {
      std::fill(iv.begin(), iv.end(), 0.);
      bin.evaluate(fg.xin[qp].data()), bin.jacobians(fg.xin[qp].data());
      if (F.kind == FF_DIRICHLET_COUPLING) {
        const double s = F.param;
        for (int i = 0; i < ls; ++i) {
          const double kgt = flux_n(kin, Kin, &bin.grads[3 * i], n, d);
          for (int j = 0; j < ls; ++j) {
            iv[i * ls + j] += -1.0 * flux_n(kin, Kin, &bin.grads[3 * j], n, d) * bin.vals[i];
            iv[i * ls + j] += -1.0 * s * bin.vals[j] * kgt;
          }
        }
      } else {
        const double nwn = n_w_n(win, Win, n, d);
        const double penalty = (F.param * nwn) / fg.diameter;
        for (int i = 0; i < ls; ++i)
          for (int j = 0; j < ls; ++j)
            iv[i * ls + j] += penalty * bin.vals[j] * bin.vals[i];
      }
      for (int t = 0; t < ls * ls; ++t)
        result[t] += iv[t] * fg.integration_element * fg.w[qp];
    }
  }

  // quadrature order of a boundary functional term: DirichletCoupling as a unary integrand,
  // laplace-ipdg.hh:292-297 (data + diffusion + max(test - 1, 0)); BoundaryPenalty and the intersection
  // product with the data as ansatz (conversion.hh:194-197): weight + test + data
  int boundary_source_order(const BoundarySource& S) const
  {
    const int k = space->fe.order;
    if (S.kind == BS_DIRICHLET_COUPLING)
      return S.g_order + 0 + std::max(k - 1, 0) + S.over_integrate;
    return 0 + k + S.g_order + S.over_integrate;
  }

  // LocalIntersectionIntegralFunctional::apply (local/functionals/integrals.hh:157-185) with
  //   DirichletCoupling (unary): -s g (kappa grad psi_i) . n            (laplace-ipdg.hh:300-320)
  //   BoundaryPenalty(sigma, w).with_ansatz(g): sigma (n.w n)/h g psi_i  (ipdg.hh:254-282, conversion.hh:199-215)
  //   LocalIntersectionProductIntegrand(w).with_ansatz(g): (w g) psi_i   (product.hh:355-377)
  void boundary_source_local(const BoundarySource& S, int64_t e, int64_t bi, const FaceGeo& fg, BoundBasis& bin,
                             std::vector<double>& result) const
  {
    const int ls = bin.fe->size, d = space->mesh->ref.dim;
    result.assign(ls, 0.);
    const double* n = fg.normal;
    const double c = S.coeff.at(e);
    const double* Ct = S.coeff_t.empty() ? nullptr : &S.coeff_t[size_t(e) * d * d];
    const size_t nq = fg.w.size();
    for (size_t qp = 0; qp < nq; ++qp) {
      bin.evaluate(fg.xin[qp].data()), bin.jacobians(fg.xin[qp].data());
      const double g = S.g.empty() ? S.g_const : S.g[bi * nq + qp];
      for (int i = 0; i < ls; ++i) {
        double v;
        if (S.kind == BS_DIRICHLET_COUPLING) {
          v = -1.0 * S.param * g * flux_n(c, Ct, &bin.grads[3 * i], n, d);
        } else if (S.kind == BS_BOUNDARY_PENALTY) {
          const double nwn = n_w_n(c, Ct, n, d);
          const double penalty = (S.param * nwn) / fg.diameter;
          v = penalty * g * bin.vals[i];
        } else {
          v = (c * g) * bin.vals[i];
        }
        result[i] += v * fg.integration_element * fg.w[qp];
      }
    }
  }

  static double esv2007_force(const double* x)
  {
    return M_PI_2 * M_PI * std::cos(M_PI_2 * x[0]) * std::cos(M_PI_2 * x[1]);
  }

  // LocalElementIntegralFunctional::apply (local/functionals/integrals.hh:72-97) with
  // LocalProductIntegrand().with_ansatz(f) (local/integrands/conversion.hh:90-117)
  void source_local(const Source& S, int64_t e, const Geometry& geo, BoundBasis& basis,
                    std::vector<double>& result) const
  {
    const int ls = space->fe_of(e).size;
    result.assign(ls, 0.);
    const int order = 0 + space->fe.order + S.f_order + S.over_integrate;
    const Quadrature q = element_rule(space->mesh->eref(e), order);
    for (size_t qp = 0; qp < q.w.size(); ++qp) {
      const double* xh = q.x[qp].data();
      basis.evaluate(xh);
      double f;
      if (S.kind == 0) {
        f = S.f.at(e);
      } else {
        double X[3];
        geo.global(xh, X);
        f = esv2007_force(X);
      }
      for (int i = 0; i < ls; ++i)
        result[i] += ((1. * basis.vals[i]) * f) * geo.integration_element * q.w[qp];
    }
  }

  void assemble(int nthreads)
  {
    const Mesh& m = *space->mesh;
    const int ls = space->fe.size, nf = m.ref.nfaces;
    std::fill(values.begin(), values.end(), 0.);
    std::fill(rhs.begin(), rhs.end(), 0.);
    const bool atomic = nthreads > 1;
    std::string err;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
      // per-thread copies, as walk(use_tbb=true) copies functors per thread
      BoundBasis basis, bout;
      std::vector<double> local, lvec, blocks[4];
#pragma omp for schedule(static)
      for (int64_t e = 0; e < m.ne; ++e) {
        try {
          const Geometry geo = m.geometry(e);
          basis.bind(&space->fe_of(e), &geo);
          const int32_t* gi = &space->indices[e * ls];  // rows: test functions of the range space
          const int32_t* si = &source->indices[e * ls]; // columns: ansatz functions of the source space
          const int le = space->fe_of(e).size;
          for (const auto& F : eforms) {
            element_local(F, e, geo, basis, local);
            for (int i = 0; i < le; ++i)
              for (int j = 0; j < le; ++j)
                add_to_entry(gi[i], si[j], F.scaling * local[i * le + j], atomic);
          }
          for (const auto& S : sources) {
            source_local(S, e, geo, basis, lvec);
            for (int i = 0; i < le; ++i) {
              if (atomic) {
#pragma omp atomic
                rhs[gi[i]] += lvec[i];
              } else {
                rhs[gi[i]] += lvec[i];
              }
            }
          }
          // boundary functionals, the element's boundary intersections in face order
          for (int f = 0; f < nf && !bsources.empty(); ++f) {
            if (m.nb[e * nf + f] != -1)
              continue;
            const int64_t bi = bface_index[e * nf + f];
            for (const auto& S : bsources) {
              if (!on_boundary_subset(S.mask, e, f))
                continue;
              const FaceGeo fg = face_geometry(e, f, -1, geo, nullptr, boundary_source_order(S));
              boundary_source_local(S, e, bi, fg, basis, lvec);
              for (int i = 0; i < le; ++i) {
                if (atomic) {
#pragma omp atomic
                  rhs[gi[i]] += lvec[i];
                } else {
                  rhs[gi[i]] += lvec[i];
                }
              }
            }
          }
          for (int f = 0; f < nf; ++f) {
            const int64_t o = m.nb[e * nf + f];
            if (o >= 0 && e < o && !cforms.empty()) {
              // ApplyOn::InnerIntersectionsOnce: visited from the element with the smaller index
              const Geometry gout = m.geometry(o);
              bout.bind(&space->fe_of(o), &gout);
              const int lo = space->fe_of(o).size;
              const int32_t* gj = &space->indices[o * ls];
              const int32_t* sj = &source->indices[o * ls];
              for (const auto& F : cforms) {
                const int order = 0 + 0 + space->fe.order + space->fe.order + F.over_integrate;
                const FaceGeo fg = face_geometry(e, f, o, geo, &gout, order);
                coupling_local(F, e, o, fg, basis, bout, blocks);
                // four scatter loops, bilinear-form-assemblers.hh:262-277
                for (int i = 0; i < le; ++i) {
                  for (int j = 0; j < le; ++j)
                    add_to_entry(gi[i], si[j], F.scaling * blocks[0][i * le + j], atomic);
                  for (int j = 0; j < lo; ++j)
                    add_to_entry(gi[i], sj[j], F.scaling * blocks[1][i * lo + j], atomic);
                }
                for (int i = 0; i < lo; ++i) {
                  for (int j = 0; j < le; ++j)
                    add_to_entry(gj[i], si[j], F.scaling * blocks[2][i * le + j], atomic);
                  for (int j = 0; j < lo; ++j)
                    add_to_entry(gj[i], sj[j], F.scaling * blocks[3][i * lo + j], atomic);
                }
              }
            } else if (o == -1 && !bforms.empty()) {
              for (const auto& F : bforms) {
                if (!on_boundary_subset(F.mask, e, f))
                  continue;
                const int order = 0 + space->fe.order + space->fe.order + F.over_integrate;
                const FaceGeo fg = face_geometry(e, f, -1, geo, nullptr, order);
                boundary_local(F, e, fg, basis, local);
                for (int i = 0; i < le; ++i)
                  for (int j = 0; j < le; ++j)
                    add_to_entry(gi[i], si[j], F.scaling * local[i * le + j], atomic);
              }
            }
          }
        } catch (const std::exception& ex) {
#pragma omp critical
          err = ex.what();
        }
      }
    }
    if (!err.empty())
      throw Error(err);
    if (dirichlet)
      apply_dirichlet();
  }

  // DirichletConstraints::apply_local (tools/dirichlet-constraints.hh:85-110) + apply(matrix, vector, only_clear,
  // ensure_symmetry) (:152-184): for every DoF [unit|clear]_col (with ensure_symmetry) then [unit|clear]_row, the
  // vector entry cleared.  Column d ends with 1 at (d, d) at most, which its row keeps unless cleared, so the
  // sequence amounts to: rows of constrained DoFs unit (cleared), with ensure_symmetry their columns zero
  void apply_dirichlet()
  {
    const Mesh& m = *space->mesh;
    const int ls = space->fe.size, nf = m.ref.nfaces;
    std::set<int32_t> dofs;
    if (space->continuous) {
      for (int64_t e = 0; e < m.ne; ++e)
        for (int f = 0; f < nf; ++f) {
          if (m.nb[e * nf + f] != -1 || !on_boundary_subset(dirichlet_faces, e, f))
            continue;
          const auto& fv = m.eref(e).faces[f];
          for (int i = 0; i < space->fe_of(e).size; ++i) {
            bool on_face = true;
            for (int lv : space->fe_of(e).node_subentity[i])
              if (std::find(fv.begin(), fv.end(), lv) == fv.end())
                on_face = false;
            if (on_face)
              dofs.insert(space->indices[e * ls + i]);
          }
        }
    }
    dirichlet_dofs.assign(dofs.begin(), dofs.end());
    std::vector<char> mask(n, 0);
    for (int32_t d : dirichlet_dofs)
      mask[d] = 1;
    for (int64_t i = 0; i < n; ++i)
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
        const int32_t j = cols[k];
        if (mask[i] || (ensure_symmetry && mask[j]))
          values[k] = (i == j && !only_clear) ? 1. : 0.;
      }
    for (int32_t d : dirichlet_dofs)
      rhs[d] = 0.;
  }
};

template <class F>
int guarded(F&& f)
{
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

} // namespace

struct OrMesh : Mesh {};
struct OrSpace : Space {};
struct OrOperator : Operator {};

extern "C" {

const char* or_last_error(void)
{
  return g_error.c_str();
}

OrMesh* or_mesh_structured(int geom, const int32_t* cells, const double* lo, const double* hi)
{
  OrMesh* out = nullptr;
  guarded([&] {
    Mesh* m = structured_mesh(geom, cells, lo, hi);
    out = new OrMesh;
    static_cast<Mesh&>(*out) = std::move(*m);
    delete m;
  });
  return out;
}

OrMesh* or_mesh_create(int geom, int64_t nv, const double* coords, int64_t ne, const int32_t* conn)
{
  OrMesh* out = nullptr;
  guarded([&] {
    auto* m = new OrMesh;
    m->ref = make_ref(geom);
    m->nv = nv, m->ne = ne;
    m->coords.assign(coords, coords + nv * m->ref.dim);
    m->conn.assign(conn, conn + ne * m->ref.nverts);
    for (int64_t k = 0; k < ne * m->ref.nverts; ++k)
      if (m->conn[k] < 0 || m->conn[k] >= nv) {
        delete m;
        throw Error("connectivity references a vertex out of range");
      }
    m->check_affine();
    m->build_neighbors();
    out = m;
  });
  return out;
}

// a mesh of triangles and quadrilaterals, element by element (GridFactory::insertElement with the element's
// geometry type, as make_mixed_grid in test/spaces/base.hh:355-376): element e of type geometry[e] has the
// vertices vertices[offsets[e] .. offsets[e + 1]); one type only gives the plain mesh of that type
OrMesh* or_mesh_mixed(int64_t nv, const double* coords, int64_t ne, const int32_t* geometry, const int64_t* offsets,
                      const int32_t* vertices)
{
  OrMesh* out = nullptr;
  guarded([&] {
    auto* m = new OrMesh;
    try {
      bool present[2] = {false, false};
      for (int64_t e = 0; e < ne; ++e) {
        if ((geometry[e] != OR_QUAD && geometry[e] != OR_TRI) ||
            offsets[e + 1] - offsets[e] != (geometry[e] == OR_QUAD ? 4 : 3))
          throw Error("mixed meshes hold triangles (3 vertices) and quadrilaterals (4 vertices)");
        present[geometry[e]] = true;
      }
      m->nv = nv, m->ne = ne;
      m->coords.assign(coords, coords + nv * 2);
      if (!(present[0] && present[1])) {
        m->ref = make_ref(present[0] ? OR_QUAD : OR_TRI);
        m->conn.assign(vertices, vertices + offsets[ne]);
      } else {
        m->mixed = true;
        m->ref = make_ref(OR_QUAD);
        m->ref.geom = OR_MIXED;
        m->refs[0] = make_ref(OR_QUAD), m->refs[1] = make_ref(OR_TRI);
        m->etype.resize(ne);
        m->conn.assign(ne * 4, -1);
        for (int64_t e = 0; e < ne; ++e) {
          m->etype[e] = int8_t(geometry[e] == OR_QUAD ? 0 : 1);
          for (int64_t t = offsets[e]; t < offsets[e + 1]; ++t)
            m->conn[e * 4 + (t - offsets[e])] = vertices[t];
        }
      }
      for (int64_t t = 0; t < offsets[ne]; ++t)
        if (vertices[t] < 0 || vertices[t] >= nv)
          throw Error("connectivity references a vertex out of range");
      m->check_affine();
      m->build_neighbors();
    } catch (...) {
      delete m;
      throw;
    }
    out = m;
  });
  return out;
}

int or_mesh_is_mixed(const OrMesh* m)
{
  return m->mixed ? 1 : 0;
}

void or_mesh_destroy(OrMesh* m)
{
  delete m;
}
int64_t or_mesh_num_vertices(const OrMesh* m)
{
  return m->nv;
}
int64_t or_mesh_num_elements(const OrMesh* m)
{
  return m->ne;
}
int or_mesh_dim(const OrMesh* m)
{
  return m->ref.dim;
}
int or_mesh_verts_per_element(const OrMesh* m)
{
  return m->ref.nverts;
}
void or_mesh_coords(const OrMesh* m, double* out)
{
  std::memcpy(out, m->coords.data(), m->coords.size() * sizeof(double));
}
void or_mesh_conn(const OrMesh* m, int32_t* out)
{
  std::memcpy(out, m->conn.data(), m->conn.size() * sizeof(int32_t));
}
int or_mesh_faces_per_element(const OrMesh* m)
{
  return m->ref.nfaces;
}
void or_mesh_neighbors(const OrMesh* m, int32_t* out)
{
  std::memcpy(out, m->nb.data(), m->nb.size() * sizeof(int32_t));
}

OrSpace* or_space_create(OrMesh* m, int continuous, int order)
{
  OrSpace* out = nullptr;
  guarded([&] {
    auto* s = new OrSpace;
    try {
      s->init(m, continuous != 0, order);
    } catch (...) {
      delete s;
      throw;
    }
    out = s;
  });
  return out;
}
void or_space_destroy(OrSpace* s)
{
  delete s;
}
int64_t or_space_size(const OrSpace* s)
{
  return s->size;
}
int or_space_local_size(const OrSpace* s)
{
  return s->fe.size;
}
void or_space_global_indices(const OrSpace* s, int32_t* out)
{
  std::memcpy(out, s->indices.data(), s->indices.size() * sizeof(int32_t));
}

OrOperator* or_operator_create(OrSpace* s, int stencil)
{
  OrOperator* out = nullptr;
  guarded([&] {
    auto* op = new OrOperator;
    op->space = op->source = s, op->stencil = stencil;
    op->build_pattern();
    out = op;
  });
  return out;
}
// make_matrix_operator(source_space, range_space, stencil) (operators/matrix-based.hh:200-216): rows from the
// range space, columns from the source space (same mesh and local finite element)
OrOperator* or_operator_create_between(OrSpace* source, OrSpace* range, int stencil)
{
  OrOperator* out = nullptr;
  guarded([&] {
    if (source->mesh != range->mesh || source->fe.size != range->fe.size || source->fe.order != range->fe.order)
      throw Error("source and range spaces need the same mesh and local finite element");
    auto* op = new OrOperator;
    op->space = range, op->source = source, op->stencil = stencil;
    op->build_pattern();
    out = op;
  });
  return out;
}
int64_t or_operator_cols(const OrOperator* op)
{
  return op->source->size;
}
void or_operator_destroy(OrOperator* op)
{
  delete op;
}
int64_t or_operator_nnz(const OrOperator* op)
{
  return op->row_ptr.back();
}
void or_operator_pattern(const OrOperator* op, int64_t* row_ptr, int32_t* cols)
{
  std::memcpy(row_ptr, op->row_ptr.data(), op->row_ptr.size() * sizeof(int64_t));
  std::memcpy(cols, op->cols.data(), op->cols.size() * sizeof(int32_t));
}

int or_operator_append_laplace(OrOperator* op, const double* kappa, double kappa_const, int over_integrate,
                               double scaling)
{
  return guarded([&] {
    ElementForm F{EF_LAPLACE, {}, over_integrate, scaling, {}, {}, 0, {}};
    F.coeff.set(kappa, kappa_const, op->space->mesh->ne);
    op->eforms.push_back(F);
  });
}
int or_operator_append_product(OrOperator* op, const double* weight, double weight_const, int over_integrate,
                               double scaling)
{
  return guarded([&] {
    ElementForm F{EF_PRODUCT, {}, over_integrate, scaling, {}, {}, 0, {}};
    F.coeff.set(weight, weight_const, op->space->mesh->ne);
    op->eforms.push_back(F);
  });
}
// element forms whose coefficient varies inside the elements: values at the points of the rules of order
// coeff_order + 2 k + over_integrate (count: their total, checked)
static int append_element_points(OrOperator* op, int kind, const double* values, int64_t count, int coeff_order,
                                 int over_integrate, double scaling)
{
  return guarded([&] {
    const Mesh& m = *op->space->mesh;
    if (!values || coeff_order < 0 || over_integrate < 0)
      throw Error("invalid coefficient values");
    ElementForm F{kind, {}, over_integrate, scaling, {}, {}, 0, {}};
    F.coeff_order = coeff_order;
    const int order = coeff_order + 2 * op->space->fe.order + over_integrate;
    F.points_off.assign(size_t(m.ne) + 1, 0);
    for (int64_t e = 0; e < m.ne; ++e)
      F.points_off[e + 1] = F.points_off[e] + int64_t(element_rule(m.eref(e), order).w.size());
    if (F.points_off[m.ne] != count)
      throw Error("coefficient values: one per point of the element rules is expected");
    F.points.assign(values, values + count);
    op->eforms.push_back(F);
  });
}
int or_operator_append_laplace_tensor(OrOperator* op, const double* tensor, int per_element, int over_integrate,
                                      double scaling)
{
  return guarded([&] {
    const Mesh& m = *op->space->mesh;
    const int d = m.ref.dim;
    if (!tensor || over_integrate < 0)
      throw Error("invalid diffusion tensor");
    ElementForm F{EF_LAPLACE, {}, over_integrate, scaling, {}, {}, 0, {}};
    F.tensor.assign(tensor, tensor + (per_element ? size_t(m.ne) * d * d : size_t(d * d)));
    op->eforms.push_back(F);
  });
}
int or_operator_append_laplace_points(OrOperator* op, const double* values, int64_t count, int coeff_order,
                                      int over_integrate, double scaling)
{
  return append_element_points(op, EF_LAPLACE, values, count, coeff_order, over_integrate, scaling);
}
int or_operator_append_product_points(OrOperator* op, const double* values, int64_t count, int coeff_order,
                                      int over_integrate, double scaling)
{
  return append_element_points(op, EF_PRODUCT, values, count, coeff_order, over_integrate, scaling);
}
static int append_face(OrOperator* op, int kind, double param, const double* kappa, double kc, const double* omega,
                       double oc, int over_integrate, double scaling)
{
  return guarded([&] {
    FaceForm F{kind, param, {}, {}, over_integrate, scaling, {}, {}, {}};
    F.kappa.set(kappa, kc, op->space->mesh->ne);
    F.omega.set(omega, oc, op->space->mesh->ne);
    if (kind == FF_INNER_COUPLING || kind == FF_INNER_PENALTY)
      op->cforms.push_back(F);
    else {
      F.mask = op->pending_mask;
      op->bforms.push_back(F);
    }
  });
}
int or_operator_append_inner_coupling(OrOperator* op, double s, const double* kappa, double kc, const double* omega,
                                      double oc, int over_integrate, double scaling)
{
  return append_face(op, FF_INNER_COUPLING, s, kappa, kc, omega, oc, over_integrate, scaling);
}
int or_operator_append_inner_penalty(OrOperator* op, double penalty, const double* omega, double oc,
                                     int over_integrate, double scaling)
{
  return append_face(op, FF_INNER_PENALTY, penalty, nullptr, 1., omega, oc, over_integrate, scaling);
}
int or_operator_append_dirichlet_coupling(OrOperator* op, double s, const double* kappa, double kc,
                                          int over_integrate, double scaling)
{
  return append_face(op, FF_DIRICHLET_COUPLING, s, kappa, kc, nullptr, 1., over_integrate, scaling);
}
int or_operator_append_boundary_penalty(OrOperator* op, double penalty, const double* omega, double oc,
                                        int over_integrate, double scaling)
{
  return append_face(op, FF_BOUNDARY_PENALTY, penalty, nullptr, 1., omega, oc, over_integrate, scaling);
}
int or_operator_append_source(OrOperator* op, int kind, const double* f, double f_const, int f_order,
                              int over_integrate)
{
  return guarded([&] {
    if (kind != 0 && kind != 1)
      throw Error("unknown source kind");
    Source S{kind, {}, f_order, over_integrate};
    S.f.set(f, f_const, op->space->mesh->ne);
    op->sources.push_back(S);
  });
}
// matrix-valued diffusion / weight of the last appended inner (boundary = 0) or boundary (1) intersection form:
// d x d per element (NULL: keep the scalar)
// matrix-valued coefficient (d x d per element) of the last appended boundary functional
int or_operator_set_boundary_source_tensor(OrOperator* op, const double* tensor)
{
  return guarded([&] {
    if (op->bsources.empty() || !tensor)
      throw Error("no boundary functional appended");
    const size_t n = size_t(op->space->mesh->ne) * op->space->mesh->ref.dim * op->space->mesh->ref.dim;
    op->bsources.back().coeff_t.assign(tensor, tensor + n);
  });
}
int or_operator_set_face_tensors(OrOperator* op, int boundary, const double* kappa_t, const double* omega_t)
{
  return guarded([&] {
    auto& forms = boundary ? op->bforms : op->cforms;
    if (forms.empty())
      throw Error("no intersection form appended");
    const size_t n = size_t(op->space->mesh->ne) * op->space->mesh->ref.dim * op->space->mesh->ref.dim;
    if (kappa_t)
      forms.back().kappa_t.assign(kappa_t, kappa_t + n);
    if (omega_t)
      forms.back().omega_t.assign(omega_t, omega_t + n);
  });
}
int or_operator_set_dirichlet_mode(OrOperator* op, int only_clear, int ensure_symmetry)
{
  op->only_clear = only_clear != 0;
  op->ensure_symmetry = ensure_symmetry != 0;
  return 0;
}
int or_operator_set_dirichlet(OrOperator* op, int enabled)
{
  if (enabled && op->source != op->space) {
    g_error = "Dirichlet constraints need an operator whose source and range spaces are the same";
    return -1;
  }
  op->dirichlet = enabled != 0;
  return 0;
}
int or_operator_assemble(OrOperator* op, int nthreads)
{
  return guarded([&] { op->assemble(nthreads); });
}
void or_operator_values(const OrOperator* op, double* out)
{
  std::memcpy(out, op->values.data(), op->values.size() * sizeof(double));
}
void or_operator_rhs(const OrOperator* op, double* out)
{
  std::memcpy(out, op->rhs.data(), op->rhs.size() * sizeof(double));
}
int64_t or_operator_num_dirichlet(const OrOperator* op)
{
  return int64_t(op->dirichlet_dofs.size());
}
void or_operator_dirichlet_dofs(const OrOperator* op, int32_t* out)
{
  std::memcpy(out, op->dirichlet_dofs.data(), op->dirichlet_dofs.size() * sizeof(int32_t));
}

// ConstMatrixOperator::apply = matrix_.mv (operators/matrix-based.hh:121-129): row-wise sum in column order
void or_spmv(int64_t n, const int64_t* row_ptr, const int32_t* cols, const double* vals, const double* x, double* y,
             int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
      s += vals[k] * x[cols[k]];
    y[i] = s;
  }
}

static double par_dot(int64_t n, const double* a, const double* b, int nthreads)
{
  double s = 0.;
  if (nthreads <= 1) {
    for (int64_t i = 0; i < n; ++i)
      s += a[i] * b[i];
    return s;
  }
#pragma omp parallel for reduction(+ : s) schedule(static) num_threads(nthreads)
  for (int64_t i = 0; i < n; ++i)
    s += a[i] * b[i];
  return s;
}

// apply_inverse with a Jacobi-preconditioned conjugate gradient method
int or_pcg_jacobi(int64_t n, const int64_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                  double* x, double rtol, int64_t max_iter, int nthreads, int64_t max_iters_run, int64_t* iters,
                  double* relres)
{
  const int nt = nthreads > 0 ? nthreads : 1;
  std::vector<double> dinv(n), r(n), z(n), p(n), q(n);
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
      if (cols[k] == i)
        d = vals[k];
    if (!(d > 0.)) {
      g_error = "Jacobi-PCG needs a positive diagonal (matrix not SPD)";
      return -2;
    }
    dinv[i] = 1. / d;
  }
  or_spmv(n, row_ptr, cols, vals, x, q.data(), nt);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i] - q[i];
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
  }
  const double bb = par_dot(n, b, b, nt);
  double rr = par_dot(n, r.data(), r.data(), nt);
  double rz = par_dot(n, r.data(), z.data(), nt);
  const double tol2 = rtol * rtol * bb;
  int64_t k = 0;
  int status = 1;
  if (rr <= tol2) {
    status = 0;
  } else {
    for (k = 1; k <= max_iter; ++k) {
      or_spmv(n, row_ptr, cols, vals, p.data(), q.data(), nt);
      const double pq = par_dot(n, p.data(), q.data(), nt);
      const double alpha = rz / pq;
#pragma omp parallel for schedule(static) num_threads(nt)
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
      }
      rr = par_dot(n, r.data(), r.data(), nt);
      if (rr <= tol2) {
        status = 0;
        break;
      }
      if (max_iters_run > 0 && k >= max_iters_run)
        break;
#pragma omp parallel for schedule(static) num_threads(nt)
      for (int64_t i = 0; i < n; ++i)
        z[i] = dinv[i] * r[i];
      const double rz_new = par_dot(n, r.data(), z.data(), nt);
      const double beta = rz_new / rz;
      rz = rz_new;
#pragma omp parallel for schedule(static) num_threads(nt)
      for (int64_t i = 0; i < n; ++i)
        p[i] = z[i] + beta * p[i];
    }
    if (k > max_iter)
      k = max_iter;
  }
  *iters = k;
  *relres = bb > 0. ? std::sqrt(rr / bb) : 0.;
  if (status != 0 && !(max_iters_run > 0 && k >= max_iters_run))
    g_error = "Jacobi-PCG did not converge";
  return status;
}

void or_laplace_evaluate(int d, int r, int rows, int cols, const double* weight, const double* test_grads,
                         const double* ansatz_grads, double* result)
{
  // local/integrands/laplace.hh:98-101
  for (int ii = 0; ii < rows; ++ii)
    for (int jj = 0; jj < cols; ++jj) {
      double acc = 0.;
      for (int rr = 0; rr < r; ++rr) {
        const double* ga = &ansatz_grads[(jj * r + rr) * d];
        const double* gt = &test_grads[(ii * r + rr) * d];
        double wg[3];
        for (int a = 0; a < d; ++a) {
          double s = 0.;
          for (int c = 0; c < d; ++c)
            s += weight[a * d + c] * ga[c];
          wg[a] = s;
        }
        acc += dot(wg, gt, d);
      }
      result[ii * cols + jj] = acc;
    }
}

void or_product_evaluate(int r, int rows, int cols, const double* weight, const double* test_values,
                         const double* ansatz_values, double* result)
{
  // local/integrands/product.hh:126-128: (weight * test) * ansatz
  for (int ii = 0; ii < rows; ++ii)
    for (int jj = 0; jj < cols; ++jj) {
      double wt[8];
      for (int a = 0; a < r; ++a) {
        double s = 0.;
        for (int c = 0; c < r; ++c)
          s += weight[a * r + c] * test_values[ii * r + c];
        wt[a] = s;
      }
      result[ii * cols + jj] = dot(wt, &ansatz_values[jj * r], r);
    }
}

int or_quadrature(int geom, int order, double* points, double* weights, int capacity)
{
  int count = -1;
  guarded([&] {
    const RefElement ref = make_ref(geom);
    const Quadrature q = element_rule(ref, order);
    count = int(q.w.size());
    if (count > capacity)
      return;
    for (int i = 0; i < count; ++i) {
      for (int a = 0; a < ref.dim; ++a)
        points[i * ref.dim + a] = q.x[i][a];
      weights[i] = q.w[i];
    }
  });
  return count;
}

double or_esv2007_h1_semi_error(const OrSpace* s, const double* u, int quad_order)
{
  if (s->mesh->mixed)
    return std::nan("");
  const Mesh& m = *s->mesh;
  const int ls = s->fe.size, d = m.ref.dim;
  const Quadrature q = element_rule(m.ref, quad_order);
  double total = 0.;
  BoundBasis basis;
  for (int64_t e = 0; e < m.ne; ++e) {
    const Geometry geo = m.geometry(e);
    basis.bind(&s->fe, &geo);
    double acc = 0.;
    for (size_t qp = 0; qp < q.w.size(); ++qp) {
      basis.jacobians(q.x[qp].data());
      double X[3];
      geo.global(q.x[qp].data(), X);
      double gu[2] = {-M_PI_2 * std::sin(M_PI_2 * X[0]) * std::cos(M_PI_2 * X[1]),
                      -M_PI_2 * std::cos(M_PI_2 * X[0]) * std::sin(M_PI_2 * X[1])};
      for (int i = 0; i < ls; ++i)
        for (int a = 0; a < d; ++a)
          gu[a] -= u[s->indices[e * ls + i]] * basis.grads[3 * i + a];
      acc += (gu[0] * gu[0] + gu[1] * gu[1]) * geo.integration_element * q.w[qp];
    }
    total += acc;
  }
  return std::sqrt(total);
}

// Squared error norms of e = u_h - g: BilinearForm(grid_view, e, e) with
// LocalElementIntegralBilinearForm(LocalElementProductIntegrand()) and with
// LocalElementIntegralBilinearForm(LocalLaplaceIntegrand(kappa)), apply2()
// (operators/bilinear-form.hh:344-452, local/bilinear-forms/integrals.hh:97-134,
// as used in examples/stationary-heat-equation.cc:120-132).  g is given by its values
// [ne][npts] and physical gradients [ne][npts][dim] at the points of element_rule(order).
int or_operator_append_boundary_source(OrOperator* op, int kind, double param, const double* coeff, double coeff_const,
                                      const double* g, double g_const, int g_order, int over_integrate)
{
  return guarded([&] {
    if (kind < BS_DIRICHLET_COUPLING || kind > BS_PRODUCT)
      throw Error("unknown boundary functional kind");
    BoundarySource S;
    S.kind = kind, S.param = param, S.g_const = g_const, S.g_order = g_order, S.over_integrate = over_integrate;
    const Mesh& m = *op->space->mesh;
    if (coeff)
      S.coeff.values.assign(coeff, coeff + m.ne);
    else
      S.coeff.constant = coeff_const;
    op->number_boundary_faces();
    if (g) {
      int64_t nb = 0;
      for (int64_t v : op->bface_index)
        nb += (v >= 0);
      const int order = op->boundary_source_order(S);
      const int64_t nq = int64_t((m.ref.dim == 3 ? (m.ref.simplex ? simplex_rule(2, order) : cube_rule(2, order))
                                                 : line_rule(order)).w.size());
      S.g.assign(g, g + nb * nq);
    }
    S.mask = op->pending_mask;
    op->bsources.push_back(S);
  });
}

static void boundary_mask(OrOperator* op, const uint8_t* mask, int64_t count, std::vector<uint8_t>& out)
{
  op->number_boundary_faces();
  int64_t nb = 0;
  for (int64_t v : op->bface_index)
    nb += (v >= 0);
  if (!mask) {
    out.clear();
    return;
  }
  if (count != nb)
    throw Error("boundary mask length differs from the number of boundary intersections");
  out.assign(mask, mask + count);
}
// ApplyOn::CustomBoundaryIntersections for the boundary forms and functionals appended from now on: one byte
// per boundary intersection in walk order (nonzero: applied); NULL: every boundary intersection
int or_operator_set_boundary_filter(OrOperator* op, const uint8_t* mask, int64_t count)
{
  return guarded([&] { boundary_mask(op, mask, count, op->pending_mask); });
}
// the Dirichlet intersections of the constraints (boundary_info.type == DirichletBoundary); NULL: all
int or_operator_set_dirichlet_faces(OrOperator* op, const uint8_t* mask, int64_t count)
{
  return guarded([&] { boundary_mask(op, mask, count, op->dirichlet_faces); });
}

// physical coordinates of the face quadrature points of the boundary intersections in walk order,
// out[nb][npts][dim] (NULL: size query); returns npts (2d grids)
int or_mesh_boundary_quadrature_points(const OrMesh* mesh, int order, double* out)
{
  int count = -1;
  guarded([&] {
    const Mesh& m = *mesh;
    if (m.ref.dim == 3) {
      // the points of Operator::face_geometry_3d, in walk order
      const Quadrature q = m.ref.simplex ? simplex_rule(2, order) : cube_rule(2, order);
      count = int(q.w.size());
      if (!out)
        return;
      const int nf = m.ref.nfaces;
      int64_t bi = 0;
      for (int64_t e = 0; e < m.ne; ++e) {
        const Geometry geo = m.geometry(e);
        for (int f = 0; f < nf; ++f) {
          if (m.nb[e * nf + f] >= 0)
            continue;
          const auto& fv = m.ref.faces[f];
          for (int qp = 0; qp < count; ++qp) {
            const double s = q.x[qp][0], t = q.x[qp][1];
            double xh[3], X[3];
            for (int c = 0; c < 3; ++c)
              xh[c] = m.ref.simplex
                          ? (1. - s - t) * m.ref.verts[fv[0]][c] + s * m.ref.verts[fv[1]][c] + t * m.ref.verts[fv[2]][c]
                          : (1. - s) * (1. - t) * m.ref.verts[fv[0]][c] + s * (1. - t) * m.ref.verts[fv[1]][c] +
                                (1. - s) * t * m.ref.verts[fv[2]][c] + s * t * m.ref.verts[fv[3]][c];
            geo.global(xh, X);
            for (int c = 0; c < 3; ++c)
              out[(bi * count + qp) * 3 + c] = X[c];
          }
          ++bi;
        }
      }
      return;
    }
    if (m.ref.dim != 2)
      throw Error("boundary quadrature points are implemented for 2d and 3d grids");
    const Quadrature q = line_rule(order);
    count = int(q.w.size());
    if (!out)
      return;
    const int nf = m.ref.nfaces;
    int64_t bi = 0;
    for (int64_t e = 0; e < m.ne; ++e) {
      const Geometry geo = m.geometry(e);
      const RefElement& r = m.eref(e);
      for (int f = 0; f < nf; ++f) {
        if (m.nb[e * nf + f] != -1)
          continue;
        const int a = r.faces[f][0], b = r.faces[f][1];
        for (int qp = 0; qp < count; ++qp) {
          const double t = q.x[qp][0];
          double xh[3] = {0, 0, 0}, X[3];
          for (int c = 0; c < 2; ++c)
            xh[c] = (1. - t) * r.verts[a][c] + t * r.verts[b][c];
          geo.global(xh, X);
          out[(bi * count + qp) * 2 + 0] = X[0];
          out[(bi * count + qp) * 2 + 1] = X[1];
        }
        ++bi;
      }
    }
  });
  return count;
}

int or_error_norms(const OrSpace* s, const double* u, int order, const double* gval, const double* ggrad,
                   const double* kappa, double kappa_const, double* out)
{
  return guarded([&] {
    const Mesh& m = *s->mesh;
    const int ls = s->fe.size, d = m.ref.dim;
    double l2 = 0., h1 = 0.;
    BoundBasis basis;
    int64_t q0 = 0; // the element's first point in gval / ggrad (every element with its own type's rule)
    for (int64_t e = 0; e < m.ne; ++e) {
      const Quadrature q = element_rule(m.eref(e), order);
      const int nq = int(q.w.size());
      const Geometry geo = m.geometry(e);
      basis.bind(&s->fe_of(e), &geo);
      const double k = kappa ? kappa[e] : kappa_const;
      double el2 = 0., eh1 = 0.;
      for (int qp = 0; qp < nq; ++qp) {
        basis.evaluate(q.x[qp].data());
        basis.jacobians(q.x[qp].data());
        double ev = 0., eg[3] = {0., 0., 0.};
        for (int i = 0; i < s->fe_of(e).size; ++i) {
          const double ui = u[s->indices[e * ls + i]];
          ev += ui * basis.vals[i];
          for (int a = 0; a < d; ++a)
            eg[a] += ui * basis.grads[3 * i + a];
        }
        if (gval)
          ev -= gval[q0 + qp];
        if (ggrad)
          for (int a = 0; a < d; ++a)
            eg[a] -= ggrad[(q0 + qp) * d + a];
        const double f = geo.integration_element * q.w[qp];
        el2 += (ev * ev) * f;
        eh1 += kgrad_dot(k, eg, eg, d) * f;
      }
      l2 += el2, h1 += eh1;
      q0 += nq;
    }
    out[0] = l2, out[1] = h1;
  });
}

// physical coordinates of the points of element_rule(order) on every element, out[ne][npts][dim]; on a mixed
// mesh (returns 0) every element's points follow the previous element's, with its own type's rule
int or_mesh_quadrature_points(const OrMesh* mesh, int order, double* out)
{
  int count = -1;
  guarded([&] {
    const Mesh& m = *mesh;
    count = m.mixed ? 0 : int(element_rule(m.ref, order).w.size());
    if (!out)
      return;
    int64_t q0 = 0;
    for (int64_t e = 0; e < m.ne; ++e) {
      const Quadrature q = element_rule(m.eref(e), order);
      const Geometry geo = m.geometry(e);
      for (size_t qp = 0; qp < q.w.size(); ++qp)
        geo.global(q.x[qp].data(), &out[(q0 + int64_t(qp)) * m.ref.dim]);
      q0 += int64_t(q.w.size());
    }
  });
  return count;
}

// where every element's points start in that layout: out[ne + 1]
int or_mesh_quadrature_offsets(const OrMesh* mesh, int order, int64_t* out)
{
  return guarded([&] {
    const Mesh& m = *mesh;
    out[0] = 0;
    for (int64_t e = 0; e < m.ne; ++e)
      out[e + 1] = out[e] + int64_t(element_rule(m.eref(e), order).w.size());
  });
}

int or_adjoint_log_kappa_gradient(const OrSpace* s, const double* kappa, const double* u, const double* lambda,
                                  double* grad)
{
  return guarded([&] {
    Operator op;
    op.space = const_cast<OrSpace*>(s);
    const Mesh& m = *s->mesh;
    const int ls = s->fe.size; // the stride of the DoF indices (the largest local size on a mixed mesh)
    ElementForm F{EF_LAPLACE, {}, 0, 1., {}, {}, 0, {}};
    F.coeff.constant = 1.;
#pragma omp parallel
    {
      BoundBasis basis;
      std::vector<double> local;
#pragma omp for schedule(static)
      for (int64_t e = 0; e < m.ne; ++e) {
        const Geometry geo = m.geometry(e);
        basis.bind(&s->fe_of(e), &geo);
        op.element_local(F, e, geo, basis, local);
        const int32_t* gi = &s->indices[e * ls];
        const int le = s->fe_of(e).size;
        double acc = 0.;
        for (int i = 0; i < le; ++i) {
          double row = 0.;
          for (int j = 0; j < le; ++j)
            row += local[i * le + j] * u[gi[j]];
          acc += lambda[gi[i]] * row;
        }
        grad[e] = -kappa[e] * acc;
      }
    }
  });
}

} // extern "C"
