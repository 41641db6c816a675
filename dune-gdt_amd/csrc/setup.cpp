// setup.cpp -- host-side set-up of libhgdt.so: reference elements, quadrature,
// grids, DoF maps, sparsity patterns, fixed-order scatter maps, Dirichlet DoFs.
//
// Integer/topological work only; every floating-point pass over the problem
// data runs in the HIP kernels of assembly.hip and solver.hip.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>

#include "hgdt_internal.hpp"

namespace hgdt {

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& body)
{
  if (n <= 0)
    return;
  int nt = int(std::thread::hardware_concurrency());
  nt = std::max(1, std::min(nt, 16));
  if (n < 4096 || nt == 1) {
    body(0, n);
    return;
  }
  std::vector<std::thread> pool;
  std::vector<std::exception_ptr> errs(nt);
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e)
      break;
    pool.emplace_back([&, t, b, e] {
      try {
        body(b, e);
      } catch (...) {
        errs[t] = std::current_exception();
      }
    });
  }
  for (auto& th : pool)
    th.join();
  for (auto& ep : errs)
    if (ep)
      std::rethrow_exception(ep);
}

// DUNE reference elements: vertex and face numbering as in dune-geometry
RefElement reference_element(int geom)
{
  RefElement r;
  r.geom = geom;
  auto setf = [&](std::initializer_list<std::initializer_list<int>> fl) {
    int f = 0;
    for (auto& face : fl) {
      int k = 0;
      for (int v : face)
        r.faces[f][k++] = v;
      r.face_nverts = k;
      ++f;
    }
    r.nfaces = f;
  };
  switch (geom) {
    case HGDT_QUAD:
      r.dim = 2, r.nverts = 4;
      for (int b = 0; b < 4; ++b)
        r.verts[b][0] = b & 1, r.verts[b][1] = (b >> 1) & 1;
      setf({{0, 2}, {1, 3}, {0, 1}, {2, 3}});
      r.volume = 1.;
      break;
    case HGDT_TRI:
      r.dim = 2, r.nverts = 3, r.simplex = true;
      r.verts[1][0] = 1., r.verts[2][1] = 1.;
      setf({{0, 1}, {0, 2}, {1, 2}});
      r.volume = 0.5;
      break;
    case HGDT_HEX:
      r.dim = 3, r.nverts = 8;
      for (int b = 0; b < 8; ++b)
        r.verts[b][0] = b & 1, r.verts[b][1] = (b >> 1) & 1, r.verts[b][2] = (b >> 2) & 1;
      setf({{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}});
      r.volume = 1.;
      break;
    case HGDT_TET:
      r.dim = 3, r.nverts = 4, r.simplex = true;
      r.verts[1][0] = 1., r.verts[2][1] = 1., r.verts[3][2] = 1.;
      setf({{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}});
      r.volume = 1. / 6.;
      break;
    default:
      HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "unknown geometry type " + std::to_string(geom));
  }
  return r;
}

std::vector<std::vector<int>> lagrange_node_vertices(const RefElement& ref, int order)
{
  std::vector<std::vector<int>> nodes;
  const int d = ref.dim;
  if (!ref.simplex) {
    const int n1 = order + 1;
    const int count = (d == 2) ? n1 * n1 : n1 * n1 * n1;
    for (int t = 0; t < count; ++t) {
      const int idx[3] = {t % n1, (t / n1) % n1, t / (n1 * n1)};
      std::vector<int> sub;
      for (int b = 0; b < ref.nverts; ++b) {
        bool ok = true;
        for (int a = 0; a < d; ++a) {
          const int bit = (b >> a) & 1;
          if ((idx[a] == 0 && bit != 0) || (idx[a] == order && bit != 1))
            ok = false;
        }
        if (ok)
          sub.push_back(b);
      }
      nodes.push_back(sub);
    }
  } else {
    for (int v = 0; v < ref.nverts; ++v)
      nodes.push_back({v});
    if (order == 2)
      for (int a = 0; a < ref.nverts; ++a)
        for (int b = a + 1; b < ref.nverts; ++b)
          nodes.push_back({a, b});
  }
  return nodes;
}

// Gauss-Legendre points on [0,1] (Newton on the three-term recurrence), ascending
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w)
{
  x.assign(n, 0.), w.assign(n, 0.);
  auto legendre = [n](double t, double& pn, double& dpn) {
    double p0 = 1., p1 = t;
    if (n == 0) {
      pn = 1., dpn = 0.;
      return;
    }
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2. * k - 1.) * t * p1 - (k - 1.) * p0) / k;
      p0 = p1, p1 = p2;
    }
    pn = p1;
    const double pnm1 = (n == 1) ? 1. : p0;
    dpn = n * (t * pn - pnm1) / (t * t - 1.);
  };
  for (int i = 0; i < n; ++i) {
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pn, dpn;
    for (int it = 0; it < 100; ++it) {
      legendre(t, pn, dpn);
      const double dt = pn / dpn;
      t -= dt;
      if (std::fabs(dt) < 1e-17)
        break;
    }
    legendre(t, pn, dpn);
    // t_i descending in i -> store ascending on [0,1]
    x[n - 1 - i] = 0.5 * (1. + t);
    w[n - 1 - i] = 1. / ((1. - t * t) * dpn * dpn);
  }
  for (int i = 0; i < n / 2; ++i) { // symmetrise against round-off
    const double xs = 0.5 * ((1. - x[n - 1 - i]) + x[i]);
    x[i] = xs, x[n - 1 - i] = 1. - xs;
    const double ws = 0.5 * (w[i] + w[n - 1 - i]);
    w[i] = ws, w[n - 1 - i] = ws;
  }
}

Quadrature line_rule(int order)
{
  Quadrature q;
  q.dim = 1;
  gauss_legendre01(order / 2 + 1, q.x, q.w);
  return q;
}

// quadrature of the faces of `ref` (integrals.hh:226-233,352-356 use the rule of the intersection's
// geometry type): Gauss-Legendre on [0, 1] in 2d, the triangle or square rule of element_rule in 3d
Quadrature face_rule(const RefElement& ref, int order)
{
  if (ref.dim == 2)
    return line_rule(order);
  return element_rule(reference_element(ref.simplex ? HGDT_TRI : HGDT_QUAD), order);
}

Quadrature element_rule(const RefElement& ref, int order)
{
  Quadrature q;
  q.dim = ref.dim;
  const int d = ref.dim;
  if (!ref.simplex) {
    std::vector<double> x, w;
    gauss_legendre01(order / 2 + 1, x, w);
    const int n = int(x.size());
    const int total = (d == 2) ? n * n : n * n * n;
    for (int t = 0; t < total; ++t) {
      const int idx[3] = {t % n, (t / n) % n, t / (n * n)};
      double wt = 1.;
      for (int a = 0; a < d; ++a)
        q.x.push_back(x[idx[a]]), wt *= w[idx[a]];
      q.w.push_back(wt);
    }
    return q;
  }
  if (order <= 1) {
    for (int a = 0; a < d; ++a)
      q.x.push_back(1. / (d + 1));
    q.w.push_back(ref.volume);
    return q;
  }
  if (order == 2) {
    if (d == 2) {
      const double a = 1. / 6., b = 2. / 3.;
      q.x = {a, a, b, a, a, b};
      q.w = {1. / 6., 1. / 6., 1. / 6.};
    } else {
      const double a = (5. + 3. * std::sqrt(5.)) / 20., b = (5. - std::sqrt(5.)) / 20.;
      q.x = {b, b, b, a, b, b, b, a, b, b, b, a};
      q.w = {1. / 24., 1. / 24., 1. / 24., 1. / 24.};
    }
    return q;
  }
  // collapsed (conical product) Gauss rule, exact to the requested order
  std::vector<double> xu, wu, xv, wv, xw, ww;
  if (d == 2) {
    gauss_legendre01((order + 3) / 2, xu, wu);
    gauss_legendre01((order + 2) / 2, xv, wv);
    for (size_t i = 0; i < xu.size(); ++i)
      for (size_t j = 0; j < xv.size(); ++j) {
        q.x.push_back(xu[i]), q.x.push_back((1. - xu[i]) * xv[j]);
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
          q.x.push_back(u), q.x.push_back((1. - u) * v), q.x.push_back((1. - u) * (1. - v) * xw[k]);
          q.w.push_back(wu[i] * wv[j] * ww[k] * (1. - u) * (1. - u) * (1. - v));
        }
  }
  return q;
}

std::vector<int64_t> element_point_offsets(const Grid& g, int order)
{
  std::vector<int64_t> off(g.ne + 1, 0);
  if (!g.mixed) {
    const int64_t nq = element_rule(g.ref, order).size();
    for (int64_t e = 0; e <= g.ne; ++e)
      off[e] = e * nq;
    return off;
  }
  std::vector<int64_t> nq;
  for (const GridPart& P : g.parts)
    nq.push_back(element_rule(P.ref, order).size());
  for (int64_t e = 0; e < g.ne; ++e)
    off[e + 1] = off[e] + nq[g.etype[e]];
  return off;
}

void Grid::check_affine() const
{
  if (ref.simplex)
    return;
  const int d = ref.dim, k = ref.nverts;
  std::atomic<bool> bad{false};
  parallel_for(ne, [&](int64_t b, int64_t e) {
    for (int64_t el = b; el < e && !bad; ++el) {
      if (eref(el).simplex)
        continue;
      const int32_t* c = &conn[el * k];
      for (int v = 0; v < k; ++v)
        for (int a = 0; a < d; ++a) {
          double expect = coords[int64_t(c[0]) * d + a];
          for (int r = 0; r < d; ++r)
            if ((v >> r) & 1)
              expect += coords[int64_t(c[1 << r]) * d + a] - coords[int64_t(c[0]) * d + a];
          const double got = coords[int64_t(c[v]) * d + a];
          if (std::fabs(expect - got) > 1e-12 * (1. + std::fabs(got)))
            bad = true;
        }
    }
  });
  if (bad)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "cube elements must be affine (parallelograms / parallelepipeds)");
}

// Intersections by face matching: faces are bucketed by their smallest vertex,
// then matched inside each bucket (O(#faces) for bounded vertex valence).
void Grid::finalize(bool neighbours_known)
{
  if (neighbours_known) {
    count_faces_and_upload();
    return;
  }
  // on a mixed grid the faces past an element's own count are absent: neighbour -2, never matched
  const int nf = ref.nfaces, fk = ref.face_nverts, k = ref.nverts;
  const int64_t F = ne * nf;
  nb.assign(F, -1);
  nbface.assign(F, -1);
  auto present = [&](int64_t ef) { return !mixed || int(ef % nf) < eref(ef / nf).nfaces; };
  std::vector<int64_t> start(nv + 1, 0);
  auto face_key = [&](int64_t ef, int32_t* key) {
    const int64_t e = ef / nf;
    const int f = int(ef % nf);
    const RefElement& r = eref(e);
    for (int a = 0; a < fk; ++a)
      key[a] = conn[e * k + r.faces[f][a]];
    std::sort(key, key + fk);
  };
  int64_t nfaces = 0;
  for (int64_t ef = 0; ef < F; ++ef) {
    if (!present(ef)) {
      nb[ef] = -2;
      continue;
    }
    int32_t key[4];
    face_key(ef, key);
    ++start[key[0] + 1];
    ++nfaces;
  }
  for (int64_t v = 0; v < nv; ++v)
    start[v + 1] += start[v];
  std::vector<int64_t> bucket(std::max<int64_t>(nfaces, 1)), fill(start.begin(), start.end() - 1);
  for (int64_t ef = 0; ef < F; ++ef) {
    if (!present(ef))
      continue;
    int32_t key[4];
    face_key(ef, key);
    bucket[fill[key[0]]++] = ef;
  }
  std::atomic<bool> nonconforming{false};
  parallel_for(nv, [&](int64_t b, int64_t e) {
    std::vector<std::array<int32_t, 4>> keys;
    for (int64_t v = b; v < e; ++v) {
      const int64_t s0 = start[v], s1 = start[v + 1];
      keys.resize(s1 - s0);
      for (int64_t i = s0; i < s1; ++i) {
        int32_t key[4] = {0, 0, 0, 0};
        face_key(bucket[i], key);
        std::copy(key, key + 4, keys[i - s0].begin());
      }
      for (int64_t i = s0; i < s1; ++i)
        for (int64_t j = i + 1; j < s1; ++j)
          if (keys[i - s0] == keys[j - s0]) {
            const int64_t a = bucket[i], c = bucket[j];
            if (nb[a] != -1 || nb[c] != -1)
              nonconforming = true;
            nb[a] = int32_t(c / nf), nbface[a] = int8_t(c % nf);
            nb[c] = int32_t(a / nf), nbface[c] = int8_t(a % nf);
          }
    }
  });
  if (nonconforming)
    HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "grid is not conforming: a face is shared by more than two elements");
  count_faces_and_upload();
}

void Grid::count_faces_and_upload()
{
  const int nf = ref.nfaces;
  const int64_t F = ne * nf;
  n_inner = 0, n_boundary = 0;
  for (int64_t ef = 0; ef < F; ++ef) {
    if (nb[ef] == -1)
      ++n_boundary;
    else if (nb[ef] > ef / nf)
      ++n_inner;
  }
  d_coords.upload(coords);
  d_conn.upload(conn);
  for (GridPart& p : parts)
    p.d_eid.upload(p.eid), p.d_conn.upload(p.conn);
  if (mixed)
    d_etype.upload(etype);
}

// ----------------------------------------------------------------------------
// sparsity pattern: make_element_sparsity_pattern / make_element_and_intersection_sparsity_pattern
// (dune/gdt/tools/sparsity-pattern.hh:33-50, 117-144) computed row by row from the
// DoF -> element adjacency, rows sorted
// ----------------------------------------------------------------------------
struct Adjacency {
  std::vector<int64_t> ptr;   // n + 1
  std::vector<int32_t> elem;  // element of each (dof, element) incidence, ascending per dof
  std::vector<int8_t> local;  // local index of the dof in that element
};

static Adjacency dof_adjacency(const Space& s)
{
  const int64_t ne = s.grid->ne, n = s.size;
  const int ls = s.local_size;
  Adjacency A;
  A.ptr.assign(n + 1, 0);
  for (int64_t i = 0; i < ne * ls; ++i)
    if (s.dofs[i] >= 0) // (mixed grids: -1 past an element's own DoFs)
      ++A.ptr[s.dofs[i] + 1];
  for (int64_t i = 0; i < n; ++i)
    A.ptr[i + 1] += A.ptr[i];
  A.elem.resize(A.ptr[n]), A.local.resize(A.ptr[n]);
  std::vector<int64_t> fill(A.ptr.begin(), A.ptr.end() - 1);
  for (int64_t e = 0; e < ne; ++e)
    for (int i = 0; i < s.elem_size(e); ++i) {
      const int64_t p = fill[s.dofs[e * ls + i]]++;
      A.elem[p] = int32_t(e), A.local[p] = int8_t(i);
    }
  return A;
}

// sparsity pattern of the stencil (make_element_sparsity_pattern and friends, tools/sparsity-pattern.hh):
// rows are the range space's DoFs, the columns of row i the source space's DoFs of the elements holding i
// (and, with an intersection stencil, of their neighbours).  On a caller's matrix the caller's pattern is
// the operator's, and it must hold every entry of the stencil's.
void Operator::build_pattern()
{
  const Space& s = *space;
  const Space& src = *source;
  const Grid& g = *s.grid;
  const int ls = s.local_size, nf = g.ref.nfaces;
  n = s.size, ncols = src.size;
  const Adjacency A = dof_adjacency(s);
  std::vector<int32_t> counts(n);
  auto row_columns = [&](int64_t i, std::vector<int32_t>& c) {
    c.clear();
    for (int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
      const int64_t e = A.elem[p];
      if (stencil != HGDT_STENCIL_INTERSECTION)
        for (int j = 0; j < src.elem_size(e); ++j)
          c.push_back(src.dofs[e * ls + j]);
      if (stencil != HGDT_STENCIL_ELEMENT)
        for (int f = 0; f < nf; ++f) {
          const int64_t o = g.nb[e * nf + f];
          if (o >= 0)
            for (int j = 0; j < src.elem_size(o); ++j)
              c.push_back(src.dofs[o * ls + j]);
        }
    }
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
  };
  if (borrowed) {
    const std::vector<int64_t>& rp = borrowed_row_ptr;
    const std::vector<int32_t>& cc = borrowed_cols;
    if (int64_t(rp.size()) != n + 1 || rp[0] != 0)
      HGDT_THROW(HGDT_ERR_SHAPES_DO_NOT_MATCH, "the matrix needs one row per DoF of the range space");
    std::atomic<int> bad{0}; // 1: row pointer / columns malformed, 2: a stencil entry is missing
    parallel_for(n, [&](int64_t b, int64_t e) {
      std::vector<int32_t> c;
      for (int64_t i = b; i < e && !bad; ++i) {
        const int64_t r0 = rp[i], r1 = rp[i + 1];
        if (r1 < r0 || r1 > int64_t(cc.size())) {
          bad = 1;
          break;
        }
        for (int64_t t = r0; t < r1; ++t)
          if (cc[t] < 0 || cc[t] >= ncols || (t > r0 && cc[t] <= cc[t - 1]))
            bad = 1;
        row_columns(i, c);
        for (int32_t col : c)
          if (!std::binary_search(cc.begin() + r0, cc.begin() + r1, col))
            bad = 2;
      }
    });
    if (bad == 1 || rp[n] != int64_t(cc.size()))
      HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT,
                 "the matrix pattern needs ascending row pointers and ascending, unique, valid columns per row");
    if (bad == 2)
      HGDT_THROW(HGDT_ERR_SHAPES_DO_NOT_MATCH, "the matrix pattern does not hold every entry of the stencil");
    row_ptr = rp, cols = cc;
    nnz = row_ptr[n];
  } else {
    row_ptr.assign(n + 1, 0);
    parallel_for(n, [&](int64_t b, int64_t e) {
      std::vector<int32_t> c;
      for (int64_t i = b; i < e; ++i) {
        row_columns(i, c);
        counts[i] = int32_t(c.size());
      }
    });
    for (int64_t i = 0; i < n; ++i)
      row_ptr[i + 1] = row_ptr[i] + counts[i];
    nnz = row_ptr[n];
    cols.resize(nnz);
    parallel_for(n, [&](int64_t b, int64_t e) {
      std::vector<int32_t> c;
      for (int64_t i = b; i < e; ++i) {
        row_columns(i, c);
        std::copy(c.begin(), c.end(), cols.begin() + row_ptr[i]);
      }
    });
  }
  d_row_ptr.upload(row_ptr);
  d_cols.upload(cols);
  if (!borrowed) {
    d_values.allocate(std::max<int64_t>(nnz, 1));
    HGDT_HIP(hipMemset(d_values.ptr, 0, nnz * sizeof(double)));
  }
  d_rhs.allocate(std::max<int64_t>(n, 1));
  HGDT_HIP(hipMemset(d_rhs.ptr, 0, n * sizeof(double)));
}

// ----------------------------------------------------------------------------
// fixed-order scatter maps.  The reference walk (XT::Grid::Walker driven by
// MatrixOperator::assemble, matrix-based.hh:496-500) visits element e, applies
// the element assemblers (bilinear-form-assemblers.hh:110-128), then the
// intersections of e in local order: inner ones once, from the element with the
// smaller index (ApplyOn::InnerIntersectionsOnce), with four block scatters
// (:262-277), boundary ones with one block (:392-395).  Every matrix slot gets
// the list of its contributions in exactly that order.
// ----------------------------------------------------------------------------
void Operator::build_scatter_maps()
{
  const Space& s = *space;   // rows (test functions)
  const Space& src = *source; // columns (ansatz functions)
  const Grid& g = *s.grid;
  const int ls = s.local_size, nf = g.ref.nfaces;
  const int64_t ne = g.ne, k2 = int64_t(ls) * ls;
  ScatterMaps& M = maps;
  const bool has_e = !eforms.empty(), has_c = !cforms.empty(), has_b = !bforms.empty();
  const bool has_s = !sources.empty(), has_bs = !bsources.empty();
  const int signature = (has_e ? 1 : 0) | (has_c ? 2 : 0) | (has_b ? 4 : 0) | (has_s ? 8 : 0) | (has_bs ? 16 : 0);
  if (M.built && M.signature == signature)
    return;
  // intersections in walk order
  std::vector<int32_t> inner_index(ne * nf, -1), bnd_index(ne * nf, -1);
  M.inner_in.clear(), M.inner_fin.clear(), M.inner_out.clear(), M.inner_fout.clear();
  M.bnd_e.clear(), M.bnd_f.clear();
  for (int64_t e = 0; e < ne; ++e)
    for (int f = 0; f < nf; ++f) {
      const int64_t o = g.nb[e * nf + f];
      if (o > e) {
        inner_index[e * nf + f] = int32_t(M.inner_in.size());
        M.inner_in.push_back(int32_t(e)), M.inner_fin.push_back(f);
        M.inner_out.push_back(int32_t(o)), M.inner_fout.push_back(g.nbface[e * nf + f]);
      } else if (o == -1) {
        bnd_index[e * nf + f] = int32_t(M.bnd_e.size());
        M.bnd_e.push_back(int32_t(e)), M.bnd_f.push_back(f);
      }
    }
  const int64_t nin = int64_t(M.inner_in.size()), nbd = int64_t(M.bnd_e.size());
  // mixed grids: the element matrices (and right-hand sides) part after part, each entry-major over the
  // part's elements with the part's local size (the intersection forms need one element type)
  const int np = std::max<int>(1, int(s.parts.size()));
  auto part_ls = [&](int p) { return s.parts.empty() ? ls : s.parts[p].local_size; };
  auto part_ne = [&](int p) { return s.parts.empty() ? ne : g.parts[p].ne(); };
  M.part_elem_base.assign(np, 0), M.part_rhs_base.assign(np, 0);
  int64_t elem_entries = 0, rhs_entries = 0;
  for (int p = 0; p < np; ++p) {
    M.part_elem_base[p] = elem_entries, M.part_rhs_base[p] = rhs_entries;
    elem_entries += part_ne(p) * (int64_t(part_ls(p)) * (part_ls(p) + 1) / 2);
    rhs_entries += part_ne(p) * part_ls(p);
  }
  M.elem_base = 0;
  M.inner_base = has_e ? elem_entries : 0;
  M.bnd_base = M.inner_base + (has_c ? nin * 4 * k2 : 0);
  const int64_t rhs_base = M.bnd_base + (has_b ? nbd * k2 : 0);
  M.rhs_base = rhs_base;
  for (int p = 0; p < np; ++p)
    M.part_elem_base[p] += M.elem_base, M.part_rhs_base[p] += rhs_base;
  M.brhs_base = rhs_base + (has_s ? rhs_entries : 0); // boundary right-hand sides, ls per boundary face
  // element staging holds the upper triangle of the (symmetric) element matrices, entry-major
  auto tri = [](int i, int j, int ls) {
    const int a = std::min(i, j), b = std::max(i, j);
    return a * ls - a * (a - 1) / 2 + (b - a);
  };
  // staging entry of element e's matrix entry (i, j) and right-hand side entry i
  auto elem_slot = [&](int64_t e, int i, int j) {
    const int p = s.parts.empty() ? 0 : g.etype[e];
    const int64_t el = s.parts.empty() ? e : g.elocal[e];
    return M.part_elem_base[p] + tri(i, j, part_ls(p)) * part_ne(p) + el;
  };
  auto rhs_slot = [&](int64_t e, int i) {
    const int p = s.parts.empty() ? 0 : g.etype[e];
    const int64_t el = s.parts.empty() ? e : g.elocal[e];
    return M.part_rhs_base[p] + int64_t(i) * part_ne(p) + el;
  };
  M.staging_size = M.brhs_base + (has_bs ? nbd * ls : 0);
  if (M.staging_size >= (int64_t(1) << 32))
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "local-matrix staging exceeds 2^32 entries");

  const Adjacency A = dof_adjacency(s);
  struct Entry {
    int64_t key;
    uint32_t staging;
    int32_t col;
  };
  auto row_entries = [&](int64_t i, std::vector<Entry>& out) {
    out.clear();
    for (int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
      const int64_t e = A.elem[p];
      const int li = A.local[p];
      const int32_t* de = &src.dofs[e * ls];
      if (has_e)
        for (int lj = 0; lj < src.elem_size(e); ++lj)
          out.push_back({e * (nf + 1), uint32_t(elem_slot(e, li, lj)), de[lj]});
      for (int f = 0; f < nf; ++f) {
        const int64_t o = g.nb[e * nf + f];
        if (o >= 0 && has_c) {
          // blocks of ls x ls entries (ls the largest local size: a mixed grid's elements use the first
          // rows and columns of theirs)
          if (e < o) {
            const int64_t fi = inner_index[e * nf + f];
            const int32_t* dout = &src.dofs[o * ls];
            const int64_t key = e * (nf + 1) + 1 + f;
            for (int lj = 0; lj < src.elem_size(e); ++lj) // in_in
              out.push_back({key, uint32_t(M.inner_base + fi * 4 * k2 + 0 * k2 + li * ls + lj), de[lj]});
            for (int lj = 0; lj < src.elem_size(o); ++lj) // in_out
              out.push_back({key, uint32_t(M.inner_base + fi * 4 * k2 + 1 * k2 + li * ls + lj), dout[lj]});
          } else {
            const int fo = g.nbface[e * nf + f];
            const int64_t fi = inner_index[o * nf + fo];
            const int32_t* din = &src.dofs[o * ls];
            const int64_t key = o * (nf + 1) + 1 + fo;
            for (int lj = 0; lj < src.elem_size(o); ++lj) // out_in
              out.push_back({key, uint32_t(M.inner_base + fi * 4 * k2 + 2 * k2 + li * ls + lj), din[lj]});
            for (int lj = 0; lj < src.elem_size(e); ++lj) // out_out
              out.push_back({key, uint32_t(M.inner_base + fi * 4 * k2 + 3 * k2 + li * ls + lj), de[lj]});
          }
        } else if (o == -1 && has_b) {
          const int64_t bi = bnd_index[e * nf + f];
          for (int lj = 0; lj < src.elem_size(e); ++lj)
            out.push_back({e * (nf + 1) + 1 + f, uint32_t(M.bnd_base + bi * k2 + li * ls + lj), de[lj]});
        }
      }
    }
    std::stable_sort(out.begin(), out.end(), [](const Entry& a, const Entry& b) { return a.key < b.key; });
  };
  // pass 1: contributions per slot
  std::vector<uint32_t> slot_count(nnz, 0);
  parallel_for(n, [&](int64_t b, int64_t e) {
    std::vector<Entry> ent;
    for (int64_t i = b; i < e; ++i) {
      row_entries(i, ent);
      const int32_t* c0 = &cols[row_ptr[i]];
      const int32_t* c1 = &cols[row_ptr[i + 1]];
      for (const auto& en : ent) {
        const int32_t* it = std::lower_bound(c0, c1, en.col);
        if (it == c1 || *it != en.col)
          throw Error(HGDT_ERR_OPERATOR, "local entry outside the sparsity pattern (stencil too small)");
        ++slot_count[row_ptr[i] + (it - c0)];
      }
    }
  });
  std::vector<uint32_t> slot_ptr(nnz + 1, 0);
  uint64_t total = 0;
  for (int64_t t = 0; t < nnz; ++t) {
    slot_ptr[t] = uint32_t(total);
    total += slot_count[t];
  }
  if (total >= (uint64_t(1) << 32))
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "number of scatter contributions exceeds 2^32");
  slot_ptr[nnz] = uint32_t(total);
  std::vector<uint32_t> contrib(std::max<uint64_t>(total, 1));
  parallel_for(n, [&](int64_t b, int64_t e) {
    std::vector<Entry> ent;
    std::vector<uint32_t> fillpos;
    for (int64_t i = b; i < e; ++i) {
      row_entries(i, ent);
      const int64_t r0 = row_ptr[i], r1 = row_ptr[i + 1];
      fillpos.assign(slot_ptr.begin() + r0, slot_ptr.begin() + r1);
      const int32_t* c0 = &cols[r0];
      const int32_t* c1 = &cols[r1];
      for (const auto& en : ent) {
        const int64_t t = std::lower_bound(c0, c1, en.col) - c0;
        contrib[fillpos[t]++] = en.staging;
      }
    }
  });
  slot_count.clear();
  slot_count.shrink_to_fit();
  M.d_slot_ptr.upload(slot_ptr);
  M.d_contrib.upload(contrib);
  // right-hand side: per DoF in walk order, for every element holding it (ascending) the element
  // functional's contribution, then those of the element's boundary intersections in face order
  if (has_s || has_bs) {
    std::vector<uint32_t> rptr(n + 1, 0), rc;
    rc.reserve((has_s ? rhs_entries : 0) + (has_bs ? nbd * ls : 0) + 1);
    for (int64_t i = 0; i < n; ++i) {
      rptr[i] = uint32_t(rc.size());
      for (int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
        const int64_t e = A.elem[p];
        const int l = A.local[p];
        if (has_s)
          rc.push_back(uint32_t(rhs_slot(e, l)));
        if (has_bs)
          for (int f = 0; f < nf; ++f)
            if (bnd_index[e * nf + f] >= 0)
              rc.push_back(uint32_t(M.brhs_base + int64_t(bnd_index[e * nf + f]) * ls + l));
      }
    }
    rptr[n] = uint32_t(rc.size());
    if (rc.empty())
      rc.push_back(0);
    M.d_rhs_ptr.upload(rptr);
    M.d_rhs_contrib.upload(rc);
  }
  std::vector<int32_t> inner_packed(std::max<int64_t>(4 * nin, 1)), bnd_packed(std::max<int64_t>(2 * nbd, 1));
  for (int64_t f = 0; f < nin; ++f) {
    inner_packed[4 * f + 0] = M.inner_in[f], inner_packed[4 * f + 1] = M.inner_fin[f];
    inner_packed[4 * f + 2] = M.inner_out[f], inner_packed[4 * f + 3] = M.inner_fout[f];
  }
  for (int64_t f = 0; f < nbd; ++f)
    bnd_packed[2 * f + 0] = M.bnd_e[f], bnd_packed[2 * f + 1] = M.bnd_f[f];
  M.d_inner.upload(inner_packed);
  M.d_bnd.upload(bnd_packed);
  d_staging.allocate(std::max<int64_t>(M.staging_size, 1));
  M.built = true;
  M.signature = signature;
}

// DirichletConstraints::apply_local (tools/dirichlet-constraints.hh:85-110): DoFs
// attached to boundary intersections (the face and its lower-dimensional
// subentities); apply(matrix, vector) with ensure_symmetry (:172-177) becomes a
// per-slot flag (keep / zero / one) applied by the scatter kernel.
void Operator::build_dirichlet()
{
  if (dirichlet_built)
    return;
  const Space& s = *space;
  const Grid& g = *s.grid;
  const int ls = s.local_size, nf = g.ref.nfaces, fk = g.ref.face_nverts;
  std::vector<uint8_t> mask(n, 0);
  if (s.kind == HGDT_CONTINUOUS_LAGRANGE) {
    // local DoFs whose subentity lies in a boundary face (the face itself and its lower-dimensional
    // subentities), for every element type of the grid
    std::vector<std::vector<std::vector<int>>> face_nodes;
    for (size_t p = 0; p < std::max<size_t>(1, g.parts.size()); ++p) {
      const RefElement& r = g.mixed ? g.parts[p].ref : g.ref;
      const auto nodes = lagrange_node_vertices(r, s.order);
      face_nodes.emplace_back(r.nfaces);
      for (int f = 0; f < r.nfaces; ++f)
        for (int i = 0; i < int(nodes.size()); ++i) {
          bool inside = true;
          for (int v : nodes[i])
            inside = inside && std::find(r.faces[f], r.faces[f] + fk, v) != r.faces[f] + fk;
          if (inside)
            face_nodes.back()[f].push_back(i);
        }
    }
    // boundary intersections in walk order; only the Dirichlet ones (all, without a boundary info) constrain
    int64_t bi = 0;
    for (int64_t e = 0; e < g.ne; ++e)
      for (int f = 0; f < nf; ++f)
        if (g.nb[e * nf + f] == -1) {
          const bool dir = dirichlet_faces.empty() || dirichlet_faces[bi] != 0;
          ++bi;
          if (dir)
            for (int i : face_nodes[g.mixed ? g.etype[e] : 0][f])
              mask[s.dofs[e * ls + i]] = 1;
        }
  }
  dirichlet_dofs.clear();
  for (int64_t i = 0; i < n; ++i)
    if (mask[i])
      dirichlet_dofs.push_back(int32_t(i));
  // slot flags: 1 the value 0, 2 the value 1; the rows of the constrained DoFs, and with ensure_symmetry
  // their columns, unit (only_clear: cleared)
  std::vector<uint8_t> flag(std::max<int64_t>(nnz, 1), 0);
  const bool clear = dirichlet_only_clear, cols_too = dirichlet_ensure_symmetry;
  parallel_for(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i)
      for (int64_t t = row_ptr[i]; t < row_ptr[i + 1]; ++t)
        if (mask[i] || (cols_too && mask[cols[t]]))
          flag[t] = (cols[t] == i && !clear) ? 2 : 1;
  });
  d_slot_flag.upload(flag);
  d_dof_mask.upload(mask);
  dirichlet_built = true;
}

} // namespace hgdt
