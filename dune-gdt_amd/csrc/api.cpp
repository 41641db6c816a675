// api.cpp -- the extern "C" boundary of libhgdt.so (declared in include/hgdt.h).
//
// No exception crosses the boundary: every entry point returns an hgdt_status and
// keeps the message of the last failure in a thread-local string.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <unordered_map>

#include "hgdt_internal.hpp"

struct hgdt_grid_s {
  std::shared_ptr<hgdt::Grid> g;
};
struct hgdt_space_s {
  std::shared_ptr<hgdt::Space> s;
};
struct hgdt_operator_s {
  std::unique_ptr<hgdt::Operator> op;
};

namespace hgdt {

thread_local std::string g_last_error;

template <class F>
hgdt_status guarded(F&& f)
{
  try {
    f();
    g_last_error.clear();
    return HGDT_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return HGDT_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return HGDT_ERR_OPERATOR;
  }
}

static void require(bool cond, hgdt_status s, const char* msg)
{
  if (!cond)
    throw Error(s, msg);
}

static void require_device()
{
  int count = 0;
  const hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count < 1)
    HGDT_THROW(HGDT_ERR_DEVICE, "no usable HIP device: libhgdt has no CPU fallback");
}

static std::shared_ptr<Grid> make_structured(int geom, const int32_t* cells, const double* lo, const double* hi)
{
  auto g = std::make_shared<Grid>();
  g->ref = reference_element(geom);
  g->structured = true;
  const int d = g->ref.dim;
  int64_t n[3] = {1, 1, 1}, nvx[3] = {1, 1, 1};
  double h[3] = {0, 0, 0};
  for (int a = 0; a < d; ++a) {
    require(cells[a] > 0, HGDT_ERR_INVALID_ARGUMENT, "cells must be positive");
    require(hi[a] > lo[a], HGDT_ERR_INVALID_ARGUMENT, "upper_right must exceed lower_left");
    n[a] = cells[a], nvx[a] = cells[a] + 1, h[a] = (hi[a] - lo[a]) / cells[a];
  }
  g->nv = nvx[0] * nvx[1] * nvx[2];
  const int per_cell = (geom == HGDT_TRI) ? 2 : (geom == HGDT_TET) ? 6 : 1;
  const int64_t ncell = n[0] * n[1] * n[2];
  g->ne = ncell * per_cell;
  require(g->nv < (int64_t(1) << 31) && g->ne * g->ref.nverts < (int64_t(1) << 40), HGDT_ERR_NOT_SUPPORTED,
          "grid too large for 32-bit vertex indices");
  g->coords.resize(g->nv * d);
  parallel_for(g->nv, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) {
      const int64_t idx[3] = {v % nvx[0], (v / nvx[0]) % nvx[1], v / (nvx[0] * nvx[1])};
      for (int a = 0; a < d; ++a)
        g->coords[v * d + a] = lo[a] + idx[a] * h[a];
    }
  });
  g->conn.resize(g->ne * g->ref.nverts);
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  parallel_for(ncell, [&](int64_t b, int64_t e) {
    for (int64_t c = b; c < e; ++c) {
      const int64_t i = c % n[0], j = (c / n[0]) % n[1], k = c / (n[0] * n[1]);
      int32_t corner[8];
      for (int bit = 0; bit < (1 << d); ++bit)
        corner[bit] = int32_t(((k + ((bit >> 2) & 1)) * nvx[1] + (j + ((bit >> 1) & 1))) * nvx[0] + i + (bit & 1));
      int32_t* out = &g->conn[c * per_cell * g->ref.nverts];
      if (geom == HGDT_QUAD || geom == HGDT_HEX) {
        for (int bit = 0; bit < (1 << d); ++bit)
          out[bit] = corner[bit];
      } else if (geom == HGDT_TRI) {
        const int32_t t[6] = {corner[0], corner[1], corner[3], corner[0], corner[2], corner[3]};
        std::copy(t, t + 6, out);
      } else {
        for (int p = 0; p < 6; ++p) {
          const int b1 = 1 << perms[p][0], b2 = b1 | (1 << perms[p][1]);
          out[4 * p + 0] = corner[0], out[4 * p + 1] = corner[b1], out[4 * p + 2] = corner[b2];
          out[4 * p + 3] = corner[7];
        }
      }
    }
  });
  g->inst_elem = {0, g->ne};
  g->inst_vert = {0, g->nv};
  Lattice L;
  for (int a = 0; a < d; ++a)
    L.n[a] = n[a], L.lo[a] = lo[a], L.h[a] = h[a];
  g->lattice = {L};
  g->finalize();
  return g;
}

// per-element input arrays an assembly reads: the distinct coefficient fields of the forms and the per-element
// sources
static int64_t per_element_inputs(const Operator& op)
{
  std::set<const double*> fields;
  for (const auto& f : op.eforms)
    if (f.field)
      fields.insert(f.field);
  for (const auto* forms : {&op.cforms, &op.bforms})
    for (const auto& f : *forms) {
      if (f.kappa)
        fields.insert(f.kappa);
      if (f.omega)
        fields.insert(f.omega);
    }
  int64_t k = int64_t(fields.size());
  for (const auto& s : op.sources)
    if (s.kind == HGDT_SOURCE_PER_ELEMENT)
      ++k;
  return k;
}

void Operator::assemble(hipStream_t stream)
{
  prepare_assembly();
  // matrix-free mode: the right-hand side only, the operator is recomputed at every application
  matrix_free_assembly = (requested_format == HGDT_FORMAT_MATRIX_FREE);
  const int64_t inputs = 8 * space->grid->ne * per_element_inputs(*this);
  if (!matrix_free_assembly && direct_applicable()) {
    if (dirichlet)
      build_dirichlet();
    assemble_direct(stream);
    // one pass over the DoFs: the per-element inputs and the Dirichlet mask read; the diagonal and 7 upper
    // diagonals, the right-hand side, D^{-1/2} and D^{1/2} written
    assembly_bytes = inputs + (88 + (dirichlet ? 1 : 0)) * n;
    return;
  }
  build_scatter_maps();
  if (dirichlet)
    build_dirichlet();
  launch_assembly(*this, stream, !matrix_free_assembly);
  fmt.valid = false;
  direct_storage = false;
  values_stale = false;
  // the compulsory bytes: the per-element inputs read, the values and the right-hand side written
  assembly_bytes = inputs + (matrix_free_assembly ? 0 : 8 * nnz) + 8 * n;
}

} // namespace hgdt

using namespace hgdt;

extern "C" {

const char* hgdt_last_error(void)
{
  return g_last_error.c_str();
}

int hgdt_version(void)
{
  return 100; // 0.1.0
}

hgdt_status hgdt_device_count(int* count)
{
  return guarded([&] {
    require(count, HGDT_ERR_INVALID_ARGUMENT, "count is NULL");
    *count = 0;
    HGDT_HIP(hipGetDeviceCount(count));
    require(*count > 0, HGDT_ERR_DEVICE, "no HIP device visible");
  });
}

hgdt_status hgdt_memcpy(void* dst, const void* src, int64_t bytes, void* stream)
{
  return guarded([&] {
    require(dst && src && bytes >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid copy");
    HGDT_HIP(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDefault, static_cast<hipStream_t>(stream)));
  });
}

hgdt_status hgdt_device_alloc(int64_t bytes, void** ptr)
{
  return guarded([&] {
    require(ptr && bytes >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    require_device();
    *ptr = nullptr;
    if (bytes == 0)
      return;
    const hipError_t e = hipMalloc(ptr, size_t(bytes));
    if (e == hipErrorOutOfMemory)
      HGDT_THROW(HGDT_ERR_OUT_OF_MEMORY, "device allocation failed");
    HGDT_HIP(e);
  });
}

hgdt_status hgdt_device_free(void* ptr)
{
  return guarded([&] {
    if (ptr)
      HGDT_HIP(hipFree(ptr));
  });
}

hgdt_status hgdt_stream_synchronize(void* stream)
{
  return guarded([&] {
    require_device();
    HGDT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  });
}

hgdt_status hgdt_grid_create_structured(int geometry, const int32_t* cells, const double* lower_left,
                                        const double* upper_right, hgdt_grid* grid)
{
  return guarded([&] {
    require(cells && lower_left && upper_right && grid, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    require_device();
    *grid = new hgdt_grid_s{make_structured(geometry, cells, lower_left, upper_right)};
  });
}

hgdt_status hgdt_grid_create(int geometry, int64_t num_vertices, const double* coords, int64_t num_elements,
                             const int32_t* connectivity, hgdt_grid* grid)
{
  return guarded([&] {
    require(coords && connectivity && grid, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    require(num_vertices > 0 && num_elements > 0, HGDT_ERR_INVALID_ARGUMENT, "empty grid");
    require(num_vertices < (int64_t(1) << 31), HGDT_ERR_NOT_SUPPORTED, "too many vertices");
    require_device();
    auto g = std::make_shared<Grid>();
    g->ref = reference_element(geometry);
    g->nv = num_vertices, g->ne = num_elements;
    g->coords.assign(coords, coords + num_vertices * g->ref.dim);
    g->conn.assign(connectivity, connectivity + num_elements * g->ref.nverts);
    for (int32_t v : g->conn)
      require(v >= 0 && v < num_vertices, HGDT_ERR_INVALID_ARGUMENT, "connectivity references a vertex out of range");
    g->check_affine();
    g->inst_elem = {0, g->ne};
    g->inst_vert = {0, g->nv};
    g->finalize();
    *grid = new hgdt_grid_s{g};
  });
}

hgdt_status hgdt_grid_create_mixed(int64_t num_vertices, const double* coords, int64_t num_elements,
                                   const int32_t* geometry, const int64_t* offsets, const int32_t* vertices,
                                   hgdt_grid* grid)
{
  return guarded([&] {
    require(coords && geometry && offsets && vertices && grid, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    require(num_vertices > 0 && num_elements > 0, HGDT_ERR_INVALID_ARGUMENT, "empty grid");
    require(num_vertices < (int64_t(1) << 31) && num_elements < (int64_t(1) << 31), HGDT_ERR_NOT_SUPPORTED,
            "too many vertices or elements");
    require(offsets[0] == 0, HGDT_ERR_INVALID_ARGUMENT, "offsets[0] must be 0");
    bool present[2] = {false, false};
    for (int64_t e = 0; e < num_elements; ++e) {
      require(geometry[e] == HGDT_QUAD || geometry[e] == HGDT_TRI, HGDT_ERR_INVALID_ARGUMENT,
              "mixed grids hold triangles and quadrilaterals");
      require(offsets[e + 1] - offsets[e] == (geometry[e] == HGDT_QUAD ? 4 : 3), HGDT_ERR_INVALID_ARGUMENT,
              "an element's vertex count does not match its geometry type");
      present[geometry[e]] = true;
    }
    for (int64_t t = 0; t < offsets[num_elements]; ++t)
      require(vertices[t] >= 0 && vertices[t] < num_vertices, HGDT_ERR_INVALID_ARGUMENT,
              "connectivity references a vertex out of range");
    require_device();
    auto g = std::make_shared<Grid>();
    g->nv = num_vertices, g->ne = num_elements;
    g->coords.assign(coords, coords + num_vertices * 2);
    if (!(present[0] && present[1])) { // one element type: the plain grid of that type
      g->ref = reference_element(present[0] ? HGDT_QUAD : HGDT_TRI);
      g->conn.assign(vertices, vertices + offsets[num_elements]);
    } else {
      // strides of the largest type (the quadrilateral), parts in ascending geometry id (quadrilaterals,
      // then triangles), each with its elements in grid order
      g->mixed = true;
      g->ref = reference_element(HGDT_QUAD);
      g->ref.geom = HGDT_MIXED;
      g->parts = std::vector<GridPart>(2);
      g->parts[0].ref = reference_element(HGDT_QUAD);
      g->parts[1].ref = reference_element(HGDT_TRI);
      g->etype.resize(num_elements), g->elocal.resize(num_elements);
      g->conn.assign(num_elements * 4, -1);
      for (int64_t e = 0; e < num_elements; ++e) {
        const int p = geometry[e] == HGDT_QUAD ? 0 : 1;
        GridPart& P = g->parts[p];
        g->etype[e] = int8_t(p), g->elocal[e] = int32_t(P.eid.size());
        P.eid.push_back(int32_t(e));
        for (int64_t t = offsets[e]; t < offsets[e + 1]; ++t) {
          P.conn.push_back(vertices[t]);
          g->conn[e * 4 + (t - offsets[e])] = vertices[t];
        }
      }
    }
    g->check_affine();
    g->inst_elem = {0, g->ne};
    g->inst_vert = {0, g->nv};
    g->finalize();
    *grid = new hgdt_grid_s{g};
  });
}

hgdt_status hgdt_grid_create_union(const hgdt_grid* grids, int32_t count, hgdt_grid* grid)
{
  return guarded([&] {
    require(grids && grid && count > 0, HGDT_ERR_INVALID_ARGUMENT, "need at least one grid");
    require_device();
    auto u = std::make_shared<Grid>();
    u->ref = grids[0]->g->ref;
    u->structured = false;
    // a union of structured grids keeps the lattices of its parts (matrix-free application)
    bool lattices = true;
    for (int i = 0; i < count; ++i)
      lattices = lattices && grids[i] && grids[i]->g->lattice.size() == 1;
    if (lattices)
      for (int i = 0; i < count; ++i)
        u->lattice.push_back(grids[i]->g->lattice[0]);
    int64_t nv = 0, ne = 0;
    for (int i = 0; i < count; ++i) {
      require(grids[i] && grids[i]->g->ref.geom == u->ref.geom, HGDT_ERR_INVALID_ARGUMENT,
              "all grids of a union must have the same geometry type");
      require(grids[i]->g->ninst() == 1, HGDT_ERR_INVALID_ARGUMENT, "grids of a union must not be unions");
      require(!grids[i]->g->mixed, HGDT_ERR_NOT_SUPPORTED, "unions of mixed grids are not provided");
      nv += grids[i]->g->nv, ne += grids[i]->g->ne;
    }
    require(nv < (int64_t(1) << 31), HGDT_ERR_NOT_SUPPORTED, "union too large for 32-bit vertex indices");
    u->nv = nv, u->ne = ne;
    u->coords.reserve(nv * u->ref.dim);
    u->conn.reserve(ne * u->ref.nverts);
    u->inst_elem = {0};
    u->inst_vert = {0};
    u->nb.reserve(ne * u->ref.nfaces);
    u->nbface.reserve(ne * u->ref.nfaces);
    for (int i = 0; i < count; ++i) {
      const Grid& g = *grids[i]->g;
      const int32_t voff = int32_t(u->inst_vert.back());
      const int32_t eoff = int32_t(u->inst_elem.back());
      u->coords.insert(u->coords.end(), g.coords.begin(), g.coords.end());
      for (int32_t v : g.conn)
        u->conn.push_back(v + voff);
      // the parts are disjoint: their intersections carry over with shifted element indices
      for (int32_t o : g.nb)
        u->nb.push_back(o < 0 ? o : o + eoff);
      u->nbface.insert(u->nbface.end(), g.nbface.begin(), g.nbface.end());
      u->inst_elem.push_back(u->inst_elem.back() + g.ne);
      u->inst_vert.push_back(u->inst_vert.back() + g.nv);
    }
    u->finalize(/*neighbours_known=*/true);
    *grid = new hgdt_grid_s{u};
  });
}

hgdt_status hgdt_grid_destroy(hgdt_grid grid)
{
  return guarded([&] { delete grid; });
}

hgdt_status hgdt_grid_info(hgdt_grid grid, int64_t* info)
{
  return guarded([&] {
    require(grid && info, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    const Grid& g = *grid->g;
    info[0] = g.ref.dim, info[1] = g.nv, info[2] = g.ne, info[3] = g.ref.nverts, info[4] = g.ref.nfaces;
    info[5] = g.n_inner, info[6] = g.n_boundary, info[7] = g.ninst(), info[8] = g.ref.geom;
  });
}

hgdt_status hgdt_grid_copy_coords(hgdt_grid grid, double* host_out)
{
  return guarded([&] {
    require(grid && host_out, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::memcpy(host_out, grid->g->coords.data(), grid->g->coords.size() * sizeof(double));
  });
}

hgdt_status hgdt_grid_copy_connectivity(hgdt_grid grid, int32_t* host_out)
{
  return guarded([&] {
    require(grid && host_out, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::memcpy(host_out, grid->g->conn.data(), grid->g->conn.size() * sizeof(int32_t));
  });
}

hgdt_status hgdt_grid_instance_elements(hgdt_grid grid, int64_t* host_out)
{
  return guarded([&] {
    require(grid && host_out, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::copy(grid->g->inst_elem.begin(), grid->g->inst_elem.end(), host_out);
  });
}

hgdt_status hgdt_grid_quadrature_points(hgdt_grid grid, int order, double* host_out, int32_t* num_points)
{
  return guarded([&] {
    require(grid && num_points && order >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    const Grid& g = *grid->g;
    *num_points = g.mixed ? 0 : element_rule(g.ref, order).size();
    if (!host_out)
      return;
    const std::vector<int64_t> off = element_point_offsets(g, order);
    std::vector<Quadrature> rules;
    for (const GridPart& P : g.parts)
      rules.push_back(element_rule(P.ref, order));
    const Quadrature q0 = element_rule(g.mixed ? g.parts[0].ref : g.ref, order);
    const int d = g.ref.dim;
    parallel_for(g.ne, [&](int64_t b, int64_t e) {
      for (int64_t el = b; el < e; ++el) {
        const RefElement& r = g.eref(el);
        const Quadrature& q = g.mixed ? rules[g.etype[el]] : q0;
        const int32_t* c = &g.conn[el * g.ref.nverts];
        double x0[3], J[3][3];
        for (int a = 0; a < d; ++a)
          x0[a] = g.coords[int64_t(c[0]) * d + a];
        for (int k = 0; k < d; ++k) {
          const int lv = r.simplex ? (k + 1) : (1 << k);
          for (int a = 0; a < d; ++a)
            J[a][k] = g.coords[int64_t(c[lv]) * d + a] - x0[a];
        }
        for (int p = 0; p < q.size(); ++p)
          for (int a = 0; a < d; ++a) {
            double s = x0[a];
            for (int k = 0; k < d; ++k)
              s += J[a][k] * q.x[p * d + k];
            host_out[(off[el] + p) * d + a] = s;
          }
      }
    });
  });
}

hgdt_status hgdt_grid_quadrature_offsets(hgdt_grid grid, int order, int64_t* host_out)
{
  return guarded([&] {
    require(grid && host_out && order >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    const std::vector<int64_t> off = element_point_offsets(*grid->g, order);
    std::copy(off.begin(), off.end(), host_out);
  });
}

hgdt_status hgdt_grid_boundary_intersections(hgdt_grid grid, double* centers, double* normals)
{
  return guarded([&] {
    require(grid, HGDT_ERR_INVALID_ARGUMENT, "NULL grid");
    const Grid& g = *grid->g;
    const int d = g.ref.dim, nf = g.ref.nfaces, nv = g.ref.nverts, fk = g.ref.face_nverts;
    int64_t bi = 0;
    for (int64_t e = 0; e < g.ne; ++e)
      for (int f = 0; f < nf; ++f) {
        if (g.nb[e * nf + f] != -1)
          continue;
        const RefElement& r = g.eref(e);
        double cen[3] = {0, 0, 0}, ecen[3] = {0, 0, 0}, v[4][3] = {};
        for (int k = 0; k < fk; ++k) {
          const int32_t vid = g.conn[e * nv + r.faces[f][k]];
          for (int a = 0; a < d; ++a)
            v[k][a] = g.coords[int64_t(vid) * d + a], cen[a] += v[k][a] / fk;
        }
        for (int k = 0; k < r.nverts; ++k)
          for (int a = 0; a < d; ++a)
            ecen[a] += g.coords[int64_t(g.conn[e * nv + k]) * d + a] / r.nverts;
        double nrm[3] = {0, 0, 0};
        if (d == 2) {
          nrm[0] = v[1][1] - v[0][1], nrm[1] = -(v[1][0] - v[0][0]);
        } else {
          const double a[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
          const double b[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
          nrm[0] = a[1] * b[2] - a[2] * b[1], nrm[1] = a[2] * b[0] - a[0] * b[2], nrm[2] = a[0] * b[1] - a[1] * b[0];
        }
        double len = 0., side = 0.;
        for (int a = 0; a < d; ++a)
          len += nrm[a] * nrm[a], side += nrm[a] * (cen[a] - ecen[a]);
        len = std::sqrt(len) * (side < 0. ? -1. : 1.); // outward: away from the element's centre
        for (int a = 0; a < d; ++a) {
          if (centers)
            centers[bi * d + a] = cen[a];
          if (normals)
            normals[bi * d + a] = nrm[a] / len;
        }
        ++bi;
      }
  });
}

hgdt_status hgdt_grid_boundary_quadrature_points(hgdt_grid grid, int order, double* host_out, int32_t* num_points)
{
  return guarded([&] {
    require(grid && num_points && order >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    const Grid& g = *grid->g;
    const int d = g.ref.dim;
    const Quadrature q = face_rule(g.ref, order);
    const int nq = q.size(), nf = g.ref.nfaces;
    *num_points = nq;
    if (!host_out)
      return;
    // the points of the device face kernels: X = P0 + s (P1 - P0) (+ t (P2 - P0) in 3d) over the corners of
    // the reference face, faces in walk order (elements ascending, faces in reference order)
    int64_t bi = 0;
    for (int64_t e = 0; e < g.ne; ++e)
      for (int f = 0; f < nf; ++f) {
        if (g.nb[e * nf + f] != -1)
          continue;
        double P[3][3];
        for (int k = 0; k < d; ++k) {
          const int32_t v = g.conn[e * g.ref.nverts + g.eref(e).faces[f][k]];
          for (int a = 0; a < d; ++a)
            P[k][a] = g.coords[int64_t(v) * d + a];
        }
        for (int p = 0; p < nq; ++p)
          for (int a = 0; a < d; ++a) {
            const double s = q.x[size_t(p) * (d - 1)];
            double X = P[0][a] + s * (P[1][a] - P[0][a]);
            if (d == 3)
              X += q.x[size_t(p) * 2 + 1] * (P[2][a] - P[0][a]);
            host_out[(bi * nq + p) * d + a] = X;
          }
        ++bi;
      }
  });
}

// the DoFs of a space on a mixed grid: as on one element type (continuous order 1: the vertices; order 2:
// one DoF per subentity in order of first appearance in the element walk; discontinuous: element offsets
// and local index), every element with the Lagrange nodes of its own type (the local finite element of its
// geometry type, ContinuousMapper / DiscontinuousMapper over the grid's types,
// spaces/mapper/continuous.hh:117-182, spaces/mapper/discontinuous.hh:102-125)
static void make_mixed_space(Space& s)
{
  const Grid& g = *s.grid;
  const int np = int(g.parts.size());
  std::vector<std::vector<std::vector<int>>> nodes(np);
  s.parts = std::vector<SpacePart>(np);
  s.local_size = 0;
  for (int p = 0; p < np; ++p) {
    nodes[p] = lagrange_node_vertices(g.parts[p].ref, s.order);
    s.parts[p].local_size = int(nodes[p].size());
    s.local_size = std::max(s.local_size, s.parts[p].local_size);
  }
  const int ls = s.local_size, nv = g.ref.nverts;
  s.dofs.assign(g.ne * ls, -1);
  if (s.kind == HGDT_CONTINUOUS_LAGRANGE && s.order == 1) {
    for (int64_t e = 0; e < g.ne; ++e)
      for (int i = 0; i < s.elem_size(e); ++i)
        s.dofs[e * ls + i] = g.conn[e * nv + nodes[g.etype[e]][i][0]];
    s.size = g.nv;
  } else if (s.kind == HGDT_CONTINUOUS_LAGRANGE) {
    std::map<std::array<int32_t, 4>, int32_t> ids;
    for (int64_t e = 0; e < g.ne; ++e)
      for (int i = 0; i < s.elem_size(e); ++i) {
        std::array<int32_t, 4> key;
        key.fill(-1);
        int c = 0;
        for (int lv : nodes[g.etype[e]][i])
          key[c++] = g.conn[e * nv + lv];
        std::sort(key.begin(), key.begin() + c);
        auto it = ids.find(key);
        if (it == ids.end())
          it = ids.emplace(key, int32_t(ids.size())).first;
        s.dofs[e * ls + i] = it->second;
      }
    s.size = int64_t(ids.size());
  } else {
    int64_t off = 0;
    for (int64_t e = 0; e < g.ne; ++e)
      for (int i = 0; i < s.elem_size(e); ++i)
        s.dofs[e * ls + i] = int32_t(off++);
    s.size = off;
    require(s.size < (int64_t(1) << 31), HGDT_ERR_NOT_SUPPORTED, "too many DoFs for 32-bit indices");
  }
  s.inst_dof = {0, s.size};
  for (int p = 0; p < np; ++p) {
    SpacePart& P = s.parts[p];
    P.dofs.reserve(g.parts[p].ne() * P.local_size);
    for (int32_t e : g.parts[p].eid)
      for (int i = 0; i < P.local_size; ++i)
        P.dofs.push_back(s.dofs[int64_t(e) * ls + i]);
    P.d_dofs.upload(P.dofs);
  }
  s.d_dofs.upload(s.dofs);
}

hgdt_status hgdt_space_create(hgdt_grid grid, int kind, int order, hgdt_space* space)
{
  return guarded([&] {
    require(grid && space, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    require(kind == HGDT_CONTINUOUS_LAGRANGE || kind == HGDT_DISCONTINUOUS_LAGRANGE, HGDT_ERR_INVALID_ARGUMENT,
            "unknown space kind");
    if (order != 1 && order != 2)
      HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "device spaces are provided for polynomial order 1 and 2");
    auto s = std::make_shared<Space>();
    s->grid = grid->g;
    s->kind = kind, s->order = order;
    const Grid& g = *s->grid;
    if (g.mixed) {
      make_mixed_space(*s);
      *space = new hgdt_space_s{s};
      return;
    }
    const auto nodes = lagrange_node_vertices(g.ref, order);
    s->local_size = int(nodes.size());
    const int ls = s->local_size;
    s->dofs.resize(g.ne * ls);
    if (kind == HGDT_CONTINUOUS_LAGRANGE && order == 1) {
      // P1/Q1 Lagrange nodes are the element vertices in reference order: global DoF = vertex index
      s->dofs = g.conn;
      s->size = g.nv;
      s->inst_dof = g.inst_vert;
    } else if (kind == HGDT_CONTINUOUS_LAGRANGE) {
      // one DoF per subentity (order 2 has one node per subentity), keyed by the sorted global vertices
      // of the subentity and numbered in order of first appearance in the element walk
      // (ContinuousMapper, spaces/mapper/continuous.hh:139-182, up to the order of its index sets)
      struct KeyHash {
        size_t operator()(const std::array<int32_t, 8>& k) const
        {
          uint64_t h = 1469598103934665603ull;
          for (int32_t v : k)
            h = (h ^ uint32_t(v)) * 1099511628211ull;
          return size_t(h);
        }
      };
      std::unordered_map<std::array<int32_t, 8>, int32_t, KeyHash> ids;
      s->inst_dof.assign(1, 0);
      for (int in = 0; in < g.ninst(); ++in) {
        for (int64_t e = g.inst_elem[in]; e < g.inst_elem[in + 1]; ++e)
          for (int i = 0; i < ls; ++i) {
            std::array<int32_t, 8> key;
            key.fill(-1);
            int c = 0;
            for (int lv : nodes[i])
              key[c++] = g.conn[e * g.ref.nverts + lv];
            std::sort(key.begin(), key.begin() + c);
            auto it = ids.find(key);
            if (it == ids.end())
              it = ids.emplace(key, int32_t(ids.size())).first;
            s->dofs[e * ls + i] = it->second;
          }
        s->inst_dof.push_back(int64_t(ids.size()));
      }
      s->size = int64_t(ids.size());
    } else {
      // DiscontinuousMapper: element offset + local index (spaces/mapper/discontinuous.hh:104-109)
      parallel_for(g.ne, [&](int64_t b, int64_t e) {
        for (int64_t el = b; el < e; ++el)
          for (int i = 0; i < ls; ++i)
            s->dofs[el * ls + i] = int32_t(el * ls + i);
      });
      s->size = g.ne * ls;
      require(s->size < (int64_t(1) << 31), HGDT_ERR_NOT_SUPPORTED, "too many DoFs for 32-bit indices");
      s->inst_dof.clear();
      for (int64_t e : g.inst_elem)
        s->inst_dof.push_back(e * ls);
    }
    s->d_dofs.upload(s->dofs);
    *space = new hgdt_space_s{s};
  });
}

hgdt_status hgdt_space_destroy(hgdt_space space)
{
  return guarded([&] { delete space; });
}

hgdt_status hgdt_space_size(hgdt_space space, int64_t* size, int32_t* local_size)
{
  return guarded([&] {
    require(space, HGDT_ERR_INVALID_ARGUMENT, "NULL space");
    if (size)
      *size = space->s->size;
    if (local_size)
      *local_size = space->s->local_size;
  });
}

hgdt_status hgdt_space_copy_global_indices(hgdt_space space, int32_t* host_out)
{
  return guarded([&] {
    require(space && host_out, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::memcpy(host_out, space->s->dofs.data(), space->s->dofs.size() * sizeof(int32_t));
  });
}

hgdt_status hgdt_space_instance_dofs(hgdt_space space, int64_t* host_out)
{
  return guarded([&] {
    require(space && host_out, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::copy(space->s->inst_dof.begin(), space->s->inst_dof.end(), host_out);
  });
}

hgdt_status hgdt_space_error_norms(hgdt_space space, const double* dofs, int order, const double* g_values,
                                   const double* g_gradients, const double* diffusion_field, double diffusion_value,
                                   double* host_out, void* stream)
{
  return guarded([&] {
    require(space && dofs && host_out, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    require_device();
    error_norms(*space->s, dofs, order, g_values, g_gradients, diffusion_field, diffusion_value, host_out,
                static_cast<hipStream_t>(stream));
  });
}

hgdt_status hgdt_space_error_norms_noalloc(hgdt_space space, const double* dofs, int order, const double* g_values,
                                           const double* g_gradients, const double* diffusion_field,
                                           double diffusion_value, double* host_out, void* stream)
{
  return guarded([&] {
    require(space && dofs && host_out, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    require_device();
    error_norms_noalloc(*space->s, dofs, order, g_values, g_gradients, diffusion_field, diffusion_value, host_out,
                        static_cast<hipStream_t>(stream));
  });
}

hgdt_status hgdt_noalloc_scratch_bytes(int64_t* bytes)
{
  return guarded([&] {
    require(bytes, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    *bytes = error_norms_scratch_bytes();
  });
}

hgdt_status hgdt_operator_create(hgdt_space space, int stencil, hgdt_operator* op)
{
  return guarded([&] {
    require(space && op, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    require(stencil >= HGDT_STENCIL_ELEMENT && stencil <= HGDT_STENCIL_ELEMENT_AND_INTERSECTION,
            HGDT_ERR_INVALID_ARGUMENT, "unknown stencil");
    auto o = std::make_unique<Operator>();
    o->space = o->source = space->s;
    o->stencil = stencil;
    o->build_pattern();
    *op = new hgdt_operator_s{std::move(o)};
  });
}

static void check_space_pair(hgdt_space source, hgdt_space range)
{
  require(source && range, HGDT_ERR_INVALID_ARGUMENT, "NULL space");
  const Space& s = *source->s;
  const Space& r = *range->s;
  require(s.grid == r.grid, HGDT_ERR_NOT_SUPPORTED, "the source and range spaces must live on the same grid");
  require(s.order == r.order && s.local_size == r.local_size, HGDT_ERR_NOT_SUPPORTED,
          "the source and range spaces must have the same local finite element (polynomial order)");
}

hgdt_status hgdt_operator_create_between(hgdt_space source, hgdt_space range, int stencil, hgdt_operator* op)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    check_space_pair(source, range);
    require(stencil >= HGDT_STENCIL_ELEMENT && stencil <= HGDT_STENCIL_ELEMENT_AND_INTERSECTION,
            HGDT_ERR_INVALID_ARGUMENT, "unknown stencil");
    auto o = std::make_unique<Operator>();
    o->space = range->s;
    o->source = source->s == range->s ? range->s : source->s;
    o->stencil = stencil;
    o->build_pattern();
    *op = new hgdt_operator_s{std::move(o)};
  });
}

hgdt_status hgdt_operator_create_on_matrix(hgdt_space source, hgdt_space range, int stencil, const int64_t* row_ptr,
                                           const int32_t* cols, int64_t rows, int64_t nnz, double* values,
                                           hgdt_operator* op)
{
  return guarded([&] {
    require(op && row_ptr && (cols || nnz == 0) && (values || nnz == 0) && rows >= 0 && nnz >= 0,
            HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    check_space_pair(source, range);
    require(stencil >= HGDT_STENCIL_ELEMENT && stencil <= HGDT_STENCIL_ELEMENT_AND_INTERSECTION,
            HGDT_ERR_INVALID_ARGUMENT, "unknown stencil");
    require(rows == range->s->size, HGDT_ERR_SHAPES_DO_NOT_MATCH, "the matrix needs one row per DoF of the range space");
    require(row_ptr[rows] == nnz, HGDT_ERR_SHAPES_DO_NOT_MATCH, "row_ptr[rows] must equal nnz");
    auto o = std::make_unique<Operator>();
    o->space = range->s;
    o->source = source->s == range->s ? range->s : source->s;
    o->stencil = stencil;
    o->borrowed = true;
    o->borrowed_row_ptr.assign(row_ptr, row_ptr + rows + 1);
    o->borrowed_cols.assign(cols, cols + nnz);
    o->d_values.borrow(values, size_t(nnz));
    o->build_pattern();
    o->borrowed_row_ptr.clear(), o->borrowed_cols.clear();
    o->borrowed_row_ptr.shrink_to_fit(), o->borrowed_cols.shrink_to_fit();
    *op = new hgdt_operator_s{std::move(o)};
  });
}

hgdt_status hgdt_operator_destroy(hgdt_operator op)
{
  return guarded([&] { delete op; });
}

hgdt_status hgdt_operator_shape(hgdt_operator op, int64_t* rows, int64_t* cols, int64_t* nnz)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    if (rows)
      *rows = op->op->n;
    if (cols)
      *cols = op->op->ncols;
    if (nnz)
      *nnz = op->op->nnz;
  });
}

hgdt_status hgdt_operator_copy_pattern(hgdt_operator op, int64_t* row_ptr, int32_t* cols)
{
  return guarded([&] {
    require(op && row_ptr && cols, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::copy(op->op->row_ptr.begin(), op->op->row_ptr.end(), row_ptr);
    std::copy(op->op->cols.begin(), op->op->cols.end(), cols);
  });
}

static void require_matrix(const Operator& o)
{
  require(!o.matrix_free_assembly, HGDT_ERR_OPERATOR,
          "the operator was assembled in matrix-free mode and holds no matrix: set another format and assemble again");
}

// CSR values of a structured assembly, computed from its storage once the assembly's work is done (the
// assembly's stream may be gone: the device is synchronized and the default stream used)
static void materialize(Operator& o)
{
  if (!o.values_stale)
    return;
  HGDT_HIP(hipDeviceSynchronize());
  o.materialize_values(nullptr);
}

static void forms_changed(Operator& o)
{
  o.tables_ready = false;
  o.direct_state = -1;
  if (o.requested_format == HGDT_FORMAT_MATRIX_FREE) // the reference matrices follow the forms
    o.fmt.structure_ready = false, o.fmt.valid = false;
}

hgdt_status hgdt_operator_append_laplace(hgdt_operator op, const double* field, double value, int32_t over_integrate,
                                         double scaling)
{
  return guarded([&] {
    require(op && over_integrate >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    op->op->eforms.push_back({EF_LAPLACE, field, value, over_integrate, scaling});
    forms_changed(*op->op);
  });
}

hgdt_status hgdt_operator_append_product(hgdt_operator op, const double* field, double value, int32_t over_integrate,
                                         double scaling)
{
  return guarded([&] {
    require(op && over_integrate >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    op->op->eforms.push_back({EF_PRODUCT, field, value, over_integrate, scaling});
    forms_changed(*op->op);
  });
}

static void append_element_points(hgdt_operator op, int kind, const double* values, int32_t coeff_order,
                                  int32_t over_integrate, double scaling)
{
  require(op && values && coeff_order >= 0 && over_integrate >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
  ElementFormDesc F{kind, nullptr, 0., over_integrate, scaling};
  F.points = values, F.coeff_order = coeff_order;
  op->op->eforms.push_back(F);
  forms_changed(*op->op);
}

hgdt_status hgdt_operator_append_laplace_tensor(hgdt_operator op, const double* tensor_field,
                                                const double* tensor_value, int32_t over_integrate, double scaling)
{
  return guarded([&] {
    require(op && (tensor_field || tensor_value) && over_integrate >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    const Grid& g = *op->op->space->grid;
    const int d = g.ref.dim;
    ElementFormDesc F{EF_LAPLACE, nullptr, 0., over_integrate, scaling};
    if (tensor_field) {
      require(tensor_field_symmetric(tensor_field, g.ne, d), HGDT_ERR_NOT_SUPPORTED,
              "the diffusion tensor must be symmetric in every element (the local matrices are symmetric)");
      F.tensor_kind = 1, F.tensor = tensor_field;
    } else {
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b)
          require(tensor_value[a * d + b] == tensor_value[b * d + a], HGDT_ERR_NOT_SUPPORTED,
                  "the diffusion tensor must be symmetric (the local matrices are symmetric)");
      F.tensor_kind = 2;
      for (int t = 0; t < d * d; ++t)
        F.tensor_value[t] = tensor_value[t];
    }
    op->op->eforms.push_back(F);
    forms_changed(*op->op);
  });
}

hgdt_status hgdt_operator_append_laplace_points(hgdt_operator op, const double* values, int32_t coeff_order,
                                                int32_t over_integrate, double scaling)
{
  return guarded([&] { append_element_points(op, EF_LAPLACE, values, coeff_order, over_integrate, scaling); });
}

hgdt_status hgdt_operator_append_product_points(hgdt_operator op, const double* values, int32_t coeff_order,
                                                int32_t over_integrate, double scaling)
{
  return guarded([&] { append_element_points(op, EF_PRODUCT, values, coeff_order, over_integrate, scaling); });
}

static void append_face(hgdt_operator op, int kind, double param, const double* kappa, double kv, const double* omega,
                        double ov, int over_integrate, double scaling)
{
  require(op && over_integrate >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
  FaceFormDesc F{kind, param, kappa, kv, omega, ov, over_integrate, scaling, nullptr};
  if (kind == FF_INNER_COUPLING || kind == FF_INNER_PENALTY) {
    if (op->op->stencil == HGDT_STENCIL_ELEMENT)
      HGDT_THROW(HGDT_ERR_OPERATOR, "coupling intersection forms need an intersection stencil");
    op->op->cforms.push_back(F);
  } else {
    F.mask = op->op->pending_mask;
    op->op->bforms.push_back(F);
  }
  forms_changed(*op->op);
}

hgdt_status hgdt_operator_set_boundary_source_tensor(hgdt_operator op, const double* tensor)
{
  return guarded([&] {
    require(op && tensor, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    auto& bs = op->op->bsources;
    require(!bs.empty(), HGDT_ERR_INVALID_ARGUMENT, "no boundary functional appended");
    require(bs.back().kind != HGDT_BOUNDARY_PRODUCT, HGDT_ERR_NOT_SUPPORTED,
            "the boundary product term takes a scalar weight");
    const Grid& g = *op->op->space->grid;
    require(tensor_field_symmetric(tensor, g.ne, g.ref.dim), HGDT_ERR_NOT_SUPPORTED,
            "the tensor must be symmetric in every element");
    bs.back().coeff_t = tensor;
    forms_changed(*op->op);
  });
}

hgdt_status hgdt_operator_set_face_tensors(hgdt_operator op, int32_t boundary, const double* diffusion_tensor,
                                           const double* weight_tensor)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    auto& forms = boundary ? op->op->bforms : op->op->cforms;
    require(!forms.empty(), HGDT_ERR_INVALID_ARGUMENT, "no intersection form of that kind appended");
    const Grid& g = *op->op->space->grid;
    for (const double* t : {diffusion_tensor, weight_tensor})
      require(!t || tensor_field_symmetric(t, g.ne, g.ref.dim), HGDT_ERR_NOT_SUPPORTED,
              "the diffusion and weight tensors must be symmetric in every element");
    if (diffusion_tensor)
      forms.back().kappa_t = diffusion_tensor;
    if (weight_tensor)
      forms.back().omega_t = weight_tensor;
    forms_changed(*op->op);
  });
}

hgdt_status hgdt_operator_append_inner_coupling(hgdt_operator op, double s, const double* kappa, double kv,
                                                const double* omega, double ov, int32_t over_integrate, double scaling)
{
  return guarded([&] { append_face(op, FF_INNER_COUPLING, s, kappa, kv, omega, ov, over_integrate, scaling); });
}

hgdt_status hgdt_operator_append_inner_penalty(hgdt_operator op, double penalty, const double* omega, double ov,
                                               int32_t over_integrate, double scaling)
{
  return guarded([&] { append_face(op, FF_INNER_PENALTY, penalty, nullptr, 1., omega, ov, over_integrate, scaling); });
}

hgdt_status hgdt_operator_append_dirichlet_coupling(hgdt_operator op, double s, const double* kappa, double kv,
                                                    int32_t over_integrate, double scaling)
{
  return guarded([&] { append_face(op, FF_DIRICHLET_COUPLING, s, kappa, kv, nullptr, 1., over_integrate, scaling); });
}

hgdt_status hgdt_operator_append_boundary_penalty(hgdt_operator op, double penalty, const double* omega, double ov,
                                                  int32_t over_integrate, double scaling)
{
  return guarded(
      [&] { append_face(op, FF_BOUNDARY_PENALTY, penalty, nullptr, 1., omega, ov, over_integrate, scaling); });
}

hgdt_status hgdt_operator_append_source(hgdt_operator op, int kind, const double* data, double value, int32_t f_order,
                                        int32_t over_integrate)
{
  return guarded([&] {
    require(op && f_order >= 0 && over_integrate >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    require(kind >= HGDT_SOURCE_CONSTANT && kind <= HGDT_SOURCE_PER_QUADRATURE_POINT, HGDT_ERR_INVALID_ARGUMENT,
            "unknown source kind");
    require(kind == HGDT_SOURCE_CONSTANT || data, HGDT_ERR_INVALID_ARGUMENT, "source data is NULL");
    op->op->sources.push_back({kind, data, value, f_order, over_integrate});
    op->op->maps.built = false; // staging layout gains the right-hand side block
    forms_changed(*op->op);
  });
}

hgdt_status hgdt_operator_append_boundary_source(hgdt_operator op, int kind, double param, const double* coeff_field,
                                                 double coeff_value, const double* data, double value, int32_t g_order,
                                                 int32_t over_integrate)
{
  return guarded([&] {
    require(op && g_order >= 0 && over_integrate >= 0, HGDT_ERR_INVALID_ARGUMENT, "invalid argument");
    require(kind >= HGDT_BOUNDARY_DIRICHLET_COUPLING && kind <= HGDT_BOUNDARY_PRODUCT, HGDT_ERR_INVALID_ARGUMENT,
            "unknown boundary functional kind");
    op->op->bsources.push_back(
        {kind, param, coeff_field, coeff_value, data, value, g_order, over_integrate, op->op->pending_mask});
    op->op->maps.built = false; // staging layout gains the boundary right-hand side block
    forms_changed(*op->op);
  });
}

static void check_boundary_mask(const Operator& o, const uint8_t* mask, int64_t count)
{
  require(!mask || count == o.space->grid->n_boundary, HGDT_ERR_SHAPES_DO_NOT_MATCH,
          "the boundary mask needs one entry per boundary intersection");
}

hgdt_status hgdt_operator_set_boundary_filter(hgdt_operator op, const uint8_t* mask, int64_t count)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    Operator& o = *op->op;
    check_boundary_mask(o, mask, count);
    if (!mask) {
      o.pending_mask = nullptr;
      return;
    }
    auto d = std::make_unique<DeviceArray<uint8_t>>();
    d->upload(std::vector<uint8_t>(mask, mask + count));
    o.pending_mask = d->ptr;
    o.face_masks.push_back(std::move(d));
  });
}

hgdt_status hgdt_operator_set_dirichlet_faces(hgdt_operator op, const uint8_t* mask, int64_t count)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    Operator& o = *op->op;
    check_boundary_mask(o, mask, count);
    if (mask)
      o.dirichlet_faces.assign(mask, mask + count);
    else
      o.dirichlet_faces.clear();
    o.dirichlet_built = false;
    forms_changed(o);
  });
}

hgdt_status hgdt_operator_set_dirichlet(hgdt_operator op, int32_t enabled)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    require(!enabled || op->op->square(), HGDT_ERR_NOT_SUPPORTED,
            "Dirichlet constraints need an operator whose source and range spaces are the same");
    op->op->dirichlet = enabled != 0;
    forms_changed(*op->op);
  });
}

hgdt_status hgdt_operator_set_dirichlet_mode(hgdt_operator op, int32_t only_clear, int32_t ensure_symmetry)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    Operator& o = *op->op;
    if (o.dirichlet_only_clear != (only_clear != 0) || o.dirichlet_ensure_symmetry != (ensure_symmetry != 0))
      o.dirichlet_built = false; // the slot flags follow the mode
    o.dirichlet_only_clear = only_clear != 0;
    o.dirichlet_ensure_symmetry = ensure_symmetry != 0;
    forms_changed(o);
  });
}

hgdt_status hgdt_operator_dirichlet_dofs(hgdt_operator op, int64_t* count, int32_t* host_out)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    op->op->build_dirichlet();
    if (count)
      *count = int64_t(op->op->dirichlet_dofs.size());
    if (host_out)
      std::copy(op->op->dirichlet_dofs.begin(), op->op->dirichlet_dofs.end(), host_out);
  });
}

hgdt_status hgdt_operator_assemble(hgdt_operator op, void* stream)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    op->op->assemble(static_cast<hipStream_t>(stream));
  });
}

hgdt_status hgdt_operator_values_device(hgdt_operator op, double** values)
{
  return guarded([&] {
    require(op && values, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    require_matrix(*op->op);
    materialize(*op->op);
    *values = op->op->d_values.ptr;
  });
}

hgdt_status hgdt_operator_rhs_device(hgdt_operator op, double** rhs)
{
  return guarded([&] {
    require(op && rhs, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    *rhs = op->op->d_rhs.ptr;
  });
}

hgdt_status hgdt_operator_copy_values(hgdt_operator op, double* host_out)
{
  return guarded([&] {
    require(op && host_out, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    require_matrix(*op->op);
    materialize(*op->op);
    HGDT_HIP(hipDeviceSynchronize());
    HGDT_HIP(hipMemcpy(host_out, op->op->d_values.ptr, op->op->nnz * sizeof(double), hipMemcpyDeviceToHost));
  });
}

hgdt_status hgdt_operator_copy_rhs(hgdt_operator op, double* host_out)
{
  return guarded([&] {
    require(op && host_out, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    HGDT_HIP(hipDeviceSynchronize());
    HGDT_HIP(hipMemcpy(host_out, op->op->d_rhs.ptr, op->op->n * sizeof(double), hipMemcpyDeviceToHost));
  });
}

hgdt_status hgdt_operator_assembly_info(hgdt_operator op, int32_t* structured, int32_t* solver_upper_diagonals)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    if (structured)
      *structured = op->op->direct_storage ? 1 : 0;
    if (solver_upper_diagonals)
      *solver_upper_diagonals = op->op->solver_upper_diagonals();
  });
}

hgdt_status hgdt_operator_assembly_bytes(hgdt_operator op, int64_t* bytes)
{
  return guarded([&] {
    require(op && bytes, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    *bytes = op->op->assembly_bytes;
  });
}

hgdt_status hgdt_operator_set_format(hgdt_operator op, int format)
{
  return guarded([&] {
    require(op && format >= HGDT_FORMAT_AUTO && format <= HGDT_FORMAT_BLOCK, HGDT_ERR_INVALID_ARGUMENT,
            "unknown format");
    // the CSR values of a structured assembly are computed from its storage before that storage is rebuilt
    materialize(*op->op);
    op->op->direct_storage = false;
    op->op->direct_state = -1;
    op->op->requested_format = (format == HGDT_FORMAT_DIA_NONSYMMETRIC) ? HGDT_FORMAT_DIA : format;
    op->op->allow_symmetric = (format != HGDT_FORMAT_DIA_NONSYMMETRIC);
    op->op->fmt.partitions_ready = false;
    op->op->fmt.structure_ready = false;
    op->op->fmt.valid = false;
  });
}

hgdt_status hgdt_operator_format_info(hgdt_operator op, int32_t* format, int64_t* spmv_bytes, int32_t* num_diagonals)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    require(op->op->fmt.structure_ready, HGDT_ERR_OPERATOR, "no apply / apply_inverse since the format was chosen");
    if (format)
      *format = (op->op->fmt.format == HGDT_FORMAT_DIA && !op->op->fmt.symmetric) ? HGDT_FORMAT_DIA_NONSYMMETRIC
                                                                                   : op->op->fmt.format;
    if (spmv_bytes)
      *spmv_bytes = op->op->fmt.bytes;
    if (num_diagonals)
      *num_diagonals = op->op->fmt.ndiag;
  });
}

hgdt_status hgdt_operator_apply(hgdt_operator op, const double* x, double* y, void* stream)
{
  return guarded([&] {
    require(op && x && y, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    op->op->apply(x, y, static_cast<hipStream_t>(stream));
  });
}

hgdt_status hgdt_solver_options_default(hgdt_solver_options* o)
{
  return guarded([&] {
    require(o, HGDT_ERR_INVALID_ARGUMENT, "NULL options");
    o->rtol = 1e-10;
    o->max_iter = 100000;
    o->check_every = 16;
    o->use_graph = 0;
    o->time_spmv = 0;
    o->zero_initial_guess = 1;
    o->true_residual = 1;
    o->restarts = 0;
  });
}

hgdt_status hgdt_operator_apply_inverse(hgdt_operator op, const double* b, double* x, const hgdt_solver_options* options,
                                        hgdt_solver_stats* stats, void* stream)
{
  return guarded([&] {
    require(op && b && x, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    hgdt_solver_options o;
    if (options)
      o = *options;
    else
      hgdt_solver_options_default(&o);
    require(o.restarts >= 0 && (o.restarts == 0 || o.true_residual), HGDT_ERR_INVALID_ARGUMENT,
            "restarts need the true residual (true_residual = 1)");
    hgdt_solver_stats st{};
    Operator& O = *op->op;
    try {
      O.apply_inverse(b, x, o, st, static_cast<hipStream_t>(stream));
      // restarts from the true residual: CG again from the returned x while an instance misses rtol (the
      // instances that meet it stop at the first stopping test); iterations add up per instance
      for (int k = 0; k < o.restarts && st.true_relres > o.rtol; ++k) {
        hgdt_solver_options r = o;
        r.zero_initial_guess = 0;
        const std::vector<int64_t> before = O.last_iters;
        const double ms = st.solve_ms;
        const int64_t iters = st.iterations;
        st = hgdt_solver_stats{};
        O.apply_inverse(b, x, r, st, static_cast<hipStream_t>(stream));
        for (size_t i = 0; i < before.size() && i < O.last_iters.size(); ++i)
          O.last_iters[i] += before[i];
        st.iterations += iters;
        st.solve_ms += ms;
        st.restarts = k + 1;
      }
    } catch (...) {
      if (stats)
        *stats = st;
      throw;
    }
    if (stats)
      *stats = st;
  });
}

hgdt_status hgdt_operator_instance_stats(hgdt_operator op, int64_t* iterations, double* relres)
{
  return guarded([&] {
    require(op, HGDT_ERR_INVALID_ARGUMENT, "NULL operator");
    if (iterations)
      std::copy(op->op->last_iters.begin(), op->op->last_iters.end(), iterations);
    if (relres)
      std::copy(op->op->last_relres.begin(), op->op->last_relres.end(), relres);
  });
}

hgdt_status hgdt_operator_instance_true_residuals(hgdt_operator op, double* relres)
{
  return guarded([&] {
    require(op && relres, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    const auto& v = op->op->last_true_relres;
    if (v.empty())
      std::fill(relres, relres + (op->op->space->inst_dof.size() - 1), -1.);
    else
      std::copy(v.begin(), v.end(), relres);
  });
}

hgdt_status hgdt_operator_adjoint_log_diffusion_gradient(hgdt_operator op, int32_t form_index, const double* u,
                                                         const double* lambda, double* grad, void* stream)
{
  return guarded([&] {
    require(op && u && lambda && grad, HGDT_ERR_INVALID_ARGUMENT, "NULL argument");
    op->op->adjoint_gradient(form_index, u, lambda, grad, static_cast<hipStream_t>(stream));
  });
}

hgdt_status hgdt_operator_set_launch(hgdt_operator op, int32_t kernel, int32_t block_size, int32_t items_per_thread)
{
  return guarded([&] {
    require(op && kernel >= 0 && kernel < 5, HGDT_ERR_INVALID_ARGUMENT, "unknown kernel id");
    const int max_items = (kernel == 3) ? 16 : 8;
    require(block_size >= 64 && block_size % 64 == 0 && items_per_thread >= 1 && items_per_thread <= max_items &&
                (items_per_thread & (items_per_thread - 1)) == 0,
            HGDT_ERR_INVALID_ARGUMENT,
            "block_size must be a multiple of 64, items_per_thread a power of two up to 8 (16 for kernel 3)");
    const int limit = (kernel <= 2) ? 256 : 1024; // assembly kernels are compiled for at most 256 lanes
    require(block_size <= limit, HGDT_ERR_INVALID_ARGUMENT, "block_size above the kernel's launch bound");
    Launch& L = op->op->launch[kernel];
    if ((kernel == 3 || kernel == 4) && (L.block != block_size || L.items != items_per_thread)) {
      op->op->fmt.partitions_ready = false;
      if (kernel == 3 && op->op->fmt.format == HGDT_FORMAT_SELL) { // SELL slices follow the A partition
        op->op->fmt.structure_ready = false;
        op->op->fmt.valid = false;
      }
    }
    L.block = block_size;
    L.items = (kernel <= 2) ? 1 : items_per_thread;
    L.user = true;
  });
}

hgdt_status hgdt_operator_set_column_staging(hgdt_operator op, int32_t mode)
{
  return guarded([&] {
    require(op && mode >= -1 && mode <= 1, HGDT_ERR_INVALID_ARGUMENT, "column staging mode must be -1, 0 or 1");
    if (op->op->column_staging != mode)
      op->op->fmt.partitions_ready = false; // rebuilt before the next solve (captured graphs too)
    op->op->column_staging = mode;
  });
}

hgdt_status hgdt_operator_get_launch(hgdt_operator op, int32_t kernel, int32_t* block_size, int32_t* items_per_thread)
{
  return guarded([&] {
    require(op && kernel >= 0 && kernel < 5, HGDT_ERR_INVALID_ARGUMENT, "unknown kernel id");
    if (block_size)
      *block_size = op->op->launch[kernel].block;
    if (items_per_thread)
      *items_per_thread = op->op->launch[kernel].items;
  });
}

} // extern "C"
