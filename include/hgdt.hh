// hgdt.hh -- header-only C++ classes over the C ABI of libhgdt.so (include/hgdt.h).
//
// The classes carry the names, argument meaning and defaults of dune-gdt's assembly + solve path so
// that code written against the reference reads the same (paths relative to the dune-gdt tree):
//
//   make_cube_grid / Grid                      XT::Grid::make_cube_grid (examples/stationary-heat-equation.cc:87)
//   ContinuousLagrangeSpace(grid, order)       spaces/h1/continuous-lagrange.hh:68-79
//   DiscontinuousLagrangeSpace(grid, order)    spaces/l2/discontinuous-lagrange.hh:78-86
//   LocalLaplaceIntegrand(diffusion)           local/integrands/laplace.hh:40-48
//   LocalElementProductIntegrand(weight)       local/integrands/product.hh:56-65 (with_ansatz: interfaces.hh:225-229)
//   LocalLaplaceIPDGIntegrands::InnerCoupling / DirichletCoupling   local/integrands/laplace-ipdg.hh:47-61,237-252
//   LocalIPDGIntegrands::InnerPenalty / BoundaryPenalty             local/integrands/ipdg.hh:59-72,201-213
//   LocalElementIntegralBilinearForm, LocalCouplingIntersectionIntegralBilinearForm,
//   LocalIntersectionIntegralBilinearForm      local/bilinear-forms/integrals.hh:52-63,169-180,305-316
//   LocalElementIntegralFunctional             local/functionals/integrals.hh:40-97
//   make_vector_functional / VectorFunctional  functionals/vector-based.hh:132-308
//   make_dirichlet_constraints                 tools/dirichlet-constraints.hh:214-220
//   MatrixOperator / make_matrix_operator      operators/matrix-based.hh:280-500,650-658
//   make_discrete_function, make_bilinear_form(...).apply2()   operators/bilinear-form.hh:344-452
//
// Template parameters of the reference (grid, matrix and vector types, range dimension) become runtime
// values: the geometry (quad, tri, hex, tet) and the space kind.  Vectors live in HIP device memory
// (DeviceVector); every computation runs in libhgdt.so on the GPU.  Failures are thrown as exceptions
// named after the reference's exception families (exceptions.hh:24-77).  Link with -lhgdt; no HIP
// headers are needed.
#ifndef HGDT_HH
#define HGDT_HH

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "hgdt.h"

namespace hgdt {

// ---------------------------------------------------------------------------------------------
// exceptions
// ---------------------------------------------------------------------------------------------
class Exception : public std::runtime_error
{
public:
  Exception(hgdt_status status, const std::string& msg)
    : std::runtime_error(msg)
    , status_(status)
  {}
  hgdt_status status() const
  {
    return status_;
  }

private:
  hgdt_status status_;
};

namespace Exceptions {
#define HGDT_EXCEPTION(name)                                                                                           \
  struct name : public Exception                                                                                       \
  {                                                                                                                    \
    using Exception::Exception;                                                                                        \
  };
HGDT_EXCEPTION(wrong_input_given)
HGDT_EXCEPTION(shapes_do_not_match)
HGDT_EXCEPTION(operator_error)
HGDT_EXCEPTION(integrand_error)
HGDT_EXCEPTION(space_error)
HGDT_EXCEPTION(mapper_error)
HGDT_EXCEPTION(linear_solver_failed)
HGDT_EXCEPTION(device_error)
HGDT_EXCEPTION(not_supported)
HGDT_EXCEPTION(out_of_memory)
#undef HGDT_EXCEPTION
} // namespace Exceptions

inline void check(hgdt_status s)
{
  if (s == HGDT_OK)
    return;
  const std::string msg = hgdt_last_error();
  switch (s) {
    case HGDT_ERR_INVALID_ARGUMENT:
      throw Exceptions::wrong_input_given(s, msg);
    case HGDT_ERR_SHAPES_DO_NOT_MATCH:
      throw Exceptions::shapes_do_not_match(s, msg);
    case HGDT_ERR_OPERATOR:
      throw Exceptions::operator_error(s, msg);
    case HGDT_ERR_INTEGRAND:
      throw Exceptions::integrand_error(s, msg);
    case HGDT_ERR_SPACE:
      throw Exceptions::space_error(s, msg);
    case HGDT_ERR_MAPPER:
      throw Exceptions::mapper_error(s, msg);
    case HGDT_ERR_SOLVER_FAILED:
      throw Exceptions::linear_solver_failed(s, msg);
    case HGDT_ERR_DEVICE:
      throw Exceptions::device_error(s, msg);
    case HGDT_ERR_NOT_SUPPORTED:
      throw Exceptions::not_supported(s, msg);
    case HGDT_ERR_OUT_OF_MEMORY:
      throw Exceptions::out_of_memory(s, msg);
    default:
      throw Exception(s, msg);
  }
}

// ---------------------------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------------------------
// FP64 vector in HIP device memory (the role of XT::LA::IstlDenseVector on this path)
class DeviceVector
{
public:
  DeviceVector() = default;
  explicit DeviceVector(int64_t n, double value = 0.)
  {
    resize(n);
    if (n > 0)
      from_host(std::vector<double>(size_t(n), value));
  }
  explicit DeviceVector(const std::vector<double>& host)
  {
    resize(int64_t(host.size()));
    from_host(host);
  }
  DeviceVector(const DeviceVector&) = delete;
  DeviceVector& operator=(const DeviceVector&) = delete;
  DeviceVector(DeviceVector&& o) noexcept
    : ptr_(o.ptr_)
    , n_(o.n_)
  {
    o.ptr_ = nullptr, o.n_ = 0;
  }
  DeviceVector& operator=(DeviceVector&& o) noexcept
  {
    std::swap(ptr_, o.ptr_), std::swap(n_, o.n_);
    return *this;
  }
  ~DeviceVector()
  {
    if (ptr_)
      hgdt_device_free(ptr_);
  }
  void resize(int64_t n)
  {
    if (n == n_)
      return;
    if (ptr_)
      check(hgdt_device_free(ptr_));
    ptr_ = nullptr, n_ = 0;
    void* p = nullptr;
    check(hgdt_device_alloc(n * int64_t(sizeof(double)), &p));
    ptr_ = static_cast<double*>(p), n_ = n;
  }
  int64_t size() const
  {
    return n_;
  }
  double* data()
  {
    return ptr_;
  }
  const double* data() const
  {
    return ptr_;
  }
  void from_host(const std::vector<double>& host)
  {
    if (int64_t(host.size()) != n_)
      throw Exceptions::shapes_do_not_match(HGDT_ERR_SHAPES_DO_NOT_MATCH, "host vector size differs");
    if (n_)
      check(hgdt_memcpy(ptr_, host.data(), n_ * int64_t(sizeof(double)), nullptr));
    check(hgdt_stream_synchronize(nullptr));
  }
  std::vector<double> to_host() const
  {
    std::vector<double> out(static_cast<size_t>(n_));
    check(hgdt_stream_synchronize(nullptr));
    if (n_)
      check(hgdt_memcpy(out.data(), ptr_, n_ * int64_t(sizeof(double)), nullptr));
    check(hgdt_stream_synchronize(nullptr));
    return out;
  }

private:
  double* ptr_ = nullptr;
  int64_t n_ = 0;
};

// ---------------------------------------------------------------------------------------------
// grids
// ---------------------------------------------------------------------------------------------
enum class Geometry
{
  quad = HGDT_QUAD,
  tri = HGDT_TRI,
  hex = HGDT_HEX,
  tet = HGDT_TET
};

class Grid
{
public:
  explicit Grid(hgdt_grid h)
    : h_(h, [](hgdt_grid g) { hgdt_grid_destroy(g); })
  {
    int64_t info[9];
    check(hgdt_grid_info(h, info));
    dim_ = int(info[0]), nv_ = info[1], ne_ = info[2], ninst_ = int(info[7]);
  }
  static Grid cube(Geometry geometry, const std::vector<double>& lower_left, const std::vector<double>& upper_right,
                   const std::vector<int32_t>& num_elements)
  {
    const size_t d = (geometry == Geometry::quad || geometry == Geometry::tri) ? 2 : 3;
    if (lower_left.size() != d || upper_right.size() != d || num_elements.size() != d)
      throw Exceptions::shapes_do_not_match(HGDT_ERR_SHAPES_DO_NOT_MATCH, "cube grid arguments of wrong dimension");
    hgdt_grid g = nullptr;
    check(hgdt_grid_create_structured(int(geometry), num_elements.data(), lower_left.data(), upper_right.data(), &g));
    return Grid(g);
  }
  // triangles and quadrilaterals, element by element as GridFactory<G>::insertElement(type, vertices) in
  // make_mixed_grid (dune/gdt/test/spaces/base.hh:355-376): coords [num_vertices][2], elements[e] the vertices of
  // element e in reference order (3: a triangle, 4: a quadrilateral)
  static Grid mixed(const std::vector<double>& coords, const std::vector<std::vector<int32_t>>& elements)
  {
    if (coords.size() % 2)
      throw Exceptions::shapes_do_not_match(HGDT_ERR_SHAPES_DO_NOT_MATCH, "mixed grids are two-dimensional");
    std::vector<int32_t> geometry, vertices;
    std::vector<int64_t> offsets{0};
    for (const auto& el : elements) {
      if (el.size() != 3 && el.size() != 4)
        throw Exceptions::wrong_input_given(HGDT_ERR_INVALID_ARGUMENT, "an element needs 3 or 4 vertices");
      geometry.push_back(el.size() == 4 ? HGDT_QUAD : HGDT_TRI);
      vertices.insert(vertices.end(), el.begin(), el.end());
      offsets.push_back(int64_t(vertices.size()));
    }
    hgdt_grid g = nullptr;
    check(hgdt_grid_create_mixed(int64_t(coords.size() / 2), coords.data(), int64_t(elements.size()),
                                 geometry.data(), offsets.data(), vertices.data(), &g));
    return Grid(g);
  }
  // disjoint union of independent grids: the batch of problem instances
  static Grid union_of(const std::vector<Grid>& grids)
  {
    std::vector<hgdt_grid> hs;
    for (const auto& g : grids)
      hs.push_back(g.handle());
    hgdt_grid u = nullptr;
    check(hgdt_grid_create_union(hs.data(), int32_t(hs.size()), &u));
    return Grid(u);
  }
  // the leaf view of the reference's grid providers; grids here are their own views
  const Grid& leaf_view() const
  {
    return *this;
  }
  int dimension() const
  {
    return dim_;
  }
  int64_t num_vertices() const
  {
    return nv_;
  }
  int64_t num_elements() const
  {
    return ne_;
  }
  int num_instances() const
  {
    return ninst_;
  }
  // physical coordinates of the element quadrature points of `order`, [num_elements][num_points][dim] (mixed
  // grids: num_points 0, element e's points from quadrature_offsets(order)[e])
  std::vector<double> quadrature_points(int order, int32_t& num_points) const
  {
    const std::vector<int64_t> off = quadrature_offsets(order);
    std::vector<double> out(size_t(off.back()) * size_t(dim_));
    check(hgdt_grid_quadrature_points(handle(), order, out.data(), &num_points));
    return out;
  }
  std::vector<int64_t> quadrature_offsets(int order) const
  {
    std::vector<int64_t> off(size_t(ne_) + 1);
    check(hgdt_grid_quadrature_offsets(handle(), order, off.data()));
    return off;
  }
  int64_t num_boundary_intersections() const
  {
    int64_t info[9];
    check(hgdt_grid_info(handle(), info));
    return info[6];
  }
  // centres and outward unit normals of the boundary intersections in walk order, [num][dim] each
  void boundary_intersections(std::vector<double>& centers, std::vector<double>& normals) const
  {
    const size_t nb = size_t(num_boundary_intersections());
    centers.resize(nb * size_t(dim_)), normals.resize(nb * size_t(dim_));
    check(hgdt_grid_boundary_intersections(handle(), centers.data(), normals.data()));
  }
  hgdt_grid handle() const
  {
    return h_.get();
  }

private:
  std::shared_ptr<hgdt_grid_s> h_;
  int dim_ = 0, ninst_ = 1;
  int64_t nv_ = 0, ne_ = 0;
};

// XT::Grid::make_cube_grid<G>(lower_left, upper_right, num_elements): one number for every direction
inline Grid make_cube_grid(Geometry geometry, double lower_left, double upper_right, int32_t num_elements)
{
  const size_t d = (geometry == Geometry::quad || geometry == Geometry::tri) ? 2 : 3;
  return Grid::cube(geometry, std::vector<double>(d, lower_left), std::vector<double>(d, upper_right),
                    std::vector<int32_t>(d, num_elements));
}
inline Grid make_cube_grid(Geometry geometry, const std::vector<double>& lower_left,
                           const std::vector<double>& upper_right, const std::vector<int32_t>& num_elements)
{
  return Grid::cube(geometry, lower_left, upper_right, num_elements);
}

// ---------------------------------------------------------------------------------------------
// functions
// ---------------------------------------------------------------------------------------------
// piecewise constant scalar (times the identity where a matrix is expected), or a constant
class GridFunction
{
public:
  GridFunction(double value = 1.)
    : value_(value)
  {}
  GridFunction(int value)
    : value_(double(value))
  {}
  // one value per element, copied to the device
  GridFunction(const std::vector<double>& per_element)
    : field_(std::make_shared<DeviceVector>(per_element))
  {}
  // one value per element, already on the device (shared, kept alive by every user)
  GridFunction(std::shared_ptr<DeviceVector> per_element)
    : field_(std::move(per_element))
  {}
  const double* field() const
  {
    return field_ ? field_->data() : nullptr;
  }
  double value() const
  {
    return value_;
  }
  const std::shared_ptr<DeviceVector>& storage() const
  {
    return field_;
  }
  int order() const
  {
    return 0;
  }

private:
  double value_ = 0.;
  std::shared_ptr<DeviceVector> field_;
};

// matrix-valued GridFunction<E, d, d> (the diffusion of LocalLaplaceIntegrand, laplace.hh:40-48): one symmetric
// d x d matrix (row-major) for every element, or one per element (num_elements * d * d values)
class TensorGridFunction
{
public:
  // d x d values for every element
  explicit TensorGridFunction(std::vector<double> value)
    : value_(std::move(value))
  {}
  // d x d values per element, copied to the device (per_element must be true)
  TensorGridFunction(const std::vector<double>& values, bool per_element)
    : field_(std::make_shared<DeviceVector>(values))
  {
    if (!per_element)
      value_ = values, field_.reset();
  }
  const double* field() const
  {
    return field_ ? field_->data() : nullptr;
  }
  const double* value() const
  {
    return value_.empty() ? nullptr : value_.data();
  }
  const std::shared_ptr<DeviceVector>& storage() const
  {
    return field_;
  }

private:
  std::vector<double> value_;
  std::shared_ptr<DeviceVector> field_;
};

// XT::Functions::GenericFunction<d>(order, evaluate, name, parameter_type, jacobian): analytic function of
// the physical coordinates, evaluated on the host at quadrature points
class GenericFunction
{
public:
  using Evaluate = std::function<double(const double* x)>;
  using Jacobian = std::function<void(const double* x, double* gradient)>;
  GenericFunction(int order, Evaluate evaluate, const std::string& name = "", Jacobian jacobian = nullptr)
    : order_(order)
    , evaluate_(std::move(evaluate))
    , jacobian_(std::move(jacobian))
    , name_(name)
  {}
  int order() const
  {
    return order_;
  }
  double evaluate(const double* x) const
  {
    return evaluate_(x);
  }
  std::vector<double> values_at(const Grid& grid, int quadrature_order) const
  {
    int32_t nq = 0;
    const std::vector<double> pts = grid.quadrature_points(quadrature_order, nq);
    const size_t d = size_t(grid.dimension()), np = pts.size() / d;
    std::vector<double> out(np);
    for (size_t p = 0; p < np; ++p)
      out[p] = evaluate_(&pts[p * d]);
    return out;
  }
  std::vector<double> gradients_at(const Grid& grid, int quadrature_order) const
  {
    if (!jacobian_)
      throw Exceptions::not_supported(HGDT_ERR_NOT_SUPPORTED, "GenericFunction '" + name_ + "' has no jacobian");
    int32_t nq = 0;
    const std::vector<double> pts = grid.quadrature_points(quadrature_order, nq);
    const size_t d = size_t(grid.dimension()), np = pts.size() / d;
    std::vector<double> out(np * d);
    for (size_t p = 0; p < np; ++p)
      jacobian_(&pts[p * d], &out[p * d]);
    return out;
  }

private:
  int order_;
  Evaluate evaluate_;
  Jacobian jacobian_;
  std::string name_;
};

// ---------------------------------------------------------------------------------------------
// spaces
// ---------------------------------------------------------------------------------------------
class SpaceInterface
{
public:
  SpaceInterface(const Grid& grid, int kind, int order)
    : grid_(grid)
    , kind_(kind)
    , order_(order)
  {
    hgdt_space s = nullptr;
    check(hgdt_space_create(grid.handle(), kind, order, &s));
    h_ = std::shared_ptr<hgdt_space_s>(s, [](hgdt_space p) { hgdt_space_destroy(p); });
    int32_t ls = 0;
    check(hgdt_space_size(s, &size_, &ls));
    local_size_ = ls;
  }
  const Grid& grid_view() const
  {
    return grid_;
  }
  int64_t size() const
  {
    return size_;
  }
  int max_local_size() const
  {
    return local_size_;
  }
  int order() const
  {
    return order_;
  }
  bool continuous() const
  {
    return kind_ == HGDT_CONTINUOUS_LAGRANGE;
  }
  // mapper().global_indices(element) for every element
  std::vector<int32_t> global_indices() const
  {
    std::vector<int32_t> out(size_t(grid_.num_elements()) * size_t(local_size_));
    check(hgdt_space_copy_global_indices(handle(), out.data()));
    return out;
  }
  hgdt_space handle() const
  {
    return h_.get();
  }

private:
  Grid grid_;
  int kind_, order_, local_size_ = 0;
  int64_t size_ = 0;
  std::shared_ptr<hgdt_space_s> h_;
};

struct ContinuousLagrangeSpace : public SpaceInterface
{
  ContinuousLagrangeSpace(const Grid& grid_view, int order)
    : SpaceInterface(grid_view, HGDT_CONTINUOUS_LAGRANGE, order)
  {}
};
struct DiscontinuousLagrangeSpace : public SpaceInterface
{
  DiscontinuousLagrangeSpace(const Grid& grid_view, int order = 1)
    : SpaceInterface(grid_view, HGDT_DISCONTINUOUS_LAGRANGE, order)
  {}
};
inline ContinuousLagrangeSpace make_continuous_lagrange_space(const Grid& grid_view, int order)
{
  return ContinuousLagrangeSpace(grid_view, order);
}
inline DiscontinuousLagrangeSpace make_discontinuous_lagrange_space(const Grid& grid_view, int order = 1)
{
  return DiscontinuousLagrangeSpace(grid_view, order);
}

// ---------------------------------------------------------------------------------------------
// integrands
// ---------------------------------------------------------------------------------------------
struct LocalLaplaceIntegrand
{
  explicit LocalLaplaceIntegrand(GridFunction diffusion_ = 1.)
    : diffusion(std::move(diffusion_))
  {}
  // a diffusion varying inside the elements, evaluated at the quadrature points (rule order: its order + 2 k)
  explicit LocalLaplaceIntegrand(const GenericFunction& diffusion_)
    : generic(std::make_shared<GenericFunction>(diffusion_))
  {}
  // a matrix-valued diffusion (symmetric)
  explicit LocalLaplaceIntegrand(const TensorGridFunction& diffusion_)
    : tensor(std::make_shared<TensorGridFunction>(diffusion_))
  {}
  GridFunction diffusion;
  std::shared_ptr<GenericFunction> generic;
  std::shared_ptr<TensorGridFunction> tensor;
};

// LocalBinaryToUnaryElementIntegrand(LocalElementProductIntegrand(1), f): the right-hand side integrand
struct LocalProductWithAnsatz
{
  enum Kind
  {
    grid_function,
    generic_function
  } kind;
  GridFunction grid_f;
  std::shared_ptr<GenericFunction> generic_f;
};

struct LocalElementProductIntegrand
{
  explicit LocalElementProductIntegrand(GridFunction weight_ = 1.)
    : weight(std::move(weight_))
  {}
  // a weight varying inside the elements, evaluated at the quadrature points (rule order: its order + 2 k)
  explicit LocalElementProductIntegrand(const GenericFunction& weight_)
    : generic(std::make_shared<GenericFunction>(weight_))
  {}
  LocalProductWithAnsatz with_ansatz(const GridFunction& f) const
  {
    require_unit_weight();
    return LocalProductWithAnsatz{LocalProductWithAnsatz::grid_function, f, nullptr};
  }
  LocalProductWithAnsatz with_ansatz(double f) const
  {
    return with_ansatz(GridFunction(f));
  }
  LocalProductWithAnsatz with_ansatz(const GenericFunction& f) const
  {
    require_unit_weight();
    return LocalProductWithAnsatz{LocalProductWithAnsatz::generic_function, GridFunction(0.),
                                  std::make_shared<GenericFunction>(f)};
  }
  GridFunction weight;
  std::shared_ptr<GenericFunction> generic;

private:
  void require_unit_weight() const
  {
    if (generic || weight.field() || weight.value() != 1.)
      throw Exceptions::not_supported(HGDT_ERR_NOT_SUPPORTED, "right-hand side functionals use the unit weight");
  }
};
using LocalProductIntegrand = LocalElementProductIntegrand;

// one term of an intersection integrand sum (LocalQuaternaryIntersectionIntegrandSum, combined.hh:329-440)
struct IntersectionTerm
{
  enum Kind
  {
    inner_coupling,
    inner_penalty,
    dirichlet_coupling,
    boundary_penalty
  } kind;
  double parameter; // symmetry prefactor or penalty
  GridFunction diffusion, weight;
};
struct IntersectionIntegrandSum
{
  std::vector<IntersectionTerm> terms;
};
inline IntersectionIntegrandSum operator+(IntersectionIntegrandSum a, const IntersectionIntegrandSum& b)
{
  a.terms.insert(a.terms.end(), b.terms.begin(), b.terms.end());
  return a;
}

namespace LocalLaplaceIPDGIntegrands {
struct InnerCoupling : public IntersectionIntegrandSum
{
  InnerCoupling(double symmetry_prefactor, GridFunction diffusion, GridFunction weight = 1.)
  {
    terms.push_back({IntersectionTerm::inner_coupling, symmetry_prefactor, std::move(diffusion), std::move(weight)});
  }
};
// in a LocalIntersectionIntegralBilinearForm the binary integrand; in a LocalIntersectionIntegralFunctional
// the unary one with the Dirichlet data (laplace-ipdg.hh:237-320)
struct DirichletCoupling : public IntersectionIntegrandSum
{
  DirichletCoupling(double symmetry_prefactor, GridFunction diffusion, double dirichlet_data = 0.)
    : data_value(dirichlet_data)
  {
    terms.push_back({IntersectionTerm::dirichlet_coupling, symmetry_prefactor, std::move(diffusion), 1.});
  }
  DirichletCoupling(double symmetry_prefactor, GridFunction diffusion, const GenericFunction& dirichlet_data)
    : data_function(std::make_shared<GenericFunction>(dirichlet_data))
  {
    terms.push_back({IntersectionTerm::dirichlet_coupling, symmetry_prefactor, std::move(diffusion), 1.});
  }
  double data_value = 0.;
  std::shared_ptr<GenericFunction> data_function;
};
} // namespace LocalLaplaceIPDGIntegrands

namespace LocalIPDGIntegrands {
struct InnerPenalty : public IntersectionIntegrandSum
{
  InnerPenalty(double penalty, GridFunction weight = 1.)
  {
    terms.push_back({IntersectionTerm::inner_penalty, penalty, 1., std::move(weight)});
  }
};
struct BoundaryPenalty;
} // namespace LocalIPDGIntegrands

// a binary boundary integrand with a function as ansatz: LocalBinaryToUnaryIntersectionIntegrand
// (local/integrands/conversion.hh:133-221)
struct BoundaryWithAnsatz
{
  int kind; // HGDT_BOUNDARY_PENALTY or HGDT_BOUNDARY_PRODUCT
  double param;
  GridFunction coeff;
  double data_value = 0.;
  std::shared_ptr<GenericFunction> data_function;
};

namespace LocalIPDGIntegrands {
struct BoundaryPenalty : public IntersectionIntegrandSum
{
  BoundaryPenalty(double penalty_, GridFunction weight_ = 1.)
    : penalty(penalty_)
    , weight(weight_)
  {
    terms.push_back({IntersectionTerm::boundary_penalty, penalty_, 1., std::move(weight_)});
  }
  BoundaryWithAnsatz with_ansatz(double g) const
  {
    return {HGDT_BOUNDARY_PENALTY, penalty, weight, g, nullptr};
  }
  BoundaryWithAnsatz with_ansatz(const GenericFunction& g) const
  {
    return {HGDT_BOUNDARY_PENALTY, penalty, weight, 0., std::make_shared<GenericFunction>(g)};
  }
  double penalty;
  GridFunction weight;
};
} // namespace LocalIPDGIntegrands

// (weight psi_i) phi_j on intersections (local/integrands/product.hh:280-380), used with a function as ansatz
struct LocalIntersectionProductIntegrand
{
  explicit LocalIntersectionProductIntegrand(GridFunction weight_ = 1.)
    : weight(std::move(weight_))
  {}
  BoundaryWithAnsatz with_ansatz(double g) const
  {
    return {HGDT_BOUNDARY_PRODUCT, 0., weight, g, nullptr};
  }
  BoundaryWithAnsatz with_ansatz(const GenericFunction& g) const
  {
    return {HGDT_BOUNDARY_PRODUCT, 0., weight, 0., std::make_shared<GenericFunction>(g)};
  }
  GridFunction weight;
};

// ---------------------------------------------------------------------------------------------
// local forms, filters, functionals, constraints
// ---------------------------------------------------------------------------------------------
struct LocalElementIntegralBilinearForm
{
  LocalElementIntegralBilinearForm(const LocalLaplaceIntegrand& integrand, int over_integrate_ = 0)
    : laplace(true)
    , coefficient(integrand.diffusion)
    , generic(integrand.generic)
    , tensor(integrand.tensor)
    , over_integrate(over_integrate_)
  {}
  LocalElementIntegralBilinearForm(const LocalElementProductIntegrand& integrand, int over_integrate_ = 0)
    : laplace(false)
    , coefficient(integrand.weight)
    , generic(integrand.generic)
    , over_integrate(over_integrate_)
  {}
  bool laplace;
  GridFunction coefficient;
  std::shared_ptr<GenericFunction> generic; // a coefficient varying inside the elements
  std::shared_ptr<TensorGridFunction> tensor; // a matrix-valued diffusion (Laplace forms)
  int over_integrate;
};
struct LocalCouplingIntersectionIntegralBilinearForm
{
  LocalCouplingIntersectionIntegralBilinearForm(IntersectionIntegrandSum integrand_, int over_integrate_ = 0)
    : integrand(std::move(integrand_))
    , over_integrate(over_integrate_)
  {}
  IntersectionIntegrandSum integrand;
  int over_integrate;
};
struct LocalIntersectionIntegralBilinearForm
{
  LocalIntersectionIntegralBilinearForm(IntersectionIntegrandSum integrand_, int over_integrate_ = 0)
    : integrand(std::move(integrand_))
    , over_integrate(over_integrate_)
  {}
  IntersectionIntegrandSum integrand;
  int over_integrate;
};
struct LocalElementIntegralFunctional
{
  LocalElementIntegralFunctional(LocalProductWithAnsatz integrand_, int over_integrate_ = 0)
    : integrand(std::move(integrand_))
    , over_integrate(over_integrate_)
  {}
  LocalProductWithAnsatz integrand;
  int over_integrate;
};

// LocalIntersectionIntegralFunctional(integrand, over_integrate) on boundary intersections
// (local/functionals/integrals.hh:108-191)
struct LocalIntersectionIntegralFunctional
{
  LocalIntersectionIntegralFunctional(const LocalLaplaceIPDGIntegrands::DirichletCoupling& integrand,
                                      int over_integrate_ = 0)
    : term{HGDT_BOUNDARY_DIRICHLET_COUPLING, integrand.terms.at(0).parameter, integrand.terms.at(0).diffusion,
           integrand.data_value, integrand.data_function}
    , over_integrate(over_integrate_)
  {}
  LocalIntersectionIntegralFunctional(const BoundaryWithAnsatz& integrand, int over_integrate_ = 0)
    : term(integrand)
    , over_integrate(over_integrate_)
  {}
  BoundaryWithAnsatz term;
  int over_integrate;
};

// XT::Grid boundary types and boundary infos: the type of every boundary intersection of a grid, in walk order
struct NoBoundary
{
  static constexpr int id = 0;
};
struct DirichletBoundary
{
  static constexpr int id = 1;
};
struct NeumannBoundary
{
  static constexpr int id = 2;
};
class BoundaryInfo
{
public:
  virtual ~BoundaryInfo() = default;
  virtual std::vector<int8_t> type_ids(const Grid& grid) const = 0;
  // one byte per boundary intersection: 1 where type(intersection) is the type
  std::vector<uint8_t> mask(const Grid& grid, int type_id) const
  {
    const auto ids = type_ids(grid);
    std::vector<uint8_t> m(ids.size());
    for (size_t i = 0; i < ids.size(); ++i)
      m[i] = ids[i] == type_id ? 1 : 0;
    return m;
  }
};
struct AllDirichletBoundaryInfo : BoundaryInfo
{
  std::vector<int8_t> type_ids(const Grid& grid) const override
  {
    return std::vector<int8_t>(size_t(grid.num_boundary_intersections()), int8_t(DirichletBoundary::id));
  }
};
struct AllNeumannBoundaryInfo : BoundaryInfo
{
  std::vector<int8_t> type_ids(const Grid& grid) const override
  {
    return std::vector<int8_t>(size_t(grid.num_boundary_intersections()), int8_t(NeumannBoundary::id));
  }
};
// XT::Grid::NormalBasedBoundaryInfo: the type registered for an intersection's outward unit normal (within tol),
// else the default
class NormalBasedBoundaryInfo : public BoundaryInfo
{
public:
  explicit NormalBasedBoundaryInfo(double tol = 1e-10, int default_type_id = NoBoundary::id)
    : tol_(tol)
    , default_(default_type_id)
  {}
  template <class Type>
  NormalBasedBoundaryInfo& register_new_normal(std::vector<double> normal, const Type&)
  {
    double len = 0.;
    for (double v : normal)
      len += v * v;
    for (double& v : normal)
      v /= std::sqrt(len);
    normals_.push_back({normal, Type::id});
    return *this;
  }
  std::vector<int8_t> type_ids(const Grid& grid) const override
  {
    std::vector<double> c, n;
    grid.boundary_intersections(c, n);
    const int d = grid.dimension();
    std::vector<int8_t> ids(n.size() / size_t(d), int8_t(default_));
    for (size_t b = 0; b < ids.size(); ++b)
      for (const auto& reg : normals_) {
        double dist = 0.;
        for (int a = 0; a < d; ++a) {
          const double r = a < int(reg.first.size()) ? reg.first[a] : 0.;
          dist += (n[b * d + a] - r) * (n[b * d + a] - r);
        }
        if (std::sqrt(dist) < tol_)
          ids[b] = int8_t(reg.second);
      }
    return ids;
  }

private:
  double tol_;
  int default_;
  std::vector<std::pair<std::vector<double>, int>> normals_;
};

namespace ApplyOn {
struct AllElements
{};
struct InnerIntersectionsOnce
{};
struct BoundaryIntersections
{};
// the boundary intersections with boundary_info.type(intersection) == type
struct CustomBoundaryIntersections
{
  CustomBoundaryIntersections() = default;
  template <class Info, class Type = DirichletBoundary>
  CustomBoundaryIntersections(const Info& info, const Type& = {})
    : boundary_info(std::make_shared<Info>(info))
    , type_id(Type::id)
  {}
  // one byte per boundary intersection, or empty for every boundary intersection
  std::vector<uint8_t> mask(const Grid& grid) const
  {
    if (!boundary_info)
      return {};
    auto m = boundary_info->mask(grid, type_id);
    for (uint8_t v : m)
      if (!v)
        return m;
    return {};
  }
  std::shared_ptr<const BoundaryInfo> boundary_info;
  int type_id = DirichletBoundary::id;
};
} // namespace ApplyOn

enum class Stencil
{
  element = HGDT_STENCIL_ELEMENT,
  intersection = HGDT_STENCIL_INTERSECTION,
  element_and_intersection = HGDT_STENCIL_ELEMENT_AND_INTERSECTION
};

// XT::Common::Parameter, reduced to the one key this path reads
struct Parameter
{
  Parameter() = default;
  Parameter(std::initializer_list<std::pair<const std::string, double>> values)
    : values_(values)
  {}
  double get(const std::string& key, double fallback) const
  {
    const auto it = values_.find(key);
    return it == values_.end() ? fallback : it->second;
  }

private:
  std::map<std::string, double> values_;
};

class MatrixOperator;

// storage of an operator's apply / apply_inverse (HGDT_FORMAT_*)
enum class Format {
  automatic = HGDT_FORMAT_AUTO,
  sell = HGDT_FORMAT_SELL,
  dia = HGDT_FORMAT_DIA,
  dia_nonsymmetric = HGDT_FORMAT_DIA_NONSYMMETRIC,
  matrix_free = HGDT_FORMAT_MATRIX_FREE,
  block = HGDT_FORMAT_BLOCK
};

class VectorFunctional
{
public:
  explicit VectorFunctional(const SpaceInterface& space)
    : space_(space)
  {}
  VectorFunctional& append(const LocalElementIntegralFunctional& local_functional)
  {
    functionals_.push_back(local_functional);
    return *this;
  }
  VectorFunctional& operator+=(const LocalElementIntegralFunctional& local_functional)
  {
    return append(local_functional);
  }
  // on the boundary intersections the filter selects (ApplyOn::CustomBoundaryIntersections; every boundary
  // intersection with ApplyOn::BoundaryIntersections)
  VectorFunctional& append(const LocalIntersectionIntegralFunctional& local_functional, const Parameter&,
                           const ApplyOn::CustomBoundaryIntersections& filter)
  {
    boundary_functionals_.push_back(local_functional);
    boundary_filters_.push_back(filter);
    return *this;
  }
  VectorFunctional& append(const LocalIntersectionIntegralFunctional& local_functional, const Parameter& param,
                           const ApplyOn::BoundaryIntersections&)
  {
    return append(local_functional, param, ApplyOn::CustomBoundaryIntersections{});
  }
  const std::vector<ApplyOn::CustomBoundaryIntersections>& boundary_filters() const
  {
    return boundary_filters_;
  }
  const std::vector<LocalElementIntegralFunctional>& local_functionals() const
  {
    return functionals_;
  }
  const std::vector<LocalIntersectionIntegralFunctional>& boundary_functionals() const
  {
    return boundary_functionals_;
  }
  const SpaceInterface& space() const
  {
    return space_;
  }

private:
  SpaceInterface space_;
  std::vector<LocalElementIntegralFunctional> functionals_;
  std::vector<LocalIntersectionIntegralFunctional> boundary_functionals_;
  std::vector<ApplyOn::CustomBoundaryIntersections> boundary_filters_;
};
inline VectorFunctional make_vector_functional(const SpaceInterface& space)
{
  return VectorFunctional(space);
}

// DirichletConstraints(boundary_info, space), tools/dirichlet-constraints.hh:69-110: the DoFs on the intersections
// with boundary_info.type(intersection) == DirichletBoundary()
struct DirichletConstraints
{
  explicit DirichletConstraints(const SpaceInterface& space_)
    : space(space_)
  {}
  template <class Info>
  DirichletConstraints(const SpaceInterface& space_, const Info& info)
    : space(space_)
    , faces(info, DirichletBoundary())
  {}
  SpaceInterface space;
  ApplyOn::CustomBoundaryIntersections faces;
  // the variant of apply(matrix, vector, only_clear, ensure_symmetry) (tools/dirichlet-constraints.hh:152-184)
  // fused into the assembly: unit (or cleared) rows, and with ensure_symmetry their columns
  bool only_clear = false, ensure_symmetry = true;
};
inline DirichletConstraints make_dirichlet_constraints(const SpaceInterface& space)
{
  return DirichletConstraints(space);
}
template <class Info>
DirichletConstraints make_dirichlet_constraints(const SpaceInterface& space, const Info& boundary_info)
{
  return DirichletConstraints(space, boundary_info);
}

// ---------------------------------------------------------------------------------------------
// the operator
// ---------------------------------------------------------------------------------------------
struct SolverOptions
{
  double precision = 1e-10; // relative residual ||r|| <= precision ||b||
  int64_t max_iter = 100000;
  int check_every = 16;
  bool use_graph = false;
  int restarts = 0; // restarts from the true residual while it misses the precision (no reference counterpart)
};

struct SolverStats
{
  bool converged = false;
  int64_t iterations = 0;
  double relres = 0.;
  double solve_ms = 0.;
  double true_relres = -1.; // |b - A x| / |b| recomputed at exit
};

// assembled CSR matrix of an operator (values in device memory)
class CsrMatrix
{
public:
  CsrMatrix(hgdt_operator op)
    : op_(op)
  {
    check(hgdt_operator_shape(op, &rows_, &cols_, &nnz_));
  }
  int64_t rows() const
  {
    return rows_;
  }
  int64_t cols() const
  {
    return cols_;
  }
  int64_t non_zeros() const
  {
    return nnz_;
  }
  const double* values_device() const
  {
    double* v = nullptr;
    check(hgdt_operator_values_device(op_, &v));
    return v;
  }
  void to_host(std::vector<int64_t>& row_ptr, std::vector<int32_t>& col_idx, std::vector<double>& values) const
  {
    row_ptr.resize(size_t(rows_ + 1)), col_idx.resize(size_t(nnz_)), values.resize(size_t(nnz_));
    check(hgdt_operator_copy_pattern(op_, row_ptr.data(), col_idx.data()));
    check(hgdt_operator_copy_values(op_, values.data()));
  }

private:
  hgdt_operator op_;
  int64_t rows_ = 0, cols_ = 0, nnz_ = 0;
};

// a caller's CSR matrix for MatrixOperator(grid_view, source_space, range_space, matrix): the pattern on the
// host (ascending, unique columns per row), the values on the device (nnz entries, owned by the caller)
struct DeviceCsrMatrix {
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col_idx;
  double* values = nullptr;
};

class MatrixOperator
{
public:
  MatrixOperator(const SpaceInterface& space, Stencil stencil = Stencil::element_and_intersection)
    : space_(space)
  {
    hgdt_operator op = nullptr;
    check(hgdt_operator_create(space.handle(), int(stencil), &op));
    h_ = std::shared_ptr<hgdt_operator_s>(op, [](hgdt_operator p) { hgdt_operator_destroy(p); });
  }
  // MatrixOperator(grid_view, source_space, range_space, pattern), matrix-based.hh:295-311: rows of the range
  // space, columns of the source space (same grid and polynomial order)
  MatrixOperator(const SpaceInterface& source_space, const SpaceInterface& range_space,
                 Stencil stencil = Stencil::element_and_intersection)
    : space_(range_space)
  {
    hgdt_operator op = nullptr;
    check(hgdt_operator_create_between(source_space.handle(), range_space.handle(), int(stencil), &op));
    h_ = std::shared_ptr<hgdt_operator_s>(op, [](hgdt_operator p) { hgdt_operator_destroy(p); });
  }
  // MatrixOperator(grid_view, source_space, range_space, matrix), matrix-based.hh:279-293: assembles into the
  // caller's matrix (its values must outlive the operator), adding to the values it holds
  MatrixOperator(const SpaceInterface& source_space, const SpaceInterface& range_space, DeviceCsrMatrix& matrix,
                 Stencil stencil = Stencil::element_and_intersection)
    : space_(range_space)
  {
    hgdt_operator op = nullptr;
    const int64_t rows = int64_t(matrix.row_ptr.size()) - 1;
    check(hgdt_operator_create_on_matrix(source_space.handle(), range_space.handle(), int(stencil),
                                         matrix.row_ptr.data(), matrix.col_idx.data(), rows,
                                         int64_t(matrix.col_idx.size()), matrix.values, &op));
    h_ = std::shared_ptr<hgdt_operator_s>(op, [](hgdt_operator p) { hgdt_operator_destroy(p); });
  }

  // append(LocalElementBilinearFormInterface, param, filter), matrix-based.hh:346-369
  MatrixOperator& append(const LocalElementIntegralBilinearForm& form, const Parameter& param = {},
                         const ApplyOn::AllElements& = {})
  {
    keep(form.coefficient);
    const double scaling = param.get("matrixoperator.scaling", 1.);
    if (form.tensor) {
      if (form.tensor->storage())
        kept_.push_back(form.tensor->storage());
      check(hgdt_operator_append_laplace_tensor(handle(), form.tensor->field(), form.tensor->value(),
                                                form.over_integrate, scaling));
      return *this;
    }
    if (form.generic) {
      // values at the points of the rule of order coefficient order + 2 k + over_integrate
      const int order = form.generic->order() + 2 * space_.order() + form.over_integrate;
      auto vals = std::make_shared<DeviceVector>(form.generic->values_at(space_.grid_view(), order));
      kept_.push_back(vals);
      if (form.laplace)
        check(hgdt_operator_append_laplace_points(handle(), vals->data(), form.generic->order(), form.over_integrate,
                                                  scaling));
      else
        check(hgdt_operator_append_product_points(handle(), vals->data(), form.generic->order(), form.over_integrate,
                                                  scaling));
      return *this;
    }
    if (form.laplace)
      check(hgdt_operator_append_laplace(handle(), form.coefficient.field(), form.coefficient.value(),
                                         form.over_integrate, scaling));
    else
      check(hgdt_operator_append_product(handle(), form.coefficient.field(), form.coefficient.value(),
                                         form.over_integrate, scaling));
    return *this;
  }
  // append(LocalCouplingIntersectionBilinearFormInterface, param, filter), matrix-based.hh:371-393
  MatrixOperator& append(const LocalCouplingIntersectionIntegralBilinearForm& form, const Parameter& param = {},
                         const ApplyOn::InnerIntersectionsOnce& = {})
  {
    const double scaling = param.get("matrixoperator.scaling", 1.);
    for (const auto& t : form.integrand.terms) {
      keep(t.diffusion), keep(t.weight);
      if (t.kind == IntersectionTerm::inner_coupling)
        check(hgdt_operator_append_inner_coupling(handle(), t.parameter, t.diffusion.field(), t.diffusion.value(),
                                                  t.weight.field(), t.weight.value(), form.over_integrate, scaling));
      else if (t.kind == IntersectionTerm::inner_penalty)
        check(hgdt_operator_append_inner_penalty(handle(), t.parameter, t.weight.field(), t.weight.value(),
                                                 form.over_integrate, scaling));
      else
        throw Exceptions::integrand_error(HGDT_ERR_INTEGRAND, "boundary integrand on inner intersections");
    }
    return *this;
  }
  // append(LocalIntersectionBilinearFormInterface, param, filter), matrix-based.hh:395-408
  MatrixOperator& append(const LocalIntersectionIntegralBilinearForm& form, const Parameter& param,
                         const ApplyOn::BoundaryIntersections&)
  {
    return append(form, param, ApplyOn::CustomBoundaryIntersections{});
  }
  MatrixOperator& append(const LocalIntersectionIntegralBilinearForm& form, const Parameter& param,
                         const ApplyOn::CustomBoundaryIntersections& filter)
  {
    const BoundaryFilter scope(*this, filter);
    const double scaling = param.get("matrixoperator.scaling", 1.);
    for (const auto& t : form.integrand.terms) {
      keep(t.diffusion), keep(t.weight);
      if (t.kind == IntersectionTerm::dirichlet_coupling)
        check(hgdt_operator_append_dirichlet_coupling(handle(), t.parameter, t.diffusion.field(), t.diffusion.value(),
                                                      form.over_integrate, scaling));
      else if (t.kind == IntersectionTerm::boundary_penalty)
        check(hgdt_operator_append_boundary_penalty(handle(), t.parameter, t.weight.field(), t.weight.value(),
                                                    form.over_integrate, scaling));
      else
        throw Exceptions::integrand_error(HGDT_ERR_INTEGRAND, "inner integrand on boundary intersections");
    }
    return *this;
  }
  // a functional assembled in the same walk (walker.append(rhs_func), examples/stationary-heat-equation.cc:100-103)
  MatrixOperator& append(const VectorFunctional& functional)
  {
    for (const auto& lf : functional.local_functionals()) {
      const auto& in = lf.integrand;
      if (in.kind == LocalProductWithAnsatz::generic_function) {
        const int order = space_.order() + in.generic_f->order() + lf.over_integrate;
        auto vals = std::make_shared<DeviceVector>(in.generic_f->values_at(space_.grid_view(), order));
        kept_.push_back(vals);
        check(hgdt_operator_append_source(handle(), HGDT_SOURCE_PER_QUADRATURE_POINT, vals->data(), 0.,
                                          in.generic_f->order(), lf.over_integrate));
      } else {
        keep(in.grid_f);
        check(hgdt_operator_append_source(handle(),
                                          in.grid_f.field() ? HGDT_SOURCE_PER_ELEMENT : HGDT_SOURCE_CONSTANT,
                                          in.grid_f.field(), in.grid_f.value(), 0, lf.over_integrate));
      }
    }
    for (size_t b = 0; b < functional.boundary_functionals().size(); ++b) {
      const auto& bf = functional.boundary_functionals()[b];
      const BoundaryFilter scope(*this, functional.boundary_filters()[b]);
      const auto& t = bf.term;
      keep(t.coeff);
      const int k = space_.order();
      const int g_order = t.data_function ? t.data_function->order() : 0;
      const int order = (t.kind == HGDT_BOUNDARY_DIRICHLET_COUPLING) ? g_order + std::max(k - 1, 0) + bf.over_integrate
                                                                      : k + g_order + bf.over_integrate;
      const double* data = nullptr;
      if (t.data_function) {
        int32_t nq = 0;
        check(hgdt_grid_boundary_quadrature_points(space_.grid_view().handle(), order, nullptr, &nq));
        int64_t info[9];
        check(hgdt_grid_info(space_.grid_view().handle(), info));
        std::vector<double> pts(size_t(info[6]) * size_t(nq) * 2);
        check(hgdt_grid_boundary_quadrature_points(space_.grid_view().handle(), order, pts.data(), &nq));
        std::vector<double> vals(size_t(info[6]) * size_t(nq));
        for (size_t p = 0; p < vals.size(); ++p)
          vals[p] = t.data_function->evaluate(&pts[2 * p]);
        auto dv = std::make_shared<DeviceVector>(vals);
        kept_.push_back(dv);
        data = dv->data();
      }
      check(hgdt_operator_append_boundary_source(handle(), t.kind, t.param, t.coeff.field(), t.coeff.value(), data,
                                                 t.data_value, g_order, bf.over_integrate));
    }
    return *this;
  }
  // matrix-valued diffusion / weight (per-element TensorGridFunctions, symmetric) of the intersection form
  // appended last (inner: boundary false), in place of its scalar coefficients (laplace-ipdg.hh:47-61)
  MatrixOperator& set_face_tensors(bool boundary, const TensorGridFunction* diffusion,
                                   const TensorGridFunction* weight = nullptr)
  {
    for (const TensorGridFunction* t : {diffusion, weight})
      if (t && t->storage())
        kept_.push_back(t->storage());
    check(hgdt_operator_set_face_tensors(handle(), boundary ? 1 : 0, diffusion ? diffusion->field() : nullptr,
                                         weight ? weight->field() : nullptr));
    return *this;
  }
  MatrixOperator& append(const DirichletConstraints& constraints)
  {
    const auto mask = constraints.faces.mask(space_.grid_view());
    check(hgdt_operator_set_dirichlet_faces(handle(), mask.empty() ? nullptr : mask.data(), int64_t(mask.size())));
    check(hgdt_operator_set_dirichlet(handle(), 1));
    check(hgdt_operator_set_dirichlet_mode(handle(), constraints.only_clear ? 1 : 0, constraints.ensure_symmetry ? 1 : 0));
    return *this;
  }
  template <class T>
  MatrixOperator& operator+=(const T& what)
  {
    return append(what);
  }

  // storage used by apply / apply_inverse; Format::matrix_free stores no matrix and recomputes the operator
  // from the coefficients at every application, as LocalizableOperator::apply does
  // (operators/localizable-operator.hh:352-379); assemble() then builds the right-hand side only
  MatrixOperator& set_format(Format format)
  {
    check(hgdt_operator_set_format(handle(), int(format)));
    return *this;
  }
  // assemble(use_tbb), matrix-based.hh:496-500; asynchronous on `stream` (NULL: default stream)
  void assemble(bool /*use_tbb*/ = false, void* stream = nullptr)
  {
    check(hgdt_operator_assemble(handle(), stream));
  }
  CsrMatrix matrix() const
  {
    return CsrMatrix(handle());
  }
  // the assembled right-hand side (device memory owned by the operator)
  const double* rhs_device() const
  {
    double* b = nullptr;
    check(hgdt_operator_rhs_device(handle(), &b));
    return b;
  }
  std::vector<double> rhs() const
  {
    int64_t rows = 0, cols = 0, nnz = 0;
    check(hgdt_operator_shape(handle(), &rows, &cols, &nnz));
    std::vector<double> out(static_cast<size_t>(rows));
    check(hgdt_operator_copy_rhs(handle(), out.data()));
    return out;
  }
  // apply(source, range), matrix-based.hh:121-129
  void apply(const DeviceVector& source, DeviceVector& range, void* stream = nullptr) const
  {
    check(hgdt_operator_apply(handle(), source.data(), range.data(), stream));
  }
  // apply_inverse(range, source, opts), matrix-based.hh:148-159: Jacobi-preconditioned CG from 0
  SolverStats apply_inverse(const double* range_device, DeviceVector& source, const SolverOptions& opts = {},
                            void* stream = nullptr)
  {
    hgdt_solver_options o;
    check(hgdt_solver_options_default(&o));
    o.rtol = opts.precision, o.max_iter = opts.max_iter, o.check_every = opts.check_every;
    o.use_graph = opts.use_graph ? 1 : 0;
    o.restarts = opts.restarts;
    hgdt_solver_stats st{};
    const hgdt_status s = hgdt_operator_apply_inverse(handle(), range_device, source.data(), &o, &st, stream);
    last_ = SolverStats{st.converged != 0, st.iterations, st.relres, st.solve_ms, st.true_relres};
    check(s);
    return last_;
  }
  SolverStats apply_inverse(const DeviceVector& range, DeviceVector& source, const SolverOptions& opts = {},
                            void* stream = nullptr)
  {
    return apply_inverse(range.data(), source, opts, stream);
  }
  const SolverStats& last_solve() const
  {
    return last_;
  }
  const SpaceInterface& source_space() const
  {
    return space_;
  }
  hgdt_operator handle() const
  {
    return h_.get();
  }

private:
  // the ApplyOn filter of the boundary forms and functionals appended while it lives (restored to every
  // boundary intersection afterwards)
  struct BoundaryFilter
  {
    BoundaryFilter(const MatrixOperator& op, const ApplyOn::CustomBoundaryIntersections& filter)
      : h(op.handle())
    {
      const auto mask = filter.mask(op.space_.grid_view());
      check(hgdt_operator_set_boundary_filter(h, mask.empty() ? nullptr : mask.data(), int64_t(mask.size())));
    }
    ~BoundaryFilter()
    {
      hgdt_operator_set_boundary_filter(h, nullptr, 0);
    }
    BoundaryFilter(const BoundaryFilter&) = delete;
    BoundaryFilter& operator=(const BoundaryFilter&) = delete;
    hgdt_operator h;
  };
  void keep(const GridFunction& f)
  {
    if (f.storage())
      kept_.push_back(f.storage());
  }
  SpaceInterface space_;
  std::shared_ptr<hgdt_operator_s> h_;
  std::vector<std::shared_ptr<DeviceVector>> kept_; // coefficient fields must outlive the operator's use
  SolverStats last_;
};

inline MatrixOperator make_matrix_operator(const SpaceInterface& space, Stencil stencil = Stencil::element_and_intersection)
{
  return MatrixOperator(space, stencil);
}
// make_matrix_operator(source_space, range_space, stencil), matrix-based.hh:200-216
inline MatrixOperator make_matrix_operator(const SpaceInterface& source_space, const SpaceInterface& range_space,
                                           Stencil stencil = Stencil::element_and_intersection)
{
  return MatrixOperator(source_space, range_space, stencil);
}
// make_matrix_operator(grid_view, source_space, range_space, matrix), matrix-based.hh:511-546
inline MatrixOperator make_matrix_operator(const SpaceInterface& source_space, const SpaceInterface& range_space,
                                           DeviceCsrMatrix& matrix, Stencil stencil = Stencil::element_and_intersection)
{
  return MatrixOperator(source_space, range_space, matrix, stencil);
}

// ---------------------------------------------------------------------------------------------
// discrete functions and bilinear forms on them
// ---------------------------------------------------------------------------------------------
class DiscreteFunction
{
public:
  explicit DiscreteFunction(const SpaceInterface& space, const std::string& name = "")
    : space_(space)
    , dofs_(std::make_shared<DeviceVector>(space.size(), 0.))
    , name_(name)
  {}
  const SpaceInterface& space() const
  {
    return space_;
  }
  DeviceVector& dofs()
  {
    return *dofs_;
  }
  const DeviceVector& dofs() const
  {
    return *dofs_;
  }

private:
  SpaceInterface space_;
  std::shared_ptr<DeviceVector> dofs_;
  std::string name_;
};
inline DiscreteFunction make_discrete_function(const SpaceInterface& space, const std::string& name = "")
{
  return DiscreteFunction(space, name);
}

// u_h - g
struct DiscreteFunctionDifference
{
  const DiscreteFunction* discrete;
  const GenericFunction* function;
};
inline DiscreteFunctionDifference operator-(const DiscreteFunction& u, const GenericFunction& g)
{
  return {&u, &g};
}

// make_bilinear_form(grid_view, e, e) += LocalElementIntegralBilinearForm(...); apply2() (bilinear-form.hh:344-452)
// for e = u_h - g: the squared error in the norm of the appended integrands, computed on the device
class BilinearForm
{
public:
  BilinearForm(const Grid& grid_view, DiscreteFunctionDifference e)
    : grid_(grid_view)
    , e_(e)
  {}
  BilinearForm& operator+=(const LocalElementIntegralBilinearForm& form)
  {
    forms_.push_back(form);
    return *this;
  }
  BilinearForm& append(const LocalElementIntegralBilinearForm& form)
  {
    return *this += form;
  }
  double apply2() const
  {
    const int order_e = std::max(e_.discrete->space().order(), e_.function ? e_.function->order() : 0);
    double total = 0.;
    const int ninst = grid_.num_instances();
    for (const auto& f : forms_) {
      const int order = f.coefficient.order() + 2 * order_e + f.over_integrate;
      DeviceVector gv, gg;
      if (e_.function) {
        gv = DeviceVector(e_.function->values_at(grid_, order));
        if (f.laplace)
          gg = DeviceVector(e_.function->gradients_at(grid_, order));
      }
      std::vector<double> out(size_t(2 * ninst));
      if (!f.laplace && f.coefficient.field())
        throw Exceptions::not_supported(HGDT_ERR_NOT_SUPPORTED, "product norms with a constant weight");
      check(hgdt_space_error_norms(e_.discrete->space().handle(), e_.discrete->dofs().data(), order,
                                   e_.function ? gv.data() : nullptr, (e_.function && f.laplace) ? gg.data() : nullptr,
                                   f.laplace ? f.coefficient.field() : nullptr,
                                   f.laplace ? f.coefficient.value() : 1., out.data(), nullptr));
      for (int i = 0; i < ninst; ++i)
        total += f.laplace ? out[size_t(2 * i + 1)] : f.coefficient.value() * out[size_t(2 * i)];
    }
    return total;
  }
  double result() const
  {
    return apply2();
  }

private:
  Grid grid_;
  DiscreteFunctionDifference e_;
  std::vector<LocalElementIntegralBilinearForm> forms_;
};
inline BilinearForm make_bilinear_form(const Grid& grid_view, DiscreteFunctionDifference e)
{
  return BilinearForm(grid_view, e);
}

} // namespace hgdt

#endif // HGDT_HH
