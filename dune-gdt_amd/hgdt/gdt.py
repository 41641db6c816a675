"""Python face of libhgdt, mirroring the dune.gdt names on the assembly + solve path.

Reference bindings mirrored (paths relative to the dune-gdt source tree):
  MatrixOperator(grid, source_space, range_space, ...)   python/dune/gdt/operators/matrix-based.hh:197-351
  op.append / op += ..., op.assemble(parallel=False)      :261-351
  op.apply / op.apply_inverse                             python/dune/gdt/operators/interfaces.hh:55-141
  ContinuousLagrangeSpace / DiscontinuousLagrangeSpace     python/dune/gdt/spaces/h1/continuous-lagrange.hh:57-61
  LocalLaplaceIntegrand(diffusion)                         python/dune/gdt/local/integrands/laplace.cc:60-83
  LocalElementIntegralBilinearForm(integrand, over_integrate=0)
                                                          python/dune/gdt/local/bilinear-forms/element_integrals.cc:83-98
  LocalIPDGInnerPenalty(penalty, weight)                   python/dune/gdt/local/integrands/ipdg_inner_penalty.cc:62-86
  LocalLaplaceIPDGInnerCoupling(symmetry_prefactor, diffusion, weight=1)
                                                          python/dune/gdt/local/integrands/laplace_ipdg_inner_coupling.cc:62-86
  DirichletConstraints                                     python/dune/gdt/tools/dirichlet-constraints.cc:54-140

Vectors are torch float64 tensors on the current HIP device; numpy arrays are copied there.
Every computation runs in libhgdt.so on the GPU; there is no host fallback.
"""
import ctypes
import enum

import numpy as np
import torch

from . import _lib as L

__all__ = [
    "Grid", "make_cube_grid", "ContinuousLagrangeSpace", "DiscontinuousLagrangeSpace", "GridFunction",
    "QuadraturePointFunction", "LocalLaplaceIntegrand", "LocalElementProductIntegrand", "LocalProductIntegrand",
    "LocalLaplaceIPDGInnerCoupling", "LocalIPDGInnerPenalty", "LocalLaplaceIPDGDirichletCoupling",
    "LocalIPDGBoundaryPenalty", "LocalElementIntegralBilinearForm", "LocalCouplingIntersectionIntegralBilinearForm",
    "LocalIntersectionIntegralBilinearForm", "LocalElementIntegralFunctional", "VectorFunctional",
    "make_vector_functional", "DirichletConstraints", "make_dirichlet_constraints", "AllDirichletBoundaryInfo",
    "AllNeumannBoundaryInfo", "NormalBasedBoundaryInfo", "BoundaryInfo", "DirichletBoundary", "NeumannBoundary",
    "NoBoundary",
    "ApplyOn", "Stencil", "MatrixOperator", "make_matrix_operator", "CsrMatrix", "GenericFunction",
    "DiscreteFunction", "make_discrete_function", "BilinearForm", "make_bilinear_form",
    "LocalIntersectionIntegralFunctional", "LocalIntersectionProductIntegrand",
]

GEOMETRIES = {"quad": L.QUAD, "cube": L.QUAD, "tri": L.TRI, "simplex": L.TRI, "hex": L.HEX, "tet": L.TET}


def _device():
    if not torch.cuda.is_available():
        raise L.device_error(8, "no HIP device visible to torch: libhgdt has no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


def _stream_ptr(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def _dvec(x, n=None, name="vector"):
    """float64 contiguous device tensor (copies numpy / host tensors)."""
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64))
    if not isinstance(x, torch.Tensor):
        raise L.wrong_input_given(1, f"{name} must be a torch tensor or numpy array")
    if x.dtype != torch.float64:
        raise L.wrong_input_given(1, f"{name} must be float64 (FP64 path)")
    if x.device.type != "cuda":
        x = x.to(_device())
    x = x.contiguous()
    if n is not None and x.numel() != n:
        raise L.shapes_do_not_match(2, f"{name} has {x.numel()} entries, expected {n}")
    return x


class Stencil(enum.IntEnum):
    """dune/gdt/type_traits.hh:55"""
    element = L.STENCIL_ELEMENT
    intersection = L.STENCIL_INTERSECTION
    element_and_intersection = L.STENCIL_ELEMENT_AND_INTERSECTION


class NoBoundary:
    """XT::Grid::NoBoundary"""
    id = 0


class DirichletBoundary:
    """XT::Grid::DirichletBoundary"""
    id = 1


class NeumannBoundary:
    """XT::Grid::NeumannBoundary"""
    id = 2


def _type_id(t):
    return (t if isinstance(t, type) else type(t)).id


class BoundaryInfo:
    """XT::Grid::BoundaryInfo: the boundary type of every boundary intersection.  type_ids(grid) gives one id per
    boundary intersection in walk order (elements ascending, their faces in reference order)."""

    def type_ids(self, grid):
        raise NotImplementedError

    def mask(self, grid, boundary_type):
        """one byte per boundary intersection: 1 where type(intersection) == boundary_type"""
        return (self.type_ids(grid) == _type_id(boundary_type)).astype(np.uint8)


class AllDirichletBoundaryInfo(BoundaryInfo):
    """XT::Grid::AllDirichletBoundaryInfo: every boundary intersection is a Dirichlet one."""

    def type_ids(self, grid):
        return np.full(grid.num_boundary_intersections, DirichletBoundary.id, np.int8)


class AllNeumannBoundaryInfo(BoundaryInfo):
    """XT::Grid::AllNeumannBoundaryInfo: every boundary intersection is a Neumann one."""

    def type_ids(self, grid):
        return np.full(grid.num_boundary_intersections, NeumannBoundary.id, np.int8)


class NormalBasedBoundaryInfo(BoundaryInfo):
    """XT::Grid::NormalBasedBoundaryInfo: the type of an intersection is the one registered for its outward unit
    normal (within tol), else the default (as used in dune/gdt/test/interpolations/boundary.hh:79-81)."""

    def __init__(self, tol=1e-10, default_boundary_type=None):
        self.tol = tol
        self.default = NoBoundary if default_boundary_type is None else default_boundary_type
        self.normals = []

    def register_new_normal(self, normal, boundary_type):
        n = np.asarray(normal, np.float64)
        self.normals.append((n / np.linalg.norm(n), boundary_type))
        return self

    def type_ids(self, grid):
        _, normals = grid.boundary_intersections()
        ids = np.full(len(normals), _type_id(self.default), np.int8)
        for n, t in self.normals:
            d = n[: normals.shape[1]]
            ids[np.linalg.norm(normals - d, axis=1) < self.tol] = _type_id(t)
        return ids


class ApplyOn:
    """Element / intersection filters of XT::Grid::ApplyOn used on this path."""

    class AllElements:
        pass

    class InnerIntersectionsOnce:
        pass

    class BoundaryIntersections:
        pass

    class CustomBoundaryIntersections:
        """the boundary intersections with boundary_info.type(intersection) == boundary_type"""

        def __init__(self, boundary_info=None, boundary_type=None):
            self.boundary_info = boundary_info
            self.boundary_type = DirichletBoundary if boundary_type is None else boundary_type

        def mask(self, grid):
            """None when every boundary intersection qualifies, else one byte per boundary intersection"""
            if self.boundary_info is None:
                return None
            m = self.boundary_info.mask(grid, self.boundary_type)
            return None if m.all() else m


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------
class Grid:
    """Grid (leaf view) held by libhgdt: vertex coordinates + connectivity on the device."""

    def __init__(self, handle, parts=()):
        self._h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle
        self._parts = tuple(parts)
        info = (ctypes.c_int64 * 9)()
        L.call("hgdt_grid_info", self._h, info)
        (self.dimension, self.num_vertices, self.num_elements, self.vertices_per_element, self.faces_per_element,
         self.num_inner_intersections, self.num_boundary_intersections, self.num_instances,
         self.geometry) = [int(v) for v in info]

    @classmethod
    def structured(cls, geometry, lower_left, upper_right, num_elements):
        geometry = GEOMETRIES.get(geometry, geometry)
        dim = 2 if geometry in (L.QUAD, L.TRI) else 3
        ll = np.broadcast_to(np.asarray(lower_left, np.float64), (dim,)).copy()
        ur = np.broadcast_to(np.asarray(upper_right, np.float64), (dim,)).copy()
        ne = np.broadcast_to(np.asarray(num_elements, np.int32), (dim,)).copy()
        h = ctypes.c_void_p()
        L.call("hgdt_grid_create_structured", geometry, ne.ctypes.data_as(L._pi32), ll.ctypes.data_as(L._pdbl),
               ur.ctypes.data_as(L._pdbl), ctypes.byref(h))
        return cls(h)

    @classmethod
    def from_arrays(cls, geometry, coords, connectivity):
        geometry = GEOMETRIES.get(geometry, geometry)
        c = np.ascontiguousarray(coords, np.float64)
        k = np.ascontiguousarray(connectivity, np.int32)
        h = ctypes.c_void_p()
        L.call("hgdt_grid_create", geometry, c.shape[0], c.ctypes.data_as(L._pdbl), k.shape[0],
               k.ctypes.data_as(L._pi32), ctypes.byref(h))
        return cls(h)

    @classmethod
    def mixed(cls, coords, elements):
        """Triangles and quadrilaterals in one grid, element by element as GridFactory<G>::insertElement(
        GeometryTypes::simplex(2) / cube(2), vertices) builds the mixed grids of
        dune/gdt/test/spaces/base.hh:355-376: `elements` lists every element's vertices in the reference
        vertex order (3: a triangle, 4: a quadrilateral)."""
        c = np.ascontiguousarray(coords, np.float64)
        if c.ndim != 2 or c.shape[1] != 2:
            raise L.shapes_do_not_match(2, "mixed grids are two-dimensional: coords must be [num_vertices, 2]")
        sizes = np.array([len(v) for v in elements], np.int64)
        if np.any((sizes != 3) & (sizes != 4)):
            raise L.wrong_input_given(1, "an element of a mixed grid has 3 (triangle) or 4 (quadrilateral) vertices")
        geometry = np.where(sizes == 4, L.QUAD, L.TRI).astype(np.int32)
        offsets = np.zeros(len(elements) + 1, np.int64)
        offsets[1:] = np.cumsum(sizes)
        verts = np.ascontiguousarray(np.concatenate([np.asarray(v, np.int32) for v in elements]), np.int32)
        h = ctypes.c_void_p()
        L.call("hgdt_grid_create_mixed", c.shape[0], c.ctypes.data_as(L._pdbl), len(elements),
               geometry.ctypes.data_as(L._pi32), offsets.ctypes.data_as(L._pi64), verts.ctypes.data_as(L._pi32),
               ctypes.byref(h))
        return cls(h)

    @property
    def is_mixed(self):
        return self.geometry == L.MIXED

    @classmethod
    def union(cls, grids):
        """Disjoint union of independent grids: the batch of problem instances."""
        arr = (ctypes.c_void_p * len(grids))(*[g._h.value for g in grids])
        h = ctypes.c_void_p()
        L.call("hgdt_grid_create_union", arr, len(grids), ctypes.byref(h))
        return cls(h, parts=grids)

    def coordinates(self):
        out = np.empty((self.num_vertices, self.dimension))
        L.call("hgdt_grid_copy_coords", self._h, out.ctypes.data_as(L._pdbl))
        return out

    def connectivity(self):
        out = np.empty((self.num_elements, self.vertices_per_element), np.int32)
        L.call("hgdt_grid_copy_connectivity", self._h, out.ctypes.data_as(L._pi32))
        return out

    def instance_elements(self):
        out = np.empty(self.num_instances + 1, np.int64)
        L.call("hgdt_grid_instance_elements", self._h, out.ctypes.data_as(L._pi64))
        return out

    def boundary_intersections(self):
        """(centres, outward unit normals) of the boundary intersections in walk order, each [nb, dim]"""
        nb, d = self.num_boundary_intersections, self.dimension
        c, n = np.empty((nb, d)), np.empty((nb, d))
        L.call("hgdt_grid_boundary_intersections", self._h, c.ctypes.data_as(L._pdbl), n.ctypes.data_as(L._pdbl))
        return c, n

    def boundary_quadrature_points(self, order):
        """Physical coordinates of the face-rule points on the boundary intersections in walk order: (nb, nq, dim)."""
        nq = ctypes.c_int32()
        L.call("hgdt_grid_boundary_quadrature_points", self._h, order, None, ctypes.byref(nq))
        out = np.empty((self.num_boundary_intersections, nq.value, self.dimension))
        L.call("hgdt_grid_boundary_quadrature_points", self._h, order, out.ctypes.data_as(L._pdbl),
               ctypes.byref(nq))
        return out

    def quadrature_points(self, order):
        """Physical coordinates of the element quadrature points of the given order: (E, nq, dim); on a mixed
        grid, whose element types have rules of different sizes, (P, dim) in element order (element e holds
        the points quadrature_offsets(order)[e] .. [e + 1])."""
        nq = ctypes.c_int32()
        L.call("hgdt_grid_quadrature_points", self._h, order, None, ctypes.byref(nq))
        if self.is_mixed:
            out = np.empty((int(self.quadrature_offsets(order)[-1]), self.dimension))
        else:
            out = np.empty((self.num_elements, nq.value, self.dimension))
        L.call("hgdt_grid_quadrature_points", self._h, order, out.ctypes.data_as(L._pdbl), ctypes.byref(nq))
        return out

    def quadrature_offsets(self, order):
        """Start of every element's quadrature points (rule of the given order) in the point layout: (E + 1,)."""
        out = np.empty(self.num_elements + 1, np.int64)
        L.call("hgdt_grid_quadrature_offsets", self._h, order, out.ctypes.data_as(L._pi64))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and L._lib is not None:
            L._lib.hgdt_grid_destroy(h)
            self._h = None


def make_cube_grid(lower_left, upper_right, num_elements, simplex=False, dim=2):
    """XT::Grid::make_cube_grid: cube cells, or dim! Kuhn simplices per cube when simplex=True."""
    n = np.atleast_1d(num_elements)
    if n.size > 1:
        dim = n.size
    geometry = (L.TRI if simplex else L.QUAD) if dim == 2 else (L.TET if simplex else L.HEX)
    return Grid.structured(geometry, lower_left, upper_right, num_elements)


# ---------------------------------------------------------------------------
# spaces
# ---------------------------------------------------------------------------
class _LagrangeSpace:
    _kind = None

    def __init__(self, grid, order=1):
        self.grid = grid  # keep-alive, as py::keep_alive in the reference bindings
        h = ctypes.c_void_p()
        L.call("hgdt_space_create", grid._h, self._kind, order, ctypes.byref(h))
        self._h = h
        self.order = order
        n = ctypes.c_int64()
        ls = ctypes.c_int32()
        L.call("hgdt_space_size", h, ctypes.byref(n), ctypes.byref(ls))
        self.num_DoFs = int(n.value)
        self.max_local_size = int(ls.value)

    @property
    def size(self):
        return self.num_DoFs

    def global_indices(self):
        out = np.empty((self.grid.num_elements, self.max_local_size), np.int32)
        L.call("hgdt_space_copy_global_indices", self._h, out.ctypes.data_as(L._pi32))
        return out

    def instance_dofs(self):
        out = np.empty(self.grid.num_instances + 1, np.int64)
        L.call("hgdt_space_instance_dofs", self._h, out.ctypes.data_as(L._pi64))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and L._lib is not None:
            L._lib.hgdt_space_destroy(h)
            self._h = None


class ContinuousLagrangeSpace(_LagrangeSpace):
    """dune/gdt/spaces/h1/continuous-lagrange.hh:49-183 (orders 1 and 2 on the device)."""
    _kind = L.CONTINUOUS_LAGRANGE


class DiscontinuousLagrangeSpace(_LagrangeSpace):
    """dune/gdt/spaces/l2/discontinuous-lagrange.hh:60-183 (orders 1 and 2 on the device)."""
    _kind = L.DISCONTINUOUS_LAGRANGE


# ---------------------------------------------------------------------------
# data functions
# ---------------------------------------------------------------------------
class GridFunction:
    """Element-wise constant scalar function (times the identity where a matrix is expected), or a constant."""

    def __init__(self, value):
        if isinstance(value, GridFunction):
            self.constant, self.field = value.constant, value.field
            return
        if isinstance(value, (int, float)):
            self.constant, self.field = float(value), None
        else:
            if isinstance(value, TensorGridFunction) or getattr(value, "ndim", 1) > 1:
                raise L.not_supported(9, "matrix-valued coefficients are taken by LocalLaplaceIntegrand only")
            self.constant, self.field = 0.0, _dvec(value, name="element field")

    def args(self, grid=None):
        if self.field is None:
            return None, self.constant
        if grid is not None and self.field.numel() != grid.num_elements:
            raise L.shapes_do_not_match(2, f"field has {self.field.numel()} values, grid has "
                                           f"{grid.num_elements} elements")
        return ctypes.c_void_p(self.field.data_ptr()), 0.0

    def order(self):
        return 0


class TensorGridFunction:
    """Matrix-valued GridFunction<E, d, d> for LocalLaplaceIntegrand (laplace.hh:40-48): one symmetric d x d
    matrix for every element (shape (d, d)) or one per element (shape (num_elements, d, d), device tensor)."""

    def __init__(self, value):
        if isinstance(value, TensorGridFunction):
            self.constant, self.field = value.constant, value.field
            return
        if isinstance(value, torch.Tensor) and value.dim() == 3:
            self.constant, self.field = None, value.to(dtype=torch.float64, device=_device()).contiguous()
        else:
            a = np.asarray(value.cpu() if isinstance(value, torch.Tensor) else value, np.float64)
            if a.ndim == 3:
                self.constant, self.field = None, torch.from_numpy(np.ascontiguousarray(a)).to(_device())
            elif a.ndim == 2 and a.shape[0] == a.shape[1]:
                self.constant, self.field = np.ascontiguousarray(a), None
            else:
                raise L.shapes_do_not_match(2, f"a diffusion tensor has shape (d, d) or (elements, d, d), not {a.shape}")


class QuadraturePointFunction:
    """Values at the element quadrature points of `order` (layout of Grid.quadrature_points)."""

    def __init__(self, grid, values_or_callable, order):
        self.grid, self.order_total = grid, int(order)
        if callable(values_or_callable):
            pts = grid.quadrature_points(order)
            vals = np.asarray(values_or_callable(pts.reshape(-1, grid.dimension).T), np.float64)
            self.values = _dvec(vals.reshape(-1))
        else:
            self.values = _dvec(values_or_callable)


class GenericFunction:
    """XT::Functions::GenericFunction(order, evaluate, name, parameter_type, jacobian) as used in
    examples/stationary-heat-equation.cc:68-85: an analytic function of the physical coordinates, evaluated
    on the host at the element quadrature points.  `evaluate(X)` and `jacobian(X)` receive X of shape
    (dim, N) and return (N,) values and (dim, N) gradients."""

    def __init__(self, order, evaluate, name="", parameter_type=None, jacobian=None):
        self.order_value, self.evaluate, self.name, self.jacobian = int(order), evaluate, name, jacobian

    def order(self):
        return self.order_value

    def values_at(self, grid, order):
        pts = grid.quadrature_points(order)
        return np.asarray(self.evaluate(pts.reshape(-1, grid.dimension).T), np.float64).reshape(-1)

    def gradients_at(self, grid, order):
        if self.jacobian is None:
            raise L.not_supported(9, f"GenericFunction {self.name!r} has no jacobian")
        pts = grid.quadrature_points(order)
        g = np.asarray(self.jacobian(pts.reshape(-1, grid.dimension).T), np.float64).reshape(grid.dimension, -1)
        return np.ascontiguousarray(g.T).reshape(-1)


def _gf(v):
    if isinstance(v, (GenericFunction, QuadraturePointFunction)):
        raise L.not_supported(9, "this integrand takes constant or per-element coefficients (element integrands "
                                 "also take coefficients varying inside the elements)")
    return v if isinstance(v, GridFunction) else GridFunction(v)


def _face_coefficient(v):
    """coefficient of an intersection integrand: a GridFunction, or a symmetric matrix-valued TensorGridFunction
    (GridFunction<E, d, d>, laplace-ipdg.hh:47-61, ipdg.hh:59-72)"""
    if isinstance(v, TensorGridFunction):
        return v
    if getattr(v, "ndim", 1) > 1 or (isinstance(v, (list, tuple)) and np.ndim(v) > 1):
        return TensorGridFunction(v)
    return _gf(v)


def _element_coefficient(v):
    """coefficient of an element integrand: a GridFunction (constant or per element), or one that varies inside
    the elements (GenericFunction, QuadraturePointFunction), evaluated at the quadrature points"""
    if isinstance(v, (GenericFunction, QuadraturePointFunction, TensorGridFunction)):
        return v
    if getattr(v, "ndim", 1) > 1 or (isinstance(v, (list, tuple)) and np.ndim(v) > 1):
        return TensorGridFunction(v)
    return _gf(v)


# ---------------------------------------------------------------------------
# integrands (dune/gdt/local/integrands/*.hh)
# ---------------------------------------------------------------------------
class _Integrand:
    kinds = ()

    def __add__(self, other):
        return _IntegrandSum(self, other)


class _IntegrandSum(_Integrand):
    """LocalQuaternaryIntersectionIntegrandSum / binary sums (local/integrands/combined.hh:168-440)."""

    def __init__(self, a, b):
        self.parts = (a.parts if isinstance(a, _IntegrandSum) else (a,)) + (
            b.parts if isinstance(b, _IntegrandSum) else (b,))


class LocalLaplaceIntegrand(_Integrand):
    """(diffusion grad phi_j) . grad psi_i, local/integrands/laplace.hh:27-109"""

    def __init__(self, diffusion=1.0, logging_prefix=""):
        self.diffusion = _element_coefficient(diffusion)


class LocalElementProductIntegrand(_Integrand):
    """(weight psi_i) phi_j, local/integrands/product.hh:43-137"""

    def __init__(self, weight=1.0, logging_prefix=""):
        self.weight = _element_coefficient(weight)

    def with_ansatz(self, function):
        """LocalBinaryToUnaryElementIntegrand (local/integrands/conversion.hh:42-124)"""
        return _ProductWithAnsatz(self, function)


LocalProductIntegrand = LocalElementProductIntegrand


class _ProductWithAnsatz(_Integrand):
    def __init__(self, product, function):
        if (not isinstance(product.weight, GridFunction) or product.weight.field is not None
                or product.weight.constant != 1.0):
            raise L.not_supported(9, "right-hand side functionals use the unit product weight")
        self.function = function


class LocalLaplaceIPDGInnerCoupling(_Integrand):
    """LocalLaplaceIPDGIntegrands::InnerCoupling, local/integrands/laplace-ipdg.hh:32-199"""

    def __init__(self, symmetry_prefactor, diffusion, weight=1.0, logging_prefix=""):
        self.symmetry_prefactor = float(symmetry_prefactor)
        self.diffusion, self.weight = _face_coefficient(diffusion), _face_coefficient(weight)


class LocalIPDGInnerPenalty(_Integrand):
    """LocalIPDGIntegrands::InnerPenalty, local/integrands/ipdg.hh:44-183 (face diameter as h)"""

    def __init__(self, penalty, weight=1.0, logging_prefix=""):
        self.penalty, self.weight = float(penalty), _face_coefficient(weight)


class LocalLaplaceIPDGDirichletCoupling(_Integrand):
    """LocalLaplaceIPDGIntegrands::DirichletCoupling, local/integrands/laplace-ipdg.hh:213-382: a binary
    integrand in a LocalIntersectionIntegralBilinearForm; in a LocalIntersectionIntegralFunctional the
    unary one with the Dirichlet data (a constant or a GenericFunction)"""

    def __init__(self, symmetry_prefactor, diffusion, dirichlet_data=0.0, logging_prefix=""):
        self.symmetry_prefactor, self.diffusion = float(symmetry_prefactor), _face_coefficient(diffusion)
        self.dirichlet_data = dirichlet_data


class LocalIPDGBoundaryPenalty(_Integrand):
    """LocalIPDGIntegrands::BoundaryPenalty, local/integrands/ipdg.hh:186-291"""

    def __init__(self, penalty, weight=1.0, logging_prefix=""):
        self.penalty, self.weight = float(penalty), _face_coefficient(weight)

    def with_ansatz(self, function):
        """LocalBinaryToUnaryIntersectionIntegrand (local/integrands/conversion.hh:133-221)"""
        return _BoundaryWithAnsatz(L.BOUNDARY_PENALTY, self.penalty, self.weight, function)


class LocalIntersectionProductIntegrand(_Integrand):
    """(weight psi_i) phi_j on intersections, local/integrands/product.hh:280-380 (inside element)"""

    def __init__(self, weight=1.0, logging_prefix=""):
        self.weight = _gf(weight)

    def with_ansatz(self, function):
        """LocalBinaryToUnaryIntersectionIntegrand (local/integrands/conversion.hh:133-221)"""
        return _BoundaryWithAnsatz(L.BOUNDARY_PRODUCT, 0.0, self.weight, function)


class _BoundaryWithAnsatz(_Integrand):
    def __init__(self, kind, param, coeff, function):
        self.kind, self.param, self.coeff, self.function = kind, float(param), coeff, function


# ---------------------------------------------------------------------------
# local forms (dune/gdt/local/bilinear-forms/integrals.hh, local/functionals/integrals.hh)
# ---------------------------------------------------------------------------
class LocalElementIntegralBilinearForm:
    def __init__(self, integrand, over_integrate=0, logging_prefix=""):
        self.integrand, self.over_integrate = integrand, int(over_integrate)


class LocalCouplingIntersectionIntegralBilinearForm:
    def __init__(self, integrand, over_integrate=0, logging_prefix=""):
        self.integrand, self.over_integrate = integrand, int(over_integrate)


class LocalIntersectionIntegralBilinearForm:
    def __init__(self, integrand, over_integrate=0, logging_prefix=""):
        self.integrand, self.over_integrate = integrand, int(over_integrate)


class LocalElementIntegralFunctional:
    def __init__(self, integrand, over_integrate=0, logging_prefix=""):
        if not isinstance(integrand, _ProductWithAnsatz):
            raise L.not_supported(9, "element functionals take LocalProductIntegrand().with_ansatz(f)")
        self.integrand, self.over_integrate = integrand, int(over_integrate)


class LocalIntersectionIntegralFunctional:
    """local/functionals/integrals.hh:108-191 on boundary intersections: the integrand is
    LocalLaplaceIPDGDirichletCoupling(s, diffusion, dirichlet_data), LocalIPDGBoundaryPenalty(...).with_ansatz(g)
    or LocalIntersectionProductIntegrand(weight).with_ansatz(g)"""

    def __init__(self, integrand, over_integrate=0, logging_prefix=""):
        if not isinstance(integrand, (LocalLaplaceIPDGDirichletCoupling, _BoundaryWithAnsatz)):
            raise L.not_supported(9, "intersection functionals take the Dirichlet coupling or a with_ansatz integrand")
        self.integrand, self.over_integrate = integrand, int(over_integrate)


class VectorFunctional:
    """VectorBasedFunctional (functionals/vector-based.hh:132-286), assembled in the operator's walk."""

    def __init__(self, space):
        self.space = space
        self.local_functionals = []
        self._op = None

    def append(self, local_functional, param=None, filter=None):
        if isinstance(local_functional, LocalIntersectionIntegralFunctional) and filter is not None and not isinstance(
                filter, (ApplyOn.CustomBoundaryIntersections, ApplyOn.BoundaryIntersections)):
            raise L.not_supported(9, "intersection functionals are applied on boundary intersections")
        local_functional._filter = filter
        self.local_functionals.append(local_functional)
        return self

    __iadd__ = append

    @property
    def vector(self):
        if self._op is None:
            raise L.operator_error(3, "functional not assembled: append it to a MatrixOperator and assemble")
        return self._op.rhs


def make_vector_functional(space):
    return VectorFunctional(space)


class DirichletConstraints:
    """DirichletConstraints (tools/dirichlet-constraints.hh:47-211), applied inside the operator's assembly: the DoFs
    on the intersections with boundary_info.type(intersection) == DirichletBoundary() (every boundary intersection
    without a boundary info)."""

    def __init__(self, space, boundary_info=None):
        self.space = space
        self.boundary_info = boundary_info
        self._op = None

    @property
    def dirichlet_DoFs(self):
        if self._op is None:
            raise L.operator_error(3, "constraints not attached: append them to a MatrixOperator")
        return self._op.dirichlet_dofs()

    def apply(self, *args, only_clear=None, ensure_symmetry=None):
        """apply(matrix[, vector], only_clear=False, ensure_symmetry=True) (tools/dirichlet-constraints.hh:123-184):
        fused into the operator's assembly (unit rows and columns, rhs entries 0, by default).  Another variant
        (cleared rows; rows without their columns) is set for the attached operator, which is assembled again
        so that its matrix holds it.  The flags may be given positionally after the matrix and vector, as in
        the reference."""
        if self._op is None or not self._op._assembled:
            raise L.operator_error(3, "constraints are applied during MatrixOperator.assemble()")
        flags = [a for a in args if isinstance(a, bool)]
        oc = bool(flags[0]) if len(flags) > 0 else False if only_clear is None else bool(only_clear)
        es = bool(flags[1]) if len(flags) > 1 else True if ensure_symmetry is None else bool(ensure_symmetry)
        op = self._op
        if (oc, es) != op._dirichlet_mode:
            L.call("hgdt_operator_set_dirichlet_mode", op._h, int(oc), int(es))
            op._dirichlet_mode = (oc, es)
            op.assemble()


def make_dirichlet_constraints(space, boundary_info=None):
    return DirichletConstraints(space, boundary_info)


class CsrMatrix:
    """Assembled matrix: sorted CSR pattern (host) + values (device tensor)."""

    def __init__(self, row_ptr, cols, values, shape):
        self.row_ptr, self.cols, self.values, self.shape = row_ptr, cols, values, shape

    @property
    def nnz(self):
        return int(self.values.numel())

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.values.cpu().numpy(), self.cols, self.row_ptr), shape=self.shape)

    def rows(self):
        return self.shape[0]

    def cols_count(self):
        return self.shape[1]


# ---------------------------------------------------------------------------
# MatrixOperator
# ---------------------------------------------------------------------------
class MatrixOperator:
    """MatrixOperator (dune/gdt/operators/matrix-based.hh:245-508) on the GPU."""

    def __init__(self, grid, source_space, range_space=None, stencil=Stencil.element_and_intersection,
                 logging_prefix="", matrix=None):
        """MatrixOperator(grid, source_space, range_space[, matrix]) (matrix-based.hh:279-311; python bindings
        :197-240): rows are the range space's DoFs, columns the source space's.  With `matrix` (a CsrMatrix
        whose values are a float64 device tensor) the operator assembles into that matrix, adding to the
        values it holds, and `matrix` is the operator's matrix; its pattern must hold the stencil's."""
        range_space = source_space if range_space is None else range_space
        self.grid, self.space, self.source_space = grid, range_space, source_space
        self._borrowed = None
        h = ctypes.c_void_p()
        if matrix is not None:
            rp = np.ascontiguousarray(matrix.row_ptr, np.int64)
            cols = np.ascontiguousarray(matrix.cols, np.int32)
            vals = matrix.values
            if not (isinstance(vals, torch.Tensor) and vals.is_cuda and vals.dtype == torch.float64
                    and vals.is_contiguous() and vals.numel() == cols.size):
                raise L.wrong_input_given(1, "matrix values: a contiguous float64 device tensor of nnz entries")
            L.call("hgdt_operator_create_on_matrix", source_space._h, range_space._h, int(stencil),
                   rp.ctypes.data_as(L._pi64), cols.ctypes.data_as(L._pi32), len(rp) - 1, cols.size,
                   ctypes.c_void_p(vals.data_ptr()), ctypes.byref(h))
            self._borrowed = matrix
        elif range_space is source_space:
            L.call("hgdt_operator_create", source_space._h, int(stencil), ctypes.byref(h))
        else:
            L.call("hgdt_operator_create_between", source_space._h, range_space._h, int(stencil), ctypes.byref(h))
        self._h = h
        r, c, nnz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        L.call("hgdt_operator_shape", h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(nnz))
        self.shape = (int(r.value), int(c.value))
        self.nnz = int(nnz.value)
        self._keep = []  # device fields referenced by the appended forms
        self._eforms = 0
        self._dirichlet_mode = (False, True)  # DirichletConstraints::apply(..., only_clear, ensure_symmetry)
        self._assembled = False
        self.last_solve = None
        self._pattern = None

    # -- append / += (matrix-based.hh:346-408, python bindings :261-346) --
    def append(self, what, param=None, filter=None):
        if isinstance(what, LocalElementIntegralBilinearForm):
            for part in _parts(what.integrand):
                self._append_element(part, what.over_integrate, _scaling(param))
        elif isinstance(what, LocalCouplingIntersectionIntegralBilinearForm):
            if filter is not None and not isinstance(filter, ApplyOn.InnerIntersectionsOnce):
                raise L.not_supported(9, "coupling forms are applied with ApplyOn.InnerIntersectionsOnce")
            for part in _parts(what.integrand):
                self._append_coupling(part, what.over_integrate, _scaling(param))
        elif isinstance(what, LocalIntersectionIntegralBilinearForm):
            self._set_boundary_filter(filter)
            try:
                for part in _parts(what.integrand):
                    self._append_boundary(part, what.over_integrate, _scaling(param))
            finally:
                self._set_boundary_filter(None)
        elif isinstance(what, VectorFunctional):
            if what.space is not self.space:
                raise L.not_supported(9, "the functional must live on the operator's space")
            for lf in what.local_functionals:
                self._append_source(lf)
            what._op = self
        elif isinstance(what, DirichletConstraints):
            mask = (None if what.boundary_info is None
                    else ApplyOn.CustomBoundaryIntersections(what.boundary_info, DirichletBoundary).mask(self.grid))
            L.call("hgdt_operator_set_dirichlet_faces", self._h,
                   None if mask is None else ctypes.c_void_p(mask.ctypes.data), 0 if mask is None else len(mask))
            L.call("hgdt_operator_set_dirichlet", self._h, 1)
            what._op = self
        elif isinstance(what, tuple):
            return self.append(*what)
        else:
            raise L.wrong_input_given(1, f"cannot append {type(what).__name__}")
        self._assembled = False
        return self

    def __iadd__(self, what):
        return self.append(what)

    def _set_boundary_filter(self, filter):
        """ApplyOn filter of the boundary forms / functionals appended next (None or BoundaryIntersections: all)"""
        mask = filter.mask(self.grid) if isinstance(filter, ApplyOn.CustomBoundaryIntersections) else None
        if filter is not None and not isinstance(filter, (ApplyOn.CustomBoundaryIntersections,
                                                          ApplyOn.BoundaryIntersections)):
            raise L.not_supported(9, "boundary forms are applied on boundary intersections")
        L.call("hgdt_operator_set_boundary_filter", self._h,
               None if mask is None else ctypes.c_void_p(mask.ctypes.data), 0 if mask is None else len(mask))

    def _tensor_field(self, t):
        """the per-element device tensor (elements, d, d) of a TensorGridFunction (a constant one expanded)"""
        d, ne = self.grid.dimension, self.grid.num_elements
        f = t.field if t.field is not None else torch.from_numpy(
            np.ascontiguousarray(np.broadcast_to(t.constant, (ne, d, d)))).to(_device())
        if tuple(f.shape) != (ne, d, d):
            raise L.shapes_do_not_match(2, f"diffusion tensor of shape {tuple(f.shape)}, expected ({ne}, {d}, {d})")
        self._keep.append(f)
        return ctypes.c_void_p(f.data_ptr())

    def _face_tensors(self, boundary, diffusion, weight):
        """matrix-valued coefficients of the intersection form just appended"""
        kt = self._tensor_field(diffusion) if isinstance(diffusion, TensorGridFunction) else None
        wt = self._tensor_field(weight) if isinstance(weight, TensorGridFunction) else None
        if kt is not None or wt is not None:
            L.call("hgdt_operator_set_face_tensors", self._h, int(boundary), kt, wt)

    def _field(self, gf):
        if isinstance(gf, TensorGridFunction):
            return None, 1.0  # replaced by the tensor (_face_tensors)
        if not isinstance(gf, GridFunction):
            raise L.not_supported(9, "intersection integrands take constant or per-element coefficients")
        ptr, val = gf.args(self.grid)
        if gf.field is not None:
            self._keep.append(gf.field)
        return ptr, val

    def _append_element(self, integrand, over, scaling):
        if isinstance(integrand, LocalLaplaceIntegrand):
            coeff, kind = integrand.diffusion, "laplace"
        elif isinstance(integrand, LocalElementProductIntegrand):
            coeff, kind = integrand.weight, "product"
        else:
            raise L.not_supported(9, f"{type(integrand).__name__} is not an element integrand of this path")
        if isinstance(coeff, TensorGridFunction):
            if kind != "laplace":
                raise L.not_supported(9, "matrix-valued coefficients are taken by LocalLaplaceIntegrand only")
            d = self.grid.dimension
            if coeff.field is not None:
                if tuple(coeff.field.shape) != (self.grid.num_elements, d, d):
                    raise L.shapes_do_not_match(2, f"diffusion tensor of shape {tuple(coeff.field.shape)}, expected "
                                                   f"({self.grid.num_elements}, {d}, {d})")
                self._keep.append(coeff.field)
                L.call("hgdt_operator_append_laplace_tensor", self._h, ctypes.c_void_p(coeff.field.data_ptr()), None,
                       over, scaling)
            else:
                if coeff.constant.shape != (d, d):
                    raise L.shapes_do_not_match(2, f"diffusion tensor of shape {coeff.constant.shape}, expected ({d}, {d})")
                L.call("hgdt_operator_append_laplace_tensor", self._h, None,
                       coeff.constant.ctypes.data_as(ctypes.c_void_p), over, scaling)
        elif isinstance(coeff, GridFunction):
            p, v = self._field(coeff)
            L.call(f"hgdt_operator_append_{kind}", self._h, p, v, over, scaling)
        else:
            # varying inside the elements: values at the points of the rule of order c + 2 k + over_integrate
            # (laplace.hh:74-79, product.hh:88-100), c the coefficient's order
            k = self.space.order
            if isinstance(coeff, GenericFunction):
                c = coeff.order()
                vals = _dvec(coeff.values_at(self.grid, c + 2 * k + over), name="coefficient values")
            else:
                c = coeff.order_total - 2 * k - over
                if c < 0:
                    raise L.wrong_input_given(1, f"values at the points of order {coeff.order_total}: the rule of "
                                                 f"this form has order >= {2 * k + over}")
                vals = coeff.values
            need = int(self.grid.quadrature_offsets(c + 2 * k + over)[-1])
            if vals.numel() != need:
                raise L.shapes_do_not_match(2, f"{vals.numel()} coefficient values, the rule of order "
                                               f"{c + 2 * k + over} has {need} points on this grid")
            self._keep.append(vals)
            L.call(f"hgdt_operator_append_{kind}_points", self._h, ctypes.c_void_p(vals.data_ptr()), c, over,
                   scaling)
        self._eforms += 1

    def _append_coupling(self, integrand, over, scaling):
        if isinstance(integrand, LocalLaplaceIPDGInnerCoupling):
            kp, kv = self._field(integrand.diffusion)
            wp, wv = self._field(integrand.weight)
            L.call("hgdt_operator_append_inner_coupling", self._h, integrand.symmetry_prefactor, kp, kv, wp, wv,
                   over, scaling)
            self._face_tensors(False, integrand.diffusion, integrand.weight)
        elif isinstance(integrand, LocalIPDGInnerPenalty):
            wp, wv = self._field(integrand.weight)
            L.call("hgdt_operator_append_inner_penalty", self._h, integrand.penalty, wp, wv, over, scaling)
            self._face_tensors(False, None, integrand.weight)
        else:
            raise L.not_supported(9, f"{type(integrand).__name__} is not a coupling integrand of this path")

    def _append_boundary(self, integrand, over, scaling):
        if isinstance(integrand, LocalLaplaceIPDGDirichletCoupling):
            kp, kv = self._field(integrand.diffusion)
            L.call("hgdt_operator_append_dirichlet_coupling", self._h, integrand.symmetry_prefactor, kp, kv, over,
                   scaling)
            self._face_tensors(True, integrand.diffusion, None)
        elif isinstance(integrand, LocalIPDGBoundaryPenalty):
            wp, wv = self._field(integrand.weight)
            L.call("hgdt_operator_append_boundary_penalty", self._h, integrand.penalty, wp, wv, over, scaling)
            self._face_tensors(True, None, integrand.weight)
        else:
            raise L.not_supported(9, f"{type(integrand).__name__} is not a boundary integrand of this path")

    def _append_boundary_source(self, lf):
        it, over = lf.integrand, lf.over_integrate
        k = self.space.order
        if isinstance(it, LocalLaplaceIPDGDirichletCoupling):
            kind, param, coeff, g = L.BOUNDARY_DIRICHLET_COUPLING, it.symmetry_prefactor, it.diffusion, it.dirichlet_data
        else:
            kind, param, coeff, g = it.kind, it.param, it.coeff, it.function
        g_order = g.order() if isinstance(g, GenericFunction) else 0
        order = (g_order + max(k - 1, 0) + over) if kind == L.BOUNDARY_DIRICHLET_COUPLING else (k + g_order + over)
        cp, cv = self._field(coeff)
        if isinstance(g, GenericFunction):
            pts = self.grid.boundary_quadrature_points(order)
            vals = _dvec(np.asarray(g.evaluate(pts.reshape(-1, self.grid.dimension).T), np.float64).reshape(-1),
                         name="boundary data")
            self._keep.append(vals)
            L.call("hgdt_operator_append_boundary_source", self._h, kind, param, cp, cv,
                   ctypes.c_void_p(vals.data_ptr()), 0.0, g_order, over)
        else:
            L.call("hgdt_operator_append_boundary_source", self._h, kind, param, cp, cv, None, float(g), 0, over)
        if isinstance(coeff, TensorGridFunction):
            L.call("hgdt_operator_set_boundary_source_tensor", self._h, self._tensor_field(coeff))

    def _append_source(self, lf):
        if isinstance(lf, LocalIntersectionIntegralFunctional):
            self._set_boundary_filter(getattr(lf, "_filter", None))
            try:
                return self._append_boundary_source(lf)
            finally:
                self._set_boundary_filter(None)
        f = lf.integrand.function
        over = lf.over_integrate
        if isinstance(f, QuadraturePointFunction):
            f_order = f.order_total - self.space.order - over
            if f_order < 0:
                raise L.wrong_input_given(1, "quadrature order lower than the functional's own order")
            expected = int(self.grid.quadrature_offsets(f.order_total)[-1])
            if f.values.numel() != expected:
                raise L.shapes_do_not_match(2, "quadrature-point values do not match the grid's rule")
            self._keep.append(f.values)
            L.call("hgdt_operator_append_source", self._h, L.SOURCE_PER_QUADRATURE_POINT,
                   ctypes.c_void_p(f.values.data_ptr()), 0.0, f_order, over)
            return
        if isinstance(f, GenericFunction):
            order_total = self.space.order + f.order() + over
            vals = _dvec(f.values_at(self.grid, order_total), name="source values")
            self._keep.append(vals)
            L.call("hgdt_operator_append_source", self._h, L.SOURCE_PER_QUADRATURE_POINT,
                   ctypes.c_void_p(vals.data_ptr()), 0.0, f.order(), over)
            return
        gf = _gf(f)
        p, v = self._field(gf)
        kind = L.SOURCE_CONSTANT if p is None else L.SOURCE_PER_ELEMENT
        L.call("hgdt_operator_append_source", self._h, kind, p, v, 0, over)

    # -- assemble (matrix-based.hh:496-500) --
    def assemble(self, parallel=False, stream=None):
        L.call("hgdt_operator_assemble", self._h, _stream_ptr(stream))
        self._assembled = True
        return self

    def pattern(self):
        if self._pattern is None:
            rp = np.empty(self.shape[0] + 1, np.int64)
            cols = np.empty(self.nnz, np.int32)
            L.call("hgdt_operator_copy_pattern", self._h, rp.ctypes.data_as(L._pi64), cols.ctypes.data_as(L._pi32))
            self._pattern = (rp, cols)
        return self._pattern

    def values_device_ptr(self):
        p = ctypes.c_void_p()
        L.call("hgdt_operator_values_device", self._h, ctypes.byref(p))
        return p

    @property
    def matrix(self):
        if self._borrowed is not None:  # the caller's matrix itself (its values are the operator's)
            return self._borrowed
        rp, cols = self.pattern()
        vals = torch.empty(self.nnz, dtype=torch.float64, device=_device())
        L.call("hgdt_memcpy", ctypes.c_void_p(vals.data_ptr()), self.values_device_ptr(), 8 * self.nnz,
               _stream_ptr())
        return CsrMatrix(rp, cols, vals, self.shape)

    @property
    def rhs(self):
        p = ctypes.c_void_p()
        L.call("hgdt_operator_rhs_device", self._h, ctypes.byref(p))
        out = torch.empty(self.shape[0], dtype=torch.float64, device=_device())
        L.call("hgdt_memcpy", ctypes.c_void_p(out.data_ptr()), p, 8 * self.shape[0], _stream_ptr())
        return out

    def dirichlet_dofs(self):
        cnt = ctypes.c_int64()
        L.call("hgdt_operator_dirichlet_dofs", self._h, ctypes.byref(cnt), None)
        out = np.empty(cnt.value, np.int32)
        L.call("hgdt_operator_dirichlet_dofs", self._h, ctypes.byref(cnt), out.ctypes.data_as(L._pi32))
        return out

    # -- operator interface (operators/interfaces.hh, matrix-based.hh:121-159) --
    def apply(self, source, range=None, stream=None):
        x = _dvec(source, self.shape[1], "source")
        y = torch.empty(self.shape[0], dtype=torch.float64, device=x.device) if range is None else range
        L.call("hgdt_operator_apply", self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
               _stream_ptr(stream))
        return y

    def invert_options(self, type=None):
        if type is None:
            return ["cg.jacobi"]
        if type != "cg.jacobi":
            raise L.operator_error(3, f"unknown inverse type {type!r}")
        return {"type": "cg.jacobi", "precision": 1e-10, "max_iter": 100000}

    def apply_inverse(self, range, source=None, opts=None, stream=None, check_every=16, use_graph=False,
                      time_spmv=False):
        """Jacobi-PCG from x0 = 0 (or from `source` when opts['initial_guess'] is true).

        `range` may be the string "rhs" to solve with the operator's own assembled right-hand side
        (no copy)."""
        if self.source_space is not self.space:
            raise L.not_supported(9, "apply_inverse needs an operator whose source and range spaces are the same")
        if isinstance(range, str) and range == "rhs":
            p = ctypes.c_void_p()
            L.call("hgdt_operator_rhs_device", self._h, ctypes.byref(p))
            b_ptr = p
            b_like = torch.empty(self.shape[0], dtype=torch.float64, device=_device())
            return self._apply_inverse(b_ptr, b_like if source is None else source, source is not None, opts,
                                       stream, check_every, use_graph, time_spmv)
        b = _dvec(range, self.shape[0], "range")
        return self._apply_inverse(ctypes.c_void_p(b.data_ptr()), torch.empty_like(b) if source is None else source,
                                   source is not None, opts, stream, check_every, use_graph, time_spmv, keep=b)

    def _apply_inverse(self, b_ptr, x, has_source, opts, stream, check_every, use_graph, time_spmv, keep=None):
        opts = dict(self.invert_options("cg.jacobi"), **(opts or {}))
        if opts.get("type", "cg.jacobi") != "cg.jacobi":
            raise L.operator_error(3, f"unknown inverse type {opts['type']!r}")
        x = _dvec(x, self.shape[1], "source")
        o = L.SolverOptions()
        L.call("hgdt_solver_options_default", ctypes.byref(o))
        o.rtol = float(opts.get("precision", 1e-10))
        o.max_iter = int(opts.get("max_iter", 100000))
        o.check_every = int(check_every)
        o.use_graph = int(bool(use_graph))
        o.time_spmv = int(time_spmv)  # True -> every application, k -> every k-th
        o.zero_initial_guess = 0 if (has_source and opts.get("initial_guess", False)) else 1
        o.restarts = int(opts.get("restarts", 0))  # restarts from the true residual (no reference counterpart)
        st = L.SolverStats()
        status = L.load().hgdt_operator_apply_inverse(self._h, b_ptr, ctypes.c_void_p(x.data_ptr()), ctypes.byref(o),
                                                      ctypes.byref(st), _stream_ptr(stream))
        self.last_solve = st.as_dict()
        L.check(status)
        return x

    def instance_stats(self):
        n = self.grid.num_instances
        it = np.empty(n, np.int64)
        rr = np.empty(n)
        L.call("hgdt_operator_instance_stats", self._h, it.ctypes.data_as(L._pi64), rr.ctypes.data_as(L._pdbl))
        return it, rr

    def adjoint_log_diffusion_gradient(self, u, lam, form_index=0, stream=None):
        """grad_e = -kappa_e lambda_e^T Khat_e u_e (no reference counterpart)."""
        u = _dvec(u, self.shape[0], "u")
        lam = _dvec(lam, self.shape[0], "lambda")
        g = torch.empty(self.grid.num_elements, dtype=torch.float64, device=u.device)
        L.call("hgdt_operator_adjoint_log_diffusion_gradient", self._h, form_index, ctypes.c_void_p(u.data_ptr()),
               ctypes.c_void_p(lam.data_ptr()), ctypes.c_void_p(g.data_ptr()), _stream_ptr(stream))
        return g

    # -- storage / launch configuration --
    _FORMATS = {"auto": L.FORMAT_AUTO, "sell64": L.FORMAT_SELL, "dia_symmetric": L.FORMAT_DIA,
                "dia": L.FORMAT_DIA_NONSYMMETRIC, "matrix_free": L.FORMAT_MATRIX_FREE, "block4": L.FORMAT_BLOCK}

    def set_format(self, fmt):
        """storage of apply / apply_inverse: "auto", "sell64", "dia_symmetric", "dia", "block4" (4 x 4
        element blocks of a discontinuous space with 4 DoFs per element, the default for such spaces) or
        "matrix_free" (no stored matrix: the operator is recomputed from the coefficient fields at every
        application, as LocalizableOperator::apply does; assemble() then builds the right-hand side only)"""
        L.call("hgdt_operator_set_format", self._h, self._FORMATS[fmt] if isinstance(fmt, str) else int(fmt))

    def format_info(self):
        f, b, d = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        L.call("hgdt_operator_format_info", self._h, ctypes.byref(f), ctypes.byref(b), ctypes.byref(d))
        return {"format": {L.FORMAT_SELL: "sell64", L.FORMAT_DIA: "dia_symmetric",
                           L.FORMAT_DIA_NONSYMMETRIC: "dia", L.FORMAT_MATRIX_FREE: "matrix_free",
                           L.FORMAT_BLOCK: "block4"}[f.value],
                "spmv_bytes": int(b.value),
                "num_diagonals": int(d.value)}

    def assembly_info(self):
        """{"structured": the last assemble() wrote the solver's storage directly (P1 Kuhn tetrahedra on
        axis-parallel lattices, Laplace forms), "solver_upper_diagonals": strict upper diagonals kept in the
        solver's scaled symmetric DIA copy (-1: other storage, or not prepared since the assembly)}"""
        st, su = ctypes.c_int32(), ctypes.c_int32()
        L.call("hgdt_operator_assembly_info", self._h, ctypes.byref(st), ctypes.byref(su))
        return {"structured": bool(st.value), "solver_upper_diagonals": int(su.value)}

    def assembly_bytes(self):
        """algorithmic HBM bytes of the last assemble(): the structured assembly's one pass, otherwise the
        compulsory bytes (per-element inputs read, values and right-hand side written)"""
        b = ctypes.c_int64()
        L.call("hgdt_operator_assembly_bytes", self._h, ctypes.byref(b))
        return b.value

    def set_launch(self, kernel, block_size, items_per_thread=1):
        L.call("hgdt_operator_set_launch", self._h, int(kernel), int(block_size), int(items_per_thread))

    def set_column_staging(self, mode):
        """p entries of the solver's symmetric DIA application staged in LDS: -1 automatic (the default:
        when the application moves more than 256 MiB), 0 never, 1 wherever the stencil allows it"""
        L.call("hgdt_operator_set_column_staging", self._h, int(mode))

    def get_launch(self, kernel):
        b, i = ctypes.c_int32(), ctypes.c_int32()
        L.call("hgdt_operator_get_launch", self._h, int(kernel), ctypes.byref(b), ctypes.byref(i))
        return b.value, i.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and L._lib is not None:
            L._lib.hgdt_operator_destroy(h)
            self._h = None


def _parts(integrand):
    return integrand.parts if isinstance(integrand, _IntegrandSum) else (integrand,)


def _scaling(param):
    if not param:
        return 1.0
    v = param.get("matrixoperator.scaling", 1.0)
    return float(v[0] if isinstance(v, (list, tuple)) else v)


def make_matrix_operator(space, stencil=Stencil.element_and_intersection, logging_prefix="", range_space=None,
                         matrix=None):
    """make_matrix_operator<M>(space, stencil) (matrix-based.hh:650-658), make_matrix_operator<M>(source_space,
    range_space, stencil) (:200-216) with `range_space`, and the variants on a given matrix (:511-546)"""
    return MatrixOperator(space.grid, space, space if range_space is None else range_space, stencil=stencil,
                          matrix=matrix)


# ---------------------------------------------------------------------------
# discrete functions and bilinear forms on them (error norms)
# ---------------------------------------------------------------------------
class DiscreteFunction:
    """DiscreteFunction(space, vector) (discretefunction/default.hh): a DoF vector on a space."""

    def __init__(self, space, vector=None, name=""):
        self.space, self.name = space, name
        self.vector = (torch.zeros(space.num_DoFs, dtype=torch.float64, device=_device()) if vector is None
                       else _dvec(vector, space.num_DoFs, "DoF vector"))

    def dofs(self):
        return self.vector

    def order(self):
        return self.space.order

    def __sub__(self, other):
        return _Difference(self, other)


def make_discrete_function(space, vector=None, name=""):
    return DiscreteFunction(space, vector, name)


class _Difference:
    def __init__(self, discrete, function):
        if not isinstance(discrete, DiscreteFunction) or not isinstance(function, (GenericFunction, type(None))):
            raise L.not_supported(9, "differences are DiscreteFunction - GenericFunction")
        self.discrete, self.function = discrete, function

    def order(self):
        return max(self.discrete.order(), self.function.order() if self.function else 0)


class BilinearForm:
    """BilinearForm(grid_view, source, range) += LocalElementIntegralBilinearForm(...); apply2()
    (operators/bilinear-form.hh:344-452): with source = range = u_h - g and the Laplace or product
    integrand this is the squared energy (H1-semi) or L2 norm of the error, computed on the device
    by one element kernel and a fixed-order sum."""

    def __init__(self, grid, source, range=None, logging_prefix=""):
        range = source if range is None else range
        if range is not source:
            raise L.not_supported(9, "apply2 is provided for source == range (norms)")
        self.grid = grid
        self.function = source if isinstance(source, _Difference) else _Difference(source, None)
        self.local_forms = []

    def append(self, local_form):
        if not isinstance(local_form, LocalElementIntegralBilinearForm) or not isinstance(
                local_form.integrand, (LocalLaplaceIntegrand, LocalElementProductIntegrand)):
            raise L.not_supported(9, "bilinear forms of element Laplace / product integrals")
        self.local_forms.append(local_form)
        return self

    __iadd__ = append

    def apply2_per_instance(self, stream=None):
        d = self.function.discrete
        space, g = d.space, self.function.function
        ninst = self.grid.num_instances
        total = np.zeros(ninst)
        for lf in self.local_forms:
            it = lf.integrand
            order = 0 + 2 * self.function.order() + lf.over_integrate
            gv = gg = None
            if g is not None:
                gv = _dvec(g.values_at(self.grid, order), name="function values")
                if isinstance(it, LocalLaplaceIntegrand):
                    gg = _dvec(g.gradients_at(self.grid, order), name="function gradients")
            out = np.empty(2 * ninst)
            coeff = it.diffusion if isinstance(it, LocalLaplaceIntegrand) else it.weight
            if not isinstance(coeff, GridFunction):
                raise L.not_supported(9, "error norms take constant or per-element coefficients")
            if isinstance(it, LocalLaplaceIntegrand):
                kp, kv = it.diffusion.args(self.grid)
                comp, factor = 1, 1.0
            else:
                if it.weight.field is not None:
                    raise L.not_supported(9, "product norms with a constant weight")
                kp, kv, comp, factor = None, 1.0, 0, it.weight.constant
            L.call("hgdt_space_error_norms", space._h, ctypes.c_void_p(d.vector.data_ptr()), order,
                   ctypes.c_void_p(gv.data_ptr()) if gv is not None else None,
                   ctypes.c_void_p(gg.data_ptr()) if gg is not None else None, kp, kv,
                   out.ctypes.data_as(L._pdbl), _stream_ptr(stream))
            total += factor * out[comp::2]
        return total

    def apply2(self, stream=None):
        return float(self.apply2_per_instance(stream).sum())

    result = apply2


def make_bilinear_form(grid, source, range=None, logging_prefix=""):
    return BilinearForm(grid, source, range, logging_prefix)
