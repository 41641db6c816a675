/*
 * hgdt.h -- C ABI of the MI355X-native assembly + solve library (libhgdt.so).
 *
 * This is the drop-in boundary for dune-gdt's global assembly and
 * operator-application path.  Each entry point names the reference interface it
 * replaces (paths relative to the dune-gdt source tree).  No C++ or torch types
 * cross this boundary: handles are opaque pointers, arrays are plain pointers
 * plus sizes, device arrays are HIP device pointers, streams are hipStream_t
 * passed as void* (NULL = the default stream).
 *
 * Conventions
 *   - every function returns an hgdt_status; 0 is success.  The message of the
 *     last failure on the calling thread is returned by hgdt_last_error().
 *     Status codes map the reference's exception families
 *     (dune/gdt/exceptions.hh:24-77).
 *   - all floating-point data is double (FP64).
 *   - geometry: HGDT_QUAD / HGDT_TRI (2d), HGDT_HEX / HGDT_TET (3d), HGDT_MIXED
 *     (2d grids of triangles and quadrilaterals).  Reference cube vertex b has
 *     coordinates given by the bits of b (x = bit 0, y = bit 1, z = bit 2);
 *     reference simplex vertex 0 is the origin and vertex k is e_k.
 *   - a handle must not be used from two host threads at the same time.
 *   - results are deterministic: no floating-point atomics anywhere; every sum
 *     is taken in a fixed order.
 */
#ifndef HGDT_H
#define HGDT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hgdt_status {
  HGDT_OK = 0,
  HGDT_ERR_INVALID_ARGUMENT = 1,   /* XT::Common::Exceptions::wrong_input_given */
  HGDT_ERR_SHAPES_DO_NOT_MATCH = 2, /* XT::Common::Exceptions::shapes_do_not_match */
  HGDT_ERR_OPERATOR = 3,           /* Exceptions::operator_error */
  HGDT_ERR_INTEGRAND = 4,          /* Exceptions::integrand_error */
  HGDT_ERR_SPACE = 5,              /* Exceptions::space_error */
  HGDT_ERR_MAPPER = 6,             /* Exceptions::mapper_error */
  HGDT_ERR_SOLVER_FAILED = 7,      /* XT::LA::Exceptions::linear_solver_failed */
  HGDT_ERR_DEVICE = 8,             /* HIP runtime failure */
  HGDT_ERR_NOT_SUPPORTED = 9,      /* combination not provided by this library */
  HGDT_ERR_OUT_OF_MEMORY = 10
} hgdt_status;

enum { HGDT_QUAD = 0, HGDT_TRI = 1, HGDT_HEX = 2, HGDT_TET = 3, HGDT_MIXED = 4 };
/* space kinds */
enum { HGDT_CONTINUOUS_LAGRANGE = 0, HGDT_DISCONTINUOUS_LAGRANGE = 1 };
/* Stencil, dune/gdt/type_traits.hh:55 */
enum { HGDT_STENCIL_ELEMENT = 0, HGDT_STENCIL_INTERSECTION = 1, HGDT_STENCIL_ELEMENT_AND_INTERSECTION = 2 };
/* source data kinds for the right-hand side functional */
enum { HGDT_SOURCE_CONSTANT = 0, HGDT_SOURCE_PER_ELEMENT = 1, HGDT_SOURCE_PER_QUADRATURE_POINT = 2 };
/* right-hand side terms on boundary intersections */
enum { HGDT_BOUNDARY_DIRICHLET_COUPLING = 0, HGDT_BOUNDARY_PENALTY = 1, HGDT_BOUNDARY_PRODUCT = 2 };
/* matrix storage used for apply / apply_inverse (HGDT_FORMAT_MATRIX_FREE: no stored matrix) */
enum {
  HGDT_FORMAT_AUTO = 0,
  HGDT_FORMAT_SELL = 1,
  HGDT_FORMAT_DIA = 2,
  HGDT_FORMAT_DIA_NONSYMMETRIC = 3,
  HGDT_FORMAT_MATRIX_FREE = 4,
  HGDT_FORMAT_BLOCK = 5
};

typedef struct hgdt_grid_s* hgdt_grid;
typedef struct hgdt_space_s* hgdt_space;
typedef struct hgdt_operator_s* hgdt_operator;

/* message of the last failure on this thread ("" if none) */
const char* hgdt_last_error(void);
/* library version as major*10000 + minor*100 + patch */
int hgdt_version(void);
/* HIP device count visible to the library; fails with HGDT_ERR_DEVICE when no GPU is usable */
hgdt_status hgdt_device_count(int* count);

/* stream-ordered device-to-device (or host<->device) copy of `bytes` bytes; utility for bindings that
 * hold device buffers the library does not own (no reference counterpart) */
hgdt_status hgdt_memcpy(void* dst, const void* src, int64_t bytes, void* stream);

/* device memory and stream utilities for bindings that hold no HIP runtime of their own, such as the
 * C++ classes of hgdt.hh (no reference counterpart: the reference's containers live in host memory) */
hgdt_status hgdt_device_alloc(int64_t bytes, void** ptr);
hgdt_status hgdt_device_free(void* ptr);
hgdt_status hgdt_stream_synchronize(void* stream);

/* ------------------------------------------------------------------------- */
/* grids                                                                     */
/* ------------------------------------------------------------------------- */

/* XT::Grid::make_cube_grid<G>(lower_left, upper_right, num_elements) as used in
 * examples/stationary-heat-equation.cc:87 and dune/gdt/test/integrands/integrands.hh:64-70.
 * cells[0..dim-1] > 0.  Simplices: every cube split into dim! Kuhn simplices
 * (v0, v0+e_p0, v0+e_p0+e_p1[, v7]) for the axis permutations p in lexicographic order. */
hgdt_status hgdt_grid_create_structured(int geometry, const int32_t* cells, const double* lower_left,
                                        const double* upper_right, hgdt_grid* grid);
/* grid from explicit host arrays: coords [num_vertices x dim], connectivity
 * [num_elements x vertices_per_element] in the reference vertex order above.
 * Cube elements must be affine (parallelograms / parallelepipeds). */
hgdt_status hgdt_grid_create(int geometry, int64_t num_vertices, const double* coords, int64_t num_elements,
                             const int32_t* connectivity, hgdt_grid* grid);
/* grid of triangles and quadrilaterals, element by element as GridFactory<G>::insertElement(
 * GeometryTypes::simplex(2) or cube(2), vertices) builds the mixed grids of
 * dune/gdt/test/spaces/base.hh:355-376 (make_mixed_grid, a UGGrid): coords [num_vertices x 2]; element e has
 * type geometry[e] (HGDT_TRI or HGDT_QUAD) and the vertices vertices[offsets[e] .. offsets[e + 1]) in the
 * reference vertex order above.  Quadrilaterals must be parallelograms.  The grid's geometry
 * (hgdt_grid_info) is HGDT_MIXED, with 4 vertices and 4 faces per element (its connectivity has 4 entries
 * per element, -1 after a triangle's third), unless every element has one type: then it is a grid of that
 * type as hgdt_grid_create makes it.  Spaces of order 1 and 2 on mixed grids take element forms, right-hand
 * sides, Dirichlet constraints, intersection integrals (faces between a triangle and a quadrilateral
 * included), boundary functionals, error norms and the adjoint kernel; matrix-free application and unions
 * need one element type. */
hgdt_status hgdt_grid_create_mixed(int64_t num_vertices, const double* coords, int64_t num_elements,
                                   const int32_t* geometry, const int64_t* offsets, const int32_t* vertices,
                                   hgdt_grid* grid);
/* disjoint union of grids of one geometry type: the batch of independent problem
 * instances.  Vertices, elements and DoFs are numbered instance by instance. */
hgdt_status hgdt_grid_create_union(const hgdt_grid* grids, int32_t count, hgdt_grid* grid);
hgdt_status hgdt_grid_destroy(hgdt_grid grid);
/* info[0..8] = dim, num_vertices, num_elements, vertices_per_element, faces_per_element,
 *              num_inner_intersections, num_boundary_intersections, num_instances, geometry */
hgdt_status hgdt_grid_info(hgdt_grid grid, int64_t* info);
hgdt_status hgdt_grid_copy_coords(hgdt_grid grid, double* host_out);
hgdt_status hgdt_grid_copy_connectivity(hgdt_grid grid, int32_t* host_out);
/* element offsets of the instances, host_out[num_instances + 1] */
hgdt_status hgdt_grid_instance_elements(hgdt_grid grid, int64_t* host_out);
/* physical coordinates of the element quadrature points of the given order,
 * host_out[num_elements x num_points x dim]; *num_points receives the rule size.
 * Pass host_out = NULL to query the size only.  On a mixed grid the elements' rules differ
 * in size: *num_points is 0 and element e's points are host_out[offsets[e] .. offsets[e + 1])
 * (x dim) with the offsets of hgdt_grid_quadrature_offsets. */
hgdt_status hgdt_grid_quadrature_points(hgdt_grid grid, int order, double* host_out, int32_t* num_points);
/* where every element's quadrature points of the given order start in the layout of
 * hgdt_grid_quadrature_points (and of per-quadrature-point data: sources, error norms):
 * host_out[num_elements + 1] */
hgdt_status hgdt_grid_quadrature_offsets(hgdt_grid grid, int order, int64_t* host_out);

/* physical coordinates of the quadrature points (face rule of the given order: Gauss-Legendre on segments
 * in 2d, the triangle or tensor-Gauss square rule on the faces of tetrahedra / hexahedra) on the boundary
 * intersections, in walk order (elements ascending, their faces in reference order):
 * host_out[num_boundary_intersections x num_points x dim]; host_out = NULL queries the size. */
hgdt_status hgdt_grid_boundary_quadrature_points(hgdt_grid grid, int order, double* host_out, int32_t* num_points);

/* centres [num_boundary_intersections x dim] and outward unit normals [num_boundary_intersections x dim] of the
 * boundary intersections in walk order (elements ascending, their faces in reference order): what a boundary
 * info (XT::Grid::BoundaryInfo::type(intersection), e.g. NormalBasedBoundaryInfo) decides on.  Either output may
 * be NULL. */
hgdt_status hgdt_grid_boundary_intersections(hgdt_grid grid, double* centers, double* normals);

/* ------------------------------------------------------------------------- */
/* spaces                                                                    */
/* ------------------------------------------------------------------------- */

/* ContinuousLagrangeSpace<GV>(grid_view, order), dune/gdt/spaces/h1/continuous-lagrange.hh:68-79
 * DiscontinuousLagrangeSpace<GV>(grid_view, order), dune/gdt/spaces/l2/discontinuous-lagrange.hh:78-86
 * order 1 (P1 on simplices, Q1 on cubes) or 2 (P2 on triangles and tetrahedra, Q2 on quadrilaterals and
 * hexahedra; the local DoF order of the Lagrange nodes: cubes by tensor index, simplices vertices then edge
 * midpoints (a, b), a < b, lexicographic; continuous order 2 numbers one DoF per subentity in order of first
 * appearance in the element walk).  Order-2 spaces take element forms, right-hand sides, intersection
 * integrals, boundary functionals and the adjoint kernel. */
hgdt_status hgdt_space_create(hgdt_grid grid, int kind, int order, hgdt_space* space);
hgdt_status hgdt_space_destroy(hgdt_space space);
/* mapper().size() and max_local_size(), spaces/mapper/continuous.hh:139-150 */
hgdt_status hgdt_space_size(hgdt_space space, int64_t* size, int32_t* local_size);
/* mapper().global_indices(element) for all elements, host_out[num_elements x local_size]
 * (spaces/mapper/continuous.hh:139-150, spaces/mapper/discontinuous.hh:102-125); on a mixed grid
 * local_size is the largest element's (max_local_size()) and a triangle's row ends in -1 entries */
hgdt_status hgdt_space_copy_global_indices(hgdt_space space, int32_t* host_out);
/* DoF offsets of the instances of a union grid, host_out[num_instances + 1] */
hgdt_status hgdt_space_instance_dofs(hgdt_space space, int64_t* host_out);

/* Squared error norms of e = u_h - g, the discrete function u_h with DoF vector `dofs` (device, [space size])
 * minus a function g given at the points of hgdt_grid_quadrature_points(order): values
 * [num_elements x num_points] and physical gradients [num_elements x num_points x dim] (device arrays;
 * NULL for a zero value / gradient).  This is
 *   make_bilinear_form(grid_view, e, e) += LocalElementIntegralBilinearForm(LocalElementProductIntegrand())
 *   make_bilinear_form(grid_view, e, e) += LocalElementIntegralBilinearForm(LocalLaplaceIntegrand(diffusion))
 * followed by apply2() (operators/bilinear-form.hh:344-452, as in examples/stationary-heat-equation.cc:120-132),
 * with the quadrature order given explicitly (integrand order + over_integrate, integrals.hh:115).
 * host_out[2*i] = ||e||_L2^2 and host_out[2*i+1] = sum_e int (kappa grad e) . grad e for every instance i.
 * Element values are summed in a fixed order (deterministic).  Blocks until the results are on the host. */
hgdt_status hgdt_space_error_norms(hgdt_space space, const double* dofs, int order, const double* g_values,
                                   const double* g_gradients, const double* diffusion_field, double diffusion_value,
                                   double* host_out, void* stream);

/* No reference counterpart: hgdt_space_error_norms for callers that cannot allocate device memory (code
 * inside a tight serving or time-stepping loop, a harness that forbids allocation).  Same arguments and the
 * same results, bit for bit; its element values, chunk sums and tables live in one fixed device array of
 * hgdt_noalloc_scratch_bytes() bytes per GPU (16.75 MiB, held until the process ends), walked in windows of
 * 1 048 576 elements, so it calls no hipMalloc, hipHostMalloc or hipFree (host-side set-up uses ordinary
 * heap memory).  Limits: at most 8192 chunk-table entries (chunks of 65 536 elements, at least one per
 * instance, plus one per window) and a quadrature table of at most 65 536 doubles; beyond them it fails with
 * HGDT_ERR_NOT_SUPPORTED naming the limit.  Calls that overlap on one GPU (from two host threads or on two
 * streams) share the array and race.  Blocks until the results are on the host. */
hgdt_status hgdt_space_error_norms_noalloc(hgdt_space space, const double* dofs, int order, const double* g_values,
                                           const double* g_gradients, const double* diffusion_field,
                                           double diffusion_value, double* host_out, void* stream);
/* size of the fixed device array behind the no-allocation entry points (per GPU) */
hgdt_status hgdt_noalloc_scratch_bytes(int64_t* bytes);

/* ------------------------------------------------------------------------- */
/* matrix operator                                                           */
/* ------------------------------------------------------------------------- */

/* make_matrix_operator<M>(space, stencil), dune/gdt/operators/matrix-based.hh:650-658: builds the
 * sparsity pattern (dune/gdt/tools/sparsity-pattern.hh:33-144) and a zero matrix on it
 * (MatrixOperator ctor, matrix-based.hh:298-311).  Test and ansatz space are the same space. */
hgdt_status hgdt_operator_create(hgdt_space space, int stencil, hgdt_operator* op);
/* make_matrix_operator<M>(source_space, range_space, stencil) (matrix-based.hh:200-216 with the pattern
 * constructor :295-311): an operator from the source space (ansatz functions, columns) to the range
 * space (test functions, rows), rows x cols = range size x source size.  Both spaces on the same grid
 * with the same local finite element (the same polynomial order; a continuous and a discontinuous space
 * may be paired).  apply() maps source vectors to range vectors; apply_inverse and Dirichlet constraints
 * need source == range (HGDT_ERR_NOT_SUPPORTED otherwise).  Stored in SELL format. */
hgdt_status hgdt_operator_create_between(hgdt_space source, hgdt_space range, int stencil, hgdt_operator* op);
/* MatrixOperator(assembly_grid_view, source_space, range_space, matrix) (matrix-based.hh:279-293,
 * make_matrix_operator(..., matrix) :511-546): assembles into the caller's CSR matrix.  Host row_ptr[rows + 1]
 * and cols[nnz] (ascending and unique per row) give its pattern, which must hold every entry of the
 * stencil's (HGDT_ERR_SHAPES_DO_NOT_MATCH otherwise); rows = the range space's size.  The device values[nnz]
 * stay the caller's and must outlive the operator.  As the reference's local assemblers
 * (add_to_entry, local/assembler/bilinear-form-assemblers.hh:126,264-276,394), assemble() adds every
 * contribution to the value held, in walk order; with Dirichlet constraints the constrained rows and
 * columns are then set (tools/dirichlet-constraints.hh:153-184).  The operator's matrix is the caller's
 * whole matrix: apply, apply_inverse and copy_values use every entry, those outside the stencil included. */
hgdt_status hgdt_operator_create_on_matrix(hgdt_space source, hgdt_space range, int stencil, const int64_t* row_ptr,
                                           const int32_t* cols, int64_t rows, int64_t nnz, double* values,
                                           hgdt_operator* op);
hgdt_status hgdt_operator_destroy(hgdt_operator op);
/* rows, cols, nnz of the pattern */
hgdt_status hgdt_operator_shape(hgdt_operator op, int64_t* rows, int64_t* cols, int64_t* nnz);
/* CSR pattern (sorted rows), host arrays row_ptr[rows + 1], cols[nnz] */
hgdt_status hgdt_operator_copy_pattern(hgdt_operator op, int64_t* row_ptr, int32_t* cols);

/* Coefficient convention of the append calls: `field` is a device pointer to one value per
 * element (piecewise constant grid function, scalar times the identity) that must stay valid
 * until the operator is destroyed, or NULL to use the constant `value`.  `scaling` is the
 * "matrixoperator.scaling" parameter (matrix-based.hh:342-365, bilinear-form-assemblers.hh:64).
 * `over_integrate` raises the quadrature order (local/bilinear-forms/integrals.hh:115). */

/* append(LocalElementIntegralBilinearForm(LocalLaplaceIntegrand(diffusion), over_integrate)),
 * matrix-based.hh:346-369, local/integrands/laplace.hh:74-102 */
hgdt_status hgdt_operator_append_laplace(hgdt_operator op, const double* diffusion_field, double diffusion_value,
                                         int32_t over_integrate, double scaling);
/* append(LocalElementIntegralBilinearForm(LocalElementProductIntegrand(weight))),
 * local/integrands/product.hh:104-130 */
hgdt_status hgdt_operator_append_product(hgdt_operator op, const double* weight_field, double weight_value,
                                         int32_t over_integrate, double scaling);
/* the same two forms with a coefficient that varies inside the elements, a GridFunction of polynomial order
 * coeff_order (LocalLaplaceIntegrand(diffusion) / LocalElementProductIntegrand(weight) with
 * local_weight_->order() = coeff_order: laplace.hh:74-79, product.hh:88-100): `values` holds its values at
 * the points of hgdt_grid_quadrature_points(order), order = coeff_order + 2 k + over_integrate (k the space
 * order), element-ordered (hgdt_grid_quadrature_offsets on mixed grids); a device array that must stay
 * valid.  Not for the matrix-free format or the adjoint gradient (their coefficients are per element). */
hgdt_status hgdt_operator_append_laplace_points(hgdt_operator op, const double* values, int32_t coeff_order,
                                                int32_t over_integrate, double scaling);
/* the Laplace form with a matrix-valued diffusion, LocalLaplaceIntegrand(GridFunction<E, d, d>)
 * (laplace.hh:40-48,94-101: (K grad phi_j) . grad psi_i): tensor_field a device array of d x d values per
 * element (row-major, num_elements * d * d), or NULL and tensor_value d x d host values for every element
 * (copied).  K must be symmetric (checked; the local matrices are then symmetric and only their upper
 * triangles are computed).  Not for the matrix-free format or the adjoint gradient. */
hgdt_status hgdt_operator_append_laplace_tensor(hgdt_operator op, const double* tensor_field,
                                                const double* tensor_value, int32_t over_integrate, double scaling);
hgdt_status hgdt_operator_append_product_points(hgdt_operator op, const double* values, int32_t coeff_order,
                                                int32_t over_integrate, double scaling);
/* append(LocalCouplingIntersectionIntegralBilinearForm(LocalLaplaceIPDGIntegrands::InnerCoupling(
 *   symmetry_prefactor, diffusion, weight)), {}, ApplyOn::InnerIntersectionsOnce),
 * matrix-based.hh:371-393, local/integrands/laplace-ipdg.hh:47-186 */
hgdt_status hgdt_operator_append_inner_coupling(hgdt_operator op, double symmetry_prefactor,
                                                const double* diffusion_field, double diffusion_value,
                                                const double* weight_field, double weight_value,
                                                int32_t over_integrate, double scaling);
/* append(LocalCouplingIntersectionIntegralBilinearForm(LocalIPDGIntegrands::InnerPenalty(penalty, weight)),
 * {}, ApplyOn::InnerIntersectionsOnce); intersection diameter = face diameter
 * (local/integrands/ipdg.hh:27-38, 59-171) */
hgdt_status hgdt_operator_append_inner_penalty(hgdt_operator op, double penalty, const double* weight_field,
                                               double weight_value, int32_t over_integrate, double scaling);
/* append(LocalIntersectionIntegralBilinearForm(LocalLaplaceIPDGIntegrands::DirichletCoupling(
 *   symmetry_prefactor, diffusion)), {}, CustomBoundaryIntersections(all Dirichlet)),
 * matrix-based.hh:395-408, local/integrands/laplace-ipdg.hh:237-368 */
hgdt_status hgdt_operator_append_dirichlet_coupling(hgdt_operator op, double symmetry_prefactor,
                                                    const double* diffusion_field, double diffusion_value,
                                                    int32_t over_integrate, double scaling);
/* append(LocalIntersectionIntegralBilinearForm(LocalIPDGIntegrands::BoundaryPenalty(penalty, weight)), ...),
 * local/integrands/ipdg.hh:201-282 */
hgdt_status hgdt_operator_append_boundary_penalty(hgdt_operator op, double penalty, const double* weight_field,
                                                  double weight_value, int32_t over_integrate, double scaling);
/* matrix-valued diffusion and weight (GridFunction<E, d, d>) of the last appended inner (boundary 0) or
 * boundary (1) intersection form, replacing its scalar coefficients (laplace-ipdg.hh:47-61,145-186:
 * delta = n . (weight n), fluxes (diffusion grad phi) . n; ipdg.hh:59-72,149-153): device arrays of d x d
 * values per element (row-major, symmetric: checked), NULL to keep a scalar.  Not for the matrix-free format;
 * the boundary functionals take theirs through hgdt_operator_set_boundary_source_tensor. */
hgdt_status hgdt_operator_set_face_tensors(hgdt_operator op, int32_t boundary, const double* diffusion_tensor,
                                           const double* weight_tensor);
/* matrix-valued coefficient of the last appended boundary functional (the diffusion of the Dirichlet coupling
 * term, (K grad psi_i) . n; the weight of the boundary penalty term, n . (W n)), replacing its scalar coefficient:
 * a device array of d x d values per element (row-major, symmetric: checked).  Not for the product term. */
hgdt_status hgdt_operator_set_boundary_source_tensor(hgdt_operator op, const double* tensor);
/* right-hand side functional assembled in the same walk:
 * VectorBasedFunctional::append(LocalElementIntegralFunctional(LocalProductIntegrand().with_ansatz(f)))
 * (functionals/vector-based.hh:132-308, local/functionals/integrals.hh:72-97,
 *  local/assembler/functional-assemblers.hh:77-86).  kind: HGDT_SOURCE_*; for PER_ELEMENT `data` has one
 * value per element, for PER_QUADRATURE_POINT one value per point of hgdt_grid_quadrature_points(order)
 * with order = 1 + f_order + over_integrate.  `data` is a device pointer that must stay valid. */
hgdt_status hgdt_operator_append_source(hgdt_operator op, int kind, const double* data, double value,
                                        int32_t f_order, int32_t over_integrate);
/* right-hand side terms on the boundary intersections, assembled in the same walk:
 * VectorBasedFunctional::append(LocalIntersectionIntegralFunctional(integrand, over_integrate), {},
 * CustomBoundaryIntersections(all)) (local/functionals/integrals.hh:108-191) with
 *   HGDT_BOUNDARY_DIRICHLET_COUPLING: LocalLaplaceIPDGIntegrands::DirichletCoupling(param, coeff, g) as a unary
 *     integrand, -param g (coeff grad psi_i) . n (local/integrands/laplace-ipdg.hh:237-320);
 *   HGDT_BOUNDARY_PENALTY: LocalIPDGIntegrands::BoundaryPenalty(param, coeff).with_ansatz(g),
 *     param (n . coeff n) / h_F g psi_i (local/integrands/ipdg.hh:201-282, conversion.hh:133-221);
 *   HGDT_BOUNDARY_PRODUCT: LocalIntersectionProductIntegrand(coeff).with_ansatz(g), (coeff g) psi_i
 *     (local/integrands/product.hh:280-380).
 * coeff: per-element device field or coeff_value (NULL field).  g: device values at the points of
 * hgdt_grid_boundary_quadrature_points(order) with order = g_order + max(k - 1, 0) + over_integrate for the
 * coupling term and k + g_order + over_integrate for the others (k the space order), or the constant
 * `value` (NULL data).  Device arrays must stay valid.  Order-1 spaces. */
hgdt_status hgdt_operator_append_boundary_source(hgdt_operator op, int kind, double param, const double* coeff_field,
                                                 double coeff_value, const double* data, double value, int32_t g_order,
                                                 int32_t over_integrate);
/* ApplyOn::CustomBoundaryIntersections(boundary_info, type) for the boundary bilinear forms and boundary
 * functionals appended from now on (test/stationary-eocstudies/diffusion-ipdg.hh:302-311): one byte per boundary
 * intersection in walk order (see hgdt_grid_boundary_intersections), nonzero where boundary_info.type(intersection)
 * is the type; NULL: every boundary intersection (the default).  The mask is copied. */
hgdt_status hgdt_operator_set_boundary_filter(hgdt_operator op, const uint8_t* mask, int64_t count);
/* the Dirichlet intersections of the constraints, boundary_info.type(intersection) == DirichletBoundary()
 * (tools/dirichlet-constraints.hh:85-110): DoFs on those intersections and their subentities are constrained; one
 * byte per boundary intersection in walk order; NULL: every boundary intersection (AllDirichletBoundaryInfo, the
 * default).  The mask is copied. */
hgdt_status hgdt_operator_set_dirichlet_faces(hgdt_operator op, const uint8_t* mask, int64_t count);
/* DirichletConstraints(space, AllDirichletBoundaryInfo) applied after assembly with
 * ensure_symmetry (tools/dirichlet-constraints.hh:85-184): unit rows and columns, rhs entry 0 */
hgdt_status hgdt_operator_set_dirichlet(hgdt_operator op, int32_t enabled);
/* the variant of DirichletConstraints::apply(matrix, vector, only_clear, ensure_symmetry)
 * (tools/dirichlet-constraints.hh:152-184) used by the next assemble(): the rows of the constrained DoFs
 * unit (only_clear 0, the default) or cleared (1), and with ensure_symmetry (1, the default) their columns
 * too; the rhs entries are 0 in every variant.  Jacobi-PCG takes the default (it refuses unit rows without
 * their columns, and cleared rows have no positive diagonal); apply() takes every variant. */
hgdt_status hgdt_operator_set_dirichlet_mode(hgdt_operator op, int32_t only_clear, int32_t ensure_symmetry);
/* number of constrained DoFs and their (sorted) indices, host_out may be NULL */
hgdt_status hgdt_operator_dirichlet_dofs(hgdt_operator op, int64_t* count, int32_t* host_out);

/* MatrixOperator::assemble(use_tbb) (matrix-based.hh:496-500): one device walk over elements and
 * intersections, deterministic fixed-order scatter, Dirichlet constraints, right-hand side. Async on `stream`. */
hgdt_status hgdt_operator_assemble(hgdt_operator op, void* stream);
/* device pointers of the assembled CSR values [nnz] and right-hand side [rows] (owned by the operator);
 * the values are refused (HGDT_ERR_OPERATOR) after an assembly in matrix-free mode */
hgdt_status hgdt_operator_values_device(hgdt_operator op, double** values);
hgdt_status hgdt_operator_rhs_device(hgdt_operator op, double** rhs);
/* synchronous copies to host memory */
hgdt_status hgdt_operator_copy_values(hgdt_operator op, double* host_out);
hgdt_status hgdt_operator_copy_rhs(hgdt_operator op, double* host_out);

/* How the last assemble() ran and what the solver holds (no reference counterpart; for tests and reports):
 * *structured = 1 when the structured assembly wrote the solver's storage directly (P1 on the Kuhn tetrahedra of
 * axis-parallel lattices with Laplace element forms and constant or per-element sources: one pass over the
 * DoFs, the CSR values computed from that storage only when they are asked for), else 0;
 * *solver_upper_diagonals = strict upper diagonals in the solver's scaled copy of a symmetric DIA matrix (the
 * diagonals whose entries are all exactly zero are left out: they add only 0 * p), -1 for other storages or
 * before the first apply / apply_inverse after an assembly.  Either pointer may be NULL. */
hgdt_status hgdt_operator_assembly_info(hgdt_operator op, int32_t* structured, int32_t* solver_upper_diagonals);

/* Algorithmic HBM bytes of the last assemble() (no reference counterpart; a measurement query like the one
 * above): for the structured assembly its one pass (the per-element coefficient fields and sources and the
 * Dirichlet mask read; 8 diagonals, the right-hand side, D^{-1/2} and D^{1/2} written: 88 B per DoF);
 * otherwise the compulsory bytes, 8 per element and per-element input, 8 per nonzero (none in matrix-free
 * mode) and 8 per DoF (right-hand side).  0 before the first assembly. */
hgdt_status hgdt_operator_assembly_bytes(hgdt_operator op, int64_t* bytes);

/* choose the storage used by apply / apply_inverse (default HGDT_FORMAT_AUTO: element blocks for a
 * discontinuous space with 4 DoFs per element, else DIA when every instance has at most 32 distinct
 * diagonals, else SELL-64).  DIA reads only the diagonals with
 * offset >= 0 when the assembled matrix is exactly (bitwise) symmetric;
 * HGDT_FORMAT_DIA_NONSYMMETRIC always reads every diagonal.  The format is rebuilt from the CSR
 * values at the next apply / apply_inverse after each assemble.
 *
 * HGDT_FORMAT_MATRIX_FREE: apply and apply_inverse recompute the operator from the coefficient fields
 * at every application, element by element and intersection by intersection, as
 * LocalizableOperator::apply does (operators/localizable-operator.hh:352-379 with the local operator
 * applicators of local/assembler/operator-applicators.hh:82-87,219-230), and no matrix is stored: assemble()
 * then builds the right-hand side and the Dirichlet DoF set only, and the CSR values are not available
 * until the operator is assembled again in another format.  Provided for order-1 spaces on structured
 * grids of quadrilaterals or hexahedra (and unions of them): continuous Lagrange with element forms and
 * Dirichlet constraints, discontinuous Lagrange on quadrilaterals with element forms and the SWIPDG
 * intersection forms.  Results agree with the assembled operator's to round-off (the sums are
 * associated differently); set_format succeeds for any operator, and the first apply / apply_inverse
 * fails with HGDT_ERR_NOT_SUPPORTED, naming the reason, for one outside this set.
 *
 * HGDT_FORMAT_BLOCK: the 4 x 4 blocks of a discontinuous space with 4 DoFs per element (Q1 on
 * quadrilaterals), one record per element (its own block, then the blocks of its neighbours in ascending
 * order).  When the assembled matrix is exactly symmetric, an application reads only the own block's
 * upper triangle and the blocks of the neighbours with larger index, and takes the others from the
 * neighbours' records (transposed).  The default for such spaces (HGDT_FORMAT_AUTO); every row sums its
 * terms in ascending column order, as DIA and SELL do, so apply gives the same bits in every format. */
hgdt_status hgdt_operator_set_format(hgdt_operator op, int format);
/* format actually in use after the last preparation (HGDT_FORMAT_SELL, HGDT_FORMAT_DIA for symmetric
 * DIA, HGDT_FORMAT_DIA_NONSYMMETRIC, HGDT_FORMAT_MATRIX_FREE, HGDT_FORMAT_BLOCK) and the algorithmic
 * bytes one application reads and writes (num_diagonals: neighbour blocks per element for
 * HGDT_FORMAT_BLOCK) */
hgdt_status hgdt_operator_format_info(hgdt_operator op, int32_t* format, int64_t* spmv_bytes, int32_t* num_diagonals);

/* ConstMatrixOperator::apply (matrix-based.hh:121-129): y = A x, device vectors of length rows */
hgdt_status hgdt_operator_apply(hgdt_operator op, const double* x, double* y, void* stream);

typedef struct hgdt_solver_options {
  double rtol;            /* stop when ||r_k||_2 <= rtol * ||b||_2 (per instance) */
  int64_t max_iter;       /* per instance */
  int32_t check_every;    /* iterations between host-side convergence checks (>= 1) */
  int32_t use_graph;      /* 1: replay the iteration block as a captured hipGraph */
  int32_t time_spmv;      /* k > 0: record HIP events around the operator application of every k-th
                             iteration */
  int32_t zero_initial_guess; /* 1: start from x = 0, else from the given x */
  int32_t true_residual;  /* 1 (default): compute |b - A x|_2 / |b|_2 of every instance at exit with the
                             unscaled operator (one more application; reported in true_relres) */
  int32_t restarts;       /* no reference counterpart; 0 (default): the reference's stopping rule on the
                             recursive residual only.  k > 0 (with true_residual): while the true residual
                             of an instance misses rtol, restart CG from the returned x (r = b - A x), at
                             most k times; iterations add up */
} hgdt_solver_options;

typedef struct hgdt_solver_stats {
  int32_t converged;      /* all instances converged */
  int32_t breakdown;      /* non-positive curvature or non-positive diagonal met */
  int64_t iterations;     /* max over instances */
  double relres;          /* max over instances of ||r_k|| / ||b|| (recursive residual) */
  double spmv_avg_ms;     /* mean duration of the timed operator applications (time_spmv > 0) */
  int64_t spmv_count;     /* applications timed */
  double solve_ms;        /* wall time of the solve measured with HIP events */
  int64_t step_bytes;     /* algorithmic bytes of the solver's operator application (the kernel timed
                             by time_spmv) */
  int64_t compactions;    /* batches: times the launches were narrowed to the blocks of the instances
                             still iterating */
  double true_relres;     /* max over instances of |b - A x| / |b| at exit, recomputed from the returned x
                             (-1 when options.true_residual is 0) */
  int32_t column_staged;  /* 1: the solver's operator application ran with column staging (see
                             hgdt_operator_set_column_staging) */
  int32_t restarts;       /* restarts from the true residual taken (options.restarts) */
  int64_t iteration_bytes; /* algorithmic bytes of one PCG iteration over all instances: the application
                              (step_bytes), the direction update (36 B per DoF on average over the two
                              parities of the deferred iterate update, none when the application does it)
                              and the update (32 B per DoF) */
} hgdt_solver_stats;

hgdt_status hgdt_solver_options_default(hgdt_solver_options* options);
/* ConstMatrixOperator::apply_inverse (matrix-based.hh:148-159) with a Jacobi-preconditioned CG
 * (x0 = 0 by default), run as CG on D^{-1/2} A D^{-1/2}; b and x are device vectors of length rows.  For a union grid every instance
 * is solved with its own scalars and its own stopping test in the same launches.  Blocks the host
 * until convergence is known.  Fails with HGDT_ERR_SOLVER_FAILED when an instance does not converge. */
hgdt_status hgdt_operator_apply_inverse(hgdt_operator op, const double* b, double* x,
                                        const hgdt_solver_options* options, hgdt_solver_stats* stats,
                                        void* stream);
/* per-instance iteration counts and relative residuals of the last apply_inverse, host arrays */
hgdt_status hgdt_operator_instance_stats(hgdt_operator op, int64_t* iterations, double* relres);
/* per-instance true relative residuals |b - A x| / |b| of the last apply_inverse (host array [num_instances];
 * -1 when it was not computed).  The reference's apply_inverse checks the solver's result the same way
 * (operators/matrix-based.hh:148-159 with XT::LA::Solver's post-check of the residual). */
hgdt_status hgdt_operator_instance_true_residuals(hgdt_operator op, double* relres);

/* Adjoint element kernel (no reference counterpart): for the Laplace form appended at position
 * `form_index` among the element forms, with per-element diffusion kappa_e,
 *   grad_e = - kappa_e * lambda_e^T Khat_e u_e
 * where Khat_e is the unit-coefficient local Laplace matrix recomputed element-locally.  This is
 * dJ/d(log kappa_e) for J(u(kappa)) when lambda solves A lambda = dJ/du.  u, lambda: device vectors
 * [rows]; grad: device [num_elements]. */
hgdt_status hgdt_operator_adjoint_log_diffusion_gradient(hgdt_operator op, int32_t form_index, const double* u,
                                                         const double* lambda, double* grad, void* stream);

/* launch configuration of the tunable kernels (used by the autotuner); kernel ids: 0 element assembly,
 * 1 intersection assembly, 2 scatter/gather, 3 operator application, 4 vector updates.
 * block_size multiple of 64 in [64, 1024] (assembly kernels: at most 256); items_per_thread = rows per
 * lane, a power of two up to 8, or up to 16 for kernel 3, which loads 1 or 2 rows of a lane together and
 * processes larger counts in chunks (fewer, larger blocks: fewer partial sums for the solver). */
hgdt_status hgdt_operator_set_launch(hgdt_operator op, int32_t kernel, int32_t block_size, int32_t items_per_thread);
hgdt_status hgdt_operator_get_launch(hgdt_operator op, int32_t kernel, int32_t* block_size, int32_t* items_per_thread);

/* Column staging of the solver's operator application (no reference counterpart; a launch setting like
 * the one above).  With symmetric DIA storage, one row per lane (kernel 3 with items_per_thread 1), at most
 * 8 strict upper diagonals and the stencil's column offsets in at most 4 windows no wider than the block,
 * the entries of p a row block needs are loaded into LDS once per block and read from there.  mode -1
 * (default): staged when the application moves more than 256 MiB (beyond the GPU's last-level cache);
 * 0: never; 1: wherever it applies.  Results are the same bits either way. */
hgdt_status hgdt_operator_set_column_staging(hgdt_operator op, int32_t mode);

#ifdef __cplusplus
}
#endif
#endif /* HGDT_H */
