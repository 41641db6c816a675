/*
 * oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a CPU restatement of dune-gdt's
 * assembly + linear-solve path and is used exclusively as the checker by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.
 * The product library (libhgdt.so) never links, loads or calls it.
 *
 * Every routine cites the reference file:line it follows.  Reference paths are
 * relative to the dune-gdt source tree.
 */
#ifndef HGDT_ORACLE_H
#define HGDT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* geometry types (match hgdt.h) */
enum { OR_QUAD = 0, OR_TRI = 1, OR_HEX = 2, OR_TET = 3, OR_MIXED = 4 };
/* sparsity stencils, dune/gdt/type_traits.hh:55 */
enum { OR_STENCIL_ELEMENT = 0, OR_STENCIL_INTERSECTION = 1, OR_STENCIL_ELEMENT_AND_INTERSECTION = 2 };

typedef struct OrMesh OrMesh;
typedef struct OrSpace OrSpace;
typedef struct OrOperator OrOperator;

const char* or_last_error(void);

/* ---- grids ---- */
OrMesh* or_mesh_structured(int geom, const int32_t* cells, const double* lo, const double* hi);
OrMesh* or_mesh_create(int geom, int64_t nv, const double* coords, int64_t ne, const int32_t* conn);
/* triangles and quadrilaterals: element e of type geometry[e] has vertices[offsets[e] .. offsets[e + 1]) */
OrMesh* or_mesh_mixed(int64_t nv, const double* coords, int64_t ne, const int32_t* geometry, const int64_t* offsets,
                      const int32_t* vertices);
int or_mesh_is_mixed(const OrMesh* m);
void or_mesh_destroy(OrMesh* m);
int64_t or_mesh_num_vertices(const OrMesh* m);
int64_t or_mesh_num_elements(const OrMesh* m);
int or_mesh_dim(const OrMesh* m);
int or_mesh_verts_per_element(const OrMesh* m);
void or_mesh_coords(const OrMesh* m, double* out);
void or_mesh_conn(const OrMesh* m, int32_t* out);
/* intersections: for element e, local face f: neighbour element (or -1) */
int or_mesh_faces_per_element(const OrMesh* m);
void or_mesh_neighbors(const OrMesh* m, int32_t* out /* ne * faces */);

/* ---- spaces ---- */
OrSpace* or_space_create(OrMesh* m, int continuous, int order);
void or_space_destroy(OrSpace* s);
int64_t or_space_size(const OrSpace* s);
int or_space_local_size(const OrSpace* s);
void or_space_global_indices(const OrSpace* s, int32_t* out /* ne * local_size */);

/* ---- matrix operator (pattern + forms + functional + Dirichlet) ---- */
OrOperator* or_operator_create(OrSpace* s, int stencil);
OrOperator* or_operator_create_between(OrSpace* source, OrSpace* range, int stencil);
int64_t or_operator_cols(const OrOperator* op);
void or_operator_destroy(OrOperator* op);
int64_t or_operator_nnz(const OrOperator* op);
void or_operator_pattern(const OrOperator* op, int64_t* row_ptr, int32_t* cols);

/* coefficient arguments: per-element array (ne values) or, if NULL, the constant */
int or_operator_append_laplace(OrOperator* op, const double* kappa, double kappa_const,
                               int over_integrate, double scaling);
int or_operator_append_product(OrOperator* op, const double* weight, double weight_const,
                               int over_integrate, double scaling);
/* the same forms with a coefficient varying inside the elements: `count` values at the points of the element
 * rules of order coeff_order + 2 k + over_integrate (k the space order), element by element */
/* the Laplace form with a matrix-valued diffusion: d x d values (row-major) per element (per_element 1) or for
 * all elements (0) */
int or_operator_append_laplace_tensor(OrOperator* op, const double* tensor, int per_element, int over_integrate,
                                      double scaling);
int or_operator_append_laplace_points(OrOperator* op, const double* values, int64_t count, int coeff_order,
                                      int over_integrate, double scaling);
int or_operator_append_product_points(OrOperator* op, const double* values, int64_t count, int coeff_order,
                                      int over_integrate, double scaling);
int or_operator_append_inner_coupling(OrOperator* op, double symmetry_prefactor, const double* kappa,
                                      double kappa_const, const double* omega, double omega_const,
                                      int over_integrate, double scaling);
int or_operator_append_inner_penalty(OrOperator* op, double penalty, const double* omega,
                                     double omega_const, int over_integrate, double scaling);
int or_operator_append_dirichlet_coupling(OrOperator* op, double symmetry_prefactor, const double* kappa,
                                          double kappa_const, int over_integrate, double scaling);
int or_operator_append_boundary_penalty(OrOperator* op, double penalty, const double* omega,
                                        double omega_const, int over_integrate, double scaling);
/* source kinds: 0 constant / per-element (f, f_const), 1 = ESV2007 force (pi^2/2) cos(pi x/2) cos(pi y/2)
 * declared with polynomial order f_order */
int or_operator_append_source(OrOperator* op, int kind, const double* f, double f_const, int f_order,
                              int over_integrate);
/* constrain every boundary DoF (AllDirichletBoundaryInfo) -- only meaningful for CG spaces */
int or_operator_set_dirichlet(OrOperator* op, int enabled);
/* DirichletConstraints::apply(matrix, vector, only_clear, ensure_symmetry), tools/dirichlet-constraints.hh:152-184 */
/* matrix-valued diffusion / weight (d x d per element, NULL: scalar) of the last appended inner (boundary 0) or
 * boundary (1) intersection form (laplace-ipdg.hh:47-61, ipdg.hh:59-72) */
/* matrix-valued coefficient (d x d per element) of the last appended boundary functional */
int or_operator_set_boundary_source_tensor(OrOperator* op, const double* tensor);
int or_operator_set_face_tensors(OrOperator* op, int boundary, const double* kappa_t, const double* omega_t);
int or_operator_set_dirichlet_mode(OrOperator* op, int only_clear, int ensure_symmetry);

int or_operator_assemble(OrOperator* op, int nthreads);
void or_operator_values(const OrOperator* op, double* out);
void or_operator_rhs(const OrOperator* op, double* out);
int64_t or_operator_num_dirichlet(const OrOperator* op);
void or_operator_dirichlet_dofs(const OrOperator* op, int32_t* out);

/* ---- algebra on CSR arrays ---- */
void or_spmv(int64_t n, const int64_t* row_ptr, const int32_t* cols, const double* vals, const double* x,
             double* y, int nthreads);
/* Jacobi-PCG from x (in/out), stop when ||r||_2 <= rtol ||b||_2; returns 0 if converged */
int or_pcg_jacobi(int64_t n, const int64_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                  double* x, double rtol, int64_t max_iter, int nthreads, int64_t max_iters_run,
                  int64_t* iters, double* relres);

/* ---- pointwise integrands (reference unit tests) ---- */
/* Laplace: result[i][j] += sum_r (W g^a_j[r]) . g^t_i[r]; W is d x d; grads are (n, r, d) row-major */
void or_laplace_evaluate(int d, int r, int rows, int cols, const double* weight, const double* test_grads,
                         const double* ansatz_grads, double* result);
/* product: result[i][j] = (W t_i) . a_j; W is r x r; values are (n, r) */
void or_product_evaluate(int r, int rows, int cols, const double* weight, const double* test_values,
                         const double* ansatz_values, double* result);
/* quadrature rule (points in reference element, weights) of at least the given order */
int or_quadrature(int geom, int order, double* points /* n*dim */, double* weights, int capacity);

/* ---- post-processing ---- */
/* broken H1-semi norm of (u_exact - u_h) for the ESV2007 exact solution cos(pi x/2) cos(pi y/2) */
double or_esv2007_h1_semi_error(const OrSpace* s, const double* u, int quad_order);
/* adjoint element kernel: g_e = - kappa_e * lambda_e^T Khat_e u_e, Khat_e the unit-coefficient
 * local Laplace matrix of element e (no reference counterpart) */
int or_adjoint_log_kappa_gradient(const OrSpace* s, const double* kappa, const double* u, const double* lambda,
                                  double* grad);

/* squared error norms of e = u_h - g (g: values [ne][npts], physical gradients [ne][npts][dim] at the
 * points of the element rule of `order`, NULL for 0): out[0] = sum_e int e^2, out[1] = sum_e int kappa |grad e|^2
 * (operators/bilinear-form.hh:344-452) */
int or_error_norms(const OrSpace* s, const double* u, int order, const double* gval, const double* ggrad,
                   const double* kappa, double kappa_const, double* out);
/* physical coordinates of the element-rule points, out[ne][npts][dim] (NULL: size query); returns npts */
int or_mesh_quadrature_points(const OrMesh* m, int order, double* out);
int or_mesh_quadrature_offsets(const OrMesh* m, int order, int64_t* out /* ne + 1 */);

/* right-hand side on the boundary intersections: VectorFunctional::append(LocalIntersectionIntegralFunctional(...))
 * kind 0: LocalLaplaceIPDGIntegrands::DirichletCoupling(param, coeff, g) as a unary integrand
 * kind 1: LocalIPDGIntegrands::BoundaryPenalty(param, coeff).with_ansatz(g)
 * kind 2: LocalIntersectionProductIntegrand(coeff).with_ansatz(g)
 * g: values at the face quadrature points of the term's order, [boundary face (walk order)][point], or g_const */
int or_operator_append_boundary_source(OrOperator* op, int kind, double param, const double* coeff, double coeff_const,
                                      const double* g, double g_const, int g_order, int over_integrate);
/* physical coordinates of the line-rule points on the boundary intersections in walk order, out[nb][npts][2]
 * (NULL: size query); returns npts */
int or_mesh_boundary_quadrature_points(const OrMesh* m, int order, double* out);
/* boundary intersections of the boundary forms / functionals appended next, and the Dirichlet intersections of
 * the constraints: one byte per boundary intersection in walk order (element, then face); NULL: all */
int or_operator_set_boundary_filter(OrOperator* op, const uint8_t* mask, int64_t count);
int or_operator_set_dirichlet_faces(OrOperator* op, const uint8_t* mask, int64_t count);

#ifdef __cplusplus
}
#endif
#endif
