// hgdt_internal.hpp -- host-side data structures of libhgdt.so.
//
// Host C++ owns the set-up (grids, DoF maps, sparsity pattern, fixed-order
// scatter maps); HIP kernels own every floating-point pass over the data
// (assembly, constraint application, operator application, solves, adjoint).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hgdt.h"

namespace hgdt {

// Error carrying an hgdt_status; the C ABI converts it into a return code.
struct Error : std::runtime_error {
  hgdt_status status;
  Error(hgdt_status s, const std::string& msg)
    : std::runtime_error(msg)
    , status(s)
  {}
};

#define HGDT_THROW(status, msg) throw ::hgdt::Error((status), (msg))

#define HGDT_HIP(call)                                                                                         \
  do {                                                                                                         \
    hipError_t hgdt_e_ = (call);                                                                               \
    if (hgdt_e_ != hipSuccess)                                                                                 \
      throw ::hgdt::Error(hgdt_e_ == hipErrorOutOfMemory ? HGDT_ERR_OUT_OF_MEMORY : HGDT_ERR_DEVICE,          \
                          std::string(#call) + ": " + hipGetErrorString(hgdt_e_));                            \
  } while (0)

// XCD-aware index of 1-d launches.  The hardware deals workgroups round-robin to the 8 XCDs of the
// chip, each with its own L2; hardware block hw runs logical block (hw % 8) * ceil(nb / 8) + hw / 8, so
// every XCD works on a contiguous eighth of the index range and neighbouring blocks share one L2.  The
// grid must have xcd_grid(n, block) blocks; lanes whose index is >= n are idle.
__device__ __forceinline__ int64_t xcd_index(int64_t n)
{
  const int64_t nb = (n + blockDim.x - 1) / blockDim.x;
  const int64_t q = (nb + 7) >> 3;
  const int64_t hw = blockIdx.x;
  const int64_t logical = (hw & 7) * q + (hw >> 3);
  return logical * blockDim.x + threadIdx.x;
}
inline unsigned xcd_grid(int64_t n, int block)
{
  const int64_t nb = (n + block - 1) / block;
  return unsigned(((nb + 7) / 8) * 8);
}

// device buffer owned by the library
template <class T>
struct DeviceArray {
  T* ptr = nullptr;
  size_t count = 0;
  bool owned = true; // false: memory of the caller (borrow), never freed here
  DeviceArray() = default;
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  ~DeviceArray()
  {
    release();
  }
  void release()
  {
    if (ptr && owned)
      (void)hipFree(ptr);
    ptr = nullptr, count = 0, owned = true;
  }
  // use the caller's device memory of n elements (it must outlive this array's use)
  void borrow(T* p, size_t n)
  {
    release();
    ptr = p, count = n, owned = false;
  }
  void allocate(size_t n)
  {
    if (n == count && ptr)
      return;
    if (!owned)
      HGDT_THROW(HGDT_ERR_OPERATOR, "borrowed device memory cannot be reallocated");
    release();
    if (n == 0)
      return;
    HGDT_HIP(hipMalloc(&ptr, n * sizeof(T)));
    count = n;
  }
  void upload(const std::vector<T>& host)
  {
    allocate(host.size());
    if (!host.empty())
      HGDT_HIP(hipMemcpy(ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  void download(T* host) const
  {
    if (count)
      HGDT_HIP(hipMemcpy(host, ptr, count * sizeof(T), hipMemcpyDeviceToHost));
  }
};

struct RefElement {
  int geom = 0, dim = 0, nverts = 0, nfaces = 0;
  bool simplex = false;
  double verts[8][3] = {};
  int face_nverts = 0;
  int faces[6][4] = {};
  double volume = 1.; // reference element measure
};
RefElement reference_element(int geom);
// Lagrange nodes of order 1 or 2 on a reference element, in the local DoF order of the shape
// functions: for every node the reference vertices spanning its subentity.  Cubes: tensor index
// t = i0 + (k+1) i1 [+ (k+1)^2 i2], node at i/k; simplices: the vertices, then (order 2) the edge
// midpoints (a, b), a < b, in lexicographic order.
std::vector<std::vector<int>> lagrange_node_vertices(const RefElement& ref, int order);

// Quadrature rule on a reference element or on [0,1]
struct Quadrature {
  int dim = 0;
  std::vector<double> x; // npts * dim
  std::vector<double> w;
  int size() const
  {
    return int(w.size());
  }
};
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w);
Quadrature line_rule(int order);
Quadrature element_rule(const RefElement& ref, int order);
Quadrature face_rule(const RefElement& ref, int order); // rule of the faces of ref (dim - 1 coordinates)

// axis-parallel lattice of one instance of a structured grid: n[a] cells of width h[a] from lo[a]
// (cells and vertices numbered lexicographically, x fastest)
struct Lattice {
  int64_t n[3] = {1, 1, 1};
  double lo[3] = {0., 0., 0.}, h[3] = {0., 0., 0.};
};

// the elements of one geometry type of a mixed grid: their global indices (ascending) and their
// connectivity in that type's reference vertex order
struct GridPart {
  RefElement ref;
  std::vector<int32_t> eid;  // global element index of every element of the part
  std::vector<int32_t> conn; // eid.size() * ref.nverts
  DeviceArray<int32_t> d_eid, d_conn;
  int64_t ne() const
  {
    return int64_t(eid.size());
  }
};

struct Grid {
  // homogeneous grids: the element type.  Mixed grids (triangles and quadrilaterals): geom HGDT_MIXED,
  // nverts / nfaces the largest of the types (the strides of conn, nb and nbface, padded with -1 vertices
  // and -2 neighbours), no faces or vertex coordinates: every element's own type is eref(e)
  RefElement ref;
  bool mixed = false;
  std::vector<int8_t> etype;    // mixed: part of every element
  std::vector<int32_t> elocal;  // mixed: index of every element in its part
  std::vector<GridPart> parts;  // mixed: one per geometry type present, ascending geometry id
  const RefElement& eref(int64_t e) const
  {
    return mixed ? parts[etype[e]].ref : ref;
  }
  bool structured = false;
  std::vector<Lattice> lattice; // one per instance when every instance is a structured grid, else empty
  int64_t nv = 0, ne = 0;
  std::vector<double> coords;  // nv * dim
  std::vector<int32_t> conn;   // ne * nverts
  std::vector<int32_t> nb;     // ne * nfaces: neighbour element or -1 (boundary)
  std::vector<int8_t> nbface;  // ne * nfaces: local face index in the neighbour
  std::vector<int64_t> inst_elem{0}, inst_vert{0};
  int64_t n_inner = 0, n_boundary = 0;
  DeviceArray<double> d_coords;
  DeviceArray<int32_t> d_conn;
  DeviceArray<int8_t> d_etype; // mixed: etype on the device (the face kernels select their type pair by it)

  int ninst() const
  {
    return int(inst_elem.size()) - 1;
  }
  void finalize(bool neighbours_known = false); // neighbours (face matching), face counts, device upload
  void count_faces_and_upload();
  void check_affine() const;
};

// where the quadrature points of every element (rule of the given order of its own type) start in the
// element-ordered layout of quadrature-point data: ne + 1 offsets
std::vector<int64_t> element_point_offsets(const Grid& g, int order);

// the DoFs of the elements of one part of a mixed grid, in the part's element order
struct SpacePart {
  int local_size = 0;
  std::vector<int32_t> dofs; // part ne * local_size
  DeviceArray<int32_t> d_dofs;
};

struct Space {
  std::shared_ptr<Grid> grid;
  int kind = HGDT_CONTINUOUS_LAGRANGE, order = 1;
  int local_size = 0;               // mixed grids: the largest of the parts'
  int64_t size = 0;
  std::vector<int32_t> dofs;        // ne * local_size (mixed grids: -1 past an element's own local size)
  std::vector<int64_t> inst_dof{0}; // DoF offsets of the instances
  DeviceArray<int32_t> d_dofs;
  std::vector<SpacePart> parts;     // mixed grids: one per grid part
  int elem_size(int64_t e) const    // local size of element e
  {
    return parts.empty() ? local_size : parts[grid->etype[e]].local_size;
  }
};

enum ElementFormKind { EF_LAPLACE = 0, EF_PRODUCT = 1 };
enum FaceFormKind { FF_INNER_COUPLING = 0, FF_INNER_PENALTY = 1, FF_DIRICHLET_COUPLING = 2, FF_BOUNDARY_PENALTY = 3 };

struct ElementFormDesc {
  int kind;
  const double* field;
  double value;
  int over_integrate;
  double scaling;
  // the coefficient at the points of the element rule of order coeff_order + 2 k + over_integrate (k the
  // space order), element-ordered (NULL: field or value)
  const double* points = nullptr;
  int coeff_order = 0;
  // Laplace forms with a matrix-valued diffusion (GridFunction<E, d, d>): 1 d x d per element (tensor, device,
  // row-major), 2 one d x d for all (tensor_value)
  int tensor_kind = 0;
  const double* tensor = nullptr;
  double tensor_value[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
};
struct FaceFormDesc {
  int kind;
  double param;
  const double* kappa;
  double kappa_value;
  const double* omega;
  double omega_value;
  int over_integrate;
  double scaling;
  // boundary forms: the boundary intersections (walk order) the form is applied on, one byte each (device,
  // owned by the operator); NULL: every boundary intersection
  const uint8_t* mask;
  // matrix-valued diffusion / weight (GridFunction<E, d, d>): d x d per element (device, row-major, symmetric);
  // NULL: the scalars above
  const double* kappa_t = nullptr;
  const double* omega_t = nullptr;
};
struct SourceDesc {
  int kind;
  const double* data;
  double value;
  int f_order;
  int over_integrate;
};

// right-hand side term on the boundary intersections (LocalIntersectionIntegralFunctional)
struct BoundarySourceDesc {
  int kind;            // HGDT_BOUNDARY_*
  double param;        // symmetry prefactor or penalty
  const double* coeff; // diffusion (coupling) or weight, per element; NULL -> coeff_value
  double coeff_value;
  const double* data;  // g at the face quadrature points [boundary face][point]; NULL -> value
  double value;
  int g_order;
  int over_integrate;
  const uint8_t* mask; // as FaceFormDesc::mask
  // matrix-valued diffusion (coupling) or weight (penalty): d x d per element (device, symmetric); NULL: coeff
  const double* coeff_t = nullptr;
};

// launch configuration of one tunable kernel
struct Launch {
  int block = 256;
  int items = 1;
  bool user = false; // set explicitly (autotuner / caller) rather than by the default heuristic
};

// Fixed-order scatter maps: every matrix slot (and right-hand side entry) lists
// its contributions, as indices into the staging buffer of local matrices, in the
// order the reference's grid walk would add them.
struct ScatterMaps {
  int64_t elem_base = 0, inner_base = 0, bnd_base = 0, rhs_base = 0, brhs_base = 0, staging_size = 0;
  // mixed grids: the element matrices and right-hand sides of every part start here (entry-major per part)
  std::vector<int64_t> part_elem_base, part_rhs_base;
  std::vector<int32_t> inner_in, inner_fin, inner_out, inner_fout; // inner intersections visited once
  std::vector<int32_t> bnd_e, bnd_f;                               // boundary intersections
  DeviceArray<uint32_t> d_slot_ptr, d_contrib;
  DeviceArray<uint32_t> d_rhs_ptr, d_rhs_contrib;
  DeviceArray<int32_t> d_inner, d_bnd; // packed (e_in, f_in, e_out, f_out) / (e, f)
  bool built = false;
  int nforms_built = -1; // which form set the maps were built for (element/face/source presence)
  int signature = -1;
};

// consecutive rows handed to the blocks of a launch; a block owns block * items rows
struct RowPartition {
  int block = 256, items = 1;
  int64_t nblocks = 0;
  std::vector<int64_t> row0;        // nblocks + 1 (row start of every block, last = n)
  std::vector<int32_t> inst;        // nblocks
  std::vector<int32_t> inst_block0; // ninst + 1
  DeviceArray<int64_t> d_row0;
  DeviceArray<int32_t> d_inst, d_inst_block0;
  void build(const std::vector<int64_t>& inst_dof, int block_size, int items_per_thread);
  // blocks given per instance (tiles of the matrix-free kernel): row0 holds the instance's first row
  void build_counts(const std::vector<int64_t>& inst_dof, const std::vector<int64_t>& blocks_per_inst,
                    int block_size, int items_per_thread);
};

// storage of the matrix used by apply / apply_inverse
struct SolverFormat {
  int format = 0;      // HGDT_FORMAT_SELL, HGDT_FORMAT_DIA, HGDT_FORMAT_BLOCK or HGDT_FORMAT_MATRIX_FREE
  bool partitions_ready = false; // row partitions A and V built for the current launch shapes
  bool structure_ready = false;  // storage layout (DIA offsets / SELL slices) built
  bool valid = false;  // values converted since the last assemble
  bool diag_positive = false;
  bool symmetric = false; // DIA read through its upper diagonals only (exactly symmetric matrix)
  int nup = 0;            // diagonals with offset >= 0
  int64_t bytes = 0;   // algorithmic bytes of one operator application
  // row-block partitions (blocks never straddle instances): A for operator application, V for the
  // vector kernels of the solver; they are tuned separately
  RowPartition A, V;
  DeviceArray<int64_t> d_inst_row0;
  // DIA
  int ndiag = 0;
  DeviceArray<int64_t> d_dia_offsets; // ninst * ndiag
  DeviceArray<double> d_dia;          // ndiag * n
  DeviceArray<int64_t> d_dia_pos;     // nnz: CSR slot -> DIA position
  // SELL-64 (slices of 64 rows starting at every block start)
  int64_t nslices = 0;
  DeviceArray<int64_t> d_slice_ptr;  // nslices + 1
  DeviceArray<int32_t> d_slice_row0; // nslices
  DeviceArray<int32_t> d_sell_cols;
  DeviceArray<double> d_sell_vals;
  DeviceArray<int64_t> d_sell_pos; // nnz: CSR slot -> SELL position
  // 4 x 4 element blocks (discontinuous spaces with 4 DoFs per element): per element its own block and
  // the blocks of its neighbours (ascending), value planes: value m of element E at m * blk_ne + E, m < 16
  // the own block (row-major), 16 + 16 s + 4 a + b entry (a, b) of the block with neighbour slot s
  int64_t blk_ne = 0;
  int blk_mn = 0;                // neighbour slots per element
  int64_t blk_up = 0, blk_all = 0; // neighbour blocks above the element / all, summed over the elements
  DeviceArray<int32_t> d_blk_nb;  // blk_mn * blk_ne: neighbour of each slot (ascending, -1 past the last)
  DeviceArray<int32_t> d_blk_mir; // blk_mn * blk_ne: slot of the element in its neighbour's list
  DeviceArray<double> d_blk;      // (16 + 16 blk_mn) * blk_ne
  DeviceArray<int64_t> d_blk_pos; // nnz: CSR slot -> block position
  // Jacobi preconditioning by symmetric scaling: h_i = 1 / sqrt(a_ii), s_i = sqrt(a_ii), and the scaled
  // matrix D^{-1/2} A D^{-1/2} in the layout of the storage above (DIA or SELL values)
  DeviceArray<double> d_dh, d_dsq;
  DeviceArray<double> d_scaled;
  int64_t step_bytes = 0; // algorithmic bytes of the CG solver's operator application
  // matrix-free application (format HGDT_FORMAT_MATRIX_FREE): per row block of A one record (first DoF
  // and first element of its instance, the instance's cells n0 n1 n2, the instance, the block's rows
  // relative to the instance), and per instance the reference local matrices of the appended forms
  int mf_kind = -1;
  int64_t mf_ref_stride = 0;
  int mf_fields = 0; // distinct coefficient arrays read by one application
  DeviceArray<int64_t> d_mf_blk;
  DeviceArray<double> d_mf_ref;
};

// device-side PCG state, one record per instance; two copies indexed by iteration parity so that
// no kernel reads a record it writes
struct PcgScalars {
  double rz, rr, bb, tol2, alpha, beta, pq;
  double alpha_prev; // alpha of the step before the one in alpha (the iterate is updated every second step)
  int64_t iters;
  int32_t done; // 0 running, 1 converged, 2 max_iter reached, 3 breakdown (set by the direction update)
  // breakdown met by the update kernel of this record's step.  Only the next direction update reads it
  // (and then sets done = 3): the update kernel's own blocks test `done`, which it never writes, so no
  // block of that launch sees a flag another block of the same launch may still be writing.
  int32_t brk;
};
// state of a record as the next direction update sees it
__host__ __device__ inline int pcg_state(const PcgScalars& s)
{
  return s.done ? s.done : (s.brk ? 3 : 0);
}

struct SolverCache; // solver.hip: events, pinned buffers, captured graph

// one instance of a structured grid of Kuhn tetrahedra for the structured assembly: its first DoF and
// element, its cells per direction, and its unit-coefficient local quantities (the same for every
// tetrahedron of an axis-parallel lattice): wT[a] = |T| / h_a^2 (the P1 Laplace matrix of a tetrahedron with
// path v0 -> v0 + e_p0 -> v0 + e_p0 + e_p1 -> v7 has diagonal wT[p0], wT[p0] + wT[p1], wT[p1] + wT[p2], wT[p2]
// and the entries -wT[p_k] between consecutive path vertices, nothing else) and |T| / 4 (the local
// right-hand side of a unit source)
struct DirectInst {
  int64_t dof0, elem0;
  int32_t nx, ny, nz, pad;
  double wT[3];
  double quarter_vol;
};

struct Operator {
  std::shared_ptr<Space> space;  // the range space: rows (test functions), right-hand side, constraints
  std::shared_ptr<Space> source; // the source space: columns (ansatz functions); the range space by default
  int stencil = HGDT_STENCIL_ELEMENT;
  int64_t n = 0, nnz = 0;
  int64_t ncols = 0; // source space size (n for a square operator)
  bool square() const { return source == space; }
  // MatrixOperator on a caller's matrix (matrix-based.hh:279-293): its CSR pattern (host copy) and device
  // values; assemble() adds to the values held (add_to_entry)
  bool borrowed = false;
  std::vector<int64_t> borrowed_row_ptr;
  std::vector<int32_t> borrowed_cols;
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> cols;
  DeviceArray<int64_t> d_row_ptr;
  DeviceArray<int32_t> d_cols;
  DeviceArray<double> d_values, d_rhs, d_staging;
  std::vector<ElementFormDesc> eforms;
  std::vector<FaceFormDesc> cforms, bforms;
  std::vector<SourceDesc> sources;
  std::vector<BoundarySourceDesc> bsources;
  // quadrature order of a boundary term (laplace-ipdg.hh:292-297; conversion.hh:194-197 with ipdg.hh:245-250
  // and product.hh:336-341)
  int boundary_source_order(const BoundarySourceDesc& S) const
  {
    const int k = space->order;
    if (S.kind == HGDT_BOUNDARY_DIRICHLET_COUPLING)
      return S.g_order + 0 + std::max(k - 1, 0) + S.over_integrate;
    return 0 + k + S.g_order + S.over_integrate;
  }
  bool dirichlet = false;
  // DirichletConstraints::apply(matrix, vector, only_clear, ensure_symmetry) (tools/dirichlet-constraints.hh:
  // 152-184): clear instead of unit rows; the columns of the constrained DoFs too
  bool dirichlet_only_clear = false, dirichlet_ensure_symmetry = true;
  // the Dirichlet intersections of the constraints (boundary_info.type == DirichletBoundary), one byte per
  // boundary intersection in walk order; empty: every boundary intersection
  std::vector<uint8_t> dirichlet_faces;
  // boundary-intersection masks of the appended boundary forms and functionals, and the one the next appends
  // take (ApplyOn::CustomBoundaryIntersections); NULL: every boundary intersection
  std::vector<std::unique_ptr<DeviceArray<uint8_t>>> face_masks;
  const uint8_t* pending_mask = nullptr;
  bool matrix_free_assembly = false; // the last assemble() ran in matrix-free mode: no CSR values
  std::vector<int32_t> dirichlet_dofs;
  DeviceArray<uint8_t> d_slot_flag; // 0 keep, 1 zero, 2 one
  DeviceArray<uint8_t> d_dof_mask;
  bool dirichlet_built = false;
  ScatterMaps maps;
  DeviceArray<double> d_qtab_elem, d_qtab_line; // quadrature tables of the appended forms
  // mixed grids: per (part, source given at quadrature points) where every element's points start
  std::vector<std::unique_ptr<DeviceArray<int64_t>>> part_qoff;
  // the same per (part, element form with its coefficient at quadrature points)
  std::vector<std::unique_ptr<DeviceArray<int64_t>>> part_eqoff;
  bool tables_ready = false;
  int requested_format = HGDT_FORMAT_AUTO;
  bool allow_symmetric = true;
  SolverFormat fmt;
  Launch launch[5];
  // p entries of the solver's symmetric DIA application staged in LDS: -1 automatic, 0 never, 1 wherever
  // the stencil allows it
  int column_staging = -1;
  // algorithmic HBM bytes of the last assembly (hgdt_operator_assembly_bytes)
  int64_t assembly_bytes = 0;
  // solver workspace
  DeviceArray<double> d_r, d_p, d_q;
  DeviceArray<double> d_p2; // second p buffer of the matrix-free solver with the fused direction update
  DeviceArray<double> d_partials; // 2 per block
  DeviceArray<PcgScalars> d_scalars;
  std::shared_ptr<SolverCache> solver_cache;
  std::vector<int64_t> last_iters;
  std::vector<double> last_relres;
  std::vector<double> last_true_relres; // |b - A x| / |b| at exit (-1: not computed)

  // Structured assembly (P1 on the Kuhn tetrahedra of axis-parallel lattices with Laplace element forms and
  // constant or per-element sources): the symmetric DIA storage of the solver (diagonal and upper diagonals),
  // D^{-1/2}, D^{1/2} and the right-hand side are written by one pass over the DoFs, and the CSR values are
  // computed from that storage only when they are asked for.
  bool direct_storage = false; // the DIA storage holds the last assembly (no CSR conversion needed)
  bool values_stale = false;   // d_values does not hold the last assembly yet
  int direct_state = -1;       // -1 not checked since the last change of forms or format, 0 no, 1 yes
  DeviceArray<DirectInst> d_direct_inst;
  bool direct_applicable();
  void assemble_direct(hipStream_t stream);
  void materialize_values(hipStream_t stream); // CSR values from the structured assembly's storage
  int solver_upper_diagonals() const;          // strict upper diagonals of the solver's symmetric DIA copy

  void build_pattern();
  void build_scatter_maps();
  void build_dirichlet();
  void prepare_assembly();
  void assemble(hipStream_t stream);
  void prepare_solver(hipStream_t stream);
  void prepare_storage(hipStream_t stream); // storage format, Jacobi scaling, scaled values
  void apply(const double* x, double* y, hipStream_t stream);
  void apply_inverse(const double* b, double* x, const hgdt_solver_options& opt, hgdt_solver_stats& st,
                     hipStream_t stream);
  void adjoint_gradient(int form_index, const double* u, const double* lambda, double* grad, hipStream_t stream);
};

// worker-thread helper for host set-up loops
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& body);

// matrix-free application (matrix_free.hip)
// MF_PCG_DIR: the solver's direction update fused into its operator application (discontinuous kernel)
enum { MF_APPLY = 0, MF_PCG = 1, MF_DIAG = 2, MF_PCG_DIR = 3 };
struct MfVectors {
  const double* x = nullptr; // MF_APPLY: x; MF_PCG: p; MF_PCG_DIR: p of the previous step
  double* y = nullptr;       // MF_APPLY: y = A x; MF_PCG(_DIR): q = D^{-1/2} A D^{-1/2} p
  const double* dh = nullptr; // MF_PCG(_DIR): D^{-1/2}
  double* part = nullptr;     // MF_PCG(_DIR): p.q per row block
  PcgScalars* S = nullptr;    // MF_PCG_DIR writes the next record (first block of every instance)
  int par = 0;
  const int32_t* list = nullptr; // MF_PCG(_DIR): blocks of unfinished instances (nb of them; NULL: all)
  int nb = -1;
  int rev = 0; // each XCD walks its blocks top down
  // MF_PCG_DIR: new p, scaled residual, scaled iterate (updated every second step with the deferred alpha p),
  // partials of the previous update (2 per V-block), first V- and A-block of every instance, max_iter
  double* p_out = nullptr;
  const double* r = nullptr;
  double* xh = nullptr;
  const double* part2 = nullptr;
  const int32_t* v_ib0 = nullptr;
  const int32_t* a_ib0 = nullptr;
  int64_t max_iter = 0;
  double* out_dh = nullptr;  // MF_DIAG: D^{-1/2}, D^{1/2} and the non-positive-diagonal flag
  double* out_dsq = nullptr;
  int* flags = nullptr;
};
// host set-up of the matrix-free data; throws HGDT_ERR_NOT_SUPPORTED (with the reason) for operators
// it does not cover
void mf_setup(Operator& op);
// per-block records of the operator partition A (after every rebuild of the partition)
void mf_block_table(Operator& op);
// tiles per instance of the discontinuous matrix-free kernel (empty when the grid has no lattices)
std::vector<int64_t> mf_tile_counts(const Operator& op);
// algorithmic bytes of one application in mode MF_APPLY, MF_PCG or MF_PCG_DIR
int64_t mf_bytes(const Operator& op, int mode);
// the solver fuses its direction update into the application (MF_PCG_DIR) for this operator
bool mf_fused_direction(const Operator& op);
void mf_launch(const Operator& op, int mode, const MfVectors& v, hipStream_t stream, hipEvent_t ev0 = nullptr,
               hipEvent_t ev1 = nullptr);

// kernel launchers (assembly.hip)
void launch_assembly(Operator& op, hipStream_t stream, bool matrix);
// squared L2 and (kappa-weighted) H1-semi norms of u_h - g per instance (host_out[2 * ninst])
// every element's d x d block of a device tensor field is symmetric (exactly)
bool tensor_field_symmetric(const double* tensor, int64_t ne, int d);

void error_norms(const Space& s, const double* dofs, int order, const double* gval, const double* ggrad,
                 const double* kappa, double kappa_value, double* host_out, hipStream_t stream);
// the same results (bit for bit) through a fixed device array instead of allocated buffers
void error_norms_noalloc(const Space& s, const double* dofs, int order, const double* gval, const double* ggrad,
                         const double* kappa, double kappa_value, double* host_out, hipStream_t stream);
int64_t error_norms_scratch_bytes();

} // namespace hgdt
