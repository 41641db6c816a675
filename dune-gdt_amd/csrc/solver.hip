// solver.hip -- operator application and Jacobi-preconditioned CG on gfx950 (FP64).
//
// Replaces ConstMatrixOperator::apply (matrix_.mv, dune/gdt/operators/matrix-based.hh:121-129)
// and apply_inverse (XT::LA::Solver, matrix-based.hh:148-159).
//
// Storage: the assembled CSR matrix is converted once per assembly into
//   * DIA  -- one value array per diagonal, no column indices, when every problem
//             instance has <= 32 distinct (col - row) offsets (structured grids), or
//   * SELL-64 -- sliced ELLPACK with slices of 64 rows (one wavefront), column-major
//             inside a slice so every load of a wave is one contiguous 512-byte segment.
// Both are HBM-bandwidth-bound; one lane owns a row, rows are summed in column order.
//
// PCG iteration = three launches, each one pass over its vectors:
//   k_pcg_spmv   : q = A p and the block partials of p.q
//   k_pcg_update : x += alpha p, r -= alpha q, z = D^{-1} r (not stored), partials of r.z and r.r
//   k_pcg_pupdate: p = D^{-1} r + beta p
// Dot products are reduced deterministically: butterfly inside a wavefront, waves in
// fixed order inside a block, then the last block of the instance (found with an
// integer ticket; the only atomic in the library, never on values) sums the block
// partials in block order and updates the instance's scalars on the device.  Every
// block belongs to exactly one problem instance, so a batch of independent problems
// (a union grid) is solved in the same launches with per-instance scalars and stopping
// tests; blocks of converged instances exit immediately.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <mutex>

#include "hgdt_internal.hpp"

namespace hgdt {
namespace {

constexpr int MAXDIAG = 32;

struct PcgParams {
  int64_t n;
  int items;
  // partition
  const int64_t* block_row0;
  const int32_t* block_inst;
  const int32_t* inst_block0;
  const int64_t* inst_row0;
  // DIA
  int ndiag;
  const int64_t* dia_off;
  const double* dia;
  // SELL
  const int64_t* block_slice0;
  const int64_t* slice_ptr;
  const int32_t* sell_cols;
  const double* sell_vals;
  // vectors
  const double* b;
  const double* ax;
  double* x;
  double* r;
  double* p;
  double* q;
  const double* dinv;
  double* partials;
  PcgScalars* S;
  double rtol2;
  int64_t max_iter;
  int zero_guess;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

// deterministic block reduction of NV values; result valid in thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
  __shared__ double lds[NV][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k)
    v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      lds[k][wave] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.;
      for (int w = 0; w < nw; ++w)
        s += lds[k][w];
      v[k] = s;
    }
  __syncthreads();
}

// store this block's partials; returns true in every thread of the instance's last block,
// which then sees all partials of the instance
template <int NV>
__device__ __forceinline__ bool publish_partials(const PcgParams& P, int b, int inst, const double (&v)[NV])
{
  __shared__ int last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      __hip_atomic_store(&P.partials[int64_t(b) * NV + k], v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const int nb = P.inst_block0[inst + 1] - P.inst_block0[inst];
    const int t = __hip_atomic_fetch_add(&P.S[inst].ticket, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == nb - 1);
  }
  __syncthreads();
  if (last)
    __threadfence();
  return last;
}

// the last block sums the instance's partials in block order (strided per thread, then block_sum)
template <int NV>
__device__ __forceinline__ void gather_partials(const PcgParams& P, int inst, double (&tot)[NV])
{
  const int b0 = P.inst_block0[inst], b1 = P.inst_block0[inst + 1];
#pragma unroll
  for (int k = 0; k < NV; ++k)
    tot[k] = 0.;
  for (int bb = b0 + threadIdx.x; bb < b1; bb += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      tot[k] += __hip_atomic_load(&P.partials[int64_t(bb) * NV + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  block_sum<NV>(tot);
}

template <int FMT>
__device__ __forceinline__ double row_apply(const PcgParams& P, int b, int inst, int64_t r0, int64_t i,
                                            const double* __restrict__ v)
{
  double acc = 0.;
  if constexpr (FMT == HGDT_FORMAT_DIA) {
    const int64_t lo = P.inst_row0[inst], hi = P.inst_row0[inst + 1];
    const int64_t* __restrict__ off = P.dia_off + int64_t(inst) * P.ndiag;
    for (int k = 0; k < P.ndiag; ++k) {
      const int64_t c = i + off[k];
      if (c >= lo && c < hi)
        acc = fma(P.dia[int64_t(k) * P.n + i], v[c], acc);
    }
  } else {
    const int64_t local = i - r0;
    const int64_t sl = P.block_slice0[b] + (local >> 6);
    const int lane = int(local & 63);
    const int64_t s0 = P.slice_ptr[sl];
    const int w = int((P.slice_ptr[sl + 1] - s0) >> 6);
    const int32_t* __restrict__ cc = P.sell_cols + s0 + lane;
    const double* __restrict__ vv = P.sell_vals + s0 + lane;
    for (int k = 0; k < w; ++k)
      acc = fma(vv[int64_t(k) * 64], v[cc[int64_t(k) * 64]], acc);
  }
  return acc;
}

template <int FMT>
__global__ void __launch_bounds__(1024) k_apply(PcgParams P, const double* __restrict__ x, double* __restrict__ y)
{
  const int b = blockIdx.x;
  const int inst = P.block_inst[b];
  const int64_t r0 = P.block_row0[b], r1 = P.block_row0[b + 1];
  for (int j = 0; j < P.items; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    if (i < r1)
      y[i] = row_apply<FMT>(P, b, inst, r0, i, x);
  }
}

__global__ void k_reset_scalars(PcgScalars* S, int ninst)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ninst) {
    PcgScalars z{};
    S[i] = z;
  }
}

__global__ void __launch_bounds__(1024) k_pcg_init(PcgParams P)
{
  const int b = blockIdx.x;
  const int inst = P.block_inst[b];
  const int64_t r0 = P.block_row0[b], r1 = P.block_row0[b + 1];
  double v[3] = {0., 0., 0.};
  for (int j = 0; j < P.items; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    if (i < r1) {
      const double bi = P.b[i];
      const double ri = P.ax ? bi - P.ax[i] : bi;
      const double zi = P.dinv[i] * ri;
      P.r[i] = ri;
      P.p[i] = zi;
      if (P.zero_guess)
        P.x[i] = 0.;
      v[0] += bi * bi, v[1] += ri * ri, v[2] += ri * zi;
    }
  }
  block_sum<3>(v);
  if (publish_partials<3>(P, b, inst, v)) {
    double tot[3];
    gather_partials<3>(P, inst, tot);
    if (threadIdx.x == 0) {
      PcgScalars& S = P.S[inst];
      S.bb = tot[0], S.rr = tot[1], S.rz = tot[2];
      S.tol2 = P.rtol2 * tot[0];
      S.iters = 0;
      S.alpha = 0., S.beta = 0., S.pq = 0.;
      S.done = (tot[1] <= S.tol2) ? 1 : 0;
      S.ticket = 0;
    }
  }
}

template <int FMT>
__global__ void __launch_bounds__(1024) k_pcg_spmv(PcgParams P)
{
  const int b = blockIdx.x;
  const int inst = P.block_inst[b];
  if (P.S[inst].done)
    return;
  const int64_t r0 = P.block_row0[b], r1 = P.block_row0[b + 1];
  double v[1] = {0.};
  for (int j = 0; j < P.items; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    if (i < r1) {
      const double qi = row_apply<FMT>(P, b, inst, r0, i, P.p);
      P.q[i] = qi;
      v[0] += P.p[i] * qi;
    }
  }
  block_sum<1>(v);
  if (publish_partials<1>(P, b, inst, v)) {
    double tot[1];
    gather_partials<1>(P, inst, tot);
    if (threadIdx.x == 0) {
      PcgScalars& S = P.S[inst];
      S.pq = tot[0];
      if (!(tot[0] > 0.)) {
        S.done = 3; // breakdown: A not positive definite on the Krylov space
      } else {
        S.alpha = S.rz / tot[0];
      }
      S.ticket = 0;
    }
  }
}

__global__ void __launch_bounds__(1024) k_pcg_update(PcgParams P)
{
  const int b = blockIdx.x;
  const int inst = P.block_inst[b];
  if (P.S[inst].done)
    return;
  const double alpha = P.S[inst].alpha;
  const int64_t r0 = P.block_row0[b], r1 = P.block_row0[b + 1];
  double v[2] = {0., 0.};
  for (int j = 0; j < P.items; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    if (i < r1) {
      P.x[i] += alpha * P.p[i];
      const double ri = P.r[i] - alpha * P.q[i];
      P.r[i] = ri;
      const double zi = P.dinv[i] * ri;
      v[0] += ri * zi, v[1] += ri * ri;
    }
  }
  block_sum<2>(v);
  if (publish_partials<2>(P, b, inst, v)) {
    double tot[2];
    gather_partials<2>(P, inst, tot);
    if (threadIdx.x == 0) {
      PcgScalars& S = P.S[inst];
      S.rr = tot[1];
      const int64_t it = S.iters + 1;
      S.iters = it;
      if (tot[1] <= S.tol2) {
        S.done = 1;
      } else if (it >= P.max_iter) {
        S.done = 2;
      } else {
        S.beta = tot[0] / S.rz;
        S.rz = tot[0];
      }
      S.ticket = 0;
    }
  }
}

__global__ void __launch_bounds__(1024) k_pcg_pupdate(PcgParams P)
{
  const int b = blockIdx.x;
  const int inst = P.block_inst[b];
  if (P.S[inst].done)
    return;
  const double beta = P.S[inst].beta;
  const int64_t r0 = P.block_row0[b], r1 = P.block_row0[b + 1];
  for (int j = 0; j < P.items; ++j) {
    const int64_t i = r0 + threadIdx.x + int64_t(j) * blockDim.x;
    if (i < r1)
      P.p[i] = P.dinv[i] * P.r[i] + beta * P.p[i];
  }
}

__global__ void k_scatter_values(int64_t nnz, const double* __restrict__ vals, const int64_t* __restrict__ pos,
                                 double* __restrict__ dst)
{
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s < nnz)
    dst[pos[s]] = vals[s];
}

__global__ void k_fill(int64_t n, double* __restrict__ dst, double v)
{
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s < n)
    dst[s] = v;
}

__global__ void k_jacobi(int64_t n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ cols,
                         const double* __restrict__ vals, double* __restrict__ dinv, int* bad)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  double d = 0.;
  for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
    if (cols[k] == i)
      d = vals[k];
  if (!(d > 0.)) {
    *bad = 1;
    d = 1.;
  }
  dinv[i] = 1. / d;
}

inline unsigned blocks_for(int64_t n, int block)
{
  return unsigned((n + block - 1) / block);
}

} // namespace

// per-operator solver workspace that the header does not need to know about
struct SolverCache {
  DeviceArray<int64_t> block_slice0;
  DeviceArray<int> bad;
  PcgScalars* host_scalars[2] = {nullptr, nullptr};
  int host_capacity = 0;
  hipEvent_t copied[2] = {nullptr, nullptr};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipStream_t own_stream = nullptr;
  hipEvent_t join = nullptr;
  hipGraphExec_t graph = nullptr;
  const double* graph_x = nullptr;
  const double* graph_b = nullptr;
  int64_t graph_maxit = 0;
  int graph_chunk = 0;
  int64_t graph_version = -1;
  int64_t version = 0;
  ~SolverCache()
  {
    for (auto& h : host_scalars)
      if (h)
        (void)hipHostFree(h);
    for (auto& e : copied)
      if (e)
        (void)hipEventDestroy(e);
    if (t0)
      (void)hipEventDestroy(t0);
    if (t1)
      (void)hipEventDestroy(t1);
    if (join)
      (void)hipEventDestroy(join);
    if (graph)
      (void)hipGraphExecDestroy(graph);
    if (own_stream)
      (void)hipStreamDestroy(own_stream);
  }
};

static std::mutex g_cache_mutex;

SolverCache& cache_of(Operator& op)
{
  if (!op.solver_cache) {
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    op.solver_cache = std::make_shared<SolverCache>();
  }
  return *op.solver_cache;
}

// row-block partition and storage structure of the operator (host set-up)
static void build_structure(Operator& op)
{
  SolverFormat& F = op.fmt;
  SolverCache& C = cache_of(op);
  const Space& s = *op.space;
  const int ninst = int(s.inst_dof.size()) - 1;
  const int block = op.launch[3].block, items = op.launch[3].items;
  const int64_t R = int64_t(block) * items;
  F.rows_per_block = int(R);
  F.block_row0.clear(), F.block_inst.clear(), F.inst_block0.assign(1, 0);
  for (int in = 0; in < ninst; ++in) {
    const int64_t a = s.inst_dof[in], e = s.inst_dof[in + 1];
    for (int64_t r = a; r < e; r += R) {
      F.block_row0.push_back(r);
      F.block_inst.push_back(in);
    }
    F.inst_block0.push_back(int32_t(F.block_row0.size()));
  }
  F.block_row0.push_back(op.n);
  F.nblocks = int64_t(F.block_inst.size());
  if (F.nblocks == 0)
    HGDT_THROW(HGDT_ERR_OPERATOR, "operator has no rows");
  F.d_block_row0.upload(F.block_row0);
  F.d_block_inst.upload(F.block_inst);
  F.d_inst_block0.upload(F.inst_block0);
  F.d_inst_row0.upload(s.inst_dof);

  // DIA detection: distinct (col - row) offsets per instance
  std::vector<std::vector<int64_t>> offs(ninst);
  bool dia_ok = op.requested_format != HGDT_FORMAT_SELL;
  for (int in = 0; in < ninst && dia_ok; ++in) {
    std::vector<int64_t>& o = offs[in];
    std::mutex m;
    std::atomic<bool> ok{true};
    parallel_for(s.inst_dof[in + 1] - s.inst_dof[in], [&](int64_t b0, int64_t b1) {
      std::vector<int64_t> loc;
      for (int64_t i = s.inst_dof[in] + b0; i < s.inst_dof[in] + b1 && ok; ++i)
        for (int64_t t = op.row_ptr[i]; t < op.row_ptr[i + 1]; ++t) {
          const int64_t d = int64_t(op.cols[t]) - i;
          auto it = std::lower_bound(loc.begin(), loc.end(), d);
          if (it == loc.end() || *it != d) {
            loc.insert(it, d);
            if (loc.size() > size_t(MAXDIAG)) {
              ok = false;
              break;
            }
          }
        }
      std::lock_guard<std::mutex> lock(m);
      for (int64_t d : loc) {
        auto it = std::lower_bound(o.begin(), o.end(), d);
        if (it == o.end() || *it != d)
          o.insert(it, d);
      }
    });
    if (!ok || o.size() > size_t(MAXDIAG))
      dia_ok = false;
  }
  if (op.requested_format == HGDT_FORMAT_DIA && !dia_ok)
    HGDT_THROW(HGDT_ERR_NOT_SUPPORTED, "DIA storage requested but an instance has more than 32 diagonals");
  F.d_dia.release(), F.d_dia_pos.release(), F.d_sell_vals.release(), F.d_sell_cols.release();
  F.d_sell_pos.release();
  if (dia_ok) {
    F.format = HGDT_FORMAT_DIA;
    int nd = 0;
    for (auto& o : offs)
      nd = std::max(nd, int(o.size()));
    F.ndiag = nd;
    std::vector<int64_t> table(int64_t(ninst) * nd, op.n); // sentinel offset: never inside an instance
    for (int in = 0; in < ninst; ++in)
      for (size_t k = 0; k < offs[in].size(); ++k)
        table[int64_t(in) * nd + k] = offs[in][k];
    std::vector<int64_t> pos(std::max<int64_t>(op.nnz, 1));
    for (int in = 0; in < ninst; ++in) {
      const auto& o = offs[in];
      parallel_for(s.inst_dof[in + 1] - s.inst_dof[in], [&](int64_t b0, int64_t b1) {
        for (int64_t i = s.inst_dof[in] + b0; i < s.inst_dof[in] + b1; ++i)
          for (int64_t t = op.row_ptr[i]; t < op.row_ptr[i + 1]; ++t) {
            const int64_t k = std::lower_bound(o.begin(), o.end(), int64_t(op.cols[t]) - i) - o.begin();
            pos[t] = k * op.n + i;
          }
      });
    }
    F.d_dia_offsets.upload(table);
    F.d_dia_pos.upload(pos);
    F.d_dia.allocate(std::max<int64_t>(int64_t(nd) * op.n, 1));
    F.bytes = int64_t(nd) * 8 * op.n + 16 * op.n;
  } else {
    F.format = HGDT_FORMAT_SELL;
    std::vector<int64_t> slice_ptr{0}, bslice0;
    for (int64_t b = 0; b < F.nblocks; ++b) {
      bslice0.push_back(int64_t(slice_ptr.size()) - 1);
      for (int64_t r = F.block_row0[b]; r < F.block_row0[b + 1]; r += 64) {
        int64_t w = 0;
        for (int64_t i = r; i < std::min(r + 64, F.block_row0[b + 1]); ++i)
          w = std::max(w, op.row_ptr[i + 1] - op.row_ptr[i]);
        slice_ptr.push_back(slice_ptr.back() + 64 * w);
      }
    }
    bslice0.push_back(int64_t(slice_ptr.size()) - 1);
    F.nslices = int64_t(slice_ptr.size()) - 1;
    const int64_t total = slice_ptr.back();
    std::vector<int32_t> scols(std::max<int64_t>(total, 1));
    std::vector<int64_t> pos(std::max<int64_t>(op.nnz, 1));
    parallel_for(F.nblocks, [&](int64_t b0, int64_t b1) {
      for (int64_t b = b0; b < b1; ++b) {
        int64_t sl = bslice0[b];
        for (int64_t r = F.block_row0[b]; r < F.block_row0[b + 1]; r += 64, ++sl) {
          const int64_t w = (slice_ptr[sl + 1] - slice_ptr[sl]) / 64;
          for (int64_t i = r; i < r + 64; ++i) {
            const int lane = int(i - r);
            const int64_t len = (i < F.block_row0[b + 1]) ? op.row_ptr[i + 1] - op.row_ptr[i] : 0;
            for (int64_t k = 0; k < w; ++k) {
              const int64_t at = slice_ptr[sl] + k * 64 + lane;
              if (k < len) {
                scols[at] = op.cols[op.row_ptr[i] + k];
                pos[op.row_ptr[i] + k] = at;
              } else {
                scols[at] = int32_t(std::min<int64_t>(i, op.n - 1)); // padding: zero value, valid column
              }
            }
          }
        }
      }
    });
    F.d_slice_ptr.upload(slice_ptr);
    C.block_slice0.upload(bslice0);
    F.d_sell_cols.upload(scols);
    F.d_sell_pos.upload(pos);
    F.d_sell_vals.allocate(std::max<int64_t>(total, 1));
    F.bytes = 12 * op.nnz + 16 * op.n;
  }
  F.structure_ready = true;
  ++C.version;
}

static PcgParams make_params(Operator& op)
{
  SolverFormat& F = op.fmt;
  SolverCache& C = cache_of(op);
  PcgParams P{};
  P.n = op.n;
  P.items = op.launch[3].items;
  P.block_row0 = F.d_block_row0.ptr;
  P.block_inst = F.d_block_inst.ptr;
  P.inst_block0 = F.d_inst_block0.ptr;
  P.inst_row0 = F.d_inst_row0.ptr;
  P.ndiag = F.ndiag;
  P.dia_off = F.d_dia_offsets.ptr;
  P.dia = F.d_dia.ptr;
  P.block_slice0 = C.block_slice0.ptr;
  P.slice_ptr = F.d_slice_ptr.ptr;
  P.sell_cols = F.d_sell_cols.ptr;
  P.sell_vals = F.d_sell_vals.ptr;
  P.dinv = F.d_dinv.ptr;
  return P;
}

void Operator::prepare_solver(hipStream_t stream)
{
  SolverFormat& F = fmt;
  SolverCache& C = cache_of(*this);
  if (!F.structure_ready)
    build_structure(*this);
  if (F.valid)
    return;
  const int B = 256;
  double* dst = (F.format == HGDT_FORMAT_DIA) ? F.d_dia.ptr : F.d_sell_vals.ptr;
  const int64_t total = (F.format == HGDT_FORMAT_DIA) ? int64_t(F.ndiag) * n : int64_t(F.d_sell_vals.count);
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(total, B)), dim3(B), 0, stream, total, dst, 0.);
  hipLaunchKernelGGL(k_scatter_values, dim3(blocks_for(nnz, B)), dim3(B), 0, stream, nnz, d_values.ptr,
                     (F.format == HGDT_FORMAT_DIA) ? F.d_dia_pos.ptr : F.d_sell_pos.ptr, dst);
  F.d_dinv.allocate(std::max<int64_t>(n, 1));
  C.bad.allocate(1);
  HGDT_HIP(hipMemsetAsync(C.bad.ptr, 0, sizeof(int), stream));
  hipLaunchKernelGGL(k_jacobi, dim3(blocks_for(n, B)), dim3(B), 0, stream, n, d_row_ptr.ptr, d_cols.ptr, d_values.ptr,
                     F.d_dinv.ptr, C.bad.ptr);
  HGDT_HIP(hipGetLastError());
  int bad = 0;
  HGDT_HIP(hipMemcpyAsync(&bad, C.bad.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
  HGDT_HIP(hipStreamSynchronize(stream));
  F.diag_positive = (bad == 0);
  F.valid = true;
}

void Operator::apply(const double* x, double* y, hipStream_t stream)
{
  prepare_solver(stream);
  PcgParams P = make_params(*this);
  const int block = launch[3].block;
  if (fmt.format == HGDT_FORMAT_DIA)
    hipLaunchKernelGGL(k_apply<HGDT_FORMAT_DIA>, dim3(unsigned(fmt.nblocks)), dim3(block), 0, stream, P, x, y);
  else
    hipLaunchKernelGGL(k_apply<HGDT_FORMAT_SELL>, dim3(unsigned(fmt.nblocks)), dim3(block), 0, stream, P, x, y);
  HGDT_HIP(hipGetLastError());
}

static void launch_iteration(const Operator& op, const PcgParams& P, hipStream_t stream, hipEvent_t ev0,
                             hipEvent_t ev1)
{
  const unsigned nb = unsigned(op.fmt.nblocks);
  const int block = op.launch[3].block;
  if (ev0)
    HGDT_HIP(hipEventRecord(ev0, stream));
  if (op.fmt.format == HGDT_FORMAT_DIA)
    hipLaunchKernelGGL(k_pcg_spmv<HGDT_FORMAT_DIA>, dim3(nb), dim3(block), 0, stream, P);
  else
    hipLaunchKernelGGL(k_pcg_spmv<HGDT_FORMAT_SELL>, dim3(nb), dim3(block), 0, stream, P);
  if (ev1)
    HGDT_HIP(hipEventRecord(ev1, stream));
  hipLaunchKernelGGL(k_pcg_update, dim3(nb), dim3(block), 0, stream, P);
  hipLaunchKernelGGL(k_pcg_pupdate, dim3(nb), dim3(block), 0, stream, P);
}

void Operator::apply_inverse(const double* b, double* x, const hgdt_solver_options& opt, hgdt_solver_stats& st,
                             hipStream_t user_stream)
{
  SolverCache& C = cache_of(*this);
  if (opt.check_every < 1 || opt.max_iter < 1 || !(opt.rtol >= 0.))
    HGDT_THROW(HGDT_ERR_INVALID_ARGUMENT, "invalid solver options");
  prepare_solver(user_stream);
  if (!fmt.diag_positive)
    HGDT_THROW(HGDT_ERR_SOLVER_FAILED, "Jacobi-PCG needs a positive diagonal: the matrix is not SPD");
  const int ninst = int(space->inst_dof.size()) - 1;
  d_r.allocate(std::max<int64_t>(n, 1));
  d_p.allocate(std::max<int64_t>(n, 1));
  d_q.allocate(std::max<int64_t>(n, 1));
  d_partials.allocate(std::max<int64_t>(3 * fmt.nblocks, 1));
  d_scalars.allocate(ninst);
  if (C.host_capacity < ninst) {
    for (auto& h : C.host_scalars) {
      if (h)
        HGDT_HIP(hipHostFree(h));
      HGDT_HIP(hipHostMalloc((void**)&h, sizeof(PcgScalars) * ninst));
    }
    C.host_capacity = ninst;
  }
  for (auto& e : C.copied)
    if (!e)
      HGDT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!C.t0)
    HGDT_HIP(hipEventCreate(&C.t0));
  if (!C.t1)
    HGDT_HIP(hipEventCreate(&C.t1));
  // graphs cannot be captured on the legacy default stream: run on an owned stream joined to the caller's
  hipStream_t stream = user_stream;
  const bool graph = opt.use_graph && !opt.time_spmv;
  if (graph) {
    if (!C.own_stream)
      HGDT_HIP(hipStreamCreateWithFlags(&C.own_stream, hipStreamNonBlocking));
    if (!C.join)
      HGDT_HIP(hipEventCreateWithFlags(&C.join, hipEventDisableTiming));
    HGDT_HIP(hipEventRecord(C.join, user_stream));
    HGDT_HIP(hipStreamWaitEvent(C.own_stream, C.join, 0));
    stream = C.own_stream;
  }
  PcgParams P = make_params(*this);
  P.b = b, P.x = x, P.r = d_r.ptr, P.p = d_p.ptr, P.q = d_q.ptr;
  P.partials = d_partials.ptr, P.S = d_scalars.ptr;
  P.rtol2 = opt.rtol * opt.rtol, P.max_iter = opt.max_iter, P.zero_guess = opt.zero_initial_guess ? 1 : 0;
  const int block = launch[3].block;
  HGDT_HIP(hipEventRecord(C.t0, stream));
  hipLaunchKernelGGL(k_reset_scalars, dim3(blocks_for(ninst, 256)), dim3(256), 0, stream, d_scalars.ptr, ninst);
  if (!opt.zero_initial_guess) {
    apply(x, d_q.ptr, stream);
    P.ax = d_q.ptr;
  }
  hipLaunchKernelGGL(k_pcg_init, dim3(unsigned(fmt.nblocks)), dim3(block), 0, stream, P);
  HGDT_HIP(hipGetLastError());
  P.ax = nullptr;

  std::vector<hipEvent_t> evs;
  const int chunk = opt.check_every;
  if (graph && (C.graph == nullptr || C.graph_x != x || C.graph_chunk != chunk || C.graph_version != C.version
                || C.graph_b != b || C.graph_maxit != opt.max_iter)) {
    if (C.graph)
      HGDT_HIP(hipGraphExecDestroy(C.graph));
    C.graph = nullptr;
    hipGraph_t g;
    HGDT_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < chunk; ++it)
      launch_iteration(*this, P, stream, nullptr, nullptr);
    HGDT_HIP(hipStreamEndCapture(stream, &g));
    HGDT_HIP(hipGraphInstantiate(&C.graph, g, nullptr, nullptr, 0));
    HGDT_HIP(hipGraphDestroy(g));
    C.graph_x = x, C.graph_b = b, C.graph_chunk = chunk, C.graph_version = C.version;
    C.graph_maxit = opt.max_iter;
  }
  auto launch_chunk = [&]() {
    if (graph) {
      HGDT_HIP(hipGraphLaunch(C.graph, stream));
    } else {
      for (int it = 0; it < chunk; ++it) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (opt.time_spmv) {
          HGDT_HIP(hipEventCreate(&e0));
          HGDT_HIP(hipEventCreate(&e1));
          evs.push_back(e0), evs.push_back(e1);
        }
        launch_iteration(*this, P, stream, e0, e1);
      }
    }
    HGDT_HIP(hipGetLastError());
  };
  auto all_done = [&](const PcgScalars* h) {
    for (int i = 0; i < ninst; ++i)
      if (h[i].done == 0)
        return false;
    return true;
  };
  // double-buffered convergence checks: chunk k+1 is queued before the flags of chunk k are read
  const int64_t max_chunks = (opt.max_iter + chunk - 1) / chunk + 1;
  int cur = 0;
  int64_t launched = 0;
  launch_chunk();
  ++launched;
  HGDT_HIP(hipMemcpyAsync(C.host_scalars[cur], d_scalars.ptr, sizeof(PcgScalars) * ninst, hipMemcpyDeviceToHost,
                          stream));
  HGDT_HIP(hipEventRecord(C.copied[cur], stream));
  while (true) {
    const int nxt = cur ^ 1;
    bool queued = false;
    if (launched < max_chunks) {
      launch_chunk();
      ++launched;
      HGDT_HIP(hipMemcpyAsync(C.host_scalars[nxt], d_scalars.ptr, sizeof(PcgScalars) * ninst,
                              hipMemcpyDeviceToHost, stream));
      HGDT_HIP(hipEventRecord(C.copied[nxt], stream));
      queued = true;
    }
    HGDT_HIP(hipEventSynchronize(C.copied[cur]));
    if (all_done(C.host_scalars[cur]) || !queued) {
      if (!queued)
        break;
      cur = nxt;
      HGDT_HIP(hipEventSynchronize(C.copied[cur]));
      break;
    }
    cur = nxt;
  }
  HGDT_HIP(hipEventRecord(C.t1, stream));
  HGDT_HIP(hipEventSynchronize(C.t1));
  if (graph) {
    HGDT_HIP(hipEventRecord(C.join, stream));
    HGDT_HIP(hipStreamWaitEvent(user_stream, C.join, 0));
  }
  const PcgScalars* h = C.host_scalars[cur];
  last_iters.assign(ninst, 0), last_relres.assign(ninst, 0.);
  st = hgdt_solver_stats{};
  st.converged = 1;
  for (int i = 0; i < ninst; ++i) {
    last_iters[i] = h[i].iters;
    last_relres[i] = h[i].bb > 0. ? std::sqrt(h[i].rr / h[i].bb) : 0.;
    st.iterations = std::max<int64_t>(st.iterations, h[i].iters);
    st.relres = std::max(st.relres, last_relres[i]);
    if (h[i].done != 1)
      st.converged = 0;
    if (h[i].done == 3)
      st.breakdown = 1;
  }
  float ms = 0.f;
  HGDT_HIP(hipEventElapsedTime(&ms, C.t0, C.t1));
  st.solve_ms = ms;
  if (opt.time_spmv && !evs.empty()) {
    // only iterations that did work: the early-exit launches after convergence are skipped
    double total = 0.;
    int64_t cnt = 0;
    const int64_t useful = st.iterations;
    for (size_t k = 0; k + 1 < evs.size(); k += 2) {
      if (int64_t(k / 2) < useful) {
        float t = 0.f;
        HGDT_HIP(hipEventElapsedTime(&t, evs[k], evs[k + 1]));
        total += t, ++cnt;
      }
    }
    for (auto e : evs)
      (void)hipEventDestroy(e);
    st.spmv_avg_ms = cnt ? total / cnt : 0.;
    st.spmv_count = cnt;
  }
  if (!st.converged) {
    if (st.breakdown)
      HGDT_THROW(HGDT_ERR_SOLVER_FAILED, "Jacobi-PCG breakdown: p.Ap <= 0 (matrix not SPD)");
    HGDT_THROW(HGDT_ERR_SOLVER_FAILED, "Jacobi-PCG did not reach the requested tolerance within max_iter");
  }
}

} // namespace hgdt
