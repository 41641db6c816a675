// block_reduce.hpp -- launch-order and reduction helpers shared by the solver's kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace hgdt {
namespace {

// XCD-aware logical block: hardware block hw runs logical block (hw % 8) * ceil(nb / 8) + hw / 8, so each
// XCD walks one contiguous eighth of the blocks (rev: the same eighth from its top down); -1 for the
// padding blocks of the last XCD
__device__ __forceinline__ int logical_block(int nb, bool rev = false)
{
  const int q = (nb + 7) >> 3;
  const int hw = blockIdx.x;
  const int s = hw >> 3;
  const int L = (hw & 7) * q + (rev ? q - 1 - s : s);
  return L < nb ? L : -1;
}

// logical block of a launch over a list of active blocks (NULL: all blocks 0 .. nb - 1)
__device__ __forceinline__ int active_block(int nb, const int32_t* __restrict__ list, bool rev = false)
{
  const int L = logical_block(nb, rev);
  return (L < 0 || !list) ? L : list[L];
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

// deterministic block reduction of NV values; the result is returned to every thread
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
  __shared__ double lds[NV][16];
  __shared__ double res[NV];
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
      res[k] = s;
    }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k)
    v[k] = res[k];
  __syncthreads();
}

// *flag = 1 when `pred` holds for an active lane of the wave: one store per wave (by its first active lane)
// when every active lane names the same flag, else one store per lane whose predicate holds.  Every store
// writes the same value, so the flag needs no atomic.
__device__ __forceinline__ void wave_flag(bool pred, int* flag)
{
  if (!__ballot(pred))
    return;
  const int leader = __ffsll((long long)__ballot(1)) - 1;
  const long long mine = (long long)(uintptr_t)flag;
  const bool same = !__ballot(mine != __shfl(mine, leader, 64));
  if (same ? int(threadIdx.x & 63) == leader : pred)
    *flag = 1;
}

// sum of an instance's block partials (written by an earlier kernel) in block order
// (each lane sums the partials b0 + lane, b0 + lane + B, ... in that order; the loads of U such
// partials are issued together, so the prologue costs one memory round trip per U * B partials:
// one round trip for up to 2048 single partials with B = 256).  Split in two so that a kernel can issue
// the first trip's loads together with its other loads and wait for all of them once:
// Partials::issue loads the first trip, Partials::sum adds it (and any further trips) and reduces the block.
template <int NV>
struct Partials {
  static constexpr int U = NV >= 8 ? 1 : 8 / NV; // partials per thread and trip: few registers, since callers
                                                 // keep their rows' loads in flight across the reduction
  double v[U][NV];
  int b0, b1;
  // the loads are unconditional (a lane past the range reads the partial b0, or 0 of an empty range, and
  // discards it in sum), so no branch separates them from the caller's other loads
  __device__ __forceinline__ void issue(const double* __restrict__ part, int first, int last)
  {
    b0 = first, b1 = last;
    const int base = b0 + int(threadIdx.x), B = blockDim.x;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int bb = base + u * B;
      const int64_t src = (bb < b1) ? bb : b0;
#pragma unroll
      for (int k = 0; k < NV; ++k)
        v[u][k] = part[src * NV + k];
    }
  }
  __device__ __forceinline__ void sum(const double* __restrict__ part, double (&tot)[NV])
  {
    const int B = blockDim.x;
#pragma unroll
    for (int k = 0; k < NV; ++k)
      tot[k] = 0.;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = b0 + int(threadIdx.x) + u * B < b1;
#pragma unroll
      for (int k = 0; k < NV; ++k)
        tot[k] += in ? v[u][k] : 0.;
    }
    for (int base = b0 + int(threadIdx.x) + U * B; base < b1; base += U * B) {
      double w[U][NV];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int bb = base + u * B;
#pragma unroll
        for (int k = 0; k < NV; ++k)
          w[u][k] = (bb < b1) ? part[int64_t(bb) * NV + k] : 0.;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int k = 0; k < NV; ++k)
          tot[k] += w[u][k];
    }
    block_sum<NV>(tot);
  }
};

template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int b0, int b1, double (&tot)[NV])
{
  Partials<NV> p;
  p.issue(part, b0, b1);
  p.sum(part, tot);
}

} // namespace
} // namespace hgdt
