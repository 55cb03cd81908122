// sunhip_common.h -- device helpers shared by the solver kernels: XCD-aware
// block order and deterministic two-level reductions finished by the last
// block to arrive.
#pragma once
#include <hip/hip_runtime.h>

#include "sunhip_internal.h"

namespace sunhip {

// Workgroups are dealt to the 8 XCDs round-robin.  Renumber them so that each
// XCD walks one contiguous run of tiles: neighbouring columns then share an
// L2, and the run is walked in Morton order.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return x * q + (x < r ? x : r) + i;
}

// The tile group (4 tiles) a block of a column kernel works on: xcd_block,
// then the grid's depth-first order within each XCD's run when it has one.
__device__ __forceinline__ int tile_group(const Geo &G, int b, int nb) {
  const int v = xcd_block(b, nb);
  return G.b4perm ? G.b4perm[v] : v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double ld_agent(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduce NV per-thread values over the whole grid.  Every block publishes its
// partial sums; the last block to take a ticket adds all partials in a fixed
// order (so the result does not depend on scheduling) and returns true with
// the totals in tot[].  The ticket counter is reset for the next launch.
template <int NV, int NT>
__device__ bool grid_reduce(const double (&v)[NV], double *part, unsigned *counter, double (&tot)[NV]) {
  constexpr int NW = NT / 64;
  __shared__ double sh[NV][NW > 0 ? NW : 1];
  __shared__ int last;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    const double s = wave_sum(v[i]);
    if (l == 0) sh[i][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double s = 0;
      for (int j = 0; j < NW; j++) s += sh[i][j];
      st_agent(&part[(size_t)blockIdx.x * NV + i], s);
    }
    const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  const int nb = gridDim.x;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = 0;
    for (int b = threadIdx.x; b < nb; b += NT) s += ld_agent(&part[(size_t)b * NV + i]);
    s = wave_sum(s);
    __syncthreads();
    if (l == 0) sh[i][w] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = 0;
    for (int j = 0; j < NW; j++) s += sh[i][j];
    tot[i] = s;
  }
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

}  // namespace sunhip
