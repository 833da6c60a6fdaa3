// K1: CELL_SORT (cellsort.f90:53-80) as a stable counting sort on the device.
// Atoms are physically reordered so that each cell's members are contiguous
// (coalesced reads in the pair kernels); within a cell they keep ascending
// caller index, which is the reference's link order.
#pragma once
#include "common.hpp"

namespace sx {

// cell id = MOD(INT((x + L/2)/Lcell) + Nc, Nc) per axis (cellsort.f90:69-72)
__global__ void k_cell_assign(Geom g, const double* __restrict__ x, int* __restrict__ cell,
                              int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  int id3[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double xx;
    {
#pragma clang fp contract(off)
      xx = x[3 * i + d] + g.hbox[d];
    }
    int c = int(xx / g.lcell[d]);
    c = (c + g.nc[d]) % g.nc[d];  // the reference's MOD(id + Nc, Nc)
    id3[d] = c < 0 ? c + g.nc[d] : c;  // positions far outside the box stay in bounds
  }
  const int c = id3[0] + id3[1] * g.nc[0] + id3[2] * g.nc[0] * g.nc[1];
  cell[i] = c;
  atomicAdd(&count[c], 1);
}

// Exclusive scan of the per-cell counts (single block; ncall is small).
__global__ void k_cell_scan(int ncall, const int* __restrict__ count, int* __restrict__ start,
                            int* __restrict__ fill, int* __restrict__ err) {
  __shared__ int part[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < ncall; base += blockDim.x) {
    const int c = base + threadIdx.x;
    const int v = c < ncall ? count[c] : 0;
    if (v > NPT_MAX) atomicCAS(err, 0, SIMOX_ERR_CELL_OVERFLOW);
    part[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < blockDim.x; o <<= 1) {
      int t = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    if (c < ncall) {
      start[c] = carry + part[threadIdx.x] - v;
      fill[c] = 0;
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += part[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) start[ncall] = carry;
}

// Unordered scatter of atom slots into their cell segments.
__global__ void k_cell_scatter(int n, const int* __restrict__ cell, const int* __restrict__ start,
                               int* __restrict__ fill, int* __restrict__ slot_src) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell[i];
  const int p = start[c] + atomicAdd(&fill[c], 1);
  slot_src[p] = i;
}

// Order each cell's members by caller index (rank counting in LDS; a cell holds
// at most NPT_MAX atoms), giving order[new_pos] = old_pos.  perm = null: the
// positions are in caller order already (the key is the position itself).
__global__ void __launch_bounds__(BLOCK) k_cell_order(const int* __restrict__ start,
                                                      const int* __restrict__ slot_src,
                                                      const int* __restrict__ perm,
                                                      int* __restrict__ order) {
  __shared__ int key[NPT_MAX];
  __shared__ int src[NPT_MAX];
  const int c = blockIdx.x;
  const int s0 = start[c];
  const int m = start[c + 1] - s0;
  if (m > NPT_MAX) {
    // reported by k_cell_scan; the slots stay unordered but every order entry
    // is written, so the kernels after the sort index in bounds
    for (int t = threadIdx.x; t < m; t += blockDim.x) order[s0 + t] = slot_src[s0 + t];
    return;
  }
  for (int t = threadIdx.x; t < m; t += blockDim.x) {
    src[t] = slot_src[s0 + t];
    key[t] = perm ? perm[src[t]] : src[t];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < m; t += blockDim.x) {
    const int k = key[t];
    int rank = 0;
    for (int u = 0; u < m; ++u) rank += key[u] < k;
    order[s0 + rank] = src[t];
  }
}

// Gather the state into the new order.
__global__ void k_reorder(int n, const int* __restrict__ order, const int* __restrict__ cell,
                          const double* __restrict__ x, const double* __restrict__ v,
                          const double* __restrict__ f, const double* __restrict__ q,
                          const int* __restrict__ id, const int* __restrict__ perm,
                          double* __restrict__ x2, double* __restrict__ v2,
                          double* __restrict__ f2, double* __restrict__ q2,
                          int* __restrict__ id2, int* __restrict__ perm2,
                          int* __restrict__ cell2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int k = order[p];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x2[3 * p + d] = x[3 * k + d];
    v2[3 * p + d] = v[3 * k + d];
    f2[3 * p + d] = f[3 * k + d];
  }
  q2[p] = q[k];
  id2[p] = id[k];
  perm2[p] = perm[k];
  cell2[p] = cell[k];
}

}  // namespace sx
