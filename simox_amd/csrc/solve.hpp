// CTIP charge-solve kernels (K10, K11), velocity Verlet (K13) and the
// deterministic reductions they use.
#pragma once
#include "common.hpp"
#include "pme.hpp"

namespace sx {

// K10: g = -dU/dq at ql (dUdq, ctip.f90:307-333) with the PME potential
// gathered from the convolved grid in the same kernel:
//   g_i = -(chi + J' q_i + c_pme phi_i + 2w (min(q-qmin,0) + max(q-qmax,0)))
//         - b_i - sum_j sym_ij q_j
// part[3*block] = sum g, part[3*block+1] = -sum g.p (Brent's f, ctip.f90:236),
// part[3*block+2] = sum g^2.
// sum_k val[k n + i] vec[col[k n + i]] over row i of the ELL matrix
// (column-major, k < c), accumulated in Tv in ascending k.  Sixteen columns
// and values are loaded before their vector entries, so each thread keeps
// many gathers in flight.
// scalar steps of the device-side CG (defined with the line-search kernels)
__device__ inline void ls_init(const double* red, double* ls);
__device__ inline void ls_update(const double* red, double* ls, bool bounds = false);
__device__ inline void beta_update(const double* red, double* ls);
template <int OP>
__device__ inline void reduce_scalar(const double* out, double* ls);

// Final reduction done by the producer's last workgroup instead of a
// k_reduce_op launch: ctr (zero between launches) counts the finished
// workgroups, out receives the row sums, op the scalar step of k_reduce_op
// (0 none, 2 ls_update, 4 projected beta); ctr == null: rows only.
constexpr int RED_GRID = 256;
// threads per workgroup of k_ncg_step_dev (four times BLOCK: its 256
// workgroups then keep 16 waves per CU in flight, 14.4 -> 13.7 us at the
// 1M-atom slab; the line search and the preconditioner's finish are faster
// with BLOCK)
constexpr int RED_NT = 1024;
struct RedTail {
  unsigned* ctr;
  double* out;
  double* ls;
  int op;
};

// The workgroup's row of NV partial sums (valid in thread 0 after
// block_sum) to part[NV*block ..].  With t.ctr, each workgroup then counts
// itself done; the one that counts last sums every row in workgroup order, as
// k_reduce_op does (fixed rows per thread, fixed shuffle tree, fixed wave
// order: deterministic), writes t.out, runs the scalar step and clears the
// counter for the next launch.  The rows are stored and read with agent-scope
// atomic operations, which are coherent across the XCDs' L2 caches, and each
// workgroup waits for its stores to be acknowledged before it counts itself
// done.  An acquire-release count would do the same with an L2 writeback and
// an L2 invalidate in every workgroup: measured ten times slower (6.9 -> 70 us
// per line-search round at C3).
// The producers run on RED_GRID workgroups at most (grid-stride over the
// atoms): at C3, 4,096 workgroups counting on one address took 40 us per
// launch, 512 took 11 us and 256 took 10 us per line-search round (the
// preconditioner's finish, whose gathers want more workgroups: 29 us at 256,
// 33 us at 1,024).
// lds: the block_sum scratch (NV * blockDim / WAVE doubles).
template <int NV>
__device__ inline void partials_out(const double (&acc)[NV], double* part, const RedTail& t,
                                    double* lds) {
  __shared__ int last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      __hip_atomic_store(&part[size_t(NV) * blockIdx.x + k], acc[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    last = 0;
    if (t.ctr) {
      __builtin_amdgcn_s_waitcnt(0);  // the row's stores acknowledged
      last = __hip_atomic_fetch_add(t.ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             gridDim.x - 1;
    }
  }
  if (!t.ctr) return;  // uniform over the launch
  __syncthreads();
  if (!last) return;   // uniform over the workgroup
  double s[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) s[v] = 0.0;
  for (int k = threadIdx.x; k < int(gridDim.x); k += blockDim.x)
#pragma unroll
    for (int v = 0; v < NV; ++v)
      s[v] += __hip_atomic_load(&part[size_t(NV) * k + v], __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double x = wave_sum(s[v]);
    if (lane == 0) lds[v * nw + w] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double x = lds[v * nw];
      for (int q = 1; q < nw; ++q) x += lds[v * nw + q];
      t.out[v] = x;
    }
    if (t.op == 2) reduce_scalar<2>(t.out, t.ls);
    if (t.op == 4) reduce_scalar<4>(t.out, t.ls);
    __hip_atomic_store(t.ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// partials_out for the line-search rounds: rows of {D1, S2, up, dn}; the last
// workgroup sums the first two in workgroup order, takes min / max of the
// others, writes t.out[0..3] and runs ls_update with the crossings.
__device__ inline void partials_out_ls(const double (&row)[4], double* part, const RedTail& t,
                                       double* lds) {
  __shared__ int last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      __hip_atomic_store(&part[size_t(4) * blockIdx.x + k], row[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // the row's stores acknowledged
    last = __hip_atomic_fetch_add(t.ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
           gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;  // uniform over the workgroup
  double s0 = 0.0, s1 = 0.0, up = 1e300, dn = -1e300;
  for (int k = threadIdx.x; k < int(gridDim.x); k += blockDim.x) {
    s0 += __hip_atomic_load(&part[size_t(4) * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s1 += __hip_atomic_load(&part[size_t(4) * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    up = fmin(up, __hip_atomic_load(&part[size_t(4) * k + 2], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT));
    dn = fmax(dn, __hip_atomic_load(&part[size_t(4) * k + 3], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT));
  }
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  up = -wave_max(-up);
  dn = wave_max(dn);
  if (lane == 0) {
    lds[w] = s0;
    lds[nw + w] = s1;
  }
  __shared__ double mm[2 * 16];
  if (lane == 0) {
    mm[2 * w] = up;
    mm[2 * w + 1] = dn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x0 = lds[0], x1 = lds[nw];
    for (int q = 1; q < nw; ++q) {
      x0 += lds[q];
      x1 += lds[nw + q];
      up = fmin(up, mm[2 * q]);
      dn = fmax(dn, mm[2 * q + 1]);
    }
    up = fmin(up, mm[0]);
    dn = fmax(dn, mm[1]);
    t.out[0] = x0;
    t.out[1] = x1;
    t.out[2] = up;
    t.out[3] = dn;
    ls_update(t.out, t.ls, true);
    __hip_atomic_store(t.ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class Tv, class Tp>
__device__ inline Tv ell_dot(int c, size_t n, int i, const int* __restrict__ col,
                             const Tv* __restrict__ val, const Tp* __restrict__ vec) {
  // batches of U entries with their loads in flight together, then the
  // rest in masked batches of UT (value 0 at column i past the row's end):
  // exact zeros added in the order an entry-by-entry tail would sum
  constexpr int U = 16, UT = 4;
  Tv s = Tv(0);
  int k = 0;
  for (; k + U <= c; k += U) {
    int cc[U];
    Tv vv[U], pp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cc[u] = col[size_t(k + u) * n + i];
      vv[u] = val[size_t(k + u) * n + i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) pp[u] = Tv(vec[cc[u]]);
#pragma unroll
    for (int u = 0; u < U; ++u) s += vv[u] * pp[u];
  }
  for (; k < c; k += UT) {
    int cc[UT];
    Tv vv[UT];
#pragma unroll
    for (int u = 0; u < UT; ++u) {
      const bool ok = k + u < c;
      const size_t e = size_t(ok ? k + u : k) * n + i;  // in bounds either way
      cc[u] = ok ? col[e] : i;
      vv[u] = ok ? val[e] : Tv(0);
    }
#pragma unroll
    for (int u = 0; u < UT; ++u) s += vv[u] * Tv(vec[cc[u]]);
  }
  return s;
}

template <class Rg, class Tv>
__global__ void __launch_bounds__(BLOCK) k_dudq(Geom g, const Tables* __restrict__ tb,
                                                const double* __restrict__ x,
                                                const int* __restrict__ id,
                                                const double* __restrict__ ql,
                                                const Rg* __restrict__ grid, double c_pme,
                                                const int* __restrict__ cnt,
                                                const int* __restrict__ col,
                                                const Tv* __restrict__ val,
                                                const double* __restrict__ b,
                                                const double* __restrict__ p,
                                                double* __restrict__ gout,
                                                double* __restrict__ phi_out,
                                                double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < g.n) {
    const double phi = c_pme * gather_phi(g, x + 3 * i, grid);
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double qi = ql[i];
    double dqmin = qi - sc.qmin, dqmax = qi - sc.qmax;
    if (dqmin >= 0.0) dqmin = 0.0;
    if (dqmax <= 0.0) dqmax = 0.0;
    double gi = -(sc.chi + sc.Jp * qi + phi + 2.0 * WPEN * (dqmin + dqmax));
    gi -= b[i];
    const int c = cnt[i];
    const size_t n = size_t(g.n);
    const Tv s = ell_dot(c, n, i, col, val, ql);
    gi -= double(s);
    gout[i] = gi;
    if (phi_out) phi_out[i] = phi;
    acc[0] = gi;
    if (p) acc[1] = -gi * p[i];
    acc[2] = gi * gi;
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = acc[0];
    part[3 * blockIdx.x + 1] = acc[1];
    part[3 * blockIdx.x + 2] = acc[2];
  }
}

// Final pass of a deterministic reduction: out[v] = sum over nparts of
// part[nv*k + v] (single block, fixed order); maxmask marks values reduced by
// max.  All nv <= RED_MAXV values of a row are read in one pass.
constexpr int RED_MAXV = 10;
__global__ void k_reduce_final(const double* __restrict__ part, int nparts, int nv, int maxmask,
                               double* __restrict__ out) {
  constexpr int NW = 1024 / WAVE;
  __shared__ double lds[RED_MAXV * NW];
  if (nv > RED_MAXV) return;  // callers reduce at most RED_MAXV values
  double s[RED_MAXV];
#pragma unroll
  for (int v = 0; v < RED_MAXV; ++v) s[v] = 0.0;
  // every value of a partial row in one pass (fixed order per thread)
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
#pragma unroll
    for (int v = 0; v < RED_MAXV; ++v) {
      if (v >= nv) break;
      const double e = part[size_t(nv) * k + v];
      s[v] = ((maxmask >> v) & 1) ? fmax(s[v], e) : s[v] + e;
    }
  }
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
  for (int v = 0; v < RED_MAXV; ++v) {
    if (v >= nv) break;
    const double t = ((maxmask >> v) & 1) ? wave_max(s[v]) : wave_sum(s[v]);
    if (lane == 0) lds[v * NW + w] = t;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    const int v = threadIdx.x;
    const bool mx = (maxmask >> v) & 1;
    const int nw = int(blockDim.x / WAVE);
    double t = lds[v * NW];
    for (int q = 1; q < nw; ++q) t = mx ? fmax(t, lds[v * NW + q]) : t + lds[v * NW + q];
    out[v] = t;
  }
}

// Nonlinear-CG direction (ctip.f90:172-181): d = g + p * gsum/gsum_old, then
// p = d - mean(d).  Pass 1 writes d and sum(d); pass 2 projects and reports max|p|.
__global__ void __launch_bounds__(BLOCK) k_cg_dir1(int n, const double* __restrict__ g,
                                                   const double* __restrict__ p, double gsum,
                                                   double gsum_old, double* __restrict__ d,
                                                   double* __restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const double di = g[i] + p[i] * gsum / gsum_old;
    d[i] = di;
    acc[0] = di;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ void __launch_bounds__(BLOCK) k_cg_dir2(int n, const double* __restrict__ d,
                                                   const double* __restrict__ dsum,
                                                   double* __restrict__ p,
                                                   double* __restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (i < n) {
    const double pi = d[i] - dsum[0] / double(n);
    p[i] = pi;
    m = fabs(pi);
  }
  m = wave_max(m);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  if (lane == 0) lds[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = lds[0];
    for (int q = 1; q < int(blockDim.x / WAVE); ++q) t = fmax(t, lds[q]);
    part[blockIdx.x] = t;
  }
}

// FP32 copy of the FP64 CTIP matrix values (rows up to their counts)
__global__ void k_ell_to_f32(int n, int cap, const int* __restrict__ cnt,
                             const double* __restrict__ v64, float* __restrict__ v32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cnt[i] < cap ? cnt[i] : cap;
  for (int k = 0; k < c; ++k) v32[size_t(k) * n + i] = float(v64[size_t(k) * n + i]);
}

__global__ void k_axpy(int n, double a, const double* __restrict__ p, double* __restrict__ q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = q[i] + a * p[i];
}

// Self / charge energy partials: part[3k] = -sum q^2, part[3k+1] = sum chi q +
// 0.5 J q^2 (J or J' = J - 2 kc a/sqrt(pi)), part[3k+2] = penalty sum with
// weight wpen (2w in PME_SELF, force.f90:494-496; w in CTIP_SOLVER,
// ctip.f90:201-205; quirk Q4).
__global__ void __launch_bounds__(BLOCK) k_self_energy(int n, const Tables* __restrict__ tb,
                                                       const int* __restrict__ id,
                                                       const double* __restrict__ q, int use_jp,
                                                       double wpen, double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double qi = q[i];
    acc[0] = -(qi * qi);
    acc[1] = sc.chi * qi + 0.5 * (use_jp ? sc.Jp : sc.J) * (qi * qi);
    double dq = qi - sc.qmin;
    if (dq < 0.0) acc[2] += wpen * (dq * dq);
    dq = qi - sc.qmax;
    if (dq > 0.0) acc[2] += wpen * (dq * dq);
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = acc[0];
    part[3 * blockIdx.x + 1] = acc[1];
    part[3 * blockIdx.x + 2] = acc[2];
  }
}

// K13: VV_NVE1 (vverlet.f90:87-93): half kick, drift, DNINT wrap.
__global__ void k_vv1(int n, const Tables* __restrict__ tb, const int* __restrict__ id,
                      double b0, double b1, double b2, double dt, double* __restrict__ x,
                      double* __restrict__ v, const double* __restrict__ f) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xm = tb->sc[id[i] - 1].xm;
  const double box[3] = {b0, b1, b2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double vn = v[3 * i + d] + 0.5 * dt * f[3 * i + d] / xm;
    v[3 * i + d] = vn;
    const double xn = x[3 * i + d] + dt * vn;
    x[3 * i + d] = xn - box[d] * round(xn / box[d]);
  }
}

// OXIDATION-mode first half (VV_NVE1_REFLECT, vverlet.f90:100-129): as
// VV_NVE1 with x, y periodic and a reflecting wall at z = box_z / 2:
// z > box_z / 2 -> z = box_z - z, v_z = -|v_z|.
__global__ void k_vv1_reflect(int n, const Tables* __restrict__ tb, const int* __restrict__ id,
                              double b0, double b1, double b2, double dt, double* __restrict__ x,
                              double* __restrict__ v, const double* __restrict__ f) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xm = tb->sc[id[i] - 1].xm;
  const double box[3] = {b0, b1, b2};
  double xn[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double vn = v[3 * i + d] + 0.5 * dt * f[3 * i + d] / xm;
    v[3 * i + d] = vn;
    xn[d] = x[3 * i + d] + dt * vn;
  }
#pragma unroll
  for (int d = 0; d < 2; ++d) x[3 * i + d] = xn[d] - box[d] * round(xn[d] / box[d]);
  if (xn[2] > b2 / 2.0) {
    xn[2] = b2 - xn[2];
    v[3 * i + 2] = -fabs(v[3 * i + 2]);
  }
  x[3 * i + 2] = xn[2];
}

// REMOVE_RIGID_MOTION (vverlet.f90:211-234), pass 1: part[4k..] = sums of
// v_x, v_y, v_z and z; pass 2 subtracts the means (v, and z of the positions).
__global__ void __launch_bounds__(BLOCK) k_rigid_sums(int n, const double* __restrict__ x,
                                                      const double* __restrict__ v,
                                                      double* __restrict__ part) {
  __shared__ double lds[4 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    acc[0] = v[3 * i]; acc[1] = v[3 * i + 1]; acc[2] = v[3 * i + 2]; acc[3] = x[3 * i + 2];
  }
  block_sum<4>(acc, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = acc[k];
}
__global__ void k_rigid_apply(int n, const double* __restrict__ sums, double* __restrict__ x,
                              double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dn = double(n);
  for (int d = 0; d < 3; ++d) v[3 * i + d] = v[3 * i + d] - sums[d] / dn;
  x[3 * i + 2] = x[3 * i + 2] - sums[3] / dn;
}

// Langevin first half (VV_NVT1_STOCH, vverlet.f90:237-257):
//   v = v (1 - alpha dt / 2m) + dt f / 2m,  x += dt v, wrapped (DNINT).
__global__ void k_vv1_stoch(int n, const Tables* __restrict__ tb, const int* __restrict__ id,
                            double b0, double b1, double b2, double dt, double alpha,
                            double* __restrict__ x, double* __restrict__ v,
                            const double* __restrict__ f) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xm = tb->sc[id[i] - 1].xm;
  const double box[3] = {b0, b1, b2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double vn = v[3 * i + d] * (1.0 - 0.5 * alpha * dt / xm) + 0.5 * dt * f[3 * i + d] / xm;
    v[3 * i + d] = vn;
    const double xn = x[3 * i + d] + dt * vn;
    x[3 * i + d] = xn - box[d] * round(xn / box[d]);
  }
}

// Counter-based normal deviates for the Langevin force: splitmix64 of
// (seed, step, atom, component) gives two uniforms in (0, 1], Box-Muller gives
// N(0, 1).  Independent of launch shape and atom order (the caller's atom
// index keys the stream).
__device__ inline unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double gauss(unsigned long long seed, unsigned long long step, int atom, int k) {
  const unsigned long long key =
      splitmix64(seed ^ splitmix64(step * 0x100000001B3ull + (unsigned long long)(3 * atom + k)));
  const unsigned long long k2 = splitmix64(key);
  const double u1 = (double((key >> 11) + 1)) * (1.0 / 9007199254740992.0);  // (0, 1]
  const double u2 = double(k2 >> 11) * (1.0 / 9007199254740992.0);         // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
}

// Langevin second half (VV_NVT2_STOCH, vverlet.f90:260-296): per component
// f += eta beta (the random force stays in f for the next first half),
// v = (v + dt f / 2m) / (1 + alpha dt / 2m), part = m v^2 sums.  alpha and
// beta are single precision there (no IMPLICIT NONE: implicit REAL), as here.
__global__ void __launch_bounds__(BLOCK) k_vv2_stoch(int n, const Tables* __restrict__ tb,
                                                     const int* __restrict__ id,
                                                     const int* __restrict__ perm, double dt,
                                                     float alpha, float beta,
                                                     unsigned long long seed,
                                                     unsigned long long step,
                                                     double* __restrict__ v,
                                                     double* __restrict__ f,
                                                     double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const double xm = tb->sc[id[i] - 1].xm;
    double m2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double eta = beta != 0.0f ? gauss(seed, step, perm[i], k) : 0.0;
      const double fk = f[3 * i + k] + eta * double(beta);
      f[3 * i + k] = fk;
      const double vk = (v[3 * i + k] + 0.5 * dt * fk / xm) / (1.0 + 0.5 * double(alpha) * dt / xm);
      v[3 * i + k] = vk;
      m2 = m2 + xm * (vk * vk);
    }
    acc[0] = m2;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// Berendsen second half (VV_NVT2_BEREND, vverlet.f90:160-208), pass 1: half
// kick and per-species sums of m v^2 (summed per component) and counts;
// part[11k + s] = sum over species s+1, part[11k + 5 + s] = count.
__global__ void __launch_bounds__(BLOCK) k_vv2_berend1(int n, const Tables* __restrict__ tb,
                                                       const int* __restrict__ id, double dt,
                                                       double* __restrict__ v,
                                                       const double* __restrict__ f,
                                                       double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double lds[10 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < n) {
    const int s = id[i] - 1;
    const double xm = tb->sc[s].xm;
    double m2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double vk = v[3 * i + k] + 0.5 * dt * f[3 * i + k] / xm;
      v[3 * i + k] = vk;
      m2 = m2 + xm * (vk * vk);
    }
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      acc[t] = (t == s) ? m2 : 0.0;
      acc[5 + t] = (t == s) ? 1.0 : 0.0;
    }
  }
  block_sum<10>(acc, lds);
  if (threadIdx.x == 0)
    for (int t = 0; t < 10; ++t) part[10 * blockIdx.x + t] = acc[t];
}

// pass 2: lambda_s = sqrt(1 + C (T_given / T0_s - 1)), T0_s = sum_s / (3 N_s)
// (C_BEREND = 0.025), v *= lambda, part = m v^2 sums.
__global__ void __launch_bounds__(BLOCK) k_vv2_berend2(int n, const Tables* __restrict__ tb,
                                                       const int* __restrict__ id,
                                                       const double* __restrict__ sums,
                                                       double t_given, double* __restrict__ v,
                                                       double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const int s = id[i] - 1;
    const double xm = tb->sc[s].xm;
    const double t0 = sums[s] / (3.0 * sums[5 + s]);
    const double lam = sqrt(1.0 + 0.025 * (t_given / t0 - 1.0));
    double m2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double vk = v[3 * i + k] * lam;
      v[3 * i + k] = vk;
      m2 = m2 + xm * (vk * vk);
    }
    acc[0] = m2;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// K13: VV_NVE2 (vverlet.f90:145-154): half kick and m v^2 partial sums.
__global__ void __launch_bounds__(BLOCK) k_vv2(int n, const Tables* __restrict__ tb,
                                               const int* __restrict__ id, double dt,
                                               double* __restrict__ v,
                                               const double* __restrict__ f,
                                               double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const double xm = tb->sc[id[i] - 1].xm;
    double t = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double vn = v[3 * i + d] + 0.5 * dt * f[3 * i + d] / xm;
      v[3 * i + d] = vn;
      t = t + vn * vn;
    }
    acc[0] = xm * t;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

}  // namespace sx

namespace sx {

// ---------------------------------------------------------------------------
// Converged charge solve (modes FP64 / MIXED).  The charge energy is convex
// and quadratic on each active set of the qmin/qmax penalty, with Hessian
//   H = diag(J' + 2w [q outside (qmin, qmax)]) + c_pme P + S
// (P: PME potential operator, S: CTIP sparse matrix).  Outer loop (FP64): the
// true gradient g = -dU/dq (k_dudq); inner loop: projected preconditioned CG
// on H d = g with sum(d) = 0, operator in the inner precision.
// ---------------------------------------------------------------------------
__device__ inline bool outside(const SpecCoef& sc, double q) {
  return (q - sc.qmin < 0.0) || (q - sc.qmax > 0.0);
}
// diagonal of the local part of H: J' + 2w on the active set
__device__ inline double diag_h(const SpecCoef& sc, double q) {
  return sc.Jp + (outside(sc, q) ? 2.0 * WPEN : 0.0);
}
// Jacobi preconditioner: J + 2w.  J' alone can be negative (J' = J - 2 kc a /
// sqrt(pi) < 0 for Al at Rcut < 7 A); the PME self-potential restores about
// 2 kc a / sqrt(pi) on the diagonal of H, so J approximates H_ii.
__device__ inline double diag_m(const SpecCoef& sc, double q) {
  return sc.J + (outside(sc, q) ? 2.0 * WPEN : 0.0);
}

// y = H p; part[2*block] = sum p.y, part[2*block+1] = sum y
template <class Rg, class Tv>
__global__ void __launch_bounds__(BLOCK) k_hmatvec(Geom g, const Tables* __restrict__ tb,
                                                   const double* __restrict__ x,
                                                   const int* __restrict__ id,
                                                   const double* __restrict__ q,
                                                   const double* __restrict__ p,
                                                   const Rg* __restrict__ grid, double c_pme,
                                                   const int* __restrict__ cnt,
                                                   const int* __restrict__ col,
                                                   const Tv* __restrict__ val,
                                                   double* __restrict__ y,
                                                   double* __restrict__ part) {
  __shared__ double lds[2 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (i < g.n) {
    const double phi = c_pme * gather_phi(g, x + 3 * i, grid);
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double pi = p[i];
    const int c = cnt[i];
    const size_t n = size_t(g.n);
    const Tv s = ell_dot(c, n, i, col, val, p);
    const double yi = diag_h(sc, q[i]) * pi + phi + double(s);
    y[i] = yi;
    acc[0] = pi * yi;
    acc[1] = yi;
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// part[block] = sum (v - sum_v/n)^2 (projected norm without cancellation)
__global__ void __launch_bounds__(BLOCK) k_center_sq(int n, const double* __restrict__ v,
                                                     const double* __restrict__ sum_v,
                                                     double* __restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const double d = v[i] - sum_v[0] / double(n);
    acc[0] = d * d;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// Start of an inner solve: r = g - mean(g) (projected), d = 0, minv = 1/M.
// part[3k..] = sum minv r, sum minv, sum r^2
__global__ void __launch_bounds__(BLOCK) k_pcg_start(int n, const Tables* __restrict__ tb,
                                                     const int* __restrict__ id,
                                                     const double* __restrict__ q,
                                                     const double* __restrict__ g,
                                                     const double* __restrict__ gsum,
                                                     double* __restrict__ r,
                                                     double* __restrict__ d,
                                                     double* __restrict__ minv,
                                                     double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const double m = 1.0 / diag_m(tb->sc[id[i] - 1], q[i]);
    const double ri = g[i] - gsum[0] / double(n);
    r[i] = ri;
    d[i] = 0.0;
    minv[i] = m;
    acc[0] = m * ri; acc[1] = m; acc[2] = ri * ri;
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = acc[k];
}

// Projected preconditioner z = M^-1 r - M^-1 (sum M^-1 r / sum M^-1), so that
// sum z = 0 (total charge conserved); s_minv_r, s_minv: device scalars.
// With first != 0 also p = z.  part[block] = sum r.z
__global__ void __launch_bounds__(BLOCK) k_pcg_z(int n, const double* __restrict__ r,
                                                 const double* __restrict__ minv,
                                                 const double* __restrict__ s_minv_r,
                                                 const double* __restrict__ s_minv, int first,
                                                 double* __restrict__ z, double* __restrict__ p,
                                                 double* __restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const double zi = minv[i] * r[i] - minv[i] * (s_minv_r[0] / s_minv[0]);
    z[i] = zi;
    if (first) p[i] = zi;
    acc[0] = r[i] * zi;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// d += a p, r -= a (Ap - mean Ap) (r stays projected); part[2k..] = sum minv r, sum r^2
__global__ void __launch_bounds__(BLOCK) k_pcg_update(int n, double a,
                                                      const double* __restrict__ p,
                                                      const double* __restrict__ Ap,
                                                      const double* __restrict__ ap_sum,
                                                      const double* __restrict__ minv,
                                                      double* __restrict__ d,
                                                      double* __restrict__ r,
                                                      double* __restrict__ part) {
  __shared__ double lds[2 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (i < n) {
    d[i] += a * p[i];
    const double ri = r[i] - a * (Ap[i] - ap_sum[0] / double(n));
    r[i] = ri;
    acc[0] = minv[i] * ri; acc[1] = ri * ri;
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 2; ++k) part[2 * blockIdx.x + k] = acc[k];
}

// p = z + b p
__global__ void k_pcg_dir(int n, double b, const double* __restrict__ z, double* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + b * p[i];
}

}  // namespace sx

namespace sx {

// ---------------------------------------------------------------------------
// Nonlinear preconditioned CG for the converged charge solve.  U(q) = U0(q) +
// sum_i pen_i(q_i) with U0 quadratic (Hessian H0 = diag(J') + c_pme P + S) and
// pen(q) = w (min(q-qmin,0)^2 + max(q-qmax,0)^2), C1 and piecewise quadratic.
// Along a direction p, with v = H0 p (one matvec),
//   phi'(a) = -g.p + a p.v + sum_i p_i (pen'(q_i + a p_i) - pen'(q_i))
// is piecewise linear and increasing, so the exact line search and the
// gradient update g(q + a p) = g - a v - dpen' need no further matvec.
// ---------------------------------------------------------------------------
__device__ inline double pen_d1(const SpecCoef& sc, double q) {
  double lo = q - sc.qmin, hi = q - sc.qmax;
  if (lo >= 0.0) lo = 0.0;
  if (hi <= 0.0) hi = 0.0;
  return 2.0 * WPEN * (lo + hi);
}

// Potential at the atoms from the real grid: phi_i = c_pme sum w Qfin
// (PHI_MESH's gather, ctip.f90:381-393), as its own pass so that the matvec
// below keeps few registers and many waves in flight.
template <class Rg>
__global__ void __launch_bounds__(BLOCK) k_phi_gather(Geom g, const double* __restrict__ x,
                                                      const Rg* __restrict__ grid, double c_pme,
                                                      double* __restrict__ phi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < g.n) phi[i] = c_pme * gather_phi(g, x + 3 * i, grid);
}

// y = H0 p (no penalty term); part[3k..] = p.y, p.gr, sum p^2 pen''(q).
// The sparse part reads pt, the direction rounded to the operator's type
// (p itself in FP64, the FP32 copy in MIXED: half the gathered bytes).
template <class Rg, class Tv>
__global__ void __launch_bounds__(BLOCK) k_h0matvec(Geom g, const Tables* __restrict__ tb,
                                                    const double* __restrict__ x,
                                                    const int* __restrict__ id,
                                                    const double* __restrict__ q,
                                                    const double* __restrict__ p,
                                                    const Tv* __restrict__ pt,
                                                    const double* __restrict__ gr,
                                                    const double* __restrict__ phi_in,
                                                    const int* __restrict__ cnt,
                                                    const int* __restrict__ col,
                                                    const Tv* __restrict__ val,
                                                    double* __restrict__ y,
                                                    double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < g.n) {
    const double phi = phi_in[i];
    const double pi = p[i];
    const int c = cnt[i];
    const size_t n = size_t(g.n);
    const Tv s = ell_dot(c, n, i, col, val, pt);  // pt = p rounded to Tv
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double yi = sc.Jp * pi + phi + double(s);
    y[i] = yi;
    acc[0] = pi * yi;
    acc[1] = pi * gr[i];
    acc[2] = outside(sc, q[i]) ? pi * pi * 2.0 * WPEN : 0.0;
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = acc[0];
    part[3 * blockIdx.x + 1] = acc[1];
    part[3 * blockIdx.x + 2] = acc[2];
  }
}

// Compacted FP16 copy of the CTIP matrix for the MIXED inner operator when
// its reciprocal part is on the coarse grid: the entries with |H0_ij| >= tau
// (the far real-space Coulomb entries fall below it; tau = 0 keeps all), in
// the same column-major layout with their own row counts.  Round to nearest;
// |H0_ij| stays far below the FP16 range.
__global__ void k_ell_to_f16(int n, int cap, const int* __restrict__ cnt,
                             const int* __restrict__ col, const double* __restrict__ v64,
                             double tau, int* __restrict__ cnt16, int* __restrict__ col16,
                             _Float16* __restrict__ v16) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cnt[i] < cap ? cnt[i] : cap;
  int m = 0;
  for (int k = 0; k < c; ++k) {
    const double v = v64[size_t(k) * n + i];
    if (fabs(v) < tau) continue;
    col16[size_t(m) * n + i] = col[size_t(k) * n + i];
    v16[size_t(m) * n + i] = _Float16(float(v));
    ++m;
  }
  cnt16[i] = m;
}

// The compacted matrix as one 32-bit word per entry when n <= 2^20: the
// column in the low 20 bits, the value in the high 12 as an FP16 number with
// its mantissa rounded to 6 bits (relative error <= 2^-7), so the word's
// bits 31..20 are those of the FP16 value (entries |H0_ij| >= tau stay far
// inside the normal FP16 range).  4 bytes per entry instead of 6.
constexpr int PK_COL_BITS = 20;
constexpr int SLICE = 64;  // rows per slice of the SELL-4 layout (one wave)
__global__ void k_ell_to_p32(int n, int cap, const int* __restrict__ cnt,
                             const int* __restrict__ col, const double* __restrict__ v64,
                             double tau, int* __restrict__ cnt16, unsigned* __restrict__ pk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cnt[i] < cap ? cnt[i] : cap;
  int m = 0;
  for (int k = 0; k < c; ++k) {
    const double v = v64[size_t(k) * n + i];
    if (fabs(v) < tau) continue;
    const unsigned hb = __builtin_bit_cast(unsigned short, _Float16(float(v)));
    const unsigned code = (hb & 0x8000u) | (((hb & 0x7FFFu) + 8u) & 0x7FF0u);
    pk[size_t(m) * n + i] = (code << 16) | unsigned(col[size_t(k) * n + i]);
    ++m;
  }
  cnt16[i] = m;
}
__device__ inline float pk_value(unsigned w) {
  return float(__builtin_bit_cast(_Float16, (unsigned short)((w >> 16) & 0xFFF0u)));
}
__device__ inline int pk_col(unsigned w) { return int(w & ((1u << PK_COL_BITS) - 1u)); }

// y = H0 p as k_h0matvec<float, float> with FP16 matrix values (6 bytes per
// entry instead of 8), FP32 products and sums; PK: the packed 32-bit entries
// of k_ell_to_p32 (pk) instead of col and val.
template <bool PK>
__global__ void __launch_bounds__(BLOCK) k_h0matvec16(Geom g, const Tables* __restrict__ tb,
                                                      const int* __restrict__ id,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ p,
                                                      const float* __restrict__ pt,
                                                      const double* __restrict__ gr,
                                                      const double* __restrict__ phi_in,
                                                      const int* __restrict__ cnt,
                                                      const int* __restrict__ col,
                                                      const _Float16* __restrict__ val,
                                                      const unsigned* __restrict__ pk,
                                                      double* __restrict__ y,
                                                      double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < g.n) {
    const size_t n = size_t(g.n);
    const int c = cnt[i];
    constexpr int U = 16;
    float s = 0.f;
    int k = 0;
    for (; k + U <= c; k += U) {
      int cc[U];
      float vv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if constexpr (PK) {
          const unsigned w = pk[size_t(k + u) * n + i];
          cc[u] = pk_col(w);
          vv[u] = pk_value(w);
        } else {
          cc[u] = col[size_t(k + u) * n + i];
          vv[u] = float(val[size_t(k + u) * n + i]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) s += vv[u] * pt[cc[u]];
    }
    constexpr int UT = 4;
    for (; k < c; k += UT) {  // masked batches of the rest, as ell_dot
      int cc[UT];
      float vv[UT];
#pragma unroll
      for (int u = 0; u < UT; ++u) {
        const bool ok = k + u < c;
        const size_t e = size_t(ok ? k + u : k) * n + i;
        if constexpr (PK) {
          const unsigned w = pk[e];
          cc[u] = ok ? pk_col(w) : i;
          vv[u] = ok ? pk_value(w) : 0.f;
        } else {
          cc[u] = ok ? col[e] : i;
          vv[u] = ok ? float(val[e]) : 0.f;
        }
      }
#pragma unroll
      for (int u = 0; u < UT; ++u) s += vv[u] * pt[cc[u]];
    }
    const double pi = p[i];
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double yi = sc.Jp * pi + phi_in[i] + double(s);
    y[i] = yi;
    acc[0] = pi * yi;
    acc[1] = pi * gr[i];
    acc[2] = outside(sc, q[i]) ? pi * pi * 2.0 * WPEN : 0.0;
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = acc[0];
    part[3 * blockIdx.x + 1] = acc[1];
    part[3 * blockIdx.x + 2] = acc[2];
  }
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// One 16-byte group of the packed matrix; NT: a non-temporal (streaming) load,
// so that the matrix stream does not evict the gathered vector from L2.
template <bool NT>
__device__ inline u32x4 pk_group(const uint4* __restrict__ a) {
  if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(a));
  return *reinterpret_cast<const u32x4*>(a);
}

// k_h0matvec16 over the SELL-4 packed matrix: each lane streams its row as
// 16-byte groups of 4 entries, U groups (4 U entries) in flight.  XCD: the
// row blocks go to the XCDs in contiguous runs (xcd_tile), so that each
// XCD's L2 holds the part of the gathered vector its rows read; the block
// partials stay indexed by row block, so every variant sums in the same order.
template <int U, bool NT, bool XCD = false>
__global__ void __launch_bounds__(BLOCK) k_h0matvec_sell4(Geom g, const Tables* __restrict__ tb,
                                                          const int* __restrict__ id,
                                                          const double* __restrict__ q,
                                                          const double* __restrict__ p,
                                                          const float* __restrict__ pt,
                                                          const double* __restrict__ gr,
                                                          const double* __restrict__ phi_in,
                                                          const int* __restrict__ slen,
                                                          const int* __restrict__ soff,
                                                          const uint4* __restrict__ pk,
                                                          double* __restrict__ y,
                                                          double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int rb = XCD ? xcd_tile(blockIdx.x, gridDim.x) : int(blockIdx.x);
  const int i = rb * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < g.n) {
    const int s = i / SLICE, lane = i % SLICE;
    const int L = slen[s];
    const uint4* row = pk + size_t(soff[s]) + lane;
    float sum = 0.f;
    int k = 0;
    for (; k + U <= L; k += U) {
      u32x4 w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) w[u] = pk_group<NT>(row + size_t(k + u) * SLICE);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        sum += pk_value(w[u].x) * pt[pk_col(w[u].x)];
        sum += pk_value(w[u].y) * pt[pk_col(w[u].y)];
        sum += pk_value(w[u].z) * pt[pk_col(w[u].z)];
        sum += pk_value(w[u].w) * pt[pk_col(w[u].w)];
      }
    }
    for (; k < L; ++k) {
      const u32x4 w = pk_group<NT>(row + size_t(k) * SLICE);
      sum += pk_value(w.x) * pt[pk_col(w.x)];
      sum += pk_value(w.y) * pt[pk_col(w.y)];
      sum += pk_value(w.z) * pt[pk_col(w.z)];
      sum += pk_value(w.w) * pt[pk_col(w.w)];
    }
    const double pi = p[i];
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double yi = sc.Jp * pi + phi_in[i] + double(sum);
    y[i] = yi;
    acc[0] = pi * yi;
    acc[1] = pi * gr[i];
    acc[2] = outside(sc, q[i]) ? pi * pi * 2.0 * WPEN : 0.0;
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[3 * rb] = acc[0];
    part[3 * rb + 1] = acc[1];
    part[3 * rb + 2] = acc[2];
  }
}

// Halo of a row block of the packed matrix: the direction entries its 256
// rows read, staged in LDS by k_h0matvec_sell4h so that the matvec's gathers
// (about 60 per row at the 1M-atom slab) read LDS instead of global memory.
// At that slab every block has one, 1,638 atoms in 80 runs on average, and
// the matvec takes 57.8 us instead of 76.5 (4,096 slots of 16 KB: a 32 KB
// halo halves the workgroups per CU and is slower; gaps up to 32 atoms give
// 2,691 atoms in 10 runs, 59.4 us).
// Each workgroup of k_pk_halo marks the columns of its rows in an LDS bitmap
// over HALO_SPAN atoms from the smallest, fills every gap of fewer than
// HALO_GAP unmarked atoms between two marked ones (so that the staged atoms
// form few runs), ranks the marked atoms (their LDS slots), and writes
//   runs[b][r]  = (first atom, first slot) of run r, in slot order;
//   hcount[b]   = (staged atoms << 10) | runs, or 0 when the block's columns
//                 span HALO_SPAN or more (the blocks at the periodic z
//                 boundary) or its halo exceeds HALO_MAX atoms or HALO_RUNS runs;
//   pkh         = the block's entries with the LDS slot in the column field
//                 (the global column where hcount is 0);
//   pk24        = (blocks with a halo) the same entries in 24 bits each: the
//                 12-bit value code over the 12-bit slot, a group of 4 in
//                 12 bytes (pk24_group), 3 bytes per entry instead of 4.
constexpr int HALO_SPAN = 1 << 17;  // bitmap of 16 KB
constexpr int HALO_MAX = 4096;      // staged atoms (16 KB of LDS)
constexpr int HALO_RUNS = 256;
constexpr int HALO_GAP = 8;         // hi - lo <= HALO_GAP for marked lo < gap < hi

// exclusive block scan of two counters (BLOCK threads), with the totals
__device__ inline void block_scan2(int a, int b, int* red, int& ea, int& eb, int& ta, int& tb) {
  const int t = threadIdx.x, l = t & (WAVE - 1);
  int ia = a, ib = b;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const int va = __shfl_up(ia, o, WAVE), vb = __shfl_up(ib, o, WAVE);
    if (l >= o) { ia += va; ib += vb; }
  }
  if (l == WAVE - 1) { red[2 * (t / WAVE)] = ia; red[2 * (t / WAVE) + 1] = ib; }
  __syncthreads();
  int ba = 0, bb = 0;
  ta = 0; tb = 0;
  for (int w = 0; w < BLOCK / WAVE; ++w) {
    if (w < t / WAVE) { ba += red[2 * w]; bb += red[2 * w + 1]; }
    ta += red[2 * w]; tb += red[2 * w + 1];
  }
  ea = ba + ia - a;
  eb = bb + ib - b;
}

// hmax: the staged-atom limit (at most HALO_MAX; lower values send more
// blocks to the global gathers, which tests use to run both kinds of block).
__global__ void __launch_bounds__(BLOCK) k_pk_halo(int n, int hmax, const int* __restrict__ slen,
                                                   const int* __restrict__ soff,
                                                   const uint4* __restrict__ pk,
                                                   uint4* __restrict__ pkh, unsigned* __restrict__ pk24,
                                                   int2* __restrict__ runs,
                                                   int* __restrict__ hcount) {
  static_assert(HALO_MAX <= 4096, "slots in 12 bits");
  constexpr int NW = HALO_SPAN / 32;  // bitmap words
  constexpr int WPT = NW / BLOCK;     // consecutive words per thread
  __shared__ unsigned bits[NW];
  __shared__ int woff[NW];            // slot of the word's first marked atom
  __shared__ int red[2 * (BLOCK / WAVE)];
  __shared__ int cmin_s, cmax_s;
  const int b = blockIdx.x, t = threadIdx.x;
  const int i = b * BLOCK + t;
  const int sl = i / SLICE, lane = i % SLICE;
  const int L = i < n ? slen[sl] : 0;
  const size_t r0 = (i < n ? size_t(soff[sl]) : 0) + lane;
  // the block's column range (padding entries hold the row's own index)
  int mn = 0x7fffffff, mx = -1;
  for (int k = 0; k < L; ++k) {
    const uint4 w = pk[r0 + size_t(k) * SLICE];
    mn = min(mn, min(min(pk_col(w.x), pk_col(w.y)), min(pk_col(w.z), pk_col(w.w))));
    mx = max(mx, max(max(pk_col(w.x), pk_col(w.y)), max(pk_col(w.z), pk_col(w.w))));
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    mn = min(mn, __shfl_xor(mn, o, WAVE));
    mx = max(mx, __shfl_xor(mx, o, WAVE));
  }
  if (t == 0) { cmin_s = 0x7fffffff; cmax_s = -1; }
#pragma unroll
  for (int k = 0; k < WPT; ++k) bits[t * WPT + k] = 0u;
  __syncthreads();
  if ((t & (WAVE - 1)) == 0) { atomicMin(&cmin_s, mn); atomicMax(&cmax_s, mx); }
  __syncthreads();
  const int c0 = cmin_s;
  const bool span_ok = cmax_s >= 0 && cmax_s - c0 < HALO_SPAN;
  if (span_ok)
    for (int k = 0; k < L; ++k) {
      const uint4 w = pk[r0 + size_t(k) * SLICE];
      const int c[4] = {pk_col(w.x) - c0, pk_col(w.y) - c0, pk_col(w.z) - c0, pk_col(w.w) - c0};
#pragma unroll
      for (int u = 0; u < 4; ++u) atomicOr(&bits[c[u] >> 5], 1u << (c[u] & 31));
    }
  __syncthreads();
  // gap filling from the marked bitmap: an unmarked atom a is filled when the
  // nearest marked atoms lo < a < hi lie within HALO_GAP of each other (both
  // then lie in the words either side of a's)
  unsigned f[WPT];
#pragma unroll
  for (int k = 0; k < WPT; ++k) {
    const int w = t * WPT + k;
    const unsigned m = bits[w];
    f[k] = m;
    if (!span_ok || m == 0u || m == ~0u) continue;  // an empty word is a gap of 32 or more
    const unsigned long long lw = (static_cast<unsigned long long>(m) << 32) | (w > 0 ? bits[w - 1] : 0u);
    const unsigned long long hw =
        (static_cast<unsigned long long>(w + 1 < NW ? bits[w + 1] : 0u) << 32) | m;
    unsigned holes = ~m;
    while (holes) {
      const int bit = __ffs(holes) - 1;
      holes &= holes - 1u;
      const unsigned below = unsigned(lw >> bit);       // bit j: atom a - 32 + j
      const unsigned above = unsigned(hw >> (bit + 1));  // bit j: atom a + 1 + j
      if (below && above && (__clz(below) + 1) + __ffs(above) <= HALO_GAP) f[k] |= 1u << bit;
    }
  }
  __syncthreads();
  int cs = 0, cr = 0;
#pragma unroll
  for (int k = 0; k < WPT; ++k) {
    bits[t * WPT + k] = f[k];
    cs += __popc(f[k]);
  }
  __syncthreads();
  // run starts: a marked atom whose predecessor is unmarked
  unsigned st[WPT];
#pragma unroll
  for (int k = 0; k < WPT; ++k) {
    const int w = t * WPT + k;
    const unsigned prev = w > 0 ? bits[w - 1] >> 31 : 0u;
    st[k] = f[k] & ~((f[k] << 1) | prev);
    cr += __popc(st[k]);
  }
  int es, er, ts, tr;
  block_scan2(cs, cr, red, es, er, ts, tr);
  const bool ok = span_ok && ts <= hmax && tr <= HALO_RUNS;
  if (ok) {
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
      const int w = t * WPT + k;
      woff[w] = es;
      unsigned s = st[k];
      while (s) {
        const int bit = __ffs(s) - 1;
        s &= s - 1u;
        runs[size_t(b) * HALO_RUNS + er] =
            make_int2(c0 + w * 32 + bit, es + __popc(f[k] & ((1u << bit) - 1u)));
        ++er;
      }
      es += __popc(f[k]);
    }
  }
  if (t == 0) hcount[b] = ok ? (ts << 10) | tr : 0;
  __syncthreads();
  for (int k = 0; k < L; ++k) {
    uint4 w = pk[r0 + size_t(k) * SLICE];
    if (ok) {
      unsigned* e = reinterpret_cast<unsigned*>(&w);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int d = pk_col(e[u]) - c0;
        const int slot = woff[d >> 5] + __popc(bits[d >> 5] & ((1u << (d & 31)) - 1u));
        e[u] = (e[u] & ~((1u << PK_COL_BITS) - 1u)) | unsigned(slot);
      }
      // 24-bit entries: value code (bits 31..20 of the packed word) << 12 | slot
      unsigned c24[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) c24[u] = ((e[u] >> PK_COL_BITS) << 12) | (e[u] & 0xFFFu);
      unsigned* g24 = pk24 + 3 * (r0 + size_t(k) * SLICE);
      g24[0] = c24[0] | (c24[1] << 24);
      g24[1] = (c24[1] >> 8) | (c24[2] << 16);
      g24[2] = (c24[2] >> 16) | (c24[3] << 8);
    }
    pkh[r0 + size_t(k) * SLICE] = w;
  }
}

// k_h0matvec_sell4 over the halo-slot copy of k_pk_halo: a block with a halo
// stages its runs of pt in LDS and gathers from there; the others gather from
// global memory as k_h0matvec_sell4 does.  Same products in the same order.
// A group of 4 24-bit entries (12 bytes at 3 g words of pk24) as four
// 32-bit packed words whose column field holds the slot (pk_value, pk_col).
template <bool NT>
__device__ inline u32x4 pk24_group(const unsigned* __restrict__ pk24, size_t g) {
  const unsigned* a = pk24 + 3 * g;
  unsigned w0, w1, w2;
  if constexpr (NT) {
    w0 = __builtin_nontemporal_load(a);
    w1 = __builtin_nontemporal_load(a + 1);
    w2 = __builtin_nontemporal_load(a + 2);
  } else {
    w0 = a[0];
    w1 = a[1];
    w2 = a[2];
  }
  const unsigned e0 = w0 & 0xFFFFFFu, e1 = (w0 >> 24) | ((w1 & 0xFFFFu) << 8);
  const unsigned e2 = (w1 >> 16) | ((w2 & 0xFFu) << 16), e3 = w2 >> 8;
  auto word = [](unsigned e) { return ((e >> 12) << PK_COL_BITS) | (e & 0xFFFu); };
  u32x4 w;
  w.x = word(e0);
  w.y = word(e1);
  w.z = word(e2);
  w.w = word(e3);
  return w;
}

template <int U, bool NT, bool P24 = false>
__global__ void __launch_bounds__(BLOCK) k_h0matvec_sell4h(Geom g, const Tables* __restrict__ tb,
                                                           const int* __restrict__ id,
                                                           const double* __restrict__ q,
                                                           const double* __restrict__ p,
                                                           const float* __restrict__ pt,
                                                           const double* __restrict__ gr,
                                                           const double* __restrict__ phi_in,
                                                           const int* __restrict__ slen,
                                                           const int* __restrict__ soff,
                                                           const uint4* __restrict__ pk,
                                                           const unsigned* __restrict__ pk24,
                                                           const int2* __restrict__ runs,
                                                           const int* __restrict__ hcount,
                                                           double* __restrict__ y,
                                                           double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  __shared__ float hv[HALO_MAX];
  __shared__ int2 rs[HALO_RUNS];
  const int b = blockIdx.x;
  const int hc = hcount[b];
  const int nh = hc >> 10, nr = hc & 1023;
  if (nh > 0) {
    // runs in slot order; slot s lies in the last run starting at or before it
    if (threadIdx.x < nr) rs[threadIdx.x] = runs[size_t(b) * HALO_RUNS + threadIdx.x];
    __syncthreads();
    for (int s = threadIdx.x; s < nh; s += BLOCK) {
      int lo = 0;
#pragma unroll
      for (int step = HALO_RUNS / 2; step > 0; step >>= 1)
        if (lo + step < nr && rs[lo + step].y <= s) lo += step;
      hv[s] = pt[rs[lo].x + (s - rs[lo].y)];
    }
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < g.n) {
    const int s = i / SLICE, lane = i % SLICE;
    const int L = slen[s];
    const size_t g0 = size_t(soff[s]) + lane;
    const uint4* row = pk + g0;
    float sum = 0.f;
    // the group of entry group k of this row: 24-bit entries of the halo copy
    // (P24, blocks with a halo), else the 32-bit words
    auto grp = [&](int kk, bool h24) {
      if constexpr (P24) {
        if (h24) return pk24_group<NT>(pk24, g0 + size_t(kk) * SLICE);
      }
      return pk_group<NT>(row + size_t(kk) * SLICE);
    };
    auto rowsum = [&](auto gat, bool h24) {
      int k = 0;
      for (; k + U <= L; k += U) {
        u32x4 w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = grp(k + u, h24);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          sum += pk_value(w[u].x) * gat(pk_col(w[u].x));
          sum += pk_value(w[u].y) * gat(pk_col(w[u].y));
          sum += pk_value(w[u].z) * gat(pk_col(w[u].z));
          sum += pk_value(w[u].w) * gat(pk_col(w[u].w));
        }
      }
      for (; k < L; ++k) {
        const u32x4 w = grp(k, h24);
        sum += pk_value(w.x) * gat(pk_col(w.x));
        sum += pk_value(w.y) * gat(pk_col(w.y));
        sum += pk_value(w.z) * gat(pk_col(w.z));
        sum += pk_value(w.w) * gat(pk_col(w.w));
      }
    };
    if (nh > 0)
      rowsum([&](int c) { return hv[c]; }, true);
    else
      rowsum([&](int c) { return pt[c]; }, false);
    const double pi = p[i];
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double yi = sc.Jp * pi + phi_in[i] + double(sum);
    y[i] = yi;
    acc[0] = pi * yi;
    acc[1] = pi * gr[i];
    acc[2] = outside(sc, q[i]) ? pi * pi * 2.0 * WPEN : 0.0;
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[3 * b] = acc[0];
    part[3 * b + 1] = acc[1];
    part[3 * b + 2] = acc[2];
  }
}

// charge predictor: out = c0 q + c1 h0 + c2 h1 + c3 h2 with the first ord+1
// terms (polynomial extrapolation in time; the coefficients sum to 1, which
// preserves the total charge)
__global__ void k_extrapolate(int n, int ord, double c0, double c1, double c2, double c3,
                              const double* __restrict__ q, const double* __restrict__ h0,
                              const double* __restrict__ h1, const double* __restrict__ h2,
                              double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = c0 * q[i];
  if (ord >= 1) v += c1 * h0[i];
  if (ord >= 2) v += c2 * h1[i];
  if (ord >= 3) v += c3 * h2[i];
  out[i] = v;
}

// The packed inner matrix in slices of 64 rows (one wave, SELL-4 layout):
// slice s holds 4 * slen[s] entries per row, slen[s] = ceil(max kept entries
// of its rows / 4), group g of lane l at uint4 index soff[s] + 64 g + l, so a
// wave reads 1 KiB per 16-byte load and each lane 4 entries.  Rows shorter
// than the slice are padded with zero values on their own column.
// The kept-entry count of each row (cnt16) comes from k_ctip_matrix's PACK
// pass, which also writes each entry's packed word.

// slen[s] = ceil(max over the slice's rows of cnt16 / 4)
__global__ void k_pk_slices(int n, const int* __restrict__ cnt16, int* __restrict__ slen) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int c = i < n ? cnt16[i] : 0;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) c = max(c, __shfl_xor(c, o, WAVE));
  const int nsl = (n + SLICE - 1) / SLICE;
  if ((threadIdx.x & (WAVE - 1)) == 0 && i / SLICE < nsl) slen[i / SLICE] = (c + 3) / 4;
}

// soff = 64 x exclusive prefix sums of slen (k_scan_block etc. on slen give
// the group offsets; this scales them to uint4 units of 64 lanes)
__global__ void k_pk_scale(int nsl, int* __restrict__ soff) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s <= nsl) soff[s] *= SLICE;
}

// The SELL-4 slices from the packed words k_ctip_matrix<..., PACK> wrote in
// the list's layout (4 bytes per entry read, not the FP64 value and the
// column): each row's kept words in list order, each lane's groups of 4
// stored whole (16 bytes), then padding groups (zero values on the row's own
// column) up to the slice's length.
__global__ void k_ell_to_sell4(int n, const int* __restrict__ cnt,
                               const unsigned* __restrict__ pkw, const int* __restrict__ slen,
                               const int* __restrict__ soff, uint4* __restrict__ pk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = i / SLICE, lane = i % SLICE;
  uint4* base = pk + size_t(soff[s]) + lane;
  const int c = cnt[i];
  const unsigned pad = unsigned(i);
  unsigned a0 = pad, a1 = pad, a2 = pad, a3 = pad;
  int m = 0;
  for (int k = 0; k < c; ++k) {
    const unsigned w = pkw[size_t(k) * n + i];
    if (w == PK_DROP) continue;
    const int r = m & 3;
    a0 = r == 0 ? w : a0;
    a1 = r == 1 ? w : a1;
    a2 = r == 2 ? w : a2;
    a3 = r == 3 ? w : a3;
    ++m;
    if (r == 3) {
      base[size_t(m / 4 - 1) * SLICE] = make_uint4(a0, a1, a2, a3);
      a0 = a1 = a2 = a3 = pad;
    }
  }
  int gdone = m / 4;
  if (m & 3) base[size_t(gdone++) * SLICE] = make_uint4(a0, a1, a2, a3);
  for (; gdone < slen[s]; ++gdone) base[size_t(gdone) * SLICE] = make_uint4(pad, pad, pad, pad);
}

// Line-search terms at step a: part[2k] = sum p (pen'(q + a p) - pen'(q)),
// part[2k+1] = sum p^2 pen''(q + a p)
__global__ void __launch_bounds__(BLOCK) k_ncg_ls(int n, const Tables* __restrict__ tb,
                                                  const int* __restrict__ id,
                                                  const double* __restrict__ q,
                                                  const double* __restrict__ p, double a,
                                                  double* __restrict__ part) {
  __shared__ double lds[2 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (i < n) {
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double pi = p[i], qi = q[i], qa = qi + a * pi;
    acc[0] = pi * (pen_d1(sc, qa) - pen_d1(sc, qi));
    acc[1] = outside(sc, qa) ? pi * pi * 2.0 * WPEN : 0.0;
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// q += a p; g_old = g; g -= a v + (pen'(q_new) - pen'(q_old)); part[k] = sum g
__global__ void __launch_bounds__(BLOCK) k_ncg_step(int n, const Tables* __restrict__ tb,
                                                    const int* __restrict__ id, double a,
                                                    const double* __restrict__ p,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ q,
                                                    double* __restrict__ g,
                                                    double* __restrict__ g_old,
                                                    double* __restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double qi = q[i], qn = qi + a * p[i];
    const double gi = g[i];
    g_old[i] = gi;
    const double gn = gi - a * v[i] - (pen_d1(sc, qn) - pen_d1(sc, qi));
    q[i] = qn;
    g[i] = gn;
    acc[0] = gn;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// Centre g (subtract its mean; the constant part is the neutrality multiplier
// and drops out of every update) and the preconditioner m = 1/(J + pen'').
// part[3k..] = sum m g, sum m, sum g^2 (of the centred g)
__global__ void __launch_bounds__(BLOCK) k_ncg_center(int n, const Tables* __restrict__ tb,
                                                      const int* __restrict__ id,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ gsum,
                                                      double* __restrict__ g,
                                                      double* __restrict__ m,
                                                      double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const double gi = g[i] - gsum[0] / double(n);
    g[i] = gi;
    const double mi = 1.0 / diag_m(tb->sc[id[i] - 1], q[i]);
    m[i] = mi;
    acc[0] = mi * gi; acc[1] = mi; acc[2] = gi * gi;
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = acc[k];
}

// z = m g - m (sum m g / sum m) (sum z = 0); part[2k..] = z.g, z.g_old
__global__ void __launch_bounds__(BLOCK) k_ncg_z(int n, const double* __restrict__ g,
                                                 const double* __restrict__ g_old,
                                                 const double* __restrict__ m,
                                                 const double* __restrict__ s_mg,
                                                 const double* __restrict__ s_m,
                                                 double* __restrict__ z,
                                                 double* __restrict__ part) {
  __shared__ double lds[2 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (i < n) {
    const double zi = m[i] * g[i] - m[i] * (s_mg[0] / s_m[0]);
    z[i] = zi;
    acc[0] = zi * g[i];
    acc[1] = g_old ? zi * g_old[i] : 0.0;
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

}  // namespace sx

namespace sx {

// ---------------------------------------------------------------------------
// Long-range (Woodbury) preconditioner for the charge solve.
// The Hessian is H = D + S + W^T Gc W: D the diagonal J (+2w where the
// penalty is active), S the short-range pair matrix, W the B-spline spread
// onto the PME grid and Gc the reciprocal-space kernel (c_pme times the PME
// influence table).  For a slab of N atoms the W^T Gc W part grows like 1/k^2
// on long wavelengths, so Jacobi CG needs iterations that grow with the box.
// Dropping S, Woodbury gives
//   (D + W^T Gc W)^-1 = D^-1 - D^-1 W^T (Gc^-1 + W D^-1 W^T)^-1 W D^-1,
// and on smooth grid fields W D^-1 W^T acts as the local density rho of 1/J
// per grid point, so the middle factor is the k-space filter
//   F(k) = Gc(k) / (1 + Gc(k) rho).
// For a smooth mode M^-1 ~ (1/J)(1 - rho_true/rho) + 1/(J + Gc rho_true):
// rho below the true density makes M indefinite on long waves (and CG
// stalls), rho above it only weakens the filter.  rho is therefore the
// maximum of the low-passed (sigma = 2 A) density of 1/J on the grid, which
// for a crystalline slab matches the bulk value, times 1.03.
// Applied as M^-1 = D^-1/2 (I - C W^T F W C) D^-1/2 with C = diag(1/sqrt(J)),
// fixed for the solve, so M stays symmetric when the penalty set changes.
// ---------------------------------------------------------------------------
// Filter tables on the half spectrum (same layout as G; values / Ng):
//   stage 0: S(k) = exp(-4 pi^2 m^2 sigma^2)            (Gaussian low-pass)
//   stage 1: F(k) = s / (1 + s rho),  s = c_pme Ng G(k)  (G holds psi B / Ng)
template <class R>
__global__ void k_filter_table(Geom g, const double* __restrict__ G64, double sigma2,
                               double c_pme, double rho, int stage, R* __restrict__ F) {
  const size_t total = size_t(g.H) * g.K[1] * g.K[2];
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const double ng = double(g.K[0]) * g.K[1] * g.K[2];
  double v;
  if (stage == 0) {
    const int kz = int(t % g.K[2]);
    const int ky = int((t / g.K[2]) % g.K[1]);
    const int kx = int(t / (size_t(g.K[2]) * g.K[1]));
    const int k3[3] = {kx, ky, kz};
    double m2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int K = g.K[d];
      const int m = (k3[d] + K / 2) % K - K / 2;
      m2 += double(m * m) / (g.box[d] * g.box[d]);
    }
    v = exp(-4.0 * PI * PI * m2 * sigma2);
  } else {
    const double sk = c_pme * G64[t] * ng;
    v = sk / (1.0 + sk * rho);
  }
  F[t] = R(v / ng);
}

// part[block] = max of grid values over the block's span
template <class R>
__global__ void __launch_bounds__(BLOCK) k_grid_max(size_t ng, const R* __restrict__ grid,
                                                    double* __restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  double m = -1e300;
  for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < ng;
       t += size_t(gridDim.x) * blockDim.x)
    m = fmax(m, double(grid[t]));
  m = wave_max(m);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  if (lane == 0) lds[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = lds[0];
    for (int q = 1; q < int(blockDim.x / WAVE); ++q) t = fmax(t, lds[q]);
    part[blockIdx.x] = t;
  }
}

__global__ void k_fill_f64(int n, double v, double* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// w_i = 1/J_i (the density spread to find rho), c_i = 1/sqrt(J_i)
__global__ void k_prec_weights(int n, const Tables* __restrict__ tb, const int* __restrict__ id,
                               double* __restrict__ w, double* __restrict__ c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double J = tb->sc[id[i] - 1].J;
    w[i] = 1.0 / J;
    c[i] = 1.0 / sqrt(J);
  }
}

// s = c g / sqrt(D)
__global__ void k_prec_scale(int n, const Tables* __restrict__ tb, const int* __restrict__ id,
                             const double* __restrict__ q, const double* __restrict__ g,
                             const double* __restrict__ c, double* __restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = c[i] * g[i] / sqrt(diag_m(tb->sc[id[i] - 1], q[i]));
}

// z = (g/sqrt(D) - c W^T grid)/sqrt(D); part[k] = sum z
template <int ORD, class Rg>
__global__ void __launch_bounds__(BLOCK) k_prec_finish(Geom g, const Tables* __restrict__ tb,
                                                       const double* __restrict__ x,
                                                       const int* __restrict__ id,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ gr,
                                                       const double* __restrict__ c,
                                                       const Rg* __restrict__ grid,
                                                       double* __restrict__ z,
                                                       double* __restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < g.n) {
    const double sd = sqrt(diag_m(tb->sc[id[i] - 1], q[i]));
    const double zi = (gr[i] / sd - c[i] * gather_ord<ORD>(g, x + 3 * i, grid)) / sd;
    z[i] = zi;
    acc[0] = zi;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// Fused tail of an inner iteration with the long-range preconditioner
// (multi-block path): k_center_scale centres g (its mean, gsum / n, is the
// neutrality multiplier's part) and scales it for the preconditioner's
// spread, s = c g / sqrt(D); k_prec_finish6 forms z and the six sums
// {z, z.g, z.g_old, g, g_old, g^2} from which k_reduce_op<4> derives the
// projected z' = z - mean(z): z'.g = z.g - mean(z) sum g, likewise z'.g_old,
// then beta and rz; k_pcg_dir_proj applies the projection with the direction
// update.  The same arithmetic as centre, scale, finish, project and beta in
// five launches fewer.
__global__ void k_center_scale(int n, const Tables* __restrict__ tb, const int* __restrict__ id,
                               const double* __restrict__ q, const double* __restrict__ gsum,
                               double* __restrict__ g, const double* __restrict__ c,
                               double* __restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double gi = g[i] - gsum[0] / double(n);
    g[i] = gi;
    s[i] = c[i] * gi / sqrt(diag_m(tb->sc[id[i] - 1], q[i]));
  }
}

// z = (g/sqrt(D) - c W^T grid)/sqrt(D); part[6k..] = sums of z, z g, z g_old,
// g, g_old, g^2
template <int ORD, class Rg>
__global__ void __launch_bounds__(BLOCK) k_prec_finish6(Geom g, const Tables* __restrict__ tb,
                                                        const double* __restrict__ x,
                                                        const int* __restrict__ id,
                                                        const double* __restrict__ q,
                                                        const double* __restrict__ gr,
                                                        const double* __restrict__ g_old,
                                                        const double* __restrict__ c,
                                                        const Rg* __restrict__ grid,
                                                        double* __restrict__ z,
                                                        double* __restrict__ part, RedTail rt) {
  __shared__ double lds[6 * BLOCK / WAVE];
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += gridDim.x * blockDim.x) {
    const double sd = sqrt(diag_m(tb->sc[id[i] - 1], q[i]));
    const double gi = gr[i], go = g_old[i];
    const double zi = (gi / sd - c[i] * gather_ord<ORD>(g, x + 3 * i, grid)) / sd;
    z[i] = zi;
    acc[0] += zi;
    acc[1] += zi * gi;
    acc[2] += zi * go;
    acc[3] += gi;
    acc[4] += go;
    acc[5] += gi * gi;
  }
  block_sum<6>(acc, lds);
  partials_out<6>(acc, part, rt, lds);
}

// k_prec_finish6<2, float> with the preconditioner grid's atom-order stencil
// records: the same FP32 weights as its 32-bit spread (the gather is the
// spread's transpose), no FP64 grid coordinates per atom.
__global__ void __launch_bounds__(BLOCK) k_prec_finish6_rec(Geom g, const Tables* __restrict__ tb,
                                                            const float4* __restrict__ rec_atom,
                                                            const int* __restrict__ id,
                                                            const double* __restrict__ q,
                                                            const double* __restrict__ gr,
                                                            const double* __restrict__ g_old,
                                                            const double* __restrict__ c,
                                                            const float* __restrict__ grid,
                                                            double* __restrict__ z,
                                                            double* __restrict__ part, RedTail rt) {
  __shared__ double lds[6 * BLOCK / WAVE];
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += gridDim.x * blockDim.x) {
    const double sd = sqrt(diag_m(tb->sc[id[i] - 1], q[i]));
    const double gi = gr[i], go = g_old[i];
    const double zi = (gi / sd - c[i] * gather_cic_rec(g, rec_atom[i], grid)) / sd;
    z[i] = zi;
    acc[0] += zi;
    acc[1] += zi * gi;
    acc[2] += zi * go;
    acc[3] += gi;
    acc[4] += go;
    acc[5] += gi * gi;
  }
  block_sum<6>(acc, lds);
  partials_out<6>(acc, part, rt, lds);
}

// p = (z - mean z) + beta p (mean z in ls[11], beta in ls[7])
__global__ void k_pcg_dir_proj(int n, const double* __restrict__ ls, const double* __restrict__ z,
                               double* __restrict__ p, float* __restrict__ p32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double v = (z[i] - ls[11]) + ls[7] * p[i];
    p[i] = v;
    if (p32) p32[i] = float(v);
  }
}

// Neutrality: z -= mean(z) (P z with P the projector on sum = 0; P M P is
// symmetric positive definite on that subspace).  part[2k..] = z.g, z.g_old
__global__ void __launch_bounds__(BLOCK) k_prec_project(int n, const double* __restrict__ zsum,
                                                        const double* __restrict__ gr,
                                                        const double* __restrict__ g_old,
                                                        double* __restrict__ z,
                                                        double* __restrict__ part) {
  __shared__ double lds[2 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (i < n) {
    const double zi = z[i] - zsum[0] / double(n);
    z[i] = zi;
    acc[0] = zi * gr[i];
    acc[1] = g_old ? zi * g_old[i] : 0.0;
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

}  // namespace sx

namespace sx {

// Device-side scalar logic of one nonlinear-CG iteration (no host round trip).
// ls layout: [0] g.p  [1] p.H0p  [2] a  [3] lo  [4] hi  [5] done  [6] rz
//            [7] beta  [8] descent-failure flag  [9] Newton rounds run
// red: reduction outputs; red[0] = p.H0p, red[1] = p.g, red[2] = sum p^2 pen''(q)
__device__ inline void ls_init(const double* red, double* ls) {
  const double pHp = red[0], gp = red[1], s2 = red[2];
  ls[0] = gp;
  ls[1] = pHp;
  ls[3] = 0.0;
  ls[4] = -1.0;
  const bool bad = !(pHp > 0.0) || !(gp > 0.0);
  ls[8] = bad ? 1.0 : 0.0;
  ls[5] = bad ? 1.0 : 0.0;
  ls[2] = bad ? 0.0 : gp / (pHp + s2);
  ls[9] = 0.0;
}

__global__ void k_ls_init(const double* __restrict__ red, double* __restrict__ ls) { ls_init(red, ls); }

// Newton step on phi'(a) = -gp + a pHp + D1(a), bracketed; red = {D1(a), S2(a)}
// bounds: red[2], red[3] = the nearest steps above and below a at which an
// atom's charge crosses a bound (k_ncg_ls_dev, each pulled towards a by its
// rounding margin).  When a and the Newton step an both lie strictly between
// them, no atom changes sides on [a, an], D1 is linear there with slope S2(a),
// and phi'(an) = phi'(a) + (an - a)(pHp + S2(a)) = 0: the search ends at an
// without the round that would evaluate it.
__device__ inline void ls_update(const double* red, double* ls, bool bounds) {
  if (ls[5] != 0.0) return;
  ls[9] += 1.0;
  const double gp = ls[0], pHp = ls[1], a = ls[2];
  const double fprime = -gp + a * pHp + red[0];
  if (fabs(fprime) <= 1e-13 * (gp + fabs(a * pHp))) { ls[5] = 1.0; return; }
  double lo = ls[3], hi = ls[4];
  if (fprime < 0.0) lo = a; else hi = a;
  double an = a - fprime / (pHp + red[1]);
  const bool newton = !(an <= lo || (hi >= 0.0 && an >= hi));
  if (!newton) an = hi >= 0.0 ? 0.5 * (lo + hi) : 2.0 * a;
  if (fabs(an - a) <= 1e-15 * fabs(a)) ls[5] = 1.0;
  if (bounds && newton && red[3] < a && a < red[2] && red[3] < an && an < red[2]) ls[5] = 1.0;
  ls[2] = an;
  ls[3] = lo;
  ls[4] = hi;
}

__global__ void k_ls_update(const double* __restrict__ red, double* __restrict__ ls) { ls_update(red, ls); }

// beta = max(0, (z.g - z.g_old)/rz); rz = z.g; red = {z.g, z.g_old}
__device__ inline void beta_update(const double* red, double* ls) {
  const double b = (red[0] - red[1]) / ls[6];
  ls[7] = b > 0.0 ? b : 0.0;
  ls[6] = red[0];
}
__global__ void k_beta(const double* __restrict__ red, double* __restrict__ ls) { beta_update(red, ls); }

// line-search terms with the step read from device memory (skipped once done)
__global__ void __launch_bounds__(BLOCK) k_ncg_ls_dev(int n, const Tables* __restrict__ tb,
                                                      const int* __restrict__ id,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ p,
                                                      const double* __restrict__ ls,
                                                      double* __restrict__ part, RedTail rt) {
  __shared__ double lds[2 * BLOCK / WAVE];
  if (ls[5] != 0.0) return;  // uniform: the search is done
  const double a = ls[2];
  double acc[2] = {0.0, 0.0};
  double up = 1e300, dn = -1e300;  // nearest crossings above and below a
  auto term = [&](int idi, double qi, double pi) {
    const SpecCoef& sc = tb->sc[idi - 1];
    const double qa = qi + a * pi;
    acc[0] += pi * (pen_d1(sc, qa) - pen_d1(sc, qi));
    acc[1] += outside(sc, qa) ? pi * pi * 2.0 * WPEN : 0.0;
    if (pi != 0.0) {
      // q + t p crosses qmin and qmax at these t; each pulled towards a by
      // far more than the rounding of q + t p (|t p| stays below the bounds'
      // scale on the interval)
      const double m = 1e-14 * (fabs(qi) + fabs(sc.qmin) + fabs(sc.qmax)) / fabs(pi);
      const double b1 = (sc.qmin - qi) / pi, b2 = (sc.qmax - qi) / pi;
      if (b1 > a) up = fmin(up, b1 - m); else dn = fmax(dn, b1 + m);
      if (b2 > a) up = fmin(up, b2 - m); else dn = fmax(dn, b2 + m);
    }
  };
  // grid-stride, four atoms' loads in flight per trip, summed in atom order
  const int S = gridDim.x * blockDim.x;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * S < n; i += 4 * S) {
    int iv[4];
    double qv[4], pv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      iv[u] = id[i + u * S];
      qv[u] = q[i + u * S];
      pv[u] = p[i + u * S];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) term(iv[u], qv[u], pv[u]);
  }
  for (; i < n; i += S) term(id[i], q[i], p[i]);
  block_sum<2>(acc, lds);
  // the block's nearest crossings (min, max: order-independent)
  __shared__ double bnd[2 * BLOCK / WAVE];
  up = -wave_max(-up);
  dn = wave_max(dn);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    bnd[2 * (threadIdx.x / WAVE)] = up;
    bnd[2 * (threadIdx.x / WAVE) + 1] = dn;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < BLOCK / WAVE; ++w) {
      up = fmin(up, bnd[2 * w]);
      dn = fmax(dn, bnd[2 * w + 1]);
    }
  const double row[4] = {acc[0], acc[1], up, dn};
  partials_out_ls(row, part, rt, lds);
}

// The whole line search in one workgroup (systems of up to LS_SMALL_MAX atoms;
// at 78 k atoms one workgroup is already slower than the two-launch rounds):
// up to 8 Newton rounds, each a fixed-order sum over the atoms by the 1024
// threads and the scalar update, instead of two launches per round.  red
// receives the last round's sums (as k_reduce_op<2> leaves them).
constexpr int LS_SMALL_MAX = 16384;
__global__ void __launch_bounds__(1024) k_ncg_ls_small(int n, const Tables* __restrict__ tb,
                                                       const int* __restrict__ id,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ p,
                                                       double* __restrict__ ls,
                                                       double* __restrict__ red) {
  __shared__ double lds[4 * 1024 / WAVE];
  __shared__ double st[10];  // ls[0..9]
  __shared__ double sums[4];
  if (threadIdx.x < 10) st[threadIdx.x] = ls[threadIdx.x];
  if (threadIdx.x < 2) sums[threadIdx.x] = red[threadIdx.x];
  __syncthreads();
  for (int round = 0; round < 8; ++round) {
    if (st[5] != 0.0) break;  // uniform: read after the barrier
    const double a = st[2];
    double acc[2] = {0.0, 0.0};
    double up = 1e300, dn = -1e300;  // nearest bound crossings, as k_ncg_ls_dev
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const SpecCoef& sc = tb->sc[id[i] - 1];
      const double pi = p[i], qi = q[i], qa = qi + a * pi;
      acc[0] += pi * (pen_d1(sc, qa) - pen_d1(sc, qi));
      acc[1] += outside(sc, qa) ? pi * pi * 2.0 * WPEN : 0.0;
      if (pi != 0.0) {
        const double m = 1e-14 * (fabs(qi) + fabs(sc.qmin) + fabs(sc.qmax)) / fabs(pi);
        const double b1 = (sc.qmin - qi) / pi, b2 = (sc.qmax - qi) / pi;
        if (b1 > a) up = fmin(up, b1 - m); else dn = fmax(dn, b1 + m);
        if (b2 > a) up = fmin(up, b2 - m); else dn = fmax(dn, b2 + m);
      }
    }
    constexpr int NW = 1024 / WAVE;
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const double w = wave_sum(acc[v]);
      if ((threadIdx.x & (WAVE - 1)) == 0) lds[v * NW + threadIdx.x / WAVE] = w;
    }
    up = -wave_max(-up);
    dn = wave_max(dn);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
      lds[2 * NW + threadIdx.x / WAVE] = up;
      lds[3 * NW + threadIdx.x / WAVE] = dn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int nw = int(blockDim.x / WAVE);
      for (int v = 0; v < 2; ++v) {
        double t = lds[v * NW];
        for (int k = 1; k < nw; ++k) t += lds[v * NW + k];
        sums[v] = t;
      }
      double u = lds[2 * NW], d = lds[3 * NW];
      for (int k = 1; k < nw; ++k) {
        u = fmin(u, lds[2 * NW + k]);
        d = fmax(d, lds[3 * NW + k]);
      }
      sums[2] = u;
      sums[3] = d;
      ls_update(sums, st, true);
    }
    __syncthreads();
  }
  if (threadIdx.x < 10) ls[threadIdx.x] = st[threadIdx.x];
  if (threadIdx.x < 2) red[threadIdx.x] = sums[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Small systems (n <= LS_SMALL_MAX): the vector work of an iteration after the
// line search in one 1024-thread workgroup each, with the global sums formed
// inside (fixed order), instead of a kernel and a reduction launch per step.
// Launch latency, not bandwidth, bounds these systems.
// ---------------------------------------------------------------------------
template <int NV>
__device__ inline void blk_allsum(double (&v)[NV], double* lds, double* tot) {
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = v[k];
  __syncthreads();
}

// Step q += a p, g -= a H0p + dpen', centring of g and m = 1/(J + pen'')
// (k_ncg_step_dev, k_ncg_center and their reductions); then, JACOBI: z, the
// PR+ beta and the new direction (k_ncg_z, beta_update, k_pcg_dir_dev);
// else s = c g / sqrt(D) for the long-range preconditioner (k_prec_scale).
// scal: the context's scalar slots ([0] sum g, [16..18] sum m g, sum m,
// sum g^2, [19..20] z.g, z.g_old).
template <bool JACOBI>
__global__ void __launch_bounds__(1024) k_small_step(int n, const Tables* __restrict__ tb,
                                                     const int* __restrict__ id,
                                                     double* __restrict__ ls,
                                                     const double* __restrict__ v,
                                                     const double* __restrict__ c,
                                                     double* __restrict__ q,
                                                     double* __restrict__ g,
                                                     double* __restrict__ g_old,
                                                     double* __restrict__ m,
                                                     double* __restrict__ z,
                                                     double* __restrict__ p,
                                                     float* __restrict__ p32,
                                                     double* __restrict__ s_out,
                                                     double* __restrict__ scal) {
  __shared__ double lds[3 * 1024 / WAVE];
  __shared__ double tot[3];
  const double a = ls[8] != 0.0 ? 0.0 : ls[2];
  double s1[1] = {0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const SpecCoef& sc = tb->sc[id[i] - 1];
    const double qi = q[i], qn = qi + a * p[i];
    const double gi = g[i];
    g_old[i] = gi;
    const double gn = gi - a * v[i] - (pen_d1(sc, qn) - pen_d1(sc, qi));
    q[i] = qn;
    g[i] = gn;
    s1[0] += gn;
  }
  blk_allsum<1>(s1, lds, tot);
  const double gmean = tot[0] / double(n);
  if (threadIdx.x == 0) scal[0] = tot[0];
  double s3[3] = {0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double gi = g[i] - gmean;
    g[i] = gi;
    const double mi = 1.0 / diag_m(tb->sc[id[i] - 1], q[i]);
    m[i] = mi;
    s3[0] += mi * gi; s3[1] += mi; s3[2] += gi * gi;
  }
  __syncthreads();
  blk_allsum<3>(s3, lds, tot);
  if (threadIdx.x == 0) { scal[16] = tot[0]; scal[17] = tot[1]; scal[18] = tot[2]; }
  if constexpr (JACOBI) {
    const double ratio = tot[0] / tot[1];
    double s2[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const double zi = m[i] * g[i] - m[i] * ratio;
      z[i] = zi;
      s2[0] += zi * g[i];
      s2[1] += zi * g_old[i];
    }
    __syncthreads();
    blk_allsum<2>(s2, lds, tot);
    if (threadIdx.x == 0) {
      scal[19] = tot[0];
      scal[20] = tot[1];
      beta_update(tot, ls);
      tot[2] = ls[7];
    }
    __syncthreads();
    const double beta = tot[2];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const double pv = z[i] + beta * p[i];
      p[i] = pv;
      if (p32) p32[i] = float(pv);
    }
  } else {
    for (int i = threadIdx.x; i < n; i += blockDim.x)
      s_out[i] = c[i] * g[i] / sqrt(diag_m(tb->sc[id[i] - 1], q[i]));
  }
}

// Long-range preconditioner's tail for small systems: z = (g/sqrt(D) - c W^T
// grid)/sqrt(D) (k_prec_finish), z -= mean z (k_prec_project), the PR+ beta
// and the new direction, in one workgroup.
template <int ORD, class Rg>
__global__ void __launch_bounds__(1024) k_small_prec_tail(Geom gp, const Tables* __restrict__ tb,
                                                          const double* __restrict__ x,
                                                          const int* __restrict__ id,
                                                          const double* __restrict__ q,
                                                          const double* __restrict__ g,
                                                          const double* __restrict__ g_old,
                                                          const double* __restrict__ c,
                                                          const Rg* __restrict__ grid,
                                                          double* __restrict__ ls,
                                                          double* __restrict__ z,
                                                          double* __restrict__ p,
                                                          float* __restrict__ p32,
                                                          double* __restrict__ scal) {
  __shared__ double lds[2 * 1024 / WAVE];
  __shared__ double tot[3];
  const int n = gp.n;
  double s1[1] = {0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double sd = sqrt(diag_m(tb->sc[id[i] - 1], q[i]));
    const double zi = (g[i] / sd - c[i] * gather_ord<ORD>(gp, x + 3 * i, grid)) / sd;
    z[i] = zi;
    s1[0] += zi;
  }
  blk_allsum<1>(s1, lds, tot);
  if (threadIdx.x == 0) scal[26] = tot[0];
  const double zmean = tot[0] / double(n);
  double s2[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double zi = z[i] - zmean;
    z[i] = zi;
    s2[0] += zi * g[i];
    s2[1] += zi * g_old[i];
  }
  __syncthreads();
  blk_allsum<2>(s2, lds, tot);
  if (threadIdx.x == 0) {
    scal[19] = tot[0];
    scal[20] = tot[1];
    beta_update(tot, ls);
    tot[2] = ls[7];
  }
  __syncthreads();
  const double beta = tot[2];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double pv = z[i] + beta * p[i];
    p[i] = pv;
    if (p32) p32[i] = float(pv);
  }
}

// Final reduction (as k_reduce_final, one block, fixed order) followed by a
// scalar step of the device-side CG on the sums, in the same launch:
// OP 1 = line-search start (ls_init), 2 = Newton update of the line search
// (ls_update; nothing once ls[5] marks it done), 3 = PR+ beta (beta_update,
// when ls != null), 4 = projected beta from k_prec_finish6's sums (mean z to
// ls[11], sum g^2 to ls[10]; ls[12] holds the atom count), 0 = none.
template <int OP>
__global__ void k_reduce_op(const double* __restrict__ part, int nparts, int nv,
                            double* __restrict__ out, double* __restrict__ ls) {
  constexpr int NW = 1024 / WAVE;
  __shared__ double lds[RED_MAXV * NW];
  if (OP == 2 && ls[5] != 0.0) return;
  if (nv > RED_MAXV) return;
  double s[RED_MAXV];
#pragma unroll
  for (int v = 0; v < RED_MAXV; ++v) s[v] = 0.0;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
#pragma unroll
    for (int v = 0; v < RED_MAXV; ++v) {
      if (v >= nv) break;
      s[v] += part[size_t(nv) * k + v];
    }
  }
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
  for (int v = 0; v < RED_MAXV; ++v) {
    if (v >= nv) break;
    const double t = wave_sum(s[v]);
    if (lane == 0) lds[v * NW + w] = t;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    const int v = threadIdx.x;
    const int nw = int(blockDim.x / WAVE);
    double t = lds[v * NW];
    for (int q = 1; q < nw; ++q) t += lds[v * NW + q];
    out[v] = t;
  }
  __syncthreads();  // out[] complete before the scalar step below reads it
  if (threadIdx.x == 0 && ls) reduce_scalar<OP>(out, ls);
}

template <int OP>
__device__ inline void reduce_scalar(const double* out, double* ls) {
  if (OP == 1) ls_init(out, ls);
  if (OP == 2) ls_update(out, ls);
  if (OP == 3) beta_update(out, ls);
  if (OP == 4) {  // out = the six sums of k_prec_finish6 (n = nparts' atom count in ls[12])
    const double zbar = out[0] / ls[12];
    const double red[2] = {out[1] - zbar * out[3], out[2] - zbar * out[4]};
    beta_update(red, ls);
    ls[10] = out[5];
    ls[11] = zbar;
  }
}

__global__ void k_reduce_if(const double* __restrict__ part, int nparts, int nv,
                            const double* __restrict__ ls, double* __restrict__ out) {
  __shared__ double lds[1024 / WAVE];
  if (ls[5] != 0.0) return;
  for (int v = 0; v < nv; ++v) {
    double s = 0.0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) s += part[size_t(nv) * k + v];
    s = wave_sum(s);
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    if (lane == 0) lds[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = lds[0];
      for (int q = 1; q < int(blockDim.x / WAVE); ++q) t += lds[q];
      out[v] = t;
    }
    __syncthreads();
  }
}

// q += a p; g_old = g; g -= a v + dpen' with a from device memory; part = sum g
__global__ void __launch_bounds__(RED_NT) k_ncg_step_dev(int n, const Tables* __restrict__ tb,
                                                        const int* __restrict__ id,
                                                        const double* __restrict__ ls,
                                                        const double* __restrict__ p,
                                                        const double* __restrict__ v,
                                                        double* __restrict__ q,
                                                        double* __restrict__ g,
                                                        double* __restrict__ g_old,
                                                        double* __restrict__ part, RedTail rt) {
  __shared__ double lds[RED_NT / WAVE];
  const double a = ls[8] != 0.0 ? 0.0 : ls[2];
  double acc[1] = {0.0};
  auto step = [&](int i, int idi, double qi, double pi, double gi, double vi) {
    const SpecCoef& sc = tb->sc[idi - 1];
    const double qn = qi + a * pi;
    g_old[i] = gi;
    const double gn = gi - a * vi - (pen_d1(sc, qn) - pen_d1(sc, qi));
    q[i] = qn;
    g[i] = gn;
    acc[0] += gn;
  };
  // grid-stride, four atoms' loads in flight per trip, summed in atom order
  const int S = gridDim.x * blockDim.x;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * S < n; i += 4 * S) {
    int iv[4];
    double qv[4], pv[4], gv[4], vv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      iv[u] = id[i + u * S];
      qv[u] = q[i + u * S];
      pv[u] = p[i + u * S];
      gv[u] = g[i + u * S];
      vv[u] = v[i + u * S];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) step(i + u * S, iv[u], qv[u], pv[u], gv[u], vv[u]);
  }
  for (; i < n; i += S) step(i, id[i], q[i], p[i], g[i], v[i]);
  block_sum<1>(acc, lds);
  partials_out<1>(acc, part, rt, lds);
}

__global__ void k_pcg_dir_dev(int n, const double* __restrict__ ls, const double* __restrict__ z,
                              double* __restrict__ p, float* __restrict__ p32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double v = z[i] + ls[7] * p[i];
    p[i] = v;
    if (p32) p32[i] = float(v);
  }
}
// p = z (restart of the direction), with its FP32 copy when p32 != null
__global__ void k_set_dir(int n, const double* __restrict__ z, double* __restrict__ p,
                          float* __restrict__ p32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    p[i] = z[i];
    if (p32) p32[i] = float(z[i]);
  }
}

}  // namespace sx
