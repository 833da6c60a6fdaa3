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
    Tv s = Tv(0);
    for (int k = 0; k < c; ++k) s += val[size_t(k) * n + i] * Tv(ql[col[size_t(k) * n + i]]);
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
// part[nv*k + v] (single block, fixed order); MAXV marks values reduced by max.
__global__ void k_reduce_final(const double* __restrict__ part, int nparts, int nv, int maxmask,
                               double* __restrict__ out) {
  __shared__ double lds[1024 / WAVE];
  for (int v = 0; v < nv; ++v) {
    const bool mx = (maxmask >> v) & 1;
    double s = mx ? 0.0 : 0.0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
      const double e = part[size_t(nv) * k + v];
      s = mx ? fmax(s, e) : s + e;
    }
    s = mx ? wave_max(s) : wave_sum(s);
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    if (lane == 0) lds[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = lds[0];
      for (int q = 1; q < int(blockDim.x / WAVE); ++q) t = mx ? fmax(t, lds[q]) : t + lds[q];
      out[v] = t;
    }
    __syncthreads();
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
    Tv s = Tv(0);
    for (int k = 0; k < c; ++k) s += val[size_t(k) * n + i] * Tv(p[col[size_t(k) * n + i]]);
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
