// Parameter gradients of the CTIP energy through the charge equilibration
// (build-only; the reference has no derivatives).
//
// E(theta) = Eself + Edir + Erec + Ectip at the converged charges q*(theta),
// theta = (chi_s, J_s) per species.  q* satisfies G(q*, theta) = mu 1,
// sum q* = Q, with G = dU/dq the solver's gradient (ctip.f90:307-333).  With
// the adjoint lambda (sum lambda = 0) of H lambda = P dE/dq (H = dG/dq on the
// active set, P the projection on sum = 0):
//   dE/dtheta = dE/dtheta|_q - lambda . dG/dtheta,
//   dG_i/dchi_s = [s_i = s],  dG_i/dJ_s = q_i [s_i = s].
// dE/dq follows the energies the MD step reports: Eself of CTIP_SOLVER's end
// (ctip.f90:194-209: chi q + J' q^2 / 2 - kc a / sqrt(pi) q^2 + w pen), Edir
// and Ectip of FORCE (force.f90:94-140, with its same-species Slater form, Q2)
// and Erec of PME, whose q-gradient is the potential phi.
#pragma once
#include "common.hpp"
#include "pair.hpp"
#include "solve.hpp"

namespace sx {

// d/dq_i of Ectip / kc for the pair (i, j) of FORCE's CTIP terms
// (derivative of e = q_i Z_j (jfi - fifj) + q_j Z_i (ifj - fifj) + q_i q_j fifj).
__device__ inline double ctip_energy_dqi(const PairCoef& pc, double r, double r2, double qj) {
  const double xi_i = pc.xi_i, xi_j = pc.xi_j, Zj = pc.Z_j;
  double jfi, fifj;
  const double ir = 1.0 / r;
  if (pc.same) {
    const double ei = exp(-2.0 * xi_i * r);
    jfi = -ei * (xi_i + ir);
    fifj = -ei * (1.0 + pc.c[1] * r + pc.c[2] * r2 + pc.c[3] * r2 * r);
  } else {
    const double ei = exp(-2.0 * xi_i * r), ej = exp(-2.0 * xi_j * r);
    jfi = -ei * (xi_i + ir);
    fifj = -pc.c[0] * ei - pc.c[1] * ej - pc.c[2] * ei / r - pc.c[3] * ej / r;
  }
  return Zj * (jfi - fifj) + qj * fifj;
}

// out_i = dE/dq_i with phi_i (PME potential of q, c_pme scaled) given.
__global__ void __launch_bounds__(BLOCK) k_denergy_dq(Geom g, const Tables* __restrict__ tb,
                                                      const double* __restrict__ x,
                                                      const int* __restrict__ id,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ phi,
                                                      const int* __restrict__ cnt,
                                                      const int* __restrict__ col,
                                                      double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int si = id[i];
  const double qi = q[i];
  const size_t n = size_t(g.n);
  double s = 0.0;
  for (int k = 0; k < cnt[i]; ++k) {
    const int j = col[size_t(k) * n + i];
    double rx[3], r2;
    pair_sep(g, x, i, j, rx, r2);
    const PairCoef& pc = tb->pc[pidx(si, id[j])];
    const double r = sqrt(r2), qj = q[j];
    s += qj * erfc(g.alpha * r) / r + ctip_energy_dqi(pc, r, r2, qj);
  }
  const SpecCoef& sc = tb->sc[si - 1];
  const double self = sc.chi + sc.Jp * qi - 2.0 * (KC * g.alpha / SQRTPI) * qi + pen_d1(sc, qi);
  out[i] = KC * s + self + phi[i];
}

// y += 2w p on the active set (the penalty part of H); part = p.y partials
__global__ void __launch_bounds__(BLOCK) k_add_penalty(int n, const Tables* __restrict__ tb,
                                                       const int* __restrict__ id,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ p,
                                                       double* __restrict__ y,
                                                       double* __restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const SpecCoef& sc = tb->sc[id[i] - 1];
    double yi = y[i];
    if (outside(sc, q[i])) yi += 2.0 * WPEN * p[i];
    y[i] = yi;
    acc[0] = p[i] * yi;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// part[3k..] = sum v, sum m, sum m v with m = 1 / diag_m (Jacobi weights)
__global__ void __launch_bounds__(BLOCK) k_adj_sums(int n, const Tables* __restrict__ tb,
                                                    const int* __restrict__ id,
                                                    const double* __restrict__ q,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const double m = 1.0 / diag_m(tb->sc[id[i] - 1], q[i]);
    acc[0] = v[i]; acc[1] = m; acc[2] = m * v[i];
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = acc[k];
}

// r = v - mean(v) (projection on sum = 0), r written in place of v
__global__ void k_adj_center(int n, const double* __restrict__ sums, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] -= sums[0] / double(n);
}

// z = M^-1 r - M^-1 1 (1' M^-1 r / 1' M^-1 1) (sum z = 0); part = r.z partials
__global__ void __launch_bounds__(BLOCK) k_adj_prec(int n, const Tables* __restrict__ tb,
                                                    const int* __restrict__ id,
                                                    const double* __restrict__ q,
                                                    const double* __restrict__ r,
                                                    const double* __restrict__ sums,
                                                    double* __restrict__ z,
                                                    double* __restrict__ part) {
  __shared__ double lds[BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const double m = 1.0 / diag_m(tb->sc[id[i] - 1], q[i]);
    const double zi = m * r[i] - m * (sums[2] / sums[1]);
    z[i] = zi;
    acc[0] = r[i] * zi;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// lambda += a p; r -= a y; part[2k..] = r.r, sum r partials
__global__ void __launch_bounds__(BLOCK) k_adj_update(int n, double a, const double* __restrict__ p,
                                                      const double* __restrict__ y,
                                                      double* __restrict__ lam,
                                                      double* __restrict__ r,
                                                      double* __restrict__ part) {
  __shared__ double lds[2 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (i < n) {
    lam[i] += a * p[i];
    const double ri = r[i] - a * y[i];
    r[i] = ri;
    acc[0] = ri * ri;
    acc[1] = ri;
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// p = z + b p
__global__ void k_adj_dir(int n, double b, const double* __restrict__ z, double* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + b * p[i];
}

// part[10k + s] = sum over species s+1 of (q - lambda) (dE/dchi),
// part[10k + 5 + s] = sum of (q^2 / 2 - lambda q) (dE/dJ)
__global__ void __launch_bounds__(BLOCK) k_param_grad(int n, const int* __restrict__ id,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ lam,
                                                      double* __restrict__ part) {
  __shared__ double lds[10 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < n) {
    const int s = id[i] - 1;
    const double qi = q[i], li = lam[i];
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      acc[t] = (t == s) ? qi - li : 0.0;
      acc[5 + t] = (t == s) ? 0.5 * qi * qi - li * qi : 0.0;
    }
  }
  block_sum<10>(acc, lds);
  if (threadIdx.x == 0)
    for (int t = 0; t < 10; ++t) part[10 * blockIdx.x + t] = acc[t];
}

// ---------------------------------------------------------------------------
// Derivatives in the Slater exponents xi (zeta) by forward-mode dual numbers
// through the same CTIP formulas (ctip_force_core, ctip_matrix_core) and the
// [fa|fb] coefficients of POTENTIAL_SETUP (datafmt.f90:201-220).
// ---------------------------------------------------------------------------
struct Dual {
  double v, d;
  __host__ __device__ Dual(double a = 0.0, double b = 0.0) : v(a), d(b) {}
};
__device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline Dual operator/(Dual a, Dual b) {
  return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
template <>
__device__ inline Dual fexp<Dual>(Dual a) {
  const double e = exp(a.v);
  return Dual(e, a.d * e);
}
__device__ inline Dual erfc(Dual a) {
  return Dual(::erfc(a.v), -a.d * 2.0 / SQRTPI * exp(-a.v * a.v));
}

// [fa|fb] coefficients of exponents a (row species) and b (column species)
template <class S>
__device__ inline void fafb_core(bool same, S a, S b, S c[4]) {
  if (same) {
    c[0] = S(1);
    c[1] = S(11) * a / S(8);
    c[2] = S(0.75) * (a * a);
    c[3] = (a * a * a) / S(6);
  } else {
    const S s = a + b, da = a - b, db = b - a;
    const S a2 = a * a, b2 = b * b, a4 = a2 * a2, b4 = b2 * b2;
    c[0] = a * b4 / (s * s) / (da * da);
    c[1] = b * a4 / (s * s) / (db * db);
    c[2] = (S(3) * a2 * b4 - b4 * b2) / (s * s * s) / (da * da * da);
    c[3] = (S(3) * b2 * a4 - a4 * a2) / (s * s * s) / (db * db * db);
  }
}

// part[10k + s] = half the sum over the block's list entries of dEctip/dxi_s
// / kc (FORCE's Q2 form) or, with u != null, of d/dxi_s of the CTIP pair
// force term dotted with (u_i - u_j) (half of d(sum u.F)/dxi_s at fixed
// charges); part[10k + 5 + s] = sum of lambda_i dG_i/dxi_s (the solver's Q3
// matrix and its b_i = sum_j Z_j asym_ij), s = species - 1.
__global__ void __launch_bounds__(BLOCK) k_xi_grad(Geom g, const Tables* __restrict__ tb,
                                                   const double* __restrict__ x,
                                                   const int* __restrict__ id,
                                                   const double* __restrict__ q,
                                                   const double* __restrict__ lam,
                                                   const double* __restrict__ u,
                                                   const int* __restrict__ cnt,
                                                   const int* __restrict__ col,
                                                   double* __restrict__ part) {
  __shared__ double lds[10 * BLOCK / WAVE];
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < g.n) {
    const int si = id[i];
    const double qi = q[i], li = lam[i];
    const size_t n = size_t(g.n);
    for (int k = 0; k < cnt[i]; ++k) {
      const int j = col[size_t(k) * n + i];
      double rx[3], r2;
      pair_sep(g, x, i, j, rx, r2);
      const double r = sqrt(r2);
      const int sj = id[j];
      const double qj = q[j];
      const PairCoef& pc = tb->pc[pidx(si, sj)];
      const bool same = pc.same != 0;
      // same species: one exponent, seeded in both places; else each in turn
      for (int pass = 0; pass < (same ? 1 : 2); ++pass) {
        const Dual xa(pc.xi_i, (same || pass == 0) ? 1.0 : 0.0);
        const Dual xb(pc.xi_j, (same || pass == 1) ? 1.0 : 0.0);
        Dual c[4];
        fafb_core<Dual>(same, xa, xb, c);
        Dual df, e, sym, asym;
        ctip_force_core<Dual>(same, xa, xb, Dual(pc.Z_i), Dual(pc.Z_j), c, Dual(r), Dual(r2),
                              Dual(qi), Dual(qj), df, e);
        ctip_matrix_core<Dual>(same, xa, xb, c, Dual(r), Dual(g.alpha), sym, asym);
        const int sp = (pass == 0 ? si : sj) - 1;
        double first = 0.5 * e.d;
        if (u)
          first = 0.5 * df.d * ((u[3 * i] - u[3 * j]) * rx[0] + (u[3 * i + 1] - u[3 * j + 1]) * rx[1] +
                                (u[3 * i + 2] - u[3 * j + 2]) * rx[2]);
#pragma unroll
        for (int t = 0; t < 5; ++t) {
          if (t != sp) continue;
          acc[t] += first;
          acc[5 + t] += li * (sym.d * qj + pc.Z_j * asym.d);
        }
      }
    }
  }
  block_sum<10>(acc, lds);
  if (threadIdx.x == 0)
    for (int t = 0; t < 10; ++t) part[10 * blockIdx.x + t] = acc[t];
}

// ---------------------------------------------------------------------------
// Charge derivative of S = sum_i u_i . F_i (vector-Jacobian product of the
// forces with respect to the charges), for the force-loss gradient in the
// CTIP parameters.
// ---------------------------------------------------------------------------
// Pair terms (real-space Ewald, force.f90:94-98, and CTIP, force.f90:101-140):
// F_i = sum_j s_ij r_ij with s symmetric, so S = sum over pairs of
// s_ij (u_i - u_j).r_ij and b_i = sum_j ds_ij/dq_i (u_i - u_j).r_ij; the CTIP
// derivative by a dual number in q_i through ctip_force_core.
__global__ void k_qvjp_pair(Geom g, const Tables* __restrict__ tb, const double* __restrict__ x,
                            const int* __restrict__ id, const double* __restrict__ q,
                            const double* __restrict__ u, const int* __restrict__ cnt,
                            const int* __restrict__ col, double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int si = id[i];
  const double qi = q[i], a = g.alpha, a2 = a * a;
  const size_t n = size_t(g.n);
  double acc = 0.0;
  for (int k = 0; k < cnt[i]; ++k) {
    const int j = col[size_t(k) * n + i];
    double rx[3], r2;
    pair_sep(g, x, i, j, rx, r2);
    const double r = sqrt(r2), qj = q[j];
    const PairCoef& pc = tb->pc[pidx(si, id[j])];
    const double wij = (u[3 * i] - u[3 * j]) * rx[0] + (u[3 * i + 1] - u[3 * j + 1]) * rx[1] +
                       (u[3 * i + 2] - u[3 * j + 2]) * rx[2];
    // d/dq_i of kc q_i q_j (2 e^{-a^2 r^2} a r / sqrt(pi) + erfc(a r)) / r^3
    const double dcoul = KC * qj * (2.0 * exp(-a2 * r2) * a * r / SQRTPI + ::erfc(a * r)) / r2 / r;
    Dual c[4] = {Dual(pc.c[0]), Dual(pc.c[1]), Dual(pc.c[2]), Dual(pc.c[3])};
    Dual df, e;
    ctip_force_core<Dual>(pc.same != 0, Dual(pc.xi_i), Dual(pc.xi_j), Dual(pc.Z_i), Dual(pc.Z_j), c,
                          Dual(r), Dual(r2), Dual(qi, 1.0), Dual(qj), df, e);
    acc += (dcoul + df.d) * wij;
  }
  b[i] = acc;
}

// PME part, step 1: t1_i = sum_p (u'_i . grad w_i(p)) Phi(p), Phi the
// convolved grid of the charges; grad w as k_gather_force takes it.
template <class Rg>
__global__ void k_qvjp_rec1(Geom g, const double* __restrict__ x, const double* __restrict__ up,
                            const Rg* __restrict__ grid, double* __restrict__ t1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  double uu[3], AX[3][4], AXX[3][4];
  int b[3];
  grid_coord(g, x + 3 * i, uu, b);
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline4(uu[d], AX[d], AXX[d]);
  const double kx[3] = {double(g.K[0]) / g.box[0], double(g.K[1]) / g.box[1],
                        double(g.K[2]) / g.box[2]};
  double acc = 0.0;
  for (int c = 0; c < 4; ++c)
    for (int bb = 0; bb < 4; ++bb) {
      const size_t row = size_t(g.K[0]) * (size_t(wrapk(b[1], bb, g.K[1])) +
                                           size_t(g.K[1]) * size_t(wrapk(b[2], c, g.K[2])));
      for (int a = 0; a < 4; ++a) {
        const double Qr = double(grid[row + wrapk(b[0], a, g.K[0])]);
        const double d0 = AXX[0][a] * AX[1][bb] * AX[2][c] * kx[0];
        const double d1 = AX[0][a] * AXX[1][bb] * AX[2][c] * kx[1];
        const double d2 = AX[0][a] * AX[1][bb] * AXX[2][c] * kx[2];
        acc += (up[3 * i] * d0 + up[3 * i + 1] * d1 + up[3 * i + 2] * d2) * Qr;
      }
    }
  t1[i] = acc;
}

// step 2: the dipole grid D(p) = sum_i q_i u'_i . grad w_i(p), accumulated in
// 64-bit fixed point at 2^40 with global integer atomics (order-independent)
constexpr double DIP_SCALE = 1099511627776.0;
__global__ void k_qvjp_dipole(Geom g, const double* __restrict__ x, const double* __restrict__ q,
                              const double* __restrict__ up, unsigned long long* __restrict__ acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  double uu[3], AX[3][4], AXX[3][4];
  int b[3];
  grid_coord(g, x + 3 * i, uu, b);
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline4(uu[d], AX[d], AXX[d]);
  const double kx[3] = {double(g.K[0]) / g.box[0], double(g.K[1]) / g.box[1],
                        double(g.K[2]) / g.box[2]};
  const double qi = q[i];
  for (int c = 0; c < 4; ++c)
    for (int bb = 0; bb < 4; ++bb) {
      const size_t row = size_t(g.K[0]) * (size_t(wrapk(b[1], bb, g.K[1])) +
                                           size_t(g.K[1]) * size_t(wrapk(b[2], c, g.K[2])));
      for (int a = 0; a < 4; ++a) {
        const double d0 = AXX[0][a] * AX[1][bb] * AX[2][c] * kx[0];
        const double d1 = AX[0][a] * AXX[1][bb] * AX[2][c] * kx[1];
        const double d2 = AX[0][a] * AX[1][bb] * AXX[2][c] * kx[2];
        const double v = qi * (up[3 * i] * d0 + up[3 * i + 1] * d1 + up[3 * i + 2] * d2);
        atomicAdd(&acc[row + wrapk(b[0], a, g.K[0])], (unsigned long long)llrint(v * DIP_SCALE));
      }
    }
}

__global__ void k_fixed_to_grid(size_t ng, const unsigned long long* __restrict__ acc,
                                double* __restrict__ grid) {
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < ng) grid[t] = double((long long)acc[t]) / DIP_SCALE;
}

__global__ void k_fill_u64(size_t ng, unsigned long long* __restrict__ a) {
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < ng) a[t] = 0ull;
}

// step 3: b_i += -cst (t1_i + sum_p w_i(p) (G * D)(p)) (scale applied by the caller)
template <class Rg>
__global__ void k_qvjp_rec2(Geom g, const double* __restrict__ x, const Rg* __restrict__ grid,
                            const double* __restrict__ t1, double cst, double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  b[i] += -cst * (t1[i] + gather_phi(g, x + 3 * i, grid));
}

// u' = u - mean(u) (the reciprocal forces have their mean removed, Q13);
// sums[0..2] holds sum u
__global__ void k_center3(int n, const double* __restrict__ u, const double* __restrict__ sums,
                          double* __restrict__ up) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) up[3 * i + d] = u[3 * i + d] - sums[d] / double(n);
}

// part[3k + d] = block sums of u_d
__global__ void __launch_bounds__(BLOCK) k_sum3(int n, const double* __restrict__ u,
                                                double* __restrict__ part) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < n)
    for (int d = 0; d < 3; ++d) acc[d] = u[3 * i + d];
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; ++d) part[3 * blockIdx.x + d] = acc[d];
}

// part[10k + s] = sum over species s+1 of mu_i, part[10k + 5 + s] = of mu_i q_i
__global__ void __launch_bounds__(BLOCK) k_mu_sums(int n, const int* __restrict__ id,
                                                   const double* __restrict__ q,
                                                   const double* __restrict__ mu,
                                                   double* __restrict__ part) {
  __shared__ double lds[10 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < n) {
    const int s = id[i] - 1;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      acc[t] = (t == s) ? mu[i] : 0.0;
      acc[5 + t] = (t == s) ? mu[i] * q[i] : 0.0;
    }
  }
  block_sum<10>(acc, lds);
  if (threadIdx.x == 0)
    for (int t = 0; t < 10; ++t) part[10 * blockIdx.x + t] = acc[t];
}

}  // namespace sx
