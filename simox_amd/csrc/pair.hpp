// Short-range pair forms and kernels over the neighbour list: EAM densities,
// pair terms and embedding (eam.f90, force.f90:256-348), the real-space Ewald
// and CTIP terms of FORCE, the CTIP sparse matrix, and the EAM spline
// gradients.  The list itself, the densities and the half-shell force pass
// are in tile.hpp.
#pragma once
#include "common.hpp"

namespace sx {

// ---------------------------------------------------------------------------
// Radial forms (eam.f90).  T = double for the FP64 paths, float for MIXED.
// ---------------------------------------------------------------------------
template <class T>
__device__ inline T fexp(T v) { return exp(v); }
template <>
__device__ inline float fexp<float>(float v) { return __expf(v); }

// a / b.  In FP32 it is a times the hardware reciprocal of b (v_rcp_f32, 1 ulp)
// instead of the ten-instruction IEEE division sequence; the FP64 paths keep
// the correctly rounded division.
template <class T>
__device__ inline T fdiv(T a, T b) { return a / b; }
template <>
__device__ inline float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <class T>
__device__ inline void dens_eval(const DensCoef& c, T r, T& f, T& dfr) {
  const T re = T(c.re), l = T(c.l), b = T(c.b), fe = T(c.fe);
  T r_re = fdiv<T>(r, re), x = r_re - l;
  T x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
  T l20 = x16 * x4, l19 = x16 * x2 * x;
  T eb = fexp<T>(-b * (r_re - T(1)));
  T den = T(1) + l20;
  f = fe * fdiv<T>(eb, den);
  dfr = fdiv<T>(fe * fdiv<T>(fdiv<T>(eb * (-b * den - T(20) * l19), re), den * den), r);
}

template <class T>
__device__ inline void phi_eval(const PhiCoef& c, T r, T& phi, T& df) {
  const T re = T(c.re);
  T r_re = fdiv<T>(r, re), r1 = r_re - T(1);
  T xk = r_re - T(c.k), xl = r_re - T(c.l);
  T k2 = xk * xk, k4 = k2 * k2, k8 = k4 * k4, k16 = k8 * k8;
  T l2 = xl * xl, l4 = l2 * l2, l8 = l4 * l4, l16 = l8 * l8;
  T k20 = k16 * k4, k19 = k16 * k2 * xk, l20 = l16 * l4, l19 = l16 * l2 * xl;
  T ea = fexp<T>(-T(c.a) * r1), eb = fexp<T>(-T(c.b) * r1);
  T dk = T(1) + k20, dl = T(1) + l20;
  T dphidr = fdiv<T>(fdiv<T>(T(c.A) * ea * (-T(c.a) * dk - T(20) * k19), re), dk * dk) -
             fdiv<T>(fdiv<T>(T(c.B) * eb * (-T(c.b) * dl - T(20) * l19), re), dl * dl);
  phi = fdiv<T>(T(c.A) * ea, dk) - fdiv<T>(T(c.B) * eb, dl);
  df = fdiv<T>(-dphidr, r);
}

// CTIP short-range Slater terms of FORCE (force.f90:101-140).  Same-species
// fifj lacks the 1/r factor (quirk Q2) and is reproduced as written.  The
// core takes the exponents and [fa|fb] coefficients as values of S (the
// working type, or a dual number for derivatives in the exponents).
template <class S>
__device__ inline void ctip_force_core(bool same, S xi_i, S xi_j, S Zi, S Zj, const S c[4], S r,
                                       S r2, S qi, S qj, S& df, S& e) {
  S jfi, ifj, djfidr, difjdr, fifj, dfifjdr;
  S ir = fdiv<S>(S(1), r);
  if (same) {
    S xir = S(2) * xi_i * r, ei = fexp<S>(-xir);
    jfi = -ei * (xi_i + ir);
    ifj = jfi;
    djfidr = ei * (S(2) * (xi_i * xi_i) + fdiv<S>(S(1) + xir, r2));
    difjdr = djfidr;
    const S c1 = c[1], c2 = c[2], c3 = c[3];
    fifj = -ei * (S(1) + c1 * r + c2 * r2 + c3 * r2 * r);
    dfifjdr = ei * ((xir + S(1)) * (fdiv<S>(S(1), r2) + c1 * ir + c2 + c3 * r) -
                    (c1 * ir + c2 * S(2) + c3 * S(3) * r));
  } else {
    S xir_i = S(2) * xi_i * r, ei = fexp<S>(-xir_i);
    S xir_j = S(2) * xi_j * r, ej = fexp<S>(-xir_j);
    jfi = -ei * (xi_i + ir);
    ifj = -ej * (xi_j + ir);
    djfidr = ei * (S(2) * (xi_i * xi_i) + fdiv<S>(S(1) + xir_i, r2));
    difjdr = ej * (S(2) * (xi_j * xi_j) + fdiv<S>(S(1) + xir_j, r2));
    const S c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
    fifj = -c0 * ei - c1 * ej - fdiv<S>(c2 * ei, r) - fdiv<S>(c3 * ej, r);
    dfifjdr = c0 * ei * S(2) * xi_i + c1 * ej * S(2) * xi_j +
              fdiv<S>(c2 * ei * (S(1) + xir_i), r2) + fdiv<S>(c3 * ej * (S(1) + xir_j), r2);
  }
  df = fdiv<S>(
      S(KC) * ((qi * Zj + qj * Zi - qi * qj) * dfifjdr - qi * Zj * djfidr - qj * Zi * difjdr), r);
  e = qi * Zj * (jfi - fifj) + qj * Zi * (ifj - fifj) + qi * qj * fifj;
}

template <class T>
__device__ inline void ctip_force_terms(const PairCoef& pc, T r, T r2, T qi, T qj, T& df,
                                        T& e) {
  const T c[4] = {T(pc.c[0]), T(pc.c[1]), T(pc.c[2]), T(pc.c[3])};
  ctip_force_core<T>(pc.same != 0, T(pc.xi_i), T(pc.xi_j), T(pc.Z_i), T(pc.Z_j), c, r, r2, qi,
                     qj, df, e);
}

// CTIP solver matrix entries for row i (ctip.f90:81-110): sym_ij and asym_ij.
// Same-species [fi|fj] is the solver's own variant (quirk Q3).
template <class S>
__device__ inline void ctip_matrix_core(bool same, S xi_i, S xi_j, const S c[4], S r, S alpha,
                                        S& sym, S& asym) {
  S ej = fexp<S>(S(-2) * xi_j * r), ej_r = ej / r;
  S afb_i, fifj;
  if (same) {
    afb_i = -xi_j * ej - ej_r;
    S xir = xi_i * r;
    fifj = -ej * (S(1) + xir * (c[1] + xir * (c[2] + xir * c[3]))) / r;
  } else {
    S ei = fexp<S>(S(-2) * xi_i * r), ei_r = ei / r;
    afb_i = -xi_i * ei - ei_r;
    fifj = -ei * (c[0] + c[2] / r) - ej * (c[1] + c[3] / r);
  }
  asym = S(KC) * (afb_i - fifj);
  sym = S(KC) * (fifj + erfc(alpha * r) / r);
}

template <class T>
__device__ inline void ctip_matrix_terms(const PairCoef& pc, T r, T alpha, T& sym, T& asym) {
  const T c[4] = {T(pc.c[0]), T(pc.c[1]), T(pc.c[2]), T(pc.c[3])};
  ctip_matrix_core<T>(pc.same != 0, T(pc.xi_i), T(pc.xi_j), c, r, alpha, sym, asym);
}

// ---------------------------------------------------------------------------
// Embedding function F(rho) and F'(rho) (force.f90:256-348)
// ---------------------------------------------------------------------------
template <class T>
__device__ inline T poly3(const double* c, T t) {
  T s = T(0), tp = T(1);
#pragma unroll
  for (int j = 0; j < 4; ++j) { s += T(c[j]) * tp; tp *= t; }
  return s;
}
template <class T>
__device__ inline T dpoly3(const double* c, T t, T scale) {
  T s = T(0), tp = T(1);
#pragma unroll
  for (int j = 1; j < 4; ++j) { s += T(c[j]) * T(j) * tp / scale; tp *= t; }
  return s;
}

template <class T>
__device__ inline void embed(const EmbedCoef& e, T rho, T& F, T& dF) {
  F = T(0); dF = T(0);
  if (e.metal) {
    if (rho < T(e.rho_n)) {
      T t = rho / T(e.rho_n) - T(1);
      F = poly3<T>(e.Fn, t); dF = dpoly3<T>(e.Fn, t, T(e.rho_n));
    } else if (rho < T(e.rho_e)) {
      T t = rho / T(e.rho_e) - T(1);
      F = poly3<T>(e.Fm, t); dF = dpoly3<T>(e.Fm, t, T(e.rho_e));
    } else if (rho < T(e.rho_0)) {
      T t = rho / T(e.rho_e) - T(1);
      F = poly3<T>(e.Fp, t); dF = dpoly3<T>(e.Fp, t, T(e.rho_e));
    } else if (rho > T(1.0e-6)) {
      T t = pow(rho / T(e.rho_s), T(e.eta));
      T lt = log(t);
      F = T(e.Fe) * (T(1) - lt) * t;
      dF = -T(e.Fe) * T(e.eta) * lt * t / rho;
    }
  } else {
    int k = 4;
    for (int s = 0; s < 4; ++s)
      if (rho < T(e.rmax[s])) { k = s; break; }
    T t = rho / T(e.re[k]) - T(1);
    F = poly3<T>(e.F[k], t); dF = dpoly3<T>(e.F[k], t, T(e.re[k]));
  }
}

// ---------------------------------------------------------------------------
// EAM in spline form (build-only, Al/O; the layout is SplineSet's).
// ---------------------------------------------------------------------------
__host__ __device__ inline int spl_dens(int si, int sj) {  // density into si from sj
  return sj == 1 ? 0 : (si == 1 ? 1 : 2);
}
__host__ __device__ inline int spl_pair(int si, int sj) {
  return (si == 1 && sj == 1) ? 3 : ((si == 4 && sj == 4) ? 5 : 4);
}
__host__ __device__ inline int spl_embed(int si) { return si == 1 ? 6 : 7; }
__host__ __device__ inline bool spl_species_ok(int s) { return s == 1 || s == 4; }

// Cubic B-spline weights of x on function f: returns the first of the four
// coefficient indices; w = weights, dw = their x-derivatives.  x is clamped to
// the valid range (derivatives 0 outside it).
template <class T>
__device__ inline int spl_basis(const SplineSet& sp, int f, T x, T w[4], T dw[4]) {
  const T h = T(sp.h[f]);
  T u = (x - T(sp.lo[f])) / h + T(1);
  const T ulo = T(1), uhi = T(sp.m - 2) - T(1e-6);
  bool clamped = false;
  if (!(u >= ulo)) { u = ulo; clamped = true; }  // also catches NaN
  if (!(u <= uhi)) { u = uhi; clamped = true; }
  int k0 = int(u);  // 1 .. m-3, also where uhi rounds up in float
  k0 = k0 < 1 ? 1 : (k0 > sp.m - 3 ? sp.m - 3 : k0);
  const T t = u - T(k0), t2 = t * t, t3 = t2 * t, s1 = T(1) - t;
  w[0] = s1 * s1 * s1 / T(6);
  w[1] = (T(3) * t3 - T(6) * t2 + T(4)) / T(6);
  w[2] = (-T(3) * t3 + T(3) * t2 + T(3) * t + T(1)) / T(6);
  w[3] = t3 / T(6);
  const T ih = clamped ? T(0) : T(1) / h;
  dw[0] = -s1 * s1 / T(2) * ih;
  dw[1] = (T(3) * t2 - T(4) * t) / T(2) * ih;
  dw[2] = (-T(3) * t2 + T(2) * t + T(1)) / T(2) * ih;
  dw[3] = t2 / T(2) * ih;
  return k0 - 1;
}

// value, first and second x-derivatives (second derivatives of the cubic
// B-splines: (1 - t), 3t - 2, 1 - 3t, t over h^2)
__device__ inline void spl_eval2(const SplineSet& sp, int f, double x, double& v, double& dv,
                                 double& d2v) {
  double w[4], dw[4];
  const int k = spl_basis<double>(sp, f, x, w, dw);
  const double h = sp.h[f];
  const double u = (x - sp.lo[f]) / h + 1.0;
  const double t = u - double(k + 1);
  const bool inside = dw[1] != 0.0 || dw[2] != 0.0 || dw[0] != 0.0 || dw[3] != 0.0;
  const double ih2 = inside ? 1.0 / (h * h) : 0.0;
  const double d2w[4] = {(1.0 - t) * ih2, (3.0 * t - 2.0) * ih2, (1.0 - 3.0 * t) * ih2, t * ih2};
  v = dv = d2v = 0.0;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const double c = sp.c[f][k + m];
    v += c * w[m];
    dv += c * dw[m];
    d2v += c * d2w[m];
  }
}

template <class T>
__device__ inline void spl_eval(const SplineSet& sp, int f, T x, T& v, T& dv) {
  T w[4], dw[4];
  const int k = spl_basis<T>(sp, f, x, w, dw);
  v = T(0);
  dv = T(0);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const T c = T(sp.c[f][k + m]);
    v += c * w[m];
    dv += c * dw[m];
  }
}

#define SX_ERR_SET(ptr, code) atomicCAS((ptr), 0, (code))

// Atom block of this workgroup in the pair kernels.  Workgroups go to the 8
// XCDs round-robin; with g.xcd the consecutive atom blocks (neighbouring
// cells, which gather the same neighbour positions) run on one XCD and share
// its L2.  Any bijection of the blocks gives the same per-atom results.
__device__ inline int pair_block(const Geom& g) {
  return g.xcd ? xcd_tile(blockIdx.x, gridDim.x) : int(blockIdx.x);
}

// Separation of sorted atoms i and j (exact reference arithmetic).
__device__ inline void pair_sep(const Geom& g, const double* __restrict__ x, int i, int j,
                                double rx[3], double& r2) {
  min_image(g, x[3 * i], x[3 * i + 1], x[3 * i + 2], x[3 * j], x[3 * j + 1], x[3 * j + 2], rx, r2);
}

// MIXED (T = float): pairs closer than R_NEAR are evaluated in FP64.  The
// EAM terms there are large (|phi'|, |F' f'| of 5-30 eV/A at 1.2-1.8 A, the
// Al-O and O-O contacts of an oxidising surface) and the FP32 forms lose
// ~1e-7 of them, i.e. ~1e-5 of an RMS force; beyond 2.4 A every term is below
// ~1.5 eV/A and FP32 keeps the error under 2e-6 of the RMS force.  Bulk Al
// (nearest neighbours 2.86 A) has no such pairs.  The density pass evaluates
// them in FP64 in place; the FP32 force pass queues them for a small FP64
// kernel, so that it keeps its register budget.
constexpr double R_NEAR2 = 2.4 * 2.4;

template <class T>
__device__ inline bool near_pair(double r2) {
  if constexpr (sizeof(T) == sizeof(float)) return r2 < R_NEAR2;
  return false;
}

// One pair of FORCE's loop (force.f90:86-140) in arithmetic type S: the
// scalar s with F_i += rx * s, and the pair's phi, q_i q_j erfc(ar)/r and
// CTIP energy.  EAM terms only inside the EAM cutoff (eam).
template <class S>
__device__ inline S pair_force(const Tables* __restrict__ tb, const PairCoef& pc, int si, int sj,
                               S r, S r2, S qi, S qj, S dFi, S dFj, bool eam, S a, S& ep, S& ed,
                               S& ec) {
  S fi_, dfi, fj_, dfj, phi, dfp;
  if (tb->sp.on) {  // spline EAM: dfi, dfj, dfp in the analytic forms' units (/r)
    spl_eval<S>(tb->sp, spl_dens(sj, si), r, fi_, dfi);
    spl_eval<S>(tb->sp, spl_dens(si, sj), r, fj_, dfj);
    spl_eval<S>(tb->sp, spl_pair(si, sj), r, phi, dfp);
    dfi = fdiv<S>(dfi, r);
    dfj = fdiv<S>(dfj, r);
    dfp = fdiv<S>(-dfp, r);
  } else {
    dens_eval<S>(pc.di, r, fi_, dfi);
    dens_eval<S>(pc.dj, r, fj_, dfj);
    phi_eval<S>(pc.ph, r, phi, dfp);
    if (pc.kind == 1) {
      // METAL_HETERO (eam.f90:90-96): phi = (phi_i f_j/f_i + phi_j f_i/f_j)/2
      // and -phi'/r, with d(f_j/f_i)/dr / r = (f_j'/r f_i - f_j f_i'/r)/f_i^2
      S phj, dpj;
      phi_eval<S>(pc.ph2, r, phj, dpj);
      const S phi_i = phi, ri = fdiv<S>(fj_, fi_), rj = fdiv<S>(fi_, fj_);
      phi = S(0.5) * (phi_i * ri + phj * rj);
      dfp = S(0.5) * (dfp * ri + dpj * rj) -
            S(0.5) * (phi_i * fdiv<S>(dfj * fi_ - fj_ * dfi, fi_ * fi_) +
                      phj * fdiv<S>(dfi * fj_ - fi_ * dfj, fj_ * fj_));
    }
  }
  // EAM: pair force and embedding force of both atoms (force.f90:86-90,349,356)
  S s = eam ? dfp - dFi * dfj - dFj * dfi : S(0);
  // real-space Ewald (force.f90:94-96)
  const S tmp = erfc(a * r);
  s += fdiv<S>(
      fdiv<S>(S(KC) * qi * qj * (fdiv<S>(S(2) * fexp<S>(-(a * a) * r2) * a * r, S(SQRTPI)) + tmp),
              r2),
      r);
  ed = fdiv<S>(qi * qj * tmp, r);
  // CTIP Slater terms (force.f90:101-140); in FP32 not those of pc.ctip64,
  // which k_force_ctip64 adds in FP64
  S dfc = S(0);
  ec = S(0);
  if (!(sizeof(S) == sizeof(float) && pc.ctip64)) ctip_force_terms<S>(pc, r, r2, qi, qj, dfc, ec);
  ep = eam ? phi : S(0);
  return s + dfc;
}

// Pair histogram of FORCE (force.f90:55,76,159,180): NRDF bins of width
// pair%dr per ordered species pair, rdf(5,5,201) from its second bin on.
constexpr int NRDF = 200;  // datafmt.f90:31
constexpr int RDF_BINS = 25 * (NRDF + 1);

// Gradient of the spline-form EAM energy E = sum_pairs phi(r) + sum_i F(rho_i)
// with respect to the spline coefficients (build-only):
//   dE/dc(pair)_k = sum over pairs of B_k(r),
//   dE/dc(F)_k    = sum over atoms of B_k(rho_i),
//   dE/dc(dens)_k = sum_i F'(rho_i) sum_j B_k(r_ij) over the sources j.
// Each block accumulates its atoms' terms in LDS in 64-bit fixed point at
// 2^32 (integer adds: order-independent, so deterministic) and writes one
// partial row of NSPL * SPL_MAX values.
constexpr double SPL_SCALE = 4294967296.0;
__global__ void __launch_bounds__(BLOCK) k_eam_spline_grad(Geom g, const Tables* __restrict__ tb,
                                                           const double* __restrict__ x,
                                                           const int* __restrict__ id,
                                                           const double* __restrict__ rho,
                                                           const int* __restrict__ cnt,
                                                           const int* __restrict__ col,
                                                           double* __restrict__ part) {
  __shared__ unsigned long long acc[NSPL * SPL_MAX];
  for (int k = threadIdx.x; k < NSPL * SPL_MAX; k += blockDim.x) acc[k] = 0ull;
  __syncthreads();
  const SplineSet& sp = tb->sp;
  auto add = [&](int f, int k0, const double w[4], double scale) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
      atomicAdd(&acc[f * SPL_MAX + k0 + m], (unsigned long long)llrint(scale * w[m] * SPL_SCALE));
  };
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < g.n) {
    const int si = id[i];
    double w[4], dw[4], F, dF;
    const int fe = spl_embed(si);
    add(fe, spl_basis<double>(sp, fe, rho[i], w, dw), w, 1.0);
    spl_eval<double>(sp, fe, rho[i], F, dF);
    const size_t n = size_t(g.n);
    for (int k = 0; k < cnt[i]; ++k) {
      const int j = col[size_t(k) * n + i];
      double rx[3], r2;
      pair_sep(g, x, i, j, rx, r2);
      if (!(r2 < g.ream2)) continue;
      const double r = sqrt(r2);
      const int sj = id[j];
      const int fp = spl_pair(si, sj), fd = spl_dens(si, sj);
      add(fp, spl_basis<double>(sp, fp, r, w, dw), w, 0.5);
      add(fd, spl_basis<double>(sp, fd, r, w, dw), w, dF);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NSPL * SPL_MAX; k += blockDim.x)
    part[size_t(blockIdx.x) * (NSPL * SPL_MAX) + k] = double((long long)acc[k]) / SPL_SCALE;
}

// Vector-Jacobian product of the spline-form EAM forces (build-only): for
// S = sum_i u_i . F_i(EAM), with W_ij = (u_i - u_j) . r_ij / r (symmetric in
// i, j) and T_i = sum_j f'_{j->i}(r_ij) W_ij,
//   dS/dc(pair)_k = -1/2 sum_i sum_j B'_k(r_ij) W_ij,
//   dS/dc(F)_k    = -sum_i B'_k(rho_i) T_i,
//   dS/dc(dens)_k = -sum_i sum_j (F''(rho_i) T_i B_k(r_ij) + F'(rho_i) W_ij B'_k(r_ij)),
// the sums over j running over the full shell (each pair from both sides).
__device__ inline double vjp_w(const double* __restrict__ u, int i, int j, const double rx[3],
                               double r) {
  return ((u[3 * i] - u[3 * j]) * rx[0] + (u[3 * i + 1] - u[3 * j + 1]) * rx[1] +
          (u[3 * i + 2] - u[3 * j + 2]) * rx[2]) / r;
}

// pass 1: T_i
__global__ void k_eam_vjp_t(Geom g, const Tables* __restrict__ tb, const double* __restrict__ x,
                            const int* __restrict__ id, const double* __restrict__ u,
                            const int* __restrict__ cnt, const int* __restrict__ col,
                            double* __restrict__ t_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int si = id[i];
  const size_t n = size_t(g.n);
  double t = 0.0;
  for (int k = 0; k < cnt[i]; ++k) {
    const int j = col[size_t(k) * n + i];
    double rx[3], r2;
    pair_sep(g, x, i, j, rx, r2);
    if (!(r2 < g.ream2)) continue;
    const double r = sqrt(r2);
    double f, df;
    spl_eval<double>(tb->sp, spl_dens(si, id[j]), r, f, df);
    t += df * vjp_w(u, i, j, rx, r);
  }
  t_out[i] = t;
}

// pass 2: the coefficient sums, 64-bit fixed point at 2^28 in LDS per block
constexpr double VJP_SCALE = 268435456.0;
__global__ void __launch_bounds__(BLOCK) k_eam_vjp(Geom g, const Tables* __restrict__ tb,
                                                   const double* __restrict__ x,
                                                   const int* __restrict__ id,
                                                   const double* __restrict__ rho,
                                                   const double* __restrict__ u,
                                                   const double* __restrict__ tv,
                                                   const int* __restrict__ cnt,
                                                   const int* __restrict__ col,
                                                   double* __restrict__ part) {
  __shared__ unsigned long long acc[NSPL * SPL_MAX];
  for (int k = threadIdx.x; k < NSPL * SPL_MAX; k += blockDim.x) acc[k] = 0ull;
  __syncthreads();
  const SplineSet& sp = tb->sp;
  auto add = [&](int f, int k0, const double w[4], double scale) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
      atomicAdd(&acc[f * SPL_MAX + k0 + m], (unsigned long long)llrint(scale * w[m] * VJP_SCALE));
  };
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < g.n) {
    const int si = id[i];
    const double ti = tv[i];
    double w[4], dw[4], F, dF, d2F;
    const int fe = spl_embed(si);
    add(fe, spl_basis<double>(sp, fe, rho[i], w, dw), dw, -ti);
    spl_eval2(sp, fe, rho[i], F, dF, d2F);
    const size_t n = size_t(g.n);
    for (int k = 0; k < cnt[i]; ++k) {
      const int j = col[size_t(k) * n + i];
      double rx[3], r2;
      pair_sep(g, x, i, j, rx, r2);
      if (!(r2 < g.ream2)) continue;
      const double r = sqrt(r2);
      const double wij = vjp_w(u, i, j, rx, r);
      const int sj = id[j];
      const int fp = spl_pair(si, sj), fd = spl_dens(si, sj);
      add(fp, spl_basis<double>(sp, fp, r, w, dw), dw, -0.5 * wij);
      const int k0 = spl_basis<double>(sp, fd, r, w, dw);
      add(fd, k0, w, -d2F * ti);
      add(fd, k0, dw, -dF * wij);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NSPL * SPL_MAX; k += blockDim.x)
    part[size_t(blockIdx.x) * (NSPL * SPL_MAX) + k] = double((long long)acc[k]) / VJP_SCALE;
}

// out[v] = sum over the nparts rows of part[row * nv + v], rows in order
__global__ void k_sum_rows(const double* __restrict__ part, int nparts, int nv,
                           double* __restrict__ out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  double s = 0.0;
  for (int r = 0; r < nparts; ++r) s += part[size_t(r) * nv + v];
  out[v] = s;
}

// K5: CTIP sparse matrix over the neighbour list: val[k*n + i] = sym_ij for
// the list's entry k.  Only sym is stored; the asym part enters the solve only
// through b_i = sum_j Z_j asym_ij, which is constant for a whole solve (the
// reference recomputes it per gradient, ctip.f90:327).
// The packed inner matrix's word of an entry (the SELL-4 slices' encoding):
// the value as FP16 with its mantissa rounded to 6 bits in bits 31..20, the
// column in bits 19..0; PK_DROP for an entry left out (|v| < tau), which no
// finite value's word equals.
constexpr unsigned PK_DROP = 0xFFFFFFFFu;
__device__ inline unsigned pk_word(double v, int j) {
  const unsigned hb = __builtin_bit_cast(unsigned short, _Float16(float(v)));
  const unsigned code = (hb & 0x8000u) | (((hb & 0x7FFFu) + 8u) & 0x7FF0u);
  return (code << 16) | unsigned(j);
}

// PACK: each entry's packed word (or PK_DROP) into pkw (the list's layout)
// and the row's kept-entry count into kept, for k_ell_to_sell4.
template <class T, int U = 1, bool PACK = false>
__global__ void __launch_bounds__(BLOCK) k_ctip_matrix(Geom g, const Tables* __restrict__ tb,
                                                       const double* __restrict__ x,
                                                       const int* __restrict__ id,
                                                       const int* __restrict__ cnt,
                                                       const int* __restrict__ col,
                                                       T* __restrict__ val,
                                                       double* __restrict__ b, double tau = 0.0,
                                                       unsigned* __restrict__ pkw = nullptr,
                                                       int* __restrict__ kept = nullptr) {
  const int i = pair_block(g) * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int si = id[i];
  const T alpha = T(g.alpha);
  const size_t n = size_t(g.n);
  const int c = cnt[i];
  double bi = 0.0;
  int nk = 0;
  if (si >= 1 && si <= NSPEC) {
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    // U entries' neighbour data loaded ahead, then taken in list order
    for (int k0 = 0; k0 < c; k0 += U) {
      int sv[U], jv[U];
      double xv[U][3];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = k0 + u < c ? col[size_t(k0 + u) * n + i] : i;
        jv[u] = j;
        sv[u] = id[j];
        xv[u][0] = x[3 * j];
        xv[u][1] = x[3 * j + 1];
        xv[u][2] = x[3 * j + 2];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u;
        if (k >= c) break;
        double rx[3], r2;
        min_image(g, xi, yi, zi, xv[u][0], xv[u][1], xv[u][2], rx, r2);
        const int sj = sv[u];
        if (sj < 1 || sj > NSPEC) {
          if constexpr (PACK) pkw[size_t(k) * n + i] = PK_DROP;
          continue;
        }
        const PairCoef& pc = tb->pc[pidx(si, sj)];
        T sym, asym;
        ctip_matrix_terms<T>(pc, T(sqrt(r2)), alpha, sym, asym);
        bi += double(T(pc.Z_j) * asym);
        val[size_t(k) * n + i] = sym;
        if constexpr (PACK) {
          const bool keep = fabs(double(sym)) >= tau;
          pkw[size_t(k) * n + i] = keep ? pk_word(double(sym), jv[u]) : PK_DROP;
          nk += keep;
        }
      }
    }
  }
  if constexpr (PACK) {
    if (!(si >= 1 && si <= NSPEC))
      for (int k = 0; k < c; ++k) pkw[size_t(k) * n + i] = PK_DROP;
    kept[i] = nk;
  }
  b[i] = bi;
}

}  // namespace sx
