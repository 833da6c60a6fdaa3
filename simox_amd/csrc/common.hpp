// Shared device code of the SIMOX hot-path kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/simox_gpu.h"

namespace sx {

constexpr double KC = 14.399644154020082;       // datafmt.f90:23
constexpr double PI = 3.1415926535897931;       // datafmt.f90:25
constexpr double SQRTPI = 1.7724538509055159;   // datafmt.f90:26
constexpr double WPEN = 20.0;                   // datafmt.f90:28
constexpr int NPT_MAX = 500;                    // datafmt.f90:20
constexpr int NNB_MAX = 500;                    // datafmt.f90:22
constexpr int NSPEC = 4;                        // Al, Ni, Fe, O handled on the device
constexpr int BLOCK = 256;
constexpr int WAVE = 64;

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): workgroup b runs on XCD b mod 8.  xcd_tile(b, nb) maps it to tile
// (b mod 8) (nb / 8) + b / 8, so consecutive tiles, which share cache lines
// or halo data, run on one XCD; the last nb mod 8 workgroups keep their index.
constexpr int NXCD = 8;
__device__ inline int xcd_tile(int b, int nb) {
  const int per = nb / NXCD;
  if (b >= per * NXCD) return b;
  return (b % NXCD) * per + b / NXCD;
}

// Fixed-point scale of the PME charge grid (2^50, quantum 8.9e-16): integer
// atomics make the spread order-independent, hence deterministic.  Grid values
// must stay below 2^13 = 8192 in magnitude (a grid point collects a few atoms'
// charges times B-spline weights <= 1).
constexpr double GRID_SCALE = 1125899906842624.0;
constexpr double GRID_INV_SCALE = 1.0 / 1125899906842624.0;

// ---------------------------------------------------------------------------
// Per species-pair coefficient table.  The EAM branches of eam.f90
// (METAL_HOMO, AL_NI, METAL_OXYGEN, OXYGEN_METAL, OXYGEN_ONLY) are one radial
// form each, METAL_HETERO a weighted mean of two:
//   f(r)   = fe e^{-b(r/re-1)} / (1 + (r/re - l)^20)             (densities)
//   phi(r) = A e^{-a(r/re-1)} / (1 + (r/re - k)^20)
//          - B e^{-b(r/re-1)} / (1 + (r/re - l)^20)             (pair term)
// with the coefficients picked per pair as the reference's branch picks them
// (force.f90:58-74).  The O density in a metal-O pair uses the pair's r_e
// (eam.f90:276,299-302, quirk Q15).
// ---------------------------------------------------------------------------
struct DensCoef { double re, fe, b, l; };
struct PhiCoef { double re, a, b, A, B, k, l; };
struct PairCoef {
  int valid;
  int same;            // id_i == id_j (CTIP branch, force.f90:101)
  int kind;            // 0: pair term ph; 1: METAL_HETERO's mean of ph (metal i) and ph2 (metal j)
  int ctip64;          // Slater terms of a pair of close xi (MIXED evaluates them in FP64)
  DensCoef di, dj;     // fj(1) and fj(2) of the reference's EAM routine for (id_i, id_j)
  PhiCoef ph, ph2;
  double xi_i, xi_j, Z_i, Z_j;
  double c[4];         // fafb(id_i, id_j, 1:4)
};

struct EmbedCoef {     // force.f90:256-348
  int metal;
  double rho_e, rho_s, rho_n, rho_0, eta, Fe;
  double Fn[4], Fm[4], Fp[4];
  double re[5], rmax[5], F[5][4];
};

struct SpecCoef {      // per-species CTIP data (ctip.f90:33-39) and mass
  double chi, J, Jp, Z, qmin, qmax, xm, xi;  // Jp = J - 2 kc a / sqrt(pi)
};

// EAM in spline form (build-only, Al/O): eight functions, each a uniform
// cubic B-spline s(x) = sum_k c_k B((x - lo)/h + 1 - k), k = 0..m-1, valid on
// [lo, lo + (m - 3) h]:
//   0 density of an Al source (into any atom), 1 density of an O source into
//   Al, 2 density of an O source into O, 3 pair term Al-Al, 4 Al-O, 5 O-O,
//   6 embedding F of Al, 7 of O.
constexpr int NSPL = 8;
constexpr int SPL_MAX = 64;
struct SplineSet {
  int on, m;
  double lo[NSPL], h[NSPL];
  double c[NSPL][SPL_MAX];
};

struct Tables {
  PairCoef pc[NSPEC * NSPEC];
  EmbedCoef ec[NSPEC];
  SpecCoef sc[NSPEC];
  SplineSet sp;
};

// Geometry passed by value to kernels.
struct Geom {
  int n;
  int nc[3];
  int ncall;
  double box[3];
  double hbox[3];
  double qbox[3];      // box/4: |dx| below it cannot wrap
  double lcell[3];
  double rcut2;
  double ream2;        // EAM cutoff squared (= rcut2 unless split)
  double alpha;        // Ewald a = 3.8/Rcut
  int K[3];            // PME grid
  int H;               // Kx/2 + 1
  int xcd;             // pair kernels: XCD-aware block-to-atom-block mapping (pair_block)
};

__host__ __device__ inline int pidx(int si, int sj) { return (si - 1) * NSPEC + (sj - 1); }

template <class T>
__device__ inline T pow20(T x) {
  T x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
  return x16 * x4;
}

// Minimum-image separation exactly as dx - box*DNINT(dx/box) (force.f90:51-53):
// no FMA contraction, IEEE division, round half away from zero.
__host__ __device__ inline void min_image(const Geom& g, double xi, double yi, double zi, double xj,
                                 double yj, double zj, double rx[3], double& r2) {
#pragma clang fp contract(off)
  double d[3] = {xi - xj, yi - yj, zi - zj};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double s = 0.0;
    if (!(fabs(d[k]) < g.qbox[k])) s = round(d[k] / g.box[k]);
    rx[k] = d[k] - g.box[k] * s;
  }
  r2 = rx[0] * rx[0] + rx[1] * rx[1] + rx[2] * rx[2];
}

__host__ __device__ inline int cell_wrap(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

// ---------------------------------------------------------------------------
// Reductions: deterministic (fixed shuffle tree per wave, fixed order across
// waves, fixed order across blocks in the final pass).
// ---------------------------------------------------------------------------
template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
template <class T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
  return v;
}

// Block-wide sum of NV values per thread; result valid in thread 0.
template <int NV, class T>
__device__ inline void block_sum(T (&v)[NV], T* lds /* >= NV * blockDim/64 */) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * nw + w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      T s = lds[k * nw];
      for (int q = 1; q < nw; ++q) s += lds[k * nw + q];
      v[k] = s;
    }
  }
  __syncthreads();
}

}  // namespace sx
