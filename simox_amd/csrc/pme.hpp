// Smooth-PME kernels: B-spline charge spreading (K6), LDS-resident 3D FFT
// passes (K7) with the influence-function multiply fused into the z pass (K8),
// and the field / force gathers (K9).
//
// Grid layout: real grid [Kz][Ky][Kx] (x fastest); half spectrum
// [Kz][Ky][H], H = Kx/2 + 1.  One PME evaluation:
//   spread -> x R2C -> y FFT -> z (FFT, x G, inverse FFT) -> y inverse -> x C2R -> gather
// which equals the reference's Qfin = ZFFT3D(-1)(G . ZFFT3D(+1)(Qbas))
// (force.f90:411-426) because G is real and even: Qfin = IFFT(G . FFT(Q)) / Ng.
#pragma once
#include <type_traits>
#include "common.hpp"

namespace sx {

template <class R> struct Cplx;
template <> struct Cplx<double> { using T = double2; };
template <> struct Cplx<float> { using T = float2; };

struct FftPlan {
  int L;
  int nst;
  int rad[16];
};

template <class C>
__device__ inline C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <class C>
__device__ inline C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <class C>
__device__ inline C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// (i * s) * a for s = +-1
template <class C>
__device__ inline C cis(C a, int s) { C r; r.x = -s * a.y; r.y = s * a.x; return r; }

// Small DFTs X_k = sum_r a_r exp(i s 2 pi r k / R)
template <class C, int R> struct Dft;
template <class C> struct Dft<C, 2> {
  __device__ static void run(C* v, int) { C a = v[0], b = v[1]; v[0] = cadd(a, b); v[1] = csub(a, b); }
};
template <class C> struct Dft<C, 4> {
  __device__ static void run(C* v, int s) {
    C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    C t2 = cadd(v[1], v[3]), t3 = cis(csub(v[1], v[3]), s);
    v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
  }
};
template <class C> struct Dft<C, 8> {
  // two 4-point DFTs (even / odd samples) joined by the twiddles W_8^{s k}
  __device__ static void run(C* v, int s) {
    using R = decltype(v[0].x);
    const R h = R(0.70710678118654752440);
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    Dft<C, 4>::run(e, s);
    Dft<C, 4>::run(o, s);
    C w1, w3;
    w1.x = h * (o[1].x - R(s) * o[1].y);
    w1.y = h * (o[1].y + R(s) * o[1].x);
    const C w2 = cis(o[2], s);
    w3.x = h * (-o[3].x - R(s) * o[3].y);
    w3.y = h * (-o[3].y + R(s) * o[3].x);
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);   v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);   v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);   v[7] = csub(e[3], w3);
  }
};
template <class C> struct Dft<C, 3> {
  __device__ static void run(C* v, int s) {
    using R = decltype(v[0].x);
    const R h = R(0.86602540378443864676);
    C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    C a = v[0];
    C m; m.x = a.x - R(0.5) * t1.x; m.y = a.y - R(0.5) * t1.y;
    C u = cis(t2, s); u.x *= h; u.y *= h;
    v[0] = cadd(a, t1); v[1] = cadd(m, u); v[2] = csub(m, u);
  }
};
template <class C> struct Dft<C, 5> {
  __device__ static void run(C* v, int s) {
    using R = decltype(v[0].x);
    const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
    const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
    C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    C t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    C a = v[0];
    C m1, m2, u1, u2;
    m1.x = a.x + c1 * t1.x + c2 * t2.x; m1.y = a.y + c1 * t1.y + c2 * t2.y;
    m2.x = a.x + c2 * t1.x + c1 * t2.x; m2.y = a.y + c2 * t1.y + c1 * t2.y;
    C w1; w1.x = s1 * t3.x + s2 * t4.x; w1.y = s1 * t3.y + s2 * t4.y;
    C w2; w2.x = s2 * t3.x - s1 * t4.x; w2.y = s2 * t3.y - s1 * t4.y;
    u1 = cis(w1, s); u2 = cis(w2, s);
    v[0] = cadd(a, cadd(t1, t2));
    v[1] = cadd(m1, u1); v[4] = csub(m1, u1);
    v[2] = cadd(m2, u2); v[3] = csub(m2, u2);
  }
};

// One Stockham stage of radix R over M lines (self-sorting, mixed radix):
// butterfly j reads src[j + r L/R], twiddles by W_L^{s r (j mod Ns) L/(Ns R)},
// and writes dst[(j/Ns) Ns R + (j mod Ns) + r Ns].
template <class C, int R>
__device__ inline void fft_stage(const C* __restrict__ src, C* __restrict__ dst, int M, int L,
                                 int stride, int Ns, const C* __restrict__ tw, int sign) {
  const int nb = L / R;
  const int total = M * nb;
  const int tws = L / (Ns * R);
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    const int line = b / nb;
    const int j = b - line * nb;
    const int k = j % Ns;
    const C* s = src + line * stride;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      C a = s[j + r * nb];
      if (r > 0 && k > 0) {
        C w = tw[r * k * tws];
        if (sign > 0) w.y = -w.y;
        a = cmul(a, w);
      }
      v[r] = a;
    }
    Dft<C, R>::run(v, sign);
    C* d = dst + line * stride + (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r * Ns] = v[r];
  }
}

// Transforms M lines of length L held in LDS; returns the buffer holding the result.
template <class C>
__device__ inline C* lds_fft(C* b0, C* b1, int M, int L, int stride, const FftPlan& pl,
                             const C* __restrict__ tw, int sign) {
  C* src = b0;
  C* dst = b1;
  int Ns = 1;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.rad[s];
    if (R == 4) fft_stage<C, 4>(src, dst, M, L, stride, Ns, tw, sign);
    else if (R == 2) fft_stage<C, 2>(src, dst, M, L, stride, Ns, tw, sign);
    else if (R == 3) fft_stage<C, 3>(src, dst, M, L, stride, Ns, tw, sign);
    else fft_stage<C, 5>(src, dst, M, L, stride, Ns, tw, sign);
    __syncthreads();
    Ns *= R;
    C* t = src; src = dst; dst = t;
  }
  return src;
}

// Compile-time plans: for the grid sizes of the workloads the line length is a
// template parameter, so every stage's radix, span and the index arithmetic
// (divisions by L/R and Ns) are constants.  Radix 8 first (3 stages for 256,
// 512, 192), then 4, 2, 3, 5.
__host__ __device__ constexpr int ct_radix(int m) {
  return (m % 8 == 0) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : 5;
}
__host__ __device__ constexpr int ct_nstages(int L) {
  int m = L, n = 0;
  while (m > 1) { m /= ct_radix(m); ++n; }
  return n;
}
__host__ __device__ constexpr int ct_radix_at(int L, int s) {
  int m = L;
  for (int k = 0; k < s; ++k) m /= ct_radix(m);
  return ct_radix(m);
}
__host__ __device__ constexpr int ct_ns_at(int L, int s) {
  int m = L, ns = 1;
  for (int k = 0; k < s; ++k) { const int r = ct_radix(m); ns *= r; m /= r; }
  return ns;
}

// Padded line layout between two stages: element i at i + i/8, so the
// strided writes of the first stages (stride R elements at Ns = 1) fall in
// different LDS banks.  PI / PO: the stage reads / writes padded lines.
__host__ __device__ constexpr int lpad(int i) { return i + (i >> 3); }

template <class C, int L, int R, int Ns, bool PI = false, bool PO = false, int NT = BLOCK>
__device__ inline void fft_stage_ct(const C* __restrict__ src, C* __restrict__ dst, int M,
                                    int stride, const C* __restrict__ tw, int sign) {
  constexpr int nb = L / R;
  constexpr int tws = L / (Ns * R);
  const int total = M * nb;
#pragma unroll 4
  for (int b = threadIdx.x; b < total; b += NT) {
    const int line = b / nb;
    const int j = b - line * nb;
    const int k = j % Ns;
    const C* s = src + line * stride;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      C a = s[PI ? lpad(j + r * nb) : j + r * nb];
      if (r > 0 && k > 0) {
        C w = tw[r * k * tws];
        if (sign > 0) w.y = -w.y;
        a = cmul(a, w);
      }
      v[r] = a;
    }
    Dft<C, R>::run(v, sign);
    C* d = dst + line * stride;
    const int o = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[PO ? lpad(o + r * Ns) : o + r * Ns] = v[r];
  }
}

// All stages on padded intermediate lines: the first stage reads and the last
// writes the caller's unpadded lines.  Line stride >= L + L/8.
template <class C, int L, int S, int NT>
__device__ inline void fft_stages_pad(C* a, C* b, int M, int stride, const C* __restrict__ tw,
                                      int sign) {
  if constexpr (S < ct_nstages(L)) {
    fft_stage_ct<C, L, ct_radix_at(L, S), ct_ns_at(L, S), (S > 0), (S + 1 < ct_nstages(L)), NT>(
        a, b, M, stride, tw, sign);
    __syncthreads();
    fft_stages_pad<C, L, S + 1, NT>(b, a, M, stride, tw, sign);
  }
}
template <class C, int L, int NT = BLOCK>
__device__ inline C* lds_fft_pad(C* b0, C* b1, int M, int stride, const C* __restrict__ tw,
                                 int sign) {
  fft_stages_pad<C, L, 0, NT>(b0, b1, M, stride, tw, sign);
  return (ct_nstages(L) % 2 == 0) ? b0 : b1;
}
__host__ __device__ constexpr int pad_stride(int L) { return L + L / 8 + 1; }

template <class C, int L, int S>
__device__ inline void fft_stages_ct(C* a, C* b, int M, int stride, const C* __restrict__ tw,
                                     int sign) {
  if constexpr (S < ct_nstages(L)) {
    fft_stage_ct<C, L, ct_radix_at(L, S), ct_ns_at(L, S)>(a, b, M, stride, tw, sign);
    __syncthreads();
    fft_stages_ct<C, L, S + 1>(b, a, M, stride, tw, sign);
  }
}

// L_CT > 0: compile-time length; 0: runtime plan
template <class C, int L_CT>
__device__ inline C* lds_fft_any(C* b0, C* b1, int M, int L, int stride, const FftPlan& pl,
                                 const C* __restrict__ tw, int sign) {
  if constexpr (L_CT > 0) {
    fft_stages_ct<C, L_CT, 0>(b0, b1, M, stride, tw, sign);
    return (ct_nstages(L_CT) % 2 == 0) ? b0 : b1;
  } else {
    return lds_fft(b0, b1, M, L, stride, pl, tw, sign);
  }
}

// ---------------------------------------------------------------------------
// B-spline weights of order 4 and their u-derivatives (force.f90:388-400)
// ---------------------------------------------------------------------------
__device__ inline void bspline4(double u, double AX[4], double AXX[4]) {
#pragma clang fp contract(off)
  double z = trunc(u) - u + 1.0;
  AX[0] = z * z * z / 6.0;
  AXX[0] = -3.0 * (z * z) / 6.0;
  z += 1.0;
  AX[1] = (-3.0 * (z * z * z) + 12.0 * (z * z) - 12.0 * z + 4.0) / 6.0;
  AXX[1] = (9.0 * (z * z) - 24.0 * z + 12.0) / 6.0;
  z += 1.0;
  AX[2] = (3.0 * (z * z * z) - 24.0 * (z * z) + 60.0 * z - 44.0) / 6.0;
  AXX[2] = (-9.0 * (z * z) + 48.0 * z - 60.0) / 6.0;
  z += 1.0;
  AX[3] = (-(z * z * z) + 12.0 * (z * z) - 48.0 * z + 64.0) / 6.0;
  AXX[3] = (3.0 * (z * z) - 24.0 * z + 48.0) / 6.0;
}

// The weights alone, with the divisions by 6 as products by 1/6 (within an
// ulp of bspline4's): the spread and the gather of the converged solves,
// where the FAITHFUL mode's exact replay of the reference's expressions does
// not apply.  Horner forms of the same cubics.
__device__ inline void bspline4_w(double u, double AX[4]) {
  constexpr double s6 = 1.0 / 6.0;
  double z = trunc(u) - u + 1.0;
  AX[0] = z * z * z * s6;
  z += 1.0;
  AX[1] = (((-3.0 * z + 12.0) * z - 12.0) * z + 4.0) * s6;
  z += 1.0;
  AX[2] = (((3.0 * z - 24.0) * z + 60.0) * z - 44.0) * s6;
  z += 1.0;
  AX[3] = (((-z + 12.0) * z - 48.0) * z + 64.0) * s6;
}

// u = K (x + L/2) / L, base = INT(u) (force.f90:388,404)
__device__ inline void grid_coord(const Geom& g, const double* xi, double u[3], int b[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    u[d] = double(g.K[d]) * (xi[d] + g.box[d] * 0.5) / g.box[d];
    b[d] = int(u[d]);
  }
}
// 0-based grid index of stencil point nn (0..3): MOD(INT(u) + nn - 1 + K, K)
// (any int b is mapped into [0, K), so positions outside the box stay in bounds)
__device__ inline int wrapk(int b, int nn, int K) {
  int v = (b + nn - 1) % K;
  return v < 0 ? v + K : v;
}

// Half-spectrum layout: spec[(kx * Ky + ky) * Kz + kz], kx = 0 .. Kx/2.
// The x passes transpose between the real grid (x fastest) and this layout,
// so the z pass works on contiguous lines and the y pass on columns of
// consecutive kz (coalesced in both).
__device__ inline size_t spec_index(const Geom& g, int kx, int ky, int kz) {
  return (size_t(kx) * g.K[1] + ky) * size_t(g.K[2]) + kz;
}

// K7a: R2C along x.  Block (bz, y) transforms the M real lines
// (y, z = bz*M .. bz*M+M-1), two per complex FFT (a + i b), and writes
// kx = 0..Kx/2 of each line, M consecutive kz per kx.
//   A[k] = (Z[k] + conj(Z[-k])) / 2,  B[k] = (Z[k] - conj(Z[-k])) / (2i).
// Element loop e = threadIdx.x, +BLOCK, ... < total in chunks of U: the U
// loads of a chunk are issued before any of them is consumed, so a thread
// keeps U global loads in flight instead of waiting on each in turn.
template <int U, class T, int NT = BLOCK, class Ld, class St>
__device__ inline void staged(int total, Ld ld, St st) {
  for (int e0 = threadIdx.x; e0 < total; e0 += U * NT) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * NT;
      if (e < total) v[u] = ld(e);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * NT;
      if (e < total) st(e, v[u]);
    }
  }
}

// Copies the twiddle table W_L^k (k < L) into LDS.
template <class C>
__device__ inline void load_twiddles(C* __restrict__ twl, const C* __restrict__ tw, int L) {
  for (int e = threadIdx.x; e < L; e += blockDim.x) twl[e] = tw[e];
}

// LDS layout of the x passes (in units of C): b0 = P (L+1), b1 = max(P (L+1),
// M (H+1)) (also the C2R staging area), then the L twiddles.
__host__ __device__ inline int xpass_b1(int M, int L) {
  const int P = M / 2, H = L / 2 + 1;
  return P * (L + 1) > M * (H + 1) ? P * (L + 1) : M * (H + 1);
}

// K7a: R2C along x.  Block (bz, y) transforms the M real lines
// (y, z = bz*M .. bz*M+M-1), two per complex FFT (a + i b), and writes
// kx = 0..Kx/2 of each line, M consecutive kz per kx.
//   A[k] = (Z[k] + conj(Z[-k])) / 2,  B[k] = (Z[k] - conj(Z[-k])) / (2i).
template <class C, int L_CT>
__global__ void __launch_bounds__(BLOCK) k_fft_x_r2c(Geom g, FftPlan pl, const C* __restrict__ tw,
                                                     const decltype(C{}.x)* __restrict__ qgrid,
                                                     C* __restrict__ spec, int M) {
  using R = decltype(C{}.x);
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = g.K[0], H = g.H, Ky = g.K[1];
  const int P = M / 2, sx = L + 1;
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + P * sx;
  C* twl = b1 + xpass_b1(M, L);
  load_twiddles(twl, tw, L);
  const int z0 = blockIdx.x * M, y = blockIdx.y;
  staged<8, C>(
      P * L,
      [&](int e) {
        const int p = e / L, xk = e - p * L;
        const size_t row = (size_t(z0 + 2 * p) * Ky + y) * L;
        C v;
        v.x = qgrid[row + xk];
        v.y = qgrid[row + size_t(Ky) * L + xk];
        return v;
      },
      [&](int e, C v) {
        const int p = e / L, xk = e - p * L;
        b0[p * sx + xk] = v;
      });
  __syncthreads();
  const C* res = lds_fft_any<C, L_CT>(b0, b1, P, L, sx, pl, twl, -1);
  for (int e = threadIdx.x; e < M * H; e += blockDim.x) {
    const int kx = e / M, t = e - kx * M;
    const int p = t >> 1;
    const C z = res[p * sx + kx];
    const C w = res[p * sx + (kx == 0 ? 0 : L - kx)];
    C v;
    if ((t & 1) == 0) {
      v.x = R(0.5) * (z.x + w.x);
      v.y = R(0.5) * (z.y - w.y);
    } else {
      v.x = R(0.5) * (z.y + w.y);
      v.y = R(0.5) * (w.x - z.x);
    }
    spec[spec_index(g, kx, y, z0 + t)] = v;
  }
}

// K7b: FFT along y of the half spectrum.  Block (bz, kx) transforms the
// columns kz = bz*M .. bz*M+M-1 (contiguous, coalesced) at that kx.
// LDS: two buffers of M (L+1), then the L twiddles.
template <class C, int L_CT>
__global__ void __launch_bounds__(BLOCK) k_fft_y(Geom g, FftPlan pl, const C* __restrict__ tw,
                                                 C* __restrict__ spec, int M, int stride, int sign) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = g.K[1], Kz = g.K[2];
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + M * stride;
  C* twl = b1 + M * stride;
  load_twiddles(twl, tw, L);
  const int z0 = blockIdx.x * M;
  const size_t base = size_t(blockIdx.y) * L * Kz + z0;
  staged<8, C>(
      M * L,
      [&](int e) {
        const int l = e / M, t = e - l * M;
        return spec[base + size_t(l) * Kz + t];
      },
      [&](int e, C v) {
        const int l = e / M, t = e - l * M;
        b0[t * stride + l] = v;
      });
  __syncthreads();
  const C* res = lds_fft_any<C, L_CT>(b0, b1, M, L, stride, pl, twl, sign);
  for (int e = threadIdx.x; e < M * L; e += blockDim.x) {
    const int l = e / M, t = e - l * M;
    spec[base + size_t(l) * Kz + t] = res[t * stride + l];
  }
}

// K7c/K8: along z (contiguous lines of the half spectrum), fused: forward
// FFT, multiply by the influence table G (same layout, 1/Ng folded in),
// accumulate sum w_kx G |X|^2 into epart[block], inverse FFT.
// LDS: two buffers of M L, then the L twiddles.
template <class C, int L_CT>
__global__ void __launch_bounds__(BLOCK) k_fft_z_fused(Geom g, FftPlan pl,
                                                       const C* __restrict__ tw,
                                                       C* __restrict__ spec, int M,
                                                       const decltype(C{}.x)* __restrict__ G,
                                                       double* __restrict__ epart) {
  using R = decltype(C{}.x);
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double lds_red[BLOCK / WAVE];
  const int L = g.K[2], Ky = g.K[1];
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + M * L;
  C* twl = b1 + M * L;
  load_twiddles(twl, tw, L);
  const int nlines = g.H * Ky;
  const int line0 = blockIdx.x * M;
  staged<8, C>(
      M * L,
      [&](int e) {
        const int t = e / L, k = e - t * L;
        C v; v.x = R(0); v.y = R(0);
        if (line0 + t < nlines) v = spec[size_t(line0 + t) * L + k];
        return v;
      },
      [&](int e, C v) { b0[e] = v; });
  __syncthreads();
  C* res = lds_fft_any<C, L_CT>(b0, b1, M, L, L, pl, twl, -1);
  double en[1] = {0.0};
  staged<8, R>(
      M * L,
      [&](int e) {
        const int line = line0 + e / L;
        return line < nlines ? G[size_t(line0) * L + e] : R(0);
      },
      [&](int e, R gv) {
        const int line = line0 + e / L;
        C v = res[e];
        if (line < nlines) {
          const int kx = line / Ky;
          const double wx = (kx == 0 || 2 * kx == g.K[0]) ? 1.0 : 2.0;
          en[0] += wx * double(gv) * (double(v.x) * double(v.x) + double(v.y) * double(v.y));
        }
        v.x *= gv; v.y *= gv;
        res[e] = v;
      });
  __syncthreads();
  C* other = (res == b0) ? b1 : b0;
  res = lds_fft_any<C, L_CT>(res, other, M, L, L, pl, twl, +1);
  block_sum<1>(en, lds_red);
  if (threadIdx.x == 0) epart[blockIdx.x] = en[0];
  for (int e = threadIdx.x; e < M * L; e += blockDim.x) {
    if (line0 + e / L < nlines) spec[size_t(line0) * L + e] = res[e];
  }
}

// Register-resident transform of one 512-point line by one wave: lane j
// holds elements j + 64 r (r < 8) on entry and on return.  The three radix-8
// Stockham stages of fft_stage_ct for L = 512 (Ns = 1, 8, 64), with the same
// twiddles and butterflies, so the result is that of the LDS transform; the
// data cross lanes twice, through the wave's LDS line x (padded by one
// element in nine against bank conflicts).
__device__ inline int pad9(int e) { return e + (e >> 3); }
template <class C>
__device__ inline void reg_fft512(C (&v)[8], C* __restrict__ x, const C* __restrict__ twl,
                                  int j, int sign) {
  Dft<C, 8>::run(v, sign);  // Ns = 1: no twiddles; output 8 j + r
#pragma unroll
  for (int r = 0; r < 8; ++r) x[pad9(8 * j + r)] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = x[pad9(j + 64 * r)];
  __syncthreads();
  {  // Ns = 8: twiddle W^{8 r k}, k = j mod 8; output (j/8) 64 + k + 8 r
    const int k = j & 7;
#pragma unroll
    for (int r = 1; r < 8; ++r)
      if (k > 0) {
        C w = twl[r * k * 8];
        if (sign > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    Dft<C, 8>::run(v, sign);
    const int o = (j >> 3) * 64 + k;
#pragma unroll
    for (int r = 0; r < 8; ++r) x[pad9(o + 8 * r)] = v[r];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = x[pad9(j + 64 * r)];
  __syncthreads();
  // Ns = 64: twiddle W^{r j}; output j + 64 r, the lane's own elements
#pragma unroll
  for (int r = 1; r < 8; ++r)
    if (j > 0) {
      C w = twl[r * j];
      if (sign > 0) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  Dft<C, 8>::run(v, sign);
}

// Register-resident transform of one 256-point line by one wave: lane j
// holds elements j + 64 r (r < 4); four radix-4 Stockham stages (Ns = 1, 4,
// 16, 64) with three exchanges through the wave's padded LDS line.  A
// different factorisation from the tiled pass (8 8 4), so FP32 results
// differ from it by rounding; used on FP32 chains only.
template <class C>
__device__ inline void reg_fft256(C (&v)[4], C* __restrict__ x, const C* __restrict__ twl,
                                  int j, int sign) {
  Dft<C, 4>::run(v, sign);  // Ns = 1: output 4 j + r
#pragma unroll
  for (int r = 0; r < 4; ++r) x[pad9(4 * j + r)] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = x[pad9(j + 64 * r)];
  __syncthreads();
#pragma unroll
  for (int st = 1; st < 4; ++st) {
    const int Ns = 1 << (2 * st);       // 4, 16, 64
    const int k = j & (Ns - 1), tws = 64 / Ns;
#pragma unroll
    for (int r = 1; r < 4; ++r)
      if (k > 0) {
        C w = twl[r * k * tws];
        if (sign > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    Dft<C, 4>::run(v, sign);
    if (st == 3) break;  // Ns = 64: output j + 64 r, the lane's own elements
    const int o = (j / Ns) * Ns * 4 + k;
#pragma unroll
    for (int r = 0; r < 4; ++r) x[pad9(o + r * Ns)] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = x[pad9(j + 64 * r)];
    __syncthreads();
  }
}

// K7c/K8 for Kz = 256 on FP32 chains: as k_fft_z512_reg with 256-point lines.
template <class C>
__global__ void __launch_bounds__(BLOCK) k_fft_z256_reg(Geom g, const C* __restrict__ tw,
                                                        C* __restrict__ spec,
                                                        const decltype(C{}.x)* __restrict__ G,
                                                        double* __restrict__ epart) {
  using R = decltype(C{}.x);
  constexpr int L = 256, NW = BLOCK / WAVE;
  __shared__ C twl[L];
  __shared__ C xb[NW][L + L / 8];
  __shared__ double lds_red[NW];
  load_twiddles(twl, tw, L);
  const int j = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int Ky = g.K[1], nlines = g.H * Ky;
  const int line = blockIdx.x * NW + w;
  const bool live = line < nlines;
  const size_t base = size_t(live ? line : 0) * L;
  C v[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    v[r].x = R(0); v[r].y = R(0);
    if (live) v[r] = spec[base + j + 64 * r];
  }
  __syncthreads();  // twiddles in LDS
  reg_fft256(v, xb[w], twl, j, -1);
  double en[1] = {0.0};
  const int kx = live ? line / Ky : 0;
  const double wx = (kx == 0 || 2 * kx == g.K[0]) ? 1.0 : 2.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const R gv = live ? G[base + j + 64 * r] : R(0);
    if (live) en[0] += wx * double(gv) * (double(v[r].x) * double(v[r].x) + double(v[r].y) * double(v[r].y));
    v[r].x *= gv; v[r].y *= gv;
  }
  reg_fft256(v, xb[w], twl, j, +1);
  block_sum<1>(en, lds_red);
  if (threadIdx.x == 0) epart[blockIdx.x] = en[0];
  if (live)
#pragma unroll
    for (int r = 0; r < 4; ++r) spec[base + j + 64 * r] = v[r];
}

// K7c/K8 for Kz = 512: k_fft_z_fused with one line per wave held in
// registers (two LDS exchanges per transform instead of the three stage
// passes through LDS of the tiled kernel).  Four lines per workgroup; the
// energy partial of each workgroup goes to epart[block].
template <class C>
__global__ void __launch_bounds__(BLOCK) k_fft_z512_reg(Geom g, const C* __restrict__ tw,
                                                        C* __restrict__ spec,
                                                        const decltype(C{}.x)* __restrict__ G,
                                                        double* __restrict__ epart) {
  using R = decltype(C{}.x);
  constexpr int L = 512, NW = BLOCK / WAVE;
  __shared__ C twl[L];
  __shared__ C xb[NW][L + L / 8];
  __shared__ double lds_red[NW];
  load_twiddles(twl, tw, L);
  const int j = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int Ky = g.K[1], nlines = g.H * Ky;
  const int line = blockIdx.x * NW + w;
  const bool live = line < nlines;
  const size_t base = size_t(live ? line : 0) * L;
  C v[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    v[r].x = R(0); v[r].y = R(0);
    if (live) v[r] = spec[base + j + 64 * r];
  }
  __syncthreads();  // twiddles in LDS
  reg_fft512(v, xb[w], twl, j, -1);
  double en[1] = {0.0};
  const int kx = live ? line / Ky : 0;
  const double wx = (kx == 0 || 2 * kx == g.K[0]) ? 1.0 : 2.0;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const R gv = live ? G[base + j + 64 * r] : R(0);
    if (live) en[0] += wx * double(gv) * (double(v[r].x) * double(v[r].x) + double(v[r].y) * double(v[r].y));
    v[r].x *= gv; v[r].y *= gv;
  }
  reg_fft512(v, xb[w], twl, j, +1);
  block_sum<1>(en, lds_red);
  if (threadIdx.x == 0) epart[blockIdx.x] = en[0];
  if (live)
#pragma unroll
    for (int r = 0; r < 8; ++r) spec[base + j + 64 * r] = v[r];
}

// K7d: C2R along x (inverse, unnormalised).  Block (bz, y) gathers the half
// lines (y, z = bz*M ..), stages them in LDS, forms Z = A + i B over the full
// Hermitian line (A[L-k] = conj(A[k])), inverse-transforms, and writes Re Z
// and Im Z as the two real lines.
template <class C, int L_CT>
__global__ void __launch_bounds__(BLOCK) k_fft_x_c2r(Geom g, FftPlan pl, const C* __restrict__ tw,
                                                     const C* __restrict__ spec,
                                                     decltype(C{}.x)* __restrict__ out, int M) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = g.K[0], H = g.H, Ky = g.K[1];
  const int P = M / 2, sx = L + 1, hs = H + 1;
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + P * sx;
  C* twl = b1 + xpass_b1(M, L);
  load_twiddles(twl, tw, L);
  const int z0 = blockIdx.x * M, y = blockIdx.y;
  staged<8, C>(
      M * H,
      [&](int e) {
        const int kx = e / M, t = e - kx * M;
        return spec[spec_index(g, kx, y, z0 + t)];
      },
      [&](int e, C v) {
        const int kx = e / M, t = e - kx * M;
        b1[t * hs + kx] = v;
      });
  __syncthreads();
  for (int e = threadIdx.x; e < P * L; e += blockDim.x) {
    const int p = e / L, k = e - p * L;
    C a, b;
    if (k < H) {
      a = b1[(2 * p) * hs + k];
      b = b1[(2 * p + 1) * hs + k];
    } else {
      a = b1[(2 * p) * hs + (L - k)]; a.y = -a.y;
      b = b1[(2 * p + 1) * hs + (L - k)]; b.y = -b.y;
    }
    C z;
    z.x = a.x - b.y;
    z.y = a.y + b.x;
    b0[p * sx + k] = z;
  }
  __syncthreads();
  const C* res = lds_fft_any<C, L_CT>(b0, b1, P, L, sx, pl, twl, +1);
  for (int e = threadIdx.x; e < P * L; e += blockDim.x) {
    const int p = e / L, xk = e - p * L;
    const size_t row = (size_t(z0 + 2 * p) * Ky + y) * L;
    const C v = res[p * sx + xk];
    out[row + xk] = v.x;
    out[row + size_t(Ky) * L + xk] = v.y;
  }
}

// ---------------------------------------------------------------------------
// Plane-resident FFT chain of the FP32 coarse grids (Kx = Ky = L <= 128): one
// workgroup holds a whole z plane in LDS, so x and y are transformed in one
// pass in each direction, and the z pass reads the plane-major half spectrum
//   spec2[(kz * Ky + ky) * H + kx]
// in tiles of NL consecutive (ky, kx) lines.  Three launches per convolution
// instead of five, each reading and writing the grid once.  The same
// transforms as the five-pass chain; the real x lines are paired along y
// (rows 2p, 2p + 1) instead of along z, so FP32 results differ from it by
// rounding.
// ---------------------------------------------------------------------------
// LDS: two buffers of H S lines (S = pad_stride(L); the x lines need P S <=
// H S), then the L twiddles (the x and y tables are equal for Kx = Ky).
__host__ __device__ constexpr int plane_buf(int L) { return (L / 2 + 1) * pad_stride(L); }

struct F2x2 { float4 a, b; };

// Forward: real plane z -> plane of the half spectrum (x R2C, then y FFT).
template <int L, int NT>
__global__ void __launch_bounds__(NT) k_fft_xy_r2c(const float2* __restrict__ tw,
                                                      const float* __restrict__ qgrid,
                                                      float2* __restrict__ spec2) {
  using C = float2;
  constexpr int P = L / 2, H = L / 2 + 1, S = pad_stride(L), L4 = L / 4;
  extern __shared__ __align__(16) unsigned char smem[];
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + plane_buf(L);
  C* twl = b1 + plane_buf(L);
  load_twiddles(twl, tw, L);
  const size_t plane = size_t(blockIdx.x) * L * L;
  const float4* q4 = reinterpret_cast<const float4*>(qgrid + plane);
  // rows 2p (real part) and 2p + 1 (imaginary part) form complex line p
  staged<8, F2x2, NT>(
      P * L4,
      [&](int e) {
        const int p = e / L4, x4 = e - p * L4;
        F2x2 v;
        v.a = q4[(2 * p) * L4 + x4];
        v.b = q4[(2 * p + 1) * L4 + x4];
        return v;
      },
      [&](int e, F2x2 v) {
        const int p = e / L4, x = 4 * (e - p * L4);
        C* d = b0 + p * S + x;
        d[0] = make_float2(v.a.x, v.b.x);
        d[1] = make_float2(v.a.y, v.b.y);
        d[2] = make_float2(v.a.z, v.b.z);
        d[3] = make_float2(v.a.w, v.b.w);
      });
  __syncthreads();
  C* res = lds_fft_pad<C, L, NT>(b0, b1, P, S, twl, -1);
  C* col = (res == b0) ? b1 : b0;
  // split the pair spectra: column kx of row y at col[kx S + y]
  //   A[k] = (Z[k] + conj(Z[-k])) / 2,  B[k] = (Z[k] - conj(Z[-k])) / (2i)
  for (int e = threadIdx.x; e < H * L; e += NT) {
    const int kx = e / L, y = e - kx * L;
    const int p = y >> 1;
    const C z = res[p * S + kx];
    const C w = res[p * S + (kx == 0 ? 0 : L - kx)];
    C v;
    if ((y & 1) == 0) {
      v.x = 0.5f * (z.x + w.x);
      v.y = 0.5f * (z.y - w.y);
    } else {
      v.x = 0.5f * (z.y + w.y);
      v.y = 0.5f * (w.x - z.x);
    }
    col[kx * S + y] = v;
  }
  __syncthreads();
  const C* out = lds_fft_pad<C, L, NT>(col, res, H, S, twl, -1);
  float2* dst = spec2 + size_t(blockIdx.x) * L * H;
  for (int e = threadIdx.x; e < L * H; e += NT) {
    const int ky = e / H, kx = e - ky * H;
    dst[e] = out[kx * S + ky];
  }
}

// Inverse: plane of the half spectrum -> real plane (y inverse FFT, then x
// C2R), unnormalised.
template <int L, int NT>
__global__ void __launch_bounds__(NT) k_fft_xy_c2r(const float2* __restrict__ tw,
                                                      const float2* __restrict__ spec2,
                                                      float* __restrict__ out) {
  using C = float2;
  constexpr int P = L / 2, H = L / 2 + 1, S = pad_stride(L), L4 = L / 4;
  constexpr int NP = L * H;  // complex values per plane (even: L is a multiple of 4)
  extern __shared__ __align__(16) unsigned char smem[];
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + plane_buf(L);
  C* twl = b1 + plane_buf(L);
  load_twiddles(twl, tw, L);
  const float4* s4 = reinterpret_cast<const float4*>(spec2 + size_t(blockIdx.x) * NP);
  staged<8, float4, NT>(
      NP / 2, [&](int e) { return s4[e]; },
      [&](int e, float4 v) {
        int ky = (2 * e) / H, kx = 2 * e - ky * H;
        b0[kx * S + ky] = make_float2(v.x, v.y);
        if (++kx == H) { kx = 0; ++ky; }
        b0[kx * S + ky] = make_float2(v.z, v.w);
      });
  __syncthreads();
  C* res = lds_fft_pad<C, L, NT>(b0, b1, H, S, twl, +1);
  C* lin = (res == b0) ? b1 : b0;
  // Z = A + i B over the full Hermitian x line, A = row 2p, B = row 2p + 1
  for (int e = threadIdx.x; e < P * L; e += NT) {
    const int p = e / L, k = e - p * L;
    C a, b;
    if (k < H) {
      a = res[k * S + 2 * p];
      b = res[k * S + 2 * p + 1];
    } else {
      a = res[(L - k) * S + 2 * p]; a.y = -a.y;
      b = res[(L - k) * S + 2 * p + 1]; b.y = -b.y;
    }
    C z;
    z.x = a.x - b.y;
    z.y = a.y + b.x;
    lin[p * S + k] = z;
  }
  __syncthreads();
  const C* r = lds_fft_pad<C, L, NT>(lin, res, P, S, twl, +1);
  float4* o4 = reinterpret_cast<float4*>(out + size_t(blockIdx.x) * L * L);
  for (int e = threadIdx.x; e < P * L4; e += NT) {
    const int p = e / L4, x = 4 * (e - p * L4);
    const C* s = r + p * S + x;
    o4[(2 * p) * L4 + x / 4] = make_float4(s[0].x, s[1].x, s[2].x, s[3].x);
    o4[(2 * p + 1) * L4 + x / 4] = make_float4(s[0].y, s[1].y, s[2].y, s[3].y);
  }
}

// z pass on the plane-major spectrum: workgroup b takes lines
// NL b .. NL b + NL - 1 of the Ky H (ky, kx) lines, forward FFT along z,
// multiply by the influence or filter table (same layout), sum
// w_kx G |X|^2 into epart[b], inverse FFT.  LDS: two buffers of NL lines of
// stride pad_stride(KZ), then the KZ twiddles.
template <int KZ, int NL>
__global__ void __launch_bounds__(BLOCK) k_fft_zp(Geom g, const float2* __restrict__ tw,
                                                  float2* __restrict__ spec2,
                                                  const float* __restrict__ G,
                                                  double* __restrict__ epart) {
  using C = float2;
  constexpr int S = pad_stride(KZ);
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double lds_red[BLOCK / WAVE];
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + NL * S;
  C* twl = b1 + NL * S;
  load_twiddles(twl, tw, KZ);
  const int H = g.H, nlines = g.K[1] * H;
  const int line0 = xcd_tile(blockIdx.x, gridDim.x) * NL;
  staged<8, C>(
      NL * KZ,
      [&](int e) {
        const int kz = e / NL, t = e - kz * NL;
        C v = make_float2(0.f, 0.f);
        if (line0 + t < nlines) v = spec2[size_t(kz) * nlines + line0 + t];
        return v;
      },
      [&](int e, C v) {
        const int kz = e / NL, t = e - kz * NL;
        b0[t * S + kz] = v;
      });
  __syncthreads();
  C* res = lds_fft_pad<C, KZ>(b0, b1, NL, S, twl, -1);
  double en[1] = {0.0};
  staged<8, float>(
      NL * KZ,
      [&](int e) {
        const int kz = e / NL, t = e - kz * NL;
        return line0 + t < nlines ? G[size_t(kz) * nlines + line0 + t] : 0.f;
      },
      [&](int e, float gv) {
        const int kz = e / NL, t = e - kz * NL;
        C v = res[t * S + kz];
        if (line0 + t < nlines) {
          const int kx = (line0 + t) % H;
          const double wx = (kx == 0 || 2 * kx == g.K[0]) ? 1.0 : 2.0;
          en[0] += wx * double(gv) * (double(v.x) * double(v.x) + double(v.y) * double(v.y));
        }
        v.x *= gv; v.y *= gv;
        res[t * S + kz] = v;
      });
  __syncthreads();
  C* other = (res == b0) ? b1 : b0;
  res = lds_fft_pad<C, KZ>(res, other, NL, S, twl, +1);
  block_sum<1>(en, lds_red);
  if (threadIdx.x == 0) epart[blockIdx.x] = en[0];
  for (int e = threadIdx.x; e < NL * KZ; e += BLOCK) {
    const int kz = e / NL, t = e - kz * NL;
    if (line0 + t < nlines) spec2[size_t(kz) * nlines + line0 + t] = res[t * S + kz];
  }
}

// Table in the half-spectrum layout [kx][ky][kz] -> plane-major [kz][ky][kx].
template <class R>
__global__ void k_to_plane_major(Geom g, const R* __restrict__ src, R* __restrict__ dst) {
  const int H = g.H, Ky = g.K[1], Kz = g.K[2];
  const size_t total = size_t(H) * Ky * Kz;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int kx = int(t % H);
  const int ky = int((t / H) % Ky);
  const int kz = int(t / (size_t(H) * Ky));
  dst[t] = src[(size_t(kx) * Ky + ky) * Kz + kz];
}

// K14: influence table on the half spectrum, G(k)/Ng with
// G = psi(k) 36^3 / prod_d (4 + 2 cos(2 pi k_d / K_d))^2 (force.f90:414-420,
// 500-514); G(0) = 0 (quirk Q5).
template <class R>
__global__ void k_influence(Geom g, double a2, R* __restrict__ G, int order) {
  const size_t total = size_t(g.H) * g.K[1] * g.K[2];
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int kz = int(t % g.K[2]);
  const int ky = int((t / g.K[2]) % g.K[1]);
  const int kx = int(t / (size_t(g.K[2]) * g.K[1]));
  const int k3[3] = {kx, ky, kz};
  int m[3];
  double zf = 1.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int K = g.K[d];
    m[d] = (k3[d] + K / 2) % K - K / 2;
    // |b(k)|^-2 of the order-4 B-spline; 1 for order 2 (linear weights)
    double z = order == 4 ? 4.0 + 2.0 * cos(2.0 * PI * double(k3[d]) / double(K)) : 6.0;
    zf *= z * z;
  }
  double val = 0.0;
  if (m[0] * m[0] + m[1] * m[1] + m[2] * m[2] != 0) {
    const double m2 = double(m[0] * m[0]) / (g.box[0] * g.box[0]) +
                      double(m[1] * m[1]) / (g.box[1] * g.box[1]) +
                      double(m[2] * m[2]) / (g.box[2] * g.box[2]);
    const double psi = exp(-m2 * (PI * PI) / a2) / m2;
    val = psi * 46656.0 / zf / (double(g.K[0]) * double(g.K[1]) * double(g.K[2]));
  }
  G[t] = R(val);
}

// K9a: reciprocal forces F_i = -c sum grad(w) Qfin (force.f90:439-456) into
// frec; fpart[3*block + d] = block sums for the mean-force removal.
template <class Rg>
__global__ void __launch_bounds__(BLOCK) k_gather_force(Geom g, const double* __restrict__ x,
                                                        const double* __restrict__ q,
                                                        const Rg* __restrict__ grid, double cst,
                                                        double* __restrict__ frec,
                                                        double* __restrict__ fpart) {
  __shared__ double lds[3 * BLOCK / WAVE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ff[3] = {0.0, 0.0, 0.0};
  if (i < g.n) {
    double u[3], AX[3][4], AXX[3][4];
    int b[3];
    grid_coord(g, x + 3 * i, u, b);
#pragma unroll
    for (int d = 0; d < 3; ++d) bspline4(u[d], AX[d], AXX[d]);
    const double kx[3] = {double(g.K[0]) / g.box[0], double(g.K[1]) / g.box[1],
                          double(g.K[2]) / g.box[2]};
    const double qi = q[i];
    int ix[4], iy[4], iz[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      ix[a] = wrapk(b[0], a, g.K[0]);
      iy[a] = wrapk(b[1], a, g.K[1]);
      iz[a] = wrapk(b[2], a, g.K[2]);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int bb = 0; bb < 4; ++bb) {
        const size_t row = size_t(g.K[0]) * (size_t(iy[bb]) + size_t(g.K[1]) * size_t(iz[c]));
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const double Qr = double(grid[row + ix[a]]);
          const double d0 = qi * AXX[0][a] * AX[1][bb] * AX[2][c] * kx[0];
          const double d1 = qi * AX[0][a] * AXX[1][bb] * AX[2][c] * kx[1];
          const double d2 = qi * AX[0][a] * AX[1][bb] * AXX[2][c] * kx[2];
          ff[0] -= cst * d0 * Qr;
          ff[1] -= cst * d1 * Qr;
          ff[2] -= cst * d2 * Qr;
        }
      }
    frec[3 * i] = ff[0];
    frec[3 * i + 1] = ff[1];
    frec[3 * i + 2] = ff[2];
  }
  block_sum<3>(ff, lds);
  if (threadIdx.x == 0) {
    fpart[3 * blockIdx.x] = ff[0];
    fpart[3 * blockIdx.x + 1] = ff[1];
    fpart[3 * blockIdx.x + 2] = ff[2];
  }
}

// f += frec - mean(frec) (force.f90:457-460); fmean holds the three sums.
__global__ void k_pme_apply(int n, const double* __restrict__ frec,
                            const double* __restrict__ fsum, double* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) f[3 * i + d] += frec[3 * i + d] - fsum[d] / double(n);
}

// phi_i = c_pme sum_w w Qfin (PHI_MESH gather, ctip.f90:381-393, scaled ctip.f90:170)
template <class Rg>
__device__ inline double gather_phi(const Geom& g, const double* xi, const Rg* __restrict__ grid) {
  double u[3], AX[3][4], AXX[3][4];
  int b[3];
  grid_coord(g, xi, u, b);
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline4(u[d], AX[d], AXX[d]);
  int ix[4], iy[4], iz[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    ix[a] = wrapk(b[0], a, g.K[0]);
    iy[a] = wrapk(b[1], a, g.K[1]);
    iz[a] = wrapk(b[2], a, g.K[2]);
  }
  double e = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const size_t gi = size_t(ix[a]) + size_t(g.K[0]) * (size_t(iy[bb]) + size_t(g.K[1]) * size_t(iz[c]));
        e += AX[0][a] * AX[1][bb] * AX[2][c] * double(grid[gi]);
      }
  return e;
}

// Linear (order-2, cloud-in-cell) assignment, used on the charge solver's
// preconditioner grid: weights (1 - f, f) at base points b and b + 1,
// b = INT(u), f = u - b.
__device__ inline void cic_weights(const Geom& g, const double* xi, double W[3][2], int ix[3][2]) {
  double u[3];
  int b[3];
  grid_coord(g, xi, u, b);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double f = u[d] - double(b[d]);
    W[d][0] = 1.0 - f;
    W[d][1] = f;
    int r = b[d] % g.K[d];
    if (r < 0) r += g.K[d];
    ix[d][0] = r;
    ix[d][1] = r + 1 == g.K[d] ? 0 : r + 1;
  }
}

template <class Rg>
__device__ inline double gather_cic(const Geom& g, const double* xi, const Rg* __restrict__ grid) {
  double W[3][2];
  int ix[3][2];
  cic_weights(g, xi, W, ix);
  double e = 0.0;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) {
      const size_t row = size_t(g.K[0]) * (size_t(ix[1][bb]) + size_t(g.K[1]) * size_t(ix[2][c]));
      e += W[1][bb] * W[2][c] * (W[0][0] * double(grid[row + ix[0][0]]) +
                                 W[0][1] * double(grid[row + ix[0][1]]));
    }
  return e;
}

// gather of the given assignment order (4: the reference's B-splines, 2: linear)
template <int ORD, class Rg>
__device__ inline double gather_ord(const Geom& g, const double* xi, const Rg* __restrict__ grid) {
  if constexpr (ORD == 4) return gather_phi(g, xi, grid);
  else return gather_cic(g, xi, grid);
}

}  // namespace sx

namespace sx {

// ---------------------------------------------------------------------------
// Brick-tiled spread.  Per set of positions, atoms are binned into bricks of
// 4x4x4 grid points by their base point INT(u) mod K (grid sizes are multiples
// of 4) and listed brick by brick in ascending atom order.  Per evaluation a
// workgroup owns one tile of TX x TY x TZ grid points: it walks the bricks
// whose atoms can reach the tile (base within [lo-2, hi+1]), accumulates
// q w into the tile in LDS with 64-bit integer atomics at the fixed-point
// scale (order-independent, hence deterministic), and writes every grid point
// of the tile exactly once.  No global atomics and no zeroing pass.
// ---------------------------------------------------------------------------
constexpr int BRK = 4;
constexpr int TX = 32, TY = 16, TZ = 8;
constexpr int TILE = TX * TY * TZ;
constexpr int MAXB = (TX / 4 + 2) * (TY / 4 + 2) * (TZ / 4 + 2);  // candidate bricks per tile
static_assert(MAXB <= BLOCK, "one thread per candidate brick");

__global__ void k_brick_assign(Geom g, const double* __restrict__ x, int* __restrict__ brick,
                               int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  double u[3];
  int b[3];
  grid_coord(g, x + 3 * i, u, b);
  int bi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int r = b[d] % g.K[d];
    if (r < 0) r += g.K[d];
    bi[d] = r / BRK;
  }
  const int nbx = g.K[0] / BRK, nby = g.K[1] / BRK;
  const int id = bi[0] + nbx * (bi[1] + nby * bi[2]);
  brick[i] = id;
  atomicAdd(&count[id], 1);
}

// slot of atom i in its brick = number of atoms of the same brick with a
// smaller index (bricks hold a few atoms); writes border[start + slot] = i.
__global__ void k_brick_place(int n, const int* __restrict__ brick, const int* __restrict__ start,
                              const int* __restrict__ fillpos, const int* __restrict__ unordered,
                              int* __restrict__ border) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = brick[i];
  const int s0 = start[b], s1 = start[b + 1];
  int rank = 0;
  for (int k = s0; k < s1; ++k) rank += unordered[k] < i;
  border[s0 + rank] = i;
  (void)fillpos;
}

__global__ void k_brick_scatter(int n, const int* __restrict__ brick, const int* __restrict__ start,
                                int* __restrict__ fill, int* __restrict__ unordered) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = brick[i];
  unordered[start[b] + atomicAdd(&fill[b], 1)] = i;
}

// Candidate slot table of a spread tile: entry f = (candidate brick k << 8)
// | offset of the tile's f-th candidate atom within brick k, when the tile has
// at most SPREAD_MAXC candidates and no brick holds more than 256 atoms (else
// the slot is found by a binary search over the prefix array bofs).
constexpr int SPREAD_MAXC = 6144;
static_assert(MAXB <= 256, "candidate brick index in 8 bits");
__device__ inline bool spread_slot_table(int ncand, int total, const int* bofs,
                                         unsigned short* ctab, int* bad) {
  if (threadIdx.x == 0) *bad = total > SPREAD_MAXC;
  __syncthreads();
  if (!*bad && int(threadIdx.x) < ncand) {
    const int k = threadIdx.x, o = bofs[k], c = bofs[k + 1] - o;
    if (c > 256) *bad = 1;
    else
      for (int j = 0; j < c; ++j) ctab[o + j] = (unsigned short)((k << 8) | j);
  }
  __syncthreads();
  return !*bad;
}
// Brick-order index (into border) of the tile's f-th candidate.
__device__ inline int spread_slot(int f, bool table, int ncand, const int* bofs, const int* bbeg,
                                  const unsigned short* ctab) {
  if (table) {
    const int e = ctab[f];
    return bbeg[e >> 8] + (e & 255);
  }
  int l = 0, h = ncand - 1;
  while (l < h) {
    const int mid = (l + h + 1) >> 1;
    if (bofs[mid] <= f) l = mid; else h = mid - 1;
  }
  return bbeg[l] + (f - bofs[l]);
}

// K6 (tiled): grid[tile] = sum over atoms of q w, written as R (double/float).
// ORD 4: the reference's order-4 B-splines (points base-1 .. base+2); ORD 2:
// linear weights (points base, base+1) for the preconditioner grid.
template <int ORD, bool F32W, class R>
__global__ void __launch_bounds__(BLOCK) k_spread_tile(Geom g, const double* __restrict__ x,
                                                       const double* __restrict__ q,
                                                       const int* __restrict__ bstart,
                                                       const int* __restrict__ border,
                                                       R* __restrict__ grid) {
  __shared__ unsigned long long acc[TILE];
  __shared__ unsigned short ctab[SPREAD_MAXC];  // candidate f: (candidate brick << 8) | offset
  __shared__ int tab_bad;
  __shared__ int bofs[MAXB + 1];
  __shared__ int bbeg[MAXB];
  __shared__ int wsum[BLOCK / WAVE];
  const int ntx = (g.K[0] + TX - 1) / TX, nty = (g.K[1] + TY - 1) / TY;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  const int lo[3] = {tx * TX, ty * TY, tz * TZ};
  const int ext[3] = {min(TX, g.K[0] - lo[0]), min(TY, g.K[1] - lo[1]), min(TZ, g.K[2] - lo[2])};
  const int nb[3] = {g.K[0] / BRK, g.K[1] / BRK, g.K[2] / BRK};
  // candidate brick ranges per axis
  int fb[3], cb[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // bases whose stencil meets the tile: lo-2 .. hi+1 (ORD 4), lo-1 .. hi (ORD 2)
    const int a = lo[d] - (ORD == 4 ? 2 : 1), b = lo[d] + ext[d] - (ORD == 4 ? 0 : 1);
    int f0 = (a >= 0) ? a / BRK : -((-a + BRK - 1) / BRK);
    int l0 = b / BRK;
    int c = l0 - f0 + 1;
    if (c >= nb[d]) { f0 = 0; c = nb[d]; }
    fb[d] = f0;
    cb[d] = c;
  }
  const int ncand = cb[0] * cb[1] * cb[2];
  for (int k = threadIdx.x; k < TILE; k += blockDim.x) acc[k] = 0ull;
  // atom counts of the candidate bricks and their prefix sum (wave scans,
  // then the totals of the preceding waves)
  int cntk = 0;
  if (threadIdx.x < ncand) {
    const int k = threadIdx.x;
    const int ix = k % cb[0], iy = (k / cb[0]) % cb[1], iz = k / (cb[0] * cb[1]);
    int bx = (fb[0] + ix) % nb[0]; if (bx < 0) bx += nb[0];
    int by = (fb[1] + iy) % nb[1]; if (by < 0) by += nb[1];
    int bz = (fb[2] + iz) % nb[2]; if (bz < 0) bz += nb[2];
    const int bid = bx + nb[0] * (by + nb[1] * bz);
    const int b0 = bstart[bid];
    bbeg[k] = b0;
    cntk = bstart[bid + 1] - b0;
  }
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const int t = __shfl_up(cntk, o, WAVE);
    if (lane >= o) cntk += t;
  }
  if (lane == WAVE - 1) wsum[wid] = cntk;
  if (threadIdx.x == 0) bofs[0] = 0;
  __syncthreads();
  for (int w = 0; w < wid; ++w) cntk += wsum[w];
  if (threadIdx.x < ncand) bofs[threadIdx.x + 1] = cntk;
  __syncthreads();
  const int total = bofs[ncand];
  const bool table = spread_slot_table(ncand, total, bofs, ctab, &tab_bad);
  for (int f = threadIdx.x; f < total; f += blockDim.x) {
    const int ai = border[spread_slot(f, table, ncand, bofs, bbeg, ctab)];
    const double qi = q[ai];
    if constexpr (ORD == 2) {
      double W[3][2];
      int ix[3][2];
      cic_weights(g, x + 3 * ai, W, ix);
      int o[3][2];
      bool in[3][2];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          int p = ix[d][a] - lo[d];
          if (p < 0) p += g.K[d];
          o[d][a] = p;
          in[d][a] = p < ext[d];
        }
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
          for (int a = 0; a < 2; ++a)
            if (in[2][c] && in[1][bb] && in[0][a]) {
              const long long v = llrint(qi * (W[0][a] * W[1][bb] * W[2][c]) * GRID_SCALE);
              atomicAdd(&acc[TX * (o[1][bb] + TY * o[2][c]) + o[0][a]], (unsigned long long)v);
            }
      continue;
    }
    double u[3], AX[3][4], AXX[3][4];
    int b[3];
    grid_coord(g, x + 3 * ai, u, b);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if constexpr (F32W) bspline4(u[d], AX[d], AXX[d]);  // the reference's expressions (Q1 replay)
      else bspline4_w(u[d], AX[d]);
    }
    int o[3][4];
    bool in[3][4];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        int p = wrapk(b[d], a, g.K[d]) - lo[d];
        if (p < 0) p += g.K[d];  // periodic distance from the tile origin
        o[d][a] = p;
        in[d][a] = p < ext[d];
      }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (!in[2][c]) continue;
#pragma unroll
      for (int bb = 0; bb < 4; ++bb) {
        if (!in[1][bb]) continue;
        const int row = TX * (o[1][bb] + TY * o[2][c]);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          if (!in[0][a]) continue;
          double w = AX[0][a] * AX[1][bb] * AX[2][c];
          if (F32W) w = double(float(w));
          const long long v = llrint(qi * w * GRID_SCALE);
          atomicAdd(&acc[row + o[0][a]], (unsigned long long)v);
        }
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < TILE; k += blockDim.x) {
    const int px = k % TX, py = (k / TX) % TY, pz = k / (TX * TY);
    if (px < ext[0] && py < ext[1] && pz < ext[2]) {
      const size_t gi = size_t(lo[0] + px) +
                        size_t(g.K[0]) * (size_t(lo[1] + py) + size_t(g.K[1]) * size_t(lo[2] + pz));
      grid[gi] = R(double((long long)acc[k]) * GRID_INV_SCALE);
    }
  }
}

// Stencil records of the 32-bit spread in brick order, once per set of
// positions: for atom border[f], z0_d = 1 - frac(u_d) in FP32 (the B-spline
// argument of k_spread_tile_i32) and the base index INT(u_d) mod K_d in 10
// bits per dimension (K_d <= 1024), so the spread reads 16 contiguous bytes
// per candidate instead of gathering x and redoing the FP64 grid coordinates.
__global__ void k_brick_records(Geom g, int n, const double* __restrict__ x,
                                const int* __restrict__ border, float4* __restrict__ rec,
                                float4* __restrict__ rec_atom) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  double u[3];
  int b[3];
  grid_coord(g, x + 3 * border[f], u, b);
  float z[3];
  unsigned pk = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    z[d] = float(trunc(u[d]) - u[d] + 1.0);
    int bm = b[d] % g.K[d];
    if (bm < 0) bm += g.K[d];
    pk |= unsigned(bm) << (10 * d);
  }
  rec[f] = make_float4(z[0], z[1], z[2], __uint_as_float(pk));
  if (rec_atom) rec_atom[border[f]] = rec[f];  // the same record in atom order
}

// FP32 stencil weights of a record (the ones the 32-bit spread uses), ORD 4:
// the B-splines of z0, ORD 2: z0, 1 - z0; base index of axis d in bm[d].
template <int ORD>
__device__ inline void rec_weights(const float4& rc, float AX[3][ORD], int bm[3]) {
  const float zr[3] = {rc.x, rc.y, rc.z};
  const unsigned pk = __float_as_uint(rc.w);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float z0 = zr[d];
    if constexpr (ORD == 4) {
      const float z1 = z0 + 1.f, z2 = z0 + 2.f, z3 = z0 + 3.f;
      constexpr float s6 = 1.f / 6.f;
      AX[d][0] = z0 * z0 * z0 * s6;
      AX[d][1] = (((-3.f * z1 + 12.f) * z1 - 12.f) * z1 + 4.f) * s6;
      AX[d][2] = (((3.f * z2 - 24.f) * z2 + 60.f) * z2 - 44.f) * s6;
      AX[d][3] = (((-z3 + 12.f) * z3 - 48.f) * z3 + 64.f) * s6;
    } else {
      AX[d][0] = z0;
      AX[d][1] = 1.f - z0;
    }
    bm[d] = int((pk >> (10 * d)) & 1023u);
  }
}

// The linear (ORD 2) gather of the preconditioner's grid at atom i from its
// atom-order record: the transpose of k_spread_tile_i32<2> (same FP32 weights).
__device__ inline double gather_cic_rec(const Geom& g, const float4& rc,
                                        const float* __restrict__ grid) {
  float AX[3][2];
  int bm[3];
  rec_weights<2>(rc, AX, bm);
  int ix[3][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ix[d][0] = bm[d];
    ix[d][1] = bm[d] + 1 == g.K[d] ? 0 : bm[d] + 1;
  }
  float e = 0.f;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) {
      const size_t row = size_t(g.K[0]) * (size_t(ix[1][bb]) + size_t(g.K[1]) * size_t(ix[2][c]));
      e += AX[1][bb] * AX[2][c] * (AX[0][0] * grid[row + ix[0][0]] + AX[0][1] * grid[row + ix[0][1]]);
    }
  return double(e);
}

// K6 in FP32 for the MIXED mode's coarse grids (the inner operator's,
// order-4 B-splines, and the preconditioner's, ORD 2: linear weights w0 = z0,
// w1 = 1 - z0 at b, b + 1): the same tiles and bricks with FP32 weights and
// 32-bit fixed point.  Each tile owns its points and writes each once, so it
// takes its own scale 2^S, from a bound that no grid point's sum can exceed:
// a point's stencil atoms have their bases in at most two bricks per axis,
// so at most 8 bricks of the tile's candidates reach it, and one atom adds at
// most wmax |v| (wmax = (2/3)^3 for the B-splines, 1 for linear weights):
// 8 maxcnt wmax max|v| 2^S < 2^30 (maxcnt: the fullest candidate brick; the
// factor 2 left to 2^31 covers the rounding of each product).  Resolution
// 2^-S, a few 1e-8 of max|v|.  Integer sums keep the grid independent of the
// order of the atomics; the LDS tile is half the size of the 64-bit one.
// Tile shape (TXg x TYg x TZg points, NT threads): 32 x 16 x 8 with 256
// threads, 32 x 32 x 16 with 1024 threads (one workgroup per CU, 2.3 halo
// bricks per owned brick instead of 3.75 for the order-4 stencil), or, for
// grids coarser than the atoms (more atoms than grid points), 16 x 16 x 8
// with 1024 threads, so that the launch still has a workgroup per CU.
template <int ORD, int TXg, int TYg, int TZg, int NT>
struct SpreadI32 {
  static constexpr int TILEg = TXg * TYg * TZg;
  static constexpr int MAXBg = (TXg / BRK + 2) * (TYg / BRK + 2) * (TZg / BRK + 2);
  // slot-table capacity: the big tile's ~10^4 candidates, the small dense
  // tile's (grids with more atoms than points: ~250 atoms per brick) ~2 10^4
  static constexpr int MAXCg = NT != 1024 ? 6144 : (TILEg < 4096 ? 20480 : 16384);
  // slot table entries: 32 bits (brick << 12 | offset) for the big tile,
  // 16 bits (brick << 8 | offset) for the small one
  static constexpr bool WIDE = NT == 1024;
  using Entry = typename std::conditional<WIDE, unsigned, unsigned short>::type;
  static constexpr int OFS_BITS = WIDE ? 12 : 8;
  static_assert(MAXBg <= NT, "one thread per candidate brick");
  static_assert(MAXBg < (1 << (WIDE ? 20 : 8)), "candidate brick index fits the entry");
  static int tiles(const Geom& g) {
    return ((g.K[0] + TXg - 1) / TXg) * ((g.K[1] + TYg - 1) / TYg) * ((g.K[2] + TZg - 1) / TZg);
  }
};

template <int ORD, int TXg, int TYg, int TZg, int NT>
__global__ void __launch_bounds__(NT) k_spread_tile_i32(Geom g, const double* __restrict__ x,
                                                        const double* __restrict__ q,
                                                        const int* __restrict__ bstart,
                                                        const int* __restrict__ border,
                                                        const float4* __restrict__ rec,
                                                        float* __restrict__ grid) {
  static_assert(ORD == 2 || ORD == 4, "linear or cubic B-spline assignment");
  using P = SpreadI32<ORD, TXg, TYg, TZg, NT>;
  constexpr int TILEg = P::TILEg, MAXBg = P::MAXBg, MAXCg = P::MAXCg;
  __shared__ int acc[TILEg];
  using Entry = typename P::Entry;
  constexpr int OB = P::OFS_BITS;
  __shared__ Entry ctab[MAXCg];  // candidate f: (candidate brick << OB) | offset
  __shared__ int tab_bad;
  __shared__ int bofs[MAXBg + 1];
  __shared__ int bbeg[MAXBg];
  __shared__ int wsum[NT / WAVE];
  __shared__ double wmax[NT / WAVE];
  __shared__ int wcnt[NT / WAVE];
  const int ntx = (g.K[0] + TXg - 1) / TXg, nty = (g.K[1] + TYg - 1) / TYg;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  const int lo[3] = {tx * TXg, ty * TYg, tz * TZg};
  const int ext[3] = {min(TXg, g.K[0] - lo[0]), min(TYg, g.K[1] - lo[1]),
                      min(TZg, g.K[2] - lo[2])};
  const int nb[3] = {g.K[0] / BRK, g.K[1] / BRK, g.K[2] / BRK};
  int fb[3], cb[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // bases whose stencil meets the tile: lo-2 .. hi+1 (ORD 4), lo-1 .. hi (ORD 2)
    const int a = lo[d] - (ORD == 4 ? 2 : 1), b = lo[d] + ext[d] - (ORD == 4 ? 0 : 1);
    int f0 = (a >= 0) ? a / BRK : -((-a + BRK - 1) / BRK);
    int l0 = b / BRK;
    int c = l0 - f0 + 1;
    if (c >= nb[d]) { f0 = 0; c = nb[d]; }
    fb[d] = f0;
    cb[d] = c;
  }
  const int ncand = cb[0] * cb[1] * cb[2];
  for (int k = threadIdx.x; k < TILEg; k += blockDim.x) acc[k] = 0;
  int cntk = 0;
  if (threadIdx.x < ncand) {
    const int k = threadIdx.x;
    const int ix = k % cb[0], iy = (k / cb[0]) % cb[1], iz = k / (cb[0] * cb[1]);
    int bx = (fb[0] + ix) % nb[0]; if (bx < 0) bx += nb[0];
    int by = (fb[1] + iy) % nb[1]; if (by < 0) by += nb[1];
    int bz = (fb[2] + iz) % nb[2]; if (bz < 0) bz += nb[2];
    const int bid = bx + nb[0] * (by + nb[1] * bz);
    const int b0 = bstart[bid];
    bbeg[k] = b0;
    cntk = bstart[bid + 1] - b0;
  }
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  int cmax = cntk;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const int v = __shfl_up(cntk, o, WAVE);
    if (lane >= o) cntk += v;
    cmax = max(cmax, __shfl_xor(cmax, o, WAVE));
  }
  if (lane == WAVE - 1) wsum[wid] = cntk;
  if (lane == 0) wcnt[wid] = cmax;
  if (threadIdx.x == 0) bofs[0] = 0;
  __syncthreads();
  for (int w = 0; w < wid; ++w) cntk += wsum[w];
  if (threadIdx.x < ncand) bofs[threadIdx.x + 1] = cntk;
  if (threadIdx.x == 0) tab_bad = 0;
  __syncthreads();
  const int total = bofs[ncand];
  // slot table (else a binary search over bofs per candidate)
  const bool table = total <= MAXCg;
  if (table && int(threadIdx.x) < ncand) {
    const int k = threadIdx.x, o = bofs[k], c = bofs[k + 1] - o;
    if (c > (1 << OB)) tab_bad = 1;
    else
      for (int j = 0; j < c; ++j) ctab[o + j] = Entry((unsigned(k) << OB) | unsigned(j));
  }
  __syncthreads();
  const bool use_tab = table && !tab_bad;
  auto slot_of = [&](int f) {
    if (use_tab) {
      const unsigned e = ctab[f];
      return bbeg[e >> OB] + int(e & ((1u << OB) - 1u));
    }
    int l = 0, h = ncand - 1;
    while (l < h) {
      const int mid = (l + h + 1) >> 1;
      if (bofs[mid] <= f) l = mid; else h = mid - 1;
    }
    return bbeg[l] + (f - bofs[l]);
  };
  auto atom_of = [&](int f) { return border[slot_of(f)]; };
  // the tile's scale from the largest |q| it visits and its fullest brick
  double m = 0.0;
  for (int f = threadIdx.x; f < total; f += blockDim.x) m = fmax(m, fabs(q[atom_of(f)]));
  m = wave_max(m);
  if (lane == 0) wmax[wid] = m;
  __syncthreads();
  m = wmax[0];
  int maxcnt = wcnt[0];
  for (int w = 1; w < NT / WAVE; ++w) {
    m = fmax(m, wmax[w]);
    maxcnt = max(maxcnt, wcnt[w]);
  }
  constexpr double WMAX = ORD == 4 ? 0.3 : 1.0;  // (2/3)^3 = 0.296, rounded up
  const double bound = 8.0 * double(maxcnt) * WMAX * m;
  int e = 0;
  (void)frexp(bound, &e);                 // bound < 2^e
  const int S = bound > 0.0 ? 30 - e : 0;  // bound 2^S < 2^30
  const double scale = ldexp(1.0, S);
  for (int f = threadIdx.x; f < total; f += blockDim.x) {
    const int sl = slot_of(f);
    const float qs = float(q[border[sl]] * scale);
    const float4 rc = rec[sl];
    const float zr[3] = {rc.x, rc.y, rc.z};
    const unsigned pk = __float_as_uint(rc.w);
    float AX[3][ORD];
    int o[3][ORD];
    bool in[3][ORD];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float z0 = zr[d];
      if constexpr (ORD == 4) {
        // order-4 B-spline weights of the fractional offset (bspline4_w in FP32)
        const float z1 = z0 + 1.f, z2 = z0 + 2.f, z3 = z0 + 3.f;
        constexpr float s6 = 1.f / 6.f;
        AX[d][0] = z0 * z0 * z0 * s6;
        AX[d][1] = (((-3.f * z1 + 12.f) * z1 - 12.f) * z1 + 4.f) * s6;
        AX[d][2] = (((3.f * z2 - 24.f) * z2 + 60.f) * z2 - 44.f) * s6;
        AX[d][3] = (((-z3 + 12.f) * z3 - 48.f) * z3 + 64.f) * s6;
      } else {
        AX[d][0] = z0;        // 1 - f at the base point (cic_weights)
        AX[d][1] = 1.f - z0;  // f at the next one
      }
      const int K = g.K[d];
      const int bm = int((pk >> (10 * d)) & 1023u);
#pragma unroll
      for (int a = 0; a < ORD; ++a) {
        int idx = bm + a - (ORD == 4 ? 1 : 0);
        idx += idx < 0 ? K : (idx >= K ? -K : 0);
        int p = idx - lo[d];
        if (p < 0) p += K;
        o[d][a] = p;
        in[d][a] = p < ext[d];
      }
    }
#pragma unroll
    for (int c = 0; c < ORD; ++c) {
      if (!in[2][c]) continue;
#pragma unroll
      for (int bb = 0; bb < ORD; ++bb) {
        if (!in[1][bb]) continue;
        const int row = TXg * (o[1][bb] + TYg * o[2][c]);
        const float w12 = qs * AX[1][bb] * AX[2][c];
#pragma unroll
        for (int a = 0; a < ORD; ++a) {
          if (!in[0][a]) continue;
          atomicAdd(&acc[row + o[0][a]], __float2int_rn(w12 * AX[0][a]));
        }
      }
    }
  }
  __syncthreads();
  const float inv = float(ldexp(1.0, -S));
  for (int k = threadIdx.x; k < TILEg; k += blockDim.x) {
    const int px = k % TXg, py = (k / TXg) % TYg, pz = k / (TXg * TYg);
    if (px < ext[0] && py < ext[1] && pz < ext[2]) {
      const size_t gi = size_t(lo[0] + px) +
                        size_t(g.K[0]) * (size_t(lo[1] + py) + size_t(g.K[1]) * size_t(lo[2] + pz));
      grid[gi] = float(acc[k]) * inv;
    }
  }
}

// Owner-computes 32-bit spread of the MIXED coarse grids (tiles that divide
// the grid).  A workgroup takes only the atoms whose base point lies in its
// tile -- its 4^3 bricks, in rows of TXg/4 bricks along x that are contiguous
// in brick order -- and adds all ORD^3 stencil points of each into an LDS
// region extended by the stencil's reach (PL points below the tile, PH
// above), so each atom is processed once (the tiled spread above processes
// every atom whose stencil meets a tile, 2.3 times each for 32 x 32 x 16
// tiles) and no point needs a range test or a periodic wrap.  Sums are 32-bit
// fixed point at the tile's own scale, from the bound of k_spread_tile_i32
// over the owned bricks.  Points that no other tile reaches (the interior)
// are final and go to the grid; the others (within PH of the tile's lower
// faces or PL of its upper ones, and the shell outside the tile) go to the
// tile's slot of ext, and k_spread_fold sums them over the tiles that reach
// each point in a fixed order, so the grid is deterministic.
template <int ORD, int TXg, int TYg, int TZg>
struct SpreadOwn {
  static constexpr int PL = ORD == 4 ? 1 : 0, PH = ORD == 4 ? 2 : 1;
  static constexpr int EX = TXg + PL + PH, EY = TYg + PL + PH, EZ = TZg + PL + PH;
  static constexpr int EXT = EX * EY * EZ;
  // the x columns of the extended region that the fold of the tile's x faces
  // reads (the NR lowest and the NR from TXg on, NR = PL + PH), also kept
  // transposed per (y, z) row: XCOL ints per tile
  static constexpr int NR = PL + PH, NXC = 2 * NR, XCOL = EY * EZ * NXC;
  __host__ __device__ static int xcol_of(int ex) { return ex < NR ? ex : NR + (ex - TXg); }
  static constexpr int NBX = TXg / BRK, NBY = TYg / BRK, ROWS = NBY * (TZg / BRK);
  static constexpr int QCAP = 8192;  // owned atoms whose FP32 charges stay in LDS
  static_assert(TXg % BRK == 0 && TYg % BRK == 0 && TZg % BRK == 0, "whole bricks");
  static bool fits(const Geom& g) {
    return g.K[0] % TXg == 0 && g.K[1] % TYg == 0 && g.K[2] % TZg == 0;
  }
  static int tiles(const Geom& g) { return (g.K[0] / TXg) * (g.K[1] / TYg) * (g.K[2] / TZg); }
  // point r (per axis, relative to the tile's first) receives no other tile's atoms
  __host__ __device__ static bool inner(int r, int T) { return r >= PH && r < T - PL; }
};

template <int ORD, int TXg, int TYg, int TZg, int NT>
__global__ void __launch_bounds__(NT) k_spread_own(Geom g, const double* __restrict__ q,
                                                   const int* __restrict__ bstart,
                                                   const int* __restrict__ border,
                                                   const float4* __restrict__ rec,
                                                   float* __restrict__ grid, int* __restrict__ ext,
                                                   int* __restrict__ xcol, float* __restrict__ tinv) {
  using P = SpreadOwn<ORD, TXg, TYg, TZg>;
  constexpr int EX = P::EX, EY = P::EY, EXT = P::EXT, ROWS = P::ROWS, QCAP = P::QCAP;
  static_assert(ROWS <= NT && ROWS <= WAVE, "one lane per row of bricks");
  __shared__ int acc[EXT];
  __shared__ float qf[QCAP];
  __shared__ int rofs[ROWS + 1];
  __shared__ int rbeg[ROWS];
  __shared__ double wmax[NT / WAVE];
  __shared__ int rcnt;
  const int ntx = g.K[0] / TXg, nty = g.K[1] / TYg;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  const int lo[3] = {tx * TXg, ty * TYg, tz * TZg};
  for (int k = threadIdx.x; k < EXT; k += NT) acc[k] = 0;
  // the owned rows of bricks (the first wave): offsets and the fullest brick
  if (threadIdx.x < WAVE) {
    const int r = threadIdx.x;
    int cnt = 0, cmax = 0;
    if (r < ROWS) {
      const int nb0 = g.K[0] / BRK, nb1 = g.K[1] / BRK;
      const int by = lo[1] / BRK + r % P::NBY, bz = lo[2] / BRK + r / P::NBY;
      const int id0 = lo[0] / BRK + nb0 * (by + nb1 * bz);
      const int b0 = bstart[id0];
      int prev = b0;
      for (int j = 1; j <= P::NBX; ++j) {
        const int bj = bstart[id0 + j];
        cmax = max(cmax, bj - prev);
        prev = bj;
      }
      rbeg[r] = b0;
      cnt = prev - b0;
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const int v = __shfl_up(incl, o, WAVE);
      if (r >= o) incl += v;
      cmax = max(cmax, __shfl_xor(cmax, o, WAVE));
    }
    if (r < ROWS) rofs[r + 1] = incl;
    if (r == 0) { rofs[0] = 0; rcnt = cmax; }
  }
  __syncthreads();
  const int total = rofs[ROWS];
  const bool stage = total <= QCAP;
  auto slot_of = [&](int f) {
    int l = 0;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1)
      if (s < ROWS && l + s < ROWS && rofs[l + s] <= f) l += s;
    return rbeg[l] + (f - rofs[l]);
  };
  // the tile's scale from the largest |q| it owns and its fullest brick
  double m = 0.0;
  for (int f0 = threadIdx.x; f0 < total; f0 += 4 * NT) {
    // four atoms' gathers in flight
    int at[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = f0 + u * NT;
      at[u] = f < total ? border[slot_of(f)] : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = f0 + u * NT;
      if (at[u] >= 0) {
        const double v = q[at[u]];
        m = fmax(m, fabs(v));
        if (stage) qf[f] = float(v);
      }
    }
  }
  m = wave_max(m);
  if ((threadIdx.x & (WAVE - 1)) == 0) wmax[threadIdx.x / WAVE] = m;
  __syncthreads();
  m = wmax[0];
  for (int w = 1; w < NT / WAVE; ++w) m = fmax(m, wmax[w]);
  constexpr double WMAX = ORD == 4 ? 0.3 : 1.0;  // (2/3)^3 = 0.296, rounded up
  const double bound = 8.0 * double(rcnt) * WMAX * m;
  int e = 0;
  (void)frexp(bound, &e);                  // bound < 2^e
  const int S = bound > 0.0 ? 30 - e : 0;  // bound 2^S < 2^30
  const double scale = ldexp(1.0, S);
  const float scalef = float(scale);
  // lanes of a wave take atoms 16 apart (NT / WAVE), in different bricks, so
  // that their atomics spread over the LDS banks
  const int tp = int((threadIdx.x % WAVE) * (NT / WAVE) + threadIdx.x / WAVE);
  for (int f = tp; f < total; f += NT) {
    const int sl = slot_of(f);
    const float qs = stage ? qf[f] * scalef : float(q[border[sl]] * scale);
    float AX[3][ORD];
    int bm[3];
    rec_weights<ORD>(rec[sl], AX, bm);
    const int e0 = bm[0] - lo[0], e1 = bm[1] - lo[1], e2 = bm[2] - lo[2];
    // the atom's base lies in the tile (its brick's); a stale record is skipped
    if (unsigned(e0) >= unsigned(TXg) || unsigned(e1) >= unsigned(TYg) ||
        unsigned(e2) >= unsigned(TZg))
      continue;
    int* base = acc + e0 + EX * (e1 + EY * e2);
#pragma unroll
    for (int c = 0; c < ORD; ++c)
#pragma unroll
      for (int bb = 0; bb < ORD; ++bb) {
        const float w12 = qs * AX[1][bb] * AX[2][c];
        int* row = base + EX * (bb + EY * c);
#pragma unroll
        for (int a = 0; a < ORD; ++a) atomicAdd(row + a, __float2int_rn(w12 * AX[0][a]));
      }
  }
  __syncthreads();
  const float inv = float(ldexp(1.0, -S));
  if (threadIdx.x == 0) tinv[t] = inv;
  int* mine = ext + size_t(t) * EXT;
  int* mcol = xcol + size_t(t) * P::XCOL;
  for (int k = threadIdx.x; k < EXT; k += NT) {
    const int ex = k % EX, eyz = k / EX;
    const int rx = ex - P::PL, ry = eyz % EY - P::PL, rz = eyz / EY - P::PL;
    if (P::inner(rx, TXg) && P::inner(ry, TYg) && P::inner(rz, TZg)) {
      grid[size_t(lo[0] + rx) +
           size_t(g.K[0]) * (size_t(lo[1] + ry) + size_t(g.K[1]) * size_t(lo[2] + rz))] =
          float(acc[k]) * inv;
    } else {
      mine[k] = acc[k];
      if (ex < P::NR || ex >= TXg) mcol[eyz * P::NXC + P::xcol_of(ex)] = acc[k];
    }
  }
}

// The grid points k_spread_own leaves to the fold: each sums the values of
// the tiles whose extended regions hold it (its own tile, and per axis the
// tile below when it lies within PH of its tile's lower face, the tile above
// within PL of the upper one; periodic), each at its tile's scale, in a fixed
// order (z, y, x options nested, own tile first).  Workgroups take the
// boundary points of one tile, enumerated without gaps: the planes with a
// boundary z, then the rows with a boundary y in the inner planes, then the
// boundary x of the inner rows.
template <int ORD, int TXg, int TYg, int TZg>
struct SpreadFold {
  using P = SpreadOwn<ORD, TXg, TYg, TZg>;
  static constexpr int NR = P::PL + P::PH;  // boundary positions per axis
  static constexpr int NIY = TYg - NR, NIZ = TZg - NR;
  static constexpr int NA = NR * TYg * TXg, NB = NIZ * NR * TXg, NC = NIZ * NIY * NR;
  static constexpr int NPT = NA + NB + NC;  // boundary points of a tile
  static constexpr int BPT = (NPT + BLOCK - 1) / BLOCK;  // workgroups per tile
  static_assert(TXg > NR && TYg > NR && TZg > NR, "one neighbour per axis at most");
  __device__ static int rmap(int k, int T) { return k < P::PH ? k : T - NR + k; }
};

template <int ORD, int TXg, int TYg, int TZg>
__global__ void __launch_bounds__(BLOCK) k_spread_fold(Geom g, const int* __restrict__ ext,
                                                       const int* __restrict__ xcol,
                                                       const float* __restrict__ tinv,
                                                       float* __restrict__ grid) {
  using P = SpreadOwn<ORD, TXg, TYg, TZg>;
  using F = SpreadFold<ORD, TXg, TYg, TZg>;
  const int t = blockIdx.x / F::BPT;
  const int j = (blockIdx.x % F::BPT) * BLOCK + threadIdx.x;
  if (j >= F::NPT) return;
  int r[3];
  if (j < F::NA) {
    r[0] = j % TXg;
    r[1] = (j / TXg) % TYg;
    r[2] = F::rmap(j / (TXg * TYg), TZg);
  } else if (j < F::NA + F::NB) {
    const int k = j - F::NA;
    r[0] = k % TXg;
    r[1] = F::rmap((k / TXg) % F::NR, TYg);
    r[2] = P::PH + k / (TXg * F::NR);
  } else {
    const int k = j - F::NA - F::NB;
    r[0] = F::rmap(k % F::NR, TXg);
    r[1] = P::PH + (k / F::NR) % F::NIY;
    r[2] = P::PH + k / (F::NR * F::NIY);
  }
  const int T[3] = {TXg, TYg, TZg};
  const int nt[3] = {g.K[0] / TXg, g.K[1] / TYg, g.K[2] / TZg};
  const int tc[3] = {t % nt[0], (t / nt[0]) % nt[1], t / (nt[0] * nt[1])};
  const size_t gidx = size_t(tc[0] * TXg + r[0]) +
                      size_t(g.K[0]) * (size_t(tc[1] * TYg + r[1]) +
                                        size_t(g.K[1]) * size_t(tc[2] * TZg + r[2]));
  if (j >= F::NA + F::NB) {
    // an x face only: the own tile and the x neighbour, both from the
    // transposed x columns (the sum as below: own, then neighbour)
    const bool low = r[0] < P::PH;
    const int tn = low ? (tc[0] == 0 ? nt[0] - 1 : tc[0] - 1) : (tc[0] + 1 == nt[0] ? 0 : tc[0] + 1);
    const int en = low ? r[0] + TXg + P::PL : r[0] - TXg + P::PL;
    const int row = (r[2] + P::PL) * P::EY + (r[1] + P::PL);
    const int ttn = tn + nt[0] * (tc[1] + nt[1] * tc[2]);
    const float c0 = float(xcol[size_t(t) * P::XCOL + row * P::NXC + P::xcol_of(r[0] + P::PL)]) *
                     tinv[t];
    const float c1 = float(xcol[size_t(ttn) * P::XCOL + row * P::NXC + P::xcol_of(en)]) *
                     tinv[ttn];
    float v = 0.f;
    v += c0;
    v += c1;
    grid[gidx] = v;
    return;
  }
  // per axis the own tile and at most one neighbour (the tiles hold more
  // than PL + PH points per axis, so no point is within reach of both)
  int ot[3][2], oe[3][2];
  bool two[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ot[d][0] = ot[d][1] = tc[d];
    oe[d][0] = oe[d][1] = r[d] + P::PL;
    two[d] = false;
    if (r[d] < P::PH) {
      ot[d][1] = tc[d] == 0 ? nt[d] - 1 : tc[d] - 1;
      oe[d][1] = r[d] + T[d] + P::PL;
      two[d] = true;
    } else if (r[d] >= T[d] - P::PL) {
      ot[d][1] = tc[d] + 1 == nt[d] ? 0 : tc[d] + 1;
      oe[d][1] = r[d] - T[d] + P::PL;
      two[d] = true;
    }
  }
  // all loads first (one memory latency, not one per tile), then the sum in
  // the fixed order
  float c[8];
#pragma unroll
  for (int iz = 0; iz < 2; ++iz)
#pragma unroll
    for (int iy = 0; iy < 2; ++iy)
#pragma unroll
      for (int ix = 0; ix < 2; ++ix) {
        const bool ok = (ix == 0 || two[0]) && (iy == 0 || two[1]) && (iz == 0 || two[2]);
        const int tt = ot[0][ix] + nt[0] * (ot[1][iy] + nt[1] * ot[2][iz]);
        const int ee = oe[0][ix] + P::EX * (oe[1][iy] + P::EY * oe[2][iz]);
        c[ix + 2 * (iy + 2 * iz)] = ok ? float(ext[size_t(tt) * P::EXT + ee]) * tinv[tt] : 0.f;
      }
  float v = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) v += c[k];
  grid[gidx] = v;
}

// K9 (tiled): out_i = scale * sum w Qfin over the order-4 stencil of atom i
// (the gather of PHI_MESH, ctip.f90:381-393).  One workgroup per tile of
// GX x GY x GZ grid points: the tile and its stencil halo (one point below,
// two above) are staged in LDS with row-contiguous loads, then the atoms whose
// base point lies in the tile's bricks (the spread's brick lists) read their
// 64 weights' points from LDS.  Same products and summation order as
// gather_phi, so the result is that of the per-atom gather.
template <class Rg, int GX, int GY, int GZ>
__global__ void __launch_bounds__(BLOCK) k_gather_tile(Geom g, const double* __restrict__ x,
                                                       const int* __restrict__ bstart,
                                                       const int* __restrict__ border,
                                                       const Rg* __restrict__ grid, double scale,
                                                       double* __restrict__ out) {
  constexpr int GLX = GX + 3, GLY = GY + 3, GLZ = GZ + 3;
  constexpr int GBRICKS = (GX / BRK) * (GY / BRK) * (GZ / BRK);
  static_assert(GBRICKS <= BLOCK, "one thread per brick of the tile");
  extern __shared__ unsigned char gather_lds[];
  Rg* tile = reinterpret_cast<Rg*>(gather_lds);
  __shared__ int bofs[GBRICKS + 1];
  __shared__ int bbeg[GBRICKS];
  __shared__ int wsum[BLOCK / WAVE];
  const int ntx = (g.K[0] + GX - 1) / GX, nty = (g.K[1] + GY - 1) / GY;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  const int lo[3] = {tx * GX, ty * GY, tz * GZ};
  const int ext[3] = {min(GX, g.K[0] - lo[0]), min(GY, g.K[1] - lo[1]), min(GZ, g.K[2] - lo[2])};
  // grid points lo-1 .. lo+ext+1 (periodic) into LDS, eight loads in flight
  // per thread (the staging is latency-bound otherwise); LDS index k =
  // px + GLX (py + GLY pz) with compile-time strides
  const int lx = ext[0] + 3, ly = ext[1] + 3, lz = ext[2] + 3;
  constexpr int GL = GLX * GLY * GLZ;
  constexpr int U = 8;
  for (int k0 = threadIdx.x; k0 < GL; k0 += U * BLOCK) {
    Rg v[U];
#pragma unroll
    for (int w = 0; w < U; ++w) {
      const int k = k0 + w * BLOCK;
      v[w] = Rg(0);
      const int px = k % GLX, py = (k / GLX) % GLY, pz = k / (GLX * GLY);
      if (k < GL && px < lx && py < ly && pz < lz) {
        int gx = lo[0] - 1 + px, gy = lo[1] - 1 + py, gz = lo[2] - 1 + pz;
        gx += gx < 0 ? g.K[0] : (gx >= g.K[0] ? -g.K[0] : 0);
        gy += gy < 0 ? g.K[1] : (gy >= g.K[1] ? -g.K[1] : 0);
        gz += gz < 0 ? g.K[2] : (gz >= g.K[2] ? -g.K[2] : 0);
        v[w] = grid[size_t(gx) + size_t(g.K[0]) * (size_t(gy) + size_t(g.K[1]) * size_t(gz))];
      }
    }
#pragma unroll
    for (int w = 0; w < U; ++w) {
      const int k = k0 + w * BLOCK;
      if (k < GL) tile[k] = v[w];
    }
  }
  // atom counts of the tile's bricks and their prefix sum
  const int nbx = g.K[0] / BRK, nby = g.K[1] / BRK;
  const int cbx = ext[0] / BRK, cby = ext[1] / BRK, cbz = ext[2] / BRK;
  const int ncand = cbx * cby * cbz;
  int cntk = 0;
  if (threadIdx.x < ncand) {
    const int k = threadIdx.x;
    const int bx = lo[0] / BRK + k % cbx, by = lo[1] / BRK + (k / cbx) % cby,
              bz = lo[2] / BRK + k / (cbx * cby);
    const int bid = bx + nbx * (by + nby * bz);
    const int b0 = bstart[bid];
    bbeg[k] = b0;
    cntk = bstart[bid + 1] - b0;
  }
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const int v = __shfl_up(cntk, o, WAVE);
    if (lane >= o) cntk += v;
  }
  if (lane == WAVE - 1) wsum[wid] = cntk;
  if (threadIdx.x == 0) bofs[0] = 0;
  __syncthreads();
  for (int w = 0; w < wid; ++w) cntk += wsum[w];
  if (threadIdx.x < ncand) bofs[threadIdx.x + 1] = cntk;
  __syncthreads();
  const int total = bofs[ncand];
  for (int f = threadIdx.x; f < total; f += blockDim.x) {
    int l = 0, h = ncand - 1;
    while (l < h) {
      const int mid = (l + h + 1) >> 1;
      if (bofs[mid] <= f) l = mid; else h = mid - 1;
    }
    const int ai = border[bbeg[l] + (f - bofs[l])];
    double u[3], AX[3][4], AXX[3][4];
    int b[3];
    grid_coord(g, x + 3 * ai, u, b);
    int o[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      bspline4_w(u[d], AX[d]);
      int r = b[d] % g.K[d];
      if (r < 0) r += g.K[d];
      o[d] = r - lo[d];  // stencil point a sits at LDS index o + a (0 .. ext + 2)
    }
    double e = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int bb = 0; bb < 4; ++bb)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          e += AX[0][a] * AX[1][bb] * AX[2][c] *
               double(tile[(o[0] + a) + GLX * ((o[1] + bb) + GLY * (o[2] + c))]);
    out[ai] = scale * e;
  }
}

// k_gather_tile on an FP32 grid with stencil records (the MIXED inner grid):
// each candidate's FP32 weights come from its brick-order record, i.e. the
// weights k_spread_tile_i32<4> spread with (the gather is the spread's exact
// transpose), instead of FP64 coordinates from x; FP32 products, FP64 sum.
template <int GX, int GY, int GZ>
__global__ void __launch_bounds__(BLOCK) k_gather_tile_rec(Geom g, const int* __restrict__ bstart,
                                                           const int* __restrict__ border,
                                                           const float4* __restrict__ rec,
                                                           const float* __restrict__ grid,
                                                           double scale,
                                                           double* __restrict__ out) {
  constexpr int GLX = GX + 3, GLY = GY + 3, GLZ = GZ + 3;
  constexpr int GBRICKS = (GX / BRK) * (GY / BRK) * (GZ / BRK);
  static_assert(GBRICKS <= BLOCK, "one thread per brick of the tile");
  extern __shared__ unsigned char gather_lds[];
  float* tile = reinterpret_cast<float*>(gather_lds);
  __shared__ int bofs[GBRICKS + 1];
  __shared__ int bbeg[GBRICKS];
  __shared__ int wsum[BLOCK / WAVE];
  const int ntx = (g.K[0] + GX - 1) / GX, nty = (g.K[1] + GY - 1) / GY;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  const int lo[3] = {tx * GX, ty * GY, tz * GZ};
  const int ext[3] = {min(GX, g.K[0] - lo[0]), min(GY, g.K[1] - lo[1]), min(GZ, g.K[2] - lo[2])};
  const int lx = ext[0] + 3, ly = ext[1] + 3, lz = ext[2] + 3;
  constexpr int GL = GLX * GLY * GLZ;
  constexpr int U = 8;
  for (int k0 = threadIdx.x; k0 < GL; k0 += U * BLOCK) {
    float v[U];
#pragma unroll
    for (int w = 0; w < U; ++w) {
      const int k = k0 + w * BLOCK;
      v[w] = 0.f;
      const int px = k % GLX, py = (k / GLX) % GLY, pz = k / (GLX * GLY);
      if (k < GL && px < lx && py < ly && pz < lz) {
        int gx = lo[0] - 1 + px, gy = lo[1] - 1 + py, gz = lo[2] - 1 + pz;
        gx += gx < 0 ? g.K[0] : (gx >= g.K[0] ? -g.K[0] : 0);
        gy += gy < 0 ? g.K[1] : (gy >= g.K[1] ? -g.K[1] : 0);
        gz += gz < 0 ? g.K[2] : (gz >= g.K[2] ? -g.K[2] : 0);
        v[w] = grid[size_t(gx) + size_t(g.K[0]) * (size_t(gy) + size_t(g.K[1]) * size_t(gz))];
      }
    }
#pragma unroll
    for (int w = 0; w < U; ++w) {
      const int k = k0 + w * BLOCK;
      if (k < GL) tile[k] = v[w];
    }
  }
  const int nbx = g.K[0] / BRK, nby = g.K[1] / BRK;
  const int cbx = ext[0] / BRK, cby = ext[1] / BRK, cbz = ext[2] / BRK;
  const int ncand = cbx * cby * cbz;
  int cntk = 0;
  if (threadIdx.x < ncand) {
    const int k = threadIdx.x;
    const int bx = lo[0] / BRK + k % cbx, by = lo[1] / BRK + (k / cbx) % cby,
              bz = lo[2] / BRK + k / (cbx * cby);
    const int bid = bx + nbx * (by + nby * bz);
    const int b0 = bstart[bid];
    bbeg[k] = b0;
    cntk = bstart[bid + 1] - b0;
  }
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const int v = __shfl_up(cntk, o, WAVE);
    if (lane >= o) cntk += v;
  }
  if (lane == WAVE - 1) wsum[wid] = cntk;
  if (threadIdx.x == 0) bofs[0] = 0;
  __syncthreads();
  for (int w = 0; w < wid; ++w) cntk += wsum[w];
  if (threadIdx.x < ncand) bofs[threadIdx.x + 1] = cntk;
  __syncthreads();
  const int total = bofs[ncand];
  for (int f = threadIdx.x; f < total; f += blockDim.x) {
    int l = 0, h = ncand - 1;
    while (l < h) {
      const int mid = (l + h + 1) >> 1;
      if (bofs[mid] <= f) l = mid; else h = mid - 1;
    }
    const int sl = bbeg[l] + (f - bofs[l]);
    float AX[3][4];
    int bm[3];
    rec_weights<4>(rec[sl], AX, bm);
    int o[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) o[d] = bm[d] - lo[d];  // stencil point a at LDS index o + a
    double e = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float ec = 0.f;
#pragma unroll
      for (int bb = 0; bb < 4; ++bb) {
        const float* rowp = tile + GLX * ((o[1] + bb) + GLY * (o[2] + c)) + o[0];
        ec += AX[1][bb] * (AX[0][0] * rowp[0] + AX[0][1] * rowp[1] + AX[0][2] * rowp[2] +
                           AX[0][3] * rowp[3]);
      }
      e += double(AX[2][c] * ec);
    }
    out[border[sl]] = scale * e;
  }
}

template <int GX, int GY, int GZ>
struct GatherTile {
  static constexpr size_t lds(size_t elem) { return size_t(GX + 3) * (GY + 3) * (GZ + 3) * elem; }
  static int tiles(const Geom& g) {
    return ((g.K[0] + GX - 1) / GX) * ((g.K[1] + GY - 1) / GY) * ((g.K[2] + GZ - 1) / GZ);
  }
};

// ql = q + step p (the trial charges of a line search)
__global__ void k_lincomb(int n, const double* __restrict__ q, const double* __restrict__ p,
                          double step, double* __restrict__ ql) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ql[i] = q[i] + step * p[i];
}

// generic exclusive scan of int counts: per-1024 block scan + block sums
__global__ void k_scan_block(int n, const int* __restrict__ in, int* __restrict__ out,
                             int* __restrict__ bsum) {
  __shared__ int s[1024];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const int v = i < n ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) bsum[blockIdx.x] = s[1023];
}
__global__ void k_scan_sums(int nb, int* __restrict__ bsum, int* __restrict__ total) {
  __shared__ int s[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int k = base + threadIdx.x;
    const int v = k < nb ? bsum[k] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (k < nb) bsum[k] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}
__global__ void k_scan_add(int n, int* __restrict__ out, const int* __restrict__ bsum,
                           const int* __restrict__ total) {
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += bsum[blockIdx.x];
  if (i == 0) out[n] = *total;
}

}  // namespace sx
