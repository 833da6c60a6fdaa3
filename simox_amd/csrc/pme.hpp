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

// K6: spread q (or q + step*p, written to ql) onto the fixed-point grid.
// F32W: FP32-rounded weights as the reference's default-kind CMPLX makes them
// (force.f90:406, ctip.f90:359; quirk Q1).
template <bool F32W>
__global__ void __launch_bounds__(BLOCK) k_spread(Geom g, const double* __restrict__ x,
                                                  const double* __restrict__ q,
                                                  const double* __restrict__ p, double step,
                                                  double* __restrict__ ql,
                                                  unsigned long long* __restrict__ grid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  double qi = q[i];
  if (p) {
    qi = qi + step * p[i];
    ql[i] = qi;
  }
  double u[3], AX[3][4], AXX[3][4];
  int b[3];
  grid_coord(g, x + 3 * i, u, b);
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline4(u[d], AX[d], AXX[d]);
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
        double w = AX[0][a] * AX[1][bb] * AX[2][c];
        if (F32W) w = double(float(w));
        const long long v = llrint(qi * w * GRID_SCALE);
        atomicAdd(&grid[row + ix[a]], (unsigned long long)v);
      }
    }
}

// K7a: R2C along x.  Reads the fixed-point grid (and zeroes it for the next
// spread), transforms M lines per block in LDS, writes kx = 0..Kx/2.
template <class C>
__global__ void __launch_bounds__(BLOCK) k_fft_x_r2c(Geom g, FftPlan pl, const C* __restrict__ tw,
                                                     long long* __restrict__ gfix,
                                                     C* __restrict__ spec, int M, int stride) {
  using R = decltype(C{}.x);
  extern __shared__ __align__(16) unsigned char smem[];
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + M * stride;
  const int L = g.K[0], H = g.H, nlines = g.K[1] * g.K[2];
  const int line0 = blockIdx.x * M;
  for (int e = threadIdx.x; e < M * L; e += blockDim.x) {
    const int t = e / L, xk = e - t * L;
    const int line = line0 + t;
    C v; v.x = R(0); v.y = R(0);
    if (line < nlines) {
      const size_t gi = size_t(line) * L + xk;
      const long long iv = gfix[gi];
      gfix[gi] = 0;
      v.x = R(double(iv) * GRID_INV_SCALE);
    }
    b0[t * stride + xk] = v;
  }
  __syncthreads();
  const C* res = lds_fft(b0, b1, M, L, stride, pl, tw, -1);
  for (int e = threadIdx.x; e < M * H; e += blockDim.x) {
    const int t = e / H, k = e - t * H;
    const int line = line0 + t;
    if (line < nlines) spec[size_t(line) * H + k] = res[t * stride + k];
  }
}

// K7b/K8: FFT along a strided axis of the half spectrum.  Block (bx, by)
// transforms the lines at x = bx*M .. bx*M+M-1 (contiguous, coalesced) and
// outer index by.  FUSE: forward FFT, multiply by the influence table G
// (same layout as spec, 1/Ng folded in), accumulate sum w_kx G |X|^2 into
// epart, inverse FFT.
template <class C, bool FUSE>
__global__ void __launch_bounds__(BLOCK) k_fft_strided(Geom g, FftPlan pl,
                                                       const C* __restrict__ tw,
                                                       C* __restrict__ spec, int es, int os,
                                                       int M, int stride, int sign,
                                                       const decltype(C{}.x)* __restrict__ G,
                                                       double* __restrict__ epart) {
  using R = decltype(C{}.x);
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double lds_red[BLOCK / WAVE];
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + M * stride;
  const int L = pl.L, H = g.H;
  const int x0 = blockIdx.x * M;
  const size_t obase = size_t(blockIdx.y) * os;
  for (int e = threadIdx.x; e < M * L; e += blockDim.x) {
    const int l = e / M, t = e - l * M;
    const int xx = x0 + t;
    C v; v.x = R(0); v.y = R(0);
    if (xx < H) v = spec[obase + size_t(l) * es + xx];
    b0[t * stride + l] = v;
  }
  __syncthreads();
  C* res = lds_fft(b0, b1, M, L, stride, pl, tw, FUSE ? -1 : sign);
  if (FUSE) {
    double en[1] = {0.0};
    for (int e = threadIdx.x; e < M * L; e += blockDim.x) {
      const int l = e / M, t = e - l * M;
      const int xx = x0 + t;
      C v = res[t * stride + l];
      R gv = R(0);
      if (xx < H) gv = G[obase + size_t(l) * es + xx];
      const double wx = (xx == 0 || 2 * xx == g.K[0]) ? 1.0 : 2.0;
      if (xx < H) en[0] += wx * double(gv) * (double(v.x) * double(v.x) + double(v.y) * double(v.y));
      v.x *= gv; v.y *= gv;
      res[t * stride + l] = v;
    }
    __syncthreads();
    C* other = (res == b0) ? b1 : b0;
    res = lds_fft(res, other, M, L, stride, pl, tw, +1);
    block_sum<1>(en, lds_red);
    if (threadIdx.x == 0) epart[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = en[0];
  }
  for (int e = threadIdx.x; e < M * L; e += blockDim.x) {
    const int l = e / M, t = e - l * M;
    const int xx = x0 + t;
    if (xx < H) spec[obase + size_t(l) * es + xx] = res[t * stride + l];
  }
}

// K7c: C2R along x (inverse, unnormalised); rebuilds the Hermitian line.
template <class C>
__global__ void __launch_bounds__(BLOCK) k_fft_x_c2r(Geom g, FftPlan pl, const C* __restrict__ tw,
                                                     const C* __restrict__ spec,
                                                     decltype(C{}.x)* __restrict__ out, int M,
                                                     int stride) {
  using R = decltype(C{}.x);
  extern __shared__ __align__(16) unsigned char smem[];
  C* b0 = reinterpret_cast<C*>(smem);
  C* b1 = b0 + M * stride;
  const int L = g.K[0], H = g.H, nlines = g.K[1] * g.K[2];
  const int line0 = blockIdx.x * M;
  for (int e = threadIdx.x; e < M * L; e += blockDim.x) {
    const int t = e / L, k = e - t * L;
    const int line = line0 + t;
    C v; v.x = R(0); v.y = R(0);
    if (line < nlines) {
      if (k < H) {
        v = spec[size_t(line) * H + k];
      } else {
        C w = spec[size_t(line) * H + (L - k)];
        v.x = w.x; v.y = -w.y;
      }
    }
    b0[t * stride + k] = v;
  }
  __syncthreads();
  const C* res = lds_fft(b0, b1, M, L, stride, pl, tw, +1);
  for (int e = threadIdx.x; e < M * L; e += blockDim.x) {
    const int t = e / L, k = e - t * L;
    const int line = line0 + t;
    if (line < nlines) out[size_t(line) * L + k] = res[t * stride + k].x;
  }
}

// K14: influence table on the half spectrum, G(k)/Ng with
// G = psi(k) 36^3 / prod_d (4 + 2 cos(2 pi k_d / K_d))^2 (force.f90:414-420,
// 500-514); G(0) = 0 (quirk Q5).
template <class R>
__global__ void k_influence(Geom g, double a2, R* __restrict__ G) {
  const size_t total = size_t(g.H) * g.K[1] * g.K[2];
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int kx = int(t % g.H);
  const int ky = int((t / g.H) % g.K[1]);
  const int kz = int(t / (size_t(g.H) * g.K[1]));
  const int k3[3] = {kx, ky, kz};
  int m[3];
  double zf = 1.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int K = g.K[d];
    m[d] = (k3[d] + K / 2) % K - K / 2;
    double z = 4.0 + 2.0 * cos(2.0 * PI * double(k3[d]) / double(K));
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
    for (int c = 0; c < 4; ++c)
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
  for (int a = 0; a < 4; ++a)
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const size_t gi = size_t(ix[a]) + size_t(g.K[0]) * (size_t(iy[bb]) + size_t(g.K[1]) * size_t(iz[c]));
        e += AX[0][a] * AX[1][bb] * AX[2][c] * double(grid[gi]);
      }
  return e;
}

}  // namespace sx
