// Pair kernels on the reference's half shell.
//
// K1b  k_nlist_tile: the neighbour list of every atom of one cell per
//      workgroup, from the 27 surrounding cells staged in LDS.  Positions are
//      staged in FP32 relative to the home cell's centre; an FP32 distance
//      test decides every candidate outside a narrow band around Rcut^2 whose
//      width bounds the FP32 error, and the band is decided by the exact FP64
//      test of force.f90:51-54.  The list of atom i holds first the entries in
//      which i is the reference's "i" (its own cell with j after i, then the
//      13 cells of the half-shell stencil, cellsort.f90:18-30, as FORCE visits
//      them, force.f90:43-48,146-155), then those in which it is "j"; cnt_fwd
//      counts the first part.  Every pair of FORCE therefore appears once
//      forward with the reference's orientation, which decides the densities
//      of AL_NI (eam.f90:101-168, force.f90:60-63).
// K2   k_density: rho_i and F(rho_i) over the whole list.
// K4   k_force_half: each pair once (forward entries), both atoms' forces in
//      64-bit fixed point: integer sums do not depend on their order, so the
//      forces are deterministic although atom j's share is added atomically.
#pragma once
#include "pair.hpp"

namespace sx {

// (dx, dy, dz) of cell(id)%pair(s + 1), s = 0..12 (cellsort.f90:18-30):
// the 9 cells of layer k-1 with i and j running -1..1 (i fastest), then
// (i-1, j-1, k), (i, j-1, k), (i+1, j-1, k), (i-1, j, k).
__host__ __device__ inline void stencil_off(int s, int& dx, int& dy, int& dz) {
  const int t = s < 9 ? s : s - 9;
  dz = s < 9 ? -1 : 0;
  dy = t / 3 - 1;
  dx = t % 3 - 1;
}

// Visits of one home cell, in list order: 0 the home cell (j after i),
// 1..13 the stencil cells, 14 the home cell (j before i), 15..27 the cells
// opposite the stencil.  Visits 0..13 are forward (i is FORCE's "i").
constexpr int NVISIT = 28;
__host__ __device__ inline void visit_off(int v, int& dx, int& dy, int& dz) {
  if (v == 0 || v == 14) { dx = dy = dz = 0; return; }
  if (v < 14) { stencil_off(v - 1, dx, dy, dz); return; }
  stencil_off(v - 15, dx, dy, dz);
  dx = -dx; dy = -dy; dz = -dz;
}

// Fixed-point force accumulation: 2^40 per eV/A (quantum 9.1e-13).  A pair
// contributes less than 2^14 eV/A per component (checked), so the sum of
// NNB_MAX = 500 < 2^9 of them stays below 2^23 * 2^40 = 2^63.
constexpr double FIX_SCALE = 1099511627776.0;
constexpr double FIX_INV = 1.0 / 1099511627776.0;
constexpr double FIX_MAX = 16384.0;
// near-pair energies (MIXED) in 64-bit fixed point, 2^32 per eV
constexpr double EFIX_SCALE = 4294967296.0;
constexpr double EFIX_INV = 1.0 / 4294967296.0;
constexpr double EFIX_MAX = 1048576.0;

// Band around Rcut^2 where the FP32 test defers to FP64.  Staged coordinates
// are within 3 Rcut of the home centre (cells are at most 2 Rcut wide): each
// rounded coordinate is off by < 3 Rcut 2^-24, their difference by
// < 12 Rcut 2^-24, and by < 24 Rcut 2^-24 after the FP32 periodic fold of
// boxes with fewer than four cells.  Near Rcut, r^2 is then off by
// < 2 sqrt(3) Rcut 24 Rcut 2^-24 + 3 Rcut^2 2^-24 < 5.1e-6 Rcut^2; the band
// is 2^-13 Rcut^2 = 1.2e-4 Rcut^2.
constexpr float TILE_BAND = 1.0f / 8192.0f;

// Cell-tile shapes (staged atoms per LDS chunk, threads per workgroup), chosen
// per context by timing at the first list build (all give the same list: the
// chunks split the visits, whose order is fixed).  The first is the default.
constexpr int NTILE_SHAPES = 6;
constexpr int TILE_SHAPES[NTILE_SHAPES][2] = {{1024, 128}, {512, 128}, {2048, 128},
                                              {1024, 64},  {1024, 256}, {512, 256}};
constexpr int TILE_CAP = 1024;  // the default shape: 16 KB of LDS, 4+ waves per SIMD
constexpr int TILE_NT = 128;    // two waves

template <int CAP>
struct TileLds {
  float4 p[CAP];                  // x, y, z relative to the home centre (w unused)
  int vcell[NVISIT];        // cell of each visit
  int vl[NVISIT + 1];       // LDS start of each visit of the chunk
  int kcnt[NPT_MAX];        // entries found so far, per home atom
  int kfwd[NPT_MAX];        // forward entries, per home atom
};

struct TileArgs {
  Geom g;
  const double* x;
  const int* start;   // cell segments of the sorted atoms (start[ncall] = n)
  int cap;
  int* cnt;
  int* cnt_fwd;
  int* col;           // ELL, entry k of atom i at k*n + i
  int* err;
};

// Cell of visit v from home cell (cx, cy, cz).
__host__ __device__ inline int visit_cell(const Geom& g, int cx, int cy, int cz, int v) {
  int dx, dy, dz;
  visit_off(v, dx, dy, dz);
  const int xx = cell_wrap(cx + dx, g.nc[0]);
  const int yy = cell_wrap(cy + dy, g.nc[1]);
  const int zz = cell_wrap(cz + dz, g.nc[2]);
  return xx + g.nc[0] * (yy + g.nc[1] * zz);
}

// FP32 coordinate of a position relative to the home centre h, taken to the
// periodic image nearest h (in FP64, then rounded).
__host__ __device__ inline float rel_coord(double x, double h, double box) {
  const double d = x - h;
  return float(d - box * rint(d / box));
}

// First visit of the chunk after [v0, ...) and the LDS layout of the chunk:
// visits are taken in order while their atoms fit in CAP.  A visit never
// exceeds NPT_MAX <= CAP atoms, so every chunk holds at least one.
template <int CAP>
__host__ __device__ inline int tile_chunk_end(const TileArgs& A, const TileLds<CAP>& L, int v0) {
  static_assert(NPT_MAX <= CAP, "a cell fits in one chunk");
  int tot = 0, v = v0;
  for (; v < NVISIT; ++v) {
    const int c = L.vcell[v];
    const int m = A.start[c + 1] - A.start[c];
    if (v > v0 && tot + m > CAP) break;
    tot += m;
  }
  return v;
}

// Stage the atoms of visits [v0, v1) (thread t of nt); vl is filled by every
// thread identically before the barrier that follows.
template <int CAP>
__host__ __device__ inline void tile_stage(const TileArgs& A, TileLds<CAP>& L, int v0, int v1, int t,
                                           int nt, const double h[3]) {
  int lo = 0;
  for (int v = v0; v < v1; ++v) {
    const int c = L.vcell[v];
    const int s0 = A.start[c], m = A.start[c + 1] - s0;
    for (int a = t; a < m; a += nt) {
      const int j = s0 + a;
      L.p[lo + a] = make_float4(rel_coord(A.x[3 * j], h[0], A.g.box[0]),
                                rel_coord(A.x[3 * j + 1], h[1], A.g.box[1]),
                                rel_coord(A.x[3 * j + 2], h[2], A.g.box[2]), 0.0f);
    }
    lo += m;
  }
}

template <int CAP>
__host__ __device__ inline void tile_layout(const TileArgs& A, TileLds<CAP>& L, int v0, int v1) {
  int lo = 0;
  for (int v = v0; v < v1; ++v) {
    L.vl[v] = lo;
    const int c = L.vcell[v];
    lo += A.start[c + 1] - A.start[c];
  }
  L.vl[v1] = lo;
}

// Candidates of visits [v0, v1) for home atom a (sorted index i = hs + a).
// Four candidates per round: their LDS reads and distances are independent,
// so their latencies overlap.  WRAP: the box has fewer than four cells along
// some axis, where the staged image need not be the pair's nearest.
template <bool WRAP, int CAP>
__host__ __device__ inline void tile_scan(const TileArgs& A, TileLds<CAP>& L, int v0, int v1, int hs,
                                          int a, const double h[3]) {
  const Geom& g = A.g;
  const int i = hs + a;
  const size_t n = size_t(g.n);
  const double xi = A.x[3 * i], yi = A.x[3 * i + 1], zi = A.x[3 * i + 2];
  const float pxi = rel_coord(xi, h[0], g.box[0]);
  const float pyi = rel_coord(yi, h[1], g.box[1]);
  const float pzi = rel_coord(zi, h[2], g.box[2]);
  const float bx = float(g.box[0]), by = float(g.box[1]), bz = float(g.box[2]);
  const float ibx = 1.0f / bx, iby = 1.0f / by, ibz = 1.0f / bz;
  // with four or more cells along an axis every staged image is the nearest
  const bool wx = g.nc[0] < 4, wy = g.nc[1] < 4, wz = g.nc[2] < 4;
  const float rc2 = float(g.rcut2);
  const float lo2 = rc2 * (1.0f - TILE_BAND), hi2 = rc2 * (1.0f + TILE_BAND);
  auto dist2 = [&](const float4& c) {
    float dx = pxi - c.x, dy = pyi - c.y, dz = pzi - c.z;
    if constexpr (WRAP) {
      if (wx) dx -= bx * rintf(dx * ibx);
      if (wy) dy -= by * rintf(dy * iby);
      if (wz) dz -= bz * rintf(dz * ibz);
    }
    return dx * dx + dy * dy + dz * dz;
  };
  int k = L.kcnt[a];
  int* out = A.col + size_t(k) * n + i;  // entry k of atom i
  // candidate j with FP32 r^2 = r2f below hi2: FP64 test in the band
  auto take = [&](int j, float r2f) {
    if (!(r2f < lo2)) {
      double rx[3], r2;
      min_image(g, xi, yi, zi, A.x[3 * j], A.x[3 * j + 1], A.x[3 * j + 2], rx, r2);
      if (!(r2 < g.rcut2)) return;
    }
    if (k < A.cap) *out = j;
    out += n;
    ++k;
  };
  for (int v = v0; v < v1; ++v) {
    if (v == 14) L.kfwd[a] = k;
    const int c = L.vcell[v];
    const int s0 = L.vl[v], s1 = L.vl[v + 1];
    const int gofs = A.start[c] - s0;  // sorted index of LDS slot s: s + gofs
    int s = s0, se = s1;
    if (v == 0) s = s0 + a + 1;  // home cell: j after i
    if (v == 14) se = s0 + a;    // home cell: j before i
    for (; s + 4 <= se; s += 4) {
      const float r0 = dist2(L.p[s]), r1 = dist2(L.p[s + 1]);
      const float r2 = dist2(L.p[s + 2]), r3 = dist2(L.p[s + 3]);
      // NaN compares false: the FP64 test would fail too
      if (r0 < hi2) take(s + gofs, r0);
      if (r1 < hi2) take(s + 1 + gofs, r1);
      if (r2 < hi2) take(s + 2 + gofs, r2);
      if (r3 < hi2) take(s + 3 + gofs, r3);
    }
    for (; s < se; ++s) {
      const float r0 = dist2(L.p[s]);
      if (r0 < hi2) take(s + gofs, r0);
    }
  }
  L.kcnt[a] = k;
}

// Home cell of workgroup b, its centre, its first atom and its size.
__host__ __device__ inline void tile_home(const TileArgs& A, int b, int& cx, int& cy, int& cz,
                                          double h[3], int& hs, int& m) {
  const Geom& g = A.g;
  cx = b % g.nc[0];
  cy = (b / g.nc[0]) % g.nc[1];
  cz = b / (g.nc[0] * g.nc[1]);
  h[0] = (cx + 0.5) * g.lcell[0] - g.hbox[0];
  h[1] = (cy + 0.5) * g.lcell[1] - g.hbox[1];
  h[2] = (cz + 0.5) * g.lcell[2] - g.hbox[2];
  hs = A.start[b];
  m = A.start[b + 1] - hs;
}

// Counts of home atom a; true when its list overflowed (reported like
// FORCE's list overflow, force.f90:79).
template <int CAP>
__host__ __device__ inline bool tile_finish(const TileArgs& A, const TileLds<CAP>& L, int hs, int a) {
  const int i = hs + a, k = L.kcnt[a], kf = L.kfwd[a];
  A.cnt[i] = k < A.cap ? k : A.cap;
  A.cnt_fwd[i] = kf < A.cap ? kf : A.cap;
  return k > NNB_MAX || k > A.cap;
}

// Every visited cell holds at most NPT_MAX atoms (a fuller cell is an error
// of the cell sort, which reports it; the workgroup then stages nothing).
template <int CAP>
__host__ __device__ inline bool tile_cells_ok(const TileArgs& A, const TileLds<CAP>& L) {
  for (int v = 0; v < NVISIT; ++v) {
    const int c = L.vcell[v];
    if (A.start[c + 1] - A.start[c] > NPT_MAX) return false;
  }
  return true;
}

// One workgroup of NT threads per cell (grid = ncall; with g.xcd the
// cells go to XCDs in runs of neighbours, which share staged cells in L2).
template <int CAP, int NT>
__global__ void __launch_bounds__(NT) k_nlist_tile(TileArgs A) {
  __shared__ TileLds<CAP> L;
  const int b = A.g.xcd ? xcd_tile(blockIdx.x, gridDim.x) : int(blockIdx.x);
  int cx, cy, cz, hs, m;
  double h[3];
  tile_home(A, b, cx, cy, cz, h, hs, m);
  if (m == 0 || m > NPT_MAX) return;  // uniform over the workgroup
  const int t = threadIdx.x;
  if (t < NVISIT) L.vcell[t] = visit_cell(A.g, cx, cy, cz, t);
  for (int a = t; a < m; a += NT) { L.kcnt[a] = 0; L.kfwd[a] = 0; }
  __syncthreads();
  if (!tile_cells_ok(A, L)) return;  // uniform: every thread reads the same counts
  const bool wrap = A.g.nc[0] < 4 || A.g.nc[1] < 4 || A.g.nc[2] < 4;
  for (int v0 = 0; v0 < NVISIT;) {
    const int v1 = tile_chunk_end(A, L, v0);
    if (v0 > 0) __syncthreads();  // the previous chunk is read by everyone
    if (t == 0) tile_layout(A, L, v0, v1);
    tile_stage(A, L, v0, v1, t, NT, h);
    __syncthreads();
    if (wrap)
      for (int a = t; a < m; a += NT) tile_scan<true>(A, L, v0, v1, hs, a, h);
    else
      for (int a = t; a < m; a += NT) tile_scan<false>(A, L, v0, v1, hs, a, h);
    v0 = v1;
  }
  for (int a = t; a < m; a += NT)
    if (tile_finish(A, L, hs, a)) atomicCAS(A.err, 0, SIMOX_ERR_NB_OVERFLOW);
}

// ---------------------------------------------------------------------------
// K2: rho_i = sum over the whole list of the density atom i receives, by the
// reference's orientation of each pair: forward entries take fj(2) of
// (s_i, s_j), the others fj(1) of (s_j, s_i) (force.f90:84-85).  rho is summed
// in FP64 in list order.  T = float (MIXED): FP32 terms, FP64 for pairs
// closer than R_NEAR; F(rho), F'(rho) in FP64 in every mode (force.f90:256-348).
// egroup_part[blockIdx.x] = sum of F over the block.
// ---------------------------------------------------------------------------
template <class T, int U = 1>
__global__ void __launch_bounds__(BLOCK) k_density(Geom g, const Tables* __restrict__ tb,
                                                   const double* __restrict__ x,
                                                   const int* __restrict__ id,
                                                   const int* __restrict__ cnt,
                                                   const int* __restrict__ cnt_fwd,
                                                   const int* __restrict__ col,
                                                   double* __restrict__ rho_out,
                                                   T* __restrict__ dF_out,
                                                   double* __restrict__ egroup_part,
                                                   int* __restrict__ err) {
  __shared__ double lds[BLOCK / WAVE];
  const int i = pair_block(g) * blockDim.x + threadIdx.x;
  double eg[1] = {0.0};
  if (i < g.n) {
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    const int si = id[i];
    const size_t n = size_t(g.n);
    const int c = cnt[i], cf = cnt_fwd[i];
    double rho = 0.0;
    bool bad = si < 1 || si > NSPEC;
    // U entries' neighbour data loaded ahead, then taken in list order
    for (int k0 = 0; k0 < c && !bad; k0 += U) {
      int sv[U];
      double xv[U][3];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = k0 + u < c ? col[size_t(k0 + u) * n + i] : i;
        sv[u] = id[j];
        xv[u][0] = x[3 * j];
        xv[u][1] = x[3 * j + 1];
        xv[u][2] = x[3 * j + 2];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u;
        if (k >= c || bad) break;
        const int sj = sv[u];
        if (sj < 1 || sj > NSPEC || !tb->pc[pidx(si, sj)].valid ||
            (tb->sp.on && !(spl_species_ok(si) && spl_species_ok(sj)))) {
          bad = true;
          break;
        }
        double rx[3], r2;
        min_image(g, xi, yi, zi, xv[u][0], xv[u][1], xv[u][2], rx, r2);
        if (!(r2 < g.ream2)) continue;  // split cutoff: EAM only inside r_eam
        const bool fwd = k < cf;
        if (tb->sp.on) {
          T f, dfr;
          spl_eval<T>(tb->sp, spl_dens(si, sj), T(sqrt(r2)), f, dfr);
          rho += double(f);
        } else {
          const DensCoef& dc = fwd ? tb->pc[pidx(si, sj)].dj : tb->pc[pidx(sj, si)].di;
          if (near_pair<T>(r2)) {
            double f, dfr;
            dens_eval<double>(dc, sqrt(r2), f, dfr);
            rho += f;
          } else {
            T f, dfr;
            dens_eval<T>(dc, T(sqrt(r2)), f, dfr);
            rho += double(f);
          }
        }
      }
    }
    if (bad) SX_ERR_SET(err, SIMOX_ERR_UNKNOWN_PAIR);
    double F = 0.0, dF = 0.0;
    if (!bad) {
      if (tb->sp.on)
        spl_eval<double>(tb->sp, spl_embed(si), rho, F, dF);
      else
        embed<double>(tb->ec[si - 1], rho, F, dF);
    }
    rho_out[i] = rho;
    dF_out[i] = T(dF);
    eg[0] = F;
  }
  block_sum<1>(eg, lds);
  if (threadIdx.x == 0) egroup_part[blockIdx.x] = eg[0];
}

// Fixed-point image of a force component; out of range sets *bad.
__device__ inline long long to_fix(double v, bool& bad) {
  if (!(fabs(v) < FIX_MAX)) { bad = true; return 0; }
  return llrint(v * FIX_SCALE);
}

__device__ inline void fix_add(unsigned long long* p, long long v) {
  atomicAdd(p, (unsigned long long)v);
}

// K4: FORCE's pair terms (force.f90:86-140) once per pair, over the forward
// entries.  Atom i's share is summed in fixed point in registers and added at
// the end; atom j's share (-rx s) is added to facc atomically.
// e_part[3*blockIdx.x + {0,1,2}] = sums of phi, q_i q_j erfc(ar)/r and the
// CTIP energy over the block's forward pairs (FORCE's Epair, Edir/kc,
// Ectip/kc), in FP64 in list order.  T = float (MIXED): FP32 pair terms;
// pairs closer than R_NEAR go to the queue nq (k_force_near_q).
// rdf != null (RDF_BINS unsigned ints of dynamic LDS): the pair histogram of
// FORCE (force.f90:55,76,159,180) with one count for each orientation of a
// pair, rdf[(bin-1)*25 + (s_b-1)*5 + (s_a-1)], so that rdf(a,b) + rdf(b,a)
// is the reference's output (util.f90:270-277).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_force_half(Geom g, const Tables* __restrict__ tb,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ q,
                                                      const int* __restrict__ id,
                                                      const int* __restrict__ cnt_fwd,
                                                      const int* __restrict__ col,
                                                      const T* __restrict__ dF,
                                                      unsigned long long* __restrict__ facc,
                                                      double* __restrict__ e_part,
                                                      unsigned long long* __restrict__ rdf,
                                                      double dr, int2* __restrict__ nq,
                                                      int* __restrict__ nq_n, int nq_cap,
                                                      int* __restrict__ err) {
  __shared__ double lds[3 * BLOCK / WAVE];
  extern __shared__ unsigned int hist[];
  if (rdf) {
    for (int k = threadIdx.x; k < RDF_BINS; k += blockDim.x) hist[k] = 0u;
    __syncthreads();
  }
  const int i = pair_block(g) * blockDim.x + threadIdx.x;
  double es[3] = {0.0, 0.0, 0.0};
  if (i < g.n) {
    const int si = id[i];
    bool bad = false;
    if (si >= 1 && si <= NSPEC) {
      const T qi = T(q[i]);
      const T dFi = dF[i];
      const T a = T(g.alpha);
      const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
      long long ax = 0, ay = 0, az = 0;
      double ep = 0.0, ed = 0.0, ec = 0.0;
      const size_t n = size_t(g.n);
      const int c = cnt_fwd[i];
      for (int k = 0; k < c; ++k) {
        const int j = col[size_t(k) * n + i];
        const int sj = id[j];
        if (sj < 1 || sj > NSPEC) continue;
        const PairCoef& pc = tb->pc[pidx(si, sj)];
        if (!pc.valid) continue;
        double rxd[3], r2d;
        min_image(g, xi, yi, zi, x[3 * j], x[3 * j + 1], x[3 * j + 2], rxd, r2d);
        if (rdf) {
          const int bin = int(sqrt(r2d) / dr);  // INT(r/pair%dr)
          if (bin >= 1 && bin <= NRDF + 1) {
            atomicAdd(&hist[(bin - 1) * 25 + (sj - 1) * 5 + (si - 1)], 1u);
            atomicAdd(&hist[(bin - 1) * 25 + (si - 1) * 5 + (sj - 1)], 1u);
          }
        }
        if (near_pair<T>(r2d)) {
          const int p = atomicAdd(nq_n, 1);
          if (p < nq_cap) nq[p] = make_int2(i, j);
          continue;
        }
        T e0, e1, e2;
        const double s = double(pair_force<T>(tb, pc, si, sj, T(sqrt(r2d)), T(r2d), qi, T(q[j]),
                                              dFi, dF[j], r2d < g.ream2, a, e0, e1, e2));
        ep += double(e0);
        ed += double(e1);
        ec += double(e2);
        const long long cxv = to_fix(rxd[0] * s, bad);
        const long long cyv = to_fix(rxd[1] * s, bad);
        const long long czv = to_fix(rxd[2] * s, bad);
        ax += cxv; ay += cyv; az += czv;
        fix_add(&facc[3 * j], -cxv);
        fix_add(&facc[3 * j + 1], -cyv);
        fix_add(&facc[3 * j + 2], -czv);
      }
      fix_add(&facc[3 * i], ax);
      fix_add(&facc[3 * i + 1], ay);
      fix_add(&facc[3 * i + 2], az);
      es[0] = ep;
      es[1] = ed;
      es[2] = ec;
    }
    if (bad) SX_ERR_SET(err, SIMOX_ERR_FORCE_RANGE);
  }
  block_sum<3>(es, lds);
  if (threadIdx.x == 0) {
    e_part[3 * blockIdx.x] = es[0];
    e_part[3 * blockIdx.x + 1] = es[1];
    e_part[3 * blockIdx.x + 2] = es[2];
  }
  if (rdf) {  // block_sum ended with a barrier: the histogram is complete
    for (int k = threadIdx.x; k < RDF_BINS; k += blockDim.x)
      if (hist[k]) atomicAdd(&rdf[k], (unsigned long long)hist[k]);
  }
}

// K4 (default): FORCE's pair terms over the whole list, each atom summing its
// own force in FP64 in list order: no atomics, deterministic, every pair
// evaluated from both of its atoms.  A backward entry is the reference's pair
// (j, i): its terms are evaluated with j as "i" (the pair table of (s_j, s_i)
// and the arguments swapped), so that AL_NI's orientation holds, and its
// separation r_ji = -r_ij gives F_i += r_ij s_ji.  Energies: half of every
// entry's.  T = float (MIXED): pairs closer than R_NEAR are skipped; their
// forward entries go to the near queue, which adds both atoms' shares in fixed
// point (facc), as k_force_ctip64 adds the FP64 Slater terms of pc.ctip64
// pairs; k_force_finish then adds facc to f.  rdf: as k_force_half, one count
// per entry.  U: list entries whose neighbour data is loaded ahead.
template <class T, int U = 1>
__global__ void __launch_bounds__(BLOCK) k_force_full(Geom g, const Tables* __restrict__ tb,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ q,
                                                      const int* __restrict__ id,
                                                      const int* __restrict__ cnt,
                                                      const int* __restrict__ cnt_fwd,
                                                      const int* __restrict__ col,
                                                      const T* __restrict__ dF,
                                                      double* __restrict__ f,
                                                      double* __restrict__ e_part,
                                                      unsigned long long* __restrict__ rdf,
                                                      double dr, int2* __restrict__ nq,
                                                      int* __restrict__ nq_n, int nq_cap) {
  __shared__ double lds[3 * BLOCK / WAVE];
  extern __shared__ unsigned int hist[];
  if (rdf) {
    for (int k = threadIdx.x; k < RDF_BINS; k += blockDim.x) hist[k] = 0u;
    __syncthreads();
  }
  const int i = pair_block(g) * blockDim.x + threadIdx.x;
  double es[3] = {0.0, 0.0, 0.0};
  if (i < g.n) {
    const int si = id[i];
    const T qi = T(q[i]);
    const T dFi = dF[i];
    const T a = T(g.alpha);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double ep = 0.0, ed = 0.0, ec = 0.0;
    const size_t n = size_t(g.n);
    if (si >= 1 && si <= NSPEC) {
      const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
      const int c = cnt[i], cf = cnt_fwd[i];
      // U entries' neighbour data loaded ahead of their use, then the
      // entries taken in list order (the same sums as one at a time)
      for (int k0 = 0; k0 < c; k0 += U) {
        int jv[U], sv[U];
        double xv[U][3];
        T qv[U], dv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) jv[u] = k0 + u < c ? col[size_t(k0 + u) * n + i] : i;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int j = jv[u];
          sv[u] = id[j];
          xv[u][0] = x[3 * j];
          xv[u][1] = x[3 * j + 1];
          xv[u][2] = x[3 * j + 2];
          qv[u] = T(q[j]);
          dv[u] = dF[j];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = k0 + u;
          if (k >= c) break;
          const int j = jv[u];
          const int sj = sv[u];
          if (sj < 1 || sj > NSPEC) continue;
          const bool fwd = k < cf;
          const PairCoef& pc = fwd ? tb->pc[pidx(si, sj)] : tb->pc[pidx(sj, si)];
          if (!pc.valid) continue;
          double rxd[3], r2d;
          min_image(g, xi, yi, zi, xv[u][0], xv[u][1], xv[u][2], rxd, r2d);
          if (rdf) {
            const int bin = int(sqrt(r2d) / dr);  // INT(r/pair%dr)
            if (bin >= 1 && bin <= NRDF + 1)
              atomicAdd(&hist[(bin - 1) * 25 + (sj - 1) * 5 + (si - 1)], 1u);
          }
          if (near_pair<T>(r2d)) {
            if (fwd) {
              const int p = atomicAdd(nq_n, 1);
              if (p < nq_cap) nq[p] = make_int2(i, j);
            }
            continue;
          }
          const T qj = qv[u], dFj = dv[u];
          // one evaluation with the pair's own "i" first (selects, no branch)
          const int sa = fwd ? si : sj, sb = fwd ? sj : si;
          const T qa = fwd ? qi : qj, qb = fwd ? qj : qi;
          const T da = fwd ? dFi : dFj, db = fwd ? dFj : dFi;
          T e0, e1, e2;
          const double s = double(pair_force<T>(tb, pc, sa, sb, T(sqrt(r2d)), T(r2d), qa, qb,
                                                da, db, r2d < g.ream2, a, e0, e1, e2));
          ep += double(e0);
          ed += double(e1);
          ec += double(e2);
          fx += rxd[0] * s;
          fy += rxd[1] * s;
          fz += rxd[2] * s;
        }
      }
    }
    f[3 * i] = fx;
    f[3 * i + 1] = fy;
    f[3 * i + 2] = fz;
    es[0] = 0.5 * ep;
    es[1] = 0.5 * ed;
    es[2] = 0.5 * ec;
  }
  block_sum<3>(es, lds);
  if (threadIdx.x == 0) {
    e_part[3 * blockIdx.x] = es[0];
    e_part[3 * blockIdx.x + 1] = es[1];
    e_part[3 * blockIdx.x + 2] = es[2];
  }
  if (rdf) {  // block_sum ended with a barrier: the histogram is complete
    for (int k = threadIdx.x; k < RDF_BINS; k += blockDim.x)
      if (hist[k]) atomicAdd(&rdf[k], (unsigned long long)hist[k]);
  }
}

// MIXED: the queued near pairs (i, j) in FP64, both atoms' shares into facc;
// their energies in fixed point (2^32 per eV), one atomic per block into
// eacc[0..2].  A grid-stride loop over min(*nq_n, nq_cap) pairs.
__global__ void __launch_bounds__(BLOCK) k_force_near_q(Geom g, const Tables* __restrict__ tb,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ q,
                                                        const int* __restrict__ id,
                                                        const float* __restrict__ dF,
                                                        const int2* __restrict__ nq,
                                                        const int* __restrict__ nq_n, int nq_cap,
                                                        unsigned long long* __restrict__ facc,
                                                        unsigned long long* __restrict__ eacc,
                                                        int* __restrict__ err) {
  __shared__ long long lds[3 * BLOCK / WAVE];
  const int total = min(atomicAdd(const_cast<int*>(nq_n), 0), nq_cap);
  long long es[3] = {0, 0, 0};
  bool bad = false;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
    const int2 e = nq[p];
    const int i = e.x, j = e.y;
    const int si = id[i], sj = id[j];
    const PairCoef& pc = tb->pc[pidx(si, sj)];
    double rxd[3], r2d;
    pair_sep(g, x, i, j, rxd, r2d);
    double e0, e1, e2;
    const double s = pair_force<double>(tb, pc, si, sj, sqrt(r2d), r2d, q[i], q[j], double(dF[i]),
                                        double(dF[j]), r2d < g.ream2, g.alpha, e0, e1, e2);
    const double ev[3] = {e0, e1, e2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const long long v = to_fix(rxd[d] * s, bad);
      fix_add(&facc[3 * i + d], v);
      fix_add(&facc[3 * j + d], -v);
      if (!(fabs(ev[d]) < EFIX_MAX)) bad = true;
      else es[d] += llrint(ev[d] * EFIX_SCALE);
    }
  }
  if (bad) SX_ERR_SET(err, SIMOX_ERR_FORCE_RANGE);
  block_sum<3>(es, lds);
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; ++d)
      if (es[d]) fix_add(&eacc[d], es[d]);
}

// MIXED: the CTIP Slater terms (force.f90:101-140) of the forward pairs whose
// species have close exponents (pc.ctip64: the metal-metal cross pairs), in
// FP64.  Their [fi|fj] coefficients fafb grow like 1/(xi_i - xi_j)^3
// (datafmt.f90:203-214) and cancel to a term a thousand times smaller, which
// FP32 cannot resolve; k_force_half leaves these terms out.  Pairs closer than
// R_NEAR are the near queue's, which evaluates them whole.  Energies in fixed
// point into eacc[2] (the CTIP energy / kc) as k_force_near_q's.
__global__ void __launch_bounds__(BLOCK) k_force_ctip64(Geom g, const Tables* __restrict__ tb,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ q,
                                                        const int* __restrict__ id,
                                                        const int* __restrict__ cnt_fwd,
                                                        const int* __restrict__ col,
                                                        unsigned long long* __restrict__ facc,
                                                        unsigned long long* __restrict__ eacc,
                                                        int* __restrict__ err) {
  __shared__ long long lds[BLOCK / WAVE];
  const int i = pair_block(g) * blockDim.x + threadIdx.x;
  long long es[1] = {0};
  bool bad = false;
  if (i < g.n) {
    const int si = id[i];
    if (si >= 1 && si <= NSPEC) {
      const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2], qi = q[i];
      const size_t n = size_t(g.n);
      const int c = cnt_fwd[i];
      long long ax = 0, ay = 0, az = 0;
      for (int k = 0; k < c; ++k) {
        const int j = col[size_t(k) * n + i];
        const int sj = id[j];
        if (sj < 1 || sj > NSPEC) continue;
        const PairCoef& pc = tb->pc[pidx(si, sj)];
        if (!pc.valid || !pc.ctip64) continue;
        double rx[3], r2;
        min_image(g, xi, yi, zi, x[3 * j], x[3 * j + 1], x[3 * j + 2], rx, r2);
        if (near_pair<float>(r2)) continue;
        double dfc, ec;
        ctip_force_terms<double>(pc, sqrt(r2), r2, qi, q[j], dfc, ec);
        if (!(fabs(ec) < EFIX_MAX)) bad = true;
        else es[0] += llrint(ec * EFIX_SCALE);
        const long long cxv = to_fix(rx[0] * dfc, bad);
        const long long cyv = to_fix(rx[1] * dfc, bad);
        const long long czv = to_fix(rx[2] * dfc, bad);
        ax += cxv; ay += cyv; az += czv;
        fix_add(&facc[3 * j], -cxv);
        fix_add(&facc[3 * j + 1], -cyv);
        fix_add(&facc[3 * j + 2], -czv);
      }
      fix_add(&facc[3 * i], ax);
      fix_add(&facc[3 * i + 1], ay);
      fix_add(&facc[3 * i + 2], az);
    }
  }
  if (bad) SX_ERR_SET(err, SIMOX_ERR_FORCE_RANGE);
  block_sum<1>(es, lds);
  if (threadIdx.x == 0 && es[0]) fix_add(&eacc[2], es[0]);
}

// f = facc in eV/A (add: f += facc, after k_force_full), and facc back to
// zero for the next pass.  Thread 0 also
// moves the near-pair energies (eacc, MIXED) into row e_row of the energy
// partials, resets eacc and the near queue.  The sums are read and cleared
// with atomic exchanges, the operations that built them: a plain load here
// has returned a stale zero for one of the eacc words.
__global__ void k_force_finish(int n, unsigned long long* __restrict__ facc,
                               double* __restrict__ f, int add,
                               unsigned long long* __restrict__ eacc, double* __restrict__ e_row,
                               int* __restrict__ nq_n, int nq_cap, int* __restrict__ err) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * n) {
    const double v = double((long long)atomicExch(&facc[t], 0ull)) * FIX_INV;
    f[t] = add ? f[t] + v : v;
  }
  if (t == 0 && e_row) {
    for (int d = 0; d < 3; ++d) e_row[d] = double((long long)atomicExch(&eacc[d], 0ull)) * EFIX_INV;
    if (atomicExch(nq_n, 0) > nq_cap) atomicCAS(err, 0, SIMOX_ERR_NB_OVERFLOW);
  }
}

}  // namespace sx
