// C-ABI and host orchestration of the SIMOX hot path on MI355X (gfx950).
// Kernels live in the headers included below; this file owns the device
// state (one context = one system), the launch sequences, the charge-solver
// control logic and the error mapping to the reference's STOP texts.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "cells.hpp"
#include "common.hpp"
#include "pair.hpp"
#include "tile.hpp"
#include "pme.hpp"
#include "solve.hpp"
#include "adjoint.hpp"

using namespace sx;

namespace {

const char* status_text(int code, int where_ctip) {
  switch (code) {
    case SIMOX_ERR_CELL_OVERFLOW: return "=== Cell particle limit crashed !!! ===";
    case SIMOX_ERR_NB_OVERFLOW:
      return where_ctip ? "=== neighbor particle limit crashed !!! === "
                        : "=== Memeory error for EAM density ===";
    case SIMOX_ERR_UNKNOWN_PAIR: return "=== Unknown particle pair ===";
    case SIMOX_ERR_BRENT: return "# **** Brent solver limit breached *** ";
    case SIMOX_ERR_FORCE_RANGE: return "pair force beyond the range of the fixed-point force sums";
    default: return "unknown error";
  }
}

int grid_blocks(long long n, int b = BLOCK) { return int((n + b - 1) / b); }

bool make_plan(int L, FftPlan* pl) {
  pl->L = L;
  pl->nst = 0;
  int m = L;
  while (m % 4 == 0 && m > 1) { pl->rad[pl->nst++] = 4; m /= 4; }
  while (m % 2 == 0 && m > 1) { pl->rad[pl->nst++] = 2; m /= 2; }
  while (m % 3 == 0 && m > 1) { pl->rad[pl->nst++] = 3; m /= 3; }
  while (m % 5 == 0 && m > 1) { pl->rad[pl->nst++] = 5; m /= 5; }
  return m == 1 && pl->nst <= 16;
}

struct TimedLaunch { int which; hipEvent_t a, b; };

}  // namespace

// Progress of a converged charge solve between host decisions.
enum { NCG_IDLE = 0, NCG_WAIT_RESTART = 1, NCG_WAIT_ITER = 2, NCG_DONE = 3 };
struct NcgRun {
  int phase = NCG_IDLE;
  int total_it = 0, nev = 0, restarts = 0, it = 0;
  int batch = 1;        // inner iterations issued before the next host check
  double rms = 0.0, target = 0.0;
  double rms0 = 0.0;    // residual at the restart that began this inner solve
  bool ok = false;
};

// Launch shapes of the FFT passes for one precision: mx real x-lines per
// R2C / C2R workgroup, my kz-columns per y workgroup, mz z-lines per fused z
// workgroup, and their dynamic LDS bytes.
struct FftCfg {
  int mx = 0, my = 0, mz = 0, zblocks = 0;
  int zreg = 0;  // 512 or 256: the register-resident z pass for that Kz (4 lines per workgroup)
  size_t shx = 0, shy = 0, shz = 0;
  // plane-resident chain (FP32 coarse grids with Kx = Ky <= 128): x and y in
  // one pass per direction, z pass over nl lines per workgroup
  int plane = 0, nl = 8, zpblocks = 0, xynt = 1024;
  size_t shp = 0, shzp = 0;
};

// One PME grid: sizes, FFT plans and launch shapes, twiddles, the real grid,
// the half spectrum, the influence and filter tables (same layout as the
// spectrum) and the brick lists of the spread (atoms grouped by the 4^3 grid
// brick of their base point).
struct PmeGrid {
  int K[3] = {0, 0, 0}, H = 0;
  int order = 4;  // assignment order: 4 (B-splines) or 2 (linear, preconditioner grid)
  FftPlan pl[3]{};
  FftCfg fc64{}, fc32{};
  double2* tw64[3] = {nullptr, nullptr, nullptr};
  float2* tw32[3] = {nullptr, nullptr, nullptr};
  double* greal64 = nullptr;
  float* greal32 = nullptr;
  double2* spec64 = nullptr;
  float2* spec32 = nullptr;
  double* G64 = nullptr;
  float* G32 = nullptr;
  double* F64 = nullptr;
  float* F32 = nullptr;
  int nbrick = 0;
  int *brick = nullptr, *bcount = nullptr, *bstart = nullptr, *bfill = nullptr;
  int *bunord = nullptr, *border = nullptr, *bsum = nullptr, *btotal = nullptr;
  bool bricks_valid = false;
  bool ready = false;
  bool spread32 = false;  // FP32 spreads in 32-bit fixed point (the MIXED inner grid)
  float4* brec = nullptr;  // its stencil records in brick order (k_brick_records)
  float4* brec_atom = nullptr;  // the same records in atom order
  // owner-computes spread (k_spread_own): each tile's extended-region values
  // that the fold sums, and each tile's 2^-S
  int* sp_ext = nullptr;
  int* sp_xcol = nullptr;  // the x-face columns of each tile's region, transposed
  size_t sp_xcol_n = 0;
  float* sp_tinv = nullptr;
  size_t sp_ext_n = 0;
  int sp_tiles = 0;
  int gshape = 0;  // tile shape of the record gather (GATHER_SHAPES), autotuned
};
constexpr int NGATHER_SHAPES = 6;

struct simox_ctx {
  int n = 0;
  int mode = SIMOX_MODE_FAITHFUL;
  Geom g{};
  double V = 0.0, c_pme = 0.0, rcut = 0.0;
  simox_params prm{};
  Tables htab{};
  Tables* dtab = nullptr;
  hipStream_t st = nullptr;
  // atom state in cell-sorted order
  double *x = nullptr, *v = nullptr, *f = nullptr, *q = nullptr;
  int *id = nullptr, *perm = nullptr, *cell = nullptr;
  double *x2 = nullptr, *v2 = nullptr, *f2 = nullptr, *q2 = nullptr;
  int *id2 = nullptr, *perm2 = nullptr, *cell2 = nullptr;
  int *count = nullptr, *start = nullptr, *fill = nullptr, *slot = nullptr, *order = nullptr;
  double *rho = nullptr, *dF64 = nullptr, *frec = nullptr;
  float* dF32 = nullptr;
  // half-shell force pass: forward entries per atom (the reference's "i"
  // side, first in each list), fixed-point force sums (3N, zero between
  // passes), near-pair energies (MIXED, 3) and the queue of MIXED pairs closer
  // than R_NEAR, evaluated in FP64 by k_force_near_q
  int* nb_fwd = nullptr;
  unsigned long long *facc = nullptr, *eacc = nullptr;
  int2* nq = nullptr;
  int* nq_n = nullptr;
  int nq_cap = 0;
  int tile_xcd = 1;  // k_nlist_tile: neighbouring cells on one XCD
  int tshape = 0;    // k_nlist_tile's cell-tile shape (TILE_SHAPES), autotuned
  bool tile_tuned = false, tile_fixed = false;
  bool ctip64_on = false;  // MIXED: species present form a pc.ctip64 pair
  bool force_half = false;  // K4 on the forward entries with fixed-point sums (SIMOX_FORCE_HALF=1)
  unsigned spec_present = 0u;  // bit s: species s among the atoms
  // PME: the reference's grid (pm) and the half-resolution grid of the charge
  // solver's long-range preconditioner (pc, set up on first use)
  PmeGrid pm, pc;
  // MIXED inner operator's reciprocal part on a coarser grid (pi, half the
  // resolution, order 4) when inner_coarse; the restarts keep the reference's
  // grid, so the converged charges are those of the reference's operator.
  // Inner solves stop at inner_rtol (inner_rtol_c with the coarse grid) x
  // the true residual of their restart.
  PmeGrid pi;
  bool inner_coarse = false;
  // launch shapes of the inner solve that the autotuner sets (tune_inner)
  int mv_variant = 10;      // packed matvec: 0/1 4 groups, 2/3 8 (odd: streaming loads); 4-7 the same with XCD-ordered row blocks; 8/9 4/8 groups streaming with the LDS halo; 10 as 8 with 24-bit entries
  bool inner_tuned = false;
  bool mv_fixed = false, gather_fixed = false, xynt_fixed = false, nl_fixed = false;
  // line-search rounds per inner iteration (multi-block path): adapted to the
  // rounds the previous iteration's search ran when ls_adapt
  bool ls_adapt = true;
  int ls_rounds = 8;
  bool inner_f16 = true;
  bool prec_spread32 = true;   // the preconditioner grid's FP32 spread in 32-bit fixed point
  bool spread_big = true;      // 32 x 32 x 16 tiles for the 32-bit spreads
  bool spread_own = true;      // owner-computes 32-bit spreads with a fold (k_spread_own)
  int force_batch = 4;         // k_force_full<float>: list entries loaded ahead per lane (C3: 4.58 -> 4.09 ms)
  int dens_batch = 4;          // k_density<float>: the same
  int ctip_batch = 2;          // k_ctip_matrix<double>: the same
  int prec_div = 2;            // the preconditioner grid: the context's grid / prec_div per axis
  int inner_div = 2;           // the coarse inner operator's grid: the context's grid / inner_div
  bool inner_spread32 = true;  // the coarse grid's spread in 32-bit fixed point  // with the coarse grid: FP16 matrix values in the inner matvec
  bool val16_valid = false;  // nb_val16 holds the current matrix
  double inner_rtol = 1e-4, inner_rtol_c = 3e-2;
  bool prec_on = false;
  bool small_path = false;  // single-workgroup vector passes (n <= LS_SMALL_MAX)
  // inner iterations per host check: predicted from the inner solve's rate,
  // at most max_batch (SIMOX_ITER_BATCH, 1 = a check after every iteration)
  int max_batch = 16;
  // Newton rounds launched beyond the last search's count (SIMOX_LS_SPARE;
  // C3: 0 -> 132 ms per step, 1 -> 134 ms)
  int ls_spare = 0;
  double last_rate = 0.0;  // residual reduction per iteration of the last inner solve
  double prec_rho = 0.0;  // density of 1/J per grid point the filter was built with
  double* spl_grad = nullptr;  // dE/dc of the spline-form EAM (NSPL * SPL_MAX)
  double* vjp_u = nullptr;     // 3N weights of the force vector-Jacobian product
  double* vjp_t = nullptr;     // N per-atom sums T_i of it
  double* vjp_up = nullptr;    // 3N centred weights (charge part)
  unsigned long long* dip_acc = nullptr;  // fixed-point dipole grid (charge part)
  double* sv = nullptr;
  double* cv = nullptr;  // 1/sqrt(J) per atom, the preconditioner's fixed scaling
  // CTIP
  int cap = NNB_MAX;
  int* nb_cnt = nullptr;
  int* nb_col = nullptr;
  double* nb_val64 = nullptr;
  float* nb_val32 = nullptr;
  bool val32_valid = false;
  // compacted FP16 matrix of the coarse-grid inner operator (MIXED)
  _Float16* nb_val16 = nullptr;
  int *nb_cnt16 = nullptr, *nb_col16 = nullptr;
  unsigned* nb_pk = nullptr;  // packed 32-bit entries (column + 12-bit value) when n <= 2^20
  unsigned* pk_words = nullptr;  // their words in the list's layout (k_ctip_matrix, PACK)
  // their SELL-4 slices of 64 rows: groups of 4 entries per row, offsets
  int *pk_slen = nullptr, *pk_soff = nullptr, *pk_bsum = nullptr, *pk_total = nullptr;
  bool inner_pack = true, pk_valid = false;
  // the LDS halo of each 256-row block (k_pk_halo): entries with LDS slots,
  // runs of staged atoms, staged-atom and run counts per block
  unsigned* nb_pkh = nullptr;
  unsigned* nb_pk24 = nullptr;  // the same entries in 24 bits (3 words per group of 4)
  int2* pk_runs = nullptr;
  int* pk_hcount = nullptr;
  bool pkh_valid = false;
  int halo_max = HALO_MAX;  // SIMOX_HALO_MAX (tests: blocks past it gather globally)
  double inner_drop = 1e-3;  // entries with |H0_ij| below this (eV/e^2) left out
  double* bvec = nullptr;
  double *qc = nullptr, *ql = nullptr, *gv = nullptr, *dv = nullptr, *pv = nullptr, *phi = nullptr;
  double *rv = nullptr, *zv = nullptr, *apv = nullptr, *mv = nullptr;
  float* pv32 = nullptr;  // FP32 copy of the CG direction (MIXED)
  // predictor history: qhist[0] = charges before the last solve, [1], [2] the
  // ones before; nhist of them valid (0 after any external change of q)
  double* qhist[3] = {nullptr, nullptr, nullptr};
  double* qtmp = nullptr;
  int nhist = 0;
  int pred_order = 1;
  NcgRun ncg;
  // captured inner iteration of the charge solve (batched stepping)
  bool use_graph = false;
  hipGraphExec_t iter_graph = nullptr;
  int part_gen = 0, iter_graph_gen = -1;
  // thermostat of the step (0 NVE, 1 Berendsen, 2 Langevin), target kB*T in eV
  int thermo = 0;
  bool reflect = false;  // OXIDATION-mode first half (VV_NVE1_REFLECT)
  double t_given = 0.0, alpha = 0.0;
  unsigned long long seed = 1, nvt_step = 0;
  // pair histogram of FORCE (RDF), accumulated while rdf_on
  bool rdf_on = false;
  double* adj = nullptr;  // adjoint work vectors b, lambda, r, z, p, y (6 n)
  unsigned long long* rdf = nullptr;
  double rdf_dr = 0.0;  // Rcut / Nrdf (main.f90:254)
  int last_inner = 0, last_outer = 0;
  bool matrix_valid = false;
  bool nlist_valid = false;
  double* part_eg = nullptr;  // per-block embedding-energy partials of the last list build
  // reductions / scalars
  double* part = nullptr;
  size_t part_cap = 0;
  double* dscal = nullptr;
  double* hscal = nullptr;
  int* derr = nullptr;
  unsigned* red_ctr = nullptr;  // finished workgroups of a fused reduction (zero between launches)
  int* herr = nullptr;
  // host staging
  double* hstage = nullptr;
  size_t hstage_cap = 0;
  double* dstage = nullptr;
  // solver settings
  double tol = 1e-10;
  int maxit = 1000;
  // energies of the last FORCE/PME/CTIP (eV)
  double Epair = 0, Edir = 0, Ectip = 0, Egroup = 0, Erec = 0, Eself = 0, mv2 = 0;
  bool have_state = false;
  std::string err;
  // timing
  bool timing = false;
  unsigned timing_mask = ~0u;  // categories timed while timing is on
  std::vector<TimedLaunch> tl;
  std::vector<hipEvent_t> evfree;  // recorded-and-resolved events for reuse
  double tms[16] = {0};
  long long tcnt[16] = {0};
};

namespace {

#define HIPCHK(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ctx->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                 __FILE__ + ":" + std::to_string(__LINE__);                   \
      return SIMOX_ERR_HIP;                                                   \
    }                                                                         \
  } while (0)

#define CHK(x)               \
  do {                       \
    int rc_ = (x);           \
    if (rc_) return rc_;     \
  } while (0)

template <class T>
int dalloc(simox_ctx* ctx, T** p, size_t count) {
  HIPCHK(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
  return 0;
}

// Events come from a per-context pool (created once, reused after each
// resolve), so a timed launch costs two event records and no event creation:
// the host then keeps ahead of the device and the pair measures the kernel,
// not a wait for its launch.
hipEvent_t tevent(simox_ctx* ctx) {
  if (!ctx->evfree.empty()) {
    hipEvent_t e = ctx->evfree.back();
    ctx->evfree.pop_back();
    return e;
  }
  // no system-scope fence at the records: the default event's cache writeback
  // and invalidation would slow the timed kernel itself (cold L2)
  hipEvent_t e{};
  (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
  return e;
}
void tstart(simox_ctx* ctx, int which, hipEvent_t* a) {
  if (!ctx->timing || !((ctx->timing_mask >> which) & 1u)) return;
  *a = tevent(ctx);
  (void)hipEventRecord(*a, ctx->st);
}
void tstop(simox_ctx* ctx, int which, hipEvent_t a) {
  if (!ctx->timing || !((ctx->timing_mask >> which) & 1u)) return;
  hipEvent_t b = tevent(ctx);
  (void)hipEventRecord(b, ctx->st);
  ctx->tl.push_back({which, a, b});
}
// Events for one kernel launched with hipExtLaunchKernel (the runtime records
// them at the kernel's start and end: the kernel alone, as a kernel trace
// times it); both null when the category is not timed or the stream is being
// captured into a graph.
void tkernel(simox_ctx* ctx, int which, hipEvent_t* a, hipEvent_t* b) {
  *a = *b = nullptr;
  if (!ctx->timing || !((ctx->timing_mask >> which) & 1u)) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  *a = tevent(ctx);
  *b = tevent(ctx);
  ctx->tl.push_back({which, *a, *b});
}
void tresolve(simox_ctx* ctx) {
  for (auto& t : ctx->tl) {
    float ms = 0.f;
    (void)hipEventSynchronize(t.b);
    (void)hipEventElapsedTime(&ms, t.a, t.b);
    ctx->tms[t.which] += ms;
    ctx->tcnt[t.which] += 1;
    ctx->evfree.push_back(t.a);
    ctx->evfree.push_back(t.b);
  }
  ctx->tl.clear();
}

bool use32(const simox_ctx* ctx) { return ctx->mode == SIMOX_MODE_MIXED; }

void build_tables(simox_ctx* ctx) {
  const simox_params& p = ctx->prm;
  Tables& t = ctx->htab;
  memset(&t, 0, sizeof t);
  const double kc_const = KC * 2.0 * ctx->g.alpha / SQRTPI;  // ctip.f90:31
  for (int s = 1; s <= NSPEC; ++s) {
    SpecCoef& sc = t.sc[s - 1];
    const simox_potn& pp = p.prm[s - 1];
    sc.chi = pp.chi; sc.J = pp.J; sc.Jp = pp.J - kc_const; sc.Z = pp.Z;
    sc.qmin = pp.qmin; sc.qmax = pp.qmax; sc.xm = pp.xm; sc.xi = pp.xi;
    EmbedCoef& ec = t.ec[s - 1];
    if (s < 4) {
      const simox_metal& m = p.eam[s - 1];
      ec.metal = 1;
      ec.rho_e = m.rho_e; ec.rho_s = m.rho_s; ec.rho_n = 0.85 * m.rho_e; ec.rho_0 = 1.15 * m.rho_e;
      ec.eta = m.eta; ec.Fe = m.Fe;
      for (int k = 0; k < 4; ++k) { ec.Fn[k] = m.Fn[k]; ec.Fm[k] = m.Fm[k]; ec.Fp[k] = m.Fp[k]; }
    } else {
      ec.metal = 0;
      for (int k = 0; k < 5; ++k) {
        ec.re[k] = p.ox.re[k]; ec.rmax[k] = p.ox.rmax[k];
        for (int j = 0; j < 4; ++j) ec.F[k][j] = p.ox.F[k][j];
      }
    }
  }
  auto metal_dens = [&](int s) {
    const simox_metal& m = p.eam[s - 1];
    return DensCoef{m.r_e, m.f_e, m.beta, m.lambda};
  };
  auto oxy_dens = [&](int partner) {  // r_e of the partner row (Q15)
    return DensCoef{p.ox.r_e[partner - 1], p.ox.fe, p.ox.gamm, p.ox.nu};
  };
  auto oxy_phi = [&](int row) {
    const int r = row - 1;
    return PhiCoef{p.ox.r_e[r], p.ox.alph[r], p.ox.beta[r], p.ox.A[r], p.ox.B[r], p.ox.kapp[r],
                   p.ox.lamb[r]};
  };
  for (int si = 1; si <= NSPEC; ++si)
    for (int sj = 1; sj <= NSPEC; ++sj) {
      PairCoef& pc = t.pc[pidx(si, sj)];
      pc.valid = 1;
      pc.same = si == sj;
      pc.xi_i = p.prm[si - 1].xi; pc.xi_j = p.prm[sj - 1].xi;
      pc.Z_i = p.prm[si - 1].Z; pc.Z_j = p.prm[sj - 1].Z;
      for (int k = 0; k < 4; ++k) pc.c[k] = p.prm[si - 1].fafb[sj - 1][k];
      auto metal_phi = [&](int s) {
        const simox_metal& m = p.eam[s - 1];
        return PhiCoef{m.r_e, m.alpha, m.beta, m.A, m.B, m.kappa, m.lambda};
      };
      pc.kind = 0;
      // [fi|fj] of two exponents within 50 % of their mean: FP64 in MIXED
      pc.ctip64 = si != sj && std::fabs(pc.xi_i - pc.xi_j) < 0.25 * (pc.xi_i + pc.xi_j);
      if (si < 4 && sj < 4 && si == sj) {  // METAL_HOMO
        pc.di = metal_dens(si);
        pc.dj = pc.di;
        pc.ph = metal_phi(si);
      } else if (si < 4 && sj < 4 && si + sj == 3) {
        // AL_NI (eam.f90:101-168): called as AL_NI(Al, Ni) for both orders of
        // the pair (force.f90:60-63), so fj(1) is Al's density and fj(2) Ni's
        // whichever atom is "i"; a fixed Al-Ni pair term (eam.f90:147-153)
        pc.di = metal_dens(1);
        pc.dj = metal_dens(2);
        pc.ph = PhiCoef{2.71579, 8.00443, 4.75970, 0.44254, 0.68349, 0.63279, 0.81777};
      } else if (si < 4 && sj < 4) {
        // METAL_HETERO (eam.f90:38-98): each metal's own density and the
        // density-weighted mean of the two metals' pair terms
        pc.kind = 1;
        pc.di = metal_dens(si);
        pc.dj = metal_dens(sj);
        pc.ph = metal_phi(si);
        pc.ph2 = metal_phi(sj);
      } else if (si < 4 && sj == 4) {  // METAL_OXYGEN
        pc.di = metal_dens(si);
        pc.dj = oxy_dens(si);
        pc.ph = oxy_phi(si);
      } else if (si == 4 && sj < 4) {  // OXYGEN_METAL
        pc.di = oxy_dens(sj);
        pc.dj = metal_dens(sj);
        pc.ph = oxy_phi(sj);
      } else {  // OXYGEN_ONLY
        pc.di = oxy_dens(4);
        pc.dj = pc.di;
        pc.ph = oxy_phi(4);
      }
    }
}

// k_force_ctip64 runs when two species present form a pair of close exponents.
void set_ctip64(simox_ctx* ctx) {
  ctx->ctip64_on = false;
  for (int a = 1; a <= NSPEC; ++a)
    for (int b = 1; b <= NSPEC; ++b)
      if (((ctx->spec_present >> a) & 1u) && ((ctx->spec_present >> b) & 1u) &&
          ctx->htab.pc[pidx(a, b)].ctip64)
        ctx->ctip64_on = true;
}

constexpr int NMV_VARIANTS = 11;  // k_h0matvec_sell4 shapes the autotuner times

int ensure_part(simox_ctx* ctx, size_t count) {
  if (count <= ctx->part_cap) return 0;
  if (ctx->part) HIPCHK(hipFree(ctx->part));
  ctx->part_cap = count;
  ++ctx->part_gen;  // a captured graph holding the old buffer is stale
  return dalloc(ctx, &ctx->part, count);
}

int ensure_stage(simox_ctx* ctx, size_t doubles) {
  if (doubles <= ctx->hstage_cap) return 0;
  if (ctx->hstage) HIPCHK(hipHostFree(ctx->hstage));
  if (ctx->dstage) HIPCHK(hipFree(ctx->dstage));
  HIPCHK(hipHostMalloc((void**)&ctx->hstage, doubles * sizeof(double), hipHostMallocDefault));
  ctx->hstage_cap = doubles;
  return dalloc(ctx, &ctx->dstage, doubles);
}

// Reduce nv interleaved partial values into dscal[off .. off+nv).
int reduce(simox_ctx* ctx, int nparts, int nv, int maxmask, int off) {
  if (nv > RED_MAXV) {
    ctx->err = "internal: reduction of more than RED_MAXV values";
    return SIMOX_ERR_ARG;
  }
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, ctx->st, ctx->part, nparts, nv,
                     maxmask, ctx->dscal + off);
  HIPCHK(hipGetLastError());
  return 0;
}

// Copy dscal[off .. off+nv) to the host (synchronises the stream).
int fetch(simox_ctx* ctx, int off, int nv, double* out) {
  HIPCHK(hipMemcpyAsync(ctx->hscal + off, ctx->dscal + off, nv * sizeof(double),
                        hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  for (int k = 0; k < nv; ++k) out[k] = ctx->hscal[off + k];
  return 0;
}

int check_err(simox_ctx* ctx, int where_ctip) {
  HIPCHK(hipMemcpyAsync(ctx->herr, ctx->derr, sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  const int e = *ctx->herr;
  if (e) {
    ctx->err = status_text(e, where_ctip);
    HIPCHK(hipMemsetAsync(ctx->derr, 0, sizeof(int), ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return e;
  }
  return 0;
}

__global__ void k_fill_int(int n, int v, int* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_iota(int n, int* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}
// dst[p*w + k] = src[perm[p]*w + k]  (caller order -> sorted order)
__global__ void k_from_caller(int n, int w, const int* __restrict__ perm,
                              const double* __restrict__ src, double* __restrict__ dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int c = perm[p];
  for (int k = 0; k < w; ++k) dst[size_t(p) * w + k] = src[size_t(c) * w + k];
}
// dst[perm[p]*w + k] = src[p*w + k]  (sorted order -> caller order)
__global__ void k_to_caller(int n, int w, const int* __restrict__ perm,
                            const double* __restrict__ src, double* __restrict__ dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int c = perm[p];
  for (int k = 0; k < w; ++k) dst[size_t(c) * w + k] = src[size_t(p) * w + k];
}
__global__ void k_to_caller_int(int n, const int* __restrict__ perm, const int* __restrict__ src,
                                int* __restrict__ dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  dst[perm[p]] = src[p];
}
__global__ void k_gather_f64(int n, const int* __restrict__ order, const double* __restrict__ a,
                             double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[order[i]];
}
__global__ void k_copy(int n, const double* __restrict__ a, double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
__global__ void k_int_to_double(int n, const int* __restrict__ a, double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = double(a[i]);
}

// ---------------------------------------------------------------------------
// K1: cell sort
// ---------------------------------------------------------------------------
int do_cell_sort(simox_ctx* ctx) {
  const int n = ctx->n, nb = grid_blocks(n);
  hipLaunchKernelGGL(k_fill_int, dim3(grid_blocks(ctx->g.ncall)), dim3(BLOCK), 0, ctx->st,
                     ctx->g.ncall, 0, ctx->count);
  hipLaunchKernelGGL(k_cell_assign, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->x, ctx->cell,
                     ctx->count);
  hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, ctx->st, ctx->g.ncall, ctx->count,
                     ctx->start, ctx->fill, ctx->derr);
  hipLaunchKernelGGL(k_cell_scatter, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->cell, ctx->start,
                     ctx->fill, ctx->slot);
  hipLaunchKernelGGL(k_cell_order, dim3(ctx->g.ncall), dim3(BLOCK), 0, ctx->st, ctx->start,
                     ctx->slot, ctx->perm, ctx->order);
  hipLaunchKernelGGL(k_reorder, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->order, ctx->cell,
                     ctx->x, ctx->v, ctx->f, ctx->q, ctx->id, ctx->perm, ctx->x2, ctx->v2,
                     ctx->f2, ctx->q2, ctx->id2, ctx->perm2, ctx->cell2);
  for (int h = 0; h < ctx->nhist; ++h) {
    hipLaunchKernelGGL(k_gather_f64, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->order,
                       ctx->qhist[h], ctx->qtmp);
    std::swap(ctx->qhist[h], ctx->qtmp);
  }
  HIPCHK(hipGetLastError());
  std::swap(ctx->x, ctx->x2); std::swap(ctx->v, ctx->v2); std::swap(ctx->f, ctx->f2);
  std::swap(ctx->q, ctx->q2); std::swap(ctx->id, ctx->id2); std::swap(ctx->perm, ctx->perm2);
  std::swap(ctx->cell, ctx->cell2);
  ctx->matrix_valid = false;
  ctx->nlist_valid = false;
  ctx->pm.bricks_valid = ctx->pc.bricks_valid = ctx->pi.bricks_valid = false;
  return 0;
}

// ---------------------------------------------------------------------------
// K2-K4: FORCE
// ---------------------------------------------------------------------------
// k_nlist_tile in cell-tile shape sh (TILE_SHAPES).
template <int CAP, int NT>
void launch_tile_t(simox_ctx* ctx, const TileArgs& ta) {
  hipLaunchKernelGGL((k_nlist_tile<CAP, NT>), dim3(ctx->g.ncall), dim3(NT), 0, ctx->st, ta);
}
void launch_tile(simox_ctx* ctx, int sh) {
  TileArgs ta{ctx->g, ctx->x, ctx->start, ctx->cap, ctx->nb_cnt, ctx->nb_fwd, ctx->nb_col,
              ctx->derr};
  ta.g.xcd = ctx->tile_xcd;
  switch (sh) {
    case 1: launch_tile_t<TILE_SHAPES[1][0], TILE_SHAPES[1][1]>(ctx, ta); break;
    case 2: launch_tile_t<TILE_SHAPES[2][0], TILE_SHAPES[2][1]>(ctx, ta); break;
    case 3: launch_tile_t<TILE_SHAPES[3][0], TILE_SHAPES[3][1]>(ctx, ta); break;
    case 4: launch_tile_t<TILE_SHAPES[4][0], TILE_SHAPES[4][1]>(ctx, ta); break;
    case 5: launch_tile_t<TILE_SHAPES[5][0], TILE_SHAPES[5][1]>(ctx, ta); break;
    default: launch_tile_t<TILE_SHAPES[0][0], TILE_SHAPES[0][1]>(ctx, ta); break;
  }
}

// Neighbour list (k_nlist_tile, one workgroup per cell) + EAM density and
// embedding (k_density), once per set of positions.
void launch_nlist(simox_ctx* ctx) {
  const int nb = grid_blocks(ctx->n);
  launch_tile(ctx, ctx->tshape);
  if (use32(ctx)) {
    const auto kd = ctx->dens_batch >= 4   ? &k_density<float, 4>
                    : ctx->dens_batch == 2 ? &k_density<float, 2>
                                           : &k_density<float, 1>;
    hipLaunchKernelGGL(kd, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab, ctx->x, ctx->id,
                       ctx->nb_cnt, ctx->nb_fwd, ctx->nb_col, ctx->rho, ctx->dF32, ctx->part_eg,
                       ctx->derr);
  } else
    hipLaunchKernelGGL(k_density<double>, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab,
                       ctx->x, ctx->id, ctx->nb_cnt, ctx->nb_fwd, ctx->nb_col, ctx->rho,
                       ctx->dF64, ctx->part_eg, ctx->derr);
}

bool autotune_on();
template <class F>
int time_launch(simox_ctx* ctx, F launch, float* ms);

// Autotuner of the cell tile: every shape of TILE_SHAPES timed on the first
// list build of the context, the fastest kept (the lists are the same).
int tune_tile(simox_ctx* ctx) {
  ctx->tile_tuned = true;
  if (ctx->tile_fixed || !autotune_on()) return 0;
  float best = 1e30f, ms = 0.f;
  int pick = ctx->tshape;
  for (int sh = 0; sh < NTILE_SHAPES; ++sh) {
    CHK(time_launch(ctx, [&] { launch_tile(ctx, sh); }, &ms));
    if (ms < best) { best = ms; pick = sh; }
  }
  ctx->tshape = pick;
  return 0;
}

int ensure_nlist(simox_ctx* ctx) {
  if (ctx->nlist_valid) return 0;
  if (!ctx->tile_tuned) CHK(tune_tile(ctx));
  hipEvent_t ev{};
  tstart(ctx, 8, &ev);
  launch_nlist(ctx);
  tstop(ctx, 8, ev);
  HIPCHK(hipGetLastError());
  ctx->nlist_valid = true;
  return 0;
}

// K4.  Default: k_force_full over the whole list (owner computes, FP64 sums);
// with force_half: k_force_half over the forward entries, both atoms' shares
// in fixed point.  MIXED adds the near-pair queue and, for systems with
// pc.ctip64 pairs, their FP64 Slater terms, in fixed point; k_force_finish
// brings the fixed-point sums into f.  Energy partials in ep rows [0, nb),
// the fixed-point energies (MIXED; zero otherwise) in row nb.
void launch_force(simox_ctx* ctx, double* ep) {
  const int nb = grid_blocks(ctx->n);
  unsigned long long* rdf = ctx->rdf_on ? ctx->rdf : nullptr;
  const size_t sh = rdf ? RDF_BINS * sizeof(unsigned) : 0;
  const double dr = ctx->rdf_dr;
  const bool f32 = use32(ctx);
  if (ctx->force_half) {
    if (f32)
      hipLaunchKernelGGL(k_force_half<float>, dim3(nb), dim3(BLOCK), sh, ctx->st, ctx->g,
                         ctx->dtab, ctx->x, ctx->q, ctx->id, ctx->nb_fwd, ctx->nb_col, ctx->dF32,
                         ctx->facc, ep, rdf, dr, ctx->nq, ctx->nq_n, ctx->nq_cap, ctx->derr);
    else
      hipLaunchKernelGGL(k_force_half<double>, dim3(nb), dim3(BLOCK), sh, ctx->st, ctx->g,
                         ctx->dtab, ctx->x, ctx->q, ctx->id, ctx->nb_fwd, ctx->nb_col, ctx->dF64,
                         ctx->facc, ep, rdf, dr, ctx->nq, ctx->nq_n, ctx->nq_cap, ctx->derr);
  } else {
    if (f32) {
      // list entries whose neighbour data each lane loads ahead (SIMOX_FORCE_BATCH)
      const auto kf = ctx->force_batch >= 4   ? &k_force_full<float, 4>
                      : ctx->force_batch == 2 ? &k_force_full<float, 2>
                                              : &k_force_full<float, 1>;
      hipLaunchKernelGGL(kf, dim3(nb), dim3(BLOCK), sh, ctx->st, ctx->g, ctx->dtab, ctx->x,
                         ctx->q, ctx->id, ctx->nb_cnt, ctx->nb_fwd, ctx->nb_col, ctx->dF32,
                         ctx->f, ep, rdf, dr, ctx->nq, ctx->nq_n, ctx->nq_cap);
    }
    else
      hipLaunchKernelGGL(k_force_full<double>, dim3(nb), dim3(BLOCK), sh, ctx->st, ctx->g,
                         ctx->dtab, ctx->x, ctx->q, ctx->id, ctx->nb_cnt, ctx->nb_fwd,
                         ctx->nb_col, ctx->dF64, ctx->f, ep, rdf, dr, ctx->nq, ctx->nq_n,
                         ctx->nq_cap);
  }
  if (f32) {
    if (ctx->ctip64_on)
      hipLaunchKernelGGL(k_force_ctip64, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab,
                         ctx->x, ctx->q, ctx->id, ctx->nb_fwd, ctx->nb_col, ctx->facc, ctx->eacc,
                         ctx->derr);
    hipLaunchKernelGGL(k_force_near_q, dim3(std::max(1, std::min(nb, 1024))), dim3(BLOCK), 0,
                       ctx->st, ctx->g, ctx->dtab, ctx->x, ctx->q, ctx->id, ctx->dF32, ctx->nq,
                       ctx->nq_n, ctx->nq_cap, ctx->facc, ctx->eacc, ctx->derr);
  }
  if (f32 || ctx->force_half)
    hipLaunchKernelGGL(k_force_finish, dim3(grid_blocks(3ll * ctx->n)), dim3(BLOCK), 0, ctx->st,
                       ctx->n, ctx->facc, ctx->f, ctx->force_half ? 0 : 1, ctx->eacc,
                       ep + size_t(3) * nb, ctx->nq_n, ctx->nq_cap, ctx->derr);
  else
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, ctx->st, 3, 0.0, ep + size_t(3) * nb);
}

int do_force(simox_ctx* ctx, bool fetch_energies) {
  const int nb = grid_blocks(ctx->n);
  CHK(ensure_nlist(ctx));
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, ctx->st, ctx->part_eg, nb, 1, 0,
                     ctx->dscal + 0);
  // block partials of k_force_half in rows [0, nb), the near pairs' in row nb
  CHK(ensure_part(ctx, size_t(nb) * 7));
  double* ep = ctx->part + nb;
  hipEvent_t ev{};
  tstart(ctx, 7, &ev);
  launch_force(ctx, ep);
  tstop(ctx, 7, ev);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, ctx->st, ep, nb + 1, 3, 0,
                     ctx->dscal + 1);
  HIPCHK(hipGetLastError());
  if (fetch_energies) {
    double e[4];
    CHK(fetch(ctx, 0, 4, e));
    ctx->Egroup = e[0];
    ctx->Epair = e[1];
    ctx->Edir = e[2] * KC;
    ctx->Ectip = e[3] * KC;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// K6-K8: one PME convolution: spread(q or q + step p) -> ... -> real grid
// ---------------------------------------------------------------------------
// Kernel pointers of the three FFT kernels for line length L (compile-time
// instantiations for the workloads' sizes, runtime plan otherwise).
template <class C, int L>
struct FftKernels {
  static constexpr auto r2c = &k_fft_x_r2c<C, L>;
  static constexpr auto c2r = &k_fft_x_c2r<C, L>;
  static constexpr auto y = &k_fft_y<C, L>;
  static constexpr auto fus = &k_fft_z_fused<C, L>;
};
template <class C>
struct FftSet {
  decltype(&k_fft_x_r2c<C, 0>) r2c;
  decltype(&k_fft_x_c2r<C, 0>) c2r;
  decltype(&k_fft_y<C, 0>) y;
  decltype(&k_fft_z_fused<C, 0>) fus;
};
template <class C, int L>
FftSet<C> fft_set_of() {
  return {FftKernels<C, L>::r2c, FftKernels<C, L>::c2r, FftKernels<C, L>::y, FftKernels<C, L>::fus};
}
template <class C>
FftSet<C> fft_set(int L) {
  switch (L) {
    case 16: return fft_set_of<C, 16>();
    case 24: return fft_set_of<C, 24>();
    case 32: return fft_set_of<C, 32>();
    case 48: return fft_set_of<C, 48>();
    case 64: return fft_set_of<C, 64>();
    case 96: return fft_set_of<C, 96>();
    case 128: return fft_set_of<C, 128>();
    case 192: return fft_set_of<C, 192>();
    case 256: return fft_set_of<C, 256>();
    case 384: return fft_set_of<C, 384>();
    case 512: return fft_set_of<C, 512>();
    default: return fft_set_of<C, 0>();
  }
}

// Geometry of the kernels on grid pm: the context's, with pm's sizes.
Geom geom_of(const simox_ctx* ctx, const PmeGrid& pm) {
  Geom g = ctx->g;
  for (int d = 0; d < 3; ++d) g.K[d] = pm.K[d];
  g.H = pm.H;
  return g;
}

// Event timing of the PME passes: categories 0-4 on the reference's grid,
// 11-15 (spread, x R2C, y, z, x C2R) on the MIXED inner operator's coarse grid.
bool timed_grid(const simox_ctx* ctx, const PmeGrid& pm) { return &pm == &ctx->pm || &pm == &ctx->pi; }
int tcat(const simox_ctx* ctx, const PmeGrid& pm, int k) { return &pm == &ctx->pi ? 11 + k : k; }

// Kernels of the plane-resident chain for plane size L (Kx = Ky) and Kz.
using XyFwd = decltype(&k_fft_xy_r2c<128, 256>);
using XyInv = decltype(&k_fft_xy_c2r<128, 256>);
using ZpFn = decltype(&k_fft_zp<256, 16>);
template <int L>
void xy_of(int nt, XyFwd* f, XyInv* b) {
  if (nt == 1024) { *f = &k_fft_xy_r2c<L, 1024>; *b = &k_fft_xy_c2r<L, 1024>; }
  else if (nt == 512) { *f = &k_fft_xy_r2c<L, 512>; *b = &k_fft_xy_c2r<L, 512>; }
  else { *f = &k_fft_xy_r2c<L, 256>; *b = &k_fft_xy_c2r<L, 256>; }
}
bool xy_kernels(int L, int nt, XyFwd* f, XyInv* b) {
  switch (L) {
    case 32: xy_of<32>(nt, f, b); return true;
    case 48: xy_of<48>(nt, f, b); return true;
    case 64: xy_of<64>(nt, f, b); return true;
    case 96: xy_of<96>(nt, f, b); return true;
    case 128: xy_of<128>(nt, f, b); return true;
    default: return false;
  }
}
template <int NL>
ZpFn zp_of(int KZ) {
  switch (KZ) {
    case 32: return &k_fft_zp<32, NL>;
    case 48: return &k_fft_zp<48, NL>;
    case 64: return &k_fft_zp<64, NL>;
    case 96: return &k_fft_zp<96, NL>;
    case 128: return &k_fft_zp<128, NL>;
    case 192: return &k_fft_zp<192, NL>;
    case 256: return &k_fft_zp<256, NL>;
    default: return nullptr;
  }
}
ZpFn zp_kernel(int KZ, int nl) {
  switch (nl) {
    case 4: return zp_of<4>(KZ);
    case 8: return zp_of<8>(KZ);
    case 11: return zp_of<11>(KZ);
    case 16: return zp_of<16>(KZ);
    default: return nullptr;
  }
}

int plane_chain(simox_ctx* ctx, PmeGrid& pm, float2* spec, float2* const* tw, const FftCfg& fc,
                const float* G, float* greal) {
  const Geom g = geom_of(ctx, pm);
  XyFwd fwd{};
  XyInv inv{};
  const ZpFn zp = zp_kernel(g.K[2], fc.nl);
  if (!xy_kernels(g.K[0], fc.xynt, &fwd, &inv) || !zp || g.K[1] != g.K[0]) {
    ctx->err = "plane FFT chain: no kernels for these grid sizes";
    return SIMOX_ERR_ARG;
  }
  const bool timed = timed_grid(ctx, pm);
  hipEvent_t ev{};
  if (timed) tstart(ctx, tcat(ctx, pm, 1), &ev);
  hipLaunchKernelGGL(fwd, dim3(g.K[2]), dim3(fc.xynt), fc.shp, ctx->st, tw[0], greal, spec);
  if (timed) tstop(ctx, tcat(ctx, pm, 1), ev);
  if (timed) tstart(ctx, tcat(ctx, pm, 3), &ev);
  hipLaunchKernelGGL(zp, dim3(fc.zpblocks), dim3(BLOCK), fc.shzp, ctx->st, g, tw[2], spec, G,
                     ctx->part);
  if (timed) tstop(ctx, tcat(ctx, pm, 3), ev);
  if (timed) tstart(ctx, tcat(ctx, pm, 4), &ev);
  hipLaunchKernelGGL(inv, dim3(g.K[2]), dim3(fc.xynt), fc.shp, ctx->st, tw[0], spec, greal);
  if (timed) tstop(ctx, tcat(ctx, pm, 4), ev);
  HIPCHK(hipGetLastError());
  return 0;
}

template <class C>
int pme_chain(simox_ctx* ctx, PmeGrid& pm, C* spec, C* const* tw, const FftCfg& fc,
              const decltype(C{}.x)* G, decltype(C{}.x)* greal) {
  if constexpr (sizeof(C) == sizeof(float2)) {
    if (fc.plane) return plane_chain(ctx, pm, spec, tw, fc, G, greal);
  }
  const Geom g = geom_of(ctx, pm);
  const int H = g.H, Ky = g.K[1], Kz = g.K[2];
  const FftSet<C> fx = fft_set<C>(g.K[0]), fy = fft_set<C>(Ky), fz = fft_set<C>(Kz);
  const bool timed = timed_grid(ctx, pm);
  hipEvent_t ev{};
  if (timed) tstart(ctx, tcat(ctx, pm, 1), &ev);
  hipLaunchKernelGGL(fx.r2c, dim3(Kz / fc.mx, Ky), dim3(BLOCK), fc.shx, ctx->st, g, pm.pl[0],
                     tw[0], greal, spec, fc.mx);
  if (timed) tstop(ctx, tcat(ctx, pm, 1), ev);
  if (timed) tstart(ctx, tcat(ctx, pm, 2), &ev);
  hipLaunchKernelGGL(fy.y, dim3(Kz / fc.my, H), dim3(BLOCK), fc.shy, ctx->st, g, pm.pl[1], tw[1],
                     spec, fc.my, Ky + 1, -1);
  if (timed) tstop(ctx, tcat(ctx, pm, 2), ev);
  if (timed) tstart(ctx, tcat(ctx, pm, 3), &ev);
  if (fc.zreg == 512)
    hipLaunchKernelGGL(k_fft_z512_reg<C>, dim3(fc.zblocks), dim3(BLOCK), 0, ctx->st, g, tw[2],
                       spec, G, ctx->part);
  else if (fc.zreg == 256)
    hipLaunchKernelGGL(k_fft_z256_reg<C>, dim3(fc.zblocks), dim3(BLOCK), 0, ctx->st, g, tw[2],
                       spec, G, ctx->part);
  else
    hipLaunchKernelGGL(fz.fus, dim3(fc.zblocks), dim3(BLOCK), fc.shz, ctx->st, g, pm.pl[2], tw[2],
                       spec, fc.mz, G, ctx->part);
  if (timed) tstop(ctx, tcat(ctx, pm, 3), ev);
  if (timed) tstart(ctx, tcat(ctx, pm, 2), &ev);
  hipLaunchKernelGGL(fy.y, dim3(Kz / fc.my, H), dim3(BLOCK), fc.shy, ctx->st, g, pm.pl[1], tw[1],
                     spec, fc.my, Ky + 1, +1);
  if (timed) tstop(ctx, tcat(ctx, pm, 2), ev);
  if (timed) tstart(ctx, tcat(ctx, pm, 4), &ev);
  hipLaunchKernelGGL(fx.c2r, dim3(Kz / fc.mx, Ky), dim3(BLOCK), fc.shx, ctx->st, g, pm.pl[0],
                     tw[0], spec, greal, fc.mx);
  if (timed) tstop(ctx, tcat(ctx, pm, 4), ev);
  HIPCHK(hipGetLastError());
  return 0;
}

// Launch shapes of the FFT passes.  LDS bytes per workgroup of each pass for
// tile size m (see the kernels' LDS layouts).
constexpr size_t FFT_LDS_CAP = 80 * 1024;    // default rule: two workgroups per CU
constexpr size_t FFT_LDS_MAX = 159 * 1024;   // gfx950: 160 KB per workgroup, less static LDS
template <class C>
size_t fft_shx(const Geom& g, int m) {
  return (size_t(m / 2) * (g.K[0] + 1) + xpass_b1(m, g.K[0]) + g.K[0]) * sizeof(C);
}
template <class C>
size_t fft_shy(const Geom& g, int m) { return (2 * size_t(m) * (g.K[1] + 1) + g.K[1]) * sizeof(C); }
template <class C>
size_t fft_shz(const Geom& g, int m) { return (2 * size_t(m) * g.K[2] + g.K[2]) * sizeof(C); }

template <class C>
void fft_finish(const Geom& g, FftCfg* c) {
  c->shx = fft_shx<C>(g, c->mx);
  c->shy = fft_shy<C>(g, c->my);
  c->shz = fft_shz<C>(g, c->mz);
  c->zblocks = (g.H * g.K[1] + c->mz - 1) / c->mz;
}

// Default shapes: the largest x / y tiles within FFT_LDS_CAP (they divide
// Kz), z tiles of at most 4 lines and 40 KB (SIMOX_FFT_MZ overrides).  The kernels may use up to the
// device's LDS per workgroup, which the tuner below can choose.
template <class C>
int choose_fft(simox_ctx* ctx, const Geom& g, FftCfg* c) {
  const int Kz = g.K[2];
  *c = FftCfg{};
  for (int m = 32; m >= 2 && !c->mx; m /= 2)
    if (Kz % m == 0 && fft_shx<C>(g, m) <= FFT_LDS_CAP) c->mx = m;
  for (int m = 32; m >= 1 && !c->my; m /= 2)
    if (Kz % m == 0 && fft_shy<C>(g, m) <= FFT_LDS_CAP) c->my = m;
  // z: small tiles (many workgroups per CU hide the latency of the fused pass)
  for (int m = 4; m >= 1 && !c->mz; m /= 2)
    if (fft_shz<C>(g, m) <= 40 * 1024) c->mz = m;
  if (const char* e = std::getenv("SIMOX_FFT_MZ")) {
    const int m = std::atoi(e);
    if (m >= 1 && m <= 64 && fft_shz<C>(g, m) <= FFT_LDS_MAX) c->mz = m;
  }
  // Kz = 512 outside the FAITHFUL mode: the register-resident z pass, four
  // lines per workgroup (its energy partials group four lines, like mz = 4)
  const char* zr = std::getenv("SIMOX_FFT_ZREG");
  if (Kz == 512 && ctx->mode != SIMOX_MODE_FAITHFUL && !(zr && std::atoi(zr) == 0)) {
    c->zreg = 512;
    c->mz = 4;
  }
  // Kz = 256 on the FP32 chain: a radix-4 register pass (its rounding differs
  // from the tiled 8 8 4 plan, so the FP64 chains keep the tiled one)
  if (Kz == 256 && sizeof(C) == sizeof(float2) && !(zr && std::atoi(zr) == 0)) {
    c->zreg = 256;
    c->mz = 4;
  }
  if (!c->mx || !c->my || !c->mz) return SIMOX_ERR_ARG;
  fft_finish<C>(g, c);
  const FftSet<C> fx = fft_set<C>(g.K[0]), fy = fft_set<C>(g.K[1]), fz = fft_set<C>(Kz);
  const void* fns[4] = {reinterpret_cast<const void*>(fx.r2c), reinterpret_cast<const void*>(fx.c2r),
                        reinterpret_cast<const void*>(fy.y), reinterpret_cast<const void*>(fz.fus)};
  for (const void* f : fns)
    HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(FFT_LDS_MAX)));
  return 0;
}

// Average time (ms) of `launch` over three runs after one warm-up run.
template <class F>
int time_launch(simox_ctx* ctx, F launch, float* ms) {
  hipEvent_t a{}, b{};
  HIPCHK(hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
  HIPCHK(hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
  launch();
  HIPCHK(hipEventRecord(a, ctx->st));
  for (int r = 0; r < 3; ++r) launch();
  HIPCHK(hipEventRecord(b, ctx->st));
  HIPCHK(hipEventSynchronize(b));
  HIPCHK(hipEventElapsedTime(ms, a, b));
  *ms /= 3.0f;
  HIPCHK(hipEventDestroy(a));
  HIPCHK(hipEventDestroy(b));
  HIPCHK(hipGetLastError());
  return 0;
}

// Autotuner of the x (R2C + C2R) and y tile sizes on this context's grid and
// buffers: every power-of-two tile that divides Kz and fits the device's LDS
// is timed, the fastest kept.  The tile sizes do not change any result (each
// line is transformed by the same arithmetic); the z tile, which groups the
// energy partial sums, keeps its default so energies stay reproducible.
template <class C>
int tune_fft(simox_ctx* ctx, PmeGrid& pm, FftCfg* c, C* spec, C* const* tw,
             decltype(C{}.x)* greal) {
  const Geom g = geom_of(ctx, pm);
  const int Ky = g.K[1], Kz = g.K[2], H = g.H;
  const FftSet<C> fx = fft_set<C>(g.K[0]), fy = fft_set<C>(Ky);
  float best = 1e30f;
  int pick = c->mx;
  for (int m = 2; m <= 64; m *= 2) {
    if (Kz % m || fft_shx<C>(g, m) > FFT_LDS_MAX) continue;
    const size_t sh = fft_shx<C>(g, m);
    float ms = 0.f;
    CHK(time_launch(ctx, [&] {
      hipLaunchKernelGGL(fx.r2c, dim3(Kz / m, Ky), dim3(BLOCK), sh, ctx->st, g, pm.pl[0], tw[0],
                         greal, spec, m);
      hipLaunchKernelGGL(fx.c2r, dim3(Kz / m, Ky), dim3(BLOCK), sh, ctx->st, g, pm.pl[0], tw[0],
                         spec, greal, m);
    }, &ms));
    if (ms < best) { best = ms; pick = m; }
  }
  c->mx = pick;
  best = 1e30f;
  pick = c->my;
  for (int m = 1; m <= 64; m *= 2) {
    if (Kz % m || fft_shy<C>(g, m) > FFT_LDS_MAX) continue;
    const size_t sh = fft_shy<C>(g, m);
    float ms = 0.f;
    CHK(time_launch(ctx, [&] {
      hipLaunchKernelGGL(fy.y, dim3(Kz / m, H), dim3(BLOCK), sh, ctx->st, g, pm.pl[1], tw[1], spec,
                         m, Ky + 1, -1);
    }, &ms));
    if (ms < best) { best = ms; pick = m; }
  }
  c->my = pick;
  fft_finish<C>(g, c);
  return 0;
}

int autotune_grid(simox_ctx* ctx, PmeGrid& pm) {
  CHK(tune_fft<double2>(ctx, pm, &pm.fc64, pm.spec64, pm.tw64, pm.greal64));
  if (ctx->mode == SIMOX_MODE_MIXED && !pm.fc32.plane)
    CHK(tune_fft<float2>(ctx, pm, &pm.fc32, pm.spec32, pm.tw32, pm.greal32));
  return 0;
}

int autotune(simox_ctx* ctx) { return autotune_grid(ctx, ctx->pm); }

bool autotune_on() {
  const char* e = std::getenv("SIMOX_AUTOTUNE");
  return !e || std::strcmp(e, "0") != 0;
}

int zpass_blocks(const simox_ctx* ctx, bool f32) {
  if (f32 && ctx->pm.fc32.plane) return ctx->pm.fc32.zpblocks;
  return f32 ? ctx->pm.fc32.zblocks : ctx->pm.fc64.zblocks;
}

// Brick binning of the atoms for the tiled spread (once per set of positions).
int ensure_bricks(simox_ctx* ctx, PmeGrid& pm) {
  if (pm.bricks_valid) return 0;
  const int n = ctx->n, nb = grid_blocks(n), B = pm.nbrick;
  hipLaunchKernelGGL(k_fill_int, dim3(grid_blocks(B)), dim3(BLOCK), 0, ctx->st, B, 0, pm.bcount);
  hipLaunchKernelGGL(k_fill_int, dim3(grid_blocks(B)), dim3(BLOCK), 0, ctx->st, B, 0, pm.bfill);
  hipLaunchKernelGGL(k_brick_assign, dim3(nb), dim3(BLOCK), 0, ctx->st, geom_of(ctx, pm), ctx->x, pm.brick,
                     pm.bcount);
  const int nsb = (B + 1023) / 1024;
  hipLaunchKernelGGL(k_scan_block, dim3(nsb), dim3(1024), 0, ctx->st, B, pm.bcount, pm.bstart,
                     pm.bsum);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, ctx->st, nsb, pm.bsum, pm.btotal);
  hipLaunchKernelGGL(k_scan_add, dim3(nsb), dim3(1024), 0, ctx->st, B, pm.bstart, pm.bsum,
                     pm.btotal);
  hipLaunchKernelGGL(k_brick_scatter, dim3(nb), dim3(BLOCK), 0, ctx->st, n, pm.brick, pm.bstart,
                     pm.bfill, pm.bunord);
  hipLaunchKernelGGL(k_brick_place, dim3(nb), dim3(BLOCK), 0, ctx->st, n, pm.brick, pm.bstart,
                     pm.bfill, pm.bunord, pm.border);
  if (pm.spread32)
    hipLaunchKernelGGL(k_brick_records, dim3(nb), dim3(BLOCK), 0, ctx->st, geom_of(ctx, pm), n,
                       ctx->x, pm.border, pm.brec, pm.brec_atom);
  HIPCHK(hipGetLastError());
  pm.bricks_valid = true;
  return 0;
}

// Owner-computes spread of tile TXg x TYg x TZg, then the fold of the points
// several tiles reach.
template <int ORD, int TXg, int TYg, int TZg, int NT>
int spread_own(simox_ctx* ctx, const Geom& g, PmeGrid& pm, const double* src) {
  using P = SpreadOwn<ORD, TXg, TYg, TZg>;
  const int nt = P::tiles(g);
  const size_t need = size_t(nt) * P::EXT, need_x = size_t(nt) * P::XCOL;
  if (pm.sp_ext_n < need || pm.sp_xcol_n < need_x || pm.sp_tiles < nt) {
    if (pm.sp_ext) (void)hipFree(pm.sp_ext);
    if (pm.sp_xcol) (void)hipFree(pm.sp_xcol);
    if (pm.sp_tinv) (void)hipFree(pm.sp_tinv);
    pm.sp_ext = nullptr;
    pm.sp_xcol = nullptr;
    pm.sp_tinv = nullptr;
    pm.sp_ext_n = pm.sp_xcol_n = 0;
    pm.sp_tiles = 0;
    CHK(dalloc(ctx, &pm.sp_ext, need));
    CHK(dalloc(ctx, &pm.sp_xcol, need_x));
    CHK(dalloc(ctx, &pm.sp_tinv, size_t(nt)));
    pm.sp_ext_n = need;
    pm.sp_xcol_n = need_x;
    pm.sp_tiles = nt;
  }
  hipLaunchKernelGGL((k_spread_own<ORD, TXg, TYg, TZg, NT>), dim3(nt), dim3(NT), 0, ctx->st, g,
                     src, pm.bstart, pm.border, pm.brec, pm.greal32, pm.sp_ext, pm.sp_xcol,
                     pm.sp_tinv);
  hipLaunchKernelGGL((k_spread_fold<ORD, TXg, TYg, TZg>),
                     dim3(nt * SpreadFold<ORD, TXg, TYg, TZg>::BPT), dim3(BLOCK), 0, ctx->st, g,
                     pm.sp_ext, pm.sp_xcol, pm.sp_tinv, pm.greal32);
  return 0;
}

// The 32-bit spread of a coarse FP32 grid: 32 x 32 x 16 tiles with 1024
// threads where the grid holds them, 16 x 16 x 8 tiles with 1024 threads on
// grids with more atoms than points (SIMOX_SPREAD_BIG=0: 32 x 16 x 8 tiles
// with 256 threads); owner-computes with the fold where the tile divides the
// grid (SIMOX_SPREAD_OWN=0: the tiled spread over the candidates of each tile).
template <int ORD>
void spread_i32(simox_ctx* ctx, const Geom& g, PmeGrid& pm, const double* src) {
  const double ng = double(g.K[0]) * g.K[1] * g.K[2];
  if (ctx->spread_own && ctx->spread_big) {
    if (double(ctx->n) > ng && SpreadOwn<ORD, 16, 16, 8>::fits(g)) {
      if (spread_own<ORD, 16, 16, 8, 1024>(ctx, g, pm, src)) ctx->spread_own = false;
      else return;
    } else if (double(ctx->n) <= ng && SpreadOwn<ORD, 32, 32, 16>::fits(g)) {
      if (spread_own<ORD, 32, 32, 16, 1024>(ctx, g, pm, src)) ctx->spread_own = false;
      else return;
    }
  }
  if (ctx->spread_big && double(ctx->n) > ng && g.K[0] >= 16 && g.K[1] >= 16 && g.K[2] >= 8) {
    using P = SpreadI32<ORD, 16, 16, 8, 1024>;
    hipLaunchKernelGGL((k_spread_tile_i32<ORD, 16, 16, 8, 1024>), dim3(P::tiles(g)), dim3(1024),
                       0, ctx->st, g, ctx->x, src, pm.bstart, pm.border, pm.brec, pm.greal32);
  } else if (ctx->spread_big && g.K[0] >= 32 && g.K[1] >= 32 && g.K[2] >= 16) {
    using P = SpreadI32<ORD, 32, 32, 16, 1024>;
    hipLaunchKernelGGL((k_spread_tile_i32<ORD, 32, 32, 16, 1024>), dim3(P::tiles(g)), dim3(1024),
                       0, ctx->st, g, ctx->x, src, pm.bstart, pm.border, pm.brec, pm.greal32);
  } else {
    using P = SpreadI32<ORD, 32, 16, 8, 256>;
    hipLaunchKernelGGL((k_spread_tile_i32<ORD, 32, 16, 8, 256>), dim3(P::tiles(g)), dim3(256), 0,
                       ctx->st, g, ctx->x, src, pm.bstart, pm.border, pm.brec, pm.greal32);
  }
}

// Spread charges q (or ql = q + step p when p != null) and convolve, with the
// FFT chain in FP32 (f32) or FP64.  FAITHFUL mode rounds the spreading weights
// to FP32 as the reference does (quirk Q1).
// filter: convolve with the grid's filter table F instead of the influence
// table G.  grid: the PME grid (default the reference's).
int pme_eval(simox_ctx* ctx, const double* q, const double* p, double step, bool f32,
             bool filter = false, PmeGrid* grid = nullptr) {
  PmeGrid& pm = grid ? *grid : ctx->pm;
  const Geom g = geom_of(ctx, pm);
  const int nb = grid_blocks(ctx->n);
  CHK(ensure_bricks(ctx, pm));
  const double* src = q;
  if (p) {
    hipLaunchKernelGGL(k_lincomb, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->n, q, p, step, ctx->ql);
    src = ctx->ql;
  }
  const bool f32w = ctx->mode == SIMOX_MODE_FAITHFUL;
  const int ntile = ((g.K[0] + TX - 1) / TX) * ((g.K[1] + TY - 1) / TY) * ((g.K[2] + TZ - 1) / TZ);
  const bool timed = timed_grid(ctx, pm);
  hipEvent_t ev{};
  if (timed) tstart(ctx, tcat(ctx, pm, 0), &ev);
  if (pm.order == 2 && f32 && pm.spread32)
    spread_i32<2>(ctx, g, pm, src);
  else if (pm.order == 2 && f32)
    hipLaunchKernelGGL((k_spread_tile<2, false, float>), dim3(ntile), dim3(BLOCK), 0, ctx->st, g,
                       ctx->x, src, pm.bstart, pm.border, pm.greal32);
  else if (pm.order == 2)
    hipLaunchKernelGGL((k_spread_tile<2, false, double>), dim3(ntile), dim3(BLOCK), 0, ctx->st, g,
                       ctx->x, src, pm.bstart, pm.border, pm.greal64);
  else if (f32 && pm.spread32)
    spread_i32<4>(ctx, g, pm, src);
  else if (f32)
    hipLaunchKernelGGL((k_spread_tile<4, false, float>), dim3(ntile), dim3(BLOCK), 0, ctx->st, g,
                       ctx->x, src, pm.bstart, pm.border, pm.greal32);
  else if (f32w)
    hipLaunchKernelGGL((k_spread_tile<4, true, double>), dim3(ntile), dim3(BLOCK), 0, ctx->st, g,
                       ctx->x, src, pm.bstart, pm.border, pm.greal64);
  else
    hipLaunchKernelGGL((k_spread_tile<4, false, double>), dim3(ntile), dim3(BLOCK), 0, ctx->st, g,
                       ctx->x, src, pm.bstart, pm.border, pm.greal64);
  if (timed) tstop(ctx, tcat(ctx, pm, 0), ev);
  HIPCHK(hipGetLastError());
  if (f32)
    return pme_chain<float2>(ctx, pm, pm.spec32, pm.tw32, pm.fc32, filter ? pm.F32 : pm.G32,
                             pm.greal32);
  return pme_chain<double2>(ctx, pm, pm.spec64, pm.tw64, pm.fc64, filter ? pm.F64 : pm.G64,
                            pm.greal64);
}

// The plane-resident FP32 chain on a coarse grid when Kx = Ky and both
// kernels exist for its sizes (SIMOX_FFT_PLANE=0 keeps the five-pass chain;
// SIMOX_FFT_PLANE=2 also takes it on the reference's grid, which tests use to
// compare the two chains on the same energies and forces; SIMOX_FFT_NL = 4, 8,
// 11 or 16 sets the z-pass lines per workgroup, default 8; SIMOX_FFT_XYNT =
// 256, 512 or 1024 the threads per workgroup of the x-y passes, default 1024).
bool plane_everywhere() {
  const char* e = std::getenv("SIMOX_FFT_PLANE");
  return e && std::atoi(e) == 2;
}
int plane_setup(simox_ctx* ctx, PmeGrid& pm) {
  FftCfg& c = pm.fc32;
  c.plane = 0;
  const char* e = std::getenv("SIMOX_FFT_PLANE");
  if (e && std::atoi(e) == 0) return 0;
  const int L = pm.K[0], KZ = pm.K[2];
  XyFwd fwd{};
  XyInv inv{};
  c.xynt = 1024;
  if (const char* s = std::getenv("SIMOX_FFT_XYNT")) {
    const int t = std::atoi(s);
    if (t == 256 || t == 512 || t == 1024) c.xynt = t;
  }
  if (pm.K[1] != L || !xy_kernels(L, c.xynt, &fwd, &inv) || !zp_kernel(KZ, 8)) return 0;
  c.nl = 8;
  if (const char* s = std::getenv("SIMOX_FFT_NL")) {
    const int m = std::atoi(s);
    if (zp_kernel(KZ, m)) c.nl = m;
  }
  c.shp = (2 * size_t(plane_buf(L)) + L) * sizeof(float2);
  c.shzp = (2 * size_t(c.nl) * pad_stride(KZ) + KZ) * sizeof(float2);
  if (c.shp > FFT_LDS_MAX || c.shzp > FFT_LDS_MAX) return 0;
  c.zpblocks = (pm.K[1] * pm.H + c.nl - 1) / c.nl;
  const void* fns[3] = {reinterpret_cast<const void*>(fwd), reinterpret_cast<const void*>(inv),
                        reinterpret_cast<const void*>(zp_kernel(KZ, c.nl))};
  for (const void* f : fns)
    HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(FFT_LDS_MAX)));
  CHK(ensure_part(ctx, size_t(c.zpblocks)));
  c.plane = 1;
  return 0;
}

// An FP32 table of grid pm from the half-spectrum layout to the plane-major
// layout of the plane chain (through the grid's FP32 spectrum buffer).
int to_plane_major(simox_ctx* ctx, PmeGrid& pm, float* tab) {
  const Geom g = geom_of(ctx, pm);
  const size_t nh = size_t(g.H) * g.K[1] * g.K[2];
  float* tmp = reinterpret_cast<float*>(pm.spec32);
  hipLaunchKernelGGL(k_to_plane_major<float>, dim3(unsigned((nh + BLOCK - 1) / BLOCK)), dim3(BLOCK),
                     0, ctx->st, g, tab, tmp);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(tab, tmp, nh * sizeof(float), hipMemcpyDeviceToDevice, ctx->st));
  return 0;
}

// Buffers, twiddles, plans and launch shapes of PME grid pm with sizes K (each
// a product of 2, 3, 5 and a multiple of BRK), and its influence table when
// with_G (the reference's grid; the preconditioner's grid has filter tables
// only).  The FP32 chain exists in MIXED mode.
int init_grid(simox_ctx* ctx, PmeGrid& pm, const int K[3], bool with_G, int order) {
  const bool f32 = ctx->mode == SIMOX_MODE_MIXED;
  for (int d = 0; d < 3; ++d) {
    if (K[d] < BRK || K[d] % BRK || !make_plan(K[d], &pm.pl[d])) return SIMOX_ERR_ARG;
    pm.K[d] = K[d];
  }
  pm.H = K[0] / 2 + 1;
  pm.order = order;
  const Geom g = geom_of(ctx, pm);
  const size_t n = size_t(ctx->n);
  const size_t ng = size_t(K[0]) * K[1] * K[2];
  const size_t nh = size_t(pm.H) * K[1] * K[2];
  pm.nbrick = (K[0] / BRK) * (K[1] / BRK) * (K[2] / BRK);
  CHK(dalloc(ctx, &pm.brick, n)); CHK(dalloc(ctx, &pm.bunord, n)); CHK(dalloc(ctx, &pm.border, n));
  CHK(dalloc(ctx, &pm.bcount, pm.nbrick)); CHK(dalloc(ctx, &pm.bfill, pm.nbrick));
  CHK(dalloc(ctx, &pm.bstart, size_t(pm.nbrick) + 1));
  CHK(dalloc(ctx, &pm.bsum, (pm.nbrick + 1023) / 1024 + 1)); CHK(dalloc(ctx, &pm.btotal, 1));
  // FP64 chain always (faithful / FP64 modes, MIXED outer residual); FP32 chain for MIXED
  CHK(dalloc(ctx, &pm.greal64, ng)); CHK(dalloc(ctx, &pm.spec64, nh)); CHK(dalloc(ctx, &pm.G64, nh));
  CHK(dalloc(ctx, &pm.F64, nh));
  if (f32) {
    CHK(dalloc(ctx, &pm.greal32, ng)); CHK(dalloc(ctx, &pm.spec32, nh));
    CHK(dalloc(ctx, &pm.brec, n));
    CHK(dalloc(ctx, &pm.brec_atom, n));
    if (with_G) CHK(dalloc(ctx, &pm.G32, nh));
    CHK(dalloc(ctx, &pm.F32, nh));
  }
  for (int d = 0; d < 3; ++d) {
    const int L = K[d];
    std::vector<double2> t64(L);
    std::vector<float2> t32(L);
    for (int k = 0; k < L; ++k) {
      const double ang = 2.0 * PI * double(k) / double(L);
      t64[k] = make_double2(std::cos(ang), -std::sin(ang));
      t32[k] = make_float2(float(std::cos(ang)), float(-std::sin(ang)));
    }
    CHK(dalloc(ctx, &pm.tw64[d], L)); CHK(dalloc(ctx, &pm.tw32[d], L));
    HIPCHK(hipMemcpy(pm.tw64[d], t64.data(), L * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pm.tw32[d], t32.data(), L * sizeof(float2), hipMemcpyHostToDevice));
  }
  CHK(choose_fft<double2>(ctx, g, &pm.fc64));
  if (f32) CHK(choose_fft<float2>(ctx, g, &pm.fc32));
  if (f32 && (&pm != &ctx->pm || plane_everywhere())) CHK(plane_setup(ctx, pm));
  // G64 is kept for every grid: the preconditioner's filter is built from it
  const double a2 = g.alpha * g.alpha;
  const int nbk = int((nh + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_influence<double>, dim3(nbk), dim3(BLOCK), 0, ctx->st, g, a2, pm.G64, order);
  if (f32 && with_G)
    hipLaunchKernelGGL(k_influence<float>, dim3(nbk), dim3(BLOCK), 0, ctx->st, g, a2, pm.G32, order);
  HIPCHK(hipGetLastError());
  if (f32 && with_G && pm.fc32.plane) CHK(to_plane_major(ctx, pm, pm.G32));
  // the coarse grids (preconditioner, MIXED inner operator) are set up on
  // first use and tuned then; the reference's grid at context creation
  if (&pm != &ctx->pm && autotune_on()) CHK(autotune_grid(ctx, pm));
  pm.bricks_valid = false;
  if (const char* e = std::getenv("SIMOX_GATHER_SHAPE")) {
    const int v = std::atoi(e);
    if (v >= 0 && v < NGATHER_SHAPES) pm.gshape = v;
  }
  pm.ready = true;
  return 0;
}

void free_grid(PmeGrid& pm) {
  void* bufs[] = {pm.brick, pm.bunord, pm.border, pm.bcount, pm.bfill, pm.bstart, pm.bsum,
                  pm.btotal, pm.greal64, pm.greal32, pm.spec64, pm.spec32, pm.G64, pm.G32,
                  pm.F64, pm.F32, pm.brec, pm.brec_atom, pm.sp_ext, pm.sp_xcol, pm.sp_tinv, pm.tw64[0], pm.tw64[1], pm.tw64[2], pm.tw32[0], pm.tw32[1],
                  pm.tw32[2]};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  pm = PmeGrid{};
}

// PME (force.f90:364-464): reciprocal forces added to ctx->f, E_rec.
int do_pme(simox_ctx* ctx, bool fetch_energy) {
  const int n = ctx->n, nb = grid_blocks(n);
  CHK(pme_eval(ctx, ctx->q, nullptr, 0.0, use32(ctx)));
  const int zb = zpass_blocks(ctx, use32(ctx));
  CHK(reduce(ctx, zb, 1, 0, 4));  // energy sum before the part buffer is reused
  CHK(ensure_part(ctx, size_t(nb) * 3));
  const double Ng = double(ctx->g.K[0]) * double(ctx->g.K[1]) * double(ctx->g.K[2]);
  const double cst = KC * Ng / ctx->V / PI;  // force.f90:436
  hipEvent_t ev{};
  tstart(ctx, 5, &ev);
  if (use32(ctx))
    hipLaunchKernelGGL(k_gather_force<float>, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->x,
                       ctx->q, ctx->pm.greal32, cst, ctx->frec, ctx->part);
  else
    hipLaunchKernelGGL(k_gather_force<double>, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->x,
                       ctx->q, ctx->pm.greal64, cst, ctx->frec, ctx->part);
  tstop(ctx, 5, ev);
  HIPCHK(hipGetLastError());
  CHK(reduce(ctx, nb, 3, 0, 5));
  hipLaunchKernelGGL(k_pme_apply, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->frec, ctx->dscal + 5,
                     ctx->f);
  HIPCHK(hipGetLastError());
  if (fetch_energy) {
    double s;
    CHK(fetch(ctx, 4, 1, &s));
    ctx->Erec = KC * Ng * s / ctx->V / PI / 2.0;  // force.f90:435,461
  }
  return 0;
}

// ---------------------------------------------------------------------------
// CTIP
// ---------------------------------------------------------------------------
// FP32 copy of the CTIP matrix (MIXED: the inner matvec on the context's own
// grid and simox_phi; the coarse inner operator reads its packed matrix)
int ensure_val32(simox_ctx* ctx) {
  if (ctx->val32_valid || !use32(ctx)) return 0;
  hipLaunchKernelGGL(k_ell_to_f32, dim3(grid_blocks(ctx->n)), dim3(BLOCK), 0, ctx->st, ctx->n,
                     ctx->cap, ctx->nb_cnt, ctx->nb_val64, ctx->nb_val32);
  HIPCHK(hipGetLastError());
  ctx->val32_valid = true;
  return 0;
}

void launch_ctip_matrix(simox_ctx* ctx) {
  // list entries whose neighbour data each lane loads ahead (SIMOX_CTIP_BATCH);
  // with the packed inner matrix (pk_words set by build_matrix) also each
  // entry's packed word and each row's kept count
  const bool pack = ctx->pk_words != nullptr;
  const auto km = ctx->ctip_batch >= 4
                      ? (pack ? &k_ctip_matrix<double, 4, true> : &k_ctip_matrix<double, 4>)
                  : ctx->ctip_batch == 2
                      ? (pack ? &k_ctip_matrix<double, 2, true> : &k_ctip_matrix<double, 2>)
                      : (pack ? &k_ctip_matrix<double, 1, true> : &k_ctip_matrix<double, 1>);
  hipLaunchKernelGGL(km, dim3(grid_blocks(ctx->n)), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab,
                     ctx->x, ctx->id, ctx->nb_cnt, ctx->nb_col, ctx->nb_val64, ctx->bvec,
                     ctx->inner_drop, ctx->pk_words, ctx->nb_cnt16);
}

int build_matrix(simox_ctx* ctx) {
  const int nb = grid_blocks(ctx->n);
  CHK(ensure_nlist(ctx));
  CHK(check_err(ctx, 1));
  const bool pack = use32(ctx) && ctx->inner_coarse && ctx->inner_f16 && ctx->inner_pack &&
                    ctx->n <= (1 << PK_COL_BITS);
  if (pack) {
    // the CTIP matrix pass also writes the packed words and the kept counts
    if (!ctx->nb_cnt16) CHK(dalloc(ctx, &ctx->nb_cnt16, size_t(ctx->n)));
    if (!ctx->pk_words) CHK(dalloc(ctx, &ctx->pk_words, size_t(ctx->cap) * ctx->n));
  }
  hipEvent_t ev{};
  tstart(ctx, 9, &ev);
  launch_ctip_matrix(ctx);
  tstop(ctx, 9, ev);
  ctx->val32_valid = false;  // the FP32 copy is made where it is read (ensure_val32)
  ctx->val16_valid = false;
  ctx->pk_valid = false;
  if (use32(ctx) && ctx->inner_coarse && ctx->inner_f16) {
    if (!ctx->nb_cnt16) CHK(dalloc(ctx, &ctx->nb_cnt16, size_t(ctx->n)));
    if (pack) {
      const int n = ctx->n, nsl = (n + SLICE - 1) / SLICE, nsb = (nsl + 1023) / 1024;
      if (!ctx->nb_pk) {
        // slices padded to whole groups of 4: at most (cap + 3) entries per row
        CHK(dalloc(ctx, &ctx->nb_pk, size_t(ctx->cap + 4) * (size_t(nsl) * SLICE)));
        CHK(dalloc(ctx, &ctx->pk_slen, size_t(nsl)));
        CHK(dalloc(ctx, &ctx->pk_soff, size_t(nsl) + 1));
        CHK(dalloc(ctx, &ctx->pk_bsum, size_t(nsb) + 1));
        CHK(dalloc(ctx, &ctx->pk_total, 1));
      }
      hipLaunchKernelGGL(k_pk_slices, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->nb_cnt16,
                         ctx->pk_slen);
      hipLaunchKernelGGL(k_scan_block, dim3(nsb), dim3(1024), 0, ctx->st, nsl, ctx->pk_slen,
                         ctx->pk_soff, ctx->pk_bsum);
      hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, ctx->st, nsb, ctx->pk_bsum,
                         ctx->pk_total);
      hipLaunchKernelGGL(k_scan_add, dim3(nsb), dim3(1024), 0, ctx->st, nsl, ctx->pk_soff,
                         ctx->pk_bsum, ctx->pk_total);
      hipLaunchKernelGGL(k_pk_scale, dim3(grid_blocks(nsl + 1)), dim3(BLOCK), 0, ctx->st, nsl,
                         ctx->pk_soff);
      hipLaunchKernelGGL(k_ell_to_sell4, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->nb_cnt,
                         ctx->pk_words, ctx->pk_slen, ctx->pk_soff,
                         reinterpret_cast<uint4*>(ctx->nb_pk));
      ctx->pk_valid = true;
      // the halo copy where the halo variants may run: before the autotuner
      // has timed them, or when it picked one
      ctx->pkh_valid = false;
      if (!ctx->inner_tuned || ctx->mv_variant >= 8) {
        if (!ctx->nb_pkh) {
          CHK(dalloc(ctx, &ctx->nb_pkh, size_t(ctx->cap + 4) * (size_t(nsl) * SLICE)));
          CHK(dalloc(ctx, &ctx->nb_pk24, 3 * (size_t(ctx->cap + 4) / 4) * (size_t(nsl) * SLICE)));
          CHK(dalloc(ctx, &ctx->pk_runs, size_t(nb) * HALO_RUNS));
          CHK(dalloc(ctx, &ctx->pk_hcount, size_t(nb)));
        }
        hipLaunchKernelGGL(k_pk_halo, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->halo_max,
                           ctx->pk_slen,
                           ctx->pk_soff, reinterpret_cast<const uint4*>(ctx->nb_pk),
                           reinterpret_cast<uint4*>(ctx->nb_pkh), ctx->nb_pk24, ctx->pk_runs,
                           ctx->pk_hcount);
        ctx->pkh_valid = true;
      }
    } else {
      if (!ctx->nb_val16) {
        CHK(dalloc(ctx, &ctx->nb_val16, size_t(ctx->cap) * ctx->n));
        CHK(dalloc(ctx, &ctx->nb_col16, size_t(ctx->cap) * ctx->n));
      }
      hipLaunchKernelGGL(k_ell_to_f16, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->n, ctx->cap,
                         ctx->nb_cnt, ctx->nb_col, ctx->nb_val64, ctx->inner_drop, ctx->nb_cnt16,
                         ctx->nb_col16, ctx->nb_val16);
    }
    ctx->val16_valid = true;
  }
  HIPCHK(hipGetLastError());
  ctx->matrix_valid = true;
  return 0;
}

// g = -dU/dq at (q + step*p) or at q; dscal[0..2] = {sum g, -sum g.p, sum g^2}.
int grad_eval(simox_ctx* ctx, const double* q, const double* p, double step, bool want_phi,
              bool f32) {
  const int nb = grid_blocks(ctx->n);
  const double* qeval = q;
  if (p) qeval = ctx->ql;
  CHK(pme_eval(ctx, q, p, step, f32));
  hipEvent_t ev{};
  tstart(ctx, 6, &ev);
  if (f32)
    hipLaunchKernelGGL((k_dudq<float, float>), dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g,
                       ctx->dtab, ctx->x, ctx->id, qeval, ctx->pm.greal32, ctx->c_pme, ctx->nb_cnt,
                       ctx->nb_col, ctx->nb_val32, ctx->bvec, p, ctx->gv,
                       want_phi ? ctx->phi : nullptr, ctx->part);
  else
    hipLaunchKernelGGL((k_dudq<double, double>), dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g,
                       ctx->dtab, ctx->x, ctx->id, qeval, ctx->pm.greal64, ctx->c_pme, ctx->nb_cnt,
                       ctx->nb_col, ctx->nb_val64, ctx->bvec, p, ctx->gv,
                       want_phi ? ctx->phi : nullptr, ctx->part);
  tstop(ctx, 6, ev);
  HIPCHK(hipGetLastError());
  return reduce(ctx, nb, 3, 0, 0);
}

int self_energy(simox_ctx* ctx, const double* q, int use_jp, double wpen, double* out3) {
  const int nb = grid_blocks(ctx->n);
  CHK(ensure_part(ctx, size_t(nb) * 3));
  hipLaunchKernelGGL(k_self_energy, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->n, ctx->dtab, ctx->id,
                     q, use_jp, wpen, ctx->part);
  HIPCHK(hipGetLastError());
  CHK(reduce(ctx, nb, 3, 0, 8));
  return fetch(ctx, 8, 3, out3);
}

// BRENT_SOLVER (ctip.f90:216-304): root of f(a) = -g(q + a p).p on [-1, 1].
int brent(simox_ctx* ctx, double* alpha, int* n_eval) {
  const double E_crit = 1.0e-8, E_end = 1.0e-11;
  double s2[2];
  double a = -1.0, b = 1.0, c, d = 0.0, e = 0.0, fa, fb, fc, xm, x, y, z;
  CHK(grad_eval(ctx, ctx->qc, ctx->pv, a, false, false));
  CHK(fetch(ctx, 0, 2, s2));
  fa = s2[1];
  CHK(grad_eval(ctx, ctx->qc, ctx->pv, b, false, false));
  CHK(fetch(ctx, 0, 2, s2));
  fb = s2[1];
  *n_eval += 2;
  c = b;
  fc = fb;
  for (int it = 1; it <= 100; ++it) {
    if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { c = a; fc = fa; d = b - a; e = d; }
    if (std::fabs(fc) < std::fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    xm = 0.5 * (c - b);
    if (std::fabs(xm) < E_crit || std::fabs(fb) < E_end) { *alpha = b; return 0; }
    if (std::fabs(e) >= E_crit && std::fabs(fa) > std::fabs(fb)) {
      const double s = fb / fa;
      if (a == c) { x = 2.0 * xm * s; y = 1.0 - s; }
      else {
        y = fa / fc; z = fb / fc;
        x = s * (2.0 * xm * y * (y - z) - (b - a) * (z - 1.0));
        y = (y - 1.0) * (z - 1.0) * (s - 1.0);
      }
      if (x > 0.0) y = -y;
      x = std::fabs(x);
      const double m1 = 3.0 * xm * y - std::fabs(E_crit * y), m2 = std::fabs(e * y);
      if (2.0 * x < std::min(m1, m2)) { e = d; d = x / y; }
      else { d = xm; e = d; }
    } else { d = xm; e = d; }
    a = b; fa = fb;
    if (std::fabs(d) > E_crit) b = b + d;
    else b = b + std::copysign(E_crit, xm);
    CHK(grad_eval(ctx, ctx->qc, ctx->pv, b, false, false));
    CHK(fetch(ctx, 0, 2, s2));
    fb = s2[1];
    *n_eval += 1;
  }
  ctx->err = status_text(SIMOX_ERR_BRENT, 1);
  return SIMOX_ERR_BRENT;
}

// CTIP_SOLVER with the reference's nonlinear CG (ctip.f90:167-190).
int ctip_faithful(simox_ctx* ctx, int* iters, int* n_eval) {
  const int n = ctx->n, nb = grid_blocks(n);
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->q, ctx->qc);
  HIPCHK(hipMemsetAsync(ctx->pv, 0, sizeof(double) * n, ctx->st));
  double Error = 1.0, g_sum = 1.0, g_sum_old, alpha = 0.0;
  int k = 0;
  *n_eval = 0;
  while (Error > 1.0e-3) {
    double s2[2], pmax;
    CHK(grad_eval(ctx, ctx->qc, nullptr, 0.0, false, false));
    CHK(fetch(ctx, 0, 2, s2));
    *n_eval += 1;
    g_sum_old = g_sum;
    g_sum = s2[0];
    CHK(ensure_part(ctx, size_t(nb)));
    hipLaunchKernelGGL(k_cg_dir1, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->gv, ctx->pv, g_sum,
                       g_sum_old, ctx->dv, ctx->part);
    CHK(reduce(ctx, nb, 1, 0, 2));
    hipLaunchKernelGGL(k_cg_dir2, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dv, ctx->dscal + 2,
                       ctx->pv, ctx->part);
    CHK(reduce(ctx, nb, 1, 1, 3));
    CHK(fetch(ctx, 3, 1, &pmax));
    CHK(brent(ctx, &alpha, n_eval));
    hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(BLOCK), 0, ctx->st, n, alpha, ctx->pv, ctx->qc);
    HIPCHK(hipGetLastError());
    Error = pmax * std::fabs(alpha);
    ++k;
    if (k >= ctx->maxit && Error > 1.0e-3) {
      // the reference loops without a cap (ctip.f90:169); bound it here
      ctx->err = "CTIP nonlinear CG did not reach max|alpha p| < 1e-3 within the iteration cap";
      return SIMOX_ERR_NOT_CONVERGED;
    }
  }
  *iters = k;
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->qc, ctx->q);
  HIPCHK(hipGetLastError());
  return 0;
}

// Assignment order on the preconditioner's own grid: linear weights (8
// points per atom, an eighth of the spread and gather work of the B-splines),
// or the B-splines with SIMOX_PREC_ORDER=4.
int prec_order(const simox_ctx*) {
  const char* e = std::getenv("SIMOX_PREC_ORDER");
  return (e && std::strcmp(e, "4") == 0) ? 4 : 2;
}

// The preconditioner's grid: half the resolution of the reference's grid in
// every dimension, set up on first use, where those sizes are valid grids
// (multiples of 4 with factors 2, 3, 5); else, or with SIMOX_PREC_FINE=1,
// the reference's grid itself.  The filter acts on long wavelengths, which
// the coarse grid resolves, and its FFTs cost an eighth.
int prec_grid(simox_ctx* ctx, PmeGrid** out) {
  *out = &ctx->pm;
  if (ctx->pc.ready) { *out = &ctx->pc; return 0; }
  const char* e = std::getenv("SIMOX_PREC_FINE");
  if (e && std::strcmp(e, "1") == 0) return 0;
  int K[3];
  FftPlan tmp;
  const int div = ctx->prec_div;
  for (int d = 0; d < 3; ++d) {
    K[d] = ctx->pm.K[d] / div;
    if (ctx->pm.K[d] % div || K[d] < 2 * BRK || K[d] % BRK || !make_plan(K[d], &tmp)) return 0;
  }
  const int rc = init_grid(ctx, ctx->pc, K, false, prec_order(ctx));
  if (rc) {
    free_grid(ctx->pc);
    if (rc == SIMOX_ERR_ARG) return 0;  // no launch shape for these sizes: fine grid
    return rc;
  }
  // MIXED: the FP32 spread of the preconditioner's pass in 32-bit fixed point
  // (stencil records with 10-bit base indices)
  ctx->pc.spread32 = ctx->prec_spread32 && ctx->pc.brec && K[0] <= 1024 && K[1] <= 1024 &&
                     K[2] <= 1024;
  *out = &ctx->pc;
  return 0;
}

// The MIXED inner operator's coarse grid: half the resolution of the
// reference's grid, order-4 B-splines, its own influence table; set up on
// first use where those sizes are valid grids (else the inner operator keeps
// the reference's grid).
int inner_grid(simox_ctx* ctx) {
  if (ctx->pi.ready) return 0;
  int K[3];
  FftPlan tmp;
  const int div = ctx->inner_div;
  for (int d = 0; d < 3; ++d) {
    K[d] = ctx->pm.K[d] / div;
    if (ctx->pm.K[d] % div || K[d] < 2 * BRK || K[d] % BRK || !make_plan(K[d], &tmp)) {
      ctx->inner_coarse = false;
      return 0;
    }
  }
  const int rc = init_grid(ctx, ctx->pi, K, true, 4);
  if (rc) {
    free_grid(ctx->pi);
    ctx->inner_coarse = false;
    return rc == SIMOX_ERR_ARG ? 0 : rc;
  }
  // the stencil records hold 10-bit base indices
  ctx->pi.spread32 = ctx->inner_spread32 && ctx->pi.brec && K[0] <= 1024 && K[1] <= 1024 &&
                     K[2] <= 1024;
  return 0;
}

// Long-range preconditioner set-up for the current positions (once per
// solve) on its grid: rho = 1.03 x the maximum over the grid of the
// low-passed density of 1/J (a spread of w = 1/J through the FFT chain with
// the Gaussian table, sigma = 2 A), then the filter table F = Gc / (1 + Gc rho)
// with Gc = c Ng G, c = kc Ng / (pi V) for that grid's Ng.
int build_prec(simox_ctx* ctx) {
  const int n = ctx->n, nb = grid_blocks(n);
  PmeGrid* pp = nullptr;
  CHK(prec_grid(ctx, &pp));
  PmeGrid& pc = *pp;
  const Geom g = geom_of(ctx, pc);
  const size_t nh = size_t(g.H) * g.K[1] * g.K[2];
  const size_t ng = size_t(g.K[0]) * g.K[1] * g.K[2];
  const int nbk = int((nh + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_prec_weights, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id,
                     ctx->sv, ctx->cv);
  hipLaunchKernelGGL(k_filter_table<double>, dim3(nbk), dim3(BLOCK), 0, ctx->st, g, pc.G64,
                     2.0 * 2.0, 0.0, 0.0, 0, pc.F64);
  HIPCHK(hipGetLastError());
  CHK(pme_eval(ctx, ctx->sv, nullptr, 0.0, false, true, &pc));
  const int nbm = int(std::min<size_t>(1024, (ng + BLOCK - 1) / BLOCK));
  CHK(ensure_part(ctx, size_t(nbm)));
  hipLaunchKernelGGL(k_grid_max<double>, dim3(nbm), dim3(BLOCK), 0, ctx->st, ng, pc.greal64,
                     ctx->part);
  HIPCHK(hipGetLastError());
  CHK(reduce(ctx, nbm, 1, 1, 27));
  double rmax;
  CHK(fetch(ctx, 27, 1, &rmax));
  if (!(rmax > 0.0)) {
    ctx->err = "charge preconditioner: empty density grid";
    return SIMOX_ERR_ARG;
  }
  const double rho = 1.03 * rmax;
  const double c_grid = KC * double(ng) / ctx->V / PI;
  hipLaunchKernelGGL(k_filter_table<double>, dim3(nbk), dim3(BLOCK), 0, ctx->st, g, pc.G64, 0.0,
                     c_grid, rho, 1, pc.F64);
  if (use32(ctx))
    hipLaunchKernelGGL(k_filter_table<float>, dim3(nbk), dim3(BLOCK), 0, ctx->st, g, pc.G64, 0.0,
                       c_grid, rho, 1, pc.F32);
  HIPCHK(hipGetLastError());
  if (use32(ctx) && pc.fc32.plane) CHK(to_plane_major(ctx, pc, pc.F32));
  ctx->prec_rho = rho;
  return 0;
}

// z = P M gv (preconditioned, projected on sum = 0); dscal[19..20] = {z.g, z.g_old};
// ls != null: also the PR+ beta into ls[7] (and rz into ls[6])
int apply_prec(simox_ctx* ctx, const double* g_old, bool f32, double* ls = nullptr) {
  const int n = ctx->n, nb = grid_blocks(n);
  if (!ctx->prec_on) {
    hipLaunchKernelGGL(k_ncg_z, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->gv, g_old, ctx->mv,
                       ctx->dscal + 16, ctx->dscal + 17, ctx->zv, ctx->part);
    hipLaunchKernelGGL(k_reduce_op<3>, dim3(1), dim3(1024), 0, ctx->st, ctx->part, nb, 2,
                       ctx->dscal + 19, ls);
    HIPCHK(hipGetLastError());
    return 0;
  }
  PmeGrid& pc = ctx->pc.ready ? ctx->pc : ctx->pm;  // chosen by build_prec
  const Geom g = geom_of(ctx, pc);
  hipLaunchKernelGGL(k_prec_scale, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id,
                     ctx->qc, ctx->gv, ctx->cv, ctx->sv);
  CHK(pme_eval(ctx, ctx->sv, nullptr, 0.0, f32, true, &pc));
  if (f32 && pc.order == 2)
    hipLaunchKernelGGL((k_prec_finish<2, float>), dim3(nb), dim3(BLOCK), 0, ctx->st, g, ctx->dtab,
                       ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->cv, pc.greal32, ctx->zv, ctx->part);
  else if (pc.order == 2)
    hipLaunchKernelGGL((k_prec_finish<2, double>), dim3(nb), dim3(BLOCK), 0, ctx->st, g, ctx->dtab,
                       ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->cv, pc.greal64, ctx->zv, ctx->part);
  else if (f32)
    hipLaunchKernelGGL((k_prec_finish<4, float>), dim3(nb), dim3(BLOCK), 0, ctx->st, g, ctx->dtab,
                       ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->cv, pc.greal32, ctx->zv, ctx->part);
  else
    hipLaunchKernelGGL((k_prec_finish<4, double>), dim3(nb), dim3(BLOCK), 0, ctx->st, g, ctx->dtab,
                       ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->cv, pc.greal64, ctx->zv, ctx->part);
  HIPCHK(hipGetLastError());
  CHK(reduce(ctx, nb, 1, 0, 26));
  hipLaunchKernelGGL(k_prec_project, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dscal + 26, ctx->gv,
                     g_old, ctx->zv, ctx->part);
  HIPCHK(hipGetLastError());
  CHK(reduce(ctx, nb, 2, 0, 19));
  if (ls) hipLaunchKernelGGL(k_beta, dim3(1), dim3(1), 0, ctx->st, ctx->dscal + 19, ls);
  HIPCHK(hipGetLastError());
  return 0;
}

// Converged charge solve (modes FP64 / MIXED): preconditioned nonlinear CG
// (Polak-Ribiere+) on the piecewise-quadratic charge energy with exact line
// searches, one H0 matvec per iteration (FP64, or FP32 operator in MIXED),
// gradient recurrence in FP64.  Restarts recompute the true gradient with the
// FP64 operator (iterative refinement); the solve ends when the RMS of the
// projected true gradient (eV/e) is below ctx->tol.
//
// The solve is a state machine with one host decision per restart or
// iteration (ncg_begin, then ncg_check after the context's stream has
// drained), so that several contexts can be advanced in turn
// (simox_batch_step) while the GPU works on the others.

// phi = c_pme W^T (reference grid) with the LDS-tiled gather (the brick
// lists are those of the spread that filled the grid).
template <int GX, int GY, int GZ>
void phi_gather_t(simox_ctx* ctx, bool f32, const PmeGrid& pm, double c_pme) {
  using T = GatherTile<GX, GY, GZ>;
  const Geom g = geom_of(ctx, pm);
  const int nt = T::tiles(g);
  if (f32 && pm.spread32)  // the inner grid: weights from the spread's records
    hipLaunchKernelGGL((k_gather_tile_rec<GX, GY, GZ>), dim3(nt), dim3(BLOCK), T::lds(4), ctx->st,
                       g, pm.bstart, pm.border, pm.brec, pm.greal32, c_pme, ctx->phi);
  else if (f32)
    hipLaunchKernelGGL((k_gather_tile<float, GX, GY, GZ>), dim3(nt), dim3(BLOCK), T::lds(4),
                       ctx->st, g, ctx->x, pm.bstart, pm.border, pm.greal32, c_pme, ctx->phi);
  else
    hipLaunchKernelGGL((k_gather_tile<double, GX, GY, GZ>), dim3(nt), dim3(BLOCK), T::lds(8),
                       ctx->st, g, ctx->x, pm.bstart, pm.border, pm.greal64, c_pme, ctx->phi);
}

template <int GX, int GY, int GZ>
int gather_attr(simox_ctx* ctx) {
  const int lds = int(GatherTile<GX, GY, GZ>::lds(8));
  HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gather_tile<double, GX, GY, GZ>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gather_tile<float, GX, GY, GZ>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gather_tile_rec<GX, GY, GZ>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  return 0;
}

// Tile of the potential gather: 64 x 8 x 8 points (measured on the 1M-atom
// slab, grid 256 x 256 x 512, FP32: 128 us; 32 x 16 x 8: 145 us, 32 x 16 x 16:
// 143 us, 64 x 16 x 8: 154 us, 32 x 32 x 8: 192 us; the per-atom gather: 203 us).
constexpr int PGX = 64, PGY = 8, PGZ = 8;

// Tile shapes of the record gather on the MIXED coarse grids (FP32 tiles),
// one of which the autotuner keeps per grid (PmeGrid::gshape; the default is
// shape 0).  Every shape sums each atom's 64 products in the same order.
constexpr int GATHER_SHAPES[NGATHER_SHAPES][3] = {{64, 8, 8},  {32, 8, 8},   {32, 16, 8},
                                                  {64, 16, 8}, {32, 16, 16}, {64, 16, 16}};

template <int GX, int GY, int GZ>
int gather_rec_attr(simox_ctx* ctx) {
  HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gather_tile_rec<GX, GY, GZ>),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             int(GatherTile<GX, GY, GZ>::lds(4))));
  return 0;
}
int gather_rec_attrs(simox_ctx* ctx) {
  CHK((gather_rec_attr<32, 8, 8>(ctx)));
  CHK((gather_rec_attr<32, 16, 8>(ctx)));
  CHK((gather_rec_attr<64, 16, 8>(ctx)));
  CHK((gather_rec_attr<32, 16, 16>(ctx)));
  CHK((gather_rec_attr<64, 16, 16>(ctx)));
  return 0;
}

template <int GX, int GY, int GZ>
void gather_rec_t(simox_ctx* ctx, const PmeGrid& pm, double c_pme) {
  using T = GatherTile<GX, GY, GZ>;
  const Geom g = geom_of(ctx, pm);
  hipLaunchKernelGGL((k_gather_tile_rec<GX, GY, GZ>), dim3(T::tiles(g)), dim3(BLOCK), T::lds(4),
                     ctx->st, g, pm.bstart, pm.border, pm.brec, pm.greal32, c_pme, ctx->phi);
}

int phi_gather(simox_ctx* ctx, bool f32, const PmeGrid* grid = nullptr) {
  const PmeGrid& pm = grid ? *grid : ctx->pm;
  const double c = grid ? KC * double(pm.K[0]) * pm.K[1] * pm.K[2] / ctx->V / PI : ctx->c_pme;
  switch ((f32 && pm.spread32) ? pm.gshape : 0) {
    case 1: gather_rec_t<32, 8, 8>(ctx, pm, c); break;
    case 2: gather_rec_t<32, 16, 8>(ctx, pm, c); break;
    case 3: gather_rec_t<64, 16, 8>(ctx, pm, c); break;
    case 4: gather_rec_t<32, 16, 16>(ctx, pm, c); break;
    case 5: gather_rec_t<64, 16, 16>(ctx, pm, c); break;
    default: phi_gather_t<PGX, PGY, PGZ>(ctx, f32, pm, c); break;
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// v = H0 p of the inner iteration (the sparse part reads the packed SELL-4
// matrix of the coarse inner operator, its FP16 copy, or the FP32 / FP64 ELL
// rows), with the block partials p.v, p.g, sum p^2 pen''(q)
// ea / eb (may be null): events the runtime records at the kernel's own start
// and end (hipExtLaunchKernel), so that they time the kernel alone.
void launch_matvec(simox_ctx* ctx, bool coarse, hipEvent_t ea = nullptr,
                   hipEvent_t eb = nullptr) {
  const int nb = grid_blocks(ctx->n);
  const bool f32 = use32(ctx);
  if (coarse && ctx->val16_valid && ctx->pk_valid) {
    // groups in flight per lane and streaming loads (mv_variant: 1 = 4 groups
    // non-temporal; 0 = 4 groups, 2 = 8, 3 = 8 non-temporal; 4-7 the same
    // with XCD-ordered row blocks; 8 and 9 = 4 and 8 groups non-temporal with
    // the gathers from the LDS halo; 10 = 8 with 24-bit entries, the default
    // (C3: 52.7 us against 57.6 for 8); the autotuner picks one,
    // SIMOX_MV_VARIANT fixes it).  At the 1M-atom slab the non-temporal
    // stream is 4 % faster than 0 and the whole step 4 % (the other kernels
    // keep their L2 lines), the halo 1.3x faster than 1.  All sum in the same
    // order.
    const int v = ctx->mv_variant;
    if (v >= 8 && ctx->pkh_valid) {
      // the LDS-halo variants: gathers from the staged halo, the same sums
      const auto kh = v == 8   ? &k_h0matvec_sell4h<4, true>
                      : v == 9 ? &k_h0matvec_sell4h<8, true>
                               : &k_h0matvec_sell4h<4, true, true>;
      hipExtLaunchKernelGGL(kh, dim3(nb), dim3(BLOCK), 0, ctx->st, ea, eb, 0, ctx->g, ctx->dtab,
                            ctx->id, ctx->qc, ctx->pv, ctx->pv32, ctx->gv, ctx->phi,
                            ctx->pk_slen, ctx->pk_soff,
                            reinterpret_cast<const uint4*>(ctx->nb_pkh), ctx->nb_pk24, ctx->pk_runs,
                            ctx->pk_hcount, ctx->apv, ctx->part);
      return;
    }
    const auto kv = v == 0   ? &k_h0matvec_sell4<4, false>
                    : v == 2 ? &k_h0matvec_sell4<8, false>
                    : v == 3 ? &k_h0matvec_sell4<8, true>
                    : v == 4 ? &k_h0matvec_sell4<4, false, true>
                    : v == 5 ? &k_h0matvec_sell4<4, true, true>
                    : v == 6 ? &k_h0matvec_sell4<8, false, true>
                    : v == 7 ? &k_h0matvec_sell4<8, true, true>
                             : &k_h0matvec_sell4<4, true>;
    hipExtLaunchKernelGGL(kv, dim3(nb), dim3(BLOCK), 0, ctx->st, ea, eb, 0, ctx->g, ctx->dtab,
                       ctx->id, ctx->qc, ctx->pv, ctx->pv32, ctx->gv, ctx->phi, ctx->pk_slen,
                       ctx->pk_soff, reinterpret_cast<const uint4*>(ctx->nb_pk), ctx->apv,
                       ctx->part);
  } else if (coarse && ctx->val16_valid) {
    hipExtLaunchKernelGGL(k_h0matvec16<false>, dim3(nb), dim3(BLOCK), 0, ctx->st, ea, eb, 0, ctx->g, ctx->dtab,
                       ctx->id, ctx->qc, ctx->pv, ctx->pv32, ctx->gv, ctx->phi, ctx->nb_cnt16,
                       ctx->nb_col16, ctx->nb_val16, ctx->nb_pk, ctx->apv, ctx->part);
  } else if (f32) {
    hipExtLaunchKernelGGL((k_h0matvec<float, float>), dim3(nb), dim3(BLOCK), 0, ctx->st, ea, eb, 0, ctx->g,
                       ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->pv, ctx->pv32, ctx->gv, ctx->phi,
                       ctx->nb_cnt, ctx->nb_col, ctx->nb_val32, ctx->apv, ctx->part);
  } else {
    hipExtLaunchKernelGGL((k_h0matvec<double, double>), dim3(nb), dim3(BLOCK), 0, ctx->st, ea, eb, 0, ctx->g,
                       ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->pv, ctx->pv, ctx->gv, ctx->phi,
                       ctx->nb_cnt, ctx->nb_col, ctx->nb_val64, ctx->apv, ctx->part);
  }
}

// The multi-block path with the long-range preconditioner runs the fused
// iteration tail (k_center_scale ... k_pcg_dir_proj).
bool fused_tail(const simox_ctx* ctx) { return !ctx->small_path && ctx->prec_on; }

// One inner iteration: H0 p, exact line search, step, new direction.
int ncg_iter_body(simox_ctx* ctx) {
  const int n = ctx->n, nb = grid_blocks(n);
  const int nbr = std::min(nb, RED_GRID);  // the fused-reduction kernels' grid
  const int nbw = std::min(grid_blocks(n, RED_NT), RED_GRID);  // the same, RED_NT threads each (the step)
  const bool f32 = use32(ctx);
  float* p32 = f32 ? ctx->pv32 : nullptr;
  double* ls = ctx->dscal + 40;
  // v = H0 p with p.v, p.g, sum p^2 pen''(q)
  const PmeGrid* ig = (f32 && ctx->inner_coarse && ctx->pi.ready) ? &ctx->pi : nullptr;
  CHK(pme_eval(ctx, ctx->pv, nullptr, 0.0, f32, false, ig ? &ctx->pi : nullptr));
  hipEvent_t ev{};
  tstart(ctx, 10, &ev);
  CHK(phi_gather(ctx, f32, ig));
  tstop(ctx, 10, ev);
  {
    hipEvent_t ka{}, kb{};
    tkernel(ctx, 6, &ka, &kb);
    launch_matvec(ctx, ig != nullptr, ka, kb);
  }
  // (the matvec's 4,096 workgroups summing in their last one, as the
  // grid-stride kernels below do, took 88 us instead of 60 at the 1M-atom
  // slab: they serialise on the completion counter)
  hipLaunchKernelGGL(k_reduce_op<1>, dim3(1), dim3(1024), 0, ctx->st, ctx->part, nb, 3,
                     ctx->dscal + 22, ls);
  // exact line search, Newton with bracketing, on the device: one workgroup
  // for small systems, else two launches per round (rounds after
  // convergence return at once)
  if (ctx->small_path) {
    // small systems: line search, step, centring and the preconditioner's
    // per-atom passes each in one workgroup (launch latency bound)
    hipLaunchKernelGGL(k_ncg_ls_small, dim3(1), dim3(1024), 0, ctx->st, n, ctx->dtab, ctx->id,
                       ctx->qc, ctx->pv, ls, ctx->dscal + 24);
    if (!ctx->prec_on) {
      hipLaunchKernelGGL(k_small_step<true>, dim3(1), dim3(1024), 0, ctx->st, n, ctx->dtab, ctx->id,
                         ls, ctx->apv, nullptr, ctx->qc, ctx->gv, ctx->rv, ctx->mv, ctx->zv,
                         ctx->pv, p32, nullptr, ctx->dscal);
      HIPCHK(hipGetLastError());
      return 0;
    }
    hipLaunchKernelGGL(k_small_step<false>, dim3(1), dim3(1024), 0, ctx->st, n, ctx->dtab, ctx->id,
                       ls, ctx->apv, ctx->cv, ctx->qc, ctx->gv, ctx->rv, ctx->mv, nullptr, ctx->pv,
                       nullptr, ctx->sv, ctx->dscal);
    HIPCHK(hipGetLastError());
    PmeGrid& pc = ctx->pc.ready ? ctx->pc : ctx->pm;  // chosen by build_prec
    CHK(pme_eval(ctx, ctx->sv, nullptr, 0.0, f32, true, &pc));
    const Geom gp = geom_of(ctx, pc);
    if (f32 && pc.order == 2)
      hipLaunchKernelGGL((k_small_prec_tail<2, float>), dim3(1), dim3(1024), 0, ctx->st, gp,
                         ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->rv, ctx->cv,
                         pc.greal32, ls, ctx->zv, ctx->pv, p32, ctx->dscal);
    else if (pc.order == 2)
      hipLaunchKernelGGL((k_small_prec_tail<2, double>), dim3(1), dim3(1024), 0, ctx->st, gp,
                         ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->rv, ctx->cv,
                         pc.greal64, ls, ctx->zv, ctx->pv, p32, ctx->dscal);
    else if (f32)
      hipLaunchKernelGGL((k_small_prec_tail<4, float>), dim3(1), dim3(1024), 0, ctx->st, gp,
                         ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->rv, ctx->cv,
                         pc.greal32, ls, ctx->zv, ctx->pv, p32, ctx->dscal);
    else
      hipLaunchKernelGGL((k_small_prec_tail<4, double>), dim3(1), dim3(1024), 0, ctx->st, gp,
                         ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->rv, ctx->cv,
                         pc.greal64, ls, ctx->zv, ctx->pv, p32, ctx->dscal);
    HIPCHK(hipGetLastError());
    return 0;
  } else {
    // rounds launched: one more than the last iteration's line search ran
    // (8 after one that did not finish, at a solve's start, and in graphs)
    const int rounds = (ctx->ls_adapt && !ctx->use_graph) ? ctx->ls_rounds : 8;
    // each round one launch: its last workgroup sums the rows and updates
    // the step (RedTail)
    for (int k = 0; k < rounds; ++k)
      hipLaunchKernelGGL(k_ncg_ls_dev, dim3(nbr), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id,
                         ctx->qc, ctx->pv, ls, ctx->part,
                         RedTail{ctx->red_ctr, ctx->dscal + 64, ls, 2});
  }
  hipLaunchKernelGGL(k_ncg_step_dev, dim3(nbw), dim3(RED_NT), 0, ctx->st, n, ctx->dtab, ctx->id,
                     ls, ctx->pv, ctx->apv, ctx->qc, ctx->gv, ctx->rv, ctx->part,
                     RedTail{ctx->red_ctr, ctx->dscal + 0, nullptr, 0});
  if (fused_tail(ctx)) {
    // centre + scale, the preconditioner's grid pass, z with its six sums,
    // projected beta, direction (ncg_issue_iter reads sum g^2 from ls[10])
    PmeGrid& pc = ctx->pc.ready ? ctx->pc : ctx->pm;  // chosen by build_prec
    const Geom gp = geom_of(ctx, pc);
    hipLaunchKernelGGL(k_center_scale, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id,
                       ctx->qc, ctx->dscal + 0, ctx->gv, ctx->cv, ctx->sv);
    CHK(pme_eval(ctx, ctx->sv, nullptr, 0.0, f32, true, &pc));
    // z, its six sums and (last workgroup) the projected beta in one launch
    const RedTail rt{ctx->red_ctr, ctx->dscal + 56, ls, 4};
    if (f32 && pc.order == 2 && pc.spread32)
      hipLaunchKernelGGL(k_prec_finish6_rec, dim3(nbr), dim3(BLOCK), 0, ctx->st, gp, ctx->dtab,
                         pc.brec_atom, ctx->id, ctx->qc, ctx->gv, ctx->rv, ctx->cv, pc.greal32,
                         ctx->zv, ctx->part, rt);
    else if (f32 && pc.order == 2)
      hipLaunchKernelGGL((k_prec_finish6<2, float>), dim3(nbr), dim3(BLOCK), 0, ctx->st, gp,
                         ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->rv, ctx->cv,
                         pc.greal32, ctx->zv, ctx->part, rt);
    else if (pc.order == 2)
      hipLaunchKernelGGL((k_prec_finish6<2, double>), dim3(nbr), dim3(BLOCK), 0, ctx->st, gp,
                         ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->rv, ctx->cv,
                         pc.greal64, ctx->zv, ctx->part, rt);
    else if (f32)
      hipLaunchKernelGGL((k_prec_finish6<4, float>), dim3(nbr), dim3(BLOCK), 0, ctx->st, gp,
                         ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->rv, ctx->cv,
                         pc.greal32, ctx->zv, ctx->part, rt);
    else
      hipLaunchKernelGGL((k_prec_finish6<4, double>), dim3(nbr), dim3(BLOCK), 0, ctx->st, gp,
                         ctx->dtab, ctx->x, ctx->id, ctx->qc, ctx->gv, ctx->rv, ctx->cv,
                         pc.greal64, ctx->zv, ctx->part, rt);
    hipLaunchKernelGGL(k_pcg_dir_proj, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ls, ctx->zv, ctx->pv,
                       p32);
    HIPCHK(hipGetLastError());
    return 0;
  }
  hipLaunchKernelGGL(k_ncg_center, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id,
                     ctx->qc, ctx->dscal + 0, ctx->gv, ctx->mv, ctx->part);
  CHK(reduce(ctx, nb, 3, 0, 16));
  CHK(apply_prec(ctx, ctx->rv, f32, ls));
  hipLaunchKernelGGL(k_pcg_dir_dev, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ls, ctx->zv, ctx->pv,
                     p32);
  HIPCHK(hipGetLastError());
  return 0;
}

// Launches one inner iteration: replays the captured graph of the body when
// ctx->use_graph (batched stepping), else enqueues the kernels.
int ncg_launch_iter(simox_ctx* ctx) {
  if (!ctx->use_graph || ctx->timing) return ncg_iter_body(ctx);
  if (!ctx->iter_graph || ctx->iter_graph_gen != ctx->part_gen) {
    if (ctx->iter_graph) HIPCHK(hipGraphExecDestroy(ctx->iter_graph));
    ctx->iter_graph = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
    const int rc = ncg_iter_body(ctx);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->st, &graph);
    if (rc) return rc;
    HIPCHK(e);
    HIPCHK(hipGraphInstantiate(&ctx->iter_graph, graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(graph));
    ctx->iter_graph_gen = ctx->part_gen;
  }
  HIPCHK(hipGraphLaunch(ctx->iter_graph, ctx->st));
  return 0;
}

// Issues r.batch inner iterations, then the scalars of the last one for
// ncg_check (one host round trip per batch).
int ncg_issue_iter(simox_ctx* ctx) {
  for (int k = 0; k < ctx->ncg.batch; ++k) CHK(ncg_launch_iter(ctx));
  if (fused_tail(ctx)) {
    // ls[5..10]: done flag, rz, beta, descent flag, Newton rounds run, sum g^2
    HIPCHK(hipMemcpyAsync(ctx->hscal + 45, ctx->dscal + 45, 6 * sizeof(double),
                          hipMemcpyDeviceToHost, ctx->st));
  } else {
    // the convergence and descent scalars of this iteration, for ncg_check
    HIPCHK(hipMemcpyAsync(ctx->hscal + 18, ctx->dscal + 18, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->st));
    // ls[5..9]: done flag, rz, beta, descent flag, Newton rounds run
    HIPCHK(hipMemcpyAsync(ctx->hscal + 45, ctx->dscal + 45, 5 * sizeof(double),
                          hipMemcpyDeviceToHost, ctx->st));
  }
  ctx->ncg.phase = NCG_WAIT_ITER;
  return 0;
}

// Inner iterations to issue before the next check: the iterations the
// solve's own rate (or the last solve's, at its start) predicts to reach the
// target, less 15 %, within [1, max_batch].  Issuing more iterations than
// needed only continues the CG past its target; the host checks the last.
int ncg_batch(const simox_ctx* ctx) {
  const NcgRun& r = ctx->ncg;
  if (ctx->max_batch <= 1) return 1;
  double rate = ctx->last_rate;
  if (r.it > 0 && r.rms < r.rms0 && r.rms0 > 0.0) rate = std::pow(r.rms / r.rms0, 1.0 / r.it);
  if (!(rate > 0.0 && rate < 1.0) || !(r.rms > r.target)) return r.it > 0 ? 1 : 4;
  const double need = std::log(r.target / r.rms) / std::log(rate);
  const int b = int(0.85 * need);
  return std::max(1, std::min(std::min(ctx->max_batch, b), ctx->maxit - r.it));
}

// true gradient (FP64 operator) and its projected RMS
int ncg_issue_restart(simox_ctx* ctx) {
  const int n = ctx->n, nb = grid_blocks(n);
  CHK(grad_eval(ctx, ctx->qc, nullptr, 0.0, false, false));
  ++ctx->ncg.nev;
  hipLaunchKernelGGL(k_ncg_center, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id,
                     ctx->qc, ctx->dscal + 0, ctx->gv, ctx->mv, ctx->part);
  CHK(reduce(ctx, nb, 3, 0, 16));  // {sum m g, sum m, sum g^2} at 16..18
  CHK(apply_prec(ctx, nullptr, use32(ctx)));  // {z.g, -} at 19..20
  HIPCHK(hipMemcpyAsync(ctx->hscal + 16, ctx->dscal + 16, 4 * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->st));
  ctx->ncg.phase = NCG_WAIT_RESTART;
  return 0;
}

// Plane-chain shapes of grid pm: threads per workgroup of the x-y passes
// (xynt) and z lines per workgroup (nl); false where no kernel exists or the
// LDS does not hold them.
bool plane_shape(simox_ctx* ctx, PmeGrid& pm, int xynt, int nl) {
  FftCfg& c = pm.fc32;
  XyFwd fwd{};
  XyInv inv{};
  const int L = pm.K[0], KZ = pm.K[2];
  if (!xy_kernels(L, xynt, &fwd, &inv) || !zp_kernel(KZ, nl)) return false;
  const size_t shzp = (2 * size_t(nl) * pad_stride(KZ) + KZ) * sizeof(float2);
  if (shzp > FFT_LDS_MAX) return false;
  const void* fns[3] = {reinterpret_cast<const void*>(fwd), reinterpret_cast<const void*>(inv),
                        reinterpret_cast<const void*>(zp_kernel(KZ, nl))};
  for (const void* f : fns)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(FFT_LDS_MAX)) !=
        hipSuccess)
      return false;
  const int zpb = (pm.K[1] * pm.H + nl - 1) / nl;
  if (ensure_part(ctx, size_t(zpb))) return false;
  c.xynt = xynt;
  c.nl = nl;
  c.shzp = shzp;
  c.zpblocks = zpb;
  return true;
}

// Autotuner of the inner solve's launch shapes, on this context's buffers at
// its first solve with the coarse inner operator (SIMOX_AUTOTUNE=0 keeps the
// defaults; SIMOX_MV_VARIANT, SIMOX_GATHER_SHAPE, SIMOX_FFT_XYNT and
// SIMOX_FFT_NL fix one each): the packed matvec's variant, the record
// gather's tile (GATHER_SHAPES) and, on both coarse grids, the plane FFT
// chain's x-y threads per workgroup and z lines per workgroup.  Each
// candidate is timed over 3 launches after a warm one.  Every candidate
// computes the same values in the same order (the gather sums each atom's
// products in one order, the matvec each row's, the FFT passes each line's;
// the z pass's partial sums, which its lines per workgroup group, are not
// read on the coarse grids), so the choice changes times only.  It runs
// before the solve's first restart, so the buffers it writes (coarse grids,
// potential, matvec output, block partials) are all rewritten before use.
// The fastest of n candidates: set(k) selects candidate k (false: it does
// not exist here), launch() runs the selected one.  Every candidate is timed
// twice, in two passes over all of them, and keeps its faster time, so that
// neither a clock still ramping up during the first pass nor one slow
// launch decides the pick (single passes of 3 launches picked shapes up to
// 30 % slower on the 64 x 64 x 128 grid).  *pick = -1 when none exists.
template <class S, class L>
int fastest(simox_ctx* ctx, int n, S set, L launch, int* pick) {
  std::vector<float> t(size_t(n), 1e30f);
  for (int pass = 0; pass < 2; ++pass)
    for (int k = 0; k < n; ++k) {
      if (!set(k)) continue;
      float ms = 0.f;
      CHK(time_launch(ctx, launch, &ms));
      t[size_t(k)] = std::min(t[size_t(k)], ms);
    }
  *pick = -1;
  for (int k = 0; k < n; ++k)
    if (t[size_t(k)] < 1e29f && (*pick < 0 || t[size_t(k)] < t[size_t(*pick)])) *pick = k;
  return 0;
}

int tune_inner(simox_ctx* ctx) {
  ctx->inner_tuned = true;
  PmeGrid& pi = ctx->pi;
  if (!autotune_on() || !(use32(ctx) && ctx->inner_coarse && pi.ready && pi.spread32 &&
                          pi.bricks_valid))
    return 0;
  int pick = -1;
  if (!ctx->mv_fixed && ctx->val16_valid && ctx->pk_valid) {
    const int v0 = ctx->mv_variant;
    CHK(fastest(ctx, NMV_VARIANTS, [&](int v) { ctx->mv_variant = v; return true; },
                [&] { launch_matvec(ctx, true); }, &pick));
    ctx->mv_variant = pick >= 0 ? pick : v0;
  }
  if (!ctx->gather_fixed) {
    const int g0 = pi.gshape;
    CHK(fastest(ctx, NGATHER_SHAPES, [&](int sh) { pi.gshape = sh; return true; },
                [&] { (void)phi_gather(ctx, true, &pi); }, &pick));
    pi.gshape = pick >= 0 ? pick : g0;
  }
  for (PmeGrid* pm : {&ctx->pi, &ctx->pc}) {
    FftCfg& c = pm->fc32;
    if (!pm->ready || !c.plane) continue;
    const float* tab = pm == &ctx->pc ? pm->F32 : pm->G32;
    auto chain = [&] { (void)plane_chain(ctx, *pm, pm->spec32, pm->tw32, c, tab, pm->greal32); };
    if (!ctx->xynt_fixed) {
      static const int NT[3] = {256, 512, 1024};
      const int t0 = c.xynt, nl0 = c.nl;
      CHK(fastest(ctx, 3, [&](int k) { return plane_shape(ctx, *pm, NT[k], nl0); }, chain, &pick));
      if (!plane_shape(ctx, *pm, pick >= 0 ? NT[pick] : t0, nl0)) return SIMOX_ERR_ARG;
    }
    if (!ctx->nl_fixed) {
      static const int NL[4] = {4, 8, 11, 16};
      const int t0 = c.xynt, nl0 = c.nl;
      CHK(fastest(ctx, 4, [&](int k) { return plane_shape(ctx, *pm, t0, NL[k]); }, chain, &pick));
      if (!plane_shape(ctx, *pm, t0, pick >= 0 ? NL[pick] : nl0)) return SIMOX_ERR_ARG;
    }
  }
  return 0;
}

int ncg_begin(simox_ctx* ctx) {
  const int n = ctx->n, nb = grid_blocks(n);
  {
    // polynomial extrapolation in time through q(t), q(t-1), ... of order
    // min(pred_order, nhist); coefficients sum to 1, so sum(q) is kept
    static const double coef[4][4] = {
        {1.0, 0.0, 0.0, 0.0}, {2.0, -1.0, 0.0, 0.0}, {3.0, -3.0, 1.0, 0.0}, {4.0, -6.0, 4.0, -1.0}};
    const int ord = std::min(ctx->pred_order, ctx->nhist);
    const double* c = coef[ord];
    hipLaunchKernelGGL(k_extrapolate, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ord, c[0], c[1], c[2],
                       c[3], ctx->q, ctx->qhist[0], ctx->qhist[1], ctx->qhist[2], ctx->qc);
  }
  HIPCHK(hipGetLastError());
  ctx->ncg = NcgRun{};
  ctx->ls_rounds = 8;
  // the fused tail's six sums per block, and ls[12] = n for its mean (set
  // before any capture of the iteration body)
  CHK(ensure_part(ctx, size_t(nb) * 6));
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, ctx->st, 1, double(n), ctx->dscal + 52);
  // the cell sort swaps the atom buffers (x, id, cell, ...): a graph captured
  // in an earlier solve may name the other buffer of each pair, so capture anew
  ctx->iter_graph_gen = -1;
  if (ctx->prec_on) CHK(build_prec(ctx));
  if (use32(ctx) && ctx->inner_coarse) CHK(inner_grid(ctx));
  // brick lists of the inner grid for these positions, here rather than in
  // the first iteration (which a batched solve captures into its graph)
  if (use32(ctx) && ctx->inner_coarse && ctx->pi.ready) CHK(ensure_bricks(ctx, ctx->pi));
  if (use32(ctx) && !(ctx->inner_coarse && ctx->pi.ready && ctx->val16_valid)) CHK(ensure_val32(ctx));
  if (!ctx->inner_tuned) CHK(tune_inner(ctx));
  return ncg_issue_restart(ctx);
}

// The accepted charges become the state; shift the predictor history.
int ncg_finish(simox_ctx* ctx) {
  const int n = ctx->n, nb = grid_blocks(n);
  std::swap(ctx->qhist[2], ctx->qhist[1]);
  std::swap(ctx->qhist[1], ctx->qhist[0]);
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->q, ctx->qhist[0]);
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->qc, ctx->q);
  HIPCHK(hipGetLastError());
  ctx->nhist = std::min(ctx->nhist + 1, 3);
  return 0;
}

// Host decision after the context's stream has drained: read the scalars of
// the restart or iteration just run and issue the next piece of work.
int ncg_check(simox_ctx* ctx) {
  NcgRun& r = ctx->ncg;
  const int n = ctx->n, nb = grid_blocks(n);
  if (r.phase == NCG_WAIT_RESTART) {
    const double rms_true = std::sqrt(ctx->hscal[18] / n);
    if (rms_true <= ctx->tol) {
      r.ok = true;
      r.phase = NCG_DONE;
      return ncg_finish(ctx);
    }
    if (r.restarts >= 60) { r.phase = NCG_DONE; return 0; }
    ++r.restarts;
    const double rt = (ctx->inner_coarse && ctx->pi.ready) ? ctx->inner_rtol_c : ctx->inner_rtol;
    r.target = use32(ctx) ? std::max(0.3 * ctx->tol, rt * rms_true) : 0.3 * ctx->tol;
    float* p32 = use32(ctx) ? ctx->pv32 : nullptr;
    hipLaunchKernelGGL(k_set_dir, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->zv, ctx->pv, p32);
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, ctx->st, 1, ctx->hscal[19],
                       ctx->dscal + 46);  // ls[6] = rz
    HIPCHK(hipGetLastError());
    r.it = 0;
    r.rms = rms_true;
    r.rms0 = rms_true;
    if (r.rms > r.target && r.it < ctx->maxit) {
      r.batch = ncg_batch(ctx);
      return ncg_issue_iter(ctx);
    }
    return ncg_issue_restart(ctx);
  }
  if (r.phase == NCG_WAIT_ITER) {
    r.it += r.batch;
    const bool ls_done = ctx->hscal[45] != 0.0;
    ctx->ls_rounds = ls_done ? std::min(8, std::max(1 + ctx->ls_spare,
                                                     int(ctx->hscal[49]) + ctx->ls_spare))
                             : 8;
    r.rms = std::sqrt((fused_tail(ctx) ? ctx->hscal[50] : ctx->hscal[18]) / n);
    const bool descent = ctx->hscal[48] == 0.0;  // else restart from the true gradient
    if (descent && r.rms > r.target && r.it < ctx->maxit) {
      r.batch = ncg_batch(ctx);
      return ncg_issue_iter(ctx);
    }
    if (r.it > 0 && r.rms < r.rms0 && r.rms0 > 0.0)
      ctx->last_rate = std::pow(r.rms / r.rms0, 1.0 / r.it);
    r.total_it += r.it;
    r.nev += r.it;
    return ncg_issue_restart(ctx);
  }
  return 0;
}

int ncg_end(simox_ctx* ctx, int* iters, int* n_eval) {
  ctx->last_inner = ctx->ncg.total_it;
  ctx->last_outer = ctx->ncg.restarts;
  if (iters) *iters = ctx->ncg.total_it;
  if (n_eval) *n_eval = ctx->ncg.nev;
  ctx->ncg.phase = NCG_IDLE;
  if (!ctx->ncg.ok) {
    ctx->err = "charge solve did not reach the requested tolerance";
    return SIMOX_ERR_NOT_CONVERGED;
  }
  return 0;
}

int ctip_ncg(simox_ctx* ctx, int* iters, int* n_eval) {
  CHK(ncg_begin(ctx));
  while (ctx->ncg.phase != NCG_DONE) {
    HIPCHK(hipStreamSynchronize(ctx->st));
    CHK(ncg_check(ctx));
  }
  return ncg_end(ctx, iters, n_eval);
}

int do_ctip(simox_ctx* ctx, int* iters, int* n_eval) {
  CHK(build_matrix(ctx));
  int it = 0, ne = 0;
  if (ctx->mode == SIMOX_MODE_FAITHFUL) CHK(ctip_faithful(ctx, &it, &ne));
  else CHK(ctip_ncg(ctx, &it, &ne));
  if (iters) *iters = it;
  if (n_eval) *n_eval = ne;
  double e3[3];
  CHK(self_energy(ctx, ctx->q, 1, WPEN, e3));  // ctip.f90:194-209
  ctx->Eself = e3[0] * KC * ctx->g.alpha / SQRTPI + (e3[1] + e3[2]);
  return 0;
}

int do_vv1(simox_ctx* ctx, double dt) {
  if (ctx->reflect)  // VV_NVE1_REFLECT
    hipLaunchKernelGGL(k_vv1_reflect, dim3(grid_blocks(ctx->n)), dim3(BLOCK), 0, ctx->st, ctx->n,
                       ctx->dtab, ctx->id, ctx->g.box[0], ctx->g.box[1], ctx->g.box[2], dt, ctx->x,
                       ctx->v, ctx->f);
  else if (ctx->thermo == 2)  // VV_NVT1_STOCH
    hipLaunchKernelGGL(k_vv1_stoch, dim3(grid_blocks(ctx->n)), dim3(BLOCK), 0, ctx->st, ctx->n,
                       ctx->dtab, ctx->id, ctx->g.box[0], ctx->g.box[1], ctx->g.box[2], dt,
                       ctx->alpha, ctx->x, ctx->v, ctx->f);
  else  // VV_NVE1
    hipLaunchKernelGGL(k_vv1, dim3(grid_blocks(ctx->n)), dim3(BLOCK), 0, ctx->st, ctx->n, ctx->dtab,
                       ctx->id, ctx->g.box[0], ctx->g.box[1], ctx->g.box[2], dt, ctx->x, ctx->v,
                       ctx->f);
  HIPCHK(hipGetLastError());
  ctx->matrix_valid = false;
  ctx->nlist_valid = false;
  ctx->pm.bricks_valid = ctx->pc.bricks_valid = ctx->pi.bricks_valid = false;
  return 0;
}

int do_vv2(simox_ctx* ctx, double dt, bool fetch_mv2) {
  const int nb = grid_blocks(ctx->n);
  CHK(ensure_part(ctx, size_t(nb) * 10));
  if (ctx->thermo == 1) {  // VV_NVT2_BEREND
    hipLaunchKernelGGL(k_vv2_berend1, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->n, ctx->dtab, ctx->id,
                       dt, ctx->v, ctx->f, ctx->part);
    CHK(reduce(ctx, nb, 10, 0, 50));
    hipLaunchKernelGGL(k_vv2_berend2, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->n, ctx->dtab, ctx->id,
                       ctx->dscal + 50, ctx->t_given, ctx->v, ctx->part);
  } else if (ctx->thermo == 2) {  // VV_NVT2_STOCH
    const float a32 = float(ctx->alpha);
    const float b32 = float(std::sqrt(2.0 * double(a32) * ctx->t_given / dt));
    hipLaunchKernelGGL(k_vv2_stoch, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->n, ctx->dtab, ctx->id,
                       ctx->perm, dt, a32, b32, ctx->seed, ctx->nvt_step, ctx->v, ctx->f, ctx->part);
    ++ctx->nvt_step;
  } else {  // VV_NVE2
    hipLaunchKernelGGL(k_vv2, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->n, ctx->dtab, ctx->id, dt,
                       ctx->v, ctx->f, ctx->part);
  }
  HIPCHK(hipGetLastError());
  CHK(reduce(ctx, nb, 1, 0, 12));
  if (fetch_mv2) CHK(fetch(ctx, 12, 1, &ctx->mv2));
  return 0;
}

int upload_caller(simox_ctx* ctx, const double* h, int w, double* dsorted) {
  const size_t cnt = size_t(ctx->n) * w;
  CHK(ensure_stage(ctx, cnt));
  memcpy(ctx->hstage, h, cnt * sizeof(double));
  HIPCHK(hipMemcpyAsync(ctx->dstage, ctx->hstage, cnt * sizeof(double), hipMemcpyHostToDevice,
                        ctx->st));
  hipLaunchKernelGGL(k_from_caller, dim3(grid_blocks(ctx->n)), dim3(BLOCK), 0, ctx->st, ctx->n, w,
                     ctx->perm, ctx->dstage, dsorted);
  HIPCHK(hipGetLastError());
  // the pinned and device staging buffers are reused by the next upload
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

int download_caller(simox_ctx* ctx, const double* dsorted, int w, double* h) {
  const size_t cnt = size_t(ctx->n) * w;
  CHK(ensure_stage(ctx, cnt));
  hipLaunchKernelGGL(k_to_caller, dim3(grid_blocks(ctx->n)), dim3(BLOCK), 0, ctx->st, ctx->n, w,
                     ctx->perm, dsorted, ctx->dstage);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ctx->hstage, ctx->dstage, cnt * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  memcpy(h, ctx->hstage, cnt * sizeof(double));
  return 0;
}

bool need_state(simox_ctx* ctx) {
  if (!ctx->have_state) {
    ctx->err = "state not set (simox_set_state with species ids first)";
    return false;
  }
  return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of simox_cell_sort_noalloc: per-atom cell index and slot, per-cell
// fill counter, error flag.  One copy per GPU, held until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address comes from hipGetSymbolAddress (noalloc_scratch) and only
// that address is passed to kernels.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_MAX_ATOMS = 1 << 21;  // 2,097,152 atoms
constexpr int NOALLOC_MAX_CELLS = 1 << 17;  // 131,072 cells
constexpr size_t NOALLOC_WORDS = 2 * size_t(NOALLOC_MAX_ATOMS) + NOALLOC_MAX_CELLS + 64;
static __device__ __align__(256) int scratch[NOALLOC_WORDS];

namespace {
constexpr int NOALLOC_MAX_DEV = 64;
std::mutex noalloc_mu;
int* noalloc_ptr[NOALLOC_MAX_DEV] = {};

// device address of `scratch` on the current device (looked up once per
// device, with that device current)
int noalloc_scratch(int** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEV) return SIMOX_ERR_HIP;
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    int* p = nullptr;
    if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return SIMOX_ERR_HIP;
    noalloc_ptr[dev] = p;
  }
  *out = noalloc_ptr[dev];
  return SIMOX_OK;
}
}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

void simox_fafb(simox_params* p) {
  // datafmt.f90:201-220: [fa|fb] coefficients from the Slater exponents xi
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      double* c = p->prm[i].fafb[j];
      const double a = p->prm[i].xi, b = p->prm[j].xi;
      if (i == j) {
        c[0] = 1.0;
        c[1] = 11.0 * a / 8.0;
        c[2] = 0.75 * (a * a);
        c[3] = (a * a * a) / 6.0;
      } else {
        const double s = a + b, da = a - b, db = b - a;
        c[0] = a * (b * b * b * b) / (s * s) / (da * da);
        c[1] = b * (a * a * a * a) / (s * s) / (db * db);
        c[2] = (3.0 * (a * a) * (b * b * b * b) - (b * b * b * b * b * b)) / (s * s * s) / (da * da * da);
        c[3] = (3.0 * (b * b) * (a * a * a * a) - (a * a * a * a * a * a)) / (s * s * s) / (db * db * db);
      }
    }
}

void simox_potential_setup(simox_params* p) {
  memset(p, 0, sizeof *p);
  // Zhou-Johnson-Wadley EAM for Al, Ni, Fe (datafmt.f90:155-178)
  static const double metal[3][22] = {
      {2.86392, 1.20378, 17.51747, 19.90041, 6.61317, 3.52702, 0.31487, 0.36555, 0.37985,
       0.75969, -2.80760, -0.30144, 1.25856, -1.24760, -2.83000, 0.0, 0.62225, -2.48824,
       -2.83000, 0.0, 0.62225, -2.48824},
      {2.48875, 2.21149, 30.37003, 30.37137, 8.38345, 4.47117, 0.42905, 0.63353, 0.44360,
       0.82066, -2.69351, -0.07644, 0.24144, -2.37563, -2.70000, 0.0, 0.26539, -0.15286,
       -2.70000, 0.0, 0.26539, 4.58568},
      {2.48199, 2.31453, 24.59573, 24.59573, 9.81827, 5.23641, 0.39281, 0.64624, 0.17031,
       0.34061, -2.53499, -0.05960, 0.19306, -2.28232, -2.54000, 0.0, 0.20027, -0.14877,
       -2.54000, 0.0, 0.20027, 6.69465}};
  static const double eta_fe[3][2] = {{0.78591, -2.82453}, {1.01318, -2.70839}, {1.18290, -2.55187}};
  for (int m = 0; m < 3; ++m) {
    simox_metal& e = p->eam[m];
    const double* t = metal[m];
    e.r_e = t[0]; e.f_e = t[1]; e.rho_e = t[2]; e.rho_s = t[3]; e.alpha = t[4]; e.beta = t[5];
    e.A = t[6]; e.B = t[7]; e.kappa = t[8]; e.lambda = t[9];
    for (int k = 0; k < 4; ++k) { e.Fn[k] = t[10 + k]; e.Fm[k] = t[14 + k]; e.Fp[k] = t[18 + k]; }
    e.eta = eta_fe[m][0]; e.Fe = eta_fe[m][1];
  }
  // CTIP: mass, qmin, qmax, chi, J, xi, Z for Al, Ni, Fe, O, H (datafmt.f90:183-197)
  static const double ctip[5][7] = {{26.981539, 0.0, 3.0, -1.47914, 9.07222, 0.968, 1.07514},
                                    {58.6934, 0.0, 2.0, -1.70804, 9.10954, 1.087, 1.44450},
                                    {55.84500, 0.0, 3.0, -1.90587, 8.99819, 1.024, 1.28612},
                                    {15.9994, -2.0, 0.0, 2.000, 14.99523, 2.144, 0.0},
                                    {1.00794, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
  for (int s = 0; s < 5; ++s) {
    simox_potn& q = p->prm[s];
    q.xm = ctip[s][0]; q.qmin = ctip[s][1]; q.qmax = ctip[s][2]; q.chi = ctip[s][3];
    q.J = ctip[s][4]; q.xi = ctip[s][5]; q.Z = ctip[s][6];
  }
  simox_fafb(p);
  // oxygen EAM: per row r_e, alph, beta, A, B, kapp, lamb (datafmt.f90:225-239)
  static const double ox[5][7] = {{2.98520, 8.49741, 4.52114, 0.09738, 0.38121, 0.18967, 0.95234},
                                  {2.95732, 7.96528, 4.42411, 0.13521, 0.25332, 0.47077, 0.65524},
                                  {3.07992, 7.52309, 4.13330, 0.17108, 0.39869, 0.22335, 0.34380},
                                  {3.64857, 5.44072, 3.59746, 0.34900, 0.57438, 0.08007, 0.39310},
                                  {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
  for (int r = 0; r < 5; ++r) {
    p->ox.r_e[r] = ox[r][0]; p->ox.alph[r] = ox[r][1]; p->ox.beta[r] = ox[r][2];
    p->ox.A[r] = ox[r][3]; p->ox.B[r] = ox[r][4]; p->ox.kapp[r] = ox[r][5]; p->ox.lamb[r] = ox[r][6];
  }
  // O embedding pieces F0..F4, coefficient index 1..4 (datafmt.f90:241-249)
  static const double F[5][4] = {{-1.56489, -1.39123, 1.77199, 1.59833},
                                 {-1.58967, 1.30636, 9.81033, 0.0},
                                 {-1.54116, 2.02821, 6.56240, 0.0},
                                 {-1.51798, 2.30979, 7.69582, 0.0},
                                 {-1.19082, 4.12936, 10.32338, 0.0}};
  for (int k = 0; k < 5; ++k)
    for (int j = 0; j < 4; ++j) p->ox.F[k][j] = F[k][j];
  static const double re[5] = {54.62910, 64.26953, 66.21202, 66.92391, 74.23105};
  static const double rmax[5] = {54.62910, 65.24078, 66.56797, 70.57748, 99.99999e9};
  for (int k = 0; k < 5; ++k) { p->ox.re[k] = re[k]; p->ox.rmax[k] = rmax[k]; }
  p->ox.fe = 1.39478; p->ox.gamm = 2.11725; p->ox.nu = 0.37457;  // datafmt.f90:257-259
}

int simox_create(int n, const double box[3], const int fft[3], double rcut, int mode,
                 const simox_params* params, simox_ctx** out) {
  if (!out) return SIMOX_ERR_ARG;
  *out = nullptr;
  simox_ctx* ctx = new simox_ctx();
  *out = ctx;
  if (n <= 0 || rcut <= 0.0 || mode < 0 || mode > 2) {
    ctx->err = "invalid n / rcut / mode";
    return SIMOX_ERR_ARG;
  }
  ctx->n = n;
  ctx->mode = mode;
  ctx->rcut = rcut;
  ctx->prec_on = mode == SIMOX_MODE_MIXED;
  {
    const char* e = std::getenv("SIMOX_SMALL_PATH");
    ctx->small_path = n <= LS_SMALL_MAX && !(e && std::strcmp(e, "0") == 0);
    const char* eb = std::getenv("SIMOX_ITER_BATCH");
    if (eb) ctx->max_batch = std::max(1, std::min(64, std::atoi(eb)));
    const char* es = std::getenv("SIMOX_LS_SPARE");
    if (es) ctx->ls_spare = std::max(0, std::min(4, std::atoi(es)));
  }
  Geom& g = ctx->g;
  g.n = n;
  for (int d = 0; d < 3; ++d) {
    g.box[d] = box[d];
    g.hbox[d] = 0.5 * box[d];
    g.qbox[d] = 0.25 * box[d];
    g.nc[d] = int(box[d] / rcut);  // main.f90:189
    g.K[d] = fft[d];
    if (g.nc[d] < 2) {
      ctx->err = "==== Error - unitcell is too small or cut-off radius is too small";
      return SIMOX_ERR_ARG;
    }
    g.lcell[d] = box[d] / double(g.nc[d]);  // main.f90:190
    if (fft[d] < 2 || !make_plan(fft[d], &ctx->pm.pl[d])) {
      ctx->err = "PME grid sizes must be products of 2, 3 and 5";
      return SIMOX_ERR_ARG;
    }
  }
  if (fft[0] % BRK || fft[1] % BRK || fft[2] % BRK) {
    ctx->err = "PME grid sizes must be multiples of 4";
    return SIMOX_ERR_ARG;
  }
  g.ncall = g.nc[0] * g.nc[1] * g.nc[2];
  g.rcut2 = rcut * rcut;
  ctx->rdf_dr = rcut / double(NRDF);
  g.ream2 = g.rcut2;
  g.alpha = 3.8 / rcut;
  g.H = fft[0] / 2 + 1;
  ctx->V = box[0] * box[1] * box[2];
  ctx->c_pme = KC * double(fft[0]) * double(fft[1]) * double(fft[2]) / ctx->V / PI;  // ctip.f90:59
  if (params) ctx->prm = *params;
  else simox_potential_setup(&ctx->prm);
  build_tables(ctx);

  HIPCHK(hipStreamCreate(&ctx->st));
  CHK(dalloc(ctx, &ctx->dtab, 1));
  HIPCHK(hipMemcpy(ctx->dtab, &ctx->htab, sizeof(Tables), hipMemcpyHostToDevice));
  const size_t n3 = size_t(n) * 3;
  CHK(dalloc(ctx, &ctx->x, n3)); CHK(dalloc(ctx, &ctx->v, n3)); CHK(dalloc(ctx, &ctx->f, n3));
  CHK(dalloc(ctx, &ctx->x2, n3)); CHK(dalloc(ctx, &ctx->v2, n3)); CHK(dalloc(ctx, &ctx->f2, n3));
  CHK(dalloc(ctx, &ctx->q, n)); CHK(dalloc(ctx, &ctx->q2, n));
  CHK(dalloc(ctx, &ctx->id, n)); CHK(dalloc(ctx, &ctx->id2, n));
  CHK(dalloc(ctx, &ctx->perm, n)); CHK(dalloc(ctx, &ctx->perm2, n));
  CHK(dalloc(ctx, &ctx->cell, n)); CHK(dalloc(ctx, &ctx->cell2, n));
  CHK(dalloc(ctx, &ctx->slot, n)); CHK(dalloc(ctx, &ctx->order, n));
  CHK(dalloc(ctx, &ctx->count, g.ncall)); CHK(dalloc(ctx, &ctx->fill, g.ncall));
  CHK(dalloc(ctx, &ctx->start, g.ncall + 1));
  CHK(dalloc(ctx, &ctx->rho, n)); CHK(dalloc(ctx, &ctx->dF64, n)); CHK(dalloc(ctx, &ctx->dF32, n));
  CHK(dalloc(ctx, &ctx->nb_fwd, n));
  CHK(dalloc(ctx, &ctx->facc, n3)); CHK(dalloc(ctx, &ctx->eacc, 3));
  HIPCHK(hipMemset(ctx->facc, 0, n3 * sizeof(unsigned long long)));
  HIPCHK(hipMemset(ctx->eacc, 0, 3 * sizeof(unsigned long long)));
  ctx->nq_cap = mode == SIMOX_MODE_MIXED ? 8 * n + 1024 : 1;
  CHK(dalloc(ctx, &ctx->nq, size_t(ctx->nq_cap))); CHK(dalloc(ctx, &ctx->nq_n, 1));
  HIPCHK(hipMemset(ctx->nq_n, 0, sizeof(int)));
  if (const char* e = std::getenv("SIMOX_TILE_XCD")) ctx->tile_xcd = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_TILE_SHAPE")) {
    const int v = std::atoi(e);
    if (v >= 0 && v < NTILE_SHAPES) { ctx->tshape = v; ctx->tile_fixed = true; }
  }
  // the half-shell force pass with fixed-point sums: measured slower at C3
  // (15.8 ms against 4.0 ms for the owner-computes pass), so off by default
  if (const char* e = std::getenv("SIMOX_FORCE_HALF")) ctx->force_half = std::atoi(e) != 0;
  CHK(dalloc(ctx, &ctx->frec, n3));
  CHK(dalloc(ctx, &ctx->qc, n)); CHK(dalloc(ctx, &ctx->ql, n)); CHK(dalloc(ctx, &ctx->gv, n));
  CHK(dalloc(ctx, &ctx->dv, n)); CHK(dalloc(ctx, &ctx->pv, n)); CHK(dalloc(ctx, &ctx->phi, n));
  CHK(dalloc(ctx, &ctx->rv, n)); CHK(dalloc(ctx, &ctx->zv, n)); CHK(dalloc(ctx, &ctx->apv, n));
  CHK(dalloc(ctx, &ctx->mv, n));
  for (int h = 0; h < 3; ++h) CHK(dalloc(ctx, &ctx->qhist[h], n));
  CHK(dalloc(ctx, &ctx->qtmp, n));
  // coarse inner grid by default in MIXED from 2^23 grid points on (measured:
  // C3, 256 x 256 x 512, 361 -> 287 ms per step; C2, 96 x 96 x 192, slower)
  ctx->inner_coarse = mode == SIMOX_MODE_MIXED &&
                      double(fft[0]) * double(fft[1]) * double(fft[2]) >= 8388608.0;
  if (const char* e = std::getenv("SIMOX_INNER_COARSE")) ctx->inner_coarse = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_HALO_MAX")) {
    const int v = std::atoi(e);
    if (v > 0 && v <= HALO_MAX) ctx->halo_max = v;
  }
  if (const char* e = std::getenv("SIMOX_MV_VARIANT")) {
    const int v = std::atoi(e);
    if (v >= 0 && v < NMV_VARIANTS) { ctx->mv_variant = v; ctx->mv_fixed = true; }
  }
  ctx->gather_fixed = std::getenv("SIMOX_GATHER_SHAPE") != nullptr;
  ctx->xynt_fixed = std::getenv("SIMOX_FFT_XYNT") != nullptr;
  ctx->nl_fixed = std::getenv("SIMOX_FFT_NL") != nullptr;
  if (const char* e = std::getenv("SIMOX_INNER_F16")) ctx->inner_f16 = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_INNER_SPREAD32")) ctx->inner_spread32 = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_PREC_SPREAD32")) ctx->prec_spread32 = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_SPREAD_BIG")) ctx->spread_big = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_SPREAD_OWN")) ctx->spread_own = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_FORCE_BATCH")) ctx->force_batch = std::atoi(e);
  if (const char* e = std::getenv("SIMOX_DENS_BATCH")) ctx->dens_batch = std::atoi(e);
  if (const char* e = std::getenv("SIMOX_CTIP_BATCH")) ctx->ctip_batch = std::atoi(e);
  // the preconditioner's grid: a quarter of the resolution from 2^23 points
  // (C3, 256 x 256 x 512: same iteration count as half, 143 -> 135 ms per
  // step), half below (C2, 96 x 96 x 192: a quarter needs more iterations)
  ctx->prec_div = double(fft[0]) * fft[1] * fft[2] >= double(1 << 23) ? 4 : 2;
  if (const char* e = std::getenv("SIMOX_PREC_DIV")) ctx->prec_div = std::atoi(e) == 4 ? 4 : 2;
  // XCD-aware atom blocks in the pair kernels: measured slower at C3 (k_force
  // 5.0 -> 6.2 ms, k_ctip_matrix 3.2 -> 4.1 ms), so off unless SIMOX_PAIR_XCD=1
  ctx->g.xcd = 0;
  if (const char* e = std::getenv("SIMOX_PAIR_XCD")) ctx->g.xcd = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_INNER_DIV")) ctx->inner_div = std::atoi(e) == 4 ? 4 : 2;
  if (const char* e = std::getenv("SIMOX_INNER_PACK")) ctx->inner_pack = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_LS_ADAPT")) ctx->ls_adapt = std::atoi(e) != 0;
  if (const char* e = std::getenv("SIMOX_INNER_DROP")) {
    const double v = std::atof(e);
    if (v >= 0.0 && v < 1.0) ctx->inner_drop = v;
  }
  if (const char* e = std::getenv("SIMOX_INNER_RTOL")) {
    const double v = std::atof(e);
    if (v > 0.0 && v < 1.0) ctx->inner_rtol = ctx->inner_rtol_c = v;
  }
  if (const char* e = std::getenv("SIMOX_PREDICTOR_ORDER")) {
    const int o = std::atoi(e);
    if (o >= 0 && o <= 3) ctx->pred_order = o;
  }
  CHK(dalloc(ctx, &ctx->bvec, n)); CHK(dalloc(ctx, &ctx->nb_cnt, n));
  CHK(dalloc(ctx, &ctx->nb_col, size_t(ctx->cap) * n));
  CHK(dalloc(ctx, &ctx->nb_val64, size_t(ctx->cap) * n));
  if (mode == SIMOX_MODE_MIXED) {
    CHK(dalloc(ctx, &ctx->nb_val32, size_t(ctx->cap) * n));
    CHK(dalloc(ctx, &ctx->pv32, n));
  }
  CHK(dalloc(ctx, &ctx->sv, n)); CHK(dalloc(ctx, &ctx->cv, n));
  {
    const int rc = init_grid(ctx, ctx->pm, fft, true, 4);
    if (rc == SIMOX_ERR_ARG) ctx->err = "PME grid dimension too large for the LDS-resident FFT";
    if (rc) return rc;
    CHK((gather_attr<PGX, PGY, PGZ>(ctx)));
    CHK(gather_rec_attrs(ctx));
  }
  CHK(dalloc(ctx, &ctx->dscal, 72));
  CHK(dalloc(ctx, &ctx->part_eg, grid_blocks(n)));
  HIPCHK(hipHostMalloc((void**)&ctx->hscal, 64 * sizeof(double), hipHostMallocDefault));
  CHK(dalloc(ctx, &ctx->derr, 1));
  HIPCHK(hipMemset(ctx->derr, 0, sizeof(int)));
  CHK(dalloc(ctx, &ctx->red_ctr, 1));
  HIPCHK(hipMemset(ctx->red_ctr, 0, sizeof(unsigned)));
  HIPCHK(hipHostMalloc((void**)&ctx->herr, sizeof(int), hipHostMallocDefault));
  {
    // partial-sum buffer sized once for every launch that uses it
    const size_t zb = size_t(std::max(ctx->pm.fc64.zblocks, ctx->pm.fc32.zblocks));
    CHK(ensure_part(ctx, std::max(size_t(grid_blocks(n)) * 10, zb) + 16));
  }
  hipLaunchKernelGGL(k_iota, dim3(grid_blocks(n)), dim3(BLOCK), 0, ctx->st, n, ctx->order);
  HIPCHK(hipGetLastError());
  {
    const char* e = std::getenv("SIMOX_AUTOTUNE");
    if (!e || std::strcmp(e, "0") != 0) CHK(autotune(ctx));
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  return SIMOX_OK;
}

int simox_autotune(simox_ctx* ctx) {
  if (!ctx) return SIMOX_ERR_ARG;
  CHK(autotune(ctx));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return SIMOX_OK;
}

int simox_fft_shapes(const simox_ctx* ctx, int shapes[6]) {
  if (!ctx || !shapes) return SIMOX_ERR_ARG;
  const FftCfg* c[2] = {&ctx->pm.fc64, &ctx->pm.fc32};
  for (int k = 0; k < 2; ++k) {
    shapes[3 * k] = c[k]->mx;
    shapes[3 * k + 1] = c[k]->my;
    shapes[3 * k + 2] = c[k]->mz;
  }
  return SIMOX_OK;
}

int simox_tile_shape(const simox_ctx* ctx, int shape[2]) {
  if (!ctx || !shape) return SIMOX_ERR_ARG;
  shape[0] = TILE_SHAPES[ctx->tshape][0];
  shape[1] = TILE_SHAPES[ctx->tshape][1];
  return SIMOX_OK;
}

int simox_inner_shapes(const simox_ctx* ctx, int shapes[6]) {
  if (!ctx || !shapes) return SIMOX_ERR_ARG;
  shapes[0] = ctx->mv_variant;
  shapes[1] = ctx->pi.gshape;
  shapes[2] = ctx->pi.fc32.plane ? ctx->pi.fc32.xynt : 0;
  shapes[3] = ctx->pi.fc32.plane ? ctx->pi.fc32.nl : 0;
  shapes[4] = ctx->pc.fc32.plane ? ctx->pc.fc32.xynt : 0;
  shapes[5] = ctx->pc.fc32.plane ? ctx->pc.fc32.nl : 0;
  return SIMOX_OK;
}

int simox_destroy(simox_ctx* ctx) {
  if (!ctx) return SIMOX_OK;
  if (ctx->st) (void)hipStreamSynchronize(ctx->st);
  tresolve(ctx);
  if (ctx->iter_graph) (void)hipGraphExecDestroy(ctx->iter_graph);
  for (hipEvent_t e : ctx->evfree) (void)hipEventDestroy(e);
  void* bufs[] = {ctx->dtab, ctx->x, ctx->v, ctx->f, ctx->q, ctx->id, ctx->perm, ctx->cell,
                  ctx->x2, ctx->v2, ctx->f2, ctx->q2, ctx->id2, ctx->perm2, ctx->cell2,
                  ctx->count, ctx->start, ctx->fill, ctx->slot, ctx->order, ctx->rho, ctx->dF64,
                  ctx->dF32, ctx->nb_fwd, ctx->nq, ctx->nq_n, ctx->facc, ctx->eacc, ctx->frec, ctx->sv, ctx->cv, ctx->spl_grad, ctx->vjp_u, ctx->vjp_t, ctx->vjp_up, ctx->dip_acc, ctx->nb_cnt, ctx->nb_col, ctx->nb_val64,
                  ctx->nb_val32, ctx->nb_val16, ctx->nb_col16, ctx->nb_cnt16, ctx->nb_pk, ctx->pk_words, ctx->nb_pkh, ctx->nb_pk24, ctx->pk_runs, ctx->pk_hcount, ctx->pk_slen, ctx->pk_soff, ctx->pk_bsum, ctx->pk_total, ctx->bvec, ctx->qc, ctx->ql, ctx->gv, ctx->dv, ctx->pv, ctx->phi,
                  ctx->rv, ctx->zv, ctx->apv, ctx->mv, ctx->qhist[0], ctx->qhist[1], ctx->qhist[2], ctx->qtmp, ctx->pv32, ctx->rdf, ctx->adj,
                  ctx->part, ctx->dscal, ctx->derr, ctx->red_ctr, ctx->dstage, ctx->part_eg};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  free_grid(ctx->pm);
  free_grid(ctx->pc);
  free_grid(ctx->pi);
  if (ctx->hscal) (void)hipHostFree(ctx->hscal);
  if (ctx->herr) (void)hipHostFree(ctx->herr);
  if (ctx->hstage) (void)hipHostFree(ctx->hstage);
  if (ctx->st) (void)hipStreamDestroy(ctx->st);
  delete ctx;
  return SIMOX_OK;
}

const char* simox_last_error(const simox_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int simox_get_geometry(const simox_ctx* ctx, int ncell[3], double lcell[3], double* alpha) {
  if (!ctx) return SIMOX_ERR_ARG;
  for (int d = 0; d < 3; ++d) {
    if (ncell) ncell[d] = ctx->g.nc[d];
    if (lcell) lcell[d] = ctx->g.lcell[d];
  }
  if (alpha) *alpha = ctx->g.alpha;
  return SIMOX_OK;
}

int simox_set_mode(simox_ctx* ctx, int mode) {
  if (!ctx) return SIMOX_ERR_ARG;
  const bool was32 = ctx->mode == SIMOX_MODE_MIXED, is32 = mode == SIMOX_MODE_MIXED;
  if (mode < 0 || mode > 2 || was32 != is32) {
    ctx->err = "mode change between FP64 and MIXED storage needs a new context";
    return SIMOX_ERR_ARG;
  }
  ctx->mode = mode;
  return SIMOX_OK;
}

int simox_set_charge_precond(simox_ctx* ctx, int on) {
  if (!ctx) return SIMOX_ERR_ARG;
  ctx->prec_on = on != 0;
  return SIMOX_OK;
}

int simox_set_inner_grid(simox_ctx* ctx, int on) {
  if (!ctx) return SIMOX_ERR_ARG;
  if (on && ctx->mode != SIMOX_MODE_MIXED) {
    ctx->err = "the coarse inner grid belongs to the MIXED mode's inner solves";
    return SIMOX_ERR_ARG;
  }
  ctx->inner_coarse = on != 0;
  return SIMOX_OK;
}

int simox_inner_entries(simox_ctx* ctx, long long* entries) {
  if (!ctx || !entries) return SIMOX_ERR_ARG;
  const int* cnt = ctx->val16_valid ? ctx->nb_cnt16 : ctx->nb_cnt;
  if (!cnt) return SIMOX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  std::vector<int> h(size_t(ctx->n));
  HIPCHK(hipMemcpy(h.data(), cnt, h.size() * sizeof(int), hipMemcpyDeviceToHost));
  long long t = 0;
  for (int c : h) t += std::min(c, ctx->cap);
  *entries = t;
  return SIMOX_OK;
}

int simox_charge_stats(simox_ctx* ctx, int* iters, int* restarts) {
  if (!ctx) return SIMOX_ERR_ARG;
  if (iters) *iters = ctx->last_inner;
  if (restarts) *restarts = ctx->last_outer;
  return SIMOX_OK;
}

int simox_set_charge_tol(simox_ctx* ctx, double tol, int maxit) {
  if (!ctx || tol <= 0.0 || maxit <= 0) return SIMOX_ERR_ARG;
  ctx->tol = tol;
  ctx->maxit = maxit;
  return SIMOX_OK;
}

int simox_set_state(simox_ctx* ctx, const int* id, const double* x, const double* v,
                    const double* f, const double* q) {
  if (!ctx) return SIMOX_ERR_ARG;
  const int n = ctx->n;
  if (id) {
    for (int i = 0; i < n; ++i) {
      if (id[i] < 1 || id[i] > 5) {
        ctx->err = "=== Particle type error ===";  // main.f90:241
        return SIMOX_ERR_ARG;
      }
      if (id[i] > NSPEC) {  // H has no pair functions (force.f90:72-73)
        ctx->err = "=== Unknown particle pair ===";
        return SIMOX_ERR_UNKNOWN_PAIR;
      }
    }
    if (!x) {
      ctx->err = "positions are required together with species ids";
      return SIMOX_ERR_ARG;
    }
    ctx->spec_present = 0u;
    for (int i = 0; i < n; ++i) ctx->spec_present |= 1u << id[i];
    set_ctip64(ctx);
    hipLaunchKernelGGL(k_iota, dim3(grid_blocks(n)), dim3(BLOCK), 0, ctx->st, n, ctx->perm);
    HIPCHK(hipMemcpyAsync(ctx->id, id, sizeof(int) * n, hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipMemsetAsync(ctx->v, 0, sizeof(double) * 3 * n, ctx->st));
    HIPCHK(hipMemsetAsync(ctx->f, 0, sizeof(double) * 3 * n, ctx->st));
    HIPCHK(hipMemsetAsync(ctx->q, 0, sizeof(double) * n, ctx->st));
    CHK(upload_caller(ctx, x, 3, ctx->x));
    if (v) CHK(upload_caller(ctx, v, 3, ctx->v));
    if (f) CHK(upload_caller(ctx, f, 3, ctx->f));
    if (q) CHK(upload_caller(ctx, q, 1, ctx->q));
    ctx->have_state = true;
    ctx->nhist = 0;
    CHK(do_cell_sort(ctx));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return check_err(ctx, 0);
  }
  if (!need_state(ctx)) return SIMOX_ERR_ARG;
  if (x) CHK(upload_caller(ctx, x, 3, ctx->x));
  if (v) CHK(upload_caller(ctx, v, 3, ctx->v));
  if (f) CHK(upload_caller(ctx, f, 3, ctx->f));
  if (q) { CHK(upload_caller(ctx, q, 1, ctx->q)); ctx->nhist = 0; }
  if (x) { ctx->matrix_valid = false; ctx->nlist_valid = false; ctx->pm.bricks_valid = ctx->pc.bricks_valid = ctx->pi.bricks_valid = false; }
  HIPCHK(hipStreamSynchronize(ctx->st));
  return SIMOX_OK;
}

int simox_get_state(simox_ctx* ctx, double* x, double* v, double* f, double* q) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  if (x) CHK(download_caller(ctx, ctx->x, 3, x));
  if (v) CHK(download_caller(ctx, ctx->v, 3, v));
  if (f) CHK(download_caller(ctx, ctx->f, 3, f));
  if (q) CHK(download_caller(ctx, ctx->q, 1, q));
  return SIMOX_OK;
}

// Adjoint of the converged charge solve: H lam = P b on sum(lam) = 0 (H the
// solver's Hessian on the active set at ctx->q), Jacobi-preconditioned
// projected CG in FP64; b = adj[0:n], lam = adj[n:2n].
int adjoint_solve(simox_ctx* ctx, double tol, int maxit, int* iters) {
  const int n = ctx->n, nb = grid_blocks(n);
  CHK(ensure_part(ctx, size_t(nb) * 10));
  double* b = ctx->adj;
  double* lam = b + n;
  double* r = b + 2 * size_t(n);
  double* z = b + 3 * size_t(n);
  double* p = b + 4 * size_t(n);
  double* y = b + 5 * size_t(n);
  double s[3];
  auto project = [&](double* v) -> int {
    hipLaunchKernelGGL(k_adj_sums, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id, ctx->q,
                       v, ctx->part);
    CHK(reduce(ctx, nb, 3, 0, 28));
    hipLaunchKernelGGL(k_adj_center, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dscal + 28, v);
    HIPCHK(hipGetLastError());
    return 0;
  };
  auto precond = [&](double* rz) -> int {  // z = P_M M^-1 r, *rz = r.z
    hipLaunchKernelGGL(k_adj_sums, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id, ctx->q,
                       r, ctx->part);
    CHK(reduce(ctx, nb, 3, 0, 28));
    hipLaunchKernelGGL(k_adj_prec, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id, ctx->q,
                       r, ctx->dscal + 28, z, ctx->part);
    CHK(reduce(ctx, nb, 1, 0, 31));
    return fetch(ctx, 31, 1, rz);
  };
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(BLOCK), 0, ctx->st, n, b, r);
  hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(BLOCK), 0, ctx->st, n, 0.0, lam);
  CHK(project(r));
  double rz = 0.0;
  CHK(precond(&rz));
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(BLOCK), 0, ctx->st, n, z, p);
  int it = 0;
  for (; it < maxit; ++it) {
    // y = P H p
    CHK(pme_eval(ctx, p, nullptr, 0.0, false));
    hipLaunchKernelGGL(k_phi_gather<double>, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->x,
                       ctx->pm.greal64, ctx->c_pme, ctx->phi);
    hipLaunchKernelGGL((k_h0matvec<double, double>), dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g,
                       ctx->dtab, ctx->x, ctx->id, ctx->q, p, p, p, ctx->phi, ctx->nb_cnt,
                       ctx->nb_col, ctx->nb_val64, y, ctx->part);
    hipLaunchKernelGGL(k_add_penalty, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dtab, ctx->id, ctx->q,
                       p, y, ctx->part);
    CHK(reduce(ctx, nb, 1, 0, 32));
    CHK(fetch(ctx, 32, 1, s));
    if (!(s[0] > 0.0)) break;
    CHK(project(y));
    const double a = rz / s[0];
    hipLaunchKernelGGL(k_adj_update, dim3(nb), dim3(BLOCK), 0, ctx->st, n, a, p, y, lam, r, ctx->part);
    CHK(reduce(ctx, nb, 2, 0, 33));
    CHK(fetch(ctx, 33, 1, s));
    if (std::sqrt(s[0] / n) <= tol) { ++it; break; }
    double rz_new = 0.0;
    CHK(precond(&rz_new));
    hipLaunchKernelGGL(k_adj_dir, dim3(nb), dim3(BLOCK), 0, ctx->st, n, rz_new / rz, z, p);
    HIPCHK(hipGetLastError());
    rz = rz_new;
  }
  if (iters) *iters = it;
  if (it >= maxit) {
    ctx->err = "adjoint charge solve did not reach the requested tolerance";
    return SIMOX_ERR_NOT_CONVERGED;
  }
  return 0;
}

int param_grad_impl(simox_ctx* ctx, double* grad, bool with_xi, double* lambda, int* iters) {
  if (!ctx || !need_state(ctx) || !grad) return SIMOX_ERR_ARG;
  if (ctx->mode == SIMOX_MODE_FAITHFUL) {
    ctx->err = "parameter gradients need a converged charge solve (modes FP64, MIXED)";
    return SIMOX_ERR_ARG;
  }
  const int n = ctx->n, nb = grid_blocks(n);
  if (!ctx->adj) CHK(dalloc(ctx, &ctx->adj, 6 * size_t(n)));
  if (!ctx->matrix_valid) CHK(build_matrix(ctx));
  // b = dE/dq at the current charges (PME potential with exact weights)
  CHK(pme_eval(ctx, ctx->q, nullptr, 0.0, false));
  hipLaunchKernelGGL(k_phi_gather<double>, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->x,
                     ctx->pm.greal64, ctx->c_pme, ctx->phi);
  hipLaunchKernelGGL(k_denergy_dq, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab, ctx->x,
                     ctx->id, ctx->q, ctx->phi, ctx->nb_cnt, ctx->nb_col, ctx->adj);
  HIPCHK(hipGetLastError());
  CHK(adjoint_solve(ctx, 1e-12, 20000, iters));
  CHK(ensure_part(ctx, size_t(nb) * 10));
  hipLaunchKernelGGL(k_param_grad, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->id, ctx->q,
                     ctx->adj + n, ctx->part);
  CHK(reduce(ctx, nb, 10, 0, 50));
  CHK(fetch(ctx, 50, 10, grad));
  if (with_xi) {
    // dE/dxi_s = kc * dEctip/dxi_s (at fixed q) - lambda . dG/dxi_s
    hipLaunchKernelGGL(k_xi_grad, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab, ctx->x,
                       ctx->id, ctx->q, ctx->adj + n, nullptr, ctx->nb_cnt, ctx->nb_col, ctx->part);
    HIPCHK(hipGetLastError());
    CHK(reduce(ctx, nb, 10, 0, 50));
    double t[10];
    CHK(fetch(ctx, 50, 10, t));
    for (int sp = 0; sp < 5; ++sp) grad[10 + sp] = KC * t[sp] - t[5 + sp];
  }
  if (lambda) CHK(download_caller(ctx, ctx->adj + n, 1, lambda));
  return SIMOX_OK;
}

int simox_param_grad(simox_ctx* ctx, double grad[10], double* lambda, int* iters) {
  return param_grad_impl(ctx, grad, false, lambda, iters);
}

int simox_param_grad_all(simox_ctx* ctx, double grad[15], double* lambda, int* iters) {
  return param_grad_impl(ctx, grad, true, lambda, iters);
}

// d(sum_i u_i . F_i)/dtheta, theta = (chi, J, xi) of the five species, F the
// forces of FORCE + PME, through the charge equilibration: b = dS/dq (pair
// terms and PME, k_qvjp_*), H mu = P b (the adjoint solve), then
// -mu . dG/dtheta plus, for xi, the direct CTIP force term.
int simox_param_force_grad(simox_ctx* ctx, const double* u, double grad[15], int* iters) {
  if (!ctx || !need_state(ctx) || !grad) return SIMOX_ERR_ARG;
  if (ctx->mode == SIMOX_MODE_FAITHFUL) {
    ctx->err = "parameter gradients need a converged charge solve (modes FP64, MIXED)";
    return SIMOX_ERR_ARG;
  }
  const int n = ctx->n, nb = grid_blocks(n);
  const size_t ng = size_t(ctx->g.K[0]) * ctx->g.K[1] * ctx->g.K[2];
  if (!ctx->adj) CHK(dalloc(ctx, &ctx->adj, 6 * size_t(n)));
  if (!ctx->vjp_t) CHK(dalloc(ctx, &ctx->vjp_t, size_t(n)));
  if (!ctx->vjp_up) CHK(dalloc(ctx, &ctx->vjp_up, 3 * size_t(n)));
  if (!ctx->dip_acc) CHK(dalloc(ctx, &ctx->dip_acc, ng));
  CHK(ensure_nlist(ctx));
  if (!ctx->matrix_valid) CHK(build_matrix(ctx));
  const double* ud = ctx->f;  // the forces of the last FORCE + PME
  if (u) {
    if (!ctx->vjp_u) CHK(dalloc(ctx, &ctx->vjp_u, 3 * size_t(n)));
    CHK(upload_caller(ctx, u, 3, ctx->vjp_u));
    ud = ctx->vjp_u;
  }
  CHK(ensure_part(ctx, size_t(nb) * 10));
  double* b = ctx->adj;
  // pair terms
  hipLaunchKernelGGL(k_qvjp_pair, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab, ctx->x,
                     ctx->id, ctx->q, ud, ctx->nb_cnt, ctx->nb_col, b);
  // PME: u' = u - mean(u); t1 with the grid of the charges; the dipole grid
  hipLaunchKernelGGL(k_sum3, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ud, ctx->part);
  HIPCHK(hipGetLastError());
  CHK(reduce(ctx, nb, 3, 0, 35));
  hipLaunchKernelGGL(k_center3, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ud, ctx->dscal + 35,
                     ctx->vjp_up);
  HIPCHK(hipGetLastError());
  CHK(pme_eval(ctx, ctx->q, nullptr, 0.0, false));
  hipLaunchKernelGGL(k_qvjp_rec1<double>, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->x,
                     ctx->vjp_up, ctx->pm.greal64, ctx->vjp_t);
  const int ngb = int((ng + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_fill_u64, dim3(ngb), dim3(BLOCK), 0, ctx->st, ng, ctx->dip_acc);
  hipLaunchKernelGGL(k_qvjp_dipole, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->x, ctx->q,
                     ctx->vjp_up, ctx->dip_acc);
  hipLaunchKernelGGL(k_fixed_to_grid, dim3(ngb), dim3(BLOCK), 0, ctx->st, ng, ctx->dip_acc,
                     ctx->pm.greal64);
  HIPCHK(hipGetLastError());
  CHK(pme_chain<double2>(ctx, ctx->pm, ctx->pm.spec64, ctx->pm.tw64, ctx->pm.fc64, ctx->pm.G64,
                         ctx->pm.greal64));
  const double cst = KC * double(ng) / ctx->V / PI;  // force.f90:436
  hipLaunchKernelGGL(k_qvjp_rec2<double>, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->x,
                     ctx->pm.greal64, ctx->vjp_t, cst, b);
  HIPCHK(hipGetLastError());
  // adjoint: H mu = P b
  CHK(adjoint_solve(ctx, 1e-12, 20000, iters));
  CHK(ensure_part(ctx, size_t(nb) * 10));
  hipLaunchKernelGGL(k_mu_sums, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->id, ctx->q, b + n,
                     ctx->part);
  HIPCHK(hipGetLastError());
  CHK(reduce(ctx, nb, 10, 0, 50));
  double t[10];
  CHK(fetch(ctx, 50, 10, t));
  for (int sp = 0; sp < 10; ++sp) grad[sp] = -t[sp];
  hipLaunchKernelGGL(k_xi_grad, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab, ctx->x,
                     ctx->id, ctx->q, b + n, ud, ctx->nb_cnt, ctx->nb_col, ctx->part);
  HIPCHK(hipGetLastError());
  CHK(reduce(ctx, nb, 10, 0, 50));
  CHK(fetch(ctx, 50, 10, t));
  for (int sp = 0; sp < 5; ++sp) grad[10 + sp] = t[sp] - t[5 + sp];
  return SIMOX_OK;
}

int simox_set_reflect(simox_ctx* ctx, int on) {
  if (!ctx) return SIMOX_ERR_ARG;
  ctx->reflect = on != 0;
  return SIMOX_OK;
}

int simox_remove_rigid_motion(simox_ctx* ctx) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  const int n = ctx->n, nb = grid_blocks(n);
  CHK(ensure_part(ctx, size_t(nb) * 4));
  hipLaunchKernelGGL(k_rigid_sums, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->x, ctx->v, ctx->part);
  CHK(reduce(ctx, nb, 4, 0, 60));
  hipLaunchKernelGGL(k_rigid_apply, dim3(nb), dim3(BLOCK), 0, ctx->st, n, ctx->dscal + 60, ctx->x,
                     ctx->v);
  HIPCHK(hipGetLastError());
  ctx->matrix_valid = false;
  ctx->nlist_valid = false;
  ctx->pm.bricks_valid = ctx->pc.bricks_valid = ctx->pi.bricks_valid = false;
  HIPCHK(hipStreamSynchronize(ctx->st));
  return SIMOX_OK;
}

int simox_set_rdf(simox_ctx* ctx, int on) {
  if (!ctx) return SIMOX_ERR_ARG;
  if (on && !ctx->rdf) {
    CHK(dalloc(ctx, &ctx->rdf, RDF_BINS));
    HIPCHK(hipMemsetAsync(ctx->rdf, 0, RDF_BINS * sizeof(unsigned long long), ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
  }
  ctx->rdf_on = on != 0;
  return SIMOX_OK;
}

int simox_get_rdf(simox_ctx* ctx, long long* hist, int reset) {
  if (!ctx || !hist) return SIMOX_ERR_ARG;
  if (!ctx->rdf) {
    for (int k = 0; k < RDF_BINS; ++k) hist[k] = 0;
    return SIMOX_OK;
  }
  HIPCHK(hipMemcpyAsync(hist, ctx->rdf, RDF_BINS * sizeof(long long), hipMemcpyDeviceToHost, ctx->st));
  if (reset) HIPCHK(hipMemsetAsync(ctx->rdf, 0, RDF_BINS * sizeof(unsigned long long), ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return SIMOX_OK;
}

int simox_set_thermostat(simox_ctx* ctx, int kind, double t_kelvin, double alpha,
                         unsigned long long seed) {
  if (!ctx || kind < 0 || kind > 2 || !(t_kelvin >= 0.0) || !(alpha >= 0.0)) return SIMOX_ERR_ARG;
  ctx->thermo = kind;
  ctx->t_given = t_kelvin * 8.617343e-5;  // main.f90:160, kB of datafmt.f90:27
  ctx->alpha = alpha;
  ctx->seed = seed;
  ctx->nvt_step = 0;
  return SIMOX_OK;
}

int simox_set_eam_cutoff(simox_ctx* ctx, double r_eam) {
  if (!ctx) return SIMOX_ERR_ARG;
  if (!(r_eam > 0.0) || r_eam * r_eam > ctx->g.rcut2) {
    ctx->err = "EAM cutoff must be in (0, Rcut]";
    return SIMOX_ERR_ARG;
  }
  ctx->g.ream2 = r_eam * r_eam;
  ctx->nlist_valid = false;
  return SIMOX_OK;
}

int simox_set_params(simox_ctx* ctx, const simox_params* params) {
  if (!ctx || !params) return SIMOX_ERR_ARG;
  ctx->prm = *params;
  build_tables(ctx);
  set_ctip64(ctx);
  HIPCHK(hipMemcpyAsync(ctx->dtab, &ctx->htab, sizeof(Tables), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->matrix_valid = false;
  ctx->nlist_valid = false;
  return SIMOX_OK;
}

int simox_cell_sort(simox_ctx* ctx) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  CHK(do_cell_sort(ctx));
  return check_err(ctx, 0);
}

int simox_get_cells(simox_ctx* ctx, int* npt, int* link, int link_stride) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  std::vector<int> start(ctx->g.ncall + 1), perm(ctx->n);
  HIPCHK(hipMemcpy(start.data(), ctx->start, sizeof(int) * (ctx->g.ncall + 1), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(perm.data(), ctx->perm, sizeof(int) * ctx->n, hipMemcpyDeviceToHost));
  for (int c = 0; c < ctx->g.ncall; ++c) {
    const int m = start[c + 1] - start[c];
    if (npt) npt[c] = m;
    if (link)
      for (int k = 0; k < m && k < link_stride; ++k) link[size_t(c) * link_stride + k] = perm[start[c] + k];
  }
  return SIMOX_OK;
}

int simox_force(simox_ctx* ctx, double energies[4]) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  CHK(do_force(ctx, true));
  CHK(check_err(ctx, 0));
  if (energies) {
    energies[0] = ctx->Epair; energies[1] = ctx->Edir; energies[2] = ctx->Ectip; energies[3] = ctx->Egroup;
  }
  return SIMOX_OK;
}

int simox_set_eam_splines(simox_ctx* ctx, int m, const double* lo, const double* h,
                          const double* coef) {
  if (!ctx) return SIMOX_ERR_ARG;
  SplineSet& sp = ctx->htab.sp;
  if (!coef) {
    sp.on = 0;
  } else {
    if (m < 4 || m > SPL_MAX || !lo || !h) {
      ctx->err = "EAM splines: 4 <= m <= 64 coefficients per function, lo and h required";
      return SIMOX_ERR_ARG;
    }
    for (int f = 0; f < NSPL; ++f)
      if (!(h[f] > 0.0)) {
        ctx->err = "EAM splines: knot spacing must be positive";
        return SIMOX_ERR_ARG;
      }
    sp.on = 1;
    sp.m = m;
    for (int f = 0; f < NSPL; ++f) {
      sp.lo[f] = lo[f];
      sp.h[f] = h[f];
      for (int k = 0; k < SPL_MAX; ++k) sp.c[f][k] = k < m ? coef[size_t(f) * m + k] : 0.0;
    }
  }
  HIPCHK(hipMemcpyAsync(ctx->dtab, &ctx->htab, sizeof(Tables), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->nlist_valid = false;  // densities and embedding follow the new functions
  return SIMOX_OK;
}

int simox_eam_spline_grad(simox_ctx* ctx, double* grad) {
  if (!ctx || !grad || !need_state(ctx)) return SIMOX_ERR_ARG;
  if (!ctx->htab.sp.on) {
    ctx->err = "EAM splines are not set";
    return SIMOX_ERR_ARG;
  }
  const int n = ctx->n, nb = grid_blocks(n), nv = NSPL * SPL_MAX;
  CHK(ensure_nlist(ctx));
  CHK(check_err(ctx, 0));
  if (!ctx->spl_grad) CHK(dalloc(ctx, &ctx->spl_grad, nv));
  CHK(ensure_part(ctx, size_t(nb) * nv));
  hipLaunchKernelGGL(k_eam_spline_grad, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab,
                     ctx->x, ctx->id, ctx->rho, ctx->nb_cnt, ctx->nb_col, ctx->part);
  hipLaunchKernelGGL(k_sum_rows, dim3((nv + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->st,
                     ctx->part, nb, nv, ctx->spl_grad);
  HIPCHK(hipGetLastError());
  std::vector<double> all(nv);
  HIPCHK(hipMemcpyAsync(all.data(), ctx->spl_grad, nv * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  const int m = ctx->htab.sp.m;
  for (int f = 0; f < NSPL; ++f)
    for (int k = 0; k < m; ++k) grad[size_t(f) * m + k] = all[size_t(f) * SPL_MAX + k];
  return SIMOX_OK;
}

int simox_eam_spline_force_vjp(simox_ctx* ctx, const double* u, double* grad) {
  if (!ctx || !grad || !need_state(ctx)) return SIMOX_ERR_ARG;
  if (!ctx->htab.sp.on) {
    ctx->err = "EAM splines are not set";
    return SIMOX_ERR_ARG;
  }
  const int n = ctx->n, nb = grid_blocks(n), nv = NSPL * SPL_MAX;
  CHK(ensure_nlist(ctx));
  CHK(check_err(ctx, 0));
  if (!ctx->spl_grad) CHK(dalloc(ctx, &ctx->spl_grad, nv));
  if (!ctx->vjp_t) CHK(dalloc(ctx, &ctx->vjp_t, size_t(n)));
  const double* ud = ctx->f;  // the forces of the last FORCE (+ PME)
  if (u) {
    if (!ctx->vjp_u) CHK(dalloc(ctx, &ctx->vjp_u, 3 * size_t(n)));
    CHK(upload_caller(ctx, u, 3, ctx->vjp_u));
    ud = ctx->vjp_u;
  }
  CHK(ensure_part(ctx, size_t(nb) * nv));
  hipLaunchKernelGGL(k_eam_vjp_t, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab, ctx->x,
                     ctx->id, ud, ctx->nb_cnt, ctx->nb_col, ctx->vjp_t);
  hipLaunchKernelGGL(k_eam_vjp, dim3(nb), dim3(BLOCK), 0, ctx->st, ctx->g, ctx->dtab, ctx->x,
                     ctx->id, ctx->rho, ud, ctx->vjp_t, ctx->nb_cnt, ctx->nb_col, ctx->part);
  hipLaunchKernelGGL(k_sum_rows, dim3((nv + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->st,
                     ctx->part, nb, nv, ctx->spl_grad);
  HIPCHK(hipGetLastError());
  std::vector<double> all(nv);
  HIPCHK(hipMemcpyAsync(all.data(), ctx->spl_grad, nv * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  const int m = ctx->htab.sp.m;
  for (int f = 0; f < NSPL; ++f)
    for (int k = 0; k < m; ++k) grad[size_t(f) * m + k] = all[size_t(f) * SPL_MAX + k];
  return SIMOX_OK;
}

int simox_get_rho(simox_ctx* ctx, double* rho) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  return download_caller(ctx, ctx->rho, 1, rho);
}

int simox_pme_self(simox_ctx* ctx, double* eself) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  double e3[3];
  CHK(self_energy(ctx, ctx->q, 0, 2.0 * WPEN, e3));  // force.f90:486-498
  ctx->Eself = e3[0] * ctx->g.alpha / SQRTPI * KC + (e3[1] + e3[2]);
  if (eself) *eself = ctx->Eself;
  return SIMOX_OK;
}

int simox_pme(simox_ctx* ctx, double* erec) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  CHK(do_pme(ctx, true));
  if (erec) *erec = ctx->Erec;
  return SIMOX_OK;
}

int simox_phi(simox_ctx* ctx, double* phi) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  if (!ctx->matrix_valid) CHK(build_matrix(ctx));
  CHK(ensure_val32(ctx));
  CHK(grad_eval(ctx, ctx->q, nullptr, 0.0, true, use32(ctx)));
  return download_caller(ctx, ctx->phi, 1, phi);
}

int simox_dudq(simox_ctx* ctx, double* g) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  if (!ctx->matrix_valid) CHK(build_matrix(ctx));
  CHK(grad_eval(ctx, ctx->q, nullptr, 0.0, false, false));
  return download_caller(ctx, ctx->gv, 1, g);
}

int simox_get_neighbors(simox_ctx* ctx, int* counts, int* cols, int stride) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  if (!ctx->matrix_valid) CHK(build_matrix(ctx));
  const int n = ctx->n;
  std::vector<int> cnt(n), perm(n);
  HIPCHK(hipMemcpy(cnt.data(), ctx->nb_cnt, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(perm.data(), ctx->perm, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<int> col;
  if (cols) {
    col.resize(size_t(ctx->cap) * n);
    HIPCHK(hipMemcpy(col.data(), ctx->nb_col, sizeof(int) * col.size(), hipMemcpyDeviceToHost));
  }
  for (int p = 0; p < n; ++p) {
    const int c = perm[p];
    if (counts) counts[c] = cnt[p];
    if (cols)
      for (int k = 0; k < cnt[p] && k < stride; ++k) cols[size_t(c) * stride + k] = perm[col[size_t(k) * n + p]];
  }
  return SIMOX_OK;
}

int simox_ctip_solve(simox_ctx* ctx, double* eself, int* iters, int* n_eval) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  CHK(do_ctip(ctx, iters, n_eval));
  if (eself) *eself = ctx->Eself;
  return SIMOX_OK;
}

int simox_vv1(simox_ctx* ctx, double dt) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  CHK(do_vv1(ctx, dt));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return SIMOX_OK;
}

int simox_vv2(simox_ctx* ctx, double dt, double* mv2) {
  if (!ctx || !need_state(ctx)) return SIMOX_ERR_ARG;
  CHK(do_vv2(ctx, dt, true));
  if (mv2) *mv2 = ctx->mv2;
  return SIMOX_OK;
}

int simox_step(simox_ctx* ctx, int nsteps, double dt, int sort_every, int ctip_every, double* stats) {
  if (!ctx || !need_state(ctx) || nsteps < 0) return SIMOX_ERR_ARG;
  for (int s = 0; s < nsteps; ++s) {
    const bool last = s == nsteps - 1;
    CHK(do_vv1(ctx, dt));
    if (sort_every > 0 && s % sort_every == 0) CHK(do_cell_sort(ctx));
    if (ctip_every > 0 && s % ctip_every == 0) CHK(do_ctip(ctx, nullptr, nullptr));
    CHK(do_force(ctx, last && stats));
    CHK(do_pme(ctx, last && stats));
    CHK(do_vv2(ctx, dt, last && stats));
  }
  CHK(check_err(ctx, 0));
  if (stats) {
    stats[0] = ctx->Eself; stats[1] = ctx->Edir; stats[2] = ctx->Erec; stats[3] = ctx->Ectip;
    stats[4] = ctx->Epair; stats[5] = ctx->Egroup; stats[6] = ctx->mv2;
    stats[7] = ctx->Eself + ctx->Edir + ctx->Erec + ctx->Ectip + ctx->Epair + ctx->Egroup + ctx->mv2 / 2.0;
  }
  return SIMOX_OK;
}

int simox_batch_step(simox_ctx** ctxs, int nrep, int nsteps, double dt, int sort_every,
                     int ctip_every, double* stats) {
  if (!ctxs || nrep < 0 || nsteps < 0) return SIMOX_ERR_ARG;
  for (int r = 0; r < nrep; ++r)
    if (!ctxs[r] || !need_state(ctxs[r])) return SIMOX_ERR_ARG;
  for (int r = 0; r < nrep; ++r)
    if (ctxs[r]->mode == SIMOX_MODE_FAITHFUL) {  // host Brent: no interleaving, step in turn
      for (int k = 0; k < nrep; ++k) {
        const int rc = simox_step(ctxs[k], nsteps, dt, sort_every, ctip_every,
                                  stats ? stats + 8 * size_t(k) : nullptr);
        if (rc) return rc;
      }
      return SIMOX_OK;
    }
  struct GraphMode {  // graph replay of the inner iteration for this call only
    simox_ctx** c; int n;
    ~GraphMode() { for (int r = 0; r < n; ++r) c[r]->use_graph = false; }
  } gm{ctxs, nrep};
  for (int r = 0; r < nrep; ++r) ctxs[r]->use_graph = true;
  for (int s = 0; s < nsteps; ++s) {
    const bool last = s == nsteps - 1;
    for (int r = 0; r < nrep; ++r) {
      simox_ctx* ctx = ctxs[r];
      CHK(do_vv1(ctx, dt));
      if (sort_every > 0 && s % sort_every == 0) CHK(do_cell_sort(ctx));
    }
    if (ctip_every > 0 && s % ctip_every == 0) {
      for (int r = 0; r < nrep; ++r) {
        simox_ctx* ctx = ctxs[r];
        CHK(build_matrix(ctx));
        CHK(ncg_begin(ctx));
      }
      // round robin: wait for one context while the others' work runs
      for (bool busy = true; busy;) {
        busy = false;
        for (int r = 0; r < nrep; ++r) {
          simox_ctx* ctx = ctxs[r];
          if (ctx->ncg.phase == NCG_DONE) continue;
          HIPCHK(hipStreamSynchronize(ctx->st));
          CHK(ncg_check(ctx));
          busy = busy || ctx->ncg.phase != NCG_DONE;
        }
      }
      for (int r = 0; r < nrep; ++r) {
        simox_ctx* ctx = ctxs[r];
        CHK(ncg_end(ctx, nullptr, nullptr));
        double e3[3];
        CHK(self_energy(ctx, ctx->q, 1, WPEN, e3));  // ctip.f90:194-209
        ctx->Eself = e3[0] * KC * ctx->g.alpha / SQRTPI + (e3[1] + e3[2]);
      }
    }
    for (int r = 0; r < nrep; ++r) {
      simox_ctx* ctx = ctxs[r];
      CHK(do_force(ctx, last && stats));
      CHK(do_pme(ctx, last && stats));
      CHK(do_vv2(ctx, dt, last && stats));
    }
  }
  for (int r = 0; r < nrep; ++r) {
    simox_ctx* ctx = ctxs[r];
    CHK(check_err(ctx, 0));
    if (stats) {
      double* st = stats + 8 * size_t(r);
      st[0] = ctx->Eself; st[1] = ctx->Edir; st[2] = ctx->Erec; st[3] = ctx->Ectip;
      st[4] = ctx->Epair; st[5] = ctx->Egroup; st[6] = ctx->mv2;
      st[7] = st[0] + st[1] + st[2] + st[3] + st[4] + st[5] + st[6] / 2.0;
    }
  }
  return SIMOX_OK;
}

int simox_timing_enable(simox_ctx* ctx, int on) {
  if (!ctx) return SIMOX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  tresolve(ctx);
  ctx->timing = on != 0;
  for (int k = 0; k < 16; ++k) { ctx->tms[k] = 0; ctx->tcnt[k] = 0; }
  return SIMOX_OK;
}

int simox_timing_select(simox_ctx* ctx, unsigned int mask) {
  if (!ctx) return SIMOX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  tresolve(ctx);
  ctx->timing_mask = mask ? mask : ~0u;
  return SIMOX_OK;
}

int simox_kernel_timing(simox_ctx* ctx, int which, double* ms, long long* launches) {
  if (!ctx || which < 0 || which > 15) return SIMOX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  tresolve(ctx);
  if (ms) *ms = ctx->tms[which];
  if (launches) *launches = ctx->tcnt[which];
  return SIMOX_OK;
}

// CELL_SORT (cellsort.f90:53-80) without a context and without allocating:
// the kernels of simox_cell_sort on caller-owned device buffers plus the
// static scratch above.  Calls on one GPU share that scratch and race if
// they overlap.
int simox_cell_sort_noalloc(int n, const double box[3], double rcut, const double* x, int* npt,
                            int* start, int* link) {
  if (n < 1 || n > NOALLOC_MAX_ATOMS || !box || !x || !npt || !start || !link || !(rcut > 0.0))
    return SIMOX_ERR_ARG;
  Geom g{};
  g.n = n;
  for (int d = 0; d < 3; ++d) {
    if (!(box[d] > 0.0)) return SIMOX_ERR_ARG;
    g.nc[d] = int(box[d] / rcut);  // main.f90:189
    if (g.nc[d] < 2) return SIMOX_ERR_ARG;  // main.f90:191-193
    g.lcell[d] = box[d] / double(g.nc[d]);  // main.f90:190
    g.box[d] = box[d];
    g.hbox[d] = 0.5 * box[d];
  }
  const long long ncall = (long long)g.nc[0] * g.nc[1] * g.nc[2];
  if (ncall > NOALLOC_MAX_CELLS) return SIMOX_ERR_ARG;
  g.ncall = int(ncall);
  int* sc = nullptr;
  if (noalloc_scratch(&sc)) return SIMOX_ERR_HIP;
  int* cell = sc;
  int* slot = sc + NOALLOC_MAX_ATOMS;
  int* fill = sc + 2 * size_t(NOALLOC_MAX_ATOMS);
  int* err = fill + NOALLOC_MAX_CELLS;
  const int nb = grid_blocks(n);
  // counters start at zero on every call (fill kernels, no memset nodes)
  hipLaunchKernelGGL(k_fill_int, dim3(grid_blocks(g.ncall)), dim3(BLOCK), 0, nullptr, g.ncall, 0,
                     npt);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, nullptr, 1, 0, err);
  hipLaunchKernelGGL(k_cell_assign, dim3(nb), dim3(BLOCK), 0, nullptr, g, x, cell, npt);
  hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, nullptr, g.ncall, npt, start, fill, err);
  hipLaunchKernelGGL(k_cell_scatter, dim3(nb), dim3(BLOCK), 0, nullptr, n, cell, start, fill, slot);
  hipLaunchKernelGGL(k_cell_order, dim3(g.ncall), dim3(BLOCK), 0, nullptr, start, slot, nullptr,
                     link);
  if (hipGetLastError() != hipSuccess) return SIMOX_ERR_HIP;
  int flag = 0;
  if (hipMemcpy(&flag, err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return SIMOX_ERR_HIP;
  return flag ? flag : SIMOX_OK;
}

// Average duration of one launch of an inner-solve kernel, reps launches back
// to back between two events on the context stream, on the buffers the last
// converged solve left (control / inspection; no counterpart in the reference).
int simox_time_kernel(simox_ctx* ctx, int which, int reps, double* ms) {
  if (!ctx || !ms || reps < 1 || ctx->mode == SIMOX_MODE_FAITHFUL) return SIMOX_ERR_ARG;
  const bool pair = which == 7 || which == 8 || which == 9;
  if (pair ? !(ctx->nlist_valid && ctx->matrix_valid) : !ctx->matrix_valid) return SIMOX_ERR_ARG;
  if (which == 7 && ctx->rdf_on) return SIMOX_ERR_ARG;  // would add to the histogram
  if (pair) CHK(ensure_part(ctx, size_t(grid_blocks(ctx->n)) * 7));
  const bool f32 = use32(ctx);
  const PmeGrid* ig = (f32 && ctx->inner_coarse && ctx->pi.ready) ? &ctx->pi : nullptr;
  if (which == 11 && !(ig && ctx->pi.spread32 && ctx->pi.bricks_valid)) return SIMOX_ERR_ARG;
  if (which != 6 && which != 10 && which != 11 && !pair) return SIMOX_ERR_ARG;
  if (which == 6 && !(ig ? ctx->val16_valid : (f32 ? ctx->val32_valid : true))) return SIMOX_ERR_ARG;
  hipEvent_t a, b;
  HIPCHK(hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
  HIPCHK(hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
  auto launch = [&]() -> int {
    // the pair kernels rewrite what they wrote last (same list, forces, matrix)
    if (which == 7) launch_force(ctx, ctx->part + grid_blocks(ctx->n));
    else if (which == 8) launch_nlist(ctx);
    else if (which == 9) launch_ctip_matrix(ctx);
    else if (which == 6) launch_matvec(ctx, ig != nullptr);
    else if (which == 10) CHK(phi_gather(ctx, f32, ig));
    else spread_i32<4>(ctx, geom_of(ctx, ctx->pi), ctx->pi, ctx->pv);
    HIPCHK(hipGetLastError());
    return 0;
  };
  // the force pass overwrites f with the pair forces alone: keep the total
  double* fkeep = nullptr;
  const size_t fbytes = size_t(3) * ctx->n * sizeof(double);
  if (which == 7) {
    HIPCHK(hipMalloc((void**)&fkeep, fbytes));
    HIPCHK(hipMemcpyAsync(fkeep, ctx->f, fbytes, hipMemcpyDeviceToDevice, ctx->st));
  }
  CHK(launch());  // warm
  HIPCHK(hipEventRecord(a, ctx->st));
  if (which == 6) {
    // the matvec alone: events recorded at each launch's start and end
    std::vector<hipEvent_t> ke(2 * size_t(reps));
    for (auto& e : ke) HIPCHK(hipEventCreate(&e));
    for (int r = 0; r < reps; ++r) launch_matvec(ctx, ig != nullptr, ke[2 * r], ke[2 * r + 1]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(b, ctx->st));
    HIPCHK(hipEventSynchronize(b));
    double tot = 0.0;
    for (int r = 0; r < reps; ++r) {
      float t1 = 0.f;
      HIPCHK(hipEventElapsedTime(&t1, ke[2 * r], ke[2 * r + 1]));
      tot += t1;
    }
    for (auto& e : ke) (void)hipEventDestroy(e);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = tot / reps;
    return SIMOX_OK;
  }
  for (int r = 0; r < reps; ++r) CHK(launch());
  HIPCHK(hipEventRecord(b, ctx->st));
  if (fkeep) HIPCHK(hipMemcpyAsync(ctx->f, fkeep, fbytes, hipMemcpyDeviceToDevice, ctx->st));
  HIPCHK(hipEventSynchronize(b));
  if (fkeep) {
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(hipFree(fkeep));
  }
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = double(t) / reps;
  return SIMOX_OK;
}

}  // extern "C"
