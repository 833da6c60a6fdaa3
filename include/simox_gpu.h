/*
 * simox_gpu.h - C-ABI of the MI355X (gfx950) implementation of the SIMOX
 * per-step MD hot path: cell list, EAM density/embedding/pair forces,
 * real-space Ewald + CTIP short-range terms, smooth-PME reciprocal forces and
 * potential, CTIP charge equilibration, velocity Verlet.
 *
 * Every entry point takes plain pointers and sizes, returns an int status
 * (SIMOX_OK = 0) and is synchronous at return unless stated otherwise.  On a
 * nonzero status simox_last_error() returns the text the reference would have
 * printed before its STOP.  One host thread and one HIP stream per context.
 *
 * Host arrays use the reference's conventions: atoms in the caller's order,
 * species ids 1..5 (Al, Ni, Fe, O, H; datafmt.f90:29-30), 3-vectors stored
 * atom-major (x[3*i+d], the layout of TYPE PTCL's xx/xv/ff components,
 * datafmt.f90:49-52), positions in [-L/2, L/2), velocities in Angstrom per
 * TFM fs, forces in eV/Angstrom, charges in e.
 *
 * Scope: Al, Ni, Fe homo-metal pairs, metal-O and O-O pairs (eam.f90:3-35,
 * 171-328).  Ni/Fe cross pairs (METAL_HETERO, AL_NI) and H return
 * SIMOX_ERR_UNKNOWN_PAIR.
 */
#ifndef SIMOX_GPU_H
#define SIMOX_GPU_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes; texts follow the reference's STOP messages ---------- */
#define SIMOX_OK 0
#define SIMOX_ERR_CELL_OVERFLOW 1 /* "=== Cell particle limit crashed !!! ===" cellsort.f90:75 */
#define SIMOX_ERR_NB_OVERFLOW 2   /* "=== Memeory error for EAM density ===" force.f90:79,
                                     "=== neighbor particle limit crashed !!! === " ctip.f90:153 */
#define SIMOX_ERR_UNKNOWN_PAIR 3  /* "=== Unknown particle pair ===" force.f90:73 */
#define SIMOX_ERR_BRENT 4         /* "# **** Brent solver limit breached *** " ctip.f90:302 */
#define SIMOX_ERR_ARG 5           /* invalid argument / geometry (main.f90:191-193) */
#define SIMOX_ERR_HIP 6           /* HIP runtime error */
#define SIMOX_ERR_NOT_CONVERGED 7 /* linear charge solve hit its iteration limit */
#define SIMOX_ERR_FORCE_RANGE 8   /* a pair force component of 16384 eV/A or more: beyond the
                                     range of the 64-bit fixed-point force sums (no reference
                                     counterpart; the reference has no such limit) */

/* ---- numerical modes ----------------------------------------------------- */
/* FP64, reproducing the reference's arithmetic including its quirks: FP32-
   rounded PME spreading weights (force.f90:406, ctip.f90:359) and the
   nonlinear CG + Brent charge solve with its 1e-3 stop (ctip.f90:167-304). */
#define SIMOX_MODE_FAITHFUL 0
/* FP64 kernels with a converged charge solve: preconditioned nonlinear CG with
   exact line searches on the same (piecewise-quadratic) charge energy, to an
   RMS projected-gradient tolerance (simox_set_charge_tol, default 1e-10 eV/e). */
#define SIMOX_MODE_FP64 1
/* FP32 pair / PME / sparse arithmetic with FP32 partial sums; FP64 positions,
   velocities, Verlet update, CG residuals, dot products, neutrality multiplier
   and energy totals.  Charge solve = iterative refinement: nonlinear CG with the
   FP32 operator, restarted from the FP64 true gradient. */
#define SIMOX_MODE_MIXED 2

/* ---- potential parameters: TYPE METL / OXID / POTN (datafmt.f90:54-74) --- */
typedef struct {
  double r_e, f_e, rho_e, rho_s, alpha, beta, A, B, kappa, lambda;
  double Fn[4], Fm[4], Fp[4], eta, Fe;
} simox_metal;

typedef struct {
  double r_e[5], alph[5], beta[5], A[5], B[5], kapp[5], lamb[5];
  double F[5][4]; /* F0..F4 */
  double re[5], rmax[5], fe, gamm, nu;
} simox_oxid;

typedef struct {
  double xm, qmin, qmax, chi, J, xi, Z, fafb[5][4];
} simox_potn;

typedef struct {
  simox_metal eam[3];
  simox_oxid ox;
  simox_potn prm[5];
} simox_params;

/* POTENTIAL_SETUP (datafmt.f90:142-262): fills the published Zhou-Johnson-
   Wadley EAM, oxygen EAM and Streitz-Mintmire CTIP parameters. */
void simox_potential_setup(simox_params* p);
/* Recomputes the Slater Coulomb-integral coefficients fafb from xi
   (datafmt.f90:201-220); call after changing any xi. */
void simox_fafb(simox_params* p);

typedef struct simox_ctx simox_ctx;

/* Context = one system: periodic box, PME grid, cutoff, atom count, mode.
   Replaces READ_INPUT's geometry set-up (main.f90:137-201): Ncell =
   INT(box/rcut), Lcell = box/Ncell, Ewald alpha = 3.8/rcut.  PME grid sizes
   must be multiples of 4 with prime factors 2, 3, 5 only.  params may be
   NULL (published parameters).  Allocates all device buffers; nothing is
   allocated by later calls except when a neighbour table must grow. */
int simox_create(int n, const double box[3], const int fft[3], double rcut, int mode,
                 const simox_params* params, simox_ctx** out);
int simox_destroy(simox_ctx* ctx);
const char* simox_last_error(const simox_ctx* ctx);
/* Replaces the potential parameters of an existing context (the caller's
   POTENTIAL_SETUP output, datafmt.f90:142-262, possibly edited) without
   touching atoms or cells: the drop-in FORCE / CTIP_SOLVER pass their eam,
   oxygn, param arguments through here (force.f90:3, ctip.f90:3). */
int simox_set_params(simox_ctx* ctx, const simox_params* params);
/* Split cutoff (build-only; the reference has one Rcut for EAM, Ewald and
   CTIP, main.f90:172-174, quirk Q9): EAM density, pair and embedding terms
   only for r < r_eam <= Rcut; real-space Ewald, CTIP and the cells keep Rcut.
   Default r_eam = Rcut. */
int simox_set_eam_cutoff(simox_ctx* ctx, double r_eam);
/* Thermostat of simox_vv1 / simox_vv2 / simox_step (main.f90:44-58):
   kind 0 = NVE (VV_NVE1/2), 1 = Berendsen (VV_NVE1 + VV_NVT2_BEREND,
   vverlet.f90:160-208, coupling 0.025), 2 = Langevin (VV_NVT1_STOCH +
   VV_NVT2_STOCH, vverlet.f90:237-296, friction alpha; alpha and the noise
   amplitude in single precision as there).  t_kelvin = T_given (K).  The
   Langevin noise is N(0,1) from a counter-based generator keyed by (seed,
   step, atom, component); the reference draws from the unseeded RANDOM_NUMBER,
   so the noise sequence is not the reference's, its distribution is. */
int simox_set_thermostat(simox_ctx* ctx, int kind, double t_kelvin, double alpha,
                         unsigned long long seed);
/* Pair histogram of FORCE (pair%rdf, force.f90:55,76,159,180; Nrdf = 200
   bins of Rcut/200).  While on, every simox_force / simox_step adds one
   count per direction of each listed pair to bin INT(r/dr).  simox_get_rdf
   returns hist[(bin-1)*25 + (sj-1)*5 + (si-1)], bin = 1..201 (the layout of
   rdf(5,5,201)); hist[i,j] + hist[j,i] (i != j) and hist[i,i] equal the
   reference's rdf(i,j) + rdf(j,i) and rdf(i,i), the sums PRINT_RDF writes
   (util.f90:270-277).  reset != 0 zeroes it after the copy. */
int simox_set_rdf(simox_ctx* ctx, int on);
/* OXIDATION mode (main.f90:62-86): on != 0 makes simox_vv1 / simox_step use
   VV_NVE1_REFLECT (vverlet.f90:100-129: x, y periodic, reflecting wall at
   z = box_z/2).  simox_remove_rigid_motion = REMOVE_RIGID_MOTION
   (vverlet.f90:211-234): subtracts the mean velocity and the mean z. */
int simox_set_reflect(simox_ctx* ctx, int on);
/* Build-only parameter gradients (the reference has none): at the converged
   charges of the last simox_ctip_solve (modes FP64, MIXED), the derivatives of
   E = Eself + Edir + Erec + Ectip (the energies the step reports) with respect
   to chi_s and J_s of the five species, through the charge equilibration by
   an adjoint solve (H lambda = P dE/dq, sum lambda = 0, FP64 PCG to an RMS
   residual of 1e-12):  grad[s] = dE/dchi_{s+1}, grad[5+s] = dE/dJ_{s+1}.
   lambda (caller order, may be NULL) receives the adjoint; *iters its CG
   iterations. */
int simox_param_grad(simox_ctx* ctx, double grad[10], double* lambda, int* iters);
/* As simox_param_grad, plus grad[10+s] = dE/dxi_{s+1}, the Slater exponent
   (zeta) of each species, with the [fa|fb] coefficients (datafmt.f90:201-220)
   following it: kc dEctip/dxi at fixed charges minus lambda . dG/dxi of the
   solver's matrix, by forward-mode dual numbers through the same formulas. */
int simox_param_grad_all(simox_ctx* ctx, double grad[15], double* lambda, int* iters);
/* Force term of a fitting loss: grad[s], grad[5+s], grad[10+s] = d(sum_i
   u_i . F_i)/dchi, /dJ, /dxi of species s+1, F the forces of FORCE + PME at
   the converged charges (which follow the parameters: adjoint solve with the
   right-hand side d(sum u.F)/dq).  u: caller order, 3 per atom; NULL uses the
   current forces, so that grad of simox_param_grad_all + 2 w this is the
   gradient of E + w sum |F|^2. */
int simox_param_force_grad(simox_ctx* ctx, const double* u, double grad[15], int* iters);
int simox_remove_rigid_motion(simox_ctx* ctx);
int simox_get_rdf(simox_ctx* ctx, long long* hist, int reset);
/* FFT launch shapes.  simox_create picks them by timing the candidate x
   (R2C + C2R) and y tile sizes on the context's own grid (environment
   SIMOX_AUTOTUNE=0 keeps the default rule); simox_autotune re-runs that
   search.  simox_fft_shapes reports {mx, my, mz} of the FP64 chain, then of
   the FP32 chain (zeros outside MIXED mode).  Shapes change no result. */
int simox_autotune(simox_ctx* ctx);
int simox_fft_shapes(const simox_ctx* ctx, int shapes[6]);
/* Launch shapes of the MIXED inner solve with the coarse inner operator,
   which the context times on its own buffers at its first solve (environment
   SIMOX_AUTOTUNE=0 keeps the defaults; SIMOX_MV_VARIANT, SIMOX_GATHER_SHAPE,
   SIMOX_FFT_XYNT and SIMOX_FFT_NL fix one each): {packed matvec variant (0-10:
   4 or 8 groups of 4 entries in flight per lane, odd: streaming loads,
   4-7: the same with the row blocks in contiguous runs per XCD, 8 and 9:
   4 and 8 groups, streaming, gathering from each row block's neighbourhood
   of the direction staged in LDS, 10: as 8 with 3-byte entries),
   record gather tile (0-5: 64x8x8, 32x8x8, 32x16x8, 64x16x8, 32x16x16,
   64x16x16), inner grid's x-y pass threads per workgroup and z-pass lines
   per workgroup, the same two of the preconditioner's grid (0: no plane
   chain)}.  Every choice gives the same results bit for bit.  Inspection
   only (no counterpart in the reference). */
int simox_inner_shapes(const simox_ctx* ctx, int shapes[6]);
/* Cell-tile shape of the neighbour-list build {atoms staged per LDS chunk,
   threads per workgroup}, which the context times on its first list build
   (SIMOX_AUTOTUNE=0 keeps the default {1024, 128}; SIMOX_TILE_SHAPE = 0-5
   fixes one of {1024,128}, {512,128}, {2048,128}, {1024,64}, {1024,256},
   {512,256}).  Every shape gives the same list.  Inspection only (no
   counterpart in the reference). */
int simox_tile_shape(const simox_ctx* ctx, int shape[2]);
int simox_get_geometry(const simox_ctx* ctx, int ncell[3], double lcell[3], double* alpha);
int simox_set_mode(simox_ctx* ctx, int mode);
/* Tolerance (RMS projected gradient, eV/e) and iteration cap of the converged
   charge solve (modes FP64, MIXED). */
int simox_set_charge_tol(simox_ctx* ctx, double tol, int maxit);
/* Preconditioner of the converged charge solve: 0 = Jacobi (the default in
   FP64 mode, the plain method), 1 = long-range Woodbury preconditioner (the
   default in MIXED mode): M^-1 = D^-1 - D^-1 W^T F W D^-1 with W the B-spline
   spread onto a PME grid of half the resolution and F = Gc / (1 + Gc rho) the
   Coulomb kernel screened by the density rho of 1/J.  One extra spread, FFT
   pair and gather on the coarse grid per iteration; about half the iterations
   for 1M-atom slabs (the Jacobi count grows with the box, this one does not). */
int simox_set_charge_precond(simox_ctx* ctx, int on);
/* Reciprocal part of the MIXED mode's inner operator on a coarse grid: half
   the resolution of the context's grid in every dimension, order-4
   B-splines, its own influence table, and the sparse part from 12-bit
   matrix values packed with their 20-bit columns in one 32-bit word (up to
   2^20 atoms; FP16 values with 4-byte columns above) without the entries
   below 1e-3 eV/e^2 (the far real-space Coulomb tail; about a quarter of
   the entries stay) (on = 1), or on the context's
   grid with all FP32 values (on = 0).  The
   inner solves then stop at 3e-2 (else 1e-4) of their restart's residual;
   every restart evaluates the true gradient with the context's grid and the
   FP64 matrix, so the converged charges meet the tolerance for the
   reference's operator either way.  Default: on in MIXED from 2^23 grid
   points (C3, 256 x 256 x 512: 361 -> 148 ms per step), off below (C2, 96 x
   96 x 192, is slower with it).  SIMOX_ERR_ARG outside MIXED. */
int simox_set_inner_grid(simox_ctx* ctx, int on);
/* Entries of the matrix the inner solves' sparse part reads: the compacted
   matrix of the coarse inner operator (entries with |H0_ij| >= 1e-3
   eV/e^2; SIMOX_INNER_DROP sets the bound, 0 keeps all), else the full list.
   Valid after a charge solve. */
int simox_inner_entries(simox_ctx* ctx, long long* entries);
/* Iterations and FP64 restarts of the last converged charge solve. */
int simox_charge_stats(simox_ctx* ctx, int* iters, int* restarts);

/* Host <-> device state of TYPE PTCL q(:) (datafmt.f90:49-52), caller order.
   Any pointer may be NULL to leave that component untouched.  set_state with
   id != NULL also re-sorts the atoms into cells (CELL_SORT). */
int simox_set_state(simox_ctx* ctx, const int* id, const double* x, const double* v,
                    const double* f, const double* q);
int simox_get_state(simox_ctx* ctx, double* x, double* v, double* f, double* q);

/* CELL_SORT (cellsort.f90:53-80): bins atoms into Ncell cells; within a cell
   atoms keep ascending caller index.  Replaces cell(:)%Npt/link. */
int simox_cell_sort(simox_ctx* ctx);
/* Cell occupancy (Npt per cell, 0-based cell id = ix + iy*Nx + iz*Nx*Ny) and
   members in link order (caller indices, 0-based); link may be NULL. */
int simox_get_cells(simox_ctx* ctx, int* npt, int* link, int link_stride);
/* No counterpart in the reference: CELL_SORT for callers that cannot allocate
   device memory (code under HIP graph capture, serving loops, harnesses that
   forbid hipMalloc).  Same binning and link order as simox_cell_sort +
   simox_get_cells, without a context: x = device pointer to 3n doubles in
   caller order; box, rcut on the host (Ncell = INT(box/rcut) >= 2 per axis);
   outputs are caller-owned device buffers: npt[ncell] atoms per cell (cell id
   ix + iy*Nx + iz*Nx*Ny), start[ncell+1] their exclusive prefix sums, link[n]
   the caller indices (0-based) in cell order, ascending within a cell.
   Allocates nothing: its scratch is a static device array of 16.5 MiB per GPU
   (held until the process ends), so n <= 2^21 and ncell <= 2^17, else
   SIMOX_ERR_ARG.  Synchronous (null stream).  Calls on one GPU share that
   scratch: calls that overlap on one GPU race. */
int simox_cell_sort_noalloc(int n, const double box[3], double rcut, const double* x, int* npt,
                            int* start, int* link);

/* FORCE (force.f90:3-361): overwrites forces with EAM + real-space Ewald +
   CTIP short-range forces.  energies[4] = Epair, Edir, Ectip, Egroup (eV),
   already multiplied by kc as sys%Edir/Ectip are (force.f90:250-252). */
int simox_force(simox_ctx* ctx, double energies[4]);
/* Per-atom EAM density rho_i (caller order) from the last simox_force. */
int simox_get_rho(simox_ctx* ctx, double* rho);

/* PME_SELF (force.f90:467-517): Eself of the current charges (penalty 2w,
   Q4) and the influence function; the k=0 term is zero. */
int simox_pme_self(simox_ctx* ctx, double* eself);
/* PME (force.f90:364-464): adds reciprocal-space forces (mean removed) to
   the forces; *erec = E_rec (eV). */
int simox_pme(simox_ctx* ctx, double* erec);
/* PHI_MESH (ctip.f90:336-396) for the current charges, scaled by
   c_pme = kc*Ng/(pi V) as CTIP_SOLVER scales it (ctip.f90:59,170). */
int simox_phi(simox_ctx* ctx, double* phi);

/* CTIP_SOLVER (ctip.f90:3-212): re-equilibrates the charges in place and
   sets *eself as CTIP_SOLVER does (penalty w, Q4).  *iters = CG iterations,
   *n_eval = gradient (PME + sparse matvec) evaluations.  Mode FAITHFUL runs
   the reference's nonlinear CG + Brent; FP64/MIXED run the converged solve. */
int simox_ctip_solve(simox_ctx* ctx, double* eself, int* iters, int* n_eval);
/* dUdq (ctip.f90:307-333) at the current charges: g = -dU/dq (caller order). */
int simox_dudq(simox_ctx* ctx, double* g);
/* Neighbour rows of the CTIP matrix (full shell, caller indices, per-atom
   order of the build's enumeration): counts[n]; cols[n*stride] (may be NULL). */
int simox_get_neighbors(simox_ctx* ctx, int* counts, int* cols, int stride);

/* VV_NVE1 (vverlet.f90:74-96) and VV_NVE2 (vverlet.f90:132-157); dt in TFM
   units; *mv2 = sum m v^2 as sys%mv2. */
int simox_vv1(simox_ctx* ctx, double dt);
int simox_vv2(simox_ctx* ctx, double dt, double* mv2);

/* One or more REGULAR NVE steps, device-resident (main.f90:41-61):
   VV_NVE1 -> [CELL_SORT] -> [CTIP_SOLVER] -> FORCE -> PME -> VV_NVE2.
   sort_every / ctip_every: run the sort / the charge solve when
   step % every == 0 (counting from 0 for this call; 0 = never).
   stats (may be NULL): 8 doubles of the last step =
   Eself, Edir, Erec, Ectip, Epair, Egroup, mv2, E_total (util.f90:12-15). */
int simox_step(simox_ctx* ctx, int nsteps, double dt, int sort_every, int ctip_every,
               double* stats);

/* EAM in spline form (build-only; no counterpart in the reference, whose EAM
   is analytic).  Eight uniform cubic B-splines of m coefficients each
   (4 <= m <= 64), coef[f*m + k], function f valid on [lo[f], lo[f] + (m-3) h[f]]:
   f = 0 density of an Al source, 1 density of an O source into Al, 2 into O,
   3 pair term Al-Al, 4 Al-O, 5 O-O, 6 embedding F of Al, 7 of O.  They replace
   the ZJW forms in FORCE's density, embedding and pair terms (Al/O systems
   only: other species return SIMOX_ERR_UNKNOWN_PAIR).  coef = NULL switches
   back to the analytic forms. */
int simox_set_eam_splines(simox_ctx* ctx, int m, const double* lo, const double* h,
                          const double* coef);
/* grad[f*m + k] = dE/dcoef[f*m + k] of E = Epair + Egroup (the EAM energy of
   FORCE) at the current positions, spline form set. */
int simox_eam_spline_grad(simox_ctx* ctx, double* grad);
/* grad[f*m + k] = dS/dcoef[f*m + k] for S = sum_i u_i . F_i, F the EAM part
   of the forces at the current positions (spline form set): the vector-
   Jacobian product of the forces.  u: caller order, 3 per atom; NULL uses the
   forces the last simox_force / simox_pme left (the total forces), so that
   dE/dcoef + 2 w grad is the gradient of E + w sum |F|^2. */
int simox_eam_spline_force_vjp(simox_ctx* ctx, const double* u, double* grad);

/* Batched replicas (build-only): nrep independent contexts (each its own box,
   grid, atoms and HIP stream) advanced by the same simox_step sequence in
   lockstep.  The converged charge solves of all replicas proceed together:
   the host waits on one replica's stream while the others' work runs, and
   each inner iteration is replayed from a captured HIP graph.  Results equal
   simox_step on each context.  stats (may be NULL): 8 doubles per replica as
   simox_step.  FAITHFUL contexts are stepped one after another. */
int simox_batch_step(simox_ctx** ctxs, int nrep, int nsteps, double dt, int sort_every,
                     int ctip_every, double* stats);

/* Per-kernel timing (HIP events recorded on the context stream around each
   launch while enabled): which = 0 spread, 1 FFT x R2C, 2 FFT y (both
   directions), 3 FFT z + influence multiply, 4 FFT x C2R, 5 PME force gather,
   6 dUdq (faithful mode: potential gather + sparse matvec; converged modes:
   the sparse matvec H0 p), 7 pair force, 8 EAM density + embedding, 9 CTIP
   matrix build, 10 tiled potential gather of the converged solve (on the
   coarse grid when the MIXED inner operator uses it), 11-15 spread, FFT x
   R2C, y, z and x C2R on that coarse grid.  Returns
   accumulated ms and launch count since the last simox_timing_enable. */
int simox_kernel_timing(simox_ctx* ctx, int which, double* ms, long long* launches);
int simox_timing_enable(simox_ctx* ctx, int on);
/* Categories timed while timing is enabled: bit `which` of mask (0: all, the
   default).  Each timed launch costs two event records on the stream, so a
   throughput measurement times only the kernel it needs. */
int simox_timing_select(simox_ctx* ctx, unsigned int mask);
/* Average duration (ms) of one launch of an inner-solve kernel, reps launches
   back to back between two HIP events on the context stream, on the buffers
   the last converged charge solve left: which = 6 the sparse matvec H0 p,
   10 the potential gather, 11 the spread onto the coarse inner grid (MIXED
   with the coarse inner operator); or of a pair kernel on the current
   positions, which rewrites what it wrote last: 7 the force pass (with its
   FP64 near-pair pass in MIXED; not while the RDF is on), 8 the neighbour
   list + density pass, 9 the CTIP matrix.  Events recorded around single launches
   inside a step (simox_kernel_timing) also count part of the preceding
   launch; this measures the kernel alone, as a kernel trace does.  Control /
   inspection only (no counterpart in the reference). */
int simox_time_kernel(simox_ctx* ctx, int which, int reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
