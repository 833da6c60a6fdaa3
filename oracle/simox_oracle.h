/*
 * CPU oracle for the SIMOX per-step MD hot path: a plain-C FP64 restatement of
 * the reference Fortran (cell list, Zhou-Johnson-Wadley EAM, CTIP charge
 * equilibration with smooth-PME, velocity Verlet).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product library (simox_amd) never links it.
 *
 * Conventions follow the reference exactly: species ids Al=1 Ni=2 Fe=3 O=4 H=5
 * (datafmt.f90:29), positions in [-L/2, L/2) (cellsort.f90:63-69), velocities in
 * Angstrom per TFM femtoseconds (main.f90:142), arrays of 3-vectors stored
 * atom-major (x[3*i+d]).  Atom indices in cell link lists are 0-based here
 * (the reference is 1-based); their order is the reference's.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - orc_force, orc_cell_sort, orc_cell_hierarchy, orc_pme_self, orc_vv1/2,
 *     orc_dudq: against the reference Fortran compiled from its own sources
 *     (oracle/_ref, fixtures in tests/golden).
 *   - orc_fft3d: numpy.fft and a direct O(N^2) DFT.
 *   - orc_pme: brute-force Ewald sum (converged k-space), B-spline partition of
 *     unity, central finite differences of E_rec.
 *   - orc_ctip_solve: KKT conditions of the reference's own charge energy
 *     (dense numpy solve on small systems), finite differences of U(q) vs dUdq,
 *     charge conservation.  The reference's own CG/Brent cannot be built here
 *     (PHI_MESH needs the absent FFTE library), so the iterate sequence is
 *     pinned piecewise by closed forms (tests/test_oracle_ctip_pin.py):
 *     orc_brent on functions where one secant or inverse-quadratic step is
 *     exact, on a step function (bisection count fixes E_crit) and at the
 *     E_end exit; orc_ncg on 2- and 3-atom quadratic energies whose first
 *     iterates are known in exact arithmetic (beta = sum g / sum g_old, mean
 *     projection, the 1e-3 stop); and orc_ctip_solve equals orc_ncg on the
 *     pinned PHI_MESH + dUdq gradient bit for bit.
 */
#ifndef SIMOX_ORACLE_H
#define SIMOX_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NATOM 5
#define ORC_NMETAL 3
#define ORC_NPT_MAX 500  /* datafmt.f90:20 */
#define ORC_NNB_MAX 500  /* datafmt.f90:22 */

typedef struct { /* TYPE METL, datafmt.f90:65-68 */
  double r_e, f_e, rho_e, rho_s, alpha, beta, A, B, kappa, lambda;
  double Fn[4], Fm[4], Fp[4], eta, Fe;
} orc_metal;

typedef struct { /* TYPE OXID, datafmt.f90:70-74 */
  double r_e[5], alph[5], beta[5], A[5], B[5], kapp[5], lamb[5];
  double F[5][4]; /* F0..F4 */
  double re[5], rmax[5], fe, gamm, nu;
} orc_oxid;

typedef struct { /* TYPE POTN, datafmt.f90:54-56 */
  double xm, qmin, qmax, chi, J, xi, Z, fafb[5][4];
} orc_potn;

typedef struct {
  orc_metal eam[3];
  orc_oxid ox;
  orc_potn prm[5];
} orc_params;

typedef struct {
  int n;            /* Npt_all */
  double box[3];
  int ncell[3];
  double lcell[3];
  int fft[3];
  double rcut, rcut2, a; /* a = Ewald alpha = 3.8/Rcut (force.f90:481) */
} orc_sys;

/* status codes shared with the product library's error mapping */
#define ORC_OK 0
#define ORC_ERR_CELL_OVERFLOW 1   /* cellsort.f90:75 */
#define ORC_ERR_NB_OVERFLOW 2     /* force.f90:79 / ctip.f90:153 */
#define ORC_ERR_UNKNOWN_PAIR 3    /* force.f90:73 */
#define ORC_ERR_BRENT 4           /* ctip.f90:302 */

void orc_potential_setup(orc_params* p);
void orc_fafb(orc_params* p); /* recompute fafb from xi, datafmt.f90:201-220 */
void orc_sys_init(orc_sys* s, int n, const double box[3], const int fft[3], double rcut);

void orc_cell_hierarchy(const int nc[3], int* pair /* ncell_all x 13, 0-based */);
int orc_cell_sort(const orc_sys* s, const double* x, int* npt, int* link /* ncell_all x 500 */);

void orc_eam_pair(const orc_params* p, int id_i, int id_j, double r,
                  double fj[2], double dfjdr[2], double* phi, double* df);
int orc_embed(const orc_params* p, int id, double rho, double* F, double* dF);

/* FORCE (force.f90:3-361).  f is overwritten.  en = {Epair, Edir, Ectip, Egroup}.
   nlist/list (optional, may be NULL): neighbour lists as FORCE builds them.
   rho (optional): EAM densities. */
int orc_force(const orc_sys* s, const orc_params* p, const int* id, const double* x,
              const double* q, const int* npt, const int* link, const int* pair,
              double* f, double* en, int* nlist, int* list, double* rho);

/* PME_SELF (force.f90:467-517): returns Eself, fills psi (fft1*fft2*fft3, x fastest). */
double orc_pme_self(const orc_sys* s, const orc_params* p, const int* id, const double* q,
                    double* psi);
/* 3D complex FFT in place, FFTE ZFFT3D convention: sign=-1 forward (unscaled),
   sign=+1 inverse scaled by 1/(nx*ny*nz).  data = interleaved re,im, x fastest. */
void orc_fft3d(double* data, int nx, int ny, int nz, int sign);
/* 1 (default): FP32-rounded spreading weights as the reference (quirk Q1);
   0: exact FP64 weights. */
void orc_set_spread_rounding(int on);
/* Build-only split cutoff (no reference counterpart; the reference uses one
   Rcut for EAM and Ewald, quirk Q9): EAM density, pair and embedding terms only
   for r < r_eam; Ewald real space and CTIP keep Rcut.  r_eam <= 0: off. */
void orc_set_eam_cutoff(double r_eam);
/* PME (force.f90:364-464): f += reciprocal forces (mean removed), returns Erec. */
double orc_pme(const orc_sys* s, const int* id, const double* x, const double* q,
               const double* psi, double* f);
/* PHI_MESH (ctip.f90:336-396) on q: phi (unscaled by c_pme). */
void orc_phi_mesh(const orc_sys* s, const double* x, const double* q, const double* psi,
                  double* phi);

/* CTIP_SOLVER (ctip.f90:3-212): q in/out, returns status; *eself, *iters, *n_eval out. */
int orc_ctip_solve(const orc_sys* s, const orc_params* p, const int* id, const double* x,
                   const int* npt, const int* link, const int* pair, const double* psi,
                   double* q, double* eself, int* iters, int* n_eval);
/* The two pieces of that solve on callbacks, so that they can be driven by
   closed-form functions: BRENT_SOLVER's scalar root search (ctip.f90:216-304)
   of f on [-1, 1], and the nonlinear CG loop (ctip.f90:167-190) on a
   gradient g(q) = -dU/dq; maxit > 0 caps the iterations (0: the reference's
   1e-3 stop only). */
typedef double (*orc_line_fn)(void* ctx, double alpha);
typedef void (*orc_grad_fn)(void* ctx, const double* q, double* g);
int orc_brent(orc_line_fn f, void* ctx, double* alpha, int* n_eval);
int orc_ncg(int n, orc_grad_fn grad, void* ctx, double* q, int maxit, int* iters, int* n_eval);
/* CTIP sparse matrices as CTIP_SOLVER builds them (ctip.f90:61-165). */
int orc_ctip_matrix(const orc_sys* s, const orc_params* p, const int* id, const double* x,
                    const int* npt, const int* link, const int* pair,
                    int* nlist, int* list, double* asym, double* sym);
/* dUdq (ctip.f90:307-333); xjz = N x 3 (chi, J - 2 kc a/sqrt(pi), Z), qmx = N x 2. */
void orc_dudq(int n, const double* xjz, const double* asym, const double* sym,
              const int* list, const int* nlist, const double* q, const double* qmx,
              const double* phi, double* g);
/* Charge energy U(q) whose gradient dUdq returns (negated), ctip.f90:194-209 plus the
   PME and sparse quadratic terms; used to pin dUdq by finite differences. */
double orc_ctip_energy(const orc_sys* s, const orc_params* p, const int* id, const double* x,
                       const int* npt, const int* link, const int* pair, const double* psi,
                       const double* q);

/* VV_NVE1 (vverlet.f90:74-96) and VV_NVE2 (vverlet.f90:132-157); returns mv2. */
void orc_vv1(const orc_sys* s, const orc_params* p, const int* id, double* x, double* v,
             const double* f, double dt);
double orc_vv2(const orc_sys* s, const orc_params* p, const int* id, double* v,
               const double* f, double dt);

#ifdef __cplusplus
}
#endif
#endif
