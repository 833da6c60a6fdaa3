/*
 * CPU oracle: FP64 restatement of the SIMOX hot path.  TEST INFRASTRUCTURE
 * ONLY (see simox_oracle.h for who may load it and what pins each function).
 *
 * Every routine follows one reference routine; loop order, operation order and
 * the reference's quirks are kept so that results agree with the reference
 * build to rounding.  Citations are file:line in the SIMOX sources.
 */
#include "simox_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* datafmt.f90:23-28 */
static const double KC = 14.399644154020082;
static const double PI_ = 3.1415926535897931;
static const double SQRTPI = 1.7724538509055159;
static const double WPEN = 20.0;

/* ------------------------------------------------------------------------ */
/* POTENTIAL_SETUP, datafmt.f90:142-262                                       */
/* ------------------------------------------------------------------------ */
static void set_metal(orc_metal* m, double r_e, double f_e, double rho_e, double rho_s,
                      double alpha, double beta, double A, double B, double kappa,
                      double lambda, const double Fn[4], const double Fm[4],
                      const double Fp[4], double eta, double Fe) {
  m->r_e = r_e; m->f_e = f_e; m->rho_e = rho_e; m->rho_s = rho_s;
  m->alpha = alpha; m->beta = beta; m->A = A; m->B = B; m->kappa = kappa; m->lambda = lambda;
  memcpy(m->Fn, Fn, sizeof m->Fn); memcpy(m->Fm, Fm, sizeof m->Fm); memcpy(m->Fp, Fp, sizeof m->Fp);
  m->eta = eta; m->Fe = Fe;
}

void orc_fafb(orc_params* p) {
  /* datafmt.f90:201-220 */
  for (int i = 0; i < ORC_NATOM; ++i) {
    for (int j = 0; j < ORC_NATOM; ++j) {
      double* c = p->prm[i].fafb[j];
      if (i == j) {
        double xi = p->prm[i].xi;
        c[0] = 1.0;
        c[1] = 11.0 * xi / 8.0;
        c[2] = 0.75 * (xi * xi);
        c[3] = (xi * xi * xi) / 6.0;
      } else {
        double a = p->prm[i].xi, b = p->prm[j].xi;
        double ab = a + b, amb = a - b, bma = b - a;
        c[0] = a * (b * b * b * b) / (ab * ab) / (amb * amb);
        c[1] = b * (a * a * a * a) / (ab * ab) / (bma * bma);
        c[2] = (3.0 * (a * a) * (b * b * b * b) - (b * b * b * b * b * b)) / (ab * ab * ab) /
               (amb * amb * amb);
        c[3] = (3.0 * (b * b) * (a * a * a * a) - (a * a * a * a * a * a)) / (ab * ab * ab) /
               (bma * bma * bma);
      }
    }
  }
}

void orc_potential_setup(orc_params* p) {
  memset(p, 0, sizeof *p);
  { const double Fn[4] = {-2.80760, -0.30144, 1.25856, -1.24760};
    const double Fm[4] = {-2.83000, 0.0, 0.62225, -2.48824};
    const double Fp[4] = {-2.83000, 0.0, 0.62225, -2.48824};
    set_metal(&p->eam[0], 2.86392, 1.20378, 17.51747, 19.90041, 6.61317, 3.52702, 0.31487,
              0.36555, 0.37985, 0.75969, Fn, Fm, Fp, 0.78591, -2.82453); }
  { const double Fn[4] = {-2.69351, -0.07644, 0.24144, -2.37563};
    const double Fm[4] = {-2.70000, 0.0, 0.26539, -0.15286};
    const double Fp[4] = {-2.70000, 0.0, 0.26539, 4.58568};
    set_metal(&p->eam[1], 2.48875, 2.21149, 30.37003, 30.37137, 8.38345, 4.47117, 0.42905,
              0.63353, 0.44360, 0.82066, Fn, Fm, Fp, 1.01318, -2.70839); }
  { const double Fn[4] = {-2.53499, -0.05960, 0.19306, -2.28232};
    const double Fm[4] = {-2.54000, 0.0, 0.20027, -0.14877};
    const double Fp[4] = {-2.54000, 0.0, 0.20027, 6.69465};
    set_metal(&p->eam[2], 2.48199, 2.31453, 24.59573, 24.59573, 9.81827, 5.23641, 0.39281,
              0.64624, 0.17031, 0.34061, Fn, Fm, Fp, 1.18290, -2.55187); }
  /* CTIP charge data, datafmt.f90:182-197 */
  const double xm[5] = {26.981539, 58.6934, 55.84500, 15.9994, 1.00794};
  const double qmin[5] = {0.0, 0.0, 0.0, -2.0, 0.0};
  const double qmax[5] = {3.0, 2.0, 3.0, 0.0, 0.0};
  const double chi[5] = {-1.47914, -1.70804, -1.90587, 2.000, 0.0};
  const double J[5] = {9.07222, 9.10954, 8.99819, 14.99523, 0.0};
  const double xi[5] = {0.968, 1.087, 1.024, 2.144, 0.0};
  const double Z[5] = {1.07514, 1.44450, 1.28612, 0.0, 0.0};
  for (int i = 0; i < 5; ++i) {
    p->prm[i].xm = xm[i]; p->prm[i].qmin = qmin[i]; p->prm[i].qmax = qmax[i];
    p->prm[i].chi = chi[i]; p->prm[i].J = J[i]; p->prm[i].xi = xi[i]; p->prm[i].Z = Z[i];
  }
  orc_fafb(p);
  /* oxygen EAM, datafmt.f90:225-259 */
  orc_oxid* o = &p->ox;
  const double r_e[5] = {2.98520, 2.95732, 3.07992, 3.64857, 1.0};
  const double alph[5] = {8.49741, 7.96528, 7.52309, 5.44072, 0.0};
  const double beta[5] = {4.52114, 4.42411, 4.13330, 3.59746, 0.0};
  const double A[5] = {0.09738, 0.13521, 0.17108, 0.34900, 0.0};
  const double B[5] = {0.38121, 0.25332, 0.39869, 0.57438, 0.0};
  const double kapp[5] = {0.18967, 0.47077, 0.22335, 0.08007, 0.0};
  const double lamb[5] = {0.95234, 0.65524, 0.34380, 0.39310, 0.0};
  memcpy(o->r_e, r_e, sizeof r_e); memcpy(o->alph, alph, sizeof alph);
  memcpy(o->beta, beta, sizeof beta); memcpy(o->A, A, sizeof A); memcpy(o->B, B, sizeof B);
  memcpy(o->kapp, kapp, sizeof kapp); memcpy(o->lamb, lamb, sizeof lamb);
  /* F0..F4 with index 1..4 of each (datafmt.f90:241-249) */
  const double F0[4] = {-1.56489, -1.39123, 1.77199, 1.59833};
  const double F1[4] = {-1.58967, 1.30636, 9.81033, 0.0};
  const double F2[4] = {-1.54116, 2.02821, 6.56240, 0.0};
  const double F3[4] = {-1.51798, 2.30979, 7.69582, 0.0};
  const double F4[4] = {-1.19082, 4.12936, 10.32338, 0.0};
  memcpy(o->F[0], F0, sizeof F0); memcpy(o->F[1], F1, sizeof F1); memcpy(o->F[2], F2, sizeof F2);
  memcpy(o->F[3], F3, sizeof F3); memcpy(o->F[4], F4, sizeof F4);
  const double re[5] = {54.62910, 64.26953, 66.21202, 66.92391, 74.23105};
  const double rmax[5] = {54.62910, 65.24078, 66.56797, 70.57748, 99.99999e9};
  memcpy(o->re, re, sizeof re); memcpy(o->rmax, rmax, sizeof rmax);
  o->fe = 1.39478; o->gamm = 2.11725; o->nu = 0.37457;
}

void orc_sys_init(orc_sys* s, int n, const double box[3], const int fft[3], double rcut) {
  s->n = n;
  for (int d = 0; d < 3; ++d) {
    s->box[d] = box[d];
    s->fft[d] = fft[d];
    s->ncell[d] = (int)(box[d] / rcut);     /* main.f90:189 */
    s->lcell[d] = box[d] / (double)s->ncell[d]; /* main.f90:190 */
  }
  s->rcut = rcut;
  s->rcut2 = rcut * rcut;  /* main.f90:173 */
  s->a = 3.8 / rcut;       /* main.f90:174, force.f90:481 */
}

/* ------------------------------------------------------------------------ */
/* Cell list, cellsort.f90                                                   */
/* ------------------------------------------------------------------------ */
static int id_cell(const int nc[3], int i, int j, int k) { /* cellsort.f90:39-50, 1-based i,j,k */
  int l = (i - 1 + nc[0]) % nc[0];
  int m = (j - 1 + nc[1]) % nc[1];
  int n = (k - 1 + nc[2]) % nc[2];
  return l + m * nc[0] + n * nc[0] * nc[1]; /* 0-based cell id */
}

void orc_cell_hierarchy(const int nc[3], int* pair) { /* cellsort.f90:3-36 */
  static const int off[13][3] = {{-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {0, 0, -1},
                                 {1, 0, -1},  {-1, 1, -1}, {0, 1, -1},  {1, 1, -1},  {-1, -1, 0},
                                 {0, -1, 0},  {1, -1, 0},  {-1, 0, 0}};
  for (int i = 1; i <= nc[0]; ++i)
    for (int j = 1; j <= nc[1]; ++j)
      for (int k = 1; k <= nc[2]; ++k) {
        int id = id_cell(nc, i, j, k);
        for (int n = 0; n < 13; ++n)
          pair[13 * id + n] = id_cell(nc, i + off[n][0], j + off[n][1], k + off[n][2]);
      }
}

int orc_cell_sort(const orc_sys* s, const double* x, int* npt, int* link) { /* cellsort.f90:53-80 */
  const int* nc = s->ncell;
  int ncall = nc[0] * nc[1] * nc[2];
  for (int c = 0; c < ncall; ++c) npt[c] = 0;
  for (int i = 0; i < s->n; ++i) {
    int id[3];
    for (int d = 0; d < 3; ++d) {
      double xx = x[3 * i + d] + 0.5 * s->box[d];
      id[d] = (int)(xx / s->lcell[d]);
      id[d] = (id[d] + nc[d]) % nc[d];
    }
    int c = id[0] + id[1] * nc[0] + id[2] * nc[0] * nc[1];
    if (npt[c] >= ORC_NPT_MAX) return ORC_ERR_CELL_OVERFLOW;
    link[ORC_NPT_MAX * c + npt[c]] = i;
    npt[c]++;
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* EAM pair functions, eam.f90                                               */
/* ------------------------------------------------------------------------ */
static double p20(double x) {
  double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
  return x16 * x4;
}

/* Zhou-Johnson-Wadley radial form and its r-derivative; used by every branch */
static void zjw(double r, double re, double fe_, double beta, double lambda, double* f,
                double* dfdr) {
  double r_re = r / re, r_re_1 = r_re - 1.0;
  double l20 = p20(r_re - lambda), l19 = l20 / (r_re - lambda);
  double eb = exp(-beta * r_re_1);
  *f = fe_ * (eb / (1.0 + l20));
  *dfdr = fe_ * (eb * (-beta * (1.0 + l20) - 20.0 * l19) / re / ((1.0 + l20) * (1.0 + l20)));
}

/* pair term A e^{-alpha(...)}/(1+k20) - B e^{-beta(...)}/(1+l20): eam.f90:287-296 */
static void zjw_pair(double r, double re, double alpha, double beta, double A, double B,
                     double kappa, double lambda, double* phi, double* dphidr) {
  double r_re = r / re, r_re_1 = r_re - 1.0;
  double l20 = p20(r_re - lambda), l19 = l20 / (r_re - lambda);
  double k20 = p20(r_re - kappa), k19 = k20 / (r_re - kappa);
  double ea = exp(-alpha * r_re_1), eb = exp(-beta * r_re_1);
  *dphidr = A * ea * (-alpha * (1.0 + k20) - 20.0 * k19) / re / ((1.0 + k20) * (1.0 + k20)) -
            B * eb * (-beta * (1.0 + l20) - 20.0 * l19) / re / ((1.0 + l20) * (1.0 + l20));
  *phi = A * ea / (1.0 + k20) - B * eb / (1.0 + l20);
}

static void metal_homo(const orc_metal* e, double r, double fj[2], double dfjdr[2], double* phi,
                       double* df) { /* eam.f90:3-35 */
  double r_re = r / e->r_e, r_re_1 = r_re - 1.0;
  double l20 = p20(r_re - e->lambda), l19 = l20 / (r_re - e->lambda);
  double k20 = p20(r_re - e->kappa), k19 = k20 / (r_re - e->kappa);
  double ea = exp(-e->alpha * r_re_1), eb = exp(-e->beta * r_re_1);
  double ftn = eb / (1.0 + l20);
  double dftndr = eb * (-e->beta * (1.0 + l20) - 20.0 * l19) / e->r_e / ((1.0 + l20) * (1.0 + l20));
  fj[0] = e->f_e * ftn; fj[1] = fj[0];
  dfjdr[0] = e->f_e * dftndr / r; dfjdr[1] = dfjdr[0];
  *phi = e->A * ea / (1.0 + k20) - e->B * ftn;
  double dphidr = e->A * ea * (-e->alpha * (1.0 + k20) - 20.0 * k19) / e->r_e /
                      ((1.0 + k20) * (1.0 + k20)) - e->B * dftndr;
  *df = -dphidr / r;
}

/* ZJW density and pair parts of one metal at r: f = fe ftn, dfdr = fe dftn/dr,
   phi = A e^{-a(r/re-1)}/(1+kappa20) - B ftn and its r-derivative (eam.f90:55-70) */
static void metal_part(const orc_metal* e, double r, double* f, double* dfdr, double* phi,
                       double* dphidr) {
  double r_re = r / e->r_e, r_re_1 = r_re - 1.0;
  double l20 = p20(r_re - e->lambda), l19 = l20 / (r_re - e->lambda);
  double k20 = p20(r_re - e->kappa), k19 = k20 / (r_re - e->kappa);
  double ea = exp(-e->alpha * r_re_1), eb = exp(-e->beta * r_re_1);
  double ftn = eb / (1.0 + l20);
  double dftndr = eb * (-e->beta * (1.0 + l20) - 20.0 * l19) / e->r_e / ((1.0 + l20) * (1.0 + l20));
  *f = e->f_e * ftn;
  *dfdr = e->f_e * dftndr;
  *phi = e->A * ea / (1.0 + k20) - e->B * ftn;
  *dphidr = e->A * ea * (-e->alpha * (1.0 + k20) - 20.0 * k19) / e->r_e /
                ((1.0 + k20) * (1.0 + k20)) - e->B * dftndr;
}

/* METAL_HETERO, eam.f90:38-98: each metal's own density, the pair term the
   density-weighted mean 0.5 (phi_i f_j / f_i + phi_j f_i / f_j) */
static void metal_hetero(const orc_metal* ei, const orc_metal* ej, double r, double fj[2],
                         double dfjdr[2], double* phi, double* df) {
  double fi, dfi, phi_i, dphi_i, fk, dfk, phi_j, dphi_j;
  metal_part(ei, r, &fi, &dfi, &phi_i, &dphi_i);
  metal_part(ej, r, &fk, &dfk, &phi_j, &dphi_j);
  fj[0] = fi; dfjdr[0] = dfi / r;
  fj[1] = fk; dfjdr[1] = dfk / r;
  *phi = 0.5 * (phi_i * fj[1] / fj[0] + phi_j * fj[0] / fj[1]);
  *df = -0.5 * ((dphi_i * fj[1] * fj[0] + phi_i * dfjdr[1] * r * fj[0] - phi_i * fj[1] * dfjdr[0] * r) /
                    (fj[0] * fj[0]) +
                (dphi_j * fj[0] * fj[1] + phi_j * dfjdr[0] * r * fj[1] - phi_j * fj[0] * dfjdr[1] * r) /
                    (fj[1] * fj[1])) / r;
}

/* AL_NI, eam.f90:101-168: fj(1) from the first argument's density, fj(2) from
   the second's, and a fixed Al-Ni pair term (constants eam.f90:147-153).  The
   caller passes (Al, Ni) for both orders of the pair, so which atom receives
   which density depends on the loop's orientation (force.f90:60-63). */
static void al_ni(const orc_metal* ea_, const orc_metal* eb_, double r, double fj[2],
                  double dfjdr[2], double* phi, double* df) {
  double f, dfdr;
  zjw(r, ea_->r_e, ea_->f_e, ea_->beta, ea_->lambda, &f, &dfdr);
  fj[0] = f; dfjdr[0] = dfdr / r;
  zjw(r, eb_->r_e, eb_->f_e, eb_->beta, eb_->lambda, &f, &dfdr);
  fj[1] = f; dfjdr[1] = dfdr / r;
  double dphidr;
  zjw_pair(r, 2.71579, 8.00443, 4.75970, 0.44254, 0.68349, 0.63279, 0.81777, phi, &dphidr);
  *df = -dphidr / r;
}

/* oxygen density e^{-gamma(...)}/(1+nu20) with r_e of the partner row (Q15), eam.f90:299-302 */
static void oxy_density(const orc_oxid* o, double re, double r, double* f, double* dfdr_over_r) {
  double r_re = r / re, r_re_1 = r_re - 1.0;
  double eg = exp(-o->gamm * r_re_1);
  double n20 = p20(r_re - o->nu), n19 = n20 / (r_re - o->nu);
  *f = o->fe * eg / (1.0 + n20);
  *dfdr_over_r = o->fe * eg * (-o->gamm * (1.0 + n20) - 20.0 * n19) / re / ((1.0 + n20) * (1.0 + n20)) / r;
}

void orc_eam_pair(const orc_params* p, int id_i, int id_j, double r, double fj[2],
                  double dfjdr[2], double* phi, double* df) {
  /* dispatch of force.f90:58-74 */
  if (id_i < 4 && id_j < 4 && id_i == id_j) {
    metal_homo(&p->eam[id_i - 1], r, fj, dfjdr, phi, df);
  } else if (id_i == 1 && id_j == 2) {
    al_ni(&p->eam[0], &p->eam[1], r, fj, dfjdr, phi, df);
  } else if (id_i == 2 && id_j == 1) {
    al_ni(&p->eam[0], &p->eam[1], r, fj, dfjdr, phi, df);
  } else if (id_i < 4 && id_j < 4) {
    metal_hetero(&p->eam[id_i - 1], &p->eam[id_j - 1], r, fj, dfjdr, phi, df);
  } else if (id_i < 4 && id_j == 4) { /* METAL_OXYGEN eam.f90:171-225 */
    const orc_metal* e = &p->eam[id_i - 1];
    double f, dfdr;
    zjw(r, e->r_e, e->f_e, e->beta, e->lambda, &f, &dfdr);
    fj[0] = f; dfjdr[0] = dfdr / r;
    const orc_oxid* o = &p->ox;
    int m = id_i - 1;
    double dphidr;
    zjw_pair(r, o->r_e[m], o->alph[m], o->beta[m], o->A[m], o->B[m], o->kapp[m], o->lamb[m], phi,
             &dphidr);
    *df = -dphidr / r;
    oxy_density(o, o->r_e[m], r, &fj[1], &dfjdr[1]);
  } else if (id_i == 4 && id_j < 4) { /* OXYGEN_METAL eam.f90:228-283 */
    const orc_metal* e = &p->eam[id_j - 1];
    double f, dfdr;
    zjw(r, e->r_e, e->f_e, e->beta, e->lambda, &f, &dfdr);
    fj[1] = f; dfjdr[1] = dfdr / r;
    const orc_oxid* o = &p->ox;
    int m = id_j - 1;
    double dphidr;
    zjw_pair(r, o->r_e[m], o->alph[m], o->beta[m], o->A[m], o->B[m], o->kapp[m], o->lamb[m], phi,
             &dphidr);
    *df = -dphidr / r;
    oxy_density(o, o->r_e[m], r, &fj[0], &dfjdr[0]);
  } else if (id_i == 4 && id_j == 4) { /* OXYGEN_ONLY eam.f90:286-328 */
    const orc_oxid* o = &p->ox;
    double dphidr;
    zjw_pair(r, o->r_e[3], o->alph[3], o->beta[3], o->A[3], o->B[3], o->kapp[3], o->lamb[3], phi,
             &dphidr);
    *df = -dphidr / r;
    oxy_density(o, o->r_e[3], r, &fj[0], &dfjdr[0]);
    fj[1] = fj[0]; dfjdr[1] = dfjdr[0];
  } else {
    /* H (id 5): FORCE stops with "Unknown particle pair"; callers reject it first */
    fj[0] = fj[1] = dfjdr[0] = dfjdr[1] = *phi = *df = NAN;
  }
}

static double poly3(const double c[4], double t) { /* sum_j c_j t^(j-1), force.f90:268-270 */
  double s = 0.0, tp = 1.0;
  for (int j = 0; j < 4; ++j) { s += c[j] * tp; tp *= t; }
  return s;
}
static double dpoly3(const double c[4], double t, double scale) { /* force.f90:271-273 */
  double s = 0.0, tp = 1.0;
  for (int j = 1; j < 4; ++j) { s += c[j] * (double)j * tp / scale; tp *= t; }
  return s;
}

int orc_embed(const orc_params* p, int id, double rho, double* F, double* dF) {
  /* force.f90:256-348 */
  *F = 0.0; *dF = 0.0;
  if (id < 4) {
    const orc_metal* e = &p->eam[id - 1];
    double rho_n = 0.85 * e->rho_e, rho_0 = 1.15 * e->rho_e;
    if (rho < rho_n) {
      double t = rho / rho_n - 1.0;
      *F = poly3(e->Fn, t); *dF = dpoly3(e->Fn, t, rho_n);
    } else if (rho < e->rho_e) {
      double t = rho / e->rho_e - 1.0;
      *F = poly3(e->Fm, t); *dF = dpoly3(e->Fm, t, e->rho_e);
    } else if (rho < rho_0) {
      double t = rho / e->rho_e - 1.0;
      *F = poly3(e->Fp, t); *dF = dpoly3(e->Fp, t, e->rho_e);
    } else if (rho > 1.0e-6) {
      double t = pow(rho / e->rho_s, e->eta);
      *F = e->Fe * (1.0 - log(t)) * t;
      *dF = -e->Fe * e->eta * log(t) * t / rho;
    }
    return ORC_OK;
  } else if (id == 4) {
    const orc_oxid* o = &p->ox;
    int k = 4;
    for (int s = 0; s < 4; ++s)
      if (rho < o->rmax[s]) { k = s; break; }
    double t = rho / o->re[k] - 1.0;
    *F = poly3(o->F[k], t); *dF = dpoly3(o->F[k], t, o->re[k]);
    return ORC_OK;
  }
  return ORC_ERR_UNKNOWN_PAIR;
}

/* ------------------------------------------------------------------------ */
/* FORCE, force.f90:3-361                                                    */
/* ------------------------------------------------------------------------ */
static double g_ream2 = -1.0;  /* EAM cutoff squared; < 0: the common Rcut */

typedef struct {
  const orc_sys* s; const orc_params* p; const int* id; const double* x; const double* q;
  double* f; double* rho; double* drhodx; int* nlist; int* list; double* drho_list;
  double Epair, Edir, Ectip; int err;
} force_ws;

static void force_pair(force_ws* w, int li, int lj) {
  const orc_sys* s = w->s; const orc_params* p = w->p;
  double rx[3], r2;
  for (int d = 0; d < 3; ++d) {
    double dx = w->x[3 * li + d] - w->x[3 * lj + d];
    rx[d] = dx - s->box[d] * round(dx / s->box[d]);
  }
  r2 = rx[0] * rx[0] + rx[1] * rx[1] + rx[2] * rx[2];
  if (!(r2 < s->rcut2)) return;
  double r = sqrt(r2);
  int id_i = w->id[li], id_j = w->id[lj];
  double q_i = w->q[li], q_j = w->q[lj];
  double xi_i = p->prm[id_i - 1].xi, Z_i = p->prm[id_i - 1].Z;
  double xi_j = p->prm[id_j - 1].xi, Z_j = p->prm[id_j - 1].Z;
  if (id_i < 1 || id_i > 4 || id_j < 1 || id_j > 4) { w->err = ORC_ERR_UNKNOWN_PAIR; return; }
  double fj[2], dfjdr[2], phi, df;
  orc_eam_pair(p, id_i, id_j, r, fj, dfjdr, &phi, &df);
  if (g_ream2 > 0.0 && !(r2 < g_ream2)) {  /* split cutoff: no EAM terms past r_eam */
    fj[0] = fj[1] = dfjdr[0] = dfjdr[1] = phi = df = 0.0;
  }
  w->nlist[li]++; w->nlist[lj]++;
  if (w->nlist[li] > ORC_NNB_MAX || w->nlist[lj] > ORC_NNB_MAX) { w->err = ORC_ERR_NB_OVERFLOW; return; }
  int ni = w->nlist[li] - 1, nj = w->nlist[lj] - 1;
  w->list[(size_t)ORC_NNB_MAX * li + ni] = lj;
  w->list[(size_t)ORC_NNB_MAX * lj + nj] = li;
  for (int d = 0; d < 3; ++d) {
    w->drho_list[((size_t)ORC_NNB_MAX * li + ni) * 3 + d] = rx[d] * dfjdr[0];
    w->drho_list[((size_t)ORC_NNB_MAX * lj + nj) * 3 + d] = -rx[d] * dfjdr[1];
  }
  w->rho[li] += fj[1];
  w->rho[lj] += fj[0];
  for (int d = 0; d < 3; ++d) {
    w->drhodx[3 * li + d] += rx[d] * dfjdr[1];
    w->drhodx[3 * lj + d] -= rx[d] * dfjdr[0];
  }
  for (int d = 0; d < 3; ++d) { w->f[3 * li + d] += rx[d] * df; w->f[3 * lj + d] -= rx[d] * df; }
  w->Epair += phi;
  /* real-space Ewald, force.f90:94-98 */
  double a = s->a, a2 = a * a;
  double tmp = erfc(a * r);
  w->Edir += q_i * q_j * tmp / r;
  df = KC * q_i * q_j * (2.0 * exp(-a2 * r2) * a * r / SQRTPI + tmp) / r2 / r;
  for (int d = 0; d < 3; ++d) { w->f[3 * li + d] += df * rx[d]; w->f[3 * lj + d] -= df * rx[d]; }
  /* CTIP Slater corrections, force.f90:101-140 (same-species fifj without 1/r: Q2) */
  double jfi, ifj, djfidr, difjdr, fifj, dfifjdr;
  const double* c = p->prm[id_i - 1].fafb[id_j - 1];
  if (id_i == id_j) {
    double xir_i = 2.0 * xi_i * r, ei = exp(-xir_i);
    jfi = -ei * (xi_i + 1.0 / r);
    ifj = jfi;
    djfidr = ei * (2.0 * (xi_i * xi_i) + (1.0 + xir_i) / r2);
    difjdr = djfidr;
    fifj = -ei * (1.0 + c[1] * r + c[2] * r2 + c[3] * r2 * r);
    dfifjdr = ei * ((xir_i + 1.0) * (1.0 / r2 + c[1] / r + c[2] + c[3] * r) -
                    (c[1] / r + c[2] * 2.0 + c[3] * 3.0 * r));
  } else {
    double xir_i = 2.0 * xi_i * r, ei = exp(-xir_i);
    double xir_j = 2.0 * xi_j * r, ej = exp(-xir_j);
    jfi = -ei * (xi_i + 1.0 / r);
    ifj = -ej * (xi_j + 1.0 / r);
    djfidr = ei * (2.0 * (xi_i * xi_i) + (1.0 + xir_i) / r2);
    difjdr = ej * (2.0 * (xi_j * xi_j) + (1.0 + xir_j) / r2);
    fifj = -c[0] * ei - c[1] * ej - c[2] * ei / r - c[3] * ej / r;
    dfifjdr = c[0] * ei * 2.0 * xi_i + c[1] * ej * 2.0 * xi_j + c[2] * ei * (1.0 + xir_i) / r2 +
              c[3] * ej * (1.0 + xir_j) / r2;
  }
  df = KC * ((q_i * Z_j + q_j * Z_i - q_i * q_j) * dfifjdr - q_i * Z_j * djfidr -
             q_j * Z_i * difjdr) / r;
  for (int d = 0; d < 3; ++d) { w->f[3 * li + d] += df * rx[d]; w->f[3 * lj + d] -= df * rx[d]; }
  w->Ectip += q_i * Z_j * (jfi - fifj) + q_j * Z_i * (ifj - fifj) + q_i * q_j * fifj;
}

int orc_force(const orc_sys* s, const orc_params* p, const int* id, const double* x,
              const double* q, const int* npt, const int* link, const int* pair, double* f,
              double* en, int* nlist_out, int* list_out, double* rho_out) {
  int n = s->n;
  int ncall = s->ncell[0] * s->ncell[1] * s->ncell[2];
  force_ws w;
  memset(&w, 0, sizeof w);
  w.s = s; w.p = p; w.id = id; w.x = x; w.q = q; w.f = f;
  w.rho = calloc((size_t)n, sizeof(double));
  w.drhodx = calloc((size_t)n * 3, sizeof(double));
  w.nlist = calloc((size_t)n, sizeof(int));
  w.list = calloc((size_t)n * ORC_NNB_MAX, sizeof(int));
  w.drho_list = calloc((size_t)n * ORC_NNB_MAX * 3, sizeof(double));
  double* dF = calloc((size_t)n, sizeof(double));
  if (!w.rho || !w.drhodx || !w.nlist || !w.list || !w.drho_list || !dF) {
    free(w.rho); free(w.drhodx); free(w.nlist); free(w.list); free(w.drho_list); free(dF);
    return -1;
  }
  for (int i = 0; i < 3 * n; ++i) f[i] = 0.0;
  for (int m = 0; m < ncall && !w.err; ++m) {
    int nl = npt[m];
    const int* lk = link + (size_t)ORC_NPT_MAX * m;
    for (int i = 0; i < nl - 1 && !w.err; ++i)
      for (int j = i + 1; j < nl && !w.err; ++j) force_pair(&w, lk[i], lk[j]);
    for (int k = 0; k < 13 && !w.err; ++k) {
      int cp = pair[13 * m + k];
      const int* lp = link + (size_t)ORC_NPT_MAX * cp;
      for (int j = 0; j < npt[cp] && !w.err; ++j)
        for (int i = 0; i < nl && !w.err; ++i) force_pair(&w, lk[i], lp[j]);
    }
  }
  int err = w.err;
  double Egroup = 0.0;
  if (!err) {
    for (int i = 0; i < n; ++i) {
      double F;
      if (orc_embed(p, id[i], w.rho[i], &F, &dF[i]) != ORC_OK) { err = ORC_ERR_UNKNOWN_PAIR; break; }
      for (int d = 0; d < 3; ++d) f[3 * i + d] -= w.drhodx[3 * i + d] * dF[i];
      Egroup += F;
    }
  }
  if (!err) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < w.nlist[i]; ++j) {
        int jj = w.list[(size_t)ORC_NNB_MAX * i + j];
        for (int d = 0; d < 3; ++d)
          f[3 * i + d] -= dF[jj] * w.drho_list[((size_t)ORC_NNB_MAX * i + j) * 3 + d];
      }
  }
  en[0] = w.Epair; en[1] = w.Edir * KC; en[2] = w.Ectip * KC; en[3] = Egroup;
  if (nlist_out) memcpy(nlist_out, w.nlist, (size_t)n * sizeof(int));
  if (list_out) memcpy(list_out, w.list, (size_t)n * ORC_NNB_MAX * sizeof(int));
  if (rho_out) memcpy(rho_out, w.rho, (size_t)n * sizeof(double));
  free(w.rho); free(w.drhodx); free(w.nlist); free(w.list); free(w.drho_list); free(dF);
  return err;
}

/* ------------------------------------------------------------------------ */
/* PME_SELF, force.f90:467-517                                               */
/* ------------------------------------------------------------------------ */
double orc_pme_self(const orc_sys* s, const orc_params* p, const int* id, const double* q,
                    double* psi) {
  double a = 3.8 / s->rcut, a2 = a * a;
  double Eself = 0.0, Esum = 0.0;
  for (int i = 0; i < s->n; ++i) {
    double qi = q[i];
    Eself -= qi * qi;
    const orc_potn* pp = &p->prm[id[i] - 1];
    Esum += pp->chi * qi + 0.5 * pp->J * (qi * qi);
    double dq = qi - pp->qmin;
    if (dq < 0.0) Esum += 2.0 * WPEN * (dq * dq);
    dq = qi - pp->qmax;
    if (dq > 0.0) Esum += 2.0 * WPEN * (dq * dq);
  }
  const int* K = s->fft;
  for (int k = 0; k < K[2]; ++k) {
    int iz = (k + K[2] / 2) % K[2] - K[2] / 2;
    for (int j = 0; j < K[1]; ++j) {
      int iy = (j + K[1] / 2) % K[1] - K[1] / 2;
      for (int i = 0; i < K[0]; ++i) {
        int ix = (i + K[0] / 2) % K[0] - K[0] / 2;
        size_t g = (size_t)i + (size_t)K[0] * ((size_t)j + (size_t)K[1] * k);
        if (ix * ix + iy * iy + iz * iz != 0) {
          double m2 = (double)(ix * ix) / (s->box[0] * s->box[0]) +
                      (double)(iy * iy) / (s->box[1] * s->box[1]) +
                      (double)(iz * iz) / (s->box[2] * s->box[2]);
          psi[g] = exp(-m2 * (PI_ * PI_) / a2) / m2;
        } else {
          psi[g] = 0.0; /* never assigned by the reference (Q5): zero here */
        }
      }
    }
  }
  return Eself * a / SQRTPI * KC + Esum;
}

/* ------------------------------------------------------------------------ */
/* 3D FFT with FFTE's ZFFT3D contract (Makefile:15-16; call sites            */
/* force.f90:412-426, ctip.f90:366-380).  Mixed-radix recursive DFT.         */
/* ------------------------------------------------------------------------ */
/* tw[k] = exp(-2 pi i k / N) for the top-level length N; a sub-transform of
   length n uses W_n^e = tw[e * (N / n)]. */
static void fft_rec(double complex* out, const double complex* in, int n, int stride, int sign,
                    double complex* tmp, const double complex* tw, int N) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = n;
  for (int f = 2; f * f <= n; ++f) if (n % f == 0) { p = f; break; }
  int m = n / p, step = N / n;
  for (int r = 0; r < p; ++r)
    fft_rec(out + (size_t)r * m, in + (size_t)r * stride, m, stride * p, sign, tmp, tw, N);
  for (int k = 0; k < m; ++k) {
    for (int qq = 0; qq < p; ++qq) {
      double complex acc = 0.0;
      for (int r = 0; r < p; ++r) {
        long e = ((long)r * (k + (long)m * qq)) % n;
        double complex wv = tw[e * step];
        if (sign > 0) wv = conj(wv);
        acc += out[(size_t)r * m + k] * wv;
      }
      tmp[qq] = acc;
    }
    for (int qq = 0; qq < p; ++qq) out[(size_t)qq * m + k] = tmp[qq];
  }
}

static void fft1d_strided(double complex* a, int n, size_t stride, int sign, double complex* lin,
                          double complex* lout, double complex* tmp, const double complex* tw) {
  for (int i = 0; i < n; ++i) lin[i] = a[i * stride];
  fft_rec(lout, lin, n, 1, sign, tmp, tw, n);
  for (int i = 0; i < n; ++i) a[i * stride] = lout[i];
}

static double complex* twiddles(int n) {
  double complex* tw = malloc(sizeof(double complex) * n);
  for (int k = 0; k < n; ++k) {
    double ang = 2.0 * PI_ * (double)k / (double)n;
    tw[k] = cos(ang) - I * sin(ang);
  }
  return tw;
}

void orc_fft3d(double* data, int nx, int ny, int nz, int sign) {
  double complex* a = (double complex*)data;
  int nmax = nx > ny ? nx : ny; nmax = nmax > nz ? nmax : nz;
  double complex* lin = malloc(sizeof(double complex) * nmax);
  double complex* lout = malloc(sizeof(double complex) * nmax);
  double complex* tmp = malloc(sizeof(double complex) * nmax);
  double complex *twx = twiddles(nx), *twy = twiddles(ny), *twz = twiddles(nz);
  int sg = sign < 0 ? -1 : 1;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      fft1d_strided(a + (size_t)nx * (j + (size_t)ny * k), nx, 1, sg, lin, lout, tmp, twx);
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i)
      fft1d_strided(a + i + (size_t)nx * ny * k, ny, (size_t)nx, sg, lin, lout, tmp, twy);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      fft1d_strided(a + i + (size_t)nx * j, nz, (size_t)nx * ny, sg, lin, lout, tmp, twz);
  if (sg > 0) {
    double s = 1.0 / ((double)nx * ny * nz);
    for (size_t g = 0; g < (size_t)nx * ny * nz; ++g) a[g] *= s;
  }
  free(lin); free(lout); free(tmp); free(twx); free(twy); free(twz);
}

/* ------------------------------------------------------------------------ */
/* PME, force.f90:364-464                                                    */
/* ------------------------------------------------------------------------ */
/* B-spline weights AX and their u-derivatives AXX, force.f90:388-400 */
static void bspline4(double u, double AX[4], double AXX[4]) {
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

static void grid_index(const orc_sys* s, const double* xi, double u[3], int base[3]) {
  for (int d = 0; d < 3; ++d) {
    u[d] = (double)s->fft[d] * (xi[d] + s->box[d] * 0.5) / s->box[d];
    base[d] = (int)u[d];
  }
}
static int wrapk(int b, int nn, int K) { return (b + nn - 1 + K) % K; } /* MOD(INT(u)+nn-2+K,K), nn 0-based */

/* Weight rounding of the spread: 1 = FP32-rounded as the reference's
   default-kind CMPLX makes them (force.f90:406, ctip.f90:359; quirk Q1),
   0 = exact FP64 weights (the build's FP64 / MIXED modes). */
static int g_round_f32 = 1;
void orc_set_spread_rounding(int on) { g_round_f32 = on ? 1 : 0; }
void orc_set_eam_cutoff(double r_eam) { g_ream2 = r_eam > 0.0 ? r_eam * r_eam : -1.0; }

static void spread(const orc_sys* s, const double* x, const double* q, double complex* Q) {
  const int* K = s->fft;
  size_t ng = (size_t)K[0] * K[1] * K[2];
  for (size_t g = 0; g < ng; ++g) Q[g] = 0.0;
  for (int i = 0; i < s->n; ++i) {
    double u[3], AX[3][4], AXX[3][4]; int b[3];
    grid_index(s, x + 3 * i, u, b);
    for (int d = 0; d < 3; ++d) bspline4(u[d], AX[d], AXX[d]);
    for (int a = 0; a < 4; ++a)
      for (int bb = 0; bb < 4; ++bb)
        for (int c = 0; c < 4; ++c) {
          int ix = wrapk(b[0], a, K[0]), iy = wrapk(b[1], bb, K[1]), iz = wrapk(b[2], c, K[2]);
          double w = AX[0][a] * AX[1][bb] * AX[2][c];
          if (g_round_f32) w = (double)(float)w;
          Q[ix + (size_t)K[0] * (iy + (size_t)K[1] * iz)] += q[i] * w;
        }
  }
}

/* Qfin = ZFFT3D(-1)( G . ZFFT3D(+1)(Qbas) ), force.f90:411-426 / ctip.f90:365-380 */
static void convolve(const orc_sys* s, const double* psi, double complex* Q) {
  const int* K = s->fft;
  orc_fft3d((double*)Q, K[0], K[1], K[2], +1);
  for (int i = 0; i < K[0]; ++i) {
    double z1 = 4.0 + 2.0 * cos(2.0 * PI_ * (double)i / (double)K[0]); z1 *= z1;
    for (int j = 0; j < K[1]; ++j) {
      double z2 = 4.0 + 2.0 * cos(2.0 * PI_ * (double)j / (double)K[1]); z2 *= z2;
      for (int k = 0; k < K[2]; ++k) {
        double z3 = 4.0 + 2.0 * cos(2.0 * PI_ * (double)k / (double)K[2]); z3 *= z3;
        size_t g = i + (size_t)K[0] * (j + (size_t)K[1] * k);
        Q[g] = Q[g] * psi[g] * (36.0 * 36.0 * 36.0) / z1 / z2 / z3;
      }
    }
  }
  orc_fft3d((double*)Q, K[0], K[1], K[2], -1);
}

double orc_pme(const orc_sys* s, const int* id, const double* x, const double* q,
               const double* psi, double* f) {
  (void)id;
  const int* K = s->fft;
  size_t ng = (size_t)K[0] * K[1] * K[2];
  double complex* Qbas = malloc(sizeof(double complex) * ng);
  double complex* Qfin = malloc(sizeof(double complex) * ng);
  double* ff = calloc((size_t)s->n * 3, sizeof(double));
  double V = s->box[0] * s->box[1] * s->box[2];
  spread(s, x, q, Qbas);
  memcpy(Qfin, Qbas, sizeof(double complex) * ng);
  convolve(s, psi, Qfin);
  double Epme = 0.0;
  for (size_t g = 0; g < ng; ++g) Epme += creal(Qbas[g] * Qfin[g]);
  Epme = Epme * (double)K[0] * (double)K[1] * (double)K[2];
  double cst = KC * (double)K[0] * (double)K[1] * (double)K[2] / V / PI_;
  double kx[3] = {(double)K[0] / s->box[0], (double)K[1] / s->box[1], (double)K[2] / s->box[2]};
  double Fsum[3] = {0, 0, 0};
  for (int i = 0; i < s->n; ++i) {
    double u[3], AX[3][4], AXX[3][4]; int b[3];
    grid_index(s, x + 3 * i, u, b);
    for (int d = 0; d < 3; ++d) bspline4(u[d], AX[d], AXX[d]);
    double qi = q[i];
    for (int a = 0; a < 4; ++a)
      for (int bb = 0; bb < 4; ++bb)
        for (int c = 0; c < 4; ++c) {
          double dQ[3];
          dQ[0] = qi * AXX[0][a] * AX[1][bb] * AX[2][c] * kx[0];
          dQ[1] = qi * AX[0][a] * AXX[1][bb] * AX[2][c] * kx[1];
          dQ[2] = qi * AX[0][a] * AX[1][bb] * AXX[2][c] * kx[2];
          int ix = wrapk(b[0], a, K[0]), iy = wrapk(b[1], bb, K[1]), iz = wrapk(b[2], c, K[2]);
          double Qr = creal(Qfin[ix + (size_t)K[0] * (iy + (size_t)K[1] * iz)]);
          for (int d = 0; d < 3; ++d) ff[3 * i + d] -= cst * dQ[d] * Qr;
        }
    for (int d = 0; d < 3; ++d) Fsum[d] += ff[3 * i + d];
  }
  for (int d = 0; d < 3; ++d) Fsum[d] /= (double)s->n;
  for (int i = 0; i < s->n; ++i)
    for (int d = 0; d < 3; ++d) f[3 * i + d] += ff[3 * i + d] - Fsum[d];
  free(Qbas); free(Qfin); free(ff);
  return KC * Epme / V / PI_ / 2.0;
}

/* PHI_MESH, ctip.f90:336-396 */
void orc_phi_mesh(const orc_sys* s, const double* x, const double* q, const double* psi,
                  double* phi) {
  const int* K = s->fft;
  size_t ng = (size_t)K[0] * K[1] * K[2];
  double complex* Q = malloc(sizeof(double complex) * ng);
  spread(s, x, q, Q);
  convolve(s, psi, Q);
  for (int i = 0; i < s->n; ++i) {
    double u[3], AX[3][4], AXX[3][4]; int b[3];
    grid_index(s, x + 3 * i, u, b);
    for (int d = 0; d < 3; ++d) bspline4(u[d], AX[d], AXX[d]);
    double e = 0.0;
    for (int a = 0; a < 4; ++a)
      for (int bb = 0; bb < 4; ++bb)
        for (int c = 0; c < 4; ++c) {
          int ix = wrapk(b[0], a, K[0]), iy = wrapk(b[1], bb, K[1]), iz = wrapk(b[2], c, K[2]);
          e += AX[0][a] * AX[1][bb] * AX[2][c] * creal(Q[ix + (size_t)K[0] * (iy + (size_t)K[1] * iz)]);
        }
    phi[i] = e;
  }
  free(Q);
}

/* ------------------------------------------------------------------------ */
/* CTIP_SOLVER, ctip.f90:3-212                                               */
/* ------------------------------------------------------------------------ */
static int ctip_pair(const orc_sys* s, const orc_params* p, const int* id, const double* x,
                     int li, int lj, int* nlist, int* list, double* asym, double* sym) {
  double rx[3], r2;
  for (int d = 0; d < 3; ++d) {
    double dx = x[3 * li + d] - x[3 * lj + d];
    rx[d] = dx - s->box[d] * round(dx / s->box[d]);
  }
  r2 = rx[0] * rx[0] + rx[1] * rx[1] + rx[2] * rx[2];
  if (!(r2 < s->rcut2)) return ORC_OK;
  double r = sqrt(r2), ar = s->a * r;
  int id_i = id[li], id_j = id[lj];
  double xi_i = p->prm[id_i - 1].xi, xi_j = p->prm[id_j - 1].xi;
  double ej = exp(-2.0 * xi_j * r), ej_r = ej / r;
  double afb_j = -xi_j * ej - ej_r, afb_i, fifj;
  const double* c = p->prm[id_i - 1].fafb[id_j - 1];
  if (id_i == id_j) {
    afb_i = afb_j;
    double xir = xi_i * r;
    /* Q3: same-species form with xi powers inside fafb, ctip.f90:90-93 */
    fifj = -ej * (1.0 + xir * (p->prm[id_i - 1].fafb[id_i - 1][1] +
                              xir * (p->prm[id_i - 1].fafb[id_i - 1][2] +
                                     xir * p->prm[id_i - 1].fafb[id_i - 1][3]))) / r;
  } else {
    double ei = exp(-2.0 * xi_i * r), ei_r = ei / r;
    afb_i = -xi_i * ei - ei_r;
    fifj = -ei * (c[0] + c[2] / r) - ej * (c[1] + c[3] / r);
  }
  nlist[li]++; nlist[lj]++;
  if (nlist[li] > ORC_NNB_MAX || nlist[lj] > ORC_NNB_MAX) return ORC_ERR_NB_OVERFLOW;
  size_t ki = (size_t)ORC_NNB_MAX * li + nlist[li] - 1;
  size_t kj = (size_t)ORC_NNB_MAX * lj + nlist[lj] - 1;
  list[ki] = lj; list[kj] = li;
  asym[ki] = KC * (afb_i - fifj);
  asym[kj] = KC * (afb_j - fifj);
  double tmp = KC * (fifj + erfc(ar) / r);
  sym[ki] = tmp; sym[kj] = tmp;
  return ORC_OK;
}

int orc_ctip_matrix(const orc_sys* s, const orc_params* p, const int* id, const double* x,
                    const int* npt, const int* link, const int* pair, int* nlist, int* list,
                    double* asym, double* sym) {
  int ncall = s->ncell[0] * s->ncell[1] * s->ncell[2];
  memset(nlist, 0, sizeof(int) * (size_t)s->n);
  for (int m = 0; m < ncall; ++m) {
    int nl = npt[m];
    const int* lk = link + (size_t)ORC_NPT_MAX * m;
    for (int i = 0; i < nl - 1; ++i)
      for (int j = i + 1; j < nl; ++j) {
        int e = ctip_pair(s, p, id, x, lk[i], lk[j], nlist, list, asym, sym);
        if (e) return e;
      }
    for (int k = 0; k < 13; ++k) {
      int cp = pair[13 * m + k];
      const int* lp = link + (size_t)ORC_NPT_MAX * cp;
      for (int j = 0; j < npt[cp]; ++j)
        for (int i = 0; i < nl; ++i) {
          int e = ctip_pair(s, p, id, x, lk[i], lp[j], nlist, list, asym, sym);
          if (e) return e;
        }
    }
  }
  return ORC_OK;
}

void orc_dudq(int n, const double* xjz, const double* asym, const double* sym, const int* list,
              const int* nlist, const double* q, const double* qmx, const double* phi, double* g) {
  /* ctip.f90:307-333 */
  for (int i = 0; i < n; ++i) {
    double dqmin = q[i] - qmx[2 * i], dqmax = q[i] - qmx[2 * i + 1];
    if (dqmin >= 0.0) dqmin = 0.0;
    if (dqmax <= 0.0) dqmax = 0.0;
    g[i] = -(xjz[3 * i] + xjz[3 * i + 1] * q[i] + phi[i] + 2.0 * WPEN * (dqmin + dqmax));
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < nlist[i]; ++j) {
      size_t k = (size_t)ORC_NNB_MAX * i + j;
      int jp = list[k];
      g[i] = g[i] - xjz[3 * jp + 2] * asym[k] - q[jp] * sym[k];
    }
}

typedef struct {
  const orc_sys* s; const double* x; const double* psi;
  const double* xjz; const double* qmx; const double* asym; const double* sym;
  const int* list; const int* nlist; double c_pme; double* phi; int n_eval;
} ctip_ws;

static void grad(ctip_ws* w, const double* q, double* g) {
  orc_phi_mesh(w->s, w->x, q, w->psi, w->phi);
  for (int i = 0; i < w->s->n; ++i) w->phi[i] *= w->c_pme;
  orc_dudq(w->s->n, w->xjz, w->asym, w->sym, w->list, w->nlist, q, w->qmx, w->phi, g);
  w->n_eval++;
}

/* BRENT_SOLVER, ctip.f90:216-304: the scalar root search of
   f(alpha) = -g(q + alpha p) . p on [-1, 1], each f a callback.  Its bracket
   bookkeeping, secant / inverse-quadratic steps and its two tolerances
   (E_crit = 1e-8 on the half bracket, E_end = 1e-11 on |f|) follow the
   reference line for line; *n_eval counts the callbacks. */
int orc_brent(orc_line_fn f, void* ctx, double* alpha, int* n_eval) {
  const double E_crit = 1.0e-8, E_end = 1.0e-11;
  double fa, fb, fc, x, y, z, a, b, c, d = 0.0, e = 0.0, xm, sgn;
  int ne = 0;
  a = -1.0;
  fa = f(ctx, a); ne++;
  b = 1.0;
  fb = f(ctx, b); ne++;
  c = b; fc = fb;
  for (int it = 1; it <= 100; ++it) {
    if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { c = a; fc = fa; d = b - a; e = d; }
    if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    xm = 0.5 * (c - b);
    if (fabs(xm) < E_crit || fabs(fb) < E_end) {
      *alpha = b;
      if (n_eval) *n_eval = ne;
      return ORC_OK;
    }
    if (fabs(e) >= E_crit && fabs(fa) > fabs(fb)) {
      double s = fb / fa;
      if (a == c) { x = 2.0 * xm * s; y = 1.0 - s; }
      else {
        y = fa / fc; z = fb / fc;
        x = s * (2.0 * xm * y * (y - z) - (b - a) * (z - 1.0));
        y = (y - 1.0) * (z - 1.0) * (s - 1.0);
      }
      if (x > 0.0) y = -y;
      x = fabs(x);
      double m1 = 3.0 * xm * y - fabs(E_crit * y), m2 = fabs(e * y);
      if (2.0 * x < (m1 < m2 ? m1 : m2)) { e = d; d = x / y; }
      else { d = xm; e = d; }
    } else { d = xm; e = d; }
    a = b; fa = fb;
    sgn = copysign(E_crit, xm);
    if (fabs(d) > E_crit) b = b + d; else b = b + sgn;
    fb = f(ctx, b); ne++;
  }
  if (n_eval) *n_eval = ne;
  return ORC_ERR_BRENT;
}

/* The nonlinear CG loop of CTIP_SOLVER, ctip.f90:167-190, on a gradient
   callback grad(ctx, q, g) that returns g = -dU/dq (dUdq's sign):
   beta = sum(g) / sum(g_old) (plain sums, not squared norms; quirk Q6),
   d = g + beta p, p = d - mean(d), alpha from BRENT_SOLVER on
   f(alpha) = -g(q + alpha p) . p, q += alpha p, stop when
   max|p| |alpha| <= 1e-3 (E_crit, ctip.f90:26).  maxit > 0 stops after that
   many iterations as well (0: no cap).  *n_eval counts gradient calls. */
typedef struct {
  int n; orc_grad_fn grad; void* ctx; const double* q; const double* p; double* ql; double* g;
  int n_eval;
} line_ws;

static double line_f(void* vw, double alpha) {
  line_ws* w = (line_ws*)vw;
  for (int i = 0; i < w->n; ++i) w->ql[i] = w->q[i] + alpha * w->p[i];
  w->grad(w->ctx, w->ql, w->g);
  w->n_eval++;
  double fb = 0.0;
  for (int i = 0; i < w->n; ++i) fb = fb - w->g[i] * w->p[i];
  return fb;
}

int orc_ncg(int n, orc_grad_fn grad, void* ctx, double* qi, int maxit, int* iters, int* n_eval) {
  double* g = malloc(sizeof(double) * n);
  double* d = malloc(sizeof(double) * n);
  double* pdir = calloc(n, sizeof(double));
  double* ql = malloc(sizeof(double) * n);
  double* gl = malloc(sizeof(double) * n);
  line_ws lw = {n, grad, ctx, qi, pdir, ql, gl, 0};
  int k = 0, err = ORC_OK, ne = 0;
  double Error = 1.0, g_sum = 1.0, g_sum_old, d_sum, alpha = 0.0;
  while (Error > 1.0e-3 && (maxit <= 0 || k < maxit)) {
    grad(ctx, qi, g); ne++;
    g_sum_old = g_sum; g_sum = 0.0;
    for (int i = 0; i < n; ++i) g_sum += g[i];
    for (int i = 0; i < n; ++i) d[i] = g[i] + pdir[i] * g_sum / g_sum_old;
    d_sum = 0.0;
    for (int i = 0; i < n; ++i) d_sum += d[i];
    d_sum = d_sum / (double)n;
    for (int i = 0; i < n; ++i) pdir[i] = d[i] - d_sum;
    int nb = 0;
    lw.n_eval = 0;
    err = orc_brent(line_f, &lw, &alpha, &nb);
    ne += lw.n_eval;
    if (err) break;
    for (int i = 0; i < n; ++i) qi[i] = qi[i] + alpha * pdir[i];
    double pm = 0.0;
    for (int i = 0; i < n; ++i) if (fabs(pdir[i]) > pm) pm = fabs(pdir[i]);
    Error = pm * fabs(alpha);
    k++;
  }
  *iters = k;
  if (n_eval) *n_eval = ne;
  free(g); free(d); free(pdir); free(ql); free(gl);
  return err;
}

/* PHI_MESH x c_pme + dUdq: the gradient CTIP_SOLVER's loop and BRENT_SOLVER
   evaluate (ctip.f90:170-171, 232-233) */
static void ctip_grad_cb(void* vw, const double* q, double* g) { grad((ctip_ws*)vw, q, g); }

static void xjz_fill(const orc_sys* s, const orc_params* p, const int* id, double* xjz,
                     double* qmx) {
  double kc_const = KC * 2.0 * s->a / SQRTPI;
  for (int i = 0; i < s->n; ++i) {
    const orc_potn* pp = &p->prm[id[i] - 1];
    xjz[3 * i] = pp->chi; xjz[3 * i + 1] = pp->J - kc_const; xjz[3 * i + 2] = pp->Z;
    qmx[2 * i] = pp->qmin; qmx[2 * i + 1] = pp->qmax;
  }
}

int orc_ctip_solve(const orc_sys* s, const orc_params* p, const int* id, const double* x,
                   const int* npt, const int* link, const int* pair, const double* psi,
                   double* q, double* eself, int* iters, int* n_eval) {
  const int n = s->n;
  double V = s->box[0] * s->box[1] * s->box[2];
  ctip_ws w;
  memset(&w, 0, sizeof w);
  double* xjz = malloc(sizeof(double) * 3 * n);
  double* qmx = malloc(sizeof(double) * 2 * n);
  int* nlist = calloc(n, sizeof(int));
  int* list = malloc(sizeof(int) * (size_t)n * ORC_NNB_MAX);
  double* asym = malloc(sizeof(double) * (size_t)n * ORC_NNB_MAX);
  double* sym = malloc(sizeof(double) * (size_t)n * ORC_NNB_MAX);
  double* phi = malloc(sizeof(double) * n);
  double* qi = malloc(sizeof(double) * n);
  xjz_fill(s, p, id, xjz, qmx);
  int err = orc_ctip_matrix(s, p, id, x, npt, link, pair, nlist, list, asym, sym);
  w.s = s; w.x = x; w.psi = psi; w.xjz = xjz; w.qmx = qmx; w.asym = asym; w.sym = sym;
  w.list = list; w.nlist = nlist; w.phi = phi;
  w.c_pme = KC * (double)s->fft[0] * (double)s->fft[1] * (double)s->fft[2] / V / PI_;
  for (int i = 0; i < n; ++i) qi[i] = q[i];
  int k = 0, ne = 0;
  if (!err) err = orc_ncg(n, ctip_grad_cb, &w, qi, 0, &k, &ne);
  double Eself = 0.0, Esum = 0.0;
  for (int i = 0; i < n; ++i) {
    q[i] = qi[i];
    Eself -= q[i] * q[i];
    Esum += xjz[3 * i] * qi[i] + 0.5 * xjz[3 * i + 1] * (qi[i] * qi[i]);
    double dq = qi[i] - qmx[2 * i];
    if (dq < 0.0) Esum += WPEN * (dq * dq);
    dq = qi[i] - qmx[2 * i + 1];
    if (dq > 0.0) Esum += WPEN * (dq * dq);
  }
  *eself = Eself * KC * s->a / SQRTPI + Esum;
  *iters = k;
  *n_eval = w.n_eval;
  free(xjz); free(qmx); free(nlist); free(list); free(asym); free(sym); free(phi); free(qi);
  return err;
}

double orc_ctip_energy(const orc_sys* s, const orc_params* p, const int* id, const double* x,
                       const int* npt, const int* link, const int* pair, const double* psi,
                       const double* q) {
  const int n = s->n;
  double V = s->box[0] * s->box[1] * s->box[2];
  double* xjz = malloc(sizeof(double) * 3 * n);
  double* qmx = malloc(sizeof(double) * 2 * n);
  int* nlist = calloc(n, sizeof(int));
  int* list = malloc(sizeof(int) * (size_t)n * ORC_NNB_MAX);
  double* asym = malloc(sizeof(double) * (size_t)n * ORC_NNB_MAX);
  double* sym = malloc(sizeof(double) * (size_t)n * ORC_NNB_MAX);
  double* phi = malloc(sizeof(double) * n);
  xjz_fill(s, p, id, xjz, qmx);
  orc_ctip_matrix(s, p, id, x, npt, link, pair, nlist, list, asym, sym);
  orc_phi_mesh(s, x, q, psi, phi);
  double c_pme = KC * (double)s->fft[0] * (double)s->fft[1] * (double)s->fft[2] / V / PI_;
  double U = 0.0;
  for (int i = 0; i < n; ++i) {
    U += xjz[3 * i] * q[i] + 0.5 * xjz[3 * i + 1] * q[i] * q[i] + 0.5 * c_pme * q[i] * phi[i];
    double dq = q[i] - qmx[2 * i];
    if (dq < 0.0) U += WPEN * dq * dq;
    dq = q[i] - qmx[2 * i + 1];
    if (dq > 0.0) U += WPEN * dq * dq;
    for (int j = 0; j < nlist[i]; ++j) {
      size_t k = (size_t)ORC_NNB_MAX * i + j;
      U += q[i] * xjz[3 * list[k] + 2] * asym[k] + 0.5 * q[i] * q[list[k]] * sym[k];
    }
  }
  free(xjz); free(qmx); free(nlist); free(list); free(asym); free(sym); free(phi);
  return U;
}

/* ------------------------------------------------------------------------ */
/* Velocity Verlet, vverlet.f90:74-96 and 132-157                            */
/* ------------------------------------------------------------------------ */
void orc_vv1(const orc_sys* s, const orc_params* p, const int* id, double* x, double* v,
             const double* f, double dt) {
  for (int i = 0; i < s->n; ++i) {
    double xm = p->prm[id[i] - 1].xm;
    for (int d = 0; d < 3; ++d) {
      v[3 * i + d] = v[3 * i + d] + 0.5 * dt * f[3 * i + d] / xm;
      double xn = x[3 * i + d] + dt * v[3 * i + d];
      x[3 * i + d] = xn - s->box[d] * round(xn / s->box[d]);
    }
  }
}

double orc_vv2(const orc_sys* s, const orc_params* p, const int* id, double* v, const double* f,
               double dt) {
  double mv2 = 0.0;
  for (int i = 0; i < s->n; ++i) {
    double xm = p->prm[id[i] - 1].xm, t = 0.0;
    for (int d = 0; d < 3; ++d) {
      v[3 * i + d] = v[3 * i + d] + 0.5 * dt * f[3 * i + d] / xm;
      t = t + v[3 * i + d] * v[3 * i + d];
    }
    mv2 = mv2 + xm * t;
  }
  return mv2;
}
