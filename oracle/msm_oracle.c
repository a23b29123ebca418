/*
 * msm_oracle.c -- CPU restatement of msmpb's B-spline multilevel summation.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path in
 * msmpb_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * Every function restates one piece of hkayabilisim/msmpb (cited as
 * file:line into that repository) in plain C, FP64, single-threaded.  The
 * reference keeps its level grids padded with ghost shells; this
 * restatement keeps them periodic and unpadded, which is the same
 * arithmetic inside the reference's valid parameter domain
 * (min(2*nLim+1, M_l) >= nu for every level l < L), and rejects inputs
 * outside it.  compat = 1 re-creates the two layout quirks of the
 * reference HEAD as explicit index maps (see energy() below), and the
 * HEAD's image range of the level kernels: kappaA sums the images p with
 * |s_a - p_a| < a_l * as_a where FF_rebuild passes the column norms of
 * A^-1 (forcefield0.c:618-620, 670, 1049-1051).  Every image within a_l
 * needs the row norm |row a of A^-1| instead; the two agree for
 * orthorhombic boxes, and in a triclinic box the column norms can drop
 * images (a sheared CsCl cell at L = 2 loses 1e-3 of its energy).
 * compat = 0 sums over max(row, column) norm: identical terms wherever the
 * HEAD's range already holds every image, the complete kernel otherwise.
 * The same holds for the pair term and the own-image constant coeff1 once
 * the cutoff reaches half a face-to-face width of a triclinic box: the
 * reference lists a pair by its fractionally nearest image and sums the
 * images within plim (column norms) of it (forcefield0.c:109-160,
 * forcefield1.c:240-321, forcefield0.c:612-643); compat = 0 lists every
 * pair with any image inside the cutoff and bounds the images by
 * max(row, column) norm, compat = 1 keeps the reference's rules.  Below half
 * the smallest width, and in orthorhombic boxes, the two coincide.
 */
#include "msm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 12

struct orc {
  int N;
  double A[3][3], Ai[3][3], detA;
  int nu, L, M[3];          /* M = top-level grid dims */
  double a0, estErr;
  double aCut[MAXL + 1];
  double relCutoff;
  int nLim;
  double tau[10];           /* nu <= 10 */
  double Q[5 * 10];         /* (nu/2) x nu piece coefficients */
  double J[11];             /* J[n + nu/2], n = -nu/2 .. nu/2 */
  double *omp[MAXL + 1][3]; /* omega' per level and axis */
  int opD[MAXL + 1][3];
  double *cL[3];
  int gd[MAXL + 1][3];      /* level grid dims */
  int sd[MAXL + 1][3];      /* stencil dims (l < L) */
  double *khat[MAXL + 1];
  double coeff1, coeff2;
  int compat;
  int fft;                  /* 0: the reference's DFT build (FF_set_FFT(false)) */
  const double *qref;       /* charges used by the error estimators */
  int pair_count;
  long *pstart;             /* CSR pair list (j > i) */
  int *plist;
  /* last evaluation's periodic grids */
  double *qg[MAXL + 1], *eg[MAXL + 1];
};

static double pi_(void) { return 4. * atan(1.); }
static int pmod(int a, int m) { int r = a % m; return r < 0 ? r + m : r; }

/* invert(): adjugate / determinant, returns |det| (forcefield0.c:709-724). */
static double inv3(const double a[3][3], double ai[3][3]) {
  double adj[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      adj[i][j] = a[(j + 1) % 3][(i + 1) % 3] * a[(j + 2) % 3][(i + 2) % 3] -
                  a[(j + 2) % 3][(i + 1) % 3] * a[(j + 1) % 3][(i + 2) % 3];
  double det = a[0][0] * adj[0][0] + a[0][1] * adj[1][0] + a[0][2] * adj[2][0];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) ai[i][j] = adj[i][j] / det;
  return fabs(det);
}

/* increment(): next 7-smooth integer >= M (forcefield0.c:219-245). */
static int smooth7(int M) {
  for (;; M++) {
    int r = M;
    for (int f = 2; f <= 7 && r > 1;) {
      if (r % f == 0) r /= f; else f++;
    }
    if (r == 1) return M;
  }
}

static double rownorm(const double m[3][3], int r) {
  return sqrt(m[r][0] * m[r][0] + m[r][1] * m[r][1] + m[r][2] * m[r][2]);
}
static double colnorm(const double m[3][3], int c) {
  return sqrt(m[0][c] * m[0][c] + m[1][c] * m[1][c] + m[2][c] * m[2][c]);
}

/* Image-range vector of the pair term and coeff1: the reference's column
 * norms of A^-1 (compat), or max(row, column) norm, which bounds every
 * lattice shift inside a radius r by |p_a| <= r |row a of A^-1| + 1/2. */
static double image_norm(const struct orc *o, int a) {
  double c = colnorm(o->Ai, a);
  return o->compat ? c : fmax(c, rownorm(o->Ai, a));
}

/* _FF_get_errEst (forcefield0.c:553-582), with C = 1. */
double orc_errest(const orc_t *o, int nu) {
  int L = o->L;
  double ax = rownorm(o->A, 0), ay = rownorm(o->A, 1), az = rownorm(o->A, 2);
  int M1[3] = {o->M[0], o->M[1], o->M[2]};
  for (int i = 1; i < L; i++) M1[0] *= 2, M1[1] *= 2, M1[2] *= 2;
  double Q2 = 0;
  for (int i = 0; i < o->N; i++) Q2 += o->qref[i] * o->qref[i];
  double Fref = (Q2 / (double)o->N) / pow(o->detA / (double)o->N, 5. / 6.);
  double h1 = fmax(fmax(ax / M1[0], ay / M1[1]), az / M1[2]);
  double dF = 1.0 * Q2 * (1 - pow(4.0, -L)) * pow(h1, nu - 1) *
              pow(o->detA, -1. / 3.) * pow(o->N, -2.0 / 3.0);
  return dF / Fref;
}

/* FF_get_deltaF (forcefield0.c:521-551). */
double orc_deltaF(const orc_t *o) {
  static const double C[4] = {0.825169, 5.16701, 73.6508, 1977.23};
  int L = o->L, nu = o->nu;
  double ax = rownorm(o->A, 0), ay = rownorm(o->A, 1), az = rownorm(o->A, 2);
  int M1[3] = {o->M[0], o->M[1], o->M[2]};
  for (int i = 1; i < L; i++) M1[0] *= 2, M1[1] *= 2, M1[2] *= 2;
  double Q2 = 0;
  for (int i = 0; i < o->N; i++) Q2 += o->qref[i] * o->qref[i];
  double h1 = fmax(fmax(ax / M1[0], ay / M1[1]), az / M1[2]);
  return C[(nu - 4) / 2] * Q2 * (1 - pow(4.0, -L)) * pow(h1, nu - 1) *
         pow(o->detA, -1. / 3.) * pow(o->N, -2.0 / 3.0) / pow(o->aCut[0], nu);
}

/* gamma(rho) = tau(rho^2 - 1) by Horner (forcefield1.c:123-127). */
static double tau_poly(const orc_t *o, double s) {
  double g = o->tau[o->nu - 1];
  for (int k = o->nu - 2; k >= 0; k--) g = o->tau[k] + s * g;
  return g;
}

/* Bspline / Bspline1 (forcefield1.c:609-637): piece i in [0,nu), t in [0,1),
 * symmetric half evaluated by reflection. */
double orc_bspline(const orc_t *o, int i, double t, double *deriv) {
  int nu = o->nu;
  double sign = 1.;
  if (i >= nu / 2) { i = nu - 1 - i; t = 1. - t; sign = -1.; }
  const double *c = o->Q + i * nu;
  double v = c[nu - 1], d = 0.;
  for (int k = nu - 2; k >= 0; k--) { d = v + t * d; v = c[k] + t * v; }
  if (deriv) *deriv = sign * d;
  return v;
}

/* Tables built in FF_build (forcefield0.c:369-416). */
static void build_tables(orc_t *o) {
  int nu = o->nu, nk = nu / 2;
  /* tau_k: Taylor coefficients of (1+s)^(-1/2) by the reference's recursion */
  o->tau[0] = 1.;
  for (int k = 1; k < nu; k++) o->tau[k] = -(1. - 0.5 / (double)k) * o->tau[k - 1];
  /* B-spline knot values of orders 1..nu, then scaled divided differences
   * turn row nu-1-k into the k-th derivative / k! of the order-nu spline */
  double V[10][5];
  memset(V, 0, sizeof V);
  V[0][0] = 1.;
  for (int i = 1; i < nu; i++) {
    V[i][0] = 0.;
    for (int j = 1; j < nk; j++)
      V[i][j] = ((double)j * V[i - 1][j] + (double)(i + 1 - j) * V[i - 1][j - 1]) / (double)i;
  }
  for (int k = 1; k < nu; k++)
    for (int i = 0; i < nu - k; i++) {
      for (int j = nk - 1; j > 0; j--) V[i][j] = (V[i][j] - V[i][j - 1]) / (double)k;
      V[i][0] = V[i][0] / (double)k;
    }
  for (int j = 0; j < nk; j++)
    for (int k = 0; k < nu; k++) o->Q[j * nu + k] = V[nu - 1 - k][j];
  /* two-scale stencil by repeated Pascal sweeps, then 2^(1-nu) */
  double *J0 = o->J + nu / 2;
  for (int n = 0; n <= nu; n++) o->J[n] = 0.;
  J0[0] = 1.;
  for (int i = 1; i <= nu / 2; i++) {
    for (int j = -i; j < i; j++) J0[j] += J0[j + 1];
    for (int j = i; j > -i; j--) J0[j] += J0[j - 1];
  }
  for (int n = -nu / 2; n <= nu / 2; n++) J0[n] *= pow(2., 1 - nu);
}

/* omegap(): anti-blurring coefficients per level and axis
 * (forcefield0.c:726-767). */
static void build_omegap(orc_t *o) {
  int nu = o->nu, h = nu / 2;
  double Phi[5] = {0};
  for (int i = 1; i < h; i++) Phi[h - i] = o->Q[i * nu];
  for (int k = 0; k < nu; k++) Phi[0] += o->Q[(h - 1) * nu + k];
  int M[3] = {o->M[0], o->M[1], o->M[2]};
  double pi = pi_();
  int opLim = 2 * o->nLim;
  for (int l = o->L; l >= 1; l--) {
    for (int a = 0; a < 3; a++) {
      int Ma = M[a];
      double *c = (double *)malloc(Ma * sizeof(double));
      for (int k = 0; k < Ma; k++) {
        double th = 2. * pi * (double)k / (double)Ma;
        double v = Phi[0];
        for (int n = 1; n <= h - 1; n++) v += 2. * Phi[n] * cos((double)n * th);
        v *= v;
        c[k] = 1. / v;
      }
      int D = Ma / 2 + 1;
      if (Ma / 2 > opLim) D = opLim + 1;
      o->opD[l][a] = D;
      o->omp[l][a] = (double *)calloc(D, sizeof(double));
      for (int m = 0; m < D; m++) {
        double w = c[0];
        for (int k = 1; k <= (Ma - 1) / 2; k++)
          w += 2. * c[k] * cos(2. * pi * (double)(k * m) / (double)Ma);
        if (Ma % 2 == 0) w += c[Ma / 2] * cos(pi * (double)m);
        o->omp[l][a][m] = w / (double)Ma;
      }
      if (l == o->L) o->cL[a] = c; else free(c);
    }
    M[0] *= 2, M[1] *= 2, M[2] *= 2;
  }
}

/* Fourier-space top-level kernel psi(k) before the cL and M^3 factors
 * (dAL, forcefield0.c:769-865). */
static double dal_psi(const orc_t *o, int kx, int ky, int kz) {
  static const double tc[4][7] = {
      {105.0, -1890.0, 83160.0, -6486480.0, 778377600.0, -132324192000.0, 30169915776000.0},
      {10395.0, -270270.0, 16216200.0, -1654052400.0, 251415964800.0, -52797352608000.0,
       14572069319808000.0},
      {2027025.0, -68918850.0, 5237832600.0, -659966907600.0, 121433910998400.0,
       -30358477749600000.0, 9836146790870400000.0},
      {654729075.0, -27498621150.0, 2529873145800.0, -379480971870000.0, 81967889923920000.0,
       -23770688077936800000.0, 8842695964992489600000.0}};
  const double(*Ai)[3] = o->Ai;
  int nu = o->nu;
  double pi = pi_(), aL = o->aCut[o->L];
  double ax = Ai[0][0] * kx + Ai[1][0] * ky + Ai[2][0] * kz;
  double ay = Ai[0][1] * kx + Ai[1][1] * ky + Ai[2][1] * kz;
  double az = Ai[0][2] * kx + Ai[1][2] * ky + Ai[2][2] * kz;
  double k2 = ax * ax + ay * ay + az * az;
  if (k2 == 0) return 0.0;
  double k = sqrt(k2), t = pi * k * aL, t2 = t * t, t4 = t2 * t2, t6 = t4 * t2, t8 = t6 * t2;
  double sf = 0, cf = 0;
  if (nu == 4) { sf = 15.0 - 6.0 * t2; cf = 15.0 - 1.0 * t2; }
  else if (nu == 6) { sf = 945.0 - 420.0 * t2 + 15.0 * t4; cf = 945.0 - 105.0 * t2 + 1.0 * t4; }
  else if (nu == 8) {
    sf = 135135.0 - 62370.0 * t2 + 3150.0 * t4 - 28.0 * t6;
    cf = 135135.0 - 17325.0 * t2 + 378.0 * t4 - 1.0 * t6;
  } else if (nu == 10) {
    sf = 34459425.0 - 16216200.0 * t2 + 945945.0 * t4 - 13860.0 * t6 + 45.0 * t8;
    cf = 34459425.0 - 4729725.0 * t2 + 135135.0 * t4 - 990.0 * t6 + 1.0 * t8;
  }
  double psi;
  if (t <= 4 * pi / 4.0) { /* tmax = 4, ntc = 7 (forcefield0.c:272-273) */
    double s = 0.0, p = 1.0;
    for (int i = 0; i < 7; i++) { s += p / tc[nu / 2 - 2][i]; p *= t2; }
    psi = s / t2;
  } else {
    psi = (sf * (sin(t) / t) - cf * cos(t));
    psi = psi / pow(t, 2.0 * nu);
  }
  double dfac = 1.0;
  for (int i = 1; i <= 2 * nu - 1; i += 2) dfac *= i;
  return psi * (dfac * pi * aL * aL / o->detA);
}

/* khat[L] (forcefield0.c:867-884). */
static void build_top_kernel(orc_t *o) {
  int Mx = o->M[0], My = o->M[1], Mz = o->M[2];
  double *kh = (double *)calloc((size_t)Mx * My * Mz, sizeof(double));
  for (int kx = 0; kx < Mx; kx++) {
    double cx = o->cL[0][kx];
    int ckx = kx < (Mx + 1) / 2 ? kx : kx - Mx;
    for (int ky = 0; ky < My; ky++) {
      double cxy = cx * o->cL[1][ky];
      int cky = ky < (My + 1) / 2.0 ? ky : ky - My;
      for (int kz = 0; kz < Mz; kz++) {
        double c = cxy * o->cL[2][kz];
        double p = kz < (Mz + 1) / 2 ? dal_psi(o, ckx, cky, kz) : dal_psi(o, -ckx, -cky, Mz - kz);
        kh[((size_t)kx * My + ky) * Mz + kz] += p * c * Mx * My * Mz;
      }
    }
  }
  o->khat[o->L] = kh;
}

/* kappaA (forcefield0.c:1041-1081): periodic level-l kernel at fractional
 * offset s, gamma_{a_{l-1}} - gamma_{a_l} summed over images within a_l. */
static double kappa_l(const orc_t *o, int l, const double s[3], const double as[3]) {
  double al = o->aCut[l], kap = 0.;
  double pmin[3], pmax[3];
  for (int a = 0; a < 3; a++) {
    pmin[a] = floor(s[a] - al * as[a] + 1.);
    pmax[a] = ceil(s[a] + al * as[a] - 1.);
  }
  for (double px = pmin[0]; px <= pmax[0]; px++)
    for (double py = pmin[1]; py <= pmax[1]; py++)
      for (double pz = pmin[2]; pz <= pmax[2]; pz++) {
        double x = s[0] - px, y = s[1] - py, z = s[2] - pz;
        double rx = o->A[0][0] * x + o->A[0][1] * y + o->A[0][2] * z;
        double ry = o->A[1][0] * x + o->A[1][1] * y + o->A[1][2] * z;
        double rz = o->A[2][0] * x + o->A[2][1] * y + o->A[2][2] * z;
        double rho = sqrt(rx * rx + ry * ry + rz * rz) / al;
        if (rho >= 1.) continue;
        double g1 = tau_poly(o, rho * rho - 1.) / al;
        rho *= 2.;
        double g2 = rho >= 1. ? 1. / rho : tau_poly(o, rho * rho - 1.);
        g2 *= 2. / al;
        kap += g2 - g1;
      }
  return kap;
}

/* kaphatA (forcefield0.c:965-1039): kappa on the s_l^3 offsets, then omega'
 * applied along z, y and x with the reference's periodic offset map. */
static void build_level_kernel(orc_t *o, int l, const double as[3]) {
  const int *g = o->gd[l], *sd = o->sd[l];
  int kx = sd[0], ky = sd[1], kz = sd[2]; /* kd == sd (both min(dmax, M_l)) */
  double *kap = (double *)malloc((size_t)kx * ky * kz * sizeof(double));
  for (int i = -kx / 2; i <= (kx - 1) / 2; i++)
    for (int j = -ky / 2; j <= (ky - 1) / 2; j++)
      for (int k = -kz / 2; k <= (kz - 1) / 2; k++) {
        double s[3] = {(double)i / g[0], (double)j / g[1], (double)k / g[2]};
        kap[((size_t)(i + kx / 2) * ky + j + ky / 2) * kz + k + kz / 2] = kappa_l(o, l, s, as);
      }
  double *const *op = o->omp[l];
  double *Fz = (double *)calloc((size_t)kx * ky * kz, sizeof(double));
  double *Gy = (double *)calloc((size_t)kx * ky * kz, sizeof(double));
  double *kh = (double *)calloc((size_t)kx * ky * kz, sizeof(double));
#define OFF(a, M) ((((a) + (3 * (M)) / 2) % (M)) - (M) / 2)
  for (int i0 = 0; i0 < kx; i0++)
    for (int j0 = 0; j0 < ky; j0++)
      for (int k1 = -kz / 2; k1 <= (kz - 1) / 2; k1++) {
        double sum = 0.0;
        for (int k0 = -kz / 2; k0 <= (kz - 1) / 2; k0++)
          sum += op[2][abs(OFF(k0 - k1, g[2]))] * kap[((size_t)i0 * ky + j0) * kz + k0 + kz / 2];
        Fz[((size_t)i0 * ky + j0) * kz + k1 + kz / 2] = sum;
      }
  for (int i0 = 0; i0 < kx; i0++)
    for (int j1 = -ky / 2; j1 <= (ky - 1) / 2; j1++)
      for (int k1 = 0; k1 < kz; k1++) {
        double sum = 0.0;
        for (int j0 = -ky / 2; j0 <= (ky - 1) / 2; j0++)
          sum += op[1][abs(OFF(j0 - j1, g[1]))] * Fz[((size_t)i0 * ky + j0 + ky / 2) * kz + k1];
        Gy[((size_t)i0 * ky + j1 + ky / 2) * kz + k1] = sum;
      }
  for (int i1 = -kx / 2; i1 <= (kx - 1) / 2; i1++)
    for (int j1 = 0; j1 < ky; j1++)
      for (int k1 = 0; k1 < kz; k1++) {
        double sum = 0.0;
        for (int i0 = -kx / 2; i0 <= (kx - 1) / 2; i0++)
          sum += op[0][abs(OFF(i0 - i1, g[0]))] * Gy[((size_t)(i0 + kx / 2) * ky + j1) * kz + k1];
        kh[((size_t)(i1 + kx / 2) * ky + j1) * kz + k1] = sum;
      }
#undef OFF
  free(kap); free(Fz); free(Gy);
  o->khat[l] = kh;
}

/* determineM (forcefield0.c:247-269): eta grid points per atom on the
 * finest level, equal spacing on every axis; eta = 5.7 (L = 1) or 0.83
 * (L >= 2) for the FFT build, 1 for the DFT build. */
static void determine_M(const orc_t *o, int L, int M[3]) {
  double eta = o->fft ? (L >= 2 ? 0.83 : 5.7) : 1.0;
  double ax = rownorm(o->A, 0), ay = rownorm(o->A, 1), az = rownorm(o->A, 2);
  double av[3] = {ax, ay, az};
  for (int a = 0; a < 3; a++)
    M[a] = (int)ceil(pow(2., 1. - L) * pow(eta * o->N, 1. / 3.) * av[a] *
                     pow(ax * ay * az, -1. / 3.));
}

orc_t *orc_build(int N, const double *q, const double *pos, const double *edges, int nu,
                 int L, const int *Mtop, double a0, double estErr, int compat, int fft,
                 char *err) {
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->N = N;
  o->qref = q;
  o->compat = compat;
  o->fft = fft;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o->A[i][j] = edges[3 * i + j];
  o->detA = inv3(o->A, o->Ai);
  /* parameter rules of FF_build (forcefield0.c:286-367) */
  o->estErr = estErr > 0 ? estErr : 0.001;
  o->L = L > 0 ? L : 1;
  if (Mtop && Mtop[0] > 0) {
    o->M[0] = Mtop[0]; o->M[1] = Mtop[1]; o->M[2] = Mtop[2];
  } else if (!fft) {
    /* DFT build (forcefield0.c:292-307): the first level count whose
     * determineM grid (eta = 1) has at most sqrt(N) points */
    for (int l = 1;; l++) {
      determine_M(o, l, o->M);
      if ((double)o->M[0] * o->M[1] * o->M[2] <= sqrt((double)N)) { o->L = l; break; }
    }
  } else {
    determine_M(o, o->L, o->M);
  }
  if (o->L > MAXL) { if (err) sprintf(err, "maxLevel > %d", MAXL); free(o); return NULL; }
  if (fft) /* FFT sizes (forcefield0.c:309-316) */
    for (int a = 0; a < 3; a++) o->M[a] = smooth7(o->M[a]);
  if (a0 <= 0) {
    if (nu <= 0) {
      double best = INFINITY;
      for (int t = 4; t <= 10; t += 2) {
        double c = pow(orc_errest(o, t) / o->estErr, 1. / t);
        if (c < best) best = c;
      }
      a0 = best;
    } else {
      a0 = pow(orc_errest(o, nu) / o->estErr, 1. / nu);
    }
  }
  if (nu <= 0) {
    double best = INFINITY;
    int bnu = 4;
    for (int t = 4; t <= 10; t += 2) {
      double e = orc_errest(o, t) / pow(a0, t);
      if (e < best) { best = e; bnu = t; }
    }
    nu = bnu;
  }
  if (nu < 4 || nu > 10 || nu % 2) {
    if (err) sprintf(err, "orderAcc must be one of 4, 6, 8, 10 (got %d)", nu);
    free(o); return NULL;
  }
  o->nu = nu;
  o->a0 = a0;
  o->aCut[0] = a0;
  for (int l = 1; l <= o->L; l++) o->aCut[l] = 2.0 * o->aCut[l - 1];
  int nbar = 0;
  for (int a = 0; a < 3; a++) {
    int nb = (int)ceil(o->aCut[o->L] * o->M[a] * rownorm(o->Ai, a));
    if (nb > nbar) nbar = nb;
  }
  o->relCutoff = nbar;
  o->nLim = (int)ceil(o->relCutoff - 1.);
  int dmax = 2 * o->nLim + 1;
  for (int a = 0; a < 3; a++) o->gd[o->L][a] = o->M[a];
  for (int l = o->L - 1; l >= 1; l--)
    for (int a = 0; a < 3; a++) {
      o->gd[l][a] = 2 * o->gd[l + 1][a];
      o->sd[l][a] = dmax < o->gd[l][a] ? dmax : o->gd[l][a];
      if (o->sd[l][a] < nu) {
        if (err)
          sprintf(err, "level %d stencil extent %d < nu=%d: outside the reference's valid domain",
                  l, o->sd[l][a], nu);
        free(o); return NULL;
      }
    }
  for (int a = 0; a < 3; a++) o->sd[o->L][a] = o->M[a];
  build_tables(o);
  build_omegap(o);
  /* FF_rebuild: constant-energy coefficients (forcefield0.c:612-643) */
  double as[3] = {colnorm(o->Ai, 0), colnorm(o->Ai, 1), colnorm(o->Ai, 2)};
  double lim[3];
  for (int a = 0; a < 3; a++) lim[a] = ceil(a0 * image_norm(o, a) - 1.);
  o->coeff1 = 0.;
  for (double x = -lim[0]; x <= lim[0]; x++)
    for (double y = -lim[1]; y <= lim[1]; y++)
      for (double z = -lim[2]; z <= lim[2]; z++) {
        if (x * x + y * y + z * z == 0) continue;
        double rx = o->A[0][0] * x + o->A[0][1] * y + o->A[0][2] * z;
        double ry = o->A[1][0] * x + o->A[1][1] * y + o->A[1][2] * z;
        double rz = o->A[2][0] * x + o->A[2][1] * y + o->A[2][2] * z;
        double r2 = rx * rx + ry * ry + rz * rz;
        if (r2 >= a0 * a0) continue;
        if (r2 != 0.) o->coeff1 += 1. / sqrt(r2);
        o->coeff1 -= tau_poly(o, r2 / (a0 * a0) - 1.) / a0;
      }
  o->coeff2 = pi_() * o->aCut[o->L] * o->aCut[o->L] / ((4 * nu + 2) * o->detA);
  build_top_kernel(o);
  double kas[3];
  for (int a = 0; a < 3; a++) kas[a] = o->compat ? as[a] : fmax(as[a], rownorm(o->Ai, a));
  for (int l = o->L - 1; l >= 1; l--) build_level_kernel(o, l, kas);
  for (int l = 1; l <= o->L; l++) {
    size_t n = (size_t)o->gd[l][0] * o->gd[l][1] * o->gd[l][2];
    o->qg[l] = (double *)calloc(n, sizeof(double));
    o->eg[l] = (double *)calloc(n, sizeof(double));
  }
  if (pos) orc_build_pairs(o, pos, 0.0); /* FF_rebuild -> neighborlist, margin 0 */
  return o;
}

void orc_free(orc_t *o) {
  if (!o) return;
  for (int l = 1; l <= o->L; l++) {
    for (int a = 0; a < 3; a++) free(o->omp[l][a]);
    free(o->khat[l]); free(o->qg[l]); free(o->eg[l]);
  }
  for (int a = 0; a < 3; a++) free(o->cL[a]);
  free(o->pstart); free(o->plist);
  free(o);
}

/* ---- neighbour list (forcefield0.c:21-164): all pairs i<j whose nearest
 * image lies within a0 + margin, stored CSR per i.  Pairs are found with a
 * cell grid of width >= a0 + margin (own enumeration; the reference hashes
 * into ceil(N^(1/3))^3 cells), which yields the same pair set.  pair_count
 * is that set's size; with compat = 0 the list also takes the pairs whose
 * nearest image lies outside but another image inside (a triclinic box
 * narrower than twice the radius; along such an axis the grid has one
 * cell, so every such pair is a candidate). */
int orc_build_pairs(orc_t *o, const double *pos, double margin) {
  int N = o->N;
  double ac = o->aCut[0] + margin, ac2 = ac * ac;
  double plim[3]; /* images around the nearest one that can lie inside ac */
  for (int a = 0; a < 3; a++) plim[a] = ceil(ac * image_norm(o, a) - 0.5);
  int nc[3];
  for (int a = 0; a < 3; a++) {
    double w = 1. / rownorm(o->Ai, a);
    nc[a] = (int)floor(w / ac);
    if (nc[a] < 1) nc[a] = 1;
  }
  int ncell = nc[0] * nc[1] * nc[2];
  int *head = (int *)malloc(ncell * sizeof(int)), *next = (int *)malloc(N * sizeof(int));
  int *cell = (int *)malloc(N * sizeof(int));
  for (int c = 0; c < ncell; c++) head[c] = -1;
  for (int i = N - 1; i >= 0; i--) {
    int ci[3];
    for (int a = 0; a < 3; a++) {
      double s = o->Ai[a][0] * pos[3 * i] + o->Ai[a][1] * pos[3 * i + 1] + o->Ai[a][2] * pos[3 * i + 2];
      s -= floor(s);
      ci[a] = (int)(s * nc[a]);
      if (ci[a] >= nc[a]) ci[a] = nc[a] - 1;
    }
    int c = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
    cell[i] = c;
    next[i] = head[c];
    head[c] = i;
  }
  free(o->pstart); free(o->plist);
  o->pstart = (long *)malloc(sizeof(long) * (N + 1));
  long cap = 1024 + 64L * N, len = 0;
  o->plist = (int *)malloc(sizeof(int) * cap);
  int pairs = 0;
  for (int i = 0; i < N; i++) {
    o->pstart[i] = len;
    int c = cell[i];
    int cx = c / (nc[1] * nc[2]), cy = (c / nc[2]) % nc[1], cz = c % nc[2];
    int lx[3], ly[3], lz[3], nx = 0, ny = 0, nz = 0;
    for (int d = -1; d <= 1; d++) { /* distinct neighbour cells per axis */
      int v = pmod(cx + d, nc[0]), dup = 0;
      for (int u = 0; u < nx; u++) dup |= lx[u] == v;
      if (!dup) lx[nx++] = v;
      v = pmod(cy + d, nc[1]); dup = 0;
      for (int u = 0; u < ny; u++) dup |= ly[u] == v;
      if (!dup) ly[ny++] = v;
      v = pmod(cz + d, nc[2]); dup = 0;
      for (int u = 0; u < nz; u++) dup |= lz[u] == v;
      if (!dup) lz[nz++] = v;
    }
    const double *ri = pos + 3 * i;
    for (int ux = 0; ux < nx; ux++)
      for (int uy = 0; uy < ny; uy++)
        for (int uz = 0; uz < nz; uz++) {
          int cn = (lx[ux] * nc[1] + ly[uy]) * nc[2] + lz[uz];
          for (int j = head[cn]; j >= 0; j = next[j]) {
            if (j <= i) continue;
            const double *rj = pos + 3 * j;
            double r[3] = {rj[0] - ri[0], rj[1] - ri[1], rj[2] - ri[2]}, p[3], d2 = 0.;
            for (int a = 0; a < 3; a++)
              p[a] = floor(o->Ai[a][0] * r[0] + o->Ai[a][1] * r[1] + o->Ai[a][2] * r[2] + 0.5);
            for (int a = 0; a < 3; a++) {
              double v = r[a] - (o->A[a][0] * p[0] + o->A[a][1] * p[1] + o->A[a][2] * p[2]);
              d2 += v * v;
            }
            int listed = d2 < ac2;
            pairs += listed;
            if (!o->compat && !listed) { /* another image inside ac */
              double v0[3];
              for (int a = 0; a < 3; a++)
                v0[a] = r[a] - (o->A[a][0] * p[0] + o->A[a][1] * p[1] + o->A[a][2] * p[2]);
              for (double px = -plim[0]; px <= plim[0] && !listed; px++)
                for (double py = -plim[1]; py <= plim[1] && !listed; py++)
                  for (double pz = -plim[2]; pz <= plim[2] && !listed; pz++) {
                    double e2 = 0.;
                    for (int a = 0; a < 3; a++) {
                      double v = v0[a] + (o->A[a][0] * px + o->A[a][1] * py + o->A[a][2] * pz);
                      e2 += v * v;
                    }
                    listed = e2 < ac2;
                  }
            }
            if (listed) {
              if (len == cap) { cap *= 2; o->plist = (int *)realloc(o->plist, sizeof(int) * cap); }
              o->plist[len++] = j;
            }
          }
        }
  }
  o->pstart[N] = len;
  o->pair_count = pairs;
  free(head); free(next); free(cell);
  return pairs;
}

/* ---- partcl2partcl (forcefield1.c:240-321): walk the pair list, sum the
 * images p in [-plim, plim]^3 around the nearest image with r < a0. */
static double short_range(orc_t *o, const double *pos, const double *q, double *F,
                          double *phi) {
  int N = o->N, nu = o->nu;
  double a0 = o->aCut[0], a02 = a0 * a0;
  double plim[3];
  for (int a = 0; a < 3; a++) plim[a] = ceil(a0 * image_norm(o, a) - 0.5);
  double a03 = pow(a0, 3);
  for (int i = 0; i < N; i++) { F[3 * i] = F[3 * i + 1] = F[3 * i + 2] = 0.; if (phi) phi[i] = 0.; }
  double energy = 0.;
  for (int i = 0; i < N; i++) {
    const double *ri = pos + 3 * i;
    for (long u = o->pstart[i]; u < o->pstart[i + 1]; u++) {
      int j = o->plist[u];
      const double *rj = pos + 3 * j;
      double r[3] = {rj[0] - ri[0], rj[1] - ri[1], rj[2] - ri[2]}, p[3], r0[3];
      for (int a = 0; a < 3; a++)
        p[a] = floor(o->Ai[a][0] * r[0] + o->Ai[a][1] * r[1] + o->Ai[a][2] * r[2] + 0.5);
      for (int a = 0; a < 3; a++)
        r0[a] = r[a] - (o->A[a][0] * p[0] + o->A[a][1] * p[1] + o->A[a][2] * p[2]);
      double sum = 0., f[3] = {0, 0, 0};
      for (double px = -plim[0]; px <= plim[0]; px++)
        for (double py = -plim[1]; py <= plim[1]; py++)
          for (double pz = -plim[2]; pz <= plim[2]; pz++) {
            double d[3];
            for (int a = 0; a < 3; a++)
              d[a] = r0[a] + (o->A[a][0] * px + o->A[a][1] * py + o->A[a][2] * pz);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 < a02) {
              double s = r2 / a02 - 1.;
              double ts = o->tau[nu - 1], t1 = 0.;
              for (int k = nu - 2; k >= 0; k--) { t1 = ts + s * t1; ts = o->tau[k] + s * ts; }
              sum += 1. / sqrt(r2) - ts / a0;
              double cf = -1. / (r2 * sqrt(r2)) - 2. * t1 / a03;
              f[0] += cf * d[0], f[1] += cf * d[1], f[2] += cf * d[2];
            }
          }
      double qq = q[i] * q[j];
      energy += qq * sum;
      for (int a = 0; a < 3; a++) { F[3 * i + a] += qq * f[a]; F[3 * j + a] -= qq * f[a]; }
      if (phi) { phi[i] += q[j] * sum; phi[j] += q[i] * sum; }
    }
  }
  return energy;
}

/* ---- anterpolate (forcefield1.c:324-368) onto the periodic level-1 grid */
static void anterpolate(const orc_t *o, const double *pos, const double *q, double *q1) {
  int nu = o->nu;
  const int *g = o->gd[1];
  memset(q1, 0, sizeof(double) * g[0] * g[1] * g[2]);
  double w[3][10];
  for (int i = 0; i < o->N; i++) {
    int m[3];
    for (int a = 0; a < 3; a++) {
      double s = o->Ai[a][0] * pos[3 * i] + o->Ai[a][1] * pos[3 * i + 1] + o->Ai[a][2] * pos[3 * i + 2];
      s = s - floor(s);
      double t = (double)g[a] * s, e = floor(t);
      t -= e;
      m[a] = (int)e;
      for (int n = -nu / 2; n < nu / 2; n++) w[a][n + nu / 2] = orc_bspline(o, n + nu / 2, t, NULL);
    }
    for (int nx = -nu / 2; nx < nu / 2; nx++)
      for (int ny = -nu / 2; ny < nu / 2; ny++)
        for (int nz = -nu / 2; nz < nu / 2; nz++) {
          int ix = pmod(m[0] - nx, g[0]), iy = pmod(m[1] - ny, g[1]), iz = pmod(m[2] - nz, g[2]);
          q1[((size_t)ix * g[1] + iy) * g[2] + iz] +=
              q[i] * w[0][nx + nu / 2] * w[1][ny + nu / 2] * w[2][nz + nu / 2];
        }
  }
}

/* ---- restrict_ (forcefield1.c:401-477): q^l_m = sum_n J J J q^{l-1}_{2m+n},
 * separable z -> y -> x, periodic. */
void orc_restrict(const orc_t *o, int l, const double *fine, double *coarse) {
  int nu = o->nu, h = nu / 2;
  const int *f = o->gd[l - 1], *c = o->gd[l];
  const double *J = o->J + h;
  double *T1 = (double *)malloc(sizeof(double) * f[0] * f[1] * c[2]);
  double *T2 = (double *)malloc(sizeof(double) * f[0] * c[1] * c[2]);
  for (int i = 0; i < f[0]; i++)
    for (int j = 0; j < f[1]; j++)
      for (int k = 0; k < c[2]; k++) {
        double s = 0.0;
        for (int n = -h; n <= h; n++) s += J[n] * fine[((size_t)i * f[1] + j) * f[2] + pmod(2 * k + n, f[2])];
        T1[((size_t)i * f[1] + j) * c[2] + k] = s;
      }
  for (int i = 0; i < f[0]; i++)
    for (int j = 0; j < c[1]; j++)
      for (int k = 0; k < c[2]; k++) {
        double s = 0.0;
        for (int n = -h; n <= h; n++) s += J[n] * T1[((size_t)i * f[1] + pmod(2 * j + n, f[1])) * c[2] + k];
        T2[((size_t)i * c[1] + j) * c[2] + k] = s;
      }
  for (int i = 0; i < c[0]; i++)
    for (int j = 0; j < c[1]; j++)
      for (int k = 0; k < c[2]; k++) {
        double s = 0.0;
        for (int n = -h; n <= h; n++) s += J[n] * T2[((size_t)pmod(2 * i + n, f[0]) * c[1] + j) * c[2] + k];
        coarse[((size_t)i * c[1] + j) * c[2] + k] = s;
      }
  free(T1); free(T2);
}

/* ---- prolongate (forcefield1.c:530-567): fine_{2m+n} += J_n coarse_m */
void orc_prolong(const orc_t *o, int l, const double *coarse, double *fine) {
  int nu = o->nu, h = nu / 2;
  const int *f = o->gd[l - 1], *c = o->gd[l];
  const double *J = o->J + h;
  double *T1 = (double *)calloc((size_t)c[0] * c[1] * f[2], sizeof(double));
  double *T2 = (double *)calloc((size_t)c[0] * f[1] * f[2], sizeof(double));
  for (int i = 0; i < c[0]; i++)
    for (int j = 0; j < c[1]; j++)
      for (int k = 0; k < c[2]; k++) {
        double v = coarse[((size_t)i * c[1] + j) * c[2] + k];
        for (int n = -h; n <= h; n++) T1[((size_t)i * c[1] + j) * f[2] + pmod(2 * k + n, f[2])] += J[n] * v;
      }
  for (int i = 0; i < c[0]; i++)
    for (int j = 0; j < c[1]; j++)
      for (int k = 0; k < f[2]; k++) {
        double v = T1[((size_t)i * c[1] + j) * f[2] + k];
        for (int n = -h; n <= h; n++) T2[((size_t)i * f[1] + pmod(2 * j + n, f[1])) * f[2] + k] += J[n] * v;
      }
  for (int i = 0; i < c[0]; i++)
    for (int j = 0; j < f[1]; j++)
      for (int k = 0; k < f[2]; k++) {
        double v = T2[((size_t)i * f[1] + j) * f[2] + k];
        for (int n = -h; n <= h; n++) fine[((size_t)pmod(2 * i + n, f[0]) * f[1] + j) * f[2] + k] += J[n] * v;
      }
  free(T1); free(T2);
}

/* ---- grid2grid (forcefield0.c:898-931): e_m += sum_k khat_k q_{m+k-s/2} */
void orc_grid2grid(const orc_t *o, int l, const double *ql, double *el) {
  const int *g = o->gd[l], *s = o->sd[l];
  const double *kh = o->khat[l];
  for (int ex = 0; ex < g[0]; ex++)
    for (int ey = 0; ey < g[1]; ey++)
      for (int ez = 0; ez < g[2]; ez++) {
        double acc = 0.;
        for (int kx = 0; kx < s[0]; kx++) {
          int qx = pmod(ex + kx - s[0] / 2, g[0]);
          for (int ky = 0; ky < s[1]; ky++) {
            int qy = pmod(ey + ky - s[1] / 2, g[1]);
            const double *krow = kh + ((size_t)kx * s[1] + ky) * s[2];
            const double *qrow = ql + ((size_t)qx * g[1] + qy) * g[2];
            for (int kz = 0; kz < s[2]; kz++) acc += krow[kz] * qrow[pmod(ez + kz - s[2] / 2, g[2])];
          }
        }
        el[((size_t)ex * g[1] + ey) * g[2] + ez] += acc;
      }
}

/* ---- top level: FFTg2g (forcefield1.c:499-528) with the FFTW transforms
 * restated as separable unnormalised DFTs (sign -1 forward, +1 backward). */
static void dft_axis(double *re, double *im, const int *g, int axis, int sign) {
  int M = g[axis];
  double *cr = (double *)malloc(sizeof(double) * M), *ci = (double *)malloc(sizeof(double) * M);
  double *tr = (double *)malloc(sizeof(double) * M), *ti = (double *)malloc(sizeof(double) * M);
  double tw = 8. * atan(1.) / M;
  for (int k = 0; k < M; k++) { cr[k] = cos(tw * k); ci[k] = sign * sin(tw * k); }
  size_t stride = axis == 0 ? (size_t)g[1] * g[2] : axis == 1 ? (size_t)g[2] : 1;
  int n1 = axis == 0 ? g[1] : g[0], n2 = axis == 2 ? g[1] : g[2];
  for (int a = 0; a < n1; a++)
    for (int b = 0; b < n2; b++) {
      size_t base;
      if (axis == 0) base = (size_t)a * g[2] + b;
      else if (axis == 1) base = (size_t)a * g[1] * g[2] + b;
      else base = ((size_t)a * g[1] + b) * g[2];
      for (int k = 0; k < M; k++) {
        double sr = 0, si = 0;
        for (int n = 0; n < M; n++) {
          int kn = (int)(((long)k * n) % M);
          double xr = re[base + n * stride], xi = im[base + n * stride];
          sr += xr * cr[kn] - xi * ci[kn];
          si += xr * ci[kn] + xi * cr[kn];
        }
        tr[k] = sr; ti[k] = si;
      }
      for (int k = 0; k < M; k++) { re[base + k * stride] = tr[k]; im[base + k * stride] = ti[k]; }
    }
  free(cr); free(ci); free(tr); free(ti);
}

void orc_top_level(const orc_t *o, const double *qL, double *eL) {
  const int *g = o->gd[o->L];
  size_t n = (size_t)g[0] * g[1] * g[2];
  double *re = (double *)malloc(sizeof(double) * n), *im = (double *)calloc(n, sizeof(double));
  memcpy(re, qL, sizeof(double) * n);
  for (int a = 0; a < 3; a++) dft_axis(re, im, g, a, -1);
  for (size_t i = 0; i < n; i++) { re[i] *= o->khat[o->L][i]; im[i] *= o->khat[o->L][i]; }
  for (int a = 0; a < 3; a++) dft_axis(re, im, g, a, +1);
  for (size_t i = 0; i < n; i++) eL[i] = re[i] / (double)(g[0] * g[1] * g[2]);
  free(re); free(im);
}

/* ---- interpolate (forcefield1.c:570-607) plus F += q E (:228-231); also
 * accumulates the grid potential into phi. */
static void interpolate(const orc_t *o, const double *pos, const double *q, const double *e1,
                        double *F, double *phi) {
  int nu = o->nu;
  const int *g = o->gd[1];
  double w[3][10], dw[3][10];
  for (int i = 0; i < o->N; i++) {
    int m[3];
    for (int a = 0; a < 3; a++) {
      double s = o->Ai[a][0] * pos[3 * i] + o->Ai[a][1] * pos[3 * i + 1] + o->Ai[a][2] * pos[3 * i + 2];
      double t = (double)g[a] * s, e = floor(t);
      t -= e;
      m[a] = (int)e;
      for (int n = -nu / 2; n < nu / 2; n++)
        w[a][n + nu / 2] = orc_bspline(o, n + nu / 2, t, &dw[a][n + nu / 2]);
    }
    double Ex = 0., Ey = 0., Ez = 0., ph = 0.;
    for (int nx = 0; nx < nu; nx++)
      for (int ny = 0; ny < nu; ny++)
        for (int nz = 0; nz < nu; nz++) {
          int ix = pmod(m[0] - (nx - nu / 2), g[0]), iy = pmod(m[1] - (ny - nu / 2), g[1]),
              iz = pmod(m[2] - (nz - nu / 2), g[2]);
          double v = e1[((size_t)ix * g[1] + iy) * g[2] + iz];
          Ex -= dw[0][nx] * w[1][ny] * w[2][nz] * v;
          Ey -= w[0][nx] * dw[1][ny] * w[2][nz] * v;
          Ez -= w[0][nx] * w[1][ny] * dw[2][nz] * v;
          ph += w[0][nx] * w[1][ny] * w[2][nz] * v;
        }
    Ex *= (double)g[0], Ey *= (double)g[1], Ez *= (double)g[2];
    /* prodT(Ai, E) (forcefield1.c:645-649) */
    double E[3] = {o->Ai[0][0] * Ex + o->Ai[1][0] * Ey + o->Ai[2][0] * Ez,
                   o->Ai[0][1] * Ex + o->Ai[1][1] * Ey + o->Ai[2][1] * Ez,
                   o->Ai[0][2] * Ex + o->Ai[1][2] * Ey + o->Ai[2][2] * Ez};
    for (int a = 0; a < 3; a++) F[3 * i + a] += q[i] * E[a];
    if (phi) phi[i] += ph;
  }
}

/* Element f (flat) of the reference's padded level-1 array at L = 1: dims
 * M + nu - 1, the periodic grid at offset nu/2 - 1, zero ghosts
 * (forcefield1.c:143, 338-346; anterpolate reduces every index modulo M, so
 * the fold at :370-397 adds only zeros). */
static double padded_l1_flat(const orc_t *o, size_t f) {
  const int *g = o->gd[1], nu = o->nu;
  int P1 = g[1] + nu - 1, P2 = g[2] + nu - 1, of = nu / 2 - 1;
  int a = (int)(f / ((size_t)P1 * P2)) - of, b = (int)((f / P2) % P1) - of, c = (int)(f % P2) - of;
  if (a < 0 || a >= g[0] || b < 0 || b >= g[1] || c < 0 || c >= g[2]) return 0.;
  return o->qg[1][((size_t)a * g[1] + b) * g[2] + c];
}

/* FF_energy (forcefield1.c:101-235). */
double orc_energy(orc_t *o, const double *pos, const double *q, double *F, double *phi,
                  double *terms) {
  int N = o->N, nu = o->nu, L = o->L;
  double qsum = 0, q2sum = 0.;
  for (int i = 0; i < N; i++) qsum += q[i], q2sum += q[i] * q[i];
  double g0 = tau_poly(o, -1.); /* gamma(0) */
  double e_const = 0.5 * q2sum * o->coeff1 - 0.5 * (qsum * qsum) * o->coeff2;
  e_const += -0.5 * q2sum * g0 / o->aCut[0];
  if (!o->pstart) orc_build_pairs(o, pos, 0.0);
  double e_sr = short_range(o, pos, q, F, phi);
  if (phi)
    for (int i = 0; i < N; i++)
      phi[i] += q[i] * o->coeff1 - qsum * o->coeff2 - q[i] * g0 / o->aCut[0];
  anterpolate(o, pos, q, o->qg[1]);
  for (int l = 2; l <= L; l++) orc_restrict(o, l, o->qg[l - 1], o->qg[l]);
  /* top level */
  const int *gt = o->gd[L];
  size_t nt = (size_t)gt[0] * gt[1] * gt[2];
  if (o->compat && L >= 2 && o->fft) {
    /* HEAD quirk, FFT build (forcefield1.c:504-515 vs :406-415): the
     * unpadded top grid sits at the start of a zeroed (M+nu-1)^3 buffer
     * that is read with the padded stride and an offset of nu/2-1.  (The
     * DFT build reads that buffer at flat indices, i.e. correctly.) */
    int P[3] = {gt[0] + nu - 1, gt[1] + nu - 1, gt[2] + nu - 1}, of = nu / 2 - 1;
    double *x = (double *)malloc(sizeof(double) * nt);
    for (int i = 0; i < gt[0]; i++)
      for (int j = 0; j < gt[1]; j++)
        for (int k = 0; k < gt[2]; k++) {
          size_t f = ((size_t)(i + of) * P[1] + (j + of)) * P[2] + (k + of);
          x[((size_t)i * gt[1] + j) * gt[2] + k] = f < nt ? o->qg[L][f] : 0.;
        }
    orc_top_level(o, x, o->eg[L]);
    free(x);
  } else if (o->compat && L == 1 && !o->fft) {
    /* HEAD quirk, DFT build (forcefield1.c:483-486 on the array of
     * :338-346): DFTg2g reads the padded level-1 array at the unpadded flat
     * indices */
    double *x = (double *)malloc(sizeof(double) * nt);
    for (size_t f = 0; f < nt; f++) x[f] = padded_l1_flat(o, f);
    orc_top_level(o, x, o->eg[L]);
    free(x);
  } else {
    orc_top_level(o, o->qg[L], o->eg[L]);
  }
  for (int l = L - 1; l >= 1; l--) {
    size_t n = (size_t)o->gd[l][0] * o->gd[l][1] * o->gd[l][2];
    memset(o->eg[l], 0, sizeof(double) * n);
    orc_prolong(o, l + 1, o->eg[l + 1], o->eg[l]);
    orc_grid2grid(o, l, o->qg[l], o->eg[l]);
  }
  const int *g1 = o->gd[1];
  size_t n1 = (size_t)g1[0] * g1[1] * g1[2];
  double gsum = 0.;
  if (o->compat && L == 1) {
    /* HEAD quirk (forcefield1.c:196-212): the padded level-1 array is read
     * at the unpadded flat indices */
    for (size_t f = 0; f < n1; f++) gsum += padded_l1_flat(o, f) * o->eg[1][f];
  } else {
    for (size_t m = 0; m < n1; m++) gsum += o->qg[1][m] * o->eg[1][m];
  }
  double e_grid = 0.5 * gsum;
  interpolate(o, pos, q, o->eg[1], F, phi);
  if (terms) { terms[0] = e_const; terms[1] = e_sr; terms[2] = e_grid; }
  return e_const + e_sr + e_grid;
}

void orc_info(const orc_t *o, int *info, double *dinfo) {
  if (info) {
    info[0] = o->nu; info[1] = o->L;
    info[2] = o->M[0]; info[3] = o->M[1]; info[4] = o->M[2];
    info[5] = o->nLim; info[6] = o->pair_count; info[7] = o->compat;
  }
  if (dinfo) {
    dinfo[0] = o->aCut[0]; dinfo[1] = o->relCutoff; dinfo[2] = o->coeff1;
    dinfo[3] = o->coeff2; dinfo[4] = o->detA;
  }
}

void orc_grid_dims(const orc_t *o, int l, int *d) { for (int a = 0; a < 3; a++) d[a] = o->gd[l][a]; }
void orc_stencil_dims(const orc_t *o, int l, int *d) { for (int a = 0; a < 3; a++) d[a] = o->sd[l][a]; }

void orc_copy_grid(const orc_t *o, int l, int which, double *out) {
  size_t n = (size_t)o->gd[l][0] * o->gd[l][1] * o->gd[l][2];
  memcpy(out, which ? o->eg[l] : o->qg[l], sizeof(double) * n);
}

void orc_copy_stencil(const orc_t *o, int l, double *out) {
  size_t n = (size_t)o->sd[l][0] * o->sd[l][1] * o->sd[l][2];
  memcpy(out, o->khat[l], sizeof(double) * n);
}

int orc_copy_table(const orc_t *o, int which, int l, int a, double *out) {
  int nu = o->nu, n = 0;
  switch (which) {
    case 0: n = nu; memcpy(out, o->tau, sizeof(double) * n); break;
    case 1: n = nu / 2 * nu; memcpy(out, o->Q, sizeof(double) * n); break;
    case 2: n = nu + 1; memcpy(out, o->J, sizeof(double) * n); break;
    case 3: n = o->opD[l][a]; memcpy(out, o->omp[l][a], sizeof(double) * n); break;
    case 4: n = o->M[a]; memcpy(out, o->cL[a], sizeof(double) * n); break;
  }
  return n;
}
