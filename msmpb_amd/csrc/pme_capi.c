/*
 * pme_capi.c -- libpme_hip.so: the C API of the reference's PME comparator
 * (hkayabilisim/msmpb code/pme.h, implemented there by code/pme0.c and
 * code/pme1.c) on the GPU kernels of libmsmpb_hip.so.
 *
 * The reference builds PME as a program of its own whose header reuses the
 * names FF, FF_new, FF_build, FF_energy ... with other signatures, so it is
 * a library of its own here too.  Each FF of pme.h holds a handle of the
 * MSM library configured for the PME method (msm_bridge.h): Ewald pairs,
 * B-spline spreading and interpolation, Fourier convolution on one grid.
 * This file only moves parameters and results between the two structs; the
 * arithmetic runs in libmsmpb_hip.so.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "msm_bridge.h"
#include "pme.h"
#include "pme_hip.h"

const char *FF_last_error(void); /* libmsmpb_hip.so (msmpb_hip.h) */

typedef struct PX {
  FF ff; /* first member: FF * and PX * alias */
  void *h;
  int precision;
} PX;

static PX *px(FF *ff) { return (PX *)ff; }

static void report(const char *what) { fprintf(stderr, "pme_hip: %s: %s\n", what, FF_last_error()); }

/* Adjugate inverse, returns |det| (pme0.c:427-442). */
static double inverse(const double a[3][3], double ai[3][3]) {
  double c[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      c[i][j] = a[(j + 1) % 3][(i + 1) % 3] * a[(j + 2) % 3][(i + 2) % 3] -
                a[(j + 2) % 3][(i + 1) % 3] * a[(j + 1) % 3][(i + 2) % 3];
  double det = a[0][0] * c[0][0] + a[0][1] * c[1][0] + a[0][2] * c[2][0];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) ai[i][j] = c[i][j] / det;
  return fabs(det);
}

/* Copies the built handle's parameters, tables and build times into the
 * public struct. */
static void publish(PX *x) {
  FF *ff = &x->ff;
  int iv[5];
  double dv[5], t[6];
  if (msmb_info(x->h, iv, dv) != 0) return;
  ff->orderAcc = iv[0];
  for (int a = 0; a < 3; a++) ff->topGridDim[a] = iv[1 + a];
  ff->nlist_interaction_count = iv[4];
  ff->cutoff = dv[0];
  ff->beta = dv[1];
  ff->coeff1 = dv[2];
  ff->coeff2 = dv[3];
  ff->detA = inverse(ff->A, ff->Ai);
  if (!(ff->tolDir > 0)) ff->tolDir = 1e-5;
  /* the reference's pair-list capacity (pme0.c:31), reported by its CLI */
  double a0 = ff->cutoff, N = ff->N;
  ff->nlist_len = (int)(3 * N + 2.1 * ((4.0 / 3.0) * M_PI * a0 * a0 * a0 * 0.5 * N * (N - 1.0) / ff->detA));
  const double *Q, *cL[3], *kh;
  if (msmb_tables(x->h, &Q, cL, &kh) == 0) {
    ff->Q = (double *)Q;
    for (int a = 0; a < 3; a++) ff->cL[a] = (double *)cL[a];
    ff->khat = (double *)kh;
  }
  if (msmb_times(x->h, t) == 0) {
    ff->time_nlist = t[0];
    ff->time_stencil = t[1];
  }
}

FF *FF_new(int N, double *q, double edges[3][3]) {
  PX *x = (PX *)calloc(1, sizeof(PX));
  x->ff.N = N;
  x->ff.q = q;
  memcpy(x->ff.A, edges, sizeof x->ff.A);
  return &x->ff;
}

void FF_set_cutoff(FF *ff, double cutoff) { ff->cutoff = cutoff; }
double FF_get_cutoff(FF *ff) { return ff->cutoff; }
void FF_set_orderAcc(FF *ff, int orderAcc) { ff->orderAcc = orderAcc; }
void FF_set_topGridDim(FF *ff, int topGridDim[3]) {
  for (int a = 0; a < 3; a++) ff->topGridDim[a] = topGridDim[a];
}
void FF_set_tolDir(FF *ff, double tolDir) { ff->tolDir = tolDir; }
int FF_get_orderAcc(FF *ff) { return ff->orderAcc; }
void FF_get_topGridDim(FF *ff, int topGridDim[3]) {
  for (int a = 0; a < 3; a++) topGridDim[a] = ff->topGridDim[a];
}
double FF_get_tolDir(FF *ff) { return ff->tolDir; }

int pme_set_precision(FF *ff, int precision) {
  if (!ff || (precision != 0 && precision != 1)) return -1;
  px(ff)->precision = precision;
  return px(ff)->h ? msmb_configure(px(ff)->h, precision, 1) : 0;
}

void FF_build(FF *ff, double (*position)[3]) {
  PX *x = px(ff);
  if (x->h) msmb_delete(x->h);
  x->h = msmb_create(ff->N, ff->q, MSM_METHOD_PME, ff->orderAcc, 1,
                     ff->topGridDim[0] > 0 ? ff->topGridDim : NULL, ff->cutoff, ff->tolDir);
  if (!x->h || msmb_configure(x->h, x->precision, 1) != 0 ||
      msmb_build(x->h, position, ff->A) != 0) {
    report("FF_build");
    if (x->h) msmb_delete(x->h);
    x->h = NULL;
    return;
  }
  publish(x);
}

void FF_rebuild(FF *ff, double edges[3][3], double (*position)[3]) {
  PX *x = px(ff);
  memcpy(ff->A, edges, sizeof ff->A);
  if (!x->h || msmb_rebuild(x->h, ff->A, position) != 0) {
    report("FF_rebuild");
    return;
  }
  publish(x);
}

double FF_energy(FF *ff, double (*force)[3], double (*position)[3], double *weight) {
  PX *x = px(ff);
  if (weight && (weight[0] != 1.0 || weight[1] != 1.0)) {
    fprintf(stderr, "pme_hip: FF_energy: weights other than 1 are not supported\n");
    return NAN;
  }
  if (!x->h) {
    fprintf(stderr, "pme_hip: FF_energy: FF_build has not completed successfully\n");
    return NAN;
  }
  double E = msmb_energy(x->h, force, position, NULL);
  if (isnan(E)) {
    report("FF_energy");
    return E;
  }
  double t[6];
  if (msmb_times(x->h, t) == 0) {
    ff->time_partcl2partcl = t[2];
    ff->time_anterpolation = t[3];
    ff->time_grid2grid = t[4];
    ff->time_interpolation = t[5];
  }
  return E;
}

/* A-priori force error of PME (pme0.c:333-359): C_nu Q2 h^(nu-2) /
 * (V^(1/3) sqrt(N) a0^(nu-1)), h the finest spacing 1/(|row a of Ai| M_a). */
double FF_get_errEst(FF *ff, int N, double *charge) {
  int nu = ff->orderAcc;
  if (nu > 10 || nu < 4) return NAN;
  static const double Mc[4] = {9., 825., 130095., 34096545.};
  static const double cbar[4] = {1. / 6., 1. / 30., 1. / 140., 1. / 630.};
  static const double kc[4] = {0.39189561, 0.150829428, 0.049632967, 0.013520855};
  int k = (nu - 4) / 2;
  double C = 4. / 3. * Mc[k] * cbar[k] * kc[k];
  double ai[3][3], det = inverse(ff->A, ai), hmin = INFINITY;
  for (int a = 0; a < 3; a++) {
    double as = sqrt(ai[a][0] * ai[a][0] + ai[a][1] * ai[a][1] + ai[a][2] * ai[a][2]);
    hmin = fmin(hmin, 1.0 / (as * ff->topGridDim[a]));
  }
  double Q2 = 0;
  for (int i = 0; i < N; i++) Q2 += charge[i] * charge[i];
  return Q2 * C * pow(hmin, nu - 2) / (pow(det, 1. / 3.) * sqrt((double)N) * pow(ff->cutoff, nu - 1));
}

void FF_delete(FF *ff) {
  if (!ff) return;
  if (px(ff)->h) msmb_delete(px(ff)->h);
  free(px(ff));
}

/* CPU-clock tic/toc stack (pme1.c:29-50). */
double msm4g_tictocmanager(int push) {
  static clock_t stack[100];
  static int top = 0;
  if (push) {
    if (top < 100) stack[top++] = clock();
    return 0.0;
  }
  if (top == 0) return 0.0;
  clock_t now = clock();
  return (double)(now - stack[--top]) / CLOCKS_PER_SEC;
}
void msm4g_tic() { msm4g_tictocmanager(1); }
double msm4g_toc() { return msm4g_tictocmanager(0); }

Vector prod(Matrix m, Vector v) {
  Vector r = {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.yx * v.x + m.yy * v.y + m.yz * v.z,
              m.zx * v.x + m.zy * v.y + m.zz * v.z};
  return r;
}

Vector prodT(Matrix m, Vector v) {
  Vector r = {m.xx * v.x + m.yx * v.y + m.zx * v.z, m.xy * v.x + m.yy * v.y + m.zy * v.z,
              m.xz * v.x + m.yz * v.y + m.zz * v.z};
  return r;
}
