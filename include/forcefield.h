/*
 * forcefield.h -- drop-in C API of the MI355X MSM library (libmsmpb_hip.so).
 *
 * Same type names, struct layout and prototypes as hkayabilisim/msmpb
 * code/forcefield.h:1-112, so C callers compiled against that header link
 * against this library unchanged.  The FFTW members of FF are kept as
 * pointer-sized placeholders (the reference declares fftw_complex* and two
 * fftw_plan, all pointers), so field offsets agree without fftw3.h.
 *
 * Each entry point below replaces the reference function cited next to it;
 * the computation runs on one AMD Instinct MI355X (gfx950) through HIP.
 * Units follow the reference: charge in e, length in Angstrom, energy in
 * e^2/Angstrom.
 */
#ifndef MSMPB_HIP_FORCEFIELD_H
#define MSMPB_HIP_FORCEFIELD_H

#include <stdbool.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Vector { double x, y, z; } Vector;
typedef struct Matrix { double xx, xy, xz, yx, yy, yz, zx, zy, zz; } Matrix;
typedef struct Triple { int x, y, z; } Triple;

/* Optional-output hook of the reference (forcefield.h:9-19).  Callers such
 * as the reference CLI define `struct O o`; this library never reads it and
 * provides a weak definition for callers that do not. */
struct O {
  char *test;
  double *khatLp1;
  double **kappa;
  double e_const;
  double *e;
  bool recX;
  clock_t time;
  int level;
};
extern struct O o;

/* Public state (forcefield.h:22-69).  Field order and types match the
 * reference; callers read N, detA, ntc, tmax, nlist_len,
 * nlist_interaction_count and the time_* fields directly. */
typedef struct FF {
  int N;
  double *q;                 /* caller-owned charges, re-read every FF_energy */
  double A[3][3], Ai[3][3];  /* box matrix (columns = edge vectors), inverse */
  double detA;
  double relCutoff;
  double errEst;
  double errEstMoore;
  int ntc;
  int tmax;
  double estErr;
  int orderAcc;
  int maxLevel;
  int topGridDim[3];
  int *nlist;                /* unused: pairs are binned on the GPU per call */
  int nlist_len;
  int nlist_interaction_count;
  bool FFT;
  void *fftw_in;             /* reference: fftw_complex * (unused here) */
  void *forward, *backward;  /* reference: fftw_plan (unused here) */
  double *tau;
  double *dsigma;
  double **sigmad;
  double *Q;
  double *J;
  double *(*omegap)[3];
  int nLim;
  double *aCut;
  double cutoff;
  double *cL[3];
  double **khat;
  double coeff1, coeff2;
  double time_partcl2partcl;
  double time_nlist;
  double time_grid2grid[10];
  double time_stencil[10];
  double time_anterpolation;
  double time_interpolation;
  double time_restriction[10];
  double time_prolongation[10];
  double time_padding[10];
} FF;

FF *FF_new(int N, double *q);                                  /* forcefield0.c:166 */
double FF_get_cutoff(FF *ff);                                   /* forcefield0.c:206 */
void FF_set_cutoff(FF *ff, double cutoff);                      /* forcefield0.c:203 */
void FF_set_orderAcc(FF *ff, int orderAcc);                     /* forcefield0.c:187 */
void FF_set_maxLevel(FF *ff, int maxLevel);                     /* forcefield0.c:190 */
void FF_set_topGridDim(FF *ff, int topGridDim[3]);              /* forcefield0.c:194 */
void FF_set_FFT(FF *ff, bool FFT);                              /* forcefield0.c:209 */
void FF_build(FF *ff, int N, double (*position)[3], double edges[3][3],
              double margin);                                   /* forcefield0.c:271 */
int FF_get_orderAcc(FF *ff);                                    /* forcefield0.c:512 */
int FF_get_maxLevel(FF *ff);                                    /* forcefield0.c:514 */
void FF_get_topGridDim(FF *ff, int topGridDim[3]);              /* forcefield0.c:516 */
bool FF_get_FFT(FF *ff);                                        /* forcefield0.c:518 */
double FF_get_errEst(FF *ff);                                   /* forcefield0.c:583 */
double FF_get_estErr(FF *ff);                                   /* forcefield0.c:200 */
void FF_set_estErr(FF *ff, double estErr);                      /* forcefield0.c:197 */
void FF_rebuild(FF *ff, double edges[3][3], double (*position)[3],
                double margin);                                 /* forcefield0.c:598 */
/* Returns the energy and overwrites force (N x 3).  weight must be NULL or
 * an all-ones array of length maxLevel + 1 (forcefield1.c:101). */
double FF_energy(FF *ff, double (*force)[3], double (*position)[3], double *weight);
void FF_delete(FF *ff);                                         /* forcefield0.c:682 */
double FF_get_deltaF(FF *ff);                                   /* forcefield0.c:521 */

double msm4g_tictocmanager(int push);                           /* forcefield1.c:35 */
void msm4g_tic(void);                                           /* forcefield1.c:50 */
double msm4g_toc(void);                                         /* forcefield1.c:54 */

double invert(Matrix *A);                                       /* forcefield0.c:709 */
void omegap(FF *ff);                                            /* forcefield0.c:726 */

#ifndef max
#define max(a, b) ({ __typeof__(a) _a = (a); __typeof__(b) _b = (b); _a > _b ? _a : _b; })
#endif
#ifndef min
#define min(a, b) ({ __typeof__(a) _a = (a); __typeof__(b) _b = (b); _a > _b ? _b : _a; })
#endif

Vector prod(Matrix m, Vector v);                                /* forcefield1.c:639 */
Vector prodT(Matrix m, Vector v);                               /* forcefield1.c:645 */

double _FF_get_errEst(FF *ff, int nu);                          /* forcefield0.c:553 */

/* grid2grid (forcefield0.c:898-931): e^l += K_l * q^l on the reference's
 * padded level-l layout (host arrays; runs on the GPU). */
void grid2grid(FF *ff, const int l, const Triple gd, double *el, const double *ql,
               const Triple sd, const double *kh);

#ifdef __cplusplus
}
#endif
#endif
