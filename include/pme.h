/*
 * pme.h -- drop-in C API of the MI355X PME comparator (libpme_hip.so).
 *
 * Same type names, struct layout and prototypes as hkayabilisim/msmpb
 * code/pme.h:1-68 (the smooth particle-mesh Ewald program the reference
 * compares its MSM against), so a caller compiled against that header,
 * such as the reference's code/pme.c, links against this library
 * unchanged.  The FFTW members are pointer-sized placeholders (the
 * reference declares an fftw_complex pointer and two fftw_plan, all
 * pointers), so field offsets agree without fftw3.h.
 *
 * libpme_hip.so evaluates on the GPU with the kernels of libmsmpb_hip.so:
 * the Ewald real-space pairs, B-spline charge spreading and interpolation
 * and the Fourier-space convolution on one grid.  Units as in the
 * reference: charge in e, length in Angstrom, energy in e^2/Angstrom.
 * Entry points that fail print a message to stderr; FF_energy then returns
 * NaN (the reference has no error reporting).  FF_last_error() of
 * msmpb_hip.h gives the message.
 */
#ifndef MSMPB_HIP_PME_H
#define MSMPB_HIP_PME_H

#include <stdbool.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Vector { double x, y, z; } Vector;
typedef struct Matrix { double xx, xy, xz, yx, yy, yz, zx, zy, zz; } Matrix;
typedef struct Triple { int x, y, z; } Triple;

/* Optional-output hook (pme.h:9-13), defined by callers; never read. */
struct O {
  char *test;
  clock_t time;
};
extern struct O o;

/* Public state (pme.h:15-44), field for field. */
typedef struct FF {
  int N;
  double *q;                 /* caller-owned charges, re-read every FF_energy */
  double A[3][3], Ai[3][3];  /* box matrix (columns = edge vectors), inverse */
  double detA;
  double cutoff;             /* real-space cutoff a0 (default 8) */
  int orderAcc;              /* B-spline order (default 6) */
  int topGridDim[3];         /* grid (default from the reference's table) */
  int *nlist;                /* unused: pairs are binned on the GPU per call */
  int nlist_len;             /* the reference's list capacity formula */
  int nlist_interaction_count; /* pairs i < j inside the cutoff */
  double tolDir;             /* erfc(beta a0) / a0 (default 1e-5) */
  void *fftw_in;             /* reference: fftw_complex * (unused here) */
  void *forward, *backward;  /* reference: fftw_plan (unused here) */
  double beta;               /* Ewald splitting parameter */
  double *Q;                 /* B-spline pieces, (nu/2) x nu */
  double *cL[3];             /* 1 / Phi^(2 pi k / M)^2 per axis */
  double *khat;              /* Fourier kernel, topGridDim product */
  double coeff1, coeff2;     /* constant-energy coefficients */
  double time_partcl2partcl;
  double time_nlist;
  double time_grid2grid;
  double time_stencil;
  double time_anterpolation;
  double time_interpolation;
} FF;

FF *FF_new(int N, double *q, double edges[3][3]);       /* pme0.c:164-170 */
void FF_set_cutoff(FF *ff, double cutoff);              /* pme0.c:177-178 */
double FF_get_cutoff(FF *ff);                           /* pme0.c:324-325 */
void FF_set_orderAcc(FF *ff, int orderAcc);             /* pme0.c:181-183 */
void FF_set_topGridDim(FF *ff, int topGridDim[3]);      /* pme0.c:184-185 */
void FF_set_tolDir(FF *ff, double tolDir);              /* pme0.c:186-187 */
/* Resolves the parameters, builds the tables and counts the pairs
 * (pme0.c:192-322). */
void FF_build(FF *ff, double (*position)[3]);
int FF_get_orderAcc(FF *ff);                            /* pme0.c:326-327 */
void FF_get_topGridDim(FF *ff, int topGridDim[3]);      /* pme0.c:328-329 */
double FF_get_tolDir(FF *ff);                           /* pme0.c:330-331 */
double FF_get_errEst(FF *ff, int N, double *charge);    /* pme0.c:333-359 */
/* New box: constants and Fourier kernel (pme0.c:365-412); the kernel is
 * generated on the GPU. */
void FF_rebuild(FF *ff, double edges[3][3], double (*position)[3]);
/* Energy; force (N x 3) overwritten (pme1.c:52-109).  weight must be NULL
 * or {1, 1}. */
double FF_energy(FF *ff, double (*force)[3], double (*position)[3], double *weight);
void FF_delete(FF *ff);                                 /* pme0.c:414-423 */

double msm4g_tictocmanager(int push);                   /* pme1.c:29-50 */
void msm4g_tic();
double msm4g_toc();

Vector prod(Matrix m, Vector v);                        /* pme1.c:300-311 */
Vector prodT(Matrix m, Vector v);

#ifdef __cplusplus
}
#endif
#endif
