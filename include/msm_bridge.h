/*
 * msm_bridge.h -- handle-based entry points of libmsmpb_hip.so for
 * libraries that define their own FF type and FF_* functions.
 *
 * No counterpart in hkayabilisim/msmpb.  The reference builds its PME
 * comparator as a separate program whose header (code/pme.h) reuses the
 * names FF, FF_new, FF_build, FF_energy ... with other signatures and
 * another struct layout.  libpme_hip.so implements that header on top of
 * the kernels of libmsmpb_hip.so through these functions; the handle is
 * the library's own forcefield.h FF, passed as void * so that a translation
 * unit including pme.h can use it.  Every int-returning call returns 0 on
 * success and a negative code on failure (message from FF_last_error()).
 */
#ifndef MSM_BRIDGE_H
#define MSM_BRIDGE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Summation methods. */
#define MSM_METHOD_MSM 0
#define MSM_METHOD_PME 1

/* A configured, unbuilt FF.  orderAcc, maxLevel, cutoff, tolDir <= 0 and
 * topGridDim NULL (or topGridDim[0] <= 0) select the method's defaults.
 * q is kept by pointer and re-read at every evaluation. */
void *msmb_create(int N, double *q, int method, int orderAcc, int maxLevel,
                  const int *topGridDim, double cutoff, double tolDir);
int msmb_build(void *h, double (*position)[3], double edges[3][3]);
int msmb_rebuild(void *h, double edges[3][3], double (*position)[3]);
/* E (NaN on failure); force (N x 3) overwritten; potential (N) optional. */
double msmb_energy(void *h, double (*force)[3], double (*position)[3], double *potential);
/* precision: MSM_PREC_*; timing: per-stage GPU times into msmb_times. */
int msmb_configure(void *h, int precision, int timing);
/* ints = orderAcc, topGridDim x, y, z, pairs i < j inside the cutoff;
 * dbl = cutoff, beta (PME; 0 for MSM), coeff1, coeff2, detA. */
int msmb_info(void *h, int ints[5], double dbl[5]);
/* Host tables owned by the handle: B-spline pieces Q ((nu/2) x nu), cL per
 * axis (topGridDim entries) and the top-level Fourier kernel (M^3). */
int msmb_tables(void *h, const double **Q, const double *cL[3], const double **khat);
/* Seconds of the last build / evaluation: nlist, stencil, partcl2partcl,
 * anterpolation, grid2grid (top level), interpolation. */
int msmb_times(void *h, double t[6]);
void msmb_delete(void *h);

#ifdef __cplusplus
}
#endif
#endif
