/*
 * pme_oracle.h -- CPU restatement of msmpb's smooth particle-mesh Ewald
 * comparator (TEST INFRASTRUCTURE).
 *
 * The checker for the PME method of the HIP library in msmpb_amd/.  Only
 * the test-suite, the smoke check and the CPU-baseline leg of the benchmark
 * may load it; the product library never links it.
 *
 * The functions restate code/pme0.c and code/pme1.c of hkayabilisim/msmpb,
 * FP64 throughout, single-threaded.
 */
#ifndef PME_ORACLE_H
#define PME_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmo pmo_t;

/* Build (pme0.c:192-322 with FF_rebuild :365-412).  nu <= 0, a0 <= 0,
 * tolDir <= 0 and M == NULL select the reference's defaults (6, 8 A, 1e-5
 * and the grid-size table).  Returns NULL and fills err (>= 256 bytes) on
 * input outside the domain. */
pmo_t *pmo_build(int N, const double *q, const double *pos /* N x 3 */,
                 const double *edges /* 3 x 3 row-major, A[i][j] */, int nu,
                 const int *M, double a0, double tolDir, char *err);

/* One evaluation (pme1.c:52-109): F (N x 3) overwritten with forces, phi
 * (N, optional) with dE/dq, terms (optional, 3) with {constant part,
 * real-space pair part, reciprocal (grid) part}.  Returns E. */
double pmo_energy(pmo_t *o, const double *pos, const double *q, double *F,
                  double *phi, double *terms);

/* info = nu, M x, y, z, pairs i < j with nearest image inside a0 (counted
 * at build); dinfo = a0, beta, coeff1, coeff2, detA, tolDir. */
void pmo_info(const pmo_t *o, int *info /* 5 */, double *dinfo /* 6 */);

/* The Fourier-space kernel khat (M^3, x-major) and the last evaluation's
 * charge grid (which = 0) or potential grid (which = 1). */
void pmo_copy_khat(const pmo_t *o, double *out);
void pmo_copy_grid(const pmo_t *o, int which, double *out);

/* B-spline weights M_nu(i + t), i = 0 .. nu-1, and their derivatives. */
void pmo_bspline(int nu, double t, double *w, double *dw);

void pmo_free(pmo_t *o);

#ifdef __cplusplus
}
#endif
#endif
