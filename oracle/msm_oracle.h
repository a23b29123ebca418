/*
 * msm_oracle.h -- CPU restatement of msmpb's B-spline MSM (TEST INFRASTRUCTURE).
 *
 * This is the checker for the HIP implementation in msmpb_amd/.  Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  It is never linked into, or called from, the product library.
 *
 * The functions restate code/forcefield0.c and code/forcefield1.c of
 * hkayabilisim/msmpb in periodic (unpadded) form, FP64 throughout.
 * compat = 0 is the corrected algorithm (the one the published energies in
 * results/ JSON files reproduce); compat = 1 reproduces the two index-layout
 * quirks of the reference HEAD (forcefield1.c:196-212 for L = 1 and
 * forcefield1.c:504-515 for L >= 2).
 */
#ifndef MSM_ORACLE_H
#define MSM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* Build: parameters follow FF_build (forcefield0.c:271-367).  Pass -1 (or
 * NULL for Mtop) to let the reference's auto-selection rules choose.
 * fft = 0 selects the reference's DFT build (FF_set_FFT(false)): its own
 * grid and level rules and, with compat, its top-level read.
 * Returns NULL and fills err (if given, >= 256 bytes) on invalid input. */
orc_t *orc_build(int N, const double *q, const double *pos /* N x 3 */,
                 const double *edges /* 3 x 3 row-major, A[i][j] */,
                 int nu, int L, const int *Mtop, double a0, double estErr,
                 int compat, int fft, char *err);

/* (Re)build the pair list like neighborlist() with the given margin
 * (forcefield0.c:21-164); orc_build does it with margin 0.  Returns the
 * number of pairs within a0 + margin. */
int orc_build_pairs(orc_t *o, const double *pos, double margin);

/* One evaluation (forcefield1.c:101-235).  F (N x 3) is overwritten with
 * forces, phi (N, optional) with dE/dq, terms (optional, 3) with
 * {constant part, short-range pair part, grid part}.  Returns E. */
double orc_energy(orc_t *o, const double *pos, const double *q, double *F,
                  double *phi, double *terms);

/* Build-time facts: info[0..] = nu, L, Mtop x,y,z, nLim, pair_count(after
 * energy), and dinfo = a0, relCutoff, coeff1, coeff2, detA. */
void orc_info(const orc_t *o, int *info /* 8 */, double *dinfo /* 5 */);

/* Level-l grid dims (l = 1..L) and stencil dims (l < L: s_l; l = L: M_L). */
void orc_grid_dims(const orc_t *o, int l, int *dims3);
void orc_stencil_dims(const orc_t *o, int l, int *dims3);
/* Copies of the last evaluation's periodic grids: which = 0 -> q^l,
 * 1 -> e^l; and the level stencils khat[l]. */
void orc_copy_grid(const orc_t *o, int l, int which, double *out);
void orc_copy_stencil(const orc_t *o, int l, double *out);
/* Tables: which = 0 tau (nu), 1 Q (nu/2 * nu), 2 J (nu+1), 3 omegap level l
 * axis a (opD; returns length), 4 cL axis a (M_L). */
int orc_copy_table(const orc_t *o, int which, int l, int a, double *out);

/* Stand-alone pieces used by the tests' closed-form checks. */
double orc_bspline(const orc_t *o, int i, double t, double *deriv);
/* Top-level periodic convolution with khat[L] on an arbitrary grid
 * (unpadded M_L^3 in, out). */
void orc_top_level(const orc_t *o, const double *qL, double *eL);
/* Restriction and prolongation operators between level l-1 and l. */
void orc_restrict(const orc_t *o, int l, const double *fine, double *coarse);
void orc_prolong(const orc_t *o, int l, const double *coarse, double *fine);
/* Level-l grid-to-grid correlation (l < L), e += K_l * q. */
void orc_grid2grid(const orc_t *o, int l, const double *ql, double *el);
/* Error estimates (forcefield0.c:521-587). */
double orc_errest(const orc_t *o, int nu);
double orc_deltaF(const orc_t *o);

void orc_free(orc_t *o);

#ifdef __cplusplus
}
#endif
#endif
