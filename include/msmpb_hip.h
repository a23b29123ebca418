/*
 * msmpb_hip.h -- entry points of libmsmpb_hip.so beyond forcefield.h.
 *
 * These have no counterpart in hkayabilisim/msmpb; they serve callers that
 * keep positions and charges in GPU memory (the PyTorch autograd binding in
 * msmpb_amd/) and callers that choose the precision or the reference HEAD
 * index layout.  All pointers named d_* are device pointers (HBM) of the
 * current HIP device; `stream` is a hipStream_t on which the call is
 * ordered (NULL = the default stream, as everywhere in HIP; PyTorch's
 * default stream is that stream).  Internally the short-range pairs and the
 * level-1 stencil pass run on two library streams forked from and joined
 * back into `stream`.
 * Every int-returning call returns 0 on success and a negative code on
 * failure, with a message available from FF_last_error().
 */
#ifndef MSMPB_HIP_EXT_H
#define MSMPB_HIP_EXT_H

#include "forcefield.h"
#include "msm_bridge.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MSM_OK = 0,
  MSM_ERR_ARG = -1,      /* invalid argument or parameter outside the domain */
  MSM_ERR_STATE = -2,    /* called before FF_build, or after a failed build */
  MSM_ERR_HIP = -3,      /* HIP runtime failure (message names the call) */
  MSM_ERR_ALLOC = -4     /* device allocation failed */
};

/* Arithmetic of the evaluation. */
enum {
  MSM_PREC_FP64 = 0,     /* FP64 throughout (parity mode) */
  MSM_PREC_MIXED = 1     /* FP32 short-range pairs with FP64 accumulation,
                            BF16 hi/lo-split MFMA level stencils with FP32
                            accumulation; FP64 elsewhere */
};

/* Index layout of the grid levels. */
enum {
  MSM_LAYOUT_FIXED = 0,  /* corrected layout (reproduces the published
                            results/ energies of the reference) */
  MSM_LAYOUT_ASIS = 1    /* reference HEAD: L = 1 grid energy read at
                            unpadded indices (forcefield1.c:196-212); L >= 2
                            top grid read padded (forcefield1.c:504-515) */
};

/* Last error message of the calling thread ("" if none). */
const char *FF_last_error(void);

/* Set before FF_build (or between evaluations). */
int FF_set_precision(FF *ff, int precision);
int FF_get_precision(FF *ff);
int FF_set_layout_compat(FF *ff, int layout);
/* Summation method, set before FF_build.  MSM_METHOD_PME (msm_bridge.h)
 * evaluates the reference's smooth particle-mesh Ewald comparator
 * (code/pme0.c, code/pme1.c) with the same kernels: one grid of topGridDim
 * points (used as given; default from the reference's spacing table), the
 * Ewald pair term erfc(beta r)/r inside the cutoff with erfc(beta cutoff)
 * = tolDir * cutoff (tolDir <= 0: 1e-5) and a Gaussian Fourier kernel.
 * Defaults: order 6, cutoff 8 A; maxLevel is ignored (one level). */
int FF_set_method(FF *ff, int method, double tolDir);
/* Ewald splitting parameter beta of a built PME FF (0 for MSM). */
double FF_get_beta(FF *ff);
/* Record per-stage GPU time into the FF time_* fields on FF_energy. */
int FF_set_stage_timing(FF *ff, int enable);

/* Energy split of the last evaluation: {constant part, short-range pairs,
 * grid part}; their sum is the returned energy. */
int FF_get_energy_terms(FF *ff, double terms[3]);

/* Same as FF_energy but also returns phi_i = dE/dq_i (N doubles, host). */
double FF_energy_potential(FF *ff, double (*force)[3], double (*position)[3],
                           double *potential);

/* Device-resident evaluation.  d_pos: N x 3 (row-major xyz), d_q: N.
 * d_energy (device, >= 4 doubles) receives {E, constant, pairs, grid}.
 * d_force (N x 3) and d_phi (N) are optional (NULL = not computed).
 * The call is stream-ordered: it returns before the GPU finishes. */
int msm_eval_device(FF *ff, const double *d_pos, const double *d_q, double *d_energy,
                    double *d_force, double *d_phi, void *stream);

/* Split evaluation for automatic differentiation.
 * msm_forward_device computes the energy and keeps state for
 * msm_adjoint_device: keep = MSM_KEEP_ALL keeps the level-1 potential grid,
 * the atom ordering and the short-range parts of -dE/dr and dE/dq (the pair
 * kernel writes them during the forward at no extra pass);
 * MSM_KEEP_PAIRS keeps the short-range parts only; MSM_KEEP_NONE nothing.
 * msm_adjoint_device writes d_force = -dE/dr and d_phi = dE/dq: the
 * long-range parts by interpolation from the level-1 grid, recomputed from
 * d_pos and d_q when it was not kept (or recompute != 0), plus the kept
 * short-range parts or, without them, a pair pass.  The kept state belongs
 * to the latest forward on this FF and to its d_pos and d_q pointers. */
#define MSM_KEEP_NONE 0
#define MSM_KEEP_ALL 1
#define MSM_KEEP_PAIRS 2
int msm_forward_device(FF *ff, const double *d_pos, const double *d_q, double *d_energy,
                       int keep, void *stream);
int msm_adjoint_device(FF *ff, const double *d_pos, const double *d_q, double *d_force,
                       double *d_phi, int recompute, void *stream);

/* Per-stage GPU time of the device-path evaluations run since the last call
 * while stage timing is on (FF_set_stage_timing): out[k] (k < n) receives
 * the summed seconds of stage k, named by msm_stage_name(k) (9 stages:
 * binning and sort, pairs, spreading, restriction, level stencils, top
 * level, prolongation, energy reduction, interpolation).  Returns the
 * number of evaluations covered (or a negative error code) and resets. */
int msm_stage_times(FF *ff, double *out, int n);
const char *msm_stage_name(int k);

/* Diagnostics for the stage-by-stage parity tests: copy the last
 * evaluation's periodic level-l grid (which = 0: charges q^l, 1:
 * potentials e^l; x-major, z fastest) or the level-l stencil to host
 * memory.  Dims via msm_grid_dims / msm_stencil_dims. */
int msm_grid_dims(FF *ff, int level, int dims[3]);
int msm_stencil_dims(FF *ff, int level, int dims[3]);
int msm_copy_grid(FF *ff, int level, int which, double *host_out);
int msm_copy_stencil(FF *ff, int level, double *host_out);
/* Number of atom pairs i < j within the cutoff (nearest image), counted on
 * the GPU for the given positions (host N x 3). */
int msm_count_pairs(FF *ff, double (*position)[3], long long *count);

/* grid2grid (forcefield.h, the reference's forcefield0.c:898-931) on device
 * pointers, allocating nothing: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  d_el (gd.x * gd.y * gd.z doubles) += the correlation of
 * the stencil d_kh (sd.x * sd.y * sd.z) with d_ql, a level grid in the
 * reference's padded layout (pad max(min(2 nLim + 1, gd), nu) per axis,
 * interior at max(pad / 2, nu / 2); nu and nLim of the built ff).  Its two
 * scratch grids live in the library's static device array: at most 524,288
 * grid points per call (larger grids: MSM_ERR_ARG).  Stream-ordered on
 * `stream`; calls that overlap on one GPU share the array and race. */
int msm_grid2grid_device_noalloc(FF *ff, int l, Triple gd, double *d_el, const double *d_ql,
                                 Triple sd, const double *d_kh, void *stream);

#ifdef __cplusplus
}
#endif
#endif
