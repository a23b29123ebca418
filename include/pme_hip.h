/*
 * pme_hip.h -- entry points of libpme_hip.so beyond pme.h.
 *
 * No counterpart in hkayabilisim/msmpb (its PME is FP64 on the CPU).
 */
#ifndef MSMPB_HIP_PME_EXT_H
#define MSMPB_HIP_PME_EXT_H

#include "pme.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the evaluation, before or after FF_build: 0 = FP64
 * throughout (parity mode), 1 = FP32 Ewald pairs with FP64 accumulation,
 * FP64 grid.  Returns 0, or -1 for an invalid value. */
int pme_set_precision(FF *ff, int precision);

/* Last error message of the calling thread ("" if none). */
const char *FF_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
