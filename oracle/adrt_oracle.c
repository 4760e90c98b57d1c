/*
 * CPU oracle for the ADRT transform core -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C (C11 + OpenMP) restatement of the reference's native routines
 * (src/adrt/adrt_cdefs_{adrt,bdrt,iadrt,fmg}.hpp), instantiated for float and
 * double.  Built with -ffp-contract=off so products in the high-pass filter
 * are rounded before they are summed, exactly as the reference computes them.
 * Entry points return 0 on success and 1 on allocation failure; shape
 * validation is the caller's job (the Python side of the oracle).
 */
#include <stddef.h>
#include <stdlib.h>

int oracle_openmp_enabled(void) {
#ifdef _OPENMP
    return 1;
#else
    return 0;
#endif
}

#define REAL float
#define SUF f32
#include "adrt_oracle_body.h"
#undef REAL
#undef SUF

#define REAL double
#define SUF f64
#include "adrt_oracle_body.h"
#undef REAL
#undef SUF
