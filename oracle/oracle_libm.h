/* oracle_libm.h -- TEST INFRASTRUCTURE.
 *
 * Built with -DORACLE_LIBM (oracle/liboracle_libm.so), the oracle calls the
 * host C library's pow / log10 / exp / log / atan / sin / tan in place of
 * ctsm_math.h's.  Both libraries are correctly rounded to within 1 ulp, so
 * the two builds differ only in last-bit choices of those functions: the
 * libm-sensitivity test runs both on the same columns and bounds how far a
 * reference built on another libm can drift.  Include after ctsm_math.h. */
#ifndef ORACLE_LIBM_H
#define ORACLE_LIBM_H
#ifdef ORACLE_LIBM
#include <math.h>
#define cm_pow pow
#define cm_log10 log10
#define cm_exp exp
#define cm_log log
#define cm_atan atan
#define cm_sin sin
#define cm_tan tan
#endif
#endif
