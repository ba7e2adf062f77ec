// common.h -- shared definitions of the gfx950 kernels.
//
// Physical constants follow shr_const_mod (tools/mksurfdata_esmf/src/
// shr_const_mod.F90:14-52) and clm_varcon.F90:50-121.  Arithmetic is IEEE
// binary64 with FMA contraction disabled (-ffp-contract=off), so every
// expression rounds exactly as the reference Fortran writes it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ctsm_gpu.h"
#include "ctsm_math.h"

namespace ctsm {

constexpr double TFRZ = 273.15;
constexpr double DENH2O = 1.000e3;
constexpr double DENICE = 0.917e3;
constexpr double CPLIQ = 4.188e3;
constexpr double CPICE = 2.11727e3;
constexpr double HFUS = 3.337e5;
constexpr double GRAV = 9.80616;
constexpr double SB = 5.67e-8;
constexpr double TKAIR = 0.023;
constexpr double TKICE = 2.290;
constexpr double TKWAT = 0.57;
constexpr double THK_BEDROCK = 3.0;
constexpr double CSOL_BEDROCK = 2.0e6;
constexpr double CAPR = 0.34;
constexpr double CNFAC = 0.5;
constexpr double THIN_SFCLAYER = 1.0e-6;
constexpr double M_TO_MM = 1.e3;

// Compile-time level bounds of the kernels (the reference's 20SL_8.5m
// structure with nlevsno = 12).  ctsm_create rejects larger configurations.
constexpr int MAX_NLEVSOI = 20;
constexpr int MAX_NLEVSNO = 12;
constexpr int MAX_NLEVGRND = 25;
constexpr int MAX_ROWS = MAX_NLEVSNO + 1 + MAX_NLEVGRND;   // banded rows: snow + h2osfc + ground

// Fortran MAX/MIN of two reals: the first argument unless the second is
// strictly larger/smaller.
__host__ __device__ inline double fmax90(double a, double b) { return (b > a) ? b : a; }
__host__ __device__ inline double fmin90(double a, double b) { return (b < a) ? b : a; }

// Per-launch view of the reference arrays: base pointers of the Fortran
// allocations plus the column / patch strides.  Element (c, j) of an array
// whose lower level bound is `lo` is base[(c - begc) + (j - lo) * ldc].
struct View {
    int begc, ldc, begp, ldp, nlevsno, nlevsoi, nlevgrnd;
    __device__ __forceinline__ long c1(int c) const { return (long)(c - begc); }
    __device__ __forceinline__ long sno(int c, int j) const { return (long)(c - begc) + (long)(j + nlevsno - 1) * ldc; }
    __device__ __forceinline__ long zi(int c, int j) const { return (long)(c - begc) + (long)(j + nlevsno) * ldc; }
    __device__ __forceinline__ long soi(int c, int j) const { return (long)(c - begc) + (long)(j - 1) * ldc; }
    __device__ __forceinline__ long p1(int p) const { return (long)(p - begp); }
    __device__ __forceinline__ long lyr(int p, int j) const { return (long)(p - begp) + (long)(j + nlevsno - 1) * ldp; }
};

// Namelist scalars the kernels read.
struct Params {
    double dtmin, verySmall, xTolerUpper, xTolerLower, e_ice;
    int lower_bc, flux_calculation, snow_thermal_cond;
};

// Error word of a launch: first failing column index (atomicMin) and a code.
struct ErrWord {
    int32_t first_col;   // INT32_MAX when no failure
    int32_t code;
};

// Upper bound on Richards trials per step: the reference loop accepts a trial
// at least every log2(dtime/dtmin)+1 trials, so a column needs at most a few
// hundred; the cap only turns non-finite input into an error instead of a
// wave that never retires.
constexpr int MAX_TRIALS = 20000;

__device__ __forceinline__ bool col_in_range(const View &V, int c) { return c >= V.begc && c < V.begc + V.ldc; }

__device__ inline void report_error(ErrWord *e, int c, int code)
{
    if (e) {
        atomicMin(&e->first_col, c);
        atomicMax(&e->code, code);
    }
}

// Orders the LDS accesses of one wavefront's lanes around it (the lanes
// exchange values through LDS; a wavefront's LDS operations complete in
// issue order).  No workgroup barrier: the exchanges here stay inside a wave.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace ctsm
