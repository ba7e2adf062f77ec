// phs_dev.h -- device pieces of plant hydraulic stress shared by phs.hip
// (the soil-side and calcstress kernels) and canopy.hip (CanopyFluxes with
// PhotosynthesisHydraulicStress): the parameter table's rows, the Weibull
// vulnerability curve plc and its derivative, and calcstress with getqflx,
// spacF, spacA and getvegwp (PhotosynthesisMod.F90:4490-5228).
#pragma once

#include "common.h"

namespace ctsm {
namespace phs {

constexpr double RPI = 3.14159265358979323846;   // SHR_CONST_PI
constexpr double C_TO_B = 2.0;                     // g biomass / g C (PhotosynthesisMod.F90:2911)
constexpr double CROOT_LATERAL_LENGTH = 0.25;      // m (:2910)

// rows of the device parameter table, CTSM_MAX_PFT values each
enum { T_KRMAX = 0, T_RADIUS = 1, T_DENSITY = 2, T_FROOT_LEAF = 3, T_PSI50 = 4, T_CK = 8, T_KMAX = 12, T_CROP = 16, T_ROWS = 17 };

// plc, vegetation_weibull branch: 2**(-(x/psi50)**ck), cut to zero below 0.005
__device__ __forceinline__ double plc(double x, double psi50, double ck)
{
    double v = cm_pow(2.0, -cm_pow(x / psi50, ck));
    if (v < 0.005) v = 0.0;
    return v;
}

// ---- calcstress (PhotosynthesisMod.F90:4490-4710) for one patch ---------------
// Segments 0 sun, 1 sha, 2 xyl, 3 root.  The helpers follow the reference's
// routines statement by statement, so every lane performs the Fortran's
// operations in the Fortran's order.
constexpr double RGAS = 6.02214e26 * 1.38065e-23;   // SHR_CONST_AVOGAD * SHR_CONST_BOLTZ
constexpr double LN2 = 0.6931471805599453;         // log(2._r8), correctly rounded
constexpr double SPVAL = 1.e36;
constexpr double TOL_LAI = 0.001;

struct Leaf {
    double laisun, laisha, elai, esai, tsai, htop, fdry, thm, rho, pbot, gb;
    double kmax[4], psi50[4], ck[4];
    const double *k, *smp, *z;   // layer 1 of k_soil_root(p, :), smp_l(c, :), z(c, :)
    long sk, ss;
    int n;
};

__device__ __forceinline__ double d1plc(double x, double psi50, double ck)
{
    return -ck * LN2 * cm_pow(2.0, -cm_pow(x / psi50, ck)) * cm_pow(x / psi50, ck) / x;
}

// getqflx (:5080-5164)
__device__ void getqflx(const Leaf &P, double qsatl, double qaf, bool havegs, double &gs_sun, double &gs_sha,
                        double &q_sun, double &q_sha)
{
    const double cf = P.pbot / (RGAS * 1.e-3 * P.thm) * 1.e6;
    const double wtl = (P.elai + P.esai) * P.gb;
    const double efpot = P.rho * wtl * (qsatl - qaf);
    if (havegs) {
        if (efpot > 0.0 && P.elai > 0.0) {
            if (gs_sun > 0.0) {
                const double rpp = P.fdry / P.gb * (P.laisun / (1.0 / P.gb + 1.0 / gs_sun)) / P.elai;
                q_sun = efpot * rpp / cf;
            } else {
                q_sun = 0.0;
            }
            if (gs_sha > 0.0) {
                const double rpp = P.fdry / P.gb * (P.laisha / (1.0 / P.gb + 1.0 / gs_sha)) / P.elai;
                q_sha = efpot * rpp / cf;
            } else {
                q_sha = 0.0;
            }
        } else {
            q_sun = 0.0;
            q_sha = 0.0;
        }
    } else {
        gs_sun = (q_sun > 0.0) ? P.gb * q_sun * cf * P.elai / (efpot * P.fdry * P.laisun - q_sun * cf * P.elai) : 0.0;
        gs_sha = (q_sha > 0.0) ? P.gb * q_sha * cf * P.elai / (efpot * P.fdry * P.laisha - q_sha * cf * P.elai) : 0.0;
    }
}

// spacF (:4898-4976)
__device__ void spacF(const Leaf &P, const double *x, double *f, double q_sun, double q_sha)
{
    const double grav1 = P.htop * 1000.0;
    const double fsto1 = plc(x[0], P.psi50[0], P.ck[0]), fsto2 = plc(x[1], P.psi50[1], P.ck[1]);
    const double fx = plc(x[2], P.psi50[2], P.ck[2]), fr = plc(x[3], P.psi50[3], P.ck[3]);
    f[0] = q_sun * fsto1 - P.laisun * P.kmax[0] * fx * (x[2] - x[0]);
    f[1] = q_sha * fsto2 - P.laisha * P.kmax[1] * fx * (x[2] - x[1]);
    f[2] = P.laisun * P.kmax[0] * fx * (x[2] - x[0]) + P.laisha * P.kmax[1] * fx * (x[2] - x[1])
           - P.tsai * P.kmax[2] / P.htop * fr * (x[3] - x[2] - grav1);
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < P.n; ++j) s1 = s1 + P.k[j * P.sk] * (x[3] + P.z[j * P.ss] * 1000.0);
    for (int j = 0; j < P.n; ++j) s2 = s2 + P.k[j * P.sk] * P.smp[j * P.ss];
    f[3] = P.tsai * P.kmax[2] / P.htop * fr * (x[3] - x[2] - grav1) + s1 - s2;
    if (P.laisha < TOL_LAI) {
        const double t = f[0];
        f[0] = f[1];
        f[1] = t;
    }
}

// spacA (:4715-4893); invA row-major; returns true when not invertible
__device__ bool spacA(const Leaf &P, const double *x, double *iA, double q_sun, double q_sha)
{
    double A[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) { A[i] = 0.0; iA[i] = 0.0; }
#define AA(i, j) A[((i) - 1) * 4 + (j) - 1]
#define IA(i, j) iA[((i) - 1) * 4 + (j) - 1]
    const double grav1 = P.htop * 1000.0;
    const double fsto1 = plc(x[0], P.psi50[0], P.ck[0]), fsto2 = plc(x[1], P.psi50[1], P.ck[1]);
    const double fx = plc(x[2], P.psi50[2], P.ck[2]), fr = plc(x[3], P.psi50[3], P.ck[3]);
    const double dfsto1 = d1plc(x[0], P.psi50[0], P.ck[0]), dfsto2 = d1plc(x[1], P.psi50[1], P.ck[1]);
    const double dfx = d1plc(x[2], P.psi50[2], P.ck[2]), dfr = d1plc(x[3], P.psi50[3], P.ck[3]);
    const double ks = P.kmax[0], kh = P.kmax[1], kx = P.kmax[2];
    AA(1, 1) = -(P.laisun * ks * fx) - q_sun * dfsto1;
    AA(1, 3) = P.laisun * ks * dfx * (x[2] - x[0]) + P.laisun * ks * fx;
    AA(2, 2) = -(P.laisha * kh * fx) - q_sha * dfsto2;
    AA(2, 3) = P.laisha * kh * dfx * (x[2] - x[1]) + P.laisha * kh * fx;
    AA(3, 1) = P.laisun * ks * fx;
    AA(3, 2) = P.laisha * kh * fx;
    AA(3, 3) = -(P.laisun * ks * dfx * (x[2] - x[0])) - P.laisun * ks * fx - P.laisha * kh * dfx * (x[2] - x[1])
               - P.laisha * kh * fx - P.tsai * kx / P.htop * fr;
    AA(3, 4) = P.tsai * kx / P.htop * dfr * (x[3] - x[2] - grav1) + P.tsai * kx / P.htop * fr;
    AA(4, 3) = P.tsai * kx / P.htop * fr;
    double sk = 0.0;
    for (int j = 0; j < P.n; ++j) sk = sk + P.k[j * P.sk];
    AA(4, 4) = -(P.tsai * kx / P.htop * fr) - P.tsai * kx / P.htop * dfr * (x[3] - x[2] - grav1) - sk;
    if (P.laisun > TOL_LAI && P.laisha > TOL_LAI) {
        const double determ = AA(4, 4) * AA(2, 2) * AA(3, 3) * AA(1, 1) - AA(4, 4) * AA(2, 2) * AA(3, 1) * AA(1, 3)
                              - AA(4, 4) * AA(3, 2) * AA(2, 3) * AA(1, 1) - AA(4, 3) * AA(1, 1) * AA(2, 2) * AA(3, 4);
        if (fabs(determ) <= 1.e-50) return true;
        const double ld = 1.0 / determ;
        IA(1, 1) = ld * AA(4, 4) * AA(2, 2) * AA(3, 3) - ld * AA(4, 4) * AA(3, 2) * AA(2, 3)
                   - ld * AA(4, 3) * AA(2, 2) * AA(3, 4);
        IA(2, 1) = ld * AA(2, 3) * AA(4, 4) * AA(3, 1);
        IA(3, 1) = -ld * AA(4, 4) * AA(2, 2) * AA(3, 1);
        IA(4, 1) = ld * AA(4, 3) * AA(2, 2) * AA(3, 1);
        IA(1, 2) = ld * AA(1, 3) * AA(4, 4) * AA(3, 2);
        IA(2, 2) = ld * AA(4, 4) * AA(3, 3) * AA(1, 1) - ld * AA(4, 4) * AA(3, 1) * AA(1, 3)
                   - ld * AA(4, 3) * AA(1, 1) * AA(3, 4);
        IA(3, 2) = -ld * AA(1, 1) * AA(4, 4) * AA(3, 2);
        IA(4, 2) = ld * AA(4, 3) * AA(1, 1) * AA(3, 2);
        IA(1, 3) = -ld * AA(1, 3) * AA(2, 2) * AA(4, 4);
        IA(2, 3) = -ld * AA(2, 3) * AA(1, 1) * AA(4, 4);
        IA(3, 3) = ld * AA(2, 2) * AA(1, 1) * AA(4, 4);
        IA(4, 3) = -ld * AA(4, 3) * AA(1, 1) * AA(2, 2);
        IA(1, 4) = ld * AA(1, 3) * AA(3, 4) * AA(2, 2);
        IA(2, 4) = ld * AA(2, 3) * AA(3, 4) * AA(1, 1);
        IA(3, 4) = -ld * AA(3, 4) * AA(1, 1) * AA(2, 2);
        IA(4, 4) = ld * AA(2, 2) * AA(3, 3) * AA(1, 1) - ld * AA(2, 2) * AA(3, 1) * AA(1, 3)
                   - ld * AA(3, 2) * AA(2, 3) * AA(1, 1);
    } else {
        if (P.laisha <= TOL_LAI) {
            AA(2, 2) = AA(1, 1);
            AA(3, 2) = AA(3, 1);
            AA(2, 3) = AA(1, 3);
        }
        const double determ = AA(2, 2) * AA(3, 3) * AA(4, 4) - AA(3, 4) * AA(2, 2) * AA(4, 3)
                              - AA(2, 3) * AA(3, 2) * AA(4, 4);
        if (fabs(determ) <= 1.e-50) return true;
        IA(2, 2) = AA(3, 3) * AA(4, 4) - AA(3, 4) * AA(4, 3);
        IA(2, 3) = -AA(2, 3) * AA(4, 4);
        IA(2, 4) = AA(3, 4) * AA(2, 3);
        IA(3, 2) = -AA(3, 2) * AA(4, 4);
        IA(3, 3) = AA(2, 2) * AA(4, 4);
        IA(3, 4) = -AA(3, 4) * AA(2, 2);
        IA(4, 2) = AA(3, 2) * AA(4, 3);
        IA(4, 3) = -AA(2, 2) * AA(4, 3);
        IA(4, 4) = AA(2, 2) * AA(3, 3) - AA(2, 3) * AA(3, 2);
        const double r = 1.0 / determ;
#pragma unroll
        for (int i = 0; i < 16; ++i) iA[i] = r * iA[i];
    }
#undef AA
#undef IA
    return false;
}

// getvegwp (:4979-5077); returns soilflux
__device__ double getvegwp(const Leaf &P, double *x, double gs_sun, double gs_sha, double qsatl, double qaf)
{
    const double grav1 = 1000.0 * P.htop;
    double q_sun, q_sha;
    getqflx(P, qsatl, qaf, true, gs_sun, gs_sha, q_sun, q_sha);
    double sk = 0.0;
    for (int j = 0; j < P.n; ++j) sk = sk + P.k[j * P.sk];
    if (fabs(sk) == 0.0) {
        double s = 0.0;
        for (int j = 0; j < P.n; ++j) s = s + (P.smp[j * P.ss] - 1000.0 * P.z[j * P.ss]);
        x[3] = s / P.n;
    } else {
        double s = 0.0;
        for (int j = 0; j < P.n; ++j) s = s + P.k[j * P.sk] * (P.smp[j * P.ss] - 1000.0 * P.z[j * P.ss]);
        x[3] = (s - q_sun - q_sha) / sk;
    }
    const double fr = plc(x[3], P.psi50[3], P.ck[3]);
    if (P.tsai > 0.0 && fr > 0.0)
        x[2] = x[3] - grav1 - (q_sun + q_sha) / (fr * P.kmax[3] / P.htop * P.tsai);
    else
        x[2] = x[3] - grav1;
    const double fx = plc(x[2], P.psi50[2], P.ck[2]);
    x[1] = (P.laisha > 0.0 && fx > 0.0) ? x[2] - (q_sha / (fx * P.kmax[2] * P.laisha)) : x[2];
    x[0] = (P.laisun > 0.0 && fx > 0.0) ? x[2] - (q_sun / (fx * P.kmax[2] * P.laisun)) : x[2];
    double soilflux = 0.0;
    for (int j = 0; j < P.n; ++j) soilflux = soilflux + P.k[j * P.sk] * (P.smp[j * P.ss] - x[3] - 1000.0 * P.z[j * P.ss]);
    return soilflux;
}


// calcstress for one patch: x (4) inout, bsun / bsha out; returns true where
// the call took the night branch (a positive x(sun) on entry), tran then
// holding its transpiration.  Used by the calcstress kernel and by
// PhotosynthesisHydraulicStress inside the CanopyFluxes iteration.
__device__ __forceinline__ bool calcstress(const Leaf &P, double *x, double gs_sun_in, double gs_sha_in, double qsatl,
                                           double qaf, double &bsun, double &bsha, double &tran)
{
    const bool night = x[0] > 0.0;
    if (night) x[0] = x[1];
    double gs0sun = gs_sun_in, gs0sha = gs_sha_in;
    double q_sun, q_sha;
    getqflx(P, qsatl, qaf, true, gs0sun, gs0sha, q_sun, q_sha);
    bool flag;
    if ((P.laisun > TOL_LAI || P.laisha > TOL_LAI) && (q_sun > 0.0 || q_sha > 0.0)) {
        int iter = 0;
        for (;;) {   // at most 51 passes: the iter > itmax exit
            iter = iter + 1;
            double f[4], iA[16], dx[4];
            spacF(P, x, f, q_sun, q_sha);
            double ff = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) ff = ff + f[k] * f[k];
            if (sqrt(ff) < 1.e-6 * (q_sun + q_sha)) { flag = false; break; }
            if (iter > 50) { flag = false; break; }
            flag = spacA(P, x, iA, q_sun, q_sha);
            if (flag) break;
            if (P.laisun > TOL_LAI && P.laisha > TOL_LAI) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double s = 0.0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) s = s + iA[r * 4 + k] * f[k];
                    dx[r] = s;
                }
            } else {
                dx[0] = 0.0;
#pragma unroll
                for (int r = 1; r < 4; ++r) {
                    double s = 0.0;
#pragma unroll
                    for (int k = 1; k < 4; ++k) s = s + iA[r * 4 + k] * f[k];
                    dx[r] = s;
                }
            }
            double m = fabs(dx[0]);
#pragma unroll
            for (int k = 1; k < 4; ++k) if (fabs(dx[k]) > m) m = fabs(dx[k]);
            if (m > 50000.0) {
#pragma unroll
                for (int k = 0; k < 4; ++k) dx[k] = 50000.0 * dx[k] / m;
            }
            if (P.laisun > TOL_LAI && P.laisha > TOL_LAI) {
#pragma unroll
                for (int k = 0; k < 4; ++k) x[k] = x[k] + dx[k];
            } else if (P.laisha > TOL_LAI) {
#pragma unroll
                for (int k = 0; k < 4; ++k) x[k] = x[k] + dx[k];
                x[0] = x[2];
            } else {
                x[2] = x[2] + dx[2];
                x[3] = x[3] + dx[3];
                x[0] = x[0] + dx[1];
                x[1] = x[2];
            }
            double dd = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) dd = dd + dx[k] * dx[k];
            if (sqrt(dd) < 1.e-9) break;
            if (x[2] > x[3]) x[2] = x[3];
            if (x[0] > x[2]) x[0] = x[2];
            if (x[1] > x[2]) x[1] = x[2];
        }
    } else {
        flag = true;
    }
    if (flag) {
        (void)getvegwp(P, x, gs0sun, gs0sha, qsatl, qaf);
        bsun = plc(x[0], P.psi50[0], P.ck[0]);
        bsha = plc(x[1], P.psi50[1], P.ck[1]);
    } else {
        double qs = q_sun * plc(x[0], P.psi50[0], P.ck[0]);
        double qh = q_sha * plc(x[1], P.psi50[1], P.ck[1]);
        getqflx(P, qsatl, qaf, false, gs0sun, gs0sha, qs, qh);
        bsun = (q_sun > 0.0) ? gs0sun / gs_sun_in : plc(x[0], P.psi50[0], P.ck[0]);
        bsha = (q_sha > 0.0) ? gs0sha / gs_sha_in : plc(x[1], P.psi50[1], P.ck[1]);
    }
    if (bsun < 0.01) bsun = 0.0;
    if (bsha < 0.01) bsha = 0.0;
    if (night) {
        gs0sun = bsun * gs_sun_in;
        gs0sha = bsha * gs_sha_in;
        double soilflux = getvegwp(P, x, gs0sun, gs0sha, qsatl, qaf);
        if (soilflux < 0.0) soilflux = 0.0;
        tran = soilflux;
    }
    return night;
}

}  // namespace phs
}  // namespace ctsm
