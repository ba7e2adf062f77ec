// canopy.hip -- CanopyFluxes (CanopyFluxesMod.F90:191-1765) for gfx950: one
// lane per patch runs the whole time step (root stress, roughness, Monin-
// Obukhov initialisation, the Newton / stability iteration with sun and shade
// photosynthesis, the post-iteration fluxes) with its state in registers.
//
// Callees restated here: FrictionVelocity / StabilityFunc1/2 / MoninObukIni
// (FrictionVelocityMod.F90:754-1209), QSat (QSatMod.F90:61-127),
// Photosynthesis / hybrid / brent / ci_func (PhotosynthesisMod.F90:1243-2701),
// quadratic (quadraticMod.F90:17-74), calc_effective_soilporosity /
// calc_volumetric_h2oliq / calc_root_moist_stress_clm45default
// (SoilMoistStressMod.F90:70-215, 312-514), PhotosynthesisTotal
// (PhotosynthesisMod.F90:2065-2151), truncate_small_values (NumericsMod.F90:
// 60-100); plus the transpiration p2c and sink (clm_driver.F90:1714-1716,
// SoilWaterPlantSinkMod.F90:332-424) in k_root_sink.
//
// Patches are independent, so the reference's shrinking filter of
// unconverged patches becomes a per-lane loop exit; the operation sequence of
// each patch is the reference's.  Leaf-temperature factors of Photosynthesis
// (ft, fth, 2**x) depend on t_veg only, so they are evaluated once per
// iteration and shared by the sun and shade phases (the same values the
// reference computes twice).
#include <climits>
#include <type_traits>

#include "common.h"
#include "phs_dev.h"
#include "trace.h"

namespace ctsm {
namespace canopy {

constexpr double VKC = 0.4;
constexpr double CPAIR = 1.00464e3;
constexpr double HVAP = 2.501e6;
constexpr double CWATER = 4.188e3;
constexpr double RGAS = 6.02214e26 * 1.38065e-23;
constexpr double PI = 3.14159265358979323846;
constexpr double SPVAL = 1.e36;
constexpr double TLSAI_CRIT = 2.0;
constexpr double ALPHA_AERO = 1.0;
constexpr double NU = 1.5e-5;
constexpr double CD1 = 7.5;
constexpr double CDRYBIO = 1400.0;
constexpr double C_TO_B = 2.0;
constexpr double MAX_CS = 1.e-06;
constexpr int ENDRUN = CTSM_E_ENDRUN;
// the Meier2022 fixed point u = a exp(u) converges or overflows to NaN
// within a few hundred steps for finite input; the cap only bounds a wave
constexpr int MAX_FIXED_POINT = 100000;

// PFT table field indices (rows of the device table, CTSM_MAX_PFT each)
enum {
    F_DLEAF, F_C3PSN, F_SLATOP, F_LEAFCN, F_FLNR, F_FNITR, F_MBBOPT, F_MEDSLOPE, F_MEDINT, F_SMPSO, F_SMPSC,
    F_THETA_CJ, F_Z0V_CR, F_Z0V_CS, F_Z0V_C, F_Z0V_CW, F_Z0V_LAIMAX, F_IS_TREE, F_IS_SHRUB, F_DBH, F_FBW,
    F_NSTEM, F_RSTEM, F_WOOD, N_PFT_FIELDS
};

struct CView {
    int begc, ldc, begp, ldp, begg, ldg, nlevsno, nlevsoi, nlevgrnd;
    __device__ __forceinline__ long c1(int c) const { return (long)(c - begc); }
    __device__ __forceinline__ long g1(int g) const { return (long)(g - begg); }
    __device__ __forceinline__ long p1(int p) const { return (long)(p - begp); }
    __device__ __forceinline__ long sno(int c, int j) const { return (long)(c - begc) + (long)(j + nlevsno - 1) * ldc; }
    __device__ __forceinline__ long soi(int c, int j) const { return (long)(c - begc) + (long)(j - 1) * ldc; }
    __device__ __forceinline__ long soip(int p, int j) const { return (long)(p - begp) + (long)(j - 1) * ldp; }
};

struct CSwitch {
    int z0param, stomatal, soil_resis, biomass, undercanopy, light_inhibit, itmax;
};

__device__ __forceinline__ double p4(double x) { const double x2 = x * x; return x2 * x2; }
__device__ __forceinline__ double p3(double x) { return (x * x) * x; }

// QSatMod.F90:61-127
__device__ __forceinline__ void qsat(double T, double p, double &qs, double &es, double &qsdT)
{
    const double td = fmin90(100.0, fmax90(-75.0, T - TFRZ));
    double e, d;
    if (td >= 0.0) {
        e = 6.11213476 + td * (0.444007856 + td * (0.143064234e-01 + td * (0.264461437e-03 + td * (0.305903558e-05
            + td * (0.196237241e-07 + td * (0.892344772e-10 + td * (-0.373208410e-12 + td * 0.209339997e-15)))))));
        d = 0.444017302 + td * (0.286064092e-01 + td * (0.794683137e-03 + td * (0.121211669e-04 + td * (0.103354611e-06
            + td * (0.404125005e-09 + td * (-0.788037859e-12 + td * (-0.114596802e-13 + td * 0.381294516e-16)))))));
    } else {
        e = 6.11123516 + td * (0.503109514 + td * (0.188369801e-01 + td * (0.420547422e-03 + td * (0.614396778e-05
            + td * (0.602780717e-07 + td * (0.387940929e-09 + td * (0.149436277e-11 + td * 0.262655803e-14)))))));
        d = 0.503277922 + td * (0.377289173e-01 + td * (0.126801703e-02 + td * (0.249468427e-04 + td * (0.313703411e-06
            + td * (0.257180651e-08 + td * (0.133268878e-10 + td * (0.394116744e-13 + td * 0.498070196e-16)))))));
    }
    es = e * 100.0;
    const double vp = 1.0 / (p - 0.378 * es);
    const double vp1 = 0.622 * vp;
    qs = es * vp1;
    const double esdT = d * 100.0;
    const double vp2 = vp1 * vp;
    qsdT = esdT * vp2 * p;
}

__device__ __forceinline__ double qsat_only(double T, double p, double &es)
{
    const double td = fmin90(100.0, fmax90(-75.0, T - TFRZ));
    double e;
    if (td >= 0.0)
        e = 6.11213476 + td * (0.444007856 + td * (0.143064234e-01 + td * (0.264461437e-03 + td * (0.305903558e-05
            + td * (0.196237241e-07 + td * (0.892344772e-10 + td * (-0.373208410e-12 + td * 0.209339997e-15)))))));
    else
        e = 6.11123516 + td * (0.503109514 + td * (0.188369801e-01 + td * (0.420547422e-03 + td * (0.614396778e-05
            + td * (0.602780717e-07 + td * (0.387940929e-09 + td * (0.149436277e-11 + td * 0.262655803e-14)))))));
    es = e * 100.0;
    const double vp = 1.0 / (p - 0.378 * es);
    return es * (0.622 * vp);
}

// quadraticMod.F90:17-74; false where the reference calls endrun
__device__ __forceinline__ bool quadratic(double a, double b, double c, double &r1, double &r2)
{
    if (a == 0.0) { r1 = r2 = 0.0; return false; }
    double root = b * b - 4.0 * a * c;
    bool ok = true;
    if (root < 0.0) {
        if (-root < 3.0 * 2.220446049250313e-16) root = 0.0;
        else { ok = false; root = 0.0; }
    }
    double q;
    if (b >= 0.0) q = -0.5 * (b + sqrt(root));
    else q = -0.5 * (b - sqrt(root));
    r1 = q / a;
    r2 = (q != 0.0) ? c / q : 1.e36;
    return ok;
}

__device__ __forceinline__ double stab1(double zeta)
{
    const double chik2 = sqrt(1.0 - 16.0 * zeta);
    const double chik = sqrt(chik2);
    return 2.0 * cm_log((1.0 + chik) * 0.5) + cm_log((1.0 + chik2) * 0.5) - 2.0 * cm_atan(chik) + PI * 0.5;
}
__device__ __forceinline__ double stab2(double zeta)
{
    const double chik2 = sqrt(1.0 - 16.0 * zeta);
    return 2.0 * cm_log((1.0 + chik2) * 0.5);
}

// the temperature / humidity profile block repeated in FrictionVelocityMod.F90:916-1060;
// log_z = log(zldis / z0), which does not change over the canopy iteration.
// The lanes of a wave sit in different stability regimes; each elementary
// function a regime needs is evaluated once for all the lanes that need it
// (the very unstable and the very stable regime share the logarithm of
// their height ratio, the two unstable ones stab2(z0 / obu)) rather than once
// per regime branch.  Every lane computes its own regime's expression with
// the same operations in the same order.
__device__ double temp_profile(double zldis, double z0, double obu, double log_z)
{
    const double zetat = 0.465;
    const double zeta = zldis / obu;
    const bool ra = zeta < -zetat, rb = !ra && zeta < 0.0, rc = !ra && !rb && zeta <= 1.0, rd = !(ra || rb || rc);
    double lg = 0.0, sa = 0.0, sb = 0.0, pw = 0.0, lz = 0.0;
    if (ra || rd) lg = cm_log(ra ? -zetat * obu / z0 : obu / z0);
    if (ra || rb) sb = stab2(z0 / obu);
    if (ra) {
        sa = stab2(-zetat);
        pw = cm_pow(zetat, -0.333) - cm_pow(-zeta, -0.333);
    }
    if (rb) sa = stab2(zeta);
    if (rd) lz = cm_log(zeta);
    double den;
    if (ra) den = lg - sa + sb + 0.8 * pw;
    else if (rb) den = log_z - sa + sb;
    else if (rc) den = log_z + 5.0 * zeta - 5.0 * z0 / obu;
    else den = lg + 5.0 - 5.0 * z0 / obu + (5.0 * lz + zeta - 1.0);
    return VKC / den;
}

__device__ __forceinline__ double fm_profile(double z)
{
    // dust-model fm of FrictionVelocityMod.F90:1072-1095 for z < 0
    const double tmp1 = cm_pow(1.0 - 16.0 * z, 0.25);
    const double tmp2 = cm_log((1.0 + tmp1 * tmp1) / 2.0);
    const double tmp3 = cm_log((1.0 + tmp1) / 2.0);
    return 2.0 * tmp3 + tmp2 - 2.0 * cm_atan(tmp1) + 1.5707963;
}

struct FV {
    double ustar, temp1, temp2, temp12m, temp22m, fm, u10, u10_clm, va, fv, vds;
};

// FrictionVelocityMod.F90:754-1117 (patch form), the part the canopy
// iteration consumes: ustar, the heat / moisture profile factors temp1 and
// temp2, and the running mean of the dust-model fm.  The logarithms of
// height over roughness length (lm: momentum, lh: heat, lq: moisture) are the
// step's constants.
struct FVLogs {
    double lm, lh, lq;
};
__device__ __forceinline__ void friction_velocity(double hu, double ht, double hq, double displa, double z0m,
                                                  double z0h, double z0q, const FVLogs &L, double obu, int iter,
                                                  double um, FV &o)
{
    const double zetam = 1.574;
    const double zldis = hu - displa;
    const double zeta = zldis / obu;
    // each elementary function once for the lanes whose regime needs it (as
    // in temp_profile)
    const bool ra = zeta < -zetam, rb = !ra && zeta < 0.0, rc = !ra && !rb && zeta <= 1.0, rd = !(ra || rb || rc);
    double lg = 0.0, sa = 0.0, sb = 0.0, pw = 0.0, lz = 0.0;
    if (ra || rd) lg = cm_log(ra ? -zetam * obu / z0m : obu / z0m);
    if (ra || rb) sb = stab1(z0m / obu);
    if (ra) {
        sa = stab1(-zetam);
        pw = cm_pow(-zeta, 0.333) - cm_pow(zetam, 0.333);
    }
    if (rb) sa = stab1(zeta);
    if (rd) lz = cm_log(zeta);
    double den;
    if (ra) den = lg - sa + sb + 1.14 * pw;
    else if (rb) den = L.lm - sa + sb;
    else if (rc) den = L.lm + 5.0 * zeta - 5.0 * z0m / obu;
    else den = lg + 5.0 - 5.0 * z0m / obu + (5.0 * lz + zeta - 1.0);
    o.ustar = VKC * um / den;
    o.temp1 = temp_profile(ht - displa, z0h, obu, L.lh);
    if (hq == ht && z0q == z0h) o.temp2 = o.temp1;
    else o.temp2 = temp_profile(hq - displa, z0q, obu, L.lq);
    const double zm1 = fmin90(zeta, 1.0);
    const double fmnew = (zm1 < 0.0) ? fm_profile(zm1) : -5.0 * zm1;
    o.fm = (iter == 1) ? fmnew : 0.5 * (o.fm + fmnew);
}

// The diagnostics FrictionVelocity computes beside them (deposition
// velocity, 10-m winds, 2-m profile factors).  Nothing in the iteration reads
// them, so they are evaluated once, from the inputs (obu, um) and the ustar
// and fm of the last iteration: the same values the reference keeps from its
// last call.
__device__ void friction_velocity_diag(double hu, double displa, double z0m, double z0h, double z0q, double obu,
                                       double ur, double um, FV &o)
{
    const double zetam = 1.574;
    const double zldis = hu - displa;
    const double zeta = zldis / obu;
    if (zeta < 0.0) o.vds = 2.e-3 * o.ustar * (1.0 + cm_pow(300.0 / (-obu), 0.666));
    else o.vds = 2.e-3 * o.ustar;
    if (zldis - z0m <= 10.0) {
        o.u10_clm = um;
    } else {
        const double z10 = 10.0 + z0m;
        // one evaluation of each elementary function for the lanes whose
        // regime needs it (as in friction_velocity)
        const bool ra = zeta < -zetam, rb = !ra && zeta < 0.0, rc = !ra && !rb && zeta <= 1.0;
        const bool rd = !(ra || rb || rc);
        const double lg = cm_log(ra ? -zetam * obu / z10 : (rd ? obu / z10 : zldis / z10));
        double sa = 0.0, sb = 0.0, pw = 0.0, lz = 0.0;
        if (ra || rb) sb = stab1(z10 / obu);
        if (ra) {
            sa = stab1(-zetam);
            pw = cm_pow(-zeta, 0.333) - cm_pow(zetam, 0.333);
        }
        if (rb) sa = stab1(zeta);
        if (rd) lz = cm_log(zeta);
        double t;
        if (ra) t = lg - sa + sb + 1.14 * pw;
        else if (rb) t = lg - sa + sb;
        else if (rc) t = lg + 5.0 * zeta - 5.0 * z10 / obu;
        else t = lg + 5.0 - 5.0 * z10 / obu + (5.0 * lz + zeta - 1.0);
        o.u10_clm = um - (o.ustar / VKC * t);
    }
    o.va = um;
    o.temp12m = temp_profile(2.0 + z0h, z0h, obu, cm_log((2.0 + z0h) / z0h));
    if (z0q == z0h) o.temp22m = o.temp12m;
    else o.temp22m = temp_profile(2.0 + z0q, z0q, obu, cm_log((2.0 + z0q) / z0q));
    double zeta10 = fmin90(10.0 / obu, 1.0);
    if (zeta == 0.0) zeta10 = 0.0;
    const double fm10 = (zeta10 < 0.0) ? fm_profile(zeta10) : -5.0 * zeta10;
    const double tmp4 = cm_log(fmax90(1.0, hu / 10.0));
    o.u10 = ur - o.ustar / VKC * (tmp4 - o.fm + fm10);
    o.fv = o.ustar;
}

// the exponents of ft and fth (PhotosynthesisMod.F90:2488-2561)
__device__ __forceinline__ double ft_arg(double tl, double ha)
{
    return ha / (RGAS * 1.e-3 * (TFRZ + 25.0)) * (1.0 - (TFRZ + 25.0) / tl);
}
__device__ __forceinline__ double fth_arg(double tl, double hd, double se)
{
    return (-hd + se * tl) / (RGAS * 1.e-3 * tl);
}
__device__ __forceinline__ double ft(double tl, double ha) { return cm_exp(ft_arg(tl, ha)); }
__device__ __forceinline__ double fth(double tl, double hd, double se, double scaleFactor)
{
    return scaleFactor / (1.0 + cm_exp(fth_arg(tl, hd, se)));
}
__device__ __forceinline__ double fth25(double hd, double se)
{
    return 1.0 + cm_exp((-hd + se * (TFRZ + 25.0)) / (RGAS * 1.e-3 * (TFRZ + 25.0)));
}

// One ci solve: the constants of ci_func and the photosyns_inst values it writes.
struct CiSolve {
    bool c3, medlyn;
    double forc_pbot, cair, oair, gb_mol, je, lmr_z, par_z, rh_can;
    double vcmax_z, tpu_z, kp_z, cp, kc, ko, qe, bbb, mbb, theta_cj, theta_ip, mslope, mint;
    double ac, aj, ap, ag, an, gs_mol;
    bool bad;
};

// PhotosynthesisMod.F90:2564-2701
__device__ double ci_func(double ci, CiSolve &s)
{
    double r1, r2;
    // the C3 Rubisco rate and the C4 PEP carboxylase rate are the one divide
    // of each pathway: a wave of both kinds takes it once, each lane on its
    // own pathway's operands
    const double q = (s.c3 ? s.vcmax_z * fmax90(ci - s.cp, 0.0) : s.kp_z * fmax90(ci, 0.0))
                     / (s.c3 ? ci + s.kc * (1.0 + s.oair / s.ko) : s.forc_pbot);
    if (s.c3) {
        s.ac = q;
        s.aj = s.je * fmax90(ci - s.cp, 0.0) / (4.0 * ci + 8.0 * s.cp);
        s.ap = 3.0 * s.tpu_z;
    } else {
        s.ac = s.vcmax_z;
        s.aj = s.qe * s.par_z * 4.6;
        s.ap = q;
    }
    if (!quadratic(s.theta_cj, -(s.ac + s.aj), s.ac * s.aj, r1, r2)) s.bad = true;
    const double ai = fmin90(r1, r2);
    if (!quadratic(s.theta_ip, -(ai + s.ap), ai * s.ap, r1, r2)) s.bad = true;
    s.ag = fmax90(0.0, fmin90(r1, r2));
    s.an = s.ag - s.lmr_z;
    if (s.an < 0.0) return 0.0;
    double cs = s.cair - 1.4 / s.gb_mol * s.an * s.forc_pbot;
    cs = fmax90(cs, MAX_CS);
    if (s.medlyn) {
        const double term = 1.6 * s.an / (cs / s.forc_pbot * 1.e06);
        const double bq = -(2.0 * (s.mint * 1.e-06 + term)
                            + (s.mslope * term) * (s.mslope * term) / (s.gb_mol * 1.e-06 * s.rh_can));
        const double cq = s.mint * s.mint * 1.e-12 + (2.0 * s.mint * 1.e-06 + term * (1.0 - s.mslope * s.mslope / s.rh_can)) * term;
        if (!quadratic(1.0, bq, cq, r1, r2)) s.bad = true;
        s.gs_mol = fmax90(r1, r2) * 1.e06;
    } else {
        const double bq = cs * (s.gb_mol - s.bbb) - s.mbb * s.an * s.forc_pbot;
        const double cq = -s.gb_mol * (cs * s.bbb + s.mbb * s.an * s.forc_pbot * s.rh_can);
        if (!quadratic(cs, bq, cq, r1, r2)) s.bad = true;
        s.gs_mol = fmax90(r1, r2);
    }
    return ci - s.cair + s.an * s.forc_pbot * (1.4 * s.gs_mol + 1.6 * s.gb_mol) / (s.gb_mol * s.gs_mol);
}

// hybrid (PhotosynthesisMod.F90:2251-2368) with its Brent fallback (brent,
// :2371-2485), as one loop with a single evaluation of ci_func per pass: each
// pass evaluates ci_func at the point the previous pass chose, then the
// stage the lane is in (the two starting points, a secant step, the Brent
// iteration, the re-evaluation at the best point) takes the result and
// chooses the next point or ends.  Each lane makes exactly the reference's
// sequence of evaluations and arithmetic; lanes in different stages share the
// evaluation (only the ci_func side effects matter to the caller).
__device__ void hybrid(double x0, CiSolve &s)
{
    const double eps = 1.e-2, eps1 = 1.e-4;
    enum { F_X0, F_X1, SECANT, BRENT, LAST };
    int stage = F_X0, iter = 0, biter = 0;
    double f0 = 0.0, x1 = 0.0, f1 = 0.0, minx = x0, minf = 0.0, tol = 0.0;
    double a = 0.0, b = 0.0, c = 0.0, fa = 0.0, fb = 0.0, fc = 0.0, d = 0.0, e = 0.0;
    double xe = x0;
    for (;;) {
        const double f = ci_func(xe, s);
        bool secant_next = false, brent_next = false;
        if (stage == F_X0) {
            f0 = f;
            if (f0 == 0.0) return;
            minx = x0;
            minf = f0;
            x1 = x0 * 0.99;
            xe = x1;
            stage = F_X1;
            continue;
        } else if (stage == F_X1) {
            f1 = f;
            if (f1 == 0.0) return;
            if (f1 < minf) {
                minx = x1;
                minf = f1;
            }
            iter = 1;
            secant_next = true;
        } else if (stage == SECANT) {
            f1 = f;
            if (f1 < minf) {
                minx = x1;
                minf = f1;
            }
            if (fabs(f1) <= eps1) return;
            if (f1 * f0 < 0.0) {
                // brent(x0, x1, f0, f1, tol): the bracket and its first state
                a = x0;
                b = x1;
                fa = f0;
                fb = f1;
                if ((fa > 0.0 && fb > 0.0) || (fa < 0.0 && fb < 0.0)) {
                    s.bad = true;
                    return;
                }
                c = b;
                fc = fb;
                biter = 0;
                brent_next = true;
            } else if (iter > 40) {
                xe = minx;
                stage = LAST;
                continue;
            } else {
                ++iter;
                secant_next = true;
            }
        } else if (stage == BRENT) {
            fb = f;
            if (fb == 0.0) return;
            ++biter;
            brent_next = true;
        } else {
            return;     // LAST: ci_func at the best point has run
        }
        if (secant_next) {
            const double dx = -f1 * (x1 - x0) / (f1 - f0);
            const double x = x1 + dx;
            tol = fabs(x) * eps;
            if (fabs(dx) < tol) return;
            x0 = x1;
            f0 = f1;
            x1 = x;
            xe = x1;
            stage = SECANT;
            continue;
        }
        // one Brent iteration up to its evaluation of ci_func
        if (biter >= 20) return;
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) {
            c = a;
            fc = fa;
            d = b - a;
            e = d;
        }
        if (fabs(fc) < fabs(fb)) {
            a = b;
            b = c;
            c = a;
            fa = fb;
            fb = fc;
            fc = fa;
        }
        const double tol1 = 2.0 * eps * fabs(b) + 0.5 * tol;
        const double xm = 0.5 * (c - b);
        if (fabs(xm) <= tol1 || fb == 0.0) return;
        if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
            const double ss = fb / fa;
            double p, q;
            if (a == c) {
                p = 2.0 * xm * ss;
                q = 1.0 - ss;
            } else {
                q = fa / fc;
                const double r = fb / fc;
                p = ss * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (ss - 1.0);
            }
            if (p > 0.0) q = -q;
            p = fabs(p);
            if (2.0 * p < fmin90(3.0 * xm * q - fabs(tol1 * q), fabs(e * q))) {
                e = d;
                d = p / q;
            } else {
                d = xm;
                e = d;
            }
        } else {
            d = xm;
            e = d;
        }
        a = b;
        fa = fb;
        if (fabs(d) > tol1) b = b + d;
        else b = b + copysign(fabs(tol1), xm);
        xe = b;
        stage = BRENT;
    }
}

// per-phase outputs of Photosynthesis
struct Phase {
    double rs, rs_z, ci_z, lmr, lmr_z, psn, psn_z, psn_wc, psn_wj, psn_wp, gs_ph, gs_ln;
};
// photosyns_inst values shared by the phases (last writer wins, as in the reference)
struct Shared {
    double vcmax_z, tpu_z, kp_z, gb_mol, ac, aj, ap, ag, an, gs_mol, rh_leaf, vpd_can;
};
// leaf-temperature factors of one iteration (each a value the reference
// computes inside Photosynthesis; products keep the reference's order)
struct TFac {
    double kc, ko, cp;
    double ftL, fthL, dL4;          // lmr: c3 ft*fth, c4 divisor
    double ftV, fthV, dV1, dV2;     // vcmax: c3 ft*fth, c4 divisors
    double ftJ, fthJ, ftT, fthT;    // jmax, tpu
    double pow2;                    // 2**((t_veg - (tfrz+25))/10)
};
// constants of the step for one patch
struct PsnConst {
    bool c3;
    double bbb, mbb, qe, lnc, vcmax25top, jmax25top, tpu25top, kp25top, lmr25top, lmrc;
    double vcmaxse, jmaxse, tpuse, vcmaxc, jmaxc, tpuc, kc25, ko25, cp25;
    double theta_cj, mslope, mint;
};

// PhotosynthesisMod.F90:1243-2062 for one patch and one phase.
template <class PRM>
__device__ void photosynthesis(const CSwitch &sw, const PRM &P, const PsnConst &K, const TFac &T,
                               int nrad, double par_z, double lai_z, double vcmaxcint, double o3v, double o3g,
                               bool noon, double forc_pbot, double tgcm, double esat_tv, double eair, double oair,
                               double cair, double rb, double btran, Phase &ph, Shared &sh, bool &bad)
{
    const bool medlyn = sw.stomatal == CTSM_STOMATA_MEDLYN2011;
    double jmax_z = 0.0;
    if (nrad >= 1) {
        const double nscaler = vcmaxcint;
        const double lmr25 = K.lmr25top * nscaler;
        if (K.c3) {
            ph.lmr_z = lmr25 * T.ftL * T.fthL;
        } else {
            ph.lmr_z = lmr25 * T.pow2;
            ph.lmr_z = ph.lmr_z / T.dL4;
        }
        if (par_z <= 0.0) {
            sh.vcmax_z = 0.0;
            jmax_z = 0.0;
            sh.tpu_z = 0.0;
            sh.kp_z = 0.0;
        } else {
            const double vcmax25 = K.vcmax25top * nscaler;
            const double jmax25 = K.jmax25top * nscaler;
            const double tpu25 = K.tpu25top * nscaler;
            const double kp25 = K.kp25top * nscaler;
            sh.vcmax_z = vcmax25 * T.ftV * T.fthV;
            jmax_z = jmax25 * T.ftJ * T.fthJ;
            sh.tpu_z = tpu25 * T.ftT * T.fthT;
            if (!K.c3) {
                sh.vcmax_z = vcmax25 * T.pow2;
                sh.vcmax_z = sh.vcmax_z / T.dV1;
                sh.vcmax_z = sh.vcmax_z / T.dV2;
            }
            sh.kp_z = kp25 * T.pow2;
        }
        sh.vcmax_z = sh.vcmax_z * btran;
        ph.lmr_z = ph.lmr_z * btran;
        if (sw.light_inhibit && par_z > 0.0) ph.lmr_z = ph.lmr_z * 0.67;
    }
    const double rsmax0 = 2.e4;
    const double cf = forc_pbot / (RGAS * 1.e-3 * tgcm) * 1.e06;
    const double gb = 1.0 / rb;
    sh.gb_mol = gb * cf;
    double psn_wc_z = 0.0, psn_wj_z = 0.0, psn_wp_z = 0.0;
    if (nrad >= 1) {
        if (par_z <= 0.0) {
            sh.ac = 0.0;
            sh.aj = 0.0;
            sh.ap = 0.0;
            sh.ag = 0.0;
            sh.an = sh.ag - ph.lmr_z;
            ph.psn_z = 0.0;
            if (!medlyn) ph.rs_z = fmin90(rsmax0, 1.0 / K.bbb * cf);
            else ph.rs_z = fmin90(rsmax0, 1.0 / K.mint * cf);
            ph.ci_z = 0.0;
            sh.rh_leaf = 0.0;
            ph.gs_ph = cf / ph.rs_z;
        } else {
            const double ceair = fmin90(eair, esat_tv);
            double rh_can;
            if (!medlyn) {
                rh_can = ceair / esat_tv;
            } else {
                rh_can = fmax90(esat_tv - ceair, 50.0) * 0.001;
                sh.vpd_can = rh_can;
            }
            const double qabs = 0.5 * (1.0 - P.fnps) * par_z * 4.6;
            double r1, r2;
            if (!quadratic(P.theta_psii, -(qabs + jmax_z), qabs * jmax_z, r1, r2)) bad = true;
            const double je = fmin90(r1, r2);
            const double ci0 = K.c3 ? 0.7 * cair : 0.4 * cair;
            CiSolve s;
            s.c3 = K.c3;
            s.medlyn = medlyn;
            s.forc_pbot = forc_pbot;
            s.cair = cair;
            s.oair = oair;
            s.gb_mol = sh.gb_mol;
            s.je = je;
            s.lmr_z = ph.lmr_z;
            s.par_z = par_z;
            s.rh_can = rh_can;
            s.vcmax_z = sh.vcmax_z;
            s.tpu_z = sh.tpu_z;
            s.kp_z = sh.kp_z;
            s.cp = T.cp;
            s.kc = T.kc;
            s.ko = T.ko;
            s.qe = K.qe;
            s.bbb = K.bbb;
            s.mbb = K.mbb;
            s.theta_cj = K.theta_cj;
            s.theta_ip = P.theta_ip;
            s.mslope = K.mslope;
            s.mint = K.mint;
            s.ac = sh.ac;
            s.aj = sh.aj;
            s.ap = sh.ap;
            s.ag = sh.ag;
            s.an = sh.an;
            s.gs_mol = sh.gs_mol;
            s.bad = false;
            hybrid(ci0, s);
            if (s.bad) bad = true;
            sh.ac = s.ac;
            sh.aj = s.aj;
            sh.ap = s.ap;
            sh.ag = s.ag;
            sh.an = s.an;
            sh.gs_mol = s.gs_mol;
            if (sh.an < 0.0) sh.gs_mol = medlyn ? K.mint : K.bbb;
            const double gs_mol = sh.gs_mol;
            ph.gs_ph = gs_mol;
            ph.gs_ln = noon ? gs_mol : SPVAL;
            const double an = sh.an;
            ph.ci_z = cair - an * forc_pbot * (1.4 * gs_mol + 1.6 * sh.gb_mol) / (sh.gb_mol * gs_mol);
            ph.ci_z = fmax90(ph.ci_z, 1.e-06);
            const double gs = gs_mol / cf;
            ph.rs_z = fmin90(1.0 / gs, rsmax0);
            ph.rs_z = ph.rs_z / o3g;
            ph.psn_z = sh.ag;
            ph.psn_z = ph.psn_z * o3v;
            if (sh.ac <= sh.aj && sh.ac <= sh.ap) psn_wc_z = ph.psn_z;
            else if (sh.aj < sh.ac && sh.aj <= sh.ap) psn_wj_z = ph.psn_z;
            else if (sh.ap < sh.ac && sh.ap < sh.aj) psn_wp_z = ph.psn_z;
            if (gs_mol < 0.0) bad = true;
            if (!medlyn) sh.rh_leaf = (sh.gb_mol * ceair + gs_mol * esat_tv) / ((sh.gb_mol + gs_mol) * esat_tv);
        }
    }
    double psncan = 0.0, psncan_wc = 0.0, psncan_wj = 0.0, psncan_wp = 0.0, lmrcan = 0.0, gscan = 0.0, laican = 0.0;
    if (nrad >= 1) {
        psncan = psncan + ph.psn_z * lai_z;
        psncan_wc = psncan_wc + psn_wc_z * lai_z;
        psncan_wj = psncan_wj + psn_wj_z * lai_z;
        psncan_wp = psncan_wp + psn_wp_z * lai_z;
        lmrcan = lmrcan + ph.lmr_z * lai_z;
        gscan = gscan + lai_z / (rb + ph.rs_z);
        laican = laican + lai_z;
    }
    if (laican > 0.0) {
        ph.psn = psncan / laican;
        ph.psn_wc = psncan_wc / laican;
        ph.psn_wj = psncan_wj / laican;
        ph.psn_wp = psncan_wp / laican;
        ph.lmr = lmrcan / laican;
        ph.rs = laican / gscan - rb;
    } else {
        ph.psn = 0.0;
        ph.psn_wc = 0.0;
        ph.psn_wj = 0.0;
        ph.psn_wp = 0.0;
        ph.lmr = 0.0;
        ph.rs = 0.0;
    }
}

#include "canopy_phs.h"

}  // namespace canopy
}  // namespace ctsm

namespace ctsm {
namespace canopy {

// constants of the time step for one patch (PhotosynthesisMod.F90:1525-1722)
// phs: PhotosynthesisHydraulicStress's constants (:3117-3300): the
// Ball-Berry minimum conductance without btran (the stress enters through
// bsun / bsha) and the top-leaf nitrogen capped at 10 g/m2
template <class PRM>
__device__ void psn_constants(const CSwitch &sw, const PRM &P, const double *__restrict__ pft, int ivt,
                              double forc_pbot, double oair, double t10, double btran, double dayl_factor,
                              PsnConst &K, bool &bad, bool phs = false)
{
    const bool medlyn = sw.stomatal == CTSM_STOMATA_MEDLYN2011;
    K.lmrc = fth25(P.lmrhd, P.lmrse);
    K.c3 = ((int)rint(pft[F_C3PSN * CTSM_MAX_PFT + ivt]) == 1);
    double bbbopt = 0.0;
    if (K.c3) {
        K.qe = 0.0;
        if (!medlyn) bbbopt = 10000.0;
    } else {
        K.qe = 0.05;
        if (!medlyn) bbbopt = 40000.0;
    }
    K.bbb = phs ? bbbopt : fmax90(bbbopt * btran, 1.0);
    K.mbb = pft[F_MBBOPT * CTSM_MAX_PFT + ivt];
    K.kc25 = P.kc25_coef * forc_pbot;
    K.ko25 = P.ko25_coef * forc_pbot;
    const double sco = 0.5 * 0.209 / P.cp25_yr2000;
    K.cp25 = 0.5 * oair / sco;
    const double slatop = pft[F_SLATOP * CTSM_MAX_PFT + ivt], leafcn = pft[F_LEAFCN * CTSM_MAX_PFT + ivt];
    if (slatop * leafcn <= 0.0) bad = true;
    K.lnc = 1.0 / (slatop * leafcn);
    if (phs) K.lnc = fmin90(K.lnc, 10.0);
    K.vcmax25top = K.lnc * pft[F_FLNR * CTSM_MAX_PFT + ivt] * P.fnr * P.act25 * dayl_factor;
    K.vcmax25top = K.vcmax25top * pft[F_FNITR * CTSM_MAX_PFT + ivt];
    const double t10c = fmin90(fmax90(t10 - TFRZ, 11.0), 35.0);
    K.jmax25top = ((2.59 - 0.035 * t10c) * K.vcmax25top) * P.jmax25top_sf;
    K.tpu25top = P.tpu25ratio * K.vcmax25top;
    K.kp25top = P.kp25ratio * K.vcmax25top;
    K.lmr25top = K.c3 ? K.vcmax25top * P.leaf_mr_vcm : K.vcmax25top * 0.025;
    K.vcmaxse = (668.39 - 1.07 * t10c) * P.vcmaxse_sf;
    K.jmaxse = (659.70 - 0.75 * t10c) * P.jmaxse_sf;
    K.tpuse = (668.39 - 1.07 * t10c) * P.tpuse_sf;
    K.vcmaxc = fth25(P.vcmaxhd, K.vcmaxse);
    K.jmaxc = fth25(P.jmaxhd, K.jmaxse);
    K.tpuc = fth25(P.tpuhd, K.tpuse);
    K.theta_cj = pft[F_THETA_CJ * CTSM_MAX_PFT + ivt];
    K.mslope = pft[F_MEDSLOPE * CTSM_MAX_PFT + ivt];
    K.mint = pft[F_MEDINT * CTSM_MAX_PFT + ivt];
}

template <class PRM>
__device__ void temp_factors(const PRM &P, const PsnConst &K, double tv, bool anyday, TFac &T)
{
    T.kc = K.kc25 * ft(tv, P.kcha);
    T.ko = K.ko25 * ft(tv, P.koha);
    T.cp = K.cp25 * ft(tv, P.cpha);
    T.pow2 = cm_pow(2.0, (tv - (TFRZ + 25.0)) / 10.0);
    // C3 and C4 lanes share a wave: the exponentials of the two pathways'
    // factors are evaluated once, each lane on its own pathway's argument
    const double eL = cm_exp(K.c3 ? ft_arg(tv, P.lmrha) : 1.3 * (tv - (TFRZ + 55.0)));
    if (K.c3) {
        T.ftL = eL;
        T.fthL = fth(tv, P.lmrhd, P.lmrse, K.lmrc);
    } else {
        T.dL4 = 1.0 + eL;
    }
    if (anyday) {
        const double eV1 = cm_exp(K.c3 ? ft_arg(tv, P.vcmaxha) : 0.2 * ((TFRZ + 15.0) - tv));
        const double eV2 = cm_exp(K.c3 ? fth_arg(tv, P.vcmaxhd, K.vcmaxse) : 0.3 * (tv - (TFRZ + 40.0)));
        if (K.c3) {
            T.ftV = eV1;
            T.fthV = K.vcmaxc / (1.0 + eV2);
        } else {
            T.dV1 = 1.0 + eV1;
            T.dV2 = 1.0 + eV2;
        }
        T.ftJ = ft(tv, P.jmaxha);
        T.fthJ = fth(tv, P.jmaxhd, K.jmaxse, K.jmaxc);
        T.ftT = ft(tv, P.tpuha);
        T.fthT = fth(tv, P.tpuhd, K.tpuse, K.tpuc);
    }
}

#define PF(f) pft[(f) * CTSM_MAX_PFT + ivt]

// Iteration state of a patch carried between rounds, indexed by the patch's
// position in the filter (structure of arrays).
struct CanopyWork {
    double *t_veg, *del, *efeb, *obuold, *nmozsgn, *um, *obu, *fm, *qsatl, *el, *qsatldT, *taf, *qaf, *delq,
        *itlef, *gs_mol;
};
constexpr int N_WORK_FIELDS = 16;
static_assert(sizeof(CanopyWork) == N_WORK_FIELDS * sizeof(double *), "one pointer per saved value");

// One patch of CanopyFluxes for at most `budget` leaf-temperature
// iterations.  first: the time step's set-up runs (root stress, roughness,
// Monin-Obukhov initialisation); otherwise the iteration state is restored
// from W.  Returns true when the patch is still iterating (its state saved in
// W[i0]); false when it finished (outputs written) or failed.
// PHS: CanopyFluxes with plant hydraulic stress, H its PhsArgs (unused
// otherwise): PhotosynthesisHydraulicStress in place of the two Photosynthesis
// calls and the branches of CanopyFluxesMod.F90:1219-1230, 1353-1356.
// The column-level half of calc_effective_soilporosity, calc_volumetric_h2oliq
// and calc_root_moist_stress (SoilMoistStressMod.F90:70-126, 190-240,
// 420-514): effective porosity, volumetric liquid water and, where the level
// takes up water, the soil matric potential of CH78 before the PFT's floor
// and eff / watsat.  The reference computes them once per patch of the filter
// (CanopyFluxesMod.F90:833-891 builds a column filter with one entry per
// patch); here the 64 lanes of a wave share the work of the runs of lanes
// whose patches share a column, one (run, level) item per lane at a time, and
// each lane's patch then reads its column's terms from LDS.  Same operations on
// the same inputs as the per-patch form, so the same bits.  kNoUptake marks a
// level that takes up nothing (no liquid, or frozen below -2 C): a signalling
// NaN, which no arithmetic produces.
constexpr long long kNoUptake = 0x7ff0dead00000001LL;

struct SoilTerms {
    double smp;     // -sucsat s^-bsw, or kNoUptake
    double ratio;   // eff_porosity / watsat
};

// Returns the index of the lane's run (-1 for a lane without a valid patch);
// every lane of the wave calls it.
template <class COL, class CAN>
__device__ int soil_terms_of_wave(const CView &V, const COL &col, const CAN &S, const int32_t *__restrict__ filter,
                                  int i0, bool active, SoilTerms *terms, int *run_col)
{
    const int lane = threadIdx.x;
    int c = -1;
    if (active) {
        // the patches canopy_patch accepts (it reports the others)
        const long pp = V.p1(filter ? filter[i0] : V.begp + i0);
        const int cp = S.column[pp], g = S.gridcell[pp], ivt = S.itype[pp];
        if (!(ivt < 0 || ivt >= CTSM_MAX_PFT || cp < V.begc || cp >= V.begc + V.ldc || g < V.begg ||
              g >= V.begg + V.ldg))
            c = cp;
    }
    const int prev = __shfl_up(c, 1);
    const bool lead = c >= 0 && (lane == 0 || prev != c);
    const unsigned long long leaders = __ballot(lead);
    if (lead) run_col[__popcll(leaders & ((1ull << lane) - 1ull))] = c;
    wave_sync();
    const int nl = V.nlevgrnd;
    const int items = __popcll(leaders) * nl;
    for (int it = lane; it < items; it += 64) {
        const int r = it / nl, j = it - r * nl + 1;
        const int rc = run_col[r];
        const long cs = V.soi(rc, j), cn = V.sno(rc, j);
        const double watsat = col.watsat[cs], dz = col.dz[cn];
        const double vol_ice = fmin90(watsat, col.h2osoi_ice[cn] / (DENICE * dz));
        const double eff = watsat - vol_ice;
        col.eff_porosity[cs] = eff;
        const double liqvol = fmin90(eff, col.h2osoi_liq[cn] / (dz * DENH2O));
        S.h2osoi_liqvol[cs] = liqvol;
        SoilTerms t{__longlong_as_double(kNoUptake), 0.0};
        if (!(liqvol <= 0.0 || col.t_soisno[cn] <= TFRZ - 2.0)) {
            const double s_node = fmax90(liqvol / eff, 0.01);
            t.smp = -col.sucsat[cs] * cm_pow(s_node, -col.bsw[cs]);
            t.ratio = eff / watsat;
        }
        terms[r * MAX_NLEVGRND + j - 1] = t;
    }
    wave_sync();
    return c >= 0 ? __popcll(leaders & ((2ull << lane) - 1ull)) - 1 : -1;
}

template <bool PHS, class PRM, class COL, class CAN, class WRK, class PHA>
__device__ __forceinline__ bool canopy_patch(const CView &V, const CSwitch &sw, const PRM &P,
                             const double *__restrict__ pft, const COL &col, const CAN &S,
                             int i0, const int32_t *__restrict__ filter, double dtime, ErrWord *err, bool first,
                             int budget, const WRK &W, const PHA &H, const SoilTerms *soil, double *rr_lane)
{
    const int p = filter ? filter[i0] : V.begp + i0;
    const long pp = V.p1(p);
    const int c = S.column[pp];
    const int g = S.gridcell[pp];
    const int ivt = S.itype[pp];
    // indices that address other arrays must lie inside their allocations
    if (ivt < 0 || ivt >= CTSM_MAX_PFT || c < V.begc || c >= V.begc + V.ldc || g < V.begg || g >= V.begg + V.ldg) {
        report_error(err, p, CTSM_E_ARG);
        return false;
    }
    const long cc = V.c1(c), gg = V.g1(g);
    const double btran0 = 0.0, zii = 1000.0, beta = 1.0, delmax = 1.0, dlemin = 0.1, dtmin = 0.01;
    const int itmin = 2;
    const double ria = 0.5;
    const double k_vert = 0.1, k_cyl_vol = 1.0, k_cyl_area = 1.0, k_internal = 0.0;
    const double min_stem_diameter = 0.05, min_lai = 0.1;
    bool bad = false;

    const double elai = S.elai[pp], esai = S.esai[pp], htop = S.htop[pp];
    const double emv = S.emv[pp], thm = S.thm[pp];
    const double forc_pbot = S.forc_pbot[cc], forc_q = S.forc_q[cc], forc_th = S.forc_th[cc];
    const double forc_rho = S.forc_rho[cc], thv = S.thv[cc], qg = S.qg[cc];
    const double t_grnd = col.t_grnd[cc], emg = col.emg[cc];

    // ---- biomass heat capacities (CanopyFluxesMod.F90:717-819)
    double frac_rad_abs_by_stem, sa_leaf, sa_stem, sa_internal, cp_leaf, cp_stem, rstem;
    double stem_biomass = 0.0, leaf_biomass = 0.0;
    if (sw.biomass) {
        frac_rad_abs_by_stem = (esai) / (elai + esai);
        if (elai > 0.0) frac_rad_abs_by_stem = k_vert * frac_rad_abs_by_stem;
        const double dbh = PF(F_DBH);
        sa_leaf = elai;
        sa_leaf = 2.0 * sa_leaf;
        sa_stem = PF(F_NSTEM) * (htop * PI * dbh);
        sa_stem = k_cyl_area * sa_stem;
        if (!((int)PF(F_IS_TREE) || (int)PF(F_IS_SHRUB)) || dbh < min_stem_diameter) {
            frac_rad_abs_by_stem = 0.0;
            sa_stem = 0.0;
            sa_leaf = sa_leaf + esai;
        } else {
            if (elai < min_lai) sa_leaf = sa_leaf + esai;
        }
        const double fbw = PF(F_FBW);
        leaf_biomass = (1.e-3 * C_TO_B / PF(F_SLATOP)) * fmax90(0.01, 0.5 * sa_leaf) / (1.0 - fbw);
        const double carea_stem = PI * ((dbh * 0.5) * (dbh * 0.5));
        stem_biomass = carea_stem * htop * k_cyl_vol * PF(F_NSTEM) * PF(F_WOOD) / (1.0 - fbw);
        sa_internal = fmin90(sa_leaf, sa_stem);
        sa_internal = k_internal * sa_internal;
        cp_leaf = leaf_biomass * (CDRYBIO * (1.0 - fbw) + (fbw) * CWATER);
        cp_stem = stem_biomass * (CDRYBIO * (1.0 - fbw) + (fbw) * CWATER);
        cp_stem = k_cyl_vol * cp_stem;
        rstem = PF(F_RSTEM) * dbh;
        S.leaf_biomass[pp] = leaf_biomass;
        S.stem_biomass[pp] = stem_biomass;
    } else {
        sa_leaf = (elai + esai);
        frac_rad_abs_by_stem = 0.0;
        sa_stem = 0.0;
        sa_internal = 0.0;
        cp_leaf = 0.0;
        cp_stem = 0.0;
        rstem = 0.0;
    }
    const double dayl = S.dayl[gg], max_dayl = S.max_dayl[gg];
    const double dayl_factor = fmin90(1.0, fmax90(0.01, (dayl * dayl) / (max_dayl * max_dayl)));

    // ---- effective porosity, volumetric liquid water, root moisture stress
    double btran = btran0;
    if (!first) {
        btran = S.btran[pp];
    } else {
        // rootfr * rresis waits in the lane's LDS column (rr_lane, stride 64)
        // until btran is known; rootr is then written once, normalised
        // (SoilMoistStressMod.F90:480-500).  A run-time level loop keeps the
        // per-level offsets out of the kernel's uniform registers.  The
        // column's terms (effective porosity, liquid water, the matric
        // potential) are the wave's, in soil[] (soil_terms_of_wave).
        const double smpsc = PF(F_SMPSC), smpso = PF(F_SMPSO);
        for (int j = 1; j <= V.nlevgrnd; ++j) {
            const SoilTerms t = soil[j - 1];
            double rr = 0.0;
            if (__double_as_longlong(t.smp) != kNoUptake) {
                double smp_node = t.smp;
                smp_node = fmax90(smpsc, smp_node);
                const double rres = fmin90(t.ratio * (smp_node - smpsc) / (smpso - smpsc), 1.0);
                S.rresis[V.soip(p, j)] = rres;
                rr = S.rootfr[V.soip(p, j)] * rres;
                btran = btran + fmax90(rr, 0.0);
            }
            rr_lane[(j - 1) * 64] = rr;
        }
        for (int j = 1; j <= V.nlevgrnd; ++j) {
            const long sp = V.soip(p, j);
            S.rootr[sp] = (btran > btran0) ? rr_lane[(j - 1) * 64] / btran : 0.0;
        }
        S.btran[pp] = btran;
    }

    // ---- aerodynamic parameters (CanopyFluxesMod.F90:897-943)
    double displa = S.displa[pp], z0mv = S.z0mv[pp];
    if (!first) {
        // already updated in place by the first round
    } else if (sw.z0param == CTSM_Z0_ZENGWANG2007) {
        const double lt = fmin90(elai + esai, TLSAI_CRIT);
        const double egvf = (1.0 - ALPHA_AERO * cm_exp(-lt)) / (1.0 - ALPHA_AERO * cm_exp(-TLSAI_CRIT));
        displa = egvf * displa;
        z0mv = cm_exp(egvf * cm_log(z0mv) + (1.0 - egvf) * cm_log(S.z0mg[cc]));
    } else {
        double lt = fmax90(1.e-5, elai + esai);
        displa = htop * (1.0 - (1.0 - cm_exp(-cm_pow(CD1 * lt, 0.5))) / cm_pow(CD1 * lt, 0.5));
        lt = fmin90(lt, PF(F_Z0V_LAIMAX));
        double delt = 2.0;
        const double zc = PF(F_Z0V_C);
        const double U_ustar_ini = cm_pow(PF(F_Z0V_CS) + PF(F_Z0V_CR) * lt * 0.5, -0.5) * zc * lt * 0.25;
        double U_ustar = U_ustar_ini;
        int guard = 0;
        while (delt > 1.e-4) {
            const double prev = U_ustar;
            U_ustar = U_ustar_ini * cm_exp(prev);
            delt = fabs(U_ustar - prev);
            if (++guard > MAX_FIXED_POINT) { bad = true; break; }
        }
        U_ustar = 4.0 * U_ustar / lt / zc;
        const double zcw = PF(F_Z0V_CW);
        z0mv = htop * (1.0 - displa / htop) * cm_exp(-VKC * U_ustar + cm_log(zcw) - 1.0 + 1.0 / zcw);
    }
    const double z0hv = z0mv, z0qv = z0mv;
    const double hu = S.forc_hgt_u[gg] + z0mv + displa;
    const double ht = S.forc_hgt_t[gg] + z0hv + displa;
    const double hq = S.forc_hgt_q[gg] + z0qv + displa;
    if (first) {
        S.displa[pp] = displa;
        S.z0mv[pp] = z0mv;
        S.z0hv[pp] = z0hv;
        S.z0qv[pp] = z0qv;
        S.forc_hgt_u_patch[pp] = hu;
        S.forc_hgt_t_patch[pp] = ht;
        S.forc_hgt_q_patch[pp] = hq;
    }

    // ---- radiation terms, leaf humidity, initial profiles (:945-1002)
    const double forc_lwrad = col.forc_lwrad[cc];
    const double air = emv * (1.0 + (1.0 - emv) * (1.0 - emg)) * forc_lwrad;
    const double bir = -(2.0 - emv * (1.0 - emg)) * emv * SB;
    const double cir = emv * emg * SB;
    double t_veg, qsatl, el, qsatldT, taf, qaf, dth = 0.0, dqh = 0.0, delq, um, obu;
    int nmozsgn;
    const double co2 = S.forc_pco2[gg], o2 = S.forc_po2[gg];
    const double fu = S.forc_u[gg], fvv = S.forc_v[gg];
    const double ur = fmax90(P.wind_min, sqrt(fu * fu + fvv * fvv));
    const double zldis = hu - displa;
    if (first) {
        t_veg = S.t_veg[pp];
        qsat(t_veg, forc_pbot, qsatl, el, qsatldT);
        nmozsgn = 0;
        taf = (t_grnd + thm) / 2.0;
        qaf = (forc_q + qg) / 2.0;
        dth = thm - taf;
        dqh = forc_q - qaf;
        delq = qg - qaf;
        const double dthv = dth * (1.0 + 0.61 * forc_q) + 0.61 * forc_th * dqh;
        if (zldis < 0.0) {
            report_error(err, p, ENDRUN);
            return false;
        }
        // FrictionVelocityMod.F90:1158-1209
        const double wc = 0.5;
        if (dthv >= 0.0) um = fmax90(ur, 0.1);
        else um = sqrt(ur * ur + wc * wc);
        const double rib = GRAV * zldis * dthv / (thv * um * um);
        double zeta0;
        if (rib >= 0.0) {
            zeta0 = rib * cm_log(zldis / z0mv) / (1.0 - 5.0 * fmin90(rib, 0.19));
            zeta0 = fmin90(P.zetamaxstable, fmax90(zeta0, 0.01));
        } else {
            zeta0 = rib * cm_log(zldis / z0mv);
            zeta0 = fmax90(-100.0, fmin90(zeta0, -0.01));
        }
        obu = zldis / zeta0;
    } else {
        t_veg = W.t_veg[i0];
        qsatl = W.qsatl[i0];
        el = W.el[i0];
        qsatldT = W.qsatldT[i0];
        taf = W.taf[i0];
        qaf = W.qaf[i0];
        delq = W.delq[i0];
        um = W.um[i0];
        obu = W.obu[i0];
        nmozsgn = (int)W.nmozsgn[i0];
    }
    const double tl_ini = S.t_veg[pp];
    double t_stem = S.t_stem[pp];
    const double ts_ini = t_stem;

    // per-step constants of Photosynthesis
    const int nrad = S.nrad[pp];
    const double parsun = S.parsun_z[pp], parsha = S.parsha_z[pp];
    const bool noon = S.near_local_noon[gg] != 0;
    const bool anyday = nrad >= 1 && (parsun > 0.0 || parsha > 0.0);
    const double lai_z_sun = S.laisun_z[pp], lai_z_sha = S.laisha_z[pp];
    const double vcint_sun = S.vcmaxcintsun[pp], vcint_sha = S.vcmaxcintsha[pp];
    const double o3v_sun = S.o3coefvsun[pp], o3v_sha = S.o3coefvsha[pp];
    const double o3g_sun = S.o3coefgsun[pp], o3g_sha = S.o3coefgsha[pp];
    PsnConst K;
    psn_constants(sw, P, pft, ivt, forc_pbot, o2, S.t_a10[pp], btran, dayl_factor, K, bad, PHS);
    // The iteration carries only its state from one iteration to the next
    // (and, through W, from round to round); every value the outputs read
    // after the last iteration is that iteration's own, so nothing else stays
    // live across the loop.  Photosynthesis writes some of its outputs only
    // on branches fixed for the step (nrad, par, the stomatal model): those
    // are stored under the same conditions (otherwise the arrays keep their
    // values, as in the reference).
    double gs_mol_c = first ? S.gs_mol[pp] : W.gs_mol[i0];
    double del = first ? 0.0 : W.del[i0], efeb = first ? 0.0 : W.efeb[i0];
    double obuold = first ? 0.0 : W.obuold[i0];
    double fm_c = first ? 0.0 : W.fm[i0];
    const double laisun = S.laisun[pp], laisha = S.laisha[pp], fdry = S.fdry[pp], fwet = S.fwet[pp];
    const double liqcan0 = S.liqcan[pp], snocan0 = S.snocan[pp];
    const double h2ocan = liqcan0 + snocan0;
    const double sabv = S.sabv[pp];
    const double fvn = (double)S.frac_veg_nosno[pp];
    const double z0mg = S.z0mg[cc];
    const double dleaf = PF(F_DLEAF);
    const int snl = col.snl[cc];
    if (snl < -V.nlevsno || snl > 0) {
        report_error(err, p, CTSM_E_ARG);
        return false;
    }
    const double frac_sno = col.frac_sno_eff[cc], frac_h2osfc = col.frac_h2osfc[cc];
    const double t_top = col.t_soisno[V.sno(c, snl + 1)], t_soi1 = col.t_soisno[V.sno(c, 1)];
    const double t_h2osfc = col.t_h2osfc[cc];
    const double lw_grnd = (frac_sno * p4(t_top) + (1.0 - frac_sno - frac_h2osfc) * p4(t_soi1)
                            + frac_h2osfc * p4(t_h2osfc));
    const double lai_dl_eff = P.lai_dl * (1.0 - fmin90(col.snow_depth[cc] / P.z_dl, 1.0));
    // step constants of the iteration
    const double w_veg = cm_exp(-(elai + esai));
    const double dl_ext = 1.0 - cm_exp(-lai_dl_eff);
    FVLogs fvlog;
    fvlog.lm = cm_log(zldis / z0mv);
    fvlog.lh = cm_log((ht - displa) / z0hv);
    fvlog.lq = cm_log((hq - displa) / z0qv);

    int itlef = first ? 0 : (int)W.itlef[i0];
    // values of the iteration that the outputs read: the loop leaves for the
    // outputs only after a whole iteration, so each holds that iteration's value
    double ustar, temp1, temp2, tlbef, del2, ram1, rah1, raw1, uaf, rb, rah2, raw2, svpts, eah, rh_af, vpd;
    double wtg, wtl0, wta0, wtstem0, wtga, wtal, lw_stem, lw_leaf, qflx_tran_veg, wtlq0, wtaq0, wtalq;
    double wtgq, eflx_sh_stem, efe, dt_veg, errp, qflx_evap_veg, eflx_sh_veg, zeta, num_iter;
    double obu_fv, um_fv;
    FV fv;
    TFac T;
    Phase sun, sha;
    Shared sh;
    for (int it = 0;; ++it) {
        if (it >= budget) {
            // still iterating: keep the state for the next round
            W.t_veg[i0] = t_veg;
            W.del[i0] = del;
            W.efeb[i0] = efeb;
            W.obuold[i0] = obuold;
            W.nmozsgn[i0] = (double)nmozsgn;
            W.um[i0] = um;
            W.obu[i0] = obu;
            W.fm[i0] = fm_c;
            W.qsatl[i0] = qsatl;
            W.el[i0] = el;
            W.qsatldT[i0] = qsatldT;
            W.taf[i0] = taf;
            W.qaf[i0] = qaf;
            W.delq[i0] = delq;
            W.itlef[i0] = (double)itlef;
            W.gs_mol[i0] = gs_mol_c;
            if (bad) report_error(err, p, ENDRUN);
            return true;
        }
        wtgq = 0.0;
        fv.fm = fm_c;
        friction_velocity(hu, ht, hq, displa, z0mv, z0hv, z0qv, fvlog, obu, itlef + 1, um, fv);
        obu_fv = obu;       // the inputs of this FrictionVelocity call
        um_fv = um;
        ustar = fv.ustar;
        temp1 = fv.temp1;
        temp2 = fv.temp2;
        tlbef = t_veg;
        del2 = del;
        ram1 = 1.0 / (ustar * ustar / um);
        rah1 = 1.0 / (temp1 * ustar);
        raw1 = 1.0 / (temp2 * ustar);
        uaf = um * sqrt(1.0 / (ram1 * um));
        const double uuc = fmin90(0.4, (0.03 * um / ustar));
        const double cf = P.cv / (sqrt(uaf) * sqrt(dleaf));
        rb = 1.0 / (cf * uaf);
        const double w = w_veg;
        const double csoilb = VKC / (P.a_coef * cm_pow(z0mg * uaf / NU, P.a_exp));
        const double ri = (GRAV * htop * (taf - t_grnd)) / (taf * (uaf * uaf));
        double csoilcn;
        if (sw.undercanopy && (taf - t_grnd) > 0.0) {
            const double ricsoilc = P.csoilc / (1.00 + ria * fmin90(ri, 10.0));
            csoilcn = csoilb * w + ricsoilc * (1.0 - w);
        } else {
            csoilcn = csoilb * w + P.csoilc * (1.0 - w);
        }
        rah2 = sw.biomass ? 1.0 / (csoilcn * uuc) : 1.0 / (csoilcn * uaf);
        raw2 = rah2;
        svpts = el;
        eah = forc_pbot * qaf / 0.622;
        rh_af = eah / svpts;
        vpd = fmax90((svpts - eah), 50.0) * 0.001;

        // Photosynthesis, sun then shade
        sun = Phase{};
        sha = Phase{};
        sh = Shared{};
        sh.gs_mol = gs_mol_c;
        temp_factors(P, K, t_veg, anyday, T);
        if constexpr (PHS) {
            // sun and shade together; btran and the transpiration come from the hydraulics
            btran = photosynthesis_phs(V, sw, P, K, T, H, S, p, c, ivt, nrad, noon, forc_pbot, thm, svpts, eah, o2,
                                       co2, rb, qsatl, qaf, sun, sha, sh, bad);
            S.btran[pp] = btran;
            qflx_tran_veg = S.qflx_tran_veg[pp];
        } else {
            // one call site for both phases: sun (k = 0), then shade
#pragma nounroll
            for (int k = 0; k < 2; ++k) {
                const bool ks = k == 0;
                Phase ph{};
                photosynthesis(sw, P, K, T, nrad, ks ? parsun : parsha, ks ? lai_z_sun : lai_z_sha,
                               ks ? vcint_sun : vcint_sha, ks ? o3v_sun : o3v_sha, ks ? o3g_sun : o3g_sha, noon,
                               forc_pbot, thm, svpts, eah, o2, co2, rb, btran, ph, sh, bad);
                if (ks) sun = ph;
                else sha = ph;
            }
        }

        // conductances and the leaf-temperature Newton step (:1119-1331)
        const double wta = 1.0 / rah1;
        const double wtl = sa_leaf / rb;
        wtg = 1.0 / rah2;
        const double wtstem = sa_stem / (rstem + rb);
        const double wtshi = 1.0 / (wta + wtl + wtstem + wtg);
        wtl0 = wtl * wtshi;
        const double wtg0 = wtg * wtshi;
        wta0 = wta * wtshi;
        wtstem0 = wtstem * wtshi;
        wtga = wta0 + wtg0 + wtstem0;
        wtal = wta0 + wtl0 + wtstem0;
        lw_stem = sa_internal * emv * SB * p4(t_stem);
        lw_leaf = sa_internal * emv * SB * p4(t_veg);
        double rppdry;
        if (fdry > 0.0) rppdry = fdry * rb * (laisun / (rb + sun.rs) + laisha / (rb + sha.rs)) / elai;
        else rppdry = 0.0;
        double efpot = forc_rho * ((elai + esai) / rb) * (qsatl - qaf);
        double rpp;
        if (PHS) {
            if (efpot > 0.0) {
                if (btran > btran0) rpp = rppdry + fwet;
                else rpp = fwet;
                rpp = fmin90(rpp, (qflx_tran_veg + h2ocan / dtime) / efpot);
            } else {
                rpp = 1.0;
            }
        } else if (efpot > 0.0) {
            if (btran > btran0) {
                qflx_tran_veg = efpot * rppdry;
                rpp = rppdry + fwet;
            } else {
                rpp = fwet;
                qflx_tran_veg = 0.0;
            }
            rpp = fmin90(rpp, (qflx_tran_veg + h2ocan / dtime) / efpot);
        } else {
            rpp = 1.0;
            qflx_tran_veg = 0.0;
        }
        const double wtaq = fvn / raw1;
        const double wtlq = fvn * (elai + esai) / rb * rpp;
        const double rdl = dl_ext / (0.004 * uaf);
        // the branches' divides share one evaluation (operands per lane)
        {
            const bool lee = sw.soil_resis == CTSM_SOILRESIS_LEEPIELKE1992, sl14 = sw.soil_resis == CTSM_SOILRESIS_SL14;
            if (delq < 0.0 || lee || sl14) {
                const bool wet = delq < 0.0;
                const double num = (wet || !lee) ? fvn : S.soilbeta[cc] * fvn;
                const double den = (wet || lee) ? raw2 + rdl : raw2 + S.soilresis[cc];
                wtgq = num / den;
            }
        }
        const double wtsqi = 1.0 / (wtaq + wtlq + wtgq);
        const double wtgq0 = wtgq * wtsqi;
        wtlq0 = wtlq * wtsqi;
        wtaq0 = wtaq * wtsqi;
        const double wtgaq = wtaq0 + wtgq0;
        wtalq = wtaq0 + wtlq0;
        const double dc1 = forc_rho * CPAIR * wtl;
        const double dc2 = HVAP * forc_rho * wtlq;
        const double efsh = dc1 * (wtga * t_veg - wtg0 * t_grnd - wta0 * thm - wtstem0 * t_stem);
        eflx_sh_stem = forc_rho * CPAIR * wtstem * ((wta0 + wtg0 + wtl0) * t_stem - wtg0 * t_grnd - wta0 * thm - wtl0 * t_veg);
        efe = dc2 * (wtgaq * qsatl - wtgq0 * qg - wtaq0 * forc_q);
        double erre = 0.0;
        if (efe * efeb < 0.0) {
            const double efeold = efe;
            efe = 0.1 * efeold;
            erre = efe - efeold;
        }
        const double fr = frac_rad_abs_by_stem;
        dt_veg = ((1.0 - fr) * (sabv + air + bir * p4(t_veg) + cir * lw_grnd)
                  - efsh - efe - lw_leaf + lw_stem - (cp_leaf / dtime) * (t_veg - tl_ini))
                 / ((1.0 - fr) * (-4.0 * bir * p3(t_veg)) + 4.0 * sa_internal * emv * SB * p3(t_veg)
                    + dc1 * wtga + dc2 * wtgaq * qsatldT + cp_leaf / dtime);
        t_veg = tlbef + dt_veg;
        const double dels = dt_veg;
        del = fabs(dels);
        errp = 0.0;
        if (del > delmax) {
            dt_veg = delmax * dels / del;
            t_veg = tlbef + dt_veg;
            errp = (1.0 - fr) * (sabv + air + bir * p3(tlbef) * (tlbef + 4.0 * dt_veg) + cir * lw_grnd)
                   - sa_internal * emv * SB * p3(tlbef) * (tlbef + 4.0 * dt_veg)
                   + lw_stem
                   - (efsh + dc1 * wtga * dt_veg) - (efe + dc2 * wtgaq * qsatldT * dt_veg)
                   - (cp_leaf / dtime) * (t_veg - tl_ini);
        }
        efpot = forc_rho * ((elai + esai) / rb) * (wtgaq * (qsatl + qsatldT * dt_veg) - wtgq0 * qg - wtaq0 * forc_q);
        qflx_evap_veg = rpp * efpot;
        if (!PHS) {
            if (efpot > 0.0 && btran > btran0) qflx_tran_veg = efpot * rppdry;
            else qflx_tran_veg = 0.0;
        }
        const double ecidif = fmax90(0.0, qflx_evap_veg - qflx_tran_veg - h2ocan / dtime);
        qflx_evap_veg = fmin90(qflx_evap_veg, qflx_tran_veg + h2ocan / dtime);
        eflx_sh_veg = efsh + dc1 * wtga * dt_veg + errp + erre + HVAP * ecidif;
        eflx_sh_stem = eflx_sh_stem + forc_rho * CPAIR * wtstem * (-wtl0 * dt_veg);
        lw_leaf = sa_internal * emv * SB * p3(tlbef) * (tlbef + 4.0 * dt_veg);
        qsat(t_veg, forc_pbot, qsatl, el, qsatldT);
        taf = wtg0 * t_grnd + wta0 * thm + wtl0 * t_veg + wtstem0 * t_stem;
        qaf = wtlq0 * qsatl + wtgq0 * qg + forc_q * wtaq0;
        dth = thm - taf;
        dqh = forc_q - qaf;
        delq = wtalq * qg - wtlq0 * qsatl - wtaq0 * forc_q;
        const double tstar = temp1 * dth;
        const double qstar = temp2 * dqh;
        const double thvstar = tstar * (1.0 + 0.61 * forc_q) + 0.61 * forc_th * qstar;
        zeta = zldis * VKC * GRAV * thvstar / ((ustar * ustar) * thv);
        if (zeta >= 0.0) {
            zeta = fmin90(P.zetamaxstable, fmax90(zeta, 0.01));
            um = fmax90(ur, 0.1);
        } else {
            zeta = fmax90(-100.0, fmin90(zeta, -0.01));
            double wc;
            if (ustar * thvstar > 0.0) wc = 0.0;
            else wc = beta * cm_pow(-GRAV * ustar * thvstar * zii / thv, 0.333);
            um = sqrt(ur * ur + wc * wc);
        }
        obu = zldis / zeta;
        if (obuold * obu < 0.0) nmozsgn = nmozsgn + 1;
        if (nmozsgn >= 4) obu = zldis / (-0.01);
        obuold = obu;

        gs_mol_c = sh.gs_mol;
        fm_c = fv.fm;
        itlef = itlef + 1;
        bool converged = false;
        num_iter = 0.0;
        if (itlef > itmin) {
            const double dele = fabs(efe - efeb);
            efeb = efe;
            const double det = fmax90(del, del2);
            num_iter = (double)itlef;
            converged = det < dtmin && dele < dlemin;
        }
        if (converged || itlef > sw.itmax) break;
    }

    // ---- the last iteration is done: the fluxes and outputs of the step
    friction_velocity_diag(hu, displa, z0mv, z0hv, z0qv, obu_fv, ur, um_fv, fv);
    const double temp12m = fv.temp12m;
    const double temp22m = fv.temp22m;

    // ---- fluxes after the iteration (:1467-1615)
    const double fr = frac_rad_abs_by_stem;
    errp = (1.0 - fr) * (sabv + air + bir * p3(tlbef) * (tlbef + 4.0 * dt_veg) + cir * lw_grnd)
           - lw_leaf + lw_stem - eflx_sh_veg - HVAP * qflx_evap_veg - ((t_veg - tl_ini) * cp_leaf / dtime);
    double dt_stem = 0.0;
    double dhsdt_canopy = 0.0;
    if (sw.biomass) {
        if (stem_biomass > 0.0) {
            dt_stem = (fr * (sabv + air + bir * p4(ts_ini) + cir * lw_grnd) - eflx_sh_stem + lw_leaf - lw_stem)
                      / (cp_stem / dtime - fr * bir * 4.0 * p3(ts_ini));
        } else {
            dt_stem = 0.0;
        }
        dhsdt_canopy = dt_stem * cp_stem / dtime + (t_veg - tl_ini) * cp_leaf / dtime;
        t_stem = t_stem + dt_stem;
    }
    const double delt = wtal * t_grnd - wtl0 * t_veg - wta0 * thm - wtstem0 * t_stem;
    S.taux[pp] = -forc_rho * fu / ram1;
    S.tauy[pp] = -forc_rho * fvv / ram1;
    S.eflx_sh_grnd[pp] = CPAIR * forc_rho * wtg * delt;
    const double delt_snow = wtal * t_top - wtl0 * t_veg - wta0 * thm - wtstem0 * t_stem;
    const double delt_soil = wtal * t_soi1 - wtl0 * t_veg - wta0 * thm - wtstem0 * t_stem;
    const double delt_h2osfc = wtal * t_h2osfc - wtl0 * t_veg - wta0 * thm - wtstem0 * t_stem;
    S.eflx_sh_snow[pp] = CPAIR * forc_rho * wtg * delt_snow;
    S.eflx_sh_soil[pp] = CPAIR * forc_rho * wtg * delt_soil;
    S.eflx_sh_h2osfc[pp] = CPAIR * forc_rho * wtg * delt_h2osfc;
    S.qflx_evap_soi[pp] = forc_rho * wtgq * delq;
    const double delq_snow = wtalq * S.qg_snow[cc] - wtlq0 * qsatl - wtaq0 * forc_q;
    S.qflx_ev_snow[pp] = forc_rho * wtgq * delq_snow;
    const double delq_soil = wtalq * S.qg_soil[cc] - wtlq0 * qsatl - wtaq0 * forc_q;
    S.qflx_ev_soil[pp] = forc_rho * wtgq * delq_soil;
    const double delq_h2osfc = wtalq * S.qg_h2osfc[cc] - wtlq0 * qsatl - wtaq0 * forc_q;
    S.qflx_ev_h2osfc[pp] = forc_rho * wtgq * delq_h2osfc;
    const double t_ref2m = thm + temp1 * dth * (1.0 / temp12m - 1.0 / temp1);
    S.t_ref2m[pp] = t_ref2m;
    S.t_ref2m_r[pp] = t_ref2m;
    const double q_ref2m = forc_q + temp2 * dqh * (1.0 / temp22m - 1.0 / temp2);
    S.q_ref2m[pp] = q_ref2m;
    double e_ref2m;
    const double qsat_ref2m = qsat_only(t_ref2m, forc_pbot, e_ref2m);
    const double rh_ref2m = fmin90(100.0, q_ref2m / qsat_ref2m * 100.0);
    S.rh_ref2m[pp] = rh_ref2m;
    S.rh_ref2m_r[pp] = rh_ref2m;
    S.vpd_ref2m[pp] = e_ref2m * (1.0 - rh_ref2m / 100.0);
    S.dlrad[pp] = (1.0 - emv) * emg * forc_lwrad
                  + emv * emg * SB * p3(tlbef) * (tlbef + 4.0 * dt_veg) * (1.0 - fr)
                  + emv * emg * SB * p3(ts_ini) * (ts_ini + 4.0 * dt_stem) * fr;
    S.ulrad[pp] = ((1.0 - emg) * (1.0 - emv) * (1.0 - emv) * forc_lwrad
                   + emv * (1.0 + (1.0 - emg) * (1.0 - emv)) * SB * p3(tlbef) * (tlbef + 4.0 * dt_veg) * (1.0 - fr)
                   + emv * (1.0 + (1.0 - emg) * (1.0 - emv)) * SB * p3(ts_ini) * (ts_ini + 4.0 * dt_stem) * fr
                   + emg * (1.0 - emv) * SB * lw_grnd);
    S.t_skin[pp] = emv * t_veg + (1.0 - emv) * sqrt(sqrt(lw_grnd));
    const double cgrnds = S.cgrnds[pp] + CPAIR * forc_rho * wtg * wtal;
    const double cgrndl = S.cgrndl[pp] + forc_rho * wtgq * wtalq * S.dqgdT[cc];
    S.cgrnds[pp] = cgrnds;
    S.cgrndl[pp] = cgrndl;
    S.cgrnd[pp] = cgrnds + cgrndl * col.htvp[cc];
    double snocan = snocan0, liqcan = liqcan0;
    const double snocan_baseline = snocan;
    if (t_veg > TFRZ) {
        if ((qflx_evap_veg - qflx_tran_veg) * dtime > liqcan)
            snocan = fmax90(0.0, snocan + liqcan + (qflx_tran_veg - qflx_evap_veg) * dtime);
        liqcan = fmax90(0.0, liqcan + (qflx_tran_veg - qflx_evap_veg) * dtime);
    } else if (t_veg <= TFRZ) {
        if ((qflx_evap_veg - qflx_tran_veg) * dtime > snocan)
            liqcan = liqcan + snocan + (qflx_tran_veg - qflx_evap_veg) * dtime;
        snocan = fmax90(0.0, snocan + (qflx_tran_veg - qflx_evap_veg) * dtime);
    }
    if (fabs(snocan) < 1.e-10 * fabs(snocan_baseline)) snocan = 0.0;
    S.snocan[pp] = snocan;
    S.liqcan[pp] = liqcan;

    // ---- PhotosynthesisTotal and the noon water-use efficiency
    const double fpsn = sun.psn * laisun + sha.psn * laisha;
    S.fpsn[pp] = fpsn;
    S.fpsn_wc[pp] = sun.psn_wc * laisun + sha.psn_wc * laisha;
    S.fpsn_wj[pp] = sun.psn_wj * laisun + sha.psn_wj * laisha;
    S.fpsn_wp[pp] = sun.psn_wp * laisun + sha.psn_wp * laisha;
    double iwue = SPVAL;
    if (noon && fpsn > 0.0) {
        const double gs = 1.e-6 * (laisun * sun.gs_ph + laisha * sha.gs_ph);
        iwue = (gs > 0.0) ? fpsn / gs : SPVAL;
    }
    S.iwue_ln[pp] = iwue;

    // ---- remaining outputs
    S.btran[pp] = btran;
    S.dhsdt_canopy[pp] = dhsdt_canopy;
    S.eflx_sh_stem[pp] = eflx_sh_stem;
    S.eflx_sh_veg[pp] = eflx_sh_veg;
    S.qflx_tran_veg[pp] = qflx_tran_veg;
    S.qflx_evap_veg[pp] = qflx_evap_veg;
    S.t_veg[pp] = t_veg;
    S.t_stem[pp] = t_stem;
    S.um[pp] = um;
    S.obu[pp] = obu;
    S.num_iter[pp] = num_iter;
    S.ustar[pp] = ustar;
    S.u10[pp] = fv.u10;
    S.u10_clm[pp] = fv.u10_clm;
    S.va[pp] = fv.va;
    S.fv[pp] = fv.fv;
    S.vds[pp] = fv.vds;
    S.ram1[pp] = ram1;
    S.rb1[pp] = rb;
    S.rah1[pp] = rah1;
    S.rah2[pp] = rah2;
    S.raw1[pp] = raw1;
    S.raw2[pp] = raw2;
    S.uaf[pp] = uaf;
    S.taf[pp] = taf;
    S.qaf[pp] = qaf;
    S.zeta[pp] = zeta;
    S.vpd[pp] = vpd;
    S.rh_af[pp] = rh_af;
    S.dleaf_patch[pp] = dleaf;
    S.canopy_err[pp] = errp;
    // photosyns_inst
    S.c3flag[pp] = K.c3 ? 1 : 0;
    S.qe[pp] = K.qe;
    if (sw.stomatal == CTSM_STOMATA_BB1987) {
        S.bbb[pp] = K.bbb;
        S.mbb[pp] = K.mbb;
    }
    S.kc[pp] = T.kc;
    S.ko[pp] = T.ko;
    S.cp[pp] = T.cp;
    S.lnca[pp] = K.lnc;
    S.gb_mol[pp] = sh.gb_mol;
    if (!PHS) S.gs_mol[pp] = sh.gs_mol;
    const bool medlyn = sw.stomatal == CTSM_STOMATA_MEDLYN2011;
    const bool lit_sun = nrad >= 1 && parsun > 0.0, lit_sha = nrad >= 1 && parsha > 0.0;
    if (nrad >= 1) {
        // written by both phases' canopy layer whether lit or not (the PHS
        // path has arrays of its own for these, written as it goes)
        if (!PHS) {
            S.vcmax_z[pp] = sh.vcmax_z;
            S.tpu_z[pp] = sh.tpu_z;
            S.kp_z[pp] = sh.kp_z;
            S.ac[pp] = sh.ac;
            S.aj[pp] = sh.aj;
            S.ap[pp] = sh.ap;
            S.ag[pp] = sh.ag;
            S.an[pp] = sh.an;
        }
        S.rssun_z[pp] = sun.rs_z;
        S.rssha_z[pp] = sha.rs_z;
        S.cisun_z[pp] = sun.ci_z;
        S.cisha_z[pp] = sha.ci_z;
        S.lmrsun_z[pp] = sun.lmr_z;
        S.lmrsha_z[pp] = sha.lmr_z;
        S.psnsun_z[pp] = sun.psn_z;
        S.psnsha_z[pp] = sha.psn_z;
        S.gs_mol_sun[pp] = sun.gs_ph;
        S.gs_mol_sha[pp] = sha.gs_ph;
        // rh_leaf: a dark phase writes 0, a lit one under Ball-Berry its
        // value; vpd_can: a lit phase under Medlyn (the last writer wins)
        if (!PHS && (!medlyn || !lit_sun || !lit_sha)) S.rh_leaf[pp] = sh.rh_leaf;
        if (!PHS && medlyn && (lit_sun || lit_sha)) S.vpd_can[pp] = sh.vpd_can;
    }
    if (!PHS && lit_sun) S.gs_mol_sun_ln[pp] = sun.gs_ln;
    if (!PHS && lit_sha) S.gs_mol_sha_ln[pp] = sha.gs_ln;
    S.rssun[pp] = sun.rs;
    S.rssha[pp] = sha.rs;
    S.lmrsun[pp] = sun.lmr;
    S.lmrsha[pp] = sha.lmr;
    S.psnsun[pp] = sun.psn;
    S.psnsha[pp] = sha.psn;
    S.psnsun_wc[pp] = sun.psn_wc;
    S.psnsun_wj[pp] = sun.psn_wj;
    S.psnsun_wp[pp] = sun.psn_wp;
    S.psnsha_wc[pp] = sha.psn_wc;
    S.psnsha_wj[pp] = sha.psn_wj;
    S.psnsha_wp[pp] = sha.psn_wp;
    if (bad) report_error(err, p, ENDRUN);
    return false;
}

// Everything one launch of k_canopy_round reads: the kernel's only argument.
struct RoundArgs {
    CView V;
    CSwitch sw;
    ctsm_canopy_params P;
    const double *pft;
    ctsm_col_state col;
    ctsm_canopy_state S;
    int n;
    const int32_t *filter;
    double dtime;
    ErrWord *err;
    int round, budget;
    const int32_t *list_in, *cnt_in;
    int32_t *list_out, *cnt_out;
    CanopyWork W;
};
// the PHS form's argument: the same, then the PHS arrays
struct RoundArgsPhs {
    RoundArgs r;
    PhsArgs phs;
};
#define KARG __attribute__((address_space(4)))

// One round of CanopyFluxes over the patches still iterating (round 0: the
// whole filter), compacting those that continue into list_out with one atomic
// per wavefront.  Blocks are one wavefront, a patch per lane.  Round 0
// (STRIDE false) covers the filter with its grid; later rounds (STRIDE true)
// stride a capped grid over the round's count.
// The argument is read in place from the kernel-argument segment, each member
// where it is used: loaded all at once on entry, its few hundred uniform
// pointers would not fit the scalar registers and spill.  In the striding
// form the segment pointer is made opaque on every pass, so those loads stay
// inside the pass instead of being hoisted in front of the loop.
// WAVES: the register budget the kernel is compiled for (waves per SIMD);
// above one wave the compiler spills what does not fit to scratch.  The
// launch picks the instance (a launch shape the autotuner times).
// PHS: CanopyFluxes with plant hydraulic stress (argument RoundArgsPhs).
template <int WAVES, bool STRIDE, bool PHS = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES)))
k_canopy_round(typename std::conditional<PHS, RoundArgsPhs, RoundArgs>::type args)
{
    const KARG RoundArgs *Ap = (const KARG RoundArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    const int round = Ap->round;
    const int count = (round == 0) ? Ap->n : *Ap->cnt_in;
    const int lane = threadIdx.x;
    // round 0 runs the step's set-up; it is always launched with STRIDE false
    // (the instance that holds the LDS of the soil terms)
    const bool first = round == 0;
    for (int base = blockIdx.x * 64; base < count; base += gridDim.x * 64) {
        const KARG RoundArgs *A = Ap;
        if (STRIDE) asm volatile("" : "+s"(A));
        const int idx = base + lane;
        bool cont = false;
        int i0 = (idx < count) ? ((round == 0) ? idx : A->list_in[idx]) : 0;
        const CView V{A->V.begc, A->V.ldc, A->V.begp, A->V.ldp, A->V.begg, A->V.ldg, A->V.nlevsno,
                      A->V.nlevsoi, A->V.nlevgrnd};
        const SoilTerms *soil = nullptr;
        double *rr_lane = nullptr;
        if constexpr (!STRIDE) {
            if (first) {
                __shared__ SoilTerms terms[64 * MAX_NLEVGRND];
                __shared__ double rr[MAX_NLEVGRND * 64];
                __shared__ int run_col[64];
                const int run = soil_terms_of_wave(V, A->col, A->S, A->filter, i0, idx < count, terms, run_col);
                soil = terms + (run < 0 ? 0 : run) * MAX_NLEVGRND;
                rr_lane = rr + lane;
            }
        }
        if (idx < count) {
            const CSwitch sw{A->sw.z0param, A->sw.stomatal, A->sw.soil_resis, A->sw.biomass, A->sw.undercanopy,
                             A->sw.light_inhibit, A->sw.itmax};
            if constexpr (PHS) {
                const KARG RoundArgsPhs *B = (const KARG RoundArgsPhs *)A;
                cont = canopy_patch<true>(V, sw, A->P, A->pft, A->col, A->S, i0, A->filter, A->dtime, A->err,
                                          first, A->budget, A->W, B->phs, soil, rr_lane);
            } else {
                cont = canopy_patch<false>(V, sw, A->P, A->pft, A->col, A->S, i0, A->filter, A->dtime, A->err,
                                           first, A->budget, A->W, *A, soil, rr_lane);
            }
        }
        const unsigned long long mask = __ballot(cont);
        if (mask) {
            const int leader = __ffsll((unsigned long long)mask) - 1;
            int pos = 0;
            if (lane == leader) pos = atomicAdd(A->cnt_out, __popcll(mask));
            pos = __shfl(pos, leader);
            if (cont) A->list_out[pos + __popcll(mask & ((1ull << lane) - 1ull))] = i0;
        }
        if (!STRIDE) break;
    }
}

__global__ void __launch_bounds__(64) k_zero_counts(int32_t *cnt, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) cnt[i] = 0;
}
#undef PF

// photosyns_inst%TimeStepInit (PhotosynthesisMod.F90:1143-1201) and
// rb1(begp:endp) = 0 (CanopyFluxesMod.F90:830): every patch of the bounds,
// the n patches from begp or, where the bounds are the filter's own patches,
// the patches filter names.
__global__ void __launch_bounds__(256) k_canopy_init(int n, const int32_t *__restrict__ filter, int begp,
                                                     ctsm_canopy_state S)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const long i = filter ? (long)(filter[f] - begp) : (long)f;
    S.psnsun[i] = 0.0; S.psnsun_wc[i] = 0.0; S.psnsun_wj[i] = 0.0; S.psnsun_wp[i] = 0.0;
    S.psnsha[i] = 0.0; S.psnsha_wc[i] = 0.0; S.psnsha_wj[i] = 0.0; S.psnsha_wp[i] = 0.0;
    S.fpsn[i] = 0.0; S.fpsn_wc[i] = 0.0; S.fpsn_wj[i] = 0.0; S.fpsn_wp[i] = 0.0;
    S.rb1[i] = 0.0;
}

// cgrnd / cgrnds / cgrndl = 0 before CanopyFluxes (BiogeophysPreFluxCalcsMod.F90:390-392)
__global__ void __launch_bounds__(256) k_zero_cgrnd(int n, const int32_t *__restrict__ filter, int begp,
                                                    ctsm_canopy_state S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long pp = (long)((filter ? filter[i] : begp + i) - begp);
    S.cgrnd[pp] = 0.0;
    S.cgrnds[pp] = 0.0;
    S.cgrndl[pp] = 0.0;
}

// p2c_1d_filter of qflx_tran_veg (clm_driver.F90:1714-1716) and
// Compute_EffecRootFrac_And_VertTranSink_Default (SoilWaterPlantSinkMod.F90:332-424)
// for one column, reading its patches straight from HBM
__device__ __forceinline__ void root_sink_column(const CView &V, const ctsm_col_state &col,
                                                 const ctsm_canopy_state &S, int c, ErrWord *err)
{
    const long cc = V.c1(c);
    const int pi = col.patchi[cc], pf = col.patchf[cc];
    if (pf >= pi && (pi < V.begp || pf >= V.begp + V.ldp)) {   // patches outside the allocation
        report_error(err, c, CTSM_E_ARG);
        return;
    }
    // the p2c sum and the sink's normaliser are the same sum in the same order
    double tcol = 0.0;
    double r[MAX_NLEVSOI];
#pragma unroll
    for (int j = 0; j < MAX_NLEVSOI; ++j) r[j] = 0.0;
    for (int p = pi; p <= pf; ++p) {
        const long q = V.p1(p);
        if (!S.active[q]) continue;
        const double qt = S.qflx_tran_veg[q], w = S.wtcol[q];
        tcol = tcol + qt * w;
        // per level the patches are summed in ascending order, as the reference's j-outer loop does
#pragma unroll
        for (int j = 1; j <= MAX_NLEVSOI; ++j)
            if (j <= V.nlevsoi) r[j - 1] = r[j - 1] + S.rootr[V.soip(p, j)] * qt * w;
    }
    S.qflx_tran_veg_col[cc] = tcol;
#pragma unroll
    for (int j = 1; j <= MAX_NLEVSOI; ++j) {
        if (j > V.nlevsoi) continue;
        double rc = r[j - 1];
        if (tcol != 0.0) rc = rc / tcol;
        S.rootr_col[V.soi(c, j)] = rc;
        S.qflx_rootsoi_col[V.soi(c, j)] = rc * tcol;
    }
}

__global__ void __launch_bounds__(256) k_root_sink(CView V, ctsm_col_state col, ctsm_canopy_state S, int n,
                                                   const int32_t *__restrict__ filter, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    root_sink_column(V, col, S, filter ? filter[i] : V.begc + i, err);
}

// The same for consecutive columns (no filter) whose patches are consecutive:
// a block stages the patch range of RS_COLS columns in LDS (weights and
// activity once, then rootr RS_GROUPS levels at a time, every element read once
// and coalesced); thread group g sums level j0 + g of its column over the
// column's patches in ascending order from LDS -- the operations and their
// order are those of root_sink_column.  A block whose patches exceed the
// staging space takes root_sink_column.
constexpr int RS_COLS = 64;
constexpr int RS_GROUPS = 4;
constexpr int RS_CAP = 16 * RS_COLS;

__global__ void __launch_bounds__(RS_COLS *RS_GROUPS) k_root_sink_staged(CView V, ctsm_col_state col,
                                                                         ctsm_canopy_state S, int n, ErrWord *err)
{
    __shared__ double s_qt[RS_CAP], s_w[RS_CAP], s_row[RS_GROUPS][RS_CAP];
    __shared__ int32_t s_act[RS_CAP];
    __shared__ int s_lo, s_hi, s_bad;
    const int t = threadIdx.x;
    const int lane = t % RS_COLS, g = t / RS_COLS;
    const int i = blockIdx.x * RS_COLS + lane;
    const int c = V.begc + i;
    const long cc = V.c1(c);
    const bool live = i < n;
    int pi = 0, pf = -1;
    if (live) {
        pi = col.patchi[cc];
        pf = col.patchf[cc];
    }
    if (t == 0) { s_lo = INT_MAX; s_hi = INT_MIN; s_bad = 0; }
    __syncthreads();
    if (g == 0 && live && pf >= pi) {
        if (pi < V.begp || pf >= V.begp + V.ldp) atomicOr(&s_bad, 1);
        atomicMin(&s_lo, pi);
        atomicMax(&s_hi, pf);
    }
    __syncthreads();
    const int lo = s_lo, hi = s_hi;
    const int np = hi >= lo ? hi - lo + 1 : 0;
    if (s_bad || np > RS_CAP) {              // block-uniform
        if (g == 0 && live) root_sink_column(V, col, S, c, err);
        return;
    }
    for (int k = t; k < np; k += RS_COLS * RS_GROUPS) {
        const long q = V.p1(lo + k);
        s_qt[k] = S.qflx_tran_veg[q];
        s_w[k] = S.wtcol[q];
        s_act[k] = S.active[q];
    }
    __syncthreads();
    // every group forms its column's total (the same operations, so the same bits)
    double tcol = 0.0;
    for (int p = pi; p <= pf; ++p) {
        const int k = p - lo;
        if (!s_act[k]) continue;
        tcol = tcol + s_qt[k] * s_w[k];
    }
    if (g == 0 && live) S.qflx_tran_veg_col[cc] = tcol;
    for (int j0 = 1; j0 <= V.nlevsoi; j0 += RS_GROUPS) {
        __syncthreads();                     // the previous levels have been summed
        for (int r = 0; r < RS_GROUPS && j0 + r <= V.nlevsoi; ++r)
            for (int k = t; k < np; k += RS_COLS * RS_GROUPS) s_row[r][k] = S.rootr[V.soip(lo + k, j0 + r)];
        __syncthreads();
        const int j = j0 + g;
        if (!live || j > V.nlevsoi) continue;
        double r = 0.0;
        for (int p = pi; p <= pf; ++p) {
            const int k = p - lo;
            if (!s_act[k]) continue;
            r = r + s_row[g][k] * s_qt[k] * s_w[k];
        }
        double rc = r;
        if (tcol != 0.0) rc = rc / tcol;
        S.rootr_col[V.soi(c, j)] = rc;
        S.qflx_rootsoi_col[V.soi(c, j)] = rc * tcol;
    }
}

}  // namespace canopy

// Rounds after the first one: enough to take a patch from first_its
// iterations to itmax + 1.
static int canopy_later_rounds(int itmax, int first_its, int round_its)
{
    const int left = itmax + 1 - first_its;
    return left > 0 ? (left + round_its - 1) / round_its : 0;
}

// Round buffer: N_WORK_FIELDS doubles per patch of the filter, two patch
// lists, one counter per round.  0 when the iteration runs in one pass.
size_t canopy_work_bytes(int n, int itmax, int first_its, int round_its)
{
    using namespace canopy;
    const int R = canopy_later_rounds(itmax, first_its, round_its);
    if (R == 0 || n <= 0) return 0;
    return (size_t)N_WORK_FIELDS * n * sizeof(double) + 2 * (size_t)n * sizeof(int32_t) +
           (size_t)(R + 1) * sizeof(int32_t);
}

void launch_canopy(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_canopy_params &prm,
                   const double *pft_dev, const ctsm_col_state &col, const ctsm_canopy_state &can, int n,
                   const int32_t *filter, double dtime, ErrWord *err, void *work, int first_its, int round_its,
                   bool zero_cgrnd, int waves, bool bounds_are_filter, const ctsm_phs_state *phs_q,
                   const ctsm_phs_canopy_state *phs_h, const double *phs_table)
{
    using namespace canopy;
    const GptlRange range("canflux");
    const bool hydrstress = phs_h != nullptr;
    auto first_kernel = waves == 3 ? k_canopy_round<3, false> : waves == 2 ? k_canopy_round<2, false>
                                                                          : k_canopy_round<1, false>;
    auto later_kernel = waves == 3 ? k_canopy_round<3, true> : waves == 2 ? k_canopy_round<2, true>
                                                                         : k_canopy_round<1, true>;
    CView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, L.begg_alloc, L.ldg, cfg.nlevsno, cfg.nlevsoi, cfg.nlevgrnd};
    CSwitch sw{cfg.z0param_method, cfg.stomatalcond_method, cfg.soil_resis_method, cfg.use_biomass_heat_storage,
               cfg.use_undercanopy_stability, cfg.light_inhibit, cfg.itmax_canopy_fluxes};
    if (bounds_are_filter && filter) {
        if (n > 0)
            hipLaunchKernelGGL(k_canopy_init, dim3((n + 255) / 256), dim3(256), 0, st, n, filter, L.begp_alloc, can);
    } else if (L.ldp > 0) {
        hipLaunchKernelGGL(k_canopy_init, dim3((L.ldp + 255) / 256), dim3(256), 0, st, L.ldp, nullptr, 0, can);
    }
    if (n <= 0) return;
    if (zero_cgrnd)
        hipLaunchKernelGGL(k_zero_cgrnd, dim3((n + 255) / 256), dim3(256), 0, st, n, filter, L.begp_alloc, can);
    const int R = work ? canopy_later_rounds(cfg.itmax_canopy_fluxes, first_its, round_its) : 0;
    const int blocks = (n + 63) / 64;
    RoundArgs a{V, sw, prm, pft_dev, col, can, n, filter, dtime, err, 0, 0, nullptr, nullptr, nullptr, nullptr,
                CanopyWork{}};
    // round 0 also runs the part of the routine before the iteration
    const GptlRange iter_range("can_iter");
    // the PHS form: one instance, built for one wave per SIMD
    RoundArgsPhs b{};
    if (hydrstress) b.phs = PhsArgs{*phs_q, *phs_h, phs_table, cfg.nlevsoi, cfg.modifyphoto_and_lmr_forcrop};
    auto launch = [&](bool later, int grid) {
        if (hydrstress) {
            b.r = a;
            if (later) hipLaunchKernelGGL((k_canopy_round<1, true, true>), dim3(grid), dim3(64), 0, st, b);
            else hipLaunchKernelGGL((k_canopy_round<1, false, true>), dim3(grid), dim3(64), 0, st, b);
        } else if (later) {
            hipLaunchKernelGGL(later_kernel, dim3(grid), dim3(64), 0, st, a);
        } else {
            hipLaunchKernelGGL(first_kernel, dim3(grid), dim3(64), 0, st, a);
        }
    };
    if (R == 0) {
        // one pass: every patch finishes (converged or itmax reached) in round 0
        a.budget = cfg.itmax_canopy_fluxes + 1;
        launch(false, blocks);
        return;
    }
    double *w = (double *)work;
    auto field = [&](int k) { return w + (size_t)k * n; };
    a.W = CanopyWork{field(0), field(1), field(2),  field(3),  field(4),  field(5),  field(6),  field(7),
                     field(8), field(9), field(10), field(11), field(12), field(13), field(14), field(15)};
    int32_t *list[2] = {(int32_t *)(w + (size_t)N_WORK_FIELDS * n), nullptr};
    list[1] = list[0] + n;
    int32_t *cnt = list[1] + n;
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, st, cnt, R + 1);
    // later rounds stride over the patches still iterating with a grid that
    // keeps every wave slot of the chip busy
    const int later_blocks = blocks < 8192 ? blocks : 8192;
    for (int r = 0; r <= R; ++r) {
        a.round = r;
        a.budget = r == 0 ? first_its : round_its;
        a.list_in = list[(r + 1) & 1];
        a.cnt_in = r ? cnt + r - 1 : nullptr;
        a.list_out = list[r & 1];
        a.cnt_out = cnt + r;
        launch(r > 0, r == 0 ? blocks : later_blocks);
    }
}

void launch_root_sink(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_col_state &col,
                      const ctsm_canopy_state &can, int n, const int32_t *filter, ErrWord *err)
{
    const GptlRange range("hydro_without_drainage");
    using namespace canopy;
    if (n <= 0) return;
    CView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, L.begg_alloc, L.ldg, cfg.nlevsno, cfg.nlevsoi, cfg.nlevgrnd};
    if (filter)
        hipLaunchKernelGGL(k_root_sink, dim3((n + 255) / 256), dim3(256), 0, st, V, col, can, n, filter, err);
    else
        hipLaunchKernelGGL(k_root_sink_staged, dim3((n + RS_COLS - 1) / RS_COLS), dim3(RS_COLS * RS_GROUPS), 0, st, V,
                           col, can, n, err);
}

}  // namespace ctsm
