/*
 * canopy_oracle.c -- CPU restatement of CTSM's CanopyFluxes path (test
 * infrastructure only: tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline call it as the checker; the library never links it).
 *
 * Follows, loop by loop and in the reference's operation order:
 *   oracle_canopyfluxes      CanopyFluxesMod.F90:191-1765 (nlevcan = 1, no FATES / CN /
 *                            plant hydraulics / LUNA / C13 / methane / human indices;
 *                            ozone coefficients are inputs)
 *     frictionvelocity       FrictionVelocityMod.F90:754-1117, StabilityFunc1/2 :1120-1155
 *     moninobukini           FrictionVelocityMod.F90:1158-1209
 *     qsat                   QSatMod.F90:61-127
 *     photosynthesis         PhotosynthesisMod.F90:1243-2062 (sun and shade phases)
 *     hybrid / brent         PhotosynthesisMod.F90:2251-2485
 *     ci_func                PhotosynthesisMod.F90:2564-2701
 *     quadratic              quadraticMod.F90:17-74
 *     root moisture stress   SoilMoistStressMod.F90:70-215, 312-514
 *                            (calc_effective_soilporosity, calc_volumetric_h2oliq,
 *                            calc_root_moist_stress_clm45default)
 *     PhotosynthesisTotal    PhotosynthesisMod.F90:2065-2151
 *     truncate_small_values  NumericsMod.F90:60-100
 *   oracle_root_sink         clm_driver.F90:1714-1716 (p2c_1d_filter, subgridAveMod.F90)
 *                            + SoilWaterPlantSinkMod.F90:332-424
 *
 * Elementary functions: the shared deterministic ctsm_math.h (cm_exp, cm_log,
 * cm_pow for real exponents, cm_atan); sqrt is IEEE sqrt.  Integer powers are
 * products (x**4 = (x*x)*(x*x), x**3 = (x*x)*x), x**2.0 = x*x and
 * x**(-1.0) = 1/x as the compiler folds them.
 *
 * Pins (tests/test_oracle_canopy.py): quadratic against the reference's own
 * known-answer tests (src/utils/test/quadratic_test/test_quadratic.pf); QSat
 * against the Flatau et al. (1992) polynomial values and a finite-difference
 * derivative; FrictionVelocity against the neutral log-law closed form and
 * the Businger-Dyer forms; brent/hybrid roots satisfy ci_func(ci) ~ 0 and the
 * Ball-Berry / Medlyn conductance identity the reference itself checks
 * (PhotosynthesisMod.F90:2018-2027); the converged leaf temperature closes the
 * canopy energy balance (err <= 0.1 W/m2, CanopyFluxesMod.F90:1742-1756);
 * root stress against hand-computed CH78 values.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ctsm_gpu.h"
#include "ctsm_math.h"
#include "oracle_libm.h"   /* -DORACLE_LIBM: the host libm instead */
#include "phs_oracle.h"

/* shr_const_mod (tools/mksurfdata_esmf/src/shr_const_mod.F90) and clm_varcon.F90 */
#define C_TFRZ     273.15
#define C_DENH2O   1.000e3
#define C_DENICE   0.917e3
#define C_GRAV     9.80616
#define C_SB       5.67e-8
#define C_VKC      0.4
#define C_CPAIR    1.00464e3
#define C_HVAP     2.501e6
#define C_CWATER   4.188e3
#define C_RGAS     (6.02214e26 * 1.38065e-23)
#define C_PI       3.14159265358979323846
#define C_SPVAL    1.e36
#define C_TLSAI_CRIT 2.0
#define C_ALPHA_AERO 1.0
#define C_NU       1.5e-5
#define C_CD1      7.5
#define C_CDRYBIO  1400.0
#define C_C_TO_B   2.0

#define ORC_OK       0
#define ORC_ENDRUN   6

/* addressing */
typedef struct {
    int begc, ldc, begp, ldp, begg, nlevsno, nlevgrnd;
} cl_t;
#define PA(a, p)      ((a)[(p) - L.begp])
#define CA(a, c)      ((a)[(c) - L.begc])
#define GA(a, g)      ((a)[(g) - L.begg])
#define SOIP(a, p, j) ((a)[((p) - L.begp) + (long)((j) - 1) * L.ldp])
#define SOIC(a, c, j) ((a)[((c) - L.begc) + (long)((j) - 1) * L.ldc])
#define SNOC(a, c, j) ((a)[((c) - L.begc) + (long)((j) + L.nlevsno - 1) * L.ldc])

static double fmax90(double a, double b) { return (b > a) ? b : a; }
static double fmin90(double a, double b) { return (b < a) ? b : a; }
static double p4(double x) { const double x2 = x * x; return x2 * x2; }
static double p3(double x) { return (x * x) * x; }

/* ---------------------------------------------------------------------- */
/* QSatMod.F90:61-127 */
void oracle_qsat(double T, double p, double *qs, double *es, double *qsdT)
{
    static const double a[9] = {6.11213476, 0.444007856, 0.143064234e-01, 0.264461437e-03, 0.305903558e-05,
                                0.196237241e-07, 0.892344772e-10, -0.373208410e-12, 0.209339997e-15};
    static const double b[9] = {0.444017302, 0.286064092e-01, 0.794683137e-03, 0.121211669e-04, 0.103354611e-06,
                                0.404125005e-09, -0.788037859e-12, -0.114596802e-13, 0.381294516e-16};
    static const double c[9] = {6.11123516, 0.503109514, 0.188369801e-01, 0.420547422e-03, 0.614396778e-05,
                                0.602780717e-07, 0.387940929e-09, 0.149436277e-11, 0.262655803e-14};
    static const double d[9] = {0.503277922, 0.377289173e-01, 0.126801703e-02, 0.249468427e-04, 0.313703411e-06,
                                0.257180651e-08, 0.133268878e-10, 0.394116744e-13, 0.498070196e-16};
    const double td = fmin90(100.0, fmax90(-75.0, T - C_TFRZ));
    const double *k = (td >= 0.0) ? a : c;
    double es_l = k[0] + td * (k[1] + td * (k[2] + td * (k[3] + td * (k[4]
                  + td * (k[5] + td * (k[6] + td * (k[7] + td * k[8])))))));
    es_l = es_l * 100.0;
    const double vp = 1.0 / (p - 0.378 * es_l);
    const double vp1 = 0.622 * vp;
    *qs = es_l * vp1;
    if (es) *es = es_l;
    if (qsdT) {
        const double *m = (td >= 0.0) ? b : d;
        double esdT = m[0] + td * (m[1] + td * (m[2] + td * (m[3] + td * (m[4]
                      + td * (m[5] + td * (m[6] + td * (m[7] + td * m[8])))))));
        esdT = esdT * 100.0;
        const double vp2 = vp1 * vp;
        *qsdT = esdT * vp2 * p;
    }
}

/* quadraticMod.F90:17-74.  Returns 0, or ORC_ENDRUN where the reference calls endrun. */
int oracle_quadratic(double a, double b, double c, double *r1, double *r2)
{
    if (a == 0.0) return ORC_ENDRUN;
    double root = b * b - 4.0 * a * c;
    if (root < 0.0) {
        if (-root < 3.0 * 2.220446049250313e-16) root = 0.0;
        else return ORC_ENDRUN;
    }
    double q;
    if (b >= 0.0) q = -0.5 * (b + sqrt(root));
    else q = -0.5 * (b - sqrt(root));
    *r1 = q / a;
    if (q != 0.0) *r2 = c / q;
    else *r2 = 1.e36;
    return ORC_OK;
}

/* FrictionVelocityMod.F90:1120-1155 */
static double stab1(double zeta)
{
    const double chik2 = sqrt(1.0 - 16.0 * zeta);
    const double chik = sqrt(chik2);
    return 2.0 * cm_log((1.0 + chik) * 0.5) + cm_log((1.0 + chik2) * 0.5) - 2.0 * cm_atan(chik) + C_PI * 0.5;
}
static double stab2(double zeta)
{
    const double chik2 = sqrt(1.0 - 16.0 * zeta);
    return 2.0 * cm_log((1.0 + chik2) * 0.5);
}

/* FrictionVelocityMod.F90:1158-1209 */
void oracle_moninobukini(double ur, double thv, double dthv, double zldis, double z0m, double zetamax,
                         double *um, double *obu)
{
    const double ustar_unused = 0.06;
    const double wc = 0.5;
    (void)ustar_unused;
    if (dthv >= 0.0) *um = fmax90(ur, 0.1);
    else *um = sqrt(ur * ur + wc * wc);
    const double rib = C_GRAV * zldis * dthv / (thv * *um * *um);
    double zeta;
    if (rib >= 0.0) {
        zeta = rib * cm_log(zldis / z0m) / (1.0 - 5.0 * fmin90(rib, 0.19));
        zeta = fmin90(zetamax, fmax90(zeta, 0.01));
    } else {
        zeta = rib * cm_log(zldis / z0m);
        zeta = fmax90(-100.0, fmin90(zeta, -0.01));
    }
    *obu = zldis / zeta;
}

/* temperature / humidity profile term (one of the four regimes), the
 * repeated block of FrictionVelocityMod.F90:916-1060 */
static double fv_temp_profile(double zldis, double z0, double obu)
{
    const double zetat = 0.465;
    const double zeta = zldis / obu;
    if (zeta < -zetat) {
        return C_VKC / (cm_log(-zetat * obu / z0) - stab2(-zetat) + stab2(z0 / obu)
                        + 0.8 * (cm_pow(zetat, -0.333) - cm_pow(-zeta, -0.333)));
    } else if (zeta < 0.0) {
        return C_VKC / (cm_log(zldis / z0) - stab2(zeta) + stab2(z0 / obu));
    } else if (zeta <= 1.0) {
        return C_VKC / (cm_log(zldis / z0) + 5.0 * zeta - 5.0 * z0 / obu);
    } else {
        return C_VKC / (cm_log(obu / z0) + 5.0 - 5.0 * z0 / obu + (5.0 * cm_log(zeta) + zeta - 1.0));
    }
}

typedef struct {
    double ustar, temp1, temp2, temp12m, temp22m, fm, u10, u10_clm, va, fv, vds;
} fv_out;

/* FrictionVelocityMod.F90:754-1117 (patch form).  fm is the running mean. */
void oracle_frictionvelocity(double forc_hgt_u_patch, double forc_hgt_t_patch, double forc_hgt_q_patch,
                             double displa, double z0m, double z0h, double z0q, double obu, int iter,
                             double ur, double um, double fm_in, double *out11)
{
    const double zetam = 1.574;
    fv_out o;
    double zldis = forc_hgt_u_patch - displa;
    double zeta = zldis / obu;
    if (zeta < -zetam) {
        o.ustar = C_VKC * um / (cm_log(-zetam * obu / z0m) - stab1(-zetam) + stab1(z0m / obu)
                                + 1.14 * (cm_pow(-zeta, 0.333) - cm_pow(zetam, 0.333)));
    } else if (zeta < 0.0) {
        o.ustar = C_VKC * um / (cm_log(zldis / z0m) - stab1(zeta) + stab1(z0m / obu));
    } else if (zeta <= 1.0) {
        o.ustar = C_VKC * um / (cm_log(zldis / z0m) + 5.0 * zeta - 5.0 * z0m / obu);
    } else {
        o.ustar = C_VKC * um / (cm_log(obu / z0m) + 5.0 - 5.0 * z0m / obu + (5.0 * cm_log(zeta) + zeta - 1.0));
    }
    if (zeta < 0.0) o.vds = 2.e-3 * o.ustar * (1.0 + cm_pow(300.0 / (-obu), 0.666));
    else o.vds = 2.e-3 * o.ustar;

    if (zldis - z0m <= 10.0) {
        o.u10_clm = um;
    } else {
        const double z10 = 10.0 + z0m;
        if (zeta < -zetam) {
            o.u10_clm = um - (o.ustar / C_VKC * (cm_log(-zetam * obu / z10) - stab1(-zetam) + stab1(z10 / obu)
                                                  + 1.14 * (cm_pow(-zeta, 0.333) - cm_pow(zetam, 0.333))));
        } else if (zeta < 0.0) {
            o.u10_clm = um - (o.ustar / C_VKC * (cm_log(zldis / z10) - stab1(zeta) + stab1(z10 / obu)));
        } else if (zeta <= 1.0) {
            o.u10_clm = um - (o.ustar / C_VKC * (cm_log(zldis / z10) + 5.0 * zeta - 5.0 * z10 / obu));
        } else {
            o.u10_clm = um - (o.ustar / C_VKC * (cm_log(obu / z10) + 5.0 - 5.0 * z10 / obu
                                                  + (5.0 * cm_log(zeta) + zeta - 1.0)));
        }
    }
    o.va = um;

    /* temperature profile */
    zldis = forc_hgt_t_patch - displa;
    o.temp1 = fv_temp_profile(zldis, z0h, obu);
    /* humidity profile */
    if (forc_hgt_q_patch == forc_hgt_t_patch && z0q == z0h) {
        o.temp2 = o.temp1;
    } else {
        zldis = forc_hgt_q_patch - displa;
        o.temp2 = fv_temp_profile(zldis, z0q, obu);
    }
    /* 2-m profiles */
    zldis = 2.0 + z0h;
    o.temp12m = fv_temp_profile(zldis, z0h, obu);
    if (z0q == z0h) {
        o.temp22m = o.temp12m;
    } else {
        zldis = 2.0 + z0q;
        o.temp22m = fv_temp_profile(zldis, z0q, obu);
    }
    /* 10-m wind for the dust model */
    zldis = forc_hgt_u_patch - displa;
    zeta = zldis / obu;
    double fmnew;
    if (fmin90(zeta, 1.0) < 0.0) {
        const double tmp1 = cm_pow(1.0 - 16.0 * fmin90(zeta, 1.0), 0.25);
        const double tmp2 = cm_log((1.0 + tmp1 * tmp1) / 2.0);
        const double tmp3 = cm_log((1.0 + tmp1) / 2.0);
        fmnew = 2.0 * tmp3 + tmp2 - 2.0 * cm_atan(tmp1) + 1.5707963;
    } else {
        fmnew = -5.0 * fmin90(zeta, 1.0);
    }
    if (iter == 1) o.fm = fmnew;
    else o.fm = 0.5 * (fm_in + fmnew);
    double zeta10 = fmin90(10.0 / obu, 1.0);
    if (zeta == 0.0) zeta10 = 0.0;
    double fm10;
    if (zeta10 < 0.0) {
        const double tmp1 = cm_pow(1.0 - 16.0 * zeta10, 0.25);
        const double tmp2 = cm_log((1.0 + tmp1 * tmp1) / 2.0);
        const double tmp3 = cm_log((1.0 + tmp1) / 2.0);
        fm10 = 2.0 * tmp3 + tmp2 - 2.0 * cm_atan(tmp1) + 1.5707963;
    } else {
        fm10 = -5.0 * zeta10;
    }
    const double tmp4 = cm_log(fmax90(1.0, forc_hgt_u_patch / 10.0));
    o.u10 = ur - o.ustar / C_VKC * (tmp4 - o.fm + fm10);
    o.fv = o.ustar;
    memcpy(out11, &o, sizeof o);
}

/* ---------------------------------------------------------------------- */
/* Photosynthesis (non-PHS), one patch, one phase; nlevcan = 1. */
typedef struct {
    /* patch constants of the call */
    int c3;                /* c3flag */
    double forc_pbot, cair, oair, gb_mol, je, lmr_z, par_z, rh_can;
    double vcmax_z, tpu_z, kp_z, cp, kc, ko, qe, bbb, mbb, theta_cj, theta_ip;
    double medlynslope, medlynintercept;
    int stomatal;
    /* state written by ci_func (photosyns_inst arrays) */
    double ac, aj, ap, ag, an, gs_mol;
    int err;
} psn_t;

#define MAX_CS 1.e-06

/* PhotosynthesisMod.F90:2564-2701 */
static void ci_func(double ci, double *fval, psn_t *s)
{
    double r1, r2, ai;
    if (s->c3) {
        s->ac = s->vcmax_z * fmax90(ci - s->cp, 0.0) / (ci + s->kc * (1.0 + s->oair / s->ko));
        s->aj = s->je * fmax90(ci - s->cp, 0.0) / (4.0 * ci + 8.0 * s->cp);
        s->ap = 3.0 * s->tpu_z;
    } else {
        s->ac = s->vcmax_z;
        s->aj = s->qe * s->par_z * 4.6;
        s->ap = s->kp_z * fmax90(ci, 0.0) / s->forc_pbot;
    }
    if (oracle_quadratic(s->theta_cj, -(s->ac + s->aj), s->ac * s->aj, &r1, &r2)) s->err = ORC_ENDRUN;
    ai = fmin90(r1, r2);
    if (oracle_quadratic(s->theta_ip, -(ai + s->ap), ai * s->ap, &r1, &r2)) s->err = ORC_ENDRUN;
    s->ag = fmax90(0.0, fmin90(r1, r2));
    s->an = s->ag - s->lmr_z;
    if (s->an < 0.0) {
        *fval = 0.0;
        return;
    }
    double cs = s->cair - 1.4 / s->gb_mol * s->an * s->forc_pbot;
    cs = fmax90(cs, MAX_CS);
    if (s->stomatal == CTSM_STOMATA_MEDLYN2011) {
        const double term = 1.6 * s->an / (cs / s->forc_pbot * 1.e06);
        const double aq = 1.0;
        const double bq = -(2.0 * (s->medlynintercept * 1.e-06 + term)
                            + (s->medlynslope * term) * (s->medlynslope * term) / (s->gb_mol * 1.e-06 * s->rh_can));
        const double cq = s->medlynintercept * s->medlynintercept * 1.e-12
                          + (2.0 * s->medlynintercept * 1.e-06
                             + term * (1.0 - s->medlynslope * s->medlynslope / s->rh_can)) * term;
        if (oracle_quadratic(aq, bq, cq, &r1, &r2)) s->err = ORC_ENDRUN;
        s->gs_mol = fmax90(r1, r2) * 1.e06;
    } else {
        const double aq = cs;
        const double bq = cs * (s->gb_mol - s->bbb) - s->mbb * s->an * s->forc_pbot;
        const double cq = -s->gb_mol * (cs * s->bbb + s->mbb * s->an * s->forc_pbot * s->rh_can);
        if (oracle_quadratic(aq, bq, cq, &r1, &r2)) s->err = ORC_ENDRUN;
        s->gs_mol = fmax90(r1, r2);
    }
    *fval = ci - s->cair + s->an * s->forc_pbot * (1.4 * s->gs_mol + 1.6 * s->gb_mol) / (s->gb_mol * s->gs_mol);
}

/* PhotosynthesisMod.F90:2371-2485 */
static void brent(double *x, double x1, double x2, double f1, double f2, double tol, psn_t *s)
{
    const int itmax = 20;
    const double eps = 1.e-2;
    double a = x1, b = x2, fa = f1, fb = f2, c, fc, d = 0.0, e = 0.0, p, q, r, ss, tol1, xm;
    int iter;
    if ((fa > 0.0 && fb > 0.0) || (fa < 0.0 && fb < 0.0)) {
        s->err = ORC_ENDRUN;
        *x = b;
        return;
    }
    c = b;
    fc = fb;
    iter = 0;
    for (;;) {
        if (iter == itmax) break;
        iter = iter + 1;
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
        tol1 = 2.0 * eps * fabs(b) + 0.5 * tol;
        xm = 0.5 * (c - b);
        if (fabs(xm) <= tol1 || fb == 0.0) {
            *x = b;
            return;
        }
        if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
            ss = fb / fa;
            if (a == c) {
                p = 2.0 * xm * ss;
                q = 1.0 - ss;
            } else {
                q = fa / fc;
                r = fb / fc;
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
        ci_func(b, &fb, s);
        if (fb == 0.0) break;
    }
    *x = b;
}

/* PhotosynthesisMod.F90:2251-2368 */
static void hybrid(double *x0p, psn_t *s)
{
    const double eps = 1.e-2, eps1 = 1.e-4;
    const int itmax = 40;
    double x0 = *x0p, x1, f0, f1, x, dx, tol, minx, minf;
    int iter;
    ci_func(x0, &f0, s);
    if (f0 == 0.0) return;
    minx = x0;
    minf = f0;
    x1 = x0 * 0.99;
    ci_func(x1, &f1, s);
    if (f1 == 0.0) {
        *x0p = x1;
        return;
    }
    if (f1 < minf) {
        minx = x1;
        minf = f1;
    }
    iter = 0;
    for (;;) {
        iter = iter + 1;
        dx = -f1 * (x1 - x0) / (f1 - f0);
        x = x1 + dx;
        tol = fabs(x) * eps;
        if (fabs(dx) < tol) {
            x0 = x;
            break;
        }
        x0 = x1;
        f0 = f1;
        x1 = x;
        ci_func(x1, &f1, s);
        if (f1 < minf) {
            minx = x1;
            minf = f1;
        }
        if (fabs(f1) <= eps1) {
            x0 = x1;
            break;
        }
        if (f1 * f0 < 0.0) {
            brent(&x, x0, x1, f0, f1, tol, s);
            x0 = x;
            break;
        }
        if (iter > itmax) {
            ci_func(minx, &f1, s);
            break;
        }
    }
    *x0p = x0;
}

/* ---------------------------------------------------------------------- */
typedef struct {
    cl_t L;
    const ctsm_config *cfg;
    const ctsm_canopy_params *prm;
    const ctsm_pft_params *pft;
    const ctsm_col_state *col;
    const ctsm_canopy_state *cs;
    /* use_hydrstress: the PHS arrays and parameters (NULL otherwise) */
    const ctsm_phs_state *phs;
    const ctsm_phs_canopy_state *phc;
    const ctsm_phs_params *pp;
    int nlevsoi;
} ctx_t;

static double ft(double tl, double ha)
{
    return cm_exp(ha / (C_RGAS * 1.e-3 * (C_TFRZ + 25.0)) * (1.0 - (C_TFRZ + 25.0) / tl));
}
static double fth(double tl, double hd, double se, double scaleFactor)
{
    return scaleFactor / (1.0 + cm_exp((-hd + se * tl) / (C_RGAS * 1.e-3 * tl)));
}
static double fth25(double hd, double se)
{
    return 1.0 + cm_exp((-hd + se * (C_TFRZ + 25.0)) / (C_RGAS * 1.e-3 * (C_TFRZ + 25.0)));
}

/* PhotosynthesisMod.F90:1243-2062 for one patch and one phase (sun = 1):
 * the patch's share of the four filter loops (parameters, canopy-layer
 * scaling, leaf-level solve, canopy sums).  Returns ORC_ENDRUN where the
 * reference calls endrun. */
static int photosynthesis(const ctx_t *X, int p, int sun, double esat_tv, double eair, double oair,
                          double cair, double rb, double btran, double dayl_factor)
{
    const cl_t L = X->L;
    const ctsm_canopy_state *S = X->cs;
    const ctsm_canopy_params *P = X->prm;
    const ctsm_pft_params *F = X->pft;
    const int stomatal = X->cfg->stomatalcond_method;
    const int c = PA(S->column, p);
    const int g = PA(S->gridcell, p);
    const int ivt = PA(S->itype, p);
    const double forc_pbot = CA(S->forc_pbot, c);
    const double t_veg = PA(S->t_veg, p);
    const double t10 = PA(S->t_a10, p);
    const double tgcm = PA(S->thm, p);
    const int nrad = PA(S->nrad, p);
    const double par_z = sun ? PA(S->parsun_z, p) : PA(S->parsha_z, p);
    const double lai_z = sun ? PA(S->laisun_z, p) : PA(S->laisha_z, p);
    const double vcmaxcint = sun ? PA(S->vcmaxcintsun, p) : PA(S->vcmaxcintsha, p);
    const double o3coefv = sun ? PA(S->o3coefvsun, p) : PA(S->o3coefvsha, p);
    const double o3coefg = sun ? PA(S->o3coefgsun, p) : PA(S->o3coefgsha, p);
    double *ci_z = sun ? S->cisun_z : S->cisha_z;
    double *rs = sun ? S->rssun : S->rssha;
    double *rs_z = sun ? S->rssun_z : S->rssha_z;
    double *lmr = sun ? S->lmrsun : S->lmrsha;
    double *lmr_z = sun ? S->lmrsun_z : S->lmrsha_z;
    double *psn = sun ? S->psnsun : S->psnsha;
    double *psn_z = sun ? S->psnsun_z : S->psnsha_z;
    double *psn_wc = sun ? S->psnsun_wc : S->psnsha_wc;
    double *psn_wj = sun ? S->psnsun_wj : S->psnsha_wj;
    double *psn_wp = sun ? S->psnsun_wp : S->psnsha_wp;
    double *gs_mol_ph = sun ? S->gs_mol_sun : S->gs_mol_sha;
    double *gs_mol_ln = sun ? S->gs_mol_sun_ln : S->gs_mol_sha_ln;
    int rc = ORC_OK;

    /* loop 1: C3/C4 parameters and kinetics (:1525-1567) */
    const double lmrc = fth25(P->lmrhd, P->lmrse);
    const int c3 = ((int)lrint(F->c3psn[ivt]) == 1);
    PA(S->c3flag, p) = c3;
    double bbbopt = 0.0;
    if (c3) {
        PA(S->qe, p) = 0.0;
        if (stomatal == CTSM_STOMATA_BB1987) bbbopt = 10000.0;
    } else {
        PA(S->qe, p) = 0.05;
        if (stomatal == CTSM_STOMATA_BB1987) bbbopt = 40000.0;
    }
    if (stomatal == CTSM_STOMATA_BB1987) {
        PA(S->bbb, p) = fmax90(bbbopt * btran, 1.0);
        PA(S->mbb, p) = F->mbbopt[ivt];
    }
    const double kc25 = P->kc25_coef * forc_pbot;
    const double ko25 = P->ko25_coef * forc_pbot;
    const double sco = 0.5 * 0.209 / P->cp25_yr2000;
    const double cp25 = 0.5 * oair / sco;
    PA(S->kc, p) = kc25 * ft(t_veg, P->kcha);
    PA(S->ko, p) = ko25 * ft(t_veg, P->koha);
    PA(S->cp, p) = cp25 * ft(t_veg, P->cpha);

    /* loop 2: canopy-top capacities and the layer loop (:1572-1846) */
    if (F->slatop[ivt] * F->leafcn[ivt] <= 0.0) return ORC_ENDRUN;
    PA(S->lnca, p) = 1.0 / (F->slatop[ivt] * F->leafcn[ivt]);
    const double lnc = PA(S->lnca, p);
    double vcmax25top = lnc * F->flnr[ivt] * P->fnr * P->act25 * dayl_factor;
    vcmax25top = vcmax25top * F->fnitr[ivt];
    const double t10c = fmin90(fmax90(t10 - C_TFRZ, 11.0), 35.0);
    const double jmax25top = ((2.59 - 0.035 * t10c) * vcmax25top) * P->jmax25top_sf;
    const double tpu25top = P->tpu25ratio * vcmax25top;
    const double kp25top = P->kp25ratio * vcmax25top;
    double lmr25top;
    if (c3) lmr25top = vcmax25top * P->leaf_mr_vcm;
    else lmr25top = vcmax25top * 0.025;
    double jmax_z = 0.0;
    if (nrad >= 1) {
        const double nscaler = vcmaxcint;
        const double lmr25 = lmr25top * nscaler;
        if (c3) {
            PA(lmr_z, p) = lmr25 * ft(t_veg, P->lmrha) * fth(t_veg, P->lmrhd, P->lmrse, lmrc);
        } else {
            PA(lmr_z, p) = lmr25 * cm_pow(2.0, (t_veg - (C_TFRZ + 25.0)) / 10.0);
            PA(lmr_z, p) = PA(lmr_z, p) / (1.0 + cm_exp(1.3 * (t_veg - (C_TFRZ + 55.0))));
        }
        if (par_z <= 0.0) {
            PA(S->vcmax_z, p) = 0.0;
            jmax_z = 0.0;
            PA(S->tpu_z, p) = 0.0;
            PA(S->kp_z, p) = 0.0;
        } else {
            const double vcmax25 = vcmax25top * nscaler;
            const double jmax25 = jmax25top * nscaler;
            const double tpu25 = tpu25top * nscaler;
            const double kp25 = kp25top * nscaler;
            const double vcmaxse = (668.39 - 1.07 * t10c) * P->vcmaxse_sf;
            const double jmaxse = (659.70 - 0.75 * t10c) * P->jmaxse_sf;
            const double tpuse = (668.39 - 1.07 * t10c) * P->tpuse_sf;
            const double vcmaxc = fth25(P->vcmaxhd, vcmaxse);
            const double jmaxc = fth25(P->jmaxhd, jmaxse);
            const double tpuc = fth25(P->tpuhd, tpuse);
            PA(S->vcmax_z, p) = vcmax25 * ft(t_veg, P->vcmaxha) * fth(t_veg, P->vcmaxhd, vcmaxse, vcmaxc);
            jmax_z = jmax25 * ft(t_veg, P->jmaxha) * fth(t_veg, P->jmaxhd, jmaxse, jmaxc);
            PA(S->tpu_z, p) = tpu25 * ft(t_veg, P->tpuha) * fth(t_veg, P->tpuhd, tpuse, tpuc);
            if (!c3) {
                PA(S->vcmax_z, p) = vcmax25 * cm_pow(2.0, (t_veg - (C_TFRZ + 25.0)) / 10.0);
                PA(S->vcmax_z, p) = PA(S->vcmax_z, p) / (1.0 + cm_exp(0.2 * ((C_TFRZ + 15.0) - t_veg)));
                PA(S->vcmax_z, p) = PA(S->vcmax_z, p) / (1.0 + cm_exp(0.3 * (t_veg - (C_TFRZ + 40.0))));
            }
            PA(S->kp_z, p) = kp25 * cm_pow(2.0, (t_veg - (C_TFRZ + 25.0)) / 10.0);
        }
        PA(S->vcmax_z, p) = PA(S->vcmax_z, p) * btran;
        PA(lmr_z, p) = PA(lmr_z, p) * btran;
        /* light inhibition tests par_z(p,1) of the phase being computed */
        if (X->cfg->light_inhibit && par_z > 0.0) PA(lmr_z, p) = PA(lmr_z, p) * 0.67;
    }

    /* loop 3: leaf-level photosynthesis and conductance (:1852-2032) */
    const double rsmax0 = 2.e4;
    const double cf = forc_pbot / (C_RGAS * 1.e-3 * tgcm) * 1.e06;
    const double gb = 1.0 / rb;
    PA(S->gb_mol, p) = gb * cf;
    const double gb_mol = PA(S->gb_mol, p);
    double psn_wc_z = 0.0, psn_wj_z = 0.0, psn_wp_z = 0.0;
    if (nrad >= 1) {
        if (par_z <= 0.0) {
            PA(S->ac, p) = 0.0;
            PA(S->aj, p) = 0.0;
            PA(S->ap, p) = 0.0;
            PA(S->ag, p) = 0.0;
            PA(S->an, p) = PA(S->ag, p) - PA(lmr_z, p);
            PA(psn_z, p) = 0.0;
            psn_wc_z = 0.0;
            psn_wj_z = 0.0;
            psn_wp_z = 0.0;
            if (stomatal == CTSM_STOMATA_BB1987) PA(rs_z, p) = fmin90(rsmax0, 1.0 / PA(S->bbb, p) * cf);
            else PA(rs_z, p) = fmin90(rsmax0, 1.0 / F->medlynintercept[ivt] * cf);
            PA(ci_z, p) = 0.0;
            PA(S->rh_leaf, p) = 0.0;
            PA(gs_mol_ph, p) = cf / PA(rs_z, p);
        } else {
            const double ceair = fmin90(eair, esat_tv);
            double rh_can;
            if (stomatal == CTSM_STOMATA_BB1987) {
                rh_can = ceair / esat_tv;
            } else {
                rh_can = fmax90(esat_tv - ceair, 50.0) * 0.001;
                PA(S->vpd_can, p) = rh_can;
            }
            const double qabs = 0.5 * (1.0 - P->fnps) * par_z * 4.6;
            double r1, r2;
            if (oracle_quadratic(P->theta_psii, -(qabs + jmax_z), qabs * jmax_z, &r1, &r2)) rc = ORC_ENDRUN;
            const double je = fmin90(r1, r2);
            if (c3) PA(ci_z, p) = 0.7 * cair;
            else PA(ci_z, p) = 0.4 * cair;
            double ciold = PA(ci_z, p);
            psn_t s;
            s.c3 = c3;
            s.forc_pbot = forc_pbot;
            s.cair = cair;
            s.oair = oair;
            s.gb_mol = gb_mol;
            s.je = je;
            s.lmr_z = PA(lmr_z, p);
            s.par_z = par_z;
            s.rh_can = rh_can;
            s.vcmax_z = PA(S->vcmax_z, p);
            s.tpu_z = PA(S->tpu_z, p);
            s.kp_z = PA(S->kp_z, p);
            s.cp = PA(S->cp, p);
            s.kc = PA(S->kc, p);
            s.ko = PA(S->ko, p);
            s.qe = PA(S->qe, p);
            s.bbb = PA(S->bbb, p);
            s.mbb = PA(S->mbb, p);
            s.theta_cj = F->theta_cj[ivt];
            s.theta_ip = P->theta_ip;
            s.medlynslope = F->medlynslope[ivt];
            s.medlynintercept = F->medlynintercept[ivt];
            s.stomatal = stomatal;
            s.ac = PA(S->ac, p);
            s.aj = PA(S->aj, p);
            s.ap = PA(S->ap, p);
            s.ag = PA(S->ag, p);
            s.an = PA(S->an, p);
            s.gs_mol = PA(S->gs_mol, p);
            s.err = 0;
            hybrid(&ciold, &s);
            if (s.err) rc = ORC_ENDRUN;
            PA(S->ac, p) = s.ac;
            PA(S->aj, p) = s.aj;
            PA(S->ap, p) = s.ap;
            PA(S->ag, p) = s.ag;
            PA(S->an, p) = s.an;
            PA(S->gs_mol, p) = s.gs_mol;
            if (PA(S->an, p) < 0.0) {
                if (stomatal == CTSM_STOMATA_BB1987) PA(S->gs_mol, p) = PA(S->bbb, p);
                else PA(S->gs_mol, p) = F->medlynintercept[ivt];
            }
            const double gs_mol = PA(S->gs_mol, p);
            PA(gs_mol_ph, p) = gs_mol;
            if (GA(S->near_local_noon, g)) PA(gs_mol_ln, p) = gs_mol;
            else PA(gs_mol_ln, p) = C_SPVAL;
            const double an = PA(S->an, p);
            double cs = cair - 1.4 / gb_mol * an * forc_pbot;
            cs = fmax90(cs, MAX_CS);
            PA(ci_z, p) = cair - an * forc_pbot * (1.4 * gs_mol + 1.6 * gb_mol) / (gb_mol * gs_mol);
            PA(ci_z, p) = fmax90(PA(ci_z, p), 1.e-06);
            const double gs = gs_mol / cf;
            PA(rs_z, p) = fmin90(1.0 / gs, rsmax0);
            PA(rs_z, p) = PA(rs_z, p) / o3coefg;
            PA(psn_z, p) = PA(S->ag, p);
            PA(psn_z, p) = PA(psn_z, p) * o3coefv;
            psn_wc_z = 0.0;
            psn_wj_z = 0.0;
            psn_wp_z = 0.0;
            const double ac = PA(S->ac, p), aj = PA(S->aj, p), ap = PA(S->ap, p);
            if (ac <= aj && ac <= ap) psn_wc_z = PA(psn_z, p);
            else if (aj < ac && aj <= ap) psn_wj_z = PA(psn_z, p);
            else if (ap < ac && ap < aj) psn_wp_z = PA(psn_z, p);
            if (gs_mol < 0.0) rc = ORC_ENDRUN;
            if (stomatal == CTSM_STOMATA_BB1987) {
                const double hs = (gb_mol * ceair + gs_mol * esat_tv) / ((gb_mol + gs_mol) * esat_tv);
                PA(S->rh_leaf, p) = hs;
            }
        }
    }

    /* loop 4: canopy sums (:2040-2059) */
    double psncan = 0.0, psncan_wc = 0.0, psncan_wj = 0.0, psncan_wp = 0.0, lmrcan = 0.0, gscan = 0.0,
           laican = 0.0;
    if (nrad >= 1) {
        psncan = psncan + PA(psn_z, p) * lai_z;
        psncan_wc = psncan_wc + psn_wc_z * lai_z;
        psncan_wj = psncan_wj + psn_wj_z * lai_z;
        psncan_wp = psncan_wp + psn_wp_z * lai_z;
        lmrcan = lmrcan + PA(lmr_z, p) * lai_z;
        gscan = gscan + lai_z / (rb + PA(rs_z, p));
        laican = laican + lai_z;
    }
    if (laican > 0.0) {
        PA(psn, p) = psncan / laican;
        PA(psn_wc, p) = psncan_wc / laican;
        PA(psn_wj, p) = psncan_wj / laican;
        PA(psn_wp, p) = psncan_wp / laican;
        PA(lmr, p) = lmrcan / laican;
        PA(rs, p) = laican / gscan - rb;
    } else {
        PA(psn, p) = 0.0;
        PA(psn_wc, p) = 0.0;
        PA(psn_wj, p) = 0.0;
        PA(psn_wp, p) = 0.0;
        PA(lmr, p) = 0.0;
        PA(rs, p) = 0.0;
    }
    return rc;
}

/* ---- plant hydraulic stress: PhotosynthesisHydraulicStress -------------- */

#define SUN 0
#define SHA 1
#define VEGA(a, p, k)  ((a)[((p) - L.begp) + (long)(k) * L.ldp])    /* k = 0 sun .. 3 root */
#define PHS2(a, p, k)  ((a)[((p) - L.begp) + (long)(k) * L.ldp])    /* k = 0 sun, 1 sha */

typedef struct {
    /* constants of the patch's call */
    int c3, stomatal;
    double forc_pbot, cair, oair, gb_mol, rh_can, theta_cj, theta_ip, qe, bbb, mbb, medlynslope, medlynintercept;
    double cp, kc, ko;
    double vcmax_z[2], tpu_z[2], kp_z[2], je[2], lmr_z[2], par_z[2];
    double qsatl, qaf;
    p_stress_t leaf;       /* calcstress's view of the patch */
    int morning;           /* get_local_time < isecspday / 2 */
    /* photosyns_inst values ci_func_PHS writes */
    double ac[2], aj[2], ap[2], ag[2], an[2];
    /* the arrays calcstress writes: its last call's values */
    int stressed;          /* a calcstress call happened */
    double pd[4];          /* vegwp_pd: x where night and morning, else spval */
    int tran_set;          /* a night call (or hybrid_PHS) set qflx_tran_veg */
    double tran;
    int err;
} phs_t;

static void phs_calcstress(phs_t *s, double *x, double *bsun, double *bsha, double gs_sun, double gs_sha)
{
    int night;
    double tran = 0.0;
    oracle_phs_calcstress_one(&s->leaf, x, gs_sun, gs_sha, s->qsatl, s->qaf, bsun, bsha, &night, &tran);
    s->stressed = 1;
    for (int k = 0; k < 4; ++k) s->pd[k] = (night && s->morning) ? x[k] : C_SPVAL;
    if (night) {
        s->tran_set = 1;
        s->tran = tran;
    }
}

/* ci_func_PHS (PhotosynthesisMod.F90:4227-4486) */
static void ci_func_phs(phs_t *s, double *x, double cisun, double cisha, double *fsun, double *fsha, double *bsun,
                        double *bsha, int bflag, double gs0sun, double gs0sha, double *gs_sun, double *gs_sha)
{
    double r1, r2, ai;
    if (bflag) phs_calcstress(s, x, bsun, bsha, gs0sun, gs0sha);
    const double ci[2] = {cisun, cisha};
    const double b[2] = {*bsun, *bsha};
    if (s->c3) {
        for (int k = 0; k < 2; ++k)
            s->ac[k] = b[k] * s->vcmax_z[k] * fmax90(ci[k] - s->cp, 0.0) / (ci[k] + s->kc * (1.0 + s->oair / s->ko));
        for (int k = 0; k < 2; ++k) s->aj[k] = s->je[k] * fmax90(ci[k] - s->cp, 0.0) / (4.0 * ci[k] + 8.0 * s->cp);
        for (int k = 0; k < 2; ++k) s->ap[k] = 3.0 * s->tpu_z[k];
    } else {
        for (int k = 0; k < 2; ++k) s->ac[k] = b[k] * s->vcmax_z[k];
        for (int k = 0; k < 2; ++k) s->aj[k] = s->qe * s->par_z[k] * 4.6;
        for (int k = 0; k < 2; ++k) s->ap[k] = s->kp_z[k] * fmax90(ci[k], 0.0) / s->forc_pbot;
    }
    for (int k = 0; k < 2; ++k) {
        if (oracle_quadratic(s->theta_cj, -(s->ac[k] + s->aj[k]), s->ac[k] * s->aj[k], &r1, &r2)) s->err = ORC_ENDRUN;
        ai = fmin90(r1, r2);
        if (oracle_quadratic(s->theta_ip, -(ai + s->ap[k]), ai * s->ap[k], &r1, &r2)) s->err = ORC_ENDRUN;
        s->ag[k] = fmax90(0.0, fmin90(r1, r2));
    }
    s->an[SUN] = s->ag[SUN] - *bsun * s->lmr_z[SUN];
    s->an[SHA] = s->ag[SHA] - *bsha * s->lmr_z[SHA];
    const double gsmin = (s->stomatal == CTSM_STOMATA_MEDLYN2011) ? s->medlynintercept : s->bbb;
    if (s->an[SUN] < 0.0) {
        *gs_sun = fmax90(*bsun * gsmin, 1.0);
        *fsun = 0.0;
    }
    if (s->an[SHA] < 0.0) {
        *gs_sha = fmax90(*bsha * gsmin, 1.0);
        *fsha = 0.0;
    }
    if (s->an[SUN] < 0.0 && s->an[SHA] < 0.0) return;
    double cs[2] = {0.0, 0.0};
    double *gs[2] = {gs_sun, gs_sha};
    for (int k = 0; k < 2; ++k) {
        if (s->an[k] < 0.0) continue;
        cs[k] = s->cair - 1.4 / s->gb_mol * s->an[k] * s->forc_pbot;
        cs[k] = fmax90(cs[k], MAX_CS);
        if (s->stomatal == CTSM_STOMATA_MEDLYN2011) {
            const double term = 1.6 * s->an[k] / (cs[k] / s->forc_pbot * 1.e06);
            const double aq = 1.0;
            const double bq = -(2.0 * (s->medlynintercept * 1.e-06 + term)
                                + (s->medlynslope * term) * (s->medlynslope * term) / (s->gb_mol * 1.e-06 * s->rh_can));
            const double cq = s->medlynintercept * s->medlynintercept * 1.e-12
                              + (2.0 * s->medlynintercept * 1.e-06
                                 + term * (1.0 - s->medlynslope * s->medlynslope / s->rh_can)) * term;
            if (oracle_quadratic(aq, bq, cq, &r1, &r2)) s->err = ORC_ENDRUN;
            *gs[k] = fmax90(r1, r2) * 1.e06;
        } else {
            const double bk = k == SUN ? *bsun : *bsha;
            const double aq = cs[k];
            const double bq = cs[k] * (s->gb_mol - fmax90(bk * s->bbb, 1.0)) - s->mbb * s->an[k] * s->forc_pbot;
            const double cq = -s->gb_mol * (cs[k] * fmax90(bk * s->bbb, 1.0)
                                            + s->mbb * s->an[k] * s->forc_pbot * s->rh_can);
            if (oracle_quadratic(aq, bq, cq, &r1, &r2)) s->err = ORC_ENDRUN;
            *gs[k] = fmax90(r1, r2);
        }
    }
    double *fv[2] = {fsun, fsha};
    for (int k = 0; k < 2; ++k) {
        if (s->an[k] < 0.0) continue;
        if (*gs[k] > 0.0)
            *fv[k] = ci[k] - s->cair + s->an[k] * s->forc_pbot * (1.4 * *gs[k] + 1.6 * s->gb_mol) / (s->gb_mol * *gs[k]);
        else
            *fv[k] = ci[k] - s->cair;
    }
}

/* brent_PHS (:4068-4223): both phases' brackets shrink together; tol is the
 * sunlit tolerance for both */
static void brent_phs(phs_t *s, double *xsun, double x1sun, double x2sun, double f1sun, double f2sun, double *xsha,
                      double x1sha, double x2sha, double f1sha, double f2sha, double tol, double *gs_sun,
                      double *gs_sha, double *bsun, double *bsha)
{
    const int itmax = 20;
    const double eps = 1.e-4;
    double a[2] = {x1sun, x1sha}, b[2] = {x2sun, x2sha}, fa[2] = {f1sun, f1sha}, fb[2] = {f2sun, f2sha};
    double c[2], fc[2], d[2] = {0.0, 0.0}, e[2] = {0.0, 0.0}, tol1[2], xm[2];
    double dummy[4] = {0.0, 0.0, 0.0, 0.0};   /* x of the calls: calcstress is not called (bflag false) */
    for (int k = 0; k < 2; ++k)
        if ((fa[k] > 0.0 && fb[k] > 0.0) || (fa[k] < 0.0 && fb[k] < 0.0)) {
            s->err = ORC_ENDRUN;
            *xsun = b[SUN];
            *xsha = b[SHA];
            return;
        }
    c[0] = b[0]; c[1] = b[1];
    fc[0] = fb[0]; fc[1] = fb[1];
    int iter = 0;
    for (;;) {
        if (iter == itmax) break;
        iter = iter + 1;
        for (int k = 0; k < 2; ++k) {
            if ((fb[k] > 0.0 && fc[k] > 0.0) || (fb[k] < 0.0 && fc[k] < 0.0)) {
                c[k] = a[k];
                fc[k] = fa[k];
                d[k] = b[k] - a[k];
                e[k] = d[k];
            }
            if (fabs(fc[k]) < fabs(fb[k])) {
                a[k] = b[k];
                b[k] = c[k];
                c[k] = a[k];
                fa[k] = fb[k];
                fb[k] = fc[k];
                fc[k] = fa[k];
            }
        }
        for (int k = 0; k < 2; ++k) tol1[k] = 2.0 * eps * fabs(b[k]) + 0.5 * tol;
        for (int k = 0; k < 2; ++k) xm[k] = 0.5 * (c[k] - b[k]);
        if ((fabs(xm[SUN]) <= tol1[SUN] || fb[SUN] == 0.0) && (fabs(xm[SHA]) <= tol1[SHA] || fb[SHA] == 0.0)) {
            *xsun = b[SUN];
            *xsha = b[SHA];
            return;
        }
        for (int k = 0; k < 2; ++k) {
            if (fabs(e[k]) >= tol1[k] && fabs(fa[k]) > fabs(fb[k])) {
                const double ss = fb[k] / fa[k];
                double p, q, r;
                if (a[k] == c[k]) {
                    p = 2.0 * xm[k] * ss;
                    q = 1.0 - ss;
                } else {
                    q = fa[k] / fc[k];
                    r = fb[k] / fc[k];
                    p = ss * (2.0 * xm[k] * q * (q - r) - (b[k] - a[k]) * (r - 1.0));
                    q = (q - 1.0) * (r - 1.0) * (ss - 1.0);
                }
                if (p > 0.0) q = -q;
                p = fabs(p);
                if (2.0 * p < fmin90(3.0 * xm[k] * q - fabs(tol1[k] * q), fabs(e[k] * q))) {
                    e[k] = d[k];
                    d[k] = p / q;
                } else {
                    d[k] = xm[k];
                    e[k] = d[k];
                }
            } else {
                d[k] = xm[k];
                e[k] = d[k];
            }
            a[k] = b[k];
            fa[k] = fb[k];
            if (fabs(d[k]) > tol1[k]) b[k] = b[k] + d[k];
            else b[k] = b[k] + copysign(fabs(tol1[k]), xm[k]);
        }
        const double gs0sun = *gs_sun, gs0sha = *gs_sha;
        ci_func_phs(s, dummy, b[SUN], b[SHA], &fb[SUN], &fb[SHA], bsun, bsha, 0, gs0sun, gs0sha, gs_sun, gs_sha);
        if (fb[SUN] == 0.0 && fb[SHA] == 0.0) break;
    }
    *xsun = b[SUN];
    *xsha = b[SHA];
}

/* hybrid_PHS (:3815-4064): the outer loop updates bsun / bsha through
 * calcstress, the inner one is the secant / Brent search for both ci;
 * vegwp (4) inout; returns soilflux (clipped at 0) */
static double hybrid_phs(phs_t *s, double *x0sun, double *x0sha, double *vegwp, double *bsun, double *bsha,
                         double *gs_sun, double *gs_sha)
{
    const double toldb = 1.e-2, eps = 1.e-2, eps1 = 1.e-4;
    const int itmax = 3;
    double x[4];
    double x1sun = *x0sun, x1sha = *x0sha, f0sun = 0.0, f0sha = 0.0, f1sun = 0.0, f1sha = 0.0;
    double xsun, xsha, dxsun, dxsha, tolsun, tolsha, minf = 0.0, minxsun = 0.0, minxsha = 0.0;
    int bflag = 0;
    double b0sun = -1.0, b0sha = -1.0, gs0sun = 0.0, gs0sha = 0.0, dbsun, dbsha;
    *bsun = 1.0;
    *bsha = 1.0;
    int iter1 = 0;
    for (;;) {
        for (int k = 0; k < 4; ++k) x[k] = vegwp[k];
        iter1 = iter1 + 1;
        int iter2 = 0;
        *x0sun = fmax90(0.1, x1sun);
        x1sun = 0.99 * x1sun;
        *x0sha = fmax90(0.1, x1sha);
        x1sha = 0.99 * x1sha;
        tolsun = fabs(x1sun) * eps;
        tolsha = fabs(x1sha) * eps;
        ci_func_phs(s, x, *x0sun, *x0sha, &f0sun, &f0sha, bsun, bsha, bflag, gs0sun, gs0sha, gs_sun, gs_sha);
        dbsun = b0sun - *bsun;
        dbsha = b0sha - *bsha;
        b0sun = *bsun;
        b0sha = *bsha;
        bflag = 0;
        ci_func_phs(s, x, x1sun, x1sha, &f1sun, &f1sha, bsun, bsha, bflag, gs0sun, gs0sha, gs_sun, gs_sha);
        for (;;) {
            if (fabs(f0sun) < eps1 && fabs(f0sha) < eps1) {
                x1sun = *x0sun;
                x1sha = *x0sha;
                break;
            }
            if (fabs(f1sun) < eps1 && fabs(f1sha) < eps1) break;
            iter2 = iter2 + 1;
            if ((f1sun - f0sun) == 0.0) dxsun = 0.5 * (x1sun + *x0sun) - x1sun;
            else dxsun = -f1sun * (x1sun - *x0sun) / (f1sun - f0sun);
            if ((f1sha - f0sha) == 0.0) dxsha = 0.5 * (x1sha + *x0sha) - x1sha;
            else dxsha = -f1sha * (x1sha - *x0sha) / (f1sha - f0sha);
            *x0sun = x1sun;
            x1sun = x1sun + dxsun;
            *x0sha = x1sha;
            x1sha = x1sha + dxsha;
            ci_func_phs(s, x, x1sun, x1sha, &f1sun, &f1sha, bsun, bsha, bflag, gs0sun, gs0sha, gs_sun, gs_sha);
            if (fabs(dxsun) < tolsun && fabs(dxsha) < tolsha) {
                *x0sun = x1sun;
                *x0sha = x1sha;
                break;
            }
            if (iter2 == 1) {
                minf = fabs(f1sun + f1sha);
                minxsun = x1sun;
                minxsha = x1sha;
            } else if (fabs(f1sun + f1sha) < minf) {
                minf = fabs(f1sun + f1sha);
                minxsun = x1sun;
                minxsha = x1sha;
            }
            if (fabs(f1sun) < eps1 && fabs(f1sha) < eps1) break;
            if (f1sun * f0sun < 0.0 && f1sha * f0sha < 0.0) {
                brent_phs(s, &xsun, *x0sun, x1sun, f0sun, f1sun, &xsha, *x0sha, x1sha, f0sha, f1sha, tolsun, gs_sun,
                          gs_sha, bsun, bsha);
                *x0sun = xsun;
                *x0sha = xsha;
                break;
            }
            if (iter2 > itmax) {
                x1sun = minxsun;
                x1sha = minxsha;
                ci_func_phs(s, x, x1sun, x1sha, &f1sun, &f1sha, bsun, bsha, bflag, gs0sun, gs0sha, gs_sun, gs_sha);
                break;
            }
        }
        if (*bsun > 0.01) gs0sun = *gs_sun / *bsun;
        if (*bsha > 0.01) gs0sha = *gs_sha / *bsha;
        bflag = 1;
        if (fabs(dbsun) < toldb && fabs(dbsha) < toldb) break;
        if (iter1 > itmax) break;
    }
    *x0sun = x1sun;
    *x0sha = x1sha;
    double soilflux = oracle_phs_getvegwp_one(&s->leaf, x, *gs_sun, *gs_sha, s->qsatl, s->qaf);
    for (int k = 0; k < 4; ++k) vegwp[k] = x[k];
    if (soilflux < 0.0) soilflux = 0.0;
    return soilflux;
}

/* PhotosynthesisHydraulicStress (:2704-3811) for patch p, the root-soil
 * conductances already in phs->k_soil_root (their loop, :3062-3114, reads
 * nothing the iteration changes).  Writes the patch's outputs and btran,
 * qflx_tran_veg where the reference does. */
static int photosynthesis_phs(const ctx_t *X, int p, double esat_tv, double eair, double oair, double cair, double rb,
                              double dayl_factor, double qsatl, double qaf)
{
    const cl_t L = X->L;
    const ctsm_canopy_state *S = X->cs;
    const ctsm_phs_canopy_state *H = X->phc;
    const ctsm_canopy_params *P = X->prm;
    const ctsm_pft_params *F = X->pft;
    const ctsm_phs_params *PP = X->pp;
    const int stomatal = X->cfg->stomatalcond_method;
    const int c = PA(S->column, p);
    const int g = PA(S->gridcell, p);
    const int ivt = PA(S->itype, p);
    const double forc_pbot = CA(S->forc_pbot, c);
    const double t_veg = PA(S->t_veg, p);
    const double t10 = PA(S->t_a10, p);
    const double tgcm = PA(S->thm, p);
    const int nrad = PA(S->nrad, p);
    const double par_z[2] = {PA(S->parsun_z, p), PA(S->parsha_z, p)};
    const double lai_z[2] = {PA(S->laisun_z, p), PA(S->laisha_z, p)};
    const double vcmaxcint[2] = {PA(S->vcmaxcintsun, p), PA(S->vcmaxcintsha, p)};
    const double o3coefv[2] = {PA(S->o3coefvsun, p), PA(S->o3coefvsha, p)};
    const double o3coefg[2] = {PA(S->o3coefgsun, p), PA(S->o3coefgsha, p)};
    double *ci_z[2] = {&PA(S->cisun_z, p), &PA(S->cisha_z, p)};
    double *rs[2] = {&PA(S->rssun, p), &PA(S->rssha, p)};
    double *rs_z[2] = {&PA(S->rssun_z, p), &PA(S->rssha_z, p)};
    double *lmr[2] = {&PA(S->lmrsun, p), &PA(S->lmrsha, p)};
    double *lmr_z[2] = {&PA(S->lmrsun_z, p), &PA(S->lmrsha_z, p)};
    double *psn[2] = {&PA(S->psnsun, p), &PA(S->psnsha, p)};
    double *psn_z[2] = {&PA(S->psnsun_z, p), &PA(S->psnsha_z, p)};
    double *psn_wc[2] = {&PA(S->psnsun_wc, p), &PA(S->psnsha_wc, p)};
    double *psn_wj[2] = {&PA(S->psnsun_wj, p), &PA(S->psnsha_wj, p)};
    double *psn_wp[2] = {&PA(S->psnsun_wp, p), &PA(S->psnsha_wp, p)};
    double *gs_mol[2] = {&PA(S->gs_mol_sun, p), &PA(S->gs_mol_sha, p)};
    double *gs_mol_ln[2] = {&PA(S->gs_mol_sun_ln, p), &PA(S->gs_mol_sha_ln, p)};
    double *an[2] = {&PA(H->an_sun, p), &PA(H->an_sha, p)};
    double *bsun = &PA(H->bsun, p), *bsha = &PA(H->bsha, p);
    double *bb[2] = {bsun, bsha};
    const int np = PP->npft;
    const int crop = PP->crop ? (PP->crop[ivt] != 0.0) : 0;
    const int modify = X->cfg->modifyphoto_and_lmr_forcrop != 0;
    int rc = ORC_OK;
    phs_t s;
    memset(&s, 0, sizeof(s));

    /* misc parameters (:3117-3160) */
    const double lmrc = fth25(P->lmrhd, P->lmrse);
    const int c3 = ((int)lrint(F->c3psn[ivt]) == 1);
    PA(S->c3flag, p) = c3;
    double bbbopt = 0.0;
    if (c3) {
        PA(S->qe, p) = 0.0;
        if (stomatal == CTSM_STOMATA_BB1987) bbbopt = 10000.0;
    } else {
        PA(S->qe, p) = 0.05;
        if (stomatal == CTSM_STOMATA_BB1987) bbbopt = 40000.0;
    }
    if (stomatal == CTSM_STOMATA_BB1987) {
        PA(S->bbb, p) = bbbopt;
        PA(S->mbb, p) = F->mbbopt[ivt];
    }
    const double kc25 = P->kc25_coef * forc_pbot;
    const double ko25 = P->ko25_coef * forc_pbot;
    const double sco = 0.5 * 0.209 / P->cp25_yr2000;
    const double cp25 = 0.5 * oair / sco;
    PA(S->kc, p) = kc25 * ft(t_veg, P->kcha);
    PA(S->ko, p) = ko25 * ft(t_veg, P->koha);
    PA(S->cp, p) = cp25 * ft(t_veg, P->cpha);

    /* canopy-top capacities and the layer loop (:3165-3480) */
    if (F->slatop[ivt] * F->leafcn[ivt] <= 0.0) return ORC_ENDRUN;
    PA(S->lnca, p) = 1.0 / (F->slatop[ivt] * F->leafcn[ivt]);
    PA(S->lnca, p) = fmin90(PA(S->lnca, p), 10.0);
    const double lnc = PA(S->lnca, p);
    double vcmax25top = lnc * F->flnr[ivt] * P->fnr * P->act25 * dayl_factor;
    vcmax25top = vcmax25top * F->fnitr[ivt];
    const double t10c = fmin90(fmax90(t10 - C_TFRZ, 11.0), 35.0);
    const double jmax25top = ((2.59 - 0.035 * t10c) * vcmax25top) * P->jmax25top_sf;
    const double tpu25top = P->tpu25ratio * vcmax25top;
    const double kp25top = P->kp25ratio * vcmax25top;
    double lmr25top;
    if (c3) lmr25top = vcmax25top * P->leaf_mr_vcm;
    else lmr25top = vcmax25top * 0.025;
    double vcmax_z[2] = {0.0, 0.0}, jmax_z[2] = {0.0, 0.0}, tpu_z[2] = {0.0, 0.0}, kp_z[2] = {0.0, 0.0};
    if (nrad >= 1) {
        for (int k = 0; k < 2; ++k) {
            const double lmr25 = lmr25top * vcmaxcint[k];
            if (c3) {
                *lmr_z[k] = lmr25 * ft(t_veg, P->lmrha) * fth(t_veg, P->lmrhd, P->lmrse, lmrc);
            } else {
                *lmr_z[k] = lmr25 * cm_pow(2.0, (t_veg - (C_TFRZ + 25.0)) / 10.0);
                *lmr_z[k] = *lmr_z[k] / (1.0 + cm_exp(1.3 * (t_veg - (C_TFRZ + 55.0))));
            }
        }
        for (int k = 0; k < 2; ++k)
            *lmr_z[k] = *lmr_z[k] * fmin90((0.2 * cm_exp(3.218 * PA(H->tlai_z, p))), 1.0);
        if (par_z[SUN] > 0.0) {
            const double vcmaxse = (668.39 - 1.07 * t10c) * P->vcmaxse_sf;
            const double jmaxse = (659.70 - 0.75 * t10c) * P->jmaxse_sf;
            const double tpuse = (668.39 - 1.07 * t10c) * P->tpuse_sf;
            const double vcmaxc = fth25(P->vcmaxhd, vcmaxse);
            const double jmaxc = fth25(P->jmaxhd, jmaxse);
            const double tpuc = fth25(P->tpuhd, tpuse);
            for (int k = 0; k < 2; ++k) {
                const double vcmax25 = vcmax25top * vcmaxcint[k];
                const double jmax25 = jmax25top * vcmaxcint[k];
                const double tpu25 = tpu25top * vcmaxcint[k];
                vcmax_z[k] = vcmax25 * ft(t_veg, P->vcmaxha) * fth(t_veg, P->vcmaxhd, vcmaxse, vcmaxc);
                jmax_z[k] = jmax25 * ft(t_veg, P->jmaxha) * fth(t_veg, P->jmaxhd, jmaxse, jmaxc);
                tpu_z[k] = tpu25 * ft(t_veg, P->tpuha) * fth(t_veg, P->tpuhd, tpuse, tpuc);
            }
            if (!c3) {
                for (int k = 0; k < 2; ++k) {
                    vcmax_z[k] = vcmax25top * vcmaxcint[k] * cm_pow(2.0, (t_veg - (C_TFRZ + 25.0)) / 10.0);
                    vcmax_z[k] = vcmax_z[k] / (1.0 + cm_exp(0.2 * ((C_TFRZ + 15.0) - t_veg)));
                    vcmax_z[k] = vcmax_z[k] / (1.0 + cm_exp(0.3 * (t_veg - (C_TFRZ + 40.0))));
                }
            }
            for (int k = 0; k < 2; ++k)
                kp_z[k] = kp25top * vcmaxcint[k] * cm_pow(2.0, (t_veg - (C_TFRZ + 25.0)) / 10.0);
        }
        for (int k = 0; k < 2; ++k)
            if (X->cfg->light_inhibit && par_z[k] > 0.0) *lmr_z[k] = *lmr_z[k] * 0.67;
        for (int k = 0; k < 2; ++k) {
            PHS2(H->vcmax_z_phs, p, k) = vcmax_z[k];
            PHS2(H->tpu_z_phs, p, k) = tpu_z[k];
            PHS2(H->kp_z_phs, p, k) = kp_z[k];
        }
    }

    /* leaf-level photosynthesis and conductance (:3485-3700) */
    const double rsmax0 = 2.e4;
    const double cf = forc_pbot / (C_RGAS * 1.e-3 * tgcm) * 1.e06;
    const double gb = 1.0 / rb;
    PA(S->gb_mol, p) = gb * cf;
    const double gb_mol = PA(S->gb_mol, p);
    double psn_wc_z[2] = {0.0, 0.0}, psn_wj_z[2] = {0.0, 0.0}, psn_wp_z[2] = {0.0, 0.0};
    if (nrad >= 1) {
        /* calcstress's view of the patch */
        const ctsm_phs_state *Q = X->phs;
        s.leaf.laisun = PA(S->laisun, p); s.leaf.laisha = PA(S->laisha, p); s.leaf.elai = PA(S->elai, p);
        s.leaf.esai = PA(S->esai, p); s.leaf.tsai = PA(Q->tsai, p); s.leaf.htop = PA(S->htop, p);
        s.leaf.fdry = PA(S->fdry, p); s.leaf.thm = tgcm; s.leaf.rho = CA(S->forc_rho, c); s.leaf.pbot = forc_pbot;
        s.leaf.gb = gb_mol;
        for (int k = 0; k < 4; ++k) {
            s.leaf.kmax[k] = PP->kmax[k * np + ivt];
            s.leaf.psi50[k] = PP->psi50[k * np + ivt];
            s.leaf.ck[k] = PP->ck[k * np + ivt];
        }
        s.leaf.k = &SOIP(Q->k_soil_root, p, 1);
        s.leaf.smp = &SOIC(Q->smp_l, c, 1);
        s.leaf.z = &SNOC(Q->z, c, 1);
        s.leaf.sk = L.ldp;
        s.leaf.ss = L.ldc;
        s.leaf.n = X->nlevsoi;
        s.qsatl = qsatl;
        s.qaf = qaf;
        s.morning = PA(H->local_morning, p) != 0;
        const double gsmin = (stomatal == CTSM_STOMATA_MEDLYN2011) ? F->medlynintercept[ivt] : PA(S->bbb, p);
        double vegwp[4];
        for (int k = 0; k < 4; ++k) vegwp[k] = VEGA(H->vegwp, p, k);
        if (par_z[SUN] <= 0.0) {
            /* night: the stress factors at the minimum conductance (:3500-3548) */
            vegwp[SUN] = 1.0;
            phs_calcstress(&s, vegwp, bsun, bsha, gsmin, gsmin);
            for (int k = 0; k < 2; ++k) {
                PHS2(H->ac_phs, p, k) = 0.0;
                PHS2(H->aj_phs, p, k) = 0.0;
                PHS2(H->ap_phs, p, k) = 0.0;
                PHS2(H->ag_phs, p, k) = 0.0;
                if (!crop || !modify) *an[k] = PHS2(H->ag_phs, p, k) - *bb[k] * *lmr_z[k];
                else *an[k] = PHS2(H->ag_phs, p, k) - *lmr_z[k];
                *psn_z[k] = 0.0;
                psn_wc_z[k] = 0.0;
                psn_wj_z[k] = 0.0;
                psn_wp_z[k] = 0.0;
                *rs_z[k] = fmin90(rsmax0, 1.0 / (fmax90(*bb[k] * gsmin, 1.0)) * cf);
                *ci_z[k] = 0.0;
            }
            for (int k = 0; k < 2; ++k) *gs_mol[k] = cf / *rs_z[k];
        } else {
            /* day (:3550-3697) */
            const double ceair = fmin90(eair, esat_tv);
            double rh_can;
            if (stomatal == CTSM_STOMATA_BB1987) {
                rh_can = ceair / esat_tv;
            } else {
                rh_can = fmax90((esat_tv - ceair), 50.0) * 0.001;
                PA(S->vpd_can, p) = rh_can;
            }
            for (int k = 0; k < 2; ++k) {
                const double qabs = 0.5 * (1.0 - P->fnps) * par_z[k] * 4.6;
                double r1, r2;
                if (oracle_quadratic(P->theta_psii, -(qabs + jmax_z[k]), qabs * jmax_z[k], &r1, &r2)) rc = ORC_ENDRUN;
                s.je[k] = fmin90(r1, r2);
            }
            for (int k = 0; k < 2; ++k) *ci_z[k] = c3 ? 0.7 * cair : 0.4 * cair;
            s.c3 = c3;
            s.stomatal = stomatal;
            s.forc_pbot = forc_pbot;
            s.cair = cair;
            s.oair = oair;
            s.gb_mol = gb_mol;
            s.rh_can = rh_can;
            s.theta_cj = F->theta_cj[ivt];
            s.theta_ip = P->theta_ip;
            s.qe = PA(S->qe, p);
            s.bbb = PA(S->bbb, p);
            s.mbb = PA(S->mbb, p);
            s.medlynslope = F->medlynslope[ivt];
            s.medlynintercept = F->medlynintercept[ivt];
            s.cp = PA(S->cp, p);
            s.kc = PA(S->kc, p);
            s.ko = PA(S->ko, p);
            for (int k = 0; k < 2; ++k) {
                s.vcmax_z[k] = vcmax_z[k];
                s.tpu_z[k] = tpu_z[k];
                s.kp_z[k] = kp_z[k];
                s.lmr_z[k] = *lmr_z[k];
                s.par_z[k] = par_z[k];
                s.ac[k] = PHS2(H->ac_phs, p, k);
                s.aj[k] = PHS2(H->aj_phs, p, k);
                s.ap[k] = PHS2(H->ap_phs, p, k);
                s.ag[k] = PHS2(H->ag_phs, p, k);
                s.an[k] = *an[k];
            }
            const double soilflux = hybrid_phs(&s, ci_z[SUN], ci_z[SHA], vegwp, bsun, bsha, gs_mol[SUN], gs_mol[SHA]);
            s.tran_set = 1;
            s.tran = soilflux;
            if (s.err) rc = ORC_ENDRUN;
            for (int k = 0; k < 2; ++k) {
                PHS2(H->ac_phs, p, k) = s.ac[k];
                PHS2(H->aj_phs, p, k) = s.aj[k];
                PHS2(H->ap_phs, p, k) = s.ap[k];
                PHS2(H->ag_phs, p, k) = s.ag[k];
                *an[k] = s.an[k];
            }
            for (int k = 0; k < 4; ++k)
                VEGA(H->vegwp_ln, p, k) = GA(S->near_local_noon, g) ? vegwp[k] : C_SPVAL;
            const double gs_slope = (stomatal == CTSM_STOMATA_MEDLYN2011) ? F->medlynslope[ivt] : PA(S->mbb, p);
            (void)gs_slope;   /* the reference's Ball-Berry check only logs */
            for (int k = 0; k < 2; ++k)
                if (*an[k] < 0.0) *gs_mol[k] = fmax90(*bb[k] * gsmin, 1.0);
            for (int k = 0; k < 2; ++k) *gs_mol_ln[k] = GA(S->near_local_noon, g) ? *gs_mol[k] : C_SPVAL;
            for (int k = 0; k < 2; ++k) {
                double cs = cair - 1.4 / gb_mol * *an[k] * forc_pbot;
                cs = fmax90(cs, MAX_CS);
                (void)cs;
                *ci_z[k] = cair - *an[k] * forc_pbot * (1.4 * *gs_mol[k] + 1.6 * gb_mol) / (gb_mol * *gs_mol[k]);
                *ci_z[k] = fmax90(*ci_z[k], 1.e-06);
            }
            for (int k = 0; k < 2; ++k) {
                const double gs = *gs_mol[k] / cf;
                *rs_z[k] = fmin90(1.0 / gs, rsmax0);
                *rs_z[k] = *rs_z[k] / o3coefg[k];
            }
            for (int k = 0; k < 2; ++k) {
                *psn_z[k] = PHS2(H->ag_phs, p, k);
                *psn_z[k] = *psn_z[k] * o3coefv[k];
                psn_wc_z[k] = 0.0;
                psn_wj_z[k] = 0.0;
                psn_wp_z[k] = 0.0;
                const double ac = PHS2(H->ac_phs, p, k), aj = PHS2(H->aj_phs, p, k), ap = PHS2(H->ap_phs, p, k);
                if (ac <= aj && ac <= ap) psn_wc_z[k] = *psn_z[k];
                else if (aj < ac && aj <= ap) psn_wj_z[k] = *psn_z[k];
                else if (ap < ac && ap < aj) psn_wp_z[k] = *psn_z[k];
            }
            /* the leaf-surface humidity and the Ball-Berry check (:3677-3695) only log */
            if (*gs_mol[SUN] < 0.0 || *gs_mol[SHA] < 0.0) rc = ORC_ENDRUN;
        }
        for (int k = 0; k < 4; ++k) VEGA(H->vegwp, p, k) = vegwp[k];
        if (s.stressed)
            for (int k = 0; k < 4; ++k) VEGA(H->vegwp_pd, p, k) = s.pd[k];
        if (s.tran_set) PA(S->qflx_tran_veg, p) = s.tran;
    }

    /* canopy sums (:3705-3800) and btran */
    double laican[2] = {0.0, 0.0};
    for (int k = 0; k < 2; ++k) {
        double psncan = 0.0, psncan_wc = 0.0, psncan_wj = 0.0, psncan_wp = 0.0, lmrcan = 0.0, gscan = 0.0;
        if (nrad >= 1) {
            psncan = psncan + *psn_z[k] * lai_z[k];
            psncan_wc = psncan_wc + psn_wc_z[k] * lai_z[k];
            psncan_wj = psncan_wj + psn_wj_z[k] * lai_z[k];
            psncan_wp = psncan_wp + psn_wp_z[k] * lai_z[k];
            if (!crop && modify) lmrcan = lmrcan + *lmr_z[k] * lai_z[k] * *bb[k];
            else lmrcan = lmrcan + *lmr_z[k] * lai_z[k];
            gscan = gscan + lai_z[k] / (rb + *rs_z[k]);
            laican[k] = laican[k] + lai_z[k];
        }
        if (laican[k] > 0.0) {
            *psn[k] = psncan / laican[k];
            *psn_wc[k] = psncan_wc / laican[k];
            *psn_wj[k] = psncan_wj / laican[k];
            *psn_wp[k] = psncan_wp / laican[k];
            *lmr[k] = lmrcan / laican[k];
            *rs[k] = laican[k] / gscan - rb;
        } else {
            *psn[k] = 0.0;
            *psn_wc[k] = 0.0;
            *psn_wj[k] = 0.0;
            *psn_wp[k] = 0.0;
            *lmr[k] = 0.0;
            *rs[k] = 0.0;
        }
    }
    if (laican[SHA] + laican[SUN] > 0.0)
        PA(S->btran, p) = *bsun * (laican[SUN] / (laican[SUN] + laican[SHA]))
                          + *bsha * (laican[SHA] / (laican[SUN] + laican[SHA]));
    else
        PA(S->btran, p) = *bsun;
    return rc;
}

/* Test access: PhotosynthesisHydraulicStress for patch p (1-based global
 * index) at the leaf temperature held in t_veg(p), with the canopy-air
 * arguments the iteration would pass; the conductances must already be in
 * phs->k_soil_root.  Writes the patch's outputs as in the iteration. */
int oracle_photosynthesis_phs_patch(const ctsm_config *cfg, const ctsm_canopy_params *prm, const ctsm_pft_params *pft,
                                    const ctsm_phs_params *pp, const ctsm_layout *lay, const ctsm_col_state *col,
                                    const ctsm_canopy_state *S, const ctsm_phs_state *phs,
                                    const ctsm_phs_canopy_state *phc, int p, double esat_tv, double eair,
                                    double oair, double cair, double rb, double dayl_factor, double qsatl,
                                    double qaf)
{
    ctx_t X;
    X.L.begc = lay->begc_alloc;
    X.L.ldc = lay->ldc;
    X.L.begp = lay->begp_alloc;
    X.L.ldp = lay->ldp;
    X.L.begg = lay->begg_alloc;
    X.L.nlevsno = cfg->nlevsno;
    X.L.nlevgrnd = cfg->nlevgrnd;
    X.cfg = cfg;
    X.prm = prm;
    X.pft = pft;
    X.col = col;
    X.cs = S;
    X.phs = phs;
    X.phc = phc;
    X.pp = pp;
    X.nlevsoi = cfg->nlevsoi;
    return photosynthesis_phs(&X, p, esat_tv, eair, oair, cair, rb, dayl_factor, qsatl, qaf);
}

/* Test access: Photosynthesis for patch p (1-based global index) and one
 * phase (sun = 1, shade = 0) at the leaf temperature held in t_veg(p), with
 * the canopy-air arguments the iteration would pass; writes the same patch
 * outputs.  Used only by the pins of the temperature kinetics and the
 * co-limitation. */
int oracle_photosynthesis_patch(const ctsm_config *cfg, const ctsm_canopy_params *prm, const ctsm_pft_params *pft,
                                const ctsm_layout *lay, const ctsm_col_state *col, const ctsm_canopy_state *S,
                                int p, int sun, double esat_tv, double eair, double oair, double cair, double rb,
                                double btran, double dayl_factor)
{
    ctx_t X;
    X.L.begc = lay->begc_alloc;
    X.L.ldc = lay->ldc;
    X.L.begp = lay->begp_alloc;
    X.L.ldp = lay->ldp;
    X.L.begg = lay->begg_alloc;
    X.L.nlevsno = cfg->nlevsno;
    X.L.nlevgrnd = cfg->nlevgrnd;
    X.cfg = cfg;
    X.prm = prm;
    X.pft = pft;
    X.col = col;
    X.cs = S;
    return photosynthesis(&X, p, sun, esat_tv, eair, oair, cair, rb, btran, dayl_factor);
}

/* Test access to the temperature functions (PhotosynthesisMod.F90:2488-2561). */
double oracle_ft(double tl, double ha) { return ft(tl, ha); }
double oracle_fth(double tl, double hd, double se, double c) { return fth(tl, hd, se, c); }
double oracle_fth25(double hd, double se) { return fth25(hd, se); }

/* ---------------------------------------------------------------------- */
/* CanopyFluxesMod.F90:191-1765.  Returns 0, or ORC_ENDRUN with *bad_patch
 * set where the reference calls endrun. */
static int canopyfluxes(const ctsm_config *cfg, const ctsm_canopy_params *prm, const ctsm_pft_params *pft,
                        const ctsm_layout *lay, const ctsm_col_state *col, const ctsm_canopy_state *S,
                        const ctsm_phs_params *pp, const ctsm_phs_state *phs, const ctsm_phs_canopy_state *phc,
                        int num_exposedvegp, const int *filter_exposedvegp, double dtime, int *bad_patch)
{
    ctx_t X;
    cl_t L;
    L.begc = lay->begc_alloc;
    L.ldc = lay->ldc;
    L.begp = lay->begp_alloc;
    L.ldp = lay->ldp;
    L.begg = lay->begg_alloc;
    L.nlevsno = cfg->nlevsno;
    L.nlevgrnd = cfg->nlevgrnd;
    X.L = L;
    X.cfg = cfg;
    X.prm = prm;
    X.pft = pft;
    X.col = col;
    X.cs = S;
    X.phs = phs;
    X.phc = phc;
    X.pp = pp;
    X.nlevsoi = cfg->nlevsoi;
    const int hydrstress = phc != NULL;
    const ctsm_canopy_params *P = prm;
    const ctsm_pft_params *F = pft;
    const int nlevgrnd = cfg->nlevgrnd;
    const double btran0 = 0.0, zii = 1000.0, beta = 1.0, delmax = 1.0, dlemin = 0.1, dtmin = 0.01;
    const int itmin = 2;
    const double ria = 0.5;
    const double k_vert = 0.1, k_cyl_vol = 1.0, k_cyl_area = 1.0, k_internal = 0.0;
    const double min_stem_diameter = 0.05, min_lai = 0.1;
    const int np = L.ldp;
    int rc = ORC_OK;
    *bad_patch = 0;

#define LOC(name) double *name = (double *)calloc((size_t)np, sizeof(double))
    LOC(zldis); LOC(dth); LOC(dthv); LOC(dqh); LOC(ur); LOC(temp1); LOC(temp12m); LOC(temp2); LOC(temp22m);
    LOC(rb); LOC(rah1); LOC(rah2); LOC(raw1); LOC(raw2); LOC(wtg); LOC(wta0); LOC(wtl0); LOC(wtstem0);
    LOC(wtal); LOC(wtga); LOC(wtgq); LOC(wtaq0); LOC(wtlq0); LOC(wtalq); LOC(el); LOC(qsatl); LOC(qsatldT);
    LOC(air); LOC(bir); LOC(cir); LOC(delq); LOC(del); LOC(del2); LOC(dele); LOC(det); LOC(efeb); LOC(efe);
    LOC(obuold); LOC(tlbef); LOC(tl_ini); LOC(ts_ini); LOC(err); LOC(co2); LOC(o2); LOC(svpts); LOC(eah);
    LOC(fm); LOC(dayl_factor); LOC(dt_veg); LOC(dbh); LOC(cp_leaf); LOC(cp_stem); LOC(rstem); LOC(dt_stem);
    LOC(frac_rad_abs_by_stem); LOC(lw_stem); LOC(lw_leaf); LOC(sa_stem); LOC(sa_leaf); LOC(sa_internal);
    LOC(uuc); LOC(snocan_baseline);
#undef LOC
    int *nmozsgn = (int *)calloc((size_t)np, sizeof(int));
    int fn = num_exposedvegp;
    int *filterp = (int *)malloc(sizeof(int) * (size_t)(fn > 0 ? fn : 1));
    int *fporig = (int *)malloc(sizeof(int) * (size_t)(fn > 0 ? fn : 1));
    int f, p, c, g, j;
    for (f = 0; f < fn; ++f) filterp[f] = filter_exposedvegp[f];

    /* photosyns_inst%TimeStepInit (PhotosynthesisMod.F90:1143-1201): every patch of the bounds */
    for (p = L.begp; p < L.begp + L.ldp; ++p) {
        PA(S->psnsun, p) = 0.0; PA(S->psnsun_wc, p) = 0.0; PA(S->psnsun_wj, p) = 0.0; PA(S->psnsun_wp, p) = 0.0;
        PA(S->psnsha, p) = 0.0; PA(S->psnsha_wc, p) = 0.0; PA(S->psnsha_wj, p) = 0.0; PA(S->psnsha_wp, p) = 0.0;
        PA(S->fpsn, p) = 0.0; PA(S->fpsn_wc, p) = 0.0; PA(S->fpsn_wj, p) = 0.0; PA(S->fpsn_wp, p) = 0.0;
    }

    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        PA(del, p) = 0.0;
        PA(efeb, p) = 0.0;
        PA(wtlq0, p) = 0.0;
        PA(wtalq, p) = 0.0;
        PA(wtgq, p) = 0.0;
        PA(wtaq0, p) = 0.0;
        PA(obuold, p) = 0.0;
        PA(S->btran, p) = btran0;
        PA(S->dhsdt_canopy, p) = 0.0;
        PA(S->eflx_sh_stem, p) = 0.0;
    }

    /* biomass heat capacities (:717-819), satellite phenology (use_cn = .false.) */
    if (cfg->use_biomass_heat_storage) {
        for (f = 0; f < fn; ++f) {
            p = filterp[f];
            const int it = PA(S->itype, p);
            const double elai = PA(S->elai, p), esai = PA(S->esai, p), htop = PA(S->htop, p);
            PA(frac_rad_abs_by_stem, p) = (esai) / (elai + esai);
            if (elai > 0.0) PA(frac_rad_abs_by_stem, p) = k_vert * PA(frac_rad_abs_by_stem, p);
            PA(dbh, p) = F->dbh[it];
            PA(sa_leaf, p) = elai;
            PA(sa_leaf, p) = 2.0 * PA(sa_leaf, p);
            PA(sa_stem, p) = F->nstem[it] * (htop * C_PI * PA(dbh, p));
            PA(sa_stem, p) = k_cyl_area * PA(sa_stem, p);
            if (!(F->is_tree[it] || F->is_shrub[it]) || PA(dbh, p) < min_stem_diameter) {
                PA(frac_rad_abs_by_stem, p) = 0.0;
                PA(sa_stem, p) = 0.0;
                PA(sa_leaf, p) = PA(sa_leaf, p) + esai;
            } else {
                if (elai < min_lai) PA(sa_leaf, p) = PA(sa_leaf, p) + esai;
            }
            PA(S->leaf_biomass, p) = (1.e-3 * C_C_TO_B / F->slatop[it]) * fmax90(0.01, 0.5 * PA(sa_leaf, p))
                                     / (1.0 - F->fbw[it]);
            const double carea_stem = C_PI * ((PA(dbh, p) * 0.5) * (PA(dbh, p) * 0.5));
            PA(S->stem_biomass, p) = carea_stem * htop * k_cyl_vol * F->nstem[it] * F->wood_density[it]
                                     / (1.0 - F->fbw[it]);
            PA(sa_internal, p) = fmin90(PA(sa_leaf, p), PA(sa_stem, p));
            PA(sa_internal, p) = k_internal * PA(sa_internal, p);
            PA(cp_leaf, p) = PA(S->leaf_biomass, p) * (C_CDRYBIO * (1.0 - F->fbw[it]) + (F->fbw[it]) * C_CWATER);
            PA(cp_stem, p) = PA(S->stem_biomass, p) * (C_CDRYBIO * (1.0 - F->fbw[it]) + (F->fbw[it]) * C_CWATER);
            PA(cp_stem, p) = k_cyl_vol * PA(cp_stem, p);
            PA(rstem, p) = F->rstem_per_dbh[it] * PA(dbh, p);
        }
    } else {
        for (f = 0; f < fn; ++f) {
            p = filterp[f];
            PA(sa_leaf, p) = (PA(S->elai, p) + PA(S->esai, p));
            PA(frac_rad_abs_by_stem, p) = 0.0;
            PA(sa_stem, p) = 0.0;
            PA(sa_internal, p) = 0.0;
            PA(cp_leaf, p) = 0.0;
            PA(cp_stem, p) = 0.0;
            PA(rstem, p) = 0.0;
        }
    }

    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        g = PA(S->gridcell, p);
        PA(dayl_factor, p) = fmin90(1.0, fmax90(0.01, (GA(S->dayl, g) * GA(S->dayl, g))
                                                          / (GA(S->max_dayl, g) * GA(S->max_dayl, g))));
    }
    for (p = L.begp; p < L.begp + L.ldp; ++p) PA(S->rb1, p) = 0.0;

    /* calc_effective_soilporosity / calc_volumetric_h2oliq over the columns of the filter */
    for (j = 1; j <= nlevgrnd; ++j) {
        for (f = 0; f < fn; ++f) {
            c = PA(S->column, filterp[f]);
            const double vol_ice = fmin90(SOIC(col->watsat, c, j),
                                          SNOC(col->h2osoi_ice, c, j) / (C_DENICE * SNOC(col->dz, c, j)));
            SOIC(col->eff_porosity, c, j) = SOIC(col->watsat, c, j) - vol_ice;
        }
    }
    for (j = 1; j <= nlevgrnd; ++j) {
        for (f = 0; f < fn; ++f) {
            c = PA(S->column, filterp[f]);
            SOIC(S->h2osoi_liqvol, c, j) = fmin90(SOIC(col->eff_porosity, c, j),
                                                  SNOC(col->h2osoi_liq, c, j) / (SNOC(col->dz, c, j) * C_DENH2O));
        }
    }
    /* calc_root_moist_stress_clm45default (perchroot = .false.) */
    for (j = 1; j <= nlevgrnd; ++j) {
        for (f = 0; f < fn; ++f) {
            p = filterp[f];
            c = PA(S->column, p);
            const int it = PA(S->itype, p);
            if (SOIC(S->h2osoi_liqvol, c, j) <= 0.0 || SNOC(col->t_soisno, c, j) <= C_TFRZ - 2.0) {
                SOIP(S->rootr, p, j) = 0.0;
            } else {
                const double s_node = fmax90(SOIC(S->h2osoi_liqvol, c, j) / SOIC(col->eff_porosity, c, j), 0.01);
                double smp_node = -SOIC(col->sucsat, c, j) * cm_pow(s_node, -SOIC(col->bsw, c, j));
                smp_node = fmax90(F->smpsc[it], smp_node);
                SOIP(S->rresis, p, j) = fmin90((SOIC(col->eff_porosity, c, j) / SOIC(col->watsat, c, j))
                                                   * (smp_node - F->smpsc[it]) / (F->smpso[it] - F->smpsc[it]),
                                               1.0);
                SOIP(S->rootr, p, j) = SOIP(S->rootfr, p, j) * SOIP(S->rresis, p, j);
                PA(S->btran, p) = PA(S->btran, p) + fmax90(SOIP(S->rootr, p, j), 0.0);
            }
        }
    }
    for (j = 1; j <= nlevgrnd; ++j) {
        for (f = 0; f < fn; ++f) {
            p = filterp[f];
            if (PA(S->btran, p) > btran0) SOIP(S->rootr, p, j) = SOIP(S->rootr, p, j) / PA(S->btran, p);
            else SOIP(S->rootr, p, j) = 0.0;
        }
    }

    /* aerodynamic parameters (:897-943) */
    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        c = PA(S->column, p);
        g = PA(S->gridcell, p);
        const int it = PA(S->itype, p);
        double lt;
        if (cfg->z0param_method == CTSM_Z0_ZENGWANG2007) {
            lt = fmin90(PA(S->elai, p) + PA(S->esai, p), C_TLSAI_CRIT);
            const double egvf = (1.0 - C_ALPHA_AERO * cm_exp(-lt)) / (1.0 - C_ALPHA_AERO * cm_exp(-C_TLSAI_CRIT));
            PA(S->displa, p) = egvf * PA(S->displa, p);
            PA(S->z0mv, p) = cm_exp(egvf * cm_log(PA(S->z0mv, p)) + (1.0 - egvf) * cm_log(CA(S->z0mg, c)));
        } else {
            const double htop = PA(S->htop, p);
            lt = fmax90(1.e-5, PA(S->elai, p) + PA(S->esai, p));
            PA(S->displa, p) = htop * (1.0 - (1.0 - cm_exp(-cm_pow(C_CD1 * lt, 0.5))) / cm_pow(C_CD1 * lt, 0.5));
            lt = fmin90(lt, F->z0v_LAImax[it]);
            double delt = 2.0;
            const double U_ustar_ini = cm_pow(F->z0v_Cs[it] + F->z0v_Cr[it] * lt * 0.5, -0.5) * F->z0v_c[it] * lt * 0.25;
            double U_ustar = U_ustar_ini, U_ustar_prev;
            while (delt > 1.e-4) {
                U_ustar_prev = U_ustar;
                U_ustar = U_ustar_ini * cm_exp(U_ustar_prev);
                delt = fabs(U_ustar - U_ustar_prev);
            }
            U_ustar = 4.0 * U_ustar / lt / F->z0v_c[it];
            PA(S->z0mv, p) = htop * (1.0 - PA(S->displa, p) / htop)
                             * cm_exp(-C_VKC * U_ustar + cm_log(F->z0v_cw[it]) - 1.0 + 1.0 / F->z0v_cw[it]);
        }
        PA(S->z0hv, p) = PA(S->z0mv, p);
        PA(S->z0qv, p) = PA(S->z0mv, p);
        PA(S->forc_hgt_u_patch, p) = GA(S->forc_hgt_u, g) + PA(S->z0mv, p) + PA(S->displa, p);
        PA(S->forc_hgt_t_patch, p) = GA(S->forc_hgt_t, g) + PA(S->z0hv, p) + PA(S->displa, p);
        PA(S->forc_hgt_q_patch, p) = GA(S->forc_hgt_q, g) + PA(S->z0qv, p) + PA(S->displa, p);
    }

    /* radiation terms, QSat, initial profiles (:945-1002) */
    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        c = PA(S->column, p);
        g = PA(S->gridcell, p);
        const double emv = PA(S->emv, p), emg = CA(col->emg, c);
        PA(air, p) = emv * (1.0 + (1.0 - emv) * (1.0 - emg)) * CA(col->forc_lwrad, c);
        PA(bir, p) = -(2.0 - emv * (1.0 - emg)) * emv * C_SB;
        PA(cir, p) = emv * emg * C_SB;
        oracle_qsat(PA(S->t_veg, p), CA(S->forc_pbot, c), &PA(qsatl, p), &PA(el, p), &PA(qsatldT, p));
        PA(co2, p) = GA(S->forc_pco2, g);
        PA(o2, p) = GA(S->forc_po2, g);
        nmozsgn[p - L.begp] = 0;
        PA(S->taf, p) = (CA(col->t_grnd, c) + PA(S->thm, p)) / 2.0;
        PA(S->qaf, p) = (CA(S->forc_q, c) + CA(S->qg, c)) / 2.0;
        PA(ur, p) = fmax90(P->wind_min, sqrt(GA(S->forc_u, g) * GA(S->forc_u, g) + GA(S->forc_v, g) * GA(S->forc_v, g)));
        PA(dth, p) = PA(S->thm, p) - PA(S->taf, p);
        PA(dqh, p) = CA(S->forc_q, c) - PA(S->qaf, p);
        PA(delq, p) = CA(S->qg, c) - PA(S->qaf, p);
        PA(dthv, p) = PA(dth, p) * (1.0 + 0.61 * CA(S->forc_q, c)) + 0.61 * CA(S->forc_th, c) * PA(dqh, p);
        PA(zldis, p) = PA(S->forc_hgt_u_patch, p) - PA(S->displa, p);
        if (PA(zldis, p) < 0.0) {
            rc = ORC_ENDRUN;
            if (!*bad_patch) *bad_patch = p;
        }
    }
    if (rc) goto done;

    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        c = PA(S->column, p);
        oracle_moninobukini(PA(ur, p), CA(S->thv, c), PA(dthv, p), PA(zldis, p), PA(S->z0mv, p), P->zetamaxstable,
                            &PA(S->um, p), &PA(S->obu, p));
        PA(S->num_iter, p) = 0.0;
        PA(tl_ini, p) = PA(S->t_veg, p);
        PA(ts_ini, p) = PA(S->t_stem, p);
    }

    int itlef = 0;
    const int fnorig = fn;
    for (f = 0; f < fn; ++f) fporig[f] = filterp[f];

    while (itlef <= cfg->itmax_canopy_fluxes && fn > 0) {
        for (f = 0; f < fn; ++f) {
            p = filterp[f];
            double o[11];
            oracle_frictionvelocity(PA(S->forc_hgt_u_patch, p), PA(S->forc_hgt_t_patch, p), PA(S->forc_hgt_q_patch, p),
                                    PA(S->displa, p), PA(S->z0mv, p), PA(S->z0hv, p), PA(S->z0qv, p), PA(S->obu, p),
                                    itlef + 1, PA(ur, p), PA(S->um, p), PA(fm, p), o);
            PA(S->ustar, p) = o[0];
            PA(temp1, p) = o[1];
            PA(temp2, p) = o[2];
            PA(temp12m, p) = o[3];
            PA(temp22m, p) = o[4];
            PA(fm, p) = o[5];
            PA(S->u10, p) = o[6];
            PA(S->u10_clm, p) = o[7];
            PA(S->va, p) = o[8];
            PA(S->fv, p) = o[9];
            PA(S->vds, p) = o[10];
        }
        for (f = 0; f < fn; ++f) {
            p = filterp[f];
            c = PA(S->column, p);
            const int it = PA(S->itype, p);
            PA(tlbef, p) = PA(S->t_veg, p);
            PA(del2, p) = PA(del, p);
            const double ustar = PA(S->ustar, p), um = PA(S->um, p);
            PA(S->ram1, p) = 1.0 / (ustar * ustar / um);
            PA(rah1, p) = 1.0 / (PA(temp1, p) * ustar);
            PA(raw1, p) = 1.0 / (PA(temp2, p) * ustar);
            PA(S->uaf, p) = um * sqrt(1.0 / (PA(S->ram1, p) * um));
            PA(uuc, p) = fmin90(0.4, (0.03 * um / ustar));
            PA(S->dleaf_patch, p) = F->dleaf[it];
            const double uaf = PA(S->uaf, p);
            const double cf = P->cv / (sqrt(uaf) * sqrt(PA(S->dleaf_patch, p)));
            PA(rb, p) = 1.0 / (cf * uaf);
            PA(S->rb1, p) = PA(rb, p);
            const double w = cm_exp(-(PA(S->elai, p) + PA(S->esai, p)));
            const double csoilb = C_VKC / (P->a_coef * cm_pow(CA(S->z0mg, c) * uaf / C_NU, P->a_exp));
            const double taf = PA(S->taf, p), tg = CA(col->t_grnd, c);
            const double ri = (C_GRAV * PA(S->htop, p) * (taf - tg)) / (taf * (uaf * uaf));
            double csoilcn;
            if (cfg->use_undercanopy_stability && (taf - tg) > 0.0) {
                const double ricsoilc = P->csoilc / (1.00 + ria * fmin90(ri, 10.0));
                csoilcn = csoilb * w + ricsoilc * (1.0 - w);
            } else {
                csoilcn = csoilb * w + P->csoilc * (1.0 - w);
            }
            if (cfg->use_biomass_heat_storage) PA(rah2, p) = 1.0 / (csoilcn * PA(uuc, p));
            else PA(rah2, p) = 1.0 / (csoilcn * uaf);
            PA(raw2, p) = PA(rah2, p);
            PA(svpts, p) = PA(el, p);
            PA(eah, p) = CA(S->forc_pbot, c) * PA(S->qaf, p) / 0.622;
            PA(S->rh_af, p) = PA(eah, p) / PA(svpts, p);
            PA(S->rah1, p) = PA(rah1, p);
            PA(S->raw1, p) = PA(raw1, p);
            PA(S->rah2, p) = PA(rah2, p);
            PA(S->raw2, p) = PA(raw2, p);
            PA(S->vpd, p) = fmax90((PA(svpts, p) - PA(eah, p)), 50.0) * 0.001;
        }
        if (hydrstress) {
            /* PhotosynthesisHydraulicStress, sun and shade together (:1133-1141) */
            for (f = 0; f < fn; ++f) {
                p = filterp[f];
                if (photosynthesis_phs(&X, p, PA(svpts, p), PA(eah, p), PA(o2, p), PA(co2, p), PA(rb, p),
                                       PA(dayl_factor, p), PA(qsatl, p), PA(S->qaf, p))) {
                    rc = ORC_ENDRUN;
                    if (!*bad_patch) *bad_patch = p;
                }
            }
        } else {
            for (f = 0; f < fn; ++f) {
                p = filterp[f];
                if (photosynthesis(&X, p, 1, PA(svpts, p), PA(eah, p), PA(o2, p), PA(co2, p), PA(rb, p),
                                   PA(S->btran, p), PA(dayl_factor, p))) {
                    rc = ORC_ENDRUN;
                    if (!*bad_patch) *bad_patch = p;
                }
            }
            for (f = 0; f < fn; ++f) {
                p = filterp[f];
                if (photosynthesis(&X, p, 0, PA(svpts, p), PA(eah, p), PA(o2, p), PA(co2, p), PA(rb, p),
                                   PA(S->btran, p), PA(dayl_factor, p))) {
                    rc = ORC_ENDRUN;
                    if (!*bad_patch) *bad_patch = p;
                }
            }
        }
        for (f = 0; f < fn; ++f) {
            p = filterp[f];
            c = PA(S->column, p);
            const double rho = CA(S->forc_rho, c);
            const double wta = 1.0 / PA(rah1, p);
            const double wtl = PA(sa_leaf, p) / PA(rb, p);
            PA(wtg, p) = 1.0 / PA(rah2, p);
            const double wtstem = PA(sa_stem, p) / (PA(rstem, p) + PA(rb, p));
            const double wtshi = 1.0 / (wta + wtl + wtstem + PA(wtg, p));
            PA(wtl0, p) = wtl * wtshi;
            const double wtg0 = PA(wtg, p) * wtshi;
            PA(wta0, p) = wta * wtshi;
            PA(wtstem0, p) = wtstem * wtshi;
            PA(wtga, p) = PA(wta0, p) + wtg0 + PA(wtstem0, p);
            PA(wtal, p) = PA(wta0, p) + PA(wtl0, p) + PA(wtstem0, p);
            const double emv = PA(S->emv, p);
            PA(lw_stem, p) = PA(sa_internal, p) * emv * C_SB * p4(PA(S->t_stem, p));
            PA(lw_leaf, p) = PA(sa_internal, p) * emv * C_SB * p4(PA(S->t_veg, p));
            double rppdry;
            const double rbp = PA(rb, p);
            if (PA(S->fdry, p) > 0.0) {
                rppdry = PA(S->fdry, p) * rbp * (PA(S->laisun, p) / (rbp + PA(S->rssun, p))
                                                 + PA(S->laisha, p) / (rbp + PA(S->rssha, p))) / PA(S->elai, p);
            } else {
                rppdry = 0.0;
            }
            double efpot = rho * ((PA(S->elai, p) + PA(S->esai, p)) / rbp) * (PA(qsatl, p) - PA(S->qaf, p));
            const double h2ocan = PA(S->liqcan, p) + PA(S->snocan, p);
            double rpp;
            if (hydrstress) {
                /* transpiration comes from the hydraulics, not from rppdry (:1219-1230) */
                if (efpot > 0.0) {
                    if (PA(S->btran, p) > btran0) rpp = rppdry + PA(S->fwet, p);
                    else rpp = PA(S->fwet, p);
                    rpp = fmin90(rpp, (PA(S->qflx_tran_veg, p) + h2ocan / dtime) / efpot);
                } else {
                    rpp = 1.0;
                }
            } else if (efpot > 0.0) {
                if (PA(S->btran, p) > btran0) {
                    PA(S->qflx_tran_veg, p) = efpot * rppdry;
                    rpp = rppdry + PA(S->fwet, p);
                } else {
                    rpp = PA(S->fwet, p);
                    PA(S->qflx_tran_veg, p) = 0.0;
                }
                rpp = fmin90(rpp, (PA(S->qflx_tran_veg, p) + h2ocan / dtime) / efpot);
            } else {
                rpp = 1.0;
                PA(S->qflx_tran_veg, p) = 0.0;
            }
            const double fvn = (double)PA(S->frac_veg_nosno, p);
            const double wtaq = fvn / PA(raw1, p);
            const double wtlq = fvn * (PA(S->elai, p) + PA(S->esai, p)) / rbp * rpp;
            const double snow_depth_c = P->z_dl;
            const double fsno_dl = CA(col->snow_depth, c) / snow_depth_c;
            const double elai_dl = P->lai_dl * (1.0 - fmin90(fsno_dl, 1.0));
            const double rdl = (1.0 - cm_exp(-elai_dl)) / (0.004 * PA(S->uaf, p));
            if (PA(delq, p) < 0.0) {
                PA(wtgq, p) = fvn / (PA(raw2, p) + rdl);
            } else {
                if (cfg->soil_resis_method == CTSM_SOILRESIS_LEEPIELKE1992)
                    PA(wtgq, p) = CA(S->soilbeta, c) * fvn / (PA(raw2, p) + rdl);
                if (cfg->soil_resis_method == CTSM_SOILRESIS_SL14)
                    PA(wtgq, p) = fvn / (PA(raw2, p) + CA(S->soilresis, c));
            }
            const double wtsqi = 1.0 / (wtaq + wtlq + PA(wtgq, p));
            const double wtgq0 = PA(wtgq, p) * wtsqi;
            PA(wtlq0, p) = wtlq * wtsqi;
            PA(wtaq0, p) = wtaq * wtsqi;
            const double wtgaq = PA(wtaq0, p) + wtgq0;
            PA(wtalq, p) = PA(wtaq0, p) + PA(wtlq0, p);
            const double dc1 = rho * C_CPAIR * wtl;
            const double dc2 = C_HVAP * rho * wtlq;
            const double tv = PA(S->t_veg, p), tg = CA(col->t_grnd, c), thm = PA(S->thm, p), ts = PA(S->t_stem, p);
            const double efsh = dc1 * (PA(wtga, p) * tv - wtg0 * tg - PA(wta0, p) * thm - PA(wtstem0, p) * ts);
            PA(S->eflx_sh_stem, p) = rho * C_CPAIR * wtstem
                                     * ((PA(wta0, p) + wtg0 + PA(wtl0, p)) * ts - wtg0 * tg - PA(wta0, p) * thm
                                        - PA(wtl0, p) * tv);
            PA(efe, p) = dc2 * (wtgaq * PA(qsatl, p) - wtgq0 * CA(S->qg, c) - PA(wtaq0, p) * CA(S->forc_q, c));
            double erre = 0.0;
            if (PA(efe, p) * PA(efeb, p) < 0.0) {
                const double efeold = PA(efe, p);
                PA(efe, p) = 0.1 * efeold;
                erre = PA(efe, p) - efeold;
            }
            const int snl = CA(col->snl, c);
            const double lw_grnd = (CA(col->frac_sno_eff, c) * p4(SNOC(col->t_soisno, c, snl + 1))
                                    + (1.0 - CA(col->frac_sno_eff, c) - CA(col->frac_h2osfc, c)) * p4(SNOC(col->t_soisno, c, 1))
                                    + CA(col->frac_h2osfc, c) * p4(CA(col->t_h2osfc, c)));
            const double fr = PA(frac_rad_abs_by_stem, p);
            PA(dt_veg, p) = ((1.0 - fr) * (PA(S->sabv, p) + PA(air, p) + PA(bir, p) * p4(tv) + PA(cir, p) * lw_grnd)
                             - efsh - PA(efe, p) - PA(lw_leaf, p) + PA(lw_stem, p)
                             - (PA(cp_leaf, p) / dtime) * (tv - PA(tl_ini, p)))
                            / ((1.0 - fr) * (-4.0 * PA(bir, p) * p3(tv))
                               + 4.0 * PA(sa_internal, p) * emv * C_SB * p3(tv)
                               + dc1 * PA(wtga, p) + dc2 * wtgaq * PA(qsatldT, p) + PA(cp_leaf, p) / dtime);
            PA(S->t_veg, p) = PA(tlbef, p) + PA(dt_veg, p);
            const double dels = PA(dt_veg, p);
            PA(del, p) = fabs(dels);
            PA(err, p) = 0.0;
            if (PA(del, p) > delmax) {
                PA(dt_veg, p) = delmax * dels / PA(del, p);
                PA(S->t_veg, p) = PA(tlbef, p) + PA(dt_veg, p);
                const double tb = PA(tlbef, p), dtv = PA(dt_veg, p);
                PA(err, p) = (1.0 - fr) * (PA(S->sabv, p) + PA(air, p) + PA(bir, p) * p3(tb) * (tb + 4.0 * dtv)
                                           + PA(cir, p) * lw_grnd)
                             - PA(sa_internal, p) * emv * C_SB * p3(tb) * (tb + 4.0 * dtv)
                             + PA(lw_stem, p)
                             - (efsh + dc1 * PA(wtga, p) * dtv) - (PA(efe, p) + dc2 * wtgaq * PA(qsatldT, p) * dtv)
                             - (PA(cp_leaf, p) / dtime) * (PA(S->t_veg, p) - PA(tl_ini, p));
            }
            const double dtv = PA(dt_veg, p);
            efpot = rho * ((PA(S->elai, p) + PA(S->esai, p)) / rbp)
                    * (wtgaq * (PA(qsatl, p) + PA(qsatldT, p) * dtv) - wtgq0 * CA(S->qg, c) - PA(wtaq0, p) * CA(S->forc_q, c));
            PA(S->qflx_evap_veg, p) = rpp * efpot;
            double ecidif = 0.0;
            if (!hydrstress) {   /* (:1353-1367) */
                if (efpot > 0.0 && PA(S->btran, p) > btran0) PA(S->qflx_tran_veg, p) = efpot * rppdry;
                else PA(S->qflx_tran_veg, p) = 0.0;
            }
            ecidif = fmax90(0.0, PA(S->qflx_evap_veg, p) - PA(S->qflx_tran_veg, p) - h2ocan / dtime);
            PA(S->qflx_evap_veg, p) = fmin90(PA(S->qflx_evap_veg, p), PA(S->qflx_tran_veg, p) + h2ocan / dtime);
            PA(S->eflx_sh_veg, p) = efsh + dc1 * PA(wtga, p) * dtv + PA(err, p) + erre + C_HVAP * ecidif;
            PA(S->eflx_sh_stem, p) = PA(S->eflx_sh_stem, p) + rho * C_CPAIR * wtstem * (-PA(wtl0, p) * dtv);
            {
                const double tb = PA(tlbef, p);
                PA(lw_leaf, p) = PA(sa_internal, p) * emv * C_SB * p3(tb) * (tb + 4.0 * dtv);
            }
            oracle_qsat(PA(S->t_veg, p), CA(S->forc_pbot, c), &PA(qsatl, p), &PA(el, p), &PA(qsatldT, p));
            PA(S->taf, p) = wtg0 * tg + PA(wta0, p) * thm + PA(wtl0, p) * PA(S->t_veg, p) + PA(wtstem0, p) * ts;
            PA(S->qaf, p) = PA(wtlq0, p) * PA(qsatl, p) + wtgq0 * CA(S->qg, c) + CA(S->forc_q, c) * PA(wtaq0, p);
            PA(dth, p) = thm - PA(S->taf, p);
            PA(dqh, p) = CA(S->forc_q, c) - PA(S->qaf, p);
            PA(delq, p) = PA(wtalq, p) * CA(S->qg, c) - PA(wtlq0, p) * PA(qsatl, p) - PA(wtaq0, p) * CA(S->forc_q, c);
            const double tstar = PA(temp1, p) * PA(dth, p);
            const double qstar = PA(temp2, p) * PA(dqh, p);
            const double thvstar = tstar * (1.0 + 0.61 * CA(S->forc_q, c)) + 0.61 * CA(S->forc_th, c) * qstar;
            const double ustar = PA(S->ustar, p);
            PA(S->zeta, p) = PA(zldis, p) * C_VKC * C_GRAV * thvstar / ((ustar * ustar) * CA(S->thv, c));
            if (PA(S->zeta, p) >= 0.0) {
                PA(S->zeta, p) = fmin90(P->zetamaxstable, fmax90(PA(S->zeta, p), 0.01));
                PA(S->um, p) = fmax90(PA(ur, p), 0.1);
            } else {
                PA(S->zeta, p) = fmax90(-100.0, fmin90(PA(S->zeta, p), -0.01));
                double wc;
                if (ustar * thvstar > 0.0) wc = 0.0;
                else wc = beta * cm_pow(-C_GRAV * ustar * thvstar * zii / CA(S->thv, c), 0.333);
                PA(S->um, p) = sqrt(PA(ur, p) * PA(ur, p) + wc * wc);
            }
            PA(S->obu, p) = PA(zldis, p) / PA(S->zeta, p);
            if (PA(obuold, p) * PA(S->obu, p) < 0.0) nmozsgn[p - L.begp] = nmozsgn[p - L.begp] + 1;
            if (nmozsgn[p - L.begp] >= 4) PA(S->obu, p) = PA(zldis, p) / (-0.01);
            PA(obuold, p) = PA(S->obu, p);
        }
        itlef = itlef + 1;
        if (itlef > itmin) {
            for (f = 0; f < fn; ++f) {
                p = filterp[f];
                PA(dele, p) = fabs(PA(efe, p) - PA(efeb, p));
                PA(efeb, p) = PA(efe, p);
                PA(det, p) = fmax90(PA(del, p), PA(del2, p));
                PA(S->num_iter, p) = (double)itlef;
            }
            const int fnold = fn;
            fn = 0;
            for (f = 0; f < fnold; ++f) {
                p = filterp[f];
                if (!(PA(det, p) < dtmin && PA(dele, p) < dlemin)) {
                    filterp[fn] = p;
                    fn = fn + 1;
                }
            }
        }
    }

    fn = fnorig;
    for (f = 0; f < fn; ++f) filterp[f] = fporig[f];

    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        c = PA(S->column, p);
        g = PA(S->gridcell, p);
        const double rho = CA(S->forc_rho, c);
        const int snl = CA(col->snl, c);
        const double lw_grnd = (CA(col->frac_sno_eff, c) * p4(SNOC(col->t_soisno, c, snl + 1))
                                + (1.0 - CA(col->frac_sno_eff, c) - CA(col->frac_h2osfc, c)) * p4(SNOC(col->t_soisno, c, 1))
                                + CA(col->frac_h2osfc, c) * p4(CA(col->t_h2osfc, c)));
        const double fr = PA(frac_rad_abs_by_stem, p);
        const double tb = PA(tlbef, p), dtv = PA(dt_veg, p), emv = PA(S->emv, p), emg = CA(col->emg, c);
        PA(err, p) = (1.0 - fr) * (PA(S->sabv, p) + PA(air, p) + PA(bir, p) * p3(tb) * (tb + 4.0 * dtv)
                                   + PA(cir, p) * lw_grnd)
                     - PA(lw_leaf, p) + PA(lw_stem, p) - PA(S->eflx_sh_veg, p) - C_HVAP * PA(S->qflx_evap_veg, p)
                     - ((PA(S->t_veg, p) - PA(tl_ini, p)) * PA(cp_leaf, p) / dtime);
        if (cfg->use_biomass_heat_storage) {
            if (PA(S->stem_biomass, p) > 0.0) {
                const double ti = PA(ts_ini, p);
                PA(dt_stem, p) = (fr * (PA(S->sabv, p) + PA(air, p) + PA(bir, p) * p4(ti) + PA(cir, p) * lw_grnd)
                                  - PA(S->eflx_sh_stem, p) + PA(lw_leaf, p) - PA(lw_stem, p))
                                 / (PA(cp_stem, p) / dtime - fr * PA(bir, p) * 4.0 * p3(ti));
            } else {
                PA(dt_stem, p) = 0.0;
            }
            PA(S->dhsdt_canopy, p) = PA(dt_stem, p) * PA(cp_stem, p) / dtime
                                     + (PA(S->t_veg, p) - PA(tl_ini, p)) * PA(cp_leaf, p) / dtime;
            PA(S->t_stem, p) = PA(S->t_stem, p) + PA(dt_stem, p);
        } else {
            PA(dt_stem, p) = 0.0;
        }
        const double tv = PA(S->t_veg, p), ts = PA(S->t_stem, p), thm = PA(S->thm, p);
        const double delt = PA(wtal, p) * CA(col->t_grnd, c) - PA(wtl0, p) * tv - PA(wta0, p) * thm - PA(wtstem0, p) * ts;
        PA(S->taux, p) = -rho * GA(S->forc_u, g) / PA(S->ram1, p);
        PA(S->tauy, p) = -rho * GA(S->forc_v, g) / PA(S->ram1, p);
        PA(S->eflx_sh_grnd, p) = C_CPAIR * rho * PA(wtg, p) * delt;
        const double delt_snow = PA(wtal, p) * SNOC(col->t_soisno, c, snl + 1) - PA(wtl0, p) * tv - PA(wta0, p) * thm
                                 - PA(wtstem0, p) * ts;
        const double delt_soil = PA(wtal, p) * SNOC(col->t_soisno, c, 1) - PA(wtl0, p) * tv - PA(wta0, p) * thm
                                 - PA(wtstem0, p) * ts;
        const double delt_h2osfc = PA(wtal, p) * CA(col->t_h2osfc, c) - PA(wtl0, p) * tv - PA(wta0, p) * thm
                                   - PA(wtstem0, p) * ts;
        PA(S->eflx_sh_snow, p) = C_CPAIR * rho * PA(wtg, p) * delt_snow;
        PA(S->eflx_sh_soil, p) = C_CPAIR * rho * PA(wtg, p) * delt_soil;
        PA(S->eflx_sh_h2osfc, p) = C_CPAIR * rho * PA(wtg, p) * delt_h2osfc;
        PA(S->qflx_evap_soi, p) = rho * PA(wtgq, p) * PA(delq, p);
        const double fq = CA(S->forc_q, c);
        const double delq_snow = PA(wtalq, p) * CA(S->qg_snow, c) - PA(wtlq0, p) * PA(qsatl, p) - PA(wtaq0, p) * fq;
        PA(S->qflx_ev_snow, p) = rho * PA(wtgq, p) * delq_snow;
        const double delq_soil = PA(wtalq, p) * CA(S->qg_soil, c) - PA(wtlq0, p) * PA(qsatl, p) - PA(wtaq0, p) * fq;
        PA(S->qflx_ev_soil, p) = rho * PA(wtgq, p) * delq_soil;
        const double delq_h2osfc = PA(wtalq, p) * CA(S->qg_h2osfc, c) - PA(wtlq0, p) * PA(qsatl, p) - PA(wtaq0, p) * fq;
        PA(S->qflx_ev_h2osfc, p) = rho * PA(wtgq, p) * delq_h2osfc;
        PA(S->t_ref2m, p) = thm + PA(temp1, p) * PA(dth, p) * (1.0 / PA(temp12m, p) - 1.0 / PA(temp1, p));
        PA(S->t_ref2m_r, p) = PA(S->t_ref2m, p);
        PA(S->q_ref2m, p) = fq + PA(temp2, p) * PA(dqh, p) * (1.0 / PA(temp22m, p) - 1.0 / PA(temp2, p));
        double qsat_ref2m, e_ref2m;
        oracle_qsat(PA(S->t_ref2m, p), CA(S->forc_pbot, c), &qsat_ref2m, &e_ref2m, NULL);
        PA(S->rh_ref2m, p) = fmin90(100.0, PA(S->q_ref2m, p) / qsat_ref2m * 100.0);
        PA(S->rh_ref2m_r, p) = PA(S->rh_ref2m, p);
        PA(S->vpd_ref2m, p) = e_ref2m * (1.0 - PA(S->rh_ref2m, p) / 100.0);
        const double flw = CA(col->forc_lwrad, c);
        const double ti = PA(ts_ini, p), dts = PA(dt_stem, p);
        PA(S->dlrad, p) = (1.0 - emv) * emg * flw
                          + emv * emg * C_SB * p3(tb) * (tb + 4.0 * dtv) * (1.0 - fr)
                          + emv * emg * C_SB * p3(ti) * (ti + 4.0 * dts) * fr;
        PA(S->ulrad, p) = ((1.0 - emg) * (1.0 - emv) * (1.0 - emv) * flw
                           + emv * (1.0 + (1.0 - emg) * (1.0 - emv)) * C_SB * p3(tb) * (tb + 4.0 * dtv) * (1.0 - fr)
                           + emv * (1.0 + (1.0 - emg) * (1.0 - emv)) * C_SB * p3(ti) * (ti + 4.0 * dts) * fr
                           + emg * (1.0 - emv) * C_SB * lw_grnd);
        PA(S->t_skin, p) = emv * tv + (1.0 - emv) * sqrt(sqrt(lw_grnd));
        PA(S->cgrnds, p) = PA(S->cgrnds, p) + C_CPAIR * rho * PA(wtg, p) * PA(wtal, p);
        PA(S->cgrndl, p) = PA(S->cgrndl, p) + rho * PA(wtgq, p) * PA(wtalq, p) * CA(S->dqgdT, c);
        PA(S->cgrnd, p) = PA(S->cgrnds, p) + PA(S->cgrndl, p) * CA(col->htvp, c);
        PA(snocan_baseline, p) = PA(S->snocan, p);
        const double qe_minus_qt = PA(S->qflx_evap_veg, p) - PA(S->qflx_tran_veg, p);
        if (tv > C_TFRZ) {
            if (qe_minus_qt * dtime > PA(S->liqcan, p)) {
                PA(S->snocan, p) = fmax90(0.0, PA(S->snocan, p) + PA(S->liqcan, p)
                                                   + (PA(S->qflx_tran_veg, p) - PA(S->qflx_evap_veg, p)) * dtime);
            }
            PA(S->liqcan, p) = fmax90(0.0, PA(S->liqcan, p) + (PA(S->qflx_tran_veg, p) - PA(S->qflx_evap_veg, p)) * dtime);
        } else if (tv <= C_TFRZ) {
            if (qe_minus_qt * dtime > PA(S->snocan, p)) {
                PA(S->liqcan, p) = PA(S->liqcan, p) + PA(S->snocan, p)
                                   + (PA(S->qflx_tran_veg, p) - PA(S->qflx_evap_veg, p)) * dtime;
            }
            PA(S->snocan, p) = fmax90(0.0, PA(S->snocan, p) + (PA(S->qflx_tran_veg, p) - PA(S->qflx_evap_veg, p)) * dtime);
        }
    }
    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        if (fabs(PA(S->snocan, p)) < 1.e-10 * fabs(PA(snocan_baseline, p))) PA(S->snocan, p) = 0.0;
    }
    /* PhotosynthesisTotal and the noon water-use efficiency (:1616-1640) */
    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        g = PA(S->gridcell, p);
        const double ls = PA(S->laisun, p), lh = PA(S->laisha, p);
        PA(S->fpsn, p) = PA(S->psnsun, p) * ls + PA(S->psnsha, p) * lh;
        PA(S->fpsn_wc, p) = PA(S->psnsun_wc, p) * ls + PA(S->psnsha_wc, p) * lh;
        PA(S->fpsn_wj, p) = PA(S->psnsun_wj, p) * ls + PA(S->psnsha_wj, p) * lh;
        PA(S->fpsn_wp, p) = PA(S->psnsun_wp, p) * ls + PA(S->psnsha_wp, p) * lh;
    }
    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        g = PA(S->gridcell, p);
        if (GA(S->near_local_noon, g) && PA(S->fpsn, p) > 0.0) {
            const double gs = 1.e-6 * (PA(S->laisun, p) * PA(S->gs_mol_sun, p) + PA(S->laisha, p) * PA(S->gs_mol_sha, p));
            if (gs > 0.0) PA(S->iwue_ln, p) = PA(S->fpsn, p) / gs;
            else PA(S->iwue_ln, p) = C_SPVAL;
        } else {
            PA(S->iwue_ln, p) = C_SPVAL;
        }
    }
    for (f = 0; f < fn; ++f) {
        p = filterp[f];
        PA(S->canopy_err, p) = PA(err, p);
    }

done:
    free(zldis); free(dth); free(dthv); free(dqh); free(ur); free(temp1); free(temp12m); free(temp2); free(temp22m);
    free(rb); free(rah1); free(rah2); free(raw1); free(raw2); free(wtg); free(wta0); free(wtl0); free(wtstem0);
    free(wtal); free(wtga); free(wtgq); free(wtaq0); free(wtlq0); free(wtalq); free(el); free(qsatl); free(qsatldT);
    free(air); free(bir); free(cir); free(delq); free(del); free(del2); free(dele); free(det); free(efeb); free(efe);
    free(obuold); free(tlbef); free(tl_ini); free(ts_ini); free(err); free(co2); free(o2); free(svpts); free(eah);
    free(fm); free(dayl_factor); free(dt_veg); free(dbh); free(cp_leaf); free(cp_stem); free(rstem); free(dt_stem);
    free(frac_rad_abs_by_stem); free(lw_stem); free(lw_leaf); free(sa_stem); free(sa_leaf); free(sa_internal);
    free(uuc); free(snocan_baseline); free(nmozsgn); free(filterp); free(fporig);
    return rc;
}

/* CanopyFluxesMod.F90:191-1765 with use_hydrstress = .false. */
int oracle_canopyfluxes(const ctsm_config *cfg, const ctsm_canopy_params *prm, const ctsm_pft_params *pft,
                        const ctsm_layout *lay, const ctsm_col_state *col, const ctsm_canopy_state *S,
                        int num_exposedvegp, const int *filter_exposedvegp, double dtime, int *bad_patch)
{
    return canopyfluxes(cfg, prm, pft, lay, col, S, NULL, NULL, NULL, num_exposedvegp, filter_exposedvegp, dtime,
                        bad_patch);
}

/* CanopyFluxesMod.F90:191-1765 with use_hydrstress = .true.: the root-soil
 * conductances of the filter's patches (PhotosynthesisMod.F90:3062-3114, the
 * loop PhotosynthesisHydraulicStress runs on every call over inputs the
 * iteration leaves alone) once, then the iteration with
 * PhotosynthesisHydraulicStress.  Returns 1 (bad_patch 0) for a patch whose
 * column or type is out of range in the conductance loop. */
int oracle_canopyfluxes_phs(const ctsm_config *cfg, const ctsm_canopy_params *prm, const ctsm_pft_params *pft,
                            const ctsm_phs_params *pp, const ctsm_layout *lay, const ctsm_col_state *col,
                            const ctsm_canopy_state *S, const ctsm_phs_state *phs, const ctsm_phs_canopy_state *phc,
                            int num_exposedvegp, const int *filter_exposedvegp, double dtime, int *bad_patch)
{
    *bad_patch = 0;
    if (oracle_phs_soil_root_conductance(cfg, lay, phs, pp, num_exposedvegp, filter_exposedvegp)) return 1;
    return canopyfluxes(cfg, prm, pft, lay, col, S, pp, phs, phc, num_exposedvegp, filter_exposedvegp, dtime,
                        bad_patch);
}

/* clm_driver.F90:1714-1716 (p2c_1d_filter of qflx_tran_veg) followed by
 * Compute_EffecRootFrac_And_VertTranSink_Default (SoilWaterPlantSinkMod.F90:332-424). */
void oracle_root_sink(const ctsm_config *cfg, const ctsm_layout *lay, const ctsm_col_state *col,
                      const ctsm_canopy_state *S, int num_filterc, const int *filterc)
{
    cl_t L;
    L.begc = lay->begc_alloc;
    L.ldc = lay->ldc;
    L.begp = lay->begp_alloc;
    L.ldp = lay->ldp;
    L.begg = lay->begg_alloc;
    L.nlevsno = cfg->nlevsno;
    L.nlevgrnd = cfg->nlevgrnd;
    const int nlevsoi = cfg->nlevsoi;
    int fc, c, p, j;
    double *temp = (double *)calloc((size_t)L.ldc, sizeof(double));
    for (fc = 0; fc < num_filterc; ++fc) {
        c = filterc[fc];
        CA(S->qflx_tran_veg_col, c) = 0.0;
        for (p = CA(col->patchi, c); p <= CA(col->patchf, c); ++p)
            if (PA(S->active, p)) CA(S->qflx_tran_veg_col, c) = CA(S->qflx_tran_veg_col, c) + PA(S->qflx_tran_veg, p) * PA(S->wtcol, p);
    }
    for (j = 1; j <= nlevsoi; ++j)
        for (fc = 0; fc < num_filterc; ++fc) SOIC(S->rootr_col, filterc[fc], j) = 0.0;
    for (j = 1; j <= nlevsoi; ++j) {
        for (fc = 0; fc < num_filterc; ++fc) {
            c = filterc[fc];
            for (p = CA(col->patchi, c); p <= CA(col->patchf, c); ++p) {
                if (PA(S->active, p))
                    SOIC(S->rootr_col, c, j) = SOIC(S->rootr_col, c, j)
                                               + SOIP(S->rootr, p, j) * PA(S->qflx_tran_veg, p) * PA(S->wtcol, p);
            }
        }
    }
    for (fc = 0; fc < num_filterc; ++fc) {
        c = filterc[fc];
        for (p = CA(col->patchi, c); p <= CA(col->patchf, c); ++p)
            if (PA(S->active, p)) CA(temp, c) = CA(temp, c) + PA(S->qflx_tran_veg, p) * PA(S->wtcol, p);
    }
    for (j = 1; j <= nlevsoi; ++j) {
        for (fc = 0; fc < num_filterc; ++fc) {
            c = filterc[fc];
            if (CA(temp, c) != 0.0) SOIC(S->rootr_col, c, j) = SOIC(S->rootr_col, c, j) / CA(temp, c);
            SOIC(S->qflx_rootsoi_col, c, j) = SOIC(S->rootr_col, c, j) * CA(S->qflx_tran_veg_col, c);
        }
    }
    free(temp);
}

/* vector wrappers used by the pin tests */
void oracle_qsat_vec(int n, const double *T, const double *p, double *qs, double *es, double *qsdT)
{
    int i;
    for (i = 0; i < n; ++i) oracle_qsat(T[i], p[i], &qs[i], &es[i], &qsdT[i]);
}
void oracle_frictionvelocity_vec(int n, const double *hu, const double *ht, const double *hq, const double *displa,
                                 const double *z0m, const double *z0h, const double *z0q, const double *obu,
                                 const double *ur, const double *um, double *out11)
{
    int i;
    for (i = 0; i < n; ++i)
        oracle_frictionvelocity(hu[i], ht[i], hq[i], displa[i], z0m[i], z0h[i], z0q[i], obu[i], 1, ur[i], um[i], 0.0,
                                out11 + 11 * (long)i);
}
