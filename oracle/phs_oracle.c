/*
 * phs_oracle.c -- CPU ORACLE (test infrastructure only) of the soil side of
 * plant hydraulic stress (use_hydrstress = .true.).
 *
 * Never linked into or called from the product library; tests/ use it as the
 * checker of the HIP kernels.  Same conventions as ctsm_oracle.c: binary64,
 * -ffp-contract=off, the Fortran's operation order, pow from ctsm_math.h.
 *
 *   oracle_phs_plc      plc, vegetation_weibull branch
 *                       (PhotosynthesisMod.F90:5167-5195)
 *   oracle_phs_d1plc    d1plc (:5199-5228); ln 2 as the correctly rounded
 *                       constant a compiler folds log(2._r8) to
 *   oracle_phs_soil_root_conductance
 *                       the root-soil interface conductance loop at the head of
 *                       PhotosynthesisHydraulicStress (:3062-3114)
 *   oracle_phs_root_sink
 *                       Compute_EffecRootFrac_And_VertTranSink_HydStress
 *                       (SoilWaterPlantSinkMod.F90:236-328), the reference's
 *                       loops as written: columns, then layers, then patches
 *   oracle_phs_calcstress
 *                       calcstress (:4490-4710) per patch of the filter, with
 *                       getqflx (:5080-5164), spacF (:4898-4976), spacA
 *                       (:4715-4893) and getvegwp (:4979-5077);
 *                       oracle_phs_spac exposes spacF / spacA to the tests
 *
 * Pins (tests/test_oracle_phs.py): plc and d1plc against the known answers of
 * the reference's test_Photosynthesis.pf (x = -25000 mm, psi50 = -150000 mm,
 * ck = 3.95, tolerance 1e-13), plc(0) = 1, plc(psi50) = 1/2, the 0.005 cut,
 * d1plc against a central difference of plc; the conductances against the
 * Gardner closed form for a root length density of 1/pi (r_soil = 1 m) and
 * against a numpy evaluation with the series law 1/k = 1/k_soil + 1/k_root, the
 * top layer and rootless layers at zero, the 1 gC/m2 biomass floor, the 1e-16
 * floors; the sink against a numpy sum over patches, zero exchange at
 * hydrostatic equilibrium (smp = vegwp(root) + 1000 z), and the exclusion of
 * inactive, snow-buried and zero-weight patches; calcstress without
 * transpiration against getvegwp's closed form (the conductance-weighted soil
 * potential, the canopy height below it), after Newton the root uptake equal
 * to the attenuated transpiration and the stress factors as the conductances
 * that carry it (the forward Penman form getqflx inverts), spacA times a
 * central-difference Jacobian of spacF equal to minus the identity (both the
 * 4x4 and the shaded-leaf-free 3x3 forms), the night outputs.
 */
#include <math.h>
#include <stdlib.h>
#include "ctsm_gpu.h"
#include "ctsm_math.h"
#include "oracle_libm.h"   /* -DORACLE_LIBM: the host libm instead */
#include "phs_oracle.h"

#define P_PI 3.14159265358979323846
#define P_LN2 0.6931471805599453
#define P_C_TO_B 2.0
#define P_CROOT 0.25

typedef struct {
    int begc, ldc, begp, ldp, nlevsno;
} p_lay_t;

#define PC(a, c)       ((a)[(c) - Lo.begc])
#define PP(a, p)       ((a)[(p) - Lo.begp])
#define PSOI(a, c, j)  ((a)[((c) - Lo.begc) + (long)((j) - 1) * Lo.ldc])
#define PSNO(a, c, j)  ((a)[((c) - Lo.begc) + (long)((j) + Lo.nlevsno - 1) * Lo.ldc])
#define PSOIP(a, p, j) ((a)[((p) - Lo.begp) + (long)((j) - 1) * Lo.ldp])
#define PVEG(a, p, k)  ((a)[((p) - Lo.begp) + (long)((k) - 1) * Lo.ldp])

static double p_max(double a, double b) { return (b > a) ? b : a; }
static double p_min(double a, double b) { return (b < a) ? b : a; }

double oracle_phs_plc(double x, double psi50, double ck)
{
    double v = cm_pow(2.0, -cm_pow(x / psi50, ck));
    if (v < 0.005) v = 0.0;
    return v;
}

double oracle_phs_d1plc(double x, double psi50, double ck)
{
    return -ck * P_LN2 * cm_pow(2.0, -cm_pow(x / psi50, ck)) * cm_pow(x / psi50, ck) / x;
}

static p_lay_t p_lay(const ctsm_config *cfg, const ctsm_layout *L)
{
    p_lay_t Lo = {L->begc_alloc, L->ldc, L->begp_alloc, L->ldp, cfg->nlevsno};
    return Lo;
}

/* returns 0, or 1 when a patch's column or type is out of range */
int oracle_phs_soil_root_conductance(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_phs_state *S,
                                     const ctsm_phs_params *P, int n, const int32_t *filter)
{
    const p_lay_t Lo = p_lay(cfg, L);
    const int np = P->npft;
    for (int f = 0; f < n; ++f) {
        const int p = filter[f];
        const int c = PP(S->column, p), ivt = PP(S->itype, p);
        if (c < Lo.begc || c >= Lo.begc + Lo.ldc || ivt < 0 || ivt >= np) return 1;
        for (int j = 1; j <= cfg->nlevsoi; ++j) {
            double root_biomass_density = P_C_TO_B * PP(S->froot_carbon, p) * PSOIP(S->rootfr, p, j)
                                          / PSNO(S->dz, c, j);
            root_biomass_density = p_max(P_C_TO_B * 1.0, root_biomass_density);
            const double root_cross_sec_area = P_PI * (P->root_radius[ivt] * P->root_radius[ivt]);
            const double root_length_density = root_biomass_density / (P->root_density[ivt] * root_cross_sec_area);
            const double rai = (PP(S->tsai, p) + PP(S->tlai, p)) * P->froot_leaf[ivt] * PSOIP(S->rootfr, p, j);
            const double r_soil = sqrt(1.0 / (P_PI * root_length_density));
            double soil_conductance = p_min(PSOI(S->hksat, c, j), PSOI(S->hk_l, c, j)) / (1.e3 * r_soil);
            const double fs = oracle_phs_plc(PSOI(S->smp_l, c, j), P->psi50[3 * np + ivt], P->ck[3 * np + ivt]);
            double root_conductance = (fs * rai * P->krmax[ivt]) / (P_CROOT + PSNO(S->z, c, j));
            soil_conductance = p_max(soil_conductance, 1.e-16);
            root_conductance = p_max(root_conductance, 1.e-16);
            PSOIP(S->root_conductance, p, j) = root_conductance;
            PSOIP(S->soil_conductance, p, j) = soil_conductance;
            const double rs_resis = 1.0 / soil_conductance + 1.0 / root_conductance;
            if (rai * PSOIP(S->rootfr, p, j) > 0.0 && j > 1)
                PSOIP(S->k_soil_root, p, j) = 1.0 / rs_resis;
            else
                PSOIP(S->k_soil_root, p, j) = 0.0;
        }
    }
    return 0;
}

int oracle_phs_root_sink(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_phs_state *S, int n,
                         const int32_t *filter)
{
    const p_lay_t Lo = p_lay(cfg, L);
    for (int f = 0; f < n; ++f) {
        const int c = filter[f];
        PC(S->qflx_phs_neg, c) = 0.0;
        for (int j = 1; j <= cfg->nlevsoi; ++j) {
            const double grav2 = PSNO(S->z, c, j) * 1000.0;
            double temp = 0.0;
            for (int p = PC(S->patchi, c); p <= PC(S->patchf, c); ++p) {
                if (j == 1) PP(S->qflx_hydr_redist, p) = 0.0;
                if (PP(S->active, p) && PP(S->frac_veg_nosno, p) > 0) {
                    if (PP(S->wtcol, p) > 0.0) {
                        const double patchflux = PSOIP(S->k_soil_root, p, j)
                                                 * (PSOI(S->smp_l, c, j) - PVEG(S->vegwp, p, 4) - grav2);
                        if (patchflux < 0.0) PP(S->qflx_hydr_redist, p) = PP(S->qflx_hydr_redist, p) + patchflux;
                        temp = temp + patchflux * PP(S->wtcol, p);
                    }
                }
            }
            PSOI(S->qflx_rootsoi, c, j) = temp;
            if (temp < 0.0) PC(S->qflx_phs_neg, c) = PC(S->qflx_phs_neg, c) + temp;
        }
    }
    return 0;
}

/* ---- calcstress (PhotosynthesisMod.F90:4490-4710) and its helpers ----------
 * Segments 1 sun, 2 sha, 3 xyl, 4 root; arrays below are 0-based (x[0] = sun). */
#define P_RGAS (6.02214e26 * 1.38065e-23)
#define P_SPVAL 1.e36
#define P_TOL_LAI 0.001

static double p_plc(const p_stress_t *P, int seg, double x) { return oracle_phs_plc(x, P->psi50[seg], P->ck[seg]); }
static double p_d1plc(const p_stress_t *P, int seg, double x) { return oracle_phs_d1plc(x, P->psi50[seg], P->ck[seg]); }

/* getqflx (:5080-5164): havegs = 1 conductances -> fluxes, 0 fluxes -> conductances */
static void p_getqflx(const p_stress_t *P, double qsatl, double qaf, int havegs, double *gs_sun, double *gs_sha,
                      double *q_sun, double *q_sha)
{
    const double cf = P->pbot / (P_RGAS * 1.e-3 * P->thm) * 1.e6;
    const double wtl = (P->elai + P->esai) * P->gb;
    const double efpot = P->rho * wtl * (qsatl - qaf);
    if (havegs) {
        if (efpot > 0.0 && P->elai > 0.0) {
            if (*gs_sun > 0.0) {
                const double rpp = P->fdry / P->gb * (P->laisun / (1.0 / P->gb + 1.0 / *gs_sun)) / P->elai;
                *q_sun = efpot * rpp / cf;
            } else {
                *q_sun = 0.0;
            }
            if (*gs_sha > 0.0) {
                const double rpp = P->fdry / P->gb * (P->laisha / (1.0 / P->gb + 1.0 / *gs_sha)) / P->elai;
                *q_sha = efpot * rpp / cf;
            } else {
                *q_sha = 0.0;
            }
        } else {
            *q_sun = 0.0;
            *q_sha = 0.0;
        }
    } else {
        if (*q_sun > 0.0)
            *gs_sun = P->gb * *q_sun * cf * P->elai / (efpot * P->fdry * P->laisun - *q_sun * cf * P->elai);
        else
            *gs_sun = 0.0;
        if (*q_sha > 0.0)
            *gs_sha = P->gb * *q_sha * cf * P->elai / (efpot * P->fdry * P->laisha - *q_sha * cf * P->elai);
        else
            *gs_sha = 0.0;
    }
}

/* spacF (:4898-4976) */
static void p_spacF(const p_stress_t *P, const double *x, double *f, double q_sun, double q_sha)
{
    const double grav1 = P->htop * 1000.0;
    const double fsto1 = p_plc(P, 0, x[0]), fsto2 = p_plc(P, 1, x[1]);
    const double fx = p_plc(P, 2, x[2]), fr = p_plc(P, 3, x[3]);
    f[0] = q_sun * fsto1 - P->laisun * P->kmax[0] * fx * (x[2] - x[0]);
    f[1] = q_sha * fsto2 - P->laisha * P->kmax[1] * fx * (x[2] - x[1]);
    f[2] = P->laisun * P->kmax[0] * fx * (x[2] - x[0]) + P->laisha * P->kmax[1] * fx * (x[2] - x[1])
           - P->tsai * P->kmax[2] / P->htop * fr * (x[3] - x[2] - grav1);
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < P->n; ++j) s1 = s1 + P->k[j * P->sk] * (x[3] + P->z[j * P->ss] * 1000.0);
    for (int j = 0; j < P->n; ++j) s2 = s2 + P->k[j * P->sk] * P->smp[j * P->ss];
    f[3] = P->tsai * P->kmax[2] / P->htop * fr * (x[3] - x[2] - grav1) + s1 - s2;
    if (P->laisha < P_TOL_LAI) {
        const double t = f[0];
        f[0] = f[1];
        f[1] = t;
    }
}

/* spacA (:4715-4893); returns 1 when the matrix is not invertible.  A and
 * invA are row-major 4x4: A[(i-1)*4 + (j-1)] = A(i,j). */
static int p_spacA(const p_stress_t *P, const double *x, double *iA, double q_sun, double q_sha)
{
    double A[16];
    for (int i = 0; i < 16; ++i) A[i] = 0.0, iA[i] = 0.0;
#define AA(i, j) A[((i) - 1) * 4 + (j) - 1]
#define IA(i, j) iA[((i) - 1) * 4 + (j) - 1]
    const double grav1 = P->htop * 1000.0;
    const double fsto1 = p_plc(P, 0, x[0]), fsto2 = p_plc(P, 1, x[1]);
    const double fx = p_plc(P, 2, x[2]), fr = p_plc(P, 3, x[3]);
    const double dfsto1 = p_d1plc(P, 0, x[0]), dfsto2 = p_d1plc(P, 1, x[1]);
    const double dfx = p_d1plc(P, 2, x[2]), dfr = p_d1plc(P, 3, x[3]);
    const double ks = P->kmax[0], kh = P->kmax[1], kx = P->kmax[2];
    AA(1, 1) = -(P->laisun * ks * fx) - q_sun * dfsto1;
    AA(1, 3) = P->laisun * ks * dfx * (x[2] - x[0]) + P->laisun * ks * fx;
    AA(2, 2) = -(P->laisha * kh * fx) - q_sha * dfsto2;
    AA(2, 3) = P->laisha * kh * dfx * (x[2] - x[1]) + P->laisha * kh * fx;
    AA(3, 1) = P->laisun * ks * fx;
    AA(3, 2) = P->laisha * kh * fx;
    AA(3, 3) = -(P->laisun * ks * dfx * (x[2] - x[0])) - P->laisun * ks * fx - P->laisha * kh * dfx * (x[2] - x[1])
               - P->laisha * kh * fx - P->tsai * kx / P->htop * fr;
    AA(3, 4) = P->tsai * kx / P->htop * dfr * (x[3] - x[2] - grav1) + P->tsai * kx / P->htop * fr;
    AA(4, 3) = P->tsai * kx / P->htop * fr;
    double sk = 0.0;
    for (int j = 0; j < P->n; ++j) sk = sk + P->k[j * P->sk];
    AA(4, 4) = -(P->tsai * kx / P->htop * fr) - P->tsai * kx / P->htop * dfr * (x[3] - x[2] - grav1) - sk;
    if (P->laisun > P_TOL_LAI && P->laisha > P_TOL_LAI) {
        const double determ = AA(4, 4) * AA(2, 2) * AA(3, 3) * AA(1, 1) - AA(4, 4) * AA(2, 2) * AA(3, 1) * AA(1, 3)
                              - AA(4, 4) * AA(3, 2) * AA(2, 3) * AA(1, 1) - AA(4, 3) * AA(1, 1) * AA(2, 2) * AA(3, 4);
        if (fabs(determ) <= 1.e-50) return 1;
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
        if (P->laisha <= P_TOL_LAI) {
            AA(2, 2) = AA(1, 1);
            AA(3, 2) = AA(3, 1);
            AA(2, 3) = AA(1, 3);
        }
        const double determ = AA(2, 2) * AA(3, 3) * AA(4, 4) - AA(3, 4) * AA(2, 2) * AA(4, 3)
                              - AA(2, 3) * AA(3, 2) * AA(4, 4);
        if (fabs(determ) <= 1.e-50) return 1;
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
        for (int i = 0; i < 16; ++i) iA[i] = r * iA[i];
    }
#undef AA
#undef IA
    return 0;
}

/* getvegwp (:4979-5077): x from the conductances; returns soilflux */
static double p_getvegwp(const p_stress_t *P, double *x, double gs_sun, double gs_sha, double qsatl, double qaf)
{
    const double grav1 = 1000.0 * P->htop;
    double q_sun, q_sha;
    p_getqflx(P, qsatl, qaf, 1, &gs_sun, &gs_sha, &q_sun, &q_sha);
    double sk = 0.0;
    for (int j = 0; j < P->n; ++j) sk = sk + P->k[j * P->sk];
    if (fabs(sk) == 0.0) {
        double s = 0.0;
        for (int j = 0; j < P->n; ++j) s = s + (P->smp[j * P->ss] - 1000.0 * P->z[j * P->ss]);
        x[3] = s / P->n;
    } else {
        double s = 0.0;
        for (int j = 0; j < P->n; ++j) s = s + P->k[j * P->sk] * (P->smp[j * P->ss] - 1000.0 * P->z[j * P->ss]);
        x[3] = (s - q_sun - q_sha) / sk;
    }
    const double fr = p_plc(P, 3, x[3]);
    if (P->tsai > 0.0 && fr > 0.0)
        x[2] = x[3] - grav1 - (q_sun + q_sha) / (fr * P->kmax[3] / P->htop * P->tsai);
    else
        x[2] = x[3] - grav1;
    const double fx = p_plc(P, 2, x[2]);
    if (P->laisha > 0.0 && fx > 0.0)
        x[1] = x[2] - (q_sha / (fx * P->kmax[2] * P->laisha));
    else
        x[1] = x[2];
    if (P->laisun > 0.0 && fx > 0.0)
        x[0] = x[2] - (q_sun / (fx * P->kmax[2] * P->laisun));
    else
        x[0] = x[2];
    double soilflux = 0.0;
    for (int j = 0; j < P->n; ++j)
        soilflux = soilflux + P->k[j * P->sk] * (P->smp[j * P->ss] - x[3] - 1000.0 * P->z[j * P->ss]);
    return soilflux;
}

/* calcstress for one patch; x (4) inout, returns bsun / bsha, night outputs */
static void p_calcstress(const p_stress_t *P, double *x, double gs_sun_in, double gs_sha_in, double qsatl, double qaf,
                         double *bsun, double *bsha, int *night, double *qflx_tran_veg)
{
    *night = x[0] > 0.0;
    if (*night) x[0] = x[1];
    double gs0sun = gs_sun_in, gs0sha = gs_sha_in;
    double q_sun, q_sha;
    p_getqflx(P, qsatl, qaf, 1, &gs0sun, &gs0sha, &q_sun, &q_sha);
    int flag;
    if ((P->laisun > P_TOL_LAI || P->laisha > P_TOL_LAI) && (q_sun > 0.0 || q_sha > 0.0)) {
        int iter = 0;
        for (;;) {
            iter = iter + 1;
            double f[4], iA[16], dx[4];
            p_spacF(P, x, f, q_sun, q_sha);
            double ff = 0.0;
            for (int i = 0; i < 4; ++i) ff = ff + f[i] * f[i];
            if (sqrt(ff) < 1.e-6 * (q_sun + q_sha)) { flag = 0; break; }
            if (iter > 50) { flag = 0; break; }
            flag = p_spacA(P, x, iA, q_sun, q_sha);
            if (flag) break;
            if (P->laisun > P_TOL_LAI && P->laisha > P_TOL_LAI) {
                for (int i = 0; i < 4; ++i) {
                    double s = 0.0;
                    for (int k = 0; k < 4; ++k) s = s + iA[i * 4 + k] * f[k];
                    dx[i] = s;
                }
            } else {
                dx[0] = 0.0;
                for (int i = 1; i < 4; ++i) {
                    double s = 0.0;
                    for (int k = 1; k < 4; ++k) s = s + iA[i * 4 + k] * f[k];
                    dx[i] = s;
                }
            }
            double m = fabs(dx[0]);
            for (int i = 1; i < 4; ++i) if (fabs(dx[i]) > m) m = fabs(dx[i]);
            if (m > 50000.0)
                for (int i = 0; i < 4; ++i) dx[i] = 50000.0 * dx[i] / m;
            if (P->laisun > P_TOL_LAI && P->laisha > P_TOL_LAI) {
                for (int i = 0; i < 4; ++i) x[i] = x[i] + dx[i];
            } else if (P->laisha > P_TOL_LAI) {
                for (int i = 0; i < 4; ++i) x[i] = x[i] + dx[i];
                x[0] = x[2];
            } else {
                x[2] = x[2] + dx[2];
                x[3] = x[3] + dx[3];
                x[0] = x[0] + dx[1];
                x[1] = x[2];
            }
            double dd = 0.0;
            for (int i = 0; i < 4; ++i) dd = dd + dx[i] * dx[i];
            if (sqrt(dd) < 1.e-9) break;
            if (x[2] > x[3]) x[2] = x[3];
            if (x[0] > x[2]) x[0] = x[2];
            if (x[1] > x[2]) x[1] = x[2];
        }
    } else {
        flag = 1;
    }
    if (flag) {
        (void)p_getvegwp(P, x, gs0sun, gs0sha, qsatl, qaf);
        *bsun = p_plc(P, 0, x[0]);
        *bsha = p_plc(P, 1, x[1]);
    } else {
        double qs = q_sun * p_plc(P, 0, x[0]);
        double qh = q_sha * p_plc(P, 1, x[1]);
        p_getqflx(P, qsatl, qaf, 0, &gs0sun, &gs0sha, &qs, &qh);
        *bsun = (q_sun > 0.0) ? gs0sun / gs_sun_in : p_plc(P, 0, x[0]);
        *bsha = (q_sha > 0.0) ? gs0sha / gs_sha_in : p_plc(P, 1, x[1]);
    }
    if (*bsun < 0.01) *bsun = 0.0;
    if (*bsha < 0.01) *bsha = 0.0;
    if (*night) {
        gs0sun = *bsun * gs_sun_in;
        gs0sha = *bsha * gs_sha_in;
        double soilflux = p_getvegwp(P, x, gs0sun, gs0sha, qsatl, qaf);
        if (soilflux < 0.0) soilflux = 0.0;
        *qflx_tran_veg = soilflux;
    }
}

void oracle_phs_calcstress_one(const p_stress_t *P, double *x, double gs_sun_in, double gs_sha_in, double qsatl,
                               double qaf, double *bsun, double *bsha, int *night, double *qflx_tran_veg)
{
    p_calcstress(P, x, gs_sun_in, gs_sha_in, qsatl, qaf, bsun, bsha, night, qflx_tran_veg);
}

double oracle_phs_getvegwp_one(const p_stress_t *P, double *x, double gs_sun, double gs_sha, double qsatl,
                               double qaf)
{
    return p_getvegwp(P, x, gs_sun, gs_sha, qsatl, qaf);
}

/* returns 0, or 1 when a patch's column or type is out of range */
int oracle_phs_calcstress(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_phs_stress_state *S,
                          const ctsm_phs_params *Pp, int n, const int32_t *filter)
{
    const p_lay_t Lo = p_lay(cfg, L);
    const int np = Pp->npft;
    for (int f = 0; f < n; ++f) {
        const int p = filter[f];
        const int c = PP(S->column, p), ivt = PP(S->itype, p);
        if (c < Lo.begc || c >= Lo.begc + Lo.ldc || ivt < 0 || ivt >= np) return 1;
        p_stress_t P;
        P.laisun = PP(S->laisun, p); P.laisha = PP(S->laisha, p); P.elai = PP(S->elai, p);
        P.esai = PP(S->esai, p); P.tsai = PP(S->tsai, p); P.htop = PP(S->htop, p); P.fdry = PP(S->fdry, p);
        P.thm = PP(S->thm, p); P.rho = PC(S->forc_rho, c); P.pbot = PC(S->forc_pbot, c); P.gb = PP(S->gb_mol, p);
        for (int k = 0; k < 4; ++k) {
            P.kmax[k] = Pp->kmax[k * np + ivt];
            P.psi50[k] = Pp->psi50[k * np + ivt];
            P.ck[k] = Pp->ck[k * np + ivt];
        }
        P.k = &PSOIP(S->k_soil_root, p, 1);
        P.smp = &PSOI(S->smp_l, c, 1);
        P.z = &PSNO(S->z, c, 1);
        P.sk = Lo.ldp;
        P.ss = Lo.ldc;
        P.n = cfg->nlevsoi;
        double x[4];
        for (int k = 0; k < 4; ++k) x[k] = PVEG(S->vegwp, p, k + 1);
        double bsun, bsha, tran = 0.0;
        int night;
        p_calcstress(&P, x, PP(S->gs_mol_sun, p), PP(S->gs_mol_sha, p), PP(S->qsatl, p), PP(S->qaf, p), &bsun,
                     &bsha, &night, &tran);
        for (int k = 0; k < 4; ++k) {
            PVEG(S->vegwp, p, k + 1) = x[k];
            PVEG(S->vegwp_pd, p, k + 1) = (night && PP(S->local_morning, p)) ? x[k] : P_SPVAL;
        }
        PP(S->bsun, p) = bsun;
        PP(S->bsha, p) = bsha;
        if (night) PP(S->qflx_tran_veg, p) = tran;
    }
    return 0;
}

/* Test hook: spacF and spacA of patch p at x (4) with the given fluxes (for the
 * check that invA inverts the Jacobian of F); returns spacA's flag. */
int oracle_phs_spac(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_phs_stress_state *S,
                    const ctsm_phs_params *Pp, int p, const double *x, double q_sun, double q_sha, double *f,
                    double *iA)
{
    const p_lay_t Lo = p_lay(cfg, L);
    const int np = Pp->npft;
    const int c = PP(S->column, p), ivt = PP(S->itype, p);
    p_stress_t P;
    P.laisun = PP(S->laisun, p); P.laisha = PP(S->laisha, p); P.elai = PP(S->elai, p);
    P.esai = PP(S->esai, p); P.tsai = PP(S->tsai, p); P.htop = PP(S->htop, p); P.fdry = PP(S->fdry, p);
    P.thm = PP(S->thm, p); P.rho = PC(S->forc_rho, c); P.pbot = PC(S->forc_pbot, c); P.gb = PP(S->gb_mol, p);
    for (int k = 0; k < 4; ++k) {
        P.kmax[k] = Pp->kmax[k * np + ivt];
        P.psi50[k] = Pp->psi50[k * np + ivt];
        P.ck[k] = Pp->ck[k * np + ivt];
    }
    P.k = &PSOIP(S->k_soil_root, p, 1);
    P.smp = &PSOI(S->smp_l, c, 1);
    P.z = &PSNO(S->z, c, 1);
    P.sk = Lo.ldp;
    P.ss = Lo.ldc;
    P.n = cfg->nlevsoi;
    p_spacF(&P, x, f, q_sun, q_sha);
    return p_spacA(&P, x, iA, q_sun, q_sha);
}
