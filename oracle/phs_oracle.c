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
 * inactive, snow-buried and zero-weight patches.
 */
#include <math.h>
#include <stdlib.h>
#include "ctsm_gpu.h"
#include "ctsm_math.h"

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
