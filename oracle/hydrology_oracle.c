/*
 * hydrology_oracle.c -- CPU ORACLE (test infrastructure only) for the surface
 * water input chain of HydrologyNoDrainage (HydrologyNoDrainageMod.F90:306-336).
 *
 * Never linked into or called from the product library; tests/ use it as the
 * checker of the HIP kernel.  Same conventions as ctsm_oracle.c: binary64,
 * -ffp-contract=off, the Fortran's operation order, exp / pow / sin from
 * ctsm_math.h, default-REAL literals kept as written (1.0e-8 of
 * SurfaceWaterMod.F90:497 is the single-precision value).  Each routine is a
 * loop over the filter in the reference's order:
 *
 *   SaturatedExcessRunoff    SaturatedExcessRunoffMod.F90:203-312 with
 *                            ComputeFsatTopmodel :315-360 (crop / hillslope
 *                            fsat switches off, no urban columns)
 *   SetQflxInputs            SoilHydrologyMod.F90:298-366
 *   InfiltrationExcessRunoff InfiltrationExcessRunoffMod.F90 (qinmax method
 *                            HKSAT, ComputeQinmaxHksat: minval over layers 1..3)
 *   RouteInfiltrationExcess  SoilHydrologyMod.F90:369-423 (soil / crop branch)
 *   UpdateH2osfc             SurfaceWaterMod.F90:345-433 with QflxH2osfcSurf
 *                            :436-505 (non-hillslope columns) and QflxH2osfcDrain
 *                            :508-560, truncate_small_values (NumericsMod.F90:50-99,
 *                            rel_epsilon 1e-13)
 *   Infiltration             SoilHydrologyMod.F90:426-457
 *   TotalSurfaceRunoff       SoilHydrologyMod.F90:460-550 (hydrologic columns)
 *
 * Pins (tests/test_oracle_hydrology.py): the water balance of the surface
 * (qflx_top_soil - evaporation terms = qflx_infl + qflx_surf + dh2osfc/dt over
 * the column), h2osfc never negative, qflx_infl never above the soil's
 * infiltration capacity plus h2osfc drainage, closed forms of fsat and qinmax,
 * the h2osfcflag = 0 path (no surface water pool), and cm_sin against mpmath.
 */
#include <math.h>
#include <stdlib.h>
#include "ctsm_gpu.h"
#include "ctsm_math.h"

#define H_PI 3.14159265358979323846

typedef struct {
    int begc, ldc;
} h_lay_t;
#define HC(a, c)      ((a)[(c) - Lo.begc])
#define HSOI(a, c, j) ((a)[((c) - Lo.begc) + (long)((j) - 1) * Lo.ldc])

static double h_max(double a, double b) { return (b > a) ? b : a; }
static double h_min(double a, double b) { return (b < a) ? b : a; }

int oracle_surface_water(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_col_state *S,
                         ctsm_surface_water_state *W, const ctsm_hydro_params *P, int num_c, const int *filter,
                         double dtime)
{
    h_lay_t Lo = { L->begc_alloc, L->ldc };
    double *h2osfc_partial = (double *)calloc((size_t)(L->ldc > 0 ? L->ldc : 1), sizeof(double));
    int fc, j;
    if (!h2osfc_partial) return 1;
#define COL(fc) (filter ? filter[fc] : L->begc_alloc + (fc))

    /* SaturatedExcessRunoff: TOPMODEL fsat, then the saturated-area runoff */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        if (HC(W->frost_table, c) > HC(W->zwt_perched, c) && HC(W->frost_table, c) <= HC(W->zwt, c))
            HC(W->fsat, c) = HC(W->wtfact, c) * cm_exp(-0.5 * P->fff * HC(W->zwt_perched, c));
        else
            HC(W->fsat, c) = HC(W->wtfact, c) * cm_exp(-0.5 * P->fff * HC(W->zwt, c));
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(W->qflx_sat_excess_surf, c) = HC(W->fsat, c) * HC(W->qflx_rain_plus_snomelt, c);
        HC(W->fcov, c) = HC(W->fsat, c);
    }

    /* SetQflxInputs */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        double fsno, qflx_evap;
        HC(W->qflx_top_soil, c) = HC(W->qflx_rain_plus_snomelt, c) + HC(W->qflx_snow_h2osfc, c) + HC(W->qflx_floodc, c);
        if (HC(S->snl, c) >= 0) {
            fsno = 0.0;
            qflx_evap = HC(W->qflx_liqevap_from_top_layer, c);
        } else {
            fsno = HC(S->frac_sno_eff, c);
            qflx_evap = HC(W->qflx_ev_soil, c);
        }
        const double fh = HC(S->frac_h2osfc, c);
        HC(W->qflx_in_soil, c) = (1.0 - fh) * (HC(W->qflx_top_soil, c) - HC(W->qflx_sat_excess_surf, c));
        HC(W->qflx_top_soil_to_h2osfc, c) = fh * (HC(W->qflx_top_soil, c) - HC(W->qflx_sat_excess_surf, c));
        HC(W->qflx_in_soil, c) = HC(W->qflx_in_soil, c) - (1.0 - fsno - fh) * qflx_evap;
        HC(W->qflx_top_soil_to_h2osfc, c) = HC(W->qflx_top_soil_to_h2osfc, c) - fh * HC(W->qflx_ev_h2osfc, c);
    }

    /* InfiltrationExcessRunoff, qinmax from hksat and ice impedance over layers 1..3 */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        double qmin = 0.0;
        for (j = 1; j <= 3; ++j) {
            const double v = cm_pow(10.0, -cfg->e_ice * HSOI(S->icefrac, c, j)) * HSOI(S->hksat, c, j);
            if (j == 1 || v < qmin) qmin = v;
        }
        HC(W->qinmax, c) = (1.0 - HC(W->fsat, c)) * qmin;
        HC(W->qflx_infl_excess, c) = h_max(0.0, (HC(W->qflx_in_soil, c) - (1.0 - HC(S->frac_h2osfc, c)) * HC(W->qinmax, c)));
    }

    /* RouteInfiltrationExcess (soil / crop landunits) */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(W->qflx_in_soil_limited, c) = HC(W->qflx_in_soil, c) - HC(W->qflx_infl_excess, c);
        if (P->h2osfcflag != 0) {
            HC(W->qflx_in_h2osfc, c) = HC(W->qflx_top_soil_to_h2osfc, c) + HC(W->qflx_infl_excess, c);
            HC(W->qflx_infl_excess_surf, c) = 0.0;
        } else {
            HC(W->qflx_in_h2osfc, c) = HC(W->qflx_top_soil_to_h2osfc, c);
            HC(W->qflx_infl_excess_surf, c) = HC(W->qflx_infl_excess, c);
        }
    }

    /* UpdateH2osfc: QflxH2osfcSurf */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        double frac_infclust = 0.0;
        if (P->h2osfcflag == 1) {
            if (HC(W->frac_h2osfc_nosnow, c) <= P->pc)
                frac_infclust = 0.0;
            else
                frac_infclust = cm_pow(HC(W->frac_h2osfc_nosnow, c) - P->pc, P->mu);
        }
        const double h = HC(W->h2osfc, c), th = HC(W->h2osfc_thresh, c);
        double q;
        if (h > th && P->h2osfcflag != 0) {
            const double k_wet = 1.0e-4 * cm_sin((H_PI / 180.0) * HC(W->topo_slope, c));
            q = k_wet * frac_infclust * (h - th);
            q = h_min(q, (h - th) / dtime);
        } else {
            q = 0.0;
        }
        if (q < (double)1.0e-8f) q = 0.0;
        HC(W->qflx_h2osfc_surf, c) = q;
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        h2osfc_partial[c - Lo.begc] = HC(W->h2osfc, c) + (HC(W->qflx_in_h2osfc, c) - HC(W->qflx_h2osfc_surf, c)) * dtime;
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        if (fabs(h2osfc_partial[c - Lo.begc]) < 1.e-13 * fabs(HC(W->h2osfc, c))) h2osfc_partial[c - Lo.begc] = 0.0;
    }
    /* QflxH2osfcDrain */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        const double hp = h2osfc_partial[c - Lo.begc];
        double d;
        if (hp < 0.0) {
            d = hp / dtime;
        } else {
            d = h_min(HC(S->frac_h2osfc, c) * HC(W->qinmax, c), hp / dtime);
            if (P->h2osfcflag == 0) d = h_max(0.0, hp / dtime);
        }
        HC(W->qflx_h2osfc_drain, c) = d;
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(W->h2osfc, c) = h2osfc_partial[c - Lo.begc] - HC(W->qflx_h2osfc_drain, c) * dtime;
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        if (fabs(HC(W->h2osfc, c)) < 1.e-13 * fabs(h2osfc_partial[c - Lo.begc])) HC(W->h2osfc, c) = 0.0;
    }

    /* Infiltration */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(W->qflx_infl, c) = HC(W->qflx_in_soil_limited, c) + HC(W->qflx_h2osfc_drain, c);
    }

    /* TotalSurfaceRunoff (hydrologically active columns) */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(W->qflx_surf, c) = HC(W->qflx_sat_excess_surf, c) + HC(W->qflx_infl_excess_surf, c)
                              + HC(W->qflx_h2osfc_surf, c);
    }
#undef COL
    free(h2osfc_partial);
    return 0;
}

/* clm_drv_patch2col for the three evaporation fluxes SetQflxInputs reads
 * (p2c_1d_filter, subgridAveMod.F90:312-318). */
void oracle_patch2col3(const ctsm_layout *L, const ctsm_col_state *S, const int32_t *active, const double *wtcol,
                       const double *pa, const double *pb, const double *pc, double *ca, double *cb, double *cc,
                       int num_c, const int *filter)
{
    int fc, p;
    for (fc = 0; fc < num_c; ++fc) {
        const int c = filter ? filter[fc] : L->begc_alloc + fc;
        const int k = c - L->begc_alloc;
        double a = 0.0, b = 0.0, d = 0.0;
        for (p = S->patchi[k]; p <= S->patchf[k]; ++p) {
            const int q = p - L->begp_alloc;
            if (!active[q]) continue;
            a = a + pa[q] * wtcol[q];
            b = b + pb[q] * wtcol[q];
            d = d + pc[q] * wtcol[q];
        }
        ca[k] = a;
        cb[k] = b;
        cc[k] = d;
    }
}
