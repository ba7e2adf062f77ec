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
 * and the calls after SoilWater (HydrologyNoDrainageMod.F90:355-373, no
 * aquifer layer) in oracle_water_table:
 *
 *   PerchedWaterTable        SoilHydrologyMod.F90:1525-1641
 *   ThetaBasedWaterTable     SoilHydrologyMod.F90:1933-2025
 *   RenewCondensation        SoilHydrologyMod.F90:2569-2678 (soil / crop columns),
 *                            truncate_small_values with tolerance 1e-12 (:71)
 *
 * and the soil-column drainage of HydrologyDrainage (HydrologyDrainageMod.F90:
 * 137-149, 160-245, no aquifer layer, no hillslopes) in oracle_drainage:
 *
 *   PerchedLateralFlow       SoilHydrologyMod.F90:1645-1929
 *   SubsurfaceLateralFlow    SoilHydrologyMod.F90:2029-2565 (lateral flow on)
 *   h2osoi_vol, qflx_runoff  HydrologyDrainageMod.F90:160-168, 236-244
 *
 * Pins (tests/test_oracle_hydrology.py): the water balance of the surface
 * (qflx_top_soil - evaporation terms = qflx_infl + qflx_surf + dh2osfc/dt over
 * the column), h2osfc never negative, qflx_infl never above the soil's
 * infiltration capacity plus h2osfc drainage, closed forms of fsat and qinmax,
 * the h2osfcflag = 0 path (no surface water pool), and cm_sin against mpmath;
 * for the water tables: saturated / dry / layered columns whose table depth is
 * fixed by construction, the interpolation's closed form, the frost-table
 * layer against a numpy search, and the water mass RenewCondensation adds.
 */
#include <math.h>
#include <stdlib.h>
#include "ctsm_gpu.h"
#include "ctsm_math.h"
#include "oracle_libm.h"   /* -DORACLE_LIBM: the host libm instead */

#define H_PI 3.14159265358979323846

typedef struct {
    int begc, ldc, nlevsno;
} h_lay_t;
#define HC(a, c)      ((a)[(c) - Lo.begc])
#define HSOI(a, c, j) ((a)[((c) - Lo.begc) + (long)((j) - 1) * Lo.ldc])
#define HSNO(a, c, j) ((a)[((c) - Lo.begc) + (long)((j) + Lo.nlevsno - 1) * Lo.ldc])
#define HZI(a, c, j)  ((a)[((c) - Lo.begc) + (long)((j) + Lo.nlevsno) * Lo.ldc])

static double h_max(double a, double b) { return (b > a) ? b : a; }
static double h_min(double a, double b) { return (b < a) ? b : a; }

int oracle_surface_water(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_col_state *S,
                         ctsm_surface_water_state *W, const ctsm_hydro_params *P, int num_c, const int *filter,
                         double dtime)
{
    h_lay_t Lo = { L->begc_alloc, L->ldc, cfg->nlevsno };
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

/* ---- water tables and condensation after SoilWater ---------------------- */

#define H_DENH2O 1000.0
#define H_DENICE 917.0
#define H_TFRZ 273.15

int oracle_water_table(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_col_state *S,
                       ctsm_water_table_state *W, int num_c, const int *filter, double dtime)
{
    h_lay_t Lo = { L->begc_alloc, L->ldc, cfg->nlevsno };
    const int nlevsoi = cfg->nlevsoi;
    int fc, k, rc = 0;
    double *ice_before = (double *)calloc((size_t)(L->ldc > 0 ? L->ldc : 1), sizeof(double));
    int *modified = (int *)calloc((size_t)(num_c > 0 ? num_c : 1), sizeof(int));
    int num_modified = 0;
    if (!ice_before || !modified) { free(ice_before); free(modified); return 1; }
#define COL(fc) (filter ? filter[fc] : L->begc_alloc + (fc))
#define VOL(c, k) (HSNO(S->h2osoi_liq, c, k) / (HSNO(S->dz, c, k) * H_DENH2O) \
                   + HSNO(S->h2osoi_ice, c, k) / (HSNO(S->dz, c, k) * H_DENICE))

    /* PerchedWaterTable */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        const double sat_lev = (double)0.9f;
        int k_frz, k_perch;
        if (HSNO(S->t_soisno, c, 1) > H_TFRZ) k_frz = nlevsoi;
        else k_frz = 1;
        for (k = 2; k <= nlevsoi; ++k) {
            if (HSNO(S->t_soisno, c, k - 1) > H_TFRZ && HSNO(S->t_soisno, c, k) <= H_TFRZ) {
                k_frz = k;
                break;
            }
        }
        HC(W->frost_table, c) = HZI(S->zi, c, k_frz - 1);
        HC(W->zwt_perched, c) = HC(W->frost_table, c);
        if (HC(W->zwt, c) < HC(W->frost_table, c) && HSNO(S->t_soisno, c, k_frz) <= H_TFRZ) {
            /* water table above the frost table */
        } else if (k_frz > 1) {
            k_perch = 1;
            for (k = k_frz; k >= 1; --k) {
                const double vol = VOL(c, k);
                if (W->h2osoi_vol) HSOI(W->h2osoi_vol, c, k) = vol;
                if (vol / HSOI(S->watsat, c, k) <= sat_lev) {
                    k_perch = k;
                    break;
                }
            }
            if (HSNO(S->t_soisno, c, k_frz) > H_TFRZ) k_perch = k_frz;
            if (k_frz > k_perch) {
                const double s1 = VOL(c, k_perch) / HSOI(S->watsat, c, k_perch);
                const double s2 = VOL(c, k_perch + 1) / HSOI(S->watsat, c, k_perch + 1);
                if (s1 > s2) {
                    HC(W->zwt_perched, c) = HZI(S->zi, c, k_perch - 1);
                } else {
                    const double m = (HSNO(S->z, c, k_perch + 1) - HSNO(S->z, c, k_perch)) / (s2 - s1);
                    const double b = HSNO(S->z, c, k_perch + 1) - m * s2;
                    HC(W->zwt_perched, c) = h_max(0.0, m * sat_lev + b);
                }
            }
        }
    }

    /* ThetaBasedWaterTable */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        const int nb = HC(S->nbedrock, c);
        const double sat_lev = (double)0.9f;
        int k_zwt = nb, sat_flag = 1;
        HC(W->zwt, c) = HZI(S->zi, c, nlevsoi);
        for (k = nb; k >= 1; --k) {
            const double vol = VOL(c, k);
            if (W->h2osoi_vol) HSOI(W->h2osoi_vol, c, k) = vol;
            if (vol / HSOI(S->watsat, c, k) <= sat_lev) {
                k_zwt = k;
                sat_flag = 0;
                break;
            }
        }
        if (sat_flag == 1) k_zwt = 1;
        if (k_zwt == 1) {
            HC(W->zwt, c) = HZI(S->zi, c, 1);
        } else if (k_zwt < nb) {
            const double s1 = VOL(c, k_zwt) / HSOI(S->watsat, c, k_zwt);
            const double s2 = VOL(c, k_zwt + 1) / HSOI(S->watsat, c, k_zwt + 1);
            const double m = (HSNO(S->z, c, k_zwt + 1) - HSNO(S->z, c, k_zwt)) / (s2 - s1);
            const double b = HSNO(S->z, c, k_zwt + 1) - m * s2;
            HC(W->zwt, c) = h_max(0.0, m * sat_lev + b);
        } else {
            HC(W->zwt, c) = HZI(S->zi, c, nb);
        }
    }

    /* RenewCondensation */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        if (HC(S->snl, c) + 1 >= 1) {
            modified[num_modified++] = c;
            HSNO(S->h2osoi_liq, c, 1) = HSNO(S->h2osoi_liq, c, 1)
                                        + (1.0 - HC(S->frac_h2osfc, c)) * HC(W->qflx_liqdew_to_top_layer, c) * dtime;
            HSNO(S->h2osoi_ice, c, 1) = HSNO(S->h2osoi_ice, c, 1)
                                        + (1.0 - HC(S->frac_h2osfc, c)) * HC(W->qflx_soliddew_to_top_layer, c) * dtime;
            ice_before[c - Lo.begc] = HSNO(S->h2osoi_ice, c, 1);
            HSNO(S->h2osoi_ice, c, 1) = HSNO(S->h2osoi_ice, c, 1)
                                        - (1.0 - HC(S->frac_h2osfc, c)) * HC(W->qflx_solidevap_from_top_layer, c) * dtime;
        }
    }
    for (fc = 0; fc < num_modified; ++fc) {
        const int c = modified[fc];
        if (fabs(HSNO(S->h2osoi_ice, c, 1)) < 1.e-12 * fabs(ice_before[c - Lo.begc])) HSNO(S->h2osoi_ice, c, 1) = 0.0;
    }
    for (fc = 0; fc < num_modified; ++fc) {
        const int c = modified[fc];
        if (HSNO(S->h2osoi_ice, c, 1) < 0.0) { rc = c; break; }
    }
#undef VOL
#undef COL
    free(ice_before);
    free(modified);
    return rc;
}

/* ---- drainage (HydrologyDrainage, soil columns) ---------------------------- */

#define H_WATMIN 0.01
#define H_PONDMX 0.0

static double h_spec_yield(const h_lay_t *Lp, const ctsm_col_state *S, const ctsm_hydro_params *P, int c, int k,
                           double zw)
{
    const h_lay_t Lo = *Lp;
    double s_y = HSOI(S->watsat, c, k) * (1. - cm_pow(1. + 1.e3 * zw / HSOI(S->sucsat, c, k), -(1. / HSOI(S->bsw, c, k))));
    return h_max(s_y, P->aq_sp_yield_min);
}

int oracle_drainage(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_col_state *S, ctsm_drainage_state *D,
                    const ctsm_hydro_params *P, int num_c, const int *filter, double dtime)
{
    h_lay_t Lo = { L->begc_alloc, L->ldc, cfg->nlevsno };
    const int nlevsoi = cfg->nlevsoi, nlevgrnd = cfg->nlevgrnd;
    const int ncol = L->ldc > 0 ? L->ldc : 1;
    int *k_frost = (int *)calloc((size_t)ncol, sizeof(int));
    int *k_perch = (int *)calloc((size_t)ncol, sizeof(int));
    int *jwt = (int *)calloc((size_t)ncol, sizeof(int));
    double *out = (double *)calloc((size_t)ncol, sizeof(double));
    double *icef = (double *)calloc((size_t)ncol * (size_t)nlevsoi, sizeof(double));
    double *imped_col = (double *)calloc((size_t)ncol, sizeof(double));
    double *drainage = (double *)calloc((size_t)ncol, sizeof(double));
    int fc, k, j, i;
    if (!k_frost || !k_perch || !jwt || !out || !icef || !imped_col || !drainage) {
        free(k_frost); free(k_perch); free(jwt); free(out); free(icef); free(imped_col); free(drainage);
        return 1;
    }
#define COL(fc) (filter ? filter[fc] : L->begc_alloc + (fc))
#define IX(c) ((c) - Lo.begc)
#define LIQ(c, k) HSNO(S->h2osoi_liq, c, k)
#define ICE(c, k) HSNO(S->h2osoi_ice, c, k)
#define DZ(c, k) HSNO(S->dz, c, k)
#define ZI(c, k) HZI(S->zi, c, k)

    /* ---- PerchedLateralFlow */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc), nb = HC(S->nbedrock, c);
        k_frost[IX(c)] = nb;
        k_perch[IX(c)] = nb;
        for (k = 1; k <= nb; ++k) {
            if (HC(D->frost_table, c) >= ZI(c, k - 1) && HC(D->frost_table, c) < ZI(c, k)) { k_frost[IX(c)] = k; break; }
        }
        for (k = 1; k <= nb; ++k) {
            if (HC(D->zwt_perched, c) >= ZI(c, k - 1) && HC(D->zwt_perched, c) < ZI(c, k)) { k_perch[IX(c)] = k; break; }
        }
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(D->qflx_drain_perched, c) = 0.0;
        out[IX(c)] = 0.0;
        if (HC(D->frost_table, c) > HC(D->zwt_perched, c)) {
            const double q_perch_max = P->perched_baseflow_scalar * cm_sin(HC(D->topo_slope, c) * (H_PI / 180.0));
            double wtsub = 0.0, q_perch = 0.0;
            for (k = k_perch[IX(c)]; k <= k_frost[IX(c)] - 1; ++k) {
                q_perch = q_perch + HSOI(S->hksat, c, k) * DZ(c, k);
                wtsub = wtsub + DZ(c, k);
            }
            if (wtsub > 0.0) q_perch = q_perch / wtsub;
            out[IX(c)] = q_perch_max * q_perch * (HC(D->frost_table, c) - HC(D->zwt_perched, c));
        }
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(D->qflx_drain_perched, c) = HC(D->qflx_drain_perched, c) + out[IX(c)];
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        double tot = HC(D->qflx_drain_perched, c) * dtime;
        for (k = k_perch[IX(c)]; k <= k_frost[IX(c)] - 1; ++k) {
            const double s_y = h_spec_yield(&Lo, S, P, c, k, HC(D->zwt_perched, c));
            double dl;
            if (k == k_perch[IX(c)]) dl = h_min(tot, (s_y * (ZI(c, k) - HC(D->zwt_perched, c)) * 1.e3));
            else dl = h_min(tot, (s_y * (DZ(c, k)) * 1.e3));
            dl = h_max(dl, 0.0);
            tot = tot - dl;
            LIQ(c, k) = LIQ(c, k) - dl;
        }
        HC(D->qflx_drain_perched, c) = HC(D->qflx_drain_perched, c) - tot / dtime;
    }

    /* ---- SubsurfaceLateralFlow */
    for (j = 1; j <= nlevsoi; ++j) {
        for (fc = 0; fc < num_c; ++fc) {
            const int c = COL(fc);
            const double vol_ice = h_min(HSOI(S->watsat, c, j), ICE(c, j) / (DZ(c, j) * H_DENICE));
            icef[(size_t)IX(c) * nlevsoi + (j - 1)] = h_min(1.0, vol_ice / HSOI(S->watsat, c, j));
            HSOI(S->icefrac, c, j) = icef[(size_t)IX(c) * nlevsoi + (j - 1)];
        }
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(D->qflx_drain, c) = 0.0;
        HC(D->qflx_rsub_sat, c) = 0.0;
        drainage[IX(c)] = 0.0;
        HC(D->qflx_latflow_in, c) = 0.0;
        HC(D->qflx_latflow_out, c) = 0.0;
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        jwt[IX(c)] = nlevsoi;
        for (j = 1; j <= nlevsoi; ++j) {
            if (HC(D->zwt, c) <= ZI(c, j)) { jwt[IX(c)] = j - 1; break; }
        }
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        double dzsum = 0.0, icefracsum = 0.0;
        for (j = (jwt[IX(c)] > 1 ? jwt[IX(c)] : 1); j <= nlevsoi; ++j) {
            const double dzmm = DZ(c, j) * 1.e3;
            dzsum = dzsum + dzmm;
            icefracsum = icefracsum + icef[(size_t)IX(c) * nlevsoi + (j - 1)] * dzmm;
        }
        imped_col[IX(c)] = cm_pow(10.0, -cfg->e_ice * (icefracsum / dzsum));
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc), nb = HC(S->nbedrock, c);
        if (HC(D->zwt, c) <= ZI(c, nb))
            HC(D->qflx_latflow_out, c) = imped_col[IX(c)] * P->baseflow_scalar * cm_tan(H_PI / 180.0 * HC(D->topo_slope, c))
                                         * cm_pow(ZI(c, nb) - HC(D->zwt, c), P->n_baseflow);
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc), nb = HC(S->nbedrock, c);
        const double net = HC(D->qflx_latflow_out, c) - HC(D->qflx_latflow_in, c);
        double tot;
        if (HC(D->zwt, c) <= ZI(c, nb)) drainage[IX(c)] = net;
        else drainage[IX(c)] = 0.0;
        tot = -drainage[IX(c)] * dtime;
        if (tot > 0.) {
            for (j = jwt[IX(c)] + 1; j >= 1; --j) {
                if (ZI(c, j) < HC(D->frost_table, c)) {
                    const double s_y = h_spec_yield(&Lo, S, P, c, j, HC(D->zwt, c));
                    double dl = h_min(tot, (s_y * DZ(c, j) * 1.e3));
                    dl = h_max(dl, 0.0);
                    LIQ(c, j) = LIQ(c, j) + dl;
                    tot = tot - dl;
                    if (tot <= 0.) {
                        HC(D->zwt, c) = HC(D->zwt, c) - dl / s_y / 1000.0;
                        break;
                    } else {
                        HC(D->zwt, c) = ZI(c, j - 1);
                    }
                }
            }
            HC(S->h2osfc, c) = HC(S->h2osfc, c) + tot;
        } else {
            for (j = jwt[IX(c)] + 1; j <= nb; ++j) {
                const double s_y = h_spec_yield(&Lo, S, P, c, j, HC(D->zwt, c));
                double dl = h_max(tot, -(s_y * (ZI(c, j) - HC(D->zwt, c)) * 1.e3));
                dl = h_min(dl, 0.0);
                LIQ(c, j) = LIQ(c, j) + dl;
                tot = tot - dl;
                if (tot >= 0.) {
                    HC(D->zwt, c) = HC(D->zwt, c) - dl / s_y / 1000.0;
                    break;
                } else {
                    HC(D->zwt, c) = ZI(c, j);
                }
            }
            drainage[IX(c)] = drainage[IX(c)] + tot / dtime;
        }
        HC(D->zwt, c) = h_max(0.0, HC(D->zwt, c));
        HC(D->zwt, c) = h_min(80.0, HC(D->zwt, c));
    }
    for (j = nlevsoi; j >= 2; --j) {
        for (fc = 0; fc < num_c; ++fc) {
            const int c = COL(fc);
            const double dzmm = DZ(c, j) * 1.e3;
            const double xsi = h_max(LIQ(c, j) - HSOI(S->eff_porosity, c, j) * dzmm, 0.0);
            LIQ(c, j) = h_min(HSOI(S->eff_porosity, c, j) * dzmm, LIQ(c, j));
            LIQ(c, j - 1) = LIQ(c, j - 1) + xsi;
        }
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        const double dzmm1 = DZ(c, 1) * 1.e3;
        double xs1 = h_max(h_max(LIQ(c, 1) - H_WATMIN, 0.0)
                           - h_max(0.0, (H_PONDMX + HSOI(S->watsat, c, 1) * dzmm1 - ICE(c, 1) - H_WATMIN)), 0.0);
        LIQ(c, 1) = LIQ(c, 1) - xs1;
        HC(S->h2osfc, c) = HC(S->h2osfc, c) + xs1;
        HC(D->qflx_rsub_sat, c) = 0.0;
        xs1 = h_max(h_max(ICE(c, 1), 0.0) - h_max(0.0, (H_PONDMX + HSOI(S->watsat, c, 1) * dzmm1 - LIQ(c, 1))), 0.0);
        ICE(c, 1) = h_min(h_max(0.0, H_PONDMX + HSOI(S->watsat, c, 1) * dzmm1 - LIQ(c, 1)), ICE(c, 1));
        HC(D->qflx_ice_runoff_xs, c) = xs1 / dtime;
    }
    for (j = 1; j <= nlevsoi - 1; ++j) {
        for (fc = 0; fc < num_c; ++fc) {
            const int c = COL(fc);
            double xs;
            if (LIQ(c, j) < H_WATMIN) {
                xs = H_WATMIN - LIQ(c, j);
                if (j == jwt[IX(c)]) HC(D->zwt, c) = HC(D->zwt, c) + xs / HSOI(S->eff_porosity, c, j) / 1000.0;
            } else {
                xs = 0.0;
            }
            LIQ(c, j) = LIQ(c, j) + xs;
            LIQ(c, j + 1) = LIQ(c, j + 1) - xs;
        }
    }
    j = nlevsoi;
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        double xs;
        if (LIQ(c, j) < H_WATMIN) {
            xs = H_WATMIN - LIQ(c, j);
            for (i = nlevsoi - 1; i >= 1; --i) {
                const double avail = h_max(LIQ(c, i) - H_WATMIN - xs, 0.0);
                if (avail >= xs) {
                    LIQ(c, j) = LIQ(c, j) + xs;
                    LIQ(c, i) = LIQ(c, i) - xs;
                    xs = 0.0;
                    break;
                } else {
                    LIQ(c, j) = LIQ(c, j) + avail;
                    LIQ(c, i) = LIQ(c, i) - avail;
                    xs = xs - avail;
                }
            }
        } else {
            xs = 0.0;
        }
        LIQ(c, j) = LIQ(c, j) + xs;
        HC(D->qflx_rsub_sat, c) = HC(D->qflx_rsub_sat, c) - xs / dtime;
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(D->qflx_drain, c) = HC(D->qflx_rsub_sat, c) + drainage[IX(c)];
        HC(D->qflx_qrgwl, c) = HC(D->qflx_snwcp_liq, c);
    }

    /* ---- HydrologyDrainage: volumetric water, runoff */
    if (D->h2osoi_vol) {
        for (j = 1; j <= nlevgrnd; ++j) {
            for (fc = 0; fc < num_c; ++fc) {
                const int c = COL(fc);
                HSOI(D->h2osoi_vol, c, j) = LIQ(c, j) / (DZ(c, j) * H_DENH2O) + ICE(c, j) / (DZ(c, j) * H_DENICE);
            }
        }
    }
    for (fc = 0; fc < num_c; ++fc) {
        const int c = COL(fc);
        HC(D->qflx_runoff, c) = HC(D->qflx_drain, c) + HC(D->qflx_surf, c) + HC(D->qflx_qrgwl, c)
                                + HC(D->qflx_drain_perched, c);
    }
#undef COL
#undef IX
#undef LIQ
#undef ICE
#undef DZ
#undef ZI
    free(k_frost); free(k_perch); free(jwt); free(out); free(icef); free(imped_col); free(drainage);
    return 0;
}
