/*
 * balance_oracle.c -- CPU ORACLE (test infrastructure only) for the column
 * water mass, the column water balance check and the solar radiation balance
 * check.
 *
 * Never linked into or called from the product library; tests/ use it as the
 * checker of the HIP kernels.  Same conventions as ctsm_oracle.c: binary64,
 * -ffp-contract=off, the Fortran's operation order (left to right within an
 * expression, accumulations in the loop order the reference runs).
 *
 *   oracle_water_mass     ComputeWaterMassNonLake / ComputeLiqIceMassNonLake /
 *                         AccumulateSoilLiqIceMassNonLake
 *                         (TotalWaterAndHeatMod.F90:92-141, 200-393) for soil
 *                         and crop columns, with the canopy water's
 *                         p2c_1d_filter (subgridAveMod.F90:292-320)
 *   oracle_water_balance  errh2o_col and its max / maxloc
 *                         (BalanceCheckMod.F90:588-620)
 *   oracle_solar_balance  errsol and its max / maxloc for non-urban patches
 *                         (BalanceCheckMod.F90:964-970, 1008-1012)
 *
 * Pins (tests/test_oracle_balance.py): every store counts once with weight 1
 * (moving water between any two stores leaves the mass unchanged, adding to
 * one adds to the mass), the gates (inactive patches, snow layers above snl,
 * the aquifer of hydrologically inactive columns, the baselines only when
 * asked), the canopy water weighted by wtcol; errh2o is zero for a column
 * whose mass changed by exactly its net input and each flux moves it with the
 * sign of a gain or a loss; errsol is zero when the absorbed and reflected
 * radiation add up to the incoming; max / maxloc against numpy.
 */
#include <math.h>
#include <stddef.h>

#include "ctsm_gpu.h"

/* (c, j) of a column array with lower level bound lo */
static long cj(const ctsm_layout *L, int c, int j, int lo) { return (long)(c - L->begc_alloc) + (long)(j - lo) * L->ldc; }

/* p2c_1d_filter for one column: 0, then + x(p) * wtcol(p) over the column's
 * active patches in patch order */
static double p2c_col(const ctsm_layout *L, const ctsm_col_state *S, const ctsm_canopy_state *Cn, const double *x,
                      int c)
{
    const long cc = c - L->begc_alloc;
    double s = 0.0;
    int p;
    for (p = S->patchi[cc]; p <= S->patchf[cc]; ++p) {
        const long pp = p - L->begp_alloc;
        if (Cn->active[pp]) s = s + x[pp] * Cn->wtcol[pp];
    }
    return s;
}

void oracle_water_mass(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_col_state *S,
                       const ctsm_canopy_state *Cn, const ctsm_water_balance_state *W, int num_c, const int *filter_c,
                       double aquifer_water_baseline, int subtract_dynbal_baselines, double *water_mass)
{
    const int nlevsno = cfg->nlevsno, nlevgrnd = cfg->nlevgrnd;
    const int lo = 1 - nlevsno;
    int fc, j;
    for (fc = 0; fc < num_c; ++fc) {
        const int c = filter_c ? filter_c[fc] : L->begc_alloc + fc;
        const long cc = c - L->begc_alloc;
        double liquid = 0.0, ice = 0.0;
        const double liqcan_col = p2c_col(L, S, Cn, Cn->liqcan, c);
        const double snocan_col = p2c_col(L, S, Cn, Cn->snocan, c);
        liquid = liquid + liqcan_col + W->total_plant_stored_h2o[cc];
        ice = ice + snocan_col;
        ice = ice + S->h2osno_no_layers[cc];
        for (j = S->snl[cc] + 1; j <= 0; ++j) {
            liquid = liquid + S->h2osoi_liq[cj(L, c, j, lo)];
            ice = ice + S->h2osoi_ice[cj(L, c, j, lo)];
        }
        if (W->hydrologically_active[cc]) liquid = liquid + (W->wa[cc] - aquifer_water_baseline);
        liquid = liquid + S->h2osfc[cc];      /* soil and crop columns: not roof / wall / impervious road */
        for (j = 1; j <= nlevgrnd; ++j) {
            liquid = liquid + S->h2osoi_liq[cj(L, c, j, lo)];
            ice = ice + S->h2osoi_ice[cj(L, c, j, lo)] + (W->excess_ice ? W->excess_ice[cj(L, c, j, 1)] : 0.0);
        }
        if (subtract_dynbal_baselines) {
            liquid = liquid - W->dynbal_baseline_liq[cc];
            ice = ice - W->dynbal_baseline_ice[cc];
        }
        water_mass[cc] = liquid + ice;
    }
}

void oracle_water_balance(const ctsm_layout *L, const ctsm_water_balance_state *W, int num_c, const int *filter_c,
                          double dtime, double *errmax, int *index)
{
    double m = -1.0;
    int idx = 0, fc;
    for (fc = 0; fc < num_c; ++fc) {
        const int c = filter_c ? filter_c[fc] : L->begc_alloc + fc;
        const long cc = c - L->begc_alloc;
        double e = 0.0;
        if (W->col_active[cc])
            e = W->endwb[cc] - W->begwb[cc]
                - (W->forc_rain[cc] + W->forc_snow[cc] + W->qflx_floodc[cc] + W->qflx_sfc_irrig[cc]
                   + W->qflx_glcice_dyn_water_flux[cc] - W->qflx_evap_tot[cc] - W->qflx_surf[cc] - W->qflx_qrgwl[cc]
                   - W->qflx_drain[cc] - W->qflx_drain_perched[cc] - W->qflx_ice_runoff[cc]
                   - W->qflx_snwcp_discarded_liq[cc] - W->qflx_snwcp_discarded_ice[cc]) * dtime;
        if (W->errh2o) W->errh2o[cc] = e;
        if (fabs(e) > m || (fabs(e) == m && c < idx)) {
            m = fabs(e);
            idx = c;
        }
    }
    *errmax = m < 0.0 ? 0.0 : m;
    *index = idx;
}

void oracle_solar_balance(const ctsm_layout *L, const ctsm_canopy_state *Cn, int num_p, const int *filter_p,
                          const double *fsa, const double *fsr, const double *forc_solad_col, const double *forc_solai,
                          double *errsol, double *errmax, int *index)
{
    double m = -1.0;
    int idx = 0, fp;
    for (fp = 0; fp < num_p; ++fp) {
        const int p = filter_p ? filter_p[fp] : L->begp_alloc + fp;
        const long pp = p - L->begp_alloc;
        double e = 0.0;
        if (Cn->active[pp]) {
            const long cc = Cn->column[pp] - L->begc_alloc;
            const long gg = Cn->gridcell[pp] - L->begg_alloc;
            e = fsa[pp] + fsr[pp]
                - (forc_solad_col[cc] + forc_solad_col[cc + L->ldc] + forc_solai[gg] + forc_solai[gg + L->ldg]);
        }
        if (errsol) errsol[fp] = e;
        if (fabs(e) > m || (fabs(e) == m && p < idx)) {
            m = fabs(e);
            idx = p;
        }
    }
    *errmax = m < 0.0 ? 0.0 : m;
    *index = idx;
}
