/*
 * soilfluxes_oracle.c -- CPU ORACLE (test infrastructure only) for SoilFluxes
 * and the soil part of the energy balance check.
 *
 * Never linked into or called from the product library; tests/ use it as the
 * checker of the HIP kernels.  Same conventions as ctsm_oracle.c: binary64,
 * -ffp-contract=off, the Fortran's operation order, integer powers by
 * squaring ((x*x)*(x*x) for x**4, (x*x)*x for x**3), default-REAL literals
 * kept as written (`1.` of :470 is exactly 1).
 *
 *   oracle_save_before      BiogeophysPreFluxCalcsMod.F90:302-331 (t_ssbef,
 *                           t_h2osfc_bef; non-urban columns)
 *   oracle_soilfluxes       SoilFluxesMod.F90:38-518 for soil / crop patches
 *                           (the urban branches are not taken), with the
 *                           p2c_1d_filter of errsoi (subgridAveMod.F90:292-320)
 *   oracle_soil_balance     BalanceCheckMod.F90:1101-1119: max |errsoi_col|
 *                           and its maxloc over the filtered columns
 *
 * Pins (tests/test_oracle_soilfluxes.py): errsoi closes over the heat solve
 * (|errsoi| < 1e-7 W/m2, the reference warns at 1e-5); the energy identities
 * eflx_soil_grnd = absorbed + net longwave - sensible - latent, and
 * qflx_evap_tot / eflx_lh_tot / eflx_sh_tot sums; the evaporation partition
 * identity qflx_ev_snow = liqevap + solidevap - liqdew - soliddew (when no
 * limit applies) and its limit cases; the longwave closure
 * eflx_lwrad_net = eflx_lwrad_out - forc_lwrad; tinc = 0 leaves the canopy
 * fluxes unchanged; the errsoi value agrees with oracle_errsoi (the earlier,
 * independently written heat-solve residual) on states with no evaporation
 * correction.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ctsm_gpu.h"

#define SF_TFRZ 273.15
#define SF_SB   5.67e-8
#define SF_HVAP 2.501e6

typedef struct {
    int begc, ldc, begp, ldp, nlevsno;
} sf_lay_t;
#define SC1(a, c)     ((a)[(c) - Lo.begc])
#define SSNO(a, c, j) ((a)[((c) - Lo.begc) + (long)((j) + Lo.nlevsno - 1) * Lo.ldc])
#define SP1(a, p)     ((a)[(p) - Lo.begp])

static double sf_max(double a, double b) { return (b > a) ? b : a; }
static double p4(double x) { const double x2 = x * x; return x2 * x2; }
static double p3(double x) { return (x * x) * x; }

/* BiogeophysPreFluxCalcsMod.F90:302-331, non-urban columns */
void oracle_save_before(const ctsm_config *cfg, const ctsm_layout *L, ctsm_col_state *S,
                        ctsm_soilflux_state *F, int num_filter, const int *filter)
{
    sf_lay_t Lo = { L->begc_alloc, L->ldc, L->begp_alloc, L->ldp, cfg->nlevsno };
    int fc, j;
    for (j = -cfg->nlevsno + 1; j <= cfg->nlevgrnd; ++j)
        for (fc = 0; fc < num_filter; ++fc) {
            const int c = filter ? filter[fc] : L->begc_alloc + fc;
            SSNO(F->t_ssbef, c, j) = SSNO(S->t_soisno, c, j);
        }
    for (fc = 0; fc < num_filter; ++fc) {
        const int c = filter ? filter[fc] : L->begc_alloc + fc;
        SC1(F->t_h2osfc_bef, c) = SC1(S->t_h2osfc, c);
    }
}

/* SoilFluxesMod.F90:38-518 (loop structure and order of the reference) */
int oracle_soilfluxes(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_col_state *S,
                      const ctsm_patch_state *P, ctsm_soilflux_state *F, int num_c, const int *filter_c,
                      int num_p, const int *filter_p, double dtime)
{
    sf_lay_t Lo = { L->begc_alloc, L->ldc, L->begp_alloc, L->ldp, cfg->nlevsno };
    double *tinc = (double *)calloc((size_t)(L->ldc > 0 ? L->ldc : 1), sizeof(double));
    double *t_grnd0 = (double *)calloc((size_t)(L->ldc > 0 ? L->ldc : 1), sizeof(double));
    int fc, fp, j;
    if (!tinc || !t_grnd0) { free(tinc); free(t_grnd0); return 1; }
#define CIDX(c) ((c) - Lo.begc)

    /* :257-273 ground temperature of the previous step and its increment */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = filter_c ? filter_c[fc] : L->begc_alloc + fc;
        const int snl = SC1(S->snl, c);
        const double fs = SC1(S->frac_sno_eff, c), fh = SC1(S->frac_h2osfc, c);
        if (snl < 0)
            t_grnd0[CIDX(c)] = fs * SSNO(F->t_ssbef, c, snl + 1) + (1 - fs - fh) * SSNO(F->t_ssbef, c, 1)
                               + fh * SC1(F->t_h2osfc_bef, c);
        else
            t_grnd0[CIDX(c)] = (1 - fh) * SSNO(F->t_ssbef, c, 1) + fh * SC1(F->t_h2osfc_bef, c);
        tinc[CIDX(c)] = SC1(S->t_grnd, c) - t_grnd0[CIDX(c)];
    }

    /* :277-294 fluxes corrected to the present ground temperature */
    for (fp = 0; fp < num_p; ++fp) {
        const int p = filter_p ? filter_p[fp] : L->begp_alloc + fp;
        const int c = SP1(P->column, p);
        const double ti = tinc[CIDX(c)];
        SP1(F->eflx_sh_grnd, p) = SP1(F->eflx_sh_grnd, p) + ti * SP1(F->cgrnds, p);
        SP1(F->qflx_evap_soi, p) = SP1(F->qflx_evap_soi, p) + ti * SP1(F->cgrndl, p);
        SP1(F->qflx_ev_snow, p) = SP1(F->qflx_ev_snow, p) + ti * SP1(F->cgrndl, p);
        SP1(F->qflx_ev_soil, p) = SP1(F->qflx_ev_soil, p) + ti * SP1(F->cgrndl, p);
        SP1(F->qflx_ev_h2osfc, p) = SP1(F->qflx_ev_h2osfc, p) + ti * SP1(F->cgrndl, p);
    }

    /* :297-358 evaporation partitioned into liquid / solid, evaporation / dew */
    for (fp = 0; fp < num_p; ++fp) {
        const int p = filter_p ? filter_p[fp] : L->begp_alloc + fp;
        const int c = SP1(P->column, p);
        const int jt = SC1(S->snl, c) + 1;
        const double ev = SP1(F->qflx_ev_snow, p);
        SP1(F->qflx_liqevap_from_top_layer, p) = 0.0;
        SP1(F->qflx_solidevap_from_top_layer, p) = 0.0;
        SP1(F->qflx_soliddew_to_top_layer, p) = 0.0;
        SP1(F->qflx_liqdew_to_top_layer, p) = 0.0;
        if (ev >= 0.0) {
            const double liq = SSNO(S->h2osoi_liq, c, jt), ice = SSNO(S->h2osoi_ice, c, jt);
            if ((liq + ice) > 0.0)
                SP1(F->qflx_liqevap_from_top_layer, p) = sf_max(ev * (liq / (liq + ice)), 0.0);
            else
                SP1(F->qflx_liqevap_from_top_layer, p) = 0.0;
            SP1(F->qflx_solidevap_from_top_layer, p) = ev - SP1(F->qflx_liqevap_from_top_layer, p);
        } else {
            if (SC1(S->t_grnd, c) < SF_TFRZ)
                SP1(F->qflx_soliddew_to_top_layer, p) = fabs(ev);
            else
                SP1(F->qflx_liqdew_to_top_layer, p) = fabs(ev);
        }
    }

    /* :360-410 evaporation limited by the water of the top layer */
    for (fp = 0; fp < num_p; ++fp) {
        const int p = filter_p ? filter_p[fp] : L->begp_alloc + fp;
        const int c = SP1(P->column, p);
        const int jt = SC1(S->snl, c) + 1;
        const double fs = SC1(S->frac_sno_eff, c), fh = SC1(S->frac_h2osfc, c);
        const double liq = SSNO(S->h2osoi_liq, c, jt), ice = SSNO(S->h2osoi_ice, c, jt);
        if (jt < 1) {
            const double lim = (ice + liq) / (fs * dtime);
            if (SP1(F->qflx_ev_snow, p) > lim) {
                const double dem = SP1(F->qflx_ev_snow, p);
                SP1(F->qflx_ev_snow, p) = lim;
                SP1(F->qflx_evap_soi, p) = SP1(F->qflx_evap_soi, p) - fs * (dem - lim);
                SP1(F->qflx_liqevap_from_top_layer, p) = sf_max(liq / (fs * dtime), 0.0);
                SP1(F->qflx_solidevap_from_top_layer, p) = sf_max(ice / (fs * dtime), 0.0);
                SP1(F->eflx_sh_grnd, p) = SP1(F->eflx_sh_grnd, p) + fs * (dem - lim) * SC1(S->htvp, c);
            }
        }
        if (jt == 1 && fh < 1.0) {
            const double lim = ice / (dtime * (1.0 - fh));
            if (SP1(F->qflx_solidevap_from_top_layer, p) >= lim) {
                const double dem = SP1(F->qflx_solidevap_from_top_layer, p);
                SP1(F->qflx_solidevap_from_top_layer, p) = lim;
                SP1(F->qflx_liqevap_from_top_layer, p) = SP1(F->qflx_liqevap_from_top_layer, p) + (dem - lim);
            }
        }
    }

    /* :412-471 ground heat flux, total and history fluxes */
    for (fp = 0; fp < num_p; ++fp) {
        const int p = filter_p ? filter_p[fp] : L->begp_alloc + fp;
        const int c = SP1(P->column, p);
        const int snl = SC1(S->snl, c);
        const double fs = SC1(S->frac_sno_eff, c), fh = SC1(S->frac_h2osfc, c), emg = SC1(S->emg, c);
        const double lw_grnd = (fs * p4(SSNO(F->t_ssbef, c, snl + 1))
                                + (1.0 - fs - fh) * p4(SSNO(F->t_ssbef, c, 1))
                                + fh * p4(SC1(F->t_h2osfc_bef, c)));
        const double tg0 = t_grnd0[CIDX(c)], ti = tinc[CIDX(c)];
        const double htvp = SC1(S->htvp, c);
        const int fvn = SP1(P->frac_veg_nosno, p);
        SP1(F->eflx_soil_grnd, p) = ((1.0 - fs) * SP1(P->sabg_soil, p) + fs * SP1(P->sabg_snow, p))
                                    + SP1(P->dlrad, p)
                                    + (double)(1 - fvn) * emg * SC1(S->forc_lwrad, c)
                                    - emg * SF_SB * lw_grnd - emg * SF_SB * p3(tg0) * (4.0 * ti)
                                    - (SP1(F->eflx_sh_grnd, p) + SP1(F->qflx_evap_soi, p) * htvp);
        SP1(F->eflx_soil_grnd_r, p) = SP1(F->eflx_soil_grnd, p);
        SP1(F->eflx_sh_tot, p) = SP1(F->eflx_sh_veg, p) + SP1(F->eflx_sh_grnd, p);
        SP1(F->eflx_sh_tot, p) = SP1(F->eflx_sh_tot, p) + SP1(F->eflx_sh_stem, p);
        SP1(F->qflx_evap_tot, p) = SP1(F->qflx_evap_veg, p) + SP1(F->qflx_evap_soi, p);
        SP1(F->eflx_lh_tot, p) = SF_HVAP * SP1(F->qflx_evap_veg, p) + htvp * SP1(F->qflx_evap_soi, p);
        SP1(F->eflx_lh_tot_r, p) = SP1(F->eflx_lh_tot, p);
        SP1(F->eflx_sh_tot_r, p) = SP1(F->eflx_sh_tot, p);
        SP1(F->qflx_evap_can, p) = SP1(F->qflx_evap_veg, p) - SP1(F->qflx_tran_veg, p);
        SP1(F->eflx_lh_vege, p) = (SP1(F->qflx_evap_veg, p) - SP1(F->qflx_tran_veg, p)) * SF_HVAP;
        SP1(F->eflx_lh_vegt, p) = SP1(F->qflx_tran_veg, p) * SF_HVAP;
        SP1(F->eflx_lh_grnd, p) = SP1(F->qflx_evap_soi, p) * htvp;
    }

    /* :476-499 soil energy conservation error */
    for (fp = 0; fp < num_p; ++fp) {
        const int p = filter_p ? filter_p[fp] : L->begp_alloc + fp;
        const int c = SP1(P->column, p);
        double e = SP1(F->eflx_soil_grnd, p) - SC1(S->xmf, c) - SC1(S->xmf_h2osfc, c)
                   - SC1(S->frac_h2osfc, c) * (SC1(S->t_h2osfc, c) - SC1(F->t_h2osfc_bef, c))
                         * (SC1(S->c_h2osfc, c) / dtime);
        e = e + SC1(S->eflx_h2osfc_to_snow, c);
        SP1(F->errsoi_patch, p) = e;
    }
    for (j = -cfg->nlevsno + 1; j <= cfg->nlevgrnd; ++j)
        for (fp = 0; fp < num_p; ++fp) {
            const int p = filter_p ? filter_p[fp] : L->begp_alloc + fp;
            const int c = SP1(P->column, p);
            const int snl = SC1(S->snl, c);
            if (j >= snl + 1 && j < 1)
                SP1(F->errsoi_patch, p) = SP1(F->errsoi_patch, p)
                    - SC1(S->frac_sno_eff, c) * (SSNO(S->t_soisno, c, j) - SSNO(F->t_ssbef, c, j))
                          / SSNO(S->fact, c, j);
            if (j >= 1)
                SP1(F->errsoi_patch, p) = SP1(F->errsoi_patch, p)
                    - (SSNO(S->t_soisno, c, j) - SSNO(F->t_ssbef, c, j)) / SSNO(S->fact, c, j);
        }

    /* :526-570 outgoing longwave, skin temperature */
    for (fp = 0; fp < num_p; ++fp) {
        const int p = filter_p ? filter_p[fp] : L->begp_alloc + fp;
        const int c = SP1(P->column, p);
        const int snl = SC1(S->snl, c);
        const double fs = SC1(S->frac_sno_eff, c), fh = SC1(S->frac_h2osfc, c), emg = SC1(S->emg, c);
        const double lw_grnd = (fs * p4(SSNO(F->t_ssbef, c, snl + 1))
                                + (1.0 - fs - fh) * p4(SSNO(F->t_ssbef, c, 1))
                                + fh * p4(SC1(F->t_h2osfc_bef, c)));
        const int fvn = SP1(P->frac_veg_nosno, p);
        const double fl = SC1(S->forc_lwrad, c);
        SP1(F->eflx_lwrad_out, p) = SP1(F->ulrad, p) + (double)(1 - fvn) * (1.0 - emg) * fl
                                    + (double)(1 - fvn) * emg * SF_SB * lw_grnd
                                    + 4.0 * emg * SF_SB * p3(t_grnd0[CIDX(c)]) * tinc[CIDX(c)];
        if (fvn == 0) SP1(F->t_skin, p) = sqrt(sqrt(lw_grnd));
        SP1(F->eflx_lwrad_net, p) = SP1(F->eflx_lwrad_out, p) - fl;
        SP1(F->eflx_lwrad_net_r, p) = SP1(F->eflx_lwrad_out, p) - fl;
        SP1(F->eflx_lwrad_out_r, p) = SP1(F->eflx_lwrad_out, p);
    }

    /* :574-576 p2c of errsoi (subgridAveMod.F90:312-318) */
    for (fc = 0; fc < num_c; ++fc) {
        const int c = filter_c ? filter_c[fc] : L->begc_alloc + fc;
        double s = 0.0;
        int p;
        for (p = SC1(S->patchi, c); p <= SC1(S->patchf, c); ++p)
            if (SP1(F->active, p)) s = s + SP1(F->errsoi_patch, p) * SP1(P->wtcol, p);
        SC1(F->errsoi_col, c) = s;
    }
#undef CIDX
    free(tinc);
    free(t_grnd0);
    return 0;
}

/* BalanceCheckMod.F90:1101-1119: max |errsoi_col| over the filter and the
 * lowest column index holding it. */
void oracle_soil_balance(const ctsm_layout *L, const ctsm_soilflux_state *F, int num_c, const int *filter_c,
                         double *errmax, int *index)
{
    double m = -1.0;
    int idx = 0, fc;
    for (fc = 0; fc < num_c; ++fc) {
        const int c = filter_c ? filter_c[fc] : L->begc_alloc + fc;
        const double a = fabs(F->errsoi_col[c - L->begc_alloc]);
        if (a > m || (a == m && c < idx)) {
            m = a;
            idx = c;
        }
    }
    *errmax = m < 0.0 ? 0.0 : m;
    *index = idx;
}

/* EnergyBalanceCheck's patch errors (BalanceCheckMod.F90:966-1003) for
 * non-urban patches: errseb = sabv + sabg_chk + forc_lwrad - eflx_lwrad_out
 * - eflx_sh_tot - eflx_lh_tot - eflx_soil_grnd - dhsdt_canopy and errlon =
 * eflx_lwrad_out - eflx_lwrad_net - forc_lwrad, left to right as written;
 * 0 for inactive patches.  Writes errseb[fp] / errlon[fp] per filter entry
 * and the max |err| with the lowest patch index holding it (ties). */
void oracle_energy_balance(const ctsm_layout *L, const ctsm_col_state *S, const ctsm_patch_state *Pt,
                           const ctsm_canopy_state *Cn, const ctsm_soilflux_state *F, int num_p, const int *filter_p,
                           double *errseb, double *errlon, double *seb_max, int *seb_patch, double *lon_max,
                           int *lon_patch)
{
    double ms = -1.0, ml = -1.0;
    int is = 0, il = 0, fp;
    for (fp = 0; fp < num_p; ++fp) {
        const int p = filter_p ? filter_p[fp] : L->begp_alloc + fp;
        const long pp = p - L->begp_alloc;
        double seb = 0.0, lon = 0.0;
        if (Cn->active[pp]) {
            const long cc = Cn->column[pp] - L->begc_alloc;
            seb = Cn->sabv[pp] + Pt->sabg_chk[pp] + S->forc_lwrad[cc] - F->eflx_lwrad_out[pp] - F->eflx_sh_tot[pp]
                  - F->eflx_lh_tot[pp] - F->eflx_soil_grnd[pp] - Cn->dhsdt_canopy[pp];
            lon = F->eflx_lwrad_out[pp] - F->eflx_lwrad_net[pp] - S->forc_lwrad[cc];
        }
        errseb[fp] = seb;
        errlon[fp] = lon;
        if (fabs(seb) > ms || (fabs(seb) == ms && p < is)) { ms = fabs(seb); is = p; }
        if (fabs(lon) > ml || (fabs(lon) == ml && p < il)) { ml = fabs(lon); il = p; }
    }
    *seb_max = ms < 0.0 ? 0.0 : ms;
    *seb_patch = is;
    *lon_max = ml < 0.0 ? 0.0 : ml;
    *lon_patch = il;
}
