// soilfluxes.hip -- SoilFluxes and the soil energy balance check on gfx950.
//
// Replaces SoilFluxes (SoilFluxesMod.F90:38-518) for soil / crop patches and
// the soil part of EnergyBalanceCheck (BalanceCheckMod.F90:1101-1119), and
// saves the before-step temperatures of BiogeophysPreFluxCalcsMod.F90:302-331.
// All of it is a streaming per-patch / per-column update (HBM-bound, no
// solve), so each kernel is one lane per patch or column with the arrays
// read and written once, column-fastest as the reference lays them out:
//
//   k_save_before     per column: t_ssbef(c, :) = t_soisno(c, :), t_h2osfc_bef
//   k_soilfluxes      per patch: the reference's patch loops :277-570 fused in
//                     their order (a patch's values depend only on its own
//                     column), ending with errsoi_patch
//   k_errsoi_p2c      per column: p2c_1d_filter of errsoi (subgridAveMod.F90:312-318)
//   k_balance_max /   max |errsoi_col| over the filter (64-bit atomic max on the
//   k_balance_loc     bits of a non-negative double, which order as the values)
//                     and its lowest column index (atomic min), as maxloc
//
// Every expression keeps the reference's operand order; integer powers by
// squaring as the oracle writes them.
#include "common.h"
#include "trace.h"

namespace ctsm {
namespace sflux {

constexpr double HVAP = 2.501e6;

struct SView {
    int begc, ldc, begp, ldp, nlevsno, nlevgrnd;
    __device__ __forceinline__ long c1(int c) const { return (long)(c - begc); }
    __device__ __forceinline__ long sno(int c, int j) const { return (long)(c - begc) + (long)(j + nlevsno - 1) * ldc; }
    __device__ __forceinline__ long p1(int p) const { return (long)(p - begp); }
};

__device__ __forceinline__ double p4(double x)
{
    const double x2 = x * x;
    return x2 * x2;
}
__device__ __forceinline__ double p3(double x) { return (x * x) * x; }

__global__ void __launch_bounds__(256) k_save_before(SView V, ctsm_col_state S, ctsm_soilflux_state F, int n,
                                                     const int32_t *__restrict__ filter)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) return;
    for (int j = -V.nlevsno + 1; j <= V.nlevgrnd; ++j) F.t_ssbef[V.sno(c, j)] = S.t_soisno[V.sno(c, j)];
    F.t_h2osfc_bef[V.c1(c)] = S.t_h2osfc[V.c1(c)];
}

// SoilFluxesMod.F90:257-570 for one patch.
__global__ void __launch_bounds__(256) k_soilfluxes(SView V, ctsm_col_state S, ctsm_patch_state P,
                                                    ctsm_soilflux_state F, int n, const int32_t *__restrict__ filter,
                                                    double dtime, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = filter ? filter[i] : V.begp + i;
    if (p < V.begp || p >= V.begp + V.ldp) { report_error(err, p, CTSM_E_ARG); return; }
    const long pp = V.p1(p);
    const int c = P.column[pp];
    if (c < V.begc || c >= V.begc + V.ldc) { report_error(err, p, CTSM_E_ARG); return; }
    const long cc = V.c1(c);
    const int snl = S.snl[cc];
    if (snl < -V.nlevsno || snl > 0) { report_error(err, p, CTSM_E_ARG); return; }
    const int jt = snl + 1;
    const double fs = S.frac_sno_eff[cc], fh = S.frac_h2osfc[cc];
    const double tb_top = F.t_ssbef[V.sno(c, jt)], tb1 = F.t_ssbef[V.sno(c, 1)], thb = F.t_h2osfc_bef[cc];

    // :257-273 previous ground temperature and its increment
    double t_grnd0;
    if (snl < 0) t_grnd0 = fs * tb_top + (1 - fs - fh) * tb1 + fh * thb;
    else t_grnd0 = (1 - fh) * tb1 + fh * thb;
    const double tinc = S.t_grnd[cc] - t_grnd0;

    // :277-294 correction to the present ground temperature
    const double cgs = F.cgrnds[pp], cgl = F.cgrndl[pp];
    double sh_grnd = F.eflx_sh_grnd[pp] + tinc * cgs;
    double evap_soi = F.qflx_evap_soi[pp] + tinc * cgl;
    double ev_snow = F.qflx_ev_snow[pp] + tinc * cgl;
    F.qflx_ev_soil[pp] = F.qflx_ev_soil[pp] + tinc * cgl;
    F.qflx_ev_h2osfc[pp] = F.qflx_ev_h2osfc[pp] + tinc * cgl;

    // :297-358 partition into liquid / solid evaporation and dew
    const double liq = S.h2osoi_liq[V.sno(c, jt)], ice = S.h2osoi_ice[V.sno(c, jt)];
    double liqevap = 0.0, solidevap = 0.0, soliddew = 0.0, liqdew = 0.0;
    if (ev_snow >= 0.0) {
        if ((liq + ice) > 0.0) liqevap = fmax90(ev_snow * (liq / (liq + ice)), 0.0);
        else liqevap = 0.0;
        solidevap = ev_snow - liqevap;
    } else {
        if (S.t_grnd[cc] < TFRZ) soliddew = fabs(ev_snow);
        else liqdew = fabs(ev_snow);
    }

    // :360-410 limits from the water of the top layer
    const double htvp = S.htvp[cc];
    if (jt < 1) {
        const double lim = (ice + liq) / (fs * dtime);
        if (ev_snow > lim) {
            const double dem = ev_snow;
            ev_snow = lim;
            evap_soi = evap_soi - fs * (dem - lim);
            liqevap = fmax90(liq / (fs * dtime), 0.0);
            solidevap = fmax90(ice / (fs * dtime), 0.0);
            sh_grnd = sh_grnd + fs * (dem - lim) * htvp;
        }
    }
    if (jt == 1 && fh < 1.0) {
        const double lim = ice / (dtime * (1.0 - fh));
        if (solidevap >= lim) {
            const double dem = solidevap;
            solidevap = lim;
            liqevap = liqevap + (dem - lim);
        }
    }
    F.eflx_sh_grnd[pp] = sh_grnd;
    F.qflx_evap_soi[pp] = evap_soi;
    F.qflx_ev_snow[pp] = ev_snow;
    F.qflx_liqevap_from_top_layer[pp] = liqevap;
    F.qflx_solidevap_from_top_layer[pp] = solidevap;
    F.qflx_soliddew_to_top_layer[pp] = soliddew;
    F.qflx_liqdew_to_top_layer[pp] = liqdew;

    // :412-471 ground heat flux, totals, history terms
    const double emg = S.emg[cc], forc_lwrad = S.forc_lwrad[cc];
    const double lw_grnd = (fs * p4(tb_top) + (1.0 - fs - fh) * p4(tb1) + fh * p4(thb));
    const double fveg = (double)(1 - P.frac_veg_nosno[pp]);
    const double soil_grnd = ((1.0 - fs) * P.sabg_soil[pp] + fs * P.sabg_snow[pp]) + P.dlrad[pp]
                             + fveg * emg * forc_lwrad - emg * SB * lw_grnd - emg * SB * p3(t_grnd0) * (4.0 * tinc)
                             - (sh_grnd + evap_soi * htvp);
    F.eflx_soil_grnd[pp] = soil_grnd;
    F.eflx_soil_grnd_r[pp] = soil_grnd;
    double sh_tot = F.eflx_sh_veg[pp] + sh_grnd;
    sh_tot = sh_tot + F.eflx_sh_stem[pp];
    F.eflx_sh_tot[pp] = sh_tot;
    F.eflx_sh_tot_r[pp] = sh_tot;
    const double evap_veg = F.qflx_evap_veg[pp], tran_veg = F.qflx_tran_veg[pp];
    F.qflx_evap_tot[pp] = evap_veg + evap_soi;
    const double lh_tot = HVAP * evap_veg + htvp * evap_soi;
    F.eflx_lh_tot[pp] = lh_tot;
    F.eflx_lh_tot_r[pp] = lh_tot;
    F.qflx_evap_can[pp] = evap_veg - tran_veg;
    F.eflx_lh_vege[pp] = (evap_veg - tran_veg) * HVAP;
    F.eflx_lh_vegt[pp] = tran_veg * HVAP;
    F.eflx_lh_grnd[pp] = evap_soi * htvp;

    // :476-499 soil energy conservation error
    double e = soil_grnd - S.xmf[cc] - S.xmf_h2osfc[cc] - fh * (S.t_h2osfc[cc] - thb) * (S.c_h2osfc[cc] / dtime);
    e = e + S.eflx_h2osfc_to_snow[cc];
    for (int j = jt; j <= 0; ++j)
        e = e - fs * (S.t_soisno[V.sno(c, j)] - F.t_ssbef[V.sno(c, j)]) / S.fact[V.sno(c, j)];
    for (int j = 1; j <= V.nlevgrnd; ++j)
        e = e - (S.t_soisno[V.sno(c, j)] - F.t_ssbef[V.sno(c, j)]) / S.fact[V.sno(c, j)];
    F.errsoi_patch[pp] = e;

    // :526-570 outgoing longwave, skin temperature of bare ground
    const double lw_out = F.ulrad[pp] + fveg * (1.0 - emg) * forc_lwrad + fveg * emg * SB * lw_grnd
                          + 4.0 * emg * SB * p3(t_grnd0) * tinc;
    F.eflx_lwrad_out[pp] = lw_out;
    if (P.frac_veg_nosno[pp] == 0) F.t_skin[pp] = sqrt(sqrt(lw_grnd));
    F.eflx_lwrad_net[pp] = lw_out - forc_lwrad;
    F.eflx_lwrad_net_r[pp] = lw_out - forc_lwrad;
    F.eflx_lwrad_out_r[pp] = lw_out;
}

// :574-576, p2c_1d_filter (subgridAveMod.F90:312-318): a column per lane.
// The lanes of a wave take the wave's patches 64 at a time (coalesced loads),
// and each column's lane adds its patches' weighted errors from LDS in patch
// order (inactive patches skipped), the same sum as a loop over the column's
// patches.
__global__ void __launch_bounds__(256) k_errsoi_p2c(SView V, ctsm_col_state S, ctsm_patch_state P,
                                                    ctsm_soilflux_state F, int n, const int32_t *__restrict__ filter,
                                                    ErrWord *err)
{
    __shared__ int s_pi[4][64], s_off[4][64];
    __shared__ double s_x[4][64];
    __shared__ int s_on[4][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = V.begc, n_own = 0, pi = 0;
    bool ok = false;
    if (i < n) {
        c = filter ? filter[i] : V.begc + i;
        if (c < V.begc || c >= V.begc + V.ldc) {
            report_error(err, c, CTSM_E_ARG);
        } else {
            const long cc = V.c1(c);
            const int p0 = S.patchi[cc], p1 = S.patchf[cc];
            if (p1 >= p0 && (p0 < V.begp || p1 >= V.begp + V.ldp)) {
                report_error(err, c, CTSM_E_ARG);
            } else {
                ok = true;
                pi = p0;
                n_own = p1 >= p0 ? p1 - p0 + 1 : 0;
            }
        }
    }
    int incl = n_own;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    const int off = incl - n_own;
    const int total = __shfl(incl, 63);
    s_pi[w][lane] = pi;
    s_off[w][lane] = off;
    double s = 0.0;
    wave_sync();
    for (int k0 = 0; k0 < total; k0 += 64) {
        const int k = k0 + lane;
        if (k < total) {
            int lo = 0, hi = 63;       // the column: the last lane whose list starts at or before k
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_off[w][mid] <= k) lo = mid;
                else hi = mid - 1;
            }
            const long pp = V.p1(s_pi[w][lo] + (k - s_off[w][lo]));
            const int on = F.active[pp];
            s_on[w][lane] = on;
            s_x[w][lane] = on ? F.errsoi_patch[pp] * P.wtcol[pp] : 0.0;
        }
        wave_sync();
        if (n_own > 0) {
            const int a = (off > k0 ? off : k0) - k0;
            const int e = off + n_own, b = (e < k0 + 64 ? e : k0 + 64) - k0;
            for (int t = a; t < b; ++t)
                if (s_on[w][t]) s = s + s_x[w][t];
        }
        wave_sync();
    }
    if (ok) F.errsoi_col[V.c1(c)] = s;
}

// Balance reduction word: [0] bits of the max |errsoi_col|, [1] its column.
__global__ void k_balance_init(unsigned long long *w)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        w[0] = 0ull;
        w[1] = 0x7fffffffull;
    }
}

__global__ void __launch_bounds__(256) k_balance_max(SView V, const double *__restrict__ errsoi_col, int n,
                                                     const int32_t *__restrict__ filter, unsigned long long *w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) return;
    const double a = fabs(errsoi_col[V.c1(c)]);
    atomicMax(&w[0], (unsigned long long)__double_as_longlong(a));
}

__global__ void __launch_bounds__(256) k_balance_loc(SView V, const double *__restrict__ errsoi_col, int n,
                                                     const int32_t *__restrict__ filter, unsigned long long *w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) return;
    const double a = fabs(errsoi_col[V.c1(c)]);
    if ((unsigned long long)__double_as_longlong(a) == w[0]) atomicMin(&w[1], (unsigned long long)c);
}

// EnergyBalanceCheck's patch errors (BalanceCheckMod.F90:966-1003), non-urban
// patches, left to right as the reference writes them; 0 when inactive.
// w[0..1]: max |errseb| (bits) and its patch, w[2..3]: the same for errlon.
__device__ __forceinline__ void energy_errors(const SView &V, const ctsm_col_state &S, const ctsm_patch_state &Pt,
                                              const ctsm_canopy_state &Cn, const ctsm_soilflux_state &F, int p,
                                              double &seb, double &lon)
{
    const long pp = V.p1(p);
    seb = 0.0;
    lon = 0.0;
    if (Cn.active[pp]) {
        const long cc = V.c1(Cn.column[pp]);
        seb = Cn.sabv[pp] + Pt.sabg_chk[pp] + S.forc_lwrad[cc] - F.eflx_lwrad_out[pp] - F.eflx_sh_tot[pp] -
              F.eflx_lh_tot[pp] - F.eflx_soil_grnd[pp] - Cn.dhsdt_canopy[pp];
        lon = F.eflx_lwrad_out[pp] - F.eflx_lwrad_net[pp] - S.forc_lwrad[cc];
    }
}

__global__ void k_energy_init(unsigned long long *w)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        w[0] = 0ull;
        w[1] = 0x7fffffffull;
        w[2] = 0ull;
        w[3] = 0x7fffffffull;
    }
}

__global__ void __launch_bounds__(256) k_energy_max(SView V, ctsm_col_state S, ctsm_patch_state Pt,
                                                    ctsm_canopy_state Cn, ctsm_soilflux_state F, int n,
                                                    const int32_t *__restrict__ filter, double *errseb,
                                                    double *errlon, unsigned long long *w, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = filter ? filter[i] : V.begp + i;
    if (p < V.begp || p >= V.begp + V.ldp) { report_error(err, p, CTSM_E_ARG); return; }
    const int c = Cn.column[V.p1(p)];
    if (Cn.active[V.p1(p)] && (c < V.begc || c >= V.begc + V.ldc)) { report_error(err, p, CTSM_E_ARG); return; }
    double seb, lon;
    energy_errors(V, S, Pt, Cn, F, p, seb, lon);
    if (errseb) errseb[i] = seb;
    if (errlon) errlon[i] = lon;
    atomicMax(&w[0], (unsigned long long)__double_as_longlong(fabs(seb)));
    atomicMax(&w[2], (unsigned long long)__double_as_longlong(fabs(lon)));
}

__global__ void __launch_bounds__(256) k_energy_loc(SView V, ctsm_col_state S, ctsm_patch_state Pt,
                                                    ctsm_canopy_state Cn, ctsm_soilflux_state F, int n,
                                                    const int32_t *__restrict__ filter, unsigned long long *w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = filter ? filter[i] : V.begp + i;
    if (p < V.begp || p >= V.begp + V.ldp) return;
    const int c = Cn.column[V.p1(p)];
    if (Cn.active[V.p1(p)] && (c < V.begc || c >= V.begc + V.ldc)) return;
    double seb, lon;
    energy_errors(V, S, Pt, Cn, F, p, seb, lon);
    if ((unsigned long long)__double_as_longlong(fabs(seb)) == w[0]) atomicMin(&w[1], (unsigned long long)p);
    if ((unsigned long long)__double_as_longlong(fabs(lon)) == w[2]) atomicMin(&w[3], (unsigned long long)p);
}


// ComputeLiqIceMassNonLake (TotalWaterAndHeatMod.F90:200-393) for a soil or
// crop column, sums in the reference's order: the canopy water's
// p2c_1d_filter (subgridAveMod.F90:312-318), plant storage, unlayered snow,
// the snow layers above snl, the aquifer, surface water, the ground layers
// with excess ice, then the dynbal baselines when subtracted.
__global__ void __launch_bounds__(256) k_water_mass(SView V, ctsm_col_state S, ctsm_canopy_state Cn,
                                                    ctsm_water_balance_state W, int n,
                                                    const int32_t *__restrict__ filter, double baseline,
                                                    int subtract, double *__restrict__ mass, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) { report_error(err, c, CTSM_E_ARG); return; }
    const long cc = V.c1(c);
    const int pi = S.patchi[cc], pf = S.patchf[cc], snl = S.snl[cc];
    if ((pf >= pi && (pi < V.begp || pf >= V.begp + V.ldp)) || snl < -V.nlevsno || snl > 0) {
        report_error(err, c, CTSM_E_ARG);
        return;
    }
    double lc = 0.0, sc = 0.0;
    for (int p = pi; p <= pf; ++p) {
        const long pp = V.p1(p);
        if (Cn.active[pp]) {
            lc = lc + Cn.liqcan[pp] * Cn.wtcol[pp];
            sc = sc + Cn.snocan[pp] * Cn.wtcol[pp];
        }
    }
    double liquid = 0.0, ice = 0.0;
    liquid = liquid + lc + W.total_plant_stored_h2o[cc];
    ice = ice + sc;
    ice = ice + S.h2osno_no_layers[cc];
    for (int j = snl + 1; j <= 0; ++j) {
        liquid = liquid + S.h2osoi_liq[V.sno(c, j)];
        ice = ice + S.h2osoi_ice[V.sno(c, j)];
    }
    if (W.hydrologically_active[cc]) liquid = liquid + (W.wa[cc] - baseline);
    liquid = liquid + S.h2osfc[cc];
    for (int j = 1; j <= V.nlevgrnd; ++j) {
        liquid = liquid + S.h2osoi_liq[V.sno(c, j)];
        const double xi = W.excess_ice ? W.excess_ice[cc + (long)(j - 1) * V.ldc] : 0.0;
        ice = ice + S.h2osoi_ice[V.sno(c, j)] + xi;
    }
    if (subtract) {
        liquid = liquid - W.dynbal_baseline_liq[cc];
        ice = ice - W.dynbal_baseline_ice[cc];
    }
    mass[cc] = liquid + ice;
}

// errh2o_col (BalanceCheckMod.F90:588-610), left to right as written; 0 for
// inactive columns.
__device__ __forceinline__ double water_error(const ctsm_water_balance_state &W, long cc, double dtime)
{
    if (!W.col_active[cc]) return 0.0;
    return W.endwb[cc] - W.begwb[cc] -
           (W.forc_rain[cc] + W.forc_snow[cc] + W.qflx_floodc[cc] + W.qflx_sfc_irrig[cc] +
            W.qflx_glcice_dyn_water_flux[cc] - W.qflx_evap_tot[cc] - W.qflx_surf[cc] - W.qflx_qrgwl[cc] -
            W.qflx_drain[cc] - W.qflx_drain_perched[cc] - W.qflx_ice_runoff[cc] - W.qflx_snwcp_discarded_liq[cc] -
            W.qflx_snwcp_discarded_ice[cc]) *
               dtime;
}

// w[0..1] as k_balance_max / k_balance_loc: bits of the max |errh2o|, its column.
__global__ void __launch_bounds__(256) k_water_balance_max(SView V, ctsm_water_balance_state W, int n,
                                                           const int32_t *__restrict__ filter, double dtime,
                                                           unsigned long long *w, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) { report_error(err, c, CTSM_E_ARG); return; }
    const long cc = V.c1(c);
    const double e = water_error(W, cc, dtime);
    if (W.errh2o) W.errh2o[cc] = e;
    atomicMax(&w[0], (unsigned long long)__double_as_longlong(fabs(e)));
}

__global__ void __launch_bounds__(256) k_water_balance_loc(SView V, ctsm_water_balance_state W, int n,
                                                           const int32_t *__restrict__ filter, double dtime,
                                                           unsigned long long *w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) return;
    const double e = water_error(W, V.c1(c), dtime);
    if ((unsigned long long)__double_as_longlong(fabs(e)) == w[0]) atomicMin(&w[1], (unsigned long long)c);
}

// errsol (BalanceCheckMod.F90:964-970) for non-urban patches; 0 when inactive.
struct SolarIn {
    const double *fsa, *fsr, *solad, *solai;
    int begg, ldg;
};

__device__ __forceinline__ double solar_error(const SView &V, const ctsm_canopy_state &Cn, const SolarIn &R, int p)
{
    const long pp = V.p1(p);
    if (!Cn.active[pp]) return 0.0;
    const long cc = V.c1(Cn.column[pp]);
    const long gg = (long)(Cn.gridcell[pp] - R.begg);
    return R.fsa[pp] + R.fsr[pp] - (R.solad[cc] + R.solad[cc + V.ldc] + R.solai[gg] + R.solai[gg + R.ldg]);
}

__device__ __forceinline__ bool solar_patch_ok(const SView &V, const ctsm_canopy_state &Cn, const SolarIn &R, int p)
{
    if (p < V.begp || p >= V.begp + V.ldp) return false;
    const long pp = V.p1(p);
    if (!Cn.active[pp]) return true;
    const int c = Cn.column[pp], g = Cn.gridcell[pp];
    return c >= V.begc && c < V.begc + V.ldc && g >= R.begg && g < R.begg + R.ldg;
}

__global__ void __launch_bounds__(256) k_solar_max(SView V, ctsm_canopy_state Cn, SolarIn R, int n,
                                                   const int32_t *__restrict__ filter, double *errsol,
                                                   unsigned long long *w, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = filter ? filter[i] : V.begp + i;
    if (!solar_patch_ok(V, Cn, R, p)) { report_error(err, p, CTSM_E_ARG); return; }
    const double e = solar_error(V, Cn, R, p);
    if (errsol) errsol[i] = e;
    atomicMax(&w[0], (unsigned long long)__double_as_longlong(fabs(e)));
}

__global__ void __launch_bounds__(256) k_solar_loc(SView V, ctsm_canopy_state Cn, SolarIn R, int n,
                                                   const int32_t *__restrict__ filter, unsigned long long *w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = filter ? filter[i] : V.begp + i;
    if (!solar_patch_ok(V, Cn, R, p)) return;
    const double e = solar_error(V, Cn, R, p);
    if ((unsigned long long)__double_as_longlong(fabs(e)) == w[0]) atomicMin(&w[1], (unsigned long long)p);
}

}  // namespace sflux

static sflux::SView sview(const ctsm_layout &L, const ctsm_config &cfg)
{
    return sflux::SView{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, cfg.nlevsno, cfg.nlevgrnd};
}

void launch_save_before(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_col_state &S,
                        const ctsm_soilflux_state &F, int n, const int32_t *filter)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(sflux::k_save_before, dim3((n + 255) / 256), dim3(256), 0, st, sview(L, cfg), S, F, n, filter);
}

void launch_soilfluxes(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_col_state &S,
                       const ctsm_patch_state &P, const ctsm_soilflux_state &F, int nc, const int32_t *filterc,
                       int np, const int32_t *filterp, double dtime, ErrWord *err)
{
    const GptlRange range("bgp2");
    const sflux::SView V = sview(L, cfg);
    if (np > 0)
        hipLaunchKernelGGL(sflux::k_soilfluxes, dim3((np + 255) / 256), dim3(256), 0, st, V, S, P, F, np, filterp,
                           dtime, err);
    if (nc > 0)
        hipLaunchKernelGGL(sflux::k_errsoi_p2c, dim3((nc + 255) / 256), dim3(256), 0, st, V, S, P, F, nc, filterc,
                           err);
}

void launch_balance(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const double *errsoi_col, int n,
                    const int32_t *filter, unsigned long long *w)
{
    const GptlRange range("balchk");
    const sflux::SView V = sview(L, cfg);
    hipLaunchKernelGGL(sflux::k_balance_init, dim3(1), dim3(64), 0, st, w);
    if (n <= 0) return;
    hipLaunchKernelGGL(sflux::k_balance_max, dim3((n + 255) / 256), dim3(256), 0, st, V, errsoi_col, n, filter, w);
    hipLaunchKernelGGL(sflux::k_balance_loc, dim3((n + 255) / 256), dim3(256), 0, st, V, errsoi_col, n, filter, w);
}

void launch_energy_balance(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_col_state &S,
                           const ctsm_patch_state &P, const ctsm_canopy_state &Cn, const ctsm_soilflux_state &F, int n,
                           const int32_t *filter, double *errseb, double *errlon, unsigned long long *w, ErrWord *err)
{
    const GptlRange range("balchk");
    const sflux::SView V = sview(L, cfg);
    hipLaunchKernelGGL(sflux::k_energy_init, dim3(1), dim3(64), 0, st, w);
    if (n <= 0) return;
    hipLaunchKernelGGL(sflux::k_energy_max, dim3((n + 255) / 256), dim3(256), 0, st, V, S, P, Cn, F, n, filter,
                       errseb, errlon, w, err);
    hipLaunchKernelGGL(sflux::k_energy_loc, dim3((n + 255) / 256), dim3(256), 0, st, V, S, P, Cn, F, n, filter, w);
}

void launch_water_mass(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_col_state &S,
                       const ctsm_canopy_state &Cn, const ctsm_water_balance_state &W, int n, const int32_t *filter,
                       double baseline, int subtract, double *mass, ErrWord *err)
{
    const GptlRange range("balchk");
    if (n <= 0) return;
    hipLaunchKernelGGL(sflux::k_water_mass, dim3((n + 255) / 256), dim3(256), 0, st, sview(L, cfg), S, Cn, W, n,
                       filter, baseline, subtract, mass, err);
}

void launch_water_balance(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg,
                          const ctsm_water_balance_state &W, int n, const int32_t *filter, double dtime,
                          unsigned long long *w, ErrWord *err)
{
    const GptlRange range("balchk");
    const sflux::SView V = sview(L, cfg);
    hipLaunchKernelGGL(sflux::k_balance_init, dim3(1), dim3(64), 0, st, w);
    if (n <= 0) return;
    hipLaunchKernelGGL(sflux::k_water_balance_max, dim3((n + 255) / 256), dim3(256), 0, st, V, W, n, filter, dtime,
                       w, err);
    hipLaunchKernelGGL(sflux::k_water_balance_loc, dim3((n + 255) / 256), dim3(256), 0, st, V, W, n, filter, dtime,
                       w);
}

void launch_solar_balance(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_canopy_state &Cn,
                          int n, const int32_t *filter, const double *fsa, const double *fsr, const double *solad,
                          const double *solai, double *errsol, unsigned long long *w, ErrWord *err)
{
    const GptlRange range("balchk");
    const sflux::SView V = sview(L, cfg);
    const sflux::SolarIn R{fsa, fsr, solad, solai, L.begg_alloc, L.ldg};
    hipLaunchKernelGGL(sflux::k_balance_init, dim3(1), dim3(64), 0, st, w);
    if (n <= 0) return;
    hipLaunchKernelGGL(sflux::k_solar_max, dim3((n + 255) / 256), dim3(256), 0, st, V, Cn, R, n, filter, errsol, w,
                       err);
    hipLaunchKernelGGL(sflux::k_solar_loc, dim3((n + 255) / 256), dim3(256), 0, st, V, Cn, R, n, filter, w);
}

}  // namespace ctsm
