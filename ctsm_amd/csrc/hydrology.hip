// hydrology.hip -- the surface water input chain of HydrologyNoDrainage on gfx950.
//
// Replaces HydrologyNoDrainageMod.F90:306-336 for soil / crop columns:
// SaturatedExcessRunoff (TOPMODEL fsat, SaturatedExcessRunoffMod.F90:203-360),
// SetQflxInputs (SoilHydrologyMod.F90:298-366), InfiltrationExcessRunoff with
// qinmax from hksat (InfiltrationExcessRunoffMod.F90), RouteInfiltrationExcess
// (SoilHydrologyMod.F90:369-423), UpdateH2osfc with QflxH2osfcSurf and
// QflxH2osfcDrain (SurfaceWaterMod.F90:345-560), Infiltration (:426-457) and
// TotalSurfaceRunoff (:460-550).  Every routine is a per-column update, so the
// chain is one lane per column reading each input once and writing each
// output once (HBM-bound, ~30 FP64 per column); the routines' order inside a
// column is the reference's, so the values are the reference's.
//
// k_water_table: the calls after SoilWater (HydrologyNoDrainageMod.F90:355-373,
// no aquifer layer): PerchedWaterTable (SoilHydrologyMod.F90:1525-1641),
// ThetaBasedWaterTable (:1933-2025) and RenewCondensation (:2569-2678), one
// lane per column, each a search over the column's layers.
//
// k_drainage: the soil-column part of HydrologyDrainage without the aquifer
// layer or hillslopes (HydrologyDrainageMod.F90:137-149, 160-245):
// PerchedLateralFlow (SoilHydrologyMod.F90:1645-1929), SubsurfaceLateralFlow
// (:2029-2565), the volumetric water and the runoff sum; one lane per column,
// the routines' column loops run in order inside the lane.
//
// k_patch2col: column values of three patch fluxes as clm_drv_patch2col forms
// them (p2c_1d_filter, subgridAveMod.F90:312-318), for the land step: the
// evaporation fluxes SetQflxInputs reads and the dew / sublimation fluxes
// RenewCondensation reads.
#include <climits>

#include "common.h"
#include "trace.h"

namespace ctsm {
namespace hydro {

constexpr double RPI = 3.14159265358979323846;

struct HView {
    int begc, ldc, begp, ldp, nlevsno, nlevsoi, nlevgrnd;
    __device__ __forceinline__ long c1(int c) const { return (long)(c - begc); }
    __device__ __forceinline__ long soi(int c, int j) const { return (long)(c - begc) + (long)(j - 1) * ldc; }
    __device__ __forceinline__ long sno(int c, int j) const { return (long)(c - begc) + (long)(j + nlevsno - 1) * ldc; }
    __device__ __forceinline__ long zi(int c, int j) const { return (long)(c - begc) + (long)(j + nlevsno) * ldc; }
    __device__ __forceinline__ long p1(int p) const { return (long)(p - begp); }
};

__global__ void __launch_bounds__(256) k_surface_water(HView V, double e_ice, ctsm_hydro_params P, ctsm_col_state S,
                                                       ctsm_surface_water_state W, int n,
                                                       const int32_t *__restrict__ filter, double dtime,
                                                       ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) { report_error(err, c, CTSM_E_ARG); return; }
    const long cc = V.c1(c);

    // SaturatedExcessRunoff, ComputeFsatTopmodel
    const double zwt = W.zwt[cc], zwtp = W.zwt_perched[cc], ft = W.frost_table[cc];
    double fsat;
    if (ft > zwtp && ft <= zwt) fsat = W.wtfact[cc] * cm_exp(-0.5 * P.fff * zwtp);
    else fsat = W.wtfact[cc] * cm_exp(-0.5 * P.fff * zwt);
    const double rain = W.qflx_rain_plus_snomelt[cc];
    const double sat_excess = fsat * rain;
    W.fsat[cc] = fsat;
    W.qflx_sat_excess_surf[cc] = sat_excess;
    W.fcov[cc] = fsat;

    // SetQflxInputs
    const double top = rain + W.qflx_snow_h2osfc[cc] + W.qflx_floodc[cc];
    double fsno, qevap;
    if (S.snl[cc] >= 0) {
        fsno = 0.0;
        qevap = W.qflx_liqevap_from_top_layer[cc];
    } else {
        fsno = S.frac_sno_eff[cc];
        qevap = W.qflx_ev_soil[cc];
    }
    const double fh = S.frac_h2osfc[cc];
    double in_soil = (1.0 - fh) * (top - sat_excess);
    double to_h2osfc = fh * (top - sat_excess);
    in_soil = in_soil - (1.0 - fsno - fh) * qevap;
    to_h2osfc = to_h2osfc - fh * W.qflx_ev_h2osfc[cc];
    W.qflx_top_soil[cc] = top;
    W.qflx_in_soil[cc] = in_soil;
    W.qflx_top_soil_to_h2osfc[cc] = to_h2osfc;

    // InfiltrationExcessRunoff, ComputeQinmaxHksat (minval over layers 1..3)
    double qmin = 0.0;
#pragma unroll
    for (int j = 1; j <= 3; ++j) {
        const double v = cm_pow(10.0, -e_ice * S.icefrac[V.soi(c, j)]) * S.hksat[V.soi(c, j)];
        if (j == 1 || v < qmin) qmin = v;
    }
    const double qinmax = (1.0 - fsat) * qmin;
    const double infl_excess = fmax90(0.0, (in_soil - (1.0 - fh) * qinmax));
    W.qinmax[cc] = qinmax;
    W.qflx_infl_excess[cc] = infl_excess;

    // RouteInfiltrationExcess (soil / crop)
    const double in_soil_limited = in_soil - infl_excess;
    double in_h2osfc, infl_excess_surf;
    if (P.h2osfcflag != 0) {
        in_h2osfc = to_h2osfc + infl_excess;
        infl_excess_surf = 0.0;
    } else {
        in_h2osfc = to_h2osfc;
        infl_excess_surf = infl_excess;
    }
    W.qflx_in_soil_limited[cc] = in_soil_limited;
    W.qflx_in_h2osfc[cc] = in_h2osfc;
    W.qflx_infl_excess_surf[cc] = infl_excess_surf;

    // UpdateH2osfc: QflxH2osfcSurf
    double frac_infclust = 0.0;
    if (P.h2osfcflag == 1) {
        const double fn = W.frac_h2osfc_nosnow[cc];
        if (fn <= P.pc) frac_infclust = 0.0;
        else frac_infclust = cm_pow(fn - P.pc, P.mu);
    }
    const double h0 = W.h2osfc[cc], th = W.h2osfc_thresh[cc];
    double surf;
    if (h0 > th && P.h2osfcflag != 0) {
        const double k_wet = 1.0e-4 * cm_sin((RPI / 180.0) * W.topo_slope[cc]);
        surf = k_wet * frac_infclust * (h0 - th);
        surf = fmin90(surf, (h0 - th) / dtime);
    } else {
        surf = 0.0;
    }
    if (surf < (double)1.0e-8f) surf = 0.0;      // default-REAL literal in SurfaceWaterMod.F90:497
    W.qflx_h2osfc_surf[cc] = surf;
    double hp = h0 + (in_h2osfc - surf) * dtime;
    if (fabs(hp) < 1.e-13 * fabs(h0)) hp = 0.0;
    // QflxH2osfcDrain
    double drain;
    if (hp < 0.0) {
        drain = hp / dtime;
    } else {
        drain = fmin90(fh * qinmax, hp / dtime);
        if (P.h2osfcflag == 0) drain = fmax90(0.0, hp / dtime);
    }
    W.qflx_h2osfc_drain[cc] = drain;
    double h1 = hp - drain * dtime;
    if (fabs(h1) < 1.e-13 * fabs(hp)) h1 = 0.0;
    W.h2osfc[cc] = h1;

    // Infiltration, TotalSurfaceRunoff
    W.qflx_infl[cc] = in_soil_limited + drain;
    W.qflx_surf[cc] = sat_excess + infl_excess_surf + surf;
}

// sat_lev of both water-table searches: `0.9` without a kind, a single-precision
// value (SoilHydrologyMod.F90:1553, 1983)
constexpr double SAT_LEV = (double)0.9f;
// tolerance of RenewCondensation's truncation (SoilHydrologyMod.F90:71)
constexpr double RENEW_TOLERANCE = 1.e-12;

// volumetric water of soil layer k
__device__ __forceinline__ double layer_vol(const HView &V, const ctsm_col_state &S, int c, int k)
{
    const long i = V.sno(c, k);
    return S.h2osoi_liq[i] / (S.dz[i] * DENH2O) + S.h2osoi_ice[i] / (S.dz[i] * DENICE);
}

// water table from the interpolated depth where s crosses sat_lev between layers k and k + 1
__device__ __forceinline__ double crossing(const HView &V, const ctsm_col_state &S, int c, int k, double s1, double s2)
{
    const double z0 = S.z[V.sno(c, k)], z1 = S.z[V.sno(c, k + 1)];
    const double m = (z1 - z0) / (s2 - s1);
    const double b = z1 - m * s2;
    return fmax90(0.0, m * SAT_LEV + b);
}

__global__ void __launch_bounds__(256) k_water_table(HView V, ctsm_col_state S, ctsm_water_table_state W, int n,
                                                     const int32_t *__restrict__ filter, double dtime, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) { report_error(err, c, CTSM_E_ARG); return; }
    const long cc = V.c1(c);
    const int nlevsoi = V.nlevsoi;

    // PerchedWaterTable: the frost table is the top of the first frozen layer
    // under an unfrozen one
    int k_frz = (S.t_soisno[V.sno(c, 1)] > TFRZ) ? nlevsoi : 1;
    for (int k = 2; k <= nlevsoi; ++k) {
        if (S.t_soisno[V.sno(c, k - 1)] > TFRZ && S.t_soisno[V.sno(c, k)] <= TFRZ) {
            k_frz = k;
            break;
        }
    }
    const double frost = S.zi[V.zi(c, k_frz - 1)];
    double zwt_perched = frost;
    const double t_frz = S.t_soisno[V.sno(c, k_frz)];
    if (W.zwt[cc] < frost && t_frz <= TFRZ) {
        // water table above the frost table: nothing to do
    } else if (k_frz > 1) {
        int k_perch = 1;
        for (int k = k_frz; k >= 1; --k) {
            const double vol = layer_vol(V, S, c, k);
            if (W.h2osoi_vol) W.h2osoi_vol[V.soi(c, k)] = vol;
            if (vol / S.watsat[V.soi(c, k)] <= SAT_LEV) {
                k_perch = k;
                break;
            }
        }
        if (t_frz > TFRZ) k_perch = k_frz;
        if (k_frz > k_perch) {
            const double s1 = layer_vol(V, S, c, k_perch) / S.watsat[V.soi(c, k_perch)];
            const double s2 = layer_vol(V, S, c, k_perch + 1) / S.watsat[V.soi(c, k_perch + 1)];
            zwt_perched = (s1 > s2) ? S.zi[V.zi(c, k_perch - 1)] : crossing(V, S, c, k_perch, s1, s2);
        }
    }
    W.frost_table[cc] = frost;
    W.zwt_perched[cc] = zwt_perched;

    // ThetaBasedWaterTable: first layer from the bedrock up below sat_lev
    const int nb = S.nbedrock[cc];
    int k_zwt = nb;
    bool all_sat = true;
    for (int k = nb; k >= 1; --k) {
        const double vol = layer_vol(V, S, c, k);
        if (W.h2osoi_vol) W.h2osoi_vol[V.soi(c, k)] = vol;
        if (vol / S.watsat[V.soi(c, k)] <= SAT_LEV) {
            k_zwt = k;
            all_sat = false;
            break;
        }
    }
    if (all_sat) k_zwt = 1;
    double zwt;
    if (k_zwt == 1) {
        zwt = S.zi[V.zi(c, 1)];
    } else if (k_zwt < nb) {
        const double s1 = layer_vol(V, S, c, k_zwt) / S.watsat[V.soi(c, k_zwt)];
        const double s2 = layer_vol(V, S, c, k_zwt + 1) / S.watsat[V.soi(c, k_zwt + 1)];
        zwt = crossing(V, S, c, k_zwt, s1, s2);
    } else {
        zwt = S.zi[V.zi(c, nb)];
    }
    W.zwt[cc] = zwt;

    // RenewCondensation: dew, frost and sublimation of a snow-free top layer
    if (S.snl[cc] + 1 >= 1) {
        const long t = V.sno(c, 1);
        const double wf = 1.0 - S.frac_h2osfc[cc];
        S.h2osoi_liq[t] = S.h2osoi_liq[t] + wf * W.qflx_liqdew_to_top_layer[cc] * dtime;
        double ice = S.h2osoi_ice[t] + wf * W.qflx_soliddew_to_top_layer[cc] * dtime;
        const double before = ice;
        ice = ice - wf * W.qflx_solidevap_from_top_layer[cc] * dtime;
        if (fabs(ice) < RENEW_TOLERANCE * fabs(before)) ice = 0.0;
        S.h2osoi_ice[t] = ice;
        if (ice < 0.0) report_error(err, c, CTSM_E_ENDRUN);
    }
}

constexpr double WATMIN = 0.01;    // clm_varcon.F90:81
constexpr double PONDMX = 0.0;     // clm_varcon.F90:116

// specific yield of layer k at water table depth zw (SoilHydrologyMod.F90:1904-1906, 2459-2461)
__device__ __forceinline__ double spec_yield(const HView &V, const ctsm_col_state &S, const ctsm_hydro_params &P,
                                             int c, int k, double zw)
{
    const long s = V.soi(c, k);
    const double sy = S.watsat[s] * (1. - cm_pow(1. + 1.e3 * zw / S.sucsat[s], -(1. / S.bsw[s])));
    return fmax90(sy, P.aq_sp_yield_min);
}

__global__ void __launch_bounds__(256) k_drainage(HView V, double e_ice, ctsm_hydro_params P, ctsm_col_state S,
                                                  ctsm_drainage_state D, int n, const int32_t *__restrict__ filter,
                                                  double dtime, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) { report_error(err, c, CTSM_E_ARG); return; }
    const long cc = V.c1(c);
    const int nlevsoi = V.nlevsoi;
    const int nb = S.nbedrock[cc];
    double *liq = S.h2osoi_liq, *ice = S.h2osoi_ice;
    auto L = [&](int k) -> double & { return liq[V.sno(c, k)]; };
    auto Z = [&](int k) { return S.zi[V.zi(c, k)]; };
    auto DZ = [&](int k) { return S.dz[V.sno(c, k)]; };

    // ---- PerchedLateralFlow (no hillslope)
    const double frost = D.frost_table[cc], zwtp = D.zwt_perched[cc];
    int k_frost = nb, k_perch = nb;
    for (int k = 1; k <= nb; ++k)
        if (frost >= Z(k - 1) && frost < Z(k)) { k_frost = k; break; }
    for (int k = 1; k <= nb; ++k)
        if (zwtp >= Z(k - 1) && zwtp < Z(k)) { k_perch = k; break; }
    double perched_out = 0.0;
    if (frost > zwtp) {
        const double q_perch_max = P.perched_baseflow_scalar * cm_sin(D.topo_slope[cc] * (RPI / 180.0));
        double wtsub = 0.0, q_perch = 0.0;
        for (int k = k_perch; k <= k_frost - 1; ++k) {
            q_perch = q_perch + S.hksat[V.soi(c, k)] * DZ(k);
            wtsub = wtsub + DZ(k);
        }
        if (wtsub > 0.0) q_perch = q_perch / wtsub;
        perched_out = q_perch_max * q_perch * (frost - zwtp);
    }
    double drain_perched = 0.0 + perched_out;
    double tot = drain_perched * dtime;
    for (int k = k_perch; k <= k_frost - 1; ++k) {
        const double s_y = spec_yield(V, S, P, c, k, zwtp);
        double dl = (k == k_perch) ? fmin90(tot, (s_y * (Z(k) - zwtp) * 1.e3)) : fmin90(tot, (s_y * DZ(k) * 1.e3));
        dl = fmax90(dl, 0.0);
        tot = tot - dl;
        L(k) = L(k) - dl;
    }
    drain_perched = drain_perched - tot / dtime;

    // ---- SubsurfaceLateralFlow (no hillslope, lateral flow on)
    double dzsum = 0.0, icefracsum = 0.0;
    double zwt = D.zwt[cc];
    int jwt = nlevsoi;
    for (int j = 1; j <= nlevsoi; ++j) {
        if (zwt <= Z(j)) { jwt = j - 1; break; }
    }
    for (int j = 1; j <= nlevsoi; ++j) {
        const long s = V.soi(c, j);
        const double vol_ice = fmin90(S.watsat[s], ice[V.sno(c, j)] / (DZ(j) * DENICE));
        const double f = fmin90(1.0, vol_ice / S.watsat[s]);
        S.icefrac[s] = f;
        if (j >= (jwt > 1 ? jwt : 1)) {
            const double dzmm = DZ(j) * 1.e3;
            dzsum = dzsum + dzmm;
            icefracsum = icefracsum + f * dzmm;
        }
    }
    const double ice_imped_col = cm_pow(10.0, -e_ice * (icefracsum / dzsum));
    const double zinb = Z(nb);
    double latflow_out = 0.0;
    if (zwt <= zinb)
        latflow_out = ice_imped_col * P.baseflow_scalar * cm_tan(RPI / 180.0 * D.topo_slope[cc]) *
                      cm_pow(zinb - zwt, P.n_baseflow);
    const double latflow_in = 0.0;
    const double net = latflow_out - latflow_in;
    double drainage = (zwt <= zinb) ? net : 0.0;
    tot = -drainage * dtime;
    double h2osfc = S.h2osfc[cc];
    if (tot > 0.) {          // rising water table
        for (int j = jwt + 1; j >= 1; --j) {
            if (Z(j) < frost) {
                const double s_y = spec_yield(V, S, P, c, j, zwt);
                double dl = fmin90(tot, (s_y * DZ(j) * 1.e3));
                dl = fmax90(dl, 0.0);
                L(j) = L(j) + dl;
                tot = tot - dl;
                if (tot <= 0.) {
                    zwt = zwt - dl / s_y / 1000.0;
                    break;
                } else {
                    zwt = Z(j - 1);
                }
            }
        }
        h2osfc = h2osfc + tot;
    } else {                 // deepening water table
        for (int j = jwt + 1; j <= nb; ++j) {
            const double s_y = spec_yield(V, S, P, c, j, zwt);
            double dl = fmax90(tot, -(s_y * (Z(j) - zwt) * 1.e3));
            dl = fmin90(dl, 0.0);
            L(j) = L(j) + dl;
            tot = tot - dl;
            if (tot >= 0.) {
                zwt = zwt - dl / s_y / 1000.0;
                break;
            } else {
                zwt = Z(j);
            }
        }
        drainage = drainage + tot / dtime;
    }
    zwt = fmax90(0.0, zwt);
    zwt = fmin90(80.0, zwt);

    // saturation excess pushed up, layer by layer
    for (int j = nlevsoi; j >= 2; --j) {
        const double cap = S.eff_porosity[V.soi(c, j)] * (DZ(j) * 1.e3);
        const double xsi = fmax90(L(j) - cap, 0.0);
        L(j) = fmin90(cap, L(j));
        L(j - 1) = L(j - 1) + xsi;
    }
    // the top layer's excess liquid to the surface water, excess ice to runoff
    const double room = PONDMX + S.watsat[V.soi(c, 1)] * (DZ(1) * 1.e3);
    double &ice1 = ice[V.sno(c, 1)];
    double xs1 = fmax90(fmax90(L(1) - WATMIN, 0.0) - fmax90(0.0, (room - ice1 - WATMIN)), 0.0);
    L(1) = L(1) - xs1;
    h2osfc = h2osfc + xs1;
    double rsub_sat = 0.0;
    xs1 = fmax90(fmax90(ice1, 0.0) - fmax90(0.0, (room - L(1))), 0.0);
    ice1 = fmin90(fmax90(0.0, room - L(1)), ice1);
    const double ice_runoff_xs = xs1 / dtime;
    // no layer below the minimum water: the deficit comes from the layer below
    for (int j = 1; j <= nlevsoi - 1; ++j) {
        double xs;
        if (L(j) < WATMIN) {
            xs = WATMIN - L(j);
            if (j == jwt) zwt = zwt + xs / S.eff_porosity[V.soi(c, j)] / 1000.0;
        } else {
            xs = 0.0;
        }
        L(j) = L(j) + xs;
        L(j + 1) = L(j + 1) - xs;
    }
    // ... and the bottom layer's from the layers above, the rest from outside
    {
        const int j = nlevsoi;
        double xs = 0.0;
        if (L(j) < WATMIN) {
            xs = WATMIN - L(j);
            for (int k = nlevsoi - 1; k >= 1; --k) {
                const double avail = fmax90(L(k) - WATMIN - xs, 0.0);
                if (avail >= xs) {
                    L(j) = L(j) + xs;
                    L(k) = L(k) - xs;
                    xs = 0.0;
                    break;
                } else {
                    L(j) = L(j) + avail;
                    L(k) = L(k) - avail;
                    xs = xs - avail;
                }
            }
        }
        L(j) = L(j) + xs;
        rsub_sat = rsub_sat - xs / dtime;
    }
    const double qdrain = rsub_sat + drainage;
    const double qrgwl = D.qflx_snwcp_liq[cc];

    // HydrologyDrainage: volumetric water, runoff
    if (D.h2osoi_vol)
        for (int j = 1; j <= V.nlevgrnd; ++j)
            D.h2osoi_vol[V.soi(c, j)] = L(j) / (DZ(j) * DENH2O) + ice[V.sno(c, j)] / (DZ(j) * DENICE);
    S.h2osfc[cc] = h2osfc;
    D.zwt[cc] = zwt;
    D.qflx_drain_perched[cc] = drain_perched;
    D.qflx_latflow_out[cc] = latflow_out;
    D.qflx_latflow_in[cc] = latflow_in;
    D.qflx_rsub_sat[cc] = rsub_sat;
    D.qflx_ice_runoff_xs[cc] = ice_runoff_xs;
    D.qflx_drain[cc] = qdrain;
    D.qflx_qrgwl[cc] = qrgwl;
    D.qflx_runoff[cc] = qdrain + D.qflx_surf[cc] + qrgwl + drain_perched;
}

// p2c_1d_filter of three patch fluxes (clm_drv_patch2col, subgridAveMod.F90:312-318)
struct P2C {
    const int32_t *active;
    const double *wtcol, *pa, *pb, *pc;
    double *ca, *cb, *cc;
};

__device__ __forceinline__ void patch2col_column(const HView &V, const ctsm_col_state &S, const P2C &F, int c,
                                                 ErrWord *err)
{
    const long cc = V.c1(c);
    const int pi = S.patchi[cc], pf = S.patchf[cc];
    if (pf >= pi && (pi < V.begp || pf >= V.begp + V.ldp)) { report_error(err, c, CTSM_E_ARG); return; }
    double a = 0.0, b = 0.0, d = 0.0;
    for (int p = pi; p <= pf; ++p) {
        const long pp = V.p1(p);
        if (!F.active[pp]) continue;
        const double w = F.wtcol[pp];
        a = a + F.pa[pp] * w;
        b = b + F.pb[pp] * w;
        d = d + F.pc[pp] * w;
    }
    F.ca[cc] = a;
    F.cb[cc] = b;
    F.cc[cc] = d;
}

__global__ void __launch_bounds__(256) k_patch2col(HView V, ctsm_col_state S, P2C F, int n,
                                                   const int32_t *__restrict__ filter, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) { report_error(err, c, CTSM_E_ARG); return; }
    patch2col_column(V, S, F, c, err);
}

// The same for consecutive columns (no filter): a block of P2C_COLS columns
// stages its consecutive patch range in LDS with coalesced reads, then each
// lane sums its own patches from LDS in ascending order (the operations of
// patch2col_column, in its order).  A block whose patches exceed the staging
// space takes patch2col_column.
constexpr int P2C_COLS = 64;
constexpr int P2C_CAP = 16 * P2C_COLS;

__global__ void __launch_bounds__(P2C_COLS) k_patch2col_staged(HView V, ctsm_col_state S, P2C F, int n, ErrWord *err)
{
    __shared__ double s_w[P2C_CAP], s_a[P2C_CAP], s_b[P2C_CAP], s_c[P2C_CAP];
    __shared__ int32_t s_act[P2C_CAP];
    __shared__ int s_lo, s_hi, s_bad;
    const int t = threadIdx.x;
    const int i = blockIdx.x * P2C_COLS + t;
    const int c = V.begc + i;
    const long cc = V.c1(c);
    const bool live = i < n;
    int pi = 0, pf = -1;
    if (live) {
        pi = S.patchi[cc];
        pf = S.patchf[cc];
    }
    if (t == 0) { s_lo = INT_MAX; s_hi = INT_MIN; s_bad = 0; }
    __syncthreads();
    if (live && pf >= pi) {
        if (pi < V.begp || pf >= V.begp + V.ldp) atomicOr(&s_bad, 1);
        atomicMin(&s_lo, pi);
        atomicMax(&s_hi, pf);
    }
    __syncthreads();
    const int lo = s_lo, hi = s_hi;
    const int np = hi >= lo ? hi - lo + 1 : 0;
    if (s_bad || np > P2C_CAP) {             // block-uniform
        if (live) patch2col_column(V, S, F, c, err);
        return;
    }
    for (int k = t; k < np; k += P2C_COLS) {
        const long q = V.p1(lo + k);
        s_act[k] = F.active[q];
        s_w[k] = F.wtcol[q];
        s_a[k] = F.pa[q];
        s_b[k] = F.pb[q];
        s_c[k] = F.pc[q];
    }
    __syncthreads();
    if (!live) return;
    double a = 0.0, b = 0.0, d = 0.0;
    for (int p = pi; p <= pf; ++p) {
        const int k = p - lo;
        if (!s_act[k]) continue;
        const double w = s_w[k];
        a = a + s_a[k] * w;
        b = b + s_b[k] * w;
        d = d + s_c[k] * w;
    }
    F.ca[cc] = a;
    F.cb[cc] = b;
    F.cc[cc] = d;
}

}  // namespace hydro

void launch_surface_water(hipStream_t st, const ctsm_layout &L, double e_ice, const ctsm_hydro_params &P,
                          const ctsm_col_state &S, const ctsm_surface_water_state &W, int n, const int32_t *filter,
                          double dtime, ErrWord *err)
{
    if (n <= 0) return;
    const GptlRange range("hydro_without_drainage");
    const hydro::HView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, 0, 0, 0};
    hipLaunchKernelGGL(hydro::k_surface_water, dim3((n + 255) / 256), dim3(256), 0, st, V, e_ice, P, S, W, n, filter,
                       dtime, err);
}

void launch_patch2col3(hipStream_t st, const ctsm_layout &L, const ctsm_col_state &S, const int32_t *active,
                       const double *wtcol, const double *pa, const double *pb, const double *pc, double *ca,
                       double *cb, double *ccol, int n, const int32_t *filter, ErrWord *err)
{
    if (n <= 0) return;
    const GptlRange range("hydro_without_drainage");
    const hydro::HView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, 0, 0, 0};
    const hydro::P2C F{active, wtcol, pa, pb, pc, ca, cb, ccol};
    if (filter)
        hipLaunchKernelGGL(hydro::k_patch2col, dim3((n + 255) / 256), dim3(256), 0, st, V, S, F, n, filter, err);
    else
        hipLaunchKernelGGL(hydro::k_patch2col_staged, dim3((n + hydro::P2C_COLS - 1) / hydro::P2C_COLS),
                           dim3(hydro::P2C_COLS), 0, st, V, S, F, n, err);
}

void launch_water_table(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_col_state &S,
                        const ctsm_water_table_state &W, int n, const int32_t *filter, double dtime, ErrWord *err)
{
    if (n <= 0) return;
    const GptlRange range("hydro_without_drainage");
    const hydro::HView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, cfg.nlevsno, cfg.nlevsoi, cfg.nlevgrnd};
    hipLaunchKernelGGL(hydro::k_water_table, dim3((n + 255) / 256), dim3(256), 0, st, V, S, W, n, filter, dtime, err);
}

void launch_drainage(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_hydro_params &P,
                     const ctsm_col_state &S, const ctsm_drainage_state &D, int n, const int32_t *filter,
                     double dtime, ErrWord *err)
{
    if (n <= 0) return;
    const GptlRange range("hydro2_drainage");
    const hydro::HView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, cfg.nlevsno, cfg.nlevsoi, cfg.nlevgrnd};
    hipLaunchKernelGGL(hydro::k_drainage, dim3((n + 255) / 256), dim3(256), 0, st, V, cfg.e_ice, P, S, D, n, filter,
                       dtime, err);
}

}  // namespace ctsm
