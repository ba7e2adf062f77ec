// soiltemp_dev.h -- per-level thermal properties shared by the fused and the
// plain soil-temperature kernels (SoilThermProp, SoilTemperatureMod.F90:602-901).
#pragma once
#include "common.h"

namespace ctsm {

// Thermal conductivity of level j (SoilThermProp :674-760), non-urban soil.
__device__ __forceinline__ double level_thk(const View &V, const Params &P, const ctsm_col_state &S,
                                            int c, int j, int nbedrock, double frac_sno)
{
    const double liq = S.h2osoi_liq[V.sno(c, j)], ice = S.h2osoi_ice[V.sno(c, j)];
    const double dz = S.dz[V.sno(c, j)];
    if (j >= 1) {
        const double ws = S.watsat[V.soi(c, j)];
        double th;
        double satw = (liq / DENH2O + ice / DENICE + 0.0 / DENICE) / (dz * ws);
        satw = fmin90(1.0, satw);
        if (satw > .1e-6) {
            double dke;
            if (S.t_soisno[V.sno(c, j)] >= TFRZ) dke = fmax90(0.0, cm_log10(satw) + 1.0);
            else dke = satw;
            const double fl = (liq / (DENH2O * dz)) /
                              (liq / (DENH2O * dz) + ice / (DENICE * dz) + 0.0 / (DENICE * dz));
            const double dksat = S.tkmg[V.soi(c, j)] * cm_pow(TKWAT, fl * ws) * cm_pow(TKICE, (1.0 - fl) * ws);
            th = dke * dksat + (1.0 - dke) * S.tkdry[V.soi(c, j)];
        } else {
            th = S.tkdry[V.soi(c, j)];
        }
        if (j > nbedrock) th = THK_BEDROCK;
        return th;
    }
    // snow layer
    const double bw = (ice + liq) / (frac_sno * dz);
    if (S.bw) S.bw[V.sno(c, j)] = bw;
    if (P.snow_thermal_cond == CTSM_SNOWTHK_JORDAN1991)
        return TKAIR + (7.75e-5 * bw + 1.105e-6 * bw * bw) * (TKICE - TKAIR);
    // Sturm1997: the reference writes these constants as default (4-byte) REALs.
    if (bw <= 156.0) return (double)0.023f + (double)0.234f * (bw / 1000.0);
    return (double)0.138f - (double)1.01f * (bw / 1000.0) + ((double)3.233f * ((bw / 1000.0) * (bw / 1000.0)));
}

// Heat capacity of level j (:826-880).
__device__ __forceinline__ double level_cv(const View &V, const ctsm_col_state &S, int c, int j,
                                           int nbedrock, double frac_sno, double h2osno_no_layers)
{
    const double liq = S.h2osoi_liq[V.sno(c, j)], ice = S.h2osoi_ice[V.sno(c, j)];
    const double dz = S.dz[V.sno(c, j)];
    if (j >= 1) {
        double cv = S.csol[V.soi(c, j)] * (1.0 - S.watsat[V.soi(c, j)]) * dz + (ice * CPICE + liq * CPLIQ) + 0.0 * CPICE;
        if (j > nbedrock) cv = CSOL_BEDROCK * dz;
        if (j == 1 && h2osno_no_layers > 0.0) cv = cv + CPICE * h2osno_no_layers;
        return cv;
    }
    if (frac_sno > 0.0) return fmax90(THIN_SFCLAYER, (CPLIQ * liq + CPICE * ice) / frac_sno);
    return THIN_SFCLAYER;
}

}  // namespace ctsm
