// watertable_bc.h -- the water-table bottom boundary of the moisture-form
// Richards solve (lower_boundary_condition = bc_waterTable):
//   - the bottom flux (SoilWaterMovementMod.F90:1767-1808, zdflag = 0): none
//     when the water table lies inside the soil column, else the flux to a
//     saturated water table at depth zwt below the lowest node (Niu et al.
//     2007), with hk at the lowest node;
//   - the drainage of a sub-step, qout + dqodw1 * dwat of the bottom layer
//     (:1376-1377);
//   - compute_qcharge (:2016-2135): when the water table is inside the soil
//     column, the recharge from the layer above it, replacing the sub-steps'
//     drainage, limited to +-10 mm per step.
// The reference requires use_bedrock = .false. for this boundary
// (SoilWaterMovementMod.F90:181-183), so every column solves nlevsoi layers.
#pragma once
#include "common.h"

namespace ctsm {

// A column whose nbedrock is not nlevsoi under the water-table boundary, or
// a state without zwt / smpmin.
constexpr int RICHARDS_BAD_COLUMN = 5;

__device__ __forceinline__ int richards_error_code(int rc)
{
    return rc == RICHARDS_BAD_COLUMN ? CTSM_E_ARG : CTSM_E_SOLVER;
}

// jwt, the layer above the water table: the first j in 1..nlevsoi with
// zwt <= zi(j) gives j - 1; nlevsoi when the table is below the column.
__device__ __forceinline__ int watertable_layer(const View &V, const ctsm_col_state &S, int c)
{
    const double zwt = S.zwt[V.c1(c)];
    for (int j = 1; j <= V.nlevsoi; ++j)
        if (zwt <= S.zi[V.zi(c, j)]) return j - 1;
    return V.nlevsoi;
}

// Flux and its derivative at the bottom of layer nlevsoi when the water table
// is below the column: smp = 0 assumed at depth zwt.
__device__ __forceinline__ void watertable_bottom_flux(double hk, double dhkdw, double smp, double dsmpdw,
                                                       double watsat, double zwt, double z, double &qout,
                                                       double &dqodw1)
{
    const double dhkds1 = dhkdw / watsat;
    const double num = -smp;
    const double den = M_TO_MM * (zwt - z);
    qout = -(hk * num / den) + hk;
    dqodw1 = (hk * dsmpdw - dhkds1 * num) / den + dhkds1;
}

// soil_hk of compute_qcharge for layer jwt+1 at the last trial's liquid:
// s1 = min(1, max(vwc_liq / watsat, 0.01)), imped * hksat * s1**(2b+3).
__device__ __forceinline__ double watertable_ka(double vwc, double watsat, double imped_hksat, double bsw)
{
    double s1 = fmax90(vwc / watsat, 0.01);
    s1 = fmin90(1.0, s1);
    return imped_hksat * cm_pow(s1, 2.0 * bsw + 3.0);
}

// qcharge of compute_qcharge for jwt < nlevsoi; smp_w is the last trial's
// smp of layer max(1, jwt).  The 1.e-3 of the jwt = 0 case is a default-real
// literal in the reference, so it enters as the binary32 value.
__device__ __forceinline__ double watertable_qcharge(const View &V, const ctsm_col_state &S, int c, int jwt,
                                                     double ka, double smp_w, double dtime)
{
    const double zwt = S.zwt[V.c1(c)];
    const double wh_zwt = -S.sucsat[V.soi(c, jwt + 1)] - zwt * M_TO_MM;
    const double smp1 = fmax90(S.smpmin[V.c1(c)], smp_w);
    const int jw = jwt > 1 ? jwt : 1;
    const double wh = smp1 - S.z[V.sno(c, jw)] * M_TO_MM;
    double q;
    if (jwt == 0) q = -(ka * (wh_zwt - wh) / ((zwt + (double)1.e-3f) * M_TO_MM));
    else q = -(ka * (wh_zwt - wh) / ((zwt - S.z[V.sno(c, jwt)]) * M_TO_MM * 2.0));
    q = fmax90(-10.0 / dtime, q);
    q = fmin90(10.0 / dtime, q);
    return q;
}

}  // namespace ctsm
