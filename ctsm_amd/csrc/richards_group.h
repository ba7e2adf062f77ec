// richards_group.h -- moisture-form Richards solve with G lanes per column.
//
// The same arithmetic as soilwater_moisture_form (SoilWaterMovementMod.F90:
// 976-1440) and as the one-lane form in soilwater_dev.h, laid out for gfx950:
// a column's layers are split over G consecutive lanes (lane g of the group
// owns layers g*K .. g*K+K-1, K = 20/G), so
//   - the Clapp-Hornberger properties (two x**y per layer) and the flux rows
//     of a trial are computed by G lanes at once;
//   - the LAPACK dgtsv elimination, which is sequential in the layer index,
//     runs as G stages: lane s eliminates its rows, then the partly
//     eliminated row is broadcast to lane s+1 (a cross-lane shuffle);
//   - the back substitution runs the stages in reverse, carrying the two
//     solution values below;
//   - the error estimate max is reduced over the group, so every lane of the
//     group takes the same accept / halve / double decision.
// A lane keeps ~K*13 doubles instead of 20*13, so nothing spills and several
// waves share a SIMD.  Values are bit-identical to the one-lane form.
// G = 0 takes the group size at run time (argument gsz: 2, 4 or 8 lanes of K
// layers each), so one kernel can give each wavefront the group size its
// columns' depth needs.
#pragma once
#include "common.h"
#include "watertable_bc.h"

namespace ctsm {

__device__ __forceinline__ double gshfl(double v, int src, int width)
{
    return __shfl(v, src, width);
}

// WT: the water-table bottom boundary (watertable_bc.h), a separate
// instantiation so the flux / zero-flux kernels keep their registers.
template <int G, bool WT = false, int K = (MAX_NLEVSOI + (G > 0 ? G : 1) - 1) / (G > 0 ? G : 1)>
__device__ __forceinline__ int richards_group(const View &V, const Params &P, const ctsm_col_state &S, int c,
                                              double dtime, int g, int gsz = G)
{
    const int GS = (G > 0) ? G : gsz;       // lanes in the group
    const int nl = S.nbedrock[V.c1(c)];
    if (nl < 2 || nl > MAX_NLEVSOI || nl > GS * K) return 1;

    double mdz[K], ws[K], bsw[K], hkc[K], suc[K], den[K], sink[K], liq[K];
    double Ud[K], Udu[K], Udl[K], Ub[K], f1[K];
    const double infl = S.qflx_infl[V.c1(c)];
    // watsat of the first layer of the next lane (for the interface below our last layer)
    double ws_nl = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = g * K + k;            // 0-based layer
        mdz[k] = ws[k] = bsw[k] = hkc[k] = suc[k] = den[k] = sink[k] = liq[k] = 0.0;
        if (j < nl) {
            const int jj = j + 1;
            mdz[k] = M_TO_MM * S.dz[V.sno(c, jj)];
            ws[k] = S.watsat[V.soi(c, jj)];
            bsw[k] = S.bsw[V.soi(c, jj)];
            suc[k] = S.sucsat[V.soi(c, jj)];
            sink[k] = S.qflx_rootsoi[V.soi(c, jj)];
            liq[k] = S.h2osoi_liq[V.sno(c, jj)];
            const double ic = S.icefrac[V.soi(c, jj)];
            double icei;
            if (j == nl - 1) {
                icei = ic;
            } else {
                icei = 0.5 * (ic + S.icefrac[V.soi(c, jj + 1)]);
                den[k] = M_TO_MM * (S.z[V.sno(c, jj + 1)] - S.z[V.sno(c, jj)]);
            }
            hkc[k] = cm_pow(10.0, -(P.e_ice * icei)) * S.hksat[V.soi(c, jj)];
        }
    }
    {
        const int jn = (g + 1) * K;
        if (jn < nl) ws_nl = S.watsat[V.soi(c, jn + 1)];
    }

    const bool zero_flux_bottom = (P.lower_bc == CTSM_BC_ZERO_FLUX);
    constexpr bool wt_bottom = WT;
    int jwt = 0;
    double zwt = 0.0, z_bot = 0.0, q_ka = 0.0, q_smp = 0.0;
    if (wt_bottom) {
        if (nl != V.nlevsoi || !S.zwt || !S.smpmin) return RICHARDS_BAD_COLUMN;
        jwt = watertable_layer(V, S, c);
        zwt = S.zwt[V.c1(c)];
        z_bot = S.z[V.sno(c, nl)];
    }
    const int bot_lane = (nl - 1) / K;
    int nsub = 0;
    double dtsub = dtime, dtdone = 0.0, qcharge = 0.0;
    int status = 0;

    for (;;) {
        nsub = nsub + 1;
        if (nsub > MAX_TRIALS) { status = 3; break; }
        const bool candidate = fabs(dtime - (dtdone + dtsub)) < P.verySmall;

        // ---- node properties of each lane's first layer; the previous lane
        //      needs them for the interface below its last layer
        double s2_0 = 0.0, smp_0 = 0.0, dsm_0 = 0.0;
        if (g * K < nl) {
            const double vwc = fmax90(liq[0], 1.0e-6) / mdz[0];
            s2_0 = fmin90(vwc / ws[0], 1.0);
            s2_0 = fmax90(0.01, s2_0);
            smp_0 = -(suc[0] * cm_pow(s2_0, -bsw[0]));
            dsm_0 = (-(bsw[0] * smp_0) / s2_0) / ws[0];
        }
        const int src_nx = (g + 1 < GS) ? g + 1 : g;
        const double s2_nx = gshfl(s2_0, src_nx, GS);
        const double smp_nx = gshfl(smp_0, src_nx, GS);
        const double dsmpdw_nx = gshfl(dsm_0, src_nx, GS);

        // ---- forward: properties, fluxes and rows of the owned layers
        double qout_last = 0.0, dqodw1_last = 0.0, dqodw2_last = 0.0;
        double hk_bot = 0.0, dhkdw_bot = 0.0;
        {
            double s2 = s2_0, smp = smp_0, dsmpdw = dsm_0;   // node values of layer k
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int j = g * K + k;
                if (j < nl) {
                    const bool bottom = (j == nl - 1);
                    double s2n, smpn, dsmpdwn, wsn;
                    if (k + 1 < K) {
                        s2n = 0.0; smpn = 0.0; dsmpdwn = 0.0; wsn = 0.0;
                        if (j + 1 < nl) {
                            const double vwcn = fmax90(liq[k + 1], 1.0e-6) / mdz[k + 1];
                            s2n = fmin90(vwcn / ws[k + 1], 1.0);
                            s2n = fmax90(0.01, s2n);
                            smpn = -(suc[k + 1] * cm_pow(s2n, -bsw[k + 1]));
                            dsmpdwn = (-(bsw[k + 1] * smpn) / s2n) / ws[k + 1];
                            wsn = ws[k + 1];
                        }
                    } else {
                        s2n = s2_nx; smpn = smp_nx; dsmpdwn = dsmpdw_nx; wsn = ws_nl;
                    }
                    double s1 = bottom ? s2 : 0.5 * (s2 + s2n);
                    s1 = fmin90(s1, 1.0);
                    s1 = fmax90(0.01, s1);
                    const double e = 2.0 * bsw[k] + 3.0;
                    const double hk = hkc[k] * cm_pow(s1, e);
                    const double dhkdw = e * hk / s1;
                    double qout, dqodw1, dqodw2;
                    if (!bottom) {
                        const double dhkds1 = 0.5 * dhkdw / ws[k];
                        const double dhkds2 = 0.5 * dhkdw / wsn;
                        const double num = smpn - smp;
                        qout = -(hk * num / den[k]) + hk;
                        dqodw1 = (hk * dsmpdw - dhkds1 * num) / den[k] + dhkds1;
                        dqodw2 = (-(hk * dsmpdwn) - dhkds2 * num) / den[k] + dhkds2;
                    } else {
                        hk_bot = hk;
                        dhkdw_bot = dhkdw;
                        if (zero_flux_bottom) { qout = 0.0; dqodw1 = 0.0; }
                        else if (wt_bottom) {
                            if (jwt < V.nlevsoi) { qout = 0.0; dqodw1 = 0.0; }
                            else watertable_bottom_flux(hk, dhkdw, smp, dsmpdw, ws[k], zwt, z_bot, qout, dqodw1);
                            // the sub-step's drainage is qout + dqodw1 * dwat
                            hk_bot = qout;
                            dhkdw_bot = dqodw1;
                        }
                        else { qout = hk; dqodw1 = dhkdw / ws[k]; }
                        dqodw2 = 0.0;
                    }
                    if (candidate && wt_bottom) {
                        // what compute_qcharge reads of the last trial
                        if (j == jwt) q_ka = watertable_ka(fmax90(liq[k], 1.0e-6) / mdz[k], ws[k], hkc[k], bsw[k]);
                        if (j == (jwt > 1 ? jwt : 1) - 1) q_smp = smp;
                    }
                    if (candidate) {
                        S.smp_l[V.soi(c, j + 1)] = smp;
                        S.hk_l[V.soi(c, j + 1)] = hk;
                        S.qout[V.soi(c, j + 1)] = qout;
                    }
                    // keep this layer's outgoing flux terms; the row needs the incoming
                    // ones, which belong to the layer above (filled in below)
                    Ud[k] = qout; Udu[k] = dqodw1; Udl[k] = dqodw2;
                    Ub[k] = dtsub / mdz[k];            // dt_dz
                    s2 = s2n; smp = smpn; dsmpdw = dsmpdwn;
                }
            }
            qout_last = Ud[K - 1]; dqodw1_last = Udu[K - 1]; dqodw2_last = Udl[K - 1];
        }
        // incoming flux of our first layer = outgoing flux of the previous lane's last layer
        const int prv = (g > 0) ? g - 1 : 0;
        const double qin_p = gshfl(qout_last, prv, GS);
        const double dqidw0_p = gshfl(dqodw1_last, prv, GS);
        const double dqidw1_p = gshfl(dqodw2_last, prv, GS);
        // rows (a in Udl, b in Ud, c in Udu, r in Ub) and start-of-step net flux f1
        {
            double qin = (g == 0) ? infl : qin_p;
            double dqidw0 = (g == 0) ? 0.0 : dqidw0_p;
            double dqidw1 = (g == 0) ? 0.0 : dqidw1_p;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int j = g * K + k;
                if (j < nl) {
                    const double qout = Ud[k], dqodw1 = Udu[k], dqodw2 = Udl[k], dtdz = Ub[k];
                    if (candidate) S.qin[V.soi(c, j + 1)] = qin;
                    const bool bottom = (j == nl - 1);
                    const double fluxNet = qin - qout - sink[k];
                    f1[k] = fluxNet;
                    Ub[k] = -(fluxNet * dtdz);                              // r
                    Udl[k] = dqidw0 * dtdz;                                 // a (unused for j = 0)
                    Ud[k] = -1.0 - (-dqidw1 + dqodw1) * dtdz;               // b
                    Udu[k] = bottom ? 0.0 : -(dqodw2 * dtdz);               // c
                    qin = qout; dqidw0 = dqodw1; dqidw1 = dqodw2;
                }
            }
        }
        // first row of the next lane (a, b, c, r) for our last elimination step
        const int nxt = (g + 1 < GS) ? g + 1 : g;
        const double a_nx = gshfl(Udl[0], nxt, GS), b_nx = gshfl(Ud[0], nxt, GS);
        const double c_nx = gshfl(Udu[0], nxt, GS), r_nx = gshfl(Ub[0], nxt, GS);

        // ---- dgtsv elimination in G stages
        double cd = 0.0, cdu = 0.0, cb = 0.0;     // current (partly eliminated) row
        bool singular = false;
#pragma unroll
        for (int s = 0; s < GS; ++s) {
            if (g == s && g * K < nl) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int j = g * K + k;
                    if (j < nl) {
                        if (j == 0) { cd = Ud[k]; cdu = Udu[k]; cb = Ub[k]; }
                        if (j == nl - 1) {
                            Ud[k] = cd; Udu[k] = 0.0; Udl[k] = 0.0; Ub[k] = cb;
                            if (cd == 0.0) singular = true;
                        } else {
                            // next row j+1
                            double dl, nd, ndu, nb;
                            if (k + 1 < K) { dl = Udl[k + 1]; nd = Ud[k + 1]; ndu = Udu[k + 1]; nb = Ub[k + 1]; }
                            else { dl = a_nx; nd = b_nx; ndu = c_nx; nb = r_nx; }
                            const bool last = (j + 1 == nl - 1);
                            if (fabs(cd) >= fabs(dl)) {
                                if (cd != 0.0) {
                                    const double fact = dl / cd;
                                    nd = nd - fact * cdu;
                                    nb = nb - fact * cb;
                                } else {
                                    singular = true;
                                }
                                if (!last) dl = 0.0;
                            } else {
                                const double fact = cd / dl;
                                const double d_i = dl;
                                const double temp = nd;
                                nd = cdu - fact * temp;
                                if (!last) { dl = ndu; ndu = -fact * dl; }
                                const double du_i = temp;
                                const double b_i = nb;
                                nb = cb - fact * nb;
                                cd = d_i; cdu = du_i; cb = b_i;
                            }
                            Ud[k] = cd; Udu[k] = cdu; Udl[k] = dl; Ub[k] = cb;
                            cd = nd; cdu = ndu; cb = nb;
                        }
                    }
                }
            }
            cd = gshfl(cd, s, GS);
            cdu = gshfl(cdu, s, GS);
            cb = gshfl(cb, s, GS);
        }
        // any lane singular -> the whole column fails
        {
            double sg = singular ? 1.0 : 0.0;
#pragma unroll
            for (int m = 1; m < GS; m <<= 1) sg = fmax90(sg, __shfl_xor(sg, m, GS));
            if (sg != 0.0) { status = 2; break; }
        }

        // ---- back substitution in reverse stages + error estimate
        double x1 = 0.0, x2 = 0.0, emax = 0.0, dwat_bot = 0.0;
        bool have_e = false;
#pragma unroll
        for (int s = GS - 1; s >= 0; --s) {
            if (g == s && g * K < nl) {
#pragma unroll
                for (int k = K - 1; k >= 0; --k) {
                    const int j = g * K + k;
                    if (j < nl) {
                        double x;
                        if (j == nl - 1)      x = Ub[k] / Ud[k];
                        else if (j == nl - 2) x = (Ub[k] - Udu[k] * x1) / Ud[k];
                        else                  x = (Ub[k] - Udu[k] * x1 - Udl[k] * x2) / Ud[k];
                        if (j == nl - 1) dwat_bot = x;
                        const double dtdz = dtsub / mdz[k];
                        const double ev = fabs(f1[k] - x / dtdz) * dtsub * 0.5;
                        if (!have_e || ev > emax) { emax = ev; have_e = true; }
                        Ub[k] = x;
                        x2 = x1;
                        x1 = x;
                    }
                }
            }
            x1 = gshfl(x1, s, GS);
            x2 = gshfl(x2, s, GS);
        }
        // group max of the error estimate (each owned layer's value; empty lanes hold 0 <= any)
        double errorMax = emax;
#pragma unroll
        for (int m = 1; m < GS; m <<= 1) errorMax = fmax90(errorMax, __shfl_xor(errorMax, m, GS));
        if (errorMax > P.xTolerUpper && dtsub > P.dtmin) {
            dtsub = fmax90(dtsub / 2.0, P.dtmin);
            continue;
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (g * K + k < nl) liq[k] = liq[k] + Ub[k] * mdz[k];
        hk_bot = gshfl(hk_bot, bot_lane, GS);
        dhkdw_bot = gshfl(dhkdw_bot, bot_lane, GS);
        dwat_bot = gshfl(dwat_bot, bot_lane, GS);
        const double qcTemp = zero_flux_bottom ? 0.0 : hk_bot + dhkdw_bot * dwat_bot;
        qcharge = qcharge + qcTemp * (dtsub / dtime);
        dtdone = dtdone + dtsub;
        if (fabs(dtime - dtdone) < P.verySmall) break;
        if (errorMax < P.xTolerLower) dtsub = dtsub * 2.0;
        dtsub = fmin90(dtsub, dtime - dtdone);
    }
    if (status) return status;

    if (wt_bottom && jwt < V.nlevsoi) {
        const double ka = gshfl(q_ka, jwt / K, GS);
        const double smp_w = gshfl(q_smp, ((jwt > 1 ? jwt : 1) - 1) / K, GS);
        qcharge = watertable_qcharge(V, S, c, jwt, ka, smp_w, dtime);
    }
    if (g == 0) {
        S.num_substeps[V.c1(c)] = (double)nsub;
        S.qcharge[V.c1(c)] = qcharge;
    }
    // over-saturated layers push water upward (:1405-1409), bottom lane first
    double carry = 0.0;
#pragma unroll
    for (int s = GS - 1; s >= 0; --s) {
        if (g == s) {
#pragma unroll
            for (int k = K - 1; k >= 0; --k) {
                const int j = g * K + k;
                if (j < nl) {
                    const double w = liq[k] + carry;
                    if (j == 0) {
                        liq[k] = w;
                    } else {
                        const double cap = S.eff_porosity[V.soi(c, j + 1)] * M_TO_MM * S.dz[V.sno(c, j + 1)];
                        carry = fmax90(w - cap, 0.0);
                        liq[k] = fmin90(cap, w);
                    }
                }
            }
        }
        carry = gshfl(carry, s, GS);
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (g * K + k < nl) S.h2osoi_liq[V.sno(c, g * K + k + 1)] = liq[k];
    return 0;
}

}  // namespace ctsm
