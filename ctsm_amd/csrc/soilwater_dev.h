// soilwater_dev.h -- the per-column moisture-form Richards solve, shared by
// the soil-water kernel and the fused soil-step kernel.  See soilwater.hip.
#pragma once
#include "common.h"
#include "watertable_bc.h"

namespace ctsm {

// Column inputs that stay constant during the sub-step loop.
struct RichardsStatic {
    double mdz;     // m_to_mm*dz(j)   (= dz*denh2o)
    double ws;      // watsat(j)
    double bsw;     // bsw(j)
    double hkc;     // imped(j)*hksat(j)
    double suc;     // sucsat(j)
    double den;     // m_to_mm*(z(j+1)-z(j))
    double sink;    // qflx_rootsoi_col(j)
};

// WT: the water-table bottom boundary (watertable_bc.h), a separate
// instantiation so the flux / zero-flux kernels keep their registers.
template <int ML, bool WT = false>
__device__ __forceinline__ int richards_column(const View &V, const Params &P, const ctsm_col_state &S,
                                               int c, double dtime)
{
    const int nl = S.nbedrock[V.c1(c)];
    if (nl < 2 || nl > ML) return 1;

    double mdz[ML], ws[ML], bsw[ML], hkc[ML], suc[ML], den[ML], sink[ML], liq[ML];
    // per-trial factorisation (U rows of dgtsv) and start-of-step net flux
    double Ud[ML], Udu[ML], Udl[ML], Ub[ML], f1[ML];

    const double infl = S.qflx_infl[V.c1(c)];
#pragma unroll
    for (int j = 0; j < ML; ++j) {
        if (j < nl) {
            const int jj = j + 1;
            const double dz = S.dz[V.sno(c, jj)];
            mdz[j] = M_TO_MM * dz;
            ws[j] = S.watsat[V.soi(c, jj)];
            bsw[j] = S.bsw[V.soi(c, jj)];
            suc[j] = S.sucsat[V.soi(c, jj)];
            sink[j] = S.qflx_rootsoi[V.soi(c, jj)];
            liq[j] = S.h2osoi_liq[V.sno(c, jj)];
            // IceImpedance (:2139-2160) of the interface below layer j (node at the bottom)
            const double ic = S.icefrac[V.soi(c, jj)];
            double icei;
            if (j == nl - 1) {
                icei = ic;
                den[j] = 0.0;
            } else {
                icei = 0.5 * (ic + S.icefrac[V.soi(c, jj + 1)]);
                den[j] = M_TO_MM * (S.z[V.sno(c, jj + 1)] - S.z[V.sno(c, jj)]);
            }
            hkc[j] = cm_pow(10.0, -(P.e_ice * icei)) * S.hksat[V.soi(c, jj)];
        }
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
    int nsub = 0;
    double dtsub = dtime, dtdone = 0.0, qcharge = 0.0;
    int status = 0;

    for (;;) {
        nsub = nsub + 1;
        if (nsub > MAX_TRIALS) { status = 3; break; }
        // A trial that would finish the step if accepted writes the per-layer
        // diagnostics (smp_l, hk_l, qin, qout); the last such trial is the
        // accepted final one, as in the reference where they keep the values
        // of the last trial.
        const bool candidate = fabs(dtime - (dtdone + dtsub)) < P.verySmall;

        // ---------------- forward sweep ----------------
        // node properties of layer 0
        double vwc = fmax90(liq[0], 1.0e-6) / mdz[0];
        double s2 = fmin90(vwc / ws[0], 1.0);
        s2 = fmax90(0.01, s2);
        double smp = -(suc[0] * cm_pow(s2, -bsw[0]));
        double dsmpdw = (-(bsw[0] * smp) / s2) / ws[0];
        double dtdz = dtsub / mdz[0];

        double qin = infl;           // bc_flux top (:1630-1637)
        double dqidw0 = 0.0, dqidw1 = 0.0;
        // dgtsv window: current row i (d, du, b) and its sub-diagonal to row i+1
        double cd = 0.0, cdu = 0.0, cb = 0.0;
        double hk_bot = 0.0, dhkdw_bot = 0.0;
        bool singular = false;

#pragma unroll
        for (int j = 0; j < ML; ++j) {
            if (j < nl) {
                // properties of the node below (lookahead)
                double s2n = 0.0, smpn = 0.0, dsmpdwn = 0.0, dtdzn = 0.0;
                if (j + 1 < ML && j + 1 < nl) {
                    const double vwcn = fmax90(liq[j + 1], 1.0e-6) / mdz[j + 1];
                    s2n = fmin90(vwcn / ws[j + 1], 1.0);
                    s2n = fmax90(0.01, s2n);
                    smpn = -(suc[j + 1] * cm_pow(s2n, -bsw[j + 1]));
                    dsmpdwn = (-(bsw[j + 1] * smpn) / s2n) / ws[j + 1];
                    dtdzn = dtsub / mdz[j + 1];
                }
                const bool bottom = (j == nl - 1);
                double s1 = bottom ? s2 : 0.5 * (s2 + s2n);
                s1 = fmin90(s1, 1.0);
                s1 = fmax90(0.01, s1);
                const double e = 2.0 * bsw[j] + 3.0;
                const double hk = hkc[j] * cm_pow(s1, e);
                const double dhkdw = e * hk / s1;
                if (candidate) {
                    S.smp_l[V.soi(c, j + 1)] = smp;
                    S.hk_l[V.soi(c, j + 1)] = hk;
                }
                double qout, dqodw1, dqodw2;
                if (!bottom) {
                    const double dhkds1 = 0.5 * dhkdw / ws[j];
                    const double dhkds2 = 0.5 * dhkdw / ws[j + 1 < ML ? j + 1 : j];
                    const double num = smpn - smp;
                    qout = -(hk * num / den[j]) + hk;
                    dqodw1 = (hk * dsmpdw - dhkds1 * num) / den[j] + dhkds1;
                    dqodw2 = (-(hk * dsmpdwn) - dhkds2 * num) / den[j] + dhkds2;
                } else {
                    hk_bot = hk;
                    dhkdw_bot = dhkdw;
                    if (zero_flux_bottom) {
                        qout = 0.0;
                        dqodw1 = 0.0;
                    } else if (wt_bottom) {
                        if (jwt < V.nlevsoi) { qout = 0.0; dqodw1 = 0.0; }
                        else watertable_bottom_flux(hk, dhkdw, smp, dsmpdw, ws[j], zwt, z_bot, qout, dqodw1);
                        // the sub-step's drainage is qout + dqodw1 * dwat
                        hk_bot = qout;
                        dhkdw_bot = dqodw1;
                    } else {
                        qout = hk;
                        dqodw1 = dhkdw / ws[j];
                    }
                    dqodw2 = 0.0;
                }
                if (candidate && wt_bottom) {
                    // what compute_qcharge reads of the last trial
                    if (j == jwt) q_ka = watertable_ka(fmax90(liq[j], 1.0e-6) / mdz[j], ws[j], hkc[j], bsw[j]);
                    if (j == (jwt > 1 ? jwt : 1) - 1) q_smp = smp;
                }
                if (candidate) {
                    S.qin[V.soi(c, j + 1)] = qin;
                    S.qout[V.soi(c, j + 1)] = qout;
                }
                // RHS / LHS row j
                const double fluxNet = qin - qout - sink[j];
                f1[j] = fluxNet;
                const double r = -(fluxNet * dtdz);
                const double a = dqidw0 * dtdz;
                const double b = -1.0 - (-dqidw1 + dqodw1) * dtdz;
                const double cc = bottom ? 0.0 : -(dqodw2 * dtdz);

                if (j == 0) {
                    cd = b; cdu = cc; cb = r;
                } else {
                    // dgtsv elimination of row j-1 against row j (dl = a).
                    // Rows j-1 < n-2 also carry the second super-diagonal fill.
                    double nd = b, ndu = cc, nb = r, dl = a;
                    const bool last = bottom;   // i = n-2 in LAPACK's 1-based N-1
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
                        if (!last) {
                            dl = ndu;
                            ndu = -fact * dl;
                        }
                        const double du_i = temp;
                        const double tb = cb;
                        const double b_i = nb;
                        nb = tb - fact * nb;
                        cd = d_i; cdu = du_i; cb = b_i;
                    }
                    Ud[j - 1] = cd; Udu[j - 1] = cdu; Udl[j - 1] = dl; Ub[j - 1] = cb;
                    cd = nd; cdu = ndu; cb = nb;
                }
                if (bottom) {
                    Ud[j] = cd; Udu[j] = 0.0; Udl[j] = 0.0; Ub[j] = cb;
                    if (cd == 0.0) singular = true;
                }
                // shift to the next layer
                qin = qout;
                dqidw0 = dqodw1;
                dqidw1 = dqodw2;
                s2 = s2n; smp = smpn; dsmpdw = dsmpdwn; dtdz = dtdzn;
            }
        }
        if (singular) { status = 2; break; }

        // ---------------- backward sweep: dgtsv back substitution + error ----------------
        double x1 = 0.0, x2 = 0.0;      // dwat of the two layers below
        double errorMax = 0.0;
        double dwat_bot = 0.0;
#pragma unroll
        for (int j = ML - 1; j >= 0; --j) {
            if (j < nl) {
                double x;
                if (j == nl - 1)      x = Ub[j] / Ud[j];
                else if (j == nl - 2) x = (Ub[j] - Udu[j] * x1) / Ud[j];
                else                  x = (Ub[j] - Udu[j] * x1 - Udl[j] * x2) / Ud[j];
                if (j == nl - 1) dwat_bot = x;
                // error estimate (inexpensive flux, :1341): |fluxNet1 - dwat/dt_dz|*dtsub/2
                const double dtdz = dtsub / mdz[j];
                const double ev = fabs(f1[j] - x / dtdz) * dtsub * 0.5;
                if (j == nl - 1 || ev > errorMax) errorMax = ev;   // maxval: ties keep the value
                Ub[j] = x;   // keep dwat
                x2 = x1;
                x1 = x;
            }
        }
        if (errorMax > P.xTolerUpper && dtsub > P.dtmin) {
            dtsub = fmax90(dtsub / 2.0, P.dtmin);
            continue;
        }
#pragma unroll
        for (int j = 0; j < ML; ++j)
            if (j < nl) liq[j] = liq[j] + Ub[j] * mdz[j];
        const double qcTemp = zero_flux_bottom ? 0.0 : hk_bot + dhkdw_bot * dwat_bot;
        qcharge = qcharge + qcTemp * (dtsub / dtime);
        dtdone = dtdone + dtsub;
        if (fabs(dtime - dtdone) < P.verySmall) break;
        if (errorMax < P.xTolerLower) dtsub = dtsub * 2.0;
        dtsub = fmin90(dtsub, dtime - dtdone);
    }
    if (status) return status;

    if (wt_bottom && jwt < V.nlevsoi) qcharge = watertable_qcharge(V, S, c, jwt, q_ka, q_smp, dtime);
    S.num_substeps[V.c1(c)] = (double)nsub;
    S.qcharge[V.c1(c)] = qcharge;
    // over-saturated layers push water upward (:1405-1409)
    double carry = 0.0;
#pragma unroll
    for (int j = ML - 1; j >= 0; --j) {
        if (j < nl) {
            double w = liq[j] + carry;
            if (j == 0) {
                liq[j] = w;
            } else {
                const double cap = S.eff_porosity[V.soi(c, j + 1)] * M_TO_MM * S.dz[V.sno(c, j + 1)];
                const double over = fmax90(w - cap, 0.0);
                liq[j] = fmin90(cap, w);
                carry = over;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < ML; ++j)
        if (j < nl) S.h2osoi_liq[V.sno(c, j + 1)] = liq[j];
    return 0;
}

// Device entry used by the fused soil step.
__device__ __forceinline__ int soilwater_fractions_column(const View &V, const ctsm_col_state &S, int c)
{
    for (int j = 1; j <= V.nlevsoi; ++j) {
        const double ws = S.watsat[V.soi(c, j)];
        const double dz_ext = S.dz[V.sno(c, j)] + 0.0 / DENICE;
        const double vol_ice = fmin90(ws, (S.h2osoi_ice[V.sno(c, j)] + 0.0) / (dz_ext * DENICE));
        S.eff_porosity[V.soi(c, j)] = fmax90(0.01, ws - vol_ice);
        S.icefrac[V.soi(c, j)] = fmin90(1.0, vol_ice / ws);
    }
    return 0;
}

}  // namespace ctsm
