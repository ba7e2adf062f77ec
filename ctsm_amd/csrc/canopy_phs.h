// canopy_phs.h -- PhotosynthesisHydraulicStress (PhotosynthesisMod.F90:
// 2704-3811) for one patch inside the CanopyFluxes iteration, with the
// coupled sun / shade ci solve hybrid_PHS, brent_PHS and ci_func_PHS
// (:3815-4486) and calcstress (phs_dev.h) in it.  Included by canopy.hip
// inside namespace ctsm::canopy, after Photosynthesis.
//
// The reference calls this once per iteration for the whole filter; here a
// lane runs its patch's share, statement by statement in the reference's
// order.  The solve's helpers are out-of-line functions (ci_func_phs has
// five call sites in hybrid_phs / brent_phs and carries calcstress's Newton
// iteration): one copy of each body.  The root-soil conductances of the
// routine's first loop (:3062-3114) come from the conductance kernel run
// before the iteration, over the same inputs.
#pragma once

constexpr int PHS_SUN = 0, PHS_SHA = 1;

// Arrays and switches of the PHS path, one launch.
struct PhsArgs {
    ctsm_phs_state Q;                // k_soil_root, tsai, z, smp_l
    ctsm_phs_canopy_state H;
    const double *table;             // the PHS parameter table (phs::T_* rows)
    int nlevsoi;
    int modify;                      // modifyphoto_and_lmr_forcrop
};

// What ci_func_PHS reads and writes, and the arrays calcstress writes (the
// last call's values).
struct PhsSolve {
    bool c3, medlyn;
    double forc_pbot, cair, oair, gb_mol, rh_can, theta_cj, theta_ip, qe, bbb, mbb, mslope, mint, cp, kc, ko;
    double qsatl, qaf;
    double vcmax_z[2], tpu_z[2], kp_z[2], je[2], lmr_z[2], par_z[2];
    double ac[2], aj[2], ap[2], ag[2], an[2];
    phs::Leaf leaf;
    bool morning;
    bool stressed, tran_set, bad;
    double pd[4], tran;
};

__device__ __noinline__ void phs_calcstress(PhsSolve &s, double *x, double &bsun, double &bsha, double gs_sun,
                                            double gs_sha)
{
    double tran = 0.0;
    const bool night = phs::calcstress(s.leaf, x, gs_sun, gs_sha, s.qsatl, s.qaf, bsun, bsha, tran);
    s.stressed = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) s.pd[k] = (night && s.morning) ? x[k] : SPVAL;
    if (night) {
        s.tran_set = true;
        s.tran = tran;
    }
}

// ci_func_PHS (:4227-4486)
__device__ __noinline__ void ci_func_phs(PhsSolve &s, double *x, double cisun, double cisha, double &fsun,
                                         double &fsha, double &bsun, double &bsha, bool bflag, double gs0sun,
                                         double gs0sha, double &gs_sun, double &gs_sha)
{
    double r1, r2;
    if (bflag) phs_calcstress(s, x, bsun, bsha, gs0sun, gs0sha);
    const double ci[2] = {cisun, cisha};
    const double b[2] = {bsun, bsha};
    if (s.c3) {
#pragma unroll
        for (int k = 0; k < 2; ++k)
            s.ac[k] = b[k] * s.vcmax_z[k] * fmax90(ci[k] - s.cp, 0.0) / (ci[k] + s.kc * (1.0 + s.oair / s.ko));
#pragma unroll
        for (int k = 0; k < 2; ++k) s.aj[k] = s.je[k] * fmax90(ci[k] - s.cp, 0.0) / (4.0 * ci[k] + 8.0 * s.cp);
#pragma unroll
        for (int k = 0; k < 2; ++k) s.ap[k] = 3.0 * s.tpu_z[k];
    } else {
#pragma unroll
        for (int k = 0; k < 2; ++k) s.ac[k] = b[k] * s.vcmax_z[k];
#pragma unroll
        for (int k = 0; k < 2; ++k) s.aj[k] = s.qe * s.par_z[k] * 4.6;
#pragma unroll
        for (int k = 0; k < 2; ++k) s.ap[k] = s.kp_z[k] * fmax90(ci[k], 0.0) / s.forc_pbot;
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (!quadratic(s.theta_cj, -(s.ac[k] + s.aj[k]), s.ac[k] * s.aj[k], r1, r2)) s.bad = true;
        const double ai = fmin90(r1, r2);
        if (!quadratic(s.theta_ip, -(ai + s.ap[k]), ai * s.ap[k], r1, r2)) s.bad = true;
        s.ag[k] = fmax90(0.0, fmin90(r1, r2));
    }
    s.an[PHS_SUN] = s.ag[PHS_SUN] - bsun * s.lmr_z[PHS_SUN];
    s.an[PHS_SHA] = s.ag[PHS_SHA] - bsha * s.lmr_z[PHS_SHA];
    const double gsmin = s.medlyn ? s.mint : s.bbb;
    if (s.an[PHS_SUN] < 0.0) {
        gs_sun = fmax90(bsun * gsmin, 1.0);
        fsun = 0.0;
    }
    if (s.an[PHS_SHA] < 0.0) {
        gs_sha = fmax90(bsha * gsmin, 1.0);
        fsha = 0.0;
    }
    if (s.an[PHS_SUN] < 0.0 && s.an[PHS_SHA] < 0.0) return;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (s.an[k] < 0.0) continue;
        double &gs = k == PHS_SUN ? gs_sun : gs_sha;
        double cs = s.cair - 1.4 / s.gb_mol * s.an[k] * s.forc_pbot;
        cs = fmax90(cs, MAX_CS);
        if (s.medlyn) {
            const double term = 1.6 * s.an[k] / (cs / s.forc_pbot * 1.e06);
            const double bq = -(2.0 * (s.mint * 1.e-06 + term)
                                + (s.mslope * term) * (s.mslope * term) / (s.gb_mol * 1.e-06 * s.rh_can));
            const double cq = s.mint * s.mint * 1.e-12
                              + (2.0 * s.mint * 1.e-06 + term * (1.0 - s.mslope * s.mslope / s.rh_can)) * term;
            if (!quadratic(1.0, bq, cq, r1, r2)) s.bad = true;
            gs = fmax90(r1, r2) * 1.e06;
        } else {
            const double bk = k == PHS_SUN ? bsun : bsha;
            const double bq = cs * (s.gb_mol - fmax90(bk * s.bbb, 1.0)) - s.mbb * s.an[k] * s.forc_pbot;
            const double cq = -s.gb_mol * (cs * fmax90(bk * s.bbb, 1.0) + s.mbb * s.an[k] * s.forc_pbot * s.rh_can);
            if (!quadratic(cs, bq, cq, r1, r2)) s.bad = true;
            gs = fmax90(r1, r2);
        }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (s.an[k] < 0.0) continue;
        const double gs = k == PHS_SUN ? gs_sun : gs_sha;
        double &fv = k == PHS_SUN ? fsun : fsha;
        if (gs > 0.0) fv = ci[k] - s.cair + s.an[k] * s.forc_pbot * (1.4 * gs + 1.6 * s.gb_mol) / (s.gb_mol * gs);
        else fv = ci[k] - s.cair;
    }
}

// brent_PHS (:4068-4223): both phases' brackets shrink together; tol is the
// sunlit tolerance for both
__device__ __noinline__ void brent_phs(PhsSolve &s, double &xsun, double x1sun, double x2sun, double f1sun,
                                       double f2sun, double &xsha, double x1sha, double x2sha, double f1sha,
                                       double f2sha, double tol, double &gs_sun, double &gs_sha, double &bsun,
                                       double &bsha)
{
    const int itmax = 20;
    const double eps = 1.e-4;
    double a[2] = {x1sun, x1sha}, b[2] = {x2sun, x2sha}, fa[2] = {f1sun, f1sha}, fb[2] = {f2sun, f2sha};
    double c[2], fc[2], d[2] = {0.0, 0.0}, e[2] = {0.0, 0.0}, tol1[2], xm[2];
    double dummy[4] = {0.0, 0.0, 0.0, 0.0};   // calcstress is not called from here (bflag false)
#pragma unroll
    for (int k = 0; k < 2; ++k)
        if ((fa[k] > 0.0 && fb[k] > 0.0) || (fa[k] < 0.0 && fb[k] < 0.0)) {
            s.bad = true;
            xsun = b[PHS_SUN];
            xsha = b[PHS_SHA];
            return;
        }
    c[0] = b[0];
    c[1] = b[1];
    fc[0] = fb[0];
    fc[1] = fb[1];
    int iter = 0;
    for (;;) {
        if (iter == itmax) break;
        iter = iter + 1;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if ((fb[k] > 0.0 && fc[k] > 0.0) || (fb[k] < 0.0 && fc[k] < 0.0)) {
                c[k] = a[k];
                fc[k] = fa[k];
                d[k] = b[k] - a[k];
                e[k] = d[k];
            }
            if (fabs(fc[k]) < fabs(fb[k])) {
                a[k] = b[k];
                b[k] = c[k];
                c[k] = a[k];
                fa[k] = fb[k];
                fb[k] = fc[k];
                fc[k] = fa[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) tol1[k] = 2.0 * eps * fabs(b[k]) + 0.5 * tol;
#pragma unroll
        for (int k = 0; k < 2; ++k) xm[k] = 0.5 * (c[k] - b[k]);
        if ((fabs(xm[PHS_SUN]) <= tol1[PHS_SUN] || fb[PHS_SUN] == 0.0) &&
            (fabs(xm[PHS_SHA]) <= tol1[PHS_SHA] || fb[PHS_SHA] == 0.0)) {
            xsun = b[PHS_SUN];
            xsha = b[PHS_SHA];
            return;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (fabs(e[k]) >= tol1[k] && fabs(fa[k]) > fabs(fb[k])) {
                const double ss = fb[k] / fa[k];
                double p, q;
                if (a[k] == c[k]) {
                    p = 2.0 * xm[k] * ss;
                    q = 1.0 - ss;
                } else {
                    q = fa[k] / fc[k];
                    const double r = fb[k] / fc[k];
                    p = ss * (2.0 * xm[k] * q * (q - r) - (b[k] - a[k]) * (r - 1.0));
                    q = (q - 1.0) * (r - 1.0) * (ss - 1.0);
                }
                if (p > 0.0) q = -q;
                p = fabs(p);
                if (2.0 * p < fmin90(3.0 * xm[k] * q - fabs(tol1[k] * q), fabs(e[k] * q))) {
                    e[k] = d[k];
                    d[k] = p / q;
                } else {
                    d[k] = xm[k];
                    e[k] = d[k];
                }
            } else {
                d[k] = xm[k];
                e[k] = d[k];
            }
            a[k] = b[k];
            fa[k] = fb[k];
            if (fabs(d[k]) > tol1[k]) b[k] = b[k] + d[k];
            else b[k] = b[k] + copysign(fabs(tol1[k]), xm[k]);
        }
        const double gs0sun = gs_sun, gs0sha = gs_sha;
        ci_func_phs(s, dummy, b[PHS_SUN], b[PHS_SHA], fb[PHS_SUN], fb[PHS_SHA], bsun, bsha, false, gs0sun, gs0sha,
                    gs_sun, gs_sha);
        if (fb[PHS_SUN] == 0.0 && fb[PHS_SHA] == 0.0) break;
    }
    xsun = b[PHS_SUN];
    xsha = b[PHS_SHA];
}

// hybrid_PHS (:3815-4064): the outer loop updates bsun / bsha through
// calcstress, the inner one is the secant / Brent search for both ci.
// vegwp (4) inout; returns soilflux clipped at 0.
__device__ __noinline__ double hybrid_phs(PhsSolve &s, double &x0sun, double &x0sha, double *vegwp, double &bsun,
                                          double &bsha, double &gs_sun, double &gs_sha)
{
    const double toldb = 1.e-2, eps = 1.e-2, eps1 = 1.e-4;
    const int itmax = 3;
    double x[4];
    double x1sun = x0sun, x1sha = x0sha, f0sun = 0.0, f0sha = 0.0, f1sun = 0.0, f1sha = 0.0;
    double minf = 0.0, minxsun = 0.0, minxsha = 0.0;
    bool bflag = false;
    double b0sun = -1.0, b0sha = -1.0, gs0sun = 0.0, gs0sha = 0.0;
    bsun = 1.0;
    bsha = 1.0;
    int iter1 = 0;
    for (;;) {
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = vegwp[k];
        iter1 = iter1 + 1;
        int iter2 = 0;
        x0sun = fmax90(0.1, x1sun);
        x1sun = 0.99 * x1sun;
        x0sha = fmax90(0.1, x1sha);
        x1sha = 0.99 * x1sha;
        const double tolsun = fabs(x1sun) * eps;
        const double tolsha = fabs(x1sha) * eps;
        ci_func_phs(s, x, x0sun, x0sha, f0sun, f0sha, bsun, bsha, bflag, gs0sun, gs0sha, gs_sun, gs_sha);
        const double dbsun = b0sun - bsun;
        const double dbsha = b0sha - bsha;
        b0sun = bsun;
        b0sha = bsha;
        bflag = false;
        ci_func_phs(s, x, x1sun, x1sha, f1sun, f1sha, bsun, bsha, bflag, gs0sun, gs0sha, gs_sun, gs_sha);
        for (;;) {
            if (fabs(f0sun) < eps1 && fabs(f0sha) < eps1) {
                x1sun = x0sun;
                x1sha = x0sha;
                break;
            }
            if (fabs(f1sun) < eps1 && fabs(f1sha) < eps1) break;
            iter2 = iter2 + 1;
            double dxsun, dxsha;
            if ((f1sun - f0sun) == 0.0) dxsun = 0.5 * (x1sun + x0sun) - x1sun;
            else dxsun = -f1sun * (x1sun - x0sun) / (f1sun - f0sun);
            if ((f1sha - f0sha) == 0.0) dxsha = 0.5 * (x1sha + x0sha) - x1sha;
            else dxsha = -f1sha * (x1sha - x0sha) / (f1sha - f0sha);
            x0sun = x1sun;
            x1sun = x1sun + dxsun;
            x0sha = x1sha;
            x1sha = x1sha + dxsha;
            ci_func_phs(s, x, x1sun, x1sha, f1sun, f1sha, bsun, bsha, bflag, gs0sun, gs0sha, gs_sun, gs_sha);
            if (fabs(dxsun) < tolsun && fabs(dxsha) < tolsha) {
                x0sun = x1sun;
                x0sha = x1sha;
                break;
            }
            if (iter2 == 1) {
                minf = fabs(f1sun + f1sha);
                minxsun = x1sun;
                minxsha = x1sha;
            } else if (fabs(f1sun + f1sha) < minf) {
                minf = fabs(f1sun + f1sha);
                minxsun = x1sun;
                minxsha = x1sha;
            }
            if (fabs(f1sun) < eps1 && fabs(f1sha) < eps1) break;
            if (f1sun * f0sun < 0.0 && f1sha * f0sha < 0.0) {
                double xsun, xsha;
                brent_phs(s, xsun, x0sun, x1sun, f0sun, f1sun, xsha, x0sha, x1sha, f0sha, f1sha, tolsun, gs_sun,
                          gs_sha, bsun, bsha);
                x0sun = xsun;
                x0sha = xsha;
                break;
            }
            if (iter2 > itmax) {
                x1sun = minxsun;
                x1sha = minxsha;
                ci_func_phs(s, x, x1sun, x1sha, f1sun, f1sha, bsun, bsha, bflag, gs0sun, gs0sha, gs_sun, gs_sha);
                break;
            }
        }
        if (bsun > 0.01) gs0sun = gs_sun / bsun;
        if (bsha > 0.01) gs0sha = gs_sha / bsha;
        bflag = true;
        if (fabs(dbsun) < toldb && fabs(dbsha) < toldb) break;
        if (iter1 > itmax) break;
    }
    x0sun = x1sun;
    x0sha = x1sha;
    double soilflux = phs::getvegwp(s.leaf, x, gs_sun, gs_sha, s.qsatl, s.qaf);
#pragma unroll
    for (int k = 0; k < 4; ++k) vegwp[k] = x[k];
    if (soilflux < 0.0) soilflux = 0.0;
    return soilflux;
}

// PhotosynthesisHydraulicStress (:3117-3811) for patch p at the iteration's
// leaf temperature: fills the outputs both paths share into sun / sha / sh
// (written after the iteration, as Photosynthesis's are), writes the PHS
// arrays, vegwp, the noon conductances, vpd_can and qflx_tran_veg where the
// reference does, and returns btran.  qsatl and qaf are the iteration's.
template <class PRM, class CAN, class PHA>
__device__ __noinline__ double photosynthesis_phs(const CView &V, const CSwitch &sw, const PRM &P, const PsnConst &K,
                                                  const TFac &T, const PHA &A, const CAN &S, int p, int c,
                                                  int ivt, int nrad, bool noon, double forc_pbot, double tgcm,
                                                  double esat_tv, double eair, double oair, double cair, double rb,
                                                  double qsatl, double qaf, Phase &sun, Phase &sha, Shared &sh,
                                                  bool &bad)
{
    const auto &H = A.H;
    const long pp = V.p1(p), cc = V.c1(c), ldp = V.ldp;
    const bool medlyn = sw.stomatal == CTSM_STOMATA_MEDLYN2011;
    const double par_z[2] = {S.parsun_z[pp], S.parsha_z[pp]};
    const double lai_z[2] = {S.laisun_z[pp], S.laisha_z[pp]};
    const double vcmaxcint[2] = {S.vcmaxcintsun[pp], S.vcmaxcintsha[pp]};
    const double o3coefv[2] = {S.o3coefvsun[pp], S.o3coefvsha[pp]};
    const double o3coefg[2] = {S.o3coefgsun[pp], S.o3coefgsha[pp]};
    const bool crop = A.table[phs::T_CROP * CTSM_MAX_PFT + ivt] != 0.0;
    const bool modify = A.modify != 0;
    Phase *ph[2] = {&sun, &sha};
    double bb[2] = {H.bsun[pp], H.bsha[pp]};
    double lmr_z[2] = {0.0, 0.0}, vcmax_z[2] = {0.0, 0.0}, jmax_z[2] = {0.0, 0.0}, tpu_z[2] = {0.0, 0.0},
           kp_z[2] = {0.0, 0.0};
    // the layer loop (:3327-3480): respiration always, capacities by day
    if (nrad >= 1) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double lmr25 = K.lmr25top * vcmaxcint[k];
            if (K.c3) {
                lmr_z[k] = lmr25 * T.ftL * T.fthL;
            } else {
                lmr_z[k] = lmr25 * T.pow2;
                lmr_z[k] = lmr_z[k] / T.dL4;
            }
        }
        const double tlai_z = H.tlai_z[pp];
#pragma unroll
        for (int k = 0; k < 2; ++k) lmr_z[k] = lmr_z[k] * fmin90((0.2 * cm_exp(3.218 * tlai_z)), 1.0);
        if (par_z[PHS_SUN] > 0.0) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                vcmax_z[k] = K.vcmax25top * vcmaxcint[k] * T.ftV * T.fthV;
                jmax_z[k] = K.jmax25top * vcmaxcint[k] * T.ftJ * T.fthJ;
                tpu_z[k] = K.tpu25top * vcmaxcint[k] * T.ftT * T.fthT;
            }
            if (!K.c3) {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    vcmax_z[k] = K.vcmax25top * vcmaxcint[k] * T.pow2;
                    vcmax_z[k] = vcmax_z[k] / T.dV1;
                    vcmax_z[k] = vcmax_z[k] / T.dV2;
                }
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) kp_z[k] = K.kp25top * vcmaxcint[k] * T.pow2;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k)
            if (sw.light_inhibit && par_z[k] > 0.0) lmr_z[k] = lmr_z[k] * 0.67;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            H.vcmax_z_phs[pp + k * ldp] = vcmax_z[k];
            H.tpu_z_phs[pp + k * ldp] = tpu_z[k];
            H.kp_z_phs[pp + k * ldp] = kp_z[k];
        }
    }

    // leaf-level photosynthesis and conductance (:3485-3700)
    const double rsmax0 = 2.e4;
    const double cf = forc_pbot / (RGAS * 1.e-3 * tgcm) * 1.e06;
    const double gb = 1.0 / rb;
    sh.gb_mol = gb * cf;
    const double gb_mol = sh.gb_mol;
    double psn_wc_z[2] = {0.0, 0.0}, psn_wj_z[2] = {0.0, 0.0}, psn_wp_z[2] = {0.0, 0.0};
    if (nrad >= 1) {
        PhsSolve s;
        s.leaf.laisun = S.laisun[pp];
        s.leaf.laisha = S.laisha[pp];
        s.leaf.elai = S.elai[pp];
        s.leaf.esai = S.esai[pp];
        s.leaf.tsai = A.Q.tsai[pp];
        s.leaf.htop = S.htop[pp];
        s.leaf.fdry = S.fdry[pp];
        s.leaf.thm = tgcm;
        s.leaf.rho = S.forc_rho[cc];
        s.leaf.pbot = forc_pbot;
        s.leaf.gb = gb_mol;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            s.leaf.kmax[k] = A.table[(phs::T_KMAX + k) * CTSM_MAX_PFT + ivt];
            s.leaf.psi50[k] = A.table[(phs::T_PSI50 + k) * CTSM_MAX_PFT + ivt];
            s.leaf.ck[k] = A.table[(phs::T_CK + k) * CTSM_MAX_PFT + ivt];
        }
        s.leaf.k = A.Q.k_soil_root + V.soip(p, 1);
        s.leaf.smp = A.Q.smp_l + V.soi(c, 1);
        s.leaf.z = A.Q.z + V.sno(c, 1);
        s.leaf.sk = ldp;
        s.leaf.ss = V.ldc;
        s.leaf.n = A.nlevsoi;
        s.qsatl = qsatl;
        s.qaf = qaf;
        s.morning = H.local_morning[pp] != 0;
        s.stressed = false;
        s.tran_set = false;
        s.bad = false;
        s.tran = 0.0;
        const double gsmin = medlyn ? K.mint : K.bbb;
        double vegwp[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) vegwp[k] = H.vegwp[pp + k * ldp];
        if (par_z[PHS_SUN] <= 0.0) {
            // night: the stress factors at the minimum conductance (:3500-3548)
            vegwp[0] = 1.0;
            phs_calcstress(s, vegwp, bb[PHS_SUN], bb[PHS_SHA], gsmin, gsmin);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                H.ac_phs[pp + k * ldp] = 0.0;
                H.aj_phs[pp + k * ldp] = 0.0;
                H.ap_phs[pp + k * ldp] = 0.0;
                H.ag_phs[pp + k * ldp] = 0.0;
                const double an = (!crop || !modify) ? 0.0 - bb[k] * lmr_z[k] : 0.0 - lmr_z[k];
                (k == PHS_SUN ? H.an_sun : H.an_sha)[pp] = an;
                ph[k]->psn_z = 0.0;
                ph[k]->rs_z = fmin90(rsmax0, 1.0 / (fmax90(bb[k] * gsmin, 1.0)) * cf);
                ph[k]->ci_z = 0.0;
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) ph[k]->gs_ph = cf / ph[k]->rs_z;
        } else {
            // day (:3550-3697)
            const double ceair = fmin90(eair, esat_tv);
            double rh_can;
            if (!medlyn) {
                rh_can = ceair / esat_tv;
            } else {
                rh_can = fmax90((esat_tv - ceair), 50.0) * 0.001;
                S.vpd_can[pp] = rh_can;
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const double qabs = 0.5 * (1.0 - P.fnps) * par_z[k] * 4.6;
                double r1, r2;
                if (!quadratic(P.theta_psii, -(qabs + jmax_z[k]), qabs * jmax_z[k], r1, r2)) bad = true;
                s.je[k] = fmin90(r1, r2);
            }
            double ci[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) ci[k] = K.c3 ? 0.7 * cair : 0.4 * cair;
            s.c3 = K.c3;
            s.medlyn = medlyn;
            s.forc_pbot = forc_pbot;
            s.cair = cair;
            s.oair = oair;
            s.gb_mol = gb_mol;
            s.rh_can = rh_can;
            s.theta_cj = K.theta_cj;
            s.theta_ip = P.theta_ip;
            s.qe = K.qe;
            s.bbb = K.bbb;
            s.mbb = K.mbb;
            s.mslope = K.mslope;
            s.mint = K.mint;
            s.cp = T.cp;
            s.kc = T.kc;
            s.ko = T.ko;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                s.vcmax_z[k] = vcmax_z[k];
                s.tpu_z[k] = tpu_z[k];
                s.kp_z[k] = kp_z[k];
                s.lmr_z[k] = lmr_z[k];
                s.par_z[k] = par_z[k];
                s.ac[k] = H.ac_phs[pp + k * ldp];
                s.aj[k] = H.aj_phs[pp + k * ldp];
                s.ap[k] = H.ap_phs[pp + k * ldp];
                s.ag[k] = H.ag_phs[pp + k * ldp];
            }
            s.an[PHS_SUN] = H.an_sun[pp];
            s.an[PHS_SHA] = H.an_sha[pp];
            double gs[2] = {S.gs_mol_sun[pp], S.gs_mol_sha[pp]};
            const double soilflux = hybrid_phs(s, ci[PHS_SUN], ci[PHS_SHA], vegwp, bb[PHS_SUN], bb[PHS_SHA],
                                               gs[PHS_SUN], gs[PHS_SHA]);
            s.tran_set = true;
            s.tran = soilflux;
            if (s.bad) bad = true;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                H.ac_phs[pp + k * ldp] = s.ac[k];
                H.aj_phs[pp + k * ldp] = s.aj[k];
                H.ap_phs[pp + k * ldp] = s.ap[k];
                H.ag_phs[pp + k * ldp] = s.ag[k];
            }
            H.an_sun[pp] = s.an[PHS_SUN];
            H.an_sha[pp] = s.an[PHS_SHA];
#pragma unroll
            for (int k = 0; k < 4; ++k) H.vegwp_ln[pp + k * ldp] = noon ? vegwp[k] : SPVAL;
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (s.an[k] < 0.0) gs[k] = fmax90(bb[k] * gsmin, 1.0);
            S.gs_mol_sun_ln[pp] = noon ? gs[PHS_SUN] : SPVAL;
            S.gs_mol_sha_ln[pp] = noon ? gs[PHS_SHA] : SPVAL;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double ciz = cair - s.an[k] * forc_pbot * (1.4 * gs[k] + 1.6 * gb_mol) / (gb_mol * gs[k]);
                ph[k]->ci_z = fmax90(ciz, 1.e-06);
                const double g = gs[k] / cf;
                double rsz = fmin90(1.0 / g, rsmax0);
                ph[k]->rs_z = rsz / o3coefg[k];
                double psnz = s.ag[k];
                ph[k]->psn_z = psnz * o3coefv[k];
                const double ac = s.ac[k], aj = s.aj[k], ap = s.ap[k];
                if (ac <= aj && ac <= ap) psn_wc_z[k] = ph[k]->psn_z;
                else if (aj < ac && aj <= ap) psn_wj_z[k] = ph[k]->psn_z;
                else if (ap < ac && ap < aj) psn_wp_z[k] = ph[k]->psn_z;
                ph[k]->gs_ph = gs[k];
            }
            // the leaf-surface humidity and the Ball-Berry check (:3677-3695) only log
            if (gs[PHS_SUN] < 0.0 || gs[PHS_SHA] < 0.0) bad = true;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) H.vegwp[pp + k * ldp] = vegwp[k];
        if (s.stressed) {
#pragma unroll
            for (int k = 0; k < 4; ++k) H.vegwp_pd[pp + k * ldp] = s.pd[k];
        }
        if (s.tran_set) S.qflx_tran_veg[pp] = s.tran;
        H.bsun[pp] = bb[PHS_SUN];
        H.bsha[pp] = bb[PHS_SHA];
#pragma unroll
        for (int k = 0; k < 2; ++k) ph[k]->lmr_z = lmr_z[k];
    }

    // canopy sums (:3705-3800) and btran
    double laican[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double psncan = 0.0, psncan_wc = 0.0, psncan_wj = 0.0, psncan_wp = 0.0, lmrcan = 0.0, gscan = 0.0;
        if (nrad >= 1) {
            psncan = psncan + ph[k]->psn_z * lai_z[k];
            psncan_wc = psncan_wc + psn_wc_z[k] * lai_z[k];
            psncan_wj = psncan_wj + psn_wj_z[k] * lai_z[k];
            psncan_wp = psncan_wp + psn_wp_z[k] * lai_z[k];
            if (!crop && modify) lmrcan = lmrcan + lmr_z[k] * lai_z[k] * bb[k];
            else lmrcan = lmrcan + lmr_z[k] * lai_z[k];
            gscan = gscan + lai_z[k] / (rb + ph[k]->rs_z);
            laican[k] = laican[k] + lai_z[k];
        }
        if (laican[k] > 0.0) {
            ph[k]->psn = psncan / laican[k];
            ph[k]->psn_wc = psncan_wc / laican[k];
            ph[k]->psn_wj = psncan_wj / laican[k];
            ph[k]->psn_wp = psncan_wp / laican[k];
            ph[k]->lmr = lmrcan / laican[k];
            ph[k]->rs = laican[k] / gscan - rb;
        } else {
            ph[k]->psn = 0.0;
            ph[k]->psn_wc = 0.0;
            ph[k]->psn_wj = 0.0;
            ph[k]->psn_wp = 0.0;
            ph[k]->lmr = 0.0;
            ph[k]->rs = 0.0;
        }
    }
    if (laican[PHS_SHA] + laican[PHS_SUN] > 0.0)
        return bb[PHS_SUN] * (laican[PHS_SUN] / (laican[PHS_SUN] + laican[PHS_SHA]))
               + bb[PHS_SHA] * (laican[PHS_SHA] / (laican[PHS_SUN] + laican[PHS_SHA]));
    return bb[PHS_SUN];
}
