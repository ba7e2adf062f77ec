// plain.hip -- the unfused baseline: one launch per reference routine.
//
// This is the straightforward GPU form of the reference's structure, kept as
// the comparison point for the fused kernels: every routine of
// SoilTemperature (SoilThermProp, ComputeGroundHeatFluxAndDeriv,
// ComputeHeatDiffFluxAndFactor, SetRHSVec, SetMatrix, BandDiagonal,
// PhaseChangeH2osfc, Phasechange, the t_grnd / flux tail) is its own kernel,
// and every intermediate the reference keeps in a (begc:endc, lev) array
// (tk, cv, fn, fn1, sabg_lyr_col, rvector, tvector, bmatrix, the dgbsv band
// matrix) lives in HBM between launches.  The Richards solve is one kernel
// per call (the reference's per-column sub-step loop) whose per-layer arrays
// (hk, smp, fluxes, amx/bmx/cmx/rmx, dwat) are HBM-resident too.  Same
// expressions and order as the reference, so results equal the fused path's.
#include "common.h"
#include "trace.h"
#include "soiltemp_dev.h"
#include "watertable_bc.h"

namespace ctsm {

namespace {

constexpr int NL = MAX_NLEVSNO + MAX_NLEVGRND + 1;   // levels -nlevsno .. nlevgrnd
// workspace slots (each slot = one double per lane)
constexpr int S_TK = 0;
constexpr int S_CV = S_TK + NL;
constexpr int S_FN = S_CV + NL;
constexpr int S_FN1 = S_FN + NL;
constexpr int S_SABG = S_FN1 + NL;
constexpr int S_RV = S_SABG + NL;
constexpr int S_TV = S_RV + NL;
constexpr int S_BM = S_TV + NL;             // 5 bands
constexpr int S_AB = S_BM + 5 * NL;         // 7 x NL band storage
constexpr int S_IPIV = S_AB + 7 * NL;
constexpr int S_SC = S_IPIV + NL;           // column scalars
enum { SC_HS_SOIL, SC_HS_H2OSFC, SC_DHSDT, SC_HS_TOP_SNOW, SC_TK_H2OSFC, SC_C_H2OSFC, SC_DZ_H2OSFC, SC_FN_H2OSFC, SC_N };
constexpr int S_TOTAL = S_SC + SC_N;

struct PW {
    double *w;
    long ld;
    __device__ __forceinline__ double &at(int lane, int slot) const { return w[(long)slot * ld + lane]; }
    __device__ __forceinline__ double &lev(int lane, int slot, int j, int nlevsno) const {
        return w[(long)(slot + j + nlevsno) * ld + lane];
    }
};

#define PLAIN_PROLOGUE                                              \
    const int i = blockIdx.x * blockDim.x + threadIdx.x;           \
    if (i >= n) return;                                              \
    const int c = filter ? filter[i] : V.begc + i;                   \
    if (!col_in_range(V, c)) return;                                 \
    const int snl = S.snl[V.c1(c)];                                  \
    const int NS = V.nlevsno, NG = V.nlevgrnd;                       \
    (void)snl; (void)NS; (void)NG;

__global__ void k_thermprop(View V, Params P, ctsm_col_state S, int n, const int32_t *filter, PW w)
{
    PLAIN_PROLOGUE
    const int nbedrock = S.nbedrock[V.c1(c)];
    const double frac_sno = S.frac_sno_eff[V.c1(c)];
    const double h2osno_no_layers = S.h2osno_no_layers[V.c1(c)];
    for (int j = snl + 1; j <= NG; ++j) S.thk[V.sno(c, j)] = level_thk(V, P, S, c, j, nbedrock, frac_sno);
    for (int j = snl + 1; j <= NG; ++j) {
        double tk;
        if (j <= NG - 1) {
            const double a = S.thk[V.sno(c, j)], b = S.thk[V.sno(c, j + 1)];
            const double zj = S.z[V.sno(c, j)], zn = S.z[V.sno(c, j + 1)], zij = S.zi[V.zi(c, j)];
            tk = a * b * (zn - zj) / (a * (zn - zij) + b * (zij - zj));
        } else {
            tk = 0.0;
        }
        w.lev(i, S_TK, j, NS) = tk;
        w.lev(i, S_CV, j, NS) = level_cv(V, S, c, j, nbedrock, frac_sno, h2osno_no_layers);
    }
    const double zh2osfc = (double)1.0e-3f * (0.5 * S.h2osfc[V.c1(c)]);
    const double th1 = S.thk[V.sno(c, 1)], z1 = S.z[V.sno(c, 1)];
    w.at(i, S_SC + SC_TK_H2OSFC) = TKWAT * th1 * (z1 + zh2osfc) / (TKWAT * z1 + th1 * zh2osfc);
}

__global__ void k_groundflux(View V, ctsm_col_state S, ctsm_patch_state Pt, int n, const int32_t *filter, PW w)
{
    PLAIN_PROLOGUE
    const double frac_sno = S.frac_sno_eff[V.c1(c)];
    const double emg = S.emg[V.c1(c)], forc_lwrad = S.forc_lwrad[V.c1(c)], htvp = S.htvp[V.c1(c)];
    const int ltop = snl + 1;
    const double tg = S.t_grnd[V.c1(c)], ts = S.t_soisno[V.sno(c, ltop)];
    const double t1 = S.t_soisno[V.sno(c, 1)], th = S.t_h2osfc[V.c1(c)];
    const double tg2 = tg * tg, ts2 = ts * ts, t12 = t1 * t1, th2 = th * th;
    const double lwrad_emit = emg * SB * (tg2 * tg2);
    const double dlwrad_emit = 4.0 * emg * SB * (tg2 * tg);
    const double lwrad_emit_snow = emg * SB * (ts2 * ts2);
    const double lwrad_emit_soil = emg * SB * (t12 * t12);
    const double lwrad_emit_h2osfc = emg * SB * (th2 * th2);
    double hs_soil = 0.0, hs_h2osfc = 0.0, dhsdT = 0.0, hs_top_snow = 0.0;
    int pi = S.patchi[V.c1(c)], pf = S.patchf[V.c1(c)];
    if (pi < V.begp || pf >= V.begp + V.ldp) { pi = 1; pf = 0; }
    for (int p = pi; p <= pf; ++p) {
        const double fveg = (double)Pt.frac_veg_nosno[V.p1(p)], wt = Pt.wtcol[V.p1(p)];
        const double lwin = (1.0 - fveg) * emg * forc_lwrad;
        const double dl = Pt.dlrad[V.p1(p)];
        Pt.eflx_gnet[V.p1(p)] = Pt.sabg[V.p1(p)] + dl + lwin - lwrad_emit -
                                (Pt.eflx_sh_grnd[V.p1(p)] + Pt.qflx_evap_soi[V.p1(p)] * htvp);
        Pt.sabg_chk[V.p1(p)] = frac_sno * Pt.sabg_snow[V.p1(p)] + (1.0 - frac_sno) * Pt.sabg_soil[V.p1(p)];
        const double egsoil = Pt.sabg_soil[V.p1(p)] + dl + lwin - lwrad_emit_soil -
                              (Pt.eflx_sh_soil[V.p1(p)] + Pt.qflx_ev_soil[V.p1(p)] * htvp);
        const double egh2o = Pt.sabg_soil[V.p1(p)] + dl + lwin - lwrad_emit_h2osfc -
                             (Pt.eflx_sh_h2osfc[V.p1(p)] + Pt.qflx_ev_h2osfc[V.p1(p)] * htvp);
        const double dg = -Pt.cgrnd[V.p1(p)] - dlwrad_emit;
        Pt.dgnetdT[V.p1(p)] = dg;
        dhsdT = dhsdT + dg * wt;
        hs_soil = hs_soil + egsoil * wt;
        hs_h2osfc = hs_h2osfc + egh2o * wt;
    }
    for (int j = -NS + 1; j <= 1; ++j) w.lev(i, S_SABG, j, NS) = 0.0;
    for (int p = pi; p <= pf; ++p) {
        const double fveg = (double)Pt.frac_veg_nosno[V.p1(p)], wt = Pt.wtcol[V.p1(p)];
        const double lwin = (1.0 - fveg) * emg * forc_lwrad;
        const double sl = Pt.sabg_lyr[V.lyr(p, ltop)];
        const double egsnow = sl + Pt.dlrad[V.p1(p)] + lwin - lwrad_emit_snow -
                              (Pt.eflx_sh_snow[V.p1(p)] + Pt.qflx_ev_snow[V.p1(p)] * htvp);
        hs_top_snow = hs_top_snow + egsnow * wt;
        for (int j = ltop; j <= 1; ++j)
            w.lev(i, S_SABG, j, NS) = w.lev(i, S_SABG, j, NS) + Pt.sabg_lyr[V.lyr(p, j)] * wt;
    }
    w.at(i, S_SC + SC_HS_SOIL) = hs_soil;
    w.at(i, S_SC + SC_HS_H2OSFC) = hs_h2osfc;
    w.at(i, S_SC + SC_DHSDT) = dhsdT;
    w.at(i, S_SC + SC_HS_TOP_SNOW) = hs_top_snow;
}

__global__ void k_diffflux(View V, ctsm_col_state S, int n, const int32_t *filter, double dtime, PW w)
{
    PLAIN_PROLOGUE
    for (int j = snl + 1; j <= NG; ++j) {
        const double cv = w.lev(i, S_CV, j, NS), tk = w.lev(i, S_TK, j, NS);
        double fact, fn;
        if (j == snl + 1) {
            const double zj = S.z[V.sno(c, j)], zn = S.z[V.sno(c, j + 1)], zim = S.zi[V.zi(c, j - 1)];
            fact = dtime / cv * S.dz[V.sno(c, j)] / (0.5 * (zj - zim + CAPR * (zn - zim)));
            fn = tk * (S.t_soisno[V.sno(c, j + 1)] - S.t_soisno[V.sno(c, j)]) / (zn - zj);
        } else if (j <= NG - 1) {
            fact = dtime / cv;
            fn = tk * (S.t_soisno[V.sno(c, j + 1)] - S.t_soisno[V.sno(c, j)]) / (S.z[V.sno(c, j + 1)] - S.z[V.sno(c, j)]);
        } else {
            fact = dtime / cv;
            fn = S.eflx_bot[V.c1(c)];
        }
        S.fact[V.sno(c, j)] = fact;
        w.lev(i, S_FN, j, NS) = fn;
    }
    const double h2osfc = S.h2osfc[V.c1(c)], fh = S.frac_h2osfc[V.c1(c)];
    double ch, dzh;
    if (h2osfc > THIN_SFCLAYER && fh > THIN_SFCLAYER) {
        ch = fmax90(THIN_SFCLAYER, CPLIQ * h2osfc / fh);
        dzh = fmax90(THIN_SFCLAYER, (double)1.0e-3f * h2osfc / fh);
    } else {
        ch = THIN_SFCLAYER;
        dzh = THIN_SFCLAYER;
    }
    S.c_h2osfc[V.c1(c)] = ch;
    w.at(i, S_SC + SC_C_H2OSFC) = ch;
    w.at(i, S_SC + SC_DZ_H2OSFC) = dzh;
}

__global__ void k_rhs(View V, ctsm_col_state S, int n, const int32_t *filter, double dtime, PW w)
{
    PLAIN_PROLOGUE
    const double frac_sno = S.frac_sno_eff[V.c1(c)], frac_h2osfc = S.frac_h2osfc[V.c1(c)];
    const double dhsdT = w.at(i, S_SC + SC_DHSDT), hs_top_snow = w.at(i, S_SC + SC_HS_TOP_SNOW);
    const double hs_soil = w.at(i, S_SC + SC_HS_SOIL), hs_h2osfc = w.at(i, S_SC + SC_HS_H2OSFC);
#define T(jj) S.t_soisno[V.sno(c, jj)]
#define FACT(jj) S.fact[V.sno(c, jj)]
#define FN(jj) w.lev(i, S_FN, jj, NS)
    for (int j = snl + 1; j <= 0; ++j) {
        double r;
        if (j == snl + 1) r = T(j) + FACT(j) * (hs_top_snow - dhsdT * T(j) + CNFAC * FN(j));
        else {
            r = T(j) + CNFAC * FACT(j) * (FN(j) - FN(j - 1));
            r = r + FACT(j) * w.lev(i, S_SABG, j, NS);
        }
        w.lev(i, S_RV, j - 1, NS) = r;
    }
    const double dzm = (0.5 * w.at(i, S_SC + SC_DZ_H2OSFC) + S.z[V.sno(c, 1)]);
    const double th = S.t_h2osfc[V.c1(c)];
    const double fn_h2osfc = w.at(i, S_SC + SC_TK_H2OSFC) * (T(1) - th) / dzm;
    w.at(i, S_SC + SC_FN_H2OSFC) = fn_h2osfc;
    w.lev(i, S_RV, 0, NS) = th + (dtime / w.at(i, S_SC + SC_C_H2OSFC)) * (hs_h2osfc - dhsdT * th + CNFAC * fn_h2osfc);
    for (int j = 1; j <= NG; ++j) {
        double r;
        if (j == snl + 1) r = T(j) + FACT(j) * (hs_top_snow - dhsdT * T(j) + CNFAC * FN(j));
        else if (j == 1) {
            r = T(j) + FACT(j) * ((1.0 - frac_sno) * (hs_soil - dhsdT * T(j)) + CNFAC * (FN(j) - frac_sno * FN(j - 1)));
            r = r + frac_sno * FACT(j) * w.lev(i, S_SABG, j, NS);
        } else if (j <= NG - 1) r = T(j) + CNFAC * FACT(j) * (FN(j) - FN(j - 1));
        else r = T(j) - CNFAC * FACT(j) * FN(j - 1) + FACT(j) * FN(j);
        w.lev(i, S_RV, j, NS) = r;
    }
    if (frac_h2osfc != 0.0)
        w.lev(i, S_RV, 1, NS) = w.lev(i, S_RV, 1, NS) -
                                frac_h2osfc * FACT(1) * ((hs_soil - dhsdT * T(1)) + CNFAC * fn_h2osfc);
}

__global__ void k_matrix(View V, ctsm_col_state S, int n, const int32_t *filter, double dtime, PW w)
{
    PLAIN_PROLOGUE
    const double frac_sno = S.frac_sno_eff[V.c1(c)], frac_h2osfc = S.frac_h2osfc[V.c1(c)];
    const double dhsdT = w.at(i, S_SC + SC_DHSDT);
    const double tkh = w.at(i, S_SC + SC_TK_H2OSFC), ch = w.at(i, S_SC + SC_C_H2OSFC);
#define BM(b, jj) w.lev(i, S_BM + (b - 1) * NL, jj, NS)
#define TK(jj) w.lev(i, S_TK, jj, NS)
#define Z(jj) S.z[V.sno(c, jj)]
    for (int j = -NS; j <= NG; ++j)
        for (int b = 1; b <= 5; ++b) BM(b, j) = 0.0;
    for (int j = snl + 1; j <= 0; ++j) {
        const double dzp = Z(j + 1) - Z(j);
        if (j == snl + 1) {
            BM(4, j - 1) = 0.0;
            BM(3, j - 1) = 1.0 + (1.0 - CNFAC) * FACT(j) * TK(j) / dzp - FACT(j) * dhsdT;
        } else {
            const double dzm = (Z(j) - Z(j - 1));
            BM(4, j - 1) = -((1.0 - CNFAC) * FACT(j) * TK(j - 1) / dzm);
            BM(3, j - 1) = 1.0 + (1.0 - CNFAC) * FACT(j) * (TK(j) / dzp + TK(j - 1) / dzm);
        }
        if (j != 0) BM(2, j - 1) = -((1.0 - CNFAC) * FACT(j) * TK(j) / dzp);
        else BM(1, j - 1) = -((1.0 - CNFAC) * FACT(j) * TK(j) / dzp);
    }
    for (int j = 1; j <= NG; ++j) {
        if (j == snl + 1) {
            const double dzp = Z(j + 1) - Z(j);
            BM(3, j) = 1.0 + (1.0 - CNFAC) * FACT(j) * TK(j) / dzp - FACT(j) * dhsdT;
            BM(2, j) = -((1.0 - CNFAC) * FACT(j) * TK(j) / dzp);
        } else if (j == 1) {
            const double dzm = (Z(j) - Z(j - 1)), dzp = (Z(j + 1) - Z(j));
            BM(2, j) = -((1.0 - CNFAC) * FACT(j) * TK(j) / dzp);
            BM(3, j) = 1.0 + (1.0 - CNFAC) * FACT(j) * (TK(j) / dzp + frac_sno * TK(j - 1) / dzm) -
                       (1.0 - frac_sno) * FACT(j) * dhsdT;
            BM(5, j) = -(frac_sno * (1.0 - CNFAC) * FACT(j) * TK(j - 1) / dzm);
        } else if (j <= NG - 1) {
            const double dzm = (Z(j) - Z(j - 1)), dzp = (Z(j + 1) - Z(j));
            BM(2, j) = -((1.0 - CNFAC) * FACT(j) * TK(j) / dzp);
            BM(3, j) = 1.0 + (1.0 - CNFAC) * FACT(j) * (TK(j) / dzp + TK(j - 1) / dzm);
            BM(4, j) = -((1.0 - CNFAC) * FACT(j) * TK(j - 1) / dzm);
        } else {
            const double dzm = (Z(j) - Z(j - 1));
            BM(2, j) = 0.0;
            BM(3, j) = 1.0 + (1.0 - CNFAC) * FACT(j) * TK(j - 1) / dzm;
            BM(4, j) = -((1.0 - CNFAC) * FACT(j) * TK(j - 1) / dzm);
        }
    }
    const double dzm = (0.5 * w.at(i, S_SC + SC_DZ_H2OSFC) + Z(1));
    if (frac_h2osfc != 0.0)
        BM(3, 1) = BM(3, 1) + frac_h2osfc * ((1.0 - CNFAC) * FACT(1) * tkh / dzm + FACT(1) * dhsdT);
    BM(3, 0) = 1.0 + (1.0 - CNFAC) * (dtime / ch) * tkh / dzm - (dtime / ch) * dhsdT;
    BM(2, 0) = -((1.0 - CNFAC) * (dtime / ch) * tkh / dzm);
    BM(4, 1) = (frac_h2osfc != 0.0) ? -(frac_h2osfc * (1.0 - CNFAC) * FACT(1) * tkh / dzm) : 0.0;
    if (snl == 0) BM(5, 1) = 0.0;
}

// BandDiagonal (BandDiagonalMod.F90:29-221): pack, LAPACK dgbsv (dgbtf2 + dgbtrs), unpack.
__global__ void k_banddiag(View V, ctsm_col_state S, int n, const int32_t *filter, PW w, ErrWord *err)
{
    PLAIN_PROLOGUE
    const int jtop = snl, jbot = NG, kl = 2, ku = 2, kv = kl + ku, m = 2 * kl + ku + 1;
    const int nn = jbot - jtop + 1;
#define AB(r, cc) w.at(i, S_AB + ((r) - 1) + ((cc) - 1) * m)
#define RES(k) w.lev(i, S_RV, jtop + (k) - 1, NS)   /* 1-based, solved in place */
    for (int k = 0; k < m * nn; ++k) w.at(i, S_AB + k) = 0.0;
    for (int k = 3; k <= nn; ++k) AB(kl + ku - 1, k) = BM(1, jtop + k - 3);
    for (int k = 2; k <= nn; ++k) AB(kl + ku + 0, k) = BM(2, jtop + k - 2);
    for (int k = 1; k <= nn; ++k) AB(kl + ku + 1, k) = BM(3, jtop + k - 1);
    for (int k = 1; k <= nn - 1; ++k) AB(kl + ku + 2, k) = BM(4, jtop + k);
    for (int k = 1; k <= nn - 2; ++k) AB(kl + ku + 3, k) = BM(5, jtop + k + 1);
    // dgbtf2
    int info = 0, ju = 1;
    for (int j = 1; j <= nn; ++j) {
        if (j + kv <= nn)
            for (int r = 1; r <= kl; ++r) AB(r, j + kv) = 0.0;
        const int km = kl < nn - j ? kl : nn - j;
        int jp = 1;
        double dmx = fabs(AB(kv + 1, j));
        for (int r = 2; r <= km + 1; ++r)
            if (fabs(AB(kv + r, j)) > dmx) { dmx = fabs(AB(kv + r, j)); jp = r; }
        w.at(i, S_IPIV + j - 1) = (double)(jp + j - 1);
        if (AB(kv + jp, j) != 0.0) {
            int jl = j + ku + jp - 1;
            if (jl > nn) jl = nn;
            if (jl > ju) ju = jl;
            if (jp != 1)
                for (int k = 0; k <= ju - j; ++k) {
                    const double t = AB(kv + jp - k, j + k);
                    AB(kv + jp - k, j + k) = AB(kv + 1 - k, j + k);
                    AB(kv + 1 - k, j + k) = t;
                }
            if (km > 0) {
                const double r = 1.0 / AB(kv + 1, j);
                for (int q = 1; q <= km; ++q) AB(kv + 1 + q, j) = AB(kv + 1 + q, j) * r;
                for (int k = 1; k <= ju - j; ++k) {
                    const double y = AB(kv + 1 - k, j + k);
                    if (y != 0.0) {
                        const double temp = -1.0 * y;
                        for (int q = 1; q <= km; ++q)
                            AB(kv + 1 + q - k, j + k) = AB(kv + 1 + q - k, j + k) + AB(kv + 1 + q, j) * temp;
                    }
                }
            }
        } else if (info == 0) {
            info = j;
        }
    }
    if (info != 0) { report_error(err, c, CTSM_E_SOLVER); return; }
    // dgbtrs
    for (int j = 1; j <= nn - 1; ++j) {
        const int lm = kl < nn - j ? kl : nn - j;
        const int l = (int)w.at(i, S_IPIV + j - 1);
        if (l != j) { const double t = RES(l); RES(l) = RES(j); RES(j) = t; }
        if (RES(j) != 0.0) {
            const double temp = -1.0 * RES(j);
            for (int q = 1; q <= lm; ++q) RES(j + q) = RES(j + q) + AB(kv + 1 + q, j) * temp;
        }
    }
    for (int j = nn; j >= 1; --j) {
        if (RES(j) != 0.0) {
            const int lo = j - kv > 1 ? j - kv : 1;
            RES(j) = RES(j) / AB(kv + 1, j);
            const double temp = RES(j);
            for (int q = j - 1; q >= lo; --q) RES(q) = RES(q) - temp * AB(kv + 1 + q - j, j);
        }
    }
    for (int j = snl + 1; j <= 0; ++j) T(j) = w.lev(i, S_RV, j - 1, NS);
    for (int j = 1; j <= NG; ++j) T(j) = w.lev(i, S_RV, j, NS);
    if (S.frac_h2osfc[V.c1(c)] == 0.0) S.t_h2osfc[V.c1(c)] = T(1);
    else S.t_h2osfc[V.c1(c)] = w.lev(i, S_RV, 0, NS);
#undef AB
#undef RES
}

__global__ void k_fn1_h2osfc(View V, ctsm_col_state S, int n, const int32_t *filter, double dtime, PW w)
{
    PLAIN_PROLOGUE
    for (int j = snl + 1; j <= NG; ++j)
        w.lev(i, S_FN1, j, NS) = (j <= NG - 1) ? TK(j) * (T(j + 1) - T(j)) / (Z(j + 1) - Z(j)) : 0.0;
    // PhaseChangeH2osfc
    const double frac_h2osfc = S.frac_h2osfc[V.c1(c)], frac_sno = S.frac_sno_eff[V.c1(c)];
    const double dhsdT = w.at(i, S_SC + SC_DHSDT), c_h2osfc = w.at(i, S_SC + SC_C_H2OSFC);
    S.xmf_h2osfc[V.c1(c)] = 0.0;
    S.qflx_h2osfc_to_ice[V.c1(c)] = 0.0;
    S.eflx_h2osfc_to_snow[V.c1(c)] = 0.0;
    double h2osno_total = S.h2osno_no_layers[V.c1(c)];
    for (int j = snl + 1; j <= 0; ++j) h2osno_total = h2osno_total + S.h2osoi_ice[V.sno(c, j)] + S.h2osoi_liq[V.sno(c, j)];
    double t_h2osfc = S.t_h2osfc[V.c1(c)];
    if (frac_h2osfc > 0.0 && t_h2osfc <= TFRZ) {
        double h2osfc = S.h2osfc[V.c1(c)], snow_depth = S.snow_depth[V.c1(c)];
        const double tinc = TFRZ - t_h2osfc;
        t_h2osfc = TFRZ;
        const double hm = frac_h2osfc * (dhsdT * tinc - tinc * c_h2osfc / dtime);
        const double xm = hm * dtime / HFUS;
        const double temp1 = h2osfc + xm;
        const double z_avg = frac_sno * snow_depth;
        double rho_avg = (z_avg > 0.0) ? fmin90(800.0, h2osno_total / z_avg) : 200.0;
        double c1, c2;
        if (temp1 >= 0.0) {
            S.int_snow[V.c1(c)] = S.int_snow[V.c1(c)] - xm;
            if (snl == 0) S.h2osno_no_layers[V.c1(c)] = S.h2osno_no_layers[V.c1(c)] - xm;
            else S.h2osoi_ice[V.sno(c, 0)] = S.h2osoi_ice[V.sno(c, 0)] - xm;
            h2osno_total = h2osno_total - xm;
            h2osfc = h2osfc + xm;
            S.xmf_h2osfc[V.c1(c)] = hm;
            S.qflx_h2osfc_to_ice[V.c1(c)] = -xm / dtime;
            if (frac_sno > 0 && snl < 0) snow_depth = h2osno_total / (rho_avg * frac_sno);
            else snow_depth = h2osno_total / DENICE;
            if (snl == 0) {
                T(0) = t_h2osfc;
                S.eflx_h2osfc_to_snow[V.c1(c)] = 0.0;
            } else {
                if (snl == -1) c1 = frac_sno * (dtime / FACT(0) - dhsdT * dtime);
                else c1 = frac_sno / FACT(0) * dtime;
                if (frac_h2osfc != 0.0) c2 = (-CPLIQ * xm - frac_h2osfc * dhsdT * dtime);
                else c2 = 0.0;
                T(0) = (c1 * T(0) + c2 * t_h2osfc) / (c1 + c2);
                S.eflx_h2osfc_to_snow[V.c1(c)] = (t_h2osfc - T(0)) * c2 / dtime;
            }
        } else {
            rho_avg = (h2osno_total * rho_avg + h2osfc * DENICE) / (h2osno_total + h2osfc);
            S.int_snow[V.c1(c)] = S.int_snow[V.c1(c)] + h2osfc;
            if (snl == 0) S.h2osno_no_layers[V.c1(c)] = S.h2osno_no_layers[V.c1(c)] + h2osfc;
            else S.h2osoi_ice[V.sno(c, 0)] = S.h2osoi_ice[V.sno(c, 0)] + h2osfc;
            h2osno_total = h2osno_total + h2osfc;
            S.qflx_h2osfc_to_ice[V.c1(c)] = h2osfc / dtime;
            t_h2osfc = t_h2osfc - temp1 * HFUS / (dtime * dhsdT - c_h2osfc);
            S.xmf_h2osfc[V.c1(c)] = (hm - frac_h2osfc * temp1 * HFUS / dtime);
            if (snl == 0) {
                T(0) = t_h2osfc;
            } else {
                if (snl == -1) c1 = frac_sno * (dtime / FACT(0) - dhsdT * dtime);
                else c1 = frac_sno / FACT(0) * dtime;
                if (frac_h2osfc != 0.0) c2 = frac_h2osfc * (c_h2osfc - dtime * dhsdT);
                else c2 = 0.0;
                T(0) = (c1 * T(0) + c2 * t_h2osfc) / (c1 + c2);
                t_h2osfc = T(0);
            }
            h2osfc = 0.0;
            if (frac_sno > 0 && snl < 0) snow_depth = h2osno_total / (rho_avg * frac_sno);
            else snow_depth = h2osno_total / DENICE;
        }
        S.h2osfc[V.c1(c)] = h2osfc;
        S.snow_depth[V.c1(c)] = snow_depth;
        S.t_h2osfc[V.c1(c)] = t_h2osfc;
    }
}

__global__ void k_phasechange(View V, ctsm_col_state S, int n, const int32_t *filter, double dtime, PW w)
{
    PLAIN_PROLOGUE
    const double frac_sno = S.frac_sno_eff[V.c1(c)], frac_h2osfc = S.frac_h2osfc[V.c1(c)];
    const double dhsdT = w.at(i, S_SC + SC_DHSDT);
    double h2osno_no_layers = S.h2osno_no_layers[V.c1(c)], snow_depth = S.snow_depth[V.c1(c)];
    double xmf = 0.0, qflx_snomelt = 0.0, qflx_snofrz = 0.0, qflx_snow_drain = 0.0;
    double snomelt_accum = S.snomelt_accum ? S.snomelt_accum[V.c1(c)] : 0.0;
    const int ltop = snl + 1;
    if (S.qflx_snomelt_lyr)
        for (int j = -NS + 1; j <= 0; ++j) { S.qflx_snomelt_lyr[V.sno(c, j)] = 0.0; S.qflx_snofrz_lyr[V.sno(c, j)] = 0.0; }
    for (int j = ltop; j <= NG; ++j) {
        double t = T(j), liq = S.h2osoi_liq[V.sno(c, j)], ice = S.h2osoi_ice[V.sno(c, j)];
        const double wice0 = ice, wmass0 = ice + liq + 0.0, fact_j = FACT(j);
        int imelt = 0;
        double tinc = 0.0, supercool = 0.0, hm = 0.0, xm = 0.0;
        if (j <= 0) {
            if (ice > 0.0 && t > TFRZ) { imelt = 1; tinc = TFRZ - t; t = TFRZ; }
            if (liq > 0.0 && t < TFRZ) { imelt = 2; tinc = TFRZ - t; t = TFRZ; }
        } else {
            if (ice > 0.0 && t > TFRZ) { imelt = 1; tinc = TFRZ - t; t = TFRZ; }
            if (t < TFRZ) {
                const double smp = HFUS * (TFRZ - t) / (GRAV * t) * 1000.0;
                supercool = S.watsat[V.soi(c, j)] * cm_pow(smp / S.sucsat[V.soi(c, j)], -1.0 / S.bsw[V.soi(c, j)]);
                supercool = supercool * S.dz[V.sno(c, j)] * 1000.0;
            }
            if (liq > supercool && t < TFRZ) { imelt = 2; tinc = TFRZ - t; t = TFRZ; }
            if (h2osno_no_layers > 0.0 && j == 1 && t > TFRZ) { imelt = 1; tinc = TFRZ - t; t = TFRZ; }
        }
        if (imelt > 0) {
            if (j == ltop) {
                if (j > 0) hm = dhsdT * tinc - tinc / fact_j;
                else hm = frac_sno * (dhsdT * tinc - tinc / fact_j);
                if (j == 1 && frac_h2osfc != 0.0) hm = hm - frac_h2osfc * (dhsdT * tinc);
            } else if (j == 1) {
                hm = (1.0 - frac_sno - frac_h2osfc) * dhsdT * tinc - tinc / fact_j;
            } else {
                if (j < 1) hm = -(frac_sno * (tinc / fact_j));
                else hm = -(tinc / fact_j);
            }
        }
        if (imelt == 1 && hm < 0.0) { hm = 0.0; imelt = 0; }
        if (imelt == 2 && hm > 0.0) { hm = 0.0; imelt = 0; }
        if (imelt > 0 && fabs(hm) > 0.0) {
            double heatr;
            xm = hm * dtime / HFUS;
            if (j == 1 && h2osno_no_layers > 0.0 && xm > 0.0) {
                const double temp1 = h2osno_no_layers;
                h2osno_no_layers = fmax90(0.0, temp1 - xm);
                const double propor = h2osno_no_layers / temp1;
                snow_depth = propor * snow_depth;
                heatr = hm - HFUS * (temp1 - h2osno_no_layers) / dtime;
                if (heatr > 0.0) { xm = heatr * dtime / HFUS; hm = heatr; }
                else { xm = 0.0; hm = 0.0; }
                qflx_snomelt = fmax90(0.0, (temp1 - h2osno_no_layers)) / dtime;
                xmf = HFUS * qflx_snomelt;
                qflx_snow_drain = qflx_snomelt;
            }
            heatr = 0.0;
            if (xm > 0.0) {
                ice = fmax90(0.0, wice0 - xm);
                if (j >= 1) heatr = hm - HFUS * (0.0 - 0.0 + wice0 - ice) / dtime;
                else heatr = hm - HFUS * (wice0 - ice) / dtime;
            } else if (xm < 0.0) {
                if (j <= 0) ice = fmin90(wmass0, wice0 - xm);
                else if (wmass0 - 0.0 < supercool) ice = 0.0;
                else ice = fmin90(wmass0 - 0.0 - supercool, wice0 - xm);
                heatr = hm - HFUS * (wice0 - ice) / dtime;
            }
            liq = fmax90(0.0, wmass0 - ice - 0.0);
            if (fabs(heatr) > 0.0) {
                if (j == ltop) {
                    if (j == 1) t = t + fact_j * heatr / (1.0 - (1.0 - frac_h2osfc) * fact_j * dhsdT);
                    else t = t + (fact_j / frac_sno) * heatr / (1.0 - fact_j * dhsdT);
                } else if (j == 1) {
                    t = t + fact_j * heatr / (1.0 - (1.0 - frac_sno - frac_h2osfc) * fact_j * dhsdT);
                } else {
                    if (j > 0) t = t + fact_j * heatr;
                    else if (frac_sno > 0.0) t = t + (fact_j / frac_sno) * heatr;
                }
                if (j <= 0 && liq * ice > 0.0) t = TFRZ;
            }
            if (j >= 1) xmf = xmf + HFUS * (wice0 - ice) / dtime + HFUS * (0.0 - 0.0) / dtime;
            else xmf = xmf + HFUS * (wice0 - ice) / dtime;
            if (imelt == 1 && j < 1) {
                const double ml = fmax90(0.0, (wice0 - ice)) / dtime;
                if (S.qflx_snomelt_lyr) S.qflx_snomelt_lyr[V.sno(c, j)] = ml;
                qflx_snomelt = qflx_snomelt + ml;
                snomelt_accum = snomelt_accum + ml * dtime * 1.e-3;
            }
            if (imelt == 2 && j < 1) {
                const double fz = fmax90(0.0, (ice - wice0)) / dtime;
                if (S.qflx_snofrz_lyr) S.qflx_snofrz_lyr[V.sno(c, j)] = fz;
                qflx_snofrz = qflx_snofrz + fz;
            }
            S.h2osoi_ice[V.sno(c, j)] = ice;
            S.h2osoi_liq[V.sno(c, j)] = liq;
        }
        T(j) = t;
        S.imelt[V.sno(c, j)] = imelt;
    }
    S.xmf[V.c1(c)] = xmf;
    S.qflx_snomelt[V.c1(c)] = qflx_snomelt;
    S.qflx_snofrz[V.c1(c)] = qflx_snofrz;
    S.qflx_snow_drain[V.c1(c)] = qflx_snow_drain;
    S.eflx_snomelt[V.c1(c)] = qflx_snomelt * HFUS;
    if (S.snomelt_accum) S.snomelt_accum[V.c1(c)] = snomelt_accum;
    S.h2osno_no_layers[V.c1(c)] = h2osno_no_layers;
    S.snow_depth[V.c1(c)] = snow_depth;
}

__global__ void k_tail(View V, ctsm_col_state S, int n, const int32_t *filter, PW w)
{
    PLAIN_PROLOGUE
    const double frac_sno = S.frac_sno_eff[V.c1(c)], fh = S.frac_h2osfc[V.c1(c)];
    const double t1 = T(1), th = S.t_h2osfc[V.c1(c)];
    double tg;
    if (snl < 0) {
        if (fh != 0.0) tg = frac_sno * T(snl + 1) + (1.0 - frac_sno - fh) * t1 + fh * th;
        else tg = frac_sno * T(snl + 1) + (1.0 - frac_sno) * t1;
    } else {
        if (fh != 0.0) tg = (1.0 - fh) * t1 + fh * th;
        else tg = t1;
    }
    S.t_grnd[V.c1(c)] = tg;
    S.eflx_fgr12[V.c1(c)] = -(CNFAC * FN(1)) - (1.0 - CNFAC) * w.lev(i, S_FN1, 1, NS);
    for (int j = 1; j <= NG; ++j)
        S.eflx_fgr[V.soi(c, j)] = (j < NG) ? -(CNFAC * FN(j)) - (1.0 - CNFAC) * w.lev(i, S_FN1, j, NS) : 0.0;
}
#undef T
#undef FACT
#undef FN
#undef BM
#undef TK
#undef Z

// ---- Richards, one kernel per call with per-layer arrays in HBM ----------
constexpr int R_HK = 0, R_SMP = 1, R_DHKDW = 2, R_DSMPDW = 3, R_IMPED = 4, R_VWC = 5, R_DTDZ = 6, R_QIN = 7, R_QOUT = 8,
              R_DQIDW0 = 9, R_DQIDW1 = 10, R_DQODW1 = 11, R_DQODW2 = 12, R_AMX = 13, R_BMX = 14, R_CMX = 15, R_RMX = 16,
              R_DWAT = 17, R_DL = 18, R_D = 19, R_DU = 20, R_NF = 21;

__global__ void k_soilwater_plain(View V, Params P, ctsm_col_state S, int n, const int32_t *filter, double dtime,
                                  PW w, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (!col_in_range(V, c)) { report_error(err, c, CTSM_E_ARG); return; }
    const int nl = S.nbedrock[V.c1(c)];
    if (nl < 2 || nl > MAX_NLEVSOI) { report_error(err, c, CTSM_E_SOLVER); return; }
    const bool wt_bottom = (P.lower_bc == CTSM_BC_WATERTABLE);
    if (wt_bottom && (nl != V.nlevsoi || !S.zwt || !S.smpmin)) { report_error(err, c, CTSM_E_ARG); return; }
    const int jwt = wt_bottom ? watertable_layer(V, S, c) : 0;
#define A(f, j) w.at(i, (f) * MAX_NLEVSOI + (j) - 1)
#define LIQ(j) S.h2osoi_liq[V.sno(c, j)]
#define DZ(j) S.dz[V.sno(c, j)]
#define WS(j) S.watsat[V.soi(c, j)]
    int nsub = 0;
    double dtsub = dtime, dtdone = 0.0, qcharge = 0.0;
    for (;;) {
        nsub = nsub + 1;
        if (nsub > MAX_TRIALS) { report_error(err, c, CTSM_E_SOLVER); return; }
        for (int j = 1; j <= nl; ++j) {
            A(R_VWC, j) = fmax90(LIQ(j), 1.0e-6) / (DZ(j) * DENH2O);
            A(R_DTDZ, j) = dtsub / (M_TO_MM * DZ(j));
        }
        // compute_hydraulic_properties
        for (int j = 1; j <= nl; ++j) {
            double s2 = fmin90(A(R_VWC, j) / WS(j), 1.0);
            A(R_SMP, j) = fmax90(0.01, s2);     // s2 kept in the smp slot until used
        }
        for (int j = 1; j <= nl; ++j) {
            double s1;
            const double bsw = S.bsw[V.soi(c, j)];
            if (j == nl) {
                s1 = A(R_SMP, j);
                A(R_IMPED, j) = cm_pow(10.0, -(P.e_ice * S.icefrac[V.soi(c, j)]));
            } else {
                s1 = 0.5 * (A(R_SMP, j) + A(R_SMP, j + 1));
                A(R_IMPED, j) = cm_pow(10.0, -(P.e_ice * (0.5 * (S.icefrac[V.soi(c, j)] + S.icefrac[V.soi(c, j + 1)]))));
            }
            s1 = fmin90(s1, 1.0);
            s1 = fmax90(0.01, s1);
            const double hk = A(R_IMPED, j) * S.hksat[V.soi(c, j)] * cm_pow(s1, 2.0 * bsw + 3.0);
            A(R_HK, j) = hk;
            A(R_DHKDW, j) = (2.0 * bsw + 3.0) * hk / s1;
        }
        for (int j = 1; j <= nl; ++j) {
            const double s2 = A(R_SMP, j), bsw = S.bsw[V.soi(c, j)];
            const double smp = -(S.sucsat[V.soi(c, j)] * cm_pow(s2, -bsw));
            A(R_DSMPDW, j) = (-(bsw * smp) / s2) / WS(j);
            A(R_SMP, j) = smp;
            S.smp_l[V.soi(c, j)] = smp;
            S.hk_l[V.soi(c, j)] = A(R_HK, j);
        }
        // compute_moisture_fluxes_and_derivs
        A(R_QIN, 1) = S.qflx_infl[V.c1(c)];
        A(R_DQIDW0, 1) = 0.0;
        A(R_DQIDW1, 1) = 0.0;
        for (int j = 1; j <= nl - 1; ++j) {
            if (j >= 2) {
                A(R_QIN, j) = A(R_QOUT, j - 1);
                A(R_DQIDW0, j) = A(R_DQODW1, j - 1);
                A(R_DQIDW1, j) = A(R_DQODW2, j - 1);
            }
            const double dhkds1 = 0.5 * A(R_DHKDW, j) / WS(j), dhkds2 = 0.5 * A(R_DHKDW, j) / WS(j + 1);
            const double num = A(R_SMP, j + 1) - A(R_SMP, j);
            const double den = M_TO_MM * (S.z[V.sno(c, j + 1)] - S.z[V.sno(c, j)]);
            const double hk = A(R_HK, j);
            A(R_QOUT, j) = -(hk * num / den) + hk;
            A(R_DQODW1, j) = (hk * A(R_DSMPDW, j) - dhkds1 * num) / den + dhkds1;
            A(R_DQODW2, j) = (-(hk * A(R_DSMPDW, j + 1)) - dhkds2 * num) / den + dhkds2;
        }
        A(R_QIN, nl) = A(R_QOUT, nl - 1);
        A(R_DQIDW0, nl) = A(R_DQODW1, nl - 1);
        A(R_DQIDW1, nl) = A(R_DQODW2, nl - 1);
        if (P.lower_bc == CTSM_BC_FLUX) {
            A(R_QOUT, nl) = A(R_HK, nl);
            A(R_DQODW1, nl) = A(R_DHKDW, nl) / WS(nl);
        } else if (wt_bottom && jwt >= V.nlevsoi) {
            double qo, dq;
            watertable_bottom_flux(A(R_HK, nl), A(R_DHKDW, nl), A(R_SMP, nl), A(R_DSMPDW, nl), WS(nl),
                                   S.zwt[V.c1(c)], S.z[V.sno(c, nl)], qo, dq);
            A(R_QOUT, nl) = qo;
            A(R_DQODW1, nl) = dq;
        } else {
            A(R_QOUT, nl) = 0.0;
            A(R_DQODW1, nl) = 0.0;
        }
        A(R_DQODW2, nl) = 0.0;
        // RHS / LHS
        for (int j = 1; j <= nl; ++j) {
            const double fluxNet = A(R_QIN, j) - A(R_QOUT, j) - S.qflx_rootsoi[V.soi(c, j)];
            A(R_RMX, j) = -(fluxNet * A(R_DTDZ, j));
            A(R_AMX, j) = (j == 1) ? 0.0 : A(R_DQIDW0, j) * A(R_DTDZ, j);
            A(R_BMX, j) = -1.0 - (-A(R_DQIDW1, j) + A(R_DQODW1, j)) * A(R_DTDZ, j);
            A(R_CMX, j) = (j == nl) ? 0.0 : -(A(R_DQODW2, j) * A(R_DTDZ, j));
        }
        // dgtsv on copies (dl = amx(2:n), d = bmx, du = cmx(1:n-1), b = rmx)
        for (int j = 1; j <= nl; ++j) {
            A(R_D, j) = A(R_BMX, j);
            A(R_DWAT, j) = A(R_RMX, j);
            if (j < nl) { A(R_DL, j) = A(R_AMX, j + 1); A(R_DU, j) = A(R_CMX, j); }
        }
        bool singular = false;
        for (int k = 1; k <= nl - 1 && !singular; ++k) {
            const bool last = (k == nl - 1);
            if (fabs(A(R_D, k)) >= fabs(A(R_DL, k))) {
                if (A(R_D, k) != 0.0) {
                    const double fact = A(R_DL, k) / A(R_D, k);
                    A(R_D, k + 1) = A(R_D, k + 1) - fact * A(R_DU, k);
                    A(R_DWAT, k + 1) = A(R_DWAT, k + 1) - fact * A(R_DWAT, k);
                } else {
                    singular = true;
                }
                if (!last) A(R_DL, k) = 0.0;
            } else {
                const double fact = A(R_D, k) / A(R_DL, k);
                A(R_D, k) = A(R_DL, k);
                const double temp = A(R_D, k + 1);
                A(R_D, k + 1) = A(R_DU, k) - fact * temp;
                if (!last) {
                    A(R_DL, k) = A(R_DU, k + 1);
                    A(R_DU, k + 1) = -fact * A(R_DL, k);
                }
                A(R_DU, k) = temp;
                const double tb = A(R_DWAT, k);
                A(R_DWAT, k) = A(R_DWAT, k + 1);
                A(R_DWAT, k + 1) = tb - fact * A(R_DWAT, k + 1);
            }
        }
        if (singular || A(R_D, nl) == 0.0) { report_error(err, c, CTSM_E_SOLVER); return; }
        A(R_DWAT, nl) = A(R_DWAT, nl) / A(R_D, nl);
        A(R_DWAT, nl - 1) = (A(R_DWAT, nl - 1) - A(R_DU, nl - 1) * A(R_DWAT, nl)) / A(R_D, nl - 1);
        for (int k = nl - 2; k >= 1; --k)
            A(R_DWAT, k) = (A(R_DWAT, k) - A(R_DU, k) * A(R_DWAT, k + 1) - A(R_DL, k) * A(R_DWAT, k + 2)) / A(R_D, k);
        // error estimate
        double errorMax = 0.0;
        for (int j = 1; j <= nl; ++j) {
            const double f0 = A(R_DWAT, j) / A(R_DTDZ, j);
            const double f1 = A(R_QIN, j) - A(R_QOUT, j) - S.qflx_rootsoi[V.soi(c, j)];
            const double e = fabs(f1 - f0) * dtsub * 0.5;
            if (j == 1 || e > errorMax) errorMax = e;
        }
        if (errorMax > P.xTolerUpper && dtsub > P.dtmin) {
            dtsub = fmax90(dtsub / 2.0, P.dtmin);
            continue;
        }
        for (int j = 1; j <= nl; ++j) LIQ(j) = LIQ(j) + A(R_DWAT, j) * (M_TO_MM * DZ(j));
        const double qcTemp = (P.lower_bc == CTSM_BC_FLUX) ? A(R_HK, nl) + A(R_DHKDW, nl) * A(R_DWAT, nl)
                              : wt_bottom ? A(R_QOUT, nl) + A(R_DQODW1, nl) * A(R_DWAT, nl) : 0.0;
        qcharge = qcharge + qcTemp * (dtsub / dtime);
        dtdone = dtdone + dtsub;
        if (fabs(dtime - dtdone) < P.verySmall) break;
        if (errorMax < P.xTolerLower) dtsub = dtsub * 2.0;
        dtsub = fmin90(dtsub, dtime - dtdone);
    }
    if (wt_bottom && jwt < V.nlevsoi) {
        // compute_qcharge on the last trial's vwc_liq, imped and smp
        const double ka = watertable_ka(A(R_VWC, jwt + 1), WS(jwt + 1), A(R_IMPED, jwt + 1) * S.hksat[V.soi(c, jwt + 1)],
                                        S.bsw[V.soi(c, jwt + 1)]);
        qcharge = watertable_qcharge(V, S, c, jwt, ka, A(R_SMP, jwt > 1 ? jwt : 1), dtime);
    }
    S.num_substeps[V.c1(c)] = (double)nsub;
    S.qcharge[V.c1(c)] = qcharge;
    for (int j = nl; j >= 2; --j) {
        const double cap = S.eff_porosity[V.soi(c, j)] * M_TO_MM * DZ(j);
        const double over = fmax90(LIQ(j) - cap, 0.0);
        LIQ(j) = fmin90(cap, LIQ(j));
        LIQ(j - 1) = LIQ(j - 1) + over;
    }
    for (int j = 1; j <= nl; ++j) {
        S.qin[V.soi(c, j)] = A(R_QIN, j);
        S.qout[V.soi(c, j)] = A(R_QOUT, j);
    }
#undef A
#undef LIQ
#undef DZ
#undef WS
}

}  // namespace

size_t soiltemp_plain_workspace_doubles(int n) { return (size_t)S_TOTAL * (size_t)n; }
size_t soilwater_plain_workspace_doubles(int n) { return (size_t)R_NF * MAX_NLEVSOI * (size_t)n; }

void launch_soiltemp_plain(hipStream_t st, const View &V, const Params &P, const ctsm_col_state &S,
                           const ctsm_patch_state &Pt, int n, const int32_t *filter, double dtime,
                           double *ws, long ld, ErrWord *err)
{
    if (n <= 0) return;
    const int bs = 128;
    const dim3 g((n + bs - 1) / bs), b(bs);
    PW w{ws, ld};
    const GptlRange range("soiltemperature");
    {
        const GptlRange r("SoilThermProp");
        hipLaunchKernelGGL(k_thermprop, g, b, 0, st, V, P, S, n, filter, w);
    }
    hipLaunchKernelGGL(k_groundflux, g, b, 0, st, V, S, Pt, n, filter, w);
    hipLaunchKernelGGL(k_diffflux, g, b, 0, st, V, S, n, filter, dtime, w);
    hipLaunchKernelGGL(k_rhs, g, b, 0, st, V, S, n, filter, dtime, w);
    hipLaunchKernelGGL(k_matrix, g, b, 0, st, V, S, n, filter, dtime, w);
    {
        const GptlRange r("SoilTempBandDiag");
        hipLaunchKernelGGL(k_banddiag, g, b, 0, st, V, S, n, filter, w, err);
    }
    hipLaunchKernelGGL(k_fn1_h2osfc, g, b, 0, st, V, S, n, filter, dtime, w);
    {
        const GptlRange r("PhaseChange");
        hipLaunchKernelGGL(k_phasechange, g, b, 0, st, V, S, n, filter, dtime, w);
    }
    hipLaunchKernelGGL(k_tail, g, b, 0, st, V, S, n, filter, w);
}

void launch_soilwater_plain(hipStream_t st, const View &V, const Params &P, const ctsm_col_state &S, int n,
                            const int32_t *filter, double dtime, double *ws, long ld, ErrWord *err)
{
    if (n <= 0) return;
    const GptlRange range("hydro_without_drainage");
    const int bs = 128;
    PW w{ws, ld};
    hipLaunchKernelGGL(k_soilwater_plain, dim3((n + bs - 1) / bs), dim3(bs), 0, st, V, P, S, n, filter, dtime, w, err);
}

}  // namespace ctsm
