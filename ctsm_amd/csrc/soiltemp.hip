// soiltemp.hip -- snow + soil heat solve with phase change on gfx950.
//
// Replaces SoilTemperature (SoilTemperatureMod.F90:92-508) for non-urban,
// non-lake columns with use_excess_ice = .false.  One column per lane, three
// sweeps over the column's rows (snow layers snl+1..0, the standing-water row,
// ground layers 1..nlevgrnd):
//
//   1. down: thermal properties (SoilThermProp :602-901), heat-diffusion flux
//      and factor (:1799-1910), the RHS and the 5-band row of each level
//      (SetRHSVec* :1913-2353, SetMatrix* :2356-2926) assembled in a 3-row
//      window, and LAPACK dgbtf2's pivoted elimination of the column two rows
//      up, with dgbtrs's swap + unit-lower solve of the RHS applied in the
//      same step (identical operations in identical order).  Final U rows go
//      to a coalesced workspace [row][entry][lane];
//   2. up:   dtbsv back substitution (upper, band 4) -> new temperatures;
//   3. down: fn1, PhaseChangeH2osfc (:904-1130), Phasechange (:1133-1540),
//      t_grnd and the soil heat fluxes (:457-504).
//
// The patch -> column ground heat flux (ComputeGroundHeatFluxAndDeriv
// :1543-1796) runs first, looping the column's patches in reference order.
#include "common.h"
#include "trace.h"
#include "soiltemp_dev.h"
#include "soilwater_dev.h"

namespace ctsm {

// Workspace per lane: U row (5 band entries + transformed RHS) per row.
constexpr int UFIELDS = 6;
// Per-level saved values for sweep 3: fn (old flux), tk.
constexpr int LFIELDS = 2;
constexpr int LEVELS = MAX_NLEVSNO + MAX_NLEVGRND + 1;
// The column's absorbed solar radiation of snow levels ltop..0 and of level 1
// (the patch sum of sabg_lyr * wtcol), index j + nlevsno - 1.
constexpr int SLEVELS = MAX_NLEVSNO + 1;
constexpr int WS_DOUBLES_PER_COL = MAX_ROWS * UFIELDS + LEVELS * LFIELDS + SLEVELS;

struct TempWork {
    double *ws;     // workspace base
    long    ld;     // lane stride (>= launch size)
    __device__ __forceinline__ double &U(int lane, int row, int f) const { return ws[((long)row * UFIELDS + f) * ld + lane]; }
    __device__ __forceinline__ double &Lv(int lane, int lev_ix, int f) const {
        return ws[((long)MAX_ROWS * UFIELDS + (long)lev_ix * LFIELDS + f) * ld + lane];
    }
    __device__ __forceinline__ double &Sab(int lane, int jj) const {
        return ws[((long)MAX_ROWS * UFIELDS + (long)LEVELS * LFIELDS + jj) * ld + lane];
    }
};

// ---- ComputeGroundHeatFluxAndDeriv (SoilTemperatureMod.F90:1543-1796) ----
// The patch -> column part: per patch the net ground heat flux and its
// derivative (eflx_gnet, dgnetdT, sabg_chk written), summed into the column
// in patch order (hs_soil, hs_h2osfc, dhsdT, then hs_top_snow, and the
// absorbed solar radiation of each snow level).  With one column per lane the
// lanes would walk their patches in step, each load touching a cache line per
// lane; instead the 64 lanes of a wave take the wave's patches in order, 64 at
// a time (coalesced), and each column's lane then adds its patches' terms from
// LDS in patch order: the same products summed in the same order, so the same
// bits as the per-column loops.
struct GroundFlux {
    double hs_soil, hs_h2osfc, dhsdT, hs_top_snow;
};

// LDS of one wave: the columns' first patch, offset in the wave's patch list,
// top level and constants (emg, forc_lwrad, htvp, frac_sno and the emitted
// longwave terms), and the terms of the 64 patches in flight.
struct GfWave {
    int pi[64], off[64], ltop[64];
    double cc[9][64];
    double x[4][64];
    double y[SLEVELS][64];
};

// Every lane of the wave calls it (valid: the lane has a column in range);
// the sums of a lane whose column soiltemp_column rejects are not used.  The
// snow levels' sums go to wk.Sab(i, .).
__device__ void ground_flux_of_wave(const View &V, const ctsm_col_state &S, const ctsm_patch_state &Pt, int c,
                                    bool valid, GfWave &L, const TempWork &wk, int i, GroundFlux &g)
{
    const int lane = threadIdx.x & 63;
    const int NS = V.nlevsno;
    int n_own = 0, pi = 0, ltop = 1;
    if (valid) {
        const long cc = V.c1(c);
        const int snl = S.snl[cc];
        const int p0 = S.patchi[cc], p1 = S.patchf[cc];
        if (!(snl < -NS || snl > 0) && !(p0 < V.begp || p1 >= V.begp + V.ldp)) {
            ltop = snl + 1;
            pi = p0;
            n_own = (p1 >= p0) ? p1 - p0 + 1 : 0;
            const double emg = S.emg[cc];
            const double tg = S.t_grnd[cc], ts = S.t_soisno[V.sno(c, ltop)];
            const double t1 = S.t_soisno[V.sno(c, 1)], th = S.t_h2osfc[cc];
            const double tg2 = tg * tg, ts2 = ts * ts, t12 = t1 * t1, th2 = th * th;
            L.cc[0][lane] = emg;
            L.cc[1][lane] = S.forc_lwrad[cc];
            L.cc[2][lane] = S.htvp[cc];
            L.cc[3][lane] = S.frac_sno_eff[cc];
            L.cc[4][lane] = emg * SB * (tg2 * tg2);          // lwrad_emit
            L.cc[5][lane] = 4.0 * emg * SB * (tg2 * tg);     // dlwrad_emit
            L.cc[6][lane] = emg * SB * (ts2 * ts2);          // lwrad_emit_snow
            L.cc[7][lane] = emg * SB * (t12 * t12);          // lwrad_emit_soil
            L.cc[8][lane] = emg * SB * (th2 * th2);          // lwrad_emit_h2osfc
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
    L.pi[lane] = pi;
    L.off[lane] = off;
    L.ltop[lane] = ltop;
    double dhsdT = 0.0, hs_soil = 0.0, hs_h2osfc = 0.0, hs_top_snow = 0.0;
    double sab[SLEVELS];
#pragma unroll
    for (int jj = 0; jj < SLEVELS; ++jj) sab[jj] = 0.0;
    const int jj_top = ltop + NS - 1;
    wave_sync();
    for (int k0 = 0; k0 < total; k0 += 64) {
        const int k = k0 + lane;
        if (k < total) {
            int lo = 0, hi = 63;       // the column: the last lane whose list starts at or before k
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (L.off[mid] <= k) lo = mid;
                else hi = mid - 1;
            }
            const int o = lo;
            const int p = L.pi[o] + (k - L.off[o]);
            const long pp = V.p1(p);
            const double emg = L.cc[0][o], forc_lwrad = L.cc[1][o], htvp = L.cc[2][o], frac_sno = L.cc[3][o];
            const double fveg = (double)Pt.frac_veg_nosno[pp], wt = Pt.wtcol[pp];
            const double lwin = (1.0 - fveg) * emg * forc_lwrad;
            const double dl = Pt.dlrad[pp];
            const double eg = Pt.sabg[pp] + dl + lwin - L.cc[4][o] -
                              (Pt.eflx_sh_grnd[pp] + Pt.qflx_evap_soi[pp] * htvp);
            Pt.eflx_gnet[pp] = eg;
            Pt.sabg_chk[pp] = frac_sno * Pt.sabg_snow[pp] + (1.0 - frac_sno) * Pt.sabg_soil[pp];
            const double egsoil = Pt.sabg_soil[pp] + dl + lwin - L.cc[7][o] -
                                  (Pt.eflx_sh_soil[pp] + Pt.qflx_ev_soil[pp] * htvp);
            const double egh2o = Pt.sabg_soil[pp] + dl + lwin - L.cc[8][o] -
                                 (Pt.eflx_sh_h2osfc[pp] + Pt.qflx_ev_h2osfc[pp] * htvp);
            const double dg = -Pt.cgrnd[pp] - L.cc[5][o];
            Pt.dgnetdT[pp] = dg;
            const int lt = L.ltop[o];
            const double sl = Pt.sabg_lyr[V.lyr(p, lt)];
            const double egsnow = sl + dl + lwin - L.cc[6][o] - (Pt.eflx_sh_snow[pp] + Pt.qflx_ev_snow[pp] * htvp);
            L.x[0][lane] = dg * wt;
            L.x[1][lane] = egsoil * wt;
            L.x[2][lane] = egh2o * wt;
            L.x[3][lane] = egsnow * wt;
            for (int j = lt; j <= 1; ++j) L.y[j + NS - 1][lane] = Pt.sabg_lyr[V.lyr(p, j)] * wt;
        }
        wave_sync();
        if (n_own > 0) {
            const int a = (off > k0 ? off : k0) - k0;
            const int e = off + n_own, b = (e < k0 + 64 ? e : k0 + 64) - k0;
            for (int s = a; s < b; ++s) {
                dhsdT = dhsdT + L.x[0][s];
                hs_soil = hs_soil + L.x[1][s];
                hs_h2osfc = hs_h2osfc + L.x[2][s];
                hs_top_snow = hs_top_snow + L.x[3][s];
#pragma unroll
                for (int jj = 0; jj < SLEVELS; ++jj)
                    if (jj >= jj_top && jj <= NS) sab[jj] = sab[jj] + L.y[jj][s];
            }
        }
        wave_sync();
    }
    g.hs_soil = hs_soil;
    g.hs_h2osfc = hs_h2osfc;
    g.dhsdT = dhsdT;
    g.hs_top_snow = hs_top_snow;
    if (valid) {
#pragma unroll
        for (int jj = 0; jj < SLEVELS; ++jj)
            if (jj >= jj_top && jj <= NS) wk.Sab(i, jj) = sab[jj];
    }
}

// The 3-row elimination window: W[a][m] = A(k+a, k+m), m = 0..4; B[a] = rhs.
struct Window {
    double W[3][5];
    double B[3];
};

__device__ __forceinline__ void window_shift(Window &w)
{
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int m = 0; m < 4; ++m) w.W[a][m] = w.W[a + 1][m + 1];
        w.W[a][4] = 0.0;
        w.B[a] = w.B[a + 1];
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) w.W[2][m] = 0.0;
    w.B[2] = 0.0;
}

// One column of dgbtf2 (kl = ku = 2) plus the matching dgbtrs L-solve step.
// km = number of sub-diagonal rows that exist.  Returns false on a zero pivot.
__device__ __forceinline__ bool eliminate(Window &w, int km)
{
    int jp = 0;
    double dmx = fabs(w.W[0][0]);
    if (km >= 1 && fabs(w.W[1][0]) > dmx) { dmx = fabs(w.W[1][0]); jp = 1; }
    if (km >= 2 && fabs(w.W[2][0]) > dmx) { dmx = fabs(w.W[2][0]); jp = 2; }
    double piv = (jp == 0) ? w.W[0][0] : ((jp == 1) ? w.W[1][0] : w.W[2][0]);
    if (piv == 0.0) return false;
    if (jp != 0) {
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const double t0 = w.W[0][m];
            const double tp = (jp == 1) ? w.W[1][m] : w.W[2][m];
            w.W[0][m] = tp;
            if (jp == 1) w.W[1][m] = t0; else w.W[2][m] = t0;
        }
        const double b0 = w.B[0];
        const double bp = (jp == 1) ? w.B[1] : w.B[2];
        w.B[0] = bp;
        if (jp == 1) w.B[1] = b0; else w.B[2] = b0;
    }
    if (km > 0) {
        const double r = 1.0 / w.W[0][0];
        double l1 = w.W[1][0] * r, l2 = (km >= 2) ? w.W[2][0] * r : 0.0;
        w.W[1][0] = l1;
        if (km >= 2) w.W[2][0] = l2;
#pragma unroll
        for (int m = 1; m < 5; ++m) {
            const double y = w.W[0][m];
            if (y != 0.0) {
                const double temp = -1.0 * y;
                w.W[1][m] = w.W[1][m] + l1 * temp;
                if (km >= 2) w.W[2][m] = w.W[2][m] + l2 * temp;
            }
        }
        if (w.B[0] != 0.0) {
            const double temp = -1.0 * w.B[0];
            w.B[1] = w.B[1] + l1 * temp;
            if (km >= 2) w.B[2] = w.B[2] + l2 * temp;
        }
    }
    return true;
}

// Full column.  `lane` indexes the workspace.  Returns 0 or an error code.
__device__ int soiltemp_column(const View &V, const Params &P, const ctsm_col_state &S,
                               const ctsm_patch_state &Pt, int c, double dtime,
                               const TempWork &wk, int lane, const GroundFlux &gf)
{
    const int nlevgrnd = V.nlevgrnd;
    const int snl = S.snl[V.c1(c)], nbedrock = S.nbedrock[V.c1(c)];
    if (snl < -V.nlevsno || snl > 0) return 1;
    const double frac_sno = S.frac_sno_eff[V.c1(c)], frac_h2osfc = S.frac_h2osfc[V.c1(c)];
    const double emg = S.emg[V.c1(c)], forc_lwrad = S.forc_lwrad[V.c1(c)], htvp = S.htvp[V.c1(c)];
    double h2osno_no_layers = S.h2osno_no_layers[V.c1(c)];
    const double h2osfc0 = S.h2osfc[V.c1(c)];
    double t_h2osfc = S.t_h2osfc[V.c1(c)];
    const int ltop = snl + 1;     // top level

    // ---- ComputeGroundHeatFluxAndDeriv: summed over the wave (ground_flux_of_wave)
    const double hs_soil = gf.hs_soil, hs_h2osfc = gf.hs_h2osfc, dhsdT = gf.dhsdT, hs_top_snow = gf.hs_top_snow;
    const int pi = S.patchi[V.c1(c)], pf = S.patchf[V.c1(c)];
    if (pi < V.begp || pf >= V.begp + V.ldp) return 1;
    // column absorbed solar of level j in [ltop, 1], summed over patches in order
    auto sabg_col = [&](int j) -> double { return wk.Sab(lane, j + V.nlevsno - 1); };

    // ---- standing water properties (:257-266, :816-821)
    double c_h2osfc, dz_h2osfc;
    if (h2osfc0 > THIN_SFCLAYER && frac_h2osfc > THIN_SFCLAYER) {
        c_h2osfc = fmax90(THIN_SFCLAYER, CPLIQ * h2osfc0 / frac_h2osfc);
        dz_h2osfc = fmax90(THIN_SFCLAYER, (double)1.0e-3f * h2osfc0 / frac_h2osfc);
    } else {
        c_h2osfc = THIN_SFCLAYER;
        dz_h2osfc = THIN_SFCLAYER;
    }
    S.c_h2osfc[V.c1(c)] = c_h2osfc;
    const double z1 = S.z[V.sno(c, 1)];
    const double dzm_h2osfc = (0.5 * dz_h2osfc + z1);

    // ---- sweep 1: levels top-down, rows assembled into the window
    const int n = nlevgrnd - snl + 1;      // rows jtop..jbot
    Window w;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int m = 0; m < 5; ++m) w.W[a][m] = 0.0;
        w.B[a] = 0.0;
    }
    int nrow = 0;        // rows assembled
    int kelim = 0;       // next column to eliminate
    bool ok = true;
    // place a row's 5 band entries (b1..b5 as the reference's bmatrix bands) and RHS
    auto put_row = [&](double b1, double b2, double b3, double b4, double b5, double r) {
        const int a = nrow - kelim;     // window slot 0..2
        // A(row, row-2..row+2) -> window columns (row-kelim)-2 .. +2
        const double e[5] = {b5, b4, b3, b2, b1};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            if (s == a) {
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    const int col = a - 2 + m;   // window column index
                    if (col >= 0 && col < 5) w.W[s][col] = e[m];
                }
                w.B[s] = r;
            }
        }
        ++nrow;
        if (nrow - kelim == 3) {
            // eliminate column kelim with 2 sub-diagonal rows
            if (!eliminate(w, 2)) ok = false;
#pragma unroll
            for (int m = 0; m < 5; ++m) wk.U(lane, kelim, m) = w.W[0][m];
            wk.U(lane, kelim, 5) = w.B[0];
            window_shift(w);
            ++kelim;
        }
    };

    double tkh2osfc = 0.0, fn_h2osfc = 0.0;
    // lookahead state for level j: thk(j), t(j), z(j)
    double thk_j = level_thk(V, P, S, c, ltop, nbedrock, frac_sno);
    double tk_prev = 0.0, fn_prev = 0.0;
    for (int j = ltop; j <= nlevgrnd; ++j) {
        const double tj = S.t_soisno[V.sno(c, j)];
        const double zj = S.z[V.sno(c, j)];
        double thk_n = 0.0, tn = 0.0, zn = 0.0, tk_j, fn_j, fact_j;
        S.thk[V.sno(c, j)] = thk_j;
        if (j < nlevgrnd) {
            thk_n = level_thk(V, P, S, c, j + 1, nbedrock, frac_sno);
            tn = S.t_soisno[V.sno(c, j + 1)];
            zn = S.z[V.sno(c, j + 1)];
            const double zij = S.zi[V.zi(c, j)];
            tk_j = thk_j * thk_n * (zn - zj) / (thk_j * (zn - zij) + thk_n * (zij - zj));
        } else {
            tk_j = 0.0;
        }
        const double cv = level_cv(V, S, c, j, nbedrock, frac_sno, h2osno_no_layers);
        if (j == ltop) {
            const double zim = S.zi[V.zi(c, j - 1)];
            fact_j = dtime / cv * S.dz[V.sno(c, j)] / (0.5 * (zj - zim + CAPR * (zn - zim)));
            fn_j = tk_j * (tn - tj) / (zn - zj);
        } else if (j <= nlevgrnd - 1) {
            fact_j = dtime / cv;
            fn_j = tk_j * (tn - tj) / (zn - zj);
        } else {
            fact_j = dtime / cv;
            fn_j = S.eflx_bot[V.c1(c)];
        }
        S.fact[V.sno(c, j)] = fact_j;
        wk.Lv(lane, j + V.nlevsno, 0) = fn_j;
        wk.Lv(lane, j + V.nlevsno, 1) = tk_j;

        // ---- rows
        if (j <= 0) {
            // snow layer j -> row of level j-1
            const double dzp = zn - zj;
            double b1 = 0.0, b2 = 0.0, b3, b4, r;
            if (j == ltop) {
                b4 = 0.0;
                b3 = 1.0 + (1.0 - CNFAC) * fact_j * tk_j / dzp - fact_j * dhsdT;
                r = tj + fact_j * (hs_top_snow - dhsdT * tj + CNFAC * fn_j);
            } else {
                const double dzm = (zj - S.z[V.sno(c, j - 1)]);
                b4 = -((1.0 - CNFAC) * fact_j * tk_prev / dzm);
                b3 = 1.0 + (1.0 - CNFAC) * fact_j * (tk_j / dzp + tk_prev / dzm);
                r = tj + CNFAC * fact_j * (fn_j - fn_prev);
                r = r + fact_j * sabg_col(j);
            }
            if (j != 0) b2 = -((1.0 - CNFAC) * fact_j * tk_j / dzp);
            else        b1 = -((1.0 - CNFAC) * fact_j * tk_j / dzp);
            put_row(b1, b2, b3, b4, 0.0, r);
            if (j == 0) {
                // level 0: standing water row, needs thk(1) (= thk_n here)
                const double zh2osfc = (double)1.0e-3f * (0.5 * h2osfc0);
                tkh2osfc = TKWAT * thk_n * (z1 + zh2osfc) / (TKWAT * z1 + thk_n * zh2osfc);
                fn_h2osfc = tkh2osfc * (tn - t_h2osfc) / dzm_h2osfc;
                const double rr = t_h2osfc + (dtime / c_h2osfc) * (hs_h2osfc - dhsdT * t_h2osfc + CNFAC * fn_h2osfc);
                const double bb3 = 1.0 + (1.0 - CNFAC) * (dtime / c_h2osfc) * tkh2osfc / dzm_h2osfc - (dtime / c_h2osfc) * dhsdT;
                const double bb2 = -((1.0 - CNFAC) * (dtime / c_h2osfc) * tkh2osfc / dzm_h2osfc);
                put_row(0.0, bb2, bb3, 0.0, 0.0, rr);
            }
        } else {
            if (j == 1 && ltop == 1) {
                // no snow: the standing-water row is the top row (level 0)
                const double zh2osfc = (double)1.0e-3f * (0.5 * h2osfc0);
                tkh2osfc = TKWAT * thk_j * (z1 + zh2osfc) / (TKWAT * z1 + thk_j * zh2osfc);
                fn_h2osfc = tkh2osfc * (tj - t_h2osfc) / dzm_h2osfc;
                const double rr = t_h2osfc + (dtime / c_h2osfc) * (hs_h2osfc - dhsdT * t_h2osfc + CNFAC * fn_h2osfc);
                const double bb3 = 1.0 + (1.0 - CNFAC) * (dtime / c_h2osfc) * tkh2osfc / dzm_h2osfc - (dtime / c_h2osfc) * dhsdT;
                const double bb2 = -((1.0 - CNFAC) * (dtime / c_h2osfc) * tkh2osfc / dzm_h2osfc);
                put_row(0.0, bb2, bb3, 0.0, 0.0, rr);
            }
            double b2, b3, b4 = 0.0, b5 = 0.0, r;
            if (j == ltop) {
                const double dzp = zn - zj;
                b3 = 1.0 + (1.0 - CNFAC) * fact_j * tk_j / dzp - fact_j * dhsdT;
                b2 = -((1.0 - CNFAC) * fact_j * tk_j / dzp);
                r = tj + fact_j * (hs_top_snow - dhsdT * tj + CNFAC * fn_j);
            } else if (j == 1) {
                const double dzm = (zj - S.z[V.sno(c, 0)]);
                const double dzp = (zn - zj);
                b2 = -((1.0 - CNFAC) * fact_j * tk_j / dzp);
                b3 = 1.0 + (1.0 - CNFAC) * fact_j * (tk_j / dzp + frac_sno * tk_prev / dzm) -
                     (1.0 - frac_sno) * fact_j * dhsdT;
                b5 = -(frac_sno * (1.0 - CNFAC) * fact_j * tk_prev / dzm);
                r = tj + fact_j * ((1.0 - frac_sno) * (hs_soil - dhsdT * tj) + CNFAC * (fn_j - frac_sno * fn_prev));
                r = r + frac_sno * fact_j * sabg_col(1);
            } else if (j <= nlevgrnd - 1) {
                const double dzm = (zj - S.z[V.sno(c, j - 1)]);
                const double dzp = (zn - zj);
                b2 = -((1.0 - CNFAC) * fact_j * tk_j / dzp);
                b3 = 1.0 + (1.0 - CNFAC) * fact_j * (tk_j / dzp + tk_prev / dzm);
                b4 = -((1.0 - CNFAC) * fact_j * tk_prev / dzm);
                r = tj + CNFAC * fact_j * (fn_j - fn_prev);
            } else {
                const double dzm = (zj - S.z[V.sno(c, j - 1)]);
                b2 = 0.0;
                b3 = 1.0 + (1.0 - CNFAC) * fact_j * tk_prev / dzm;
                b4 = -((1.0 - CNFAC) * fact_j * tk_prev / dzm);
                r = tj - CNFAC * fact_j * fn_prev + fact_j * fn_j;
            }
            if (j == 1) {
                if (frac_h2osfc != 0.0) {
                    r = r - frac_h2osfc * fact_j * ((hs_soil - dhsdT * tj) + CNFAC * fn_h2osfc);
                    b3 = b3 + frac_h2osfc * ((1.0 - CNFAC) * fact_j * tkh2osfc / dzm_h2osfc + fact_j * dhsdT);
                    b4 = -(frac_h2osfc * (1.0 - CNFAC) * fact_j * tkh2osfc / dzm_h2osfc);
                } else {
                    b4 = 0.0;
                }
            }
            put_row(0.0, b2, b3, b4, b5, r);
        }
        tk_prev = tk_j;
        fn_prev = fn_j;
        thk_j = thk_n;
    }
    // remaining columns: n-2 (one sub-diagonal row), n-1 (none)
    if (n >= 2) {
        if (!eliminate(w, 1)) ok = false;
#pragma unroll
        for (int m = 0; m < 5; ++m) wk.U(lane, kelim, m) = w.W[0][m];
        wk.U(lane, kelim, 5) = w.B[0];
        window_shift(w);
        ++kelim;
    }
    if (!eliminate(w, 0)) ok = false;
#pragma unroll
    for (int m = 0; m < 5; ++m) wk.U(lane, kelim, m) = w.W[0][m];
    wk.U(lane, kelim, 5) = w.B[0];
    if (!ok) return 2;

    // ---- sweep 2: dtbsv back substitution, rows n-1 .. 0
    double x1 = 0.0, x2 = 0.0, x3 = 0.0, x4 = 0.0;   // x(i+1..i+4)
    for (int i = n - 1; i >= 0; --i) {
        double x = wk.U(lane, i, 5);
        const double u1 = wk.U(lane, i, 1), u2 = wk.U(lane, i, 2), u3 = wk.U(lane, i, 3), u4 = wk.U(lane, i, 4);
        // contributions in dtbsv's order: column i+4 first, then i+3, i+2, i+1
        if (i + 4 <= n - 1 && x4 != 0.0) x = x - x4 * u4;
        if (i + 3 <= n - 1 && x3 != 0.0) x = x - x3 * u3;
        if (i + 2 <= n - 1 && x2 != 0.0) x = x - x2 * u2;
        if (i + 1 <= n - 1 && x1 != 0.0) x = x - x1 * u1;
        if (x != 0.0) x = x / wk.U(lane, i, 0);
        x4 = x3; x3 = x2; x2 = x1; x1 = x;
        // row i -> level snl + i
        const int lev = snl + i;
        if (lev < 0) S.t_soisno[V.sno(c, lev + 1)] = x;
        else if (lev == 0) t_h2osfc = x;
        else S.t_soisno[V.sno(c, lev)] = x;
    }
    if (frac_h2osfc == 0.0) t_h2osfc = S.t_soisno[V.sno(c, 1)];

    // ---- sweep 3a: fn1 and soil heat fluxes need pre-phase-change temperatures
    double fn1_1 = 0.0;
    {
        double tcur = S.t_soisno[V.sno(c, ltop)];
        for (int j = ltop; j <= nlevgrnd; ++j) {
            double fn1;
            double tnext = 0.0;
            if (j <= nlevgrnd - 1) {
                tnext = S.t_soisno[V.sno(c, j + 1)];
                fn1 = wk.Lv(lane, j + V.nlevsno, 1) * (tnext - tcur) / (S.z[V.sno(c, j + 1)] - S.z[V.sno(c, j)]);
            } else {
                fn1 = 0.0;
            }
            const double fnj = wk.Lv(lane, j + V.nlevsno, 0);
            if (j == 1) fn1_1 = fn1;
            if (j >= 1) {
                if (j < nlevgrnd) S.eflx_fgr[V.soi(c, j)] = -(CNFAC * fnj) - (1.0 - CNFAC) * fn1;
                else S.eflx_fgr[V.soi(c, j)] = 0.0;
            }
            tcur = tnext;
        }
        S.eflx_fgr12[V.c1(c)] = -(CNFAC * wk.Lv(lane, 1 + V.nlevsno, 0)) - (1.0 - CNFAC) * fn1_1;
    }

    // ---- PhaseChangeH2osfc (:904-1130)
    double snow_depth = S.snow_depth[V.c1(c)];
    double xmf_h2osfc = 0.0, qflx_h2osfc_to_ice = 0.0, eflx_h2osfc_to_snow = 0.0;
    double h2osfc = h2osfc0;
    if (frac_h2osfc > 0.0 && t_h2osfc <= TFRZ) {
        double h2osno_total = h2osno_no_layers;
        for (int j = ltop; j <= 0; ++j)
            h2osno_total = h2osno_total + S.h2osoi_ice[V.sno(c, j)] + S.h2osoi_liq[V.sno(c, j)];
        const double tinc = TFRZ - t_h2osfc;
        t_h2osfc = TFRZ;
        const double hm = frac_h2osfc * (dhsdT * tinc - tinc * c_h2osfc / dtime);
        const double xm = hm * dtime / HFUS;
        const double temp1 = h2osfc + xm;
        const double z_avg = frac_sno * snow_depth;
        double rho_avg = (z_avg > 0.0) ? fmin90(800.0, h2osno_total / z_avg) : 200.0;
        const double fact0 = (snl < 0) ? S.fact[V.sno(c, 0)] : 0.0;
        double c1, c2;
        if (temp1 >= 0.0) {
            S.int_snow[V.c1(c)] = S.int_snow[V.c1(c)] - xm;
            if (snl == 0) h2osno_no_layers = h2osno_no_layers - xm;
            else S.h2osoi_ice[V.sno(c, 0)] = S.h2osoi_ice[V.sno(c, 0)] - xm;
            h2osno_total = h2osno_total - xm;
            h2osfc = h2osfc + xm;
            xmf_h2osfc = hm;
            qflx_h2osfc_to_ice = -xm / dtime;
            if (frac_sno > 0 && snl < 0) snow_depth = h2osno_total / (rho_avg * frac_sno);
            else snow_depth = h2osno_total / DENICE;
            if (snl == 0) {
                S.t_soisno[V.sno(c, 0)] = t_h2osfc;
                eflx_h2osfc_to_snow = 0.0;
            } else {
                if (snl == -1) c1 = frac_sno * (dtime / fact0 - dhsdT * dtime);
                else c1 = frac_sno / fact0 * dtime;
                if (frac_h2osfc != 0.0) c2 = (-CPLIQ * xm - frac_h2osfc * dhsdT * dtime);
                else c2 = 0.0;
                const double t0 = (c1 * S.t_soisno[V.sno(c, 0)] + c2 * t_h2osfc) / (c1 + c2);
                S.t_soisno[V.sno(c, 0)] = t0;
                eflx_h2osfc_to_snow = (t_h2osfc - t0) * c2 / dtime;
            }
        } else {
            rho_avg = (h2osno_total * rho_avg + h2osfc * DENICE) / (h2osno_total + h2osfc);
            S.int_snow[V.c1(c)] = S.int_snow[V.c1(c)] + h2osfc;
            if (snl == 0) h2osno_no_layers = h2osno_no_layers + h2osfc;
            else S.h2osoi_ice[V.sno(c, 0)] = S.h2osoi_ice[V.sno(c, 0)] + h2osfc;
            h2osno_total = h2osno_total + h2osfc;
            qflx_h2osfc_to_ice = h2osfc / dtime;
            t_h2osfc = t_h2osfc - temp1 * HFUS / (dtime * dhsdT - c_h2osfc);
            xmf_h2osfc = (hm - frac_h2osfc * temp1 * HFUS / dtime);
            if (snl == 0) {
                S.t_soisno[V.sno(c, 0)] = t_h2osfc;
            } else {
                if (snl == -1) c1 = frac_sno * (dtime / fact0 - dhsdT * dtime);
                else c1 = frac_sno / fact0 * dtime;
                if (frac_h2osfc != 0.0) c2 = frac_h2osfc * (c_h2osfc - dtime * dhsdT);
                else c2 = 0.0;
                const double t0 = (c1 * S.t_soisno[V.sno(c, 0)] + c2 * t_h2osfc) / (c1 + c2);
                S.t_soisno[V.sno(c, 0)] = t0;
                t_h2osfc = t0;
            }
            h2osfc = 0.0;
            if (frac_sno > 0 && snl < 0) snow_depth = h2osno_total / (rho_avg * frac_sno);
            else snow_depth = h2osno_total / DENICE;
        }
    }
    S.h2osfc[V.c1(c)] = h2osfc;
    S.xmf_h2osfc[V.c1(c)] = xmf_h2osfc;
    S.qflx_h2osfc_to_ice[V.c1(c)] = qflx_h2osfc_to_ice;
    S.eflx_h2osfc_to_snow[V.c1(c)] = eflx_h2osfc_to_snow;

    // ---- Phasechange (:1133-1540), levels top-down
    double xmf = 0.0, qflx_snomelt = 0.0, qflx_snofrz = 0.0, qflx_snow_drain = 0.0;
    double snomelt_accum = S.snomelt_accum ? S.snomelt_accum[V.c1(c)] : 0.0;
    if (S.qflx_snomelt_lyr)
        for (int j = -V.nlevsno + 1; j <= 0; ++j) { S.qflx_snomelt_lyr[V.sno(c, j)] = 0.0; S.qflx_snofrz_lyr[V.sno(c, j)] = 0.0; }
    for (int j = ltop; j <= nlevgrnd; ++j) {
        double t = S.t_soisno[V.sno(c, j)];
        double liq = S.h2osoi_liq[V.sno(c, j)], ice = S.h2osoi_ice[V.sno(c, j)];
        const double wice0 = ice, wmass0 = ice + liq + 0.0;
        const double fact_j = S.fact[V.sno(c, j)];
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
            if (h2osno_no_layers > 0.0 && j == 1) {
                if (t > TFRZ) { imelt = 1; tinc = TFRZ - t; t = TFRZ; }
            }
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
            if (j == 1) {
                if (h2osno_no_layers > 0.0 && xm > 0.0) {
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
            }
            heatr = 0.0;
            if (xm > 0.0) {
                ice = fmax90(0.0, wice0 - xm);
                if (j >= 1) heatr = hm - HFUS * (0.0 - 0.0 + wice0 - ice) / dtime;
                else heatr = hm - HFUS * (wice0 - ice) / dtime;
            } else if (xm < 0.0) {
                if (j <= 0) {
                    ice = fmin90(wmass0, wice0 - xm);
                } else {
                    if (wmass0 - 0.0 < supercool) ice = 0.0;
                    else ice = fmin90(wmass0 - 0.0 - supercool, wice0 - xm);
                }
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
                if (j <= 0) {
                    if (liq * ice > 0.0) t = TFRZ;
                }
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
        S.t_soisno[V.sno(c, j)] = t;
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
    S.t_h2osfc[V.c1(c)] = t_h2osfc;

    // ---- t_grnd (:457-477)
    const double t1 = S.t_soisno[V.sno(c, 1)];
    double tg;
    if (snl < 0) {
        const double ttop = S.t_soisno[V.sno(c, ltop)];
        if (frac_h2osfc != 0.0) tg = frac_sno * ttop + (1.0 - frac_sno - frac_h2osfc) * t1 + frac_h2osfc * t_h2osfc;
        else tg = frac_sno * ttop + (1.0 - frac_sno) * t1;
    } else {
        if (frac_h2osfc != 0.0) tg = (1.0 - frac_h2osfc) * t1 + frac_h2osfc * t_h2osfc;
        else tg = t1;
    }
    S.t_grnd[V.c1(c)] = tg;
    return 0;
}

// The lane's column (filter entry i) and its ground heat flux sums; every
// lane of the block takes part (the sums are the wave's work).  Returns false
// for a lane with nothing to solve (past the filter, or a column out of range:
// reported).
template <int BLOCK>
__device__ __forceinline__ bool column_and_ground_flux(const View &V, const ctsm_col_state &S,
                                                       const ctsm_patch_state &Pt, int n,
                                                       const int32_t *__restrict__ filter, const TempWork &wk,
                                                       ErrWord *err, int i, int &c, GroundFlux &gf)
{
    __shared__ GfWave gfl[BLOCK / 64];
    const bool active = i < n;
    c = active ? (filter ? filter[i] : V.begc + i) : V.begc;
    const bool in_range = active && col_in_range(V, c);
    ground_flux_of_wave(V, S, Pt, c, in_range, gfl[threadIdx.x >> 6], wk, i, gf);
    if (active && !in_range) report_error(err, c, CTSM_E_ARG);
    return in_range;
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_soiltemp(View V, Params P, ctsm_col_state S, ctsm_patch_state Pt, int n,
           const int32_t *__restrict__ filter, double dtime, TempWork wk, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c;
    GroundFlux gf;
    if (!column_and_ground_flux<64>(V, S, Pt, n, filter, wk, err, i, c, gf)) return;
    const int rc = soiltemp_column(V, P, S, Pt, c, dtime, wk, i, gf);
    if (rc) report_error(err, c, CTSM_E_SOLVER);
}


// SoilTemperature followed by SetSoilWaterFractions, one column per lane: the
// first kernel of the two-kernel soil step (the Richards solve follows with G
// lanes per column).
// BLOCK: 64 or 128 threads (the LDS of the ground heat flux is per wave).
// Built for three waves per SIMD (168 VGPRs, 60 B/lane of scratch): on
// 1,048,576 binned columns 3.48 ms against 4.22 ms at two waves (190 VGPRs);
// 64-thread blocks, 4.22 against 6.37 ms for 128.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_soiltemp_fractions(View V, Params P, ctsm_col_state S, ctsm_patch_state Pt, int n,
                     const int32_t *__restrict__ filter, double dtime, TempWork wk, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c;
    GroundFlux gf;
    if (!column_and_ground_flux<BLOCK>(V, S, Pt, n, filter, wk, err, i, c, gf)) return;
    int rc = soiltemp_column(V, P, S, Pt, c, dtime, wk, i, gf);
    if (rc == 0) rc = soilwater_fractions_column(V, S, c);
    if (rc) report_error(err, c, CTSM_E_SOLVER);
}

// The fused soil step in one kernel: SoilTemperature -> SetSoilWaterFractions -> SoilWater.
template <bool WT>
__global__ void __launch_bounds__(64)
k_soil_step(View V, Params P, ctsm_col_state S, ctsm_patch_state Pt, int n,
            const int32_t *__restrict__ filter, double dtime, TempWork wk, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c;
    GroundFlux gf;
    if (!column_and_ground_flux<64>(V, S, Pt, n, filter, wk, err, i, c, gf)) return;
    int rc = soiltemp_column(V, P, S, Pt, c, dtime, wk, i, gf);
    if (rc == 0) rc = soilwater_fractions_column(V, S, c);
    if (rc == 0) rc = richards_column<MAX_NLEVSOI, WT>(V, P, S, c, dtime);
    if (rc) report_error(err, c, richards_error_code(rc));
}

size_t soiltemp_workspace_doubles(int n) { return (size_t)WS_DOUBLES_PER_COL * (size_t)n; }

void launch_soiltemp(hipStream_t st, const View &V, const Params &P, const ctsm_col_state &S,
                     const ctsm_patch_state &Pt, int n, const int32_t *filter, double dtime,
                     double *ws, long ld, ErrWord *err)
{
    if (n <= 0) return;
    const GptlRange range("soiltemperature");
    const int bs = 64;
    TempWork wk{ws, ld};
    hipLaunchKernelGGL(k_soiltemp, dim3((n + bs - 1) / bs), dim3(bs), 0, st, V, P, S, Pt, n, filter, dtime, wk, err);
}

void launch_soil_step(hipStream_t st, const View &V, const Params &P, const ctsm_col_state &S,
                      const ctsm_patch_state &Pt, int n, const int32_t *filter, double dtime,
                      double *ws, long ld, ErrWord *err)
{
    if (n <= 0) return;
    const GptlRange range("soiltemperature+hydro_without_drainage");
    const int bs = 64;
    TempWork wk{ws, ld};
    if (P.lower_bc == CTSM_BC_WATERTABLE)
        hipLaunchKernelGGL(k_soil_step<true>, dim3((n + bs - 1) / bs), dim3(bs), 0, st, V, P, S, Pt, n, filter, dtime, wk,
                           err);
    else
        hipLaunchKernelGGL(k_soil_step<false>, dim3((n + bs - 1) / bs), dim3(bs), 0, st, V, P, S, Pt, n, filter, dtime,
                           wk, err);
}

void launch_soiltemp_fractions(hipStream_t st, const View &V, const Params &P, const ctsm_col_state &S,
                               const ctsm_patch_state &Pt, int n, const int32_t *filter, double dtime,
                               double *ws, long ld, ErrWord *err, int block)
{
    if (n <= 0) return;
    const GptlRange range("soiltemperature");
    TempWork wk{ws, ld};
    if (block == 64)
        hipLaunchKernelGGL(k_soiltemp_fractions<64>, dim3((n + 63) / 64), dim3(64), 0, st, V, P, S, Pt, n, filter,
                           dtime, wk, err);
    else
        hipLaunchKernelGGL(k_soiltemp_fractions<128>, dim3((n + 127) / 128), dim3(128), 0, st, V, P, S, Pt, n,
                           filter, dtime, wk, err);
}

}  // namespace ctsm
