/*
 * phs_oracle.h -- CPU ORACLE (test infrastructure only): the per-patch pieces
 * of plant hydraulic stress that the oracle's CanopyFluxes shares with
 * oracle_phs_calcstress (phs_oracle.c).  Segments 1 sun, 2 sha, 3 xyl,
 * 4 root; arrays are 0-based (x[0] = sun).
 */
#ifndef CTSM_PHS_ORACLE_H
#define CTSM_PHS_ORACLE_H

typedef struct {
    double laisun, laisha, elai, esai, tsai, htop, fdry, thm, rho, pbot, gb;
    double kmax[4], psi50[4], ck[4];
    const double *k;     /* k_soil_root(p, 1..n), stride sk */
    const double *smp;   /* smp_l(c, 1..n), stride ss */
    const double *z;     /* z(c, 1..n), stride ss */
    long sk, ss;
    int n;
} p_stress_t;

double oracle_phs_plc(double x, double psi50, double ck);
/* calcstress (PhotosynthesisMod.F90:4490-4710) for one patch: x (4) inout;
 * bsun / bsha out; night = the call took the night branch, then
 * *qflx_tran_veg is its transpiration */
void oracle_phs_calcstress_one(const p_stress_t *P, double *x, double gs_sun_in, double gs_sha_in, double qsatl,
                               double qaf, double *bsun, double *bsha, int *night, double *qflx_tran_veg);
/* getvegwp (:4979-5077): x (4) from the conductances; returns soilflux */
double oracle_phs_getvegwp_one(const p_stress_t *P, double *x, double gs_sun, double gs_sha, double qsatl,
                               double qaf);
int oracle_phs_soil_root_conductance(const ctsm_config *cfg, const ctsm_layout *L, const ctsm_phs_state *S,
                                     const ctsm_phs_params *P, int n, const int32_t *filter);

#endif
